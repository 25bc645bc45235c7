"""GPU star-product cascade (emu_star_cascade_batch via batch.calc_scat_many)
against Stack.calc_scat on the same layers: a 7-layer stack of three
structured layers with synthetic Bloch-mode data (S-matrix blocks of the
size of the benchmark: 49 plane waves per polarisation, 60 modes) between
films, for several wavelength points at once, and the reference's case 1
(tests/test_case_1_simple_struct.py) with GPU-computed modes.  Tolerances:
1e-9 relative on every a/r/t entry and on R_net, T_net (two FP64 evaluations
of the same formulas with different solve orders); case 1 within 1e-6 of the
reference's files and 1e-10 of the numpy Stack."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class _FakeSim:
    """A structured layer's data as Simmo holds it (k_z, J, J_dag)."""

    def __init__(self, structure, light, rng, nm, P2):
        from emustack_amd.mode_calcs import Modes
        self.structure = structure
        self.light = light
        self.k_z = rng.standard_normal(nm) * 5 + 1j * np.abs(rng.standard_normal(nm))
        self.J = np.asmatrix((rng.standard_normal((P2, nm)) + 1j * rng.standard_normal((P2, nm))) * 0.05)
        self.J_dag = np.asmatrix((rng.standard_normal((nm, P2)) + 1j * rng.standard_normal((nm, P2))) * 0.05)

    def air_ref(self):
        return self.light._air_ref(self.structure.period, self.structure.period_y, self.structure.world_1d)


def _stacks(npts, seed=0):
    from emustack_amd import objects, materials, Stack
    from emustack_amd.mode_calcs import Simmo
    rng = np.random.default_rng(seed)
    stacks = []
    for k in range(npts):
        wl = 500.0 + 40 * k
        light = objects.Light(wl, max_order_PWs=4, theta=0.0, phi=0.0)
        sub = objects.ThinFilm(500, height_nm='semi_inf', material=materials.Material(1.5 + 0j), loss=False)
        sup = objects.ThinFilm(500, height_nm='semi_inf', material=materials.Air, loss=False)
        film = objects.ThinFilm(500, height_nm=50, material=materials.Material(1.46 + 0.01j))
        ns = objects.NanoStruct('2D_array', 500, 200, height_nm=120, make_mesh_now=False, mesh_file='x.mail')
        a_sub, a_sup, a_f = sub.calc_modes(light), sup.calc_modes(light), film.calc_modes(light)
        P2 = len(a_sup.k_z)
        sims = []
        for h in (100, 150, 120):
            s = _FakeSim(ns, light, rng, 60, P2)
            s.__class__ = type('FakeSimmo', (Simmo,), {'air_ref': _FakeSim.air_ref})
            s.structure = objects.NanoStruct('2D_array', 500, 200, height_nm=h, make_mesh_now=False,
                                             mesh_file='x.mail')
            sims.append(s)
        stacks.append(Stack((a_sub, sims[0], a_f, sims[1], a_f, sims[2], a_sup)))
    return stacks


def test_cascade_matches_numpy_stack(gpu_ctx):
    from emustack_amd.batch import calc_scat_many
    ref = _stacks(4)
    for st in ref:
        st.calc_scat(pol='TE')
    got = calc_scat_many(_stacks(4), pol='TE', want_nets=True)
    for r, g in zip(ref, got):
        np.testing.assert_allclose(np.array(g.a_list, float), np.array(r.a_list, float), rtol=1e-9)
        np.testing.assert_allclose(np.array(g.t_list, float), np.array(r.t_list, float), rtol=1e-9)
        np.testing.assert_allclose(float(g.r_list[-1]), float(r.r_list[-1]), rtol=1e-9)
        np.testing.assert_allclose(np.asarray(g.R_net), np.asarray(r.R_net), rtol=1e-9,
                                   atol=1e-9 * np.abs(r.R_net).max())
        np.testing.assert_allclose(np.asarray(g.T_net), np.asarray(r.T_net), rtol=1e-9,
                                   atol=1e-9 * np.abs(r.T_net).max())


def test_case1_through_gpu_cascade(gpu_ctx):
    import json
    import os
    from conftest import GOLDEN, MESH_600, load_txt
    from emustack_amd import objects, materials, Stack
    from emustack_amd.batch import calc_modes_many, calc_scat_many
    d = json.load(open(os.path.join(GOLDEN, 'materials.json')))
    wls = (500.0, 1000.0)
    lights = [objects.Light(wl, max_order_PWs=2, theta=0.0, phi=0.0) for wl in wls]
    si = materials.Material(complex(*d['Si_c']['500.0']))
    stacks = []
    for wl, light in zip(wls, lights):
        si = materials.Material(complex(*d['Si_c'][repr(wl)]))
        sio2 = materials.Material(complex(*d['SiO2'][repr(wl)]))
        nw = objects.NanoStruct('2D_array', 600, 120, height_nm=2330, inclusion_a=si, background=materials.Air,
                                loss=True, make_mesh_now=False, mesh_file=MESH_600)
        sim = calc_modes_many(nw, [light], num_BMs=40)[0]
        sup = objects.ThinFilm(period=600, height_nm='semi_inf', material=materials.Air, loss=False)
        sub = objects.ThinFilm(period=600, height_nm='semi_inf', material=sio2, loss=False)
        stacks.append(Stack((sub.calc_modes(light), sim, sup.calc_modes(light))))
    calc_scat_many(stacks, pol='TE')
    for k, fn in ((lambda s: s.a_list[-1], 'Absorptance_stack0001.txt'),
                  (lambda s: s.r_list[-1], 'Reflectance_stack0001.txt'),
                  (lambda s: s.t_list[-1], 'Transmittance_stack0001.txt')):
        ref = load_txt('case_1', fn)[:, 1]
        np.testing.assert_allclose([float(k(s)) for s in stacks], ref, rtol=1e-6, atol=1e-6, err_msg=fn)
    for s in stacks:
        got = (float(s.a_list[-1]), float(s.r_list[-1]), float(s.t_list[-1]))
        s.calc_scat(pol='TE')
        np.testing.assert_allclose(got, (float(s.a_list[-1]), float(s.r_list[-1]), float(s.t_list[-1])),
                                   rtol=0, atol=1e-10)
