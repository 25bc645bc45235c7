"""GPU star-product cascade (emu_star_cascade_batch via batch.calc_scat_many)
against Stack.calc_scat on the same layers: a 7-layer stack of three
structured layers (GPU-computed modes of the case-1 cell) between films, for
two wavelength points at once, and the reference's case 1
(tests/test_case_1_simple_struct.py) with GPU-computed modes.  Tolerances:
1e-9 relative on every a/r/t entry and on R_net, T_net (two FP64 evaluations
of the same formulas with different solve orders); case 1 within 1e-8 of the
reference's files and 1e-10 of the numpy Stack; case 0 (films only) within
1e-8 of its files; a thickness sweep over shared layers against the numpy
Stack."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _stacks(wls=(500.0, 700.0)):
    """Seven layers bottom to top: SiO2 substrate, three structured layers
    (the case-1 nanowire cell with GPU-computed modes, heights 200 / 150 /
    100 nm) separated by two lossy 50 nm films, air superstrate."""
    import json
    import os
    from conftest import GOLDEN, MESH_600
    from emustack_amd import objects, materials, Stack
    from emustack_amd.batch import calc_modes_many
    d = json.load(open(os.path.join(GOLDEN, 'materials.json')))
    out = []
    for wl in wls:
        light = objects.Light(wl, max_order_PWs=2, theta=0.0, phi=0.0)
        si = materials.Material(complex(*d['Si_c']['500.0']))
        sims = []
        for h in (200, 150, 100):
            nw = objects.NanoStruct('2D_array', 600, 120, height_nm=h, inclusion_a=si, background=materials.Air,
                                    loss=True, make_mesh_now=False, mesh_file=MESH_600)
            sims.append(calc_modes_many(nw, [light], num_BMs=40)[0])
        sub = objects.ThinFilm(600, height_nm='semi_inf', material=materials.Material(1.45 + 0j), loss=False)
        sup = objects.ThinFilm(600, height_nm='semi_inf', material=materials.Air, loss=False)
        film = objects.ThinFilm(600, height_nm=50, material=materials.Material(1.46 + 0.02j))
        f = film.calc_modes(light)
        out.append(Stack((sub.calc_modes(light), sims[0], f, sims[1], f, sims[2], sup.calc_modes(light))))
    return out


def test_cascade_matches_numpy_stack(gpu_ctx):
    from emustack_amd.batch import calc_scat_many
    stacks = _stacks()
    ref = []
    for st in stacks:
        st.calc_scat(pol='TE')
        ref.append((list(st.a_list), list(st.t_list), st.r_list[-1], np.asarray(st.R_net), np.asarray(st.T_net)))
    calc_scat_many(stacks, pol='TE', want_nets=True)
    flt = lambda xs: np.array([float(np.asarray(x).real.ravel()[0]) for x in xs])
    for (a, t, r, R, T), g in zip(ref, stacks):
        np.testing.assert_allclose(flt(g.a_list), flt(a), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(flt(g.t_list), flt(t), rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(float(g.r_list[-1]), float(r), rtol=1e-9)
        np.testing.assert_allclose(np.asarray(g.R_net), R, rtol=0, atol=1e-9 * np.abs(R).max())
        np.testing.assert_allclose(np.asarray(g.T_net), T, rtol=0, atol=1e-9 * np.abs(T).max())


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
        np.testing.assert_allclose([float(np.real(k(s))) for s in stacks], ref, rtol=0, atol=1e-8, err_msg=fn)
    for s in stacks:
        got = (float(s.a_list[-1]), float(s.r_list[-1]), float(s.t_list[-1]))
        s.calc_scat(pol='TE')
        np.testing.assert_allclose(got, (float(s.a_list[-1]), float(s.r_list[-1]), float(s.t_list[-1])),
                                   rtol=0, atol=1e-10)


def test_heights_sweep_with_shared_layers(gpu_ctx):
    """Stacks built from the same layer objects with different heights_nm
    (a thickness sweep): every stack keeps its own heights in the batched
    cascade.  Against Stack.calc_scat at 1e-6 (the 5 and 10 nm films make
    the air-gap inversions ill-conditioned: two FP64 evaluations of the
    same formulas differ by ~1e-7 here), far below the spread between the
    stacks."""
    from emustack_amd import Stack
    from emustack_amd.batch import calc_scat_many
    base = _stacks(wls=(600.0,))[0]
    sweeps = ([200, 50, 150, 50, 100], [120, 80, 60, 30, 240], [10, 5, 300, 5, 20])
    stacks = [Stack(base.layers, heights_nm=h) for h in sweeps]
    ref = []
    for st in stacks:
        st.calc_scat(pol='TE')
        ref.append((float(st.a_list[-1]), float(st.r_list[-1]), float(st.t_list[-1])))
    calc_scat_many(stacks, pol='TE')
    got = [(float(st.a_list[-1]), float(st.r_list[-1]), float(st.t_list[-1])) for st in stacks]
    np.testing.assert_allclose(got, ref, rtol=0, atol=1e-6)
    assert min(abs(ref[i][k] - ref[j][k]) for i in range(3) for j in range(i) for k in (0, 2)) > 1e-2


def test_case0_films_through_gpu_cascade(gpu_ctx, mats):
    """The reference's case 0 (tests/test_case_0_thin_film_multistack.py:
    five analytic films, period 1 nm, max_order_PWs 1, TM, 400/700/1000 nm)
    through the batched GPU cascade, against its golden files at 1e-8."""
    from conftest import load_txt
    from emustack_amd import objects, materials, Stack
    from emustack_amd.batch import calc_scat_many
    stacks = []
    for wl in (400.0, 700.0, 1000.0):
        light = objects.Light(wl, max_order_PWs=1, theta=0.0, phi=0.0)
        sup = objects.ThinFilm(period=1, height_nm='semi_inf', material=materials.Material(3.5 + 0.0j), loss=False)
        f1 = objects.ThinFilm(period=1, height_nm=50, material=materials.Material(3.6 + 0.27j), loss=True)
        f2 = objects.ThinFilm(period=1, height_nm=200, material=materials.Material(mats('Si_c', wl)), loss=True)
        mirror = objects.ThinFilm(period=1, height_nm=100, material=materials.Material(mats('Ag', wl)), loss=True)
        sub = objects.ThinFilm(period=1, height_nm='semi_inf', material=materials.Air, loss=False)
        s1, s2 = f1.calc_modes(light), f2.calc_modes(light)
        stacks.append(Stack((sub.calc_modes(light), mirror.calc_modes(light), s1, s2, s1, s2,
                             sup.calc_modes(light))))
    calc_scat_many(stacks, pol='TM')
    checks = {'Absorptance': lambda s: s.a_list[-1], 'Reflectance': lambda s: s.r_list[-1],
              'Transmittance': lambda s: s.t_list[-1]}
    for i in range(5):
        checks['Lay_Absorb_%d' % i] = (lambda i: lambda s: s.a_list[i])(i)
    for name, get in checks.items():
        ref = load_txt('case_0', name + '_stack0001.txt')[:, 1]
        got = np.array([np.real(np.asarray(get(s))).ravel()[0] for s in stacks])
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-8, err_msg=name)
