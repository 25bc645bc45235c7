"""GPU parity tests: every stage of the HIP path against the oracle on the
same inputs, and the full drop-in API against the reference's golden R/T/A.
Tolerances: assembly 1e-13 relative (same FP64 arithmetic, different
summation order); LU residual 1e-8; beta 1e-9 relative; J J^dagger 1e-8 of
its largest entry; R/T/A 1e-8 against the reference files (the reference
itself compares at 1e-6) and 1e-9 against the oracle with the same
normalisation of near-degenerate mode groups."""
import os

import numpy as np
import pytest

from conftest import MESH_600, GOLDEN, k_parallel_norm, load_txt
from oracle import fem2d, stack_ref

pytestmark = pytest.mark.gpu


def _params(wl, period, n_incl, theta=0.0, phi=0.0):
    wln = wl / period
    k0 = 2 * np.pi / wln
    n_eff = np.array([1.0, n_incl])
    kpn = k_parallel_norm(wl, theta, phi, period)
    shift = 1.1 * np.real(n_eff).max() ** 2 * k0 ** 2 - kpn @ kpn
    return wln, k0, n_eff, kpn, shift


@pytest.fixture(scope='module')
def mesh600(gpu_ctx):
    from emustack_amd import _lib
    return _lib.MeshHandle(MESH_600, gpu_ctx), fem2d.build_mesh(MESH_600)


def test_assembly_matches_oracle(gpu_ctx, mesh600):
    mh, m = mesh600
    rng = np.random.default_rng(0)
    cases = []
    for i in range(3):
        wln = 0.7 + 0.4 * i
        k0 = 2 * np.pi / wln
        eps = np.array([1.0 + 0.1 * i, (3.0 + 0.3 * i + 0.02j) ** 2])
        bl = rng.uniform(-1, 1, 2)
        cases.append((np.ones(2, complex), eps * k0 ** 2, 1.1 * 3.6 ** 2 * k0 ** 2 + 0.5j * i, bl))
    m1, m2 = gpu_ctx.assemble(mh, [c[0] for c in cases], [c[1] for c in cases], [c[2] for c in cases],
                              [c[3] for c in cases])
    for i, (pp, qq, sh, bl) in enumerate(cases):
        A, B = fem2d.assemble(m, pp, qq, sh, bl)
        A.sort_indices()
        B.sort_indices()
        assert np.abs(m1[i] - A.data).max() <= 1e-13 * np.abs(A.data).max()
        assert np.abs(m2[i] - B.data).max() <= 1e-13 * np.abs(B.data).max()


def test_lu_solve_residual(gpu_ctx, mesh600):
    mh, m = mesh600
    mats = []
    for wl in (500.0, 900.0):
        wln, k0, n_eff, kpn, shift = _params(wl, 600., 3.5 + 0.01j)
        A, _ = fem2d.assemble(m, np.ones(2, complex), n_eff ** 2 * k0 ** 2, shift, kpn)
        A.sort_indices()
        mats.append(A)
    rng = np.random.default_rng(1)
    b = rng.standard_normal((2, 3, m.neq)) + 1j * rng.standard_normal((2, 3, m.neq))
    x, npert = gpu_ctx.lu_solve(mh, np.stack([A.data for A in mats]), b)
    assert npert == 0
    for i, A in enumerate(mats):
        for q in range(3):
            r = A @ x[i, q] - b[i, q]
            assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(b[i, q])


@pytest.fixture(scope='module')
def case1_500(gpu_ctx, mesh600):
    from conftest import mats as _m  # noqa: F401
    import json
    d = json.load(open(os.path.join(GOLDEN, 'materials.json')))
    nsi = complex(*d['Si_c']['500.0'])
    wln, k0, n_eff, kpn, shift = _params(500.0, 600., nsi)
    mh, m = mesh600
    g = gpu_ctx.calc_modes_batch(mh, [wln], 40, 2, [n_eff], [kpn], [shift], 13)
    # the oracle with the GPU's normalisation of near-degenerate groups
    o = fem2d.calc_modes_2d(wln, 40, 2, m, n_eff, kpn, shift, 13, biorth='cluster')
    return g, o, (wln, kpn)


def test_beta_matches_oracle(case1_500):
    (beta, J, Jd, st), (bo, Jo, Jdo), _ = case1_500
    assert st[2] == 0  # no pivot perturbations
    np.testing.assert_allclose(beta[0], bo, rtol=1e-9, atol=1e-9 * np.abs(bo).max())


def test_mode_projectors_match_oracle(case1_500):
    """J J^dagger is invariant to rotations inside (near-)degenerate mode
    groups and to mode scaling, so it compares the mode subspaces."""
    (beta, J, Jd, st), (bo, Jo, Jdo), _ = case1_500
    Pg = J[0] @ Jd[0]
    Po = Jo @ Jdo
    assert np.abs(Pg - Po).max() <= 1e-8 * np.abs(Po).max()
    # non-degenerate modes agree column by column (eigenvector accuracy of a
    # 1e-12 Ritz residual divided by the eigenvalue gap)
    gaps = np.abs(np.diff(bo))
    iso = [i for i in range(1, len(bo) - 1) if min(gaps[i - 1], gaps[i]) > 1e-3 * abs(bo[i])]
    assert len(iso) > 5
    for i in iso:
        np.testing.assert_allclose(J[0][:, i], Jo[:, i], rtol=0, atol=1e-5 * np.abs(Jo[:, i]).max())


def test_rta_matches_reference_and_oracle(case1_500, mats):
    (beta, J, Jd, st), (bo, Jo, Jdo), (wln, kpn) = case1_500
    res = {}
    for name, (b, JJ, JJd) in (('gpu', (beta[0], J[0], Jd[0])), ('oracle', (bo, Jo, Jdo))):
        layers = [dict(kind='tf', n=mats('SiO2', 500.0).real, h=None),
                  dict(kind='ns', kz=b, J=JJ, Jd=JJd, h=2330 / 600.), dict(kind='tf', n=1.0, h=None)]
        o = stack_ref.calc_scat(layers, wln, kpn, 2, pol='TE')
        res[name] = np.array([o['a_list'][-1], o['r_list'][-1], o['t_list'][-1]], float)
    ref = np.array([load_txt('case_1', f)[0, 1] for f in ('Absorptance_stack0001.txt', 'Reflectance_stack0001.txt',
                                                         'Transmittance_stack0001.txt')])
    # the golden files carry the per-mode normalisation's ~5e-9 ambiguity of
    # this normal-incidence cell (its near-degenerate pairs)
    np.testing.assert_allclose(res['gpu'], ref, rtol=0, atol=1e-8)
    np.testing.assert_allclose(res['gpu'], res['oracle'], rtol=0, atol=1e-9)


def test_batched_equals_single(gpu_ctx, mesh600, mats):
    mh, m = mesh600
    wls = [500.0, 1000.0]
    P = [_params(wl, 600., mats('Si_c', wl), 20, 40) for wl in wls]
    bb, JJ, JJd, _ = gpu_ctx.calc_modes_batch(mh, [p[0] for p in P], 20, 1, [p[2] for p in P], [p[3] for p in P],
                                              [p[4] for p in P], 5)
    for i, p in enumerate(P):
        b1, J1, Jd1, _ = gpu_ctx.calc_modes_batch(mh, [p[0]], 20, 1, [p[2]], [p[3]], [p[4]], 5)
        np.testing.assert_allclose(bb[i], b1[0], rtol=1e-9)
        np.testing.assert_allclose(JJ[i] @ JJd[i], J1[0] @ Jd1[0], rtol=0, atol=1e-8 * np.abs(J1[0] @ Jd1[0]).max())


def test_homogeneous_layer_raises(gpu_ctx, mesh600):
    mh, m = mesh600
    from emustack_amd import _lib
    with pytest.raises(_lib.EmuError) as e:
        gpu_ctx.calc_modes_batch(mh, [1.0], 20, 1, [[1.5, 1.5]], [[1e-7, 1e-7]], [10.0], 5)
    assert e.value.code == -5


def test_nonconvergence_is_an_error_and_the_context_stays_usable(gpu_ctx, mesh600, mats):
    """py_calc_modes.f:647-656 stops the process when ARPACK does not
    converge; here the call returns EMU_E_NOCONV (itermax = 0 with a Krylov
    dimension just above 40 wanted modes cannot converge in one expansion)
    and the next call on the same context and mesh gives the converged
    result."""
    from emustack_amd import _lib
    mh, m = mesh600
    p = _params(500.0, 600., mats('Si_c', 500.0), 20, 40)
    with pytest.raises(_lib.EmuError) as e:
        gpu_ctx.calc_modes_batch(mh, [p[0]], 40, 1, [p[2]], [p[3]], [p[4]], 5, itermax=0, ncv=48)
    assert e.value.code == -4
    b1, _, _, st = gpu_ctx.calc_modes_batch(mh, [p[0]], 20, 1, [p[2]], [p[3]], [p[4]], 5)
    b2, _, _, _ = gpu_ctx.calc_modes_batch(mh, [p[0]], 20, 1, [p[2]], [p[3]], [p[4]], 5)
    np.testing.assert_array_equal(b1, b2)
    assert st[1] <= 30


def _api_stack(wl, theta, phi, nbm, order, pol):
    from emustack_amd import objects, materials, Stack
    import json
    d = json.load(open(os.path.join(GOLDEN, 'materials.json')))
    si = materials.Material(complex(*d['Si_c'][repr(wl)]))
    sio2 = materials.Material(complex(*d['SiO2'][repr(wl)]))
    light = objects.Light(wl, max_order_PWs=order, theta=theta, phi=phi)
    nw = objects.NanoStruct('2D_array', 600, 120, height_nm=2330, inclusion_a=si, background=materials.Air,
                            loss=True, make_mesh_now=False, mesh_file=MESH_600)
    sup = objects.ThinFilm(period=600, height_nm='semi_inf', material=materials.Air, loss=False)
    sub = objects.ThinFilm(period=600, height_nm='semi_inf', material=sio2, loss=False)
    st = Stack((sub.calc_modes(light), nw.calc_modes(light, num_BMs=nbm), sup.calc_modes(light)))
    st.calc_scat(pol=pol)
    return st


def test_dropin_api_case1():
    """tests/test_case_1_simple_struct.py through the drop-in objects."""
    stacks = [_api_stack(wl, 0, 0, 40, 2, 'TE') for wl in (500.0, 1000.0)]
    for fn, get in (('Absorptance_stack0001.txt', lambda s: s.a_list[-1]),
                    ('Lay_Absorb_0_stack0001.txt', lambda s: s.a_list[0]),
                    ('Lay_Trans_0_stack0001.txt', lambda s: s.t_list[0]),
                    ('Reflectance_stack0001.txt', lambda s: s.r_list[-1]),
                    ('Transmittance_stack0001.txt', lambda s: s.t_list[-1])):
        ref = load_txt('case_1', fn)[:, 1]
        np.testing.assert_allclose([np.real(np.asarray(get(s))).ravel()[0] for s in stacks], ref, rtol=0, atol=1e-8,
                                   err_msg=fn)


def test_dropin_api_case2():
    """tests/test_case_2_simple_non_norm_inc.py through the drop-in objects."""
    s = _api_stack(500.0, 20, 40, 20, 1, 'TE')
    for fn, v in (('Absorptance_stack0001.txt', s.a_list[-1]), ('Reflectance_stack0001.txt', s.r_list[-1]),
                  ('Transmittance_stack0001.txt', s.t_list[-1])):
        np.testing.assert_allclose(float(np.real(v)), load_txt('case_2', fn)[1], rtol=0, atol=1e-8, err_msg=fn)


def test_stack_batch_equals_per_layer(gpu_ctx):
    """batch.calc_modes_stack (all layers' sub-batches on one thread pool, one
    context per thread) gives what calc_modes_many gives layer by layer."""
    from emustack_amd import objects, materials
    from emustack_amd.batch import calc_modes_many, calc_modes_stack
    import json
    d = json.load(open(os.path.join(GOLDEN, 'materials.json')))
    lights = [objects.Light(wl, max_order_PWs=1, theta=0.0, phi=0.0) for wl in (500.0, 1000.0)]
    structs = [objects.NanoStruct('2D_array', 600, 120, height_nm=h, background=materials.Air, loss=True,
                                  inclusion_a=materials.Material(complex(*d['Si_c']['500.0'])) if h < 300 else
                                  materials.Material(3.0 + 0.01j), make_mesh_now=False, mesh_file=MESH_600)
               for h in (200.0, 400.0)]
    many = [calc_modes_many(s, lights, num_BMs=20, groups=2) for s in structs]
    from emustack_amd.batch import preload_meshes
    preload_meshes(structs, groups=2)  # every pool thread's context holds the mesh beforehand
    stack = calc_modes_stack(structs, lights, num_BMs=20, groups=2)
    for a, b in zip(many, stack):
        for sa, sb in zip(a, b):
            np.testing.assert_allclose(np.asarray(sb.k_z), np.asarray(sa.k_z), rtol=1e-9)
            P1, P2 = np.asarray(sa.J) @ np.asarray(sa.J_dag), np.asarray(sb.J) @ np.asarray(sb.J_dag)
            np.testing.assert_allclose(P2, P1, rtol=0, atol=1e-8 * np.abs(P1).max())


@pytest.mark.parametrize('block', [1, 2, 4])
def test_small_blocks_match_oracle(case1_500, gpu_ctx, mesh600, block):
    """Block sizes below 8 take the non-delayed block CGS2 path of the
    Krylov-Schur expansion; the eigenpairs must not depend on it."""
    (beta8, J8, Jd8, _), (bo, Jo, Jdo), (wln, kpn) = case1_500
    mh, m = mesh600
    import json
    d = json.load(open(os.path.join(GOLDEN, 'materials.json')))
    nsi = complex(*d['Si_c']['500.0'])
    _, _, n_eff, _, shift = _params(500.0, 600., nsi)
    beta, J, Jd, st = gpu_ctx.calc_modes_batch(mh, [wln], 40, 2, [n_eff], [kpn], [shift], 13, block=block)
    print('block %d: OP applications %d, restarts %d' % (block, st[0], st[1]), flush=True)
    np.testing.assert_allclose(beta[0], bo, rtol=1e-9, atol=1e-9 * np.abs(bo).max())
    Po = Jo @ Jdo
    assert np.abs(J[0] @ Jd[0] - Po).max() <= 1e-8 * np.abs(Po).max()


def test_submitted_batches_equal_blocking_calls(gpu_ctx, mesh600, mats):
    """emu_calc_modes_2d_submit / emu_job_wait: two batches queued on one
    context and one on a second context give what the blocking calls give."""
    from emustack_amd import _lib
    mh, m = mesh600
    wls = [500.0, 700.0, 1000.0]
    P = [_params(wl, 600., mats('Si_c', wl), 20, 40) for wl in wls]
    ref = [gpu_ctx.calc_modes_batch(mh, [p[0]], 20, 1, [p[2]], [p[3]], [p[4]], 5) for p in P]
    ctx2 = _lib.Context(0)
    mh2 = _lib.MeshHandle(MESH_600, ctx2)
    jobs = [gpu_ctx.calc_modes_batch_submit(mh, [P[0][0]], 20, 1, [P[0][2]], [P[0][3]], [P[0][4]], 5),
            gpu_ctx.calc_modes_batch_submit(mh, [P[1][0]], 20, 1, [P[1][2]], [P[1][3]], [P[1][4]], 5),
            ctx2.calc_modes_batch_submit(mh2, [P[2][0]], 20, 1, [P[2][2]], [P[2][3]], [P[2][4]], 5)]
    for j, r in zip(jobs, ref):
        b, J, Jd, st = j.wait()
        np.testing.assert_allclose(b[0], r[0][0], rtol=1e-12)
        np.testing.assert_allclose(J[0] @ Jd[0], r[1][0] @ r[2][0], rtol=0,
                                   atol=1e-10 * np.abs(r[1][0] @ r[2][0]).max())
        assert st[1] == r[3][1]
    with pytest.raises(RuntimeError):
        jobs[0].wait()
