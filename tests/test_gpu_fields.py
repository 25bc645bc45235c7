"""GPU parity of the plotting-field outputs (emu_calc_modes_2d_fields,
emu_calc_modes_1d_fields) against the oracle on an oblique-incidence cell
whose modes are mostly non-degenerate.

A mode's field is fixed up to the eigensolver's accuracy divided by the
distance to the next eigenvalue, so the columns compared are those of modes
isolated by more than 1e-3 |beta| (tolerance 1e-6 of the field's largest
value, which array_sol.f's normalisation makes 1); mode_pol, a ratio of
integrals of such fields, at 1e-6.  The normalisation itself (largest P2-node
component exactly 1) and the shapes are checked for every mode."""
import numpy as np
import pytest

from conftest import MESH_600, k_parallel_norm
from oracle import fem1d, fem2d

pytestmark = pytest.mark.gpu


def _isolated(beta):
    b = np.asarray(beta)
    d = np.abs(b[:, None] - b[None, :]) + np.eye(len(b)) * 1e300
    return np.where(d.min(axis=1) > 1e-3 * np.abs(b))[0]


@pytest.fixture(scope='module')
def case2d(gpu_ctx):
    from emustack_amd import _lib
    wl, period = 700.0, 600.0
    wln = wl / period
    k0 = 2 * np.pi / wln
    n_eff = np.array([1.0, 3.7 + 0.02j])
    kpn = k_parallel_norm(wl, 20.0, 40.0, period)
    shift = 1.1 * np.real(n_eff).max() ** 2 * k0 ** 2 - kpn @ kpn
    mh = _lib.MeshHandle(MESH_600, gpu_ctx)
    g = gpu_ctx.calc_modes_2d_fields(mh, wln, 20, 1, n_eff, kpn, shift, 5)
    m = fem2d.build_mesh(MESH_600)
    o = fem2d.calc_modes_2d(wln, 20, 1, m, n_eff, kpn, shift, 5, return_debug=True, biorth='cluster')
    return g, o, m


def test_sol1_matches_oracle(case2d):
    (beta, J, Jd, sol, pol), (bo, Jo, Jdo, dbg), m = case2d
    assert sol.shape == (3, 13, 20, m.nel)
    np.testing.assert_allclose(beta, bo, rtol=1e-9, atol=1e-9 * np.abs(bo).max())
    # array_sol.f's normalisation: the largest P2-node component is 1
    top = np.abs(sol[:, :6]).max(axis=(0, 1, 3))
    np.testing.assert_allclose(top, 1.0, rtol=0, atol=1e-12)
    so = dbg['sol1']
    iso = _isolated(bo)
    print('isolated modes: %d of %d' % (len(iso), len(bo)), flush=True)
    assert len(iso) >= 10
    err = max(np.abs(sol[:, :, k] - so[:, :, k]).max() for k in iso)
    print('sol1 max err over isolated modes %.2e' % err, flush=True)
    assert err <= 1e-6


def test_mode_pol_matches_oracle(case2d):
    (beta, J, Jd, sol, pol), (bo, Jo, Jdo, dbg), m = case2d
    po = dbg['mode_pol']
    assert pol.shape == (4, 20)
    b2 = dbg['res'][2][0]
    iso = _isolated(b2)
    err = np.abs(pol[:, iso] - po[:, iso]).max()
    print('mode_pol max err over %d isolated adjoint modes %.2e' % (len(iso), err), flush=True)
    assert err <= 1e-6
    np.testing.assert_allclose(pol[:3].sum(axis=0), 1.0, rtol=0, atol=1e-12)


def test_dropin_plot_modes_2d(gpu_ctx, case2d):
    """EMUstack.calc_modes_2d with plot_modes = 1 returns the same arrays as
    the entry point, in the reference's tuple positions."""
    import os
    from emustack_amd import EMUstack
    (beta, J, Jd, sol, pol), _, m = case2d
    wl, period = 700.0, 600.0
    wln = wl / period
    k0 = 2 * np.pi / wln
    n_eff = np.array([1.0, 3.7 + 0.02j])
    kpn = k_parallel_norm(wl, 20.0, 40.0, period)
    shift = 1.1 * np.real(n_eff).max() ** 2 * k0 ** 2 - kpn @ kpn
    d, f = os.path.split(MESH_600)
    out = EMUstack.calc_modes_2d(wln, 20, 1, 0, d + '/', f, m.npt, m.nel, 2, n_eff, kpn, shift, 1, 2, 30, 1, 1, 0,
                                 0, 5, 0, 0, 0)
    assert out[3].shape == (3, 13, 20, m.nel) and out[4].shape == (4, 20)
    np.testing.assert_allclose(out[0], beta, rtol=1e-12)
    np.testing.assert_allclose(out[3], sol, rtol=0, atol=1e-9)
    np.testing.assert_allclose(out[4], pol, rtol=0, atol=1e-9)
    plain = EMUstack.calc_modes_2d(wln, 20, 1, 0, d + '/', f, m.npt, m.nel, 2, n_eff, kpn, shift, 1, 2, 30, 0, 1,
                                   0, 0, 5, 0, 0, 0)
    assert plain[3] is None and plain[4] is None


def test_sol_p2_1d_matches_oracle(gpu_ctx):
    from emustack_amd import _lib
    period, wl, order = 760.0, 800.0, 6
    x, tn, te = fem1d.grating_mesh_1d(period, [190, 190], 0.005)
    wln = wl / period
    k0 = 2 * np.pi / wln
    th, ph = np.radians(20.0), np.radians(0.0)
    kp = 2 * np.pi / wl * np.sin(th) * np.array([np.cos(ph), np.sin(ph)]) * period
    neff = np.array([3.61, 1.46, 1.46], complex)
    shift = 1.1 * np.real(neff).max() ** 2 * k0 ** 2 - kp @ kp
    mh = _lib.MeshHandle1D(x, tn, te, gpu_ctx)
    beta, J, Jd, _, _, sol = gpu_ctx.calc_modes_1d_fields(mh, wln, 30, order, neff, kp, shift, 2 * order + 1)
    m = fem1d.build_mesh_1d(x, tn, te)
    bo, Jo, Jdo, so = fem1d.calc_modes_1d(wln, 30, order, m, neff, kp[0], kp[1], shift, biorth='cluster',
                                          return_sol=True)
    assert sol.shape == (3, 3, 30, m.nel)
    np.testing.assert_allclose(beta, bo, rtol=1e-9, atol=1e-9 * np.abs(bo).max())
    iso = _isolated(bo)
    print('isolated 1D modes: %d of %d' % (len(iso), len(bo)), flush=True)
    assert len(iso) >= 15
    scale = np.abs(so).max()
    err = max(np.abs(sol[:, :, k] - so[:, :, k]).max() for k in iso) / scale
    print('sol_P2 max rel err over isolated modes %.2e' % err, flush=True)
    assert err <= 1e-6
