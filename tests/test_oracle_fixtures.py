"""The oracle against the reference's own golden R/T/A files
(tests/ref/case_{0,1,2}/*.txt of the reference, format %18.11f, compared
there at rtol = atol = 1e-6)."""
import numpy as np
import pytest

from conftest import MESH_600, k_parallel_norm, load_txt
from oracle import fem2d, stack_ref


def test_case0_thin_film_multistack(mats):
    """Five analytic films (tests/test_case_0_thin_film_multistack.py:43-104)."""
    period = 1.0
    A, R, T = [], [], []
    for wl in (400.0, 700.0, 1000.0):
        kpn = k_parallel_norm(wl, 0, 0, period)
        nag, nsi = mats('Ag', wl), mats('Si_c', wl)
        f1 = dict(kind='tf', n=3.6 + 0.27j, h=50 / period)
        f2 = dict(kind='tf', n=nsi, h=200 / period)
        L = [dict(kind='tf', n=1.0, h=None), dict(kind='tf', n=nag, h=100 / period), f1, f2, f1, f2,
             dict(kind='tf', n=3.5, h=None)]
        out = stack_ref.calc_scat(L, wl / period, kpn, 1, pol='TM')
        A.append(out['a_list'])
        R.append(out['r_list'][-1])
        T.append(out['t_list'][-1])
    A = np.array(A, float)
    np.testing.assert_allclose(A[:, -1], load_txt('case_0', 'Absorptance_stack0001.txt')[:, 1], 0, 1e-8)
    np.testing.assert_allclose(R, load_txt('case_0', 'Reflectance_stack0001.txt')[:, 1], 0, 1e-8)
    np.testing.assert_allclose(T, load_txt('case_0', 'Transmittance_stack0001.txt')[:, 1], 0, 1e-8)
    for i in range(5):
        ref = load_txt('case_0', 'Lay_Absorb_%d_stack0001.txt' % i)[:, 1]
        np.testing.assert_allclose(A[:, i], ref, 0, 1e-8)


@pytest.fixture(scope='module')
def mesh600():
    return fem2d.build_mesh(MESH_600)


def _nanowire(mesh, mats, wl, theta, phi, nval, order, pol):
    period = 600.0
    wln = wl / period
    kpn = k_parallel_norm(wl, theta, phi, period)
    n_eff = np.array([1.0, mats('Si_c', wl)])
    k0 = 2 * np.pi / wln
    shift = 1.1 * np.real(n_eff).max() ** 2 * k0 ** 2 - kpn[0] ** 2 - kpn[1] ** 2
    P = len(fem2d.pw_list(order))
    # group normalisation of near-degenerate modes: with the reference's
    # per-mode one, R/T/A of this normal-incidence cell depends at ~1e-8 on
    # the basis ARPACK's random start vector picks inside its near-degenerate
    # pairs, which the golden files (11 decimals) resolve
    beta, J, Jd = fem2d.calc_modes_2d(wln, nval, order, mesh, n_eff, kpn, shift, P, biorth='cluster')
    layers = [dict(kind='tf', n=mats('SiO2', wl).real, h=None),
              dict(kind='ns', kz=beta, J=J, Jd=Jd, h=2330 / period), dict(kind='tf', n=1.0, h=None)]
    return stack_ref.calc_scat(layers, wln, kpn, order, pol=pol)


def test_case1_simple_struct(mesh600, mats):
    """Si nanowire array, 518-triangle fixture mesh, 40 Bloch modes
    (tests/test_case_1_simple_struct.py:47-118)."""
    outs = [_nanowire(mesh600, mats, wl, 0, 0, 40, 2, 'TE') for wl in (500.0, 1000.0)]
    got = {'Absorptance': [o['a_list'][-1] for o in outs], 'Reflectance': [o['r_list'][-1] for o in outs],
           'Transmittance': [o['t_list'][-1] for o in outs], 'Lay_Absorb_0': [o['a_list'][0] for o in outs],
           'Lay_Trans_0': [o['t_list'][0] for o in outs]}
    for name, vals in got.items():
        fn = name + '_stack0001.txt' if not name.startswith('Lay') else name + '_stack0001.txt'
        ref = load_txt('case_1', fn)[:, 1]
        np.testing.assert_allclose(np.array(vals, float), ref, 0, 1e-8, err_msg=name)


def test_case2_oblique(mesh600, mats):
    """Same mesh at theta=20, phi=40, 20 Bloch modes, order 1
    (tests/test_case_2_simple_non_norm_inc.py:40-97)."""
    o = _nanowire(mesh600, mats, 500.0, 20, 40, 20, 1, 'TE')
    np.testing.assert_allclose(o['a_list'][-1], load_txt('case_2', 'Absorptance_stack0001.txt')[1], 0, 1e-8)
    np.testing.assert_allclose(o['r_list'][-1], load_txt('case_2', 'Reflectance_stack0001.txt')[1], 0, 1e-8)
    np.testing.assert_allclose(o['t_list'][-1], load_txt('case_2', 'Transmittance_stack0001.txt')[1], 0, 1e-8)
