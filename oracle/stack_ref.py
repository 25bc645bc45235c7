"""CPU oracle: numpy restatement of EMUstack's thin-film modes, interface
scattering matrices and Redheffer star-product cascade with energy fluxes.

TEST INFRASTRUCTURE ONLY (used by tests/, smoke() and bench.py's
cpu_baseline leg as the checker; the product package never imports it).

Restates:
  Anallo.calc_kz / Z / specular_incidence   mode_calcs.py:110-250
  calc_2d_grating_orders                    mode_calcs.py:61-74
  r_t_mat_anallo                            mode_calcs.py:451-477
  r_t_mat_tf_ns                             mode_calcs.py:480-515
  Stack.calc_scat (cascade + fluxes)        stack.py:162-427 (shears=None)

Layers are plain dicts: thin film {'kind':'tf','n':complex,'h':h_norm|None}
and structured {'kind':'ns','kz':beta[nval],'J':(2P x nval),'Jd':(nval x 2P),
'h':h_norm}.  The list is ordered bottom (substrate) to top (superstrate).
"""
import numpy as np


def grating_orders_2d(max_order, d=1.0, dy=1.0):
    """mode_calcs.py:61-74: orders with (px/d)^2+(py/dy)^2 <= (p/max(d,dy))^2,
    px outer, py inner."""
    pxs = np.arange(-max_order, max_order + 1)
    pys_mesh, pxs_mesh = np.meshgrid(pxs, pxs)
    low = (pxs_mesh / d) ** 2 + (pys_mesh / dy) ** 2 <= (max_order / max(d, dy)) ** 2
    return pxs_mesh[low], pys_mesh[low]


def grating_orders(max_order, d_ratio=1.0, world_1d=False):
    """1D orders (mode_calcs.py:55-59, no y orders) or the 2D set."""
    if world_1d:
        pxs = np.arange(-max_order, max_order + 1)
        return pxs, np.zeros_like(pxs)
    return grating_orders_2d(max_order, 1.0, d_ratio)


def air_sort(k_pll_norm, wl_norm, max_order, d_ratio=1.0, world_1d=False):
    """Sort order of the air reference (mode_calcs.py:164-170)."""
    pxs, pys = grating_orders(max_order, d_ratio, world_1d)
    a0, b0 = k_pll_norm
    al = a0 + pxs * 2 * np.pi
    be = b0 + pys * 2 * np.pi / d_ratio
    k = np.complex128(2 * np.pi * 1.0 / wl_norm)
    kz = np.sqrt(k ** 2 - al ** 2 - be ** 2)
    if np.count_nonzero(kz) != kz.size:
        a0, b0 = (1 - 1e-9) * np.asarray(k_pll_norm)
        al = a0 + pxs * 2 * np.pi
        be = b0 + pys * 2 * np.pi / d_ratio
        kz = np.sqrt(k ** 2 - al ** 2 - be ** 2)
    s = np.argsort(-1 * kz.real + kz.imag)
    return s, pxs, pys


def film_kz(n, k_pll_norm, wl_norm, max_order, sort_order, d_ratio=1.0, world_1d=False):
    """Anallo.calc_kz (mode_calcs.py:116-189) with the air-reference sort."""
    pxs, pys = grating_orders(max_order, d_ratio, world_1d)
    a0, b0 = k_pll_norm
    al = a0 + pxs * 2 * np.pi
    be = b0 + pys * 2 * np.pi / d_ratio
    k = np.complex128(2 * np.pi * n / wl_norm)
    kz = np.sqrt(k ** 2 - al ** 2 - be ** 2)
    if np.count_nonzero(kz) != kz.size:
        a0, b0 = (1 - 1e-9) * np.asarray(k_pll_norm)
        al = a0 + pxs * 2 * np.pi
        be = b0 + pys * 2 * np.pi / d_ratio
        kz = np.sqrt(k ** 2 - al ** 2 - be ** 2)
    s = sort_order
    spec = np.nonzero((pxs[s] == 0) * (pys[s] == 0))[0][0]
    nprop = int((kz.imag == 0).sum())
    return kz[s], spec, nprop, k


def film_Z(n, kz, k):
    """Anallo.Z (mode_calcs.py:198-218)."""
    Zcr = 1.0 / n
    kk = k / kz
    return np.concatenate((Zcr * kk, Zcr / kk))


def r_t_anallo(Z1, Z2):
    """r_t_mat_anallo (mode_calcs.py:451-477)."""
    R12 = np.diag((Z2 - Z1) / (Z2 + Z1))
    T12 = np.diag(2. * np.sqrt(Z2) * np.sqrt(Z1) / (Z2 + Z1))
    return R12, T12, -R12, T12


def r_t_tf_ns(Z1, J, Jd):
    """r_t_mat_tf_ns (mode_calcs.py:480-515); J already area-normalised."""
    zs = np.sqrt(1 / Z1)
    A = zs[:, None] * J
    B = Jd * zs[None, :]
    nv = B.shape[0]
    D = np.eye(nv) + B @ A
    DB = np.linalg.solve(D, B)
    R12 = -np.eye(A.shape[0]) + 2 * A @ DB
    T12 = 2 * DB
    R21 = np.linalg.solve(D, np.eye(nv) - B @ A)
    T21 = 2 * A @ np.linalg.inv(D)
    return R12, T12, R21, T21


def calc_scat(layers, wl_norm, k_pll_norm, max_order, pol='TE', d_ratio=1.0, world_1d=False):
    """Stack.calc_scat (stack.py:162-427) for shears=None, calc_fluxes=True.
    world_1d: thin films expand in the 1D grating orders (1D stacks).
    Returns dict with R_net, T_net, a_list, r_list, t_list."""
    s, _, _ = air_sort(k_pll_norm, wl_norm, max_order, d_ratio, world_1d)
    kz_air, spec_air, nprop_air, k_air = film_kz(1.0, k_pll_norm, wl_norm, max_order, s, d_ratio, world_1d)
    Zair = film_Z(1.0, kz_air, k_air)
    r12, r21, t12, t21, kzs = [], [], [], [], []
    nprop = []
    spec = []
    for lay in layers:
        if lay['kind'] == 'tf':
            kz, sp_, npw, k = film_kz(lay['n'], k_pll_norm, wl_norm, max_order, s, d_ratio, world_1d)
            R12, T12, R21, T21 = r_t_anallo(Zair, film_Z(lay['n'], kz, k))
            kzs.append(np.append(kz, kz))
            nprop.append(npw)
            spec.append(sp_)
        else:
            R12, T12, R21, T21 = r_t_tf_ns(Zair, lay['J'], lay['Jd'])
            kzs.append(lay['kz'])
            nprop.append(None)
            spec.append(None)
        r12.append(R12)
        t12.append(T12)
        r21.append(R21)
        t21.append(T21)
    PWp = r12[0].shape[0]
    neq_PW = PWp // 2
    I = np.eye(PWp, dtype=complex)
    tnet, rnet = t12[0], r12[0]
    tnet_list, rnet_list = [tnet], [rnet]
    inv_t21_list, inv_t12_list, P_list = [], [], []
    for i in range(1, len(layers) - 1):
        it21 = np.linalg.solve(I - r12[i] @ rnet, t21[i])
        tnet = tnet @ it21
        rnet = r21[i] + t12[i] @ rnet @ it21
        inv_t21_list.append(it21)
        tnet_list.append(tnet)
        rnet_list.append(rnet)
        P = np.diag(np.exp(1j * kzs[i] * layers[i]['h']))
        ITF = np.eye(len(P), dtype=complex)
        it12 = np.linalg.solve(ITF - r21[i] @ P @ rnet @ P, t12[i])
        Pit12 = P @ it12
        tnet = tnet @ Pit12
        rnet = r12[i] + t21[i] @ P @ rnet @ Pit12
        P_list.append(P)
        inv_t12_list.append(it12)
        tnet_list.append(tnet)
        rnet_list.append(rnet)
    it21 = np.linalg.solve(I - r12[-1] @ rnet, t21[-1])
    tnet = tnet @ it21
    rnet = r21[-1] + t12[-1] @ rnet @ it21
    inv_t21_list.append(it21)
    tnet_list.append(tnet)
    rnet_list.append(rnet)
    out = dict(R_net=rnet, T_net=tnet)

    num_prop_air = nprop_air
    num_prop_in = nprop[-1]
    U = np.zeros((2 * PWp, 2 * PWp), complex)
    for i in range(num_prop_air):
        U[i, i] = 1.0
        U[neq_PW + i, neq_PW + i] = 1.0
        U[PWp + i, PWp + i] = -1.0
        U[PWp + neq_PW + i, PWp + neq_PW + i] = -1.0
    for i in range(num_prop_air, neq_PW):
        U[i, PWp + i] = -1.0j
        U[neq_PW + i, PWp + neq_PW + i] = 1.0j
        U[PWp + i, i] = 1.0j
        U[PWp + neq_PW + i, neq_PW + i] = -1.0j
    d_minus = np.zeros((PWp, 1), complex)
    sTE, sTM = spec[-1], spec[-1] + neq_PW
    if pol == 'TE':
        d_minus[sTE] = 1
    elif pol == 'TM':
        d_minus[sTM] = 1
    elif pol == 'un':
        d_minus[sTE] = d_minus[sTM] = 1 / np.sqrt(2.)
    elif pol == 'R Circ':
        d_minus[sTE], d_minus[sTM] = 1 / np.sqrt(2.), 1j / np.sqrt(2.)
    elif pol == 'L Circ':
        d_minus[sTE], d_minus[sTM] = 1 / np.sqrt(2.), -1j / np.sqrt(2.)
    down, up = [], []
    nrm = lambda v: np.linalg.norm(v) ** 2
    down.append(nrm(d_minus[:num_prop_in]) + nrm(d_minus[neq_PW:neq_PW + num_prop_in]))
    d_plus = rnet_list[-1] @ d_minus
    up.append(nrm(d_plus[:num_prop_in]) + nrm(d_plus[neq_PW:neq_PW + num_prop_in]))
    f1m = inv_t21_list[-1] @ d_minus
    for i in range(len(layers) - 2):
        f1p = rnet_list[-2 * i - 2] @ f1m
        f = np.concatenate((f1m, f1p))
        down.append((f.conj().T @ U @ f)[0, 0])
        f2m = inv_t12_list[-i - 1] @ f1m
        f1m = inv_t21_list[-i - 2] @ P_list[-i - 1] @ f2m
    f2m = tnet_list[0] @ f1m
    a_list, r_list, t_list = [], [], []
    num_prop_out = nprop[0]
    if num_prop_out is None or num_prop_out != 0:
        if num_prop_out is not None:
            down.append(nrm(f2m[:num_prop_out]) + nrm(f2m[neq_PW:neq_PW + num_prop_out]))
        for i in range(1, len(down) - 1):
            a_list.append(abs(abs(down[i]) - abs(down[i + 1])))
        a_list.append(abs(down[0] - down[-1] - up[0]))
        for i in range(1, len(up) - 1):
            r_list.append(abs(up[i]) / abs(down[i]))
        r_list.append(abs(up[0] / down[0]))
        for i in range(0, len(down) - 2):
            t_list.append(abs(abs(down[i + 2]) / abs(down[i])))
        t_list.append(abs(down[-1] / down[0]))
    out.update(a_list=a_list, r_list=r_list, t_list=t_list, down=down, up=up)
    return out
