"""CPU restatement of EMUstack's 1D-grating FEM mode path.

TEST INFRASTRUCTURE ONLY: this module is the checker for the GPU path.  Only
tests/, __graft_entry__.smoke() and the CPU-baseline leg of bench.py may
import it; the product never does.

Follows (reference paths relative to the reference repository):
  grating_mesh_1d  backend/objects.py:1054-1530 (uniform P2 mesh, material
                   type by element position; one or two inclusions).  The
                   centre inclusion's rule has two forms: 'code' is the
                   current objects.py:1502-1503 test (elements starting up
                   to one element width left of 0.5 - r count as inclusion),
                   'exact' takes the elements inside [0.5 - r, 0.5 + r]; the
                   reference's case-4 golden R/T files were generated with
                   the 'exact' form (reproduced to 1e-10 in tests, the
                   'code' form is 2e-2 away)
  dof_1d           backend/fortran/periodic_cond_1d.f:1-220
  element_mats_1d  backend/fortran/matrix_{mxx,kxy,kyx,kyy,myy}_1d.f
                   (integrals of Lagrange P2/P3 bases, computed here by
                   Gauss-Legendre quadrature; pinned against the reference's
                   rational tables in tests)
  assemble_1d      backend/fortran/asmbly_1d.f:1-359 (dense K - shift M, M)
  eigensolve_1d    backend/fortran/valpr_lapack_1d.f:43-150 (dense LU,
                   shift-invert Arnoldi for the nval largest |mu|)
  fields_1d        backend/fortran/array_sol_1d.f:1-301
  sol_p2_1d        backend/fortran/array_sol_P2_1d.f:13-61
  overlap_1d       backend/fortran/orthogonal_1d.f:1-252
  normalise_1d     backend/fortran/normalisation_1d.f:1-34
  pw_order_inv_1d  backend/fortran/pw_ordering_1d.f:1-80
  j_overlaps_1d    backend/fortran/J_overlap_1d.f, J_dagger_overlap_1d.f,
                   with the exponential integrals of vector_p{2,3}_exp_1d.f
                   computed by quadrature
  pw_1d_to_2d      backend/fortran/pw_matrix_1d_to_2d.f:1-200
  calc_modes_1d    backend/fortran/py_calc_modes_1d.f:1-707
"""
import numpy as np
import scipy.linalg as sla
import scipy.sparse.linalg as spla

from .fem2d import beta_from_mu, biorth_clusters, nr_indexx, pair_swap

# ------------------------------------------------------------------ mesh


def _inside(x1, x2, lo, hi):
    return lo <= x1 <= hi and lo <= x2 <= hi


def _element_type(x1, x2, dx, N, r, d, small_space, edge_spacing, split):
    """Material type of the element [x1, x2] (objects.py:1054-1506): the
    reference's if-chain for N inclusions, r[k] the normalised half-widths
    (r[0] unused), d the diameters in nm."""
    if x1 <= 0.5 + r[1] and 0.5 - r[1] - dx <= x1 and x2 <= 0.5 + r[1] and 0.5 - r[1] - dx <= x2:
        return 2
    if N == 1:
        return 1
    if N == 2 or (split and not edge_spacing):
        if not split:  # objects.py:1472-1489
            if small_space is None:
                sm = ld = 0.5 - r[1] - r[2]
            else:
                sm, ld = small_space, 1.0 - small_space - (2 * r[1]) - (2 * r[2])
            if x1 <= 0.5 - r[1] - sm and x2 <= 0.5 - r[1] - sm:
                return 3
            if x1 >= 0.5 + ld + r[1] and x2 >= 0.5 + ld + r[1]:
                return 3
            return 1
        # touching inclusions (objects.py:1128-1166, 1234-1264, 1327-1352,
        # 1408-1425, 1456-1471); the element-width slack only for N <= 3
        sl = dx if N <= 3 else 0.0
        a, b = 0.5 + r[1], 0.5 - r[1]
        for k in range(2, N + 1):
            if x1 <= a + r[k] and a - sl <= x1 and x2 <= a + r[k] and a - sl <= x2:
                return k + 1
            if x1 <= b and b - r[k] - sl <= x1 and x2 <= b and b - r[k] - sl <= x2:
                return k + 1
            a, b = a + r[k], b - r[k]
        return 1
    if edge_spacing:
        if N == 3:  # objects.py:1387-1407: the gap from diameters in nm
            i_d = (1.0 - d[1] - d[2] - d[3]) / 3.0
            if _inside(x1, x2, 0.5 - i_d - 2.0 * r[2] - r[1], 0.5 - i_d - r[1]):
                return 3
            if _inside(x1, x2, 0.5 + i_d + r[1], 0.5 + i_d + 2.0 * r[3] + r[1]):
                return 4
            return 1
        if N == 4:  # objects.py:1304-1326
            i_d = 2.0 * (0.5 - r[1] - r[2] - r[3] - r[4]) / 4.0
            if _inside(x1, x2, 0.5 - i_d - r[1] - 2.0 * r[2], 0.5 - i_d - r[1]):
                return 3
            if _inside(x1, x2, 0.5 + i_d + r[1], 0.5 + i_d + r[1] + 2.0 * r[3]):
                return 4
            if x1 >= 0.5 + 2.0 * i_d + r[1] + 2.0 * r[3] and x2 >= 0.5 + 2.0 * i_d + r[1] + 2.0 * r[3]:
                return 4
            if x1 <= 0.5 - 2.0 * i_d - r[1] - 2.0 * r[2] and x2 <= 0.5 - 2.0 * i_d - r[1] - 2.0 * r[2]:
                return 3
            return 1
        # N = 5 (objects.py:1210-1233) and N = 6 (objects.py:1098-1127)
        if N == 5:
            i_d = 2.0 * (0.5 - r[1] - r[2] - r[3] - r[4] - r[5]) / 5.0
        else:
            i_d = 2.0 * (0.5 - r[1] - r[2] - r[3] - r[4] - r[5] - r[6]) / 6.0
        if _inside(x1, x2, 0.5 - i_d - r[1] - 2.0 * r[2], 0.5 - i_d - r[1]):
            return 3
        if _inside(x1, x2, 0.5 + i_d + r[1], 0.5 + i_d + 2.0 * r[3] + r[1]):
            return 4
        if _inside(x1, x2, 0.5 - 2.0 * i_d - 2.0 * r[2] - 2.0 * r[4] - r[1], 0.5 - 2.0 * i_d - 2.0 * r[2] - r[1]):
            return 5
        if _inside(x1, x2, 0.5 + 2.0 * i_d + 2.0 * r[3] + r[1], 0.5 + 2.0 * i_d + 2.0 * r[3] + 2.0 * r[5] + r[1]):
            return 4 if N == 5 else 6
        if N == 6:
            top = 0.5 - 3.0 * i_d - r[1] - 2.0 * r[2] - 2.0 * r[4]
            if _inside(x1, x2, top - r[5], top):
                return 6
            bot = 0.5 + 3.0 * i_d + r[1] + 2.0 * r[3] + 2.0 * r[5]
            if _inside(x1, x2, bot, bot + 2.0 * r[6]):
                return 7
        return 1
    # evenly spaced centres 0.5 -/+ i_d, 0.5 -/+ 2 i_d (objects.py:1167-1193,
    # 1265-1285, 1353-1372, 1426-1444)
    i_d = 1.0 / N
    c = {2: 0.5 - i_d, 3: 0.5 + i_d, 4: 0.5 - 2.0 * i_d, 5: 0.5 + 2.0 * i_d}
    for k in range(2, N + 1):
        if k == N and N in (4, 6):  # the last inclusion across the cell edges
            if (x1 <= r[k] and x2 <= r[k]) or (x1 >= 1.0 - r[k] and x2 >= 1.0 - r[k]):
                return k + 1
        elif _inside(x1, x2, c[k] - r[k], c[k] + r[k]):
            return k + 1
    return 1


def grating_mesh_1d(period, diameters, lc_bkg, small_space=None, split_touching_incs=False,
                    inclusion_edge='code', edge_spacing=False):
    """Uniform mesh of the unit cell [0, 1] with nel = round(1/lc_bkg) P2
    elements and the element material types of objects.py:1054-1506 for up
    to six inclusions (N = index of the last positive diameter).  Returns x
    (npt,), table_nod (nel, 3) 1-based [left, right, mid], type_el (nel,)
    1-based.  inclusion_edge 'exact' drops the element-width slack of the
    reference's rules (the rule of its tests/ref/case_4 golden files)."""
    d = [0.0] + [float(v) for v in diameters] + [0.0] * (6 - len(diameters))
    pos = [k for k in range(1, 7) if d[k] > 0]
    if not pos:
        raise ValueError('Must have at least one grating of nonzero width.')
    N = pos[-1]
    if inclusion_edge not in ('code', 'exact'):
        raise ValueError(inclusion_edge)
    nel = int(np.round(1.0 / lc_bkg))
    npt = 2 * nel + 1
    delta_x = (1.0 - 0.0) / nel
    ls_x = np.zeros(npt + 1)
    for i_el in range(1, nel + 1):
        x = 0.0 + (i_el - 1) * delta_x
        ls_x[2 * i_el - 1] = x
        ls_x[2 * i_el] = x + delta_x / 2.0
    ls_x[2 * nel + 1] = 0.0 + nel * delta_x
    table = np.array([[2 * i - 1, 2 * i + 1, 2 * i] for i in range(1, nel + 1)], dtype=np.int64)
    r = [0.0] + [d[k] / (2.0 * period) for k in range(1, 7)]
    type_el = np.ones(nel, dtype=np.int64)
    for e in range(nel):
        x1 = ls_x[2 * (e + 1) - 1]
        x2 = ls_x[2 * (e + 1) + 1]
        dx = x2 - x1 if inclusion_edge == 'code' else 0.0
        type_el[e] = _element_type(x1, x2, dx, N, r, d, small_space, edge_spacing, split_touching_incs)
    return ls_x[1:], table, type_el


class Mesh1D:
    pass


def dof_1d(x, table_nod):
    """Degrees of freedom of periodic_cond_1d.f: per element 11 local DOFs
    [P2 x: left, right, mid | P3 y: left, right, 1/3, 2/3 | P3 z: same],
    9 new equations per element (the right-end P3 nodes belong to the next
    element; the last element's wrap to the first).  Returns a Mesh1D with
    eq (nel, 11) 0-based equation numbers, xd (nel, 11) DOF coordinates,
    wrap (nel, 11) True for the DOFs that carry the Bloch phase, and the
    period measured as in asmbly_1d.f:108-122."""
    nel = table_nod.shape[0]
    m = Mesh1D()
    m.nel = nel
    m.neq = 9 * nel
    m.x = np.asarray(x, float)
    m.table_nod = table_nod
    eq = np.zeros((nel, 11), dtype=np.int64)
    xd = np.zeros((nel, 11))
    wrap = np.zeros((nel, 11), dtype=bool)
    for e in range(nel):
        xmin = m.x[table_nod[e, 0] - 1]
        xmax = m.x[table_nod[e, 1] - 1]
        dx = xmax - xmin
        b = 9 * e
        nb = 9 * (e + 1) if e + 1 < nel else 0
        eq[e] = [b, b + 1, b + 2, b + 3, nb + 3, b + 4, b + 5, b + 6, nb + 6, b + 7, b + 8]
        xd[e] = [xmin, xmax, xmin + dx / 2.0, xmin, xmax, xmin + dx / 3.0, xmin + 2.0 * dx / 3.0,
                 xmin, xmax, xmin + dx / 3.0, xmin + 2.0 * dx / 3.0]
    wrap[nel - 1, 4] = wrap[nel - 1, 8] = True
    m.eq, m.xd, m.wrap = eq, xd, wrap
    m.period_x = xd[nel - 1, 4] - xd[0, 3]
    return m


def build_mesh_1d(x, table_nod, type_el):
    m = dof_1d(x, table_nod)
    m.type_el = np.asarray(type_el, dtype=np.int64)
    return m


# ------------------------------------------------------- element matrices
_P2_NODES = np.array([0.0, 1.0, 0.5])
_P3_NODES = np.array([0.0, 1.0, 1.0 / 3.0, 2.0 / 3.0])


def lagrange(nodes, t):
    """Values and derivatives (d/dt) of the Lagrange basis on `nodes` at t."""
    t = np.atleast_1d(np.asarray(t, float))
    n = len(nodes)
    val = np.ones((n, t.size))
    der = np.zeros((n, t.size))
    for i in range(n):
        others = [nodes[j] for j in range(n) if j != i]
        den = np.prod([nodes[i] - o for o in others])
        val[i] = np.prod([t - o for o in others], axis=0) / den
        for k in range(len(others)):
            rest = [others[j] for j in range(len(others)) if j != k]
            der[i] += np.prod([t - o for o in rest], axis=0) / den if rest else 1.0 / den
    return val, der


def gauss01(nq):
    z, w = np.polynomial.legendre.leggauss(nq)
    return 0.5 * (z + 1.0), 0.5 * w


def element_mats_1d(h):
    """Mxx = int P2 P2, Kxy = int P2 (P3)', Kyx = Kxy^T, Kyy = int P3' P3',
    Myy = int P3 P3 on an element of length h (x = xmin + h t)."""
    t, w = gauss01(8)
    p2, _ = lagrange(_P2_NODES, t)
    p3, d3 = lagrange(_P3_NODES, t)
    Mxx = h * (p2 * w) @ p2.T
    Kxy = (p2 * w) @ d3.T
    Kyy = (d3 * w) @ d3.T / h
    Myy = h * (p3 * w) @ p3.T
    return Mxx, Kxy, Kxy.T.copy(), Kyy, Myy


def element_blocks_1d(h, pp, qq, ky):
    """11 x 11 element blocks (K, M) of asmbly_1d.f:139-258, rows = trial."""
    Mxx, Kxy, Kyx, Kyy, Myy = element_mats_1d(h)
    K = np.zeros((11, 11), complex)
    M = np.zeros((11, 11), complex)
    X, Y, Z = slice(0, 3), slice(3, 7), slice(7, 11)
    K[X, X] = (pp * ky ** 2 - qq) * Mxx
    M[X, X] = -pp * Mxx
    K[Y, X] = -1j * ky * pp * Kyx
    K[Z, X] = -pp * Kyx
    K[X, Y] = 1j * ky * pp * Kxy
    K[Y, Y] = pp * Kyy - qq * Myy
    M[Y, Y] = -pp * Myy
    K[Z, Y] = 1j * ky * pp * Myy
    M[X, Z] = pp * Kxy
    M[Y, Z] = 1j * ky * pp * Myy
    K[Z, Z] = pp * Kyy + (pp * ky ** 2 - qq) * Myy
    return K, M


def phase_1d(m, kx):
    """val_exp of asmbly_1d.f:123-137 / array_sol_1d.f:146-153 (nel, 11)."""
    ve = np.ones((m.nel, 11), complex)
    dxw = m.xd[m.nel - 1, 4] - m.xd[0, 3]
    ve[m.wrap] = np.exp(1j * kx * dxw)
    return ve


def assemble_1d(m, pp, qq, shift, kx, ky):
    """Dense matrix_1 = K - shift M and matrix_2 = M (asmbly_1d.f:262-283)."""
    n = m.neq
    A = np.zeros((n, n), complex)
    B = np.zeros((n, n), complex)
    ve = phase_1d(m, kx)
    for e in range(m.nel):
        h = m.x[m.table_nod[e, 1] - 1] - m.x[m.table_nod[e, 0] - 1]
        t = m.type_el[e] - 1
        K, M = element_blocks_1d(h, pp[t], qq[t], ky)
        ph = np.conj(ve[e])[:, None] * ve[e][None, :]
        K = K * ph
        M = M * ph
        idx = m.eq[e]
        np.add.at(A, (idx[:, None], idx[None, :]), K - shift * M)
        np.add.at(B, (idx[:, None], idx[None, :]), M)
    return A, B


# ------------------------------------------------------------ eigensolve


def eigensolve_1d(A, B, nval, ncv):
    """nval largest-|mu| eigenpairs of OP = A^{-1} B (valpr_lapack_1d.f)."""
    lu = sla.lu_factor(A)
    n = A.shape[0]
    op = spla.LinearOperator((n, n), matvec=lambda v: sla.lu_solve(lu, B @ v), dtype=complex)
    mu, V = spla.eigs(op, k=nval, ncv=min(ncv, n - 1), which='LM', tol=0, maxiter=30 * n)
    return mu, V


# ------------------------------------------------------- post-processing


def fields_1d(m, V, order, kx, n_k):
    """array_sol_1d.f: element DOF values (nel, 11, nval) of the sorted
    modes, Bloch phase on the wrapped DOFs, normalised by the largest value
    near the cell centre, else on the right (prime, n_k = 1) or left
    (adjoint, n_k = 2) half."""
    ve = phase_1d(m, kx)
    sol = V[m.eq][:, :, order] * ve[:, :, None]
    xs = m.xd.ravel()
    xmin, xmax = xs.min(), xs.max()
    xmid = (xmin + xmax) / 2.0
    dx = (xmax - xmin) / m.nel
    flat = sol.reshape(m.nel * 11, -1)
    ab = np.abs(flat)

    def first_max(mask):
        a = np.where(mask[:, None], ab, -1.0)
        i = np.argmax(a, axis=0)
        v = flat[i, np.arange(flat.shape[1])]
        return np.where(a[i, np.arange(flat.shape[1])] > 0, v, 0.0)
    zmax = first_max(np.ones(len(xs), bool))
    zl = first_max(xs < xmid)
    zr = first_max(xs > xmid)
    zm = first_max(np.abs(xs - xmid) < dx / 4.0)
    with np.errstate(divide='ignore', invalid='ignore'):
        use_mid = np.abs(zm / zmax) >= 0.95
    den = np.where(use_mid, zm, zr if n_k == 1 else zl)
    return sol / den[None, None, :]


def sol_p2_1d(sol):
    """sol_P2 (3, 3, nval, nel) of array_sol_P2_1d.f:13-61 from the element
    DOF values sol (nel, 11, nval): E_x at the three P2 nodes (DOFs 0-2); E_y
    and E_z (DOFs 3-6 and 7-10, P3 nodes at the two ends, then 1/3 and 2/3)
    at the ends, and at the mid-node the P3 interpolant, whose nodal weights
    at x = 1/2 are (-1, -1, 9, 9) / 16."""
    w = np.array([-1.0, -1.0, 9.0, 9.0]) / 16.0
    nel, _, nval = sol.shape
    out = np.zeros((3, 3, nval, nel), complex)
    out[0] = np.transpose(sol[:, 0:3, :], (1, 2, 0))
    for c, base in ((1, 3), (2, 7)):
        out[c, 0] = sol[:, base].T
        out[c, 1] = sol[:, base + 1].T
        out[c, 2] = np.einsum('j,ejn->ne', w, sol[:, base:base + 4])
    return out


def overlap_1d(m, pp, ky, beta1, sol1, sol2, k0):
    """mat_overlap(i, j) = sum_e sol2_e(:, i) . M2_e (beta1_j sol1_e(:, j)) / k0
    (orthogonal_1d.f:108-189)."""
    nval = sol1.shape[2]
    L = np.zeros((nval, nval), complex)
    X, Y, Z = slice(0, 3), slice(3, 7), slice(7, 11)
    for e in range(m.nel):
        h = m.x[m.table_nod[e, 1] - 1] - m.x[m.table_nod[e, 0] - 1]
        Mxx, Kxy, Kyx, Kyy, Myy = element_mats_1d(h)
        p = pp[m.type_el[e] - 1]
        M2 = np.zeros((11, 11), complex)
        M2[X, X] = p * Mxx
        M2[Y, Y] = p * Myy
        M2[X, Z] = -p * Kxy
        M2[Y, Z] = -1j * ky * p * Myy
        L += sol2[e].T @ (M2 @ (sol1[e] * beta1[None, :])) / k0
    return L


def normalise_1d(L, sol1, sol2):
    """normalisation_1d.f: both families scaled by 1/sqrt(L_ii)."""
    z = np.sqrt(np.diag(L))
    f = np.where(np.abs(z) > 1e-8, 1.0 / np.where(z == 0, 1, z), 1.0)
    return sol1 * f[None, None, :], sol2 * f[None, None, :]


def pw_order_inv_1d(ordre, period_x, kx, ky, k0):
    """index_pw_inv of pw_ordering_1d.f (0-based positions) and alphas."""
    px = np.arange(-ordre, ordre + 1)
    al = kx + 2.0 * np.pi / period_x * px
    kz = np.sqrt((k0 ** 2 - al ** 2 - ky ** 2).astype(complex))
    idx = nr_indexx(-(kz ** 2).real)
    inv = np.empty(len(px), dtype=np.int64)
    inv[idx] = np.arange(len(px))
    return inv, al


def exp_integrals_1d(xmin, xmax, alpha):
    """int_{xmin}^{xmax} phi_i(x) exp(i alpha x) dx for the P2 and P3 bases
    (the quantities of vector_p2_exp_1d.f / vector_p3_exp_1d.f)."""
    t, w = gauss01(16)
    h = xmax - xmin
    e = np.exp(1j * alpha * (xmin + h * t)) * w * h
    p2, _ = lagrange(_P2_NODES, t)
    p3, _ = lagrange(_P3_NODES, t)
    return p2 @ e, p3 @ e


def j_overlaps_1d(m, sol1, sol2, ordre, kx, ky, k0):
    """J (2P x nval) from sol1 with exp(-i alpha x) and J^dagger (nval x 2P)
    from sol2 with exp(+i alpha x); rows E-type then K-type, in PW order."""
    inv, al = pw_order_inv_1d(ordre, m.period_x, kx, ky, k0)
    P = len(al)
    nval = sol1.shape[2]
    JE = np.zeros((P, nval), complex)
    JK = np.zeros((P, nval), complex)
    DE = np.zeros((P, nval), complex)
    DK = np.zeros((P, nval), complex)
    for s in range(P):
        a = al[s]
        if a != 0.0 or ky != 0.0:
            nrm = 1.0 / np.sqrt(a ** 2 + ky ** 2)
            Kv = (a * nrm, ky * nrm)
            Ev = (ky * nrm, -a * nrm)
        else:
            Kv, Ev = (0.0, 1.0), (1.0, 0.0)
        s2 = inv[s]
        for e in range(m.nel):
            xmin = m.x[m.table_nod[e, 0] - 1]
            xmax = m.x[m.table_nod[e, 1] - 1]
            for sgn, sol, (TE, TK) in ((-1.0, sol1, (JE, JK)), (1.0, sol2, (DE, DK))):
                v2, v3 = exp_integrals_1d(xmin, xmax, sgn * a)
                fx = v2 @ sol[e, 0:3, :]
                fy = v3 @ sol[e, 3:7, :]
                TK[s2] += fx * Kv[0] + fy * Kv[1]
                TE[s2] += fx * Ev[0] + fy * Ev[1]
    J = np.vstack([JE, JK])
    Jd = np.vstack([DE, DK]).T
    return J, Jd


def pw_1d_to_2d(J, Jd, ordre, period_x, kx, ky, k0):
    """Embed the 1D plane-wave basis into the 2D one (pw_matrix_1d_to_2d.f)."""
    P1 = 2 * ordre + 1
    inv1, _ = pw_order_inv_1d(ordre, period_x, kx, ky, k0)
    kz2, pos_py0 = [], {}
    s = 0
    for px in range(-ordre, ordre + 1):
        for py in range(-ordre, ordre + 1):
            if (px / period_x) ** 2 + (py / period_x) ** 2 <= (ordre / max(period_x, period_x)) ** 2:
                a = kx + 2 * np.pi / period_x * px
                b = ky + 2 * np.pi / period_x * py
                kz2.append(k0 ** 2 - a ** 2 - b ** 2)
                if py == 0:
                    pos_py0[px + ordre] = s
                s += 1
    P2 = s
    idx = nr_indexx(-np.real(np.array(kz2)))
    inv2 = np.empty(P2, dtype=np.int64)
    inv2[idx] = np.arange(P2)
    nval = J.shape[1]
    J2 = np.zeros((2 * P2, nval), complex)
    Jd2 = np.zeros((nval, 2 * P2), complex)
    for ipw in range(P1):
        s1 = inv1[ipw]
        s2 = inv2[pos_py0[ipw]]
        J2[s2] = J[s1]
        J2[s2 + P2] = J[s1 + P1]
        Jd2[:, s2] = Jd[:, s1]
        Jd2[:, s2 + P2] = Jd[:, s1 + P1]
    return J2, Jd2


# ----------------------------------------------------------------- driver


def calc_modes_1d(wl_norm, nval, ordre, m, n_eff, kx, ky, shift, world_1d=True, biorth='diag',
                  return_sol=False):
    """beta (nval,), J (2P x nval), J^dagger (nval x 2P) as returned by
    calc_modes_1d (py_calc_modes_1d.f); P = 2 ordre + 1, or the 2D count
    when world_1d is False.  return_sol: also return sol_P2 of the prime
    modes after normalisation_1d (py_calc_modes_1d.f:547-551) as the last
    element of the tuple."""
    n_eff = np.asarray(n_eff, complex)
    eps = n_eff ** 2
    if np.all(eps[:-1] == eps[1:]):
        raise ValueError('FEM routine cannot handle homogeneous layers.')
    k0 = 2.0 * np.pi / wl_norm
    pp = np.ones(len(eps), complex)
    qq = eps * k0 ** 2
    nvect = 2 * nval + nval // 2 + 3
    out = []
    for n_k, (bx, by) in ((1, (kx, ky)), (2, (-kx, -ky))):
        A, B = assemble_1d(m, pp, qq, shift, bx, by)
        mu, V = eigensolve_1d(A, B, nval, nvect)
        beta = beta_from_mu(mu, shift)
        order = np.array(nr_indexx(-(beta ** 2).real))
        beta = beta[order]
        sol = fields_1d(m, V, order, bx, n_k)
        out.append((beta, sol))
    (b1, s1), (b2, s2) = out
    L = overlap_1d(m, pp, ky, b1, s1, s2, k0)
    for j, j2 in pair_swap(L):
        b2[[j, j2]] = b2[[j2, j]]
        s2[:, :, [j, j2]] = s2[:, :, [j2, j]]
    L = overlap_1d(m, pp, ky, b1, s1, s2, k0)
    s1, s2 = normalise_1d(L, s1, s2)
    J, Jd = j_overlaps_1d(m, s1, s2, ordre, kx, ky, k0)
    Jdc = None
    if biorth in ('cluster', 'both'):
        # after the reference's 1/sqrt(L_ii) scaling of both families the
        # overlap is L_ij / sqrt(L_ii L_jj); invert it within each group of
        # coupled modes (identity for isolated modes)
        z = np.sqrt(np.diag(L))
        ok = np.abs(z) > 1e-8
        Ln = L / np.outer(np.where(ok, z, 1), np.where(ok, z, 1))
        N = np.eye(len(b1), dtype=complex)
        for g in biorth_clusters(L):
            if g.size > 1 and np.all(ok[g]):
                N[np.ix_(g, g)] = np.linalg.inv(Ln[np.ix_(g, g)])
        Jdc = N @ Jd
    if not world_1d:
        J2, Jd = pw_1d_to_2d(J, Jd, ordre, m.period_x, kx, ky, k0)
        if Jdc is not None:
            Jdc = pw_1d_to_2d(J, Jdc, ordre, m.period_x, kx, ky, k0)[1]
        J = J2
    out = (b1, J, Jd, Jdc) if biorth == 'both' else (b1, J, (Jdc if biorth == 'cluster' else Jd))
    return out + (sol_p2_1d(s1),) if return_sol else out
