"""CPU oracle: FP64 numpy restatement of EMUstack's 2D FEM Bloch-mode path.

TEST INFRASTRUCTURE ONLY.  Nothing in the product package imports this file;
only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it,
and only as the checker / CPU baseline.

It restates, step by step, the f2py Fortran driver ``calc_modes_2d``
(backend/fortran/py_calc_modes.f:1-958) of the reference:

  read .mail            geometry.f:27-61
  lattice vectors       lattice_vec.f:4-91
  topology / DOFs       list_face.f, list_edge.f:2-77, list_node_P3.f:2-134,
                        type_node_edge_face.f:4-79, get_coord_p3.f:13-140
  periodic merge        periodic_node.f:4-184, periodic_N_E_F.f:4-131,
                        periodic_cond.f:7-178
  sparse pattern        csr_length.f:5-178, sort_csr.f:5-53
  assembly              asmbly.f:3-341 with basis_ls.f, basis_vec.f,
                        phi{1,2,3}_2d_mat.f, jacobian_p1_2d.f, quad_triangle.f
  eigen-solve           valpr_umf_64.f (OP = (K - shift M)^-1 M, ARPACK 'LM')
  beta branch / sort    py_calc_modes.f:664-683, z_indexx.f:27
  fields + normalise    array_sol.f:9-347
  mode_pol              mode_energy.f:10-141 (adjoint family, py_calc_modes.f:709-714)
  orthogonality         orthogonal.f:3-337 (incl. conjugate-pair swap)
  adjoint normalise     normalisation.f:2-35
  PW ordering           pw_ordering.f:2-92
  J, J^dagger           J_overlap.f:3-246, J_dagger_overlap.f:3-246

The sparse LU + Arnoldi of the reference (UMFPACK + ARPACK) are replaced here
by scipy's SuperLU + ARPACK (scipy.sparse.linalg.eigs) on the same operator:
any backward-stable LU and converged Arnoldi give the same eigenpairs.

Parity pins (see tests/test_oracle_*.py): R/T/A text fixtures of the
reference's test cases 1 and 2 at 1e-6, brute-force dense eigenvalues of the
assembled pencil on small meshes, and element-matrix checks against
closed-form integrals.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

# ---------------------------------------------------------------------------
# Reference-element data
# ---------------------------------------------------------------------------


def quad_triangle():
    """16-point degree-8 rule on the unit triangle (quad_triangle.f:42-145)."""
    wbar = 1.443156076777862e-1 / 2.0
    w1 = 2.852749028018549e-1 / 6.0
    w2 = 9.737549286959440e-2 / 6.0
    w3 = 3.096521116041552e-1 / 6.0
    w4 = 1.633818850466092e-1 / 1.2e1
    cb = 1.0 / 3.0
    a1, b1 = 4.592925882927229e-1, 8.141482341455413e-2
    a2, b2 = 5.054722831703103e-2, 8.989055433659379e-1
    a3, b3 = 1.705693077517601e-1, 6.588613844964797e-1
    a4, b4, c4 = 7.284923929554041e-1, 2.63112829634638689e-1, 8.394777409957211e-3
    pts = [(cb, cb, wbar),
           (a1, a1, w1), (a1, b1, w1), (b1, a1, w1),
           (a2, a2, w2), (a2, b2, w2), (b2, a2, w2),
           (a3, a3, w3), (a3, b3, w3), (b3, a3, w3),
           (a4, b4, w4), (b4, a4, w4), (b4, c4, w4),
           (c4, b4, w4), (c4, a4, w4), (a4, c4, w4)]
    a = np.array(pts)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


def phi1(x, y):
    """P1 values/gradients on the reference triangle (phi1_2d_mat.f)."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    one = np.ones_like(x)
    zero = np.zeros_like(x)
    phi = np.stack([1 - x - y, x, y], -1)
    gx = np.stack([-one, one, zero], -1)
    gy = np.stack([-one, zero, one], -1)
    return phi, gx, gy


def phi2(x, y):
    """P2 values/gradients (phi2_2d_mat.f); node order 1,2,3 vertices,
    4=mid(1,2), 5=mid(2,3), 6=mid(3,1)."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    phi = np.stack([(-1 + x + y) * (-1 + 2 * x + 2 * y),
                    x * (-1 + 2 * x),
                    y * (-1 + 2 * y),
                    -4 * x * (-1 + x + y),
                    4 * x * y,
                    -4 * y * (-1 + x + y)], -1)
    gx = np.stack([-3 + 4 * x + 4 * y, -1 + 4 * x, 0 * x,
                   -4 * (-1 + 2 * x + y), 4 * y, -4 * y], -1)
    gy = np.stack([-3 + 4 * x + 4 * y, 0 * x, -1 + 4 * y,
                   -4 * x, 4 * x, -4 * (-1 + x + 2 * y)], -1)
    return phi, gx, gy


def phi3(x, y):
    """P3 values/gradients (phi3_2d_mat.f); node order: 3 vertices,
    2 nodes on each edge (1->2, 2->3, 3->1), 1 interior."""
    x = np.asarray(x, float)
    y = np.asarray(y, float)
    phi = np.stack([
        -((-1 + x + y) * (-2 + 3 * x + 3 * y) * (-1 + 3 * x + 3 * y)) / 2.,
        (x * (-2 + 3 * x) * (-1 + 3 * x)) / 2.,
        (y * (-2 + 3 * y) * (-1 + 3 * y)) / 2.,
        (9 * x * (-1 + x + y) * (-2 + 3 * x + 3 * y)) / 2.,
        (-9 * x * (-1 + 3 * x) * (-1 + x + y)) / 2.,
        (9 * x * (-1 + 3 * x) * y) / 2.,
        (9 * x * y * (-1 + 3 * y)) / 2.,
        (-9 * y * (-1 + x + y) * (-1 + 3 * y)) / 2.,
        (9 * y * (-1 + x + y) * (-2 + 3 * x + 3 * y)) / 2.,
        -27 * x * y * (-1 + x + y)], -1)
    g1 = (-11 - 27 * x ** 2 + x * (36 - 54 * y) + 36 * y - 27 * y ** 2) / 2.
    gx = np.stack([
        g1,
        1 - 9 * x + (27 * x ** 2) / 2.,
        0 * x,
        (9 * (2 + 9 * x ** 2 - 5 * y + 3 * y ** 2 + 2 * x * (-5 + 6 * y))) / 2.,
        (-9 * (1 + 9 * x ** 2 - y + x * (-8 + 6 * y))) / 2.,
        (9 * (-1 + 6 * x) * y) / 2.,
        (9 * y * (-1 + 3 * y)) / 2.,
        (-9 * y * (-1 + 3 * y)) / 2.,
        (9 * y * (-5 + 6 * x + 6 * y)) / 2.,
        -27 * y * (-1 + 2 * x + y)], -1)
    gy = np.stack([
        g1,
        0 * x,
        1 - 9 * y + (27 * y ** 2) / 2.,
        (9 * x * (-5 + 6 * x + 6 * y)) / 2.,
        (-9 * x * (-1 + 3 * x)) / 2.,
        (9 * x * (-1 + 3 * x)) / 2.,
        (9 * x * (-1 + 6 * y)) / 2.,
        (-9 * (1 - 8 * y + 9 * y ** 2 + x * (-1 + 6 * y))) / 2.,
        (9 * (2 + 3 * x ** 2 - 10 * y + 9 * y ** 2 + x * (-5 + 12 * y))) / 2.,
        -27 * x * (-1 + x + 2 * y)], -1)
    return phi, gx, gy


# ---------------------------------------------------------------------------
# Mesh and topology
# ---------------------------------------------------------------------------


class Mesh2D:
    """Everything the reference computes once per mesh (py_calc_modes.f:297-522)."""


def read_mail(path):
    """Read a .mail mesh (geometry.f:27-61).  Returns x[npt,2], type_nod[npt],
    table_nod[nel,6] (1-based node numbers), type_el[nel]."""
    with open(path) as f:
        toks = f.read().replace('D', 'E').replace('d', 'e').split()
    npt, nel = int(toks[0]), int(toks[1])
    p = 2
    body = np.array(toks[p:p + 4 * npt], dtype=float).reshape(npt, 4)
    x = body[:, 1:3].copy()
    type_nod = body[:, 3].astype(np.int64)
    p += 4 * npt
    el = np.array(toks[p:p + 8 * nel], dtype=float).astype(np.int64).reshape(nel, 8)
    table_nod = el[:, 1:7].copy()
    type_el = el[:, 7].copy()
    return x, type_nod, table_nod, type_el


def lattice_vec(x):
    """lattice_vec.f:4-91."""
    tol = 1e-6
    ymin, ymax = x[:, 1].min(), x[:, 1].max()
    bot = np.abs(x[:, 1] - ymin) < tol
    top = np.abs(x[:, 1] - ymax) < tol
    v1 = np.array([x[bot, 0].max() - x[bot, 0].min(), 0.0])
    v2 = np.array([x[top, 0].min() - x[bot, 0].min(), ymax - ymin])
    return np.stack([v1, v2], 1)  # lat[:, 0] = v1, lat[:, 1] = v2


def _lattice_shifts(lat):
    return np.array([ix * lat[:, 0] + iy * lat[:, 1]
                     for ix in (-1, 0, 1) for iy in (-1, 0, 1)])


def _periodic_pairs(pos, active, same_class, lat):
    """Shared pairing rule of periodic_node.f:58-106 and periodic_N_E_F.f:36-90:
    ascending i over active entries; every later active j (not yet paired)
    at a lattice offset of i joins i's class; the class representative is its
    smallest member.  Returns ip_period (1-based representative, 0 = none)."""
    n = len(pos)
    ip = np.zeros(n, dtype=np.int64)
    shifts = _lattice_shifts(lat)
    idx = np.nonzero(active)[0]
    tol = 1e-6
    for a, i in enumerate(idx[:-1]):
        js = idx[a + 1:]
        js = js[ip[js] == 0]
        if same_class is not None:
            js = js[same_class[js] == same_class[i]]
        if len(js) == 0:
            continue
        d = pos[js] - pos[i]
        hit = np.zeros(len(js), bool)
        for s in shifts:
            hit |= (np.abs(d[:, 0] - s[0]) + np.abs(d[:, 1] - s[1])) < tol
        js = js[hit]
        if len(js) == 0:
            continue
        if ip[i] == 0:
            ip[i] = i + 1
            ip[js] = i + 1
        else:
            ip[js] = ip[i]
    return ip


def build_mesh(path):
    """Per-mesh preprocessing of calc_modes_2d (py_calc_modes.f:297-522)."""
    m = Mesh2D()
    x, type_nod, table_nod, type_el = read_mail(path)
    npt, nel = len(x), len(table_nod)
    m.x, m.type_nod, m.table_nod, m.type_el = x, type_nod, table_nod, type_el
    m.npt, m.nel = npt, nel
    m.lat = lattice_vec(x)
    tn = table_nod - 1  # 0-based node numbers

    # list_face / list_edge (list_edge.f:37-69).  Entities are numbered
    # faces 1..nel, edges nel+1..nel+n_edge, P3 nodes after (1-based).
    T = np.zeros((nel, 14), dtype=np.int64)
    T[:, 0] = np.arange(1, nel + 1)
    visite = np.zeros(npt, dtype=np.int64)
    n_edge = 0
    for i in range(nel):
        for j in range(3, 6):
            j1 = tn[i, j]
            k = visite[j1]
            if k == 0:
                n_edge += 1
                visite[j1] = n_edge
                T[i, j - 2] = n_edge + nel
            else:
                T[i, j - 2] = k + nel
    # list_node_P3 (list_node_P3.f:29-128)
    visite[:] = 0
    mm2 = n_edge + nel
    nn = 0
    for iel in range(nel):
        nod = tn[iel]
        for inod in range(3):
            k = nod[inod]
            if visite[k] == 0:
                visite[k] = iel + 1
                nn += 1
                T[iel, inod + 4] = nn + mm2
            else:
                iel2 = visite[k] - 1
                inod2 = int(np.nonzero(tn[iel2, :3] == k)[0][0])
                T[iel, inod + 4] = T[iel2, inod2 + 4]
        for inod in range(3, 6):  # mid-edge nodes, edge inod-3: (a, b)
            k = nod[inod]
            a, b = inod - 3, (inod - 2) % 3
            v1 = (nod[a], nod[b])
            base = 7 + 2 * (inod - 3)  # 0-based column of first P3 edge node
            if visite[k] == 0:
                visite[k] = iel + 1
                for jj in range(2):
                    nn += 1
                    T[iel, base + jj] = nn + mm2
            else:
                iel2 = visite[k] - 1
                inod2 = int(np.nonzero(tn[iel2, 3:6] == k)[0][0]) + 3
                a2, b2 = inod2 - 3, (inod2 - 2) % 3
                v2 = (tn[iel2, a2], tn[iel2, b2])
                base2 = 7 + 2 * (inod2 - 3)
                for jj in range(2):
                    if v2 == v1:
                        T[iel, base + jj] = T[iel2, base2 + jj]
                    elif v2 == (v1[1], v1[0]):
                        T[iel, base + (1 - jj)] = T[iel2, base2 + jj]
                    else:
                        raise RuntimeError("list_node_P3: edge endpoints mismatch")
        nn += 1
        T[iel, 13] = nn + mm2
    npt_p3 = nn
    n_ddl = n_edge + nel + npt_p3
    m.n_edge, m.npt_p3, m.n_ddl = n_edge, npt_p3, n_ddl
    m.T = T

    # type_node_edge_face + get_coord_p3
    xE = np.zeros((n_ddl, 2))
    tE = np.zeros((n_ddl, 2), dtype=np.int64)  # (boundary type, dimension)
    xel = x[tn]  # nel,6,2
    xE[T[:, 0] - 1] = xel[:, :3].mean(1)
    tE[T[:, 0] - 1] = (0, 2)
    for j in range(3):
        e = T[:, j + 1] - 1
        xE[e] = xel[:, j + 3]
        tE[e, 0] = type_nod[tn[:, j + 3]]
        tE[e, 1] = 1
    for inod in range(3):
        e = T[:, inod + 4] - 1
        xE[e] = xel[:, inod]
        tE[e, 0] = type_nod[tn[:, inod]]
        tE[e, 1] = 0
    # P3 edge nodes: set from the first element that visits the edge
    # (get_coord_p3.f:84-109); values agree for every element by construction
    seen = np.zeros(npt, bool)
    ipn = [1, 2, 0]
    for iel in range(nel):
        for inod in range(3, 6):
            k = tn[iel, inod]
            if seen[k]:
                continue
            seen[k] = True
            p1 = x[tn[iel, inod - 3]]
            p2 = x[tn[iel, ipn[inod - 3]]]
            d = (p2 - p1) / 3.0
            for jj in range(2):
                e = T[iel, 7 + 2 * (inod - 3) + jj] - 1
                xE[e] = p1 + (jj + 1) * d
                tE[e] = (type_nod[k], 0)
    e = T[:, 13] - 1
    xE[e] = xel[:, :3].sum(1) * (1.0 / 3.0)
    tE[e] = (0, 0)
    m.xE, m.tE = xE, tE

    # periodic merge
    m.ip_period_N = _periodic_pairs(x, type_nod != 0, None, m.lat)
    if np.any((type_nod != 0) & (m.ip_period_N == 0)):
        raise RuntimeError("periodic_node: mesh is not periodic")
    m.ip_period_E = _periodic_pairs(xE, tE[:, 0] != 0, tE[:, 1], m.lat)
    if np.any((tE[:, 0] != 0) & (m.ip_period_E == 0)):
        raise RuntimeError("periodic_N_E_F: DOF set is not periodic")

    # periodic_cond: equation numbers ineq[n_ddl,3] (1-based, 0 = none)
    ineq = np.zeros((n_ddl, 3), dtype=np.int64)
    neq = 0
    for pas in (0, 1):
        for i in range(n_ddl):
            b, dim = tE[i]
            if pas == 0 and b != 0:
                continue
            if pas == 1 and (b == 0 or m.ip_period_E[i] != i + 1):
                continue
            if dim in (1, 2):
                ineq[i] = (neq + 1, neq + 2, neq + 3)
                neq += 3
            else:
                ineq[i, 0] = neq + 1
                neq += 1
    for i in range(n_ddl):
        if tE[i, 0] != 0 and m.ip_period_E[i] != i + 1:
            ineq[i] = ineq[m.ip_period_E[i] - 1]
    m.ineq, m.neq = ineq, neq

    # element -> 22 global equation numbers (0-based): 4 entities x 3 + 10 P3
    ld = np.zeros((nel, 22), dtype=np.int64)
    for jt in range(4):
        for q in range(3):
            ld[:, 3 * jt + q] = ineq[T[:, jt] - 1, q] - 1
    for jt in range(10):
        ld[:, 12 + jt] = ineq[T[:, 4 + jt] - 1, 0] - 1
    m.ldof = ld
    # local entity of each local dof (for the Bloch phase)
    m.lent = np.array([0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3] + list(range(4, 14)))

    # periodic representative node numbers used for edge orientation
    nod_p = table_nod.copy()
    rep = m.ip_period_N[table_nod - 1]
    nod_p[rep != 0] = rep[rep != 0]
    m.nod_p = nod_p
    # phase offsets x_E - x_rep per element entity
    dx = np.zeros((nel, 14, 2))
    ent = T - 1
    rp = m.ip_period_E[ent]
    has = rp != 0
    dx[has] = xE[ent[has]] - xE[rp[has] - 1]
    m.dphase = dx
    m.has_period = has

    # sparse pattern: union over elements of (ldof x ldof), sorted columns
    rows = np.repeat(ld, 22, axis=1).ravel()
    cols = np.tile(ld, (1, 22)).ravel()
    P = sp.csr_matrix((np.ones(len(rows)), (rows, cols)), shape=(neq, neq))
    P.sum_duplicates()
    P.sort_indices()
    m.row_ptr, m.col_ind = P.indptr.astype(np.int64), P.indices.astype(np.int64)
    m.nnz = P.nnz
    return m


# ---------------------------------------------------------------------------
# Element kernel and assembly (asmbly.f)
# ---------------------------------------------------------------------------


def element_geometry(m):
    """Affine P1 map per element (jacobian_p1_2d.f:45-101; curved elements are
    treated as affine, curved_elem_tri.f:63)."""
    xv = m.x[m.table_nod[:, :3] - 1]  # nel,3,2
    B = np.stack([xv[:, 1] - xv[:, 0], xv[:, 2] - xv[:, 0]], axis=2)  # B[:, j, i]
    det = B[:, 0, 0] * B[:, 1, 1] - B[:, 0, 1] * B[:, 1, 0]
    if np.any(np.abs(det) < 1e-10):
        raise RuntimeError("asmbly: det = 0")
    Tm = np.linalg.inv(B)  # mat_T
    return xv, B, det, Tm


def edge_orientation(m):
    """basis_ls.f:45-86: for each local edge (1,2),(2,3),(1,3) the local
    vertex with smaller (periodic-representative) global number first."""
    ends = [(0, 1), (1, 2), (0, 2)]
    o = np.zeros((m.nel, 3, 2), dtype=np.int64)
    for e, (a, b) in enumerate(ends):
        na, nb = m.nod_p[:, a], m.nod_p[:, b]
        if np.any(na == nb):
            raise RuntimeError("basis_ls: degenerate edge")
        sw = nb < na
        o[:, e, 0] = np.where(sw, b, a)
        o[:, e, 1] = np.where(sw, a, b)
    return o


def basis_at(xq, yq, Tm, orient):
    """Transverse (12) and longitudinal (10) basis at one reference point for
    all elements (basis_vec.f:17-71, asmbly.f:194-266).
    Returns vec[nel,12,2], curl[nel,12], zval[10], zgrad[nel,10,2]."""
    nel = Tm.shape[0]
    p1, g1x, g1y = phi1(xq, yq)
    p2, g2x, g2y = phi2(xq, yq)
    p3, g3x, g3y = phi3(xq, yq)

    def phys(gx, gy):  # grad_phys = mat_T^T grad_ref
        return np.stack([Tm[:, 0, 0, None] * gx + Tm[:, 1, 0, None] * gy,
                         Tm[:, 0, 1, None] * gx + Tm[:, 1, 1, None] * gy], -1)
    G1 = phys(g1x, g1y)  # nel,3,2
    G2 = phys(g2x, g2y)  # nel,6,2
    G3 = phys(g3x, g3y)  # nel,10,2
    vec = np.zeros((nel, 12, 2))
    curl = np.zeros((nel, 12))
    ar = np.arange(nel)
    opp = [2, 0, 1]
    for j in range(3):  # face functions: p2(mid j) grad p1(opposite vertex)
        mnode, n1 = j + 3, opp[j]
        vec[:, j] = p2[mnode] * G1[:, n1]
        curl[:, j] = G2[:, mnode, 0] * G1[:, n1, 1] - G2[:, mnode, 1] * G1[:, n1, 0]
    for e in range(3):  # edge functions
        j1, j2 = orient[:, e, 0], orient[:, e, 1]
        b = 3 + 3 * e
        for q, (mn, gp) in enumerate([(j1, G1[ar, j2]), (j2, G1[ar, j1]),
                                      (np.full(nel, e + 3), G1[ar, j1] - G1[ar, j2])]):
            ph = p2[mn]
            g2 = G2[ar, mn]
            vec[:, b + q] = ph[:, None] * gp
            curl[:, b + q] = g2[:, 0] * gp[:, 1] - g2[:, 1] * gp[:, 0]
    return vec, curl, p3, G3


def element_matrices(m, pp, qq, shift, bloch):
    """Per element 22x22 blocks of mat1 = K - shift*M and mat2 = M with the
    Bloch phase applied (asmbly.f:107-322); E[e, i, j] is the contribution to
    global entry (row = ldof[e,i], col = ldof[e,j])."""
    xq, yq, wq = quad_triangle()
    xv, B, det, Tm = element_geometry(m)
    orient = edge_orientation(m)
    nel = m.nel
    typ = m.type_el - 1
    ppe = np.asarray(pp)[typ]
    qqe = np.asarray(qq)[typ]
    K = np.zeros((nel, 22, 22), complex)
    M = np.zeros((nel, 22, 22), complex)
    adet = np.abs(det)
    for q in range(len(wq)):
        vec, curl, zv, zg = basis_at(xq[q], yq[q], Tm, orient)
        w = wq[q] * adet
        dot_tt = np.einsum('eia,eja->eij', vec, vec)
        cc = curl[:, :, None] * curl[:, None, :]
        # rows i (trial), cols j (test)
        K[:, :12, :12] += w[:, None, None] * (cc * ppe[:, None, None] - dot_tt * qqe[:, None, None])
        M[:, :12, :12] += w[:, None, None] * (-dot_tt * ppe[:, None, None])
        # row transverse i, col z j: M_tz = (grad z_j . vec_i) pp
        M[:, :12, 12:] += w[:, None, None] * np.einsum('eia,eja->eij', vec, zg) * ppe[:, None, None]
        # row z i, col transverse j: K_zt = (vec_j . grad z_i) pp
        K[:, 12:, :12] += w[:, None, None] * np.einsum('eia,eja->eij', zg, vec) * ppe[:, None, None]
        # z-z: K_zz = -(grad_j.grad_i) pp + z_j z_i qq
        gg = np.einsum('eia,eja->eij', zg, zg)
        zz = np.outer(zv, zv)
        K[:, 12:, 12:] += w[:, None, None] * (-gg * ppe[:, None, None] + zz[None] * qqe[:, None, None])
    # Bloch phase val_exp(j) conj(val_exp(i)) (asmbly.f:121-146, :234)
    ph = np.exp(1j * (m.dphase @ np.asarray(bloch, float)))  # nel,14
    ph[~m.has_period] = 1.0
    phl = ph[:, m.lent]  # nel,22
    fac = phl[:, None, :] * np.conj(phl)[:, :, None]
    K *= fac
    M *= fac
    return K - shift * M, M


def assemble(m, pp, qq, shift, bloch):
    """Global mat1 = K - shift M and mat2 = M as scipy CSR matrices
    (row = ind_ip, col = ind_jp as in asmbly.f:304-314)."""
    E1, E2 = element_matrices(m, pp, qq, shift, bloch)
    rows = np.repeat(m.ldof, 22, axis=1).ravel()
    cols = np.tile(m.ldof, (1, 22)).ravel()
    A = sp.csr_matrix((E1.ravel(), (rows, cols)), shape=(m.neq, m.neq))
    Bm = sp.csr_matrix((E2.ravel(), (rows, cols)), shape=(m.neq, m.neq))
    A.sum_duplicates()
    Bm.sum_duplicates()
    return A, Bm


# ---------------------------------------------------------------------------
# Eigen-solve and post-processing
# ---------------------------------------------------------------------------


def nr_indexx(key):
    """Numerical-Recipes quicksort index (z_indexx.f / sort_n.f), ascending
    key, 0-based indices.  Kept verbatim in behaviour because it is not a
    stable sort and ties must break as in the reference."""
    key = np.asarray(key, float)
    n = len(key)
    indx = list(range(n))
    M = 7
    istack = []
    l, ir = 0, n - 1
    while True:
        if ir - l < M:
            for j in range(l + 1, ir + 1):
                indxt = indx[j]
                a = key[indxt]
                i = j - 1
                while i >= l:
                    if key[indx[i]] <= a:
                        break
                    indx[i + 1] = indx[i]
                    i -= 1
                indx[i + 1] = indxt
            if not istack:
                return np.array(indx, dtype=np.int64)
            ir = istack.pop()
            l = istack.pop()
        else:
            k = (l + ir) // 2
            indx[k], indx[l + 1] = indx[l + 1], indx[k]
            if key[indx[l]] > key[indx[ir]]:
                indx[l], indx[ir] = indx[ir], indx[l]
            if key[indx[l + 1]] > key[indx[ir]]:
                indx[l + 1], indx[ir] = indx[ir], indx[l + 1]
            if key[indx[l]] > key[indx[l + 1]]:
                indx[l], indx[l + 1] = indx[l + 1], indx[l]
            i, j = l + 1, ir
            indxt = indx[l + 1]
            a = key[indxt]
            while True:
                i += 1
                while key[indx[i]] < a:
                    i += 1
                j -= 1
                while key[indx[j]] > a:
                    j -= 1
                if j < i:
                    break
                indx[i], indx[j] = indx[j], indx[i]
            indx[l + 1] = indx[j]
            indx[j] = indxt
            if ir - i + 1 >= j - l:
                istack.extend([i, ir])
                ir = j - 1
            else:
                istack.extend([l, j - 1])
                l = i


def beta_from_mu(mu, shift):
    """beta = sqrt(1/mu + shift) with the branch rule of py_calc_modes.f:664-679."""
    b = np.sqrt(1.0 / mu + shift)
    prop = np.abs(b.imag) < 1e-8
    flip = np.where(prop, b.real < 0, b.imag < 0)
    return np.where(flip, -b, b)


def eigensolve(A, Bm, nval, nvect, itermax=30, tol=0.0, seed=None):
    """nval largest-|mu| eigenpairs of OP = A^-1 Bm (valpr_umf_64.f:291-459):
    scipy SuperLU + ARPACK ('LM', mode 1 on the user OP)."""
    lu = spla.splu(A.tocsc(), permc_spec='MMD_AT_PLUS_A')
    n = A.shape[0]
    op = spla.LinearOperator((n, n), matvec=lambda v: lu.solve(Bm @ v), dtype=complex)
    v0 = None
    if seed is not None:
        r = np.random.default_rng(seed)
        v0 = r.standard_normal(n) + 1j * r.standard_normal(n)
    mu, V = spla.eigs(op, k=nval, ncv=min(nvect, n), which='LM', tol=tol,
                      maxiter=max(itermax, 1) * 10, v0=v0)
    return mu, V


def fields_at_nodes(m, V, bloch):
    """DOF vectors -> per-element fields sol[3, 13, nval, nel] (array_sol.f:120-263)
    before normalisation.  Also returns the normalisation factors z_sol_max."""
    nel = m.nel
    nval = V.shape[1]
    xv, B, det, Tm = element_geometry(m)
    orient = edge_orientation(m)
    ph = np.exp(1j * (m.dphase @ np.asarray(bloch, float)))
    ph[~m.has_period] = 1.0
    phl = ph[:, m.lent]
    coef = V[m.ldof] * phl[:, :, None]  # nel,22,nval
    sol = np.zeros((3, 13, nval, nel), complex)
    xn = [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)]
    # map: transverse basis function index -> its P2 node m
    for inod, (a, b) in enumerate(xn):
        vec, curl, zv, zg = basis_at(a, b, Tm, orient)
        # basis node of each transverse function
        mnode = np.zeros((nel, 12), dtype=np.int64)
        mnode[:, 0:3] = [3, 4, 5]
        for e in range(3):
            mnode[:, 3 + 3 * e] = orient[:, e, 0]
            mnode[:, 4 + 3 * e] = orient[:, e, 1]
            mnode[:, 5 + 3 * e] = e + 3
        sel = (mnode == inod).astype(float)  # only functions attached to node inod
        for c in range(2):
            w = vec[:, :, c] * sel  # nel,12
            sol[c, inod] = np.einsum('ei,ein->ne', w, coef[:, :12])
        sol[2, inod] = np.einsum('i,ein->ne', zv, coef[:, 12:])
    for k in range(7):  # P3 edge + interior nodes carry Ez only
        sol[2, 6 + k] = coef[:, 12 + 3 + k].T
    # z_sol_max: first largest |component| scanning iel, inod(1..6), comp
    z = np.zeros(nval, complex)
    blk = sol[:, :6]  # 3,6,nval,nel
    scan = np.transpose(blk, (2, 3, 1, 0)).reshape(nval, -1)  # nval, nel*6*3
    ia = np.argmax(np.abs(scan), axis=1)  # first occurrence of the maximum
    z = scan[np.arange(nval), ia]
    return sol, z


def core_type(eps):
    """n_core of py_calc_modes.f:546-551: material 1 when Re eps_1 > Re eps_2,
    else material 2 (1-based)."""
    eps = np.asarray(eps, complex)
    return 1 if len(eps) < 2 or eps[0].real > eps[1].real else 2


def mode_energy(m, sol, beta, eps):
    """mode_pol (4, nval) of mode_energy.f:60-141: on the 16-point rule,
    w |det| eps(type) |E_c|^2 with E the P2 interpolant of the six node values
    of sol (3, 13, nval, nel) and E_z scaled by beta; rows x, y, z and the
    sum over the core material, divided by the x + y + z total."""
    eps = np.asarray(eps, complex)
    xq, yq, wq = quad_triangle()
    phi, _, _ = phi2(xq, yq)  # nq, 6
    _, _, det, _ = element_geometry(m)
    coef = np.abs(det) * eps[m.type_el - 1]  # nel
    core = m.type_el == core_type(eps)
    nval = sol.shape[2]
    pol = np.zeros((4, nval), complex)
    for q in range(len(wq)):
        v = np.einsum('i,cine->cne', phi[q], sol[:, :6])  # 3, nval, nel
        v[2] = v[2] * np.asarray(beta)[:, None]
        e2 = wq[q] * coef[None, None, :] * np.abs(v) ** 2
        pol[:3] += e2.sum(axis=2)
        pol[3] += e2[:, :, core].sum(axis=(0, 2))
    return pol / (pol[0] + pol[1] + pol[2])[None, :]


def overlap_matrix(m, sol1, sol2, beta1, k0, pp):
    """orthogonal.f:95-248: overlap_L(i,j) = sum_e sol2_e(i)^T S_e sol1_e(j) beta1(j)."""
    xq, yq, wq = quad_triangle()
    xv, B, det, Tm = element_geometry(m)
    nel = m.nel
    typ = m.type_el - 1
    ppe = np.asarray(pp)[typ]
    S = np.zeros((nel, 12, 22), complex)
    for q in range(len(wq)):
        p2, _, _ = phi2(xq[q], yq[q])
        _, g3x, g3y = phi3(xq[q], yq[q])
        G3 = np.stack([Tm[:, 0, 0, None] * g3x + Tm[:, 1, 0, None] * g3y,
                       Tm[:, 0, 1, None] * g3x + Tm[:, 1, 1, None] * g3y], -1)
        c = wq[q] * np.abs(det) * ppe / k0
        for it in range(6):
            for ie in range(2):
                ip = ie + 2 * it
                for jt in range(6):
                    S[:, ip, ie + 2 * jt] += c * p2[it] * p2[jt]
                S[:, ip, 12:] += (c * p2[it])[:, None] * (-G3[:, :, ie])
    # sol_el_2 (12) and sol_el_1 (22)
    s2 = np.transpose(sol2[:2, :6], (3, 1, 0, 2)).reshape(nel, 12, -1)  # e, (node,comp), n
    s1t = np.transpose(sol1[:2, :6], (3, 1, 0, 2)).reshape(nel, 12, -1)
    s1z = np.concatenate([sol1[2, 0:3], sol1[2, 6:13]], axis=0)  # 10, nval, nel
    s1 = np.concatenate([s1t, np.transpose(s1z, (2, 0, 1))], axis=1) * beta1[None, None, :]
    W = np.einsum('eij,ejn->ein', S, s1)
    return np.einsum('eim,ein->mn', s2, W)


def pair_swap(L):
    """Conjugate-pair reordering of orthogonal.f:251-301; returns the list of
    (j, j2) swaps applied in order."""
    nval = L.shape[0]
    swaps = []
    j = 0
    skip = 0
    while j < nval:
        if abs(L[j, j]) > 1e-7:
            j += 1
        elif skip == 0:
            for j2 in range(j + 1, nval):
                if abs(L[j2, j]) > 1e-7:
                    swaps.append((j, j2))
            skip = 1
            j += 1
        else:
            skip = 0
            j += 1
    return swaps


def pw_list(ordre, d=1.0, dy=1.0):
    """PW orders in the reference's loop order (px outer, py inner) and the
    circular truncation of pw_ordering.f:54-67 / mode_calcs.py:61-74."""
    out = []
    for px in range(-ordre, ordre + 1):
        for py in range(-ordre, ordre + 1):
            if (px / d) ** 2 + (py / dy) ** 2 <= (ordre / max(d, dy)) ** 2:
                out.append((px, py))
    return np.array(out, dtype=np.int64).reshape(-1, 2)


def pw_order_inv(ordre, lat, bloch, k0):
    """index_pw_inv of pw_ordering.f:48-78 (0-based positions)."""
    d, dy = lat[0, 0], lat[1, 1]
    pw = pw_list(ordre, d, dy)
    al = bloch[0] + 2 * np.pi / d * pw[:, 0]
    be = bloch[1] + 2 * np.pi / dy * pw[:, 1]
    kz = np.sqrt((k0 ** 2 - al ** 2 - be ** 2).astype(complex))
    idx = nr_indexx(-(kz ** 2).real)
    inv = np.empty(len(pw), dtype=np.int64)
    inv[idx] = np.arange(len(pw))
    return inv, pw, al, be


def j_overlaps(m, sol1, sol2, ordre, bloch, k0):
    """J (2P x nval) and J^dagger (nval x 2P) of J_overlap.f / J_dagger_overlap.f."""
    xq, yq, wq = quad_triangle()
    xv, B, det, Tm = element_geometry(m)
    inv, pw, al, be = pw_order_inv(ordre, m.lat, bloch, k0)
    P = len(pw)
    norm = 1.0 / np.sqrt(al ** 2 + be ** 2)
    nval = sol1.shape[2]
    oK = np.zeros((P, nval), complex)
    oE = np.zeros((P, nval), complex)
    dK = np.zeros((P, nval), complex)
    dE = np.zeros((P, nval), complex)
    for q in range(len(wq)):
        p2, _, _ = phi2(xq[q], yq[q])
        xg = xv[:, 0] + B[:, :, 0] * xq[q] + B[:, :, 1] * yq[q]  # nel,2
        c = wq[q] * np.abs(det)
        f1 = np.einsum('k,ckne->cne', p2, sol1[:2, :6])  # 2,nval,nel
        f2 = np.einsum('k,ckne->cne', p2, sol2[:2, :6])
        r = np.outer(al, xg[:, 0]) + np.outer(be, xg[:, 1])  # P,nel
        em = np.exp(-1j * r) * norm[:, None] * c[None, :]
        ep = np.exp(1j * r) * norm[:, None] * c[None, :]
        # K-type (alpha, beta), E-type (beta, -alpha)
        oK += em @ f1[0].T * al[:, None] + em @ f1[1].T * be[:, None]
        oE += em @ f1[0].T * be[:, None] - em @ f1[1].T * al[:, None]
        dK += ep @ f2[0].T * al[:, None] + ep @ f2[1].T * be[:, None]
        dE += ep @ f2[0].T * be[:, None] - ep @ f2[1].T * al[:, None]
    J = np.zeros((2 * P, nval), complex)
    Jd = np.zeros((nval, 2 * P), complex)
    J[inv] = oE
    J[P + inv] = oK
    Jd[:, inv] = dE.T
    Jd[:, P + inv] = dK.T
    return J, Jd


def biorth_clusters(L, rel=1e-10):
    """Groups of modes coupled through the off-diagonal of the overlap
    matrix L (|L_ij| > rel sqrt(|L_ii L_jj|)): near-degenerate modes whose
    prime and adjoint vectors are not biorthogonal to working precision.
    Returns a list of index arrays (singletons included)."""
    n = L.shape[0]
    d = np.sqrt(np.abs(np.diag(L)))
    A = np.abs(L) > rel * np.outer(d, d)
    np.fill_diagonal(A, False)
    A = A | A.T
    parent = list(range(n))

    def find(i):
        while parent[i] != i:
            parent[i] = parent[parent[i]]
            i = parent[i]
        return i
    for i, j in zip(*np.nonzero(A)):
        a, b = find(i), find(j)
        if a != b:
            parent[max(a, b)] = min(a, b)
    groups = {}
    for i in range(n):
        groups.setdefault(find(i), []).append(i)
    return [np.array(g) for g in sorted(groups.values())]


def adjoint_normaliser(L, biorth='diag'):
    """Matrix N with J^dagger = N J^dagger_raw.  'diag' is the reference's
    normalisation.f:15-31 (N = diag(1/L_ii), entries with |sqrt(L_ii)| <=
    1e-8 left unscaled); 'cluster' inverts L within each group of
    biorthogonally coupled modes (the same as 'diag' for isolated modes),
    which gives the projector the exact eigenvectors would give."""
    n = L.shape[0]
    N = np.zeros((n, n), complex)
    dg = np.sqrt(np.diag(L))
    for k in range(n):
        N[k, k] = 1.0 / dg[k] ** 2 if abs(dg[k]) > 1e-8 else 1.0
    if biorth == 'diag':
        return N
    for g in biorth_clusters(L):
        if g.size > 1 and np.all(np.abs(dg[g]) > 1e-8):
            N[np.ix_(g, g)] = np.linalg.inv(L[np.ix_(g, g)])
    return N


def calc_modes_2d(wl_norm, nval, ordre_ls, mesh, n_eff, bloch_vec, shift,
                  neq_PW, itermax=30, E_H_field=1, eig_tol=0.0, return_debug=False, biorth='diag'):
    """Oracle of EMUstack.calc_modes_2d (py_calc_modes.f:1-958).

    mesh: a Mesh2D from build_mesh().  Returns (beta1, J, J_dagger) where J is
    (2 neq_PW x nval) and J_dagger (nval x 2 neq_PW), before the Python-side
    area normalisation (mode_calcs.py:384-386)."""
    m = mesh
    n_eff = np.asarray(n_eff, complex)
    eps = n_eff ** 2
    k0 = 2.0 * np.pi / wl_norm
    if E_H_field == 1:
        qq, pp = eps * k0 ** 2, np.ones_like(eps)
    else:
        qq, pp = np.full_like(eps, k0 ** 2), 1.0 / eps
    if len(eps) > 1 and np.all(eps == eps[0]):
        raise ValueError("FEM routine cannot handle homogeneous layers")
    nvect = 2 * nval + nval // 2 + 3
    bloch_vec = np.asarray(bloch_vec, float)
    res = {}
    for n_k, bv in ((1, bloch_vec), (2, -bloch_vec)):
        A, Bm = assemble(m, pp, qq, shift, bv)
        mu, V = eigensolve(A, Bm, nval, nvect, itermax, eig_tol)
        beta = beta_from_mu(mu, shift)
        idx = nr_indexx(-(beta ** 2).real)
        beta = beta[idx]
        V = V[:, idx]
        sol, zmax = fields_at_nodes(m, V, bv)
        sol = sol / zmax[None, None, :, None]
        res[n_k] = (beta, sol, V / zmax[None, :], A, Bm)
    beta1, sol1 = res[1][0], res[1][1]
    beta2, sol2 = res[2][0].copy(), res[2][1].copy()
    L = overlap_matrix(m, sol1, sol2, beta1, k0, pp)
    for (j, j2) in pair_swap(L):
        beta2[[j, j2]] = beta2[[j2, j]]
        sol2[:, :, [j, j2]] = sol2[:, :, [j2, j]]
    L = overlap_matrix(m, sol1, sol2, beta1, k0, pp)
    J, Jd_raw = j_overlaps(m, sol1, sol2, ordre_ls, bloch_vec, k0)
    Jd = adjoint_normaliser(L, biorth) @ Jd_raw
    if return_debug:
        # mode_pol: the reference's last writer is mode_energy on the adjoint
        # family in its own sorted order (before the pair swap)
        pol = mode_energy(m, res[2][1], res[2][0], eps)
        return beta1, J, Jd, dict(res=res, L=L, sol1=sol1, sol2=sol2, k0=k0, Jd_raw=Jd_raw, mode_pol=pol)
    return beta1, J, Jd
