"""Periodic P2 triangular meshes of a unit cell in the reference's .mail format.

Replaces the reference's gmsh step (backend/fortran/conv_gmsh.f:58-61 and the
msh/*.geo templates), which needs an external gmsh.  Conventions of the
reference's meshes (checked on its fixture meshes):
  * coordinates normalised by the period: x in [0, 1], y in [-period_y/period, 0];
  * node boundary tags 1 = top (y = 0), 2 = bottom, 3 = left (x = 0),
    4 = right (x = 1) (corners carry the left/right tag), interior 0;
  * element nodes (v1, v2, v3, mid(1,2), mid(2,3), mid(3,1));
  * element material tag 1 = background, 2 = inclusion.
Opposite sides carry identical point positions, so the periodic pairing of
periodic_node.f:58-106 succeeds.

Points: boundary points at spacing lc, inclusion outline at spacing lc,
interior points on a seeded jittered triangular lattice; the triangulation is
scipy's Delaunay.  Host-side preprocessing only (not on the GPU path).
"""
import numpy as np
from scipy.spatial import Delaunay

__all__ = ['ellipse_cell', 'write_mail', 'lc_for_triangles', 'ellipse_cell_ntri']


def _side(n):
    return np.linspace(0.0, 1.0, n + 1)


def ellipse_cell(period, period_y, a_nm, b_nm, rotation_deg=0.0, lc=0.09, seed=0, center=(0.5, 0.5)):
    """Mesh of one unit cell with an elliptical inclusion of semi-axes a_nm,
    b_nm (nm) rotated by rotation_deg.  Returns (x[npt,2], type_nod[npt],
    table_nod[nel,6] 1-based, type_el[nel])."""
    dy = period_y / period
    a, b = a_nm / period, b_nm / period
    rng = np.random.default_rng(seed)
    nx = max(2, int(round(1.0 / lc)))
    ny = max(2, int(round(dy / lc)))
    sx = _side(nx)
    sy = -dy * _side(ny)
    pts = []
    # boundary: corners + sides (identical positions on opposite sides)
    for xv in sx:
        pts.append((xv, 0.0))
        pts.append((xv, -dy))
    for yv in sy[1:-1]:
        pts.append((0.0, yv))
        pts.append((1.0, yv))
    nb = len(pts)
    cx, cy = center[0], -center[1] * dy
    th = np.deg2rad(rotation_deg)
    ct, st = np.cos(th), np.sin(th)
    have_inc = a > 0 and b > 0
    if have_inc:
        # outline at arc-length spacing ~ lc (Ramanujan perimeter)
        h = ((a - b) / (a + b)) ** 2
        perim = np.pi * (a + b) * (1 + 3 * h / (10 + np.sqrt(4 - 3 * h)))
        nt = max(8, int(round(perim / (0.8 * lc))))
        t = np.linspace(0, 2 * np.pi, nt, endpoint=False)
        ex, ey = a * np.cos(t), b * np.sin(t)
        ox = cx + ct * ex - st * ey
        oy = cy + st * ex + ct * ey
        if ox.min() <= lc * 0.5 or ox.max() >= 1 - lc * 0.5 or oy.max() >= -lc * 0.5 or oy.min() <= -dy + lc * 0.5:
            raise ValueError("inclusion does not fit inside the unit cell")
        outline = np.stack([ox, oy], 1)
        pts.extend(map(tuple, outline))
    # interior: jittered triangular lattice
    hy = lc * np.sqrt(3) / 2
    rows = int(np.ceil(dy / hy)) + 1
    cols = int(np.ceil(1.0 / lc)) + 1
    grid = []
    for r in range(rows):
        for c in range(cols):
            xv = (c + 0.5 * (r % 2)) * lc
            yv = -r * hy
            grid.append((xv, yv))
    grid = np.array(grid) + rng.uniform(-0.15, 0.15, (len(grid), 2)) * lc
    P = np.array(pts)
    keep = (grid[:, 0] > 0.45 * lc) & (grid[:, 0] < 1 - 0.45 * lc) & (grid[:, 1] < -0.45 * lc) & \
           (grid[:, 1] > -dy + 0.45 * lc)
    grid = grid[keep]
    if have_inc:
        # distance to the outline (in the ellipse frame, approximate)
        dxg, dyg = grid[:, 0] - cx, grid[:, 1] - cy
        u = ct * dxg + st * dyg
        v = -st * dxg + ct * dyg
        r = np.sqrt((u / a) ** 2 + (v / b) ** 2)
        gap = np.abs(r - 1.0) * min(a, b)
        grid = grid[gap > 0.45 * lc]
    allp = np.vstack([P, grid])
    tri = Delaunay(allp)
    simp = tri.simplices.copy()
    if len(np.unique(simp)) != len(allp):
        raise RuntimeError("mesher: triangulation dropped points")
    v = allp[simp]
    det = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 2, 0] - v[:, 0, 0]) * (v[:, 1, 1] - v[:, 0, 1])
    good = np.abs(det) > 1e-12
    simp = simp[good]
    det = det[good]
    neg = det < 0
    simp[neg] = simp[neg][:, [0, 2, 1]]
    # material by centroid
    cen = allp[simp].mean(1)
    typ = np.ones(len(simp), np.int64)
    if have_inc:
        dxg, dyg = cen[:, 0] - cx, cen[:, 1] - cy
        u = ct * dxg + st * dyg
        vv = -st * dxg + ct * dyg
        typ[(u / a) ** 2 + (vv / b) ** 2 < 1.0] = 2
    # boundary tags of vertices
    tol = 1e-12

    def tag(p):
        t = np.zeros(len(p), np.int64)
        t[np.abs(p[:, 1]) < tol] = 1
        t[np.abs(p[:, 1] + dy) < tol] = 2
        t[np.abs(p[:, 0]) < tol] = 3
        t[np.abs(p[:, 0] - 1) < tol] = 4
        return t
    vtag = tag(allp)
    # P2 mid-edge nodes
    edges = np.concatenate([simp[:, [0, 1]], simp[:, [1, 2]], simp[:, [2, 0]]])
    key = np.sort(edges, axis=1)
    uniq, inv = np.unique(key, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    mids = 0.5 * (allp[uniq[:, 0]] + allp[uniq[:, 1]])
    mtag = tag(mids)
    # a mid node is on the boundary only if both ends lie on that side
    ends_b = (vtag[uniq[:, 0]] != 0) & (vtag[uniq[:, 1]] != 0)
    mtag[~ends_b] = 0
    nv = len(allp)
    x = np.vstack([allp, mids])
    tnod = np.concatenate([vtag, mtag])
    nel = len(simp)
    table = np.zeros((nel, 6), np.int64)
    table[:, :3] = simp
    table[:, 3] = nv + inv[0:nel]
    table[:, 4] = nv + inv[nel:2 * nel]
    table[:, 5] = nv + inv[2 * nel:3 * nel]
    # renumber nodes along a space-filling order for locality
    order = np.lexsort((x[:, 0], np.round(x[:, 1] / (lc * 2)).astype(np.int64)))
    newid = np.empty(len(x), np.int64)
    newid[order] = np.arange(len(x))
    x = x[order]
    tnod = tnod[order]
    table = newid[table]
    ecen = x[table[:, :3]].mean(1)
    eorder = np.lexsort((ecen[:, 0], np.round(ecen[:, 1] / (lc * 2)).astype(np.int64)))
    return x, tnod, table[eorder] + 1, typ[eorder]


def write_mail(path, mesh):
    """Write (x, type_nod, table_nod, type_el) as a .mail file (geometry.f:27-61)."""
    x, tnod, table, typ = mesh
    with open(path, 'w') as f:
        f.write('%12d %12d\n' % (len(x), len(table)))
        for i, (p, t) in enumerate(zip(x, tnod)):
            f.write('%10d %24.16E %24.16E %6d\n' % (i + 1, p[0], p[1], t))
        for i, (row, t) in enumerate(zip(table, typ)):
            f.write('%10d ' % (i + 1) + ' '.join('%10d' % v for v in row) + ' %6d\n' % t)


def lc_for_triangles(ntri, period=1.0, period_y=None):
    """Target edge length giving about ntri triangles in the cell."""
    dy = (period_y or period) / period
    return float(np.sqrt(4.0 * dy / (np.sqrt(3) * ntri)))


def ellipse_cell_ntri(period, period_y, a_nm, b_nm, ntri, rotation_deg=0.0, seed=0):
    """Mesh with about ntri triangles (lc adjusted once on the measured count)."""
    lc = lc_for_triangles(ntri, period, period_y)
    m = ellipse_cell(period, period_y, a_nm, b_nm, rotation_deg, lc, seed)
    lc *= np.sqrt(len(m[2]) / ntri)
    return ellipse_cell(period, period_y, a_nm, b_nm, rotation_deg, lc, seed)


# --------------------------------------------------------------- 1D gratings

def _dec_float_str(v):
    """The reference's mesh-name number format: str(v) with '.' -> '_'."""
    return str(v).replace('.', '_')


def grating_layout_1d(period, diameters, edge_spacing=False, split_touching_incs=False, small_space=None):
    """Material-type rules of a 1D grating cell with up to six inclusions,
    as intervals of the unit cell (objects.py:1054-1506 of the reference).

    Returns a list of (type, lo, hi, slack): an element [x1, x2] takes the
    type of the first rule with lo - slack * dx <= x1, x2 <= hi (dx the
    element width; lo = -inf / hi = +inf for one-sided rules); elements no
    rule takes are background (type 1).  The number of inclusions N is the
    index of the last positive diameter (earlier zero diameters stay in
    place with zero width).  Layouts: inclusion 1 (type 2) is centred on
    0.5; then
      * N = 2: inclusion 2 (type 3) touching inclusion 1 on both sides
        (split_touching_incs), or at the cell edges after a gap small_space
        on the left and the rest of the cell on the right;
      * N >= 3, edge_spacing: inclusions alternate left / right of the
        centre with equal gaps i_d between them;
      * N >= 3, split_touching_incs: inclusion k touches k - 1 on both sides;
      * N >= 3 otherwise: inclusion k centred at 0.5 -/+ (k // 2) / N,
        alternating left / right, for N = 4 and N = 6 the last inclusion
        split across the cell edges.
    The bounds are evaluated in the reference's order of operations, so that
    an element end that meets a bound compares the same way.  What the
    reference's rules do, including the irregular ones, is kept:
      * N = 3 edge_spacing: i_d = (1 - d1 - d2 - d3) / 3 from the diameters
        in nm (objects.py:1388), not the normalised widths;
      * N = 4 edge_spacing: the fourth inclusion is not placed; elements past
        the second gap on either side take types 4 and 3 (objects.py:1318-1323);
      * N = 5 edge_spacing: the fifth inclusion takes type 4 (objects.py:1230-1232);
      * N = 6 edge_spacing: a type-6 band of width r5 left of inclusion 4 and
        inclusion 6 (type 7) on the right (objects.py:1118-1126);
      * the slack dx widens the centre rule always and the touching rules
        for N = 2 and N = 3 only.
    """
    d = [float(v) for v in diameters] + [0.0] * (6 - len(diameters))
    if len(d) > 6:
        raise ValueError("1D gratings take at most six inclusions (diameter1..diameter6)")
    pos = [k for k in range(6) if d[k] > 0]
    if not pos:
        raise ValueError("Must have at least one grating of nonzero width.")
    N = pos[-1] + 1
    r = [0.0] + [v / (2.0 * period) for v in d]  # r[k]: half-width of inclusion k, unit cell
    inf = np.inf
    rules = [(2, 0.5 - r[1], 0.5 + r[1], 1)]
    if N == 1:
        return rules
    # touching layout: cumulative edges, accumulated term by term
    if split_touching_incs and (N == 2 or not edge_spacing):
        slack = 1 if N <= 3 else 0
        right, left = 0.5 + r[1], 0.5 - r[1]
        for k in range(2, N + 1):
            rules.append((k + 1, right, right + r[k], slack))
            rules.append((k + 1, left - r[k], left, slack))
            right, left = right + r[k], left - r[k]
        return rules
    if N == 2:
        if small_space is None:
            small = large = 0.5 - r[1] - r[2]
        else:
            small = small_space
            large = 1.0 - small_space - (2 * r[1]) - (2 * r[2])
        rules.append((3, -inf, 0.5 - r[1] - small, 0))
        rules.append((3, 0.5 + large + r[1], inf, 0))
        return rules
    if edge_spacing:
        if N == 3:
            g = (1.0 - d[0] - d[1] - d[2]) / 3.0
            rules += [(3, 0.5 - g - 2.0 * r[2] - r[1], 0.5 - g - r[1], 0),
                      (4, 0.5 + g + r[1], 0.5 + g + 2.0 * r[3] + r[1], 0)]
            return rules
        rem = 0.5
        for k in range(1, N + 1):
            rem = rem - r[k]
        g = 2.0 * rem / float(N)
        if N == 4:
            rules += [(3, 0.5 - g - r[1] - 2.0 * r[2], 0.5 - g - r[1], 0),
                      (4, 0.5 + g + r[1], 0.5 + g + r[1] + 2.0 * r[3], 0),
                      (4, 0.5 + 2.0 * g + r[1] + 2.0 * r[3], inf, 0),
                      (3, -inf, 0.5 - 2.0 * g - r[1] - 2.0 * r[2], 0)]
            return rules
        rules += [(3, 0.5 - g - r[1] - 2.0 * r[2], 0.5 - g - r[1], 0),
                  (4, 0.5 + g + r[1], 0.5 + g + 2.0 * r[3] + r[1], 0),
                  (5, 0.5 - 2.0 * g - 2.0 * r[2] - 2.0 * r[4] - r[1], 0.5 - 2.0 * g - 2.0 * r[2] - r[1], 0),
                  (4 if N == 5 else 6, 0.5 + 2.0 * g + 2.0 * r[3] + r[1],
                   0.5 + 2.0 * g + 2.0 * r[3] + 2.0 * r[5] + r[1], 0)]
        if N == 6:
            h = 0.5 - 3.0 * g - r[1] - 2.0 * r[2] - 2.0 * r[4]
            right = 0.5 + 3.0 * g + r[1] + 2.0 * r[3] + 2.0 * r[5]
            rules += [(6, h - r[5], h, 0), (7, right, right + 2.0 * r[6], 0)]
        return rules
    g = 1.0 / float(N)
    centres = [None, None, 0.5 - g, 0.5 + g, 0.5 - 2.0 * g, 0.5 + 2.0 * g]
    last = N if N in (3, 5) else N - 1  # N = 4, 6: the last inclusion sits at the cell edges
    for k in range(2, last + 1):
        rules.append((k + 1, centres[k] - r[k], centres[k] + r[k], 0))
    if last < N:
        rules += [(N + 1, -inf, r[N], 0), (N + 1, 1.0 - r[N], inf, 0)]
    return rules


def grating_1d(period, diameters, lc_bkg, small_space=None, split_touching_incs=False, inclusion_edge='code',
               edge_spacing=False):
    """Uniform P2 mesh of a 1D grating cell with up to six inclusions
    (objects.py:1054-1530 of the reference): nel = round(1/lc_bkg) elements
    on [0, 1]; nodes 2i-1 (left), 2i (mid), 2i+1 (right) of element i; the
    material type of each element from grating_layout_1d's rules.  Returns
    x_arr (2 nel + 1,), table_nod (3, nel) 1-based, type_el (nel,) 1-based
    and the reference's mesh name.

    inclusion_edge: 'code' applies the reference's current tests, whose
    centre rule (objects.py:1502-1503) also takes an element whose left end
    lies up to one element width left of 0.5 - r; 'exact' drops that slack
    (only elements inside the inclusion's interval), the rule with which the
    reference's stacked-gratings golden R/T files (tests/ref/case_4) were
    generated."""
    if inclusion_edge not in ('code', 'exact'):
        raise ValueError("inclusion_edge must be 'code' or 'exact'")
    rules = grating_layout_1d(period, diameters, edge_spacing, split_touching_incs, small_space)
    nel = int(np.round(1.0 / lc_bkg))
    delta_x = (1.0 - 0.0) / nel
    x = np.zeros(2 * nel + 1)
    left = 0.0 + np.arange(nel) * delta_x
    x[0:2 * nel:2] = left
    x[1:2 * nel:2] = left + delta_x / 2.0
    x[2 * nel] = 0.0 + nel * delta_x
    table = np.empty((3, nel), dtype=np.int64, order='F')
    table[0] = 2 * np.arange(1, nel + 1) - 1
    table[1] = 2 * np.arange(1, nel + 1) + 1
    table[2] = 2 * np.arange(1, nel + 1)
    x1, x2 = x[0:2 * nel:2], x[2::2]
    dx = x2 - x1 if inclusion_edge == 'code' else np.zeros(nel)
    type_el = np.ones(nel, dtype=np.int64)
    free = np.ones(nel, dtype=bool)
    for t, lo, hi, slack in rules:
        lo_e = lo - dx if slack else lo
        hit = free & (x1 <= hi) & (lo_e <= x1) & (x2 <= hi) & (lo_e <= x2)
        type_el[hit] = t
        free &= ~hit
    n_inc = max(k + 1 for k in range(len(diameters)) if diameters[k] > 0)
    name = ('1D_' if n_inc <= 2 else '') + '_'.join(_dec_float_str(v) for v in [period] + list(diameters[:n_inc]))
    return x, table, type_el, name
