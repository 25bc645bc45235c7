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

def grating_1d(period, diameters, lc_bkg, small_space=None, split_touching_incs=False, inclusion_edge='code'):
    """Uniform P2 mesh of a 1D grating cell (objects.py:1054-1530 of the
    reference, one or two inclusions): nel = round(1/lc_bkg) elements on
    [0, 1]; nodes 2i-1 (left), 2i (mid), 2i+1 (right) of element i; the
    material type of each element follows from its end points exactly as the
    reference decides it (inclusion a centred, inclusion b at the cell edges
    or touching a).  Returns x_arr (2 nel + 1,), table_nod (3, nel) 1-based,
    type_el (nel,) 1-based and the reference's mesh name.

    inclusion_edge: 'code' applies the reference's current test for the
    centre inclusion (objects.py:1502-1503: an element whose left end lies up
    to one element width left of 0.5 - r still counts as inclusion); 'exact'
    takes only the elements inside [0.5 - r, 0.5 + r], the rule with which
    the reference's stacked-gratings golden R/T files (tests/ref/case_4)
    were generated."""
    ds = [d for d in diameters if d > 0]
    if not ds:
        raise ValueError("Must have at least one grating of nonzero width.")
    if len(ds) > 2:
        raise NotImplementedError("1D gratings with more than two inclusions")
    nel = int(np.round(1.0 / lc_bkg))
    delta_x = (1.0 - 0.0) / nel
    x = np.zeros(2 * nel + 1)
    for i in range(nel):
        x0 = 0.0 + i * delta_x
        x[2 * i] = x0
        x[2 * i + 1] = x0 + delta_x / 2.0
    x[2 * nel] = 0.0 + nel * delta_x
    table = np.empty((3, nel), dtype=np.int64, order='F')
    table[0] = 2 * np.arange(1, nel + 1) - 1
    table[1] = 2 * np.arange(1, nel + 1) + 1
    table[2] = 2 * np.arange(1, nel + 1)
    r1 = ds[0] / (2.0 * period)
    r2 = ds[1] / (2.0 * period) if len(ds) > 1 else 0.0
    if len(ds) > 1:
        small = 0.5 - r1 - r2 if small_space is None else small_space
        large = 0.5 - r1 - r2 if small_space is None else 1.0 - small_space - 2 * r1 - 2 * r2
    if inclusion_edge not in ('code', 'exact'):
        raise ValueError("inclusion_edge must be 'code' or 'exact'")
    type_el = np.ones(nel, dtype=np.int64)
    for i in range(nel):
        x1, x2 = x[2 * i], x[2 * i + 2]
        dx = x2 - x1 if inclusion_edge == 'code' else 0.0
        if (x1 <= 0.5 + r1 and 0.5 - r1 - dx <= x1 and x2 <= 0.5 + r1 and 0.5 - r1 - dx <= x2):
            type_el[i] = 2
        elif len(ds) == 2 and split_touching_incs:
            if ((x1 <= 0.5 + r1 + r2 and 0.5 + r1 - dx <= x1 and x2 <= 0.5 + r1 + r2 and 0.5 + r1 - dx <= x2)
                    or (x1 <= 0.5 - r1 and 0.5 - r1 - r2 - dx <= x1 and x2 <= 0.5 - r1
                        and 0.5 - r1 - r2 - dx <= x2)):
                type_el[i] = 3
        elif len(ds) == 2:
            if (x1 <= 0.5 - r1 - small and x2 <= 0.5 - r1 - small) or (x1 >= 0.5 + large + r1
                                                                        and x2 >= 0.5 + large + r1):
                type_el[i] = 3
    fmt = lambda v: ('%.6f' % v).rstrip('0').rstrip('.') if v != int(v) else str(int(v))
    name = '1D_' + '_'.join(fmt(v) for v in [period] + ds)
    return x, table, type_el, name
