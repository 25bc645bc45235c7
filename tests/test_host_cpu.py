"""Host-side (no GPU) checks of the native library: it loads and exports
every entry point of include/emustack_mi355x.h; its per-mesh preprocessing
reproduces the oracle's DOF numbering and sparse pattern bit for bit; the
symbolic multifrontal structure factors and solves the assembled pencil
when walked in numpy; failure paths return the documented status codes."""
import os
import re

import numpy as np
import pytest
import scipy.linalg as sla

from conftest import MESH_600, MESH_940, ROOT, k_parallel_norm
from oracle import fem2d
from emustack_amd import _lib, mesher


def test_exports_match_header():
    hdr = open(os.path.join(ROOT, 'include', 'emustack_mi355x.h')).read()
    decl = set(re.findall(r'\b(emu_[a-z0-9_]+)\s*\(', hdr))
    L = _lib.lib()
    for name in decl:
        assert hasattr(L, name), name
    assert decl == set(_lib.EXPORTS)


@pytest.mark.parametrize('path', [MESH_600, MESH_940])
def test_mesh_preprocessing_matches_oracle(path):
    mh = _lib.MeshHandle(path)
    m = fem2d.build_mesh(path)
    a = mh.arrays()
    assert (mh.npt, mh.nel, mh.n_ddl, mh.neq, mh.nnz) == (m.npt, m.nel, m.n_ddl, m.neq, m.nnz)
    np.testing.assert_array_equal(a['row_ptr'], m.row_ptr)
    np.testing.assert_array_equal(a['col_ind'], m.col_ind)
    np.testing.assert_array_equal(a['ldof'], m.ldof)
    np.testing.assert_allclose(a['lat'], m.lat.T.ravel())
    assert sorted(a['perm']) == list(range(mh.neq))
    tn, te, x = mh.geometry()
    np.testing.assert_array_equal(tn.T, m.table_nod)
    np.testing.assert_array_equal(te, m.type_el)


def _walk_multifrontal(mh, A, b, inverted=False):
    """Factor and solve A x = b in numpy with the library's symbolic data
    (fronts, update index sets, extend-add maps, entry scatter map).  With
    inverted=True the solves follow the device scheme: each front rewritten
    as Fwd = [L11^-1; -L21 L11^-1], Bwd = [U11^-1, -U11^-1 U12], the pivot
    order applied as a gather, children's update vectors pulled per row."""
    S = mh.symbolic()
    fr = S['fronts']
    arr = mh.arrays()
    pool = np.zeros(mh.pool, complex)
    pool[S['nnz2front']] = A.data
    perm = arr['perm']
    pivs = {}

    def front(t):
        f = fr[t][2]
        return pool[S['foff'][t]:S['foff'][t] + f * f].reshape(f, f, order='F')
    for lvl in range(fr[:, 5].max(), -1, -1):
        for t in np.nonzero(fr[:, 5] == lvl)[0]:
            s, u, f, _, par = fr[t][:5]
            F = front(t).copy()
            piv = np.arange(s)
            for k in range(s):
                p = k + np.argmax(np.abs(F[k:s, k].real) + np.abs(F[k:s, k].imag))
                piv[k] = p
                F[[k, p], :] = F[[p, k], :]
                F[k + 1:, k] /= F[k, k]
                F[k + 1:, k + 1:] -= np.outer(F[k + 1:, k], F[k, k + 1:])
            pivs[t] = piv
            pool[S['foff'][t]:S['foff'][t] + f * f] = F.reshape(-1, order='F')
            if par >= 0 and u > 0:
                fp = fr[par][2]
                P = front(par).copy()
                mp = S['pmap'][S['uoff'][t]:S['uoff'][t] + u]
                P[np.ix_(mp, mp)] += F[s:, s:]
                pool[S['foff'][par]:S['foff'][par] + fp * fp] = P.reshape(-1, order='F')
    bn = b[perm]
    if inverted:
        return _solve_inverted(mh, S, fr, front, pivs, bn, perm)
    y = np.zeros(mh.neq, complex)
    wv = {}
    for lvl in range(fr[:, 5].max(), -1, -1):
        for t in np.nonzero(fr[:, 5] == lvl)[0]:
            s, u, f, b0, par, _, c0, c1 = fr[t]
            F = front(t)
            r = np.zeros(f, complex)
            r[:s] = bn[b0:b0 + s]
            for c in (c0, c1):
                if c >= 0:
                    uc = fr[c][1]
                    r[S['pmap'][S['uoff'][c]:S['uoff'][c] + uc]] += wv[c]
            for k in range(s):
                p = pivs[t][k]
                r[[k, p]] = r[[p, k]]
            for k in range(s):
                r[k + 1:] -= F[k + 1:, k] * r[k]
            y[b0:b0 + s] = r[:s]
            wv[t] = r[s:].copy()
    x = np.zeros(mh.neq, complex)
    for lvl in range(0, fr[:, 5].max() + 1):
        for t in np.nonzero(fr[:, 5] == lvl)[0]:
            s, u, f, b0 = fr[t][:4]
            F = front(t)
            ui = S['uidx'][S['uoff'][t]:S['uoff'][t] + u]
            x[b0:b0 + s] = sla.solve_triangular(F[:s, :s], y[b0:b0 + s] - F[:s, s:] @ x[ui], lower=False)
    xo = np.zeros_like(x)
    xo[perm] = x
    return xo


def _solve_inverted(mh, S, fr, front, pivs, bn, perm):
    nf = len(fr)
    fvoff = np.concatenate([[0], np.cumsum(fr[:, 2])])
    cmap = -np.ones((fvoff[-1], 2), np.int64)
    for t in range(nf):
        s, u, f, b0, par, _, c0, c1 = fr[t]
        for slot, c in enumerate((c0, c1)):
            if c >= 0:
                uc = fr[c][1]
                mp = S['pmap'][S['uoff'][c]:S['uoff'][c] + uc]
                cmap[fvoff[t] + mp, slot] = S['uoff'][c] + np.arange(uc)
    w = np.zeros(max(mh.nw, 1), complex)
    y = np.zeros(mh.neq, complex)
    eye = np.eye
    fwd, bwd, sig = {}, {}, {}
    for t in range(nf):
        s, u, f = fr[t][:3]
        F = front(t)
        L11 = np.tril(F[:s, :s], -1) + eye(s)
        Li = sla.solve_triangular(L11, eye(s), lower=True, unit_diagonal=True)
        Ui = sla.solve_triangular(np.triu(F[:s, :s]), eye(s), lower=False)
        fwd[t] = np.vstack([Li, -F[s:, :s] @ Li])
        bwd[t] = np.hstack([Ui, -Ui @ F[:s, s:]])
        g = np.arange(s)
        for k in range(s):
            p = pivs[t][k]
            g[[k, p]] = g[[p, k]]
        sig[t] = g

    def pulled(t, rows):
        v = np.zeros(len(rows), complex)
        for slot in range(2):
            ci = cmap[fvoff[t] + rows, slot]
            v[ci >= 0] += w[ci[ci >= 0]]
        return v
    for lvl in range(fr[:, 5].max(), -1, -1):
        for t in np.nonzero(fr[:, 5] == lvl)[0]:
            s, u, f, b0 = fr[t][:4]
            xin = bn[b0 + sig[t]] + pulled(t, sig[t])
            out = fwd[t] @ xin
            y[b0:b0 + s] = out[:s]
            w[S['uoff'][t]:S['uoff'][t] + u] = out[s:] + pulled(t, np.arange(s, f))
    x = np.zeros(mh.neq, complex)
    for lvl in range(0, fr[:, 5].max() + 1):
        for t in np.nonzero(fr[:, 5] == lvl)[0]:
            s, u, f, b0 = fr[t][:4]
            ui = S['uidx'][S['uoff'][t]:S['uoff'][t] + u]
            x[b0:b0 + s] = bwd[t] @ np.concatenate([y[b0:b0 + s], x[ui]])
    xo = np.zeros_like(x)
    xo[perm] = x
    return xo


def test_symbolic_multifrontal_solves_pencil():
    mh = _lib.MeshHandle(MESH_600)
    m = fem2d.build_mesh(MESH_600)
    wln = 500 / 600
    k0 = 2 * np.pi / wln
    eps = np.array([1.0, (3.5 + 0.01j) ** 2])
    A, _ = fem2d.assemble(m, np.ones(2, complex), eps * k0 ** 2, 1.1 * 3.5 ** 2 * k0 ** 2,
                          k_parallel_norm(500, 0, 0, 600))
    A.sort_indices()
    rng = np.random.default_rng(5)
    b = rng.standard_normal(m.neq) + 1j * rng.standard_normal(m.neq)
    x = _walk_multifrontal(mh, A, b)
    assert np.linalg.norm(A @ x - b) / np.linalg.norm(b) < 1e-9
    xi = _walk_multifrontal(mh, A, b, inverted=True)
    assert np.linalg.norm(A @ xi - b) / np.linalg.norm(b) < 1e-9
    # every front variable is owned exactly once, update sets are above the front
    fr = mh.symbolic()['fronts']
    assert fr[:, 0].sum() == mh.neq
    assert (fr[:, 4] == -1).sum() == 1


def test_generated_meshes_are_periodic(tmp_path):
    for ntri, a, b, rot in [(600, 120., 120., 0.), (900, 200., 90., 35.)]:
        p = str(tmp_path / ('g%d.mail' % ntri))
        mesher.write_mail(p, mesher.ellipse_cell_ntri(600., 600., a, b, ntri, rotation_deg=rot, seed=4))
        mh = _lib.MeshHandle(p)
        m = fem2d.build_mesh(p)
        assert mh.neq == m.neq == 12 * mh.nel
        assert mh.nnz == 411 * mh.nel
        assert abs(mh.nel - ntri) < 0.15 * ntri


def test_non_periodic_mesh_is_rejected(tmp_path):
    x, tn, table, te = fem2d.read_mail(MESH_600)
    x = x.copy()
    i = np.nonzero(tn == 3)[0][3]
    x[i, 1] += 1e-3   # break the left/right pairing of one node
    p = str(tmp_path / 'bad.mail')
    mesher.write_mail(p, (x, tn, table, te))
    with pytest.raises(_lib.EmuError) as e:
        _lib.MeshHandle(p)
    assert e.value.code == -2
    with pytest.raises(RuntimeError):
        fem2d.build_mesh(p)


def test_missing_mesh_file():
    with pytest.raises(_lib.EmuError) as e:
        _lib.MeshHandle('/nonexistent/mesh.mail')
    assert e.value.code == -2


def test_simmo_rejects_too_few_modes(tmp_path):
    from emustack_amd import objects, materials
    ns = objects.NanoStruct('2D_array', 600, 120, make_mesh_now=False, mesh_file=MESH_600,
                            inclusion_a=materials.Material(3.5))
    light = objects.Light(500, max_order_PWs=2, theta=0, phi=0)
    from emustack_amd.mode_calcs import Simmo
    with pytest.raises(AssertionError):
        Simmo(ns, light).fem_inputs(num_BMs=26)
