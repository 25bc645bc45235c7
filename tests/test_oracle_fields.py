"""Pins of the oracle's plotting-field restatements against what the
mathematics fixes (no GPU):

* mode_energy (mode_energy.f): the 16-point rule applied to |P2 interpolant|^2
  equals the closed-form P2 mass matrix (area / 180 times the integer matrix
  below), so the per-element energies must equal s^H M s with the field's
  node values s; E_z is scaled by beta, the fourth row sums the core material
  (py_calc_modes.f:546-551 picks it) and the rows divide by the x + y + z total.
* array_sol (fields_at_nodes): E_z at the P2 nodes comes from the P3 DOFs, so
  at a vertex it equals that vertex's P3 value and at an edge midpoint the
  cubic through the edge's four P3 nodes, (-1, -1, 9, 9) / 16 of the two ends
  and the two edge nodes stored at sol[2, 6..11].
* array_sol_P2_1d (sol_p2_1d): the mid-node value of the y / z components is
  the cubic through the element's four P3 values (ends, 1/3, 2/3) at 1/2.
"""
import numpy as np
import pytest

from conftest import MESH_600
from oracle import fem1d, fem2d

# P2 mass matrix on a triangle of area A, times 180 / A; nodes: vertices 1-3,
# then the midpoints of edges 1-2, 2-3, 3-1
_K = np.array([[6, -1, -1, 0, -4, 0],
               [-1, 6, -1, 0, 0, -4],
               [-1, -1, 6, -4, 0, 0],
               [0, 0, -4, 32, 16, 16],
               [-4, 0, 0, 16, 32, 16],
               [0, -4, 0, 16, 16, 32]], float)


@pytest.fixture(scope='module')
def mesh600():
    return fem2d.build_mesh(MESH_600)


def test_core_type_rule():
    assert fem2d.core_type([1.0, 12.0]) == 2
    assert fem2d.core_type([12.0 + 1j, 2.0]) == 1
    assert fem2d.core_type([4.0, 4.0 + 1j]) == 2  # ties go to material 2


def test_mode_energy_matches_p2_mass_matrix(mesh600):
    m = mesh600
    rng = np.random.default_rng(3)
    nval = 5
    sol = rng.standard_normal((3, 13, nval, m.nel)) + 1j * rng.standard_normal((3, 13, nval, m.nel))
    beta = rng.standard_normal(nval) + 1j * rng.standard_normal(nval)
    eps = np.array([2.25 + 0.1j, 12.0 + 0.5j])
    pol = fem2d.mode_energy(m, sol, beta, eps)
    _, _, det, _ = fem2d.element_geometry(m)
    area = np.abs(det) / 2.0
    ew = area / 180.0 * eps[m.type_el - 1]
    s = sol[:, :6].copy()
    s[2] *= beta[:, None]
    quad = np.einsum('cine,ij,cjne->cne', s.conj(), _K, s).real  # 3, nval, nel
    parts = np.einsum('e,cne->cn', ew, quad)
    core = m.type_el == 2
    ref4 = np.einsum('e,ne->n', ew[core], quad[:, :, core].sum(axis=0))
    tot = parts.sum(axis=0)
    np.testing.assert_allclose(pol[:3], parts / tot, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(pol[3], ref4 / tot, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(pol[:3].sum(axis=0), 1.0, rtol=0, atol=1e-13)
    # a uniform field in one uniform material: the fractions are the squared
    # components and the core row is the core material's area fraction
    one = np.zeros((3, 13, 1, m.nel), complex)
    one[:, :6, 0, :] = np.array([1.0, 2.0j, 0.5])[:, None, None]
    pol1 = fem2d.mode_energy(m, one, np.array([2.0]), np.array([1.0, 1.0 + 0.5]))
    c2 = np.array([1.0, 4.0, 1.0])
    w = np.where(m.type_el == 2, 1.5, 1.0) * area
    np.testing.assert_allclose(pol1[:3, 0], c2 / c2.sum(), rtol=1e-12)
    np.testing.assert_allclose(pol1[3, 0], w[core].sum() / w.sum(), rtol=1e-12)


def test_array_sol_ez_is_the_p3_interpolant(mesh600):
    m = mesh600
    rng = np.random.default_rng(4)
    V = rng.standard_normal((m.neq, 3)) + 1j * rng.standard_normal((m.neq, 3))
    sol, zmax = fem2d.fields_at_nodes(m, V, np.array([0.3, -0.7]))
    ez = sol[2]
    # vertices: phi3 is nodal there, so E_z at vertex i is a value the edge
    # interpolants below share
    w = np.array([-1.0, -1.0, 9.0, 9.0]) / 16.0
    for mid, (a, b), (p, q) in ((3, (0, 1), (6, 7)), (4, (1, 2), (8, 9)), (5, (2, 0), (10, 11))):
        cubic = w[0] * ez[a] + w[1] * ez[b] + w[2] * ez[p] + w[3] * ez[q]
        np.testing.assert_allclose(ez[mid], cubic, rtol=0, atol=1e-12 * np.abs(ez).max())
    # the transverse components are not carried at the P3-only nodes
    assert np.all(sol[:2, 6:] == 0)
    # normalisation value: the first largest |component| over the P2 nodes
    blk = np.abs(sol[:, :6]).max(axis=(0, 1, 3))
    np.testing.assert_allclose(np.abs(zmax), blk, rtol=1e-15)


def test_sol_p2_1d_mid_node_is_the_p3_interpolant():
    rng = np.random.default_rng(5)
    nel, nval = 7, 3
    coef = rng.standard_normal((2, 4, nval, nel)) + 1j * rng.standard_normal((2, 4, nval, nel))

    def cubic(c, t):
        return c[0] + c[1] * t + c[2] * t ** 2 + c[3] * t ** 3

    sol = np.zeros((nel, 11, nval), complex)
    sol[:, 0:3] = np.transpose(rng.standard_normal((3, nval, nel)), (2, 0, 1))
    for c, base in ((0, 3), (1, 7)):
        for j, t in enumerate((0.0, 1.0, 1.0 / 3.0, 2.0 / 3.0)):
            sol[:, base + j] = cubic(coef[c], t).T
    out = fem1d.sol_p2_1d(sol)
    assert out.shape == (3, 3, nval, nel)
    np.testing.assert_array_equal(out[0], np.transpose(sol[:, 0:3], (1, 2, 0)))
    for c in (0, 1):
        np.testing.assert_allclose(out[c + 1, 0], cubic(coef[c], 0.0), rtol=1e-14)
        np.testing.assert_allclose(out[c + 1, 1], cubic(coef[c], 1.0), rtol=1e-14)
        np.testing.assert_allclose(out[c + 1, 2], cubic(coef[c], 0.5), rtol=0, atol=1e-13)
