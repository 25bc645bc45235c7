// Per-mesh preprocessing of a 1D grating (uniform-period cell [0, 1] cut
// into P2 elements): DOF numbering, periodic closure, sparse pattern, and
// the reference-element integral tables.
//
// Follows backend/fortran/periodic_cond_1d.f:1-220 (11 local DOFs per
// element: discontinuous P2 x-component, continuous P3 y- and
// z-components; 9 new equations per element; the last element's right-end
// P3 nodes close onto the first element's left-end nodes) and the element
// integrals of matrix_{mxx,kxy,kyx,kyy,myy}_1d.f, computed here by
// Gauss-Legendre quadrature of the Lagrange bases.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "emu_host.h"

namespace emu {

namespace {

// Lagrange basis on `nodes` (size n) at t: values and d/dt
void lagrange(const double *nodes, int n, double t, double *val, double *der) {
  for (int i = 0; i < n; ++i) {
    double den = 1.0, v = 1.0, d = 0.0;
    for (int j = 0; j < n; ++j)
      if (j != i) den *= nodes[i] - nodes[j];
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      v *= t - nodes[j];
      double p = 1.0;
      for (int k = 0; k < n; ++k)
        if (k != i && k != j) p *= t - nodes[k];
      d += p;
    }
    val[i] = v / den;
    der[i] = d / den;
  }
}

// 8-point Gauss-Legendre rule on [0, 1] (exact to degree 15)
void gauss8(double *t, double *w) {
  const double z[4] = {0.1834346424956498, 0.5255324099163290, 0.7966664774136267, 0.9602898564975363};
  const double c[4] = {0.3626837833783620, 0.3137066458778873, 0.2223810344533745, 0.1012285362903763};
  for (int i = 0; i < 4; ++i) {
    t[2 * i] = 0.5 * (1.0 - z[i]);
    t[2 * i + 1] = 0.5 * (1.0 + z[i]);
    w[2 * i] = w[2 * i + 1] = 0.5 * c[i];
  }
}

}  // namespace

void element_tables_1d(double *tab) {
  // tab: Mxx (3x3, per unit length), Kxy (3x4), Kyy (4x4, times length),
  // Myy (4x4, per unit length); row-major [i][j]
  const double p2n[3] = {0.0, 1.0, 0.5};
  const double p3n[4] = {0.0, 1.0, 1.0 / 3.0, 2.0 / 3.0};
  double t[8], w[8];
  gauss8(t, w);
  std::fill(tab, tab + kTab1d, 0.0);
  for (int q = 0; q < 8; ++q) {
    double v2[3], d2[3], v3[4], d3[4];
    lagrange(p2n, 3, t[q], v2, d2);
    lagrange(p3n, 4, t[q], v3, d3);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) tab[kMxx + 3 * i + j] += w[q] * v2[i] * v2[j];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 4; ++j) tab[kKxy + 4 * i + j] += w[q] * v2[i] * d3[j];
    for (int i = 0; i < 4; ++i)
      for (int j = 0; j < 4; ++j) {
        tab[kKyy + 4 * i + j] += w[q] * d3[i] * d3[j];
        tab[kMyy + 4 * i + j] += w[q] * v3[i] * v3[j];
      }
  }
}

void build_mesh_1d(Mesh &m, int64_t npt, int64_t nel, const int64_t *table_nod, const int64_t *type_el,
                   const double *x) {
  if (nel < 2 || npt != 2 * nel + 1)
    throw std::runtime_error("1D mesh: npt must equal 2 nel + 1 (py_calc_modes_1d.f)");
  m.dim = 1;
  m.path = "<1D grating>";
  m.npt = npt;
  m.nel = nel;
  m.neq = 9 * nel;
  m.n_ddl = 9 * nel + 2;
  m.x.assign(x, x + npt);
  m.table_nod.assign(table_nod, table_nod + 3 * nel);
  m.type_el.assign(type_el, type_el + nel);
  m.h1d.resize(nel);
  m.xd1d.resize((size_t)11 * nel);
  m.ldof.resize((size_t)11 * nel);
  for (int64_t e = 0; e < nel; ++e) {
    if (m.type_el[e] < 1) throw std::runtime_error("1D mesh: type_el(i) < 1");
    int64_t a = m.table_nod[3 * e], b = m.table_nod[3 * e + 1];
    if (a < 1 || a > npt || b < 1 || b > npt) throw std::runtime_error("1D mesh: node index out of range");
    double xmin = x[a - 1], xmax = x[b - 1], dx = xmax - xmin;
    if (!(dx > 0)) throw std::runtime_error("1D mesh: element with xmax <= xmin");
    m.h1d[e] = dx;
    const double xs[11] = {xmin, xmax, xmin + dx / 2.0, xmin, xmax, xmin + dx / 3.0, xmin + 2.0 * dx / 3.0,
                           xmin, xmax, xmin + dx / 3.0, xmin + 2.0 * dx / 3.0};
    std::copy(xs, xs + 11, &m.xd1d[(size_t)11 * e]);
    const int32_t base = (int32_t)(9 * e), nb = (int32_t)(e + 1 < nel ? 9 * (e + 1) : 0);
    const int32_t eqs[11] = {base, base + 1, base + 2, base + 3, nb + 3, base + 4, base + 5,
                             base + 6, nb + 6, base + 7, base + 8};
    std::copy(eqs, eqs + 11, &m.ldof[(size_t)11 * e]);
  }
  // Bloch phase distance of the wrapped DOFs (asmbly_1d.f:108-122)
  m.period_x = m.xd1d[(size_t)11 * (nel - 1) + 4] - m.xd1d[3];
  // pattern: union of the element cliques
  const int64_t neq = m.neq;
  std::vector<std::vector<int32_t>> rows(neq);
  for (int64_t e = 0; e < nel; ++e)
    for (int i = 0; i < 11; ++i)
      for (int j = 0; j < 11; ++j) rows[m.ldof[11 * e + i]].push_back(m.ldof[11 * e + j]);
  m.row_ptr.assign(neq + 1, 0);
  for (int64_t i = 0; i < neq; ++i) {
    auto &r = rows[i];
    std::sort(r.begin(), r.end());
    r.erase(std::unique(r.begin(), r.end()), r.end());
    m.row_ptr[i + 1] = m.row_ptr[i] + (int64_t)r.size();
  }
  m.nnz = m.row_ptr[neq];
  m.col_ind.resize(m.nnz);
  for (int64_t i = 0; i < neq; ++i) std::copy(rows[i].begin(), rows[i].end(), m.col_ind.begin() + m.row_ptr[i]);
  // per-nnz contributor lists (e * 121 + i * 11 + j)
  m.slot.assign((size_t)121 * nel, 0);
  std::vector<int32_t> cnt(m.nnz + 1, 0);
  for (int64_t e = 0; e < nel; ++e)
    for (int i = 0; i < 11; ++i) {
      int32_t r = m.ldof[11 * e + i];
      const int32_t *cb = m.col_ind.data() + m.row_ptr[r];
      const int32_t *ce = m.col_ind.data() + m.row_ptr[r + 1];
      for (int j = 0; j < 11; ++j) {
        const int32_t *it = std::lower_bound(cb, ce, m.ldof[11 * e + j]);
        int32_t k = (int32_t)(it - m.col_ind.data());
        m.slot[(size_t)121 * e + 11 * i + j] = k;
        cnt[k + 1]++;
      }
    }
  m.cptr.assign(m.nnz + 1, 0);
  for (int64_t k = 0; k < m.nnz; ++k) m.cptr[k + 1] = m.cptr[k] + cnt[k + 1];
  m.citem.assign(m.cptr[m.nnz], 0);
  std::vector<int32_t> fill(m.cptr.begin(), m.cptr.end() - 1);
  for (int64_t e = 0; e < nel; ++e)
    for (int ij = 0; ij < 121; ++ij) {
      int32_t k = m.slot[(size_t)121 * e + ij];
      m.citem[fill[k]++] = (int32_t)(121 * e + ij);
    }
  element_tables_1d(m.tab1d);
}

}  // namespace emu
