// Per-mesh preprocessing: .mail reader, DOF topology, periodic merge,
// equation numbering and sparse pattern.
//
// Follows the reference's per-mesh steps of calc_modes_2d
// (backend/fortran/py_calc_modes.f:297-522): geometry.f, lattice_vec.f,
// list_edge.f, list_node_P3.f, type_node_edge_face.f, get_coord_p3.f,
// periodic_node.f, periodic_N_E_F.f, periodic_cond.f, csr_length.f.
// The reference redoes all of this for every wavelength; here it runs once
// per mesh and the result is cached with the mesh handle.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>

#include "emu_host.h"

namespace emu {

static void read_mail(Mesh &m, const std::string &path) {
  std::ifstream f(path);
  if (!f) throw std::runtime_error("cannot open mesh file " + path);
  std::stringstream ss;
  ss << f.rdbuf();
  std::string s = ss.str();
  for (auto &c : s)
    if (c == 'D' || c == 'd') c = 'E';  // Fortran exponent letter
  const char *p = s.c_str();
  char *end = nullptr;
  auto next_d = [&]() {
    double v = std::strtod(p, &end);
    if (end == p) throw std::runtime_error("malformed mesh file " + path);
    p = end;
    return v;
  };
  m.npt = (int64_t)next_d();
  m.nel = (int64_t)next_d();
  if (m.npt <= 0 || m.nel <= 0) throw std::runtime_error("empty mesh " + path);
  m.x.resize(2 * m.npt);
  m.type_nod.resize(m.npt);
  for (int64_t i = 0; i < m.npt; ++i) {
    next_d();
    m.x[2 * i] = next_d();
    m.x[2 * i + 1] = next_d();
    m.type_nod[i] = (int64_t)next_d();
  }
  m.table_nod.resize(6 * m.nel);
  m.type_el.resize(m.nel);
  for (int64_t i = 0; i < m.nel; ++i) {
    next_d();
    for (int j = 0; j < 6; ++j) {
      int64_t v = (int64_t)next_d();
      if (v < 1 || v > m.npt) throw std::runtime_error("mesh: node index out of range");
      m.table_nod[6 * i + j] = v;
    }
    int64_t t = (int64_t)next_d();
    if (t < 1) throw std::runtime_error("geometry: type_el(i) < 1");
    m.type_el[i] = t;
  }
}

// lattice_vec.f:4-91
static void lattice_vec(Mesh &m) {
  const double tol = 1e-6;
  double ymin = m.x[1], ymax = m.x[1];
  for (int64_t i = 0; i < m.npt; ++i) {
    ymin = std::min(ymin, m.x[2 * i + 1]);
    ymax = std::max(ymax, m.x[2 * i + 1]);
  }
  double bmin = 1e300, bmax = -1e300, tmin = 1e300;
  for (int64_t i = 0; i < m.npt; ++i) {
    double xx = m.x[2 * i], yy = m.x[2 * i + 1];
    if (std::fabs(yy - ymin) < tol) { bmin = std::min(bmin, xx); bmax = std::max(bmax, xx); }
    if (std::fabs(yy - ymax) < tol) tmin = std::min(tmin, xx);
  }
  m.lat[0] = bmax - bmin;
  m.lat[1] = 0.0;
  m.lat[2] = tmin - bmin;
  m.lat[3] = ymax - ymin;
}

// Shared pairing rule of periodic_node.f:58-106 / periodic_N_E_F.f:36-90.
static std::vector<int64_t> periodic_pairs(const double *pos, const std::vector<char> &active,
                                           const int64_t *cls, const double lat[4], int64_t n) {
  std::vector<int64_t> ip(n, 0);
  std::vector<int64_t> idx;
  for (int64_t i = 0; i < n; ++i)
    if (active[i]) idx.push_back(i);
  double sh[9][2];
  int k = 0;
  for (int ix = -1; ix <= 1; ++ix)
    for (int iy = -1; iy <= 1; ++iy) {
      sh[k][0] = ix * lat[0] + iy * lat[2];
      sh[k][1] = ix * lat[1] + iy * lat[3];
      ++k;
    }
  const double tol = 1e-6;
  for (size_t a = 0; a + 1 < idx.size(); ++a) {
    int64_t i = idx[a];
    for (size_t b = a + 1; b < idx.size(); ++b) {
      int64_t j = idx[b];
      if (ip[j] != 0) continue;
      if (cls && cls[j] != cls[i]) continue;
      double dx = pos[2 * j] - pos[2 * i], dy = pos[2 * j + 1] - pos[2 * i + 1];
      bool hit = false;
      for (int s = 0; s < 9 && !hit; ++s)
        hit = (std::fabs(dx - sh[s][0]) + std::fabs(dy - sh[s][1])) < tol;
      if (!hit) continue;
      if (ip[i] == 0) {
        ip[i] = i + 1;
        ip[j] = i + 1;
      } else {
        ip[j] = ip[i];
      }
    }
  }
  return ip;
}

void build_mesh(Mesh &m, const std::string &path) {
  m.path = path;
  read_mail(m, path);
  lattice_vec(m);
  const int64_t nel = m.nel, npt = m.npt;
  auto tn = [&](int64_t e, int j) { return m.table_nod[6 * e + j] - 1; };
  m.T.assign(14 * nel, 0);
  auto T = [&](int64_t e, int j) -> int64_t & { return m.T[14 * e + j]; };
  // list_face / list_edge (list_edge.f:37-69)
  std::vector<int64_t> visite(npt, 0);
  int64_t n_edge = 0;
  for (int64_t i = 0; i < nel; ++i) {
    T(i, 0) = i + 1;
    for (int j = 3; j < 6; ++j) {
      int64_t j1 = tn(i, j);
      int64_t k = visite[j1];
      if (k == 0) {
        ++n_edge;
        visite[j1] = n_edge;
        T(i, j - 2) = n_edge + nel;
      } else {
        T(i, j - 2) = k + nel;
      }
    }
  }
  // list_node_P3 (list_node_P3.f:29-128)
  std::fill(visite.begin(), visite.end(), 0);
  const int64_t mm2 = n_edge + nel;
  int64_t nn = 0;
  for (int64_t iel = 0; iel < nel; ++iel) {
    for (int inod = 0; inod < 3; ++inod) {
      int64_t k = tn(iel, inod);
      if (visite[k] == 0) {
        visite[k] = iel + 1;
        ++nn;
        T(iel, inod + 4) = nn + mm2;
      } else {
        int64_t iel2 = visite[k] - 1;
        int inod2 = -1;
        for (int j = 0; j < 3; ++j)
          if (tn(iel2, j) == k) inod2 = j;
        if (inod2 < 0) throw std::runtime_error("list_node_P3: problem with a vertex");
        T(iel, inod + 4) = T(iel2, inod2 + 4);
      }
    }
    for (int inod = 3; inod < 6; ++inod) {
      int64_t k = tn(iel, inod);
      int a = inod - 3, b = (inod - 2) % 3;
      int64_t v1a = tn(iel, a), v1b = tn(iel, b);
      int base = 7 + 2 * (inod - 3);
      if (visite[k] == 0) {
        visite[k] = iel + 1;
        for (int jj = 0; jj < 2; ++jj) T(iel, base + jj) = (++nn) + mm2;
      } else {
        int64_t iel2 = visite[k] - 1;
        int inod2 = -1;
        for (int j = 3; j < 6; ++j)
          if (tn(iel2, j) == k) inod2 = j;
        if (inod2 < 0) throw std::runtime_error("list_node_P3: problem with a node");
        int a2 = inod2 - 3, b2 = (inod2 - 2) % 3;
        int64_t v2a = tn(iel2, a2), v2b = tn(iel2, b2);
        int base2 = 7 + 2 * (inod2 - 3);
        for (int jj = 0; jj < 2; ++jj) {
          if (v2a == v1a && v2b == v1b)
            T(iel, base + jj) = T(iel2, base2 + jj);
          else if (v2a == v1b && v2b == v1a)
            T(iel, base + (1 - jj)) = T(iel2, base2 + jj);
          else
            throw std::runtime_error("list_node_P3: check the edge endpoints");
        }
      }
    }
    T(iel, 13) = (++nn) + mm2;
  }
  m.n_edge = n_edge;
  m.npt_p3 = nn;
  m.n_ddl = n_edge + nel + nn;
  const int64_t n_ddl = m.n_ddl;

  // type_node_edge_face.f + get_coord_p3.f
  m.xE.assign(2 * n_ddl, 0.0);
  m.tE.assign(2 * n_ddl, 0);
  auto X = [&](int64_t node, int c) { return m.x[2 * node + c]; };
  for (int64_t e = 0; e < nel; ++e) {
    int64_t f = T(e, 0) - 1;
    for (int c = 0; c < 2; ++c)
      m.xE[2 * f + c] = (X(tn(e, 0), c) + X(tn(e, 1), c) + X(tn(e, 2), c)) / 3.0;
    m.tE[2 * f] = 0;
    m.tE[2 * f + 1] = 2;
    for (int j = 0; j < 3; ++j) {
      int64_t ed = T(e, j + 1) - 1;
      int64_t mid = tn(e, j + 3);
      for (int c = 0; c < 2; ++c) m.xE[2 * ed + c] = X(mid, c);
      m.tE[2 * ed] = m.type_nod[mid];
      m.tE[2 * ed + 1] = 1;
    }
    for (int j = 0; j < 3; ++j) {
      int64_t v = T(e, j + 4) - 1;
      for (int c = 0; c < 2; ++c) m.xE[2 * v + c] = X(tn(e, j), c);
      m.tE[2 * v] = m.type_nod[tn(e, j)];
      m.tE[2 * v + 1] = 0;
    }
  }
  {
    std::vector<char> seen(npt, 0);
    const int ipn[3] = {1, 2, 0};
    for (int64_t e = 0; e < nel; ++e) {
      for (int inod = 3; inod < 6; ++inod) {
        int64_t k = tn(e, inod);
        if (seen[k]) continue;
        seen[k] = 1;
        int64_t p1 = tn(e, inod - 3), p2 = tn(e, ipn[inod - 3]);
        double dx = (X(p2, 0) - X(p1, 0)) / 3.0, dy = (X(p2, 1) - X(p1, 1)) / 3.0;
        for (int jj = 0; jj < 2; ++jj) {
          int64_t d = T(e, 7 + 2 * (inod - 3) + jj) - 1;
          m.xE[2 * d] = X(p1, 0) + (jj + 1) * dx;
          m.xE[2 * d + 1] = X(p1, 1) + (jj + 1) * dy;
          m.tE[2 * d] = m.type_nod[k];
          m.tE[2 * d + 1] = 0;
        }
      }
      int64_t d = T(e, 13) - 1;
      double t3 = 1.0 / 3.0;
      m.xE[2 * d] = X(tn(e, 0), 0) * t3 + X(tn(e, 1), 0) * t3 + X(tn(e, 2), 0) * t3;
      m.xE[2 * d + 1] = X(tn(e, 0), 1) * t3 + X(tn(e, 1), 1) * t3 + X(tn(e, 2), 1) * t3;
      m.tE[2 * d] = 0;
      m.tE[2 * d + 1] = 0;
    }
  }

  // periodic merge
  {
    std::vector<char> act(npt);
    for (int64_t i = 0; i < npt; ++i) act[i] = m.type_nod[i] != 0;
    m.ip_period_N = periodic_pairs(m.x.data(), act, nullptr, m.lat, npt);
    for (int64_t i = 0; i < npt; ++i)
      if (act[i] && m.ip_period_N[i] == 0)
        throw std::runtime_error("periodic_node: mesh is not periodic");
    std::vector<char> actE(n_ddl);
    std::vector<int64_t> dim(n_ddl);
    for (int64_t i = 0; i < n_ddl; ++i) {
      actE[i] = m.tE[2 * i] != 0;
      dim[i] = m.tE[2 * i + 1];
    }
    m.ip_period_E = periodic_pairs(m.xE.data(), actE, dim.data(), m.lat, n_ddl);
    for (int64_t i = 0; i < n_ddl; ++i)
      if (actE[i] && m.ip_period_E[i] == 0)
        throw std::runtime_error("periodic_N_E_F: DOF set is not periodic");
  }
  // periodic_cond.f:103-175: interior entities first, then representatives
  m.ineq.assign(3 * n_ddl, 0);
  int64_t neq = 0;
  for (int pas = 0; pas < 2; ++pas)
    for (int64_t i = 0; i < n_ddl; ++i) {
      int64_t b = m.tE[2 * i], dim = m.tE[2 * i + 1];
      if (pas == 0 && b != 0) continue;
      if (pas == 1 && (b == 0 || m.ip_period_E[i] != i + 1)) continue;
      if (dim == 1 || dim == 2) {
        for (int q = 0; q < 3; ++q) m.ineq[3 * i + q] = ++neq;
      } else {
        m.ineq[3 * i] = ++neq;
      }
    }
  for (int64_t i = 0; i < n_ddl; ++i)
    if (m.tE[2 * i] != 0 && m.ip_period_E[i] != i + 1)
      for (int q = 0; q < 3; ++q) m.ineq[3 * i + q] = m.ineq[3 * (m.ip_period_E[i] - 1) + q];
  m.neq = neq;

  // element -> 22 equation numbers
  m.ldof.assign(22 * nel, 0);
  for (int64_t e = 0; e < nel; ++e) {
    for (int jt = 0; jt < 4; ++jt)
      for (int q = 0; q < 3; ++q) {
        int64_t v = m.ineq[3 * (T(e, jt) - 1) + q] - 1;
        if (v < 0) throw std::runtime_error("periodic_cond: missing equation number");
        m.ldof[22 * e + 3 * jt + q] = (int32_t)v;
      }
    for (int jt = 0; jt < 10; ++jt) {
      int64_t v = m.ineq[3 * (T(e, 4 + jt) - 1)] - 1;
      if (v < 0) throw std::runtime_error("periodic_cond: missing equation number");
      m.ldof[22 * e + 12 + jt] = (int32_t)v;
    }
  }
  // edge orientation by periodic-representative node number (basis_ls.f:45-86)
  m.orient.assign(6 * nel, 0);
  const int ends[3][2] = {{0, 1}, {1, 2}, {0, 2}};
  for (int64_t e = 0; e < nel; ++e) {
    int64_t np[3];
    for (int j = 0; j < 3; ++j) {
      int64_t nd = m.table_nod[6 * e + j];
      int64_t r = m.ip_period_N[nd - 1];
      np[j] = r != 0 ? r : nd;
    }
    for (int ed = 0; ed < 3; ++ed) {
      int a = ends[ed][0], b = ends[ed][1];
      if (np[a] == np[b]) throw std::runtime_error("basis_ls: degenerate edge");
      bool sw = np[b] < np[a];
      m.orient[6 * e + 2 * ed] = (int8_t)(sw ? b : a);
      m.orient[6 * e + 2 * ed + 1] = (int8_t)(sw ? a : b);
    }
  }
  // Bloch phase offsets (asmbly.f:121-146)
  m.dphase.assign(28 * nel, 0.0);
  for (int64_t e = 0; e < nel; ++e)
    for (int j = 0; j < 14; ++j) {
      int64_t ent = T(e, j) - 1;
      int64_t r = m.ip_period_E[ent];
      if (r != 0) {
        m.dphase[28 * e + 2 * j] = m.xE[2 * ent] - m.xE[2 * (r - 1)];
        m.dphase[28 * e + 2 * j + 1] = m.xE[2 * ent + 1] - m.xE[2 * (r - 1) + 1];
      }
    }
  // affine geometry (jacobian_p1_2d.f:45-101)
  m.geo.assign(8 * nel, 0.0);
  m.det.assign(nel, 0.0);
  for (int64_t e = 0; e < nel; ++e) {
    double x0 = X(tn(e, 0), 0), y0 = X(tn(e, 0), 1);
    double B00 = X(tn(e, 1), 0) - x0, B10 = X(tn(e, 1), 1) - y0;
    double B01 = X(tn(e, 2), 0) - x0, B11 = X(tn(e, 2), 1) - y0;
    double det = B00 * B11 - B01 * B10;
    if (std::fabs(det) < 1e-10) throw std::runtime_error("asmbly: det = 0");
    double *g = &m.geo[8 * e];
    g[0] = B00; g[1] = B10; g[2] = B01; g[3] = B11;
    g[4] = B11 / det; g[5] = -B10 / det; g[6] = -B01 / det; g[7] = B00 / det;
    m.det[e] = det;
  }

  // sparse pattern: union of element cliques, sorted columns (csr_length.f)
  {
    std::vector<std::vector<int32_t>> rows(neq);
    for (int64_t e = 0; e < nel; ++e)
      for (int i = 0; i < 22; ++i) {
        auto &r = rows[m.ldof[22 * e + i]];
        for (int j = 0; j < 22; ++j) r.push_back(m.ldof[22 * e + j]);
      }
    m.row_ptr.assign(neq + 1, 0);
    for (int64_t i = 0; i < neq; ++i) {
      auto &r = rows[i];
      std::sort(r.begin(), r.end());
      r.erase(std::unique(r.begin(), r.end()), r.end());
      m.row_ptr[i + 1] = m.row_ptr[i] + (int64_t)r.size();
    }
    m.nnz = m.row_ptr[neq];
    m.col_ind.resize(m.nnz);
    for (int64_t i = 0; i < neq; ++i)
      std::copy(rows[i].begin(), rows[i].end(), m.col_ind.begin() + m.row_ptr[i]);
  }
  // element slot map and per-nnz contributor lists (gather assembly)
  m.slot.assign((size_t)484 * nel, 0);
  std::vector<int32_t> cnt(m.nnz + 1, 0);
  for (int64_t e = 0; e < nel; ++e)
    for (int i = 0; i < 22; ++i) {
      int32_t r = m.ldof[22 * e + i];
      const int32_t *cb = m.col_ind.data() + m.row_ptr[r];
      const int32_t *ce = m.col_ind.data() + m.row_ptr[r + 1];
      for (int j = 0; j < 22; ++j) {
        int32_t c = m.ldof[22 * e + j];
        const int32_t *it = std::lower_bound(cb, ce, c);
        int32_t k = (int32_t)(it - m.col_ind.data());
        m.slot[(size_t)484 * e + 22 * i + j] = k;
        cnt[k + 1]++;
      }
    }
  m.cptr.assign(m.nnz + 1, 0);
  for (int64_t k = 0; k < m.nnz; ++k) m.cptr[k + 1] = m.cptr[k] + cnt[k + 1];
  m.citem.assign(m.cptr[m.nnz], 0);
  std::vector<int32_t> fill(m.cptr.begin(), m.cptr.end() - 1);
  for (int64_t e = 0; e < nel; ++e)
    for (int ij = 0; ij < 484; ++ij) {
      int32_t k = m.slot[(size_t)484 * e + ij];
      m.citem[fill[k]++] = (int32_t)(484 * e + ij);
    }
}

}  // namespace emu
