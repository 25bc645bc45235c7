// Symbolic analysis for the multifrontal sparse LU of K - shift*M.
//
// The reference factors the shifted pencil with UMFPACK (umf4zsym/umf4znum,
// backend/fortran/valpr_umf_64.f:155-199), redoing the symbolic phase for
// every wavelength.  Here the pattern is fixed per mesh, so the ordering,
// assembly tree, front structures and all scatter maps are computed once and
// reused by every (wavelength, Bloch vector) instance on the GPU.
//
// Ordering: nested dissection with BFS level-set separators.  Every
// separator and every leaf sub-domain becomes one dense front (a relaxed
// supernode), which makes the numeric phase a sequence of batched dense
// kernels over the levels of a binary tree.
#include <algorithm>
#include <cstdio>
#include <numeric>
#include <stdexcept>

#include "emu_host.h"

namespace emu {

namespace {

struct NDNode {
  std::vector<int32_t> own;  // old indices, in elimination order
  int32_t child[2] = {-1, -1};
  int32_t nchild = 0;
};

struct ND {
  const int64_t *rp;
  const int32_t *ci;
  int64_t n;
  std::vector<int32_t> part;   // subset id of each vertex
  std::vector<int32_t> level;  // BFS level scratch
  std::vector<NDNode> nodes;
  int leaf;
  int32_t next_id = 1;

  // BFS over vertices with part[v]==id; returns vertices in BFS order and
  // fills level[] for them.
  void bfs(int32_t root, int32_t id, std::vector<int32_t> &order) {
    order.clear();
    order.push_back(root);
    level[root] = 0;
    part[root] = -id;  // visited marker
    for (size_t h = 0; h < order.size(); ++h) {
      int32_t v = order[h];
      for (int64_t k = rp[v]; k < rp[v + 1]; ++k) {
        int32_t w = ci[k];
        if (part[w] == id) {
          part[w] = -id;
          level[w] = level[v] + 1;
          order.push_back(w);
        }
      }
    }
    for (int32_t v : order) part[v] = id;
  }

  int32_t build(std::vector<int32_t> verts) {
    int32_t id = next_id++;
    for (int32_t v : verts) part[v] = id;
    int32_t me = (int32_t)nodes.size();
    nodes.emplace_back();
    if ((int)verts.size() <= leaf) {
      nodes[me].own = verts;
      return me;
    }
    std::vector<int32_t> order;
    // connected components
    bfs(verts[0], id, order);
    if (order.size() < verts.size()) {
      // disconnected: split components into two balanced groups, empty separator
      std::vector<std::vector<int32_t>> comps;
      std::vector<char> seen_local;
      int32_t cid = next_id++;
      for (int32_t v : order) part[v] = cid;
      comps.push_back(order);
      for (int32_t v : verts) {
        if (part[v] == id) {
          bfs(v, id, order);
          int32_t c2 = next_id++;
          for (int32_t w : order) part[w] = c2;
          comps.push_back(order);
        }
      }
      std::sort(comps.begin(), comps.end(),
                [](const std::vector<int32_t> &a, const std::vector<int32_t> &b) { return a.size() > b.size(); });
      std::vector<int32_t> g[2];
      for (auto &c : comps) {
        auto &t = g[0].size() <= g[1].size() ? g[0] : g[1];
        t.insert(t.end(), c.begin(), c.end());
      }
      int32_t c0 = build(g[0]);
      int32_t c1 = build(g[1]);
      nodes[me].child[0] = c0;
      nodes[me].child[1] = c1;
      nodes[me].nchild = 2;
      return me;
    }
    // pseudo-peripheral root
    int32_t root = verts[0];
    int32_t ecc = -1;
    for (int it = 0; it < 5; ++it) {
      bfs(root, id, order);
      int32_t last = level[order.back()];
      if (last <= ecc) break;
      ecc = last;
      int32_t best = order.back();
      int64_t bdeg = 1 << 30;
      for (auto it2 = order.rbegin(); it2 != order.rend() && level[*it2] == last; ++it2) {
        int64_t d = rp[*it2 + 1] - rp[*it2];
        if (d < bdeg) { bdeg = d; best = *it2; }
      }
      root = best;
    }
    bfs(root, id, order);
    int32_t nlev = level[order.back()] + 1;
    if (nlev < 3) {  // dense cluster: no useful separator
      nodes[me].own = verts;
      return me;
    }
    std::vector<int64_t> cnt(nlev, 0);
    for (int32_t v : order) cnt[level[v]]++;
    int64_t half = (int64_t)verts.size() / 2, acc = 0;
    int32_t L = 1;
    for (int32_t l = 0; l < nlev; ++l) {
      acc += cnt[l];
      if (acc >= half) { L = l; break; }
    }
    L = std::max<int32_t>(1, std::min<int32_t>(L, nlev - 2));
    // classify: 0 = A (< L), 1 = B (> L), 2 = S (== L)
    std::vector<int32_t> A, B, S;
    std::vector<int8_t> cls(0);
    for (int32_t v : order) {
      int32_t l = level[v];
      level[v] = l < L ? 0 : (l > L ? 1 : 2);
    }
    // refine separator: S vertices without a B neighbour go to A, then
    // S vertices without an A neighbour go to B
    for (int pass = 0; pass < 2; ++pass) {
      int32_t other = pass == 0 ? 1 : 0;
      for (int32_t v : order) {
        if (level[v] != 2) continue;
        bool touch = false;
        for (int64_t k = rp[v]; k < rp[v + 1] && !touch; ++k) {
          int32_t w = ci[k];
          if (part[w] == id && level[w] == other) touch = true;
        }
        if (!touch) level[v] = pass == 0 ? 0 : 1;
      }
    }
    for (int32_t v : order) {
      if (level[v] == 0) A.push_back(v);
      else if (level[v] == 1) B.push_back(v);
      else S.push_back(v);
    }
    if (A.empty() || B.empty()) {
      nodes[me].own = verts;
      return me;
    }
    std::sort(S.begin(), S.end());
    nodes[me].own = S;
    int32_t c0 = build(A);
    int32_t c1 = build(B);
    nodes[me].child[0] = c0;
    nodes[me].child[1] = c1;
    nodes[me].nchild = 2;
    return me;
  }
};

}  // namespace

void build_symbolic(Mesh &m, int leaf_size) {
  Symbolic &sy = m.sym;
  const int64_t n = m.neq;
  sy.n = n;
  ND nd;
  nd.rp = m.row_ptr.data();
  nd.ci = m.col_ind.data();
  nd.n = n;
  nd.part.assign(n, 0);
  nd.level.assign(n, 0);
  nd.leaf = leaf_size;
  std::vector<int32_t> all(n);
  std::iota(all.begin(), all.end(), 0);
  int32_t root = nd.build(all);
  // post-order numbering: children first, then own variables
  const int32_t nn = (int32_t)nd.nodes.size();
  std::vector<int32_t> post;
  post.reserve(nn);
  {
    std::vector<std::pair<int32_t, int>> st;
    st.push_back({root, 0});
    while (!st.empty()) {
      auto &t = st.back();
      NDNode &nd_ = nd.nodes[t.first];
      if (t.second < nd_.nchild) {
        int32_t c = nd_.child[t.second++];
        st.push_back({c, 0});
      } else {
        post.push_back(t.first);
        st.pop_back();
      }
    }
  }
  sy.perm.assign(n, 0);
  sy.iperm.assign(n, 0);
  std::vector<int32_t> fid(nn, -1);  // ND node -> front index (post order)
  sy.fronts.assign(nn, FrontInfo());
  int32_t pos = 0;
  for (int32_t i = 0; i < nn; ++i) {
    int32_t t = post[i];
    fid[t] = i;
    FrontInfo &F = sy.fronts[i];
    F.b = pos;
    F.s = (int32_t)nd.nodes[t].own.size();
    for (int32_t v : nd.nodes[t].own) {
      sy.perm[pos] = v;
      sy.iperm[v] = pos;
      ++pos;
    }
  }
  if (pos != n) throw std::runtime_error("symbolic: ordering does not cover all variables");
  for (int32_t i = 0; i < nn; ++i) {
    int32_t t = post[i];
    FrontInfo &F = sy.fronts[i];
    F.nchild = nd.nodes[t].nchild;
    for (int c = 0; c < F.nchild; ++c) {
      int32_t ci = fid[nd.nodes[t].child[c]];
      F.child[c] = ci;
      sy.fronts[ci].parent = i;
      sy.fronts[ci].slot = c;
    }
  }
  // owner front of each variable (new numbering)
  std::vector<int32_t> owner(n);
  for (int32_t i = 0; i < nn; ++i)
    for (int32_t k = 0; k < sy.fronts[i].s; ++k) owner[sy.fronts[i].b + k] = i;
  // update index sets, bottom-up (post order = children before parents)
  std::vector<std::vector<int32_t>> U(nn);
  std::vector<int32_t> mark(n, -1);
  for (int32_t i = 0; i < nn; ++i) {
    FrontInfo &F = sy.fronts[i];
    int32_t e = F.b + F.s;
    std::vector<int32_t> &u = U[i];
    for (int32_t k = 0; k < F.s; ++k) {
      int32_t vold = sy.perm[F.b + k];
      for (int64_t q = m.row_ptr[vold]; q < m.row_ptr[vold + 1]; ++q) {
        int32_t w = sy.iperm[m.col_ind[q]];
        if (w >= e && mark[w] != i) { mark[w] = i; u.push_back(w); }
      }
    }
    for (int c = 0; c < F.nchild; ++c)
      for (int32_t w : U[F.child[c]])
        if (w >= e && mark[w] != i) { mark[w] = i; u.push_back(w); }
    std::sort(u.begin(), u.end());
    F.u = (int32_t)u.size();
    F.f = F.s + F.u;
  }
  // levels (depth from root) and storage offsets
  int32_t rootf = fid[root];
  sy.fronts[rootf].level = 0;
  int32_t maxlev = 0;
  for (int32_t i = nn - 1; i >= 0; --i) {
    FrontInfo &F = sy.fronts[i];
    if (F.parent >= 0) F.level = sy.fronts[F.parent].level + 1;
    maxlev = std::max(maxlev, F.level);
  }
  sy.levels.assign(maxlev + 1, LevelPlan());
  int64_t off = 0, uoff = 0, pivoff = 0;
  sy.max_f = 0;
  sy.flops = 0;
  sy.factor_entries = 0;
  for (int32_t i = 0; i < nn; ++i) {
    FrontInfo &F = sy.fronts[i];
    F.off = off;
    off += (int64_t)F.f * F.f;
    F.uoff = uoff;
    F.moff = uoff;
    F.woff = uoff;
    uoff += F.u;
    F.pivoff = pivoff;
    pivoff += F.s;
    sy.max_f = std::max(sy.max_f, F.f);
    LevelPlan &L = sy.levels[F.level];
    L.fronts.push_back(i);
    L.max_s = std::max(L.max_s, F.s);
    L.max_f = std::max(L.max_f, F.f);
    double fs = F.f, ss = F.s;
    sy.flops += 8.0 * (fs * fs * ss - fs * ss * ss + ss * ss * ss / 3.0);
    sy.factor_entries += (int64_t)F.s * F.s + 2 * (int64_t)F.s * F.u;
  }
  sy.pool = off;
  sy.npiv = pivoff;
  sy.nw = uoff;
  sy.uidx.assign(uoff, 0);
  sy.pmap.assign(uoff, 0);
  for (int32_t i = 0; i < nn; ++i) {
    FrontInfo &F = sy.fronts[i];
    std::copy(U[i].begin(), U[i].end(), sy.uidx.begin() + F.uoff);
  }
  // extend-add maps: child U entry -> parent-local position
  for (int32_t i = 0; i < nn; ++i) {
    FrontInfo &F = sy.fronts[i];
    if (F.parent < 0) {
      if (F.u != 0) throw std::runtime_error("symbolic: root front has update rows");
      continue;
    }
    const FrontInfo &P = sy.fronts[F.parent];
    const int32_t *pu = sy.uidx.data() + P.uoff;
    for (int32_t k = 0; k < F.u; ++k) {
      int32_t w = sy.uidx[F.uoff + k];
      int32_t loc;
      if (w >= P.b && w < P.b + P.s) {
        loc = w - P.b;
      } else {
        const int32_t *it = std::lower_bound(pu, pu + P.u, w);
        if (it == pu + P.u || *it != w) throw std::runtime_error("symbolic: extend-add index missing in parent");
        loc = P.s + (int32_t)(it - pu);
      }
      sy.pmap[F.moff + k] = loc;
    }
  }
  // scatter map: every nnz of A -> one position of one front
  sy.nnz2front.assign(m.nnz, -1);
  auto local_of = [&](const FrontInfo &F, int32_t w) -> int32_t {
    if (w >= F.b && w < F.b + F.s) return w - F.b;
    const int32_t *pu = sy.uidx.data() + F.uoff;
    const int32_t *it = std::lower_bound(pu, pu + F.u, w);
    if (it == pu + F.u || *it != w) return -1;
    return F.s + (int32_t)(it - pu);
  };
  for (int64_t r = 0; r < n; ++r)
    for (int64_t q = m.row_ptr[r]; q < m.row_ptr[r + 1]; ++q) {
      int32_t pi = sy.iperm[r], pj = sy.iperm[m.col_ind[q]];
      const FrontInfo &F = sy.fronts[owner[std::min(pi, pj)]];
      int32_t li = local_of(F, pi), lj = local_of(F, pj);
      if (li < 0 || lj < 0) throw std::runtime_error("symbolic: entry outside its front");
      sy.nnz2front[q] = F.off + li + (int64_t)lj * F.f;
    }
}

}  // namespace emu
