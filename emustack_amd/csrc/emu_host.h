// Host-side data structures of the MI355X FEM/S-matrix library.
//
// Per-mesh work (topology, periodic merge, sparse pattern, LU symbolic
// analysis) is done once on the host in C++ and cached in a Mesh object;
// everything that depends on the wavelength / Bloch vector runs on the GPU.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace emu {

// One node of the multifrontal assembly tree (nested-dissection order).
struct FrontInfo {
  int32_t s = 0;       // number of own (fully summed) variables
  int32_t u = 0;       // number of update variables (ancestor rows/cols)
  int32_t f = 0;       // front order s + u
  int32_t b = 0;       // first own variable in the new numbering
  int32_t parent = -1;
  int32_t child[2] = {-1, -1};
  int32_t nchild = 0;
  int32_t slot = 0;    // 0/1: which child of its parent
  int32_t level = 0;   // depth from the root
  int64_t off = 0;     // offset of the f x f column-major front in the pool
  int64_t uoff = 0;    // offset of the u update indices (new numbering)
  int64_t moff = 0;    // offset of the u parent-local positions (extend-add map)
  int64_t pivoff = 0;  // offset of the s pivot rows
  int64_t woff = 0;    // offset of the u-long forward-solve update vector
};

// Launch schedule of one level: fronts and tile tasks (precomputed once).
struct LevelPlan {
  std::vector<int32_t> fronts;   // fronts of this level
  int32_t max_s = 0;
  int32_t max_f = 0;
};

struct Symbolic {
  int64_t n = 0;
  std::vector<int32_t> perm;   // new -> old
  std::vector<int32_t> iperm;  // old -> new
  std::vector<FrontInfo> fronts;
  std::vector<int32_t> uidx;   // concatenated update indices (new numbering)
  std::vector<int32_t> pmap;   // concatenated parent-local positions of U
  std::vector<int64_t> nnz2front;  // per nnz of A: offset into the front pool
  std::vector<LevelPlan> levels;   // levels[0] = root level
  int64_t pool = 0;    // total complex entries of all fronts
  int64_t npiv = 0;    // total own variables (== n)
  int64_t nw = 0;      // total update-vector entries
  int32_t max_f = 0;
  double flops = 0;    // real flops of one numeric factorisation
  int64_t factor_entries = 0;  // entries of L and U actually used by a solve
};

// reference-element integral tables of the 1D grating (Lagrange P2 / P3):
// Mxx = int P2 P2, Kxy = int P2 (P3)', Kyy = int (P3)' (P3)', Myy = int P3 P3
// on [0, 1], row-major; physical element of length h: Mxx, Myy times h,
// Kyy divided by h, Kxy unchanged
constexpr int kMxx = 0, kKxy = 9, kKyy = 21, kMyy = 37, kTab1d = 53;

struct Mesh {
  int dim = 2;                    // 2: triangular .mail mesh, 1: 1D grating
  std::string path;
  int64_t npt = 0, nel = 0, n_edge = 0, npt_p3 = 0, n_ddl = 0, neq = 0, nnz = 0;
  double lat[4] = {0, 0, 0, 0};  // v1 = (lat[0], lat[1]), v2 = (lat[2], lat[3])
  std::vector<double> x;          // npt x 2
  std::vector<int64_t> type_nod;  // npt
  std::vector<int64_t> table_nod; // nel x 6, 1-based
  std::vector<int64_t> type_el;   // nel, 1-based material index
  std::vector<int64_t> T;         // nel x 14 entity numbers, 1-based
  std::vector<double> xE;         // n_ddl x 2
  std::vector<int64_t> tE;        // n_ddl x 2 (boundary type, dimension)
  std::vector<int64_t> ip_period_N, ip_period_E;  // 1-based reps, 0 = none
  std::vector<int64_t> ineq;      // n_ddl x 3, 1-based, 0 = none
  std::vector<int32_t> ldof;      // nel x 22, 0-based equation numbers
  std::vector<int8_t> orient;     // nel x 3 x 2 local vertex order per edge
  std::vector<double> dphase;     // nel x 14 x 2 (x_E - x_rep), 0 if none
  std::vector<int64_t> row_ptr;   // neq + 1
  std::vector<int32_t> col_ind;   // nnz, sorted per row
  std::vector<int32_t> slot;      // nel x 22 x 22 -> nnz index
  std::vector<int32_t> cptr;      // nnz + 1: contributors per nnz
  std::vector<int32_t> citem;     // element-entry index e*484 + i*22 + j
  // element geometry: affine map
  std::vector<double> geo;        // nel x 8: B00,B10,B01,B11, T00,T10,T01,T11 (col-major)
  std::vector<double> det;        // nel
  // 1D gratings (dim == 1): ldof is nel x 11, slot nel x 121, citem e*121+i*11+j
  std::vector<double> h1d;        // nel element lengths
  std::vector<double> xd1d;       // nel x 11 DOF coordinates
  double period_x = 1.0;          // distance carried by the Bloch phase
  double tab1d[kTab1d] = {};
  Symbolic sym;
};

// Build all per-mesh data from a .mail file.  Throws std::runtime_error.
void build_mesh(Mesh &m, const std::string &path);
// Build a 1D grating mesh from the reference's arrays (table_nod 3 x nel,
// 1-based [left, right, mid]; type_el 1-based; x of the 2 nel + 1 nodes).
void build_mesh_1d(Mesh &m, int64_t npt, int64_t nel, const int64_t *table_nod, const int64_t *type_el,
                   const double *x);
void element_tables_1d(double *tab);
// Nested-dissection multifrontal symbolic analysis of the pattern.
void build_symbolic(Mesh &m, int leaf_size);

}  // namespace emu
