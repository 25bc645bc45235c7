// Device-side state of a mesh and of a batch of solve instances, and the
// kernel launchers of the FEM / LU / Krylov / overlap stages.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "common.h"
#include "emu_host.h"

namespace emu {

template <typename T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  DBuf() = default;
  DBuf(const DBuf &) = delete;
  DBuf &operator=(const DBuf &) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&p, count * sizeof(T));
    if (e != hipSuccess) {
      p = nullptr;
      throw HipError("hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
    }
    n = count;
  }
  // synchronous: the host source may be a temporary
  void upload(const T *h, size_t count, hipStream_t s) {
    alloc(count);
    EMU_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    EMU_HIP(hipStreamSynchronize(s));
  }
  void upload(const std::vector<T> &v, hipStream_t s) { upload(v.data(), v.size(), s); }
};

// Start/stop events of one timed dispatch (both null: not timed).  They are
// handed to hipExtLaunchKernelGGL, which stamps them at the start and end of
// exactly that kernel, so the measured time excludes launch gaps and the
// neighbouring kernels of a stage.
struct KTime {
  hipEvent_t a = nullptr, b = nullptr;
};

// Live kernel timing (enabled by emu_profile).  A "sweep" is one solve or one
// Gram-Schmidt pass; every `every`-th sweep is recorded.  Classes: 0 forward
// solve (k_fwd), 1 backward solve (k_bwd), 2 Gram-Schmidt projection
// (k_atb_mfma16; k_atb_tall below 8-column blocks), 3 Gram-Schmidt update
// (k_combine; k_sub_vh_blk below 8-column blocks), 4 sparse M product of the
// Krylov block (k_spmm_il; k_spmm below 8 columns); per class: dispatches,
// milliseconds, algorithmic bytes.
struct Profiler {
  bool on = false;
  int every = 8;
  int64_t sweep = 0;
  bool active = false;
  struct Pend {
    int cls;
    double bytes;
    int a, b;
  };
  std::vector<hipEvent_t> ev;
  size_t used = 0;
  std::vector<Pend> pend;
  double stat[8][3] = {};
  void start_sweep() { active = on && (sweep++ % every == 0); }
  int slot() {
    if (used == ev.size()) {
      hipEvent_t e;
      EMU_HIP(hipEventCreate(&e));
      ev.push_back(e);
    }
    return (int)used++;
  }
  // events for the next dispatch of class `cls` moving `bytes` algorithmic
  // bytes; null events when this sweep is not recorded
  KTime time(int cls, double bytes) {
    if (!active) return {};
    int a = slot(), b = slot();
    pend.push_back({cls, bytes, a, b});
    return {ev[a], ev[b]};
  }
  // call after the stream has been synchronised
  void collect() {
    for (auto &p : pend) {
      float ms = 0;
      EMU_HIP(hipEventElapsedTime(&ms, ev[p.a], ev[p.b]));
      stat[p.cls][0] += 1;
      stat[p.cls][1] += ms;
      stat[p.cls][2] += p.bytes;
    }
    pend.clear();
    used = 0;
  }
  void reset() {
    for (auto &r : stat) r[0] = r[1] = r[2] = 0;
    pend.clear();
    used = 0;
    sweep = 0;
  }
  ~Profiler() {
    for (auto e : ev) (void)hipEventDestroy(e);
  }
};

// One solve task, 64 bytes, read by a single scalar load per workgroup: the
// tile (front t, first row r0, tiles per workgroup tpw; first column kb and
// partial-sum slot of a split task, else -1) with the front's sizes and
// offsets, so that no further lookup precedes the workgroup's first loads.
struct SolveTask {
  int32_t t, r0, tpw, kb, slot, s, f, fb;
  int64_t ioff, pivoff, fvoff, uoff;
};
static_assert(sizeof(SolveTask) == 64, "SolveTask is 16 words");

// Solve tasks of one level and direction: SolveTask records (off: word
// offset into the task array, a multiple of 16), partial-sum reductions
// (front, row0, first slot, count), the number of 64-row partial-sum slots
// and the column span of split tasks.
struct SolveTasks {
  int64_t off = 0, cnt = 0;
  int64_t roff = 0, rcnt = 0;
  int32_t nslot = 0, span = 0;
};

// Kernel task lists of one factorisation / solve level.
struct StepPlan {
  int32_t k0;
  int64_t panel_off, panel_cnt;  // front ids
  int64_t rows_off, rows_cnt;    // (front, row0): update rows s..f-1 of the panel, 64 per task
  int64_t trsm_off, trsm_cnt;    // (front, col0)
  int64_t gemm_off, gemm_cnt;    // (front, row0, col0)
};
struct LevelSched {
  std::vector<StepPlan> steps;
  int64_t fronts_off, fronts_cnt;     // fronts of the level
  int64_t ea_off[2], ea_cnt[2];       // extend-add tasks (child front, col0) by slot
  // forward / backward solve task lists, [0] without column splitting and
  // [1] with the long tiles split (chosen per solve by the grid size)
  SolveTasks fwd[2], bwd[2];
  int32_t max_f;
  // algorithmic bytes of one solve sweep of the level, per instance: factor
  // bytes (counted once per solve, whatever the RHS chunking: re-reads of a
  // front by several chunks are not algorithmic) and vector bytes (per RHS)
  double fwd_mat = 0, fwd_vec = 0, bwd_mat = 0, bwd_vec = 0;
};

struct DevMesh {
  Mesh host;
  int64_t ntyp = 0;
  // FEM constants
  DBuf<double> geo, det;       // nel x 8, nel
  DBuf<signed char> orient;    // nel x 6
  DBuf<double> dphase;         // nel x 28
  DBuf<int32_t> type_el;       // nel (0-based)
  DBuf<int32_t> ldof;          // nel x 22
  DBuf<int32_t> cptr, citem;   // assembly gather lists
  DBuf<double> eka, ekc, emd;  // element integrals nel x 484 (K pp-part, K qq-part, M pp-part)
  DBuf<int64_t> row_ptr;       // CSR of the pattern (original numbering)
  DBuf<int32_t> col_ind;
  // permuted CSR of M (new numbering) and map from original nnz
  DBuf<int64_t> prow_ptr;
  DBuf<int32_t> pcol_ind, pnnz_src;  // pnnz_src[k'] = original nnz index
  // transverse DOF contributor lists for the plane-wave load (J)
  DBuf<int32_t> dcptr, dcitem;       // per dof: (e*12 + l)
  DBuf<double> gt;                   // nel x 12 x 12: transverse field at P2 nodes from coefs
  // symbolic LU
  DBuf<int32_t> perm, iperm;
  DBuf<int32_t> fs, fu, ff, fb, fparent, fslot;
  DBuf<int64_t> foff, fuoff, fpivoff;
  DBuf<int32_t> uidx, pmap;
  DBuf<int64_t> nnz2front;
  DBuf<int64_t> fioff, ffvoff;       // inverted-form offset (prefix of 2 f s), prefix of f
  DBuf<int32_t> cmap;                // per front row: update-vector index of child slot 0 / 1, or -1
  int64_t isz = 0;                   // inverted-form entries per instance
  int64_t part_rhs = 0;              // partial-sum entries per RHS (largest level)
  DBuf<int32_t> tasks;               // all task lists
  int64_t tri_off = 0, tri_cnt = 0;  // triangular inversion tasks (front, J, kind)
  int64_t wz_off = 0, wz_cnt = 0;    // off-diagonal inverted blocks (front, kind, row0, col0)
  std::vector<LevelSched> sched;     // sched[0] = deepest level
  int32_t max_f = 0;
  // geometry for the plane-wave integrals
  // 1D gratings
  DBuf<double> h1d, xd1d, tab1d;     // element lengths, DOF coordinates (nel x 11), integral tables
  DBuf<double> xnod;                 // npt x 2
  DBuf<int64_t> table_nod;           // nel x 6 (1-based)
};

// A device context: one HIP stream and named, grow-only work buffers.
// Buffers are allocated with 1/16 headroom and grow by at least a quarter:
// batches on meshes of nearly equal size then reuse them, and hipFree (which
// waits for the whole device, stalling every other context's stream) stays
// out of the steady state.
struct Ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  Profiler prof;
  std::vector<void *> bufs;
  std::vector<size_t> sizes;
  std::vector<std::string> names;
  std::map<int, DevMesh *> dense;  // single-front LU handles of dense n x n batches
  void *get(const std::string &name, size_t bytes) {
    for (size_t i = 0; i < names.size(); ++i)
      if (names[i] == name) {
        if (sizes[i] >= bytes) return bufs[i];
        (void)hipFree(bufs[i]);
        bufs[i] = nullptr;
        const size_t grow = std::max(bytes + bytes / 16, sizes[i] + sizes[i] / 4);
        if (hipMalloc(&bufs[i], grow) != hipSuccess) throw HipError("hipMalloc failed for " + name);
        sizes[i] = grow;
        return bufs[i];
      }
    void *p = nullptr;
    const size_t first = std::max<size_t>(bytes + bytes / 16, 256);
    if (hipMalloc(&p, first) != hipSuccess) throw HipError("hipMalloc failed for " + name);
    names.push_back(name);
    bufs.push_back(p);
    sizes.push_back(first);
    return p;
  }
  template <typename T>
  T *buf(const std::string &name, size_t count) {
    return reinterpret_cast<T *>(get(name, count * sizeof(T)));
  }
  ~Ctx() {
    for (auto &kv : dense) delete kv.second;
    for (void *p : bufs) (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

void devmesh_build(DevMesh &dm, hipStream_t s);
// batched dense solve X = M^-1 R (R = identity when null): M n x n, R n x r,
// column-major, point-major; factorised by the multifrontal kernels as one front
void solve_dense(Ctx *c, int nb, int n, const zd *M, const zd *R, int r, zd *X);

// ---------------- FEM kernels (kernels_fem.hip) ----------------
void launch_elem_integrals(DevMesh &dm, hipStream_t s);
// ve[nb][nel][14]
void launch_phase(DevMesh &dm, int nb, const double *bloch_dev, zd *ve, hipStream_t s);
// mat1, mat2 [nb][nnz]
void launch_gather(DevMesh &dm, int nb, int ntyp, const zd *pp, const zd *qq, const zd *shift,
                   const zd *ve, zd *mat1, zd *mat2, hipStream_t s);
void launch_permute_vals(DevMesh &dm, int nb, const zd *src, zd *dst, hipStream_t s);

// ---------------- LU kernels (kernels_lu.hip) ----------------
// numeric factorisation into pool (piv: LAPACK-style row swaps, pflag: zero
// pivots perturbed, sig: pivot-order gather), then the inverted-factor form
// inv [nb][isz] and the forward gather records gx [nb][npiv][4] (per pivot:
// right-hand-side index, the two child update entries) that the solves read
void lu_factor(DevMesh &dm, int nb, const zd *mat1, zd *pool, int32_t *piv, int32_t *pflag, int32_t *sig,
               int32_t *gx, zd *inv, hipStream_t s);
void lu_invert(DevMesh &dm, int nb, const zd *pool, zd *inv, hipStream_t s);
// x = A^{-1} b, both in the new (nested-dissection) numbering, [nb][nrhs][n];
// y (same shape) and w ([nb][nrhs][solve_ws(dm)]) are work vectors.  sx:
// instance stride of x (default nrhs n), so that x can be a column slot of a
// larger per-instance array (the Krylov basis)
int64_t solve_ws(const DevMesh &dm);
void lu_solve(DevMesh &dm, int nb, int nrhs, const zd *inv, const int32_t *gx, const zd *b, zd *y, zd *x,
              zd *w, hipStream_t s, Profiler *prof = nullptr, int64_t sx = -1);
// the same with the complex64 copy of the inverted factor (lu_narrow)
void lu_solve(DevMesh &dm, int nb, int nrhs, const float2 *inv32, const int32_t *gx, const zd *b, zd *y, zd *x,
              zd *w, hipStream_t s, Profiler *prof = nullptr, int64_t sx = -1);
void lu_narrow(const DevMesh &dm, int nb, const zd *inv, float2 *inv32, hipStream_t s);
void launch_swap_f2(int64_t n, float2 *a, float2 *b, hipStream_t s);
// permute vectors old <-> new numbering: dst[new] = src[perm[new]] or inverse
void launch_permute_vec(DevMesh &dm, int nb, int ncol, const zd *src, zd *dst, bool to_new,
                        hipStream_t s);

// ---------------- dense / vector kernels (kernels_krylov.hip) ----------------
// y[b][r] = sum_k A[b][r][k] x[b][col k] for CSR (n x n), ncol columns
// M times one 8-column block X (column-major, ld n) through the
// row-interleaved scratch xi (nb x 8 n): y column-major (ld n).  xi_ready:
// xi already holds X's interleaved copy (written by launch_trmm_right)
void launch_spmm8_il(int nb, int64_t n, const int64_t *rp, const int32_t *ci, const zd *vals, int64_t vstride,
                     const zd *x, int64_t xstride, zd *xi, zd *y, int64_t ystride, hipStream_t s, KTime kt = {},
                     bool xi_ready = false);
// xi (optional, nb x n x ncol): scratch for the row-major copy that the
// many-column form (ncol > 8) works from
void launch_spmm(int nb, int64_t n, int ncol, const int64_t *rp, const int32_t *ci, const zd *vals,
                 int64_t vstride, const zd *x, int64_t xstride, zd *y, int64_t ystride,
                 hipStream_t s, KTime kt = {}, zd *xi = nullptr);
// C[b] (M x N, ldc) = op(A[b])^T B[b] with K rows; op = conj if conjA, split
// over K (partial products per K chunk in scratch, then a reduction).  N in
// {1, 2, 4, 8, 16}: the tall-skinny kernel (the one kt times); 8 x 8 Gram
// matrices of a block: a dedicated kernel; other N: 64 x 64 MFMA output tiles.
void launch_atb(int nb, int64_t K, int M, int N, const zd *A, int64_t lda, int64_t sA,
                const zd *B, int64_t ldb, int64_t sB, zd *C, int64_t ldc, int64_t sC,
                bool conjA, zd *scratch, size_t scratch_elems, hipStream_t s, KTime kt = {});
// C[b] (M x N) = alpha * A[b] (M x K) * B[b] (K x N) + beta*C[b], column-major
// (k_zgemm_nn: 64 x 64 tiles on the FP64 matrix cores); C must not alias A or B
void launch_gemm_nn(int nb, int64_t M, int N, int K, zd alpha, const zd *A, int64_t lda, int64_t sA,
                    const zd *B, int64_t ldb, int64_t sB, zd beta, zd *C, int64_t ldc, int64_t sC,
                    hipStream_t s);
// w[b] -= V[b][:, 0:k] h[b]  (h device, [b][hstride])
void launch_sub_vh(int nb, int64_t n, int k, const zd *V, int64_t ldv, int64_t sV, const zd *h,
                   int64_t sh, zd *w, int64_t sw, hipStream_t s);
// nrm[b] = ||w[b]||; then w[b] *= 1/nrm[b] written to out[b]
void launch_norm_scale(int nb, int64_t n, const zd *w, int64_t sw, zd *out, int64_t so,
                       double *nrm, double *scratch, hipStream_t s);
void launch_axpy_h(int nb, int k, const zd *h1, const zd *h2, int64_t sh, zd *H, int64_t sH,
                   hipStream_t s);
void launch_copy(int nb, int64_t n, const zd *src, int64_t ss, zd *dst, int64_t sd, hipStream_t s);
// exchange two ranges of n elements (two batch slots' data)
void launch_swap(int64_t n, zd *a, zd *b, hipStream_t s);
void launch_swap_i32(int64_t n, int32_t *a, int32_t *b, hipStream_t s);
// block Gram-Schmidt: W[:, :bq] -= V[:, :k] H[:k, :bq]
void launch_sub_vh_blk(int nb, int64_t n, int k, int bq, const zd *V, int64_t ldv, int64_t sV, const zd *H,
                       int64_t ldh, int64_t sH, zd *W, int64_t ldw, int64_t sW, hipStream_t s, KTime kt = {});
// Delayed block Gram-Schmidt (block size 8): W (n x nq, in place) = W L - V[:, :k] E
// (E row-major, k x nq per instance)
void launch_combine(int nb, int64_t n, int k, int nq, const zd *V, int64_t ldv, int64_t sV, const zd *E,
                    int64_t sE, const zd *L, zd *W, int64_t ldw, int64_t sW, hipStream_t s, KTime kt = {});
// ... and the coefficients of one step (fin: reorthogonalise the last block only)
void launch_dcgs_coef(int nb, bool fin, int j, int c0, const zd *P, int64_t ldp, int64_t sP, const zd *A0,
                      int64_t lda, int64_t sA, zd *H, int64_t ldh, int64_t sH, zd *E, int64_t sE, zd *L, int32_t *bad,
                      hipStream_t s);
// Cholesky-QR of a bq-column block: G = R^H R, Rinv = R^-1, Racc = R * Rprev
void launch_cholqr(int nb, int bq, const zd *G, zd *Rinv, const zd *Rprev, zd *Racc, int32_t *bad,
                   hipStream_t s);
// O = W * Rinv (bq x bq upper triangular), in place allowed; for bq = 8 and
// xi given, O is also written row-interleaved to xi (nb x 8 n, the layout
// launch_spmm8_il reads)
void launch_trmm_right(int nb, int64_t n, int bq, const zd *W, int64_t ldw, int64_t sW, const zd *Ri, zd *O,
                       int64_t ldo, int64_t sO, hipStream_t s, zd *xi = nullptr);
void launch_block_copy(int nb, int rows, int cols, const zd *S, int64_t lds, int64_t sS, zd *D, int64_t ldd,
                       int64_t sD, bool add, hipStream_t s);
void launch_fill(zd *p, int64_t n, zd v, hipStream_t s);

// ---------------- 1D gratings (kernels_1d.hip) ----------------
// which = 0: out1 = K - shift M, out2 = M of asmbly_1d.f; which = 1: out1 =
// the overlap form of orthogonal_1d.f (out2 unused).  bloch [nb][2].
void launch_gather_1d(DevMesh &dm, int nb, int ntyp, const zd *pp, const zd *qq, const zd *shift,
                      const double *bloch, int which, zd *out1, zd *out2, hipStream_t s);
// G[b] (n x 2P, col-major): plane-wave load of J (sign -1, exp(-i alpha x))
// or J^dagger (sign +1); alpha[b][P] in sorted plane-wave order, ky[b] the
// prime k_y, phase from bloch[b][0]
void launch_pw_load_1d(DevMesh &dm, int nb, int P, const double *alpha, const double *ky, const int32_t *sign,
                       const double *bloch, zd *G, hipStream_t s);
// plotting fields (kernels_fields.hip): 13-node fields of one instance's
// modes; their mode_pol fractions (beta: nval, eps: per material type;
// part: 4 nval nel scratch); sol_P2 of a 1D instance
void launch_fields_2d(DevMesh &dm, int nval, const zd *X, int64_t ldx, const zd *ve, zd *sol, hipStream_t s);
void launch_mode_pol_2d(DevMesh &dm, int nval, const zd *X, int64_t ldx, const zd *ve, const zd *beta,
                        const zd *eps, int core_t, zd *part, zd *pol, hipStream_t s);
void launch_sol_p2_1d(DevMesh &dm, int nval, const zd *X, int64_t ldx, double kx, zd *out, hipStream_t s);
// array_sol_1d.f normalisation value per mode: den[b][k]
void launch_field_max_1d(DevMesh &dm, int nb, int nval, const zd *X, int64_t sX, const double *bloch, double xmid,
                         double dq, zd *den, hipStream_t s);

// ---------------- post-processing (kernels_post.hip) ----------------
// coefficient fields: max |field| over (e, node, comp) per mode, first index
void launch_field_max(DevMesh &dm, int nb, int nval, const zd *X, int64_t ldx, int64_t sX,
                      const zd *ve, bool conj_phase, zd *zmax, double *scratch, hipStream_t s);
void launch_scale_cols(int nb, int64_t n, int ncol, zd *X, int64_t ldx, int64_t sX, const zd *f,
                       int64_t sf, bool invert, hipStream_t s);
// plane-wave load matrices G[b] (n x 2P, col-major): J = G^T X
// per instance b: sign[b] = -1 (J, exp(-ikx)) or +1 (J^dagger, exp(+ikx)); the
// DOF phase is the instance's own ve[b]
void launch_pw_load(DevMesh &dm, int nb, int P, const double *alpha, const double *beta,
                    const int32_t *pwpos, const int32_t *sign, const zd *ve, zd *JB, zd *G,
                    hipStream_t s);
void launch_permute_cols(int nb, int64_t n, int ncol, const zd *src, int64_t ss, const int32_t *idx,
                         zd *dst, int64_t sd, hipStream_t s);

}  // namespace emu
