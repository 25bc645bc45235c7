// Multifrontal sparse LU of K - shift*M on the GPU, batched over instances
// that share one mesh (replaces UMFPACK umf4znum / umf4zsol, reference
// backend/fortran/valpr_umf_64.f:198 and :340).
//
// Every front is a dense column-major f x f block [S | U] (S = own,
// fully-summed variables, U = update variables).  The numeric phase walks
// the assembly tree level by level (deepest first); within a level all
// fronts of all instances are processed by the same launches:
//   panel  : NB columns of partial-pivoting LU (pivots restricted to S rows)
//   trsm   : U12 block row  <- L11^{-1} U12
//   gemm   : trailing update F22 -= L21 U12 (64 x 64 tiles)
//   extend : child Schur complement added into its parent (slot 0, then 1:
//            no two tasks of one launch touch the same parent -> no atomics)
// The factors are then rewritten into an inverted-factor form and the solves
// walk the same tree as dense matrix-vector products: forward elimination
// bottom-up with update vectors passed to the parent (deterministic,
// atomic-free), back substitution top-down.
#include "engine.h"

namespace emu {

constexpr int NB = 32;     // panel width
constexpr int TILE = 64;   // gemm tile
constexpr int EA_COLS = 8; // columns per extend-add task

struct FrontArrays {
  const int32_t *s, *u, *f, *b, *parent;
  const int64_t *off, *uoff, *pivoff, *ioff, *fvoff;
  const int32_t *uidx, *pmap, *cmap;
};

static FrontArrays fa(DevMesh &dm) {
  FrontArrays a;
  a.s = dm.fs.p; a.u = dm.fu.p; a.f = dm.ff.p; a.b = dm.fb.p; a.parent = dm.fparent.p;
  a.off = dm.foff.p; a.uoff = dm.fuoff.p; a.pivoff = dm.fpivoff.p;
  a.ioff = dm.fioff.p; a.fvoff = dm.ffvoff.p;
  a.uidx = dm.uidx.p; a.pmap = dm.pmap.p; a.cmap = dm.cmap.p;
  return a;
}

__global__ void k_fill(zd *__restrict__ p, int64_t n, zd v) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[i] = v;
}

void launch_fill(zd *p, int64_t n, zd v, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(256), 0, s, p, n, v);
  EMU_LAUNCH_CHECK();
}

// pool[b][nnz2front[k]] = mat1[b][k]
__global__ void k_scatter_fronts(int64_t nnz, int64_t pool, const int64_t *__restrict__ map,
                                 const zd *__restrict__ mat1, zd *__restrict__ P) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (k >= nnz) return;
  P[b * pool + map[k]] = mat1[b * nnz + k];
}

__global__ __launch_bounds__(256) void k_panel(const int32_t *__restrict__ tasks, int k0,
                                               FrontArrays A, zd *__restrict__ pool, int64_t psz,
                                               int32_t *__restrict__ piv, int64_t npiv,
                                               int32_t *__restrict__ pflag, int32_t *__restrict__ sig,
                                               double pert) {
  const int t = tasks[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  zd *F = pool + b * psz + A.off[t];
  int32_t *pv = piv + b * npiv + A.pivoff[t];
  int32_t *pf = pflag + b * npiv + A.pivoff[t];
  int32_t *sg = sig + b * npiv + A.pivoff[t];
  const int w = min(NB, s - k0);
  if (k0 == 0) {
    for (int i = tid; i < s; i += 256) sg[i] = i;
    __syncthreads();
  }
  __shared__ double redv[4];
  __shared__ int redi[4];
  __shared__ zd urow[NB];
  const int lane = tid & 63, wave = tid >> 6;
  for (int jj = 0; jj < w; ++jj) {
    const int k = k0 + jj;
    double best = -1.0;
    int bi = k;
    for (int r = k + tid; r < s; r += 256) {
      double a = zabs1(F[r + (int64_t)k * f]);
      if (a > best) { best = a; bi = r; }
    }
    // pivot search: largest |.|_1, smallest row on ties (wave shuffles, then
    // the four waves' winners through LDS: two barriers per column)
#pragma unroll
    for (int m = 32; m > 0; m >>= 1) {
      const double a = __shfl_xor(best, m, 64);
      const int ai = __shfl_xor(bi, m, 64);
      if (a > best || (a == best && ai < bi)) {
        best = a;
        bi = ai;
      }
    }
    if (lane == 0) {
      redv[wave] = best;
      redi[wave] = bi;
    }
    __syncthreads();
    int p = redi[0];
    {
      double pb = redv[0];
      for (int q = 1; q < 4; ++q)
        if (redv[q] > pb || (redv[q] == pb && redi[q] < p)) {
          pb = redv[q];
          p = redi[q];
        }
    }
    if (p != k)
      for (int c = tid; c < f; c += 256) {
        zd a = F[k + (int64_t)c * f];
        F[k + (int64_t)c * f] = F[p + (int64_t)c * f];
        F[p + (int64_t)c * f] = a;
      }
    __syncthreads();
    zd d = F[k + (int64_t)k * f];
    const bool zero = (d.x == 0.0 && d.y == 0.0);
    if (zero) d = zmk(pert, 0.0);
    if (tid == 0) {
      pv[k] = p;
      pf[k] = zero ? 1 : 0;
      if (p != k) {
        const int32_t a = sg[k];
        sg[k] = sg[p];
        sg[p] = a;
      }
      if (zero) F[k + (int64_t)k * f] = d;
    }
    const zd dinv = zdiv(zmk(1.0, 0.0), d);
    const int nrest = w - jj - 1;
    for (int c = tid; c < nrest; c += 256) urow[c] = F[k + (int64_t)(k + 1 + c) * f];
    __syncthreads();
    for (int r = k + 1 + tid; r < s; r += 256) {
      zd l = zmul(F[r + (int64_t)k * f], dinv);
      F[r + (int64_t)k * f] = l;
      // rank-1 update of the row's remaining panel columns, eight at a time
      // (all eight loads in flight before the updates)
      zd *pr = &F[r + (int64_t)(k + 1) * f];
      for (int c0 = 0; c0 < nrest; c0 += 8) {
        zd v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = pr[(int64_t)min(c0 + u, nrest - 1) * f];
#pragma unroll
        for (int u = 0; u < 8; ++u)
          if (c0 + u < nrest) pr[(int64_t)(c0 + u) * f] = zfms(l, urow[c0 + u], v[u]);
      }
    }
    __syncthreads();
  }
}

// The panel's update rows (s <= r < f: not pivot candidates) by a row-wise
// triangular solve against the factored panel block U (w x w, unit-free
// upper triangle incl. the possibly perturbed diagonal): x U = a_r.  The
// operations are those k_panel's column steps would apply to the row, in the
// same order (l = x_j / U_jj via the reciprocal, then x_c -= l U_jc), so the
// result is the same; the rows are independent, so they run on all CUs
// instead of inside the panel's single workgroup.  One thread per row.
__global__ __launch_bounds__(64) void k_panel_rows(const int32_t *__restrict__ tasks, int k0, FrontArrays A,
                                                   zd *__restrict__ pool, int64_t psz) {
  const int t = tasks[2 * blockIdx.x], r0 = tasks[2 * blockIdx.x + 1];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  zd *F = pool + b * psz + A.off[t];
  const int w = min(NB, s - k0);
  __shared__ zd U[NB][NB + 1];
  __shared__ zd dinv[NB];
  for (int idx = tid; idx < w * w; idx += 64) {
    const int i = idx % w, j = idx / w;
    U[i][j] = F[(k0 + i) + (int64_t)(k0 + j) * f];
  }
  __syncthreads();
  if (tid < w) dinv[tid] = zdiv(zmk(1.0, 0.0), U[tid][tid]);
  __syncthreads();
  const int r = r0 + tid;
  if (r >= f) return;
  zd *row = F + r + (int64_t)k0 * f;
  zd x[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) x[j] = row[(int64_t)min(j, w - 1) * f];
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < w) {
      const zd l = zmul(x[j], dinv[j]);
      x[j] = l;
#pragma unroll
      for (int c = j + 1; c < NB; ++c)
        if (c < w) x[c] = zfms(l, U[j][c], x[c]);
    }
  }
#pragma unroll
  for (int j = 0; j < NB; ++j)
    if (j < w) row[(int64_t)j * f] = x[j];
}

__global__ __launch_bounds__(64) void k_trsm(const int32_t *__restrict__ tasks, int k0, FrontArrays A,
                                             zd *__restrict__ pool, int64_t psz) {
  const int t = tasks[2 * blockIdx.x], c0 = tasks[2 * blockIdx.x + 1];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  zd *F = pool + b * psz + A.off[t];
  const int w = min(NB, s - k0);
  __shared__ zd L[NB][NB + 1];
  for (int idx = tid; idx < w * w; idx += 64) {
    int i = idx % w, j = idx / w;
    L[i][j] = F[(k0 + i) + (int64_t)(k0 + j) * f];
  }
  __syncthreads();
  const int c = c0 + tid;
  if (c >= f) return;
  zd *col = F + (int64_t)c * f + k0;
  zd x[NB];
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    const zd v = col[min(i, w - 1)];
    x[i] = i < w ? v : zmk(0, 0);
  }
#pragma unroll
  for (int i = 1; i < NB; ++i) {
    if (i < w) {
      zd acc = x[i];
#pragma unroll
      for (int j = 0; j < i; ++j) acc = zfms(L[i][j], x[j], acc);
      x[i] = acc;
    }
  }
#pragma unroll
  for (int i = 0; i < NB; ++i)
    if (i < w) col[i] = x[i];
}

// C -= L21 U12 on one 64 x 64 tile of the trailing block.  The panel's K = w
// <= 32 columns are staged through LDS in two halves of 16 (32 KB of LDS, so
// that four workgroups share a CU); thread (tx, ty) owns rows tx + 16 i and
// columns ty + 16 j, so a wave's C accesses cover 16 consecutive rows of 4
// columns; the 16 reads of C are issued together from clamped addresses
// before any is used, the stores are masked.
constexpr int GK = 16;  // K staged per round
__global__ __launch_bounds__(256) void k_gemm_update(const int32_t *__restrict__ tasks, int k0,
                                                     FrontArrays A, zd *__restrict__ pool,
                                                     int64_t psz) {
  const int t = tasks[3 * blockIdx.x], r0 = tasks[3 * blockIdx.x + 1], c0 = tasks[3 * blockIdx.x + 2];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  zd *F = pool + b * psz + A.off[t];
  const int w = min(NB, s - k0);
  __shared__ zd As[GK][TILE];
  __shared__ zd Bs[GK][TILE];
  const int tx = tid % 16, ty = tid / 16;
  zd acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = zmk(0, 0);
  for (int kb = 0; kb < w; kb += GK) {
    if (kb > 0) __syncthreads();
    // staging: unconditional loads from clamped addresses (a load inside a
    // branch is waited for before the branch merges, which would serialise
    // the tile's loads), out-of-range entries zeroed afterwards
#pragma unroll
    for (int u = 0; u < GK * TILE / 256; ++u) {
      const int idx = tid + 256 * u, i = idx % TILE, kk = idx / TILE, r = r0 + i;
      const zd v = F[min(r, f - 1) + (int64_t)(k0 + min(kb + kk, w - 1)) * f];
      As[kk][i] = (kb + kk < w && r < f) ? v : zmk(0, 0);
    }
#pragma unroll
    for (int u = 0; u < GK * TILE / 256; ++u) {
      const int idx = tid + 256 * u, kk = idx % GK, jj = idx / GK, c = c0 + jj;
      const zd v = F[(k0 + min(kb + kk, w - 1)) + (int64_t)min(c, f - 1) * f];
      Bs[kk][jj] = (kb + kk < w && c < f) ? v : zmk(0, 0);
    }
    __syncthreads();
    const int kn = min(GK, w - kb);
    for (int kk = 0; kk < kn; ++kk) {
      zd a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = zfma(a[i], bb[j], acc[i][j]);
    }
  }
  // in two halves of 8 entries (fewer live registers)
#pragma unroll
  for (int jh = 0; jh < 4; jh += 2) {
    zd cv[4][2];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int i = 0; i < 4; ++i)
        cv[i][j] = F[min(r0 + tx + 16 * i, f - 1) + (int64_t)min(c0 + ty + 16 * (jh + j), f - 1) * f];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int c = c0 + ty + 16 * (jh + j);
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = r0 + tx + 16 * i;
        if (r < f && c < f) F[r + (int64_t)c * f] = zsub(cv[i][j], acc[i][jh + j]);
      }
    }
  }
}

__global__ __launch_bounds__(256) void k_extend_add(const int32_t *__restrict__ tasks, FrontArrays A,
                                                    zd *__restrict__ pool, int64_t psz) {
  const int t = tasks[2 * blockIdx.x], c0 = tasks[2 * blockIdx.x + 1];
  const int b = blockIdx.y;
  const int s = A.s[t], u = A.u[t], f = A.f[t];
  const int p = A.parent[t];
  const int fp = A.f[p];
  const zd *C = pool + b * psz + A.off[t];
  zd *P = pool + b * psz + A.off[p];
  const int32_t *mp = A.pmap + A.uoff[t];
  for (int cc = 0; cc < EA_COLS; ++cc) {
    int c = c0 + cc;
    if (c >= u) break;
    int64_t pc = (int64_t)mp[c] * fp;
    const zd *src = C + (int64_t)(s + c) * f + s;
    for (int r = threadIdx.x; r < u; r += blockDim.x) {
      zd *d = P + pc + mp[r];
      *d = zadd(*d, src[r]);
    }
  }
}

// Forward gather records, one per pivot: the right-hand-side index of the
// pivot's row (fb + j, j = sig[k]) and the two child update entries that add
// into that row (-1: none).  The solves of one factorisation then fetch a
// front's inputs with one dependent load instead of three.
__global__ __launch_bounds__(64) void k_gather_index(FrontArrays A, const int32_t *__restrict__ sig, int64_t npiv,
                                                     int4 *__restrict__ gx) {
  const int t = blockIdx.x, b = blockIdx.y;
  const int s = A.s[t], fb = A.b[t];
  const int64_t po = A.pivoff[t], fvo = A.fvoff[t];
  const int32_t *sg = sig + b * npiv + po;
  int4 *g = gx + b * npiv + po;
  for (int k = threadIdx.x; k < s; k += 64) {
    const int j = sg[k];
    g[k] = make_int4(fb + j, A.cmap[2 * (fvo + j)], A.cmap[2 * (fvo + j) + 1], 0);
  }
}

void lu_factor(DevMesh &dm, int nb, const zd *mat1, zd *pool, int32_t *piv, int32_t *pflag, int32_t *sig,
               int32_t *gx, zd *inv, hipStream_t s) {
  const Symbolic &sy = dm.host.sym;
  const int64_t psz = sy.pool;
  FrontArrays A = fa(dm);
  launch_fill(pool, psz * nb, zmk(0, 0), s);
  {
    int64_t nnz = dm.host.nnz;
    dim3 g((unsigned)((nnz + 255) / 256), nb);
    hipLaunchKernelGGL(k_scatter_fronts, g, dim3(256), 0, s, nnz, psz, dm.nnz2front.p, mat1, pool);
    EMU_LAUNCH_CHECK();
  }
  const double pert = 1e-13;
  for (const LevelSched &L : dm.sched) {
    for (const StepPlan &st : L.steps) {
      if (st.panel_cnt)
        hipLaunchKernelGGL(k_panel, dim3((unsigned)st.panel_cnt, nb), dim3(256), 0, s,
                           dm.tasks.p + st.panel_off, st.k0, A, pool, psz, piv, sy.npiv, pflag, sig, pert);
      if (st.rows_cnt)
        hipLaunchKernelGGL(k_panel_rows, dim3((unsigned)st.rows_cnt, nb), dim3(64), 0, s,
                           dm.tasks.p + st.rows_off, st.k0, A, pool, psz);
      if (st.trsm_cnt)
        hipLaunchKernelGGL(k_trsm, dim3((unsigned)st.trsm_cnt, nb), dim3(64), 0, s,
                           dm.tasks.p + st.trsm_off, st.k0, A, pool, psz);
      if (st.gemm_cnt)
        hipLaunchKernelGGL(k_gemm_update, dim3((unsigned)st.gemm_cnt, nb), dim3(256), 0, s,
                           dm.tasks.p + st.gemm_off, st.k0, A, pool, psz);
    }
    for (int sl = 0; sl < 2; ++sl)
      if (L.ea_cnt[sl])
        hipLaunchKernelGGL(k_extend_add, dim3((unsigned)L.ea_cnt[sl], nb), dim3(256), 0, s,
                           dm.tasks.p + L.ea_off[sl], A, pool, psz);
    EMU_LAUNCH_CHECK();
  }
  lu_invert(dm, nb, pool, inv, s);
  const int nf = (int)sy.fronts.size();
  if (nf)
    hipLaunchKernelGGL(k_gather_index, dim3((unsigned)nf, nb), dim3(64), 0, s, A, sig, sy.npiv,
                       reinterpret_cast<int4 *>(gx));
  EMU_LAUNCH_CHECK();
}

// ------------------------------------------------- inverted-factor form
//
// After the numeric factorisation every front is rewritten, out of place,
// into the form the solves use:
//   forward block  (f x s, ld f):  [ L11^{-1} ; -L21 L11^{-1} ]
//   backward block (s x f, ld s):  [ U11^{-1} , -U11^{-1} U12 ]
// so that one front's forward step is y = Fwd * (P r_S) (+ r_U on the update
// rows) and its backward step is x_S = Bwd * [y_S ; x_U]: dense matrix-vector
// products that split over row tiles across the whole GPU instead of a
// column-sequential substitution inside one workgroup.  P is the front's
// row-pivot permutation, applied as a gather (the gx records).

constexpr int TB = 32;  // triangular block of the inversion

__device__ __forceinline__ zd shfl_z(zd v, int lane) {
  return zmk(__shfl(v.x, lane, 64), __shfl(v.y, lane, 64));
}

// One workgroup per (front, column block J, kind, instance): column block J
// of L11^{-1} (kind 0, unit lower) or U11^{-1} (kind 1, upper) by blocked
// substitution against identity columns.  Column blocks are independent.
__global__ __launch_bounds__(256) void k_triinv(const int32_t *__restrict__ tasks, FrontArrays A,
                                                const zd *__restrict__ pool, int64_t psz, zd *__restrict__ inv,
                                                int64_t isz) {
  const int t = tasks[3 * blockIdx.x], J = tasks[3 * blockIdx.x + 1], kind = tasks[3 * blockIdx.x + 2];
  const int b = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int s = A.s[t], f = A.f[t];
  const zd *F = pool + b * psz + A.off[t];
  zd *O = inv + b * isz + A.ioff[t] + (kind ? (int64_t)f * s : 0);
  const int64_t ldo = kind ? s : f;
  const int nbk = (s + TB - 1) / TB;
  const int j0 = J * TB, wj = min(TB, s - j0);
  __shared__ zd Ts[TB][TB + 1];
  __shared__ zd Xs[TB][TB + 1];
  const int r = tid & 31, cg = tid >> 5;  // rows r, columns cg + 8m
  // structural zeros of this column block
  {
    const int z0 = kind ? j0 + wj : 0, z1 = kind ? s : j0;
    for (int idx = tid; idx < (z1 - z0) * wj; idx += 256) {
      int i = z0 + idx % (z1 - z0), c = idx / (z1 - z0);
      O[i + (int64_t)(j0 + c) * ldo] = zmk(0, 0);
    }
  }
  const int Ib = J, Ie = kind ? -1 : nbk, dI = kind ? -1 : 1;
  for (int I = Ib; I != Ie; I += dI) {
    const int i0 = I * TB, wi = min(TB, s - i0);
    zd acc[4];
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int c = cg + 8 * m;
      acc[m] = (I == J && r == c) ? zmk(1, 0) : zmk(0, 0);
    }
    const int Kb = kind ? I + 1 : J, Ke = kind ? J + 1 : I;
    for (int K = Kb; K < Ke; ++K) {
      const int k0 = K * TB, wk = min(TB, s - k0);
#pragma unroll
      for (int u = 0; u < TB * TB / 256; ++u) {
        const int idx = tid + 256 * u, i = idx % TB, kk = idx / TB;
        const zd a = F[(i0 + min(i, wi - 1)) + (int64_t)(k0 + min(kk, wk - 1)) * f];
        const zd x = O[(k0 + min(i, wk - 1)) + (int64_t)(j0 + min(kk, wj - 1)) * ldo];
        Ts[i][kk] = (i < wi && kk < wk) ? a : zmk(0, 0);
        Xs[i][kk] = (i < wk && kk < wj) ? x : zmk(0, 0);
      }
      __syncthreads();
#pragma unroll 4
      for (int kk = 0; kk < TB; ++kk) {
        const zd a = Ts[r][kk];
#pragma unroll
        for (int m = 0; m < 4; ++m) acc[m] = zfms(a, Xs[kk][cg + 8 * m], acc[m]);
      }
      __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < TB * TB / 256; ++u) {
      const int idx = tid + 256 * u, i = idx % TB, j = idx / TB;
      const zd a = F[(i0 + min(i, wi - 1)) + (int64_t)(i0 + min(j, wi - 1)) * f];
      Ts[i][j] = (i < wi && j < wi) ? a : zmk(0, 0);
    }
#pragma unroll
    for (int m = 0; m < 4; ++m) Xs[r][cg + 8 * m] = acc[m];
    __syncthreads();
    for (int c = wave; c < wj; c += 4) {
      zd v = lane < wi ? Xs[lane][c] : zmk(0, 0);
      if (kind == 0) {
        for (int k = 0; k < wi - 1; ++k) {
          const zd xk = shfl_z(v, k);
          if (lane > k && lane < wi) v = zfms(Ts[lane][k], xk, v);
        }
      } else {
        for (int k = wi - 1; k >= 0; --k) {
          const zd xk = zdiv(shfl_z(v, k), Ts[k][k]);
          if (lane == k) v = xk;
          else if (lane < k) v = zfms(Ts[lane][k], xk, v);
        }
      }
      if (lane < wi) Xs[lane][c] = v;
    }
    __syncthreads();
    for (int idx = tid; idx < wi * wj; idx += 256) {
      int i = idx % wi, c = idx / wi;
      O[(i0 + i) + (int64_t)(j0 + c) * ldo] = Xs[i][c];
    }
    __syncthreads();
  }
}

// Off-diagonal blocks of the inverted form (64 x 64 output tiles):
//   kind 0:  Fwd[s+i, j] = -sum_{k>=j} L21[i,k] Linv[k,j]
//   kind 1:  Bwd[i, s+c] = -sum_{k>=i} Uinv[i,k] U12[k,c]
__global__ __launch_bounds__(256) void k_wz(const int32_t *__restrict__ tasks, FrontArrays A,
                                            const zd *__restrict__ pool, int64_t psz, zd *__restrict__ inv,
                                            int64_t isz) {
  const int t = tasks[4 * blockIdx.x], kind = tasks[4 * blockIdx.x + 1];
  const int r0 = tasks[4 * blockIdx.x + 2], c0 = tasks[4 * blockIdx.x + 3];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  const zd *F = pool + b * psz + A.off[t];
  zd *Ofw = inv + b * isz + A.ioff[t];
  zd *Obw = Ofw + (int64_t)f * s;
  const int M = kind ? s : A.u[t], N = kind ? A.u[t] : s;
  const int kbeg = kind ? r0 : c0;
  // K staged in rounds of GK (32 KB of LDS, four workgroups per CU); the next
  // round's operands are fetched into registers while this one is used, from
  // clamped addresses, and masked when stored to LDS (a select right after
  // the load would wait for it at once)
  __shared__ zd As[GK][TILE];
  __shared__ zd Bs[GK][TILE];
  const int tx = tid % 16, ty = tid / 16;  // rows tx + 16 i, columns ty + 16 j
  zd acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = zmk(0, 0);
  constexpr int NL = GK * TILE / 256;
  zd ra[NL], rb[NL];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int idx = tid + 256 * u, i = idx % TILE, kk = idx / TILE;
      const int rc = min(r0 + i, M - 1), kc = min(k0 + kk, s - 1);
      ra[u] = kind ? Obw[rc + (int64_t)kc * s] : F[(s + rc) + (int64_t)kc * f];
    }
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int idx = tid + 256 * u, kk = idx % GK, jj = idx / GK;
      const int ccc = min(c0 + jj, N - 1), kc = min(k0 + kk, s - 1);
      rb[u] = kind ? F[kc + (int64_t)(s + ccc) * f] : Ofw[kc + (int64_t)ccc * f];
    }
  };
  if (kbeg < s) fetch(kbeg);
  for (int k0 = kbeg; k0 < s; k0 += GK) {
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int idx = tid + 256 * u, i = idx % TILE, kk = idx / TILE;
      As[kk][i] = (r0 + i < M && k0 + kk < s) ? ra[u] : zmk(0, 0);
    }
#pragma unroll
    for (int u = 0; u < NL; ++u) {
      const int idx = tid + 256 * u, kk = idx % GK, jj = idx / GK;
      Bs[kk][jj] = (c0 + jj < N && k0 + kk < s) ? rb[u] : zmk(0, 0);
    }
    __syncthreads();
    if (k0 + GK < s) fetch(k0 + GK);
#pragma unroll 2
    for (int kk = 0; kk < GK; ++kk) {
      zd a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = zfma(a[i], bb[j], acc[i][j]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int cc = c0 + ty + 16 * j;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int rr = r0 + tx + 16 * i;
      if (rr < M && cc < N) {
        const zd v = zmk(-acc[i][j].x, -acc[i][j].y);
        if (kind) Obw[rr + (int64_t)(s + cc) * s] = v;
        else Ofw[(s + rr) + (int64_t)cc * f] = v;
      }
    }
  }
}

void lu_invert(DevMesh &dm, int nb, const zd *pool, zd *inv, hipStream_t s) {
  const Symbolic &sy = dm.host.sym;
  FrontArrays A = fa(dm);
  if (dm.tri_cnt)
    hipLaunchKernelGGL(k_triinv, dim3((unsigned)dm.tri_cnt, nb), dim3(256), 0, s, dm.tasks.p + dm.tri_off, A, pool,
                       sy.pool, inv, dm.isz);
  EMU_LAUNCH_CHECK();
  if (dm.wz_cnt)
    hipLaunchKernelGGL(k_wz, dim3((unsigned)dm.wz_cnt, nb), dim3(256), 0, s, dm.tasks.p + dm.wz_off, A, pool,
                       sy.pool, inv, dm.isz);
  EMU_LAUNCH_CHECK();
}

// ---------------------------------------------------------------- solves
//
// Rows of a front are cut into 64-row tiles, one per wavefront.  A workgroup
// (4 waves) takes `tpw` consecutive tiles of one front (task (front, r0,
// tpw), tpw = 4, 2 or 1 chosen per front on the host); the 4 / tpw waves that
// share a tile split its column range and add their partial sums through
// LDS.  Small-front levels thus keep all waves busy on distinct rows, the
// few large fronts of the top levels get the column split.  The front's
// input vector is staged through LDS in chunks of KC entries, gathered on
// the fly (forward: b at the pivot rows in pivot order plus the children's
// update vectors; backward: y at the own rows, x at the ancestors' rows).
// Structural zeros of the triangular inverses are skipped per tile.  Tasks
// of one level never write what another reads, so a level is one launch.

constexpr int KC = 64;  // staged input chunk

// Storage of the inverted factor read by the solves: complex128 (zd), or
// complex64 (float2, the mixed-precision form), widened to double on load;
// the arithmetic is FP64 either way.
template <typename FT>
struct FactorLd;
template <>
struct FactorLd<zd> {
  static constexpr uint32_t kBytes = 16;
  static __device__ __forceinline__ zd ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    return buf_ld_stream(r, voff, soff);
  }
};
template <>
struct FactorLd<float2> {
  static constexpr uint32_t kBytes = 8;
  static __device__ __forceinline__ zd ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 v = __builtin_bit_cast(f2, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 2));
    return zmk((double)v.x, (double)v.y);
  }
};

template <int RC, typename FT>
__global__ __launch_bounds__(256) void k_fwd(const SolveTask *__restrict__ tasks, const int32_t *__restrict__ cmap,
                                             const FT *__restrict__ inv, int64_t isz, const int4 *__restrict__ gx,
                                             int64_t npiv, int nrhs, int64_t n, int64_t nw,
                                             const zd *__restrict__ bv, zd *__restrict__ y, zd *__restrict__ wbuf,
                                             zd *__restrict__ part, int nslot, int span) {
  const SolveTask T = tasks[blockIdx.x];
  const int b = blockIdx.y;
  const int q0 = blockIdx.z * RC;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: column offsets in SGPRs
  const int ks = 4 / T.tpw, tile = wave / ks, kp = wave - tile * ks;
  const int s = T.s, f = T.f;
  const FT *Fw = inv + b * isz + T.ioff;
  constexpr uint32_t EB = FactorLd<FT>::kBytes;
  const int4 *gi = gx + b * npiv + T.pivoff;
  const int nq = min(RC, nrhs - q0);
  const int64_t col0 = (int64_t)b * nrhs + q0;
  const int row0 = T.r0 + 64 * tile;  // first row of this wave's tile
  const int row = row0 + lane;
  const int rowc = min(row, f - 1);   // clamped: addresses stay inside the front
  const uint32_t rowb = (uint32_t)rowc * EB;  // byte offset of the row in a column
  // columns of this task: the workgroup's range, cut to [kb, kb + span)
  // for a split task; lower-triangular rows need k <= row
  int kbeg = 0, kend = min(s, T.r0 + 64 * T.tpw);
  if (T.kb >= 0) {
    kbeg = T.kb;
    kend = min(kend, T.kb + span);
  }
  const int kend_w = row0 < f ? min(kend, min(s, row0 + 64)) : 0;
  // child update entries of this row (static), added by the epilogue on the
  // update rows; requested now so that the epilogue waits for one load only
  const int2 cr = reinterpret_cast<const int2 *>(cmap)[T.fvoff + rowc];
  __shared__ zd xin[KC][RC];
  __shared__ zd red[3][64][RC];
  zd acc[RC];
#pragma unroll
  for (int q = 0; q < RC; ++q) acc[q] = zmk(0, 0);
  // Load order per chunk: gather records, then this wave's first 8 columns,
  // then the input values (which depend on the records); the column loads are
  // thus in flight during the gather, and each later batch of 8 is requested
  // before the previous one is used.  All loads are unconditional from
  // clamped addresses, unused values zeroed or discarded afterwards.
  for (int kc = kbeg; kc < kend; kc += KC) {
    const int k = tid & (KC - 1), qg = tid / KC;
    const int kk = kc + k;
    const bool okk = kk < kend;
    const int4 g = gi[okk ? kk : kbeg];
    const int kl = min(KC, kend_w - kc);
    const int nk = kl > kp ? (kl - kp + ks - 1) / ks : 0;  // this wave's columns kp, kp + ks, ...
    // column c of the wave's rows: byte c f EB + rowb of the front's block
    const __amdgpu_buffer_rsrc_t Fcol = buf_view(Fw, (uint32_t)f * (uint32_t)s * EB);
    // column indices clamped into [0, s): an idle wave (nk = 0) reads a
    // valid column and discards it, so the loads need no branch (a branch
    // would make the gather below wait for them)
    const int cbase = kc + kp, clast = min(kc + kp + max(nk - 1, 0) * ks, s - 1);
    zd a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = FactorLd<FT>::ld(Fcol, rowb, (uint32_t)(min(cbase + u * ks, clast) * f) * EB);
    constexpr int NG = (RC + 3) / 4;  // right-hand sides gathered per thread
    zd gv[NG], g0[NG], g1[NG];
#pragma unroll
    for (int u = 0; u < NG; ++u) {
      const int64_t cq = col0 + min(min(qg + 4 * u, RC - 1), nq - 1);
      gv[u] = bv[cq * n + g.x];
      g0[u] = wbuf[cq * nw + max(g.y, 0)];
      g1[u] = wbuf[cq * nw + max(g.z, 0)];
    }
#pragma unroll
    for (int u = 0; u < NG; ++u) {
      const int qq = qg + 4 * u;
      // selects, not branches: a branch would let the compiler sink the
      // load into it and wait there for every load in flight
      const zd v = zadd(zadd(gv[u], g.y >= 0 ? g0[u] : zmk(0, 0)), g.z >= 0 ? g1[u] : zmk(0, 0));
      if (qq < RC) xin[k][qq] = (okk && qq < nq) ? v : zmk(0, 0);
    }
    __syncthreads();
    // full batches: the next batch is requested before this one is used (no
    // branch around the loads: a branch join would wait for them at once)
    int t0 = 0;
    for (; t0 + 8 < nk; t0 += 8) {
      zd nx[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) nx[u] = FactorLd<FT>::ld(Fcol, rowb, (uint32_t)(min(cbase + (t0 + 8 + u) * ks, clast) * f) * EB);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kx = kp + (t0 + u) * ks;
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zfma(a[u], xin[kx][q], acc[q]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = nx[u];
    }
    // the last, possibly partial, batch
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t0 + u < nk) {
        const int kx = kp + (t0 + u) * ks;
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zfma(a[u], xin[kx][q], acc[q]);
      }
    __syncthreads();
  }
  if (ks > 1) {
    if (kp > 0)
#pragma unroll
      for (int q = 0; q < RC; ++q) red[wave - 1][lane][q] = acc[q];
    __syncthreads();
    if (kp == 0)
      for (int j = 1; j < ks; ++j)
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zadd(acc[q], red[wave + j - 1][lane][q]);
  }
  if (kp != 0) return;
  if (T.slot >= 0) {  // split task: partial sums of the tile, added by k_fwd_red
#pragma unroll
    for (int q = 0; q < RC; ++q)
      if (q < nq) part[((col0 + q) * nslot + T.slot) * 64 + lane] = acc[q];
    return;
  }
  if (row < f) {
    if (row < s) {
#pragma unroll
      for (int q = 0; q < RC; ++q)
        if (q < nq) y[(col0 + q) * n + T.fb + row] = acc[q];
    } else {
      zd w0[RC], w1[RC];
#pragma unroll
      for (int q = 0; q < RC; ++q) {
        const int64_t cq = col0 + min(q, nq - 1);
        w0[q] = wbuf[cq * nw + max(cr.x, 0)];
        w1[q] = wbuf[cq * nw + max(cr.y, 0)];
      }
#pragma unroll
      for (int q = 0; q < RC; ++q) {
        const zd v = zadd(zadd(acc[q], cr.x >= 0 ? w0[q] : zmk(0, 0)), cr.y >= 0 ? w1[q] : zmk(0, 0));
        if (q < nq) wbuf[(col0 + q) * nw + T.uoff + (row - s)] = v;
      }
    }
  }
}

// Sum of the partial sums of a split forward tile (task: front, r0, first
// slot, count; one lane per row, the four waves take every fourth
// right-hand side), then the forward epilogue of k_fwd.  The slot loads of a
// right-hand side are issued eight at a time from clamped slots and added in
// slot order.
template <int RC>
__global__ __launch_bounds__(256) void k_fwd_red(const int32_t *__restrict__ tasks, FrontArrays A, int nrhs,
                                                 int64_t n, int64_t nw, const zd *__restrict__ part, int nslot,
                                                 zd *__restrict__ y, zd *__restrict__ wbuf) {
  const int32_t *tk = tasks + 4 * blockIdx.x;
  const int t = tk[0], r0 = tk[1], p0 = tk[2], np = tk[3];
  const int b = blockIdx.y, q0 = blockIdx.z * RC;
  const int lane = threadIdx.x & 63, qg = threadIdx.x >> 6;
  const int s = A.s[t], f = A.f[t], fb = A.b[t];
  const int32_t *cm = A.cmap + 2 * A.fvoff[t];
  const int64_t uo = A.uoff[t];
  const int nq = min(RC, nrhs - q0);
  const int64_t col0 = (int64_t)b * nrhs + q0;
  const int row = r0 + lane;
  if (row >= f) return;
  const int c0i = row >= s ? cm[2 * row] : -1, c1i = row >= s ? cm[2 * row + 1] : -1;
#pragma unroll
  for (int u = 0; u < (RC + 3) / 4; ++u) {
    const int q = qg + 4 * u;
    if (q >= nq) break;
    const zd *pp = part + ((col0 + q) * nslot + p0) * 64 + lane;
    zd v = zmk(0, 0);
    for (int c = 0; c < np; c += 8) {
      zd h[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = pp[(int64_t)min(c + i, np - 1) * 64];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (c + i < np) v = zadd(v, h[i]);
    }
    if (row < s) {
      y[(col0 + q) * n + fb + row] = v;
    } else {
      const zd w0 = wbuf[(col0 + q) * nw + max(c0i, 0)], w1 = wbuf[(col0 + q) * nw + max(c1i, 0)];
      if (c0i >= 0) v = zadd(v, w0);
      if (c1i >= 0) v = zadd(v, w1);
      wbuf[(col0 + q) * nw + uo + (row - s)] = v;
    }
  }
}

template <int RC, typename FT>
__global__ __launch_bounds__(256) void k_bwd(const SolveTask *__restrict__ tasks, const int32_t *__restrict__ uidx,
                                             const FT *__restrict__ inv, int64_t isz, int nrhs, int64_t n,
                                             const zd *__restrict__ y, zd *__restrict__ x, int64_t sx,
                                             zd *__restrict__ part, int nslot, int span) {
  const SolveTask T = tasks[blockIdx.x];
  const int b = blockIdx.y;
  const int q0 = blockIdx.z * RC;
  const int tid = threadIdx.x, lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform: column offsets in SGPRs
  const int ks = 4 / T.tpw, tile = wave / ks, kp = wave - tile * ks;
  const int s = T.s, f = T.f;
  const FT *Bw = inv + b * isz + T.ioff + (int64_t)f * s;
  constexpr uint32_t EB = FactorLd<FT>::kBytes;
  const int32_t *ui = uidx + T.uoff;
  const int nq = min(RC, nrhs - q0);
  const int64_t col0 = (int64_t)b * nrhs + q0;
  // this instance's input y (nrhs columns of n) and output x (instance
  // stride sx: the solve may write straight into a slot of a larger array)
  const zd *yb = y + (int64_t)b * nrhs * n;
  zd *xb = x + (int64_t)b * sx;
  const int row = T.r0 + 64 * tile;  // first row of this wave's tile
  const int rowl = row + lane;
  const int rowc = min(rowl, s - 1);
  const uint32_t rowb = (uint32_t)max(rowc, 0) * EB;  // byte offset of the row in a column
  // columns: [r0, f) for the workgroup (upper-triangular rows need k >= row),
  // cut to [kb, kb + span) for a split task
  int kbeg = T.r0, kend = f;
  if (T.kb >= 0) {
    kbeg = T.kb;
    kend = min(f, T.kb + span);
  }
  __shared__ zd xin[KC][RC];
  __shared__ zd red[3][64][RC];
  zd acc[RC];
#pragma unroll
  for (int q = 0; q < RC; ++q) acc[q] = zmk(0, 0);
  // load order as in k_fwd: input index, first 8 columns, input values
  for (int kc = kbeg; kc < kend; kc += KC) {
    const int k = tid & (KC - 1), qg = tid / KC;
    const int kk = kc + k;
    const bool okk = kk < kend;
    const int kkc = okk ? kk : kbeg;
    // input row kkc: y of this front or x of an ancestor
    const int64_t off = kkc < s ? (int64_t)T.fb + kkc : (int64_t)ui[kkc - s];
    const int klo = max(0, row - kc), khi = min(KC, kend - kc);
    const int kst = klo + kp;
    const int nk = (row < s && khi > kst) ? (khi - kst + ks - 1) / ks : 0;
    // column c of the wave's rows: byte c s EB + rowb of the front's block
    const __amdgpu_buffer_rsrc_t Fcol = buf_view(Bw, (uint32_t)f * (uint32_t)s * EB);
    // column indices clamped into [0, f) (see k_fwd)
    const int cbase = kc + kst, clast = min(kc + kst + max(nk - 1, 0) * ks, f - 1);
    zd a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = FactorLd<FT>::ld(Fcol, rowb, (uint32_t)(min(cbase + u * ks, clast) * s) * EB);
    const zd *src = kkc < s ? yb : xb;
    constexpr int NG = (RC + 3) / 4;
    zd gv[NG];
#pragma unroll
    for (int u = 0; u < NG; ++u) gv[u] = src[(q0 + min(min(qg + 4 * u, RC - 1), nq - 1)) * n + off];
#pragma unroll
    for (int u = 0; u < NG; ++u) {
      const int qq = qg + 4 * u;
      if (qq < RC) xin[k][qq] = (okk && qq < nq) ? gv[u] : zmk(0, 0);
    }
    __syncthreads();
    // full batches: the next batch is requested before this one is used (no
    // branch around the loads: a branch join would wait for them at once)
    int t0 = 0;
    for (; t0 + 8 < nk; t0 += 8) {
      zd nx[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) nx[u] = FactorLd<FT>::ld(Fcol, rowb, (uint32_t)(min(cbase + (t0 + 8 + u) * ks, clast) * s) * EB);
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int kx = kst + (t0 + u) * ks;
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zfma(a[u], xin[kx][q], acc[q]);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) a[u] = nx[u];
    }
    // the last, possibly partial, batch
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (t0 + u < nk) {
        const int kx = kst + (t0 + u) * ks;
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zfma(a[u], xin[kx][q], acc[q]);
      }
    __syncthreads();
  }
  if (ks > 1) {
    if (kp > 0)
#pragma unroll
      for (int q = 0; q < RC; ++q) red[wave - 1][lane][q] = acc[q];
    __syncthreads();
    if (kp == 0)
      for (int j = 1; j < ks; ++j)
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zadd(acc[q], red[wave + j - 1][lane][q]);
  }
  if (kp != 0) return;
  if (T.slot >= 0) {  // split task: partial sums, added by k_bwd_red
#pragma unroll
    for (int q = 0; q < RC; ++q)
      if (q < nq) part[((col0 + q) * nslot + T.slot) * 64 + lane] = acc[q];
    return;
  }
  if (rowl < s)
#pragma unroll
    for (int q = 0; q < RC; ++q)
      if (q < nq) xb[(int64_t)(q0 + q) * n + T.fb + rowl] = acc[q];
}

template <int RC>
__global__ __launch_bounds__(256) void k_bwd_red(const int32_t *__restrict__ tasks, FrontArrays A, int nrhs,
                                                 int64_t n, const zd *__restrict__ part, int nslot,
                                                 zd *__restrict__ x, int64_t sx) {
  const int32_t *tk = tasks + 4 * blockIdx.x;
  const int t = tk[0], r0 = tk[1], p0 = tk[2], np = tk[3];
  const int b = blockIdx.y, q0 = blockIdx.z * RC;
  const int lane = threadIdx.x & 63, qg = threadIdx.x >> 6;
  const int s = A.s[t], fb = A.b[t];
  const int nq = min(RC, nrhs - q0);
  const int64_t col0 = (int64_t)b * nrhs + q0;
  const int row = r0 + lane;
  if (row >= s) return;
#pragma unroll
  for (int u = 0; u < (RC + 3) / 4; ++u) {
    const int q = qg + 4 * u;
    if (q >= nq) break;
    const zd *pp = part + ((col0 + q) * nslot + p0) * 64 + lane;
    zd v = zmk(0, 0);
    for (int c = 0; c < np; c += 8) {
      zd h[8];
#pragma unroll
      for (int i = 0; i < 8; ++i) h[i] = pp[(int64_t)min(c + i, np - 1) * 64];
#pragma unroll
      for (int i = 0; i < 8; ++i)
        if (c + i < np) v = zadd(v, h[i]);
    }
    x[(int64_t)b * sx + (int64_t)(q0 + q) * n + fb + row] = v;
  }
}

// RHS columns per workgroup: 8 when the level has enough workgroups to fill
// the GPU; fewer (more workgroups, each re-reading the front from cache) on
// the top levels whose few large fronts would otherwise leave most CUs idle
static int pick_rc(int nrhs, int64_t tasks, int nb) {
  int rc = 8;
  while (rc > nrhs) rc >>= 1;
  while (rc > 1 && tasks * nb * ((nrhs + rc - 1) / rc) < 1024) rc >>= 1;
  return rc;
}

template <int RC, typename FT>
static void launch_fwd(const SolveTasks &T, DevMesh &dm, FrontArrays A, int nb, int nrhs, const FT *inv,
                       const int32_t *gx, const zd *bnew, zd *y, zd *w, zd *part, hipStream_t s, KTime kt) {
  const Symbolic &sy = dm.host.sym;
  const unsigned nz = (unsigned)((nrhs + RC - 1) / RC);
  const int64_t nw = std::max<int64_t>(sy.nw, 1);
  hipExtLaunchKernelGGL(k_fwd<RC, FT>, dim3((unsigned)T.cnt, nb, nz), dim3(256), 0, s, kt.a, kt.b, 0,
                        reinterpret_cast<const SolveTask *>(dm.tasks.p + T.off), dm.cmap.p, inv, dm.isz,
                        reinterpret_cast<const int4 *>(gx), sy.npiv, nrhs, sy.n, nw, bnew, y, w, part,
                        (int)T.nslot, (int)std::min<int64_t>(T.span, INT32_MAX));
  if (T.rcnt > 0)
    hipLaunchKernelGGL(k_fwd_red<RC>, dim3((unsigned)T.rcnt, nb, nz), dim3(256), 0, s, dm.tasks.p + T.roff, A, nrhs,
                       sy.n, nw, part, (int)T.nslot, y, w);
}

template <int RC, typename FT>
static void launch_bwd(const SolveTasks &T, DevMesh &dm, FrontArrays A, int nb, int nrhs, const FT *inv,
                       const zd *y, zd *x, int64_t sx, zd *part, hipStream_t s, KTime kt) {
  const Symbolic &sy = dm.host.sym;
  const unsigned nz = (unsigned)((nrhs + RC - 1) / RC);
  hipExtLaunchKernelGGL(k_bwd<RC, FT>, dim3((unsigned)T.cnt, nb, nz), dim3(256), 0, s, kt.a, kt.b, 0,
                        reinterpret_cast<const SolveTask *>(dm.tasks.p + T.off), dm.uidx.p, inv, dm.isz, nrhs,
                        sy.n, y, x, sx, part, (int)T.nslot, (int)std::min<int64_t>(T.span, INT32_MAX));
  if (T.rcnt > 0)
    hipLaunchKernelGGL(k_bwd_red<RC>, dim3((unsigned)T.rcnt, nb, nz), dim3(256), 0, s, dm.tasks.p + T.roff, A, nrhs,
                       sy.n, part, (int)T.nslot, x, sx);
}

int64_t solve_ws(const DevMesh &dm) { return std::max<int64_t>(dm.host.sym.nw, 1) + dm.part_rhs; }

// The split task lists for the Krylov solves (one block of at most 8
// right-hand sides: the grid has no RHS dimension to fill the GPU with); the
// plain ones for many right-hand sides (dense solves of the cascade), where
// RHS chunks already give a large grid and partial sums would only add
// traffic.
static const SolveTasks &pick_tasks(const SolveTasks *v, int nrhs) { return nrhs <= 8 ? v[1] : v[0]; }

template <typename FT>
static void lu_solve_t(DevMesh &dm, int nb, int nrhs, const FT *inv, const int32_t *gx, const zd *bnew, zd *y,
                       zd *x, zd *w, hipStream_t s, Profiler *prof, int64_t sx) {
  FrontArrays A = fa(dm);
  if (sx < 0) sx = (int64_t)nrhs * dm.host.sym.n;
  const double vecs = (double)nb * nrhs;
  const double fscale = FactorLd<FT>::kBytes / 16.0;  // factor bytes per entry relative to complex128
  // partial sums of split tasks live behind the update vectors in w
  zd *part = w + (int64_t)nb * nrhs * std::max<int64_t>(dm.host.sym.nw, 1);
  for (const LevelSched &L : dm.sched) {
    const SolveTasks &T = pick_tasks(L.fwd, nrhs);
    const int rc = pick_rc(nrhs, T.cnt, nb);
    KTime kt = prof ? prof->time(0, L.fwd_mat * fscale * nb + L.fwd_vec * vecs) : KTime{};
    switch (rc) {
      case 8: launch_fwd<8>(T, dm, A, nb, nrhs, inv, gx, bnew, y, w, part, s, kt); break;
      case 4: launch_fwd<4>(T, dm, A, nb, nrhs, inv, gx, bnew, y, w, part, s, kt); break;
      case 2: launch_fwd<2>(T, dm, A, nb, nrhs, inv, gx, bnew, y, w, part, s, kt); break;
      default: launch_fwd<1>(T, dm, A, nb, nrhs, inv, gx, bnew, y, w, part, s, kt); break;
    }
  }
  EMU_LAUNCH_CHECK();
  for (auto it = dm.sched.rbegin(); it != dm.sched.rend(); ++it) {
    const SolveTasks &T = pick_tasks(it->bwd, nrhs);
    const int rc = pick_rc(nrhs, T.cnt, nb);
    KTime kt = prof ? prof->time(1, it->bwd_mat * fscale * nb + it->bwd_vec * vecs) : KTime{};
    switch (rc) {
      case 8: launch_bwd<8>(T, dm, A, nb, nrhs, inv, y, x, sx, part, s, kt); break;
      case 4: launch_bwd<4>(T, dm, A, nb, nrhs, inv, y, x, sx, part, s, kt); break;
      case 2: launch_bwd<2>(T, dm, A, nb, nrhs, inv, y, x, sx, part, s, kt); break;
      default: launch_bwd<1>(T, dm, A, nb, nrhs, inv, y, x, sx, part, s, kt); break;
    }
  }
  EMU_LAUNCH_CHECK();
}

void lu_solve(DevMesh &dm, int nb, int nrhs, const zd *inv, const int32_t *gx, const zd *bnew, zd *y, zd *x,
              zd *w, hipStream_t s, Profiler *prof, int64_t sx) {
  lu_solve_t<zd>(dm, nb, nrhs, inv, gx, bnew, y, x, w, s, prof, sx);
}

void lu_solve(DevMesh &dm, int nb, int nrhs, const float2 *inv32, const int32_t *gx, const zd *bnew, zd *y, zd *x,
              zd *w, hipStream_t s, Profiler *prof, int64_t sx) {
  lu_solve_t<float2>(dm, nb, nrhs, inv32, gx, bnew, y, x, w, s, prof, sx);
}

// complex64 copy of the inverted factor (round to nearest)
__global__ void k_narrow(int64_t n, const zd *__restrict__ src, float2 *__restrict__ dst) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const zd v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}

void lu_narrow(const DevMesh &dm, int nb, const zd *inv, float2 *inv32, hipStream_t s) {
  const int64_t n = (int64_t)nb * dm.isz;
  if (n <= 0) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_narrow, dim3(g), dim3(256), 0, s, n, inv, inv32);
  EMU_LAUNCH_CHECK();
}

// dst[b][c][new] = src[b][c][perm[new]]  (to_new) or dst[b][c][perm[new]] = src[b][c][new]
__global__ void k_permute_vec(int64_t n, int ncol, const int32_t *__restrict__ perm,
                              const zd *__restrict__ src, zd *__restrict__ dst, int to_new) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t bc = blockIdx.y;
  if (i >= n) return;
  const zd *sp = src + bc * n;
  zd *dp = dst + bc * n;
  if (to_new) dp[i] = sp[perm[i]];
  else dp[perm[i]] = sp[i];
}

void launch_permute_vec(DevMesh &dm, int nb, int ncol, const zd *src, zd *dst, bool to_new,
                        hipStream_t s) {
  int64_t n = dm.host.neq;
  dim3 g((unsigned)((n + 255) / 256), nb * ncol);
  hipLaunchKernelGGL(k_permute_vec, g, dim3(256), 0, s, n, ncol, dm.perm.p, src, dst, to_new ? 1 : 0);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu
