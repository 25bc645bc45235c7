// Dense / vector kernels of the Krylov-Schur eigensolver and of the overlap
// contractions, batched over instances (blockIdx.y / blockIdx.z = instance).
#include <map>
#include <mutex>
#include <string>

#include "engine.h"

namespace emu {

// --------------------------------------------------------------- SpMM (CSR)
// y[b][c][r] = sum_k A[b][r][k] x[b][c][k] for a block of SM_C columns.
// Eight lanes share a row: each lane walks every eighth nonzero (coalesced
// value / index reads, one read of the row per column block), accumulates
// the SM_C column sums, and the eight partial sums are combined with
// cross-lane shuffles; lane j then writes column j.  A wavefront covers 8
// rows, a workgroup 32.  Workgroups are dealt to the 8 XCDs round-robin by
// launch index, so the kernel maps launch index p to row block
// (p % 8) * per + p / 8: each XCD works through one contiguous eighth of the
// rows and, the rows being in nested-dissection order, gathers x from a
// compact range that stays in its own L2 instead of every L2 fetching all of
// x.
constexpr int SM_C = 8;

__device__ __forceinline__ zd shfl_xor_z(zd v, int m) {
  return zmk(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

__global__ __launch_bounds__(256) void k_spmm(int64_t n, int ncol, const int64_t *__restrict__ rp,
                                              const int32_t *__restrict__ ci, const zd *__restrict__ vals,
                                              int64_t vstride, const zd *__restrict__ x, int64_t xstride,
                                              zd *__restrict__ y, int64_t ystride, int per) {
  const int lane8 = threadIdx.x & 7;
  const int64_t blk = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int64_t r = blk * 32 + (threadIdx.x >> 3);
  const int b = blockIdx.y;
  const int c0 = blockIdx.z * SM_C;
  const int nc = min(SM_C, ncol - c0);
  const zd *v = vals + b * vstride;
  const zd *xb = x + b * xstride + (int64_t)c0 * n;
  zd acc[SM_C];
#pragma unroll
  for (int c = 0; c < SM_C; ++c) acc[c] = zmk(0, 0);
  const int64_t k0 = r < n ? rp[r] : 0, k1 = r < n ? rp[r + 1] : 0;
  if (k1 > k0) {  // rows >= n and empty rows read nothing
    // The value and column index of the lane's next nonzero are requested
    // before this one's gathers (indices clamped to the row), so each step
    // waits for one round of loads; the gathers are unconditional (clamped
    // column, unused columns multiply zero) so that all are in flight together.
    int64_t k = k0 + lane8;
    const int64_t klast = max(k1 - 1, k0);
    zd a = ld_stream(v + min(k, klast));
    int64_t col = ld_stream(ci + min(k, klast));
    for (; k < k1; k += 8) {
      const zd an = ld_stream(v + min(k + 8, klast));
      const int64_t coln = ld_stream(ci + min(k + 8, klast));
      zd xv[SM_C];
#pragma unroll
      for (int c = 0; c < SM_C; ++c) xv[c] = xb[(int64_t)min(c, nc - 1) * n + col];
      // keep the scheduler from interleaving each gather with its use (it
      // otherwise reuses one register quad and waits for every load in turn)
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = 0; c < SM_C; ++c) acc[c] = zfma(c < nc ? a : zmk(0, 0), xv[c], acc[c]);
      a = an;
      col = coln;
    }
  }
#pragma unroll
  for (int c = 0; c < SM_C; ++c) {
    acc[c] = zadd(acc[c], shfl_xor_z(acc[c], 1));
    acc[c] = zadd(acc[c], shfl_xor_z(acc[c], 2));
    acc[c] = zadd(acc[c], shfl_xor_z(acc[c], 4));
  }
  zd out = acc[0];
#pragma unroll
  for (int c = 1; c < SM_C; ++c)
    if (lane8 == c) out = acc[c];
  if (r < n && lane8 < nc) y[b * ystride + (int64_t)(c0 + lane8) * n + r] = out;
}

// ------------------------------------------ SpMM with many columns (modes)
// Post-processing multiplies M (or K - shift M) by all nval mode vectors at
// once.  The mode block is first transposed to row-major (xi[col][c], one
// pass through LDS tiles), so that a nonzero's nval values are one
// contiguous run; a wavefront then owns a row and its lanes the columns
// (lane, lane + 64, ...), reads the row's values and column indices once
// (wave-uniform) and streams the column runs coalesced, four nonzeros at a
// time.  A workgroup's 8 rows are gathered in LDS and written back column by
// column (128-byte runs).  Column chunks of 256 (grid z) bound the LDS tile.
constexpr int SR_ROWS = 8, SR_COLS = 256;

__global__ __launch_bounds__(256) void k_transpose_rm(int64_t n, int ncol, const zd *__restrict__ X, int64_t ldx,
                                                      int64_t sX, zd *__restrict__ xi, int64_t sI) {
  __shared__ zd T[32][33];
  const int b = blockIdx.z;
  const int64_t r0 = (int64_t)blockIdx.x * 32;
  const int c0 = blockIdx.y * 32;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  const zd *x = X + b * sX;
  for (int k = ty; k < 32; k += 8) {
    const int c = c0 + k;
    const int64_t r = r0 + tx;
    if (c < ncol && r < n) T[k][tx] = x[r + (int64_t)c * ldx];
  }
  __syncthreads();
  zd *o = xi + b * sI;
  for (int k = ty; k < 32; k += 8) {
    const int64_t r = r0 + k;
    const int c = c0 + tx;
    if (c < ncol && r < n) o[r * ncol + c] = T[tx][k];
  }
}

__global__ __launch_bounds__(256) void k_spmm_rows(int64_t n, int ncol, const int64_t *__restrict__ rp,
                                                   const int32_t *__restrict__ ci, const zd *__restrict__ vals,
                                                   int64_t vstride, const zd *__restrict__ xi, int64_t sI,
                                                   zd *__restrict__ y, int64_t ldy, int64_t sY) {
  __shared__ zd Ys[SR_ROWS][SR_COLS + 1];
  const int b = blockIdx.y;
  const int c0 = blockIdx.z * SR_COLS;
  const int nc = min(SR_COLS, ncol - c0);
  const int64_t rbase = (int64_t)blockIdx.x * SR_ROWS;
  const int tid = threadIdx.x, lane = tid & 63;
  const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
  const zd *v = vals + b * vstride;
  const zd *xb = xi + b * sI + c0;
  for (int i = 0; i < SR_ROWS / 4; ++i) {
    const int rr = (SR_ROWS / 4) * w + i;
    const int64_t r = rbase + rr;
    zd acc[SR_COLS / 64];
#pragma unroll
    for (int q = 0; q < SR_COLS / 64; ++q) acc[q] = zmk(0, 0);
    if (r < n) {
      const int64_t k0 = rp[r], k1 = rp[r + 1];
      // four nonzeros per step: their column runs are all requested before
      // the first is used
      for (int64_t kb = k0; kb < k1; kb += 4) {
        zd a[4], xv[4][SR_COLS / 64];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int64_t k = min(kb + u, k1 - 1);
          a[u] = kb + u < k1 ? v[k] : zmk(0, 0);
          const zd *xr = xb + (int64_t)ci[k] * ncol;
#pragma unroll
          for (int q = 0; q < SR_COLS / 64; ++q) xv[u][q] = xr[min(lane + 64 * q, nc - 1)];
        }
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
          for (int q = 0; q < SR_COLS / 64; ++q) acc[q] = zfma(a[u], xv[u][q], acc[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < SR_COLS / 64; ++q) Ys[rr][lane + 64 * q] = acc[q];
  }
  __syncthreads();
  zd *yb = y + b * sY + (int64_t)c0 * ldy;
  for (int idx = tid; idx < SR_ROWS * nc; idx += 256) {
    const int c = idx / SR_ROWS, rr = idx % SR_ROWS;
    const int64_t r = rbase + rr;
    if (r < n) yb[(int64_t)c * ldy + r] = Ys[rr][c];
  }
}

void launch_spmm(int nb, int64_t n, int ncol, const int64_t *rp, const int32_t *ci, const zd *vals,
                 int64_t vstride, const zd *x, int64_t xstride, zd *y, int64_t ystride, hipStream_t s, KTime kt,
                 zd *xi) {
  if (xi && ncol > 8) {  // many columns: row-major copy, then one wavefront per row
    dim3 gt((unsigned)((n + 31) / 32), (unsigned)((ncol + 31) / 32), nb);
    hipLaunchKernelGGL(k_transpose_rm, gt, dim3(256), 0, s, n, ncol, x, n, xstride, xi, n * ncol);
    EMU_LAUNCH_CHECK();
    dim3 g((unsigned)((n + SR_ROWS - 1) / SR_ROWS), nb, (unsigned)((ncol + SR_COLS - 1) / SR_COLS));
    hipExtLaunchKernelGGL(k_spmm_rows, g, dim3(256), 0, s, kt.a, kt.b, 0, n, ncol, rp, ci, vals, vstride, xi,
                          n * ncol, y, n, ystride);
    EMU_LAUNCH_CHECK();
    return;
  }
  // 8 * per launch indices cover the (n + 31) / 32 row blocks; rows >= n idle
  const int per = (int)(((n + 31) / 32 + 7) / 8);
  dim3 g((unsigned)(8 * per), nb, (unsigned)((ncol + SM_C - 1) / SM_C));
  hipExtLaunchKernelGGL(k_spmm, g, dim3(256), 0, s, kt.a, kt.b, 0, n, ncol, rp, ci, vals, vstride, x, xstride, y,
                        ystride, per);
  EMU_LAUNCH_CHECK();
}


// ------------------------------------------- SpMM on a row-interleaved block
// The Krylov loop multiplies M by one 8-column block per step.  Gathering
// x[col] of eight columns stored n apart touches eight cache lines per
// nonzero; with the block stored row-interleaved (xi[8 col + c]) the eight
// values of a nonzero's column are one 128-byte line, read whole.  The block
// is interleaved by one pass (k_interleave8, 2 x 16 n bytes per column
// block) before the product.
__global__ __launch_bounds__(256) void k_interleave8(int64_t n, const zd *__restrict__ X, int64_t sX,
                                                     zd *__restrict__ xi, int64_t sI) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (r >= n) return;
  const zd *x = X + b * sX;
  zd v[8];
#pragma unroll
  for (int c = 0; c < 8; ++c) v[c] = ld_stream(x + (int64_t)c * n + r);
  zd *o = xi + b * sI + 8 * r;
#pragma unroll
  for (int c = 0; c < 8; ++c) o[c] = v[c];
}

__global__ __launch_bounds__(256) void k_spmm_il(int64_t n, const int64_t *__restrict__ rp,
                                                 const int32_t *__restrict__ ci, const zd *__restrict__ vals,
                                                 int64_t vstride, const zd *__restrict__ xi, int64_t xstride,
                                                 zd *__restrict__ y, int64_t ystride, int per) {
  // Eight lanes share a row and lane c owns output column c.  Per step the
  // group takes the row's next eight nonzeros (lane j loads value and column
  // of nonzero j, coalesced), then for each of them lane c loads x[col][c]:
  // the group reads one 128-byte line per nonzero, eight lines per wavefront
  // load, and accumulates without a final cross-lane reduction.
  const int lane8 = threadIdx.x & 7;
  const int gbase = threadIdx.x & 63 & ~7;  // first lane of the group in the wavefront
  const int64_t blk = (int64_t)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
  const int64_t r = blk * 32 + (threadIdx.x >> 3);
  const int b = blockIdx.y;
  const zd *v = vals + b * vstride;
  const zd *xb = xi + b * xstride;
  zd acc = zmk(0, 0);
  const int64_t k0 = r < n ? rp[r] : 0, k1 = r < n ? rp[r + 1] : 0;
  for (int64_t kb = k0; kb < k1; kb += 8) {
    const int64_t k = kb + lane8;
    const bool ok = k < k1;
    const int64_t kc = ok ? k : k1 - 1;
    zd a = ld_stream(v + kc);
    const int col = ld_stream(ci + kc);
    if (!ok) a = zmk(0, 0);
    const int cnt = (int)min((int64_t)8, k1 - kb);
    zd xv[8], av[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int cq = __shfl(col, gbase + q, 64);
      av[q] = zmk(__shfl(a.x, gbase + q, 64), __shfl(a.y, gbase + q, 64));
      xv[q] = xb[8 * (int64_t)cq + lane8];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < 8; ++q)
      if (q < cnt) acc = zfma(av[q], xv[q], acc);
  }
  if (r < n) y[b * ystride + (int64_t)lane8 * n + r] = acc;
}

void launch_spmm8_il(int nb, int64_t n, const int64_t *rp, const int32_t *ci, const zd *vals, int64_t vstride,
                     const zd *x, int64_t xstride, zd *xi, zd *y, int64_t ystride, hipStream_t s, KTime kt,
                     bool xi_ready) {
  if (!xi_ready) {
    hipLaunchKernelGGL(k_interleave8, dim3((unsigned)((n + 255) / 256), nb), dim3(256), 0, s, n, x, xstride, xi,
                       8 * n);
    EMU_LAUNCH_CHECK();
  }
  const int per = (int)(((n + 31) / 32 + 7) / 8);
  hipExtLaunchKernelGGL(k_spmm_il, dim3((unsigned)(8 * per), nb), dim3(256), 0, s, kt.a, kt.b, 0, n, rp, ci, vals,
                        vstride, xi, 8 * n, y, ystride, per);
  EMU_LAUNCH_CHECK();
}

typedef double v4d __attribute__((ext_vector_type(4)));

// -------------------------------- C = op(A)^T B, wide, on the matrix cores
// Mode overlaps and plane-wave contractions: M, N of tens to hundreds, K the
// DOF count.  A workgroup owns a 64 x 64 output tile and one K chunk; per
// 32-row step it stages the 32 x 64 tiles of A and B in LDS (the next pair is
// fetched into registers meanwhile), and wave w accumulates output columns
// 16 w .. 16 w + 15 against all 64 rows with FP64 16x16x4 MFMAs (four real
// products per complex one).  Partials per chunk: part[b][chunk][n][m].
template <bool CONJ>
__global__ __launch_bounds__(256) void k_atb_mfma_wide(int64_t K, int M, int N, int64_t kchunk,
                                                       const zd *__restrict__ A, int64_t lda, int64_t sA,
                                                       const zd *__restrict__ B, int64_t ldb, int64_t sB,
                                                       zd *__restrict__ part, int tilesM) {
  constexpr int TT = 64, TK = 32;
  // The output tiles of one (K chunk, instance) read the same rows of A and
  // B: the linear workgroup id is renumbered (a bijection of [0, T)) so that
  // each of the 8 XCDs, which take workgroups round-robin and have an L2
  // each, runs a contiguous range of the tile order.
  const unsigned T = gridDim.x * gridDim.y * gridDim.z;
  const unsigned Lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned xcd = Lin & 7u, q8 = T >> 3, r8 = T & 7u;
  const unsigned Lg = xcd * q8 + min(xcd, r8) + (Lin >> 3);
  const int bx = (int)(Lg % gridDim.x);
  const int tm = bx % tilesM, tn = bx / tilesM;
  const int m0 = tm * TT, n0 = tn * TT;
  const int chunk = (int)((Lg / gridDim.x) % gridDim.y), nchunk = gridDim.y;
  const int b = (int)(Lg / (gridDim.x * gridDim.y));
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  __shared__ zd As[TK][TT + 1];
  __shared__ zd Bs[TK][TT + 1];
  const int64_t kb = chunk * kchunk, ke = min(K, kb + kchunk);
  const int lr = tid & 31, lc = tid >> 5;  // staging: row lr, columns lc + 8 i
  zd ra[8], rb[8];
  auto fetch = [&](int64_t k0) {
    const int64_t k = k0 + lr;
    const int64_t kc = k < ke ? k : ke - 1;
#pragma unroll
    for (int i = 0; i < 8; ++i) ra[i] = Ab[kc + (int64_t)min(m0 + lc + 8 * i, M - 1) * lda];
#pragma unroll
    for (int i = 0; i < 8; ++i) rb[i] = Bb[kc + (int64_t)min(n0 + lc + 8 * i, N - 1) * ldb];
  };
  v4d pr[4], pi[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    pr[t] = (v4d){0, 0, 0, 0};
    pi[t] = (v4d){0, 0, 0, 0};
  }
  if (kb < ke) fetch(kb);
  const int fi = lane & 15, fk = lane >> 4;
  for (int64_t k0 = kb; k0 < ke; k0 += TK) {
    __syncthreads();
    const bool okk = k0 + lr < ke;
#pragma unroll
    for (int i = 0; i < 8; ++i) As[lr][lc + 8 * i] = (okk && m0 + lc + 8 * i < M) ? ra[i] : zmk(0, 0);
#pragma unroll
    for (int i = 0; i < 8; ++i) Bs[lr][lc + 8 * i] = (okk && n0 + lc + 8 * i < N) ? rb[i] : zmk(0, 0);
    __syncthreads();
    if (k0 + TK < ke) fetch(k0 + TK);
#pragma unroll
    for (int ks = 0; ks < TK / 4; ++ks) {
      const int kk = ks * 4 + fk;
      const zd xb = Bs[kk][16 * wv + fi];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const zd va = As[kk][16 * t + fi];
        pr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb.x, pr[t], 0, 0, 0);
        pi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb.y, pi[t], 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const zd va = As[kk][16 * t + fi];
        // CONJ: Re += ar br + ai bi, Im += ar bi - ai br
        pr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? va.y : -va.y, xb.y, pr[t], 0, 0, 0);
        pi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? -va.y : va.y, xb.x, pi[t], 0, 0, 0);
      }
    }
  }
  const int nn = n0 + 16 * wv + fi;
  if (nn < N) {
    zd *pp = part + (((int64_t)b * nchunk + chunk) * N + nn) * M;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 16 * t + fk + 4 * r;
        if (m < M) pp[m] = zmk(pr[t][r], pi[t][r]);
      }
  }
}

// ------------------------------------------ C = op(A)^T B for N <= 8 columns
// The Gram-Schmidt projection h = V^H w: A (K x M) is the Krylov basis, read
// once from HBM; B (K x NQ) is the new block.  A workgroup owns 64 columns of
// A and one K chunk.  Per 32-row step it stages the 32 x 64 tile of A (read
// along the rows, coalesced) and the 32 x NQ tile of B in LDS; lane c of each
// wave then accumulates column c against all NQ columns of B over the rows of
// its wave (rows w, w + 4, ...; the B reads are wave-uniform LDS broadcasts).
// The next tile is fetched into registers while the current one is used.  At
// the end the four waves' sums are combined in LDS and written as the chunk's
// partial product; k_atb_reduce adds the chunks.
// CPT columns of A per thread: a tile is 64 CPT columns x 32 / CPT rows (the
// same 2048 elements); with CPT = 2 every B value read from LDS feeds two
// columns, halving the LDS traffic per FMA for N = 16.
template <int NQ, int CPT>
__global__ __launch_bounds__(256) void k_atb_tall(int64_t K, int M, int64_t kchunk, const zd *__restrict__ A,
                                                  int64_t lda, int64_t sA, const zd *__restrict__ B, int64_t ldb,
                                                  int64_t sB, int conjA, zd *__restrict__ part) {
  static_assert(NQ >= 1 && NQ <= 16, "tall kernel covers N <= 16");
  constexpr int TM = 64 * CPT, TK = 32 / CPT, CG = 256 / TK;  // CG column groups in the staging map
  const int m0 = blockIdx.x * TM;
  const int chunk = blockIdx.y, nchunk = gridDim.y;
  const int b = blockIdx.z;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  __shared__ zd As[TK][TM + 1];
  __shared__ zd Bs[TK][NQ];
  const int64_t kb = chunk * kchunk, ke = min(K, kb + kchunk);
  const int lr = tid % TK, lc = tid / TK;  // staging: row lr, columns lc + CG i
  zd ra[8], rb[(TK * NQ + 255) / 256];
  // Loads are unconditional from clamped (in-range) addresses, and the
  // out-of-range values are zeroed afterwards: a load inside a branch makes
  // the compiler wait for it before the branch merges, which serialises the
  // tile's loads instead of keeping them all in flight.
  auto fetch = [&](int64_t k0) {
    const int64_t k = k0 + lr;
    const bool okk = k < ke;
    const int64_t kc = okk ? k : ke - 1;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int m = m0 + lc + CG * i;
      const zd v = Ab[kc + (int64_t)min(m, M - 1) * lda];
      ra[i] = (okk && m < M) ? v : zmk(0, 0);
    }
#pragma unroll
    for (int i = 0; i < (TK * NQ + 255) / 256; ++i) {
      const int idx = tid + 256 * i, r = idx % TK, q = idx / TK;
      const bool ok = q < NQ && k0 + r < ke;
      const zd v = Bb[(ok ? k0 + r : ke - 1) + (int64_t)min(q, NQ - 1) * ldb];
      rb[i] = ok ? v : zmk(0, 0);
    }
  };
  zd acc[CPT][NQ];
#pragma unroll
  for (int c = 0; c < CPT; ++c)
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[c][q] = zmk(0, 0);
  if (kb < ke) fetch(kb);
  for (int64_t k0 = kb; k0 < ke; k0 += TK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) As[lr][lc + CG * i] = conjA ? zconj(ra[i]) : ra[i];
#pragma unroll
    for (int i = 0; i < (TK * NQ + 255) / 256; ++i) {
      const int idx = tid + 256 * i;
      if (idx < TK * NQ) Bs[idx % TK][idx / TK] = rb[i];
    }
    __syncthreads();
    if (k0 + TK < ke) fetch(k0 + TK);
#pragma unroll
    for (int kk = wv; kk < TK; kk += 4) {
      zd a[CPT];
#pragma unroll
      for (int c = 0; c < CPT; ++c) a[c] = As[kk][lane + 64 * c];
#pragma unroll
      for (int q = 0; q < NQ; ++q) {
        const zd bq = Bs[kk][q];
#pragma unroll
        for (int c = 0; c < CPT; ++c) acc[c][q] = zfma(a[c], bq, acc[c][q]);
      }
    }
  }
  __syncthreads();
  // combine the four waves, 64 columns and eight columns of B at a time (LDS
  // of the A tile reused: 4 x 64 x 8 <= TK x (TM + 1))
  zd *red = &As[0][0];
  constexpr int QH = NQ < 8 ? NQ : 8;
#pragma unroll
  for (int c = 0; c < CPT; ++c)
#pragma unroll
    for (int h = 0; h < NQ; h += QH) {
#pragma unroll
      for (int q = 0; q < QH; ++q) red[(wv * 64 + lane) * QH + q] = acc[c][h + q];
      __syncthreads();
      for (int o = tid; o < 64 * QH; o += 256) {
        const int cc = o % 64, q = o / 64;
        zd sum = red[cc * QH + q];
#pragma unroll
        for (int w = 1; w < 4; ++w) sum = zadd(sum, red[(w * 64 + cc) * QH + q]);
        const int m = m0 + 64 * c + cc;
        if (m < M) part[(((int64_t)b * nchunk + chunk) * NQ + h + q) * M + m] = sum;
      }
      __syncthreads();
    }
}

// Exchange two equal ranges (the data of two batch slots).
template <typename T>
__global__ __launch_bounds__(256) void k_swap(int64_t n, T *__restrict__ a, T *__restrict__ b) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const T t = a[i];
    a[i] = b[i];
    b[i] = t;
  }
}

void launch_swap(int64_t n, zd *a, zd *b, hipStream_t s) {
  if (n <= 0 || a == b) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_swap<zd>, dim3(g), dim3(256), 0, s, n, a, b);
  EMU_LAUNCH_CHECK();
}

void launch_swap_f2(int64_t n, float2 *a, float2 *b, hipStream_t s) {
  if (n <= 0 || a == b) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_swap<float2>, dim3(g), dim3(256), 0, s, n, a, b);
  EMU_LAUNCH_CHECK();
}

void launch_swap_i32(int64_t n, int32_t *a, int32_t *b, hipStream_t s) {
  if (n <= 0 || a == b) return;
  const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_swap<int32_t>, dim3(g), dim3(256), 0, s, n, a, b);
  EMU_LAUNCH_CHECK();
}


// The 16-column projection on the matrix cores: op(A)^T B with B of 16
// columns, as four real v_mfma_f64_16x16x4f64 per 16 x 16 x 4 complex block
// (real and imaginary parts).  A workgroup stages a 32-row x 64-column tile
// of A and the 32 x 16 tile of B in LDS (as k_atb_tall); wave w multiplies
// rows 8w .. 8w + 7 of the tile (two k-steps of 4) into four 16 x 16 output
// tiles and writes its own partial product (four partials per K chunk, added
// by k_atb_reduce).  Fragment layouts (FP64 16x16x4): A[i][k] in lane
// i + 16 k, B[k][j] in lane j + 16 k, D[i][j] in lane j + 16 (i % 4),
// register i / 4.

template <bool CONJ>
// three waves per SIMD (164 registers, no spills): more loads in flight
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
void k_atb_mfma16(int64_t K, int M, int64_t kchunk, const zd *__restrict__ A,
                                                    int64_t lda, int64_t sA, const zd *__restrict__ B, int64_t ldb,
                                                    int64_t sB, zd *__restrict__ part) {
  constexpr int TM = 64, TK = 32, NQ = 16;
  // The column tiles of one (K chunk, instance) all read the same rows of B.
  // Workgroups go to the 8 XCDs round-robin by linear id, each XCD with its
  // own L2, so the linear id is renumbered to give consecutive tiles to one
  // XCD (XCD x takes a contiguous range of the tile order, in order): B's
  // rows are then fetched from HBM once per chunk instead of once per tile.
  // The renumbering is a bijection of [0, T) whatever the hardware's
  // placement, so it affects only locality.
  const unsigned T = gridDim.x * gridDim.y * gridDim.z;
  const unsigned Lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
  const unsigned xcd = Lin & 7u, q = T >> 3, r = T & 7u;
  const unsigned Lg = xcd * q + min(xcd, r) + (Lin >> 3);
  const int m0 = (int)(Lg % gridDim.x) * TM;
  const int chunk = (int)((Lg / gridDim.x) % gridDim.y), nchunk = gridDim.y;
  const int b = (int)(Lg / (gridDim.x * gridDim.y));
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  __shared__ zd As[TK][TM + 1];
  __shared__ zd Bs[TK][NQ + 1];
  const int64_t kb = chunk * kchunk, ke = min(K, kb + kchunk);
  const int lr = tid & 31, lc = tid >> 5;  // staging: row lr, columns lc + 8 i
  zd ra[8], rb[2];
  // unconditional loads from clamped addresses (see k_atb_tall); the
  // out-of-range values are zeroed when they are stored to LDS, not here: a
  // select right after the load would make the wave wait for the prefetch
  // before the matrix-core phase instead of during it
  auto fetch = [&](int64_t k0) {
    const int64_t k = k0 + lr;
    const int64_t kc = k < ke ? k : ke - 1;
#pragma unroll
    for (int i = 0; i < 8; ++i) ra[i] = Ab[kc + (int64_t)min(m0 + lc + 8 * i, M - 1) * lda];
#pragma unroll
    for (int i = 0; i < 2; ++i) rb[i] = Bb[kc + (int64_t)(lc + 8 * i) * ldb];
  };
  v4d pr[4], pi[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    pr[t] = (v4d){0, 0, 0, 0};
    pi[t] = (v4d){0, 0, 0, 0};
  }
  if (kb < ke) fetch(kb);
  const int fi = lane & 15, fk = lane >> 4;
  for (int64_t k0 = kb; k0 < ke; k0 += TK) {
    __syncthreads();
    const bool okk = k0 + lr < ke;
#pragma unroll
    for (int i = 0; i < 8; ++i) As[lr][lc + 8 * i] = (okk && m0 + lc + 8 * i < M) ? ra[i] : zmk(0, 0);
#pragma unroll
    for (int i = 0; i < 2; ++i) Bs[lr][lc + 8 * i] = okk ? rb[i] : zmk(0, 0);
    __syncthreads();
    if (k0 + TK < ke) fetch(k0 + TK);
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int kk = wv * 8 + ks * 4 + fk;
      const zd xb = Bs[kk][fi];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const zd va = As[kk][16 * t + fi];
        // CONJ: Re += ar br + ai bi, Im += ar bi - ai br; else Re += ar br - ai bi, Im += ar bi + ai br
        pr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb.x, pr[t], 0, 0, 0);
        pr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? va.y : -va.y, xb.y, pr[t], 0, 0, 0);
        pi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb.y, pi[t], 0, 0, 0);
        pi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(CONJ ? -va.y : va.y, xb.x, pi[t], 0, 0, 0);
      }
    }
  }
  const int64_t pc = ((int64_t)b * nchunk + chunk) * 4 + wv;  // this wave's partial
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int m = m0 + 16 * t + fk + 4 * r;
      if (m < M) part[(pc * NQ + fi) * M + m] = zmk(pr[t][r], pi[t][r]);
    }
}

// Gram matrix W^H W of an 8-column block (K x 8, ld ldw), partial sums per K
// chunk (layout of k_atb_tall's partials).  A 256-row tile of W is staged in
// LDS; thread (o, g) owns output o = (p, q) and sums rows g, g + 4, ... of the
// tile, so every lane does useful work (the general tall kernel would keep
// 56 of its 64 columns idle).
constexpr int GR_TK = 256;

__global__ __launch_bounds__(256) void k_gram8(int64_t K, int64_t kchunk, const zd *__restrict__ W, int64_t ldw,
                                               int64_t sW, zd *__restrict__ part) {
  constexpr int NQ = 8;
  __shared__ zd Ws[GR_TK][NQ + 1];
  const int chunk = blockIdx.x, nchunk = gridDim.x, b = blockIdx.y, tid = threadIdx.x;
  const int o = tid % (NQ * NQ), g = tid / (NQ * NQ), p = o % NQ, q = o / NQ;
  const zd *Wb = W + b * sW;
  const int64_t kb = chunk * kchunk, ke = min(K, kb + kchunk);
  zd rw[NQ];
  auto fetch = [&](int64_t k0) {
    const int64_t k = k0 + tid;
    const bool ok = k < ke;
    const int64_t kc = ok ? k : ke - 1;
#pragma unroll
    for (int c = 0; c < NQ; ++c) {
      const zd v = Wb[kc + (int64_t)c * ldw];
      rw[c] = ok ? v : zmk(0, 0);
    }
  };
  zd acc = zmk(0, 0);
  if (kb < ke) fetch(kb);
  for (int64_t k0 = kb; k0 < ke; k0 += GR_TK) {
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NQ; ++c) Ws[tid][c] = rw[c];
    __syncthreads();
    if (k0 + GR_TK < ke) fetch(k0 + GR_TK);
#pragma unroll 8
    for (int kk = g; kk < GR_TK; kk += 4) acc = zfma(zconj(Ws[kk][p]), Ws[kk][q], acc);
  }
  __syncthreads();
  zd *red = &Ws[0][0];
  red[tid] = acc;
  __syncthreads();
  if (tid < NQ * NQ) {
    const zd v = zadd(zadd(red[tid], red[tid + 64]), zadd(red[tid + 128], red[tid + 192]));
    part[(((int64_t)b * nchunk + chunk) * NQ + q) * NQ + p] = v;
  }
}

// C[b][m + nn ldc] = sum over chunks of part[b][chunk][nn][m]: 64 outputs per
// workgroup, its four waves summing every fourth chunk, combined in LDS.
__global__ __launch_bounds__(256) void k_atb_reduce(int M, int N, int nchunk, const zd *__restrict__ part,
                                                    zd *__restrict__ C, int64_t ldc, int64_t sC) {
  __shared__ zd red[4][64];
  const int b = blockIdx.y;
  const int lane = threadIdx.x & 63, g = threadIdx.x >> 6;
  const int64_t idx = blockIdx.x * 64ll + lane, tot = (int64_t)M * N;
  zd acc = zmk(0, 0);
  if (idx < tot)
    for (int c = g; c < nchunk; c += 4) acc = zadd(acc, part[((int64_t)b * nchunk + c) * tot + idx]);
  red[g][lane] = acc;
  __syncthreads();
  if (g == 0 && idx < tot) {
    const int m = (int)(idx % M), nn = (int)(idx / M);
    C[b * sC + m + (int64_t)nn * ldc] = zadd(zadd(red[0][lane], red[1][lane]), zadd(red[2][lane], red[3][lane]));
  }
}

void launch_atb(int nb, int64_t K, int M, int N, const zd *A, int64_t lda, int64_t sA, const zd *B,
                int64_t ldb, int64_t sB, zd *C, int64_t ldc, int64_t sC, bool conjA, zd *scratch,
                size_t scratch_elems, hipStream_t s, KTime kt) {
  if (M <= 0 || N <= 0 || nb <= 0) return;
  const bool tall = N == 1 || N == 2 || N == 4 || N == 8 || N == 16;
  if (!tall) {  // wide products (mode overlaps, J contractions): 64 x 64 MFMA tiles, split over K
    const int tilesM = (M + 63) / 64, tiles = tilesM * ((N + 63) / 64);
    const int64_t per = (int64_t)nb * M * N;
    const int64_t maxchunk = std::max<int64_t>(1, (int64_t)(scratch_elems / std::max<int64_t>(per, 1)));
    const int64_t want = std::max<int64_t>(1, 2048 / std::max(1, tiles * nb));
    int64_t nchunk = std::min<int64_t>(std::min<int64_t>(want, maxchunk), (K + 31) / 32);
    nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, 65535));
    int64_t kchunk = (K + nchunk - 1) / nchunk;
    kchunk = ((kchunk + 31) / 32) * 32;
    nchunk = (K + kchunk - 1) / kchunk;
    if ((size_t)(nchunk * per) > scratch_elems) throw HipError("atb: scratch too small");
    // the kernel renumbers the linear id of the whole grid in 32 bits
    if ((int64_t)tiles * nchunk * nb > 0xffffffffLL) throw HipError("atb: grid too large");
    dim3 g((unsigned)tiles, (unsigned)nchunk, nb);
    if (conjA)
      hipLaunchKernelGGL(k_atb_mfma_wide<true>, g, dim3(256), 0, s, K, M, N, kchunk, A, lda, sA, B, ldb, sB, scratch,
                         tilesM);
    else
      hipLaunchKernelGGL(k_atb_mfma_wide<false>, g, dim3(256), 0, s, K, M, N, kchunk, A, lda, sA, B, ldb, sB,
                         scratch, tilesM);
    EMU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_atb_reduce, dim3((unsigned)((per / nb + 63) / 64), nb), dim3(256), 0, s, M, N, (int)nchunk,
                       scratch, C, ldc, sC);
    EMU_LAUNCH_CHECK();
    return;
  }
  if (M == 8 && N == 8 && conjA && A == B && lda == ldb && sA == sB) {  // Gram matrix of a block
    const int64_t per = (int64_t)nb * 64;
    int64_t nchunk = std::max<int64_t>(1, std::min<int64_t>(1024 / std::max(nb, 1), (K + GR_TK - 1) / GR_TK));
    nchunk = std::min<int64_t>(nchunk, std::max<int64_t>(1, (int64_t)(scratch_elems / per)));
    int64_t kchunk = (K + nchunk - 1) / nchunk;
    kchunk = ((kchunk + GR_TK - 1) / GR_TK) * GR_TK;
    nchunk = (K + kchunk - 1) / kchunk;
    hipExtLaunchKernelGGL(k_gram8, dim3((unsigned)nchunk, nb), dim3(256), 0, s, kt.a, kt.b, 0, K, kchunk, A, lda, sA,
                          scratch);
    EMU_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_atb_reduce, dim3(1, nb), dim3(256), 0, s, 8, 8, (int)nchunk, scratch, C, ldc, sC);
    EMU_LAUNCH_CHECK();
    return;
  }
  if (N == 16) {  // matrix cores (four partials per K chunk, one per wave)
    const int tiles = (M + 63) / 64;
    const int64_t per = (int64_t)nb * M * N * 4;
    const int64_t maxchunk = std::max<int64_t>(1, (int64_t)(scratch_elems / std::max<int64_t>(per, 1)));
    int64_t nchunk = std::min<int64_t>(std::min<int64_t>(std::max<int64_t>(1, 2048 / std::max(1, tiles * nb)),
                                                         maxchunk), (K + 31) / 32);
    nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, 256));
    int64_t kchunk = (K + nchunk - 1) / nchunk;
    kchunk = ((kchunk + 31) / 32) * 32;
    nchunk = (K + kchunk - 1) / kchunk;
    if ((size_t)(nchunk * per) > scratch_elems) throw HipError("atb: scratch too small");
    // the kernel renumbers the linear id of the whole grid in 32 bits
    if ((int64_t)tiles * nchunk * nb > 0xffffffffLL) throw HipError("atb: grid too large");
    dim3 g((unsigned)tiles, (unsigned)nchunk, nb), t(256);
    if (conjA)
      hipExtLaunchKernelGGL(k_atb_mfma16<true>, g, t, 0, s, kt.a, kt.b, 0, K, M, kchunk, A, lda, sA, B, ldb, sB, scratch);
    else
      hipExtLaunchKernelGGL(k_atb_mfma16<false>, g, t, 0, s, kt.a, kt.b, 0, K, M, kchunk, A, lda, sA, B, ldb, sB, scratch);
    EMU_LAUNCH_CHECK();
    const int64_t tot = (int64_t)M * N;
    hipLaunchKernelGGL(k_atb_reduce, dim3((unsigned)((tot + 63) / 64), nb), dim3(256), 0, s, M, N, (int)(nchunk * 4),
                       scratch, C, ldc, sC);
    EMU_LAUNCH_CHECK();
    return;
  }
  const int cpt = 1;
  const int kstep = 32 / cpt;
  const int tiles = (M + 64 * cpt - 1) / (64 * cpt);
  int64_t per = (int64_t)nb * M * N;
  int64_t maxchunk = std::max<int64_t>(1, (int64_t)(scratch_elems / std::max<int64_t>(per, 1)));
  int64_t want = std::max<int64_t>(1, 2048 / std::max(1, tiles * nb));
  int64_t nchunk = std::min<int64_t>(std::min<int64_t>(want, maxchunk), (K + kstep - 1) / kstep);
  nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, 256));
  int64_t kchunk = (K + nchunk - 1) / nchunk;
  kchunk = ((kchunk + kstep - 1) / kstep) * kstep;
  nchunk = (K + kchunk - 1) / kchunk;
  if ((size_t)(nchunk * per) > scratch_elems) throw HipError("atb: scratch too small");
  const int cj = conjA ? 1 : 0;
  dim3 g((unsigned)tiles, (unsigned)nchunk, nb), t(256);
  switch (N) {
    case 1: hipExtLaunchKernelGGL(k_atb_tall<1, 1>, g, t, 0, s, kt.a, kt.b, 0, K, M, kchunk, A, lda, sA, B, ldb, sB, cj, scratch); break;
    case 2: hipExtLaunchKernelGGL(k_atb_tall<2, 1>, g, t, 0, s, kt.a, kt.b, 0, K, M, kchunk, A, lda, sA, B, ldb, sB, cj, scratch); break;
    case 4: hipExtLaunchKernelGGL(k_atb_tall<4, 1>, g, t, 0, s, kt.a, kt.b, 0, K, M, kchunk, A, lda, sA, B, ldb, sB, cj, scratch); break;
    default: hipExtLaunchKernelGGL(k_atb_tall<8, 1>, g, t, 0, s, kt.a, kt.b, 0, K, M, kchunk, A, lda, sA, B, ldb, sB, cj, scratch); break;
  }
  EMU_LAUNCH_CHECK();
  const int64_t tot = (int64_t)M * N;
  hipLaunchKernelGGL(k_atb_reduce, dim3((unsigned)((tot + 63) / 64), nb), dim3(256), 0, s, M, N, (int)nchunk,
                     scratch, C, ldc, sC);
  EMU_LAUNCH_CHECK();
}


// ------------------------------------------------ C = alpha A B + beta C (NN)
// Batched complex GEMM on the FP64 matrix cores, for the restart combination
// V Q, the Ritz vectors V W, H S in the delayed Gram-Schmidt and the
// star-product cascade.  A workgroup owns a 64 x 64 tile of C; per 16-deep K
// step it stages the 64 x 16 tile of A (loaded along the columns of A,
// coalesced) and the 16 x 64 tile of B in LDS (33 KB, four workgroups per
// CU), and wave w accumulates the 64 x 16 slice of columns 16 w .. 16 w + 15
// with 16x16x4 FP64 MFMAs (four real products per complex one, the two
// products of each accumulator issued apart).  The next pair of tiles is fetched into
// registers while the current one is multiplied.  Launch order runs the N
// tiles of one M tile next to each other, so A's rows are shared in L2.
template <bool BETA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3)))
void k_zgemm_nn(int64_t M, int N, int K, zd alpha, const zd *__restrict__ A,
                                                  int64_t lda, int64_t sA, const zd *__restrict__ B, int64_t ldb,
                                                  int64_t sB, zd beta, zd *__restrict__ C, int64_t ldc, int64_t sC,
                                                  int tilesN) {
  constexpr int TT = 64, TK = 16;
  const int64_t tm = blockIdx.x / tilesN;
  const int tn = blockIdx.x % tilesN;
  const int64_t m0 = tm * TT;
  const int n0 = tn * TT;
  const int b = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  __shared__ zd As[TK][TT + 1];
  __shared__ zd Bs[TK][TT + 1];
  // A tile: row am = tid & 63 of the tile, K columns ak + 4 i; B tile: K row
  // bk = tid & 15, columns bn + 16 i
  const int am = tid & 63, ak = tid >> 6;
  const int bk = tid & 15, bn = tid >> 4;
  const int64_t arow = min(m0 + am, M - 1);
  zd ra[4], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 4; ++i) ra[i] = Ab[arow + (int64_t)min(k0 + ak + 4 * i, K - 1) * lda];
    const int kb = min(k0 + bk, K - 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) rb[i] = Bb[kb + (int64_t)min(n0 + bn + 16 * i, N - 1) * ldb];
  };
  v4d pr[4], pi[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    pr[t] = (v4d){0, 0, 0, 0};
    pi[t] = (v4d){0, 0, 0, 0};
  }
  fetch(0);
  const int fi = lane & 15, fk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += TK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int kk = ak + 4 * i;
      As[kk][am] = (k0 + kk < K && m0 + am < M) ? ra[i] : zmk(0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) Bs[bk][bn + 16 * i] = (k0 + bk < K && n0 + bn + 16 * i < N) ? rb[i] : zmk(0, 0);
    __syncthreads();
    if (k0 + TK < K) fetch(k0 + TK);
#pragma unroll
    for (int ks = 0; ks < TK / 4; ++ks) {
      const int kk = ks * 4 + fk;
      const zd xb = Bs[kk][16 * wv + fi];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const zd va = As[kk][16 * t + fi];
        pr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb.x, pr[t], 0, 0, 0);
        pi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb.y, pi[t], 0, 0, 0);
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const zd va = As[kk][16 * t + fi];
        pr[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.y, xb.y, pr[t], 0, 0, 0);
        pi[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.y, xb.x, pi[t], 0, 0, 0);
      }
    }
  }
  const int nn = n0 + 16 * wv + fi;
  if (nn >= N) return;
  zd *Cb = C + b * sC + (int64_t)nn * ldc;
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t m = m0 + 16 * t + fk + 4 * r;
      if (m < M) {
        zd v = zmul(alpha, zmk(pr[t][r], pi[t][r]));
        if (BETA) v = zfma(beta, Cb[m], v);
        Cb[m] = v;
      }
    }
}

// Tall products (M = the DOF count: restart combination, Ritz vectors): a
// 128 x 64 tile per workgroup, waves in a 2 x 2 grid of 64 x 32 slices, so
// that every A fragment read from LDS feeds two column blocks (six LDS reads
// per 32 MFMAs instead of five per 16) and B is staged once per 128 rows.
template <bool BETA>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_zgemm_nn_tall(int64_t M, int N, int K, zd alpha, const zd *__restrict__ A,
                                                       int64_t lda, int64_t sA, const zd *__restrict__ B,
                                                       int64_t ldb, int64_t sB, zd beta, zd *__restrict__ C,
                                                       int64_t ldc, int64_t sC, int tilesN) {
  constexpr int TM = 128, TN = 64, TK = 16;
  // The tilesN column tiles of a row tile read the same 128 rows of A: the
  // linear workgroup id is renumbered (a bijection of [0, T)) so that each
  // of the 8 XCDs, which take workgroups round-robin and have an L2 each,
  // runs a contiguous range of the tile order, row tile by row tile.
  const unsigned T = gridDim.x * gridDim.y;
  const unsigned Lin = blockIdx.x + gridDim.x * blockIdx.y;
  const unsigned xcd = Lin & 7u, q = T >> 3, r = T & 7u;
  const unsigned Lg = xcd * q + min(xcd, r) + (Lin >> 3);
  const unsigned bx = Lg % gridDim.x;
  const int64_t tm = bx / tilesN;
  const int tn = bx % tilesN;
  const int64_t m0 = tm * TM;
  const int n0 = tn * TN;
  const int b = (int)(Lg / gridDim.x);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = wv & 1, wn = wv >> 1;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  __shared__ zd As[TK][TM + 1];
  __shared__ zd Bs[TK][TN + 1];
  // A tile: row am = tid & 127, K columns ak + 2 i; B tile: K row bk =
  // tid & 15, columns bn + 16 i
  const int am = tid & 127, ak = tid >> 7;
  const int bk = tid & 15, bn = tid >> 4;
  const int64_t arow = min(m0 + am, M - 1);
  zd ra[8], rb[4];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int i = 0; i < 8; ++i) ra[i] = Ab[arow + (int64_t)min(k0 + ak + 2 * i, K - 1) * lda];
    const int kb = min(k0 + bk, K - 1);
#pragma unroll
    for (int i = 0; i < 4; ++i) rb[i] = Bb[kb + (int64_t)min(n0 + bn + 16 * i, N - 1) * ldb];
  };
  v4d pr[4][2], pi[4][2];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      pr[t][j] = (v4d){0, 0, 0, 0};
      pi[t][j] = (v4d){0, 0, 0, 0};
    }
  fetch(0);
  const int fi = lane & 15, fk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += TK) {
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int kk = ak + 2 * i;
      As[kk][am] = (k0 + kk < K && m0 + am < M) ? ra[i] : zmk(0, 0);
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) Bs[bk][bn + 16 * i] = (k0 + bk < K && n0 + bn + 16 * i < N) ? rb[i] : zmk(0, 0);
    __syncthreads();
    if (k0 + TK < K) fetch(k0 + TK);
#pragma unroll
    for (int ks = 0; ks < TK / 4; ++ks) {
      const int kk = ks * 4 + fk;
      zd xb[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) xb[j] = Bs[kk][32 * wn + 16 * j + fi];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const zd va = As[kk][64 * wm + 16 * t + fi];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          pr[t][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb[j].x, pr[t][j], 0, 0, 0);
          pi[t][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.x, xb[j].y, pi[t][j], 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          pr[t][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-va.y, xb[j].y, pr[t][j], 0, 0, 0);
          pi[t][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(va.y, xb[j].x, pi[t][j], 0, 0, 0);
        }
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int nn = n0 + 32 * wn + 16 * j + fi;
    if (nn >= N) continue;
    zd *Cb = C + b * sC + (int64_t)nn * ldc;
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t m = m0 + 64 * wm + 16 * t + fk + 4 * r;
        if (m < M) {
          zd v = zmul(alpha, zmk(pr[t][j][r], pi[t][j][r]));
          if (BETA) v = zfma(beta, Cb[m], v);
          Cb[m] = v;
        }
      }
  }
}

void launch_gemm_nn(int nb, int64_t M, int N, int K, zd alpha, const zd *A, int64_t lda, int64_t sA,
                    const zd *B, int64_t ldb, int64_t sB, zd beta, zd *C, int64_t ldc, int64_t sC,
                    hipStream_t s) {
  if (M <= 0 || N <= 0 || nb <= 0) return;
  if (K <= 0) throw HipError("gemm_nn: K must be positive");
  if (M >= 4096 && N >= 64) {  // tall: 128 x 64 tiles
    const int tilesN = (N + 63) / 64;
    const int64_t tiles = ((M + 127) / 128) * tilesN;
    // the kernel renumbers the linear id of the whole grid in 32 bits
    if (tiles * nb > 0xffffffffLL || nb > 65535) throw HipError("gemm_nn: grid too large");
    dim3 g((unsigned)tiles, nb), t(256);
    if (beta.x == 0.0 && beta.y == 0.0)
      hipLaunchKernelGGL(k_zgemm_nn_tall<false>, g, t, 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                         sC, tilesN);
    else
      hipLaunchKernelGGL(k_zgemm_nn_tall<true>, g, t, 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc,
                         sC, tilesN);
    EMU_LAUNCH_CHECK();
    return;
  }
  const int tilesN = (N + 63) / 64;
  const int64_t tiles = ((M + 63) / 64) * tilesN;
  if (tiles > 0x7fffffff || nb > 65535) throw HipError("gemm_nn: grid too large");
  dim3 g((unsigned)tiles, nb), t(256);
  if (beta.x == 0.0 && beta.y == 0.0)
    hipLaunchKernelGGL(k_zgemm_nn<false>, g, t, 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                       tilesN);
  else
    hipLaunchKernelGGL(k_zgemm_nn<true>, g, t, 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC,
                       tilesN);
  EMU_LAUNCH_CHECK();
}

// ------------------------------------------------------- w -= V[:, :k] h
__global__ __launch_bounds__(256) void k_sub_vh(int64_t n, int k, const zd *__restrict__ V, int64_t ldv,
                                                int64_t sV, const zd *__restrict__ h, int64_t sh,
                                                zd *__restrict__ w, int64_t sw) {
  __shared__ zd hs[1024];
  const int b = blockIdx.y;
  for (int i = threadIdx.x; i < k && i < 1024; i += blockDim.x) hs[i] = h[b * sh + i];
  __syncthreads();
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  const zd *Vb = V + b * sV;
  zd acc = w[b * sw + r];
  for (int c = 0; c < k; ++c) acc = zfms(Vb[r + (int64_t)c * ldv], hs[c], acc);
  w[b * sw + r] = acc;
}

void launch_sub_vh(int nb, int64_t n, int k, const zd *V, int64_t ldv, int64_t sV, const zd *h, int64_t sh,
                   zd *w, int64_t sw, hipStream_t s) {
  if (k > 1024) throw HipError("sub_vh: Krylov dimension above 1024");
  dim3 g((unsigned)((n + 255) / 256), nb);
  hipLaunchKernelGGL(k_sub_vh, g, dim3(256), 0, s, n, k, V, ldv, sV, h, sh, w, sw);
  EMU_LAUNCH_CHECK();
}

// -------------------------------------------------- norm and normalisation
__global__ __launch_bounds__(256) void k_norm_partial(int64_t n, const zd *__restrict__ w, int64_t sw,
                                                      double *__restrict__ part) {
  __shared__ double red[256];
  const int b = blockIdx.y;
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    zd v = w[b * sw + i];
    acc += v.x * v.x + v.y * v.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[b * gridDim.x + blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_norm_scale(int64_t n, int nparts, const double *__restrict__ part,
                                                    const zd *__restrict__ w, int64_t sw, zd *__restrict__ out,
                                                    int64_t so, double *__restrict__ nrm) {
  __shared__ double tot;
  const int b = blockIdx.y;
  if (threadIdx.x == 0) {
    double acc = 0;
    for (int i = 0; i < nparts; ++i) acc += part[b * nparts + i];
    tot = sqrt(acc);
    if (blockIdx.x == 0) nrm[b] = tot;
  }
  __syncthreads();
  const double inv = tot > 0 ? 1.0 / tot : 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    out[b * so + i] = zscale(w[b * sw + i], inv);
}

void launch_norm_scale(int nb, int64_t n, const zd *w, int64_t sw, zd *out, int64_t so, double *nrm,
                       double *scratch, hipStream_t s) {
  const int nparts = 64;
  hipLaunchKernelGGL(k_norm_partial, dim3(nparts, nb), dim3(256), 0, s, n, w, sw, scratch);
  EMU_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_norm_scale, dim3(nparts, nb), dim3(256), 0, s, n, nparts, scratch, w, sw, out, so, nrm);
  EMU_LAUNCH_CHECK();
}

// H[b][0:k] = h1 + h2
__global__ void k_axpy_h(int k, const zd *__restrict__ h1, const zd *__restrict__ h2, int64_t sh,
                         zd *__restrict__ H, int64_t sH) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (i < k) H[b * sH + i] = zadd(h1[b * sh + i], h2[b * sh + i]);
}

void launch_axpy_h(int nb, int k, const zd *h1, const zd *h2, int64_t sh, zd *H, int64_t sH, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_h, dim3((k + 255) / 256, nb), dim3(256), 0, s, k, h1, h2, sh, H, sH);
  EMU_LAUNCH_CHECK();
}

__global__ void k_copy(int64_t n, const zd *__restrict__ src, int64_t ss, zd *__restrict__ dst, int64_t sd) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (i < n) dst[b * sd + i] = src[b * ss + i];
}

void launch_copy(int nb, int64_t n, const zd *src, int64_t ss, zd *dst, int64_t sd, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((n + 255) / 256), nb), dim3(256), 0, s, n, src, ss, dst, sd);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu

namespace emu {

// ---------------------------------------------- block Gram-Schmidt helpers
// W[b][:, q] -= sum_{c<k} V[b][:, c] H[b][c, q],  q < bq  (thread per row,
// H staged through LDS in chunks of 64 rows)
template <int BQ>
__global__ __launch_bounds__(256) void k_sub_vh_blk(int64_t n, int k, const zd *__restrict__ V, int64_t ldv,
                                                    int64_t sV, const zd *__restrict__ H, int64_t ldh, int64_t sH,
                                                    zd *__restrict__ W, int64_t ldw, int64_t sW) {
  __shared__ zd hs[64][BQ];
  const int b = blockIdx.y;
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const zd *Vb = V + b * sV;
  const zd *Hb = H + b * sH;
  zd acc[BQ];
#pragma unroll
  for (int q = 0; q < BQ; ++q) acc[q] = (r < n) ? W[b * sW + q * ldw + r] : zmk(0, 0);
  for (int c0 = 0; c0 < k; c0 += 64) {
    const int cn = min(64, k - c0);
    __syncthreads();
    for (int idx = threadIdx.x; idx < 64 * BQ; idx += blockDim.x) {
      int c = idx / BQ, q = idx - c * BQ;
      hs[c][q] = c < cn ? Hb[(c0 + c) + (int64_t)q * ldh] : zmk(0, 0);
    }
    __syncthreads();
    if (r < n)
      for (int c = 0; c < cn; ++c) {
        const zd v = Vb[r + (int64_t)(c0 + c) * ldv];
#pragma unroll
        for (int q = 0; q < BQ; ++q) acc[q] = zfms(v, hs[c][q], acc[q]);
      }
  }
  if (r < n)
#pragma unroll
    for (int q = 0; q < BQ; ++q) W[b * sW + q * ldw + r] = acc[q];
}

void launch_sub_vh_blk(int nb, int64_t n, int k, int bq, const zd *V, int64_t ldv, int64_t sV, const zd *H,
                       int64_t ldh, int64_t sH, zd *W, int64_t ldw, int64_t sW, hipStream_t s, KTime kt) {
  dim3 g((unsigned)((n + 255) / 256), nb);
  dim3 t(256);
  switch (bq) {
    case 1: hipExtLaunchKernelGGL(k_sub_vh_blk<1>, g, t, 0, s, kt.a, kt.b, 0, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    case 2: hipExtLaunchKernelGGL(k_sub_vh_blk<2>, g, t, 0, s, kt.a, kt.b, 0, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    case 4: hipExtLaunchKernelGGL(k_sub_vh_blk<4>, g, t, 0, s, kt.a, kt.b, 0, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    case 8: hipExtLaunchKernelGGL(k_sub_vh_blk<8>, g, t, 0, s, kt.a, kt.b, 0, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    default: throw HipError("sub_vh_blk: block size must be 1, 2, 4 or 8");
  }
  EMU_LAUNCH_CHECK();
}

// ------------------------------- delayed block Gram-Schmidt (one pass per block)
// Block Arnoldi where the second Gram-Schmidt pass of block j is delayed into
// the projection of block j + 1 (the low-synchronisation "DCGS2" scheme of
// Swirydowicz et al., Numer. Linear Algebra Appl. 28 (2021), here in block
// form): the operator is applied to the provisional block Z (projected once),
// and a single read of the basis Vt = V[:, :j] gives both Vt^H Z (Z's
// reorthogonalisation) and Vt^H W (W = OP Z's first projection), so each
// block costs two passes over the basis instead of four.  The Arnoldi
// relation OP V[:, :j] = V[:, :j+8] H[:j+8, :j] is kept exact by correcting
// row block j of H when Z is reorthogonalised and by removing OP Vt S (known
// from H) from W.

// W[b] (n x NQ, in place) = W[b] L[b] - V[b][:, :k] E[b] (L: NQ x NQ, column-
// major; E: k x NQ, row-major so that the NQ coefficients of basis column c
// are contiguous).  Thread per row.  The coefficients of a column are the
// same for every lane, so they are read with scalar loads straight into
// SGPRs and used as FMA operands: no LDS traffic in the main loop.
template <int NQ>
__global__ __launch_bounds__(256) void k_combine(int64_t n, int k, const zd *__restrict__ V, int64_t ldv,
                                                 int64_t sV, const zd *__restrict__ E, int64_t sE,
                                                 const zd *__restrict__ L, zd *__restrict__ W, int64_t ldw,
                                                 int64_t sW) {
  __shared__ zd ls[NQ][NQ];
  const int b = blockIdx.y;
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t rr = r < n ? r : n - 1;  // clamped row: loads stay in bounds, the store is skipped
  const zd *Vb = V + b * sV + rr;
  const zd *Eb = E + b * sE;
  for (int idx = threadIdx.x; idx < NQ * NQ; idx += blockDim.x) ls[idx % NQ][idx / NQ] = L[(int64_t)b * NQ * NQ + idx];
  __syncthreads();
  zd acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) acc[q] = zmk(0, 0);
  // acc = W L, W read in halves of NQ / 2 columns, each half's loads all in
  // flight before the first use (the scheduler would otherwise wait for each
  // in turn); halves keep the live registers within four waves per SIMD
#pragma unroll
  for (int h = 0; h < NQ; h += NQ / 2) {
    zd x[NQ / 2];
#pragma unroll
    for (int p = 0; p < NQ / 2; ++p) x[p] = W[b * sW + (h + p) * ldw + rr];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int q = 0; q < NQ; ++q)
#pragma unroll
      for (int p = 0; p < NQ / 2; ++p) acc[q] = zfma(x[p], ls[h + p][q], acc[q]);
  }
#pragma unroll 2
  for (int c = 0; c < k; ++c) {
    const zd v = Vb[(int64_t)c * ldv];
    const zd *ec = Eb + (int64_t)c * NQ;
#pragma unroll
    for (int q = 0; q < NQ; ++q) acc[q] = zfms(v, ec[q], acc[q]);
  }
  if (r < n)
#pragma unroll
    for (int q = 0; q < NQ; ++q) W[b * sW + q * ldw + r] = acc[q];
}

void launch_combine(int nb, int64_t n, int k, int nq, const zd *V, int64_t ldv, int64_t sV, const zd *E,
                    int64_t sE, const zd *L, zd *W, int64_t ldw, int64_t sW, hipStream_t s, KTime kt) {
  dim3 g((unsigned)((n + 255) / 256), nb), t(256);
  switch (nq) {
    case 8: hipExtLaunchKernelGGL(k_combine<8>, g, t, 0, s, kt.a, kt.b, 0, n, k, V, ldv, sV, E, sE, L, W, ldw, sW); break;
    case 16: hipExtLaunchKernelGGL(k_combine<16>, g, t, 0, s, kt.a, kt.b, 0, n, k, V, ldv, sV, E, sE, L, W, ldw, sW); break;
    default: throw HipError("combine: 8 or 16 columns");
  }
  EMU_LAUNCH_CHECK();
}

// Coefficients of one delayed step, one workgroup per instance (block size 8).
//   P = [Vt Z W]^H [Z W] ((j + 16) x 16, ld ldp): S = P[:j, :8], T = P[:j, 8:],
//   G = P[j:j+16, :] = [Z W]^H [Z W];  A0 = H[:j, :j] S (j x 8, ld lda), old H
//   row block j of H (rows j..j+7) is nonzero in columns [c0, j)
// Z's reorthogonalisation: C = Z^H Z - S^H S = R^H R, Q = (Z - Vt S) R^-1,
// H[:j, c0:j] += S H[j:j+8, c0:j], H[j:j+8, c0:j] = R H[j:j+8, c0:j].
// W's first projection against [Vt, Q] with OP Vt S = [Vt, Q] H S removed:
//   u = H_old[j:j+8, :j] S, a = A0 + S u, b = R u,
//   cV = (T - a) R^-1, cQ = R^-H (Z^H W - S^H T) R^-1 - b R^-1  (column j of H),
//   D = b R^-1 + cQ, E = [S R^-1, a R^-1 + cV - S R^-1 D],
//   L = [[R^-1, -R^-1 D], [0, R^-1]], so that [Q, Zraw] = [Z W] L - Vt E.
// FIN (no W; the last provisional block of an expansion; P = [Vt Z]^H Z):
// only Z's reorthogonalisation, E = S R^-1 (j x 8), L = R^-1.
template <bool FIN>
__global__ __launch_bounds__(256) void k_dcgs_coef(int j, int c0, const zd *__restrict__ P, int64_t ldp, int64_t sP,
                                                   const zd *__restrict__ A0, int64_t lda,
                                                   int64_t sA, zd *__restrict__ H, int64_t ldh, int64_t sH,
                                                   zd *__restrict__ E, int64_t sE, zd *__restrict__ L,
                                                   int32_t *__restrict__ bad) {
  constexpr int B = 8;
  constexpr int NG = FIN ? B : 2 * B;
  const int b = blockIdx.x, tid = threadIdx.x;
  const zd *Pb = P + b * sP;
  zd *Hb = H + b * sH;
  zd *Eb = E + b * sE;
  __shared__ zd SS[B][B], ST[B][B], U[B][B], R[B][B], Ri[B][B], Bm[B][B], CQ[B][B], Dm[B][B], T1[B][B];
  __shared__ zd Gs[NG][NG];
  for (int idx = tid; idx < NG * NG; idx += blockDim.x) Gs[idx % NG][idx / NG] = Pb[(j + idx % NG) + (idx / NG) * ldp];
  // 1. S^H S, S^H T over the j basis rows: 32-row chunks of P[:, :16] are
  // staged in LDS by all threads (row tid % 32, columns tid / 32 and
  // tid / 32 + 8), so a chunk costs one round of loads; one thread per
  // output then sums the chunk's rows in order (the same order, hence the
  // same rounding, as a single loop over i).  u = Hrow S after the loop.
  __shared__ zd Pc[32][2 * B + 1];
  {
    const int o = tid % (B * B), kind = tid / (B * B), p = o % B, q = o / B;
    const int lr = tid & 31, lc = tid >> 5;
    zd acc = zmk(0, 0);
    for (int i0 = 0; i0 < j; i0 += 32) {
      const int i = i0 + lr;
      const bool ok = i < j;
      const zd v0 = ok ? Pb[i + lc * ldp] : zmk(0, 0);
      const zd v1 = (ok && !FIN) ? Pb[i + (lc + B) * ldp] : zmk(0, 0);
      __syncthreads();  // the previous chunk has been consumed
      Pc[lr][lc] = v0;
      Pc[lr][lc + B] = v1;
      __syncthreads();
      if (kind < 2) {
        const int nr = min(32, j - i0), cq = kind == 0 ? q : B + q;
        for (int rr = 0; rr < nr; ++rr) acc = zfma(zconj(Pc[rr][p]), Pc[rr][cq], acc);
      }
    }
    if (kind == 0) SS[p][q] = acc;
    if (kind == 1) ST[p][q] = FIN ? zmk(0, 0) : acc;
  }
  if (tid >= 2 * B * B && tid < 3 * B * B) {
    const int o = tid % (B * B), p = o % B, q = o / B;
    zd acc = zmk(0, 0);
    {
      if (!FIN) {
#pragma unroll 8
        for (int c = c0; c < j; ++c) acc = zfma(Hb[(j + p) + (int64_t)c * ldh], Pb[c + q * ldp], acc);
      }
      U[p][q] = acc;
    }
  }
  __syncthreads();
  // 2. Cholesky of C = Z^H Z - S^H S and the inverse of R (one thread)
  if (tid == 0) {
    double gmax = 0;
    for (int i = 0; i < B; ++i) gmax = fmax(gmax, Gs[i][i].x);
    for (int c = 0; c < B; ++c) {
      for (int i = 0; i < B; ++i) { R[i][c] = zmk(0, 0); Ri[i][c] = zmk(0, 0); }
    }
    for (int c = 0; c < B; ++c) {
      for (int i = 0; i < c; ++i) {
        zd acc = zsub(Gs[i][c], SS[i][c]);
        for (int k = 0; k < i; ++k) acc = zsub(acc, zmul(zconj(R[k][i]), R[k][c]));
        R[i][c] = zdiv(acc, R[i][i]);
      }
      double d = Gs[c][c].x - SS[c][c].x;
      for (int k = 0; k < c; ++k) d -= R[k][c].x * R[k][c].x + R[k][c].y * R[k][c].y;
      if (!(d > 1e-28 * gmax)) {
        d = 1e-28 * gmax + 1e-300;
        bad[b] += 1;
      }
      R[c][c] = zmk(sqrt(d), 0);
    }
    for (int c = 0; c < B; ++c) {
      Ri[c][c] = zdiv(zmk(1, 0), R[c][c]);
      for (int i = c - 1; i >= 0; --i) {
        zd acc = zmk(0, 0);
        for (int k = i + 1; k <= c; ++k) acc = zadd(acc, zmul(R[i][k], Ri[k][c]));
        Ri[i][c] = zsub(zmk(0, 0), zdiv(acc, R[i][i]));
      }
    }
  }
  __syncthreads();
  // 3. b = R u; T1 = Z^H W - S^H T; CQ = Ri^H T1 Ri - b Ri; D = Ri^H T1 Ri (64 threads, one entry each)
  const int p8 = tid % B, q8 = (tid / B) % B;
  if (!FIN) {
    if (tid < B * B) {
      zd acc = zmk(0, 0);
      for (int k = p8; k < B; ++k) acc = zfma(R[p8][k], U[k][q8], acc);
      Bm[p8][q8] = acc;
      T1[p8][q8] = zsub(Gs[p8][B + q8], ST[p8][q8]);
    }
    __syncthreads();
    if (tid < B * B) {  // SS reused: T1 Ri
      zd acc = zmk(0, 0);
      for (int k = 0; k <= q8; ++k) acc = zfma(T1[p8][k], Ri[k][q8], acc);
      SS[p8][q8] = acc;
    }
    __syncthreads();
    if (tid < B * B) {
      zd acc = zmk(0, 0), bri = zmk(0, 0);
      for (int k = 0; k <= p8; ++k) acc = zfma(zconj(Ri[k][p8]), SS[k][q8], acc);
      for (int k = 0; k <= q8; ++k) bri = zfma(Bm[p8][k], Ri[k][q8], bri);
      CQ[p8][q8] = zsub(acc, bri);
      Dm[p8][q8] = acc;  // D = b R^-1 + cQ = R^-H T1 R^-1
    }
    __syncthreads();
  }
  // 4. per basis row i < j: a, cV (column j of H), E.  32 rows at a time, one
  // thread per (row, column); the row vectors a and S R^-1 go through LDS.
  __shared__ zd sa[32][B], se[32][B];
  {
    const int rr = tid / B, q = tid % B;
    for (int i0 = 0; i0 < j; i0 += 32) {
      const int i = i0 + rr;
      const bool ok = i < j;
      zd e1 = zmk(0, 0), av = zmk(0, 0);
      if (ok) {
        for (int k = 0; k <= q; ++k) e1 = zfma(Pb[i + k * ldp], Ri[k][q], e1);
        Eb[(int64_t)i * NG + q] = e1;
        if (!FIN) {
          av = A0[b * sA + i + (int64_t)q * lda];
          for (int k = 0; k < B; ++k) av = zfma(Pb[i + k * ldp], U[k][q], av);
        }
      }
      se[rr][q] = e1;
      sa[rr][q] = av;
      __syncthreads();
      if (!FIN && ok) {
        zd cv = zmk(0, 0), e2 = zmk(0, 0);
        for (int k = 0; k <= q; ++k) {
          cv = zfma(zsub(Pb[i + (B + k) * ldp], sa[rr][k]), Ri[k][q], cv);
          e2 = zfma(sa[rr][k], Ri[k][q], e2);
        }
        for (int k = 0; k < B; ++k) e2 = zfms(se[rr][k], Dm[k][q], e2);
        Hb[i + (int64_t)(j + q) * ldh] = cv;
        Eb[(int64_t)i * NG + B + q] = zadd(e2, cv);
      }
      __syncthreads();
    }
  }
  const int w = j - c0;
  for (int64_t idx = tid; idx < (int64_t)j * w; idx += blockDim.x) {
    const int i = (int)(idx % j), c = c0 + (int)(idx / j);
    zd acc = Hb[i + (int64_t)c * ldh];
    for (int k = 0; k < B; ++k) acc = zfma(Pb[i + k * ldp], Hb[(j + k) + (int64_t)c * ldh], acc);
    Hb[i + (int64_t)c * ldh] = acc;
  }
  __syncthreads();
  // 5. Hrow <- R Hrow, column j's diagonal block, L
  for (int c = c0 + tid; c < j; c += blockDim.x) {
    zd h[B];
#pragma unroll
    for (int k = 0; k < B; ++k) h[k] = Hb[(j + k) + (int64_t)c * ldh];
#pragma unroll
    for (int p = 0; p < B; ++p) {
      zd acc = zmk(0, 0);
#pragma unroll
      for (int k = p; k < B; ++k) acc = zfma(R[p][k], h[k], acc);
      Hb[(j + p) + (int64_t)c * ldh] = acc;
    }
  }
  zd *Lb = L + (int64_t)b * NG * NG;
  if (FIN) {
    if (tid < B * B) Lb[p8 + q8 * B] = Ri[p8][q8];
  } else {
    if (tid < B * B) {
      Hb[(j + p8) + (int64_t)(j + q8) * ldh] = CQ[p8][q8];
      zd rd = zmk(0, 0);
      for (int k = p8; k < B; ++k) rd = zfma(Ri[p8][k], Dm[k][q8], rd);
      Lb[p8 + q8 * NG] = Ri[p8][q8];
      Lb[p8 + (B + q8) * NG] = zsub(zmk(0, 0), rd);
      Lb[(B + p8) + q8 * NG] = zmk(0, 0);
      Lb[(B + p8) + (B + q8) * NG] = Ri[p8][q8];
    }
  }
}

void launch_dcgs_coef(int nb, bool fin, int j, int c0, const zd *P, int64_t ldp, int64_t sP, const zd *A0,
                      int64_t lda, int64_t sA, zd *H, int64_t ldh, int64_t sH, zd *E, int64_t sE, zd *L, int32_t *bad,
                      hipStream_t s) {
  if (fin)
    hipLaunchKernelGGL(k_dcgs_coef<true>, dim3(nb), dim3(256), 0, s, j, c0, P, ldp, sP, A0, lda, sA, H, ldh, sH, E, sE,
                       L, bad);
  else
    hipLaunchKernelGGL(k_dcgs_coef<false>, dim3(nb), dim3(256), 0, s, j, c0, P, ldp, sP, A0, lda, sA, H, ldh, sH, E,
                       sE, L, bad);
  EMU_LAUNCH_CHECK();
}

// Cholesky of the bq x bq Gram matrix G = R^H R (R upper), its inverse, and
// Racc = R * Rprev (Rprev = identity when null).  One 64-thread workgroup per
// instance: the matrices are loaded into LDS in parallel, one thread runs the
// (tiny, sequential) factorisation there, and the products are written back
// one entry per thread.  A non-positive pivot (rank-deficient block) is
// replaced by a tiny value and counted in bad[b].
__global__ __launch_bounds__(64) void k_cholqr(int bq, const zd *__restrict__ G, zd *__restrict__ Rinv,
                                               const zd *__restrict__ Rprev, zd *__restrict__ Racc,
                                               int32_t *__restrict__ bad) {
  const int b = blockIdx.x, t = threadIdx.x;
  const int i = t % 8, j = t / 8;
  const bool in = i < bq && j < bq;
  const int64_t o = (int64_t)b * bq * bq + i + j * bq;
  __shared__ zd g[8][8], R[8][8], Ri[8][8], rp[8][8];
  g[i][j] = in ? G[o] : zmk(0, 0);
  rp[i][j] = (in && Rprev) ? Rprev[o] : zmk(i == j ? 1.0 : 0.0, 0);
  R[i][j] = zmk(0, 0);
  Ri[i][j] = zmk(0, 0);
  __syncthreads();
  if (t == 0) {
    double gmax = 0;
    for (int k = 0; k < bq; ++k) gmax = fmax(gmax, g[k][k].x);
    for (int c = 0; c < bq; ++c) {
      for (int r = 0; r < c; ++r) {
        zd acc = g[r][c];
        for (int k = 0; k < r; ++k) acc = zsub(acc, zmul(zconj(R[k][r]), R[k][c]));
        R[r][c] = zdiv(acc, R[r][r]);
      }
      double d = g[c][c].x;
      for (int k = 0; k < c; ++k) d -= R[k][c].x * R[k][c].x + R[k][c].y * R[k][c].y;
      if (!(d > 1e-28 * gmax)) {
        d = 1e-28 * gmax + 1e-300;
        bad[b] += 1;
      }
      R[c][c] = zmk(sqrt(d), 0);
    }
    for (int c = 0; c < bq; ++c) {
      Ri[c][c] = zdiv(zmk(1, 0), R[c][c]);
      for (int r = c - 1; r >= 0; --r) {
        zd acc = zmk(0, 0);
        for (int k = r + 1; k <= c; ++k) acc = zadd(acc, zmul(R[r][k], Ri[k][c]));
        Ri[r][c] = zsub(zmk(0, 0), zdiv(acc, R[r][r]));
      }
    }
  }
  __syncthreads();
  if (!in) return;
  Rinv[o] = Ri[i][j];
  zd acc = zmk(0, 0);
  for (int k = i; k < bq; ++k) acc = zfma(R[i][k], rp[k][j], acc);
  Racc[o] = acc;
}

void launch_cholqr(int nb, int bq, const zd *G, zd *Rinv, const zd *Rprev, zd *Racc, int32_t *bad,
                   hipStream_t s) {
  if (bq > 8) throw HipError("cholqr: block size above 8");
  hipLaunchKernelGGL(k_cholqr, dim3(nb), dim3(64), 0, s, bq, G, Rinv, Rprev, Racc, bad);
  EMU_LAUNCH_CHECK();
}

// Wout[b][:, 0:bq] = W[b][:, 0:bq] * Rinv[b]  (bq x bq upper triangular);
// with xi, the row-interleaved copy of Wout as well (xi[b][8 r + j], the
// operand of the next Krylov block's SpMM, saving k_interleave8's pass)
template <int BQ>
__global__ __launch_bounds__(256) void k_trmm_right(int64_t n, const zd *W, int64_t ldw, int64_t sW,
                                                    const zd *__restrict__ Ri, zd *O, int64_t ldo,
                                                    int64_t sO, zd *__restrict__ xi) {
  __shared__ zd rs[BQ][BQ];
  const int b = blockIdx.y;
  for (int idx = threadIdx.x; idx < BQ * BQ; idx += blockDim.x) rs[idx % BQ][idx / BQ] = Ri[(int64_t)b * BQ * BQ + idx];
  __syncthreads();
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  zd w[BQ];
#pragma unroll
  for (int q = 0; q < BQ; ++q) w[q] = W[b * sW + q * ldw + r];
  zd o[BQ];
#pragma unroll
  for (int j = 0; j < BQ; ++j) {
    zd acc = zmk(0, 0);
#pragma unroll
    for (int k = 0; k <= j; ++k) acc = zfma(w[k], rs[k][j], acc);
    o[j] = acc;
    O[b * sO + j * ldo + r] = acc;
  }
  if (xi) {
    zd *xr = xi + (int64_t)b * BQ * n + (int64_t)BQ * r;
#pragma unroll
    for (int j = 0; j < BQ; ++j) xr[j] = o[j];
  }
}

void launch_trmm_right(int nb, int64_t n, int bq, const zd *W, int64_t ldw, int64_t sW, const zd *Ri, zd *O,
                       int64_t ldo, int64_t sO, hipStream_t s, zd *xi) {
  dim3 g((unsigned)((n + 255) / 256), nb);
  if (xi && bq != 8) throw HipError("trmm_right: the interleaved copy is for 8-column blocks");
  switch (bq) {
    case 1: hipLaunchKernelGGL(k_trmm_right<1>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO, nullptr); break;
    case 2: hipLaunchKernelGGL(k_trmm_right<2>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO, nullptr); break;
    case 4: hipLaunchKernelGGL(k_trmm_right<4>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO, nullptr); break;
    case 8: hipLaunchKernelGGL(k_trmm_right<8>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO, xi); break;
    default: throw HipError("trmm_right: block size must be 1, 2, 4 or 8");
  }
  EMU_LAUNCH_CHECK();
}

// Dst[b][i + j*ldd] (+)= Src[b][i + j*lds] for i < rows, j < cols
__global__ void k_block_copy(int rows, int cols, const zd *__restrict__ S, int64_t lds, int64_t sS,
                             zd *__restrict__ D, int64_t ldd, int64_t sD, int add) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= rows * cols) return;
  const int i = idx % rows, j = idx / rows;
  zd v = S[b * sS + i + (int64_t)j * lds];
  zd *d = &D[b * sD + i + (int64_t)j * ldd];
  *d = add ? zadd(*d, v) : v;
}

void launch_block_copy(int nb, int rows, int cols, const zd *S, int64_t lds, int64_t sS, zd *D, int64_t ldd,
                       int64_t sD, bool add, hipStream_t s) {
  if (rows <= 0 || cols <= 0) return;
  hipLaunchKernelGGL(k_block_copy, dim3((rows * cols + 255) / 256, nb), dim3(256), 0, s, rows, cols, S, lds, sS,
                     D, ldd, sD, add ? 1 : 0);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu
