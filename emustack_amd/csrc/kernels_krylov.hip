// Dense / vector kernels of the Krylov-Schur eigensolver and of the overlap
// contractions, batched over instances (blockIdx.y / blockIdx.z = instance).
#include "engine.h"

namespace emu {

// --------------------------------------------------------------- SpMM (CSR)
// y[b][c][r] = sum_k A[b][r][k] x[b][c][k] for a block of SM_C columns.
// Eight lanes share a row: each lane walks every eighth nonzero (coalesced
// value / index reads, one read of the row per column block), accumulates
// the SM_C column sums, and the eight partial sums are combined with
// cross-lane shuffles; lane j then writes column j.  A wavefront covers 8
// rows, a workgroup 32.
constexpr int SM_C = 8;

__device__ __forceinline__ zd shfl_xor_z(zd v, int m) {
  return zmk(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

__global__ __launch_bounds__(256) void k_spmm(int64_t n, int ncol, const int64_t *__restrict__ rp,
                                              const int32_t *__restrict__ ci, const zd *__restrict__ vals,
                                              int64_t vstride, const zd *__restrict__ x, int64_t xstride,
                                              zd *__restrict__ y, int64_t ystride) {
  const int lane8 = threadIdx.x & 7;
  const int64_t r = (int64_t)blockIdx.x * 32 + (threadIdx.x >> 3);
  const int b = blockIdx.y;
  const int c0 = blockIdx.z * SM_C;
  const int nc = min(SM_C, ncol - c0);
  const zd *v = vals + b * vstride;
  const zd *xb = x + b * xstride + (int64_t)c0 * n;
  zd acc[SM_C];
#pragma unroll
  for (int c = 0; c < SM_C; ++c) acc[c] = zmk(0, 0);
  if (r < n) {
    const int64_t k0 = rp[r], k1 = rp[r + 1];
    for (int64_t k = k0 + lane8; k < k1; k += 8) {
      const zd a = v[k];
      const int64_t col = ci[k];
#pragma unroll
      for (int c = 0; c < SM_C; ++c)
        if (c < nc) acc[c] = zfma(a, xb[(int64_t)c * n + col], acc[c]);
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

void launch_spmm(int nb, int64_t n, int ncol, const int64_t *rp, const int32_t *ci, const zd *vals,
                 int64_t vstride, const zd *x, int64_t xstride, zd *y, int64_t ystride, hipStream_t s) {
  dim3 g((unsigned)((n + 31) / 32), nb, (unsigned)((ncol + SM_C - 1) / SM_C));
  hipLaunchKernelGGL(k_spmm, g, dim3(256), 0, s, n, ncol, rp, ci, vals, vstride, x, xstride, y, ystride);
  EMU_LAUNCH_CHECK();
}

// ----------------------------------------------- C = op(A)^T B, split over K
constexpr int AT = 32;     // output tile
constexpr int AKC = 32;    // K sub-chunk staged in LDS

__global__ __launch_bounds__(256) void k_atb_partial(int64_t K, int M, int N, int64_t kchunk,
                                                     const zd *__restrict__ A, int64_t lda, int64_t sA,
                                                     const zd *__restrict__ B, int64_t ldb, int64_t sB,
                                                     int conjA, zd *__restrict__ part, int tilesM) {
  const int tile = blockIdx.x;
  const int tm = tile % tilesM, tn = tile / tilesM;
  const int chunk = blockIdx.y;
  const int b = blockIdx.z;
  const int nchunk = gridDim.y;
  const int tid = threadIdx.x;
  const int m0 = tm * AT, n0 = tn * AT;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  __shared__ zd As[AKC][AT + 1];
  __shared__ zd Bs[AKC][AT + 1];
  const int tx = tid % 16, ty = tid / 16;  // outputs (ty + 16 i, tx + 16 j), i, j in {0,1}
  zd acc[2][2] = {{zmk(0, 0), zmk(0, 0)}, {zmk(0, 0), zmk(0, 0)}};
  const int64_t kb = chunk * kchunk, ke = min(K, kb + kchunk);
  for (int64_t k0 = kb; k0 < ke; k0 += AKC) {
    for (int idx = tid; idx < AKC * AT; idx += 256) {
      int kk = idx % AKC, mm = idx / AKC;
      int64_t k = k0 + kk;
      zd a = (k < ke && m0 + mm < M) ? Ab[k + (int64_t)(m0 + mm) * lda] : zmk(0, 0);
      As[kk][mm] = conjA ? zconj(a) : a;
      zd bb = (k < ke && n0 + mm < N) ? Bb[k + (int64_t)(n0 + mm) * ldb] : zmk(0, 0);
      Bs[kk][mm] = bb;
    }
    __syncthreads();
#pragma unroll 8
    for (int kk = 0; kk < AKC; ++kk) {
      zd a0 = As[kk][ty], a1 = As[kk][ty + 16];
      zd b0 = Bs[kk][tx], b1 = Bs[kk][tx + 16];
      acc[0][0] = zfma(a0, b0, acc[0][0]);
      acc[0][1] = zfma(a0, b1, acc[0][1]);
      acc[1][0] = zfma(a1, b0, acc[1][0]);
      acc[1][1] = zfma(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      int m = m0 + ty + 16 * i, nn = n0 + tx + 16 * j;
      if (m < M && nn < N) part[(((int64_t)b * nchunk + chunk) * N + nn) * M + m] = acc[i][j];
    }
}

__global__ void k_atb_reduce(int M, int N, int nchunk, const zd *__restrict__ part, zd *__restrict__ C,
                             int64_t ldc, int64_t sC) {
  int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (idx >= (int64_t)M * N) return;
  int m = idx % M, nn = idx / M;
  zd acc = zmk(0, 0);
  for (int c = 0; c < nchunk; ++c) acc = zadd(acc, part[(((int64_t)b * nchunk + c) * N + nn) * M + m]);
  C[b * sC + m + (int64_t)nn * ldc] = acc;
}

void launch_atb(int nb, int64_t K, int M, int N, const zd *A, int64_t lda, int64_t sA, const zd *B,
                int64_t ldb, int64_t sB, zd *C, int64_t ldc, int64_t sC, bool conjA, zd *scratch,
                size_t scratch_elems, hipStream_t s) {
  int tilesM = (M + AT - 1) / AT, tilesN = (N + AT - 1) / AT;
  int64_t per = (int64_t)nb * M * N;
  int64_t maxchunk = std::max<int64_t>(1, (int64_t)(scratch_elems / std::max<int64_t>(per, 1)));
  int64_t want = std::max<int64_t>(1, 2048 / std::max(1, tilesM * tilesN * nb));
  int64_t nchunk = std::min<int64_t>(std::min<int64_t>(want, maxchunk), (K + AKC - 1) / AKC);
  nchunk = std::max<int64_t>(1, std::min<int64_t>(nchunk, 65535));
  int64_t kchunk = (K + nchunk - 1) / nchunk;
  kchunk = ((kchunk + AKC - 1) / AKC) * AKC;
  nchunk = (K + kchunk - 1) / kchunk;
  if ((size_t)(nchunk * per) > scratch_elems) throw HipError("atb: scratch too small");
  hipLaunchKernelGGL(k_atb_partial, dim3(tilesM * tilesN, (unsigned)nchunk, nb), dim3(256), 0, s, K, M,
                     N, kchunk, A, lda, sA, B, ldb, sB, conjA ? 1 : 0, scratch, tilesM);
  EMU_LAUNCH_CHECK();
  int64_t tot = (int64_t)M * N;
  hipLaunchKernelGGL(k_atb_reduce, dim3((unsigned)((tot + 255) / 256), nb), dim3(256), 0, s, M, N,
                     (int)nchunk, scratch, C, ldc, sC);
  EMU_LAUNCH_CHECK();
}

// ------------------------------------------------ C = alpha A B + beta C (NN)
constexpr int GT = 64;
constexpr int GK = 16;

__global__ __launch_bounds__(256) void k_gemm_nn(int64_t M, int N, int K, zd alpha, const zd *__restrict__ A,
                                                 int64_t lda, int64_t sA, const zd *__restrict__ B,
                                                 int64_t ldb, int64_t sB, zd beta, zd *__restrict__ C,
                                                 int64_t ldc, int64_t sC) {
  const int64_t m0 = (int64_t)blockIdx.x * GT;
  const int n0 = blockIdx.y * GT;
  const int b = blockIdx.z;
  const int tid = threadIdx.x;
  const zd *Ab = A + b * sA;
  const zd *Bb = B + b * sB;
  zd *Cb = C + b * sC;
  __shared__ zd As[GK][GT];
  __shared__ zd Bs[GK][GT + 1];
  const int tx = tid % 16, ty = tid / 16;
  zd acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = zmk(0, 0);
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int idx = tid; idx < GK * GT; idx += 256) {
      int i = idx % GT, kk = idx / GT;
      int64_t m = m0 + i;
      int k = k0 + kk;
      As[kk][i] = (m < M && k < K) ? Ab[m + (int64_t)k * lda] : zmk(0, 0);
    }
    for (int idx = tid; idx < GK * GT; idx += 256) {
      int kk = idx % GK, j = idx / GK;
      int k = k0 + kk, nn = n0 + j;
      Bs[kk][j] = (k < K && nn < N) ? Bb[k + (int64_t)nn * ldb] : zmk(0, 0);
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; ++kk) {
      zd a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][tx + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = zfma(a[i], bb[j], acc[i][j]);
    }
    __syncthreads();
  }
  const bool bz = (beta.x == 0.0 && beta.y == 0.0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int nn = n0 + tx + 16 * j;
    if (nn >= N) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int64_t m = m0 + ty + 16 * i;
      if (m < M) {
        zd *p = &Cb[m + (int64_t)nn * ldc];
        zd v = zmul(alpha, acc[i][j]);
        *p = bz ? v : zfma(beta, *p, v);
      }
    }
  }
}

void launch_gemm_nn(int nb, int64_t M, int N, int K, zd alpha, const zd *A, int64_t lda, int64_t sA,
                    const zd *B, int64_t ldb, int64_t sB, zd beta, zd *C, int64_t ldc, int64_t sC,
                    hipStream_t s) {
  if (M <= 0 || N <= 0) return;
  dim3 g((unsigned)((M + GT - 1) / GT), (unsigned)((N + GT - 1) / GT), nb);
  hipLaunchKernelGGL(k_gemm_nn, g, dim3(256), 0, s, M, N, K, alpha, A, lda, sA, B, ldb, sB, beta, C, ldc, sC);
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
                       int64_t ldh, int64_t sH, zd *W, int64_t ldw, int64_t sW, hipStream_t s) {
  dim3 g((unsigned)((n + 255) / 256), nb);
  switch (bq) {
    case 1: hipLaunchKernelGGL(k_sub_vh_blk<1>, g, dim3(256), 0, s, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    case 2: hipLaunchKernelGGL(k_sub_vh_blk<2>, g, dim3(256), 0, s, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    case 4: hipLaunchKernelGGL(k_sub_vh_blk<4>, g, dim3(256), 0, s, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    case 8: hipLaunchKernelGGL(k_sub_vh_blk<8>, g, dim3(256), 0, s, n, k, V, ldv, sV, H, ldh, sH, W, ldw, sW); break;
    default: throw HipError("sub_vh_blk: block size must be 1, 2, 4 or 8");
  }
  EMU_LAUNCH_CHECK();
}

// Cholesky of the bq x bq Gram matrix G = R^H R (R upper), its inverse, and
// Racc = R * Rprev (Rprev = identity when null).  One thread per instance.
// A non-positive pivot (rank-deficient block) is replaced by a tiny value
// and counted in bad[b].
__global__ void k_cholqr(int nb, int bq, const zd *__restrict__ G, zd *__restrict__ Rinv,
                         const zd *__restrict__ Rprev, zd *__restrict__ Racc, int32_t *__restrict__ bad) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const zd *g = G + (int64_t)b * bq * bq;
  zd R[8][8], Ri[8][8];
  for (int i = 0; i < bq; ++i)
    for (int j = 0; j < bq; ++j) { R[i][j] = zmk(0, 0); Ri[i][j] = zmk(0, 0); }
  double gmax = 0;
  for (int i = 0; i < bq; ++i) gmax = fmax(gmax, g[i + i * bq].x);
  for (int j = 0; j < bq; ++j) {
    // R[i][j] for i < j, then R[j][j]
    for (int i = 0; i < j; ++i) {
      zd acc = g[i + j * bq];  // G(i, j)
      for (int k = 0; k < i; ++k) acc = zsub(acc, zmul(zconj(R[k][i]), R[k][j]));
      R[i][j] = zdiv(acc, R[i][i]);
    }
    double d = g[j + j * bq].x;
    for (int k = 0; k < j; ++k) d -= R[k][j].x * R[k][j].x + R[k][j].y * R[k][j].y;
    if (!(d > 1e-28 * gmax)) {
      d = 1e-28 * gmax + 1e-300;
      bad[b] += 1;
    }
    R[j][j] = zmk(sqrt(d), 0);
  }
  // inverse of upper triangular R
  for (int j = 0; j < bq; ++j) {
    Ri[j][j] = zdiv(zmk(1, 0), R[j][j]);
    for (int i = j - 1; i >= 0; --i) {
      zd acc = zmk(0, 0);
      for (int k = i + 1; k <= j; ++k) acc = zadd(acc, zmul(R[i][k], Ri[k][j]));
      Ri[i][j] = zsub(zmk(0, 0), zdiv(acc, R[i][i]));
    }
  }
  zd *ri = Rinv + (int64_t)b * bq * bq;
  zd *ra = Racc + (int64_t)b * bq * bq;
  const zd *rp = Rprev ? Rprev + (int64_t)b * bq * bq : nullptr;
  for (int i = 0; i < bq; ++i)
    for (int j = 0; j < bq; ++j) {
      ri[i + j * bq] = Ri[i][j];
      zd acc = zmk(0, 0);
      if (rp) {
        for (int k = 0; k < bq; ++k) acc = zadd(acc, zmul(R[i][k], rp[k + j * bq]));
      } else {
        acc = R[i][j];
      }
      ra[i + j * bq] = acc;
    }
}

void launch_cholqr(int nb, int bq, const zd *G, zd *Rinv, const zd *Rprev, zd *Racc, int32_t *bad,
                   hipStream_t s) {
  if (bq > 8) throw HipError("cholqr: block size above 8");
  hipLaunchKernelGGL(k_cholqr, dim3((nb + 63) / 64), dim3(64), 0, s, nb, bq, G, Rinv, Rprev, Racc, bad);
  EMU_LAUNCH_CHECK();
}

// Wout[b][:, 0:bq] = W[b][:, 0:bq] * Rinv[b]  (bq x bq upper triangular)
template <int BQ>
__global__ __launch_bounds__(256) void k_trmm_right(int64_t n, const zd *W, int64_t ldw, int64_t sW,
                                                    const zd *__restrict__ Ri, zd *O, int64_t ldo,
                                                    int64_t sO) {
  __shared__ zd rs[BQ][BQ];
  const int b = blockIdx.y;
  for (int idx = threadIdx.x; idx < BQ * BQ; idx += blockDim.x) rs[idx % BQ][idx / BQ] = Ri[(int64_t)b * BQ * BQ + idx];
  __syncthreads();
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (r >= n) return;
  zd w[BQ];
#pragma unroll
  for (int q = 0; q < BQ; ++q) w[q] = W[b * sW + q * ldw + r];
#pragma unroll
  for (int j = 0; j < BQ; ++j) {
    zd acc = zmk(0, 0);
#pragma unroll
    for (int k = 0; k <= j; ++k) acc = zfma(w[k], rs[k][j], acc);
    O[b * sO + j * ldo + r] = acc;
  }
}

void launch_trmm_right(int nb, int64_t n, int bq, const zd *W, int64_t ldw, int64_t sW, const zd *Ri, zd *O,
                       int64_t ldo, int64_t sO, hipStream_t s) {
  dim3 g((unsigned)((n + 255) / 256), nb);
  switch (bq) {
    case 1: hipLaunchKernelGGL(k_trmm_right<1>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO); break;
    case 2: hipLaunchKernelGGL(k_trmm_right<2>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO); break;
    case 4: hipLaunchKernelGGL(k_trmm_right<4>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO); break;
    case 8: hipLaunchKernelGGL(k_trmm_right<8>, g, dim3(256), 0, s, n, W, ldw, sW, Ri, O, ldo, sO); break;
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
