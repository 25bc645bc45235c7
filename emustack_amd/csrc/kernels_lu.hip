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
// The solve uses the same tree: forward elimination bottom-up with
// update vectors passed to the parent (deterministic, atomic-free), back
// substitution top-down.
#include "engine.h"

namespace emu {

constexpr int NB = 32;     // panel width
constexpr int TILE = 64;   // gemm tile
constexpr int EA_COLS = 8; // columns per extend-add task

struct FrontArrays {
  const int32_t *s, *u, *f, *b, *parent;
  const int64_t *off, *uoff, *pivoff;
  const int32_t *uidx, *pmap;
};

static FrontArrays fa(DevMesh &dm) {
  FrontArrays a;
  a.s = dm.fs.p; a.u = dm.fu.p; a.f = dm.ff.p; a.b = dm.fb.p; a.parent = dm.fparent.p;
  a.off = dm.foff.p; a.uoff = dm.fuoff.p; a.pivoff = dm.fpivoff.p;
  a.uidx = dm.uidx.p; a.pmap = dm.pmap.p;
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
                                               int32_t *__restrict__ pflag, double pert) {
  const int t = tasks[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  zd *F = pool + b * psz + A.off[t];
  int32_t *pv = piv + b * npiv + A.pivoff[t];
  int32_t *pf = pflag + b * npiv + A.pivoff[t];
  const int w = min(NB, s - k0);
  __shared__ double redv[256];
  __shared__ int redi[256];
  __shared__ zd urow[NB];
  for (int jj = 0; jj < w; ++jj) {
    const int k = k0 + jj;
    double best = -1.0;
    int bi = k;
    for (int r = k + tid; r < s; r += 256) {
      double a = zabs1(F[r + (int64_t)k * f]);
      if (a > best) { best = a; bi = r; }
    }
    redv[tid] = best;
    redi[tid] = bi;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (tid < st) {
        double a = redv[tid + st];
        int ai = redi[tid + st];
        if (a > redv[tid] || (a == redv[tid] && ai < redi[tid])) {
          redv[tid] = a;
          redi[tid] = ai;
        }
      }
      __syncthreads();
    }
    const int p = redi[0];
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
      if (zero) F[k + (int64_t)k * f] = d;
    }
    const zd dinv = zdiv(zmk(1.0, 0.0), d);
    const int nrest = w - jj - 1;
    for (int c = tid; c < nrest; c += 256) urow[c] = F[k + (int64_t)(k + 1 + c) * f];
    __syncthreads();
    for (int r = k + 1 + tid; r < f; r += 256) {
      zd l = zmul(F[r + (int64_t)k * f], dinv);
      F[r + (int64_t)k * f] = l;
      for (int c = 0; c < nrest; ++c) {
        zd *pr = &F[r + (int64_t)(k + 1 + c) * f];
        *pr = zfms(l, urow[c], *pr);
      }
    }
    __syncthreads();
  }
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
  for (int i = 0; i < NB; ++i) x[i] = i < w ? col[i] : zmk(0, 0);
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

__global__ __launch_bounds__(256) void k_gemm_update(const int32_t *__restrict__ tasks, int k0,
                                                     FrontArrays A, zd *__restrict__ pool,
                                                     int64_t psz) {
  const int t = tasks[3 * blockIdx.x], r0 = tasks[3 * blockIdx.x + 1], c0 = tasks[3 * blockIdx.x + 2];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], f = A.f[t];
  zd *F = pool + b * psz + A.off[t];
  const int w = min(NB, s - k0);
  __shared__ zd As[NB][TILE];
  __shared__ zd Bs[NB][TILE];
  for (int idx = tid; idx < NB * TILE; idx += 256) {
    int i = idx % TILE, kk = idx / TILE;
    int r = r0 + i;
    As[kk][i] = (kk < w && r < f) ? F[r + (int64_t)(k0 + kk) * f] : zmk(0, 0);
  }
  for (int idx = tid; idx < NB * TILE; idx += 256) {
    int kk = idx % NB, j = idx / NB;
    int c = c0 + j;
    Bs[kk][j] = (kk < w && c < f) ? F[(k0 + kk) + (int64_t)c * f] : zmk(0, 0);
  }
  __syncthreads();
  const int tx = tid % 16, ty = tid / 16;
  zd acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = zmk(0, 0);
  for (int kk = 0; kk < w; ++kk) {
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
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int c = c0 + tx + 16 * j;
    if (c >= f) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int r = r0 + ty + 16 * i;
      if (r < f) {
        zd *p = &F[r + (int64_t)c * f];
        *p = zsub(*p, acc[i][j]);
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

void lu_factor(DevMesh &dm, int nb, const zd *mat1, zd *pool, int32_t *piv, int32_t *pflag,
               hipStream_t s) {
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
                           dm.tasks.p + st.panel_off, st.k0, A, pool, psz, piv, sy.npiv, pflag, pert);
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
}

// ---------------------------------------------------------------- solves
//
// One workgroup per (front, instance, chunk of RC right-hand sides).  The
// front's RHS block r[f][RC] lives in LDS.  Triangular parts are processed in
// SB-column blocks: the SB x SB diagonal block is solved by one wavefront per
// RHS column (register shuffles, no barrier), then all 256 lanes apply the
// block to the remaining rows (coalesced column reads of the factor).  Two
// barriers per SB columns instead of one per column.

constexpr int SB = 32;

__device__ __forceinline__ zd shfl_z(zd v, int lane) {
  return zmk(__shfl(v.x, lane, 64), __shfl(v.y, lane, 64));
}

template <int RC>
__global__ __launch_bounds__(256) void k_fwd(const int32_t *__restrict__ fronts, FrontArrays A,
                                             const int32_t *__restrict__ child0,
                                             const int32_t *__restrict__ child1, const zd *__restrict__ pool,
                                             int64_t psz, const int32_t *__restrict__ piv, int64_t npiv,
                                             int nrhs, int64_t n, int64_t nw, const zd *__restrict__ bv,
                                             zd *__restrict__ y, zd *__restrict__ wbuf) {
  extern __shared__ zd r[];
  const int t = fronts[blockIdx.x];
  const int b = blockIdx.y;
  const int q0 = blockIdx.z * RC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int s = A.s[t], u = A.u[t], f = A.f[t], fb = A.b[t];
  const zd *F = pool + b * psz + A.off[t];
  const int32_t *pv = piv + b * npiv + A.pivoff[t];
  const int nq = min(RC, nrhs - q0);
  const int64_t col0 = (int64_t)b * nrhs + q0;
  for (int idx = tid; idx < f * RC; idx += 256) {
    int i = idx / RC, q = idx - i * RC;
    r[idx] = (i < s && q < nq) ? bv[(col0 + q) * n + fb + i] : zmk(0, 0);
  }
  __syncthreads();
  for (int c = 0; c < 2; ++c) {
    const int cf = c == 0 ? child0[t] : child1[t];
    if (cf < 0) continue;
    const int uc = A.u[cf];
    const int32_t *mp = A.pmap + A.uoff[cf];
    for (int idx = tid; idx < uc * RC; idx += 256) {
      int k = idx / RC, q = idx - k * RC;
      if (q < nq) {
        zd *d = &r[mp[k] * RC + q];
        *d = zadd(*d, wbuf[(col0 + q) * nw + A.uoff[cf] + k]);
      }
    }
    __syncthreads();
  }
  if (tid < nq)
    for (int k = 0; k < s; ++k) {
      int p = pv[k];
      if (p != k) {
        zd a = r[k * RC + tid];
        r[k * RC + tid] = r[p * RC + tid];
        r[p * RC + tid] = a;
      }
    }
  __syncthreads();
  __shared__ zd D[SB][SB + 1];
  for (int kb = 0; kb < s; kb += SB) {
    const int w = min(SB, s - kb);
    for (int idx = tid; idx < w * w; idx += 256) {
      int i = idx % w, j = idx / w;
      D[i][j] = F[(kb + i) + (int64_t)(kb + j) * f];
    }
    __syncthreads();
    for (int q = wave; q < nq; q += 4) {
      zd v = lane < w ? r[(kb + lane) * RC + q] : zmk(0, 0);
      for (int k = 0; k < w - 1; ++k) {
        zd yk = shfl_z(v, k);
        if (lane > k && lane < w) v = zfms(D[lane][k], yk, v);
      }
      if (lane < w) r[(kb + lane) * RC + q] = v;
    }
    __syncthreads();
    for (int i = kb + w + tid; i < f; i += 256) {
      zd acc[RC];
#pragma unroll
      for (int q = 0; q < RC; ++q) acc[q] = r[i * RC + q];
      for (int kk = 0; kk < w; ++kk) {
        const zd l = F[i + (int64_t)(kb + kk) * f];
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zfms(l, r[(kb + kk) * RC + q], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < RC; ++q) r[i * RC + q] = acc[q];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < f * RC; idx += 256) {
    int i = idx / RC, q = idx - i * RC;
    if (q >= nq) continue;
    if (i < s) y[(col0 + q) * n + fb + i] = r[idx];
    else wbuf[(col0 + q) * nw + A.uoff[t] + (i - s)] = r[idx];
  }
}

template <int RC>
__global__ __launch_bounds__(256) void k_bwd(const int32_t *__restrict__ fronts, FrontArrays A,
                                             const zd *__restrict__ pool, int64_t psz, int nrhs, int64_t n,
                                             const zd *__restrict__ y, zd *__restrict__ x) {
  extern __shared__ zd r[];
  const int t = fronts[blockIdx.x];
  const int b = blockIdx.y;
  const int q0 = blockIdx.z * RC;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int s = A.s[t], u = A.u[t], f = A.f[t], fb = A.b[t];
  const zd *F = pool + b * psz + A.off[t];
  const int32_t *ui = A.uidx + A.uoff[t];
  const int nq = min(RC, nrhs - q0);
  const int64_t col0 = (int64_t)b * nrhs + q0;
  for (int idx = tid; idx < f * RC; idx += 256) {
    int i = idx / RC, q = idx - i * RC;
    zd v = zmk(0, 0);
    if (q < nq) v = i < s ? y[(col0 + q) * n + fb + i] : x[(col0 + q) * n + ui[i - s]];
    r[idx] = v;
  }
  __syncthreads();
  for (int i = tid; i < s; i += 256) {
    zd acc[RC];
#pragma unroll
    for (int q = 0; q < RC; ++q) acc[q] = r[i * RC + q];
    for (int c = 0; c < u; ++c) {
      const zd uv = F[i + (int64_t)(s + c) * f];
#pragma unroll
      for (int q = 0; q < RC; ++q) acc[q] = zfms(uv, r[(s + c) * RC + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < RC; ++q) r[i * RC + q] = acc[q];
  }
  __syncthreads();
  __shared__ zd D[SB][SB + 1];
  for (int bi = (s + SB - 1) / SB - 1; bi >= 0; --bi) {
    const int kb = bi * SB, w = min(SB, s - kb);
    for (int idx = tid; idx < w * w; idx += 256) {
      int i = idx % w, j = idx / w;
      D[i][j] = F[(kb + i) + (int64_t)(kb + j) * f];
    }
    __syncthreads();
    for (int q = wave; q < nq; q += 4) {
      zd v = lane < w ? r[(kb + lane) * RC + q] : zmk(0, 0);
      for (int k = w - 1; k >= 0; --k) {
        const zd xk = zdiv(shfl_z(v, k), D[k][k]);
        if (lane == k) v = xk;
        else if (lane < k) v = zfms(D[lane][k], xk, v);
      }
      if (lane < w) r[(kb + lane) * RC + q] = v;
    }
    __syncthreads();
    for (int i = tid; i < kb; i += 256) {
      zd acc[RC];
#pragma unroll
      for (int q = 0; q < RC; ++q) acc[q] = r[i * RC + q];
      for (int kk = 0; kk < w; ++kk) {
        const zd uv = F[i + (int64_t)(kb + kk) * f];
#pragma unroll
        for (int q = 0; q < RC; ++q) acc[q] = zfms(uv, r[(kb + kk) * RC + q], acc[q]);
      }
#pragma unroll
      for (int q = 0; q < RC; ++q) r[i * RC + q] = acc[q];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < s * RC; idx += 256) {
    int i = idx / RC, q = idx - i * RC;
    if (q < nq) x[(col0 + q) * n + fb + i] = r[idx];
  }
}

static const size_t kMaxLds = 160 * 1024 - 16 * 1024;  // dynamic part (static D block + margin)

template <int RC>
static void set_attr_once() {
  static bool attr = false;
  if (!attr) {
    EMU_HIP(hipFuncSetAttribute((const void *)k_fwd<RC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds));
    EMU_HIP(hipFuncSetAttribute((const void *)k_bwd<RC>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kMaxLds));
    attr = true;
  }
}

// largest RHS chunk (8, 4, 2, 1) not above nrhs whose LDS block fits
static int pick_rc(int nrhs, int max_f) {
  const size_t per = (size_t)max_f * sizeof(zd);
  for (int rc : {8, 4, 2}) if (nrhs >= rc && rc * per <= kMaxLds) return rc;
  return 1;
}

template <int RC>
static void launch_fwd(const LevelSched &L, DevMesh &dm, FrontArrays A, const int32_t *c0, const int32_t *c1, int nb,
                       int nrhs, const zd *pool, const int32_t *piv, const zd *bnew, zd *y, zd *w, hipStream_t s) {
  set_attr_once<RC>();
  const Symbolic &sy = dm.host.sym;
  const unsigned nz = (unsigned)((nrhs + RC - 1) / RC);
  size_t lds = (size_t)L.max_f * RC * sizeof(zd);
  hipLaunchKernelGGL(k_fwd<RC>, dim3((unsigned)L.fronts_cnt, nb, nz), dim3(256), lds, s, dm.tasks.p + L.fronts_off,
                     A, c0, c1, pool, sy.pool, piv, sy.npiv, nrhs, sy.n, sy.nw, bnew, y, w);
}

template <int RC>
static void launch_bwd(const LevelSched &L, DevMesh &dm, FrontArrays A, int nb, int nrhs, const zd *pool,
                       const zd *y, zd *x, hipStream_t s) {
  set_attr_once<RC>();
  const Symbolic &sy = dm.host.sym;
  const unsigned nz = (unsigned)((nrhs + RC - 1) / RC);
  size_t lds = (size_t)L.max_f * RC * sizeof(zd);
  hipLaunchKernelGGL(k_bwd<RC>, dim3((unsigned)L.fronts_cnt, nb, nz), dim3(256), lds, s, dm.tasks.p + L.fronts_off,
                     A, pool, sy.pool, nrhs, sy.n, y, x);
}

void lu_solve(DevMesh &dm, int nb, int nrhs, const zd *pool, const int32_t *piv, const zd *bnew,
              zd *y, zd *x, zd *w, hipStream_t s, Profiler *prof) {
  if ((size_t)dm.max_f * sizeof(zd) > kMaxLds) throw HipError("front too large for the LDS-resident solve");
  FrontArrays A = fa(dm);
  const int32_t *c0 = dm.tasks.p + dm.child_off[0];
  const int32_t *c1 = dm.tasks.p + dm.child_off[1];
  const double scale = (double)nb * nrhs;
  for (const LevelSched &L : dm.sched) {
    int ev = prof ? prof->begin(s) : -1;
    switch (pick_rc(nrhs, L.max_f)) {
      case 8: launch_fwd<8>(L, dm, A, c0, c1, nb, nrhs, pool, piv, bnew, y, w, s); break;
      case 4: launch_fwd<4>(L, dm, A, c0, c1, nb, nrhs, pool, piv, bnew, y, w, s); break;
      case 2: launch_fwd<2>(L, dm, A, c0, c1, nb, nrhs, pool, piv, bnew, y, w, s); break;
      default: launch_fwd<1>(L, dm, A, c0, c1, nb, nrhs, pool, piv, bnew, y, w, s); break;
    }
    if (prof) prof->end(s, ev, 0, L.fwd_bytes * scale);
  }
  EMU_LAUNCH_CHECK();
  for (auto it = dm.sched.rbegin(); it != dm.sched.rend(); ++it) {
    int ev = prof ? prof->begin(s) : -1;
    switch (pick_rc(nrhs, it->max_f)) {
      case 8: launch_bwd<8>(*it, dm, A, nb, nrhs, pool, y, x, s); break;
      case 4: launch_bwd<4>(*it, dm, A, nb, nrhs, pool, y, x, s); break;
      case 2: launch_bwd<2>(*it, dm, A, nb, nrhs, pool, y, x, s); break;
      default: launch_bwd<1>(*it, dm, A, nb, nrhs, pool, y, x, s); break;
    }
    if (prof) prof->end(s, ev, 1, it->bwd_bytes * scale);
  }
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
