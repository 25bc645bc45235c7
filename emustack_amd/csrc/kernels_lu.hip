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

__global__ __launch_bounds__(256) void k_fwd(const int32_t *__restrict__ fronts, FrontArrays A,
                                             const int32_t *__restrict__ child0,
                                             const int32_t *__restrict__ child1, const zd *__restrict__ pool,
                                             int64_t psz, const int32_t *__restrict__ piv, int64_t npiv,
                                             int nrhs, int64_t n, int64_t nw, const zd *__restrict__ bv,
                                             zd *__restrict__ y, zd *__restrict__ wbuf) {
  extern __shared__ zd r[];
  const int t = fronts[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], u = A.u[t], f = A.f[t], fb = A.b[t];
  const zd *F = pool + b * psz + A.off[t];
  const int32_t *pv = piv + b * npiv + A.pivoff[t];
  const int ch[2] = {child0[t], child1[t]};
  for (int q = 0; q < nrhs; ++q) {
    const zd *bb = bv + ((int64_t)b * nrhs + q) * n;
    zd *yy = y + ((int64_t)b * nrhs + q) * n;
    zd *ww = wbuf + ((int64_t)b * nrhs + q) * nw;
    for (int i = tid; i < f; i += 256) r[i] = i < s ? bb[fb + i] : zmk(0, 0);
    __syncthreads();
    for (int c = 0; c < 2; ++c) {
      int cf = ch[c];
      if (cf < 0) continue;
      const int uc = A.u[cf];
      const int32_t *mp = A.pmap + A.uoff[cf];
      const zd *wc = ww + A.uoff[cf];
      for (int k = tid; k < uc; k += 256) r[mp[k]] = zadd(r[mp[k]], wc[k]);
      __syncthreads();
    }
    if (tid == 0)
      for (int k = 0; k < s; ++k) {
        int p = pv[k];
        if (p != k) { zd a = r[k]; r[k] = r[p]; r[p] = a; }
      }
    __syncthreads();
    for (int k = 0; k < s; ++k) {
      const zd yk = r[k];
      const zd *col = F + (int64_t)k * f;
      for (int i = k + 1 + tid; i < f; i += 256) r[i] = zfms(col[i], yk, r[i]);
      __syncthreads();
    }
    for (int i = tid; i < s; i += 256) yy[fb + i] = r[i];
    for (int i = tid; i < u; i += 256) ww[A.uoff[t] + i] = r[s + i];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_bwd(const int32_t *__restrict__ fronts, FrontArrays A,
                                             const zd *__restrict__ pool, int64_t psz, int nrhs, int64_t n,
                                             const zd *__restrict__ y, zd *__restrict__ x) {
  extern __shared__ zd r[];
  const int t = fronts[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int s = A.s[t], u = A.u[t], f = A.f[t], fb = A.b[t];
  const zd *F = pool + b * psz + A.off[t];
  const int32_t *ui = A.uidx + A.uoff[t];
  for (int q = 0; q < nrhs; ++q) {
    const zd *yy = y + ((int64_t)b * nrhs + q) * n;
    zd *xx = x + ((int64_t)b * nrhs + q) * n;
    for (int i = tid; i < f; i += 256) r[i] = i < s ? yy[fb + i] : xx[ui[i - s]];
    __syncthreads();
    for (int i = tid; i < s; i += 256) {
      zd acc = r[i];
      for (int c = 0; c < u; ++c) acc = zfms(F[i + (int64_t)(s + c) * f], r[s + c], acc);
      r[i] = acc;
    }
    __syncthreads();
    for (int k = s - 1; k >= 0; --k) {
      const zd xk = zdiv(r[k], F[k + (int64_t)k * f]);
      const zd *col = F + (int64_t)k * f;
      for (int i = tid; i < k; i += 256) r[i] = zfms(col[i], xk, r[i]);
      __syncthreads();
      if (tid == 0) r[k] = xk;
    }
    __syncthreads();
    for (int i = tid; i < s; i += 256) xx[fb + i] = r[i];
    __syncthreads();
  }
}

void lu_solve(DevMesh &dm, int nb, int nrhs, const zd *pool, const int32_t *piv, const zd *bnew,
              zd *y, zd *x, zd *w, hipStream_t s, Profiler *prof) {
  const Symbolic &sy = dm.host.sym;
  FrontArrays A = fa(dm);
  size_t lds = (size_t)dm.max_f * sizeof(zd);
  const int32_t *c0 = dm.tasks.p + dm.child_off[0];
  const int32_t *c1 = dm.tasks.p + dm.child_off[1];
  if (prof) prof->start_sweep();
  const double scale = (double)nb * nrhs;
  for (const LevelSched &L : dm.sched) {
    int ev = prof ? prof->begin(s) : -1;
    hipLaunchKernelGGL(k_fwd, dim3((unsigned)L.fronts_cnt, nb), dim3(256), lds, s,
                       dm.tasks.p + L.fronts_off, A, c0, c1, pool, sy.pool, piv, sy.npiv, nrhs,
                       sy.n, sy.nw, bnew, y, w);
    if (prof) prof->end(s, ev, 0, L.fwd_bytes * scale);
  }
  EMU_LAUNCH_CHECK();
  for (auto it = dm.sched.rbegin(); it != dm.sched.rend(); ++it) {
    int ev = prof ? prof->begin(s) : -1;
    hipLaunchKernelGGL(k_bwd, dim3((unsigned)it->fronts_cnt, nb), dim3(256), lds, s,
                       dm.tasks.p + it->fronts_off, A, pool, sy.pool, nrhs, sy.n, y, x);
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
