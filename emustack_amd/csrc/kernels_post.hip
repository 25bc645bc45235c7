// Mode post-processing on the GPU: field normalisation (array_sol.f:120-339),
// and plane-wave load matrices for the J / J^dagger overlaps
// (J_overlap.f:117-186, J_dagger_overlap.f:115-186).
//
// The reference interpolates every mode to per-element 13-node field arrays
// (sol, nel*13*3*nval complex) and then integrates each plane wave against
// them element by element.  Here the modes stay as DOF vectors X (neq x nval):
//  * the normalisation constant of a mode is the first largest field value
//    at the P2 nodes (scan order element, node, component), computed directly
//    from the DOF coefficients with a per-element 12x12 interpolation block;
//  * J = G^T X where G (neq x 2P) is the plane-wave "load" of every DOF,
//    assembled once per instance from element quadrature; this turns the
//    overlap into one tall-skinny contraction.
#include "engine.h"
#include "fem_basis.h"

namespace emu {

__device__ __forceinline__ int ent_of(int l) { return l < 12 ? l / 3 : 4 + (l - 12); }

struct MaxRec {
  double a;
  int64_t idx;
  zd v;
};

__device__ __forceinline__ void max_merge(MaxRec &x, const MaxRec &y) {
  if (y.a > x.a || (y.a == x.a && y.idx < x.idx)) x = y;
}

// phi3 at the six P2 nodes
__constant__ double c_p3node[6][10];

// One workgroup per 64-element chunk of one instance: the chunk's
// interpolation blocks (64 x 144 doubles), Bloch phases and DOF indices are
// staged in LDS once, then wave q walks modes q, q + 4, ...; lane l holds
// element e0 + l, computes its 18 P2-node values of the mode and the wave
// reduces them to the chunk's first largest one (scan order element, node,
// component).  Each mode's partials over the chunks are merged by
// k_field_max_final.
constexpr int FM_E = 64;

__device__ __forceinline__ void max_shfl(MaxRec &x, int m) {
  MaxRec y;
  y.a = __shfl_xor(x.a, m, 64);
  y.idx = __shfl_xor(x.idx, m, 64);
  y.v = zmk(__shfl_xor(x.v.x, m, 64), __shfl_xor(x.v.y, m, 64));
  max_merge(x, y);
}

__global__ __launch_bounds__(256) void k_field_max_partial(int64_t nel, int nval, const int32_t *__restrict__ ldof,
                                                           const double *__restrict__ gt, const zd *__restrict__ X,
                                                           int64_t ldx, int64_t sX, const zd *__restrict__ ve,
                                                           int conj_phase, MaxRec *__restrict__ part) {
  __shared__ double Gs[FM_E][145];
  __shared__ zd Ps[FM_E][15];
  __shared__ int32_t Ls[FM_E][23];
  const int chunk = blockIdx.x, nparts = gridDim.x;
  const int b = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wq = tid >> 6;
  const int64_t e0 = (int64_t)chunk * FM_E;
  const int ne = (int)min((int64_t)FM_E, nel - e0);
  for (int i = tid; i < ne * 144; i += 256) Gs[i / 144][i % 144] = gt[144 * e0 + i];
  const zd *vb = ve + ((int64_t)b * nel + e0) * 14;
  for (int i = tid; i < ne * 14; i += 256) {
    zd p = vb[i];
    if (conj_phase) p = zconj(p);
    Ps[i / 14][i % 14] = p;
  }
  for (int i = tid; i < ne * 22; i += 256) Ls[i / 22][i % 22] = ldof[22 * e0 + i];
  __syncthreads();
  const bool live = lane < ne;
  const int el = live ? lane : 0;
  const zd *xb = X + b * sX;
  for (int c = wq; c < nval; c += 4) {
    const zd *x = xb + (int64_t)c * ldx;
    MaxRec best;
    best.a = -1.0;
    best.idx = INT64_MAX;
    best.v = zmk(0, 0);
    if (live) {
      zd co[22];
#pragma unroll
      for (int l = 0; l < 22; ++l) co[l] = zmul(x[Ls[el][l]], Ps[el][ent_of(l)]);
      const int64_t e = e0 + el;
      for (int inod = 0; inod < 6; ++inod) {
        zd f0 = zmk(0, 0), f1 = zmk(0, 0), fz = zmk(0, 0);
#pragma unroll
        for (int l = 0; l < 12; ++l) {
          f0 = zfma(zmk(Gs[el][(2 * inod) * 12 + l], 0), co[l], f0);
          f1 = zfma(zmk(Gs[el][(2 * inod + 1) * 12 + l], 0), co[l], f1);
        }
#pragma unroll
        for (int m = 0; m < 10; ++m) fz = zfma(zmk(c_p3node[inod][m], 0), co[12 + m], fz);
        zd fv[3] = {f0, f1, fz};
        for (int cc = 0; cc < 3; ++cc) {
          double a = hypot(fv[cc].x, fv[cc].y);
          if (a > best.a) {
            best.a = a;
            best.idx = e * 18 + inod * 3 + cc;
            best.v = fv[cc];
          }
        }
      }
    }
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) max_shfl(best, m);
    if (lane == 0) part[((int64_t)b * nval + c) * nparts + chunk] = best;
  }
}

__global__ void k_field_max_final(int nparts, int nval, const MaxRec *__restrict__ part, zd *__restrict__ zmax) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (c >= nval) return;
  MaxRec best = part[((int64_t)b * nval + c) * nparts];
  for (int i = 1; i < nparts; ++i) max_merge(best, part[((int64_t)b * nval + c) * nparts + i]);
  zmax[(int64_t)b * nval + c] = best.v;
}

void launch_field_max(DevMesh &dm, int nb, int nval, const zd *X, int64_t ldx, int64_t sX, const zd *ve,
                      bool conj_phase, zd *zmax, double *scratch, hipStream_t s) {
  static bool init = false;
  if (!init) {
    double tab[6][10];
    const double xn[6][2] = {{0, 0}, {1, 0}, {0, 1}, {0.5, 0}, {0.5, 0.5}, {0, 0.5}};
    for (int i = 0; i < 6; ++i) {
      double p[10], gx[10], gy[10];
      ref_p3(xn[i][0], xn[i][1], p, gx, gy);
      for (int m = 0; m < 10; ++m) tab[i][m] = p[m];
    }
    EMU_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_p3node), tab, sizeof(tab)));
    init = true;
  }
  int64_t nel = dm.host.nel;
  int nparts = (int)((nel + FM_E - 1) / FM_E);
  MaxRec *part = reinterpret_cast<MaxRec *>(scratch);
  hipLaunchKernelGGL(k_field_max_partial, dim3(nparts, nb), dim3(256), 0, s, nel, nval, dm.ldof.p, dm.gt.p, X, ldx,
                     sX, ve, conj_phase ? 1 : 0, part);
  EMU_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_field_max_final, dim3((nval + 63) / 64, nb), dim3(64), 0, s, nparts, nval, part, zmax);
  EMU_LAUNCH_CHECK();
}

// X[b][:, c] *= f[b][c]  (or /= if invert)
__global__ void k_scale_cols(int64_t n, const zd *__restrict__ f, int64_t sf, int invert, zd *__restrict__ X,
                             int64_t ldx, int64_t sX) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int c = blockIdx.y, b = blockIdx.z;
  if (r >= n) return;
  zd g = f[b * sf + c];
  zd *p = X + b * sX + (int64_t)c * ldx + r;
  *p = invert ? zdiv(*p, g) : zmul(*p, g);
}

void launch_scale_cols(int nb, int64_t n, int ncol, zd *X, int64_t ldx, int64_t sX, const zd *f, int64_t sf,
                       bool invert, hipStream_t s) {
  hipLaunchKernelGGL(k_scale_cols, dim3((unsigned)((n + 255) / 256), ncol, nb), dim3(256), 0, s, n, f, sf,
                     invert ? 1 : 0, X, ldx, sX);
  EMU_LAUNCH_CHECK();
}

// Element plane-wave integrals JB[b][e][l][s2] for the 12 transverse functions:
//   sum_q w_q |det| exp(-/+ i (alpha x_q + beta y_q)) / |k_t| (PW_s . phi_l(x_q))
// E-type PW (beta, -alpha) at s2 = pos[s], K-type (alpha, beta) at P + pos[s].
__global__ __launch_bounds__(256) void k_pw_elem(int64_t nel, int P, const double *__restrict__ geo,
                                                 const double *__restrict__ detv, const double *__restrict__ x,
                                                 const int64_t *__restrict__ table_nod,
                                                 const signed char *__restrict__ orient,
                                                 const double *__restrict__ alpha, const double *__restrict__ beta,
                                                 const int32_t *__restrict__ pos, const int32_t *__restrict__ signs,
                                                 zd *__restrict__ JB) {
  const int64_t e = blockIdx.x;
  const int b = blockIdx.y;
  __shared__ double bv[16][12][2];
  __shared__ double xq[16][2];
  __shared__ double wq[16];
  const double *Bm = geo + 8 * e;
  const double *T = Bm + 4;
  const double adet = fabs(detv[e]);
  const int64_t n0 = table_nod[6 * e] - 1;
  const double x0 = x[2 * n0], y0 = x[2 * n0 + 1];
  for (int t = threadIdx.x; t < 16 * 12; t += blockDim.x) {
    int q = t / 12, l = t % 12;
    double xr, yr, w;
    quad16(q, xr, yr, w);
    double vx, vy, c;
    transverse_basis(l, xr, yr, T, orient + 6 * e, vx, vy, c);
    bv[q][l][0] = vx;
    bv[q][l][1] = vy;
    if (l == 0) {
      wq[q] = w * adet;
      xq[q][0] = x0 + Bm[0] * xr + Bm[2] * yr;
      xq[q][1] = y0 + Bm[1] * xr + Bm[3] * yr;
    }
  }
  __syncthreads();
  const double sign = (double)signs[b];
  const double *al = alpha + (int64_t)b * P;
  const double *be = beta + (int64_t)b * P;
  const int32_t *ps = pos + (int64_t)b * P;
  zd *out = JB + ((int64_t)b * nel + e) * 12 * 2 * P;
  for (int t = threadIdx.x; t < P * 12; t += blockDim.x) {
    int sidx = t / 12, l = t % 12;
    double a = al[sidx], bb = be[sidx];
    double nrm = 1.0 / sqrt(a * a + bb * bb);
    zd accE = zmk(0, 0), accK = zmk(0, 0);
    for (int q = 0; q < 16; ++q) {
      double arg = a * xq[q][0] + bb * xq[q][1];
      double sn, cs;
      sincos(arg, &sn, &cs);
      zd ex = zmk(cs * nrm * wq[q], sign * sn * nrm * wq[q]);
      double vx = bv[q][l][0], vy = bv[q][l][1];
      accE = zfma(zmk(bb * vx - a * vy, 0), ex, accE);
      accK = zfma(zmk(a * vx + bb * vy, 0), ex, accK);
    }
    int s2 = ps[sidx];
    out[l * 2 * P + s2] = accE;
    out[l * 2 * P + P + s2] = accK;
  }
}

// G[b][s2][dof] = sum over (e, l) of dof: JB[b][e][l][s2] * phase(e, l)
__global__ __launch_bounds__(256) void k_pw_gather(int64_t n, int64_t nel, int P2, const int32_t *__restrict__ dcptr,
                                                   const int32_t *__restrict__ dcitem, const zd *__restrict__ JB,
                                                   const zd *__restrict__ ve, zd *__restrict__ G) {
  int64_t dof = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int s2 = blockIdx.y;
  int b = blockIdx.z;
  if (dof >= n) return;
  zd acc = zmk(0, 0);
  for (int32_t c = dcptr[dof]; c < dcptr[dof + 1]; ++c) {
    int32_t item = dcitem[c];
    int32_t e = item / 12, l = item - 12 * e;
    zd p = ve[((int64_t)b * nel + e) * 14 + l / 3];
    acc = zfma(JB[(((int64_t)b * nel + e) * 12 + l) * P2 + s2], p, acc);
  }
  G[((int64_t)b * P2 + s2) * n + dof] = acc;
}

void launch_pw_load(DevMesh &dm, int nb, int P, const double *alpha, const double *beta, const int32_t *pwpos,
                    const int32_t *sign, const zd *ve, zd *JB, zd *G, hipStream_t s) {
  int64_t nel = dm.host.nel, n = dm.host.neq;
  hipLaunchKernelGGL(k_pw_elem, dim3((unsigned)nel, nb), dim3(256), 0, s, nel, P, dm.geo.p, dm.det.p, dm.xnod.p,
                     dm.table_nod.p, dm.orient.p, alpha, beta, pwpos, sign, JB);
  EMU_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_pw_gather, dim3((unsigned)((n + 255) / 256), 2 * P, nb), dim3(256), 0, s, n, nel, 2 * P,
                     dm.dcptr.p, dm.dcitem.p, JB, ve, G);
  EMU_LAUNCH_CHECK();
}

// dst[b][:, c] = src[b][:, idx[b*ncol + c]]
__global__ void k_permute_cols(int64_t n, int ncol, const zd *__restrict__ src, int64_t ss,
                               const int32_t *__restrict__ idx, zd *__restrict__ dst, int64_t sd) {
  int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int c = blockIdx.y, b = blockIdx.z;
  if (r >= n) return;
  dst[b * sd + (int64_t)c * n + r] = src[b * ss + (int64_t)idx[b * ncol + c] * n + r];
}

void launch_permute_cols(int nb, int64_t n, int ncol, const zd *src, int64_t ss, const int32_t *idx, zd *dst,
                         int64_t sd, hipStream_t s) {
  hipLaunchKernelGGL(k_permute_cols, dim3((unsigned)((n + 255) / 256), ncol, nb), dim3(256), 0, s, n, ncol, src, ss,
                     idx, dst, sd);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu
