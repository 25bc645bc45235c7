// Kernels of the 1D-grating mode path, batched over instances (blockIdx.y):
// gather assembly of the pencil and of the overlap form
// (backend/fortran/asmbly_1d.f:139-283, orthogonal_1d.f:108-189), the
// plane-wave load matrices of J / J^dagger (J_overlap_1d.f,
// J_dagger_overlap_1d.f) and the field normalisation of array_sol_1d.f.
// The eigensolve and the contractions reuse the sparse LU / Krylov-Schur /
// dense kernels of the 2D path on the 1D pattern.
#include "engine.h"

namespace emu {

namespace {

// Local DOFs of a 1D element: 0..2 x (P2: left, right, mid), 3..6 y and
// 7..10 z (P3: left, right, 1/3, 2/3).
__device__ __forceinline__ int grp(int l) { return l < 3 ? 0 : (l < 7 ? 1 : 2); }
__device__ __forceinline__ int loc(int l) { return l < 3 ? l : (l < 7 ? l - 3 : l - 7); }

// entry (i, j) of the element blocks (rows = trial functions); which = 0:
// (K, M) of asmbly_1d.f:139-258, which = 1: K = overlap form of
// orthogonal_1d.f:108-160
__device__ void blk1d(int i, int j, double h, zd pp, zd qq, double ky, const double *__restrict__ tab, int which,
                      zd &K, zd &M) {
  const int gi = grp(i), gj = grp(j), a = loc(i), b = loc(j);
  auto Mxx = [&](int r, int c) { return h * tab[kMxx + 3 * r + c]; };
  auto Kxy = [&](int r, int c) { return tab[kKxy + 4 * r + c]; };  // r: P2, c: P3
  auto Kyy = [&](int r, int c) { return tab[kKyy + 4 * r + c] / h; };
  auto Myy = [&](int r, int c) { return h * tab[kMyy + 4 * r + c]; };
  const zd ikypp = zmul(zmk(0, ky), pp);
  const zd ppk2mq = zsub(zscale(pp, ky * ky), qq);
  K = zmk(0, 0);
  M = zmk(0, 0);
  if (which == 0) {
    if (gi == 0 && gj == 0) {
      K = zscale(ppk2mq, Mxx(a, b));
      M = zscale(pp, -Mxx(a, b));
    } else if (gi == 1 && gj == 0) {
      K = zscale(ikypp, -Kxy(b, a));
    } else if (gi == 2 && gj == 0) {
      K = zscale(pp, -Kxy(b, a));
    } else if (gi == 0 && gj == 1) {
      K = zscale(ikypp, Kxy(a, b));
    } else if (gi == 1 && gj == 1) {
      K = zsub(zscale(pp, Kyy(a, b)), zscale(qq, Myy(a, b)));
      M = zscale(pp, -Myy(a, b));
    } else if (gi == 2 && gj == 1) {
      K = zscale(ikypp, Myy(a, b));
    } else if (gi == 0 && gj == 2) {
      M = zscale(pp, Kxy(a, b));
    } else if (gi == 1 && gj == 2) {
      M = zscale(ikypp, Myy(a, b));
    } else if (gi == 2 && gj == 2) {
      K = zadd(zscale(pp, Kyy(a, b)), zscale(ppk2mq, Myy(a, b)));
    }
  } else {
    if (gi == 0 && gj == 0) K = zscale(pp, Mxx(a, b));
    else if (gi == 1 && gj == 1) K = zscale(pp, Myy(a, b));
    else if (gi == 0 && gj == 2) K = zscale(pp, -Kxy(a, b));
    else if (gi == 1 && gj == 2) K = zscale(ikypp, -Myy(a, b));
  }
}

__device__ __forceinline__ zd phase1d(int64_t e, int l, int64_t nel, double kx, double dxw) {
  if (e == nel - 1 && (l == 4 || l == 8)) {
    double sn, cs;
    sincos(kx * dxw, &sn, &cs);
    return zmk(cs, sn);
  }
  return zmk(1, 0);
}

}  // namespace

__global__ __launch_bounds__(256) void k_gather_1d(int64_t nnz, int64_t nel, int ntyp,
                                                   const int32_t *__restrict__ cptr, const int32_t *__restrict__ citem,
                                                   const int32_t *__restrict__ type_el, const double *__restrict__ h1d,
                                                   const double *__restrict__ tab, const zd *__restrict__ pp,
                                                   const zd *__restrict__ qq, const zd *__restrict__ shift,
                                                   const double *__restrict__ bloch, double dxw, int which,
                                                   zd *__restrict__ out1, zd *__restrict__ out2) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (k >= nnz) return;
  const double kx = bloch[2 * b], ky = bloch[2 * b + 1];
  const zd sh = shift[b];
  zd a1 = zmk(0, 0), a2 = zmk(0, 0);
  for (int32_t c = cptr[k]; c < cptr[k + 1]; ++c) {
    const int32_t it = citem[c];
    const int64_t e = it / 121;
    const int ij = it - (int)(e * 121), i = ij / 11, j = ij - 11 * (ij / 11);
    const int t = type_el[e];
    zd K, M;
    blk1d(i, j, h1d[e], pp[(int64_t)b * ntyp + t], qq[(int64_t)b * ntyp + t], ky, tab, which, K, M);
    const zd ph = zmul(zconj(phase1d(e, i, nel, kx, dxw)), phase1d(e, j, nel, kx, dxw));
    if (which == 0) {
      a1 = zadd(a1, zmul(zsub(K, zmul(sh, M)), ph));
      a2 = zadd(a2, zmul(M, ph));
    } else {
      a1 = zadd(a1, zmul(K, ph));
    }
  }
  out1[(int64_t)b * nnz + k] = a1;
  if (which == 0) out2[(int64_t)b * nnz + k] = a2;
}

void launch_gather_1d(DevMesh &dm, int nb, int ntyp, const zd *pp, const zd *qq, const zd *shift,
                      const double *bloch, int which, zd *out1, zd *out2, hipStream_t s) {
  const Mesh &m = dm.host;
  dim3 g((unsigned)((m.nnz + 255) / 256), nb);
  hipLaunchKernelGGL(k_gather_1d, g, dim3(256), 0, s, m.nnz, m.nel, ntyp, dm.cptr.p, dm.citem.p, dm.type_el.p,
                     dm.h1d.p, dm.tab1d.p, pp, qq, shift, bloch, m.period_x, which, out1, out2);
  EMU_LAUNCH_CHECK();
}

// ------------------------------------------------- plane-wave load matrices
namespace {

// 16-point Gauss-Legendre rule on [0, 1]
__constant__ double c_gl16_t[16] = {
    0.0052995325041750, 0.0277124884633837, 0.0671843988060841, 0.1222977958224985,
    0.1910618777986781, 0.2709916111713863, 0.3591982246103705, 0.4524937450811813,
    0.5475062549188188, 0.6408017753896295, 0.7290083888286136, 0.8089381222013219,
    0.8777022041775016, 0.9328156011939159, 0.9722875115366163, 0.9947004674958250};
__constant__ double c_gl16_w[16] = {
    0.0135762297058770, 0.0311267619693240, 0.0475792558412464, 0.0623144856277669,
    0.0747979944082884, 0.0845782596975013, 0.0913017075224618, 0.0947253052275342,
    0.0947253052275342, 0.0913017075224618, 0.0845782596975013, 0.0747979944082884,
    0.0623144856277669, 0.0475792558412464, 0.0311267619693240, 0.0135762297058770};

__device__ __forceinline__ double lag1d(int l, double t) {
  // P2 on (0, 1, 1/2) for l < 3; P3 on (0, 1, 1/3, 2/3) for l in 3..6
  switch (l) {
    case 0: return 2.0 * (t - 1.0) * (t - 0.5);
    case 1: return 2.0 * t * (t - 0.5);
    case 2: return -4.0 * t * (t - 1.0);
    case 3: return -4.5 * (t - 1.0 / 3.0) * (t - 2.0 / 3.0) * (t - 1.0);
    case 4: return 4.5 * t * (t - 1.0 / 3.0) * (t - 2.0 / 3.0);
    case 5: return 13.5 * t * (t - 2.0 / 3.0) * (t - 1.0);
    default: return -13.5 * t * (t - 1.0 / 3.0) * (t - 1.0);
  }
}

}  // namespace

// thread per (dof r, plane wave s) of instance b: E-type column s, K-type P+s
__global__ __launch_bounds__(256) void k_pw_load_1d(int64_t n, int P, int64_t nel, const int32_t *__restrict__ dcptr,
                                                    const int32_t *__restrict__ dcitem,
                                                    const double *__restrict__ xd1d, const double *__restrict__ h1d,
                                                    const double *__restrict__ alpha, const double *__restrict__ kyv,
                                                    const int32_t *__restrict__ sign,
                                                    const double *__restrict__ bloch, double dxw,
                                                    zd *__restrict__ G) {
  const int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int s = blockIdx.z;
  const int b = blockIdx.y;
  if (r >= n) return;
  const double a = alpha[(int64_t)b * P + s], ky = kyv[b], sg = sign[b];
  double Ex, Ey, Kx, Ky;
  if (a != 0.0 || ky != 0.0) {
    const double nr = 1.0 / sqrt(a * a + ky * ky);
    Kx = a * nr; Ky = ky * nr; Ex = ky * nr; Ey = -a * nr;
  } else {
    Kx = 0; Ky = 1; Ex = 1; Ey = 0;
  }
  zd gE = zmk(0, 0), gK = zmk(0, 0);
  for (int32_t c = dcptr[r]; c < dcptr[r + 1]; ++c) {
    const int32_t it = dcitem[c];
    const int64_t e = it / 11;
    const int l = it - (int)(e * 11);
    const double xmin = xd1d[e * 11], h = h1d[e];
    zd acc = zmk(0, 0);
    for (int q = 0; q < 16; ++q) {
      const double t = c_gl16_t[q];
      double sn, cs;
      sincos(sg * a * (xmin + h * t), &sn, &cs);
      const double w = c_gl16_w[q] * h * lag1d(l, t);
      acc = zadd(acc, zmk(w * cs, w * sn));
    }
    acc = zmul(acc, phase1d(e, l, nel, bloch[2 * b], dxw));
    gE = zadd(gE, zscale(acc, l < 3 ? Ex : Ey));
    gK = zadd(gK, zscale(acc, l < 3 ? Kx : Ky));
  }
  zd *Gb = G + (int64_t)b * 2 * P * n;
  Gb[(int64_t)s * n + r] = gE;
  Gb[(int64_t)(P + s) * n + r] = gK;
}

void launch_pw_load_1d(DevMesh &dm, int nb, int P, const double *alpha, const double *ky, const int32_t *sign,
                       const double *bloch, zd *G, hipStream_t s) {
  const Mesh &m = dm.host;
  dim3 g((unsigned)((m.neq + 255) / 256), nb, P);
  hipLaunchKernelGGL(k_pw_load_1d, g, dim3(256), 0, s, m.neq, P, m.nel, dm.dcptr.p, dm.dcitem.p, dm.xd1d.p,
                     dm.h1d.p, alpha, ky, sign, bloch, m.period_x, G);
  EMU_LAUNCH_CHECK();
}

// ---------------------------------------------------- field normalisation
// One workgroup per (mode k, instance b): first-occurrence maxima of |value|
// over the element DOFs in (element, local) order -- overall, left half,
// right half, and within dq of the centre -- then the array_sol_1d.f choice:
// the centre value when it is at least 0.95 of the overall maximum, else the
// right-half maximum (prime, even b) or the left-half one (adjoint, odd b).
struct Max1 {
  double a;
  int64_t i;
  zd v;
};

__device__ __forceinline__ void upd(Max1 &m, double a, int64_t i, zd v) {
  if (a > m.a || (a == m.a && i < m.i)) {
    m.a = a;
    m.i = i;
    m.v = v;
  }
}

__global__ __launch_bounds__(256) void k_field_max_1d(int64_t nel, int64_t n, int nval, const int32_t *__restrict__ ldof,
                                                      const double *__restrict__ xd1d, const zd *__restrict__ X,
                                                      int64_t sX, const double *__restrict__ bloch, double dxw,
                                                      double xmid, double dq, zd *__restrict__ den) {
  const int k = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const zd *x = X + (int64_t)b * sX + (int64_t)k * n;
  const double kx = bloch[2 * b];
  Max1 m[4];
  for (int q = 0; q < 4; ++q) m[q] = {0.0, INT64_MAX, zmk(0, 0)};
  for (int64_t idx = tid; idx < nel * 11; idx += 256) {
    const int64_t e = idx / 11;
    const int l = (int)(idx - e * 11);
    const zd v = zmul(x[ldof[idx]], phase1d(e, l, nel, kx, dxw));
    const double a = sqrt(v.x * v.x + v.y * v.y);
    const double xx = xd1d[idx];
    if (a > 0) {
      upd(m[0], a, idx, v);
      if (xx < xmid) upd(m[1], a, idx, v);
      if (xx > xmid) upd(m[2], a, idx, v);
      if (fabs(xx - xmid) < dq) upd(m[3], a, idx, v);
    }
  }
  __shared__ double sa[4][256];
  __shared__ int64_t si[4][256];
  __shared__ zd sv[4][256];
  for (int q = 0; q < 4; ++q) {
    sa[q][tid] = m[q].a;
    si[q][tid] = m[q].i;
    sv[q][tid] = m[q].v;
  }
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st)
      for (int q = 0; q < 4; ++q) {
        const double a2 = sa[q][tid + st];
        const int64_t i2 = si[q][tid + st];
        if (a2 > sa[q][tid] || (a2 == sa[q][tid] && i2 < si[q][tid])) {
          sa[q][tid] = a2;
          si[q][tid] = i2;
          sv[q][tid] = sv[q][tid + st];
        }
      }
    __syncthreads();
  }
  if (tid == 0) {
    const zd zl = sv[1][0], zr = sv[2][0], zm = sv[3][0];
    zd d = (b & 1) ? zl : zr;
    if (sa[0][0] > 0 && sa[3][0] / sa[0][0] >= 0.95) d = zm;
    if (d.x == 0.0 && d.y == 0.0) d = zmk(1, 0);
    den[(int64_t)b * nval + k] = d;
  }
}

void launch_field_max_1d(DevMesh &dm, int nb, int nval, const zd *X, int64_t sX, const double *bloch, double xmid,
                         double dq, zd *den, hipStream_t s) {
  const Mesh &m = dm.host;
  hipLaunchKernelGGL(k_field_max_1d, dim3(nval, nb), dim3(256), 0, s, m.nel, m.neq, nval, dm.ldof.p, dm.xd1d.p, X,
                     sX, bloch, m.period_x, xmid, dq, den);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu
