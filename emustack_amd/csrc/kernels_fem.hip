// FEM assembly kernels of K - shift*M and M (reference: asmbly.f:3-341).
//
// Design: the element integrals depend only on the geometry, so they are
// computed once per mesh as three real 22x22 blocks per element,
//     K_e = Ka*pp + Kc*qq,   M_e = Md*pp,
// and every (wavelength, Bloch vector) instance only applies the material
// constants and the Bloch phase while gathering into the CSR values.  The
// gather is atomic-free: each output nonzero sums its own precomputed list
// of element contributions, so the result is deterministic and every CSR
// value is written exactly once (coalesced).
#include "engine.h"
#include "fem_basis.h"

namespace emu {

// One workgroup per element.  rec[q][l] = (vx, vy, curl) for transverse l<12
// and (phi, gx, gy) for the P3 functions l>=12.
__global__ __launch_bounds__(256) void k_elem_integrals(int64_t nel, const double *__restrict__ geo,
                                                        const double *__restrict__ detv,
                                                        const signed char *__restrict__ orient,
                                                        double *__restrict__ eka,
                                                        double *__restrict__ ekc,
                                                        double *__restrict__ emd) {
  __shared__ double rec[16][22][3];
  __shared__ double wq[16];
  const int64_t e = blockIdx.x;
  if (e >= nel) return;
  const double *T = geo + 8 * e + 4;
  const signed char *ori = orient + 6 * e;
  const double adet = fabs(detv[e]);
  for (int t = threadIdx.x; t < 16 * 22; t += blockDim.x) {
    int q = t / 22, l = t % 22;
    double x, y, w;
    quad16(q, x, y, w);
    if (l == 0) wq[q] = w * adet;
    if (l < 12) {
      double vx, vy, c;
      transverse_basis(l, x, y, T, ori, vx, vy, c);
      rec[q][l][0] = vx;
      rec[q][l][1] = vy;
      rec[q][l][2] = c;
    } else {
      double p[10], gx[10], gy[10];
      ref_p3(x, y, p, gx, gy);
      int m = l - 12;
      double ox, oy;
      phys_grad(T, gx[m], gy[m], ox, oy);
      rec[q][l][0] = p[m];
      rec[q][l][1] = ox;
      rec[q][l][2] = oy;
    }
  }
  __syncthreads();
  for (int t = threadIdx.x; t < 484; t += blockDim.x) {
    int i = t / 22, j = t % 22;  // row i (trial), column j (test)
    double ka = 0, kc = 0, md = 0;
    for (int q = 0; q < 16; ++q) {
      double w = wq[q];
      const double *ri = rec[q][i];
      const double *rj = rec[q][j];
      if (i < 12 && j < 12) {
        double dot = ri[0] * rj[0] + ri[1] * rj[1];
        ka += w * (ri[2] * rj[2]);  // curl_j curl_i pp
        kc -= w * dot;              // - (phi_j . phi_i) qq
        md -= w * dot;              // M_tt = -(phi_j . phi_i) pp
      } else if (i < 12) {          // M_tz = (grad z_j . phi_i) pp
        md += w * (rj[1] * ri[0] + rj[2] * ri[1]);
      } else if (j < 12) {          // K_zt = (phi_j . grad z_i) pp
        ka += w * (rj[0] * ri[1] + rj[1] * ri[2]);
      } else {                      // K_zz = -(grad_j.grad_i) pp + z_j z_i qq
        ka -= w * (ri[1] * rj[1] + ri[2] * rj[2]);
        kc += w * (ri[0] * rj[0]);
      }
    }
    eka[484 * e + t] = ka;
    ekc[484 * e + t] = kc;
    emd[484 * e + t] = md;
  }
}

void launch_elem_integrals(DevMesh &dm, hipStream_t s) {
  int64_t nel = dm.host.nel;
  dm.eka.alloc(484 * nel);
  dm.ekc.alloc(484 * nel);
  dm.emd.alloc(484 * nel);
  hipLaunchKernelGGL(k_elem_integrals, dim3((unsigned)nel), dim3(256), 0, s, nel, dm.geo.p,
                     dm.det.p, dm.orient.p, dm.eka.p, dm.ekc.p, dm.emd.p);
  EMU_LAUNCH_CHECK();
}

// Bloch phase per element entity: exp(i bloch . (x_E - x_rep)) (asmbly.f:121-146)
__global__ void k_phase(int64_t nel, int nb, const double *__restrict__ bloch,
                        const double *__restrict__ dphase, zd *__restrict__ ve) {
  int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (t >= nel * 14) return;
  double kx = bloch[2 * b], ky = bloch[2 * b + 1];
  double arg = kx * dphase[2 * t] + ky * dphase[2 * t + 1];
  double sn, cs;
  sincos(arg, &sn, &cs);
  ve[b * nel * 14 + t] = zmk(cs, sn);
}

void launch_phase(DevMesh &dm, int nb, const double *bloch_dev, zd *ve, hipStream_t s) {
  int64_t n = dm.host.nel * 14;
  dim3 g((unsigned)((n + 255) / 256), nb);
  hipLaunchKernelGGL(k_phase, g, dim3(256), 0, s, dm.host.nel, nb, bloch_dev, dm.dphase.p, ve);
  EMU_LAUNCH_CHECK();
}

__device__ __forceinline__ int local_entity(int l) { return l < 12 ? l / 3 : 4 + (l - 12); }

// One thread per (nonzero, instance).
__global__ __launch_bounds__(256) void k_gather(int64_t nnz, int64_t nel, int ntyp,
                                                const int32_t *__restrict__ cptr,
                                                const int32_t *__restrict__ citem,
                                                const int32_t *__restrict__ type_el,
                                                const double *__restrict__ eka,
                                                const double *__restrict__ ekc,
                                                const double *__restrict__ emd,
                                                const zd *__restrict__ pp, const zd *__restrict__ qq,
                                                const zd *__restrict__ shift, const zd *__restrict__ ve,
                                                zd *__restrict__ mat1, zd *__restrict__ mat2) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (k >= nnz) return;
  const zd sh = shift[b];
  const zd *veb = ve + (int64_t)b * nel * 14;
  zd a1 = zmk(0, 0), a2 = zmk(0, 0);
  for (int32_t c = cptr[k]; c < cptr[k + 1]; ++c) {
    int32_t item = citem[c];
    int32_t e = item / 484, ij = item - 484 * e;
    int i = ij / 22, j = ij - 22 * i;
    int t = type_el[e];
    zd p = pp[b * ntyp + t], q = qq[b * ntyp + t];
    double ka = eka[item], kc = ekc[item], md = emd[item];
    zd K = zmk(ka * p.x + kc * q.x, ka * p.y + kc * q.y);
    zd M = zmk(md * p.x, md * p.y);
    zd fac = zmul(veb[14 * (int64_t)e + local_entity(j)], zconj(veb[14 * (int64_t)e + local_entity(i)]));
    zd m1 = zsub(K, zmul(sh, M));
    a1 = zfma(m1, fac, a1);
    a2 = zfma(M, fac, a2);
  }
  mat1[b * nnz + k] = a1;
  mat2[b * nnz + k] = a2;
}

void launch_gather(DevMesh &dm, int nb, int ntyp, const zd *pp, const zd *qq, const zd *shift,
                   const zd *ve, zd *mat1, zd *mat2, hipStream_t s) {
  int64_t nnz = dm.host.nnz;
  dim3 g((unsigned)((nnz + 255) / 256), nb);
  hipLaunchKernelGGL(k_gather, g, dim3(256), 0, s, nnz, dm.host.nel, ntyp, dm.cptr.p, dm.citem.p,
                     dm.type_el.p, dm.eka.p, dm.ekc.p, dm.emd.p, pp, qq, shift, ve, mat1, mat2);
  EMU_LAUNCH_CHECK();
}

// dst[b][k'] = src[b][pnnz_src[k']] (values of the ND-permuted CSR)
__global__ void k_permute_vals(int64_t nnz, const int32_t *__restrict__ srcidx,
                               const zd *__restrict__ src, zd *__restrict__ dst) {
  int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int b = blockIdx.y;
  if (k >= nnz) return;
  dst[b * nnz + k] = src[b * nnz + srcidx[k]];
}

void launch_permute_vals(DevMesh &dm, int nb, const zd *src, zd *dst, hipStream_t s) {
  int64_t nnz = dm.host.nnz;
  dim3 g((unsigned)((nnz + 255) / 256), nb);
  hipLaunchKernelGGL(k_permute_vals, g, dim3(256), 0, s, nnz, dm.pnnz_src.p, src, dst);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu
