// Mode field arrays for plotting (the `sol1` / `mode_pol` outputs of
// calc_modes_2d and `sol_P2` of calc_modes_1d), produced from the DOF vectors
// only when the caller asks for them.
//
// 2D sol (array_sol.f:120-263): sol(3, 13, nval, nel) holds the field at the
// 6 P2 nodes (Ex, Ey from the 12 edge DOFs through the per-element
// interpolation block gt, Ez from the 10 P3 DOFs through phi3 at the node)
// and Ez at the 7 remaining P3 nodes.
// 2D mode_pol (mode_energy.f:60-121): per mode, the integrals over the unit
// cell of eps |E_x|^2, eps |E_y|^2, eps |beta E_z|^2 and of eps |E|^2 over the
// core material, with E the P2 interpolant of the node values, on the
// 16-point rule, each divided by the sum of the first three.
#include "engine.h"
#include "fem_basis.h"

namespace emu {

__constant__ double c_p3at[6][10];  // phi3 at the six P2 nodes
__constant__ double c_p2q[16][7];   // phi2 at the 16 quadrature points, weight in [6]

// Field of mode column x at the 6 P2 nodes of element e: f[node][component]
__device__ __forceinline__ void node_fields(int64_t e, const int32_t *__restrict__ ldof,
                                            const double *__restrict__ gt, const zd *__restrict__ x,
                                            const zd *__restrict__ ve, zd f[6][3], zd co[22]) {
  const zd *vb = ve + e * 14;
#pragma unroll
  for (int l = 0; l < 22; ++l) {
    const int ent = l < 12 ? l / 3 : 4 + (l - 12);
    co[l] = zmul(x[ldof[22 * e + l]], vb[ent]);
  }
  const double *G = gt + 144 * e;
  for (int inod = 0; inod < 6; ++inod) {
    zd f0 = zmk(0, 0), f1 = zmk(0, 0), fz = zmk(0, 0);
#pragma unroll
    for (int l = 0; l < 12; ++l) {
      f0 = zfma(zmk(G[(2 * inod) * 12 + l], 0), co[l], f0);
      f1 = zfma(zmk(G[(2 * inod + 1) * 12 + l], 0), co[l], f1);
    }
#pragma unroll
    for (int m = 0; m < 10; ++m) fz = zfma(zmk(c_p3at[inod][m], 0), co[12 + m], fz);
    f[inod][0] = f0;
    f[inod][1] = f1;
    f[inod][2] = fz;
  }
}

// One thread per (element, mode): the 13-node field of the element
__global__ __launch_bounds__(256) void k_fields_2d(int64_t nel, int nval, const int32_t *__restrict__ ldof,
                                                   const double *__restrict__ gt, const zd *__restrict__ X,
                                                   int64_t ldx, const zd *__restrict__ ve, zd *__restrict__ sol) {
  const int64_t e = blockIdx.y * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.x;  // modes fastest: the element range's data stays in L2
  if (e >= nel) return;
  zd f[6][3], co[22];
  node_fields(e, ldof, gt, X + (int64_t)k * ldx, ve, f, co);
  zd *out = sol + 39 * ((int64_t)k + (int64_t)nval * e);
  for (int inod = 0; inod < 6; ++inod)
    for (int c = 0; c < 3; ++c) out[3 * inod + c] = f[inod][c];
  for (int j = 0; j < 7; ++j) {
    out[3 * (6 + j) + 0] = zmk(0, 0);
    out[3 * (6 + j) + 1] = zmk(0, 0);
    out[3 * (6 + j) + 2] = co[12 + 3 + j];
  }
}

// One thread per (element, mode): eps |det| sum_q w_q |E_c(x_q)|^2 for
// c = x, y, z (E_z scaled by beta) and their sum on core elements
__global__ __launch_bounds__(256) void k_energy_2d(int64_t nel, int nval, const int32_t *__restrict__ ldof,
                                                   const double *__restrict__ gt, const double *__restrict__ det,
                                                   const int32_t *__restrict__ type_el, int core_t,
                                                   const zd *__restrict__ X, int64_t ldx, const zd *__restrict__ ve,
                                                   const zd *__restrict__ beta, const zd *__restrict__ eps,
                                                   zd *__restrict__ part) {
  const int64_t e = blockIdx.y * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.x;  // modes fastest: the element range's data stays in L2
  if (e >= nel) return;
  zd f[6][3], co[22];
  node_fields(e, ldof, gt, X + (int64_t)k * ldx, ve, f, co);
  double acc[3] = {0, 0, 0};
  for (int q = 0; q < 16; ++q) {
    zd v[3] = {zmk(0, 0), zmk(0, 0), zmk(0, 0)};
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
      for (int c = 0; c < 3; ++c) v[c] = zfma(zmk(c_p2q[q][i], 0), f[i][c], v[c]);
    v[2] = zmul(v[2], beta[k]);
#pragma unroll
    for (int c = 0; c < 3; ++c) acc[c] += c_p2q[q][6] * (v[c].x * v[c].x + v[c].y * v[c].y);
  }
  const int t = type_el[e];
  const zd w = zmul(zmk(fabs(det[e]), 0), eps[t]);
  zd *p = part + 4 * ((int64_t)k * nel + e);
  p[0] = zmul(w, zmk(acc[0], 0));
  p[1] = zmul(w, zmk(acc[1], 0));
  p[2] = zmul(w, zmk(acc[2], 0));
  p[3] = t == core_t ? zmul(w, zmk(acc[0] + acc[1] + acc[2], 0)) : zmk(0, 0);
}

// mode_pol(:, k) = element sums of the parts divided by the x + y + z total
__global__ __launch_bounds__(256) void k_mode_pol(int64_t nel, const zd *__restrict__ part, zd *__restrict__ pol) {
  const int k = blockIdx.x, tid = threadIdx.x;
  zd acc[4] = {zmk(0, 0), zmk(0, 0), zmk(0, 0), zmk(0, 0)};
  for (int64_t e = tid; e < nel; e += 256)
    for (int q = 0; q < 4; ++q) acc[q] = zadd(acc[q], part[4 * ((int64_t)k * nel + e) + q]);
  __shared__ zd red[4][256];
  for (int q = 0; q < 4; ++q) red[q][tid] = acc[q];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st)
      for (int q = 0; q < 4; ++q) red[q][tid] = zadd(red[q][tid], red[q][tid + st]);
    __syncthreads();
  }
  if (tid == 0) {
    const zd tot = zadd(zadd(red[0][0], red[1][0]), red[2][0]);
    const bool ok = tot.x != 0.0 || tot.y != 0.0;
    for (int q = 0; q < 4; ++q) pol[4 * k + q] = ok ? zdiv(red[q][0], tot) : zmk(0, 0);
  }
}

static void fields_tables() {
  static bool init = false;
  if (init) return;
  double tab[6][10];
  const double xn[6][2] = {{0, 0}, {1, 0}, {0, 1}, {0.5, 0}, {0.5, 0.5}, {0, 0.5}};
  for (int i = 0; i < 6; ++i) {
    double p[10], gx[10], gy[10];
    ref_p3(xn[i][0], xn[i][1], p, gx, gy);
    for (int m = 0; m < 10; ++m) tab[i][m] = p[m];
  }
  EMU_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_p3at), tab, sizeof(tab)));
  double q2[16][7];
  for (int q = 0; q < 16; ++q) {
    double x, y, w, p[6], gx[6], gy[6];
    quad16(q, x, y, w);
    ref_p2(x, y, p, gx, gy);
    for (int i = 0; i < 6; ++i) q2[q][i] = p[i];
    q2[q][6] = w;
  }
  EMU_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_p2q), q2, sizeof(q2)));
  init = true;
}

void launch_fields_2d(DevMesh &dm, int nval, const zd *X, int64_t ldx, const zd *ve, zd *sol, hipStream_t s) {
  fields_tables();
  const int64_t nel = dm.host.nel;
  hipLaunchKernelGGL(k_fields_2d, dim3(nval, (unsigned)((nel + 255) / 256)), dim3(256), 0, s, nel, nval, dm.ldof.p,
                     dm.gt.p, X, ldx, ve, sol);
  EMU_LAUNCH_CHECK();
}

void launch_mode_pol_2d(DevMesh &dm, int nval, const zd *X, int64_t ldx, const zd *ve, const zd *beta,
                        const zd *eps, int core_t, zd *part, zd *pol, hipStream_t s) {
  fields_tables();
  const int64_t nel = dm.host.nel;
  hipLaunchKernelGGL(k_energy_2d, dim3(nval, (unsigned)((nel + 255) / 256)), dim3(256), 0, s, nel, nval, dm.ldof.p,
                     dm.gt.p, dm.det.p, dm.type_el.p, core_t, X, ldx, ve, beta, eps, part);
  EMU_LAUNCH_CHECK();
  hipLaunchKernelGGL(k_mode_pol, dim3(nval), dim3(256), 0, s, nel, part, pol);
  EMU_LAUNCH_CHECK();
}

// 1D: sol_P2(3, 3, nval, nel) of array_sol_P2_1d.f: x at the three P2 nodes
// (left, right, mid), y and z at the ends from the P3 end values and at the
// mid node from the four P3 values with weights (-1, -1, 9, 9)/16.
__global__ __launch_bounds__(256) void k_sol_p2_1d(int64_t nel, int nval, const int32_t *__restrict__ ldof,
                                                   const zd *__restrict__ X, int64_t ldx, double kx, double dxw,
                                                   zd *__restrict__ out) {
  const int64_t e = blockIdx.y * (int64_t)blockDim.x + threadIdx.x;
  const int k = blockIdx.x;  // modes fastest: the element range's data stays in L2
  if (e >= nel) return;
  const zd *x = X + (int64_t)k * ldx;
  zd v[11];
  for (int l = 0; l < 11; ++l) {
    zd ph = zmk(1, 0);
    if (e == nel - 1 && (l == 4 || l == 8)) {  // right end wraps to the first element's left end
      double sn, cs;
      sincos(kx * dxw, &sn, &cs);
      ph = zmk(cs, sn);
    }
    v[l] = zmul(x[ldof[11 * e + l]], ph);
  }
  const double w[4] = {-1.0 / 16.0, -1.0 / 16.0, 9.0 / 16.0, 9.0 / 16.0};
  zd ym = zmk(0, 0), zm = zmk(0, 0);
  for (int j = 0; j < 4; ++j) {
    ym = zfma(zmk(w[j], 0), v[3 + j], ym);
    zm = zfma(zmk(w[j], 0), v[7 + j], zm);
  }
  zd *o = out + 9 * ((int64_t)k + (int64_t)nval * e);
  o[0] = v[0];
  o[3] = v[1];
  o[6] = v[2];
  o[1] = v[3];
  o[2] = v[7];
  o[4] = v[4];
  o[5] = v[8];
  o[7] = ym;
  o[8] = zm;
}

void launch_sol_p2_1d(DevMesh &dm, int nval, const zd *X, int64_t ldx, double kx, zd *out, hipStream_t s) {
  const int64_t nel = dm.host.nel;
  hipLaunchKernelGGL(k_sol_p2_1d, dim3(nval, (unsigned)((nel + 255) / 256)), dim3(256), 0, s, nel, nval, dm.ldof.p, X,
                     ldx, kx, dm.host.period_x, out);
  EMU_LAUNCH_CHECK();
}

}  // namespace emu
