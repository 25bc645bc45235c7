// Refractive index n(lambda) of many (material, wavelength) queries in one
// launch (reference backend/materials.py:129-186): constant indices, tables
// interpolated linearly, Drude metals.  One thread per query; the tables of
// all materials are concatenated (xs: wavelengths or Drude parameters, ys:
// complex indices) with per-material offsets.
#include "common.h"
#include "engine_api.h"
#include "engine.h"

namespace emu {

// principal square root (branch cut on the negative real axis)
__device__ __forceinline__ zd zsqrt_principal(zd z) {
  const double ax = fabs(z.x), ay = fabs(z.y);
  if (ax == 0.0 && ay == 0.0) return zmk(0.0, z.y);
  const double r = hypot(ax, ay);
  const double t = sqrt(0.5 * (ax + r));
  if (z.x >= 0.0) return zmk(t, z.y / (2.0 * t));
  return zmk(ay / (2.0 * t), copysign(t, z.y));
}

__global__ void k_material_index(int64_t nq, const int64_t *__restrict__ kind, const int64_t *__restrict__ off,
                                 const double *__restrict__ xs, const zd *__restrict__ ys,
                                 const int64_t *__restrict__ qm, const double *__restrict__ qw,
                                 zd *__restrict__ out) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nq) return;
  const int64_t m = qm[q];
  const double wl = qw[q];
  const int64_t a = off[m], b = off[m + 1];
  const int64_t k = kind[m];
  if (k == 0) {  // constant
    out[q] = ys[a];
    return;
  }
  if (k == 2) {  // Drude: omega_p, omega_gamma, eps_inf (rad/s)
    const double c = 299792458.0;
    const double wp = xs[a], wg = xs[a + 1], eps_inf = xs[a + 2];
    const double om = 2.0 * 3.141592653589793 * c / (wl * 1e-9);
    // eps = eps_inf - wp^2 / (om^2 + i wg om)
    const zd den = zmk(om * om, wg * om);
    const zd frac = zdiv(zmk(wp * wp, 0.0), den);
    out[q] = zsqrt_principal(zmk(eps_inf - frac.x, -frac.y));
    return;
  }
  // table: first point at or above wl is the upper end of the segment,
  // clamped to [a + 1, b - 1] (the segment choice of scipy's interp1d)
  int64_t lo = a, hi = b;  // lower_bound over [a, b)
  while (lo < hi) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (xs[mid] < wl) lo = mid + 1;
    else hi = mid;
  }
  int64_t h = lo < a + 1 ? a + 1 : (lo > b - 1 ? b - 1 : lo);
  const int64_t l = h - 1;
  const double dx = xs[h] - xs[l];
  const zd y0 = ys[l], y1 = ys[h];
  const zd slope = zdiv(zsub(y1, y0), zmk(dx, 0.0));
  const double t = wl - xs[l];
  out[q] = zmk(slope.x * t + y0.x, slope.y * t + y0.y);
}

void material_index(Ctx *c, int64_t nmat, const int64_t *kind, const int64_t *off, const double *xs,
                    const double *ys, int64_t nq, const int64_t *qm, const double *qw, double *out) {
  EMU_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int64_t npts = off[nmat];
  for (int64_t m = 0; m < nmat; ++m) {
    const int64_t len = off[m + 1] - off[m];
    if (kind[m] == 0 && len < 1) throw std::invalid_argument("material_index: constant needs one value");
    if (kind[m] == 1 && len < 2) throw std::invalid_argument("material_index: a table needs two points");
    if (kind[m] == 2 && len < 3) throw std::invalid_argument("material_index: Drude needs 3 parameters");
    if (kind[m] < 0 || kind[m] > 2) throw std::invalid_argument("material_index: kind must be 0, 1 or 2");
  }
  for (int64_t q = 0; q < nq; ++q)
    if (qm[q] < 0 || qm[q] >= nmat) throw std::invalid_argument("material_index: material index out of range");
  int64_t *dk = c->buf<int64_t>("mi_kind", (size_t)nmat);
  int64_t *doff = c->buf<int64_t>("mi_off", (size_t)nmat + 1);
  double *dxs = c->buf<double>("mi_xs", (size_t)std::max<int64_t>(npts, 1));
  zd *dys = c->buf<zd>("mi_ys", (size_t)std::max<int64_t>(npts, 1));
  int64_t *dqm = c->buf<int64_t>("mi_qm", (size_t)std::max<int64_t>(nq, 1));
  double *dqw = c->buf<double>("mi_qw", (size_t)std::max<int64_t>(nq, 1));
  zd *dout = c->buf<zd>("mi_out", (size_t)std::max<int64_t>(nq, 1));
  EMU_HIP(hipMemcpyAsync(dk, kind, sizeof(int64_t) * nmat, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(doff, off, sizeof(int64_t) * (nmat + 1), hipMemcpyHostToDevice, s));
  if (npts > 0) {
    EMU_HIP(hipMemcpyAsync(dxs, xs, sizeof(double) * npts, hipMemcpyHostToDevice, s));
    EMU_HIP(hipMemcpyAsync(dys, ys, sizeof(zd) * npts, hipMemcpyHostToDevice, s));
  }
  if (nq > 0) {
    EMU_HIP(hipMemcpyAsync(dqm, qm, sizeof(int64_t) * nq, hipMemcpyHostToDevice, s));
    EMU_HIP(hipMemcpyAsync(dqw, qw, sizeof(double) * nq, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_material_index, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, nq, dk, doff, dxs,
                       dys, dqm, dqw, dout);
    EMU_LAUNCH_CHECK();
    EMU_HIP(hipMemcpyAsync(out, dout, sizeof(zd) * nq, hipMemcpyDeviceToHost, s));
  }
  EMU_HIP(hipStreamSynchronize(s));
}

// Analytic plane-wave modes of homogeneous films (reference mode_calcs.py:
// 110-186 Anallo.calc_modes / calc_kz, 195-212 Anallo.k / Anallo.Z) for nb (film, light) instances, one
// thread per (instance, retained order): k = 2 pi n / wl_norm, k_z =
// sqrt(k^2 - alpha^2 - beta^2) (principal branch) of order order[b][i],
// stored at position i (the air reference's order) of both polarisation
// blocks; Z = (1/n) k / k_z (TE block), (1/n) k_z / k (TM block).  An order
// with k_z exactly 0 (grazing) gets Z = 0 in place of the infinite
// impedance; the caller finds it in k_z, moves k_parallel and evaluates
// again, as Anallo does.
__global__ void k_film_modes(int64_t nb, int64_t npw, const zd *__restrict__ n, const double *__restrict__ wl_norm,
                             const double *__restrict__ alpha, const double *__restrict__ beta,
                             const int64_t *__restrict__ order, zd *__restrict__ kz, zd *__restrict__ Z) {
  const int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nb * npw) return;
  const int64_t b = q / npw, i = q % npw;
  const int64_t o = order[q];
  const zd nb_ = n[b];
  const double w = wl_norm[b], tp = 2.0 * 3.141592653589793;
  const zd k = zmk(tp * nb_.x / w, tp * nb_.y / w);
  const double a = alpha[b * npw + o], be = beta[b * npw + o];
  const zd k2 = zmul(k, k);
  const zd kzv = zsqrt_principal(zmk(k2.x - a * a - be * be, k2.y));
  zd *kb = kz + b * 2 * npw;
  zd *Zb = Z + b * 2 * npw;
  kb[i] = kzv;
  kb[npw + i] = kzv;
  if (kzv.x == 0.0 && kzv.y == 0.0) {
    Zb[i] = Zb[npw + i] = zmk(0.0, 0.0);
    return;
  }
  const zd inv_n = zdiv(zmk(1.0, 0.0), nb_);
  const zd ratio = zdiv(k, kzv);
  Zb[i] = zmul(inv_n, ratio);
  Zb[npw + i] = zdiv(inv_n, ratio);
}

void film_modes(Ctx *c, int64_t nb, int64_t npw, const double *n, const double *wl_norm, const double *alpha,
                const double *beta, const int64_t *order, double *kz, double *Z) {
  EMU_HIP(hipSetDevice(c->device));
  for (int64_t b = 0; b < nb; ++b) {
    if (!(wl_norm[b] > 0.0)) throw std::invalid_argument("film_modes: wl_norm must be positive");
    if (n[2 * b] == 0.0 && n[2 * b + 1] == 0.0) throw std::invalid_argument("film_modes: zero refractive index");
    std::vector<char> seen(npw, 0);
    for (int64_t i = 0; i < npw; ++i) {
      const int64_t o = order[b * npw + i];
      if (o < 0 || o >= npw || seen[o]) throw std::invalid_argument("film_modes: order must be a permutation");
      seen[o] = 1;
    }
  }
  hipStream_t s = c->stream;
  const size_t nq = (size_t)(nb * npw);
  zd *dn = c->buf<zd>("fm_n", (size_t)nb);
  double *dw = c->buf<double>("fm_wl", (size_t)nb);
  double *da = c->buf<double>("fm_alpha", nq);
  double *dbeta = c->buf<double>("fm_beta", nq);
  int64_t *dord = c->buf<int64_t>("fm_order", nq);
  zd *dkz = c->buf<zd>("fm_kz", 2 * nq);
  zd *dZ = c->buf<zd>("fm_Z", 2 * nq);
  EMU_HIP(hipMemcpyAsync(dn, n, sizeof(zd) * nb, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dw, wl_norm, sizeof(double) * nb, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(da, alpha, sizeof(double) * nq, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dbeta, beta, sizeof(double) * nq, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dord, order, sizeof(int64_t) * nq, hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(k_film_modes, dim3((unsigned)((nq + 255) / 256)), dim3(256), 0, s, nb, npw, dn, dw, da, dbeta,
                     dord, dkz, dZ);
  EMU_LAUNCH_CHECK();
  EMU_HIP(hipMemcpyAsync(kz, dkz, sizeof(zd) * 2 * nq, hipMemcpyDeviceToHost, s));
  EMU_HIP(hipMemcpyAsync(Z, dZ, sizeof(zd) * 2 * nq, hipMemcpyDeviceToHost, s));
  EMU_HIP(hipStreamSynchronize(s));
}

}  // namespace emu
