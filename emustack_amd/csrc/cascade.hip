// Redheffer star-product cascade of a layer stack with energy fluxes, batched
// over wavelength points, on the GPU.  Replaces the numpy of
// backend/mode_calcs.py:451-515 (r_t_mat_anallo, r_t_mat_tf_ns) and
// backend/stack.py:162-427 (Stack.calc_scat, shears = None) for stacks whose
// top and bottom layers are thin films.
//
// Every "solve(M, R)" of the reference is a batched dense LU + solve: a
// dense n x n matrix is one front of the multifrontal kernels (factorisation,
// inverted-factor form, row-tiled solves), one instance per wavelength
// point; every product is the batched GEMM.  Column-major storage; point b of
// an (r x c) quantity at offset b * r * c.
#include <cmath>
#include <complex>
#include <map>
#include <string>
#include <stdexcept>

#include "emu_host.h"
#include "engine.h"
#include "engine_api.h"

namespace emu {

typedef std::complex<double> cplx;

// ------------------------------------------------------ dense LU handles
static DevMesh *make_dense(Ctx *c, int n) {
  DevMesh *dm = new DevMesh();
  try {
    Mesh &m = dm->host;
    m.dim = 0;
    m.path = "<dense>";
    m.neq = n;
    m.nnz = (int64_t)n * n;
    m.row_ptr.resize(n + 1);
    m.col_ind.resize((size_t)n * n);
    for (int i = 0; i <= n; ++i) m.row_ptr[i] = (int64_t)i * n;
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) m.col_ind[(size_t)i * n + j] = j;
    Symbolic &sy = m.sym;
    sy.n = n;
    sy.perm.resize(n);
    sy.iperm.resize(n);
    for (int i = 0; i < n; ++i) sy.perm[i] = sy.iperm[i] = i;
    FrontInfo F;
    F.s = n;
    F.u = 0;
    F.f = n;
    F.b = 0;
    sy.fronts.push_back(F);
    // the values are handed over column-major: entry (i, j) at i + j n, which
    // is also its place in the single front
    sy.nnz2front.resize((size_t)n * n);
    for (int64_t k = 0; k < (int64_t)n * n; ++k) sy.nnz2front[k] = k;
    LevelPlan L;
    L.fronts.push_back(0);
    L.max_s = n;
    L.max_f = n;
    sy.levels.push_back(L);
    sy.pool = (int64_t)n * n;
    sy.npiv = n;
    sy.nw = 0;
    sy.max_f = n;
    sy.factor_entries = (int64_t)n * n;
    sy.flops = 8.0 * n * (double)n * n / 3.0;
    devmesh_build(*dm, c->stream);
  } catch (...) {
    delete dm;
    throw;
  }
  return dm;
}

static DevMesh &dense(Ctx *c, int n) {
  auto it = c->dense.find(n);
  if (it != c->dense.end()) return *it->second;
  DevMesh *dm = make_dense(c, n);
  c->dense[n] = dm;
  return *dm;
}

// ----------------------------------------------------- elementwise kernels
// Y = alpha diag(dl) X diag(dr) + beta I   (rows x cols per point; dl / dr may be null)
__global__ void k_affine(int rows, int cols, const zd *__restrict__ X, int64_t sX, const zd *__restrict__ dl,
                         int64_t sl, const zd *__restrict__ dr, int64_t sr, zd alpha, zd beta, zd *__restrict__ Y,
                         int64_t sY) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= (int64_t)rows * cols) return;
  const int i = (int)(idx % rows), j = (int)(idx / rows);
  zd v = X ? X[b * sX + idx] : zmk(0, 0);
  if (dl) v = zmul(dl[b * sl + i], v);
  if (dr) v = zmul(v, dr[b * sr + j]);
  v = zmul(alpha, v);
  if (i == j) v = zadd(v, beta);
  Y[b * sY + idx] = v;
}

static void affine(int nb, int rows, int cols, const zd *X, int64_t sX, const zd *dl, int64_t sl, const zd *dr,
                   int64_t sr, zd alpha, zd beta, zd *Y, int64_t sY, hipStream_t s) {
  const int64_t tot = (int64_t)rows * cols;
  hipLaunchKernelGGL(k_affine, dim3((unsigned)((tot + 255) / 256), nb), dim3(256), 0, s, rows, cols, X, sX, dl, sl,
                     dr, sr, alpha, beta, Y, sY);
  EMU_LAUNCH_CHECK();
}

// principal square root with the C99 csqrt conventions (signed zeros)
__device__ __forceinline__ zd zsqrt(zd z) {
  if (z.x == 0.0 && z.y == 0.0) return zmk(0.0, z.y);
  const double t = sqrt(0.5 * (fabs(z.x) + hypot(z.x, z.y)));
  if (z.x >= 0.0) return zmk(t, z.y / (2.0 * t));
  return zmk(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// film interface against air (r_t_mat_anallo): diagonal R12 / T12 as dense
// matrices; R21 = -R12, T21 = T12.  Z1 = air, Z2 = film, per point.
__global__ void k_film_rt(int d, const zd *__restrict__ Z1, const zd *__restrict__ Z2, zd *__restrict__ R12,
                          zd *__restrict__ T12, zd *__restrict__ R21, zd *__restrict__ T21) {
  const int64_t idx = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (idx >= (int64_t)d * d) return;
  const int i = (int)(idx % d), j = (int)(idx / d);
  zd r = zmk(0, 0), t = zmk(0, 0);
  if (i == j) {
    const zd z1 = Z1[(int64_t)b * d + i], z2 = Z2[(int64_t)b * d + i];
    const zd sum = zadd(z2, z1);
    r = zdiv(zsub(z2, z1), sum);
    t = zdiv(zscale(zmul(zsqrt(z2), zsqrt(z1)), 2.0), sum);
  }
  const int64_t o = (int64_t)b * d * d + idx;
  R12[o] = r;
  T12[o] = t;
  R21[o] = zmk(-r.x, -r.y);
  T21[o] = t;
}

// propagation phases exp(i kz h) per point; h = hts[b * L + layer], the
// layer's height (over the period) in stack b
__global__ void k_prop(int d, const zd *__restrict__ kz, const double *__restrict__ hts, int L, int layer,
                       zd *__restrict__ P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int b = blockIdx.y;
  if (i >= d) return;
  const double h = hts[(int64_t)b * L + layer];
  const zd k = kz[(int64_t)b * d + i];
  // exp(i (kr + i ki) h) = exp(-ki h) (cos kr h + i sin kr h)
  const double m = exp(-k.y * h);
  double sn, cs;
  sincos(k.x * h, &sn, &cs);
  P[(int64_t)b * d + i] = zmk(m * cs, m * sn);
}

// flux forms per point: out[b] = f^H U f (f = [f_minus; f_plus], U of
// stack.py:296-314 with num_prop_air propagating orders) when U_form, else the
// propagating power ||f[0:np]||^2 + ||f[P:P+np]||^2 of f_minus with np = nprop[b]
__global__ void k_flux(int nb, int P, const zd *__restrict__ fm, const zd *__restrict__ fp, int64_t sf,
                       const int32_t *__restrict__ nprop, int U_form, zd *__restrict__ out) {
  const int b = blockIdx.x;
  const int tid = threadIdx.x;
  __shared__ double sr[256], si[256];
  const zd *m = fm + (int64_t)b * sf;
  const zd *p = fp ? fp + (int64_t)b * sf : nullptr;
  const int np = nprop[b];
  double ar = 0, ai = 0;
  for (int q = tid; q < 2 * P; q += 256) {
    const int i = q % P;  // order index within the polarisation
    const zd a = m[q];
    if (!U_form) {
      if (i < np) ar += a.x * a.x + a.y * a.y;
      continue;
    }
    const zd c = p[q];
    const bool te = q < P;
    if (i < np) {
      ar += (a.x * a.x + a.y * a.y) - (c.x * c.x + c.y * c.y);
    } else {
      // conj(a) (s1 c) + conj(c) (s2 a) with (s1, s2) = (-i, +i) for TE and (+i, -i) for TM
      const zd s1 = te ? zmk(0, -1) : zmk(0, 1), s2 = te ? zmk(0, 1) : zmk(0, -1);
      const zd t1 = zmul(zconj(a), zmul(s1, c)), t2 = zmul(zconj(c), zmul(s2, a));
      ar += t1.x + t2.x;
      ai += t1.y + t2.y;
    }
  }
  sr[tid] = ar;
  si[tid] = ai;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      sr[tid] += sr[tid + st];
      si[tid] += si[tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) out[b] = zmk(sr[0], si[0]);
}

// ------------------------------------------------------------- the cascade
struct Lay {
  int kind, dim;        // 0 film (dim 2P), 1 structured (dim = modes)
  zd *R12, *T12, *R21, *T21, *P;
};

void star_cascade(Ctx *c, const CascadeArgs &a) {
  EMU_HIP(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  const int nb = (int)a.nbatch, L = (int)a.nlayers, P = (int)a.neq_pw, d = 2 * P;
  if (nb < 1 || L < 2 || P < 1) throw std::invalid_argument("star_cascade: bad sizes");
  if (a.kind[0] != 0 || a.kind[L - 1] != 0)
    throw std::invalid_argument("star_cascade: top and bottom layers must be thin films");
  auto up = [&](const std::string &name, const double *src, size_t count) {
    zd *p = c->buf<zd>(name, count);
    EMU_HIP(hipMemcpyAsync(p, src, sizeof(zd) * count, hipMemcpyHostToDevice, s));
    return p;
  };
  const zd one = zmk(1, 0), zero = zmk(0, 0);
  zd *Zair = up("cs_zair", a.Zair, (size_t)nb * d);
  double *hts = c->buf<double>("cs_heights", (size_t)nb * L);
  EMU_HIP(hipMemcpyAsync(hts, a.heights, sizeof(double) * nb * L, hipMemcpyHostToDevice, s));
  // 1. interface matrices of every layer against air
  std::vector<Lay> lay(L);
  for (int i = 0; i < L; ++i) {
    Lay &l = lay[i];
    l.kind = (int)a.kind[i];
    l.dim = l.kind == 0 ? d : (int)a.nmodes[i];
    const int m = l.dim;
    const std::string t = "cs_l" + std::to_string(i) + "_";
    l.R12 = c->buf<zd>(t + "r12", (size_t)nb * d * d);
    l.T12 = c->buf<zd>(t + "t12", (size_t)nb * m * d);
    l.R21 = c->buf<zd>(t + "r21", (size_t)nb * m * m);
    l.T21 = c->buf<zd>(t + "t21", (size_t)nb * d * m);
    l.P = c->buf<zd>(t + "p", (size_t)nb * m);
    zd *kz = up(t + "kz", a.kz[i], (size_t)nb * m);
    if (i > 0 && i < L - 1) {
      hipLaunchKernelGGL(k_prop, dim3((m + 255) / 256, nb), dim3(256), 0, s, m, kz, hts, L, i, l.P);
      EMU_LAUNCH_CHECK();
    }
    if (l.kind == 0) {
      zd *Z = up(t + "z", a.Z[i], (size_t)nb * d);
      hipLaunchKernelGGL(k_film_rt, dim3((unsigned)(((int64_t)d * d + 255) / 256), nb), dim3(256), 0, s, d, Zair,
                         Z, l.R12, l.T12, l.R21, l.T21);
      EMU_LAUNCH_CHECK();
      continue;
    }
    // r_t_mat_tf_ns: A = diag(zs) J, B = J^dagger diag(zs), D = I + B A
    zd *J = up(t + "J", a.J[i], (size_t)nb * d * m);
    zd *Jd = up(t + "Jd", a.Jd[i], (size_t)nb * m * d);
    zd *zs = c->buf<zd>("cs_zs", (size_t)nb * d);
    {
      std::vector<cplx> zh((size_t)nb * d);
      for (size_t q = 0; q < zh.size(); ++q) zh[q] = std::sqrt(1.0 / cplx(a.Zair[2 * q], a.Zair[2 * q + 1]));
      EMU_HIP(hipMemcpyAsync(zs, zh.data(), sizeof(zd) * zh.size(), hipMemcpyHostToDevice, s));
      EMU_HIP(hipStreamSynchronize(s));
    }
    zd *A = c->buf<zd>("cs_A", (size_t)nb * d * m);
    zd *B = c->buf<zd>("cs_B", (size_t)nb * m * d);
    zd *D = c->buf<zd>("cs_D", (size_t)nb * m * m);
    zd *Dinv = c->buf<zd>("cs_Dinv", (size_t)nb * m * m);
    affine(nb, d, m, J, (int64_t)d * m, zs, d, nullptr, 0, one, zero, A, (int64_t)d * m, s);
    affine(nb, m, d, Jd, (int64_t)m * d, nullptr, 0, zs, d, one, zero, B, (int64_t)m * d, s);
    launch_gemm_nn(nb, m, m, d, one, B, m, (int64_t)m * d, A, d, (int64_t)d * m, zero, D, m, (int64_t)m * m, s);
    affine(nb, m, m, D, (int64_t)m * m, nullptr, 0, nullptr, 0, one, one, D, (int64_t)m * m, s);
    solve_dense(c, nb, m, D, nullptr, m, Dinv);  // Dinv = D^-1
    // T12 = 2 Dinv B ; R12 = -I + A T12 ; R21 = 2 Dinv - I ; T21 = 2 A Dinv
    launch_gemm_nn(nb, m, d, m, zmk(2, 0), Dinv, m, (int64_t)m * m, B, m, (int64_t)m * d, zero, l.T12, m,
                   (int64_t)m * d, s);
    affine(nb, d, d, nullptr, 0, nullptr, 0, nullptr, 0, zero, zmk(-1, 0), l.R12, (int64_t)d * d, s);
    launch_gemm_nn(nb, d, d, m, one, A, d, (int64_t)d * m, l.T12, m, (int64_t)m * d, one, l.R12, d, (int64_t)d * d,
                   s);
    affine(nb, m, m, Dinv, (int64_t)m * m, nullptr, 0, nullptr, 0, zmk(2, 0), zmk(-1, 0), l.R21, (int64_t)m * m, s);
    launch_gemm_nn(nb, d, m, m, zmk(2, 0), A, d, (int64_t)d * m, Dinv, m, (int64_t)m * m, zero, l.T21, d,
                   (int64_t)d * m, s);
  }
  // 2. net matrices bottom-up through the infinitesimal air gaps (stack.py:204-278)
  int tr = d, tc = d;  // tnet shape
  zd *tnet = c->buf<zd>("cs_tnet0", (size_t)nb * d * d);
  EMU_HIP(hipMemcpyAsync(tnet, lay[0].T12, sizeof(zd) * nb * d * d, hipMemcpyDeviceToDevice, s));
  std::vector<zd *> rnet_list, it21_list, it12_list;
  std::vector<int> rdim;
  zd *rnet = c->buf<zd>("cs_rnet0", (size_t)nb * d * d);
  EMU_HIP(hipMemcpyAsync(rnet, lay[0].R12, sizeof(zd) * nb * d * d, hipMemcpyDeviceToDevice, s));
  rnet_list.push_back(rnet);
  rdim.push_back(d);
  int rd = d;  // rnet dimension
  zd *tnet0 = tnet;
  int step = 0;
  auto fresh = [&](const char *what, size_t count) {
    return c->buf<zd>(std::string("cs_") + what + std::to_string(step), count);
  };
  auto air_gap = [&](int i) {
    const Lay &l = lay[i];
    const int m = l.dim;
    ++step;
    // it21 = (I - r12 rnet)^-1 t21  (d x m)
    zd *Mx = c->buf<zd>("cs_M", (size_t)nb * d * d);
    launch_gemm_nn(nb, d, d, d, zmk(-1, 0), l.R12, d, (int64_t)d * d, rnet, d, (int64_t)d * d, zero, Mx, d,
                   (int64_t)d * d, s);
    affine(nb, d, d, Mx, (int64_t)d * d, nullptr, 0, nullptr, 0, one, one, Mx, (int64_t)d * d, s);
    zd *it21 = fresh("it21_", (size_t)nb * d * m);
    solve_dense(c, nb, d, Mx, l.T21, m, it21);
    // tnet = tnet it21 ; rnet = r21 + t12 rnet it21
    zd *tn = fresh("tnet", (size_t)nb * tr * m);
    launch_gemm_nn(nb, tr, m, d, one, tnet, tr, (int64_t)tr * tc, it21, d, (int64_t)d * m, zero, tn, tr,
                   (int64_t)tr * m, s);
    tnet = tn;
    tc = m;
    zd *tmp = c->buf<zd>("cs_tmp", (size_t)nb * m * d);
    launch_gemm_nn(nb, m, d, d, one, l.T12, m, (int64_t)m * d, rnet, d, (int64_t)d * d, zero, tmp, m,
                   (int64_t)m * d, s);
    zd *rn = fresh("rnet", (size_t)nb * m * m);
    EMU_HIP(hipMemcpyAsync(rn, l.R21, sizeof(zd) * nb * m * m, hipMemcpyDeviceToDevice, s));
    launch_gemm_nn(nb, m, m, d, one, tmp, m, (int64_t)m * d, it21, d, (int64_t)d * m, one, rn, m, (int64_t)m * m,
                   s);
    rnet = rn;
    rd = m;
    rnet_list.push_back(rnet);
    rdim.push_back(m);
    it21_list.push_back(it21);
  };
  for (int i = 1; i < L - 1; ++i) {
    air_gap(i);
    const Lay &l = lay[i];
    const int m = l.dim;
    ++step;
    // it12 = (I - r21 P rnet P)^-1 t12  (m x d)
    zd *PrP = c->buf<zd>("cs_PrP", (size_t)nb * m * m);
    affine(nb, m, m, rnet, (int64_t)m * m, l.P, m, l.P, m, one, zero, PrP, (int64_t)m * m, s);
    zd *Mx = c->buf<zd>("cs_Mm", (size_t)nb * m * m);
    launch_gemm_nn(nb, m, m, m, zmk(-1, 0), l.R21, m, (int64_t)m * m, PrP, m, (int64_t)m * m, zero, Mx, m,
                   (int64_t)m * m, s);
    affine(nb, m, m, Mx, (int64_t)m * m, nullptr, 0, nullptr, 0, one, one, Mx, (int64_t)m * m, s);
    zd *it12 = fresh("it12_", (size_t)nb * m * d);
    solve_dense(c, nb, m, Mx, l.T12, d, it12);
    // P it12, tnet = tnet P it12 ; rnet = r12 + t21 (P rnet P) it12
    zd *Pit = c->buf<zd>("cs_Pit", (size_t)nb * m * d);
    affine(nb, m, d, it12, (int64_t)m * d, l.P, m, nullptr, 0, one, zero, Pit, (int64_t)m * d, s);
    zd *tn = fresh("tnet", (size_t)nb * tr * d);
    launch_gemm_nn(nb, tr, d, m, one, tnet, tr, (int64_t)tr * tc, Pit, m, (int64_t)m * d, zero, tn, tr,
                   (int64_t)tr * d, s);
    tnet = tn;
    tc = d;
    zd *tmp = c->buf<zd>("cs_tmp2", (size_t)nb * d * m);
    launch_gemm_nn(nb, d, m, m, one, l.T21, d, (int64_t)d * m, PrP, m, (int64_t)m * m, zero, tmp, d,
                   (int64_t)d * m, s);
    zd *rn = fresh("rnet", (size_t)nb * d * d);
    EMU_HIP(hipMemcpyAsync(rn, l.R12, sizeof(zd) * nb * d * d, hipMemcpyDeviceToDevice, s));
    launch_gemm_nn(nb, d, d, m, one, tmp, d, (int64_t)d * m, it12, m, (int64_t)m * d, one, rn, d, (int64_t)d * d,
                   s);
    rnet = rn;
    rd = d;
    rnet_list.push_back(rnet);
    rdim.push_back(d);
    it12_list.push_back(it12);
  }
  air_gap(L - 1);
  (void)rd;
  if (a.R_net) EMU_HIP(hipMemcpyAsync(a.R_net, rnet, sizeof(zd) * nb * d * d, hipMemcpyDeviceToHost, s));
  if (a.T_net) EMU_HIP(hipMemcpyAsync(a.T_net, tnet, sizeof(zd) * nb * d * d, hipMemcpyDeviceToHost, s));
  // 3. fluxes (stack.py:285-427)
  zd *dm_ = up("cs_dminus", a.d_minus, (size_t)nb * d);
  std::vector<int32_t> npa(nb), npi(nb), npo(nb);
  for (int b = 0; b < nb; ++b) {
    npa[b] = (int32_t)a.nprop[3 * b];
    npi[b] = (int32_t)a.nprop[3 * b + 1];
    npo[b] = (int32_t)a.nprop[3 * b + 2];
  }
  int32_t *dnpa = c->buf<int32_t>("cs_npa", nb), *dnpi = c->buf<int32_t>("cs_npi", nb),
          *dnpo = c->buf<int32_t>("cs_npo", nb);
  EMU_HIP(hipMemcpyAsync(dnpa, npa.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dnpi, npi.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
  EMU_HIP(hipMemcpyAsync(dnpo, npo.data(), sizeof(int32_t) * nb, hipMemcpyHostToDevice, s));
  const int nflux = L;  // down[0..L-1]
  zd *dflux = c->buf<zd>("cs_down", (size_t)nflux * nb);
  zd *uflux = c->buf<zd>("cs_up", nb);
  auto matvec = [&](const zd *M, int r, int k, const zd *v, zd *out) {
    launch_gemm_nn(nb, r, 1, k, one, M, r, (int64_t)r * k, v, k, k, zero, out, r, r, s);
  };
  zd *vplus = c->buf<zd>("cs_vplus", (size_t)nb * d);
  hipLaunchKernelGGL(k_flux, dim3(nb), dim3(256), 0, s, nb, P, dm_, nullptr, (int64_t)d, dnpi, 0, dflux);
  matvec(rnet_list.back(), d, d, dm_, vplus);
  hipLaunchKernelGGL(k_flux, dim3(nb), dim3(256), 0, s, nb, P, vplus, nullptr, (int64_t)d, dnpi, 0, uflux);
  EMU_LAUNCH_CHECK();
  int maxm = d;
  for (auto &l : lay) maxm = std::max(maxm, l.dim);
  zd *f1m = c->buf<zd>("cs_f1m", (size_t)nb * maxm), *f1p = c->buf<zd>("cs_f1p", (size_t)nb * maxm);
  zd *f2m = c->buf<zd>("cs_f2m", (size_t)nb * maxm), *tv = c->buf<zd>("cs_tv", (size_t)nb * maxm);
  // f1_minus = it21_list[-1] d_minus : the last air gap maps into the top film (dim d)
  matvec(it21_list.back(), d, d, dm_, f1m);
  const int nI = L - 2;
  for (int i = 0; i < nI; ++i) {
    // f1_plus = rnet_list[-2i-2] f1_minus ; down = f^H U f
    const int ri = (int)rnet_list.size() - 2 * i - 2;
    matvec(rnet_list[ri], rdim[ri], rdim[ri], f1m, f1p);
    hipLaunchKernelGGL(k_flux, dim3(nb), dim3(256), 0, s, nb, P, f1m, f1p, (int64_t)d, dnpa, 1,
                       dflux + (int64_t)(i + 1) * nb);
    EMU_LAUNCH_CHECK();
    // f2_minus = it12_list[-i-1] f1_minus (dim m of layer L-2-i)
    const int li = L - 2 - i;
    const int m = lay[li].dim;
    matvec(it12_list[it12_list.size() - 1 - i], m, d, f1m, f2m);
    // f1_minus = it21_list[-i-2] P f2_minus
    affine(nb, m, 1, f2m, m, lay[li].P, m, nullptr, 0, one, zero, tv, m, s);
    matvec(it21_list[it21_list.size() - 2 - i], d, m, tv, f1m);
  }
  // f2_minus = tnet_list[0] f1_minus ; bottom film flux
  matvec(tnet0, d, d, f1m, f2m);
  hipLaunchKernelGGL(k_flux, dim3(nb), dim3(256), 0, s, nb, P, f2m, nullptr, (int64_t)d, dnpo, 0,
                     dflux + (int64_t)(L - 1) * nb);
  EMU_LAUNCH_CHECK();
  std::vector<zd> down((size_t)nflux * nb), upv(nb);
  EMU_HIP(hipMemcpyAsync(down.data(), dflux, sizeof(zd) * down.size(), hipMemcpyDeviceToHost, s));
  EMU_HIP(hipMemcpyAsync(upv.data(), uflux, sizeof(zd) * nb, hipMemcpyDeviceToHost, s));
  EMU_HIP(hipStreamSynchronize(s));
  // tally (stack.py:392-427, bottom film with propagating orders)
  for (int b = 0; b < nb; ++b) {
    auto dn = [&](int i) { return cplx(down[(size_t)i * nb + b].x, down[(size_t)i * nb + b].y); };
    const cplx u0(upv[b].x, upv[b].y);
    double *A_ = a.a_list + (size_t)b * (L - 1), *T_ = a.t_list + (size_t)b * (L - 1);
    if (npo[b] == 0) {
      for (int i = 0; i < L - 1; ++i) A_[i] = T_[i] = 0;
      a.r_list[b] = 0;
      continue;
    }
    for (int i = 1; i < L - 1; ++i) A_[i - 1] = std::abs(std::abs(dn(i)) - std::abs(dn(i + 1)));
    A_[L - 2] = std::abs(dn(0) - dn(L - 1) - u0);
    a.r_list[b] = std::abs(u0 / dn(0));
    for (int i = 0; i < L - 2; ++i) T_[i] = std::abs(std::abs(dn(i + 2)) / std::abs(dn(i)));
    T_[L - 2] = std::abs(dn(L - 1) / dn(0));
  }
}

// batched dense solve X = M^-1 R (R = identity when null); n x n, R n x r
void solve_dense(Ctx *c, int nb, int n, const zd *M, const zd *R, int r, zd *X) {
  hipStream_t s = c->stream;
  DevMesh &dm = dense(c, n);
  const Symbolic &sy = dm.host.sym;
  zd *pool = c->buf<zd>("ds_pool", (size_t)nb * sy.pool);
  int32_t *piv = c->buf<int32_t>("ds_piv", (size_t)nb * n);
  int32_t *pfl = c->buf<int32_t>("ds_pfl", (size_t)nb * n);
  int32_t *sig = c->buf<int32_t>("ds_sig", (size_t)nb * n);
  int32_t *gx = c->buf<int32_t>("ds_gx", (size_t)nb * n * 4);
  zd *inv = c->buf<zd>("ds_inv", (size_t)nb * dm.isz);
  lu_factor(dm, nb, M, pool, piv, pfl, sig, gx, inv, s);
  const zd *rhs = R;
  if (!R) {
    zd *I = c->buf<zd>("ds_eye", (size_t)nb * n * n);
    affine(nb, n, n, nullptr, 0, nullptr, 0, nullptr, 0, zmk(0, 0), zmk(1, 0), I, (int64_t)n * n, s);
    rhs = I;
    r = n;
  }
  zd *y = c->buf<zd>("ds_y", (size_t)nb * n * r);
  zd *w = c->buf<zd>("ds_w", (size_t)nb * r * solve_ws(dm));
  lu_solve(dm, nb, r, inv, gx, rhs, y, X, w, s);
}

}  // namespace emu
