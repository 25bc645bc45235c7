// extern "C" entry points of libemu_mi355x.so (declared in
// include/emustack_mi355x.h).  Converts C++ exceptions to status codes and a
// thread-local message.
#include <condition_variable>
#include <cstring>
#include <deque>
#include <exception>
#include <functional>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/emustack_mi355x.h"
#include "common.h"
#include "engine_api.h"

// A context runs one engine call at a time (run_mu).  Submitted jobs queue on
// the context's worker thread, started on the first submission and joined
// (after draining the queue) by emu_ctx_destroy.
struct emu_ctx {
  emu::Ctx *c = nullptr;
  std::mutex run_mu;
  std::mutex q_mu;
  std::condition_variable q_cv;
  std::deque<std::function<void()>> q;
  bool stop = false;
  std::thread worker;
};

// A submitted batch: owned copies of the inputs, its status and message
struct emu_job {
  std::mutex mu;
  std::condition_variable cv;
  bool done = false;
  int status = 0;
  std::string err;
  std::vector<double> lambda, n_eff, bloch, shift, opts;
};

static void worker_loop(emu_ctx *ctx) {
  for (;;) {
    std::function<void()> f;
    {
      std::unique_lock<std::mutex> lk(ctx->q_mu);
      ctx->q_cv.wait(lk, [&] { return ctx->stop || !ctx->q.empty(); });
      if (ctx->q.empty()) return;  // stop requested and nothing left
      f = std::move(ctx->q.front());
      ctx->q.pop_front();
    }
    f();
  }
}
struct emu_mesh {
  emu::DevMesh *m;
};

static thread_local std::string g_err = "no error";

static int fail(int code, const std::string &msg) {
  g_err = msg;
  return code;
}

#define EMU_GUARD(body)                                                       \
  try {                                                                       \
    body;                                                                     \
    return EMU_OK;                                                            \
  } catch (const emu::NoConvergence &e) {                                     \
    return fail(EMU_E_NOCONV, e.what());                                      \
  } catch (const emu::Homogeneous &e) {                                       \
    return fail(EMU_E_HOMOGENEOUS, e.what());                                 \
  } catch (const emu::LapackMissing &e) {                                     \
    return fail(EMU_E_LAPACK, e.what());                                      \
  } catch (const emu::HipError &e) {                                          \
    std::string w = e.what();                                                 \
    return fail(w.find("hipMalloc") != std::string::npos ? EMU_E_OOM : EMU_E_HIP, w); \
  } catch (const std::invalid_argument &e) {                                  \
    return fail(EMU_E_ARG, e.what());                                         \
  } catch (const std::bad_alloc &) {                                          \
    return fail(EMU_E_OOM, "host allocation failed");                         \
  } catch (const std::exception &e) {                                         \
    std::string w = e.what();                                                 \
    bool mesh = w.find("periodic") != std::string::npos || w.find("mesh") != std::string::npos || \
                w.find("list_") != std::string::npos || w.find("asmbly") != std::string::npos ||  \
                w.find("basis_ls") != std::string::npos || w.find("geometry") != std::string::npos; \
    return fail(mesh ? EMU_E_MESH : EMU_E_LAPACK, w);                         \
  }

extern "C" {

const char *emu_last_error(void) { return g_err.c_str(); }

int emu_ctx_create(int device, emu_ctx **out) {
  if (!out) return fail(EMU_E_ARG, "emu_ctx_create: out is NULL");
  EMU_GUARD({
    emu_ctx *c = new emu_ctx;
    try {
      c->c = emu::ctx_create(device);
    } catch (...) {
      delete c;
      throw;
    }
    *out = c;
  })
}

void emu_ctx_destroy(emu_ctx *ctx) {
  if (!ctx) return;
  {
    std::lock_guard<std::mutex> lk(ctx->q_mu);
    ctx->stop = true;
  }
  ctx->q_cv.notify_all();
  if (ctx->worker.joinable()) ctx->worker.join();
  emu::ctx_destroy(ctx->c);
  delete ctx;
}

int emu_profile(emu_ctx *ctx, int enable, int every, double *out) {
  if (!ctx) return fail(EMU_E_ARG, "emu_profile: NULL context");
  EMU_GUARD(emu::profile(ctx->c, enable, every, out))
}

int emu_set_lapack(void *zgees, void *ztrsen, void *ztrevc) {
  if (!zgees || !ztrsen || !ztrevc) return fail(EMU_E_ARG, "emu_set_lapack: NULL routine");
  emu::set_lapack(zgees, ztrsen, ztrevc);
  return EMU_OK;
}

int emu_mesh_load(emu_ctx *ctx, const char *path, int leaf_size, emu_mesh **out) {
  if (!ctx || !path || !out) return fail(EMU_E_ARG, "emu_mesh_load: NULL argument");
  EMU_GUARD({
    emu_mesh *m = new emu_mesh;
    try {
      m->m = emu::mesh_load(ctx->c, path, leaf_size);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  })
}

int emu_mesh_analyze(const char *path, int leaf_size, emu_mesh **out) {
  if (!path || !out) return fail(EMU_E_ARG, "emu_mesh_analyze: NULL argument");
  EMU_GUARD({
    emu_mesh *m = new emu_mesh;
    try {
      m->m = emu::mesh_analyze(path, leaf_size);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  })
}

int emu_mesh_symbolic(const emu_mesh *mesh, int32_t *fronts, int64_t *foff, int64_t *uoff, int32_t *uidx,
                      int32_t *pmap, int64_t *nnz2front) {
  if (!mesh) return fail(EMU_E_ARG, "emu_mesh_symbolic: NULL mesh");
  EMU_GUARD(emu::mesh_symbolic(mesh->m, fronts, foff, uoff, uidx, pmap, nnz2front))
}

void emu_mesh_free(emu_mesh *mesh) {
  if (!mesh) return;
  emu::mesh_free(mesh->m);
  delete mesh;
}

int emu_mesh_info(const emu_mesh *mesh, int64_t *info) {
  if (!mesh || !info) return fail(EMU_E_ARG, "emu_mesh_info: NULL argument");
  EMU_GUARD(emu::mesh_info(mesh->m, info))
}

int emu_mesh_arrays(const emu_mesh *mesh, int64_t *row_ptr, int32_t *col_ind, int32_t *ldof, int32_t *perm,
                    double *lat) {
  if (!mesh) return fail(EMU_E_ARG, "emu_mesh_arrays: NULL mesh");
  EMU_GUARD(emu::mesh_arrays(mesh->m, row_ptr, col_ind, ldof, perm, lat))
}

int emu_mesh_geometry(const emu_mesh *mesh, int64_t *table_nod, int64_t *type_el, double *x_arr) {
  if (!mesh) return fail(EMU_E_ARG, "emu_mesh_geometry: NULL mesh");
  EMU_GUARD(emu::mesh_geometry(mesh->m, table_nod, type_el, x_arr))
}

}  // extern "C"

static int modes2d_impl(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda, int64_t nval,
                            int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                            const double *shift, int64_t E_H_field, int64_t itermax, int64_t neq_PW,
                            const double *opts, double *beta1, double *overlap_J, double *overlap_J_dagger,
                            int64_t *stats, double *sol, double *mode_pol) {
  if (!ctx || !mesh || !lambda || !n_eff || !bloch_vec || !shift || !beta1 || !overlap_J || !overlap_J_dagger)
    return fail(EMU_E_ARG, "emu_calc_modes_2d_batch: NULL argument");
  if (nbatch < 1 || nval < 1 || nb_typ_el < 1 || neq_PW < 1 || itermax < 0)
    return fail(EMU_E_ARG, "emu_calc_modes_2d_batch: invalid size");
  if (E_H_field != 1 && E_H_field != 2) return fail(EMU_E_ARG, "E_H_field must be 1 or 2");
  for (int64_t i = 0; i < nbatch; ++i)
    if (!(lambda[i] > 0)) return fail(EMU_E_ARG, "lambda must be positive");
  std::lock_guard<std::mutex> run(ctx->run_mu);
  EMU_GUARD({
    emu::ModesArgs a;
    a.nl = nbatch;
    a.lambda = lambda;
    a.nval = nval;
    a.ordre = ordre_ls;
    a.ntyp = nb_typ_el;
    a.eh = E_H_field;
    a.itermax = itermax;
    a.neq_pw = neq_PW;
    a.n_eff = n_eff;
    a.bloch = bloch_vec;
    a.shift = shift;
    if (opts && opts[0] > 0) {
      a.tol = opts[0];
      a.tol_set = true;
    }
    if (opts && opts[1] != 0) a.precision = (int64_t)opts[1];
    if (a.precision != 0 && a.precision != 1) return fail(EMU_E_ARG, "precision must be 0 (FP64) or 1 (mixed)");
    if (opts && opts[2] > 0) a.block = (int64_t)opts[2];
    if (opts && opts[3] > 0) a.ncv = (int64_t)opts[3];
    if (a.block != 1 && a.block != 2 && a.block != 4 && a.block != 8)
      return fail(EMU_E_ARG, "Krylov block size must be 1, 2, 4 or 8");
    a.beta = beta1;
    a.J = overlap_J;
    a.Jd = overlap_J_dagger;
    a.stats = stats;
    a.sol = sol;
    a.mode_pol = mode_pol;
    emu::calc_modes(ctx->c, *mesh->m, a);
  })
}

extern "C" {

int emu_calc_modes_2d_batch(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda, int64_t nval,
                            int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                            const double *shift, int64_t E_H_field, int64_t itermax, int64_t neq_PW,
                            const double *opts, double *beta1, double *overlap_J, double *overlap_J_dagger,
                            int64_t *stats) {
  return modes2d_impl(ctx, mesh, nbatch, lambda, nval, ordre_ls, nb_typ_el, n_eff, bloch_vec, shift, E_H_field,
                      itermax, neq_PW, opts, beta1, overlap_J, overlap_J_dagger, stats, nullptr, nullptr);
}

int emu_calc_modes_2d(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                      int64_t nb_typ_el, const double *n_eff, const double *bloch_vec, const double *shift,
                      int64_t E_H_field, int64_t itermax, int64_t neq_PW, double *beta1, double *overlap_J,
                      double *overlap_J_dagger) {
  return emu_calc_modes_2d_batch(ctx, mesh, 1, &lambda, nval, ordre_ls, nb_typ_el, n_eff, bloch_vec, shift,
                                 E_H_field, itermax, neq_PW, nullptr, beta1, overlap_J, overlap_J_dagger, nullptr);
}

int emu_calc_modes_2d_fields(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                             int64_t nb_typ_el, const double *n_eff, const double *bloch_vec, const double *shift,
                             int64_t E_H_field, int64_t itermax, int64_t neq_PW, double *beta1, double *overlap_J,
                             double *overlap_J_dagger, double *sol1, double *mode_pol) {
  if (!sol1 && !mode_pol) return fail(EMU_E_ARG, "emu_calc_modes_2d_fields: no field output requested");
  return modes2d_impl(ctx, mesh, 1, &lambda, nval, ordre_ls, nb_typ_el, n_eff, bloch_vec, shift, E_H_field, itermax,
                      neq_PW, nullptr, beta1, overlap_J, overlap_J_dagger, nullptr, sol1, mode_pol);
}

int emu_calc_modes_2d_submit(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda, int64_t nval,
                             int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                             const double *shift, int64_t E_H_field, int64_t itermax, int64_t neq_PW,
                             const double *opts, double *beta1, double *overlap_J, double *overlap_J_dagger,
                             int64_t *stats, emu_job **out) {
  if (!out) return fail(EMU_E_ARG, "emu_calc_modes_2d_submit: out is NULL");
  *out = nullptr;
  if (!ctx || !mesh || !lambda || !n_eff || !bloch_vec || !shift || !beta1 || !overlap_J || !overlap_J_dagger)
    return fail(EMU_E_ARG, "emu_calc_modes_2d_submit: NULL argument");
  if (nbatch < 1 || nb_typ_el < 1) return fail(EMU_E_ARG, "emu_calc_modes_2d_submit: invalid size");
  EMU_GUARD({
    emu_job *job = new emu_job;
    job->lambda.assign(lambda, lambda + nbatch);
    job->n_eff.assign(n_eff, n_eff + 2 * nbatch * nb_typ_el);  // complex per (instance, material)
    job->bloch.assign(bloch_vec, bloch_vec + 2 * nbatch);
    job->shift.assign(shift, shift + 2 * nbatch);
    if (opts) job->opts.assign(opts, opts + 4);
    auto run = [=]() {
      const int st = modes2d_impl(ctx, mesh, nbatch, job->lambda.data(), nval, ordre_ls, nb_typ_el, job->n_eff.data(),
                                  job->bloch.data(), job->shift.data(), E_H_field, itermax, neq_PW,
                                  job->opts.empty() ? nullptr : job->opts.data(), beta1, overlap_J,
                                  overlap_J_dagger, stats, nullptr, nullptr);
      std::string msg = st == EMU_OK ? std::string() : g_err;
      {
        std::lock_guard<std::mutex> lk(job->mu);
        job->status = st;
        job->err = msg;
        job->done = true;
      }
      job->cv.notify_all();
    };
    {
      std::lock_guard<std::mutex> lk(ctx->q_mu);
      if (!ctx->worker.joinable()) ctx->worker = std::thread(worker_loop, ctx);
      ctx->q.push_back(run);
    }
    ctx->q_cv.notify_one();
    *out = job;
  })
}

int emu_job_wait(emu_job *job) {
  if (!job) return fail(EMU_E_ARG, "emu_job_wait: NULL job");
  int st;
  std::string msg;
  {
    std::unique_lock<std::mutex> lk(job->mu);
    job->cv.wait(lk, [&] { return job->done; });
    st = job->status;
    msg = job->err;
  }
  delete job;
  if (st != EMU_OK) return fail(st, msg);
  return EMU_OK;
}

int emu_mesh_create_1d(emu_ctx *ctx, int64_t npt_P2, int64_t nel, const int64_t *table_nod,
                       const int64_t *type_el, const double *x_P2, int leaf_size, emu_mesh **out) {
  if (!ctx || !table_nod || !type_el || !x_P2 || !out) return fail(EMU_E_ARG, "emu_mesh_create_1d: NULL argument");
  *out = nullptr;
  EMU_GUARD({
    emu_mesh *h = new emu_mesh;
    try {
      h->m = emu::mesh_create_1d(ctx->c, npt_P2, nel, table_nod, type_el, x_P2, leaf_size);
    } catch (...) {
      delete h;
      throw;
    }
    *out = h;
  })
}

}  // extern "C"

static int modes1d_impl(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda, int64_t nval,
                            int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                            const double *shift, int64_t itermax, int64_t neq_PW, int64_t neq_PW_2d,
                            int64_t world_1d, const double *opts, double *beta1, double *overlap_J,
                            double *overlap_J_dagger, double *overlap_J_2d, double *overlap_J_dagger_2d,
                            int64_t *stats, double *sol) {
  if (!ctx || !mesh || !lambda || !n_eff || !bloch_vec || !shift || !beta1 || !overlap_J || !overlap_J_dagger)
    return fail(EMU_E_ARG, "emu_calc_modes_1d_batch: NULL argument");
  if (nbatch < 1 || nval < 1 || nb_typ_el < 1 || neq_PW < 1 || itermax < 0 || ordre_ls < 0)
    return fail(EMU_E_ARG, "emu_calc_modes_1d_batch: invalid size");
  if (world_1d != 0 && world_1d != 1) return fail(EMU_E_ARG, "world_1d must be 0 or 1");
  if (world_1d == 0 && (!overlap_J_2d || !overlap_J_dagger_2d || neq_PW_2d < 1))
    return fail(EMU_E_ARG, "world_1d = 0 needs the 2D overlap outputs and neq_PW_2d");
  for (int64_t i = 0; i < nbatch; ++i)
    if (!(lambda[i] > 0)) return fail(EMU_E_ARG, "lambda must be positive");
  std::lock_guard<std::mutex> run(ctx->run_mu);
  EMU_GUARD({
    emu::ModesArgs a;
    a.nl = nbatch;
    a.lambda = lambda;
    a.nval = nval;
    a.ordre = ordre_ls;
    a.ntyp = nb_typ_el;
    a.eh = 1;
    a.itermax = itermax;
    a.neq_pw = neq_PW;
    a.n_eff = n_eff;
    a.bloch = bloch_vec;
    a.shift = shift;
    if (opts && opts[0] > 0) {
      a.tol = opts[0];
      a.tol_set = true;
    }
    if (opts && opts[1] != 0) a.precision = (int64_t)opts[1];
    if (a.precision != 0 && a.precision != 1) return fail(EMU_E_ARG, "precision must be 0 (FP64) or 1 (mixed)");
    a.block = 1;  // single vectors on the small banded 1D pencils
    if (opts && opts[2] > 0) a.block = (int64_t)opts[2];
    if (opts && opts[3] > 0) a.ncv = (int64_t)opts[3];
    if (a.block != 1 && a.block != 2 && a.block != 4 && a.block != 8)
      return fail(EMU_E_ARG, "Krylov block size must be 1, 2, 4 or 8");
    a.beta = beta1;
    a.J = overlap_J;
    a.Jd = overlap_J_dagger;
    a.world_1d = world_1d;
    a.neq_pw_2d = neq_PW_2d;
    a.J2d = overlap_J_2d;
    a.Jd2d = overlap_J_dagger_2d;
    a.stats = stats;
    a.sol = sol;
    emu::calc_modes_1d(ctx->c, *mesh->m, a);
  })
}

extern "C" {

int emu_calc_modes_1d_batch(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda, int64_t nval,
                            int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                            const double *shift, int64_t itermax, int64_t neq_PW, int64_t neq_PW_2d,
                            int64_t world_1d, const double *opts, double *beta1, double *overlap_J,
                            double *overlap_J_dagger, double *overlap_J_2d, double *overlap_J_dagger_2d,
                            int64_t *stats) {
  return modes1d_impl(ctx, mesh, nbatch, lambda, nval, ordre_ls, nb_typ_el, n_eff, bloch_vec, shift, itermax, neq_PW,
                      neq_PW_2d, world_1d, opts, beta1, overlap_J, overlap_J_dagger, overlap_J_2d, overlap_J_dagger_2d,
                      stats, nullptr);
}

int emu_calc_modes_1d(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                      int64_t nb_typ_el, const double *n_eff, double bloch_vec_x, double bloch_vec_y,
                      const double *shift, int64_t itermax, int64_t neq_PW, int64_t neq_PW_2d, int64_t world_1d,
                      double *beta1, double *overlap_J, double *overlap_J_dagger, double *overlap_J_2d,
                      double *overlap_J_dagger_2d) {
  const double bl[2] = {bloch_vec_x, bloch_vec_y};
  return emu_calc_modes_1d_batch(ctx, mesh, 1, &lambda, nval, ordre_ls, nb_typ_el, n_eff, bl, shift, itermax,
                                 neq_PW, neq_PW_2d, world_1d, nullptr, beta1, overlap_J, overlap_J_dagger,
                                 overlap_J_2d, overlap_J_dagger_2d, nullptr);
}

int emu_calc_modes_1d_fields(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                             int64_t nb_typ_el, const double *n_eff, double bloch_vec_x, double bloch_vec_y,
                             const double *shift, int64_t itermax, int64_t neq_PW, int64_t neq_PW_2d,
                             int64_t world_1d, double *beta1, double *overlap_J, double *overlap_J_dagger,
                             double *overlap_J_2d, double *overlap_J_dagger_2d, double *sol_P2) {
  if (!sol_P2) return fail(EMU_E_ARG, "emu_calc_modes_1d_fields: sol_P2 is NULL");
  const double bl[2] = {bloch_vec_x, bloch_vec_y};
  return modes1d_impl(ctx, mesh, 1, &lambda, nval, ordre_ls, nb_typ_el, n_eff, bl, shift, itermax, neq_PW, neq_PW_2d,
                      world_1d, nullptr, beta1, overlap_J, overlap_J_dagger, overlap_J_2d, overlap_J_dagger_2d,
                      nullptr, sol_P2);
}

int emu_star_cascade_batch(emu_ctx *ctx, int64_t nbatch, int64_t nlayers, int64_t neq_PW, const int64_t *kind,
                           const int64_t *nmodes, const double *const *kz, const double *const *Z,
                           const double *const *J, const double *const *J_dagger, const double *heights,
                           const double *Zair, const double *d_minus, const int64_t *nprop, double *R_net,
                           double *T_net, double *a_list, double *r_list, double *t_list) {
  if (!ctx || !kind || !nmodes || !kz || !heights || !Zair || !d_minus || !nprop || !a_list || !r_list || !t_list)
    return fail(EMU_E_ARG, "emu_star_cascade_batch: NULL argument");
  if (nbatch < 1 || nlayers < 2 || neq_PW < 1) return fail(EMU_E_ARG, "emu_star_cascade_batch: invalid size");
  for (int64_t l = 0; l < nlayers; ++l) {
    if (kind[l] != 0 && kind[l] != 1) return fail(EMU_E_ARG, "layer kind must be 0 (film) or 1 (structured)");
    if (!kz[l]) return fail(EMU_E_ARG, "kz missing for a layer");
    if (kind[l] == 0 && (!Z || !Z[l] || nmodes[l] != 2 * neq_PW))
      return fail(EMU_E_ARG, "film layers need Z and nmodes = 2 neq_PW");
    if (kind[l] == 1 && (!J || !J_dagger || !J[l] || !J_dagger[l] || nmodes[l] < 1))
      return fail(EMU_E_ARG, "structured layers need J, J_dagger and nmodes");
  }
  EMU_GUARD({
    emu::CascadeArgs a;
    a.nbatch = nbatch;
    a.nlayers = nlayers;
    a.neq_pw = neq_PW;
    a.kind = kind;
    a.nmodes = nmodes;
    a.kz = kz;
    a.Z = Z;
    a.J = J;
    a.Jd = J_dagger;
    a.heights = heights;
    a.Zair = Zair;
    a.d_minus = d_minus;
    a.nprop = nprop;
    a.R_net = R_net;
    a.T_net = T_net;
    a.a_list = a_list;
    a.r_list = r_list;
    a.t_list = t_list;
    emu::star_cascade(ctx->c, a);
  })
}

int emu_material_index(emu_ctx *ctx, int64_t nmat, const int64_t *kind, const int64_t *offset,
                       const double *xs, const double *ys, int64_t nq, const int64_t *material,
                       const double *wl_nm, double *n_out) {
  if (!ctx || !kind || !offset || (nq > 0 && (!material || !wl_nm || !n_out)))
    return fail(EMU_E_ARG, "emu_material_index: NULL argument");
  if (nmat < 1 || nq < 0) return fail(EMU_E_ARG, "emu_material_index: invalid size");
  if (offset[0] != 0) return fail(EMU_E_ARG, "emu_material_index: offset[0] must be 0");
  for (int64_t m = 0; m < nmat; ++m)
    if (offset[m + 1] < offset[m]) return fail(EMU_E_ARG, "emu_material_index: offsets must not decrease");
  if (offset[nmat] > 0 && (!xs || !ys)) return fail(EMU_E_ARG, "emu_material_index: NULL table");
  EMU_GUARD(emu::material_index(ctx->c, nmat, kind, offset, xs, ys, nq, material, wl_nm, n_out))
}

int emu_film_modes_batch(emu_ctx *ctx, int64_t nbatch, int64_t npw, const double *n, const double *wl_norm,
                         const double *alpha, const double *beta, const int64_t *order, double *k_z, double *Z) {
  if (!ctx || !n || !wl_norm || !alpha || !beta || !order || !k_z || !Z)
    return fail(EMU_E_ARG, "emu_film_modes_batch: NULL argument");
  if (nbatch < 1 || npw < 1) return fail(EMU_E_ARG, "emu_film_modes_batch: invalid size");
  EMU_GUARD(emu::film_modes(ctx->c, nbatch, npw, n, wl_norm, alpha, beta, order, k_z, Z))
}

int emu_assemble(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, int64_t nb_typ_el, const double *pp,
                 const double *qq, const double *shift, const double *bloch, double *mat1, double *mat2) {
  if (!ctx || !mesh || !pp || !qq || !shift || !bloch) return fail(EMU_E_ARG, "emu_assemble: NULL argument");
  if (nbatch < 1 || nb_typ_el < 1) return fail(EMU_E_ARG, "emu_assemble: invalid size");
  EMU_GUARD(emu::assemble(ctx->c, *mesh->m, (int)nbatch, (int)nb_typ_el, pp, qq, shift, bloch, mat1, mat2))
}

int emu_assemble_noalloc(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, int64_t nb_typ_el, const double *pp,
                         const double *qq, const double *shift, const double *bloch, double *mat1, double *mat2,
                         void *stream) {
  if (!ctx || !mesh || !pp || !qq || !shift || !bloch || !mat1 || !mat2)
    return fail(EMU_E_ARG, "emu_assemble_noalloc: NULL argument");
  if (nbatch < 1 || nb_typ_el < 1) return fail(EMU_E_ARG, "emu_assemble_noalloc: invalid size");
  EMU_GUARD(emu::assemble_noalloc(ctx->c, *mesh->m, (int)nbatch, (int)nb_typ_el, pp, qq, shift, bloch, mat1, mat2,
                                  stream))
}

int emu_lu_solve(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *mat1, int64_t nrhs, const double *b,
                 double *x, int64_t *stats) {
  if (!ctx || !mesh || !mat1 || !b || !x) return fail(EMU_E_ARG, "emu_lu_solve: NULL argument");
  if (nbatch < 1 || nrhs < 1) return fail(EMU_E_ARG, "emu_lu_solve: invalid size");
  EMU_GUARD(emu::lu_solve_host(ctx->c, *mesh->m, (int)nbatch, mat1, (int)nrhs, b, x, stats))
}

}  // extern "C"
