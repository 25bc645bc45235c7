// Internal interface between the C-ABI layer (capi.cpp) and the engine.
#pragma once
#include <cstdint>
#include <stdexcept>
#include <string>

namespace emu {

struct Ctx;
struct DevMesh;

struct NoConvergence : std::runtime_error {
  explicit NoConvergence(const std::string &s) : std::runtime_error(s) {}
};
struct Homogeneous : std::runtime_error {
  explicit Homogeneous(const std::string &s) : std::runtime_error(s) {}
};
struct LapackMissing : std::runtime_error {
  explicit LapackMissing(const std::string &s) : std::runtime_error(s) {}
};

void set_lapack(void *zgees, void *ztrsen, void *ztrevc);
bool lapack_ready();
Ctx *ctx_create(int device);
void ctx_destroy(Ctx *c);
void profile(Ctx *c, int enable, int every, double *out);
DevMesh *mesh_load(Ctx *c, const std::string &path, int leaf);
void mesh_free(DevMesh *m);
DevMesh *mesh_analyze(const std::string &path, int leaf);
void mesh_symbolic(const DevMesh *m, int32_t *fronts, int64_t *foff, int64_t *uoff, int32_t *uidx, int32_t *pmap,
                   int64_t *nnz2front);
void mesh_info(const DevMesh *m, int64_t *info);
void mesh_arrays(const DevMesh *m, int64_t *row_ptr, int32_t *col_ind, int32_t *ldof, int32_t *perm, double *lat);
void mesh_geometry(const DevMesh *m, int64_t *table_nod, int64_t *type_el, double *x_arr);

void assemble(Ctx *c, DevMesh &dm, int nb, int ntyp, const double *pp, const double *qq, const double *shift,
              const double *bloch, double *mat1, double *mat2);
void lu_solve_host(Ctx *c, DevMesh &dm, int nb, const double *mat1, int nrhs, const double *b, double *x,
                   int64_t *stats);

struct ModesArgs {
  int64_t nl = 1;           // number of (lambda, bloch) points
  const double *lambda = nullptr;
  int64_t nval = 0, ordre = 0, ntyp = 0, eh = 1, itermax = 30, neq_pw = 0;
  const double *n_eff = nullptr;   // nl x ntyp complex
  const double *bloch = nullptr;   // nl x 2
  const double *shift = nullptr;   // nl complex
  double tol = 1e-12;  // relative residual of every wanted Ritz pair (see ks_project)
  int64_t block = 8;               // Krylov block size (1, 2, 4 or 8)
  // 0: complex128 everywhere; 1: mixed: the Krylov solves read a complex64
  // copy of the (FP64-computed) inverted factor, and every converged pair's
  // beta is corrected by the FP64 two-sided Rayleigh quotient of the pencil
  int64_t precision = 0;
  bool tol_set = false;            // tol given by the caller (else the precision's default)
  int64_t ncv = 0;                 // Krylov dimension (0: the reference's 2 nval + nval/2 + 3)
  double *beta = nullptr;          // nl x nval complex
  double *J = nullptr;             // nl x (2P x nval) complex
  double *Jd = nullptr;            // nl x (nval x 2P) complex
  int64_t *stats = nullptr;
  // 1D gratings: world_1d = 0 also fills the 2D plane-wave embedding
  int64_t world_1d = 1, neq_pw_2d = 0;
  double *J2d = nullptr;           // nl x (2 P2d x nval) complex
  double *Jd2d = nullptr;          // nl x (nval x 2 P2d) complex
  // plotting fields (host, may be null): 2D sol1 nl x (3 x 13 x nval x nel)
  // and mode_pol nl x (4 x nval); 1D sol_P2 nl x (3 x 3 x nval x nel)
  double *sol = nullptr;
  double *mode_pol = nullptr;
};
void calc_modes(Ctx *c, DevMesh &dm, const ModesArgs &a);
DevMesh *mesh_create_1d(Ctx *c, int64_t npt, int64_t nel, const int64_t *table_nod, const int64_t *type_el,
                        const double *x, int leaf);
void calc_modes_1d(Ctx *c, DevMesh &dm, const ModesArgs &a);

// Star-product cascade of a batch of stacks with identical layer structure
// (bottom to top; layers 0 and nlayers-1 are thin films).  Complex arrays
// interleaved, point-major.
struct CascadeArgs {
  int64_t nbatch = 1, nlayers = 0, neq_pw = 0;
  const int64_t *kind = nullptr;     // per layer: 0 thin film, 1 structured
  const int64_t *nmodes = nullptr;   // per layer: Bloch modes (structured) or 2 neq_pw
  const double *const *kz = nullptr; // per layer: nbatch x dim
  const double *const *Z = nullptr;  // per film layer: nbatch x 2neq_pw impedances
  const double *const *J = nullptr;  // per structured layer: nbatch x (2neq_pw x modes), F-order
  const double *const *Jd = nullptr; // per structured layer: nbatch x (modes x 2neq_pw)
  const double *heights = nullptr;   // [nbatch * nlayers]: per stack and layer, over the period
  const double *Zair = nullptr;      // nbatch x 2neq_pw: air reference impedances
  const double *d_minus = nullptr;   // nbatch x 2neq_pw incident amplitudes (top)
  const int64_t *nprop = nullptr;    // nbatch x 3: propagating orders of air, top film, bottom film
  double *R_net = nullptr, *T_net = nullptr;  // nbatch x (2neq_pw)^2 (optional)
  double *a_list = nullptr, *r_list = nullptr, *t_list = nullptr;  // nbatch x (nlayers-1), nbatch, nbatch x (nlayers-1)
};
void star_cascade(Ctx *c, const CascadeArgs &a);

// Allocation-free assembly (noalloc.hip): scratch in a fixed device array
constexpr size_t kNoallocScratchBytes = (size_t)16 << 20;  // 16 MiB per GPU
constexpr int kMaxDevices = 64;
void *noalloc_scratch(int device);
size_t assemble_noalloc_scratch(const DevMesh &dm, int64_t nb);
// device pointers in and out (complex interleaved), stream NULL = the context's
void assemble_noalloc(Ctx *c, DevMesh &dm, int nb, int ntyp, const double *pp, const double *qq,
                      const double *shift, const double *bloch, double *mat1, double *mat2, void *stream);

// n(lambda) of nq (material, wavelength) queries (kernels_material.hip)
void material_index(Ctx *c, int64_t nmat, const int64_t *kind, const int64_t *off, const double *xs,
                    const double *ys, int64_t nq, const int64_t *qm, const double *qw, double *out);
void film_modes(Ctx *c, int64_t nb, int64_t npw, const double *n, const double *wl_norm, const double *alpha,
                const double *beta, const int64_t *order, double *kz, double *Z);

}  // namespace emu
