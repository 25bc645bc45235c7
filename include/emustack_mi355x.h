/* C-ABI of libemu_mi355x.so: EMUstack's per-wavelength FEM + scattering
 * path on one MI355X (gfx950).
 *
 * The reference exposes this path to Python as the f2py module
 * `fortran.EMUstack` (backend/fortran/Makefile:16-18), imported at
 * backend/mode_calcs.py:30 and backend/objects.py:29.  Its hot entry point
 * is EMUstack.calc_modes_2d (backend/fortran/py_calc_modes.f:1-11), called
 * positionally at backend/mode_calcs.py:371-378.  The entry points below
 * replace it.  Plain pointers and sizes only; complex numbers are
 * interleaved (re, im) doubles; matrices are column-major (Fortran order)
 * exactly as f2py hands them to Python.
 *
 * Every function returns 0 on success or a negative EMU_E_* status; the
 * message of the last failure on the calling thread is emu_last_error().
 * A context owns one HIP stream on one device; calls on one context are
 * serialised by the caller (one context per host thread).
 */
#ifndef EMUSTACK_MI355X_H
#define EMUSTACK_MI355X_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  EMU_OK = 0,
  EMU_E_ARG = -1,         /* invalid argument (sizes, null pointers) */
  EMU_E_MESH = -2,        /* unreadable / non-periodic mesh (periodic_node.f:108-128) */
  EMU_E_HIP = -3,         /* HIP runtime error or no device */
  EMU_E_NOCONV = -4,      /* eigensolver did not converge (py_calc_modes.f:647-656) */
  EMU_E_HOMOGENEOUS = -5, /* all materials equal (py_calc_modes.f:553-567) */
  EMU_E_LAPACK = -6,      /* host LAPACK for the projected problem missing/failed */
  EMU_E_OOM = -7          /* device allocation failed */
};

typedef struct emu_ctx emu_ctx;
typedef struct emu_mesh emu_mesh;

/* Message of the last failure on this thread (never NULL). */
const char *emu_last_error(void);

/* Context on HIP device `device` with its own stream. */
int emu_ctx_create(int device, emu_ctx **out);
void emu_ctx_destroy(emu_ctx *ctx);

/* Host LAPACK routines (zgees, ztrsen, ztrevc; LAPACK calling convention,
 * all arguments by pointer, no hidden string lengths) used for the small
 * ncv x ncv projected Schur problem of the Krylov-Schur eigensolver -- the
 * role zlahqr/ztrevc play inside ARPACK (zarpack.f:3960 zneigh).  Must be set
 * once per process before the first eigensolve. */
int emu_set_lapack(void *zgees, void *ztrsen, void *ztrevc);

/* Load a .mail mesh: topology, DOF numbering, periodic merge, sparse
 * pattern (py_calc_modes.f:297-522), nested-dissection symbolic LU analysis
 * and upload of the per-mesh element integrals.  The reference redoes this
 * per wavelength inside calc_modes_2d; here it is cached in the handle.
 * leaf_size <= 0 selects the default. */
int emu_mesh_load(emu_ctx *ctx, const char *path, int leaf_size, emu_mesh **out);
void emu_mesh_free(emu_mesh *mesh);

/* Host-only analysis of a .mail mesh (no device needed): the same
 * preprocessing and symbolic analysis as emu_mesh_load, without upload.
 * The handle serves emu_mesh_info / _arrays / _geometry / _symbolic only. */
int emu_mesh_analyze(const char *path, int leaf_size, emu_mesh **out);

/* Symbolic multifrontal structure (any pointer may be NULL):
 * fronts[nfronts*8] = (s, u, f, b, parent, level, child0, child1) per front,
 * foff[nfronts] = offset of the f x f column-major front in the pool,
 * uoff[nfronts] = offset of its update indices, uidx/pmap[total update
 * entries] = update rows (nested-dissection numbering) and their positions
 * in the parent front, nnz2front[nnz] = pool offset of every CSR entry. */
int emu_mesh_symbolic(const emu_mesh *mesh, int32_t *fronts, int64_t *foff, int64_t *uoff,
                      int32_t *uidx, int32_t *pmap, int64_t *nnz2front);

/* info[0..15]: npt, nel, n_ddl, neq, nnz, nfronts, nlevels, pool entries,
 * max front order, factor entries, LU flops (as int64), n_edge, npt_p3,
 * total update entries, 0, 0 */
int emu_mesh_info(const emu_mesh *mesh, int64_t *info);
/* Host copies of per-mesh arrays (any pointer may be NULL):
 * row_ptr[neq+1], col_ind[nnz], ldof[nel*22], perm[neq] (new -> old),
 * lat[4] (lattice vectors v1, v2). */
int emu_mesh_arrays(const emu_mesh *mesh, int64_t *row_ptr, int32_t *col_ind,
                    int32_t *ldof, int32_t *perm, double *lat);
/* Mesh arrays returned by f2py's calc_modes_2d: table_nod[6*nel] (1-based,
 * column-major 6 x nel), type_el[nel], x_arr[2*npt]. */
int emu_mesh_geometry(const emu_mesh *mesh, int64_t *table_nod, int64_t *type_el,
                      double *x_arr);

/* Drop-in for EMUstack.calc_modes_2d (py_calc_modes.f:1-11, E-field,
 * periodic boundary).  Inputs as the f2py call at mode_calcs.py:371-378:
 * lambda = wl/period, nval = num_BMs, ordre_ls = max_order_PWs,
 * n_eff[2*nb_typ_el], bloch_vec[2], shift[2] (complex), E_H_field (1 or 2),
 * itermax (restart limit), neq_PW.  Outputs (caller-allocated):
 * beta1[2*nval], overlap_J[2*(2*neq_PW*nval)] (2neq_PW x nval, col-major),
 * overlap_J_dagger[2*(nval*2*neq_PW)] (nval x 2neq_PW, col-major). */
int emu_calc_modes_2d(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval,
                      int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff,
                      const double *bloch_vec, const double *shift, int64_t E_H_field,
                      int64_t itermax, int64_t neq_PW, double *beta1, double *overlap_J,
                      double *overlap_J_dagger);

/* emu_calc_modes_2d plus the plotting fields the reference fills when
 * plot_modes = 1 (py_calc_modes.f:1-11 outputs sol1, mode_pol; array_sol.f:
 * 120-339): sol1[2*(3*13*nval*nel)] complex, Fortran order (component, node,
 * mode, element): the field of every prime mode at the 6 P2 nodes (Ex, Ey,
 * Ez) and Ez at the 7 remaining P3 nodes, normalised as array_sol.f leaves
 * it (largest P2-node component = 1); mode_pol[2*(4*nval)] complex (real
 * values): fractions of |E|^2 in x, y, z and in the core material, of the
 * adjoint modes as the reference's last array_sol call leaves them.  Either
 * output may be NULL (not both). */
int emu_calc_modes_2d_fields(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                             int64_t nb_typ_el, const double *n_eff, const double *bloch_vec, const double *shift,
                             int64_t E_H_field, int64_t itermax, int64_t neq_PW, double *beta1, double *overlap_J,
                             double *overlap_J_dagger, double *sol1, double *mode_pol);

/* Batched form: nbatch independent (lambda, n_eff, bloch, shift) instances on
 * one mesh, solved together (one launch sequence; prime and adjoint problems
 * of every instance share the symbolic factorisation).  Arrays are
 * instance-major: lambda[nbatch], n_eff[nbatch*2*nb_typ_el],
 * bloch_vec[nbatch*2], shift[nbatch*2]; outputs beta1[nbatch*2*nval],
 * overlap_J[nbatch*2*2neq_PW*nval], overlap_J_dagger[nbatch*2*nval*2neq_PW].
 * opts (may be NULL; 4 entries): opts[0] = eigen tolerance: a Ritz pair is
 * converged when its residual estimate is <= tol max(eps^(2/3), |mu|), the
 * form of ARPACK's test (zarpack.f:2037-2047).  <= 0: 1e-12 for precision 0,
 * 1e-8 for precision 1.  (The reference asks ARPACK for machine precision,
 * py_calc_modes.f:174; with this build's block Krylov-Schur the relative
 * 1e-12 converges inside the reference's 30 restarts and gives beta within
 * 1e-9 of ARPACK's; near-degenerate modes, whose vectors neither method
 * fixes to working precision, are normalised by their overlap block.)
 * opts[1] = precision: 0 complex128 throughout; 1 mixed (the Krylov solves
 * read a complex64 copy of the FP64 inverted factor, and each converged
 * pair's beta is the FP64 two-sided Rayleigh quotient of the pencil);
 * opts[2] = Krylov
 * block size (1, 2, 4 or 8; <= 0: default 8); opts[3] = Krylov dimension
 * (<= 0: the reference's 2 nval + nval/2 + 3, rounded up to whole blocks).  stats[0..7] (may be NULL): OP applications,
 * restarts (max over instances), pivot perturbations, Arnoldi steps,
 * rank repairs, modes whose beta came from the FP64 two-sided Rayleigh
 * quotient, modes that kept the Ritz value (quotient disagreed by > 1e-4),
 * 0. */
int emu_calc_modes_2d_batch(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch,
                            const double *lambda, int64_t nval, int64_t ordre_ls,
                            int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                            const double *shift, int64_t E_H_field, int64_t itermax,
                            int64_t neq_PW, const double *opts, double *beta1,
                            double *overlap_J, double *overlap_J_dagger, int64_t *stats);

/* Asynchronous form of emu_calc_modes_2d_batch (the batch submit / wait pair
 * of the drop-in boundary; the reference's Python runs one blocking f2py call
 * per point, mode_calcs.py:371-378).  Copies the inputs, queues the batch on
 * the context's worker thread and returns at once with a job handle; the
 * outputs (caller-owned, kept alive until the wait) are complete when
 * emu_job_wait returns.  A context runs its jobs, and any other calculation
 * called on it, one at a time in submission order; batches on different
 * contexts run concurrently.  Every job must be waited on exactly once;
 * emu_job_wait returns the batch's status (and sets emu_last_error) and
 * frees the handle. */
typedef struct emu_job emu_job;
int emu_calc_modes_2d_submit(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda, int64_t nval,
                             int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff, const double *bloch_vec,
                             const double *shift, int64_t E_H_field, int64_t itermax, int64_t neq_PW,
                             const double *opts, double *beta1, double *overlap_J, double *overlap_J_dagger,
                             int64_t *stats, emu_job **out);
int emu_job_wait(emu_job *job);

/* 1D gratings.  Create a mesh handle from the arrays the reference passes to
 * calc_modes_1d (py_calc_modes_1d.f:2-12, built by objects.py:1054-1530):
 * x_P2[npt_P2] node coordinates in [0, 1], table_nod[3*nel] (element-major,
 * 1-based: left, right, mid node), type_el[nel] (1-based material index).
 * Requires npt_P2 = 2 nel + 1 (py_calc_modes_1d.f:186-191).  The handle
 * caches the DOF numbering of periodic_cond_1d.f, the sparse pattern and
 * the symbolic LU; free it with emu_mesh_free. */
int emu_mesh_create_1d(emu_ctx *ctx, int64_t npt_P2, int64_t nel, const int64_t *table_nod,
                       const int64_t *type_el, const double *x_P2, int leaf_size, emu_mesh **out);

/* Replaces calc_modes_1d (py_calc_modes_1d.f:2-12, called at
 * mode_calcs.py:333-342).  Same meaning as the reference arguments:
 * lambda = wl / period, n_eff[2*nb_typ_el] (complex interleaved),
 * bloch_vec_x / bloch_vec_y, shift[2] (complex), neq_PW = 2 ordre_ls + 1,
 * neq_PW_2d = the 2D circular plane-wave count (used when world_1d = 0).
 * Outputs: beta1[2*nval]; overlap_J[2*2neq_PW*nval] (F-order 2neq_PW x nval);
 * overlap_J_dagger[2*nval*2neq_PW] (F-order nval x 2neq_PW); when
 * world_1d = 0 also overlap_J_2d (2neq_PW_2d x nval) and overlap_J_dagger_2d
 * (nval x 2neq_PW_2d), else these two may be NULL.  The E-field formulation
 * is the only one (py_calc_modes_1d.f:146). */
int emu_calc_modes_1d(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                      int64_t nb_typ_el, const double *n_eff, double bloch_vec_x, double bloch_vec_y,
                      const double *shift, int64_t itermax, int64_t neq_PW, int64_t neq_PW_2d,
                      int64_t world_1d, double *beta1, double *overlap_J, double *overlap_J_dagger,
                      double *overlap_J_2d, double *overlap_J_dagger_2d);

/* Batched form of emu_calc_modes_1d, instance-major like
 * emu_calc_modes_2d_batch: bloch_vec[nbatch*2] = (x, y) per instance; the
 * 2D outputs are nbatch blocks of 2*2neq_PW_2d*nval doubles.  opts and
 * stats as for emu_calc_modes_2d_batch, except that the default Krylov block
 * size is 1: on these small banded pencils single vectors need about half
 * the operator applications of 8-column blocks and give the deepest
 * evanescent modes' beta several times more accurately. */
int emu_calc_modes_1d_batch(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *lambda,
                            int64_t nval, int64_t ordre_ls, int64_t nb_typ_el, const double *n_eff,
                            const double *bloch_vec, const double *shift, int64_t itermax,
                            int64_t neq_PW, int64_t neq_PW_2d, int64_t world_1d, const double *opts,
                            double *beta1, double *overlap_J, double *overlap_J_dagger,
                            double *overlap_J_2d, double *overlap_J_dagger_2d, int64_t *stats);

/* emu_calc_modes_1d plus sol_P2 (py_calc_modes_1d.f output when plot_modes =
 * 1; array_sol_P2_1d.f:20-60): sol_P2[2*(3*3*nval*nel)] complex, Fortran order
 * (component, node, mode, element), the prime modes after
 * normalisation_1d.f at the element's left, right and mid nodes. */
int emu_calc_modes_1d_fields(emu_ctx *ctx, emu_mesh *mesh, double lambda, int64_t nval, int64_t ordre_ls,
                             int64_t nb_typ_el, const double *n_eff, double bloch_vec_x, double bloch_vec_y,
                             const double *shift, int64_t itermax, int64_t neq_PW, int64_t neq_PW_2d,
                             int64_t world_1d, double *beta1, double *overlap_J, double *overlap_J_dagger,
                             double *overlap_J_2d, double *overlap_J_dagger_2d, double *sol_P2);

/* Star-product cascade with energy fluxes for nbatch stacks of identical
 * layer structure (Stack.calc_scat, stack.py:162-427, with the interface
 * matrices of r_t_mat_anallo / r_t_mat_tf_ns, mode_calcs.py:451-515), on the
 * GPU as batched dense LU + GEMM.  Layers bottom to top; layers 0 and
 * nlayers-1 must be thin films.  Per layer l: kind[l] (0 thin film, 1
 * structured), nmodes[l] (structured: Bloch modes; films: 2 neq_PW),
 * kz[l] (nbatch x nmodes[l] complex), Z[l] (films: nbatch x 2neq_PW complex
 * impedances, Anallo.Z; NULL for structured), J[l] / J_dagger[l] (structured:
 * nbatch x (2neq_PW x nmodes) and nbatch x (nmodes x 2neq_PW), F-order,
 * J already divided by area_norm as Simmo.J holds it; NULL for films),
 * heights[b * nlayers + l] (height / period of layer l in stack b; the
 * entries of layers 0 and nlayers-1 are ignored).
 * Zair[nbatch*2neq_PW]: impedances of the air reference; d_minus[nbatch *
 * 2neq_PW]: incident amplitudes in the top layer (specular_incidence);
 * nprop[nbatch*3]: propagating orders per polarisation of the air
 * reference, the top and the bottom film.  Outputs (complex interleaved
 * where complex): R_net, T_net [nbatch * (2neq_PW)^2] (may be NULL), a_list
 * and t_list [nbatch*(nlayers-1)], r_list [nbatch] (the reference's
 * a_list, t_list, r_list of each stack). */
int emu_star_cascade_batch(emu_ctx *ctx, int64_t nbatch, int64_t nlayers, int64_t neq_PW, const int64_t *kind,
                           const int64_t *nmodes, const double *const *kz, const double *const *Z,
                           const double *const *J, const double *const *J_dagger, const double *heights,
                           const double *Zair, const double *d_minus, const int64_t *nprop, double *R_net,
                           double *T_net, double *a_list, double *r_list, double *t_list);

/* Refractive indices of nq (material, wavelength) queries in one launch
 * (replaces Material.n of backend/materials.py:129-170 evaluated once per
 * wavelength point: linear interpolation of tables with interp1d's segment
 * choice, Drude n = sqrt(eps_inf - wp^2/(w^2 + i wg w)), w = 2 pi c/lambda).
 * Material m (kind[m]: 0 constant, 1 table, 2 Drude) owns entries
 * offset[m] .. offset[m+1]-1 of xs (table wavelengths in nm, increasing;
 * Drude: wp, wg, eps_inf in rad/s) and ys (complex indices, interleaved;
 * constant: ys[offset[m]] is the index).  Queries: material[q], wl_nm[q]
 * (inside the table's range; the caller checks).  n_out[2*nq] complex. */
int emu_material_index(emu_ctx *ctx, int64_t nmat, const int64_t *kind, const int64_t *offset,
                       const double *xs, const double *ys, int64_t nq, const int64_t *material,
                       const double *wl_nm, double *n_out);

/* Analytic plane-wave modes of homogeneous films for nbatch (film, light)
 * instances in one launch (replaces Anallo.calc_modes and Anallo.Z of
 * backend/mode_calcs.py:110-186, 195-212, evaluated per layer and
 * wavelength in numpy).  Per instance b: n[2*b] complex index (real part
 * only for a lossless film, as the caller decides), wl_norm[b] = lambda /
 * period; alpha, beta[b*npw + o] in-plane wavevector of order o (units of
 * 1/period); order[b*npw + i] (0-based) the order stored at position i (the
 * free-space reference's propagation order).  Output per instance, complex
 * interleaved, TE block then TM block (2*npw each): k_z[i] = k_z[npw + i] =
 * sqrt(k^2 - alpha^2 - beta^2), k = 2 pi n / wl_norm, principal branch;
 * Z = k / (n k_z) (TE), k_z / (n k) (TM).  An order with k_z exactly 0
 * (grazing) gets Z = 0: the caller moves k_parallel (Anallo scales it by
 * 1 - 1e-9) and calls again. */
int emu_film_modes_batch(emu_ctx *ctx, int64_t nbatch, int64_t npw, const double *n, const double *wl_norm,
                         const double *alpha, const double *beta, const int64_t *order, double *k_z, double *Z);

/* Live kernel timing with HIP events on the context's stream (benchmarks).
 * enable = 1 resets the counters and starts recording every `every`-th
 * solve sweep; enable = 0 stops.  out[k*3 + {0,1,2}] (may be NULL) for kernel
 * class k: launches timed, total milliseconds, total algorithmic bytes of
 * those launches.  Classes: 0 forward-solve level sweep (k_fwd), 1 backward
 * level sweep (k_bwd), 2 Gram-Schmidt projections (k_atb_mfma16 for the
 * 8-column blocks, k_atb_tall for smaller ones), 3 Gram-Schmidt updates
 * (k_combine; k_sub_vh_blk for blocks below 8 columns), 4 mass-matrix
 * product of the Krylov block (k_spmm_il; k_spmm below 8 columns),
 * 5..7 reserved. */
int emu_profile(emu_ctx *ctx, int enable, int every, double *out);

/* Stage entry points (parity tests and benchmarks). */

/* Assembly of mat1 = K - shift*M and mat2 = M (asmbly.f:3-341) for nbatch
 * instances: pp, qq [nbatch*2*nb_typ_el], shift [nbatch*2], bloch [nbatch*2].
 * Outputs mat1, mat2 [nbatch*2*nnz] in the CSR order of emu_mesh_arrays
 * (entry (row, col_ind[k]) at position k; the reference's CSC with
 * row = ind_ip, col = ind_jp). */
int emu_assemble(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, int64_t nb_typ_el,
                 const double *pp, const double *qq, const double *shift,
                 const double *bloch, double *mat1, double *mat2);

/* Allocation-free variant of emu_assemble.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments and results as emu_assemble, except that every
 * array is a DEVICE pointer (pp, qq [nbatch*2*nb_typ_el], shift [nbatch*2],
 * bloch [nbatch*2] in; mat1, mat2 [nbatch*2*nnz] out, allocated by the
 * caller) and the call is asynchronous on `stream` (NULL: the context's
 * stream).  The kernels' only scratch (per-element Bloch phases, nbatch x
 * nel x 224 bytes for 2D meshes, none for 1D) lives in a fixed 16 MiB
 * device array per GPU; larger requests fail with EMU_E_ARG before any
 * launch.  Calls that overlap on one GPU share that array and race. */
int emu_assemble_noalloc(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, int64_t nb_typ_el, const double *pp,
                         const double *qq, const double *shift, const double *bloch, double *mat1, double *mat2,
                         void *stream);

/* Factor mat1 (host, nbatch*2*nnz, CSR order) with the GPU multifrontal LU and
 * solve mat1 x = b for nrhs right-hand sides per instance
 * (b, x: nbatch*nrhs*2*neq, column-major per instance).  stats[0] = pivot
 * perturbations. */
int emu_lu_solve(emu_ctx *ctx, emu_mesh *mesh, int64_t nbatch, const double *mat1,
                 int64_t nrhs, const double *b, double *x, int64_t *stats);

#ifdef __cplusplus
}
#endif
#endif
