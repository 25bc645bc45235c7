"""Drop-in for the reference's f2py module `fortran.EMUstack`
(backend/fortran/Makefile:16-18, imported at mode_calcs.py:30).

calc_modes_2d and calc_modes_1d keep the positional signatures of
py_calc_modes.f:1-11 and py_calc_modes_1d.f:2-12 and return the same tuples;
the work runs on the GPU through libemu_mi355x.so.
Meshes are loaded once per path and cached with their symbolic analysis.
There is no CPU fallback: without the built library or a GPU this raises.
"""
import os
import threading

import numpy as np

from . import _lib

_state = threading.local()
_mesh_cache = {}
_cache_lock = threading.Lock()


_contexts = []


def context(device=None):
    """The calling thread's GPU context (device from EMU_DEVICE, default 0)."""
    if getattr(_state, 'ctx', None) is None:
        dev = int(os.environ.get('EMU_DEVICE', '0')) if device is None else device
        _state.ctx = _lib.Context(dev)
        with _cache_lock:
            _contexts.append(_state.ctx)
    return _state.ctx


def all_contexts():
    """Every context created so far (one per host thread that used the GPU)."""
    with _cache_lock:
        return list(_contexts)


def mesh(path, leaf_size=0):
    """Loaded mesh handle for `path` (cached per context and path).  The load
    (file, symbolic analysis, upload) runs outside the cache lock, so that
    other threads' contexts and cached meshes stay available meanwhile; a
    handle is only ever used by its own context's thread."""
    ctx = context()
    key = (id(ctx), os.path.abspath(path), leaf_size)
    with _cache_lock:
        h = _mesh_cache.get(key)
    if h is None:
        h = _lib.MeshHandle(os.path.abspath(path), ctx, leaf_size)
        with _cache_lock:
            h = _mesh_cache.setdefault(key, h)
    return h


def calc_modes_2d(lambda_, nval, ordre_ls, debug, msh_path, mesh_file, npt, nel, nb_typ_el, n_eff, bloch_vec,
                  shift, E_H_field, i_cond, itermax, plot_modes, plot_real, plot_imag, plot_abs, neq_PW,
                  cmplx_max, real_max, int_max):
    """Bloch modes of a 2D-periodic layer (py_calc_modes.f:1-958).

    Returns (beta1[nval], overlap_J[2 neq_PW, nval], overlap_J_dagger[nval, 2 neq_PW],
    sol1, mode_pol, table_nod[6, nel], type_el[nel], x_arr[2, npt]).  With
    plot_modes == 1, sol1[3, 13, nval, nel] holds the prime modes' fields at
    the 6 P2 nodes and Ez at the 7 further P3 nodes (array_sol.f:120-339),
    and mode_pol[4, nval] the x / y / z / core fractions of eps |E|^2 of the
    adjoint modes (mode_energy.f:60-141); otherwise both are None (the
    reference fills them with values only plotting reads,
    mode_calcs.py:396-408).  plot_real / plot_imag / plot_abs select the
    reference's gmsh post-processing files and are not used.
    cmplx_max, real_max and int_max size the reference's Fortran work arrays
    and are accepted and ignored.
    """
    if i_cond != 2:
        raise ValueError("only periodic boundary conditions (i_cond=2) are supported")
    h = mesh(msh_path + mesh_file)
    if h.npt != npt or h.nel != nel:
        raise ValueError("mesh header (%d, %d) does not match npt, nel = (%d, %d)" % (h.npt, h.nel, npt, nel))
    n_eff = np.asarray(n_eff, np.complex128).reshape(-1)
    if n_eff.size != nb_typ_el:
        raise ValueError("n_eff must have nb_typ_el entries")
    table_nod, type_el, x_arr = h.geometry()
    if plot_modes == 1:
        beta, J, Jd, sol1, pol = context().calc_modes_2d_fields(
            h, lambda_, int(nval), int(ordre_ls), n_eff, np.asarray(bloch_vec, float), complex(shift), int(neq_PW),
            E_H_field=int(E_H_field), itermax=int(itermax))
        return (beta, np.asfortranarray(J), np.asfortranarray(Jd), np.asfortranarray(sol1), np.asfortranarray(pol),
                table_nod, type_el, x_arr)
    beta, J, Jd, _ = context().calc_modes_batch(h, [lambda_], int(nval), int(ordre_ls), [n_eff],
                                                [np.asarray(bloch_vec, float)], [complex(shift)], int(neq_PW),
                                                E_H_field=int(E_H_field), itermax=int(itermax))
    return (beta[0], np.asfortranarray(J[0]), np.asfortranarray(Jd[0]), None, None, table_nod, type_el, x_arr)


def mesh_1d(x_P2, table_nod, type_el, leaf_size=0):
    """Mesh handle of a 1D grating (cached per context and element arrays)."""
    ctx = context()
    x = np.ascontiguousarray(np.asarray(x_P2, np.float64))
    tn = np.asarray(table_nod, np.int64)
    te = np.ascontiguousarray(np.asarray(type_el, np.int64))
    key = (id(ctx), '1d', x.tobytes(), np.ascontiguousarray(tn).tobytes(), te.tobytes(), leaf_size)
    with _cache_lock:
        h = _mesh_cache.get(key)
        if h is None:
            h = _lib.MeshHandle1D(x, tn, te, ctx, leaf_size)
            _mesh_cache[key] = h
        return h


def calc_modes_1d(lambda_, nval, ordre_ls, nb_typ_el, npt_P2, nel, table_nod, type_el, x_P2, itermax, debug,
                  mesh_file, n_eff, bloch_vec_x, bloch_vec_y, shift, plot_modes, plot_real, plot_imag, plot_abs,
                  neq_PW, neq_PW_2d, world_1d):
    """Bloch modes of a 1D grating (py_calc_modes_1d.f:2-12).

    Returns (beta_1[nval], overlap_J[2 neq_PW, nval], overlap_J_dagger[nval, 2 neq_PW],
    overlap_J_2d, overlap_J_dagger_2d, sol_P2); the 2D overlaps are filled when
    world_1d is 0 (else zero arrays of the reference's shapes).  With
    plot_modes == 1, sol_P2[3, 3, nval, nel] holds the prime modes after
    normalisation_1d at each element's two ends and mid-node
    (array_sol_P2_1d.f); otherwise it is None."""
    tn = np.asarray(table_nod, np.int64)
    te = np.asarray(type_el, np.int64).reshape(-1)
    x = np.asarray(x_P2, np.float64).reshape(-1)
    if te.size != nel or x.size != npt_P2:
        raise ValueError("nel / npt_P2 do not match the element arrays")
    h = mesh_1d(x, tn, te)
    n_eff = np.asarray(n_eff, np.complex128).reshape(-1)
    if n_eff.size != nb_typ_el:
        raise ValueError("n_eff must have nb_typ_el entries")
    w1 = bool(world_1d)
    if plot_modes == 1:
        beta, J, Jd, J2, Jd2, sol = context().calc_modes_1d_fields(
            h, lambda_, int(nval), int(ordre_ls), n_eff, [bloch_vec_x, bloch_vec_y], complex(shift), int(neq_PW),
            neq_pw_2d=int(neq_PW_2d), world_1d=w1, itermax=int(itermax))
        sol = np.asfortranarray(sol)
    else:
        beta, J, Jd, J2, Jd2, _ = context().calc_modes_1d_batch(
            h, [lambda_], int(nval), int(ordre_ls), [n_eff], [[bloch_vec_x, bloch_vec_y]], [complex(shift)],
            int(neq_PW), neq_pw_2d=int(neq_PW_2d), world_1d=w1, itermax=int(itermax))
        beta, J, Jd = beta[0], J[0], Jd[0]
        if not w1:
            J2, Jd2 = J2[0], Jd2[0]
        sol = None
    if w1:
        J2 = np.zeros((2 * neq_PW_2d, nval), np.complex128, order='F')
        Jd2 = np.zeros((nval, 2 * neq_PW_2d), np.complex128, order='F')
    else:
        J2, Jd2 = np.asfortranarray(J2), np.asfortranarray(Jd2)
    return beta, np.asfortranarray(J), np.asfortranarray(Jd), J2, Jd2, sol


def calc_modes_2d_batch(lambdas, nval, ordre_ls, mesh_path, n_effs, bloch_vecs, shifts, neq_PW, E_H_field=1,
                        itermax=30, tol=0.0, ncv=0, precision=0):
    """Batched form: many (lambda, n_eff, bloch_vec, shift) points on one mesh
    solved together.  Returns beta [nl, nval], J [nl, 2P, nval],
    J_dagger [nl, nval, 2P] and the stats vector.  ncv: Krylov dimension
    (0: the reference's rule 2 nval + nval/2 + 3); precision 0 (complex128)
    or 1 (mixed: complex64 factor in the Krylov solves, FP64 beta)."""
    h = mesh(mesh_path)
    return context().calc_modes_batch(h, lambdas, int(nval), int(ordre_ls), n_effs, bloch_vecs, shifts,
                                      int(neq_PW), E_H_field=int(E_H_field), itermax=int(itermax), tol=tol,
                                      ncv=int(ncv), precision=int(precision))
