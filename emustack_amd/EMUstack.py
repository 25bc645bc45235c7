"""Drop-in for the reference's f2py module `fortran.EMUstack`
(backend/fortran/Makefile:16-18, imported at mode_calcs.py:30).

calc_modes_2d keeps the positional signature of py_calc_modes.f:1-11 and
returns the same tuple; the work runs on the GPU through libemu_mi355x.so.
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


def context(device=None):
    """The calling thread's GPU context (device from EMU_DEVICE, default 0)."""
    if getattr(_state, 'ctx', None) is None:
        dev = int(os.environ.get('EMU_DEVICE', '0')) if device is None else device
        _state.ctx = _lib.Context(dev)
    return _state.ctx


def mesh(path, leaf_size=0):
    """Loaded mesh handle for `path` (cached per context and path)."""
    ctx = context()
    key = (id(ctx), os.path.abspath(path), leaf_size)
    with _cache_lock:
        h = _mesh_cache.get(key)
        if h is None:
            h = _lib.MeshHandle(os.path.abspath(path), ctx, leaf_size)
            _mesh_cache[key] = h
        return h


def calc_modes_2d(lambda_, nval, ordre_ls, debug, msh_path, mesh_file, npt, nel, nb_typ_el, n_eff, bloch_vec,
                  shift, E_H_field, i_cond, itermax, plot_modes, plot_real, plot_imag, plot_abs, neq_PW,
                  cmplx_max, real_max, int_max):
    """Bloch modes of a 2D-periodic layer (py_calc_modes.f:1-958).

    Returns (beta1[nval], overlap_J[2 neq_PW, nval], overlap_J_dagger[nval, 2 neq_PW],
    sol1, mode_pol, table_nod[6, nel], type_el[nel], x_arr[2, npt]).  The field
    arrays sol1 and mode_pol are only used for plotting by the reference
    (mode_calcs.py:396-408) and are returned as None.
    cmplx_max, real_max and int_max size the reference's Fortran work arrays
    and are accepted and ignored.
    """
    if i_cond != 2:
        raise ValueError("only periodic boundary conditions (i_cond=2) are supported")
    if plot_modes == 1:
        raise NotImplementedError("field plotting output is not produced by the GPU path")
    h = mesh(msh_path + mesh_file)
    if h.npt != npt or h.nel != nel:
        raise ValueError("mesh header (%d, %d) does not match npt, nel = (%d, %d)" % (h.npt, h.nel, npt, nel))
    n_eff = np.asarray(n_eff, np.complex128).reshape(-1)
    if n_eff.size != nb_typ_el:
        raise ValueError("n_eff must have nb_typ_el entries")
    beta, J, Jd, _ = context().calc_modes_batch(h, [lambda_], int(nval), int(ordre_ls), [n_eff],
                                                [np.asarray(bloch_vec, float)], [complex(shift)], int(neq_PW),
                                                E_H_field=int(E_H_field), itermax=int(itermax))
    table_nod, type_el, x_arr = h.geometry()
    return (beta[0], np.asfortranarray(J[0]), np.asfortranarray(Jd[0]), None, None, table_nod, type_el, x_arr)


def calc_modes_2d_batch(lambdas, nval, ordre_ls, mesh_path, n_effs, bloch_vecs, shifts, neq_PW, E_H_field=1,
                        itermax=30, tol=0.0):
    """Batched form: many (lambda, n_eff, bloch_vec, shift) points on one mesh
    solved together.  Returns beta [nl, nval], J [nl, 2P, nval],
    J_dagger [nl, nval, 2P] and the stats vector."""
    h = mesh(mesh_path)
    return context().calc_modes_batch(h, lambdas, int(nval), int(ordre_ls), n_effs, bloch_vecs, shifts,
                                      int(neq_PW), E_H_field=int(E_H_field), itermax=int(itermax), tol=tol)
