"""ctypes binding of libemu_mi355x.so (C-ABI in include/emustack_mi355x.h).

The library is built in-tree by emustack_amd.build; importing this module
never builds and never falls back to anything else: if the library is
missing the error says how to build it.
"""
import contextlib
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libemu_mi355x.so')

EMU_OK = 0
ERRORS = {
    -1: ValueError,
    -2: RuntimeError,
    -3: RuntimeError,
    -4: RuntimeError,
    -5: ValueError,
    -6: RuntimeError,
    -7: MemoryError,
}

_lib = None
_lock = threading.Lock()

c_i64 = ctypes.c_int64
c_dp = ctypes.POINTER(ctypes.c_double)
c_i64p = ctypes.POINTER(ctypes.c_int64)
c_i32p = ctypes.POINTER(ctypes.c_int32)
c_vp = ctypes.c_void_p

EXPORTS = [
    'emu_last_error', 'emu_ctx_create', 'emu_ctx_destroy', 'emu_set_lapack', 'emu_mesh_load',
    'emu_mesh_analyze', 'emu_mesh_symbolic', 'emu_mesh_free', 'emu_mesh_info', 'emu_mesh_arrays',
    'emu_mesh_geometry', 'emu_calc_modes_2d', 'emu_calc_modes_2d_batch', 'emu_assemble', 'emu_lu_solve',
    'emu_profile', 'emu_mesh_create_1d', 'emu_calc_modes_1d', 'emu_calc_modes_1d_batch', 'emu_star_cascade_batch',
    'emu_material_index', 'emu_assemble_noalloc', 'emu_calc_modes_2d_fields', 'emu_calc_modes_1d_fields',
    'emu_calc_modes_2d_submit', 'emu_job_wait', 'emu_film_modes_batch',
]


class EmuError(RuntimeError):
    """Failure reported by the native library (status code + message)."""

    def __init__(self, code, msg):
        super().__init__('%s (status %d)' % (msg, code))
        self.code = code


_blas_lock = threading.Lock()
_blas_users = 0
_blas_limit = None


@contextlib.contextmanager
def _host_lapack_threads(nl):
    """The library projects the Krylov-Schur problem of each batch instance
    on its own host thread; the registered LAPACK (scipy's OpenBLAS) then
    runs single-threaded, so that concurrent calls do not contend for one
    BLAS thread pool (a threaded m = 504 zgees measured 5x slower than a
    single-threaded one on the GPU host).  The limit is process-wide, so it
    is reference-counted: concurrent library calls (sub-batches on several
    contexts) share it and the last one to leave restores the original."""
    global _blas_users, _blas_limit
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        yield
        return
    with _blas_lock:
        if _blas_users == 0:
            _blas_limit = threadpool_limits(limits=1, user_api='blas')
        _blas_users += 1
    try:
        yield
    finally:
        with _blas_lock:
            _blas_users -= 1
            if _blas_users == 0:
                _blas_limit.restore_original_limits()
                _blas_limit = None


def _lapack_pointers():
    """Addresses of zgees/ztrsen/ztrevc from scipy's LAPACK build (used for
    the small projected Schur problem of the eigensolver, on the host)."""
    import scipy.linalg.cython_lapack as cl
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    out = []
    for name in ('zgees', 'ztrsen', 'ztrevc'):
        cap = cl.__pyx_capi__[name]
        out.append(get_ptr(cap, get_name(cap)))
    return out


def lib():
    """Load the library once (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError('%s is missing: build it with `python -m emustack_amd.build` '
                              '(or __graft_entry__.build())' % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        L.emu_last_error.restype = ctypes.c_char_p
        L.emu_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(c_vp)]
        L.emu_ctx_destroy.argtypes = [c_vp]
        L.emu_ctx_destroy.restype = None
        L.emu_set_lapack.argtypes = [c_vp, c_vp, c_vp]
        L.emu_mesh_load.argtypes = [c_vp, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(c_vp)]
        L.emu_mesh_analyze.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(c_vp)]
        L.emu_mesh_symbolic.argtypes = [c_vp, c_i32p, c_i64p, c_i64p, c_i32p, c_i32p, c_i64p]
        L.emu_mesh_free.argtypes = [c_vp]
        L.emu_mesh_free.restype = None
        L.emu_mesh_info.argtypes = [c_vp, c_i64p]
        L.emu_mesh_arrays.argtypes = [c_vp, c_i64p, c_i32p, c_i32p, c_i32p, c_dp]
        L.emu_mesh_geometry.argtypes = [c_vp, c_i64p, c_i64p, c_dp]
        L.emu_calc_modes_2d.argtypes = [c_vp, c_vp, ctypes.c_double, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp,
                                        c_i64, c_i64, c_i64, c_dp, c_dp, c_dp]
        L.emu_calc_modes_2d_batch.argtypes = [c_vp, c_vp, c_i64, c_dp, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp,
                                              c_i64, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_i64p]
        L.emu_assemble.argtypes = [c_vp, c_vp, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_dp, c_dp]
        L.emu_lu_solve.argtypes = [c_vp, c_vp, c_i64, c_dp, c_i64, c_dp, c_dp, c_i64p]
        L.emu_profile.argtypes = [c_vp, ctypes.c_int, ctypes.c_int, c_dp]
        L.emu_mesh_create_1d.argtypes = [c_vp, c_i64, c_i64, c_i64p, c_i64p, c_dp, ctypes.c_int,
                                         ctypes.POINTER(c_vp)]
        L.emu_calc_modes_1d.argtypes = [c_vp, c_vp, ctypes.c_double, c_i64, c_i64, c_i64, c_dp, ctypes.c_double,
                                        ctypes.c_double, c_dp, c_i64, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp,
                                        c_dp]
        c_dpp = ctypes.POINTER(c_dp)
        L.emu_star_cascade_batch.argtypes = [c_vp, c_i64, c_i64, c_i64, c_i64p, c_i64p, c_dpp, c_dpp, c_dpp, c_dpp,
                                             c_dp, c_dp, c_dp, c_i64p, c_dp, c_dp, c_dp, c_dp, c_dp]
        L.emu_calc_modes_1d_batch.argtypes = [c_vp, c_vp, c_i64, c_dp, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp,
                                              c_i64, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_dp, c_dp,
                                              c_i64p]
        L.emu_material_index.argtypes = [c_vp, c_i64, c_i64p, c_i64p, c_dp, c_dp, c_i64, c_i64p, c_dp, c_dp]
        L.emu_film_modes_batch.argtypes = [c_vp, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_i64p, c_dp, c_dp]
        L.emu_assemble_noalloc.argtypes = [c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]
        L.emu_calc_modes_2d_submit.argtypes = [c_vp, c_vp, c_i64, c_dp, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp, c_i64,
                                               c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_i64p,
                                               ctypes.POINTER(ctypes.c_void_p)]
        L.emu_job_wait.argtypes = [c_vp]
        L.emu_calc_modes_2d_fields.argtypes = [c_vp, c_vp, ctypes.c_double, c_i64, c_i64, c_i64, c_dp, c_dp, c_dp,
                                               c_i64, c_i64, c_i64, c_dp, c_dp, c_dp, c_dp, c_dp]
        L.emu_calc_modes_1d_fields.argtypes = [c_vp, c_vp, ctypes.c_double, c_i64, c_i64, c_i64, c_dp,
                                               ctypes.c_double, ctypes.c_double, c_dp, c_i64, c_i64, c_i64, c_i64,
                                               c_dp, c_dp, c_dp, c_dp, c_dp, c_dp]
        for f in EXPORTS:
            getattr(L, f)
        rc = L.emu_set_lapack(*_lapack_pointers())
        if rc != EMU_OK:
            raise EmuError(rc, L.emu_last_error().decode())
        _lib = L
        return L


def check(rc):
    if rc != EMU_OK:
        msg = lib().emu_last_error().decode()
        raise EmuError(rc, msg)


def ptr(a, kind=c_dp):
    if a is None:
        return None
    return a.ctypes.data_as(kind)


def as_f64(a):
    """Complex array -> contiguous interleaved float64 view."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


class Job:
    """A batch queued with Context.calc_modes_batch_submit: owns the output
    arrays (and keeps the mesh handle alive) until wait() returns them."""

    def __init__(self, nl, nval, neq_pw, mesh):
        self.beta = np.zeros((nl, nval), np.complex128)
        self.J = np.zeros((nl, nval, 2 * neq_pw), np.complex128)
        self.Jd = np.zeros((nl, 2 * neq_pw, nval), np.complex128)
        self.st = np.zeros(8, np.int64)
        self.mesh = mesh
        self.h = None

    def wait(self):
        if self.h is None:
            raise RuntimeError('job already waited on (or never submitted)')
        h, self.h = self.h, None
        try:
            check(lib().emu_job_wait(h))
        finally:
            self._blas.__exit__(None, None, None)
        return self.beta, self.J.transpose(0, 2, 1), self.Jd.transpose(0, 2, 1), self.st


class Context:
    """One HIP device + stream (emu_ctx)."""

    def __init__(self, device=0):
        L = lib()
        h = c_vp()
        check(L.emu_ctx_create(int(device), ctypes.byref(h)))
        self.h = h
        self.device = device

    def assemble(self, mesh, pp, qq, shift, bloch):
        """mat1 = K - shift*M, mat2 = M for nb instances (CSR order of mesh.arrays())."""
        pp = np.atleast_2d(np.asarray(pp, np.complex128))
        qq = np.atleast_2d(np.asarray(qq, np.complex128))
        nb, ntyp = pp.shape
        shift = np.asarray(shift, np.complex128).reshape(nb)
        bloch = np.ascontiguousarray(np.asarray(bloch, np.float64).reshape(nb, 2))
        m1 = np.zeros((nb, mesh.nnz), np.complex128)
        m2 = np.zeros((nb, mesh.nnz), np.complex128)
        check(lib().emu_assemble(self.h, mesh.h, nb, ntyp, ptr(as_f64(pp)), ptr(as_f64(qq)), ptr(as_f64(shift)),
                                 ptr(bloch), ptr(m1.view(np.float64)), ptr(m2.view(np.float64))))
        return m1, m2

    def lu_solve(self, mesh, mat1, b):
        """Factor mat1 [nb, nnz] on the GPU and solve for b [nb, nrhs, neq]."""
        mat1 = np.ascontiguousarray(np.atleast_2d(mat1), np.complex128)
        nb = mat1.shape[0]
        b = np.ascontiguousarray(np.asarray(b, np.complex128).reshape(nb, -1, mesh.neq))
        x = np.zeros_like(b)
        st = np.zeros(8, np.int64)
        check(lib().emu_lu_solve(self.h, mesh.h, nb, ptr(mat1.view(np.float64)), b.shape[1],
                                 ptr(b.view(np.float64)), ptr(x.view(np.float64)), ptr(st, c_i64p)))
        return x, int(st[0])

    def calc_modes_batch(self, mesh, lam, nval, ordre, n_eff, bloch, shift, neq_pw, E_H_field=1, itermax=30,
                         tol=0.0, block=0, ncv=0, precision=0):
        """Batched calc_modes_2d: returns beta [nl, nval], J [nl, 2P, nval], Jd [nl, nval, 2P], stats.
        precision 0: complex128 throughout; 1: mixed (complex64 factor in the
        Krylov solves, FP64 two-sided Rayleigh-quotient beta)."""
        lam = np.ascontiguousarray(np.atleast_1d(np.asarray(lam, np.float64)))
        nl = lam.shape[0]
        n_eff = np.ascontiguousarray(np.asarray(n_eff, np.complex128).reshape(nl, -1))
        ntyp = n_eff.shape[1]
        bloch = np.ascontiguousarray(np.asarray(bloch, np.float64).reshape(nl, 2))
        shift = np.ascontiguousarray(np.asarray(shift, np.complex128).reshape(nl))
        beta = np.zeros((nl, nval), np.complex128)
        J = np.zeros((nl, nval, 2 * neq_pw), np.complex128)    # C-order view of F-order (2P x nval)
        Jd = np.zeros((nl, 2 * neq_pw, nval), np.complex128)   # C-order view of F-order (nval x 2P)
        opts = np.array([tol, float(precision), float(block), float(ncv)])
        st = np.zeros(8, np.int64)
        with _host_lapack_threads(nl):
            check(lib().emu_calc_modes_2d_batch(self.h, mesh.h, nl, ptr(lam), nval, ordre, ntyp,
                                                ptr(n_eff.view(np.float64)), ptr(bloch),
                                                ptr(shift.view(np.float64)), E_H_field, itermax, neq_pw, ptr(opts),
                                                ptr(beta.view(np.float64)), ptr(J.view(np.float64)),
                                                ptr(Jd.view(np.float64)), ptr(st, c_i64p)))
        return beta, J.transpose(0, 2, 1), Jd.transpose(0, 2, 1), st

    def calc_modes_batch_submit(self, mesh, lam, nval, ordre, n_eff, bloch, shift, neq_pw, E_H_field=1,
                                itermax=30, tol=0.0, block=0, ncv=0, precision=0):
        """calc_modes_batch queued on this context's worker thread; returns a
        Job whose wait() gives (beta, J, Jd, stats) as calc_modes_batch does."""
        lam = np.ascontiguousarray(np.atleast_1d(np.asarray(lam, np.float64)))
        nl = lam.shape[0]
        n_eff = np.ascontiguousarray(np.asarray(n_eff, np.complex128).reshape(nl, -1))
        ntyp = n_eff.shape[1]
        bloch = np.ascontiguousarray(np.asarray(bloch, np.float64).reshape(nl, 2))
        shift = np.ascontiguousarray(np.asarray(shift, np.complex128).reshape(nl))
        job = Job(nl, nval, neq_pw, mesh)
        job.ctx = self  # the context (and its worker thread) must outlive the job
        opts = np.array([tol, float(precision), float(block), float(ncv)])
        h = ctypes.c_void_p()
        # single-threaded host LAPACK from now until the job is waited on: the
        # library's worker thread may reach its Schur steps at any time
        job._blas = _host_lapack_threads(nl)
        job._blas.__enter__()
        try:
            check(lib().emu_calc_modes_2d_submit(self.h, mesh.h, nl, ptr(lam), nval, ordre, ntyp,
                                                 ptr(n_eff.view(np.float64)), ptr(bloch),
                                                 ptr(shift.view(np.float64)), E_H_field, itermax, neq_pw, ptr(opts),
                                                 ptr(job.beta.view(np.float64)), ptr(job.J.view(np.float64)),
                                                 ptr(job.Jd.view(np.float64)), ptr(job.st, c_i64p),
                                                 ctypes.byref(h)))
        except BaseException:
            job._blas.__exit__(None, None, None)
            raise
        job.h = h
        return job

    def calc_modes_2d_fields(self, mesh, lam, nval, ordre, n_eff, bloch, shift, neq_pw, E_H_field=1, itermax=30):
        """One point with the plotting fields: beta (nval,), J (2P, nval),
        Jd (nval, 2P), sol1 (3, 13, nval, nel) and mode_pol (4, nval)."""
        n_eff = np.ascontiguousarray(np.asarray(n_eff, np.complex128).reshape(-1))
        bloch = np.ascontiguousarray(np.asarray(bloch, np.float64).reshape(2))
        sh = np.array([complex(shift)])
        beta = np.zeros(nval, np.complex128)
        J = np.zeros((nval, 2 * neq_pw), np.complex128)
        Jd = np.zeros((2 * neq_pw, nval), np.complex128)
        sol = np.zeros((mesh.nel, nval, 13, 3), np.complex128)
        pol = np.zeros((nval, 4), np.complex128)
        with _host_lapack_threads(1):
            check(lib().emu_calc_modes_2d_fields(self.h, mesh.h, float(lam), nval, ordre, n_eff.size,
                                                 ptr(n_eff.view(np.float64)), ptr(bloch), ptr(sh.view(np.float64)),
                                                 E_H_field, itermax, neq_pw, ptr(beta.view(np.float64)),
                                                 ptr(J.view(np.float64)), ptr(Jd.view(np.float64)),
                                                 ptr(sol.view(np.float64)), ptr(pol.view(np.float64))))
        return beta, J.T, Jd.T, sol.transpose(3, 2, 1, 0), pol.T

    def calc_modes_1d_fields(self, mesh, lam, nval, ordre, n_eff, bloch, shift, neq_pw, neq_pw_2d=1, world_1d=True,
                             itermax=30):
        """One 1D point with sol_P2 (3, 3, nval, nel): beta, J, Jd, J2d, Jd2d,
        sol_P2 (J2d / Jd2d None for a 1D world)."""
        n_eff = np.ascontiguousarray(np.asarray(n_eff, np.complex128).reshape(-1))
        sh = np.array([complex(shift)])
        beta = np.zeros(nval, np.complex128)
        J = np.zeros((nval, 2 * neq_pw), np.complex128)
        Jd = np.zeros((2 * neq_pw, nval), np.complex128)
        J2 = np.zeros((nval, 2 * neq_pw_2d), np.complex128)
        Jd2 = np.zeros((2 * neq_pw_2d, nval), np.complex128)
        sol = np.zeros((mesh.nel, nval, 3, 3), np.complex128)
        with _host_lapack_threads(1):
            check(lib().emu_calc_modes_1d_fields(self.h, mesh.h, float(lam), nval, ordre, n_eff.size,
                                                 ptr(n_eff.view(np.float64)), float(bloch[0]), float(bloch[1]),
                                                 ptr(sh.view(np.float64)), itermax, neq_pw, neq_pw_2d,
                                                 1 if world_1d else 0, ptr(beta.view(np.float64)),
                                                 ptr(J.view(np.float64)), ptr(Jd.view(np.float64)),
                                                 ptr(J2.view(np.float64)), ptr(Jd2.view(np.float64)),
                                                 ptr(sol.view(np.float64))))
        if world_1d:
            return beta, J.T, Jd.T, None, None, sol.transpose(3, 2, 1, 0)
        return beta, J.T, Jd.T, J2.T, Jd2.T, sol.transpose(3, 2, 1, 0)

    def calc_modes_1d_batch(self, mesh, lam, nval, ordre, n_eff, bloch, shift, neq_pw, neq_pw_2d=0,
                            world_1d=True, itermax=30, tol=0.0, block=0, ncv=0, precision=0):
        """Batched calc_modes_1d on one MeshHandle1D.  Returns beta (nl, nval),
        J (nl, 2P, nval), Jd (nl, nval, 2P), J2d, Jd2d (None for a 1D world)
        and stats."""
        lam = np.ascontiguousarray(np.atleast_1d(np.asarray(lam, np.float64)))
        nl = lam.size
        n_eff = np.ascontiguousarray(np.asarray(n_eff, np.complex128).reshape(nl, -1))
        ntyp = n_eff.shape[1]
        bloch = np.ascontiguousarray(np.asarray(bloch, np.float64).reshape(nl, 2))
        shift = np.ascontiguousarray(np.asarray(shift, np.complex128).reshape(nl))
        beta = np.zeros((nl, nval), np.complex128)
        J = np.zeros((nl, nval, 2 * neq_pw), np.complex128)
        Jd = np.zeros((nl, 2 * neq_pw, nval), np.complex128)
        w2 = 0 if world_1d else 1
        J2 = np.zeros((nl, nval, 2 * neq_pw_2d), np.complex128) if w2 else None
        Jd2 = np.zeros((nl, 2 * neq_pw_2d, nval), np.complex128) if w2 else None
        opts = np.array([tol, float(precision), float(block), float(ncv)])
        st = np.zeros(8, np.int64)
        with _host_lapack_threads(nl):
            check(lib().emu_calc_modes_1d_batch(
                self.h, mesh.h, nl, ptr(lam), nval, ordre, ntyp, ptr(n_eff.view(np.float64)), ptr(bloch),
                ptr(shift.view(np.float64)), itermax, neq_pw, neq_pw_2d, 0 if w2 else 1, ptr(opts),
                ptr(beta.view(np.float64)), ptr(J.view(np.float64)), ptr(Jd.view(np.float64)),
                ptr(J2.view(np.float64)) if w2 else None, ptr(Jd2.view(np.float64)) if w2 else None,
                ptr(st, c_i64p)))
        out2 = (J2.transpose(0, 2, 1), Jd2.transpose(0, 2, 1)) if w2 else (None, None)
        return beta, J.transpose(0, 2, 1), Jd.transpose(0, 2, 1), out2[0], out2[1], st

    def star_cascade(self, kinds, kz, Z, J, Jd, heights, Zair, d_minus, nprop, want_nets=False):
        """Batched star-product cascade (emu_star_cascade_batch).  Per layer
        (bottom to top) arrays with a leading batch axis: kz[l] (nb, dim),
        Z[l] (nb, 2P) for films, J[l] (nb, 2P, modes), Jd[l] (nb, modes, 2P)
        for structured layers (None elsewhere); heights (nb, L) over the
        period (or (L,) shared by every stack).  Returns a_list (nb, L-1),
        r_list (nb,), t_list (nb, L-1) and, if asked, R_net, T_net."""
        L = len(kinds)
        nb = np.asarray(Zair).shape[0]
        d = np.asarray(Zair).shape[1]
        keep = []

        def cptr(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, np.complex128)
            keep.append(a)
            return a.view(np.float64).ctypes.data_as(c_dp)

        def fptr(a):  # (nb, r, c) -> per point F-order
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, np.complex128).transpose(0, 2, 1))
            keep.append(a)
            return a.view(np.float64).ctypes.data_as(c_dp)
        arr = lambda xs: (c_dp * L)(*xs)
        kz_p = arr([cptr(k) for k in kz])
        Z_p = arr([cptr(z) for z in Z])
        J_p = arr([fptr(j) for j in J])
        Jd_p = arr([fptr(j) for j in Jd])
        kinds = np.asarray(kinds, np.int64)
        nmodes = np.array([np.asarray(k).shape[1] for k in kz], np.int64)
        heights = np.asarray(heights, np.float64)
        if heights.ndim == 1:
            heights = np.broadcast_to(heights, (nb, L))
        if heights.shape != (nb, L):
            raise ValueError('heights must have shape (nbatch, nlayers) or (nlayers,)')
        heights = np.ascontiguousarray(heights)
        nprop = np.ascontiguousarray(nprop, np.int64)
        a_l = np.zeros((nb, L - 1))
        r_l = np.zeros(nb)
        t_l = np.zeros((nb, L - 1))
        R = np.zeros((nb, d, d), np.complex128) if want_nets else None
        T = np.zeros((nb, d, d), np.complex128) if want_nets else None
        check(lib().emu_star_cascade_batch(self.h, nb, L, d // 2, ptr(kinds, c_i64p), ptr(nmodes, c_i64p), kz_p,
                                           Z_p, J_p, Jd_p, ptr(heights), cptr(Zair), cptr(d_minus),
                                           ptr(nprop, c_i64p), ptr(R.view(np.float64)) if want_nets else None,
                                           ptr(T.view(np.float64)) if want_nets else None, ptr(a_l), ptr(r_l),
                                           ptr(t_l)))
        if want_nets:  # stored per point column-major
            R, T = R.transpose(0, 2, 1), T.transpose(0, 2, 1)
        return a_l, r_l, t_l, R, T

    def material_index(self, kind, offset, xs, ys, qmat, qwl):
        """emu_material_index: complex n of every query (qmat[q], qwl[q])."""
        kind = np.ascontiguousarray(kind, np.int64)
        offset = np.ascontiguousarray(offset, np.int64)
        xs = np.ascontiguousarray(xs, np.float64)
        ys = np.ascontiguousarray(ys, np.complex128)
        qmat = np.ascontiguousarray(qmat, np.int64)
        qwl = np.ascontiguousarray(qwl, np.float64)
        out = np.zeros(qmat.size, np.complex128)
        check(lib().emu_material_index(self.h, kind.size, ptr(kind, c_i64p), ptr(offset, c_i64p),
                                       ptr(xs) if xs.size else None, ptr(ys.view(np.float64)) if ys.size else None,
                                       qmat.size, ptr(qmat, c_i64p), ptr(qwl), ptr(out.view(np.float64))))
        return out

    def film_modes(self, n, wl_norm, alpha, beta, order):
        """emu_film_modes_batch: k_z and Z (nbatch x 2 npw complex, TE block
        then TM block) of homogeneous films; n, wl_norm per instance, alpha,
        beta (nbatch x npw) per order, order (nbatch x npw) the stored order."""
        n = np.ascontiguousarray(n, np.complex128).reshape(-1)
        wl_norm = np.ascontiguousarray(wl_norm, np.float64).reshape(-1)
        alpha = np.ascontiguousarray(alpha, np.float64)
        beta = np.ascontiguousarray(beta, np.float64)
        order = np.ascontiguousarray(order, np.int64)
        nb, npw = alpha.shape
        if n.size != nb or wl_norm.size != nb or beta.shape != (nb, npw) or order.shape != (nb, npw):
            raise ValueError("film_modes: inconsistent shapes")
        kz = np.zeros((nb, 2 * npw), np.complex128)
        Z = np.zeros((nb, 2 * npw), np.complex128)
        check(lib().emu_film_modes_batch(self.h, nb, npw, ptr(n.view(np.float64)), ptr(wl_norm), ptr(alpha),
                                         ptr(beta), ptr(order, c_i64p), ptr(kz.view(np.float64)),
                                         ptr(Z.view(np.float64))))
        return kz, Z

    PROFILE_CLASSES = ('solve_fwd', 'solve_bwd', 'gs_project', 'gs_update', 'spmv_M')

    def profile(self, enable=None, every=8):
        """enable=True: reset and start timing; False: stop; None: read.
        Returns {class: (launches, ms, algorithmic bytes)}."""
        out = np.zeros(24)
        flag = -1 if enable is None else (1 if enable else 0)
        check(lib().emu_profile(self.h, flag, int(every), ptr(out)))
        return {name: tuple(out[3 * k:3 * k + 3]) for k, name in enumerate(self.PROFILE_CLASSES)}

    def close(self):
        if self.h:
            lib().emu_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MeshHandle1D:
    """A 1D grating (arrays as objects.py builds them) with cached DOF
    numbering, pattern and symbolic LU on a context."""

    def __init__(self, x, table_nod, type_el, ctx, leaf_size=0):
        L = lib()
        x = np.ascontiguousarray(np.asarray(x, np.float64))
        tn = np.asarray(table_nod, np.int64)
        if tn.shape[0] == 3 and tn.ndim == 2:
            tn = tn.T
        tn = np.ascontiguousarray(tn)
        te = np.ascontiguousarray(np.asarray(type_el, np.int64))
        h = c_vp()
        check(L.emu_mesh_create_1d(ctx.h, x.size, te.size, ptr(tn, c_i64p), ptr(te, c_i64p), ptr(x),
                                   int(leaf_size), ctypes.byref(h)))
        self.h = h
        self.ctx = ctx
        info = np.zeros(16, np.int64)
        check(L.emu_mesh_info(h, ptr(info, c_i64p)))
        (self.npt, self.nel, self.n_ddl, self.neq, self.nnz, self.nfronts, self.nlevels, self.pool,
         self.max_f, self.factor_entries, self.lu_flops, self.n_edge, self.npt_p3, self.nw) = [int(v) for v in info[:14]]

    def close(self):
        if self.h:
            lib().emu_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MeshHandle:
    """A loaded .mail mesh with cached topology, pattern and symbolic LU."""

    def __init__(self, path, ctx=None, leaf_size=0):
        L = lib()
        h = c_vp()
        if ctx is None:
            check(L.emu_mesh_analyze(path.encode(), int(leaf_size), ctypes.byref(h)))
        else:
            check(L.emu_mesh_load(ctx.h, path.encode(), int(leaf_size), ctypes.byref(h)))
        self.h = h
        self.path = path
        self.ctx = ctx
        info = np.zeros(16, np.int64)
        check(L.emu_mesh_info(h, ptr(info, c_i64p)))
        (self.npt, self.nel, self.n_ddl, self.neq, self.nnz, self.nfronts, self.nlevels, self.pool,
         self.max_f, self.factor_entries, self.lu_flops, self.n_edge, self.npt_p3, self.nw) = [int(v) for v in info[:14]]

    def arrays(self):
        L = lib()
        rp = np.zeros(self.neq + 1, np.int64)
        ci = np.zeros(self.nnz, np.int32)
        ld = np.zeros(self.nel * 22, np.int32)
        pm = np.zeros(self.neq, np.int32)
        lat = np.zeros(4)
        check(L.emu_mesh_arrays(self.h, ptr(rp, c_i64p), ptr(ci, c_i32p), ptr(ld, c_i32p), ptr(pm, c_i32p),
                                ptr(lat)))
        return dict(row_ptr=rp, col_ind=ci, ldof=ld.reshape(self.nel, 22), perm=pm, lat=lat)

    def geometry(self):
        L = lib()
        tn = np.zeros(6 * self.nel, np.int64)
        te = np.zeros(self.nel, np.int64)
        x = np.zeros(2 * self.npt)
        check(L.emu_mesh_geometry(self.h, ptr(tn, c_i64p), ptr(te, c_i64p), ptr(x)))
        return tn.reshape(self.nel, 6).T.copy(order='F'), te, x.reshape(self.npt, 2).T.copy(order='F')

    def symbolic(self):
        L = lib()
        fr = np.zeros(8 * self.nfronts, np.int32)
        fo = np.zeros(self.nfronts, np.int64)
        uo = np.zeros(self.nfronts, np.int64)
        ui = np.zeros(max(self.nw, 1), np.int32)
        pmp = np.zeros(max(self.nw, 1), np.int32)
        nf = np.zeros(self.nnz, np.int64)
        check(L.emu_mesh_symbolic(self.h, ptr(fr, c_i32p), ptr(fo, c_i64p), ptr(uo, c_i64p), ptr(ui, c_i32p),
                                  ptr(pmp, c_i32p), ptr(nf, c_i64p)))
        return dict(fronts=fr.reshape(-1, 8), foff=fo, uoff=uo, uidx=ui[:self.nw], pmap=pmp[:self.nw],
                    nnz2front=nf)

    def close(self):
        if self.h:
            lib().emu_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
