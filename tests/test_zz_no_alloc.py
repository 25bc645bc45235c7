"""emu_assemble_noalloc (allocation-free assembly into caller device
buffers, scratch in a fixed per-GPU device array) against emu_assemble on
the same inputs: bit-identical mat1 = K - shift M and mat2 = M, on two calls
in a row (the second with fewer instances and another mesh), for a 2D and a
1D mesh; requests beyond the fixed scratch are rejected before any launch.
The test allocates its device buffers with hipMalloc (through ctypes, the
caller's side); the library call goes through the C-ABI."""
import numpy as np
import pytest

from conftest import MESH_600, MESH_940

pytestmark = pytest.mark.gpu


def _inputs(nb, ntyp, seed):
    rng = np.random.default_rng(seed)
    pp = (rng.uniform(0.5, 1.5, (nb, ntyp)) + 1j * rng.uniform(0, 0.1, (nb, ntyp))).astype(np.complex128)
    qq = (rng.uniform(10, 200, (nb, ntyp)) + 1j * rng.uniform(0, 5, (nb, ntyp))).astype(np.complex128)
    shift = (rng.uniform(100, 300, nb) + 0j).astype(np.complex128)
    bloch = rng.uniform(-0.5, 0.5, (nb, 2))
    return pp, qq, shift, bloch


class _Hip:
    """hipMalloc / hipMemcpy / hipFree / hipDeviceSynchronize of the caller."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        try:
            self.L = ctypes.CDLL('libamdhip64.so')
        except OSError:
            self.L = ctypes.CDLL('/opt/rocm/lib/libamdhip64.so')
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes):
        p = self.ct.c_void_p()
        assert self.L.hipMalloc(self.ct.byref(p), max(int(nbytes), 8)) == 0
        self.ptrs.append(p)
        return p

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        return p

    def get(self, p, shape, dtype):
        out = np.zeros(shape, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.L.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.ptrs:
            self.L.hipFree(p)
        self.ptrs = []


def _noalloc(ctx, mesh, pp, qq, shift, bloch):
    from emustack_amd import _lib
    nb, ntyp = pp.shape
    h = _Hip()
    try:
        dpp, dqq, dsh, dbl = h.put(pp), h.put(qq), h.put(shift), h.put(bloch)
        m1 = h.alloc(nb * mesh.nnz * 16)
        m2 = h.alloc(nb * mesh.nnz * 16)
        _lib.check(_lib.lib().emu_assemble_noalloc(ctx.h, mesh.h, nb, ntyp, dpp, dqq, dsh, dbl, m1, m2, None))
        h.sync()
        return h.get(m1, (nb, mesh.nnz), np.complex128), h.get(m2, (nb, mesh.nnz), np.complex128)
    finally:
        h.free()


def test_noalloc_matches_assemble_twice(gpu_ctx):
    from emustack_amd import _lib
    big = _lib.MeshHandle(MESH_940, gpu_ctx)
    small = _lib.MeshHandle(MESH_600, gpu_ctx)
    for mesh, nb, seed in ((big, 3, 1), (small, 1, 2)):
        pp, qq, shift, bloch = _inputs(nb, 2, seed)
        ref1, ref2 = gpu_ctx.assemble(mesh, pp, qq, shift, bloch)
        got1, got2 = _noalloc(gpu_ctx, mesh, pp, qq, shift, bloch)
        np.testing.assert_array_equal(got1, ref1)
        np.testing.assert_array_equal(got2, ref2)


def test_noalloc_1d_mesh(gpu_ctx):
    from emustack_amd import _lib, mesher
    x, tn, te, _ = mesher.grating_1d(760.0, [190.0], 0.01)
    mesh = _lib.MeshHandle1D(x, tn, te, gpu_ctx)
    pp, qq, shift, bloch = _inputs(2, 2, 3)
    ref1, ref2 = gpu_ctx.assemble(mesh, pp, qq, shift, bloch)
    got1, got2 = _noalloc(gpu_ctx, mesh, pp, qq, shift, bloch)
    np.testing.assert_array_equal(got1, ref1)
    np.testing.assert_array_equal(got2, ref2)


def test_noalloc_rejects_oversized_batch(gpu_ctx):
    from emustack_amd import _lib
    mesh = _lib.MeshHandle(MESH_940, gpu_ctx)
    nb = (16 << 20) // (mesh.nel * 224) + 1
    h = _Hip()
    try:
        p = h.alloc(256)
        rc = _lib.lib().emu_assemble_noalloc(gpu_ctx.h, mesh.h, nb, 2, p, p, p, p, p, p, None)
    finally:
        h.free()
    assert rc == -1  # EMU_E_ARG, before any launch
    assert b'scratch' in _lib.lib().emu_last_error()
