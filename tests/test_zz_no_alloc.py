"""peec_apply_K_noalloc (the allocation-free K x, FFT work grids in a static
device array) against peec_apply_K_batched on the same inputs, through the
C ABI: two calls in a row, the second with fewer right-hand sides, in
complex128 and complex64.  Both run the same kernels on the same tuned
launch shapes, so the results must be bitwise equal.  Device memory for
the operands comes from the HIP runtime (hipMalloc / hipMemcpy through
ctypes, synchronous, so ordered with the context's stream by
peec_synchronize); the entry point itself allocates nothing."""
import ctypes

import numpy as np
import pytest

from solver_amd import _lib, configs, mesher, peec, sweep

from test_gpu_operator import pself

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def hip():
    h = ctypes.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DeviceBuffer:
    def __init__(self, nbytes):
        self.h = hip()
        self.ptr = ctypes.c_void_p()
        self.nbytes = nbytes
        assert self.h.hipMalloc(ctypes.byref(self.ptr), nbytes) == 0

    def upload(self, a):
        a = np.ascontiguousarray(a)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D) == 0

    def download(self, dtype, shape):
        a = np.empty(shape, dtype)
        assert a.nbytes == self.nbytes
        assert self.h.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, D2H) == 0
        return a

    def free(self):
        if self.ptr:
            assert self.h.hipFree(self.ptr) == 0
            self.ptr = ctypes.c_void_p()


def noalloc_apply(ctx, X, rps, fp32):
    N, nrhs = X.shape
    xt = np.ascontiguousarray(X.T)                        # column-major N x nrhs
    dx, dy = DeviceBuffer(xt.nbytes), DeviceBuffer(xt.nbytes)
    try:
        dx.upload(xt)
        dy.upload(np.full_like(xt, complex(np.nan, np.nan)))   # garbage the call must overwrite
        _lib.check(_lib.lib().peec_synchronize(ctx.handle), "peec_synchronize")
        rc = _lib.lib().peec_apply_K_noalloc(ctx.handle, nrhs, int(rps), int(bool(fp32)), dx.ptr, dy.ptr)
        _lib.check(rc, "peec_apply_K_noalloc")
        _lib.check(_lib.lib().peec_synchronize(ctx.handle), "peec_synchronize")
        return np.ascontiguousarray(dy.download(np.complex128, (nrhs, N)).T)
    finally:
        dx.free()
        dy.free()


@pytest.fixture(scope="module")
def geometry():
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(20, 14, 6, cell, 4, 0.45, 4.4, 0.02, seed=3)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), configs._x_face_ports(g, cell, 4), p_self_fn=pself)
    ctx = peec.Context(0)
    S = sweep.Solver(M, ctx=ctx, device_lu=True)  # two frequency slots need the device LU analysis
    peec.frequency_slots(2, ctx=ctx)
    for slot, f in enumerate((3e7, 6e9)):
        w = 2 * np.pi * f * M.escalings["freq"]
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=ctx)
    yield M, S, ctx
    ctx.close()


@pytest.mark.parametrize("fp32", [False, True])
def test_noalloc_matches_batched_entry_point(geometry, fp32):
    M, S, ctx = geometry
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(3)
    for nrhs, rps in ((4, 2), (1, 1)):
        X = rng.normal(size=(N, nrhs)) + 1j * rng.normal(size=(N, nrhs))
        ref = np.ascontiguousarray(peec.apply_K_batched(X, rps, fp32=fp32, ctx=ctx))
        got = noalloc_apply(ctx, X, rps, fp32)
        same = np.array_equal(got.view(np.float64), ref.view(np.float64))
        print("fp32=%s nrhs %d: bitwise %s, rel diff %.2e" % (fp32, nrhs, same,
                                                              np.linalg.norm(got - ref) / np.linalg.norm(ref)),
              flush=True)
        assert same


def test_noalloc_rejects_oversized_calls(geometry):
    """More right-hand sides than the static work grids hold is an error,
    not a fault (and nothing is launched)."""
    M, S, ctx = geometry
    d = DeviceBuffer(16)
    try:
        rc = _lib.lib().peec_apply_K_noalloc(ctx.handle, 1 << 20, 1 << 20, 0, d.ptr, d.ptr)
    finally:
        d.free()
    assert rc == -1 and b"above" in _lib.lib().peec_last_error()
