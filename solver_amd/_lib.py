"""Loader of libpeec_hip.so (ctypes).  Fails loudly when the library is
missing: there is no CPU fallback for any entry point."""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpeec_hip.so")

PEEC_OK = 0
PEEC_STOPPED = 1

_lib = None

i64 = ctypes.c_int64
dp = ctypes.POINTER(ctypes.c_double)
i64p = ctypes.POINTER(ctypes.c_int64)


class C128(ctypes.Structure):
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


c128p = ctypes.POINTER(C128)


class CscF64(ctypes.Structure):
    _fields_ = [("colptr", i64p), ("rowval", i64p), ("nzval", dp), ("nrow", i64), ("ncol", i64)]


class CscC128(ctypes.Structure):
    _fields_ = [("colptr", i64p), ("rowval", i64p), ("nzval", c128p), ("nrow", i64), ("ncol", i64)]


class PeecError(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PeecError("libpeec_hip.so is not built (%s); run python solver_amd/build.py" % LIB_PATH)
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        L.peec_last_error.restype = ctypes.c_char_p
        L.peec_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(vp)]
        L.peec_ctx_destroy.argtypes = [vp]
        L.peec_set_stop_flag.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
        L.peec_lp_row.argtypes = [vp, dp, dp, i64] + [ctypes.c_double] * 6 + [ctypes.c_int, dp]
        L.peec_p_row.argtypes = [vp, dp, dp, i64] + [ctypes.c_double] * 3 + [ctypes.c_int, ctypes.c_int, dp]
        L.peec_lp_self.argtypes = [ctypes.c_double] * 3 + [dp]
        d = ctypes.c_double
        L.peec_kernels_build.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d, d, d]
        L.peec_kernel_row.argtypes = [vp, ctypes.c_int, dp, i64p]
        L.peec_kernels_build_rcc.argtypes = [vp, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, d, d, d, d, d]
        L.peec_rcc_spectra.argtypes = [vp, d]
        L.peec_operator_set.argtypes = [vp, i64, i64, i64, ctypes.POINTER(CscF64), ctypes.POINTER(CscF64),
                                        ctypes.POINTER(CscF64), ctypes.POINTER(CscF64), dp]
        L.peec_frequency_set.argtypes = [vp, d, c128p, ctypes.POINTER(CscC128)]
        L.peec_lu_set_factors.argtypes = [vp, i64, ctypes.POINTER(CscC128), ctypes.POINTER(CscC128), i64p, i64p, dp]
        L.peec_lu_set_factors_blocked.argtypes = [vp, i64, ctypes.POINTER(CscC128), ctypes.POINTER(CscC128), i64p,
                                                  i64p, dp, i64, i64p]
        L.peec_lu_analyze.argtypes = [vp, i64, i64p, i64, i64p, i64p, i64, i64p]
        L.peec_lu_factor.argtypes = [vp]
        L.peec_lu_front.argtypes = [vp, i64, i64p, i64p, i64p, c128p, i64p]
        ip = ctypes.POINTER(ctypes.c_int)
        L.peec_frequency_slots.argtypes = [vp, ctypes.c_int]
        L.peec_frequency_set_slot.argtypes = [vp, ctypes.c_int, ctypes.c_double, c128p, ctypes.POINTER(CscC128)]
        L.peec_lu_factor_slot.argtypes = [vp, ctypes.c_int]
        L.peec_lu_factor_slots.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.peec_lu_info.argtypes = [vp, i64p]
        L.peec_lu_set_solve.argtypes = [vp, ctypes.c_int]
        L.peec_lu_solve_info.argtypes = [vp, i64p]
        L.peec_gmres_batched.argtypes = [vp, ctypes.c_int, ctypes.c_int, c128p, ctypes.c_double, ctypes.c_int,
                                         ctypes.c_int, ctypes.c_int, ctypes.c_double, c128p, ip, dp, ip, ip]
        L.peec_precond_rhs_batched.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp, c128p]
        L.peec_sweep_batch.argtypes = [vp, ctypes.c_int, i64p, ctypes.c_double, c128p, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_int, ctypes.c_double, c128p, c128p, ip, dp, ip, ip]
        L.peec_sweep_state.argtypes = [vp, ctypes.c_int, c128p]
        L.peec_edge_terms_set.argtypes = [vp, i64, dp, c128p, c128p, dp, ctypes.c_void_p, ctypes.c_double,
                                          ctypes.c_double]
        L.peec_frequency_set_slot_edges.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.POINTER(CscC128)]
        L.peec_frequency_dz.argtypes = [vp, ctypes.c_int, c128p]
        L.peec_apply_K_batched.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c128p, c128p]
        L.peec_apply_K_noalloc.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                           ctypes.c_void_p]
        L.peec_apply_KH_batched.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, c128p, c128p]
        L.peec_precond_rhs_batched_adjoint.argtypes = [vp, ctypes.c_int, ctypes.c_int, c128p, c128p]
        L.peec_gmres_batched_adjoint.argtypes = [vp, ctypes.c_int, ctypes.c_int, c128p, ctypes.c_double,
                                                 ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, c128p,
                                                 ip, dp, ip, ip]
        L.peec_fft_tune.argtypes = [vp, ctypes.c_int, ctypes.c_int, dp]
        L.peec_set_graphs.argtypes = [vp, ctypes.c_int]
        L.peec_synchronize.argtypes = [vp]
        L.peec_copy_reference.argtypes = [vp, i64]
        L.peec_fft_set_choice.argtypes = [vp] + [ctypes.c_int] * 7
        L.peec_fft_choice.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ip]
        L.peec_matvec.argtypes = [vp, c128p, c128p]
        L.peec_apply_K.argtypes = [vp, c128p, c128p]
        L.peec_precond_rhs.argtypes = [vp, dp, c128p]
        L.peec_gmres.argtypes = [vp, c128p, d, ctypes.c_int, c128p, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(d), i64p, dp, i64p]
        L.peec_test_lp_row_host.argtypes = [dp, dp, i64] + [d] * 6 + [ctypes.c_int, dp]
        L.peec_test_p_row_host.argtypes = [dp, dp, i64] + [d] * 3 + [ctypes.c_int, ctypes.c_int, dp]
        L.peec_test_grid_row_host.argtypes = [ctypes.c_int, dp, ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
            [d] * 4 + [i64p, i64, dp, i64p]
        L.peec_profile.argtypes = [vp, ctypes.c_int, dp, dp, i64p, ctypes.c_int]
        L.peec_test_probe.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, i64, dp]
        _lib = L
    return _lib


def check(rc, what=""):
    if rc == PEEC_OK or rc == PEEC_STOPPED:
        return rc
    msg = lib().peec_last_error()
    raise PeecError("%s failed (%d): %s" % (what, rc, msg.decode() if msg else ""))


def as_dp(a):
    return a.ctypes.data_as(dp)


def as_c128p(a):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(c128p)
