"""ORACLE -- test infrastructure only (see oracle/__init__.py).

ctypes front end of the C restatement in peec_integrals_oracle.c:
  lp_row   <- compute_Lp_Voxels      (compute_Lp_Voxels.jl:1-63)
  p_row    <- compute_row_P_sup      (compute_row_P_sup.jl:1-104)
  lp_self  <- Lp_self                (compute_diagonals.jl:136-155)
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libpeec_oracle.so")
_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_lp_row.argtypes = [dp, dp, ctypes.c_int64] + [ctypes.c_double] * 6 + [ctypes.c_int, dp]
        L.oracle_lp_row.restype = None
        L.oracle_p_row.argtypes = [dp, dp, ctypes.c_int64] + [ctypes.c_double] * 3 + [ctypes.c_int, ctypes.c_int, dp]
        L.oracle_p_row.restype = None
        L.oracle_lp_self.argtypes = [ctypes.c_double] * 3
        L.oracle_lp_self.restype = ctypes.c_double
        L.oracle_closed_form.argtypes = [ctypes.c_int, dp]
        L.oracle_closed_form.restype = ctypes.c_double
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def lp_row(src, centers, sx, sy, sz, sx2, sy2, sz2, dc):
    """One Lp row (without the Lp escaling): out[n] = 1e-7/S * Integ."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(3)
    c = np.asfortranarray(np.asarray(centers, dtype=np.float64).reshape(-1, 3))
    N = c.shape[0]
    out = np.empty(N, dtype=np.float64)
    cf = np.ascontiguousarray(c.T).reshape(-1)  # column-major N x 3
    lib().oracle_lp_row(_dp(src), _dp(cf), N, sx, sy, sz, sx2, sy2, sz2, int(dc), _dp(out))
    return out


def p_row(src, centers, sx, sy, sz, caser, caser2):
    """One P row (without the P escaling)."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(3)
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    N = c.shape[0]
    out = np.empty(N, dtype=np.float64)
    cf = np.ascontiguousarray(c.T).reshape(-1)
    lib().oracle_p_row(_dp(src), _dp(cf), N, sx, sy, sz, int(caser), int(caser2), _dp(out))
    return out


def lp_self(l, W, T):
    return lib().oracle_lp_self(l, W, T)


def closed_form(fid, args):
    a = np.ascontiguousarray(args, dtype=np.float64)
    return lib().oracle_closed_form(int(fid), _dp(a))
