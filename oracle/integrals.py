"""ORACLE -- test infrastructure only (see oracle/__init__.py).

ctypes front end of the C restatement in peec_integrals_oracle.c:
  lp_row   <- compute_Lp_Voxels      (compute_Lp_Voxels.jl:1-63)
  lp_row_qs <- compute_Lp_Voxels_QS  (compute_Lp_Voxels_QS.jl:9-42, the Rcc rows)
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
        ip = ctypes.POINTER(ctypes.c_int32)
        L.oracle_lp_row_forms.argtypes = [dp, dp, ctypes.c_int64] + [ctypes.c_double] * 6 + [ctypes.c_int, dp, ip]
        L.oracle_lp_row_forms.restype = None
        L.oracle_p_row_forms.argtypes = [dp, dp, ctypes.c_int64] + [ctypes.c_double] * 3 + \
            [ctypes.c_int, ctypes.c_int, dp, ip]
        L.oracle_p_row_forms.restype = None
        L.oracle_lp_row_qs.argtypes = L.oracle_lp_row.argtypes
        L.oracle_lp_row_qs.restype = None
        L.oracle_lp_row_qs_forms.argtypes = L.oracle_lp_row_forms.argtypes
        L.oracle_lp_row_qs_forms.restype = None
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


# closed-form ids of the routing trace (oracle_closed_form's ids, plus the
# far-field point-point term and the two collinear-segment forms)
FORM_NAMES = {0: "vol_vol", 1: "vol_surf", 2: "lp_surf_surf_para", 3: "lp_surf_surf_ortho", 4: "line_vol",
              5: "line_surf_para", 6: "lp_line_surf_ortho", 7: "point_vol", 8: "line_line_ortho_xy",
              9: "lp_line_line_parall", 10: "point_sup", 11: "line_point", 12: "qs_surf_surf_para",
              13: "p_surf_surf_ortho", 14: "p_line_surf_ortho", 15: "p_line_line_parall", 16: "far_field",
              17: "lp_line_line_sp", 18: "p_line_line_sp", 19: "qs_line_line_sp", 20: "qs_line_point",
              21: "qs_line_surf_ortho", 22: "qs_line_surf_para", 23: "qs_line_vol", 24: "qs_vol_surf",
              25: "qs_surf_surf_ortho_call_error"}
# the forms that do not evaluate the integral their call sites use them for
DEVIANT_FORMS = (1, 3, 6, 17, 18)


def lp_row_qs(src, centers, sx, sy, sz, sx2, sy2, sz2, dc):
    """One row of compute_Lp_Voxels_QS (compute_Lp_Voxels_QS.jl:9-42), the
    closed-form copy of the retarded (Rcc) rows; without the Lp escaling."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(3)
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    N = c.shape[0]
    out = np.empty(N, dtype=np.float64)
    cf = np.ascontiguousarray(c.T).reshape(-1)
    lib().oracle_lp_row_qs(_dp(src), _dp(cf), N, sx, sy, sz, sx2, sy2, sz2, int(dc), _dp(out))
    return out


def lp_row_qs_forms(src, centers, sx, sy, sz, sx2, sy2, sz2, dc):
    """(values, forms) of one compute_Lp_Voxels_QS row, as lp_row_forms."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(3)
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    N = c.shape[0]
    out = np.empty(N, dtype=np.float64)
    form = np.empty(N, dtype=np.int32)
    cf = np.ascontiguousarray(c.T).reshape(-1)
    lib().oracle_lp_row_qs_forms(_dp(src), _dp(cf), N, sx, sy, sz, sx2, sy2, sz2, int(dc), _dp(out),
                                 form.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out, form


def lp_row_forms(src, centers, sx, sy, sz, sx2, sy2, sz2, dc):
    """(values, forms) of one Lp row: forms[n] is the closed-form id entry n
    routes to, + 100 where a suppression test sat at its threshold."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(3)
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    N = c.shape[0]
    out = np.empty(N, dtype=np.float64)
    form = np.empty(N, dtype=np.int32)
    cf = np.ascontiguousarray(c.T).reshape(-1)
    lib().oracle_lp_row_forms(_dp(src), _dp(cf), N, sx, sy, sz, sx2, sy2, sz2, int(dc), _dp(out),
                              form.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out, form


def p_row_forms(src, centers, sx, sy, sz, caser, caser2):
    """(values, forms) of one P row, as lp_row_forms."""
    src = np.ascontiguousarray(src, dtype=np.float64).reshape(3)
    c = np.asarray(centers, dtype=np.float64).reshape(-1, 3)
    N = c.shape[0]
    out = np.empty(N, dtype=np.float64)
    form = np.empty(N, dtype=np.int32)
    cf = np.ascontiguousarray(c.T).reshape(-1)
    lib().oracle_p_row_forms(_dp(src), _dp(cf), N, sx, sy, sz, int(caser), int(caser2), _dp(out),
                             form.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    return out, form


def lp_self(l, W, T):
    return lib().oracle_lp_self(l, W, T)


def closed_form(fid, args):
    a = np.ascontiguousarray(args, dtype=np.float64)
    return lib().oracle_closed_form(int(fid), _dp(a))
