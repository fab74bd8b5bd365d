"""CPU-side checks of libpeec_hip: the library loads, exports every entry
point declared in include/peec_hip.h, and the integral kernels' per-entry
code (run on the host through a test hook) equals the oracle bit for bit.
No GPU call is made here."""
import ctypes
import os
import re

import numpy as np
import pytest

from oracle import integrals as OI
from oracle.circulant import P_CASES, centers

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "peec_hip.h")).read()
    return sorted(set(re.findall(r"\b(peec_[a-z0-9_]+)\s*\(", txt)))


@pytest.fixture(scope="module")
def lib():
    from solver_amd import _lib
    return _lib.lib()


def test_every_declared_entry_point_is_exported(lib):
    names = declared_symbols()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


# the reference functions julia/PeecHip.jl redefines, with their arities
# (compute_Lp_Voxels.jl:1, compute_row_P_sup.jl:1, compute_diagonals.jl:136,
# gmres_custom.jl:1, compute_FFT_mutual_coupling_mats.jl:1,
# FFT_solver_QS_S_type.jl:1, compute_Matrix_vector_fft.jl:1,
# FFT_solver_QS_S_type.jl:237)
JULIA_SEAM = {"compute_Lp_Voxels": 11, "compute_row_P_sup": 8, "Lp_self": 3, "gmres_custom": 23,
              "compute_FFT_mutual_coupling_mats": 8, "FFT_solver_QS_S_type": 16, "ComputeMatrixVectorFTT": 16,
              "precond_3_3_Kt_fft!": 10}


def _julia_functions(txt):
    """name -> argument count of every top-level `function name(...)`."""
    out = {}
    for m in re.finditer(r"^function ([A-Za-z_][A-Za-z0-9_!]*)\((.*?)\)\s*$", txt, flags=re.M | re.S):
        args = re.sub(r"\s+", " ", m.group(2)).split(";")[0]
        out[m.group(1)] = len([a for a in args.split(",") if a.strip()])
    return out


def test_julia_shim_redefines_the_seam_and_calls_only_declared_symbols(lib):
    txt = open(os.path.join(ROOT, "julia", "PeecHip.jl")).read()
    funcs = _julia_functions(txt)
    for name, arity in JULIA_SEAM.items():
        assert funcs.get(name) == arity, (name, funcs.get(name), arity)
    called = set(re.findall(r"ccall\(\(:(peec_[a-z0-9_]+), LIBPEEC\)", txt))
    assert called and called <= set(declared_symbols()), called - set(declared_symbols())
    assert all(hasattr(lib, n) for n in called)


def _host_lp(lib, src, c, s, dc):
    from solver_amd import _lib as L
    cf = np.ascontiguousarray(c.T).reshape(-1)
    out = np.empty(len(c))
    src = np.ascontiguousarray(src, dtype=np.float64)
    lib.peec_test_lp_row_host(L.as_dp(src), L.as_dp(cf), len(c), *s, *s, dc, L.as_dp(out))
    return out


def _host_p(lib, src, c, s, c1, c2):
    from solver_amd import _lib as L
    cf = np.ascontiguousarray(c.T).reshape(-1)
    out = np.empty(len(c))
    src = np.ascontiguousarray(src, dtype=np.float64)
    lib.peec_test_p_row_host(L.as_dp(src), L.as_dp(cf), len(c), *s, c1, c2, L.as_dp(out))
    return out


@pytest.mark.parametrize("grid", [(24, 24, 4, 0.5e-3, 0.5e-3, 0.1e-3), (30, 20, 6, 1e-3, 1e-3, 1e-4)])
def test_kernel_entry_code_matches_oracle_bitwise(lib, grid):
    """The table-driven classifier and closed forms of integrals.hip, run on
    the CPU, reproduce the oracle's nested-branch restatement exactly."""
    Nx, Ny, Nz, sx, sy, sz = grid
    s = (sx, sy, sz)
    for kind in range(3):
        c = centers(kind, (0.0, 0.0, 0.0), Nx, Ny, Nz, s)
        assert np.array_equal(_host_lp(lib, c[0], c, s, kind + 1), OI.lp_row(c[0], c, *s, *s, kind + 1))
    for r, (c1, c2) in enumerate(P_CASES):
        c = centers(3 + r, (1e-3, 0.0, 0.0), Nx, Ny, Nz, s)[::3]
        assert np.array_equal(_host_p(lib, c[0], c, s, c1, c2), OI.p_row(c[0], c, *s, c1, c2))


def test_kernel_entry_code_all_branches_bitwise(lib):
    rng = np.random.default_rng(3)
    for trial in range(60):
        s1 = tuple(rng.choice([1e-6, 1e-5, 1e-4, 1e-3], size=3))
        src = rng.uniform(-1e-3, 1e-3, 3)
        dist = 10 ** rng.uniform(-4, -0.5, 100)
        dirs = rng.normal(size=(100, 3))
        dirs /= np.linalg.norm(dirs, axis=1)[:, None]
        tgt = src + dirs * dist[:, None]
        dc = int(rng.integers(1, 4))
        assert np.array_equal(_host_lp(lib, src, tgt, s1, dc), OI.lp_row(src, tgt, *s1, *s1, dc))
        c1, c2 = int(rng.integers(1, 7)), int(rng.integers(1, 7))
        assert np.array_equal(_host_p(lib, src, tgt, s1, c1, c2), OI.p_row(src, tgt, *s1, c1, c2))


def test_deterministic_math_within_one_ulp_of_libm(lib):
    from solver_amd import _lib as L
    rng = np.random.default_rng(0)
    n = 50000
    x = np.ascontiguousarray(10 ** rng.uniform(-300, 300, n))
    o = np.empty(n)
    lib.peec_test_probe(0, 1, L.as_dp(x), L.as_dp(x), n, L.as_dp(o))
    r = np.log(x)
    assert np.max(np.abs(o - r) / np.spacing(np.abs(r))) <= 1.0
    y = np.ascontiguousarray(rng.normal(size=n) * 10 ** rng.uniform(-10, 10, n))
    xx = np.ascontiguousarray(rng.normal(size=n) * 10 ** rng.uniform(-10, 10, n))
    lib.peec_test_probe(0, 2, L.as_dp(y), L.as_dp(xx), n, L.as_dp(o))
    r = np.arctan2(y, xx)
    assert np.max(np.abs(o - r) / np.spacing(np.abs(r))) <= 1.0


def test_rcc_row_entries_on_host_match_oracle(lib):
    """The retarded rows' per-entry device code (static rows from the
    compute_Lp_Voxels_QS closed forms, centre distances), run on the host
    through the test hook, against the oracle on a strided sample of the
    64x64x16 grid."""
    from solver_amd import _lib as L
    from oracle import rcc
    Nx, Ny, Nz, s = 64, 64, 16, (0.25e-3, 0.25e-3, 0.05e-3)
    org = np.zeros(3)
    rows = rcc.rcc_rows(org, Nx, Ny, Nz, *s)
    for k in range(9):
        for variant in (1, 2):
            ref = rows[k][variant - 1]
            idx = np.ascontiguousarray(np.arange(0, ref.size, 37, dtype=np.int64))
            out = np.empty(idx.size)
            n = ctypes.c_int64()
            rc = lib.peec_test_grid_row_host(k + 16 * variant, L.as_dp(org), Nx, Ny, Nz, *s, 1.0,
                                             idx.ctypes.data_as(L.i64p), idx.size, L.as_dp(out), ctypes.byref(n))
            assert rc == 0 and n.value == ref.size
            assert np.array_equal(out, ref[idx]), (k, variant)
