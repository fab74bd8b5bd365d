"""FFT-convolution launch shapes, the fused y-z pass and the autotuner.

Every launch shape of the LDS pass kernel (lines per tile, lanes per line,
right-hand sides per block, hence also the prefetching of the next
right-hand side) runs the same butterflies in the same order,
so K x must come out bit for bit the same under each of them, in complex128
and in complex64.  The complex64 fused pipeline (x pass into planes, one y
forward / z convolution / y inverse pass per plane, x pass back), which the
tuner may choose, runs the same stages again: it must equal the five-pass
path bit for bit, also with the 48- and 64-point z kernels.  The complex64
radix schedules change rounding only: they stay within complex64 accuracy
of the complex128 result.  The complex128 result itself is held to the
oracle's K x (1e-12, relative).  The lane-split passes (a line spread over
Q lanes, cross-lane butterflies by DPP and swizzles or through LDS) compute
the same transforms in another operation order: complex128 within 1e-13 of the LDS passes and 1e-12 of
the oracle, complex64 within 1e-5 of complex128.
"""
import numpy as np
import pytest

from oracle import fft_solver as OF
from solver_amd import _lib, configs, mesher, peec, sweep

from test_gpu_operator import pself, rel

pytestmark = pytest.mark.gpu

NRHS = 3
N_VARIANTS = 12  # 11 LDS tile shapes + the register z kernel (fused z pass only)
LANE_BASE, N_LANE = 12, 5  # lane-split passes with 2, 4, .. 32 lanes per line (y and z passes)
LANE_XL_BASE = LANE_BASE + N_LANE  # the same with the exchanges through LDS, 4 and 8 lanes per line
PASS_YZ = 6      # tuning-table pass of the fused y-z pass: variant 0 fused, 1 five passes


@pytest.fixture(scope="module")
def batched():
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(20, 14, 6, cell, 4, 0.45, 4.4, 0.02, seed=3)
    ports = configs._x_face_ports(g, cell, 4)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), ports, p_self_fn=pself)
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    ws = [2 * np.pi * f * M.escalings["freq"] for f in (3e7, 1e9, 8e9)]
    peec.frequency_slots(NRHS, ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(11)
    X = rng.normal(size=(N, NRHS)) + 1j * rng.normal(size=(N, NRHS))
    dims = sorted({n for k in range(9) for n in OF.kernel_dims(k, M.Nx, M.Ny, M.Nz)})
    print("mesh n=%d m=%d ns=%d; compact kernel dims %s" % (M.n, M.m, M.ns, dims), flush=True)
    return M, S, ws, X


def bitwise_equal(a, b):
    return np.array_equal(np.ascontiguousarray(a).view(np.float64), np.ascontiguousarray(b).view(np.float64))


def set_all(S, fp32, variant, rpb, style, nrhs=NRHS):
    """Force one launch shape on every one of the five passes where it fits
    (and, in complex64, the five-pass path instead of the fused one);
    returns how many passes took it."""
    took = 0
    for k in range(9):
        for p in range(5):
            try:
                peec.fft_set_choice(k, p, fp32, nrhs, variant, rpb, style, ctx=S.ctx)
                took += 1
            except _lib.PeecError:
                peec.fft_set_choice(k, p, fp32, nrhs, -1, ctx=S.ctx)
        if fp32:
            set_fused(S, k, False, nrhs=nrhs)
    return took


def set_fused(S, k, on, rpb=1, nrhs=NRHS):
    """Fused y-z pipeline on or off for kernel k (complex64); False when it
    does not apply."""
    try:
        peec.fft_set_choice(k, PASS_YZ, True, nrhs, 0 if on else 1, rpb, 0, ctx=S.ctx)
        return True
    except _lib.PeecError:
        return False


def clear_all(S, nrhs=NRHS):
    for fp32 in (False, True):
        for k in range(9):
            for p in range(8 if fp32 else 5):
                peec.fft_set_choice(k, p, fp32, nrhs, -1, ctx=S.ctx)


def test_default_shapes_match_oracle(batched):
    M, S, ws, X = batched
    clear_all(S)
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    Y = peec.apply_K_batched(X, 1, ctx=S.ctx)
    for r, w in enumerate(ws):
        op = OF.Operator(M, spec)
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle)
        e = rel(Y[:, r], np.concatenate(op.K(X[:, r])))
        print("slot %d: K x vs oracle %.2e" % (r, e), flush=True)
        assert e < 1e-12


@pytest.mark.parametrize("fp32", [False, True])
def test_every_launch_shape_is_bitwise_identical(batched, fp32):
    M, S, ws, X = batched
    clear_all(S)
    ref = peec.apply_K_batched(X, 1, fp32=fp32, ctx=S.ctx)
    assert np.all(np.isfinite(ref))
    for v in range(N_VARIANTS):
        for rpb in (1, 2, 3):
            took = set_all(S, fp32, v, rpb, 0)
            if took == 0:
                continue
            Y = peec.apply_K_batched(X, 1, fp32=fp32, ctx=S.ctx)
            same = bitwise_equal(Y, ref)
            print("fp32=%s variant %d rpb %d on %d passes: bitwise %s (rel diff %.2e)" % (fp32, v, rpb, took, same,
                                                                                      rel(Y, ref)), flush=True)
            assert same
    clear_all(S)


def test_fp32_radix_schedules_and_tuner(batched):
    M, S, ws, X = batched
    clear_all(S)
    y64 = peec.apply_K_batched(X, 1, fp32=False, ctx=S.ctx)
    for style in (0, 1, 2):
        assert set_all(S, True, 8, 1, style) > 0
        y32 = peec.apply_K_batched(X, 1, fp32=True, ctx=S.ctx)
        e = rel(y32, y64)
        print("complex64 radix schedule %d: rel err vs complex128 %.2e" % (style, e), flush=True)
        assert e < 1e-5
    clear_all(S)
    ms = peec.fft_tune(NRHS, ctx=S.ctx)
    print("tuned: summed best pass time %.3f ms" % ms, flush=True)
    for k in range(9):
        print("kernel %d fp64" % k, [peec.fft_choice(k, p, False, NRHS, ctx=S.ctx) for p in range(5)],
              "fp32", [peec.fft_choice(k, p, True, NRHS, ctx=S.ctx) for p in range(5)], flush=True)
    # the tuner may pick lane-split passes, whose butterflies round differently
    y64t = peec.apply_K_batched(X, 1, fp32=False, ctx=S.ctx)
    assert rel(y64t, y64) < 1e-13
    y32t = peec.apply_K_batched(X, 1, fp32=True, ctx=S.ctx)
    assert rel(y32t, y64) < 1e-5


def fused_profile(S, X, rps):
    """K x in complex64 with the per-kind launch counts of that one call."""
    nk = 10
    import ctypes
    ms, by, cnt = (ctypes.c_double * nk)(), (ctypes.c_double * nk)(), (ctypes.c_int64 * nk)()
    _lib.check(_lib.lib().peec_profile(S.ctx.handle, 1, ms, by, cnt, nk), "peec_profile")
    Y = peec.apply_K_batched(X, rps, fp32=True, ctx=S.ctx)
    _lib.check(_lib.lib().peec_profile(S.ctx.handle, 0, ms, by, cnt, nk), "peec_profile")
    return Y, list(cnt)


def test_fused_yz_matches_five_passes(batched):
    M, S, ws, X = batched
    clear_all(S)
    five, cnt5 = fused_profile(S, X, 1)
    assert cnt5[9] == 0 and cnt5[1] > 0  # untuned: the five passes
    for k in range(9):
        assert set_fused(S, k, True)
    fused, cnt = fused_profile(S, X, 1)
    print("fused complex64 launches per kind", cnt, flush=True)
    assert cnt[9] > 0 and cnt[1] == 0  # fused y-z passes ran, no separate y pass
    print("fused vs five passes: bitwise %s (rel diff %.2e)" % (bitwise_equal(fused, five), rel(fused, five)),
          flush=True)
    assert bitwise_equal(fused, five)
    for rpb in (1, 2, 3):
        for k in range(9):
            set_fused(S, k, True, rpb)
        assert bitwise_equal(peec.apply_K_batched(X, 1, fp32=True, ctx=S.ctx), five)
    clear_all(S)


def test_fused_yz_long_z_lines():
    """24 z cells: the 48- and 64-point register z kernels inside the fused
    pass (real and complex spectra), 16 right-hand sides over 2 slots."""
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(12, 10, 24, cell, 18, 0.5, 4.4, 0.02, seed=5)
    ports = configs._x_face_ports(g, cell, 18)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), ports, p_self_fn=pself)
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    nrhs, rps = 16, 8
    peec.frequency_slots(2, ctx=S.ctx)
    for slot, f in enumerate((1e8, 5e9)):
        w = 2 * np.pi * f * M.escalings["freq"]
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(23)
    X = rng.normal(size=(N, nrhs)) + 1j * rng.normal(size=(N, nrhs))
    y64 = peec.apply_K_batched(X, rps, fp32=False, ctx=S.ctx)
    five = peec.apply_K_batched(X, rps, fp32=True, ctx=S.ctx)
    for k in range(9):
        assert set_fused(S, k, True, nrhs=nrhs)
    fused, cnt = fused_profile(S, X, rps)
    assert cnt[9] > 0
    print("nz 24: fused vs five passes bitwise %s (rel diff %.2e); vs complex128 %.2e"
          % (bitwise_equal(fused, five), rel(fused, five), rel(fused, y64)), flush=True)
    assert bitwise_equal(fused, five)
    assert rel(fused, y64) < 1e-5
    ms = peec.fft_tune(nrhs, ctx=S.ctx)
    fused_on = [peec.fft_choice(k, PASS_YZ, True, nrhs, ctx=S.ctx)[0] for k in range(9)]
    print("tuned %.3f ms; fused choice per kernel %s" % (ms, fused_on), flush=True)
    # the tuner may pick other complex64 radix schedules: rounding changes
    assert rel(peec.apply_K_batched(X, rps, fp32=True, ctx=S.ctx), y64) < 1e-5


@pytest.fixture(scope="module")
def lanes_geometry():
    """20 x 64 x 24: 128-point y lines and 48-/64-point z lines, the lengths
    the lane-split kernels are built for."""
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(20, 64, 24, cell, 18, 0.45, 4.4, 0.02, seed=7)
    ports = configs._x_face_ports(g, cell, 18)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), ports, p_self_fn=pself)
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    ws = [2 * np.pi * f * M.escalings["freq"] for f in (3e7, 8e9)]
    peec.frequency_slots(2, ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(19)
    X = rng.normal(size=(N, 4)) + 1j * rng.normal(size=(N, 4))
    return M, S, ws, X


def set_lanes(S, fp32, Q, rpb, nrhs, xl=False):
    """Force the lane-split kernel with Q lanes per line (xl: exchanges
    through LDS) on the y and z passes of every kernel where it is built;
    returns (passes taken, z passes taken)."""
    took = tz = 0
    v = LANE_XL_BASE + int(np.log2(Q)) - 2 if xl else LANE_BASE + int(np.log2(Q)) - 1
    for k in range(9):
        if fp32:
            set_fused(S, k, False, nrhs=nrhs)
        for p in (1, 2, 3):
            try:
                peec.fft_set_choice(k, p, fp32, nrhs, v, rpb, 0, ctx=S.ctx)
                took += 1
                tz += p == 2
            except _lib.PeecError:
                peec.fft_set_choice(k, p, fp32, nrhs, -1, ctx=S.ctx)
    return took, tz


@pytest.mark.parametrize("fp32", [False, True])
def test_lane_split_passes_match(lanes_geometry, fp32):
    M, S, ws, X = lanes_geometry
    nrhs, rps = X.shape[1], 2
    clear_all(S, nrhs)
    y64 = peec.apply_K_batched(X, rps, fp32=False, ctx=S.ctx)
    ref = peec.apply_K_batched(X, rps, fp32=fp32, ctx=S.ctx)
    total = 0
    for Q, xl in ((4, False), (8, False), (16, False), (4, True), (8, True)):
        for rpb in (1, 3):
            took, tz = set_lanes(S, fp32, Q, rpb, nrhs, xl)
            if took == 0:
                continue
            total += took
            Y = peec.apply_K_batched(X, rps, fp32=fp32, ctx=S.ctx)
            e, e64 = rel(Y, ref), rel(Y, y64)
            print("fp32=%s Q %d%s rpb %d: lane passes on %d passes (%d z); vs LDS passes %.2e, vs complex128 %.2e"
                  % (fp32, Q, " (LDS exchanges)" if xl else "", rpb, took, tz, e, e64), flush=True)
            assert e64 < (1e-5 if fp32 else 1e-13)
        clear_all(S, nrhs)
    assert total > 0
    if not fp32:
        rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
        spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
        for Q, xl in ((8, False), (4, True)):
            set_lanes(S, False, Q, 1, nrhs, xl)
            Y = peec.apply_K_batched(X, rps, fp32=False, ctx=S.ctx)
            for r in range(nrhs):
                w = ws[r // rps]
                op = OF.Operator(M, spec)
                Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"],
                                   50.0)
                op.set_frequency(w, OF.frequency_terms(M, w), Yle)
                e = rel(Y[:, r], np.concatenate(op.K(X[:, r])))
                print("lane passes Q %d%s, column %d: K x vs oracle %.2e" % (Q, " (LDS exchanges)" if xl else "", r, e),
                      flush=True)
                assert e < 1e-12
            clear_all(S, nrhs)


@pytest.fixture(scope="module")
def xlanes_geometry():
    """96 x 10 x 6: 192-point x lines, where the x passes can run the
    lane-split kernel with their scatter and gather maps."""
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(96, 10, 6, cell, 4, 0.45, 4.4, 0.02, seed=11)
    ports = configs._x_face_ports(g, cell, 4)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), ports, p_self_fn=pself)
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    ws = [2 * np.pi * f * M.escalings["freq"] for f in (3e7, 8e9)]
    peec.frequency_slots(2, ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(23)
    X = rng.normal(size=(N, 4)) + 1j * rng.normal(size=(N, 4))
    return M, S, ws, X


def test_x_pass_lanes_match(xlanes_geometry):
    """Every lane-split x pass built for 192 points (scatter MODE 0 and
    gather MODE 1, lane exchanges and LDS exchanges, 1 and 3 right-hand
    sides per block) against complex128 and the oracle's K x."""
    M, S, ws, X = xlanes_geometry
    nrhs, rps = X.shape[1], 2
    clear_all(S, nrhs)
    y64 = peec.apply_K_batched(X, rps, fp32=False, ctx=S.ctx)
    ref = peec.apply_K_batched(X, rps, fp32=True, ctx=S.ctx)
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    op = OF.Operator(M, OF.spectra(rows, M.Nx, M.Ny, M.Nz))
    yo = np.empty_like(X)
    for r in range(nrhs):
        w = ws[r // rps]
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle)
        yo[:, r] = np.concatenate(op.K(X[:, r]))
    assert rel(y64, yo) < 1e-12
    total = 0
    for Q, xl in ((8, False), (16, False), (8, True)):
        v = LANE_XL_BASE + int(np.log2(Q)) - 2 if xl else LANE_BASE + int(np.log2(Q)) - 1
        for rpb in (1, 3):
            took = 0
            for k in range(9):
                set_fused(S, k, False, nrhs=nrhs)
                for p in (0, 4):
                    try:
                        peec.fft_set_choice(k, p, True, nrhs, v, rpb, 0, ctx=S.ctx)
                        took += 1
                    except _lib.PeecError:
                        peec.fft_set_choice(k, p, True, nrhs, -1, ctx=S.ctx)
            if took == 0:
                continue
            total += took
            Y = peec.apply_K_batched(X, rps, fp32=True, ctx=S.ctx)
            e, e64, eo = rel(Y, ref), rel(Y, y64), rel(Y, yo)
            print("x passes Q %d%s rpb %d on %d passes: vs LDS x passes %.2e, vs complex128 %.2e, vs oracle %.2e"
                  % (Q, " (LDS exchanges)" if xl else "", rpb, took, e, e64, eo), flush=True)
            assert e64 < 1e-5 and eo < 1e-5
        clear_all(S, nrhs)
    assert total >= 12
