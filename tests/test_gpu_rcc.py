"""GPU parity of the retarded (Rcc, QS_Rcc_FW = 2) kernels against the
oracle (oracle/rcc.py).

- The static rows of the retarded path (the Lp rows from
  compute_Lp_Voxels_QS.jl's forms, the P rows) and the centre distances,
  built on the device from the grid: bitwise equal to the oracle's on the
  24x24x4 and 64x64x16 grids (the latter reaches the real-log
  parallel-surface form of the QS copy).
- K x with the retarded spectra at 1e8, 5e9 and 2e10 Hz: the complex
  phasors are the only device/host libm difference (sincos), so K x agrees
  to 1e-12, including the half-cell-reduced off-diagonal P blocks, whose
  two directions have separate spectra.
- The 24x24x4 crossing-bar sweep with QS_Rcc_FW = 2 at 4 frequencies: the
  same GMRES flags and iteration counts as the oracle's retarded sweep, and
  Z within max(1e-9, cond(SS) eps).  The quasi-static sweep holds 1e-12;
  the retarded operator near its first resonance (the 12 mm crossbar at
  5 GHz is a fifth of a wavelength) amplifies the 1e-16 matvec rounding
  through the Krylov recurrence more (measured 6.4e-11), still two orders
  below the GMRES tolerance 1e-8.  At 2e10 Hz the retarded and
  quasi-static impedances differ by more than 1e-4, so the test sees the
  retardation.
"""
import numpy as np
import pytest

from oracle import fft_solver as OF
from oracle import rcc
from test_gpu_operator import crossbar_mesh, rel, small_mesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("grid", [(24, 24, 4, 0.5e-3, 0.5e-3, 0.1e-3), (64, 64, 16, 0.25e-3, 0.25e-3, 0.05e-3)])
def test_device_rcc_rows_and_distances_match_oracle(grid):
    from solver_amd import peec
    Nx, Ny, Nz, sx, sy, sz = grid
    origin = (0.0, 0.0, 0.0)
    with peec.Context(0) as ctx:
        o = np.zeros(3)
        peec.compute_rows_Rcc(o, Nx, Ny, Nz, sx, sy, sz, dict(Lp=1e6, P=1e-12), ctx=ctx)
        rows = rcc.rcc_rows(origin, Nx, Ny, Nz, sx, sy, sz)
        for k in range(9):
            g_row = peec.kernel_row(k, ctx=ctx, variant=1)
            g_dist = peec.kernel_row(k, ctx=ctx, variant=2)
            o_row, o_dist = rows[k]
            same_row = np.array_equal(g_row, o_row)
            same_dist = np.array_equal(g_dist, o_dist)
            print("grid %s kernel %d: rows bitwise %s, distances bitwise %s" % (grid[:3], k, same_row, same_dist),
                  flush=True)
            assert same_row and same_dist


@pytest.fixture(scope="module")
def small_rcc():
    from solver_amd import peec, sweep
    M = small_mesh()
    rows = rcc.rcc_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    ctx = peec.Context(0)
    S = sweep.Solver(M, ctx=ctx, QS_Rcc_FW=2)
    yield M, rows, S
    ctx.close()


@pytest.mark.parametrize("f", [1e8, 5e9, 2e10])
def test_rcc_K_matches_oracle(small_rcc, f):
    from solver_amd import peec
    M, rows, S = small_rcc
    w = 2 * np.pi * f * M.escalings["freq"]
    S.set_frequency(w)
    spec = rcc.rcc_spectra(rows, M.Nx, M.Ny, M.Nz, w / (2 * np.pi * M.escalings["freq"]), M.escalings["Lp"],
                           M.escalings["P"])
    op = OF.Operator(M, spec)
    Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
    op.set_frequency(w, OF.frequency_terms(M, w), Yle, order=S.order)
    rng = np.random.default_rng(int(f) % 997)
    N = M.m + M.ns + M.n
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    k_gpu = peec.apply_K(x, ctx=S.ctx)
    k_ora = np.concatenate(op.K(x))
    # the quasi-static operator differs (the test sees the retardation)
    static = OF.Operator(M, OF.spectra(OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz), M.Nx, M.Ny,
                                       M.Nz))
    static.set_frequency(w, OF.frequency_terms(M, w), Yle, F=False)
    d_static = rel(np.concatenate(static.K(x)), k_ora)
    print("Rcc K at %.1e Hz: GPU vs oracle %.2e; retarded vs quasi-static %.2e" % (f, rel(k_gpu, k_ora), d_static),
          flush=True)
    assert rel(k_gpu, k_ora) < 1e-12
    assert d_static > 1e-9


def test_rcc_crossbar_sweep_matches_oracle():
    from solver_amd import ordering, peec, sweep
    M, cfg = crossbar_mesh()
    freqs = np.asarray([1e7, 5e8, 5e9, 2e10])
    rows = rcc.rcc_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec_of = lambda f: rcc.rcc_spectra(rows, M.Nx, M.Ny, M.Nz, f, M.escalings["Lp"], M.escalings["P"])
    rec_o, rec_g = [], []
    out_o = OF.fft_solver_qs(freqs, M, spec_of, 300, 1e-8, record=rec_o, order=ordering.mesh_order(M)[0])
    with peec.Context(0) as ctx:
        S = sweep.Solver(M, ctx=ctx, QS_Rcc_FW=2)
        out_g = sweep.FFT_solver_QS_S_type(freqs, M, 300, 1e-8, solver=S, record=rec_g)
        Sq = sweep.Solver(M, ctx=ctx)
        out_q = sweep.FFT_solver_QS_S_type(freqs[-1:], M, 300, 1e-8, solver=Sq)
    for a, b in zip(rec_g, rec_o):
        assert a["flag"] == b["flag"] and a["iter"] == b["iter"], (a, b)
    for k, f in enumerate(freqs):
        w = 2 * np.pi * f * M.escalings["freq"]
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, f, 50.0)
        op = OF.Operator(M, None)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle, F=False)
        cond = np.linalg.cond(op.SS.toarray())
        tol = max(1e-9, cond * np.finfo(float).eps)
        err = np.max(np.abs(out_g["Z"][..., k] - out_o["Z"][..., k]) / np.abs(out_o["Z"][..., k]))
        print("Rcc crossbar f %.1e: Z rel err %.2e (bar %.1e)" % (f, err, tol), flush=True)
        assert err < tol
    dq = np.max(np.abs(out_q["Z"][..., 0] - out_g["Z"][..., -1]) / np.abs(out_g["Z"][..., -1]))
    print("Rcc vs quasi-static Z at %.1e Hz: %.2e" % (freqs[-1], dq), flush=True)
    assert dq > 1e-4
