"""GPU parity of the circulant kernels, the FFT operator, the preconditioner,
GMRES and the full QS sweep against the oracle (plain FP64 path).

Tolerances: kernel rows and matvecs 1e-12 relative (norm-wise; rows also
per entry against the row's largest entry); port impedances 1e-12 relative
on the same GMRES trajectory (the GPU path takes the same iteration counts
and flags as the oracle, checked here).
"""
import numpy as np
import pytest

from oracle import fft_solver as OF
from oracle import integrals as OI
from solver_amd import configs, mesher

pytestmark = pytest.mark.gpu


def pself(sx, sy, sz):
    c = np.zeros((1, 3))
    return [OI.p_row(c[0], c, sx, sy, sz, 1, 1)[0], OI.p_row(c[0], c, sx, sy, sz, 3, 3)[0],
            OI.p_row(c[0], c, sx, sy, sz, 5, 5)[0]]


def small_mesh():
    g = np.zeros((6, 5, 4), dtype=np.int64)
    g[:, :, :2] = 2
    g[1:5, 2, 2] = 1
    g[4, 1:4, 2:4] = 1
    g[0, 0, 0] = 0
    mats = [mesher.Material("copper", 5.8e7), mesher.Material("diel", 0.0, 4.4, 0.02)]
    cell = (0.4e-3, 0.3e-3, 0.1e-3)
    ports = [((1.5 * 0.4e-3, 2.5 * 0.3e-3, 2.5e-4), (1.5 * 0.4e-3, 2.5 * 0.3e-3, 1.5e-4)),
             ((4.5 * 0.4e-3, 3.5 * 0.3e-3, 3.5e-4), (4.5 * 0.4e-3, 3.5 * 0.3e-3, 0.5e-4))]
    le = [dict(start=(0.5 * 0.4e-3, 0.5 * 0.3e-3, 1.5e-4), end=(5.5 * 0.4e-3, 4.5 * 0.3e-3, 1.5e-4),
               type=2, R=10.0, L=1e-9, C=1e-12)]
    return mesher.mesh(g, mats, cell, (1e-3, 2e-3, 0.0), ports, le, p_self_fn=pself)


def crossbar_mesh():
    cfg = configs.crossbar_24x24x4()
    return mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"], p_self_fn=pself), cfg


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


@pytest.fixture(scope="module")
def small():
    from solver_amd import sweep
    M = small_mesh()
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    S = sweep.Solver(M)
    return M, rows, spec, S


def test_kernel_rows_match_oracle(small):
    from solver_amd import peec
    M, rows, spec, S = small
    for k in range(9):
        g = peec.kernel_row(k, ctx=S.ctx)
        o = rows[k]
        assert g.shape == o.shape
        if o.size:
            err = np.max(np.abs(g - o)) / np.max(np.abs(o))
            print("kernel", k, "err", err, flush=True)
            assert err <= 1e-12


def _setfreq(M, spec, S, f):
    w = 2 * np.pi * f * M.escalings["freq"]
    S.set_frequency(w)
    op = OF.Operator(M, spec)
    Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
    op.set_frequency(w, OF.frequency_terms(M, w), Yle, order=S.order)
    return op


@pytest.mark.parametrize("f", [1e7, 3e8, 5e9])
def test_K_and_matvec_match_oracle(small, f):
    """K x to 1e-12.  M^-1 K x to 1e-12 or, where SS is ill-conditioned
    (low frequency), to cond(SS) * eps, the forward-error scale of a
    backward-stable solve: both sides factor SS in the same elimination
    order, but the oracle runs the reference's three LU solves and the GPU
    one combined solve with blocked triangular sweeps, and the rounding of
    either is amplified by cond(SS)."""
    from solver_amd import peec
    M, rows, spec, S = small
    op = _setfreq(M, spec, S, f)
    rng = np.random.default_rng(int(f) % 1000)
    N = M.m + M.ns + M.n
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    k_gpu = peec.apply_K(x, ctx=S.ctx)
    k_ora = np.concatenate(op.K(x))
    print("K", f, rel(k_gpu, k_ora), flush=True)
    assert rel(k_gpu, k_ora) < 1e-12
    y_gpu = peec.ComputeMatrixVectorFTT(x, ctx=S.ctx)
    y_ora = op.matvec(x)
    cond = np.linalg.cond(op.SS.toarray())
    tol = max(1e-12, cond * np.finfo(float).eps)
    print("M^-1 K", f, rel(y_gpu, y_ora), "cond(SS) %.2e tol %.1e" % (cond, tol), flush=True)
    assert rel(y_gpu, y_ora) < tol


def test_precond_rhs_and_gmres_match_oracle(small):
    from solver_amd import peec
    M, rows, spec, S = small
    op = _setfreq(M, spec, S, 2e8)
    is_ = np.zeros(M.n)
    is_[M.port_nodes[0, 0]] = 1e3
    is_[M.port_nodes[0, 1]] = -1e3
    tn_g = peec.precond_3_3_Kt_fft(is_, ctx=S.ctx)
    tn_o = op.precond_rhs(is_)
    assert rel(tn_g, tn_o) < 1e-12
    for tol, maxit in [(1e-8, 200), (1e-12, 200), (1e-14, 4)]:
        xg, fg, rg, ig, resg = peec.gmres_custom(tn_g, False, tol, maxit, None, ctx=S.ctx)
        xo, fo, ro, io, reso = OF.gmres_custom(tn_o, tol, maxit, None, op.matvec)
        print("gmres", tol, maxit, fg, fo, ig, io, rel(xg, xo), flush=True)
        assert fg == fo and list(ig) == list(io) and len(resg) == len(reso)
        assert rel(xg, xo) < 1e-10
        assert np.max(np.abs(resg - reso)) < 1e-10 * reso[0]


def test_gmres_stop_flag(small):
    from solver_amd import peec
    M, rows, spec, S = small
    _setfreq(M, spec, S, 2e8)
    b = np.ones(M.m + M.ns + M.n, dtype=np.complex128)
    S.ctx.request_stop(True)
    out = peec.gmres_custom(b, False, 1e-8, 10, None, ctx=S.ctx)
    S.ctx.request_stop(False)
    assert out[1] == 99


def test_crossbar_sweep_matches_oracle():
    """The 24x24x4 crossing-bar configuration, 10 frequencies 1e6..1e9 Hz,
    tol 1e-8: same GMRES flags and iteration counts, and port impedance to
    1e-12 relative or, at the low frequencies where SS is ill-conditioned,
    to the same cond(SS) * eps bound as the preconditioned matvec (the
    impedance inherits the preconditioner's rounding)."""
    from solver_amd import ordering, sweep
    M, cfg = crossbar_mesh()
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    rec_o, rec_g = [], []
    out_o = OF.fft_solver_qs(cfg["freqs"], M, spec, 300, cfg["tol"], record=rec_o,
                             order=ordering.mesh_order(M)[0])
    out_g = sweep.FFT_solver_QS_S_type(cfg["freqs"], M, 300, cfg["tol"], record=rec_g)
    for a, b in zip(rec_g, rec_o):
        assert a["flag"] == b["flag"] and a["iter"] == b["iter"], (a, b)
    Zg, Zo = out_g["Z"], out_o["Z"]
    for k, f in enumerate(cfg["freqs"]):
        w = 2 * np.pi * f * M.escalings["freq"]
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, f, 50.0)
        op = OF.Operator(M, spec)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle, F=False)
        cond = np.linalg.cond(op.SS.toarray())
        tol = max(1e-12, cond * np.finfo(float).eps)
        err = np.max(np.abs(Zg[..., k] - Zo[..., k]) / np.abs(Zo[..., k]))
        print("crossbar f %.3e Z rel err %.2e cond(SS) %.2e tol %.1e" % (f, err, cond, tol), flush=True)
        assert err < tol
        erry = np.max(np.abs(out_g["Y"][..., k] - out_o["Y"][..., k]) / np.abs(out_o["Y"][..., k]))
        assert erry < 10 * tol
