"""Device factorization of SS (peec_lu_analyze / peec_lu_factor) against
SuperLU in the same elimination order, and the solver built on it against
the oracle.

Both factor SS[order, order] without pivoting, so the factors agree to
rounding (checked entry-wise, 1e-10 relative to the largest entry of each
factor column block); the preconditioned matvec and the sweep are held to
the same max(1e-12, cond(SS) * eps) bar as the host-factor path."""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

from oracle import fft_solver as OF
from solver_amd import ordering, sweep

from test_gpu_operator import crossbar_mesh, rel, small_mesh

pytestmark = pytest.mark.gpu


def test_device_lu_matches_superlu_small():
    from solver_amd import peec
    M = small_mesh()
    S = sweep.Solver(M, device_lu=True)
    w = 2 * np.pi * 3e8 * M.escalings["freq"]
    S.set_frequency(w)
    Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
    SS = sweep.SS_of(M, w, sweep.DZ_of(M, w), Yle)
    D = S.dissection
    Lf, Uf, p, q = sweep.lu_factor(SS, D.order)
    Ld, Ud = Lf.toarray(), Uf.toarray()
    n = M.n
    for K in range(len(D.front_ptr) - 1):
        s, pos, F = peec.lu_front(K, ctx=S.ctx)
        first = D.front_ptr[K]
        assert s == D.front_ptr[K + 1] - first and np.array_equal(pos[:s], np.arange(first, first + s))
        # L: unit lower in the pivot block, then the boundary rows; U: the
        # pivot rows over all columns of the front
        Lk = np.tril(F[:, :s], -1)
        Lk[:s, :s] += np.eye(s)
        Uk = np.triu(F[:s, :])
        ref_L = Ld[np.ix_(pos, pos[:s])]
        ref_U = Ud[np.ix_(pos[:s], pos)]
        eL = np.abs(Lk - ref_L).max() / max(np.abs(ref_L).max(), 1e-300)
        eU = np.abs(Uk - ref_U).max() / max(np.abs(ref_U).max(), 1e-300)
        print("front", K, "f", F.shape[0], "s", s, "L err %.2e U err %.2e" % (eL, eU), flush=True)
        assert eL <= 1e-10 and eU <= 1e-10
    # nothing of SuperLU's factors lies outside the fronts
    covered = np.zeros((n, n), dtype=bool)
    for K in range(len(D.front_ptr) - 1):
        s, pos, F = peec.lu_front(K, ctx=S.ctx)
        covered[np.ix_(pos, pos[:s])] = True
        covered[np.ix_(pos[:s], pos)] = True
    assert not np.any((np.abs(Ld) + np.abs(Ud) > 0) & ~covered)
    assert n == D.order.size


@pytest.mark.parametrize("f", [1e7, 3e8, 5e9])
def test_device_lu_matvec_matches_oracle(f):
    from solver_amd import peec
    M = small_mesh()
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    S = sweep.Solver(M, device_lu=True)
    w = 2 * np.pi * f * M.escalings["freq"]
    S.set_frequency(w)
    op = OF.Operator(M, spec)
    Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
    op.set_frequency(w, OF.frequency_terms(M, w), Yle, order=S.order)
    rng = np.random.default_rng(7)
    N = M.m + M.ns + M.n
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    y_gpu = peec.ComputeMatrixVectorFTT(x, ctx=S.ctx)
    y_ora = op.matvec(x)
    cond = np.linalg.cond(op.SS.toarray())
    tol = max(1e-12, cond * np.finfo(float).eps)
    print("device LU: M^-1 K", f, rel(y_gpu, y_ora), "cond(SS) %.2e tol %.1e" % (cond, tol), flush=True)
    assert rel(y_gpu, y_ora) < tol


def test_device_lu_crossbar_sweep_matches_oracle():
    M, cfg = crossbar_mesh()
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    order = ordering.mesh_order(M)[0]
    rec_o, rec_g = [], []
    out_o = OF.fft_solver_qs(cfg["freqs"], M, spec, 300, cfg["tol"], record=rec_o, order=order)
    S = sweep.Solver(M, device_lu=True)
    out_g = sweep.FFT_solver_QS_S_type(cfg["freqs"], M, 300, cfg["tol"], solver=S, record=rec_g)
    for a, b in zip(rec_g, rec_o):
        assert a["flag"] == b["flag"] and a["iter"] == b["iter"], (a, b)
    for k, f in enumerate(cfg["freqs"]):
        w = 2 * np.pi * f * M.escalings["freq"]
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, f, 50.0)
        op = OF.Operator(M, spec)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle, F=False)
        cond = np.linalg.cond(op.SS.toarray())
        tol = max(1e-12, cond * np.finfo(float).eps)
        err = np.max(np.abs(out_g["Z"][..., k] - out_o["Z"][..., k]) / np.abs(out_o["Z"][..., k]))
        print("device LU crossbar f %.3e Z rel err %.2e tol %.1e" % (f, err, tol), flush=True)
        assert err < tol


@pytest.mark.parametrize("rps,fs", [(9, [2e8, 4e9]), (3, [2e8, 4e9]), (1, [1e8, 6e8, 2e9, 7e9])])
def test_panel_sums_match_oracle_preconditioner(rps, fs):
    """With the device factors every U block is a panel block (its rows are
    one front's pivot rows and reach the same columns), so the backward
    sweep sums them as panels (profile kind 10) and the forward sweep's
    rows through the per-row gathers (kind 5).  M^-1 x over len(fs) slots x
    rps right-hand sides -- 9: a pass of 8 and a remainder of 1 per slot;
    3: one pass of the 4-wide variant; 1 (4 slots): the 1-wide variant --
    matches the oracle's precond_3_3_vector column by column."""
    import ctypes
    from solver_amd import _lib, configs, mesher, peec
    from test_gpu_operator import pself
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(12, 10, 6, cell, 4, 0.45, 4.4, 0.02, seed=3)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), configs._x_face_ports(g, cell, 4), p_self_fn=pself)
    S = sweep.Solver(M, device_lu=True)  # 526 nodes in 16 fronts
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    ws = [2 * np.pi * f * M.escalings["freq"] for f in fs]
    peec.frequency_slots(len(ws), ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
        peec.lu_factor_slot(slot, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(17)
    X = rng.normal(size=(N, rps * len(ws))) + 1j * rng.normal(size=(N, rps * len(ws)))
    nk = 11
    ms, by, cnt = (ctypes.c_double * nk)(), (ctypes.c_double * nk)(), (ctypes.c_int64 * nk)()
    _lib.check(_lib.lib().peec_profile(S.ctx.handle, 1, ms, by, cnt, nk), "peec_profile")
    Y = peec.apply_K_batched(X, rps, precond=2, ctx=S.ctx)
    _lib.check(_lib.lib().peec_profile(S.ctx.handle, 0, ms, by, cnt, nk), "peec_profile")
    print("launches per kind", list(cnt), flush=True)
    assert cnt[10] > 0 and cnt[5] > 0
    m, ns = M.m, M.ns
    for slot, w in enumerate(ws):
        op = OF.Operator(M, spec)
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle, order=S.order)
        cond = np.linalg.cond(op.SS.toarray())
        tol = max(1e-12, cond * np.finfo(float).eps)
        for c in range(slot * rps, (slot + 1) * rps):
            x = X[:, c]
            e = rel(Y[:, c], op.precond(x[:m], x[m:m + ns], x[m + ns:]))
            print("slot %d column %d: M^-1 x vs oracle %.2e (tol %.1e)" % (slot, c, e, tol), flush=True)
            assert e < tol


def test_zero_or_nonfinite_pivot_is_an_error():
    """The device LU does not pivot: a non-finite SS entry (here a NaN
    edge impedance) must come back as PEEC_ERR_NUMERIC, not as factors."""
    from solver_amd import _lib, peec
    M = small_mesh()
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    w = 2 * np.pi * 3e8 * M.escalings["freq"]
    Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
    DZ = sweep.DZ_of(M, w)
    DZ[0] = np.nan
    peec.frequency_slots(2, ctx=S.ctx)
    peec.frequency_set_slot(0, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    peec.frequency_set_slot(1, w, DZ, Yle, ctx=S.ctx)
    with pytest.raises(_lib.PeecError, match="pivot"):
        peec.lu_factor_slots(0, 2, ctx=S.ctx)
    peec.lu_factor_slot(0, ctx=S.ctx)  # the good slot alone still factors
    S.ctx.close()


def test_umfpack_convention_factors_with_pivoting_and_row_scaling():
    """peec_lu_set_factors takes F = lu(SS) in UMFPACK's convention, which
    Julia exposes as F.L, F.U, F.p, F.q, F.Rs: L * U = (Rs .* SS)[p, q].
    Here SuperLU factors Rs .* SS with partial pivoting (p != q) and a
    row scaling Rs far from 1; M^-1 x from those factors must match the
    oracle's precond_3_3_vector (its own unpivoted LU) to rounding times
    the conditioning of SS."""
    from solver_amd import peec
    M = small_mesh()
    S = sweep.Solver(M, ctx=peec.Context(0), nested_dissection=False, device_lu=False)
    for f in (3e7, 2e9):
        w = 2 * np.pi * f * M.escalings["freq"]
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        DZ = sweep.DZ_of(M, w)
        peec.frequency_set(w, DZ, Yle, ctx=S.ctx)
        SS = sweep.SS_of(M, w, DZ, Yle).tocsc()
        rng = np.random.default_rng(4)
        Rs = 10.0 ** rng.uniform(-3, 3, SS.shape[0])
        A = SS.multiply(Rs[:, None]).tocsc()
        F = spla.splu(A, permc_spec="COLAMD", diag_pivot_thresh=1.0)
        p, q = np.argsort(F.perm_r), np.argsort(F.perm_c)
        assert not np.array_equal(p, q)
        LU = (F.L @ F.U).toarray()
        assert np.abs(LU - A.toarray()[np.ix_(p, q)]).max() <= 1e-12 * np.abs(A).max()
        peec.lu_set_factors(F.L, F.U, p, q, Rs=Rs, ctx=S.ctx)
        N = M.m + M.ns + M.n
        x = rng.normal(size=N) + 1j * rng.normal(size=N)
        got = peec.apply_K_batched(x[:, None], 1, precond=2, ctx=S.ctx)[:, 0]
        op = OF.Operator(M, None)
        op.set_frequency(w, OF.frequency_terms(M, w), OF.build_Yle(M.lumped_elements, M.port_nodes,
                                                                  M.escalings["Yle"], M.n, w / M.escalings["freq"],
                                                                  50.0))
        ref = op.precond(x[:M.m], x[M.m:M.m + M.ns], x[M.m + M.ns:])
        cond = np.linalg.cond(SS.toarray())
        e = rel(got, ref)
        print("f %.3g: M^-1 x with pivoted, row-scaled factors vs oracle %.2e (cond(SS) %.1e)" % (f, e, cond),
              flush=True)
        assert e <= max(1e-12, 10 * cond * np.finfo(float).eps)
    S.ctx.close()


def _layered_solver(nx, ny, nz, seed=3):
    from solver_amd import configs, mesher
    from test_gpu_operator import pself
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = configs._layered(nx, ny, nz, cell, 4, 0.45, 4.4, 0.02, seed=seed)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), configs._x_face_ports(g, cell, 4), p_self_fn=pself)
    return M, sweep.Solver(M, device_lu=True)


def _set_slots(M, S, fs):
    from solver_amd import peec
    ws = [2 * np.pi * f * M.escalings["freq"] for f in fs]
    peec.frequency_slots(len(ws), ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    return ws


def _profiled_precond(S, X, rps):
    """M^-1 X and the launch count per profile kind (11: multifrontal
    small fronts, 12: GEMV tasks, 13: assembly / finishing)."""
    import ctypes
    from solver_amd import _lib, peec
    nk = 14
    ms, by, cnt = (ctypes.c_double * nk)(), (ctypes.c_double * nk)(), (ctypes.c_int64 * nk)()
    _lib.check(_lib.lib().peec_profile(S.ctx.handle, 1, ms, by, cnt, nk), "peec_profile")
    Y = peec.apply_K_batched(X, rps, precond=2, ctx=S.ctx)
    _lib.check(_lib.lib().peec_profile(S.ctx.handle, 0, ms, by, cnt, nk), "peec_profile")
    return Y, list(cnt)


@pytest.mark.parametrize("nxyz,rps,fs", [((12, 10, 6), 9, [2e8, 4e9]), ((12, 10, 6), 3, [2e8, 4e9, 1e7]),
                                         ((32, 30, 8), 8, [1e7, 3e9]), ((32, 30, 8), 1, [5e8])])
def test_multifrontal_solve_matches_oracle(nxyz, rps, fs):
    """The multifrontal solves (every slot factored at once, so the front
    pool holds them all with their pivot-block inverses): M^-1 x per slot
    matches the oracle's precond_3_3_vector column by column, to the
    max(1e-12, cond(SS) eps) bar of the level-scheduled solves.  rps 9: a
    pass of 8 right-hand sides and a remainder of 1 per slot; 3: the 4-wide
    kernels; 1: the 1-wide ones.  12 x 10 x 6 has only small fronts (one
    fused launch per tree height); 32 x 30 x 8 has separators of more than
    64 pivots, whose pivots go through the GEMV tasks (and their
    L11^-1 / U11^-1 from the block-row inversion)."""
    from solver_amd import peec
    M, S = _layered_solver(*nxyz)
    ws = _set_slots(M, S, fs)
    peec.lu_factor_slots(0, len(ws), ctx=S.ctx)
    info = peec.lu_solve_info(ctx=S.ctx)
    print("solve info", info, flush=True)
    assert info["multifrontal"] == 1
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(23)
    X = rng.normal(size=(N, rps * len(ws))) + 1j * rng.normal(size=(N, rps * len(ws)))
    Y, cnt = _profiled_precond(S, X, rps)
    print("launches per kind", cnt, flush=True)
    assert cnt[11] > 0 and cnt[5] == 0 and cnt[10] == 0
    if nxyz[0] > 12:
        assert cnt[12] > 0 and cnt[13] > 0
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    m, ns = M.m, M.ns
    for slot, w in enumerate(ws):
        op = OF.Operator(M, spec)
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
        op.set_frequency(w, OF.frequency_terms(M, w), Yle, order=S.order)
        # 1-norm condition estimate from the oracle's factors; SS is complex
        # symmetric, so SS^-H v = conj(SS^-1 conj(v))
        F = op.F
        inv = spla.LinearOperator(op.SS.shape, matvec=lambda v: F.solve(np.asarray(v, dtype=complex)),
                                  rmatvec=lambda v: np.conj(F.solve(np.conj(np.asarray(v, dtype=complex)))),
                                  dtype=complex)
        cond = spla.onenormest(op.SS) * spla.onenormest(inv)
        tol = max(1e-12, cond * np.finfo(float).eps)
        for c in range(slot * rps, (slot + 1) * rps):
            x = X[:, c]
            e = rel(Y[:, c], op.precond(x[:m], x[m:m + ns], x[m + ns:]))
            print("slot %d column %d: M^-1 x vs oracle %.2e (tol %.1e)" % (slot, c, e, tol), flush=True)
            assert e < tol


def test_multifrontal_matches_level_solve():
    """The same factors through both solves (mode 0: level-scheduled
    blocks, mode 1: multifrontal): M^-1 x agrees to 1e-11 relative, and the
    multifrontal solve gives bitwise the same result when run again."""
    from solver_amd import peec
    M, S = _layered_solver(32, 30, 8, seed=5)
    ws = _set_slots(M, S, [3e7, 1e9, 8e9])
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(29)
    X = rng.normal(size=(N, 8 * len(ws))) + 1j * rng.normal(size=(N, 8 * len(ws)))
    peec.lu_set_solve(0, ctx=S.ctx)
    peec.lu_factor_slots(0, len(ws), ctx=S.ctx)
    assert peec.lu_solve_info(ctx=S.ctx)["multifrontal"] == 0
    Y0 = peec.apply_K_batched(X, 8, precond=2, ctx=S.ctx)
    peec.lu_set_solve(1, ctx=S.ctx)
    assert peec.lu_solve_info(ctx=S.ctx)["multifrontal"] == 0  # needs a factorization with the inverses
    peec.lu_factor_slots(0, len(ws), ctx=S.ctx)
    assert peec.lu_solve_info(ctx=S.ctx)["multifrontal"] == 1
    Y1 = peec.apply_K_batched(X, 8, precond=2, ctx=S.ctx)
    Y2 = peec.apply_K_batched(X, 8, precond=2, ctx=S.ctx)
    e = max(rel(Y1[:, c], Y0[:, c]) for c in range(X.shape[1]))
    print("multifrontal vs level-scheduled: max column rel %.2e; repeat bitwise %s" % (e, np.array_equal(Y1, Y2)),
          flush=True)
    assert np.array_equal(Y1, Y2)
    assert e < 1e-11
