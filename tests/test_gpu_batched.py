"""Batched frequency x port solves (peec_gmres_batched) against the oracle.

The batched path runs restarted CGS2 GMRES inside FP64 iterative
refinement, so it does not follow the reference's iteration sequence; its
impedances are checked against the oracle's sweep at tight tolerances:
plain FP64 to 1e-8 relative (both solved to a 1e-12 residual), mixed
precision (FP32 FFT convolutions inside the cycles) to the 1e-6 relative
impedance error the mixed path is held to."""
import numpy as np
import pytest

from oracle import fft_solver as OF
from solver_amd import peec, sweep

from test_gpu_operator import crossbar_mesh, rel, small_mesh

pytestmark = pytest.mark.gpu


def test_batched_matvec_matches_single():
    """M^-1 K applied to 3 slots x 2 ports at once equals the one-slot
    matvecs (same factors: a slot factored on the device at the same w)."""
    M = small_mesh()
    S = sweep.Solver(M, device_lu=True)
    fs = [1e8, 1e9, 3e9]
    ws = [2 * np.pi * f * M.escalings["freq"] for f in fs]
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(5)
    X = rng.normal(size=(N, 6)) + 1j * rng.normal(size=(N, 6))
    single = []
    for w in ws:
        S.set_frequency(w)
        single.append(np.stack([peec.ComputeMatrixVectorFTT(X[:, len(single) * 2 + c], ctx=S.ctx)
                                for c in range(2)], axis=1))
    peec.frequency_slots(3, ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
        peec.lu_factor_slot(slot, ctx=S.ctx)
    # the batched path applies A only inside GMRES: solve A x = b with
    # b = A x from the one-slot matvecs and check that it returns x (every
    # column against its own slot's operator)
    B = np.concatenate(single, axis=1)
    Xs, flag, relres, iters, cycles = peec.gmres_batched(B, 2, 1e-11, None, restart=80, max_cycles=10, ctx=S.ctx)
    print("batched solve of A x: flags", flag, "relres", relres, "iters", iters, flush=True)
    assert np.all(relres < 1e-10)
    for c in range(6):
        print("column", c, "rel err", rel(Xs[:, c], X[:, c]), flush=True)
        assert rel(Xs[:, c], X[:, c]) < 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_batched_crossbar_sweep(mixed):
    M, cfg = crossbar_mesh()
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    out_o = OF.fft_solver_qs(cfg["freqs"], M, spec, 400, 1e-12)
    rec = []
    out_b = sweep.FFT_solver_QS_batched(cfg["freqs"], M, 1e-12 if not mixed else 1e-10, batch=4, mixed=mixed,
                                        record=rec)
    for r in rec:
        print(r, flush=True)
    assert all(r["flag"] == 0 for r in rec)
    err = np.max(np.abs(out_b["Z"] - out_o["Z"]) / np.abs(out_o["Z"]))
    print("batched (mixed=%s) crossbar Z max rel err %.3e" % (mixed, err), flush=True)
    assert err < (1e-6 if mixed else 1e-8)


@pytest.mark.parametrize("mixed", [False, True])
def test_device_sweep_matches_host_driven(mixed):
    """peec_sweep_batch (excitations, preconditioned right-hand sides,
    solves and the warm start chained across batches, all on the device)
    against the same sweep driven from the host (keep=True: right-hand
    sides and solutions cross PCIe): identical inputs to the same GMRES, so
    identical impedances and final warm start."""
    M = small_mesh()
    S = sweep.Solver(M, device_lu=True)
    fs = np.logspace(7, 10, 6)
    kw = dict(batch=3, mixed=mixed, restart=40, solver=S)
    rec_h, rec_d = [], []
    host = sweep.FFT_solver_QS_batched(fs, M, 1e-10, keep=True, record=rec_h, **kw)
    dev = sweep.FFT_solver_QS_batched(fs, M, 1e-10, record=rec_d, **kw)
    e_z = float(np.max(np.abs(dev["Z"] - host["Z"]) / np.abs(host["Z"])))
    e_v = rel(dev["Vrest"], host["Vrest"])
    print("mixed=%s: Z device vs host %.2e (bitwise %s), Vrest %.2e; iterations %s vs %s"
          % (mixed, e_z, np.array_equal(dev["Z"], host["Z"]), e_v, [r["iters"] for r in rec_d],
             [r["iters"] for r in rec_h]), flush=True)
    assert [r["iters"] for r in rec_d] == [r["iters"] for r in rec_h]
    assert all(r["flag"] == 0 for r in rec_d)
    assert e_z < 1e-13 and e_v < 1e-13
    # a chained second call continues from the kept warm start
    dev2 = sweep.FFT_solver_QS_batched(fs[:3], M, 1e-10, Vrest=dev["Vrest"], **kw)
    assert np.all(np.isfinite(dev2["Z"]))


def test_device_edge_impedances_match_host():
    """DZ computed on the device (peec_frequency_set_slot_edges) against
    DZ_of on the host, over frequencies on both sides of the skin-effect
    switch and with per-voxel conductivities."""
    from solver_amd import configs, mesher
    M = small_mesh()
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, sg = configs._layered(12, 10, 6, cell, 4, 0.5, 4.4, 0.02, seed=2, sigma_range=(3e7, 6e7))
    M2 = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), configs._x_face_ports(g, cell, 4), p_self_fn=pself_fn(),
                     sigma_grid=sg)
    S2 = sweep.Solver(M2, ctx=peec.Context(0), device_lu=True)
    for mesh, solver, sigma in ((M, S, None), (M2, S2, M2.sigma_vox * 1.3)):
        fs = np.logspace(5, 10.5, 8)
        peec.frequency_slots(len(fs), ctx=solver.ctx)
        sweep.set_slots(solver, [2 * np.pi * f * mesh.escalings["freq"] for f in fs], sigma=sigma)
        worst, exact = 0.0, 0
        for slot, f in enumerate(fs):
            w = 2 * np.pi * f * mesh.escalings["freq"]
            ref = sweep.DZ_of(mesh, w, sigma)
            dz = peec.frequency_dz(slot, mesh.m, ctx=solver.ctx)
            worst = max(worst, float(np.max(np.abs(dz - ref) / np.abs(ref))))
            exact += int(np.sum(dz == ref))
        print("m=%d: device DZ vs DZ_of max rel %.2e, %d of %d bitwise" % (mesh.m, worst, exact, mesh.m * len(fs)),
              flush=True)
        assert worst < 1e-15


def pself_fn():
    from test_gpu_operator import pself
    return pself


def test_batched_lu_equals_per_slot():
    """peec_lu_factor_slots (all slots through each launch of the
    multifrontal schedule together) gives the same factors, bit for bit,
    as factoring the slots one at a time: M^-1 x on every slot agrees.
    Both apply the factors through the level-scheduled solves (the
    multifrontal solves need every slot factored at once)."""
    M = small_mesh()
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    peec.lu_set_solve(0, ctx=S.ctx)
    fs = [1e7, 3e8, 2e9, 9e9]
    ws = [2 * np.pi * f * M.escalings["freq"] for f in fs]
    peec.frequency_slots(len(ws), ctx=S.ctx)
    for slot, w in enumerate(ws):
        Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
        peec.frequency_set_slot(slot, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(17)
    X = rng.normal(size=(N, 8)) + 1j * rng.normal(size=(N, 8))
    for slot in range(len(ws)):
        peec.lu_factor_slot(slot, ctx=S.ctx)
    one = peec.apply_K_batched(X, 2, precond=2, ctx=S.ctx)
    peec.lu_factor_slots(0, len(ws), ctx=S.ctx)
    together = peec.apply_K_batched(X, 2, precond=2, ctx=S.ctx)
    print("batched LU vs per-slot: bitwise %s (rel %.2e)" % (np.array_equal(one, together), rel(together, one)),
          flush=True)
    assert np.array_equal(one, together)


def test_batched_gmres_nonfinite_and_zero_rhs():
    """A NaN right-hand side comes back flagged 4 (not converged, relres
    NaN) without touching the other columns; a zero right-hand side returns
    the zero solution with flag 0 even from a non-zero start, as
    gmres_custom does (gmres_custom.jl:20-28)."""
    M = small_mesh()
    S = sweep.Solver(M, device_lu=True)
    w = 2 * np.pi * 1e9 * M.escalings["freq"]
    peec.frequency_slots(1, ctx=S.ctx)
    Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
    peec.frequency_set_slot(0, w, sweep.DZ_of(M, w), Yle, ctx=S.ctx)
    peec.lu_factor_slot(0, ctx=S.ctx)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(9)
    B = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    B[5, 1] = np.nan
    B[:, 2] = 0.0
    X0 = np.zeros((N, 3), dtype=np.complex128)
    X0[:, 2] = 1.0
    ref, fl_ref, _, _, _ = peec.gmres_batched(B[:, :1], 3, 1e-10, None, restart=60, ctx=S.ctx)
    X, flag, relres, iters, cycles = peec.gmres_batched(B, 3, 1e-10, X0, restart=60, ctx=S.ctx)
    print("flags", flag, "relres", relres, "iters", iters, flush=True)
    assert list(flag) == [0, 4, 0]
    assert np.isnan(relres[1]) and relres[2] == 0.0 and relres[0] < 1e-10
    assert np.all(X[:, 2] == 0)
    assert rel(X[:, 0], ref[:, 0]) < 1e-12
