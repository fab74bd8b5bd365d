"""The adjoint path on the GPU: K^H, M^-H, the adjoint GMRES and the
conductivity gradient of a port impedance.

  - <y, K x> = <K^H y, x> and <y, M^-1 x> = <M^-H y, x> per frequency slot
    (an adjoint with a wrong conjugate, sign or block fails these);
  - the adjoint GMRES solution lambda against the reciprocity shortcut
    lambda_I = -sum_j A_jp V_j / Is, computed from the forward solutions;
  - the gradient of sum_f |Z_pp|^2 against the dense CPU computation of the
    same gradient (tests/test_gradient_cpu.py pins that one by central
    differences and by the forward sensitivity), and against central
    differences of the plain FP64 path (gmres_custom, one port at a time).
"""
import numpy as np
import pytest

from oracle import fft_solver as OF
from solver_amd import configs, mesher, peec, sweep

import test_gradient_cpu as TC
from test_gpu_operator import pself

pytestmark = pytest.mark.gpu


def inner(a, b):
    return np.sum(np.conj(a) * b, axis=0)


@pytest.fixture(scope="module")
def layered():
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, sig = configs._layered(20, 14, 6, cell, 4, 0.45, 4.4, 0.02, seed=3, sigma_range=(3e7, 6e7))
    ports = configs._x_face_ports(g, cell, 4)
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), ports, p_self_fn=pself, sigma_grid=sig)
    # a context of its own: other tests of this module use the default one
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    return M, S


def test_adjoint_identities(layered):
    M, S = layered
    fs = [3e7, 1e9, 8e9]
    ws = 2 * np.pi * np.array(fs) * M.escalings["freq"]
    peec.frequency_slots(len(fs), ctx=S.ctx)
    sweep.set_slots(S, ws)
    N = M.m + M.ns + M.n
    rng = np.random.default_rng(3)
    X = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    Y = rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))
    # mode 0: K against K^H; mode 2: M^-1 against M^-H
    for mode, name in ((0, "K"), (2, "M^-1")):
        AX = peec.apply_K_batched(X, 1, precond=mode, ctx=S.ctx)
        AHY = peec.apply_K_batched(Y, 1, precond=mode, adjoint=True, ctx=S.ctx)
        a, b = inner(Y, AX), inner(AHY, X)
        err = np.max(np.abs(a - b) / np.abs(a))
        print("%s: <y, A x> vs <A^H y, x> per slot, max rel diff %.2e" % (name, err), flush=True)
        assert err < 1e-12
    # the adjoint GMRES operator M^-H K^H is the composition of the two
    B = peec.apply_K_batched(Y, 1, precond=True, adjoint=True, ctx=S.ctx)
    C = peec.apply_K_batched(peec.apply_K_batched(Y, 1, adjoint=True, ctx=S.ctx), 1, precond=2, adjoint=True,
                             ctx=S.ctx)
    assert np.linalg.norm(B - C) / np.linalg.norm(C) < 1e-14


def test_adjoint_gmres_matches_reciprocity(layered):
    M, S = layered
    fs = np.array([5e7, 2e9])
    ws = 2 * np.pi * fs * M.escalings["freq"]
    fw = sweep.FFT_solver_QS_batched(fs, M, 1e-11, batch=2, mixed=False, solver=S, keep=True)
    P = M.port_nodes.shape[0]
    peec.frequency_slots(len(fs), ctx=S.ctx)
    sweep.set_slots(S, ws)
    port = 0
    Am = [np.linalg.inv(np.eye(P) - fw["S"][:, :, k]) for k in range(len(fs))]
    G = np.stack([np.conj(sweep._port_vector(M, Am[k][:, port])) for k in range(len(fs))], axis=1)
    TN = peec.precond_rhs_batched_adjoint(G, 1, ctx=S.ctx)
    eta, flag, relres, it, cyc = peec.gmres_batched(TN, 1, 1e-10, None, 80, 20, False, ctx=S.ctx, adjoint=True)
    print("adjoint GMRES: flags", flag, "relres", relres, "iterations", it, flush=True)
    assert np.all(flag == 0)
    # K^H eta = conj(g) in the node block, zero elsewhere
    R = peec.apply_K_batched(eta, 1, adjoint=True, ctx=S.ctx)
    rhs = np.zeros_like(R)
    rhs[M.m + M.ns:, :] = G
    print("||K^H eta - rhs|| / ||rhs||:", np.linalg.norm(R - rhs, axis=0) / np.linalg.norm(rhs, axis=0), flush=True)
    for k in range(len(fs)):
        lam = np.conj(eta[:M.m, k])
        lam_rec = -(fw["V"][k][:M.m, :] @ Am[k][:, port]) / M.escalings["Is"]
        e = np.linalg.norm(lam - lam_rec) / np.linalg.norm(lam_rec)
        print("f=%.3g: adjoint solve vs reciprocity %.2e" % (fs[k], e), flush=True)
        assert e < 1e-7


def test_gradient_matches_dense_cpu_gradient():
    """The tiny mesh of the CPU gradient tests: the GPU adjoint gradient of
    sum_f |Z_11|^2 equals the dense CPU one."""
    M, spec = TC.tiny_mesh()
    fs = np.array(TC.FREQS)
    out = sweep.impedance_objective_gradient(fs, M, port=0, tol=1e-12, batch=2, method="adjoint")
    ref = np.zeros(M.n)
    Jref = 0.0
    for f in fs:
        w = 2 * np.pi * f * M.escalings["freq"]
        K = TC.dense_K(M, spec, w, None)
        V, S = TC.port_solves(M, K)
        A = np.linalg.inv(np.eye(2) - S)
        gt = np.zeros(K.shape[0], dtype=np.complex128)
        gt[M.m + M.ns:] = sweep._port_vector(M, A[:, 0])
        lam = np.linalg.solve(K.T, gt)
        z, dZ = sweep.impedance_sensitivity(M, w, S, V, lam[:M.m], 0)
        Jref += abs(z) ** 2
        ref += 2 * np.real(np.conj(z) * dZ)
    e = np.max(np.abs(out["grad"] - ref)) / np.max(np.abs(ref))
    print("J %.12e vs %.12e; gradient max err / max %.2e" % (out["J"], Jref, e), flush=True)
    assert abs(out["J"] - Jref) / Jref < 1e-9
    assert e < 1e-7


def test_gradient_matches_plain_path_central_differences(layered):
    """Central differences need an impedance that the conductivity moves:
    the 24 x 24 x 4 crossing bars (port from end face to end face of the
    x-bar, Z ~ R + j w L of the bar) with per-voxel sigma.  On the layered
    geometry, whose ports close through the dielectric, one voxel moves
    sum |Z_11|^2 by ~1e-13 of itself per percent of sigma -- below what any
    iterative solve resolves -- so there it is checked against the
    reciprocity gradient only."""
    M2, S2 = layered
    fs2 = np.array([1e8, 3e9])
    out2 = sweep.impedance_objective_gradient(fs2, M2, port=0, tol=1e-12, batch=2, solver=S2, method="adjoint")
    rec2 = sweep.impedance_objective_gradient(fs2, M2, port=0, tol=1e-12, batch=2, solver=S2, method="reciprocity")
    e_rec = np.max(np.abs(out2["grad"] - rec2["grad"])) / np.max(np.abs(out2["grad"]))
    print("layered: adjoint vs reciprocity gradient %.2e" % e_rec, flush=True)
    assert e_rec < 1e-7

    cfg = configs.crossbar_24x24x4()
    rng = np.random.default_rng(5)
    sig = np.where(cfg["grid"] == 1, rng.uniform(3e7, 6e7, cfg["grid"].shape), 0.0)
    M = mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"], p_self_fn=pself,
                    sigma_grid=sig)
    S = sweep.Solver(M, ctx=peec.Context(0), device_lu=True)
    fs = np.array([1e7, 1e9])
    out = sweep.impedance_objective_gradient(fs, M, port=0, tol=1e-12, batch=2, solver=S, method="adjoint")
    pla = sweep.impedance_objective_gradient(fs, M, port=0, tol=1e-12, solver=S, method="plain")
    g = out["grad"]
    e_pla = np.max(np.abs(g - pla["grad"])) / np.max(np.abs(g))
    print("crossbar: J %.10e (plain %.10e); adjoint vs plain-path gradient %.2e" % (out["J"], pla["J"], e_pla),
          flush=True)
    assert e_pla < 1e-7
    cond = np.nonzero(M.sigma_vox > 0)[0]
    top = cond[np.argsort(-np.abs(g[cond]))[:3]]

    def cd(v, h):
        J = []
        for sgn in (1, -1):
            sg = M.sigma_vox.copy()
            sg[v] += sgn * h
            J.append(sweep.impedance_objective(fs, M, port=0, sigma=sg, tol=1e-12, solver=S)[0])
        return (J[0] - J[1]) / (2 * h)

    for v in top:
        h = 1e-3 * M.sigma_vox[v]
        fd = cd(v, h)
        err = abs(fd - g[v]) / abs(g[v])
        print("voxel %d: dJ/dsigma adjoint %.6e, plain-path central difference %.6e, rel err %.2e"
              % (v, g[v], fd, err), flush=True)
        assert err < 1e-3
