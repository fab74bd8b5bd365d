"""Conductivity gradients of port impedances, checked on the CPU with a
dense operator (the oracle's K applied to unit vectors) on a tiny mesh.

Pins, each against an independent computation:
  - dDZ/dsigma (sweep.dDZ_dsigma) against central differences of DZ_of;
  - the reciprocity similarity K^T = D1 K D2 (D1 = diag(I, j w, -I),
    D2 = diag(I, 1/(j w), -I)) on the dense K;
  - the adjoint sensitivity (sweep.impedance_sensitivity with lambda from a
    dense transposed solve) against central differences of Z_pp;
  - lambda_I = -sum_j A_jp V_j / Is (the reciprocity shortcut) against the
    transposed solve.
"""
import numpy as np
import pytest

from oracle import fft_solver as OF
from oracle import integrals as OI
from solver_amd import mesher, sweep


def pself(sx, sy, sz):
    c = np.zeros((1, 3))
    return [OI.p_row(c[0], c, sx, sy, sz, cs, cs)[0] for cs in (1, 3, 5)]


@pytest.fixture(scope="module")
def tiny():
    return tiny_mesh()


def tiny_mesh():
    g = np.zeros((6, 4, 3), dtype=np.int64)
    g[:, :, 0] = 2                      # dielectric layer
    g[:, 1:3, 1] = 1                    # copper bar along x
    g[2:4, :, 1] = 1                    # and across y
    g[1, 1, 2] = 1
    cell = (0.4e-3, 0.3e-3, 0.1e-3)
    mats = [mesher.Material("copper", 5.8e7), mesher.Material("diel", 0.0, 4.4, 0.02)]
    rng = np.random.default_rng(7)
    sig = np.where(g == 1, rng.uniform(3e7, 6e7, g.shape), 0.0)
    ports = [((0.5 * cell[0], 1.5 * cell[1], 1.5 * cell[2]), (0.5 * cell[0], 1.5 * cell[1], 0.5 * cell[2])),
             ((5.5 * cell[0], 2.5 * cell[1], 1.5 * cell[2]), (5.5 * cell[0], 2.5 * cell[1], 0.5 * cell[2]))]
    M = mesher.mesh(g, mats, cell, (0.0, 0.0, 0.0), ports, p_self_fn=pself, sigma_grid=sig)
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    return M, spec


def dense_K(M, spec, w, sigma):
    op = OF.Operator(M, spec)
    Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
    op.set_frequency(w, sweep.DZ_of(M, w, sigma), Yle)
    N = M.m + M.ns + M.n
    K = np.zeros((N, N), dtype=np.complex128)
    for i in range(N):
        e = np.zeros(N, dtype=np.complex128)
        e[i] = 1.0
        K[:, i] = np.concatenate(op.K(e))
    return K


def port_solves(M, K):
    N = K.shape[0]
    P = M.port_nodes.shape[0]
    B = np.zeros((N, P), dtype=np.complex128)
    for c, (n1, n2) in enumerate(M.port_nodes):
        B[M.m + M.ns + n1, c] = M.escalings["Is"]
        B[M.m + M.ns + n2, c] = -M.escalings["Is"]
    V = np.linalg.solve(K, B)
    S = np.zeros((P, P), dtype=np.complex128)
    phi = M.m + M.ns
    for c1 in range(P):
        for c2 in range(P):
            n3, n4 = M.port_nodes[c2]
            S[c1, c2] = (2 * (V[phi + n3, c1] - V[phi + n4, c1]) - (50.0 if c1 == c2 else 0.0)) / 50.0
    return V, S


FREQS = [2e8, 6e9]


@pytest.mark.parametrize("f", FREQS)
def test_dDZ_matches_central_differences(tiny, f):
    M, _ = tiny
    w = 2 * np.pi * f * M.escalings["freq"]
    D = sweep.dDZ_dsigma(M, w).tocsc()
    cond = np.nonzero(M.sigma_vox > 0)[0]
    def cd(v, h):
        sp_, sm = M.sigma_vox.copy(), M.sigma_vox.copy()
        sp_[v] += h
        sm[v] -= h
        return (sweep.DZ_of(M, w, sp_) - sweep.DZ_of(M, w, sm)) / (2 * h)

    for v in cond[:: max(1, len(cond) // 8)]:
        # central differences with Richardson extrapolation (DZ carries the
        # large j w Lp, so the step cannot be small)
        h = 1e-3 * M.sigma_vox[v]
        fd = (4 * cd(v, h / 2) - cd(v, h)) / 3
        an = D[:, v].toarray().ravel()
        err = np.max(np.abs(fd - an)) / max(np.max(np.abs(an)), 1e-300)
        print("voxel %d: |dDZ| %.3e, rel err vs FD %.2e" % (v, np.max(np.abs(an)), err))
        assert np.max(np.abs(an)) > 0
        assert err < 1e-7


@pytest.mark.parametrize("f", FREQS)
def test_reciprocity_similarity(tiny, f):
    M, spec = tiny
    w = 2 * np.pi * f * M.escalings["freq"]
    K = dense_K(M, spec, w, None)
    m, ns, n = M.m, M.ns, M.n
    d1 = np.concatenate([np.ones(m), np.full(ns, 1j * w), -np.ones(n)])
    d2 = np.concatenate([np.ones(m), np.full(ns, 1 / (1j * w)), -np.ones(n)])
    err = np.max(np.abs(K.T - d1[:, None] * K * d2[None, :])) / np.max(np.abs(K))
    print("K^T vs D1 K D2: %.2e" % err)
    assert err < 1e-13


@pytest.mark.parametrize("f", FREQS)
def test_adjoint_sensitivity_matches_central_differences(tiny, f):
    M, spec = tiny
    w = 2 * np.pi * f * M.escalings["freq"]
    port = 0
    K = dense_K(M, spec, w, None)
    V, S = port_solves(M, K)
    A = np.linalg.inv(np.eye(2) - S)
    gt = np.zeros(K.shape[0], dtype=np.complex128)
    gt[M.m + M.ns:] = sweep._port_vector(M, A[:, port])
    lam = np.linalg.solve(K.T, gt)
    # the reciprocity shortcut
    lam_rec = -(V[:M.m, :] @ A[:, port]) / M.escalings["Is"]
    e_rec = np.max(np.abs(lam[:M.m] - lam_rec)) / np.max(np.abs(lam[:M.m]))
    print("lambda_I: transposed solve vs reciprocity %.2e" % e_rec)
    assert e_rec < 1e-9
    z, dZ = sweep.impedance_sensitivity(M, w, S, V, lam[:M.m], port)
    cond = np.nonzero(M.sigma_vox > 0)[0]
    scale = np.max(np.abs(dZ))
    # forward sensitivity: dV_i = -K^-1 (dDZ_v .* V_i) per port, then
    # dZ = 2 R0 A dS A -- independent algebra, same result to roundoff
    D = sweep.dDZ_dsigma(M, w).tocsc()
    phi = M.m + M.ns
    worst = 0.0
    for v in cond:
        dk = D[:, v].toarray().ravel()
        R = np.zeros((K.shape[0], 2), dtype=np.complex128)
        R[:M.m, :] = dk[:, None] * V[:M.m, :]
        dV = -np.linalg.solve(K, R)
        dS = np.zeros((2, 2), dtype=np.complex128)
        for i in range(2):
            for j in range(2):
                n3, n4 = M.port_nodes[j]
                dS[i, j] = 2 * (dV[phi + n3, i] - dV[phi + n4, i]) / 50.0
        dZf = (2 * 50.0 * A @ dS @ A)[port, port]
        worst = max(worst, abs(dZf - dZ[v]) / scale)
    print("adjoint vs forward sensitivity, all %d conductor voxels: %.2e of the largest" % (len(cond), worst))
    assert worst < 1e-9

    def cd(v, h):
        zz = []
        for sgn in (1, -1):
            sg = M.sigma_vox.copy()
            sg[v] += sgn * h
            Vp, Sp = port_solves(M, dense_K(M, spec, w, sg))
            zz.append((50.0 * np.linalg.solve(np.eye(2) - Sp, np.eye(2) + Sp))[port, port])
        return (zz[0] - zz[1]) / (2 * h)

    # central differences where the sensitivity stands out of the roundoff
    # of Z (one voxel moves Z_pp by ~1e-7 of itself per percent of sigma):
    # the 3 most sensitive voxels, step 2 % Richardson-extrapolated
    for v in cond[np.argsort(-np.abs(dZ[cond]))[:3]]:
        h = 2e-2 * M.sigma_vox[v]
        fd = (4 * cd(v, h / 2) - cd(v, h)) / 3
        err = abs(fd - dZ[v]) / abs(dZ[v])
        print("voxel %d: dZ/dsigma adjoint %s FD %s, rel err %.2e" % (v, dZ[v], fd, err))
        assert err < 1e-3
