"""Conductivity gradient of a port-impedance objective on the adjoint
workload: the 64 x 64 x 16 two-material geometry with per-voxel copper
conductivity drawn uniform in [3e7, 6e7] S/m from the seed, objective
J = sum over 32 frequencies (1e7..1e10 Hz, log-spaced) of |Z_11|^2.

  1. the full dJ/dsigma by the adjoint method (batched FP64 forward solves
     of both ports, one adjoint GMRES per frequency on M^-H K^H), timed;
  2. the same gradient from the reciprocity shortcut (no adjoint solve),
     and the plain FP64 path's gradient (forward solutions from
     gmres_custom, one port at a time -- the reference's own solve);
  3. central differences with relative step 1e-4 at 16 seeded conductor
     voxels, each J from the batched FP64 path -- or, with --fd-path plain,
     from the plain FP64 path.

Prints one JSON line with the timings and the agreement.

    python -m solver_amd.gradcheck [--nfreq 32] [--nvox 16] [--tol 1e-12]
"""
import argparse
import json
import sys
import time
from types import SimpleNamespace

import numpy as np

from . import configs, mesher, peec, sweep


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nfreq", type=int, default=32)
    ap.add_argument("--nvox", type=int, default=16)
    ap.add_argument("--tol", type=float, default=1e-12)
    ap.add_argument("--step", type=float, default=1e-4, help="relative central-difference step")
    ap.add_argument("--fd-path", choices=["plain", "batched"], default="batched")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    print(json.dumps(run(a.nfreq, a.nvox, a.tol, a.step, a.fd_path, a.seed)), flush=True)


def run(nfreq=32, nvox=16, tol=1e-12, step=1e-4, fd_path="batched", seed=0, ctx=None):
    """The check above; returns its record (timings, agreements, and per
    checked voxel the adjoint and central-difference derivatives)."""
    a = SimpleNamespace(nfreq=nfreq, nvox=nvox, tol=tol, step=step, fd_path=fd_path, seed=seed)

    def log(*x):
        print(*x, file=sys.stderr, flush=True)

    cfg = configs.adjoint_64x64x16(a.seed)

    def pself(sx, sy, sz):
        c = np.zeros((1, 3))
        return [peec.compute_row_P_sup(c[0], c, sx, sy, sz, cs, cs)[0, 0] for cs in (1, 3, 5)]

    M = mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"], p_self_fn=pself,
                    sigma_grid=cfg["sigma"])
    fs = np.asarray(cfg["freqs"])[np.unique(np.linspace(0, len(cfg["freqs"]) - 1, a.nfreq).round().astype(int))]
    log("adjoint workload: N=%d, %d conductor voxels, %d frequencies" % (M.m + M.ns + M.n,
                                                                        int(np.sum(M.sigma_vox > 0)), len(fs)))
    t0 = time.perf_counter()
    S = sweep.Solver(M, ctx=ctx, device_lu=True)
    t_setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    adj = sweep.impedance_objective_gradient(fs, M, port=0, tol=a.tol, solver=S, method="adjoint")
    t_adj = time.perf_counter() - t0
    t0 = time.perf_counter()
    rec = sweep.impedance_objective_gradient(fs, M, port=0, tol=a.tol, solver=S, method="reciprocity")
    t_rec = time.perf_counter() - t0
    t0 = time.perf_counter()
    pla = sweep.impedance_objective_gradient(fs, M, port=0, tol=a.tol, solver=S, method="plain")
    t_pla = time.perf_counter() - t0
    g = adj["grad"]
    e_rec = float(np.max(np.abs(g - rec["grad"])) / np.max(np.abs(g)))
    e_pla = float(np.max(np.abs(g - pla["grad"])) / np.max(np.abs(g)))
    log("plain-path gradient %.2f s, max diff / max %.2e; J plain %.12e" % (t_pla, e_pla, pla["J"]))
    log("J = %.12e; adjoint gradient %.2f s (flags %s), reciprocity %.2f s, max diff / max %.2e"
        % (adj["J"], t_adj, sorted(set(adj["adjoint_flags"])), t_rec, e_rec))
    cond = np.nonzero(M.sigma_vox > 0)[0]
    rng = np.random.default_rng(a.seed + 16)
    vox = np.sort(rng.choice(cond, size=min(a.nvox, cond.size), replace=False))
    gmax = float(np.max(np.abs(g)))
    rows = []
    t0 = time.perf_counter()
    for v in vox:
        h = a.step * M.sigma_vox[v]
        J = []
        for sgn in (1, -1):
            sg = M.sigma_vox.copy()
            sg[v] += sgn * h
            if a.fd_path == "plain":
                J.append(sweep.impedance_objective(fs, M, port=0, sigma=sg, tol=a.tol, solver=S)[0])
            else:
                out = sweep.FFT_solver_QS_batched(fs, M, a.tol, mixed=False, solver=S, sigma=sg)
                J.append(float(np.sum(np.abs(out["Z"][0, 0, :]) ** 2)))
        fd = (J[0] - J[1]) / (2 * h)
        rows.append(dict(voxel=int(v), adjoint=float(g[v]), fd=float(fd), err_rel=float(abs(fd - g[v]) / abs(g[v])),
                         err_of_max=float(abs(fd - g[v]) / gmax)))
        log("voxel %6d: adjoint %+.6e  FD %+.6e  rel %.2e  of max %.2e" % (v, g[v], fd, rows[-1]["err_rel"],
                                                                          rows[-1]["err_of_max"]))
    t_fd = time.perf_counter() - t0
    return dict(workload=cfg["name"], unknowns=int(M.m + M.ns + M.n), frequencies=len(fs),
                conductor_voxels=int(cond.size), J=adj["J"], setup_s=t_setup, adjoint_gradient_s=t_adj,
                reciprocity_gradient_s=t_rec, adjoint_vs_reciprocity=e_rec, plain_gradient_s=t_pla,
                adjoint_vs_plain=e_pla, J_plain=pla["J"], fd_path=a.fd_path,
                fd_step=a.step, fd_s=t_fd, fd=rows, adjoint_flags=sorted(set(adj["adjoint_flags"])),
                fd_max_err_of_max=max(r["err_of_max"] for r in rows) if rows else None)


if __name__ == "__main__":
    main()
