"""Benchmark: frequency-point solves per second of the voxel-PEEC FFT solver.

One step = one frequency point of the sweep: SS is factored -- on the GPU
(peec_lu_factor, the default) or with SuperLU on the host as the reference
does with UMFPACK (--lu host) -- then every port right-hand side is solved
by GMRES on the GPU (FFT convolutions, blocked triangular solves,
Householder GMRES), warm-started from the previous frequency like
FFT_solver_QS_S_type.jl:140,196.  Inputs are resident on the device before
the timed region; the one-time geometry work (integrals, spectra, operator
upload) is reported separately as setup_s.

Kernel times for the roofline come from a separate profiled pass after the
timed steps (HIP events around every launch on the context's stream), so
the timed steps carry no event overhead.

Prints one JSON line (rank 0, N = 1).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
KIND_NAMES = ["fft_x_fwd_scatter", "fft_y_fwd", "fft_z_conv", "fft_y_inv", "fft_x_inv_gather", "trsv_gather",
              "trsv_block_L", "trsv_block_U", "lu_factor"]
# kernel symbol behind each profile kind (FFT passes are instantiated per
# mode -- 0 forward, 1 inverse, 2 forward * spectrum * inverse -- and per
# line length class; the z pass of every kernel at the two-material
# workload is k_fft_pass<double, 2, 16, 128>)
KERNEL_SYMBOLS = {"fft_x_fwd_scatter": "k_fft_pass<double, 0, TPL, LMAX>", "fft_y_fwd": "k_fft_pass<double, 0, TPL, LMAX>",
                  "fft_z_conv": "k_fft_pass<double, 2, TPL, LMAX>", "fft_y_inv": "k_fft_pass<double, 1, TPL, LMAX>",
                  "fft_x_inv_gather": "k_fft_pass<double, 1, TPL, LMAX>", "trsv_gather": "k_trsv_gather",
                  "trsv_block_L": "k_trsv_block<true>", "trsv_block_U": "k_trsv_block<false>",
                  "lu_factor": "peec_lu_factor (all its kernels)"}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def build_mesh(cfg, p_self):
    from solver_amd import mesher
    return mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"], p_self_fn=p_self)


def gpu_p_self(sx, sy, sz):
    from solver_amd import peec
    c = np.zeros((1, 3))
    return [peec.compute_row_P_sup(c[0], c, sx, sy, sz, cs, cs)[0, 0] for cs in (1, 3, 5)]


def profile_read(ctx, enable):
    import ctypes
    from solver_amd import _lib as L
    nk = len(KIND_NAMES)
    ms = (ctypes.c_double * nk)()
    by = (ctypes.c_double * nk)()
    cnt = (ctypes.c_int64 * nk)()
    L.check(L.lib().peec_profile(ctx.handle, int(enable), ms, by, cnt, nk), "peec_profile")
    return [dict(name=KIND_NAMES[k], ms=ms[k], bytes=by[k], count=cnt[k]) for k in range(nk)]


def cpu_baseline(mesh, cfg, gpu_iters, sample_matvecs):
    """The oracle (numpy/scipy restatement) on the host: one LU of SS plus
    `sample_matvecs` preconditioned matvecs, extrapolated to one frequency
    point with the GPU run's average GMRES matvec count per frequency."""
    from oracle import fft_solver as OF
    t0 = time.perf_counter()
    rows = OF.kernel_rows(mesh.origin, mesh.Nx, mesh.Ny, mesh.Nz, mesh.sx, mesh.sy, mesh.sz)
    spec = OF.spectra(rows, mesh.Nx, mesh.Ny, mesh.Nz)
    t_setup = time.perf_counter() - t0
    f = cfg["freqs"][len(cfg["freqs"]) // 2]
    w = 2 * np.pi * f * mesh.escalings["freq"]
    t0 = time.perf_counter()
    Yle = OF.build_Yle(mesh.lumped_elements, mesh.port_nodes, mesh.escalings["Yle"], mesh.n, w / mesh.escalings["freq"], 50.0)
    op = OF.Operator(mesh, spec)
    op.set_frequency(w, OF.frequency_terms(mesh, w), Yle)
    t_lu = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    N = mesh.m + mesh.ns + mesh.n
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    t0 = time.perf_counter()
    for _ in range(sample_matvecs):
        x = op.matvec(x)
        x /= np.linalg.norm(x)
    t_mv = (time.perf_counter() - t0) / sample_matvecs
    per_freq = t_lu + gpu_iters * t_mv
    return dict(value=1.0 / per_freq, unit="freq-point solves/s", cores=1, kind="port",
                sample="oracle on %s: one SS assembly + LU (%.2f s) and %d preconditioned FFT matvecs "
                       "(%.3f s each), extrapolated to %.1f matvecs per frequency point (the GPU run's GMRES "
                       "average); one-time integrals + spectra %.1f s not included"
                       % (cfg["name"], t_lu, sample_matvecs, t_mv, gpu_iters, t_setup))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="two_material_64x64x16")
    ap.add_argument("--inner", type=int, default=400)
    ap.add_argument("--cpu-matvecs", type=int, default=12)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--prof-matvecs", type=int, default=20)
    ap.add_argument("--lu", choices=["device", "host"], default="device",
                    help="factor SS per frequency on the GPU (default) or with SuperLU on the host")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB device to device (a known-size "
                         "reference dispatch for scaling FETCH_SIZE / WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: every box in this setup has a single MI355X, and the sweep shards "
              "by frequency across GPUs (see DESIGN.md); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from solver_amd import configs, sweep
    cfg = configs.CONFIGS[args.workload]()
    t0 = time.perf_counter()
    mesh = build_mesh(cfg, gpu_p_self)
    t_mesh = time.perf_counter() - t0
    log("mesh %s: n=%d m=%d ns=%d N=%d ports=%d (%.1f s)" % (cfg["name"], mesh.n, mesh.m, mesh.ns,
                                                              mesh.m + mesh.ns + mesh.n, len(mesh.port_nodes), t_mesh))
    t0 = time.perf_counter()
    solver = sweep.Solver(mesh, device_lu=args.lu == "device")
    t_setup = time.perf_counter() - t0
    log("setup (integrals, spectra, operator upload): %.2f s" % t_setup)

    esc = mesh.escalings
    freqs = np.asarray(cfg["freqs"]) * esc["freq"]
    P = mesh.port_nodes.shape[0]
    N = mesh.m + mesh.ns + mesh.n
    Vrest = np.zeros((N, P), dtype=np.complex128)
    from solver_amd import peec
    iters = []

    def step(k, record):
        w = 2 * np.pi * freqs[k % len(freqs)]
        solver.set_frequency(w)
        is_ = np.zeros(mesh.n)
        for c1 in range(P):
            n1, n2 = mesh.port_nodes[c1]
            is_[:] = 0
            is_[n1] = esc["Is"]
            is_[n2] = -esc["Is"]
            tn = peec.precond_3_3_Kt_fft(is_, ctx=solver.ctx)
            V, flag, relres, it, resvec = peec.gmres_custom(tn, False, cfg["tol"], args.inner, Vrest[:, c1],
                                                            ctx=solver.ctx)
            Vrest[:, c1] = V
            if record:
                iters.append(len(resvec) - 1 + 2)  # + residual check and precond_rhs
                log("  f=%.3e Hz port %d flag %d iter %s relres %.2e" % (freqs[k % len(freqs)] / esc["freq"], c1,
                                                                       flag, it, relres))

    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    for k in range(args.warmup):
        step(k, False)
    sync()
    t0 = time.perf_counter()
    for k in range(args.warmup, args.warmup + args.steps):
        step(k, True)
        log("step %d done" % k)
    sync()
    dt = time.perf_counter() - t0
    value = args.steps / dt

    # profiled pass: preconditioned matvecs at the last frequency
    x = np.ascontiguousarray(Vrest[:, 0])
    peec.ComputeMatrixVectorFTT(x, ctx=solver.ctx)
    profile_read(solver.ctx, True)
    for _ in range(args.prof_matvecs):
        peec.ComputeMatrixVectorFTT(x, ctx=solver.ctx)
    prof = profile_read(solver.ctx, False)
    for r in prof:
        r["avg_us"] = 1e3 * r["ms"] / max(r["count"], 1)
        r["gbs"] = r["bytes"] / max(r["ms"], 1e-12) / 1e6
        log("  kernel %-18s launches %6d  time %8.3f ms  avg %8.2f us  bytes %.3e  %.1f GB/s" % (
            r["name"], r["count"], r["ms"], r["avg_us"], r["bytes"], r["gbs"]))
    # the roofline is for the matvec's kernels (the profiled pass runs no LU)
    dom = max((r for r in prof if r["name"] != "lu_factor"), key=lambda r: r["ms"])
    ach = dom["gbs"]
    mv_ms = sum(r["ms"] for r in prof) / args.prof_matvecs
    fft_ms = sum(r["ms"] for r in prof if r["name"].startswith("fft")) / args.prof_matvecs
    fft_bytes = sum(r["bytes"] for r in prof if r["name"].startswith("fft")) / args.prof_matvecs

    cpu = None
    if not args.no_cpu_baseline:
        per_step = float(np.sum(iters)) / args.steps if iters else 50.0
        cpu = cpu_baseline(mesh, cfg, per_step, args.cpu_matvecs)

    out = {
        "metric": "freq-point solves/sec at 1e-6 rel. impedance err; FFT-matvec HBM GB/s",
        "value": value,
        "unit": "freq-point solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic (seeded geometry, %s)" % cfg["name"],
        "config": {"workload": cfg["name"], "grid": list(cfg["grid"].shape), "unknowns": int(N),
                   "ports": int(P), "gmres_tol": cfg["tol"], "path": "plain FP64",
                   "setup_s": round(t_setup, 3), "matvec_kernel_ms": round(mv_ms, 4),
                   "fft_matvec_ms": round(fft_ms, 4), "fft_matvec_GBs": round(fft_bytes / fft_ms / 1e6, 1),
                   "gmres_matvecs_per_step": float(np.sum(iters)) / args.steps if iters else None},
        "roofline": {"bound": "hbm", "kernel": dom["name"], "symbol": KERNEL_SYMBOLS[dom["name"]], "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": ach / PEAK_HBM_GBS, "traffic": None, "avg_launch_us": dom["avg_us"],
                     "bytes_per_launch": dom["bytes"] / max(dom["count"], 1)},
        "kernels": {r["name"]: {"launches_per_matvec": r["count"] / args.prof_matvecs, "avg_us": round(r["avg_us"], 3),
                                "GBs": round(r["gbs"], 1)} for r in prof},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    if args.copy_mib > 0:
        import torch
        a = torch.empty(args.copy_mib << 20, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        log("copied %d MiB device to device 3 times" % args.copy_mib)


if __name__ == "__main__":
    main()
