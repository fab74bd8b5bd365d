"""Write the sweep fixtures: port voltages, S and Z of the oracle's
quasi-static FFT solver (FFT_solver_QS_S_type.jl:60-232) on the seeded
benchmark geometries meshed by the reference mesher (configs.mesh), at
three frequencies each.

Run from the repository root, in the build container:

    python tests/golden/make_sweeps.py two_material   # ~30 min, 1 process
    python tests/golden/make_sweeps.py batched [j]    # ~1 h, j processes (default 4)

two_material (64x64x16, 2 ports): the reference's plain FP64 sweep at
1e7, 3.0e8 and 1e10 Hz (the nearest points of its 32-point band), GMRES
tolerance 1e-10, warm-started from the previous frequency's solution per
port (Vrest, :140,196), LU of SS in the nested-dissection order the GPU
path uses.  Records Z, S, the port voltage differences, and each solve's
flag, iteration count and final relative residual: the GPU plain path runs
the same Krylov iteration and is held to these.

batched (96x96x24, 8 ports): the same solver, run tightly (tolerance 1e-12)
at 1e6, 1.0e8 and 2e10 Hz (first, nearest, last of the 128-point band), to
hold the batched mixed-precision path's per-port impedances.  Ports are
split over j worker processes; each solves its ports' rows of S over the
three frequencies with its own warm-start chain.

Output: sweep_<config>.npz (numpy arrays, no pickles) beside this script.
"""
import os
import sys
import time
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fft_solver as OF  # noqa: E402
from oracle import integrals as OI  # noqa: E402
from solver_amd import configs, ordering  # noqa: E402

R0 = 50.0
SETS = {
    "two_material": dict(config="two_material_64x64x16", targets=(1e7, 3e8, 1e10), tol=1e-10, inner=400),
    "batched": dict(config="batched_96x96x24", targets=(1e6, 1e8, 2e10), tol=1e-12, inner=300),
}


def p_self(sx, sy, sz):
    c = np.zeros((1, 3))
    return [OI.p_row(c[0], c, sx, sy, sz, cs, cs)[0] for cs in (1, 3, 5)]


def setup(name):
    s = SETS[name]
    cfg = configs.CONFIGS[s["config"]]()
    M = configs.mesh(cfg, p_self_fn=p_self)
    fall = np.asarray(cfg["freqs"])
    idx = sorted({int(np.argmin(np.abs(np.log(fall / t)))) for t in s["targets"]})
    return s, cfg, M, fall[idx], np.asarray(idx)


def solve_ports(args):
    """Rows c1 in `ports` of the port-voltage matrix at every frequency:
    Vp[c1, c2, k] = V[m+ns+n3] - V[m+ns+n4] of the solution for port c1
    (FFT_solver_QS_S_type.jl:121-207), with the per-port warm start."""
    name, ports = args
    s, cfg, M, freqs, _ = setup(name)
    esc = M.escalings
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    order = ordering.mesh_order(M)[0]
    op = OF.Operator(M, spec)
    m, n, ns = M.m, M.n, M.ns
    P = M.port_nodes.shape[0]
    Vp = np.zeros((P, P, len(freqs)), dtype=np.complex128)
    rec = np.zeros((P, len(freqs), 3))
    Vrest = np.zeros((m + ns + n, P), dtype=np.complex128)
    is_ = np.zeros(n)
    for k, f in enumerate(freqs):
        w = 2 * np.pi * f * esc["freq"]
        Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, esc["Yle"], n, w / esc["freq"], R0)
        t0 = time.time()
        op.set_frequency(w, OF.frequency_terms(M, w), Yle, order=order)
        for c1 in ports:
            n1, n2 = M.port_nodes[c1]
            is_[n1], is_[n2] = esc["Is"], -esc["Is"]
            tn = op.precond_rhs(is_)
            V, flag, relres, it, resvec = OF.gmres_custom(tn, s["tol"], s["inner"], Vrest[:, c1], op.matvec)
            is_[n1] = is_[n2] = 0
            Vrest[:, c1] = V
            rec[c1, k] = (flag, it[1], relres)
            for c2 in range(P):
                n3, n4 = M.port_nodes[c2]
                Vp[c1, c2, k] = V[m + ns + n3] - V[m + ns + n4]
            print("%s f=%.4g port %d: flag %d, %d iterations, relres %.2e (%.0f s)"
                  % (name, f, c1, flag, it[1], relres, time.time() - t0), flush=True)
    return ports, Vp, rec


def s_of(Vp):
    """S from the port voltage differences (FFT_solver_QS_S_type.jl:199-207)."""
    P = Vp.shape[0]
    S = 2 * Vp / R0
    for c in range(P):
        S[c, c] -= 1.0
    return S


def make(name, jobs):
    s, cfg, M, freqs, idx = setup(name)
    P = M.port_nodes.shape[0]
    groups = [list(range(P))[i::jobs] for i in range(jobs)] if jobs > 1 else [list(range(P))]
    groups = [g for g in groups if g]
    if len(groups) > 1:
        with Pool(len(groups)) as pool:
            parts = pool.map(solve_ports, [(name, g) for g in groups])
    else:
        parts = [solve_ports((name, groups[0]))]
    Vp = np.zeros((P, P, len(freqs)), dtype=np.complex128)
    rec = np.zeros((P, len(freqs), 3))
    for ports, v, r in parts:
        Vp[ports] = v[ports]
        rec[ports] = r[ports]
    # S as the reference assembles it: rows c1 from port c1's solution, the
    # upper triangle mirrored (S[c2, c1] = S[c1, c2] for c2 > c1)
    S = np.zeros_like(Vp)
    for k in range(len(freqs)):
        Sk = s_of(Vp[:, :, k])
        S[:, :, k] = np.triu(Sk) + np.triu(Sk, 1).T
    Z = OF.s2z(S, R0)
    np.savez_compressed(os.path.join(HERE, "sweep_%s.npz" % s["config"]), freqs=freqs, freq_index=idx, Vp=Vp,
                        S=S, Z=Z, flag=rec[:, :, 0].astype(np.int32), iters=rec[:, :, 1].astype(np.int32),
                        relres=rec[:, :, 2], tol=np.float64(s["tol"]), inner=np.int32(s["inner"]),
                        port_nodes=M.port_nodes.astype(np.int64))
    print("wrote sweep_%s.npz" % s["config"], flush=True)


if __name__ == "__main__":
    which = sys.argv[1]
    make(which, int(sys.argv[2]) if len(sys.argv) > 2 else (1 if which == "two_material" else 4))
