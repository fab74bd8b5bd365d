"""Write the request-sweep fixture: the oracle's quasi-static FFT solver
(FFT_solver_QS_S_type.jl:60-232) on the reference's three request fixtures
(requests/*.json.gz), meshed by the reference mesher restatement, at every
frequency each request lists, with the request's own GMRES settings
(innerIteration, convergenceThreshold) and the GPU path's elimination order.

Run from the repository root, in the build container (about 2 minutes):

    python tests/golden/make_requests.py

Output: requests_oracle.npz beside this script (numpy arrays, no pickles):
per request <name>_Z, _S, _Y (ports x ports x frequencies), _f, and
_flag / _iters (ports x frequencies) of each gmres_custom solve.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import fft_solver as OF  # noqa: E402
from oracle import integrals as OI  # noqa: E402
from solver_amd import ordering, request  # noqa: E402

NAMES = ("first_run_data", "first_run_data_2_materials", "crossbar")


def p_self(sx, sy, sz):
    c = np.zeros((1, 3))
    return [OI.p_row(c[0], c, sx, sy, sz, cs, cs)[0] for cs in (1, 3, 5)]


def main():
    out = {}
    for name in NAMES:
        R = request.parse(os.path.join(HERE, "requests", name + ".json.gz"))
        M = request.mesh_of(R, p_self_fn=p_self)
        rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
        spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
        rec = []
        res = OF.fft_solver_qs(R.freqs, M, spec, R.inner_iter, R.tol, R0=R.R0, record=rec,
                               order=ordering.mesh_order(M)[0])
        P, F = M.port_nodes.shape[0], len(R.freqs)
        flag = np.zeros((P, F), dtype=np.int32)
        iters = np.zeros((P, F), dtype=np.int32)
        for r in rec:
            flag[r["port"], r["k"]] = r["flag"]
            iters[r["port"], r["k"]] = r["iter"][1]
        for key in ("Z", "S", "Y", "f"):
            out["%s_%s" % (name, key)] = res[key]
        out[name + "_flag"] = flag
        out[name + "_iters"] = iters
        print("%s: N=%d, flags %s, iterations %s" % (name, M.m + M.ns + M.n, flag.ravel().tolist(),
                                                    iters.ravel().tolist()), flush=True)
    np.savez_compressed(os.path.join(HERE, "requests_oracle.npz"), **out)
    print("wrote requests_oracle.npz")


if __name__ == "__main__":
    main()
