"""The reference's own requests, end to end on the GPU: request JSON ->
reference mesher -> kernel rows and spectra on the device -> the plain
FP64 sweep (FFT_solver_QS_S_type.jl:60-232) -> dump_json_data's result.

Held to the oracle's sweep of the same requests (tests/golden/
make_requests.py, run on the CPU with each request's innerIteration and
convergenceThreshold): the same GMRES flags and iteration counts per port
and frequency, and Z, S and Y within 1e-9 of the largest entry of each
port row (the GPU runs the same Krylov iteration; what differs is rounding
in the FFT convolutions and the LU, amplified by the conditioning of SS at
these frequencies).
"""
import os

import numpy as np
import pytest

from solver_amd import peec, request

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
NAMES = ("first_run_data", "first_run_data_2_materials", "crossbar")


@pytest.fixture(scope="module")
def oracle():
    return np.load(os.path.join(GOLDEN, "requests_oracle.npz"))


def row_err(A, B):
    """max over entries of |A - B| / max |B| of its port row, per frequency."""
    scale = np.max(np.abs(B), axis=1, keepdims=True)
    return float(np.max(np.abs(A - B) / scale))


@pytest.mark.parametrize("name", NAMES)
def test_request_sweep_matches_oracle(name, oracle):
    path = os.path.join(GOLDEN, "requests", name + ".json.gz")
    R = request.parse(path)
    P, F = len(R.ports), len(R.freqs)
    print("%s: %s grid, %d materials, %d ports, %d frequencies" % (name, R.grids[0].shape, len(R.grids), P, F),
          flush=True)
    rec = []
    with peec.Context(0) as ctx:
        out = request.solve(path, id_=name, ctx=ctx, record=rec)
    assert out is not None and out["id"] == name
    flag = np.zeros((P, F), dtype=np.int32)
    iters = np.zeros((P, F), dtype=np.int32)
    for r in rec:
        flag[r["port"], r["k"]] = r["flag"]
        iters[r["port"], r["k"]] = r["iter"][1]
    assert np.array_equal(flag, oracle[name + "_flag"])
    assert np.array_equal(iters, oracle[name + "_iters"])
    for key, mat in (("Z", "matrix_Z"), ("S", "matrix_S"), ("Y", "matrix_Y")):
        A = request.parse_matrix(out["matrices"][mat], P, F)
        e = row_err(A, oracle["%s_%s" % (name, key)])
        print("%s %s: max row-relative error vs oracle %.2e" % (name, key, e), flush=True)
        assert e < 1e-9
