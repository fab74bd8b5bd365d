"""Kernel-row fixtures (tests/golden/kernel_rows_*): the oracle's nine
circulant rows on the grids of the benchmark configurations, as SHA-256
hashes of the whole rows plus sampled entries (every entry sitting at a
suppression threshold, every entry routed to a deviant reference form, the
first 64 and a strided sample).

CPU side:
  - the oracle still reproduces the fixtures (whole rows on the two small
    grids, the sampled entries on the two large ones);
  - the product's per-entry code of k_grid_row -- the device centre
    arithmetic, the classifier and the closed forms -- run on the host
    through a test hook, is bitwise equal to the oracle on every sampled
    entry of every grid, including the threshold ties, and on whole rows of
    the 24x24x4 and 64x64x16 grids.
The GPU test (test_gpu_kernel_rows.py) compares the device rows bitwise
with the fixtures' hashes.
"""
import ctypes

import numpy as np
import pytest

from golden import kernel_rows as KR
from oracle import integrals as OI
from oracle.circulant import P_CASES, centers


def host_rows(meta, k, idx):
    from solver_amd import _lib as L
    Nx, Ny, Nz = meta["grid"]
    sx, sy, sz = meta["cell"]
    org = np.asarray(meta["origin"], dtype=np.float64)
    idx = np.ascontiguousarray(idx, dtype=np.int64)
    out = np.empty(idx.size)
    n = ctypes.c_int64()
    scale = KR.ESC_LP if k < 3 else KR.ESC_P
    rc = L.lib().peec_test_grid_row_host(k, L.as_dp(org), Nx, Ny, Nz, sx, sy, sz, scale,
                                         idx.ctypes.data_as(L.i64p), idx.size, L.as_dp(out), ctypes.byref(n))
    assert rc == 0
    if k < 3:
        out[idx == 0] = 0.0
    return out, n.value


def oracle_entries(meta, k, idx):
    Nx, Ny, Nz = meta["grid"]
    s = tuple(meta["cell"])
    c = centers(k, tuple(meta["origin"]), Nx, Ny, Nz, s)
    if k < 3:
        v = KR.ESC_LP * OI.lp_row(c[0], c[idx], *s, *s, k + 1)
        v[np.asarray(idx) == 0] = 0.0
    else:
        c1, c2 = P_CASES[k - 3]
        v = KR.ESC_P * OI.p_row(c[0], c[idx], *s, c1, c2)
    return v, len(c)


@pytest.mark.parametrize("name", KR.GRIDS)
def test_oracle_reproduces_sampled_entries(name):
    meta, smp = KR.load(name)
    for k in range(9):
        idx, val, form = smp[k]
        v, N = oracle_entries(meta, k, idx)
        assert N == meta["kernels"][k]["N"]
        assert np.array_equal(v.view(np.int64), val.view(np.int64)), (name, k)


@pytest.mark.parametrize("name", KR.GRIDS[:2])
def test_oracle_reproduces_whole_row_hashes(name):
    meta, _ = KR.load(name)
    for k in range(9):
        v, N = oracle_entries(meta, k, np.arange(meta["kernels"][k]["N"]))
        assert KR.sha256(v) == meta["kernels"][k]["sha256"], (name, k)


@pytest.mark.parametrize("name", KR.GRIDS)
def test_product_entry_code_matches_fixture_samples(name):
    meta, smp = KR.load(name)
    nties = 0
    for k in range(9):
        idx, val, form = smp[k]
        v, N = host_rows(meta, k, idx)
        assert N == meta["kernels"][k]["N"]
        bad = np.nonzero(v.view(np.int64) != val.view(np.int64))[0]
        assert bad.size == 0, (name, k, idx[bad[:5]], v[bad[:5]], val[bad[:5]], form[bad[:5]])
        nties += int(np.sum(form >= 100))
    assert nties == sum(kk["ties"] for kk in meta["kernels"])


@pytest.mark.parametrize("name", KR.GRIDS[:2])
def test_product_entry_code_whole_rows(name):
    meta, _ = KR.load(name)
    for k in range(9):
        N = meta["kernels"][k]["N"]
        v, _ = host_rows(meta, k, np.arange(N))
        assert KR.sha256(v) == meta["kernels"][k]["sha256"], (name, k)


def test_fixture_routing_statistics():
    """The routing the fixtures record: no suppression anywhere on the
    24x24x4 grid; on the 0.05 mm-layer grids the collinear P form
    (a deviant reference form) is reached by the P11 / P22 rows only; the
    192x192x48 grid reaches no deviant form."""
    meta, _ = KR.load("crossbar_24x24x4")
    for kk in meta["kernels"]:
        assert kk["ties"] == 0 and kk["deviant"] == 0
        assert set(kk["forms"]) <= {"vol_vol", "qs_surf_surf_para", "p_surf_surf_ortho"}
    for name in ("two_material_64x64x16", "batched_96x96x24"):
        meta, _ = KR.load(name)
        dev = {kk["kind"]: kk["deviant_forms"] for kk in meta["kernels"] if kk["deviant"]}
        assert set(dev) == {3, 4} and all(set(d) == {"p_line_line_sp"} for d in dev.values())
    meta, _ = KR.load("large_192x192x48")
    assert all(kk["deviant"] == 0 for kk in meta["kernels"])
