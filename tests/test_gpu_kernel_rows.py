"""The device's circulant rows (k_grid_row: centres generated on the
device, classifier, closed forms) against the oracle fixtures of the
benchmark grids, bit for bit: the SHA-256 of every whole row and every
sampled entry (the threshold ties and the deviant-form entries among them).
The fixtures come from tests/golden/make_kernel_rows.py (oracle, CPU)."""
import numpy as np
import pytest

from golden import kernel_rows as KR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name", KR.GRIDS)
def test_device_kernel_rows_match_fixtures(name):
    from solver_amd import peec
    meta, smp = KR.load(name)
    Nx, Ny, Nz = meta["grid"]
    sx, sy, sz = meta["cell"]
    print("grid %s: building kernel rows and spectra" % name, flush=True)
    with peec.Context(0) as ctx:
        assert peec.compute_FFT_mutual_coupling_mats(np.asarray(meta["origin"]), Nx, Ny, Nz, sx, sy, sz,
                                                     {"Lp": KR.ESC_LP, "P": KR.ESC_P}, ctx=ctx)
        for k in range(9):
            row = peec.kernel_row(k, ctx=ctx)
            kk = meta["kernels"][k]
            idx, val, form = smp[k]
            assert row.size == kk["N"], (name, k)
            bad = np.nonzero(row[idx].view(np.int64) != val.view(np.int64))[0]
            assert bad.size == 0, (name, k, idx[bad[:5]], row[idx][bad[:5]], val[bad[:5]], form[bad[:5]])
            assert KR.sha256(row) == kk["sha256"], (name, k)
            print("  kernel %d: %d entries bitwise equal (ties %d, deviant %d)" % (k, row.size, kk["ties"],
                                                                                kk["deviant"]), flush=True)
