"""Write the kernel-row fixtures: the oracle's nine circulant rows on the
grids of the benchmark configurations (24x24x4, 64x64x16 and 96x96x24 at
0.25 x 0.25 x 0.05 mm cells, 192x192x48 at 0.1 mm cubes).

Run from the repository root, in the build container (needs oracle/'s C
restatement, built by `make -C oracle`):

    python tests/golden/make_kernel_rows.py [grid ...]

For each grid and kernel k (0..2 Lp x/y/z, 3..8 P rows of the caser pairs
(1,1) (3,3) (5,5) (3,2) (5,2) (5,4), compute_FFT_mutual_coupling_mats.jl:
36-128) it records
  - the row length and the SHA-256 of the scaled row's float64 bytes (the
    row as the spectra are built from: Lp x 1e6 with the self entry zeroed,
    P x 1e-12);
  - which closed form every entry routes to (a histogram), and how many
    entries sat within 1e-9 of a suppression threshold;
  - the index and value of every threshold-tie entry, of every entry that
    routes to one of the deviant reference forms, of the first 64 entries
    and of a strided sample of about 1024 more.
Outputs kernel_rows_<grid>.json (metadata) and kernel_rows_<grid>.npz
(index / value arrays; numpy, no pickles) beside this script.
"""
import hashlib
import json
import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import integrals as OI  # noqa: E402
from oracle.circulant import P_CASES, centers  # noqa: E402

GRIDS = {
    "crossbar_24x24x4": (24, 24, 4, 0.5e-3, 0.5e-3, 0.1e-3),
    "two_material_64x64x16": (64, 64, 16, 0.25e-3, 0.25e-3, 0.05e-3),
    "batched_96x96x24": (96, 96, 24, 0.25e-3, 0.25e-3, 0.05e-3),
    "large_192x192x48": (192, 192, 48, 0.1e-3, 0.1e-3, 0.1e-3),
}
ESC_LP, ESC_P = 1e6, 1e-12
CHUNK = 1 << 18
NSAMPLE = 1024


def _chunk(args):
    kind, src, c, s = args
    if kind < 3:
        return OI.lp_row_forms(src, c, *s, *s, kind + 1)
    c1, c2 = P_CASES[kind - 3]
    return OI.p_row_forms(src, c, *s, c1, c2)


def row_and_forms(kind, Nx, Ny, Nz, s, pool):
    c = centers(kind, (0.0, 0.0, 0.0), Nx, Ny, Nz, s)
    parts = pool.map(_chunk, [(kind, c[0], c[i:i + CHUNK], s) for i in range(0, len(c), CHUNK)])
    v = np.concatenate([p[0] for p in parts])
    f = np.concatenate([p[1] for p in parts])
    if kind < 3:
        v = ESC_LP * v
        v[0] = 0.0
    else:
        v = ESC_P * v
    return v, f


def make(name, pool):
    Nx, Ny, Nz, sx, sy, sz = GRIDS[name]
    meta = dict(grid=[Nx, Ny, Nz], cell=[sx, sy, sz], origin=[0.0, 0.0, 0.0], kernels=[])
    arrays = {}
    for k in range(9):
        v, f = row_and_forms(k, Nx, Ny, Nz, (sx, sy, sz), pool)
        N = v.size
        tie = f >= 100
        base = f % 100
        dev = np.isin(base, OI.DEVIANT_FORMS)
        stride = max(1, N // NSAMPLE)
        pick = np.unique(np.concatenate([np.nonzero(tie)[0], np.nonzero(dev)[0], np.arange(min(64, N)),
                                         np.arange(0, N, stride)]))
        u, cnt = np.unique(base, return_counts=True)
        meta["kernels"].append(dict(
            kind=k, N=int(N), sha256=hashlib.sha256(np.ascontiguousarray(v, dtype="<f8").tobytes()).hexdigest(),
            forms={OI.FORM_NAMES[int(a)]: int(n) for a, n in zip(u, cnt)},
            ties=int(tie.sum()), deviant=int(dev.sum()),
            tie_forms={OI.FORM_NAMES[int(a)]: int(n) for a, n in zip(*np.unique(base[tie], return_counts=True))},
            deviant_forms={OI.FORM_NAMES[int(a)]: int(n)
                           for a, n in zip(*np.unique(base[dev], return_counts=True))}))
        arrays["idx%d" % k] = pick.astype(np.int64)
        arrays["val%d" % k] = v[pick]
        arrays["form%d" % k] = f[pick].astype(np.int32)
        print("%s kernel %d: N %d, ties %d, deviant %d, %s" % (name, k, N, tie.sum(), dev.sum(),
                                                                meta["kernels"][-1]["forms"]), flush=True)
    with open(os.path.join(HERE, "kernel_rows_%s.json" % name), "w") as fh:
        json.dump(meta, fh, indent=1)
    np.savez_compressed(os.path.join(HERE, "kernel_rows_%s.npz" % name), **arrays)


if __name__ == "__main__":
    names = sys.argv[1:] or list(GRIDS)
    with Pool(min(8, os.cpu_count() or 1)) as pool:
        for n in names:
            make(n, pool)
