"""Loader of the kernel-row fixtures written by make_kernel_rows.py."""
import hashlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GRIDS = ["crossbar_24x24x4", "two_material_64x64x16", "batched_96x96x24", "large_192x192x48"]
ESC_LP, ESC_P = 1e6, 1e-12


def load(name):
    """(meta, samples): meta as written (grid, cell, origin, per-kernel N,
    sha256, form histogram, tie and deviant counts); samples[k] =
    (indices, values, forms) of kernel k's recorded entries."""
    with open(os.path.join(HERE, "kernel_rows_%s.json" % name)) as fh:
        meta = json.load(fh)
    z = np.load(os.path.join(HERE, "kernel_rows_%s.npz" % name))
    samples = {k: (z["idx%d" % k], z["val%d" % k], z["form%d" % k]) for k in range(9)}
    return meta, samples


def sha256(row):
    return hashlib.sha256(np.ascontiguousarray(row, dtype="<f8").tobytes()).hexdigest()
