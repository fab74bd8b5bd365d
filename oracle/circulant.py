"""ORACLE -- test infrastructure only (see oracle/__init__.py).

Centre sets of the circulant rows, restated from the reference.

build_center_air_Voxels.jl:1-26, build_centers_Lp_with_air.jl:1-60 and
build_center_P_Voxels.jl:1-111: every set is a regular grid (x fastest) and
row k of a circulant is the integral from the set's first centre to all
centres of the set.
"""
import numpy as np


def _vox(mn, s, i1):
    return mn + s * (i1 - 1) + s / 2


def _half(mn, s, i1):
    return mn + s * (np.ceil(i1 / 2) - 1) + s / 2


def grid_dims(kind, Nx, Ny, Nz):
    return {
        0: (Nx - 1, Ny, Nz), 1: (Nx, Ny - 1, Nz), 2: (Nx, Ny, Nz - 1),
        3: (Nx, Ny + 1, Nz), 4: (Nx + 1, Ny, Nz), 5: (Nx, Ny, Nz + 1),
        6: (2 * (Nx + 1) - 1, 2 * (Ny + 1) - 1, Nz),
        7: (Nx, 2 * (Ny + 1) - 1, 2 * (Nz + 1) - 1),
        8: (2 * (Nx + 1) - 1, Ny, 2 * (Nz + 1) - 1),
    }[kind]


def centers(kind, origin, Nx, Ny, Nz, s):
    """N x 3 centres (x fastest) of circulant set `kind` (0..2 Lp x/y/z,
    3..8 P p12_se, p34_se, p56_se, p1234, p1256, p3456)."""
    d = grid_dims(kind, Nx, Ny, Nz)
    i, j, k = np.meshgrid(np.arange(1, d[0] + 1), np.arange(1, d[1] + 1), np.arange(1, d[2] + 1), indexing="ij")
    I = [i.transpose(2, 1, 0).reshape(-1), j.transpose(2, 1, 0).reshape(-1), k.transpose(2, 1, 0).reshape(-1)]
    out = np.empty((I[0].size, 3))
    if kind <= 2:
        for a in range(3):
            A = _vox(origin[a], s[a], I[a])
            B = _vox(origin[a], s[a], I[a] + (1 if a == kind else 0))
            out[:, a] = 0.5 * (A + B)
    elif kind <= 5:
        for a in range(3):
            out[:, a] = _vox(origin[a], s[a], I[a])
    else:
        half = {6: (1, 1, 0), 7: (0, 1, 1), 8: (1, 0, 1)}[kind]
        for a in range(3):
            out[:, a] = _half(origin[a], s[a], I[a]) if half[a] else _vox(origin[a], s[a], I[a])
    return out


# (caser, caser2) of the six P rows, compute_FFT_mutual_coupling_mats.jl:96-121
P_CASES = [(1, 1), (3, 3), (5, 5), (3, 2), (5, 2), (5, 4)]
