"""ORACLE -- test infrastructure only (see oracle/__init__.py).

numpy restatement of the retarded (Rcc, QS_Rcc_FW = 2) kernels:

  compute_rows_Rcc_Lp            compute_rows_Rcc_Lp.jl:1-105
  compute_rows_Rcc_P             compute_rows_Rcc_P.jl:1-64
  compute_Voxels_Rcc             compute_Voxels_Rcc.jl:1-13
  compute_P_vox_Rcc              compute_P_vox_Rcc.jl:1-87
  build_row_vox_Rcc              build_row_vox_Rcc.jl:1-10
  compute_Circulant_Lp_Rcc /     compute_Circulant_Lp_Rcc.jl:1-25,
  compute_Circulant_P_sup_Rcc    compute_Circulant_P_sup_Rcc.jl:1-24
  store_circulant(_P)            compute_Circulant_Lp_Rcc.jl:28-57

Per geometry the rows hold the static (quasi-static) integrals -- the Lp
rows of compute_Lp_Voxels_QS.jl's closed forms with the self entry zeroed,
the P rows of compute_row_P_sup -- and the centre-to-centre distances R.
Per frequency every row becomes escaling * row .* exp(-j beta R), beta =
2 pi f sqrt(eps0 mu0), embedded even-symmetrically and transformed; the
matvec is the quasi-static one with these complex spectra (the sweep passes
them to gmres_custom in place of FFTCP / FFTCLp, FFT_solver_QS_S_type.jl:61-64,
178-179).
"""
import numpy as np

from . import integrals as OI
from .circulant import P_CASES, centers, grid_dims

EPS0 = 8.854187816997944e-12
MU0 = 4 * np.pi * 1e-7


def _norm3(d):
    """LinearAlgebra.norm of 3-vectors (rows of d) without rescaling:
    sqrt((d1^2 + d2^2) + d3^2)."""
    return np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])


def voxels_rcc(c0, c):
    """compute_Voxels_Rcc.jl:1-13: |c0 - c[n]|."""
    return _norm3(np.asarray(c0, dtype=np.float64)[None, :] - c)


def _face(caser, c, sx, sy, sz):
    """Face centre of voxel centres c (N x 3) for a caser code
    (compute_P_vox_Rcc.jl:3-28 for the source, :35-80 for the targets)."""
    f = np.array(c, dtype=np.float64, copy=True)
    if caser in (1, 2):
        f[:, 1] = c[:, 1] - sy / 2 if caser == 1 else c[:, 1] + sy / 2
    elif caser in (3, 4):
        f[:, 0] = c[:, 0] - sx / 2 if caser == 3 else c[:, 0] + sx / 2
    else:
        f[:, 2] = c[:, 2] - sz / 2 if caser == 5 else c[:, 2] + sz / 2
    return f


def p_vox_rcc(c0, c, sx, sy, sz, caser, caser2):
    """compute_P_vox_Rcc.jl:1-87: |face(c0, caser) - face(c[k], caser2)|."""
    src = _face(caser, np.asarray(c0, dtype=np.float64)[None, :], sx, sy, sz)
    return _norm3(src - _face(caser2, c, sx, sy, sz))


def rcc_rows(origin, Nx, Ny, Nz, sx, sy, sz):
    """compute_rows_Rcc_Lp / compute_rows_Rcc_P: {kind: (static row, R)}
    for the 9 kinds (0..2 Lp x/y/z, 3..8 P 11, 22, 33, 12, 13, 23), the rows
    without escaling, Lp self entries zeroed (compute_rows_Rcc_Lp.jl:41)."""
    out = {}
    for d in range(3):
        c = centers(d, origin, Nx, Ny, Nz, (sx, sy, sz))
        if len(c) == 0:
            out[d] = (np.zeros(0), np.zeros(0))
            continue
        r = OI.lp_row_qs(c[0], c, sx, sy, sz, sx, sy, sz, d + 1)
        r[0] = 0.0
        out[d] = (r, voxels_rcc(c[0], c))
    for i, (c1, c2) in enumerate(P_CASES):
        c = centers(3 + i, origin, Nx, Ny, Nz, (sx, sy, sz))
        out[3 + i] = (OI.p_row(c[0], c, sx, sy, sz, c1, c2), p_vox_rcc(c[0], c, sx, sy, sz, c1, c2))
    return out


def beta_of(freq_hz):
    """build_row_vox_Rcc.jl:3-5."""
    return 2 * np.pi * freq_hz * np.sqrt(EPS0 * MU0)


def row_vox_rcc(row, R, freq_hz):
    """build_row_vox_Rcc.jl:1-10: row .* exp(-j beta R)."""
    th = -beta_of(freq_hz) * R
    return row * np.cos(th) + 1j * (row * np.sin(th))


def store_circulant(row, Nx, Ny, Nz):
    """compute_Circulant_Lp_Rcc.jl:28-57 (and store_circulant_P): the row on
    an Nx x Ny x Nz grid (x fastest), mirrored into 2N per axis with the
    entry N (0-based) of each axis zero."""
    C = np.asarray(row, dtype=np.complex128).reshape(Nz, Ny, Nx).transpose(2, 1, 0)
    out = np.zeros((2 * Nx, 2 * Ny, 2 * Nz), dtype=np.complex128)
    ix = [np.arange(Nx), np.arange(Nx + 1, 2 * Nx)]
    rx = [np.arange(Nx), np.arange(Nx - 1, 0, -1)]
    iy = [np.arange(Ny), np.arange(Ny + 1, 2 * Ny)]
    ry = [np.arange(Ny), np.arange(Ny - 1, 0, -1)]
    iz = [np.arange(Nz), np.arange(Nz + 1, 2 * Nz)]
    rz = [np.arange(Nz), np.arange(Nz - 1, 0, -1)]
    for a in range(2):
        for b in range(2):
            for c in range(2):
                out[np.ix_(ix[a], iy[b], iz[c])] = C[np.ix_(rx[a], ry[b], rz[c])]
    return out


def rcc_spectra(rows, Nx, Ny, Nz, freq_hz, esc_Lp=1e6, esc_P=1e-12):
    """compute_Circulant_Lp_Rcc + compute_Circulant_P_sup_Rcc at one
    frequency (Hz): {kind: fft(store_circulant(escaling * row .* exp(-j beta R)))}."""
    out = {}
    for k in range(9):
        n = grid_dims(k, Nx, Ny, Nz)
        if min(n) <= 0:
            out[k] = None
            continue
        row, R = rows[k]
        esc = esc_Lp if k < 3 else esc_P
        out[k] = np.fft.fftn(store_circulant(esc * row_vox_rcc(row, R, freq_hz), *n))
    return out
