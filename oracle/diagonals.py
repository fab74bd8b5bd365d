"""ORACLE -- test infrastructure only (see oracle/__init__.py).

Restatement of compute_diagonals (compute_diagonals.jl:1-134): the
half-voxel resistances R and dielectric capacitances Cd of every edge's four
slots (the two half voxels of the edge, then the two border half voxels) and
the self partial inductance of every edge, lengthened by half a cell per
border end.

Inputs follow the reference's meaning:
  materials   list of dicts with "sigmar" and "epsr" (complex epsr, as
              mesher_FFT.jl:21 builds it), 1-based material ids;
  li_mat      (m, 2) material id of each edge's two half voxels, per axis;
  li_border   (m, 2) material id of a half voxel on a border end, 0 inside.
Lp_self is the C restatement (compute_diagonals.jl:136-155).
"""
import numpy as np

from . import integrals as OI

EPS0 = 8.854187816997944e-12


def material_terms(mat, sx, sy, sz):
    """compute_diagonals.jl:6-32: (Rx, Ry, Rz, Cx, Cy, Cz) of one material."""
    sigmar, epsr = mat["sigmar"], mat["epsr"]
    if sigmar != 0:
        Rx = 0.5 * sx / (sigmar * sy * sz)
        Ry = 0.5 * sy / (sigmar * sx * sz)
        Rz = 0.5 * sz / (sigmar * sy * sx)
        if epsr == 1:
            Cx = Cy = Cz = 0
        else:
            Cx = EPS0 * (epsr - 1) * sy * sz / (0.5 * sx)
            Cy = EPS0 * (epsr - 1) * sx * sz / (0.5 * sy)
            Cz = EPS0 * (epsr - 1) * sy * sx / (0.5 * sz)
    else:
        Rx = Ry = Rz = 0
        Cx = EPS0 * (epsr - 1) * sy * sz / (0.5 * sx)
        Cy = EPS0 * (epsr - 1) * sx * sz / (0.5 * sy)
        Cz = EPS0 * (epsr - 1) * sy * sx / (0.5 * sz)
    return Rx, Ry, Rz, Cx, Cy, Cz


def compute_diagonals(escalings, materials, sx, sy, sz, lix_mat, liy_mat, liz_mat, lix_border, liy_border,
                      liz_border):
    """compute_diagonals.jl:1-134.  Returns dict R (m x 4), Cd (m x 4,
    complex), Lp (m) and fc_Lp (m), scaled by the escalings."""
    terms = [material_terms(mat, sx, sy, sz) for mat in materials]
    blocks = []
    for axis, (mat2, bord2) in enumerate(((lix_mat, lix_border), (liy_mat, liy_border), (liz_mat, liz_border))):
        mat2 = np.asarray(mat2)
        bord2 = np.asarray(bord2)
        R = np.zeros((mat2.shape[0], 4))
        C = np.zeros((mat2.shape[0], 4), dtype=np.complex128)
        # :39-104, one material at a time, its R then its C of this axis
        for cont, t in enumerate(terms, start=1):
            Rv, Cv = t[axis], t[3 + axis]
            for col, src in ((0, mat2[:, 0]), (1, mat2[:, 1]), (2, bord2[:, 0]), (3, bord2[:, 1])):
                ind = src == cont
                if Rv != 0:
                    R[ind, col] = Rv
                if Cv != 0:
                    C[ind, col] = Cv
        blocks.append((R, C))
    # :105-125: border count per edge, ceil(id / 100) per end
    lens = ((sx, sy, sz), (sy, sx, sz), (sz, sy, sx))
    Lp, Lp0 = [], []
    for axis, bord2 in enumerate((lix_border, liy_border, liz_border)):
        bord2 = np.asarray(bord2)
        aux = np.ceil(bord2[:, 0] / 100).astype(np.int64) + np.ceil(bord2[:, 1] / 100).astype(np.int64)
        l, W, T = lens[axis]
        s0 = OI.lp_self(l, W, T)
        s1 = OI.lp_self(l + l / 2, W, T)
        s2 = OI.lp_self(2 * l, W, T)
        d = np.full(aux.size, s0)
        d[aux == 1] = s1
        d[aux == 2] = s2
        Lp.append(d)
        Lp0.append(np.full(aux.size, s0))
    Lp = np.concatenate(Lp)
    Lp0 = np.concatenate(Lp0)
    return dict(R=escalings["R"] * np.concatenate([b[0] for b in blocks]),
                Cd=escalings["Cd"] * np.concatenate([b[1] for b in blocks]),
                Lp=escalings["Lp"] * Lp, fc_Lp=escalings["Lp"] * (Lp - Lp0))
