"""Pin the oracle's closed-form integrals against independent quadrature.

The oracle is the C restatement in oracle/peec_integrals_oracle.c.  Each
closed form is compared with tests/quadrature.py's Gauss-Legendre evaluation
of the integral it stands for; Lp_self (Ruehli's published formula,
compute_diagonals.jl:136-155) is compared with the coincident vol-vol form;
rows are checked for translation invariance, exchange symmetry and the
lambda^5 / lambda^3 scaling laws, and the far-field branch is checked against
V1*V2/R.  A dropped term, a wrong sign or a transposed argument in any form
moves the result by far more than the tolerances used here.
"""
import numpy as np
import pytest

from oracle import integrals as OI
from quadrature import box_integral

S = 5e-4
T = 1e-4


def iv(c, h):
    return (c - h / 2, c + h / 2)


def pt(c):
    return (c, c)


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


# (closed-form id, packed args, box1, box2)
def _cases():
    cs = []
    X1, Y1, Z1 = iv(0, S), iv(0, S), iv(0, T)
    for off in [(S, 0, 0), (0, S, T), (2 * S, S, 0), (S, 3 * S, 2 * T)]:
        X2, Y2, Z2 = iv(off[0], S), iv(off[1], S), iv(off[2], T)
        cs.append((0, [*X1, *Y1, *Z1, *X2, *Y2, *Z2], (X1, Y1, Z1), (X2, Y2, Z2)))
    # vol_surf: vol1 full, rect2 at z2
    for z2 in (T, 3 * T, 2 * S):
        X2, Y2 = iv(S, S), iv(0.5 * S, S)
        cs.append((1, [*X1, *Y1, *Z1, *X2, *Y2, z2], (X1, Y1, Z1), (X2, Y2, pt(z2))))
    # surf_surf_para (Lp copy id 2, QS copy id 12)
    for off in [(S, 0, T), (0.5 * S, 2 * S, 3 * T), (2 * S, S, 0.0)]:
        A = (iv(0, S), iv(0, T), pt(0.0))
        B = (iv(off[0], S), iv(off[1], T), pt(off[2]))
        args = [*A[0], *A[1], 0.0, *B[0], *B[1], off[2]]
        cs.append((2, args, A, B))
        cs.append((12, args, A, B))
    # surf_surf_ortho: rect1 (x,y) at z1 ; rect2 (x, z) at y2 (Lp copy 3, P copy 13)
    for y2, z0 in [(S, T), (2 * S, -S), (0.7 * S, 0.3 * S)]:
        A = (iv(0, S), iv(0, S), pt(0.0))
        B = (iv(S, S), pt(y2), iv(z0, S))
        args = [*A[0], *A[1], 0.0, *B[0], y2, *B[2]]
        cs.append((3, args, A, B))
        cs.append((13, args, A, B))
    # line_vol: vol1, line2 along x at (y2, z2)
    for y2, z2 in [(S, T), (2 * S, 3 * S)]:
        A = (iv(0, S), iv(0, S), iv(0, T))
        B = (iv(2 * S, S), pt(y2), pt(z2))
        cs.append((4, [*A[0], *A[1], *A[2], *B[0], y2, z2], A, B))
    # line_surf_para: rect1 (x, y) at z1, line2 along x at (y2, z2)
    for y2, z2 in [(S, T), (0.2 * S, 2 * S)]:
        A = (iv(0, S), iv(0, S), pt(0.0))
        B = (iv(S, S), pt(y2), pt(z2))
        cs.append((5, [*A[0], *A[1], 0.0, *B[0], y2, z2], A, B))
    # line_surf_ortho: line1 along x at (y1, z1), rect2 (y, z) at x2 (Lp 6, P 14)
    for x2, y1 in [(2 * S, 0.0), (S, 0.4 * S)]:
        A = (iv(0, S), pt(y1), pt(T))
        B = (pt(x2), iv(S, S), iv(0, S))
        args = [*A[0], y1, T, x2, *B[1], *B[2]]
        cs.append((6, args, A, B))
        cs.append((14, args, A, B))
    # point_vol
    for p in [(2 * S, 0.0, 0.0), (S, S, 3 * T)]:
        A = (pt(p[0]), pt(p[1]), pt(p[2]))
        B = (iv(0, S), iv(0, S), iv(0, T))
        cs.append((7, [*p, *B[0], *B[1], *B[2]], A, B))
    # line_line_ortho_xy: line1 along x at (y1, z1); line2 along y at (x2, z2)
    for x2, z2 in [(2 * S, 0.0), (S, T)]:
        A = (iv(0, S), pt(0.0), pt(0.0))
        B = (pt(x2), iv(S, S), pt(z2))
        cs.append((8, [*A[0], 0.0, 0.0, x2, *B[1], z2], A, B))
    # line_line_parall (Lp 9, P 15), off-axis
    for dy, dz in [(S, 0.0), (S, T)]:
        A = (iv(0, S), pt(0.0), pt(0.0))
        B = (iv(2 * S, S), pt(dy), pt(dz))
        args = [*A[0], 0.0, 0.0, *B[0], dy, dz]
        cs.append((9, args, A, B))
        cs.append((15, args, A, B))
    # collinear segments (line_line_sp branch), Lp 9 and P 15
    A = (iv(0, S), pt(0.0), pt(0.0))
    B = (iv(3 * S, S), pt(0.0), pt(0.0))
    cs.append((109, [*A[0], 0.0, 0.0, *B[0], 0.0, 0.0], A, B))
    cs.append((115, [*A[0], 0.0, 0.0, *B[0], 0.0, 0.0], A, B))
    # point_sup
    for p in [(0.0, 0.0, S), (2 * S, S, T)]:
        A = (pt(p[0]), pt(p[1]), pt(p[2]))
        B = (iv(0, S), iv(0, S), pt(0.0))
        cs.append((10, [*p, *B[0], *B[1], 0.0], A, B))
    # line_point: line along x at (y3, z3), point (x1, y1, z1)
    for p in [(2 * S, S, 0.0), (0.0, S, T)]:
        A = (iv(0, S), pt(0.0), pt(0.0))
        cs.append((11, [*A[0], 0.0, 0.0, *p], A, (pt(p[0]), pt(p[1]), pt(p[2]))))
    return cs


CASES = _cases()

# Closed forms whose reference text is not the integral their call sites use
# it for.  The oracle reproduces the reference; these are measured, not
# pinned, and the deviation is recorded here (see test below).
# 109 / 115 are ids 9 / 15 evaluated on collinear segments, which routes to
# integ_line_line_sp_fft (compute_Lp_Voxels.jl:514-529) and
# integ_line_line_sp_P (compute_row_P_sup.jl:461-476).
KNOWN_DEVIANT = {1, 3, 6, 109, 115}


def _cf(fid, args):
    return OI.closed_form(fid % 100, np.array(args, dtype=np.float64))


@pytest.mark.parametrize("k", [i for i, c in enumerate(CASES) if c[0] not in KNOWN_DEVIANT])
def test_closed_form_vs_quadrature(k):
    fid, args, A, B = CASES[k]
    cf = _cf(fid, args)
    q = box_integral(A, B, n=28)
    assert rel(cf, q) < 1e-8, (fid, cf, q)


def test_known_deviant_forms_really_deviate():
    """Three Lp-copy closed forms do not evaluate the integral their call
    sites use them for: integ_vol_surf_fft (compute_Lp_Voxels.jl:728-764),
    integ_surf_surf_ortho_fft (:699-726; the P copy at
    compute_row_P_sup.jl:557-590 is correct) and integ_line_surf_ortho_fft
    (:584-607; the P copy at compute_row_P_sup.jl:500-529 is correct); and
    both collinear-segment forms integ_line_line_sp_fft (:514-529) and
    integ_line_line_sp_P (compute_row_P_sup.jl:461-476).  They
    are only reached when a dimension is suppressed (far, thin voxels).  The
    oracle reproduces the reference text; this records the deviation so a
    silent 'fix' in the oracle would be caught."""
    for fid, args, A, B in CASES:
        if fid in KNOWN_DEVIANT:
            cf = _cf(fid, args)
            q = box_integral(A, B, n=28)
            assert rel(cf, q) > 1e-2, fid


def test_self_vol_vol_matches_ruehli_lp_self():
    # coincident boxes: 1e-7/S * Integ == Lp_self(l, W, T) for an x-edge
    for l, W, Tt in [(S, S, T), (S, 0.5 * S, 0.2 * S), (1e-3, 2e-4, 5e-5)]:
        row = OI.lp_row([0, 0, 0], np.array([[0.0, 0.0, 0.0]]), l, W, Tt, l, W, Tt, 1)
        assert rel(row[0], OI.lp_self(l, W, Tt)) < 1e-12


def test_self_vol_vol_quadrature():
    A = (iv(0, S), iv(0, S), iv(0, T))
    q = box_integral(A, A, n=40)
    cf = OI.closed_form(0, np.array([*A[0], *A[1], *A[2], *A[0], *A[1], *A[2]]))
    assert rel(cf, q) < 1e-8


def _grid_centers(nx, ny, nz, s=(S, S, T), origin=(0.0, 0.0, 0.0)):
    i, j, k = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    c = np.stack([origin[0] + s[0] * i + s[0] / 2, origin[1] + s[1] * j + s[1] / 2,
                  origin[2] + s[2] * k + s[2] / 2], axis=-1)
    return c.transpose(2, 1, 0, 3).reshape(-1, 3)  # x fastest


def test_lp_row_translation_and_symmetry():
    c = _grid_centers(6, 5, 3)
    r0 = OI.lp_row(c[0], c, S, S, T, S, S, T, 1)
    sh = np.array([7 * S, -3 * S, 11 * T])
    r1 = OI.lp_row(c[0] + sh, c + sh, S, S, T, S, S, T, 1)
    # coordinates are rounded to 1e-14 m, so translation changes the last bits
    assert np.max(np.abs(r1 - r0)) < 1e-9 * np.max(np.abs(r0))
    # exchange symmetry: L(a, b) == L(b, a)
    for j in (1, 7, 20, len(c) - 1):
        back = OI.lp_row(c[j], c[:1], S, S, T, S, S, T, 1)[0]
        assert rel(back, r0[j]) < 1e-9


def test_lp_scaling_law():
    # Lp = 1e-7/S * I with I ~ lambda^5 and S ~ lambda^4 => Lp ~ lambda
    c = _grid_centers(4, 3, 2)
    r0 = OI.lp_row(c[0], c, S, S, T, S, S, T, 2)
    lam = 3.0
    r1 = OI.lp_row(c[0] * lam, c * lam, S * lam, S * lam, T * lam, S * lam, S * lam, T * lam, 2)
    assert np.max(np.abs(r1 - lam * r0)) < 1e-9 * np.max(np.abs(lam * r0))


def test_p_row_scaling_and_symmetry():
    c = _grid_centers(5, 4, 3)
    for (c1, c2) in [(1, 1), (3, 3), (5, 5), (3, 2), (5, 2), (5, 4)]:
        r0 = OI.p_row(c[0], c, S, S, T, c1, c2)
        lam = 2.0
        r1 = OI.p_row(c[0] * lam, c * lam, S * lam, S * lam, T * lam, c1, c2)
        # P = I / (4 pi eps0 A1 A2), I ~ lambda^3, A ~ lambda^2 => P ~ 1/lambda
        assert np.max(np.abs(r1 - r0 / lam)) < 1e-9 * np.max(np.abs(r0)), (c1, c2)
        assert np.all(np.isfinite(r0))


def test_p_diag_rows_vs_quadrature():
    # P(face_a, face_b) * 4 pi eps0 A1 A2 == integral over the two faces
    eps0 = 8.854187816997944e-12
    c = _grid_centers(4, 4, 3)
    for j in (1, 5, 17, 30):
        cj = c[j]
        r = OI.p_row(c[0], c[j:j + 1], S, S, T, 5, 5)[0]
        A1 = A2 = S * S
        f1 = (iv(c[0][0], S), iv(c[0][1], S), pt(c[0][2] - T / 2))
        f2 = (iv(cj[0], S), iv(cj[1], S), pt(cj[2] - T / 2))
        q = box_integral(f1, f2, n=28) / (4 * np.pi * eps0 * A1 * A2)
        assert rel(r, q) < 1e-8


def test_collinear_p_form_depends_on_segment_order():
    """integ_line_line_sp_P (compute_row_P_sup.jl:461-476) evaluates the
    collinear-segment integral exactly when the source segment lies after
    the target one (x1 > x2 at every corner); when it lies before, its text
    gives about minus the integral.  The rows of the benchmark grids reach it
    only in the second order: in the P rows of the (1,1) and (3,3) face pairs
    on the 0.25 x 0.25 x 0.05 mm grids, the faces 41 or more cells along x
    (resp. y) from the source lose their z extent and become collinear
    segments (23 entries per row at 64x64x16, 55 at 96x96x24).  The oracle
    follows the text there, so those entries are pinned by the exact order
    only; their measured value is -1.02 x the face integral."""
    A = (iv(0, S), pt(0.0), pt(0.0))
    for off in (1.5, 3.0, 41.0):
        B_after = (iv(-off * S, S), pt(0.0), pt(0.0))     # source segment follows
        q = box_integral(A, B_after, n=28)
        assert rel(_cf(115, [*A[0], 0.0, 0.0, *B_after[0], 0.0, 0.0]), q) < 1e-10
        B_before = (iv(off * S, S), pt(0.0), pt(0.0))     # source segment precedes
        v = _cf(115, [*A[0], 0.0, 0.0, *B_before[0], 0.0, 0.0])
        assert -1.25 < v / box_integral(A, B_before, n=28) < -0.99
    # the benchmark-grid entries, through the oracle's P row
    sx, sy, sz = 0.25e-3, 0.25e-3, 0.05e-3
    src = np.array([0.0, 0.0, 0.0])
    eps0 = 8.854187816997944e-12
    for k in (41, 44, 63):
        tgt = np.array([[k * sx, 0.0, 0.0]])
        v, form = OI.p_row_forms(src, tgt, sx, sy, sz, 1, 1)
        assert int(form[0]) % 100 == 18
        f1, A1 = _p_face(1, src, sx, sy, sz)
        f2, A2 = _p_face(1, tgt[0], sx, sy, sz)
        q = box_integral(f1, f2, n=24) / (4 * np.pi * eps0 * A1 * A2)
        assert -1.03 < v[0] / q < -1.0


def _p_face(caser, c, sx, sy, sz):
    """The face patch of voxel centre c for a caser code and its area
    (compute_row_P_sup.jl:3-51): 1/2 the -y/+y face, 3/4 the -x/+x face,
    5/6 the -z/+z face."""
    x, y, z = c
    if caser in (1, 2):
        return (iv(x, sx), pt(y - sy / 2 if caser == 1 else y + sy / 2), iv(z, sz)), sx * sz
    if caser in (3, 4):
        return (pt(x - sx / 2 if caser == 3 else x + sx / 2), iv(y, sy), iv(z, sz)), sy * sz
    return (iv(x, sx), iv(y, sy), pt(z - sz / 2 if caser == 5 else z + sz / 2)), sx * sy


# target offsets in cells: touching and near neighbours (the full closed
# forms), then 40..1500 cells away, where the suppression tests of
# Integ_sup_sup_P (compute_row_P_sup.jl:106-394) drop extents and the entry
# routes to a line/line, line/surface or point/point form
P_OFFSETS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 1, 0), (1, 2, 1), (3, 0, 2), (0, 3, 1), (40, 3, 1), (5, 60, 2),
             (3, 4, 150), (300, 20, 7), (20, 700, 3), (1500, 900, 40)]
P_PAIRS = [(1, 1), (3, 3), (5, 5), (3, 2), (5, 2), (5, 4), (2, 1), (4, 3), (6, 5), (1, 3), (1, 5), (3, 5)]
FULL_P_FORMS = (12, 13)  # qs_surf_surf_para, p_surf_surf_ortho


@pytest.mark.parametrize("pair", P_PAIRS)
def test_p_pair_rows_vs_quadrature(pair):
    """Every face pair of the P rows, at every routing distance, against the
    Gauss-Legendre integral over the two faces / (4 pi eps0 A1 A2).  Full
    surface-surface forms hold 1e-7 (the closed forms lose digits to
    cancellation at 40+ cells: measured <= 1.9e-8); coincident faces are the
    singular self terms, held at 1e-7 with a 64-point rule; entries that a
    suppression test routes to a line or point form hold the suppression
    error, 5e-6 (measured <= 2.1e-6).  A source face on the wrong side of
    its voxel (+s/2 for -s/2) moves the near entries of every pair by 26% or
    more (measured with the oracle's caser 2/4/6 rows against the 1/3/5
    faces), and a wrong area in the prefactor moves every entry by S/T."""
    c1, c2 = pair
    eps0 = 8.854187816997944e-12
    src = np.zeros(3)
    reached = set()
    for o in [(0, 0, 0)] + P_OFFSETS:
        cj = np.array([o[0] * S, o[1] * S, o[2] * T])
        f1, A1 = _p_face(c1, src, S, S, T)
        f2, A2 = _p_face(c2, cj, S, S, T)
        coincident = f1 == f2
        if o == (0, 0, 0) and not coincident:
            continue  # two faces of one voxel that share an edge: covered by the neighbour offsets
        v, form = OI.p_row_forms(src, cj[None, :], S, S, T, c1, c2)
        fid = int(form[0]) % 100
        assert int(form[0]) < 100, ("threshold tie", pair, o)
        assert fid not in OI.DEVIANT_FORMS, (pair, o, OI.FORM_NAMES[fid])
        reached.add(fid)
        q = box_integral(f1, f2, n=64 if coincident else 24) / (4 * np.pi * eps0 * A1 * A2)
        bar = 1e-7 if fid in FULL_P_FORMS else 5e-6
        assert rel(v[0], q) < bar, (pair, o, OI.FORM_NAMES[fid], rel(v[0], q))
    # the offsets reach both the full form of the pair and a suppressed one
    assert reached & set(FULL_P_FORMS) and reached - set(FULL_P_FORMS), reached


def test_far_field_branch_is_point_point():
    # very distant thin voxels take the sum_supp == 6 branch: V1 V2 / R
    src = np.array([0.0, 0.0, 0.0])
    far = np.array([[0.5, 0.3, 0.0]])
    s = (1e-4, 1e-4, 1e-5)
    row = OI.lp_row(src, far, *s, *s, 1)[0]
    V = s[0] * s[1] * s[2]
    expect = 1e-7 / (s[1] * s[2] * s[1] * s[2]) * V * V / np.linalg.norm(far[0])
    assert rel(row, expect) < 1e-11  # coordinates are rounded to 1e-14 m
    # and it is within the suppression error of the exact 64-corner value
    X1, Y1, Z1 = iv(0, s[0]), iv(0, s[1]), iv(0, s[2])
    X2, Y2, Z2 = iv(0.5, s[0]), iv(0.3, s[1]), iv(0.0, s[2])
    q = box_integral((X1, Y1, Z1), (X2, Y2, Z2), n=8)
    assert rel(V * V / np.linalg.norm(far[0]), q) < 1e-6
