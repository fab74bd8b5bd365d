"""Independent numerical pin for the closed-form 1/R integrals.

Every closed form of the reference (compute_Lp_Voxels.jl:449-805,
compute_row_P_sup.jl:396-590, compute_Lp_Voxels_QS.jl:695-730) is the double
integral of 1/|r1 - r2| over two axis-aligned sets (a box, a rectangle, a
segment or a point).  Writing r1 - r2 = d, the integral over each axis
reduces to a one-dimensional weight w(d) -- a trapezoid when both sets are
extended along that axis, a box function when one is, a Dirac when neither
is -- so

    I = \\int w_x(dx) w_y(dy) w_z(dz) / |d|   (over the continuous axes).

That integral is evaluated here with tensor Gauss-Legendre rules on the
piecewise-linear pieces of w, with a graded u^3 substitution on pieces that
end at d = 0 so that the 1/|d| corner singularity of touching sets is
integrated to full precision.  Nothing here shares code or formulas with the
closed forms.
"""
import numpy as np

_GL_CACHE = {}


def _gl(n):
    if n not in _GL_CACHE:
        _GL_CACHE[n] = np.polynomial.legendre.leggauss(n)
    return _GL_CACHE[n]


def _axis_weight(i1, i2):
    """Pieces of w(d) for d = x1 - x2, x1 in i1, x2 in i2.

    Returns ("dirac", d0) or ("pieces", [(d_lo, d_hi, w_lo, w_hi), ...])."""
    a1, a2 = min(i1), max(i1)
    b1, b2 = min(i2), max(i2)
    la, lb = a2 - a1, b2 - b1
    if la == 0 and lb == 0:
        return ("dirac", a1 - b1)
    if la == 0 or lb == 0:
        lo, hi = (a1 - b2, a1 - b1) if la == 0 else (a1 - b1, a2 - b1)
        return ("pieces", [(lo, hi, 1.0, 1.0)])
    # trapezoid: overlap length of [a1,a2] and [b1+d, b2+d]
    k = sorted([a1 - b2, a1 - b1, a2 - b2, a2 - b1])
    m = min(la, lb)
    pieces = []
    w_at = {k[0]: 0.0, k[1]: m, k[2]: m, k[3]: 0.0}
    for lo, hi in ((k[0], k[1]), (k[1], k[2]), (k[2], k[3])):
        if hi > lo:
            pieces.append((lo, hi, w_at[lo], w_at[hi]))
    return ("pieces", pieces)


def _split_at_zero(pieces):
    out = []
    for lo, hi, wl, wh in pieces:
        if lo < 0.0 < hi:
            w0 = wl + (wh - wl) * (0.0 - lo) / (hi - lo)
            out.append((lo, 0.0, wl, w0))
            out.append((0.0, hi, w0, wh))
        else:
            out.append((lo, hi, wl, wh))
    return out


def _nodes(piece, n, graded):
    lo, hi, wl, wh = piece
    u, gw = _gl(n)
    u = 0.5 * (u + 1.0)
    gw = 0.5 * gw
    L = hi - lo
    if graded and (lo == 0.0 or hi == 0.0):
        # cluster nodes at the end that touches 0: t = u^3
        if lo == 0.0:
            t = u ** 3
            jac = 3 * u ** 2
        else:
            t = 1.0 - u ** 3
            jac = 3 * u ** 2
        d = lo + L * t
        w = gw * jac * L
    else:
        d = lo + L * u
        w = gw * L
    wd = wl + (wh - wl) * (d - lo) / L
    return d, w * wd


def box_integral(box1, box2, n=24, beta=0.0):
    """\\int_{box1} \\int_{box2} 1/|r1-r2| for degenerate-or-not boxes, or
    with beta != 0 the retarded kernel exp(-j beta |r1-r2|) / |r1-r2|.

    box = ((x_lo, x_hi), (y_lo, y_hi), (z_lo, z_hi)); lo == hi collapses an
    axis (surface, segment or point)."""
    axes = [_axis_weight(box1[k], box2[k]) for k in range(3)]
    dirac = [a[1] for a in axes if a[0] == "dirac"]
    cont = [a[1] for a in axes if a[0] == "pieces"]
    d0sq = sum(v * v for v in dirac)
    graded = d0sq == 0.0
    cont = [_split_at_zero(p) if graded else p for p in cont]
    total = 0.0
    if not cont:
        r = np.sqrt(d0sq)
        return np.exp(-1j * beta * r) / r if beta else 1.0 / r
    import itertools
    for combo in itertools.product(*cont):
        grids = [_nodes(pc, n, graded) for pc in combo]
        ds = np.meshgrid(*[g[0] for g in grids], indexing="ij")
        ws = np.meshgrid(*[g[1] for g in grids], indexing="ij")
        r2 = d0sq + sum(d * d for d in ds)
        w = np.ones_like(r2)
        for wi in ws:
            w = w * wi
        with np.errstate(divide="ignore", invalid="ignore"):
            r = np.sqrt(r2)
            f = np.where(r2 > 0, w / r, 0.0)
            if beta:
                f = f * np.exp(-1j * beta * r)
        total += f.sum()
    return total
