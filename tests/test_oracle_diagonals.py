"""compute_diagonals (compute_diagonals.jl:1-134): the oracle's restatement
(oracle/diagonals.py) pinned by what the physics fixes, and the product
mesher's diagonal terms (the R, Cd and self-Lp every solve uploads) held to
it on the benchmark meshes.

Pins:
  - the two half-voxel resistances of an edge inside one conductor add up
    to the bar resistance l / (sigma W T);
  - the two half-voxel capacitances of an edge inside one dielectric, in
    series, are the parallel-plate capacitance eps0 (epsr - 1) W T / l;
  - the self inductance of an edge with one / two border ends is that of a
    bar 1.5 / 2 cells long: Lp_self(1.5 l) and Lp_self(2 l) equal the
    coincident vol-vol integral of such a bar (itself pinned by quadrature in
    test_oracle_integrals.py), 1e-7 / (W T)^2 times it;
  - border ends count through ceil(id / 100), so any material id 1..100
    marks one end.
"""
import numpy as np
import pytest

from oracle import diagonals as OD
from oracle import integrals as OI
from solver_amd import configs, mesher

EPS0 = 8.854187816997944e-12


def rel(a, b):
    return np.max(np.abs(np.asarray(a) - np.asarray(b))) / max(np.max(np.abs(np.asarray(b))), 1e-300)


def _mesh(cfg):
    return mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"])


def _oracle_of(M, materials):
    mats = [dict(sigmar=m.sigmar, epsr=m.epsr) for m in materials]
    sx, sy, sz = M.sx, M.sy, M.sz
    a, b = M.mx, M.mx + M.my
    return OD.compute_diagonals(M.escalings, mats, sx, sy, sz, M.li_mat[:a], M.li_mat[a:b], M.li_mat[b:],
                                M.li_border[:a], M.li_border[a:b], M.li_border[b:])


@pytest.mark.parametrize("name", ["crossbar_24x24x4", "two_material_64x64x16"])
def test_mesher_diagonals_match_oracle(name):
    cfg = configs.CONFIGS[name]()
    M = _mesh(cfg)
    D = _oracle_of(M, cfg["materials"])
    assert np.array_equal(M.diag_R, D["R"])
    assert np.array_equal(M.diag_Cd, D["Cd"])
    assert rel(M.diag_Lp, D["Lp"]) < 1e-14
    # every slot kind is exercised: conductor halves, dielectric halves, borders
    assert np.count_nonzero(D["R"][:, 2:]) and np.count_nonzero(D["R"][:, :2])
    if name == "two_material_64x64x16":
        assert np.count_nonzero(D["Cd"][:, :2]) and np.count_nonzero(D["Cd"][:, 2:])
    assert np.count_nonzero(D["fc_Lp"] > 0)


def _one_edge(materials, sx, sy, sz, mat_ids, border_ids, axis):
    e = [np.zeros((0, 2), dtype=np.int64)] * 3
    bd = [np.zeros((0, 2), dtype=np.int64)] * 3
    e = list(e)
    bd = list(bd)
    e[axis] = np.array([mat_ids], dtype=np.int64)
    bd[axis] = np.array([border_ids], dtype=np.int64)
    esc = dict(R=1.0, Cd=1.0, Lp=1.0)
    return OD.compute_diagonals(esc, materials, sx, sy, sz, *e, *bd)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_half_voxel_resistances_add_to_the_bar(axis):
    sx, sy, sz = 0.25e-3, 0.3e-3, 0.05e-3
    sig = 5.8e7
    D = _one_edge([dict(sigmar=sig, epsr=1.0)], sx, sy, sz, (1, 1), (0, 0), axis)
    l = (sx, sy, sz)[axis]
    A = sx * sy * sz / l
    assert abs(D["R"][0, 0] + D["R"][0, 1] - l / (sig * A)) <= 1e-15 * l / (sig * A)
    assert np.all(D["Cd"] == 0) and np.all(D["R"][0, 2:] == 0)
    # a conductor with a relative permittivity also gets the capacitances
    D = _one_edge([dict(sigmar=sig, epsr=3.0)], sx, sy, sz, (1, 1), (0, 0), axis)
    c = D["Cd"][0, :2]
    assert abs(1 / (1 / c[0] + 1 / c[1]) - EPS0 * 2.0 * A / l) <= 1e-14 * abs(EPS0 * 2.0 * A / l)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_half_voxel_capacitances_in_series_are_the_plate(axis):
    sx, sy, sz = 0.25e-3, 0.3e-3, 0.05e-3
    epsr = 4.4 - 1j * 4.4 * 0.02
    D = _one_edge([dict(sigmar=0.0, epsr=epsr)], sx, sy, sz, (1, 1), (1, 0), axis)
    l = (sx, sy, sz)[axis]
    A = sx * sy * sz / l
    c = D["Cd"][0]
    plate = EPS0 * (epsr - 1) * A / l
    assert abs(1 / (1 / c[0] + 1 / c[1]) - plate) <= 1e-14 * abs(plate)
    assert c[2] == c[0] and c[3] == 0 and np.all(D["R"] == 0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_border_ends_lengthen_the_self_inductance(axis):
    sx, sy, sz = 0.25e-3, 0.3e-3, 0.05e-3
    l = (sx, sy, sz)[axis]
    W, T = ((sy, sz), (sx, sz), (sy, sx))[axis]
    mats = [dict(sigmar=5.8e7, epsr=1.0)] * 100
    for borders, cells in (((0, 0), 1.0), ((7, 0), 1.5), ((0, 100), 1.5), ((1, 42), 2.0)):
        D = _one_edge(mats, sx, sy, sz, (1, 1), borders, axis)
        L = cells * l
        # the bar of length L against itself, through the quadrature-pinned vol-vol form
        box = [-L / 2, L / 2, -W / 2, W / 2, -T / 2, T / 2]
        vv = 1e-7 / (W * T) ** 2 * OI.closed_form(0, np.array(box + box))
        assert abs(D["Lp"][0] - vv) <= 1e-12 * vv, (borders, D["Lp"][0], vv)
        assert abs(D["fc_Lp"][0] - (vv - OI.lp_self(l, W, T))) <= 1e-12 * vv
