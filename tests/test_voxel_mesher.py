"""The reference mesher (mesher_FFT.jl and the create_* files) and the
request wire formats, on the CPU.

The product mesher (solver_amd/voxel_mesher.py, vectorised) is held to the
oracle's loop-for-loop restatement (oracle/mesher.py) on the reference's
three request fixtures (tests/golden/requests/, copies of the reference's
first_run_data.json, first_run_data_2_materials.json and crossbar.json)
and on seeded random multi-material grids, overlapping grids included:
node lattice ids, A, Gamma, the nine P expansion maps, the Lp maps, the
edge materials and borders, R, Cd and Zs equal bit for bit; the self
inductances (two implementations of the closed form) to 1e-15.

Neither restatement can be run against the reference (Julia is absent);
what pins them beyond their agreement is worked out by hand from the
reference's rules on small boxes below: node positions on the lattice,
edge and surface counts, the single-layer surface rule, the interface node
sharing of a conductor on a dielectric, and the skin-effect terms.
"""
import json
import os

import numpy as np
import pytest

from oracle import mesher as OM
from solver_amd import request as RQ
from solver_amd import voxel_mesher as VM
from solver_amd.mesher import Material

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "requests")
FIXTURES = ["first_run_data", "first_run_data_2_materials", "crossbar"]


def dense(a):
    return a.toarray() if hasattr(a, "toarray") else np.asarray(a)


def oracle_of(grids, materials, cell, origin, dominant, ports=(), les=()):
    mats = [dict(sigmar=m.sigmar, eps_re=m.eps_re, tan_D=m.tan_D) for m in materials]
    dom = [k + 1 for k in range(len(grids)) if dominant[k]]
    return OM.mesh(grids, mats, *cell, origin, dom, ports, les)


def assert_same(M, O):
    assert np.array_equal(M.nodes_red, O["nodes_red"].astype(np.int64))
    for name in ("A", "Gamma"):
        assert np.array_equal(dense(getattr(M, name)), dense(O[name])), name
    for a in range(3):
        assert np.array_equal(dense(M.mat_map_Lp[a]), dense(O["mat_map_Lp"][a])), a
        for b in range(3):
            assert np.array_equal(dense(M.exp_P[a][b]), dense(O["exp_P"][a][b])), (a, b)
    assert np.array_equal(M.li_mat, np.concatenate(O["li_mat"]))
    assert np.array_equal(M.li_border, np.concatenate(O["li_border"]))
    assert (M.N1, M.N2, M.N3) == (O["N1"], O["N2"], O["N3"])
    assert np.array_equal(M.Zs, O["Zs"])
    assert np.array_equal(M.diag_R, O["diag_R"])
    assert np.array_equal(M.diag_Cd, O["diag_Cd"])
    assert np.allclose(M.diag_Lp, O["diag_Lp"], rtol=1e-15, atol=0)
    assert np.array_equal(M.port_nodes, O["port_nodes"])
    assert [tuple(e["nodes"]) for e in M.lumped_elements] == [tuple(e["nodes"]) for e in O["lumped_elements"]]


@pytest.mark.parametrize("name", FIXTURES)
def test_request_fixtures_product_equals_oracle(name):
    R = RQ.parse(os.path.join(GOLDEN, name + ".json.gz"))
    M = RQ.mesh_of(R)
    O = oracle_of(R.grids, R.materials, R.cell, R.origin, R.dominant, R.ports, R.lumped_elements)
    assert_same(M, O)
    # every edge joins two distinct nodes, every surface sits on one node
    A = M.A.tocsr()
    assert np.all(np.diff(A.indptr) == 2) and np.all(np.asarray(A.sum(axis=1)).ravel() == 0)
    assert np.all(np.asarray(M.Gamma.sum(axis=0)).ravel() == 1)
    assert M.ns == M.N1 + M.N2 + M.N3


def random_case(seed):
    rng = np.random.default_rng(seed)
    K = int(rng.integers(1, 4))
    shape = tuple(int(v) for v in rng.integers(1, 7, size=3))
    owner = rng.integers(0, K + 1, size=shape)
    grids = [owner == k + 1 for k in range(K)]
    if K > 1 and rng.random() < 0.3:
        grids[0] = grids[0] | (rng.random(shape) < 0.2)  # a voxel in two grids
    mats = [Material("m%d" % k, sigma=(5.8e7, 0.0, 1e6)[k % 3], eps_re=(1.0, 4.4, 2.0)[k % 3],
                     tan_d=(2.7, 0.02, 0.0)[k % 3]) for k in range(K)]
    cell = (1e-3, 2e-3, 0.5e-3)
    full = np.argwhere(np.logical_or.reduce(grids))
    ports, les = [], []
    if len(full) > 1:
        for _ in range(2):
            a, b = full[rng.integers(len(full))], full[rng.integers(len(full))]
            ports.append((tuple((a + 0.5) * cell), tuple((b + 0.5) * cell)))
        les = [dict(start=ports[0][0], end=ports[0][1], type=1, R=50.0, L=0.0, C=0.0)]
    return grids, mats, cell, [m.sigmar > 0 for m in mats], ports, les


@pytest.mark.parametrize("seed", range(40))
def test_random_grids_product_equals_oracle(seed):
    grids, mats, cell, dom, ports, les = random_case(seed)
    if not np.any(np.logical_or.reduce(grids)):
        pytest.skip("empty grid")
    M = VM.mesh_request(grids, mats, cell, (0.0, 0.0, 0.0), dom, ports, les)
    O = oracle_of(grids, mats, cell, (0.0, 0.0, 0.0), dom, ports, les)
    assert_same(M, O)


CU = Material("cu", sigma=5.8e7, eps_re=1.0, tan_d=0.0)
DIEL = Material("fr4", sigma=0.0, eps_re=4.4, tan_d=0.02)


def lattice(i, j, k, Nx, Ny):
    """1-based lattice id of the 0-based lattice point (i, j, k)."""
    return 1 + i + 3 * Nx * j + 9 * Nx * Ny * k


def test_full_copper_box_by_hand():
    """3 x 2 x 2 copper: every voxel lies on the box in y and z, the end
    ones also in x, so each node moves to the corner or edge point of its
    boundary faces; 8 x-, 6 y- and 6 z-edges; the faces are the 32 box
    faces (12 y-normal, 8 x-normal, 12 z-normal)."""
    g = np.ones((3, 2, 2), dtype=bool)
    M = VM.mesh_request([g], [CU], (1e-3, 1e-3, 1e-3), (0, 0, 0), [True])
    expect = set()
    for i in range(3):
        for j in range(2):
            for k in range(2):
                dx = -1 if i == 0 else 1 if i == 2 else 0
                dy = -1 if j == 0 else 1
                dz = -1 if k == 0 else 1
                expect.add(lattice(2 * i + 1 + dx, 2 * j + 1 + dy, 2 * k + 1 + dz, 3, 2))
    assert set(M.nodes_red.tolist()) == expect
    assert (M.n, M.mx, M.my, M.mz) == (12, 8, 6, 6)
    assert (M.N1, M.N2, M.N3) == (12, 8, 12)


def test_single_layer_has_no_faces_normal_to_it():
    """A slab one voxel thick in y has both y faces on its boundary, so the
    reference's first and third blocks (create_mapping_Gamma_no_rep.jl) make
    no y-normal surface for it."""
    g = np.ones((3, 1, 2), dtype=bool)
    M = VM.mesh_request([g], [CU], (1e-3, 1e-3, 1e-3), (0, 0, 0), [True])
    assert M.N1 == 0 and M.N2 == 4 and M.N3 == 6


def test_conductor_on_dielectric_shares_interface_nodes():
    """A copper layer (z = 3) on three dielectric layers (z = 0..2), 4 x 1 in
    x, y.  The copper voxels' -z faces touch the dielectric, whose +z faces
    are on its boundary, so the copper nodes move down onto the shared face
    points (create_nodes_ref.jl:84-116; the contact test on the lower face
    runs from the third layer on).  The top dielectric voxels find those
    nodes among their 26 lattice neighbours and take them: 4 shared nodes
    and 4 for each lower dielectric layer (12), 12 x edges (3 per layer),
    8 z edges inside the dielectric and none between the materials.

    With the copper on the second of two layers the same contact test does
    not run (index 2), the copper nodes stay off the interface and nothing
    is shared."""
    def build(nz, cu_z):
        cu = np.zeros((4, 1, nz), dtype=bool)
        cu[:, :, cu_z] = True
        d = np.zeros((4, 1, nz), dtype=bool)
        d[:, :, :cu_z] = True
        return VM.mesh_request([d, cu], [DIEL, CU], (1e-3, 1e-3, 1e-3), (0, 0, 0), [False, True])

    M = build(4, 3)
    vn = M.vox_node.reshape(4, 4)  # voxels z-major: rows z = 0..3
    assert np.array_equal(vn[2], vn[3])
    assert M.n == 12 and (M.mx, M.my, M.mz) == (12, 0, 8)
    # the shared nodes are the face points under the copper voxels (the x
    # ends also on their x faces): lattice z index 2*3 + 1 - 1 = 6
    z_of = (M.nodes_red[vn[3]] - 1) // (9 * 4 * 1)
    assert np.all(z_of == 6)
    M2 = build(2, 1)
    assert M2.n == 8 and not set(M2.vox_node[:4]) & set(M2.vox_node[4:])


def test_skin_terms_by_hand():
    """A 3 x 1 x 1 copper bar: the two x edges touch the box on every side,
    so each half edge has both face counts positive, the bar ends are
    borders (doubling the end halves), and Zs of an edge is the two halves'
    parallel face impedances summed (compute_Zs_with_indices.jl)."""
    sx, sy, sz = 1e-3, 2e-3, 0.5e-3
    g = np.ones((3, 1, 1), dtype=bool)
    M = VM.mesh_request([g], [CU], (sx, sy, sz), (0, 0, 0), [True])
    assert M.mx == 2 and M.my == 0 and M.mz == 0
    base = (1 + 1j) * np.sqrt(2 * np.pi * 1e-7 / 5.8e7)
    za, zb = 0.5 * sx / sy * base, 0.5 * sx / sz * base

    def par(a, b):
        return a * b / (a + b)

    # edge 0: lower half on the -x border (doubled), upper half inside;
    # edge 1: the mirror image
    e0 = par(2 * za, 2 * zb) + par(za, zb)
    e1 = par(za, zb) + par(2 * za, 2 * zb)
    assert np.allclose(M.Zs, [e0, e1], rtol=1e-15)
    assert np.array_equal(M.li_border, [[1, 0], [0, 1]])


def test_request_parsing_and_result_format():
    R = RQ.parse(os.path.join(GOLDEN, "first_run_data_2_materials.json.gz"))
    # grids pair with materials by name; the lengths take the unit (mm),
    # the origin does not
    assert R.names == ["AntennaDielMaterial", "AntennaCondMaterial"]
    assert [m.name for m in R.materials] == R.names
    assert R.dominant == [False, True]
    assert np.allclose(R.cell, (0.5e-3, 2.5e-3, 0.024444444974263513e-3))
    assert R.materials[0].epsr == 5 - 1j * 5 * 2.7
    assert len(R.lumped_elements) == 2 and R.lumped_elements[0]["R"] == 50.0
    assert np.allclose(R.ports[0][1], (1.5e-3, 0.0, 1.1e-3))
    assert (R.inner_iter, R.tol, R.R0) == (100, 1e-4, 50.0)
    P, F = 2, 3
    rng = np.random.default_rng(1)
    Z, S, Y = (rng.normal(size=(P, P, F)) + 1j * rng.normal(size=(P, P, F)) for _ in range(3))
    out = RQ.dump_json_data(Z, S, Y, P, "sim-1")
    assert set(out["matrices"]) == {"matrix_Z", "matrix_S", "matrix_Y"}
    assert out["id"] == "sim-1" and out["partial"] is False and out["isStopped"] is False
    # [[ [re, im] per frequency ]] per port pair, (1, 1), (1, 2), (2, 1), (2, 2)
    raw = json.loads(out["matrices"]["matrix_Z"])
    assert len(raw) == P * P and raw[1][0][2] == [Z[0, 1, 2].real, Z[0, 1, 2].imag]
    for key, A in (("matrix_Z", Z), ("matrix_S", S), ("matrix_Y", Y)):
        assert np.array_equal(RQ.parse_matrix(out["matrices"][key], P, F), A)


def test_request_rejects_unlisted_material():
    d = RQ.load(os.path.join(GOLDEN, "first_run_data.json.gz"))
    d["solverInput"]["materials"][0]["name"] = "other"
    with pytest.raises(ValueError):
        RQ.parse(d)


def test_renumbered_edges_are_the_same_operator_in_x_fastest_order():
    """voxel_mesher.renumber_edges (applied by configs.mesh) only relabels
    the edges: each direction's edges keep their direction group and move
    to x-fastest order of their lower node; A's rows, every per-edge array
    and the Lp position map's columns move together, so A'A, the set of
    (Lp position, edge data) pairs and the ports are unchanged."""
    import copy
    import scipy.sparse as sp
    from solver_amd import configs
    cfg = configs.two_material_64x64x16()
    g = np.asarray(cfg["grid"])[:24, :20, :8]
    grids = [g == k + 1 for k in range(len(cfg["materials"]))]
    M0 = VM.mesh_request(grids, cfg["materials"], cfg["cell"], cfg["origin"],
                         [m.sigmar > 0 for m in cfg["materials"]], [])
    M, perm = VM.renumber_edges(copy.deepcopy(M0))
    assert sorted(perm.tolist()) == list(range(M0.m))
    b = (0, M0.mx, M0.mx + M0.my, M0.m)
    for d in range(3):
        assert np.all((perm[b[d]:b[d + 1]] >= b[d]) & (perm[b[d]:b[d + 1]] < b[d + 1]))
    A0, A1 = sp.csr_matrix(M0.A), sp.csr_matrix(M.A)
    assert (A1 - A0[perm]).nnz == 0
    assert abs((A0.T @ A0) - (A1.T @ A1)).max() == 0
    for k in ("Zs", "diag_Lp", "diag_R", "diag_Cd", "li_border", "li_mat"):
        assert np.array_equal(np.asarray(getattr(M, k)), np.asarray(getattr(M0, k))[perm]), k
    # the surface edges are the same edges under their new numbers
    assert np.array_equal(np.sort(perm[M.surface_edges]), M0.surface_edges)
    ijk = np.asarray(M.node_ijk)
    for d in range(3):
        P0 = sp.csc_matrix(M0.mat_map_Lp[d])
        P1 = sp.csc_matrix(M.mat_map_Lp[d])
        local = perm[b[d]:b[d + 1]] - b[d]
        assert (P1 - P0[:, local]).nnz == 0
        # x-fastest: the lower nodes' lattice keys increase along the group
        lower = np.array([A1.indices[A1.indptr[e]:A1.indptr[e + 1]].min() for e in range(b[d], b[d + 1])])
        key = ijk[lower][:, 2] * 10 ** 8 + ijk[lower][:, 1] * 10 ** 4 + ijk[lower][:, 0]
        assert np.all(np.diff(key) > 0)
