"""Seeded benchmark geometries (CPU): the recipes the workloads are named
for, checked on the generated grids themselves."""
import numpy as np
import scipy.ndimage as ndi

from solver_amd import configs


def _vox(cell, p):
    return tuple(int(np.floor(c / s)) for c, s in zip(p, cell))


def test_layered_copper_is_one_6_connected_component():
    d = configs.two_material_64x64x16(0)
    g = d["grid"]
    assert g.shape == (64, 64, 16)
    assert np.all(g[:, :, :12] == 2)                      # dielectric slab z = 0..11
    cu = g == 1
    assert cu[:, :, :12].sum() == 0 and cu.sum() > 0      # copper only in z = 12..15
    _, nlab = ndi.label(cu)
    assert nlab == 1


def test_adjoint_ports_span_the_component_and_reach_the_slab():
    """Port 1 joins the copper component's -x and +x face voxels (a
    conduction path, so |Z_11|^2 depends on the conductivity at first
    order); port 2 joins the +x face voxel to the slab voxel below it."""
    d = configs.adjoint_64x64x16(0)
    g, cell = d["grid"], d["cell"]
    (a1, b1), (a2, b2) = d["ports"]
    va, vb = _vox(cell, a1), _vox(cell, b1)
    cu = np.argwhere(g == 1)
    assert g[va] == 1 and g[vb] == 1
    assert va[0] == cu[:, 0].min() and vb[0] == cu[:, 0].max()
    assert _vox(cell, a2) == vb
    w = _vox(cell, b2)
    assert g[w] == 2 and w[:2] == vb[:2] and w[2] == 11   # top slab layer under the port
    sig = d["sigma"]
    assert np.all((sig > 0) == (g == 1))
    assert sig[g == 1].min() >= 3e7 and sig[g == 1].max() <= 6e7


def test_batched_ports_are_seeded_and_on_copper():
    d1, d2 = configs.batched_96x96x24(0), configs.batched_96x96x24(0)
    assert d1["ports"] == d2["ports"] and len(d1["ports"]) == 8
    g, cell = d1["grid"], d1["cell"]
    for a, b in d1["ports"]:
        assert g[_vox(cell, a)] == 1 and g[_vox(cell, b)] == 2


def test_batched_ports_sit_on_exposed_faces_of_the_component():
    """Every port's copper voxel owns at least one exposed face of the
    largest component (a face whose neighbour is not copper), the voxels
    are distinct, and the slab end is the top dielectric layer under it."""
    d = configs.batched_96x96x24(0)
    g, cell = d["grid"], d["cell"]
    faces = configs.boundary_faces(g)
    owners = {f[0] for f in faces}
    vox = [_vox(cell, a) for a, _ in d["ports"]]
    assert len(set(vox)) == 8
    for (a, b), v in zip(d["ports"], vox):
        assert v in owners
        w = _vox(cell, b)
        assert w[:2] == v[:2] and w[2] == 17 and g[w] == 2


def test_boundary_faces_of_a_single_voxel_and_a_bar():
    g = np.zeros((3, 3, 3), dtype=np.int64)
    g[1, 1, 1] = 1
    assert len(configs.boundary_faces(g)) == 6
    g[:, 1, 1] = 1          # a 3-voxel bar along x: 3 * 4 side faces + 2 end faces
    f = configs.boundary_faces(g)
    assert len(f) == 14
    assert sum(1 for v, ax, s in f if ax == 0) == 2
