"""Seeded geometries of the five benchmark configurations.

Each builder returns a dict with the material grid (0 = empty, k = material
k), the materials, cell size, port coordinates, frequencies and GMRES
tolerance.  Ports are given as coordinates (metres) and resolved to voxels
by the mesher the way find_voxels_port_pp.jl does (nearest voxel centre).

mesh() builds the operator the reference would build for the geometry:
the reference mesher (voxel_mesher, mesher_FFT.jl) on one grid per
material.  The adjoint workload carries a conductivity per voxel, which the
reference's per-material mesher has no place for; it is meshed by
mesher.mesh, whose conductor slots take per-voxel conductivities.
"""
import numpy as np
import scipy.ndimage as ndi

from . import mesher as _mesher
from . import voxel_mesher
from .mesher import Material

COPPER = 5.8e7


def crossbar_24x24x4():
    """24x24x4 grid, 0.5 x 0.5 x 0.1 mm cells, copper x-bar (y = 10..13) and
    y-bar (x = 10..13) in layers z = 0..1; one port from the -x end face to
    the +x end face of the x-bar; 10 frequencies 1e6..1e9 Hz; tol 1e-8."""
    Nx, Ny, Nz = 24, 24, 4
    cell = (0.5e-3, 0.5e-3, 0.1e-3)
    g = np.zeros((Nx, Ny, Nz), dtype=np.int64)
    g[:, 10:14, 0:2] = 1
    g[10:14, :, 0:2] = 1
    yc = 12 * cell[1]
    zc = 1 * cell[2]
    ports = [((0.0, yc, zc), (Nx * cell[0], yc, zc))]
    return dict(name="crossbar_24x24x4", grid=g, materials=[Material("copper", COPPER)], cell=cell,
                origin=(0.0, 0.0, 0.0), ports=ports, freqs=np.logspace(6, 9, 10), tol=1e-8)


def _largest_component(mask):
    lab, nlab = ndi.label(mask)  # 6-connectivity
    if nlab == 0:
        return np.zeros_like(mask)
    sizes = np.bincount(lab.reshape(-1))
    sizes[0] = 0
    return lab == np.argmax(sizes)


def _layered(Nx, Ny, Nz, cell, z_diel, p_cu, eps_r, tan_d, seed, sigma_range=None):
    rng = np.random.default_rng(seed)
    g = np.zeros((Nx, Ny, Nz), dtype=np.int64)
    g[:, :, :z_diel] = 2
    cu = rng.random((Nx, Ny, Nz - z_diel)) < p_cu
    cu = _largest_component(cu)
    g[:, :, z_diel:][cu] = 1
    mats = [Material("copper", COPPER), Material("dielectric", 0.0, eps_r, tan_d)]
    sigma = None
    if sigma_range is not None:
        sigma = np.where(g == 1, rng.uniform(sigma_range[0], sigma_range[1], g.shape), 0.0)
    return g, mats, sigma


def _port_to_slab(g, cell, vox, z_diel):
    """Port from a copper voxel (given by index) to the slab voxel under its
    column at the top dielectric layer (coordinates of the two centres)."""
    i, j, k = vox
    a = ((i + 0.5) * cell[0], (j + 0.5) * cell[1], (k + 0.5) * cell[2])
    b = ((i + 0.5) * cell[0], (j + 0.5) * cell[1], (z_diel - 0.5) * cell[2])
    return (a, b)


def _x_face_voxels(g):
    cu = np.argwhere(g == 1)
    lo = cu[np.lexsort((cu[:, 2], cu[:, 1], cu[:, 0]))][0]
    hi = cu[np.lexsort((cu[:, 2], cu[:, 1], -cu[:, 0]))][0]
    return lo, hi


def _x_face_ports(g, cell, z_diel):
    lo, hi = _x_face_voxels(g)
    return [_port_to_slab(g, cell, lo, z_diel), _port_to_slab(g, cell, hi, z_diel)]


def _centre(cell, vox):
    return tuple((v + 0.5) * c for v, c in zip(vox, cell))


def two_material_64x64x16(seed=0):
    """64x64x16, 0.25 x 0.25 x 0.05 mm; z = 0..11 dielectric (eps_r 4.4,
    tan d 0.02); z = 12..15 copper Bernoulli(0.35) reduced to its largest
    6-connected component; two ports on the component's -x and +x faces
    (each from the face voxel to the slab voxel below); 32 frequencies
    1e7..1e10 Hz; tol 1e-10."""
    Nx, Ny, Nz = 64, 64, 16
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = _layered(Nx, Ny, Nz, cell, 12, 0.35, 4.4, 0.02, seed)
    return dict(name="two_material_64x64x16", grid=g, materials=mats, cell=cell, origin=(0.0, 0.0, 0.0),
                ports=_x_face_ports(g, cell, 12), freqs=np.logspace(7, 10, 32), tol=1e-10)


def boundary_faces(g, mat=1):
    """Exposed faces of the voxels of material `mat`: (voxel, axis, side)
    for every face whose neighbour across it is not `mat` (outside the grid
    counts as not `mat`), in voxel order (x fastest), then axis, then side."""
    cu = g == mat
    out = []
    for axis in range(3):
        for side in (-1, 1):
            nb = np.zeros_like(cu)
            sl_dst = [slice(None)] * 3
            sl_src = [slice(None)] * 3
            if side < 0:
                sl_dst[axis], sl_src[axis] = slice(1, None), slice(None, -1)
            else:
                sl_dst[axis], sl_src[axis] = slice(None, -1), slice(1, None)
            nb[tuple(sl_dst)] = cu[tuple(sl_src)]
            for v in np.argwhere(cu & ~nb):
                out.append((tuple(int(t) for t in v), axis, side))
    out.sort(key=lambda f: (f[0][2], f[0][1], f[0][0], f[1], f[2]))
    return out


def batched_96x96x24(seed=0, nports=8):
    """96x96x24, same recipe with copper Bernoulli(0.3); 8 ports on boundary
    faces of the largest component, chosen uniformly by the seed among the
    component's exposed faces (distinct voxels); each port runs from the
    voxel owning the chosen face to the slab voxel under it; 128
    frequencies 1e6..2e10 Hz."""
    Nx, Ny, Nz = 96, 96, 24
    cell = (0.25e-3, 0.25e-3, 0.05e-3)
    g, mats, _ = _layered(Nx, Ny, Nz, cell, 18, 0.3, 4.4, 0.02, seed)
    rng = np.random.default_rng(seed + 1000)
    faces = boundary_faces(g)
    pick, seen = [], set()
    for f in rng.permutation(len(faces)):
        v = faces[int(f)][0]
        if v in seen:
            continue
        seen.add(v)
        pick.append(v)
        if len(pick) == nports:
            break
    pick.sort(key=lambda v: (v[2], v[1], v[0]))
    ports = [_port_to_slab(g, cell, v, 18) for v in pick]
    return dict(name="batched_96x96x24", grid=g, materials=mats, cell=cell, origin=(0.0, 0.0, 0.0), ports=ports,
                freqs=np.logspace(6, np.log10(2e10), 128), tol=1e-6)


def adjoint_64x64x16(seed=0):
    """The two-material 64x64x16 geometry with per-voxel copper conductivity
    uniform in [3e7, 6e7] S/m from the seed.  Port 1 spans the copper
    component from its -x face voxel to its +x face voxel (a conduction
    path, so Z_11 = R + j w L depends on the conductivity at first order);
    port 2 runs from the +x face voxel to the slab below it.  A trace-to-slab
    port 1 would make Z_11 a series capacitance whose relative sensitivity to
    one voxel's conductivity is ~1e-17, below what central differences can
    resolve in FP64."""
    d = two_material_64x64x16(seed)
    g, mats, sigma = _layered(64, 64, 16, d["cell"], 12, 0.35, 4.4, 0.02, seed, sigma_range=(3e7, 6e7))
    lo, hi = _x_face_voxels(g)
    ports = [(_centre(d["cell"], lo), _centre(d["cell"], hi)), _port_to_slab(g, d["cell"], hi, 12)]
    d.update(name="adjoint_64x64x16", sigma=sigma, ports=ports)
    return d


def large_192x192x48(seed=0):
    """192x192x48 cubic 0.1 mm cells; eps_r 3.5 slab z = 0..35, copper
    Bernoulli(0.25) in z = 36..47 (largest component); one port; 16
    frequencies 1e8..1e10 Hz."""
    Nx, Ny, Nz = 192, 192, 48
    cell = (0.1e-3, 0.1e-3, 0.1e-3)
    g, mats, _ = _layered(Nx, Ny, Nz, cell, 36, 0.25, 3.5, 0.0, seed)
    return dict(name="large_192x192x48", grid=g, materials=mats, cell=cell, origin=(0.0, 0.0, 0.0),
                ports=_x_face_ports(g, cell, 36)[:1], freqs=np.logspace(8, 10, 16), tol=1e-6)


CONFIGS = {
    "crossbar_24x24x4": crossbar_24x24x4,
    "two_material_64x64x16": two_material_64x64x16,
    "batched_96x96x24": batched_96x96x24,
    "adjoint_64x64x16": adjoint_64x64x16,
    "large_192x192x48": large_192x192x48,
}


def mesh(cfg, p_self_fn=None, renumber=True):
    """The operator inputs of a configuration: the reference mesher on one
    grid per material (its edges renumbered x-fastest in each direction,
    voxel_mesher.renumber_edges, unless renumber is false), or (per-voxel
    conductivities) mesher.mesh."""
    if cfg.get("sigma") is not None:
        return _mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"],
                            p_self_fn=p_self_fn, sigma_grid=cfg["sigma"])
    g = np.asarray(cfg["grid"])
    grids = [g == k + 1 for k in range(len(cfg["materials"]))]
    M = voxel_mesher.mesh_request(grids, cfg["materials"], cfg["cell"], cfg["origin"],
                                  [m.sigmar > 0 for m in cfg["materials"]], cfg["ports"],
                                  cfg.get("lumped_elements", ()), p_self_fn=p_self_fn)
    return voxel_mesher.renumber_edges(M)[0] if renumber else M
