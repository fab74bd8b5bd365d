"""ORACLE -- test infrastructure only (see oracle/__init__.py).

Loop-for-loop restatement of the reference's voxel mesher for the FFT
solver: mesher_FFT.jl:1-112 and the files it calls.  It keeps the
reference's loop orders, 1-based indices and tie rules, and so its quirks:

  * the contact tests of create_nodes_ref.jl for the lower faces start at
    index 3 (`cont2 - 1 > 1`, :40,58,76), the upper ones at 2;
  * the skin-effect face counts of create_A_mats_and_find_borders_with_map_
    Zs.jl count a box face once per half edge and skip the opposite side
    (:34-37), and the z edges test the lower z level for their upper half
    (:347, :392, the second also recording into slot 3);
  * compute_Zs_with_indices.jl uses only whether a face count is positive.

It is slow (Python loops over voxels) and meant for the small request
fixtures and random grids the product mesher is checked against.

Grids are a list of boolean arrays g[k][x, y, z] (0-based storage, accessed
1-based here); materials are dicts with "sigmar", "eps_re", "tan_D" (as
material() of format_input_output_solver_functions.jl:38-64 builds them).
"""
import numpy as np
import scipy.sparse as sp

from . import diagonals as OD

EPS0 = 8.854187816997944e-12


def from_3D_to_1D(i, j, k, M, N):
    """From_3D_to_1D.jl:1-3 (1-based)."""
    return (k - 1) * M * N + (j - 1) * M + i


def bin_search(num, A):
    """bin_search.jl:1-7: 1-based position of num in the sorted A, 0 if absent."""
    i = int(np.searchsorted(A, num, side="left"))
    if i >= len(A) or A[i] != num:
        return 0
    return i + 1


class Grids:
    def __init__(self, grids):
        self.g = [np.asarray(x, dtype=bool) for x in grids]
        self.K = len(self.g)
        self.Nx, self.Ny, self.Nz = self.g[0].shape

    def __call__(self, k, i, j, l):
        """grids[k][i][j][l] with 1-based k, i, j, l."""
        return bool(self.g[k - 1][i - 1, j - 1, l - 1])


def create_volumes_mapping(G):
    """create_volumes_mapping.jl:1-22."""
    mapping = np.zeros(G.Nx * G.Ny * G.Nz)
    num = 0
    for c3 in range(1, G.Nz + 1):
        for c2 in range(1, G.Ny + 1):
            for c in range(1, G.Nx + 1):
                for k in range(1, G.K + 1):
                    if G(k, c, c2, c3):
                        num += 1
                        mapping[from_3D_to_1D(c, c2, c3, G.Nx, G.Ny) - 1] = num
                        break
    return mapping, num


def create_volume_centers(G, mapping, num, sx, sy, sz, origin):
    """create_volume_centers.jl:1-25."""
    centri = np.zeros((num, 3))
    id_mat = np.zeros(num, dtype=np.int64)
    for c3 in range(1, G.Nz + 1):
        for c2 in range(1, G.Ny + 1):
            for c in range(1, G.Nx + 1):
                for k in range(1, G.K + 1):
                    if G(k, c, c2, c3):
                        pos = int(mapping[from_3D_to_1D(c, c2, c3, G.Nx, G.Ny) - 1])
                        centri[pos - 1] = (origin[0] + sx * (c - 1) + sx / 2, origin[1] + sy * (c2 - 1) + sy / 2,
                                           origin[2] + sz * (c3 - 1) + sz / 2)
                        id_mat[pos - 1] = k
                        break
    return centri, id_mat


def create_grids_externals(G):
    """create_Grids_externals.jl:1-261: ext[k][f] (f = 1..6: -y, +y, -x, +x,
    -z, +z) is true where voxel (k) has face f on its material's boundary
    or on the box."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    ext = [[None] + [np.zeros((Nx, Ny, Nz), dtype=bool) for _ in range(6)] for _ in range(G.K)]

    def setf(k, f, i, j, l):
        ext[k - 1][f][i - 1, j - 1, l - 1] = True

    for c3 in range(1, Nz + 1):
        c = Nx
        for c2 in range(1, Ny + 1):
            for k in range(1, G.K + 1):
                if G(k, 1, c2, c3):
                    setf(k, 3, 1, c2, c3)
                    if Nx > 1 and not G(k, 2, c2, c3):
                        setf(k, 4, 1, c2, c3)
                if G(k, c, c2, c3):
                    setf(k, 4, c, c2, c3)
                    if Nx > 1 and not G(k, c - 1, c2, c3):
                        setf(k, 3, c, c2, c3)
    for c3 in range(1, Nz + 1):
        c2 = Ny
        for c in range(1, Nx + 1):
            for k in range(1, G.K + 1):
                if G(k, c, 1, c3):
                    setf(k, 1, c, 1, c3)
                    if Ny > 1 and not G(k, c, 2, c3):
                        setf(k, 2, c, 1, c3)
                if G(k, c, c2, c3):
                    setf(k, 2, c, c2, c3)
                    if Ny > 1 and not G(k, c, c2 - 1, c3):
                        setf(k, 1, c, c2, c3)
    for c2 in range(1, Ny + 1):
        c3 = Nz
        for c in range(1, Nx + 1):
            for k in range(1, G.K + 1):
                if G(k, c, c2, 1):
                    setf(k, 5, c, c2, 1)
                    if Nz > 1 and not G(k, c, c2, 2):
                        setf(k, 6, c, c2, 1)
                if G(k, c, c2, c3):
                    setf(k, 6, c, c2, c3)
                    if Nz > 1 and not G(k, c, c2, c3 - 1):
                        setf(k, 5, c, c2, c3)
    for c3 in range(1, Nz + 1):
        for c2 in range(1, Ny + 1):
            for c in range(1, Nx + 1):
                for k in range(1, G.K + 1):
                    if G(k, c, c2, c3):
                        if c != 1 and c != Nx:
                            if not G(k, c - 1, c2, c3):
                                setf(k, 3, c, c2, c3)
                            if not G(k, c + 1, c2, c3):
                                setf(k, 4, c, c2, c3)
                        if c2 != 1 and c2 != Ny:
                            if not G(k, c, c2 - 1, c3):
                                setf(k, 1, c, c2, c3)
                            if not G(k, c, c2 + 1, c3):
                                setf(k, 2, c, c2, c3)
                        if c3 != 1 and c3 != Nz:
                            if not G(k, c, c2, c3 - 1):
                                setf(k, 5, c, c2, c3)
                            if not G(k, c, c2, c3 + 1):
                                setf(k, 6, c, c2, c3)
    return ext


def _ext(ext, k, f, i, j, l):
    return bool(ext[k - 1][f][i - 1, j - 1, l - 1])


# lattice offset of a node moved to the faces (f1 .. f6) it lies on:
# f1 -y, f2 +y, f3 -x, f4 +x, f5 -z, f6 +z (create_nodes_ref.jl:117-196)
_FACE_OFF = {1: (0, -1, 0), 2: (0, 1, 0), 3: (-1, 0, 0), 4: (1, 0, 0), 5: (0, 0, -1), 6: (0, 0, 1)}


def _moved(is_f, c1, c2, c3, NX3, NY3):
    """The lattice node of create_nodes_ref.jl:117-196: the centre moved to
    the one face, edge or corner point of the set faces (at most one per
    axis); None for a combination the reference does not list."""
    on = [f for f in range(1, 7) if is_f[f]]
    if len(on) > 3:
        return None
    d = [0, 0, 0]
    for f in on:
        o = _FACE_OFF[f]
        d = [d[0] + o[0], d[1] + o[1], d[2] + o[2]]
    return float(from_3D_to_1D(c1 + d[0], c2 + d[1], c3 + d[2], NX3, NY3))


def _face_flags(G, ext, k, c, c2, c3):
    """is_f1 .. is_f6 of create_nodes_ref.jl:16-116 for voxel (c, c2, c3) of grid k."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    f = [None] + [_ext(ext, k, i, c, c2, c3) for i in range(1, 7)]
    fc = [None] + [False] * 6
    others = [k2 for k2 in range(1, G.K + 1) if k2 != k]
    if c2 - 1 > 1:
        fc[1] = any(_ext(ext, k2, 2, c, c2 - 1, c3) for k2 in others)
    if c2 + 1 <= Ny:
        fc[2] = any(_ext(ext, k2, 1, c, c2 + 1, c3) for k2 in others)
    if c - 1 > 1:
        fc[3] = any(_ext(ext, k2, 4, c - 1, c2, c3) for k2 in others)
    if c + 1 <= Nx:
        fc[4] = any(_ext(ext, k2, 3, c + 1, c2, c3) for k2 in others)
    if c3 - 1 > 1:
        fc[5] = any(_ext(ext, k2, 6, c, c2, c3 - 1) for k2 in others)
    if c3 + 1 <= Nz:
        fc[6] = any(_ext(ext, k2, 5, c, c2, c3 + 1) for k2 in others)
    is_f = list(f)
    for a, b in ((1, 2), (3, 4), (5, 6)):
        if f[a] and f[b]:
            is_f[a] = is_f[b] = False
            if fc[a] and not fc[b]:
                is_f[a] = True
            elif not fc[a] and fc[b]:
                is_f[b] = True
    return is_f


def _build_nodes(c1, c2, c3, NX3, NY3):
    """build_nodes of create_nodes_ref.jl:433-488: the 26 lattice neighbours,
    x offset slowest, then y, then z (the centre skipped)."""
    out = []
    for dy in (-1, 0, 1):
        for dx in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if dx == 0 and dy == 0 and dz == 0:
                    continue
                out.append(float(from_3D_to_1D(c1 + dx, c2 + dy, c3 + dz, NX3, NY3)))
    return out


def create_nodes_ref(G, num_full_vox, ext, mapping, dominant_list):
    """create_nodes_ref.jl:1-418 (dominant_list: 1-based grid ids of the
    conductors)."""
    Nx, Ny = G.Nx, G.Ny
    NX3, NY3 = 3 * Nx, 3 * Ny
    nodes = np.zeros(num_full_vox)
    lista = []
    for k in dominant_list:
        for c in range(1, Nx + 1):
            for c2 in range(1, Ny + 1):
                for c3 in range(1, G.Nz + 1):
                    if not G(k, c, c2, c3):
                        continue
                    c1_, c2_, c3_ = 2 * c, 2 * c2, 2 * c3
                    v = int(mapping[from_3D_to_1D(c, c2, c3, Nx, Ny) - 1]) - 1
                    is_f = _face_flags(G, ext, k, c, c2, c3)
                    if any(is_f[1:]):
                        node = _moved(is_f, c1_, c2_, c3_, NX3, NY3)
                        if node is not None:
                            nodes[v] = node
                            lista.append(node)
                    else:
                        nodes[v] = float(from_3D_to_1D(c1_, c2_, c3_, NX3, NY3))
    nodes_red_dom = np.unique(np.asarray(lista, dtype=np.float64))
    non_dom = [k for k in range(1, G.K + 1) if k not in dominant_list]
    reused = []
    for k in non_dom:
        for c in range(1, Nx + 1):
            for c2 in range(1, Ny + 1):
                for c3 in range(1, G.Nz + 1):
                    if not G(k, c, c2, c3):
                        continue
                    c1_, c2_, c3_ = 2 * c, 2 * c2, 2 * c3
                    v = int(mapping[from_3D_to_1D(c, c2, c3, Nx, Ny) - 1]) - 1
                    is_f = _face_flags(G, ext, k, c, c2, c3)
                    if any(is_f[1:]):
                        shared = 0.0
                        for cand in _build_nodes(c1_, c2_, c3_, NX3, NY3):
                            if bin_search(cand, nodes_red_dom) != 0:
                                shared = cand
                                break
                        if shared != 0.0:
                            nodes[v] = shared
                            reused.append(shared)
                        else:
                            node = _moved(is_f, c1_, c2_, c3_, NX3, NY3)
                            if node is not None:
                                nodes[v] = node
                    else:
                        nodes[v] = float(from_3D_to_1D(c1_, c2_, c3_, NX3, NY3))
    return nodes, np.unique(nodes), np.unique(np.asarray(reused, dtype=np.float64))


def _edge_loops(axis, G):
    """(cont, cont2, cont3) of the lower voxel of each candidate edge, in the
    loop order of create_mapping_A{x,y,z}_v2.jl (x: y, z, x; y: z, x, y;
    z: x, y, z; last fastest)."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    if axis == 0:
        for c2 in range(1, Ny + 1):
            for c3 in range(1, Nz + 1):
                for c in range(1, Nx):
                    yield c, c2, c3
    elif axis == 1:
        for c3 in range(1, Nz + 1):
            for c in range(1, Nx + 1):
                for c2 in range(1, Ny):
                    yield c, c2, c3
    else:
        for c in range(1, Nx + 1):
            for c2 in range(1, Ny + 1):
                for c3 in range(1, Nz):
                    yield c, c2, c3


def _edge_key(axis, c, c2, c3, G):
    if axis == 0:
        return from_3D_to_1D(c, c2, c3, G.Nx - 1, G.Ny)
    if axis == 1:
        return from_3D_to_1D(c, c2, c3, G.Nx, G.Ny - 1)
    return from_3D_to_1D(c, c2, c3, G.Nx, G.Ny)


def _up(axis, c, c2, c3, n=1):
    return (c + n, c2, c3) if axis == 0 else (c, c2 + n, c3) if axis == 1 else (c, c2, c3 + n)


def _node_of(G, mapping, nodes, c, c2, c3):
    return nodes[int(mapping[from_3D_to_1D(c, c2, c3, G.Nx, G.Ny) - 1]) - 1]


def create_mapping_A(axis, G, mapping, nodes, nodes_red):
    """create_mapping_A{x,y,z}_v2.jl: edge number of each lower voxel of an
    edge between two voxels of one grid whose nodes differ."""
    size = ((G.Nx - 1) * G.Ny * G.Nz, G.Nx * (G.Ny - 1) * G.Nz, G.Nx * G.Ny * (G.Nz - 1))[axis]
    amap = np.zeros(size, dtype=np.int64)
    num = 0
    for c, c2, c3 in _edge_loops(axis, G):
        u = _up(axis, c, c2, c3)
        for k in range(1, G.K + 1):
            if G(k, c, c2, c3) and G(k, *u):
                nn1 = bin_search(_node_of(G, mapping, nodes, c, c2, c3), nodes_red)
                nn2 = bin_search(_node_of(G, mapping, nodes, *u), nodes_red)
                if abs(nn1 - nn2) > 1e-8:
                    key = _edge_key(axis, c, c2, c3, G)
                    if amap[key - 1] == 0:
                        num += 1
                        amap[key - 1] = num
                    break
    return amap, num


def _a_loops(axis, G):
    """Loop order of create_A_mats_and_find_borders_with_map_Zs.jl (x: z, y,
    x; y: z, x, y; z: x, y, z)."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    if axis == 0:
        for c3 in range(1, Nz + 1):
            for c2 in range(1, Ny + 1):
                for c in range(1, Nx):
                    yield c, c2, c3
    else:
        yield from _edge_loops(axis, G)


def _first_other(G, k, i, j, l):
    for k2 in range(1, G.K + 1):
        if k2 != k and G(k2, i, j, l):
            return k2
    return None


def create_A_mats(G, mapping, amaps, nums, materials, nodes, nodes_red):
    """create_A_mats_and_find_borders_with_map_Zs.jl:1-425: A (m x n),
    li_mat and li_border per axis, and the skin-effect face maps (per axis
    two planes, columns: count of half 1, count of half 2, other material
    at half 1, other material at half 2)."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    li_mat, li_border, rows, cols, vals = [], [], [], [], []
    zs = {}
    names = (("x_xy", "x_zx"), ("y_xy", "y_yz"), ("z_zx", "z_yz"))
    off = 0
    for axis in range(3):
        NA = nums[axis]
        lm = np.zeros((NA, 2), dtype=np.int64)
        lb = np.zeros((NA, 2), dtype=np.int64)
        zA = np.zeros((NA, 4), dtype=np.int64)
        zB = np.zeros((NA, 4), dtype=np.int64)
        for c, c2, c3 in _a_loops(axis, G):
            u = _up(axis, c, c2, c3)
            for k in range(1, G.K + 1):
                if not (G(k, c, c2, c3) and G(k, *u)):
                    continue
                n1 = _node_of(G, mapping, nodes, c, c2, c3)
                n2 = _node_of(G, mapping, nodes, *u)
                if not abs(n1 - n2) > 1e-8:
                    continue
                nn1, nn2 = bin_search(n1, nodes_red), bin_search(n2, nodes_red)
                pos = int(amaps[axis][_edge_key(axis, c, c2, c3, G) - 1])
                rows += [off + pos, off + pos]
                cols += [nn1, nn2]
                vals += [-1.0, 1.0]
                lm[pos - 1] = (k, k)
                if materials[k - 1]["sigmar"] != 0:
                    _skin_counts(axis, G, k, c, c2, c3, zA[pos - 1], zB[pos - 1])
                # borders: the half voxel beyond each end of the edge
                lo = {0: c, 1: c2, 2: c3}[axis]
                Nax = (Nx, Ny, Nz)[axis]
                if lo > 1:
                    if not G(k, *_up(axis, c, c2, c3, -1)):
                        lb[pos - 1, 0] = k
                else:
                    lb[pos - 1, 0] = k
                if lo + 1 == Nax:
                    lb[pos - 1, 1] = k
                elif lo + 2 <= Nax:
                    if not G(k, *_up(axis, c, c2, c3, 2)):
                        lb[pos - 1, 1] = k
                break
        zs[names[axis][0]] = zA
        zs[names[axis][1]] = zB
        li_mat.append(lm)
        li_border.append(lb)
        off += NA
    m = off
    A = sp.csc_matrix((vals, (np.asarray(rows, dtype=np.int64) - 1, np.asarray(cols, dtype=np.int64) - 1)),
                      shape=(m, len(nodes_red)))
    return A, li_mat, li_border, zs


def _skin_counts(axis, G, k, c, c2, c3, zA, zB):
    """The face counts of one conductor edge: plane A (normal to the next
    axis of (x: z, y: z, z: x)) then plane B, with the reference's index
    choices, including the z-edge ones that look at the lower z level."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    if axis == 0:
        planes = (("z", c3, Nz), ("y", c2, Ny))
    elif axis == 1:
        planes = (("z", c3, Nz), ("x", c, Nx))
    else:
        planes = (("x", c, Nx), ("y", c2, Ny))
    for pi, (nrm, idx, Nn) in enumerate(planes):
        z = zA if pi == 0 else zB
        if idx == 1 or idx == Nn:
            z[0] += 1
            z[1] += 1
            continue

        def at(half, d):
            # voxel of the given half (0 lower, 1 upper along the edge axis)
            # moved by d along the plane normal
            p = [c, c2, c3]
            p[axis] += half
            p[{"x": 0, "y": 1, "z": 2}[nrm]] += d
            return p

        # neighbour below (d = -1) then above (d = +1)
        for d in (-1, 1):
            if (d == -1 and idx > 1) or (d == 1 and idx < Nn):
                p1 = at(0, d)
                if not G(k, *p1):
                    z[0] += 1
                    o = _first_other(G, k, *p1)
                    if o is not None:
                        z[2] = o
                # the upper half: the z edges test the lower z level on
                # their +x / +y side (:347, :392), and the +y one records
                # into slot 3 (:392)
                if axis == 2 and d == 1:
                    p2 = at(0, d)
                    slot = 3 if nrm == "x" else 2
                else:
                    p2 = at(1, d)
                    slot = 3
                if not G(k, *p2):
                    z[1] += 1
                    o = _first_other(G, k, *p2)
                    if o is not None:
                        z[slot] = o


def compute_Zs(materials, li_mat, li_border, zs, sx, sy, sz):
    """compute_Zs_with_indices.jl:1-411: the surface impedance per edge
    (frequency factor sqrt(omega) applied by the solver) and the surface
    edges (1-based in the reference; 0-based here)."""
    ratios = (0.5 * sx / sy, 0.5 * sx / sz, 0.5 * sy / sx, 0.5 * sy / sz, 0.5 * sz / sx, 0.5 * sz / sy)
    planes = (("x_xy", "x_zx", 0, 1), ("y_xy", "y_yz", 2, 3), ("z_zx", "z_yz", 4, 5))
    out = []
    for axis in range(3):
        pa, pb, ra, rb = planes[axis]
        lm, lb = li_mat[axis], li_border[axis]
        NA = lm.shape[0]
        Zsw = np.zeros((NA, 4), dtype=np.complex128)
        for cont, mat in enumerate(materials, start=1):
            if mat["sigmar"] == 0:
                continue
            base = (1 + 1j) * np.sqrt(2 * np.pi * 1e-7 / mat["sigmar"])
            ca, cb = ratios[ra] * base, ratios[rb] * base
            on = lm[:, 0] == cont
            c1 = on & (zs[pa][:, 0] > 0)
            c3 = on & (zs[pa][:, 1] > 0)
            c2 = on & (zs[pb][:, 0] > 0)
            c4 = on & (zs[pb][:, 1] > 0)
            Zsw[c1, 0] = ca
            Zsw[c3, 2] = ca
            Zsw[c2, 1] = cb
            Zsw[c4, 3] = cb
            b1 = lb[:, 0] == cont
            b2 = lb[:, 1] == cont
            Zsw[c1 & b1, 0] *= 2
            Zsw[c2 & b1, 1] *= 2
            Zsw[c3 & b2, 2] *= 2
            Zsw[c4 & b2, 3] *= 2
        z1, z2, z3, z4 = (Zsw[:, i].copy() for i in range(4))

        def par(a, b):
            s = a + b
            both = (np.abs(a) != 0) & (np.abs(b) != 0)
            s[both] = a[both] * b[both] / (a[both] + b[both])
            return s

        out.append(par(z1, z2) + par(z3, z4))
    Zs = np.concatenate(out)
    return Zs, np.nonzero(np.abs(Zs) != 0)[0]


def create_expansion_ind_Lp(axis, G, amap, nA, mapping, nodes, nodes_red):
    """create_expansion_ind_Lp_{x,y,z}_grids_v2.jl: (Lp grid position,
    edge number) per edge, in the loop order x: z, y, x; y: z, x, y; z: x,
    y, z (1-based)."""
    out = []
    for c, c2, c3 in _a_loops(axis, G):
        u = _up(axis, c, c2, c3)
        for k in range(1, G.K + 1):
            if G(k, c, c2, c3) and G(k, *u):
                nn1 = bin_search(_node_of(G, mapping, nodes, c, c2, c3), nodes_red)
                nn2 = bin_search(_node_of(G, mapping, nodes, *u), nodes_red)
                if abs(nn1 - nn2) > 1e-8:
                    key = _edge_key(axis, c, c2, c3, G)
                    out.append((key, int(amap[key - 1])))
                    break
    return np.asarray(out, dtype=np.int64).reshape(-1, 2)


def create_mapping_Gamma_no_rep(G, mapping, nodes, nodes_red, ext):
    """create_mapping_Gamma_no_rep.jl:1-421: the surfaces (y-normal 12, then
    x-normal 34, then z-normal 56; four blocks each), Gamma (n x ns) and the
    expansion maps exp_P[a][b] of the P grids (0-based [a][b] = the
    reference's [a+1, b+1])."""
    Nx, Ny, Nz = G.Nx, G.Ny, G.Nz
    NX1, NY1, NZ1 = 2 * (Nx + 1) - 1, 2 * (Ny + 1) - 1, 2 * (Nz + 1) - 1
    m12 = np.zeros(Nx * (Ny + 1) * Nz, dtype=np.int64)
    m34 = np.zeros((Nx + 1) * Ny * Nz, dtype=np.int64)
    m56 = np.zeros(Nx * Ny * (Nz + 1), dtype=np.int64)
    m1234 = np.zeros(NX1 * NY1 * Nz, dtype=np.int64)
    m1256 = np.zeros(Nx * NY1 * NZ1, dtype=np.int64)
    m3456 = np.zeros(NX1 * Ny * NZ1, dtype=np.int64)
    gr, gc = [], []
    cnt = {"12": 0, "34": 0, "56": 0}

    def other_at(k, i, j, l, valid):
        return valid and any(G(k2, i, j, l) for k2 in range(1, G.K + 1) if k2 != k)

    def record(t, k, c, c2, c3, face):
        # face: the surface's own index along the normal (c2 or c2 + 1 ..)
        cnt[t] += 1
        if t == "12":
            sid = cnt["12"]
            m12[from_3D_to_1D(c, face, c3, Nx, Ny + 1) - 1] = sid
            m1234[from_3D_to_1D(c * 2, face * 2 - 1, c3, NX1, NY1) - 1] = sid
            m1256[from_3D_to_1D(c, face * 2 - 1, c3 * 2, Nx, NY1) - 1] = sid
        elif t == "34":
            sid = cnt["12"] + cnt["34"]
            m34[from_3D_to_1D(face, c2, c3, Nx + 1, Ny) - 1] = sid
            m1234[from_3D_to_1D(face * 2 - 1, c2 * 2, c3, NX1, NY1) - 1] = sid
            m3456[from_3D_to_1D(face * 2 - 1, c2, c3 * 2, NX1, Ny) - 1] = sid
        else:
            sid = cnt["12"] + cnt["34"] + cnt["56"]
            m56[from_3D_to_1D(c, c2, face, Nx, Ny) - 1] = sid
            m1256[from_3D_to_1D(c, c2 * 2, face * 2 - 1, Nx, NY1) - 1] = sid
            m3456[from_3D_to_1D(c * 2, c2, face * 2 - 1, NX1, Ny) - 1] = sid
        gr.append(bin_search(_node_of(G, mapping, nodes, c, c2, c3), nodes_red))
        gc.append(sid)

    # per type: (normal axis, outer loops in order, lower/upper externals)
    def blocks(t, ax, f_lo_ext, f_hi_ext):
        Nax = (Nx, Ny, Nz)[ax]

        def loops(fixed=None, lo=1, hi=None):
            rng = {0: range(1, Nx + 1), 1: range(1, Ny + 1), 2: range(1, Nz + 1)}
            r = dict(rng)
            r[ax] = [fixed] if fixed is not None else range(lo, hi + 1)
            if t == "34" and fixed == 1:
                # the first x-normal block runs y, then z (cont fixed)
                for c2 in r[1]:
                    for c3 in r[2]:
                        yield 1, c2, c3
                return
            for c in r[0]:
                for c2 in r[1]:
                    for c3 in r[2]:
                        yield c, c2, c3

        def nb(p, d):
            q = list(p)
            q[ax] += d
            return q

        # 1: lower faces on the box
        for p in loops(fixed=1):
            for k in range(1, G.K + 1):
                idx = p[ax]
                check_om = other_at(k, *nb(p, 1), idx + 1 <= Nax)
                if G(k, *p) and (not _ext(ext, k, f_hi_ext, *p) or check_om):
                    record(t, k, *p, face=idx)
                    break
        # 2: lower faces inside
        for p in loops(lo=2, hi=Nax):
            for k in range(1, G.K + 1):
                idx = p[ax]
                check_om = other_at(k, *nb(p, 1), idx + 1 <= Nax)
                if G(k, *p) and not G(k, *nb(p, -1)) and (not _ext(ext, k, f_hi_ext, *p) or check_om):
                    record(t, k, *p, face=idx)
                    break
        # 3: upper faces on the box
        for p in loops(fixed=Nax):
            for k in range(1, G.K + 1):
                idx = p[ax]
                check_om = other_at(k, *nb(p, -1), idx - 1 >= 1)
                if G(k, *p) and (not _ext(ext, k, f_lo_ext, *p) or check_om):
                    record(t, k, *p, face=idx + 1)
                    break
        # 4: upper faces inside, unless another material sits beyond
        for p in loops(lo=1, hi=Nax - 1):
            for k in range(1, G.K + 1):
                idx = p[ax]
                check_om = other_at(k, *nb(p, -1), idx - 1 >= 1)
                if G(k, *p) and not G(k, *nb(p, 1)) and (not _ext(ext, k, f_lo_ext, *p) or check_om):
                    if not any(G(k2, *nb(p, 1)) for k2 in range(1, G.K + 1) if k2 != k):
                        record(t, k, *p, face=idx + 1)
                    break

    blocks("12", 1, 1, 2)
    blocks("34", 0, 3, 4)
    blocks("56", 2, 5, 6)
    n12, n34 = cnt["12"], cnt["34"]
    ns = n12 + n34 + cnt["56"]
    Gamma = sp.csc_matrix((np.ones(len(gr)), (np.asarray(gr, dtype=np.int64) - 1, np.asarray(gc, dtype=np.int64) - 1)),
                          shape=(len(nodes_red), ns))

    def sel(mp, keep):
        inn = np.nonzero(keep)[0]
        return sp.csc_matrix((np.ones(inn.size), (inn, mp[inn] - 1)), shape=(mp.size, ns))

    exp = [[None] * 3 for _ in range(3)]
    exp[0][0] = sel(m12, m12 != 0)
    exp[1][1] = sel(m34, m34 != 0)
    exp[2][2] = sel(m56, m56 != 0)
    exp[1][0] = sel(m1234, (m1234 != 0) & (m1234 <= n12))
    exp[0][1] = sel(m1234, m1234 > n12)
    exp[2][0] = sel(m1256, (m1256 != 0) & (m1256 <= n12))
    exp[0][2] = sel(m1256, m1256 > n12 + n34)
    exp[2][1] = sel(m3456, (m3456 != 0) & (m3456 <= n12 + n34))
    exp[1][2] = sel(m3456, m3456 > n12 + n34)
    N = (int(np.count_nonzero(m12)), int(np.count_nonzero(m34)), int(np.count_nonzero(m56)))
    return exp, Gamma, N


def nodes_find_rev(point, centri, skip):
    """find_voxels_le_pp.jl:14-37: nearest voxel centre (stable order), the
    next one when the nearest is `skip` (1-based voxel ids)."""
    d = np.sqrt(((centri - np.asarray(point, dtype=np.float64)) ** 2).sum(axis=1))
    o = np.argsort(d, kind="stable") + 1
    return int(o[0]) if o[0] != skip else int(o[1])


def mesh(grids, materials, sx, sy, sz, origin, dominant_list, ports=(), lumped_elements=(), p_self=None):
    """mesher_FFT.jl:1-112 on the given grids (materials[k-1] belongs to
    grid k; dominant_list: 1-based conductor grids).  ports: list of (start,
    end) points; lumped_elements: list of dicts with start, end, type, R,
    L, C.  p_self: the three self P values (mesher_FFT.jl:114-138),
    unscaled.  Returns a dict of the reference's objects, 0-based."""
    G = Grids(grids)
    esc = {"Lp": 1e6, "P": 1e-12, "R": 1e-3, "Cd": 1e12, "Is": 1e3, "Yle": 1e3, "freq": 1e-9, "time": 1e9}
    mats = [dict(m, epsr=m["eps_re"] - 1j * m["eps_re"] * m["tan_D"]) for m in materials]
    mapping, num = create_volumes_mapping(G)
    centri, id_mat = create_volume_centers(G, mapping, num, sx, sy, sz, origin)
    ext = create_grids_externals(G)
    # mesher_FFT.jl:24-31 sizes the node array by the full voxels of every
    # grid (a voxel in two grids counts twice)
    num_full_vox = int(sum(np.count_nonzero(g) for g in G.g))
    nodes, nodes_red, nodes_reused = create_nodes_ref(G, num_full_vox, ext, mapping, dominant_list)
    exp, Gamma, (N1, N2, N3) = create_mapping_Gamma_no_rep(G, mapping, nodes, nodes_red, ext)
    amaps, nums = [], []
    for axis in range(3):
        a, nA = create_mapping_A(axis, G, mapping, nodes, nodes_red)
        amaps.append(a)
        nums.append(nA)
    A, li_mat, li_border, zs = create_A_mats(G, mapping, amaps, nums, mats, nodes, nodes_red)
    Zs, surface_edges = compute_Zs(mats, li_mat, li_border, zs, sx, sy, sz)
    mat_map_Lp = []
    sizes = ((G.Nx - 1) * G.Ny * G.Nz, G.Nx * (G.Ny - 1) * G.Nz, G.Nx * G.Ny * (G.Nz - 1))
    for axis in range(3):
        ind = create_expansion_ind_Lp(axis, G, amaps[axis], nums[axis], mapping, nodes, nodes_red)
        mat_map_Lp.append(sp.csc_matrix((np.ones(ind.shape[0]), (ind[:, 0] - 1, ind[:, 1] - 1)),
                                        shape=(sizes[axis], nums[axis])))
    diag = OD.compute_diagonals(esc, mats, sx, sy, sz, *li_mat, *li_border)
    out = dict(escalings=esc, nodes=nodes, nodes_red=nodes_red, nodes_reused=nodes_reused, A=A, Gamma=Gamma,
               exp_P=exp, N1=N1, N2=N2, N3=N3, li_mat=li_mat, li_border=li_border, maps_Zs=zs, Zs=Zs,
               surface_edges=surface_edges, mat_map_Lp=mat_map_Lp, mx=nums[0], my=nums[1], mz=nums[2],
               diag_R=diag["R"], diag_Cd=diag["Cd"], diag_Lp=diag["Lp"], centri=centri, id_mat=id_mat)
    if p_self is not None:
        out["diag_P"] = esc["P"] * np.concatenate([np.full(N1, p_self[0]), np.full(N2, p_self[1]), np.full(N3, p_self[2])])
    pn = []
    for a, b in ports:
        v1 = nodes_find_rev(a, centri, -1)
        v2 = nodes_find_rev(b, centri, v1)
        pn.append((bin_search(nodes[v1 - 1], nodes_red) - 1, bin_search(nodes[v2 - 1], nodes_red) - 1))
    out["port_nodes"] = np.asarray(pn, dtype=np.int64).reshape(-1, 2)
    le = []
    for e in lumped_elements:
        v1 = nodes_find_rev(e["start"], centri, -1)
        v2 = nodes_find_rev(e["end"], centri, v1)
        le.append(dict(e, nodes=(bin_search(nodes[v1 - 1], nodes_red) - 1, bin_search(nodes[v2 - 1], nodes_red) - 1)))
    out["lumped_elements"] = le
    return out
