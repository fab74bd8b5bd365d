"""The reference's voxel mesher for the FFT solver, vectorised.

mesher_FFT.jl:1-112 and the files it calls, over numpy arrays instead of
voxel loops, producing the objects the hot path consumes (the same Mesh
fields as mesher.mesh()):

  nodes         create_nodes_ref.jl: every voxel's node is a point of the
                lattice of voxel centres, face centres, edge midpoints and
                corners (3 Nx x 3 Ny x 3 Nz, the centre of voxel (i, j, k)
                at (2i, 2j, 2k), 1-based).  A conductor voxel on its
                material's boundary moves its node to the face, edge or
                corner point of its boundary faces (one per axis; where
                both faces of an axis are boundary faces, the one touching
                another material, else neither).  A voxel of another
                material on its boundary takes the first conductor node
                among its 26 neighbouring points (build_nodes order), else
                moves the same way.  Voxels sharing a point share the node.
  edges (A)     create_mapping_A{x,y,z}_v2.jl and create_A_mats_...jl:
                between face neighbours of one material whose nodes differ,
                numbered in the loop orders (y, z, x) / (z, x, y) / (x, y, z).
  surfaces      create_mapping_Gamma_no_rep.jl: the 12 (y-normal), 34 and 56
                faces, four blocks each, with the reference's conditions on
                boundary faces and neighbouring materials.
  Zs            create_A_mats_...jl face counts and compute_Zs_with_indices.jl
  diagonals     compute_diagonals.jl; self P: mesher_FFT.jl:114-138.
  ports, lumped elements: find_voxels_port_pp.jl / find_voxels_le_pp.jl.

The reference's index quirks are kept (they define the operator it builds):
the lower-face contact tests of create_nodes_ref.jl start at the third
layer, the skin-effect counts treat a box face as one exposed face per half
edge, the z edges test the lower z level for the upper half on their +x/+y
sides (the +y one recording into slot 3), and the surface impedance uses
only whether a count is positive.

Materials pair with grids by name (the reference pairs the request's
material list with the Dict iteration order of mesher_matrices,
do_solving_fft.jl:17-28, which follows the keys' hashes).
"""
import numpy as np
import scipy.sparse as sp

from .mesher import ESCALINGS, EPS0, Mesh, _lp_self

AX = ("x", "y", "z")


def _lin(i, j, k, M, N):
    """0-based linear index, x fastest."""
    return i + M * (j + N * k)


def _shift(a, axis, d, fill=False):
    """b[p] = a[p + d e_axis] (fill outside the box)."""
    b = np.full_like(a, fill)
    n = a.shape[axis]
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if d > 0:
        src[axis] = slice(d, n)
        dst[axis] = slice(0, n - d)
    else:
        src[axis] = slice(0, n + d)
        dst[axis] = slice(-d, n)
    b[tuple(dst)] = a[tuple(src)]
    return b


def _index_grids(shape):
    return np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), np.arange(shape[2]), indexing="ij")


def externals(grids):
    """create_Grids_externals.jl: ext[k][f], f = 0..5 for faces -y, +y, -x,
    +x, -z, +z: the voxel is in grid k and that face lies on the box or
    borders a voxel outside grid k."""
    out = []
    for g in grids:
        e = []
        for axis in (1, 0, 2):
            e.append(g & ~_shift(g, axis, -1))
            e.append(g & ~_shift(g, axis, 1))
        out.append(e)
    return out


def _face_flags(grids, ext, k):
    """is_f1 .. is_f6 (create_nodes_ref.jl:16-116) of every voxel of grid k."""
    Nx, Ny, Nz = grids[0].shape
    X, Y, Z = _index_grids(grids[0].shape)
    others = [k2 for k2 in range(len(grids)) if k2 != k]

    def any_other(f, axis, d, valid):
        acc = np.zeros(grids[0].shape, dtype=bool)
        for k2 in others:
            acc |= _shift(ext[k2][f], axis, d)
        return acc & valid

    f = ext[k]
    fc = [any_other(1, 1, -1, Y >= 2), any_other(0, 1, 1, Y <= Ny - 2),
          any_other(3, 0, -1, X >= 2), any_other(2, 0, 1, X <= Nx - 2),
          any_other(5, 2, -1, Z >= 2), any_other(4, 2, 1, Z <= Nz - 2)]
    is_f = [x.copy() for x in f]
    for a, b in ((0, 1), (2, 3), (4, 5)):
        both = f[a] & f[b]
        is_f[a] = np.where(both, fc[a] & ~fc[b], f[a])
        is_f[b] = np.where(both, ~fc[a] & fc[b], f[b])
    return is_f


def _lattice(grids, is_f):
    """Lattice node id (1-based, create_nodes_ref.jl) of each voxel's centre
    moved to its boundary faces."""
    Nx, Ny, Nz = grids[0].shape
    X, Y, Z = _index_grids(grids[0].shape)
    c1 = 2 * (X + 1) + is_f[3].astype(np.int64) - is_f[2].astype(np.int64)
    c2 = 2 * (Y + 1) + is_f[1].astype(np.int64) - is_f[0].astype(np.int64)
    c3 = 2 * (Z + 1) + is_f[5].astype(np.int64) - is_f[4].astype(np.int64)
    return c1 + 3 * Nx * (c2 - 1) + 9 * Nx * Ny * (c3 - 1)


def create_nodes(grids, ext, vox_id, dominant):
    """create_nodes_ref.jl:1-418: the lattice node id of every voxel (vox_id
    numbering), the sorted distinct ids, and the ids a non-conductor voxel
    took from a conductor."""
    Nx, Ny, Nz = grids[0].shape
    nvox = int(sum(np.count_nonzero(g) for g in grids))
    nodes = np.zeros(nvox, dtype=np.int64)
    dom_ids = []
    for k in [k for k in range(len(grids)) if dominant[k]]:
        g = grids[k]
        is_f = _face_flags(grids, ext, k)
        node = _lattice(grids, is_f)
        nodes[vox_id[g]] = node[g]
        moved = g & np.logical_or.reduce(is_f)
        dom_ids.append(node[moved])
    dom = np.unique(np.concatenate(dom_ids)) if dom_ids else np.zeros(0, dtype=np.int64)
    reused = []
    X, Y, Z = _index_grids(grids[0].shape)
    centre = 2 * (X + 1) + 3 * Nx * (2 * (Y + 1) - 1) + 9 * Nx * Ny * (2 * (Z + 1) - 1)
    for k in [k for k in range(len(grids)) if not dominant[k]]:
        g = grids[k]
        is_f = _face_flags(grids, ext, k)
        node = _lattice(grids, is_f)
        on_b = g & np.logical_or.reduce(is_f)
        # the first of the 26 neighbouring points that is a conductor node
        # (build_nodes order: y offset slowest, then x, then z)
        shared = np.zeros(g.shape, dtype=np.int64)
        todo = on_b.copy()
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    if dx == 0 and dy == 0 and dz == 0:
                        continue
                    cand = centre + dx + 3 * Nx * dy + 9 * Nx * Ny * dz
                    hit = todo & np.isin(cand, dom)
                    shared[hit] = cand[hit]
                    todo &= ~hit
        take = on_b & (shared != 0)
        node = np.where(take, shared, node)
        nodes[vox_id[g]] = node[g]
        reused.append(shared[take])
    nodes_red = np.unique(nodes)
    return nodes, nodes_red, (np.unique(np.concatenate(reused)) if reused else np.zeros(0, dtype=np.int64))


def _node_index(nodes, nodes_red, vox):
    """0-based position of each voxel's node in nodes_red."""
    return np.searchsorted(nodes_red, nodes[vox])


def mesh_request(grids, materials, cell, origin, dominant, ports=(), lumped_elements=(), p_self_fn=None):
    """The reference mesher on `grids` (list of boolean (Nx, Ny, Nz) arrays,
    grid k made of materials[k], a mesher.Material), cell (sx, sy, sz) in
    metres, origin as given (the reference does not scale it), dominant[k]:
    grid k is a conductor.  ports: (start, end) points; lumped_elements:
    dicts with start, end, type, R, L, C.  p_self_fn(sx, sy, sz): the three
    self P values, unscaled.  Returns a mesher.Mesh."""
    grids = [np.asarray(g, dtype=bool) for g in grids]
    K = len(grids)
    Nx, Ny, Nz = grids[0].shape
    sx, sy, sz = (float(c) for c in cell)
    full = np.logical_or.reduce(grids)
    M = Mesh()
    M.Nx, M.Ny, M.Nz = Nx, Ny, Nz
    M.sx, M.sy, M.sz = sx, sy, sz
    M.origin = tuple(float(o) for o in origin)
    M.materials = list(materials)
    M.escalings = dict(ESCALINGS)

    # voxels: numbered z, y, x with x fastest (create_volumes_mapping.jl);
    # a voxel belongs to the first grid holding it
    owner = np.argmax(np.stack(grids), axis=0)
    order = np.nonzero(full.transpose(2, 1, 0).reshape(-1))[0]
    flat = -np.ones(Nx * Ny * Nz, dtype=np.int64)
    flat[order] = np.arange(order.size)
    vox_id = flat.reshape(Nz, Ny, Nx).transpose(2, 1, 0)
    nv = order.size
    X, Y, Z = _index_grids(full.shape)
    M.vox_centers = np.zeros((nv, 3))
    M.vox_centers[vox_id[full], 0] = origin[0] + sx * X[full] + sx / 2
    M.vox_centers[vox_id[full], 1] = origin[1] + sy * Y[full] + sy / 2
    M.vox_centers[vox_id[full], 2] = origin[2] + sz * Z[full] + sz / 2
    M.vox_material = np.zeros(nv, dtype=np.int64)
    M.vox_material[vox_id[full]] = owner[full] + 1

    ext = externals(grids)
    nodes, nodes_red, _ = create_nodes(grids, ext, vox_id, dominant)
    M.nodes, M.nodes_red = nodes, nodes_red
    n = nodes_red.size
    M.n = n
    # node coordinates on the lattice (0-based), for the elimination order
    t = nodes_red - 1
    M.node_ijk = np.stack([t % (3 * Nx), (t // (3 * Nx)) % (3 * Ny), t // (9 * Nx * Ny)], axis=1)
    nidx = np.full(full.shape, -1, dtype=np.int64)
    nidx[full] = _node_index(nodes, nodes_red, vox_id[full])
    M.vox_node = np.full(nv, -1, dtype=np.int64)
    M.vox_node[vox_id[full]] = nidx[full]

    # edges: face neighbours in one grid (the first holding both) with
    # different nodes
    E = []
    for axis in range(3):
        both = [g & _shift(g, axis, 1) for g in grids]
        anyb = np.logical_or.reduce(both)
        kfirst = np.argmax(np.stack(both), axis=0)
        up = _shift(nidx, axis, 1, -1)
        ok = anyb & (nidx != up)
        dims = list(full.shape)
        dims[axis] -= 1
        ok = ok[:dims[0], :dims[1], :dims[2]]
        ii, jj, kk = np.nonzero(ok)
        # numbering: x (y, z, x), y (z, x, y), z (x, y, z), last fastest
        key = np.lexsort((ii, kk, jj)) if axis == 0 else np.lexsort((jj, ii, kk)) if axis == 1 else np.lexsort((kk, jj, ii))
        ii, jj, kk = ii[key], jj[key], kk[key]
        off = [0, 0, 0]
        off[axis] = 1
        E.append(dict(i=ii, j=jj, k=kk, mat=kfirst[ii, jj, kk], n1=nidx[ii, jj, kk],
                      n2=nidx[ii + off[0], jj + off[1], kk + off[2]], pos=_lin(ii, jj, kk, dims[0], dims[1]),
                      grid=int(np.prod(dims))))
    M.mx, M.my, M.mz = (e["i"].size for e in E)
    m = M.mx + M.my + M.mz
    M.m = m
    n1 = np.concatenate([e["n1"] for e in E])
    n2 = np.concatenate([e["n2"] for e in E])
    rows = np.concatenate([np.arange(m), np.arange(m)])
    M.A = sp.csc_matrix((np.concatenate([-np.ones(m), np.ones(m)]), (rows, np.concatenate([n1, n2]))), shape=(m, n))
    M.mat_map_Lp = [sp.csc_matrix((np.ones(e["i"].size), (e["pos"], np.arange(e["i"].size))), shape=(e["grid"], e["i"].size))
                    for e in E]
    M.N_ind = [e["grid"] for e in E]

    # materials and borders per edge (1-based material ids, 0 = none)
    li_mat, li_border, zs = [], [], []
    for axis, e in enumerate(E):
        k1 = e["mat"] + 1
        li_mat.append(np.stack([k1, k1], axis=1))
        g_of = np.stack(grids)
        lo = (e["i"], e["j"], e["k"])[axis]
        Nax = (Nx, Ny, Nz)[axis]

        def at(d):
            p = [e["i"].copy(), e["j"].copy(), e["k"].copy()]
            p[axis] = p[axis] + d
            inside = (p[axis] >= 0) & (p[axis] < Nax)
            val = np.zeros(p[0].size, dtype=bool)
            val[inside] = g_of[e["mat"][inside], p[0][inside], p[1][inside], p[2][inside]]
            return val

        b0 = np.where(lo > 0, np.where(~at(-1), k1, 0), k1)
        b1 = np.where(lo + 1 == Nax - 1, k1, np.where((lo + 2 <= Nax - 1) & ~at(2), k1, 0))
        li_border.append(np.stack([b0, b1], axis=1))
        zs.append(_skin_counts(axis, e, grids, materials))
    M.li_mat = np.concatenate(li_mat)
    M.li_border = np.concatenate(li_border)

    # surfaces
    _surfaces(M, grids, ext, nidx)

    # diagonals (compute_diagonals.jl:1-134)
    esc = M.escalings
    R = np.zeros((m, 4))
    Cd = np.zeros((m, 4), dtype=np.complex128)
    lens = [(sx, sy, sz), (sy, sx, sz), (sz, sy, sx)]
    dl = []
    start = 0
    for axis, e in enumerate(E):
        cnt = e["i"].size
        sl = slice(start, start + cnt)
        l, W, T = lens[axis]
        for cont, mat in enumerate(materials, start=1):
            if mat.sigmar != 0:
                Rv = 0.5 * l / (mat.sigmar * W * T)
                Cv = 0 if mat.epsr == 1 else EPS0 * (mat.epsr - 1) * W * T / (0.5 * l)
            else:
                Rv = 0
                Cv = EPS0 * (mat.epsr - 1) * W * T / (0.5 * l)
            for col, src in ((0, M.li_mat[sl, 0]), (1, M.li_mat[sl, 1]), (2, M.li_border[sl, 0]),
                             (3, M.li_border[sl, 1])):
                ind = start + np.nonzero(src == cont)[0]
                if Rv != 0:
                    R[ind, col] = Rv
                if Cv != 0:
                    Cd[ind, col] = Cv
        aux = np.ceil(M.li_border[sl, 0] / 100).astype(np.int64) + np.ceil(M.li_border[sl, 1] / 100).astype(np.int64)
        s0, s1, s2 = _lp_self(l, W, T), _lp_self(l + l / 2, W, T), _lp_self(2 * l, W, T)
        dl.append(np.where(aux == 2, s2, np.where(aux == 1, s1, s0)))
        start += cnt
    M.diag_R = esc["R"] * R
    M.diag_Cd = esc["Cd"] * Cd
    M.diag_Lp = esc["Lp"] * np.concatenate(dl) if dl else np.zeros(0)
    if p_self_fn is not None:
        ps = p_self_fn(sx, sy, sz)
        M.diag_P = esc["P"] * np.concatenate([np.full(M.N1, ps[0]), np.full(M.N2, ps[1]), np.full(M.N3, ps[2])])
    else:
        M.diag_P = None

    # surface impedance (compute_Zs_with_indices.jl)
    M.Zs = _surface_impedance(materials, li_mat, li_border, zs, sx, sy, sz)
    M.surface_edges = np.nonzero(np.abs(M.Zs) != 0)[0]
    M.sigma_vox = np.array([0.0] + [mt.sigmar for mt in materials])[M.vox_material]
    M.cond = None

    # ports and lumped elements: nearest voxel centres (stable order), the
    # end skipping the start voxel, then that voxel's node
    def nearest(pt, skip):
        d = np.sqrt(((M.vox_centers - np.asarray(pt, dtype=np.float64)) ** 2).sum(axis=1))
        o = np.argsort(d, kind="stable")
        return o[0] if o[0] != skip else o[1]

    pn = []
    for a, b in ports:
        v1 = nearest(a, -1)
        v2 = nearest(b, v1)
        pn.append((M.vox_node[v1], M.vox_node[v2]))
    M.port_nodes = np.array(pn, dtype=np.int64).reshape(-1, 2)
    le = []
    for L_ in lumped_elements:
        v1 = nearest(L_["start"], -1)
        v2 = nearest(L_["end"], v1)
        le.append(dict(L_, nodes=(int(M.vox_node[v1]), int(M.vox_node[v2]))))
    M.lumped_elements = le
    return M


def _skin_counts(axis, e, grids, materials):
    """Face counts of create_A_mats_and_find_borders_with_map_Zs.jl for the
    conductor edges of one axis: two planes (x: normals z then y; y: z then
    x; z: x then y), each (NA, 4): count of the lower half, count of the
    upper half, other material beside the lower half, beside the upper half
    (1-based, 0 none)."""
    g_of = np.stack(grids)
    K = len(grids)
    NA = e["i"].size
    shape = grids[0].shape
    cond = np.array([materials[k].sigmar != 0 for k in range(K)])[e["mat"]]
    planes = {0: (2, 1), 1: (2, 0), 2: (0, 1)}[axis]
    out = []
    for pi, nrm in enumerate(planes):
        z = np.zeros((NA, 4), dtype=np.int64)
        idx = (e["i"], e["j"], e["k"])[nrm]
        Nn = shape[nrm]
        box = cond & ((idx == 0) | (idx == Nn - 1))
        z[box, 0] += 1
        z[box, 1] += 1
        inner = cond & ~box

        def probe(half, d):
            p = [e["i"].copy(), e["j"].copy(), e["k"].copy()]
            p[axis] = p[axis] + half
            p[nrm] = p[nrm] + d
            ok = (p[nrm] >= 0) & (p[nrm] < Nn)
            own = np.zeros(NA, dtype=bool)
            own[ok] = g_of[e["mat"][ok], p[0][ok], p[1][ok], p[2][ok]]
            # first other grid holding the probed voxel (1-based), 0 none
            other = np.zeros(NA, dtype=np.int64)
            for k2 in range(K - 1, -1, -1):
                hit = np.zeros(NA, dtype=bool)
                hit[ok] = g_of[k2, p[0][ok], p[1][ok], p[2][ok]]
                hit &= e["mat"] != k2
                other[hit] = k2 + 1
            return own, other

        for d in (-1, 1):
            lo_hit, lo_other = probe(0, d)
            sel = inner & ~lo_hit
            z[sel, 0] += 1
            rec = sel & (lo_other > 0)
            z[rec, 2] = lo_other[rec]
            if axis == 2 and d == 1:
                # the reference's upper-half tests on the +x / +y sides of a
                # z edge look at the lower z level (and the +y one records
                # into slot 3)
                up_hit, up_other = lo_hit, lo_other
                slot = 3 if nrm == 0 else 2
            else:
                up_hit, up_other = probe(1, d)
                slot = 3
            sel = inner & ~up_hit
            z[sel, 1] += 1
            rec = sel & (up_other > 0)
            z[rec, slot] = up_other[rec]
        out.append(z)
    return out


def _surface_impedance(materials, li_mat, li_border, zs, sx, sy, sz):
    """compute_Zs_with_indices.jl: per edge, the two half edges' face
    impedances in parallel per half, summed; a face count only counts when
    positive, and doubles on a border half voxel."""
    ratios = ((0.5 * sx / sy, 0.5 * sx / sz), (0.5 * sy / sx, 0.5 * sy / sz), (0.5 * sz / sx, 0.5 * sz / sy))
    out = []
    for axis in range(3):
        lm, lb = li_mat[axis], li_border[axis]
        za, zb = zs[axis]
        Zsw = np.zeros((lm.shape[0], 4), dtype=np.complex128)
        for cont, mat in enumerate(materials, start=1):
            if mat.sigmar == 0:
                continue
            base = (1 + 1j) * np.sqrt(2 * np.pi * 1e-7 / mat.sigmar)
            ca, cb = ratios[axis][0] * base, ratios[axis][1] * base
            on = lm[:, 0] == cont
            s1, s3 = on & (za[:, 0] > 0), on & (za[:, 1] > 0)
            s2, s4 = on & (zb[:, 0] > 0), on & (zb[:, 1] > 0)
            Zsw[s1, 0] = ca
            Zsw[s3, 2] = ca
            Zsw[s2, 1] = cb
            Zsw[s4, 3] = cb
            b1, b2 = lb[:, 0] == cont, lb[:, 1] == cont
            Zsw[s1 & b1, 0] *= 2
            Zsw[s2 & b1, 1] *= 2
            Zsw[s3 & b2, 2] *= 2
            Zsw[s4 & b2, 3] *= 2

        def par(a, b):
            s = a + b
            both = (np.abs(a) != 0) & (np.abs(b) != 0)
            s[both] = a[both] * b[both] / (a[both] + b[both])
            return s

        out.append(par(Zsw[:, 0], Zsw[:, 1]) + par(Zsw[:, 2], Zsw[:, 3]))
    return np.concatenate(out)


def _surfaces(M, grids, ext, nidx):
    """create_mapping_Gamma_no_rep.jl: surfaces per type (12: y-normal, 34:
    x-normal, 56: z-normal), each in four blocks -- lower faces on the box,
    lower faces inside, upper faces on the box, upper faces inside -- with
    the reference's conditions, Gamma (n x ns) and the P expansion maps."""
    Nx, Ny, Nz = grids[0].shape
    K = len(grids)
    NX1, NY1, NZ1 = 2 * (Nx + 1) - 1, 2 * (Ny + 1) - 1, 2 * (Nz + 1) - 1
    G = np.stack(grids)
    X, Y, Z = _index_grids(grids[0].shape)
    gsize = {"12": Nx * (Ny + 1) * Nz, "34": (Nx + 1) * Ny * Nz, "56": Nx * Ny * (Nz + 1),
             "1234": NX1 * NY1 * Nz, "1256": Nx * NY1 * NZ1, "3456": NX1 * Ny * NZ1}
    maps = {k: np.zeros(v, dtype=np.int64) for k, v in gsize.items()}
    g_rows, g_cols = [], []
    count = {"12": 0, "34": 0, "56": 0}
    # (type, normal axis, external face id of the lower / upper face)
    for t, ax, f_lo, f_hi in (("12", 1, 0, 1), ("34", 0, 2, 3), ("56", 2, 4, 5)):
        Nax = (Nx, Ny, Nz)[ax]
        idx = (X, Y, Z)[ax]
        for block in range(4):
            lower = block < 2
            d = 1 if lower else -1
            # another material at the neighbour on the far side (check_om)
            other_far = [np.zeros(G.shape[1:], dtype=bool) for _ in range(K)]
            beyond = [np.zeros(G.shape[1:], dtype=bool) for _ in range(K)]
            for k in range(K):
                for k2 in range(K):
                    if k2 != k:
                        other_far[k] |= _shift(grids[k2], ax, d)
                        beyond[k] |= _shift(grids[k2], ax, 1)
            if block == 0:
                region = idx == 0
            elif block == 1:
                region = idx >= 1
            elif block == 2:
                region = idx == Nax - 1
            else:
                region = idx <= Nax - 2
            chosen = np.full(G.shape[1:], -1, dtype=np.int64)
            for k in range(K - 1, -1, -1):
                c = region & grids[k] & (~ext[k][f_hi if lower else f_lo] | other_far[k])
                if block == 1:
                    c &= ~_shift(grids[k], ax, -1)
                elif block == 3:
                    c &= ~_shift(grids[k], ax, 1)
                chosen[c] = k
            rec = chosen >= 0
            if block == 3:
                kk = np.where(rec, chosen, 0)
                blocked = np.take_along_axis(np.stack(beyond), kk[None], axis=0)[0]
                rec &= ~blocked
            ii, jj, ll = np.nonzero(rec)
            # loop order: the first x-normal block runs (y, z); all others
            # (x, y, z), z fastest
            key = np.lexsort((ll, jj)) if (t == "34" and block == 0) else np.lexsort((ll, jj, ii))
            ii, jj, ll = ii[key], jj[key], ll[key]
            cnt = ii.size
            base = count["12"] + count["34"] + count["56"]
            sid = base + 1 + np.arange(cnt)
            count[t] += cnt
            face = (ii, jj, ll)[ax] + (0 if lower else 1)  # 0-based face index along the normal
            if t == "12":
                maps["12"][_lin(ii, face, ll, Nx, Ny + 1)] = sid
                maps["1234"][_lin(2 * ii + 1, 2 * face, ll, NX1, NY1)] = sid
                maps["1256"][_lin(ii, 2 * face, 2 * ll + 1, Nx, NY1)] = sid
            elif t == "34":
                maps["34"][_lin(face, jj, ll, Nx + 1, Ny)] = sid
                maps["1234"][_lin(2 * face, 2 * jj + 1, ll, NX1, NY1)] = sid
                maps["3456"][_lin(2 * face, jj, 2 * ll + 1, NX1, Ny)] = sid
            else:
                maps["56"][_lin(ii, jj, face, Nx, Ny)] = sid
                maps["1256"][_lin(ii, 2 * jj + 1, 2 * face, Nx, NY1)] = sid
                maps["3456"][_lin(2 * ii + 1, jj, 2 * face, NX1, Ny)] = sid
            g_rows.append(nidx[ii, jj, ll])
            g_cols.append(sid - 1)
    n12, n34 = count["12"], count["34"]
    ns = n12 + n34 + count["56"]
    M.ns = ns
    r = np.concatenate(g_rows) if g_rows else np.zeros(0, dtype=np.int64)
    c = np.concatenate(g_cols) if g_cols else np.zeros(0, dtype=np.int64)
    M.Gamma = sp.csc_matrix((np.ones(r.size), (r, c)), shape=(M.n, ns))

    def sel(key, keep):
        mp = maps[key]
        inn = np.nonzero(keep)[0]
        return sp.csc_matrix((np.ones(inn.size), (inn, mp[inn] - 1)), shape=(mp.size, ns))

    exp = [[None] * 3 for _ in range(3)]
    exp[0][0] = sel("12", maps["12"] != 0)
    exp[1][1] = sel("34", maps["34"] != 0)
    exp[2][2] = sel("56", maps["56"] != 0)
    exp[1][0] = sel("1234", (maps["1234"] != 0) & (maps["1234"] <= n12))
    exp[0][1] = sel("1234", maps["1234"] > n12)
    exp[2][0] = sel("1256", (maps["1256"] != 0) & (maps["1256"] <= n12))
    exp[0][2] = sel("1256", maps["1256"] > n12 + n34)
    exp[2][1] = sel("3456", (maps["3456"] != 0) & (maps["3456"] <= n12 + n34))
    exp[1][2] = sel("3456", maps["3456"] > n12 + n34)
    M.exp_P = exp
    M.N1 = int(np.count_nonzero(maps["12"]))
    M.N2 = int(np.count_nonzero(maps["34"]))
    M.N3 = int(np.count_nonzero(maps["56"]))


def renumber_edges(M):
    """The same mesh with each direction's edges in x-fastest order of their
    lower node's lattice position (z slowest).  The reference mesher lists
    the y edges y-fastest and the z edges z-fastest (mesher_FFT.jl), so the
    nodes of consecutive y / z edges lie a grid row / plane apart; in
    x-fastest order they are neighbours, like the nodes themselves, and the
    gathers between edge and node values (A, A') read contiguous runs.

    The operator is the same up to the order of its edge unknowns: the rows
    of A and of every per-edge array, and the columns of each direction's
    Lp position map, move together; nodes, surfaces and ports keep their
    numbers, so port impedances are unchanged.  Returns the mesh and the
    permutation (new edge k is old edge perm[k])."""
    A = M.A.tocsr()
    ijk = np.asarray(M.node_ijk)
    bounds = (0, M.mx, M.mx + M.my, M.m)
    perm = []
    for d in range(3):
        lo, hi = bounds[d], bounds[d + 1]
        rows = np.arange(lo, hi)
        lower = np.minimum.reduceat(A.indices[A.indptr[lo]:A.indptr[hi]], A.indptr[lo:hi] - A.indptr[lo])
        p = ijk[lower]
        perm.append(lo + np.lexsort((p[:, 0], p[:, 1], p[:, 2])))
        assert np.all(np.diff(A.indptr[lo:hi + 1]) > 0)
    perm = np.concatenate(perm)
    M.A = A[perm].tocsc()
    for k in ("Zs", "diag_Lp", "diag_R", "diag_Cd", "li_border", "li_mat"):
        v = getattr(M, k, None)
        if v is not None:
            setattr(M, k, np.asarray(v)[perm])
    M.mat_map_Lp = [sp.csc_matrix(M.mat_map_Lp[d])[:, perm[bounds[d]:bounds[d + 1]] - bounds[d]]
                    for d in range(3)]
    M.surface_edges = np.nonzero(np.abs(M.Zs) != 0)[0]
    return M, perm
