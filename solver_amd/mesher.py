"""Voxel topology front end: the inputs of the FFT hot path.

This is a compact, vectorised restatement of the reference's mesher
(mesher_FFT.jl:1-112 and the create_* / find_voxels_* files it calls) that
produces the same objects in the same layouts:

  incidence_selection["A"]      m x n  edge-node incidence, -1 at the lower
                                       voxel, +1 at the upper (create_A_mats_
                                       and_find_borders_with_map_Zs.jl:24-31)
  incidence_selection["Gamma"]  n x ns surface-node selection
  expansions["mat_map_Lp"][d]   N_ind[d] x m_d edge -> Lp grid position
  expansions["exp_P"][i][j]     P grid position <-> surface (create_mapping_
                                       Gamma_no_rep.jl:380-421 layouts)
  diagonals R, Cd (m x 4), Lp (m), P (ns)       (compute_diagonals.jl,
                                                 mesher_FFT.jl:114-138)
  Zs_info["Zs"], ["surface_edges"]             (compute_Zs_with_indices.jl)
  ports["port_nodes"], lumped_elements["le_nodes"]

Numbering follows the reference: x-edges in (y, z, x) loop order, y-edges
in (z, x, y), z-edges in (x, y, z) (create_mapping_A{x,y,z}_v2.jl); surfaces
are the y-normal ("12"), then x-normal ("34"), then z-normal ("56") faces,
each in the four-block order of create_mapping_Gamma_no_rep.jl.

Simplifications against the reference, stated once here:
  * every full voxel owns one node at its centre; the reference moves
    conductor nodes onto faces shared with another material
    (create_nodes_ref.jl:84-407).  Adjacent voxels of different materials
    are joined by an edge whose two halves carry each material's R and Cd
    (the four-slot layout of compute_diagonals.jl:170-236 already allows it);
  * a surface is a voxel face that borders empty space or the bounding box;
  * the skin-effect count of a half edge is the number of its voxel's
    exposed faces parallel to the edge.
The hot path takes these objects as given, so both the GPU path and the
oracle consume identical inputs.
"""
import numpy as np
import scipy.sparse as sp

ESCALINGS = {"Lp": 1e6, "P": 1e-12, "R": 1e-3, "Cd": 1e12, "Is": 1e3, "Yle": 1e3, "freq": 1e-9, "time": 1e9}
EPS0 = 8.854187816997944e-12


def lin(i, j, k, M, N):
    """0-based linear index, x fastest (From_3D_to_1D.jl:1-3 minus one)."""
    return i + M * (j + N * k)


class Material:
    def __init__(self, name, sigma=0.0, eps_re=1.0, tan_d=0.0):
        self.name = name
        self.sigmar = float(sigma)
        self.eps_re = float(eps_re)
        self.tan_D = float(tan_d)
        # mesher_FFT.jl:21
        self.epsr = self.eps_re - 1j * self.eps_re * self.tan_D


class Mesh:
    """Output of mesh(): every object the FFT solver hot path consumes."""


def conductor_terms(M, sigma_vox):
    """(diag_R scaled m x 4, Zs m) of mesh M with voxel conductivities
    sigma_vox in place of M.sigma_vox (compute_diagonals.jl:10-16,
    compute_Zs_with_indices.jl); every other slot keeps its value."""
    c = M.cond
    sig = np.asarray(sigma_vox, dtype=np.float64)
    R = M.diag_R / M.escalings["R"]
    on = c["slot_vox"] >= 0
    R = R.copy()
    R[on] = c["slot_lh"][on] / (sig[c["slot_vox"][on]] * c["slot_wt"][on])
    Zs = np.zeros(M.m, dtype=np.complex128)
    for h in range(2):
        v = c["zs_vox"][:, h]
        ok = v >= 0
        Zs[ok] += c["zs_geo"][ok, h] * np.sqrt(2 * np.pi * 1e-7 / sig[v[ok]])
    return M.escalings["R"] * R, Zs


def _lp_self(l, W, T):
    # compute_diagonals.jl:274-293, host copy for the diagonal terms
    w = W / l
    t = T / l
    r = np.sqrt(w ** 2 + t ** 2)
    aw = np.sqrt(w ** 2 + 1)
    at = np.sqrt(t ** 2 + 1)
    ar = np.sqrt(w ** 2 + t ** 2 + 1)
    mu0 = 4 * np.pi * 1e-7
    return 2 * mu0 * l / np.pi * (1 / 4 * (1 / w * np.arcsinh(w / at) + 1 / t * np.arcsinh(t / aw) + np.arcsinh(1 / r)) +
        1 / 24 * (t ** 2 / w * np.arcsinh(w / (t * at * (r + ar))) + w ** 2 / t * np.arcsinh(t / (w * aw * (r + ar))) +
        t ** 2 / w ** 2 * np.arcsinh(w ** 2 / (t * r * (at + ar))) + w ** 2 / t ** 2 * np.arcsinh(t ** 2 / (w * r * (aw + ar))) +
        1 / (w * t ** 2) * np.arcsinh(w * t ** 2 / (at * (aw + ar))) + 1 / (t * w ** 2) * np.arcsinh(t * w ** 2 / (aw * (at + ar)))) -
        1 / 6 * (1 / (w * t) * np.arctan(w * t / ar) + t / w * np.arctan(w / (t * ar)) + w / t * np.arctan(t / (w * ar))) -
        1 / 60 * ((ar + r + t + at) * t ** 2 / ((ar + r) * (r + t) * (t + at) * (at + ar)) +
        (ar + r + w + aw) * w ** 2 / ((ar + r) * (r + w) * (w + aw) * (aw + ar)) +
        (ar + aw + 1 + at) / ((ar + aw) * (aw + 1) * (1 + at) * (at + ar))) -
        1 / 20 * (1 / (r + ar) + 1 / (aw + ar) + 1 / (at + ar)))


def _sel(rows, cols, nrow, ncol):
    return sp.csc_matrix((np.ones(len(rows)), (np.asarray(rows, dtype=np.int64), np.asarray(cols, dtype=np.int64))),
                         shape=(nrow, ncol))


def mesh(mat_grid, materials, cell, origin=(0.0, 0.0, 0.0), ports=(), lumped_elements=(), p_self_fn=None,
         sigma_grid=None):
    """mat_grid: int array (Nx, Ny, Nz), 0 = empty, k >= 1 = materials[k-1].
    ports: list of (start_xyz, end_xyz) in metres; lumped_elements: list of
    dicts with start, end, type, R, L, C.  p_self_fn(sx, sy, sz) -> (self P of
    y-, x-, z-faces) unscaled (compute_diagonal_P_v2, mesher_FFT.jl:114-138).
    sigma_grid (Nx, Ny, Nz), optional: the conductivity of each conductor
    voxel in place of its material's (the reference assigns sigma per
    material, compute_diagonals.jl:5-33; per-voxel values are what the
    conductivity gradients are taken against)."""
    g = np.asarray(mat_grid, dtype=np.int64)
    Nx, Ny, Nz = g.shape
    sx, sy, sz = (float(c) for c in cell)
    full = g > 0
    M = Mesh()
    M.Nx, M.Ny, M.Nz = Nx, Ny, Nz
    M.sx, M.sy, M.sz = sx, sy, sz
    M.origin = tuple(float(o) for o in origin)
    M.materials = list(materials)
    M.escalings = dict(ESCALINGS)

    # voxels / nodes: numbered x fastest (create_volumes_mapping.jl:65-77)
    vox_id = -np.ones((Nx, Ny, Nz), dtype=np.int64)
    order = np.nonzero(full.transpose(2, 1, 0).reshape(-1))[0]  # z, y, x loops with x fastest
    flat = -np.ones(Nx * Ny * Nz, dtype=np.int64)
    flat[order] = np.arange(order.size)
    vox_id = flat.reshape(Nz, Ny, Nx).transpose(2, 1, 0)
    n = order.size
    M.n = n
    xs, ys, zs = np.meshgrid(np.arange(Nx), np.arange(Ny), np.arange(Nz), indexing="ij")
    M.vox_centers = np.zeros((n, 3))
    sel = full
    M.vox_centers[vox_id[sel], 0] = origin[0] + sx * xs[sel] + sx / 2
    M.vox_centers[vox_id[sel], 1] = origin[1] + sy * ys[sel] + sy / 2
    M.vox_centers[vox_id[sel], 2] = origin[2] + sz * zs[sel] + sz / 2
    M.vox_material = np.zeros(n, dtype=np.int64)
    M.vox_material[vox_id[sel]] = g[sel]
    # conductivity per voxel (0 for non-conductors)
    mat_sigma = np.array([0.0] + [mt.sigmar for mt in materials])
    M.sigma_vox = mat_sigma[M.vox_material]
    if sigma_grid is not None:
        sg = np.asarray(sigma_grid, dtype=np.float64)[sel]
        cond = M.sigma_vox[vox_id[sel]] != 0
        M.sigma_vox[vox_id[sel][cond]] = sg[cond]
        if np.any(M.sigma_vox[vox_id[sel][cond]] <= 0):
            raise ValueError("sigma_grid must be positive on conductor voxels")

    # edges
    def edges(axis):
        a = [slice(None)] * 3
        b = [slice(None)] * 3
        a[axis] = slice(0, -1)
        b[axis] = slice(1, None)
        both = full[tuple(a)] & full[tuple(b)]
        dims = both.shape
        ii, jj, kk = np.nonzero(both)
        # reference loop orders: x (y,z,x); y (z,x,y); z (x,y,z) -- last index fastest
        if axis == 0:
            key = np.lexsort((ii, kk, jj))
        elif axis == 1:
            key = np.lexsort((jj, ii, kk))
        else:
            key = np.lexsort((kk, jj, ii))
        ii, jj, kk = ii[key], jj[key], kk[key]
        v1 = vox_id[ii, jj, kk]
        off = [0, 0, 0]
        off[axis] = 1
        v2 = vox_id[ii + off[0], jj + off[1], kk + off[2]]
        pos = lin(ii, jj, kk, dims[0], dims[1])
        return dict(v1=v1, v2=v2, i=ii, j=jj, k=kk, pos=pos, grid=np.prod(dims), dims=dims)

    E = [edges(0), edges(1), edges(2)]
    M.mx, M.my, M.mz = (e["v1"].size for e in E)
    m = M.mx + M.my + M.mz
    M.m = m
    v1 = np.concatenate([e["v1"] for e in E])
    v2 = np.concatenate([e["v2"] for e in E])
    rows = np.concatenate([np.arange(m), np.arange(m)])
    M.A = sp.csc_matrix((np.concatenate([-np.ones(m), np.ones(m)]), (rows, np.concatenate([v1, v2]))), shape=(m, n))
    M.mat_map_Lp = [_sel(E[d]["pos"], np.arange(E[d]["v1"].size), E[d]["grid"], E[d]["v1"].size) for d in range(3)]
    M.N_ind = [E[d]["grid"] for d in range(3)]

    # edge materials and borders
    li_mat = np.stack([M.vox_material[v1], M.vox_material[v2]], axis=1)
    border = np.zeros((m, 2), dtype=np.int64)
    start = 0
    for d in range(3):
        e = E[d]
        cnt = e["v1"].size
        off = [0, 0, 0]
        off[d] = 1
        i, j, k = e["i"], e["j"], e["k"]
        lo = np.array([i - off[0], j - off[1], k - off[2]])
        hi = np.array([i + 2 * off[0], j + 2 * off[1], k + 2 * off[2]])

        def has(c):
            ok = (c[0] >= 0) & (c[0] < Nx) & (c[1] >= 0) & (c[1] < Ny) & (c[2] >= 0) & (c[2] < Nz)
            out = np.zeros(c.shape[1], dtype=bool)
            out[ok] = full[c[0][ok], c[1][ok], c[2][ok]]
            return out

        border[start:start + cnt, 0] = np.where(~has(lo), li_mat[start:start + cnt, 0], 0)
        border[start:start + cnt, 1] = np.where(~has(hi), li_mat[start:start + cnt, 1], 0)
        start += cnt
    M.li_mat = li_mat
    M.li_border = border

    # surfaces: faces bordering empty space / the box, four-block order per type
    surf_vox, surf_pos = [], {}
    Gp = {}
    maps = {k: [] for k in ("12", "34", "56", "1234", "1256", "3456")}
    Nx1, Ny1, Nz1 = 2 * (Nx + 1) - 1, 2 * (Ny + 1) - 1, 2 * (Nz + 1) - 1

    def exposed(i, j, k, di, dj, dk):
        ni, nj, nk = i + di, j + dj, k + dk
        out = (ni < 0) | (ni >= Nx) | (nj < 0) | (nj >= Ny) | (nk < 0) | (nk >= Nz)
        inside = ~out
        res = out.copy()
        res[inside] = ~full[ni[inside], nj[inside], nk[inside]]
        return res

    def add_type(axis):
        res = []
        # blocks: lower faces at index 0, lower faces at 1.., upper faces at N-1, upper at 0..N-2
        N_ax = (Nx, Ny, Nz)[axis]
        for upper in (False, True):
            for first_block in (True, False):
                ii, jj, kk = np.nonzero(full)
                c = (ii, jj, kk)[axis]
                if not upper:
                    m_ = (c == 0) if first_block else (c >= 1)
                else:
                    m_ = (c == N_ax - 1) if first_block else (c <= N_ax - 2)
                ii, jj, kk = ii[m_], jj[m_], kk[m_]
                d = [0, 0, 0]
                d[axis] = 1 if upper else -1
                ex = exposed(ii, jj, kk, *d)
                ii, jj, kk = ii[ex], jj[ex], kk[ex]
                # loop orders: y-faces (x, y, z), x-faces (x, y, z) [block 1: (y, z)],
                # z-faces (x, y, z); z fastest in every block
                key = np.lexsort((kk, jj, ii))
                res.append((ii[key], jj[key], kk[key], upper))
        return res

    blocks = {1: add_type(1), 0: add_type(0), 2: add_type(2)}
    s = 0
    rows = {k: [] for k in maps}
    cols = {k: [] for k in maps}
    for axis, tname in ((1, "12"), (0, "34"), (2, "56")):
        for ii, jj, kk, upper in blocks[axis]:
            cnt = ii.size
            sid = np.arange(s, s + cnt)
            s += cnt
            surf_vox.append(vox_id[ii, jj, kk])
            c = [ii + 1, jj + 1, kk + 1]  # 1-based cont, cont2, cont3
            c[axis] = c[axis] + (1 if upper else 0)
            cx, cy, cz = c
            if tname == "12":
                rows["12"].append(lin(cx - 1, cy - 1, cz - 1, Nx, Ny + 1))
                rows["1234"].append(lin(2 * cx - 1, 2 * cy - 2, cz - 1, Nx1, Ny1))
                rows["1256"].append(lin(cx - 1, 2 * cy - 2, 2 * cz - 1, Nx, Ny1))
                for k_ in ("12", "1234", "1256"):
                    cols[k_].append(sid)
            elif tname == "34":
                rows["34"].append(lin(cx - 1, cy - 1, cz - 1, Nx + 1, Ny))
                rows["1234"].append(lin(2 * cx - 2, 2 * cy - 1, cz - 1, Nx1, Ny1))
                rows["3456"].append(lin(2 * cx - 2, cy - 1, 2 * cz - 1, Nx1, Ny))
                for k_ in ("34", "1234", "3456"):
                    cols[k_].append(sid)
            else:
                rows["56"].append(lin(cx - 1, cy - 1, cz - 1, Nx, Ny))
                rows["1256"].append(lin(cx - 1, 2 * cy - 1, 2 * cz - 2, Nx, Ny1))
                rows["3456"].append(lin(2 * cx - 1, cy - 1, 2 * cz - 2, Nx1, Ny))
                for k_ in ("56", "1256", "3456"):
                    cols[k_].append(sid)
        if tname == "12":
            M.N1 = s
        elif tname == "34":
            M.N2 = s - M.N1
    ns = s
    M.N3 = ns - M.N1 - M.N2
    M.ns = ns
    surf_vox = np.concatenate(surf_vox) if surf_vox else np.zeros(0, dtype=np.int64)
    M.Gamma = _sel(surf_vox, np.arange(ns), n, ns)
    gsize = {"12": Nx * (Ny + 1) * Nz, "34": (Nx + 1) * Ny * Nz, "56": Nx * Ny * (Nz + 1),
             "1234": Nx1 * Ny1 * Nz, "1256": Nx * Ny1 * Nz1, "3456": Nx1 * Ny * Nz1}

    def build(key, lo, hi):
        r = np.concatenate(rows[key]) if rows[key] else np.zeros(0, dtype=np.int64)
        c = np.concatenate(cols[key]) if cols[key] else np.zeros(0, dtype=np.int64)
        keep = (c >= lo) & (c < hi)
        return _sel(r[keep], c[keep], gsize[key], ns)

    N1, N2 = M.N1, M.N2
    exp = [[None] * 3 for _ in range(3)]
    exp[0][0] = build("12", 0, N1)
    exp[1][1] = build("34", N1, N1 + N2)
    exp[2][2] = build("56", N1 + N2, ns)
    exp[1][0] = build("1234", 0, N1)          # exp_P[2,1]: 12 surfaces on p1234
    exp[0][1] = build("1234", N1, N1 + N2)    # exp_P[1,2]: 34 surfaces on p1234
    exp[2][0] = build("1256", 0, N1)          # exp_P[3,1]
    exp[0][2] = build("1256", N1 + N2, ns)    # exp_P[1,3]
    exp[2][1] = build("3456", N1, N1 + N2)    # exp_P[3,2]
    exp[1][2] = build("3456", N1 + N2, ns)    # exp_P[2,3]
    M.exp_P = exp

    # diagonals (compute_diagonals.jl:1-134)
    esc = M.escalings
    R = np.zeros((m, 4))
    Cd = np.zeros((m, 4), dtype=np.complex128)
    lens = [(sx, sy, sz), (sy, sx, sz), (sz, sy, sx)]
    # conductor slots: voxel, half length and cross-section (R = lh / (sigma wt))
    slot_vox = -np.ones((m, 4), dtype=np.int64)
    slot_lh = np.zeros((m, 4))
    slot_wt = np.zeros((m, 4))
    start = 0
    dl = []
    for d, (E_d, cnt) in enumerate(zip(E, (M.mx, M.my, M.mz))):
        sl = slice(start, start + cnt)
        l, W, T = lens[d]
        for slot, mats in ((0, li_mat[sl, 0]), (1, li_mat[sl, 1]), (2, border[sl, 0]), (3, border[sl, 1])):
            vox = (v1 if slot in (0, 2) else v2)[sl]
            for kmat in np.unique(mats):
                if kmat == 0:
                    continue
                mat = materials[kmat - 1]
                loc = np.nonzero(mats == kmat)[0]
                idx = loc + start
                if mat.sigmar != 0:
                    R[idx, slot] = 0.5 * l / (M.sigma_vox[vox[loc]] * W * T)
                    slot_vox[idx, slot] = vox[loc]
                    slot_lh[idx, slot] = 0.5 * l
                    slot_wt[idx, slot] = W * T
                    if mat.epsr != 1:
                        Cd[idx, slot] = EPS0 * (mat.epsr - 1) * W * T / (0.5 * l)
                else:
                    Cd[idx, slot] = EPS0 * (mat.epsr - 1) * W * T / (0.5 * l)
        aux = (border[sl, 0] > 0).astype(int) + (border[sl, 1] > 0).astype(int)
        s0, s1, s2 = _lp_self(l, W, T), _lp_self(l + l / 2, W, T), _lp_self(2 * l, W, T)
        dl.append(np.where(aux == 2, s2, np.where(aux == 1, s1, s0)))
        start += cnt
    M.diag_R = esc["R"] * R
    M.diag_Cd = esc["Cd"] * Cd
    M.diag_Lp = esc["Lp"] * np.concatenate(dl) if dl else np.zeros(0)
    if p_self_fn is not None:
        pself = p_self_fn(sx, sy, sz)
        M.diag_P = esc["P"] * np.concatenate([np.full(M.N1, pself[0]), np.full(M.N2, pself[1]), np.full(M.N3, pself[2])])
    else:
        M.diag_P = None

    # skin-effect surface impedance per edge (compute_Zs_with_indices.jl):
    # Zs = sum over the two halves of zs_geo * sqrt(2 pi 1e-7 / sigma)
    Zs = np.zeros(m, dtype=np.complex128)
    zs_vox = -np.ones((m, 2), dtype=np.int64)
    zs_geo = np.zeros((m, 2), dtype=np.complex128)
    start = 0
    for d, e in enumerate(E):
        cnt = e["v1"].size
        l = lens[d][0]
        others = [a for a in range(3) if a != d]
        for half, vk in ((0, e["v1"]), (1, e["v2"])):
            mat_ids = M.vox_material[vk]
            ii = e["i"] + (e["i"] * 0 if half == 0 else (1 if d == 0 else 0))
            jj = e["j"] + (0 if half == 0 else (1 if d == 1 else 0))
            kk = e["k"] + (0 if half == 0 else (1 if d == 2 else 0))
            for nrm in others:
                width = (sx, sy, sz)[3 - d - nrm]
                cntf = np.zeros(cnt)
                for sgn in (-1, 1):
                    dd = [0, 0, 0]
                    dd[nrm] = sgn
                    cntf += exposed(ii, jj, kk, *dd)
                for kmat in np.unique(mat_ids):
                    mat = materials[kmat - 1]
                    if mat.sigmar == 0:
                        continue
                    sel_ = mat_ids == kmat
                    rows_ = start + np.nonzero(sel_)[0]
                    geo = cntf[sel_] * 0.5 * l / width * (1 + 1j)
                    Zs[rows_] += geo * np.sqrt(2 * np.pi * 1e-7 / M.sigma_vox[vk[sel_]])
                    zs_geo[rows_, half] += geo
                    zs_vox[rows_, half] = vk[sel_]
        start += cnt
    M.Zs = Zs
    M.cond = dict(slot_vox=slot_vox, slot_lh=slot_lh, slot_wt=slot_wt, zs_vox=zs_vox, zs_geo=zs_geo)
    M.surface_edges = np.nonzero(np.abs(Zs) != 0)[0]

    # ports / lumped elements: nearest voxel centre, the end skipping the
    # start voxel (find_voxels_port_pp.jl, nodes_find_rev :21-29)
    def nearest(pt, skip):
        dist = np.sqrt(((M.vox_centers - np.asarray(pt, dtype=float)) ** 2).sum(axis=1))
        o = np.argsort(dist, kind="stable")
        return o[0] if o[0] != skip else o[1]

    pn = []
    for a, b in ports:
        v_a = nearest(a, -1)
        v_b = nearest(b, v_a)
        pn.append((v_a, v_b))
    M.port_nodes = np.array(pn, dtype=np.int64).reshape(-1, 2)
    le = []
    for L_ in lumped_elements:
        v_a = nearest(L_["start"], -1)
        v_b = nearest(L_["end"], v_a)
        le.append(dict(L_, nodes=(v_a, v_b)))
    M.lumped_elements = le
    return M
