"""Nested-dissection elimination order for the node system SS.

SS = Yle + A' Z^-1 A + j w Gamma P^-1 Gamma' (FFT_solver_QS_S_type.jl:118)
couples a node only to the nodes its edges reach (A' Z^-1 A; the Gamma
term is diagonal because P_d is), plus the port and lumped-element node
pairs through Yle, which may lie anywhere.  Each dissection step cuts the
nodes by a coordinate plane and takes as separator the nodes on one side
that an edge joins to the other side (the smaller such set), so the two
halves are uncoupled whatever the mesher's node placement: one node per
voxel centre, or the reference mesher's lattice points moved onto faces,
edges and corners (two nodes an edge joins may then lie four lattice
steps apart).  Without the adjacency the separator is the plane itself,
exact for one node per voxel centre.  The reference
lets UMFPACK choose a fill-reducing order (lu(SS), :119); any order gives
the same solution, and this one is chosen for the GPU's triangular solves:
the factor rows come in contiguous blocks -- leaves and separators of the
dissection -- whose dependency levels are wide near the bottom of the tree,
so each level's blocks are solved in parallel.

The order is a function of the geometry only and is computed once per mesh.
"""
import numpy as np


def _touching(nodes, adj, side, other):
    """The nodes with a neighbour (in adj) whose side is `other`."""
    if nodes.size == 0:
        return nodes
    ptr = adj.indptr
    starts = ptr[nodes]
    counts = ptr[nodes + 1] - starts
    total = int(counts.sum())
    if total == 0:
        return nodes[:0]
    offs = np.repeat(starts - np.concatenate(([0], np.cumsum(counts)[:-1])), counts) + np.arange(total)
    hit = side[adj.indices[offs]] == other
    return np.unique(np.repeat(nodes, counts)[hit])


def _split(idx, ijk, leaf, fronts, adj=None, side=None):
    """Append the fronts of a dissection of idx (postorder: children before
    parents) and return the index of its top front."""
    if idx.size <= leaf:
        fronts.append([idx, []])
        return len(fronts) - 1
    c = ijk[idx]
    lo, hi = c.min(0), c.max(0)
    ax = int(np.argmax(hi - lo))
    if hi[ax] == lo[ax]:  # coincident nodes: no plane separates them
        fronts.append([idx, []])
        return len(fronts) - 1
    vals = c[:, ax]
    m = int(np.floor(np.median(vals)))
    m = min(max(m, int(lo[ax]) + 1), int(hi[ax]))
    if adj is None:
        left, sep, right = idx[vals < m], idx[vals == m], idx[vals > m]
    else:
        left, right = idx[vals < m], idx[vals >= m]
        side[left] = 0
        side[right] = 1
        sl = _touching(left, adj, side, 1)
        sr = _touching(right, adj, side, 0)
        side[idx] = -1
        if sl.size <= sr.size:
            sep, left = sl, np.setdiff1d(left, sl, assume_unique=True)
        else:
            sep, right = sr, np.setdiff1d(right, sr, assume_unique=True)
    kids = []
    if left.size:
        kids.append(_split(left, ijk, leaf, fronts, adj, side))
    if right.size:
        kids.append(_split(right, ijk, leaf, fronts, adj, side))
    fronts.append([sep, kids])
    return len(fronts) - 1


class Dissection:
    """A nested-dissection elimination order with its assembly tree.

    order[k]      node eliminated k-th
    front_ptr     fronts are order[front_ptr[f] : front_ptr[f+1]]: leaves
                  of the dissection, separators and (last) the special
                  nodes; every front follows its descendants
    front_parent  parent front (-1 for the root)
    block_ptr     the order cut into blocks of at most `cap` rows, each
                  inside one front (rows of the triangular solves)
    """

    def __init__(self, order, front_ptr, front_parent, block_ptr):
        self.order, self.front_ptr, self.front_parent, self.block_ptr = order, front_ptr, front_parent, block_ptr


def dissect(ijk, special=(), leaf=64, cap=64, adj=None):
    """Nested dissection of the nodes with integer lattice coordinates ijk
    (n x 3).  Nodes in `special` (ports, lumped elements: coupled through
    Yle to arbitrary nodes) form the root front.  adj: the symmetric node
    adjacency (CSR, n x n) the separators are taken from; None: coordinate
    planes."""
    ijk = np.asarray(ijk, dtype=np.int64)
    n = ijk.shape[0]
    special = np.unique(np.asarray(special, dtype=np.int64).reshape(-1))
    rest = np.setdiff1d(np.arange(n, dtype=np.int64), special)
    fronts = []
    side = None
    if adj is not None:
        adj = adj.tocsr()
        side = np.full(n, -1, dtype=np.int8)
    tops = [_split(rest, ijk, leaf, fronts, adj, side)] if rest.size else []
    if special.size or not tops:
        fronts.append([special, tops])
    # drop empty fronts (a separator can be empty when a side is empty)
    parent = [-1] * len(fronts)
    for f, (_, kids) in enumerate(fronts):
        for k in kids:
            parent[k] = f
    keep = [f for f in range(len(fronts)) if fronts[f][0].size]
    remap = {}
    for f in keep:
        remap[f] = len(remap)

    def up(f):
        while f >= 0 and f not in remap:
            f = parent[f]
        return f

    front_parent = np.array([remap[up(parent[f])] if up(parent[f]) >= 0 else -1 for f in keep], dtype=np.int64)
    sizes = [fronts[f][0].size for f in keep]
    order = np.concatenate([fronts[f][0] for f in keep]) if keep else np.zeros(0, dtype=np.int64)
    front_ptr = np.zeros(len(keep) + 1, dtype=np.int64)
    front_ptr[1:] = np.cumsum(sizes)
    bp = [0]
    for f in range(len(keep)):
        a, b = int(front_ptr[f]), int(front_ptr[f + 1])
        bp.extend(range(a + cap, b, cap))
        bp.append(b)
    assert order.size == n and np.array_equal(np.sort(order), np.arange(n))
    return Dissection(order, front_ptr, front_parent, np.array(bp, dtype=np.int64))


def nested_dissection(ijk, special=(), leaf=64, cap=64):
    """(order, block_ptr) of dissect(): the elimination order and its cut
    into blocks of at most `cap` rows."""
    d = dissect(ijk, special, leaf, cap)
    return d.order, d.block_ptr


def mesh_ijk(mesh):
    """Integer coordinates of the nodes: their lattice points where the
    mesher records them, else the voxel centres (one node per voxel)."""
    if getattr(mesh, "node_ijk", None) is not None:
        return np.asarray(mesh.node_ijk, dtype=np.int64)
    return np.rint((np.asarray(mesh.vox_centers) - np.asarray(mesh.origin)) /
                   np.array([mesh.sx, mesh.sy, mesh.sz]) - 0.5).astype(np.int64)


def mesh_special(mesh):
    special = [np.asarray(mesh.port_nodes, dtype=np.int64).reshape(-1)]
    special += [np.asarray(le["nodes"], dtype=np.int64).reshape(-1) for le in mesh.lumped_elements]
    return np.concatenate(special)


def mesh_adjacency(mesh):
    """Node adjacency through the edges: the pattern of A' A off the diagonal."""
    import scipy.sparse as sp

    A = mesh.A.tocsr()
    rows = np.repeat(np.arange(A.shape[0]), np.diff(A.indptr))
    E = sp.csr_matrix((np.ones(A.indices.size), (rows, A.indices)), shape=A.shape)
    G = (E.T @ E).tocsr()
    G.setdiag(0)
    G.eliminate_zeros()
    return G


def mesh_dissection(mesh, leaf=64, cap=64):
    """dissect() for a voxel mesh; port and lumped-element nodes are
    special.  Meshes whose nodes are the reference mesher's lattice points
    (node_ijk) take their separators from the edge adjacency; one node per
    voxel centre takes coordinate planes."""
    adj = mesh_adjacency(mesh) if getattr(mesh, "node_ijk", None) is not None else None
    return dissect(mesh_ijk(mesh), mesh_special(mesh), leaf, cap, adj)


def mesh_order(mesh, leaf=64, cap=64):
    """nested_dissection for a voxel mesh (see mesh_dissection)."""
    d = mesh_dissection(mesh, leaf, cap)
    return d.order, d.block_ptr
