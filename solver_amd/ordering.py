"""Nested-dissection elimination order for the node system SS.

SS = Yle + A' Z^-1 A + j w Gamma P^-1 Gamma' (FFT_solver_QS_S_type.jl:118)
couples a node only to the nodes of face-adjacent voxels (A' Z^-1 A; the
Gamma term is diagonal because P_d is), plus the port and lumped-element
node pairs through Yle, which may lie anywhere.  Cutting the voxel lattice
by coordinate planes therefore gives exact separators.  The reference
lets UMFPACK choose a fill-reducing order (lu(SS), :119); any order gives
the same solution, and this one is chosen for the GPU's triangular solves:
the factor rows come in contiguous blocks -- leaves and separators of the
dissection -- whose dependency levels are wide near the bottom of the tree,
so each level's blocks are solved in parallel.

The order is a function of the geometry only and is computed once per mesh.
"""
import numpy as np


def _split(idx, ijk, leaf, cap, out):
    if idx.size <= leaf:
        out.append(idx)
        return
    c = ijk[idx]
    lo, hi = c.min(0), c.max(0)
    ax = int(np.argmax(hi - lo))
    if hi[ax] == lo[ax]:  # coincident nodes: no plane separates them
        for s in range(0, idx.size, leaf):
            out.append(idx[s:s + leaf])
        return
    vals = c[:, ax]
    m = int(np.floor(np.median(vals)))
    m = min(max(m, int(lo[ax]) + 1), int(hi[ax]))
    left, sep, right = idx[vals < m], idx[vals == m], idx[vals > m]
    _split(left, ijk, leaf, cap, out)
    if right.size:
        _split(right, ijk, leaf, cap, out)
    for s in range(0, sep.size, cap):
        out.append(sep[s:s + cap])


def nested_dissection(ijk, special=(), leaf=64, cap=64):
    """Order the nodes with integer lattice coordinates ijk (n x 3).

    Nodes in `special` (ports, lumped elements: coupled through Yle to
    arbitrary nodes) are placed last.  Returns (order, block_ptr): order[k]
    is the node eliminated k-th; block_ptr (nblocks + 1, from 0 to n) cuts
    the order into leaves of at most `leaf` nodes and separator chunks of
    at most `cap` nodes."""
    ijk = np.asarray(ijk, dtype=np.int64)
    n = ijk.shape[0]
    special = np.unique(np.asarray(special, dtype=np.int64).reshape(-1))
    rest = np.setdiff1d(np.arange(n, dtype=np.int64), special)
    parts = []
    if rest.size:
        _split(rest, ijk, leaf, cap, parts)
    for s in range(0, special.size, cap):
        parts.append(special[s:s + cap])
    parts = [p for p in parts if p.size]
    order = np.concatenate(parts) if parts else np.zeros(0, dtype=np.int64)
    block_ptr = np.zeros(len(parts) + 1, dtype=np.int64)
    block_ptr[1:] = np.cumsum([p.size for p in parts])
    assert order.size == n and np.array_equal(np.sort(order), np.arange(n))
    return order, block_ptr


def mesh_order(mesh, leaf=64, cap=64):
    """nested_dissection for a voxel mesh: nodes are voxel centres; port
    and lumped-element nodes are special."""
    ijk = np.rint((np.asarray(mesh.vox_centers) - np.asarray(mesh.origin)) /
                  np.array([mesh.sx, mesh.sy, mesh.sz]) - 0.5).astype(np.int64)
    special = [np.asarray(mesh.port_nodes, dtype=np.int64).reshape(-1)]
    special += [np.asarray(le["nodes"], dtype=np.int64).reshape(-1) for le in mesh.lumped_elements]
    return nested_dissection(ijk, np.concatenate(special), leaf, cap)
