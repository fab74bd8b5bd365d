"""Nested-dissection order of the node system and the blocked triangular
solve it feeds, checked on the CPU.

The blocked solve is restated here in numpy (gather from earlier blocks,
then column-by-column substitution inside the block) and compared with
scipy's sparse triangular solve; the level schedule is checked to respect
every block dependency of the factor."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import integrals as OI
from solver_amd import configs, mesher, ordering, sweep


def pself(sx, sy, sz):
    c = np.zeros((1, 3))
    return [OI.p_row(c[0], c, sx, sy, sz, cs, cs)[0] for cs in (1, 3, 5)]


@pytest.fixture(scope="module")
def crossbar():
    cfg = configs.crossbar_24x24x4()
    M = mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"], p_self_fn=pself)
    w = 2 * np.pi * 1e8 * M.escalings["freq"]
    Yle = sweep.build_Yle_FFT(M.lumped_elements, M.port_nodes, M.escalings, M.n, w / M.escalings["freq"], 50.0)
    SS = sweep.SS_of(M, w, sweep.DZ_of(M, w), Yle)
    return M, SS


def test_order_is_permutation_with_valid_blocks(crossbar):
    M, SS = crossbar
    order, bp = ordering.mesh_order(M, leaf=16, cap=24)
    assert np.array_equal(np.sort(order), np.arange(M.n))
    assert bp[0] == 0 and bp[-1] == M.n and np.all(np.diff(bp) > 0)
    assert np.max(np.diff(bp)) <= 24
    # port nodes are eliminated last
    pn = np.unique(M.port_nodes)
    assert set(order[-pn.size:]) == set(pn)


def test_fill_stays_in_ancestor_blocks(crossbar):
    """Factoring SS[order, order] keeps the order (no pivoting, no column
    postorder), and every entry (r, c) of L has block(r) >= block(c): the
    fill of a leaf reaches only the separators above it."""
    M, SS = crossbar
    order, bp = ordering.mesh_order(M, leaf=16, cap=10 ** 9)
    blk = np.searchsorted(bp, np.arange(M.n), side="right") - 1
    F = spla.splu(SS[order][:, order].tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(SymmetricMode=True))
    assert np.array_equal(F.perm_c, np.arange(M.n)) and np.array_equal(F.perm_r, np.arange(M.n))
    Lc = F.L.tocoo()
    assert np.all(blk[Lc.row] >= blk[Lc.col])
    # two sibling leaves never touch: the leaves of the first split are
    # solved in parallel
    nb = len(bp) - 1
    dep = sp.coo_matrix((np.ones(Lc.nnz), (blk[Lc.row], blk[Lc.col])), shape=(nb, nb)).tocsr()
    dep.setdiag(0)
    dep.eliminate_zeros()
    first_level = [B for B in range(nb) if dep[B].nnz == 0]
    assert len(first_level) >= 4


def _blocked_solve(T, bp, lower, b):
    """The GPU kernel's algorithm, in numpy, block by block in level order."""
    n = T.shape[0]
    nb = len(bp) - 1
    blk = np.searchsorted(bp, np.arange(n), side="right") - 1
    T = T.tocsr()
    Tc = T.tocsc()
    level = np.zeros(nb, dtype=int)
    seq = range(nb) if lower else range(nb - 1, -1, -1)
    for B in seq:
        rows = np.arange(bp[B], bp[B + 1])
        lv = 0
        for r in rows:
            cols = T.indices[T.indptr[r]:T.indptr[r + 1]]
            other = blk[cols][blk[cols] != B]
            if other.size:
                lv = max(lv, level[other].max() + 1)
        level[B] = lv
    x = np.zeros(n, dtype=complex)
    done = np.zeros(n, dtype=bool)
    for lv in range(level.max() + 1):
        for B in np.nonzero(level == lv)[0]:
            r0, r1 = bp[B], bp[B + 1]
            acc = np.array(b[r0:r1], dtype=complex)
            for t in range(r0, r1):
                cols = T.indices[T.indptr[t]:T.indptr[t + 1]]
                vals = T.data[T.indptr[t]:T.indptr[t + 1]]
                sel = (blk[cols] != B) & (cols != t)
                assert done[cols[sel]].all(), "level schedule read an unsolved row"
                acc[t - r0] -= vals[sel] @ x[cols[sel]]
            js = range(r0, r1) if lower else range(r1 - 1, r0 - 1, -1)
            for j in js:
                y = acc[j - r0] / T[j, j]
                ks = slice(Tc.indptr[j], Tc.indptr[j + 1])
                for t, v in zip(Tc.indices[ks], Tc.data[ks]):
                    if t != j and blk[t] == B:
                        acc[t - r0] -= v * y
            x[r0:r1] = acc / T.diagonal()[r0:r1]
            done[r0:r1] = True
    return x


def test_blocked_triangular_solve_matches_scipy(crossbar):
    M, SS = crossbar
    order, bp = ordering.mesh_order(M, leaf=16, cap=24)
    Lf, Uf, p, q = sweep.lu_factor(SS, order)
    assert np.array_equal(p, order) and np.array_equal(q, order)
    rng = np.random.default_rng(3)
    b = rng.normal(size=M.n) + 1j * rng.normal(size=M.n)
    y = _blocked_solve(Lf, bp, True, b)
    y_ref = spla.spsolve_triangular(Lf.tocsr(), b, lower=True)
    assert np.linalg.norm(y - y_ref) <= 1e-12 * np.linalg.norm(y_ref)
    z = _blocked_solve(Uf, bp, False, y)
    z_ref = spla.spsolve_triangular(Uf.tocsr(), y, lower=False)
    assert np.linalg.norm(z - z_ref) <= 1e-12 * np.linalg.norm(z_ref)
    # and the two sweeps solve SS (backward error; SS is ill-conditioned)
    x = np.empty_like(z)
    x[q] = z
    rhs = np.empty_like(b)
    rhs[p] = b
    nS = spla.norm(SS, 1)
    assert np.linalg.norm(SS @ x - rhs, 1) <= 1e-13 * (nS * np.linalg.norm(x, 1) + np.linalg.norm(rhs, 1))
