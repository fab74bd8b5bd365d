"""Pin the oracle's circulant matvec, preconditioner, GMRES and sweep.

  * the FFT matvec against a dense brute-force assembly of the same
    Toeplitz blocks from the same kernel rows (no FFT involved);
  * the block-Schur preconditioner against a dense solve with the block
    matrix it inverts, [[diag DZ, 0, A], [0, diag P, -Gamma'],
    [-A', j w Gamma, Yle]];
  * GMRES against a dense direct solve of M^-1 K x = b;
  * the whole sweep against the DC resistance of the crossing bars computed
    from the resistor network alone.
"""
import numpy as np
import pytest
import scipy.sparse as sp

from oracle import fft_solver as OF
from oracle import integrals as OI
from solver_amd import configs, mesher


def pself(sx, sy, sz):
    c = np.zeros((1, 3))
    return [OI.p_row(c[0], c, sx, sy, sz, 1, 1)[0], OI.p_row(c[0], c, sx, sy, sz, 3, 3)[0],
            OI.p_row(c[0], c, sx, sy, sz, 5, 5)[0]]


def small_mesh():
    """6 x 5 x 4 grid: a dielectric slab under a bent copper trace, with
    holes so that every surface type and border case appears."""
    g = np.zeros((6, 5, 4), dtype=np.int64)
    g[:, :, :2] = 2
    g[1:5, 2, 2] = 1
    g[4, 1:4, 2:4] = 1
    g[0, 0, 0] = 0
    mats = [mesher.Material("copper", 5.8e7), mesher.Material("diel", 0.0, 4.4, 0.02)]
    cell = (0.4e-3, 0.3e-3, 0.1e-3)
    ports = [((1.5 * 0.4e-3, 2.5 * 0.3e-3, 2.5e-4), (1.5 * 0.4e-3, 2.5 * 0.3e-3, 1.5e-4)),
             ((4.5 * 0.4e-3, 3.5 * 0.3e-3, 3.5e-4), (4.5 * 0.4e-3, 3.5 * 0.3e-3, 0.5e-4))]
    le = [dict(start=(0.5 * 0.4e-3, 0.5 * 0.3e-3, 1.5e-4), end=(5.5 * 0.4e-3, 4.5 * 0.3e-3, 1.5e-4),
               type=2, R=10.0, L=1e-9, C=1e-12)]
    return mesher.mesh(g, mats, cell, (1e-3, 2e-3, 0.0), ports, le, p_self_fn=pself)


@pytest.fixture(scope="module")
def setup():
    M = small_mesh()
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    return M, rows, spec


def _pos(mapmat):
    """column -> grid position of a 0/1 selection (or -1)."""
    c = mapmat.tocsc()
    out = -np.ones(c.shape[1], dtype=np.int64)
    for j in range(c.shape[1]):
        if c.indptr[j + 1] > c.indptr[j]:
            out[j] = c.indices[c.indptr[j]]
    return out


def _toeplitz(row, n, pos_out, pos_in):
    """Dense block: entry = row[|dx| + nx (|dy| + ny |dz|)] between the grid
    positions of an output and an input unknown (-1 = not on the grid)."""
    nx, ny, nz = n

    def ijk(p):
        return p % nx, (p // nx) % ny, p // (nx * ny)

    B = np.zeros((len(pos_out), len(pos_in)))
    vo = pos_out >= 0
    vi = pos_in >= 0
    io, jo, ko = ijk(pos_out[vo])
    ii, ji, ki = ijk(pos_in[vi])
    d = np.abs(io[:, None] - ii[None, :]) + nx * (np.abs(jo[:, None] - ji[None, :]) + ny * np.abs(ko[:, None] - ki[None, :]))
    B[np.ix_(vo, vi)] = row[d]
    return B


def dense_K(M, rows, w, DZ, Yle):
    m, ns, n = M.m, M.ns, M.n
    Lp = np.zeros((m, m))
    offs = [0, M.mx, M.mx + M.my]
    sizes = [M.mx, M.my, M.mz]
    for d in range(3):
        nd = OF.kernel_dims(d, M.Nx, M.Ny, M.Nz)
        p = _pos(M.mat_map_Lp[d])
        sl = slice(offs[d], offs[d] + sizes[d])
        Lp[sl, sl] = _toeplitz(rows[d], nd, p, p)
    P = np.zeros((ns, ns))
    kk = {(0, 0): 3, (1, 1): 4, (2, 2): 5, (0, 1): 6, (0, 2): 7, (1, 2): 8}
    for (c1, c2), k in kk.items():
        nd = OF.kernel_dims(k, M.Nx, M.Ny, M.Nz)
        P += _toeplitz(rows[k], nd, _pos(M.exp_P[c2][c1]), _pos(M.exp_P[c1][c2]))
        if c1 != c2:
            P += _toeplitz(rows[k], nd, _pos(M.exp_P[c1][c2]), _pos(M.exp_P[c2][c1]))
    A = M.A.toarray()
    G = M.Gamma.toarray()
    K = np.zeros((m + ns + n, m + ns + n), dtype=np.complex128)
    K[:m, :m] = 1j * w * Lp + np.diag(DZ)
    K[:m, m + ns:] = A
    K[m:m + ns, m:m + ns] = P
    K[m:m + ns, m + ns:] = -G.T
    K[m + ns:, :m] = -A.T
    K[m + ns:, m:m + ns] = 1j * w * G
    K[m + ns:, m + ns:] = Yle.toarray()
    return K


def dense_M(M, w, DZ, Yle):
    m, ns, n = M.m, M.ns, M.n
    A = M.A.toarray()
    G = M.Gamma.toarray()
    B = np.zeros((m + ns + n, m + ns + n), dtype=np.complex128)
    B[:m, :m] = np.diag(DZ)
    B[:m, m + ns:] = A
    B[m:m + ns, m:m + ns] = np.diag(M.diag_P)
    B[m:m + ns, m + ns:] = -G.T
    B[m + ns:, :m] = -A.T
    B[m + ns:, m:m + ns] = 1j * w * G
    B[m + ns:, m + ns:] = Yle.toarray()
    return B


def _freq_data(M, f):
    w = 2 * np.pi * f * M.escalings["freq"]
    Yle = OF.build_Yle(M.lumped_elements, M.port_nodes, M.escalings["Yle"], M.n, w / M.escalings["freq"], 50.0)
    DZ = OF.frequency_terms(M, w)
    return w, DZ, Yle


def test_fft_matvec_matches_dense_toeplitz(setup):
    M, rows, spec = setup
    w, DZ, Yle = _freq_data(M, 3e8)
    op = OF.Operator(M, spec)
    op.set_frequency(w, DZ, Yle)
    rng = np.random.default_rng(0)
    N = M.m + M.ns + M.n
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    y = np.concatenate(op.K(x))
    yd = dense_K(M, rows, w, DZ, Yle) @ x
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) < 1e-13


def test_preconditioner_is_the_block_schur_inverse(setup):
    M, rows, spec = setup
    w, DZ, Yle = _freq_data(M, 1e9)
    op = OF.Operator(M, spec)
    op.set_frequency(w, DZ, Yle)
    rng = np.random.default_rng(1)
    N = M.m + M.ns + M.n
    X = rng.normal(size=N) + 1j * rng.normal(size=N)
    y = op.precond(X[:M.m], X[M.m:M.m + M.ns], X[M.m + M.ns:])
    yd = np.linalg.solve(dense_M(M, w, DZ, Yle), X)
    assert np.linalg.norm(y - yd) / np.linalg.norm(yd) < 1e-11


def test_gmres_matches_direct_solve(setup):
    M, rows, spec = setup
    w, DZ, Yle = _freq_data(M, 2e8)
    op = OF.Operator(M, spec)
    op.set_frequency(w, DZ, Yle)
    N = M.m + M.ns + M.n
    is_ = np.zeros(M.n)
    is_[M.port_nodes[0, 0]] = 1e3
    is_[M.port_nodes[0, 1]] = -1e3
    b = op.precond_rhs(is_)
    x, flag, relres, it, resvec = OF.gmres_custom(b, 1e-12, N, None, op.matvec)
    assert flag == 0 and relres <= 1e-12
    Kd = dense_K(M, rows, w, DZ, Yle)
    xd = np.linalg.solve(Kd, np.concatenate([np.zeros(M.m + M.ns), is_]))
    assert np.linalg.norm(x - xd) / np.linalg.norm(xd) < 1e-9
    # resvec is the reference's residual history: first entry ||b - A x0||
    assert abs(resvec[0] - np.linalg.norm(b)) <= 1e-12 * np.linalg.norm(b)


def test_gmres_stagnation_and_iteration_cap(setup):
    M, rows, spec = setup
    w, DZ, Yle = _freq_data(M, 2e8)
    op = OF.Operator(M, spec)
    op.set_frequency(w, DZ, Yle)
    is_ = np.zeros(M.n)
    is_[M.port_nodes[1, 0]] = 1e3
    is_[M.port_nodes[1, 1]] = -1e3
    b = op.precond_rhs(is_)
    # too few iterations: flag 1, minimum-residual iterate returned
    x, flag, relres, it, resvec = OF.gmres_custom(b, 1e-14, 3, None, op.matvec)
    assert flag == 1 and len(resvec) == 4 and relres > 1e-14
    # stop request -> flag 99
    out = OF.gmres_custom(b, 1e-8, 10, None, op.matvec, stop=lambda: True)
    assert out[1] == 99


def test_sweep_dc_limit_of_crossbar():
    """With the skin-effect switch off, Re Z11 of the crossing bars at a
    frequency where omega L << R equals the DC resistance of the resistor
    network between the port voxels (S -> Z de-embeds the 50 ohm port)."""
    cfg = configs.crossbar_24x24x4()
    M = mesher.mesh(cfg["grid"], cfg["materials"], cfg["cell"], cfg["origin"], cfg["ports"], p_self_fn=pself)
    M.Zs = np.zeros_like(M.Zs)
    M.surface_edges = np.zeros(0, dtype=np.int64)
    rows = OF.kernel_rows(M.origin, M.Nx, M.Ny, M.Nz, M.sx, M.sy, M.sz)
    spec = OF.spectra(rows, M.Nx, M.Ny, M.Nz)
    out = OF.fft_solver_qs([50.0, 1e9], M, spec, 300, 1e-12)
    Z = out["Z"][0, 0, :]
    # DC network: conductances 1/R_edge (R = sum of the half-cell slots, unscaled)
    R = M.diag_R.sum(axis=1) / M.escalings["R"]
    A = M.A.tocsr()
    Gl = (A.T @ sp.diags(1.0 / R) @ A).tolil()
    n1, n2 = M.port_nodes[0]
    Gl[n2, :] = 0
    Gl[n2, n2] = 1.0
    rhs = np.zeros(M.n)
    rhs[n1] = 1.0
    v = sp.linalg.spsolve(Gl.tocsc(), rhs)
    Rdc = v[n1] - v[n2]
    assert abs(Z[0].real - Rdc) / Rdc < 1e-3, (Z[0], Rdc)
    # inductive: Im Z grows with frequency
    assert Z[1].imag > 100 * abs(Z[0].imag)
