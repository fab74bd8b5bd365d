"""Host-side mirror of the reference's FFT-solver functions over libpeec_hip.

Each function keeps the name, argument order and meaning of the Julia
function it replaces (Draco112358/Solver, src/lib/solverFFT/) and forwards
to the C ABI declared in include/peec_hip.h.  All arithmetic of the hot path
runs on the GPU; the host only moves arguments and results.
"""
import ctypes

import numpy as np

from . import _lib as L


class Context:
    """One device context (one HIP stream), the analogue of one simulation
    task of the reference (solver_start.jl:187)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        L.check(L.lib().peec_ctx_create(int(device), ctypes.byref(self._h)), "peec_ctx_create")
        self._stop = ctypes.c_int32(0)
        L.check(L.lib().peec_set_stop_flag(self._h, ctypes.byref(self._stop)), "peec_set_stop_flag")

    @property
    def handle(self):
        return self._h

    def request_stop(self, on=True):
        """stopComputation[id][] = true (solver_start.jl:345-367)."""
        self._stop.value = 1 if on else 0

    def close(self):
        if self._h:
            L.lib().peec_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context():
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


def _centers_colmajor(c):
    c = np.asarray(c, dtype=np.float64).reshape(-1, 3)
    return np.ascontiguousarray(c.T).reshape(-1), c.shape[0]


def compute_Lp_Voxels(centri_m, centri_n, sx, sy, sz, sx2, sy2, sz2, dc, is_sym=False, id=None, ctx=None):
    """compute_Lp_Voxels.jl:1-63 with M = 1 (the hot-path call).

    Returns a 1 x N matrix, or None when a stop was requested (the
    reference's `return nothing`, :58)."""
    if is_sym:
        raise ValueError("is_sym=true is not on the FFT hot path (compute_FFT_mutual_coupling_mats.jl:46)")
    ctx = ctx or default_context()
    src = np.ascontiguousarray(np.asarray(centri_m, dtype=np.float64).reshape(-1)[:3])
    cf, N = _centers_colmajor(centri_n)
    out = np.empty(N, dtype=np.float64)
    rc = L.check(L.lib().peec_lp_row(ctx.handle, L.as_dp(src), L.as_dp(cf), N, sx, sy, sz, sx2, sy2, sz2,
                                     int(dc), L.as_dp(out)), "peec_lp_row")
    if rc == L.PEEC_STOPPED:
        return None
    return out.reshape(1, N)


def compute_row_P_sup(centers, centers2, sx, sy, sz, caser, caser2, id=None, ctx=None):
    """compute_row_P_sup.jl:1-104.  Returns an N x 1 matrix, or NaN when a
    stop was requested (the reference's `return NaN`, :59)."""
    ctx = ctx or default_context()
    src = np.ascontiguousarray(np.asarray(centers, dtype=np.float64).reshape(-1)[:3])
    cf, N = _centers_colmajor(centers2)
    out = np.empty(N, dtype=np.float64)
    rc = L.check(L.lib().peec_p_row(ctx.handle, L.as_dp(src), L.as_dp(cf), N, sx, sy, sz, int(caser), int(caser2),
                                    L.as_dp(out)), "peec_p_row")
    if rc == L.PEEC_STOPPED:
        return float("nan")
    return out.reshape(N, 1)


def Lp_self(l, W, T):
    """compute_diagonals.jl:136-155."""
    out = ctypes.c_double()
    L.check(L.lib().peec_lp_self(l, W, T, ctypes.byref(out)), "peec_lp_self")
    return out.value


# ---------------------------------------------------------------- operator
def _csc_f64(A):
    """Julia-layout CSC (1-based Int64) view of a scipy matrix; the arrays
    are returned too so they stay alive during the call."""
    A = A.tocsc()
    A.sort_indices()
    colptr = np.ascontiguousarray(A.indptr.astype(np.int64) + 1)
    rowval = np.ascontiguousarray(A.indices.astype(np.int64) + 1)
    nz = np.ascontiguousarray(A.data.astype(np.float64))
    s = L.CscF64(colptr.ctypes.data_as(L.i64p), rowval.ctypes.data_as(L.i64p), nz.ctypes.data_as(L.dp),
                 A.shape[0], A.shape[1])
    return s, (colptr, rowval, nz)


def _csc_c128(A):
    A = A.tocsc()
    A.sort_indices()
    colptr = np.ascontiguousarray(A.indptr.astype(np.int64) + 1)
    rowval = np.ascontiguousarray(A.indices.astype(np.int64) + 1)
    nz = np.ascontiguousarray(A.data.astype(np.complex128))
    s = L.CscC128(colptr.ctypes.data_as(L.i64p), rowval.ctypes.data_as(L.i64p), nz.ctypes.data_as(L.c128p),
                  A.shape[0], A.shape[1])
    return s, (colptr, rowval, nz)


def compute_FFT_mutual_coupling_mats(origin, Nx, Ny, Nz, sx, sy, sz, escalings, ctx=None):
    """compute_FFT_mutual_coupling_mats.jl:1-33 (QS): the 9 kernel spectra,
    kept on the device inside ctx.  Returns False when a stop was requested
    (the reference returns nothing)."""
    ctx = ctx or default_context()
    o = np.ascontiguousarray(origin, dtype=np.float64)
    rc = L.check(L.lib().peec_kernels_build(ctx.handle, L.as_dp(o), int(Nx), int(Ny), int(Nz), sx, sy, sz,
                                            float(escalings["Lp"]), float(escalings["P"])), "peec_kernels_build")
    return rc != L.PEEC_STOPPED


def compute_rows_Rcc(origin, Nx, Ny, Nz, sx, sy, sz, escalings, ctx=None):
    """compute_FFT_mutual_coupling_mats.jl:18-27 (QS_Rcc_FW = 2):
    compute_rows_Rcc_P + compute_rows_Rcc_Lp -- the static rows and centre
    distances of the retarded kernels, kept on the device inside ctx (with
    the quasi-static spectra).  From then on the convolutions of ctx use the
    retarded spectra of the frequency last given to
    compute_Circulant_Rcc.  Returns False when a stop was requested."""
    ctx = ctx or default_context()
    o = np.ascontiguousarray(origin, dtype=np.float64)
    rc = L.check(L.lib().peec_kernels_build_rcc(ctx.handle, L.as_dp(o), int(Nx), int(Ny), int(Nz), sx, sy, sz,
                                                float(escalings["Lp"]), float(escalings["P"])),
                 "peec_kernels_build_rcc")
    return rc != L.PEEC_STOPPED


def compute_Circulant_Rcc(freq_hz, ctx=None):
    """compute_Circulant_Lp_Rcc + compute_Circulant_P_sup_Rcc
    (FFT_solver_QS_S_type.jl:61-64) at one frequency in Hz: the complex
    spectra of escaling * row .* exp(-j beta Rcc), on the device."""
    ctx = ctx or default_context()
    L.check(L.lib().peec_rcc_spectra(ctx.handle, float(freq_hz)), "peec_rcc_spectra")


def kernel_row(kind, ctx=None, variant=0):
    """Kernel row `kind` (0..8) read back from the device: variant 0 the
    scaled quasi-static row, 1 the retarded path's static row, 2 its centre
    distances."""
    kind = int(kind) + 16 * int(variant)
    ctx = ctx or default_context()
    n = ctypes.c_int64()
    L.check(L.lib().peec_kernel_row(ctx.handle, int(kind), None, ctypes.byref(n)), "peec_kernel_row")
    out = np.empty(n.value)
    if n.value:
        L.check(L.lib().peec_kernel_row(ctx.handle, int(kind), L.as_dp(out), ctypes.byref(n)), "peec_kernel_row")
    return out


def operator_set(mesh, ctx=None):
    """Upload A, Gamma, the expansion maps and 1 ./ diag(P) of one geometry."""
    ctx = ctx or default_context()
    keep = []
    A, k = _csc_f64(mesh.A)
    keep.append(k)
    G, k = _csc_f64(mesh.Gamma)
    keep.append(k)
    lp = (L.CscF64 * 3)()
    for d in range(3):
        s, k = _csc_f64(mesh.mat_map_Lp[d])
        lp[d] = s
        keep.append(k)
    pm = (L.CscF64 * 9)()
    for i in range(3):
        for j in range(3):
            s, k = _csc_f64(mesh.exp_P[i][j])
            pm[3 * j + i] = s
            keep.append(k)
    dP = np.ascontiguousarray(mesh.diag_P, dtype=np.float64)
    L.check(L.lib().peec_operator_set(ctx.handle, int(mesh.mx), int(mesh.my), int(mesh.mz), ctypes.byref(A),
                                      ctypes.byref(G), lp, pm, L.as_dp(dP)), "peec_operator_set")
    set_ctx_size(ctx, int(mesh.m + mesh.ns + mesh.n))


def frequency_set(w, DZ, Yle, ctx=None):
    ctx = ctx or default_context()
    dz = np.ascontiguousarray(DZ, dtype=np.complex128)
    Y, keep = _csc_c128(Yle)
    L.check(L.lib().peec_frequency_set(ctx.handle, float(w), L.as_c128p(dz), ctypes.byref(Y)), "peec_frequency_set")


def frequency_slots(nslots, ctx=None):
    """Allocate nslots frequency slots for batched solves (clears frequency
    and factor state)."""
    ctx = ctx or default_context()
    L.check(L.lib().peec_frequency_slots(ctx.handle, int(nslots)), "peec_frequency_slots")


def frequency_set_slot(slot, w, DZ, Yle, ctx=None):
    ctx = ctx or default_context()
    dz = np.ascontiguousarray(DZ, dtype=np.complex128)
    Y, keep = _csc_c128(Yle)
    L.check(L.lib().peec_frequency_set_slot(ctx.handle, int(slot), float(w), L.as_c128p(dz), ctypes.byref(Y)),
            "peec_frequency_set_slot")


def edge_terms_set(R, Cd, Zs0, Lp, surface, esc_R, esc_freq, ctx=None):
    """Per-edge data from which frequency_set_slot_edges computes DZ on the
    device: R (m x 4), Cd (m x 4, complex for lossy dielectrics), Zs0 (m,
    surface impedance at unit scaled frequency), Lp (m), surface (m flags),
    and the R / freq escalings."""
    ctx = ctx or default_context()
    R = np.ascontiguousarray(R, dtype=np.float64)
    Cd = np.ascontiguousarray(Cd, dtype=np.complex128)
    zs = np.ascontiguousarray(Zs0, dtype=np.complex128)
    lp = np.ascontiguousarray(Lp, dtype=np.float64)
    sf = np.ascontiguousarray(np.asarray(surface) != 0, dtype=np.int8)
    m = lp.shape[0]
    if R.shape != (m, 4) or Cd.shape != (m, 4) or zs.shape != (m,) or sf.shape != (m,):
        raise ValueError("edge_terms_set: R, Cd must be m x 4 and Zs0, surface length m")
    L.check(L.lib().peec_edge_terms_set(ctx.handle, m, L.as_dp(R), L.as_c128p(Cd), L.as_c128p(zs), L.as_dp(lp),
                                        sf.ctypes.data_as(ctypes.c_void_p), float(esc_R), float(esc_freq)),
            "peec_edge_terms_set")


def frequency_set_slot_edges(slot, w, Yle, ctx=None):
    """frequency_set_slot with DZ (and 1/DZ) computed on the device from
    the edge data of edge_terms_set."""
    ctx = ctx or default_context()
    Y, keep = _csc_c128(Yle)
    L.check(L.lib().peec_frequency_set_slot_edges(ctx.handle, int(slot), float(w), ctypes.byref(Y)),
            "peec_frequency_set_slot_edges")


def frequency_dz(slot, m, ctx=None):
    """DZ of a frequency slot (m edges), read back from the device."""
    ctx = ctx or default_context()
    out = np.zeros(int(m), dtype=np.complex128)
    L.check(L.lib().peec_frequency_dz(ctx.handle, int(slot), L.as_c128p(out)), "peec_frequency_dz")
    return out


def lu_factor_slot(slot, ctx=None):
    ctx = ctx or default_context()
    L.check(L.lib().peec_lu_factor_slot(ctx.handle, int(slot)), "peec_lu_factor_slot")


def lu_factor_slots(slot0, count, ctx=None):
    """Factor the SS of frequency slots slot0 .. slot0 + count - 1 on the
    device, all of them through each launch together."""
    ctx = ctx or default_context()
    L.check(L.lib().peec_lu_factor_slots(ctx.handle, int(slot0), int(count)), "peec_lu_factor_slots")


def precond_rhs_batched(IS, rhs_per_slot, ctx=None):
    """M^-1 [0; 0; is_r] for the columns of IS (n x nrhs)."""
    ctx = ctx or default_context()
    ISt = np.ascontiguousarray(np.asarray(IS, dtype=np.float64).T)  # column-major n x nrhs
    nrhs = ISt.shape[0]
    out = np.zeros((nrhs, ctx_size(ctx)), dtype=np.complex128)
    L.check(L.lib().peec_precond_rhs_batched(ctx.handle, nrhs, int(rhs_per_slot), L.as_dp(ISt), L.as_c128p(out)),
            "peec_precond_rhs_batched")
    return out.T


def gmres_batched(B, rhs_per_slot, tol, X0=None, restart=60, max_cycles=20, mixed=False, inner_tol=0.0, ctx=None,
                  adjoint=False):
    """Solve (M^-1 K) x_r = b_r (adjoint: (M^-H K^H) x_r = b_r) for the
    columns of B (N x nrhs), column r against frequency slot
    r / rhs_per_slot.  Returns (X, flag, relres, iters, cycles)."""
    ctx = ctx or default_context()
    Bt = np.ascontiguousarray(np.asarray(B, dtype=np.complex128).T)  # column-major N x nrhs
    nrhs, N = Bt.shape
    Xt = np.zeros((nrhs, N), dtype=np.complex128) if X0 is None else \
        np.ascontiguousarray(np.asarray(X0, dtype=np.complex128).T).copy()
    flag = np.zeros(nrhs, dtype=np.int32)
    relres = np.zeros(nrhs)
    iters = np.zeros(nrhs, dtype=np.int32)
    cycles = np.zeros(nrhs, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    fn = L.lib().peec_gmres_batched_adjoint if adjoint else L.lib().peec_gmres_batched
    L.check(fn(ctx.handle, nrhs, int(rhs_per_slot), L.as_c128p(Bt), float(tol), int(restart), int(max_cycles),
               int(bool(mixed)), float(inner_tol), L.as_c128p(Xt), flag.ctypes.data_as(ip), L.as_dp(relres),
               iters.ctypes.data_as(ip), cycles.ctypes.data_as(ip)), "peec_gmres_batched")
    return Xt.T, flag, relres, iters, cycles


def sweep_batch(port_nodes, is_value, nslots, x0=None, tol=1e-9, restart=100, max_cycles=20, mixed=True,
                inner_tol=0.0, want_x=False, ctx=None):
    """One batch of the quasi-static sweep on the device (peec_sweep_batch):
    for every frequency slot s < nslots and port c, the port excitation, its
    preconditioned right-hand side and the GMRES solve, warm-started from x0
    (N x P) or, when x0 is None, from the last slot of the previous call.
    Returns (Vp, X, flag, relres, iters, cycles): Vp[s, c1, c2] is the node
    voltage difference of port c2 in the solution for (slot s, port c1); X
    (N x nslots P) only when want_x."""
    ctx = ctx or default_context()
    pn = np.ascontiguousarray(np.asarray(port_nodes, dtype=np.int64).reshape(-1, 2))
    P = pn.shape[0]
    nrhs, N = int(nslots) * P, ctx_size(ctx)
    x0t = None if x0 is None else np.ascontiguousarray(np.asarray(x0, dtype=np.complex128).T)
    vp = np.zeros((nrhs, P), dtype=np.complex128)
    Xt = np.zeros((nrhs, N), dtype=np.complex128) if want_x else None
    flag = np.zeros(nrhs, dtype=np.int32)
    relres = np.zeros(nrhs)
    iters = np.zeros(nrhs, dtype=np.int32)
    cycles = np.zeros(nrhs, dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    L.check(L.lib().peec_sweep_batch(ctx.handle, P, pn.ctypes.data_as(L.i64p), float(is_value),
                                     None if x0t is None else L.as_c128p(x0t), float(tol), int(restart),
                                     int(max_cycles), int(bool(mixed)), float(inner_tol), L.as_c128p(vp),
                                     None if Xt is None else L.as_c128p(Xt), flag.ctypes.data_as(ip), L.as_dp(relres),
                                     iters.ctypes.data_as(ip), cycles.ctypes.data_as(ip)), "peec_sweep_batch")
    return (vp.reshape(int(nslots), P, P), None if Xt is None else Xt.T, flag, relres, iters, cycles)


def sweep_state(nports, ctx=None):
    """The last-slot solutions peec_sweep_batch keeps (N x nports)."""
    ctx = ctx or default_context()
    out = np.zeros((int(nports), ctx_size(ctx)), dtype=np.complex128)
    L.check(L.lib().peec_sweep_state(ctx.handle, int(nports), L.as_c128p(out)), "peec_sweep_state")
    return out.T


def apply_K_batched(X, rhs_per_slot=1, fp32=False, precond=False, adjoint=False, ctx=None):
    """K x (precond False / 0), M^-1 K x (precond True / 1) or M^-1 x
    (precond 2) for the columns of X (N x nrhs), column r against frequency
    slot r / rhs_per_slot, with complex128 or (fp32) complex64 FFT
    convolutions; adjoint: K^H x, M^-H K^H x or M^-H x."""
    ctx = ctx or default_context()
    Xt = np.ascontiguousarray(np.asarray(X, dtype=np.complex128).T)
    Yt = np.zeros_like(Xt)
    fn = L.lib().peec_apply_KH_batched if adjoint else L.lib().peec_apply_K_batched
    L.check(fn(ctx.handle, Xt.shape[0], int(rhs_per_slot), int(bool(fp32)), int(precond), L.as_c128p(Xt),
               L.as_c128p(Yt)), "peec_apply_K%s_batched" % ("H" if adjoint else ""))
    return Yt.T


def precond_rhs_batched_adjoint(G, rhs_per_slot, ctx=None):
    """M^-H [0; 0; g_r] for the columns of the complex node excitations G
    (n x nrhs)."""
    ctx = ctx or default_context()
    Gt = np.ascontiguousarray(np.asarray(G, dtype=np.complex128).T)
    nrhs, n = Gt.shape
    N = ctx_size(ctx)
    T = np.zeros((nrhs, N), dtype=np.complex128)
    L.check(L.lib().peec_precond_rhs_batched_adjoint(ctx.handle, nrhs, int(rhs_per_slot), L.as_c128p(Gt),
                                                     L.as_c128p(T)), "peec_precond_rhs_batched_adjoint")
    return T.T


def set_graphs(on, ctx=None):
    """Replay GMRES matvecs from HIP graphs (default) or launch kernels one by one."""
    ctx = ctx or default_context()
    L.check(L.lib().peec_set_graphs(ctx.handle, int(bool(on))), "peec_set_graphs")


def fft_tune(nrhs, fp64=True, fp32=True, verbose=False, progress=False, ctx=None):
    """Autotune the FFT-convolution passes for batched calls with nrhs
    right-hand sides; returns the summed best pass time in ms."""
    ctx = ctx or default_context()
    ms = ctypes.c_double(0.0)
    flags = int(fp64) | (int(fp32) << 1) | (int(verbose) << 2) | (int(progress) << 3)
    L.check(L.lib().peec_fft_tune(ctx.handle, int(nrhs), flags, ctypes.byref(ms)),
            "peec_fft_tune")
    return ms.value


def fft_set_choice(kernel, pass_, fp32, nrhs, variant, rpb=1, style=0, ctx=None):
    ctx = ctx or default_context()
    L.check(L.lib().peec_fft_set_choice(ctx.handle, int(kernel), int(pass_), int(bool(fp32)), int(nrhs),
                                        int(variant), int(rpb), int(style)), "peec_fft_set_choice")


def fft_choice(kernel, pass_, fp32, nrhs, ctx=None):
    """(variant, lines per tile, lanes per line, rhs per block, radix schedule)."""
    ctx = ctx or default_context()
    out = (ctypes.c_int * 5)()
    L.check(L.lib().peec_fft_choice(ctx.handle, int(kernel), int(pass_), int(bool(fp32)), int(nrhs), out),
            "peec_fft_choice")
    return tuple(out)


def lu_set_factors(Lf, Uf, p, q, Rs=None, ctx=None, blocks=None):
    """F = lu(SS) in Julia's convention L*U = (Rs .* SS)[p, q] (0-based p, q).
    blocks: optional block_ptr cutting the factor order into row blocks
    for the triangular solves (peec_lu_set_factors_blocked)."""
    ctx = ctx or default_context()
    Ls, k1 = _csc_c128(Lf)
    Us, k2 = _csc_c128(Uf)
    pp = np.ascontiguousarray(p, dtype=np.int64)
    qq = np.ascontiguousarray(q, dtype=np.int64)
    rs = None if Rs is None else np.ascontiguousarray(Rs, dtype=np.float64)
    rsp = None if rs is None else L.as_dp(rs)
    if blocks is None:
        L.check(L.lib().peec_lu_set_factors(ctx.handle, int(Lf.shape[0]), ctypes.byref(Ls), ctypes.byref(Us),
                                            pp.ctypes.data_as(L.i64p), qq.ctypes.data_as(L.i64p), rsp),
                "peec_lu_set_factors")
    else:
        bp = np.ascontiguousarray(blocks, dtype=np.int64)
        L.check(L.lib().peec_lu_set_factors_blocked(ctx.handle, int(Lf.shape[0]), ctypes.byref(Ls),
                                                    ctypes.byref(Us), pp.ctypes.data_as(L.i64p),
                                                    qq.ctypes.data_as(L.i64p), rsp, int(bp.size - 1),
                                                    bp.ctypes.data_as(L.i64p)), "peec_lu_set_factors_blocked")
    del k1, k2


def lu_analyze(dissection, ctx=None):
    """Symbolic factorization of SS for the device LU (peec_lu_analyze):
    `dissection` has order, front_ptr, front_parent and block_ptr
    (ordering.Dissection).  Once per geometry, after operator_set."""
    ctx = ctx or default_context()
    arrs = [np.ascontiguousarray(getattr(dissection, k), dtype=np.int64)
            for k in ("order", "front_ptr", "front_parent", "block_ptr")]
    order, fptr, fpar, bptr = arrs
    L.check(L.lib().peec_lu_analyze(ctx.handle, int(order.size), order.ctypes.data_as(L.i64p), int(fptr.size - 1),
                                    fptr.ctypes.data_as(L.i64p), fpar.ctypes.data_as(L.i64p), int(bptr.size - 1),
                                    bptr.ctypes.data_as(L.i64p)), "peec_lu_analyze")


def lu_factor(ctx=None):
    """Device LU of SS at the frequency of the last frequency_set
    (peec_lu_factor); replaces the host lu(SS) of FFT_solver_QS_S_type.jl:119."""
    ctx = ctx or default_context()
    L.check(L.lib().peec_lu_factor(ctx.handle), "peec_lu_factor")


def lu_info(ctx=None):
    """Sizes of the device factorization (peec_lu_info)."""
    ctx = ctx or default_context()
    out = (ctypes.c_int64 * 8)()
    L.check(L.lib().peec_lu_info(ctx.handle, out), "peec_lu_info")
    keys = ("n", "fronts", "max_front", "pool_per_slot", "pool_with_scratch", "ss_entries", "l_entries", "u_entries")
    return dict(zip(keys, (int(v) for v in out)))


def lu_set_solve(mode, ctx=None):
    """Triangular solves of the preconditioner (peec_lu_set_solve): 1 =
    multifrontal from the device factorization's fronts (default; takes
    effect with the next factorization of every slot at once), 0 = the
    level-scheduled 64-row blocks (at once)."""
    ctx = ctx or default_context()
    L.check(L.lib().peec_lu_set_solve(ctx.handle, int(mode)), "peec_lu_set_solve")


def lu_solve_info(ctx=None):
    """{mode, multifrontal (in use), height (assembly tree), max_tasks}
    (peec_lu_solve_info)."""
    ctx = ctx or default_context()
    out = (ctypes.c_int64 * 4)()
    L.check(L.lib().peec_lu_solve_info(ctx.handle, out), "peec_lu_solve_info")
    return dict(zip(("mode", "multifrontal", "height", "max_tasks"), (int(v) for v in out)))


def lu_front(K, ctx=None):
    """(s, positions, F): front K of the last device factorization -- its
    pivot count, the factor position of each row/column, and the f x f
    matrix (L below and U on/above the diagonal in the first s columns /
    rows, the update matrix in the rest)."""
    ctx = ctx or default_context()
    f, s, first = (ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64())
    L.check(L.lib().peec_lu_front(ctx.handle, int(K), ctypes.byref(f), ctypes.byref(s), ctypes.byref(first), None,
                                  None), "peec_lu_front")
    out = np.zeros(f.value * f.value, dtype=np.complex128)
    pos = np.zeros(f.value, dtype=np.int64)
    L.check(L.lib().peec_lu_front(ctx.handle, int(K), None, None, None, L.as_c128p(out), pos.ctypes.data_as(L.i64p)),
            "peec_lu_front")
    return s.value, pos, out.reshape(f.value, f.value, order="F")


def ComputeMatrixVectorFTT(x, ctx=None):
    """compute_Matrix_vector_fft.jl:1-81 (M^-1 K x) for the operator,
    frequency and LU factors loaded into ctx."""
    ctx = ctx or default_context()
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(xx)
    L.check(L.lib().peec_matvec(ctx.handle, L.as_c128p(xx), L.as_c128p(y)), "peec_matvec")
    return y


def apply_K(x, ctx=None):
    ctx = ctx or default_context()
    xx = np.ascontiguousarray(x, dtype=np.complex128)
    y = np.empty_like(xx)
    L.check(L.lib().peec_apply_K(ctx.handle, L.as_c128p(xx), L.as_c128p(y)), "peec_apply_K")
    return y


def precond_3_3_Kt_fft(is_, ctx=None):
    """FFT_solver_QS_S_type.jl:237-262: M^-1 [0; 0; is]."""
    ctx = ctx or default_context()
    v = np.ascontiguousarray(is_, dtype=np.float64)
    N = ctx_size(ctx)
    tn = np.empty(N, dtype=np.complex128)
    L.check(L.lib().peec_precond_rhs(ctx.handle, L.as_dp(v), L.as_c128p(tn)), "peec_precond_rhs")
    return tn


def gmres_custom(b, restarted, tol, maxit, x, ctx=None):
    """gmres_custom.jl:1-331.  Returns (x, flag, relres, iter, resvec);
    flag 99 means a stop was requested."""
    if restarted:
        raise ValueError("the restarted branch of gmres_custom is dead code in the reference (gmres_custom.jl:4-10)")
    ctx = ctx or default_context()
    bb = np.ascontiguousarray(b, dtype=np.complex128).reshape(-1)
    n = bb.size
    xx = np.ascontiguousarray(x if x is not None and len(x) else np.zeros(n), dtype=np.complex128).reshape(-1).copy()
    flag = ctypes.c_int()
    relres = ctypes.c_double()
    it = (ctypes.c_int64 * 2)()
    resvec = np.zeros(min(maxit, n) + 1)
    nres = ctypes.c_int64()
    rc = L.check(L.lib().peec_gmres(ctx.handle, L.as_c128p(bb), float(tol), int(maxit), L.as_c128p(xx),
                                    ctypes.byref(flag), ctypes.byref(relres), it, L.as_dp(resvec), ctypes.byref(nres)),
                 "peec_gmres")
    if rc == L.PEEC_STOPPED:
        return xx, 99, 0.0, [it[0], it[1]], 0
    return xx, flag.value, relres.value, [it[0], it[1]], resvec[:nres.value]


_sizes = {}


def ctx_size(ctx):
    return _sizes[id(ctx)]


def set_ctx_size(ctx, N):
    _sizes[id(ctx)] = N
