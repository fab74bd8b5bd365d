"""FFT_solver_QS_S_type on the GPU path (FFT_solver_QS_S_type.jl:1-285).

The host runs the per-frequency bookkeeping of the reference -- Yle
(build_Yle_FFT.jl), Z_self (compute_Z_self.jl), the skin-effect switch and
DZ (FFT_solver_QS_S_type.jl:105-114) -- and the numeric LU of SS with
SuperLU (the reference uses UMFPACK on the host too, :118-119).  The kernel
spectra, the FFT matvec, the preconditioner's triangular solves and GMRES
run on the GPU through libpeec_hip.
"""
import hashlib
import time

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import mesher, ordering, peec


def lumped_admittance(typ, R, L, C, w):
    """get_lumped_elements_admittance_fft (build_Yle_FFT.jl:128-209)."""
    # numpy scalars: a zero C or L gives an infinite or NaN term (not a
    # Python ZeroDivisionError), which the fallbacks catch
    jw = np.complex128(1j * w)
    R, L, C = np.float64(R), np.float64(L), np.float64(C)
    with np.errstate(divide="ignore", invalid="ignore"):
        def ok(y):
            return not (np.isnan(y) or np.isinf(y) or abs(y) < 1e-16)

        if typ == 1:
            z = R + jw * L + 1 / (jw * C)
            if not (not (np.isnan(z) or np.isinf(z)) and abs(z) >= 1e-16):
                z = R + jw * L
            y = 1 / z
        else:
            chains = {
                2: [lambda: 1 / (R + jw * L) + jw * C, lambda: 1 / R + jw * C, lambda: 1 / (R + jw * L),
                    lambda: 1 / R, lambda: 1 / (jw * L) + jw * C, lambda: jw * C],
                3: [lambda: 1 / (R + 1 / (jw * C)) + 1 / (jw * L), lambda: 1 / (R + 1 / (jw * C)),
                    lambda: 1 / R + 1 / (jw * L), lambda: 1 / R, lambda: 1 / (1 / (jw * C)) + 1 / (jw * L),
                    lambda: 1 / (jw * L), lambda: 1 / (1 / (jw * C))],
                4: [lambda: 1 / R + jw * C + 1 / (jw * L), lambda: 1 / R + jw * C, lambda: jw * C + 1 / (jw * L),
                    lambda: jw * C],
                5: [lambda: 1 / R + 1 / (jw * L + 1 / (jw * C)), lambda: 1 / R + 1 / (1 / (jw * C)),
                    lambda: 1 / R + 1 / (jw * L), lambda: 1 / R, lambda: 1 / (jw * L + 1 / (jw * C)),
                    lambda: 1 / (jw * L), lambda: 1 / (1 / (jw * C))],
            }[typ]
            y = chains[0]()
            for f in chains[1:]:
                if not ok(y):
                    y = f()
        if np.isnan(y) or np.isinf(y):
            y = 0.0
    return complex(y)


def build_Yle_FFT(lumped_elements, port_nodes, escalings, n, w, val_chiusura):
    """build_Yle_FFT.jl:1-126 (w unscaled, values scaled by escalings['Yle'])."""
    acc = {}
    for le in lumped_elements:
        n1, n2 = le["nodes"]
        y = lumped_admittance(le["type"], le["R"], le["L"], le["C"], w)
        for (a, b), s in (((n1, n1), 1), ((n2, n2), 1), ((n1, n2), -1), ((n2, n1), -1)):
            acc[(a, b)] = acc.get((a, b), 0.0) + s * y
    for n1, n2 in port_nodes:
        for (a, b), s in (((n1, n1), 1), ((n2, n2), 1), ((n1, n2), -1), ((n2, n1), -1)):
            acc[(a, b)] = acc.get((a, b), 0.0) + s / val_chiusura
    if not acc:
        return sp.csc_matrix((n, n), dtype=np.complex128)
    ks = list(acc)
    return sp.csc_matrix(([acc[k] * escalings["Yle"] for k in ks], ([k[0] for k in ks], [k[1] for k in ks])),
                         shape=(n, n), dtype=np.complex128)


def compute_Z_self(R, Cd, w):
    """compute_Z_self.jl:1-15."""
    Z = np.zeros(R.shape[0], dtype=np.complex128)
    with np.errstate(divide="ignore", invalid="ignore"):
        for a in range(4):
            r, c = R[:, a], Cd[:, a]
            both = (r != 0) & (c != 0)
            only_r = (r != 0) & (c == 0)
            only_c = (r == 0) & (c != 0)
            Z[both] += 1 / (1 / r[both] + 1j * w * c[both])
            Z[only_r] += r[only_r]
            Z[only_c] += 1 / (1j * w * c[only_c])
    return Z


def DZ_of(mesh, w, sigma=None):
    """FFT_solver_QS_S_type.jl:106-114 with the scaled angular frequency w
    (sigma: per-voxel conductivities in place of the mesh's)."""
    esc = mesh.escalings
    diag_R, Zs0 = (mesh.diag_R, mesh.Zs) if sigma is None else mesher.conductor_terms(mesh, sigma)
    Z_self = compute_Z_self(diag_R, mesh.diag_Cd, w)
    Zs = esc["R"] * (Zs0 * np.sqrt(w / esc["freq"]))
    se = mesh.surface_edges
    d = np.real(Zs[se]) - np.real(Z_self[se])
    a, b = se[d > 0], se[d < 0]
    Z_self[a] = 1j * np.imag(Z_self[a])
    Zs[b] = 1j * np.imag(Zs[b])
    return Z_self + np.real(Zs) + 1j * w * mesh.diag_Lp


def dDZ_dsigma(mesh, w, sigma=None):
    """d DZ_e / d sigma_v (m x n sparse, complex) at the scaled angular
    frequency w, on the branch of the skin-effect switch DZ_of takes
    (FFT_solver_QS_S_type.jl:107-114): the half-voxel resistances
    R = l/2 / (sigma W T) (compute_diagonals.jl:10-16) enter Z_self, the
    surface impedance Zs ~ sigma^-1/2 (compute_Zs_with_indices.jl) enters
    Re DZ where it wins the switch.  The switch itself is not differentiable
    at its ties (Re Zs = Re Z_self)."""
    esc = mesh.escalings
    c = mesh.cond
    sig = mesh.sigma_vox if sigma is None else np.asarray(sigma, dtype=np.float64)
    diag_R, Zs0 = mesher.conductor_terms(mesh, sig)
    Cd = mesh.diag_Cd
    m = mesh.m
    Z_self = compute_Z_self(diag_R, Cd, w)
    Zs = esc["R"] * (Zs0 * np.sqrt(w / esc["freq"]))
    se = mesh.surface_edges
    d = np.real(Zs[se]) - np.real(Z_self[se])
    skin = np.zeros(m, dtype=bool)      # Re Zs replaces Re Z_self
    own = np.ones(m, dtype=bool)        # Z_self keeps its real part
    zs_on = np.zeros(m, dtype=bool)     # Re Zs is added
    skin[se[d > 0]] = True
    own[se[d > 0]] = False
    zs_on[se[d >= 0]] = True
    rows, cols, vals = [], [], []
    for slot in range(4):
        v = c["slot_vox"][:, slot]
        e = np.nonzero(v >= 0)[0]
        if e.size == 0:
            continue
        R = diag_R[e, slot]
        C = Cd[e, slot]
        # d f / d R with f = R (no Cd) or 1 / (1/R + j w Cd) (compute_Z_self.jl:4-12)
        with np.errstate(divide="ignore", invalid="ignore"):
            f = np.where(C != 0, 1 / (1 / R + 1j * w * C), R)
            dfdR = np.where(C != 0, (f / R) ** 2, 1.0)
        dval = dfdR * (-R / sig[v[e]])
        dval = np.where(own[e], dval, 1j * np.imag(dval))
        rows.append(e)
        cols.append(v[e])
        vals.append(dval)
    for h in range(2):
        v = c["zs_vox"][:, h]
        e = np.nonzero((v >= 0) & zs_on)[0]
        if e.size == 0:
            continue
        zh = esc["R"] * c["zs_geo"][e, h] * np.sqrt(2 * np.pi * 1e-7 / sig[v[e]]) * np.sqrt(w / esc["freq"])
        rows.append(e)
        cols.append(v[e])
        vals.append(np.real(-zh / (2 * sig[v[e]])) + 0j)
    if not rows:
        return sp.csr_matrix((m, mesh.n), dtype=np.complex128)
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(m, mesh.n))


def SS_of(mesh, w, DZ, Yle):
    """SS = Yle + A' Z^-1 A + j w Gamma P^-1 Gamma' (FFT_solver_QS_S_type.jl:118)."""
    invZ = sp.diags(1.0 / DZ)
    invP = sp.diags(1.0 / mesh.diag_P)
    A = mesh.A.tocsr()
    G = mesh.Gamma.tocsr()
    return (Yle + A.T @ (invZ @ A) + 1j * w * (G @ invP) @ G.T).tocsc()


def lu_factor(SS, order=None):
    """Host LU of SS (the reference's lu(SS), :119) with SuperLU, no pivoting
    (SS is complex symmetric and diagonally weighted), returned in Julia's
    convention L*U = SS[p, q].  With `order` (a nested-dissection order,
    ordering.mesh_order) SS[order, order] is factored as it stands;
    otherwise SuperLU orders by minimum degree on A'+A."""
    if order is None:
        F = spla.splu(SS, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))
        return F.L, F.U, np.argsort(F.perm_r), np.argsort(F.perm_c)
    F = spla.splu(SS[order][:, order].tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                  options=dict(SymmetricMode=True))
    return F.L, F.U, order[np.argsort(F.perm_r)], order[np.argsort(F.perm_c)]


class Solver:
    """Holds one geometry on one device context."""

    def __init__(self, mesh, ctx=None, nested_dissection=True, device_lu=False, QS_Rcc_FW=1):
        """device_lu: factor SS on the GPU (peec_lu_factor) instead of with
        SuperLU on the host; both use the nested-dissection order.
        QS_Rcc_FW = 2: retarded (Rcc) kernels, whose spectra are rebuilt per
        frequency (FFT_solver_QS_S_type.jl:61-64)."""
        if QS_Rcc_FW not in (1, 2):
            raise ValueError("QS_Rcc_FW must be 1 (quasi-static) or 2 (retarded Rcc)")
        self.mesh = mesh
        self.ctx = ctx or peec.default_context()
        self.device_lu = device_lu
        self.rcc = QS_Rcc_FW == 2
        self.dissection = ordering.mesh_dissection(mesh) if (nested_dissection or device_lu) else None
        self.order, self.blocks = ((self.dissection.order, self.dissection.block_ptr) if self.dissection is not None
                                   else (None, None))
        build = peec.compute_rows_Rcc if self.rcc else peec.compute_FFT_mutual_coupling_mats
        ok = build(mesh.origin, mesh.Nx, mesh.Ny, mesh.Nz, mesh.sx, mesh.sy, mesh.sz, mesh.escalings, ctx=self.ctx)
        if not ok:
            raise RuntimeError("stopped")
        peec.operator_set(mesh, ctx=self.ctx)
        peec.set_ctx_size(self.ctx, mesh.m + mesh.ns + mesh.n)
        if device_lu:
            peec.lu_analyze(self.dissection, ctx=self.ctx)
        self.tuned = {}

    def tune(self, nrhs, fp64=True, fp32=True, verbose=False, progress=False):
        """Autotune the FFT passes once for batched calls with nrhs
        right-hand sides (peec_fft_tune); returns the summed best pass ms."""
        key = (nrhs, fp64, fp32)
        if key not in self.tuned:
            self.tuned[key] = peec.fft_tune(nrhs, fp64=fp64, fp32=fp32, verbose=verbose, progress=progress,
                                            ctx=self.ctx)
        return self.tuned[key]

    def set_frequency(self, w, R0=50.0, sigma=None):
        m = self.mesh
        esc = m.escalings
        if self.rcc:
            peec.compute_Circulant_Rcc(w / (2 * np.pi * esc["freq"]), ctx=self.ctx)
        Yle = build_Yle_FFT(m.lumped_elements, m.port_nodes, esc, m.n, w / esc["freq"], R0)
        DZ = DZ_of(m, w, sigma)
        peec.frequency_set(w, DZ, Yle, ctx=self.ctx)
        if self.device_lu:
            peec.lu_factor(ctx=self.ctx)
            return
        Lf, Uf, p, q = lu_factor(SS_of(m, w, DZ, Yle), self.order)
        # the blocks hold when SuperLU kept the order (no pivoting, no
        # column postorder); otherwise the library finds its own
        blocks = self.blocks if self.order is not None and np.array_equal(p, self.order) and \
            np.array_equal(q, self.order) else None
        peec.lu_set_factors(Lf, Uf, p, q, ctx=self.ctx, blocks=blocks)

    def matvec(self, x):
        return peec.ComputeMatrixVectorFTT(x, ctx=self.ctx)


def FFT_solver_QS_S_type(freq, mesh, inner_iter, tol, ports_scatter_value=50.0, solver=None, record=None,
                         sigma=None, keep=False):
    """FFT_solver_QS_S_type.jl:1-234.  The solver's kernels decide between
    QS_Rcc_FW = 1 and 2 (Solver(..., QS_Rcc_FW=2) rebuilds the retarded
    spectra per frequency in set_frequency, :61-64).  Returns dict(S, Z, Y,
    f), or None when a stop was requested."""
    solver = solver or Solver(mesh)
    esc = mesh.escalings
    freq = np.asarray(freq, dtype=np.float64) * esc["freq"]
    tol = np.broadcast_to(np.asarray(tol, dtype=np.float64), freq.shape)
    m, n, ns = mesh.m, mesh.n, mesh.ns
    w = 2 * np.pi * freq
    P = mesh.port_nodes.shape[0]
    R0 = ports_scatter_value
    S = np.zeros((P, P, len(freq)), dtype=np.complex128)
    Vrest = np.zeros((m + n + ns, P), dtype=np.complex128)
    is_ = np.zeros(n)
    Vkeep = []
    for k in range(len(freq)):
        if solver.ctx._stop.value:
            return None
        solver.set_frequency(w[k], R0, sigma)
        for c1 in range(P):
            n1, n2 = mesh.port_nodes[c1]
            is_[n1] = esc["Is"]
            is_[n2] = -1 * esc["Is"]
            tn = peec.precond_3_3_Kt_fft(is_, ctx=solver.ctx)
            V, flag, relres, it, resvec = peec.gmres_custom(tn, False, tol[k], inner_iter, Vrest[:, c1],
                                                            ctx=solver.ctx)
            if flag == 99:
                return None
            if record is not None:
                record.append(dict(k=k, port=c1, flag=flag, iter=list(it), relres=relres, nres=len(resvec)))
            Vrest[:, c1] = V
            is_[n1] = 0
            is_[n2] = 0
            for c2 in range(c1, P):
                n3, n4 = mesh.port_nodes[c2]
                if c1 == c2:
                    S[c1, c2, k] = (2 * (V[m + ns + n3] - V[m + ns + n4]) - R0) / R0
                else:
                    S[c1, c2, k] = (2 * (V[m + ns + n3] - V[m + ns + n4])) / R0
                    S[c2, c1, k] = S[c1, c2, k]
        if keep:
            Vkeep.append(Vrest.copy())
    out = dict(S=S, Z=s2z_fft(S, R0), Y=s2y_fft(S, R0), f=freq / esc["freq"])
    if keep:
        out["V"] = Vkeep
    return out


def s2z_fft(S, Zo):
    """FFT_solver_QS_S_type.jl:265-274."""
    P = S.shape[0]
    I = np.eye(P)
    return np.stack([Zo * np.linalg.solve(I - S[:, :, k], I + S[:, :, k]) for k in range(S.shape[2])], axis=2)


def s2y_fft(S, Zo):
    """FFT_solver_QS_S_type.jl:276-285, reproduced as written (it multiplies
    by Zo where an admittance would divide)."""
    P = S.shape[0]
    I = np.eye(P)
    return np.stack([Zo * np.linalg.solve(I + S[:, :, k], I - S[:, :, k]) for k in range(S.shape[2])], axis=2)


def _device_edges(solver, sigma):
    """Upload the per-edge data DZ_of works from (for this sigma) once, so
    that every frequency's DZ is computed on the device."""
    mesh = solver.mesh
    key = None if sigma is None else hashlib.sha1(np.ascontiguousarray(sigma, dtype=np.float64)).hexdigest()
    if getattr(solver, "_edge_key", "unset") == key:
        return
    diag_R, Zs0 = (mesh.diag_R, mesh.Zs) if sigma is None else mesher.conductor_terms(mesh, sigma)
    surf = np.zeros(mesh.m, dtype=np.int8)
    surf[mesh.surface_edges] = 1
    peec.edge_terms_set(diag_R, mesh.diag_Cd, Zs0, mesh.diag_Lp, surf, mesh.escalings["R"], mesh.escalings["freq"],
                        ctx=solver.ctx)
    solver._edge_key = key


def set_slots(solver, ws, R0=50.0, timings=None, sigma=None):
    """Fill frequency slots 0 .. len(ws)-1 (already allocated) with the
    per-frequency terms at scaled angular frequencies ws and factor each
    slot's SS on the device (FFT_solver_QS_S_type.jl:105-119 per slot).
    DZ (DZ_of) is computed on the device from per-edge data uploaded once."""
    mesh = solver.mesh
    esc = mesh.escalings
    tm = timings if timings is not None else {}
    _device_edges(solver, sigma)
    t0 = time.perf_counter()
    for slot, w in enumerate(ws):
        Yle = build_Yle_FFT(mesh.lumped_elements, mesh.port_nodes, esc, mesh.n, w / esc["freq"], R0)
        peec.frequency_set_slot_edges(slot, w, Yle, ctx=solver.ctx)
    t1 = time.perf_counter()
    peec.lu_factor_slots(0, len(ws), ctx=solver.ctx)
    tm["freq_terms"] = tm.get("freq_terms", 0.0) + t1 - t0
    tm["lu"] = tm.get("lu", 0.0) + time.perf_counter() - t1


def FFT_solver_QS_batched(freq, mesh, tol, batch=8, mixed=True, restart=100, max_cycles=20, inner_tol=0.0,
                          ports_scatter_value=50.0, solver=None, record=None, timings=None, tune=False, Vrest=None,
                          sigma=None, keep=False):
    """The QS sweep of FFT_solver_QS_S_type.jl:60-232 with its frequency x
    port loop batched: `batch` frequencies at a time, every port of each,
    go through one peec_gmres_batched call (one batched matvec per
    iteration; factors per frequency slot on the device).  Warm starts
    chain across batches per port, like Vrest (:140, :196).  mixed = True
    runs the inner cycles with FP32 FFT convolutions.  Returns dict(S, Z, Y,
    f) as FFT_solver_QS_S_type.  `timings`, a dict, accumulates the seconds
    spent per phase (host frequency terms, device LU, preconditioned
    right-hand sides, GMRES); every library call returns synchronised.
    Without keep, each batch is one peec_sweep_batch call: excitations,
    preconditioned right-hand sides, solves and the warm start chained
    from batch to batch stay on the device (timed as "gmres"), and only
    the port voltages come back.
    tune = True autotunes the FFT passes for each batch width first
    (timed as "tune").  Vrest (N x P) is the warm start of the first batch
    (zeros by default); the result's "Vrest" is the last frequency's
    solution, to chain a sweep split over several calls.  sigma and keep as
    in FFT_solver_QS_S_type."""
    solver = solver or Solver(mesh, device_lu=True)
    tm = timings if timings is not None else {}
    for key in ("tune", "freq_terms", "lu", "precond_rhs", "gmres"):
        tm.setdefault(key, 0.0)
    if not solver.device_lu:
        raise ValueError("the batched sweep needs a Solver with device_lu=True")
    esc = mesh.escalings
    R0 = ports_scatter_value
    freq = np.asarray(freq, dtype=np.float64) * esc["freq"]
    F = freq.size
    m, n, ns = mesh.m, mesh.n, mesh.ns
    N = m + ns + n
    P = mesh.port_nodes.shape[0]
    w = 2 * np.pi * freq
    S = np.zeros((P, P, F), dtype=np.complex128)
    Xprev = np.zeros((N, P), dtype=np.complex128) if Vrest is None else np.asarray(Vrest, dtype=np.complex128)
    nslots = None
    Vkeep = []
    for g0 in range(0, F, batch):
        ks = list(range(g0, min(F, g0 + batch)))
        nb = len(ks)
        if nslots != nb:
            peec.frequency_slots(nb, ctx=solver.ctx)
            nslots = nb
        if tune:
            t0 = time.perf_counter()
            solver.tune(nb * P, fp64=True, fp32=mixed)
            tm["tune"] += time.perf_counter() - t0
        set_slots(solver, w[ks], R0, tm, sigma)
        if not keep:
            # right-hand sides, solves and warm starts stay on the device
            t0 = time.perf_counter()
            Vp, _, flag, relres, iters, cycles = peec.sweep_batch(mesh.port_nodes, esc["Is"], nb,
                                                                  Xprev if g0 == 0 else None, tol, restart,
                                                                  max_cycles, mixed, inner_tol, ctx=solver.ctx)
            tm["gmres"] += time.perf_counter() - t0
            for slot, k in enumerate(ks):
                for c1 in range(P):
                    if record is not None:
                        r = slot * P + c1
                        record.append(dict(k=k, port=c1, flag=int(flag[r]), relres=float(relres[r]),
                                           iters=int(iters[r]), cycles=int(cycles[r])))
                    for c2 in range(c1, P):
                        dv = Vp[slot, c1, c2]
                        S[c1, c2, k] = (2 * dv - R0) / R0 if c1 == c2 else 2 * dv / R0
                        S[c2, c1, k] = S[c1, c2, k]
            continue
        t0 = time.perf_counter()
        IS = np.zeros((n, nb * P))
        for slot in range(nb):
            for c1 in range(P):
                n1, n2 = mesh.port_nodes[c1]
                IS[n1, slot * P + c1] = esc["Is"]
                IS[n2, slot * P + c1] = -esc["Is"]
        TN = peec.precond_rhs_batched(IS, P, ctx=solver.ctx)
        X0 = np.tile(Xprev, (1, nb))
        t1 = time.perf_counter()
        X, flag, relres, iters, cycles = peec.gmres_batched(TN, P, tol, X0, restart, max_cycles, mixed, inner_tol,
                                                            ctx=solver.ctx)
        tm["precond_rhs"] += t1 - t0
        tm["gmres"] += time.perf_counter() - t1
        Xprev = X[:, (nb - 1) * P:nb * P]
        for slot, k in enumerate(ks):
            V = X[:, slot * P:(slot + 1) * P]
            if keep:
                Vkeep.append(np.array(V))
            for c1 in range(P):
                if record is not None:
                    r = slot * P + c1
                    record.append(dict(k=k, port=c1, flag=int(flag[r]), relres=float(relres[r]),
                                       iters=int(iters[r]), cycles=int(cycles[r])))
                for c2 in range(c1, P):
                    n3, n4 = mesh.port_nodes[c2]
                    dv = V[m + ns + n3, c1] - V[m + ns + n4, c1]
                    S[c1, c2, k] = (2 * dv - R0) / R0 if c1 == c2 else 2 * dv / R0
                    S[c2, c1, k] = S[c1, c2, k]
    if not keep and F > 0:
        Xprev = peec.sweep_state(P, ctx=solver.ctx)
    out = dict(S=S, Z=s2z_fft(S, R0), Y=s2y_fft(S, R0), f=freq / esc["freq"], Vrest=Xprev)
    if keep:
        out["V"] = Vkeep
    return out


# ---------------------------------------------------------------- gradients
def _port_vector(mesh, weights):
    """Node excitation sum_j weights[j] (e_{n1_j} - e_{n2_j}) (n, complex)."""
    g = np.zeros(mesh.n, dtype=np.complex128)
    for j, (n1, n2) in enumerate(mesh.port_nodes):
        g[n1] += weights[j]
        g[n2] -= weights[j]
    return g


def impedance_objective_gradient(freq, mesh, port=0, sigma=None, tol=1e-10, batch=8, restart=100, max_cycles=20,
                                 mixed=False, ports_scatter_value=50.0, solver=None, method="adjoint", record=None):
    """J = sum_f |Z_pp(f)|^2 and dJ/dsigma_v for every voxel (zero off the
    conductors), by the adjoint method.

    With Z = R0 (I - S)^-1 (I + S) and A = (I - S)^-1 per frequency,
    dZ_pp = 2 R0 sum_ij A_pi dS_ij A_jp; S_ij = 2 g_j' V_i / R0 - delta_ij
    (FFT_solver_QS_S_type.jl:203-205, V_i the solution for port i, g_j the
    node-voltage difference of port j) and dV_i = -K^-1 dK V_i with dK =
    diag(dDZ) on the edge block, so
        dZ_pp = -4 lambda_I' (dDZ .* U),  U = sum_i A_pi V_i (edge block),
    where K^T lambda = sum_j A_jp g_j.  method = "adjoint" solves
    K^H conj(lambda) = conj(sum_j A_jp g_j) on the device
    (peec_gmres_batched_adjoint: explicit K^H, M^-H from the forward
    factors), one right-hand side per frequency; method = "reciprocity"
    uses lambda_I = -sum_j A_jp V_j / Is (K^T = D1 K D2, no extra solve);
    method = "plain" takes the forward solutions from the plain FP64 path
    (gmres_custom per port, the reference's solve) with the reciprocity
    adjoint.

    Forward solves: the batched path (FP64, or mixed), all ports, keeping
    the solutions.  Returns dict(J, grad (n), Z, f, adjoint_flags)."""
    solver = solver or Solver(mesh, device_lu=True)
    esc = mesh.escalings
    R0 = ports_scatter_value
    if method == "plain":
        # the plain FP64 path's own solutions (gmres_custom per port) with
        # the reciprocity adjoint
        fw = FFT_solver_QS_S_type(freq, mesh, 400, tol, ports_scatter_value=R0, solver=solver, record=record,
                                  sigma=sigma, keep=True)
        method = "reciprocity"
    else:
        fw = FFT_solver_QS_batched(freq, mesh, tol, batch=batch, mixed=mixed, restart=restart, max_cycles=max_cycles,
                                   ports_scatter_value=R0, solver=solver, record=record, sigma=sigma, keep=True)
    S, Z = fw["S"], fw["Z"]
    F = len(freq)
    P = mesh.port_nodes.shape[0]
    m = mesh.m
    ws = 2 * np.pi * np.asarray(freq, dtype=np.float64) * esc["freq"]
    I = np.eye(P)
    Am = [np.linalg.inv(I - S[:, :, k]) for k in range(F)]
    lam = [None] * F
    flags = []
    if method == "adjoint":
        for g0 in range(0, F, batch):
            ks = list(range(g0, min(F, g0 + batch)))
            peec.frequency_slots(len(ks), ctx=solver.ctx)
            set_slots(solver, ws[ks], R0, sigma=sigma)
            G = np.stack([np.conj(_port_vector(mesh, Am[k][:, port])) for k in ks], axis=1)
            TN = peec.precond_rhs_batched_adjoint(G, 1, ctx=solver.ctx)
            eta, fl, relres, it, cyc = peec.gmres_batched(TN, 1, tol, None, restart, max_cycles, mixed,
                                                          ctx=solver.ctx, adjoint=True)
            flags += [int(x) for x in fl]
            for slot, k in enumerate(ks):
                lam[k] = np.conj(eta[:m, slot])
    elif method == "reciprocity":
        for k in range(F):
            lam[k] = -(fw["V"][k][:m, :] @ Am[k][:, port]) / esc["Is"]
    else:
        raise ValueError("method must be 'adjoint', 'reciprocity' or 'plain'")
    J = 0.0
    grad = np.zeros(mesh.n)
    for k in range(F):
        z, dZ = impedance_sensitivity(mesh, ws[k], S[:, :, k], fw["V"][k], lam[k], port, sigma)
        J += abs(z) ** 2
        grad += 2.0 * np.real(np.conj(z) * dZ)
    return dict(J=J, grad=grad, Z=Z, f=np.asarray(freq), adjoint_flags=flags)


def impedance_sensitivity(mesh, w, S, V, lam_I, port, sigma=None, R0=50.0):
    """(Z_pp, dZ_pp/dsigma (n)) at one frequency from the port scattering
    matrix S (P x P), the forward solutions V (N x P) and the edge block of
    the adjoint solution lam_I (K^T lambda = sum_j A_jp g_j):
    dZ_pp = -4 lambda_I' (dDZ .* sum_i A_pi V_i)."""
    P = S.shape[0]
    A = np.linalg.inv(np.eye(P) - S)
    Zm = R0 * A @ (np.eye(P) + S)
    U = V[:mesh.m, :] @ A[port, :]
    dZ = -4.0 * (dDZ_dsigma(mesh, w, sigma).T @ (lam_I * U))
    return Zm[port, port], dZ


def impedance_objective(freq, mesh, port=0, sigma=None, tol=1e-10, solver=None, inner_iter=400):
    """J = sum_f |Z_pp(f)|^2 with the plain FP64 path (one port at a time,
    gmres_custom; the reference's own solve), for finite differences."""
    out = FFT_solver_QS_S_type(freq, mesh, inner_iter, tol, solver=solver, sigma=sigma)
    return float(np.sum(np.abs(out["Z"][port, port, :]) ** 2)), out["Z"]
