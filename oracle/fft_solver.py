"""ORACLE -- test infrastructure only (see oracle/__init__.py).

numpy/scipy restatement of the reference's FFT-solver hot path, following
the reference text operation by operation:

  store_circulant_fft            compute_FFT_mutual_coupling_mats.jl:130-155
  kernel spectra (3 Lp, 6 P)     compute_FFT_mutual_coupling_mats.jl:36-128
  ComputeMatrixVectorFTT         compute_Matrix_vector_fft.jl:1-81
  precond_3_3_vector             compute_Matrix_vector_fft.jl:84-146 (3 LU solves)
  precond_3_3_Kt_fft!            FFT_solver_QS_S_type.jl:237-262
  gmres_custom                   gmres_custom.jl:1-331, scalarsignfft :661-668
  FFT_solver_QS_S_type           FFT_solver_QS_S_type.jl:1-234
  s2z_fft / s2y_fft              FFT_solver_QS_S_type.jl:265-285
  build_Yle_FFT                  build_Yle_FFT.jl:1-209
  compute_Z_self                 compute_Z_self.jl:1-15

FFTW is replaced by numpy's pocketfft and UMFPACK by scipy's SuperLU; both
give the same results up to roundoff.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import integrals as OI
from .circulant import P_CASES, centers


# ---------------------------------------------------------------- circulants
def kernel_dims(kind, Nx, Ny, Nz):
    from .circulant import grid_dims
    return grid_dims(kind, Nx, Ny, Nz)


def store_circulant_fft(row, Nx, Ny, Nz):
    """Even-symmetric 2x embedding; C[n] (0-based) stays zero."""
    C = np.asarray(row, dtype=np.float64).reshape(Nz, Ny, Nx).transpose(2, 1, 0)
    out = np.zeros((2 * Nx, 2 * Ny, 2 * Nz), dtype=np.complex128)
    ix = [np.arange(Nx), np.arange(Nx + 1, 2 * Nx)]
    rx = [np.arange(Nx), np.arange(Nx - 1, 0, -1)]
    iy = [np.arange(Ny), np.arange(Ny + 1, 2 * Ny)]
    ry = [np.arange(Ny), np.arange(Ny - 1, 0, -1)]
    iz = [np.arange(Nz), np.arange(Nz + 1, 2 * Nz)]
    rz = [np.arange(Nz), np.arange(Nz - 1, 0, -1)]
    for a in range(2):
        for b in range(2):
            for c in range(2):
                out[np.ix_(ix[a], iy[b], iz[c])] = C[np.ix_(rx[a], ry[b], rz[c])]
    return out


def kernel_rows(origin, Nx, Ny, Nz, sx, sy, sz, esc_Lp=1e6, esc_P=1e-12):
    """The 9 scaled circulant rows (3 Lp with the self entry zeroed, 6 P)."""
    rows = {}
    for d in range(3):
        c = centers(d, origin, Nx, Ny, Nz, (sx, sy, sz))
        if len(c) == 0:
            rows[d] = np.zeros(0)
            continue
        r = esc_Lp * OI.lp_row(c[0], c, sx, sy, sz, sx, sy, sz, d + 1)
        r[0] = 0.0
        rows[d] = r
    for i, (c1, c2) in enumerate(P_CASES):
        c = centers(3 + i, origin, Nx, Ny, Nz, (sx, sy, sz))
        rows[3 + i] = esc_P * OI.p_row(c[0], c, sx, sy, sz, c1, c2)
    return rows


def spectra(rows, Nx, Ny, Nz):
    out = {}
    for k in range(9):
        n = kernel_dims(k, Nx, Ny, Nz)
        if min(n) <= 0:
            out[k] = None
            continue
        out[k] = np.fft.fftn(store_circulant_fft(rows[k], *n))
    return out


# FFT worker threads of the convolutions (None: numpy's single-threaded
# pocketfft; an integer: scipy.fft with that many threads, the reference's
# multithreaded FFTW plan, compute_FFT_mutual_coupling_mats.jl:4).  Both
# compute every 1-D transform the same way; the threads split the batch.
FFT_WORKERS = None


def set_fft_workers(n):
    global FFT_WORKERS
    FFT_WORKERS = n


def conv(spec, vec, n):
    """ifft(spec .* fft(pad(vec))), first octant (compute_Matrix_vector_fft.jl:149-161)."""
    nx, ny, nz = n
    pad = np.zeros(spec.shape, dtype=np.complex128)
    pad[:nx, :ny, :nz] = vec.reshape(nz, ny, nx).transpose(2, 1, 0)
    if FFT_WORKERS:
        import scipy.fft as sfft
        out = sfft.ifftn(spec * sfft.fftn(pad, workers=FFT_WORKERS), workers=FFT_WORKERS)
    else:
        out = np.fft.ifftn(spec * np.fft.fftn(pad))
    return out[:nx, :ny, :nz].transpose(2, 1, 0).reshape(-1)


# ---------------------------------------------------------------- operator
class _OrderedLU:
    """F.solve(b) for F the LU of SS[order, order]."""

    def __init__(self, SS, order):
        self.order = np.asarray(order)
        self.F = spla.splu(SS[self.order][:, self.order].tocsc(), permc_spec="NATURAL", diag_pivot_thresh=0.0,
                           options=dict(SymmetricMode=True))

    def solve(self, b):
        x = np.empty_like(b)
        x[self.order] = self.F.solve(b[self.order])
        return x


def lu(SS, order=None):
    """Sparse LU of SS (UMFPACK's role in the reference, FFT_solver_QS_S_type.jl:119).

    SS is ill-conditioned at low frequency (cond ~ 1e9 at 10 MHz for the test
    geometries), so two different exact factorizations already disagree at
    the cond * eps level.  The oracle therefore takes the elimination order
    as an input and factors without pivoting, like the GPU path's host
    factorization: minimum degree on A'+A when no order is given, else
    SS[order, order] as it stands."""
    if order is not None:
        return _OrderedLU(SS, order)
    return spla.splu(SS, permc_spec="MMD_AT_PLUS_A", diag_pivot_thresh=0.0, options=dict(SymmetricMode=True))


class Operator:
    """Everything ComputeMatrixVectorFTT needs for one geometry + frequency."""

    def __init__(self, mesh, spec):
        self.mesh = mesh
        self.spec = spec
        self.m, self.ns, self.n = mesh.m, mesh.ns, mesh.n
        self.A = mesh.A.tocsr()
        self.Gamma = mesh.Gamma.tocsr()
        self.invP = sp.diags(1.0 / mesh.diag_P)

    def set_frequency(self, w, DZ, Yle, F=None, order=None):
        self.w = w
        self.DZ = np.asarray(DZ).reshape(-1)
        self.Yle = Yle.tocsr()
        self.invZ = sp.diags(1.0 / self.DZ)
        SS = self.Yle + self.A.T @ (self.invZ @ self.A) + 1j * w * (self.Gamma @ self.invP) @ self.Gamma.T
        self.SS = SS.tocsc()
        self.F = F if F is not None else lu(self.SS, order)

    def solve(self, b):
        return self.F.solve(np.asarray(b, dtype=np.complex128))

    def K(self, x):
        """The operator before preconditioning (compute_Matrix_vector_fft.jl:1-77)."""
        M = self.mesh
        m, ns = self.m, self.ns
        I, Q, Phi = x[:m], x[m:m + ns], x[m + ns:]
        Y1 = np.zeros(m, dtype=np.complex128)
        offs = [0, M.mx, M.mx + M.my]
        sizes = [M.mx, M.my, M.mz]
        for d in range(3):
            if sizes[d] == 0:
                continue
            n = kernel_dims(d, M.Nx, M.Ny, M.Nz)
            Ired = I[offs[d]:offs[d] + sizes[d]]
            e = M.mat_map_Lp[d] @ Ired
            Y1[offs[d]:offs[d] + sizes[d]] += M.mat_map_Lp[d].T @ conv(self.spec[d], e, n)
        Y1 = 1j * self.w * Y1 + self.DZ * I + self.A @ Phi
        Y2 = np.zeros(ns, dtype=np.complex128)
        kk = {(0, 0): 3, (1, 1): 4, (2, 2): 5, (0, 1): 6, (0, 2): 7, (1, 2): 8}
        for c1 in range(3):
            for c2 in range(c1, 3):
                k = kk[(c1, c2)]
                n = kernel_dims(k, M.Nx, M.Ny, M.Nz)
                Y2 = Y2 + M.exp_P[c2][c1].T @ conv(self.spec[k], M.exp_P[c1][c2] @ Q, n)
                if c1 != c2:
                    Y2 = Y2 + M.exp_P[c1][c2].T @ conv(self.spec[k], M.exp_P[c2][c1] @ Q, n)
        Y2 = Y2 - self.Gamma.T @ Phi
        Y3 = -(self.A.T @ I) + self.Yle @ Phi + 1j * self.w * (self.Gamma @ Q)
        return Y1, Y2, Y3

    def precond(self, X1, X2, X3):
        """precond_3_3_vector with its three LU solves."""
        w = self.w
        Y1 = self.invZ @ X1
        M2 = self.solve(self.A.T @ (self.invZ @ X1))
        invP_X2 = self.invP @ X2
        Y2 = invP_X2.copy()
        M4 = self.solve(self.Gamma @ invP_X2)
        M5 = self.solve(X3)
        Y1 = Y1 - self.invZ @ (self.A @ M2)
        Y1 = Y1 + 1j * w * (self.invZ @ (self.A @ M4))
        Y1 = Y1 - self.invZ @ (self.A @ M5)
        Y2 = Y2 + self.invP @ (self.Gamma.T @ M2)
        Y2 = Y2 - 1j * w * (self.invP @ (self.Gamma.T @ M4))
        Y2 = Y2 + self.invP @ (self.Gamma.T @ M5)
        Y3 = M2 - 1j * w * M4 + M5
        return np.concatenate([Y1, Y2, Y3])

    def matvec(self, x):
        return self.precond(*self.K(np.asarray(x, dtype=np.complex128)))

    def precond_rhs(self, is_):
        M5 = self.solve(np.asarray(is_, dtype=np.complex128))
        Y1 = -(self.invZ @ (self.A @ M5))
        Y2 = self.invP @ (self.Gamma.T @ M5)
        return np.concatenate([Y1, Y2, M5])


# ---------------------------------------------------------------- GMRES
def scalarsign(d):
    a = abs(d)
    return 1.0 + 0j if a == 0 else d / a


def gmres_custom(b, tol, maxit, x0, matvec, stop=lambda: False):
    """gmres_custom.jl:1-331 with restarted = false.
    Returns (x, flag, relres, iter, resvec)."""
    b = np.asarray(b, dtype=np.complex128)
    n = b.size
    outer = 1
    if maxit > n:
        maxit = n
    inner = maxit
    n2b = np.linalg.norm(b)
    if n2b == 0:
        return np.zeros(n, dtype=np.complex128), 0, 0.0, [0, 0], np.array([0.0])
    x = np.array(x0, dtype=np.complex128) if x0 is not None and len(x0) else np.zeros(n, dtype=np.complex128)
    flag = 1
    xmin = x.copy()
    imin = jmin = 0
    tolb = tol * n2b
    xm = x.copy()
    evalxm = 0
    stag = 0
    moresteps = 0
    maxmsteps = int(min(np.floor(n / 50), 5, n - maxit))
    maxstagsteps = 3
    minupdated = 0
    r = b - matvec(x)
    normr = np.linalg.norm(r)
    if normr <= tolb:
        return xmin, 0, normr / n2b, [0, 0], np.array([normr])
    n2minv_b = np.linalg.norm(b)
    tolb = tol * n2minv_b
    resvec = np.zeros(inner * outer + 1)
    resvec[0] = normr
    normrmin = normr
    J = np.zeros((2, inner), dtype=np.complex128)
    U = np.zeros((n, inner + 1), dtype=np.complex128)
    R = np.zeros((inner, inner), dtype=np.complex128)
    w = np.zeros(inner + 1, dtype=np.complex128)
    outitercount = 0
    initercount = 0
    normr_act = normr
    it = [0, 0]
    for outiter in range(1, outer + 1):
        outitercount += 1
        u = r.copy()
        beta = scalarsign(r[0]) * normr
        u[0] += beta
        u = u / np.linalg.norm(u)
        U[:, 0] = u
        w[0] = -beta
        for initer in range(1, inner + 1):
            if stop():
                return x, 99, 0, [outiter, initer], 0
            initercount += 1
            v = -2 * np.conj(u[initer - 1]) * u
            v[initer - 1] += 1
            for k in range(initer - 2, -1, -1):
                v = v - U[:, k] * (2 * np.vdot(U[:, k], v))
            v = v / np.linalg.norm(v)
            v = matvec(v)
            for k in range(0, initer):
                v = v - U[:, k] * (2 * np.vdot(U[:, k], v))
            if initer != len(v):
                u = v.copy()
                u[:initer] = 0
                alpha = np.linalg.norm(u)
                if alpha != 0:
                    alpha = scalarsign(v[initer]) * alpha
                    u[initer] += alpha
                    u = u / np.linalg.norm(u)
                    U[:, initer] = u
                    v[initer + 1:] = 0
                    v[initer] = -alpha
            for colJ in range(0, initer - 1):
                tmpv = v[colJ]
                v[colJ] = np.conj(J[0, colJ]) * v[colJ] + np.conj(J[1, colJ]) * v[colJ + 1]
                v[colJ + 1] = -J[1, colJ] * tmpv + J[0, colJ] * v[colJ + 1]
            if initer != len(v):
                rho = np.linalg.norm(v[initer - 1:initer + 1])
                J[:, initer - 1] = v[initer - 1:initer + 1] / rho
                w[initer] = -J[1, initer - 1] * w[initer - 1]
                w[initer - 1] = np.conj(J[0, initer - 1]) * w[initer - 1]
                v[initer - 1] = rho
                v[initer] = 0
            R[:, initer - 1] = v[:inner]
            normr = abs(w[initer])
            resvec[(outiter - 1) * inner + initer] = normr
            normr_act = normr
            if normr <= tolb or stag >= maxstagsteps or moresteps == 1:
                if evalxm == 0:
                    ytmp = _triu_solve(R[:initer, :initer], w[:initer])
                    additive = U[:, initer - 1] * (-2 * ytmp[initer - 1] * np.conj(U[initer - 1, initer - 1]))
                    additive[initer - 1] += ytmp[initer - 1]
                    for k in range(initer - 2, -1, -1):
                        additive[k] += ytmp[k]
                        additive = additive - U[:, k] * (2 * np.vdot(U[:, k], additive))
                    if np.linalg.norm(additive) < np.finfo(float).eps * np.linalg.norm(x):
                        stag += 1
                    else:
                        stag = 0
                    xm = x + additive
                    evalxm = 1
                elif evalxm == 1:
                    last = w[initer - 1] / R[initer - 1, initer - 1]
                    addvc = np.concatenate([-_triu_solve(R[:initer - 1, :initer - 1], R[:initer - 1, initer - 1]) * last,
                                            [last]])
                    if np.linalg.norm(addvc) < np.finfo(float).eps * np.linalg.norm(xm):
                        stag += 1
                    else:
                        stag = 0
                    additive = U[:, initer - 1] * (-2 * addvc[initer - 1] * np.conj(U[initer - 1, initer - 1]))
                    additive[initer - 1] += addvc[initer - 1]
                    for k in range(initer - 2, -1, -1):
                        additive[k] += addvc[k]
                        additive = additive - U[:, k] * (2 * np.vdot(U[:, k], additive))
                    xm = xm + additive
                r = b - matvec(xm)
                if np.linalg.norm(r) <= tol * n2b:
                    x = xm
                    flag = 0
                    it = [outiter, initer]
                    break
                normr_act = np.linalg.norm(r)
                resvec[(outiter - 1) * inner + initer] = normr_act
                if normr_act <= normrmin:
                    normrmin = normr_act
                    imin, jmin = outiter, initer
                    xmin = xm.copy()
                    minupdated = 1
                if normr_act <= tolb:
                    x = xm
                    flag = 0
                    it = [outiter, initer]
                    break
                else:
                    if stag >= maxstagsteps and moresteps == 0:
                        stag = 0
                    moresteps += 1
                    if moresteps >= maxmsteps:
                        flag = 3
                        it = [outiter, initer]
                        break
            if normr_act <= normrmin:
                normrmin = normr_act
                imin, jmin = outiter, initer
                minupdated = 1
            if stag >= maxstagsteps:
                flag = 3
                break
        evalxm = 0
        if flag != 0:
            idx = jmin if minupdated == 1 else initercount
            if idx > 0:
                y = _triu_solve(R[:idx, :idx], w[:idx])
                additive = U[:, idx - 1] * (-2 * y[idx - 1] * np.conj(U[idx - 1, idx - 1]))
                additive[idx - 1] += y[idx - 1]
                for k in range(idx - 2, -1, -1):
                    additive[k] += y[k]
                    additive = additive - U[:, k] * (2 * np.vdot(U[:, k], additive))
                x = x + additive
            xmin = x.copy()
            r = b - matvec(x)
            normr_act = np.linalg.norm(r)
        if normr_act <= normrmin:
            xmin = x.copy()
            normrmin = normr_act
            imin, jmin = outiter, initercount
        if flag == 3:
            break
        if normr_act <= tolb:
            flag = 0
            it = [outitercount, initercount]
            break
        minupdated = 0
    if flag == 0:
        relres = normr_act / n2minv_b
    else:
        x = xmin
        it = [imin, jmin]
        relres = normr_act / n2minv_b
    resvec = resvec[:max(outitercount - 1, 0) * inner + initercount + 1]
    return x, flag, relres, it, resvec


def _triu_solve(R, b):
    n = len(b)
    y = np.zeros(n, dtype=np.complex128)
    for i in range(n - 1, -1, -1):
        y[i] = (b[i] - R[i, i + 1:n] @ y[i + 1:n]) / R[i, i]
    return y


# ---------------------------------------------------------------- sweep
def lumped_admittance(typ, R, L, C, w):
    """get_lumped_elements_admittance_fft, build_Yle_FFT.jl:128-209."""
    with np.errstate(divide="ignore", invalid="ignore"):
        def bad(y):
            return np.isnan(y) or np.isinf(y) or abs(y) < 1e-16

        # numpy scalars: a zero C or L gives an infinite or NaN term, as in
        # Julia, which the fallbacks below catch
        jw = np.complex128(1j * w)
        R, L, C = np.float64(R), np.float64(L), np.float64(C)
        if typ == 1:
            z = R + jw * L + 1 / (jw * C)
            if np.isnan(z) or np.isinf(z) or abs(z) < 1e-16:
                z = R + jw * L
            y = 1 / z
        elif typ == 2:
            y = 1 / (R + jw * L) + jw * C
            for alt in (lambda: 1 / R + jw * C, lambda: 1 / (R + jw * L), lambda: 1 / R,
                        lambda: 1 / (jw * L) + jw * C, lambda: jw * C):
                if bad(y):
                    y = alt()
        elif typ == 3:
            y = 1 / (R + 1 / (jw * C)) + 1 / (jw * L)
            for alt in (lambda: 1 / (R + 1 / (jw * C)), lambda: 1 / R + 1 / (jw * L), lambda: 1 / R,
                        lambda: 1 / (1 / (jw * C)) + 1 / (jw * L), lambda: 1 / (jw * L), lambda: 1 / (1 / (jw * C))):
                if bad(y):
                    y = alt()
        elif typ == 4:
            y = 1 / R + jw * C + 1 / (jw * L)
            for alt in (lambda: 1 / R + jw * C, lambda: jw * C + 1 / (jw * L), lambda: jw * C):
                if bad(y):
                    y = alt()
        else:
            y = 1 / R + 1 / (jw * L + 1 / (jw * C))
            for alt in (lambda: 1 / R + 1 / (1 / (jw * C)), lambda: 1 / R + 1 / (jw * L), lambda: 1 / R,
                        lambda: 1 / (jw * L + 1 / (jw * C)), lambda: 1 / (jw * L), lambda: 1 / (1 / (jw * C))):
                if bad(y):
                    y = alt()
        if np.isnan(y) or np.isinf(y):
            y = 0.0
    return complex(y)


def build_Yle(lumped, port_nodes, esc_Yle, n, w, R0):
    entries = {}

    def add(i, j, v):
        entries[(i, j)] = entries.get((i, j), 0.0) + v

    for le in lumped:
        n1, n2 = le["nodes"]
        y = lumped_admittance(le["type"], le["R"], le["L"], le["C"], w)
        add(n1, n1, y)
        add(n2, n2, y)
        add(n1, n2, -y)
        add(n2, n1, -y)
    for n1, n2 in port_nodes:
        add(n1, n1, 1 / R0)
        add(n2, n2, 1 / R0)
        add(n1, n2, -1 / R0)
        add(n2, n1, -1 / R0)
    if not entries:
        return sp.csc_matrix((n, n), dtype=np.complex128)
    keys = list(entries.keys())
    r = [k[0] for k in keys]
    c = [k[1] for k in keys]
    v = [entries[k] * esc_Yle for k in keys]
    return sp.csc_matrix((v, (r, c)), shape=(n, n), dtype=np.complex128)


def compute_Z_self(R, Cd, w):
    Z = np.zeros(R.shape[0], dtype=np.complex128)
    with np.errstate(divide="ignore", invalid="ignore"):
        for a in range(4):
            r = R[:, a]
            c = Cd[:, a]
            both = (r != 0) & (c != 0)
            onlyr = (r != 0) & ~both
            onlyc = (c != 0) & ~both
            Z[both] += 1 / (1 / r[both] + 1j * w * c[both])
            Z[onlyr] += r[onlyr]
            Z[onlyc] += 1 / (1j * w * c[onlyc])
    return Z


def frequency_terms(mesh, w):
    """DZ of FFT_solver_QS_S_type.jl:106-114 (w scaled by escalings['freq'])."""
    esc = mesh.escalings
    Z_self = compute_Z_self(mesh.diag_R, mesh.diag_Cd, w)
    Zs = esc["R"] * (mesh.Zs * np.sqrt(w / esc["freq"]))
    se = mesh.surface_edges
    d = np.real(Zs[se]) - np.real(Z_self[se])
    put0_self = se[d > 0]
    put0_zs = se[d < 0]
    Z_self[put0_self] = 0 + 1j * np.imag(Z_self[put0_self])
    Zs[put0_zs] = 0 + 1j * np.imag(Zs[put0_zs])
    DZ = Z_self + np.real(Zs)
    DZ = DZ + 1j * w * mesh.diag_Lp
    return DZ


def s2z(S, Zo):
    P, _, F = S.shape
    Z = np.zeros_like(S)
    I = np.eye(P)
    for k in range(F):
        Z[:, :, k] = Zo * np.linalg.solve(I - 1.0 * S[:, :, k], I + S[:, :, k])
    return Z


def s2y(S, Zo):
    """As written at FFT_solver_QS_S_type.jl:276-285 (multiplies by Zo)."""
    P, _, F = S.shape
    Y = np.zeros_like(S)
    I = np.eye(P)
    for k in range(F):
        Y[:, :, k] = Zo * np.linalg.solve(I + S[:, :, k], I - 1.0 * S[:, :, k])
    return Y


def fft_solver_qs(freq, mesh, spec, inner_iter, tol, R0=50.0, stop=lambda: False, record=None, order=None):
    """FFT_solver_QS_S_type.jl:1-234.  `spec`: the kernel spectra
    (QS_Rcc_FW = 1), or a function of the frequency in Hz returning them
    (QS_Rcc_FW = 2: the retarded spectra rebuilt per frequency, :61-64).
    `order`: elimination order for the LU of SS (see lu)."""
    esc = mesh.escalings
    freq = np.asarray(freq, dtype=np.float64) * esc["freq"]
    tol = np.broadcast_to(np.asarray(tol, dtype=np.float64), freq.shape)
    m, n, ns = mesh.m, mesh.n, mesh.ns
    w = 2 * np.pi * freq
    P = mesh.port_nodes.shape[0]
    S = np.zeros((P, P, len(freq)), dtype=np.complex128)
    Vrest = np.zeros((m + n + ns, P), dtype=np.complex128)
    op = Operator(mesh, None if callable(spec) else spec)
    is_ = np.zeros(n)
    for k in range(len(freq)):
        if stop():
            return None
        if callable(spec):
            op.spec = spec(freq[k] / esc["freq"])
        Yle = build_Yle(mesh.lumped_elements, mesh.port_nodes, esc["Yle"], n, w[k] / esc["freq"], R0)
        DZ = frequency_terms(mesh, w[k])
        op.set_frequency(w[k], DZ, Yle, order=order)
        for c1 in range(P):
            n1, n2 = mesh.port_nodes[c1]
            is_[n1] = esc["Is"]
            is_[n2] = -1 * esc["Is"]
            tn = op.precond_rhs(is_)
            V, flag, relres, it, resvec = gmres_custom(tn, tol[k], inner_iter, Vrest[:, c1], op.matvec, stop)
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
    return dict(S=S, Z=s2z(S, R0), Y=s2y(S, R0), f=freq / esc["freq"])
