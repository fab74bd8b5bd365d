// Batched GMRES for a frequency sweep: many right-hand sides -- ports x
// frequency slots -- solved in lockstep, each against its own slot's
// preconditioned operator M^-1 K, with one batched matvec per iteration.
//
// The reference solves one port at one frequency at a time with
// unrestarted Householder GMRES (gmres_custom.jl, called from
// FFT_solver_QS_S_type.jl:140).  Here every right-hand side runs restarted
// GMRES with classical Gram-Schmidt (streaming passes of V' w and
// w - V h over the basis instead of j dependent ones), repeated where a pass
// cancels more than 1 - 1/sqrt(2) of ||w|| (the DGKS criterion), wrapped in iterative
// refinement: each cycle solves M^-1 K d = r to a relative tolerance, adds
// d to x, and recomputes the true residual r = M^-1 (b - K x) with the FP64
// operator.  With mixed = 1 the inner cycles apply K with the FFT
// convolutions in complex FP32 (FP32 spectra, FP32 grids); the triangular
// solves, the sparse terms, the orthogonalisation, the Givens least-squares
// and the refinement residual stay in FP64, so the converged x satisfies
// the FP64 residual test ||r|| <= tol ||b|| like the plain path.
//
// Right-hand side r uses frequency slot r / rhs_per_slot.  All basis
// vectors and iterates live in HBM, and so does the least squares: the
// Hessenberg columns, their Givens rotations, the rotated residual vector,
// the back-substitution and the per-column convergence flags are computed
// by one thread per right-hand side.  The host reads only the nrhs flags
// after each step, to shrink the next matvec's batch to the columns still
// iterating (reading them a step late would cost a wasted matvec of every
// converged column instead).
#include "context.h"

#include <algorithm>
#include <cmath>
#include <vector>



namespace {

typedef double2_t cd;

constexpr int kGx = 128;  // row blocks per right-hand side in the reductions

// part[(r * gx + bx) * 4 + wave] = partial sum of |v_r|^2
__global__ void __launch_bounds__(256) k_bnorm_part(const cd *v, int64_t ld, int64_t n, double *part)
{
    const int r = blockIdx.y;
    const cd *vr = v + (int64_t)r * ld;
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
        s += vr[i].x * vr[i].x + vr[i].y * vr[i].y;
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) part[((int64_t)r * gridDim.x + blockIdx.x) * 4 + (threadIdx.x >> 6)] = s;
}

// out[r] = sqrt(sum of the np partials of r)
__global__ void __launch_bounds__(256) k_bnorm_final(const double *part, int np, double *out)
{
    __shared__ double sh[4];
    const int r = blockIdx.x;
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[(int64_t)r * np + i];
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) out[r] = sqrt((sh[0] + sh[1]) + (sh[2] + sh[3]));
}

// part[((r * gx + bx) * 4 + wave) * ncol + k] = partial of conj(V_r[:, k]) . w_r
// part[((r * gx + bx) * 4 + wave) * ncol + k] = partial of conj(V_r[:, k]) . w_r
// Columns k go 8 at a time, so w is read once per 8 basis vectors.  Right-
// hand sides with mask[r] == 0 are skipped (mask may be null).
__global__ void __launch_bounds__(256) k_bdot_part(const cd *V, int64_t vstride, int ncol, const cd *w, int64_t wld,
                                                   int64_t n, cd *part, const int32_t *mask)
{
    const int r = blockIdx.y, lane = threadIdx.x & 63;
    if (mask && !mask[r]) return;
    const int64_t slot = ((int64_t)r * gridDim.x + blockIdx.x) * 4 + (threadIdx.x >> 6);
    const cd *wr = w + (int64_t)r * wld;
    const cd *Vr = V + (int64_t)r * vstride;
    for (int k0 = 0; k0 < ncol; k0 += 8) {
        double sx[8], sy[8];
#pragma unroll
        for (int q = 0; q < 8; q++) sx[q] = sy[q] = 0.0;
        for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
            const cd b = wr[i];
#pragma unroll
            for (int q = 0; q < 8; q++) {
                if (k0 + q < ncol) {
                    const cd a = Vr[(int64_t)(k0 + q) * n + i];
                    sx[q] += a.x * b.x + a.y * b.y;
                    sy[q] += a.x * b.y - a.y * b.x;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 8; q++) {
            if (k0 + q >= ncol) break;
            double x = sx[q], y = sy[q];
            for (int o = 32; o > 0; o >>= 1) {
                x += __shfl_xor(x, o, 64);
                y += __shfl_xor(y, o, 64);
            }
            if (lane == 0) part[slot * ncol + k0 + q] = {x, y};
        }
    }
}

// h[r * hs + k] = sum of the np partials of (r, k); grid (ncol, nrhs); zero
// for right-hand sides the mask skips
__global__ void __launch_bounds__(256) k_bdot_final(const cd *part, int np, int ncol, cd *h, int hs,
                                                    const int32_t *mask)
{
    __shared__ double shx[4], shy[4];
    const int k = blockIdx.x, r = blockIdx.y;
    if (mask && !mask[r]) {
        if (threadIdx.x == 0) h[(int64_t)r * hs + k] = {0.0, 0.0};
        return;
    }
    double sx = 0.0, sy = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) {
        const cd p = part[((int64_t)r * np + i) * ncol + k];
        sx += p.x;
        sy += p.y;
    }
    for (int o = 32; o > 0; o >>= 1) {
        sx += __shfl_xor(sx, o, 64);
        sy += __shfl_xor(sy, o, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        shx[threadIdx.x >> 6] = sx;
        shy[threadIdx.x >> 6] = sy;
    }
    __syncthreads();
    if (threadIdx.x == 0) h[(int64_t)r * hs + k] = {(shx[0] + shx[1]) + (shx[2] + shx[3]), (shy[0] + shy[1]) + (shy[2] + shy[3])};
}

// w_r += sign * V_r[:, 0:ncol) h_r (right-hand sides the mask skips untouched)
__global__ void k_bgemvn(cd *w, int64_t wld, const cd *V, int64_t vstride, int ncol, const cd *h, int hs, int64_t n,
                         double sign, const int32_t *mask)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n || (mask && !mask[r])) return;
    const cd *Vr = V + (int64_t)r * vstride;
    const cd *hr = h + (int64_t)r * hs;
    double ax = 0.0, ay = 0.0;
    for (int k = 0; k < ncol; k++) {
        const cd u = Vr[i + (int64_t)k * n], c = hr[k];
        ax += u.x * c.x - u.y * c.y;
        ay += u.x * c.y + u.y * c.x;
    }
    cd *p = w + (int64_t)r * wld + i;
    *p = {p->x + sign * ax, p->y + sign * ay};
}

// Selective reorthogonalisation (Daniel, Gragg, Kaufman, Stewart 1976): a
// second Gram-Schmidt pass only where the first one cancelled more than
// 1 - 1/sqrt(2) of ||w||.  mask2[r] = live[r] && ||w_after|| < ||w_before|| / sqrt(2)
__global__ void k_dgks_mask(const int32_t *live, const double *before, const double *after, int32_t *mask2, int nrhs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < nrhs) mask2[r] = (live[r] && after[r] < 0.70710678118654752 * before[r]) ? 1 : 0;
}

// dst_r = src_r / nrm_r (0 where nrm_r = 0)
__global__ void k_bscale(cd *dst, int64_t dld, const cd *src, int64_t sld, const double *nrm, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double s = nrm[r] > 0.0 ? 1.0 / nrm[r] : 0.0;
    const cd a = src[(int64_t)r * sld + i];
    dst[(int64_t)r * dld + i] = {a.x * s, a.y * s};
}

// dst_r = src_r (column copy with strides)
__global__ void k_bcopy(cd *dst, int64_t dld, const cd *src, int64_t sld, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i < n) dst[(int64_t)r * dld + i] = src[(int64_t)r * sld + i];
}

// dst[:, i] = src[:, col[i]] (stride sld) for the listed columns
__global__ void k_bgather(cd *dst, int64_t dld, const cd *src, int64_t sld, const int32_t *col, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i < n) dst[(int64_t)r * dld + i] = src[(int64_t)col[r] * sld + i];
}

// dst[:, col[i]] = src[:, i]
__global__ void k_bscatter(cd *dst, int64_t dld, const cd *src, int64_t sld, const int32_t *col, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i < n) dst[(int64_t)col[r] * dld + i] = src[(int64_t)r * sld + i];
}

// r = b - y
// Port excitations of a sweep batch: column r holds Is at node n1 and -Is at
// node n2 of port r % nports (node block at offset off); the rest is zero.
__global__ void k_port_rhs(cd *B, int64_t N, int64_t off, const int64_t *pn, int nports, int nrhs, double Is)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs) return;
    const int c = r % nports;
    cd *b = B + (int64_t)r * N + off;
    b[pn[2 * c]] = {Is, 0.0};
    b[pn[2 * c + 1]] = {-Is, 0.0};
}

// v[r * nports + c2] = x_r[n3] - x_r[n4] for the nodes (n3, n4) of port c2.
__global__ void k_port_volts(const cd *X, int64_t N, int64_t off, const int64_t *pn, int nports, int nrhs, cd *v)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrhs * nports) return;
    const int r = i / nports, c2 = i % nports;
    const cd *x = X + (int64_t)r * N + off;
    const cd a = x[pn[2 * c2]], b = x[pn[2 * c2 + 1]];
    v[i] = {a.x - b.x, a.y - b.y};
}

// x[0 .. n) = 0 (the solution of a zero right-hand side)
__global__ void k_bzero(cd *x, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = {0.0, 0.0};
}

__global__ void k_bsub(cd *r, const cd *b, const cd *y, int64_t total)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < total) r[i] = {b[i].x - y[i].x, b[i].y - y[i].y};
}

// ---- Arnoldi least squares on the device (one thread per right-hand side)
//
// Complex helpers in FP64.
__device__ __forceinline__ cd zmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd zconj(cd a) { return {a.x, -a.y}; }
__device__ __forceinline__ cd zadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ double zabs(cd a) { return hypot(a.x, a.y); }
// a / b
__device__ __forceinline__ cd zdiv(cd a, cd b)
{
    const double d = b.x * b.x + b.y * b.y;
    return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}

// Start of a restart cycle: g = beta e_1 for every right-hand side.
__global__ void k_ls_init(cd *G, int mp1, const double *beta, int nrhs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs) return;
    cd *g = G + (int64_t)r * mp1;
    g[0] = {beta[r], 0.0};
    for (int i = 1; i < mp1; i++) g[i] = {0.0, 0.0};
}

// Step j of the Givens QR of the Hessenberg matrix (gmres_custom.jl:148-163
// restated for Gram-Schmidt columns): column j is the sum of the two
// Gram-Schmidt passes (hbuf rows 0 .. j and m+1 .. m+1+j) with ||w|| below
// it; the previous rotations are applied, a new one zeroes the subdiagonal,
// g is rotated, and the right-hand side leaves the iteration (live[r] = 0)
// when |g_(j+1)| <= thr[r] (the cycle's target times its start residual)
// or the Krylov space is exhausted (||w|| = 0).
__global__ void k_ls_step(int j, int m, const cd *hbuf, int hs, const double *hn, cd *H, cd *CS, cd *SN, cd *G,
                          const double *thr, int32_t *live, int nrhs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs || !live[r]) return;
    const int mp1 = m + 1;
    const cd *h = hbuf + (int64_t)r * hs;
    cd *col = H + ((int64_t)r * m + j) * mp1;
    cd *cs = CS + (int64_t)r * m, *sn = SN + (int64_t)r * m, *g = G + (int64_t)r * mp1;
    for (int k = 0; k <= j; k++) col[k] = zadd(h[k], h[mp1 + k]);
    col[j + 1] = {hn[r], 0.0};
    for (int k = 0; k < j; k++) {
        const cd t = zadd(zmul(cs[k], col[k]), zmul(sn[k], col[k + 1]));
        const cd u = zadd(zmul({-sn[k].x, sn[k].y}, col[k]), zmul(zconj(cs[k]), col[k + 1]));
        col[k] = t;
        col[k + 1] = u;
    }
    const double a = zabs(col[j]), bb = zabs(col[j + 1]);
    cd c = {1.0, 0.0}, sv = {0.0, 0.0};
    if (bb > 0) {
        const double rr = hypot(a, bb);
        if (a == 0) {
            c = {0.0, 0.0};
            const cd cj = zconj(col[j + 1]);
            sv = {cj.x / bb, cj.y / bb};
        } else {
            c = {a / rr, 0.0};
            const cd ph = {col[j].x / a, col[j].y / a};
            const cd t = zmul(ph, zconj(col[j + 1]));
            sv = {t.x / rr, t.y / rr};
        }
    }
    cs[j] = c;
    sn[j] = sv;
    col[j] = zadd(zmul(c, col[j]), zmul(sv, col[j + 1]));
    col[j + 1] = {0.0, 0.0};
    const cd gj = g[j];
    g[j + 1] = zmul({-sv.x, sv.y}, gj);
    g[j] = zmul(c, gj);
    if (zabs(g[j + 1]) <= thr[r] || hn[r] == 0.0) live[r] = 0;
}

// y = R \ g over the kdone[r] leading rows, written into hbuf row r (zero
// beyond them and for right-hand sides with kdone[r] == 0), ready for
// x += V y.
__global__ void k_ls_solve(int m, const cd *H, const cd *G, const int32_t *kdone, cd *hbuf, int hs, int nrhs)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nrhs) return;
    const int mp1 = m + 1, k = kdone[r];
    const cd *Hr = H + (int64_t)r * m * mp1, *g = G + (int64_t)r * mp1;
    cd *y = hbuf + (int64_t)r * hs;
    for (int i = 0; i < hs; i++) y[i] = {0.0, 0.0};
    for (int i = k - 1; i >= 0; i--) {
        cd t = g[i];
        for (int l = i + 1; l < k; l++) {
            const cd p = zmul(Hr[(int64_t)l * mp1 + i], y[l]);
            t = {t.x - p.x, t.y - p.y};
        }
        y[i] = zdiv(t, Hr[(int64_t)i * mp1 + i]);
    }
}

inline unsigned nb256(int64_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

struct BWork {
    peec_ctx *ctx;
    int nrhs, rps, m;
    int64_t N;
    DevBuf<cd> V, B, X, R, W, T, Wc, hbuf, part;
    DevBuf<double> nrm, npart;
    DevBuf<int32_t> cmap;  // compacted columns: col[nrhs], rslot[nrhs], slot_ptr[nslots + 1]
    DevBuf<int32_t> masks; // live[nrhs] (right-hand sides still iterating), second-pass mask[nrhs]
    DevBuf<double> nrm2;   // ||w|| before and after the first Gram-Schmidt pass
    // least squares of the cycle: Hessenberg columns (m+1) x m, rotations,
    // rotated g, per-column exit thresholds and steps taken
    DevBuf<cd> Hd, CSd, SNd, Gd;
    DevBuf<double> thr;
    DevBuf<int32_t> kd;
    int alloc()
    {
        const size_t nr = (size_t)nrhs, n = (size_t)N;
        if (!Hd.alloc(nr * (size_t)m * (size_t)(m + 1)) || !CSd.alloc(nr * (size_t)m) || !SNd.alloc(nr * (size_t)m) ||
            !Gd.alloc(nr * (size_t)(m + 1)) || !thr.alloc(nr) || !kd.alloc(nr)) {
            peec_set_error("peec_gmres_batched: out of device memory for the least-squares state");
            return PEEC_ERR_NOMEM;
        }
        if (!V.alloc(nr * (size_t)(m + 1) * n) || !B.alloc(nr * n) || !X.alloc(nr * n) || !R.alloc(nr * n) ||
            !W.alloc(nr * n) || !T.alloc(nr * n) || !Wc.alloc(nr * n) || !masks.alloc(2 * nr) || !nrm2.alloc(2 * nr) ||
            !cmap.alloc(2 * nr + (size_t)ctx->nslots + 1) || !hbuf.alloc(nr * 2 * (size_t)(m + 1)) ||
            !part.alloc(nr * kGx * 4 * (size_t)(m + 1)) || !nrm.alloc(nr) || !npart.alloc(nr * kGx * 4)) {
            peec_set_error("peec_gmres_batched: out of device memory for %d right-hand sides x %d basis vectors", nrhs,
                           m + 1);
            return PEEC_ERR_NOMEM;
        }
        return PEEC_OK;
    }
    // nrm[r] = ||v_r||, v with stride ld; copies to host
    int norms(const cd *v, int64_t ld, std::vector<double> &out)
    {
        hipLaunchKernelGGL(k_bnorm_part, dim3(kGx, nrhs), dim3(256), 0, ctx->stream, v, ld, N, npart.p);
        hipLaunchKernelGGL(k_bnorm_final, dim3(nrhs), dim3(256), 0, ctx->stream, npart.p, kGx * 4, nrm.p);
        PEEC_HIP_TRY(hipGetLastError());
        out.resize((size_t)nrhs);
        PEEC_HIP_TRY(hipMemcpyAsync(out.data(), nrm.p, (size_t)nrhs * 8, hipMemcpyDeviceToHost, ctx->stream));
        PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
        return PEEC_OK;
    }
    // h = V[:, 0:ncol)' w into hbuf at column offset hoff (stride hs = 2(m+1))
    void dot(int ncol, const cd *w, int64_t wld, int hoff, const int32_t *mask = nullptr)
    {
        hipLaunchKernelGGL(k_bdot_part, dim3(kGx, nrhs), dim3(256), 0, ctx->stream, V.p, (int64_t)(m + 1) * N, ncol, w,
                           wld, N, part.p, mask);
        hipLaunchKernelGGL(k_bdot_final, dim3(ncol, nrhs), dim3(256), 0, ctx->stream, part.p, kGx * 4, ncol,
                           hbuf.p + hoff, 2 * (m + 1), mask);
    }
    void gemv(cd *w, int64_t wld, int ncol, const cd *h, int hs, double sign, const int32_t *mask = nullptr)
    {
        hipLaunchKernelGGL(k_bgemvn, dim3(nb256(N), nrhs), dim3(256), 0, ctx->stream, w, wld, V.p,
                           (int64_t)(m + 1) * N, ncol, h, hs, N, sign, mask);
    }
    // nrm_out[r] = ||v_r|| on the device (no copy to the host)
    void norms_device(const cd *v, int64_t ld, double *nrm_out)
    {
        hipLaunchKernelGGL(k_bnorm_part, dim3(kGx, nrhs), dim3(256), 0, ctx->stream, v, ld, N, npart.p);
        hipLaunchKernelGGL(k_bnorm_final, dim3(nrhs), dim3(256), 0, ctx->stream, npart.p, kGx * 4, nrm_out);
    }
    // W = A x (x with stride N), in FP64 or with FP32 convolutions; A =
    // M^-1 K, or M^-H K^H for the adjoint solve; replayed from HIP graphs
    bool adjoint = false;
    MatvecGraphs graphs;
    std::vector<int32_t> cols;  // the compacted columns now uploaded
    int wmax = 0;               // most of them in one slot
    void graphs_init()
    {
        if (!graphs.ctx) {
            graphs.ctx = ctx;
            graphs.nrhs = nrhs;
            graphs.rps = rps;
            graphs.ld = N;
            graphs.adjoint = adjoint;
        }
    }
    int matvec(const cd *x, bool fp32)
    {
        graphs_init();
        return graphs.run(x, W.p, fp32);
    }
    // W[:, c] = A x[:, c] for the listed columns c only (x with stride
    // xld): the columns are gathered into T, multiplied as one smaller
    // batch with an explicit column -> slot map, and scattered back; W of
    // the other columns is left as it was.  Uses every column when most are
    // listed.
    int matvec_cols(const cd *x, int64_t xld, const std::vector<int32_t> &c, bool fp32)
    {
        const int nc = (int)c.size();
        if (nc == 0) return PEEC_OK;
        if (nc * 8 > nrhs * 7) {  // most columns: no compaction
            if (x != T.p || xld != N)
                hipLaunchKernelGGL(k_bcopy, dim3(nb256(N), nrhs), dim3(256), 0, ctx->stream, T.p, N, x, xld, N);
            return matvec(T.p, fp32);
        }
        graphs_init();
        const int nsl = ctx->nslots;
        if (c != cols) {
            std::vector<int32_t> h((size_t)(2 * nrhs + nsl + 1), 0);
            for (int i = 0; i < nc; i++) {
                h[(size_t)i] = c[(size_t)i];
                h[(size_t)(nrhs + i)] = c[(size_t)i] / rps;
            }
            // columns grouped by slot (c is increasing): slot_ptr by counting
            for (int i = 0; i < nc; i++) h[(size_t)(2 * nrhs + c[(size_t)i] / rps + 1)]++;
            for (int sl = 0; sl < nsl; sl++) h[(size_t)(2 * nrhs + sl + 1)] += h[(size_t)(2 * nrhs + sl)];
            PEEC_HIP_TRY(hipMemcpyAsync(cmap.p, h.data(), h.size() * 4, hipMemcpyHostToDevice, ctx->stream));
            cols = c;
            wmax = 0;
            for (int sl = 0; sl < nsl; sl++)
                wmax = std::max(wmax, h[(size_t)(2 * nrhs + sl + 1)] - h[(size_t)(2 * nrhs + sl)]);
        }
        ColMap cm;
        cm.rslot = cmap.p + nrhs;
        cm.slot_ptr = cmap.p + 2 * nrhs;
        cm.rps = rps;
        cm.nsl = nsl;
        cm.wmax = wmax;
        hipLaunchKernelGGL(k_bgather, dim3(nb256(N), nc), dim3(256), 0, ctx->stream, T.p, N, x, xld, cmap.p, N);
        PEEC_HIP_TRY(hipGetLastError());
        int rc = graphs.run_cols(T.p, Wc.p, fp32, nc, cm);
        if (rc) return rc;
        hipLaunchKernelGGL(k_bscatter, dim3(nb256(N), nc), dim3(256), 0, ctx->stream, W.p, N, Wc.p, N, cmap.p, N);
        PEEC_HIP_TRY(hipGetLastError());
        return PEEC_OK;
    }
};

void bwork_free(BWork *w) { delete w; }

// The context's batched workspace for (nrhs, restart, N), reallocated only
// when the shape changes (the old one is freed first).  Matvec graphs are
// captured per call: buffers of the context may move between calls.
static BWork *bwork_acquire(peec_ctx *ctx, bool adjoint, int nrhs, int rps, int restart, int64_t N, int *rc)
{
    BWork *W = ctx->bwork.get();
    if (!W || W->nrhs != nrhs || W->m != restart || W->N != N) {
        ctx->bwork.reset();
        ctx->bwork.reset(new BWork());
        W = ctx->bwork.get();
        W->nrhs = nrhs;
        W->m = restart;
        W->N = N;
    }
    W->ctx = ctx;
    W->adjoint = adjoint;
    W->rps = rps;
    W->graphs.clear();
    W->graphs.ctx = nullptr;
    W->cols.clear();
    W->wmax = 0;
    if ((*rc = W->alloc())) {  // sizes unchanged: no reallocation
        ctx->bwork.reset();
        return nullptr;
    }
    return W;
}

static int gmres_batched_impl(peec_ctx *ctx, bool adjoint, int nrhs, int rhs_per_slot, const peec_c128 *b,
                              double tol, int restart, int max_cycles, int mixed, double inner_tol, peec_c128 *x,
                              int *flag, double *relres, int *iters, int *cycles);

extern "C" int peec_gmres_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, const peec_c128 *b, double tol,
                                  int restart, int max_cycles, int mixed, double inner_tol, peec_c128 *x, int *flag,
                                  double *relres, int *iters, int *cycles)
{
    return gmres_batched_impl(ctx, false, nrhs, rhs_per_slot, b, tol, restart, max_cycles, mixed, inner_tol, x, flag,
                              relres, iters, cycles);
}

extern "C" int peec_gmres_batched_adjoint(peec_ctx *ctx, int nrhs, int rhs_per_slot, const peec_c128 *b, double tol,
                                          int restart, int max_cycles, int mixed, double inner_tol, peec_c128 *x,
                                          int *flag, double *relres, int *iters, int *cycles)
{
    return gmres_batched_impl(ctx, true, nrhs, rhs_per_slot, b, tol, restart, max_cycles, mixed, inner_tol, x, flag,
                              relres, iters, cycles);
}

// Flag of a right-hand side whose norm or residual is not finite (the
// batched solve's analogue of MATLAB gmres flag 4: a scalar quantity became
// too small or too large to continue).
constexpr int kFlagNonFinite = 4;

// Per right-hand side results of a batched solve.
struct BOut {
    std::vector<int> fl, it, cyc;
    std::vector<double> nb, beta;
    void report(int nrhs, int *flag, double *relres, int *iters, int *cycles) const
    {
        for (int r = 0; r < nrhs; r++) {
            if (flag) flag[r] = fl[(size_t)r];
            const double b = nb[(size_t)r];
            if (relres) relres[r] = b > 0 && std::isfinite(b) ? beta[(size_t)r] / b : b == 0.0 ? 0.0 : NAN;
            if (iters) iters[r] = it[(size_t)r];
            if (cycles) cycles[r] = cyc[(size_t)r];
        }
    }
};

// The batched solve on W.B and W.X (device, N x nrhs at stride N): W.X holds
// the start and receives the solution.
static int bgmres_solve(peec_ctx *ctx, BWork &W, double tol, int max_cycles, int mixed, double inner_tol, BOut &o)
{
    const int nrhs = W.nrhs, restart = W.m;
    const int64_t N = W.N;
    int rc = PEEC_OK;
    std::vector<double> nb, beta;
    if ((rc = W.norms(W.B.p, N, nb))) return rc;
    // r = b - A x for the listed columns (the others keep r and beta)
    auto residual = [&](const std::vector<int32_t> &cl) -> int {
        int e = W.matvec_cols(W.X.p, N, cl, false);
        if (e) return e;
        hipLaunchKernelGGL(k_bsub, dim3(nb256((int64_t)nrhs * N)), dim3(256), 0, ctx->stream, W.R.p, W.B.p, W.W.p,
                           (int64_t)nrhs * N);
        const std::vector<double> old = beta;
        if ((e = W.norms(W.R.p, N, beta))) return e;
        if (!old.empty()) {
            std::vector<char> in((size_t)nrhs, 0);
            for (int32_t c : cl) in[(size_t)c] = 1;
            for (int r = 0; r < nrhs; r++)
                if (!in[(size_t)r]) beta[(size_t)r] = old[(size_t)r];
        }
        return PEEC_OK;
    };
    std::vector<int32_t> allc((size_t)nrhs);
    for (int r = 0; r < nrhs; r++) allc[(size_t)r] = r;
    if ((rc = residual(allc))) return rc;
    std::vector<char> active((size_t)nrhs);
    std::vector<int> it((size_t)nrhs, 0), cyc((size_t)nrhs, 0), fl((size_t)nrhs, 1), stall((size_t)nrhs, 0);
    for (int r = 0; r < nrhs; r++) {
        const double b = nb[(size_t)r], be = beta[(size_t)r];
        if (!std::isfinite(b) || !std::isfinite(be)) {
            // a non-finite right-hand side or start residual: flag 4 (a
            // quantity became too large or too small to continue), not a
            // convergence
            active[(size_t)r] = 0;
            fl[(size_t)r] = kFlagNonFinite;
        } else if (b == 0.0) {
            // a zero right-hand side has the zero solution (gmres_custom.jl:20-28)
            hipLaunchKernelGGL(k_bzero, dim3(nb256(N)), dim3(256), 0, ctx->stream, W.X.p + (int64_t)r * N, N);
            active[(size_t)r] = 0;
            fl[(size_t)r] = 0;
            beta[(size_t)r] = 0.0;
        } else {
            active[(size_t)r] = be > tol * b;
            if (!active[(size_t)r]) fl[(size_t)r] = 0;
        }
    }
    PEEC_HIP_TRY(hipGetLastError());
    const int m = restart, hs = 2 * (m + 1);
    std::vector<int32_t> flags((size_t)nrhs), kdone((size_t)nrhs);
    std::vector<double> thr((size_t)nrhs);
    int ncycle = 0;
    while (std::any_of(active.begin(), active.end(), [](char c) { return c != 0; }) && ncycle < max_cycles) {
        if (ctx->stop_requested()) return PEEC_STOPPED;
        ncycle++;
        // V[:, 0] = r / beta
        PEEC_HIP_TRY(hipMemcpyAsync(W.nrm.p, beta.data(), (size_t)nrhs * 8, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_bscale, dim3(nb256(N), nrhs), dim3(256), 0, ctx->stream, W.V.p, (int64_t)(m + 1) * N,
                           W.R.p, N, W.nrm.p, N);
        // g = beta e_1; exit threshold of each column: the final tolerance,
        // or (mixed) at most inner_tol (default 1e-5) per cycle, times beta
        hipLaunchKernelGGL(k_ls_init, dim3((nrhs + 63) / 64), dim3(64), 0, ctx->stream, W.Gd.p, m + 1, W.nrm.p, nrhs);
        std::vector<char> inner((size_t)nrhs);
        for (int r = 0; r < nrhs; r++) {
            inner[(size_t)r] = active[(size_t)r];
            flags[(size_t)r] = active[(size_t)r] ? 1 : 0;
            kdone[(size_t)r] = 0;
            const double target = tol * nb[(size_t)r] / std::max(beta[(size_t)r], 1e-300);
            const double eta = mixed ? std::max(target, inner_tol > 0 ? inner_tol : 1e-5) : std::max(target, 1e-15);
            thr[(size_t)r] = eta * beta[(size_t)r];
        }
        PEEC_HIP_TRY(hipMemcpyAsync(W.thr.p, thr.data(), (size_t)nrhs * 8, hipMemcpyHostToDevice, ctx->stream));
        // the device's live flags (masks[0 .. nrhs)) drive the Gram-Schmidt
        // passes and the least squares; the host reads them back after each
        // step to shrink the next matvec's batch
        PEEC_HIP_TRY(hipMemcpyAsync(W.masks.p, flags.data(), (size_t)nrhs * 4, hipMemcpyHostToDevice, ctx->stream));
        std::vector<int32_t> live;
        int jlast = 0;
        for (int j = 0; j < m; j++) {
            // w = A_in v_j for the right-hand sides still iterating (the
            // batch shrinks as they converge)
            live.clear();
            for (int r = 0; r < nrhs; r++)
                if (inner[(size_t)r]) live.push_back(r);
            if ((rc = W.matvec_cols(W.V.p + (int64_t)j * N, (int64_t)(m + 1) * N, live, mixed != 0))) return rc;
            // classical Gram-Schmidt h = V' w, w -= V h on the live right-
            // hand sides, and a second pass where it cancelled too much
            W.norms_device(W.W.p, N, W.nrm2.p);
            W.dot(j + 1, W.W.p, N, 0, W.masks.p);
            W.gemv(W.W.p, N, j + 1, W.hbuf.p, hs, -1.0, W.masks.p);
            W.norms_device(W.W.p, N, W.nrm2.p + nrhs);
            hipLaunchKernelGGL(k_dgks_mask, dim3((nrhs + 255) / 256), dim3(256), 0, ctx->stream, W.masks.p, W.nrm2.p,
                               W.nrm2.p + nrhs, W.masks.p + nrhs, nrhs);
            W.dot(j + 1, W.W.p, N, m + 1, W.masks.p + nrhs);
            W.gemv(W.W.p, N, j + 1, W.hbuf.p + (m + 1), hs, -1.0, W.masks.p + nrhs);
            // v_(j+1) = w / ||w||, then the least-squares step
            W.norms_device(W.W.p, N, W.nrm.p);
            hipLaunchKernelGGL(k_bscale, dim3(nb256(N), nrhs), dim3(256), 0, ctx->stream,
                               W.V.p + (int64_t)(j + 1) * N, (int64_t)(m + 1) * N, W.W.p, N, W.nrm.p, N);
            hipLaunchKernelGGL(k_ls_step, dim3((nrhs + 63) / 64), dim3(64), 0, ctx->stream, j, m, W.hbuf.p, hs,
                               W.nrm.p, W.Hd.p, W.CSd.p, W.SNd.p, W.Gd.p, W.thr.p, W.masks.p, nrhs);
            PEEC_HIP_TRY(hipGetLastError());
            PEEC_HIP_TRY(hipMemcpyAsync(flags.data(), W.masks.p, (size_t)nrhs * 4, hipMemcpyDeviceToHost, ctx->stream));
            PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
            bool any = false;
            for (int r : live) {
                kdone[(size_t)r] = j + 1;
                it[(size_t)r]++;
                inner[(size_t)r] = flags[(size_t)r] != 0;
                any = any || inner[(size_t)r];
            }
            jlast = j + 1;
            if (!any) break;
        }
        // x += V y with y = R \ g per right-hand side (zero where inactive);
        // only basis columns 0 .. jlast-1 were written this cycle
        for (int r = 0; r < nrhs; r++)
            if (!active[(size_t)r]) kdone[(size_t)r] = 0;
        PEEC_HIP_TRY(hipMemcpyAsync(W.kd.p, kdone.data(), (size_t)nrhs * 4, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_ls_solve, dim3((nrhs + 63) / 64), dim3(64), 0, ctx->stream, m, W.Hd.p, W.Gd.p, W.kd.p,
                           W.hbuf.p, hs, nrhs);
        if (jlast > 0) W.gemv(W.X.p, N, jlast, W.hbuf.p, hs, 1.0);
        PEEC_HIP_TRY(hipGetLastError());
        const std::vector<double> beta_old = beta;
        std::vector<int32_t> act;
        for (int r = 0; r < nrhs; r++)
            if (active[(size_t)r]) act.push_back(r);
        if ((rc = residual(act))) return rc;
        for (int r = 0; r < nrhs; r++) {
            if (!active[(size_t)r]) continue;
            cyc[(size_t)r]++;
            if (!std::isfinite(beta[(size_t)r])) {
                active[(size_t)r] = 0;
                fl[(size_t)r] = kFlagNonFinite;
            } else if (beta[(size_t)r] <= tol * nb[(size_t)r]) {
                active[(size_t)r] = 0;
                fl[(size_t)r] = 0;
            } else if (beta[(size_t)r] > 0.5 * beta_old[(size_t)r]) {
                // refinement no longer pays: two cycles without halving the residual
                if (++stall[(size_t)r] >= 2) {
                    active[(size_t)r] = 0;
                    fl[(size_t)r] = 3;
                }
            } else {
                stall[(size_t)r] = 0;
            }
        }
    }
    o.fl = fl;
    o.it = it;
    o.cyc = cyc;
    o.nb = nb;
    o.beta = beta;
    return PEEC_OK;
}

static int gmres_batched_impl(peec_ctx *ctx, bool adjoint, int nrhs, int rhs_per_slot, const peec_c128 *b,
                              double tol, int restart, int max_cycles, int mixed, double inner_tol, peec_c128 *x,
                              int *flag, double *relres, int *iters, int *cycles)
{
    if (!ctx || nrhs < 1 || rhs_per_slot < 1 || !b || !x || restart < 1 || max_cycles < 1) {
        peec_set_error("peec_gmres_batched: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready || !ctx->lu_ready) {
        peec_set_error("peec_gmres_batched: operator, frequency slots and factors must be set first");
        return PEEC_ERR_STATE;
    }
    if ((nrhs + rhs_per_slot - 1) / rhs_per_slot > ctx->nslots) {
        peec_set_error("peec_gmres_batched: %d right-hand sides at %d per slot need more than the %d slots", nrhs,
                       rhs_per_slot, ctx->nslots);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t N = ctx->m + ctx->ns + ctx->n;
    int rc = PEEC_OK;
    BWork *Wp = bwork_acquire(ctx, adjoint, nrhs, rhs_per_slot, restart, N, &rc);
    if (!Wp) return rc;
    BWork &W = *Wp;
    const size_t bytes = (size_t)nrhs * (size_t)N * 16;
    PEEC_HIP_TRY(hipMemcpyAsync(W.B.p, b, bytes, hipMemcpyHostToDevice, ctx->stream));
    PEEC_HIP_TRY(hipMemcpyAsync(W.X.p, x, bytes, hipMemcpyHostToDevice, ctx->stream));
    BOut o;
    if ((rc = bgmres_solve(ctx, W, tol, max_cycles, mixed, inner_tol, o))) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(x, W.X.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    o.report(nrhs, flag, relres, iters, cycles);
    return PEEC_OK;
}

// One batch of the quasi-static sweep kept on the device: for every
// frequency slot and port, the port excitation, its preconditioned
// right-hand side and the GMRES solve (warm-started from x0, or from the
// previous call's last slot), returning only the port voltages.
extern "C" int peec_sweep_batch(peec_ctx *ctx, int nports, const int64_t *port_nodes, double is_value,
                                const peec_c128 *x0, double tol, int restart, int max_cycles, int mixed,
                                double inner_tol, peec_c128 *vports, peec_c128 *x_out, int *flag, double *relres,
                                int *iters, int *cycles)
{
    if (!ctx || nports < 1 || !port_nodes || !vports || restart < 1 || max_cycles < 1) {
        peec_set_error("peec_sweep_batch: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready || !ctx->lu_ready) {
        peec_set_error("peec_sweep_batch: operator, frequency slots and factors must be set first");
        return PEEC_ERR_STATE;
    }
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n, N = m + ns + n, off = m + ns;
    for (int i = 0; i < 2 * nports; i++)
        if (port_nodes[i] < 0 || port_nodes[i] >= n) {
            peec_set_error("peec_sweep_batch: port node %lld outside 0 .. %lld", (long long)port_nodes[i],
                           (long long)(n - 1));
            return PEEC_ERR_ARG;
        }
    const int nsl = ctx->nslots, nrhs = nsl * nports;
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    int rc = PEEC_OK;
    BWork *Wp = bwork_acquire(ctx, false, nrhs, nports, restart, N, &rc);
    if (!Wp) return rc;
    BWork &W = *Wp;
    DevBuf<int64_t> pn;
    DevBuf<cd> vd;
    if (!pn.alloc((size_t)(2 * nports)) || !vd.alloc((size_t)nrhs * nports)) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpyAsync(pn.p, port_nodes, (size_t)(2 * nports) * 8, hipMemcpyHostToDevice, ctx->stream));
    // b = M^-1 [0; 0; is] for every column, in place
    PEEC_HIP_TRY(hipMemsetAsync(W.B.p, 0, (size_t)nrhs * N * 16, ctx->stream));
    hipLaunchKernelGGL(k_port_rhs, dim3((unsigned)((nrhs + 63) / 64)), dim3(64), 0, ctx->stream, W.B.p, N, off, pn.p,
                       nports, nrhs, is_value);
    PEEC_HIP_TRY(hipGetLastError());
    if ((rc = apply_precond(ctx, W.B.p, nrhs, N, uniform_cols(nrhs, nports)))) return rc;
    // warm start: x0, else the previous batch's last slot, else zero
    const size_t blk = (size_t)N * nports * 16;
    if (x0) {
        if (!ctx->sweep_x0.alloc((size_t)N * nports)) return PEEC_ERR_NOMEM;
        PEEC_HIP_TRY(hipMemcpyAsync(ctx->sweep_x0.p, x0, blk, hipMemcpyHostToDevice, ctx->stream));
        ctx->sweep_nports = nports;
    }
    if (ctx->sweep_x0.p && ctx->sweep_nports == nports) {
        for (int s = 0; s < nsl; s++)
            PEEC_HIP_TRY(hipMemcpyAsync(W.X.p + (size_t)s * N * nports, ctx->sweep_x0.p, blk,
                                        hipMemcpyDeviceToDevice, ctx->stream));
    } else {
        PEEC_HIP_TRY(hipMemsetAsync(W.X.p, 0, (size_t)nrhs * N * 16, ctx->stream));
    }
    BOut o;
    if ((rc = bgmres_solve(ctx, W, tol, max_cycles, mixed, inner_tol, o))) return rc;
    if (!ctx->sweep_x0.alloc((size_t)N * nports)) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpyAsync(ctx->sweep_x0.p, W.X.p + (size_t)(nsl - 1) * N * nports, blk,
                                hipMemcpyDeviceToDevice, ctx->stream));
    ctx->sweep_nports = nports;
    hipLaunchKernelGGL(k_port_volts, dim3((unsigned)((nrhs * nports + 255) / 256)), dim3(256), 0, ctx->stream, W.X.p,
                       N, off, pn.p, nports, nrhs, vd.p);
    PEEC_HIP_TRY(hipGetLastError());
    PEEC_HIP_TRY(hipMemcpyAsync(vports, vd.p, (size_t)nrhs * nports * 16, hipMemcpyDeviceToHost, ctx->stream));
    if (x_out)
        PEEC_HIP_TRY(hipMemcpyAsync(x_out, W.X.p, (size_t)nrhs * N * 16, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    o.report(nrhs, flag, relres, iters, cycles);
    return PEEC_OK;
}

// The warm start peec_sweep_batch keeps (N x nports): copied out to x.
extern "C" int peec_sweep_state(peec_ctx *ctx, int nports, peec_c128 *x)
{
    if (!ctx || nports < 1 || !x) {
        peec_set_error("peec_sweep_state: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->sweep_x0.p || ctx->sweep_nports != nports) {
        peec_set_error("peec_sweep_state: no kept solution for %d ports", nports);
        return PEEC_ERR_STATE;
    }
    const size_t bytes = (size_t)(ctx->m + ctx->ns + ctx->n) * nports * 16;
    PEEC_HIP_TRY(hipMemcpyAsync(x, ctx->sweep_x0.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

// tn_r = M^-1 [0; 0; is_r] for each right-hand side (precond_3_3_Kt_fft!,
// FFT_solver_QS_S_type.jl:237-262, batched over slots)
extern "C" int peec_precond_rhs_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, const double *is, peec_c128 *tn)
{
    if (!ctx || nrhs < 1 || rhs_per_slot < 1 || !is || !tn) {
        peec_set_error("peec_precond_rhs_batched: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready || !ctx->lu_ready) {
        peec_set_error("peec_precond_rhs_batched: operator, frequency slots and factors must be set first");
        return PEEC_ERR_STATE;
    }
    if ((nrhs + rhs_per_slot - 1) / rhs_per_slot > ctx->nslots) {
        peec_set_error("peec_precond_rhs_batched: not enough frequency slots");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n, N = m + ns + n;
    std::vector<cd> h((size_t)(N * nrhs), cd{0.0, 0.0});
    for (int r = 0; r < nrhs; r++)
        for (int64_t v = 0; v < n; v++) h[(size_t)(r * N + m + ns + v)] = {is[r * n + v], 0.0};
    DevBuf<cd> Y;
    if (!Y.alloc(h.size())) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpyAsync(Y.p, h.data(), h.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    int rc = apply_precond(ctx, Y.p, nrhs, N, uniform_cols(nrhs, rhs_per_slot));
    if (rc) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(tn, Y.p, h.size() * 16, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

// tn_r = M^-H [0; 0; g_r] for complex node excitations g (n x nrhs,
// column-major): the right-hand side of the preconditioned adjoint system
// M^-H K^H lambda = M^-H [0; 0; g] (peec_gmres_batched_adjoint).
extern "C" int peec_precond_rhs_batched_adjoint(peec_ctx *ctx, int nrhs, int rhs_per_slot, const peec_c128 *g,
                                                peec_c128 *tn)
{
    if (!ctx || nrhs < 1 || rhs_per_slot < 1 || !g || !tn) {
        peec_set_error("peec_precond_rhs_batched_adjoint: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready || !ctx->lu_ready) {
        peec_set_error("peec_precond_rhs_batched_adjoint: operator, frequency slots and factors must be set first");
        return PEEC_ERR_STATE;
    }
    if ((nrhs + rhs_per_slot - 1) / rhs_per_slot > ctx->nslots) {
        peec_set_error("peec_precond_rhs_batched_adjoint: not enough frequency slots");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n, N = m + ns + n;
    std::vector<cd> h((size_t)(N * nrhs), cd{0.0, 0.0});
    for (int r = 0; r < nrhs; r++)
        for (int64_t v = 0; v < n; v++) h[(size_t)(r * N + m + ns + v)] = {g[r * n + v].re, g[r * n + v].im};
    DevBuf<cd> Y;
    if (!Y.alloc(h.size())) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpyAsync(Y.p, h.data(), h.size() * 16, hipMemcpyHostToDevice, ctx->stream));
    int rc = apply_precond_adj(ctx, Y.p, nrhs, N, uniform_cols(nrhs, rhs_per_slot));
    if (rc) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(tn, Y.p, h.size() * 16, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}
