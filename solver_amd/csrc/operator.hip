// The preconditioned PEEC operator y = M^-1 K x on the GPU.
//
// K x follows ComputeMatrixVectorFTT (compute_Matrix_vector_fft.jl:1-81):
//   Y1 = j w (Lp I) + DZ .* I + A Phi
//   Y2 = P Q - Gamma' Phi
//   Y3 = -A' I + Yle Phi + j w Gamma Q
// with Lp I and P Q as 12 FFT convolutions (fft.hip).  M^-1 follows
// precond_3_3_vector (:84-146).  The reference runs three LU solves
// (M2 = F\(A' Z^-1 X1), M4 = F\(Gamma P^-1 X2), M5 = F\X3) and only ever uses
// the combination M2 - j w M4 + M5, so by linearity one solve of
//   SS M = A' Z^-1 X1 - j w Gamma P^-1 X2 + X3
// gives the same vector; then
//   Y1 = Z^-1 (X1 - A M),  Y2 = P^-1 (X2 + Gamma' M),  Y3 = M.
//
// The LU factors of SS come from the host; the triangular solves are the
// blocked, level-scheduled ones of trsv.hip.
#include "context.h"

#include <algorithm>
#include <cstring>
#include <mutex>


void peec_ctx::release_operator()
{
    for (auto &a : apps) {
        a.map_in.reset();
        a.map_out.reset();
        a.kernel = -1;
        a.nnz_in = a.nnz_out = 0;
    }
    napps = 0;
    op_ready = false;
    freq_ready = false;
    lu_ready = false;
    // frequency slots and factor layouts are sized for the old geometry:
    // the next peec_frequency_set / peec_frequency_slots re-creates them
    nslots = 1;
    slot_set.clear();
    slot_lu.clear();
    hYle.clear();
    hw.clear();
    lu_layout_from_plan = false;
    lu.ready = false;
    mf_ready = false;
    mf.ready = false;
    mf_w.reset();
    mf_u.reset();
    mf_part.reset();
    fft_tune.clear();
    sweep_x0.reset();
    sweep_nports = 0;
    bwork.reset();
    edges_ready = false;
    for (auto &k : kern) {
        k.poly = false;
        k.pspec64.reset();
        k.pspec32.reset();
        k.fspec32.reset();
        k.pfspec32.reset();
    }
}

int conv_apply(peec_ctx *ctx, const ConvApp &app, bool fp32, const double2_t *vec_in, double2_t *vec_out, int nrhs,
               int64_t ld, void *w1, void *w2);
int64_t conv_work_elems(const ConvApp &app);
int poly_spectrum_build(peec_ctx *ctx, int k, const int delta[3]);
int yz_prepare(peec_ctx *ctx);
extern const int kHalfAxes[9][3];

namespace {

typedef double2_t cd;

__device__ __forceinline__ cd cadd(cd a, cd b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cd csub(cd a, cd b) { return {a.x - b.x, a.y - b.y}; }
__device__ __forceinline__ cd cmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd cscal(cd a, double s) { return {a.x * s, a.y * s}; }
__device__ __forceinline__ cd jw(cd a, double w) { return {-w * a.y, w * a.x}; }  // (j w) * a
__device__ __forceinline__ cd cdiv(cd a, cd b)
{
    double d = b.x * b.x + b.y * b.y;
    return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}

__global__ void k_zero(cd *y, int64_t n, int nrhs, int64_t ld)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    y[i + (int64_t)blockIdx.y * ld] = {0.0, 0.0};  // one right-hand side per grid row
}

// The sparse terms and the permutation scatter / gather run one thread per
// row and frequency slot (blockIdx.y = slot s; the columns of a batch are
// grouped by slot, cm.first(s) .. cm.end(s)).  Rows go to the XCDs in
// contiguous ranges (xcd_tile): neighbouring rows gather the same cache
// lines of the other sections, which then stay in one L2.  The thread takes the row's
// index data and the slot's DZ / 1/DZ / w once and gathers kRhsChunk
// right-hand sides per index, with one accumulator each, so the loads of a
// chunk are independent.  Each right-hand side's arithmetic is the
// per-column formula, in the same order.
constexpr int kRhsChunk = 8;

// Y1 = j w Y1 + DZ .* I + A Phi           (compute_Matrix_vector_fft.jl:29-31)
__global__ void k_y1(cd *y, const cd *x, const cd *DZ, const int32_t *ap, const int32_t *ai, const double *av,
                     const double *wv, int64_t m, int64_t phi0, int nrhs, int64_t ld, ColMap cm)
{
    const int64_t e = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    const int s = (int)blockIdx.y;
    const int rb = cm.first(s), re = cm.end(s, nrhs);
    const double w = wv[s];
    const cd dz = DZ[e + s * m];
    const int k0 = ap[e], k1 = ap[e + 1];
    for (int r0 = rb; r0 < re; r0 += kRhsChunk) {
        const int cnt = min(kRhsChunk, re - r0);
        const cd *x0 = x + r0 * ld;
        cd acc[kRhsChunk];
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) acc[j] = {0.0, 0.0};
        for (int k = k0; k < k1; k++) {
            const int64_t c = phi0 + ai[k];
            const double a = av[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) acc[j] = cadd(acc[j], cscal(x0[c + j * ld], a));
        }
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) {
            if (j < cnt) {
                const int64_t o = e + (r0 + j) * ld;
                cd v = jw(y[o], w);
                v = cadd(v, cmul(dz, x[o]));
                y[o] = cadd(v, acc[j]);
            }
        }
    }
}

// Y2 = Y2 - Gamma' Phi                    (:66-68)
__global__ void k_y2(cd *y, const cd *x, const int32_t *gp, const int32_t *gi, const double *gv, int64_t m,
                     int64_t ns, int64_t phi0, int nrhs, int64_t ld)
{
    int64_t s = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    {
        const int r = (int)blockIdx.y;  // one right-hand side per grid row
        const cd *xr = x + r * ld;
        cd a = {0.0, 0.0};
        for (int k = gp[s]; k < gp[s + 1]; k++) a = cadd(a, cscal(xr[phi0 + gi[k]], gv[k]));
        y[m + s + r * ld] = csub(y[m + s + r * ld], a);
    }
}

// Y3 = -A' I + Yle Phi + j w Gamma Q      (:69-77)
__global__ void k_y3(cd *y, const cd *x, const int32_t *atp, const int32_t *ati, const double *atv, const int32_t *yp0,
                     const int32_t *yi, const cd *yv, const int32_t *gp, const int32_t *gi, const double *gv,
                     const double *wv, int64_t m, int64_t ns, int64_t n, int nrhs, int64_t ld, ColMap cm)
{
    const int64_t v = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int64_t phi0 = m + ns;
    const int s = (int)blockIdx.y;
    const int rb = cm.first(s), re = cm.end(s, nrhs);
    const int32_t *yp = yp0 + s * (n + 1);
    const double w = wv[s];
    const int a0 = atp[v], a1 = atp[v + 1], y0 = yp[v], y1 = yp[v + 1], g0 = gp[v], g1 = gp[v + 1];
    for (int r0 = rb; r0 < re; r0 += kRhsChunk) {
        const int cnt = min(kRhsChunk, re - r0);
        const cd *x0 = x + r0 * ld;
        cd t[kRhsChunk], acc[kRhsChunk];
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) t[j] = {0.0, 0.0};
        for (int k = a0; k < a1; k++) {
            const int64_t c = ati[k];
            const double a = atv[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) t[j] = cadd(t[j], cscal(x0[c + j * ld], a));
        }
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) {
            t[j] = {-t[j].x, -t[j].y};
            acc[j] = {0.0, 0.0};
        }
        for (int k = y0; k < y1; k++) {
            const int64_t c = phi0 + yi[k];
            const cd a = yv[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) acc[j] = cadd(acc[j], cmul(a, x0[c + j * ld]));
        }
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) {
            t[j] = cadd(t[j], acc[j]);
            acc[j] = {0.0, 0.0};
        }
        for (int k = g0; k < g1; k++) {
            const int64_t c = m + gi[k];
            const double a = gv[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) acc[j] = cadd(acc[j], cscal(x0[c + j * ld], a));
        }
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++)
            if (j < cnt) y[phi0 + v + (r0 + j) * ld] = cadd(t[j], jw(acc[j], w));
    }
}

// t = A' (Z^-1 X1) - j w Gamma (P^-1 X2) + X3, scattered through the row
// permutation and row scaling of the LU: b[i] = Rs[p[i]] t[p[i]]
__global__ void k_pre(const cd *y, const cd *invZ0, const double *invP, const int32_t *atp, const int32_t *ati,
                      const double *atv, const int32_t *gp, const int32_t *gi, const double *gv, const int32_t *pinv,
                      const double *Rs, cd *b, const double *wv, int64_t m, int64_t ns, int64_t n, int nrhs,
                      int64_t ld, ColMap cm)
{
    const int64_t v = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int64_t phi0 = m + ns;
    const int s = (int)blockIdx.y;
    const int rb = cm.first(s), re = cm.end(s, nrhs);
    const cd *invZ = invZ0 + s * m;
    const double w = wv[s];
    const int a0 = atp[v], a1 = atp[v + 1], g0 = gp[v], g1 = gp[v + 1];
    const double rs = Rs ? Rs[v] : 1.0;
    const int64_t dst = pinv[v];
    for (int r0 = rb; r0 < re; r0 += kRhsChunk) {
        const int cnt = min(kRhsChunk, re - r0);
        const cd *y0 = y + r0 * ld;
        cd a[kRhsChunk], c[kRhsChunk];
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) a[j] = c[j] = {0.0, 0.0};
        for (int k = a0; k < a1; k++) {
            const int64_t e = ati[k];
            const cd iz = invZ[e];
            const double av = atv[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) a[j] = cadd(a[j], cscal(cmul(iz, y0[e + j * ld]), av));
        }
        for (int k = g0; k < g1; k++) {
            const int64_t q = gi[k];
            const double f = invP[q] * gv[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) c[j] = cadd(c[j], cscal(y0[m + q + j * ld], f));
        }
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) {
            if (j < cnt) {
                cd t = csub(a[j], jw(c[j], w));
                t = cadd(t, y0[phi0 + v + j * ld]);
                if (Rs) t = cscal(t, rs);
                b[dst + (r0 + j) * n] = t;
            }
        }
    }
}

// Y1 = Z^-1 (X1 - A M), Y2 = P^-1 (X2 + Gamma' M), Y3 = M (M = Q z, i.e.
// M[q[j]] = z[j]; read through the inverse column permutation)
__global__ void k_post(cd *y, const cd *z, const int32_t *qinv, const cd *invZ0, const double *invP, const int32_t *ap,
                       const int32_t *ai, const double *av, const int32_t *gtp, const int32_t *gti, const double *gtv,
                       int64_t m, int64_t ns, int64_t n, int nrhs, int64_t ld, ColMap cm)
{
    const int64_t i = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    const int64_t N = m + ns + n;
    if (i >= N) return;
    const int s = (int)blockIdx.y;
    const int rb = cm.first(s), re = cm.end(s, nrhs);
    if (i >= m + ns) {
        const int64_t src = qinv[i - m - ns];
        for (int r = rb; r < re; r++) y[i + r * ld] = z[src + r * n];
        return;
    }
    const bool edge = i < m;
    const int64_t s_ = i - m;
    const int k0 = edge ? ap[i] : gtp[s_], k1 = edge ? ap[i + 1] : gtp[s_ + 1];
    const int32_t *idx = edge ? ai : gti;
    const double *val = edge ? av : gtv;
    for (int r0 = rb; r0 < re; r0 += kRhsChunk) {
        const int cnt = min(kRhsChunk, re - r0);
        const cd *z0 = z + r0 * n;
        cd acc[kRhsChunk];
#pragma unroll
        for (int j = 0; j < kRhsChunk; j++) acc[j] = {0.0, 0.0};
        for (int k = k0; k < k1; k++) {
            const int64_t c = qinv[idx[k]];
            const double a = val[k];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++)
                if (j < cnt) acc[j] = cadd(acc[j], cscal(z0[c + j * n], a));
        }
        if (edge) {
            const cd iz = invZ0[s * m + i];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++) {
                if (j < cnt) {
                    const int64_t o = i + (r0 + j) * ld;
                    y[o] = cmul(iz, csub(y[o], acc[j]));
                }
            }
        } else {
            const double ip = invP[s_];
#pragma unroll
            for (int j = 0; j < kRhsChunk; j++) {
                if (j < cnt) {
                    const int64_t o = i + (r0 + j) * ld;
                    y[o] = cscal(cadd(y[o], acc[j]), ip);
                }
            }
        }
    }
}

}  // namespace

// ------------------------------------------------------------------ adjoint
// The conjugate transpose of K (the operator above):
//   Y1 = -j w (Lp I) + conj(DZ) .* I - A Phi
//   Y2 = P Q - j w Gamma' Phi
//   Y3 = A' I - Gamma Q + conj(Yle) Phi
// Lp and P are real and symmetric (each off-diagonal P block is applied in
// both directions), so the 12 convolutions are those of K; only the sparse
// terms and the frequency factors change.  Yle is symmetric.

// Y1 = -j w Y1 + conj(DZ) .* I - A Phi
__global__ void k_y1h(cd *y, const cd *x, const cd *DZ, const int32_t *ap, const int32_t *ai, const double *av,
                      const double *wv, int64_t m, int64_t phi0, int nrhs, int64_t ld, ColMap cm)
{
    int64_t e = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (e >= m) return;
    {
        const int r = (int)blockIdx.y;  // one right-hand side per grid row
        const int sl = cm.slot(r);
        const cd *xr = x + r * ld;
        cd s = {0.0, 0.0};
        for (int k = ap[e]; k < ap[e + 1]; k++) s = cadd(s, cscal(xr[phi0 + ai[k]], av[k]));
        const cd dz = DZ[e + sl * m];
        cd v = jw(y[e + r * ld], -wv[sl]);
        v = cadd(v, cmul(cd{dz.x, -dz.y}, xr[e]));
        y[e + r * ld] = csub(v, s);
    }
}

// Y2 = Y2 - j w Gamma' Phi
__global__ void k_y2h(cd *y, const cd *x, const int32_t *gp, const int32_t *gi, const double *gv, const double *wv,
                      int64_t m, int64_t ns, int64_t phi0, int nrhs, int64_t ld, ColMap cm)
{
    int64_t s = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (s >= ns) return;
    {
        const int r = (int)blockIdx.y;  // one right-hand side per grid row
        const cd *xr = x + r * ld;
        cd a = {0.0, 0.0};
        for (int k = gp[s]; k < gp[s + 1]; k++) a = cadd(a, cscal(xr[phi0 + gi[k]], gv[k]));
        y[m + s + r * ld] = csub(y[m + s + r * ld], jw(a, wv[cm.slot(r)]));
    }
}

// Y3 = A' I - Gamma Q + conj(Yle) Phi
__global__ void k_y3h(cd *y, const cd *x, const int32_t *atp, const int32_t *ati, const double *atv,
                      const int32_t *yp0, const int32_t *yi, const cd *yv, const int32_t *gp, const int32_t *gi,
                      const double *gv, int64_t m, int64_t ns, int64_t n, int nrhs, int64_t ld, ColMap cm)
{
    int64_t v = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (v >= n) return;
    const int64_t phi0 = m + ns;
    {
        const int r = (int)blockIdx.y;  // one right-hand side per grid row
        const int sl = cm.slot(r);
        const int32_t *yp = yp0 + sl * (n + 1);
        const cd *xr = x + r * ld;
        cd a = {0.0, 0.0};
        for (int k = atp[v]; k < atp[v + 1]; k++) a = cadd(a, cscal(xr[ati[k]], atv[k]));
        cd b = {0.0, 0.0};
        for (int k = yp[v]; k < yp[v + 1]; k++) b = cadd(b, cmul(cd{yv[k].x, -yv[k].y}, xr[phi0 + yi[k]]));
        cd c = {0.0, 0.0};
        for (int k = gp[v]; k < gp[v + 1]; k++) c = cadd(c, cscal(xr[m + gi[k]], gv[k]));
        y[phi0 + v + r * ld] = cadd(csub(a, c), b);
    }
}

// The diagonal similarity behind reciprocity: K^T = D1 K D2 and M^T = D1 M
// D2 with D1 = diag(I, j w, -I), D2 = diag(I, 1/(j w), -I) on the (I, Q, Phi)
// blocks, so M^-H y = conj(D2^-1 M^-1 D1^-1 conj(y)).  stage 0 applies
// y <- D1^-1 conj(y), stage 1 y <- conj(D2^-1 y).
__global__ void k_adj_xform(cd *y, const double *wv, int64_t m, int64_t ns, int64_t n, int nrhs, int64_t ld, ColMap cm,
                            int stage)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m + ns + n) return;
    {
        const int r = (int)blockIdx.y;  // one right-hand side per grid row
        const double w = wv[cm.slot(r)];
        cd a = y[i + r * ld];
        if (stage == 0) {
            a.y = -a.y;
            if (i >= m + ns) a = {-a.x, -a.y};
            else if (i >= m) a = {a.y / w, -a.x / w};  // a / (j w)
        } else {
            if (i >= m + ns) a = {-a.x, -a.y};
            else if (i >= m) a = jw(a, w);
            a.y = -a.y;
        }
        y[i + r * ld] = a;
    }
}

// ------------------------------------------------------------------ host helpers
static inline unsigned nblk(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// Julia CSC (1-based Int64) -> device CSR (0-based int32) of the matrix or of
// its transpose.
static int csc_to_csr(const peec_csc_f64 *a, bool transpose, CsrR *out, const char *what)
{
    if (!a || !a->colptr || (a->colptr[a->ncol] - 1 > 0 && (!a->rowval || !a->nzval))) {
        peec_set_error("operator_set: %s missing", what);
        return PEEC_ERR_ARG;
    }
    int64_t nr = transpose ? a->ncol : a->nrow, nc = transpose ? a->nrow : a->ncol;
    int64_t nnz = a->colptr[a->ncol] - 1;
    std::vector<int32_t> ptr((size_t)nr + 1, 0), idx((size_t)std::max<int64_t>(nnz, 1));
    std::vector<double> val((size_t)std::max<int64_t>(nnz, 1));
    for (int64_t c = 0; c < a->ncol; c++)
        for (int64_t k = a->colptr[c] - 1; k < a->colptr[c + 1] - 1; k++) {
            int64_t r = a->rowval[k] - 1;
            if (r < 0 || r >= a->nrow) {
                peec_set_error("operator_set: %s row index out of range", what);
                return PEEC_ERR_ARG;
            }
            ptr[(size_t)(transpose ? c : r) + 1]++;
        }
    for (int64_t i = 0; i < nr; i++) ptr[(size_t)i + 1] += ptr[(size_t)i];
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t c = 0; c < a->ncol; c++)
        for (int64_t k = a->colptr[c] - 1; k < a->colptr[c + 1] - 1; k++) {
            int64_t r = a->rowval[k] - 1;
            int64_t row = transpose ? c : r, col = transpose ? r : c;
            int32_t p = pos[(size_t)row]++;
            idx[(size_t)p] = (int32_t)col;
            val[(size_t)p] = a->nzval[k];
        }
    out->nrow = nr;
    out->ncol = nc;
    if (!out->ptr.alloc(ptr.size()) || !out->idx.alloc(idx.size()) || !out->val.alloc(val.size())) {
        peec_set_error("operator_set: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    PEEC_HIP_TRY(hipMemcpy(out->ptr.p, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(out->idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(out->val.p, val.data(), val.size() * 8, hipMemcpyHostToDevice));
    return PEEC_OK;
}

// Parity of the grid positions a 0/1 expansion matrix selects along each
// half-cell axis of a kernel grid with dims nf: 0 or 1, -1 when mixed, -2
// when it selects nothing.
static void map_parity(const peec_csc_f64 *a, const int nf[3], const int half[3], int par[3])
{
    for (int d = 0; d < 3; d++) par[d] = -2;
    const int64_t nnz = a->colptr[a->ncol] - 1;
    for (int64_t k = 0; k < nnz; k++) {
        const int64_t r = a->rowval[k] - 1;
        const int64_t f[3] = {r % nf[0], (r / nf[0]) % nf[1], r / ((int64_t)nf[0] * nf[1])};
        for (int d = 0; d < 3; d++) {
            if (!half[d]) continue;
            const int p = (int)(f[d] & 1);
            if (par[d] == -2) par[d] = p;
            else if (par[d] != p) par[d] = -1;
        }
    }
}

// 0/1 expansion matrix (rows = compact kernel grid, one entry per column) ->
// position map; columns are local unknown indices [col0, col0 + ncols).
// With coarse (dims nf of the fine grid, half-cell axes, coarse box nb)
// every fine position is mapped to its coarse box position (f / 2 along the
// half-cell axes).
struct CoarseMap {
    int nf[3], half[3], nb[3];
};
static int map_from_csc(const peec_csc_f64 *a, int64_t grid, int64_t col0, int64_t ncols, DevBuf<int32_t> *out,
                        int64_t *nnz, const char *what, const CoarseMap *coarse = nullptr)
{
    if (a->nrow != grid) {
        peec_set_error("operator_set: %s has %lld rows, kernel grid has %lld", what, (long long)a->nrow,
                       (long long)grid);
        return PEEC_ERR_ARG;
    }
    const int64_t cells = coarse ? (int64_t)coarse->nb[0] * coarse->nb[1] * coarse->nb[2] : grid;
    std::vector<int32_t> map((size_t)std::max<int64_t>(cells, 1), -1);
    int64_t cnt = 0;
    for (int64_t c = 0; c < a->ncol; c++) {
        int64_t k0 = a->colptr[c] - 1, k1 = a->colptr[c + 1] - 1;
        if (k1 - k0 > 1) {
            peec_set_error("operator_set: %s column %lld has %lld entries (expected <= 1)", what, (long long)c + 1,
                           (long long)(k1 - k0));
            return PEEC_ERR_ARG;
        }
        for (int64_t k = k0; k < k1; k++) {
            int64_t r = a->rowval[k] - 1;
            if (coarse) {
                const CoarseMap &C = *coarse;
                int64_t f[3] = {r % C.nf[0], (r / C.nf[0]) % C.nf[1], r / ((int64_t)C.nf[0] * C.nf[1])};
                for (int d = 0; d < 3; d++)
                    if (C.half[d]) f[d] >>= 1;
                r = f[0] + (int64_t)C.nb[0] * (f[1] + (int64_t)C.nb[1] * f[2]);
            }
            if (a->nzval[k] != 1.0) {
                peec_set_error("operator_set: %s is not a 0/1 selection", what);
                return PEEC_ERR_ARG;
            }
            if (c < col0 || c >= col0 + ncols) {
                peec_set_error("operator_set: %s selects column %lld outside its section", what, (long long)c + 1);
                return PEEC_ERR_ARG;
            }
            if (map[(size_t)r] >= 0) {
                peec_set_error("operator_set: %s maps two unknowns to grid position %lld", what, (long long)r + 1);
                return PEEC_ERR_ARG;
            }
            map[(size_t)r] = (int32_t)(c - col0);
            cnt++;
        }
    }
    if (!out->alloc(map.size())) {
        peec_set_error("operator_set: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    PEEC_HIP_TRY(hipMemcpy(out->p, map.data(), map.size() * 4, hipMemcpyHostToDevice));
    *nnz = cnt;
    return PEEC_OK;
}

extern "C" int peec_operator_set(peec_ctx *ctx, int64_t mx, int64_t my, int64_t mz, const peec_csc_f64 *A,
                                 const peec_csc_f64 *Gamma, const peec_csc_f64 lp_map[3], const peec_csc_f64 p_map[9],
                                 const double *diag_P)
{
    if (!ctx || !A || !Gamma || !lp_map || !p_map || !diag_P) {
        peec_set_error("peec_operator_set: null argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->kernels_ready) {
        peec_set_error("peec_operator_set: call peec_kernels_build first");
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->release_operator();
    const int64_t m = A->nrow, n = A->ncol, ns = Gamma->ncol;
    if (mx + my + mz != m || Gamma->nrow != n) {
        peec_set_error("peec_operator_set: inconsistent sizes (m=%lld, mx+my+mz=%lld, Gamma %lldx%lld, n=%lld)",
                       (long long)m, (long long)(mx + my + mz), (long long)Gamma->nrow, (long long)ns, (long long)n);
        return PEEC_ERR_ARG;
    }
    if (m + ns + n > INT32_MAX) {
        peec_set_error("peec_operator_set: %lld unknowns exceed the int32 index range", (long long)(m + ns + n));
        return PEEC_ERR_ARG;
    }
    ctx->m = m;
    ctx->n = n;
    ctx->ns = ns;
    ctx->mx = mx;
    ctx->my = my;
    ctx->mz = mz;
    int rc;
    if ((rc = csc_to_csr(A, false, &ctx->A, "A")) || (rc = csc_to_csr(A, true, &ctx->At, "A'")) ||
        (rc = csc_to_csr(Gamma, false, &ctx->G, "Gamma")) || (rc = csc_to_csr(Gamma, true, &ctx->Gt, "Gamma'")))
        return rc;
    std::vector<double> ip((size_t)std::max<int64_t>(ns, 1), 0.0);
    for (int64_t s = 0; s < ns; s++) ip[(size_t)s] = 1.0 / diag_P[s];
    // host copies for the device LU (lu.hip)
    ctx->hA_ptr.assign((size_t)m + 1, 0);
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = A->colptr[c] - 1; k < A->colptr[c + 1] - 1; k++) ctx->hA_ptr[(size_t)A->rowval[k]]++;
    for (int64_t e = 0; e < m; e++) ctx->hA_ptr[(size_t)e + 1] += ctx->hA_ptr[(size_t)e];
    ctx->hA_idx.assign((size_t)ctx->hA_ptr[(size_t)m], 0);
    ctx->hA_val.assign((size_t)ctx->hA_ptr[(size_t)m], 0.0);
    {
        std::vector<int64_t> at(ctx->hA_ptr.begin(), ctx->hA_ptr.end() - 1);
        for (int64_t c = 0; c < n; c++)
            for (int64_t k = A->colptr[c] - 1; k < A->colptr[c + 1] - 1; k++) {
                const int64_t p = at[(size_t)(A->rowval[k] - 1)]++;
                ctx->hA_idx[(size_t)p] = (int32_t)c;
                ctx->hA_val[(size_t)p] = A->nzval[k];
            }
    }
    ctx->hG_ptr.assign((size_t)ns + 1, 0);
    ctx->hG_idx.clear();
    ctx->hG_val.clear();
    for (int64_t c = 0; c < ns; c++) {
        for (int64_t k = Gamma->colptr[c] - 1; k < Gamma->colptr[c + 1] - 1; k++) {
            ctx->hG_idx.push_back((int32_t)(Gamma->rowval[k] - 1));
            ctx->hG_val.push_back(Gamma->nzval[k]);
        }
        ctx->hG_ptr[(size_t)c + 1] = (int64_t)ctx->hG_idx.size();
    }
    ctx->h_invP.assign(ip.begin(), ip.begin() + ns);
    ctx->lu.ready = false;
    if (!ctx->invP.alloc(ip.size())) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpy(ctx->invP.p, ip.data(), ip.size() * 8, hipMemcpyHostToDevice));

    // Lp applications (compute_Matrix_vector_fft.jl:14-28)
    const int64_t moff[3] = {0, mx, mx + my}, msz[3] = {mx, my, mz};
    int na = 0;
    int64_t wmax = 1;
    for (int d = 0; d < 3; d++) {
        if (msz[d] == 0) continue;
        ConvApp &a = ctx->apps[na];
        a.kernel = d;
        a.in_off = moff[d];
        a.out_off = moff[d];
        const char *nm[3] = {"mat_map_Lp[1,1]", "mat_map_Lp[2,1]", "mat_map_Lp[3,1]"};
        if (lp_map[d].ncol != msz[d]) {
            peec_set_error("peec_operator_set: %s has %lld columns, expected %lld", nm[d], (long long)lp_map[d].ncol,
                           (long long)msz[d]);
            return PEEC_ERR_ARG;
        }
        if ((rc = map_from_csc(&lp_map[d], ctx->kern[d].compact(), 0, msz[d], &a.map_in, &a.nnz_in, nm[d]))) return rc;
        if ((rc = map_from_csc(&lp_map[d], ctx->kern[d].compact(), 0, msz[d], &a.map_out, &a.nnz_out, nm[d]))) return rc;
        for (int q = 0; q < 3; q++) {
            a.n[q] = ctx->kern[d].n[q];
            a.L[q] = ctx->kern[d].L[q];
        }
        a.poly = a.conj = false;
        wmax = std::max(wmax, conv_work_elems(a));
        na++;
    }
    // P applications (:35-65): pairs (c1, c2), c2 >= c1; off-diagonal pairs twice
    const int pk[3][3] = {{3, 6, 7}, {6, 4, 8}, {7, 8, 5}};
    for (int c1 = 0; c1 < 3; c1++)
        for (int c2 = c1; c2 < 3; c2++) {
            int k = pk[c1][c2];
            for (int pass = 0; pass < (c1 == c2 ? 1 : 2); pass++) {
                int i_in = pass == 0 ? c1 : c2, j_in = pass == 0 ? c2 : c1;
                const peec_csc_f64 *min_ = &p_map[3 * j_in + i_in];  // exp_P[i_in, j_in]
                const peec_csc_f64 *mout = &p_map[3 * i_in + j_in];  // exp_P[j_in, i_in]
                ConvApp &a = ctx->apps[na];
                Kernel &K = ctx->kern[k];
                a.kernel = k;
                a.in_off = m;
                a.out_off = m;
                for (int q = 0; q < 3; q++) {
                    a.n[q] = K.n[q];
                    a.L[q] = K.L[q];
                }
                a.poly = a.conj = false;
                // Half-cell reduction of an off-diagonal block: when every
                // scattered and every gathered position keeps one parity
                // per half-cell axis, convolve on the coarse box
                if (c1 != c2 && min_->nrow == K.compact() && mout->nrow == K.compact()) {
                    int pin[3], pout[3], delta[3] = {0, 0, 0};
                    map_parity(min_, K.n, kHalfAxes[k], pin);
                    map_parity(mout, K.n, kHalfAxes[k], pout);
                    bool ok = true;
                    for (int q = 0; q < 3; q++)
                        if (kHalfAxes[k][q]) {
                            if (pin[q] < 0 || pout[q] < 0) ok = false;
                            else delta[q] = pout[q] - pin[q];
                        }
                    if (ok && !K.poly && (rc = poly_spectrum_build(ctx, k, delta))) return rc;
                    bool same = ok, flip = ok;
                    for (int q = 0; q < 3 && ok; q++) {
                        same = same && delta[q] == K.delta[q];
                        flip = flip && delta[q] == -K.delta[q];
                    }
                    if (ok && (same || flip)) {
                        a.poly = true;
                        a.conj = !same;
                        for (int q = 0; q < 3; q++) {
                            a.n[q] = K.pn[q];
                            a.L[q] = K.pL[q];
                        }
                    }
                }
                CoarseMap cm;
                for (int q = 0; q < 3; q++) {
                    cm.nf[q] = K.n[q];
                    cm.half[q] = kHalfAxes[k][q];
                    cm.nb[q] = K.pn[q];
                }
                const CoarseMap *cmp = a.poly ? &cm : nullptr;
                if ((rc = map_from_csc(min_, K.compact(), 0, ns, &a.map_in, &a.nnz_in, "exp_P (scatter)", cmp)))
                    return rc;
                if ((rc = map_from_csc(mout, K.compact(), 0, ns, &a.map_out, &a.nnz_out, "exp_P (gather)", cmp)))
                    return rc;
                wmax = std::max(wmax, conv_work_elems(a));
                na++;
            }
        }
    ctx->napps = na;
    // work buffers: two grids (one rhs) plus vectors
    if (!ctx->w1.alloc((size_t)wmax) || !ctx->w2.alloc((size_t)wmax)) {
        peec_set_error("peec_operator_set: out of device memory for FFT work grids");
        return PEEC_ERR_NOMEM;
    }
    const int64_t N = m + ns + n;
    if (!ctx->vx.alloc((size_t)N) || !ctx->vy.alloc((size_t)N) || !ctx->vt1.alloc((size_t)std::max<int64_t>(n, 1)) ||
        !ctx->vt2.alloc((size_t)std::max<int64_t>(n, 1))) {
        peec_set_error("peec_operator_set: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    if ((rc = yz_prepare(ctx))) return rc;
    ctx->op_ready = true;
    ctx->rcc_ready = false;  // the half-cell reductions may have changed: rebuild retarded spectra
    return PEEC_OK;
}

// All slots' Yle as consecutive CSRs (ptr offsets absolute).
int peec_ctx::upload_yle()
{
    std::vector<int32_t> ptr, idx;
    std::vector<double2_t> val;
    for (int sl = 0; sl < nslots; sl++) {
        const HostCsr &h = hYle[(size_t)sl];
        const int32_t base = (int32_t)idx.size();
        for (int64_t i = 0; i <= n; i++) ptr.push_back(base + (h.ptr.empty() ? 0 : h.ptr[(size_t)i]));
        if (!h.ptr.empty()) {
            idx.insert(idx.end(), h.idx.begin(), h.idx.begin() + h.ptr[(size_t)n]);
            val.insert(val.end(), h.val.begin(), h.val.begin() + h.ptr[(size_t)n]);
        }
    }
    if (idx.empty()) {
        idx.push_back(0);
        val.push_back({0.0, 0.0});
    }
    Yle.nrow = Yle.ncol = n;
    if (!Yle.ptr.alloc(ptr.size()) || !Yle.idx.alloc(idx.size()) || !Yle.val.alloc(val.size())) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpy(Yle.ptr.p, ptr.data(), ptr.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(Yle.idx.p, idx.data(), idx.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(Yle.val.p, val.data(), val.size() * 16, hipMemcpyHostToDevice));
    return PEEC_OK;
}

extern "C" int peec_frequency_slots(peec_ctx *ctx, int nslots)
{
    if (!ctx || nslots < 1) {
        peec_set_error("peec_frequency_slots: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready) {
        peec_set_error("peec_frequency_slots: call peec_operator_set first");
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int64_t m = std::max<int64_t>(ctx->m, 1);
    if (nslots > 1 && !ctx->lu_layout_from_plan) {
        peec_set_error("peec_frequency_slots: several slots need the device factorization (peec_lu_analyze)");
        return PEEC_ERR_STATE;
    }
    if (!ctx->wv.alloc((size_t)nslots) || !ctx->DZ.alloc((size_t)(m * nslots)) ||
        !ctx->invZ.alloc((size_t)(m * nslots))) {
        peec_set_error("peec_frequency_slots: out of device memory for %d slots", nslots);
        return PEEC_ERR_NOMEM;
    }
    if (nslots > 1 && ctx->rcc) {
        peec_set_error("peec_frequency_slots: the retarded kernels have one set of spectra (one frequency)");
        return PEEC_ERR_STATE;
    }
    ctx->nslots = nslots;
    ctx->hw.assign((size_t)nslots, 0.0);
    ctx->hYle.assign((size_t)nslots, peec_ctx::HostCsr());
    ctx->slot_set.assign((size_t)nslots, 0);
    ctx->slot_lu.assign((size_t)nslots, 0);
    ctx->mf_ready = false;
    ctx->freq_ready = false;
    ctx->lu_ready = false;
    int rc;
    if (ctx->lu_layout_from_plan) {
        if ((rc = tri_set_slots(&ctx->Lf, nslots)) || (rc = tri_set_slots(&ctx->Uf, nslots))) return rc;
        // L is unit lower triangular: 1 / L[i, i] = 1 in every slot
        std::vector<double2_t> ones((size_t)(ctx->n * nslots), double2_t{1.0, 0.0});
        PEEC_HIP_TRY(hipMemcpy(ctx->Lf.diag.p, ones.data(), ones.size() * 16, hipMemcpyHostToDevice));
    }
    return PEEC_OK;
}

static int slot_set_rest(peec_ctx *ctx, int slot, double w, const peec_csc_c128 *Yle);

extern "C" int peec_frequency_set_slot(peec_ctx *ctx, int slot, double w, const peec_c128 *DZ,
                                       const peec_csc_c128 *Yle)
{
    if (!ctx || !DZ || !Yle) {
        peec_set_error("peec_frequency_set: null argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready) {
        peec_set_error("peec_frequency_set: call peec_operator_set first");
        return PEEC_ERR_STATE;
    }
    if (slot < 0 || slot >= ctx->nslots || (int)ctx->slot_set.size() != ctx->nslots) {
        peec_set_error("peec_frequency_set: slot %d outside the %d slots of peec_frequency_slots", slot, ctx->nslots);
        return PEEC_ERR_ARG;
    }
    if (Yle->nrow != ctx->n || Yle->ncol != ctx->n) {
        peec_set_error("peec_frequency_set: Yle must be n x n");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    const int64_t m = ctx->m;
    std::vector<double2_t> dz((size_t)m), iz((size_t)m);
    for (int64_t e = 0; e < m; e++) {
        dz[(size_t)e] = {DZ[e].re, DZ[e].im};
        double d = DZ[e].re * DZ[e].re + DZ[e].im * DZ[e].im;
        if (d == 0.0) {
            peec_set_error("peec_frequency_set: DZ[%lld] is zero", (long long)e + 1);
            return PEEC_ERR_NUMERIC;
        }
        // 1 ./ DZ (FFT_solver_QS_S_type.jl:115)
        iz[(size_t)e] = {DZ[e].re / d, -DZ[e].im / d};
    }
    PEEC_HIP_TRY(hipMemcpy(ctx->DZ.p + (int64_t)slot * m, dz.data(), (size_t)m * 16, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->invZ.p + (int64_t)slot * m, iz.data(), (size_t)m * 16, hipMemcpyHostToDevice));
    return slot_set_rest(ctx, slot, w, Yle);
}

// The rest of setting a slot once its DZ and 1/DZ are on the device: w and
// the admittance matrix Yle.
static int slot_set_rest(peec_ctx *ctx, int slot, double w, const peec_csc_c128 *Yle)
{
    const int64_t n = ctx->n;
    ctx->hw[(size_t)slot] = w;
    if (slot == 0) ctx->w = w;
    PEEC_HIP_TRY(hipMemcpy(ctx->wv.p, ctx->hw.data(), ctx->hw.size() * 8, hipMemcpyHostToDevice));
    // Yle CSC -> CSR
    int64_t nnz = Yle->colptr[n] - 1;
    std::vector<int32_t> ptr((size_t)n + 1, 0), idx((size_t)std::max<int64_t>(nnz, 1));
    std::vector<double2_t> val((size_t)std::max<int64_t>(nnz, 1));
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = Yle->colptr[c] - 1; k < Yle->colptr[c + 1] - 1; k++) ptr[(size_t)Yle->rowval[k]]++;
    for (int64_t i = 0; i < n; i++) ptr[(size_t)i + 1] += ptr[(size_t)i];
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = Yle->colptr[c] - 1; k < Yle->colptr[c + 1] - 1; k++) {
            int64_t r = Yle->rowval[k] - 1;
            int32_t p = pos[(size_t)r]++;
            idx[(size_t)p] = (int32_t)c;
            val[(size_t)p] = {Yle->nzval[k].re, Yle->nzval[k].im};
        }
    peec_ctx::HostCsr &h = ctx->hYle[(size_t)slot];
    h.ptr = std::move(ptr);
    h.idx = std::move(idx);
    h.val = std::move(val);
    int rc = ctx->upload_yle();
    if (rc) return rc;
    ctx->slot_set[(size_t)slot] = 1;
    ctx->slot_lu[(size_t)slot] = 0;
    ctx->freq_ready = std::all_of(ctx->slot_set.begin(), ctx->slot_set.end(), [](char c) { return c != 0; });
    ctx->lu_ready = false;
    return PEEC_OK;
}

// Edge impedances at one frequency on the device (FFT_solver_QS_S_type.jl
// :106-115): Z_self from the four half-edge R/Cd pairs (compute_Z_self.jl:
// 1-15: R || Cd, R alone or Cd alone, summed in order; Cd is complex for
// lossy dielectrics, j w Cd = (-w Im Cd, w Re Cd)), the surface
// impedance Zs = esc_R * (Zs0 * sqrt(w / esc_freq)), the skin switch on
// surface edges (the larger real part wins; the loser keeps only its
// imaginary part), then DZ = Z_self + Re(Zs) + j w Lp and 1 / DZ.  The
// complex reciprocals follow Smith's method, the operation order the host
// restatement uses, so results match it to the last bit or so.
__global__ void k_edge_dz(const double *__restrict__ R, const double2_t *__restrict__ Cd,
                          const double2_t *__restrict__ Zs0, const double *__restrict__ Lp,
                          const int8_t *__restrict__ surf, int64_t m, double w, double esc_R, double esc_freq,
                          double2_t *__restrict__ dz, double2_t *__restrict__ iz)
{
#pragma clang fp contract(off)
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m) return;
    // 1 / (a + j b) by Smith's method
    auto recip = [](double a, double b, double &x, double &y) {
        if (fabs(a) >= fabs(b)) {
            const double rat = b / a, scl = 1.0 / (a + b * rat);
            x = (1.0 + 0.0 * rat) * scl;
            y = (0.0 - 1.0 * rat) * scl;
        } else {
            const double rat = a / b, scl = 1.0 / (b + a * rat);
            x = (1.0 * rat + 0.0) * scl;
            y = (0.0 * rat - 1.0) * scl;
        }
    };
    double zr = 0.0, zi = 0.0;
    for (int a = 0; a < 4; a++) {
        const double r = R[4 * e + a];
        const double2_t c = Cd[4 * e + a];
        const bool hc = c.x != 0.0 || c.y != 0.0;
        // j w c, as (0 + j w) * c
        const double jr = 0.0 * c.x - w * c.y, ji = 0.0 * c.y + w * c.x;
        double tr, ti;
        if (r != 0.0 && hc) {
            recip(1.0 / r + jr, 0.0 + ji, tr, ti);
        } else if (r != 0.0) {
            tr = r;
            ti = 0.0;
        } else if (hc) {
            recip(jr, ji, tr, ti);
        } else {
            continue;
        }
        zr = zr + tr;
        zi = zi + ti;
    }
    const double s = sqrt(w / esc_freq);
    const double2_t z0 = Zs0[e];
    double sr = esc_R * (z0.x * s), si = esc_R * (z0.y * s);
    if (surf[e]) {
        const double d = sr - zr;
        if (d > 0.0) zr = 0.0;
        else if (d < 0.0) sr = 0.0;
    }
    (void)si;
    const double dr = zr + sr + 0.0 * Lp[e], di = zi + w * Lp[e];
    dz[e] = {dr, di};
    const double q = dr * dr + di * di;
    iz[e] = {dr / q, -di / q};
}

extern "C" int peec_edge_terms_set(peec_ctx *ctx, int64_t m, const double *R, const peec_c128 *Cd, const peec_c128 *Zs0,
                                   const double *Lp, const int8_t *surface, double esc_R, double esc_freq)
{
    if (!ctx || !R || !Cd || !Zs0 || !Lp || !surface || !(esc_freq > 0)) {
        peec_set_error("peec_edge_terms_set: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || m != ctx->m) {
        peec_set_error("peec_edge_terms_set: call peec_operator_set first; m must be %lld", (long long)ctx->m);
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    if (!ctx->eR.alloc((size_t)(4 * m)) || !ctx->eCd.alloc((size_t)(4 * m)) || !ctx->eLp.alloc((size_t)m) ||
        !ctx->eZs.alloc((size_t)m) || !ctx->eSurf.alloc((size_t)m))
        return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpy(ctx->eR.p, R, (size_t)(4 * m) * 8, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->eCd.p, Cd, (size_t)(4 * m) * 16, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->eLp.p, Lp, (size_t)m * 8, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->eZs.p, Zs0, (size_t)m * 16, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->eSurf.p, surface, (size_t)m, hipMemcpyHostToDevice));
    ctx->esc_R = esc_R;
    ctx->esc_freq = esc_freq;
    ctx->edges_ready = true;
    return PEEC_OK;
}

extern "C" int peec_frequency_set_slot_edges(peec_ctx *ctx, int slot, double w, const peec_csc_c128 *Yle)
{
    if (!ctx || !Yle || !(w > 0)) {
        peec_set_error("peec_frequency_set_slot_edges: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->edges_ready) {
        peec_set_error("peec_frequency_set_slot_edges: call peec_operator_set and peec_edge_terms_set first");
        return PEEC_ERR_STATE;
    }
    if (slot < 0 || slot >= ctx->nslots || (int)ctx->slot_set.size() != ctx->nslots) {
        peec_set_error("peec_frequency_set_slot_edges: slot %d outside the %d slots", slot, ctx->nslots);
        return PEEC_ERR_ARG;
    }
    if (Yle->nrow != ctx->n || Yle->ncol != ctx->n) {
        peec_set_error("peec_frequency_set_slot_edges: Yle must be n x n");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t m = ctx->m;
    hipLaunchKernelGGL(k_edge_dz, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, ctx->stream, ctx->eR.p, ctx->eCd.p,
                       ctx->eZs.p, ctx->eLp.p, ctx->eSurf.p, m, w, ctx->esc_R, ctx->esc_freq,
                       ctx->DZ.p + (int64_t)slot * m, ctx->invZ.p + (int64_t)slot * m);
    PEEC_HIP_TRY(hipGetLastError());
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return slot_set_rest(ctx, slot, w, Yle);
}

extern "C" int peec_frequency_dz(peec_ctx *ctx, int slot, peec_c128 *out)
{
    if (!ctx || !out || slot < 0 || slot >= ctx->nslots || (int)ctx->slot_set.size() != ctx->nslots ||
        !ctx->slot_set[(size_t)slot]) {
        peec_set_error("peec_frequency_dz: slot %d is not set", slot);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    PEEC_HIP_TRY(hipMemcpy(out, ctx->DZ.p + (int64_t)slot * ctx->m, (size_t)ctx->m * 16, hipMemcpyDeviceToHost));
    return PEEC_OK;
}

extern "C" int peec_frequency_set(peec_ctx *ctx, double w, const peec_c128 *DZ, const peec_csc_c128 *Yle)
{
    if (!ctx) {
        peec_set_error("peec_frequency_set: null context");
        return PEEC_ERR_ARG;
    }
    if (ctx->nslots != 1 || (int)ctx->slot_set.size() != 1) {
        int rc = peec_frequency_slots(ctx, 1);
        if (rc) return rc;
    }
    return peec_frequency_set_slot(ctx, 0, w, DZ, Yle);
}

extern "C" int peec_lu_set_factors(peec_ctx *ctx, int64_t n, const peec_csc_c128 *L, const peec_csc_c128 *U,
                                   const int64_t *p, const int64_t *q, const double *Rs)
{
    return peec_lu_set_factors_blocked(ctx, n, L, U, p, q, Rs, 0, nullptr);
}

extern "C" int peec_lu_set_factors_blocked(peec_ctx *ctx, int64_t n, const peec_csc_c128 *L, const peec_csc_c128 *U,
                                           const int64_t *p, const int64_t *q, const double *Rs, int64_t nblocks,
                                           const int64_t *block_ptr)
{
    if (!ctx || !L || !U || !p || !q) {
        peec_set_error("peec_lu_set_factors: null argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->freq_ready) {
        peec_set_error("peec_lu_set_factors: call peec_frequency_set first");
        return PEEC_ERR_STATE;
    }
    if (n != ctx->n || L->ncol != n || U->ncol != n) {
        peec_set_error("peec_lu_set_factors: factors must be n x n with n = %lld", (long long)ctx->n);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (n > INT32_MAX) {
        peec_set_error("peec_lu_set_factors: n exceeds the int32 index range");
        return PEEC_ERR_ARG;
    }
    if (ctx->nslots != 1) {
        peec_set_error("peec_lu_set_factors: host factors serve one frequency slot (%d set)", ctx->nslots);
        return PEEC_ERR_STATE;
    }
    std::vector<int32_t> blk;
    if (nblocks > 0 && block_ptr) {
        if (block_ptr[0] != 0 || block_ptr[nblocks] != n) {
            peec_set_error("peec_lu_set_factors: block_ptr must run from 0 to n");
            return PEEC_ERR_ARG;
        }
        blk.resize((size_t)nblocks + 1);
        for (int64_t B = 0; B <= nblocks; B++) {
            if (B > 0 && block_ptr[B] <= block_ptr[B - 1]) {
                peec_set_error("peec_lu_set_factors: empty or decreasing block");
                return PEEC_ERR_ARG;
            }
            blk[(size_t)B] = (int32_t)block_ptr[B];
        }
    } else {
        blk = tri_chain_blocks(L, kTrsvMaxBlock);
    }
    int rc;
    if ((rc = tri_build(L, true, blk, &ctx->Lf, "L")) || (rc = tri_build(U, false, blk, &ctx->Uf, "U"))) return rc;
    // pinv[p[i]] = i (scatter of the permuted rhs), qinv[q[j]] = j
    std::vector<int32_t> pinv((size_t)n, -1), qinv((size_t)n, -1);
    for (int64_t i = 0; i < n; i++) {
        if (p[i] < 0 || p[i] >= n || pinv[(size_t)p[i]] >= 0 || q[i] < 0 || q[i] >= n || qinv[(size_t)q[i]] >= 0) {
            peec_set_error("peec_lu_set_factors: p and q must be 0-based permutations");
            return PEEC_ERR_ARG;
        }
        pinv[(size_t)p[i]] = (int32_t)i;
        qinv[(size_t)q[i]] = (int32_t)i;
    }
    if (!ctx->perm_r.alloc((size_t)n) || !ctx->perm_c.alloc((size_t)n)) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpy(ctx->perm_r.p, pinv.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->perm_c.p, qinv.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    if (Rs) {
        if (!ctx->Rs.alloc((size_t)n)) return PEEC_ERR_NOMEM;
        PEEC_HIP_TRY(hipMemcpy(ctx->Rs.p, Rs, (size_t)n * 8, hipMemcpyHostToDevice));
        ctx->has_Rs = true;
    } else {
        ctx->has_Rs = false;
    }
    ctx->slot_lu.assign(1, 1);
    ctx->lu_ready = true;
    ctx->lu_layout_from_plan = false;
    ctx->mf_ready = false;
    return PEEC_OK;
}

// ------------------------------------------------------------------ apply
// In place on y (length N per rhs, stride ld): y <- M^-1 y
int apply_precond(peec_ctx *ctx, double2_t *y, int nrhs, int64_t ld, const ColMap &cm)
{
    if (nrhs <= 0) return PEEC_OK;
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n;
    if (nrhs * n > (int64_t)ctx->vt1.n) {
        if (!ctx->vt1.alloc((size_t)(nrhs * n)) || !ctx->vt2.alloc((size_t)(nrhs * n))) return PEEC_ERR_NOMEM;
    }
    double2_t *b = ctx->vt1.p, *z = ctx->vt2.p;
    hipLaunchKernelGGL(k_pre, dim3(nblk(n), (unsigned)cm.nsl), dim3(256), 0, ctx->stream, y, ctx->invZ.p, ctx->invP.p, ctx->At.ptr.p,
                       ctx->At.idx.p, ctx->At.val.p, ctx->G.ptr.p, ctx->G.idx.p, ctx->G.val.p, ctx->perm_r.p,
                       ctx->has_Rs ? ctx->Rs.p : nullptr, b, ctx->wv.p, m, ns, n, nrhs, ld, cm);
    PEEC_HIP_TRY(hipGetLastError());
    int rc;
    if (ctx->mf_ready && ctx->lu_solve == 1 && ctx->lu_layout_from_plan) {
        // multifrontal, from the device factorization's fronts
        if ((rc = mf_solve(ctx, true, b, z, nrhs, cm))) return rc;
        if ((rc = mf_solve(ctx, false, z, b, nrhs, cm))) return rc;
    } else {
        if ((rc = trsv_solve(ctx, ctx->Lf, true, b, z, nrhs, cm))) return rc;   // L c = b   (c in z)
        if ((rc = trsv_solve(ctx, ctx->Uf, false, z, b, nrhs, cm))) return rc;  // U z' = c  (z' in b)
    }
    hipLaunchKernelGGL(k_post, dim3(nblk(m + ns + n), (unsigned)cm.nsl), dim3(256), 0, ctx->stream, y, b, ctx->perm_c.p, ctx->invZ.p,
                       ctx->invP.p, ctx->A.ptr.p, ctx->A.idx.p, ctx->A.val.p, ctx->Gt.ptr.p, ctx->Gt.idx.p,
                       ctx->Gt.val.p, m, ns, n, nrhs, ld, cm);
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// y = K x (before the preconditioner), device vectors
// complex128 elements of each FFT work grid that nrhs right-hand sides need
static size_t work_grid_elems(const peec_ctx *ctx, int nrhs)
{
    size_t need = 0;
    for (int a = 0; a < ctx->napps; a++) need = std::max<size_t>(need, (size_t)conv_work_elems(ctx->apps[a]));
    return need * (size_t)nrhs;
}

// K x with the two FFT work grids w1, w2 given (work_grid_elems each);
// launches only, on the context's stream
static int apply_K_work(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32,
                        const ColMap &cm, double2_t *w1, double2_t *w2);

int apply_K(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32, const ColMap &cm)
{
    if (nrhs <= 0) return PEEC_OK;
    const size_t need = work_grid_elems(ctx, nrhs);
    if (need > ctx->w1.n) {
        if (!ctx->w1.alloc(need) || !ctx->w2.alloc(need)) {
            peec_set_error("out of device memory for %d-rhs FFT work grids", nrhs);
            return PEEC_ERR_NOMEM;
        }
    }
    return apply_K_work(ctx, x, y, nrhs, ld, fp32, cm, ctx->w1.p, ctx->w2.p);
}

static int apply_K_work(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32,
                        const ColMap &cm, double2_t *w1, double2_t *w2)
{
    if (nrhs <= 0) return PEEC_OK;
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n, N = m + ns + n;
    hipLaunchKernelGGL(k_zero, dim3(nblk(m + ns), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, m + ns, nrhs, ld);
    PEEC_HIP_TRY(hipGetLastError());
    for (int a = 0; a < ctx->napps; a++) {
        int rc = conv_apply(ctx, ctx->apps[a], fp32, x, y, nrhs, ld, w1, w2);
        if (rc) return rc;
    }
    const int64_t phi0 = m + ns;
    hipLaunchKernelGGL(k_y1, dim3(nblk(m), (unsigned)cm.nsl), dim3(256), 0, ctx->stream, y, x, ctx->DZ.p, ctx->A.ptr.p, ctx->A.idx.p,
                       ctx->A.val.p, ctx->wv.p, m, phi0, nrhs, ld, cm);
    hipLaunchKernelGGL(k_y2, dim3(nblk(ns), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, x, ctx->Gt.ptr.p, ctx->Gt.idx.p,
                       ctx->Gt.val.p, m, ns, phi0, nrhs, ld);
    hipLaunchKernelGGL(k_y3, dim3(nblk(n), (unsigned)cm.nsl), dim3(256), 0, ctx->stream, y, x, ctx->At.ptr.p, ctx->At.idx.p,
                       ctx->At.val.p, ctx->Yle.ptr.p, ctx->Yle.idx.p, ctx->Yle.val.p, ctx->G.ptr.p, ctx->G.idx.p,
                       ctx->G.val.p, ctx->wv.p, m, ns, n, nrhs, ld, cm);
    PEEC_HIP_TRY(hipGetLastError());
    (void)N;
    return PEEC_OK;
}

int sync_stream(peec_ctx *ctx)
{
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

// device-vector matvec used by GMRES: y = M^-1 K x
int matvec_device(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32, const ColMap &cm)
{
    int rc = apply_K(ctx, x, y, nrhs, ld, fp32, cm);
    if (rc) return rc;
    return apply_precond(ctx, y, nrhs, ld, cm);
}

// y = K^H x (device vectors)
int apply_KH(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32, const ColMap &cm)
{
    if (nrhs <= 0) return PEEC_OK;
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n;
    hipLaunchKernelGGL(k_zero, dim3(nblk(m + ns), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, m + ns, nrhs, ld);
    PEEC_HIP_TRY(hipGetLastError());
    size_t need = 0;
    for (int a = 0; a < ctx->napps; a++) need = std::max<size_t>(need, (size_t)conv_work_elems(ctx->apps[a]));
    need *= (size_t)nrhs;
    if (need > ctx->w1.n && (!ctx->w1.alloc(need) || !ctx->w2.alloc(need))) {
        peec_set_error("out of device memory for %d-rhs FFT work grids", nrhs);
        return PEEC_ERR_NOMEM;
    }
    for (int a = 0; a < ctx->napps; a++) {
        int rc = conv_apply(ctx, ctx->apps[a], fp32, x, y, nrhs, ld, ctx->w1.p, ctx->w2.p);
        if (rc) return rc;
    }
    const int64_t phi0 = m + ns;
    hipLaunchKernelGGL(k_y1h, dim3(nblk(m), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, x, ctx->DZ.p, ctx->A.ptr.p, ctx->A.idx.p,
                       ctx->A.val.p, ctx->wv.p, m, phi0, nrhs, ld, cm);
    hipLaunchKernelGGL(k_y2h, dim3(nblk(ns), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, x, ctx->Gt.ptr.p, ctx->Gt.idx.p,
                       ctx->Gt.val.p, ctx->wv.p, m, ns, phi0, nrhs, ld, cm);
    hipLaunchKernelGGL(k_y3h, dim3(nblk(n), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, x, ctx->At.ptr.p, ctx->At.idx.p,
                       ctx->At.val.p, ctx->Yle.ptr.p, ctx->Yle.idx.p, ctx->Yle.val.p, ctx->G.ptr.p, ctx->G.idx.p,
                       ctx->G.val.p, m, ns, n, nrhs, ld, cm);
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// In place: y <- M^-H y, through the forward factors of SS
int apply_precond_adj(peec_ctx *ctx, double2_t *y, int nrhs, int64_t ld, const ColMap &cm)
{
    if (nrhs <= 0) return PEEC_OK;
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n;
    hipLaunchKernelGGL(k_adj_xform, dim3(nblk(m + ns + n), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, ctx->wv.p, m, ns, n, nrhs,
                       ld, cm, 0);
    PEEC_HIP_TRY(hipGetLastError());
    int rc = apply_precond(ctx, y, nrhs, ld, cm);
    if (rc) return rc;
    hipLaunchKernelGGL(k_adj_xform, dim3(nblk(m + ns + n), (unsigned)nrhs), dim3(256), 0, ctx->stream, y, ctx->wv.p, m, ns, n, nrhs,
                       ld, cm, 1);
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// adjoint GMRES operator: y = M^-H K^H x
int matvec_device_adj(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32,
                      const ColMap &cm)
{
    int rc = apply_KH(ctx, x, y, nrhs, ld, fp32, cm);
    if (rc) return rc;
    return apply_precond_adj(ctx, y, nrhs, ld, cm);
}

// Allocate every work buffer a matvec of nrhs right-hand sides uses, so
// that a captured matvec allocates nothing.
static int matvec_prepare(peec_ctx *ctx, int nrhs)
{
    size_t need = 0;
    for (int a = 0; a < ctx->napps; a++) need = std::max<size_t>(need, (size_t)conv_work_elems(ctx->apps[a]));
    need *= (size_t)nrhs;
    if (need > ctx->w1.n && (!ctx->w1.alloc(need) || !ctx->w2.alloc(need))) {
        peec_set_error("out of device memory for %d-rhs FFT work grids", nrhs);
        return PEEC_ERR_NOMEM;
    }
    if ((size_t)(nrhs * ctx->n) > ctx->vt1.n &&
        (!ctx->vt1.alloc((size_t)(nrhs * ctx->n)) || !ctx->vt2.alloc((size_t)(nrhs * ctx->n)))) {
        peec_set_error("out of device memory for %d-rhs triangular-solve vectors", nrhs);
        return PEEC_ERR_NOMEM;
    }
    const size_t part = (size_t)std::max<int64_t>(std::max<int64_t>(ctx->Lf.nseg, ctx->Uf.nseg), 1) * (size_t)nrhs;
    if (part > ctx->trsv_part.n && !ctx->trsv_part.alloc(part)) {
        peec_set_error("out of device memory for triangular-solve partial sums");
        return PEEC_ERR_NOMEM;
    }
    if (ctx->mf_ready) return mf_prepare(ctx, nrhs);
    return PEEC_OK;
}

MatvecGraphs::~MatvecGraphs() { clear(); }

void MatvecGraphs::clear()
{
    for (auto &e : entries) (void)hipGraphExecDestroy(e.exec);
    entries.clear();
}

int MatvecGraphs::run(const double2_t *x, double2_t *y, bool fp32)
{
    return run_cols(x, y, fp32, nrhs, uniform_cols(nrhs, rps));
}

int MatvecGraphs::run_cols(const double2_t *x, double2_t *y, bool fp32, int nrhs_c, const ColMap &cm)
{
    auto direct = [&]() {
        return adjoint ? matvec_device_adj(ctx, x, y, nrhs_c, ld, fp32, cm)
                       : matvec_device(ctx, x, y, nrhs_c, ld, fp32, cm);
    };
    if (!ctx->use_graphs || ctx->prof_on) return direct();
    for (auto &e : entries)
        if (e.x == x && e.y == y && e.fp32 == (int)fp32 && e.nrhs == nrhs_c && e.rslot == cm.rslot) {
            PEEC_HIP_TRY(hipGraphLaunch(e.exec, ctx->stream));
            return PEEC_OK;
        }
    int rc = matvec_prepare(ctx, nrhs_c);
    if (rc) return rc;
    PEEC_HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    rc = direct();
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ec != hipSuccess) {
        peec_set_error("matvec graph capture failed: %s", hipGetErrorString(ec));
        return PEEC_ERR_HIP;
    }
    Entry e{x, y, (int)fp32, nrhs_c, cm.rslot, nullptr};
    const hipError_t ei = hipGraphInstantiate(&e.exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        peec_set_error("matvec graph instantiation failed: %s", hipGetErrorString(ei));
        return PEEC_ERR_HIP;
    }
    entries.push_back(e);
    PEEC_HIP_TRY(hipGraphLaunch(e.exec, ctx->stream));
    return PEEC_OK;
}

// dst = src, one 16-byte element per thread (a known-size dispatch)
__global__ void k_copy_reference(double2_t *__restrict__ dst, const double2_t *__restrict__ src, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" int peec_copy_reference(peec_ctx *ctx, int64_t bytes)
{
    if (!ctx || bytes < 16) {
        peec_set_error("peec_copy_reference: invalid argument");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t n = bytes / 16;
    DevBuf<double2_t> a, b;
    if (!a.alloc((size_t)n) || !b.alloc((size_t)n)) {
        peec_set_error("peec_copy_reference: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_zero, dim3(nblk(n), 1), dim3(256), 0, ctx->stream, a.p, n, 1, n);
    for (int i = 0; i < 3; i++)
        hipLaunchKernelGGL(k_copy_reference, dim3(nblk(n)), dim3(256), 0, ctx->stream, b.p, a.p, n);
    PEEC_HIP_TRY(hipGetLastError());
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

extern "C" int peec_synchronize(peec_ctx *ctx)
{
    if (!ctx) {
        peec_set_error("peec_synchronize: null context");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    PEEC_HIP_TRY(hipDeviceSynchronize());
    return PEEC_OK;
}

extern "C" int peec_set_graphs(peec_ctx *ctx, int on)
{
    if (!ctx) {
        peec_set_error("peec_set_graphs: null context");
        return PEEC_ERR_ARG;
    }
    ctx->use_graphs = on != 0;
    return PEEC_OK;
}

extern "C" int peec_lu_set_solve(peec_ctx *ctx, int mode)
{
    if (!ctx || mode < 0 || mode > 1) {
        peec_set_error("peec_lu_set_solve: mode must be 0 (level-scheduled blocks) or 1 (multifrontal)");
        return PEEC_ERR_ARG;
    }
    ctx->lu_solve = mode;
    // the multifrontal solve needs the pivot-block inverses of the current
    // factorization: mode 1 takes effect with the next one that covers
    // every slot, mode 0 at once (with the layouts' values gathered from the
    // pool if the factorization left them to the multifrontal solve)
    ctx->mf_ready = mode == 1 && ctx->mf_ready;
    if (mode == 0) return lu_gather_deferred(ctx);
    return PEEC_OK;
}

extern "C" int peec_lu_solve_info(peec_ctx *ctx, int64_t *out4)
{
    if (!ctx || !out4) {
        peec_set_error("peec_lu_solve_info: invalid argument");
        return PEEC_ERR_ARG;
    }
    out4[0] = ctx->lu_solve;
    out4[1] = ctx->mf_ready && ctx->lu_solve == 1 && ctx->lu_layout_from_plan ? 1 : 0;
    out4[2] = ctx->mf.ready ? ctx->mf.H : 0;
    out4[3] = ctx->mf.ready ? ctx->mf.max_tasks : 0;
    return PEEC_OK;
}

static int ready(peec_ctx *ctx, const char *what)
{
    if (!ctx || !ctx->op_ready || !ctx->freq_ready || !ctx->lu_ready) {
        peec_set_error("%s: operator, frequency and LU factors must be set first", what);
        return PEEC_ERR_STATE;
    }
    return PEEC_OK;
}

extern "C" int peec_matvec(peec_ctx *ctx, const peec_c128 *x, peec_c128 *y)
{
    int rc = ready(ctx, "peec_matvec");
    if (rc) return rc;
    if (!x || !y) {
        peec_set_error("peec_matvec: null vector");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t N = ctx->m + ctx->ns + ctx->n;
    PEEC_HIP_TRY(hipMemcpyAsync(ctx->vx.p, x, (size_t)N * 16, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = matvec_device(ctx, ctx->vx.p, ctx->vy.p, 1, N, false, uniform_cols(1, 1)))) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(y, ctx->vy.p, (size_t)N * 16, hipMemcpyDeviceToHost, ctx->stream));
    return sync_stream(ctx);
}

// K x only (no preconditioner); exported for parity tests of the operator
extern "C" int peec_apply_K(peec_ctx *ctx, const peec_c128 *x, peec_c128 *y)
{
    if (!ctx || !ctx->op_ready || !ctx->freq_ready) {
        peec_set_error("peec_apply_K: operator and frequency must be set first");
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t N = ctx->m + ctx->ns + ctx->n;
    PEEC_HIP_TRY(hipMemcpyAsync(ctx->vx.p, x, (size_t)N * 16, hipMemcpyHostToDevice, ctx->stream));
    int rc = apply_K(ctx, ctx->vx.p, ctx->vy.p, 1, N, false, uniform_cols(1, 1));
    if (rc) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(y, ctx->vy.p, (size_t)N * 16, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

static int apply_batched_impl(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, int precond, bool adjoint,
                              const peec_c128 *x, peec_c128 *y);

extern "C" int peec_apply_K_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, int precond,
                                    const peec_c128 *x, peec_c128 *y)
{
    return apply_batched_impl(ctx, nrhs, rhs_per_slot, fp32, precond, false, x, y);
}

extern "C" int peec_apply_KH_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, int precond,
                                     const peec_c128 *x, peec_c128 *y)
{
    return apply_batched_impl(ctx, nrhs, rhs_per_slot, fp32, precond, true, x, y);
}

// ---- K x without allocating (peec_apply_K_noalloc)
//
// The only device memory K x needs beyond its operands and the operator
// data set up earlier is the pair of FFT work grids.  Here they are one
// file-scope device array of kNoAllocGridElems complex128 entries per grid
// (64 MiB in all), which every GPU the library runs on holds until the
// process ends.  In host code `scratch` is a stand-in for the symbol, not a
// device address: the address is looked up once per device
// (hipGetSymbolAddress, that device current) and only that pointer is ever
// passed to kernels.  All calls on one GPU share the array; calls that
// overlap on one GPU (two contexts, two streams) race.  Every grid entry a
// convolution reads is written by its own earlier pass, so nothing needs
// clearing between calls.
constexpr size_t kNoAllocGridElems = (size_t)1 << 21;
static __device__ __align__(256) double2_t scratch[2 * kNoAllocGridElems];

constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static double2_t *g_scratch_ptr[kNoAllocMaxDevices] = {};

// The device address of `scratch` on `device` (which the caller has made
// current).
static int scratch_address(int device, double2_t **out)
{
    if (device < 0 || device >= kNoAllocMaxDevices) {
        peec_set_error("peec_apply_K_noalloc: device %d outside the %d the library tracks", device,
                       kNoAllocMaxDevices);
        return PEEC_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            peec_set_error("peec_apply_K_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                           hipGetErrorString(e));
            return PEEC_ERR_HIP;
        }
        g_scratch_ptr[device] = (double2_t *)p;
    }
    *out = g_scratch_ptr[device];
    return PEEC_OK;
}

extern "C" int peec_apply_K_noalloc(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, const peec_c128 *d_x,
                                    peec_c128 *d_y)
{
    if (!ctx || nrhs < 1 || rhs_per_slot < 1 || !d_x || !d_y) {
        peec_set_error("peec_apply_K_noalloc: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready) {
        peec_set_error("peec_apply_K_noalloc: operator and frequency slots must be set first");
        return PEEC_ERR_STATE;
    }
    if (ctx->prof_on) {
        peec_set_error("peec_apply_K_noalloc: kernel profiling (peec_profile) creates events; switch it off first");
        return PEEC_ERR_STATE;
    }
    if ((nrhs + rhs_per_slot - 1) / rhs_per_slot > ctx->nslots) {
        peec_set_error("peec_apply_K_noalloc: %d right-hand sides at %d per slot need more than the %d slots", nrhs,
                       rhs_per_slot, ctx->nslots);
        return PEEC_ERR_ARG;
    }
    const size_t need = work_grid_elems(ctx, nrhs);
    if (need > kNoAllocGridElems) {
        peec_set_error("peec_apply_K_noalloc: %d right-hand sides need FFT work grids of %zu complex128 entries, "
                       "above the %zu this entry point holds; use fewer right-hand sides per call or "
                       "peec_apply_K_batched",
                       nrhs, need, kNoAllocGridElems);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    double2_t *w = nullptr;
    int rc = scratch_address(ctx->device, &w);
    if (rc) return rc;
    const int64_t N = ctx->m + ctx->ns + ctx->n;
    return apply_K_work(ctx, (const double2_t *)d_x, (double2_t *)d_y, nrhs, N, fp32 != 0,
                        uniform_cols(nrhs, rhs_per_slot), w, w + kNoAllocGridElems);
}

static int apply_batched_impl(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, int precond, bool adjoint,
                              const peec_c128 *x, peec_c128 *y)
{
    if (!ctx || nrhs < 1 || rhs_per_slot < 1 || !x || !y) {
        peec_set_error("peec_apply_K_batched: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready || !ctx->freq_ready || (precond && !ctx->lu_ready)) {
        peec_set_error("peec_apply_K_batched: operator, frequency slots%s must be set first",
                       precond ? " and factors" : "");
        return PEEC_ERR_STATE;
    }
    if ((nrhs + rhs_per_slot - 1) / rhs_per_slot > ctx->nslots) {
        peec_set_error("peec_apply_K_batched: %d right-hand sides at %d per slot need more than the %d slots", nrhs,
                       rhs_per_slot, ctx->nslots);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t N = ctx->m + ctx->ns + ctx->n;
    const size_t bytes = (size_t)N * (size_t)nrhs * 16;
    DevBuf<double2_t> X, Y;
    if (!X.alloc((size_t)(N * nrhs)) || !Y.alloc((size_t)(N * nrhs))) {
        peec_set_error("peec_apply_K_batched: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    PEEC_HIP_TRY(hipMemcpyAsync(X.p, x, bytes, hipMemcpyHostToDevice, ctx->stream));
    const ColMap cm = uniform_cols(nrhs, rhs_per_slot);
    int rc;
    if (precond == 2) {
        // the preconditioner alone: M^-1 x or M^-H x
        PEEC_HIP_TRY(hipMemcpyAsync(Y.p, X.p, bytes, hipMemcpyDeviceToDevice, ctx->stream));
        rc = adjoint ? apply_precond_adj(ctx, Y.p, nrhs, N, cm) : apply_precond(ctx, Y.p, nrhs, N, cm);
    } else if (adjoint) {
        rc = precond ? matvec_device_adj(ctx, X.p, Y.p, nrhs, N, fp32 != 0, cm)
                     : apply_KH(ctx, X.p, Y.p, nrhs, N, fp32 != 0, cm);
    } else {
        rc = precond ? matvec_device(ctx, X.p, Y.p, nrhs, N, fp32 != 0, cm)
                     : apply_K(ctx, X.p, Y.p, nrhs, N, fp32 != 0, cm);
    }
    if (rc) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(y, Y.p, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

extern "C" int peec_precond_rhs(peec_ctx *ctx, const double *is, peec_c128 *tn)
{
    int rc = ready(ctx, "peec_precond_rhs");
    if (rc) return rc;
    if (!is || !tn) {
        peec_set_error("peec_precond_rhs: null vector");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t m = ctx->m, ns = ctx->ns, n = ctx->n, N = m + ns + n;
    // precond_3_3_Kt_fft! is M^-1 applied to [0; 0; is]
    std::vector<double2_t> h((size_t)N, {0.0, 0.0});
    for (int64_t v = 0; v < n; v++) h[(size_t)(m + ns + v)] = {is[v], 0.0};
    PEEC_HIP_TRY(hipMemcpyAsync(ctx->vy.p, h.data(), (size_t)N * 16, hipMemcpyHostToDevice, ctx->stream));
    if ((rc = apply_precond(ctx, ctx->vy.p, 1, N, uniform_cols(1, 1)))) return rc;
    PEEC_HIP_TRY(hipMemcpyAsync(tn, ctx->vy.p, (size_t)N * 16, hipMemcpyDeviceToHost, ctx->stream));
    return sync_stream(ctx);
}
