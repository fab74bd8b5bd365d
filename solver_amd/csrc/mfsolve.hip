// Multifrontal triangular solves with the device LU's fronts.
//
// The preconditioner solves SS M = r once per matvec (precond_3_3_vector,
// compute_Matrix_vector_fft.jl:84-146; one solve of the combined right-hand
// side, as in trsv.hip).  When the device factorization holds every
// frequency slot's fronts in its pool, the solves read the factors there, in
// place, front by front along the assembly tree:
//
//   forward (L c = r), heights 0, 1, ..., H-1: front K gathers its working
//     rows w = [r at its pivots; 0 at its boundary] + the updates u of its
//     children (extend-add), solves its pivots y = L11^-1 w_P, and passes
//     u = w_B - L21 y to its parent;
//   backward (U x = c), heights H-1, ..., 0: front K reads x at its
//     boundary (pivots of ancestors, already solved), forms
//     t = c_P - U12 x_B and solves x_P = U11^-1 t.
//
// Fronts of one height are independent.  The factorization leaves explicit
// inverses of every pivot block L11 and U11 in the pool (lu.hip), so a
// front's pivots are one dense product instead of a chain of 64-row
// substitutions: the critical path is the tree height (15 at 96 x 96 x 24),
// not the number of 64-row blocks on it (about 85).  Every dense operand --
// L11^-1, L21, U12, U11^-1 -- is read once per frequency slot and serves
// all of that slot's right-hand sides.
//
// Small fronts (s <= 64 pivots, b <= kMfSmallB = 256 boundary rows) take
// one workgroup per (front, slot) for the whole front.  Large fronts are cut
// into GEMV tasks of at most 64 rows x kMfKc columns; each task writes its
// partial sums and a finishing launch adds them in a fixed order, so the
// results are the same run to run.
#include "context.h"

#include <algorithm>

namespace {

typedef double2_t cd;

constexpr int kMfKc = 256;       // columns of A per GEMV task (B staged in LDS)
constexpr int kMfSmallB = 256;   // small fronts: s <= 64 and b <= kMfSmallB

// acc += a b and acc -= a b, complex, as FMA chains
__device__ __forceinline__ void cfma(cd &acc, cd a, cd b)
{
    acc.x = __builtin_fma(a.x, b.x, acc.x);
    acc.x = __builtin_fma(-a.y, b.y, acc.x);
    acc.y = __builtin_fma(a.x, b.y, acc.y);
    acc.y = __builtin_fma(a.y, b.x, acc.y);
}
__device__ __forceinline__ void cfms(cd &acc, cd a, cd b)
{
    acc.x = __builtin_fma(-a.x, b.x, acc.x);
    acc.x = __builtin_fma(a.y, b.y, acc.x);
    acc.y = __builtin_fma(-a.x, b.y, acc.y);
    acc.y = __builtin_fma(-a.y, b.x, acc.y);
}

// Sum over the 4 lanes of a quad (lanes 4q .. 4q+3), in the same order in
// every lane: DPP quad swaps at distance 1, then 2.
template <int CTRL>
__device__ __forceinline__ double mf_dpp(double v)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ cd quad_sum(cd v)
{
    v.x += mf_dpp<0xB1>(v.x);  // quad_perm [1, 0, 3, 2]
    v.y += mf_dpp<0xB1>(v.y);
    v.x += mf_dpp<0x4E>(v.x);  // quad_perm [2, 3, 0, 1]
    v.y += mf_dpp<0x4E>(v.y);
    return v;
}

// acc[c] += sum over the columns k = ph, ph + 4, ... < kend of
// A[k * lda] * B[k * Q + c] (A already at the lane's row), 8 loads in flight.
template <int Q>
__device__ __forceinline__ void mf_row_phase(cd (&acc)[Q], const cd *A, int64_t lda, const cd *B, int ph, int kend)
{
    int k = ph;
    for (; k + 28 < kend; k += 32) {
        cd a[8];
#pragma unroll
        for (int u = 0; u < 8; u++) a[u] = A[(int64_t)(k + 4 * u) * lda];
#pragma unroll
        for (int u = 0; u < 8; u++)
#pragma unroll
            for (int c = 0; c < Q; c++) cfma(acc[c], a[u], B[(k + 4 * u) * Q + c]);
    }
    for (; k < kend; k += 4) {
        const cd a = A[(int64_t)k * lda];
#pragma unroll
        for (int c = 0; c < Q; c++) cfma(acc[c], a, B[k * Q + c]);
    }
}

// Small fronts: one workgroup per (front list[blockIdx.x], slot
// blockIdx.y), Q right-hand sides per pass.  A wavefront takes 16 rows, a
// quad of lanes per row, each lane every fourth column; the quad's sums
// are added by DPP.  The pivot-block inverses are stored whole (zeros in the
// other triangle), so every row runs over all s columns.
template <int Q>
__global__ void __launch_bounds__(256) k_mf_fwd_small(const int32_t *list, const MfFront *fr, const cd *pool0,
                                                      int64_t stride, const int64_t *cptr, const int64_t *csrc,
                                                      const cd *in, cd *out, cd *ubuf, int64_t utot, int64_t n,
                                                      int nrhs, ColMap cm)
{
    __shared__ cd sw[64 * Q];
    __shared__ cd sy[64 * Q];
    const int slot = (int)blockIdx.y;
    const MfFront F = fr[list[blockIdx.x]];
    const cd *pool = pool0 + (int64_t)slot * stride;
    const cd *Li = pool + F.inv;           // L11^-1, s x s
    const cd *L21 = pool + F.off + F.s;    // column k at k * f
    const int s = F.s, b = F.f - F.s;
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int i = threadIdx.x & 63, cg = threadIdx.x >> 6;
    const int ph = threadIdx.x & 3, rr = (int)(threadIdx.x >> 2);  // quad: row rr (0..63), column phase ph
    for (int r0 = rbeg; r0 < rend; r0 += Q) {
        const int nc = min(Q, rend - r0);
        // w at the pivots: r plus the children's updates
        for (int c = cg; c < Q; c += 4) {
            cd w = {0.0, 0.0};
            if (i < s && c < nc) {
                const int64_t r = r0 + c;
                w = in[r * n + F.first + i];
                for (int64_t e = cptr[F.crow + i]; e < cptr[F.crow + i + 1]; e++) {
                    const cd v = ubuf[r * utot + csrc[e]];
                    w.x += v.x;
                    w.y += v.y;
                }
            }
            sw[i * Q + c] = w;
        }
        __syncthreads();
        // y = L11^-1 w
        {
            cd acc[Q];
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = {0.0, 0.0};
            mf_row_phase<Q>(acc, Li + min(rr, s - 1), s, sw, ph, s);
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = quad_sum(acc[c]);
            if (ph == 0) {
#pragma unroll
                for (int c = 0; c < Q; c++) {
                    const bool live = rr < s && c < nc;
                    if (live) out[(int64_t)(r0 + c) * n + F.first + rr] = acc[c];
                    sy[rr * Q + c] = live ? acc[c] : cd{0.0, 0.0};
                }
            }
        }
        __syncthreads();
        // u = w_B - L21 y, 64 boundary rows at a time
        for (int jb = 0; jb < b; jb += 64) {
            const int j = jb + rr;
            cd acc[Q];
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = {0.0, 0.0};
            mf_row_phase<Q>(acc, L21 + min(j, b - 1), F.f, sy, ph, s);
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = quad_sum(acc[c]);
            if (ph == 0 && j < b) {
                const int64_t e0 = cptr[F.crow + s + j], e1 = cptr[F.crow + s + j + 1];
#pragma unroll
                for (int c = 0; c < Q; c++) {
                    if (c >= nc) continue;
                    cd w = {0.0, 0.0};
                    for (int64_t e = e0; e < e1; e++) {
                        const cd v = ubuf[(int64_t)(r0 + c) * utot + csrc[e]];
                        w.x += v.x;
                        w.y += v.y;
                    }
                    ubuf[(int64_t)(r0 + c) * utot + F.uoff + j] = {w.x - acc[c].x, w.y - acc[c].y};
                }
            }
        }
        __syncthreads();
    }
}

// Backward sweep of the small fronts of one height (same grid and lanes).
template <int Q>
__global__ void __launch_bounds__(256) k_mf_bwd_small(const int32_t *list, const MfFront *fr, const cd *pool0,
                                                      int64_t stride, const int32_t *bnd, const cd *in, cd *out,
                                                      int64_t n, int nrhs, ColMap cm)
{
    __shared__ cd sx[kMfSmallB * Q];  // x at the boundary
    __shared__ cd st[64 * Q];
    const int slot = (int)blockIdx.y;
    const MfFront F = fr[list[blockIdx.x]];
    const cd *pool = pool0 + (int64_t)slot * stride;
    const int s = F.s, b = F.f - F.s;
    const cd *Ui = pool + F.inv + (int64_t)s * s;      // U11^-1, s x s
    const cd *U12 = pool + F.off + (int64_t)s * F.f;   // column k at k * f, rows 0 .. s
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int ph = threadIdx.x & 3, rr = (int)(threadIdx.x >> 2);
    const int row = min(rr, s - 1);  // rows past s read row s - 1 and are not written
    for (int r0 = rbeg; r0 < rend; r0 += Q) {
        const int nc = min(Q, rend - r0);
        for (int e = threadIdx.x; e < b * Q; e += 256) {
            const int k = e / Q, c = e % Q;
            sx[e] = c < nc ? out[(int64_t)(r0 + c) * n + bnd[F.bptr + k]] : cd{0.0, 0.0};
        }
        __syncthreads();
        // t = c_P - U12 x_B
        {
            cd acc[Q];
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = {0.0, 0.0};
            mf_row_phase<Q>(acc, U12 + row, F.f, sx, ph, b);
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = quad_sum(acc[c]);
            if (ph == 0) {
#pragma unroll
                for (int c = 0; c < Q; c++) {
                    cd t = {0.0, 0.0};
                    if (rr < s && c < nc) {
                        const cd v = in[(int64_t)(r0 + c) * n + F.first + rr];
                        t = {v.x - acc[c].x, v.y - acc[c].y};
                    }
                    st[rr * Q + c] = t;
                }
            }
        }
        __syncthreads();
        // x_P = U11^-1 t
        {
            cd acc[Q];
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = {0.0, 0.0};
            mf_row_phase<Q>(acc, Ui + row, s, st, ph, s);
#pragma unroll
            for (int c = 0; c < Q; c++) acc[c] = quad_sum(acc[c]);
            if (ph == 0 && rr < s) {
#pragma unroll
                for (int c = 0; c < Q; c++)
                    if (c < nc) out[(int64_t)(r0 + c) * n + F.first + rr] = acc[c];
            }
        }
        __syncthreads();
    }
}

// Working rows w of the large fronts of one height: rows[e] = (front, row).
__global__ void __launch_bounds__(256) k_mf_assemble(const int2 *rows, int64_t nrows, const MfFront *fr,
                                                     const int64_t *cptr, const int64_t *csrc, const cd *in, cd *wbuf,
                                                     int64_t wtot, const cd *ubuf, int64_t utot, int64_t n, int nrhs,
                                                     ColMap cm)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nrows) return;
    const int slot = (int)blockIdx.y;
    const int2 fj = rows[e];
    const MfFront F = fr[fj.x];
    const int j = fj.y;
    const int64_t e0 = cptr[F.crow + j], e1 = cptr[F.crow + j + 1];
    const int rend = cm.end(slot, nrhs);
    for (int64_t r = cm.first(slot); r < rend; r++) {
        cd w = j < F.s ? in[r * n + F.first + j] : cd{0.0, 0.0};
        for (int64_t c = e0; c < e1; c++) {
            const cd v = ubuf[r * utot + csrc[c]];
            w.x += v.x;
            w.y += v.y;
        }
        wbuf[r * wtot + F.woff + j] = w;
    }
}

// D += A (16 rows from Ar, columns 0 .. kend, leading dimension lda) times
// B (rows 0 .. kend of an LDS block of 8 right-hand sides, [k][c]), on the
// matrix cores as in k_mf_gemv8m: lane l loads A[l % 16][4 s + l / 16] whole
// and feeds its real and imaginary parts to two v_mfma_f64_16x16x4f64.
// Ar already points at the lane's row (clamped by the caller).
typedef double v4d __attribute__((ext_vector_type(4)));
__device__ __forceinline__ void mf_mfma_rows(v4d &acc, const cd *Ar, int64_t lda, const cd *sB, int kend, int q,
                                             int bc, bool im_out)
{
    const int ksteps = (kend + 3) / 4;
    int st = 0;
    for (; st + 4 <= ksteps; st += 4) {
        cd a[4], b[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int kk = 4 * (st + u) + q;
            a[u] = kk < kend ? Ar[kk * lda] : cd{0.0, 0.0};
            b[u] = kk < kend ? sB[kk * 8 + bc] : cd{0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, im_out ? b[u].y : b[u].x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].y, im_out ? b[u].x : -b[u].y, acc, 0, 0, 0);
        }
    }
    for (; st < ksteps; st++) {
        const int kk = 4 * st + q;
        const cd a = kk < kend ? Ar[kk * lda] : cd{0.0, 0.0};
        const cd b = kk < kend ? sB[kk * 8 + bc] : cd{0.0, 0.0};
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, im_out ? b.y : b.x, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, im_out ? b.x : -b.y, acc, 0, 0, 0);
    }
}

// Small fronts, passes of 8 right-hand sides, on the matrix cores: the same
// steps as k_mf_fwd_small / k_mf_bwd_small, each product as 16-row blocks
// (wavefront w takes rows 16 w .. 16 w + 15 of a 64-row block).  Lane l
// holds row 16 w + l / 16 + 4 j, right-hand side l % 8, real (l % 16 < 8) or
// imaginary part of the result.
__global__ void __launch_bounds__(256) k_mf_fwd_small8m(const int32_t *list, const MfFront *fr, const cd *pool0,
                                                        int64_t stride, const int64_t *cptr, const int64_t *csrc,
                                                        const cd *in, cd *out, cd *ubuf, int64_t utot, int64_t n,
                                                        int nrhs, ColMap cm)
{
    __shared__ cd sw[64 * 8];
    __shared__ cd sy[64 * 8];
    const int slot = (int)blockIdx.y;
    const MfFront F = fr[list[blockIdx.x]];
    const cd *pool = pool0 + (int64_t)slot * stride;
    const int s = F.s, b = F.f - F.s;
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int i = threadIdx.x & 63, cg = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, c16 = lane & 15, bc = c16 & 7;
    const int part = c16 >> 3;
    const bool im_out = part == 1;
    double *outd = reinterpret_cast<double *>(out);
    double *ud = reinterpret_cast<double *>(ubuf);
    double *syd = reinterpret_cast<double *>(sy);
    for (int r0 = rbeg; r0 < rend; r0 += 8) {
        const int nc = min(8, rend - r0);
        // w at the pivots: r plus the children's updates
        for (int c = cg; c < 8; c += 4) {
            cd w = {0.0, 0.0};
            if (i < s && c < nc) {
                const int64_t r = r0 + c;
                w = in[r * n + F.first + i];
                for (int64_t e = cptr[F.crow + i]; e < cptr[F.crow + i + 1]; e++) {
                    const cd v = ubuf[r * utot + csrc[e]];
                    w.x += v.x;
                    w.y += v.y;
                }
            }
            sw[i * 8 + c] = w;
        }
        __syncthreads();
        // y = L11^-1 w (stored whole: zeros above the diagonal)
        if (16 * wave < s) {
            v4d acc = {0.0, 0.0, 0.0, 0.0};
            mf_mfma_rows(acc, pool + F.inv + min(16 * wave + c16, s - 1), s, sw, s, q, bc, im_out);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = 16 * wave + q + 4 * j;
                const bool live = row < s && bc < nc;
                if (live) outd[((int64_t)(r0 + bc) * n + F.first + row) * 2 + part] = acc[j];
                syd[(row * 8 + bc) * 2 + part] = live ? acc[j] : 0.0;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) syd[((16 * wave + q + 4 * j) * 8 + bc) * 2 + part] = 0.0;
        }
        __syncthreads();
        // u = w_B - L21 y, 64 boundary rows at a time
        for (int jb = 0; jb < b; jb += 64) {
            if (jb + 16 * wave >= b) continue;  // wave-uniform: no barrier inside
            v4d acc = {0.0, 0.0, 0.0, 0.0};
            mf_mfma_rows(acc, pool + F.off + F.s + min(jb + 16 * wave + c16, b - 1), F.f, sy, s, q, bc, im_out);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = jb + 16 * wave + q + 4 * j;
                if (row >= b || bc >= nc) continue;
                double w = 0.0;
                for (int64_t e = cptr[F.crow + s + row]; e < cptr[F.crow + s + row + 1]; e++)
                    w += ud[((int64_t)(r0 + bc) * utot + csrc[e]) * 2 + part];
                ud[((int64_t)(r0 + bc) * utot + F.uoff + row) * 2 + part] = w - acc[j];
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_mf_bwd_small8m(const int32_t *list, const MfFront *fr, const cd *pool0,
                                                        int64_t stride, const int32_t *bnd, const cd *in, cd *out,
                                                        int64_t n, int nrhs, ColMap cm)
{
    __shared__ cd sx[kMfSmallB * 8];  // x at the boundary
    __shared__ cd st[64 * 8];
    const int slot = (int)blockIdx.y;
    const MfFront F = fr[list[blockIdx.x]];
    const cd *pool = pool0 + (int64_t)slot * stride;
    const int s = F.s, b = F.f - F.s;
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, q = lane >> 4, c16 = lane & 15, bc = c16 & 7;
    const int part = c16 >> 3;
    const bool im_out = part == 1;
    const bool rows_here = 16 * wave < s;  // wave-uniform
    const int arow = min(16 * wave + c16, s - 1);
    const double *ind = reinterpret_cast<const double *>(in);
    double *outd = reinterpret_cast<double *>(out);
    double *std_ = reinterpret_cast<double *>(st);
    for (int r0 = rbeg; r0 < rend; r0 += 8) {
        const int nc = min(8, rend - r0);
        for (int e = threadIdx.x; e < b * 8; e += 256) {
            const int k = e >> 3, c = e & 7;
            sx[e] = c < nc ? out[(int64_t)(r0 + c) * n + bnd[F.bptr + k]] : cd{0.0, 0.0};
        }
        __syncthreads();
        // t = c_P - U12 x_B
        if (rows_here) {
            v4d acc = {0.0, 0.0, 0.0, 0.0};
            mf_mfma_rows(acc, pool + F.off + (int64_t)s * F.f + arow, F.f, sx, b, q, bc, im_out);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = 16 * wave + q + 4 * j;
                double t = 0.0;
                if (row < s && bc < nc) t = ind[((int64_t)(r0 + bc) * n + F.first + row) * 2 + part] - acc[j];
                std_[(row * 8 + bc) * 2 + part] = t;
            }
        } else {
#pragma unroll
            for (int j = 0; j < 4; j++) std_[((16 * wave + q + 4 * j) * 8 + bc) * 2 + part] = 0.0;
        }
        __syncthreads();
        // x_P = U11^-1 t (stored whole: zeros below the diagonal)
        if (rows_here) {
            v4d acc = {0.0, 0.0, 0.0, 0.0};
            mf_mfma_rows(acc, pool + F.inv + (int64_t)s * s + arow, s, st, s, q, bc, im_out);
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const int row = 16 * wave + q + 4 * j;
                if (row < s && bc < nc) outd[((int64_t)(r0 + bc) * n + F.first + row) * 2 + part] = acc[j];
            }
        }
        __syncthreads();
    }
}

// x at the boundary of the large fronts of one height, into their w rows
// s .. f (the backward sweep's right-hand side of U12 x_B).
__global__ void __launch_bounds__(256) k_mf_gather_b(const int2 *rows, int64_t nrows, const MfFront *fr,
                                                     const int32_t *bnd, const cd *x, cd *wbuf, int64_t wtot,
                                                     int64_t n, int nrhs, ColMap cm)
{
    const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= nrows) return;
    const int slot = (int)blockIdx.y;
    const int2 fj = rows[e];
    const MfFront F = fr[fj.x];
    if (fj.y < F.s) return;
    const int64_t pos = bnd[F.bptr + fj.y - F.s];
    const int rend = cm.end(slot, nrhs);
    for (int64_t r = cm.first(slot); r < rend; r++) wbuf[r * wtot + F.woff + fj.y] = x[r * n + pos];
}

// GEMV tasks of the large fronts: partial sums of rows r0 .. r0 + m of
// A (kc columns) times rows k0 .. k0 + kc of the front's working rows w.
// The slot's right-hand sides go Q per pass; their rows of w are staged in
// LDS and each of the four wavefronts takes a quarter of the columns, one
// row per lane, with 8 loads of A in flight.
template <int Q>
__global__ void __launch_bounds__(256) k_mf_gemv(const MfTask *tasks, const MfFront *fr, const cd *pool0,
                                                 int64_t stride, const cd *wbuf, int64_t wtot, cd *part, int nrhs,
                                                 ColMap cm)
{
    __shared__ cd sb[kMfKc * Q];  // B rows [k][c]; then the four wavefronts' sums [wave][row][c]
    const int slot = (int)blockIdx.y;
    const MfTask t = tasks[blockIdx.x];
    const MfFront F = fr[t.front];
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // lanes past m read row m - 1 (in bounds) and their sums are not written
    const cd *Ap = pool0 + (int64_t)slot * stride + t.a + min(lane, t.m - 1);
    const int64_t lda = t.lda;
    // partials [task][row][column]: a slot's tasks write its own columns
    cd *pout = part + (int64_t)blockIdx.x * 64 * nrhs;
    const int kb = wave * 64, ke = min(t.kc, kb + 64);
    for (int r0 = rbeg; r0 < rend; r0 += Q) {
        const int nc = min(Q, rend - r0);
        __syncthreads();
        for (int e = threadIdx.x; e < t.kc * Q; e += 256) {
            const int k = e / Q, c = e % Q;
            // columns past the pass read its last one (in bounds) and store 0
            const cd v = wbuf[(int64_t)(r0 + min(c, nc - 1)) * wtot + F.woff + t.k0 + k];
            const bool keep = c < nc;
            sb[e] = {keep ? v.x : 0.0, keep ? v.y : 0.0};
        }
        __syncthreads();
        cd acc[Q];
#pragma unroll
        for (int c = 0; c < Q; c++) acc[c] = {0.0, 0.0};
        int k = kb;
        for (; k + 8 <= ke; k += 8) {
            cd a[8];
#pragma unroll
            for (int u = 0; u < 8; u++) a[u] = Ap[(int64_t)(k + u) * lda];
#pragma unroll
            for (int u = 0; u < 8; u++)
#pragma unroll
                for (int c = 0; c < Q; c++) cfma(acc[c], a[u], sb[(k + u) * Q + c]);
        }
        for (; k < ke; k++) {
            const cd a = Ap[(int64_t)k * lda];
#pragma unroll
            for (int c = 0; c < Q; c++) cfma(acc[c], a, sb[k * Q + c]);
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < Q; c++) sb[(wave * 64 + lane) * Q + c] = acc[c];
        __syncthreads();
        for (int e = threadIdx.x; e < 64 * Q; e += 256) {
            const int i = e / Q, c = e % Q;
            if (i < t.m && c < nc) {
                const cd p0 = sb[i * Q + c], p1 = sb[(64 + i) * Q + c], p2 = sb[(128 + i) * Q + c],
                         p3 = sb[(192 + i) * Q + c];
                pout[(int64_t)i * nrhs + r0 + c] = {(p0.x + p1.x) + (p2.x + p3.x), (p0.y + p1.y) + (p2.y + p3.y)};
            }
        }
    }
}

// The same tasks for passes of 8 right-hand sides on the matrix cores.  The
// complex product is the real one C' = A' B' with A' = [Ar Ai] and, for the
// rows of A' holding column k's real / imaginary parts, B' rows [Br Bi] /
// [-Bi Br] (16 columns: the 8 right-hand sides' real parts, then their
// imaginary parts).  Wavefront w owns rows 16 w .. 16 w + 15 of the task over
// all its columns.  Lane l loads the whole complex A[16 w + l % 16][4 s + l / 16]
// and feeds its real part to one v_mfma_f64_16x16x4f64 and its imaginary part
// to the next (lane l holds A'[l % 16][l / 16], B'[l / 16][l % 16] and
// D[l / 16 + 4 j][l % 16]): four complex columns per pair of instructions, no
// wavefront waits on another, and the LDS serves only the staged rows of w.
__global__ void __launch_bounds__(256) k_mf_gemv8m(const MfTask *tasks, const MfFront *fr, const cd *pool0,
                                                   int64_t stride, const cd *wbuf, int64_t wtot, cd *part, int nrhs,
                                                   ColMap cm)
{
    __shared__ cd sb[kMfKc * 8];
    const int slot = (int)blockIdx.y;
    const MfTask t = tasks[blockIdx.x];
    const MfFront F = fr[t.front];
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int q = lane >> 4, c16 = lane & 15;
    // A fragment: row 16 w + l % 16 (rows past m read row m - 1 and are not
    // written), column 4 s + q
    const int arow = min(wave * 16 + c16, t.m - 1);
    const cd *Ap = pool0 + (int64_t)slot * stride + t.a + arow;
    const int64_t lda = t.lda;
    // B' fragment: right-hand side c16 & 7, output part c16 >> 3 (0 real,
    // 1 imaginary); with A's real part B' = [Br Bi], with its imaginary part
    // B' = [-Bi Br]
    const int bc = c16 & 7;
    const bool im_out = c16 >= 8;
    double *pout = reinterpret_cast<double *>(part + (int64_t)blockIdx.x * 64 * nrhs);
    const int ksteps = (t.kc + 3) / 4;
    for (int r0 = rbeg; r0 < rend; r0 += 8) {
        const int nc = min(8, rend - r0);
        __syncthreads();
        for (int e = threadIdx.x; e < t.kc * 8; e += 256) {
            const int k = e >> 3, c = e & 7;
            const cd v = wbuf[(int64_t)(r0 + min(c, nc - 1)) * wtot + F.woff + t.k0 + k];
            const bool keep = c < nc;
            sb[e] = {keep ? v.x : 0.0, keep ? v.y : 0.0};
        }
        __syncthreads();
        v4d acc = {0.0, 0.0, 0.0, 0.0};
        int st = 0;
        for (; st + 4 <= ksteps; st += 4) {
            cd a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int kk = 4 * (st + u) + q;
                a[u] = kk < t.kc ? Ap[kk * lda] : cd{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int kk = 4 * (st + u) + q;
                b[u] = kk < t.kc ? sb[kk * 8 + bc] : cd{0.0, 0.0};
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].x, im_out ? b[u].y : b[u].x, acc, 0, 0, 0);
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[u].y, im_out ? b[u].x : -b[u].y, acc, 0, 0, 0);
            }
        }
        for (; st < ksteps; st++) {
            const int kk = 4 * st + q;
            const cd a = kk < t.kc ? Ap[kk * lda] : cd{0.0, 0.0};
            const cd b = kk < t.kc ? sb[kk * 8 + bc] : cd{0.0, 0.0};
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.x, im_out ? b.y : b.x, acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a.y, im_out ? b.x : -b.y, acc, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int row = wave * 16 + q + 4 * j;
            if (row < t.m && bc < nc) pout[((int64_t)row * nrhs + r0 + bc) * 2 + (c16 >> 3)] = acc[j];
        }
    }
}

// Sums of a large front's task partials, in task order, and what they make:
// KIND 0: y (forward pivots) -> out and w;  1: u = w_B - sum -> u area;
// KIND 2: t = in - sum (backward pivots) -> w;  3: x -> out.
template <int KIND>
__global__ void __launch_bounds__(256) k_mf_finish(const MfFin *fins, const MfFront *fr, const cd *part,
                                                   const cd *in, cd *out, cd *wbuf, int64_t wtot, cd *ubuf,
                                                   int64_t utot, int64_t n, int nrhs, ColMap cm)
{
    const int slot = (int)blockIdx.y;
    const MfFin fi = fins[blockIdx.x];
    const int i = threadIdx.x & 63;
    if (i >= fi.m) return;  // no barrier below
    const MfFront F = fr[fi.front];
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    const int row = fi.r0 + i;
    for (int r = rbeg + (int)(threadIdx.x >> 6); r < rend; r += 4) {
        cd sum = {0.0, 0.0};
        for (int p = 0; p < fi.np; p++) {
            const cd v = part[((int64_t)(fi.p0 + p) * 64 + i) * nrhs + r];
            sum.x += v.x;
            sum.y += v.y;
        }
        const int64_t rr = r;
        if (KIND == 0) {
            out[rr * n + F.first + row] = sum;
            wbuf[rr * wtot + F.woff + row] = sum;
        } else if (KIND == 1) {
            const cd w = wbuf[rr * wtot + F.woff + F.s + row];
            ubuf[rr * utot + F.uoff + row] = {w.x - sum.x, w.y - sum.y};
        } else if (KIND == 2) {
            const cd v = in[rr * n + F.first + row];
            wbuf[rr * wtot + F.woff + row] = {v.x - sum.x, v.y - sum.y};
        } else {
            out[rr * n + F.first + row] = sum;
        }
    }
}

template <typename T>
int mf_upload(DevBuf<T> &d, const std::vector<T> &h)
{
    if (!d.alloc(std::max<size_t>(h.size(), 1))) {
        peec_set_error("peec_lu_analyze: out of device memory for the multifrontal solve plan");
        return PEEC_ERR_NOMEM;
    }
    if (!h.empty()) PEEC_HIP_TRY(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return PEEC_OK;
}

// Columns of the widest slot of a call.
int mf_cols(int nrhs, const ColMap &cm)
{
    return cm.rslot ? (cm.wmax > 0 ? cm.wmax : nrhs) : std::min(cm.rps, nrhs);
}

}  // namespace

int mf_analyze(peec_ctx *ctx)
{
    MfPlan &M = ctx->mf;
    const LuPlan &P = ctx->lu;
    M.ready = false;
    const int64_t nf = (int64_t)P.fronts.size();
    std::vector<int32_t> height((size_t)nf, 0);
    std::vector<std::vector<int32_t>> kids((size_t)nf);
    int H = 0;
    for (int64_t K = 0; K < nf; K++)
        if (P.fronts[(size_t)K].parent >= 0) kids[(size_t)P.fronts[(size_t)K].parent].push_back((int32_t)K);
    for (int64_t K = 0; K < nf; K++) {
        for (int32_t c : kids[(size_t)K]) height[(size_t)K] = std::max(height[(size_t)K], height[(size_t)c] + 1);
        H = std::max(H, height[(size_t)K] + 1);
    }
    M.H = H;
    std::vector<MfFront> fr((size_t)nf);
    std::vector<char> small((size_t)nf, 0);
    int64_t uoff = 0, woff = 0, crow = 0;
    for (int64_t K = 0; K < nf; K++) {
        const LuFrontInfo &F = P.fronts[(size_t)K];
        const int b = F.f - F.s;
        MfFront m{};
        m.off = F.off;
        m.inv = F.inv;
        m.f = F.f;
        m.s = F.s;
        m.first = F.first;
        m.bptr = P.bnd_ptr[(size_t)K];
        m.uoff = uoff;
        uoff += b;
        small[(size_t)K] = F.s <= 64 && b <= kMfSmallB;
        m.woff = small[(size_t)K] ? 0 : woff;
        if (!small[(size_t)K]) woff += F.f;
        m.crow = crow;
        crow += F.f;
        fr[(size_t)K] = m;
    }
    M.utot = std::max<int64_t>(uoff, 1);
    M.wtot = std::max<int64_t>(woff, 1);
    // child K's u row i adds into row loc of its parent's working rows;
    // a row's contributions in child order
    auto local = [&](int32_t K, int32_t q) -> int {
        const LuFrontInfo &F = P.fronts[(size_t)K];
        if (q >= F.first && q < F.first + F.s) return q - F.first;
        const int32_t *b0 = P.bnd.data() + P.bnd_ptr[(size_t)K], *b1 = P.bnd.data() + P.bnd_ptr[(size_t)K + 1];
        const int32_t *it = std::lower_bound(b0, b1, q);
        return it == b1 || *it != q ? -1 : F.s + (int)(it - b0);
    };
    std::vector<int64_t> cptr((size_t)crow + 1, 0);
    std::vector<int32_t> loc(P.bnd.size(), -1);
    for (int64_t K = 0; K < nf; K++) {
        const int32_t par = P.fronts[(size_t)K].parent;
        for (int64_t t = P.bnd_ptr[(size_t)K]; t < P.bnd_ptr[(size_t)K + 1]; t++) {
            const int l = par >= 0 ? local(par, P.bnd[(size_t)t]) : -1;
            if (l < 0) {
                peec_set_error("peec_lu_analyze: boundary of front %lld is not inside its parent", (long long)K);
                return PEEC_ERR_ARG;
            }
            loc[(size_t)t] = l;
            cptr[(size_t)(fr[(size_t)par].crow + l + 1)]++;
        }
    }
    for (int64_t r = 0; r < crow; r++) cptr[(size_t)r + 1] += cptr[(size_t)r];
    std::vector<int64_t> csrc((size_t)std::max<int64_t>(cptr[(size_t)crow], 1));
    {
        std::vector<int64_t> at(cptr.begin(), cptr.end() - 1);
        for (int64_t K = 0; K < nf; K++) {
            const int32_t par = P.fronts[(size_t)K].parent;
            if (par < 0) continue;
            for (int64_t t = P.bnd_ptr[(size_t)K]; t < P.bnd_ptr[(size_t)K + 1]; t++)
                csrc[(size_t)at[(size_t)(fr[(size_t)par].crow + loc[(size_t)t])]++] =
                    fr[(size_t)K].uoff + (t - P.bnd_ptr[(size_t)K]);
        }
    }
    // per height: small fronts, large-front rows, GEMV tasks and finishes
    std::vector<int32_t> slist;
    std::vector<int2> rows;
    std::vector<MfTask> tasks;
    std::vector<MfFin> fins;
    M.small_ptr.assign((size_t)H + 1, 0);
    M.row_ptr.assign((size_t)H + 1, 0);
    M.task_ptr.assign((size_t)H * 4 + 1, 0);
    M.fin_ptr.assign((size_t)H * 4 + 1, 0);
    M.small_fwd_bytes.assign((size_t)H, 0.0);
    M.small_bwd_bytes.assign((size_t)H, 0.0);
    M.task_bytes.assign((size_t)H * 4, 0.0);
    M.max_tasks = 1;
    std::vector<std::vector<int32_t>> lev((size_t)H);
    for (int64_t K = 0; K < nf; K++) lev[(size_t)height[(size_t)K]].push_back((int32_t)K);
    for (int h = 0; h < H; h++) {
        for (int32_t K : lev[(size_t)h])
            if (small[(size_t)K]) {
                slist.push_back(K);
                const double s = P.fronts[(size_t)K].s, b = P.fronts[(size_t)K].f - s;
                M.small_fwd_bytes[(size_t)h] += 16.0 * (s * (s + 1) / 2 + s * b);
                M.small_bwd_bytes[(size_t)h] += 16.0 * (s * (s + 1) / 2 + s * b);
            }
        M.small_ptr[(size_t)h + 1] = (int64_t)slist.size();
        for (int32_t K : lev[(size_t)h])
            if (!small[(size_t)K])
                for (int j = 0; j < P.fronts[(size_t)K].f; j++) rows.push_back(make_int2(K, j));
        M.row_ptr[(size_t)h + 1] = (int64_t)rows.size();
        for (int ph = 0; ph < 4; ph++) {
            const int64_t t0 = (int64_t)tasks.size();
            for (int32_t K : lev[(size_t)h]) {
                if (small[(size_t)K]) continue;
                const LuFrontInfo &F = P.fronts[(size_t)K];
                const int s = F.s, b = F.f - F.s;
                const int nrow = ph == 1 ? b : s;
                for (int r0 = 0; r0 < nrow; r0 += 64) {
                    const int m = std::min(64, nrow - r0);
                    // column range of this row chunk
                    int kbeg = 0, kend = 0;
                    if (ph == 0) kend = r0 + m;            // L11^-1 is lower triangular
                    else if (ph == 1) kend = s;             // L21
                    else if (ph == 2) kend = b;             // U12
                    else kbeg = r0, kend = s;               // U11^-1 is upper triangular
                    MfFin fi{};
                    fi.front = K;
                    fi.r0 = r0;
                    fi.m = m;
                    fi.p0 = (int32_t)((int64_t)tasks.size() - t0);
                    for (int k0 = kbeg; k0 < kend; k0 += kMfKc) {
                        MfTask tk{};
                        const int kc = std::min(kMfKc, kend - k0);
                        if (ph == 0) {
                            tk.a = F.inv + (int64_t)k0 * s + r0;
                            tk.lda = s;
                        } else if (ph == 1) {
                            tk.a = F.off + (int64_t)k0 * F.f + s + r0;
                            tk.lda = F.f;
                        } else if (ph == 2) {
                            tk.a = F.off + (int64_t)(s + k0) * F.f + r0;
                            tk.lda = F.f;
                        } else {
                            tk.a = F.inv + (int64_t)s * s + (int64_t)k0 * s + r0;
                            tk.lda = s;
                        }
                        tk.m = m;
                        tk.kc = kc;
                        tk.front = K;
                        tk.r0 = r0;
                        tk.k0 = ph == 2 ? s + k0 : k0;  // first row of w it multiplies
                        tasks.push_back(tk);
                        M.task_bytes[(size_t)(h * 4 + ph)] += 16.0 * m * kc;
                    }
                    fi.np = (int32_t)((int64_t)tasks.size() - t0) - fi.p0;
                    // the backward t of a front without boundary is c itself:
                    // its finish still runs, with no partials
                    if (fi.np > 0 || ph == 2) fins.push_back(fi);
                }
            }
            M.task_ptr[(size_t)(h * 4 + ph) + 1] = (int64_t)tasks.size();
            M.fin_ptr[(size_t)(h * 4 + ph) + 1] = (int64_t)fins.size();
            M.max_tasks = std::max<int64_t>(M.max_tasks, (int64_t)tasks.size() - t0);
        }
    }
    int rc;
    if ((rc = mf_upload(M.fr, fr)) || (rc = mf_upload(M.small_list, slist)) || (rc = mf_upload(M.bnd, P.bnd)) ||
        (rc = mf_upload(M.rows, rows)) || (rc = mf_upload(M.tasks, tasks)) || (rc = mf_upload(M.fins, fins)) ||
        (rc = mf_upload(M.cptr, cptr)) || (rc = mf_upload(M.csrc, csrc)))
        return rc;
    M.ready = true;
    return PEEC_OK;
}

int mf_prepare(peec_ctx *ctx, int nrhs)
{
    const MfPlan &M = ctx->mf;
    const size_t nw = (size_t)(M.wtot * nrhs), nu = (size_t)(M.utot * nrhs);
    const size_t np = (size_t)M.max_tasks * 64 * (size_t)nrhs;
    if ((nw > ctx->mf_w.n && !ctx->mf_w.alloc(nw)) || (nu > ctx->mf_u.n && !ctx->mf_u.alloc(nu)) ||
        (np > ctx->mf_part.n && !ctx->mf_part.alloc(np))) {
        peec_set_error("out of device memory for the multifrontal solve (%d right-hand sides)", nrhs);
        return PEEC_ERR_NOMEM;
    }
    return PEEC_OK;
}

// lower: L c = b (forward sweep), else U x = b (backward); b and x hold
// nrhs vectors of n positions (factor order), column r at r * n.
int mf_solve(peec_ctx *ctx, bool lower, const double2_t *b, double2_t *x, int nrhs, const ColMap &cm)
{
    if (nrhs <= 0) return PEEC_OK;
    const MfPlan &M = ctx->mf;
    const LuPlan &P = ctx->lu;
    const int nsl = cm.nsl;
    if (!ctx->mf_ready || !M.ready || P.pool_slot0 != 0 || nsl > P.pool_ns) {
        peec_set_error("multifrontal solve: the front pool does not hold the factors of slots 0..%d", nsl - 1);
        return PEEC_ERR_STATE;
    }
    int rc = mf_prepare(ctx, nrhs);
    if (rc) return rc;
    const int cq = mf_cols(nrhs, cm);
    const int Q = cq >= 5 ? 8 : cq >= 2 ? 4 : 1;
    const int64_t stride = P.pool_total, n = P.n;
    const cd *pool = P.pool.p;
    cd *w = ctx->mf_w.p, *u = ctx->mf_u.p, *part = ctx->mf_part.p;
    const hipStream_t st = ctx->stream;
    const unsigned ny = (unsigned)nsl;
    const double vbytes = 16.0 * nrhs;  // one row of every right-hand side
    auto gemv = [&](int h, int ph) {
        const int64_t t0 = M.task_ptr[(size_t)(h * 4 + ph)], nt = M.task_ptr[(size_t)(h * 4 + ph) + 1] - t0;
        if (nt <= 0) return;
        int pi = ctx->prof_begin(PK_MF_GEMV, M.task_bytes[(size_t)(h * 4 + ph)] * nsl + (double)nt * 64 * 2 * vbytes);
        const MfTask *tk = M.tasks.p + t0;
        const dim3 grid((unsigned)nt, ny);
#define MF_GEMV(QQ)                                                                                                  \
    hipLaunchKernelGGL((k_mf_gemv<QQ>), grid, dim3(256), 0, st, tk, M.fr.p, pool, stride, w, M.wtot, part, nrhs, cm)
        if (Q == 8)
            hipLaunchKernelGGL(k_mf_gemv8m, grid, dim3(256), 0, st, tk, M.fr.p, pool, stride, w, M.wtot, part, nrhs, cm);
        else if (Q == 4) MF_GEMV(4);
        else MF_GEMV(1);
#undef MF_GEMV
        ctx->prof_end(pi);
    };
    auto finish = [&](int h, int ph) {
        const int64_t f0 = M.fin_ptr[(size_t)(h * 4 + ph)], nfin = M.fin_ptr[(size_t)(h * 4 + ph) + 1] - f0;
        if (nfin <= 0) return;
        int pi = ctx->prof_begin(PK_MF_AUX, (double)nfin * 64 * 3 * vbytes);
        const MfFin *fi = M.fins.p + f0;
        const dim3 grid((unsigned)nfin, ny);
#define MF_FIN(KK)                                                                                                   \
    hipLaunchKernelGGL(k_mf_finish<KK>, grid, dim3(256), 0, st, fi, M.fr.p, part, b, x, w, M.wtot, u, M.utot, n, \
                       nrhs, cm)
        if (ph == 0) MF_FIN(0);
        else if (ph == 1) MF_FIN(1);
        else if (ph == 2) MF_FIN(2);
        else MF_FIN(3);
#undef MF_FIN
        ctx->prof_end(pi);
    };
    for (int step = 0; step < M.H; step++) {
        const int h = lower ? step : M.H - 1 - step;
        const int64_t s0 = M.small_ptr[(size_t)h], nsm = M.small_ptr[(size_t)h + 1] - s0;
        if (nsm > 0) {
            const dim3 grid((unsigned)nsm, ny);
            const int32_t *ls = M.small_list.p + s0;
            if (lower) {
                int pi = ctx->prof_begin(PK_MF_SMALL, M.small_fwd_bytes[(size_t)h] * nsl);
#define MF_FWD(QQ)                                                                                                   \
    hipLaunchKernelGGL(k_mf_fwd_small<QQ>, grid, dim3(256), 0, st, ls, M.fr.p, pool, stride, M.cptr.p, M.csrc.p, b, x, \
                       u, M.utot, n, nrhs, cm)
                if (Q == 8)
                    hipLaunchKernelGGL(k_mf_fwd_small8m, grid, dim3(256), 0, st, ls, M.fr.p, pool, stride, M.cptr.p,
                                       M.csrc.p, b, x, u, M.utot, n, nrhs, cm);
                else if (Q == 4) MF_FWD(4);
                else MF_FWD(1);
#undef MF_FWD
                ctx->prof_end(pi);
            } else {
                int pi = ctx->prof_begin(PK_MF_SMALL, M.small_bwd_bytes[(size_t)h] * nsl);
#define MF_BWD(QQ)                                                                                                   \
    hipLaunchKernelGGL(k_mf_bwd_small<QQ>, grid, dim3(256), 0, st, ls, M.fr.p, pool, stride, M.bnd.p, b, x, n, nrhs, \
                       cm)
                if (Q == 8)
                    hipLaunchKernelGGL(k_mf_bwd_small8m, grid, dim3(256), 0, st, ls, M.fr.p, pool, stride, M.bnd.p, b,
                                       x, n, nrhs, cm);
                else if (Q == 4) MF_BWD(4);
                else MF_BWD(1);
#undef MF_BWD
                ctx->prof_end(pi);
            }
        }
        const int64_t r0 = M.row_ptr[(size_t)h], nr = M.row_ptr[(size_t)h + 1] - r0;
        if (nr <= 0) continue;
        if (lower) {
            int pi = ctx->prof_begin(PK_MF_AUX, (double)nr * 2 * vbytes);
            hipLaunchKernelGGL(k_mf_assemble, dim3((unsigned)((nr + 255) / 256), ny), dim3(256), 0, st, M.rows.p + r0,
                               nr, M.fr.p, M.cptr.p, M.csrc.p, b, w, M.wtot, u, M.utot, n, nrhs, cm);
            ctx->prof_end(pi);
            gemv(h, 0);    // L11^-1 w_P
            finish(h, 0);  // y -> x and w
            gemv(h, 1);    // L21 y
            finish(h, 1);  // u = w_B - L21 y
        } else {
            int pi = ctx->prof_begin(PK_MF_AUX, (double)nr * 2 * vbytes);
            hipLaunchKernelGGL(k_mf_gather_b, dim3((unsigned)((nr + 255) / 256), ny), dim3(256), 0, st, M.rows.p + r0,
                               nr, M.fr.p, M.bnd.p, x, w, M.wtot, n, nrhs, cm);
            ctx->prof_end(pi);
            gemv(h, 2);    // U12 x_B
            finish(h, 2);  // t = c_P - U12 x_B
            gemv(h, 3);    // U11^-1 t
            finish(h, 3);  // x_P
        }
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}
