// Blocked sparse triangular solves with the LU factors of SS.
//
// The preconditioner solves SS M = r once per matvec (precond_3_3_vector,
// compute_Matrix_vector_fft.jl:84-146, whose three F\ solves the operator
// combines into one), i.e. one forward sweep with L and one backward sweep
// with U.  The factor order is cut into row blocks of at most 64 rows, the
// leaves and separator chunks of a nested dissection when the host gives
// them.  A block depends on the blocks its rows read; blocks with no
// pending dependency form a level, and the levels are solved one after the
// other, each in two launches:
//
//   gather  -- every row of the level subtracts its entries in other
//              (already solved) blocks; long rows are cut into segments of
//              at most kTrsvSeg entries, one wavefront per segment, so a
//              separator row with thousands of entries is spread over many
//              CUs.  Each segment writes one partial sum per right-hand side.
//   block   -- one workgroup per block loads the block's dense s x s
//              triangle into LDS; a wavefront per right-hand side holds one
//              row per lane, sums the row's partials and substitutes column
//              by column, broadcasting each solved value by a lane shuffle.
//
// Kernel boundaries order the levels; there is no waiting between
// workgroups.  The sums are taken in a fixed order, so results are
// reproducible run to run.
#include "context.h"

#include <algorithm>

namespace {

typedef double2_t cd;

__device__ __forceinline__ cd tcmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd tcdiv(cd a, cd b)
{
    double d = b.x * b.x + b.y * b.y;
    return {(a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d};
}

// part[s + r * pstride] = sum over the entries of segment s of T[row, c] * x[c]
// (right-hand side r uses the value set of its slot, ob_n values each)
__global__ void __launch_bounds__(256) k_trsv_gather(const int64_t *seg_k0, const int64_t *seg_k1, int64_t s0,
                                                     int64_t nseg, const int32_t *idx, const cd *val, int64_t ob_n,
                                                     const cd *x, cd *part, int64_t pstride, int64_t n, int nrhs,
                                                     ColMap cm)
{
    // segments of one block's rows read the same columns of x: keep them on
    // one XCD (one L2)
    const int64_t sg = s0 + (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (sg >= s0 + nseg) return;  // whole wavefront: the kernel has no barrier
    const int lane = threadIdx.x & 63;
    const int64_t k0 = seg_k0[sg], k1 = seg_k1[sg];
    // a segment holds at most kTrsvSeg = 4 * 64 entries: 4 per lane
    int c[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t k = k0 + lane + 64 * i;
        c[i] = k < k1 ? idx[k] : -1;
    }
    cd v[4];
    int slot = -1;
    for (int r = 0; r < nrhs; r++) {
        if (cm.slot(r) != slot) {
            slot = cm.slot(r);
            const cd *vs = val + (int64_t)slot * ob_n;
#pragma unroll
            for (int i = 0; i < 4; i++) v[i] = c[i] >= 0 ? vs[k0 + lane + 64 * i] : cd{0.0, 0.0};
        }
        const cd *xr = x + (int64_t)r * n;
        double ax = 0.0, ay = 0.0;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            if (c[i] >= 0) {
                const cd xv = xr[c[i]];
                ax += v[i].x * xv.x - v[i].y * xv.y;
                ay += v[i].x * xv.y + v[i].y * xv.x;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            ax += __shfl_xor(ax, o, 64);
            ay += __shfl_xor(ay, o, 64);
        }
        if (lane == 0) part[sg + (int64_t)r * pstride] = {ax, ay};
    }
}

// The same sums for many right-hand sides, one frequency slot per
// blockIdx.y: the values of the slot are loaded once, then its right-hand
// sides go 4 per pass, with the 16 x loads of a
// pass (4 right-hand sides x 4 entries per lane) all in flight, and the 4
// lane sums reduced together by recursive halving -- exchanges at distance
// 32 and 16 keep 2, then 1 of the sums per lane, then a 16-lane all-reduce
// finishes it: 7 exchanges for 4 complex sums instead of 24.
//
// The loads carry no branches: a lane past the segment's end reads column
// 0 against a zero value, and right-hand sides past the slot's end re-read
// the last one (their sums are not written).  The complex products are
// FMA chains.  The 16-lane all-reduce uses DPP lane moves (row mirror,
// half-row mirror, then quad swaps 1 and 2: after the four steps every lane
// of the row holds the sum of all 16), the two cross-row exchanges
// ds_bpermute.
__device__ __forceinline__ double xsh(double v, int d) { return __shfl_xor(v, d, 64); }

template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(b >> 32), CTRL, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
constexpr int kDppRowMirror = 0x140, kDppRowHalfMirror = 0x141, kDppQuadSwap1 = 0xB1, kDppQuadSwap2 = 0x4E;

__device__ __forceinline__ double row16_sum(double v)
{
    v += dpp_move<kDppRowMirror>(v);
    v += dpp_move<kDppRowHalfMirror>(v);
    v += dpp_move<kDppQuadSwap1>(v);
    v += dpp_move<kDppQuadSwap2>(v);
    return v;
}

__global__ void __launch_bounds__(256) k_trsv_gather4(const int64_t *seg_k0, const int64_t *seg_k1, int64_t s0,
                                                      int64_t nseg, const int32_t *idx, const cd *val, int64_t ob_n,
                                                      const cd *x, cd *part, int64_t pstride, int64_t n, int nrhs,
                                                      ColMap cm)
{
    // segments of one block's rows read the same columns of x: keep them on
    // one XCD (one L2)
    const int64_t sg = s0 + (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (sg >= s0 + nseg) return;  // whole wavefront: the kernel has no barrier
    const int lane = threadIdx.x & 63;
    const int64_t k0 = seg_k0[sg], k1 = seg_k1[sg];
    int c[4];
    bool ok[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int64_t k = k0 + lane + 64 * i;
        ok[i] = k < k1;
        c[i] = ok[i] ? idx[k] : 0;
    }
    const bool h32 = lane & 32, h16 = lane & 16;
    const int gsel = (h32 ? 2 : 0) + (h16 ? 1 : 0);
    {
        // one slot per blockIdx.y: the slots' sums run side by side, so a
        // level with few segments (the separator chain) is not a chain of
        // dependent loads slot after slot
        const int slot = (int)blockIdx.y;
        const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
        if (rbeg >= rend) return;
        cd v[4];
        const cd *vs = val + (int64_t)slot * ob_n;
#pragma unroll
        for (int i = 0; i < 4; i++) v[i] = ok[i] ? vs[k0 + lane + 64 * i] : cd{0.0, 0.0};
        for (int r0 = rbeg; r0 < rend; r0 += 4) {
            cd xv[4][4];
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const cd *xr = x + (int64_t)min(r0 + g, rend - 1) * n;
#pragma unroll
                for (int i = 0; i < 4; i++) xv[g][i] = xr[c[i]];
            }
            double sx[4], sy[4];
#pragma unroll
            for (int g = 0; g < 4; g++) {
                double ax = 0.0, ay = 0.0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    ax = __builtin_fma(v[i].x, xv[g][i].x, ax);
                    ax = __builtin_fma(-v[i].y, xv[g][i].y, ax);
                    ay = __builtin_fma(v[i].x, xv[g][i].y, ay);
                    ay = __builtin_fma(v[i].y, xv[g][i].x, ay);
                }
                sx[g] = ax;
                sy[g] = ay;
            }
            // recursive halving: after the exchange at distance d a lane
            // keeps the half of its sums selected by bit d of its index
            double tx[2], ty[2];
#pragma unroll
            for (int g = 0; g < 2; g++) {
                const double kx = h32 ? sx[g + 2] : sx[g], ky = h32 ? sy[g + 2] : sy[g];
                const double ox = h32 ? sx[g] : sx[g + 2], oy = h32 ? sy[g] : sy[g + 2];
                tx[g] = kx + xsh(ox, 32);
                ty[g] = ky + xsh(oy, 32);
            }
            const double wx = row16_sum((h16 ? tx[1] : tx[0]) + xsh(h16 ? tx[0] : tx[1], 16));
            const double wy = row16_sum((h16 ? ty[1] : ty[0]) + xsh(h16 ? ty[0] : ty[1], 16));
            if ((lane & 15) == 0 && r0 + gsel < rend) part[sg + (int64_t)(r0 + gsel) * pstride] = {wx, wy};
        }
    }
}

// a + b summed over the lane pairs (l, l + 32): lanes 0..31 return the sum
// of a, lanes 32..63 that of b.  v_permlane32_swap exchanges the upper half
// of its first operand with the lower half of its second, so after the swap
// each lane holds both terms of its own sum and no lane selects.
__device__ __forceinline__ double swap32_add(double a, double b)
{
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const auto lo = __builtin_amdgcn_permlane32_swap((uint32_t)ua, (uint32_t)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
    const double na = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0]);
    const double nb = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1]);
    return na + nb;
}

// The same over the lane pairs (l, l + 16) of each 32-lane half:
// v_permlane16_swap exchanges the odd 16-lane rows of its first operand
// with the even rows of its second.  Even rows return the sum of a, odd
// rows that of b.
__device__ __forceinline__ double swap16_add(double a, double b)
{
    const uint64_t ua = __builtin_bit_cast(uint64_t, a), ub = __builtin_bit_cast(uint64_t, b);
    const auto lo = __builtin_amdgcn_permlane16_swap((uint32_t)ua, (uint32_t)ub, false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), false, false);
    const double na = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[0] << 32) | (uint32_t)lo[0]);
    const double nb = __builtin_bit_cast(double, ((uint64_t)(uint32_t)hi[1] << 32) | (uint32_t)lo[1]);
    return na + nb;
}

// Sixteen lane sums reduced together by recursive halving: after the steps
// at lane distance 32, 16, 8 (row mirror) and 4 (half-row mirror) each lane
// keeps one of them, two quad swaps finish it, and lane l holds the total of
// s[(l >> 2) & 15] (so do lanes l ^ 1, l ^ 2, l ^ 3).  24 lane exchanges
// for 16 sums instead of 96.
__device__ __forceinline__ double reduce16(const double (&s)[16], int lane)
{
    double t[8], u[4], w[2];
#pragma unroll
    for (int j = 0; j < 8; j++) t[j] = swap32_add(s[j], s[j + 8]);
#pragma unroll
    for (int j = 0; j < 4; j++) u[j] = swap16_add(t[j], t[j + 4]);
    const bool h8 = lane & 8, h4 = lane & 4;
#pragma unroll
    for (int j = 0; j < 2; j++) w[j] = (h8 ? u[j + 2] : u[j]) + dpp_move<kDppRowMirror>(h8 ? u[j] : u[j + 2]);
    double y = (h4 ? w[1] : w[0]) + dpp_move<kDppRowHalfMirror>(h4 ? w[0] : w[1]);
    y += dpp_move<kDppQuadSwap1>(y);
    y += dpp_move<kDppQuadSwap2>(y);
    return y;
}

// One right-hand side: the complex sum (s[0], s[1]) over the wavefront;
// lanes 0..31 return the real part, lanes 32..63 the imaginary part.
__device__ __forceinline__ double reduce2(const double (&s)[2])
{
    double t = swap32_add(s[0], s[1]);
    t = swap16_add(t, t);
    return row16_sum(t);
}

// Four right-hand sides: the 8 sums s[2g + part] reduced together (halving
// at lane distance 32, 16 and 8, then a row all-reduce); lane l holds the
// total of s[(l >> 3) & 7].
__device__ __forceinline__ double reduce8(const double (&s)[8], int lane)
{
    double t[4], u[2];
#pragma unroll
    for (int j = 0; j < 4; j++) t[j] = swap32_add(s[j], s[j + 4]);
#pragma unroll
    for (int j = 0; j < 2; j++) u[j] = swap16_add(t[j], t[j + 2]);
    const bool h8 = lane & 8;
    double y = (h8 ? u[1] : u[0]) + dpp_move<kDppRowMirror>(h8 ? u[0] : u[1]);
    y += dpp_move<kDppRowHalfMirror>(y);
    y += dpp_move<kDppQuadSwap1>(y);
    y += dpp_move<kDppQuadSwap2>(y);
    return y;
}

// Panel blocks: the rows of a block reach the same columns, so a wavefront
// loads x at one chunk of at most kTrsvPanelCols columns (2 per lane) for R
// right-hand sides of a slot (8, 4 or 1, by the slots' widths) once and sums
// the chunk for every row of its item against it; only the values are read
// per row, and the next row's are in flight while a row is reduced.
// items[w] = (segment of the first row's chunk, rows, segment stride between
// rows, entries per row); row rho's chunk is segment item.x + rho * item.z
// and its values start item.w entries after the previous row's.  Wavefront
// w of the launch takes item w % nitems of slot w / nitems.
template <int R>
__global__ void __launch_bounds__(256) k_trsv_panel(const int4 *items, int64_t nitems, const int64_t *seg_k0,
                                                    const int64_t *seg_k1, const int32_t *idx, const cd *val,
                                                    int64_t ob_n, const cd *x, cd *part, int64_t pstride, int64_t n,
                                                    int nrhs, ColMap cm)
{
    static_assert(R == 8 || R == 4 || R == 1, "8, 4 or 1 right-hand sides per pass");
    // items of one chunk (same columns of x) are adjacent: keep them on one XCD
    const int64_t w = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (w >= nitems * cm.nsl) return;  // whole wavefront: the kernel has no barrier
    const int slot = (int)(w / nitems);
    const int4 it = items[w % nitems];
    const int rbeg = cm.first(slot), rend = cm.end(slot, nrhs);
    if (rbeg >= rend) return;
    const int lane = threadIdx.x & 63;
    const int64_t c0 = seg_k0[it.x];
    const int cnt = (int)(seg_k1[it.x] - c0);
    int c[2];
    bool ok[2];
#pragma unroll
    for (int i = 0; i < 2; i++) {
        ok[i] = lane + 64 * i < cnt;
        c[i] = ok[i] ? idx[c0 + lane + 64 * i] : 0;
    }
    const cd *vs = val + (int64_t)slot * ob_n + c0;
    double *pd = reinterpret_cast<double *>(part);
    // the lanes that hold a total after the reduction, and which one
    constexpr int kShift = R == 8 ? 2 : R == 4 ? 3 : 5;
    const bool writer = (lane & ((1 << kShift) - 1)) == 0;
    const int vi = (lane >> kShift) & (2 * R - 1), gi = vi >> 1;
    for (int r0 = rbeg; r0 < rend; r0 += R) {
        cd xv[R][2];
#pragma unroll
        for (int g = 0; g < R; g++) {
            const cd *xr = x + (int64_t)min(r0 + g, rend - 1) * n;
#pragma unroll
            for (int i = 0; i < 2; i++) xv[g][i] = xr[c[i]];
        }
        cd v[2];
#pragma unroll
        for (int i = 0; i < 2; i++) v[i] = ok[i] ? vs[lane + 64 * i] : cd{0.0, 0.0};
        for (int rho = 0; rho < it.y; rho++) {
            cd vn[2];
            const int64_t kn = (int64_t)min(rho + 1, it.y - 1) * it.w;
#pragma unroll
            for (int i = 0; i < 2; i++) vn[i] = ok[i] ? vs[kn + lane + 64 * i] : cd{0.0, 0.0};
            double s[2 * R];
#pragma unroll
            for (int g = 0; g < R; g++) {
                double ax = 0.0, ay = 0.0;
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    ax = __builtin_fma(v[i].x, xv[g][i].x, ax);
                    ax = __builtin_fma(-v[i].y, xv[g][i].y, ax);
                    ay = __builtin_fma(v[i].x, xv[g][i].y, ay);
                    ay = __builtin_fma(v[i].y, xv[g][i].x, ay);
                }
                s[2 * g] = ax;
                s[2 * g + 1] = ay;
            }
            double tot;
            if constexpr (R == 8)
                tot = reduce16(s, lane);
            else if constexpr (R == 4)
                tot = reduce8(s, lane);
            else
                tot = reduce2(s);
            const int64_t sg = it.x + (int64_t)rho * it.z;
            if (writer && r0 + gi < rend) pd[2 * (sg + (int64_t)(r0 + gi) * pstride) + (vi & 1)] = tot;
#pragma unroll
            for (int i = 0; i < 2; i++) v[i] = vn[i];
        }
    }
}

// Value of v in lane l (l uniform across the wavefront).
__device__ __forceinline__ double lane_value(double v, int l)
{
    const uint64_t b = __builtin_bit_cast(uint64_t, v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)b, l);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(b >> 32), l);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// One workgroup per (block of the level, value set, group of right-hand
// sides): binfo lists the level's blocks in launch order; blockIdx.y =
// slot * gps + g, and the workgroup solves the right-hand sides
// first + g * 4 + wave (stride 4 gps) of that slot.  inv_diag holds
// 1 / T[i, i].
template <bool LOWER>
__global__ void __launch_bounds__(256) k_trsv_block(const int4 *binfo, const cd *dense, int64_t dense_n,
                                                    const cd *inv_diag, const int2 *row_seg, const cd *part,
                                                    int64_t pstride, const cd *b, cd *x, int64_t n, int nrhs, ColMap cm,
                                                    int gps)
{
    __shared__ cd tile[64 * 64];
    const int4 bi = binfo[blockIdx.x];  // r0, s, dense offset (lo, hi)
    const int r0 = bi.x, s = bi.y;
    const int slot = (int)blockIdx.y / gps, g = (int)blockIdx.y % gps;
    const cd *T = dense + (int64_t)slot * dense_n + (int64_t)(((uint64_t)(uint32_t)bi.w << 32) | (uint32_t)bi.z);
    inv_diag += (int64_t)slot * n;
    const int ss = s * s;
    {
        // all 16 loads in flight before the first LDS store
        cd t[16];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int e = threadIdx.x + 256 * i;
            t[i] = e < ss ? T[e] : cd{0.0, 0.0};
        }
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int e = threadIdx.x + 256 * i;
            if (e < ss) tile[e] = t[i];
        }
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool mine = lane < s;
    const int row = r0 + (mine ? lane : 0);
    const cd dinv = mine ? inv_diag[row] : cd{1.0, 0.0};
    const int2 gs = mine ? row_seg[row] : int2{0, 0};
    const int g0 = gs.x, g1 = gs.y;
    __syncthreads();
    const int rend = cm.end(slot, nrhs);
    for (int r = cm.first(slot) + g * 4 + wave; r < rend; r += gps * 4) {
        cd acc = {0.0, 0.0};
        if (mine) {
            acc = b[row + (int64_t)r * n];
            const cd *pr = part + (int64_t)r * pstride;
            int g = g0;
            for (; g + 4 <= g1; g += 4) {
                const cd p0 = pr[g], p1 = pr[g + 1], p2 = pr[g + 2], p3 = pr[g + 3];
                acc.x -= ((p0.x + p1.x) + (p2.x + p3.x));
                acc.y -= ((p0.y + p1.y) + (p2.y + p3.y));
            }
            for (; g < g1; g++) {
                acc.x -= pr[g].x;
                acc.y -= pr[g].y;
            }
        }
        for (int q = 0; q < s; q++) {
            const int j = LOWER ? q : s - 1 - q;
            const cd t = tile[j * s + lane < 64 * 64 ? j * s + lane : 0];
            const double ax = lane_value(acc.x, j), ay = lane_value(acc.y, j);
            const double dx = lane_value(dinv.x, j), dy = lane_value(dinv.y, j);
            const double yx = ax * dx - ay * dy, yy = ax * dy + ay * dx;
            if (mine && (LOWER ? lane > j : lane < j)) {
                acc.x -= t.x * yx - t.y * yy;
                acc.y -= t.x * yy + t.y * yx;
            }
        }
        if (mine) x[row + (int64_t)r * n] = {acc.x * dinv.x - acc.y * dinv.y, acc.x * dinv.y + acc.y * dinv.x};
    }
}

}  // namespace

// Blocks when the caller gives none: chains of the elimination tree of L
// (column j + 1 is the parent of column j), at most `cap` rows each.
std::vector<int32_t> tri_chain_blocks(const peec_csc_c128 *L, int cap)
{
    const int64_t n = L->ncol;
    std::vector<int32_t> blk{0};
    for (int64_t j = 0; j < n; j++) {
        int64_t parent = n;
        for (int64_t k = L->colptr[j] - 1; k < L->colptr[j + 1] - 1; k++) {
            int64_t r = L->rowval[k] - 1;
            if (r > j && r < parent && (L->nzval[k].re != 0.0 || L->nzval[k].im != 0.0)) parent = r;
        }
        bool cont = parent == j + 1 && (j + 1 - blk.back()) < cap;
        if (!cont) blk.push_back((int32_t)(j + 1));
    }
    if (blk.back() != n) blk.push_back((int32_t)n);
    return blk;
}

int tri_layout(int64_t n, bool lower, const std::vector<int32_t> &blk, const std::vector<TriEntry> &ent,
               const std::vector<int64_t> &diag_src, TriFactor *f, TriSources *src, const char *what)
{
    const int nb = (int)blk.size() - 1;
    if (nb < 0 || blk.front() != 0 || blk.back() != n) {
        peec_set_error("%s: blocks do not cover the factor", what);
        return PEEC_ERR_ARG;
    }
    std::vector<int32_t> bof((size_t)n);
    std::vector<int64_t> doff((size_t)nb + 1, 0);
    for (int B = 0; B < nb; B++) {
        const int s = blk[(size_t)B + 1] - blk[(size_t)B];
        if (s < 1 || s > kTrsvMaxBlock) {
            peec_set_error("%s: block %d has %d rows (1..%d allowed)", what, B, s, kTrsvMaxBlock);
            return PEEC_ERR_ARG;
        }
        for (int i = blk[(size_t)B]; i < blk[(size_t)B + 1]; i++) bof[(size_t)i] = B;
        doff[(size_t)B + 1] = doff[(size_t)B] + (int64_t)s * s;
    }
    src->dense.assign((size_t)doff[(size_t)nb], -1);
    src->diag = diag_src;
    std::vector<int64_t> optr((size_t)n + 1, 0);
    std::vector<double> ibn((size_t)std::max(nb, 1), 0.0);
    for (const TriEntry &e : ent) {
        if (e.r < 0 || e.r >= n || e.c < 0 || e.c >= n || (lower ? e.r <= e.c : e.r >= e.c)) {
            peec_set_error("%s: entry (%d, %d) outside the %s triangle", what, e.r + 1, e.c + 1, lower ? "lower" : "upper");
            return PEEC_ERR_ARG;
        }
        const int Br = bof[(size_t)e.r], Bc = bof[(size_t)e.c];
        if (Br == Bc) {
            const int r0 = blk[(size_t)Br], s = blk[(size_t)Br + 1] - r0;
            src->dense[(size_t)(doff[(size_t)Br] + (int64_t)(e.c - r0) * s + (e.r - r0))] = e.src;
            ibn[(size_t)Br] += 1.0;
        } else {
            optr[(size_t)e.r + 1]++;
        }
    }
    for (int64_t i = 0; i < n; i++) optr[(size_t)i + 1] += optr[(size_t)i];
    const int64_t nob = optr[(size_t)n];
    std::vector<int32_t> oidx((size_t)std::max<int64_t>(nob, 1), 0);
    src->ob.assign(oidx.size(), -1);
    {
        // entries of a row in increasing column order
        std::vector<int64_t> opos(optr.begin(), optr.end() - 1);
        for (const TriEntry &e : ent) {
            if (bof[(size_t)e.r] == bof[(size_t)e.c]) continue;
            const int64_t p = opos[(size_t)e.r]++;
            oidx[(size_t)p] = e.c;
            src->ob[(size_t)p] = e.src;
        }
        for (int64_t i = 0; i < n; i++) {
            const int64_t a = optr[(size_t)i], b = optr[(size_t)i + 1];
            if (b - a < 2) continue;
            std::vector<std::pair<int32_t, int64_t>> row((size_t)(b - a));
            for (int64_t k = a; k < b; k++) row[(size_t)(k - a)] = {oidx[(size_t)k], src->ob[(size_t)k]};
            std::sort(row.begin(), row.end());
            for (int64_t k = a; k < b; k++) {
                oidx[(size_t)k] = row[(size_t)(k - a)].first;
                src->ob[(size_t)k] = row[(size_t)(k - a)].second;
            }
        }
    }
    // levels in solve order: a block follows every block its rows read
    std::vector<int32_t> level((size_t)std::max(nb, 1), 0);
    for (int step = 0; step < nb; step++) {
        const int B = lower ? step : nb - 1 - step;
        int lv = 0;
        for (int i = blk[(size_t)B]; i < blk[(size_t)B + 1]; i++)
            for (int64_t k = optr[(size_t)i]; k < optr[(size_t)i + 1]; k++)
                lv = std::max(lv, level[(size_t)bof[(size_t)oidx[(size_t)k]]] + 1);
        level[(size_t)B] = lv;
    }
    int nlev = 0;
    for (int B = 0; B < nb; B++) nlev = std::max(nlev, level[(size_t)B] + 1);
    std::vector<int32_t> lptr((size_t)nlev + 1, 0), lblk((size_t)std::max(nb, 1));
    for (int B = 0; B < nb; B++) lptr[(size_t)level[(size_t)B] + 1]++;
    for (int l = 0; l < nlev; l++) lptr[(size_t)l + 1] += lptr[(size_t)l];
    {
        std::vector<int32_t> lpos(lptr.begin(), lptr.end() - 1);
        for (int B = 0; B < nb; B++) lblk[(size_t)lpos[(size_t)level[(size_t)B]]++] = B;
    }
    // panel blocks: two or more rows, all with the same off-block columns
    std::vector<char> panel((size_t)std::max(nb, 1), 0);
    for (int B = 0; B < nb; B++) {
        const int r0 = blk[(size_t)B], r1 = blk[(size_t)B + 1];
        const int64_t K = optr[(size_t)r0 + 1] - optr[(size_t)r0];
        bool same = r1 - r0 >= 2 && K > 0;
        for (int i = r0 + 1; same && i < r1; i++)
            same = optr[(size_t)i + 1] - optr[(size_t)i] == K &&
                   std::equal(oidx.begin() + optr[(size_t)i], oidx.begin() + optr[(size_t)i] + K,
                              oidx.begin() + optr[(size_t)r0]);
        panel[(size_t)B] = same;
    }
    f->lev_ob.assign((size_t)nlev, 0.0);
    f->lev_pob.assign((size_t)nlev, 0.0);
    f->lev_ib.assign((size_t)nlev, 0.0);
    f->lev_rows.assign((size_t)nlev, 0.0);
    // segments, numbered level by level: in each level first the rows
    // outside panel blocks, then the panel blocks row by row
    std::vector<int64_t> sk0, sk1, slev((size_t)nlev + 1, 0), snp((size_t)nlev, 0), ilev((size_t)nlev + 1, 0);
    std::vector<int32_t> rs((size_t)n * 2, 0), items;
    std::vector<int32_t> binfo((size_t)std::max(nb, 1) * 4, 0);
    for (int l = 0; l < nlev; l++) {
        for (int e = lptr[(size_t)l]; e < lptr[(size_t)l + 1]; e++) {
            const int B = lblk[(size_t)e];
            binfo[(size_t)e * 4 + 0] = blk[(size_t)B];
            binfo[(size_t)e * 4 + 1] = blk[(size_t)B + 1] - blk[(size_t)B];
            binfo[(size_t)e * 4 + 2] = (int32_t)(uint32_t)(uint64_t)doff[(size_t)B];
            binfo[(size_t)e * 4 + 3] = (int32_t)(uint32_t)((uint64_t)doff[(size_t)B] >> 32);
            f->lev_ib[(size_t)l] += ibn[(size_t)B];
            f->lev_rows[(size_t)l] += blk[(size_t)B + 1] - blk[(size_t)B];
            if (panel[(size_t)B]) continue;
            for (int i = blk[(size_t)B]; i < blk[(size_t)B + 1]; i++) {
                f->lev_ob[(size_t)l] += (double)(optr[(size_t)i + 1] - optr[(size_t)i]);
                rs[(size_t)i * 2] = (int32_t)sk0.size();
                for (int64_t k = optr[(size_t)i]; k < optr[(size_t)i + 1]; k += kTrsvSeg) {
                    sk0.push_back(k);
                    sk1.push_back(std::min<int64_t>(k + kTrsvSeg, optr[(size_t)i + 1]));
                }
                rs[(size_t)i * 2 + 1] = (int32_t)sk0.size();
            }
        }
        snp[(size_t)l] = (int64_t)sk0.size();
        for (int e = lptr[(size_t)l]; e < lptr[(size_t)l + 1]; e++) {
            const int B = lblk[(size_t)e];
            if (!panel[(size_t)B]) continue;
            const int r0 = blk[(size_t)B], s = blk[(size_t)B + 1] - r0;
            const int64_t K = optr[(size_t)r0 + 1] - optr[(size_t)r0];
            const int64_t nch = (K + kTrsvPanelCols - 1) / kTrsvPanelCols;
            const int64_t base = (int64_t)sk0.size();
            f->lev_pob[(size_t)l] += (double)K * s;
            for (int i = r0; i < r0 + s; i++) {
                rs[(size_t)i * 2] = (int32_t)sk0.size();
                for (int64_t k = optr[(size_t)i]; k < optr[(size_t)i + 1]; k += kTrsvPanelCols) {
                    sk0.push_back(k);
                    sk1.push_back(std::min<int64_t>(k + kTrsvPanelCols, optr[(size_t)i + 1]));
                }
                rs[(size_t)i * 2 + 1] = (int32_t)sk0.size();
            }
            for (int64_t c = 0; c < nch; c++)
                for (int g = 0; g < s; g += kTrsvPanelRows) {
                    items.push_back((int32_t)(base + g * nch + c));
                    items.push_back(std::min(kTrsvPanelRows, s - g));
                    items.push_back((int32_t)nch);
                    items.push_back((int32_t)K);
                }
        }
        slev[(size_t)l + 1] = (int64_t)sk0.size();
        ilev[(size_t)l + 1] = (int64_t)(items.size() / 4);
    }
    if (sk0.size() > (size_t)INT32_MAX) {
        peec_set_error("%s: too many segments", what);
        return PEEC_ERR_ARG;
    }
    f->n = n;
    f->nblocks = nb;
    f->level_ptr = lptr;
    f->seg_level_ptr = slev;
    f->seg_level_np = snp;
    f->item_level_ptr = ilev;
    f->nitems = (int64_t)(items.size() / 4);
    f->nseg = (int64_t)sk0.size();
    f->max_block = 0;
    for (int B = 0; B < nb; B++) f->max_block = std::max(f->max_block, blk[(size_t)B + 1] - blk[(size_t)B]);
    f->nnz = (double)ent.size() + (double)n;
    if (sk0.empty()) {
        sk0.push_back(0);
        sk1.push_back(0);
    }
    if (items.empty()) items.assign(4, 0);
    f->nslots = 1;
    f->ob_n = (int64_t)oidx.size();
    f->dense_n = (int64_t)std::max<size_t>(src->dense.size(), 1);
    if (!f->binfo.alloc(binfo.size()) || !f->seg_k0.alloc(sk0.size()) || !f->seg_k1.alloc(sk1.size()) ||
        !f->row_seg.alloc(rs.size()) || !f->ob_idx.alloc(oidx.size()) || !f->ob_val.alloc((size_t)f->ob_n) ||
        !f->dense.alloc((size_t)f->dense_n) || !f->diag.alloc((size_t)std::max<int64_t>(n, 1)) ||
        !f->pitem.alloc(items.size())) {
        peec_set_error("%s: out of device memory", what);
        return PEEC_ERR_NOMEM;
    }
    PEEC_HIP_TRY(hipMemcpy(f->binfo.p, binfo.data(), binfo.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->pitem.p, items.data(), items.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->seg_k0.p, sk0.data(), sk0.size() * 8, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->seg_k1.p, sk1.data(), sk1.size() * 8, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->row_seg.p, rs.data(), rs.size() * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->ob_idx.p, oidx.data(), oidx.size() * 4, hipMemcpyHostToDevice));
    return PEEC_OK;
}

int tri_set_slots(TriFactor *f, int nslots)
{
    if (nslots == f->nslots && f->ob_val.p) return PEEC_OK;
    if (!f->ob_val.alloc((size_t)(f->ob_n * nslots)) || !f->dense.alloc((size_t)(f->dense_n * nslots)) ||
        !f->diag.alloc((size_t)(std::max<int64_t>(f->n, 1) * nslots))) {
        peec_set_error("out of device memory for %d factor value sets", nslots);
        return PEEC_ERR_NOMEM;
    }
    f->nslots = nslots;
    return PEEC_OK;
}

int tri_build(const peec_csc_c128 *T, bool lower, const std::vector<int32_t> &blk, TriFactor *f, const char *what)
{
    const int64_t n = T->ncol;
    if (T->nrow != n) {
        peec_set_error("lu_set_factors: %s is not square", what);
        return PEEC_ERR_ARG;
    }
    std::vector<TriEntry> ent;
    std::vector<int64_t> dsrc((size_t)n, -1);
    for (int64_t c = 0; c < n; c++)
        for (int64_t k = T->colptr[c] - 1; k < T->colptr[c + 1] - 1; k++) {
            const int64_t r = T->rowval[k] - 1;
            if (r < 0 || r >= n) {
                peec_set_error("lu_set_factors: %s row index out of range", what);
                return PEEC_ERR_ARG;
            }
            if (r == c) {
                dsrc[(size_t)c] = k;
                continue;
            }
            if (lower ? r < c : r > c) {
                if (T->nzval[k].re != 0.0 || T->nzval[k].im != 0.0) {
                    peec_set_error("lu_set_factors: %s is not %s triangular", what, lower ? "lower" : "upper");
                    return PEEC_ERR_ARG;
                }
                continue;
            }
            ent.push_back({(int32_t)r, (int32_t)c, k});
        }
    std::vector<double2_t> diag((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const int64_t k = dsrc[(size_t)i];
        const double2_t d = k >= 0 ? double2_t{T->nzval[k].re, T->nzval[k].im} : double2_t{0.0, 0.0};
        if (d.x == 0.0 && d.y == 0.0) {
            peec_set_error("lu_set_factors: %s has a zero diagonal at %lld", what, (long long)i + 1);
            return PEEC_ERR_NUMERIC;
        }
        const double m2 = d.x * d.x + d.y * d.y;
        diag[(size_t)i] = {d.x / m2, -d.y / m2};  // stored as 1 / d
    }
    TriSources src;
    int rc = tri_layout(n, lower, blk, ent, dsrc, f, &src, "lu_set_factors");
    if (rc) return rc;
    std::vector<double2_t> oval(src.ob.size()), dense(std::max<size_t>(src.dense.size(), 1), double2_t{0.0, 0.0});
    for (size_t k = 0; k < src.ob.size(); k++) {
        const int64_t s = src.ob[k];
        oval[k] = s >= 0 ? double2_t{T->nzval[s].re, T->nzval[s].im} : double2_t{0.0, 0.0};
    }
    for (size_t e = 0; e < src.dense.size(); e++) {
        const int64_t s = src.dense[e];
        if (s >= 0) dense[e] = {T->nzval[s].re, T->nzval[s].im};
    }
    PEEC_HIP_TRY(hipMemcpy(f->ob_val.p, oval.data(), oval.size() * 16, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->dense.p, dense.data(), dense.size() * 16, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(f->diag.p, diag.data(), (size_t)n * 16, hipMemcpyHostToDevice));
    return PEEC_OK;
}

int trsv_solve(peec_ctx *ctx, const TriFactor &f, bool lower, const double2_t *b, double2_t *x, int nrhs,
               const ColMap &cm)
{
    const int64_t pstride = std::max<int64_t>(f.nseg, 1);
    if ((size_t)(pstride * nrhs) > ctx->trsv_part.n) {
        if (!ctx->trsv_part.alloc((size_t)(pstride * nrhs))) {
            peec_set_error("out of device memory for triangular-solve partial sums");
            return PEEC_ERR_NOMEM;
        }
    }
    const int nsl = cm.nsl;
    if (nsl > f.nslots) {
        peec_set_error("triangular solve: %d right-hand sides need %d factor slots, %d set", nrhs, nsl, f.nslots);
        return PEEC_ERR_STATE;
    }
    cd *part = ctx->trsv_part.p;
    // columns of the widest slot; workgroups per slot: enough for them, 4
    // per workgroup
    const int cols = cm.rslot ? (cm.wmax > 0 ? cm.wmax : (nrhs + nsl - 1) / nsl) : cm.rps;
    const int gps = std::min((cols + 3) / 4, 16);
    const int nlev = (int)f.level_ptr.size() - 1;
    // with fewer than 4 right-hand sides a panel wavefront has too little
    // work per row to cover the loads: every segment of the level (the
    // panel segments follow the others) goes through the per-row gather
    const bool per_row = nrhs < 4;
    for (int l = 0; l < nlev; l++) {
        const int64_t s0 = f.seg_level_ptr[(size_t)l];
        const int64_t ns = (per_row ? f.seg_level_ptr[(size_t)l + 1] : f.seg_level_np[(size_t)l]) - s0;
        const int64_t nps = f.seg_level_ptr[(size_t)l + 1] - f.seg_level_np[(size_t)l];
        const int64_t i0 = f.item_level_ptr[(size_t)l], ni = per_row ? 0 : f.item_level_ptr[(size_t)l + 1] - i0;
        // algorithmic bytes of a gather: its entries (value + index) and the
        // partial sums written; of a block solve: the in-block entries, the
        // diagonal, b and x, and the partial sums read
        if (ni > 0) {
            int pi = ctx->prof_begin(PK_TRSV_PANEL, f.lev_pob[(size_t)l] * (4.0 + 16.0 * nsl) + (double)nps * 16.0 * nrhs);
            const int64_t waves = ni * nsl;
            const int4 *it = reinterpret_cast<const int4 *>(f.pitem.p) + i0;
            // right-hand sides per pass: a slot's columns in one or two passes
            if (cols >= 5)
                hipLaunchKernelGGL(k_trsv_panel<8>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, ctx->stream, it,
                                   ni, f.seg_k0.p, f.seg_k1.p, f.ob_idx.p, f.ob_val.p, f.ob_n, (const cd *)x, part,
                                   pstride, f.n, nrhs, cm);
            else if (cols >= 2)
                hipLaunchKernelGGL(k_trsv_panel<4>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, ctx->stream, it,
                                   ni, f.seg_k0.p, f.seg_k1.p, f.ob_idx.p, f.ob_val.p, f.ob_n, (const cd *)x, part,
                                   pstride, f.n, nrhs, cm);
            else
                hipLaunchKernelGGL(k_trsv_panel<1>, dim3((unsigned)((waves + 3) / 4)), dim3(256), 0, ctx->stream, it,
                                   ni, f.seg_k0.p, f.seg_k1.p, f.ob_idx.p, f.ob_val.p, f.ob_n, (const cd *)x, part,
                                   pstride, f.n, nrhs, cm);
            ctx->prof_end(pi);
        }
        if (ns > 0) {
            const double ent = f.lev_ob[(size_t)l] + (per_row ? f.lev_pob[(size_t)l] : 0.0);
            int pi = ctx->prof_begin(PK_TRSV_GATHER, ent * (4.0 + 16.0 * nsl) + (double)ns * 16.0 * nrhs);
            if (nrhs >= 4)
                hipLaunchKernelGGL(k_trsv_gather4, dim3((unsigned)((ns + 3) / 4), (unsigned)nsl), dim3(256), 0, ctx->stream,
                                   f.seg_k0.p, f.seg_k1.p, s0, ns, f.ob_idx.p, f.ob_val.p, f.ob_n, (const cd *)x, part,
                                   pstride, f.n, nrhs, cm);
            else
                hipLaunchKernelGGL(k_trsv_gather, dim3((unsigned)((ns + 3) / 4)), dim3(256), 0, ctx->stream,
                                   f.seg_k0.p, f.seg_k1.p, s0, ns, f.ob_idx.p, f.ob_val.p, f.ob_n, (const cd *)x,
                                   part, pstride, f.n, nrhs, cm);
            ctx->prof_end(pi);
        }
        int pi = ctx->prof_begin(lower ? PK_TRSV_BLOCK_L : PK_TRSV_BLOCK_U,
                                 (f.lev_ib[(size_t)l] + f.lev_rows[(size_t)l]) * 16.0 * nsl +
                                     f.lev_rows[(size_t)l] * 32.0 * nrhs +
                                     (double)(f.seg_level_ptr[(size_t)l + 1] - s0) * 16.0 * nrhs);
        const int nbl = f.level_ptr[(size_t)l + 1] - f.level_ptr[(size_t)l];
        const int4 *bi = reinterpret_cast<const int4 *>(f.binfo.p) + f.level_ptr[(size_t)l];
        const int2 *rsg = reinterpret_cast<const int2 *>(f.row_seg.p);
        const dim3 grid(nbl, nsl * gps);
        if (lower)
            hipLaunchKernelGGL(k_trsv_block<true>, grid, dim3(256), 0, ctx->stream, bi, f.dense.p, f.dense_n, f.diag.p,
                               rsg, (const cd *)part, pstride, b, x, f.n, nrhs, cm, gps);
        else
            hipLaunchKernelGGL(k_trsv_block<false>, grid, dim3(256), 0, ctx->stream, bi, f.dense.p, f.dense_n,
                               f.diag.p, rsg, (const cd *)part, pstride, b, x, f.n, nrhs, cm, gps);
        ctx->prof_end(pi);
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}
