// Numeric LU of SS on the GPU (multifrontal, no pivoting).
//
// The reference factors SS = Yle + A' Z^-1 A + j w Gamma P^-1 Gamma' with
// UMFPACK on the host at every frequency (FFT_solver_QS_S_type.jl:105-119).
// Its sparsity pattern does not depend on the frequency, so the symbolic
// work is done once per geometry (peec_lu_analyze) from a nested-dissection
// assembly tree given by the caller: every front of the tree (a leaf of the
// dissection, a separator, or the port/lumped-element nodes) becomes a
// dense f x f matrix in one device pool -- its s pivot rows/columns
// followed by its boundary, the positions of later fronts its pivots
// couple to.  Each frequency (peec_lu_factor) then runs entirely on the
// device:
//
//   1. SS entries are computed from 1/DZ, the frequency-independent
//      Gamma P^-1 Gamma' terms and Yle, and written straight into the front
//      that owns them (each stored entry has one target);
//   2. fronts are processed level by level from the leaves: each child's
//      update matrix is added into its parent (one launch per child rank,
//      so no two writers meet), then the front is factored in panels of 64
//      pivots -- the 64 x 64 diagonal block by one workgroup in LDS, which
//      also forms its two triangular inverses; the panel's L and U blocks
//      as GEMMs with those inverses; the trailing update as 64 x 64 GEMM
//      tiles, blocked over groups of kLuGroup pivots -- so the dense work of
//      every level is spread over all CUs;
//   3. the factors are gathered into the blocked layout of the triangular
//      solves (trsv.hip).
//
// The elimination order is the caller's, with no pivoting, as in the host
// path (SuperLU with diag_pivot_thresh = 0): the factors agree with the
// host ones to rounding.  FP64 arithmetic on the VALU throughout.
#include "context.h"

#include <algorithm>
#include <unordered_map>

namespace {

typedef double2_t cd;

__device__ __forceinline__ cd lcmul(cd a, cd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cd lcinv(cd a)
{
    const double d = a.x * a.x + a.y * a.y;
    return {a.x / d, -a.y / d};
}

// The element-wise kernels below loop with a grid stride: their counts (the
// front pool holds more than 2^32 entries at 192 x 192 x 48) can exceed
// the 2^32 work-items one dispatch dimension addresses, so they launch at
// most kMaxBlocks1d blocks (grid1d) and each thread walks its share.
__device__ __forceinline__ int64_t gtid() { return (int64_t)blockIdx.x * blockDim.x + threadIdx.x; }
__device__ __forceinline__ int64_t gstride() { return (int64_t)gridDim.x * blockDim.x; }

// The factorization kernels run for several frequency slots at once:
// blockIdx.y selects the slot, whose fronts live at pool + y * stride.
__global__ void k_lu_zero(cd *p, int64_t n, int64_t stride)
{
    cd *q = p + (int64_t)blockIdx.y * stride;
    for (int64_t i = gtid(); i < n; i += gstride()) q[i] = {0.0, 0.0};
}

// SS entry k = sum_t coef_t / DZ[edge_t] + j w gsum_k, written to its front
// (slot slot0 + y: its 1/DZ at invZ + (slot0 + y) m, its w at wv[slot0 + y])
__global__ void k_lu_ss(const int64_t *tgt, const int32_t *cptr, const int32_t *cedge, const double *ccoef,
                        const double *gsum, const cd *invZ0, const double *wv, int slot0, int64_t m, cd *pool0,
                        int64_t nss, int64_t stride)
{
    const int slot = slot0 + (int)blockIdx.y;
    const cd *invZ = invZ0 + (int64_t)slot * m;
    const double w = wv[slot];
    cd *pool = pool0 + (int64_t)blockIdx.y * stride;
    for (int64_t k = gtid(); k < nss; k += gstride()) {
        double vx = 0.0, vy = 0.0;
        for (int32_t t = cptr[k]; t < cptr[k + 1]; t++) {
            const cd z = invZ[cedge[t]];
            vx += ccoef[t] * z.x;
            vy += ccoef[t] * z.y;
        }
        vy += w * gsum[k];
        pool[tgt[k]] = {vx, vy};
    }
}

// pool[tgt[i]] += val[i] (Yle entries; each target once per slot; tgt
// includes the slot's pool offset)
__global__ void k_lu_add(const int64_t *tgt, const cd *val, int64_t cnt, cd *pool)
{
    for (int64_t i = gtid(); i < cnt; i += gstride()) {
        const cd a = pool[tgt[i]], b = val[i];
        pool[tgt[i]] = {a.x + b.x, a.y + b.y};
    }
}

// Column j of child c's update matrix added into its parent.
__global__ void __launch_bounds__(256) k_lu_extend_add(const int2 *items, const LuFrontInfo *fr, const int64_t *rel_off,
                                                       const int32_t *rel, cd *pool0, int64_t stride)
{
    cd *pool = pool0 + (int64_t)blockIdx.y * stride;
    const int2 it = items[blockIdx.x];
    const LuFrontInfo C = fr[it.x];
    const LuFrontInfo P = fr[C.parent];
    const int b = C.f - C.s;
    const int32_t *r = rel + rel_off[it.x];
    const cd *src = pool + C.off + (int64_t)(C.s + it.y) * C.f + C.s;
    cd *dst = pool + P.off + (int64_t)r[it.y] * P.f;
    for (int i = threadIdx.x; i < b; i += 256) {
        const cd a = dst[r[i]], s = src[i];
        dst[r[i]] = {a.x + s.x, a.y + s.y};
    }
}

// LU (no pivoting) of the w x w diagonal block at pivot k0 of a front, in
// LDS; the factors go back to the front, the inverses of U and of the unit
// lower L to the item's scratch slot (64 x 64 each, column-major) and to the
// diagonal blocks of the front's pivot-block inverses.
__global__ void __launch_bounds__(256) k_lu_diag(const int4 *items, const LuFrontInfo *fr, cd *pool0, int64_t scratch0,
                                                 int64_t stride)
{
    __shared__ cd A[64 * 64];
    __shared__ cd X[64 * 64];
    cd *pool = pool0 + (int64_t)blockIdx.y * stride;
    const int4 it = items[blockIdx.x];
    const LuFrontInfo F = fr[it.x];
    const int k0 = it.z, w = min(64, F.s - k0), ld = F.f;
    const int tid = threadIdx.x;
    cd *base = pool + F.off + (int64_t)k0 * ld + k0;
    for (int e = tid; e < w * w; e += 256) {
        const int i = e % w, j = e / w;
        A[j * 64 + i] = base[(int64_t)j * ld + i];
    }
    __syncthreads();
    for (int k = 0; k < w; k++) {
        const cd inv = lcinv(A[k * 64 + k]);
        for (int i = k + 1 + tid; i < w; i += 256) A[k * 64 + i] = lcmul(A[k * 64 + i], inv);
        __syncthreads();
        const int r = w - k - 1;
        for (int e = tid; e < r * r; e += 256) {
            const int i = k + 1 + e % r, j = k + 1 + e / r;
            const cd t = lcmul(A[k * 64 + i], A[j * 64 + k]);
            A[j * 64 + i].x -= t.x;
            A[j * 64 + i].y -= t.y;
        }
        __syncthreads();
    }
    for (int e = tid; e < w * w; e += 256) {
        const int i = e % w, j = e / w;
        base[(int64_t)j * ld + i] = A[j * 64 + i];
    }
    cd *Ui = pool + scratch0 + (int64_t)it.y * 8192;
    cd *Li = Ui + 4096;
    // X = U^-1, row by row from the bottom: X[i][j] = -(sum_{k=i+1..j} U[i][k] X[k][j]) / U[i][i]
    __shared__ cd red[4][64];
    const int j = tid & 63, q = tid >> 6;
    for (int e = tid; e < 64 * 64; e += 256) X[e] = {0.0, 0.0};
    __syncthreads();
    for (int i = w - 1; i >= 0; i--) {
        cd s = {0.0, 0.0};
        if (j > i && j < w)
            for (int k = i + 1 + q; k <= j; k += 4) {
                const cd t = lcmul(A[k * 64 + i], X[j * 64 + k]);
                s.x += t.x;
                s.y += t.y;
            }
        red[q][j] = s;
        __syncthreads();
        if (q == 0 && j < w) {
            const cd di = lcinv(A[i * 64 + i]);
            if (j == i) {
                X[i * 64 + i] = di;
            } else if (j > i) {
                const cd t = {((red[0][j].x + red[1][j].x) + (red[2][j].x + red[3][j].x)),
                              ((red[0][j].y + red[1][j].y) + (red[2][j].y + red[3][j].y))};
                const cd v = lcmul(t, di);
                X[j * 64 + i] = {-v.x, -v.y};
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < 64 * 64; e += 256) Ui[e] = X[e];
    // the w x w block also goes to the front's U11^-1 (diagonal block at k0)
    cd *Uinv = pool + F.inv + (int64_t)F.s * F.s + (int64_t)k0 * F.s + k0;
    for (int e = tid; e < w * w; e += 256) Uinv[(int64_t)(e / w) * F.s + e % w] = X[(e / w) * 64 + e % w];
    __syncthreads();
    // X = L^-1 (unit lower), row by row from the top: X[i][j] = -sum_{k=j..i-1} L[i][k] X[k][j]
    for (int e = tid; e < 64 * 64; e += 256) X[e] = {0.0, 0.0};
    __syncthreads();
    for (int i = 0; i < w; i++) {
        cd s = {0.0, 0.0};
        if (j < i)
            for (int k = j + q; k < i; k += 4) {
                const cd t = lcmul(A[k * 64 + i], X[j * 64 + k]);
                s.x += t.x;
                s.y += t.y;
            }
        red[q][j] = s;
        __syncthreads();
        if (q == 0 && j < w) {
            if (j == i) {
                X[i * 64 + i] = {1.0, 0.0};
            } else if (j < i) {
                X[j * 64 + i] = {-((red[0][j].x + red[1][j].x) + (red[2][j].x + red[3][j].x)),
                                 -((red[0][j].y + red[1][j].y) + (red[2][j].y + red[3][j].y))};
            }
        }
        __syncthreads();
    }
    for (int e = tid; e < 64 * 64; e += 256) Li[e] = X[e];
    cd *Linv = pool + F.inv + (int64_t)k0 * F.s + k0;
    for (int e = tid; e < w * w; e += 256) Linv[(int64_t)(e / w) * F.s + e % w] = X[(e / w) * 64 + e % w];
}

// GEMM tiles (see LuGemmItem) on the matrix cores.  The complex product is
// the real one C' = A' B' with A' = [Ar Ai] and, for the rows of A' holding
// column k's real / imaginary parts, B' rows [Br Bi] / [-Bi Br]; each
// v_mfma_f64_16x16x4f64 takes 16 rows of A against 8 complex columns of B
// (lane l holds A'[l % 16][l / 16], B'[l / 16][l % 16] and
// D[l / 16 + 4 j][l % 16]).  Wavefront w owns rows 16 w .. 16 w + 15 of
// the tile and all 64 columns (8 accumulator blocks).  k goes in chunks of
// 16: B's chunk is staged through LDS, A is read straight into registers
// (lane l: row 16 w + l % 16, columns k0 + 4 g + l / 16), and both of the
// next chunk are loaded while this one is multiplied.  A may alias C (a
// wavefront reads only its own rows of A, before it writes them) and B may
// alias C (every read of B ends before the last barrier).
typedef double v4d __attribute__((ext_vector_type(4)));
__global__ void __launch_bounds__(256) k_lu_gemm(const LuGemmItem *items, cd *pool0, int64_t stride)
{
    cd *pool = pool0 + (int64_t)blockIdx.y * stride;
    __shared__ cd Bs[16][66];
    const LuGemmItem it = items[blockIdx.x];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int q = lane >> 4, c16 = lane & 15, bc = c16 & 7;
    const bool im_out = c16 >= 8;
    const bool rows_here = 16 * wave < it.m;  // wave-uniform
    const cd *Ap = pool + it.a + min(16 * wave + c16, it.m - 1);
    const cd *Bp = pool + it.b;
    auto load_a = [&](int k0, cd (&a)[4]) {
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int kk = k0 + 4 * g + q;
            a[g] = rows_here && kk < it.k ? Ap[(int64_t)kk * it.lda] : cd{0.0, 0.0};
        }
    };
    // B staging: thread tid takes elements e = tid + 256 i of the 16 x 64
    // chunk, k fastest (kb = e & 15, column e >> 4)
    auto load_b = [&](int k0, cd (&bv)[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = tid + 256 * i, kb = e & 15, c = e >> 4;
            bv[i] = c < it.n && k0 + kb < it.k ? Bp[(int64_t)c * it.ldb + k0 + kb] : cd{0.0, 0.0};
        }
    };
    v4d acc[8];
#pragma unroll
    for (int bl = 0; bl < 8; bl++) acc[bl] = {0.0, 0.0, 0.0, 0.0};
    cd a[4], bv[4];
    load_a(0, a);
    load_b(0, bv);
    for (int k0 = 0; k0 < it.k; k0 += 16) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int e = tid + 256 * i;
            Bs[e & 15][e >> 4] = bv[i];
        }
        __syncthreads();
        cd an[4], bn[4];
        if (k0 + 16 < it.k) {
            load_a(k0 + 16, an);
            load_b(k0 + 16, bn);
        }
        if (rows_here) {
#pragma unroll
            for (int g = 0; g < 4; g++)
#pragma unroll
                for (int bl = 0; bl < 8; bl++) {
                    const cd b = Bs[4 * g + q][8 * bl + bc];
                    acc[bl] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[g].x, im_out ? b.y : b.x, acc[bl], 0, 0, 0);
                    acc[bl] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[g].y, im_out ? b.x : -b.y, acc[bl], 0, 0, 0);
                }
        }
        __syncthreads();
        if (k0 + 16 < it.k) {
#pragma unroll
            for (int i = 0; i < 4; i++) {
                a[i] = an[i];
                bv[i] = bn[i];
            }
        }
    }
    if (!rows_here) return;
    double *Cd = reinterpret_cast<double *>(pool + it.c);
    const int part = c16 >> 3;
#pragma unroll
    for (int bl = 0; bl < 8; bl++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int r = 16 * wave + q + 4 * j, c = 8 * bl + bc;
            if (r < it.m && c < it.n) {
                double *p = Cd + ((int64_t)c * it.ldc + r) * 2 + part;
                if (it.mode == 1)
                    *p -= acc[bl][j];
                else if (it.mode == 2)
                    *p = -acc[bl][j];
                else
                    *p = acc[bl][j];
            }
        }
}

// dst[i] = pool[src[i]] (or 1 / pool[src[i]] when invert), 0 where src < 0
__global__ void k_lu_gather(const int64_t *src, cd *dst, int64_t cnt, const cd *pool, int invert)
{
    for (int64_t i = gtid(); i < cnt; i += gstride()) {
        const int64_t s = src[i];
        cd v = s >= 0 ? pool[s] : cd{0.0, 0.0};
        if (invert) v = lcinv(v);
        dst[i] = v;
    }
}

// bad[y] = 1 when slot y's inverted pivots 1/U[i,i] hold a non-finite
// value (a zero pivot inverts to Inf / NaN; a non-finite SS propagates);
// bad[] is cleared by k_lu_clear_flags before.  Benign races: every writer
// stores the same value.
__global__ void k_lu_pivot_check(const cd *diag0, int64_t n, int32_t *bad)
{
    const cd *diag = diag0 + (int64_t)blockIdx.y * n;
    bool ok = true;
    for (int64_t i = gtid(); i < n; i += gstride()) {
        const cd v = diag[i];
        ok = ok && isfinite(v.x) && isfinite(v.y) && (v.x != 0.0 || v.y != 0.0);
    }
    if (!ok) bad[blockIdx.y] = 1;
}

__global__ void k_lu_clear_flags(int32_t *bad, int n)
{
    for (int i = threadIdx.x; i < n; i += blockDim.x) bad[i] = 0;
}

// Pivots per group of the blocked trailing update (a multiple of the
// 64-pivot panel).
constexpr int kLuGroup = 256;

// Blocks of 256 for a grid-stride kernel over n elements: enough to fill
// the chip, never more than kMaxBlocks1d (n > 2^32 included).
constexpr int64_t kMaxBlocks1d = 1 << 16;
inline unsigned grid1d(int64_t n) { return (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, kMaxBlocks1d)); }

template <typename T>
int upload(DevBuf<T> &d, const std::vector<T> &h)
{
    if (!d.alloc(std::max<size_t>(h.size(), 1))) {
        peec_set_error("peec_lu_analyze: out of device memory");
        return PEEC_ERR_NOMEM;
    }
    if (!h.empty()) PEEC_HIP_TRY(hipMemcpy(d.p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return PEEC_OK;
}

}  // namespace

// Local index of factor position q in front K (-1 if not in the front).
static int lu_local(const LuPlan &P, int K, int32_t q)
{
    const LuFrontInfo &F = P.fronts[(size_t)K];
    if (q >= F.first && q < F.first + F.s) return q - F.first;
    const int32_t *b0 = P.bnd.data() + P.bnd_ptr[(size_t)K], *b1 = P.bnd.data() + P.bnd_ptr[(size_t)K + 1];
    const int32_t *it = std::lower_bound(b0, b1, q);
    if (it == b1 || *it != q) return -1;
    return F.s + (int)(it - b0);
}

extern "C" int peec_lu_analyze(peec_ctx *ctx, int64_t n, const int64_t *order, int64_t nfront, const int64_t *front_ptr,
                               const int64_t *front_parent, int64_t nblocks, const int64_t *block_ptr)
{
    if (!ctx || !order || !front_ptr || !front_parent || !block_ptr || nfront < 1 || nblocks < 1) {
        peec_set_error("peec_lu_analyze: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->op_ready) {
        peec_set_error("peec_lu_analyze: call peec_operator_set first");
        return PEEC_ERR_STATE;
    }
    if (n != ctx->n || n > INT32_MAX) {
        peec_set_error("peec_lu_analyze: n = %lld, operator has %lld nodes", (long long)n, (long long)ctx->n);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    LuPlan &P = ctx->lu;
    P.ready = false;
    P.n = n;
    // ---- order, fronts, blocks
    P.pos.assign((size_t)n, -1);
    for (int64_t k = 0; k < n; k++) {
        const int64_t v = order[k];
        if (v < 0 || v >= n || P.pos[(size_t)v] >= 0) {
            peec_set_error("peec_lu_analyze: order is not a permutation of 0..n-1");
            return PEEC_ERR_ARG;
        }
        P.pos[(size_t)v] = (int32_t)k;
    }
    if (front_ptr[0] != 0 || front_ptr[nfront] != n) {
        peec_set_error("peec_lu_analyze: front_ptr must run from 0 to n");
        return PEEC_ERR_ARG;
    }
    P.fronts.assign((size_t)nfront, LuFrontInfo{0, 0, 0, 0, -1});
    P.front_of.assign((size_t)n, 0);
    std::vector<std::vector<int32_t>> kids((size_t)nfront);
    for (int64_t K = 0; K < nfront; K++) {
        if (front_ptr[K + 1] <= front_ptr[K]) {
            peec_set_error("peec_lu_analyze: front %lld is empty", (long long)K);
            return PEEC_ERR_ARG;
        }
        const int64_t par = front_parent[K];
        if (par != -1 && (par <= K || par >= nfront)) {
            peec_set_error("peec_lu_analyze: front %lld has parent %lld (parents follow their children)", (long long)K,
                           (long long)par);
            return PEEC_ERR_ARG;
        }
        P.fronts[(size_t)K].first = (int32_t)front_ptr[K];
        P.fronts[(size_t)K].s = (int32_t)(front_ptr[K + 1] - front_ptr[K]);
        P.fronts[(size_t)K].parent = (int32_t)par;
        if (par >= 0) kids[(size_t)par].push_back((int32_t)K);
        for (int64_t q = front_ptr[K]; q < front_ptr[K + 1]; q++) P.front_of[(size_t)q] = (int32_t)K;
    }
    std::vector<int32_t> blk((size_t)nblocks + 1);
    if (block_ptr[0] != 0 || block_ptr[nblocks] != n) {
        peec_set_error("peec_lu_analyze: block_ptr must run from 0 to n");
        return PEEC_ERR_ARG;
    }
    for (int64_t B = 0; B <= nblocks; B++) {
        if (B > 0 && (block_ptr[B] <= block_ptr[B - 1] || block_ptr[B] - block_ptr[B - 1] > kTrsvMaxBlock ||
                      P.front_of[(size_t)block_ptr[B - 1]] != P.front_of[(size_t)(block_ptr[B] - 1)])) {
            peec_set_error("peec_lu_analyze: block %lld is empty, longer than %d rows or crosses a front", (long long)B,
                           kTrsvMaxBlock);
            return PEEC_ERR_ARG;
        }
        blk[(size_t)B] = (int32_t)block_ptr[B];
    }
    // ---- adjacency of SS in factor positions (A' Z^-1 A couples the nodes of
    // an edge; Gamma P^-1 Gamma' the nodes of a surface)
    const int64_t m = ctx->m, ns = ctx->ns;
    std::vector<std::pair<int32_t, int32_t>> pairs;
    auto add_group = [&](const int32_t *nodes, int64_t cnt) {
        for (int64_t a = 0; a < cnt; a++)
            for (int64_t b = 0; b < cnt; b++)
                if (a != b) pairs.push_back({P.pos[(size_t)nodes[a]], P.pos[(size_t)nodes[b]]});
    };
    for (int64_t e = 0; e < m; e++)
        add_group(ctx->hA_idx.data() + ctx->hA_ptr[(size_t)e], ctx->hA_ptr[(size_t)e + 1] - ctx->hA_ptr[(size_t)e]);
    for (int64_t s = 0; s < ns; s++)
        add_group(ctx->hG_idx.data() + ctx->hG_ptr[(size_t)s], ctx->hG_ptr[(size_t)s + 1] - ctx->hG_ptr[(size_t)s]);
    std::vector<int64_t> aptr((size_t)n + 1, 0);
    for (auto &pr : pairs) aptr[(size_t)pr.first + 1]++;
    for (int64_t i = 0; i < n; i++) aptr[(size_t)i + 1] += aptr[(size_t)i];
    std::vector<int32_t> adj(pairs.size());
    {
        std::vector<int64_t> at(aptr.begin(), aptr.end() - 1);
        for (auto &pr : pairs) adj[(size_t)at[(size_t)pr.first]++] = pr.second;
    }
    pairs.clear();
    pairs.shrink_to_fit();
    // ---- boundaries (symbolic factorization on the assembly tree)
    P.bnd_ptr.assign((size_t)nfront + 1, 0);
    P.bnd.clear();
    {
        std::vector<int32_t> mark((size_t)n, -1);
        std::vector<int32_t> tmp;
        for (int64_t K = 0; K < nfront; K++) {
            const int32_t first = P.fronts[(size_t)K].first, last = first + P.fronts[(size_t)K].s;
            tmp.clear();
            for (int32_t q = first; q < last; q++)
                for (int64_t t = aptr[(size_t)q]; t < aptr[(size_t)q + 1]; t++) {
                    const int32_t r = adj[(size_t)t];
                    if (r >= last && mark[(size_t)r] != (int32_t)K) {
                        mark[(size_t)r] = (int32_t)K;
                        tmp.push_back(r);
                    }
                    if (r < first) {
                        // an earlier position must belong to a descendant front
                        int32_t a = P.front_of[(size_t)r];
                        while (a >= 0 && a != (int32_t)K) a = P.fronts[(size_t)a].parent;
                        if (a != (int32_t)K) {
                            peec_set_error("peec_lu_analyze: SS couples position %d to %d outside the assembly tree",
                                           q, r);
                            return PEEC_ERR_ARG;
                        }
                    }
                }
            for (int32_t c : kids[(size_t)K])
                for (int64_t t = P.bnd_ptr[(size_t)c]; t < P.bnd_ptr[(size_t)c + 1]; t++) {
                    const int32_t r = P.bnd[(size_t)t];
                    if (r >= last && mark[(size_t)r] != (int32_t)K) {
                        mark[(size_t)r] = (int32_t)K;
                        tmp.push_back(r);
                    }
                }
            std::sort(tmp.begin(), tmp.end());
            P.bnd.insert(P.bnd.end(), tmp.begin(), tmp.end());
            P.bnd_ptr[(size_t)K + 1] = (int64_t)P.bnd.size();
        }
    }
    // fronts: sizes and pool offsets; flops of the partial factorizations
    int64_t off = 0;
    P.flops = 0.0;
    for (int64_t K = 0; K < nfront; K++) {
        LuFrontInfo &F = P.fronts[(size_t)K];
        const int64_t b = P.bnd_ptr[(size_t)K + 1] - P.bnd_ptr[(size_t)K];
        if (F.s + b > INT32_MAX / 2) {
            peec_set_error("peec_lu_analyze: front %lld too large", (long long)K);
            return PEEC_ERR_ARG;
        }
        F.f = (int32_t)(F.s + b);
        F.off = off;
        off += (int64_t)F.f * F.f;
        const double s = F.s, bb = (double)b;
        P.flops += 8.0 * (s * s * s / 3.0 + s * s * bb + s * bb * bb);
    }
    P.pool_fronts = off;
    // pivot-block inverses after the fronts: L11^-1 then U11^-1, s x s each
    P.inv0 = off;
    for (int64_t K = 0; K < nfront; K++) {
        LuFrontInfo &F = P.fronts[(size_t)K];
        F.inv = off;
        off += 2 * (int64_t)F.s * F.s;
    }
    P.inv_n = off - P.inv0;
    // child boundary -> parent local index
    std::vector<int64_t> rel_off((size_t)nfront, 0);
    std::vector<int32_t> rel;
    for (int64_t K = 0; K < nfront; K++) {
        rel_off[(size_t)K] = (int64_t)rel.size();
        const int32_t par = P.fronts[(size_t)K].parent;
        for (int64_t t = P.bnd_ptr[(size_t)K]; t < P.bnd_ptr[(size_t)K + 1]; t++) {
            const int loc = par >= 0 ? lu_local(P, par, P.bnd[(size_t)t]) : -1;
            if (loc < 0) {
                peec_set_error("peec_lu_analyze: boundary of front %lld is not inside its parent", (long long)K);
                return PEEC_ERR_ARG;
            }
            rel.push_back(loc);
        }
    }
    // ---- stored entries of SS: targets and terms
    std::unordered_map<uint64_t, int64_t> key;
    key.reserve((size_t)(n + adj.size()));
    std::vector<int64_t> tgt;
    auto entry = [&](int32_t pi, int32_t pj, int64_t *id) -> int {
        const uint64_t kk = ((uint64_t)(uint32_t)pi << 32) | (uint32_t)pj;
        auto f = key.find(kk);
        if (f != key.end()) {
            *id = f->second;
            return PEEC_OK;
        }
        const int K = P.front_of[(size_t)std::min(pi, pj)];
        const int li = lu_local(P, K, pi), lj = lu_local(P, K, pj);
        if (li < 0 || lj < 0) {
            peec_set_error("peec_lu_analyze: entry (%d, %d) outside its front", pi, pj);
            return PEEC_ERR_ARG;
        }
        const LuFrontInfo &F = P.fronts[(size_t)K];
        *id = (int64_t)tgt.size();
        key.emplace(kk, *id);
        tgt.push_back(F.off + (int64_t)lj * F.f + li);
        return PEEC_OK;
    };
    int rc;
    int64_t id;
    for (int64_t i = 0; i < n; i++)
        if ((rc = entry((int32_t)i, (int32_t)i, &id))) return rc;
    struct Term {
        int64_t id;
        int32_t edge;
        double coef;
    };
    std::vector<Term> terms;
    for (int64_t e = 0; e < m; e++)
        for (int64_t a = ctx->hA_ptr[(size_t)e]; a < ctx->hA_ptr[(size_t)e + 1]; a++)
            for (int64_t b = ctx->hA_ptr[(size_t)e]; b < ctx->hA_ptr[(size_t)e + 1]; b++) {
                if ((rc = entry(P.pos[(size_t)ctx->hA_idx[(size_t)a]], P.pos[(size_t)ctx->hA_idx[(size_t)b]], &id)))
                    return rc;
                terms.push_back({id, (int32_t)e, ctx->hA_val[(size_t)a] * ctx->hA_val[(size_t)b]});
            }
    std::vector<double> gsum;
    std::vector<std::pair<int64_t, double>> gterms;
    for (int64_t s = 0; s < ns; s++)
        for (int64_t a = ctx->hG_ptr[(size_t)s]; a < ctx->hG_ptr[(size_t)s + 1]; a++)
            for (int64_t b = ctx->hG_ptr[(size_t)s]; b < ctx->hG_ptr[(size_t)s + 1]; b++) {
                if ((rc = entry(P.pos[(size_t)ctx->hG_idx[(size_t)a]], P.pos[(size_t)ctx->hG_idx[(size_t)b]], &id)))
                    return rc;
                gterms.push_back({id, ctx->hG_val[(size_t)a] * ctx->hG_val[(size_t)b] * ctx->h_invP[(size_t)s]});
            }
    P.nss = (int64_t)tgt.size();
    gsum.assign((size_t)P.nss, 0.0);
    for (auto &g : gterms) gsum[(size_t)g.first] += g.second;
    std::vector<int32_t> cptr((size_t)P.nss + 1, 0), cedge(std::max<size_t>(terms.size(), 1));
    std::vector<double> ccoef(cedge.size());
    for (auto &t : terms) cptr[(size_t)t.id + 1]++;
    for (int64_t k = 0; k < P.nss; k++) cptr[(size_t)k + 1] += cptr[(size_t)k];
    {
        std::vector<int32_t> at(cptr.begin(), cptr.end() - 1);
        for (auto &t : terms) {
            const int32_t p = at[(size_t)t.id]++;
            cedge[(size_t)p] = t.edge;
            ccoef[(size_t)p] = t.coef;
        }
    }
    // ---- schedule: levels by height in the tree
    std::vector<int32_t> height((size_t)nfront, 0);
    int H = 0;
    for (int64_t K = 0; K < nfront; K++) {
        for (int32_t c : kids[(size_t)K]) height[(size_t)K] = std::max(height[(size_t)K], height[(size_t)c] + 1);
        H = std::max(H, height[(size_t)K] + 1);
    }
    std::vector<std::vector<int32_t>> lev((size_t)H);
    for (int64_t K = 0; K < nfront; K++) lev[(size_t)height[(size_t)K]].push_back((int32_t)K);
    std::vector<int2> ea;
    std::vector<int4> di;
    std::vector<LuGemmItem> gi;
    P.sched.clear();
    int64_t max_slots = 1;
    for (int h = 0; h < H; h++) {
        size_t maxk = 0;
        for (int32_t K : lev[(size_t)h]) maxk = std::max(maxk, kids[(size_t)K].size());
        for (size_t r = 0; r < maxk; r++) {
            const int64_t first = (int64_t)ea.size();
            for (int32_t K : lev[(size_t)h]) {
                if (kids[(size_t)K].size() <= r) continue;
                const int32_t c = kids[(size_t)K][r];
                const int b = P.fronts[(size_t)c].f - P.fronts[(size_t)c].s;
                for (int j = 0; j < b; j++) ea.push_back(make_int2(c, j));
            }
            if ((int64_t)ea.size() > first) P.sched.push_back({0, first, (int64_t)ea.size() - first});
        }
        int maxs = 0;
        for (int32_t K : lev[(size_t)h]) maxs = std::max(maxs, P.fronts[(size_t)K].s);
        for (int k0 = 0; k0 < maxs; k0 += 64) {
            const int64_t d0 = (int64_t)di.size();
            std::vector<std::pair<int32_t, int32_t>> act;  // (front, slot)
            for (int32_t K : lev[(size_t)h])
                if (P.fronts[(size_t)K].s > k0) {
                    act.push_back({K, (int32_t)act.size()});
                    di.push_back(make_int4(K, (int32_t)(act.size() - 1), k0, 0));
                }
            max_slots = std::max<int64_t>(max_slots, (int64_t)act.size());
            P.sched.push_back({1, d0, (int64_t)di.size() - d0});
            // panel L and U blocks through the inverses (scratch offsets are
            // filled in below, once the scratch base is known)
            const int64_t g0 = (int64_t)gi.size();
            for (auto &a : act) {
                const LuFrontInfo &F = P.fronts[(size_t)a.first];
                const int w = std::min(64, F.s - k0), rest = F.f - k0 - w;
                const int64_t slot = -1 - (int64_t)a.second;  // negative: scratch slot, resolved later
                for (int r = 0; r < rest; r += 64) {
                    LuGemmItem g{};
                    g.a = F.off + (int64_t)k0 * F.f + (k0 + w + r);
                    g.lda = F.f;
                    g.b = slot;  // U^-1
                    g.ldb = 64;
                    g.c = g.a;
                    g.ldc = F.f;
                    g.m = std::min(64, rest - r);
                    g.n = w;
                    g.k = w;
                    g.mode = 0;
                    g.pad = 1;  // b refers to the slot's U^-1
                    gi.push_back(g);
                }
                for (int c = 0; c < rest; c += 64) {
                    LuGemmItem g{};
                    g.a = slot;  // L^-1
                    g.lda = 64;
                    g.b = F.off + (int64_t)(k0 + w + c) * F.f + k0;
                    g.ldb = F.f;
                    g.c = g.b;
                    g.ldc = F.f;
                    g.m = w;
                    g.n = std::min(64, rest - c);
                    g.k = w;
                    g.mode = 0;
                    g.pad = 2;  // a refers to the slot's L^-1
                    gi.push_back(g);
                }
            }
            if ((int64_t)gi.size() > g0) P.sched.push_back({2, g0, (int64_t)gi.size() - g0});
            // trailing update.  Pivots go in groups of kLuGroup: inside a
            // group a panel updates only what the group's later panels read
            // (rows below it in the group's columns, and the group's rows
            // right of the group); the group's last panel then applies the
            // whole group's L and U panels to the rest in one pass of
            // k = group width, so the large trailing matrices are read and
            // written once per group instead of once per 64 pivots
            const int64_t t0 = (int64_t)gi.size();
            for (auto &a : act) {
                const LuFrontInfo &F = P.fronts[(size_t)a.first];
                const int w = std::min(64, F.s - k0), next = k0 + w;
                const int g0 = k0 / kLuGroup * kLuGroup, gend = std::min(g0 + kLuGroup, F.s);
                // C(r0.., c0..) -= L(r0.., kb .. kb + kk) U(kb .. kb + kk, c0..)
                auto tile = [&](int r0, int c0, int m, int n, int kb, int kk) {
                    LuGemmItem g{};
                    g.a = F.off + (int64_t)kb * F.f + r0;
                    g.lda = F.f;
                    g.b = F.off + (int64_t)c0 * F.f + kb;
                    g.ldb = F.f;
                    g.c = F.off + (int64_t)c0 * F.f + r0;
                    g.ldc = F.f;
                    g.m = m;
                    g.n = n;
                    g.k = kk;
                    g.mode = 1;
                    gi.push_back(g);
                };
                if (next < gend) {
                    for (int c = next; c < gend; c += 64)
                        for (int r = next; r < F.f; r += 64) tile(r, c, std::min(64, F.f - r), std::min(64, gend - c), k0, w);
                    for (int c = gend; c < F.f; c += 64)
                        for (int r = next; r < gend; r += 64)
                            tile(r, c, std::min(64, gend - r), std::min(64, F.f - c), k0, w);
                } else {
                    for (int c = gend; c < F.f; c += 64)
                        for (int r = gend; r < F.f; r += 64)
                            tile(r, c, std::min(64, F.f - r), std::min(64, F.f - c), g0, gend - g0);
                }
            }
            if ((int64_t)gi.size() > t0) P.sched.push_back({2, t0, (int64_t)gi.size() - t0});
        }
    }
    P.scratch0 = P.inv0 + P.inv_n;
    P.pool_total = P.scratch0 + max_slots * 8192;
    for (auto &g : gi) {
        if (g.pad == 1) g.b = P.scratch0 + (-1 - g.b) * 8192;         // U^-1 of the slot
        if (g.pad == 2) g.a = P.scratch0 + (-1 - g.a) * 8192 + 4096;  // L^-1 of the slot
        g.pad = 0;
    }
    // ---- off-diagonal 64 x 64 blocks of the pivot-block inverses, block row
    // by block row (k_lu_diag gives the diagonal ones).  X = L11^-1:
    // X_ij = -Linv_ii sum_{k=j..i-1} L_ik X_kj for j < i, rows i = 1, 2, ...;
    // X = U11^-1: X_ij = -Uinv_ii sum_{k=i+1..j} U_ik X_kj for j > i, rows
    // i = nb-2, nb-3, ...  Step t does row t of every L11^-1 and row nb-1-t
    // of every U11^-1: one launch forms the sums in place, the next
    // multiplies them by -Linv_ii / -Uinv_ii (B is C, read before written).
    std::vector<LuGemmItem> ii;
    P.inv_sched.clear();
    {
        int maxnb = 0;
        for (const auto &F : P.fronts) maxnb = std::max(maxnb, (F.s + 63) / 64);
        for (int t = 1; t < maxnb; t++) {
            for (int pass = 0; pass < 2; pass++) {
                const int64_t a0 = (int64_t)ii.size();
                for (const auto &F : P.fronts) {
                    const int s = F.s, nb = (s + 63) / 64;
                    if (nb <= t) continue;
                    const int64_t ss = (int64_t)s * s;
                    {
                        const int i = t, wi = std::min(64, s - 64 * i);
                        for (int j = 0; j < i; j++) {
                            LuGemmItem g{};
                            g.c = F.inv + (int64_t)64 * j * s + 64 * i;
                            g.ldc = s;
                            g.m = wi;
                            g.n = 64;
                            if (pass == 0) {
                                g.a = F.off + (int64_t)64 * j * F.f + 64 * i;
                                g.lda = F.f;
                                g.b = F.inv + (int64_t)64 * j * s + 64 * j;
                                g.ldb = s;
                                g.k = 64 * (i - j);
                                g.mode = 0;
                            } else {
                                g.a = F.inv + (int64_t)64 * i * s + 64 * i;
                                g.lda = s;
                                g.b = g.c;
                                g.ldb = s;
                                g.k = wi;
                                g.mode = 2;
                            }
                            ii.push_back(g);
                        }
                    }
                    {
                        const int i = nb - 1 - t;
                        for (int j = i + 1; j < nb; j++) {
                            const int wj = std::min(64, s - 64 * j);
                            LuGemmItem g{};
                            g.c = F.inv + ss + (int64_t)64 * j * s + 64 * i;
                            g.ldc = s;
                            g.m = 64;
                            g.n = wj;
                            if (pass == 0) {
                                g.a = F.off + (int64_t)64 * (i + 1) * F.f + 64 * i;
                                g.lda = F.f;
                                g.b = F.inv + ss + (int64_t)64 * j * s + 64 * (i + 1);
                                g.ldb = s;
                                g.k = 64 * j + wj - 64 * (i + 1);
                                g.mode = 0;
                            } else {
                                g.a = F.inv + ss + (int64_t)64 * i * s + 64 * i;
                                g.lda = s;
                                g.b = g.c;
                                g.ldb = s;
                                g.k = 64;
                                g.mode = 2;
                            }
                            ii.push_back(g);
                        }
                    }
                }
                if ((int64_t)ii.size() > a0) P.inv_sched.push_back({2, a0, (int64_t)ii.size() - a0});
            }
        }
    }
    // ---- layouts of the triangular solves, with pool sources
    auto gpos = [&](int K, int li) -> int32_t {
        const LuFrontInfo &F = P.fronts[(size_t)K];
        return li < F.s ? F.first + li : P.bnd[(size_t)(P.bnd_ptr[(size_t)K] + li - F.s)];
    };
    TriSources ls, us;
    {
        std::vector<TriEntry> ent;
        for (int64_t K = 0; K < nfront; K++) {
            const LuFrontInfo &F = P.fronts[(size_t)K];
            for (int lj = 0; lj < F.s; lj++)
                for (int li = lj + 1; li < F.f; li++)
                    ent.push_back({gpos((int)K, li), F.first + lj, F.off + (int64_t)lj * F.f + li});
        }
        std::vector<int64_t> dsrc((size_t)n, -1);
        if ((rc = tri_layout(n, true, blk, ent, dsrc, &ctx->Lf, &ls, "peec_lu_analyze (L)"))) return rc;
    }
    {
        std::vector<TriEntry> ent;
        std::vector<int64_t> dsrc((size_t)n, -1);
        for (int64_t K = 0; K < nfront; K++) {
            const LuFrontInfo &F = P.fronts[(size_t)K];
            for (int li = 0; li < F.s; li++) {
                dsrc[(size_t)(F.first + li)] = F.off + (int64_t)li * F.f + li;
                for (int lj = li + 1; lj < F.f; lj++)
                    ent.push_back({F.first + li, gpos((int)K, lj), F.off + (int64_t)lj * F.f + li});
            }
        }
        if ((rc = tri_layout(n, false, blk, ent, dsrc, &ctx->Uf, &us, "peec_lu_analyze (U)"))) return rc;
    }
    // one value set per frequency slot; L is unit lower triangular, so its
    // inverse diagonal is 1 in every slot for good
    if ((rc = tri_set_slots(&ctx->Lf, ctx->nslots)) || (rc = tri_set_slots(&ctx->Uf, ctx->nslots))) return rc;
    {
        std::vector<double2_t> ones((size_t)(n * ctx->nslots), double2_t{1.0, 0.0});
        PEEC_HIP_TRY(hipMemcpy(ctx->Lf.diag.p, ones.data(), ones.size() * 16, hipMemcpyHostToDevice));
    }
    // ---- uploads
    if ((rc = upload(P.l_ob_src, ls.ob)) || (rc = upload(P.l_dense_src, ls.dense)) || (rc = upload(P.u_ob_src, us.ob)) ||
        (rc = upload(P.u_dense_src, us.dense)) || (rc = upload(P.u_diag_src, us.diag)) ||
        (rc = upload(P.d_fronts, P.fronts)) || (rc = upload(P.ss_tgt, tgt)) || (rc = upload(P.ss_cptr, cptr)) ||
        (rc = upload(P.ss_cedge, cedge)) || (rc = upload(P.ss_ccoef, ccoef)) || (rc = upload(P.ss_gsum, gsum)) ||
        (rc = upload(P.ea_items, ea)) || (rc = upload(P.rel_off, rel_off)) || (rc = upload(P.rel, rel)) ||
        (rc = upload(P.diag_items, di)) || (rc = upload(P.gemm_items, gi)) || (rc = upload(P.inv_items, ii)))
        return rc;
    if (!P.pool.alloc((size_t)P.pool_total)) {
        peec_set_error("peec_lu_analyze: out of device memory for %lld front entries", (long long)P.pool_total);
        return PEEC_ERR_NOMEM;
    }
    // permutations: the factor order is the caller's order on both sides
    if (!ctx->perm_r.alloc((size_t)n) || !ctx->perm_c.alloc((size_t)n)) return PEEC_ERR_NOMEM;
    PEEC_HIP_TRY(hipMemcpy(ctx->perm_r.p, P.pos.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(ctx->perm_c.p, P.pos.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    ctx->has_Rs = false;
    ctx->lu_ready = false;
    ctx->slot_lu.assign((size_t)ctx->nslots, 0);
    ctx->lu_layout_from_plan = true;
    P.ready = true;
    P.pool_slot0 = -1;
    P.pool_ns = 0;
    P.pool_inv = false;
    P.layouts_deferred = false;
    ctx->mf_ready = false;
    return mf_analyze(ctx);
}

// The factor values of slots slot0 .. slot0 + ns - 1 (pool copies 0 ..
// ns - 1) into the level-scheduled layouts: U's inverse diagonal always
// (the pivot check reads it), the off-block and dense values when `all`.
static void lu_gather_values(peec_ctx *ctx, int slot0, int ns, bool all)
{
    LuPlan &P = ctx->lu;
    TriFactor &Lf = ctx->Lf, &Uf = ctx->Uf;
    for (int y = 0; y < ns; y++) {
        const int slot = slot0 + y;
        const cd *pl = P.pool.p + (int64_t)y * P.pool_total;
        if (all) {
            hipLaunchKernelGGL(k_lu_gather, dim3(grid1d(Lf.ob_n)), dim3(256), 0, ctx->stream, P.l_ob_src.p,
                               Lf.ob_val.p + slot * Lf.ob_n, Lf.ob_n, pl, 0);
            hipLaunchKernelGGL(k_lu_gather, dim3(grid1d(Lf.dense_n)), dim3(256), 0, ctx->stream, P.l_dense_src.p,
                               Lf.dense.p + slot * Lf.dense_n, Lf.dense_n, pl, 0);
            hipLaunchKernelGGL(k_lu_gather, dim3(grid1d(Uf.ob_n)), dim3(256), 0, ctx->stream, P.u_ob_src.p,
                               Uf.ob_val.p + slot * Uf.ob_n, Uf.ob_n, pl, 0);
            hipLaunchKernelGGL(k_lu_gather, dim3(grid1d(Uf.dense_n)), dim3(256), 0, ctx->stream, P.u_dense_src.p,
                               Uf.dense.p + slot * Uf.dense_n, Uf.dense_n, pl, 0);
        }
        hipLaunchKernelGGL(k_lu_gather, dim3(grid1d(P.n)), dim3(256), 0, ctx->stream, P.u_diag_src.p,
                           Uf.diag.p + slot * P.n, P.n, pl, 1);
    }
}

int lu_gather_deferred(peec_ctx *ctx)
{
    LuPlan &P = ctx->lu;
    if (!P.layouts_deferred) return PEEC_OK;
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    lu_gather_values(ctx, P.pool_slot0, P.pool_ns, true);
    PEEC_HIP_TRY(hipGetLastError());
    P.layouts_deferred = false;
    return PEEC_OK;
}

// Factor slots slot0 .. slot0 + ns - 1 together: every launch of the
// schedule covers all of them (blockIdx.y = slot), each slot's fronts in
// its own copy of the pool.
static int lu_factor_range(peec_ctx *ctx, int slot0, int ns)
{
    LuPlan &P = ctx->lu;
    const int64_t stride = P.pool_total;
    // a factorization that does not cover every slot overwrites pool copies
    // that may still hold the only values of other slots: gather them first
    if (P.layouts_deferred && !(slot0 == 0 && ns == ctx->nslots)) {
        int rc = lu_gather_deferred(ctx);
        if (rc) return rc;
    }
    if (P.pool.n < (size_t)(stride * ns) && !P.pool.alloc((size_t)(stride * ns))) {
        peec_set_error("peec_lu_factor: out of device memory for %d x %lld front entries", ns, (long long)stride);
        return PEEC_ERR_NOMEM;
    }
    // Yle entries -> pool targets (host; a handful of port / lumped entries per slot)
    std::vector<int64_t> ytgt;
    std::vector<double2_t> yval;
    for (int y = 0; y < ns; y++) {
        const peec_ctx::HostCsr &Y = ctx->hYle[(size_t)(slot0 + y)];
        for (int64_t i = 0; i < P.n; i++)
            for (int32_t k = Y.ptr[(size_t)i]; k < Y.ptr[(size_t)i + 1]; k++) {
                const int32_t pi = P.pos[(size_t)i], pj = P.pos[(size_t)Y.idx[(size_t)k]];
                const int K = P.front_of[(size_t)std::min(pi, pj)];
                const int li = lu_local(P, K, pi), lj = lu_local(P, K, pj);
                if (li < 0 || lj < 0) {
                    peec_set_error("peec_lu_factor: Yle couples nodes %lld and %d that share no front; put port and "
                                   "lumped-element nodes in the root front",
                                   (long long)i + 1, Y.idx[(size_t)k] + 1);
                    return PEEC_ERR_ARG;
                }
                const LuFrontInfo &F = P.fronts[(size_t)K];
                ytgt.push_back((int64_t)y * stride + F.off + (int64_t)lj * F.f + li);
                yval.push_back(Y.val[(size_t)k]);
            }
    }
    int pi = ctx->prof_begin(PK_LU, 0.0);
    cd *pool = P.pool.p;
    const unsigned ny = (unsigned)ns;
    hipLaunchKernelGGL(k_lu_zero, dim3(grid1d(P.pool_fronts), ny), dim3(256), 0, ctx->stream, pool, P.pool_fronts,
                       stride);
    hipLaunchKernelGGL(k_lu_ss, dim3(grid1d(P.nss), ny), dim3(256), 0, ctx->stream, P.ss_tgt.p, P.ss_cptr.p,
                       P.ss_cedge.p, P.ss_ccoef.p, P.ss_gsum.p, ctx->invZ.p, ctx->wv.p, slot0, ctx->m, pool, P.nss,
                       stride);
    DevBuf<int64_t> dyt;
    DevBuf<double2_t> dyv;
    if (!ytgt.empty()) {
        if (!dyt.alloc(ytgt.size()) || !dyv.alloc(yval.size())) return PEEC_ERR_NOMEM;
        PEEC_HIP_TRY(hipMemcpyAsync(dyt.p, ytgt.data(), ytgt.size() * 8, hipMemcpyHostToDevice, ctx->stream));
        PEEC_HIP_TRY(hipMemcpyAsync(dyv.p, yval.data(), yval.size() * 16, hipMemcpyHostToDevice, ctx->stream));
        hipLaunchKernelGGL(k_lu_add, dim3(grid1d((int64_t)ytgt.size())), dim3(256), 0, ctx->stream, dyt.p, dyv.p,
                           (int64_t)ytgt.size(), pool);
    }
    for (const auto &L : P.sched) {
        if (L.count <= 0) continue;
        const dim3 grid((unsigned)L.count, ny);
        if (L.kind == 0)
            hipLaunchKernelGGL(k_lu_extend_add, grid, dim3(256), 0, ctx->stream, P.ea_items.p + L.first, P.d_fronts.p,
                               P.rel_off.p, P.rel.p, pool, stride);
        else if (L.kind == 1)
            hipLaunchKernelGGL(k_lu_diag, grid, dim3(256), 0, ctx->stream, P.diag_items.p + L.first, P.d_fronts.p, pool,
                               P.scratch0, stride);
        else
            hipLaunchKernelGGL(k_lu_gemm, grid, dim3(256), 0, ctx->stream, P.gemm_items.p + L.first, pool, stride);
    }
    // the pivot-block inverses of the multifrontal solves, when the pool
    // holds every slot's fronts (the solves read them there)
    P.pool_slot0 = slot0;
    P.pool_ns = ns;
    P.pool_inv = ctx->lu_solve == 1 && ctx->mf.ready && slot0 == 0 && ns == ctx->nslots;
    ctx->mf_ready = false;
    if (P.pool_inv)
        for (const auto &L : P.inv_sched)
            hipLaunchKernelGGL(k_lu_gemm, dim3((unsigned)L.count, ny), dim3(256), 0, ctx->stream,
                               P.inv_items.p + L.first, pool, stride);
    PEEC_HIP_TRY(hipGetLastError());
    // gather into the slots' value sets of the triangular-solve layouts
    TriFactor &Lf = ctx->Lf, &Uf = ctx->Uf;
    if (Lf.nslots < slot0 + ns || Uf.nslots < slot0 + ns) {
        peec_set_error("peec_lu_factor: factor layouts hold %d value sets, slots up to %d requested", Lf.nslots,
                       slot0 + ns - 1);
        return PEEC_ERR_STATE;
    }
    // (the multifrontal solves read the pool itself: the level-scheduled
    // layouts are gathered only if peec_lu_set_solve(0) asks for them)
    P.layouts_deferred = P.pool_inv;
    lu_gather_values(ctx, slot0, ns, !P.layouts_deferred);
    PEEC_HIP_TRY(hipGetLastError());
    // no pivoting: a zero or non-finite pivot makes the factors unusable
    if (P.bad_pivot.n < (size_t)ns && !P.bad_pivot.alloc((size_t)std::max(ns, 64))) {
        peec_set_error("peec_lu_factor: out of device memory for the pivot flags");
        return PEEC_ERR_NOMEM;
    }
    hipLaunchKernelGGL(k_lu_clear_flags, dim3(1), dim3(256), 0, ctx->stream, P.bad_pivot.p, ns);
    hipLaunchKernelGGL(k_lu_pivot_check, dim3(grid1d(P.n), ny), dim3(256), 0, ctx->stream,
                       Uf.diag.p + (int64_t)slot0 * P.n, P.n, P.bad_pivot.p);
    PEEC_HIP_TRY(hipGetLastError());
    ctx->prof_end(pi);
    // the Yle staging buffers are freed on return: wait for their use
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<int32_t> bad((size_t)ns, 0);
    PEEC_HIP_TRY(hipMemcpy(bad.data(), P.bad_pivot.p, (size_t)ns * 4, hipMemcpyDeviceToHost));
    for (int y = 0; y < ns; y++)
        if (bad[(size_t)y]) {
            for (int z = 0; z < ns; z++) ctx->slot_lu[(size_t)(slot0 + z)] = 0;
            ctx->lu_ready = false;
            peec_set_error("peec_lu_factor: zero or non-finite pivot in the factorization of frequency slot %d "
                           "(the device LU does not pivot)", slot0 + y);
            return PEEC_ERR_NUMERIC;
        }
    for (int y = 0; y < ns; y++) ctx->slot_lu[(size_t)(slot0 + y)] = 1;
    ctx->lu_ready = std::all_of(ctx->slot_lu.begin(), ctx->slot_lu.end(), [](char c) { return c != 0; });
    ctx->mf_ready = P.pool_inv;
    return PEEC_OK;
}

static int lu_factor_check(peec_ctx *ctx, int slot0, int ns)
{
    if (!ctx) {
        peec_set_error("peec_lu_factor: null context");
        return PEEC_ERR_ARG;
    }
    LuPlan &P = ctx->lu;
    if (!P.ready || !ctx->lu_layout_from_plan || !ctx->op_ready) {
        peec_set_error("peec_lu_factor: call peec_lu_analyze first (after peec_operator_set, and again after "
                       "peec_lu_set_factors)");
        return PEEC_ERR_STATE;
    }
    for (int slot = slot0; slot < slot0 + ns; slot++)
        if (slot < 0 || slot >= ctx->nslots || (int)ctx->slot_set.size() != ctx->nslots ||
            !ctx->slot_set[(size_t)slot]) {
            peec_set_error("peec_lu_factor: frequency slot %d is not set (peec_frequency_set / _slot)", slot);
            return PEEC_ERR_STATE;
        }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    return PEEC_OK;
}

extern "C" int peec_lu_factor_slot(peec_ctx *ctx, int slot)
{
    int rc = lu_factor_check(ctx, slot, 1);
    return rc ? rc : lu_factor_range(ctx, slot, 1);
}

// Slots slot0 .. slot0 + count - 1, in groups whose front pools fit kLuBatchBytes.
static constexpr int64_t kLuBatchBytes = (int64_t)80 << 30;
extern "C" int peec_lu_factor_slots(peec_ctx *ctx, int slot0, int count)
{
    if (count < 1) {
        peec_set_error("peec_lu_factor_slots: count must be positive");
        return PEEC_ERR_ARG;
    }
    int rc = lu_factor_check(ctx, slot0, count);
    if (rc) return rc;
    const int64_t per = std::max<int64_t>(ctx->lu.pool_total, 1) * 16;
    const int group = (int)std::max<int64_t>(1, std::min<int64_t>(count, kLuBatchBytes / per));
    for (int s0 = slot0; s0 < slot0 + count; s0 += group)
        if ((rc = lu_factor_range(ctx, s0, std::min(group, slot0 + count - s0)))) return rc;
    return PEEC_OK;
}

extern "C" int peec_lu_factor(peec_ctx *ctx) { return peec_lu_factor_slot(ctx, 0); }

// Sizes of the device factorization (after peec_lu_analyze): out[0] n,
// [1] fronts, [2] largest front order f, [3] front-pool entries per slot,
// [4] pool entries per slot with the inverse scratch, [5] stored SS
// entries, [6] L entries (off-block + dense tiles), [7] U entries.
extern "C" int peec_lu_info(peec_ctx *ctx, int64_t *out8)
{
    if (!ctx || !out8 || !ctx->lu.ready) {
        peec_set_error("peec_lu_info: call peec_lu_analyze first");
        return PEEC_ERR_STATE;
    }
    const LuPlan &P = ctx->lu;
    int64_t fmax = 0;
    for (const auto &F : P.fronts) fmax = std::max<int64_t>(fmax, F.f);
    out8[0] = P.n;
    out8[1] = (int64_t)P.fronts.size();
    out8[2] = fmax;
    out8[3] = P.pool_fronts;
    out8[4] = P.pool_total;
    out8[5] = P.nss;
    out8[6] = ctx->Lf.ob_n + ctx->Lf.dense_n;
    out8[7] = ctx->Uf.ob_n + ctx->Uf.dense_n;
    return PEEC_OK;
}

// Read back one front of the last factorization (tests): f x f column-major,
// and the factor position of each of its f rows/columns.
extern "C" int peec_lu_front(peec_ctx *ctx, int64_t K, int64_t *f, int64_t *s, int64_t *first, peec_c128 *out,
                             int64_t *positions)
{
    if (!ctx || !ctx->lu.ready || K < 0 || K >= (int64_t)ctx->lu.fronts.size()) {
        peec_set_error("peec_lu_front: no such front");
        return PEEC_ERR_ARG;
    }
    const LuFrontInfo &F = ctx->lu.fronts[(size_t)K];
    if (f) *f = F.f;
    if (s) *s = F.s;
    if (first) *first = F.first;
    if (positions)
        for (int li = 0; li < F.f; li++)
            positions[li] = li < F.s ? F.first + li : ctx->lu.bnd[(size_t)(ctx->lu.bnd_ptr[(size_t)K] + li - F.s)];
    if (out) {
        PEEC_HIP_TRY(hipSetDevice(ctx->device));
        PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
        PEEC_HIP_TRY(hipMemcpy(out, ctx->lu.pool.p + F.off, (size_t)F.f * F.f * 16, hipMemcpyDeviceToHost));
    }
    return PEEC_OK;
}
