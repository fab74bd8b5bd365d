// Context object behind the opaque peec_ctx handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <vector>

#include "peec_internal.h"

// A device array that frees itself.
template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { reset(); }
    void reset()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    bool alloc(size_t count)
    {
        if (count == n && p) return true;
        reset();
        if (count == 0) return true;
        if (hipMalloc(&p, count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return false;
        }
        n = count;
        return true;
    }
};

struct double2_t {
    double x, y;
};

// Workgroup order that keeps runs of consecutive tiles on one XCD.  The
// dispatcher deals workgroups to the 8 XCDs round-robin by linear id and
// each XCD has its own L2, so neighbouring tiles that read the same lines
// would fetch them once per XCD.  xcd_tile maps linear id lin (of n) to a
// tile so that XCD k walks the k-th contiguous range of tiles; it is a
// bijection on [0, n) for every n.
__device__ __forceinline__ unsigned xcd_tile(unsigned lin, unsigned n)
{
    const unsigned q = n / 8, r = n % 8, xcd = lin % 8, loc = lin / 8;
    return xcd < r ? xcd * (q + 1) + loc : r * (q + 1) + (xcd - r) * q + loc;
}

// Which frequency slot each right-hand side (column) of a batched call
// uses.  Uniform: column r -> slot r / rps.  Explicit (rslot non-null): the
// columns are grouped by slot in increasing order, rslot[r] is the slot of
// column r and the columns of slot s are slot_ptr[s] .. slot_ptr[s+1]-1
// (device arrays) -- the compacted batches of GMRES.
struct ColMap {
    const int32_t *rslot = nullptr;
    const int32_t *slot_ptr = nullptr;
    int rps = 1;
    int nsl = 1;   // slots spanned
    int wmax = 0;  // host: most columns of one slot when slot_ptr is set (0: not known)
    __host__ __device__ int slot(int r) const { return rslot ? rslot[r] : r / rps; }
    __host__ __device__ int first(int s) const { return slot_ptr ? slot_ptr[s] : s * rps; }
    __host__ __device__ int end(int s, int nrhs) const
    {
        return slot_ptr ? slot_ptr[s + 1] : (s + 1) * rps < nrhs ? (s + 1) * rps : nrhs;
    }
};
inline ColMap uniform_cols(int nrhs, int rps)
{
    ColMap c;
    c.rps = rps;
    c.nsl = (nrhs + rps - 1) / rps;
    return c;
}

// One circulant kernel (3 Lp + 6 P): compact dims, embedded dims and its
// real spectrum (already divided by Lx*Ly*Lz) in FP64 and FP32.
struct Kernel {
    int n[3] = {0, 0, 0};
    int L[3] = {0, 0, 0};
    DevBuf<double> row;      // compact row (scaled), kept for inspection
    DevBuf<double> spec64;   // L0*L1*L2
    DevBuf<float> spec32;
    int64_t compact() const { return (int64_t)n[0] * n[1] * n[2]; }
    int64_t embedded() const { return (int64_t)L[0] * L[1] * L[2]; }
    // Half-cell reduction of an off-diagonal P kernel (built by
    // peec_operator_set when every scattered and gathered position keeps
    // one parity per half-cell axis): coarse box pn, embedding pL, parity
    // offset delta = out - in per axis, complex spectrum (FP64 pairs and
    // interleaved FP32) with 1/(pL0 pL1 pL2) folded in.
    bool poly = false;
    int pn[3] = {0, 0, 0}, pL[3] = {0, 0, 0}, delta[3] = {0, 0, 0};
    DevBuf<double2_t> pspec64;
    DevBuf<float> pspec32;
    // FP32 spectra in plane order [x][z][y] for the fused y-z pass (built
    // for the applications that use it): even kernel (real) and half-cell
    // reduction (complex)
    DevBuf<float> fspec32, pfspec32;
    // Retarded (Rcc) kernels: the static row without escaling
    // (compute_rows_Rcc_Lp.jl / compute_rows_Rcc_P.jl) and the centre
    // distances Rcc, once per geometry; per frequency the complex spectra of
    // escaling * row .* exp(-j beta Rcc) on the full embedding L, and for a
    // half-cell reduction one per direction ([0]: parity offset delta,
    // [1]: -delta; the reduced kernel of a complex row is not conjugate
    // symmetric), 1/G folded in.
    DevBuf<double> rrow, rdist;
    DevBuf<double2_t> rspec64, rpspec64[2];
    DevBuf<float> rspec32, rpspec32[2];
};

// One FFT-convolution application of the matvec: scatter unknowns of one
// section of x into the kernel's compact grid, convolve, gather into a
// section of y.  map_in/map_out: compact position -> unknown (or -1).
struct ConvApp {
    int kernel = -1;
    int64_t in_off = 0, out_off = 0;  // section offsets inside x / y
    DevBuf<int32_t> map_in, map_out;
    int64_t nnz_in = 0, nnz_out = 0;
    // geometry of this application: the kernel's compact grid and
    // embedding, or the coarse box of its half-cell reduction (poly), whose
    // complex spectrum is applied as is or conjugated (conj: the reverse
    // direction, parity offset -delta)
    int n[3] = {0, 0, 0}, L[3] = {0, 0, 0};
    bool poly = false, conj = false;
    int64_t compact() const { return (int64_t)n[0] * n[1] * n[2]; }
    int64_t embedded() const { return (int64_t)L[0] * L[1] * L[2]; }
};

// One front of the multifrontal factorization (lu.hip): an f x f
// column-major dense matrix at pool offset `off`; its first s rows/columns
// are the pivots (factor positions first .. first + s - 1), the rest its
// boundary.
struct LuFrontInfo {
    int64_t off;
    int32_t f, s, first, parent;
    int64_t inv;  // pool offset of the inverses of the pivot block: L11^-1 (s x s), then U11^-1
};

// A complex GEMM tile C(m x n) = A(m x k) B(k x n) (mode 0), C -= A B
// (mode 1) or C = -A B (mode 2); offsets into the pool, column-major,
// m, n <= 64.
struct LuGemmItem {
    int64_t a, b, c;
    int32_t lda, ldb, ldc, m, n, k, mode, pad;
};

// Symbolic factorization of SS for the device LU (peec_lu_analyze) and the
// launch schedule of its numeric factorization (peec_lu_factor).
struct LuPlan {
    bool ready = false;
    int64_t n = 0;
    std::vector<int32_t> pos;                  // node -> factor position
    std::vector<LuFrontInfo> fronts;           // host copy
    std::vector<int32_t> front_of;             // factor position -> front
    std::vector<int64_t> bnd_ptr;              // boundary positions of front K: bnd[bnd_ptr[K] ..)
    std::vector<int32_t> bnd;
    int64_t pool_fronts = 0, pool_total = 0;   // complex entries
    DevBuf<double2_t> pool;
    DevBuf<LuFrontInfo> d_fronts;
    // SS = Yle + A' Z^-1 A + j w Gamma P^-1 Gamma': per stored entry its pool
    // target, its A' Z^-1 A terms (edge, coefficient) and its Gamma term
    int64_t nss = 0;
    DevBuf<int64_t> ss_tgt;
    DevBuf<int32_t> ss_cptr, ss_cedge;
    DevBuf<double> ss_ccoef, ss_gsum;
    // schedule
    struct Launch {
        int kind;               // 0 extend-add, 1 panel factor, 2 GEMM tiles
        int64_t first, count;   // item range
    };
    std::vector<Launch> sched;
    DevBuf<int2> ea_items;      // (child, column of its update matrix)
    DevBuf<int64_t> rel_off;    // per child: offset of its boundary -> parent map
    DevBuf<int32_t> rel;
    DevBuf<int4> diag_items;    // (front, scratch slot, first pivot, -)
    DevBuf<LuGemmItem> gemm_items;
    int64_t scratch0 = 0;       // pool offset of the per-slot 64 x 64 inverse pairs
    // inverses of every front's pivot block (the multifrontal solves): the
    // 64 x 64 diagonal blocks come from k_lu_diag, the others from GEMM tiles
    // in inv_sched after the factorization
    int64_t inv0 = 0, inv_n = 0;
    std::vector<Launch> inv_sched;
    DevBuf<LuGemmItem> inv_items;
    int pool_slot0 = -1, pool_ns = 0;  // slots whose fronts the pool holds now
    bool pool_inv = false;             // ... with their inverses
    bool layouts_deferred = false;     // Lf/Uf values not gathered from the pool yet
    // values of the triangular-solve layouts, gathered from the pool
    DevBuf<int64_t> l_ob_src, l_dense_src, u_ob_src, u_dense_src, u_diag_src;
    DevBuf<int32_t> bad_pivot;  // per slot of a factorization group: 1 = a zero or non-finite pivot
    double flops = 0.0;
};

// Multifrontal triangular solves (mfsolve.hip) with the device LU's fronts,
// read in place from the pool: per front its pool offsets and where its
// vectors live.  u (the update a front passes to its parent, b rows) is kept
// for every front; w (the front's working rows, f) only for large fronts.
struct MfFront {
    int64_t off, inv;    // pool offsets: the front (f x f), its pivot-block inverses
    int64_t uoff, woff;  // offsets of its u rows / w rows in one right-hand side's u / w area
    int64_t crow, bptr;  // first contribution row (cptr) / first boundary position (bnd)
    int32_t f, s, first, pad;
};
// A GEMV task of a large front: rows r0 .. r0 + m of an output segment from
// kc columns of A (pool offset a of its first element, leading dimension
// lda) against rows k0 .. k0 + kc of the front's working rows w.
struct MfTask {
    int64_t a;
    int32_t lda, m, kc, front, r0, k0;
};
// Sums the partials of tasks p0 .. p0 + np of one launch for rows r0 .. r0 + m.
struct MfFin {
    int32_t front, r0, m, p0, np, pad;
};
struct MfPlan {
    bool ready = false;
    int H = 0;                           // tree heights
    int64_t utot = 0, wtot = 0;          // u / w rows per right-hand side
    int64_t max_tasks = 1;               // most tasks in one GEMV launch
    std::vector<int64_t> small_ptr;      // per height: fronts small_list[small_ptr[h] ..)
    std::vector<int64_t> row_ptr;        // per height: large-front rows rows[row_ptr[h] ..)
    std::vector<int64_t> task_ptr;       // per (height, phase 0..3): tasks[task_ptr[4h + p] ..)
    std::vector<int64_t> fin_ptr;        // per (height, phase 0..3): fins[fin_ptr[4h + p] ..)
    std::vector<double> small_fwd_bytes, small_bwd_bytes, task_bytes;  // algorithmic factor bytes per slot
    DevBuf<MfFront> fr;
    DevBuf<int32_t> small_list, bnd;
    DevBuf<int2> rows;                   // (front, row)
    DevBuf<MfTask> tasks;
    DevBuf<MfFin> fins;
    DevBuf<int64_t> cptr, csrc;          // per front row: the u rows of its children that add into it
};

// Sparse matrix in CSR (0-based, int32 indices) with real or complex values.
struct CsrR {
    int64_t nrow = 0, ncol = 0;
    DevBuf<int32_t> ptr, idx;
    DevBuf<double> val;
};
struct CsrC {
    int64_t nrow = 0, ncol = 0;
    DevBuf<int32_t> ptr, idx;
    DevBuf<double2_t> val;
};

constexpr int kTrsvMaxBlock = 64;   // rows of one block (one per lane)
constexpr int kTrsvSeg = 256;       // entries summed by one wavefront (4 per lane)
constexpr int kTrsvPanelCols = 128; // panel blocks: columns of one segment (2 per lane)
constexpr int kTrsvPanelRows = 16;  // panel blocks: rows one wavefront sums against one x load

// One triangular factor laid out for the blocked solve (trsv.hip).  The
// factor order is cut into row blocks of at most 64 rows; blocks are
// grouped into dependency levels (solve order: forward for L, backward for
// U).  Entries that reach into other blocks are kept per row (CSR, ob_*)
// and summed in segments of at most kTrsvSeg entries, one wavefront per
// segment; the entries inside a block are kept as a dense s x s
// column-major tile.  A block whose rows all reach the same columns (the
// rows of one front's panel) is a panel block: its segments are chunks of
// kTrsvPanelCols columns, and one wavefront sums a chunk for up to
// kTrsvPanelRows rows against one load of x.  In each level the segments
// of the other rows come first, then those of the panel blocks.
struct TriFactor {
    int64_t n = 0;
    int nblocks = 0, max_block = 0;
    int nslots = 1;                          // value sets (one per frequency slot)
    int64_t ob_n = 0, dense_n = 0;           // values per slot in ob_val / dense (diag: n)
    int64_t nseg = 0;
    double nnz = 0.0;                        // stored off-diagonal entries + diagonal
    std::vector<int32_t> level_ptr;          // host: blocks of level l are binfo entries level_ptr[l] ..
    std::vector<int64_t> seg_level_ptr;      // host: segments of level l
    std::vector<int64_t> seg_level_np;       // host: end of level l's segments outside panel blocks
    std::vector<int64_t> item_level_ptr;     // host: panel items of level l
    std::vector<double> lev_ob, lev_pob, lev_ib, lev_rows;  // host, per level: off-block entries (outside
                                                            // and inside panel blocks), in-block entries, rows
    int64_t nitems = 0;
    DevBuf<int32_t> pitem;                   // per panel item: first segment, rows, segment stride, columns
    DevBuf<int32_t> binfo;                   // per block, in level order: r0, rows, dense offset (2 x 32 bit)
    DevBuf<int64_t> seg_k0, seg_k1;          // entry range of each segment
    DevBuf<int32_t> row_seg;                 // per row: first and end segment
    DevBuf<int32_t> ob_idx;
    DevBuf<double2_t> ob_val;                // nslots x ob_n
    DevBuf<double2_t> dense, diag;           // nslots x dense_n, nslots x n; diag holds 1 / T[i, i]
};

// Launch shape of one FFT pass chosen by peec_fft_tune: variant of the
// launch-shape table, right-hand sides per block (fused z pass) and radix
// schedule.  Keyed by (kernel, pass, fp32, nrhs).
struct FftChoice {
    int variant = -1, rpb = 1, style = 0;
};
struct FftTuneKey {
    int kernel, pass, fp32, nrhs;
    bool operator<(const FftTuneKey &o) const
    {
        if (kernel != o.kernel) return kernel < o.kernel;
        if (pass != o.pass) return pass < o.pass;
        if (fp32 != o.fp32) return fp32 < o.fp32;
        return nrhs < o.nrhs;
    }
};

// Per-kernel timing with HIP events on the context's stream.
struct ProfRec {
    int kind;
    double bytes;
    hipEvent_t a, b;
};
enum ProfKind {
    PK_FFT_X_FWD = 0, PK_FFT_Y_FWD, PK_FFT_Z_CONV, PK_FFT_Y_INV, PK_FFT_X_INV, PK_TRSV_GATHER, PK_TRSV_BLOCK_L, PK_TRSV_BLOCK_U, PK_LU,
    PK_FFT_YZ_CONV, PK_TRSV_PANEL, PK_MF_SMALL, PK_MF_GEMV, PK_MF_AUX, PK_COUNT
};

// The batched GMRES workspace, defined in bgmres.hip.
struct BWork;
void bwork_free(BWork *w);
struct BWorkDeleter {
    void operator()(BWork *w) const { bwork_free(w); }
};

struct peec_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    volatile const int32_t *stop_flag = nullptr;

    static constexpr int kScratchSlots = 8;
    void *scratch_ptr[kScratchSlots] = {};
    size_t scratch_bytes[kScratchSlots] = {};
    void *scratch(size_t bytes, int slot);
    bool stop_requested() const;
    void release_all();

    // ---- kernels
    bool kernels_ready = false;
    Kernel kern[9];
    void release_kernels();
    // retarded (Rcc) kernels (peec_kernels_build_rcc): the convolutions use
    // the complex spectra of frequency rcc_freq once peec_rcc_spectra has
    // built them (rcc_ready)
    bool rcc = false, rcc_ready = false;
    double rcc_freq = 0.0, rcc_esc_Lp = 1.0, rcc_esc_P = 1.0;

    // ---- FFT work buffers (complex, sized for the largest application)
    DevBuf<double2_t> w1, w2;
    DevBuf<double2_t> twiddles;             // concatenated exp(-2 pi i t / L) tables
    std::map<int, int64_t> twiddle_off;     // L -> offset into twiddles
    int64_t twiddle_offset(int L);          // builds the table on first use
    std::map<FftTuneKey, FftChoice> fft_tune;  // peec_fft_tune results for the current geometry

    // ---- operator
    bool op_ready = false;
    int64_t m = 0, ns = 0, n = 0, mx = 0, my = 0, mz = 0;
    ConvApp apps[12];
    int napps = 0;
    CsrR A, At, G, Gt;           // A (m x n), A' (n x m), Gamma (n x ns), Gamma' (ns x n)
    DevBuf<double> invP;         // ns
    // host copies for the device LU: A by edge (CSR), Gamma by surface, 1/P
    std::vector<int64_t> hA_ptr, hG_ptr;
    std::vector<int32_t> hA_idx, hG_idx;
    std::vector<double> hA_val, hG_val, h_invP;
    void release_operator();

    // ---- frequency: one or more slots; right-hand side r of a batched call
    // uses slot r / rhs_per_slot
    bool freq_ready = false;        // every slot set
    int nslots = 1;
    double w = 0.0;                 // slot 0
    std::vector<double> hw;         // per slot
    DevBuf<double> wv;              // per slot
    DevBuf<double2_t> DZ, invZ;     // nslots x m
    // per-edge data for DZ on the device (peec_edge_terms_set): half-edge
    // R and Cd (4 per edge), unit-frequency surface impedance, Lp, surface
    // flags, and the escalings
    DevBuf<double> eR, eLp;
    DevBuf<double2_t> eCd, eZs;
    DevBuf<int8_t> eSurf;
    double esc_R = 0.0, esc_freq = 0.0;
    bool edges_ready = false;
    CsrC Yle;                       // nslots consecutive CSRs: ptr nslots x (n + 1), offsets into idx / val
    struct HostCsr {
        std::vector<int32_t> ptr, idx;
        std::vector<double2_t> val;
    };
    std::vector<HostCsr> hYle;      // per slot
    std::vector<char> slot_set, slot_lu;
    int upload_yle();
    // ---- LU
    bool lu_ready = false;          // every slot factored
    TriFactor Lf, Uf;
    DevBuf<int32_t> perm_r, perm_c;  // inverse row / column permutations (0-based)
    DevBuf<double> Rs;               // optional row scaling of SS
    bool has_Rs = false;
    LuPlan lu;                       // device factorization (lu.hip)
    bool lu_layout_from_plan = false;  // Lf/Uf hold the plan's layouts
    // triangular solves: 1 = multifrontal from the pool when it holds every
    // slot's fronts and inverses (mf_ready), 0 = the level-scheduled blocks
    int lu_solve = 1;
    bool mf_ready = false;
    MfPlan mf;
    DevBuf<double2_t> mf_w, mf_u, mf_part;

    // ---- vectors
    DevBuf<double2_t> vx, vy, vt1, vt2, vt3;  // N-sized work vectors
    // peec_sweep_batch: the last slot's port solutions (N x sweep_nports),
    // the warm start of the next batch
    DevBuf<double2_t> sweep_x0;
    int sweep_nports = 0;
    DevBuf<double2_t> trsv_part;              // segment partial sums of the triangular solves

    // ---- HIP graphs: GMRES replays its captured matvec (peec_set_graphs)
    bool use_graphs = true;

    // ---- batched GMRES workspace (bgmres.hip): the Krylov basis and the
    // N x nrhs vectors, kept across peec_gmres_batched / peec_sweep_batch
    // calls of the same shape (nrhs, restart, N) and freed with the operator
    std::unique_ptr<BWork, BWorkDeleter> bwork;

    // ---- profiling
    bool prof_on = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> event_pool;
    int prof_begin(int kind, double bytes);
    void prof_end(int idx);
};

// trsv.hip
// An off-diagonal entry of a triangular factor (factor order) and where its
// value comes from (an index into the caller's value array).
struct TriEntry {
    int32_t r, c;
    int64_t src;
};
// Source index of every value slot of a TriFactor (-1: structural zero).
struct TriSources {
    std::vector<int64_t> ob, dense, diag;
};
int tri_layout(int64_t n, bool lower, const std::vector<int32_t> &blk, const std::vector<TriEntry> &ent,
               const std::vector<int64_t> &diag_src, TriFactor *f, TriSources *src, const char *what);
int tri_set_slots(TriFactor *f, int nslots);  // reallocates the value sets (contents undefined)
std::vector<int32_t> tri_chain_blocks(const peec_csc_c128 *L, int cap);
int tri_build(const peec_csc_c128 *T, bool lower, const std::vector<int32_t> &blk, TriFactor *f, const char *what);
int lu_gather_deferred(peec_ctx *ctx);  // lu.hip: Lf/Uf values from the pool if a factorization skipped them
int mf_analyze(peec_ctx *ctx);  // mfsolve.hip: the plan of the multifrontal solves (after the LU plan)
int mf_prepare(peec_ctx *ctx, int nrhs);  // allocates the solve buffers for nrhs right-hand sides
int mf_solve(peec_ctx *ctx, bool lower, const double2_t *b, double2_t *x, int nrhs, const ColMap &cm);
int trsv_solve(peec_ctx *ctx, const TriFactor &f, bool lower, const double2_t *b, double2_t *x, int nrhs,
               const ColMap &cm);

// The matvec of one GMRES call (M^-1 K x, or M^-H K^H x) captured once per
// (x, y, precision) into a HIP graph and replayed: nothing the matvec reads
// or writes moves while a GMRES call runs, so a cache scoped to the call
// never goes stale.  Falls back to plain launches while profiling (events
// are not recorded into graphs) or when graphs are switched off.
struct MatvecGraphs {
    struct Entry {
        const void *x;
        void *y;
        int fp32, nrhs;
        const int32_t *rslot;
        hipGraphExec_t exec;
    };
    peec_ctx *ctx = nullptr;
    int nrhs = 1, rps = 1;
    int64_t ld = 0;
    bool adjoint = false;
    std::vector<Entry> entries;
    ~MatvecGraphs();
    void clear();  // destroys the captured graphs
    int run(const double2_t *x, double2_t *y, bool fp32);
    // the same with nrhs_c columns grouped by an explicit column map (the
    // map's device arrays may change between calls; their pointers may not)
    int run_cols(const double2_t *x, double2_t *y, bool fp32, int nrhs_c, const ColMap &cm);
};

int apply_K(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32, const ColMap &cm);
int apply_KH(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32, const ColMap &cm);
int apply_precond(peec_ctx *ctx, double2_t *y, int nrhs, int64_t ld, const ColMap &cm);
int apply_precond_adj(peec_ctx *ctx, double2_t *y, int nrhs, int64_t ld, const ColMap &cm);
int matvec_device(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32, const ColMap &cm);
int matvec_device_adj(peec_ctx *ctx, const double2_t *x, double2_t *y, int nrhs, int64_t ld, bool fp32,
                      const ColMap &cm);
