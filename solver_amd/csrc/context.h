// Context object behind the opaque peec_ctx handle.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
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
};

// One FFT-convolution application of the matvec: scatter unknowns of one
// section of x into the kernel's compact grid, convolve, gather into a
// section of y.  map_in/map_out: compact position -> unknown (or -1).
struct ConvApp {
    int kernel = -1;
    int64_t in_off = 0, out_off = 0;  // section offsets inside x / y
    DevBuf<int32_t> map_in, map_out;
    int64_t nnz_in = 0, nnz_out = 0;
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

// Triangular factor prepared for the sync-free solve: rows in solve order,
// each row's off-diagonal entries plus its diagonal.
constexpr int kTrsvMaxBlock = 64;   // rows of one block (one per lane)
constexpr int kTrsvSeg = 256;       // entries summed by one wavefront (4 per lane)

// One triangular factor laid out for the blocked solve (trsv.hip).  The
// factor order is cut into row blocks of at most 64 rows; blocks are
// grouped into dependency levels (solve order: forward for L, backward for
// U).  Entries that reach into other blocks are kept per row (CSR, ob_*)
// and summed in segments of at most kTrsvSeg entries, one wavefront per
// segment; the entries inside a block are kept as a dense s x s
// column-major tile.
struct TriFactor {
    int64_t n = 0;
    int nblocks = 0, max_block = 0;
    int64_t nseg = 0;
    double nnz = 0.0;                        // stored off-diagonal entries + diagonal
    std::vector<int32_t> level_ptr;          // host: blocks of level l are binfo entries level_ptr[l] ..
    std::vector<int64_t> seg_level_ptr;      // host: segments of level l
    std::vector<double> lev_ob, lev_ib, lev_rows;  // host, per level: off-block and in-block entries, rows
    DevBuf<int32_t> binfo;                   // per block, in level order: r0, rows, dense offset (2 x 32 bit)
    DevBuf<int64_t> seg_k0, seg_k1;          // entry range of each segment
    DevBuf<int32_t> row_seg;                 // per row: first and end segment
    DevBuf<int32_t> ob_idx;
    DevBuf<double2_t> ob_val;
    DevBuf<double2_t> dense, diag;           // diag holds 1 / T[i, i]
};

// Per-kernel timing with HIP events on the context's stream.
struct ProfRec {
    int kind;
    double bytes;
    hipEvent_t a, b;
};
enum ProfKind {
    PK_FFT_X_FWD = 0, PK_FFT_Y_FWD, PK_FFT_Z_CONV, PK_FFT_Y_INV, PK_FFT_X_INV, PK_TRSV_GATHER, PK_TRSV_BLOCK_L, PK_TRSV_BLOCK_U, PK_COUNT
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

    // ---- FFT work buffers (complex, sized for the largest application)
    DevBuf<double2_t> w1, w2;
    DevBuf<double2_t> twiddles;             // concatenated exp(-2 pi i t / L) tables
    std::map<int, int64_t> twiddle_off;     // L -> offset into twiddles
    int64_t twiddle_offset(int L);          // builds the table on first use

    // ---- operator
    bool op_ready = false;
    int64_t m = 0, ns = 0, n = 0, mx = 0, my = 0, mz = 0;
    ConvApp apps[12];
    int napps = 0;
    CsrR A, At, G, Gt;           // A (m x n), A' (n x m), Gamma (n x ns), Gamma' (ns x n)
    DevBuf<double> invP;         // ns
    void release_operator();

    // ---- frequency
    bool freq_ready = false;
    double w = 0.0;
    DevBuf<double2_t> DZ, invZ;  // m
    CsrC Yle;                    // n x n

    // ---- LU
    bool lu_ready = false;
    TriFactor Lf, Uf;
    DevBuf<int32_t> perm_r, perm_c;  // inverse row / column permutations (0-based)
    DevBuf<double> Rs;               // optional row scaling of SS
    bool has_Rs = false;

    // ---- vectors
    DevBuf<double2_t> vx, vy, vt1, vt2, vt3;  // N-sized work vectors
    DevBuf<double2_t> trsv_part;              // segment partial sums of the triangular solves

    // ---- profiling
    bool prof_on = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> event_pool;
    int prof_begin(int kind, double bytes);
    void prof_end(int idx);
};

// trsv.hip
std::vector<int32_t> tri_chain_blocks(const peec_csc_c128 *L, int cap);
int tri_build(const peec_csc_c128 *T, bool lower, const std::vector<int32_t> &blk, TriFactor *f, const char *what);
int trsv_solve(peec_ctx *ctx, const TriFactor &f, bool lower, const double2_t *b, double2_t *x, int nrhs);
