/*
 * libpeec_hip -- C ABI of the MI355X (gfx950) voxel-PEEC hot path.
 *
 * Every entry point replaces one function of the reference's FFT solver
 * (Draco112358/Solver, Julia, src/lib/solverFFT/).  The reference has no FFI:
 * a Julia shim included after the reference files redefines each function
 * name with the same positional signature and forwards here with ccall.
 *
 * Conventions
 *  - Host pointers in and out unless a name says d_ (device).  The library
 *    never keeps a host pointer after a call returns.
 *  - Arrays follow Julia's column-major layout; an N x 3 matrix of centres
 *    is x[0..N), y[0..N), z[0..N).
 *  - Sparse matrices are Julia SparseMatrixCSC{Tv,Int64}: colptr[ncol+1] and
 *    rowval[nnz] 1-based Int64, nzval[nnz].
 *  - Complex numbers are interleaved {re, im} doubles (bit-compatible with
 *    Julia ComplexF64).
 *  - Return value: PEEC_OK, PEEC_STOPPED (the registered stop flag was set;
 *    the shim returns the reference's stop sentinel), or a negative error
 *    code; peec_last_error() then describes it.
 *  - One context = one HIP stream.  Calls on distinct contexts may run from
 *    distinct threads; calls on one context must not overlap.
 */
#ifndef PEEC_HIP_H
#define PEEC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct peec_ctx peec_ctx;

typedef struct {
    double re, im;
} peec_c128;

enum {
    PEEC_OK = 0,
    PEEC_STOPPED = 1,
    PEEC_ERR_ARG = -1,
    PEEC_ERR_HIP = -2,
    PEEC_ERR_STATE = -3,
    PEEC_ERR_NUMERIC = -4,
    PEEC_ERR_NOMEM = -5
};

/* ---- context ----------------------------------------------------------- */
int peec_ctx_create(int device, peec_ctx **out);
int peec_ctx_destroy(peec_ctx *ctx);
const char *peec_last_error(void);
/* Cooperative stop (is_stop_requested, solver_start.jl:370-374): the
 * library polls *flag between GMRES iterations and integral tiles. */
int peec_set_stop_flag(peec_ctx *ctx, volatile const int32_t *flag);
/* Waits for the context's stream and the whole device (for timing). */
int peec_synchronize(peec_ctx *ctx);

/* ---- partial-element rows (FP64) --------------------------------------- */
/* compute_Lp_Voxels(centri_m, centri_n, sx, sy, sz, sx2, sy2, sz2, dc,
 *                   is_sym=false, id)            compute_Lp_Voxels.jl:1-63
 * centri_m: 3 doubles (one source centre); centri_n: N x 3; out: N. */
int peec_lp_row(peec_ctx *ctx, const double *centri_m, const double *centri_n, int64_t N,
                double sx, double sy, double sz, double sx2, double sy2, double sz2,
                int dc, double *out);

/* compute_row_P_sup(centers, centers2, sx, sy, sz, caser, caser2, id)
 *                                                compute_row_P_sup.jl:1-104 */
int peec_p_row(peec_ctx *ctx, const double *centers, const double *centers2, int64_t N,
               double sx, double sy, double sz, int caser, int caser2, double *out);

/* Lp_self(l, W, T)                               compute_diagonals.jl:136-155 */
int peec_lp_self(double l, double W, double T, double *out);

/* ---- circulant kernels -------------------------------------------------- */
/* compute_FFT_mutual_coupling_mats(circulant_centers, escalings, Nx, Ny, Nz,
 *   QS_Rcc_FW=1, id, chan)             compute_FFT_mutual_coupling_mats.jl:1-33
 * The circulant centres are generated on the device from the voxel grid
 * (build_center_P_Voxels.jl:1-111, build_centers_Lp_with_air.jl:1-60,
 * build_center_air_Voxels.jl:1-26).  The 3 Lp and 6 P rows are embedded and
 * transformed into device-resident real spectra owned by the context. */
int peec_kernels_build(peec_ctx *ctx, const double origin[3], int Nx, int Ny, int Nz,
                       double sx, double sy, double sz, double esc_Lp, double esc_P);
/* Copy one scaled kernel row back (kind 0..2 = Lp x,y,z; 3..8 = P 11, 22,
 * 33, 12, 13, 23).  out holds the compact grid (x fastest); *n receives its
 * length.  Pass out = NULL to query the length.  kind + 16 reads the
 * retarded path's static row (no escaling), kind + 32 its centre distances
 * (after peec_kernels_build_rcc). */
int peec_kernel_row(peec_ctx *ctx, int kind, double *out, int64_t *n);

/* ---- retarded (Rcc) kernels ---------------------------------------------- */
/* compute_FFT_mutual_coupling_mats(..., QS_Rcc_FW=2, ...)
 *   = compute_rows_Rcc_P + compute_rows_Rcc_Lp
 *                         compute_FFT_mutual_coupling_mats.jl:18-27,
 *                         compute_rows_Rcc_Lp.jl:1-105, compute_rows_Rcc_P.jl:1-64
 * Builds the quasi-static kernels (as peec_kernels_build) and, per kernel,
 * the static row of the retarded path -- the Lp rows from the closed forms
 * of compute_Lp_Voxels_QS.jl, self entry zeroed, no escaling -- and its
 * centre distances Rcc (compute_Voxels_Rcc.jl, compute_P_vox_Rcc.jl).
 * From then on every convolution uses the retarded spectra, which
 * peec_rcc_spectra builds per frequency.  Needs one frequency slot. */
int peec_kernels_build_rcc(peec_ctx *ctx, const double origin[3], int Nx, int Ny, int Nz,
                           double sx, double sy, double sz, double esc_Lp, double esc_P);

/* compute_Circulant_Lp_Rcc(rows_Lp, escalings, f) +
 * compute_Circulant_P_sup_Rcc(rows_P, escalings, f)
 *                         FFT_solver_QS_S_type.jl:61-64,
 *                         compute_Circulant_Lp_Rcc.jl:1-25,
 *                         compute_Circulant_P_sup_Rcc.jl:1-24,
 *                         build_row_vox_Rcc.jl:1-10
 * The complex spectra of escaling * row .* exp(-j beta Rcc), beta = 2 pi f
 * sqrt(eps0 mu0), at frequency freq_hz (Hz, unscaled), for every kernel
 * (both directions of a half-cell-reduced off-diagonal P block).  Call
 * after peec_operator_set, and again for each frequency of the sweep. */
int peec_rcc_spectra(peec_ctx *ctx, double freq_hz);

/* ---- operator (per geometry) -------------------------------------------- */
/* The sparse data of ComputeMatrixVectorFTT (compute_Matrix_vector_fft.jl:1-81):
 *   A       m x n   (incidence_selection["A"], +-1)
 *   Gamma   n x ns  (incidence_selection["Gamma"], 0/1)
 *   lp_map[d] (d = 0..2)  N_ind[d] x m_d   (expansions["mat_map_Lp"][d,1])
 *   p_map[3*j + i] (i, j = 0..2) rows(grid) x ns  (expansions["exp_P"][i,j])
 *   diag_P  ns      (diagonals["P"]; invP = 1 ./ diag_P)
 * All maps are 0/1 with one entry per column (asserted). */
typedef struct {
    const int64_t *colptr;
    const int64_t *rowval;
    const double *nzval;
    int64_t nrow, ncol;
} peec_csc_f64;

typedef struct {
    const int64_t *colptr;
    const int64_t *rowval;
    const peec_c128 *nzval;
    int64_t nrow, ncol;
} peec_csc_c128;

int peec_operator_set(peec_ctx *ctx, int64_t mx, int64_t my, int64_t mz,
                      const peec_csc_f64 *A, const peec_csc_f64 *Gamma,
                      const peec_csc_f64 lp_map[3], const peec_csc_f64 p_map[9],
                      const double *diag_P);

/* Per-frequency data of FFT_solver_QS_S_type.jl:105-119: w (scaled angular
 * frequency), DZ (m), Yle (n x n CSC). */
int peec_frequency_set(peec_ctx *ctx, double w, const peec_c128 *DZ, const peec_csc_c128 *Yle);

/* Frequency slots for batched solves (no counterpart in the reference,
 * whose sweep loop handles one frequency at a time,
 * FFT_solver_QS_S_type.jl:100-232).  peec_frequency_slots allocates
 * nslots slots (clearing all frequency and factor state; more than one
 * slot needs the device factorization, peec_lu_analyze);
 * peec_frequency_set_slot fills one, with the arguments of
 * peec_frequency_set; peec_lu_factor_slot factors its SS on the device.
 * Right-hand side r of a batched call uses slot r / rhs_per_slot.
 * peec_frequency_set and peec_lu_factor are the one-slot forms (slot 0). */
int peec_frequency_slots(peec_ctx *ctx, int nslots);
int peec_frequency_set_slot(peec_ctx *ctx, int slot, double w, const peec_c128 *DZ, const peec_csc_c128 *Yle);

/* DZ on the device (no counterpart in the reference, which evaluates the
 * edge impedances on the host per frequency, FFT_solver_QS_S_type.jl:
 * 106-115 with compute_Z_self.jl:1-15).  peec_edge_terms_set uploads the
 * frequency-independent edge data once per geometry (and conductivity):
 * R (m x 4, row-major: the four half-edge resistances, scaled), Cd (m x 4
 * complex, the half-edge capacitances), Zs0 (m, the surface impedance at
 * unit scaled frequency), Lp (m, the self partial inductances), surface
 * (m flags: edges whose Zs is not zero) and the R and freq escalings.
 * peec_frequency_set_slot_edges then fills a slot as peec_frequency_set_slot
 * does, with DZ = Z_self + Re(Zs) + j w Lp (after the skin-effect switch)
 * and 1 ./ DZ computed on the device.  peec_frequency_dz reads a slot's DZ
 * back (m values). */
int peec_edge_terms_set(peec_ctx *ctx, int64_t m, const double *R, const peec_c128 *Cd, const peec_c128 *Zs0,
                        const double *Lp, const int8_t *surface, double esc_R, double esc_freq);
int peec_frequency_set_slot_edges(peec_ctx *ctx, int slot, double w, const peec_csc_c128 *Yle);
int peec_frequency_dz(peec_ctx *ctx, int slot, peec_c128 *out);
int peec_lu_factor_slot(peec_ctx *ctx, int slot);
/* Factors the SS of slots slot0 .. slot0 + count - 1 (no counterpart in the
 * reference): the slots go through each launch of the multifrontal schedule
 * together (one copy of the front pool per slot, in groups of at most
 * 80 GiB of pools), so the small fronts near the leaves run as one wide
 * launch instead of one per slot.  Same arithmetic as
 * peec_lu_factor_slot on each slot. */
int peec_lu_factor_slots(peec_ctx *ctx, int slot0, int count);

/* LU factors of SS = Yle + A' Z^-1 A + j w Gamma P^-1 Gamma'
 * (F = lu(SS), FFT_solver_QS_S_type.jl:118-119) in UMFPACK's convention,
 * which Julia exposes as F.L, F.U, F.p, F.q, F.Rs:
 *     L * U = (Rs .* SS)[p, q]
 * L lower and U upper triangular (n x n CSC, diagonals stored), p and q
 * 0-based permutations, Rs the row scaling (NULL = no scaling).  The
 * triangular solves run on the GPU. */
int peec_lu_set_factors(peec_ctx *ctx, int64_t n, const peec_csc_c128 *L, const peec_csc_c128 *U,
                        const int64_t *p, const int64_t *q, const double *Rs);
/* Same, with the factor order cut into nblocks contiguous row blocks
 * (block_ptr[0] = 0 ... block_ptr[nblocks] = n, at most 64 rows each),
 * e.g. the leaves and separator chunks of a nested-dissection ordering.
 * All blocks of one dependency level are solved together, one workgroup
 * per block.  No counterpart in the reference (UMFPACK solves on the host);
 * peec_lu_set_factors derives the blocks from chains of the elimination
 * tree of L instead. */
int peec_lu_set_factors_blocked(peec_ctx *ctx, int64_t n, const peec_csc_c128 *L, const peec_csc_c128 *U,
                                const int64_t *p, const int64_t *q, const double *Rs, int64_t nblocks,
                                const int64_t *block_ptr);

/* Device factorization of SS (no counterpart in the reference, which
 * calls UMFPACK's lu(SS) on the host at every frequency,
 * FFT_solver_QS_S_type.jl:118-119).  peec_lu_analyze does the symbolic work
 * once per geometry, after peec_operator_set: order is the elimination
 * order (0-based node ids, n), cut into nfront fronts
 * order[front_ptr[f] .. front_ptr[f+1]) that form an assembly tree
 * (front_parent[f] > f, or -1 for a root) -- e.g. the leaves and separators
 * of a nested dissection, with the port and lumped-element nodes (the only
 * ones Yle couples) in the root front; block_ptr cuts the order into blocks
 * of at most 64 rows inside single fronts for the triangular solves.
 * peec_lu_factor then factors SS at the frequency of the last
 * peec_frequency_set entirely on the device (no pivoting, like the host
 * path) and makes it the preconditioner's factorization.  A later
 * peec_lu_set_factors replaces it; peec_lu_analyze must then run again. */
int peec_lu_analyze(peec_ctx *ctx, int64_t n, const int64_t *order, int64_t nfront, const int64_t *front_ptr,
                    const int64_t *front_parent, int64_t nblocks, const int64_t *block_ptr);
int peec_lu_factor(peec_ctx *ctx);
/* The device factorization has no pivoting: peec_lu_factor, _slot and
 * _slots return PEEC_ERR_NUMERIC (and leave the slots unfactored) when a
 * pivot is zero or non-finite, where UMFPACK would pivot or fail.
 * peec_lu_info (no counterpart in the reference) reports its sizes after
 * peec_lu_analyze: out8 = {n, fronts, largest front order, front-pool
 * entries per slot, pool entries per slot with scratch, stored SS entries,
 * L entries, U entries}. */
int peec_lu_info(peec_ctx *ctx, int64_t *out8);
/* Front K of the last device factorization, f x f column-major (tests),
 * and the factor position of each of its rows/columns (positions, f).
 * out and positions may be NULL to query f, s (pivots) and first (first
 * pivot position). */
int peec_lu_front(peec_ctx *ctx, int64_t K, int64_t *f, int64_t *s, int64_t *first, peec_c128 *out,
                  int64_t *positions);
/* Which triangular solves apply the device factorization inside
 * precond_3_3_vector (compute_Matrix_vector_fft.jl:103-114, where the
 * reference calls F\ three times with UMFPACK's factors; no counterpart for
 * the choice itself).  mode 1 (default): multifrontal, front by front along
 * the assembly tree, reading L11^-1, L21, U12 and U11^-1 of each front in
 * place from the factorization's pool; the factorization then also forms
 * the pivot-block inverses.  It applies when one peec_lu_factor_slots call
 * (or peec_lu_factor with one slot) factored every slot at once, so that
 * the pool holds all of them; otherwise, and with mode 0, the solves use
 * the level-scheduled 64-row blocks.  Mode 1 takes effect with the next
 * such factorization, mode 0 at once.  peec_lu_solve_info reports out4 =
 * {mode, multifrontal in use, tree height, most GEMV tasks in a launch}. */
int peec_lu_set_solve(peec_ctx *ctx, int mode);
int peec_lu_solve_info(peec_ctx *ctx, int64_t *out4);

/* ComputeMatrixVectorFTT(x, w, ...) including the precond_3_3_vector step
 * (compute_Matrix_vector_fft.jl:1-146): y = M^-1 K x, x and y of length
 * m + ns + n. */
int peec_matvec(peec_ctx *ctx, const peec_c128 *x, peec_c128 *y);

/* K x alone, without M^-1 (the first part of ComputeMatrixVectorFTT,
 * compute_Matrix_vector_fft.jl:1-77): exposes the operator for checks. */
int peec_apply_K(peec_ctx *ctx, const peec_c128 *x, peec_c128 *y);

/* precond_3_3_Kt_fft!(F, invZ, invP, A, Gamma, m, ns, is, tn, resProd)
 *                                         FFT_solver_QS_S_type.jl:237-262 */
int peec_precond_rhs(peec_ctx *ctx, const double *is, peec_c128 *tn);

/* gmres_custom(b, restarted=false, tol, maxit, x, ...)  gmres_custom.jl:1-331
 * x is the warm start on input and the solution on output.  resvec must
 * hold maxit + 1 doubles; *nres receives its used length. */
int peec_gmres(peec_ctx *ctx, const peec_c128 *b, double tol, int maxit, peec_c128 *x,
               int *flag, double *relres, int64_t iter[2], double *resvec, int64_t *nres);

/* Batched solves for a frequency sweep (no counterpart in the reference,
 * which calls gmres_custom once per port and frequency,
 * FFT_solver_QS_S_type.jl:120-206).  nrhs right-hand sides b (N x nrhs,
 * column-major, N = m + ns + n), right-hand side r against slot
 * r / rhs_per_slot, each solved for (M^-1 K) x = b by restarted GMRES
 * (restart basis vectors, classical Gram-Schmidt twice) inside iterative
 * refinement until the FP64 residual satisfies ||b - M^-1 K x|| <=
 * tol ||b|| or max_cycles refinement cycles ran.  mixed = 1 applies K with
 * complex-FP32 FFT convolutions inside the cycles (FP64 everywhere else);
 * inner_tol is the residual reduction a mixed cycle aims for (0: 1e-5).
 * x holds the warm starts on input and the solutions on output.  Per
 * right-hand side: flag (0 converged, 1 not within max_cycles, 3 the
 * refinement stalled, 4 its norm or residual is not finite -- NaN or Inf
 * in b, the start or the operator), relres (final FP64 relative
 * residual; NaN with flag 4), iters (inner iterations) and cycles.  A zero
 * right-hand side returns the zero solution with flag 0 and relres 0, as
 * gmres_custom.jl:20-28 does. */
int peec_gmres_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, const peec_c128 *b, double tol, int restart,
                       int max_cycles, int mixed, double inner_tol, peec_c128 *x, int *flag, double *relres, int *iters,
                       int *cycles);
/* tn_r = M^-1 [0; 0; is_r] for nrhs right-hand sides (is: n x nrhs real,
 * column-major; slot r / rhs_per_slot): precond_3_3_Kt_fft!
 * (FFT_solver_QS_S_type.jl:237-262) batched. */
int peec_precond_rhs_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, const double *is, peec_c128 *tn);

/* One batch of the quasi-static sweep kept on the device (no counterpart in
 * the reference; the body of its frequency x port loop,
 * FFT_solver_QS_S_type.jl:120-206, for every frequency slot at once).  For
 * slot s and port c (right-hand side r = s * nports + c), the excitation
 * is = is_value (e_n1 - e_n2) with port_nodes[2c], port_nodes[2c + 1]
 * (0-based node indices, :124-130), its preconditioned right-hand side
 * M^-1 [0; 0; is] (peec_precond_rhs_batched) and the solve of
 * peec_gmres_batched, all without leaving the device.  The warm start of
 * every slot is x0 (N x nports, host) when given, else the previous call's
 * last-slot solutions, which the context keeps (:140, :196), else zero.
 * Returns vports[r * nports + c2] = x_r[n1(c2)] - x_r[n2(c2)] (the node
 * voltage differences s2z needs, :200-206), optionally all solutions
 * (x_out: N x nrhs, or NULL), and per right-hand side flag, relres, iters
 * and cycles as peec_gmres_batched. */
int peec_sweep_batch(peec_ctx *ctx, int nports, const int64_t *port_nodes, double is_value, const peec_c128 *x0,
                     double tol, int restart, int max_cycles, int mixed, double inner_tol, peec_c128 *vports,
                     peec_c128 *x_out, int *flag, double *relres, int *iters, int *cycles);
/* The last-slot solutions peec_sweep_batch keeps (N x nports) copied to x:
 * the warm start to hand to a later call or another context. */
int peec_sweep_state(peec_ctx *ctx, int nports, peec_c128 *x);

/* FFT-pass autotuner (no counterpart in the reference, whose FFTW plans are
 * made with FFTW's defaults, compute_Matrix_vector_fft.jl:20-24).  For every
 * convolution kernel of the current operator and each precision in
 * `precisions` (bit 0: complex128, bit 1: complex64; bit 2 prints every
 * timing to stderr, bit 3 one line per tuned kernel), times the five passes
 * of one convolution over nrhs right-hand sides under each launch shape of
 * the library's table (lines per tile, lanes per line; for the fused z pass
 * also variant 11, the kernel that holds a whole z line in registers), then
 * the right-hand sides per block, then (complex64 only) the radix schedule;
 * in complex64 it also times the fused pipeline (x pass into planes, one
 * y-z pass per plane, x pass back) where it applies, and keeps the faster
 * of the two for batched calls with exactly nrhs right-hand sides.  Launch shapes do not change any result; a
 * complex64 radix schedule changes rounding only.  *report_ms (optional)
 * receives the summed best pass times.  The choices last until the next
 * peec_kernels_build or peec_operator_set. */
int peec_fft_tune(peec_ctx *ctx, int nrhs, int precisions, double *report_ms);
/* The choice peec_fft_tune made for pass `pass` of kernel `kernel` (0..8):
 * out5 = {variant (-1 untuned), lines per tile, lanes per line, right-hand
 * sides per block, radix schedule}.  Passes 0..4 are the five passes of a
 * convolution; 5..7 (complex64) the fused pipeline: 5 x forward into
 * planes, 6 the fused y-z pass (variant 0: the pipeline is used, 1: the
 * five passes are), 7 x inverse from planes. */
int peec_fft_choice(peec_ctx *ctx, int kernel, int pass, int fp32, int nrhs, int *out5);

/* Sets the choice of one pass by hand (a tuning table kept from an earlier
 * run, or a test of every launch shape): variant -1 restores the default
 * (for pass 6: the five passes).  Non-zero radix schedules, and passes
 * 5..7, are accepted for complex64 only.  Variants 12..16 (2..32 lanes per
 * line, lane exchanges) and 17..18 (4 and 8 lanes, LDS exchanges) are the
 * lane-split kernel: on the y and z passes in both precisions, on the x
 * passes 0 and 4 (with their maps) in complex64, for the lengths it is
 * built for; not on passes 5..7. */
int peec_fft_set_choice(peec_ctx *ctx, int kernel, int pass, int fp32, int nrhs, int variant, int rpb, int style);

/* K x (precond = 0), M^-1 K x (precond = 1) or M^-1 x (precond = 2, the
 * preconditioner alone) for nrhs vectors (N x nrhs,
 * column-major), vector r against frequency slot r / rhs_per_slot, with
 * the FFT convolutions in complex128 (fp32 = 0) or complex64 (fp32 = 1):
 * the batched form of ComputeMatrixVectorFTT (compute_Matrix_vector_fft.jl
 * :1-146) that peec_gmres_batched applies, exposed for checks. */
int peec_apply_K_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, int precond, const peec_c128 *x,
                         peec_c128 *y);

/* K x for nrhs device vectors (N x nrhs column-major, vector r against
 * frequency slot r / rhs_per_slot, complex128 or (fp32 = 1) complex64 FFT
 * convolutions), allocating nothing.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory -- code under HIP
 * graph capture, serving loops, harnesses that forbid hipMalloc.  d_x and d_y
 * are device pointers the caller owns; the FFT work grids come from
 * one static device array of 2 x 2^21 complex128 entries (64 MiB) per GPU,
 * held until the process ends.  Calls needing larger grids (work grid
 * entries per right-hand side x nrhs above 2^21; e.g. more than one
 * right-hand side at 200 x 200 x 24 grids) return PEEC_ERR_ARG.  The call
 * is asynchronous on the context's stream (peec_synchronize waits for it);
 * calls that overlap on one GPU share the array and race.  Operator and
 * frequency slots must be set (peec_operator_set, peec_frequency_set*);
 * profiling must be off.  Same kernels and results as
 * peec_apply_K_batched(..., precond = 0, ...). */
int peec_apply_K_noalloc(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, const peec_c128 *d_x,
                         peec_c128 *d_y);

/* ---- adjoint (conductivity gradients) ------------------------------------
 * No counterpart in the reference, which has no derivatives.  K^H is
 * applied explicitly: the 12 convolutions of K (Lp and P are real and
 * symmetric) with the conjugate-transposed sparse blocks and frequency
 * factors (-j w Lp I + conj(DZ) I - A Phi; P Q - j w Gamma' Phi;
 * A' I - Gamma Q + conj(Yle) Phi).  The adjoint preconditioner M^-H reuses
 * the forward factors of SS through the reciprocity similarity
 * M^T = D1 M D2, D1 = diag(I, j w, -I), D2 = diag(I, 1/(j w), -I):
 * M^-H y = conj(D2^-1 M^-1 D1^-1 conj(y)).
 *
 * peec_apply_KH_batched: K^H x (precond = 0), M^-H K^H x (precond = 1) or
 * M^-H x (precond = 2), as peec_apply_K_batched.
 * peec_precond_rhs_batched_adjoint: tn_r = M^-H [0; 0; g_r], g complex node
 * excitations (n x nrhs, column-major).
 * peec_gmres_batched_adjoint: peec_gmres_batched on M^-H K^H lambda = tn. */
int peec_apply_KH_batched(peec_ctx *ctx, int nrhs, int rhs_per_slot, int fp32, int precond, const peec_c128 *x,
                          peec_c128 *y);
int peec_precond_rhs_batched_adjoint(peec_ctx *ctx, int nrhs, int rhs_per_slot, const peec_c128 *g, peec_c128 *tn);
int peec_gmres_batched_adjoint(peec_ctx *ctx, int nrhs, int rhs_per_slot, const peec_c128 *b, double tol,
                               int restart, int max_cycles, int mixed, double inner_tol, peec_c128 *x, int *flag,
                               double *relres, int *iters, int *cycles);

/* HIP graphs (no counterpart in the reference).  With on = 1 (the default)
 * peec_gmres and peec_gmres_batched capture their matvec (FFT passes,
 * sparse terms, triangular solves: a few hundred launches) into a HIP graph
 * at its first use in a call and replay it for every iteration; on = 0
 * launches the kernels one by one.  Results are identical either way. */
int peec_set_graphs(peec_ctx *ctx, int on);

/* Copies `bytes` bytes device to device three times with a plain kernel
 * (k_copy_reference, 16 B per thread): a dispatch of known size for scaling
 * hardware traffic counters (FETCH_SIZE / WRITE_SIZE).  No counterpart in
 * the reference. */
int peec_copy_reference(peec_ctx *ctx, int64_t bytes);

/* Kernel timing (no counterpart in the reference, which only prints @time,
 * do_solving_fft.jl:133).  Reads and clears the per-kind totals recorded
 * since the previous call -- kernel milliseconds from HIP events on the
 * context's stream, algorithmic bytes and launch counts -- and switches
 * recording on (enable = 1) or off.  Kinds: 0..4 the five FFT-convolution
 * passes (x fwd + scatter, y fwd, z fwd * spectrum * z inv, y inv, x inv +
 * gather), 5 triangular-solve gathers, 6 and 7 triangular-solve block
 * substitutions with L and with U (one record per launch), 8 device
 * factorizations (peec_lu_factor, one record per call), 9 the fused
 * y forward / z convolution / y inverse pass (complex64 applications that
 * take it; their x passes count under 0 and 4), 10 triangular-solve panel
 * sums (blocks whose rows share their columns; solves with 4 or more
 * right-hand sides). */
int peec_profile(peec_ctx *ctx, int enable, double *ms, double *bytes, int64_t *count, int nkinds);

#ifdef __cplusplus
}
#endif
#endif /* PEEC_HIP_H */
