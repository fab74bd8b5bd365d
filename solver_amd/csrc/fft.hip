// LDS-staged mixed-radix FFT passes and the circulant convolution pipeline.
//
// The reference convolves every Toeplitz block with FFTW over a 2x
// zero-padded grid (compute_Matrix_vector_fft.jl:14-65, 149-161) and builds
// the kernel spectra with fft(store_circulant_fft(row)) (compute_FFT_mutual_
// coupling_mats.jl:57-125, 130-155).  Here:
//   - each axis uses its own smooth length L >= 2n-2 (2^a 3^b 5^c); any such
//     length gives the same linear convolution as the reference's 2n (the
//     kernels are even, or half-cell reductions spanning 2n-2 lags);
//   - the even-symmetric embedding has a real spectrum, stored as a real
//     array (FP64 and FP32 copies) with 1/(Lx Ly Lz) folded in;
//   - one convolution = five passes over HBM: x-FFT (fused with the 0/1
//     scatter from the unknown vector, zero padding pruned), y-FFT, a fused
//     z-FFT * spectrum * z-IFFT, y-IFFT (output pruned to the compact rows)
//     and x-IFFT fused with the 0/1 gather back into the result vector.
//
// A pass loads a tile of TB lines into LDS (coalesced along whichever axis
// is contiguous), runs a Stockham autosort FFT with TPL threads per line
// (radix 8/4/2/3/5 butterflies held in registers between barriers), and
// stores the tile back.
#include "context.h"
#include "fft_common.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <cstdio>

// ------------------------------------------------------------------ plans
int fft_z_length(int n);
int fft_smooth_length(int nmin)
{
    if (nmin <= 1) return 1;
    for (int L = nmin;; L++) {
        int r = L;
        while (r % 2 == 0) r /= 2;
        while (r % 3 == 0) r /= 3;
        while (r % 5 == 0) r /= 5;
        if (r == 1) return L;
    }
}

bool fft_make_plan(int L, FftPlan1D *p)
{
    p->L = L;
    p->nstages = 0;
    int r = L;
    while (r > 1) {
        int R;
        if (r % 8 == 0) R = 8;
        else if (r % 4 == 0) R = 4;
        else if (r % 2 == 0) R = 2;
        else if (r % 3 == 0) R = 3;
        else if (r % 5 == 0) R = 5;
        else return false;
        if (p->nstages >= 16) return false;
        p->radix[p->nstages++] = R;
        r /= R;
    }
    return true;
}

int64_t peec_ctx::twiddle_offset(int L)
{
    auto it = twiddle_off.find(L);
    if (it != twiddle_off.end()) return it->second;
    // rebuild the concatenated table with the new length appended
    std::vector<int> lens;
    for (auto &kv : twiddle_off) lens.push_back(kv.first);
    lens.push_back(L);
    int64_t total = 0;
    for (int l : lens) total += l;
    std::vector<double2_t> h((size_t)total);
    std::map<int, int64_t> offs;
    int64_t o = 0;
    for (int l : lens) {
        offs[l] = o;
        for (int t = 0; t < l; t++) {
            long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)t / (long double)l;
            h[(size_t)(o + t)] = {(double)cosl(a), (double)sinl(a)};
        }
        o += l;
    }
    DevBuf<double2_t> nb;
    if (!nb.alloc((size_t)total)) return -1;
    if (hipMemcpy(nb.p, h.data(), sizeof(double2_t) * (size_t)total, hipMemcpyHostToDevice) != hipSuccess) return -1;
    std::swap(twiddles.p, nb.p);
    std::swap(twiddles.n, nb.n);
    twiddle_off = offs;
    return twiddle_off[L];
}

// Register an FFT length's twiddles: the table is rebuilt (and its offsets
// move) whenever a length appears, so every length is registered before the
// first launch.
static bool register_fft_length(peec_ctx *ctx, int L) { return L <= 0 || ctx->twiddle_offset(L) >= 0; }

void peec_ctx::release_kernels()
{
    for (auto &k : kern) {
        k.row.reset();
        k.spec64.reset();
        k.spec32.reset();
        k.n[0] = k.n[1] = k.n[2] = 0;
        k.poly = false;
        k.pspec64.reset();
        k.pspec32.reset();
        k.rrow.reset();
        k.rdist.reset();
        k.rspec64.reset();
        k.rspec32.reset();
        for (int d = 0; d < 2; d++) {
            k.rpspec64[d].reset();
            k.rpspec32[d].reset();
        }
    }
    fft_tune.clear();
    kernels_ready = false;
    rcc = rcc_ready = false;
}

// ------------------------------------------------------------------ device
namespace {

using namespace fftk;

// One Stockham stage over a tile of W = (256 / TPL) * LPT lines held in LDS
// (line stride LS), of which the first nl are transformed.  Lane `lane` of
// line group `lg` owns butterflies k = lane + q * TPL of lines
// lg + lp * (256 / TPL).  twl: the L twiddles exp(-2 pi i t / L) in LDS.
template <typename T, int DIR, int R, int TPL, int LPT, int LMAX>
__device__ __forceinline__ void stockham_stage(C2<T> *tile, int LS, int L, int Ns, int lg, int lane,
                                               const C2<T> *twl, int nl)
{
    constexpr int G = 256 / TPL;
    constexpr int MAXB = (LMAX + R * TPL - 1) / (R * TPL);
    const int nb = L / R;
    const int stride = L / (Ns * R);
    C2<T> v[LPT][MAXB][R];
#pragma unroll
    for (int lp = 0; lp < LPT; lp++) {
        const C2<T> *buf = tile + (lg + lp * G) * LS;
#pragma unroll
        for (int q = 0; q < MAXB; q++) {
            const int k = lane + q * TPL;
            if (k < nb && lg + lp * G < nl) {
#pragma unroll
                for (int r = 0; r < R; r++) v[lp][q][r] = buf[k + r * nb];
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int lp = 0; lp < LPT; lp++) {
        C2<T> *buf = tile + (lg + lp * G) * LS;
#pragma unroll
        for (int q = 0; q < MAXB; q++) {
            const int k = lane + q * TPL;
            if (k < nb && lg + lp * G < nl) {
                const int kk = k % Ns;
                if (Ns > 1) {
#pragma unroll
                    for (int r = 1; r < R; r++) {
                        const C2<T> w = twl[kk * r * stride];
                        const C2<T> wt = {w.x, DIR < 0 ? w.y : -w.y};
                        v[lp][q][r] = cmul(v[lp][q][r], wt);
                    }
                }
                dftR<T, DIR, R>(v[lp][q]);
                const int base = (k - kk) * R + kk;
#pragma unroll
                for (int r = 0; r < R; r++) buf[base + r * Ns] = v[lp][q][r];
            }
        }
    }
    __syncthreads();
}

template <typename T, int DIR, int TPL, int LPT, int LMAX>
__device__ void run_fft(C2<T> *tile, int LS, int L, const PassDesc &d, int lg, int lane, const C2<T> *twl,
                        int nl = 256 / TPL * LPT)
{
    int Ns = 1;
    for (int s = 0; s < d.nstage; s++) {
        const int R = d.radix[s];
        switch (R) {
        case 8: stockham_stage<T, DIR, 8, TPL, LPT, LMAX>(tile, LS, L, Ns, lg, lane, twl, nl); break;
        case 4: stockham_stage<T, DIR, 4, TPL, LPT, LMAX>(tile, LS, L, Ns, lg, lane, twl, nl); break;
        case 2: stockham_stage<T, DIR, 2, TPL, LPT, LMAX>(tile, LS, L, Ns, lg, lane, twl, nl); break;
        case 3: stockham_stage<T, DIR, 3, TPL, LPT, LMAX>(tile, LS, L, Ns, lg, lane, twl, nl); break;
        default: stockham_stage<T, DIR, 5, TPL, LPT, LMAX>(tile, LS, L, Ns, lg, lane, twl, nl); break;
        }
        Ns *= R;
    }
}

// e / n for 0 <= e < 2^17 and 1 <= n <= 1024: the float quotient of e + 1/2
// is at least 1/(2n) away from an integer, far beyond the rounding error.
__device__ __forceinline__ int qdiv(int e, float inv) { return (int)(((float)e + 0.5f) * inv); }

// Input elements a thread carries from one right-hand side to the next:
// forward passes read at most (L + 1) / 2 of L (the zero padding of the
// circulant embedding), inverse passes the whole line.
template <int MODE, int W, int LMAX>
constexpr int fft_prefetch_slots()
{
    return MODE == 1 ? (W * LMAX + 255) / 256 : (W * (LMAX / 2 + 1) + 255) / 256;  // MODE 0, 2, 3
}

// MODE: 0 forward, 1 inverse, 2 forward * real spectrum * inverse, 3
// forward * complex spectrum (or its conjugate) * inverse.
// A block (tile (bx, by, bz) after the XCD remap) transforms W consecutive
// lines (positions b0 = bx W .. b0+W-1 along the tiled axis) at line index
// l2 = by, for right-hand sides bz * rpb .. + rpb - 1.  Lines contiguous in memory (sj == 1) are
// read and written line after line; strided lines are read and written
// row by row across the W lines (coalesced along the tiled axis).  The LDS
// line stride is L + 1 so that both access orders avoid bank conflicts.
//
// Latency hiding: the input of the next right-hand side is loaded into
// registers while the current one is transformed (the positions, and for
// the scatter pass the unknown indices, are computed once per block); the
// twiddles of the line length sit in LDS; for MODE 2 the block's spectrum
// tile is read once into registers and reused for all its right-hand
// sides.
template <typename T, int MODE, int TPL, int LPT, int LMAX>
__global__ void __launch_bounds__(256) k_fft_pass(PassDesc d, const C2<T> *src, C2<T> *dst,
                                                  const double2_t *__restrict__ twall)
{
    constexpr int G = 256 / TPL;
    constexpr int W = G * LPT;
    constexpr bool CONV = MODE >= 2;
    constexpr int NSPEC = CONV ? (W * LMAX + 255) / 256 : 1;
    constexpr int NPF = fft_prefetch_slots<MODE, W, LMAX>();
    extern __shared__ char smem_raw[];
    const int L = d.L, LS = L + 1;
    C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *tile = twl + L;
    const int tid = threadIdx.x;
    unsigned bx, by, bz;
    xcd_block(bx, by, bz);
    const int b0 = (int)bx * W;
    const int l2 = (int)by;
    const int nbt = min(W, d.nb - b0);
    const int lg = tid / TPL, lane = tid % TPL;
    const bool in_contig = (d.in_sj == 1), out_contig = (d.out_sj == 1);
    const float invNin = 1.0f / (float)d.nin, invNout = 1.0f / (float)d.nout;
    const int nin_tile = W * d.nin;
    const bool pf = nin_tile <= 256 * NPF;  // else: load straight into LDS

    for (int t = tid; t < L; t += 256) {
        const double2_t w = twall[d.tw_off + t];
        twl[t] = {(T)w.x, (T)w.y};
    }

    // spectrum tile: real (MODE 2) or complex (MODE 3, imaginary parts in svi)
    T sv[NSPEC], svi[MODE == 3 ? NSPEC : 1];
    if (CONV) {
        const int64_t sbase = (int64_t)b0 * d.spec_sb + (int64_t)l2 * d.spec_s2;
#pragma unroll
        for (int i = 0; i < NSPEC; i++) {
            const int e = tid + 256 * i;
            const int b = e % W, j = e / W;
            const bool in = b < nbt && j < L;
            const int64_t q = sbase + (int64_t)b * d.spec_sb + (int64_t)j * d.spec_sj;
            if (MODE == 2) {
                sv[i] = in ? reinterpret_cast<const T *>(d.spec)[q] : (T)0;
            } else {
                const C2<T> c = in ? reinterpret_cast<const C2<T> *>(d.spec)[q] : C2<T>{(T)0, (T)0};
                sv[i] = c.x;
                svi[MODE == 3 ? i : 0] = d.spec_conj ? -c.y : c.y;
            }
        }
    }

    // input positions of this thread's slots: offset inside one right-hand
    // side's tile, or the unknown index for the scatter pass; -1 = zero
    const int64_t ibase = (int64_t)b0 * d.in_sb + (int64_t)l2 * d.in_s2;
    int32_t poff[NPF];
    C2<T> pv[NPF];
    auto slot_bj = [&](int e, int &b, int &j) {
        if (in_contig) {
            b = qdiv(e, invNin);
            j = e - b * d.nin;
        } else {
            b = e % W;
            j = e / W;
        }
    };
    if (pf) {
#pragma unroll
        for (int i = 0; i < NPF; i++) {
            const int e = tid + 256 * i;
            int o = -1;
            if (e < nin_tile) {
                int b, j;
                slot_bj(e, b, j);
                if (b < nbt) {
                    const int64_t p = (int64_t)b * d.in_sb + (int64_t)j * d.in_sj;
                    o = d.map_in ? d.map_in[ibase + p] : (int32_t)p;
                }
            }
            poff[i] = o;
        }
    }
    auto fetch = [&](int rhs) {
#pragma unroll
        for (int i = 0; i < NPF; i++) {
            C2<T> v = {(T)0, (T)0};
            if (poff[i] >= 0) {
                if (d.map_in) {
                    const double2_t x = d.vec_in[(int64_t)poff[i] + (int64_t)rhs * d.vec_in_ld];
                    v = {(T)x.x, (T)x.y};
                } else {
                    v = src[ibase + (int64_t)rhs * d.in_rhs + poff[i]];
                }
            }
            pv[i] = v;
        }
    };

    const int r0 = (int)bz * d.rpb;
    const int r_end = min(d.nrhs, r0 + d.rpb);
    if (pf && r0 < r_end) fetch(r0);
    for (int rhs = r0; rhs < r_end; rhs++) {
        // ---- fill the tile: the input rows, zeros beyond nin
        if (pf) {
#pragma unroll
            for (int i = 0; i < NPF; i++) {
                const int e = tid + 256 * i;
                if (e < nin_tile) {
                    int b, j;
                    slot_bj(e, b, j);
                    tile[b * LS + j] = pv[i];
                }
            }
            if (rhs + 1 < r_end) fetch(rhs + 1);
        } else {
            for (int e = tid; e < nin_tile; e += 256) {
                int b, j;
                slot_bj(e, b, j);
                C2<T> v = {(T)0, (T)0};
                if (b < nbt) {
                    const int64_t p = ibase + (int64_t)b * d.in_sb + (int64_t)j * d.in_sj;
                    if (d.map_in) {
                        const int u = d.map_in[p];
                        if (u >= 0) {
                            const double2_t x = d.vec_in[(int64_t)u + (int64_t)rhs * d.vec_in_ld];
                            v = {(T)x.x, (T)x.y};
                        }
                    } else {
                        v = src[p + (int64_t)rhs * d.in_rhs];
                    }
                }
                tile[b * LS + j] = v;
            }
        }
        for (int e = tid; e < W * (L - d.nin); e += 256) tile[(e % W) * LS + d.nin + e / W] = {(T)0, (T)0};
        __syncthreads();

        // ---- transform
        if (MODE == 1) {
            run_fft<T, 1, TPL, LPT, LMAX>(tile, LS, L, d, lg, lane, twl);
        } else {
            run_fft<T, -1, TPL, LPT, LMAX>(tile, LS, L, d, lg, lane, twl);
            if (CONV) {
#pragma unroll
                for (int i = 0; i < NSPEC; i++) {
                    const int e = tid + 256 * i;
                    const int b = e % W, j = e / W;
                    if (j < L) {
                        if (MODE == 2)
                            tile[b * LS + j] = cscale(tile[b * LS + j], sv[i]);
                        else
                            tile[b * LS + j] = cmul(tile[b * LS + j], C2<T>{sv[i], svi[MODE == 3 ? i : 0]});
                    }
                }
                __syncthreads();
                run_fft<T, 1, TPL, LPT, LMAX>(tile, LS, L, d, lg, lane, twl);
            }
        }

        // ---- store (first nout elements of each line)
        const int64_t obase = (int64_t)b0 * d.out_sb + (int64_t)l2 * d.out_s2 + (int64_t)rhs * d.out_rhs;
        if (out_contig) {
            for (int e = tid; e < W * d.nout; e += 256) {
                const int b = qdiv(e, invNout), j = e - b * d.nout;
                if (b >= nbt) continue;
                const C2<T> v = tile[b * LS + j];
                const int64_t p = obase + (int64_t)b * d.out_sb + j;
                if (d.map_out) {
                    const int u = d.map_out[p - (int64_t)rhs * d.out_rhs];
                    if (u >= 0) {
                        double2_t *o = d.vec_out + (int64_t)u + (int64_t)rhs * d.vec_out_ld;
                        double2_t a = *o;
                        a.x += (double)v.x;
                        a.y += (double)v.y;
                        *o = a;
                    }
                } else {
                    dst[p] = v;
                }
            }
        } else {
            for (int e = tid; e < W * d.nout; e += 256) {
                const int b = e % W, j = e / W;
                if (b < nbt) dst[obase + (int64_t)b * d.out_sb + (int64_t)j * d.out_sj] = tile[b * LS + j];
            }
        }
        __syncthreads();
    }
}

// ---- z convolution with the whole line in registers
//
// For short z lines one thread owns one line: it loads the nin compact
// rows (strided by Lx Ly, coalesced across the threads' consecutive x),
// runs the forward FFT, the spectrum product and the inverse FFT entirely
// in registers (compile-time length: every index below is a constant after
// unrolling, so the arrays live in VGPRs), and stores the nout rows.  No
// LDS and no barriers; the block's spectrum stays in registers for its rpb
// right-hand sides.  Same Stockham stages and twiddles as the LDS passes.
// MODE 2: real spectrum, MODE 3: complex spectrum (conjugated with spec_conj).
// Lines: x = blockIdx.x * 256 + threadIdx.x < nb, y = blockIdx.y; right-hand
// sides blockIdx.z * rpb ...; src and dst may alias (in place).
template <typename T, int MODE, int L>
__global__ void __launch_bounds__(256) k_zconv_reg(PassDesc d, const C2<T> *src, C2<T> *dst,
                                                   const double2_t *__restrict__ twall)
{
    const int x = blockIdx.x * 256 + threadIdx.x;
    if (x >= d.nb) return;  // no barrier in this kernel
    const int y = blockIdx.y;
    const double2_t *tw = twall + d.tw_off;
    const int64_t sbase = (int64_t)x * d.spec_sb + (int64_t)y * d.spec_s2;
    T sr[L], si[MODE == 3 ? L : 1];
#pragma unroll
    for (int j = 0; j < L; j++) {
        if (MODE == 2) {
            sr[j] = reinterpret_cast<const T *>(d.spec)[sbase + (int64_t)j * d.spec_sj];
        } else {
            const C2<T> c = reinterpret_cast<const C2<T> *>(d.spec)[sbase + (int64_t)j * d.spec_sj];
            sr[j] = c.x;
            si[MODE == 3 ? j : 0] = d.spec_conj ? -c.y : c.y;
        }
    }
    const int64_t ib = (int64_t)x * d.in_sb + (int64_t)y * d.in_s2;
    const int64_t ob = (int64_t)x * d.out_sb + (int64_t)y * d.out_s2;
    const int r_end = min(d.nrhs, ((int)blockIdx.z + 1) * d.rpb);
    for (int rhs = (int)blockIdx.z * d.rpb; rhs < r_end; rhs++) {
        C2<T> a[L];
        const C2<T> *in = src + ib + (int64_t)rhs * d.in_rhs;
#pragma unroll
        for (int j = 0; j < L; j++) a[j] = j < d.nin ? in[(int64_t)j * d.in_sj] : C2<T>{(T)0, (T)0};
        reg_stages<T, -1, L, 0>(a, tw);
#pragma unroll
        for (int j = 0; j < L; j++) {
            if (MODE == 2) a[j] = cscale(a[j], sr[j]);
            else a[j] = cmul(a[j], C2<T>{sr[j], si[MODE == 3 ? j : 0]});
        }
        reg_stages<T, 1, L, 0>(a, tw);
        C2<T> *out = dst + ob + (int64_t)rhs * d.out_rhs;
#pragma unroll
        for (int j = 0; j < L; j++)
            if (j < d.nout) out[(int64_t)j * d.out_sj] = a[j];
    }
}

// ---- fused y forward, z convolution, y inverse over one x' plane
//
// The x pass writes its output plane by plane (x' slowest, then z, then y;
// ny * nz elements per plane per right-hand side).  One block owns one x'
// plane for rpb right-hand sides and keeps it in LDS as nz lines of length
// Ly (line stride Ly + 1):
//   1. read the plane's ny x nz compact values, zero-pad each line to Ly;
//   2. forward y FFT of the nz lines (Stockham stages, as k_fft_pass);
//   3. z convolution: thread t owns column y' = t, reads its nz values out
//      of the tile, zero-pads to LZ, and runs forward z FFT * spectrum *
//      inverse z FFT in registers (as k_zconv_reg), then writes back the
//      nz kept values;
//   4. inverse y FFT of the nz lines, keep ny values per line;
//   5. write the plane back (in place).
// Three HBM passes (y forward, z, y inverse) become one, over the compact
// plane only.  The spectrum is stored plane by plane too, [x'][kz][y']
// (real for MODE 2, complex for MODE 3), so step 3 reads it coalesced; it
// stays in registers for the block's right-hand sides.  The same stages and
// twiddles as the unfused passes give bitwise-identical results.
// d: L = Ly, nin = nout = ny, nb = number of planes, n2 = nz, in_rhs =
// out_rhs = plane stride * planes, in_s2 = plane size, tw_off (y),
// tw_off_z (z).
template <typename T, int MODE, int TPL, int LPT, int LMAXY, int LZ>
__global__ void __launch_bounds__(256) k_yz_conv(PassDesc d, const C2<T> *src, C2<T> *dst,
                                                 const double2_t *__restrict__ twall)
{
    constexpr int G = 256 / TPL;
    extern __shared__ char smem_raw[];
    const int Ly = d.L, LS = Ly + 1, ny = d.nin, nz = d.n2;
    C2<T> *twl = reinterpret_cast<C2<T> *>(smem_raw);
    C2<T> *tile = twl + Ly;
    const int tid = threadIdx.x;
    unsigned bx, by, bz;
    xcd_block(bx, by, bz);
    const int xp = (int)bx;
    const int lg = tid / TPL, lane = tid % TPL;
    for (int t = tid; t < Ly; t += 256) {
        const double2_t w = twall[d.tw_off + t];
        twl[t] = {(T)w.x, (T)w.y};
    }
    const double2_t *twz = twall + d.tw_off_z;
    const bool col = tid < Ly;
    T sr[LZ], si[MODE == 3 ? LZ : 1];
    {
        const int64_t sb = (int64_t)xp * LZ * Ly + (col ? tid : 0);
#pragma unroll
        for (int j = 0; j < LZ; j++) {
            if (MODE == 2) {
                sr[j] = col ? reinterpret_cast<const T *>(d.spec)[sb + (int64_t)j * Ly] : (T)0;
            } else {
                const C2<T> c = col ? reinterpret_cast<const C2<T> *>(d.spec)[sb + (int64_t)j * Ly] : C2<T>{(T)0, (T)0};
                sr[j] = c.x;
                si[MODE == 3 ? j : 0] = d.spec_conj ? -c.y : c.y;
            }
        }
    }
    const float invNy = 1.0f / (float)ny, invPad = Ly > ny ? 1.0f / (float)(Ly - ny) : 0.0f;
    const int r0 = (int)by * d.rpb, r_end = min(d.nrhs, r0 + d.rpb);
    // the plane of the next right-hand side is loaded into registers while
    // the current one is transformed (nz * ny <= 8 LPT * 128 elements)
    constexpr int NPF = 4 * LPT;
    const int npl = nz * ny;
    C2<T> pv[NPF];
    auto fetch = [&](int rhs) {
        const int64_t pb = (int64_t)rhs * d.in_rhs + (int64_t)xp * d.in_s2;
#pragma unroll
        for (int i = 0; i < NPF; i++) {
            const int e = tid + 256 * i;
            pv[i] = e < npl ? src[pb + e] : C2<T>{(T)0, (T)0};
        }
    };
    if (r0 < r_end) fetch(r0);
    for (int rhs = r0; rhs < r_end; rhs++) {
        const int64_t pbase = (int64_t)rhs * d.in_rhs + (int64_t)xp * d.in_s2;
        // 1. the plane, zero-padded lines
#pragma unroll
        for (int i = 0; i < NPF; i++) {
            const int e = tid + 256 * i;
            if (e < npl) {
                const int z = qdiv(e, invNy);
                tile[z * LS + e - z * ny] = pv[i];
            }
        }
        for (int e = tid; e < nz * (Ly - ny); e += 256) {
            const int z = qdiv(e, invPad);
            tile[z * LS + ny + e - z * (Ly - ny)] = C2<T>{(T)0, (T)0};
        }
        if (rhs + 1 < r_end) fetch(rhs + 1);
        __syncthreads();
        // 2. forward y FFT
        run_fft<T, -1, TPL, LPT, LMAXY>(tile, LS, Ly, d, lg, lane, twl, nz);
        // 3. z convolution of column y' = tid
        if (col) {
            C2<T> a[LZ];
#pragma unroll
            for (int j = 0; j < LZ; j++) a[j] = j < nz ? tile[j * LS + tid] : C2<T>{(T)0, (T)0};
            reg_stages<T, -1, LZ, 0>(a, twz);
#pragma unroll
            for (int j = 0; j < LZ; j++) {
                if (MODE == 2) a[j] = cscale(a[j], sr[j]);
                else a[j] = cmul(a[j], C2<T>{sr[j], si[MODE == 3 ? j : 0]});
            }
            reg_stages<T, 1, LZ, 0>(a, twz);
#pragma unroll
            for (int j = 0; j < LZ; j++)
                if (j < nz) tile[j * LS + tid] = a[j];
        }
        __syncthreads();
        // 4. inverse y FFT
        run_fft<T, 1, TPL, LPT, LMAXY>(tile, LS, Ly, d, lg, lane, twl, nz);
        // 5. the plane's ny kept values per line
        for (int e = tid; e < nz * ny; e += 256) {
            const int z = qdiv(e, invNy), y = e - z * ny;
            dst[pbase + (int64_t)z * ny + y] = tile[z * LS + y];
        }
        __syncthreads();
    }
    (void)G;
}

// Spectrum in plane order for k_yz_conv: out[(x * Lz + z) * Ly + y] =
// in[(z * Ly + y) * Lx + x], NC reals per element (1: real, 2: complex).
template <typename T, int NC>
__global__ void k_spec_planes(const T *__restrict__ in, int Lx, int Ly, int Lz, T *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t G = (int64_t)Lx * Ly * Lz;
    if (i >= G) return;
    const int y = (int)(i % Ly);
    const int64_t r = i / Ly;
    const int z = (int)(r % Lz), x = (int)(r / Lz);
    const int64_t j = ((int64_t)z * Ly + y) * Lx + x;
#pragma unroll
    for (int c = 0; c < NC; c++) out[i * NC + c] = in[j * NC + c];
}

// Even-symmetric embedding of a compact row into an L grid
// (store_circulant_fft, compute_FFT_mutual_coupling_mats.jl:130-155, at
// lengths L >= 2n - 2 instead of 2n: at 2n - 2 the offsets n-1 and -(n-1)
// share index n - 1, where both hold row(n-1)).
__global__ void k_embed(const double *__restrict__ row, int nx, int ny, int nz, int Lx, int Ly, int Lz,
                        double2_t *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t G = (int64_t)Lx * Ly * Lz;
    if (i >= G) return;
    int x = (int)(i % Lx);
    int64_t r = i / Lx;
    int y = (int)(r % Ly);
    int z = (int)(r / Ly);
    int sx = x < nx ? x : (Lx - x < nx ? Lx - x : -1);
    int sy = y < ny ? y : (Ly - y < ny ? Ly - y : -1);
    int sz = z < nz ? z : (Lz - z < nz ? Lz - z : -1);
    double v = 0.0;
    if (sx >= 0 && sy >= 0 && sz >= 0) v = row[sx + (int64_t)nx * (sy + (int64_t)ny * sz)];
    out[i] = {v, 0.0};
}

// Half-cell reduction of an off-diagonal P kernel.  Along a half-cell axis
// the fine row holds the kernel at every half-cell offset f = 0 .. nf-1; a
// block that scatters from positions of one parity and gathers at another
// only ever uses the offsets f = |2 m + delta| (m the coarse offset, delta
// = out parity - in parity), so the convolution runs on the coarse box
// nb = (nf + 1) / 2 with the kernel Kc[m] = row(|2 m + delta|).  Kc is not
// even, so it is embedded as a general circulant (offset m at m mod L,
// m = -(nb-1) .. nb-1) and its spectrum is complex.  Axes without half
// cells keep the even embedding (delta = 0, f = |m|).
//
// At L = 2 nb - 2 the offsets nb-1 and -(nb-1) share i = L/2.  On a half-cell
// axis the parity-0 positions number nb and the parity-1 ones nb - 1, so a
// block with delta = +1 (-1) uses the offsets -(nb-1) .. nb-2 (-(nb-2) ..
// nb-1): 2 nb - 2 of them, and the one it does not use lies outside the row
// (|2m + delta| = 2 nb - 1 = nf).  With delta = 0, or without half cells,
// both give the same value.  So i = L/2 takes whichever lies inside.
__device__ __forceinline__ int poly_offset(int i, int L, int nb, int half, int delta, int nf)
{
    if (i <= nb - 1) {
        const int f = half ? abs(2 * i + delta) : i;
        if (f < nf) return f;
    }
    if (L - i <= nb - 1) {
        const int m = i - L;
        const int f = half ? abs(2 * m + delta) : -m;
        if (f < nf) return f;
    }
    return -1;
}

__global__ void k_embed_poly(const double *__restrict__ row, int nfx, int nfy, int nfz, int hx, int hy, int hz,
                             int dx, int dy, int dz, int nbx, int nby, int nbz, int Lx, int Ly, int Lz,
                             double2_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t G = (int64_t)Lx * Ly * Lz;
    if (i >= G) return;
    const int x = (int)(i % Lx);
    const int64_t r = i / Lx;
    const int y = (int)(r % Ly);
    const int z = (int)(r / Ly);
    const int fx = poly_offset(x, Lx, nbx, hx, dx, nfx);
    const int fy = poly_offset(y, Ly, nby, hy, dy, nfy);
    const int fz = poly_offset(z, Lz, nbz, hz, dz, nfz);
    double v = 0.0;
    if (fx >= 0 && fy >= 0 && fz >= 0) v = row[fx + (int64_t)nfx * (fy + (int64_t)nfy * fz)];
    out[i] = {v, 0.0};
}

// escaling * row .* exp(-j beta R) at compact index j (build_row_vox_Rcc.jl:
// row times the unit phasor, then the escaling, compute_Circulant_*_Rcc.jl)
__device__ __forceinline__ double2_t rcc_value(const double *__restrict__ row, const double *__restrict__ dist,
                                               int64_t j, double mbeta, double esc)
{
    double s, c;
    sincos(mbeta * dist[j], &s, &c);
    const double r = row[j];
    return {esc * (r * c), esc * (r * s)};
}

// Even-symmetric embedding of a retarded (complex) row, as k_embed.
__global__ void k_embed_rcc(const double *__restrict__ row, const double *__restrict__ dist, double mbeta, double esc,
                            int nx, int ny, int nz, int Lx, int Ly, int Lz, double2_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t G = (int64_t)Lx * Ly * Lz;
    if (i >= G) return;
    const int x = (int)(i % Lx);
    const int64_t r = i / Lx;
    const int y = (int)(r % Ly);
    const int z = (int)(r / Ly);
    const int sx = x < nx ? x : (Lx - x < nx ? Lx - x : -1);
    const int sy = y < ny ? y : (Ly - y < ny ? Ly - y : -1);
    const int sz = z < nz ? z : (Lz - z < nz ? Lz - z : -1);
    double2_t v = {0.0, 0.0};
    if (sx >= 0 && sy >= 0 && sz >= 0) v = rcc_value(row, dist, sx + (int64_t)nx * (sy + (int64_t)ny * sz), mbeta, esc);
    out[i] = v;
}

// Half-cell reduction of a retarded row for parity offset (dx, dy, dz), as
// k_embed_poly.
__global__ void k_embed_poly_rcc(const double *__restrict__ row, const double *__restrict__ dist, double mbeta,
                                 double esc, int nfx, int nfy, int nfz, int hx, int hy, int hz, int dx, int dy, int dz,
                                 int nbx, int nby, int nbz, int Lx, int Ly, int Lz, double2_t *__restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t G = (int64_t)Lx * Ly * Lz;
    if (i >= G) return;
    const int x = (int)(i % Lx);
    const int64_t r = i / Lx;
    const int y = (int)(r % Ly);
    const int z = (int)(r / Ly);
    const int fx = poly_offset(x, Lx, nbx, hx, dx, nfx);
    const int fy = poly_offset(y, Ly, nby, hy, dy, nfy);
    const int fz = poly_offset(z, Lz, nbz, hz, dz, nfz);
    double2_t v = {0.0, 0.0};
    if (fx >= 0 && fy >= 0 && fz >= 0) v = rcc_value(row, dist, fx + (int64_t)nfx * (fy + (int64_t)nfy * fz), mbeta, esc);
    out[i] = v;
}

__global__ void k_spec_extract_c(const double2_t *__restrict__ g, int64_t G, double scale,
                                 double2_t *__restrict__ s64, float *__restrict__ s32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    const double2_t v = {g[i].x * scale, g[i].y * scale};
    s64[i] = v;
    s32[2 * i] = (float)v.x;
    s32[2 * i + 1] = (float)v.y;
}

__global__ void k_spec_extract(const double2_t *__restrict__ g, int64_t G, double scale, double *__restrict__ s64,
                               float *__restrict__ s32)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G) return;
    double v = g[i].x * scale;
    s64[i] = v;
    s32[i] = (float)v;
}

}  // namespace

// ------------------------------------------------------------------ host side
// lane-split passes (fft_lanes.hip)
bool lane_fits(const fftk::PassDesc &d, int Q, int mode, bool xl, bool fp32);
bool lane_built(int L, int Q, int mode, bool xl, bool map);
int launch_lane_pass(peec_ctx *ctx, bool fp32, int mode, fftk::PassDesc &d, const void *src, void *dst, int Q,
                     bool xl);

struct PassGeom {
    int L, nin, nout, nb, n2;
    int64_t in_sb, in_s2, in_sj, out_sb, out_s2, out_sj;
};

// Launch shapes of k_fft_pass: TPL lanes per line, LPT lines per lane group,
// LMAX the longest line; a tile holds W = 256 / TPL * LPT lines.
struct FftVariant {
    int tpl, lpt, lmax;
};
static constexpr FftVariant kFftVariants[] = {
    {4, 1, 64},  {8, 1, 64},   {8, 2, 128}, {8, 1, 128}, {16, 1, 128}, {16, 2, 256},
    {16, 1, 256}, {32, 1, 256}, {32, 2, 512}, {32, 1, 512}, {64, 1, 1024},
};
static constexpr int kNumFftVariants = (int)(sizeof(kFftVariants) / sizeof(kFftVariants[0]));
static constexpr size_t kFftLdsMax = 64 * 1024;

static int variant_lines(int v) { return 256 / kFftVariants[v].tpl * kFftVariants[v].lpt; }
static size_t variant_lds(int v, int L, bool fp32)
{
    return ((size_t)variant_lines(v) * (size_t)(L + 1) + (size_t)L) * (fp32 ? 8 : 16);
}
static bool variant_fits(int v, int L, bool fp32)
{
    return kFftVariants[v].lmax >= L && variant_lds(v, L, fp32) <= kFftLdsMax;
}

// Default shape: the shortest LMAX that holds L, then the widest tile
// whose LDS stays within 40 KiB (four blocks per CU), at least 4 lines.
static int default_variant(int L, bool fp32)
{
    int best = -1;
    for (int v = 0; v < kNumFftVariants; v++) {
        if (!variant_fits(v, L, fp32)) continue;
        if (best < 0) {
            best = v;
            continue;
        }
        const int lm = kFftVariants[v].lmax, lb = kFftVariants[best].lmax;
        const bool ok_v = variant_lds(v, L, fp32) <= 40 * 1024, ok_b = variant_lds(best, L, fp32) <= 40 * 1024;
        if (lm < lb && ok_v) best = v;
        else if (lm == lb && ok_v && (!ok_b || variant_lines(v) > variant_lines(best))) best = v;
        else if (!ok_b && ok_v) best = v;
    }
    return best;
}

#define PEEC_FFT_CASE(V, TPL_, LPT_, LMAX_)                                                                       \
    case V:                                                                                                       \
        hipLaunchKernelGGL((k_fft_pass<T, MODE, TPL_, LPT_, LMAX_>), grid, dim3(256), lds, ctx->stream, d, s, o, tw); \
        break;

template <typename T, int MODE>
static int launch_pass_t(peec_ctx *ctx, PassDesc &d, const void *src, void *dst, int v)
{
    const int W = variant_lines(v);
    const size_t lds = variant_lds(v, d.L, sizeof(T) == 4);
    const int rpb = std::max(1, d.rpb);
    dim3 grid((unsigned)((d.nb + W - 1) / W), (unsigned)d.n2, (unsigned)((d.nrhs + rpb - 1) / rpb));
    const double2_t *tw = ctx->twiddles.p;
    const C2<T> *s = (const C2<T> *)src;
    C2<T> *o = (C2<T> *)dst;
    switch (v) {
        PEEC_FFT_CASE(0, 4, 1, 64)
        PEEC_FFT_CASE(1, 8, 1, 64)
        PEEC_FFT_CASE(2, 8, 2, 128)
        PEEC_FFT_CASE(3, 8, 1, 128)
        PEEC_FFT_CASE(4, 16, 1, 128)
        PEEC_FFT_CASE(5, 16, 2, 256)
        PEEC_FFT_CASE(6, 16, 1, 256)
        PEEC_FFT_CASE(7, 32, 1, 256)
        PEEC_FFT_CASE(8, 32, 2, 512)
        PEEC_FFT_CASE(9, 32, 1, 512)
        PEEC_FFT_CASE(10, 64, 1, 1024)
    default:
        peec_set_error("FFT launch variant %d out of range", v);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}
#undef PEEC_FFT_CASE

// Shortest circulant of a compact extent n: 2n - 2.  The kernels are even
// along every axis the operator convolves in full (lags -(n-1) and n-1 share
// index n-1 and hold the same value), and a half-cell reduction spans 2n - 2
// lags (poly_offset), so any length >= 2n - 2 gives the linear convolution.
int fft_embed_min(int n) { return n <= 1 ? 1 : 2 * n - 2; }

// z lengths the register kernel is built for (complex64; complex128 up to
// 32), and the z embedding length of a compact extent n: the smallest of
// them >= 2n - 2, else the smallest multiple of 16 of the form
// 16 * 2^a 3^b 5^c.
static const int kRegLengths[] = {8, 16, 32, 48, 64};
static bool reg_length(int L, bool fp32)
{
    for (int l : kRegLengths)
        if (l == L) return fp32 || L <= 32;
    return false;
}
int fft_z_length(int n)
{
    for (int l : kRegLengths)
        if (l >= fft_embed_min(n)) return l;
    // longer z lines: a multiple of 16 (16 * 2^a 3^b 5^c), so that the lane
    // kernels can split them over 8 or 16 lanes with an even share per lane
    for (int L = fft_embed_min(n);; L++)
        if (L % 16 == 0 && fft_smooth_length(L / 16) == L / 16) return L;
}

template <typename T, int MODE>
static int launch_zreg_t(peec_ctx *ctx, PassDesc &d, const void *src, void *dst)
{
    const int rpb = std::max(1, d.rpb);
    dim3 grid((unsigned)((d.nb + 255) / 256), (unsigned)d.n2, (unsigned)((d.nrhs + rpb - 1) / rpb));
    const C2<T> *s = (const C2<T> *)src;
    C2<T> *o = (C2<T> *)dst;
    const double2_t *tw = ctx->twiddles.p;
    switch (d.L) {
    case 8: hipLaunchKernelGGL((k_zconv_reg<T, MODE, 8>), grid, dim3(256), 0, ctx->stream, d, s, o, tw); break;
    case 16: hipLaunchKernelGGL((k_zconv_reg<T, MODE, 16>), grid, dim3(256), 0, ctx->stream, d, s, o, tw); break;
    case 32: hipLaunchKernelGGL((k_zconv_reg<T, MODE, 32>), grid, dim3(256), 0, ctx->stream, d, s, o, tw); break;
    case 48:
        if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_zconv_reg<T, MODE, 48>), grid, dim3(256), 0, ctx->stream, d, s, o, tw);
        break;
    case 64:
        if constexpr (sizeof(T) == 4) hipLaunchKernelGGL((k_zconv_reg<T, MODE, 64>), grid, dim3(256), 0, ctx->stream, d, s, o, tw);
        break;
    default:
        peec_set_error("register z pass: no kernel for length %d", d.L);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// Radix schedules: 0 = radix 8 first (fewest stages), 1 = radix 4 first,
// 2 = radix 3 and 5 first.
static bool make_plan_style(int L, int style, FftPlan1D *p)
{
    if (style == 0) return fft_make_plan(L, p);
    p->L = L;
    p->nstages = 0;
    int r = L;
    int odd[8], nodd = 0, pow2 = 1;
    while (r % 2 == 0) {
        r /= 2;
        pow2 *= 2;
    }
    while (r % 3 == 0) { r /= 3; odd[nodd++] = 3; if (nodd >= 8) return false; }
    while (r % 5 == 0) { r /= 5; odd[nodd++] = 5; if (nodd >= 8) return false; }
    if (r != 1) return false;
    int p2[16], np2 = 0;
    while (pow2 > 1) {
        const int R = style == 1 ? (pow2 % 4 == 0 ? 4 : 2) : (pow2 % 8 == 0 ? 8 : (pow2 % 4 == 0 ? 4 : 2));
        p2[np2++] = R;
        pow2 /= R;
    }
    if (np2 + nodd > 16) return false;
    if (style == 2)
        for (int i = 0; i < nodd; i++) p->radix[p->nstages++] = odd[i];
    for (int i = 0; i < np2; i++) p->radix[p->nstages++] = p2[i];
    if (style != 2)
        for (int i = 0; i < nodd; i++) p->radix[p->nstages++] = odd[i];
    return true;
}

static int make_desc(peec_ctx *ctx, const PassGeom &g, PassDesc *d, int style = 0)
{
    if (g.L > 1024) {
        peec_set_error("FFT length %d exceeds the 1024 supported per axis", g.L);
        return PEEC_ERR_ARG;
    }
    FftPlan1D p;
    if (!make_plan_style(g.L, style, &p)) {
        peec_set_error("FFT length %d is not 2^a 3^b 5^c", g.L);
        return PEEC_ERR_ARG;
    }
    memset(d, 0, sizeof(*d));
    d->L = g.L;
    d->nin = g.nin;
    d->nout = g.nout;
    d->nstage = p.nstages;
    for (int i = 0; i < p.nstages; i++) d->radix[i] = p.radix[i];
    // every offset is read after the last registration (a new length moves them)
    if (!register_fft_length(ctx, g.L)) {
        peec_set_error("twiddle table allocation failed");
        return PEEC_ERR_NOMEM;
    }
    d->tw_off = ctx->twiddle_offset(g.L);
    d->nb = g.nb;
    d->n2 = g.n2;
    d->in_sb = g.in_sb; d->in_s2 = g.in_s2; d->in_sj = g.in_sj;
    d->out_sb = g.out_sb; d->out_s2 = g.out_s2; d->out_sj = g.out_sj;
    d->nrhs = 1;
    d->rpb = 1;
    return PEEC_OK;
}

static int RegFirstRadix(int L) { return L % 8 == 0 ? 8 : L % 4 == 0 ? 4 : L % 2 == 0 ? 2 : L % 3 == 0 ? 3 : 5; }

// Launch shape kRegVariant: the register z kernel (fused z passes only).
static constexpr int kRegVariant = kNumFftVariants;
// Launch shapes kLaneBase + k: the lane-split kernel with Q = 2 << k lanes
// per line (strided passes: y and the fused z pass), in fft_lanes.hip, its
// exchanges across lanes by DPP and swizzles (k < kNumLaneQ) or, for
// kLaneBase + kNumLaneQ + j, through LDS with Q = 4 << j.
static constexpr int kLaneBase = kRegVariant + 1, kNumLaneQ = 5, kNumLaneX = 2, kNumLaneV = kNumLaneQ + kNumLaneX;
static bool is_lane_variant(int v) { return v >= kLaneBase && v < kLaneBase + kNumLaneV; }
static bool lane_xl(int v) { return v >= kLaneBase + kNumLaneQ; }
static int lane_q(int v) { return lane_xl(v) ? 4 << (v - kLaneBase - kNumLaneQ) : 2 << (v - kLaneBase); }

static int launch_pass(peec_ctx *ctx, bool fp32, int mode, PassDesc &d, const void *src, void *dst, int v = -1)
{
    if (d.nb <= 0 || d.n2 <= 0 || d.nrhs <= 0) return PEEC_OK;
    if (is_lane_variant(v)) {
        const int Q = lane_q(v);
        if (!lane_fits(d, Q, mode, lane_xl(v), fp32)) {
            peec_set_error("lane-split FFT pass does not apply (mode %d, length %d, %d lanes)", mode, d.L, Q);
            return PEEC_ERR_ARG;
        }
        return launch_lane_pass(ctx, fp32, mode, d, src, dst, Q, lane_xl(v));
    }
    if (v == kRegVariant) {
        // radix-8-first schedule only (the plan the kernel unrolls)
        if (mode < 2 || !reg_length(d.L, fp32) || d.nstage < 1 || d.radix[0] != RegFirstRadix(d.L)) {
            peec_set_error("register z pass does not apply (mode %d, length %d)", mode, d.L);
            return PEEC_ERR_ARG;
        }
        if (fp32) return mode == 2 ? launch_zreg_t<float, 2>(ctx, d, src, dst) : launch_zreg_t<float, 3>(ctx, d, src, dst);
        return mode == 2 ? launch_zreg_t<double, 2>(ctx, d, src, dst) : launch_zreg_t<double, 3>(ctx, d, src, dst);
    }
    if (v < 0) v = default_variant(d.L, fp32);
    if (v < 0 || !variant_fits(v, d.L, fp32)) {
        peec_set_error("no FFT launch shape for length %d", d.L);
        return PEEC_ERR_ARG;
    }
    // the kernel keeps element offsets inside one tile in 32 bits
    if ((int64_t)variant_lines(v) * d.in_sb + (int64_t)d.L * d.in_sj >= (int64_t)INT32_MAX) {
        peec_set_error("FFT tile offsets exceed 32 bits (line stride %lld)", (long long)d.in_sj);
        return PEEC_ERR_ARG;
    }
    if (fp32) {
        if (mode == 0) return launch_pass_t<float, 0>(ctx, d, src, dst, v);
        if (mode == 1) return launch_pass_t<float, 1>(ctx, d, src, dst, v);
        if (mode == 2) return launch_pass_t<float, 2>(ctx, d, src, dst, v);
        return launch_pass_t<float, 3>(ctx, d, src, dst, v);
    }
    if (mode == 0) return launch_pass_t<double, 0>(ctx, d, src, dst, v);
    if (mode == 1) return launch_pass_t<double, 1>(ctx, d, src, dst, v);
    if (mode == 2) return launch_pass_t<double, 2>(ctx, d, src, dst, v);
    return launch_pass_t<double, 3>(ctx, d, src, dst, v);
}

// Fused y-z pass (k_yz_conv): complex64; y lines up to 256, z lengths of
// the register kernel, at most 32 z lines, the plane within 64 KiB of LDS.
static constexpr int kYzMaxLy = 256, kYzMaxLines = 32;
static size_t yz_lds(int Ly, int nz) { return ((size_t)Ly + (size_t)nz * (size_t)(Ly + 1)) * sizeof(C2<float>); }
static bool yz_fits(const ConvApp &app, bool fp32)
{
    return fp32 && app.L[1] <= kYzMaxLy && app.n[2] >= 1 && app.n[2] <= kYzMaxLines && app.n[2] <= app.L[2] &&
           reg_length(app.L[2], true) && yz_lds(app.L[1], app.n[2]) <= kFftLdsMax;
}

#define PEEC_YZ_CASE(LZ_)                                                                                           \
    case LZ_:                                                                                                       \
        hipLaunchKernelGGL((k_yz_conv<float, MODE, 32, LPT, 256, LZ_>), grid, dim3(256), lds, ctx->stream, d, s, o, \
                           tw);                                                                                     \
        break;

template <int MODE, int LPT>
static int launch_yz_t(peec_ctx *ctx, PassDesc &d, const void *src, void *dst, int Lz)
{
    const int rpb = std::max(1, d.rpb);
    dim3 grid((unsigned)d.nb, (unsigned)((d.nrhs + rpb - 1) / rpb));
    const size_t lds = yz_lds(d.L, d.n2);
    const C2<float> *s = (const C2<float> *)src;
    C2<float> *o = (C2<float> *)dst;
    const double2_t *tw = ctx->twiddles.p;
    switch (Lz) {
        PEEC_YZ_CASE(8)
        PEEC_YZ_CASE(16)
        PEEC_YZ_CASE(32)
        PEEC_YZ_CASE(48)
        PEEC_YZ_CASE(64)
    default:
        peec_set_error("fused y-z pass: no kernel for z length %d", Lz);
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}
#undef PEEC_YZ_CASE

// d: the fused pass descriptor (conv_passes_fused); mode 2 (real
// spectrum) or 3 (complex); Lz the z embedding length.
static int launch_yz(peec_ctx *ctx, int mode, PassDesc &d, const void *src, void *dst, int Lz)
{
    if (d.nb <= 0 || d.n2 <= 0 || d.nrhs <= 0) return PEEC_OK;
    // (the kernel's plane prefetch holds nz * ny <= 8 LPT * 128 elements)
    if (d.L > kYzMaxLy || d.n2 > kYzMaxLines || d.nin > kYzMaxLy / 2 || yz_lds(d.L, d.n2) > kFftLdsMax ||
        !reg_length(Lz, true) || d.n2 > Lz || !d.spec) {
        peec_set_error("fused y-z pass does not apply (Ly %d, %d z lines, z length %d)", d.L, d.n2, Lz);
        return PEEC_ERR_ARG;
    }
    const bool sixteen = d.n2 <= 16;  // tile of 16 lines (LPT 2) or 32 (LPT 4), 32 lanes per line
    if (mode == 2) return sixteen ? launch_yz_t<2, 2>(ctx, d, src, dst, Lz) : launch_yz_t<2, 4>(ctx, d, src, dst, Lz);
    return sixteen ? launch_yz_t<3, 2>(ctx, d, src, dst, Lz) : launch_yz_t<3, 4>(ctx, d, src, dst, Lz);
}

// Plane-order FP32 spectrum of an application that the fused pass can
// take (the even kernel's real spectrum or the half-cell reduction's
// complex one).
static int yz_spectrum_build(peec_ctx *ctx, const ConvApp &app)
{
    Kernel &K = ctx->kern[app.kernel];
    const int64_t G = (int64_t)app.L[0] * app.L[1] * app.L[2];
    const unsigned nb = (unsigned)((G + 255) / 256);
    if (app.poly) {
        if (!K.pfspec32.alloc((size_t)(2 * G))) {
            peec_set_error("operator_set: out of device memory for a plane-order spectrum");
            return PEEC_ERR_NOMEM;
        }
        hipLaunchKernelGGL((k_spec_planes<float, 2>), dim3(nb), dim3(256), 0, ctx->stream, K.pspec32.p, app.L[0],
                           app.L[1], app.L[2], K.pfspec32.p);
    } else {
        if (!K.fspec32.alloc((size_t)G)) {
            peec_set_error("operator_set: out of device memory for a plane-order spectrum");
            return PEEC_ERR_NOMEM;
        }
        hipLaunchKernelGGL((k_spec_planes<float, 1>), dim3(nb), dim3(256), 0, ctx->stream, K.spec32.p, app.L[0],
                           app.L[1], app.L[2], K.fspec32.p);
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// After peec_operator_set has laid out the applications: plane-order
// spectra (and the z twiddles) for every one the fused pass can take.
int yz_prepare(peec_ctx *ctx)
{
    for (int a = 0; a < ctx->napps; a++) {
        const ConvApp &app = ctx->apps[a];
        if (!yz_fits(app, true) || app.compact() <= 0) continue;
        if (ctx->twiddle_offset(app.L[2]) < 0) {
            peec_set_error("operator_set: twiddle table allocation failed");
            return PEEC_ERR_NOMEM;
        }
        int rc = yz_spectrum_build(ctx, app);
        if (rc) return rc;
    }
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}

// compact dims of each circulant kernel
static void kernel_dims(int kind, int Nx, int Ny, int Nz, int d[3])
{
    switch (kind) {
    case 0: d[0] = Nx - 1; d[1] = Ny; d[2] = Nz; break;
    case 1: d[0] = Nx; d[1] = Ny - 1; d[2] = Nz; break;
    case 2: d[0] = Nx; d[1] = Ny; d[2] = Nz - 1; break;
    case 3: d[0] = Nx; d[1] = Ny + 1; d[2] = Nz; break;
    case 4: d[0] = Nx + 1; d[1] = Ny; d[2] = Nz; break;
    case 5: d[0] = Nx; d[1] = Ny; d[2] = Nz + 1; break;
    case 6: d[0] = 2 * (Nx + 1) - 1; d[1] = 2 * (Ny + 1) - 1; d[2] = Nz; break;
    case 7: d[0] = Nx; d[1] = 2 * (Ny + 1) - 1; d[2] = 2 * (Nz + 1) - 1; break;
    default: d[0] = 2 * (Nx + 1) - 1; d[1] = Ny; d[2] = 2 * (Nz + 1) - 1; break;
    }
}

// Forward 3-D FFT in place over a full L grid (spectrum construction).
static int fft3_forward_inplace(peec_ctx *ctx, double2_t *g, const int L[3])
{
    int64_t Lx = L[0], Ly = L[1], Lz = L[2];
    PassDesc d;
    int rc;
    // x: lines contiguous, b = y, l2 = z
    PassGeom gx = {L[0], L[0], L[0], L[1], L[2], Lx, Lx * Ly, 1, Lx, Lx * Ly, 1};
    if ((rc = make_desc(ctx, gx, &d))) return rc;
    if (L[0] > 1 && (rc = launch_pass(ctx, false, 0, d, g, g))) return rc;
    PassGeom gy = {L[1], L[1], L[1], L[0], L[2], 1, Lx * Ly, Lx, 1, Lx * Ly, Lx};
    if ((rc = make_desc(ctx, gy, &d))) return rc;
    if (L[1] > 1 && (rc = launch_pass(ctx, false, 0, d, g, g))) return rc;
    PassGeom gz = {L[2], L[2], L[2], L[0], L[1], 1, Lx, Lx * Ly, 1, Lx, Lx * Ly};
    if ((rc = make_desc(ctx, gz, &d))) return rc;
    if (L[2] > 1 && (rc = launch_pass(ctx, false, 0, d, g, g))) return rc;
    return PEEC_OK;
}

// Axes along which a kernel's circulant grid has half-cell spacing: the
// merged face grids of the off-diagonal P blocks (p1234, p1256, p3456,
// build_center_P_Voxels.jl:53-111).
extern const int kHalfAxes[9][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0},
                             {0, 0, 0}, {1, 1, 0}, {0, 1, 1}, {1, 0, 1}};

// The complex spectrum of kernel k's half-cell reduction for parity offset
// delta (out - in, per half-cell axis).
int poly_spectrum_build(peec_ctx *ctx, int k, const int delta[3])
{
    Kernel &K = ctx->kern[k];
    for (int a = 0; a < 3; a++) {
        K.pn[a] = kHalfAxes[k][a] ? (K.n[a] + 1) / 2 : K.n[a];
        K.pL[a] = a == 2 ? fft_z_length(K.pn[a]) : fft_smooth_length(fft_embed_min(K.pn[a]));
        if (!register_fft_length(ctx, K.pL[a])) {
            peec_set_error("operator_set: twiddle table allocation failed");
            return PEEC_ERR_NOMEM;
        }
    }
    const int64_t G = (int64_t)K.pL[0] * K.pL[1] * K.pL[2];
    DevBuf<double2_t> grid;
    if (!grid.alloc((size_t)G) || !K.pspec64.alloc((size_t)G) || !K.pspec32.alloc((size_t)(2 * G))) {
        peec_set_error("operator_set: out of device memory for a reduced P spectrum");
        return PEEC_ERR_NOMEM;
    }
    const unsigned nbk = (unsigned)((G + 255) / 256);
    hipLaunchKernelGGL(k_embed_poly, dim3(nbk), dim3(256), 0, ctx->stream, K.row.p, K.n[0], K.n[1], K.n[2],
                       kHalfAxes[k][0], kHalfAxes[k][1], kHalfAxes[k][2], delta[0], delta[1], delta[2], K.pn[0],
                       K.pn[1], K.pn[2], K.pL[0], K.pL[1], K.pL[2], grid.p);
    PEEC_HIP_TRY(hipGetLastError());
    int rc = fft3_forward_inplace(ctx, grid.p, K.pL);
    if (rc) return rc;
    hipLaunchKernelGGL(k_spec_extract_c, dim3(nbk), dim3(256), 0, ctx->stream, grid.p, G, 1.0 / (double)G,
                       K.pspec64.p, K.pspec32.p);
    PEEC_HIP_TRY(hipGetLastError());
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int a = 0; a < 3; a++) K.delta[a] = delta[a];
    K.poly = true;
    return PEEC_OK;
}

int peec_kernels_build_impl(peec_ctx *ctx, const double origin[3], int Nx, int Ny, int Nz, double sx, double sy,
                            double sz, double esc_Lp, double esc_P)
{
    ctx->release_kernels();
    int64_t Gmax = 0;
    for (int k = 0; k < 9; k++) {
        Kernel &K = ctx->kern[k];
        kernel_dims(k, Nx, Ny, Nz, K.n);
        for (int a = 0; a < 3; a++)
            K.L[a] = K.n[a] > 0 ? (a == 2 ? fft_z_length(K.n[a]) : fft_smooth_length(fft_embed_min(K.n[a]))) : 0;
        if (K.compact() > 0) Gmax = std::max(Gmax, K.embedded());
    }
    // register every FFT length before the first launch: the twiddle table
    // is reallocated when a new length appears
    for (int k = 0; k < 9; k++)
        for (int a = 0; a < 3; a++)
            if (!register_fft_length(ctx, ctx->kern[k].L[a])) {
                peec_set_error("kernels_build: twiddle table allocation failed");
                return PEEC_ERR_NOMEM;
            }
    DevBuf<double2_t> grid;
    if (!grid.alloc((size_t)Gmax)) {
        peec_set_error("kernels_build: cannot allocate %lld-point work grid", (long long)Gmax);
        return PEEC_ERR_NOMEM;
    }
    for (int k = 0; k < 9; k++) {
        if (ctx->stop_requested()) return PEEC_STOPPED;
        Kernel &K = ctx->kern[k];
        K.fspec32.reset();  // plane-order copy: rebuilt by the next peec_operator_set
        int64_t g = K.compact();
        if (g <= 0) continue;
        if (!K.row.alloc((size_t)g) || !K.spec64.alloc((size_t)K.embedded()) || !K.spec32.alloc((size_t)K.embedded())) {
            peec_set_error("kernels_build: out of device memory");
            return PEEC_ERR_NOMEM;
        }
        int64_t n;
        int dims[3];
        int rc = peec_grid_row_device(ctx->stream, k, origin, Nx, Ny, Nz, sx, sy, sz, k <= 2 ? esc_Lp : esc_P, K.row.p,
                                      &n, dims);
        if (rc) return rc;
        int64_t G = K.embedded();
        hipLaunchKernelGGL(k_embed, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, ctx->stream, K.row.p, K.n[0],
                           K.n[1], K.n[2], K.L[0], K.L[1], K.L[2], grid.p);
        PEEC_HIP_TRY(hipGetLastError());
        if ((rc = fft3_forward_inplace(ctx, grid.p, K.L))) return rc;
        hipLaunchKernelGGL(k_spec_extract, dim3((unsigned)((G + 255) / 256)), dim3(256), 0, ctx->stream, grid.p, G,
                           1.0 / (double)G, K.spec64.p, K.spec32.p);
        PEEC_HIP_TRY(hipGetLastError());
    }
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->kernels_ready = true;
    return PEEC_OK;
}

// ---- retarded (Rcc) kernels
//
// compute_rows_Rcc_Lp.jl / compute_rows_Rcc_P.jl (compute_FFT_mutual_coupling_mats.jl
// with QS_Rcc_FW = 2): per geometry the nine static rows (the Lp rows from
// compute_Lp_Voxels_QS.jl's closed forms, no escaling, self entry zeroed)
// and their centre distances.  The quasi-static rows and spectra are built
// as well: the operator set-up reads them.
int peec_kernels_build_rcc_impl(peec_ctx *ctx, const double origin[3], int Nx, int Ny, int Nz, double sx, double sy,
                                double sz, double esc_Lp, double esc_P)
{
    int rc = peec_kernels_build_impl(ctx, origin, Nx, Ny, Nz, sx, sy, sz, esc_Lp, esc_P);
    if (rc) return rc;
    for (int k = 0; k < 9; k++) {
        Kernel &K = ctx->kern[k];
        const int64_t g = K.compact();
        if (g <= 0) continue;
        if (!K.rrow.alloc((size_t)g) || !K.rdist.alloc((size_t)g)) {
            peec_set_error("kernels_build_rcc: out of device memory");
            return PEEC_ERR_NOMEM;
        }
        int64_t n;
        int dims[3];
        if ((rc = peec_grid_row_device(ctx->stream, k, origin, Nx, Ny, Nz, sx, sy, sz, 1.0, K.rrow.p, &n, dims, 1)) ||
            (rc = peec_grid_row_device(ctx->stream, k, origin, Nx, Ny, Nz, sx, sy, sz, 1.0, K.rdist.p, &n, dims, 2)))
            return rc;
    }
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->rcc = true;
    ctx->rcc_ready = false;
    ctx->rcc_esc_Lp = esc_Lp;
    ctx->rcc_esc_P = esc_P;
    return PEEC_OK;
}

// compute_Circulant_Lp_Rcc.jl / compute_Circulant_P_sup_Rcc.jl at one
// frequency (Hz): the complex spectrum of escaling * row .* exp(-j beta Rcc)
// per kernel, beta = 2 pi f sqrt(eps0 mu0) (build_row_vox_Rcc.jl:3-5), on
// the full embedding and, for a kernel the operator applies through a
// half-cell reduction, on the reduced box for both directions.
int peec_rcc_spectra_impl(peec_ctx *ctx, double freq_hz)
{
    if (!ctx->rcc || !ctx->kernels_ready) {
        peec_set_error("peec_rcc_spectra: build the retarded kernels first (peec_kernels_build_rcc)");
        return PEEC_ERR_STATE;
    }
    if (!ctx->op_ready) {
        peec_set_error("peec_rcc_spectra: set the operator first (peec_operator_set)");
        return PEEC_ERR_STATE;
    }
    const double pi = 3.141592653589793, eps0 = 8.854187816997944e-12, mu0 = 4 * pi * 1e-7;
    const double beta = 2 * pi * freq_hz * sqrt(eps0 * mu0);
    int64_t Gmax = 0;
    for (int k = 0; k < 9; k++) {
        const Kernel &K = ctx->kern[k];
        if (K.compact() <= 0) continue;
        Gmax = std::max(Gmax, K.embedded());
        if (K.poly) Gmax = std::max(Gmax, (int64_t)K.pL[0] * K.pL[1] * K.pL[2]);
    }
    DevBuf<double2_t> grid;
    if (!grid.alloc((size_t)std::max<int64_t>(Gmax, 1))) {
        peec_set_error("peec_rcc_spectra: cannot allocate a %lld-point work grid", (long long)Gmax);
        return PEEC_ERR_NOMEM;
    }
    int rc;
    for (int k = 0; k < 9; k++) {
        Kernel &K = ctx->kern[k];
        if (K.compact() <= 0) continue;
        const double esc = k <= 2 ? ctx->rcc_esc_Lp : ctx->rcc_esc_P;
        const int64_t G = K.embedded();
        if (!K.rspec64.alloc((size_t)G) || !K.rspec32.alloc((size_t)(2 * G))) {
            peec_set_error("peec_rcc_spectra: out of device memory");
            return PEEC_ERR_NOMEM;
        }
        const unsigned nb = (unsigned)((G + 255) / 256);
        hipLaunchKernelGGL(k_embed_rcc, dim3(nb), dim3(256), 0, ctx->stream, K.rrow.p, K.rdist.p, -beta, esc, K.n[0],
                           K.n[1], K.n[2], K.L[0], K.L[1], K.L[2], grid.p);
        PEEC_HIP_TRY(hipGetLastError());
        if ((rc = fft3_forward_inplace(ctx, grid.p, K.L))) return rc;
        hipLaunchKernelGGL(k_spec_extract_c, dim3(nb), dim3(256), 0, ctx->stream, grid.p, G, 1.0 / (double)G,
                           K.rspec64.p, K.rspec32.p);
        PEEC_HIP_TRY(hipGetLastError());
        if (!K.poly) continue;
        const int64_t Gp = (int64_t)K.pL[0] * K.pL[1] * K.pL[2];
        const unsigned nbp = (unsigned)((Gp + 255) / 256);
        for (int dir = 0; dir < 2; dir++) {
            const int sg = dir == 0 ? 1 : -1;
            if (!K.rpspec64[dir].alloc((size_t)Gp) || !K.rpspec32[dir].alloc((size_t)(2 * Gp))) {
                peec_set_error("peec_rcc_spectra: out of device memory");
                return PEEC_ERR_NOMEM;
            }
            hipLaunchKernelGGL(k_embed_poly_rcc, dim3(nbp), dim3(256), 0, ctx->stream, K.rrow.p, K.rdist.p, -beta,
                               esc, K.n[0], K.n[1], K.n[2], kHalfAxes[k][0], kHalfAxes[k][1], kHalfAxes[k][2],
                               sg * K.delta[0], sg * K.delta[1], sg * K.delta[2], K.pn[0], K.pn[1], K.pn[2], K.pL[0],
                               K.pL[1], K.pL[2], grid.p);
            PEEC_HIP_TRY(hipGetLastError());
            if ((rc = fft3_forward_inplace(ctx, grid.p, K.pL))) return rc;
            hipLaunchKernelGGL(k_spec_extract_c, dim3(nbp), dim3(256), 0, ctx->stream, grid.p, Gp, 1.0 / (double)Gp,
                               K.rpspec64[dir].p, K.rpspec32[dir].p);
            PEEC_HIP_TRY(hipGetLastError());
        }
    }
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->rcc_ready = true;
    ctx->rcc_freq = freq_hz;
    return PEEC_OK;
}

int64_t conv_work_elems(const ConvApp &app);

// The five pass descriptors of one convolution application for nrhs
// right-hand sides (vectors at stride ld), with their algorithmic bytes.
struct ConvPasses {
    PassDesc d[5];
    const void *src[5];
    void *dst[5];
    double bytes[5];
};

static int conv_passes(peec_ctx *ctx, const ConvApp &app, bool fp32, const double2_t *vec_in, double2_t *vec_out,
                       int nrhs, int64_t ld, void *w1, void *w2, const int style[5], ConvPasses *cp)
{
    const Kernel &K = ctx->kern[app.kernel];
    const int nx = app.n[0], ny = app.n[1], nz = app.n[2];
    const int64_t Lx = app.L[0], Ly = app.L[1], Lz = app.L[2];
    const int64_t s1 = Lx * ny * nz, s2 = Lx * Ly * nz;  // per-rhs grid sizes
    const double e = fp32 ? 8.0 : 16.0, es = (fp32 ? 4.0 : 8.0) * (app.poly || ctx->rcc ? 2.0 : 1.0);
    const double g = (double)nx * ny * nz, G1 = (double)s1 * nrhs, G2 = (double)s2 * nrhs;
    int rc;
    // 1: x forward with scatter; lines (y, z)
    PassGeom g1 = {(int)Lx, nx, (int)Lx, ny, nz, nx, (int64_t)nx * ny, 1, Lx, Lx * ny, 1};
    PassDesc &d1 = cp->d[0];
    if ((rc = make_desc(ctx, g1, &d1, style[0]))) return rc;
    d1.map_in = app.map_in.p;
    d1.vec_in = vec_in + app.in_off;
    d1.vec_in_ld = ld;
    d1.nrhs = nrhs;
    d1.in_rhs = (int64_t)nx * ny * nz;  // only used to strip the rhs offset from map positions
    d1.out_rhs = s1;
    cp->src[0] = nullptr;
    cp->dst[0] = w1;
    // algorithmic bytes: every pass reads its input once and writes its
    // output once; the scatter/gather passes also read the 0/1 position map
    // and the unknowns it selects (the gather read-modify-writes them)
    cp->bytes[0] = 4.0 * g + 16.0 * app.nnz_in * nrhs + e * G1;
    // 2: y forward; lines (x', z)
    PassGeom g2 = {(int)Ly, ny, (int)Ly, (int)Lx, nz, 1, Lx * ny, Lx, 1, Lx * Ly, Lx};
    PassDesc &d2 = cp->d[1];
    if ((rc = make_desc(ctx, g2, &d2, style[1]))) return rc;
    d2.nrhs = nrhs;
    d2.in_rhs = s1;
    d2.out_rhs = s2;
    cp->src[1] = w1;
    cp->dst[1] = w2;
    cp->bytes[1] = e * (G1 + G2);
    // 3: z forward * spectrum * z inverse, in place; lines (x', y')
    PassGeom g3 = {(int)Lz, nz, nz, (int)Lx, (int)Ly, 1, Lx, Lx * Ly, 1, Lx, Lx * Ly};
    PassDesc &d3 = cp->d[2];
    if ((rc = make_desc(ctx, g3, &d3, style[2]))) return rc;
    d3.nrhs = nrhs;
    d3.in_rhs = s2;
    d3.out_rhs = s2;
    if (ctx->rcc) {
        // retarded kernels: complex spectra, one per direction of a reduction
        const int dir = app.conj ? 1 : 0;
        if (app.poly)
            d3.spec = fp32 ? (const void *)K.rpspec32[dir].p : (const void *)K.rpspec64[dir].p;
        else
            d3.spec = fp32 ? (const void *)K.rspec32.p : (const void *)K.rspec64.p;
        d3.spec_conj = 0;
    } else {
        if (app.poly)
            d3.spec = fp32 ? (const void *)K.pspec32.p : (const void *)K.pspec64.p;
        else
            d3.spec = fp32 ? (const void *)K.spec32.p : (const void *)K.spec64.p;
        d3.spec_conj = app.conj ? 1 : 0;
    }
    d3.spec_sb = 1;
    d3.spec_s2 = Lx;
    d3.spec_sj = Lx * Ly;
    cp->src[2] = w2;
    cp->dst[2] = w2;
    cp->bytes[2] = 2.0 * e * G2 + es * (double)app.embedded();
    // 4: y inverse, keep ny rows; lines (x', z)
    PassGeom g4 = {(int)Ly, (int)Ly, ny, (int)Lx, nz, 1, Lx * Ly, Lx, 1, Lx * ny, Lx};
    PassDesc &d4 = cp->d[3];
    if ((rc = make_desc(ctx, g4, &d4, style[3]))) return rc;
    d4.nrhs = nrhs;
    d4.in_rhs = s2;
    d4.out_rhs = s1;
    cp->src[3] = w2;
    cp->dst[3] = w1;
    cp->bytes[3] = e * (G1 + G2);
    // 5: x inverse with gather; lines (y, z)
    PassGeom g5 = {(int)Lx, (int)Lx, nx, ny, nz, Lx, Lx * ny, 1, nx, (int64_t)nx * ny, 1};
    PassDesc &d5 = cp->d[4];
    if ((rc = make_desc(ctx, g5, &d5, style[4]))) return rc;
    d5.nrhs = nrhs;
    d5.in_rhs = s1;
    d5.out_rhs = (int64_t)nx * ny * nz;
    d5.map_out = app.map_out.p;
    d5.vec_out = vec_out + app.out_off;
    d5.vec_out_ld = ld;
    cp->src[4] = w1;
    cp->dst[4] = nullptr;
    cp->bytes[4] = e * G1 + 4.0 * g + 32.0 * app.nnz_out * nrhs;
    return PEEC_OK;
}

// The three pass descriptors of one application through the fused y-z
// pass (complex64): x forward with scatter writing planes [x'][z][y], the
// fused pass in place on the planes, x inverse with gather reading them.
static int conv_passes_fused(peec_ctx *ctx, const ConvApp &app, const double2_t *vec_in, double2_t *vec_out,
                             int nrhs, int64_t ld, void *w1, const int style[3], ConvPasses *cp)
{
    const Kernel &K = ctx->kern[app.kernel];
    const int nx = app.n[0], ny = app.n[1], nz = app.n[2];
    const int64_t Lx = app.L[0], Ly = app.L[1];
    const int64_t pl = (int64_t)ny * nz, s1 = Lx * pl;  // plane size, per-rhs grid size
    const double e = 8.0, es = 4.0 * (app.poly ? 2.0 : 1.0);
    const double g = (double)nx * ny * nz, G1 = (double)s1 * nrhs;
    int rc;
    PassGeom g1 = {(int)Lx, nx, (int)Lx, ny, nz, nx, (int64_t)nx * ny, 1, 1, ny, pl};
    PassDesc &d1 = cp->d[0];
    if ((rc = make_desc(ctx, g1, &d1, style[0]))) return rc;
    d1.map_in = app.map_in.p;
    d1.vec_in = vec_in + app.in_off;
    d1.vec_in_ld = ld;
    d1.nrhs = nrhs;
    d1.in_rhs = (int64_t)nx * ny * nz;
    d1.out_rhs = s1;
    cp->src[0] = nullptr;
    cp->dst[0] = w1;
    cp->bytes[0] = 4.0 * g + 16.0 * app.nnz_in * nrhs + e * G1;
    PassGeom g2 = {(int)Ly, ny, ny, (int)Lx, nz, 0, pl, 0, 0, pl, 0};
    PassDesc &d2 = cp->d[1];
    if ((rc = make_desc(ctx, g2, &d2, style[1]))) return rc;
    d2.nrhs = nrhs;
    d2.in_rhs = s1;
    d2.out_rhs = s1;
    d2.spec = app.poly ? (const void *)K.pfspec32.p : (const void *)K.fspec32.p;
    d2.spec_conj = app.conj ? 1 : 0;
    d2.tw_off_z = ctx->twiddle_offset(app.L[2]);
    if (d2.tw_off_z < 0) {
        peec_set_error("twiddle table allocation failed");
        return PEEC_ERR_NOMEM;
    }
    cp->src[1] = w1;
    cp->dst[1] = w1;
    cp->bytes[1] = 2.0 * e * G1 + es * (double)app.embedded();
    PassGeom g3 = {(int)Lx, (int)Lx, nx, ny, nz, 1, ny, pl, nx, (int64_t)nx * ny, 1};
    PassDesc &d3 = cp->d[2];
    if ((rc = make_desc(ctx, g3, &d3, style[2]))) return rc;
    d3.nrhs = nrhs;
    d3.in_rhs = s1;
    d3.out_rhs = (int64_t)nx * ny * nz;
    d3.map_out = app.map_out.p;
    d3.vec_out = vec_out + app.out_off;
    d3.vec_out_ld = ld;
    cp->src[2] = w1;
    cp->dst[2] = nullptr;
    cp->bytes[2] = e * G1 + 4.0 * g + 32.0 * app.nnz_out * nrhs;
    return PEEC_OK;
}

// Right-hand sides per block when untuned: enough blocks (about 2048) to
// fill the chip; the rest are pipelined through each block (and share its
// spectrum tile in the fused z pass).
static int default_rpb(const PassDesc &d, int v)
{
    const int64_t tiles = (int64_t)((d.nb + variant_lines(v) - 1) / variant_lines(v)) * d.n2;
    return (int)std::max<int64_t>(1, std::min<int64_t>(16, tiles * d.nrhs / 2048));
}

static int default_lane_rpb(const PassDesc &d, int Q)
{
    const int lpb = 256 / Q;
    const int64_t tiles = (int64_t)((d.nb + lpb - 1) / lpb) * d.n2;
    return (int)std::max<int64_t>(1, std::min<int64_t>(16, tiles * d.nrhs / 4096));
}

// The tuned choice for exactly nrhs right-hand sides, else the one tuned for
// the nearest width within a factor 4 (GMRES applies the matvec to the
// columns still iterating, whose count falls below the tuned width): its launch
// shape and radix schedule, with right-hand sides per block scaled to keep
// the number of blocks.  Launch shapes do not change results.
static FftChoice tuned_choice(const peec_ctx *ctx, int kernel, int pass, bool fp32, int nrhs)
{
    auto it = ctx->fft_tune.find(FftTuneKey{kernel, pass, fp32 ? 1 : 0, nrhs});
    if (it != ctx->fft_tune.end()) return it->second;
    const FftChoice *best = nullptr;
    int bw = 0;
    for (const auto &kv : ctx->fft_tune) {
        const FftTuneKey &k = kv.first;
        if (k.kernel != kernel || k.pass != pass || k.fp32 != (fp32 ? 1 : 0)) continue;
        if (!best || std::abs(k.nrhs - nrhs) < std::abs(bw - nrhs) || (std::abs(k.nrhs - nrhs) == std::abs(bw - nrhs) && k.nrhs > bw)) {
            best = &kv.second;
            bw = k.nrhs;
        }
    }
    if (!best || nrhs * 4 < bw || nrhs > 4 * bw) return FftChoice();
    FftChoice c = *best;
    c.rpb = std::max(1, (int)(((int64_t)c.rpb * nrhs + bw / 2) / std::max(bw, 1)));
    return c;
}

static const int kProfKind[5] = {PK_FFT_X_FWD, PK_FFT_Y_FWD, PK_FFT_Z_CONV, PK_FFT_Y_INV, PK_FFT_X_INV};
static const int kPassMode[5] = {0, 0, 2, 1, 1};
// the z pass multiplies by a complex spectrum for a half-cell reduction and
// for every retarded kernel
static int pass_mode(const peec_ctx *ctx, const ConvApp &app, int p)
{
    return p == 2 && (app.poly || ctx->rcc) ? 3 : kPassMode[p];
}

// Tuning-table passes of the fused pipeline: 5 x forward (plane output),
// 6 the fused y-z pass (variant 0: use the fused pipeline, 1: the five
// passes), 7 x inverse (plane input).
static constexpr int kPassXFwdPlanes = 5, kPassYz = 6, kPassXInvPlanes = 7;

static int default_yz_rpb(const PassDesc &d)
{
    return (int)std::max<int64_t>(1, std::min<int64_t>(16, (int64_t)d.nb * d.nrhs / 2048));
}

// The fused pipeline runs when it fits, its plane-order spectrum exists,
// and the tuner (or peec_fft_set_choice) chose it for this nrhs; untuned,
// the five passes run.  On the 96 x 96 x 24 workload (z lengths 48 and 64,
// 64 right-hand sides) it measured 1.3-2.1x slower than the five passes:
// its z registers (spectrum + line) hold it to one wave per SIMD, and a
// block's LDS stages and barriers then leave HBM idle.
static bool use_fused(const peec_ctx *ctx, const ConvApp &app, bool fp32, int nrhs)
{
    if (ctx->rcc || !yz_fits(app, fp32)) return false;  // the fused pass reads the static spectra
    const Kernel &K = ctx->kern[app.kernel];
    if (app.poly ? !K.pfspec32.p : !K.fspec32.p) return false;
    return tuned_choice(ctx, app.kernel, kPassYz, fp32, nrhs).variant == 0;
}

static int conv_apply_fused(peec_ctx *ctx, const ConvApp &app, const double2_t *vec_in, double2_t *vec_out, int nrhs,
                            int64_t ld, void *w1)
{
    FftChoice ch[3];
    int style[3];
    for (int p = 0; p < 3; p++) {
        ch[p] = tuned_choice(ctx, app.kernel, kPassXFwdPlanes + p, true, nrhs);
        style[p] = ch[p].style;
    }
    ConvPasses cp;
    int rc = conv_passes_fused(ctx, app, vec_in, vec_out, nrhs, ld, w1, style, &cp);
    if (rc) return rc;
    for (int p = 0; p < 3; p += 2) {
        PassDesc &d = cp.d[p];
        const int v = ch[p].variant >= 0 ? ch[p].variant : default_variant(d.L, true);
        d.rpb = ch[p].variant >= 0 ? ch[p].rpb : default_rpb(d, std::max(v, 0));
        const int pi = ctx->prof_begin(p == 0 ? PK_FFT_X_FWD : PK_FFT_X_INV, cp.bytes[p]);
        if ((rc = launch_pass(ctx, true, p == 0 ? 0 : 1, d, cp.src[p], cp.dst[p], v))) return rc;
        ctx->prof_end(pi);
        if (p == 0) {
            PassDesc &y = cp.d[1];
            y.rpb = ch[1].variant >= 0 ? ch[1].rpb : default_yz_rpb(y);
            const int pj = ctx->prof_begin(PK_FFT_YZ_CONV, cp.bytes[1]);
            if ((rc = launch_yz(ctx, pass_mode(ctx, app, 2), y, cp.src[1], cp.dst[1], app.L[2]))) return rc;
            ctx->prof_end(pj);
        }
    }
    return PEEC_OK;
}

// One convolution application: vec_out[section] += gather(conv(scatter(vec_in)))
// for nrhs right-hand sides (vectors at stride ld).  w1/w2 are grid work
// buffers sized by conv_work_elems.  Launch shapes and radix schedules come
// from peec_fft_tune when it has run for this nrhs.
int conv_apply(peec_ctx *ctx, const ConvApp &app, bool fp32, const double2_t *vec_in, double2_t *vec_out,
               int nrhs, int64_t ld, void *w1, void *w2)
{
    if (app.compact() <= 0 || app.nnz_in == 0 || app.nnz_out == 0) return PEEC_OK;
    if (ctx->rcc && !ctx->rcc_ready) {
        peec_set_error("retarded kernels: build the spectra of a frequency first (peec_rcc_spectra)");
        return PEEC_ERR_STATE;
    }
    if (use_fused(ctx, app, fp32, nrhs)) return conv_apply_fused(ctx, app, vec_in, vec_out, nrhs, ld, w1);
    FftChoice ch[5];
    int style[5];
    for (int p = 0; p < 5; p++) {
        ch[p] = tuned_choice(ctx, app.kernel, p, fp32, nrhs);
        style[p] = ch[p].style;
    }
    ConvPasses cp;
    int rc = conv_passes(ctx, app, fp32, vec_in, vec_out, nrhs, ld, w1, w2, style, &cp);
    if (rc) return rc;
    for (int p = 0; p < 5; p++) {
        PassDesc &d = cp.d[p];
        const bool reg_ok = p == 2 && reg_length(d.L, fp32) && d.radix[0] == RegFirstRadix(d.L);
        const int v = ch[p].variant >= 0 ? ch[p].variant : (reg_ok ? kRegVariant : default_variant(d.L, fp32));
        d.rpb = ch[p].variant >= 0 ? ch[p].rpb : default_rpb(d, v == kRegVariant ? 0 : std::max(v, 0));
        const int pi = ctx->prof_begin(kProfKind[p], cp.bytes[p]);
        if ((rc = launch_pass(ctx, fp32, pass_mode(ctx, app, p), d, cp.src[p], cp.dst[p], v))) return rc;
        ctx->prof_end(pi);
    }
    return PEEC_OK;
}

// Time `reps` launches of one pass (after one warm-up launch), in ms each.
static int time_pass(peec_ctx *ctx, bool fp32, int mode, PassDesc d, const void *src, void *dst, int v, int reps,
                     hipEvent_t a, hipEvent_t b, float *ms)
{
    int rc;
    if ((rc = launch_pass(ctx, fp32, mode, d, src, dst, v))) return rc;
    PEEC_HIP_TRY(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < reps; i++)
        if ((rc = launch_pass(ctx, fp32, mode, d, src, dst, v))) return rc;
    PEEC_HIP_TRY(hipEventRecord(b, ctx->stream));
    PEEC_HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    PEEC_HIP_TRY(hipEventElapsedTime(&t, a, b));
    *ms = t / (float)reps;
    return PEEC_OK;
}

static int time_yz(peec_ctx *ctx, int mode, PassDesc d, const void *src, void *dst, int Lz, int reps,
                   hipEvent_t a, hipEvent_t b, float *ms)
{
    int rc;
    if ((rc = launch_yz(ctx, mode, d, src, dst, Lz))) return rc;
    PEEC_HIP_TRY(hipEventRecord(a, ctx->stream));
    for (int i = 0; i < reps; i++)
        if ((rc = launch_yz(ctx, mode, d, src, dst, Lz))) return rc;
    PEEC_HIP_TRY(hipEventRecord(b, ctx->stream));
    PEEC_HIP_TRY(hipEventSynchronize(b));
    float t = 0.f;
    PEEC_HIP_TRY(hipEventElapsedTime(&t, a, b));
    *ms = t / (float)reps;
    return PEEC_OK;
}

// Tune one pass: 1) launch shape at the default radix schedule and rhs
// grouping (and the register z kernel where allowed), 2) right-hand sides
// per block, 3) radix schedule (FP32 passes only: the FP64 path keeps one
// fixed operation order; the register kernel unrolls the radix-8-first
// schedule).
static int tune_pass(peec_ctx *ctx, bool fp32, int mode, PassDesc d, const void *src, void *dst, double bytes,
                     bool allow_reg, bool verbose, int kernel, int pass, hipEvent_t ea, hipEvent_t eb, FftChoice *out,
                     float *tout)
{
    FftChoice best;
    float tbest = 1e30f;
    int rc = PEEC_OK;
    for (int v = 0; v < kNumFftVariants && rc == PEEC_OK; v++) {
        if (!variant_fits(v, d.L, fp32)) continue;
        d.rpb = default_rpb(d, v);
        float t;
        if ((rc = time_pass(ctx, fp32, mode, d, src, dst, v, 3, ea, eb, &t))) break;
        if (verbose)
            fprintf(stderr, "peec_fft_tune: kernel %d pass %d %s L %d nin %d nout %d: shape %d (W %d TPL %d) "
                            "rpb %d: %.1f us, %.0f GB/s\n", kernel, pass, fp32 ? "c64" : "c128", d.L, d.nin, d.nout, v,
                    variant_lines(v), kFftVariants[v].tpl, d.rpb, 1e3 * t, bytes / (1e6 * t));
        if (t < tbest) {
            tbest = t;
            best.variant = v;
            best.rpb = d.rpb;
            best.style = 0;
        }
    }
    if (rc) return rc;
    if (allow_reg && mode >= 2 && reg_length(d.L, fp32) && d.radix[0] == RegFirstRadix(d.L)) {
        d.rpb = default_rpb(d, 0);
        float t;
        if ((rc = time_pass(ctx, fp32, mode, d, src, dst, kRegVariant, 3, ea, eb, &t))) return rc;
        if (verbose)
            fprintf(stderr, "peec_fft_tune: kernel %d pass %d %s L %d: register z kernel rpb %d: %.1f us, %.0f GB/s\n",
                    kernel, pass, fp32 ? "c64" : "c128", d.L, d.rpb, 1e3 * t, bytes / (1e6 * t));
        if (t < tbest) {
            tbest = t;
            best.variant = kRegVariant;
            best.rpb = d.rpb;
            best.style = 0;
        }
    }
    // (the x passes into and out of planes, passes 5 and 7, keep the LDS
    // kernel: the lane kernel is tested on the five-pass layout only)
    for (int k = 0; k < kNumLaneV && pass != kPassXFwdPlanes && pass != kPassXInvPlanes; k++) {
        const int Q = lane_q(kLaneBase + k);
        const bool xl = lane_xl(kLaneBase + k);
        if (!lane_fits(d, Q, mode, xl, fp32)) continue;
        d.rpb = default_lane_rpb(d, Q);
        float t;
        if ((rc = time_pass(ctx, fp32, mode, d, src, dst, kLaneBase + k, 3, ea, eb, &t))) return rc;
        if (verbose)
            fprintf(stderr, "peec_fft_tune: kernel %d pass %d %s L %d: lane-split kernel Q %d%s rpb %d: %.1f us, %.0f GB/s\n",
                    kernel, pass, fp32 ? "c64" : "c128", d.L, Q, xl ? " (LDS exchanges)" : "", d.rpb, 1e3 * t,
                    bytes / (1e6 * t));
        if (t < tbest) {
            tbest = t;
            best.variant = kLaneBase + k;
            best.rpb = d.rpb;
            best.style = 0;
        }
    }
    if (best.variant < 0) {
        peec_set_error("peec_fft_tune: no launch shape for length %d", d.L);
        return PEEC_ERR_ARG;
    }
    for (int r = 1; r <= 32 && r <= d.nrhs; r *= 2) {
        d.rpb = r;
        float t;
        if ((rc = time_pass(ctx, fp32, mode, d, src, dst, best.variant, 3, ea, eb, &t))) return rc;
        if (verbose) fprintf(stderr, "peec_fft_tune:   rpb %d: %.1f us, %.0f GB/s\n", r, 1e3 * t, bytes / (1e6 * t));
        if (t < tbest) {
            tbest = t;
            best.rpb = r;
        }
    }
    if (fp32 && best.variant != kRegVariant && !is_lane_variant(best.variant)) {
        for (int st = 1; st <= 2; st++) {
            PassDesc ds;
            PassGeom gg = {d.L, d.nin, d.nout, d.nb, d.n2, d.in_sb, d.in_s2, d.in_sj, d.out_sb, d.out_s2, d.out_sj};
            if ((rc = make_desc(ctx, gg, &ds, st))) return rc;
            if (ds.nstage == d.nstage && !memcmp(ds.radix, d.radix, sizeof(int) * (size_t)d.nstage)) continue;
            PassDesc dd = d;
            dd.nstage = ds.nstage;
            memcpy(dd.radix, ds.radix, sizeof(dd.radix));
            dd.rpb = best.rpb;
            float t;
            if ((rc = time_pass(ctx, fp32, mode, dd, src, dst, best.variant, 3, ea, eb, &t))) return rc;
            if (t < tbest) {
                tbest = t;
                best.style = st;
            }
        }
    }
    *out = best;
    *tout = tbest;
    return PEEC_OK;
}

extern "C" int peec_fft_tune(peec_ctx *ctx, int nrhs, int precisions, double *report_ms)
{
    if (!ctx || nrhs < 1 || (precisions & 3) == 0) {
        peec_set_error("peec_fft_tune: invalid argument");
        return PEEC_ERR_ARG;
    }
    const bool verbose = (precisions & 4) != 0, progress = (precisions & 12) != 0;
    if (!ctx->op_ready) {
        peec_set_error("peec_fft_tune: call peec_operator_set first");
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    const int64_t N = ctx->m + ctx->ns + ctx->n;
    size_t need = 0;
    for (int a = 0; a < ctx->napps; a++) need = std::max<size_t>(need, (size_t)conv_work_elems(ctx->apps[a]));
    need *= (size_t)nrhs;
    if (need > ctx->w1.n && (!ctx->w1.alloc(need) || !ctx->w2.alloc(need))) {
        peec_set_error("peec_fft_tune: out of device memory for %d-rhs FFT work grids", nrhs);
        return PEEC_ERR_NOMEM;
    }
    DevBuf<double2_t> vin, vout;
    if (!vin.alloc((size_t)(N * nrhs)) || !vout.alloc((size_t)(N * nrhs))) {
        peec_set_error("peec_fft_tune: out of device memory for %d right-hand sides", nrhs);
        return PEEC_ERR_NOMEM;
    }
    PEEC_HIP_TRY(hipMemsetAsync(vin.p, 0, vin.n * sizeof(double2_t), ctx->stream));
    PEEC_HIP_TRY(hipMemsetAsync(vout.p, 0, vout.n * sizeof(double2_t), ctx->stream));
    PEEC_HIP_TRY(hipMemsetAsync(ctx->w1.p, 0, ctx->w1.n * sizeof(double2_t), ctx->stream));
    PEEC_HIP_TRY(hipMemsetAsync(ctx->w2.p, 0, ctx->w2.n * sizeof(double2_t), ctx->stream));
    hipEvent_t ea, eb;
    PEEC_HIP_TRY(hipEventCreate(&ea));
    PEEC_HIP_TRY(hipEventCreate(&eb));
    double total_best = 0.0;
    int rc = PEEC_OK;
    bool seen[9] = {};
    for (int a = 0; a < ctx->napps && rc == PEEC_OK; a++) {
        const ConvApp &app = ctx->apps[a];
        if (seen[app.kernel] || app.compact() <= 0 || app.nnz_in == 0 || app.nnz_out == 0) continue;
        seen[app.kernel] = true;
        for (int prec = 0; prec < 2 && rc == PEEC_OK; prec++) {
            if (!(precisions & (1 << prec))) continue;
            const bool fp32 = prec == 1;
            const int style0[5] = {0, 0, 0, 0, 0};
            ConvPasses cp;
            if ((rc = conv_passes(ctx, app, fp32, vin.p, vout.p, nrhs, N, ctx->w1.p, ctx->w2.p, style0, &cp))) break;
            float t5 = 0.f;
            for (int p = 0; p < 5 && rc == PEEC_OK; p++) {
                FftChoice best;
                float tbest = 1e30f;
                if ((rc = tune_pass(ctx, fp32, pass_mode(ctx, app, p), cp.d[p], cp.src[p], cp.dst[p], cp.bytes[p], p == 2,
                                    verbose, app.kernel, p, ea, eb, &best, &tbest)))
                    break;
                ctx->fft_tune[FftTuneKey{app.kernel, p, fp32 ? 1 : 0, nrhs}] = best;
                t5 += tbest;
            }
            if (rc) break;
            float tsum = t5;
            // the fused pipeline (x forward into planes, fused y-z, x inverse)
            const Kernel &K = ctx->kern[app.kernel];
            if (!ctx->rcc && yz_fits(app, fp32) && (app.poly ? K.pfspec32.p : K.fspec32.p)) {
                const int st3[3] = {0, 0, 0};
                ConvPasses cf;
                if ((rc = conv_passes_fused(ctx, app, vin.p, vout.p, nrhs, N, ctx->w1.p, st3, &cf))) break;
                FftChoice bx[2], by;
                float tx[2] = {1e30f, 1e30f}, ty = 1e30f;
                for (int q = 0; q < 2 && rc == PEEC_OK; q++)
                    rc = tune_pass(ctx, true, q == 0 ? 0 : 1, cf.d[2 * q], cf.src[2 * q], cf.dst[2 * q],
                                   cf.bytes[2 * q], false, verbose, app.kernel, q == 0 ? kPassXFwdPlanes : kPassXInvPlanes,
                                   ea, eb, &bx[q], &tx[q]);
                if (rc) break;
                for (int r = 1; r <= 32 && r <= nrhs && rc == PEEC_OK; r *= 2) {
                    PassDesc d = cf.d[1];
                    d.rpb = r;
                    float t;
                    if ((rc = time_yz(ctx, pass_mode(ctx, app, 2), d, cf.src[1], cf.dst[1], app.L[2], 3, ea, eb, &t))) break;
                    if (verbose)
                        fprintf(stderr, "peec_fft_tune: kernel %d fused y-z Ly %d nz %d Lz %d rpb %d: %.1f us, %.0f GB/s\n",
                                app.kernel, d.L, d.n2, app.L[2], r, 1e3 * t, cf.bytes[1] / (1e6 * t));
                    if (t < ty) {
                        ty = t;
                        by.rpb = r;
                    }
                }
                if (rc) break;
                const float t3 = tx[0] + ty + tx[1];
                by.variant = t3 < t5 ? 0 : 1;
                ctx->fft_tune[FftTuneKey{app.kernel, kPassXFwdPlanes, 1, nrhs}] = bx[0];
                ctx->fft_tune[FftTuneKey{app.kernel, kPassYz, 1, nrhs}] = by;
                ctx->fft_tune[FftTuneKey{app.kernel, kPassXInvPlanes, 1, nrhs}] = bx[1];
                if (verbose)
                    fprintf(stderr, "peec_fft_tune: kernel %d c64: five passes %.1f us, fused pipeline %.1f us -> %s\n",
                            app.kernel, 1e3 * t5, 1e3 * t3, t3 < t5 ? "fused" : "five passes");
                tsum = std::min(t3, t5);
            }
            total_best += tsum;
            if (progress && rc == PEEC_OK)
                fprintf(stderr, "peec_fft_tune: kernel %d %s tuned\n", app.kernel, fp32 ? "c64" : "c128");
        }
    }
    (void)hipEventDestroy(ea);
    (void)hipEventDestroy(eb);
    if (rc) return rc;
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (report_ms) *report_ms = total_best;
    return PEEC_OK;
}

extern "C" int peec_fft_set_choice(peec_ctx *ctx, int kernel, int pass, int fp32, int nrhs, int variant, int rpb,
                                   int style)
{
    const bool fused = pass >= kPassXFwdPlanes;
    if (!ctx || kernel < 0 || kernel > 8 || pass < 0 || pass > kPassXInvPlanes || nrhs < 1 || variant < -1 ||
        variant >= kLaneBase + kNumLaneV || rpb < 1 || style < 0 || style > 2 || (style != 0 && !fp32) ||
        (fused && !fp32) || (variant == kRegVariant && (pass != 2 || style != 0)) ||
        (is_lane_variant(variant) && (fused || style != 0)) || (pass == kPassYz && variant > 1)) {
        peec_set_error("peec_fft_set_choice: invalid argument");
        return PEEC_ERR_ARG;
    }
    const FftTuneKey key{kernel, pass, fp32 ? 1 : 0, nrhs};
    if (variant < 0) {
        ctx->fft_tune.erase(key);
        return PEEC_OK;
    }
    const ConvApp *app = nullptr;
    for (int a = 0; a < ctx->napps && !app; a++)
        if (ctx->apps[a].kernel == kernel) app = &ctx->apps[a];
    if (pass == kPassYz) {
        if (!app || (variant == 0 && !yz_fits(*app, true))) {
            peec_set_error("peec_fft_set_choice: the fused y-z pass does not apply to kernel %d", kernel);
            return PEEC_ERR_ARG;
        }
        ctx->fft_tune[key] = FftChoice{variant, rpb, style};
        return PEEC_OK;
    }
    const int axis = (pass == 0 || pass == 4 || fused) ? 0 : (pass == 2 ? 2 : 1);
    const int L = app ? app->L[axis] : 0;
    const bool fits = !app ? false
                      : variant == kRegVariant ? reg_length(L, fp32 != 0) && L % RegFirstRadix(L) == 0
                      : is_lane_variant(variant) ? (axis != 0 || fp32 != 0) && L % lane_q(variant) == 0 &&
                                                       lane_built(L, lane_q(variant), axis == 2 ? 2 : 0, lane_xl(variant),
                                                                  axis == 0)
                      : variant_fits(variant, L, fp32 != 0);
    if (!fits) {
        peec_set_error("peec_fft_set_choice: launch variant %d does not fit length %d", variant, L);
        return PEEC_ERR_ARG;
    }
    ctx->fft_tune[key] = FftChoice{variant, rpb, style};
    return PEEC_OK;
}

// Launch shape chosen for one pass (for reports): variant lines, LMAX, TPL,
// rhs per block and radix schedule; -1 in variant when untuned.
extern "C" int peec_fft_choice(peec_ctx *ctx, int kernel, int pass, int fp32, int nrhs, int *out5)
{
    if (!ctx || kernel < 0 || kernel > 8 || pass < 0 || pass > kPassXInvPlanes || !out5) {
        peec_set_error("peec_fft_choice: invalid argument");
        return PEEC_ERR_ARG;
    }
    const FftChoice c = tuned_choice(ctx, kernel, pass, fp32 != 0, nrhs);
    const bool table = c.variant >= 0 && c.variant < kNumFftVariants && pass != kPassYz;
    out5[0] = c.variant;
    const bool lane = is_lane_variant(c.variant) && pass != kPassYz;
    out5[1] = table ? variant_lines(c.variant) : lane ? 256 / lane_q(c.variant) : (c.variant == kRegVariant ? 256 : 0);
    out5[2] = table ? kFftVariants[c.variant].tpl : lane ? lane_q(c.variant) : (c.variant == kRegVariant ? 1 : 0);
    out5[3] = c.rpb;
    out5[4] = c.style;
    return PEEC_OK;
}

int64_t conv_work_elems(const ConvApp &app) { return (int64_t)app.L[0] * app.L[1] * app.n[2]; }

extern "C" int peec_kernels_build(peec_ctx *ctx, const double origin[3], int Nx, int Ny, int Nz, double sx, double sy,
                                  double sz, double esc_Lp, double esc_P)
{
    if (!ctx || !origin || Nx < 1 || Ny < 1 || Nz < 1 || !(sx > 0) || !(sy > 0) || !(sz > 0)) {
        peec_set_error("peec_kernels_build: invalid argument");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    return peec_kernels_build_impl(ctx, origin, Nx, Ny, Nz, sx, sy, sz, esc_Lp, esc_P);
}

extern "C" int peec_kernels_build_rcc(peec_ctx *ctx, const double origin[3], int Nx, int Ny, int Nz, double sx,
                                      double sy, double sz, double esc_Lp, double esc_P)
{
    if (!ctx || !origin || Nx < 1 || Ny < 1 || Nz < 1 || !(sx > 0) || !(sy > 0) || !(sz > 0)) {
        peec_set_error("peec_kernels_build_rcc: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (ctx->nslots > 1) {
        peec_set_error("peec_kernels_build_rcc: the retarded spectra are per frequency; use one frequency slot");
        return PEEC_ERR_STATE;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    return peec_kernels_build_rcc_impl(ctx, origin, Nx, Ny, Nz, sx, sy, sz, esc_Lp, esc_P);
}

extern "C" int peec_rcc_spectra(peec_ctx *ctx, double freq_hz)
{
    if (!ctx || !(freq_hz >= 0) || !std::isfinite(freq_hz)) {
        peec_set_error("peec_rcc_spectra: invalid argument");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipSetDevice(ctx->device));
    return peec_rcc_spectra_impl(ctx, freq_hz);
}

extern "C" int peec_kernel_row(peec_ctx *ctx, int kind, double *out, int64_t *n)
{
    const int variant = kind >= 0 ? kind / 16 : -1;
    kind = kind >= 0 ? kind % 16 : kind;
    if (!ctx || kind < 0 || kind > 8 || variant < 0 || variant > 2 || !n) {
        peec_set_error("peec_kernel_row: invalid argument");
        return PEEC_ERR_ARG;
    }
    if (!ctx->kernels_ready || (variant > 0 && !ctx->rcc)) {
        peec_set_error("peec_kernel_row: call peec_kernels_build%s first", variant > 0 ? "_rcc" : "");
        return PEEC_ERR_STATE;
    }
    const Kernel &K = ctx->kern[kind];
    *n = K.compact();
    if (!out || *n <= 0) return PEEC_OK;
    const double *src = variant == 0 ? K.row.p : variant == 1 ? K.rrow.p : K.rdist.p;
    PEEC_HIP_TRY(hipMemcpyAsync(out, src, sizeof(double) * (size_t)*n, hipMemcpyDeviceToHost, ctx->stream));
    PEEC_HIP_TRY(hipStreamSynchronize(ctx->stream));
    return PEEC_OK;
}
