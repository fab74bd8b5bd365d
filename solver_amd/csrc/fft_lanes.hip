// Lane-split FFT passes: strided lines (the y passes and the fused z
// convolution) and the contiguous x lines with their scatter / gather maps,
// transformed in registers with Q lanes per line and cross-lane
// butterflies.  Part of the circulant convolution pipeline of fft.hip,
// which picks these kernels through its autotuner.
#include "fft_common.h"

#include <algorithm>
#include <type_traits>

namespace {

using namespace fftk;

// ---- strided-line passes with each line split over Q lanes
//
// Lines whose elements lie a plane or a row apart (the z and y passes; a
// block's tile runs along the contiguous x axis) are transformed in
// registers by Q lanes per line of length L = Q M: lane q holds the M
// elements n = q + Q m.  Forward: an M-point DFT per lane (Stockham stages,
// radix 8 first), the twiddle W_L^(q k2) of its output k2, then a Q-point
// DFT across the Q lanes by radix-2 decimation-in-frequency exchanges (DPP
// quad permutations for lane distances 1 and 2, ds_swizzle for 4 .. 16).
// Lane p then holds the transform's entries k2 + M rev(p) (rev: bit
// reversal over log2 Q bits), where MODE 2/3 take the spectrum product in
// place.  The inverse runs the mirror image -- decimation-in-time exchanges
// with conjugate twiddles, the conjugate W_L^(q k2), the inverse M-point
// DFT -- and ends in natural order.  MODE 0 (forward only) stores from the
// bit-reversed layout, MODE 1 (inverse only) loads into it.
//
// Variant XL replaces the exchanges by an LDS transpose per Q frequencies
// and Q-point DFTs in registers (xlds), which takes fewer VALU instructions.
//
// Registers and LDS: the twiddles of the per-lane DFT are compile-time
// constants (cx_twiddle, folded into the instructions); the lane twiddles
// are read from an LDS table filled once per block (the only barrier);
// addresses are a wave-uniform base plus one 32-bit lane offset.  What stays
// in registers for the block's rpb right-hand sides is the spectrum, and
// the next right-hand side's elements where they fit.
// exp(-2 pi i t / L) at compile time: octant reduction of t / L, Taylor
// series of sin and cos on [0, pi/4] (within 2 ulp of a long-double
// evaluation for every t, L <= 1024).
struct CxTw {
    double re, im;
};
constexpr double kCxPi = 3.141592653589793238462643383279502884;
constexpr double cx_sin_q(double x)
{
    double x2 = x * x, t = x, s = x;
    for (int n = 1; n < 12; n++) {
        t *= -x2 / (double)((2 * n) * (2 * n + 1));
        s += t;
    }
    return s;
}
constexpr double cx_cos_q(double x)
{
    double x2 = x * x, t = 1.0, s = 1.0;
    for (int n = 1; n < 12; n++) {
        t *= -x2 / (double)((2 * n - 1) * (2 * n));
        s += t;
    }
    return s;
}
constexpr CxTw cx_twiddle(int t, int L)
{
    const long long r = ((long long)t % L + L) % L;
    const long long a = 8 * r, o = a / L, f = a - o * L;  // angle 2 pi r / L = (pi / 4) (o + f / L)
    double c = 0.0, s = 0.0;
    if ((o & 1) == 0) {
        const double ph = (kCxPi / 4) * ((double)f / (double)L), cp = cx_cos_q(ph), sp = cx_sin_q(ph);
        switch ((o / 2) & 3) {
        case 0: c = cp; s = sp; break;
        case 1: c = -sp; s = cp; break;
        case 2: c = -cp; s = -sp; break;
        default: c = sp; s = -cp; break;
        }
    } else {
        const double ph = (kCxPi / 4) * ((double)(L - f) / (double)L), cp = cx_cos_q(ph), sp = cx_sin_q(ph);
        switch (((o + 1) / 2) & 3) {
        case 0: c = cp; s = -sp; break;
        case 1: c = sp; s = cp; break;
        case 2: c = -cp; s = sp; break;
        default: c = -sp; s = -cp; break;
        }
    }
    return {c, -s};
}

// exp(-2 pi i t / L) for t = 0 .. L-1, evaluated at compile time
template <int L>
struct CxTable {
    double re[L], im[L];
    constexpr CxTable() : re(), im()
    {
        for (int t = 0; t < L; t++) {
            const CxTw w = cx_twiddle(t, L);
            re[t] = w.re;
            im[t] = w.im;
        }
    }
};
template <typename T, int L>
__device__ __forceinline__ C2<T> cx_table(int t)
{
    static constexpr CxTable<L> tab{};
    return {(T)tab.re[t], (T)tab.im[t]};
}

// W_L^E (DIR < 0) or its conjugate as a constant of type T: the template
// arguments force the front end to evaluate cx_twiddle.
template <typename T, int DIR, int L, int E>
__device__ __forceinline__ C2<T> cxc()
{
    constexpr CxTw w = cx_twiddle(E, L);
    return {(T)w.re, (T)(DIR < 0 ? w.im : -w.im)};
}

template <typename T>
__device__ __forceinline__ T fma_t(T a, T b, T c) { return fma(a, b, c); }
template <>
__device__ __forceinline__ float fma_t<float>(float a, float b, float c) { return fmaf(a, b, c); }

// a * b and a * conj(b), each part one product and one fused multiply-add
template <typename T>
__device__ __forceinline__ C2<T> fmul(C2<T> a, C2<T> b) { return {fma_t(a.x, b.x, -(a.y * b.y)), fma_t(a.x, b.y, a.y * b.x)}; }
template <typename T>
__device__ __forceinline__ C2<T> fmulc(C2<T> a, C2<T> b) { return {fma_t(a.x, b.x, a.y * b.y), fma_t(a.y, b.x, -(a.x * b.y))}; }

// a * W_8^1 and a * W_8^3 (DIR < 0) or their conjugates: W_8^1 = h (1 - i),
// W_8^3 = -h (1 + i), h = 1/sqrt(2)
template <typename T, int DIR>
__device__ __forceinline__ C2<T> mul_w8_1(C2<T> a)
{
    const T h = (T)0.70710678118654752440084436210485;
    return DIR < 0 ? C2<T>{(a.x + a.y) * h, (a.y - a.x) * h} : C2<T>{(a.x - a.y) * h, (a.x + a.y) * h};
}
template <typename T, int DIR>
__device__ __forceinline__ C2<T> mul_w8_3(C2<T> a)
{
    const T h = (T)0.70710678118654752440084436210485;
    return DIR < 0 ? C2<T>{(a.y - a.x) * h, -((a.x + a.y) * h)} : C2<T>{-((a.x + a.y) * h), (a.x - a.y) * h};
}

// a * W_L^E (DIR < 0) or its conjugate for a compile-time exponent; the
// powers 0, L/4, L/2 and 3L/4 are exact sign changes and swaps, the powers
// L/8 and 3L/8 the two-product forms above
template <typename T, int DIR, int L, int E>
__device__ __forceinline__ C2<T> tw_mul(C2<T> a)
{
    constexpr int e = ((E % L) + L) % L;
    if constexpr (e == 0) return a;
    else if constexpr (4 * e == L) return rot<T, DIR>(a);
    else if constexpr (2 * e == L) return C2<T>{-a.x, -a.y};
    else if constexpr (4 * e == 3 * L) return rot<T, -DIR>(a);
    else if constexpr (8 * e == L) return mul_w8_1<T, DIR>(a);
    else if constexpr (8 * e == 3 * L) return mul_w8_3<T, DIR>(a);
    else return fmul(a, cxc<T, DIR, L, e>());
}

// v[r] *= W_L^(r BASE), r = 1 .. R-1
template <typename T, int DIR, int L, int R, int BASE, int r = 1>
__device__ __forceinline__ void cx_tw_apply(C2<T> (&v)[R])
{
    if constexpr (BASE > 0 && r < R) {
        v[r] = tw_mul<T, DIR, L, r * BASE>(v[r]);
        cx_tw_apply<T, DIR, L, R, BASE, r + 1>(v);
    }
}

// butterflies of radix 3, 5 and 8 with fused multiply-adds (radix 2 and 4
// take additions only)
template <typename T, int DIR>
__device__ __forceinline__ void fdft3(C2<T> *v)
{
    const T s = (T)0.86602540378443864676372317075294;
    const C2<T> b = cadd(v[1], v[2]);
    const C2<T> d = rot<T, DIR>(cscale(csub(v[1], v[2]), s));
    const C2<T> t = {fma_t((T)-0.5, b.x, v[0].x), fma_t((T)-0.5, b.y, v[0].y)};
    v[0] = cadd(v[0], b);
    v[1] = cadd(t, d);
    v[2] = csub(t, d);
}

template <typename T, int DIR>
__device__ __forceinline__ void fdft5(C2<T> *v)
{
    const T c1 = (T)0.30901699437494742410229341718282, c2 = (T)-0.80901699437494742410229341718282;
    const T s1 = (T)0.95105651629515357211643933337938, s2 = (T)0.58778525229247312916870595463907;
    const C2<T> b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    const C2<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    const C2<T> a0 = v[0];
    const C2<T> r1 = {fma_t(c2, b2.x, fma_t(c1, b1.x, a0.x)), fma_t(c2, b2.y, fma_t(c1, b1.y, a0.y))};
    const C2<T> r2 = {fma_t(c1, b2.x, fma_t(c2, b1.x, a0.x)), fma_t(c1, b2.y, fma_t(c2, b1.y, a0.y))};
    const C2<T> i1 = rot<T, DIR>(C2<T>{fma_t(s2, d2.x, s1 * d1.x), fma_t(s2, d2.y, s1 * d1.y)});
    const C2<T> i2 = rot<T, DIR>(C2<T>{fma_t(-s1, d2.x, s2 * d1.x), fma_t(-s1, d2.y, s2 * d1.y)});
    v[0] = cadd(a0, cadd(b1, b2));
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
}

template <typename T, int DIR>
__device__ __forceinline__ void fdft8(C2<T> *v)
{
    C2<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft4<T, DIR>(e);
    dft4<T, DIR>(o);
    o[1] = mul_w8_1<T, DIR>(o[1]);
    o[2] = rot<T, DIR>(o[2]);
    o[3] = mul_w8_3<T, DIR>(o[3]);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <typename T, int DIR, int R>
__device__ __forceinline__ void fdftR(C2<T> *v)
{
    if constexpr (R == 2) dft2<T, DIR>(v);
    else if constexpr (R == 3) fdft3<T, DIR>(v);
    else if constexpr (R == 4) dft4<T, DIR>(v);
    else if constexpr (R == 5) fdft5<T, DIR>(v);
    else fdft8<T, DIR>(v);
}

// DFT of radix R (2, 4 or 8) whose inputs v[R/2 .. R-1] are zero: the
// operations of the radix-R butterfly that do not involve those zeros
template <typename T, int DIR, int R>
__device__ __forceinline__ void dft_halfin(C2<T> *v)
{
    if constexpr (R == 2) {
        v[1] = v[0];
    } else if constexpr (R == 4) {
        const C2<T> a = v[0], b = v[1], rb = rot<T, DIR>(b);
        v[0] = cadd(a, b);
        v[2] = csub(a, b);
        v[1] = cadd(a, rb);
        v[3] = csub(a, rb);
    } else {
        C2<T> e[4] = {v[0], v[2], v[0], v[0]}, o[4] = {v[1], v[3], v[1], v[1]};
        dft_halfin<T, DIR, 4>(e);
        dft_halfin<T, DIR, 4>(o);
        o[1] = mul_w8_1<T, DIR>(o[1]);
        o[2] = rot<T, DIR>(o[2]);
        o[3] = mul_w8_3<T, DIR>(o[3]);
#pragma unroll
        for (int k = 0; k < 4; k++) {
            v[k] = cadd(e[k], o[k]);
            v[k + 4] = csub(e[k], o[k]);
        }
    }
}

// butterfly K of Stockham stage S of a length-L register FFT, a -> b; HIN:
// the inputs a[L/2 ..] are zero (stage 0 then runs half-input butterflies)
template <typename T, int DIR, int L, int S, int K, bool HIN>
__device__ __forceinline__ void cx_butterflies(const C2<T> (&a)[L], C2<T> (&b)[L])
{
    constexpr int R = RegPlan<L, S>::r, Ns = RegPlan<L, S>::ns, nb = L / R, stride = L / (Ns * R);
    if constexpr (K < nb) {
        constexpr int kk = K % Ns;
        C2<T> v[R];
        if constexpr (HIN && S == 0 && R % 2 == 0) {
#pragma unroll
            for (int r = 0; r < R / 2; r++) v[r] = a[K + r * nb];
            dft_halfin<T, DIR, R>(v);
        } else {
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = a[K + r * nb];
            cx_tw_apply<T, DIR, L, R, kk * stride>(v);
            fdftR<T, DIR, R>(v);
        }
#pragma unroll
        for (int r = 0; r < R; r++) b[(K - kk) * R + kk + r * Ns] = v[r];
        cx_butterflies<T, DIR, L, S, K + 1, HIN>(a, b);
    }
}

// Stockham stages of a length-L register FFT with compile-time twiddles
// (the schedule of RegPlan: radix 8 first, then 4, 2, 3, 5); HIN: a[L/2 ..]
// are zero on entry (L even, so the first radix is even)
template <typename T, int DIR, int L, int S, bool HIN = false>
__device__ __forceinline__ void cx_stages(C2<T> (&a)[L])
{
    if constexpr (RegPlan<L, S>::r > 1) {
        C2<T> b[L];
        cx_butterflies<T, DIR, L, S, 0, HIN>(a, b);
#pragma unroll
        for (int j = 0; j < L; j++) a[j] = b[j];
        cx_stages<T, DIR, L, S + 1, false>(a);
    }
}

// a[k2] *= W_L^(q k2) (DIR < 0) or its conjugate, k2 = 1 .. M-1, from the
// LDS table laid out k2-major: tw[k2 Q] = W_L^(q k2) for this lane's q (one
// LDS read with an immediate offset per factor)
template <typename T, int DIR, int M, int Q>
__device__ __forceinline__ void lane_twiddle(C2<T> (&a)[M], const C2<T> *tw)
{
#pragma unroll
    for (int k2 = 1; k2 < M; k2++) {
        const C2<T> w = tw[k2 * Q];
        a[k2] = DIR < 0 ? fmul(a[k2], w) : fmulc(a[k2], w);
    }
}

template <int D>
__device__ __forceinline__ int xlane_i(int v)
{
    if constexpr (D == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, true);  // quad_perm [1,0,3,2]
    else if constexpr (D == 2) return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, true);  // quad_perm [2,3,0,1]
    else return __builtin_amdgcn_ds_swizzle(v, (D << 10) | 0x1F);  // bit mode: lane ^ D inside 32 lanes
}
template <int D>
__device__ __forceinline__ float xlane(float v) { return __int_as_float(xlane_i<D>(__float_as_int(v))); }
template <int D>
__device__ __forceinline__ double xlane(double v)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)xlane_i<D>((int)(unsigned)(b & 0xffffffffull));
    const unsigned hi = (unsigned)xlane_i<D>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
template <typename T, int D>
__device__ __forceinline__ C2<T> xlane2(C2<T> v) { return {xlane<D>(v.x), xlane<D>(v.y)}; }

template <int H>
constexpr int log2c() { return H <= 1 ? 0 : 1 + log2c<H / 2>(); }

// p + sg * a with sg = +-1: one rounding, the same as p + a or p - a
template <typename T>
__device__ __forceinline__ C2<T> padd(C2<T> p, T sg, C2<T> a) { return {fma_t(sg, a.x, p.x), fma_t(sg, a.y, p.y)}; }

// Twiddles of the exchanges at lane distance H >= 4: wx[log2(H) - 2] =
// W_(2H)^(q mod H) in the upper lane of a pair, 1 in the lower one.
template <typename T, int L, int Q>
struct XTw {
    static constexpr int N = log2c<Q>() > 2 ? log2c<Q>() - 2 : 1;
    C2<T> w[N];
    __device__ __forceinline__ explicit XTw(int q)
    {
#pragma unroll
        for (int j = 0; j < N; j++) {
            const int H = 4 << j;
            w[j] = (H < Q && (q & H)) ? cx_table<T, L>((q & (H - 1)) * (L / (2 * H))) : C2<T>{(T)1, (T)0};
        }
    }
};

// forward exchanges, lane distance H = Q/2, Q/4, .., 1: the lower lane of a
// pair keeps the sum, the upper one the difference times W_(2H)^(q mod H)
// (1 for H = 1, 1 or -i for H = 2)
template <typename T, int M, int H, int L, int Q>
__device__ __forceinline__ void xl_fwd(C2<T> (&a)[M], int q, const XTw<T, L, Q> &xt)
{
    if constexpr (H >= 1) {
        const bool up = (q & H) != 0;
        const T sg = up ? (T)-1 : (T)1;
        if constexpr (H == 1) {
#pragma unroll
            for (int m = 0; m < M; m++) a[m] = padd(xlane2<T, H>(a[m]), sg, a[m]);
        } else if constexpr (H == 2) {
            const bool mi = up && (q & 1);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const C2<T> t = padd(xlane2<T, H>(a[m]), sg, a[m]);
                a[m] = mi ? C2<T>{t.y, -t.x} : t;
            }
        } else {
            const C2<T> w = xt.w[log2c<H>() - 2];
#pragma unroll
            for (int m = 0; m < M; m++) a[m] = fmul(padd(xlane2<T, H>(a[m]), sg, a[m]), w);
        }
        xl_fwd<T, M, H / 2, L, Q>(a, q, xt);
    }
}

// inverse exchanges, lane distance H = 1, 2, .., Q/2: the upper lane first
// takes its value times conj(W_(2H)^(q mod H)) (1 for H = 1, 1 or +i for
// H = 2), then the lower lane keeps the sum, the upper one the difference
template <typename T, int M, int H, int L, int Q>
__device__ __forceinline__ void xl_inv(C2<T> (&a)[M], int q, const XTw<T, L, Q> &xt)
{
    if constexpr (H < Q) {
        const bool up = (q & H) != 0;
        const T sg = up ? (T)-1 : (T)1;
        if constexpr (H == 1) {
#pragma unroll
            for (int m = 0; m < M; m++) a[m] = padd(xlane2<T, H>(a[m]), sg, a[m]);
        } else if constexpr (H == 2) {
            const bool pi = up && (q & 1);
#pragma unroll
            for (int m = 0; m < M; m++) {
                const C2<T> t = pi ? C2<T>{-a[m].y, a[m].x} : a[m];
                a[m] = padd(xlane2<T, H>(t), sg, t);
            }
        } else {
            const C2<T> w = xt.w[log2c<H>() - 2];
#pragma unroll
            for (int m = 0; m < M; m++) {
                const C2<T> t = fmulc(a[m], w);
                a[m] = padd(xlane2<T, H>(t), sg, t);
            }
        }
        xl_inv<T, M, 2 * H, L, Q>(a, q, xt);
    }
}

// The lanes of a line live in one wavefront, and a wavefront's LDS
// operations complete in order; the wait also keeps the compiler from
// moving LDS accesses across it.
__device__ __forceinline__ void lds_wave_fence() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// The Q-point DFTs across the lanes of a line through a wave-private LDS
// buffer xs (Q rows of Q + 1 entries for this line), Q frequencies k2 at a
// time.  Forward: for chunk c, lane q writes Y_q[cQ + j] (j < Q) to row j,
// lane p reads row p -- the Q lanes' values at k2 = cQ + p -- and takes
// their Q-point DFT, so that slot cQ + k1 holds X[cQ + p + M k1].  Inverse:
// the mirror image (inverse DFT over k1, row p written, column q read),
// ending with slot k2 = cQ + j of lane q holding its natural value again.
template <typename T, int DIR, int M, int Q>
__device__ __forceinline__ void xlds(C2<T> (&a)[M], C2<T> *xs, int q)
{
    constexpr int R = Q + 1;
#pragma unroll
    for (int c = 0; c < M / Q; c++) {
        C2<T> v[Q];
        if constexpr (DIR < 0) {
#pragma unroll
            for (int j = 0; j < Q; j++) xs[j * R + q] = a[c * Q + j];
            lds_wave_fence();
#pragma unroll
            for (int i = 0; i < Q; i++) v[i] = xs[q * R + i];
            lds_wave_fence();
            fdftR<T, -1, Q>(v);
#pragma unroll
            for (int k1 = 0; k1 < Q; k1++) a[c * Q + k1] = v[k1];
        } else {
#pragma unroll
            for (int k1 = 0; k1 < Q; k1++) v[k1] = a[c * Q + k1];
            fdftR<T, 1, Q>(v);
#pragma unroll
            for (int i = 0; i < Q; i++) xs[q * R + i] = v[i];
            lds_wave_fence();
#pragma unroll
            for (int j = 0; j < Q; j++) a[c * Q + j] = xs[j * R + q];
            lds_wave_fence();
        }
    }
}

// Frequency offset of register slot s past the lane's first frequency:
// slot s itself after the DPP exchanges (lane p holds k2 + M rev(p)), and
// (s / Q) Q + M (s % Q) after the LDS exchanges (lane p holds cQ + p + M k1).
template <int M, int Q, bool XL>
__host__ __device__ constexpr int slot_off(int s) { return XL ? (s / Q) * Q + M * (s % Q) : s; }

template <typename T>
__device__ __forceinline__ C2<T> ld_at(const char *base, uint32_t off) { return *(const C2<T> *)(base + off); }
template <typename T>
__device__ __forceinline__ void st_at(char *base, uint32_t off, C2<T> v) { *(C2<T> *)(base + off) = v; }

// Lines: b = tile * 256/Q + lane / Q along the tiled axis (< nb), l2 =
// blockIdx.y, right-hand sides blockIdx.z * rpb ..; src and dst may alias
// (every lane loads its line before any lane of it stores).  The host
// checks that the line offsets fit the 32-bit lane offsets.
//
// MAP: the x passes, whose lines are contiguous.  MODE 0 reads its elements
// from the FP64 unknown vector through the scatter map (d.map_in), MODE 1
// adds its outputs into it through the gather map (d.map_out); the unknown
// indices of a lane's elements are read once per block.  The Q lanes of a
// line then read (MODE 0) and update (MODE 1) Q consecutive unknowns per
// instruction where the map runs along x, and with the LDS exchanges (XL)
// also write (MODE 0) and read (MODE 1) Q consecutive frequencies.
template <typename T, int MODE, int L, int Q, bool XL, bool MAP>
__global__ void __launch_bounds__(256) k_lines(PassDesc d, const C2<T> *src, C2<T> *dst)
{
    constexpr int M = L / Q, LPB = 256 / Q, LQ = log2c<Q>();
    // forward inputs and inverse outputs live in the lower half of the line
    // and element L/2 (nin, nout <= L/2 + 1, checked on the host): with M
    // even the lower half is m < M/2 in every lane, so the zero half is
    // never loaded or transformed and the unused half of the inverse is
    // never formed
    constexpr bool HALF = M % 2 == 0;
    constexpr bool CONV = MODE >= 2;
    constexpr uint32_t ES = sizeof(C2<T>);
    // lane twiddles, k2-major: s_tw[k2 Q + q] = W_L^(q k2)
    __shared__ C2<T> s_tw[L];
    for (int t = (int)threadIdx.x; t < L; t += 256) s_tw[t] = cx_table<T, L>((t % Q) * (t / Q));
    __syncthreads();
    unsigned bx, by, bz;
    xcd_block(bx, by, bz);
    const int q = (int)threadIdx.x % Q;
    const int lb = (int)threadIdx.x / Q;
    const int b0 = (int)bx * LPB;
    const bool live = b0 + lb < d.nb;
    const int l2 = (int)by;
    const int rq = (int)(__brev((unsigned)q) >> (32 - LQ));
    const XTw<T, L, Q> xt(q);
    // exchange buffer of this lane's line (LDS exchanges)
    __shared__ C2<T> s_x[XL ? 256 * (Q + 1) : 1];
    C2<T> *xs = s_x + (XL ? lb * Q * (Q + 1) : 0);
    // the lane's first frequency: its bit-reversed index times M after the
    // DPP exchanges, its own index after the LDS exchanges
    const int lane_k0 = XL ? q : M * rq;
    // 32-bit lane offsets (bytes) of element 0 of this lane's stride pattern
    const uint32_t in_nat = (uint32_t)(lb * d.in_sb + (int64_t)q * d.in_sj) * ES;
    const uint32_t in_rev = (uint32_t)(lb * d.in_sb + (int64_t)lane_k0 * d.in_sj) * ES;
    const uint32_t out_nat = (uint32_t)(lb * d.out_sb + (int64_t)q * d.out_sj) * ES;
    const uint32_t out_rev = (uint32_t)(lb * d.out_sb + (int64_t)lane_k0 * d.out_sj) * ES;
    // Lines of n = L/2 + 1 elements (the circulant embedding of length
    // 2n - 2, exact for an even kernel and for a half-cell reduction, whose
    // 2n - 2 lags fit it): element L/2 lies outside the lower half the
    // lanes transform.  Forward, it adds x[L/2] (-1)^k to every frequency k
    // (with M and Q even, the parity of k is that of the slot after the lane
    // exchanges and that of the lane after the LDS exchanges).  Inverse,
    // y[L/2] is lane 0's element m = M/2, formed from its M-point
    // coefficients.
    const bool ext_in = HALF && MODE != 1 && d.nin > L / 2;
    const bool ext_out = HALF && MODE != 0 && d.nout > L / 2;
    const uint32_t in_mid = (uint32_t)(lb * d.in_sb + (int64_t)(L / 2) * d.in_sj) * ES;
    const uint32_t out_mid = (uint32_t)(lb * d.out_sb + (int64_t)(L / 2) * d.out_sj) * ES;
    auto odd_slot = [&](int i) { return XL ? (q & 1) != 0 : (i & 1) != 0; };
    // MAP: unknown index of element n = q + Q i (and of element L/2), -1
    // where the position holds no unknown or lies past the line
    constexpr int NMAP = MAP ? (HALF ? M / 2 : M) : 1;
    int32_t mi[NMAP], mext = -1;
    if constexpr (MAP) {
        const int32_t *map = MODE == 0 ? d.map_in : d.map_out;
        const int64_t sb = MODE == 0 ? d.in_sb : d.out_sb, s2 = MODE == 0 ? d.in_s2 : d.out_s2;
        const int64_t sj = MODE == 0 ? d.in_sj : d.out_sj;
        const int nlim = MODE == 0 ? d.nin : d.nout;
        const int32_t *mp = map + (int64_t)(b0 + lb) * sb + (int64_t)l2 * s2;
#pragma unroll
        for (int i = 0; i < NMAP; i++) {
            const int n = q + Q * i;
            mi[i] = live && n < nlim ? mp[(int64_t)n * sj] : -1;
        }
        if (live && HALF && L / 2 < nlim) mext = mp[(int64_t)(L / 2) * sj];
    }
    // MODE 3 with the lane exchanges and at most 48 KiB of it per block: the
    // complex spectrum waits in LDS, lane-private and k2-major (conflict-free
    // 8-byte reads), instead of in 2M VGPRs, so that three waves fit a SIMD
    constexpr bool SPEC_LDS = MODE == 3 && !XL && M * 256 * (int)sizeof(C2<T>) <= 48 * 1024;
    __shared__ C2<T> s_spec[SPEC_LDS ? M * 256 : 1];
    T sr[CONV && !SPEC_LDS ? M : 1], si[MODE == 3 && !SPEC_LDS ? M : 1];
    if (CONV) {
        const int64_t sbase = (int64_t)(b0 + lb) * d.spec_sb + (int64_t)l2 * d.spec_s2 + (int64_t)lane_k0 * d.spec_sj;
#pragma unroll
        for (int k2 = 0; k2 < M; k2++) {
            const int64_t s = sbase + (int64_t)slot_off<M, Q, XL>(k2) * d.spec_sj;
            if (MODE == 2) {
                sr[k2] = live ? reinterpret_cast<const T *>(d.spec)[s] : (T)0;
            } else {
                const C2<T> c = live ? reinterpret_cast<const C2<T> *>(d.spec)[s] : C2<T>{(T)0, (T)0};
                if constexpr (SPEC_LDS) {
                    s_spec[k2 * 256 + (int)threadIdx.x] = {c.x, d.spec_conj ? -c.y : c.y};
                } else {
                    sr[SPEC_LDS ? 0 : k2] = c.x;
                    si[SPEC_LDS ? 0 : k2] = d.spec_conj ? -c.y : c.y;
                }
            }
        }
    }
    // elements a lane loads per right-hand side: the lower half of its
    // stride pattern (forward input) or all M (inverse-only input)
    constexpr int NL = MODE == 1 ? M : (HALF ? M / 2 : M);
    // the next right-hand side's elements are loaded before this one is
    // transformed (memory-level parallelism through the whole loop) when
    // they take at most 48 more VGPRs
    constexpr bool PREF = NL * (int)sizeof(C2<T>) <= 192;
    // Element addresses: a wave-uniform pointer stepped by the element
    // stride (laundered per call, so that the per-element pointers are not
    // hoisted out of the loop into registers) plus the lane's 32-bit offset.
    auto load = [&](C2<T>(&v)[NL], C2<T> &ve, int rhs) {
        if constexpr (MAP && MODE == 0) {
            const double2_t *vb = d.vec_in + (int64_t)rhs * d.vec_in_ld;
#pragma unroll
            for (int i = 0; i < NL; i++) {
                const double2_t x = mi[i] >= 0 ? vb[mi[i]] : double2_t{0.0, 0.0};
                v[i] = {(T)x.x, (T)x.y};
            }
            const double2_t x = ext_in && mext >= 0 ? vb[mext] : double2_t{0.0, 0.0};
            ve = {(T)x.x, (T)x.y};
            return;
        }
        const char *in = (const char *)(src + (int64_t)b0 * d.in_sb + (int64_t)l2 * d.in_s2 + (int64_t)rhs * d.in_rhs);
        ve = ext_in && live ? ld_at<T>(in, in_mid) : C2<T>{(T)0, (T)0};
        int64_t step = (MODE == 1 ? 1 : Q) * d.in_sj * (int64_t)ES;
        asm volatile("" : "+s"(step));
        if constexpr (MODE == 1 && XL) {
            // slots at frequencies cQ + q + M k1: Q rows apart within a chunk
            // of slots, one chunk Q rows further
            const int64_t sm = M * step, sq = Q * step;
#pragma unroll
            for (int c = 0; c < M / Q; c++) {
                const char *pk = in + c * sq;
#pragma unroll
                for (int k1 = 0; k1 < Q; k1++, pk += sm) {
                    const int i = c * Q + k1;
                    v[i] = live && slot_off<M, Q, XL>(i) + lane_k0 < d.nin ? ld_at<T>(pk, in_rev) : C2<T>{(T)0, (T)0};
                }
            }
        } else {
#pragma unroll
            for (int i = 0; i < NL; i++, in += step) {
                const int k = MODE == 1 ? i + lane_k0 : q + Q * i;
                v[i] = live && k < d.nin ? ld_at<T>(in, MODE == 1 ? in_rev : in_nat) : C2<T>{(T)0, (T)0};
            }
        }
    };
    const int r_end = min(d.nrhs, ((int)bz + 1) * d.rpb);
    C2<T> nxt[NL], nxt_e;
    if (PREF && (int)bz * d.rpb < r_end) load(nxt, nxt_e, (int)bz * d.rpb);
    for (int rhs = (int)bz * d.rpb; rhs < r_end; rhs++) {
        char *out = (char *)(dst + (int64_t)b0 * d.out_sb + (int64_t)l2 * d.out_s2 + (int64_t)rhs * d.out_rhs);
        // the table pointer is laundered per right-hand side so that the
        // twiddles are read from LDS where they are used instead of being
        // hoisted into registers for the whole loop
        int twq = q;
        asm volatile("" : "+v"(twq));
        const C2<T> *twr = s_tw + twq;
        C2<T> a[M];
        if (!PREF) load(nxt, nxt_e, rhs);
#pragma unroll
        for (int i = 0; i < NL; i++) a[i] = nxt[i];
        const C2<T> xe = nxt_e;
        if (PREF && rhs + 1 < r_end) load(nxt, nxt_e, rhs + 1);
        if (MODE != 1) {
            cx_stages<T, -1, M, 0, HALF>(a);
            lane_twiddle<T, -1, M, Q>(a, twr);
            if constexpr (XL) xlds<T, -1, M, Q>(a, xs, q);
            else xl_fwd<T, M, Q / 2, L, Q>(a, q, xt);
            if (ext_in) {
                const C2<T> xn = {-xe.x, -xe.y};
#pragma unroll
                for (int i = 0; i < M; i++) a[i] = cadd(a[i], odd_slot(i) ? xn : xe);
            }
            if (MODE == 0) {
                if (live) {
                    int64_t step = d.out_sj * (int64_t)ES;
                    asm volatile("" : "+s"(step));
                    if constexpr (XL) {
                        const int64_t sm = M * step, sq = Q * step;
#pragma unroll
                        for (int c = 0; c < M / Q; c++) {
                            char *pk = out + c * sq;
#pragma unroll
                            for (int k1 = 0; k1 < Q; k1++, pk += sm) {
                                const int i = c * Q + k1;
                                if (slot_off<M, Q, XL>(i) + lane_k0 < d.nout) st_at<T>(pk, out_rev, a[i]);
                            }
                        }
                    } else {
#pragma unroll
                        for (int k2 = 0; k2 < M; k2++, out += step)
                            if (k2 + lane_k0 < d.nout) st_at<T>(out, out_rev, a[k2]);
                    }
                }
                continue;
            }
            if constexpr (SPEC_LDS) {
                // laundered per right-hand side: read where used, not hoisted
                int so = (int)threadIdx.x;
                asm volatile("" : "+v"(so));
#pragma unroll
                for (int k2 = 0; k2 < M; k2++) a[k2] = fmul(a[k2], s_spec[k2 * 256 + so]);
            } else {
#pragma unroll
                for (int k2 = 0; k2 < M; k2++) {
                    if (MODE == 2) a[k2] = cscale(a[k2], sr[k2]);
                    else a[k2] = fmul(a[k2], C2<T>{sr[k2], si[MODE == 3 ? k2 : 0]});
                }
            }
        }
        if constexpr (XL) xlds<T, 1, M, Q>(a, xs, q);
        else xl_inv<T, M, 1, L, Q>(a, q, xt);
        // element L/2 is m = M/2 of lane 0, whose lane twiddles are 1: the
        // (-1)^k2-weighted sum of its M-point coefficients
        C2<T> ye = {(T)0, (T)0};
        if (ext_out) {
#pragma unroll
            for (int k2 = 0; k2 < M; k2++) ye = (k2 & 1) ? csub(ye, a[k2]) : cadd(ye, a[k2]);
        }
        // laundered again: the forward twiddles are not kept live until here
        asm volatile("" : "+v"(twq));
        lane_twiddle<T, 1, M, Q>(a, s_tw + twq);
        cx_stages<T, 1, M, 0>(a);
        if constexpr (MAP && MODE == 1) {
            double2_t *vb = d.vec_out + (int64_t)rhs * d.vec_out_ld;
            auto acc = [&](int32_t u, C2<T> v) {
                double2_t t = vb[u];
                t.x += (double)v.x;
                t.y += (double)v.y;
                vb[u] = t;
            };
#pragma unroll
            for (int m = 0; m < NMAP; m++)
                if (mi[m] >= 0) acc(mi[m], a[m]);
            if (ext_out && q == 0 && mext >= 0) acc(mext, ye);
            continue;
        }
        if (live) {
            if (ext_out && q == 0) st_at<T>(out, out_mid, ye);
            int64_t step = Q * d.out_sj * (int64_t)ES;
            asm volatile("" : "+s"(step));
#pragma unroll
            for (int m = 0; m < (HALF ? M / 2 : M); m++, out += step)
                if (q + Q * m < d.nout) st_at<T>(out, out_nat, a[m]);
        }
    }
}

}  // namespace

// (L, Q) pairs k_lines is built for: z convolutions (MODE 2, 3) and y
// passes (MODE 0, 1) of the workloads' lengths, at most 64 points per lane;
// with the LDS exchanges (XL) where Q <= 8 divides M = L / Q; the x passes
// with their maps (complex64 only).
#define PEEC_LANES_Z(X) X(16, 4) X(32, 2) X(32, 4) X(32, 8) X(48, 2) X(48, 4) X(48, 8) X(64, 2) X(64, 4) X(64, 8) \
    X(64, 16) X(96, 4) X(96, 8) X(96, 16) X(128, 4) X(128, 8) X(128, 16)
#define PEEC_LANES_Y(X) X(128, 4) X(128, 8) X(128, 16) X(192, 4) X(192, 8) X(192, 16) X(200, 4) X(200, 8) X(384, 8) \
    X(384, 16) X(384, 32) X(400, 8) X(400, 16)
#define PEEC_LANES_ZX(X) X(32, 4) X(48, 4) X(64, 4) X(64, 8) X(96, 4) X(128, 4) X(128, 8)
#define PEEC_LANES_YX(X) X(128, 4) X(128, 8) X(192, 4) X(192, 8) X(384, 8)
#define PEEC_LANES_XM(X) X(128, 8) X(128, 16) X(192, 8) X(192, 16) X(384, 16) X(384, 32)
#define PEEC_LANES_XMX(X) X(128, 8) X(192, 8) X(384, 8)

bool lane_built(int L, int Q, int mode, bool xl, bool map)
{
    bool ok = false;
#define PEEC_LANE_HAS(L_, Q_) ok = ok || (L == L_ && Q == Q_);
    if (map) {
        if (mode >= 2) return false;
        if (xl) { PEEC_LANES_XMX(PEEC_LANE_HAS) }
        else { PEEC_LANES_XM(PEEC_LANE_HAS) }
    } else if (xl) {
        if (mode >= 2) { PEEC_LANES_ZX(PEEC_LANE_HAS) }
        else { PEEC_LANES_YX(PEEC_LANE_HAS) }
    } else {
        if (mode >= 2) { PEEC_LANES_Z(PEEC_LANE_HAS) }
        else { PEEC_LANES_Y(PEEC_LANE_HAS) }
    }
#undef PEEC_LANE_HAS
    return ok;
}

// whether the pass is an x pass with its map (forward scatter, inverse gather)
static bool lane_map(const PassDesc &d, int mode)
{
    return (mode == 0 && d.map_in != nullptr) || (mode == 1 && d.map_out != nullptr);
}

// The lane kernel applies to a pass whose tile axis is contiguous and whose
// lines are strided, or to an x pass with its map (complex64); the forward
// input (modes 0, 2, 3) and the inverse output (modes 1, 2, 3) lie in the
// lower half of the line and element L/2.
bool lane_fits(const PassDesc &d, int Q, int mode, bool xl, bool fp32)
{
    const bool map = lane_map(d, mode);
    // the kernel's lane offsets are 32-bit byte offsets within a tile
    const int64_t sb = std::max(d.in_sb, d.out_sb), sj = std::max(d.in_sj, d.out_sj);
    const int64_t span = ((int64_t)(256 / std::max(Q, 1)) * sb + (int64_t)d.L * sj + 256) * 16;
    // an even share per lane (M = L / Q) carries a line of L/2 + 1 elements
    // (its last one outside the lanes' lower halves); an odd share L/2
    const int lim = d.L / 2 + ((d.L / std::max(Q, 1)) % 2 == 0 ? 1 : 0);
    const bool layout = map ? fp32 && (mode == 0 ? d.vec_in != nullptr : d.vec_out != nullptr)
                            : d.in_sb == 1 && d.out_sb == 1 && d.in_sj > 1 && d.out_sj > 1 && !d.map_in && !d.map_out;
    return Q >= 2 && d.L % Q == 0 && lane_built(d.L, Q, mode, xl, map) && layout && d.nin <= d.L &&
           d.nout <= d.L && span < ((int64_t)1 << 31) && (mode == 1 || d.nin <= lim) && (mode == 0 || d.nout <= lim);
}

template <typename T, int MODE, bool XL>
static int launch_lines_t(peec_ctx *ctx, PassDesc &d, const void *src, void *dst, int Q)
{
    const int rpb = std::max(1, d.rpb), lpb = 256 / Q;
    dim3 grid((unsigned)((d.nb + lpb - 1) / lpb), (unsigned)d.n2, (unsigned)((d.nrhs + rpb - 1) / rpb));
    const C2<T> *s = (const C2<T> *)src;
    C2<T> *o = (C2<T> *)dst;
    const bool map = lane_map(d, MODE);
    bool done = false;
#define PEEC_LANE_LAUNCH(L_, Q_)                                                                                   \
    if (!done && d.L == L_ && Q == Q_) {                                                                           \
        hipLaunchKernelGGL((k_lines<T, MODE, L_, Q_, XL, false>), grid, dim3(256), 0, ctx->stream, d, s, o);   \
        done = true;                                                                                               \
    }
#define PEEC_LANE_LAUNCH_MAP(L_, Q_)                                                                               \
    if (!done && d.L == L_ && Q == Q_) {                                                                           \
        hipLaunchKernelGGL((k_lines<T, MODE, L_, Q_, XL, true>), grid, dim3(256), 0, ctx->stream, d, s, o);    \
        done = true;                                                                                               \
    }
    if (map) {
        if constexpr (MODE < 2 && std::is_same<T, float>::value) {
            if constexpr (XL) { PEEC_LANES_XMX(PEEC_LANE_LAUNCH_MAP) }
            else { PEEC_LANES_XM(PEEC_LANE_LAUNCH_MAP) }
        }
    } else if constexpr (XL) {
        if constexpr (MODE >= 2) { PEEC_LANES_ZX(PEEC_LANE_LAUNCH) }
        else { PEEC_LANES_YX(PEEC_LANE_LAUNCH) }
    } else {
        if constexpr (MODE >= 2) { PEEC_LANES_Z(PEEC_LANE_LAUNCH) }
        else { PEEC_LANES_Y(PEEC_LANE_LAUNCH) }
    }
#undef PEEC_LANE_LAUNCH
#undef PEEC_LANE_LAUNCH_MAP
    if (!done) {
        peec_set_error("lane-split FFT pass: no kernel for length %d with %d lanes per line%s%s", d.L, Q,
                       XL ? " (LDS exchanges)" : "", map ? " (x pass with map)" : "");
        return PEEC_ERR_ARG;
    }
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

template <bool XL>
static int launch_lane_pass_t(peec_ctx *ctx, bool fp32, int mode, PassDesc &d, const void *src, void *dst, int Q)
{
    if (fp32) {
        if (mode == 0) return launch_lines_t<float, 0, XL>(ctx, d, src, dst, Q);
        if (mode == 1) return launch_lines_t<float, 1, XL>(ctx, d, src, dst, Q);
        if (mode == 2) return launch_lines_t<float, 2, XL>(ctx, d, src, dst, Q);
        return launch_lines_t<float, 3, XL>(ctx, d, src, dst, Q);
    }
    if (mode == 0) return launch_lines_t<double, 0, XL>(ctx, d, src, dst, Q);
    if (mode == 1) return launch_lines_t<double, 1, XL>(ctx, d, src, dst, Q);
    if (mode == 2) return launch_lines_t<double, 2, XL>(ctx, d, src, dst, Q);
    return launch_lines_t<double, 3, XL>(ctx, d, src, dst, Q);
}

int launch_lane_pass(peec_ctx *ctx, bool fp32, int mode, PassDesc &d, const void *src, void *dst, int Q, bool xl)
{
    return xl ? launch_lane_pass_t<true>(ctx, fp32, mode, d, src, dst, Q)
              : launch_lane_pass_t<false>(ctx, fp32, mode, d, src, dst, Q);
}
