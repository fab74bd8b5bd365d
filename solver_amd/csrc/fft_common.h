// Device helpers shared by the FFT translation units (fft.hip: the LDS
// and register passes; fft_lanes.hip: the lane-split passes): complex
// arithmetic, radix 2/3/4/5/8 butterflies, the pass descriptor, the XCD
// tile order and the register Stockham stages.
#pragma once

#include "context.h"

namespace fftk {

template <typename T>
struct C2 {
    T x, y;
};

template <typename T>
__device__ __forceinline__ C2<T> cadd(C2<T> a, C2<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> csub(C2<T> a, C2<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ C2<T> cmul(C2<T> a, C2<T> b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
template <typename T>
__device__ __forceinline__ C2<T> cscale(C2<T> a, T s) { return {a.x * s, a.y * s}; }
// multiply by -i (forward) or +i (inverse)
template <typename T, int DIR>
__device__ __forceinline__ C2<T> rot(C2<T> a) { return DIR < 0 ? C2<T>{a.y, -a.x} : C2<T>{-a.y, a.x}; }

template <typename T, int DIR>
__device__ __forceinline__ void dft2(C2<T> *v)
{
    C2<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
}

template <typename T, int DIR>
__device__ __forceinline__ void dft4(C2<T> *v)
{
    C2<T> s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    C2<T> s13 = cadd(v[1], v[3]), d13 = rot<T, DIR>(csub(v[1], v[3]));
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
}

template <typename T, int DIR>
__device__ __forceinline__ void dft8(C2<T> *v)
{
    C2<T> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    dft4<T, DIR>(e);
    dft4<T, DIR>(o);
    const T h = (T)0.70710678118654752440084436210485;
    // o[k] *= exp(DIR * 2 pi i k / 8)
    C2<T> w1 = DIR < 0 ? C2<T>{h, -h} : C2<T>{h, h};
    C2<T> w3 = DIR < 0 ? C2<T>{-h, -h} : C2<T>{-h, h};
    o[1] = cmul(o[1], w1);
    o[2] = rot<T, DIR>(o[2]);
    o[3] = cmul(o[3], w3);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        v[k] = cadd(e[k], o[k]);
        v[k + 4] = csub(e[k], o[k]);
    }
}

template <typename T, int DIR>
__device__ __forceinline__ void dft3(C2<T> *v)
{
    const T s = (T)0.86602540378443864676372317075294;
    C2<T> b = cadd(v[1], v[2]);
    C2<T> d = rot<T, DIR>(cscale(csub(v[1], v[2]), s));
    C2<T> t = csub(v[0], cscale(b, (T)0.5));
    v[0] = cadd(v[0], b);
    v[1] = cadd(t, d);
    v[2] = csub(t, d);
}

template <typename T, int DIR>
__device__ __forceinline__ void dft5(C2<T> *v)
{
    const T c1 = (T)0.30901699437494742410229341718282, c2 = (T)-0.80901699437494742410229341718282;
    const T s1 = (T)0.95105651629515357211643933337938, s2 = (T)0.58778525229247312916870595463907;
    C2<T> b1 = cadd(v[1], v[4]), b2 = cadd(v[2], v[3]);
    C2<T> d1 = csub(v[1], v[4]), d2 = csub(v[2], v[3]);
    C2<T> a0 = v[0];
    C2<T> r1 = cadd(a0, cadd(cscale(b1, c1), cscale(b2, c2)));
    C2<T> r2 = cadd(a0, cadd(cscale(b1, c2), cscale(b2, c1)));
    C2<T> i1 = rot<T, DIR>(cadd(cscale(d1, s1), cscale(d2, s2)));
    C2<T> i2 = rot<T, DIR>(csub(cscale(d1, s2), cscale(d2, s1)));
    v[0] = cadd(a0, cadd(b1, b2));
    v[1] = cadd(r1, i1);
    v[4] = csub(r1, i1);
    v[2] = cadd(r2, i2);
    v[3] = csub(r2, i2);
}

template <typename T, int DIR, int R>
__device__ __forceinline__ void dftR(C2<T> *v)
{
    if (R == 2) dft2<T, DIR>(v);
    else if (R == 3) dft3<T, DIR>(v);
    else if (R == 4) dft4<T, DIR>(v);
    else if (R == 5) dft5<T, DIR>(v);
    else dft8<T, DIR>(v);
}

struct PassDesc {
    int L, nin, nout;
    int nstage;
    int radix[16];
    int64_t tw_off;
    int nb, n2;                          // lines: nb consecutive positions (W per block) x n2
    int64_t in_sb, in_s2, in_sj;
    int64_t out_sb, out_s2, out_sj;
    int64_t in_rhs, out_rhs;             // stride between right-hand sides
    int nrhs;
    int rpb;                             // right-hand sides per block (blockIdx.z covers rpb of them)
    const int32_t *map_in;               // optional 0/1 scatter (compact position -> unknown)
    const double2_t *vec_in;
    int64_t vec_in_ld;                   // stride between right-hand sides in vec_in
    const int32_t *map_out;              // optional 0/1 gather
    double2_t *vec_out;
    int64_t vec_out_ld;
    const void *spec;                    // spectrum for the fused z pass: real T (MODE 2), C2<T> (MODE 3)
    int64_t spec_sb, spec_s2, spec_sj;
    int spec_conj;                       // MODE 3: multiply by the conjugate spectrum
    int64_t tw_off_z;                    // fused y-z pass: z twiddles (z length is a template argument)
};


// Tile (bx, by, bz) of this workgroup in XCD order (xcd_tile over the x
// fastest, then y, then z tile order).
__device__ __forceinline__ void xcd_block(unsigned &bx, unsigned &by, unsigned &bz)
{
    const unsigned gx = gridDim.x, gy = gridDim.y;
    const unsigned t = xcd_tile(blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z), gx * gy * gridDim.z);
    bx = t % gx;
    by = (t / gx) % gy;
    bz = t / (gx * gy);
}

// The radix schedule of a register FFT of compile-time length L: radix 8
// first, then 4, 2, 3, 5 (the stages and twiddles of the LDS passes).
template <int L, int S>
struct RegPlan {
    // radix of stage S (8 first, then 4, 2, 3, 5), and the stages' product
    static constexpr int rem(int l, int s)
    {
        return s == 0 ? l : rem(l / (l % 8 == 0 ? 8 : l % 4 == 0 ? 4 : l % 2 == 0 ? 2 : l % 3 == 0 ? 3 : 5), s - 1);
    }
    static constexpr int r = rem(L, S) == 1 ? 1
                             : rem(L, S) % 8 == 0 ? 8
                             : rem(L, S) % 4 == 0 ? 4
                             : rem(L, S) % 2 == 0 ? 2
                             : rem(L, S) % 3 == 0 ? 3 : 5;
    static constexpr int ns = L / rem(L, S);  // product of the radices of stages 0 .. S-1
};

template <typename T, int DIR, int L, int S>
__device__ __forceinline__ void reg_stages(C2<T> (&a)[L], const double2_t *__restrict__ tw)
{
    constexpr int R = RegPlan<L, S>::r;
    if constexpr (R > 1) {
        constexpr int Ns = RegPlan<L, S>::ns;
        constexpr int nb = L / R;
        constexpr int stride = L / (Ns * R);
        C2<T> b[L];
#pragma unroll
        for (int k = 0; k < nb; k++) {
            const int kk = k % Ns;
            C2<T> v[R];
#pragma unroll
            for (int r = 0; r < R; r++) v[r] = a[k + r * nb];
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    const double2_t w = tw[kk * r * stride];
                    v[r] = cmul(v[r], C2<T>{(T)w.x, (T)(DIR < 0 ? w.y : -w.y)});
                }
            }
            dftR<T, DIR, R>(v);
#pragma unroll
            for (int r = 0; r < R; r++) b[(k - kk) * R + kk + r * Ns] = v[r];
        }
#pragma unroll
        for (int j = 0; j < L; j++) a[j] = b[j];
        reg_stages<T, DIR, L, S + 1>(a, tw);
    }
}

}  // namespace fftk
