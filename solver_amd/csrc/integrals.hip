// FP64 VALU kernels for the partial-element rows of the voxel PEEC hot path.
//
// compute_Lp_Voxels.jl:1-63 (Lp rows), compute_row_P_sup.jl:1-104 (P rows),
// compute_diagonals.jl:136-155 (Lp_self).  One thread evaluates one row
// entry; the branch ("dimension suppression") classifier is table driven:
// every branch of Integ_vol_vol_fft / Integ_sup_sup_P reduces to
//     prefactor * F(object1, object2, axis permutation)
// for one of twelve closed forms F, so the device code carries each closed
// form once and the classifier only picks (F, object order, permutation).
//
// Arithmetic matches the reference's Julia semantics: x^2 = x*x, x^3 =
// x*x*x, x^4 = compensated square-of-square (Base.Math.pow_body), the real
// part of a complex log of a real number = 0.5*log(a*a), ties-to-even
// coordinate rounding to 1e-14 m, NaN/Inf terms zeroed.  log, log1p and
// atan2 are the deterministic fdlibm-algorithm versions of detmath.h, so
// the cancelling corner sums come out identical on the CPU and the GPU.
// This file is built
// with -ffp-contract=off so no multiply-add is fused behind the formulas'
// backs.
#include "peec_internal.h"
#include "detmath.h"

#include <cmath>

#define HD __host__ __device__

namespace {

// ------------------------------------------------------------- scalars
HD inline bool bad(double t) { return __builtin_isnan(t) || __builtin_isinf(t); }

HD inline double pow4(double x)
{
    double x2 = x * x;
    double e1 = fma(x, x, -x2);
    double err = x2 * 2.0 * e1;
    double x4 = x2 * x2;
    double e2 = fma(x2, x2, -x4);
    double lo = e2 + err;
    return (__builtin_isfinite(x4) && __builtin_isfinite(lo)) ? fma(x4, 1.0, lo) : x4;
}

// real(log(complex(a)))
HD inline double clog_re(double a)
{
    double b = fabs(a);
    if (b == 0.0) return -INFINITY;
    if (__builtin_isinf(b)) return INFINITY;
    if (b < 1e-140 || b > 1e140) return det_log(b);
    double rho = a * a;
    if (0.5 < b * b && (b <= 1.25 || rho < 3.0)) return det_log1p((b - 1.0) * (b + 1.0)) / 2.0;
    return det_log(rho) / 2.0;
}

HD inline double round14(double v) { return rint(v * 1e14) / 1e14; }
HD inline double cube_abs(double v) { double a = fabs(v); return a * a * a; }
HD inline double sgn(int k) { return (k & 1) ? -1.0 : 1.0; }

// An axis-aligned set: interval [lo, hi] per axis, its centre, and the
// length used in prefactors (P faces carry the reference's fake unit
// thickness there, compute_row_P_sup.jl:128-146).
struct Box {
    double lo[3], hi[3], c[3], len[3];
};

// ------------------------------------------------------------- closed forms
// Each form is written against "form axes" (x, y, z); the caller passes
// coordinates already permuted.  I = interval (2 values), C = centre.
// Corner loops run 0-based, so every sign exponent carries the parity of
// the loop count relative to the reference's 1-based (-1)^(c1+...) form.

// compute_Lp_Voxels.jl:766-805   (I I I | I I I)
HD double f_vol_vol(const double *x1v, const double *y1v, const double *z1v,
                            const double *x2v, const double *y2v, const double *z2v)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 64; c++) {
        // loop nest order c1..c6 = x1,y1,z1,x2,y2,z2 with c6 fastest
        int c1 = (c >> 5) & 1, c2 = (c >> 4) & 1, c3 = (c >> 3) & 1;
        int c4 = (c >> 2) & 1, c5 = (c >> 1) & 1, c6 = c & 1;
        double dx = x1v[c1] - x2v[c4], dy = y1v[c2] - y2v[c5], dz = z1v[c3] - z2v[c6];
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
        double dx4 = pow4(dx), dy4 = pow4(dy), dz4 = pow4(dz);
        double t1 = (dx4 + dy4 + dz4 - 3 * dx2 * dy2 - 3 * dy2 * dz2 - 3 * dx2 * dz2) * R / 60;
        double t2 = (dy2 * dz2 / 4 - dy4 / 24 - dz4 / 24) * dx * clog_re(dx + R);
        if (bad(t2)) t2 = 0.0;
        double t3 = (dy2 * dx2 / 4 - dy4 / 24 - dx4 / 24) * dz * clog_re(dz + R);
        if (bad(t3)) t3 = 0.0;
        double t4 = (dz2 * dx2 / 4 - dz4 / 24 - dx4 / 24) * dy * clog_re(dy + R);
        if (bad(t4)) t4 = 0.0;
        double t5 = -cube_abs(dx) * dy * dz / 6 * det_atan2(dy * dz, (fabs(dx) * R));
        double t6 = -dx * cube_abs(dy) * dz / 6 * det_atan2(dx * dz, (fabs(dy) * R));
        double t7 = -dx * dy * cube_abs(dz) / 6 * det_atan2(dx * dy, (fabs(dz) * R));
        sol = sol + sgn(c1 + c2 + c3 + c4 + c5 + c6 + 1) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
    }
    return sol;
}

// :728-764   (I I I | I I C)
HD double f_vol_surf(const double *x1v, const double *y1v, const double *z1v,
                             const double *x2v, const double *y2v, double z2)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 32; c++) {
        int c1 = (c >> 4) & 1, c2 = (c >> 3) & 1, c3 = (c >> 2) & 1, c4 = (c >> 1) & 1, c5 = c & 1;
        double dx = x1v[c1] - x2v[c4], dy = y1v[c2] - y2v[c5], dz = z1v[c3] - z2;
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = ((dz * dz) / 12 - ((dx * dx) + (dy * dy)) / 8) * dz * R;
        double t2 = ((dy * dy) / 2 - (dz * dz) / 6) * dx * clog_re(dx + R);
        if (bad(t2)) t2 = 0.0;
        double t3 = ((dx * dx) / 2 - (dz * dz) / 6) * dy * clog_re(dy + R);
        if (bad(t3)) t3 = 0.0;
        double t4 = (-pow4(dx) / 24 - pow4(dy) / 24 + ((dy * dy) * (dx * dx)) / 4) * clog_re(dz + R);
        if (bad(t4)) t4 = 0.0;
        double t5 = -cube_abs(dx) * dy / 6 * det_atan2(dy * dz, (fabs(dx) * R));
        double t6 = -dx * cube_abs(dy) / 6 * det_atan2(dx * dz, (fabs(dy) * R));
        double t7 = -dx * dy * fabs(dz) * dz / 2 * det_atan2(dx * dy, (fabs(dz) * R));
        sol = sol + sgn(c1 + c2 + c3 + c4 + c5 + 1) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
    }
    return sol;
}

// :670-697 (complex log) and compute_Lp_Voxels_QS.jl:695-730 (real log)
// (I I C | I I C), loop order x1, y1, y2, x2
template <bool REAL_LOG>
HD double f_surf_surf_para(const double *x1v, const double *y1v, double z1,
                                   const double *x2v, const double *y2v, double z2)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 16; c++) {
        int c1 = (c >> 3) & 1, c2 = (c >> 2) & 1, c3 = (c >> 1) & 1, c4 = c & 1;
        double dx = x1v[c1] - x2v[c4], dy = y1v[c2] - y2v[c3], dz = z1 - z2;
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = -1.0 / 6 * (dx * dx + dy * dy - 2 * (dz * dz)) * R;
        double t2 = -dx * dy * fabs(dz) * det_atan2(dx * dy, (fabs(dz) * R));
        double t3 = 1.0 / 2 * (dx * dx - dz * dz) * dy * (REAL_LOG ? det_log(dy + R) : clog_re(dy + R));
        if (bad(t3)) t3 = 0.0;
        double t4 = 1.0 / 2 * (dy * dy - dz * dz) * dx * (REAL_LOG ? det_log(dx + R) : clog_re(dx + R));
        if (bad(t4)) t4 = 0.0;
        sol = sol + sgn(c1 + c2 + c3 + c4) * (t1 + t2 + t3 + t4);
    }
    return sol;
}

// Lp copy :699-726 (four terms) and P copy compute_row_P_sup.jl:557-590
// (seven terms).  (I I C | I C I), loop order x1, y1, z2, x2
template <bool P_COPY>
HD double f_surf_surf_ortho(const double *x1v, const double *y1v, double z1,
                                    const double *x2v, double y2, const double *z2v)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 16; c++) {
        int c1 = (c >> 3) & 1, c2 = (c >> 2) & 1, c3 = (c >> 1) & 1, c4 = c & 1;
        double dx = x1v[c1] - x2v[c4], dy = y1v[c2] - y2, dz = z1 - z2v[c3];
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = -1.0 / 3 * dy * dz * R;
        double t2 = -1.0 / 2 * dx * fabs(dz) * dz * det_atan2(dx * dy, (fabs(dz) * R));
        double acc;
        if (P_COPY) {
            double t3 = -1.0 / 2 * dx * fabs(dy) * dy * det_atan2(dx * dz, fabs(dy) * R);
            double t4 = -1.0 / 6 * cube_abs(dx) * det_atan2(dy * dz, fabs(dx) * R);
            double t5 = dx * dy * dz * clog_re(dx + R);
            if (bad(t5)) t5 = 0.0;
            double t6 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dz * dz)) * dz * clog_re(dy + R);
            if (bad(t6)) t6 = 0.0;
            double t7 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dy * dy)) * dy * clog_re(dz + R);
            if (bad(t7)) t7 = 0.0;
            acc = t1 + t2 + t3 + t4 + t5 + t6 + t7;
        } else {
            double t3 = 1.0 / 2 * (dx * dx - dz * dz) * dy * clog_re(dy + R);
            if (bad(t3)) t3 = 0.0;
            double t4 = 1.0 / 2 * (dx * dx - dy * dy) * dx * clog_re(dx + R);
            if (bad(t4)) t4 = 0.0;
            acc = t1 + t2 + t3 + t4;
        }
        sol = sol + sgn(c1 + c2 + c3 + c4) * acc;
    }
    return sol;
}

// :635-668   (I I I | I C C), loop x1, y1, z1, x2
HD double f_line_vol(const double *x1v, const double *y1v, const double *z1v,
                             const double *x2v, double y2, double z2)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 16; c++) {
        int c1 = (c >> 3) & 1, c2 = (c >> 2) & 1, c3 = (c >> 1) & 1, c4 = c & 1;
        double dx = x1v[c1] - x2v[c4], dy = y1v[c2] - y2, dz = z1v[c3] - z2;
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = -1.0 / 3 * dy * dz * R;
        double t2 = -1.0 / 2 * dx * fabs(dz) * dz * det_atan2(dx * dy, (fabs(dz) * R));
        double t3 = -1.0 / 2 * dx * fabs(dy) * dy * det_atan2(dx * dz, (fabs(dy) * R));
        double t4 = -1.0 / 6 * cube_abs(dx) * det_atan2(dy * dz, (fabs(dx) * R));
        double t5 = dx * dy * dz * clog_re(dx + R);
        if (bad(t5)) t5 = 0.0;
        double t6 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dz * dz)) * dz * clog_re(dy + R);
        if (bad(t6)) t6 = 0.0;
        double t7 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dy * dy)) * dy * clog_re(dz + R);
        if (bad(t7)) t7 = 0.0;
        sol = sol + sgn(c1 + c2 + c3 + c4 + 1) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
    }
    return sol;
}

// :609-633 / P :531-555   (I I C | I C C), loop x1, y1, x2
HD double f_line_surf_para(const double *x1v, const double *y1v, double z1,
                                   const double *x2v, double y2, double z2)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 8; c++) {
        int c1 = (c >> 2) & 1, c2 = (c >> 1) & 1, c3 = c & 1;
        double dx = x1v[c1] - x2v[c3], dy = y1v[c2] - y2, dz = z1 - z2;
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = dx * dy * clog_re(dx + R);
        if (bad(t1)) t1 = 0.0;
        double t2 = (dx * dx - dz * dz) / 2 * clog_re(dy + R);
        if (bad(t2)) t2 = 0.0;
        double t3 = -dx * fabs(dz) * det_atan2(dx * dy, (fabs(dz) * R));
        double t4 = -dy / 2 * R;
        sol = sol + sgn(c1 + c2 + c3) * (t1 + t2 + t3 + t4);
    }
    return sol;
}

// Lp copy :584-607 (three terms) and P copy :500-529 (six terms).
// (I C C | C I I), loop x1, y2, z2
template <bool P_COPY>
HD double f_line_surf_ortho(const double *x1v, double y1, double z1,
                                    double x2, const double *y2v, const double *z2v)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 8; c++) {
        int c1 = (c >> 2) & 1, c2 = (c >> 1) & 1, c3 = c & 1;
        double dx = x1v[c1] - x2, dy = y1 - y2v[c2], dz = z1 - z2v[c3];
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = dy * dz * clog_re(dx + R);
        if (bad(t1)) t1 = 0.0;
        double t2 = dx * dz * clog_re(dy + R);
        if (bad(t2)) t2 = 0.0;
        double acc;
        if (P_COPY) {
            double t3 = dy * dx * clog_re(dz + R);
            if (bad(t3)) t3 = 0.0;
            double t4 = -1.0 / 2 * fabs(dz) * dz * det_atan2(dx * dy, (fabs(dz) * R));
            double t5 = -1.0 / 2 * fabs(dy) * dy * det_atan2(dx * dz, (fabs(dy) * R));
            double t6 = -1.0 / 2 * fabs(dx) * dx * det_atan2(dy * dz, (fabs(dx) * R));
            acc = t1 + t2 + t3 + t4 + t5 + t6;
        } else {
            double t3 = -fabs(dz) * det_atan2(dx * dy, (fabs(dz) * R));
            acc = t1 + t2 + t3;
        }
        sol = sol + sgn(c1 + c2 + c3 + 1) * acc;
    }
    return sol;
}

// :553-582   (C C C | I I I), loop x2, y2, z2
HD double f_point_vol(double x1, double y1, double z1,
                              const double *x2v, const double *y2v, const double *z2v)
{
    double sol = 0.0;
#pragma unroll 1
    for (int c = 0; c < 8; c++) {
        int c1 = (c >> 2) & 1, c2 = (c >> 1) & 1, c3 = c & 1;
        double dx = x1 - x2v[c1], dy = y1 - y2v[c2], dz = z1 - z2v[c3];
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = dy * dz * clog_re(dx + R);
        if (bad(t1)) t1 = 0.0;
        double t2 = dx * dz * clog_re(dy + R);
        if (bad(t2)) t2 = 0.0;
        double t3 = dy * dx * clog_re(dz + R);
        if (bad(t3)) t3 = 0.0;
        double t4 = -1.0 / 2 * fabs(dz) * dz * det_atan2(dx * dy, (fabs(dz) * R));
        double t5 = -1.0 / 2 * fabs(dy) * dy * det_atan2(dx * dz, (fabs(dy) * R));
        double t6 = -1.0 / 2 * fabs(dx) * dx * det_atan2(dy * dz, (fabs(dx) * R));
        sol = sol + sgn(c1 + c2 + c3) * (t1 + t2 + t3 + t4 + t5 + t6);
    }
    return sol;
}

// :531-551 (I C C | C I C) with sign offset so, loop x1, y2.
// The same arithmetic is point_sup (:469-489) with so = 0 and x1v/y2v
// swapped roles, so one body serves both.
HD double f_two_axis(const double *xv, double xfix, const double *yv, double yfix, double dz,
                             bool x_from_1, bool y_from_1, int so)
{
    double sol = 0.0;
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int c1 = (c >> 1) & 1, c2 = c & 1;
        double dx = x_from_1 ? (xv[c1] - xfix) : (xfix - xv[c1]);
        double dy = y_from_1 ? (yv[c2] - yfix) : (yfix - yv[c2]);
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = dx * clog_re(dy + R);
        if (bad(t1)) t1 = 0.0;
        double t2 = dy * clog_re(dx + R);
        if (bad(t2)) t2 = 0.0;
        double t3 = -fabs(dz) * det_atan2(dx * dy, (fabs(dz) * R));
        sol = sol + sgn(c1 + c2 + so) * (t1 + t2 + t3);
    }
    return sol;
}

// :491-529 / P :438-476 / compute_Lp_Voxels_QS.jl:474-519  (I C C | I C C)
// COPY 0: the Lp copy, 1: the P copy (x1*log|x1-x2| added outside the
// (x1-x2)/R factor), 2: the QS copy (added inside it).
template <int COPY>
HD double f_line_line_parall(const double *x1v, double y1, double z1,
                                     const double *x2v, double y2, double z2)
{
    double dy = y1 - y2, dz = z1 - z2;
    double sol = 0.0;
    if (fabs(dy) < 1e-10 && fabs(dz) < 1e-10) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
            int c1 = (c >> 1) & 1, c2 = c & 1;
            double x1 = x1v[c1], x2 = x2v[c2];
            double R = sqrt((x1 - x2) * (x1 - x2));
            double t1 = COPY == 2 ? (x1 - x2) / R * (x1 - x2 * clog_re(x2 - x1) + x1 * clog_re(x1 - x2))
                                  : (x1 - x2) / R * (x1 - x2 * clog_re(x2 - x1));
            if (COPY == 1) t1 = t1 + x1 * clog_re(x1 - x2);
            if (bad(t1)) t1 = 0.0;
            sol = sol + sgn(c1 + c2 + 1) * t1;
        }
        return sol;
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
        int c1 = (c >> 1) & 1, c2 = c & 1;
        double dx = x1v[c1] - x2v[c2];
        double R = sqrt(dx * dx + dy * dy + dz * dz);
        double t1 = dx * clog_re(dx + R);
        if (bad(t1)) t1 = 0.0;
        sol = sol + sgn(c1 + c2 + 1) * (t1 - R);
    }
    return sol;
}

// :449-467  (I C C | C C C)
HD double f_line_point(const double *x1v, double y3, double z3, double x1, double y1, double z1)
{
    double x3 = x1v[0], x4 = x1v[1];
    double check = 1.0 / (x1 - x3);
    if (bad(check)) x1 = x1 - 1e-8;
    check = 1.0 / (x1 - x4);
    if (__builtin_isnan(check) || (__builtin_isinf(check) && x1 > 0)) x1 = x1 + 1e-8;
    double R1 = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double R2 = sqrt((x1 - x4) * (x1 - x4) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double Ip = det_log((x1 - x3) + R1) - det_log((x1 - x4) + R2);
    if (bad(Ip))
        Ip = (x1 - x3) / fabs(x1 - x3) * det_log(fabs(x1 - x3)) - (x1 - x4) / fabs(x1 - x4) * det_log(fabs(x1 - x4));
    return Ip;
}

// ------------------------------------------------------------- dispatch
// Form ids used by the classifier tables.
enum Form : int {
    VOL_VOL, VOL_SURF, SSP_C, SSP_R, SSO_LP, SSO_P, LINE_VOL, LSP, LSO_LP, LSO_P,
    POINT_VOL, LLO, LLP_LP, LLP_P, POINT_SUP, LINE_POINT, LLP_QS
};

// Axis permutations: form axis k reads physical axis PERM[p][k].
// p = 0..5 <-> (0,1,2) (0,2,1) (1,0,2) (1,2,0) (2,0,1) (2,1,0)
HD inline void perm_axes(int p, int pm[3])
{
    pm[0] = p >> 1;
    int r0 = pm[0] == 0 ? 1 : 0, r1 = pm[0] == 2 ? 1 : 2;
    pm[1] = (p & 1) ? r1 : r0;
    pm[2] = (p & 1) ? r0 : r1;
}
HD inline int perm_id(int a, int b)
{
    // permutation whose first two axes are (a, b)
    return a == 0 ? (b == 1 ? 0 : 1) : a == 1 ? (b == 0 ? 2 : 3) : (b == 0 ? 4 : 5);
}

struct Call {
    int form;
    int perm;
    bool swap;  // object1 of the form is box 2
};

HD double eval_form(const Call &k, const Box &b1, const Box &b2)
{
    const Box &P = k.swap ? b2 : b1;
    const Box &Q = k.swap ? b1 : b2;
    int pm[3];
    perm_axes(k.perm, pm);
    double p[3][2], q[3][2], pc[3], qc[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
        p[a][0] = P.lo[pm[a]]; p[a][1] = P.hi[pm[a]]; pc[a] = P.c[pm[a]];
        q[a][0] = Q.lo[pm[a]]; q[a][1] = Q.hi[pm[a]]; qc[a] = Q.c[pm[a]];
    }
    switch (k.form) {
    case VOL_VOL: return f_vol_vol(p[0], p[1], p[2], q[0], q[1], q[2]);
    case VOL_SURF: return f_vol_surf(p[0], p[1], p[2], q[0], q[1], qc[2]);
    case SSP_C: return f_surf_surf_para<false>(p[0], p[1], pc[2], q[0], q[1], qc[2]);
    case SSP_R: return f_surf_surf_para<true>(p[0], p[1], pc[2], q[0], q[1], qc[2]);
    case SSO_LP: return f_surf_surf_ortho<false>(p[0], p[1], pc[2], q[0], qc[1], q[2]);
    case SSO_P: return f_surf_surf_ortho<true>(p[0], p[1], pc[2], q[0], qc[1], q[2]);
    case LINE_VOL: return f_line_vol(p[0], p[1], p[2], q[0], qc[1], qc[2]);
    case LSP: return f_line_surf_para(p[0], p[1], pc[2], q[0], qc[1], qc[2]);
    case LSO_LP: return f_line_surf_ortho<false>(p[0], pc[1], pc[2], qc[0], q[1], q[2]);
    case LSO_P: return f_line_surf_ortho<true>(p[0], pc[1], pc[2], qc[0], q[1], q[2]);
    case POINT_VOL: return f_point_vol(pc[0], pc[1], pc[2], q[0], q[1], q[2]);
    case LLO: return f_two_axis(p[0], qc[0], q[1], pc[1], pc[2] - qc[2], true, false, 1);
    case LLP_LP: return f_line_line_parall<0>(p[0], pc[1], pc[2], q[0], qc[1], qc[2]);
    case LLP_P: return f_line_line_parall<1>(p[0], pc[1], pc[2], q[0], qc[1], qc[2]);
    case LLP_QS: return f_line_line_parall<2>(p[0], pc[1], pc[2], q[0], qc[1], qc[2]);
    case POINT_SUP: return f_two_axis(q[0], pc[0], q[1], pc[1], pc[2] - qc[2], false, false, 0);
    case LINE_POINT: return f_line_point(p[0], pc[1], pc[2], qc[0], qc[1], qc[2]);
    }
    return 0.0;
}

// Shared branch structure of Integ_vol_vol_fft (compute_Lp_Voxels.jl:174-433)
// and Integ_sup_sup_P (compute_row_P_sup.jl:194-379).  s1/s2: suppressed
// axes of box 1/2 (bit a = axis a).  Writes the call and the prefactor.
HD void classify(int s1, int s2, const Box &b1, const Box &b2, bool is_p, Call *k, double *pre,
                         bool *pp)
{
    const double *L1 = b1.len, *L2 = b2.len;
    int n1 = (s1 & 1) + ((s1 >> 1) & 1) + ((s1 >> 2) & 1);
    int n2 = (s2 & 1) + ((s2 >> 1) & 1) + ((s2 >> 2) & 1);
    int sum = n1 + n2;
    *pp = false;
    if (sum == 6) { *pp = true; return; }
    if (sum == 5) {
        // point + line: the line axis is the one free axis of the other box
        bool v1pt = (n1 == 3);
        const Box &ln = v1pt ? b2 : b1;
        int ls = v1pt ? s2 : s1;
        int ax = (ls & 1) == 0 ? 0 : (ls & 2) == 0 ? 1 : 2;
        int o1 = ax == 0 ? 1 : 0, o2 = ax == 2 ? 1 : 2;
        *k = Call{LINE_POINT, perm_id(ax, o1), v1pt};
        (void)o2;
        double V1 = is_p ? L1[0] * L1[1] * L1[2] : L1[0] * L1[1] * L1[2];
        double V2 = L2[0] * L2[1] * L2[2];
        *pre = V1 * V2 / ln.len[ax];
        return;
    }
    if (sum == 4) {
        if (n1 == 3 || n2 == 3) {
            // point + surface: surface normal = the suppressed axis of the surface
            bool v1pt = (n1 == 3);
            int ss = v1pt ? s2 : s1;
            const double *Ls = v1pt ? L2 : L1;
            const double *Lp = v1pt ? L1 : L2;
            int nrm = (ss & 1) ? 0 : (ss & 2) ? 1 : 2;
            // form axes (a, b, nrm): normal x -> (z, y, x), y -> (x, z, y), z -> (x, y, z)
            int pid = nrm == 0 ? 5 : nrm == 1 ? 1 : 0;
            *k = Call{POINT_SUP, pid, !v1pt};
            double V = is_p ? Lp[0] * Lp[1] * Lp[2] : Lp[0] * Lp[1] * Lp[2];
            *pre = V * Ls[nrm];
            return;
        }
        // line-line: line 1 along a1 (its only free axis)
        int a1 = (~s1) & 7, a2 = (~s2) & 7;
        a1 = a1 == 1 ? 0 : a1 == 2 ? 1 : 2;
        a2 = a2 == 1 ? 0 : a2 == 2 ? 1 : 2;
        // ((l1 * l1') * l2) * l2', each box's two cross lengths in axis order
        const int u1 = a1 == 0 ? 1 : 0, v1 = a1 == 2 ? 1 : 2;
        const int u2 = a2 == 0 ? 1 : 0, v2 = a2 == 2 ? 1 : 2;
        if (a1 == a2) {
            // parallel: form axes (a1, the two others in increasing order)
            int o = a1 == 0 ? 1 : 0;
            *k = Call{is_p ? LLP_P : LLP_LP, perm_id(a1, o), false};
        } else {
            // orthogonal: form x = a1, form y = a2
            *k = Call{LLO, perm_id(a1, a2), false};
        }
        *pre = L1[u1] * L1[v1] * L2[u2] * L2[v2];
        return;
    }
    if (sum == 3) {
        if (n1 == 3 || n2 == 3) {
            bool v1pt = (n1 == 3);
            const double *Lp = v1pt ? L1 : L2;
            *k = Call{POINT_VOL, 0, !v1pt};
            *pre = Lp[0] * Lp[1] * Lp[2];
            return;
        }
        // surface + line
        bool v1surf = (n1 == 1);
        int ss = v1surf ? s1 : s2;   // one suppressed axis: the normal
        int ls = v1surf ? s2 : s1;   // two suppressed axes: the free one is the line
        const double *Lsu = v1surf ? L1 : L2, *Lln = v1surf ? L2 : L1;
        int nrm = (ss & 1) ? 0 : (ss & 2) ? 1 : 2;
        int lax = (ls & 1) == 0 ? 0 : (ls & 2) == 0 ? 1 : 2;
        const int lu = lax == 0 ? 1 : 0, lv = lax == 2 ? 1 : 2;
        *pre = Lsu[nrm] * Lln[lu] * Lln[lv];
        if (lax == nrm) {
            // line orthogonal to the surface: form (lax, then the surface axes)
            int o = lax == 0 ? 1 : 0;
            *k = Call{is_p ? LSO_P : LSO_LP, perm_id(lax, o), v1surf};
        } else {
            // line parallel to the surface: form x = lax, y = other in-plane axis, z = nrm
            int other = 3 - lax - nrm;
            *k = Call{LSP, perm_id(lax, other), !v1surf};
        }
        return;
    }
    if (sum == 2) {
        if (n1 == 2 || n2 == 2) {
            // line + volume (Lp only)
            bool v1line = (n1 == 2);
            int ls = v1line ? s1 : s2;
            const double *Lln = v1line ? L1 : L2;
            int lax = (ls & 1) == 0 ? 0 : (ls & 2) == 0 ? 1 : 2;
            double fl = 1.0;
            for (int a = 0; a < 3; a++) if (a != lax) fl *= Lln[a];
            int o = lax == 0 ? 1 : 0;
            *k = Call{LINE_VOL, perm_id(lax, o), v1line};
            *pre = fl;
            return;
        }
        // surface + surface
        int n1a = (s1 & 1) ? 0 : (s1 & 2) ? 1 : 2;
        int n2a = (s2 & 1) ? 0 : (s2 & 2) ? 1 : 2;
        *pre = L1[n1a] * L2[n2a];
        if (n1a == n2a) {
            // parallel: in-plane axes in increasing order, normal last
            int u = n1a == 0 ? 1 : 0, v = n1a == 2 ? 1 : 2;
            *k = Call{is_p ? SSP_R : SSP_C, perm_id(u, v), false};
        } else {
            // orthogonal: form x = shared in-plane axis, y = normal of box 2,
            // z = normal of box 1
            int shared = 3 - n1a - n2a;
            *k = Call{is_p ? SSO_P : SSO_LP, perm_id(shared, n2a), false};
        }
        return;
    }
    if (sum == 1) {
        // surface + volume: the surface is the box with the suppressed axis
        bool v1surf = (n1 == 1);
        int ss = v1surf ? s1 : s2;
        int nrm = (ss & 1) ? 0 : (ss & 2) ? 1 : 2;
        int u = nrm == 0 ? 1 : 0, v = nrm == 2 ? 1 : 2;
        *k = Call{VOL_SURF, perm_id(u, v), v1surf};
        *pre = (v1surf ? L1 : L2)[nrm];
        return;
    }
    *k = Call{VOL_VOL, 0, false};
    *pre = 1.0;
}

HD inline double max4(double a, double b, double c, double d) { return fmax(fmax(a, b), fmax(c, d)); }
HD inline double min4(double a, double b, double c, double d) { return fmin(fmin(a, b), fmin(c, d)); }

HD void axis_gaps(const Box &b1, const Box &b2, double mx[3], double mn[3])
{
#pragma unroll
    for (int a = 0; a < 3; a++) {
        double g0 = fabs(b1.lo[a] - b2.lo[a]), g1 = fabs(b1.lo[a] - b2.hi[a]);
        double g2 = fabs(b1.hi[a] - b2.lo[a]), g3 = fabs(b1.hi[a] - b2.hi[a]);
        mx[a] = max4(g0, g1, g2, g3);
        mn[a] = min4(g0, g1, g2, g3);
    }
}

// check_condition_fft, compute_Lp_Voxels.jl:437-447
HD inline int lp_cond(double V1, double V2, double max_d, double min_R, double sd, double o1, double o2)
{
    double m = min_R + 1e-15;
    double a = V1 * V2 * max_d / (m * m * m);
    double f = sd / m;
    double b = sd / fmax(o1, o2);
    return ((b <= 1e-3 || f < 5e-3) && a < 1e-3) ? 1 : 0;
}

// suppression mask of one box (:97-170): try the thinnest axis first
HD int lp_mask(const double *L, const double mx[3], double V1, double V2, double min_R)
{
    int t = (L[0] <= L[1] && L[0] <= L[2]) ? 0 : (L[1] <= L[0] && L[1] <= L[2]) ? 1 : 2;
    int u = t == 0 ? 1 : 0, v = t == 2 ? 1 : 2;
    int m = 0;
    if (lp_cond(V1, V2, mx[t], min_R, L[t], L[u], L[v])) {
        m |= 1 << t;
        if (fmax(L[u], L[v]) / (min_R + 1e-15) < 3e-1) {
            // the reference evaluates the two remaining axes in x, y, z order
            // with (size, other, other) in (x: a,b,c) (y: b,a,c) (z: c,a,b)
            int oa[3][2] = {{1, 2}, {0, 2}, {0, 1}};
            if (lp_cond(V1, V2, mx[u], min_R, L[u], L[oa[u][0]], L[oa[u][1]])) m |= 1 << u;
            if (lp_cond(V1, V2, mx[v], min_R, L[v], L[oa[v][0]], L[oa[v][1]])) m |= 1 << v;
        }
    }
    return m;
}

// qs: compute_Lp_Voxels_QS.jl's copy (the retarded rows).  Its branches
// differ from compute_Lp_Voxels.jl's in the forms they call; the rows of
// the circulants compare two alike boxes, whose suppression is symmetric, so
// only the vol-vol (unchanged: the QS dispatcher calls Integ_vol_vol_fft,
// :418), parallel-surface (real log), parallel-line and far-field branches
// are reached, and the others return NaN in this mode.
HD double integ_vol_vol(const Box &b1, const Box &b2, bool qs = false)
{
    double mx[3], mn[3];
    axis_gaps(b1, b2, mx, mn);
    double min_R = sqrt(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
    double V1 = b1.len[0] * b1.len[1] * b1.len[2], V2 = b2.len[0] * b2.len[1] * b2.len[2];
    int s1 = lp_mask(b1.len, mx, V1, V2, min_R);
    int s2 = lp_mask(b2.len, mx, V1, V2, min_R);
    Call k;
    double pre;
    bool pp;
    classify(s1, s2, b1, b2, false, &k, &pre, &pp);
    if (pp) {
        double dx = b1.c[0] - b2.c[0], dy = b1.c[1] - b2.c[1], dz = b1.c[2] - b2.c[2];
        return V1 * V2 / sqrt(dx * dx + dy * dy + dz * dz);
    }
    if (qs && k.form != VOL_VOL) {
        if (k.form == SSP_C) k.form = SSP_R;
        else if (k.form == LLP_LP) k.form = LLP_QS;
        else return __builtin_nan("");
    }
    if (k.form == VOL_VOL) return eval_form(k, b1, b2);
    return pre * eval_form(k, b1, b2);
}

// check_condition_P_P, compute_row_P_sup.jl:383-394
HD inline int p_cond(double S1, double S2, double max_d, double min_R, double sd, double o1, double o2)
{
    double mo = o1 * o2;
    double m = min_R + 1e-15;
    double cS = fmax(mo, sd) / m;
    double a = S1 * S2 * max_d / (m * m * m);
    double f = sd / m;
    double b = sd / mo;
    return (cS < 3e-1 && (b <= 1e-3 || f < 5e-3) && a < 1e-3) ? 1 : 0;
}

HD double integ_sup_sup(Box b1, Box b2)
{
    // the flat axis of each face carries a unit fake thickness (:128-146)
    int f1 = (b1.len[0] <= b1.len[1] && b1.len[0] <= b1.len[2]) ? 0 : (b1.len[1] <= b1.len[0] && b1.len[1] <= b1.len[2]) ? 1 : 2;
    int f2 = (b2.len[0] <= b2.len[1] && b2.len[0] <= b2.len[2]) ? 0 : (b2.len[1] <= b2.len[0] && b2.len[1] <= b2.len[2]) ? 1 : 2;
    b1.len[f1] = 1.0;
    b2.len[f2] = 1.0;
    double sup1 = b1.len[0] * b1.len[1] * b1.len[2], sup2 = b2.len[0] * b2.len[1] * b2.len[2];
    double mx[3], mn[3];
    axis_gaps(b1, b2, mx, mn);
    double min_R = sqrt(mn[0] * mn[0] + mn[1] * mn[1] + mn[2] * mn[2]);
    const int oa[3][2] = {{1, 2}, {0, 2}, {0, 1}};
    int s1 = 0, s2 = 0;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (p_cond(sup1, sup2, mx[a], min_R, b1.len[a], b1.len[oa[a][0]], b1.len[oa[a][1]])) s1 |= 1 << a;
        if (p_cond(sup1, sup2, mx[a], min_R, b2.len[a], b2.len[oa[a][0]], b2.len[oa[a][1]])) s2 |= 1 << a;
    }
    s1 |= 1 << f1;
    s2 |= 1 << f2;
    Call k;
    double pre;
    bool pp;
    classify(s1, s2, b1, b2, true, &k, &pre, &pp);
    if (pp) {
        double dx = b1.c[0] - b2.c[0], dy = b1.c[1] - b2.c[1], dz = b1.c[2] - b2.c[2];
        return sup1 * sup2 / sqrt(dx * dx + dy * dy + dz * dz);
    }
    return pre * eval_form(k, b1, b2);
}

HD Box lp_box(double cx, double cy, double cz, double sx, double sy, double sz)
{
    Box b;
    double c[3] = {cx, cy, cz}, s[3] = {sx, sy, sz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        b.lo[a] = round14(c[a] - s[a] / 2);
        b.hi[a] = round14(c[a] + s[a] / 2);
        b.c[a] = 0.5 * (b.lo[a] + b.hi[a]);
        b.len[a] = fabs(b.hi[a] - b.lo[a]);
    }
    return b;
}

// face of a centre for caser 1..6 (compute_row_P_sup.jl:3-51); area in *A
HD Box p_face(int caser, double cx, double cy, double cz, double sx, double sy, double sz, double *t1,
                      double *t2)
{
    Box b;
    double c[3] = {cx, cy, cz};
    double s[3] = {sx, sy, sz};
    int nrm = caser <= 2 ? 1 : (caser <= 4 ? 0 : 2);  // caser 1,2: y-faces; 3,4: x-faces; 5,6: z-faces
    double sg = (caser & 1) ? -1.0 : 1.0;
    c[nrm] = c[nrm] + sg * (s[nrm] / 2);
    s[nrm] = 0.0;
    // tangential lengths in increasing axis order: (sy,sz) (sx,sz) (sx,sy)
    *t1 = nrm == 0 ? sy : sx;
    *t2 = nrm == 2 ? sy : sz;
#pragma unroll
    for (int a = 0; a < 3; a++) {
        b.lo[a] = c[a] - s[a] / 2;
        b.hi[a] = c[a] + s[a] / 2;
        b.c[a] = c[a];
        b.len[a] = s[a];
    }
    return b;
}

// Reference centre arithmetic (x, y, z of entry i of a circulant centre set).
//   voxel centre  min + s*(i-1) + s/2          build_center_air_Voxels.jl:238-240
//   Lp edge       0.5*(A + B)                  build_centers_Lp_with_air.jl:179
//   half-cell     min + s*(ceil(i/2)-1) + s/2  build_center_P_Voxels.jl:63-77
HD inline double vox_c(double mn, double s, int i1) { return mn + s * (double)(i1 - 1) + s / 2; }
HD inline double half_c(double mn, double s, int i1) { return mn + s * (double)((i1 + 1) / 2 - 1) + s / 2; }

// variant 0: the static row of the quasi-static path; 1: the static row of
// the retarded path (compute_rows_Rcc_Lp.jl / compute_rows_Rcc_P.jl: the Lp
// rows from compute_Lp_Voxels_QS.jl's closed forms); 2: the centre
// distances Rcc of the retarded path (compute_Voxels_Rcc.jl,
// compute_P_vox_Rcc.jl).
struct GridRow {
    int kind;
    int variant;
    int dims[3];
    double origin[3];
    double s[3];
    double scale;
};

HD void grid_center(const GridRow &g, int64_t idx, double c[3])
{
    int i = (int)(idx % g.dims[0]);
    int64_t r = idx / g.dims[0];
    int j = (int)(r % g.dims[1]);
    int k = (int)(r / g.dims[1]);
    int I[3] = {i + 1, j + 1, k + 1};
    if (g.kind <= 2) {
        int d = g.kind;
#pragma unroll
        for (int a = 0; a < 3; a++) {
            double A = vox_c(g.origin[a], g.s[a], I[a]);
            double B = vox_c(g.origin[a], g.s[a], I[a] + (a == d ? 1 : 0));
            c[a] = 0.5 * (A + B);
        }
    } else if (g.kind <= 5) {
#pragma unroll
        for (int a = 0; a < 3; a++) c[a] = vox_c(g.origin[a], g.s[a], I[a]);
    } else {
        // 6: p1234 half-cells in x, y; 7: p1256 in y, z; 8: p3456 in x, z
        bool h[3] = {g.kind == 6 || g.kind == 8, g.kind == 6 || g.kind == 7, g.kind == 7 || g.kind == 8};
#pragma unroll
        for (int a = 0; a < 3; a++) c[a] = h[a] ? half_c(g.origin[a], g.s[a], I[a]) : vox_c(g.origin[a], g.s[a], I[a]);
    }
}

// (caser, caser2) of the six P rows, compute_FFT_mutual_coupling_mats.jl:96-121
HD inline void p_cases(int r, int *c1, int *c2)
{
    const int a[6] = {1, 3, 5, 3, 5, 5}, b[6] = {1, 3, 5, 2, 2, 4};
    *c1 = a[r];
    *c2 = b[r];
}

HD inline double lp_entry(const Box &src, double cx, double cy, double cz, double sx2, double sy2, double sz2,
                          double pref, double scale)
{
    Box b2 = lp_box(cx, cy, cz, sx2, sy2, sz2);
    return scale * (pref * integ_vol_vol(src, b2));
}

HD inline double p_entry(const Box &src, double A1, double cx, double cy, double cz, double sx, double sy, double sz,
                         int caser2, double scale)
{
    double u1, u2;
    Box b2 = p_face(caser2, cx, cy, cz, sx, sy, sz, &u1, &u2);
    const double pi = 3.141592653589793, eps0 = 8.854187816997944e-12;
    return scale * (1 / (4 * pi * eps0 * A1 * u1 * u2) * integ_sup_sup(src, b2));
}

__global__ void __launch_bounds__(256) k_lp_row(Box src, const double *__restrict__ cen, int64_t N,
                                                double sx2, double sy2, double sz2, double pref,
                                                double scale, double *__restrict__ out)
{
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = lp_entry(src, cen[n], cen[n + N], cen[n + 2 * N], sx2, sy2, sz2, pref, scale);
}

__global__ void __launch_bounds__(256) k_p_row(Box src, const double *__restrict__ cen, int64_t N, double sx,
                                               double sy, double sz, int caser2, double A1, double scale,
                                               double *__restrict__ out)
{
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = p_entry(src, A1, cen[n], cen[n + N], cen[n + 2 * N], sx, sy, sz, caser2, scale);
}

// Entry n of circulant row g.kind, centres generated from the grid (the
// device kernel below and the host test hook share this code).
HD double grid_row_entry(const GridRow &g, int64_t n)
{
    double c0[3], c[3];
    grid_center(g, 0, c0);
    grid_center(g, n, c);
    double v;
    if (g.variant == 2) {
        // |c0 - c| (Lp centres) or |face(c0, caser) - face(c, caser2)| (P),
        // LinearAlgebra.norm of the 3-vector: sqrt((d1^2 + d2^2) + d3^2)
        if (g.kind > 2) {
            int cs[2];
            p_cases(g.kind - 3, &cs[0], &cs[1]);
            int n1 = cs[0] <= 2 ? 1 : (cs[0] <= 4 ? 0 : 2), n2 = cs[1] <= 2 ? 1 : (cs[1] <= 4 ? 0 : 2);
            c0[n1] = (cs[0] & 1) ? c0[n1] - g.s[n1] / 2 : c0[n1] + g.s[n1] / 2;
            c[n2] = (cs[1] & 1) ? c[n2] - g.s[n2] / 2 : c[n2] + g.s[n2] / 2;
        }
        const double d0 = c0[0] - c[0], d1 = c0[1] - c[1], d2 = c0[2] - c[2];
        return g.scale * sqrt((d0 * d0 + d1 * d1) + d2 * d2);
    }
    if (g.kind <= 2) {
        Box b1 = lp_box(c0[0], c0[1], c0[2], g.s[0], g.s[1], g.s[2]);
        Box b2 = lp_box(c[0], c[1], c[2], g.s[0], g.s[1], g.s[2]);
        int d = g.kind;
        double S = d == 0 ? g.s[1] * g.s[2] * g.s[1] * g.s[2]
                 : d == 1 ? g.s[0] * g.s[2] * g.s[0] * g.s[2] : g.s[0] * g.s[1] * g.s[0] * g.s[1];
        v = 1e-7 / S * integ_vol_vol(b1, b2, g.variant == 1);
        if (n == 0) v = 0.0;  // the self entry is carried by DZ (compute_FFT_mutual_coupling_mats.jl:56)
    } else {
        int cs[2];
        p_cases(g.kind - 3, &cs[0], &cs[1]);
        double a1, a2, u1, u2;
        Box b1 = p_face(cs[0], c0[0], c0[1], c0[2], g.s[0], g.s[1], g.s[2], &a1, &a2);
        Box b2 = p_face(cs[1], c[0], c[1], c[2], g.s[0], g.s[1], g.s[2], &u1, &u2);
        const double pi = 3.141592653589793, eps0 = 8.854187816997944e-12;
        v = 1 / (4 * pi * eps0 * (a1 * a2) * u1 * u2) * integ_sup_sup(b1, b2);
    }
    return g.scale * v;
}

__global__ void __launch_bounds__(256) k_grid_row(GridRow g, int64_t N, double *__restrict__ out)
{
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    out[n] = grid_row_entry(g, n);
}

// Test probe: scalar building blocks evaluated on the device.
// fn 0 clog_re(x), 1 log(x), 2 atan2(x, y), 3 sqrt(x), 4 pow4(x),
// 5 lp self-style entry with box sizes (x, y, x*0.2) and offset y
__global__ void k_probe(int fn, const double *x, const double *y, int64_t n, double *out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = x[i], b = y[i];
    double r = 0.0;
    switch (fn) {
    case 0: r = clog_re(a); break;
    case 1: r = det_log(a); break;
    case 2: r = det_atan2(a, b); break;
    case 3: r = sqrt(a); break;
    case 4: r = pow4(a); break;
    }
    out[i] = r;
}

HD double probe_host(int fn, double a, double b)
{
    switch (fn) {
    case 0: return clog_re(a);
    case 1: return det_log(a);
    case 2: return det_atan2(a, b);
    case 3: return sqrt(a);
    case 4: return pow4(a);
    }
    return 0.0;
}

}  // namespace

extern "C" int peec_test_probe(int device_side, int fn, const double *x, const double *y, int64_t n, double *out)
{
    if (!device_side) {
        for (int64_t i = 0; i < n; i++) out[i] = probe_host(fn, x[i], y[i]);
        return PEEC_OK;
    }
    double *dx, *dy, *dout;
    PEEC_HIP_TRY(hipMalloc(&dx, n * sizeof(double)));
    PEEC_HIP_TRY(hipMalloc(&dy, n * sizeof(double)));
    PEEC_HIP_TRY(hipMalloc(&dout, n * sizeof(double)));
    PEEC_HIP_TRY(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
    PEEC_HIP_TRY(hipMemcpy(dy, y, n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, fn, dx, dy, n, dout);
    PEEC_HIP_TRY(hipGetLastError());
    PEEC_HIP_TRY(hipMemcpy(out, dout, n * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dout);
    return PEEC_OK;
}

static Box host_lp_box(const double *c, double sx, double sy, double sz)
{
    Box b;
    double s[3] = {sx, sy, sz};
    for (int a = 0; a < 3; a++) {
        b.lo[a] = rint((c[a] - s[a] / 2) * 1e14) / 1e14;
        b.hi[a] = rint((c[a] + s[a] / 2) * 1e14) / 1e14;
        b.c[a] = 0.5 * (b.lo[a] + b.hi[a]);
        b.len[a] = fabs(b.hi[a] - b.lo[a]);
    }
    return b;
}

static Box host_p_face(int caser, const double *cc, double sx, double sy, double sz, double *A)
{
    Box b;
    double c[3] = {cc[0], cc[1], cc[2]}, s[3] = {sx, sy, sz};
    int nrm = caser <= 2 ? 1 : (caser <= 4 ? 0 : 2);  // caser 1,2: y-faces; 3,4: x-faces; 5,6: z-faces
    double sg = (caser & 1) ? -1.0 : 1.0;
    c[nrm] = c[nrm] + sg * (s[nrm] / 2);
    s[nrm] = 0.0;
    *A = nrm == 0 ? sy * sz : nrm == 1 ? sx * sz : sx * sy;
    for (int a = 0; a < 3; a++) {
        b.lo[a] = c[a] - s[a] / 2;
        b.hi[a] = c[a] + s[a] / 2;
        b.c[a] = c[a];
        b.len[a] = s[a];
    }
    return b;
}

int peec_lp_row_device(hipStream_t s, const double src[3], const double *d_centers, int64_t N, double sx,
                       double sy, double sz, double sx2, double sy2, double sz2, int dc, double scale,
                       double *d_out)
{
    if (N <= 0) return PEEC_OK;
    double S = sy * sz * sy2 * sz2;
    if (dc == 2) S = sx * sz * sx2 * sz2;
    else if (dc == 3) S = sx * sy * sx2 * sy2;
    Box b1 = host_lp_box(src, sx, sy, sz);
    int64_t nb = (N + 255) / 256;
    hipLaunchKernelGGL(k_lp_row, dim3((unsigned)nb), dim3(256), 0, s, b1, d_centers, N, sx2, sy2, sz2,
                       1e-7 / S, scale, d_out);
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

int peec_p_row_device(hipStream_t s, const double src[3], const double *d_centers, int64_t N, double sx,
                      double sy, double sz, int caser, int caser2, double scale, double *d_out)
{
    if (N <= 0) return PEEC_OK;
    double A1;
    Box b1 = host_p_face(caser, src, sx, sy, sz, &A1);
    int64_t nb = (N + 255) / 256;
    hipLaunchKernelGGL(k_p_row, dim3((unsigned)nb), dim3(256), 0, s, b1, d_centers, N, sx, sy, sz, caser2, A1,
                       scale, d_out);
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// compact grid dims of each circulant row (compute_FFT_mutual_coupling_mats.jl:57-125)
static void grid_dims(int kind, int Nx, int Ny, int Nz, int d[3])
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

static GridRow make_grid_row(int kind, const double origin[3], int Nx, int Ny, int Nz, double sx, double sy,
                             double sz, double scale, int variant = 0)
{
    GridRow g;
    g.kind = kind;
    g.variant = variant;
    grid_dims(kind, Nx, Ny, Nz, g.dims);
    for (int a = 0; a < 3; a++) g.origin[a] = origin[a];
    g.s[0] = sx; g.s[1] = sy; g.s[2] = sz;
    g.scale = scale;
    return g;
}

int peec_grid_row_device(hipStream_t s, int kind, const double origin[3], int Nx, int Ny, int Nz, double sx,
                         double sy, double sz, double scale, double *d_out, int64_t *n_out, int dims_out[3],
                         int variant)
{
    if (variant < 0 || variant > 2) return PEEC_ERR_ARG;
    GridRow g = make_grid_row(kind, origin, Nx, Ny, Nz, sx, sy, sz, scale, variant);
    int64_t N = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
    if (n_out) *n_out = N;
    if (dims_out) for (int a = 0; a < 3; a++) dims_out[a] = g.dims[a];
    if (!d_out || N <= 0) return PEEC_OK;
    int64_t nb = (N + 255) / 256;
    hipLaunchKernelGGL(k_grid_row, dim3((unsigned)nb), dim3(256), 0, s, g, N, d_out);
    PEEC_HIP_TRY(hipGetLastError());
    return PEEC_OK;
}

// Test hooks: the kernels' per-entry code run on the host CPU, so tests can
// separate classifier logic from device libm differences.  Not used by any
// product path.
extern "C" int peec_test_lp_row_host(const double *src, const double *cen, int64_t N, double sx, double sy,
                                     double sz, double sx2, double sy2, double sz2, int dc, double *out)
{
    double S = sy * sz * sy2 * sz2;
    if (dc == 2) S = sx * sz * sx2 * sz2;
    else if (dc == 3) S = sx * sy * sx2 * sy2;
    Box b1 = host_lp_box(src, sx, sy, sz);
    for (int64_t n = 0; n < N; n++) out[n] = lp_entry(b1, cen[n], cen[n + N], cen[n + 2 * N], sx2, sy2, sz2, 1e-7 / S, 1.0);
    return PEEC_OK;
}

// Entries idx[0..count) of circulant row `kind` (0..8) on an Nx x Ny x Nz
// grid, evaluated by k_grid_row's per-entry code (device centre arithmetic
// included) on the host; returns the row length in *n_out.  kind + 16 *
// variant selects the retarded path's static rows (variant 1) or distances
// (variant 2).
extern "C" int peec_test_grid_row_host(int kind, const double *origin, int Nx, int Ny, int Nz, double sx,
                                       double sy, double sz, double scale, const int64_t *idx, int64_t count,
                                       double *out, int64_t *n_out)
{
    const int variant = kind / 16;
    kind %= 16;
    if (kind < 0 || kind > 8 || variant > 2 || !origin || Nx < 1 || Ny < 1 || Nz < 1) return PEEC_ERR_ARG;
    GridRow g = make_grid_row(kind, origin, Nx, Ny, Nz, sx, sy, sz, scale, variant);
    const int64_t N = (int64_t)g.dims[0] * g.dims[1] * g.dims[2];
    if (n_out) *n_out = N;
    for (int64_t i = 0; i < count; i++) {
        if (idx[i] < 0 || idx[i] >= N) return PEEC_ERR_ARG;
        out[i] = grid_row_entry(g, idx[i]);
    }
    return PEEC_OK;
}

extern "C" int peec_test_p_row_host(const double *src, const double *cen, int64_t N, double sx, double sy, double sz,
                                    int caser, int caser2, double *out)
{
    double A1;
    Box b1 = host_p_face(caser, src, sx, sy, sz, &A1);
    for (int64_t n = 0; n < N; n++) out[n] = p_entry(b1, A1, cen[n], cen[n + N], cen[n + 2 * N], sx, sy, sz, caser2, 1.0);
    return PEEC_OK;
}

extern "C" int peec_lp_self(double l, double W, double T, double *out)
{
    if (!out || !(l > 0) || !(W > 0) || !(T > 0)) {
        peec_set_error("peec_lp_self: lengths must be positive");
        return PEEC_ERR_ARG;
    }
    // compute_diagonals.jl:274-293 (Ruehli's rectangular-bar self inductance)
    const double pi = 3.141592653589793;
    double w = W / l, t = T / l;
    double r = sqrt(w * w + t * t), aw = sqrt(w * w + 1), at = sqrt(t * t + 1), ar = sqrt(w * w + t * t + 1);
    double mu0 = 4 * pi * 1e-7;
    double A = 1.0 / 4 * (1 / w * asinh(w / at) + 1 / t * asinh(t / aw) + asinh(1 / r));
    double B = 1.0 / 24 * ((t * t) / w * asinh(w / (t * at * (r + ar))) + (w * w) / t * asinh(t / (w * aw * (r + ar))) +
                           (t * t) / (w * w) * asinh((w * w) / (t * r * (at + ar))) +
                           (w * w) / (t * t) * asinh((t * t) / (w * r * (aw + ar))) +
                           1 / (w * (t * t)) * asinh(w * (t * t) / (at * (aw + ar))) +
                           1 / (t * (w * w)) * asinh(t * (w * w) / (aw * (at + ar))));
    double C = 1.0 / 6 * (1 / (w * t) * atan(w * t / ar) + t / w * atan(w / (t * ar)) + w / t * atan(t / (w * ar)));
    double D = 1.0 / 60 * ((ar + r + t + at) * (t * t) / ((ar + r) * (r + t) * (t + at) * (at + ar)) +
                           (ar + r + w + aw) * (w * w) / ((ar + r) * (r + w) * (w + aw) * (aw + ar)) +
                           (ar + aw + 1 + at) / ((ar + aw) * (aw + 1) * (1 + at) * (at + ar)));
    double E = 1.0 / 20 * (1 / (r + ar) + 1 / (aw + ar) + 1 / (at + ar));
    *out = 2 * mu0 * l / pi * (A + B - C - D - E);
    return PEEC_OK;
}
