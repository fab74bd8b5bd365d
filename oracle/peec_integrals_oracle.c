/*
 * ORACLE -- test infrastructure only.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this; the product path never does.
 *
 * Plain-C restatement of the reference's closed-form partial-element
 * integrals (Draco112358/Solver, Julia):
 *   - Lp rows:  src/lib/solverFFT/compute_Lp_Voxels.jl (whole file)
 *   - P rows:   src/lib/solverFFT/compute_row_P_sup.jl (whole file)
 *               + integ_surf_surf_para from compute_Lp_Voxels_QS.jl:695-730
 *   - Lp_self:  src/lib/solverFFT/compute_diagonals.jl:136-155
 *
 * Julia arithmetic details reproduced on purpose:
 *   x^2, x^3 -> x*x, x*x*x   (Base.literal_pow)
 *   x^4      -> compensated power-by-squaring (Base.Math.pow_body)
 *   real(or_det_log(complex(a))) -> 0.5*or_det_log(a*a) (+log1p branch near |a|=1), i.e.
 *             Julia's Complex log real part for an imaginary part of zero
 *   round(x*10^14)/10^14 with ties-to-even (rint)
 *   atan(y, x) -> atan2
 * log, log1p and atan2 are the deterministic fdlibm-algorithm versions in
 * detmath_oracle.h (the GPU kernels evaluate the same operation sequence),
 * so far entries -- sums of corner terms that cancel by many orders of
 * magnitude -- come out identical on both sides.
 * Built with -ffp-contract=off so that no multiply-add is fused.
 */
#include <math.h>
#include <stdint.h>

#include "detmath_oracle.h"

#define ISBAD(t) (isnan(t) || isinf(t))

/* ---- routing trace (fixtures and branch statistics only) ---------------
 * g_form: id of the last closed form entered (the ids of
 * oracle_closed_form below; 16 = the point-point far-field term of
 * sum_supp == 6, 17 / 18 = the Lp / P collinear-segment forms).
 * g_tie: set when a suppression test's quantity lies within a relative
 * 1e-9 of its threshold, i.e. where the branch taken depends on the last
 * bits of the geometry arithmetic. */
static int g_form = -1;
static int g_tie = 0;
/* 1 while a row of compute_Lp_Voxels_QS.jl (the closed-form copy the
 * retarded Rcc rows use) is evaluated: the dispatcher then resolves to that
 * file's forms (see the QS block below) */
static int g_qs = 0;

static int near_thr(double v, double thr) { return fabs(v - thr) <= 1e-9 * thr; }

/* ---- Julia scalar helpers ------------------------------------------- */

/* Base.Math.pow_body(x, 4): two_mul based compensated squaring. */
static double jl_pow4(double x)
{
    double x2 = x * x;
    double e1 = fma(x, x, -x2);
    double err = x2 * 2.0 * e1;
    double x4 = x2 * x2;
    double e2 = fma(x2, x2, -x4);
    double lo = e2 + err;
    if (isfinite(x4) && isfinite(lo))
        return fma(x4, 1.0, lo);
    return x4;
}

/* real(or_det_log(complex(a))) for real a (Base or_det_log(::Complex{Float64})). */
static double jl_clog_re(double a)
{
    double b = fabs(a);
    if (b == 0.0)
        return -INFINITY;
    if (isinf(b))
        return INFINITY;
    if (b < 1e-140 || b > 1e140)
        return or_det_log(b);
    double rho = a * a;
    if (0.5 < b * b && (b <= 1.25 || rho < 3.0))
        return or_det_log1p((b - 1.0) * (b + 1.0)) / 2.0;
    return or_det_log(rho) / 2.0;
}

/* round_ud_fft(v, 14), compute_Lp_Voxels.jl:807-814 */
static double jl_round14(double v)
{
    return rint(v * 1e14) / 1e14;
}

static double sgnpow(int k) { return (k & 1) ? -1.0 : 1.0; }

/* ---- Lp closed forms: compute_Lp_Voxels.jl:449-805 -------------------- */

/* :449-467 (note the precedence of `isnan || isinf && x1>0` at :457) */
static double lp_line_point(const double *x1v, double y3, double z3,
                            double x1, double y1, double z1)
{
    g_form = 11;
    double x3 = x1v[0], x4 = x1v[1];
    double check = 1.0 / (x1 - x3);
    if (ISBAD(check))
        x1 = x1 - 1e-8;
    check = 1.0 / (x1 - x4);
    if (isnan(check) || (isinf(check) && x1 > 0))
        x1 = x1 + 1e-8;
    double R1 = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double R2 = sqrt((x1 - x4) * (x1 - x4) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double Ip = or_det_log((x1 - x3) + R1) - or_det_log((x1 - x4) + R2);
    if (ISBAD(Ip))
        Ip = (x1 - x3) / fabs(x1 - x3) * or_det_log(fabs(x1 - x3)) - (x1 - x4) / fabs(x1 - x4) * or_det_log(fabs(x1 - x4));
    return Ip;
}

/* :469-489 and (identical arithmetic) compute_row_P_sup.jl:416-436 */
static double point_sup(double x1, double y1, double z1,
                        const double *x2v, const double *y2v, double z2)
{
    g_form = 10;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x2 = x2v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y2 = y2v[c2 - 1];
            double R = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
            double t1 = (x1 - x2) * jl_clog_re((y1 - y2) + R);
            if (ISBAD(t1)) t1 = 0.0;
            double t2 = (y1 - y2) * jl_clog_re((x1 - x2) + R);
            if (ISBAD(t2)) t2 = 0.0;
            double t3 = -fabs(z1 - z2) * or_det_atan2((x1 - x2) * (y1 - y2), (fabs(z1 - z2) * R));
            sol = sol + sgnpow(c1 + c2) * (t1 + t2 + t3);
        }
    }
    return sol;
}

/* Lp copy :514-529 */
static double lp_line_line_sp(const double *x1v, const double *x2v)
{
    g_form = 17;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double x2 = x2v[c2 - 1];
            double R = sqrt((x1 - x2) * (x1 - x2));
            double t1 = (x1 - x2) / R * (x1 - x2 * jl_clog_re(x2 - x1));
            if (ISBAD(t1)) t1 = 0.0;
            sol = sol + sgnpow(c1 + c2 + 1) * t1;
        }
    }
    return sol;
}

/* P copy compute_row_P_sup.jl:461-476 (adds x1*log|x1-x2|) */
static double p_line_line_sp(const double *x1v, const double *x2v)
{
    g_form = 18;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double x2 = x2v[c2 - 1];
            double R = sqrt((x1 - x2) * (x1 - x2));
            double t1 = (x1 - x2) / R * (x1 - x2 * jl_clog_re(x2 - x1)) + x1 * jl_clog_re(x1 - x2);
            if (ISBAD(t1)) t1 = 0.0;
            sol = sol + sgnpow(c1 + c2 + 1) * t1;
        }
    }
    return sol;
}

/* ---- compute_Lp_Voxels_QS.jl: the closed forms that differ from
 * compute_Lp_Voxels.jl.  Every real(log((d) + R)) of these forms is the real
 * log of a non-negative real (R >= |d|), not the real part of a complex log;
 * the forms not listed here are the same arithmetic in both files. ------ */

/* QS :501-519: x1*log|x1-x2| inside the (x1-x2)/R factor */
static double qs_line_line_sp(const double *x1v, const double *x2v)
{
    g_form = 19;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double x2 = x2v[c2 - 1];
            double R = sqrt((x1 - x2) * (x1 - x2));
            double t1 = (x1 - x2) / R * (x1 - x2 * jl_clog_re(x2 - x1) + x1 * jl_clog_re(x1 - x2));
            if (ISBAD(t1)) t1 = 0.0;
            sol += sgnpow(c1 + c2 + 1) * t1;
        }
    }
    return sol;
}

/* QS :418-442: `(isnan(check) || isinf(check)) && x1 > 0` */
static double qs_line_point(const double *x1v, double y3, double z3,
                            double x1, double y1, double z1)
{
    g_form = 20;
    double x3 = x1v[0], x4 = x1v[1];
    double check = 1.0 / (x1 - x3);
    if (ISBAD(check))
        x1 -= 1e-8;
    check = 1.0 / (x1 - x4);
    if (ISBAD(check) && x1 > 0)
        x1 += 1e-8;
    double R1 = sqrt((x1 - x3) * (x1 - x3) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double R2 = sqrt((x1 - x4) * (x1 - x4) + (y1 - y3) * (y1 - y3) + (z1 - z3) * (z1 - z3));
    double Ip = (or_det_log((x1 - x3) + R1)) - (or_det_log((x1 - x4) + R2));
    if (ISBAD(Ip))
        Ip = (x1 - x3) / fabs(x1 - x3) * or_det_log(fabs(x1 - x3)) - (x1 - x4) / fabs(x1 - x4) * or_det_log(fabs(x1 - x4));
    return Ip;
}

/* QS :595-625: three log and three atan terms */
static double qs_line_surf_ortho(const double *x1v, double y1, double z1,
                                 double x2, const double *y2v, const double *z2v)
{
    g_form = 21;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y2 = y2v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z2 = z2v[c3 - 1];
                double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                double R = sqrt(dx * dx + dy * dy + dz * dz);
                double t1 = dy * dz * or_det_log(dx + R);
                if (ISBAD(t1)) t1 = 0.0;
                double t2 = dx * dz * or_det_log(dy + R);
                if (ISBAD(t2)) t2 = 0.0;
                double t3 = dy * dx * or_det_log(dz + R);
                if (ISBAD(t3)) t3 = 0.0;
                double t4 = -0.5 * fabs(dz) * dz * or_det_atan2(dx * dy, fabs(dz) * R);
                double t5 = -0.5 * fabs(dy) * dy * or_det_atan2(dx * dz, fabs(dy) * R);
                double t6 = -0.5 * fabs(dx) * dx * or_det_atan2(dy * dz, fabs(dx) * R);
                sol += sgnpow(c1 + c2 + c3) * (t1 + t2 + t3 + t4 + t5 + t6);
            }
        }
    }
    return sol;
}

/* QS :627-654 */
static double qs_line_surf_para(const double *x1v, const double *y1v, double z1,
                                const double *x2v, double y2, double z2)
{
    g_form = 22;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double x2 = x2v[c3 - 1];
                double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                double R = sqrt(dx * dx + dy * dy + dz * dz);
                double t1 = dx * dy * or_det_log(dx + R);
                if (ISBAD(t1)) t1 = 0.0;
                double t2 = (dx * dx - dz * dz) / 2 * or_det_log(dy + R);
                if (ISBAD(t2)) t2 = 0.0;
                double t3 = -dx * fabs(dz) * or_det_atan2(dx * dy, fabs(dz) * R);
                double t4 = -dy / 2 * R;
                sol += sgnpow(c1 + c2 + c3 + 1) * (t1 + t2 + t3 + t4);
            }
        }
    }
    return sol;
}

/* QS :656-693 */
static double qs_line_vol(const double *x1v, const double *y1v, const double *z1v,
                          const double *x2v, double y2, double z2)
{
    g_form = 23;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z1 = z1v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    double t1 = -1.0 / 3 * dy * dz * R;
                    double t2 = -1.0 / 2 * dx * fabs(dz) * dz * or_det_atan2(dx * dy, fabs(dz) * R);
                    double t3 = -1.0 / 2 * dx * fabs(dy) * dy * or_det_atan2(dx * dz, fabs(dy) * R);
                    double t4 = -1.0 / 6 * (fabs(dx) * fabs(dx) * fabs(dx)) * or_det_atan2(dy * dz, fabs(dx) * R);
                    double t5 = dx * dy * dz * or_det_log(dx + R);
                    if (ISBAD(t5)) t5 = 0.0;
                    double t6 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dz * dz)) * dz * or_det_log(dy + R);
                    if (ISBAD(t6)) t6 = 0.0;
                    double t7 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dy * dy)) * dy * or_det_log(dz + R);
                    if (ISBAD(t7)) t7 = 0.0;
                    sol += sgnpow(c1 + c2 + c3 + c4 + 1) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
                }
            }
        }
    }
    return sol;
}

/* QS :777-817: terms 2 and 3 carry a (z1 - z2) factor */
static double qs_vol_surf(const double *x1v, const double *y1v, const double *z1v,
                          const double *x2v, const double *y2v, double z2)
{
    g_form = 24;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z1 = z1v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    for (int c5 = 1; c5 <= 2; c5++) {
                        double y2 = y2v[c5 - 1];
                        double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                        double R = sqrt(dx * dx + dy * dy + dz * dz);
                        double t1 = ((dz * dz) / 12 - ((dx * dx) + (dy * dy)) / 8) * dz * R;
                        double t2 = ((dy * dy) / 2 - (dz * dz) / 6) * dx * dz * or_det_log(dx + R);
                        if (ISBAD(t2)) t2 = 0.0;
                        double t3 = ((dx * dx) / 2 - (dz * dz) / 6) * dy * dz * or_det_log(dy + R);
                        if (ISBAD(t3)) t3 = 0.0;
                        double t4 = (-jl_pow4(dx) / 24 - jl_pow4(dy) / 24 + ((dy * dy) * (dx * dx)) / 4) * or_det_log(dz + R);
                        if (ISBAD(t4)) t4 = 0.0;
                        double t5 = -(fabs(dx) * fabs(dx) * fabs(dx)) * dy / 6 * or_det_atan2(dy * dz, fabs(dx) * R);
                        double t6 = -dx * (fabs(dy) * fabs(dy) * fabs(dy)) / 6 * or_det_atan2(dx * dz, fabs(dy) * R);
                        double t7 = -dx * dy * fabs(dz) * dz / 2 * or_det_atan2(dx * dy, fabs(dz) * R);
                        sol += sgnpow(c1 + c2 + c3 + c4 + c5) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
                    }
                }
            }
        }
    }
    return sol;
}

/* :491-512 / compute_row_P_sup.jl:438-459 */
static double line_line_parall(const double *x1v, double y1, double z1,
                               const double *x2v, double y2, double z2, int is_p)
{
    g_form = is_p ? 15 : 9;
    double dy = y1 - y2, dz = z1 - z2;
    if (fabs(dy) < 1e-10 && fabs(dz) < 1e-10)
        return is_p ? p_line_line_sp(x1v, x2v) : g_qs ? qs_line_line_sp(x1v, x2v) : lp_line_line_sp(x1v, x2v);
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double x2 = x2v[c2 - 1];
            double R = sqrt((x1 - x2) * (x1 - x2) + dy * dy + dz * dz);
            double t1 = (x1 - x2) * jl_clog_re((x1 - x2) + R);
            if (ISBAD(t1)) t1 = 0.0;
            sol = sol + sgnpow(c1 + c2 + 1) * (t1 - R);
        }
    }
    return sol;
}

/* :531-551 / compute_row_P_sup.jl:478-498 */
static double line_line_ortho_xy(const double *x1v, double y1, double z1,
                                 double x2, const double *y2v, double z2)
{
    g_form = 8;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y2 = y2v[c2 - 1];
            double R = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
            double t1 = (x1 - x2) * jl_clog_re((y1 - y2) + R);
            if (ISBAD(t1)) t1 = 0.0;
            double t2 = (y1 - y2) * jl_clog_re((x1 - x2) + R);
            if (ISBAD(t2)) t2 = 0.0;
            double t3 = -fabs(z1 - z2) * or_det_atan2((x1 - x2) * (y1 - y2), (fabs(z1 - z2) * R));
            sol = sol + sgnpow(c1 + c2 + 1) * (t1 + t2 + t3);
        }
    }
    return sol;
}

/* :553-582 */
static double lp_point_vol(double x1, double y1, double z1,
                           const double *x2v, const double *y2v, const double *z2v)
{
    g_form = 7;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x2 = x2v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y2 = y2v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z2 = z2v[c3 - 1];
                double R = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
                double t1 = (y1 - y2) * (z1 - z2) * jl_clog_re((x1 - x2) + R);
                if (ISBAD(t1)) t1 = 0.0;
                double t2 = (x1 - x2) * (z1 - z2) * jl_clog_re((y1 - y2) + R);
                if (ISBAD(t2)) t2 = 0.0;
                double t3 = (y1 - y2) * (x1 - x2) * jl_clog_re((z1 - z2) + R);
                if (ISBAD(t3)) t3 = 0.0;
                double t4 = -1.0 / 2 * fabs(z1 - z2) * (z1 - z2) * or_det_atan2((x1 - x2) * (y1 - y2), (fabs(z1 - z2) * R));
                double t5 = -1.0 / 2 * fabs(y1 - y2) * (y1 - y2) * or_det_atan2((x1 - x2) * (z1 - z2), (fabs(y1 - y2) * R));
                double t6 = -1.0 / 2 * fabs(x1 - x2) * (x1 - x2) * or_det_atan2((y1 - y2) * (z1 - z2), (fabs(x1 - x2) * R));
                sol = sol + sgnpow(c1 + c2 + c3 + 1) * (t1 + t2 + t3 + t4 + t5 + t6);
            }
        }
    }
    return sol;
}

/* :584-607 (P copy :500-529 has the same arithmetic except the sign
 * exponent; both are (-1)^(c1+c2+c3)) */
static double line_surf_ortho(const double *x1v, double y1, double z1,
                              double x2, const double *y2v, const double *z2v)
{
    g_form = 6;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y2 = y2v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z2 = z2v[c3 - 1];
                double R = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
                double t1 = (y1 - y2) * (z1 - z2) * jl_clog_re((x1 - x2) + R);
                if (ISBAD(t1)) t1 = 0.0;
                double t2 = (x1 - x2) * (z1 - z2) * jl_clog_re((y1 - y2) + R);
                if (ISBAD(t2)) t2 = 0.0;
                double t3 = -fabs(z1 - z2) * or_det_atan2((x1 - x2) * (y1 - y2), (fabs(z1 - z2) * R));
                sol = sol + sgnpow(c1 + c2 + c3) * (t1 + t2 + t3);
            }
        }
    }
    return sol;
}

/* compute_row_P_sup.jl:500-529: the P copy of line_surf_ortho carries three
 * log terms and three atan terms (a different formula from the Lp copy). */
static double p_line_surf_ortho(const double *x1v, double y1, double z1,
                                double x2, const double *y2v, const double *z2v)
{
    g_form = 14;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y2 = y2v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z2 = z2v[c3 - 1];
                double R = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
                double t1 = (y1 - y2) * (z1 - z2) * jl_clog_re((x1 - x2) + R);
                if (ISBAD(t1)) t1 = 0.0;
                double t2 = (x1 - x2) * (z1 - z2) * jl_clog_re((y1 - y2) + R);
                if (ISBAD(t2)) t2 = 0.0;
                double t3 = (y1 - y2) * (x1 - x2) * jl_clog_re((z1 - z2) + R);
                if (ISBAD(t3)) t3 = 0.0;
                double t4 = -1.0 / 2 * fabs(z1 - z2) * (z1 - z2) * or_det_atan2((x1 - x2) * (y1 - y2), (fabs(z1 - z2) * R));
                double t5 = -1.0 / 2 * fabs(y1 - y2) * (y1 - y2) * or_det_atan2((x1 - x2) * (z1 - z2), (fabs(y1 - y2) * R));
                double t6 = -1.0 / 2 * fabs(x1 - x2) * (x1 - x2) * or_det_atan2((y1 - y2) * (z1 - z2), (fabs(x1 - x2) * R));
                sol = sol + sgnpow(c1 + c2 + c3) * (t1 + t2 + t3 + t4 + t5 + t6);
            }
        }
    }
    return sol;
}

/* :609-633 / compute_row_P_sup.jl:531-555 */
static double line_surf_para(const double *x1v, const double *y1v, double z1,
                             const double *x2v, double y2, double z2)
{
    g_form = 5;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double x2 = x2v[c3 - 1];
                double R = sqrt((x1 - x2) * (x1 - x2) + (y1 - y2) * (y1 - y2) + (z1 - z2) * (z1 - z2));
                double t1 = (x1 - x2) * (y1 - y2) * jl_clog_re((x1 - x2) + R);
                if (ISBAD(t1)) t1 = 0.0;
                double t2 = ((x1 - x2) * (x1 - x2) - (z1 - z2) * (z1 - z2)) / 2 * jl_clog_re((y1 - y2) + R);
                if (ISBAD(t2)) t2 = 0.0;
                double t3 = -(x1 - x2) * fabs(z1 - z2) * or_det_atan2((x1 - x2) * (y1 - y2), (fabs(z1 - z2) * R));
                double t4 = -(y1 - y2) / 2 * R;
                sol = sol + sgnpow(c1 + c2 + c3 + 1) * (t1 + t2 + t3 + t4);
            }
        }
    }
    return sol;
}

/* :635-668 */
static double lp_line_vol(const double *x1v, const double *y1v, const double *z1v,
                          const double *x2v, double y2, double z2)
{
    g_form = 4;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z1 = z1v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    double t1 = -1.0 / 3 * dy * dz * R;
                    double t2 = -1.0 / 2 * dx * fabs(dz) * dz * or_det_atan2(dx * dy, (fabs(dz) * R));
                    double t3 = -1.0 / 2 * dx * fabs(dy) * dy * or_det_atan2(dx * dz, (fabs(dy) * R));
                    double t4 = -1.0 / 6 * (fabs(dx) * fabs(dx) * fabs(dx)) * or_det_atan2(dy * dz, (fabs(dx) * R));
                    double t5 = dx * dy * dz * jl_clog_re(dx + R);
                    if (ISBAD(t5)) t5 = 0.0;
                    double t6 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dz * dz)) * dz * jl_clog_re(dy + R);
                    if (ISBAD(t6)) t6 = 0.0;
                    double t7 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dy * dy)) * dy * jl_clog_re(dz + R);
                    if (ISBAD(t7)) t7 = 0.0;
                    sol = sol + sgnpow(c1 + c2 + c3 + c4 + 1) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
                }
            }
        }
    }
    return sol;
}

/* :670-697 (Lp copy; complex log) */
static double lp_surf_surf_para(const double *x1v, const double *y1v, double z1,
                                const double *x2v, const double *y2v, double z2)
{
    g_form = 2;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double y2 = y2v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    double t1 = -1.0 / 6 * (dx * dx + dy * dy - 2 * (dz * dz)) * R;
                    double t2 = -dx * dy * fabs(dz) * or_det_atan2(dx * dy, (fabs(dz) * R));
                    double t3 = 1.0 / 2 * (dx * dx - dz * dz) * dy * jl_clog_re(dy + R);
                    if (ISBAD(t3)) t3 = 0.0;
                    double t4 = 1.0 / 2 * (dy * dy - dz * dz) * dx * jl_clog_re(dx + R);
                    if (ISBAD(t4)) t4 = 0.0;
                    sol = sol + sgnpow(c1 + c2 + c3 + c4) * (t1 + t2 + t3 + t4);
                }
            }
        }
    }
    return sol;
}

/* compute_Lp_Voxels_QS.jl:695-730: real log of a real argument (>= 0). */
static double qs_surf_surf_para(const double *x1v, const double *y1v, double z1,
                                const double *x2v, const double *y2v, double z2)
{
    g_form = 12;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double y2 = y2v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    double t1 = -1.0 / 6 * (dx * dx + dy * dy - 2 * (dz * dz)) * R;
                    double t2 = -dx * dy * fabs(dz) * or_det_atan2(dx * dy, fabs(dz) * R);
                    double t3 = 1.0 / 2 * (dx * dx - dz * dz) * dy * or_det_log(dy + R);
                    if (ISBAD(t3)) t3 = 0.0;
                    double t4 = 1.0 / 2 * (dy * dy - dz * dz) * dx * or_det_log(dx + R);
                    if (ISBAD(t4)) t4 = 0.0;
                    sol += sgnpow(c1 + c2 + c3 + c4) * (t1 + t2 + t3 + t4);
                }
            }
        }
    }
    return sol;
}

/* :699-726 (Lp copy) */
static double lp_surf_surf_ortho(const double *x1v, const double *y1v, double z1,
                                 const double *x2v, double y2, const double *z2v)
{
    g_form = 3;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z2 = z2v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    double t1 = -1.0 / 3 * dy * dz * R;
                    double t2 = -1.0 / 2 * dx * fabs(dz) * dz * or_det_atan2(dx * dy, (fabs(dz) * R));
                    double t3 = 1.0 / 2 * (dx * dx - dz * dz) * dy * jl_clog_re(dy + R);
                    if (ISBAD(t3)) t3 = 0.0;
                    double t4 = 1.0 / 2 * (dx * dx - dy * dy) * dx * jl_clog_re(dx + R);
                    if (ISBAD(t4)) t4 = 0.0;
                    sol = sol + sgnpow(c1 + c2 + c3 + c4) * (t1 + t2 + t3 + t4);
                }
            }
        }
    }
    return sol;
}

/* compute_row_P_sup.jl:557-590 */
static double p_surf_surf_ortho(const double *x1v, const double *y1v, double z1,
                                const double *x2v, double y2, const double *z2v)
{
    g_form = 13;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z2 = z2v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                    double R = sqrt(dx * dx + dy * dy + dz * dz);
                    double t1 = -1.0 / 3 * dy * dz * R;
                    double t2 = -1.0 / 2 * dx * fabs(dz) * dz * or_det_atan2(dx * dy, fabs(dz) * R);
                    double t3 = -1.0 / 2 * dx * fabs(dy) * dy * or_det_atan2(dx * dz, fabs(dy) * R);
                    double t4 = -1.0 / 6 * (fabs(dx) * fabs(dx) * fabs(dx)) * or_det_atan2(dy * dz, fabs(dx) * R);
                    double t5 = dx * dy * dz * jl_clog_re(dx + R);
                    if (ISBAD(t5)) t5 = 0.0;
                    double t6 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dz * dz)) * dz * jl_clog_re(dy + R);
                    if (ISBAD(t6)) t6 = 0.0;
                    double t7 = (1.0 / 2 * (dx * dx) - 1.0 / 6 * (dy * dy)) * dy * jl_clog_re(dz + R);
                    if (ISBAD(t7)) t7 = 0.0;
                    sol = sol + sgnpow(c1 + c2 + c3 + c4) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
                }
            }
        }
    }
    return sol;
}

/* :728-764 */
static double lp_vol_surf(const double *x1v, const double *y1v, const double *z1v,
                          const double *x2v, const double *y2v, double z2)
{
    g_form = 1;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z1 = z1v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    for (int c5 = 1; c5 <= 2; c5++) {
                        double y2 = y2v[c5 - 1];
                        double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                        double R = sqrt(dx * dx + dy * dy + dz * dz);
                        double t1 = ((dz * dz) / 12 - ((dx * dx) + (dy * dy)) / 8) * dz * R;
                        double t2 = ((dy * dy) / 2 - (dz * dz) / 6) * dx * jl_clog_re(dx + R);
                        if (ISBAD(t2)) t2 = 0.0;
                        double t3 = ((dx * dx) / 2 - (dz * dz) / 6) * dy * jl_clog_re(dy + R);
                        if (ISBAD(t3)) t3 = 0.0;
                        double t4 = (-jl_pow4(dx) / 24 - jl_pow4(dy) / 24 + ((dy * dy) * (dx * dx)) / 4) * jl_clog_re(dz + R);
                        if (ISBAD(t4)) t4 = 0.0;
                        double t5 = -(fabs(dx) * fabs(dx) * fabs(dx)) * dy / 6 * or_det_atan2(dy * dz, (fabs(dx) * R));
                        double t6 = -dx * (fabs(dy) * fabs(dy) * fabs(dy)) / 6 * or_det_atan2(dx * dz, (fabs(dy) * R));
                        double t7 = -dx * dy * fabs(dz) * dz / 2 * or_det_atan2(dx * dy, (fabs(dz) * R));
                        sol = sol + sgnpow(c1 + c2 + c3 + c4 + c5) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
                    }
                }
            }
        }
    }
    return sol;
}

/* :766-805 -- the 64-corner exact vol-vol integral */
static double lp_vol_vol(const double *x1v, const double *y1v, const double *z1v,
                         const double *x2v, const double *y2v, const double *z2v)
{
    g_form = 0;
    double sol = 0.0;
    for (int c1 = 1; c1 <= 2; c1++) {
        double x1 = x1v[c1 - 1];
        for (int c2 = 1; c2 <= 2; c2++) {
            double y1 = y1v[c2 - 1];
            for (int c3 = 1; c3 <= 2; c3++) {
                double z1 = z1v[c3 - 1];
                for (int c4 = 1; c4 <= 2; c4++) {
                    double x2 = x2v[c4 - 1];
                    for (int c5 = 1; c5 <= 2; c5++) {
                        double y2 = y2v[c5 - 1];
                        for (int c6 = 1; c6 <= 2; c6++) {
                            double z2 = z2v[c6 - 1];
                            double dx = x1 - x2, dy = y1 - y2, dz = z1 - z2;
                            double R = sqrt(dx * dx + dy * dy + dz * dz);
                            double t1 = (jl_pow4(dx) + jl_pow4(dy) + jl_pow4(dz) - 3 * (dx * dx) * (dy * dy)
                                         - 3 * (dy * dy) * (dz * dz) - 3 * (dx * dx) * (dz * dz)) * R / 60;
                            double t2 = ((dy * dy) * (dz * dz) / 4 - jl_pow4(dy) / 24 - jl_pow4(dz) / 24) * dx * jl_clog_re(dx + R);
                            if (ISBAD(t2)) t2 = 0.0;
                            double t3 = ((dy * dy) * (dx * dx) / 4 - jl_pow4(dy) / 24 - jl_pow4(dx) / 24) * dz * jl_clog_re(dz + R);
                            if (ISBAD(t3)) t3 = 0.0;
                            double t4 = ((dz * dz) * (dx * dx) / 4 - jl_pow4(dz) / 24 - jl_pow4(dx) / 24) * dy * jl_clog_re(dy + R);
                            if (ISBAD(t4)) t4 = 0.0;
                            double t5 = -(fabs(dx) * fabs(dx) * fabs(dx)) * dy * dz / 6 * or_det_atan2(dy * dz, (fabs(dx) * R));
                            double t6 = -dx * (fabs(dy) * fabs(dy) * fabs(dy)) * dz / 6 * or_det_atan2(dx * dz, (fabs(dy) * R));
                            double t7 = -dx * dy * (fabs(dz) * fabs(dz) * fabs(dz)) / 6 * or_det_atan2(dx * dy, (fabs(dz) * R));
                            sol = sol + sgnpow(c1 + c2 + c3 + c4 + c5 + c6 + 1) * (t1 + t2 + t3 + t4 + t5 + t6 + t7);
                        }
                    }
                }
            }
        }
    }
    return sol;
}

/* ---- Lp classifier: compute_Lp_Voxels.jl:65-447 ----------------------- */

static int lp_check(double e1, double e2, double e3, double V1, double V2, double max_d,
                    double min_R, double sd, double o1, double o2)
{
    double max_oth = fmax(o1, o2);
    double m = min_R + 1e-15;
    double condX1a = V1 * V2 * max_d / (m * m * m);
    double condX1f = sd / m;
    double condX1b = sd / max_oth;
    if (near_thr(condX1b, e3) || near_thr(condX1f, e1) || near_thr(condX1a, e2)) g_tie = 1;
    return ((condX1b <= e3 || condX1f < e1) && condX1a < e2) ? 1 : 0;
}

/* the second-stage test of lp_supp_box (:102, :115, :128) */
static int lp_e4(double other, double min_R, double e4)
{
    double v = other / (min_R + 1e-15);
    if (near_thr(v, e4)) g_tie = 1;
    return v < e4;
}

static double max4(const double *a) { double m = a[0]; for (int i = 1; i < 4; i++) if (a[i] > m) m = a[i]; return m; }
static double min4(const double *a) { double m = a[0]; for (int i = 1; i < 4; i++) if (a[i] < m) m = a[i]; return m; }

/* dimension-suppression pass for one box (the three if/elseif blocks of
 * :97-170 share this shape) */
static void lp_supp_box(double a, double b, double c, const double *ax, const double *ay, const double *az,
                        double V1, double V2, double min_R, int *sx, int *sy, int *sz)
{
    const double e1 = 5e-3, e2 = 1e-3, e3 = 1e-3, e4 = 3e-1;
    if (a <= b && a <= c) {
        *sx = lp_check(e1, e2, e3, V1, V2, max4(ax), min_R, a, b, c);
        if (*sx == 1 && lp_e4(fmax(b, c), min_R, e4)) {
            *sy = lp_check(e1, e2, e3, V1, V2, max4(ay), min_R, b, a, c);
            *sz = lp_check(e1, e2, e3, V1, V2, max4(az), min_R, c, a, b);
        }
    } else if (b <= a && b <= c) {
        *sy = lp_check(e1, e2, e3, V1, V2, max4(ay), min_R, b, a, c);
        if (*sy == 1 && lp_e4(fmax(a, c), min_R, e4)) {
            *sx = lp_check(e1, e2, e3, V1, V2, max4(ax), min_R, a, b, c);
            *sz = lp_check(e1, e2, e3, V1, V2, max4(az), min_R, c, a, b);
        }
    } else {
        *sz = lp_check(e1, e2, e3, V1, V2, max4(az), min_R, c, a, b);
        if (*sz == 1 && lp_e4(fmax(a, b), min_R, e4)) {
            *sx = lp_check(e1, e2, e3, V1, V2, max4(ax), min_R, a, b, c);
            *sy = lp_check(e1, e2, e3, V1, V2, max4(ay), min_R, b, a, c);
        }
    }
}

/* The Lp dispatcher's forms: compute_Lp_Voxels.jl's, or while g_qs is set
 * compute_Lp_Voxels_QS.jl's (whose Integ_vol_vol, :44-423, has the same
 * branch structure). */
static double d_line_point(const double *x1v, double y3, double z3, double x1, double y1, double z1)
{
    return g_qs ? qs_line_point(x1v, y3, z3, x1, y1, z1) : lp_line_point(x1v, y3, z3, x1, y1, z1);
}
static double d_line_surf_ortho(const double *x1v, double y1, double z1, double x2, const double *y2v, const double *z2v)
{
    return g_qs ? qs_line_surf_ortho(x1v, y1, z1, x2, y2v, z2v) : line_surf_ortho(x1v, y1, z1, x2, y2v, z2v);
}
static double d_line_surf_para(const double *x1v, const double *y1v, double z1, const double *x2v, double y2, double z2)
{
    return g_qs ? qs_line_surf_para(x1v, y1v, z1, x2v, y2, z2) : line_surf_para(x1v, y1v, z1, x2v, y2, z2);
}
static double d_line_vol(const double *x1v, const double *y1v, const double *z1v, const double *x2v, double y2,
                         double z2)
{
    return g_qs ? qs_line_vol(x1v, y1v, z1v, x2v, y2, z2) : lp_line_vol(x1v, y1v, z1v, x2v, y2, z2);
}
static double d_surf_surf_para(const double *x1v, const double *y1v, double z1, const double *x2v,
                               const double *y2v, double z2)
{
    return g_qs ? qs_surf_surf_para(x1v, y1v, z1, x2v, y2v, z2) : lp_surf_surf_para(x1v, y1v, z1, x2v, y2v, z2);
}
/* QS :345-361 pass an interval as the scalar argument y2 and a scalar as
 * the interval z2v of integ_surf_surf_ortho (:732): Julia raises a
 * MethodError there, restated as NaN (form 25). */
static double d_surf_surf_ortho(const double *x1v, const double *y1v, double z1, const double *x2v, double y2,
                                const double *z2v)
{
    if (g_qs) {
        g_form = 25;
        return NAN;
    }
    return lp_surf_surf_ortho(x1v, y1v, z1, x2v, y2, z2v);
}
static double d_vol_surf(const double *x1v, const double *y1v, const double *z1v, const double *x2v,
                         const double *y2v, double z2)
{
    return g_qs ? qs_vol_surf(x1v, y1v, z1v, x2v, y2v, z2) : lp_vol_surf(x1v, y1v, z1v, x2v, y2v, z2);
}

double oracle_integ_vol_vol(const double *x1v, const double *y1v, const double *z1v,
                            const double *x2v, const double *y2v, const double *z2v)
{
    double xc1 = 0.5 * (x1v[0] + x1v[1]), yc1 = 0.5 * (y1v[0] + y1v[1]), zc1 = 0.5 * (z1v[0] + z1v[1]);
    double xc2 = 0.5 * (x2v[0] + x2v[1]), yc2 = 0.5 * (y2v[0] + y2v[1]), zc2 = 0.5 * (z2v[0] + z2v[1]);
    double a1 = fabs(x1v[1] - x1v[0]), b1 = fabs(y1v[1] - y1v[0]), c1 = fabs(z1v[1] - z1v[0]);
    double a2 = fabs(x2v[1] - x2v[0]), b2 = fabs(y2v[1] - y2v[0]), c2 = fabs(z2v[1] - z2v[0]);
    double V1 = a1 * b1 * c1, V2 = a2 * b2 * c2;
    int sx1 = 0, sy1 = 0, sz1 = 0, sx2 = 0, sy2 = 0, sz2 = 0;
    double ax[4] = {fabs(x1v[0] - x2v[0]), fabs(x1v[0] - x2v[1]), fabs(x1v[1] - x2v[0]), fabs(x1v[1] - x2v[1])};
    double ay[4] = {fabs(y1v[0] - y2v[0]), fabs(y1v[0] - y2v[1]), fabs(y1v[1] - y2v[0]), fabs(y1v[1] - y2v[1])};
    double az[4] = {fabs(z1v[0] - z2v[0]), fabs(z1v[0] - z2v[1]), fabs(z1v[1] - z2v[0]), fabs(z1v[1] - z2v[1])};
    double mx = min4(ax), my = min4(ay), mz = min4(az);
    double min_R = sqrt(mx * mx + my * my + mz * mz);
    lp_supp_box(a1, b1, c1, ax, ay, az, V1, V2, min_R, &sx1, &sy1, &sz1);
    lp_supp_box(a2, b2, c2, ax, ay, az, V1, V2, min_R, &sx2, &sy2, &sz2);
    int sum_supp = sx1 + sy1 + sz1 + sx2 + sy2 + sz2;
    double integ = 0.0;
    double X1[2] = {x1v[0], x1v[1]}, Y1[2] = {y1v[0], y1v[1]}, Z1[2] = {z1v[0], z1v[1]};
    double X2[2] = {x2v[0], x2v[1]}, Y2[2] = {y2v[0], y2v[1]}, Z2[2] = {z2v[0], z2v[1]};

    if (sum_supp == 6) {
        g_form = 16;
        integ = V1 * V2 / sqrt((xc1 - xc2) * (xc1 - xc2) + (yc1 - yc2) * (yc1 - yc2) + (zc1 - zc2) * (zc1 - zc2));
    } else if (sum_supp == 5) {
        int is_point_v1 = (sx1 + sy1 + sz1 == 3);
        if (is_point_v1) {
            if (sx2 == 0) integ = V1 * V2 / a2 * d_line_point(X2, yc2, zc2, xc1, yc1, zc1);
            else if (sy2 == 0) integ = V1 * V2 / b2 * d_line_point(Y2, xc2, zc2, yc1, xc1, zc1);
            else integ = V1 * V2 / c2 * d_line_point(Z2, xc2, yc2, zc1, xc1, yc1);
        } else {
            if (sx1 == 0) integ = V1 * V2 / a1 * d_line_point(X1, yc1, zc1, xc2, yc2, zc2);
            else if (sy1 == 0) integ = V1 * V2 / b1 * d_line_point(Y1, xc1, zc1, yc2, xc2, zc2);
            else integ = V1 * V2 / c1 * d_line_point(Z1, xc1, yc1, zc2, xc2, yc2);
        }
    } else if (sum_supp == 4) {
        int is_point_v1 = (sx1 + sy1 + sz1 == 3);
        int is_point_v2 = (sx2 + sy2 + sz2 == 3);
        if (is_point_v1) {
            if (sx2 == 1) integ = V1 * a2 * point_sup(zc1, yc1, xc1, Z2, Y2, xc2);
            else if (sy2 == 1) integ = V1 * b2 * point_sup(xc1, zc1, yc1, X2, Z2, yc2);
            else integ = V1 * c2 * point_sup(xc1, yc1, zc1, X2, Y2, zc2);
        } else if (is_point_v2) {
            if (sx1 == 1) integ = V2 * a1 * point_sup(zc2, yc2, xc2, Z1, Y1, xc1);
            else if (sy1 == 1) integ = V2 * b1 * point_sup(xc2, zc2, yc2, X1, Z1, yc1);
            else integ = V2 * c1 * point_sup(xc2, yc2, zc2, X1, Y1, zc1);
        } else {
            if (sy1 == 1 && sz1 == 1) {
                if (sy2 == 1 && sz2 == 1) integ = b1 * c1 * b2 * c2 * line_line_parall(X1, yc1, zc1, X2, yc2, zc2, 0);
                else if (sx2 == 1 && sz2 == 1) integ = b1 * c1 * a2 * c2 * line_line_ortho_xy(X1, yc1, zc1, xc2, Y2, zc2);
                else integ = b1 * c1 * a2 * b2 * line_line_ortho_xy(X1, zc1, yc1, xc2, Z2, yc2);
            } else if (sx1 == 1 && sz1 == 1) {
                if (sx2 == 1 && sz2 == 1) integ = a1 * c1 * a2 * c2 * line_line_parall(Y1, xc1, zc1, Y2, xc2, zc2, 0);
                else if (sx2 == 1 && sy2 == 1) integ = a1 * c1 * a2 * b2 * line_line_ortho_xy(Y1, zc1, xc1, yc2, Z2, xc2);
                else integ = a1 * c1 * b2 * c2 * line_line_ortho_xy(Y1, xc1, zc1, yc2, X2, zc2);
            } else {
                if (sx2 == 1 && sy2 == 1) integ = a1 * b1 * a2 * b2 * line_line_parall(Z1, xc1, yc1, Z2, xc2, yc2, 0);
                else if (sx2 == 1 && sz2 == 1) integ = a1 * b1 * a2 * c2 * line_line_ortho_xy(Z1, yc1, xc1, zc2, Y2, xc2);
                else integ = a1 * b1 * b2 * c2 * line_line_ortho_xy(Z1, xc1, yc1, zc2, X2, yc2);
            }
        }
    } else if (sum_supp == 3) {
        int is_point_v1 = (sx1 + sy1 + sz1 == 3);
        int is_point_v2 = (sx2 + sy2 + sz2 == 3);
        int is_surf_v1 = (sx1 + sy1 + sz1 == 1);
        if (is_point_v1) {
            integ = a1 * b1 * c1 * lp_point_vol(xc1, yc1, zc1, X2, Y2, Z2);
        } else if (is_point_v2) {
            integ = a2 * b2 * c2 * lp_point_vol(xc2, yc2, zc2, X1, Y1, Z1);
        } else if (is_surf_v1) {
            if (sx1 == 1) {
                if (sx2 == 0) integ = a1 * b2 * c2 * d_line_surf_ortho(X2, yc2, zc2, xc1, Y1, Z1);
                else if (sy2 == 0) integ = a1 * a2 * c2 * d_line_surf_para(Y1, Z1, xc1, Y2, zc2, xc2);
                else integ = a1 * a2 * b2 * d_line_surf_para(Z1, Y1, xc1, Z2, yc2, xc2);
            } else if (sy1 == 1) {
                if (sx2 == 0) integ = b1 * b2 * c2 * d_line_surf_para(X1, Z1, yc1, X2, zc2, yc2);
                else if (sy2 == 0) integ = b1 * a2 * c2 * d_line_surf_ortho(Y2, xc2, zc2, yc1, X1, Z1);
                else integ = b1 * a2 * b2 * d_line_surf_para(Z1, X1, yc1, Z2, xc2, yc2);
            } else {
                if (sx2 == 0) integ = c1 * b2 * c2 * d_line_surf_para(X1, Y1, zc1, X2, yc2, zc2);
                else if (sy2 == 0) integ = c1 * a2 * c2 * d_line_surf_para(Y1, X1, zc1, Y2, xc2, zc2);
                else integ = c1 * a2 * b2 * d_line_surf_ortho(Z2, xc2, yc2, zc1, X1, Y1);
            }
        } else {
            if (sx2 == 1) {
                if (sx1 == 0) integ = a2 * b1 * c1 * d_line_surf_ortho(X1, yc1, zc1, xc2, Y2, Z2);
                else if (sy1 == 0) integ = a2 * a1 * c1 * d_line_surf_para(Y2, Z2, xc2, Y1, zc1, xc1);
                else integ = a2 * a1 * b1 * d_line_surf_para(Z2, Y2, xc2, Z1, yc1, xc1);
            } else if (sy2 == 1) {
                if (sx1 == 0) integ = b2 * b1 * c1 * d_line_surf_para(X2, Z2, yc2, X1, zc1, yc1);
                else if (sy1 == 0) integ = b2 * a1 * c1 * d_line_surf_ortho(Y1, xc1, zc1, yc2, X2, Z2);
                else integ = b2 * a1 * b1 * d_line_surf_para(Z2, X2, yc2, Z1, xc1, yc1);
            } else {
                if (sx1 == 0) integ = c2 * b1 * c1 * d_line_surf_para(X2, Y2, zc2, X1, yc1, zc1);
                else if (sy1 == 0) integ = c2 * a1 * c1 * d_line_surf_para(Y2, X2, zc2, Y1, xc1, zc1);
                else integ = c2 * a1 * b1 * d_line_surf_ortho(Z1, xc1, yc1, zc2, X2, Y2);
            }
        }
    } else if (sum_supp == 2) {
        int is_line_v1 = (sx1 + sy1 + sz1 == 2);
        int is_line_v2 = (sx2 + sy2 + sz2 == 2);
        if (is_line_v1) {
            if (sx1 == 0) integ = b1 * c1 * d_line_vol(X2, Y2, Z2, X1, yc1, zc1);
            else if (sy1 == 0) integ = a1 * c1 * d_line_vol(Y2, X2, Z2, Y1, xc1, zc1);
            else integ = a1 * b1 * d_line_vol(Z2, X2, Y2, Z1, xc1, yc1);
        } else if (is_line_v2) {
            if (sx2 == 0) integ = b2 * c2 * d_line_vol(X1, Y1, Z1, X2, yc2, zc2);
            else if (sy2 == 0) integ = a2 * c2 * d_line_vol(Y1, X1, Z1, Y2, xc2, zc2);
            else integ = a2 * b2 * d_line_vol(Z1, X1, Y1, Z2, xc2, yc2);
        } else {
            if (sx1 == 1) {
                if (sx2 == 1) integ = a1 * a2 * d_surf_surf_para(Y1, Z1, xc1, Y2, Z2, xc2);
                else if (sy2 == 1) integ = a1 * b2 * d_surf_surf_ortho(Z1, Y1, xc1, Z2, yc2, X2);
                else integ = a1 * c2 * d_surf_surf_ortho(Y1, Z1, xc1, Y2, zc2, X2);
            } else if (sy1 == 1) {
                if (sx2 == 1) integ = b1 * a2 * d_surf_surf_ortho(Z1, X1, yc1, Z2, xc2, Y2);
                else if (sy2 == 1) integ = b1 * b2 * d_surf_surf_para(X1, Z1, yc1, X2, Z2, yc2);
                else integ = b1 * c2 * d_surf_surf_ortho(X1, Z1, yc1, X2, zc2, Y2);
            } else {
                if (sx2 == 1) integ = c1 * a2 * d_surf_surf_ortho(Y1, X1, zc1, Y2, xc2, Z2);
                else if (sy2 == 1) integ = c1 * b2 * d_surf_surf_ortho(X1, Y1, zc1, X2, yc2, Z2);
                else integ = c1 * c2 * d_surf_surf_para(X1, Y1, zc1, X2, Y2, zc2);
            }
        }
    } else if (sum_supp == 1) {
        if (sx1 == 1) integ = a1 * d_vol_surf(Y2, Z2, X2, Y1, Z1, xc1);
        else if (sy1 == 1) integ = b1 * d_vol_surf(X2, Z2, Y2, X1, Z1, yc1);
        else if (sz1 == 1) integ = c1 * d_vol_surf(X2, Y2, Z2, X1, Y1, zc1);
        else if (sx2 == 1) integ = a2 * d_vol_surf(Y1, Z1, X1, Y2, Z2, xc2);
        else if (sy2 == 1) integ = b2 * d_vol_surf(X1, Z1, Y1, X2, Z2, yc2);
        else if (sz2 == 1) integ = c2 * d_vol_surf(X1, Y1, Z1, X2, Y2, zc2);
    } else {
        integ = lp_vol_vol(X1, Y1, Z1, X2, Y2, Z2);
    }
    return integ;
}

/* compute_Lp_Voxels.jl:1-63 with M = 1 and is_sym = false (the hot-path
 * call of compute_FFT_mutual_coupling_mats.jl:46-52).  centers: N x 3
 * column-major (Julia layout).  out[n] = 1e-7/S * Integ_vol_vol_fft. */
void oracle_lp_row(const double *src, const double *centers, int64_t N,
                   double sx, double sy, double sz, double sx2, double sy2, double sz2,
                   int dc, double *out)
{
    double S = sy * sz * sy2 * sz2;
    if (dc == 2) S = sx * sz * sx2 * sz2;
    else if (dc == 3) S = sx * sy * sx2 * sy2;
    double x1v[2] = {jl_round14(src[0] - sx / 2), jl_round14(src[0] + sx / 2)};
    double y1v[2] = {jl_round14(src[1] - sy / 2), jl_round14(src[1] + sy / 2)};
    double z1v[2] = {jl_round14(src[2] - sz / 2), jl_round14(src[2] + sz / 2)};
    for (int64_t n = 0; n < N; n++) {
        double cx = centers[n], cy = centers[n + N], cz = centers[n + 2 * N];
        double x2v[2] = {jl_round14(cx - sx2 / 2), jl_round14(cx + sx2 / 2)};
        double y2v[2] = {jl_round14(cy - sy2 / 2), jl_round14(cy + sy2 / 2)};
        double z2v[2] = {jl_round14(cz - sz2 / 2), jl_round14(cz + sz2 / 2)};
        out[n] = 1e-7 / S * oracle_integ_vol_vol(x1v, y1v, z1v, x2v, y2v, z2v);
    }
}

/* compute_Lp_Voxels_QS.jl:9-42 (the row of compute_rows_Rcc_Lp.jl:28-32):
 * the same corner rounding (round_ud(., 14), :3-6) and prefactor; entries
 * without suppression evaluate compute_Lp_Voxels.jl's dispatcher (the QS
 * dispatcher calls Integ_vol_vol_fft there, :418), the others the QS forms. */
void oracle_lp_row_qs(const double *src, const double *centers, int64_t N,
                      double sx, double sy, double sz, double sx2, double sy2, double sz2,
                      int dc, double *out)
{
    g_qs = 1;
    oracle_lp_row(src, centers, N, sx, sy, sz, sx2, sy2, sz2, dc, out);
    g_qs = 0;
}

/* ---- P classifier: compute_row_P_sup.jl:106-394 ----------------------- */

static int p_check(double e1, double e2, double e3, double e4, double S1, double S2,
                   double max_d, double min_R, double sd, double o1, double o2)
{
    double max_oth = o1 * o2;
    double m = min_R + 1e-15;
    double condS = fmax(max_oth, sd) / m;
    double condX1a = S1 * S2 * max_d / (m * m * m);
    double condX1f = sd / m;
    double condX1b = sd / max_oth;
    if (near_thr(condS, e4) || near_thr(condX1b, e3) || near_thr(condX1f, e1) || near_thr(condX1a, e2)) g_tie = 1;
    return (condS < e4 && (condX1b <= e3 || condX1f < e1) && condX1a < e2) ? 1 : 0;
}

double oracle_integ_sup_sup_P(double xc1, double yc1, double zc1, double xc2, double yc2, double zc2,
                              double a1, double b1, double c1, double a2, double b2, double c2)
{
    const double e1 = 5e-3, e2 = 1e-3, e3 = 1e-3, e4 = 3e-1;
    double X1[2] = {xc1 - a1 / 2, xc1 + a1 / 2}, Y1[2] = {yc1 - b1 / 2, yc1 + b1 / 2}, Z1[2] = {zc1 - c1 / 2, zc1 + c1 / 2};
    double X2[2] = {xc2 - a2 / 2, xc2 + a2 / 2}, Y2[2] = {yc2 - b2 / 2, yc2 + b2 / 2}, Z2[2] = {zc2 - c2 / 2, zc2 + c2 / 2};
    int p1yz = 0, p1xz = 0, p2yz = 0, p2xz = 0;
    if (a1 <= b1 && a1 <= c1) { p1yz = 1; a1 = 1; }
    else if (b1 <= a1 && b1 <= c1) { p1xz = 1; b1 = 1; }
    else c1 = 1;
    if (a2 <= b2 && a2 <= c2) { p2yz = 1; a2 = 1; }
    else if (b2 <= a2 && b2 <= c2) { p2xz = 1; b2 = 1; }
    else c2 = 1;
    double sup1 = a1 * b1 * c1, sup2 = a2 * b2 * c2;
    double ax[4] = {fabs(X1[0] - X2[0]), fabs(X1[0] - X2[1]), fabs(X1[1] - X2[0]), fabs(X1[1] - X2[1])};
    double ay[4] = {fabs(Y1[0] - Y2[0]), fabs(Y1[0] - Y2[1]), fabs(Y1[1] - Y2[0]), fabs(Y1[1] - Y2[1])};
    double az[4] = {fabs(Z1[0] - Z2[0]), fabs(Z1[0] - Z2[1]), fabs(Z1[1] - Z2[0]), fabs(Z1[1] - Z2[1])};
    double mx = min4(ax), my = min4(ay), mz = min4(az);
    double min_R = sqrt(mx * mx + my * my + mz * mz);
    double md = max4(ax);
    int sx1 = p_check(e1, e2, e3, e4, sup1, sup2, md, min_R, a1, b1, c1);
    int sx2 = p_check(e1, e2, e3, e4, sup1, sup2, md, min_R, a2, b2, c2);
    md = max4(ay);
    int sy1 = p_check(e1, e2, e3, e4, sup1, sup2, md, min_R, b1, a1, c1);
    int sy2 = p_check(e1, e2, e3, e4, sup1, sup2, md, min_R, b2, a2, c2);
    md = max4(az);
    int sz1 = p_check(e1, e2, e3, e4, sup1, sup2, md, min_R, c1, a1, b1);
    int sz2 = p_check(e1, e2, e3, e4, sup1, sup2, md, min_R, c2, a2, b2);
    if (p1yz) sx1 = 1; else if (p1xz) sy1 = 1; else sz1 = 1;
    if (p2yz) sx2 = 1; else if (p2xz) sy2 = 1; else sz2 = 1;
    int sum_supp = sx1 + sy1 + sz1 + sx2 + sy2 + sz2;
    double integ = 0.0;

    if (sum_supp == 6) {
        g_form = 16;
        integ = sup1 * sup2 / sqrt((xc1 - xc2) * (xc1 - xc2) + (yc1 - yc2) * (yc1 - yc2) + (zc1 - zc2) * (zc1 - zc2));
    } else if (sum_supp == 5) {
        int is_point_v1 = (sx1 + sy1 + sz1 == 3);
        if (is_point_v1) {
            if (sx2 == 0) integ = sup1 * sup2 / a2 * lp_line_point(X2, yc2, zc2, xc1, yc1, zc1);
            else if (sy2 == 0) integ = sup1 * sup2 / b2 * lp_line_point(Y2, xc2, zc2, yc1, xc1, zc1);
            else integ = sup1 * sup2 / c2 * lp_line_point(Z2, xc2, yc2, zc1, xc1, yc1);
        } else {
            if (sx1 == 0) integ = sup1 * sup2 / a1 * lp_line_point(X1, yc1, zc1, xc2, yc2, zc2);
            else if (sy1 == 0) integ = sup1 * sup2 / b1 * lp_line_point(Y1, xc1, zc1, yc2, xc2, zc2);
            else integ = sup1 * sup2 / c1 * lp_line_point(Z1, xc1, yc1, zc2, xc2, yc2);
        }
    } else if (sum_supp == 4) {
        int is_point_v1 = (sx1 + sy1 + sz1 == 3);
        int is_point_v2 = (sx2 + sy2 + sz2 == 3);
        if (is_point_v1) {
            if (sx2 == 1) integ = sup1 * a2 * point_sup(zc1, yc1, xc1, Z2, Y2, xc2);
            else if (sy2 == 1) integ = sup1 * b2 * point_sup(xc1, zc1, yc1, X2, Z2, yc2);
            else integ = sup1 * c2 * point_sup(xc1, yc1, zc1, X2, Y2, zc2);
        } else if (is_point_v2) {
            if (sx1 == 1) integ = sup2 * a1 * point_sup(zc2, yc2, xc2, Z1, Y1, xc1);
            else if (sy1 == 1) integ = sup2 * b1 * point_sup(xc2, zc2, yc2, X1, Z1, yc1);
            else integ = sup2 * c1 * point_sup(xc2, yc2, zc2, X1, Y1, zc1);
        } else {
            if (sy1 == 1 && sz1 == 1) {
                if (sy2 == 1 && sz2 == 1) integ = b1 * c1 * b2 * c2 * line_line_parall(X1, yc1, zc1, X2, yc2, zc2, 1);
                else if (sx2 == 1 && sz2 == 1) integ = b1 * c1 * a2 * c2 * line_line_ortho_xy(X1, yc1, zc1, xc2, Y2, zc2);
                else integ = b1 * c1 * a2 * b2 * line_line_ortho_xy(X1, zc1, yc1, xc2, Z2, yc2);
            } else if (sx1 == 1 && sz1 == 1) {
                if (sx2 == 1 && sz2 == 1) integ = a1 * c1 * a2 * c2 * line_line_parall(Y1, xc1, zc1, Y2, xc2, zc2, 1);
                else if (sx2 == 1 && sy2 == 1) integ = a1 * c1 * a2 * b2 * line_line_ortho_xy(Y1, zc1, xc1, yc2, Z2, xc2);
                else integ = a1 * c1 * b2 * c2 * line_line_ortho_xy(Y1, xc1, zc1, yc2, X2, zc2);
            } else {
                if (sx2 == 1 && sy2 == 1) integ = a1 * b1 * a2 * b2 * line_line_parall(Z1, xc1, yc1, Z2, xc2, yc2, 1);
                else if (sx2 == 1 && sz2 == 1) integ = a1 * b1 * a2 * c2 * line_line_ortho_xy(Z1, yc1, xc1, zc2, Y2, xc2);
                else integ = a1 * b1 * b2 * c2 * line_line_ortho_xy(Z1, xc1, yc1, zc2, X2, yc2);
            }
        }
    } else if (sum_supp == 3) {
        int is_surf_v1 = (sx1 + sy1 + sz1 == 1);
        if (is_surf_v1) {
            if (sx1 == 1) {
                if (sx2 == 0) integ = a1 * b2 * c2 * p_line_surf_ortho(X2, yc2, zc2, xc1, Y1, Z1);
                else if (sy2 == 0) integ = a1 * a2 * c2 * line_surf_para(Y1, Z1, xc1, Y2, zc2, xc2);
                else integ = a1 * a2 * b2 * line_surf_para(Z1, Y1, xc1, Z2, yc2, xc2);
            } else if (sy1 == 1) {
                if (sx2 == 0) integ = b1 * b2 * c2 * line_surf_para(X1, Z1, yc1, X2, zc2, yc2);
                else if (sy2 == 0) integ = b1 * a2 * c2 * p_line_surf_ortho(Y2, xc2, zc2, yc1, X1, Z1);
                else integ = b1 * a2 * b2 * line_surf_para(Z1, X1, yc1, Z2, xc2, yc2);
            } else {
                if (sx2 == 0) integ = c1 * b2 * c2 * line_surf_para(X1, Y1, zc1, X2, yc2, zc2);
                else if (sy2 == 0) integ = c1 * a2 * c2 * line_surf_para(Y1, X1, zc1, Y2, xc2, zc2);
                else integ = c1 * a2 * b2 * p_line_surf_ortho(Z2, xc2, yc2, zc1, X1, Y1);
            }
        } else {
            if (sx2 == 1) {
                if (sx1 == 0) integ = a2 * b1 * c1 * p_line_surf_ortho(X1, yc1, zc1, xc2, Y2, Z2);
                else if (sy1 == 0) integ = a2 * a1 * c1 * line_surf_para(Y2, Z2, xc2, Y1, zc1, xc1);
                else integ = a2 * a1 * b1 * line_surf_para(Z2, Y2, xc2, Z1, yc1, xc1);
            } else if (sy2 == 1) {
                if (sx1 == 0) integ = b2 * b1 * c1 * line_surf_para(X2, Z2, yc2, X1, zc1, yc1);
                else if (sy1 == 0) integ = b2 * a1 * c1 * p_line_surf_ortho(Y1, xc1, zc1, yc2, X2, Z2);
                else integ = b2 * a1 * b1 * line_surf_para(Z2, X2, yc2, Z1, xc1, yc1);
            } else {
                if (sx1 == 0) integ = c2 * b1 * c1 * line_surf_para(X2, Y2, zc2, X1, yc1, zc1);
                else if (sy1 == 0) integ = c2 * a1 * c1 * line_surf_para(Y2, X2, zc2, Y1, xc1, zc1);
                else integ = c2 * a1 * b1 * p_line_surf_ortho(Z1, xc1, yc1, zc2, X2, Y2);
            }
        }
    } else {
        if (sx1 == 1) {
            if (sx2 == 1) integ = a1 * a2 * qs_surf_surf_para(Y1, Z1, xc1, Y2, Z2, xc2);
            else if (sy2 == 1) integ = a1 * b2 * p_surf_surf_ortho(Z1, Y1, xc1, Z2, yc2, X2);
            else integ = a1 * c2 * p_surf_surf_ortho(Y1, Z1, xc1, Y2, zc2, X2);
        } else if (sy1 == 1) {
            if (sx2 == 1) integ = b1 * a2 * p_surf_surf_ortho(Z1, X1, yc1, Z2, xc2, Y2);
            else if (sy2 == 1) integ = b1 * b2 * qs_surf_surf_para(X1, Z1, yc1, X2, Z2, yc2);
            else integ = b1 * c2 * p_surf_surf_ortho(X1, Z1, yc1, X2, zc2, Y2);
        } else {
            if (sx2 == 1) integ = c1 * a2 * p_surf_surf_ortho(Y1, X1, zc1, Y2, xc2, Z2);
            else if (sy2 == 1) integ = c1 * b2 * p_surf_surf_ortho(X1, Y1, zc1, X2, yc2, Z2);
            else integ = c1 * c2 * qs_surf_surf_para(X1, Y1, zc1, X2, Y2, zc2);
        }
    }
    return integ;
}

/* face patch of a centre for a caser code, compute_row_P_sup.jl:3-51 */
static void p_face(int caser, const double *c, double sx, double sy, double sz,
                   double *xc, double *yc, double *zc, double *a, double *b, double *cc, double *A)
{
    *xc = c[0]; *yc = c[1]; *zc = c[2];
    switch (caser) {
    case 1: *A = sx * sz; *yc = c[1] - sy / 2; *a = sx; *b = 0; *cc = sz; break;
    case 2: *A = sx * sz; *yc = c[1] + sy / 2; *a = sx; *b = 0; *cc = sz; break;
    case 3: *A = sy * sz; *xc = c[0] - sx / 2; *a = 0; *b = sy; *cc = sz; break;
    case 4: *A = sy * sz; *xc = c[0] + sx / 2; *a = 0; *b = sy; *cc = sz; break;
    case 5: *A = sx * sy; *zc = c[2] - sz / 2; *a = sx; *b = sy; *cc = 0; break;
    default: *A = sx * sy; *zc = c[2] + sz / 2; *a = sx; *b = sy; *cc = 0; break;
    }
}

/* compute_row_P_sup.jl:1-104 (one source centre, N targets). */
void oracle_p_row(const double *src, const double *centers, int64_t N,
                  double sx, double sy, double sz, int caser, int caser2, double *out)
{
    const double eps0 = 8.854187816997944e-12;
    const double pi = 3.141592653589793;
    double xc1, yc1, zc1, a1, b1, c1, A1;
    p_face(caser, src, sx, sy, sz, &xc1, &yc1, &zc1, &a1, &b1, &c1, &A1);
    for (int64_t k = 0; k < N; k++) {
        double cx = centers[k], cy = centers[k + N], cz = centers[k + 2 * N];
        double v;
        switch (caser2) {
        case 1: v = 1 / (4 * pi * eps0 * A1 * sx * sz) * oracle_integ_sup_sup_P(xc1, yc1, zc1, cx, cy - sy / 2, cz, a1, b1, c1, sx, 0, sz); break;
        case 2: v = 1 / (4 * pi * eps0 * A1 * sx * sz) * oracle_integ_sup_sup_P(xc1, yc1, zc1, cx, cy + sy / 2, cz, a1, b1, c1, sx, 0, sz); break;
        case 3: v = 1 / (4 * pi * eps0 * A1 * sy * sz) * oracle_integ_sup_sup_P(xc1, yc1, zc1, cx - sx / 2, cy, cz, a1, b1, c1, 0, sy, sz); break;
        case 4: v = 1 / (4 * pi * eps0 * A1 * sy * sz) * oracle_integ_sup_sup_P(xc1, yc1, zc1, cx + sx / 2, cy, cz, a1, b1, c1, 0, sy, sz); break;
        case 5: v = 1 / (4 * pi * eps0 * A1 * sx * sy) * oracle_integ_sup_sup_P(xc1, yc1, zc1, cx, cy, cz - sz / 2, a1, b1, c1, sx, sy, 0); break;
        default: v = 1 / (4 * pi * eps0 * A1 * sx * sy) * oracle_integ_sup_sup_P(xc1, yc1, zc1, cx, cy, cz + sz / 2, a1, b1, c1, sx, sy, 0); break;
        }
        out[k] = v;
    }
}

/* Routing of every entry of a row: form[n] = id of the closed form the
 * entry evaluates (g_form ids), + 100 when one of its suppression tests sat
 * within a relative 1e-9 of its threshold.  Same arguments as the row
 * functions; values are written to out as well. */
void oracle_lp_row_forms(const double *src, const double *centers, int64_t N,
                         double sx, double sy, double sz, double sx2, double sy2, double sz2,
                         int dc, double *out, int32_t *form)
{
    for (int64_t n = 0; n < N; n++) {
        double c[3] = {centers[n], centers[n + N], centers[n + 2 * N]};
        g_form = -1;
        g_tie = 0;
        oracle_lp_row(src, c, 1, sx, sy, sz, sx2, sy2, sz2, dc, out + n);
        form[n] = g_form + (g_tie ? 100 : 0);
    }
}

void oracle_lp_row_qs_forms(const double *src, const double *centers, int64_t N,
                            double sx, double sy, double sz, double sx2, double sy2, double sz2,
                            int dc, double *out, int32_t *form)
{
    g_qs = 1;
    oracle_lp_row_forms(src, centers, N, sx, sy, sz, sx2, sy2, sz2, dc, out, form);
    g_qs = 0;
}

void oracle_p_row_forms(const double *src, const double *centers, int64_t N,
                        double sx, double sy, double sz, int caser, int caser2, double *out, int32_t *form)
{
    for (int64_t n = 0; n < N; n++) {
        double c[3] = {centers[n], centers[n + N], centers[n + 2 * N]};
        g_form = -1;
        g_tie = 0;
        oracle_p_row(src, c, 1, sx, sy, sz, caser, caser2, out + n);
        form[n] = g_form + (g_tie ? 100 : 0);
    }
}

/* compute_diagonals.jl:136-155 */
double oracle_lp_self(double l, double W, double T)
{
    double w = W / l, t = T / l;
    double r = sqrt(w * w + t * t), aw = sqrt(w * w + 1), at = sqrt(t * t + 1), ar = sqrt(w * w + t * t + 1);
    double mu0 = 4 * 3.141592653589793 * 1e-7;
    double pi = 3.141592653589793;
    return 2 * mu0 * l / pi * (1.0 / 4 * (1 / w * asinh(w / at) + 1 / t * asinh(t / aw) + asinh(1 / r)) +
        1.0 / 24 * ((t * t) / w * asinh(w / (t * at * (r + ar))) + (w * w) / t * asinh(t / (w * aw * (r + ar))) +
        (t * t) / (w * w) * asinh((w * w) / (t * r * (at + ar))) + (w * w) / (t * t) * asinh((t * t) / (w * r * (aw + ar))) +
        1 / (w * (t * t)) * asinh(w * (t * t) / (at * (aw + ar))) + 1 / (t * (w * w)) * asinh(t * (w * w) / (aw * (at + ar)))) -
        1.0 / 6 * (1 / (w * t) * atan(w * t / ar) + t / w * atan(w / (t * ar)) + w / t * atan(t / (w * ar))) -
        1.0 / 60 * ((ar + r + t + at) * (t * t) / ((ar + r) * (r + t) * (t + at) * (at + ar)) +
        (ar + r + w + aw) * (w * w) / ((ar + r) * (r + w) * (w + aw) * (aw + ar)) +
        (ar + aw + 1 + at) / ((ar + aw) * (aw + 1) * (1 + at) * (at + ar))) -
        1.0 / 20 * (1 / (r + ar) + 1 / (aw + ar) + 1 / (at + ar)));
}

/* ---- test hook: every closed form by id, args packed as doubles -------- */
/* id: 0 vol_vol(x1v,y1v,z1v,x2v,y2v,z2v)      12 args
 *     1 vol_surf(x1v,y1v,z1v,x2v,y2v,z2)      11
 *     2 lp_surf_surf_para(x1v,y1v,z1,x2v,y2v,z2) 10
 *     3 lp_surf_surf_ortho(x1v,y1v,z1,x2v,y2,z2v) 10
 *     4 line_vol(x1v,y1v,z1v,x2v,y2,z2)       10
 *     5 line_surf_para(x1v,y1v,z1,x2v,y2,z2)  8
 *     6 line_surf_ortho(x1v,y1,z1,x2,y2v,z2v) 8  (Lp copy)
 *     7 point_vol(x1,y1,z1,x2v,y2v,z2v)       9
 *     8 line_line_ortho_xy(x1v,y1,z1,x2,y2v,z2) 7
 *     9 line_line_parall(x1v,y1,z1,x2v,y2,z2) 8  (Lp copy)
 *    10 point_sup(x1,y1,z1,x2v,y2v,z2)        8
 *    11 line_point(x1v,y3,z3,x1,y1,z1)        6
 *    12 qs_surf_surf_para(x1v,y1v,z1,x2v,y2v,z2) 10
 *    13 p_surf_surf_ortho(x1v,y1v,z1,x2v,y2,z2v) 10
 *    14 p_line_surf_ortho(x1v,y1,z1,x2,y2v,z2v) 8
 *    15 line_line_parall (P copy)             8
 *    19 qs_line_line_sp(x1v,x2v)              4
 *    20 qs_line_point                         6
 *    21 qs_line_surf_ortho                    8
 *    22 qs_line_surf_para                     8
 *    23 qs_line_vol                           10
 *    24 qs_vol_surf                           11
 */
double oracle_closed_form(int id, const double *a)
{
    switch (id) {
    case 0: return lp_vol_vol(a, a + 2, a + 4, a + 6, a + 8, a + 10);
    case 1: return lp_vol_surf(a, a + 2, a + 4, a + 6, a + 8, a[10]);
    case 2: return lp_surf_surf_para(a, a + 2, a[4], a + 5, a + 7, a[9]);
    case 3: return lp_surf_surf_ortho(a, a + 2, a[4], a + 5, a[7], a + 8);
    case 4: return lp_line_vol(a, a + 2, a + 4, a + 6, a[8], a[9]);
    case 5: return line_surf_para(a, a + 2, a[4], a + 5, a[7], a[8]);
    case 6: return line_surf_ortho(a, a[2], a[3], a[4], a + 5, a + 7);
    case 7: return lp_point_vol(a[0], a[1], a[2], a + 3, a + 5, a + 7);
    case 8: return line_line_ortho_xy(a, a[2], a[3], a[4], a + 5, a[7]);
    case 9: return line_line_parall(a, a[2], a[3], a + 4, a[6], a[7], 0);
    case 10: return point_sup(a[0], a[1], a[2], a + 3, a + 5, a[7]);
    case 11: return lp_line_point(a, a[2], a[3], a[4], a[5], a[6]);
    case 12: return qs_surf_surf_para(a, a + 2, a[4], a + 5, a + 7, a[9]);
    case 13: return p_surf_surf_ortho(a, a + 2, a[4], a + 5, a[7], a + 8);
    case 14: return p_line_surf_ortho(a, a[2], a[3], a[4], a + 5, a + 7);
    case 15: return line_line_parall(a, a[2], a[3], a + 4, a[6], a[7], 1);
    case 19: return qs_line_line_sp(a, a + 2);
    case 20: return qs_line_point(a, a[2], a[3], a[4], a[5], a[6]);
    case 21: return qs_line_surf_ortho(a, a[2], a[3], a[4], a + 5, a + 7);
    case 22: return qs_line_surf_para(a, a + 2, a[4], a + 5, a[7], a[8]);
    case 23: return qs_line_vol(a, a + 2, a + 4, a + 6, a[8], a[9]);
    case 24: return qs_vol_surf(a, a + 2, a + 4, a + 6, a + 8, a[10]);
    }
    return NAN;
}
