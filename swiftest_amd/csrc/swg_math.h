// swg_math.h -- per-body / per-pair arithmetic of the hot path, device side.
//
// Floating-point rule (the library is compiled with -ffp-contract=off, so the
// only fused operations are the fma() calls written here):
//   (1) 3-vector squared norms and dot products: fma(z1,z2, fma(y1,y2, x1*x2));
//   (2) pair-sum accumulations acc +/- f*x, velocity kicks v += a*dt and
//       linear drifts r += p*dt: one fma;
//   (3) everything else rounded separately, in the reference's order.
// sqrt and '/' are the IEEE operations.  With this rule every kernel is
// bit-reproducible against a strict CPU evaluation of the same expressions.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace swg {

// src/globals/globals_module.f90:33-38
constexpr double PIBY2 = 1.570796326794896619231321691639751442099;
constexpr double PI3BY2 = 4.712388980384689857693965074919254326296;
constexpr double TWOPI = 6.283185307179586476925286766559005768394;
constexpr double THIRD = 0.333333333333333333333333333333333333333;
constexpr double SIXTH = 0.166666666666666666666666666666666666667;
// src/swiftest/swiftest_drift.f90:11-16
constexpr double E2MAX = 0.36, DM2MAX = 0.16, E2DM2MAX = 0.0016, DANBYB = 1.0E-13;
constexpr int NLAG1 = 50, NLAG2 = 40;
// src/symba/symba_module.f90:20-23
constexpr int NTENC = 3;
constexpr double RHSCALE = 6.5, RSHELL = 0.48075;
// src/swiftest/swiftest_orbel.f90:11
constexpr double TINYVALUE = 4.0e-15;
constexpr double DBL_TINY = 2.2250738585072014e-308;   // tiny(1.0_DP)
constexpr double VSMALL = 1.4916681462400413e-154;     // sqrt(tiny(1.0_DP)), globals_module.f90:135

// status codes, globals_module.f90:86-103
constexpr int ACTIVE = 0, INACTIVE = 1, DISCARDED_DRIFTERR = -7, COLLIDED = -14;
constexpr int DISCARDED_RMAX = -1, DISCARDED_RMIN = -2, DISCARDED_RMAXU = -3, DISCARDED_PERI = -4,
              DISCARDED_PLR = -5;

__host__ __device__ inline double dot3(double x1, double y1, double z1, double x2, double y2, double z2)
{
    return fma(z1, z2, fma(y1, y2, x1 * x2));
}

// ---------------------------------------------------------------------------
// The pair factor Gm / (r2 * sqrt(r2)) of the direct sums
// ---------------------------------------------------------------------------

// For gfx950 this compiler lowers IEEE sqrt and '/' to v_rsq / v_rcp followed
// by Newton and correction steps, wrapped in range scaling (ldexp for tiny
// operands, v_div_scale / v_div_fmas) and special-value fix-ups.  The two
// cores below are those same steps in the same order, without the wrapping.
// Where the wrapping is the identity they give the IEEE results bit for bit:
//   sqrt_core(x): normal x >= 2^-767 (below that the operand is rescaled);
//   div_core(a, b): a = 0, or a and b normal with a, b, a/b and 1/b far from
//   the exponent limits and exponent(a) - exponent(b) < 768.
__device__ inline double sqrt_core(double x)
{
    double y = __builtin_amdgcn_rsq(x);
    double g = x * y;
    double h = y * 0.5;
    double r = fma(-h, g, 0.5);
    g = fma(g, r, g);
    h = fma(h, r, h);
    double d = fma(-g, g, x);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}

__device__ inline double div_core(double a, double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    e = fma(-b, y, 1.0);
    y = fma(y, e, y);
    double q = a * y;
    double r = fma(-b, q, a);
    return fma(r, y, q);
}

// g / (r2 * sqrt(r2)), bit for bit.  The cores are used when r2 lies in
// [2^-400, 2^400] and gok says g is 0 or within [2^-500, 2^500] (so r2*sqrt(r2)
// is within [2^-600, 2^600] and every condition above holds); otherwise the
// IEEE operations run.  Separations of an N-body system in AU are far inside
// that range.
constexpr double PAIR_R2_LO = 0x1p-400, PAIR_R2_HI = 0x1p400;
constexpr double PAIR_G_LO = 0x1p-500, PAIR_G_HI = 0x1p500;

__host__ __device__ inline bool pair_g_ok(double g) { return g == 0.0 || (g >= PAIR_G_LO && g <= PAIR_G_HI); }

__device__ inline double gm_over_r3(double g, double r2, bool gok)
{
    if (gok && r2 >= PAIR_R2_LO && r2 <= PAIR_R2_HI) return div_core(g, r2 * sqrt_core(r2));
    return g / (r2 * sqrt(r2));
}

// g * r2^-1.5 for the fast pair mode (SWG_PAIR_FAST), not bit-identical to
// g / (r2 * sqrt(r2)) but within a few ulp of it.  The hardware reciprocal
// square root y = v_rsq_f64(r2) = r2^-1/2 (1 + d) has |d| < 2^-23 on gfx950
// (measured: at most 2.5e8 ulp over 4M operands).  With e = 1 - r2 y^2 =
// -(2d + d^2), r2^-1.5 = y^3 (1 - e)^-3/2 = y^3 (1 + 3/2 e + 15/8 e^2 +
// O(e^3)), and e^3 < 2^-66, so the result carries only the roundings of its
// seven operations.  The rounding of y*y enters e as well and cancels to
// first order.  Outside 2^-1020 < r2 < 2^1020 y*y leaves the normal range
// and the factor degrades where the IEEE one overflows or underflows anyway.
// c15 = 1.5 and c1875 = 1.875 come from FastConst: held in VGPRs, the
// polynomial is one v_fma_f64 (gfx9 VOP3 takes no literal, and as an
// accumulator the constant would be re-materialized for every pair).
__device__ inline double gm_over_r3_fast(double g, double r2, double c15, double c1875)
{
    const double y = __builtin_amdgcn_rsq(r2);
    const double yy = y * y;
    const double e = fma(-r2, yy, 1.0);
    const double p = fma(e, c1875, c15);
    const double g3 = (g * y) * yy;
    return fma(g3 * e, p, g3);
}

// The two polynomial constants as loop-invariant vector registers.  The
// empty asm statement (no instruction, no memory access) only hides their
// values from constant folding.
struct FastConst {
    double c15 = 1.5, c1875 = 1.875;
    __device__ FastConst()
    {
        asm volatile("" : "+v"(c15), "+v"(c1875));
    }
};

__device__ inline double gm_over_r3_fast(double g, double r2)
{
    return gm_over_r3_fast(g, r2, 1.5, 1.875);
}

// ---------------------------------------------------------------------------
// Danby / universal-variable Kepler drift (swiftest_drift.f90:111-580).
// ---------------------------------------------------------------------------

// swiftest_orbel_scget (swiftest_orbel.f90:147-172)
__device__ inline void scget(double angle, double &sx, double &cx)
{
    int nper = (int)(angle / TWOPI);
    double x = angle - nper * TWOPI;
    if (x < 0.0) x = x + TWOPI;
    sx = sin(x);
    cx = sqrt(1.0 - sx * sx);
    if ((x > PIBY2) && (x < PI3BY2)) cx = -cx;
}

// x / c for a constant integer c > 0, c < 2^26, not a power of two, given
// r = 1/c rounded: q = x*r lies within 1.5 ulp of x/c, the residual
// fma(-q, c, x) is exact, and q + residual*r lies within 2^-52 ulp of x/c.
// x/c is never a rounding midpoint (c has an odd factor above 1) and lies at
// least ulp/(2c) from every midpoint, so the corrected quotient is the IEEE
// quotient x/c whenever that quotient is a normal number.  Three dependent
// operations instead of the division sequence.
__host__ __device__ inline double cdiv(double x, double c, double r)
{
    const double q = x * r;
    return fma(fma(-q, c, x), r, q);
}

// Stumpff c0..c3 with /4 argument reduction (:536-580).  The series
// quotients use cdiv: each one feeds 1 - q, and a quotient too small to be
// normal (|x| < 2^-1014) leaves 1 - q = 1 whichever way it rounds, so
// c0..c3 equal the IEEE evaluation for every finite x.
__host__ __device__ inline void stumpff(double x, double &c0, double &c1, double &c2, double &c3)
{
    int n = 0;
    while (fabs(x) >= 0.1) {
        n += 1;
        x = x / 4.0;
    }
    double t = 1.0 - cdiv(x, 182.0, 1.0 / 182.0);
    t = 1.0 - cdiv(x * t, 132.0, 1.0 / 132.0);
    t = 1.0 - cdiv(x * t, 90.0, 1.0 / 90.0);
    t = 1.0 - cdiv(x * t, 56.0, 1.0 / 56.0);
    t = 1.0 - cdiv(x * t, 30.0, 1.0 / 30.0);
    c2 = (1.0 - cdiv(x * t, 12.0, 1.0 / 12.0)) / 2.0;
    t = 1.0 - cdiv(x, 210.0, 1.0 / 210.0);
    t = 1.0 - cdiv(x * t, 156.0, 1.0 / 156.0);
    t = 1.0 - cdiv(x * t, 110.0, 1.0 / 110.0);
    t = 1.0 - cdiv(x * t, 72.0, 1.0 / 72.0);
    t = 1.0 - cdiv(x * t, 42.0, 1.0 / 42.0);
    c3 = cdiv(1.0 - cdiv(x * t, 20.0, 1.0 / 20.0), 6.0, 1.0 / 6.0);
    c1 = 1.0 - x * c3;
    c0 = 1.0 - x * c2;
    for (int i = n; i >= 1; i--) {
        c3 = (c2 + c0 * c3) / 4.0;
        c2 = (c1 * c1) / 2.0;
        c1 = c0 * c1;
        c0 = (2.0 * c0) * c0 - 1.0;
    }
}

// Kepler's equation in difference form for small dm and e (:234-276)
__device__ inline void kepmd(double dm, double es, double ec, double &x, double &s, double &c)
{
    const double a0 = 39916800.0, a1 = 6652800.0, a2 = 332640.0, a3 = 7920.0, a4 = 110.0;
    double fac1 = 1.0 / (1.0 - ec);
    double q = fac1 * dm;
    double fac2 = (es * es) * fac1 - cdiv(ec, 3.0, 1.0 / 3.0);  // ec is 0 or normal
    x = q * (1.0 - ((0.5 * fac1) * q) * (es - q * fac2));
    double y = x * x;
    s = cdiv(x * (a0 - y * (a1 - y * (a2 - y * (a3 - y * (a4 - y))))), a0, 1.0 / a0);
    c = sqrt(1.0 - s * s);
    double f = ((x - ec * s) + es * (1.0 - c)) - dm;
    double fp = (1.0 - ec * c) + es * s;
    double fpp = ec * s + es * c;
    double fppp = ec * c - es * s;
    double dx = -f / fp;
    dx = -f / (fp + (dx * fpp) / 2.0);
    dx = -f / ((fp + (dx * fpp) / 2.0) + ((dx * dx) * fppp) * SIXTH);
    x = x + dx;
    y = x * x;
    s = cdiv(x * (a0 - y * (a1 - y * (a2 - y * (a3 - y * (a4 - y))))), a0, 1.0 / a0);
    c = sqrt(1.0 - s * s);
}

// f(s) of Kepler's equation in universal variables (:307-334)
__device__ inline double kepu_fchk(double dt, double r0, double mu, double alpha, double u, double s)
{
    double c0, c1, c2, c3;
    stumpff((s * s) * alpha, c0, c1, c2, c3);
    c1 = c1 * s;
    c2 = c2 * (s * s);
    c3 = c3 * ((s * s) * s);
    return ((r0 * c1 + u * c2) + mu * c3) - dt;
}

// 1/b to about 1 ulp for b in [1, 16]: the hardware reciprocal and two
// Newton steps on the device (the host divides).
__host__ __device__ inline double rcp_nr(double b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    double y = __builtin_amdgcn_rcp(b);
    y = fma(y, fma(-b, y, 1.0), y);
    return fma(y, fma(-b, y, 1.0), y);
#else
    return 1.0 / b;
#endif
}

// x**THIRD (THIRD = the double nearest 1/3 = 1/3 - 2^-54/3) for x in
// [2^-1000, 2^1000], correctly rounded except where the exact power lies
// within about 2^-90 (relative) of a rounding midpoint:
//  1. x = M 2^(3k), M in [0.5, 4); a float log2/exp2 seed for M^(1/3)
//     (relative error ~2^-21) and one Halley step y (y^3 + 2M) / (2y^3 + M)
//     give y within ~2^-50 of M^(1/3).
//  2. The residual M - y^3 is formed to ~2^-102 (y^2 and y*y^2 split
//     exactly by fma, M - y*y^2 exact by Sterbenz), and the Newton step
//     d = (M - y^3) / (3y^2) leaves M^(1/3) - (y + d) = O(d^2 / y) ~ 2^-100 y.
//  3. x**THIRD = x^(1/3) exp(-(2^-54/3) ln x); the correction y (2^-54/3)
//     ln x is at most 2^-46 y, so ln x is needed to ~2^-35 absolute: the
//     exponent times ln 2 plus 2 atanh((m-1)/(m+1)) for m in [sqrt(1/2),
//     sqrt(2)), summed to s^13 (truncation ~2^-41).
// The rounded sum y + (d - corr) is then the correctly rounded power (the
// general pow of the device library is not: it differs from it in about a
// quarter of the cube roots), at about half the latency of that pow.
__host__ __device__ inline bool pow_third_ok(double x) { return x >= 0x1p-1000 && x <= 0x1p1000; }

__host__ __device__ inline double pow_third_core(double x)
{
    int e;
    const double m = frexp(x, &e);  // x = m 2^e, m in [0.5, 1)
    const int k = (e >= 0 ? e : e - 2) / 3;  // floor(e / 3)
    const double M = ldexp(m, e - 3 * k);
#if defined(__HIP_DEVICE_COMPILE__)
    double y = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)M) * (1.0f / 3.0f));
#else
    double y = (double)exp2f(log2f((float)M) * (1.0f / 3.0f));
#endif
    const double y3 = (y * y) * y;
    y = y * ((y3 + 2.0 * M) * rcp_nr(2.0 * y3 + M));
    const double s2 = y * y, s2l = fma(y, y, -s2);
    const double p = y * s2, pl = fma(y, s2, -p);
    const double r = fma(-y, s2l, (M - p) - pl);
    const double d = r * rcp_nr(3.0 * s2);
    const bool lo = m < 0.70710678118654752;
    const double m2 = lo ? 2.0 * m : m;
    const double s = (m2 - 1.0) * rcp_nr(m2 + 1.0), z = s * s;
    double t = 1.0 / 13.0;
    t = fma(t, z, 1.0 / 11.0);
    t = fma(t, z, 1.0 / 9.0);
    t = fma(t, z, 1.0 / 7.0);
    t = fma(t, z, 1.0 / 5.0);
    t = fma(t, z, 1.0 / 3.0);
    t = fma(t, z, 1.0);
    const double lnx = (double)(lo ? e - 1 : e) * 0.69314718055994531 + (2.0 * s) * t;
    const double corr = y * (lnx * (0x1p-54 / 3.0));
    return ldexp(y + (d - corr), k);
}

__host__ __device__ inline double pow_third(double x)
{
    if (!pow_third_ok(x)) return pow(x, THIRD);
    return pow_third_core(x);
}

// real root of the cubic for the hyperbolic first guess (:486-533)
__device__ inline int kepu_p3solve(double dt, double r0, double mu, double alpha, double u, double &s)
{
    double denom = (mu - alpha * r0) * SIXTH;
    double a2 = (0.5 * u) / denom;
    double a1 = r0 / denom;
    double a0 = -dt / denom;
    double q = (a1 - (a2 * a2) * THIRD) * THIRD;
    double r = ((a1 * a2 - 3.0 * a0) * SIXTH) - ((a2 * a2) * a2) / 27.0;
    double sq2 = (q * q) * q + r * r;
    if (sq2 >= 0.0) {
        double sq = sqrt(sq2);
        // t <= 0 ? -(-t)**THIRD : t**THIRD, with -t == |t| in the first arm
        const double t1 = r + sq, t2 = r - sq;
        const double w1 = pow_third(fabs(t1)), w2 = pow_third(fabs(t2));
        double p1 = (t1 <= 0.0) ? -w1 : w1;
        double p2 = (t2 <= 0.0) ? -w2 : w2;
        s = (p1 + p2) - a2 * THIRD;
        return 0;
    }
    s = 0.0;
    return 1;
}

// initial guess (:337-378)
__device__ inline double kepu_guess(double dt, double r0, double mu, double alpha, double u)
{
    double s;
    if (alpha > 0.0) {
        if (dt / r0 <= 0.4) {
            s = dt / r0 - ((dt * dt) * u) / (((2.0 * r0) * r0) * r0);
        } else {
            double a = mu / alpha;
            double en = sqrt(mu / ((a * a) * a));
            double ec = 1.0 - r0 / a;
            double es = u / ((en * a) * a);
            double e = sqrt(ec * ec + es * es);
            double y = en * dt - es;
            double sy, cy;
            scget(y, sy, cy);
            double sigma = copysign(1.0, es * cy + ec * sy);
            double x = y + (sigma * 0.85) * e;
            s = x / sqrt(alpha);
        }
    } else {
        if (kepu_p3solve(dt, r0, mu, alpha, u, s) != 0) s = dt / r0;
    }
    return s;
}

// Newton with Halley/third-order corrections, at most 7 iterations (:435-483)
__device__ inline int kepu_new(double &s, double dt, double r0, double mu, double alpha, double u, double &fp,
                               double &c1, double &c2, double &c3)
{
    for (int nc = 0; nc <= 6; nc++) {
        double c0;
        stumpff((s * s) * alpha, c0, c1, c2, c3);
        c1 = c1 * s;
        c2 = (c2 * s) * s;
        c3 = ((c3 * s) * s) * s;
        double f = ((r0 * c1 + u * c2) + mu * c3) - dt;
        fp = (r0 * c0 + u * c1) + mu * c2;
        double fpp = (-(r0 * alpha) + mu) * c1 + u * c0;
        double fppp = (-(r0 * alpha) + mu) * c0 - (u * alpha) * c1;
        double ds = -f / fp;
        ds = -f / (fp + (ds * fpp) / 2.0);
        ds = -f / ((fp + (ds * fpp) / 2.0) + cdiv((ds * ds) * fppp, 6.0, 1.0 / 6.0));
        s = s + ds;
        double fdt = f / dt;
        if (fdt * fdt < DANBYB * DANBYB) return 0;
    }
    return 1;
}

// Laguerre fallback, Danby (1988) (:381-432)
__device__ inline int kepu_lag(double &s, double dt, double r0, double mu, double alpha, double u, double &fp,
                               double &c1, double &c2, double &c3)
{
    const double ln = 5.0;
    int ncmax = (alpha < 0.0) ? NLAG2 : NLAG1;
    for (int nc = 0; nc <= ncmax; nc++) {
        double c0;
        stumpff((s * s) * alpha, c0, c1, c2, c3);
        c1 = c1 * s;
        c2 = (c2 * s) * s;
        c3 = ((c3 * s) * s) * s;
        double f = ((r0 * c1 + u * c2) + mu * c3) - dt;
        fp = (r0 * c0 + u * c1) + mu * c2;
        double fpp = (-(r0 * alpha) + mu) * c1 + u * c0;
        double disc = fabs((((ln - 1.0) * (ln - 1.0)) * fp) * fp - (((ln - 1.0) * ln) * f) * fpp);
        double ds = -((ln * f) / (fp + copysign(1.0, fp) * sqrt(disc)));
        s = s + ds;
        double fdt = f / dt;
        if (fdt * fdt < DANBYB * DANBYB) return 0;
    }
    return 2;
}

// one Danby drift of (r, v) over dt0 (:141-231); returns iflag
__device__ __attribute__((always_inline)) inline int drift_dan(double mu, double &rx0, double &ry0, double &rz0, double &vx0, double &vy0,
                                double &vz0, double dt0)
{
    double dt = dt0;
    double r0 = sqrt(dot3(rx0, ry0, rz0, rx0, ry0, rz0));
    double v0s = dot3(vx0, vy0, vz0, vx0, vy0, vz0);
    double u = dot3(rx0, ry0, rz0, vx0, vy0, vz0);
    double alpha = (2.0 * mu) / r0 - v0s;
    double f, g, fdot, gdot;
    if (alpha > 0.0) {
        double a = mu / alpha;
        double asq = a * a;
        double en = sqrt(mu / (a * asq));
        double ec = 1.0 - r0 / a;
        double es = u / (en * asq);
        double esq = ec * ec + es * es;
        double dm = dt * en - (double)((int)((dt * en) / TWOPI)) * TWOPI;
        dt = dm / en;
        if ((esq < E2MAX) && (dm * dm < DM2MAX) && (esq * (dm * dm) < E2DM2MAX)) {
            double xkep, s, c;
            kepmd(dm, es, ec, xkep, s, c);
            double fchk = ((xkep - ec * s) + es * (1.0 - c)) - dm;
            if (fchk * fchk > DANBYB * DANBYB) return 1;
            double fp = (1.0 - ec * c) + es * s;
            f = (a / r0) * (c - 1.0) + 1.0;
            g = dt + (s - xkep) / en;
            fdot = ((-(a / (r0 * fp))) * en) * s;
            gdot = (c - 1.0) / fp + 1.0;
            double rx = rx0 * f + vx0 * g, ry = ry0 * f + vy0 * g, rz = rz0 * f + vz0 * g;
            double vx = rx0 * fdot + vx0 * gdot, vy = ry0 * fdot + vy0 * gdot, vz = rz0 * fdot + vz0 * gdot;
            rx0 = rx; ry0 = ry; rz0 = rz;
            vx0 = vx; vy0 = vy; vz0 = vz;
            return 0;
        }
    }
    double s = kepu_guess(dt, r0, mu, alpha, u);
    double st = s, fp, c1, c2, c3;
    int iflag = kepu_new(s, dt, r0, mu, alpha, u, fp, c1, c2, c3);
    if (iflag != 0) {
        double fo = kepu_fchk(dt, r0, mu, alpha, u, st);
        double fn = kepu_fchk(dt, r0, mu, alpha, u, s);
        if (fabs(fo) < fabs(fn)) s = st;
        iflag = kepu_lag(s, dt, r0, mu, alpha, u, fp, c1, c2, c3);
    }
    if (iflag == 0) {
        f = 1.0 - (mu / r0) * c2;
        g = dt - mu * c3;
        fdot = -((mu / (fp * r0)) * c1);
        gdot = 1.0 - (mu / fp) * c2;
        double rx = rx0 * f + vx0 * g, ry = ry0 * f + vy0 * g, rz = rz0 * f + vz0 * g;
        double vx = rx0 * fdot + vx0 * gdot, vy = ry0 * fdot + vy0 * gdot, vz = rz0 * fdot + vz0 * gdot;
        rx0 = rx; ry0 = ry; rz0 = rz;
        vx0 = vx; vy0 = vy; vz0 = vz;
    }
    return iflag;
}

// swiftest_drift_one (:111-138): on failure retry as ten dt/10 substeps,
// stopping at the first substep that fails.  Written with a single call of
// drift_dan so that the compiler inlines it (a call would save and restore
// the caller's live registers around every drift).
__device__ inline int drift_one(double mu, double &rx, double &ry, double &rz, double &vx, double &vy, double &vz,
                                double dt)
{
    double h = dt;
    int i = -1, fl;
    for (;;) {
        fl = drift_dan(mu, rx, ry, rz, vx, vy, vz, h);
        if (i < 0) {  // the full step
            if (fl == 0) break;
            h = 0.1 * dt;
            i = 0;
            continue;
        }
        if (fl != 0 || ++i >= 10) break;
    }
    return fl;
}

// ---------------------------------------------------------------------------
// Encounter / collision tests
// ---------------------------------------------------------------------------

// encounter_check_one (encounter_check.f90:573-621); returns lencounter and
// writes lvdotr.
__host__ __device__ inline bool encounter_one(double xr, double yr, double zr, double vxr, double vyr, double vzr,
                                              double renc, double dt, bool &lvdotr)
{
    double r2 = dot3(xr, yr, zr, xr, yr, zr);
    double r2crit = renc * renc;
    double vdotr, r2min;
    if (r2 > r2crit) {
        vdotr = dot3(vxr, vyr, vzr, xr, yr, zr);
        if (vdotr > 0.0) {
            r2min = r2;
        } else {
            double v2 = dot3(vxr, vyr, vzr, vxr, vyr, vzr);
            if (v2 <= VSMALL) {
                r2min = r2;
            } else {
                double tmin = -vdotr / v2;
                if (tmin < dt)
                    r2min = r2 - (vdotr * vdotr) / v2;
                else
                    r2min = (r2 + (2.0 * vdotr) * dt) + v2 * (dt * dt);
            }
        }
    } else {
        vdotr = -1.0;
        r2min = r2;
    }
    lvdotr = (vdotr < 0.0);
    return lvdotr && (r2min <= r2crit);
}

// swiftest_orbel_xv2aeq (swiftest_orbel.f90:700-764): semi-major axis and
// pericentre
__host__ __device__ inline void xv2aq(double mu, double rx, double ry, double rz, double vx, double vy, double vz, double &a,
                             double &q)
{
    a = 0.0;
    q = 0.0;
    double r = sqrt(dot3(rx, ry, rz, rx, ry, rz));
    double v2 = dot3(vx, vy, vz, vx, vy, vz);
    double hx = ry * vz - rz * vy, hy = rz * vx - rx * vz, hz = rx * vy - ry * vx;
    double h2 = dot3(hx, hy, hz, hx, hy, hz);
    if (h2 < DBL_TINY) return;
    double energy = 0.5 * v2 - mu / r;
    double e = 0.0, fac = 0.0;
    int type;
    if (fabs((energy * r) / mu) < sqrt(TINYVALUE)) {
        type = 0;
    } else {
        a = (-0.5 * mu) / energy;
        if (a < 0.0) {
            fac = -h2 / (mu * a);
            type = (fac > TINYVALUE) ? 1 : 0;
        } else {
            type = -1;
        }
    }
    if (type == -1) {
        fac = 1.0 - h2 / (mu * a);
        if (fac > TINYVALUE) e = sqrt(fac);
        q = a * (1.0 - e);
    } else if (type == 0) {
        a = (0.5 * h2) / mu;
        q = a;
    } else {
        e = sqrt(1.0 + fac);
        q = a * (1.0 - e);
    }
}

// the pericentre alone
__host__ __device__ inline double xv2q(double mu, double rx, double ry, double rz, double vx, double vy, double vz)
{
    double a, q;
    xv2aq(mu, rx, ry, rz, vx, vy, vz, a, q);
    return q;
}

// collision_check_one (collision_check.f90:16-50)
__host__ __device__ inline bool collision_one(double xr, double yr, double zr, double vxr, double vyr, double vzr,
                                     double Gmtot, double rlim, double dt, bool lvdotr)
{
    double r2 = dot3(xr, yr, zr, xr, yr, zr);
    if (r2 <= rlim * rlim) return true;
    double vdotr = dot3(xr, yr, zr, vxr, vyr, vzr);
    if (lvdotr && (vdotr > 0.0)) {
        double tcr2 = r2 / dot3(vxr, vyr, vzr, vxr, vyr, vzr);
        if (tcr2 <= dt * dt) return xv2q(Gmtot, xr, yr, zr, vxr, vyr, vzr) < rlim;
    }
    return false;
}

// ---------------------------------------------------------------------------
// SyMBA shell kick factor (symba_kick.f90:187-208)
// ---------------------------------------------------------------------------

// x**n for a run-time integer n by binary powering (what RSHELL**(2*irecl)
// compiles to)
__host__ __device__ constexpr double powi(double x, int m)
{
    unsigned n = m < 0 ? -(unsigned)m : (unsigned)m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

// r2**(-1.5), correctly rounded via a double-double x^(3/2) and a corrected
// reciprocal
__host__ __device__ inline double pow_m1p5(double x)
{
    double s = sqrt(x);
    double sl = fma(-s, s, x) / (2.0 * s);
    double ph = x * s;
    double pl = fma(x, s, -ph) + x * sl;
    double yh = 1.0 / ph;
    double e = fma(-yh, ph, 1.0) - yh * pl;
    return fma(yh, e, yh);
}

// fac of one listed pair at the two shell radii^2 rp (sign +1) and rm (sign
// -1, when minus): ok is false when r2 < ri * RSHELL^2 (the pair is dropped
// for that kick); inside ri the factor is r2^-1.5 times the switch
// polynomial, outside it 1 / (r2 sqrt(r2)) (symba_kick.f90:163-196).  Both
// forms are computed for both radii and selected, with no branch: the two
// shells' chains then overlap, and sqrt(r2) and r2^-1.5 are formed once.
// Each selected factor is the expression the reference evaluates for it.
__device__ inline void shell_facs(double rp, double rm, double r2, bool minus, bool &okp, double &facp, bool &okm,
                                  double &facm)
{
    const double r = sqrt(r2);
    const double out = 1.0 / (r2 * r);
    const double p15 = pow_m1p5(r2);
    const double sp = sqrt(rp), sm = sqrt(rm);
    const double rrp = (sp - r) / (sp * (1.0 - RSHELL)), rrm = (sm - r) / (sm * (1.0 - RSHELL));
    const double inp = p15 * ((1.0 - 3.0 * (rrp * rrp)) + 2.0 * ((rrp * rrp) * rrp));
    const double inm = p15 * ((1.0 - 3.0 * (rrm * rrm)) + 2.0 * ((rrm * rrm) * rrm));
    okp = !(r2 < rp * (RSHELL * RSHELL));
    okm = minus && !(r2 < rm * (RSHELL * RSHELL));
    facp = r2 < rp ? inp : out;
    facm = r2 < rm ? inm : out;
}

// renc at recursion depth irec: rhill * RHSCALE * RSHELL**irec with the
// power built by repeated multiplication (symba_util.f90:245-267)
__host__ __device__ constexpr double rshell_pow(int irec)
{
    double r = 1.0;
    for (int i = 1; i <= irec; i++) r = r * RSHELL;
    return r;
}

}  // namespace swg
