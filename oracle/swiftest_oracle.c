/*
 * swiftest_oracle.c -- CPU restatement of Swiftest's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle and the CPU baseline); never linked
 * into the product.  See swiftest_oracle.h for the floating-point rule.
 * All citations are to carlislewishard/swiftest v2023.10.2 (src/...).
 *
 * Parity pinning (what checks each part, besides the HIP path itself) lives
 * in tests/test_oracle.py: closed forms for the pair kernels and the encounter
 * test, analytic two-body propagation and invariants for the Danby drift,
 * Stumpff closed forms, shell-polynomial continuity, and the reference's own
 * conservation thresholds (tests/test_swiftest.py:118-121) on whole runs;
 * the collision merges in tests/test_oracle_merge.py (conservation laws, the
 * closed form of the velocity torque, the overlap shift).  Parity unpinned:
 * the rearray's carry-over of old entries by list position and the cut of
 * the list to its old length after a merge (swiftest_util.f90:1744-1805),
 * compared only with the device's restatement.
 */
#include "swiftest_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* src/globals/globals_module.f90:33-38 */
static const double PI = 3.141592653589793238462643383279502884197;
static const double PIBY2 = 1.570796326794896619231321691639751442099;
static const double PI3BY2 = 4.712388980384689857693965074919254326296;
static const double TWOPI = 6.283185307179586476925286766559005768394;
static const double THIRD = 0.333333333333333333333333333333333333333;
static const double SIXTH = 0.166666666666666666666666666666666666667;
/* src/swiftest/swiftest_drift.f90:11-16 */
static const double E2MAX = 0.36, DM2MAX = 0.16, E2DM2MAX = 0.0016, DANBYB = 1.0E-13;
#define NLAG1 50
#define NLAG2 40
/* src/symba/symba_module.f90:20-23 */
#define NTENC 3
static const double RHSCALE = 6.5, RSHELL = 0.48075;
/* src/swiftest/swiftest_orbel.f90:11 */
static const double TINYVALUE = 4.0e-15;

/* rule (1): 3-vector dot product as an fma chain */
static inline double dot3(double x1, double y1, double z1, double x2, double y2, double z2)
{
    return fma(z1, z2, fma(y1, y2, x1 * x2));
}

static int omp_threads(int n) { return n > 0 ? n : 1; }

void swo_free(void *p) { free(p); }

/* ======================================================================
 * Pair kernels -- src/swiftest/swiftest_kick.f90
 * ==================================================================== */

/* swiftest_kick_getacch_int_one_pl, swiftest_kick.f90:418-446 */
static inline void int_one_pl(double rji2, double xr, double yr, double zr, double Gmi, double Gmj,
                              double *ai, double *aj)
{
    double irij3 = 1.0 / (rji2 * sqrt(rji2));
    double faci = Gmi * irij3;
    double facj = Gmj * irij3;
    ai[0] = fma(facj, xr, ai[0]);
    ai[1] = fma(facj, yr, ai[1]);
    ai[2] = fma(facj, zr, ai[2]);
    aj[0] = fma(-faci, xr, aj[0]);
    aj[1] = fma(-faci, yr, aj[1]);
    aj[2] = fma(-faci, zr, aj[2]);
}

/* swiftest_kick_getacch_int_all_tri_rad_pl (:165-271), and the _norad
 * variant (:274-371) when radius == NULL.  The lmtiny branch (:190-216) is the
 * serial (one-thread) order of its OpenMP array reduction. */
void swo_kick_getacch_int_all_tri(int npl, int nplm, const double *r, const double *Gmass,
                                  const double *radius, double *acc)
{
    int nplt = npl - nplm;
    if (nplt > nplm) {
        double *ahi = calloc((size_t)3 * npl, sizeof(double));
        double *ahj = calloc((size_t)3 * npl, sizeof(double));
        for (int i = 0; i < nplm; i++) {
            for (int j = i + 1; j < npl; j++) {
                double rx = r[3 * j] - r[3 * i];
                double ry = r[3 * j + 1] - r[3 * i + 1];
                double rz = r[3 * j + 2] - r[3 * i + 2];
                double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                if (radius) {
                    double s = radius[i] + radius[j];
                    if (!(rji2 > s * s)) continue;
                }
                int_one_pl(rji2, rx, ry, rz, Gmass[i], Gmass[j], &ahi[3 * i], &ahj[3 * j]);
            }
        }
        for (int i = 0; i < npl; i++)
            for (int d = 0; d < 3; d++) acc[3 * i + d] = (acc[3 * i + d] + ahi[3 * i + d]) + ahj[3 * i + d];
        free(ahi);
        free(ahj);
        return;
    }
    /* full rows (:219-240): i = 1..nplm, j = 1..npl ascending, j /= i */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nplm; i++) {
        double ax = acc[3 * i], ay = acc[3 * i + 1], az = acc[3 * i + 2];
        double xi = r[3 * i], yi = r[3 * i + 1], zi = r[3 * i + 2];
        for (int j = 0; j < npl; j++) {
            if (j == i) continue;
            double rx = r[3 * j] - xi, ry = r[3 * j + 1] - yi, rz = r[3 * j + 2] - zi;
            double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            if (radius) {
                double s = radius[i] + radius[j];
                if (!(rji2 > s * s)) continue;
            }
            double fac = Gmass[j] / (rji2 * sqrt(rji2));
            ax = fma(fac, rx, ax);
            ay = fma(fac, ry, ay);
            az = fma(fac, rz, az);
        }
        acc[3 * i] = ax;
        acc[3 * i + 1] = ay;
        acc[3 * i + 2] = az;
    }
    /* partially interacting rows (:243-262): i = nplm+1..npl, j = 1..nplm */
    for (int i = nplm; i < npl; i++) {
        double ax = acc[3 * i], ay = acc[3 * i + 1], az = acc[3 * i + 2];
        for (int j = 0; j < nplm; j++) {
            double rx = r[3 * j] - r[3 * i], ry = r[3 * j + 1] - r[3 * i + 1], rz = r[3 * j + 2] - r[3 * i + 2];
            double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            if (radius) {
                double s = radius[i] + radius[j];
                if (!(rji2 > s * s)) continue;
            }
            double fac = Gmass[j] / (rji2 * sqrt(rji2));
            ax = fma(fac, rx, ax);
            ay = fma(fac, ry, ay);
            az = fma(fac, rz, az);
        }
        acc[3 * i] = ax;
        acc[3 * i + 1] = ay;
        acc[3 * i + 2] = az;
    }
}

/* swiftest_kick_getacch_int_all_flat_rad_pl (:69-115) / _norad (:118-162)
 * over a 1-based (2, nplpl) pair list; one-thread order of the reduction. */
void swo_kick_getacch_int_all_flat(int npl, int64_t nplpl, const int32_t *k_plpl, const double *r,
                                   const double *Gmass, const double *radius, double *acc)
{
    double *ahi = calloc((size_t)3 * (npl > 0 ? npl : 1), sizeof(double));
    double *ahj = calloc((size_t)3 * (npl > 0 ? npl : 1), sizeof(double));
    for (int64_t k = 0; k < nplpl; k++) {
        int i = k_plpl[2 * k] - 1, j = k_plpl[2 * k + 1] - 1;
        double rx = r[3 * j] - r[3 * i], ry = r[3 * j + 1] - r[3 * i + 1], rz = r[3 * j + 2] - r[3 * i + 2];
        double rji2 = dot3(rx, ry, rz, rx, ry, rz);
        if (radius) {
            double s = radius[i] + radius[j];
            if (!(rji2 > s * s)) continue;
        }
        int_one_pl(rji2, rx, ry, rz, Gmass[i], Gmass[j], &ahi[3 * i], &ahj[3 * j]);
    }
    for (int i = 0; i < npl; i++)
        for (int d = 0; d < 3; d++) acc[3 * i + d] = (acc[3 * i + d] + ahi[3 * i + d]) + ahj[3 * i + d];
    free(ahi);
    free(ahj);
}

/* swiftest_kick_getacch_int_all_tp (:374-415) with int_one_tp (:449-470).
 * OpenMP reduction(+:acc): each particle's private sum starts at zero and is
 * added to the incoming acc once at the end. */
void swo_kick_getacch_int_all_tp(int ntp, int npl, const double *rtp, const double *rpl,
                                 const double *GMpl, const int32_t *lmask, double *acc)
{
#pragma omp parallel for schedule(static)
    for (int i = 0; i < ntp; i++) {
        if (!lmask[i]) continue;
        double px = 0.0, py = 0.0, pz = 0.0;
        double xi = rtp[3 * i], yi = rtp[3 * i + 1], zi = rtp[3 * i + 2];
        for (int j = 0; j < npl; j++) {
            double rx = xi - rpl[3 * j], ry = yi - rpl[3 * j + 1], rz = zi - rpl[3 * j + 2];
            double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            double fac = GMpl[j] / (rji2 * sqrt(rji2));
            px = fma(-fac, rx, px);
            py = fma(-fac, ry, py);
            pz = fma(-fac, rz, pz);
        }
        acc[3 * i] = acc[3 * i] + px;
        acc[3 * i + 1] = acc[3 * i + 1] + py;
        acc[3 * i + 2] = acc[3 * i + 2] + pz;
    }
}

/* ======================================================================
 * Danby drift -- src/swiftest/swiftest_drift.f90
 * ==================================================================== */

/* swiftest_orbel_scget, src/swiftest/swiftest_orbel.f90:147-172 */
static void orbel_scget(double angle, double *sx, double *cx)
{
    int nper = (int)(angle / TWOPI);
    double x = angle - nper * TWOPI;
    if (x < 0.0) x = x + TWOPI;
    *sx = sin(x);
    *cx = sqrt(1.0 - (*sx) * (*sx));
    if ((x > PIBY2) && (x < PI3BY2)) *cx = -*cx;
}

/* swiftest_drift_kepu_stumpff, :536-580 */
void swo_drift_kepu_stumpff(double x, double *c0o, double *c1o, double *c2o, double *c3o)
{
    int n = 0;
    const double xm = 0.1;
    while (fabs(x) >= xm) {
        n = n + 1;
        x = x / 4.0;
    }
    double t;
    t = 1.0 - x / 182.0;
    t = 1.0 - (x * t) / 132.0;
    t = 1.0 - (x * t) / 90.0;
    t = 1.0 - (x * t) / 56.0;
    t = 1.0 - (x * t) / 30.0;
    double c2 = (1.0 - (x * t) / 12.0) / 2.0;
    t = 1.0 - x / 210.0;
    t = 1.0 - (x * t) / 156.0;
    t = 1.0 - (x * t) / 110.0;
    t = 1.0 - (x * t) / 72.0;
    t = 1.0 - (x * t) / 42.0;
    double c3 = (1.0 - (x * t) / 20.0) / 6.0;
    double c1 = 1.0 - x * c3;
    double c0 = 1.0 - x * c2;
    for (int i = n; i >= 1; i--) {
        c3 = (c2 + c0 * c3) / 4.0;
        c2 = (c1 * c1) / 2.0;
        c1 = c0 * c1;
        c0 = (2.0 * c0) * c0 - 1.0;
        x = x * 4.0;
    }
    *c0o = c0;
    *c1o = c1;
    *c2o = c2;
    *c3o = c3;
}

/* swiftest_drift_kepmd, :234-276 */
static void drift_kepmd(double dm, double es, double ec, double *xo, double *so, double *co)
{
    const double a0 = 39916800.0, a1 = 6652800.0, a2 = 332640.0, a3 = 7920.0, a4 = 110.0;
    double fac1 = 1.0 / (1.0 - ec);
    double q = fac1 * dm;
    double fac2 = (es * es) * fac1 - ec / 3.0;
    double x = q * (1.0 - ((0.5 * fac1) * q) * (es - q * fac2));
    double y = x * x;
    double s = (x * (a0 - y * (a1 - y * (a2 - y * (a3 - y * (a4 - y)))))) / a0;
    double c = sqrt(1.0 - s * s);
    double f = ((x - ec * s) + es * (1.0 - c)) - dm;
    double fp = (1.0 - ec * c) + es * s;
    double fpp = ec * s + es * c;
    double fppp = ec * c - es * s;
    double dx = -f / fp;
    dx = -f / (fp + (dx * fpp) / 2.0);
    dx = -f / ((fp + (dx * fpp) / 2.0) + ((dx * dx) * fppp) * SIXTH);
    x = x + dx;
    y = x * x;
    s = (x * (a0 - y * (a1 - y * (a2 - y * (a3 - y * (a4 - y)))))) / a0;
    c = sqrt(1.0 - s * s);
    *xo = x;
    *so = s;
    *co = c;
}

/* swiftest_drift_kepu_fchk, :307-334 */
static double drift_kepu_fchk(double dt, double r0, double mu, double alpha, double u, double s)
{
    double x = (s * s) * alpha;
    double c0, c1, c2, c3;
    swo_drift_kepu_stumpff(x, &c0, &c1, &c2, &c3);
    c1 = c1 * s;
    c2 = c2 * (s * s);
    c3 = c3 * ((s * s) * s);
    return ((r0 * c1 + u * c2) + mu * c3) - dt;
}

/* swiftest_drift_kepu_p3solve, :486-533 */
static void drift_kepu_p3solve(double dt, double r0, double mu, double alpha, double u, double *s,
                               int *iflag)
{
    double denom = (mu - alpha * r0) * SIXTH;
    double a2 = (0.5 * u) / denom;
    double a1 = r0 / denom;
    double a0 = -dt / denom;
    double q = (a1 - (a2 * a2) * THIRD) * THIRD;
    double r = ((a1 * a2 - 3.0 * a0) * SIXTH) - ((a2 * a2) * a2) / 27.0;
    double sq2 = (q * q) * q + r * r;
    if (sq2 >= 0.0) {
        double sq = sqrt(sq2), p1, p2;
        if ((r + sq) <= 0.0)
            p1 = -pow(-(r + sq), THIRD);
        else
            p1 = pow(r + sq, THIRD);
        if ((r - sq) <= 0.0)
            p2 = -pow(-(r - sq), THIRD);
        else
            p2 = pow(r - sq, THIRD);
        *iflag = 0;
        *s = (p1 + p2) - a2 * THIRD;
    } else {
        *iflag = 1;
        *s = 0.0;
    }
}

/* swiftest_drift_kepu_guess, :337-378 */
static double drift_kepu_guess(double dt, double r0, double mu, double alpha, double u)
{
    const double thresh = 0.4, danbyk = 0.85;
    double s;
    if (alpha > 0.0) {
        if (dt / r0 <= thresh) {
            s = dt / r0 - ((dt * dt) * u) / (((2.0 * r0) * r0) * r0);
        } else {
            double a = mu / alpha;
            double en = sqrt(mu / ((a * a) * a));
            double ec = 1.0 - r0 / a;
            double es = u / ((en * a) * a);
            double e = sqrt(ec * ec + es * es);
            double y = en * dt - es;
            double sy, cy;
            orbel_scget(y, &sy, &cy);
            double sigma = copysign(1.0, es * cy + ec * sy);
            double x = y + (sigma * danbyk) * e;
            s = x / sqrt(alpha);
        }
    } else {
        int iflag;
        drift_kepu_p3solve(dt, r0, mu, alpha, u, &s, &iflag);
        if (iflag != 0) s = dt / r0;
    }
    return s;
}

/* swiftest_drift_kepu_new, :435-483 */
static void drift_kepu_new(double *s, double dt, double r0, double mu, double alpha, double u, double *fp,
                           double *c1, double *c2, double *c3, int *iflag)
{
    for (int nc = 0; nc <= 6; nc++) {
        double x = ((*s) * (*s)) * alpha;
        double c0;
        swo_drift_kepu_stumpff(x, &c0, c1, c2, c3);
        *c1 = (*c1) * (*s);
        *c2 = ((*c2) * (*s)) * (*s);
        *c3 = (((*c3) * (*s)) * (*s)) * (*s);
        double f = ((r0 * (*c1) + u * (*c2)) + mu * (*c3)) - dt;
        *fp = (r0 * c0 + u * (*c1)) + mu * (*c2);
        double fpp = (-(r0 * alpha) + mu) * (*c1) + u * c0;
        double fppp = (-(r0 * alpha) + mu) * c0 - (u * alpha) * (*c1);
        double ds = -f / (*fp);
        ds = -f / ((*fp) + (ds * fpp) / 2.0);
        ds = -f / (((*fp) + (ds * fpp) / 2.0) + ((ds * ds) * fppp) / 6.0);
        *s = (*s) + ds;
        double fdt = f / dt;
        if (fdt * fdt < DANBYB * DANBYB) {
            *iflag = 0;
            return;
        }
    }
    *iflag = 1;
}

/* swiftest_drift_kepu_lag, :381-432 */
static void drift_kepu_lag(double *s, double dt, double r0, double mu, double alpha, double u, double *fp,
                           double *c1, double *c2, double *c3, int *iflag)
{
    const double ln = 5.0;
    int ncmax = (alpha < 0.0) ? NLAG2 : NLAG1;
    for (int nc = 0; nc <= ncmax; nc++) {
        double x = ((*s) * (*s)) * alpha;
        double c0;
        swo_drift_kepu_stumpff(x, &c0, c1, c2, c3);
        *c1 = (*c1) * (*s);
        *c2 = ((*c2) * (*s)) * (*s);
        *c3 = (((*c3) * (*s)) * (*s)) * (*s);
        double f = ((r0 * (*c1) + u * (*c2)) + mu * (*c3)) - dt;
        *fp = (r0 * c0 + u * (*c1)) + mu * (*c2);
        double fpp = (-(r0 * alpha) + mu) * (*c1) + u * c0;
        double disc = fabs((((ln - 1.0) * (ln - 1.0)) * (*fp)) * (*fp) - (((ln - 1.0) * ln) * f) * fpp);
        double ds = -((ln * f) / ((*fp) + copysign(1.0, *fp) * sqrt(disc)));
        *s = (*s) + ds;
        double fdt = f / dt;
        if (fdt * fdt < DANBYB * DANBYB) {
            *iflag = 0;
            return;
        }
    }
    *iflag = 2;
}

/* swiftest_drift_kepu, :279-304 */
static void drift_kepu(double dt, double r0, double mu, double alpha, double u, double *fp, double *c1,
                       double *c2, double *c3, int *iflag)
{
    double s = drift_kepu_guess(dt, r0, mu, alpha, u);
    double st = s;
    drift_kepu_new(&s, dt, r0, mu, alpha, u, fp, c1, c2, c3, iflag);
    if (*iflag != 0) {
        double fo = drift_kepu_fchk(dt, r0, mu, alpha, u, st);
        double fn = drift_kepu_fchk(dt, r0, mu, alpha, u, s);
        if (fabs(fo) < fabs(fn)) s = st;
        drift_kepu_lag(&s, dt, r0, mu, alpha, u, fp, c1, c2, c3, iflag);
    }
}

/* swiftest_drift_dan, :141-231 */
static void drift_dan(double mu, double *rx0, double *ry0, double *rz0, double *vx0, double *vy0,
                      double *vz0, double dt0, int *iflag)
{
    *iflag = 0;
    double dt = dt0;
    double r0 = sqrt(dot3(*rx0, *ry0, *rz0, *rx0, *ry0, *rz0));
    double v0s = dot3(*vx0, *vy0, *vz0, *vx0, *vy0, *vz0);
    double u = dot3(*rx0, *ry0, *rz0, *vx0, *vy0, *vz0);
    double alpha = (2.0 * mu) / r0 - v0s;
    double f, g, fdot, gdot, fp, c1, c2, c3;
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
            drift_kepmd(dm, es, ec, &xkep, &s, &c);
            double fchk = ((xkep - ec * s) + es * (1.0 - c)) - dm;
            if (fchk * fchk > DANBYB * DANBYB) {
                *iflag = 1;
                return;
            }
            fp = (1.0 - ec * c) + es * s;
            f = (a / r0) * (c - 1.0) + 1.0;
            g = dt + (s - xkep) / en;
            fdot = ((-(a / (r0 * fp))) * en) * s;
            gdot = (c - 1.0) / fp + 1.0;
            double rx = *rx0 * f + *vx0 * g, ry = *ry0 * f + *vy0 * g, rz = *rz0 * f + *vz0 * g;
            double vx = *rx0 * fdot + *vx0 * gdot, vy = *ry0 * fdot + *vy0 * gdot, vz = *rz0 * fdot + *vz0 * gdot;
            *rx0 = rx;
            *ry0 = ry;
            *rz0 = rz;
            *vx0 = vx;
            *vy0 = vy;
            *vz0 = vz;
            *iflag = 0;
            return;
        }
    }
    drift_kepu(dt, r0, mu, alpha, u, &fp, &c1, &c2, &c3, iflag);
    if (*iflag == 0) {
        f = 1.0 - (mu / r0) * c2;
        g = dt - mu * c3;
        fdot = -((mu / (fp * r0)) * c1);
        gdot = 1.0 - (mu / fp) * c2;
        double rx = *rx0 * f + *vx0 * g, ry = *ry0 * f + *vy0 * g, rz = *rz0 * f + *vz0 * g;
        double vx = *rx0 * fdot + *vx0 * gdot, vy = *ry0 * fdot + *vy0 * gdot, vz = *rz0 * fdot + *vz0 * gdot;
        *rx0 = rx;
        *ry0 = ry;
        *rz0 = rz;
        *vx0 = vx;
        *vy0 = vy;
        *vz0 = vz;
    }
}

/* swiftest_drift_one, :111-138 -- retry as 10 substeps of dt/10, stopping at
 * the first failing substep (which leaves a partial advance). */
void swo_drift_one(double mu, double *rx, double *ry, double *rz, double *vx, double *vy, double *vz,
                   double dt, int32_t *iflag)
{
    int fl;
    drift_dan(mu, rx, ry, rz, vx, vy, vz, dt, &fl);
    if (fl != 0) {
        double dttmp = 0.1 * dt;
        for (int i = 0; i < 10; i++) {
            drift_dan(mu, rx, ry, rz, vx, vy, vz, dttmp, &fl);
            if (fl != 0) break;
        }
    }
    *iflag = fl;
}

/* swiftest_drift_all, :60-108.  GR time dilation (:84-94) uses norm2(x),
 * restated as sqrt of the squared norm (GR is off in every workload here;
 * that branch is parity unpinned). */
void swo_drift_all(const double *mu, double *x, double *v, int n, double dt, const int32_t *lmask,
                   int32_t *iflag, int lgr, double inv_c2)
{
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
        if (!lmask[i]) continue;
        double dtp = dt;
        if (lgr) {
            double rmag = sqrt(dot3(x[3 * i], x[3 * i + 1], x[3 * i + 2], x[3 * i], x[3 * i + 1], x[3 * i + 2]));
            double vmag2 = dot3(v[3 * i], v[3 * i + 1], v[3 * i + 2], v[3 * i], v[3 * i + 1], v[3 * i + 2]);
            double energy = 0.5 * vmag2 - mu[i] / rmag;
            dtp = dt * (1.0 + (3.0 * inv_c2) * energy);
        }
        swo_drift_one(mu[i], &x[3 * i], &x[3 * i + 1], &x[3 * i + 2], &v[3 * i], &v[3 * i + 1], &v[3 * i + 2], dtp,
                      &iflag[i]);
    }
}

/* ======================================================================
 * Encounter detection -- src/encounter/encounter_check.f90
 * ==================================================================== */

/* encounter_check_one, :573-621 */
void swo_encounter_check_one(double xr, double yr, double zr, double vxr, double vyr, double vzr,
                             double renc, double dt, int32_t *lencounter, int32_t *lvdotr)
{
    const double vsmall = sqrt(DBL_MIN); /* globals_module.f90:135 */
    double r2 = dot3(xr, yr, zr, xr, yr, zr);
    double r2crit = renc * renc;
    double vdotr, r2min;
    if (r2 > r2crit) {
        vdotr = dot3(vxr, vyr, vzr, xr, yr, zr);
        if (vdotr > 0.0) {
            r2min = r2;
        } else {
            double v2 = dot3(vxr, vyr, vzr, vxr, vyr, vzr);
            if (v2 <= vsmall) {
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
    *lvdotr = (vdotr < 0.0);
    *lencounter = (*lvdotr) && (r2min <= r2crit);
}

typedef struct {
    int64_t n, cap;
    int32_t *i1, *i2, *lv;
} rlist;

static void rlist_push(rlist *l, int32_t a, int32_t b, int32_t lv)
{
    if (l->n == l->cap) {
        l->cap = l->cap ? 2 * l->cap : 64;
        l->i1 = realloc(l->i1, l->cap * sizeof(int32_t));
        l->i2 = realloc(l->i2, l->cap * sizeof(int32_t));
        l->lv = realloc(l->lv, l->cap * sizeof(int32_t));
    }
    l->i1[l->n] = a;
    l->i2[l->n] = b;
    l->lv[l->n] = lv;
    l->n++;
}

/* Concatenate per-row lists in row order -- encounter_check_collapse_ragged_list, :624-673 */
static int64_t collapse(rlist *rows, int nrow, int32_t **index1, int32_t **index2, int32_t **lvdotr)
{
    int64_t n = 0;
    for (int i = 0; i < nrow; i++) n += rows[i].n;
    *index1 = malloc((n ? n : 1) * sizeof(int32_t));
    *index2 = malloc((n ? n : 1) * sizeof(int32_t));
    *lvdotr = malloc((n ? n : 1) * sizeof(int32_t));
    int64_t o = 0;
    for (int i = 0; i < nrow; i++) {
        if (rows[i].n) {
            memcpy(*index1 + o, rows[i].i1, rows[i].n * sizeof(int32_t));
            memcpy(*index2 + o, rows[i].i2, rows[i].n * sizeof(int32_t));
            memcpy(*lvdotr + o, rows[i].lv, rows[i].n * sizeof(int32_t));
            o += rows[i].n;
        }
        free(rows[i].i1);
        free(rows[i].i2);
        free(rows[i].lv);
    }
    return n;
}

/* encounter_check_all_triangular_plpl (:436-475) + triangular_one (:384-433):
 * pairs i < j, list ordered i ascending then j ascending. */
int64_t swo_encounter_check_all_plpl(int npl, const double *r, const double *v, const double *renc,
                                     double dt, int32_t **index1, int32_t **index2, int32_t **lvdotr)
{
    return swo_encounter_check_all_plplm(npl, npl, r, v, renc, dt, index1, index2, lvdotr);
}

/* encounter_check_all_plplm (:42-109) with both groups triangular
 * (triangular_plpl :436-475 over the first nplm bodies, triangular_plplm
 * :478-522 over those against the rest).  The reference concatenates the two
 * lists and re-sorts them by index1 with its quicksort (base_module.f90:
 * 1603-1700), which leaves pairs of equal index1 in an order set by the
 * partitioning; here they are listed index2 ascending (the order the single
 * triangular sweep gives, and the one the device list has).  With nplm = npl
 * this is encounter_check_all_triangular_plpl. */
int64_t swo_encounter_check_all_plplm(int npl, int nplm, const double *r, const double *v, const double *renc,
                                      double dt, int32_t **index1, int32_t **index2, int32_t **lvdotr)
{
    rlist *rows = calloc(npl > 0 ? npl : 1, sizeof(rlist));
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nplm; i++) {
        for (int j = i + 1; j < npl; j++) {
            double xr = r[3 * j] - r[3 * i], yr = r[3 * j + 1] - r[3 * i + 1], zr = r[3 * j + 2] - r[3 * i + 2];
            double vxr = v[3 * j] - v[3 * i], vyr = v[3 * j + 1] - v[3 * i + 1], vzr = v[3 * j + 2] - v[3 * i + 2];
            int32_t le, lv;
            swo_encounter_check_one(xr, yr, zr, vxr, vyr, vzr, renc[i] + renc[j], dt, &le, &lv);
            if (le) rlist_push(&rows[i], i + 1, j + 1, lv);
        }
    }
    int64_t n = collapse(rows, npl, index1, index2, lvdotr);
    free(rows);
    return n;
}

/* encounter_check_all_triangular_pltp (:525-570): every pl x tp pair, tp
 * encounter radius 0, list ordered pl ascending then tp ascending. */
int64_t swo_encounter_check_all_pltp(int npl, int ntp, const double *rpl, const double *vpl,
                                     const double *rtp, const double *vtp, const double *renc,
                                     double dt, int32_t **index1, int32_t **index2, int32_t **lvdotr)
{
    rlist *rows = calloc(npl > 0 ? npl : 1, sizeof(rlist));
#pragma omp parallel for schedule(dynamic, 4)
    for (int i = 0; i < npl; i++) {
        for (int j = 0; j < ntp; j++) {
            double xr = rtp[3 * j] - rpl[3 * i], yr = rtp[3 * j + 1] - rpl[3 * i + 1], zr = rtp[3 * j + 2] - rpl[3 * i + 2];
            double vxr = vtp[3 * j] - vpl[3 * i], vyr = vtp[3 * j + 1] - vpl[3 * i + 1], vzr = vtp[3 * j + 2] - vpl[3 * i + 2];
            int32_t le, lv;
            swo_encounter_check_one(xr, yr, zr, vxr, vyr, vzr, renc[i] + 0.0, dt, &le, &lv);
            if (le) rlist_push(&rows[i], i + 1, j + 1, lv);
        }
    }
    int64_t n = collapse(rows, npl, index1, index2, lvdotr);
    free(rows);
    return n;
}

/* ======================================================================
 * Orbital elements used by the step layer
 * ==================================================================== */

/* swiftest_orbel_xv2aeq, src/swiftest/swiftest_orbel.f90:700-764 */
void swo_orbel_xv2aeq(double mu, double rx, double ry, double rz, double vx, double vy, double vz,
                      double *a, double *e, double *q)
{
    *a = 0.0;
    *e = 0.0;
    *q = 0.0;
    double r = sqrt(dot3(rx, ry, rz, rx, ry, rz));
    double v2 = dot3(vx, vy, vz, vx, vy, vz);
    double hx = ry * vz - rz * vy, hy = rz * vx - rx * vz, hz = rx * vy - ry * vx;
    double h2 = dot3(hx, hy, hz, hx, hy, hz);
    if (h2 < DBL_MIN) return;
    double energy = 0.5 * v2 - mu / r;
    int type; /* -1 ellipse, 0 parabola, 1 hyperbola */
    double fac = 0.0;
    if (fabs((energy * r) / mu) < sqrt(TINYVALUE)) {
        type = 0;
    } else {
        *a = (-0.5 * mu) / energy;
        if (*a < 0.0) {
            fac = -h2 / (mu * (*a));
            type = (fac > TINYVALUE) ? 1 : 0;
        } else {
            type = -1;
        }
    }
    if (type == -1) {
        fac = 1.0 - h2 / (mu * (*a));
        if (fac > TINYVALUE) *e = sqrt(fac);
        *q = (*a) * (1.0 - *e);
    } else if (type == 0) {
        *a = (0.5 * h2) / mu;
        *e = 1.0;
        *q = *a;
    } else {
        *e = sqrt(1.0 + fac);
        *q = (*a) * (1.0 - *e);
    }
}

/* Semi-major axis as computed by swiftest_orbel_xv2el (:913-1072); only the
 * `a` output feeds the Hill radius (swiftest_util.f90:2260-2280). */
double swo_orbel_xv2a(double mu, double rx, double ry, double rz, double vx, double vy, double vz)
{
    double hx = ry * vz - rz * vy, hy = rz * vx - rx * vz, hz = rx * vy - ry * vx;
    double h2 = dot3(hx, hy, hz, hx, hy, hz);
    if (h2 <= 10.0 * DBL_MIN) return 0.0;
    double r = sqrt(dot3(rx, ry, rz, rx, ry, rz));
    double v2 = dot3(vx, vy, vz, vx, vy, vz);
    double energy = 0.5 * v2 - mu / r;
    if (fabs((energy * r) / mu) < sqrt(TINYVALUE)) return (0.5 * h2) / mu;
    double a = (-0.5 * mu) / energy;
    if (a < 0.0) {
        double fac = -h2 / (mu * a);
        if (!(fac > TINYVALUE)) return (0.5 * h2) / mu;
    }
    return a;
}

/* collision_check_one, src/collision/collision_check.f90:16-50 */
void swo_collision_check_one(double xr, double yr, double zr, double vxr, double vyr, double vzr,
                             double Gmtot, double rlim, double dt, int32_t lvdotr,
                             int32_t *lcollision, int32_t *lclosest)
{
    double r2 = dot3(xr, yr, zr, xr, yr, zr);
    double rlim2 = rlim * rlim;
    *lclosest = 0;
    if (r2 <= rlim2) {
        *lcollision = 1;
    } else {
        *lcollision = 0;
        double vdotr = dot3(xr, yr, zr, vxr, vyr, vzr);
        if (lvdotr && (vdotr > 0.0)) {
            double tcr2 = r2 / dot3(vxr, vyr, vzr, vxr, vyr, vzr);
            double dt2 = dt * dt;
            if (tcr2 <= dt2) {
                double a, e, q;
                swo_orbel_xv2aeq(Gmtot, xr, yr, zr, vxr, vyr, vzr, &a, &e, &q);
                *lcollision = (q < rlim);
            }
            *lclosest = !*lcollision;
        }
    }
}

/* x**n for a run-time integer exponent, as GCC's __powidf2 evaluates it
 * (binary powering), which is what `RSHELL**(2*irecl)` compiles to. */
double swo_powi(double x, int m)
{
    unsigned n = m < 0 ? -(unsigned)m : (unsigned)m;
    double y = (n % 2) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n % 2) y = y * x;
    }
    return m < 0 ? 1.0 / y : y;
}

/* x**(-1.5), correctly rounded: x^(3/2) in double-double from a corrected
 * square root, then a corrected reciprocal, then one final rounding.  Used for
 * `r2**(-1.5_DP)` (symba_kick.f90:205, :317). */
double swo_pow_m1p5(double x)
{
    double s = sqrt(x);
    double sl = fma(-s, s, x) / (2.0 * s);     /* sqrt(x) = s + sl */
    double ph = x * s;
    double pl = fma(x, s, -ph) + x * sl;       /* x^1.5 = ph + pl */
    double yh = 1.0 / ph;
    double e = fma(-yh, ph, 1.0) - yh * pl;    /* 1 - yh*(ph+pl) */
    return fma(yh, e, yh);
}

/* ======================================================================
 * Whole-system state and the step layer (WHM, HELIO, SyMBA)
 * ==================================================================== */

typedef struct {
    int64_t n, cap;
    int32_t *i1, *i2, *lvdotr, *status, *level;
} enclist;

struct swo_sys {
    int integ, npl, ntp;
    int nplm; /* fully interacting massive bodies (the first nplm after the mass sort) */
    double GMcb, Rcb;
    double cb_vb[3], ptbeg[3], ptend[3];
    double *rh, *vh, *vb, *ah, *rbeg, *rend, *xj, *vj;
    double *Gmass, *radius, *rhill, *renc, *mu, *eta, *muj, *ir3h, *ir3j;
    int32_t *status, *levelg, *levelm, *lmask, *id, *iflag;
    double *trh, *tvh, *tvb, *tah;
    int32_t *tstatus, *tlevelg, *tlevelm, *tlmask, *tid, *tiflag;
    enclist plpl, pltp;
    int lfirst, tlfirst, lfirstkick;
    int irec;
    int err;
    int col_finish, col_seen; /* swo_sys_set_collision_mode; a collision flagged in this step */
    /* swo_sys_set_merge: collisions resolved by merging inside the step */
    int merge;
    double GU;     /* G in the system's units: Gmass = GU * mass */
    double *mass;  /* body masses (Gmass / GU unless given) */
    int32_t maxid; /* largest body id handed out (nbody_system%maxid) */
    int64_t nmerge;
    /* pericentre-passage state of the discard checks */
    int32_t *pisperi, *tisperi;
    double *patp, *pperi, *tatp, *tperi;
    int lfirst_peri, dis_ready;
    swo_params p;
    swo_stats st;
    int64_t iloop, iout;
};

static void list_reset(enclist *l) { l->n = 0; }
static void list_resize(enclist *l, int64_t n)
{
    if (n > l->cap) {
        l->cap = n;
        l->i1 = realloc(l->i1, n * sizeof(int32_t));
        l->i2 = realloc(l->i2, n * sizeof(int32_t));
        l->lvdotr = realloc(l->lvdotr, n * sizeof(int32_t));
        l->status = realloc(l->status, n * sizeof(int32_t));
        l->level = realloc(l->level, n * sizeof(int32_t));
    }
    l->n = n;
}

static double *dalloc(int64_t n) { return calloc(n > 0 ? n : 1, sizeof(double)); }
static int32_t *ialloc(int64_t n) { return calloc(n > 0 ? n : 1, sizeof(int32_t)); }

/* swiftest_util_set_rhill, swiftest_util.f90:2260-2280 (a from xv2el with
 * mu = GM_cb + Gm, from rh and vh) */
static void set_rhill(swo_sys *s)
{
    if (s->GMcb <= DBL_MIN) return;
    for (int i = 0; i < s->npl; i++) {
        double mu = s->GMcb + s->Gmass[i];
        double *r = &s->rh[3 * i], *v = &s->vh[3 * i];
        double a = swo_orbel_xv2a(mu, r[0], r[1], r[2], v[0], v[1], v[2]);
        double fac = pow((s->Gmass[i] / s->GMcb) / 3.0, THIRD);
        if (a > 0.0)
            s->rhill[i] = a * fac;
        else
            s->rhill[i] = sqrt(dot3(r[0], r[1], r[2], r[0], r[1], r[2])) * fac;
    }
}

/* symba_util_set_renc, symba_util.f90:245-267 */
static void set_renc(swo_sys *s, int scale)
{
    double rshell_irec = 1.0;
    for (int i = 1; i <= scale; i++) rshell_irec = rshell_irec * RSHELL;
    for (int i = 0; i < s->npl; i++) s->renc[i] = (s->rhill[i] * RHSCALE) * rshell_irec;
}

/* swiftest_util_coord_vh2vb_pl, swiftest_util.f90:424-459 */
void swo_coord_vh2vb(int npl, double GMcb, const double *Gmass, const double *vh, double *vb, double *cb_vb)
{
    if (npl == 0) return;
    double sm = 0.0;
    for (int i = 0; i < npl; i++) sm = sm + Gmass[i];
    double Gmtot = GMcb + sm;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < npl; i++)
        for (int d = 0; d < 3; d++) c[d] = c[d] - Gmass[i] * vh[3 * i + d];
    for (int d = 0; d < 3; d++) cb_vb[d] = c[d] / Gmtot;
    for (int i = 0; i < npl; i++)
        for (int d = 0; d < 3; d++) vb[3 * i + d] = vh[3 * i + d] + cb_vb[d];
}

static void vh2vb_pl(swo_sys *s) { swo_coord_vh2vb(s->npl, s->GMcb, s->Gmass, s->vh, s->vb, s->cb_vb); }

/* swiftest_util_coord_vb2vh_pl, swiftest_util.f90:363-395 (reverse order) */
static void vb2vh_pl(swo_sys *s)
{
    if (s->npl == 0) return;
    double c[3] = {0.0, 0.0, 0.0};
    for (int i = s->npl - 1; i >= 0; i--) {
        if (s->status[i] == SWO_INACTIVE) continue;
        for (int d = 0; d < 3; d++) c[d] = c[d] - (s->Gmass[i] * s->vb[3 * i + d]) / s->GMcb;
    }
    for (int d = 0; d < 3; d++) s->cb_vb[d] = c[d];
    for (int i = 0; i < s->npl; i++)
        for (int d = 0; d < 3; d++) s->vh[3 * i + d] = s->vb[3 * i + d] - s->cb_vb[d];
}

/* swiftest_util_coord_vh2vb_tp / vb2vh_tp, swiftest_util.f90:398-485 */
static void vh2vb_tp(swo_sys *s, const double *vbcb)
{
    for (int i = 0; i < s->ntp; i++)
        if (s->tlmask[i])
            for (int d = 0; d < 3; d++) s->tvb[3 * i + d] = s->tvh[3 * i + d] + vbcb[d];
}
static void vb2vh_tp(swo_sys *s, const double *vbcb)
{
    for (int i = 0; i < s->ntp; i++)
        if (s->tlmask[i])
            for (int d = 0; d < 3; d++) s->tvh[3 * i + d] = s->tvb[3 * i + d] - vbcb[d];
}

/* helio_drift_linear_pl, helio_drift.f90:129-163 */
static void lindrift_pl(swo_sys *s, double dt, int lbeg)
{
    if (s->npl == 0) return;
    double pt[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < 3; d++) {
        double sm = 0.0;
        for (int i = 0; i < s->npl; i++)
            if (s->lmask[i]) sm = sm + s->Gmass[i] * s->vb[3 * i + d];
        pt[d] = sm;
    }
    for (int d = 0; d < 3; d++) pt[d] = pt[d] / s->GMcb;
    for (int i = 0; i < s->npl; i++)
        if (s->lmask[i])
            for (int d = 0; d < 3; d++) s->rh[3 * i + d] = fma(pt[d], dt, s->rh[3 * i + d]);
    memcpy(lbeg ? s->ptbeg : s->ptend, pt, sizeof pt);
}

/* helio_drift_linear_tp, helio_drift.f90:166-200 */
static void lindrift_tp(swo_sys *s, double dt, int lbeg)
{
    const double *pt = lbeg ? s->ptbeg : s->ptend;
    for (int i = 0; i < s->ntp; i++)
        if (s->tlmask[i])
            for (int d = 0; d < 3; d++) s->trh[3 * i + d] = fma(pt[d], dt, s->trh[3 * i + d]);
}

/* symba_kick_getacch_pl (symba_kick.f90:36-76): full rows, then remove the
 * listed encounter pairs with the flat kernel.  For HELIO
 * (swiftest_kick_getacch_int_pl, swiftest_kick.f90:12-43) the radius test
 * follows CHK_CLOSE and there is no pair list. */
static void getacch_pl_helio(swo_sys *s)
{
    const double *rad = (s->integ == SWO_SYMBA || s->p.lclose) ? s->radius : NULL;
    swo_kick_getacch_int_all_tri(s->npl, s->nplm, s->rh, s->Gmass, rad, s->ah);
    if (s->integ == SWO_SYMBA && s->plpl.n > 0) {
        double *ah_enc = dalloc(3 * (int64_t)s->npl);
        int32_t *k = ialloc(2 * s->plpl.n);
        for (int64_t q = 0; q < s->plpl.n; q++) {
            k[2 * q] = s->plpl.i1[q];
            k[2 * q + 1] = s->plpl.i2[q];
        }
        swo_kick_getacch_int_all_flat(s->npl, s->plpl.n, k, s->rh, s->Gmass, s->radius, ah_enc);
        for (int64_t q = 0; q < 3 * (int64_t)s->npl; q++) s->ah[q] = s->ah[q] - ah_enc[q];
        free(ah_enc);
        free(k);
    }
}

/* helio_kick_vb_pl, helio_kick.f90:91-132 */
static void kick_vb_pl(swo_sys *s, double dt, int lbeg)
{
    if (s->npl == 0) return;
    memset(s->ah, 0, sizeof(double) * 3 * s->npl);
    getacch_pl_helio(s);
    memcpy(lbeg ? s->rbeg : s->rend, s->rh, sizeof(double) * 3 * s->npl);
    for (int i = 0; i < s->npl; i++)
        if (s->lmask[i])
            for (int d = 0; d < 3; d++) s->vb[3 * i + d] = fma(s->ah[3 * i + d], dt, s->vb[3 * i + d]);
}

/* helio_kick_vb_tp (helio_kick.f90:135-169) via symba_kick_getacch_tp
 * (symba_kick.f90:79-123): pl positions are rbeg / rend, then the listed
 * pl-tp encounter pairs are added back in list order. */
static void kick_vb_tp(swo_sys *s, double dt, int lbeg)
{
    if (s->ntp == 0) return;
    memset(s->tah, 0, sizeof(double) * 3 * s->ntp);
    const double *rp = lbeg ? s->rbeg : s->rend;
    if (s->npl > 0) swo_kick_getacch_int_all_tp(s->ntp, s->npl, s->trh, rp, s->Gmass, s->tlmask, s->tah);
    if (s->integ == SWO_SYMBA) {
        for (int64_t k = 0; k < s->pltp.n; k++) {
            int i = s->pltp.i1[k] - 1, j = s->pltp.i2[k] - 1;
            if (!s->tlmask[j]) continue;
            double dx = s->trh[3 * j] - rp[3 * i], dy = s->trh[3 * j + 1] - rp[3 * i + 1], dz = s->trh[3 * j + 2] - rp[3 * i + 2];
            double rjj = dot3(dx, dy, dz, dx, dy, dz);
            double fac = s->Gmass[i] / (rjj * sqrt(rjj));
            s->tah[3 * j] = fma(fac, dx, s->tah[3 * j]);
            s->tah[3 * j + 1] = fma(fac, dy, s->tah[3 * j + 1]);
            s->tah[3 * j + 2] = fma(fac, dz, s->tah[3 * j + 2]);
        }
    }
    for (int i = 0; i < s->ntp; i++)
        if (s->tlmask[i])
            for (int d = 0; d < 3; d++) s->tvb[3 * i + d] = fma(s->tah[3 * i + d], dt, s->tvb[3 * i + d]);
}

/* symba_drift_pl / symba_drift_tp (symba_drift.f90:14-61) over
 * helio_drift_body (helio_drift.f90:14-54): mu = GM_cb, (rh, vb), bodies at
 * the current recursion level only. */
static void drift_pl(swo_sys *s, double dt)
{
    if (s->npl == 0) return;
    double *mu = dalloc(s->npl);
    for (int i = 0; i < s->npl; i++) {
        mu[i] = s->GMcb;
        s->iflag[i] = 0;
        s->lmask[i] = (s->status[i] != SWO_INACTIVE) && (s->integ != SWO_SYMBA || s->levelg[i] == s->irec);
    }
    swo_drift_all(mu, s->rh, s->vb, s->npl, dt, s->lmask, s->iflag, 0, 0.0);
    for (int i = 0; i < s->npl; i++) {
        if (s->iflag[i] != 0) {
            s->status[i] = SWO_DISCARDED_DRIFTERR;
            s->st.ndrifterr++;
            s->err = -2;
        }
        s->lmask[i] = (s->status[i] != SWO_INACTIVE);
    }
    free(mu);
}

static void drift_tp(swo_sys *s, double dt)
{
    if (s->ntp == 0) return;
    double *mu = dalloc(s->ntp);
    for (int i = 0; i < s->ntp; i++) {
        mu[i] = s->GMcb;
        s->tiflag[i] = 0;
        s->tlmask[i] = (s->tstatus[i] == SWO_ACTIVE) && (s->integ != SWO_SYMBA || s->tlevelg[i] == s->irec);
    }
    swo_drift_all(mu, s->trh, s->tvb, s->ntp, dt, s->tlmask, s->tiflag, 0, 0.0);
    for (int i = 0; i < s->ntp; i++) {
        if (s->tiflag[i] != 0) {
            s->tstatus[i] = SWO_DISCARDED_DRIFTERR;
            s->st.ndrifterr++;
        }
        s->tlmask[i] = (s->tstatus[i] == SWO_ACTIVE);
    }
    free(mu);
}

/* helio_step_pl (helio_step.f90:37-76) and helio_step_tp (:79-119) */
static void helio_step(swo_sys *s)
{
    double dt = s->p.dt, dth = 0.5 * dt;
    s->tlfirst = s->lfirst; /* whm_step_system, whm_step.f90:26 */
    if (s->npl > 0) {
        if (s->lfirst) {
            vh2vb_pl(s);
            s->lfirst = 0;
        }
        lindrift_pl(s, dth, 1);
        kick_vb_pl(s, dth, 1);
        drift_pl(s, dt);
        kick_vb_pl(s, dth, 0);
        lindrift_pl(s, dth, 0);
        vb2vh_pl(s);
    }
    if (s->ntp > 0) {
        if (s->tlfirst) {
            double vbcb[3] = {-s->ptbeg[0], -s->ptbeg[1], -s->ptbeg[2]};
            vh2vb_tp(s, vbcb);
            s->tlfirst = 0;
        }
        lindrift_tp(s, dth, 1);
        kick_vb_tp(s, dth, 1);
        drift_tp(s, dt);
        kick_vb_tp(s, dth, 0);
        lindrift_tp(s, dth, 0);
        double vbcb[3] = {-s->ptend[0], -s->ptend[1], -s->ptend[2]};
        vb2vh_tp(s, vbcb);
    }
}

/* symba_step_reset_system, symba_step.f90:265-331 */
static void symba_reset(swo_sys *s)
{
    list_reset(&s->plpl);
    list_reset(&s->pltp);
    for (int i = 0; i < s->npl; i++) {
        s->levelg[i] = -1;
        s->levelm[i] = -1;
        s->lmask[i] = 1;
    }
    if (s->npl > 0) set_renc(s, 0);
    for (int i = 0; i < s->ntp; i++) {
        s->tlevelg[i] = -1;
        s->tlevelm[i] = -1;
        s->tlmask[i] = (s->tstatus[i] == SWO_ACTIVE);
    }
    s->tlfirst = s->lfirstkick;
    s->lfirst = s->lfirstkick;
}

/* symba_encounter_check_pl, symba_encounter_check.f90:14-87 */
static int symba_check_pl(swo_sys *s, double dt, int irec)
{
    if (s->npl == 0) return 0;
    set_renc(s, irec);
    int32_t *i1, *i2, *lv;
    /* symba_encounter_check.f90:44-50: plplm when some bodies are below GMTINY */
    int64_t n = swo_encounter_check_all_plplm(s->npl, s->nplm, s->rh, s->vb, s->renc, dt, &i1, &i2, &lv);
    list_resize(&s->plpl, n);
    for (int64_t k = 0; k < n; k++) {
        s->plpl.i1[k] = i1[k];
        s->plpl.i2[k] = i2[k];
        s->plpl.lvdotr[k] = lv[k];
        s->plpl.status[k] = SWO_ACTIVE;
        s->plpl.level[k] = irec;
        int i = i1[k] - 1, j = i2[k] - 1;
        s->levelg[i] = irec;
        s->levelm[i] = irec;
        s->levelg[j] = irec;
        s->levelm[j] = irec;
    }
    free(i1);
    free(i2);
    free(lv);
    s->st.nplpl_enc = n;
    return n > 0;
}

/* symba_encounter_check_tp, symba_encounter_check.f90:238-296 (tp whose
 * status is not ACTIVE are treated as removed) */
static int symba_check_tp(swo_sys *s, double dt, int irec)
{
    if (s->ntp == 0) return 0;
    set_renc(s, irec);
    int32_t *i1, *i2, *lv;
    double *rt = dalloc(3 * (int64_t)s->ntp), *vt = dalloc(3 * (int64_t)s->ntp);
    const double far = 1e300; /* removed particles never register */
    for (int j = 0; j < s->ntp; j++)
        for (int d = 0; d < 3; d++) {
            rt[3 * j + d] = s->tstatus[j] == SWO_ACTIVE ? s->trh[3 * j + d] : far;
            vt[3 * j + d] = s->tstatus[j] == SWO_ACTIVE ? s->tvb[3 * j + d] : 0.0;
        }
    int64_t n = swo_encounter_check_all_pltp(s->npl, s->ntp, s->rh, s->vb, rt, vt, s->renc, dt, &i1, &i2, &lv);
    free(rt);
    free(vt);
    list_resize(&s->pltp, n);
    for (int64_t k = 0; k < n; k++) {
        s->pltp.i1[k] = i1[k];
        s->pltp.i2[k] = i2[k];
        s->pltp.lvdotr[k] = lv[k];
        s->pltp.status[k] = SWO_ACTIVE;
        s->pltp.level[k] = irec;
        int i = i1[k] - 1, j = i2[k] - 1;
        s->levelg[i] = irec;
        s->levelm[i] = irec;
        s->tlevelg[j] = irec;
        s->tlevelm[j] = irec;
    }
    free(i1);
    free(i2);
    free(lv);
    s->st.npltp_enc = n;
    return n > 0;
}

/* symba_encounter_check_list_plpl, symba_encounter_check.f90:90-160 */
static int list_check_plpl(swo_sys *s, double dt, int irec)
{
    enclist *l = &s->plpl;
    if (l->n == 0) return 0;
    int64_t nm = 0;
    for (int64_t k = 0; k < l->n; k++)
        if (l->status[k] == SWO_ACTIVE && l->level[k] == irec - 1) nm++;
    if (nm == 0) return 0;
    set_renc(s, irec);
    char *flag = calloc(l->n, 1);
    int any = 0;
    for (int64_t k = 0; k < l->n; k++) {
        if (!(l->status[k] == SWO_ACTIVE && l->level[k] == irec - 1)) continue;
        int i = l->i1[k] - 1, j = l->i2[k] - 1;
        double xr[3], vr[3];
        for (int d = 0; d < 3; d++) {
            xr[d] = s->rh[3 * j + d] - s->rh[3 * i + d];
            vr[d] = s->vb[3 * j + d] - s->vb[3 * i + d];
        }
        int32_t le;
        swo_encounter_check_one(xr[0], xr[1], xr[2], vr[0], vr[1], vr[2], s->renc[i] + s->renc[j], dt, &le, &l->lvdotr[k]);
        if (le) {
            double sr = s->radius[i] + s->radius[j];
            double rji2 = dot3(xr[0], xr[1], xr[2], xr[0], xr[1], xr[2]);
            le = rji2 > sr * sr;
        }
        flag[k] = (char)le;
        any |= le;
    }
    if (any) {
        for (int64_t k = 0; k < l->n; k++) {
            if (!flag[k]) continue;
            int i = l->i1[k] - 1, j = l->i2[k] - 1;
            s->levelg[i] = irec;
            if (irec > s->levelm[i]) s->levelm[i] = irec;
            s->levelg[j] = irec;
            if (irec > s->levelm[j]) s->levelm[j] = irec;
            l->level[k] = irec;
        }
    }
    free(flag);
    return any;
}

/* symba_encounter_check_list_pltp, symba_encounter_check.f90:163-235 */
static int list_check_pltp(swo_sys *s, double dt, int irec)
{
    enclist *l = &s->pltp;
    if (l->n == 0) return 0;
    int64_t nm = 0;
    for (int64_t k = 0; k < l->n; k++)
        if (l->status[k] == SWO_ACTIVE && l->level[k] == irec - 1) nm++;
    if (nm == 0) return 0;
    set_renc(s, irec);
    char *flag = calloc(l->n, 1);
    int any = 0;
    for (int64_t k = 0; k < l->n; k++) {
        if (!(l->status[k] == SWO_ACTIVE && l->level[k] == irec - 1)) continue;
        int i = l->i1[k] - 1, j = l->i2[k] - 1;
        double xr[3], vr[3];
        for (int d = 0; d < 3; d++) {
            xr[d] = s->trh[3 * j + d] - s->rh[3 * i + d];
            vr[d] = s->tvb[3 * j + d] - s->vb[3 * i + d];
        }
        int32_t le;
        swo_encounter_check_one(xr[0], xr[1], xr[2], vr[0], vr[1], vr[2], s->renc[i], dt, &le, &l->lvdotr[k]);
        if (le) {
            double rji2 = dot3(xr[0], xr[1], xr[2], xr[0], xr[1], xr[2]);
            le = rji2 > s->radius[i] * s->radius[i];
        }
        flag[k] = (char)le;
        any |= le;
    }
    if (any) {
        for (int64_t k = 0; k < l->n; k++) {
            if (!flag[k]) continue;
            int i = l->i1[k] - 1, j = l->i2[k] - 1;
            s->levelg[i] = irec;
            if (irec > s->levelm[i]) s->levelm[i] = irec;
            s->tlevelg[j] = irec;
            if (irec > s->tlevelm[j]) s->tlevelm[j] = irec;
            l->level[k] = irec;
        }
    }
    free(flag);
    return any;
}

/* Planted mistakes for the tests that check the oracle's own pins
 * (tests/test_oracle_symba_pins.py builds this file with -DSWO_MUTANT=k into
 * a temporary directory and expects a pin to fail): 1 drops the shell taper
 * polynomial, 2 drops the -1 shell kicks of the recursion.  The oracle the
 * tests compare against is always built with SWO_MUTANT = 0. */
#ifndef SWO_MUTANT
#define SWO_MUTANT 0
#endif

/* shell factor of symba_kick_list_plpl / _pltp (symba_kick.f90:187-208,
 * :299-318); returns 0 when r2 < rim1 (pair dropped) */
static int shell_fac(double ri, double r2, double *fac)
{
    double rim1 = ri * (RSHELL * RSHELL);
    if (r2 < rim1) return 0;
    if (r2 < ri) {
#if SWO_MUTANT == 1
        *fac = swo_pow_m1p5(r2);
#else
        double ris = sqrt(ri);
        double r = sqrt(r2);
        double rr = (ris - r) / (ris * (1.0 - RSHELL));
        *fac = swo_pow_m1p5(r2) * ((1.0 - 3.0 * (rr * rr)) + 2.0 * ((rr * rr) * rr));
#endif
    } else {
        *fac = 1.0 / (r2 * sqrt(r2));
    }
    return 1;
}

int swo_shell_fac(double ri, double r2, double *fac) { return shell_fac(ri, r2, fac); }

/* symba_kick_list_plpl, symba_kick.f90:126-232 */
static void kick_list_plpl(swo_sys *s, double dt, int irec, int sgn)
{
    enclist *l = &s->plpl;
    if (l->n == 0 || s->npl == 0) return;
    for (int i = 0; i < s->npl; i++) s->lmask[i] = s->status[i] != SWO_INACTIVE;
    int irm1 = irec - 1;
    int irecl = sgn < 0 ? irec - 1 : irec;
    char *good = calloc(l->n, 1);
    int64_t ngood = 0;
    for (int64_t k = 0; k < l->n; k++) {
        int i = l->i1[k] - 1, j = l->i2[k] - 1;
        good[k] = (s->levelg[i] >= irm1) && (s->levelg[j] >= irm1) && (l->status[k] == SWO_ACTIVE);
        ngood += good[k];
    }
    if (ngood > 0) {
        const double shellpow = swo_powi(RSHELL, 2 * irecl);
        for (int64_t k = 0; k < l->n; k++) {
            if (!good[k]) continue;
            int i = l->i1[k] - 1, j = l->i2[k] - 1;
            memset(&s->ah[3 * i], 0, 3 * sizeof(double));
            memset(&s->ah[3 * j], 0, 3 * sizeof(double));
        }
        for (int64_t k = 0; k < l->n; k++) {
            if (!good[k]) continue;
            int i = l->i1[k] - 1, j = l->i2[k] - 1;
            double sh = s->rhill[i] + s->rhill[j];
            double ri = ((sh * sh) * (RHSCALE * RHSCALE)) * shellpow;
            double dx[3];
            for (int d = 0; d < 3; d++) dx[d] = s->rh[3 * j + d] - s->rh[3 * i + d];
            double r2 = dot3(dx[0], dx[1], dx[2], dx[0], dx[1], dx[2]);
            double fac;
            if (!shell_fac(ri, r2, &fac)) {
                good[k] = 0;
                continue;
            }
            double faci = fac * s->Gmass[i], facj = fac * s->Gmass[j];
            for (int d = 0; d < 3; d++) {
                s->ah[3 * i + d] = fma(facj, dx[d], s->ah[3 * i + d]);
                s->ah[3 * j + d] = fma(-faci, dx[d], s->ah[3 * j + d]);
            }
        }
        const double sdt = (double)sgn * dt;
        for (int64_t k = 0; k < l->n; k++) {
            if (!good[k]) continue;
            int i = l->i1[k] - 1, j = l->i2[k] - 1;
            for (int d = 0; d < 3; d++) {
                s->vb[3 * i + d] = fma(sdt, s->ah[3 * i + d], s->vb[3 * i + d]);
                s->vb[3 * j + d] = fma(sdt, s->ah[3 * j + d], s->vb[3 * j + d]);
            }
            memset(&s->ah[3 * i], 0, 3 * sizeof(double));
            memset(&s->ah[3 * j], 0, 3 * sizeof(double));
        }
    }
    free(good);
}

/* symba_kick_list_pltp, symba_kick.f90:235-340 */
static void kick_list_pltp(swo_sys *s, double dt, int irec, int sgn)
{
    enclist *l = &s->pltp;
    if (l->n == 0 || s->npl == 0 || s->ntp == 0) return;
    for (int i = 0; i < s->npl; i++) s->lmask[i] = s->status[i] != SWO_INACTIVE;
    for (int i = 0; i < s->ntp; i++) s->tlmask[i] = s->tstatus[i] == SWO_ACTIVE;
    int irm1 = irec - 1;
    int irecl = sgn < 0 ? irec - 1 : irec;
    char *good = calloc(l->n, 1);
    int64_t ngood = 0;
    for (int64_t k = 0; k < l->n; k++) {
        int i = l->i1[k] - 1, j = l->i2[k] - 1;
        good[k] = (s->levelg[i] >= irm1) && (s->tlevelg[j] >= irm1) && (l->status[k] == SWO_ACTIVE);
        ngood += good[k];
    }
    if (ngood > 0) {
        const double shellpow = swo_powi(RSHELL, 2 * irecl);
        for (int64_t k = 0; k < l->n; k++)
            if (good[k]) memset(&s->tah[3 * (l->i2[k] - 1)], 0, 3 * sizeof(double));
        for (int64_t k = 0; k < l->n; k++) {
            if (!good[k]) continue;
            int i = l->i1[k] - 1, j = l->i2[k] - 1;
            double ri = ((s->rhill[i] * s->rhill[i]) * (RHSCALE * RHSCALE)) * shellpow;
            double dx[3];
            for (int d = 0; d < 3; d++) dx[d] = s->trh[3 * j + d] - s->rh[3 * i + d];
            double r2 = dot3(dx[0], dx[1], dx[2], dx[0], dx[1], dx[2]);
            double fac;
            if (!shell_fac(ri, r2, &fac)) {
                good[k] = 0;
                continue;
            }
            double faci = fac * s->Gmass[i];
            for (int d = 0; d < 3; d++) s->tah[3 * j + d] = fma(-faci, dx[d], s->tah[3 * j + d]);
        }
        const double sdt = (double)sgn * dt;
        for (int64_t k = 0; k < l->n; k++) {
            if (!good[k]) continue;
            int j = l->i2[k] - 1;
            for (int d = 0; d < 3; d++) s->tvb[3 * j + d] = fma(sdt, s->tah[3 * j + d], s->tvb[3 * j + d]);
            memset(&s->tah[3 * j], 0, 3 * sizeof(double));
        }
    }
    free(good);
}

/* collision_check_plpl / _pltp (collision_check.f90:53-~230) -- detection
 * only; the merge / discard that follows is host work not modelled here. */
static int collision_check(swo_sys *s, double dt, int irec)
{
    int any = 0;
    enclist *l = &s->plpl;
    for (int64_t k = 0; k < l->n; k++) {
        int i = l->i1[k] - 1, j = l->i2[k] - 1;
        if (!(l->status[k] == SWO_ACTIVE && s->levelg[i] >= irec)) continue;
        double xr[3], vr[3];
        for (int d = 0; d < 3; d++) {
            xr[d] = s->rh[3 * i + d] - s->rh[3 * j + d];
            vr[d] = s->vb[3 * i + d] - s->vb[3 * j + d];
        }
        int32_t lc, lcl;
        swo_collision_check_one(xr[0], xr[1], xr[2], vr[0], vr[1], vr[2], s->Gmass[i] + s->Gmass[j],
                                s->radius[i] + s->radius[j], dt, l->lvdotr[k], &lc, &lcl);
        if (lc) {
            l->status[k] = SWO_COLLIDED;
            any = 1;
        }
    }
    l = &s->pltp;
    for (int64_t k = 0; k < l->n; k++) {
        int i = l->i1[k] - 1, j = l->i2[k] - 1;
        if (!(l->status[k] == SWO_ACTIVE && s->tlevelg[j] >= irec)) continue;
        double xr[3], vr[3];
        for (int d = 0; d < 3; d++) {
            xr[d] = s->rh[3 * i + d] - s->trh[3 * j + d];
            vr[d] = s->vb[3 * i + d] - s->tvb[3 * j + d];
        }
        int32_t lc, lcl;
        swo_collision_check_one(xr[0], xr[1], xr[2], vr[0], vr[1], vr[2], s->Gmass[i], s->radius[i], dt,
                                l->lvdotr[k], &lc, &lcl);
        if (lc) {
            l->status[k] = SWO_COLLIDED;
            any = 1;
        }
    }
    if (any) s->st.ncollision++;
    return any;
}

/* symba_step_set_recur_levels_system, symba_step.f90:118-160 */
static void set_recur_levels(swo_sys *s, int ireci)
{
    int irecp = ireci + 1;
    for (int i = 0; i < s->npl; i++)
        if (s->levelg[i] == irecp) s->levelg[i] = ireci;
    for (int i = 0; i < s->ntp; i++)
        if (s->tlevelg[i] == irecp) s->tlevelg[i] = ireci;
    for (int64_t k = 0; k < s->plpl.n; k++)
        if (s->plpl.level[k] == irecp) s->plpl.level[k] = ireci;
    for (int64_t k = 0; k < s->pltp.n; k++)
        if (s->pltp.level[k] == irecp) s->pltp.level[k] = ireci;
    s->irec = ireci;
}

/* ---- collisions resolved by merging (collision model MERGE) ----
 *
 * collision_resolve_plpl (collision_resolve.f90:521-662) with
 * collision_generate_merge (collision_generate.f90:158-259),
 * collision_resolve_mergeaddsub (collision_resolve.f90:326-520) and
 * swiftest_util_rearray_pl (swiftest_util.f90:1638-1811), for massive bodies
 * without rotation, without test particles and without the ENERGY
 * bookkeeping (whose extra torque, collision_resolve.f90:624-641, is not
 * applied).  A body in two colliding pairs of one check (a collision family,
 * collision_resolve_make_impactors_pl) is refused with -4. */

static void v3_cross(const double *a, const double *b, double *c) /* operator_cross_dp */
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double v3_mag(const double *a) { return swo_norm2(a[0], a[1], a[2]); } /* .mag. = norm2 */

static void v3_unit(const double *a, double *u) /* operator_unit_dp */
{
    double m = v3_mag(a);
    for (int d = 0; d < 3; d++) u[d] = m > DBL_MIN ? a[d] / m : 0.0;
}

/* r**3 as gfortran expands an integer power */
static double cube(double r) { return (r * r) * r; }

/* swiftest_util_coord_rh2rb_pl, swiftest_util.f90:488-521 (cb%rb only) */
static void rh2rb_cb(const swo_sys *s, double *cb_rb)
{
    double Gmtot = s->GMcb, x[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < s->npl; i++) {
        if (s->status[i] == SWO_INACTIVE) continue;
        Gmtot = Gmtot + s->Gmass[i];
        for (int d = 0; d < 3; d++) x[d] = x[d] + s->Gmass[i] * s->rh[3 * i + d];
    }
    for (int d = 0; d < 3; d++) cb_rb[d] = -x[d] / Gmtot;
}

/* collision_util_velocity_torque, collision_util.f90:1001-1029 */
static void velocity_torque(const double *dL, double mass, const double *r, double *v)
{
    double dLu[3], ru[3], q[3], c[3];
    v3_unit(dL, dLu);
    v3_unit(r, ru);
    double rmag = v3_mag(r), vmag = v3_mag(v);
    for (int d = 0; d < 3; d++) q[d] = dL[d] / (rmag * mass);
    double vapply = v3_mag(q);
    if (vapply / vmag > DBL_EPSILON) {
        v3_cross(dLu, ru, c);
        for (int d = 0; d < 3; d++) v[d] = v[d] + vapply * c[d];
    }
}

typedef struct {
    double rh[3], vh[3], vb[3], mass, Gmass, radius;
    int32_t id, levelg, levelm;
} newbody;

/* One merge, as a function of the two impactors alone: consolidate_impactors
 * (collision_resolve.f90:15-177, no children), set_coordinate_impactors
 * (collision_util.f90:609-670), the merged body of generate_merge
 * (collision_generate.f90:199-228) and its coordinates in mergeaddsub
 * (collision_resolve.f90:356-380).  m, Gm, R: 2 each; rh, vb: 2 x 3;
 * cb_rb, cb_vb: the central body's barycentric position and velocity.
 * out: rh[3] vh[3] vb[3] mass Gmass radius.  Returns the index (0 or 1) of
 * the larger impactor, whose levels and metadata the new body takes. */
int swo_merge_bodies(const double *m, const double *Gm, const double *R_in, const double *rh, const double *vb_in,
                     const double *cb_rb, const double *cb_vb, double GU, double *out)
{
    double R[2], vol[2], rb[2][3], vb[2][3];
    for (int k = 0; k < 2; k++) {
        vol[k] = ((4.0 / 3.0) * PI) * cube(R_in[k]);
        for (int d = 0; d < 3; d++) {
            rb[k][d] = rh[3 * k + d] + cb_rb[d];
            vb[k][d] = vb_in[3 * k + d];
        }
        R[k] = pow((3.0 * vol[k]) / (4.0 * PI), 1.0 / 3.0);
    }
    double mtot = m[0] + m[1];
    /* shift the impactors apart when they overlap (:142-157) */
    double rlim = R[0] + R[1], rrel[3];
    for (int d = 0; d < 3; d++) rrel[d] = rb[1][d] - rb[0][d];
    double rrel_mag = v3_mag(rrel);
    if (rrel_mag < rlim) {
        double vrel[3], vu[3], dr[3];
        for (int d = 0; d < 3; d++) vrel[d] = vb[1][d] - vb[0][d];
        v3_unit(vrel, vu);
        double vdotr = (vu[0] * rrel[0] + vu[1] * rrel[1]) + vu[2] * rrel[2];
        double root = sqrt(((rlim * rlim) - (rrel_mag * rrel_mag)) + (vdotr * vdotr));
        double t = vdotr - copysign(1.0, vdotr) * root;
        for (int d = 0; d < 3; d++) {
            dr[d] = (-t) * vu[d];
            dr[d] = (1.0 + 2.0 * DBL_EPSILON) * dr[d];
        }
        for (int d = 0; d < 3; d++) {
            rb[0][d] = rb[0][d] - (dr[d] * m[1]) / mtot;
            rb[1][d] = rb[1][d] + (dr[d] * m[0]) / mtot;
        }
    }
    /* barycentre of the pair and the orbital angular momentum about it */
    double rbcom[3], vbcom[3], L[3], Lk[2][3];
    for (int d = 0; d < 3; d++) {
        rbcom[d] = (m[0] * rb[0][d] + m[1] * rb[1][d]) / mtot;
        vbcom[d] = (m[0] * vb[0][d] + m[1] * vb[1][d]) / mtot;
    }
    for (int k = 0; k < 2; k++) {
        double rc[3], vc[3];
        for (int d = 0; d < 3; d++) {
            rc[d] = rb[k][d] - rbcom[d];
            vc[d] = vb[k][d] - vbcom[d];
        }
        Lk[k][0] = m[k] * (rc[1] * vc[2] - rc[2] * vc[1]);
        Lk[k][1] = m[k] * (rc[2] * vc[0] - rc[0] * vc[2]);
        Lk[k][2] = m[k] * (rc[0] * vc[1] - rc[1] * vc[0]);
    }
    for (int d = 0; d < 3; d++) L[d] = Lk[0][d] + Lk[1][d];
    /* the merged body: the lost orbital angular momentum (before minus the
     * after phase's zero) goes back in as a velocity torque */
    const double mass = m[0] + m[1];
    const double volume = ((4.0 / 3.0) * PI) * (cube(R[0]) + cube(R[1]));
    double v[3] = {vbcom[0], vbcom[1], vbcom[2]}, dL[3];
    for (int d = 0; d < 3; d++) dL[d] = -(0.0 - L[d]);
    velocity_torque(dL, mass, rbcom, v);
    for (int d = 0; d < 3; d++) {
        out[d] = rbcom[d] - cb_rb[d];
        out[3 + d] = v[d] - cb_vb[d];
        out[6 + d] = v[d];
    }
    out[9] = mass;
    out[10] = GU * mass;
    out[11] = pow((3.0 * volume) / (4.0 * PI), THIRD);
    return Gm[1] > Gm[0] ? 1 : 0; /* maxloc: the first of equal masses */
}

/* One merge of bodies i and j (0-based) inside the step: the new body, the
 * encounter entries of the smaller impactor renamed to the larger
 * (collision_generate.f90:232-246), the next id (mergeaddsub :447-449). */
static void merge_pair(swo_sys *s, int i, int j, const double *cb_rb, int32_t *nid1, int32_t *nid2, newbody *nb)
{
    const int idx[2] = {i, j};
    double m[2], Gm[2], R[2], rh[6], vb[6], out[12];
    for (int k = 0; k < 2; k++) {
        m[k] = s->mass[idx[k]];
        Gm[k] = s->Gmass[idx[k]];
        R[k] = s->radius[idx[k]];
        for (int d = 0; d < 3; d++) {
            rh[3 * k + d] = s->rh[3 * idx[k] + d];
            vb[3 * k + d] = s->vb[3 * idx[k] + d];
        }
    }
    const int ib = idx[swo_merge_bodies(m, Gm, R, rh, vb, cb_rb, s->cb_vb, s->GU, out)];
    enclist *l = &s->plpl;
    for (int64_t k = 0; k < l->n; k++)
        for (int q = 0; q < 2; q++) {
            const int b = idx[q];
            if (b == ib) continue;
            if (nid1[k] == s->id[b]) {
                nid1[k] = s->id[ib];
                l->i1[k] = b + 1;
            }
            if (nid2[k] == s->id[b]) {
                nid2[k] = s->id[ib];
                l->i2[k] = b + 1;
            }
            if (nid1[k] == nid2[k]) l->status[k] = SWO_INACTIVE;
        }
    s->maxid = s->maxid + 1;
    nb->id = s->maxid;
    for (int d = 0; d < 3; d++) {
        nb->rh[d] = out[d];
        nb->vh[d] = out[3 + d];
        nb->vb[d] = out[6 + d];
    }
    nb->mass = out[9];
    nb->Gmass = out[10];
    nb->radius = out[11];
    nb->levelg = s->levelg[ib];
    nb->levelm = s->levelm[ib];
}

/* swiftest_util_rearray_pl (swiftest_util.f90:1638-1811) after the merges:
 * drop the discards (keep[] = 0), append the new bodies, mass sort, rhill,
 * the level-irec encounter check over the full step, and the old entries'
 * state carried over where the same pair sits at the same list position. */
static void rearray(swo_sys *s, const char *keep, const newbody *nb, int nnew, const int32_t *oid1,
                    const int32_t *oid2, int irec)
{
    const int nold = s->npl;
    int n = 0;
    int *src = malloc(sizeof(int) * (nold + nnew + 1));
    for (int i = 0; i < nold; i++)
        if (keep[i]) src[n++] = i;
    for (int q = 0; q < nnew; q++) src[n++] = -1 - q;
    /* mass sort, descending (swiftest_util_sort_pl); ties keep their order */
    double *key = dalloc(n);
    for (int q = 0; q < n; q++) key[q] = src[q] >= 0 ? s->mass[src[q]] : nb[-1 - src[q]].mass;
    int *ord = malloc(sizeof(int) * (n + 1));
    for (int q = 0; q < n; q++) ord[q] = q;
    for (int q = 1; q < n; q++) {
        int v = ord[q], p = q - 1;
        while (p >= 0 && key[ord[p]] < key[v]) {
            ord[p + 1] = ord[p];
            p--;
        }
        ord[p + 1] = v;
    }
    double *rh = dalloc(3 * n), *vh = dalloc(3 * n), *vb = dalloc(3 * n), *Gm = dalloc(n), *ms = dalloc(n),
           *R = dalloc(n);
    int32_t *id = ialloc(n), *lg = ialloc(n), *lm = ialloc(n), *isp = ialloc(n);
    double *atp = dalloc(n), *peri = dalloc(n);
    for (int q = 0; q < n; q++) {
        const int a = src[ord[q]];
        if (a >= 0) {
            memcpy(&rh[3 * q], &s->rh[3 * a], 3 * sizeof(double));
            memcpy(&vh[3 * q], &s->vh[3 * a], 3 * sizeof(double));
            memcpy(&vb[3 * q], &s->vb[3 * a], 3 * sizeof(double));
            Gm[q] = s->Gmass[a];
            ms[q] = s->mass[a];
            R[q] = s->radius[a];
            id[q] = s->id[a];
            lg[q] = s->levelg[a];
            lm[q] = s->levelm[a];
            if (s->dis_ready) {
                isp[q] = s->pisperi[a];
                atp[q] = s->patp[a];
                peri[q] = s->pperi[a];
            }
        } else {
            const newbody *b = &nb[-1 - a];
            memcpy(&rh[3 * q], b->rh, 3 * sizeof(double));
            memcpy(&vh[3 * q], b->vh, 3 * sizeof(double));
            memcpy(&vb[3 * q], b->vb, 3 * sizeof(double));
            Gm[q] = b->Gmass;
            ms[q] = b->mass;
            R[q] = b->radius;
            id[q] = b->id;
            lg[q] = b->levelg;
            lm[q] = b->levelm;
            isp[q] = 1; /* swiftest_util_setup_body: isperi = 1 */
            atp[q] = 0.0;
            peri[q] = 0.0;
        }
    }
    s->npl = n;
    memcpy(s->rh, rh, 3 * sizeof(double) * n);
    memcpy(s->vh, vh, 3 * sizeof(double) * n);
    memcpy(s->vb, vb, 3 * sizeof(double) * n);
    memcpy(s->Gmass, Gm, sizeof(double) * n);
    memcpy(s->mass, ms, sizeof(double) * n);
    memcpy(s->radius, R, sizeof(double) * n);
    memcpy(s->id, id, sizeof(int32_t) * n);
    memcpy(s->levelg, lg, sizeof(int32_t) * n);
    memcpy(s->levelm, lm, sizeof(int32_t) * n);
    if (s->dis_ready) {
        memcpy(s->pisperi, isp, sizeof(int32_t) * n);
        memcpy(s->patp, atp, sizeof(double) * n);
        memcpy(s->pperi, peri, sizeof(double) * n);
    }
    for (int i = 0; i < n; i++) {
        s->status[i] = SWO_ACTIVE;
        s->lmask[i] = 1;
        s->mu[i] = s->GMcb + s->Gmass[i];
    }
    free(rh);
    free(vh);
    free(vb);
    free(Gm);
    free(ms);
    free(R);
    free(id);
    free(lg);
    free(lm);
    free(isp);
    free(atp);
    free(peri);
    free(key);
    free(ord);
    free(src);
    if (s->integ == SWO_SYMBA && s->p.gmtiny > 0.0) {
        s->nplm = 0;
        for (int i = 0; i < n; i++) s->nplm += !(s->Gmass[i] < s->p.gmtiny);
    } else {
        s->nplm = n;
    }
    set_rhill(s);
    /* the old list, as the merges left it */
    enclist *l = &s->plpl;
    const int64_t nenc_old = l->n;
    int32_t *olv = ialloc(nenc_old), *ost = ialloc(nenc_old), *olev = ialloc(nenc_old);
    memcpy(olv, l->lvdotr, sizeof(int32_t) * nenc_old);
    memcpy(ost, l->status, sizeof(int32_t) * nenc_old);
    memcpy(olev, l->level, sizeof(int32_t) * nenc_old);
    /* symba_encounter_check_pl at level irec over the full step: the list is
     * replaced only when the check finds an encounter */
    int32_t *nid1, *nid2;
    int64_t nnew_enc = 0;
    {
        set_renc(s, irec);
        int32_t *i1, *i2, *lv;
        nnew_enc = swo_encounter_check_all_plplm(s->npl, s->nplm, s->rh, s->vb, s->renc, s->p.dt, &i1, &i2, &lv);
        if (nnew_enc > 0) {
            list_resize(l, nnew_enc);
            nid1 = ialloc(nnew_enc);
            nid2 = ialloc(nnew_enc);
            for (int64_t k = 0; k < nnew_enc; k++) {
                l->i1[k] = i1[k];
                l->i2[k] = i2[k];
                l->lvdotr[k] = lv[k];
                l->status[k] = SWO_ACTIVE;
                l->level[k] = irec;
                int a = i1[k] - 1, b = i2[k] - 1;
                nid1[k] = s->id[a];
                nid2[k] = s->id[b];
                s->levelg[a] = irec;
                s->levelm[a] = irec;
                s->levelg[b] = irec;
                s->levelm[b] = irec;
            }
        } else {
            nid1 = ialloc(nenc_old);
            nid2 = ialloc(nenc_old);
            memcpy(nid1, oid1, sizeof(int32_t) * nenc_old);
            memcpy(nid2, oid2, sizeof(int32_t) * nenc_old);
        }
        free(i1);
        free(i2);
        free(lv);
    }
    if (nenc_old > 0) {
        int64_t nmin = l->n < nenc_old ? l->n : nenc_old;
        l->n = nmin;
        int bad = 0;
        for (int64_t k = 0; k < nmin; k++) {
            if ((nid1[k] == oid1[k] && nid2[k] == oid2[k]) || (nid1[k] == oid2[k] && nid2[k] == oid1[k])) {
                l->lvdotr[k] = olv[k];
                l->status[k] = ost[k];
                l->level[k] = olev[k];
            }
            int a = 0, b = 0;
            for (int q = 0; q < s->npl && !(a && b); q++) {
                if (!a && s->id[q] == nid1[k]) a = q + 1;
                if (!b && s->id[q] == nid2[k]) b = q + 1;
            }
            l->i1[k] = a;
            l->i2[k] = b;
            bad |= (a == 0 || b == 0);
        }
        if (bad) { /* pack the entries whose bodies are gone */
            int64_t o = 0;
            for (int64_t k = 0; k < nmin; k++) {
                if (l->i1[k] == 0 || l->i2[k] == 0) continue;
                l->i1[o] = l->i1[k];
                l->i2[o] = l->i2[k];
                l->lvdotr[o] = l->lvdotr[k];
                l->status[o] = l->status[k];
                l->level[o] = l->level[k];
                o++;
            }
            l->n = o;
        }
    }
    free(nid1);
    free(nid2);
    free(olv);
    free(ost);
    free(olev);
}

/* collision_resolve_plpl (collision_resolve.f90:521-662) */
static int resolve_plpl(swo_sys *s, double dt, int irec)
{
    vb2vh_pl(s); /* :557-558: vh and cb%vb, then cb%rb, from the current state */
    double cb_rb[3];
    rh2rb_cb(s, cb_rb);
    for (int loop = 1; loop <= 1000; loop++) {
        enclist *l = &s->plpl;
        /* collision_resolve_extract_plpl (:180-250): the colliding entries
         * leave the list, in list order */
        int64_t nc = 0;
        for (int64_t k = 0; k < l->n; k++) nc += l->status[k] == SWO_COLLIDED;
        if (nc == 0) return 0;
        int32_t *ci = ialloc(nc), *cj = ialloc(nc);
        char *seen = calloc(s->npl > 0 ? s->npl : 1, 1);
        int64_t o = 0, c = 0;
        int family = 0;
        for (int64_t k = 0; k < l->n; k++) {
            if (l->status[k] == SWO_COLLIDED) {
                ci[c] = l->i1[k] - 1;
                cj[c] = l->i2[k] - 1;
                family |= seen[ci[c]] | seen[cj[c]];
                seen[ci[c]] = seen[cj[c]] = 1;
                c++;
                continue;
            }
            l->i1[o] = l->i1[k];
            l->i2[o] = l->i2[k];
            l->lvdotr[o] = l->lvdotr[k];
            l->status[o] = l->status[k];
            l->level[o] = l->level[k];
            o++;
        }
        l->n = o;
        if (family) {
            free(ci);
            free(cj);
            free(seen);
            return -4;
        }
        /* the list's ids (id1/id2 of the entries), and a copy of them as
         * they stand before the merges rename the impactors */
        int32_t *nid1 = ialloc(l->n), *nid2 = ialloc(l->n);
        for (int64_t k = 0; k < l->n; k++) {
            nid1[k] = s->id[l->i1[k] - 1];
            nid2[k] = s->id[l->i2[k] - 1];
        }
        newbody *nb = calloc(nc, sizeof(newbody));
        char *keep = malloc(s->npl > 0 ? s->npl : 1);
        memset(keep, 1, s->npl > 0 ? s->npl : 1);
        for (int64_t q = 0; q < nc; q++) {
            merge_pair(s, ci[q], cj[q], cb_rb, nid1, nid2, &nb[q]);
            keep[ci[q]] = keep[cj[q]] = 0;
            s->nmerge++;
        }
        rearray(s, keep, nb, (int)nc, nid1, nid2, irec);
        free(nid1);
        free(nid2);
        free(nb);
        free(keep);
        free(ci);
        free(cj);
        free(seen);
        /* a new body may itself be colliding (:646-648) */
        if (!collision_check(s, dt, irec)) return 0;
    }
    return -5; /* a runaway collisional cascade (:649-654) */
}

/* symba_step_recur_system, symba_step.f90:163-262 */
static int recur(swo_sys *s, int ireci)
{
    s->irec = ireci;
    if (ireci > s->st.max_depth) s->st.max_depth = ireci;
    double p3 = 1.0;
    for (int q = 0; q < ireci; q++) p3 *= NTENC; /* NTENC**ireci, exact */
    double dtl = s->p.dt / p3;
    double dth = 0.5 * dtl;
    if (dtl / s->p.dt < DBL_EPSILON) return -3;
    int irecp = ireci + 1;
    int nloops = (ireci == 0) ? 1 : NTENC;
    for (int j = 1; j <= nloops; j++) {
        s->st.nsubsteps++;
        int lenc = list_check_plpl(s, dtl, irecp);
        if (!lenc || !s->p.short_circuit) lenc = list_check_pltp(s, dtl, irecp) || lenc;
        kick_list_plpl(s, dth, irecp, 1);
        kick_list_pltp(s, dth, irecp, 1);
        if (ireci != 0 && SWO_MUTANT != 2) {
            kick_list_plpl(s, dth, irecp, -1);
            kick_list_pltp(s, dth, irecp, -1);
        }
        drift_pl(s, dtl);
        drift_tp(s, dtl);
        if (lenc) {
            int rc = recur(s, irecp);
            if (rc) return rc;
        }
        s->irec = ireci;
        kick_list_plpl(s, dth, irecp, 1);
        kick_list_pltp(s, dth, irecp, 1);
        if (ireci != 0 && SWO_MUTANT != 2) {
            kick_list_plpl(s, dth, irecp, -1);
            kick_list_pltp(s, dth, irecp, -1);
        }
        if (s->p.lclose && collision_check(s, dtl, ireci)) {
            if (s->merge) {
                int rc = resolve_plpl(s, dtl, ireci);
                if (rc) return rc;
            } else {
                if (!s->col_finish) return -1;
                s->col_seen = 1;
            }
        }
        set_recur_levels(s, ireci);
    }
    return 0;
}

/* symba_step_interp_system, symba_step.f90:55-115 */
static int symba_interp(swo_sys *s)
{
    double dt = s->p.dt, dth = 0.5 * dt;
    s->irec = -1;
    if (s->lfirst) vh2vb_pl(s);
    lindrift_pl(s, dth, 1);
    kick_vb_pl(s, dth, 1);
    drift_pl(s, dt);
    if (s->ntp > 0) {
        if (s->tlfirst) {
            double vbcb[3] = {-s->ptbeg[0], -s->ptbeg[1], -s->ptbeg[2]};
            vh2vb_tp(s, vbcb);
        }
        lindrift_tp(s, dth, 1);
        kick_vb_tp(s, dth, 1);
        drift_tp(s, dt);
    }
    int rc = recur(s, 0);
    if (rc) return rc;
    s->irec = -1;
    kick_vb_pl(s, dth, 0);
    lindrift_pl(s, dth, 0);
    vb2vh_pl(s);
    if (s->ntp > 0) {
        kick_vb_tp(s, dth, 0);
        lindrift_tp(s, dth, 0);
        double vbcb[3] = {-s->ptend[0], -s->ptend[1], -s->ptend[2]};
        vb2vh_tp(s, vbcb);
    }
    return 0;
}

/* symba_step_system, symba_step.f90:14-52 */
static int symba_step(swo_sys *s)
{
    symba_reset(s);
    s->st.nplpl_enc = 0;
    s->st.npltp_enc = 0;
    s->st.max_depth = -1;
    s->st.nsubsteps = 0;
    double dt = s->p.dt;
    int lenc = symba_check_pl(s, dt, 0);
    if (!lenc || !s->p.short_circuit) lenc = symba_check_tp(s, dt, 0) || lenc;
    int rc = 0;
    s->col_seen = 0;
    if (lenc) {
        rc = symba_interp(s);
    } else {
        s->irec = -1;
        helio_step(s);
    }
    s->lfirstkick = s->lfirst;
    if (!rc && s->col_seen) rc = -1;
    return rc;
}

/* ---- WHM (src/whm) ---- */

/* whm_coord_h2j_pl, whm_coord.f90:14-51 */
static void whm_h2j(swo_sys *s)
{
    int n = s->npl;
    if (n == 0) return;
    double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) {
        s->xj[d] = s->rh[d];
        s->vj[d] = s->vh[d];
    }
    for (int i = 1; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            sx[d] = sx[d] + s->Gmass[i - 1] * s->rh[3 * (i - 1) + d];
            sv[d] = sv[d] + s->Gmass[i - 1] * s->vh[3 * (i - 1) + d];
            double cap = sx[d] / s->eta[i - 1], capv = sv[d] / s->eta[i - 1];
            s->xj[3 * i + d] = s->rh[3 * i + d] - cap;
            s->vj[3 * i + d] = s->vh[3 * i + d] - capv;
        }
    }
}

/* whm_coord_j2h_pl, whm_coord.f90:54-90 */
static void whm_j2h(swo_sys *s)
{
    int n = s->npl;
    if (n == 0) return;
    double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) {
        s->rh[d] = s->xj[d];
        s->vh[d] = s->vj[d];
    }
    for (int i = 1; i < n; i++) {
        for (int d = 0; d < 3; d++) {
            sx[d] = sx[d] + (s->Gmass[i - 1] * s->xj[3 * (i - 1) + d]) / s->eta[i - 1];
            sv[d] = sv[d] + (s->Gmass[i - 1] * s->vj[3 * (i - 1) + d]) / s->eta[i - 1];
            s->rh[3 * i + d] = s->xj[3 * i + d] + sx[d];
            s->vh[3 * i + d] = s->vj[3 * i + d] + sv[d];
        }
    }
}

/* whm_coord_vh2vj_pl, whm_coord.f90:93-124 */
static void whm_vh2vj(swo_sys *s)
{
    int n = s->npl;
    if (n == 0) return;
    double sv[3] = {0, 0, 0};
    for (int d = 0; d < 3; d++) s->vj[d] = s->vh[d];
    for (int i = 1; i < n; i++)
        for (int d = 0; d < 3; d++) {
            sv[d] = sv[d] + s->Gmass[i - 1] * s->vh[3 * (i - 1) + d];
            s->vj[3 * i + d] = s->vh[3 * i + d] - sv[d] / s->eta[i - 1];
        }
}

/* whm_kick_getacch_ah0, whm_kick.f90:125-149 */
static void whm_ah0(const double *mu, const double *r, int n, double *ah0)
{
    ah0[0] = ah0[1] = ah0[2] = 0.0;
    for (int i = 0; i < n; i++) {
        const double *x = &r[3 * i];
        double r2 = dot3(x[0], x[1], x[2], x[0], x[1], x[2]);
        double ir3h = 1.0 / (r2 * sqrt(r2));
        double fac = mu[i] * ir3h;
        for (int d = 0; d < 3; d++) ah0[d] = fma(-fac, x[d], ah0[d]);
    }
}

/* whm_kick_getacch_pl, whm_kick.f90:14-67 (set_ir3j whm_util.f90:117-140,
 * ah1 :152-179, ah2 :182-215, then accel_int) */
static void whm_getacch_pl(swo_sys *s)
{
    int n = s->npl;
    for (int i = 0; i < n; i++) {
        const double *x = &s->rh[3 * i];
        double r2 = dot3(x[0], x[1], x[2], x[0], x[1], x[2]);
        double ir = 1.0 / sqrt(r2);
        s->ir3h[i] = ir / r2;
        x = &s->xj[3 * i];
        r2 = dot3(x[0], x[1], x[2], x[0], x[1], x[2]);
        ir = 1.0 / sqrt(r2);
        s->ir3j[i] = ir / r2;
    }
    double ah0[3];
    whm_ah0(s->Gmass + 1, s->rh + 3, n - 1, ah0);
    for (int i = 0; i < n; i++)
        for (int d = 0; d < 3; d++) s->ah[3 * i + d] = s->ah[3 * i + d] + ah0[d];
    for (int i = 1; i < n; i++) {
        if (!s->lmask[i]) continue;
        for (int d = 0; d < 3; d++) {
            double a1j = s->xj[3 * i + d] * s->ir3j[i];
            double a1h = s->rh[3 * i + d] * s->ir3h[i];
            s->ah[3 * i + d] = s->ah[3 * i + d] + s->GMcb * (a1j - a1h);
        }
    }
    double etaj = s->GMcb, ah2o[3] = {0, 0, 0};
    for (int i = 1; i < n; i++) {
        if (!s->lmask[i]) continue;
        etaj = etaj + s->Gmass[i - 1];
        double fac = ((s->Gmass[i] * s->GMcb) * s->ir3j[i]) / etaj;
        for (int d = 0; d < 3; d++) {
            double ah2 = ah2o[d] + fac * s->xj[3 * i + d];
            s->ah[3 * i + d] = s->ah[3 * i + d] + ah2;
            ah2o[d] = ah2;
        }
    }
    swo_kick_getacch_int_all_tri(n, n, s->rh, s->Gmass, s->p.lclose ? s->radius : NULL, s->ah);
}

/* whm_kick_getacch_tp, whm_kick.f90:70-122 */
static void whm_getacch_tp(swo_sys *s, int lbeg)
{
    if (s->ntp == 0 || s->npl == 0) return;
    const double *rp = lbeg ? s->rbeg : s->rend;
    double ah0[3];
    whm_ah0(s->Gmass, rp, s->npl, ah0);
    for (int i = 0; i < s->ntp; i++)
        if (s->tlmask[i])
            for (int d = 0; d < 3; d++) s->tah[3 * i + d] = s->tah[3 * i + d] + ah0[d];
    swo_kick_getacch_int_all_tp(s->ntp, s->npl, s->trh, rp, s->Gmass, s->tlmask, s->tah);
}

/* whm_step_pl / whm_step_tp with whm_kick_vh_pl / _tp (whm_step.f90:38-104,
 * whm_kick.f90:218-318) */
static int whm_step(swo_sys *s)
{
    double dt = s->p.dt, dth = 0.5 * dt;
    int n = s->npl;
    s->tlfirst = s->lfirst;
    if (n > 0) {
        if (s->lfirst) {
            whm_h2j(s);
            memset(s->ah, 0, sizeof(double) * 3 * n);
            whm_getacch_pl(s);
            s->lfirst = 0;
        }
        memcpy(s->rbeg, s->rh, sizeof(double) * 3 * n);
        for (int i = 0; i < n; i++)
            if (s->lmask[i])
                for (int d = 0; d < 3; d++) s->vh[3 * i + d] = fma(s->ah[3 * i + d], dth, s->vh[3 * i + d]);
        whm_vh2vj(s);
        for (int i = 0; i < n; i++) s->iflag[i] = 0;
        swo_drift_all(s->muj, s->xj, s->vj, n, dt, s->lmask, s->iflag, 0, 0.0);
        for (int i = 0; i < n; i++)
            if (s->iflag[i]) {
                s->status[i] = SWO_DISCARDED_DRIFTERR;
                s->st.ndrifterr++;
                return -2; /* whm_drift.f90:38-53: fatal */
            }
        whm_j2h(s);
        memset(s->ah, 0, sizeof(double) * 3 * n);
        whm_getacch_pl(s);
        memcpy(s->rend, s->rh, sizeof(double) * 3 * n);
        for (int i = 0; i < n; i++)
            if (s->lmask[i])
                for (int d = 0; d < 3; d++) s->vh[3 * i + d] = fma(s->ah[3 * i + d], dth, s->vh[3 * i + d]);
    }
    if (s->ntp > 0) {
        if (s->tlfirst) {
            for (int i = 0; i < s->ntp; i++)
                if (s->tlmask[i]) memset(&s->tah[3 * i], 0, 3 * sizeof(double));
            whm_getacch_tp(s, 1);
            s->tlfirst = 0;
        }
        for (int i = 0; i < s->ntp; i++)
            if (s->tlmask[i])
                for (int d = 0; d < 3; d++) s->tvh[3 * i + d] = fma(s->tah[3 * i + d], dth, s->tvh[3 * i + d]);
        double *mu = dalloc(s->ntp);
        for (int i = 0; i < s->ntp; i++) {
            mu[i] = s->GMcb;
            s->tiflag[i] = 0;
        }
        swo_drift_all(mu, s->trh, s->tvh, s->ntp, dt, s->tlmask, s->tiflag, 0, 0.0);
        free(mu);
        for (int i = 0; i < s->ntp; i++)
            if (s->tiflag[i]) {
                s->tstatus[i] = SWO_DISCARDED_DRIFTERR;
                s->tlmask[i] = 0;
                s->st.ndrifterr++;
            }
        for (int i = 0; i < s->ntp; i++)
            if (s->tlmask[i]) memset(&s->tah[3 * i], 0, 3 * sizeof(double));
        whm_getacch_tp(s, 0);
        for (int i = 0; i < s->ntp; i++)
            if (s->tlmask[i])
                for (int d = 0; d < 3; d++) s->tvh[3 * i + d] = fma(s->tah[3 * i + d], dth, s->tvh[3 * i + d]);
    }
    return 0;
}

/* ---- construction, snapshot, driver loop ---- */

static void sort_pl(swo_sys *s, const double *key)
{
    /* stable descending sort of the pl arrays by key */
    int n = s->npl;
    int *idx = malloc(sizeof(int) * (n > 0 ? n : 1));
    for (int i = 0; i < n; i++) idx[i] = i;
    for (int i = 1; i < n; i++) { /* insertion sort on an index (n is modest) */
        int v = idx[i], j = i - 1;
        while (j >= 0 && key[idx[j]] < key[v]) {
            idx[j + 1] = idx[j];
            j--;
        }
        idx[j + 1] = v;
    }
    double *t3 = dalloc(3 * (int64_t)n), *t1 = dalloc(n);
    int32_t *ti = ialloc(n);
#define PERM3(a)                                                                     \
    do {                                                                             \
        for (int q = 0; q < n; q++) memcpy(&t3[3 * q], &(a)[3 * idx[q]], 3 * sizeof(double)); \
        memcpy((a), t3, 3 * sizeof(double) * n);                                     \
    } while (0)
#define PERM1(a)                                             \
    do {                                                     \
        for (int q = 0; q < n; q++) t1[q] = (a)[idx[q]];     \
        memcpy((a), t1, sizeof(double) * n);                 \
    } while (0)
    PERM3(s->rh);
    PERM3(s->vh);
    PERM1(s->Gmass);
    PERM1(s->radius);
    PERM1(s->rhill);
    for (int q = 0; q < n; q++) ti[q] = s->id[idx[q]];
    memcpy(s->id, ti, sizeof(int32_t) * n);
#undef PERM3
#undef PERM1
    free(t3);
    free(t1);
    free(ti);
    free(idx);
}

swo_sys *swo_sys_new(int integrator, int npl, int ntp, double GMcb, double Rcb, const double *pl_rh,
                     const double *pl_vh, const double *Gmass, const double *radius, const double *rhill_in,
                     const double *tp_rh, const double *tp_vh, const swo_params *prm)
{
    swo_sys *s = calloc(1, sizeof(swo_sys));
    s->integ = integrator;
    s->npl = npl;
    s->ntp = ntp;
    s->GMcb = GMcb;
    s->Rcb = Rcb;
    s->p = *prm;
#ifdef _OPENMP
    omp_set_num_threads(omp_threads(prm->nthreads));
#endif
    int64_t n3 = 3 * (int64_t)npl, t3 = 3 * (int64_t)ntp;
    s->rh = dalloc(n3);
    s->vh = dalloc(n3);
    s->vb = dalloc(n3);
    s->ah = dalloc(n3);
    s->rbeg = dalloc(n3);
    s->rend = dalloc(n3);
    s->xj = dalloc(n3);
    s->vj = dalloc(n3);
    s->Gmass = dalloc(npl);
    s->radius = dalloc(npl);
    s->rhill = dalloc(npl);
    s->renc = dalloc(npl);
    s->mu = dalloc(npl);
    s->eta = dalloc(npl);
    s->muj = dalloc(npl);
    s->ir3h = dalloc(npl);
    s->ir3j = dalloc(npl);
    s->status = ialloc(npl);
    s->levelg = ialloc(npl);
    s->levelm = ialloc(npl);
    s->lmask = ialloc(npl);
    s->id = ialloc(npl);
    s->iflag = ialloc(npl);
    s->trh = dalloc(t3);
    s->tvh = dalloc(t3);
    s->tvb = dalloc(t3);
    s->tah = dalloc(t3);
    s->tstatus = ialloc(ntp);
    s->tlevelg = ialloc(ntp);
    s->tlevelm = ialloc(ntp);
    s->tlmask = ialloc(ntp);
    s->tid = ialloc(ntp);
    s->tiflag = ialloc(ntp);
    memcpy(s->rh, pl_rh, sizeof(double) * n3);
    memcpy(s->vh, pl_vh, sizeof(double) * n3);
    memcpy(s->Gmass, Gmass, sizeof(double) * npl);
    memcpy(s->radius, radius, sizeof(double) * npl);
    if (ntp) {
        memcpy(s->trh, tp_rh, sizeof(double) * t3);
        memcpy(s->tvh, tp_vh, sizeof(double) * t3);
    }
    for (int i = 0; i < npl; i++) {
        s->id[i] = i + 1;
        s->lmask[i] = 1;
        s->levelg[i] = s->levelm[i] = -1;
    }
    for (int i = 0; i < ntp; i++) {
        s->tid[i] = npl + i + 1;
        s->tlmask[i] = 1;
        s->tlevelg[i] = s->tlevelm[i] = -1;
    }
    /* swiftest_util_setup_initialize_system (swiftest_util.f90:2415-2462) */
    if (rhill_in && prm->rhill_present)
        memcpy(s->rhill, rhill_in, sizeof(double) * npl);
    else
        set_rhill(s);
    s->lfirst = s->tlfirst = s->lfirstkick = 1;
    s->nplm = npl;
    if (integrator == SWO_WHM) {
        /* whm_util_setup_initialize_system (whm_util.f90:201-228): sort by
         * ir3h descending, then eta / muj (whm_util_set_mu_eta_pl :175-198) */
        double *key = dalloc(npl);
        for (int i = 0; i < npl; i++) {
            const double *x = &s->rh[3 * i];
            double r2 = dot3(x[0], x[1], x[2], x[0], x[1], x[2]);
            key[i] = (1.0 / sqrt(r2)) / r2;
        }
        sort_pl(s, key);
        free(key);
        if (npl > 0) {
            s->eta[0] = GMcb + s->Gmass[0];
            s->muj[0] = s->eta[0];
            for (int i = 1; i < npl; i++) {
                s->eta[i] = s->eta[i - 1] + s->Gmass[i];
                s->muj[i] = (GMcb * s->eta[i]) / s->eta[i - 1];
            }
        }
    } else {
        /* helio_util_setup_initialize_system (helio_util.f90:14-43): sort by
         * mass descending, vh2vb, tp h2b */
        double *key = dalloc(npl);
        memcpy(key, s->Gmass, sizeof(double) * npl);
        sort_pl(s, key);
        free(key);
        /* symba_util_flatten_eucl_plpl (symba_util.f90:191-202), lmtiny_pl
         * for SyMBA (swiftest_io.f90:2663): lmtiny = Gmass < GMTINY */
        s->nplm = npl;
        if (integrator == SWO_SYMBA && prm->gmtiny > 0.0) {
            s->nplm = 0;
            for (int i = 0; i < npl; i++) s->nplm += !(s->Gmass[i] < prm->gmtiny);
        }
        vh2vb_pl(s);
        for (int i = 0; i < ntp; i++)
            for (int d = 0; d < 3; d++) s->tvb[3 * i + d] = s->tvh[3 * i + d] + s->cb_vb[d];
    }
    for (int i = 0; i < npl; i++) s->mu[i] = GMcb + s->Gmass[i];
    s->st.max_depth = -1;
    return s;
}

/* swiftest_util_snapshot_system coordinate refresh, swiftest_util.f90:2663-2682 */
static void snapshot(swo_sys *s)
{
    if (s->integ == SWO_WHM) {
        whm_h2j(s);
    } else {
        vh2vb_pl(s);
        double vbcb[3] = {-s->ptbeg[0], -s->ptbeg[1], -s->ptbeg[2]};
        vh2vb_tp(s, vbcb);
    }
    if (!s->p.rhill_present) set_rhill(s);
}

int swo_sys_step(swo_sys *s)
{
    int rc;
    if (s->integ == SWO_WHM)
        rc = whm_step(s);
    else if (s->integ == SWO_HELIO) {
        s->irec = -1;
        helio_step(s);
        rc = s->err;
    } else
        rc = symba_step(s);
    if (rc == 0 && s->err) rc = s->err;
    return rc;
}

/* swiftest_driver time loop (swiftest_driver.f90:130-160): step, then a
 * snapshot every istep_out steps.  The initial snapshot before the loop is
 * taken on the first call. */
int swo_sys_run(swo_sys *s, int64_t nsteps)
{
    if (s->iloop == 0 && s->p.istep_out > 0) snapshot(s);
    for (int64_t k = 0; k < nsteps; k++) {
        int rc = swo_sys_step(s);
        s->iloop++;
        if (rc) return rc;
        if (s->p.istep_out > 0) {
            s->iout++;
            if (s->iout == s->p.istep_out) {
                s->iout = 0;
                snapshot(s);
            }
        }
    }
    return 0;
}

void swo_sys_free(swo_sys *s)
{
    if (!s) return;
    void *ptrs[] = {s->rh, s->vh, s->vb, s->ah, s->rbeg, s->rend, s->xj, s->vj, s->Gmass, s->radius, s->rhill,
                    s->renc, s->mu, s->eta, s->muj, s->ir3h, s->ir3j, s->status, s->levelg, s->levelm, s->lmask,
                    s->id, s->iflag, s->trh, s->tvh, s->tvb, s->tah, s->tstatus, s->tlevelg, s->tlevelm,
                    s->tlmask, s->tid, s->tiflag, s->plpl.i1, s->plpl.i2, s->plpl.lvdotr, s->plpl.status,
                    s->plpl.level, s->pltp.i1, s->pltp.i2, s->pltp.lvdotr, s->pltp.status, s->pltp.level,
                    s->pisperi, s->tisperi, s->patp, s->pperi, s->tatp, s->tperi, s->mass};
    for (size_t q = 0; q < sizeof ptrs / sizeof ptrs[0]; q++) free(ptrs[q]);
    free(s);
}

void swo_sys_get(const swo_sys *s, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh,
                 double *tp_vb, int32_t *pl_status, int32_t *tp_status)
{
    int64_t n3 = 3 * (int64_t)s->npl, t3 = 3 * (int64_t)s->ntp;
    if (pl_rh) memcpy(pl_rh, s->rh, n3 * sizeof(double));
    if (pl_vh) memcpy(pl_vh, s->vh, n3 * sizeof(double));
    if (pl_vb) memcpy(pl_vb, s->vb, n3 * sizeof(double));
    if (tp_rh) memcpy(tp_rh, s->trh, t3 * sizeof(double));
    if (tp_vh) memcpy(tp_vh, s->tvh, t3 * sizeof(double));
    if (tp_vb) memcpy(tp_vb, s->tvb, t3 * sizeof(double));
    if (pl_status) memcpy(pl_status, s->status, s->npl * sizeof(int32_t));
    if (tp_status) memcpy(tp_status, s->tstatus, s->ntp * sizeof(int32_t));
}

void swo_sys_get_misc(const swo_sys *s, double *rhill, double *cb_vb, double *ptbeg, double *ptend,
                      int32_t *lfirst)
{
    if (rhill) memcpy(rhill, s->rhill, s->npl * sizeof(double));
    if (cb_vb) memcpy(cb_vb, s->cb_vb, sizeof s->cb_vb);
    if (ptbeg) memcpy(ptbeg, s->ptbeg, sizeof s->ptbeg);
    if (ptend) memcpy(ptend, s->ptend, sizeof s->ptend);
    if (lfirst) *lfirst = s->lfirstkick;
}

void swo_sys_get_order(const swo_sys *s, int32_t *pl_id) { memcpy(pl_id, s->id, s->npl * sizeof(int32_t)); }
void swo_sys_stats(const swo_sys *s, swo_stats *st) { *st = s->st; }
void swo_sys_get_levels(const swo_sys *s, int32_t *levelg, int32_t *levelm, int32_t *tlevelg, int32_t *tlevelm)
{
    if (levelg) memcpy(levelg, s->levelg, s->npl * sizeof(int32_t));
    if (levelm) memcpy(levelm, s->levelm, s->npl * sizeof(int32_t));
    if (tlevelg) memcpy(tlevelg, s->tlevelg, s->ntp * sizeof(int32_t));
    if (tlevelm) memcpy(tlevelm, s->tlevelm, s->ntp * sizeof(int32_t));
}
int swo_sys_npl(const swo_sys *s) { return s->npl; }
int swo_sys_ntp(const swo_sys *s) { return s->ntp; }

int64_t swo_sys_list(const swo_sys *s, int which, int32_t *index1, int32_t *index2, int32_t *lvdotr)
{
    const enclist *l = which == 0 ? &s->plpl : &s->pltp;
    if (index1) memcpy(index1, l->i1, l->n * sizeof(int32_t));
    if (index2) memcpy(index2, l->i2, l->n * sizeof(int32_t));
    if (lvdotr) memcpy(lvdotr, l->lvdotr, l->n * sizeof(int32_t));
    return l->n;
}

int64_t swo_sys_list_status(const swo_sys *s, int which, int32_t *status)
{
    const enclist *l = which == 0 ? &s->plpl : &s->pltp;
    if (status) memcpy(status, l->status, l->n * sizeof(int32_t));
    return l->n;
}

void swo_sys_set_collision_mode(swo_sys *s, int finish_step) { s->col_finish = finish_step != 0; }

/* Collisions resolved by merging inside the step (SyMBA without test
 * particles).  pl_mass in input order, or NULL for Gmass / GU.  Body ids are
 * 1..npl in input order; merged bodies take the next ids. */
int swo_sys_set_merge(swo_sys *s, double GU, const double *pl_mass)
{
    if (s->integ != SWO_SYMBA || s->ntp > 0 || !(GU > 0.0)) return -1;
    free(s->mass);
    s->mass = dalloc(s->npl);
    for (int i = 0; i < s->npl; i++) s->mass[i] = pl_mass ? pl_mass[s->id[i] - 1] : s->Gmass[i] / GU;
    s->GU = GU;
    s->merge = 1;
    s->maxid = s->npl + s->ntp;
    return 0;
}

void swo_sys_get_bodies(const swo_sys *s, int32_t *id, double *mass, double *Gmass, double *radius)
{
    for (int i = 0; i < s->npl; i++) {
        if (id) id[i] = s->id[i];
        if (mass) mass[i] = s->mass ? s->mass[i] : 0.0;
        if (Gmass) Gmass[i] = s->Gmass[i];
        if (radius) radius[i] = s->radius[i];
    }
}

int64_t swo_sys_nmerge(const swo_sys *s) { return s->nmerge; }

/* ======================================================================
 * Energy and angular momentum
 * ====================================================================== */

/* norm2 (Fortran 2008 intrinsic) of a 3-vector.  gfortran evaluates it as a
 * scaled sum of squares: scale = max |x| so far, result accumulates
 * (x/scale)^2 and is rescaled when a larger component appears; the norm is
 * scale * sqrt(result). */
double swo_norm2(double x, double y, double z)
{
    const double c[3] = {x, y, z};
    double scale = 1.0, result = 0.0;
    for (int k = 0; k < 3; k++) {
        if (c[k] == 0.0) continue;
        const double ax = fabs(c[k]);
        if (scale < ax) {
            const double val = scale / ax;
            result = 1.0 + result * val * val;
            scale = ax;
        } else {
            const double val = ax / scale;
            result += val * val;
        }
    }
    return scale * sqrt(result);
}

void swo_energy(int npl, double GMcb, double mcb, double Rcb, const double *cb_vb, const double *rh,
                const double *vb, const double *Gmass, const double *mass, const double *radius,
                const int32_t *status, int lclose, double *out, double *rb_out)
{
    double *rb = (double *)calloc(3 * (size_t)(npl > 0 ? npl : 1), sizeof(double));
    /* swiftest_util_coord_rh2rb_pl (swiftest_util.f90:505-516) */
    double cb_rb[3] = {0.0, 0.0, 0.0};
    if (npl > 0) {
        double Gmtot = GMcb, xtmp[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < npl; i++) {
            if (status[i] == SWO_INACTIVE) continue;
            Gmtot = Gmtot + Gmass[i];
            for (int d = 0; d < 3; d++) xtmp[d] = xtmp[d] + Gmass[i] * rh[3 * i + d];
        }
        for (int d = 0; d < 3; d++) cb_rb[d] = -xtmp[d] / Gmtot;
        for (int i = 0; i < npl; i++) {
            if (status[i] == SWO_INACTIVE) continue;
            for (int d = 0; d < 3; d++) rb[3 * i + d] = rh[3 * i + d] + cb_rb[d];
        }
    }
    /* swiftest_util_get_energy_and_momentum_system (:1192-1288) */
    double GMtot = GMcb;
    for (int i = 0; i < npl; i++)
        if (status[i] != SWO_INACTIVE) GMtot = GMtot + Gmass[i];
    const double kecb = mcb * dot3(cb_vb[0], cb_vb[1], cb_vb[2], cb_vb[0], cb_vb[1], cb_vb[2]);
    const double be_cb = -(3.0 * GMcb * mcb) / (5.0 * Rcb);
    double Lcb[3];
    Lcb[0] = mcb * (cb_rb[1] * cb_vb[2] - cb_rb[2] * cb_vb[1]);
    Lcb[1] = mcb * (cb_rb[2] * cb_vb[0] - cb_rb[0] * cb_vb[2]);
    Lcb[2] = mcb * (cb_rb[0] * cb_vb[1] - cb_rb[1] * cb_vb[0]);
    double kesum = 0.0, Lsum[3] = {0.0, 0.0, 0.0};
    for (int i = 0; i < npl; i++) {
        if (status[i] == SWO_INACTIVE) continue;
        const double *r = rb + 3 * i, *v = vb + 3 * i;
        const double h0 = r[1] * v[2] - r[2] * v[1];
        const double h1 = r[2] * v[0] - r[0] * v[2];
        const double h2 = r[0] * v[1] - r[1] * v[0];
        Lsum[0] = Lsum[0] + mass[i] * h0;
        Lsum[1] = Lsum[1] + mass[i] * h1;
        Lsum[2] = Lsum[2] + mass[i] * h2;
        kesum = kesum + mass[i] * dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
    }
    /* swiftest_util_get_potential_energy_triangular (:1350-1400) */
    double pe = 0.0;
    for (int i = 0; i < npl; i++) {
        if (status[i] == SWO_INACTIVE) continue;
        double row = 0.0;
        for (int j = i + 1; j < npl; j++) {
            if (status[j] == SWO_INACTIVE) continue;
            const double dr = swo_norm2(rb[3 * i] - rb[3 * j], rb[3 * i + 1] - rb[3 * j + 1],
                                        rb[3 * i + 2] - rb[3 * j + 2]);
            row = row + (-(Gmass[i] * mass[j]) / dr);
        }
        pe = pe + row;
    }
    double pecb = 0.0;
    for (int i = 0; i < npl; i++) {
        if (status[i] == SWO_INACTIVE) continue;
        pecb = pecb + (-(GMcb * mass[i]) / swo_norm2(rb[3 * i], rb[3 * i + 1], rb[3 * i + 2]));
    }
    pe = pe + pecb;
    double be = 0.0;
    if (lclose)
        for (int i = 0; i < npl; i++)
            if (status[i] != SWO_INACTIVE) be = be + (-(3.0 * Gmass[i] * mass[i]) / (5.0 * radius[i]));
    for (int k = 0; k < SWO_ENERGY_NOUT; k++) out[k] = 0.0;
    out[SWO_E_KE_ORBIT] = 0.5 * (kecb + kesum);
    out[SWO_E_PE] = pe;
    out[SWO_E_BE] = be;
    out[SWO_E_BE_CB] = be_cb;
    out[SWO_E_TE] = out[SWO_E_KE_ORBIT] + out[SWO_E_KE_SPIN] + pe + be;
    for (int d = 0; d < 3; d++) {
        out[SWO_E_L_ORBIT + d] = Lcb[d] + Lsum[d];
        out[SWO_E_L_TOTAL + d] = out[SWO_E_L_ORBIT + d] + out[SWO_E_L_SPIN + d];
    }
    out[SWO_E_GMTOT] = GMtot;
    if (rb_out) memcpy(rb_out, rb, 3 * (size_t)npl * sizeof(double));
    free(rb);
}

/* ======================================================================
 * Discard checks
 * ====================================================================== */

/* swiftest_util_peri (swiftest_util.f90:1565-1603) for one body */
static void peri_one(double m, const double *r, const double *v, int32_t *isperi, double *atp, double *q)
{
    const double vdotr = dot3(r[0], r[1], r[2], v[0], v[1], v[2]);
    if (*isperi == -1) {
        if (vdotr >= 0.0) {
            double e;
            *isperi = 0;
            swo_orbel_xv2aeq(m, r[0], r[1], r[2], v[0], v[1], v[2], atp, &e, q);
        }
    } else {
        *isperi = vdotr > 0.0 ? -1 : 1;
    }
}

/* swiftest_discard_pl_close (swiftest_discard.f90:291-337) */
static int pl_close(const double *dx, const double *dv, double dt, double r2crit)
{
    const double r2 = dot3(dx[0], dx[1], dx[2], dx[0], dx[1], dx[2]);
    if (r2 <= r2crit) return 1;
    const double vdotr = dot3(dx[0], dx[1], dx[2], dv[0], dv[1], dv[2]);
    if (vdotr > 0.0) return 0;
    const double v2 = dot3(dv[0], dv[1], dv[2], dv[0], dv[1], dv[2]);
    const double tmin = -vdotr / v2;
    double r2min;
    if (tmin < dt)
        r2min = r2 - (vdotr * vdotr) / v2;
    else
        r2min = (r2 + (2.0 * vdotr) * dt) + v2 * (dt * dt);
    r2min = fmin(r2min, r2);
    return r2min <= r2crit;
}

void swo_discard(int integrator, int npl, int ntp, double GMcb, double Rcb, double *cb_vb, const double *pl_rh,
                 const double *pl_vh, double *pl_vb, const double *Gmass, const double *radius, const double *rhill,
                 int32_t *pl_status, const int32_t *nplenc, int32_t *pl_isperi, double *pl_atp, double *pl_peri,
                 int32_t *lfirst_peri, const double *tp_rh, const double *tp_vh, double *tp_vb, int32_t *tp_status,
                 int32_t *tp_isperi, double *tp_atp, double *tp_peri, const double *prm, int qmin_coord, int lclose,
                 int64_t *counts)
{
    const double rmin = prm[0], rmax = prm[1], rmaxu = prm[2], qmin = prm[3], alo = prm[4], ahi = prm[5];
    const double dt = prm[6];
    const int lcb = rmin >= 0.0 || rmax >= 0.0 || rmaxu >= 0.0;
    counts[0] = counts[1] = 0;
    /* set_msys (swiftest_util.f90:2115) */
    double sm = 0.0;
    for (int i = 0; i < npl; i++)
        if (pl_status[i] != SWO_INACTIVE) sm = sm + Gmass[i];
    const double Gmtot = GMcb + sm;
    double *rb = (double *)calloc(3 * (size_t)(npl > 0 ? npl : 1), sizeof(double));
    /* bodies discarded here leave the arrays (pl%rearray) before the test
     * particles are checked */
    uint8_t *gone = (uint8_t *)calloc((size_t)(npl > 0 ? npl : 1), 1);
    double cb_rb[3] = {0.0, 0.0, 0.0};

    /* ---- massive bodies (SyMBA) ---- */
    if (integrator == SWO_SYMBA && npl > 0) {
        /* pl%vb2vh repeats the step's last conversion on the same inputs, so
         * it changes nothing; pl%rh2rb (swiftest_util.f90:505-516): */
        double g = GMcb, xt[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < npl; i++) {
            if (pl_status[i] == SWO_INACTIVE) continue;
            g = g + Gmass[i];
            for (int d = 0; d < 3; d++) xt[d] = xt[d] + Gmass[i] * pl_rh[3 * i + d];
        }
        for (int d = 0; d < 3; d++) cb_rb[d] = -xt[d] / g;
        for (int i = 0; i < npl; i++)
            if (pl_status[i] != SWO_INACTIVE)
                for (int d = 0; d < 3; d++) rb[3 * i + d] = pl_rh[3 * i + d] + cb_rb[d];
        if (lcb) { /* symba_discard_cb_pl (symba_discard.f90:36-101) */
            const double rmin2 = rmin * rmin, rmax2 = rmax * rmax, rmaxu2 = rmaxu * rmaxu;
            for (int i = 0; i < npl; i++) {
                if (pl_status[i] != SWO_ACTIVE) continue;
                const double *r = pl_rh + 3 * i, *b = rb + 3 * i, *v = pl_vb + 3 * i;
                const double rh2 = dot3(r[0], r[1], r[2], r[0], r[1], r[2]);
                int32_t st = SWO_ACTIVE;
                if (rmax >= 0.0 && rh2 > rmax2) {
                    st = SWO_DISCARDED_RMAX;
                } else if (rmin >= 0.0 && rh2 < rmin2) {
                    st = SWO_DISCARDED_RMIN;
                } else if (rmaxu >= 0.0) {
                    const double rb2 = dot3(b[0], b[1], b[2], b[0], b[1], b[2]);
                    const double vb2 = dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
                    const double energy = 0.5 * vb2 - Gmtot / sqrt(rb2);
                    if (energy > 0.0 && rb2 > rmaxu2) st = SWO_DISCARDED_RMAXU;
                }
                if (st != SWO_ACTIVE) {
                    pl_status[i] = st;
                    gone[i] = 1;
                    counts[0]++;
                }
            }
        }
        if (qmin >= 0.0) { /* symba_discard_peri_pl (symba_discard.f90:311-336) */
            const int first = *lfirst_peri;
            for (int i = 0; i < npl; i++) {
                if (first) pl_isperi[i] = 0; /* swiftest_util_peri_body, swiftest_util.f90:1625 */
                if (qmin_coord == 0)
                    peri_one(GMcb + Gmass[i], pl_rh + 3 * i, pl_vh + 3 * i, &pl_isperi[i], &pl_atp[i], &pl_peri[i]);
                else
                    peri_one(Gmtot, rb + 3 * i, pl_vb + 3 * i, &pl_isperi[i], &pl_atp[i], &pl_peri[i]);
            }
            if (first) {
                *lfirst_peri = 0;
            } else {
                for (int i = 0; i < npl; i++)
                    if (pl_status[i] == SWO_ACTIVE && pl_isperi[i] == 0 && nplenc[i] == 0 && pl_atp[i] >= alo &&
                        pl_atp[i] <= ahi && pl_peri[i] <= qmin) {
                        pl_status[i] = SWO_DISCARDED_PERI;
                        gone[i] = 1;
                        counts[0]++;
                    }
            }
        }
    }

    /* ---- test particles (swiftest_discard_tp) ---- */
    if (ntp > 0) {
        double *trb = (double *)calloc(3 * (size_t)ntp, sizeof(double));
        if (lcb || (qmin >= 0.0 && qmin_coord == 1)) {
            int nleft = 0;
            for (int i = 0; i < npl; i++) nleft += !gone[i];
            if (nleft > 0) { /* pl%h2b (swiftest_util.f90:228-265) */
                double g = GMcb, xt[3] = {0.0, 0.0, 0.0}, vt[3] = {0.0, 0.0, 0.0};
                for (int i = 0; i < npl; i++) {
                    if (gone[i] || pl_status[i] == SWO_INACTIVE) continue;
                    g = g + Gmass[i];
                    for (int d = 0; d < 3; d++) {
                        xt[d] = xt[d] + Gmass[i] * pl_rh[3 * i + d];
                        vt[d] = vt[d] + Gmass[i] * pl_vh[3 * i + d];
                    }
                }
                for (int d = 0; d < 3; d++) {
                    cb_rb[d] = -xt[d] / g;
                    cb_vb[d] = -vt[d] / g;
                }
                for (int i = 0; i < npl; i++)
                    if (!gone[i] && pl_status[i] != SWO_INACTIVE)
                        for (int d = 0; d < 3; d++) {
                            rb[3 * i + d] = pl_rh[3 * i + d] + cb_rb[d];
                            pl_vb[3 * i + d] = pl_vh[3 * i + d] + cb_vb[d];
                        }
            }
            for (int t = 0; t < ntp; t++) /* tp%h2b (:268-296) */
                if (tp_status[t] != SWO_INACTIVE)
                    for (int d = 0; d < 3; d++) {
                        trb[3 * t + d] = tp_rh[3 * t + d] + cb_rb[d];
                        tp_vb[3 * t + d] = tp_vh[3 * t + d] + cb_vb[d];
                    }
        }
        if (lcb) { /* swiftest_discard_cb_tp (swiftest_discard.f90:136-188) */
            const double rmin2 = fmax(rmin * rmin, Rcb * Rcb), rmax2 = rmax * rmax, rmaxu2 = rmaxu * rmaxu;
            for (int t = 0; t < ntp; t++) {
                if (tp_status[t] != SWO_ACTIVE) continue;
                const double *r = tp_rh + 3 * t, *b = trb + 3 * t, *v = tp_vb + 3 * t;
                const double rh2 = dot3(r[0], r[1], r[2], r[0], r[1], r[2]);
                int32_t st = SWO_ACTIVE;
                if (rmax >= 0.0 && rh2 > rmax2) {
                    st = SWO_DISCARDED_RMAX;
                } else if (rmin >= 0.0 && rh2 < rmin2) {
                    st = SWO_DISCARDED_RMIN;
                } else if (rmaxu >= 0.0) {
                    const double rb2 = dot3(b[0], b[1], b[2], b[0], b[1], b[2]);
                    const double vb2 = dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
                    const double energy = 0.5 * vb2 - Gmtot / sqrt(rb2);
                    if (energy > 0.0 && rb2 > rmaxu2) st = SWO_DISCARDED_RMAXU;
                }
                if (st != SWO_ACTIVE) {
                    tp_status[t] = st;
                    counts[1]++;
                }
            }
        }
        if (qmin >= 0.0) { /* swiftest_discard_peri_tp (:209-240) */
            for (int t = 0; t < ntp; t++) {
                if (qmin_coord == 0)
                    peri_one(GMcb, tp_rh + 3 * t, tp_vh + 3 * t, &tp_isperi[t], &tp_atp[t], &tp_peri[t]);
                else
                    peri_one(Gmtot, trb + 3 * t, tp_vb + 3 * t, &tp_isperi[t], &tp_atp[t], &tp_peri[t]);
            }
            for (int t = 0; t < ntp; t++) {
                if (tp_status[t] != SWO_ACTIVE || tp_isperi[t] != 0) continue;
                int ih = 1;
                for (int j = 0; j < npl; j++) {
                    if (gone[j]) continue;
                    double dx[3];
                    for (int d = 0; d < 3; d++) dx[d] = tp_rh[3 * t + d] - pl_rh[3 * j + d];
                    if (dot3(dx[0], dx[1], dx[2], dx[0], dx[1], dx[2]) <= rhill[j] * rhill[j]) ih = 0;
                }
                if (ih && tp_atp[t] >= alo && tp_atp[t] <= ahi && tp_peri[t] <= qmin) {
                    tp_status[t] = SWO_DISCARDED_PERI;
                    counts[1]++;
                }
            }
        }
        if (lclose) { /* swiftest_discard_pl_tp (:260-285) */
            for (int t = 0; t < ntp; t++) {
                if (tp_status[t] != SWO_ACTIVE) continue;
                for (int j = 0; j < npl; j++) {
                    if (gone[j]) continue;
                    double dx[3], dv[3];
                    for (int d = 0; d < 3; d++) {
                        dx[d] = tp_rh[3 * t + d] - pl_rh[3 * j + d];
                        dv[d] = tp_vh[3 * t + d] - pl_vh[3 * j + d];
                    }
                    if (pl_close(dx, dv, dt, radius[j] * radius[j])) {
                        tp_status[t] = SWO_DISCARDED_PLR;
                        counts[1]++;
                        break;
                    }
                }
            }
        }
        free(trb);
    }
    free(gone);
    free(rb);
}

int swo_sys_discard(swo_sys *s, const double *prm6, int qmin_coord, int64_t *counts)
{
    if (!s->dis_ready) { /* swiftest_util_setup_body (swiftest_util.f90:2537): isperi = 1 */
        const size_t np = s->npl > 0 ? s->npl : 1, nt = s->ntp > 0 ? s->ntp : 1;
        s->pisperi = (int32_t *)malloc(np * sizeof(int32_t));
        s->tisperi = (int32_t *)malloc(nt * sizeof(int32_t));
        s->patp = (double *)calloc(np, sizeof(double));
        s->pperi = (double *)calloc(np, sizeof(double));
        s->tatp = (double *)calloc(nt, sizeof(double));
        s->tperi = (double *)calloc(nt, sizeof(double));
        for (size_t i = 0; i < np; i++) s->pisperi[i] = 1;
        for (size_t i = 0; i < nt; i++) s->tisperi[i] = 1;
        s->lfirst_peri = 1;
        s->dis_ready = 1;
    }
    int32_t *nplenc = (int32_t *)calloc((size_t)(s->npl > 0 ? s->npl : 1), sizeof(int32_t));
    for (int64_t k = 0; k < s->plpl.n; k++) { /* symba_encounter_check.f90:79-80 */
        nplenc[s->plpl.i1[k] - 1]++;
        nplenc[s->plpl.i2[k] - 1]++;
    }
    double prm[7];
    for (int k = 0; k < 6; k++) prm[k] = prm6[k];
    prm[6] = s->p.dt;
    swo_discard(s->integ, s->npl, s->ntp, s->GMcb, s->Rcb, s->cb_vb, s->rh, s->vh, s->vb, s->Gmass, s->radius,
                s->rhill, s->status, nplenc, s->pisperi, s->patp, s->pperi, &s->lfirst_peri, s->trh, s->tvh, s->tvb,
                s->tstatus, s->tisperi, s->tatp, s->tperi, prm, qmin_coord, s->p.lclose, counts);
    /* discarded particles leave the arrays (tp%spill, swiftest_discard.f90:113-116) */
    for (int i = 0; i < s->ntp; i++)
        if (s->tstatus[i] != SWO_ACTIVE) s->tlmask[i] = 0;
    free(nplenc);
    return 0;
}
