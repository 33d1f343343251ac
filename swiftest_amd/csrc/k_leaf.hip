// k_leaf.hip -- leaf-operator entry points (the reference's numeric leaves
// with their host-array contract) and the WHM integrator.
//
// Leaf operators: swiftest_kick.f90:69-415, swiftest_drift.f90:60-138,
// encounter_check.f90:14-140.  WHM: whm_step.f90:14-104, whm_kick.f90:14-318,
// whm_coord.f90:14-124, whm_util.f90:117-228.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>
#include <vector>

#include <mutex>
#include <string>

#include "swg_engine.h"
#include "swg_math.h"

namespace swg {

static inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

#define HIPCHK(x)                                    \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return hipfail(e_, #x); \
    } while (0)

// ---------------------------------------------------------------------------
// leaf kernels (AoS in HBM, as the reference lays its (3,n) arrays out)
// ---------------------------------------------------------------------------

// int_one_pl contribution of pair (i, j) to body b's ahi (b == i) or ahj (b == j)
__device__ inline void one_pl_half(double rji2, double xr, double yr, double zr, double Gm_other, bool b_is_i,
                                   double &ax, double &ay, double &az, int fast)
{
    double fac;
    if (fast) {
        fac = gm_over_r3_fast(Gm_other, rji2);
    } else {
        double irij3 = 1.0 / (rji2 * sqrt(rji2));
        fac = Gm_other * irij3;
    }
    if (b_is_i) {
        ax = fma(fac, xr, ax);
        ay = fma(fac, yr, ay);
        az = fma(fac, zr, az);
    } else {
        ax = fma(-fac, xr, ax);
        ay = fma(-fac, yr, ay);
        az = fma(-fac, zr, az);
    }
}

// swiftest_kick_getacch_int_all_tri_{rad,norad}_pl (swiftest_kick.f90:165-371)
__global__ void k_leaf_tri(int npl, int nplm, const double *r, const double *Gm, const double *R, double *acc, int fast)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= npl) return;
    const int nplt = npl - nplm;
    const double xb = r[3 * b], yb = r[3 * b + 1], zb = r[3 * b + 2];
    if (nplt > nplm) {
        // triangular loops with the (ahi, ahj) reduction, one-thread order
        double hix = 0, hiy = 0, hiz = 0, hjx = 0, hjy = 0, hjz = 0;
        for (int i = 0; i < min(b, nplm); i++) {  // pairs (i, b): b plays j
            double xr = xb - r[3 * i], yr = yb - r[3 * i + 1], zr = zb - r[3 * i + 2];
            double rji2 = dot3(xr, yr, zr, xr, yr, zr);
            if (R) {
                double s = R[i] + R[b];
                if (!(rji2 > s * s)) continue;
            }
            one_pl_half(rji2, xr, yr, zr, Gm[i], false, hjx, hjy, hjz, fast);
        }
        if (b < nplm) {
            for (int j = b + 1; j < npl; j++) {  // pairs (b, j)
                double xr = r[3 * j] - xb, yr = r[3 * j + 1] - yb, zr = r[3 * j + 2] - zb;
                double rji2 = dot3(xr, yr, zr, xr, yr, zr);
                if (R) {
                    double s = R[b] + R[j];
                    if (!(rji2 > s * s)) continue;
                }
                one_pl_half(rji2, xr, yr, zr, Gm[j], true, hix, hiy, hiz, fast);
            }
        }
        acc[3 * b] = (acc[3 * b] + hix) + hjx;
        acc[3 * b + 1] = (acc[3 * b + 1] + hiy) + hjy;
        acc[3 * b + 2] = (acc[3 * b + 2] + hiz) + hjz;
        return;
    }
    double ax = acc[3 * b], ay = acc[3 * b + 1], az = acc[3 * b + 2];
    const int jend = b < nplm ? npl : nplm;
    for (int j = 0; j < jend; j++) {
        if (j == b) continue;
        double rx = r[3 * j] - xb, ry = r[3 * j + 1] - yb, rz = r[3 * j + 2] - zb;
        double rji2 = dot3(rx, ry, rz, rx, ry, rz);
        if (R) {
            double s = R[b] + R[j];
            if (!(rji2 > s * s)) continue;
        }
        double fac = fast ? gm_over_r3_fast(Gm[j], rji2) : gm_over_r3(Gm[j], rji2, pair_g_ok(Gm[j]));
        ax = fma(fac, rx, ax);
        ay = fma(fac, ry, ay);
        az = fma(fac, rz, az);
    }
    acc[3 * b] = ax;
    acc[3 * b + 1] = ay;
    acc[3 * b + 2] = az;
}

// swiftest_kick_getacch_int_all_flat_{rad,norad}_pl (swiftest_kick.f90:69-162)
// gathered per body: rows = pair ids with index1 == b, cols = index2 == b,
// each in list order (built on the host).
__global__ void k_leaf_flat(int npl, const int32_t *k, const double *r, const double *Gm, const double *R,
                            const int64_t *rs, const int64_t *cs, const int64_t *rowl, const int64_t *coll,
                            double *acc, int fast)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= npl) return;
    double hix = 0, hiy = 0, hiz = 0, hjx = 0, hjy = 0, hjz = 0;
    for (int64_t q = rs[b]; q < rs[b + 1]; q++) {
        int64_t p = rowl[q];
        int i = k[2 * p] - 1, j = k[2 * p + 1] - 1;
        double xr = r[3 * j] - r[3 * i], yr = r[3 * j + 1] - r[3 * i + 1], zr = r[3 * j + 2] - r[3 * i + 2];
        double rji2 = dot3(xr, yr, zr, xr, yr, zr);
        if (R) {
            double s = R[i] + R[j];
            if (!(rji2 > s * s)) continue;
        }
        one_pl_half(rji2, xr, yr, zr, Gm[j], true, hix, hiy, hiz, fast);
    }
    for (int64_t q = cs[b]; q < cs[b + 1]; q++) {
        int64_t p = coll[q];
        int i = k[2 * p] - 1, j = k[2 * p + 1] - 1;
        double xr = r[3 * j] - r[3 * i], yr = r[3 * j + 1] - r[3 * i + 1], zr = r[3 * j + 2] - r[3 * i + 2];
        double rji2 = dot3(xr, yr, zr, xr, yr, zr);
        if (R) {
            double s = R[i] + R[j];
            if (!(rji2 > s * s)) continue;
        }
        one_pl_half(rji2, xr, yr, zr, Gm[i], false, hjx, hjy, hjz, fast);
    }
    acc[3 * b] = (acc[3 * b] + hix) + hjx;
    acc[3 * b + 1] = (acc[3 * b + 1] + hiy) + hjy;
    acc[3 * b + 2] = (acc[3 * b + 2] + hiz) + hjz;
}

// swiftest_kick_getacch_int_all_tp (swiftest_kick.f90:374-415): particle sum
// from zero over pl ascending, added to acc once (OpenMP reduction).
template <int BLK>
__global__ __launch_bounds__(BLK) void k_leaf_tp(int ntp, int npl, const double *rtp, const double *rpl,
                                                 const double *Gm, const int32_t *lmask, double *acc, int fast)
{
    __shared__ double sx[BLK], sy[BLK], sz[BLK], sg[BLK];
    int t = blockIdx.x * BLK + threadIdx.x;
    bool valid = t < ntp && lmask[t] != 0;
    double xt = 0, yt = 0, zt = 0;
    if (valid) {
        xt = rtp[3 * t];
        yt = rtp[3 * t + 1];
        zt = rtp[3 * t + 2];
    }
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int j0 = 0; j0 < npl; j0 += BLK) {
        int j = j0 + threadIdx.x;
        if (j < npl) {
            sx[threadIdx.x] = rpl[3 * j];
            sy[threadIdx.x] = rpl[3 * j + 1];
            sz[threadIdx.x] = rpl[3 * j + 2];
            sg[threadIdx.x] = Gm[j];
        }
        __syncthreads();
        int jn = min(BLK, npl - j0);
        if (valid) {
            for (int q = 0; q < jn; q++) {
                double rx = xt - sx[q], ry = yt - sy[q], rz = zt - sz[q];
                double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                double fac = fast ? gm_over_r3_fast(sg[q], rji2) : gm_over_r3(sg[q], rji2, pair_g_ok(sg[q]));
                px = fma(-fac, rx, px);
                py = fma(-fac, ry, py);
                pz = fma(-fac, rz, pz);
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    acc[3 * t] = acc[3 * t] + px;
    acc[3 * t + 1] = acc[3 * t + 1] + py;
    acc[3 * t + 2] = acc[3 * t + 2] + pz;
}

// swiftest_drift_all (swiftest_drift.f90:60-108), GR off
__global__ __launch_bounds__(256) void k_leaf_drift(const double *mu, double *x, double *v, int n, double dt, const int32_t *lmask,
                             int32_t *iflag)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !lmask[i]) return;
    double rx = x[3 * i], ry = x[3 * i + 1], rz = x[3 * i + 2], vx = v[3 * i], vy = v[3 * i + 1], vz = v[3 * i + 2];
    iflag[i] = drift_one(mu[i], rx, ry, rz, vx, vy, vz, dt);
    x[3 * i] = rx;
    x[3 * i + 1] = ry;
    x[3 * i + 2] = rz;
    v[3 * i] = vx;
    v[3 * i + 1] = vy;
    v[3 * i + 2] = vz;
}

__global__ void k_aos_to_soa(const double *a, Vec3p s, int n)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    s.x[i] = a[3 * i];
    s.y[i] = a[3 * i + 1];
    s.z[i] = a[3 * i + 2];
}

__global__ void k_fill_i32(int32_t *p, int n, int32_t v)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_copy_i32_plus1(const int32_t *a, int32_t *b, const int8_t *c, int32_t *cc, int64_t n)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    b[i] = a[i] + 1;
    if (c) cc[i] = c[i];
}

// ---------------------------------------------------------------------------
// WHM (Jacobi) integrator: massive bodies are advanced by one thread per
// system (the reference's coordinate maps and ah2 are serial recurrences);
// test particles in parallel.
// ---------------------------------------------------------------------------
struct WhmArgs {
    PlDev pl;
    SysDev sy;
    WhmDev w;
    int32_t *tlfirst;
    double *ah0;  // 3*nsys
    int lclose;
    int fast;     // fast pair mode for the test-particle sums
};

// whm_coord_h2j_pl (whm_coord.f90:14-51)
__device__ void whm_h2j(const WhmArgs &a, int p0, int n)
{
    const PlDev &pl = a.pl;
    const WhmDev &w = a.w;
    w.xj.x[p0] = pl.rh.x[p0];
    w.xj.y[p0] = pl.rh.y[p0];
    w.xj.z[p0] = pl.rh.z[p0];
    w.vj.x[p0] = pl.vh.x[p0];
    w.vj.y[p0] = pl.vh.y[p0];
    w.vj.z[p0] = pl.vh.z[p0];
    double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (int q = 1; q < n; q++) {
        int i = p0 + q;
        const double *rh[3] = {pl.rh.x, pl.rh.y, pl.rh.z}, *vh[3] = {pl.vh.x, pl.vh.y, pl.vh.z};
        double *xj[3] = {w.xj.x, w.xj.y, w.xj.z}, *vj[3] = {w.vj.x, w.vj.y, w.vj.z};
        for (int d = 0; d < 3; d++) {
            sx[d] = sx[d] + pl.Gm[i - 1] * rh[d][i - 1];
            sv[d] = sv[d] + pl.Gm[i - 1] * vh[d][i - 1];
            double cap = sx[d] / w.eta[i - 1], capv = sv[d] / w.eta[i - 1];
            xj[d][i] = rh[d][i] - cap;
            vj[d][i] = vh[d][i] - capv;
        }
    }
}

// whm_coord_j2h_pl (whm_coord.f90:54-90)
__device__ void whm_j2h(const WhmArgs &a, int p0, int n)
{
    const PlDev &pl = a.pl;
    const WhmDev &w = a.w;
    pl.rh.x[p0] = w.xj.x[p0];
    pl.rh.y[p0] = w.xj.y[p0];
    pl.rh.z[p0] = w.xj.z[p0];
    pl.vh.x[p0] = w.vj.x[p0];
    pl.vh.y[p0] = w.vj.y[p0];
    pl.vh.z[p0] = w.vj.z[p0];
    double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
    for (int q = 1; q < n; q++) {
        int i = p0 + q;
        double *rh[3] = {pl.rh.x, pl.rh.y, pl.rh.z}, *vh[3] = {pl.vh.x, pl.vh.y, pl.vh.z};
        const double *xj[3] = {w.xj.x, w.xj.y, w.xj.z}, *vj[3] = {w.vj.x, w.vj.y, w.vj.z};
        for (int d = 0; d < 3; d++) {
            sx[d] = sx[d] + (pl.Gm[i - 1] * xj[d][i - 1]) / w.eta[i - 1];
            sv[d] = sv[d] + (pl.Gm[i - 1] * vj[d][i - 1]) / w.eta[i - 1];
            rh[d][i] = xj[d][i] + sx[d];
            vh[d][i] = vj[d][i] + sv[d];
        }
    }
}

// whm_kick_getacch_pl (whm_kick.f90:14-67): ir3h/ir3j, ah0, ah1, ah2, then the
// direct sum (swiftest_kick_getacch_int_pl, swiftest_kick.f90:12-43)
__device__ void whm_getacch_pl(const WhmArgs &a, int s, int p0, int n)
{
    const PlDev &pl = a.pl;
    const WhmDev &w = a.w;
    const double GMcb = a.sy.GMcb[s];
    for (int q = 0; q < n; q++) {
        int i = p0 + q;
        double r2 = dot3(pl.rh.x[i], pl.rh.y[i], pl.rh.z[i], pl.rh.x[i], pl.rh.y[i], pl.rh.z[i]);
        double ir = 1.0 / sqrt(r2);
        w.ir3h[i] = ir / r2;
        r2 = dot3(w.xj.x[i], w.xj.y[i], w.xj.z[i], w.xj.x[i], w.xj.y[i], w.xj.z[i]);
        ir = 1.0 / sqrt(r2);
        w.ir3j[i] = ir / r2;
    }
    double h0[3] = {0, 0, 0};
    for (int q = 1; q < n; q++) {
        int i = p0 + q;
        double x = pl.rh.x[i], y = pl.rh.y[i], z = pl.rh.z[i];
        double r2 = dot3(x, y, z, x, y, z);
        double fac = pl.Gm[i] * (1.0 / (r2 * sqrt(r2)));
        h0[0] = fma(-fac, x, h0[0]);
        h0[1] = fma(-fac, y, h0[1]);
        h0[2] = fma(-fac, z, h0[2]);
    }
    double *ah[3] = {w.ah.x, w.ah.y, w.ah.z};
    const double *xj[3] = {w.xj.x, w.xj.y, w.xj.z}, *rh[3] = {pl.rh.x, pl.rh.y, pl.rh.z};
    for (int q = 0; q < n; q++)
        for (int d = 0; d < 3; d++) ah[d][p0 + q] = ah[d][p0 + q] + h0[d];
    for (int q = 1; q < n; q++) {
        int i = p0 + q;
        if (pl.status[i] == INACTIVE) continue;
        for (int d = 0; d < 3; d++) {
            double a1j = xj[d][i] * w.ir3j[i];
            double a1h = rh[d][i] * w.ir3h[i];
            ah[d][i] = ah[d][i] + GMcb * (a1j - a1h);
        }
    }
    double etaj = GMcb, ah2o[3] = {0, 0, 0};
    for (int q = 1; q < n; q++) {
        int i = p0 + q;
        if (pl.status[i] == INACTIVE) continue;
        etaj = etaj + pl.Gm[i - 1];
        double fac = ((pl.Gm[i] * GMcb) * w.ir3j[i]) / etaj;
        for (int d = 0; d < 3; d++) {
            double ah2 = ah2o[d] + fac * xj[d][i];
            ah[d][i] = ah[d][i] + ah2;
            ah2o[d] = ah2;
        }
    }
    for (int q = 0; q < n; q++) {
        int i = p0 + q;
        double ax = ah[0][i], ay = ah[1][i], az = ah[2][i];
        for (int u = 0; u < n; u++) {
            int j = p0 + u;
            if (j == i) continue;
            double rx = rh[0][j] - rh[0][i], ry = rh[1][j] - rh[1][i], rz = rh[2][j] - rh[2][i];
            double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            if (a.lclose) {
                double rs = pl.R[i] + pl.R[j];
                if (!(rji2 > rs * rs)) continue;
            }
            double fac = gm_over_r3(pl.Gm[j], rji2, pair_g_ok(pl.Gm[j]));
            ax = fma(fac, rx, ax);
            ay = fma(fac, ry, ay);
            az = fma(fac, rz, az);
        }
        ah[0][i] = ax;
        ah[1][i] = ay;
        ah[2][i] = az;
    }
}

// whm_step_pl with whm_kick_vh_pl and whm_drift_pl (whm_step.f90:38-69,
// whm_kick.f90:218-262, whm_drift.f90:14-58); one thread per system.
__global__ void k_whm_pl(WhmArgs a, int nsys, double dt)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    const PlDev &pl = a.pl;
    const WhmDev &w = a.w;
    int p0 = a.sy.pl_off[s], n = a.sy.pl_off[s + 1] - p0;
    a.tlfirst[s] = a.sy.lfirst[s];
    if (n == 0) return;
    const double dth = 0.5 * dt;
    if (a.sy.lfirst[s]) {
        whm_h2j(a, p0, n);
        for (int q = 0; q < n; q++) w.ah.x[p0 + q] = w.ah.y[p0 + q] = w.ah.z[p0 + q] = 0.0;
        whm_getacch_pl(a, s, p0, n);
        a.sy.lfirst[s] = 0;
    }
    for (int q = 0; q < n; q++) {
        int i = p0 + q;
        pl.rbeg.x[i] = pl.rh.x[i];
        pl.rbeg.y[i] = pl.rh.y[i];
        pl.rbeg.z[i] = pl.rh.z[i];
        if (pl.status[i] != INACTIVE) {
            pl.vh.x[i] = fma(w.ah.x[i], dth, pl.vh.x[i]);
            pl.vh.y[i] = fma(w.ah.y[i], dth, pl.vh.y[i]);
            pl.vh.z[i] = fma(w.ah.z[i], dth, pl.vh.z[i]);
        }
    }
    // whm_coord_vh2vj_pl (whm_coord.f90:93-124)
    w.vj.x[p0] = pl.vh.x[p0];
    w.vj.y[p0] = pl.vh.y[p0];
    w.vj.z[p0] = pl.vh.z[p0];
    {
        double sv[3] = {0, 0, 0};
        const double *vh[3] = {pl.vh.x, pl.vh.y, pl.vh.z};
        double *vj[3] = {w.vj.x, w.vj.y, w.vj.z};
        for (int q = 1; q < n; q++) {
            int i = p0 + q;
            for (int d = 0; d < 3; d++) {
                sv[d] = sv[d] + pl.Gm[i - 1] * vh[d][i - 1];
                vj[d][i] = vh[d][i] - sv[d] / w.eta[i - 1];
            }
        }
    }
    for (int q = 0; q < n; q++) {
        int i = p0 + q;
        if (pl.status[i] == INACTIVE) continue;
        double rx = w.xj.x[i], ry = w.xj.y[i], rz = w.xj.z[i], vx = w.vj.x[i], vy = w.vj.y[i], vz = w.vj.z[i];
        int fl = drift_one(w.muj[i], rx, ry, rz, vx, vy, vz, dt);
        w.xj.x[i] = rx;
        w.xj.y[i] = ry;
        w.xj.z[i] = rz;
        w.vj.x[i] = vx;
        w.vj.y[i] = vy;
        w.vj.z[i] = vz;
        if (fl) {
            pl.status[i] = DISCARDED_DRIFTERR;
            a.sy.err[s] |= 2;
            return;
        }
    }
    whm_j2h(a, p0, n);
    for (int q = 0; q < n; q++) w.ah.x[p0 + q] = w.ah.y[p0 + q] = w.ah.z[p0 + q] = 0.0;
    whm_getacch_pl(a, s, p0, n);
    for (int q = 0; q < n; q++) {
        int i = p0 + q;
        pl.rend.x[i] = pl.rh.x[i];
        pl.rend.y[i] = pl.rh.y[i];
        pl.rend.z[i] = pl.rh.z[i];
        if (pl.status[i] != INACTIVE) {
            pl.vh.x[i] = fma(w.ah.x[i], dth, pl.vh.x[i]);
            pl.vh.y[i] = fma(w.ah.y[i], dth, pl.vh.y[i]);
            pl.vh.z[i] = fma(w.ah.z[i], dth, pl.vh.z[i]);
        }
    }
}

// ah0 of the test particles (whm_kick.f90:82-104): -sum Gm r/|r|^3 over all pl
// at rbeg or rend, one thread per system
__global__ void k_whm_tp_ah0(WhmArgs a, int nsys, int lbeg)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    int p0 = a.sy.pl_off[s], n = a.sy.pl_off[s + 1] - p0;
    Vec3p rp = lbeg ? a.pl.rbeg : a.pl.rend;
    double h0[3] = {0, 0, 0};
    for (int q = 0; q < n; q++) {
        int i = p0 + q;
        double x = rp.x[i], y = rp.y[i], z = rp.z[i];
        double r2 = dot3(x, y, z, x, y, z);
        double fac = a.pl.Gm[i] * (1.0 / (r2 * sqrt(r2)));
        h0[0] = fma(-fac, x, h0[0]);
        h0[1] = fma(-fac, y, h0[1]);
        h0[2] = fma(-fac, z, h0[2]);
    }
    for (int d = 0; d < 3; d++) a.ah0[3 * s + d] = h0[d];
}

// whm_kick_vh_tp (whm_kick.f90:265-318): at the start of a step the
// accelerations of the previous step's end are reused, except on the first
// step; ah = (0 + ah0) + sum over pl; vh += ah*dt.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_whm_tp_kick(WhmArgs a, TpDev tp, const Tile *tiles, double dt, int lbeg)
{
    __shared__ double sx[BLK], sy[BLK], sz[BLK], sg[BLK];
    const Tile tl = tiles[blockIdx.x];
    const int s = tl.s;
    const int p0 = a.sy.pl_off[s], p1 = a.sy.pl_off[s + 1];
    const int j = tl.b0 + (int)threadIdx.x;
    const bool valid = j < tl.b1 && tp.status[j] == ACTIVE;
    const bool recompute = !lbeg || a.tlfirst[s];
    if (!recompute) {
        if (valid) {
            tp.vh.x[j] = fma(a.w.tah.x[j], dt, tp.vh.x[j]);
            tp.vh.y[j] = fma(a.w.tah.y[j], dt, tp.vh.y[j]);
            tp.vh.z[j] = fma(a.w.tah.z[j], dt, tp.vh.z[j]);
        }
        return;
    }
    Vec3p rp = lbeg ? a.pl.rbeg : a.pl.rend;
    double xt = 0, yt = 0, zt = 0;
    if (valid) {
        xt = tp.rh.x[j];
        yt = tp.rh.y[j];
        zt = tp.rh.z[j];
    }
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int it = p0; it < p1; it += BLK) {
        int q0 = it + (int)threadIdx.x;
        if (q0 < p1) {
            sx[threadIdx.x] = rp.x[q0];
            sy[threadIdx.x] = rp.y[q0];
            sz[threadIdx.x] = rp.z[q0];
            sg[threadIdx.x] = a.pl.Gm[q0];
        }
        __syncthreads();
        const int qn = min(BLK, p1 - it);
        if (valid) {
            for (int q = 0; q < qn; q++) {
                double rx = xt - sx[q], ry = yt - sy[q], rz = zt - sz[q];
                double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                double fac = a.fast ? gm_over_r3_fast(sg[q], rji2) : gm_over_r3(sg[q], rji2, pair_g_ok(sg[q]));
                px = fma(-fac, rx, px);
                py = fma(-fac, ry, py);
                pz = fma(-fac, rz, pz);
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    double ax = (0.0 + a.ah0[3 * s]) + px, ay = (0.0 + a.ah0[3 * s + 1]) + py, az = (0.0 + a.ah0[3 * s + 2]) + pz;
    a.w.tah.x[j] = ax;
    a.w.tah.y[j] = ay;
    a.w.tah.z[j] = az;
    tp.vh.x[j] = fma(ax, dt, tp.vh.x[j]);
    tp.vh.y[j] = fma(ay, dt, tp.vh.y[j]);
    tp.vh.z[j] = fma(az, dt, tp.vh.z[j]);
}

__global__ void k_whm_h2j_all(WhmArgs a, int nsys)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    int p0 = a.sy.pl_off[s], n = a.sy.pl_off[s + 1] - p0;
    if (n) whm_h2j(a, p0, n);
}

// ---------------------------------------------------------------------------
// Whole WHM steps in one launch (systems of at most WR_N massive bodies).
// One workgroup of two wavefronts per tile: a system and up to 64 of its
// test particles.  The first wavefront integrates the system's massive
// bodies, one lane per body (every tile of a system computes the same values
// from the same inputs; its first tile writes them back); the second
// integrates the tile's particles, one lane each.  They run one step apart:
// while the particles take step k, which needs the bodies' positions at the
// start and at the end of step k, the bodies already take step k+1.  The
// positions after each step pass through three LDS buffers; one workgroup
// barrier per step.  Each sum over the bodies in order (the Jacobi maps'
// prefixes, ah0, ah2) is formed by each lane from the terms the bodies' lanes
// put in LDS, added in the order of whm_coord.f90 / whm_kick.f90, so every
// value equals the one the per-step kernels above (and the oracle) produce.
// ---------------------------------------------------------------------------
constexpr int WR_N = 64;

// ordering of LDS accesses between the lanes of one wavefront
__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ __launch_bounds__(128) void k_whm_run(WhmArgs a, TpDev tp, const int4 *tiles, double dt, int nsteps)
{
    __shared__ double sx[WR_N], sy[WR_N], sz[WR_N], su[WR_N], sv[WR_N], sw[WR_N];
    __shared__ double sg[WR_N], seta[WR_N], sf[WR_N], sr[WR_N];
    __shared__ int sact[WR_N];
    // positions (and ah0 factors) of the bodies after j steps, in slot j % 3
    __shared__ double bx[3][WR_N], by[3][WR_N], bz[3][WR_N], bf[3][WR_N];
    __shared__ int sfail;
    const int4 tl = tiles[blockIdx.x];
    const int s = tl.x, lane = threadIdx.x & 63;
    const bool plw = threadIdx.x < 64;  // the bodies' wavefront
    const bool writer = tl.w != 0;
    const PlDev &pl = a.pl;
    const WhmDev &w = a.w;
    const int p0 = a.sy.pl_off[s], n = a.sy.pl_off[s + 1] - p0;
    const double GMcb = a.sy.GMcb[s], dth = 0.5 * dt;
    const int lfirst0 = a.sy.lfirst[s];
    if (threadIdx.x == 0) sfail = 0;

    // ---- the bodies' wavefront: state in registers, one lane per body ----
    const bool isb = plw && lane < n;
    const int i = p0 + lane;
    double rh[3] = {0, 0, 0}, vh[3] = {0, 0, 0}, xj[3] = {0, 0, 0}, vj[3] = {0, 0, 0}, ah[3] = {0, 0, 0};
    double rb[3] = {0, 0, 0}, re[3] = {0, 0, 0};
    double gm = 0.0, muj = 0.0, ir3h = 0.0, ir3j = 0.0;
    int st = INACTIVE, lfirst = lfirst0;
    if (isb) {
        rh[0] = pl.rh.x[i], rh[1] = pl.rh.y[i], rh[2] = pl.rh.z[i];
        vh[0] = pl.vh.x[i], vh[1] = pl.vh.y[i], vh[2] = pl.vh.z[i];
        xj[0] = w.xj.x[i], xj[1] = w.xj.y[i], xj[2] = w.xj.z[i];
        vj[0] = w.vj.x[i], vj[1] = w.vj.y[i], vj[2] = w.vj.z[i];
        ah[0] = w.ah.x[i], ah[1] = w.ah.y[i], ah[2] = w.ah.z[i];
        rb[0] = pl.rbeg.x[i], rb[1] = pl.rbeg.y[i], rb[2] = pl.rbeg.z[i];
        re[0] = pl.rend.x[i], re[1] = pl.rend.y[i], re[2] = pl.rend.z[i];
        gm = pl.Gm[i];
        muj = w.muj[i];
        ir3h = w.ir3h[i];
        ir3j = w.ir3j[i];
        st = pl.status[i];
        sg[lane] = gm;
        seta[lane] = w.eta[i];
        sr[lane] = pl.R[i];
        sact[lane] = st != INACTIVE;
    }
    // positions and ah0 factor of this lane's body into slot q
    auto publish = [&](int q) __attribute__((always_inline)) {
        if (!isb) return;
        double r2 = dot3(rh[0], rh[1], rh[2], rh[0], rh[1], rh[2]);
        bx[q][lane] = rh[0];
        by[q][lane] = rh[1];
        bz[q][lane] = rh[2];
        bf[q][lane] = gm * (1.0 / (r2 * sqrt(r2)));
    };
    // ah of the massive bodies at rh / xj, added to ah (whm_kick_getacch_pl,
    // as whm_getacch_pl above)
    auto getacch = [&]() __attribute__((always_inline)) {
        double r2 = dot3(rh[0], rh[1], rh[2], rh[0], rh[1], rh[2]);
        double ir = 1.0 / sqrt(r2);
        ir3h = ir / r2;
        double r2j = dot3(xj[0], xj[1], xj[2], xj[0], xj[1], xj[2]);
        ir = 1.0 / sqrt(r2j);
        ir3j = ir / r2j;
        wave_sync();
        if (isb) {
            sf[lane] = gm * (1.0 / (r2 * sqrt(r2)));
            sx[lane] = rh[0];
            sy[lane] = rh[1];
            sz[lane] = rh[2];
        }
        wave_sync();
        double fac = 0.0;
        if (isb) {
            double h0[3] = {0, 0, 0};
#pragma unroll 8
            for (int q = 1; q < n; q++) {
                h0[0] = fma(-sf[q], sx[q], h0[0]);
                h0[1] = fma(-sf[q], sy[q], h0[1]);
                h0[2] = fma(-sf[q], sz[q], h0[2]);
            }
            for (int d = 0; d < 3; d++) ah[d] = ah[d] + h0[d];
            if (lane >= 1 && st != INACTIVE)
                for (int d = 0; d < 3; d++) {
                    double a1j = xj[d] * ir3j;
                    double a1h = rh[d] * ir3h;
                    ah[d] = ah[d] + GMcb * (a1j - a1h);
                }
            // ah2: the running sums over the active bodies 1..lane
            double etaj = GMcb;
            for (int k = 1; k <= lane; k++)
                if (sact[k]) etaj = etaj + sg[k - 1];
            fac = ((gm * GMcb) * ir3j) / etaj;
            su[lane] = fac * xj[0];
            sv[lane] = fac * xj[1];
            sw[lane] = fac * xj[2];
        }
        wave_sync();
        if (!isb) return;
        if (lane >= 1 && st != INACTIVE) {
            double ah2o[3] = {0, 0, 0};
            for (int k = 1; k <= lane; k++) {
                if (!sact[k]) continue;
                ah2o[0] = ah2o[0] + su[k];
                ah2o[1] = ah2o[1] + sv[k];
                ah2o[2] = ah2o[2] + sw[k];
            }
            for (int d = 0; d < 3; d++) ah[d] = ah[d] + ah2o[d];
        }
        double ax = ah[0], ay = ah[1], az = ah[2];
#pragma unroll 8
        for (int u = 0; u < n; u++) {
            if (u == lane) continue;
            double rx = sx[u] - rh[0], ry = sy[u] - rh[1], rz = sz[u] - rh[2];
            double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            if (a.lclose) {
                double rs = sr[lane] + sr[u];
                if (!(rji2 > rs * rs)) continue;
            }
            double f = gm_over_r3(sg[u], rji2, pair_g_ok(sg[u]));
            ax = fma(f, rx, ax);
            ay = fma(f, ry, ay);
            az = fma(f, rz, az);
        }
        ah[0] = ax;
        ah[1] = ay;
        ah[2] = az;
    };
    // one step of the bodies (whm_step_pl); publishes the end positions into
    // slot q; returns false on a failed drift (fatal: whm_drift.f90:38-53,
    // the first failing body is marked and the step stops there)
    auto pl_step = [&](int q) __attribute__((always_inline)) -> bool {
        if (lfirst) {
            // whm_coord_h2j_pl
            wave_sync();
            if (isb) {
                sx[lane] = gm * rh[0], sy[lane] = gm * rh[1], sz[lane] = gm * rh[2];
                su[lane] = gm * vh[0], sv[lane] = gm * vh[1], sw[lane] = gm * vh[2];
            }
            wave_sync();
            if (isb) {
                for (int d = 0; d < 3; d++) xj[d] = rh[d], vj[d] = vh[d];
                if (lane > 0) {
                    double cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
                    for (int k = 0; k < lane; k++) {
                        cx[0] = cx[0] + sx[k], cx[1] = cx[1] + sy[k], cx[2] = cx[2] + sz[k];
                        cv[0] = cv[0] + su[k], cv[1] = cv[1] + sv[k], cv[2] = cv[2] + sw[k];
                    }
                    for (int d = 0; d < 3; d++) {
                        double cap = cx[d] / seta[lane - 1], capv = cv[d] / seta[lane - 1];
                        xj[d] = rh[d] - cap;
                        vj[d] = vh[d] - capv;
                    }
                }
                ah[0] = ah[1] = ah[2] = 0.0;
            }
            getacch();
            lfirst = 0;
        }
        if (isb) {
            for (int d = 0; d < 3; d++) rb[d] = rh[d];
            if (st != INACTIVE)
                for (int d = 0; d < 3; d++) vh[d] = fma(ah[d], dth, vh[d]);
        }
        // whm_coord_vh2vj_pl
        wave_sync();
        if (isb) su[lane] = gm * vh[0], sv[lane] = gm * vh[1], sw[lane] = gm * vh[2];
        wave_sync();
        int fl = 0;
        if (isb) {
            if (lane == 0) {
                for (int d = 0; d < 3; d++) vj[d] = vh[d];
            } else {
                double cv[3] = {0, 0, 0};
                for (int k = 0; k < lane; k++) cv[0] = cv[0] + su[k], cv[1] = cv[1] + sv[k], cv[2] = cv[2] + sw[k];
                for (int d = 0; d < 3; d++) vj[d] = vh[d] - cv[d] / seta[lane - 1];
            }
            if (st != INACTIVE) fl = drift_one(muj, xj[0], xj[1], xj[2], vj[0], vj[1], vj[2], dt);
        }
        const unsigned long long bad = __ballot(fl != 0);
        if (bad) {
            if (lane == (int)__builtin_ctzll(bad)) st = DISCARDED_DRIFTERR;
            return false;
        }
        // whm_coord_j2h_pl
        wave_sync();
        if (isb) {
            su[lane] = (gm * xj[0]) / seta[lane];
            sv[lane] = (gm * xj[1]) / seta[lane];
            sw[lane] = (gm * xj[2]) / seta[lane];
        }
        wave_sync();
        if (isb) {
            double cx[3] = {0, 0, 0};
            for (int k = 0; k < lane; k++) cx[0] = cx[0] + su[k], cx[1] = cx[1] + sv[k], cx[2] = cx[2] + sw[k];
            if (lane == 0) {
                for (int d = 0; d < 3; d++) rh[d] = xj[d], vh[d] = vj[d];
            } else {
                for (int d = 0; d < 3; d++) rh[d] = xj[d] + cx[d];
            }
        }
        wave_sync();
        if (isb) {
            su[lane] = (gm * vj[0]) / seta[lane];
            sv[lane] = (gm * vj[1]) / seta[lane];
            sw[lane] = (gm * vj[2]) / seta[lane];
        }
        wave_sync();
        if (isb && lane > 0) {
            double cv[3] = {0, 0, 0};
            for (int k = 0; k < lane; k++) cv[0] = cv[0] + su[k], cv[1] = cv[1] + sv[k], cv[2] = cv[2] + sw[k];
            for (int d = 0; d < 3; d++) vh[d] = vj[d] + cv[d];
        }
        if (isb) ah[0] = ah[1] = ah[2] = 0.0;
        getacch();
        if (isb) {
            for (int d = 0; d < 3; d++) re[d] = rh[d];
            if (st != INACTIVE)
                for (int d = 0; d < 3; d++) vh[d] = fma(ah[d], dth, vh[d]);
        }
        publish(q);
        return true;
    };

    // ---- the particles' wavefront ----
    const int j = tl.y + lane;
    const bool ist = !plw && j < tl.z;
    double trh[3] = {0, 0, 0}, tvh[3] = {0, 0, 0}, tah[3] = {0, 0, 0};
    int tst = INACTIVE;
    if (ist) {
        trh[0] = tp.rh.x[j], trh[1] = tp.rh.y[j], trh[2] = tp.rh.z[j];
        tvh[0] = tp.vh.x[j], tvh[1] = tp.vh.y[j], tvh[2] = tp.vh.z[j];
        tah[0] = w.tah.x[j], tah[1] = w.tah.y[j], tah[2] = w.tah.z[j];
        tst = tp.status[j];
    }
    // kick with ah0 and the direct sum at the positions in slot q (as
    // k_whm_tp_ah0 / k_whm_tp_kick)
    auto tp_kick = [&](int q, bool recompute) __attribute__((always_inline)) {
        if (!ist || tst != ACTIVE) return;
        if (recompute) {
            double h0[3] = {0, 0, 0};
#pragma unroll 8
            for (int u = 0; u < n; u++) {
                h0[0] = fma(-bf[q][u], bx[q][u], h0[0]);
                h0[1] = fma(-bf[q][u], by[q][u], h0[1]);
                h0[2] = fma(-bf[q][u], bz[q][u], h0[2]);
            }
            double px = 0.0, py = 0.0, pz = 0.0;
#pragma unroll 8
            for (int u = 0; u < n; u++) {
                double rx = trh[0] - bx[q][u], ry = trh[1] - by[q][u], rz = trh[2] - bz[q][u];
                double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                double f = a.fast ? gm_over_r3_fast(sg[u], rji2) : gm_over_r3(sg[u], rji2, pair_g_ok(sg[u]));
                px = fma(-f, rx, px);
                py = fma(-f, ry, py);
                pz = fma(-f, rz, pz);
            }
            tah[0] = (0.0 + h0[0]) + px;
            tah[1] = (0.0 + h0[1]) + py;
            tah[2] = (0.0 + h0[2]) + pz;
        }
        for (int d = 0; d < 3; d++) tvh[d] = fma(tah[d], dth, tvh[d]);
    };

    // The two roles branch once, so that the registers of one are not held
    // by the other; both pass the same workgroup barriers: one before the
    // first particle step, one after each.
    if (plw) {
        // prologue: the bodies' positions before step 0 (slot 0) and after
        // it (slot 1)
        if (n && nsteps > 0) {
            publish(0);
            if (!pl_step(1) && lane == 0) sfail = 1;
        }
        __syncthreads();
        int done = 0;  // particle steps taken
        for (int k = 0; k < nsteps && !sfail; k++) {
            if (n && k + 1 < nsteps && !pl_step((k + 2) % 3) && lane == 0) sfail = 2;
            done = k + 1;
            __syncthreads();
        }
        const int fail = sfail;
        if (writer) {
            if (isb) {
                pl.rh.x[i] = rh[0], pl.rh.y[i] = rh[1], pl.rh.z[i] = rh[2];
                pl.vh.x[i] = vh[0], pl.vh.y[i] = vh[1], pl.vh.z[i] = vh[2];
                w.xj.x[i] = xj[0], w.xj.y[i] = xj[1], w.xj.z[i] = xj[2];
                w.vj.x[i] = vj[0], w.vj.y[i] = vj[1], w.vj.z[i] = vj[2];
                w.ah.x[i] = ah[0], w.ah.y[i] = ah[1], w.ah.z[i] = ah[2];
                pl.rbeg.x[i] = rb[0], pl.rbeg.y[i] = rb[1], pl.rbeg.z[i] = rb[2];
                pl.rend.x[i] = re[0], pl.rend.y[i] = re[1], pl.rend.z[i] = re[2];
                w.ir3h[i] = ir3h;
                w.ir3j[i] = ir3j;
                pl.status[i] = st;
            }
            if (lane == 0 && nsteps > 0) {
                a.sy.lfirst[s] = lfirst;
                // the first-kick flag of the last step begun: lfirst0 for
                // step 0 (and for every step when no body clears lfirst)
                const int last = fail == 1 ? 0 : (fail == 2 ? done : done - 1);
                a.tlfirst[s] = (n == 0 || last == 0) ? lfirst0 : 0;
                if (fail) a.sy.err[s] |= 2;
            }
        }
    } else {
        __syncthreads();
        for (int k = 0; k < nsteps && !sfail; k++) {
            // the particles' step k: kick at the start positions (reusing the
            // last step's accelerations unless this is the first kick),
            // drift, kick at the end positions
            const bool tlf = (k == 0 || n == 0) ? lfirst0 != 0 : false;
            tp_kick(k % 3, tlf);
            if (ist && tst == ACTIVE) {
                int f = drift_one(GMcb, trh[0], trh[1], trh[2], tvh[0], tvh[1], tvh[2], dt);
                if (f) tst = DISCARDED_DRIFTERR;
            }
            tp_kick((k + 1) % 3, true);
            __syncthreads();
        }
        if (ist) {
            tp.rh.x[j] = trh[0], tp.rh.y[j] = trh[1], tp.rh.z[j] = trh[2];
            tp.vh.x[j] = tvh[0], tp.vh.y[j] = tvh[1], tp.vh.z[j] = tvh[2];
            w.tah.x[j] = tah[0], w.tah.y[j] = tah[1], w.tah.z[j] = tah[2];
            tp.status[j] = tst;
        }
    }
}

// ---------------------------------------------------------------------------
// Ensemble: WHM set-up, step and snapshot
// ---------------------------------------------------------------------------
static WhmArgs whm_args(Ensemble &e)
{
    WhmArgs a;
    a.pl = e.pl;
    a.sy = e.sy;
    a.w = e.whm;
    a.tlfirst = e.sy.flag_pt;  // per-system scratch flag
    a.ah0 = e.sy.ptend;        // per-system 3-vector scratch
    a.lclose = e.prm.lclose;
    a.fast = e.tune.pair_fast;
    return a;
}

int Ensemble::whm_setup()
{
    if (alloc_vec(whm.xj, npl) || alloc_vec(whm.vj, npl) || alloc_vec(whm.ah, npl) || alloc_vec(whm.tah, ntp) ||
        own(&whm.eta, npl) || own(&whm.muj, npl) || own(&whm.ir3h, npl) || own(&whm.ir3j, npl))
        return SWG_ERR_HIP;
    // whm_util_set_mu_eta_pl (whm_util.f90:175-198), per system
    std::vector<double> eta(npl), muj(npl);
    for (int s = 0; s < nsys; s++) {
        int a = h_pl_off[s], n = h_pl_off[s + 1] - a;
        if (!n) continue;
        eta[a] = h_GMcb[s] + h_Gm[a];
        muj[a] = eta[a];
        for (int q = 1; q < n; q++) {
            eta[a + q] = eta[a + q - 1] + h_Gm[a + q];
            muj[a + q] = (h_GMcb[s] * eta[a + q]) / eta[a + q - 1];
        }
    }
    HIPCHK(copy_wait(stream, whm.eta, eta.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, whm.muj, muj.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> z(3 * (size_t)std::max(npl, ntp) + 3, 0.0);
    for (double *p : {whm.ah.x, whm.ah.y, whm.ah.z, whm.xj.x, whm.xj.y, whm.xj.z, whm.vj.x, whm.vj.y, whm.vj.z})
        HIPCHK(copy_wait(stream, p, z.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    for (double *p : {whm.tah.x, whm.tah.y, whm.tah.z})
        HIPCHK(copy_wait(stream, p, z.data(), ntp * sizeof(double), hipMemcpyHostToDevice));
    // tiles of the whole-step kernel: every system with at most WR_N massive
    // bodies qualifies; each gets ceil(ntp/64) tiles (one if it has none),
    // the first of which writes the massive bodies back.  SWG_WHM_STEPWISE=1
    // keeps the per-step kernels.
    whm_fused = true;
    for (int s = 0; s < nsys; s++) whm_fused = whm_fused && (h_pl_off[s + 1] - h_pl_off[s]) <= WR_N;
    const char *sw = getenv("SWG_WHM_STEPWISE");
    if (sw && sw[0] == '1') whm_fused = false;
    if (whm_fused) {
        std::vector<int4> t;
        for (int s = 0; s < nsys; s++) {
            const int b = h_tp_off[s], e = h_tp_off[s + 1];
            int k = 0;
            do {
                t.push_back(make_int4(s, b + 64 * k, std::min(b + 64 * (k + 1), e), k == 0));
                k++;
            } while (b + 64 * k < e);
        }
        HIPCHK(hipMalloc((void **)&d_tiles_whm, t.size() * sizeof(int4)));
        owned.push_back((void *)d_tiles_whm);
        HIPCHK(copy_wait(stream, d_tiles_whm, t.data(), t.size() * sizeof(int4), hipMemcpyHostToDevice));
        ntiles_whm = (int)t.size();
    }
    return 0;
}

// nsteps whole WHM steps in one launch (k_whm_run)
int Ensemble::run_whm_fused(int nsteps)
{
    WhmArgs a = whm_args(*this);
    k_whm_run<<<ntiles_whm, 128, 0, stream>>>(a, tp, d_tiles_whm, prm.dt, nsteps);
    HIPCHK(hipGetLastError());
    return 0;
}

int Ensemble::step_whm()
{
    hipStream_t st = stream;
    const double dt = prm.dt, dth = 0.5 * dt;
    WhmArgs a = whm_args(*this);
    k_whm_pl<<<cdiv(nsys, 64), 64, 0, st>>>(a, nsys, dt);
    k_whm_tp_ah0<<<cdiv(nsys, 64), 64, 0, st>>>(a, nsys, 1);
    if (ntiles_tp) k_whm_tp_kick<256><<<ntiles_tp, 256, 0, st>>>(a, tp, d_tiles_tp, dth, 1);
    launch_drift(st, tp.rh, tp.vh, tp.status, tp.levelg, tp.sys, sy.GMcb, sy.err, sy.frozen, ntp, dt, -1, 0, 1);
    k_whm_tp_ah0<<<cdiv(nsys, 64), 64, 0, st>>>(a, nsys, 0);
    if (ntiles_tp) k_whm_tp_kick<256><<<ntiles_tp, 256, 0, st>>>(a, tp, d_tiles_tp, dth, 0);
    HIPCHK(hipGetLastError());
    return 0;
}

int Ensemble::whm_snapshot()
{
    WhmArgs a = whm_args(*this);
    k_whm_h2j_all<<<cdiv(nsys, 64), 64, 0, stream>>>(a, nsys);
    HIPCHK(hipGetLastError());
    if (!prm.rhill_present) {
        std::vector<double> rh(3 * (size_t)npl), vh(3 * (size_t)npl), rhill(npl);
        HIPCHK(hipStreamSynchronize(stream));
        std::vector<double> cx(npl), cy(npl), cz(npl);
        for (int which = 0; which < 2; which++) {
            Vec3p v = which == 0 ? pl.rh : pl.vh;
            std::vector<double> &o = which == 0 ? rh : vh;
            HIPCHK(copy_wait(stream, cx.data(), v.x, npl * sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK(copy_wait(stream, cy.data(), v.y, npl * sizeof(double), hipMemcpyDeviceToHost));
            HIPCHK(copy_wait(stream, cz.data(), v.z, npl * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 0; i < npl; i++) {
                o[3 * i] = cx[i];
                o[3 * i + 1] = cy[i];
                o[3 * i + 2] = cz[i];
            }
        }
        for (int s = 0; s < nsys; s++) {
            int b = h_pl_off[s], n = h_pl_off[s + 1] - b;
            host_set_rhill(h_GMcb[s], n, &rh[3 * (size_t)b], &vh[3 * (size_t)b], &h_Gm[b], &rhill[b]);
        }
        HIPCHK(copy_wait(stream, pl.rhill, rhill.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
}

// ---------------------------------------------------------------------------
// leaf scratch
// ---------------------------------------------------------------------------
LeafState::~LeafState()
{
    for (void *p : bufs)
        if (p) (void)hipFree(p);
    void *q[] = {L.words, L.wpre, L.woff, L.nb, L.rowcnt, L.rowstart, L.i1, L.i2, L.lvdotr, L.status, L.level};
    for (void *p : q)
        if (p) (void)hipFree(p);
}

void *LeafState::get(int slot, size_t bytes, int *rc)
{
    if ((int)bufs.size() <= slot) {
        bufs.resize(slot + 1, nullptr);
        sizes.resize(slot + 1, 0);
    }
    if (sizes[slot] < bytes) {
        if (bufs[slot]) (void)hipFree(bufs[slot]);
        bufs[slot] = nullptr;
        sizes[slot] = 0;
        hipError_t e = hipMalloc(&bufs[slot], bytes ? bytes : 8);
        if (e != hipSuccess) {
            *rc = hipfail(e, "hipMalloc (leaf scratch)");
            return nullptr;
        }
        sizes[slot] = bytes;
    }
    return bufs[slot];
}

// ---------------------------------------------------------------------------
// The encounter check that allocates nothing (swg_encounter_check_all_plpl_
// noalloc): every device buffer of the check lives in one file-scope device
// array, laid out for at most SWG_NOALLOC_MAX_NPL bodies and
// SWG_NOALLOC_MAX_PAIRS listed pairs.
// ---------------------------------------------------------------------------

constexpr size_t na_align(size_t b) { return (b + 255) / 256 * 256; }
constexpr int NA_NPL = SWG_NOALLOC_MAX_NPL;
constexpr int64_t NA_PAIRS = SWG_NOALLOC_MAX_PAIRS;
constexpr int NA_NB = (NA_NPL + 63) / 64;
constexpr int NA_TILES = NA_NB * (NA_NB + 1) / 2;
// byte offsets of the regions
constexpr size_t NA_AOS = 0;                                             // 7 doubles per body, as given
constexpr size_t NA_SOA = NA_AOS + na_align(sizeof(double) * 7 * NA_NPL);  // rh.x/y/z, vb.x/y/z
constexpr size_t NA_SYS = NA_SOA + na_align(sizeof(double) * 6 * NA_NPL);
constexpr size_t NA_SMALL = NA_SYS + na_align(sizeof(int32_t) * NA_NPL);  // offs[4], nb, woff, total
constexpr size_t NA_TILE = NA_SMALL + 256;
constexpr size_t NA_SCAN = NA_TILE + na_align(sizeof(int4) * NA_TILES);
constexpr size_t NA_SCAN_BYTES = 4096;
constexpr size_t NA_WORDS = NA_SCAN + NA_SCAN_BYTES;
constexpr size_t NA_WPRE = NA_WORDS + na_align(sizeof(uint64_t) * NA_NPL * NA_NB);
constexpr size_t NA_RCNT = NA_WPRE + na_align(sizeof(int32_t) * NA_NPL * NA_NB);
constexpr size_t NA_RST = NA_RCNT + na_align(sizeof(int32_t) * NA_NPL);
constexpr size_t NA_I1 = NA_RST + na_align(sizeof(int64_t) * NA_NPL);
constexpr size_t NA_I2 = NA_I1 + na_align(sizeof(int32_t) * NA_PAIRS);
constexpr size_t NA_LV = NA_I2 + na_align(sizeof(int32_t) * NA_PAIRS);
constexpr size_t NA_ST = NA_LV + na_align(NA_PAIRS);
constexpr size_t NA_LEVEL = NA_ST + na_align(NA_PAIRS);
constexpr size_t NA_BYTES = NA_LEVEL + na_align(NA_PAIRS);
static_assert(NA_BYTES == SWG_NOALLOC_BYTES, "the header states the region's size");

// The region.  In host code `scratch` names a host-side stand-in, not the
// device memory: host code uses only the address hipGetSymbolAddress gives
// for the current device (noalloc_ptr).
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

constexpr int NA_MAX_DEVICES = 64;
static void *g_noalloc_ptr[NA_MAX_DEVICES] = {};
static std::mutex g_noalloc_mu;

// Once per device, with that device current (swg_init): look up the device
// address of the region.
int noalloc_setup(int device)
{
    if (device < 0 || device >= NA_MAX_DEVICES) return fail(SWG_ERR_ARG, "device index beyond the no-allocation table");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return 0;
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(SWG_ERR_HIP, "hipGetSymbolAddress returned no address for the no-allocation region");
    g_noalloc_ptr[device] = p;
    return 0;
}

static void *noalloc_ptr(int device)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return (device >= 0 && device < NA_MAX_DEVICES) ? g_noalloc_ptr[device] : nullptr;
}

// the triangular 64x64 tiles (0, ib, jb), ib <= jb < nb, in the order the
// allocating path lists them
__global__ void k_na_tiles(int4 *tiles, int nb)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb * (nb + 1) / 2) return;
    int ib = 0, rest = t;
    while (rest >= nb - ib) {
        rest -= nb - ib;
        ib++;
    }
    tiles[t] = make_int4(0, ib, ib + rest, 0);
}

// the staged AoS inputs (r, v as (3,n), then renc) into SoA
__global__ void k_na_soa(const double *r, const double *v, double *soa, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    for (int d = 0; d < 3; d++) {
        soa[d * n + i] = r[3 * i + d];
        soa[(3 + d) * n + i] = v[3 * i + d];
    }
}

// offs = {0, npl, 0, 0}, nb, woff = 0
__global__ void k_na_small(int32_t *offs, int32_t *nbp, int64_t *woff, int npl, int nb)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        offs[0] = 0;
        offs[1] = npl;
        offs[2] = 0;
        offs[3] = 0;
        *nbp = nb;
        *woff = 0;
    }
}

int noalloc_encounter_plpl(swg_ctx *ctx, int32_t npl, const double *r, const double *v, const double *renc, double dt,
                           int64_t cap, int64_t *nenc, int32_t *index1, int32_t *index2, int32_t *lvdotr)
{
    *nenc = 0;
    if (npl > NA_NPL)
        return fail(SWG_ERR_ARG, "npl " + std::to_string(npl) + " exceeds the no-allocation limit of " +
                                     std::to_string(NA_NPL) + " bodies");
    if (npl == 0) return 0;
    unsigned char *base = (unsigned char *)noalloc_ptr(ctx->device);
    if (!base) return fail(SWG_ERR_HIP, "the no-allocation region of this device was not set up");
    if (scan_scratch_bytes(npl) > NA_SCAN_BYTES) return fail(SWG_ERR_ARG, "scan partials exceed their region");
    hipStream_t st = ctx->stream;
    double *aos = (double *)(base + NA_AOS), *soa = (double *)(base + NA_SOA);
    int32_t *psys = (int32_t *)(base + NA_SYS);
    int32_t *offs = (int32_t *)(base + NA_SMALL), *nbp = offs + 4;
    int64_t *woff = (int64_t *)(base + NA_SMALL + 64), *tot = woff + 1;
    int4 *tiles = (int4 *)(base + NA_TILE);
    const size_t nd = (size_t)npl;
    HIPCHK(hipMemcpyAsync(aos, r, 3 * nd * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(aos + 3 * nd, v, 3 * nd * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(aos + 6 * nd, renc, nd * sizeof(double), hipMemcpyHostToDevice, st));
    const int nb = (npl + 63) / 64, ntiles = nb * (nb + 1) / 2;
    k_na_soa<<<cdiv(npl, 256), 256, 0, st>>>(aos, aos + 3 * nd, soa, npl);
    k_fill_i32<<<cdiv(npl, 256), 256, 0, st>>>(psys, npl, 0);
    k_na_small<<<1, 64, 0, st>>>(offs, nbp, woff, npl, nb);
    k_na_tiles<<<cdiv(ntiles, 256), 256, 0, st>>>(tiles, nb);
    PlDev pl{};
    SysDev sy{};
    pl.rh = Vec3p{soa, soa + nd, soa + 2 * nd};
    pl.vb = Vec3p{soa + 3 * nd, soa + 4 * nd, soa + 5 * nd};
    pl.renc = aos + 6 * nd;
    pl.sys = psys;
    sy.pl_off = offs;
    sy.tp_off = offs + 2;
    ListDev L{};
    L.words = (uint64_t *)(base + NA_WORDS);
    L.wpre = (int32_t *)(base + NA_WPRE);
    L.rowcnt = (int32_t *)(base + NA_RCNT);
    L.rowstart = (int64_t *)(base + NA_RST);
    L.woff = woff;
    L.nb = nbp;
    L.i1 = (int32_t *)(base + NA_I1);
    L.i2 = (int32_t *)(base + NA_I2);
    L.lvdotr = (int8_t *)(base + NA_LV);
    L.status = (int8_t *)(base + NA_ST);
    L.level = (int8_t *)(base + NA_LEVEL);
    launch_pp_tiles(st, pl, sy, L, tiles, ntiles, dt);
    launch_rows_count(st, L, psys, 1, offs, nullptr, npl);
    launch_scan_i32_i64(st, L.rowcnt, L.rowstart, npl, tot, base + NA_SCAN);
    HIPCHK(hipGetLastError());
    int64_t n = 0;
    HIPCHK(copy_wait(st, &n, tot, sizeof n, hipMemcpyDeviceToHost));
    *nenc = n;
    if (n > NA_PAIRS)
        return fail(SWG_ERR_ARG, std::to_string(n) + " encounter pairs exceed the no-allocation limit of " +
                                     std::to_string(NA_PAIRS));
    if (n > cap) return fail(SWG_ERR_ARG, "the caller's list arrays hold " + std::to_string(cap) + " pairs, the list has " +
                                              std::to_string(n));
    if (!n) return 0;
    launch_rows_fill(st, L, psys, 1, offs, offs, npl);
    HIPCHK(hipGetLastError());
    // raw 0-based indices and 1-byte flags straight into the caller's arrays,
    // then made 1-based / widened on the host in place
    if (index1) HIPCHK(hipMemcpyAsync(index1, L.i1, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (index2) HIPCHK(hipMemcpyAsync(index2, L.i2, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (lvdotr) HIPCHK(hipMemcpyAsync(lvdotr, L.lvdotr, (size_t)n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (int64_t k = 0; k < n; k++) {
        if (index1) index1[k] += 1;
        if (index2) index2[k] += 1;
    }
    if (lvdotr) {
        // widen the n bytes at the start of the array to n int32 values, from
        // the end: value k goes to bytes 4k..4k+3, past every byte still unread
        const int8_t *b = (const int8_t *)lvdotr;
        for (int64_t k = n - 1; k >= 0; k--) {
            const int32_t x = b[k];
            lvdotr[k] = x;
        }
    }
    return 0;
}

}  // namespace swg

// ===========================================================================
// C ABI: leaf operators
// ===========================================================================
using namespace swg;

#define GETBUF(var, T, slot, count)                                                  \
    T *var = (T *)ctx->leaf->get(slot, sizeof(T) * (size_t)(count), &rc);            \
    if (!var) return rc;

extern "C" {

int swg_kick_getacch_int_all_tri_pl(swg_ctx *ctx, int32_t npl, int32_t nplm, const double *r, const double *Gmass,
                                    const double *radius, double *acc)
{
    if (!ctx || npl < 0 || nplm < 0 || nplm > npl) return fail(SWG_ERR_ARG, "bad argument");
    if (npl == 0) return 0;
    if (!r || !Gmass || !acc) return fail(SWG_ERR_ARG, "NULL array");
    int rc = 0;
    hipStream_t st = ctx->stream;
    GETBUF(dr, double, 0, 3 * npl);
    GETBUF(dg, double, 1, npl);
    GETBUF(da, double, 2, 3 * npl);
    double *dR = nullptr;
    if (radius) {
        GETBUF(t, double, 3, npl);
        dR = t;
        HIPCHK(hipMemcpyAsync(dR, radius, npl * sizeof(double), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(dr, r, 3 * npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dg, Gmass, npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(da, acc, 3 * npl * sizeof(double), hipMemcpyHostToDevice, st));
    k_leaf_tri<<<cdiv(npl, 128), 128, 0, st>>>(npl, nplm, dr, dg, dR, da, ctx->pair_mode);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(acc, da, 3 * npl * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int swg_kick_getacch_int_all_flat_pl(swg_ctx *ctx, int32_t npl, int64_t nplpl, const int32_t *k_plpl,
                                     const double *r, const double *Gmass, const double *radius, double *acc)
{
    if (!ctx || npl < 0 || nplpl < 0) return fail(SWG_ERR_ARG, "bad argument");
    if (npl == 0) return 0;
    if (!r || !Gmass || !acc || (nplpl && !k_plpl)) return fail(SWG_ERR_ARG, "NULL array");
    // per-body pair lists in list order (host counting sort)
    std::vector<int64_t> rs(npl + 1, 0), cs(npl + 1, 0), rowl(nplpl ? nplpl : 1), coll(nplpl ? nplpl : 1);
    for (int64_t p = 0; p < nplpl; p++) {
        int i = k_plpl[2 * p] - 1, j = k_plpl[2 * p + 1] - 1;
        if (i < 0 || i >= npl || j < 0 || j >= npl) return fail(SWG_ERR_ARG, "pair index out of range");
        rs[i + 1]++;
        cs[j + 1]++;
    }
    for (int b = 0; b < npl; b++) {
        rs[b + 1] += rs[b];
        cs[b + 1] += cs[b];
    }
    {
        std::vector<int64_t> fr(rs.begin(), rs.end() - 1), fc(cs.begin(), cs.end() - 1);
        for (int64_t p = 0; p < nplpl; p++) {
            rowl[fr[k_plpl[2 * p] - 1]++] = p;
            coll[fc[k_plpl[2 * p + 1] - 1]++] = p;
        }
    }
    int rc = 0;
    hipStream_t st = ctx->stream;
    GETBUF(dr, double, 0, 3 * npl);
    GETBUF(dg, double, 1, npl);
    GETBUF(da, double, 2, 3 * npl);
    GETBUF(dk, int32_t, 4, 2 * (nplpl ? nplpl : 1));
    GETBUF(drs, int64_t, 5, npl + 1);
    GETBUF(dcs, int64_t, 6, npl + 1);
    GETBUF(drl, int64_t, 7, rowl.size());
    GETBUF(dcl, int64_t, 8, coll.size());
    double *dR = nullptr;
    if (radius) {
        GETBUF(t, double, 3, npl);
        dR = t;
        HIPCHK(hipMemcpyAsync(dR, radius, npl * sizeof(double), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(dr, r, 3 * npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dg, Gmass, npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(da, acc, 3 * npl * sizeof(double), hipMemcpyHostToDevice, st));
    if (nplpl) HIPCHK(hipMemcpyAsync(dk, k_plpl, 2 * nplpl * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(drs, rs.data(), (npl + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dcs, cs.data(), (npl + 1) * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(drl, rowl.data(), rowl.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dcl, coll.data(), coll.size() * sizeof(int64_t), hipMemcpyHostToDevice, st));
    k_leaf_flat<<<cdiv(npl, 128), 128, 0, st>>>(npl, dk, dr, dg, dR, drs, dcs, drl, dcl, da, ctx->pair_mode);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(acc, da, 3 * npl * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int swg_kick_getacch_int_all_tp(swg_ctx *ctx, int32_t ntp, int32_t npl, const double *rtp, const double *rpl,
                                const double *GMpl, const int32_t *lmask, double *acc)
{
    if (!ctx || ntp < 0 || npl < 0) return fail(SWG_ERR_ARG, "bad argument");
    if (ntp == 0 || npl == 0) return 0;  // swiftest_kick.f90:58
    if (!rtp || !rpl || !GMpl || !lmask || !acc) return fail(SWG_ERR_ARG, "NULL array");
    int rc = 0;
    hipStream_t st = ctx->stream;
    GETBUF(dt_, double, 0, 3 * ntp);
    GETBUF(dp, double, 1, 3 * npl);
    GETBUF(dg, double, 2, npl);
    GETBUF(dm, int32_t, 3, ntp);
    GETBUF(da, double, 4, 3 * ntp);
    HIPCHK(hipMemcpyAsync(dt_, rtp, 3 * ntp * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dp, rpl, 3 * npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dg, GMpl, npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dm, lmask, ntp * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(da, acc, 3 * ntp * sizeof(double), hipMemcpyHostToDevice, st));
    k_leaf_tp<256><<<cdiv(ntp, 256), 256, 0, st>>>(ntp, npl, dt_, dp, dg, dm, da, ctx->pair_mode);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(acc, da, 3 * ntp * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

int swg_drift_all(swg_ctx *ctx, const double *mu, double *x, double *v, int32_t n, double dt, const int32_t *lmask,
                  int32_t *iflag)
{
    if (!ctx || n < 0) return fail(SWG_ERR_ARG, "bad argument");
    if (n == 0) return 0;
    if (!mu || !x || !v || !lmask || !iflag) return fail(SWG_ERR_ARG, "NULL array");
    int rc = 0;
    hipStream_t st = ctx->stream;
    GETBUF(dmu, double, 0, n);
    GETBUF(dx, double, 1, 3 * n);
    GETBUF(dv, double, 2, 3 * n);
    GETBUF(dm, int32_t, 3, n);
    GETBUF(df, int32_t, 4, n);
    HIPCHK(hipMemcpyAsync(dmu, mu, n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dx, x, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dv, v, 3 * n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dm, lmask, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(df, iflag, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    k_leaf_drift<<<cdiv(n, 128), 128, 0, st>>>(dmu, dx, dv, n, dt, dm, df);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(x, dx, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(v, dv, 3 * n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(iflag, df, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// Level-0 triangular encounter check of one body set against itself
// (which = 0) or of npl bodies against ntp particles (which = 1), using the
// resident-path kernels on a one-system layout.
static int leaf_encounter(swg_ctx *ctx, int which, int32_t npl, int32_t ntp, const double *rpl, const double *vpl,
                          const double *rtp, const double *vtp, const double *renc, double dt, int64_t *nenc)
{
    int rc = 0;
    hipStream_t st = ctx->stream;
    LeafState &ls = *ctx->leaf;
    *nenc = 0;
    ls.nenc = 0;
    if (npl == 0 || (which == 1 && ntp == 0)) return 0;
    PlDev pl{};
    TpDev tp{};
    SysDev sy{};
    GETBUF(prx, double, 10, 3 * npl);
    GETBUF(pvx, double, 11, 3 * npl);
    GETBUF(pre, double, 12, npl);
    GETBUF(psys, int32_t, 13, std::max(npl, ntp));
    GETBUF(offs, int32_t, 14, 4);
    GETBUF(woff, int64_t, 15, 1);
    GETBUF(nbp, int32_t, 16, 1);
    GETBUF(tot, int64_t, 17, 1);
    int ncol = which == 0 ? npl : ntp;
    GETBUF(tiles, int4, 18, which == 0 ? ((npl + 63) / 64) * ((npl + 63) / 64 + 1) / 2 : (ntp + 63) / 64);
    GETBUF(scan, char, 19, scan_scratch_bytes(std::max(npl, ntp)));
    pl.rh = Vec3p{prx, prx + npl, prx + 2 * npl};
    pl.vb = Vec3p{pvx, pvx + npl, pvx + 2 * npl};
    pl.renc = pre;
    pl.sys = psys;
    std::vector<double> tmp(3 * (size_t)std::max(npl, ntp));
    auto soa = [&](const double *aos, double *dst, int n) -> int {
        for (int i = 0; i < n; i++) {
            tmp[i] = aos[3 * i];
            tmp[n + i] = aos[3 * i + 1];
            tmp[2 * n + i] = aos[3 * i + 2];
        }
        HIPCHK(copy_wait(st, dst, tmp.data(), 3 * n * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    };
    if ((rc = soa(rpl, prx, npl)) || (rc = soa(vpl, pvx, npl))) return rc;
    HIPCHK(copy_wait(st, pre, renc, npl * sizeof(double), hipMemcpyHostToDevice));
    k_fill_i32<<<cdiv(std::max(npl, ntp), 256), 256, 0, st>>>(psys, std::max(npl, ntp), 0);
    int32_t ho[4] = {0, npl, 0, ntp};
    HIPCHK(hipMemcpyAsync(offs, ho, sizeof ho, hipMemcpyHostToDevice, st));
    sy.pl_off = offs;
    sy.tp_off = offs + 2;
    if (which == 1) {
        GETBUF(trx, double, 20, 3 * ntp);
        GETBUF(tvx, double, 21, 3 * ntp);
        GETBUF(tst, int32_t, 22, ntp);
        tp.rh = Vec3p{trx, trx + ntp, trx + 2 * ntp};
        tp.vb = Vec3p{tvx, tvx + ntp, tvx + 2 * ntp};
        tp.status = tst;
        tp.sys = psys;
        if ((rc = soa(rtp, trx, ntp)) || (rc = soa(vtp, tvx, ntp))) return rc;
        k_fill_i32<<<cdiv(ntp, 256), 256, 0, st>>>(tst, ntp, ACTIVE);
    }
    ListDev &L = ls.L;
    int nb = (ncol + 63) / 64;
    int64_t words = (int64_t)npl * nb;
    GETBUF(w, uint64_t, 23, words);
    GETBUF(wp, int32_t, 24, words);
    GETBUF(rcnt, int32_t, 25, npl);
    GETBUF(rst, int64_t, 26, npl);
    L.words = w;
    L.wpre = wp;
    L.rowcnt = rcnt;
    L.rowstart = rst;
    L.woff = woff;
    L.nb = nbp;
    int64_t z = 0;
    HIPCHK(hipMemcpyAsync(woff, &z, sizeof z, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(nbp, &nb, sizeof nb, hipMemcpyHostToDevice, st));
    std::vector<int4> ht;
    if (which == 0) {
        for (int ib = 0; ib < nb; ib++)
            for (int jb = ib; jb < nb; jb++) ht.push_back(make_int4(0, ib, jb, 0));
    } else {
        for (int tb = 0; tb < nb; tb++) ht.push_back(make_int4(0, tb, 0, 0));
    }
    HIPCHK(hipMemcpyAsync(tiles, ht.data(), ht.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    if (which == 0)
        launch_pp_tiles(st, pl, sy, L, tiles, (int)ht.size(), dt);
    else
        launch_pt_tiles(st, pl, tp, sy, L, tiles, (int)ht.size(), dt);
    launch_rows_count(st, L, psys, which == 0, offs, nullptr, npl);
    launch_scan_i32_i64(st, rcnt, rst, npl, tot, scan);
    int64_t n = 0;
    HIPCHK(hipMemcpyAsync(&n, tot, sizeof n, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (n > ls.cap) {
        for (void *q : {(void *)L.i1, (void *)L.i2, (void *)L.lvdotr, (void *)L.status, (void *)L.level})
            if (q) (void)hipFree(q);
        int64_t c = n + n / 2 + 64;
        HIPCHK(hipMalloc((void **)&L.i1, c * sizeof(int32_t)));
        HIPCHK(hipMalloc((void **)&L.i2, c * sizeof(int32_t)));
        HIPCHK(hipMalloc((void **)&L.lvdotr, c));
        HIPCHK(hipMalloc((void **)&L.status, c));
        HIPCHK(hipMalloc((void **)&L.level, c));
        ls.cap = c;
    }
    if (n) launch_rows_fill(st, L, psys, which == 0, offs, offs + (which == 0 ? 0 : 2), npl);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    ls.nenc = n;
    *nenc = n;
    return 0;
}

int swg_encounter_check_all_plpl(swg_ctx *ctx, int32_t npl, const double *r, const double *v, const double *renc,
                                 double dt, int64_t *nenc)
{
    if (!ctx || !nenc || npl < 0) return fail(SWG_ERR_ARG, "bad argument");
    if (npl && (!r || !v || !renc)) return fail(SWG_ERR_ARG, "NULL array");
    return leaf_encounter(ctx, 0, npl, 0, r, v, nullptr, nullptr, renc, dt, nenc);
}

int swg_encounter_check_all_pltp(swg_ctx *ctx, int32_t npl, int32_t ntp, const double *rpl, const double *vpl,
                                 const double *rtp, const double *vtp, const double *renc, double dt, int64_t *nenc)
{
    if (!ctx || !nenc || npl < 0 || ntp < 0) return fail(SWG_ERR_ARG, "bad argument");
    if (npl && ntp && (!rpl || !vpl || !rtp || !vtp || !renc)) return fail(SWG_ERR_ARG, "NULL array");
    return leaf_encounter(ctx, 1, npl, ntp, rpl, vpl, rtp, vtp, renc, dt, nenc);
}

int swg_encounter_check_all_plpl_noalloc(swg_ctx *ctx, int32_t npl, const double *r, const double *v,
                                         const double *renc, double dt, int64_t cap, int64_t *nenc,
                                         int32_t *index1, int32_t *index2, int32_t *lvdotr)
{
    if (!ctx || !nenc || npl < 0 || cap < 0) return fail(SWG_ERR_ARG, "bad argument");
    if (npl && (!r || !v || !renc)) return fail(SWG_ERR_ARG, "NULL array");
    HIPCHK(hipSetDevice(ctx->device));
    return noalloc_encounter_plpl(ctx, npl, r, v, renc, dt, cap, nenc, index1, index2, lvdotr);
}

int swg_encounter_list_copy(swg_ctx *ctx, int32_t *index1, int32_t *index2, int32_t *lvdotr)
{
    if (!ctx) return fail(SWG_ERR_ARG, "NULL context");
    LeafState &ls = *ctx->leaf;
    int64_t n = ls.nenc;
    if (!n) return 0;
    int rc = 0;
    hipStream_t st = ctx->stream;
    GETBUF(a, int32_t, 30, n);
    GETBUF(b, int32_t, 31, n);
    GETBUF(c, int32_t, 32, n);
    k_copy_i32_plus1<<<cdiv(n, 256), 256, 0, st>>>(ls.L.i1, a, ls.L.lvdotr, c, n);
    k_copy_i32_plus1<<<cdiv(n, 256), 256, 0, st>>>(ls.L.i2, b, nullptr, nullptr, n);
    HIPCHK(hipGetLastError());
    if (index1) HIPCHK(hipMemcpyAsync(index1, a, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (index2) HIPCHK(hipMemcpyAsync(index2, b, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    if (lvdotr) HIPCHK(hipMemcpyAsync(lvdotr, c, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

}  // extern "C"
