// swg_level.h -- SyMBA recursion working sets and the per-body / per-pair
// operations of one recursion substep, shared by the whole-device level
// kernels (k_level.hip) and the persistent per-system kernel (k_subtree.hip).
//
// While the recursion is at level d (symba_step_recur_system with ireci = d,
// symba_step.f90:163-262) only bodies with levelg >= d and pairs whose two
// bodies both have levelg >= d can change.  A LevelView lists them: bodies in
// ascending index order, and for each body its pairs in list order (so a
// per-body sum over its entries follows the reference's serial order).
#pragma once
#include "swg_math.h"
#include "swg_state.h"

namespace swg {

constexpr int MAXLEV = 36;  // deeper than the recursion depth guard (3^-33 < 2^-52)

// The powers of RSHELL the recursion uses at level k: powi(RSHELL, 2k), the
// shell radius factor of the kicks, and rshell_pow(k), the encounter radius
// factor of the checks.  They are evaluated at compile time by the same
// operations in the same order as at run time (round to nearest), so a
// table read gives the run-time bits in one load instead of a loop of
// dependent multiplications.
struct ShellPowers {
    double sq[MAXLEV + 2], r[MAXLEV + 2];
    constexpr ShellPowers() : sq(), r()
    {
        for (int k = 0; k < MAXLEV + 2; k++) {
            sq[k] = powi(RSHELL, 2 * k);
            r[k] = rshell_pow(k);
        }
    }
};
static __constant__ const ShellPowers kShellPow{};
__device__ inline double shell_sq(int k) { return kShellPow.sq[k]; }  // powi(RSHELL, 2k), 0 <= k <= MAXLEV+1
__device__ inline double shell_r(int k) { return kShellPow.r[k]; }    // rshell_pow(k), 0 <= k <= MAXLEV+1

// One pl entry: x = (pair << 1) | role (role 1: the body is index1), y = the
// partner body.  One tp entry: x = pair, y = the massive body.
struct LevelView {
    const int32_t *B, *O, *Bt, *Ot;  // bodies; entry offsets (into E / Et), n+1 each
    const int2 *E, *Et;
    int nBpl, nBtp;
};

// Word layout of one stored level, from an 8-byte aligned base:
//   B[nB] O[nB+1] E[nE] Bt[nBt] Ot[nBt+1] Et[nEt]
// with each entry array (int2) starting at an even word.
struct LevelLayout {
    int64_t oB, oO, oE, oBt, oOt, oEt, words;
};

__host__ __device__ inline LevelLayout level_layout(int64_t nB, int64_t nE, int64_t nBt, int64_t nEt)
{
    LevelLayout L;
    int64_t w = 0;
    L.oB = w;
    w += nB;
    L.oO = w;
    w += nB + 1;
    w += w & 1;
    L.oE = w;
    w += 2 * nE;
    L.oBt = w;
    w += nBt;
    L.oOt = w;
    w += nBt + 1;
    w += w & 1;
    L.oEt = w;
    w += 2 * nEt;
    L.words = w;
    return L;
}

struct RecState {
    PlDev pl;
    TpDev tp;
    ListDev pp, pt;
    SysDev sy;
};

// ---- per-pair arithmetic on values (shared by the HBM and LDS paths) --------

// The level-(d+1) test of one pl-pl pair (symba_encounter_check.f90:121-154):
// relative motion of j with respect to i, the pair's shell radius at d+1 and
// the physical-overlap exclusion.  Writes lvdotr; returns true when the pair
// enters level d+1.
__device__ inline bool check_pp_v(double xi, double yi, double zi, double vxi, double vyi, double vzi, double hi,
                                  double Ri, double xj, double yj, double zj, double vxj, double vyj, double vzj,
                                  double hj, double Rj, int irec, double dtl, bool &lv)
{
    double xr = xj - xi, yr = yj - yi, zr = zj - zi;
    double vxr = vxj - vxi, vyr = vyj - vyi, vzr = vzj - vzi;
    const double rsh = shell_r(irec);
    double rci = (hi * RHSCALE) * rsh, rcj = (hj * RHSCALE) * rsh;
    bool le = encounter_one(xr, yr, zr, vxr, vyr, vzr, rci + rcj, dtl, lv);
    if (le) {
        double rji2 = dot3(xr, yr, zr, xr, yr, zr);
        double rs = Ri + Rj;
        le = rji2 > rs * rs;
    }
    return le;
}

// The level-(d+1) test of one pl-tp pair (symba_encounter_check.f90:195-228)
__device__ inline bool check_pt_v(double xi, double yi, double zi, double vxi, double vyi, double vzi, double hi,
                                  double Ri, double xt, double yt, double zt, double vxt, double vyt, double vzt,
                                  int irec, double dtl, bool &lv)
{
    double xr = xt - xi, yr = yt - yi, zr = zt - zi;
    double vxr = vxt - vxi, vyr = vyt - vyi, vzr = vzt - vzi;
    bool le = encounter_one(xr, yr, zr, vxr, vyr, vzr, (hi * RHSCALE) * shell_r(irec), dtl, lv);
    if (le) {
        double rji2 = dot3(xr, yr, zr, xr, yr, zr);
        le = rji2 > Ri * Ri;
    }
    return le;
}

// One pl-pl shell term of body b (symba_kick.f90:150-200): the pair's
// contribution f*(dx,dy,dz) to b's sum at the outer shell (sign +1, irecl =
// d+1) and, when `minus`, at the inner shell (sign -1, irecl = d).  tp / tm
// tell whether the pair lies inside the respective shell.  bi: b is index1;
// (xp, yp, zp, hp, gp) is the partner.
struct ShellTerm {
    double dx, dy, dz, fp, fm;
    bool tp, tm;
};

__device__ inline void pl_shell_term_v(bool bi, double xb, double yb, double zb, double hb, double xp, double yp,
                                       double zp, double hp, double gp, double spp, double spm, bool minus,
                                       ShellTerm &t)
{
    double sh = bi ? hb + hp : hp + hb;
    double s2 = (sh * sh) * (RHSCALE * RHSCALE);
    if (bi) {
        t.dx = xp - xb;
        t.dy = yp - yb;
        t.dz = zp - zb;
    } else {
        t.dx = xb - xp;
        t.dy = yb - yp;
        t.dz = zb - zp;
    }
    double r2 = dot3(t.dx, t.dy, t.dz, t.dx, t.dy, t.dz);
    bool okp, okm;
    double facp, facm;
    shell_facs(s2 * spp, s2 * spm, r2, minus, okp, facp, okm, facm);
    if (okp) {
        t.tp = true;
        t.fp = bi ? facp * gp : -(facp * gp);
    }
    if (okm) {
        t.tm = true;
        t.fm = bi ? facm * gp : -(facm * gp);
    }
}

// One pl-tp shell term of a test particle (symba_kick.f90:255-300)
__device__ inline void tp_shell_term_v(double xt, double yt, double zt, double xi, double yi, double zi, double hi,
                                       double gi, double spp, double spm, bool minus, ShellTerm &t)
{
    double s2 = (hi * hi) * (RHSCALE * RHSCALE);
    t.dx = xt - xi;
    t.dy = yt - yi;
    t.dz = zt - zi;
    double r2 = dot3(t.dx, t.dy, t.dz, t.dx, t.dy, t.dz);
    bool okp, okm;
    double facp, facm;
    shell_facs(s2 * spp, s2 * spm, r2, minus, okp, facp, okm, facm);
    if (okp) {
        t.tp = true;
        t.fp = -(facp * gi);
    }
    if (okm) {
        t.tm = true;
        t.fm = -(facm * gi);
    }
}

// ---- the same on the HBM-resident state --------------------------------------

// symba_encounter_check_list_plpl for one pair; returns true when the pair
// enters level d+1
__device__ inline bool rec_check_pp(const RecState &a, int k, int d, double dtl)
{
    if (a.pp.status[k] != ACTIVE || a.pp.level[k] != d) return false;
    const int irec = d + 1;
    const int i = a.pp.i1[k], j = a.pp.i2[k];
    const PlDev &pl = a.pl;
    bool lv;
    bool le = check_pp_v(pl.rh.x[i], pl.rh.y[i], pl.rh.z[i], pl.vb.x[i], pl.vb.y[i], pl.vb.z[i], pl.rhill[i],
                         pl.R[i], pl.rh.x[j], pl.rh.y[j], pl.rh.z[j], pl.vb.x[j], pl.vb.y[j], pl.vb.z[j],
                         pl.rhill[j], pl.R[j], irec, dtl, lv);
    a.pp.lvdotr[k] = lv;
    if (le) {
        a.pp.level[k] = (int8_t)irec;
        pl.levelg[i] = irec;
        pl.levelg[j] = irec;
        atomicMax(&pl.levelm[i], irec);
        atomicMax(&pl.levelm[j], irec);
    }
    return le;
}

// symba_encounter_check_list_pltp for one pair
__device__ inline bool rec_check_pt(const RecState &a, int k, int d, double dtl)
{
    if (a.pt.status[k] != ACTIVE || a.pt.level[k] != d) return false;
    const int irec = d + 1;
    const int i = a.pt.i1[k], j = a.pt.i2[k];
    const PlDev &pl = a.pl;
    bool lv;
    bool le = check_pt_v(pl.rh.x[i], pl.rh.y[i], pl.rh.z[i], pl.vb.x[i], pl.vb.y[i], pl.vb.z[i], pl.rhill[i],
                         pl.R[i], a.tp.rh.x[j], a.tp.rh.y[j], a.tp.rh.z[j], a.tp.vb.x[j], a.tp.vb.y[j],
                         a.tp.vb.z[j], irec, dtl, lv);
    a.pt.lvdotr[k] = lv;
    if (le) {
        a.pt.level[k] = (int8_t)irec;
        a.pl.levelg[i] = irec;
        atomicMax(&a.pl.levelm[i], irec);
        a.tp.levelg[j] = irec;
        atomicMax(&a.tp.levelm[j], irec);
    }
    return le;
}

__device__ inline void pl_shell_term(const RecState &a, int2 en, double xb, double yb, double zb, double hb,
                                     double spp, double spm, bool minus, ShellTerm &t)
{
    t.tp = t.tm = false;
    const int k = (int)((uint32_t)en.x >> 1);
    if (a.pp.status[k] != ACTIVE) return;
    const int p = en.y;
    const PlDev &pl = a.pl;
    pl_shell_term_v(en.x & 1, xb, yb, zb, hb, pl.rh.x[p], pl.rh.y[p], pl.rh.z[p], pl.rhill[p], pl.Gm[p], spp, spm,
                    minus, t);
}

__device__ inline void tp_shell_term(const RecState &a, int2 en, double xt, double yt, double zt, double spp,
                                     double spm, bool minus, ShellTerm &t)
{
    t.tp = t.tm = false;
    if (a.pt.status[en.x] != ACTIVE) return;
    const int i = en.y;
    tp_shell_term_v(xt, yt, zt, a.pl.rh.x[i], a.pl.rh.y[i], a.pl.rh.z[i], a.pl.rhill[i], a.pl.Gm[i], spp, spm,
                    minus, t);
}

// vb += dth * (outer-shell sum), then vb -= dth * (inner-shell sum), each only
// when its sum had a term (symba_kick.f90:210-230)
__device__ inline void apply_shell_kicks(Vec3p v, int b, double dth, bool anyp, const double3 &p, bool anym,
                                         const double3 &m)
{
    if (!(anyp || anym)) return;
    double vx = v.x[b], vy = v.y[b], vz = v.z[b];
    if (anyp) {
        vx = fma(dth, p.x, vx);
        vy = fma(dth, p.y, vy);
        vz = fma(dth, p.z, vz);
    }
    if (anym) {
        vx = fma(-dth, m.x, vx);
        vy = fma(-dth, m.y, vy);
        vz = fma(-dth, m.z, vz);
    }
    v.x[b] = vx;
    v.y[b] = vy;
    v.z[b] = vz;
}

// The shell kicks of level d on massive body b (symba_kick_list_plpl,
// symba_kick.f90:126-232): at d >= 1 the call with sgn = +1 at irecl = d+1
// and the one with sgn = -1 at irecl = d; at d = 0 only the first.  Both sums
// are gathered in one pass over b's pairs in list order (the reference's
// serial order) and applied in the reference's order.  They read positions
// only, so fusing the two calls changes no result.
__device__ inline void rec_kick_pl(const RecState &a, int b, const int2 *E, int e0, int e1, int d, double dth)
{
    const double spp = shell_sq(d + 1), spm = shell_sq(d);
    const bool minus = d > 0;
    const double xb = a.pl.rh.x[b], yb = a.pl.rh.y[b], zb = a.pl.rh.z[b], hb = a.pl.rhill[b];
    double3 p = make_double3(0.0, 0.0, 0.0), m = p;
    bool anyp = false, anym = false;
    for (int u = e0; u < e1; u++) {
        ShellTerm t;
        pl_shell_term(a, E[u], xb, yb, zb, hb, spp, spm, minus, t);
        if (t.tp) {
            anyp = true;
            p.x = fma(t.fp, t.dx, p.x);
            p.y = fma(t.fp, t.dy, p.y);
            p.z = fma(t.fp, t.dz, p.z);
        }
        if (t.tm) {
            anym = true;
            m.x = fma(t.fm, t.dx, m.x);
            m.y = fma(t.fm, t.dy, m.y);
            m.z = fma(t.fm, t.dz, m.z);
        }
    }
    apply_shell_kicks(a.pl.vb, b, dth, anyp, p, anym, m);
}

// The shell kicks of level d on test particle t (symba_kick_list_pltp,
// symba_kick.f90:235-340)
__device__ inline void rec_kick_tp(const RecState &a, int t, const int2 *E, int e0, int e1, int d, double dth)
{
    const double spp = shell_sq(d + 1), spm = shell_sq(d);
    const bool minus = d > 0;
    const double xt = a.tp.rh.x[t], yt = a.tp.rh.y[t], zt = a.tp.rh.z[t];
    double3 p = make_double3(0.0, 0.0, 0.0), m = p;
    bool anyp = false, anym = false;
    for (int u = e0; u < e1; u++) {
        ShellTerm q;
        tp_shell_term(a, E[u], xt, yt, zt, spp, spm, minus, q);
        if (q.tp) {
            anyp = true;
            p.x = fma(q.fp, q.dx, p.x);
            p.y = fma(q.fp, q.dy, p.y);
            p.z = fma(q.fp, q.dz, p.z);
        }
        if (q.tm) {
            anym = true;
            m.x = fma(q.fm, q.dx, m.x);
            m.y = fma(q.fm, q.dy, m.y);
            m.z = fma(q.fm, q.dz, m.z);
        }
    }
    apply_shell_kicks(a.tp.vb, t, dth, anyp, p, anym, m);
}

// level-d drift of one body (symba_drift.f90:14-61 over swiftest_drift_one)
__device__ inline void rec_drift(const RecState &a, int g, bool ispl, int d, double dtl, int s)
{
    Vec3p r = ispl ? a.pl.rh : a.tp.rh, v = ispl ? a.pl.vb : a.tp.vb;
    int32_t *status = ispl ? a.pl.status : a.tp.status;
    const int32_t *levelg = ispl ? a.pl.levelg : a.tp.levelg;
    bool m = ispl ? status[g] != INACTIVE : status[g] == ACTIVE;
    if (!m || levelg[g] != d) return;
    double rx = r.x[g], ry = r.y[g], rz = r.z[g], vx = v.x[g], vy = v.y[g], vz = v.z[g];
    int fl = drift_one(a.sy.GMcb[s], rx, ry, rz, vx, vy, vz, dtl);
    r.x[g] = rx;
    r.y[g] = ry;
    r.z[g] = rz;
    v.x[g] = vx;
    v.y[g] = vy;
    v.z[g] = vz;
    if (fl) {
        status[g] = DISCARDED_DRIFTERR;
        if (ispl) atomicOr(&a.sy.err[s], 2);
    }
}

// collision detection of the pl-pl pairs whose index1 is body i
// (collision_check_plpl, collision_check.f90:53-140)
__device__ inline void rec_collide_pl(const RecState &a, int i, double dtl, int s)
{
    const PlDev &pl = a.pl;
    const int64_t r0 = a.pp.rowstart[i], r1 = r0 + a.pp.rowcnt[i];
    for (int64_t k = r0; k < r1; k++) {
        if (a.pp.status[k] != ACTIVE) continue;
        int j = a.pp.i2[k];
        bool c = collision_one(pl.rh.x[i] - pl.rh.x[j], pl.rh.y[i] - pl.rh.y[j], pl.rh.z[i] - pl.rh.z[j],
                               pl.vb.x[i] - pl.vb.x[j], pl.vb.y[i] - pl.vb.y[j], pl.vb.z[i] - pl.vb.z[j],
                               pl.Gm[i] + pl.Gm[j], pl.R[i] + pl.R[j], dtl, a.pp.lvdotr[k] != 0);
        if (c) {
            a.pp.status[k] = (int8_t)COLLIDED;
            atomicOr(&a.sy.err[s], 1);
        }
    }
}

// collision detection of the pl-tp pairs of particle j (collision_check_pltp)
__device__ inline void rec_collide_tp(const RecState &a, int j, double dtl, int s)
{
    const int64_t c0 = a.pt.colstart[j], c1 = c0 + a.pt.colcnt[j];
    for (int64_t c = c0; c < c1; c++) {
        int k = a.pt.col[c];
        if (a.pt.status[k] != ACTIVE) continue;
        int i = a.pt.i1[k];
        bool col = collision_one(a.pl.rh.x[i] - a.tp.rh.x[j], a.pl.rh.y[i] - a.tp.rh.y[j],
                                 a.pl.rh.z[i] - a.tp.rh.z[j], a.pl.vb.x[i] - a.tp.vb.x[j],
                                 a.pl.vb.y[i] - a.tp.vb.y[j], a.pl.vb.z[i] - a.tp.vb.z[j], a.pl.Gm[i], a.pl.R[i],
                                 dtl, a.pt.lvdotr[k] != 0);
        if (col) {
            a.pt.status[k] = (int8_t)COLLIDED;
            atomicOr(&a.sy.err[s], 1);
        }
    }
}

// ---- building the next level's sets ----------------------------------------

// entries of a body at level 0: its whole adjacency, index2 side first
__device__ inline int l0_pl_count(const RecState &a, int has_pp, int g)
{
    return has_pp ? a.pp.colcnt[g] + a.pp.rowcnt[g] : 0;
}
__device__ inline int2 l0_pl_entry(const RecState &a, int g, int q)
{
    int cc = a.pp.colcnt[g];
    if (q < cc) {
        int k = a.pp.col[a.pp.colstart[g] + q];
        return make_int2((int)((uint32_t)k << 1), a.pp.i1[k]);
    }
    int k = (int)(a.pp.rowstart[g] + (q - cc));
    return make_int2((int)(((uint32_t)k << 1) | 1u), a.pp.i2[k]);
}
__device__ inline int l0_tp_count(const RecState &a, int has_pt, int g) { return has_pt ? a.pt.colcnt[g] : 0; }
__device__ inline int2 l0_tp_entry(const RecState &a, int g, int q)
{
    int k = a.pt.col[a.pt.colstart[g] + q];
    return make_int2(k, a.pt.i1[k]);
}

}  // namespace swg
