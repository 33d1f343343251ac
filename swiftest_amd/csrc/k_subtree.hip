// k_subtree.hip -- the SyMBA recursion below level 0, one persistent
// workgroup per system.
//
// symba_step_recur_system (symba_step.f90:163-262) is entered at level 1 from
// the single level-0 substep.  One workgroup executes the whole subtree of a
// system with an explicit level stack: list check (symba_encounter_check.f90:
// 90-235), shell kicks (symba_kick.f90:126-340), level-masked drift
// (symba_drift.f90:14-61), descent, kicks, collision detection
// (collision_check.f90:53-230) and the level reset (symba_step.f90:118-160),
// separated by workgroup barriers, with no host round trip.
//
// Working sets.  While the recursion is at level d only bodies with
// levelg >= d and pairs whose two bodies both have levelg >= d can change
// (B_d, P_d).  Both sets are fixed for the whole call at level d and shrink
// with depth.  On descent the level-(d+1) sets are built by filtering the
// level-d ones with order-preserving workgroup scans, so each body's pairs
// stay in list order and every per-body shell-kick sum keeps the reference's
// serial order.  The sets live in a per-system arena in HBM.
//
// Arena layout of one level (int32 words):
//   Bpl[nBpl] Opl[nBpl+1] Epl[nEpl] Btp[nBtp] Otp[nBtp+1] Etp[nEtp]
// Epl entries are (pair << 1) | role, role 1 when the body is index1.
// Building a level puts 2*(nBpl+nBtp) words of temporaries after the source
// level and the new level after them.
#include <cstdio>
#include <cstdlib>

#include "swg_kernels.h"
#include "swg_math.h"

namespace swg {

constexpr int ST_T = 1024;        // threads of the persistent workgroup
constexpr int ST_W = ST_T / 64;  // wavefronts
constexpr int MAXLEV = 36;       // deeper than the recursion depth guard

struct LevelRec {
    int64_t base;  // arena word offset of Bpl; < 0 for level 0 (full adjacency)
    int nBpl, nEpl, nBtp, nEtp;
};

__device__ inline int64_t level_words(const LevelRec &L)
{
    return (int64_t)L.nBpl * 2 + 1 + L.nEpl + (int64_t)L.nBtp * 2 + 1 + L.nEtp;
}

__device__ inline int st_wave_scan(int v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

// workgroup exclusive scan of one value per thread; *tot = sum (all threads)
__device__ inline int st_block_scan(int v, int *sh, int *tot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = st_wave_scan(v);
    if (lane == 63) sh[w] = incl;
    __syncthreads();
    if (w == 0) {
        int t = lane < ST_W ? sh[lane] : 0;
        int ti = st_wave_scan(t);
        if (lane < ST_W) sh[lane] = ti - t;
        if (lane == ST_W - 1) sh[ST_W] = ti;
    }
    __syncthreads();
    int r = sh[w] + incl - v;
    *tot = sh[ST_W];
    __syncthreads();
    return r;
}

struct SubArgs {
    PlDev pl;
    TpDev tp;
    ListDev pp, pt;
    SysDev sy;
    double dt;  // param%dt
    int lclose, short_circuit, has_pp, has_pt;
    int32_t *arena;
    const int64_t *arena_off;  // nsys+1 word offsets
    int32_t *stats;            // [0] substeps, [1] max depth, [2] overflow
    // optional per-level profile (NPROF words per level): substeps, sum nBpl,
    // sum nEpl, then wall-clock ticks in check, kicks, drift, build, levels
    unsigned long long *prof;
};
constexpr int NPROF = 8;

// ---- entries of one body ----------------------------------------------------

__device__ inline int pl_partner(const SubArgs &a, uint32_t e)
{
    int k = (int)(e >> 1);
    return (e & 1u) ? a.pp.i2[k] : a.pp.i1[k];
}

// level 0: the whole adjacency (index2 side, then index1 side = list order)
__device__ inline int l0_pl_count(const SubArgs &a, int g)
{
    return a.has_pp ? a.pp.colcnt[g] + a.pp.rowcnt[g] : 0;
}
__device__ inline uint32_t l0_pl_entry(const SubArgs &a, int g, int q)
{
    int cc = a.pp.colcnt[g];
    if (q < cc) return ((uint32_t)a.pp.col[a.pp.colstart[g] + q]) << 1;
    return (((uint32_t)(a.pp.rowstart[g] + (q - cc))) << 1) | 1u;
}

struct LevelPtr {
    int32_t *B, *O, *E, *Bt, *Ot, *Et;
};

__device__ inline LevelPtr level_ptr(const SubArgs &a, const LevelRec &L)
{
    LevelPtr p;
    p.B = a.arena + L.base;
    p.O = p.B + L.nBpl;
    p.E = p.O + L.nBpl + 1;
    p.Bt = p.E + L.nEpl;
    p.Ot = p.Bt + L.nBtp;
    p.Et = p.Ot + L.nBtp + 1;
    return p;
}

// Build level dnew from src (src.base < 0: level 0 of system s).  Returns
// false (uniformly) when the arena is too small.
__device__ bool build_level(const SubArgs &a, int s, const LevelRec &src, int dnew, LevelRec &dst, int *sh)
{
    const int tid = threadIdx.x;
    const bool l0 = src.base < 0;
    const int p0 = a.sy.pl_off[s], t0 = a.sy.tp_off[s];
    const int nBpl = l0 ? a.sy.pl_off[s + 1] - p0 : src.nBpl;
    const int nBtp = l0 ? (a.has_pt ? a.sy.tp_off[s + 1] - t0 : 0) : src.nBtp;
    LevelPtr sp{};
    if (!l0) sp = level_ptr(a, src);
    const int64_t tbase = l0 ? a.arena_off[s] : src.base + level_words(src);
    const int64_t aend = a.arena_off[s + 1];
    if (tbase + 2 * (int64_t)(nBpl + nBtp) > aend) return false;
    int32_t *posP = a.arena + tbase, *offP = posP + nBpl, *posT = offP + nBpl, *offT = posT + nBtp;
    // pass 1: kept-entry counts (-1: body dropped)
    for (int q = tid; q < nBpl; q += ST_T) {
        int g = l0 ? p0 + q : sp.B[q];
        int c = -1;
        if (a.pl.levelg[g] >= dnew) {
            c = 0;
            int n = l0 ? l0_pl_count(a, g) : sp.O[q + 1] - sp.O[q];
            for (int u = 0; u < n; u++) {
                uint32_t e = l0 ? l0_pl_entry(a, g, u) : (uint32_t)sp.E[sp.O[q] + u];
                if (a.pl.levelg[pl_partner(a, e)] >= dnew) c++;
            }
        }
        offP[q] = c;
    }
    for (int q = tid; q < nBtp; q += ST_T) {
        int g = l0 ? t0 + q : sp.Bt[q];
        int c = -1;
        if (a.tp.levelg[g] >= dnew) {
            c = 0;
            int n = l0 ? a.pt.colcnt[g] : sp.Ot[q + 1] - sp.Ot[q];
            for (int u = 0; u < n; u++) {
                int k = l0 ? a.pt.col[a.pt.colstart[g] + u] : sp.Et[sp.Ot[q] + u];
                if (a.pl.levelg[a.pt.i1[k]] >= dnew) c++;
            }
        }
        offT[q] = c;
    }
    __syncthreads();
    // scans: new body positions and entry offsets
    int nb = 0, ne = 0;
    for (int c0 = 0; c0 < nBpl; c0 += ST_T) {
        int q = c0 + tid;
        int c = q < nBpl ? offP[q] : -1;
        int tb, te;
        int pb = st_block_scan(c >= 0 ? 1 : 0, sh, &tb);
        int pe = st_block_scan(c > 0 ? c : 0, sh, &te);
        if (q < nBpl) {
            posP[q] = c >= 0 ? nb + pb : -1;
            offP[q] = ne + pe;
        }
        nb += tb;
        ne += te;
    }
    int nbt = 0, net = 0;
    for (int c0 = 0; c0 < nBtp; c0 += ST_T) {
        int q = c0 + tid;
        int c = q < nBtp ? offT[q] : -1;
        int tb, te;
        int pb = st_block_scan(c >= 0 ? 1 : 0, sh, &tb);
        int pe = st_block_scan(c > 0 ? c : 0, sh, &te);
        if (q < nBtp) {
            posT[q] = c >= 0 ? nbt + pb : -1;
            offT[q] = net + pe;
        }
        nbt += tb;
        net += te;
    }
    dst.base = tbase + 2 * (int64_t)(nBpl + nBtp);
    dst.nBpl = nb;
    dst.nEpl = ne;
    dst.nBtp = nbt;
    dst.nEtp = net;
    if (dst.base + level_words(dst) > aend) return false;
    __syncthreads();
    // pass 2: fill
    LevelPtr dp = level_ptr(a, dst);
    for (int q = tid; q < nBpl; q += ST_T) {
        int pos = posP[q];
        if (pos < 0) continue;
        int g = l0 ? p0 + q : sp.B[q];
        dp.B[pos] = g;
        int o = offP[q];
        dp.O[pos] = o;
        int n = l0 ? l0_pl_count(a, g) : sp.O[q + 1] - sp.O[q];
        for (int u = 0; u < n; u++) {
            uint32_t e = l0 ? l0_pl_entry(a, g, u) : (uint32_t)sp.E[sp.O[q] + u];
            if (a.pl.levelg[pl_partner(a, e)] >= dnew) dp.E[o++] = (int32_t)e;
        }
    }
    for (int q = tid; q < nBtp; q += ST_T) {
        int pos = posT[q];
        if (pos < 0) continue;
        int g = l0 ? t0 + q : sp.Bt[q];
        dp.Bt[pos] = g;
        int o = offT[q];
        dp.Ot[pos] = o;
        int n = l0 ? a.pt.colcnt[g] : sp.Ot[q + 1] - sp.Ot[q];
        for (int u = 0; u < n; u++) {
            int k = l0 ? a.pt.col[a.pt.colstart[g] + u] : sp.Et[sp.Ot[q] + u];
            if (a.pl.levelg[a.pt.i1[k]] >= dnew) dp.Et[o++] = k;
        }
    }
    if (tid == 0) {
        dp.O[nb] = ne;
        dp.Ot[nbt] = net;
    }
    __syncthreads();
    return true;
}

// ---- phases of one substep at level d ------------------------------------------

// list checks at level d+1 (pairs at level d); returns lenc (uniform)
__device__ bool phase_check(const SubArgs &a, int s, const LevelRec &L, int d, double dtl, int *flags)
{
    const int tid = threadIdx.x;
    const int irec = d + 1;
    LevelPtr lp = level_ptr(a, L);
    if (tid == 0) {
        flags[0] = 0;
        flags[1] = 0;
    }
    __syncthreads();
    const double rsh = rshell_pow(irec);
    for (int q = tid; q < L.nEpl; q += ST_T) {
        uint32_t e = (uint32_t)lp.E[q];
        if (!(e & 1u)) continue;  // each pair once, from its index1 body
        int k = (int)(e >> 1);
        if (a.pp.status[k] != ACTIVE || a.pp.level[k] != d) continue;
        const int i = a.pp.i1[k], j = a.pp.i2[k];
        double xr = a.pl.rh.x[j] - a.pl.rh.x[i], yr = a.pl.rh.y[j] - a.pl.rh.y[i], zr = a.pl.rh.z[j] - a.pl.rh.z[i];
        double vxr = a.pl.vb.x[j] - a.pl.vb.x[i], vyr = a.pl.vb.y[j] - a.pl.vb.y[i],
               vzr = a.pl.vb.z[j] - a.pl.vb.z[i];
        double rci = (a.pl.rhill[i] * RHSCALE) * rsh, rcj = (a.pl.rhill[j] * RHSCALE) * rsh;
        bool lv;
        bool le = encounter_one(xr, yr, zr, vxr, vyr, vzr, rci + rcj, dtl, lv);
        a.pp.lvdotr[k] = lv;
        if (le) {
            double rji2 = dot3(xr, yr, zr, xr, yr, zr);
            double rs = a.pl.R[i] + a.pl.R[j];
            le = rji2 > rs * rs;
        }
        if (le) {
            a.pp.level[k] = (int8_t)irec;
            a.pl.levelg[i] = irec;
            a.pl.levelg[j] = irec;
            atomicMax(&a.pl.levelm[i], irec);
            atomicMax(&a.pl.levelm[j], irec);
            flags[0] = 1;
        }
    }
    __syncthreads();
    if (!(a.short_circuit && flags[0])) {
        for (int q = tid; q < L.nEtp; q += ST_T) {
            int k = lp.Et[q];
            if (a.pt.status[k] != ACTIVE || a.pt.level[k] != d) continue;
            const int i = a.pt.i1[k], j = a.pt.i2[k];
            double xr = a.tp.rh.x[j] - a.pl.rh.x[i], yr = a.tp.rh.y[j] - a.pl.rh.y[i],
                   zr = a.tp.rh.z[j] - a.pl.rh.z[i];
            double vxr = a.tp.vb.x[j] - a.pl.vb.x[i], vyr = a.tp.vb.y[j] - a.pl.vb.y[i],
                   vzr = a.tp.vb.z[j] - a.pl.vb.z[i];
            bool lv;
            bool le = encounter_one(xr, yr, zr, vxr, vyr, vzr, (a.pl.rhill[i] * RHSCALE) * rsh, dtl, lv);
            a.pt.lvdotr[k] = lv;
            if (le) {
                double rji2 = dot3(xr, yr, zr, xr, yr, zr);
                le = rji2 > a.pl.R[i] * a.pl.R[i];
            }
            if (le) {
                a.pt.level[k] = (int8_t)irec;
                a.pl.levelg[i] = irec;
                atomicMax(&a.pl.levelm[i], irec);
                a.tp.levelg[j] = irec;
                atomicMax(&a.tp.levelm[j], irec);
                flags[1] = 1;
            }
        }
    }
    __syncthreads();
    bool lenc = flags[0] || flags[1];
    __syncthreads();
    return lenc;
}

// one shell kick of pl body slot q (symba_kick_list_plpl) with sign sgn
__device__ inline void kick_pl_body(const SubArgs &a, const LevelPtr &lp, int q, int irec, int sgn, double dth)
{
    const int b = lp.B[q];
    const int irecl = sgn < 0 ? irec - 1 : irec;
    const double shellpow = powi(RSHELL, 2 * irecl);
    const double xb = a.pl.rh.x[b], yb = a.pl.rh.y[b], zb = a.pl.rh.z[b], hb = a.pl.rhill[b];
    double ax = 0.0, ay = 0.0, az = 0.0;
    bool any = false;
    for (int u = lp.O[q]; u < lp.O[q + 1]; u++) {
        uint32_t e = (uint32_t)lp.E[u];
        int k = (int)(e >> 1);
        if (a.pp.status[k] != ACTIVE) continue;
        if (e & 1u) {  // b = index1
            int j = a.pp.i2[k];
            double sh = hb + a.pl.rhill[j];
            double ri = ((sh * sh) * (RHSCALE * RHSCALE)) * shellpow;
            double dx = a.pl.rh.x[j] - xb, dy = a.pl.rh.y[j] - yb, dz = a.pl.rh.z[j] - zb;
            double r2 = dot3(dx, dy, dz, dx, dy, dz);
            double fac;
            if (!shell_fac(ri, r2, fac)) continue;
            any = true;
            double facj = fac * a.pl.Gm[j];
            ax = fma(facj, dx, ax);
            ay = fma(facj, dy, ay);
            az = fma(facj, dz, az);
        } else {  // b = index2
            int p = a.pp.i1[k];
            double sh = a.pl.rhill[p] + hb;
            double ri = ((sh * sh) * (RHSCALE * RHSCALE)) * shellpow;
            double dx = xb - a.pl.rh.x[p], dy = yb - a.pl.rh.y[p], dz = zb - a.pl.rh.z[p];
            double r2 = dot3(dx, dy, dz, dx, dy, dz);
            double fac;
            if (!shell_fac(ri, r2, fac)) continue;
            any = true;
            double faci = fac * a.pl.Gm[p];
            ax = fma(-faci, dx, ax);
            ay = fma(-faci, dy, ay);
            az = fma(-faci, dz, az);
        }
    }
    if (any) {
        const double sdt = (double)sgn * dth;
        a.pl.vb.x[b] = fma(sdt, ax, a.pl.vb.x[b]);
        a.pl.vb.y[b] = fma(sdt, ay, a.pl.vb.y[b]);
        a.pl.vb.z[b] = fma(sdt, az, a.pl.vb.z[b]);
    }
}

// one shell kick of test particle slot q (symba_kick_list_pltp)
__device__ inline void kick_tp_body(const SubArgs &a, const LevelPtr &lp, int q, int irec, int sgn, double dth)
{
    const int t = lp.Bt[q];
    const int irecl = sgn < 0 ? irec - 1 : irec;
    const double shellpow = powi(RSHELL, 2 * irecl);
    const double xt = a.tp.rh.x[t], yt = a.tp.rh.y[t], zt = a.tp.rh.z[t];
    double ax = 0.0, ay = 0.0, az = 0.0;
    bool any = false;
    for (int u = lp.Ot[q]; u < lp.Ot[q + 1]; u++) {
        int k = lp.Et[u];
        if (a.pt.status[k] != ACTIVE) continue;
        int i = a.pt.i1[k];
        double hi = a.pl.rhill[i];
        double ri = ((hi * hi) * (RHSCALE * RHSCALE)) * shellpow;
        double dx = xt - a.pl.rh.x[i], dy = yt - a.pl.rh.y[i], dz = zt - a.pl.rh.z[i];
        double r2 = dot3(dx, dy, dz, dx, dy, dz);
        double fac;
        if (!shell_fac(ri, r2, fac)) continue;
        any = true;
        double faci = fac * a.pl.Gm[i];
        ax = fma(-faci, dx, ax);
        ay = fma(-faci, dy, ay);
        az = fma(-faci, dz, az);
    }
    if (any) {
        const double sdt = (double)sgn * dth;
        a.tp.vb.x[t] = fma(sdt, ax, a.tp.vb.x[t]);
        a.tp.vb.y[t] = fma(sdt, ay, a.tp.vb.y[t]);
        a.tp.vb.z[t] = fma(sdt, az, a.tp.vb.z[t]);
    }
}

// kicks of symba_step.f90:212-217 / :229-234 at level d (d >= 1: +1 then -1)
__device__ void phase_kicks(const SubArgs &a, const LevelRec &L, int d, double dth)
{
    const int tid = threadIdx.x;
    LevelPtr lp = level_ptr(a, L);
    const int irec = d + 1;
    for (int q = tid; q < L.nBpl; q += ST_T) {
        kick_pl_body(a, lp, q, irec, 1, dth);
        kick_pl_body(a, lp, q, irec, -1, dth);
    }
    for (int q = tid; q < L.nBtp; q += ST_T) {
        kick_tp_body(a, lp, q, irec, 1, dth);
        kick_tp_body(a, lp, q, irec, -1, dth);
    }
    __syncthreads();
}

// drift of the bodies at level d (symba_drift.f90:14-61)
__device__ void phase_drift(const SubArgs &a, int s, const LevelRec &L, int d, double dtl)
{
    const int tid = threadIdx.x;
    LevelPtr lp = level_ptr(a, L);
    const double mu = a.sy.GMcb[s];
    for (int q = tid; q < L.nBpl + L.nBtp; q += ST_T) {
        const bool ispl = q < L.nBpl;
        const int g = ispl ? lp.B[q] : lp.Bt[q - L.nBpl];
        Vec3p r = ispl ? a.pl.rh : a.tp.rh, v = ispl ? a.pl.vb : a.tp.vb;
        int32_t *status = ispl ? a.pl.status : a.tp.status;
        const int32_t *levelg = ispl ? a.pl.levelg : a.tp.levelg;
        bool m = ispl ? status[g] != INACTIVE : status[g] == ACTIVE;
        if (!m || levelg[g] != d) continue;
        double rx = r.x[g], ry = r.y[g], rz = r.z[g], vx = v.x[g], vy = v.y[g], vz = v.z[g];
        int fl = drift_one(mu, rx, ry, rz, vx, vy, vz, dtl);
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
    __syncthreads();
}

// collision detection at level d over the pairs whose index1 (pl-pl) or
// particle (pl-tp) is at levelg >= d
__device__ void phase_collision(const SubArgs &a, int s, const LevelRec &L, int d, double dtl)
{
    const int tid = threadIdx.x;
    LevelPtr lp = level_ptr(a, L);
    for (int q = tid; q < L.nBpl; q += ST_T) {
        const int i = lp.B[q];
        if (!a.has_pp) break;
        const int64_t r0 = a.pp.rowstart[i], r1 = r0 + a.pp.rowcnt[i];
        for (int64_t k = r0; k < r1; k++) {
            if (a.pp.status[k] != ACTIVE) continue;
            int j = a.pp.i2[k];
            bool c = collision_one(a.pl.rh.x[i] - a.pl.rh.x[j], a.pl.rh.y[i] - a.pl.rh.y[j],
                                   a.pl.rh.z[i] - a.pl.rh.z[j], a.pl.vb.x[i] - a.pl.vb.x[j],
                                   a.pl.vb.y[i] - a.pl.vb.y[j], a.pl.vb.z[i] - a.pl.vb.z[j],
                                   a.pl.Gm[i] + a.pl.Gm[j], a.pl.R[i] + a.pl.R[j], dtl, a.pp.lvdotr[k] != 0);
            if (c) {
                a.pp.status[k] = (int8_t)COLLIDED;
                atomicOr(&a.sy.err[s], 1);
            }
        }
    }
    for (int q = tid; q < L.nBtp; q += ST_T) {
        const int j = lp.Bt[q];
        const int64_t c0 = a.pt.colstart[j], c1 = c0 + a.pt.colcnt[j];
        for (int64_t c = c0; c < c1; c++) {
            int k = a.pt.col[c];
            if (a.pt.status[k] != ACTIVE) continue;
            int i = a.pt.i1[k];
            bool col = collision_one(a.pl.rh.x[i] - a.tp.rh.x[j], a.pl.rh.y[i] - a.tp.rh.y[j],
                                     a.pl.rh.z[i] - a.tp.rh.z[j], a.pl.vb.x[i] - a.tp.vb.x[j],
                                     a.pl.vb.y[i] - a.tp.vb.y[j], a.pl.vb.z[i] - a.tp.vb.z[j], a.pl.Gm[i],
                                     a.pl.R[i], dtl, a.pt.lvdotr[k] != 0);
            if (col) {
                a.pt.status[k] = (int8_t)COLLIDED;
                atomicOr(&a.sy.err[s], 1);
            }
        }
    }
    __syncthreads();
}

// symba_step_set_recur_levels_system at level d
__device__ void phase_set_levels(const SubArgs &a, const LevelRec &L, int d)
{
    const int tid = threadIdx.x;
    LevelPtr lp = level_ptr(a, L);
    const int irecp = d + 1;
    for (int q = tid; q < L.nBpl; q += ST_T)
        if (a.pl.levelg[lp.B[q]] == irecp) a.pl.levelg[lp.B[q]] = d;
    for (int q = tid; q < L.nBtp; q += ST_T)
        if (a.tp.levelg[lp.Bt[q]] == irecp) a.tp.levelg[lp.Bt[q]] = d;
    for (int q = tid; q < L.nEpl; q += ST_T) {
        uint32_t e = (uint32_t)lp.E[q];
        if (!(e & 1u)) continue;
        int k = (int)(e >> 1);
        if (a.pp.level[k] == irecp) a.pp.level[k] = (int8_t)d;
    }
    for (int q = tid; q < L.nEtp; q += ST_T) {
        int k = lp.Et[q];
        if (a.pt.level[k] == irecp) a.pt.level[k] = (int8_t)d;
    }
    __syncthreads();
}

// symba_step_recur_system from level 1 down, one workgroup per system
// The argument block is read through a pointer to HBM: taking the address
// of a by-value kernel argument this large would copy it to scratch.
__global__ __launch_bounds__(ST_T) void k_subtree(const SubArgs *__restrict__ ap)
{
    const SubArgs &a = *ap;
    __shared__ int sh[ST_W + 2];
    __shared__ int flags[2];
    __shared__ LevelRec lev[MAXLEV + 2];
    __shared__ int jst[MAXLEV + 2];
    const int s = blockIdx.x;
    if (!(a.sy.flag_pp[s] || a.sy.flag_pt[s])) return;
    const int tid = threadIdx.x;
    LevelRec l0;
    l0.base = -1;
    l0.nBpl = l0.nEpl = l0.nBtp = l0.nEtp = 0;
    LevelRec cur;
    if (!build_level(a, s, l0, 1, cur, sh)) {
        if (tid == 0) {
            atomicOr(&a.stats[2], 1);
            atomicOr(&a.sy.err[s], 4);
        }
        return;
    }
    int d = 1, j = 1;
    bool resume = false;
    int nsub = 0, maxd = 1;
    for (;;) {
        double p3 = 1.0;
        for (int q = 0; q < d; q++) p3 *= NTENC;
        const double dtl = a.dt / p3, dth = 0.5 * dtl;
        if (!resume) {
            if (j > NTENC) {  // level finished
                if (d == 1) break;
                d--;
                cur = lev[d];
                j = jst[d];
                resume = true;
                continue;
            }
            if (dtl / a.dt < 2.220446049250313e-16) {  // symba_step.f90:196-201
                if (tid == 0) atomicOr(&a.sy.err[s], 8);
                break;
            }
            nsub++;
            const bool prof = a.prof && tid == 0;
            unsigned long long t0 = prof ? wall_clock64() : 0, t1, t2, t3;
            if (prof) {
                a.prof[d * NPROF + 0] += 1;
                a.prof[d * NPROF + 1] += cur.nBpl;
                a.prof[d * NPROF + 2] += cur.nEpl;
            }
            bool lenc = phase_check(a, s, cur, d, dtl, flags);
            if (prof) t1 = wall_clock64();
            phase_kicks(a, cur, d, dth);
            if (prof) t2 = wall_clock64();
            phase_drift(a, s, cur, d, dtl);
            if (prof) {
                t3 = wall_clock64();
                a.prof[d * NPROF + 3] += t1 - t0;
                a.prof[d * NPROF + 4] += t2 - t1;
                a.prof[d * NPROF + 5] += t3 - t2;
            }
            if (lenc) {
                if (d + 1 > MAXLEV) {
                    if (tid == 0) atomicOr(&a.sy.err[s], 8);
                    break;
                }
                LevelRec nxt;
                unsigned long long tb = prof ? wall_clock64() : 0;
                bool ok = build_level(a, s, cur, d + 1, nxt, sh);
                if (prof) a.prof[d * NPROF + 6] += wall_clock64() - tb;
                if (!ok) {
                    if (tid == 0) {
                        atomicOr(&a.stats[2], 1);
                        atomicOr(&a.sy.err[s], 4);
                    }
                    break;
                }
                if (tid == 0) {
                    lev[d] = cur;
                    jst[d] = j;
                }
                __syncthreads();
                d++;
                if (d > maxd) maxd = d;
                cur = nxt;
                j = 1;
                continue;
            }
        }
        resume = false;
        {
            const bool prof = a.prof && tid == 0;
            unsigned long long t0 = prof ? wall_clock64() : 0;
            phase_kicks(a, cur, d, dth);
            unsigned long long t1 = prof ? wall_clock64() : 0;
            if (a.lclose) phase_collision(a, s, cur, d, dtl);
            phase_set_levels(a, cur, d);
            if (prof) {
                a.prof[d * NPROF + 4] += t1 - t0;
                a.prof[d * NPROF + 7] += wall_clock64() - t1;
            }
        }
        j++;
    }
    if (tid == 0) {
        atomicAdd(&a.stats[0], nsub);
        atomicMax(&a.stats[1], maxd);
    }
}

// per-system arena sizes in words: room for the deepest chain of nested
// levels, each at most the size of level 1
__global__ void k_arena_caps(ListDev pp, ListDev pt, SysDev sy, int nsys, int has_pp, int has_pt, int32_t *caps,
                             int mult)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1], t0 = sy.tp_off[s], t1 = sy.tp_off[s + 1];
    int64_t npp = 0, npt = 0;
    if (has_pp && p1 > p0) npp = pp.rowstart[p1 - 1] + pp.rowcnt[p1 - 1] - pp.rowstart[p0];
    if (has_pt && p1 > p0) npt = pt.rowstart[p1 - 1] + pt.rowcnt[p1 - 1] - pt.rowstart[p0];
    int64_t w = 0;
    if (sy.flag_pp[s] || sy.flag_pt[s])
        w = (int64_t)mult * (2 * npp + npt + 3 * (int64_t)(p1 - p0) + 3 * (int64_t)(t1 - t0) + 8) + 64;
    caps[s] = (int32_t)(w > 0x7fffffff ? 0x7fffffff : w);
}

void launch_arena_caps(hipStream_t st, ListDev pp, ListDev pt, SysDev sy, int nsys, int has_pp, int has_pt,
                       int32_t *caps, int mult)
{
    k_arena_caps<<<(nsys + 255) / 256, 256, 0, st>>>(pp, pt, sy, nsys, has_pp, has_pt, caps, mult);
}

size_t subtree_args_bytes() { return sizeof(SubArgs); }

// diagnostic profile of the last launch (SWG_PROF_SUBTREE=1), printed to stderr
static unsigned long long *g_prof = nullptr;
void subtree_profile_report(hipStream_t st)
{
    if (!g_prof) return;
    unsigned long long h[NPROF * (MAXLEV + 2)];
    (void)hipStreamSynchronize(st);
    (void)hipMemcpy(h, g_prof, sizeof h, hipMemcpyDeviceToHost);
    const double tick_us = 0.01;  // wall_clock64 runs at 100 MHz
    fprintf(stderr, "[swg] level substeps   avg_nB   avg_nE  check_ms  kicks_ms  drift_ms  build_ms levels_ms\n");
    for (int d = 1; d <= MAXLEV; d++) {
        const unsigned long long *p = h + d * NPROF;
        if (!p[0]) continue;
        fprintf(stderr, "[swg] %5d %8llu %8.0f %8.0f %9.2f %9.2f %9.2f %9.2f %9.2f\n", d, p[0], (double)p[1] / p[0],
                (double)p[2] / p[0], p[3] * tick_us * 1e-3, p[4] * tick_us * 1e-3, p[5] * tick_us * 1e-3,
                p[6] * tick_us * 1e-3, p[7] * tick_us * 1e-3);
    }
}

// h_stage: pinned host memory, d_args: device memory, both of
// subtree_args_bytes(); the staging buffer must not be rewritten before the
// copy has executed.
void launch_subtree(hipStream_t st, PlDev pl, TpDev tp, ListDev pp, ListDev pt, SysDev sy, int nsys, double dt,
                    int lclose, int short_circuit, int has_pp, int has_pt, int32_t *arena, const int64_t *arena_off,
                    int32_t *stats, void *h_stage, void *d_args)
{
    SubArgs &a = *(SubArgs *)h_stage;
    a.pl = pl;
    a.tp = tp;
    a.pp = pp;
    a.pt = pt;
    a.sy = sy;
    a.dt = dt;
    a.lclose = lclose;
    a.short_circuit = short_circuit;
    a.has_pp = has_pp;
    a.has_pt = has_pt;
    a.arena = arena;
    a.arena_off = arena_off;
    a.stats = stats;
    a.prof = nullptr;
    if (getenv("SWG_PROF_SUBTREE") && getenv("SWG_PROF_SUBTREE")[0] == '1') {
        if (!g_prof) (void)hipMalloc((void **)&g_prof, sizeof(unsigned long long) * NPROF * (MAXLEV + 2));
        (void)hipMemsetAsync(g_prof, 0, sizeof(unsigned long long) * NPROF * (MAXLEV + 2), st);
        a.prof = g_prof;
    }
    (void)hipMemcpyAsync(d_args, h_stage, sizeof(SubArgs), hipMemcpyHostToDevice, st);
    k_subtree<<<nsys, ST_T, 0, st>>>((const SubArgs *)d_args);
}

}  // namespace swg
