// k_subtree.hip -- the SyMBA recursion from a level d0 >= 1 down, one
// persistent workgroup per system.
//
// Once the whole-device working set of level d0 is small (k_level.hip), each
// system's slice of it is handed to one workgroup, which executes the rest of
// that system's recursion (symba_step_recur_system, symba_step.f90:163-262)
// with an explicit level stack: list check (symba_encounter_check.f90:90-235),
// shell kicks (symba_kick.f90:126-340), level-masked drift (symba_drift.f90:
// 14-61), descent, kicks, collision detection (collision_check.f90:53-230)
// and the level reset (symba_step.f90:118-160), separated by workgroup
// barriers, with no host round trip.
//
// On descent the level-(d+1) set is built by filtering the level-d one with
// order-preserving workgroup scans, so each body's pairs stay in list order
// and every per-body shell-kick sum keeps the reference's serial order.  The
// nested sets live in a per-system arena in HBM, in the layout of
// level_layout(); building a level puts 2*(nB+nBt) words of temporaries after
// the source level and the new level after them.
#include <cstdio>
#include <cstdlib>

#include "swg_kernels.h"
#include "swg_level.h"

namespace swg {

constexpr int ST_T = 1024;        // threads of the persistent workgroup
constexpr int ST_W = ST_T / 64;  // wavefronts

// A level of one system: its working set (a slice of a whole-device level or
// a level stored in the system's arena) and the arena word offset after it.
struct LevelRec {
    const int32_t *B, *O, *Bt, *Ot;
    const int2 *E, *Et;
    int nB, nBt;
    int64_t end;
};

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

// Slices small enough for the LDS-resident kernel (k_subtree_lds)
constexpr int L_T = 256;                  // threads of its workgroup
constexpr int L_NB = 512, L_NE = 4096;    // pl bodies, pl entries (pairs = entries / 2)
constexpr int L_NP = L_NE / 2;
constexpr int L_NBT = 256, L_NET = 1024;  // tp bodies, tp entries

__device__ inline bool lds_eligible(int nB, int nE, int nBt, int nEt)
{
    return nB <= L_NB && nE <= L_NE && nBt <= L_NBT && nEt <= L_NET;
}

struct SubArgs {
    RecState st;
    double dt;  // param%dt
    int lclose, short_circuit, has_pp, has_pt;
    int lds;             // 0: never use the LDS-resident kernel
    int wave;            // 0: never hand small levels to one wavefront
    LevelView start;    // whole-device level d0; system s owns a slice of it
    const int32_t *rng;  // per system: pl begin, pl end, tp begin, tp end in `start`
    int d0;
    int32_t *arena;
    const int64_t *arena_off;  // nsys+1 word offsets
    int32_t *stats;            // [0] substeps, [1] max depth, [2] overflow
    // optional per-level profile (NPROF words per level): substeps, sum nB,
    // sum nE, then wall-clock ticks in check, kicks, drift, build, levels; the
    // LDS kernel adds kick evaluate and fold ticks, sum of the largest entry
    // count of a kicked body, sum of kicked bodies, bodies drifted, substeps
    // where a thread drifted more than one body, and list builds
    unsigned long long *prof;
};
constexpr int NPROF = 16;

// Build level dnew of system s from src into the arena after src.  Returns
// false (uniformly) when the arena is too small.
__device__ bool build_level(const SubArgs &A, int s, const LevelRec &src, int dnew, LevelRec &dst, int *sh)
{
    const RecState &a = A.st;
    const int tid = threadIdx.x;
    const int nBpl = src.nB, nBtp = src.nBt;
    const int64_t tbase = src.end;
    const int64_t aend = A.arena_off[s + 1];
    if (tbase + 2 * (int64_t)(nBpl + nBtp) > aend) return false;
    int32_t *posP = A.arena + tbase, *offP = posP + nBpl, *posT = offP + nBpl, *offT = posT + nBtp;
    // pass 1: kept-entry counts (-1: body dropped)
    for (int q = tid; q < nBpl; q += ST_T) {
        int c = -1;
        if (a.pl.levelg[src.B[q]] >= dnew) {
            c = 0;
            for (int u = src.O[q]; u < src.O[q + 1]; u++) c += a.pl.levelg[src.E[u].y] >= dnew;
        }
        offP[q] = c;
    }
    for (int q = tid; q < nBtp; q += ST_T) {
        int c = -1;
        if (a.tp.levelg[src.Bt[q]] >= dnew) {
            c = 0;
            for (int u = src.Ot[q]; u < src.Ot[q + 1]; u++) c += a.pl.levelg[src.Et[u].y] >= dnew;
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
    int64_t base = tbase + 2 * (int64_t)(nBpl + nBtp);
    base += base & 1;
    const LevelLayout L = level_layout(nb, ne, nbt, net);
    if (base + L.words > aend) return false;
    int32_t *w = A.arena + base;
    int32_t *B = w + L.oB, *O = w + L.oO, *Bt = w + L.oBt, *Ot = w + L.oOt;
    int2 *E = (int2 *)(w + L.oE), *Et = (int2 *)(w + L.oEt);
    dst.B = B;
    dst.O = O;
    dst.E = E;
    dst.Bt = Bt;
    dst.Ot = Ot;
    dst.Et = Et;
    dst.nB = nb;
    dst.nBt = nbt;
    dst.end = base + L.words;
    __syncthreads();
    // pass 2: fill
    for (int q = tid; q < nBpl; q += ST_T) {
        int pos = posP[q];
        if (pos < 0) continue;
        B[pos] = src.B[q];
        int o = offP[q];
        O[pos] = o;
        for (int u = src.O[q]; u < src.O[q + 1]; u++) {
            int2 en = src.E[u];
            if (a.pl.levelg[en.y] >= dnew) E[o++] = en;
        }
    }
    for (int q = tid; q < nBtp; q += ST_T) {
        int pos = posT[q];
        if (pos < 0) continue;
        Bt[pos] = src.Bt[q];
        int o = offT[q];
        Ot[pos] = o;
        for (int u = src.Ot[q]; u < src.Ot[q + 1]; u++) {
            int2 en = src.Et[u];
            if (a.pl.levelg[en.y] >= dnew) Et[o++] = en;
        }
    }
    if (tid == 0) {
        O[nb] = ne;
        Ot[nbt] = net;
    }
    __syncthreads();
    return true;
}

// ---- phases of one substep at level d ------------------------------------------

// list checks at level d+1 (pairs at level d); returns lenc (uniform)
__device__ bool phase_check(const SubArgs &A, const LevelRec &L, int d, double dtl, int *flags)
{
    const RecState &a = A.st;
    const int tid = threadIdx.x;
    if (tid == 0) {
        flags[0] = 0;
        flags[1] = 0;
    }
    __syncthreads();
    for (int u = L.O[0] + tid; u < L.O[L.nB]; u += ST_T) {
        const int2 en = L.E[u];
        if (!(en.x & 1)) continue;  // each pair once, from its index1 body
        if (rec_check_pp(a, (int)((uint32_t)en.x >> 1), d, dtl)) flags[0] = 1;
    }
    __syncthreads();
    if (!(A.short_circuit && flags[0])) {  // symba_step.f90:210 (.or. short-circuited)
        for (int u = L.Ot[0] + tid; u < L.Ot[L.nBt]; u += ST_T)
            if (rec_check_pt(a, L.Et[u].x, d, dtl)) flags[1] = 1;
    }
    __syncthreads();
    bool lenc = flags[0] || flags[1];
    __syncthreads();
    return lenc;
}

// kicks of symba_step.f90:212-217 / :229-234 at level d (+1 then -1, fused)
__device__ void phase_kicks(const SubArgs &A, const LevelRec &L, int d, double dth)
{
    const int tid = threadIdx.x;
    for (int q = tid; q < L.nB; q += ST_T) rec_kick_pl(A.st, L.B[q], L.E, L.O[q], L.O[q + 1], d, dth);
    for (int q = tid; q < L.nBt; q += ST_T) rec_kick_tp(A.st, L.Bt[q], L.Et, L.Ot[q], L.Ot[q + 1], d, dth);
    __syncthreads();
}

// drift of the bodies at level d (symba_drift.f90:14-61)
__device__ void phase_drift(const SubArgs &A, int s, const LevelRec &L, int d, double dtl)
{
    for (int q = threadIdx.x; q < L.nB + L.nBt; q += ST_T) {
        const bool ispl = q < L.nB;
        rec_drift(A.st, ispl ? L.B[q] : L.Bt[q - L.nB], ispl, d, dtl, s);
    }
    __syncthreads();
}

// collision detection at level d over the pairs whose index1 (pl-pl) or
// particle (pl-tp) is at levelg >= d
__device__ void phase_collision(const SubArgs &A, int s, const LevelRec &L, double dtl)
{
    for (int q = threadIdx.x; q < L.nB + L.nBt; q += ST_T) {
        if (q < L.nB) {
            if (A.has_pp) rec_collide_pl(A.st, L.B[q], dtl, s);
        } else {
            rec_collide_tp(A.st, L.Bt[q - L.nB], dtl, s);
        }
    }
    __syncthreads();
}

// symba_step_set_recur_levels_system at level d
__device__ void phase_set_levels(const SubArgs &A, const LevelRec &L, int d)
{
    const RecState &a = A.st;
    const int tid = threadIdx.x;
    const int irecp = d + 1;
    for (int q = tid; q < L.nB; q += ST_T)
        if (a.pl.levelg[L.B[q]] == irecp) a.pl.levelg[L.B[q]] = d;
    for (int q = tid; q < L.nBt; q += ST_T)
        if (a.tp.levelg[L.Bt[q]] == irecp) a.tp.levelg[L.Bt[q]] = d;
    for (int u = L.O[0] + tid; u < L.O[L.nB]; u += ST_T) {
        const int2 en = L.E[u];
        const int k = (int)((uint32_t)en.x >> 1);
        if ((en.x & 1) && a.pp.level[k] == irecp) a.pp.level[k] = (int8_t)d;
    }
    for (int u = L.Ot[0] + tid; u < L.Ot[L.nBt]; u += ST_T) {
        const int k = L.Et[u].x;
        if (a.pt.level[k] == irecp) a.pt.level[k] = (int8_t)d;
    }
    __syncthreads();
}

// symba_step_recur_system from level d0 down, one workgroup per system.  The
// argument block is read through a pointer to HBM: taking the address of a
// by-value kernel argument this large would copy it to scratch.
__global__ __launch_bounds__(ST_T) void k_subtree(const SubArgs A)
{
    __shared__ int sh[ST_W + 2];
    __shared__ int flags[2];
    __shared__ LevelRec lev[MAXLEV + 2];
    __shared__ int jst[MAXLEV + 2];
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int32_t *r = A.rng + 4 * s;
    LevelRec cur;
    cur.B = A.start.B + r[0];
    cur.O = A.start.O + r[0];
    cur.E = A.start.E;
    cur.nB = r[1] - r[0];
    cur.Bt = A.start.Bt + r[2];
    cur.Ot = A.start.Ot + r[2];
    cur.Et = A.start.Et;
    cur.nBt = r[3] - r[2];
    cur.end = A.arena_off[s];
    cur.end += cur.end & 1;
    if (cur.nB + cur.nBt == 0) return;
    if (A.lds && lds_eligible(cur.nB, cur.O[cur.nB] - cur.O[0], cur.nBt, cur.Ot[cur.nBt] - cur.Ot[0])) return;
    const int d0 = A.d0;
    int d = d0, j = 1;
    bool resume = false;
    int nsub = 0, maxd = d0;
    for (;;) {
        double p3 = 1.0;
        for (int q = 0; q < d; q++) p3 *= NTENC;
        const double dtl = A.dt / p3, dth = 0.5 * dtl;
        if (!resume) {
            if (j > NTENC) {  // level finished
                if (d == d0) break;
                d--;
                cur = lev[d];
                j = jst[d];
                resume = true;
                continue;
            }
            if (dtl / A.dt < 2.220446049250313e-16) {  // symba_step.f90:196-201
                if (tid == 0) atomicOr(&A.st.sy.err[s], 8);
                break;
            }
            nsub++;
            const bool prof = A.prof && tid == 0;
            unsigned long long t0 = prof ? wall_clock64() : 0, t1, t2, t3;
            if (prof) {
                A.prof[d * NPROF + 0] += 1;
                A.prof[d * NPROF + 1] += cur.nB;
                A.prof[d * NPROF + 2] += cur.O[cur.nB] - cur.O[0];
            }
            bool lenc = phase_check(A, cur, d, dtl, flags);
            if (prof) t1 = wall_clock64();
            phase_kicks(A, cur, d, dth);
            if (prof) t2 = wall_clock64();
            phase_drift(A, s, cur, d, dtl);
            if (prof) {
                t3 = wall_clock64();
                A.prof[d * NPROF + 3] += t1 - t0;
                A.prof[d * NPROF + 4] += t2 - t1;
                A.prof[d * NPROF + 5] += t3 - t2;
            }
            if (lenc) {
                if (d + 1 > MAXLEV) {
                    if (tid == 0) atomicOr(&A.st.sy.err[s], 8);
                    break;
                }
                LevelRec nxt;
                unsigned long long tb = prof ? wall_clock64() : 0;
                bool ok = build_level(A, s, cur, d + 1, nxt, sh);
                if (prof) A.prof[d * NPROF + 6] += wall_clock64() - tb;
                if (!ok) {
                    if (tid == 0) {
                        atomicOr(&A.stats[2], 1);
                        atomicOr(&A.st.sy.err[s], 4);
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
            const bool prof = A.prof && tid == 0;
            unsigned long long t0 = prof ? wall_clock64() : 0;
            phase_kicks(A, cur, d, dth);
            unsigned long long t1 = prof ? wall_clock64() : 0;
            if (A.lclose) phase_collision(A, s, cur, dtl);
            phase_set_levels(A, cur, d);
            if (prof) {
                A.prof[d * NPROF + 4] += t1 - t0;
                A.prof[d * NPROF + 7] += wall_clock64() - t1;
            }
        }
        j++;
    }
    if (tid == 0) {
        atomicAdd(&A.stats[0], nsub);
        atomicMax(&A.stats[1], maxd);
    }
}

// ---- the LDS-resident variant -------------------------------------------------
//
// A system whose slice of the level-d0 set fits in LDS is run from LDS: the
// bodies' state, the entries (with local indices) and the pairs' status,
// level and lvdotr are loaded once, the whole subtree runs on them, and they
// are written back at the end.  No nested set is built: during level d the
// working set is exactly the bodies with levelg >= d and the entries whose
// partner has levelg >= d, so every phase filters the d0 slice on the fly.
// Entry word: (local pair << 10) | (role << 9) | local partner.

// one evaluated shell term: separation, the +1 and -1 factors (0 where that
// shell contributes nothing) and which of the two contribute (bits 0 and 1)
struct KTerm {
    double dx, dy, dz, fp, fm;
    int fl;
};
// A body's shell-kick sums kept from the closing kicks of a substep for the
// opening kicks of the next one at the same level.  Between the two only the
// level reset, the encounter check and (not in wave mode) the collision test
// run: positions, masses, Hill radii and pair statuses are unchanged (the
// collision test is off when they are kept), and no body's level crosses the
// current one, so the reference computes the same sums again
// (symba_step.f90:208-247, symba_kick.f90:126-230).
struct KCache {
    double3 p, m;
    int fl;
};

constexpr int L_KC = 512;                              // entry terms per kick chunk
constexpr int L_KS = (L_NB + L_NBT + L_T - 1) / L_T;  // bodies folded per lane
static_assert(L_KS * L_T == L_NB + L_NBT, "the kick-list arrays cover every lane slot");

struct LdsSet {
    double px[L_NB], py[L_NB], pz[L_NB], pvx[L_NB], pvy[L_NB], pvz[L_NB], ph[L_NB], pgm[L_NB], pR[L_NB];
    int pg[L_NB], plev[L_NB], plm[L_NB], pst[L_NB], pO[L_NB + 1];
    uint32_t pE[L_NE];
    int16_t ppi[L_NP], ppj[L_NP];
    int ppg[L_NP];
    int8_t pps[L_NP], ppl[L_NP], ppv[L_NP];
    double tx[L_NBT], ty[L_NBT], tz[L_NBT], tvx[L_NBT], tvy[L_NBT], tvz[L_NBT];
    int tg[L_NBT], tlev[L_NBT], tlm[L_NBT], tst[L_NBT], tO[L_NBT + 1];
    int16_t tpl[L_NET], tpt[L_NET];
    int tpg[L_NET];
    int8_t tps[L_NET], tpl_lev[L_NET], tpv[L_NET];
    // shell kicks: the level's list (kicked bodies, their stretches of the
    // kept entries) and one chunk of evaluated terms
    int16_t act[L_NB + L_NBT], kst[L_NB + L_NBT], ken[L_NB + L_NBT];
    int16_t lst[L_NE + L_NET];
    union {
        KTerm kt[L_KC];
        int16_t posu[L_NE + L_NET];  // while the list is built
        int lab[L_NB + L_NBT];       // connected-component labels, profiling only
    };
    int jst[MAXLEV + 2];
    int wsum[L_T / 64 + 1];
    // wave mode (one wavefront runs a small sub-recursion): the level it
    // started at, that level's bodies (pl, then tp as nB + t), pl pairs, tp
    // entries, and the kick entries of its bodies grouped by body
    int16_t wB[64], wP[64], wT[64], wE[64];
    int8_t wEs[64], wEe[64], wH[64];  // wH: slots of the bodies the second wavefront drifts
    int wnB, wnP, wnT, wnE, wstop, wnH, wlenc;
    int npair, flag_pp, flag_pt, bad, nact, nlst, lst_level, kmax;
    unsigned long long prof[(MAXLEV + 2) * NPROF];  // per-level phase ticks when profiling
};

__device__ inline int lds_lower_bound(const int *a, int n, int key)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (a[m] < key) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// index of the segment [o[q], o[q+1]) that holds u
__device__ inline int lds_owner(const int *o, int n, int u)
{
    int lo = 0, hi = n;  // o[0] = 0 <= u < o[n]
    while (hi - lo > 1) {
        int m = (lo + hi) >> 1;
        if (o[m] <= u) lo = m;
        else hi = m;
    }
    return lo;
}

__device__ bool lds_check(LdsSet &S, const SubArgs &A, int nB, int nEt, int d, double dtl)
{
    const int tid = threadIdx.x, irec = d + 1;
    if (tid == 0) {
        S.flag_pp = 0;
        S.flag_pt = 0;
    }
    __syncthreads();
    for (int lp = tid; lp < S.npair; lp += L_T) {
        if (S.pps[lp] != ACTIVE || S.ppl[lp] != d) continue;
        const int i = S.ppi[lp], j = S.ppj[lp];
        bool lv;
        bool le = check_pp_v(S.px[i], S.py[i], S.pz[i], S.pvx[i], S.pvy[i], S.pvz[i], S.ph[i], S.pR[i], S.px[j],
                             S.py[j], S.pz[j], S.pvx[j], S.pvy[j], S.pvz[j], S.ph[j], S.pR[j], irec, dtl, lv);
        S.ppv[lp] = lv;
        if (le) {
            S.ppl[lp] = (int8_t)irec;
            S.plev[i] = irec;
            S.plev[j] = irec;
            atomicMax(&S.plm[i], irec);
            atomicMax(&S.plm[j], irec);
            S.flag_pp = 1;
        }
    }
    __syncthreads();
    if (!(A.short_circuit && S.flag_pp)) {  // symba_step.f90:210 (.or. short-circuited)
        for (int u = tid; u < nEt; u += L_T) {
            if (S.tps[u] != ACTIVE || S.tpl_lev[u] != d) continue;
            const int i = S.tpl[u], t = S.tpt[u];
            bool lv;
            bool le = check_pt_v(S.px[i], S.py[i], S.pz[i], S.pvx[i], S.pvy[i], S.pvz[i], S.ph[i], S.pR[i], S.tx[t],
                                 S.ty[t], S.tz[t], S.tvx[t], S.tvy[t], S.tvz[t], irec, dtl, lv);
            S.tpv[u] = lv;
            if (le) {
                S.tpl_lev[u] = (int8_t)irec;
                S.plev[i] = irec;
                atomicMax(&S.plm[i], irec);
                S.tlev[t] = irec;
                atomicMax(&S.tlm[t], irec);
                S.flag_pt = 1;
            }
        }
    }
    __syncthreads();
    const bool lenc = S.flag_pp || S.flag_pt;
    __syncthreads();
    return lenc;
}

__device__ inline void lds_apply(double &vx, double &vy, double &vz, double dth, bool anyp, const double3 &p,
                                 bool anym, const double3 &m)
{
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
}

// Fold terms [u, hi) of a chunk into the running sums, in order.  A term
// that does not add to a sum holds a zero factor for it, and fma(0, d, s) is
// s bit for bit: d is finite, and a sum that starts at +0 is never -0 (an fma
// gives -0 only from -0 operands).  So every term is folded without a
// select, and the loads run two terms ahead of the arithmetic.
__device__ inline void fold_terms(const KTerm *kt, int u, int hi, double3 &p, double3 &m, int &fl)
{
    if (u >= hi) return;
    KTerm a = kt[u], b = kt[min(u + 1, hi - 1)];
    for (; u < hi; u++) {
        const KTerm c = kt[min(u + 2, hi - 1)];
        p.x = fma(a.fp, a.dx, p.x);
        p.y = fma(a.fp, a.dy, p.y);
        p.z = fma(a.fp, a.dz, p.z);
        m.x = fma(a.fm, a.dx, m.x);
        m.y = fma(a.fm, a.dy, m.y);
        m.z = fma(a.fm, a.dz, m.z);
        fl |= a.fl;
        a = b;
        b = c;
    }
}

// workgroup exclusive scan of one value per thread of L_T; *tot = the sum
__device__ inline int lds_block_scan(LdsSet &S, int v, int *tot)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int incl = v;
    for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) S.wsum[w] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < L_T / 64; k++) {
            const int t = S.wsum[k];
            S.wsum[k] = acc;
            acc += t;
        }
        S.wsum[L_T / 64] = acc;
    }
    __syncthreads();
    const int r = S.wsum[w] + incl - v;
    *tot = S.wsum[L_T / 64];
    __syncthreads();
    return r;
}

// whether entry u (pl entries, then test-particle entries) takes part in the
// shell kicks of level d: its body and its partner are both in the level
__device__ inline bool lds_entry_in_level(const LdsSet &S, int u, int nE, int d)
{
    if (u < nE) {
        const uint32_t e = S.pE[u];
        const int lp = (int)(e >> 10), pr = (int)(e & 511u);
        const int q = ((e >> 9) & 1u) ? S.ppi[lp] : S.ppj[lp];
        return S.plev[q] >= d && S.plev[pr] >= d;
    }
    const int v = u - nE;
    return S.tlev[S.tpt[v]] >= d && S.plev[S.tpl[v]] >= d;
}

// The kick list of level d: the entries that take part, in entry order
// (grouped by body, since each body's entries are contiguous), and the bodies
// that have any, with their stretch of the list.  Which bodies and entries
// take part depends only on the levels being >= d, and no body's level
// crosses d while level d is being stepped (the list checks only raise levels
// above d, and the level reset only lowers d+1 back to d), so the list is
// built when level d is entered or resumed and serves all its kicks.  Pair
// statuses can change (collisions) and are read when a term is evaluated.
__device__ void lds_level_list(LdsSet &S, int nB, int nBt, int nE, int nEt, int d)
{
    const int tid = threadIdx.x, N = nE + nEt;
    const int seg = (N + L_T - 1) / L_T, u0 = min(N, tid * seg), u1 = min(N, u0 + seg);
    int cnt = 0;
    for (int u = u0; u < u1; u++) {
        const bool k = lds_entry_in_level(S, u, nE, d);
        S.posu[u] = (int16_t)k;
        cnt += k;
    }
    int tot;
    int pos = lds_block_scan(S, cnt, &tot);
    for (int u = u0; u < u1; u++) {
        const bool k = S.posu[u] != 0;
        S.posu[u] = (int16_t)pos;
        if (k) S.lst[pos++] = (int16_t)u;
    }
    if (tid == 0) {
        S.nact = 0;
        S.nlst = tot;
    }
    __syncthreads();
    for (int q = tid; q < nB + nBt; q += L_T) {
        const bool ispl = q < nB;
        if ((ispl ? S.plev[q] : S.tlev[q - nB]) < d) continue;
        const int e0 = ispl ? S.pO[q] : nE + S.tO[q - nB], e1 = ispl ? S.pO[q + 1] : nE + S.tO[q - nB + 1];
        if (e1 <= e0) continue;
        const int f0 = S.posu[e0], f1 = e1 < N ? S.posu[e1] : tot;
        if (f1 <= f0) continue;
        const int ai = atomicAdd(&S.nact, 1);
        S.act[ai] = (int16_t)q;
        S.kst[ai] = (int16_t)f0;
        S.ken[ai] = (int16_t)f1;
    }
    if (tid == 0) S.lst_level = d;
    __syncthreads();
}

// The +1 / -1 shell kicks of level d (the sums of rec_kick_pl / rec_kick_tp).
// The level's kick list is taken in chunks of L_KC entries: all threads
// evaluate the chunk's terms at once into LDS, then every listed body folds
// the terms of its own entries in the chunk, in entry order, on one lane.  The
// running sums stay in registers across chunks (a lane serves up to L_KS
// bodies), so each body's sum is the serial sum over its entries.  Entries
// left out of the list (partner below the level) contribute no term, and a
// body with no listed entry gets no kick, as in the reference.
// mode 0: evaluate; 1: evaluate and keep each folded body's sums in kc; 2:
// apply the sums kept in kc by the previous call, the closing kicks of the
// substep before at this level (the same lane folds the same bodies while the
// level's list stands).
__device__ __attribute__((always_inline)) inline void lds_kicks(LdsSet &S, int nB, int nBt, int nE, int nEt, int d,
                                                                double dth, bool prof, int mode, KCache (&kc)[L_KS])
{
    const int tid = threadIdx.x;
    if (mode == 2) {
        const int nact = S.nact;
#pragma unroll
        for (int k = 0; k < L_KS; k++) {
            const int ai = tid + k * L_T;
            if (ai >= nact) continue;
            const int q = S.act[ai];
            const bool ap = (kc[k].fl & 1) != 0, am = (kc[k].fl & 2) != 0;
            if (q < nB) lds_apply(S.pvx[q], S.pvy[q], S.pvz[q], dth, ap, kc[k].p, am, kc[k].m);
            else lds_apply(S.tvx[q - nB], S.tvy[q - nB], S.tvz[q - nB], dth, ap, kc[k].p, am, kc[k].m);
        }
        __syncthreads();
        return;
    }
    if (S.lst_level != d) {
        lds_level_list(S, nB, nBt, nE, nEt, d);
        if (prof && tid == 0) S.prof[d * NPROF + 14] += 1;
    }
    const int nact = S.nact, nl = S.nlst;
    const bool pr = prof && tid == 0;
    unsigned long long tc = 0, tf = 0, tt = pr ? wall_clock64() : 0;
    const double spp = shell_sq(d + 1), spm = shell_sq(d);
    int fe0[L_KS], fe1[L_KS], kmax = 0;
    double3 p[L_KS], m[L_KS];
    bool anyp[L_KS], anym[L_KS];
#pragma unroll
    for (int k = 0; k < L_KS; k++) {
        // read unconditionally (every ai < L_KS * L_T is inside the arrays),
        // so the loads issue together
        const int ai = tid + k * L_T;
        const int f0 = S.kst[ai], f1 = S.ken[ai];
        fe0[k] = ai < nact ? f0 : 0;
        fe1[k] = ai < nact ? f1 : 0;
        kmax = max(kmax, fe1[k] - fe0[k]);
        p[k] = m[k] = make_double3(0.0, 0.0, 0.0);
        anyp[k] = anym[k] = false;
    }
    for (int c0 = 0; c0 < nl; c0 += L_KC) {
        for (int w = tid; w < L_KC && c0 + w < nl; w += L_T) {
            const int u = S.lst[c0 + w];
            ShellTerm t;
            t.tp = t.tm = false;
            t.dx = t.dy = t.dz = 0.0;
            if (u < nE) {
                const uint32_t e = S.pE[u];
                const int lp = (int)(e >> 10), pr = (int)(e & 511u);
                const bool bi = (e >> 9) & 1u;
                const int q = bi ? S.ppi[lp] : S.ppj[lp];
                if (S.pps[lp] == ACTIVE)
                    pl_shell_term_v(bi, S.px[q], S.py[q], S.pz[q], S.ph[q], S.px[pr], S.py[pr], S.pz[pr], S.ph[pr],
                                    S.pgm[pr], spp, spm, true, t);
            } else {
                const int v = u - nE, i = S.tpl[v], tt = S.tpt[v];
                if (S.tps[v] == ACTIVE)
                    tp_shell_term_v(S.tx[tt], S.ty[tt], S.tz[tt], S.px[i], S.py[i], S.pz[i], S.ph[i], S.pgm[i], spp,
                                    spm, true, t);
            }
            KTerm &K = S.kt[w];
            K.dx = t.dx;
            K.dy = t.dy;
            K.dz = t.dz;
            K.fp = t.tp ? t.fp : 0.0;
            K.fm = t.tm ? t.fm : 0.0;
            K.fl = (t.tp ? 1 : 0) | (t.tm ? 2 : 0);
        }
        __syncthreads();
        if (pr) {
            const unsigned long long now = wall_clock64();
            tc += now - tt;
            tt = now;
        }
#pragma unroll
        for (int k = 0; k < L_KS; k++) {
            int fl = 0;
            fold_terms(S.kt, max(fe0[k], c0) - c0, min(fe1[k], c0 + L_KC) - c0, p[k], m[k], fl);
            anyp[k] |= (fl & 1) != 0;
            anym[k] |= (fl & 2) != 0;
        }
        __syncthreads();
        if (pr) {
            const unsigned long long now = wall_clock64();
            tf += now - tt;
            tt = now;
        }
    }
#pragma unroll
    for (int k = 0; k < L_KS; k++) {
        const int ai = tid + k * L_T;
        if (ai >= nact) continue;
        const int q = S.act[ai];
        if (q < nB) lds_apply(S.pvx[q], S.pvy[q], S.pvz[q], dth, anyp[k], p[k], anym[k], m[k]);
        else lds_apply(S.tvx[q - nB], S.tvy[q - nB], S.tvz[q - nB], dth, anyp[k], p[k], anym[k], m[k]);
        if (mode == 1) {
            kc[k].p = p[k];
            kc[k].m = m[k];
            kc[k].fl = (anyp[k] ? 1 : 0) | (anym[k] ? 2 : 0);
        }
    }
    if (prof) {
        if (tid == 0) S.kmax = 0;
        __syncthreads();
        if (kmax) atomicMax(&S.kmax, kmax);
        __syncthreads();
        if (pr) {
            unsigned long long *P = S.prof + d * NPROF;
            P[8] += tc;
            P[9] += tf;
            P[10] += S.kmax;
            P[11] += nact;
        }
    }
    __syncthreads();
}

__device__ void lds_drift(LdsSet &S, const SubArgs &A, int nB, int nBt, int d, double dtl, int s, bool prof)
{
    const double mu = A.st.sy.GMcb[s];
    int ndr = 0;
    unsigned long long tdr = 0;
    for (int q = threadIdx.x; q < nB + nBt; q += L_T) {
        const bool ispl = q < nB;
        const int t = q - nB;
        const bool on = ispl ? (S.pst[q] != INACTIVE && S.plev[q] == d) : (S.tst[t] == ACTIVE && S.tlev[t] == d);
        if (!on) continue;
        ndr++;
        double *x = ispl ? S.px : S.tx, *y = ispl ? S.py : S.ty, *z = ispl ? S.pz : S.tz;
        double *vx = ispl ? S.pvx : S.tvx, *vy = ispl ? S.pvy : S.tvy, *vz = ispl ? S.pvz : S.tvz;
        const int k = ispl ? q : t;
        double rx = x[k], ry = y[k], rz = z[k], ux = vx[k], uy = vy[k], uz = vz[k];
        const unsigned long long t0 = prof ? wall_clock64() : 0;
        const int fl = drift_one(mu, rx, ry, rz, ux, uy, uz, dtl);
        if (prof) tdr += wall_clock64() - t0;
        x[k] = rx;
        y[k] = ry;
        z[k] = rz;
        vx[k] = ux;
        vy[k] = uy;
        vz[k] = uz;
        if (fl) {
            if (ispl) {
                S.pst[q] = DISCARDED_DRIFTERR;
                atomicOr(&A.st.sy.err[s], 2);
            } else {
                S.tst[t] = DISCARDED_DRIFTERR;
            }
        }
    }
    if (prof && ndr) {
        unsigned long long *P = S.prof + d * NPROF;
        atomicAdd(&P[12], (unsigned long long)ndr);
        if (ndr > 1) atomicAdd(&P[13], 1ull);
        atomicAdd(&P[15], tdr);
    }
    __syncthreads();
}

// positions, velocities and pair flags to HBM (before the collision test
// and at the end)
__device__ void lds_store(LdsSet &S, const SubArgs &A, int nB, int nBt, int nEt, bool all)
{
    const RecState &a = A.st;
    for (int q = threadIdx.x; q < nB; q += L_T) {
        const int g = S.pg[q];
        a.pl.rh.x[g] = S.px[q];
        a.pl.rh.y[g] = S.py[q];
        a.pl.rh.z[g] = S.pz[q];
        a.pl.vb.x[g] = S.pvx[q];
        a.pl.vb.y[g] = S.pvy[q];
        a.pl.vb.z[g] = S.pvz[q];
        if (all) {
            a.pl.levelg[g] = S.plev[q];
            a.pl.levelm[g] = S.plm[q];
            a.pl.status[g] = S.pst[q];
        }
    }
    for (int q = threadIdx.x; q < nBt; q += L_T) {
        const int g = S.tg[q];
        a.tp.rh.x[g] = S.tx[q];
        a.tp.rh.y[g] = S.ty[q];
        a.tp.rh.z[g] = S.tz[q];
        a.tp.vb.x[g] = S.tvx[q];
        a.tp.vb.y[g] = S.tvy[q];
        a.tp.vb.z[g] = S.tvz[q];
        if (all) {
            a.tp.levelg[g] = S.tlev[q];
            a.tp.levelm[g] = S.tlm[q];
            a.tp.status[g] = S.tst[q];
        }
    }
    for (int lp = threadIdx.x; lp < S.npair; lp += L_T) {
        const int k = S.ppg[lp];
        a.pp.status[k] = S.pps[lp];
        a.pp.lvdotr[k] = S.ppv[lp];
        if (all) a.pp.level[k] = S.ppl[lp];
    }
    for (int u = threadIdx.x; u < nEt; u += L_T) {
        const int k = S.tpg[u];
        a.pt.status[k] = S.tps[u];
        a.pt.lvdotr[k] = S.tpv[u];
        if (all) a.pt.level[k] = S.tpl_lev[u];
    }
}

// collision detection at level d: the test reads whole rows, including
// partners outside the slice, so it runs on HBM after a store; the pair
// status it may change is then reloaded
__device__ void lds_collision(LdsSet &S, const SubArgs &A, int nB, int nBt, int nEt, int d, double dtl, int s)
{
    lds_store(S, A, nB, nBt, nEt, false);
    __syncthreads();
    for (int q = threadIdx.x; q < nB + nBt; q += L_T) {
        if (q < nB) {
            if (A.has_pp && S.plev[q] >= d) rec_collide_pl(A.st, S.pg[q], dtl, s);
        } else if (S.tlev[q - nB] >= d) {
            rec_collide_tp(A.st, S.tg[q - nB], dtl, s);
        }
    }
    __syncthreads();
    for (int lp = threadIdx.x; lp < S.npair; lp += L_T) S.pps[lp] = A.st.pp.status[S.ppg[lp]];
    for (int u = threadIdx.x; u < nEt; u += L_T) S.tps[u] = A.st.pt.status[S.tpg[u]];
    __syncthreads();
}

__device__ void lds_set_levels(LdsSet &S, int nB, int nBt, int nEt, int d)
{
    const int tid = threadIdx.x, irecp = d + 1;
    for (int q = tid; q < nB; q += L_T)
        if (S.plev[q] == irecp) S.plev[q] = d;
    for (int q = tid; q < nBt; q += L_T)
        if (S.tlev[q] == irecp) S.tlev[q] = d;
    for (int lp = tid; lp < S.npair; lp += L_T)
        if (S.ppl[lp] == irecp) S.ppl[lp] = (int8_t)d;
    for (int u = tid; u < nEt; u += L_T)
        if (S.tpl_lev[u] == irecp) S.tpl_lev[u] = (int8_t)d;
    __syncthreads();
}

// ---- wave mode -------------------------------------------------------------
// Deep in the recursion a level holds a handful of bodies and pairs.  When a
// new level's bodies, pl pairs, tp entries and the kick entries among its
// bodies each number at most 64, the first wavefront runs that level and
// every level below it alone, one item per lane, with no workgroup barrier
// and no pass over the whole slice; the other wavefronts wait.  Which items
// can change is fixed while the recursion stays at or below the level (the
// checks raise levels above it and the level resets lower them back), so the
// sets are collected once; each phase filters them by the current level.
// The phases are those of the workgroup path with the same arithmetic in the
// same order: the pair checks, the kicks (each body folds its entries in
// entry order), the drifts and the level resets.  Collision checks are not
// run in wave mode (the workgroup path keeps them).

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Collect the sets of level dn with the whole workgroup; true when all fit.
__device__ bool wave_sets(LdsSet &S, int nB, int nBt, int nE, int nEt, int dn)
{
    const int tid = threadIdx.x;
    if (tid == 0) S.wnB = S.wnP = S.wnT = S.wnE = 0;
    __syncthreads();
    for (int q = tid; q < nB + nBt; q += L_T)
        if ((q < nB ? S.plev[q] : S.tlev[q - nB]) >= dn) {
            const int k = atomicAdd(&S.wnB, 1);
            if (k < 64) S.wB[k] = (int16_t)q;
        }
    for (int lp = tid; lp < S.npair; lp += L_T)
        if (S.ppl[lp] >= dn) {
            const int k = atomicAdd(&S.wnP, 1);
            if (k < 64) S.wP[k] = (int16_t)lp;
        }
    for (int u = tid; u < nEt; u += L_T)
        if (S.tpl_lev[u] >= dn) {
            const int k = atomicAdd(&S.wnT, 1);
            if (k < 64) S.wT[k] = (int16_t)u;
        }
    __syncthreads();
    if (S.wnB > 64 || S.wnP > 64 || S.wnT > 64) return false;
    if (tid < 64) {
        // each body's entries whose partner is in the level, in entry order
        const int k = tid;
        int e0 = 0, e1 = 0, q = 0;
        bool ispl = true;
        if (k < S.wnB) {
            q = S.wB[k];
            ispl = q < nB;
            e0 = ispl ? S.pO[q] : nE + S.tO[q - nB];
            e1 = ispl ? S.pO[q + 1] : nE + S.tO[q - nB + 1];
        }
        int c = 0;
        for (int u = e0; u < e1; u++) c += (ispl ? S.plev[S.pE[u] & 511u] : S.plev[S.tpl[u - nE]]) >= dn;
        int incl = c;
        for (int o = 1; o < 64; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (k >= o) incl += t;
        }
        const int tot = __shfl(incl, 63, 64);
        if (tot <= 64) {
            int w = incl - c;
            if (k < S.wnB) {
                S.wEs[k] = (int8_t)w;
                for (int u = e0; u < e1; u++)
                    if ((ispl ? S.plev[S.pE[u] & 511u] : S.plev[S.tpl[u - nE]]) >= dn) S.wE[w++] = (int16_t)u;
                S.wEe[k] = (int8_t)w;
            }
        }
        if (k == 0) S.wnE = tot;
    }
    __syncthreads();
    return S.wnE <= 64;
}

// the pair checks of level d (lds_check); returns lenc
__device__ bool wave_check(LdsSet &S, const SubArgs &A, int d, double dtl)
{
    const int k = threadIdx.x, irec = d + 1;
    bool lp_enc = false;
    if (k < S.wnP) {
        const int lp = S.wP[k];
        if (S.pps[lp] == ACTIVE && S.ppl[lp] == d) {
            const int i = S.ppi[lp], j = S.ppj[lp];
            bool lv;
            lp_enc = check_pp_v(S.px[i], S.py[i], S.pz[i], S.pvx[i], S.pvy[i], S.pvz[i], S.ph[i], S.pR[i], S.px[j],
                                S.py[j], S.pz[j], S.pvx[j], S.pvy[j], S.pvz[j], S.ph[j], S.pR[j], irec, dtl, lv);
            S.ppv[lp] = lv;
            if (lp_enc) {
                S.ppl[lp] = (int8_t)irec;
                S.plev[i] = irec;
                S.plev[j] = irec;
                atomicMax(&S.plm[i], irec);
                atomicMax(&S.plm[j], irec);
            }
        }
    }
    const bool fpp = __ballot(lp_enc) != 0;
    wave_sync();
    bool tp_enc = false;
    if (!(A.short_circuit && fpp) && k < S.wnT) {  // symba_step.f90:210 (.or. short-circuited)
        const int u = S.wT[k];
        if (S.tps[u] == ACTIVE && S.tpl_lev[u] == d) {
            const int i = S.tpl[u], t = S.tpt[u];
            bool lv;
            tp_enc = check_pt_v(S.px[i], S.py[i], S.pz[i], S.pvx[i], S.pvy[i], S.pvz[i], S.ph[i], S.pR[i], S.tx[t],
                                S.ty[t], S.tz[t], S.tvx[t], S.tvy[t], S.tvz[t], irec, dtl, lv);
            S.tpv[u] = lv;
            if (tp_enc) {
                S.tpl_lev[u] = (int8_t)irec;
                S.plev[i] = irec;
                atomicMax(&S.plm[i], irec);
                S.tlev[t] = irec;
                atomicMax(&S.tlm[t], irec);
            }
        }
    }
    const bool fpt = __ballot(tp_enc) != 0;
    wave_sync();
    return fpp || fpt;
}

// the +1 / -1 shell kicks of level d (lds_kicks).  mode 0: evaluate; 1:
// evaluate and keep the sums in c; 2: apply the sums kept in c.
__device__ void wave_kicks(LdsSet &S, int nB, int nE, int d, double dth, int mode, KCache &c)
{
    const int k = threadIdx.x;
    if (mode == 2) {
        if (k < S.wnB) {
            const int q = S.wB[k];
            const bool ispl = q < nB;
            if ((ispl ? S.plev[q] : S.tlev[q - nB]) >= d) {
                if (ispl) lds_apply(S.pvx[q], S.pvy[q], S.pvz[q], dth, (c.fl & 1) != 0, c.p, (c.fl & 2) != 0, c.m);
                else
                    lds_apply(S.tvx[q - nB], S.tvy[q - nB], S.tvz[q - nB], dth, (c.fl & 1) != 0, c.p, (c.fl & 2) != 0,
                              c.m);
            }
        }
        wave_sync();
        return;
    }
    const double spp = shell_sq(d + 1), spm = shell_sq(d);
    if (k < S.wnE) {
        const int u = S.wE[k];
        ShellTerm t;
        t.tp = t.tm = false;
        t.dx = t.dy = t.dz = 0.0;
        if (u < nE) {
            const uint32_t e = S.pE[u];
            const int lp = (int)(e >> 10), pr = (int)(e & 511u);
            const bool bi = (e >> 9) & 1u;
            const int q = bi ? S.ppi[lp] : S.ppj[lp];
            if (S.plev[q] >= d && S.plev[pr] >= d && S.pps[lp] == ACTIVE)
                pl_shell_term_v(bi, S.px[q], S.py[q], S.pz[q], S.ph[q], S.px[pr], S.py[pr], S.pz[pr], S.ph[pr],
                                S.pgm[pr], spp, spm, true, t);
        } else {
            const int v = u - nE, i = S.tpl[v], tt = S.tpt[v];
            if (S.tlev[tt] >= d && S.plev[i] >= d && S.tps[v] == ACTIVE)
                tp_shell_term_v(S.tx[tt], S.ty[tt], S.tz[tt], S.px[i], S.py[i], S.pz[i], S.ph[i], S.pgm[i], spp, spm,
                                true, t);
        }
        KTerm &K = S.kt[k];
        K.dx = t.dx;
        K.dy = t.dy;
        K.dz = t.dz;
        K.fp = t.tp ? t.fp : 0.0;
        K.fm = t.tm ? t.fm : 0.0;
        K.fl = (t.tp ? 1 : 0) | (t.tm ? 2 : 0);
    }
    wave_sync();
    if (k < S.wnB) {
        const int q = S.wB[k];
        const bool ispl = q < nB;
        if ((ispl ? S.plev[q] : S.tlev[q - nB]) >= d) {
            double3 p = make_double3(0.0, 0.0, 0.0), m = p;
            int fl = 0;
            fold_terms(S.kt, S.wEs[k], S.wEe[k], p, m, fl);
            if (ispl) lds_apply(S.pvx[q], S.pvy[q], S.pvz[q], dth, (fl & 1) != 0, p, (fl & 2) != 0, m);
            else lds_apply(S.tvx[q - nB], S.tvy[q - nB], S.tvz[q - nB], dth, (fl & 1) != 0, p, (fl & 2) != 0, m);
            if (mode == 1) {
                c.p = p;
                c.m = m;
                c.fl = fl;
            }
        }
    }
    wave_sync();
}

// The drifts of level d (lds_drift), called by the whole workgroup.  A
// wavefront runs a bound body's series solution (kepmd) and an unbound one's
// iteration (kepu) one after the other, so the first wavefront keeps the bodies
// that look bound and hands the others to the second, and the two drift side
// by side.  The test only places the work (each drift is drift_one either
// way): alpha > 0 of drift_dan is 2 mu / r0 > v0^2, taken here without the
// square root and division as 4 mu^2 > v0^4 r0^2.
// Returns S.wlenc, read between the two barriers (it is rewritten only after
// the second).
__device__ bool wave_drift(LdsSet &S, const SubArgs &A, int nB, int d, double dtl, int s)
{
    const int tid = threadIdx.x;
    const double mu = A.st.sy.GMcb[s];
    bool on = false, far = false;
    int q = 0;
    if (tid < 64) {
        if (tid < S.wnB) {
            q = S.wB[tid];
            const bool ispl = q < nB;
            const int t = q - nB;
            on = ispl ? (S.pst[q] != INACTIVE && S.plev[q] == d) : (S.tst[t] == ACTIVE && S.tlev[t] == d);
            if (on) {
                const double rx = ispl ? S.px[q] : S.tx[t], ry = ispl ? S.py[q] : S.ty[t], rz = ispl ? S.pz[q] : S.tz[t];
                const double vx = ispl ? S.pvx[q] : S.tvx[t], vy = ispl ? S.pvy[q] : S.tvy[t],
                             vz = ispl ? S.pvz[q] : S.tvz[t];
                const double v2 = dot3(vx, vy, vz, vx, vy, vz);
                far = !(4.0 * mu * mu > (v2 * v2) * dot3(rx, ry, rz, rx, ry, rz));
            }
        }
        const uint64_t m = __ballot(far);
        if (far) S.wH[__popcll(m & ((1ull << tid) - 1))] = (int8_t)tid;
        if (tid == 0) S.wnH = __popcll(m);
    }
    __syncthreads();
    const bool lenc = S.wlenc != 0;
    int k = -1;
    if (tid < 64) {
        if (on && !far) k = tid;
    } else if (tid < 128 && tid - 64 < S.wnH) {
        k = S.wH[tid - 64];
        q = S.wB[k];
    }
    if (k >= 0) {
        const bool ispl = q < nB;
        const int t = q - nB;
        double *x = ispl ? S.px : S.tx, *y = ispl ? S.py : S.ty, *z = ispl ? S.pz : S.tz;
        double *vx = ispl ? S.pvx : S.tvx, *vy = ispl ? S.pvy : S.tvy, *vz = ispl ? S.pvz : S.tvz;
        const int i = ispl ? q : t;
        double rx = x[i], ry = y[i], rz = z[i], ux = vx[i], uy = vy[i], uz = vz[i];
        const int fl = drift_one(mu, rx, ry, rz, ux, uy, uz, dtl);
        x[i] = rx;
        y[i] = ry;
        z[i] = rz;
        vx[i] = ux;
        vy[i] = uy;
        vz[i] = uz;
        if (fl) {
            if (ispl) {
                S.pst[q] = DISCARDED_DRIFTERR;
                atomicOr(&A.st.sy.err[s], 2);
            } else {
                S.tst[t] = DISCARDED_DRIFTERR;
            }
        }
    }
    __syncthreads();
    return lenc;
}

// the level reset after a substep of level d (lds_set_levels)
__device__ void wave_set_levels(LdsSet &S, int nB, int d)
{
    const int k = threadIdx.x, irecp = d + 1;
    if (k < S.wnB) {
        const int q = S.wB[k];
        if (q < nB) {
            if (S.plev[q] == irecp) S.plev[q] = d;
        } else if (S.tlev[q - nB] == irecp) {
            S.tlev[q - nB] = d;
        }
    }
    if (k < S.wnP && S.ppl[S.wP[k]] == irecp) S.ppl[S.wP[k]] = (int8_t)d;
    if (k < S.wnT && S.tpl_lev[S.wT[k]] == irecp) S.tpl_lev[S.wT[k]] = (int8_t)d;
    wave_sync();
}

// Levels dw and below in wave mode, from the first substep of dw to the end
// of dw (the loop of k_subtree_lds), called by the whole workgroup: the first
// wavefront runs the checks, kicks and level resets, the drifts take the
// second one as well, and the others only follow the loop.  Adds its substeps
// to nsub and raises maxd; sets S.wstop when the recursion must stop (error
// flagged).
__device__ void wave_recursion(LdsSet &S, const SubArgs &A, int nB, int nE, int dw, int s, int &nsub, int &maxd)
{
    const bool w0 = threadIdx.x < 64;
    const bool prof = A.prof && threadIdx.x == 0;
    int d = dw, j = 1;
    bool resume = false;
    KCache kc{};
    int kcd = -1;  // the level whose closing-kick sums kc holds (-1: none)
    for (;;) {
        double p3 = 1.0;
        for (int q = 0; q < d; q++) p3 *= NTENC;
        const double dtl = A.dt / p3, dth = 0.5 * dtl;
        if (!resume) {
            if (j > NTENC) {
                if (d == dw) break;
                d--;
                j = S.jst[d];
                resume = true;
                continue;
            }
            if (dtl / A.dt < 2.220446049250313e-16) {  // symba_step.f90:196-201
                if (threadIdx.x == 0) {
                    atomicOr(&A.st.sy.err[s], 8);
                    S.wstop = 1;
                }
                break;
            }
            nsub++;
            const unsigned long long t0 = prof ? wall_clock64() : 0;
            if (w0) {
                const bool le = wave_check(S, A, d, dtl);
                if (threadIdx.x == 0) S.wlenc = le;
            }
            const unsigned long long t1 = prof ? wall_clock64() : 0;
            if (w0) wave_kicks(S, nB, nE, d, dth, kcd == d ? 2 : 0, kc);
            kcd = -1;
            const unsigned long long t2 = prof ? wall_clock64() : 0;
            const bool lenc = wave_drift(S, A, nB, d, dtl, s);
            if (prof) {
                unsigned long long *P = S.prof + d * NPROF;
                P[0] += 1;
                P[1] += S.wnB;
                P[2] += S.wnE;
                P[3] += t1 - t0;
                P[4] += t2 - t1;
                P[5] += wall_clock64() - t2;
            }
            if (lenc) {
                if (d + 1 > MAXLEV) {
                    if (threadIdx.x == 0) {
                        atomicOr(&A.st.sy.err[s], 8);
                        S.wstop = 1;
                    }
                    break;
                }
                if (threadIdx.x == 0) S.jst[d] = j;
                wave_sync();
                d++;
                if (d > maxd) maxd = d;
                j = 1;
                continue;
            }
        }
        resume = false;
        const unsigned long long t0 = prof ? wall_clock64() : 0;
        if (w0) wave_kicks(S, nB, nE, d, dth, 1, kc);
        kcd = d;
        const unsigned long long t1 = prof ? wall_clock64() : 0;
        if (w0) wave_set_levels(S, nB, d);
        if (prof) {
            S.prof[d * NPROF + 6] += t1 - t0;
            S.prof[d * NPROF + 7] += wall_clock64() - t1;
        }
        j++;
    }
}

__global__ __launch_bounds__(L_T) void k_subtree_lds(const SubArgs A)
{
    const RecState &a = A.st;
    __shared__ LdsSet S;
    const int s = blockIdx.x;
    const int tid = threadIdx.x;
    const int32_t *r = A.rng + 4 * s;
    const int nB = r[1] - r[0], nBt = r[3] - r[2];
    if (nB + nBt == 0) return;
    const int32_t *gB = A.start.B + r[0], *gO = A.start.O + r[0];
    const int32_t *gBt = A.start.Bt + r[2], *gOt = A.start.Ot + r[2];
    const int e0 = gO[0], nE = gO[nB] - e0, et0 = gOt[0], nEt = gOt[nBt] - et0;
    if (!A.lds || !lds_eligible(nB, nE, nBt, nEt)) return;
    // ---- load the slice ----
    for (int q = tid; q < nB; q += L_T) {
        const int g = gB[q];
        S.pg[q] = g;
        S.px[q] = a.pl.rh.x[g];
        S.py[q] = a.pl.rh.y[g];
        S.pz[q] = a.pl.rh.z[g];
        S.pvx[q] = a.pl.vb.x[g];
        S.pvy[q] = a.pl.vb.y[g];
        S.pvz[q] = a.pl.vb.z[g];
        S.ph[q] = a.pl.rhill[g];
        S.pgm[q] = a.pl.Gm[g];
        S.pR[q] = a.pl.R[g];
        S.plev[q] = a.pl.levelg[g];
        S.plm[q] = a.pl.levelm[g];
        S.pst[q] = a.pl.status[g];
        S.pO[q] = gO[q] - e0;
    }
    for (int q = tid; q < nBt; q += L_T) {
        const int g = gBt[q];
        S.tg[q] = g;
        S.tx[q] = a.tp.rh.x[g];
        S.ty[q] = a.tp.rh.y[g];
        S.tz[q] = a.tp.rh.z[g];
        S.tvx[q] = a.tp.vb.x[g];
        S.tvy[q] = a.tp.vb.y[g];
        S.tvz[q] = a.tp.vb.z[g];
        S.tlev[q] = a.tp.levelg[g];
        S.tlm[q] = a.tp.levelm[g];
        S.tst[q] = a.tp.status[g];
        S.tO[q] = gOt[q] - et0;
    }
    if (tid == 0) {
        S.pO[nB] = nE;
        S.tO[nBt] = nEt;
        S.npair = 0;
        S.bad = 0;
        S.lst_level = -1;
    }
    if (A.prof)
        for (int k = tid; k < (MAXLEV + 2) * NPROF; k += L_T) S.prof[k] = 0;
    __syncthreads();
    // pl entries: local partners; each pair gets a local id from its index1
    // entry.  A partner outside the slice or a pair without both entries
    // cannot occur in a level set; if one did, the system reports an error
    // before touching any state.
    for (int u = tid; u < nE; u += L_T) {
        const int2 en = A.start.E[e0 + u];
        const uint32_t pr = (uint32_t)lds_lower_bound(S.pg, nB, en.y);
        if (pr >= (uint32_t)nB || S.pg[pr] != en.y) {
            S.bad = 1;
            S.pE[u] = 0;
            continue;
        }
        if (en.x & 1) {
            const int k = (int)((uint32_t)en.x >> 1);
            const int lp = atomicAdd(&S.npair, 1);
            if (lp >= L_NP) {
                S.bad = 1;
                S.pE[u] = 0;
                continue;
            }
            S.ppg[lp] = k;
            S.ppi[lp] = (int16_t)lds_owner(S.pO, nB, u);
            S.ppj[lp] = (int16_t)pr;
            S.pps[lp] = a.pp.status[k];
            S.ppl[lp] = a.pp.level[k];
            S.ppv[lp] = a.pp.lvdotr[k];
            S.pE[u] = ((uint32_t)lp << 10) | (1u << 9) | pr;
        } else {
            S.pE[u] = pr;
        }
    }
    for (int u = tid; u < nEt; u += L_T) {
        const int2 en = A.start.Et[et0 + u];
        const int pl = lds_lower_bound(S.pg, nB, en.y);
        if (pl >= nB || S.pg[pl] != en.y) {
            S.bad = 1;
            S.tpl[u] = 0;
        } else {
            S.tpl[u] = (int16_t)pl;
        }
        S.tpt[u] = (int16_t)lds_owner(S.tO, nBt, u);
        S.tpg[u] = en.x;
        S.tps[u] = a.pt.status[en.x];
        S.tpl_lev[u] = a.pt.level[en.x];
        S.tpv[u] = a.pt.lvdotr[en.x];
    }
    __syncthreads();
    // index2-side entries take the local id of the same pair from the index1
    // body's entries
    for (int u = tid; u < nE; u += L_T) {
        if (S.bad) break;
        const uint32_t e = S.pE[u];
        if (e & (1u << 9)) continue;
        const int k = (int)((uint32_t)A.start.E[e0 + u].x >> 1);
        const int p = (int)(e & 511u);
        bool found = false;
        for (int w = S.pO[p]; w < S.pO[p + 1]; w++) {
            const uint32_t f = S.pE[w];
            if ((f & (1u << 9)) && S.ppg[f >> 10] == k) {
                S.pE[u] = (f & ~1023u) | e;
                found = true;
                break;
            }
        }
        if (!found) S.bad = 1;
    }
    __syncthreads();
    if (S.bad) {
        if (tid == 0) atomicOr(&a.sy.err[s], 4);
        return;
    }
    if (A.prof) {
        // diagnostic: connected components of the slice (bodies joined by
        // their pairs), accumulated in the profile's last row
        for (int q = tid; q < nB + nBt; q += L_T) S.lab[q] = q;
        __syncthreads();
        for (int it = 0; it < nB + nBt; it++) {
            if (tid == 0) S.flag_pp = 0;
            __syncthreads();
            for (int lp = tid; lp < S.npair; lp += L_T) {
                const int i = S.ppi[lp], j = S.ppj[lp];
                const int m = min(S.lab[i], S.lab[j]);
                if (S.lab[i] > m) {
                    atomicMin(&S.lab[i], m);
                    S.flag_pp = 1;
                }
                if (S.lab[j] > m) {
                    atomicMin(&S.lab[j], m);
                    S.flag_pp = 1;
                }
            }
            for (int u = tid; u < nEt; u += L_T) {
                const int i = S.tpl[u], t = nB + S.tpt[u];
                const int m = min(S.lab[i], S.lab[t]);
                if (S.lab[i] > m) {
                    atomicMin(&S.lab[i], m);
                    S.flag_pp = 1;
                }
                if (S.lab[t] > m) {
                    atomicMin(&S.lab[t], m);
                    S.flag_pp = 1;
                }
            }
            __syncthreads();
            const bool again = S.flag_pp;
            __syncthreads();
            if (!again) break;
        }
        if (tid == 0) {
            int ncomp = 0, big = 0;
            for (int q = 0; q < nB + nBt; q++) {
                if (S.lab[q] != q) continue;
                ncomp++;
                int sz = 0;
                for (int w = q; w < nB + nBt; w++) sz += S.lab[w] == q;
                big = max(big, sz);
            }
            unsigned long long *P = A.prof + (MAXLEV + 1) * NPROF;
            atomicAdd(&P[0], 1ull);
            atomicAdd(&P[1], (unsigned long long)ncomp);
            atomicAdd(&P[2], (unsigned long long)big);
            atomicAdd(&P[3], (unsigned long long)(nB + nBt));
            atomicMax(&P[4], (unsigned long long)big);
        }
        __syncthreads();
        if (tid == 0) S.flag_pp = 0;
        __syncthreads();
    }
    // ---- the subtree ----
    const int d0 = A.d0;
    int d = d0, j = 1;
    bool resume = false;
    int nsub = 0, maxd = d0;
    KCache kc[L_KS];
    int kcd = -1;  // the level whose closing-kick sums kc holds (-1: none)
    for (;;) {
        double p3 = 1.0;
        for (int q = 0; q < d; q++) p3 *= NTENC;
        const double dtl = A.dt / p3, dth = 0.5 * dtl;
        if (!resume) {
            if (j > NTENC) {  // level finished
                if (d == d0) break;
                d--;
                j = S.jst[d];
                resume = true;
                continue;
            }
            if (dtl / A.dt < 2.220446049250313e-16) {  // symba_step.f90:196-201
                if (tid == 0) atomicOr(&a.sy.err[s], 8);
                break;
            }
            nsub++;
            const bool prof = A.prof && tid == 0;
            unsigned long long t0 = prof ? wall_clock64() : 0, t1 = 0, t2 = 0, t3 = 0;
            const bool lenc = lds_check(S, A, nB, nEt, d, dtl);
            if (prof) t1 = wall_clock64();
            lds_kicks(S, nB, nBt, nE, nEt, d, dth, A.prof != nullptr, kcd == d ? 2 : 0, kc);
            kcd = -1;
            if (prof) t2 = wall_clock64();
            lds_drift(S, A, nB, nBt, d, dtl, s, A.prof != nullptr);
            if (prof) {
                t3 = wall_clock64();
                unsigned long long *P = S.prof + d * NPROF;
                P[0] += 1;
                P[1] += nB;
                P[2] += nE;
                P[3] += t1 - t0;
                P[4] += t2 - t1;
                P[5] += t3 - t2;
            }
            if (lenc) {
                if (d + 1 > MAXLEV) {
                    if (tid == 0) atomicOr(&a.sy.err[s], 8);
                    break;
                }
                // the kick list of level d stays: a deeper level in wave mode
                // leaves it alone, one in this path rebuilds it for its own
                // level (and lst_level tells)
                if (tid == 0) {
                    S.jst[d] = j;
                    S.wstop = 0;
                }
                __syncthreads();
                d++;
                if (d > maxd) maxd = d;
                j = 1;
                if (!A.lclose && A.wave && wave_sets(S, nB, nBt, nE, nEt, d)) {
                    // level d and below on the first wavefront; then level d
                    // is finished and the workgroup resumes level d - 1
                    wave_recursion(S, A, nB, nE, d, s, nsub, maxd);
                    __syncthreads();
                    if (S.wstop) break;
                    j = NTENC + 1;
                }
                continue;
            }
        }
        resume = false;
        {
            const bool prof = A.prof && tid == 0;
            unsigned long long t0 = prof ? wall_clock64() : 0;
            lds_kicks(S, nB, nBt, nE, nEt, d, dth, A.prof != nullptr, A.lclose ? 0 : 1, kc);
            kcd = A.lclose ? -1 : d;
            unsigned long long t1 = prof ? wall_clock64() : 0;
            if (A.lclose) lds_collision(S, A, nB, nBt, nEt, d, dtl, s);
            lds_set_levels(S, nB, nBt, nEt, d);
            if (prof) {
                S.prof[d * NPROF + 6] += t1 - t0;
                S.prof[d * NPROF + 7] += wall_clock64() - t1;
            }
        }
        j++;
    }
    __syncthreads();
    lds_store(S, A, nB, nBt, nEt, true);
    if (tid == 0) {
        atomicAdd(&A.stats[0], nsub);
        atomicMax(&A.stats[1], maxd);
    }
    if (A.prof)
        for (int k = tid; k < (MAXLEV + 2) * NPROF; k += L_T)
            if (S.prof[k]) atomicAdd(&A.prof[k], S.prof[k]);
}


static bool lds_mode()
{
    const char *e = getenv("SWG_SUBTREE_LDS");
    return !(e && e[0] == '0');
}

bool subtree_lds_fits(int nb, int ne, int nbt, int net)
{
    return lds_mode() && nb <= L_NB && ne <= L_NE && nbt <= L_NBT && net <= L_NET;
}

// diagnostic profile of the launches since the last report
// (SWG_PROF_SUBTREE=1), printed to stderr
static unsigned long long *g_prof = nullptr;
void subtree_profile_report(hipStream_t st)
{
    if (!g_prof) return;
    unsigned long long h[NPROF * (MAXLEV + 2)];
    (void)hipStreamSynchronize(st);
    (void)copy_wait(st, h, g_prof, sizeof h, hipMemcpyDeviceToHost);
    const double tick_us = 0.01;  // wall_clock64 runs at 100 MHz
    fprintf(stderr, "[swg] level substeps   avg_nB   avg_nE  check_ms  kick1_ms  drift_ms kick2/bld levels_ms"
                    "   eval_ms   fold_ms kick_nmax kick_nact  drifted drift>1 lists drift1_us\n");
    for (int d = 1; d <= MAXLEV; d++) {
        const unsigned long long *p = h + d * NPROF;
        if (!p[0]) continue;
        fprintf(stderr, "[swg] %5d %8llu %8.0f %8.0f %9.2f %9.2f %9.2f %9.2f %9.2f %9.2f %9.2f %9.1f %9.1f %8.2f %7llu %5llu %9.3f\n",
                d, p[0], (double)p[1] / p[0], (double)p[2] / p[0], p[3] * tick_us * 1e-3, p[4] * tick_us * 1e-3,
                p[5] * tick_us * 1e-3, p[6] * tick_us * 1e-3, p[7] * tick_us * 1e-3, p[8] * tick_us * 1e-3,
                p[9] * tick_us * 1e-3, p[10] / (2.0 * p[0]), p[11] / (2.0 * p[0]), (double)p[12] / p[0], p[13],
                p[14], p[12] ? p[15] * tick_us / p[12] : 0.0);
    }
    const unsigned long long *c = h + (MAXLEV + 1) * NPROF;
    if (c[0])
        fprintf(stderr, "[swg] LDS slices %llu: avg bodies %.1f, avg components %.1f, avg largest %.1f, max largest %llu\n",
                c[0], (double)c[3] / c[0], (double)c[1] / c[0], (double)c[2] / c[0], c[4]);
    (void)hipMemsetAsync(g_prof, 0, sizeof h, st);
}

void launch_subtree(hipStream_t st, const LevelArgs &la, const LevelView &start, const int32_t *rng, int d0,
                    double dt, int lclose, int short_circuit, int32_t *arena, const int64_t *arena_off,
                    int32_t *stats, bool all_lds)
{
    SubArgs a{};  // passed by value in the kernel arguments
    a.st.pl = la.pl;
    a.st.tp = la.tp;
    a.st.pp = la.pp;
    a.st.pt = la.pt;
    a.st.sy = la.sy;
    a.dt = dt;
    a.lclose = lclose;
    a.short_circuit = short_circuit;
    a.has_pp = la.has_pp;
    a.has_pt = la.has_pt;
    {
        const char *e = getenv("SWG_SUBTREE_LDS");
        a.lds = !(e && e[0] == '0');
        const char *w = getenv("SWG_SUBTREE_WAVE");
        a.wave = !(w && w[0] == '0');
    }
    a.start = start;
    a.rng = rng;
    a.d0 = d0;
    a.arena = arena;
    a.arena_off = arena_off;
    a.stats = stats;
    a.prof = nullptr;
    if (getenv("SWG_PROF_SUBTREE") && getenv("SWG_PROF_SUBTREE")[0] == '1') {
        if (!g_prof && hipMalloc((void **)&g_prof, sizeof(unsigned long long) * NPROF * (MAXLEV + 2)) == hipSuccess)
            (void)hipMemsetAsync(g_prof, 0, sizeof(unsigned long long) * NPROF * (MAXLEV + 2), st);
        if (g_prof) a.prof = g_prof;
    }
    // each system runs in the kernel whose limits its set fits; all_lds: every
    // system's set fits the LDS-resident one
    if (a.lds) k_subtree_lds<<<la.nsys, L_T, 0, st>>>(a);
    if (!(a.lds && all_lds)) k_subtree<<<la.nsys, ST_T, 0, st>>>(a);
}

}  // namespace swg
