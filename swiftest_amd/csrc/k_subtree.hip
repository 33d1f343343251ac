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

struct SubArgs {
    RecState st;
    double dt;  // param%dt
    int lclose, short_circuit, has_pp, has_pt;
    LevelView start;    // whole-device level d0; system s owns a slice of it
    const int32_t *rng;  // per system: pl begin, pl end, tp begin, tp end in `start`
    int d0;
    int32_t *arena;
    const int64_t *arena_off;  // nsys+1 word offsets
    int32_t *stats;            // [0] substeps, [1] max depth, [2] overflow
    // optional per-level profile (NPROF words per level): substeps, sum nB,
    // sum nE, then wall-clock ticks in check, kicks, drift, build, levels
    unsigned long long *prof;
};
constexpr int NPROF = 8;

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
__global__ __launch_bounds__(ST_T) void k_subtree(const SubArgs *__restrict__ ap)
{
    const SubArgs &A = *ap;
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
void launch_subtree(hipStream_t st, const LevelArgs &la, const LevelView &start, const int32_t *rng, int d0,
                    double dt, int lclose, int short_circuit, int32_t *arena, const int64_t *arena_off,
                    int32_t *stats, void *h_stage, void *d_args)
{
    SubArgs &a = *(SubArgs *)h_stage;
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
    a.start = start;
    a.rng = rng;
    a.d0 = d0;
    a.arena = arena;
    a.arena_off = arena_off;
    a.stats = stats;
    a.prof = nullptr;
    if (getenv("SWG_PROF_SUBTREE") && getenv("SWG_PROF_SUBTREE")[0] == '1') {
        if (!g_prof) (void)hipMalloc((void **)&g_prof, sizeof(unsigned long long) * NPROF * (MAXLEV + 2));
        if (g_prof) {
            (void)hipMemsetAsync(g_prof, 0, sizeof(unsigned long long) * NPROF * (MAXLEV + 2), st);
            a.prof = g_prof;
        }
    }
    (void)hipMemcpyAsync(d_args, h_stage, sizeof(SubArgs), hipMemcpyHostToDevice, st);
    k_subtree<<<la.nsys, ST_T, 0, st>>>((const SubArgs *)d_args);
}

}  // namespace swg
