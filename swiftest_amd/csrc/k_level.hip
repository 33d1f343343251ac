// k_level.hip -- SyMBA recursion levels with large working sets, run on the
// whole device, all systems of an ensemble in lockstep.
//
// A level d >= 1 (symba_step_recur_system with ireci = d, symba_step.f90:
// 163-262) is one stored working set (swg_level.h) and, per substep, one
// launch each of: list check (pl-pl then pl-tp, symba_encounter_check.f90:
// 90-235), shell kicks (symba_kick.f90:126-340), level-masked drift
// (symba_drift.f90:14-61), shell kicks, collision detection (collision_check.
// f90:53-230) and the level reset (symba_step.f90:118-160).  The host reads
// one flag per substep to decide whether to descend.
//
// The next level's set is built from the current one (or from the level-0
// adjacency) by an order-preserving compaction: count, scan, fill.  Massive
// bodies carry hundreds of pairs each in crowded systems; their builds run
// one wavefront per body and their shell kicks stage the pair terms of a
// block of bodies through LDS.  Test particles carry few pairs and are
// handled one thread each.
#include <algorithm>

#include "swg_kernels.h"
#include "swg_level.h"

namespace swg {

static inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

__device__ inline int wave_sum(int v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline uint64_t lanes_below()
{
    const int lane = threadIdx.x & 63;
    return lane ? (~0ull >> (64 - lane)) : 0ull;
}

// ---- building the next level ---------------------------------------------------

// Source of a build: the level-0 adjacency (l0) or a stored level.
struct BuildSrc {
    LevelView v;
    int l0, has_pp, has_pt;
    int nB, nBt;  // source bodies
};

__device__ inline int src_pl_body(const BuildSrc &s, int q) { return s.l0 ? q : s.v.B[q]; }
__device__ inline int src_tp_body(const BuildSrc &s, int q) { return s.l0 ? q : s.v.Bt[q]; }

__global__ __launch_bounds__(256) void k_build_count_pl(RecState a, BuildSrc src, int dnew, int32_t *keep,
                                                        int32_t *cnt)
{
    const int q = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (q >= src.nB) return;
    const int g = src_pl_body(src, q);
    int c = 0;
    const bool kb = a.pl.levelg[g] >= dnew;
    if (kb) {
        if (src.l0) {
            const int n = l0_pl_count(a, src.has_pp, g);
            for (int u = lane; u < n; u += 64) c += a.pl.levelg[l0_pl_entry(a, g, u).y] >= dnew;
        } else {
            const int e0 = src.v.O[q], e1 = src.v.O[q + 1];
            for (int u = e0 + lane; u < e1; u += 64) c += a.pl.levelg[src.v.E[u].y] >= dnew;
        }
        c = wave_sum(c);
    }
    if (lane == 0) {
        keep[q] = kb;
        cnt[q] = c;
    }
}

__global__ __launch_bounds__(256) void k_build_count_tp(RecState a, BuildSrc src, int dnew, int32_t *keep,
                                                        int32_t *cnt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= src.nBt) return;
    const int g = src_tp_body(src, q);
    int c = 0;
    const bool kb = a.tp.levelg[g] >= dnew;
    if (kb) {
        if (src.l0) {
            const int n = l0_tp_count(a, src.has_pt, g);
            for (int u = 0; u < n; u++) c += a.pl.levelg[l0_tp_entry(a, g, u).y] >= dnew;
        } else {
            for (int u = src.v.Ot[q]; u < src.v.Ot[q + 1]; u++) c += a.pl.levelg[src.v.Et[u].y] >= dnew;
        }
    }
    keep[q] = kb;
    cnt[q] = c;
}

// Per system: the new level's body ranges (rng[4s .. 4s+3] = pl begin, pl
// end, tp begin, tp end), the arena words its persistent workgroup needs for
// the nested levels below (caps), and over the systems: the largest set
// (maxsz[0], bodies + entries), the largest pl bodies, pl entries, tp bodies
// and tp entries (maxsz[1..4]) and the number of systems with a set (maxsz[5]).
__device__ inline int lower_bound_i32(const int32_t *a, int n, int key)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        int m = (lo + hi) >> 1;
        if (a[m] < key) lo = m + 1;
        else hi = m;
    }
    return lo;
}

__device__ inline int64_t scan_at(const int64_t *x, int n, int i, int64_t total) { return i < n ? x[i] : total; }

__device__ inline void build_range_sys(int s, SysDev sy, const BuildSrc &src, int dnew, const int64_t *posB,
                                       const int64_t *posE, const int64_t *posBt, const int64_t *posEt,
                                       const int64_t *tot, int32_t *rng, int32_t *caps, int32_t *maxsz)
{
    int a0, a1, t0, t1;
    if (src.l0) {
        a0 = sy.pl_off[s];
        a1 = sy.pl_off[s + 1];
        t0 = src.has_pt ? sy.tp_off[s] : 0;
        t1 = src.has_pt ? sy.tp_off[s + 1] : 0;
    } else {
        a0 = lower_bound_i32(src.v.B, src.nB, sy.pl_off[s]);
        a1 = lower_bound_i32(src.v.B, src.nB, sy.pl_off[s + 1]);
        t0 = lower_bound_i32(src.v.Bt, src.nBt, sy.tp_off[s]);
        t1 = lower_bound_i32(src.v.Bt, src.nBt, sy.tp_off[s + 1]);
    }
    const int64_t b0 = scan_at(posB, src.nB, a0, tot[0]), b1 = scan_at(posB, src.nB, a1, tot[0]);
    const int64_t e = scan_at(posE, src.nB, a1, tot[1]) - scan_at(posE, src.nB, a0, tot[1]);
    const int64_t bt0 = scan_at(posBt, src.nBt, t0, tot[2]), bt1 = scan_at(posBt, src.nBt, t1, tot[2]);
    const int64_t et = scan_at(posEt, src.nBt, t1, tot[3]) - scan_at(posEt, src.nBt, t0, tot[3]);
    rng[4 * s + 0] = (int32_t)b0;
    rng[4 * s + 1] = (int32_t)b1;
    rng[4 * s + 2] = (int32_t)bt0;
    rng[4 * s + 3] = (int32_t)bt1;
    const int64_t nb = b1 - b0, nbt = bt1 - bt0;
    // each nested level is a subset of this one: (levels left) x (its words +
    // the build temporaries + alignment)
    int64_t w = 0;
    if (nb + nbt > 0) w = (int64_t)(MAXLEV - dnew) * (level_layout(nb, e, nbt, et).words + 2 * (nb + nbt) + 4) + 8;
    caps[s] = (int32_t)(w > 0x7ffffffe ? 0x7ffffffe : w);
    const int64_t sz = nb + e + nbt + et;
    auto cl = [](int64_t x) { return (int32_t)(x > 0x7fffffff ? 0x7fffffff : x); };
    atomicMax(&maxsz[0], cl(sz));
    atomicMax(&maxsz[1], cl(nb));
    atomicMax(&maxsz[2], cl(e));
    atomicMax(&maxsz[3], cl(nbt));
    atomicMax(&maxsz[4], cl(et));
    if (nb + nbt > 0) atomicAdd(&maxsz[5], 1);
}

__global__ void k_build_ranges(SysDev sy, BuildSrc src, int nsys, int dnew, const int64_t *posB,
                               const int64_t *posE, const int64_t *posBt, const int64_t *posEt,
                               const int64_t *tot, int32_t *rng, int32_t *caps, int32_t *maxsz,
                               const int32_t *any, int64_t *any_out)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s == 0 && any) *any_out = *any;  // the encounter flag travels with the counts
    if (s >= nsys) return;
    build_range_sys(s, sy, src, dnew, posB, posE, posBt, posEt, tot, rng, caps, maxsz);
}

// The same for at most RS_T systems in one workgroup, followed by the
// exclusive scan of their arena sizes into arena_off (arena_off[nsys] and
// *arena_tot: the total), which otherwise takes a launch of its own.
constexpr int RS_T = 1024;
__global__ __launch_bounds__(RS_T) void k_build_ranges_scan(SysDev sy, BuildSrc src, int nsys, int dnew,
                                                            const int64_t *posB, const int64_t *posE,
                                                            const int64_t *posBt, const int64_t *posEt,
                                                            const int64_t *tot, int32_t *rng, int32_t *caps,
                                                            int32_t *maxsz, const int32_t *any, int64_t *any_out,
                                                            int64_t *arena_off, int64_t *arena_tot)
{
    __shared__ int64_t wsum[RS_T / 64 + 1];
    const int s = threadIdx.x, lane = s & 63, w = s >> 6, nw = (blockDim.x + 63) >> 6;
    if (s == 0 && any) *any_out = *any;
    int64_t c = 0;
    if (s < nsys) {
        build_range_sys(s, sy, src, dnew, posB, posE, posBt, posEt, tot, rng, caps, maxsz);
        c = caps[s];
    }
    int64_t incl = c;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t u = __shfl_up(incl, o, 64);
        if (lane >= o) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    if (s == 0) {
        int64_t acc = 0;
        for (int k = 0; k < nw; k++) {
            const int64_t t = wsum[k];
            wsum[k] = acc;
            acc += t;
        }
        wsum[RS_T / 64] = acc;
    }
    __syncthreads();
    if (s < nsys) arena_off[s] = wsum[w] + incl - c;
    if (s == 0) {
        arena_off[nsys] = wsum[RS_T / 64];
        *arena_tot = wsum[RS_T / 64];
    }
}

// Destination of a build (pointers into a stored level)
struct BuildDst {
    int32_t *B, *O, *Bt, *Ot;
    int2 *E, *Et;
};

__global__ __launch_bounds__(256) void k_build_fill_pl(RecState a, BuildSrc src, int dnew, const int32_t *keep,
                                                       const int64_t *posB, const int64_t *posE,
                                                       const int64_t *tot, BuildDst dst)
{
    const int q = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6);
    const int lane = threadIdx.x & 63;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        dst.O[tot[0]] = (int32_t)tot[1];
        dst.Ot[tot[2]] = (int32_t)tot[3];
    }
    if (q >= src.nB || !keep[q]) return;
    const int g = src_pl_body(src, q);
    const int nb = (int)posB[q];
    int o = (int)posE[q];
    if (lane == 0) {
        dst.B[nb] = g;
        dst.O[nb] = o;
    }
    int e0, n;
    if (src.l0) {
        e0 = 0;
        n = l0_pl_count(a, src.has_pp, g);
    } else {
        e0 = src.v.O[q];
        n = src.v.O[q + 1] - e0;
    }
    for (int c0 = 0; c0 < n; c0 += 64) {
        const int u = c0 + lane;
        bool k = false;
        int2 en = make_int2(0, 0);
        if (u < n) {
            en = src.l0 ? l0_pl_entry(a, g, u) : src.v.E[e0 + u];
            k = a.pl.levelg[en.y] >= dnew;
        }
        const uint64_t m = __ballot(k);
        if (k) dst.E[o + __popcll(m & lanes_below())] = en;
        o += __popcll(m);
    }
}

__global__ __launch_bounds__(256) void k_build_fill_tp(RecState a, BuildSrc src, int dnew, const int32_t *keep,
                                                       const int64_t *posB, const int64_t *posE, BuildDst dst)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= src.nBt || !keep[q]) return;
    const int g = src_tp_body(src, q);
    const int nb = (int)posB[q];
    int o = (int)posE[q];
    dst.Bt[nb] = g;
    dst.Ot[nb] = o;
    if (src.l0) {
        const int n = l0_tp_count(a, src.has_pt, g);
        for (int u = 0; u < n; u++) {
            int2 en = l0_tp_entry(a, g, u);
            if (a.pl.levelg[en.y] >= dnew) dst.Et[o++] = en;
        }
    } else {
        for (int u = src.v.Ot[q]; u < src.v.Ot[q + 1]; u++) {
            int2 en = src.v.Et[u];
            if (a.pl.levelg[en.y] >= dnew) dst.Et[o++] = en;
        }
    }
}

// ---- one substep at level d -------------------------------------------------------

__global__ __launch_bounds__(256) void k_lv_check_pp(RecState a, const int2 *E, int nE, int d, double dtl,
                                                     int32_t *any)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nE) return;
    const int2 en = E[u];
    if (!(en.x & 1)) return;  // each pair once, from its index1 body
    const int k = (int)((uint32_t)en.x >> 1);
    if (rec_check_pp(a, k, d, dtl)) {
        a.sy.flag_pp[a.pl.sys[a.pp.i1[k]]] = 1;
        *any = 1;
    }
}

__global__ __launch_bounds__(256) void k_lv_check_pt(RecState a, const int2 *Et, int nEt, int d, double dtl,
                                                     int short_circuit, int32_t *any)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nEt) return;
    const int2 en = Et[u];
    const int s = a.pl.sys[en.y];
    if (short_circuit && a.sy.flag_pp[s]) return;  // symba_step.f90:210 (.or. short-circuited)
    if (rec_check_pt(a, en.x, d, dtl)) {
        a.sy.flag_pt[s] = 1;
        *any = 1;
    }
}

// Shell kicks of the massive bodies, TB bodies per workgroup of KT_T threads.
// The workgroup walks its bodies' entries in chunks of K = KT_C / TB entries
// per body.  Every thread evaluates KT_R terms of a chunk into LDS, with the
// loads of all of them issued before any term is computed; then lane tb of the
// first wavefront folds body tb's terms of the chunk in entry order, so each
// body's sum is the reference's serial sum.  A body's fold is serial, so the
// folds of many bodies have to run side by side: the launch picks TB so that
// the level's bodies make enough workgroups to fill the device.  At level 0
// the source is the full adjacency and only the +1 kick is made.
//
// A listed pair leaves ACTIVE only when collision detection marks it, which
// also sets bit 0 of its system's error word; the pair statuses (a random
// gather over the whole list) are read only for the bodies of systems where
// that has happened.
constexpr int KT_T = 256, KT_C = 1024, KT_R = KT_C / KT_T;

template <int TB, bool L0>
__global__ __launch_bounds__(KT_T) void k_lv_kick_pl(RecState a, LevelView v, int nB, int d, double dth, int has_pp,
                                                     double *kc)
{
    constexpr int K = KT_C / TB, LD = TB + 1;  // LDS slot of (body tb, entry k): k * LD + tb
    __shared__ double sdx[K * LD], sdy[K * LD], sdz[K * LD], sfp[K * LD], sfm[K * LD];
    __shared__ uint8_t sfl[K * LD];
    __shared__ int sb[TB], se0[TB], sn[TB], sck[TB];
    __shared__ double sxb[TB], syb[TB], szb[TB], shb[TB];
    __shared__ int smax;
    const int tid = threadIdx.x;
    const int q0 = blockIdx.x * TB;
    if (tid == 0) smax = 0;
    __syncthreads();
    if (tid < TB) {
        const int q = q0 + tid;
        int n = 0;
        if (q < nB) {
            const int b = L0 ? q : v.B[q];
            const int e0 = L0 ? 0 : v.O[q];
            n = L0 ? l0_pl_count(a, has_pp, b) : v.O[q + 1] - e0;
            sb[tid] = b;
            se0[tid] = e0;
            sck[tid] = a.sy.err[a.pl.sys[b]] & 1;
            sxb[tid] = a.pl.rh.x[b];
            syb[tid] = a.pl.rh.y[b];
            szb[tid] = a.pl.rh.z[b];
            shb[tid] = a.pl.rhill[b];
        }
        sn[tid] = n;
        atomicMax(&smax, n);
    }
    __syncthreads();
    const int nmax = smax;
    if (nmax == 0) {
        if (kc && tid < TB && q0 + tid < nB) kc[8 * (size_t)(q0 + tid) + 6] = 0.0;
        return;
    }
    const double spp = shell_sq(d + 1), spm = shell_sq(d);
    const bool minus = d > 0;
    const PlDev &pl = a.pl;
    double3 p = make_double3(0.0, 0.0, 0.0), m = p;
    bool anyp = false, anym = false;
    for (int c0 = 0; c0 < nmax; c0 += K) {
        int2 en[KT_R];
        bool live[KT_R];
#pragma unroll
        for (int r = 0; r < KT_R; r++) {
            const int sidx = tid + r * KT_T, tb = sidx / K, u = c0 + sidx % K;
            live[r] = u < sn[tb];
            en[r] = make_int2(0, 0);
            if (live[r]) en[r] = L0 ? l0_pl_entry(a, sb[tb], u) : v.E[se0[tb] + u];
        }
        int8_t stt[KT_R];
        double xp[KT_R], yp[KT_R], zp[KT_R], hp[KT_R], gp[KT_R];
#pragma unroll
        for (int r = 0; r < KT_R; r++) {
            stt[r] = INACTIVE;
            xp[r] = yp[r] = zp[r] = hp[r] = gp[r] = 0.0;
            if (live[r]) {
                const int k = (int)((uint32_t)en[r].x >> 1), q = en[r].y;
                stt[r] = sck[(tid + r * KT_T) / K] ? a.pp.status[k] : (int8_t)ACTIVE;
                xp[r] = pl.rh.x[q];
                yp[r] = pl.rh.y[q];
                zp[r] = pl.rh.z[q];
                hp[r] = pl.rhill[q];
                gp[r] = pl.Gm[q];
            }
        }
#pragma unroll
        for (int r = 0; r < KT_R; r++) {
            const int sidx = tid + r * KT_T, tb = sidx / K, k = sidx % K;
            ShellTerm t;
            t.tp = t.tm = false;
            if (live[r] && stt[r] == ACTIVE)
                pl_shell_term_v(en[r].x & 1, sxb[tb], syb[tb], szb[tb], shb[tb], xp[r], yp[r], zp[r], hp[r], gp[r],
                                spp, spm, minus, t);
            const int w = k * LD + tb;
            sfl[w] = (t.tp ? 1 : 0) | (t.tm ? 2 : 0);
            sdx[w] = t.tp || t.tm ? t.dx : 0.0;
            sdy[w] = t.tp || t.tm ? t.dy : 0.0;
            sdz[w] = t.tp || t.tm ? t.dz : 0.0;
            sfp[w] = t.tp ? t.fp : 0.0;
            sfm[w] = t.tm ? t.fm : 0.0;
        }
        __syncthreads();
        if (tid < TB) {
            // A term that does not add to a sum holds a zero factor for it,
            // and fma(0, d, s) is s bit for bit (d is finite; a sum that
            // starts at +0 is never -0), so every slot is folded without a
            // select, with the loads two slots ahead of the arithmetic.
            const int kn = min(K, sn[tid] - c0);
            int fl = 0;
            if (kn > 0) {
                int w = tid;
                double ap = sfp[w], am = sfm[w], ax = sdx[w], ay = sdy[w], az = sdz[w];
                int af = sfl[w];
                w = min(1, kn - 1) * LD + tid;
                double bp = sfp[w], bm = sfm[w], bx = sdx[w], by = sdy[w], bz = sdz[w];
                int bf = sfl[w];
                for (int k = 0; k < kn; k++) {
                    w = min(k + 2, kn - 1) * LD + tid;
                    const double cp = sfp[w], cm = sfm[w], cx = sdx[w], cy = sdy[w], cz = sdz[w];
                    const int cf = sfl[w];
                    p.x = fma(ap, ax, p.x);
                    p.y = fma(ap, ay, p.y);
                    p.z = fma(ap, az, p.z);
                    m.x = fma(am, ax, m.x);
                    m.y = fma(am, ay, m.y);
                    m.z = fma(am, az, m.z);
                    fl |= af;
                    ap = bp, am = bm, ax = bx, ay = by, az = bz, af = bf;
                    bp = cp, bm = cm, bx = cx, by = cy, bz = cz, bf = cf;
                }
            }
            anyp |= (fl & 1) != 0;
            anym |= (fl & 2) != 0;
        }
        __syncthreads();
    }
    if (tid < TB && q0 + tid < nB) {
        apply_shell_kicks(a.pl.vb, sb[tid], dth, anyp, p, anym, m);
        if (kc) {
            double *c = kc + 8 * (size_t)(q0 + tid);
            c[0] = p.x;
            c[1] = p.y;
            c[2] = p.z;
            c[3] = m.x;
            c[4] = m.y;
            c[5] = m.z;
            c[6] = (anyp ? 1.0 : 0.0) + (anym ? 2.0 : 0.0);
        }
    }
}

// The opening shell kicks of a substep from the sums kept by the closing
// kicks of the substep before it at the same level (kc, 8 words per body):
// positions, masses, Hill radii and pair statuses have not changed in
// between and no body's level has crossed d, so the reference computes the
// same sums (symba_step.f90:208-247, symba_kick.f90:126-230).
__global__ __launch_bounds__(256) void k_lv_kick_pl_kept(RecState a, LevelView v, int nB, double dth,
                                                          const double *kc)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nB) return;
    const double *c = kc + 8 * (size_t)q;
    const int fl = (int)c[6];
    if (!fl) return;
    apply_shell_kicks(a.pl.vb, v.B[q], dth, (fl & 1) != 0, make_double3(c[0], c[1], c[2]), (fl & 2) != 0,
                      make_double3(c[3], c[4], c[5]));
}

template <bool L0>
static void launch_kick_pl_tiles(hipStream_t st, const RecState &a, const LevelView &v, int nB, int d, double dth,
                                 int has_pp, double *kc)
{
    // bodies per workgroup: the largest tile that still gives two workgroups
    // per compute unit
    constexpr int WANT = 512;
    if (nB >= 64 * WANT) k_lv_kick_pl<64, L0><<<cdiv(nB, 64), KT_T, 0, st>>>(a, v, nB, d, dth, has_pp, kc);
    else if (nB >= 32 * WANT) k_lv_kick_pl<32, L0><<<cdiv(nB, 32), KT_T, 0, st>>>(a, v, nB, d, dth, has_pp, kc);
    else if (nB >= 16 * WANT) k_lv_kick_pl<16, L0><<<cdiv(nB, 16), KT_T, 0, st>>>(a, v, nB, d, dth, has_pp, kc);
    else if (nB >= 8 * WANT) k_lv_kick_pl<8, L0><<<cdiv(nB, 8), KT_T, 0, st>>>(a, v, nB, d, dth, has_pp, kc);
    else k_lv_kick_pl<4, L0><<<cdiv(nB, 4), KT_T, 0, st>>>(a, v, nB, d, dth, has_pp, kc);
}

// Shell kicks of one test particle per thread
template <bool L0>
__global__ __launch_bounds__(256) void k_lv_kick_tp(RecState a, LevelView v, int nBt, int d, double dth, int has_pt)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nBt) return;
    if (L0) {
        const int t = q, n = l0_tp_count(a, has_pt, t);
        if (n == 0) return;
        const double spp = shell_sq(d + 1);
        const double xt = a.tp.rh.x[t], yt = a.tp.rh.y[t], zt = a.tp.rh.z[t];
        double3 p = make_double3(0.0, 0.0, 0.0);
        bool anyp = false;
        for (int u = 0; u < n; u++) {
            ShellTerm s;
            tp_shell_term(a, l0_tp_entry(a, t, u), xt, yt, zt, spp, spp, false, s);
            if (s.tp) {
                anyp = true;
                p.x = fma(s.fp, s.dx, p.x);
                p.y = fma(s.fp, s.dy, p.y);
                p.z = fma(s.fp, s.dz, p.z);
            }
        }
        apply_shell_kicks(a.tp.vb, t, dth, anyp, p, false, p);
    } else {
        rec_kick_tp(a, v.Bt[q], v.Et, v.Ot[q], v.Ot[q + 1], d, dth);
    }
}

__global__ __launch_bounds__(256) void k_lv_drift(RecState a, LevelView v, int nB, int nBt, int d, double dtl)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= nB + nBt) return;
    const bool ispl = q < nB;
    const int g = ispl ? v.B[q] : v.Bt[q - nB];
    rec_drift(a, g, ispl, d, dtl, ispl ? a.pl.sys[g] : a.tp.sys[g]);
}

__global__ __launch_bounds__(256) void k_lv_collision(RecState a, LevelView v, int nB, int nBt, double dtl,
                                                      int has_pp)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < nB) {
        if (has_pp) rec_collide_pl(a, v.B[q], dtl, a.pl.sys[v.B[q]]);
    } else if (q < nB + nBt) {
        const int j = v.Bt[q - nB];
        rec_collide_tp(a, j, dtl, a.tp.sys[j]);
    }
}

// symba_step_set_recur_levels_system (symba_step.f90:118-160) at level d
// any (non-null): also clear the encounter flags (per system and any) for
// the next substep's check, which follows with nothing in between
__global__ __launch_bounds__(256) void k_lv_set_levels(RecState a, LevelView v, int nB, int nBt, int nE, int nEt,
                                                       int d, int nsys, int32_t *any)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (any && blockIdx.x == 0) {
        for (int s = threadIdx.x; s < nsys; s += blockDim.x) {
            a.sy.flag_pp[s] = 0;
            a.sy.flag_pt[s] = 0;
        }
        if (threadIdx.x == 0) *any = 0;
    }
    const int irecp = d + 1;
    if (q < nB && a.pl.levelg[v.B[q]] == irecp) a.pl.levelg[v.B[q]] = d;
    if (q < nBt && a.tp.levelg[v.Bt[q]] == irecp) a.tp.levelg[v.Bt[q]] = d;
    if (q < nE) {
        const int2 en = v.E[q];
        const int k = (int)((uint32_t)en.x >> 1);
        if ((en.x & 1) && a.pp.level[k] == irecp) a.pp.level[k] = (int8_t)d;
    }
    if (q < nEt) {
        const int k = v.Et[q].x;
        if (a.pt.level[k] == irecp) a.pt.level[k] = (int8_t)d;
    }
}

// ---- launchers ------------------------------------------------------------------

static RecState rec_state(PlDev pl, TpDev tp, ListDev pp, ListDev pt, SysDev sy)
{
    RecState r;
    r.pl = pl;
    r.tp = tp;
    r.pp = pp;
    r.pt = pt;
    r.sy = sy;
    return r;
}

LevelView level_view(void *base, const LevelCounts &c)
{
    LevelLayout L = level_layout(c.nB, c.nE, c.nBt, c.nEt);
    int32_t *w = (int32_t *)base;
    LevelView v;
    v.B = w + L.oB;
    v.O = w + L.oO;
    v.E = (const int2 *)(w + L.oE);
    v.Bt = w + L.oBt;
    v.Ot = w + L.oOt;
    v.Et = (const int2 *)(w + L.oEt);
    v.nBpl = (int)c.nB;
    v.nBtp = (int)c.nBt;
    return v;
}

int64_t level_bytes(const LevelCounts &c) { return 4 * level_layout(c.nB, c.nE, c.nBt, c.nEt).words; }

bool launch_build_count(hipStream_t st, const LevelArgs &la, const LevelView *src, int npl, int ntp, int dnew,
                        const BuildScratch &w, const int32_t *any, int64_t *arena_off)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    BuildSrc s{};
    s.l0 = src == nullptr;
    if (src) s.v = *src;
    s.has_pp = la.has_pp;
    s.has_pt = la.has_pt;
    s.nB = src ? src->nBpl : npl;
    s.nBt = src ? src->nBtp : (la.has_pt ? ntp : 0);
    if (s.nB) k_build_count_pl<<<cdiv((int64_t)s.nB * 64, 256), 256, 0, st>>>(a, s, dnew, w.keepB, w.cntE);
    if (s.nBt) k_build_count_tp<<<cdiv(s.nBt, 256), 256, 0, st>>>(a, s, dnew, w.keepBt, w.cntEt);
    if (s.nB <= SCAN_SMALL_MAX && s.nBt <= SCAN_SMALL_MAX) {
        // the deep levels: one launch for the four scans and the counters
        ScanSmall sc{};
        const int32_t *in[4] = {w.keepB, w.cntE, w.keepBt, w.cntEt};
        int64_t *out[4] = {w.posB, w.posE, w.posBt, w.posEt};
        const int n[4] = {s.nB, s.nB, s.nBt, s.nBt};
        for (int k = 0; k < 4; k++) {
            sc.in[k] = in[k];
            sc.out[k] = out[k];
            sc.n[k] = n[k];
            sc.total[k] = w.tot + k;
        }
        sc.zero = w.maxsz;
        sc.nzero = 8;
        launch_scan_small(st, sc, 4);
    } else {
        launch_scan_i32_i64(st, w.keepB, w.posB, s.nB, w.tot + 0, w.scan);
        launch_scan_i32_i64(st, w.cntE, w.posE, s.nB, w.tot + 1, w.scan);
        launch_scan_i32_i64(st, w.keepBt, w.posBt, s.nBt, w.tot + 2, w.scan);
        launch_scan_i32_i64(st, w.cntEt, w.posEt, s.nBt, w.tot + 3, w.scan);
        (void)hipMemsetAsync(w.maxsz, 0, 8 * sizeof(int32_t), st);
    }
    if (arena_off && la.nsys <= RS_T) {
        const int nt = la.nsys <= 64 ? 64 : RS_T;
        k_build_ranges_scan<<<1, nt, 0, st>>>(la.sy, s, la.nsys, dnew, w.posB, w.posE, w.posBt, w.posEt, w.tot,
                                               w.rng, w.caps, w.maxsz, any, w.tot + 5, arena_off, w.tot + 4);
        return true;
    }
    k_build_ranges<<<cdiv(la.nsys, 256), 256, 0, st>>>(la.sy, s, la.nsys, dnew, w.posB, w.posE, w.posBt, w.posEt,
                                                       w.tot, w.rng, w.caps, w.maxsz, any, w.tot + 5);
    return false;
}

void launch_build_fill(hipStream_t st, const LevelArgs &la, const LevelView *src, int npl, int ntp, int dnew,
                       const BuildScratch &w, void *dst_base, const LevelCounts &c)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    BuildSrc s{};
    s.l0 = src == nullptr;
    if (src) s.v = *src;
    s.has_pp = la.has_pp;
    s.has_pt = la.has_pt;
    s.nB = src ? src->nBpl : npl;
    s.nBt = src ? src->nBtp : (la.has_pt ? ntp : 0);
    LevelView v = level_view(dst_base, c);
    BuildDst d;
    d.B = (int32_t *)v.B;
    d.O = (int32_t *)v.O;
    d.E = (int2 *)v.E;
    d.Bt = (int32_t *)v.Bt;
    d.Ot = (int32_t *)v.Ot;
    d.Et = (int2 *)v.Et;
    k_build_fill_pl<<<cdiv((int64_t)(s.nB > 0 ? s.nB : 1) * 64, 256), 256, 0, st>>>(a, s, dnew, w.keepB, w.posB,
                                                                                   w.posE, w.tot, d);
    if (s.nBt) k_build_fill_tp<<<cdiv(s.nBt, 256), 256, 0, st>>>(a, s, dnew, w.keepBt, w.posBt, w.posEt, d);
}

void launch_level_check(hipStream_t st, const LevelArgs &la, const LevelView &v, const LevelCounts &c, int d,
                        double dtl, int short_circuit, int32_t *any)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    if (c.nE) k_lv_check_pp<<<cdiv(c.nE, 256), 256, 0, st>>>(a, v.E, (int)c.nE, d, dtl, any);
    if (c.nEt) k_lv_check_pt<<<cdiv(c.nEt, 256), 256, 0, st>>>(a, v.Et, (int)c.nEt, d, dtl, short_circuit, any);
}

void launch_level_kicks(hipStream_t st, const LevelArgs &la, const LevelView *v, int nB, int nBt, int d, double dth,
                        int keep, double *kc)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    LevelView vv{};
    if (v) vv = *v;
    if (v) {
        if (nB && keep == 2) k_lv_kick_pl_kept<<<cdiv(nB, 256), 256, 0, st>>>(a, vv, nB, dth, kc);
        else if (nB) launch_kick_pl_tiles<false>(st, a, vv, nB, d, dth, la.has_pp, keep == 1 ? kc : nullptr);
        if (nBt) k_lv_kick_tp<false><<<cdiv(nBt, 256), 256, 0, st>>>(a, vv, nBt, d, dth, la.has_pt);
    } else {  // level 0: the full adjacency of every body
        if (nB && la.has_pp) launch_kick_pl_tiles<true>(st, a, vv, nB, d, dth, la.has_pp, nullptr);
        if (nBt && la.has_pt) k_lv_kick_tp<true><<<cdiv(nBt, 256), 256, 0, st>>>(a, vv, nBt, d, dth, la.has_pt);
    }
}

void launch_level_drift(hipStream_t st, const LevelArgs &la, const LevelView &v, int nB, int nBt, int d, double dtl)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    if (nB + nBt) k_lv_drift<<<cdiv(nB + nBt, 256), 256, 0, st>>>(a, v, nB, nBt, d, dtl);
}

void launch_level_collision(hipStream_t st, const LevelArgs &la, const LevelView &v, int nB, int nBt, double dtl)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    if (nB + nBt) k_lv_collision<<<cdiv(nB + nBt, 256), 256, 0, st>>>(a, v, nB, nBt, dtl, la.has_pp);
}

void launch_level_set_levels(hipStream_t st, const LevelArgs &la, const LevelView &v, const LevelCounts &c, int d,
                             int32_t *any)
{
    RecState a = rec_state(la.pl, la.tp, la.pp, la.pt, la.sy);
    int64_t n = std::max(std::max(c.nB, c.nBt), std::max(c.nE, c.nEt));
    if (n)
        k_lv_set_levels<<<cdiv(n, 256), 256, 0, st>>>(a, v, (int)c.nB, (int)c.nBt, (int)c.nE, (int)c.nEt, d,
                                                      la.nsys, any);
    else if (any)
        launch_clear_flags(st, la.sy, la.nsys, any);
}

}  // namespace swg
