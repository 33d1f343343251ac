// k_step.hip -- kernels of the HELIO / SyMBA step outside the recursion:
// frame changes, linear drift, velocity kicks with the direct-sum
// accelerations, and the Danby drift.
//
// Reference behaviour: helio_step.f90:37-119, symba_step.f90:55-115,
// helio_kick.f90:91-169, helio_drift.f90:14-200, swiftest_util.f90:363-485,
// swiftest_kick.f90:69-415, symba_kick.f90:36-123, swiftest_drift.f90:60-138.
#include "swg_kernels.h"
#include "swg_math.h"


namespace swg {

static inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

// pl-pl and pl-tp terms formed together per lane before they are folded in
// order
constexpr int KPL_G = 4;
#ifndef SWG_KTP_G
#define SWG_KTP_G 4
#endif
constexpr int KTP_G = SWG_KTP_G;

__device__ inline bool pl_active(const PlDev &pl, int i) { return pl.status[i] != INACTIVE; }
__device__ inline bool tp_active(const TpDev &tp, int j) { return tp.status[j] == ACTIVE; }

// symba_step_reset_system (symba_step.f90:265-331): recursion levels to -1
__global__ void k_reset_levels(PlDev pl, TpDev tp, int npl, int ntp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < npl) {
        pl.levelg[i] = -1;
        pl.levelm[i] = -1;
        pl.renc[i] = (pl.rhill[i] * RHSCALE) * 1.0;  // set_renc(0), symba_util.f90:245-267
    }
    if (i < ntp) {
        tp.levelg[i] = -1;
        tp.levelm[i] = -1;
    }
}

// Per-system sums, one workgroup per system.  Each sum is a serial chain in
// the reference's order.  The whole workgroup stages a chunk of the system's
// bodies in LDS as the terms of the sums (products and quotients formed in
// parallel, each exactly as the serial expression forms it), and one lane
// per sum adds them in order (a chain streaming from HBM would wait on every
// load, one forming its own terms on its serial path).
constexpr int SUM_T = 256, SUM_CH = 1024;

struct SumStage {
    double t[4][SUM_CH];  // the terms of up to four sums
};

// The serial chains over a staged chunk, in order: acc + t[0] + ... + t[n-1]
// (SUB false) or acc - t[0] - ... - t[n-1] (SUB true), walking t[n-1] down to
// t[0] when BACK.  The reads of the next eight terms are issued before the
// adds of the current eight.  A body a masked sum skips stages +0.0
// instead: the chains start at +0.0, a sum that starts at +0 is never -0
// (x + (-x) = +0 in round-to-nearest), and s + (+0) = s, s - (+0) = s bit
// for bit for every other s, so the sums are the masked ones exactly.
template <bool SUB, bool BACK>
__device__ inline double chain(double acc, const double *t, int n)
{
    auto at = [&](int k) { return BACK ? t[n - 1 - k] : t[k]; };
    auto step = [](double s, double x) { return SUB ? s - x : s + x; };
    int k = 0;
    if (n >= 8) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; u++) v[u] = at(u);
        for (k = 8; k + 8 <= n; k += 8) {
            double w[8];
#pragma unroll
            for (int u = 0; u < 8; u++) w[u] = at(k + u);
#pragma unroll
            for (int u = 0; u < 8; u++) acc = step(acc, v[u]);
#pragma unroll
            for (int u = 0; u < 8; u++) v[u] = w[u];
        }
#pragma unroll
        for (int u = 0; u < 8; u++) acc = step(acc, v[u]);
    }
    for (; k < n; k++) acc = step(acc, at(k));
    return acc;
}

// The sums of vh2vb (swiftest_util.f90:424-459) over bodies [p0, p1): lane 0
// returns sum(Gm), lanes 1-3 the components of -sum(Gm vh), in body order.
__device__ inline double sums_gm_vh(SumStage &S, const PlDev &pl, int p0, int p1)
{
    const int tid = threadIdx.x;
    double acc = 0.0;
    for (int c0 = p0; c0 < p1; c0 += SUM_CH) {
        const int n = min(SUM_CH, p1 - c0);
        for (int k = tid; k < n; k += SUM_T) {
            const int i = c0 + k;
            const double g = pl.Gm[i];
            S.t[0][k] = g;
            S.t[1][k] = g * pl.vh.x[i];
            S.t[2][k] = g * pl.vh.y[i];
            S.t[3][k] = g * pl.vh.z[i];
        }
        __syncthreads();
        if (tid == 0)
            acc = chain<false, false>(acc, S.t[0], n);
        else if (tid <= 3)
            acc = chain<true, false>(acc, S.t[tid], n);
        __syncthreads();
    }
    return acc;
}

// Start of a step:
//   vh->vb (swiftest_util.f90:424-459) when lfirst: cb%vb = -sum(Gm vh) / (GM_cb + sum(Gm))
//   lindrift pt (helio_drift.f90:129-163): pt = sum_masked(Gm vb) / GM_cb
// Lane 0 sums Gm, lanes 1-3 the components of sum(Gm vh); then lanes 0-2 the
// components of pt.
__global__ __launch_bounds__(SUM_T) void k_sums_begin(PlDev pl, SysDev sy, int nsys, int force)
{
    __shared__ SumStage S;
    __shared__ double shr[4], scb[3];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    if (p1 == p0 || sy.frozen[s]) return;
    const bool first = force || sy.lfirst[s];
    if (first) {
        const double acc = sums_gm_vh(S, pl, p0, p1);
        if (tid < 4) shr[tid] = acc;
        __syncthreads();
        if (tid < 3) {
            const double cb = shr[tid + 1] / (sy.GMcb[s] + shr[0]);
            scb[tid] = cb;
            sy.cbvb[3 * s + tid] = cb;
        }
        __syncthreads();
    }
    const Vec3p v = first ? pl.vh : pl.vb;
    double pt = 0.0;
    for (int c0 = p0; c0 < p1; c0 += SUM_CH) {
        const int n = min(SUM_CH, p1 - c0);
        for (int k = tid; k < n; k += SUM_T) {
            const int i = c0 + k;
            const double g = pl.Gm[i];
            const bool act = pl.status[i] != INACTIVE;
            const double wx = first ? v.x[i] + scb[0] : v.x[i];
            const double wy = first ? v.y[i] + scb[1] : v.y[i];
            const double wz = first ? v.z[i] + scb[2] : v.z[i];
            S.t[0][k] = act ? g * wx : 0.0;
            S.t[1][k] = act ? g * wy : 0.0;
            S.t[2][k] = act ? g * wz : 0.0;
        }
        __syncthreads();
        if (tid < 3) pt = chain<false, false>(pt, S.t[tid], n);
        __syncthreads();
    }
    if (tid < 3) sy.ptbeg[3 * s + tid] = pt / sy.GMcb[s];
}

// vb = vh + cb%vb when lfirst; then the linear drift rh += pt*dt
__global__ void k_pl_begin_apply(PlDev pl, SysDev sy, int npl, double dth, int force)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npl) return;
    int s = pl.sys[i];
    if (sy.frozen[s]) return;
    if (force || sy.lfirst[s]) {
        pl.vb.x[i] = pl.vh.x[i] + sy.cbvb[3 * s];
        pl.vb.y[i] = pl.vh.y[i] + sy.cbvb[3 * s + 1];
        pl.vb.z[i] = pl.vh.z[i] + sy.cbvb[3 * s + 2];
    }
    if (pl_active(pl, i)) {
        pl.rh.x[i] = fma(sy.ptbeg[3 * s], dth, pl.rh.x[i]);
        pl.rh.y[i] = fma(sy.ptbeg[3 * s + 1], dth, pl.rh.y[i]);
        pl.rh.z[i] = fma(sy.ptbeg[3 * s + 2], dth, pl.rh.z[i]);
    }
}

// Massive-body half kick (helio_kick_vb_pl, helio_kick.f90:91-132):
//   ah = full-row direct sum over j ascending, j /= i, with the radius overlap
//        test (swiftest_kick.f90:219-240, symba_kick.f90:14-33);
//   ah -= (0 + ahi) + ahj over the pl-pl encounter list (symba_kick.f90:61-70,
//        swiftest_kick.f90:69-115): ahi over pairs where i is index1, ahj over
//        pairs where i is index2, each in list order;
//   rbeg/rend = rh;  vb += ah*dt.
// One lane owns body i and walks j in order, so each sum keeps the
// reference's sequence; the j bodies are staged through LDS in blocks of BLK.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_kick_pl(PlDev pl, SysDev sy, ListDev pp, const Tile *tiles, double dt,
                                                 int lbeg, int use_radius, int do_pairs)
{
    static_assert(BLK % KPL_G == 0, "the grouped row walk reads LDS in whole groups");
    __shared__ double sx[BLK], syy[BLK], sz[BLK], sgm[BLK], sr[BLK];
    const Tile tl = tiles[blockIdx.x];
    const int p0 = sy.pl_off[tl.s], p1 = sy.pl_off[tl.s + 1];
    const int i = tl.b0 + (int)threadIdx.x;
    const bool valid = i < tl.b1;
    double xi = 0.0, yi = 0.0, zi = 0.0, Ri = 0.0;
    if (valid) {
        xi = pl.rh.x[i];
        yi = pl.rh.y[i];
        zi = pl.rh.z[i];
        Ri = pl.R[i];
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int jt = p0; jt < p1; jt += BLK) {
        int j = jt + (int)threadIdx.x;
        bool gok = true;
        if (j < p1) {
            sx[threadIdx.x] = pl.rh.x[j];
            syy[threadIdx.x] = pl.rh.y[j];
            sz[threadIdx.x] = pl.rh.z[j];
            sgm[threadIdx.x] = pl.Gm[j];
            sr[threadIdx.x] = pl.R[j];
            gok = pair_g_ok(pl.Gm[j]);
        }
        const bool tile_gok = __syncthreads_and(gok);
        const int jn = min(BLK, p1 - jt);
        if (valid) {
            // Groups of KPL_G bodies: the factors of a group are formed
            // without branches (the exact cores, garbage where a term is
            // not taken or its r2 is outside the cores' range), the IEEE
            // operations redo the rare out-of-range terms under a
            // wave-uniform test, and the terms are folded in j order with
            // selects (a term not taken leaves the sum as it is).  BLK is a
            // multiple of KPL_G, so the LDS reads stay inside the tile.
            for (int q0 = 0; q0 < jn; q0 += KPL_G) {
                double rx[KPL_G], ry[KPL_G], rz[KPL_G], r2[KPL_G], fac[KPL_G];
                bool take[KPL_G], need[KPL_G];
                bool any_need = false;
#pragma unroll
                for (int u = 0; u < KPL_G; u++) {
                    const int q = q0 + u;
                    rx[u] = sx[q] - xi;
                    ry[u] = syy[q] - yi;
                    rz[u] = sz[q] - zi;
                    r2[u] = dot3(rx[u], ry[u], rz[u], rx[u], ry[u], rz[u]);
                    const double rs = Ri + sr[q];
                    // & and | (not && and ||): no short-circuit, no branches
                    take[u] = (q < jn) & (jt + q != i) & (!use_radius | (r2[u] > rs * rs));
                    fac[u] = div_core(sgm[q], r2[u] * sqrt_core(r2[u]));
                    need[u] = take[u] & !(tile_gok & (r2[u] >= PAIR_R2_LO) & (r2[u] <= PAIR_R2_HI));
                    any_need = any_need | need[u];
                }
                if (__ballot(any_need)) {
#pragma unroll
                    for (int u = 0; u < KPL_G; u++)
                        if (need[u]) fac[u] = sgm[q0 + u] / (r2[u] * sqrt(r2[u]));
                }
#pragma unroll
                for (int u = 0; u < KPL_G; u++) {
                    ax = take[u] ? fma(fac[u], rx[u], ax) : ax;
                    ay = take[u] ? fma(fac[u], ry[u], ay) : ay;
                    az = take[u] ? fma(fac[u], rz[u], az) : az;
                }
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    if (do_pairs) {
        double hix = 0.0, hiy = 0.0, hiz = 0.0, hjx = 0.0, hjy = 0.0, hjz = 0.0;
        // Both walks gather KPL_G entries at a time (indices past the end
        // repeat the walk's current entry, whose term is then not taken) so
        // that the gathers overlap, and fold the terms in list order.
        const int64_t c0 = pp.colstart[i], c1 = c0 + pp.colcnt[i];
        for (int64_t c = c0; c < c1; c += KPL_G) {  // pairs (p, i), p < i
            double rx[KPL_G], ry[KPL_G], rz[KPL_G], fac[KPL_G];
            bool take[KPL_G];
#pragma unroll
            for (int u = 0; u < KPL_G; u++) {
                const int64_t cu = c + u < c1 ? c + u : c;
                const int p = pp.i1[pp.col[cu]];
                rx[u] = xi - pl.rh.x[p];
                ry[u] = yi - pl.rh.y[p];
                rz[u] = zi - pl.rh.z[p];
                const double rji2 = dot3(rx[u], ry[u], rz[u], rx[u], ry[u], rz[u]);
                const double rs = pl.R[p] + Ri;
                take[u] = (c + u < c1) & (rji2 > rs * rs);
                const double irij3 = 1.0 / (rji2 * sqrt(rji2));
                fac[u] = pl.Gm[p] * irij3;
            }
#pragma unroll
            for (int u = 0; u < KPL_G; u++) {
                hjx = take[u] ? fma(-fac[u], rx[u], hjx) : hjx;
                hjy = take[u] ? fma(-fac[u], ry[u], hjy) : hjy;
                hjz = take[u] ? fma(-fac[u], rz[u], hjz) : hjz;
            }
        }
        const int64_t r0 = pp.rowstart[i], r1 = r0 + pp.rowcnt[i];
        for (int64_t k = r0; k < r1; k += KPL_G) {  // pairs (i, j), j > i
            double rx[KPL_G], ry[KPL_G], rz[KPL_G], fac[KPL_G];
            bool take[KPL_G];
#pragma unroll
            for (int u = 0; u < KPL_G; u++) {
                const int64_t ku = k + u < r1 ? k + u : k;
                const int j = pp.i2[ku];
                rx[u] = pl.rh.x[j] - xi;
                ry[u] = pl.rh.y[j] - yi;
                rz[u] = pl.rh.z[j] - zi;
                const double rji2 = dot3(rx[u], ry[u], rz[u], rx[u], ry[u], rz[u]);
                const double rs = Ri + pl.R[j];
                take[u] = (k + u < r1) & (rji2 > rs * rs);
                const double irij3 = 1.0 / (rji2 * sqrt(rji2));
                fac[u] = pl.Gm[j] * irij3;
            }
#pragma unroll
            for (int u = 0; u < KPL_G; u++) {
                hix = take[u] ? fma(fac[u], rx[u], hix) : hix;
                hiy = take[u] ? fma(fac[u], ry[u], hiy) : hiy;
                hiz = take[u] ? fma(fac[u], rz[u], hiz) : hiz;
            }
        }
        ax = ax - ((0.0 + hix) + hjx);
        ay = ay - ((0.0 + hiy) + hjy);
        az = az - ((0.0 + hiz) + hjz);
    }
    Vec3p rb = lbeg ? pl.rbeg : pl.rend;
    rb.x[i] = xi;
    rb.y[i] = yi;
    rb.z[i] = zi;
    if (pl_active(pl, i)) {
        pl.vb.x[i] = fma(ax, dt, pl.vb.x[i]);
        pl.vb.y[i] = fma(ay, dt, pl.vb.y[i]);
        pl.vb.z[i] = fma(az, dt, pl.vb.z[i]);
    }
}

// ---------------------------------------------------------------------------
// The same massive-body half kick with the row's terms evaluated by KE
// wavefronts and folded in j order by one more: a tile of 64 bodies per
// workgroup, the row walked in chunks of KPLS_C partners.  Evaluation
// wavefront e forms the factors of its KPLS_C / KE partners of a chunk for
// all 64 bodies (the exact cores, grouped and branch-free as in k_kick_pl)
// into an LDS matrix, a zero where a term is not taken; the folding
// wavefront, one lane per body, then adds the chunk's terms in j order,
// forming each separation again from the same operands.  fma(0, d, s) = s
// bit for bit (d finite; a sum that starts at +0 is never -0), so every sum
// is k_kick_pl's bit for bit, while a 32768-body system gets KE + 1
// wavefronts per 64 bodies instead of one.  The matrix is double-buffered:
// the fold of chunk c overlaps the evaluation of chunk c + 1.  The two
// encounter-pair walks run on evaluation wavefronts 0 and 1 side by side.
// Partners are read at wave-uniform indices through the scalar cache.
// ---------------------------------------------------------------------------

typedef const __attribute__((address_space(4))) double cdouble;
__device__ inline const cdouble *cview(const double *p) { return (const cdouble *)p; }

constexpr int KPLS_C = 64;  // partners per chunk

template <int KE>
__global__ __launch_bounds__(64 * (KE + 3)) void k_kick_pl_split(PlDev pl, SysDev sy, ListDev pp, const Tile *tiles,
                                                                  double dt, int lbeg, int use_radius, int do_pairs)
{
    static_assert(KPLS_C % (4 * KE) == 0, "each evaluation wavefront takes whole groups of four partners");
    static_assert(KPLS_C == 64, "one partner per lane when a chunk is staged");
    __shared__ double facb[2][KPLS_C][64];
    __shared__ double hpart[2][3][64];
    __shared__ double fsum[3][64];  // each component's fold
    // the partners of chunk c (x, y, z, Gm, R) in slot c % 3: the chunk being
    // folded, the one being evaluated and the next one
    __shared__ double pj[3][5][KPLS_C];
    const Tile tl = tiles[blockIdx.x];
    const int p0 = sy.pl_off[tl.s], p1 = sy.pl_off[tl.s + 1];
    const int lane = (int)(threadIdx.x & 63u);
    // wavefronts 0..KE-1 evaluate; KE, KE+1, KE+2 fold the x, y, z sums
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int i = tl.b0 + lane;
    const bool valid = i < tl.b1;
    double xi = 0.0, yi = 0.0, zi = 0.0, Ri = 0.0;
    if (valid) {
        xi = pl.rh.x[i];
        yi = pl.rh.y[i];
        zi = pl.rh.z[i];
        Ri = pl.R[i];
    }
    const int n = p1 - p0, nch = (n + KPLS_C - 1) / KPLS_C;
    constexpr int PER = KPLS_C / KE;
    // staging of chunk c by the first wavefront, one partner per lane
    auto stage = [&](int c) {
        const int jt = p0 + c * KPLS_C, jn = min(KPLS_C, p1 - jt);
        if (lane < jn) {
            const int j = jt + lane;
            double(*d)[KPLS_C] = pj[c % 3];
            d[0][lane] = pl.rh.x[j];
            d[1][lane] = pl.rh.y[j];
            d[2][lane] = pl.rh.z[j];
            d[3][lane] = pl.Gm[j];
            d[4][lane] = pl.R[j];
        }
    };
    // evaluation of chunk c into buffer c & 1 (wavefronts 0 .. KE-1)
    auto evaluate = [&](int c) {
        const int jt = p0 + c * KPLS_C, jn = min(KPLS_C, p1 - jt);
        const double(*d)[KPLS_C] = pj[c % 3];
        // every mass of the chunk inside the exact cores' range: one column per lane
        const bool gok_lane = lane >= jn || pair_g_ok(d[3][lane]);
        const bool tile_gok = __all(gok_lane);
        double(*fb)[64] = facb[c & 1];
        for (int q0 = w * PER; q0 < (w + 1) * PER; q0 += 4) {
            double r2[4], fac[4], rx[4], ry[4], rz[4];
            bool take[4], need[4];
            bool any_need = false;
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const int q = q0 + u;
                const int qq = min(q, jn - 1);  // in range; a column past the end is not taken
                rx[u] = d[0][qq] - xi;
                ry[u] = d[1][qq] - yi;
                rz[u] = d[2][qq] - zi;
                r2[u] = dot3(rx[u], ry[u], rz[u], rx[u], ry[u], rz[u]);
                const double rs = Ri + d[4][qq];
                take[u] = (q < jn) & (jt + q != i) & (!use_radius | (r2[u] > rs * rs));
                fac[u] = div_core(d[3][qq], r2[u] * sqrt_core(r2[u]));
                need[u] = take[u] & !(tile_gok & (r2[u] >= PAIR_R2_LO) & (r2[u] <= PAIR_R2_HI));
                any_need = any_need | need[u];
            }
            if (__ballot(any_need)) {
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (need[u]) fac[u] = d[3][min(q0 + u, jn - 1)] / (r2[u] * sqrt(r2[u]));
            }
#pragma unroll
            for (int u = 0; u < 4; u++) fb[q0 + u][lane] = take[u] ? fac[u] : 0.0;
        }
    };
    double acc = 0.0;  // a fold wavefront's component sum
    const int comp = w - KE;
    const double ci = comp == 0 ? xi : (comp == 1 ? yi : zi);
    if (w == 0) {
        if (nch > 0) stage(0);
        if (nch > 1) stage(1);
    }
    __syncthreads();
    if (w < KE && nch > 0) evaluate(0);
    __syncthreads();
    for (int c = 0; c < nch; c++) {
        if (w < KE) {
            if (c + 1 < nch) evaluate(c + 1);
            if (w == 0 && c + 2 < nch) stage(c + 2);
        } else {
            // fold chunk c in j order, one component per wavefront
            const int jt = p0 + c * KPLS_C, jn = min(KPLS_C, p1 - jt);
            const double(*fb)[64] = facb[c & 1];
            const double *dc = pj[c % 3][comp];
            for (int q = 0; q < jn; q++) acc = fma(fb[q][lane], dc[q] - ci, acc);
        }
        __syncthreads();
    }
    if (w >= KE) fsum[comp][lane] = acc;
    if (do_pairs && w < 2 && w < KE && valid) {
        // w = 0: ahj over pairs (p, i), p < i; w = 1: ahi over pairs (i, j)
        double hx = 0.0, hy = 0.0, hz = 0.0;
        if (w == 0) {
            const int64_t c0 = pp.colstart[i], c1 = c0 + pp.colcnt[i];
            for (int64_t c = c0; c < c1; c++) {
                const int p = pp.i1[pp.col[c]];
                const double rx = xi - pl.rh.x[p], ry = yi - pl.rh.y[p], rz = zi - pl.rh.z[p];
                const double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                const double rs = pl.R[p] + Ri;
                if (rji2 > rs * rs) {
                    const double fac = pl.Gm[p] * (1.0 / (rji2 * sqrt(rji2)));
                    hx = fma(-fac, rx, hx);
                    hy = fma(-fac, ry, hy);
                    hz = fma(-fac, rz, hz);
                }
            }
        } else {
            const int64_t r0 = pp.rowstart[i], r1 = r0 + pp.rowcnt[i];
            for (int64_t k = r0; k < r1; k++) {
                const int q = pp.i2[k];
                const double rx = pl.rh.x[q] - xi, ry = pl.rh.y[q] - yi, rz = pl.rh.z[q] - zi;
                const double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                const double rs = Ri + pl.R[q];
                if (rji2 > rs * rs) {
                    const double fac = pl.Gm[q] * (1.0 / (rji2 * sqrt(rji2)));
                    hx = fma(fac, rx, hx);
                    hy = fma(fac, ry, hy);
                    hz = fma(fac, rz, hz);
                }
            }
        }
        hpart[w][0][lane] = hx;
        hpart[w][1][lane] = hy;
        hpart[w][2][lane] = hz;
    }
    __syncthreads();
    if (w != KE || !valid) return;
    double ax = fsum[0][lane], ay = fsum[1][lane], az = fsum[2][lane];
    if (do_pairs) {
        const double hjx = hpart[0][0][lane], hjy = hpart[0][1][lane], hjz = hpart[0][2][lane];
        const double hix = KE > 1 ? hpart[1][0][lane] : 0.0, hiy = KE > 1 ? hpart[1][1][lane] : 0.0,
                     hiz = KE > 1 ? hpart[1][2][lane] : 0.0;
        ax = ax - ((0.0 + hix) + hjx);
        ay = ay - ((0.0 + hiy) + hjy);
        az = az - ((0.0 + hiz) + hjz);
    }
    Vec3p rb = lbeg ? pl.rbeg : pl.rend;
    rb.x[i] = xi;
    rb.y[i] = yi;
    rb.z[i] = zi;
    if (pl_active(pl, i)) {
        pl.vb.x[i] = fma(ax, dt, pl.vb.x[i]);
        pl.vb.y[i] = fma(ay, dt, pl.vb.y[i]);
        pl.vb.z[i] = fma(az, dt, pl.vb.z[i]);
    }
}

// ---------------------------------------------------------------------------
// The massive-body half kick of an ensemble with semi-interacting bodies
// (GMTINY): in system s the first nplm bodies interact with every body, the
// others only with those nplm (swiftest_kick.f90:165-271 via
// symba_kick.f90:14-33).  One lane per body; partners at wave-uniform
// indices through the scalar cache.
//  - nplt <= nplm: full rows j /= i for the fully interacting bodies
//    (:219-240), rows over j < nplm for the others (:243-262), with the
//    factor Gm_j / (r2 sqrt r2);
//  - nplt > nplm: the triangular loop over pairs i < nplm, i < j (:190-216)
//    accumulating ahi(i) += (Gm_j irij3) r and ahj(j) -= (Gm_i irij3) r with
//    irij3 = 1 / (r2 sqrt r2) (swiftest_kick.f90:418-446), each body's ahi
//    over j ascending and ahj over i ascending (the one-thread order of the
//    OpenMP reduction), acc = (0 + ahi) + ahj.
// The encounter-pair removal and the kick follow as in k_kick_pl.  fast: the
// factors from gm_over_r3_fast.
// ---------------------------------------------------------------------------
__device__ inline double irij3_of(double r2, int fast)
{
    return fast ? gm_over_r3_fast(1.0, r2) : 1.0 / (r2 * sqrt(r2));
}

__global__ __launch_bounds__(256) void k_kick_pl_tiny(PlDev pl, SysDev sy, ListDev pp, const Tile *tiles, double dt,
                                                      int lbeg, int use_radius, int do_pairs, int fast)
{
    const Tile tl = tiles[blockIdx.x];
    const int p0 = sy.pl_off[tl.s], p1 = sy.pl_off[tl.s + 1];
    const int nm = sy.nplm[tl.s], n = p1 - p0;
    const int i = tl.b0 + (int)threadIdx.x;
    const bool valid = i < tl.b1;
    if (!valid) return;
    const double xi = pl.rh.x[i], yi = pl.rh.y[i], zi = pl.rh.z[i], Ri = pl.R[i];
    const cdouble *X = cview(pl.rh.x), *Y = cview(pl.rh.y), *Z = cview(pl.rh.z), *G = cview(pl.Gm),
                  *RR = cview(pl.R);
    const int il = i - p0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (n - nm > nm) {
        double hix = 0.0, hiy = 0.0, hiz = 0.0, hjx = 0.0, hjy = 0.0, hjz = 0.0;
        // pairs (i, j), j > i, when i is fully interacting
        const int jb = il < nm ? i + 1 : p1;
        for (int j = jb; j < p1; j++) {
            const double rx = X[j] - xi, ry = Y[j] - yi, rz = Z[j] - zi;
            const double r2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = Ri + RR[j];
            if (use_radius && !(r2 > rs * rs)) continue;
            const double facj = G[j] * irij3_of(r2, fast);
            hix = fma(facj, rx, hix);
            hiy = fma(facj, ry, hiy);
            hiz = fma(facj, rz, hiz);
        }
        // pairs (p, i), p < i, p fully interacting
        const int pe = p0 + min(il, nm);
        for (int q = p0; q < pe; q++) {
            const double rx = xi - X[q], ry = yi - Y[q], rz = zi - Z[q];
            const double r2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = RR[q] + Ri;
            if (use_radius && !(r2 > rs * rs)) continue;
            const double faci = G[q] * irij3_of(r2, fast);
            hjx = fma(-faci, rx, hjx);
            hjy = fma(-faci, ry, hjy);
            hjz = fma(-faci, rz, hjz);
        }
        ax = (0.0 + hix) + hjx;
        ay = (0.0 + hiy) + hjy;
        az = (0.0 + hiz) + hjz;
    } else {
        const int je = il < nm ? p1 : p0 + nm;
        for (int j = p0; j < je; j++) {
            if (j == i) continue;
            const double rx = X[j] - xi, ry = Y[j] - yi, rz = Z[j] - zi;
            const double r2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = Ri + RR[j];
            if (use_radius && !(r2 > rs * rs)) continue;
            const double g = G[j];
            const double fac = fast ? gm_over_r3_fast(g, r2) : gm_over_r3(g, r2, pair_g_ok(g));
            ax = fma(fac, rx, ax);
            ay = fma(fac, ry, ay);
            az = fma(fac, rz, az);
        }
    }
    if (do_pairs) {
        double hix = 0.0, hiy = 0.0, hiz = 0.0, hjx = 0.0, hjy = 0.0, hjz = 0.0;
        const int64_t c0 = pp.colstart[i], c1 = c0 + pp.colcnt[i];
        for (int64_t c = c0; c < c1; c++) {  // pairs (p, i), p < i
            const int p = pp.i1[pp.col[c]];
            const double rx = xi - pl.rh.x[p], ry = yi - pl.rh.y[p], rz = zi - pl.rh.z[p];
            const double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = pl.R[p] + Ri;
            if (!(rji2 > rs * rs)) continue;
            const double fac = fast ? gm_over_r3_fast(pl.Gm[p], rji2) : pl.Gm[p] * (1.0 / (rji2 * sqrt(rji2)));
            hjx = fma(-fac, rx, hjx);
            hjy = fma(-fac, ry, hjy);
            hjz = fma(-fac, rz, hjz);
        }
        const int64_t r0 = pp.rowstart[i], r1 = r0 + pp.rowcnt[i];
        for (int64_t k = r0; k < r1; k++) {  // pairs (i, j), j > i
            const int j = pp.i2[k];
            const double rx = pl.rh.x[j] - xi, ry = pl.rh.y[j] - yi, rz = pl.rh.z[j] - zi;
            const double rji2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = Ri + pl.R[j];
            if (!(rji2 > rs * rs)) continue;
            const double fac = fast ? gm_over_r3_fast(pl.Gm[j], rji2) : pl.Gm[j] * (1.0 / (rji2 * sqrt(rji2)));
            hix = fma(fac, rx, hix);
            hiy = fma(fac, ry, hiy);
            hiz = fma(fac, rz, hiz);
        }
        ax = ax - ((0.0 + hix) + hjx);
        ay = ay - ((0.0 + hiy) + hjy);
        az = az - ((0.0 + hiz) + hjz);
    }
    Vec3p rb = lbeg ? pl.rbeg : pl.rend;
    rb.x[i] = xi;
    rb.y[i] = yi;
    rb.z[i] = zi;
    if (pl_active(pl, i)) {
        pl.vb.x[i] = fma(ax, dt, pl.vb.x[i]);
        pl.vb.y[i] = fma(ay, dt, pl.vb.y[i]);
        pl.vb.z[i] = fma(az, dt, pl.vb.z[i]);
    }
}

// ---------------------------------------------------------------------------
// Fast pair mode (SWG_PAIR_FAST): the same sums in the same j order, with the
// pair factor from the hardware reciprocal square root (gm_over_r3_fast,
// within a few ulp of the IEEE factor).  The partners' positions, masses and
// radii are read at wave-uniform indices through a constant-address view:
// the loads go through the scalar data cache into SGPRs that every lane of
// the wavefront reads, so no LDS staging, no workgroup barrier in the loop,
// and every VALU slot goes to the pair arithmetic (13 operations, one of
// them quarter-rate, and 7 of the factor).
// ---------------------------------------------------------------------------

// Massive bodies: one tile of KPL_FAST_TILE bodies of one system per
// workgroup of K wavefronts.  Wavefront w sums body i's row over its share of
// the partners, j in [p0 + w L, p0 + (w+1) L) with L a multiple of 4, and
// takes every K-th entry of the body's encounter-pair walks; the K partial
// accelerations are added in wavefront order through LDS.  K wavefronts per
// 64 bodies give a 32768-body system 4096 wavefronts (4 per SIMD at K = 8)
// where one lane per body gives 512.
template <int K>
__global__ __launch_bounds__(64 * K) void k_kick_pl_fast(PlDev pl, SysDev sy, ListDev pp, const Tile *tiles,
                                                         double dt, int lbeg, int use_radius, int do_pairs)
{
    static_assert(KPL_FAST_TILE == 64, "one lane per body in each wavefront");
    __shared__ double part[K][3][64];
    const Tile tl = tiles[blockIdx.x];
    const int p0 = sy.pl_off[tl.s], p1 = sy.pl_off[tl.s + 1];
    const int lane = (int)(threadIdx.x & 63u);
    const int w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int i = tl.b0 + lane;
    const bool valid = i < tl.b1;
    double xi = 0.0, yi = 0.0, zi = 0.0, Ri = 0.0;
    if (valid) {
        xi = pl.rh.x[i];
        yi = pl.rh.y[i];
        zi = pl.rh.z[i];
        Ri = pl.R[i];
    }
    const cdouble *X = cview(pl.rh.x), *Y = cview(pl.rh.y), *Z = cview(pl.rh.z), *G = cview(pl.Gm),
                  *RR = cview(pl.R);
    const FastConst fc;
    const int n = p1 - p0;
    const int L = ((n + K - 1) / K + 3) & ~3;
    const int j0 = min(p1, p0 + w * L), j1 = min(p1, j0 + L);
    double ax = 0.0, ay = 0.0, az = 0.0;
    int j = j0;
    for (; j + 4 <= j1; j += 4) {
        // the four partners' data first, unconditionally, so that each
        // array's four values come in one scalar load
        double xs[4], ys[4], zs[4], gs[4], rr[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            xs[u] = X[j + u];
            ys[u] = Y[j + u];
            zs[u] = Z[j + u];
            gs[u] = G[j + u];
            rr[u] = RR[j + u];
        }
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const double rx = xs[u] - xi, ry = ys[u] - yi, rz = zs[u] - zi;
            const double r2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = Ri + rr[u];
            // a self pair has r2 = 0, so the radius test leaves it out too
            const bool take = use_radius ? (r2 > rs * rs) : (j + u != i);
            const double f = take ? gm_over_r3_fast(gs[u], r2, fc.c15, fc.c1875) : 0.0;
            ax = fma(f, rx, ax);
            ay = fma(f, ry, ay);
            az = fma(f, rz, az);
        }
    }
    for (; j < j1; j++) {
        const double rx = X[j] - xi, ry = Y[j] - yi, rz = Z[j] - zi;
        const double r2 = dot3(rx, ry, rz, rx, ry, rz);
        const double rs = Ri + RR[j];
        const bool take = use_radius ? (r2 > rs * rs) : (j != i);
        const double f = take ? gm_over_r3_fast(G[j], r2, fc.c15, fc.c1875) : 0.0;
        ax = fma(f, rx, ax);
        ay = fma(f, ry, ay);
        az = fma(f, rz, az);
    }
    if (valid && do_pairs) {
        // this wavefront's share of the encounter pairs removed again
        // (symba_kick.f90:61-70)
        double hx = 0.0, hy = 0.0, hz = 0.0;
        const int64_t c0 = pp.colstart[i], c1 = c0 + pp.colcnt[i];
        for (int64_t c = c0 + w; c < c1; c += K) {  // pairs (p, i), p < i
            const int p = pp.i1[pp.col[c]];
            const double rx = xi - pl.rh.x[p], ry = yi - pl.rh.y[p], rz = zi - pl.rh.z[p];
            const double r2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = pl.R[p] + Ri;
            if (r2 > rs * rs) {
                const double f = gm_over_r3_fast(pl.Gm[p], r2);
                hx = fma(-f, rx, hx);
                hy = fma(-f, ry, hy);
                hz = fma(-f, rz, hz);
            }
        }
        const int64_t r0 = pp.rowstart[i], r1 = r0 + pp.rowcnt[i];
        for (int64_t k = r0 + w; k < r1; k += K) {  // pairs (i, j), j > i
            const int q = pp.i2[k];
            const double rx = pl.rh.x[q] - xi, ry = pl.rh.y[q] - yi, rz = pl.rh.z[q] - zi;
            const double r2 = dot3(rx, ry, rz, rx, ry, rz);
            const double rs = Ri + pl.R[q];
            if (r2 > rs * rs) {
                const double f = gm_over_r3_fast(pl.Gm[q], r2);
                hx = fma(f, rx, hx);
                hy = fma(f, ry, hy);
                hz = fma(f, rz, hz);
            }
        }
        ax = ax - hx;
        ay = ay - hy;
        az = az - hz;
    }
    part[w][0][lane] = ax;
    part[w][1][lane] = ay;
    part[w][2][lane] = az;
    __syncthreads();
    if (w != 0 || !valid) return;
    ax = part[0][0][lane];
    ay = part[0][1][lane];
    az = part[0][2][lane];
#pragma unroll
    for (int q = 1; q < K; q++) {
        ax = ax + part[q][0][lane];
        ay = ay + part[q][1][lane];
        az = az + part[q][2][lane];
    }
    Vec3p rb = lbeg ? pl.rbeg : pl.rend;
    rb.x[i] = xi;
    rb.y[i] = yi;
    rb.z[i] = zi;
    if (pl_active(pl, i)) {
        pl.vb.x[i] = fma(ax, dt, pl.vb.x[i]);
        pl.vb.y[i] = fma(ay, dt, pl.vb.y[i]);
        pl.vb.z[i] = fma(az, dt, pl.vb.z[i]);
    }
}

// Test particles: one lane per particle, the massive bodies at wave-uniform
// indices, four per iteration.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_kick_tp_fast(TpDev tp, PlDev pl, SysDev sy, ListDev pt, const Tile *tiles,
                                                      double dt, int lbeg, int do_pairs)
{
    const Tile tl = tiles[blockIdx.x];
    const int p0 = sy.pl_off[tl.s], p1 = sy.pl_off[tl.s + 1];
    const int j = tl.b0 + (int)threadIdx.x;
    const bool valid = (j < tl.b1) && tp_active(tp, j);
    const Vec3p rp = lbeg ? pl.rbeg : pl.rend;
    double xt = 0.0, yt = 0.0, zt = 0.0;
    if (valid) {
        xt = tp.rh.x[j];
        yt = tp.rh.y[j];
        zt = tp.rh.z[j];
    }
    const cdouble *X = cview(rp.x), *Y = cview(rp.y), *Z = cview(rp.z), *G = cview(pl.Gm);
    const FastConst fc;
    double px = 0.0, py = 0.0, pz = 0.0;
    int q = p0;
    if (q + 4 <= p1) {
        // the next four partners are loaded while the current four are
        // applied, so the scalar loads' latency hides behind the arithmetic;
        // the last batch reloads itself (index clamped to p1 - 4 >= q)
        double cx[4], cy[4], cz[4], cg[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            cx[u] = X[q + u];
            cy[u] = Y[q + u];
            cz[u] = Z[q + u];
            cg[u] = G[q + u];
        }
        for (; q + 4 <= p1; q += 4) {
            const int qn = min(q + 4, p1 - 4);
            double nx[4], ny[4], nz[4], ng[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                nx[u] = X[qn + u];
                ny[u] = Y[qn + u];
                nz[u] = Z[qn + u];
                ng[u] = G[qn + u];
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double rx = xt - cx[u], ry = yt - cy[u], rz = zt - cz[u];
                const double f = gm_over_r3_fast(cg[u], dot3(rx, ry, rz, rx, ry, rz), fc.c15, fc.c1875);
                px = fma(-f, rx, px);
                py = fma(-f, ry, py);
                pz = fma(-f, rz, pz);
            }
#pragma unroll
            for (int u = 0; u < 4; u++) {
                cx[u] = nx[u];
                cy[u] = ny[u];
                cz[u] = nz[u];
                cg[u] = ng[u];
            }
        }
    }
    for (; q < p1; q++) {
        const double rx = xt - X[q], ry = yt - Y[q], rz = zt - Z[q];
        const double f = gm_over_r3_fast(G[q], dot3(rx, ry, rz, rx, ry, rz), fc.c15, fc.c1875);
        px = fma(-f, rx, px);
        py = fma(-f, ry, py);
        pz = fma(-f, rz, pz);
    }
    if (!valid) return;
    double ax = 0.0 + px, ay = 0.0 + py, az = 0.0 + pz;
    if (do_pairs) {
        const int64_t c0 = pt.colstart[j], c1 = c0 + pt.colcnt[j];
        for (int64_t c = c0; c < c1; c++) {
            const int i = pt.i1[pt.col[c]];
            const double dx = xt - rp.x[i], dy = yt - rp.y[i], dz = zt - rp.z[i];
            const double f = gm_over_r3_fast(pl.Gm[i], dot3(dx, dy, dz, dx, dy, dz));
            ax = fma(f, dx, ax);
            ay = fma(f, dy, ay);
            az = fma(f, dz, az);
        }
    }
    tp.vb.x[j] = fma(ax, dt, tp.vb.x[j]);
    tp.vb.y[j] = fma(ay, dt, tp.vb.y[j]);
    tp.vb.z[j] = fma(az, dt, tp.vb.z[j]);
}

// Danby drift of the masked bodies (helio_drift_body, helio_drift.f90:14-54,
// with the SyMBA level mask of symba_drift.f90:14-61): mu = GM_cb, (rh, vb).
__global__ __launch_bounds__(256) void k_drift(Vec3p r, Vec3p v, int32_t *status, const int32_t *levelg, const int32_t *sys,
                        const double *GMcb, int32_t *err, const int32_t *frozen, int n, double dt, int irec,
                        int use_level, int is_tp)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int st = status[i];
    bool m = is_tp ? (st == ACTIVE) : (st != INACTIVE);
    if (use_level) m = m && (levelg[i] == irec);
    if (!m) return;
    int s = sys[i];
    if (frozen[s]) return;
    double rx = r.x[i], ry = r.y[i], rz = r.z[i], vx = v.x[i], vy = v.y[i], vz = v.z[i];
    int fl = drift_one(GMcb[s], rx, ry, rz, vx, vy, vz, dt);
    r.x[i] = rx;
    r.y[i] = ry;
    r.z[i] = rz;
    v.x[i] = vx;
    v.y[i] = vy;
    v.z[i] = vz;
    if (fl != 0) {
        status[i] = DISCARDED_DRIFTERR;
        if (!is_tp) atomicOr(&err[s], 2);
    }
}

// Test particles at the start of a step: vb = vh - ptbeg when lfirst
// (swiftest_util.f90:398-421), then rh += ptbeg*dt (helio_drift.f90:166-200).
__global__ void k_tp_pre(TpDev tp, SysDev sy, int ntp, double dth)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntp || !tp_active(tp, j)) return;
    int s = tp.sys[j];
    if (sy.frozen[s]) return;
    double px = sy.ptbeg[3 * s], py = sy.ptbeg[3 * s + 1], pz = sy.ptbeg[3 * s + 2];
    if (sy.lfirst[s]) {
        tp.vb.x[j] = tp.vh.x[j] + (-px);
        tp.vb.y[j] = tp.vh.y[j] + (-py);
        tp.vb.z[j] = tp.vh.z[j] + (-pz);
    }
    tp.rh.x[j] = fma(px, dth, tp.rh.x[j]);
    tp.rh.y[j] = fma(py, dth, tp.rh.y[j]);
    tp.rh.z[j] = fma(pz, dth, tp.rh.z[j]);
}

// Test-particle half kick (helio_kick_vb_tp, helio_kick.f90:135-169):
//   p = sum over pl j ascending of -Gm_j (r_t - r_j)/|.|^3 at pl rbeg/rend,
//       started from zero and added to ah = 0 once (the OpenMP reduction of
//       swiftest_kick.f90:394-411);
//   ah += Gm_i dx/|dx|^3 for the particle's listed encounter pairs in list
//       order (symba_kick.f90:106-119);  vb += ah*dt.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_kick_tp(TpDev tp, PlDev pl, SysDev sy, ListDev pt, const Tile *tiles,
                                                 double dt, int lbeg, int do_pairs)
{
    __shared__ double sx[BLK], syy[BLK], sz[BLK], sgm[BLK];
    const Tile tl = tiles[blockIdx.x];
    const int p0 = sy.pl_off[tl.s], p1 = sy.pl_off[tl.s + 1];
    const int j = tl.b0 + (int)threadIdx.x;
    const bool valid = (j < tl.b1) && tp_active(tp, j);
    const Vec3p rp = lbeg ? pl.rbeg : pl.rend;
    double xt = 0.0, yt = 0.0, zt = 0.0;
    if (valid) {
        xt = tp.rh.x[j];
        yt = tp.rh.y[j];
        zt = tp.rh.z[j];
    }
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int it = p0; it < p1; it += BLK) {
        int q0 = it + (int)threadIdx.x;
        bool gok = true;
        if (q0 < p1) {
            sx[threadIdx.x] = rp.x[q0];
            syy[threadIdx.x] = rp.y[q0];
            sz[threadIdx.x] = rp.z[q0];
            sgm[threadIdx.x] = pl.Gm[q0];
            gok = pair_g_ok(pl.Gm[q0]);
        }
        const bool tile_gok = __syncthreads_and(gok);
        const int qn = min(BLK, p1 - it);
        if (valid) {
            // Whole groups of KTP_G bodies: the factors are formed without
            // branches (exact cores; the rare out-of-range terms redone with
            // the IEEE operations under a wave-uniform test) and folded in
            // j order.  The tail of a partial tile takes the plain path.
            int q = 0;
            for (; q + KTP_G <= qn; q += KTP_G) {
                double rx[KTP_G], ry[KTP_G], rz[KTP_G], r2[KTP_G], fac[KTP_G];
                bool need[KTP_G];
                bool any_need = false;
#pragma unroll
                for (int u = 0; u < KTP_G; u++) {
                    rx[u] = xt - sx[q + u];
                    ry[u] = yt - syy[q + u];
                    rz[u] = zt - sz[q + u];
                    r2[u] = dot3(rx[u], ry[u], rz[u], rx[u], ry[u], rz[u]);
                    fac[u] = div_core(sgm[q + u], r2[u] * sqrt_core(r2[u]));
                    need[u] = !(tile_gok & (r2[u] >= PAIR_R2_LO) & (r2[u] <= PAIR_R2_HI));
                    any_need = any_need | need[u];
                }
                if (__ballot(any_need)) {
#pragma unroll
                    for (int u = 0; u < KTP_G; u++)
                        if (need[u]) fac[u] = sgm[q + u] / (r2[u] * sqrt(r2[u]));
                }
#pragma unroll
                for (int u = 0; u < KTP_G; u++) {
                    px = fma(-fac[u], rx[u], px);
                    py = fma(-fac[u], ry[u], py);
                    pz = fma(-fac[u], rz[u], pz);
                }
            }
            for (; q < qn; q++) {
                double rx = xt - sx[q], ry = yt - syy[q], rz = zt - sz[q];
                double rji2 = dot3(rx, ry, rz, rx, ry, rz);
                double fac = gm_over_r3(sgm[q], rji2, tile_gok);
                px = fma(-fac, rx, px);
                py = fma(-fac, ry, py);
                pz = fma(-fac, rz, pz);
            }
        }
        __syncthreads();
    }
    if (!valid) return;
    double ax = 0.0 + px, ay = 0.0 + py, az = 0.0 + pz;
    if (do_pairs) {
        const int64_t c0 = pt.colstart[j], c1 = c0 + pt.colcnt[j];
        for (int64_t c = c0; c < c1; c++) {
            int i = pt.i1[pt.col[c]];
            double dx = xt - rp.x[i], dy = yt - rp.y[i], dz = zt - rp.z[i];
            double rjj = dot3(dx, dy, dz, dx, dy, dz);
            double fac = pl.Gm[i] / (rjj * sqrt(rjj));
            ax = fma(fac, dx, ax);
            ay = fma(fac, dy, ay);
            az = fma(fac, dz, az);
        }
    }
    tp.vb.x[j] = fma(ax, dt, tp.vb.x[j]);
    tp.vb.y[j] = fma(ay, dt, tp.vb.y[j]);
    tp.vb.z[j] = fma(az, dt, tp.vb.z[j]);
}

// Per-system sums at the end of a step:
//   lindrift pt = sum_masked(Gm vb) / GM_cb -> ptend (forward, lanes 0-2)
//   vb->vh (swiftest_util.f90:363-395): cb%vb -= Gm vb / GM_cb over
//   i = npl..1, active only (backward, lanes 0-2)
__global__ __launch_bounds__(SUM_T) void k_sums_end(PlDev pl, SysDev sy, int nsys)
{
    __shared__ SumStage S;
    const int s = blockIdx.x, tid = threadIdx.x;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    if (p1 == p0 || sy.frozen[s]) return;
    const double gm = sy.GMcb[s];
    double pt = 0.0;
    for (int c0 = p0; c0 < p1; c0 += SUM_CH) {
        const int n = min(SUM_CH, p1 - c0);
        for (int k = tid; k < n; k += SUM_T) {
            const int i = c0 + k;
            const double g = pl.Gm[i];
            const bool act = pl.status[i] != INACTIVE;
            S.t[0][k] = act ? g * pl.vb.x[i] : 0.0;
            S.t[1][k] = act ? g * pl.vb.y[i] : 0.0;
            S.t[2][k] = act ? g * pl.vb.z[i] : 0.0;
        }
        __syncthreads();
        if (tid < 3) pt = chain<false, false>(pt, S.t[tid], n);
        __syncthreads();
    }
    if (tid < 3) sy.ptend[3 * s + tid] = pt / gm;
    double c = 0.0;
    for (int c1 = p1; c1 > p0; c1 -= SUM_CH) {
        const int c0 = max(p0, c1 - SUM_CH), n = c1 - c0;
        for (int k = tid; k < n; k += SUM_T) {
            const int i = c0 + k;
            const double g = pl.Gm[i];
            const bool act = pl.status[i] != INACTIVE;
            S.t[0][k] = act ? (g * pl.vb.x[i]) / gm : 0.0;
            S.t[1][k] = act ? (g * pl.vb.y[i]) / gm : 0.0;
            S.t[2][k] = act ? (g * pl.vb.z[i]) / gm : 0.0;
        }
        __syncthreads();
        if (tid < 3) c = chain<true, true>(c, S.t[tid], n);
        __syncthreads();
    }
    if (tid < 3) sy.cbvb[3 * s + tid] = c;
}

// rh += ptend*dt (masked); vh = vb - cb%vb (all bodies)
__global__ void k_pl_end_apply(PlDev pl, SysDev sy, int npl, double dth)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npl) return;
    int s = pl.sys[i];
    if (sy.frozen[s]) return;
    if (pl_active(pl, i)) {
        pl.rh.x[i] = fma(sy.ptend[3 * s], dth, pl.rh.x[i]);
        pl.rh.y[i] = fma(sy.ptend[3 * s + 1], dth, pl.rh.y[i]);
        pl.rh.z[i] = fma(sy.ptend[3 * s + 2], dth, pl.rh.z[i]);
    }
    pl.vh.x[i] = pl.vb.x[i] - sy.cbvb[3 * s];
    pl.vh.y[i] = pl.vb.y[i] - sy.cbvb[3 * s + 1];
    pl.vh.z[i] = pl.vb.z[i] - sy.cbvb[3 * s + 2];
}

// Test particles at the end of a step: rh += ptend*dt, vh = vb + ptend
// (vb2vh with vbcb = -ptend)
__global__ void k_tp_post(TpDev tp, SysDev sy, int ntp, double dth)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntp || !tp_active(tp, j)) return;
    int s = tp.sys[j];
    if (sy.frozen[s]) return;
    double px = sy.ptend[3 * s], py = sy.ptend[3 * s + 1], pz = sy.ptend[3 * s + 2];
    tp.rh.x[j] = fma(px, dth, tp.rh.x[j]);
    tp.rh.y[j] = fma(py, dth, tp.rh.y[j]);
    tp.rh.z[j] = fma(pz, dth, tp.rh.z[j]);
    tp.vh.x[j] = tp.vb.x[j] - (-px);
    tp.vh.y[j] = tp.vb.y[j] - (-py);
    tp.vh.z[j] = tp.vb.z[j] - (-pz);
}

// The HELIO branch clears lfirst after its vh->vb; the encounter branch keeps
// it (helio_step.f90:58-61 vs symba_step.f90:81, then symba_step.f90:47).
__global__ void k_lfirst_update(SysDev sy, int nsys)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys || sy.frozen[s]) return;
    if (sy.lfirst[s] && !(sy.enc_pp[s] || sy.enc_pt[s])) sy.lfirst[s] = 0;
}

// Snapshot coordinate refresh (swiftest_util.f90:2663-2682): pl vh->vb, then
// tp vb = vh - ptbeg for active particles.
__global__ void k_snapshot_tp(TpDev tp, SysDev sy, int ntp)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntp || !tp_active(tp, j)) return;
    int s = tp.sys[j];
    if (sy.frozen[s]) return;
    tp.vb.x[j] = tp.vh.x[j] + (-sy.ptbeg[3 * s]);
    tp.vb.y[j] = tp.vh.y[j] + (-sy.ptbeg[3 * s + 1]);
    tp.vb.z[j] = tp.vh.z[j] + (-sy.ptbeg[3 * s + 2]);
}

__global__ __launch_bounds__(SUM_T) void k_vh2vb_sum(PlDev pl, SysDev sy, int nsys)
{
    __shared__ SumStage S;
    __shared__ double shr[4];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    if (p1 == p0 || sy.frozen[s]) return;
    const double acc = sums_gm_vh(S, pl, p0, p1);
    if (tid < 4) shr[tid] = acc;
    __syncthreads();
    if (tid < 3) sy.cbvb[3 * s + tid] = shr[tid + 1] / (sy.GMcb[s] + shr[0]);
}

__global__ void k_vh2vb_apply(PlDev pl, SysDev sy, int npl)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npl) return;
    int s = pl.sys[i];
    if (sy.frozen[s]) return;
    pl.vb.x[i] = pl.vh.x[i] + sy.cbvb[3 * s];
    pl.vb.y[i] = pl.vh.y[i] + sy.cbvb[3 * s + 1];
    pl.vb.z[i] = pl.vh.z[i] + sy.cbvb[3 * s + 2];
}

// The initialize sequence of one system loaded into slot s of a running
// ensemble (the per-system part of k_vh2vb_sum, k_vh2vb_apply and
// k_init_tp_vb): cb%vb from vh over the slot's bodies, vb = vh + cb%vb for its
// massive bodies and test particles.  Padding bodies have Gm = 0 and vh = 0,
// so they add exactly nothing to the sums.  One workgroup.
__global__ __launch_bounds__(SUM_T) void k_init_one(PlDev pl, TpDev tp, SysDev sy, int s)
{
    __shared__ SumStage S;
    __shared__ double shr[4];
    const int tid = threadIdx.x;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    const double acc = sums_gm_vh(S, pl, p0, p1);
    if (tid < 4) shr[tid] = acc;
    __syncthreads();
    if (tid < 3) sy.cbvb[3 * s + tid] = p1 > p0 ? shr[tid + 1] / (sy.GMcb[s] + shr[0]) : 0.0;
    __syncthreads();
    const double cx = sy.cbvb[3 * s], cy = sy.cbvb[3 * s + 1], cz = sy.cbvb[3 * s + 2];
    for (int i = p0 + tid; i < p1; i += SUM_T) {
        pl.vb.x[i] = pl.vh.x[i] + cx;
        pl.vb.y[i] = pl.vh.y[i] + cy;
        pl.vb.z[i] = pl.vh.z[i] + cz;
    }
    for (int j = sy.tp_off[s] + tid; j < sy.tp_off[s + 1]; j += SUM_T) {
        tp.vb.x[j] = tp.vh.x[j] + cx;
        tp.vb.y[j] = tp.vh.y[j] + cy;
        tp.vb.z[j] = tp.vh.z[j] + cz;
    }
}

// Initialization (helio_util.f90:26-28): tp vb = vh + cb%vb
__global__ void k_init_tp_vb(TpDev tp, SysDev sy, int ntp)
{
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= ntp) return;
    int s = tp.sys[j];
    tp.vb.x[j] = tp.vh.x[j] + sy.cbvb[3 * s];
    tp.vb.y[j] = tp.vh.y[j] + sy.cbvb[3 * s + 1];
    tp.vb.z[j] = tp.vh.z[j] + sy.cbvb[3 * s + 2];
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_reset_levels(hipStream_t st, PlDev pl, TpDev tp, int npl, int ntp)
{
    int n = npl > ntp ? npl : ntp;
    if (n) k_reset_levels<<<cdiv(n, 256), 256, 0, st>>>(pl, tp, npl, ntp);
}
void launch_sums_begin(hipStream_t st, PlDev pl, SysDev sy, int nsys, int force)
{
    k_sums_begin<<<nsys, SUM_T, 0, st>>>(pl, sy, nsys, force);
}
void launch_pl_begin_apply(hipStream_t st, PlDev pl, SysDev sy, int npl, double dth, int force)
{
    if (npl) k_pl_begin_apply<<<cdiv(npl, 256), 256, 0, st>>>(pl, sy, npl, dth, force);
}
void launch_kick_pl(hipStream_t st, const Tuning &tu, PlDev pl, SysDev sy, ListDev pp, const Tile *tiles,
                    int ntiles, double dt, int lbeg, int use_radius, int do_pairs)
{
    if (!ntiles) return;
    if (tu.any_tiny) {
        k_kick_pl_tiny<<<ntiles, 256, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs, tu.pair_fast);
        return;
    }
    if (tu.pair_fast) {
        switch (tu.kick_pl_split) {
        case 1: k_kick_pl_fast<1><<<ntiles, 64, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        case 2: k_kick_pl_fast<2><<<ntiles, 128, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        case 4: k_kick_pl_fast<4><<<ntiles, 256, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        case 16: k_kick_pl_fast<16><<<ntiles, 1024, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        default: k_kick_pl_fast<8><<<ntiles, 512, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        }
        return;
    }
    if (tu.kick_pl_eval) {
        switch (tu.kick_pl_eval) {
        case 2: k_kick_pl_split<2><<<ntiles, 64 * 5, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        case 8: k_kick_pl_split<8><<<ntiles, 64 * 11, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        default: k_kick_pl_split<4><<<ntiles, 64 * 7, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
        }
        return;
    }
    switch (tu.kick_pl_block) {
    case 64: k_kick_pl<64><<<ntiles, 64, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
    case 128: k_kick_pl<128><<<ntiles, 128, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
    default: k_kick_pl<256><<<ntiles, 256, 0, st>>>(pl, sy, pp, tiles, dt, lbeg, use_radius, do_pairs); break;
    }
}
void launch_drift(hipStream_t st, Vec3p r, Vec3p v, int32_t *status, int32_t *levelg, const int32_t *sys,
                  const double *GMcb, int32_t *err, const int32_t *frozen, int n, double dt, int irec, int use_level,
                  int is_tp)
{
    if (n)
        k_drift<<<cdiv(n, 128), 128, 0, st>>>(r, v, status, levelg, sys, GMcb, err, frozen, n, dt, irec, use_level,
                                              is_tp);
}
void launch_tp_pre(hipStream_t st, TpDev tp, SysDev sy, int ntp, double dth)
{
    if (ntp) k_tp_pre<<<cdiv(ntp, 256), 256, 0, st>>>(tp, sy, ntp, dth);
}
void launch_kick_tp(hipStream_t st, const Tuning &tu, TpDev tp, PlDev pl, SysDev sy, ListDev pt,
                    const Tile *tiles, int ntiles, double dt, int lbeg, int do_pairs)
{
    if (!ntiles) return;
    if (tu.pair_fast) {
        switch (tu.kick_tp_block) {
        case 64: k_kick_tp_fast<64><<<ntiles, 64, 0, st>>>(tp, pl, sy, pt, tiles, dt, lbeg, do_pairs); break;
        case 128: k_kick_tp_fast<128><<<ntiles, 128, 0, st>>>(tp, pl, sy, pt, tiles, dt, lbeg, do_pairs); break;
        default: k_kick_tp_fast<256><<<ntiles, 256, 0, st>>>(tp, pl, sy, pt, tiles, dt, lbeg, do_pairs); break;
        }
        return;
    }
    switch (tu.kick_tp_block) {
    case 64: k_kick_tp<64><<<ntiles, 64, 0, st>>>(tp, pl, sy, pt, tiles, dt, lbeg, do_pairs); break;
    case 128: k_kick_tp<128><<<ntiles, 128, 0, st>>>(tp, pl, sy, pt, tiles, dt, lbeg, do_pairs); break;
    default: k_kick_tp<256><<<ntiles, 256, 0, st>>>(tp, pl, sy, pt, tiles, dt, lbeg, do_pairs); break;
    }
}
void launch_sums_end(hipStream_t st, PlDev pl, SysDev sy, int nsys)
{
    k_sums_end<<<nsys, SUM_T, 0, st>>>(pl, sy, nsys);
}
void launch_pl_end_apply(hipStream_t st, PlDev pl, SysDev sy, int npl, double dth)
{
    if (npl) k_pl_end_apply<<<cdiv(npl, 256), 256, 0, st>>>(pl, sy, npl, dth);
}
void launch_tp_post(hipStream_t st, TpDev tp, SysDev sy, int ntp, double dth)
{
    if (ntp) k_tp_post<<<cdiv(ntp, 256), 256, 0, st>>>(tp, sy, ntp, dth);
}
void launch_lfirst_update(hipStream_t st, SysDev sy, int nsys)
{
    k_lfirst_update<<<cdiv(nsys, 256), 256, 0, st>>>(sy, nsys);
}
void launch_snapshot(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, int nsys, int npl, int ntp)
{
    k_vh2vb_sum<<<nsys, SUM_T, 0, st>>>(pl, sy, nsys);
    if (npl) k_vh2vb_apply<<<cdiv(npl, 256), 256, 0, st>>>(pl, sy, npl);
    if (ntp) k_snapshot_tp<<<cdiv(ntp, 256), 256, 0, st>>>(tp, sy, ntp);
}
void launch_vh2vb(hipStream_t st, PlDev pl, SysDev sy, int nsys, int npl)
{
    k_vh2vb_sum<<<nsys, SUM_T, 0, st>>>(pl, sy, nsys);
    if (npl) k_vh2vb_apply<<<cdiv(npl, 256), 256, 0, st>>>(pl, sy, npl);
}
void launch_init_one(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, int s)
{
    k_init_one<<<1, SUM_T, 0, st>>>(pl, tp, sy, s);
}
void launch_init_tp_vb(hipStream_t st, TpDev tp, SysDev sy, int ntp)
{
    if (ntp) k_init_tp_vb<<<cdiv(ntp, 256), 256, 0, st>>>(tp, sy, ntp);
}

}  // namespace swg
