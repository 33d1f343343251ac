// k_sas.hip -- sort-and-sweep broad phase for the level-0 pl-pl encounter
// check (ENCOUNTER_CHECK SORTSWEEP, encounter_check.f90:178-185, :795-988).
//
// The reference sorts radial extents and sweeps them, ignoring the motion over
// the step and skipping intervals that hold fewer than two foreign markers
// (encounter_check.f90:951), so it can drop pairs.  This broad phase is
// conservative instead, so that after the same exact narrow phase as the
// triangular check
// (encounter_check_one, encounter_check.f90:573-621) the list is identical to
// the TRIANGULAR list:
//   1. each body's x coordinate swept over [0, dt] with its velocity, widened
//      by its encounter radius and a relative margin, gives [lo, hi]; a pair
//      whose intervals are disjoint stays farther apart than renc_i + renc_j
//      over the whole step and cannot be an encounter;
//   2. a segmented device radix sort orders each system's bodies by lo; the
//      partners j of body i with lo_j in (lo_i, hi_i] are then the contiguous
//      run of the sorted order after i, found with one binary search, and
//      every overlapping pair is produced exactly once;
//   3. the narrow phase runs on every candidate in parallel; the survivors are
//      compacted and radix-sorted by (index1, index2), which is the triangular
//      list order, and the index2-side adjacency is a stable sort by index2.
// The sorts and the compaction are this file's own kernels: a stable LSD
// radix sort (8-bit digits; per pass a tile histogram, one scan, and a
// scatter that ranks equal digits with wavefront ballots) and an ordered
// ballot compaction.
#include "swg_kernels.h"
#include "swg_math.h"

namespace swg {

static inline int cdiv64(int64_t a, int b) { return (int)((a + b - 1) / b); }

// relative widening of the extents: covers the rounding of x + v*dt and the
// narrow phase's own rounding of r2min against R^2 many times over
constexpr double SAS_REL = 1e-12;

__device__ inline uint64_t sortable(double x);

__global__ void k_sas_extents(PlDev pl, int npl, double dt, uint64_t *lokey, double *hi, int32_t *idx)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npl) return;
    const double x = pl.rh.x[g], x1 = x + pl.vb.x[g] * dt, r = pl.renc[g];
    const double m = SAS_REL * (fabs(x) + fabs(x1) + r) + 1e-300;
    lokey[g] = sortable(fmin(x, x1) - r - m);
    hi[g] = fmax(x, x1) + r + m;
    idx[g] = g;
}

__global__ void k_sas_rank(const int32_t *sidx, int npl, int32_t *rank)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < npl) rank[sidx[r]] = r;
}

// candidates of body g: sorted positions rank+1 .. (first lo > hi_g) - 1
__global__ void k_sas_count(const double *slo, const double *hi, const int32_t *rank, const int32_t *sys,
                            const int32_t *pl_off, const int32_t *active, int npl, int32_t *cnt)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npl) return;
    const int s = sys[g];
    if (active && !active[s]) {
        cnt[g] = 0;
        return;
    }
    int lo = rank[g] + 1, hiend = pl_off[s + 1];
    const double h = hi[g];
    while (lo < hiend) {  // upper bound of h in slo[rank+1, end)
        int m = (lo + hiend) >> 1;
        if (slo[m] <= h) lo = m + 1;
        else hiend = m;
    }
    cnt[g] = lo - rank[g] - 1;
}

// the bodies' data in sorted order, so that a body's candidates (a run of
// the sorted order) are read contiguously
__global__ void k_sas_permute(PlDev pl, const int32_t *sidx, int npl, double *srt, int8_t *stiny)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= npl) return;
    const int i = sidx[r];
    srt[r] = pl.rh.x[i];
    srt[(size_t)npl + r] = pl.rh.y[i];
    srt[2 * (size_t)npl + r] = pl.rh.z[i];
    srt[3 * (size_t)npl + r] = pl.vb.x[i];
    srt[4 * (size_t)npl + r] = pl.vb.y[i];
    srt[5 * (size_t)npl + r] = pl.vb.z[i];
    srt[6 * (size_t)npl + r] = pl.renc[i];
    stiny[r] = (pl.tiny && pl.tiny[i]) ? 1 : 0;
}

// narrow phase (the pair test of k_pp_tiles with index1 < index2): one
// wavefront per body g, its lanes over g's candidates, the bodies at sorted
// positions rank[g] + 1 .. rank[g] + cnt[g]; results at coff[g] + k
__global__ __launch_bounds__(256) void k_sas_narrow(PlDev pl, const int64_t *coff, const int32_t *cnt,
                                                    const int32_t *rank, const int32_t *sidx, const double *srt,
                                                    const int8_t *stiny, int npl, double dt, uint64_t *key,
                                                    uint8_t *flag)
{
    const int g = blockIdx.x * 4 + (int)(threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (g >= npl) return;
    const int n = cnt[g];
    if (n == 0) return;
    const int64_t o = coff[g];
    const int r0 = rank[g] + 1;
    const double gx = pl.rh.x[g], gy = pl.rh.y[g], gz = pl.rh.z[g];
    const double gvx = pl.vb.x[g], gvy = pl.vb.y[g], gvz = pl.vb.z[g], gr = pl.renc[g];
    const bool gtiny = pl.tiny && pl.tiny[g];
    const size_t N = (size_t)npl;
    for (int k = lane; k < n; k += 64) {
        const int r = r0 + k;
        const int j = sidx[r];
        const double jx = srt[r], jy = srt[N + r], jz = srt[2 * N + r];
        const double jvx = srt[3 * N + r], jvy = srt[4 * N + r], jvz = srt[5 * N + r], jr = srt[6 * N + r];
        const bool gfirst = g < j;
        const int a = gfirst ? g : j, b = gfirst ? j : g;
        // differences b - a, as the pair test forms them
        const double dx = gfirst ? jx - gx : gx - jx, dy = gfirst ? jy - gy : gy - jy, dz = gfirst ? jz - gz : gz - jz;
        const double dvx = gfirst ? jvx - gvx : gvx - jvx, dvy = gfirst ? jvy - gvy : gvy - jvy,
                     dvz = gfirst ? jvz - gvz : gvz - jvz;
        const double rsum = gfirst ? gr + jr : jr + gr;
        const bool atiny = gfirst ? gtiny : stiny[r] != 0;
        bool lv;
        // a semi-interacting index1 (GMTINY): the pair is not checked
        const bool f = !atiny && encounter_one(dx, dy, dz, dvx, dvy, dvz, rsum, dt, lv);
        key[o + k] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
        flag[o + k] = f;
    }
}

// rows of the sorted pair keys: rowstart[g] = first key >= g<<32
__global__ void k_sas_rows(const uint64_t *keys, int64_t n, int npl, int32_t *rowcnt, int64_t *rowstart)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npl) return;
    auto lb = [&](uint64_t k) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            int64_t m = (lo + hi) >> 1;
            if (keys[m] < k) lo = m + 1;
            else hi = m;
        }
        return lo;
    };
    const int64_t r0 = lb((uint64_t)(uint32_t)g << 32), r1 = lb((uint64_t)(uint32_t)(g + 1) << 32);
    rowstart[g] = r0;
    rowcnt[g] = (int32_t)(r1 - r0);
}

// the list itself, as k_rows_fill writes it
__global__ void k_sas_fill(ListDev L, const uint64_t *keys, int64_t n, int32_t *levg, int32_t *levm, int32_t *pid)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const int a = (int)(keys[k] >> 32), b = (int)(keys[k] & 0xffffffffu);
    L.i1[k] = a;
    L.i2[k] = b;
    L.lvdotr[k] = 1;
    L.status[k] = (int8_t)ACTIVE;
    L.level[k] = 0;
    levg[a] = 0;
    levm[a] = 0;
    levg[b] = 0;
    levm[b] = 0;
    pid[k] = (int32_t)k;
}

// index2-side adjacency from the pairs sorted by index2 (stable: list order)
__global__ void k_sas_cols(const int32_t *sorted_i2, int64_t n, int npl, int32_t *colcnt, int64_t *colstart)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npl) return;
    auto lb = [&](int32_t k) {
        int64_t lo = 0, hi = n;
        while (lo < hi) {
            int64_t m = (lo + hi) >> 1;
            if (sorted_i2[m] < k) lo = m + 1;
            else hi = m;
        }
        return lo;
    };
    const int64_t c0 = lb(g), c1 = lb(g + 1);
    colstart[g] = c0;
    colcnt[g] = (int32_t)(c1 - c0);
}

// ---- stable LSD radix sort ------------------------------------------------
// Keys K (uint64_t or uint32_t) with optional uint32 values, digits of 8 bits
// from begin_bit to end_bit; then, when seg is given, digits of seg[value]
// (the system of each body), so that the result is ordered by (seg, key) and
// stable.  A tile of RS_TILE items is handled by one workgroup of 4
// wavefronts; each wavefront takes a contiguous quarter in chunks of 64.
constexpr int RS_WG = 256, RS_TILE = 4096, RS_BINS = 256;

template <class K>
__device__ inline int rs_digit(K key, uint32_t val, const int32_t *seg, int shift)
{
    return seg ? (int)(((uint32_t)seg[val] >> shift) & 255u) : (int)((key >> shift) & (K)255);
}

template <class K>
__global__ __launch_bounds__(RS_WG) void k_rs_hist(const K *keys, const uint32_t *vals, const int32_t *seg, int64_t n,
                                                   int shift, int ntiles, int32_t *hist)
{
    __shared__ int32_t h[RS_BINS];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int k = threadIdx.x; k < RS_TILE; k += RS_WG) {
        const int64_t i = base + k;
        if (i < n) atomicAdd(&h[rs_digit<K>(keys[i], vals ? vals[i] : 0u, seg, shift)], 1);
    }
    __syncthreads();
    hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

template <class K>
__global__ __launch_bounds__(RS_WG) void k_rs_scatter(const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                                                      const int32_t *seg, int64_t n, int shift, int ntiles,
                                                      const int64_t *offs)
{
    __shared__ int32_t wcnt[4][RS_BINS];
    __shared__ int64_t boff[RS_BINS];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * RS_TILE + (int64_t)wv * (RS_TILE / 4);
    for (int b = lane; b < RS_BINS; b += 64) wcnt[wv][b] = 0;
    __syncthreads();
    for (int c = 0; c < RS_TILE / 4; c += 64) {
        const int64_t i = base + c + lane;
        if (i < n) atomicAdd(&wcnt[wv][rs_digit<K>(kin[i], vin ? vin[i] : 0u, seg, shift)], 1);
    }
    __syncthreads();
    {
        const int b = threadIdx.x;
        boff[b] = offs[(int64_t)b * ntiles + blockIdx.x];
        int run = 0;
        for (int q = 0; q < 4; q++) {
            const int c = wcnt[q][b];
            wcnt[q][b] = run;
            run += c;
        }
    }
    __syncthreads();
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int c = 0; c < RS_TILE / 4; c += 64) {
        const int64_t i = base + c + lane;
        const bool valid = i < n;
        K key = 0;
        uint32_t val = 0;
        int d = 0;
        if (valid) {
            key = kin[i];
            val = vin ? vin[i] : 0u;
            d = rs_digit<K>(key, val, seg, shift);
        }
        // lanes holding the same digit: 8 ballots
        uint64_t m = __ballot(valid);
        for (int bit = 0; bit < 8; bit++) {
            const bool set = (d >> bit) & 1;
            const uint64_t bb = __ballot(set);
            m &= set ? bb : ~bb;
        }
        const int rank = __popcll(m & below), cnt = __popcll(m);
        const int prev = valid ? wcnt[wv][d] : 0;
        __builtin_amdgcn_wave_barrier();
        if (valid) {
            const int64_t pos = boff[d] + prev + rank;
            kout[pos] = key;
            if (vout) vout[pos] = val;
            if (rank == 0) wcnt[wv][d] = prev + cnt;
        }
        __builtin_amdgcn_wave_barrier();
    }
}

// ---- ordered compaction of flagged keys ------------------------------------
constexpr int CP_TILE = 4096;

__global__ __launch_bounds__(RS_WG) void k_cp_count(const uint8_t *flag, int64_t n, int32_t *cnt)
{
    __shared__ int32_t wsum[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * CP_TILE + (int64_t)wv * (CP_TILE / 4);
    int c = 0;
    for (int k = 0; k < CP_TILE / 4; k += 64) {
        const int64_t i = base + k + lane;
        c += __popcll(__ballot(i < n && flag[i]));
    }
    if (lane == 0) wsum[wv] = c;
    __syncthreads();
    if (threadIdx.x == 0) cnt[blockIdx.x] = wsum[0] + wsum[1] + wsum[2] + wsum[3];
}

__global__ __launch_bounds__(RS_WG) void k_cp_scatter(const uint64_t *key, const uint8_t *flag, int64_t n,
                                                      const int64_t *boff, uint64_t *out)
{
    __shared__ int32_t wsum[4];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int64_t base = (int64_t)blockIdx.x * CP_TILE + (int64_t)wv * (CP_TILE / 4);
    int c = 0;
    for (int k = 0; k < CP_TILE / 4; k += 64) {
        const int64_t i = base + k + lane;
        c += __popcll(__ballot(i < n && flag[i]));
    }
    if (lane == 0) wsum[wv] = c;
    __syncthreads();
    int64_t pos = boff[blockIdx.x];
    for (int q = 0; q < wv; q++) pos += wsum[q];
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int k = 0; k < CP_TILE / 4; k += 64) {
        const int64_t i = base + k + lane;
        const bool f = i < n && flag[i];
        const uint64_t m = __ballot(f);
        if (f) out[pos + __popcll(m & below)] = key[i];
        pos += __popcll(m);
    }
}

// transformed doubles: the unsigned order of the bits is the numeric order
__device__ inline uint64_t sortable(double x)
{
    const uint64_t u = (uint64_t)__double_as_longlong(x);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ inline double unsortable(uint64_t u)
{
    return __longlong_as_double((long long)((u >> 63) ? (u & ~(1ull << 63)) : ~u));
}

__global__ void k_sas_unkey(const uint64_t *skey, int npl, double *slo)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < npl) slo[r] = unsortable(skey[r]);
}

// ---- host side ---------------------------------------------------------------

static int grow(void **p, size_t *cap, size_t need)
{
    if (need <= *cap) return 0;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    size_t c = need + need / 4 + 4096;
    if (hipMalloc(p, c) != hipSuccess) return -1;
    *cap = c;
    return 0;
}

#define SASCHK(x)                                        \
    do {                                                 \
        if ((x) != hipSuccess) return hipGetLastError(); \
    } while (0)

SasWork::~SasWork()
{
    for (void *p : {(void *)lokey, (void *)slokey, (void *)hi, (void *)slo, (void *)idx, (void *)sidx, (void *)rank,
                    (void *)cnt, (void *)coff, (void *)key, (void *)flag, (void *)sel, (void *)skey, (void *)nsel,
                    (void *)pid, (void *)spid, (void *)si2, rtk, rtv, rhist, roffs, rscan, rtotal, (void *)srt,
                    (void *)stiny})
        if (p) (void)hipFree(p);
}

// Stable sort of n keys (and values) by their bits [begin_bit, end_bit), then
// by seg[value] over seg_bits bits when seg is given.  kin / vin are not
// written; the result is in kout / vout.
template <class K>
static hipError_t radix_sort(hipStream_t st, SasWork &w, const K *kin, const uint32_t *vin, K *kout, uint32_t *vout,
                             int64_t n, int begin_bit, int end_bit, const int32_t *seg, int seg_bits)
{
    if (n <= 0) return hipSuccess;
    const int ntiles = (int)((n + RS_TILE - 1) / RS_TILE);
    const int kp = (end_bit - begin_bit + 7) / 8, sp = seg ? (seg_bits + 7) / 8 : 0, np = kp + sp;
    if (np == 0) {
        SASCHK(hipMemcpyAsync(kout, kin, n * sizeof(K), hipMemcpyDeviceToDevice, st));
        if (vin) SASCHK(hipMemcpyAsync(vout, vin, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, st));
        return hipSuccess;
    }
    const int nh = RS_BINS * ntiles;
    if (grow(&w.rtk, &w.rtk_cap, n * sizeof(K)) || (vin && grow(&w.rtv, &w.rtv_cap, n * sizeof(uint32_t))) ||
        grow(&w.rhist, &w.rhist_cap, nh * sizeof(int32_t)) || grow(&w.roffs, &w.roffs_cap, nh * sizeof(int64_t)) ||
        grow(&w.rscan, &w.rscan_cap, scan_scratch_bytes(nh)) || grow(&w.rtotal, &w.rtotal_cap, sizeof(int64_t)))
        return hipErrorOutOfMemory;
    const K *sk = kin;
    const uint32_t *sv = vin;
    for (int p = 0; p < np; p++) {
        const bool last_to_out = (np - 1 - p) % 2 == 0;
        K *dk = last_to_out ? kout : (K *)w.rtk;
        uint32_t *dv = vin ? (last_to_out ? vout : (uint32_t *)w.rtv) : nullptr;
        const bool segpass = p >= kp;
        const int shift = segpass ? 8 * (p - kp) : begin_bit + 8 * p;
        const int32_t *sg = segpass ? seg : nullptr;
        k_rs_hist<K><<<ntiles, RS_WG, 0, st>>>(sk, sv, sg, n, shift, ntiles, (int32_t *)w.rhist);
        launch_scan_i32_i64(st, (const int32_t *)w.rhist, (int64_t *)w.roffs, nh, (int64_t *)w.rtotal, w.rscan);
        k_rs_scatter<K><<<ntiles, RS_WG, 0, st>>>(sk, sv, dk, dv, sg, n, shift, ntiles, (const int64_t *)w.roffs);
        sk = dk;
        sv = dv;
    }
    return hipGetLastError();
}

// the flagged keys, in order, into out; their number into *nsel_dev
static hipError_t compact(hipStream_t st, SasWork &w, const uint64_t *key, const uint8_t *flag, int64_t n,
                          uint64_t *out, int64_t *nsel_dev)
{
    const int nt = (int)((n + CP_TILE - 1) / CP_TILE);
    if (grow(&w.rhist, &w.rhist_cap, (size_t)(nt ? nt : 1) * sizeof(int32_t)) ||
        grow(&w.roffs, &w.roffs_cap, (size_t)(nt ? nt : 1) * sizeof(int64_t)) ||
        grow(&w.rscan, &w.rscan_cap, scan_scratch_bytes(nt)))
        return hipErrorOutOfMemory;
    if (nt) k_cp_count<<<nt, RS_WG, 0, st>>>(flag, n, (int32_t *)w.rhist);
    launch_scan_i32_i64(st, (const int32_t *)w.rhist, (int64_t *)w.roffs, nt, nsel_dev, w.rscan);
    if (nt) k_cp_scatter<<<nt, RS_WG, 0, st>>>(key, flag, n, (const int64_t *)w.roffs, out);
    return hipGetLastError();
}

hipError_t sas_candidates(hipStream_t st, SasWork &w, PlDev pl, SysDev sy, int npl, int nsys,
                          const int32_t *active, double dt, int64_t *ncand_dev, void *scan_scratch)
{
    if (w.n < npl) {
        for (void *p : {(void *)w.lokey, (void *)w.slokey, (void *)w.hi, (void *)w.slo, (void *)w.idx,
                        (void *)w.sidx, (void *)w.rank, (void *)w.cnt, (void *)w.coff, (void *)w.srt,
                        (void *)w.stiny})
            if (p) (void)hipFree(p);
        const size_t n = (size_t)npl + 1;
        SASCHK(hipMalloc(&w.lokey, n * sizeof(uint64_t)));
        SASCHK(hipMalloc(&w.slokey, n * sizeof(uint64_t)));
        SASCHK(hipMalloc(&w.hi, n * sizeof(double)));
        SASCHK(hipMalloc(&w.slo, n * sizeof(double)));
        SASCHK(hipMalloc(&w.idx, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.sidx, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.rank, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.cnt, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.coff, n * sizeof(int64_t)));
        SASCHK(hipMalloc(&w.srt, 7 * n * sizeof(double)));
        SASCHK(hipMalloc(&w.stiny, n));
        w.n = npl;
    }
    k_sas_extents<<<cdiv64(npl, 256), 256, 0, st>>>(pl, npl, dt, w.lokey, w.hi, w.idx);
    // each system's bodies by lower extent: the keys' 64 bits, then the system
    int sb = 0;
    while ((1ll << sb) < nsys) sb++;
    if (hipError_t e = radix_sort<uint64_t>(st, w, w.lokey, (const uint32_t *)w.idx, w.slokey, (uint32_t *)w.sidx, npl,
                                            0, 64, nsys > 1 ? pl.sys : nullptr, sb))
        return e;
    k_sas_unkey<<<cdiv64(npl, 256), 256, 0, st>>>(w.slokey, npl, w.slo);
    k_sas_rank<<<cdiv64(npl, 256), 256, 0, st>>>(w.sidx, npl, w.rank);
    k_sas_permute<<<cdiv64(npl, 256), 256, 0, st>>>(pl, w.sidx, npl, w.srt, w.stiny);
    k_sas_count<<<cdiv64(npl, 256), 256, 0, st>>>(w.slo, w.hi, w.rank, pl.sys, sy.pl_off, active, npl, w.cnt);
    launch_scan_i32_i64(st, w.cnt, w.coff, npl, ncand_dev, scan_scratch);
    return hipGetLastError();
}

hipError_t sas_narrow(hipStream_t st, SasWork &w, PlDev pl, int npl, int64_t ncand, double dt, int64_t *nsel_host)
{
    if (grow((void **)&w.key, &w.key_cap, (size_t)ncand * 8) || grow((void **)&w.flag, &w.flag_cap, (size_t)ncand) ||
        grow((void **)&w.sel, &w.sel_cap, (size_t)ncand * 8))
        return hipErrorOutOfMemory;
    if (!w.nsel) SASCHK(hipMalloc(&w.nsel, sizeof(int64_t)));
    if (ncand && npl)
        k_sas_narrow<<<cdiv64(npl, 4), 256, 0, st>>>(pl, w.coff, w.cnt, w.rank, w.sidx, w.srt, w.stiny, npl, dt, w.key,
                                                     w.flag);
    if (hipError_t e = compact(st, w, w.key, w.flag, ncand, w.sel, w.nsel)) return e;
    SASCHK(hipMemcpyAsync(nsel_host, w.nsel, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    SASCHK(hipStreamSynchronize(st));
    return hipGetLastError();
}

hipError_t sas_fill(hipStream_t st, SasWork &w, ListDev &pp, PlDev pl, int npl, int64_t npp)
{
    if (grow((void **)&w.skey, &w.skey_cap, (size_t)(npp ? npp : 1) * 8) ||
        grow((void **)&w.pid, &w.pid_cap, (size_t)(npp ? npp : 1) * 4) ||
        grow((void **)&w.spid, &w.spid_cap, (size_t)(npp ? npp : 1) * 4) ||
        grow((void **)&w.si2, &w.si2_cap, (size_t)(npp ? npp : 1) * 4))
        return hipErrorOutOfMemory;
    // keys are index1 << 32 | index2 with both below npl: only the low bits of
    // each half vary
    int kb = 1;
    while ((1ll << kb) < npl) kb++;
    if (hipError_t e = radix_sort<uint64_t>(st, w, w.sel, nullptr, w.skey, nullptr, npp, 0, 32 + kb, nullptr, 0))
        return e;
    k_sas_rows<<<cdiv64(npl, 256), 256, 0, st>>>(w.skey, npp, npl, pp.rowcnt, pp.rowstart);
    if (npp) k_sas_fill<<<cdiv64(npp, 256), 256, 0, st>>>(pp, w.skey, npp, pl.levelg, pl.levelm, w.pid);
    return hipGetLastError();
}

hipError_t sas_adjacency(hipStream_t st, SasWork &w, ListDev &pp, int npl, int64_t npp)
{
    if (!npp) return hipSuccess;
    int kb = 1;
    while ((1ll << kb) < npl) kb++;
    if (hipError_t e = radix_sort<uint32_t>(st, w, (const uint32_t *)pp.i2, (const uint32_t *)w.pid,
                                            (uint32_t *)w.si2, (uint32_t *)pp.col, npp, 0, kb, nullptr, 0))
        return e;
    k_sas_cols<<<cdiv64(npl, 256), 256, 0, st>>>(w.si2, npp, npl, pp.colcnt, pp.colstart);
    return hipGetLastError();
}

}  // namespace swg
