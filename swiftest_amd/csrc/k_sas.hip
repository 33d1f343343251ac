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
#include <hipcub/hipcub.hpp>

#include "swg_kernels.h"
#include "swg_math.h"

namespace swg {

static inline int cdiv64(int64_t a, int b) { return (int)((a + b - 1) / b); }

// relative widening of the extents: covers the rounding of x + v*dt and the
// narrow phase's own rounding of r2min against R^2 many times over
constexpr double SAS_REL = 1e-12;

__global__ void k_sas_extents(PlDev pl, int npl, double dt, double *lo, double *hi, int32_t *idx)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= npl) return;
    const double x = pl.rh.x[g], x1 = x + pl.vb.x[g] * dt, r = pl.renc[g];
    const double m = SAS_REL * (fabs(x) + fabs(x1) + r) + 1e-300;
    lo[g] = fmin(x, x1) - r - m;
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

// narrow phase of candidate c (the pair test of k_pp_tiles with index1 < index2)
__global__ void k_sas_narrow(PlDev pl, const int64_t *coff, const int32_t *rank, const int32_t *sidx, int npl,
                             int64_t ncand, double dt, uint64_t *key, uint8_t *flag)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncand) return;
    int lo = 0, hi = npl;  // owner: last g with coff[g] <= c
    while (hi - lo > 1) {
        int m = (lo + hi) >> 1;
        if (coff[m] <= c) lo = m;
        else hi = m;
    }
    const int g = lo;
    const int j = sidx[rank[g] + 1 + (int)(c - coff[g])];
    const int a = min(g, j), b = max(g, j);
    bool lv;
    // a semi-interacting index1 (GMTINY): the pair is not checked
    const bool f = !(pl.tiny && pl.tiny[a]) && encounter_one(pl.rh.x[b] - pl.rh.x[a], pl.rh.y[b] - pl.rh.y[a], pl.rh.z[b] - pl.rh.z[a],
                                 pl.vb.x[b] - pl.vb.x[a], pl.vb.y[b] - pl.vb.y[a], pl.vb.z[b] - pl.vb.z[a],
                                 pl.renc[a] + pl.renc[b], dt, lv);
    key[c] = ((uint64_t)(uint32_t)a << 32) | (uint32_t)b;
    flag[c] = f;
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
    for (void *p : {(void *)lo, (void *)hi, (void *)slo, (void *)idx, (void *)sidx, (void *)rank, (void *)cnt,
                    (void *)coff, (void *)key, (void *)flag, (void *)sel, (void *)skey, (void *)nsel, temp,
                    (void *)pid, (void *)spid, (void *)si2})
        if (p) (void)hipFree(p);
}

hipError_t sas_candidates(hipStream_t st, SasWork &w, PlDev pl, SysDev sy, int npl, int nsys,
                          const int32_t *active, double dt, int64_t *ncand_dev, void *scan_scratch)
{
    if (w.n < npl) {
        for (void *p : {(void *)w.lo, (void *)w.hi, (void *)w.slo, (void *)w.idx, (void *)w.sidx, (void *)w.rank,
                        (void *)w.cnt, (void *)w.coff})
            if (p) (void)hipFree(p);
        const size_t n = (size_t)npl + 1;
        SASCHK(hipMalloc(&w.lo, n * sizeof(double)));
        SASCHK(hipMalloc(&w.hi, n * sizeof(double)));
        SASCHK(hipMalloc(&w.slo, n * sizeof(double)));
        SASCHK(hipMalloc(&w.idx, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.sidx, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.rank, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.cnt, n * sizeof(int32_t)));
        SASCHK(hipMalloc(&w.coff, n * sizeof(int64_t)));
        w.n = npl;
    }
    k_sas_extents<<<cdiv64(npl, 256), 256, 0, st>>>(pl, npl, dt, w.lo, w.hi, w.idx);
    size_t bytes = 0;
    SASCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(nullptr, bytes, w.lo, w.slo, w.idx, w.sidx, npl, nsys,
                                                       sy.pl_off, sy.pl_off + 1, 0, 64, st));
    if (grow(&w.temp, &w.temp_bytes, bytes)) return hipErrorOutOfMemory;
    bytes = w.temp_bytes;
    SASCHK(hipcub::DeviceSegmentedRadixSort::SortPairs(w.temp, bytes, w.lo, w.slo, w.idx, w.sidx, npl, nsys,
                                                       sy.pl_off, sy.pl_off + 1, 0, 64, st));
    k_sas_rank<<<cdiv64(npl, 256), 256, 0, st>>>(w.sidx, npl, w.rank);
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
    if (ncand) k_sas_narrow<<<cdiv64(ncand, 256), 256, 0, st>>>(pl, w.coff, w.rank, w.sidx, npl, ncand, dt, w.key, w.flag);
    size_t bytes = 0;
    SASCHK(hipcub::DeviceSelect::Flagged(nullptr, bytes, w.key, w.flag, w.sel, w.nsel, ncand, st));
    if (grow(&w.temp, &w.temp_bytes, bytes)) return hipErrorOutOfMemory;
    bytes = w.temp_bytes;
    SASCHK(hipcub::DeviceSelect::Flagged(w.temp, bytes, w.key, w.flag, w.sel, w.nsel, ncand, st));
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
    const int endbit = 32 + kb;
    size_t bytes = 0;
    SASCHK(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, w.sel, w.skey, npp, 0, endbit, st));
    if (grow(&w.temp, &w.temp_bytes, bytes)) return hipErrorOutOfMemory;
    bytes = w.temp_bytes;
    SASCHK(hipcub::DeviceRadixSort::SortKeys(w.temp, bytes, w.sel, w.skey, npp, 0, endbit, st));
    k_sas_rows<<<cdiv64(npl, 256), 256, 0, st>>>(w.skey, npp, npl, pp.rowcnt, pp.rowstart);
    if (npp) k_sas_fill<<<cdiv64(npp, 256), 256, 0, st>>>(pp, w.skey, npp, pl.levelg, pl.levelm, w.pid);
    return hipGetLastError();
}

hipError_t sas_adjacency(hipStream_t st, SasWork &w, ListDev &pp, int npl, int64_t npp)
{
    if (!npp) return hipSuccess;
    size_t bytes = 0;
    SASCHK(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, pp.i2, w.si2, w.pid, pp.col, npp, 0, 32, st));
    if (grow(&w.temp, &w.temp_bytes, bytes)) return hipErrorOutOfMemory;
    bytes = w.temp_bytes;
    SASCHK(hipcub::DeviceRadixSort::SortPairs(w.temp, bytes, pp.i2, w.si2, w.pid, pp.col, npp, 0, 32, st));
    k_sas_cols<<<cdiv64(npl, 256), 256, 0, st>>>(w.si2, npp, npl, pp.colcnt, pp.colstart);
    return hipGetLastError();
}

}  // namespace swg
