// k_enc.hip -- close-encounter detection and the SyMBA recursion kernels.
//
// Level-0 detection (encounter_check.f90:436-570) is a brute-force sweep of
// every pair with the exact narrow-phase test (encounter_check_one :573-621).
// A wavefront owns 64 candidate partners; for each body of the row block it
// evaluates the 64 tests and __ballot() packs them into one 64-bit word of a
// per-system bit matrix.  Rows are then counted (popcount + a wave scan that
// also records each word's prefix), scanned, and filled, which yields the
// reference's list order (index1 ascending, then index2) with no sort.  The
// same bit matrix gives every body its index2-side adjacency in list order.
//
// The level-0 substep of the recursion runs over the whole lists: list checks
// (symba_encounter_check_list_*, symba_encounter_check.f90:90-235),
// collision detection (collision_check.f90:16-230) and the level bookkeeping
// of symba_step.f90:118-160; its shell kicks are in k_level.hip.
#include "swg_kernels.h"
#include "swg_math.h"

namespace swg {

static inline int cdiv(int64_t a, int b) { return (int)((a + b - 1) / b); }

// ---------------------------------------------------------------------------
// pl-pl pair tiles: tile (s, ib, jb), jb >= ib, one wavefront each
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_pp_tiles(PlDev pl, SysDev sy, ListDev pp, const int4 *tiles, double dt)
{
    __shared__ double si[7][64];
    const int4 t = tiles[blockIdx.x];
    const int s = t.x, ib = t.y, jb = t.z;
    const int p0 = sy.pl_off[s], n = sy.pl_off[s + 1] - p0, nb = pp.nb[s];
    // with semi-interacting bodies only pairs whose index1 is among the
    // first nplm are checked (encounter_check_all_plplm,
    // encounter_check.f90:42-109)
    const int nm = sy.nplm ? sy.nplm[s] : n;
    const int lane = threadIdx.x;
    const int il0 = ib * 64 + lane;
    if (il0 < n) {
        int g = p0 + il0;
        si[0][lane] = pl.rh.x[g];
        si[1][lane] = pl.rh.y[g];
        si[2][lane] = pl.rh.z[g];
        si[3][lane] = pl.vb.x[g];
        si[4][lane] = pl.vb.y[g];
        si[5][lane] = pl.vb.z[g];
        si[6][lane] = pl.renc[g];
    }
    const int jl = jb * 64 + lane;
    const bool vj = jl < n;
    double xj = 0, yj = 0, zj = 0, vxj = 0, vyj = 0, vzj = 0, rj = 0;
    if (vj) {
        int g = p0 + jl;
        xj = pl.rh.x[g];
        yj = pl.rh.y[g];
        zj = pl.rh.z[g];
        vxj = pl.vb.x[g];
        vyj = pl.vb.y[g];
        vzj = pl.vb.z[g];
        rj = pl.renc[g];
    }
    __syncthreads();
    uint64_t *wrow = pp.words + pp.woff[s] + jb;
    const int imax = min(64, n - ib * 64);
    for (int q = 0; q < imax; q++) {
        const int il = ib * 64 + q;
        bool f = false;
        if (vj && jl > il && il < nm) {
            bool lv;
            f = encounter_one(xj - si[0][q], yj - si[1][q], zj - si[2][q], vxj - si[3][q], vyj - si[4][q],
                              vzj - si[5][q], si[6][q] + rj, dt, lv);
        }
        uint64_t m = __ballot(f);
        if (lane == 0) wrow[(int64_t)il * nb] = m;
    }
}

// ---------------------------------------------------------------------------
// pl-tp pair tiles: tile (s, tb), one wavefront = 64 particles, all pl rows
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_pt_tiles(PlDev pl, TpDev tp, SysDev sy, ListDev pt, const int4 *tiles,
                                                 double dt)
{
    __shared__ double si[7][64];
    const int4 t = tiles[blockIdx.x];
    const int s = t.x, tb = t.y;
    const int p0 = sy.pl_off[s], npl = sy.pl_off[s + 1] - p0;
    const int t0 = sy.tp_off[s], ntp = sy.tp_off[s + 1] - t0;
    const int nb = pt.nb[s];
    const int lane = threadIdx.x;
    const int jl = tb * 64 + lane;
    bool vj = jl < ntp;
    double xj = 0, yj = 0, zj = 0, vxj = 0, vyj = 0, vzj = 0;
    if (vj) {
        int g = t0 + jl;
        vj = tp.status[g] == ACTIVE;
        xj = tp.rh.x[g];
        yj = tp.rh.y[g];
        zj = tp.rh.z[g];
        vxj = tp.vb.x[g];
        vyj = tp.vb.y[g];
        vzj = tp.vb.z[g];
    }
    uint64_t *wcol = pt.words + pt.woff[s] + tb;
    for (int ib = 0; ib < npl; ib += 64) {
        __syncthreads();
        if (ib + lane < npl) {
            int g = p0 + ib + lane;
            si[0][lane] = pl.rh.x[g];
            si[1][lane] = pl.rh.y[g];
            si[2][lane] = pl.rh.z[g];
            si[3][lane] = pl.vb.x[g];
            si[4][lane] = pl.vb.y[g];
            si[5][lane] = pl.vb.z[g];
            si[6][lane] = pl.renc[g];
        }
        __syncthreads();
        const int qn = min(64, npl - ib);
        for (int q = 0; q < qn; q++) {
            bool f = false;
            if (vj) {
                bool lv;
                f = encounter_one(xj - si[0][q], yj - si[1][q], zj - si[2][q], vxj - si[3][q], vyj - si[4][q],
                                  vzj - si[5][q], si[6][q] + 0.0, dt, lv);
            }
            uint64_t m = __ballot(f);
            if (lane == 0) wcol[(int64_t)(ib + q) * nb] = m;
        }
    }
}

// ---------------------------------------------------------------------------
// row counts: one wavefront per row; records each word's exclusive prefix
// ---------------------------------------------------------------------------
__device__ inline int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & 63;
    for (int o = 1; o < 64; o <<= 1) {
        int u = __shfl_up(v, o, 64);
        if (lane >= o) v += u;
    }
    return v;
}

__global__ __launch_bounds__(64) void k_rows_count(ListDev L, const int32_t *row_sys, int upper,
                                                   const int32_t *pl_off, const int32_t *run, int nrows)
{
    const int row = blockIdx.x;
    if (row >= nrows) return;
    const int s = row_sys[row];
    const int lane = threadIdx.x;
    if (run && !run[s]) {
        if (lane == 0) L.rowcnt[row] = 0;
        return;
    }
    const int il = row - pl_off[s];
    const int nb = L.nb[s];
    const int b0 = upper ? il / 64 : 0;
    const int64_t base = L.woff[s] + (int64_t)il * nb;
    int running = 0;
    for (int c = b0; c < nb; c += 64) {
        int b = c + lane;
        int pc = 0;
        if (b < nb) pc = __popcll(L.words[base + b]);
        int incl = wave_incl_scan(pc);
        if (b < nb) L.wpre[base + b] = running + incl - pc;
        running += __shfl(incl, 63, 64);
    }
    if (lane == 0) L.rowcnt[row] = running;
}

// ---------------------------------------------------------------------------
// exclusive scan int32 -> int64 (reduce, scan partials, apply)
// ---------------------------------------------------------------------------
constexpr int SCAN_T = 256, SCAN_ITEMS = 4, SCAN_CHUNK = SCAN_T * SCAN_ITEMS;

__global__ __launch_bounds__(SCAN_T) void k_scan_reduce(const int32_t *in, int n, int64_t *part)
{
    __shared__ int64_t sh[SCAN_T];
    int64_t v = 0;
    int base = blockIdx.x * SCAN_CHUNK;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k * SCAN_T + threadIdx.x;
        if (i < n) v += in[i];
    }
    sh[threadIdx.x] = v;
    __syncthreads();
    for (int o = SCAN_T / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ __launch_bounds__(SCAN_T) void k_scan_partials(int64_t *part, int np, int64_t *total)
{
    __shared__ int64_t sh[SCAN_T];
    int64_t carry = 0;
    for (int c = 0; c < np; c += SCAN_T) {
        int i = c + threadIdx.x;
        int64_t v = i < np ? part[i] : 0;
        sh[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < SCAN_T; o <<= 1) {
            int64_t u = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
            __syncthreads();
            sh[threadIdx.x] += u;
            __syncthreads();
        }
        if (i < np) part[i] = carry + sh[threadIdx.x] - v;
        carry += sh[SCAN_T - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0 && total) *total = carry;
}

__global__ __launch_bounds__(SCAN_T) void k_scan_apply(const int32_t *in, int n, const int64_t *part, int64_t *out)
{
    __shared__ int64_t sh[SCAN_T];
    int base = blockIdx.x * SCAN_CHUNK + threadIdx.x * SCAN_ITEMS;
    int64_t loc[SCAN_ITEMS];
    int64_t sum = 0;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k;
        loc[k] = sum;
        if (i < n) sum += in[i];
    }
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (int o = 1; o < SCAN_T; o <<= 1) {
        int64_t u = (int)threadIdx.x >= o ? sh[threadIdx.x - o] : 0;
        __syncthreads();
        sh[threadIdx.x] += u;
        __syncthreads();
    }
    int64_t off = part[blockIdx.x] + sh[threadIdx.x] - sum;
    for (int k = 0; k < SCAN_ITEMS; k++) {
        int i = base + k;
        if (i < n) out[i] = off + loc[k];
    }
}

size_t scan_scratch_bytes(int n) { return sizeof(int64_t) * (size_t)(cdiv(n > 0 ? n : 1, SCAN_CHUNK) + 1); }

// Up to four short exclusive scans in one launch, one workgroup each; block
// 0 also clears an int32 array (the counters the next kernel accumulates).
constexpr int SMALL_T = 1024, SMALL_ITEMS = SCAN_SMALL_MAX / SMALL_T;
static_assert(SCAN_SMALL_MAX % SMALL_T == 0, "items per thread");

__global__ __launch_bounds__(SMALL_T) void k_scan_small(ScanSmall a)
{
    __shared__ int64_t sh[SMALL_T];
    const int b = blockIdx.x, tid = threadIdx.x;
    if (b == 0 && a.zero)
        for (int i = tid; i < a.nzero; i += SMALL_T) a.zero[i] = 0;
    const int n = a.n[b];
    const int32_t *in = a.in[b];
    const int base = tid * SMALL_ITEMS;
    int64_t loc[SMALL_ITEMS];
    int64_t sum = 0;
#pragma unroll
    for (int k = 0; k < SMALL_ITEMS; k++) {
        loc[k] = sum;
        if (base + k < n) sum += in[base + k];
    }
    sh[tid] = sum;
    __syncthreads();
    for (int o = 1; o < SMALL_T; o <<= 1) {
        const int64_t u = tid >= o ? sh[tid - o] : 0;
        __syncthreads();
        sh[tid] += u;
        __syncthreads();
    }
    const int64_t off = sh[tid] - sum;
    int64_t *out = a.out[b];
#pragma unroll
    for (int k = 0; k < SMALL_ITEMS; k++)
        if (base + k < n) out[base + k] = off + loc[k];
    if (tid == SMALL_T - 1) {
        if (a.total[b]) *a.total[b] = sh[tid];
        if (a.total2[b]) *a.total2[b] = sh[tid];
    }
}

void launch_scan_small(hipStream_t st, const ScanSmall &a, int na)
{
    k_scan_small<<<na, SMALL_T, 0, st>>>(a);
}

void launch_scan_i32_i64(hipStream_t st, const int32_t *in, int64_t *out, int n, int64_t *total, void *scratch)
{
    int64_t *part = (int64_t *)scratch;
    int nb = cdiv(n > 0 ? n : 1, SCAN_CHUNK);
    k_scan_reduce<<<nb, SCAN_T, 0, st>>>(in, n, part);
    k_scan_partials<<<1, SCAN_T, 0, st>>>(part, nb, total);
    if (n) k_scan_apply<<<nb, SCAN_T, 0, st>>>(in, n, part, out);
}

// ---------------------------------------------------------------------------
// ordered fill: one wavefront per row, each lane emits the bits of one word
// ---------------------------------------------------------------------------
// Also marks both bodies of every listed pair as being at recursion level 0
// (symba_encounter_check.f90:74-79, :286-290).
__global__ __launch_bounds__(64) void k_rows_fill(ListDev L, const int32_t *row_sys, int upper,
                                                  const int32_t *pl_off, const int32_t *col_off, int nrows,
                                                  int32_t *lev1g, int32_t *lev1m, int32_t *lev2g, int32_t *lev2m)
{
    const int row = blockIdx.x;
    if (row >= nrows) return;
    const int cnt = L.rowcnt[row];
    if (cnt == 0) return;
    const int s = row_sys[row];
    const int il = row - pl_off[s];
    const int nb = L.nb[s];
    const int b0 = upper ? il / 64 : 0;
    const int64_t base = L.woff[s] + (int64_t)il * nb;
    const int64_t start = L.rowstart[row];
    const int c0 = col_off[s];
    for (int b = b0 + (int)threadIdx.x; b < nb; b += 64) {
        uint64_t w = L.words[base + b];
        int64_t pos = start + L.wpre[base + b];
        while (w) {
            int bit = __ffsll((unsigned long long)w) - 1;
            w &= w - 1;
            const int col = c0 + b * 64 + bit;
            L.i1[pos] = row;
            L.i2[pos] = col;
            if (lev2g) {
                lev2g[col] = 0;
                lev2m[col] = 0;
            }
            L.lvdotr[pos] = 1;
            L.status[pos] = (int8_t)ACTIVE;
            L.level[pos] = 0;
            pos++;
        }
    }
    if (lev1g && threadIdx.x == 0) {
        lev1g[row] = 0;
        lev1m[row] = 0;
    }
}

// per-system "any pair listed" flags
__global__ void k_sys_flags(ListDev L, const int32_t *pl_off, int32_t *flag, int nsys)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsys) return;
    int p0 = pl_off[s], p1 = pl_off[s + 1];
    int64_t n = 0;
    if (p1 > p0) n = L.rowstart[p1 - 1] + L.rowcnt[p1 - 1] - L.rowstart[p0];
    flag[s] = n > 0;
}

// index2-side adjacency: for column c (a pl body for pl-pl, a particle for
// pl-tp), the pairs (p, c) in list order, i.e. rows p ascending.  One
// wavefront per 64-column block of a system (cb_off: each system's first
// block): lane l holds row p0 + l's word of the block, and 64 ballots
// transpose the 64 x 64 bit tile, so that lane l gets the rows of column
// block*64 + l with a set bit, 64 rows at a time.  Chunks whose words are all
// zero are skipped.  With upper (pl-pl) only rows p < column count.
template <bool FILL>
__global__ __launch_bounds__(64) void k_col_t(ListDev L, const int32_t *row_off, const int32_t *col_off,
                                              const int32_t *cb_off, int nsys, const int32_t *run, int upper)
{
    const int blk = blockIdx.x, lane = threadIdx.x;
    int lo = 0, hi = nsys;  // system: last s with cb_off[s] <= blk
    while (hi - lo > 1) {
        const int m = (lo + hi) >> 1;
        if (cb_off[m] <= blk) lo = m;
        else hi = m;
    }
    const int s = lo, cb = blk - cb_off[s];
    const int ncol = col_off[s + 1] - col_off[s], nrow_s = row_off[s + 1] - row_off[s];
    const int cl = cb * 64 + lane;
    const bool mine = cl < ncol;
    const int c = col_off[s] + cl;
    if (run && !run[s]) {
        if (!FILL && mine) L.colcnt[c] = 0;
        return;
    }
    const int nrow = upper ? min(nrow_s, cb * 64 + 63) : nrow_s;
    const int nb = L.nb[s];
    const int64_t wbase = L.woff[s] + cb;
    const int bit = lane;
    const uint64_t lowm = (bit == 0) ? 0ull : ((~0ull) >> (64 - bit));
    int cnt = 0;
    int64_t o = (FILL && mine) ? L.colstart[c] : 0;
    for (int p0 = 0; p0 < nrow; p0 += 64) {
        const int p = p0 + lane;
        const uint64_t w = p < nrow ? L.words[wbase + (int64_t)p * nb] : 0ull;
        if (__ballot(w != 0ull) == 0ull) continue;
        uint64_t mycol = 0ull;
        for (int b = 0; b < 64; b++) {
            const uint64_t m = __ballot((w >> b) & 1ull);
            if (b == lane) mycol = m;
        }
        if (!mine) continue;
        if (upper) {
            const int lim = cl - p0;  // rows p0 + r < cl
            mycol &= lim >= 64 ? ~0ull : (lim <= 0 ? 0ull : ((~0ull) >> (64 - lim)));
        }
        if (!FILL) {
            cnt += __popcll(mycol);
        } else {
            while (mycol) {
                const int r = __builtin_ctzll(mycol);
                mycol &= mycol - 1;
                const int64_t widx = wbase + (int64_t)(p0 + r) * nb;
                const uint64_t wr = L.words[widx];
                const int64_t k = L.rowstart[row_off[s] + p0 + r] + L.wpre[widx] + __popcll(wr & lowm);
                L.col[o++] = (int32_t)k;
            }
        }
    }
    if (!FILL && mine) L.colcnt[c] = cnt;
}

// ---------------------------------------------------------------------------
// SyMBA recursion
// ---------------------------------------------------------------------------

// symba_encounter_check_list_plpl (symba_encounter_check.f90:90-160)
__global__ void k_list_check_pp(ListDev pp, PlDev pl, SysDev sy, int64_t n, int irec, double dt, int32_t *any)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (pp.status[k] != ACTIVE || pp.level[k] != irec - 1) return;
    const int i = pp.i1[k], j = pp.i2[k];
    double xr = pl.rh.x[j] - pl.rh.x[i], yr = pl.rh.y[j] - pl.rh.y[i], zr = pl.rh.z[j] - pl.rh.z[i];
    double vxr = pl.vb.x[j] - pl.vb.x[i], vyr = pl.vb.y[j] - pl.vb.y[i], vzr = pl.vb.z[j] - pl.vb.z[i];
    const double rsh = shell_r(irec);
    double rci = (pl.rhill[i] * RHSCALE) * rsh, rcj = (pl.rhill[j] * RHSCALE) * rsh;
    bool lv;
    bool le = encounter_one(xr, yr, zr, vxr, vyr, vzr, rci + rcj, dt, lv);
    pp.lvdotr[k] = lv;
    if (le) {
        double rji2 = dot3(xr, yr, zr, xr, yr, zr);
        double rs = pl.R[i] + pl.R[j];
        le = rji2 > rs * rs;
    }
    if (le) {
        pp.level[k] = (int8_t)irec;
        pl.levelg[i] = irec;
        pl.levelg[j] = irec;
        atomicMax(&pl.levelm[i], irec);
        atomicMax(&pl.levelm[j], irec);
        sy.flag_pp[pl.sys[i]] = 1;
        *any = 1;
    }
}

// symba_encounter_check_list_pltp (symba_encounter_check.f90:163-235); with
// short_circuit the check is skipped in systems whose pl-pl check of this
// substep found an encounter (gfortran evaluates `a .or. b` left to right and
// stops at the first .true.).
__global__ void k_list_check_pt(ListDev pt, PlDev pl, TpDev tp, SysDev sy, int64_t n, int irec, double dt,
                                int short_circuit, int32_t *any)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    if (pt.status[k] != ACTIVE || pt.level[k] != irec - 1) return;
    const int i = pt.i1[k], j = pt.i2[k];
    const int s = pl.sys[i];
    if (short_circuit && sy.flag_pp[s]) return;
    double xr = tp.rh.x[j] - pl.rh.x[i], yr = tp.rh.y[j] - pl.rh.y[i], zr = tp.rh.z[j] - pl.rh.z[i];
    double vxr = tp.vb.x[j] - pl.vb.x[i], vyr = tp.vb.y[j] - pl.vb.y[i], vzr = tp.vb.z[j] - pl.vb.z[i];
    double rci = (pl.rhill[i] * RHSCALE) * shell_r(irec);
    bool lv;
    bool le = encounter_one(xr, yr, zr, vxr, vyr, vzr, rci, dt, lv);
    pt.lvdotr[k] = lv;
    if (le) {
        double rji2 = dot3(xr, yr, zr, xr, yr, zr);
        le = rji2 > pl.R[i] * pl.R[i];
    }
    if (le) {
        pt.level[k] = (int8_t)irec;
        pl.levelg[i] = irec;
        atomicMax(&pl.levelm[i], irec);
        tp.levelg[j] = irec;
        atomicMax(&tp.levelm[j], irec);
        sy.flag_pt[s] = 1;
        *any = 1;
    }
}

// collision_check_plpl / _pltp detection (collision_check.f90:53-230)
__global__ void k_collision(ListDev pp, ListDev pt, PlDev pl, TpDev tp, SysDev sy, int64_t npp, int64_t npt,
                            int irec, double dt, int32_t *any)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < npp) {
        if (pp.status[k] != ACTIVE) return;
        int i = pp.i1[k], j = pp.i2[k];
        if (pl.levelg[i] < irec) return;
        bool c = collision_one(pl.rh.x[i] - pl.rh.x[j], pl.rh.y[i] - pl.rh.y[j], pl.rh.z[i] - pl.rh.z[j],
                               pl.vb.x[i] - pl.vb.x[j], pl.vb.y[i] - pl.vb.y[j], pl.vb.z[i] - pl.vb.z[j],
                               pl.Gm[i] + pl.Gm[j], pl.R[i] + pl.R[j], dt, pp.lvdotr[k] != 0);
        if (c) {
            pp.status[k] = (int8_t)COLLIDED;
            atomicOr(&sy.err[pl.sys[i]], 1);
            *any = 1;
        }
        return;
    }
    k -= npp;
    if (k >= npt) return;
    if (pt.status[k] != ACTIVE) return;
    int i = pt.i1[k], j = pt.i2[k];
    if (tp.levelg[j] < irec) return;
    bool c = collision_one(pl.rh.x[i] - tp.rh.x[j], pl.rh.y[i] - tp.rh.y[j], pl.rh.z[i] - tp.rh.z[j],
                           pl.vb.x[i] - tp.vb.x[j], pl.vb.y[i] - tp.vb.y[j], pl.vb.z[i] - tp.vb.z[j], pl.Gm[i],
                           pl.R[i], dt, pt.lvdotr[k] != 0);
    if (c) {
        pt.status[k] = (int8_t)COLLIDED;
        atomicOr(&sy.err[pl.sys[i]], 1);
        *any = 1;
    }
}

// symba_step_set_recur_levels_system (symba_step.f90:118-160)
__global__ void k_set_levels(PlDev pl, TpDev tp, ListDev pp, ListDev pt, int npl, int ntp, int64_t npp, int64_t npt,
                             int ireci)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int irecp = ireci + 1;
    if (k < npl && pl.levelg[k] == irecp) pl.levelg[k] = ireci;
    if (k < ntp && tp.levelg[k] == irecp) tp.levelg[k] = ireci;
    if (k < npp && pp.level[k] == irecp) pp.level[k] = (int8_t)ireci;
    if (k < npt && pt.level[k] == irecp) pt.level[k] = (int8_t)ireci;
}

__global__ void k_clear_flags(SysDev sy, int nsys, int32_t *any)
{
    int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsys) {
        sy.flag_pp[s] = 0;
        sy.flag_pt[s] = 0;
    }
    if (s == 0 && any) *any = 0;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
void launch_pp_tiles(hipStream_t st, PlDev pl, SysDev sy, ListDev pp, const int4 *tiles, int ntiles, double dt)
{
    if (ntiles) k_pp_tiles<<<ntiles, 64, 0, st>>>(pl, sy, pp, tiles, dt);
}
void launch_pt_tiles(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, ListDev pt, const int4 *tiles, int ntiles,
                     double dt)
{
    if (ntiles) k_pt_tiles<<<ntiles, 64, 0, st>>>(pl, tp, sy, pt, tiles, dt);
}
void launch_rows_count(hipStream_t st, ListDev L, const int32_t *row_sys, int upper,
                       const int32_t *pl_off, const int32_t *run, int nrows)
{
    if (nrows) k_rows_count<<<nrows, 64, 0, st>>>(L, row_sys, upper, pl_off, run, nrows);
}
void launch_rows_fill(hipStream_t st, ListDev L, const int32_t *row_sys, int upper, const int32_t *pl_off,
                      const int32_t *col_off, int nrows, int32_t *lev1g, int32_t *lev1m, int32_t *lev2g,
                      int32_t *lev2m)
{
    if (nrows)
        k_rows_fill<<<nrows, 64, 0, st>>>(L, row_sys, upper, pl_off, col_off, nrows, lev1g, lev1m, lev2g, lev2m);
}
void launch_sys_flags(hipStream_t st, ListDev L, const int32_t *pl_off, int32_t *flag, int nsys)
{
    k_sys_flags<<<cdiv(nsys, 256), 256, 0, st>>>(L, pl_off, flag, nsys);
}
void launch_col_t(hipStream_t st, ListDev L, const int32_t *row_off, const int32_t *col_off, const int32_t *cb_off,
                  int nsys, int nblocks, const int32_t *run, int upper, int fill)
{
    if (!nblocks) return;
    if (fill) k_col_t<true><<<nblocks, 64, 0, st>>>(L, row_off, col_off, cb_off, nsys, run, upper);
    else k_col_t<false><<<nblocks, 64, 0, st>>>(L, row_off, col_off, cb_off, nsys, run, upper);
}
// The same index2-side adjacency of a short pl-pl list in one workgroup
// (npp <= COLS_MAX, npl <= COLS_NPL): the counts per column, their exclusive
// scan, and each column's entries in list order, placed by their rank among
// the earlier entries of the column.  Two launches over every 64-column block
// and a device-wide scan cost more than this when a few hundred pairs are
// listed among a few thousand bodies.
constexpr int COLS_T = 1024;
__global__ __launch_bounds__(COLS_T) void k_col_small(ListDev L, int npl, int npp)
{
    __shared__ int32_t s2[COLS_MAX];
    __shared__ int32_t cnt[COLS_NPL];  // counts, then each column's first slot
    __shared__ int32_t part[COLS_T];
    const int tid = threadIdx.x;
    for (int i = tid; i < npl; i += COLS_T) cnt[i] = 0;
    for (int k = tid; k < npp; k += COLS_T) s2[k] = L.i2[k];
    __syncthreads();
    for (int k = tid; k < npp; k += COLS_T) atomicAdd(&cnt[s2[k]], 1);
    __syncthreads();
    // exclusive scan: each thread a contiguous run of columns
    const int per = (npl + COLS_T - 1) / COLS_T, a = min(npl, tid * per), b = min(npl, a + per);
    int32_t s = 0;
    for (int i = a; i < b; i++) s += cnt[i];
    part[tid] = s;
    __syncthreads();
    if (tid == 0) {
        int32_t r = 0;
        for (int q = 0; q < COLS_T; q++) {
            const int32_t v = part[q];
            part[q] = r;
            r += v;
        }
    }
    __syncthreads();
    int32_t r = part[tid];
    for (int i = a; i < b; i++) {
        const int32_t c = cnt[i];
        L.colcnt[i] = c;
        L.colstart[i] = r;
        cnt[i] = r;
        r += c;
    }
    __syncthreads();
    for (int k = tid; k < npp; k += COLS_T) {
        const int c = s2[k];
        int rank = 0;
        for (int q = 0; q < k; q++) rank += s2[q] == c;
        L.col[cnt[c] + rank] = k;
    }
}

void launch_col_small(hipStream_t st, ListDev L, int npl, int npp)
{
    if (npl > 0) k_col_small<<<1, COLS_T, 0, st>>>(L, npl, npp);
}

void launch_list_check_pp(hipStream_t st, ListDev pp, PlDev pl, SysDev sy, int64_t n, int irec, double dt,
                          int32_t *any)
{
    if (n) k_list_check_pp<<<cdiv(n, 256), 256, 0, st>>>(pp, pl, sy, n, irec, dt, any);
}
void launch_list_check_pt(hipStream_t st, ListDev pt, PlDev pl, TpDev tp, SysDev sy, int64_t n, int irec, double dt,
                          int short_circuit, int32_t *any)
{
    if (n) k_list_check_pt<<<cdiv(n, 256), 256, 0, st>>>(pt, pl, tp, sy, n, irec, dt, short_circuit, any);
}
void launch_collision(hipStream_t st, ListDev pp, ListDev pt, PlDev pl, TpDev tp, SysDev sy, int64_t npp,
                      int64_t npt, int irec, double dt, int32_t *any)
{
    if (npp + npt) k_collision<<<cdiv(npp + npt, 256), 256, 0, st>>>(pp, pt, pl, tp, sy, npp, npt, irec, dt, any);
}
void launch_set_levels(hipStream_t st, PlDev pl, TpDev tp, ListDev pp, ListDev pt, int npl, int ntp, int64_t npp,
                       int64_t npt, int ireci)
{
    int64_t n = npl;
    if (ntp > n) n = ntp;
    if (npp > n) n = npp;
    if (npt > n) n = npt;
    if (n) k_set_levels<<<cdiv(n, 256), 256, 0, st>>>(pl, tp, pp, pt, npl, ntp, npp, npt, ireci);
}
void launch_clear_flags(hipStream_t st, SysDev sy, int nsys, int32_t *any)
{
    k_clear_flags<<<cdiv(nsys, 256), 256, 0, st>>>(sy, nsys, any);
}

}  // namespace swg
