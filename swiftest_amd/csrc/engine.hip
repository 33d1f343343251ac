// engine.hip -- host side of the resident-system path: device allocation,
// the reference's initialize sequence, the step sequence of SyMBA / HELIO,
// the recursion driver, snapshots, and the C ABI of include/swiftest_hip.h.
//
// Step sequence (symba_step.f90:14-115 with the HELIO branch of
// helio_step.f90:37-119 folded in): the HELIO branch is the encounter branch
// with empty lists, except that it clears lfirst; each system keeps its own
// lfirst.  All systems of an ensemble descend the recursion together; a level
// on which a system has no listed pair is a no-op for that system, so the
// per-system result is the reference's.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <chrono>
#include <cstdlib>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/swiftest_hip.h"
#include "swg_engine.h"
#include "swg_kernels.h"
#include "swg_math.h"

namespace swg {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

int hipfail(hipError_t e, const char *what)
{
    g_err = std::string(what) + ": " + hipGetErrorString(e);
    return SWG_ERR_HIP;
}

static int hipfail_or_ok(hipError_t e) { return e == hipSuccess ? 0 : hipfail(e, "slot pool set-up"); }

#define HIPCHK(x)                                    \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return hipfail(e_, #x); \
    } while (0)

template <class T>
static hipError_t dalloc(T **p, size_t n)
{
    return hipMalloc((void **)p, sizeof(T) * (n > 0 ? n : 1));
}

// ---------------------------------------------------------------------------
// host arithmetic of the initialize / snapshot sequence
// ---------------------------------------------------------------------------

// semi-major axis of swiftest_orbel_xv2el (swiftest_orbel.f90:913-1072)
static double xv2a(double mu, const double *r, const double *v)
{
    double hx = r[1] * v[2] - r[2] * v[1], hy = r[2] * v[0] - r[0] * v[2], hz = r[0] * v[1] - r[1] * v[0];
    double h2 = dot3(hx, hy, hz, hx, hy, hz);
    if (h2 <= 10.0 * DBL_MIN) return 0.0;
    double rr = std::sqrt(dot3(r[0], r[1], r[2], r[0], r[1], r[2]));
    double v2 = dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
    double energy = 0.5 * v2 - mu / rr;
    if (std::fabs((energy * rr) / mu) < std::sqrt(TINYVALUE)) return (0.5 * h2) / mu;
    double a = (-0.5 * mu) / energy;
    if (a < 0.0) {
        double fac = -h2 / (mu * a);
        if (!(fac > TINYVALUE)) return (0.5 * h2) / mu;
    }
    return a;
}

// swiftest_util_set_rhill (swiftest_util.f90:2260-2280): mu = GM_cb + Gm
void host_set_rhill(double GMcb, int n, const double *rh, const double *vh, const double *Gm, double *rhill)
{
    if (GMcb <= DBL_MIN) return;
    for (int i = 0; i < n; i++) {
        double a = xv2a(GMcb + Gm[i], &rh[3 * i], &vh[3 * i]);
        double fac = std::pow((Gm[i] / GMcb) / 3.0, THIRD);
        if (a > 0.0)
            rhill[i] = a * fac;
        else
            rhill[i] = std::sqrt(dot3(rh[3 * i], rh[3 * i + 1], rh[3 * i + 2], rh[3 * i], rh[3 * i + 1], rh[3 * i + 2])) * fac;
    }
}

// ---------------------------------------------------------------------------
// Ensemble
// ---------------------------------------------------------------------------

static void free_list(ListDev &L)
{
    void *p[] = {L.words, L.wpre, L.woff, L.nb, L.rowcnt, L.rowstart, L.i1, L.i2, L.lvdotr, L.status, L.level,
                 L.colcnt, L.colstart, L.col};
    for (void *q : p)
        if (q) (void)hipFree(q);
    L = ListDev{};
}

Ensemble::~Ensemble()
{
    drop_graphs();
    for (void *p : owned)
        if (p) (void)hipFree(p);
    free_list(pp);
    free_list(pt);
    if (h_small) (void)hipHostFree(h_small);
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (ev_any) (void)hipEventDestroy(ev_any);
    for (hipEvent_t e : ev_tp) (void)hipEventDestroy(e);
    for (hipEvent_t e : ev_pl) (void)hipEventDestroy(e);
    if (d_arena) (void)hipFree(d_arena);
    for (void *q : lv_buf)
        if (q) (void)hipFree(q);
}

template <class T>
int Ensemble::own(T **p, size_t n)
{
    HIPCHK(dalloc(p, n));
    owned.push_back((void *)*p);
    return 0;
}

int Ensemble::alloc_vec(Vec3p &v, size_t n)
{
    if (own(&v.x, n) || own(&v.y, n) || own(&v.z, n)) return SWG_ERR_HIP;
    return 0;
}

static int upload_aos(hipStream_t st, Vec3p v, const double *aos, size_t n, std::vector<double> &tmp)
{
    tmp.resize(3 * (n ? n : 1));
    double *cx = tmp.data(), *cy = cx + n, *cz = cy + n;
    for (size_t i = 0; i < n; i++) {
        cx[i] = aos[3 * i];
        cy[i] = aos[3 * i + 1];
        cz[i] = aos[3 * i + 2];
    }
    HIPCHK(hipMemcpyAsync(v.x, cx, n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(v.y, cy, n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(v.z, cz, n * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

static int download_aos(hipStream_t st, Vec3p v, double *aos, size_t n)
{
    std::vector<double> tmp(3 * (n ? n : 1));
    double *cx = tmp.data(), *cy = cx + n, *cz = cy + n;
    HIPCHK(hipMemcpyAsync(cx, v.x, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cy, v.y, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(cz, v.z, n * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    for (size_t i = 0; i < n; i++) {
        aos[3 * i] = cx[i];
        aos[3 * i + 1] = cy[i];
        aos[3 * i + 2] = cz[i];
    }
    return 0;
}

int Ensemble::create(const swg_params &p, int ns, const int32_t *npl_s, const int32_t *ntp_s, const double *GMcb,
                     const double *pl_rh, const double *pl_vh, const double *Gmass, const double *radius,
                     const double *rhill_in, const double *tp_rh, const double *tp_vh)
{
    prm = p;
    if (!(p.dt > 0.0)) return fail(SWG_ERR_ARG, "dt must be positive");
    if (p.integrator != SWG_HELIO && p.integrator != SWG_SYMBA && p.integrator != SWG_WHM)
        return fail(SWG_ERR_ARG, "unknown integrator");
    if (ns <= 0) return fail(SWG_ERR_ARG, "nsys must be positive");
    if (p.pair_mode != SWG_PAIR_EXACT && p.pair_mode != SWG_PAIR_FAST)
        return fail(SWG_ERR_ARG, "pair_mode must be SWG_PAIR_EXACT or SWG_PAIR_FAST");
    if (!(p.gmtiny >= 0.0)) return fail(SWG_ERR_ARG, "gmtiny must be >= 0");
    tune.pair_fast = p.pair_mode == SWG_PAIR_FAST;
    if (const char *e = getenv("SWG_KICK_PL_EVAL")) {
        // exact-mode massive-body kernel before any autotuning: 0 one lane
        // per body, 2 / 4 / 8 evaluating wavefronts per 64-body tile
        const int k = atoi(e);
        if (k == 0 || k == 2 || k == 4 || k == 8) tune.kick_pl_eval = k;
    }
    nsys = ns;
    {
        // exact-mode massive-body kick before any autotuning: with fewer
        // than two wavefronts per SIMD in the one-lane-per-body kernel, the
        // terms go to 8 evaluating wavefronts per 64-body tile instead
        int64_t npl_all = 0;
        for (int s = 0; s < ns; s++) npl_all += npl_s[s] > 0 ? npl_s[s] : 0;
        if (!getenv("SWG_KICK_PL_EVAL") && (npl_all + 63) / 64 < 2048) tune.kick_pl_eval = 8;
    }
    h_pl_off.assign(nsys + 1, 0);
    h_tp_off.assign(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        if (npl_s[s] < 0 || ntp_s[s] < 0) return fail(SWG_ERR_ARG, "negative body count");
        h_pl_off[s + 1] = h_pl_off[s] + npl_s[s];
        h_tp_off[s + 1] = h_tp_off[s] + ntp_s[s];
    }
    npl = h_pl_off[nsys];
    ntp = h_tp_off[nsys];
    h_GMcb.assign(GMcb, GMcb + nsys);
    h_nreal_pl.assign(npl_s, npl_s + nsys);
    h_nreal_tp.assign(ntp_s, ntp_s + nsys);

    // ---- the reference's initialize sequence on the host side ----
    // set_rhill (unless given), then sort each system's massive bodies by mass,
    // descending (helio_util.f90:26), carrying every per-body array along.
    std::vector<double> rh(pl_rh, pl_rh + 3 * (size_t)npl), vh(pl_vh, pl_vh + 3 * (size_t)npl);
    std::vector<double> gm(Gmass, Gmass + npl), rad(radius, radius + npl), rhill(npl);
    for (int s = 0; s < nsys; s++) {
        int a = h_pl_off[s], n = h_pl_off[s + 1] - a;
        if (rhill_in && p.rhill_present)
            std::copy(rhill_in + a, rhill_in + a + n, rhill.begin() + a);
        else
            host_set_rhill(GMcb[s], n, &rh[3 * (size_t)a], &vh[3 * (size_t)a], &gm[a], &rhill[a]);
    }
    h_pl_id.resize(npl);
    {
        std::vector<double> rh2(rh.size()), vh2(vh.size()), gm2(npl), rad2(npl), rhill2(npl);
        for (int s = 0; s < nsys; s++) {
            int a = h_pl_off[s], n = h_pl_off[s + 1] - a;
            std::vector<int> idx(n);
            std::iota(idx.begin(), idx.end(), 0);
            // WHM sorts by 1/|rh|^3 (whm_util.f90:213-215), HELIO/SyMBA by mass
            std::vector<double> key(n);
            for (int q = 0; q < n; q++) {
                if (p.integrator == SWG_WHM) {
                    const double *x = &rh[3 * (size_t)(a + q)];
                    double r2 = dot3(x[0], x[1], x[2], x[0], x[1], x[2]);
                    key[q] = (1.0 / std::sqrt(r2)) / r2;
                } else {
                    key[q] = gm[a + q];
                }
            }
            std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) { return key[u] > key[v]; });
            for (int q = 0; q < n; q++) {
                int src = a + idx[q], dst = a + q;
                for (int d = 0; d < 3; d++) {
                    rh2[3 * (size_t)dst + d] = rh[3 * (size_t)src + d];
                    vh2[3 * (size_t)dst + d] = vh[3 * (size_t)src + d];
                }
                gm2[dst] = gm[src];
                rad2[dst] = rad[src];
                rhill2[dst] = rhill[src];
                h_pl_id[dst] = idx[q] + 1;
            }
        }
        rh.swap(rh2);
        vh.swap(vh2);
        gm.swap(gm2);
        rad.swap(rad2);
        rhill.swap(rhill2);
    }
    h_Gm = gm;

    // ---- device allocation ----
    HIPCHK(hipHostMalloc((void **)&h_small, 4096 + 16 * sizeof(int32_t) * (size_t)nsys));
    if (alloc_vec(pl.rh, npl) || alloc_vec(pl.vh, npl) || alloc_vec(pl.vb, npl) || alloc_vec(pl.rbeg, npl) ||
        alloc_vec(pl.rend, npl) || own(&pl.Gm, npl) || own(&pl.R, npl) || own(&pl.rhill, npl) || own(&pl.renc, npl) ||
        own(&pl.status, npl) || own(&pl.levelg, npl) || own(&pl.levelm, npl) || own(&pl.sys, npl) ||
        own(&pl.tiny, npl))
        return SWG_ERR_HIP;
    if (alloc_vec(tp.rh, ntp) || alloc_vec(tp.vh, ntp) || alloc_vec(tp.vb, ntp) || own(&tp.status, ntp) ||
        own(&tp.levelg, ntp) || own(&tp.levelm, ntp) || own(&tp.sys, ntp))
        return SWG_ERR_HIP;
    if (own(&sy.pl_off, nsys + 1) || own(&sy.tp_off, nsys + 1) || own(&sy.GMcb, nsys) || own(&sy.cbvb, 3 * nsys) ||
        own(&sy.ptbeg, 3 * nsys) || own(&sy.ptend, 3 * nsys) || own(&sy.lfirst, nsys) || own(&sy.enc_pp, nsys) ||
        own(&sy.enc_pt, nsys) || own(&sy.flag_pp, nsys) || own(&sy.flag_pt, nsys) || own(&sy.err, nsys) ||
        own(&sy.frozen, nsys) || own(&d_active, nsys) || own(&d_any, 4) || own(&d_total, 4) || own(&d_pt_run, nsys) ||
        own(&sy.nplm, nsys) || own(&d_cb_pl, nsys + 1) || own(&d_cb_tp, nsys + 1))
        return SWG_ERR_HIP;
    hipStream_t st = stream;
    {
        // each system's first 64-column block of the pl-pl and pl-tp bit
        // matrices (index2-side adjacency kernels)
        std::vector<int32_t> cbp(nsys + 1, 0), cbt(nsys + 1, 0);
        for (int s = 0; s < nsys; s++) {
            cbp[s + 1] = cbp[s] + (h_pl_off[s + 1] - h_pl_off[s] + 63) / 64;
            cbt[s + 1] = cbt[s] + (h_tp_off[s + 1] - h_tp_off[s] + 63) / 64;
        }
        nblk_pl = cbp[nsys];
        nblk_tp = cbt[nsys];
        HIPCHK(copy_wait(stream, d_cb_pl, cbp.data(), (nsys + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(stream, d_cb_tp, cbt.data(), (nsys + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    std::vector<double> tmp;
    if (upload_aos(st, pl.rh, rh.data(), npl, tmp) || upload_aos(st, pl.vh, vh.data(), npl, tmp)) return SWG_ERR_HIP;
    if (upload_aos(st, tp.rh, tp_rh, ntp, tmp) || upload_aos(st, tp.vh, tp_vh, ntp, tmp)) return SWG_ERR_HIP;
    HIPCHK(copy_wait(stream, pl.Gm, gm.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, pl.R, rad.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, pl.rhill, rhill.data(), npl * sizeof(double), hipMemcpyHostToDevice));
    std::vector<int32_t> zeros(std::max(std::max(npl, ntp), 3 * nsys) + 1, 0);
    std::vector<int32_t> sys_pl(npl), sys_tp(ntp), ones(nsys, 1);
    for (int s = 0; s < nsys; s++) {
        for (int i = h_pl_off[s]; i < h_pl_off[s + 1]; i++) sys_pl[i] = s;
        for (int i = h_tp_off[s]; i < h_tp_off[s + 1]; i++) sys_tp[i] = s;
    }
    HIPCHK(copy_wait(stream, pl.status, zeros.data(), npl * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, tp.status, zeros.data(), ntp * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, pl.sys, sys_pl.data(), npl * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, tp.sys, sys_tp.data(), ntp * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, sy.pl_off, h_pl_off.data(), (nsys + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, sy.tp_off, h_tp_off.data(), (nsys + 1) * sizeof(int32_t), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, sy.GMcb, GMcb, nsys * sizeof(double), hipMemcpyHostToDevice));
    std::vector<double> dz(3 * (size_t)nsys, 0.0);
    HIPCHK(copy_wait(stream, sy.ptbeg, dz.data(), 3 * nsys * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, sy.ptend, dz.data(), 3 * nsys * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, sy.cbvb, dz.data(), 3 * nsys * sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(stream, sy.lfirst, ones.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice));
    h_frozen.assign(nsys, 0);
    for (int32_t *q : {sy.enc_pp, sy.enc_pt, sy.flag_pp, sy.flag_pt, sy.err, sy.frozen})
        HIPCHK(copy_wait(stream, q, zeros.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice));
    // semi-interacting bodies (symba_util.f90:191-202 with lmtiny_pl for
    // SyMBA, swiftest_io.f90:2663): after the mass sort the fully interacting
    // bodies of a system come first, nplm of them
    {
        std::vector<int32_t> tiny(std::max(npl, 1), 0), nplm(nsys);
        for (int s = 0; s < nsys; s++) {
            int a = h_pl_off[s], n = h_pl_off[s + 1] - a, m = n;
            if (prm.integrator == SWG_SYMBA && prm.gmtiny > 0.0) {
                m = 0;
                for (int q = 0; q < n; q++) {
                    tiny[a + q] = h_Gm[a + q] < prm.gmtiny;
                    m += !tiny[a + q];
                }
            }
            nplm[s] = m;
            any_tiny |= m < n;
        }
        HIPCHK(copy_wait(stream, pl.tiny, tiny.data(), npl * sizeof(int32_t), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(stream, sy.nplm, nplm.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice));
        tune.any_tiny = any_tiny;
    }

    // ---- tiles ----
    if (int rc = build_tiles()) return rc;

    // ---- pl-pl bit matrix ----
    {
        std::vector<int64_t> woff(nsys);
        std::vector<int32_t> nb(nsys);
        int64_t w = 0;
        for (int s = 0; s < nsys; s++) {
            int n = h_pl_off[s + 1] - h_pl_off[s];
            nb[s] = (n + 63) / 64;
            woff[s] = w;
            w += (int64_t)n * nb[s];
        }
        // the sort-and-sweep check needs no bit matrix
        if (prm.encounter_check == 1) w = 1;
        pp_words = w;
        if (own2(&pp.words, w) || own2(&pp.wpre, w) || own2(&pp.woff, nsys) || own2(&pp.nb, nsys) ||
            own2(&pp.rowcnt, npl) || own2(&pp.rowstart, npl) || own2(&pp.colcnt, npl) || own2(&pp.colstart, npl))
            return SWG_ERR_HIP;
        HIPCHK(copy_wait(stream, pp.woff, woff.data(), nsys * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(stream, pp.nb, nb.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    {
        std::vector<int64_t> woff(nsys);
        std::vector<int32_t> nb(nsys);
        int64_t w = 0;
        for (int s = 0; s < nsys; s++) {
            int n = h_pl_off[s + 1] - h_pl_off[s], m = h_tp_off[s + 1] - h_tp_off[s];
            nb[s] = (m + 63) / 64;
            woff[s] = w;
            w += (int64_t)n * nb[s];
        }
        pt_words = w;
        if (own2(&pt.woff, nsys) || own2(&pt.nb, nsys) || own2(&pt.rowcnt, npl) || own2(&pt.rowstart, npl) ||
            own2(&pt.colcnt, ntp) || own2(&pt.colstart, ntp))
            return SWG_ERR_HIP;
        HIPCHK(copy_wait(stream, pt.woff, woff.data(), nsys * sizeof(int64_t), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(stream, pt.nb, nb.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    int mx = std::max(npl, ntp);
    HIPCHK(hipMalloc(&d_scan, scan_scratch_bytes(mx)));
    owned.push_back(d_scan);
    HIPCHK(hipMalloc(&d_scan_sys, scan_scratch_bytes(nsys)));
    owned.push_back(d_scan_sys);
    if (own(&bw.keepB, npl) || own(&bw.cntE, npl) || own(&bw.posB, npl) || own(&bw.posE, npl) ||
        own(&bw.keepBt, ntp) || own(&bw.cntEt, ntp) || own(&bw.posBt, ntp) || own(&bw.posEt, ntp) ||
        own(&bw.tot, 10) || own(&bw.rng, 4 * (size_t)nsys) || own(&bw.caps, nsys) || own(&d_kc, 8 * (size_t)npl) ||
        own(&d_arena_off, nsys + 1) || own(&d_sub_stats, 4))
        return SWG_ERR_HIP;
    bw.scan = d_scan;
    bw.maxsz = (int32_t *)(bw.tot + 6);
    lv_buf.assign(MAXLEV + 2, nullptr);
    lv_cap.assign(MAXLEV + 2, 0);
    lvv.assign(MAXLEV + 2, LevelView{});
    lvc.assign(MAXLEV + 2, LevelCounts{});
    if (const char *e = getenv("SWG_SUBTREE_MAX")) {
        subtree_max = strtoll(e, nullptr, 10);
        subtree_forced = true;
    }
    HIPCHK(hipEventCreate(&ev0));
    HIPCHK(hipEventCreate(&ev1));
    HIPCHK(hipEventCreateWithFlags(&ev_any, hipEventDisableTiming));

    // ---- device part of the initialize sequence: vh->vb, tp vb = vh + cb%vb ----
    if (prm.integrator == SWG_WHM) {
        if (int rc = whm_setup()) return rc;
    } else {
        launch_snapshot(st, pl, tp, sy, nsys, npl, 0);
        launch_init_tp_vb(st, tp, sy, ntp);
    }
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipGetLastError());
    st_.max_depth = -1;
    return 0;
}

template <class T>
int Ensemble::own2(T **p, size_t n)
{
    HIPCHK(dalloc(p, n));
    return 0;
}

int Ensemble::build_tiles()
{
    // kick tiles: blocks of tuning.kick_*_block bodies of one system
    std::vector<Tile> tpl, ttp;
    const int pl_tile = (tune.pair_fast || tune.kick_pl_eval) ? KPL_FAST_TILE : tune.kick_pl_block;
    for (int s = 0; s < nsys; s++) {
        if (h_frozen[s]) continue;
        // rows of the bodies a system holds; a slot's padding (inactive, at
        // the end of its range) is only ever a partner, with zero mass
        const int pe = h_pl_off[s] + h_nreal_pl[s], te = h_tp_off[s] + h_nreal_tp[s];
        for (int b = h_pl_off[s]; b < pe; b += pl_tile) tpl.push_back(Tile{s, b, std::min(b + pl_tile, pe), 0});
        for (int b = h_tp_off[s]; b < te; b += tune.kick_tp_block)
            ttp.push_back(Tile{s, b, std::min(b + tune.kick_tp_block, te), 0});
    }
    std::vector<int4> tpp;
    for (int s = 0; s < nsys; s++) {
        if (h_frozen[s]) continue;
        int n = h_pl_off[s + 1] - h_pl_off[s], nb = (n + 63) / 64;
        for (int ib = 0; ib < nb; ib++)
            for (int jb = ib; jb < nb; jb++) tpp.push_back(make_int4(s, ib, jb, 0));
    }
    // A tile array is published (its count set) only once it holds the
    // tiles: after a failed allocation or copy the count stays 0, so no
    // kernel is launched over a missing array, the error is returned, and
    // steps refuse to run (tiles_ok) until a later rebuild succeeds.
    // SWG_FAIL_TILE_ALLOC=1 makes the allocation fail, for tests of this path.
    tiles_ok = false;
    const char *inject = getenv("SWG_FAIL_TILE_ALLOC");
    auto put = [&](auto &vec, auto **dptr, int &cnt) -> int {
        using T = typename std::remove_reference<decltype(vec)>::type::value_type;
        cnt = 0;
        if (*dptr) (void)hipFree(*dptr);
        *dptr = nullptr;
        if (inject && inject[0] == '1')
            return fail(SWG_ERR_HIP, "hipMalloc (tile arrays): out of memory (injected by SWG_FAIL_TILE_ALLOC)");
        HIPCHK(hipMalloc((void **)dptr, sizeof(T) * (vec.size() ? vec.size() : 1)));
        if (!vec.empty()) HIPCHK(copy_wait(stream, *dptr, vec.data(), sizeof(T) * vec.size(), hipMemcpyHostToDevice));
        cnt = (int)vec.size();
        return 0;
    };
    if (int rc = put(tpl, &d_tiles_pl, ntiles_pl)) return rc;
    if (int rc = put(ttp, &d_tiles_tp, ntiles_tp)) return rc;
    if (int rc = put(tpp, &d_tiles_pp, ntiles_pp)) return rc;
    if (!d_tiles_pt) {
        int total = 0;
        for (int s = 0; s < nsys; s++) total += (h_tp_off[s + 1] - h_tp_off[s] + 63) / 64;
        HIPCHK(hipMalloc((void **)&d_tiles_pt, sizeof(int4) * (total ? total : 1)));
    }
    tiles_ok = true;
    return 0;
}

int Ensemble::ensure_list(ListDev &L, int64_t n, int64_t &cap)
{
    if (n <= cap) return 0;
    int64_t c = std::max<int64_t>(n + n / 2, 1024);
    for (void *q : {(void *)L.i1, (void *)L.i2, (void *)L.lvdotr, (void *)L.status, (void *)L.level, (void *)L.col})
        if (q) (void)hipFree(q);
    HIPCHK(dalloc(&L.i1, c));
    HIPCHK(dalloc(&L.i2, c));
    HIPCHK(dalloc(&L.lvdotr, c));
    HIPCHK(dalloc(&L.status, c));
    HIPCHK(dalloc(&L.level, c));
    HIPCHK(dalloc(&L.col, c));
    cap = c;
    return 0;
}

// Level-0 pl-pl detection: bit matrix, row counts, scan, ordered fill.
int Ensemble::detect_pp(double dt)
{
    hipStream_t st = stream;
    launch_pp_tiles(st, pl, sy, pp, d_tiles_pp, ntiles_pp, dt);
    launch_rows_count(st, pp, pl.sys, 1, sy.pl_off, nfrozen ? d_active : nullptr, npl);
    launch_scan_i32_i64(st, pp.rowcnt, pp.rowstart, npl, d_total, d_scan);
    launch_sys_flags(st, pp, sy.pl_off, sy.enc_pp, nsys);
    int64_t *ht = (int64_t *)h_small;
    int32_t *hf = (int32_t *)(h_small + 1024);
    HIPCHK(hipMemcpyAsync(ht, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hf, sy.enc_pp, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hf + nsys, sy.err, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    npp = npl ? ht[0] : 0;
    h_enc_pp.assign(hf, hf + nsys);
    h_err.assign(hf + nsys, hf + 2 * nsys);
    if (int rc = ensure_list(pp, npp, cap_pp)) return rc;
    if (npp)
        launch_rows_fill(st, pp, pl.sys, 1, sy.pl_off, sy.pl_off, npl, pl.levelg, pl.levelm, pl.levelg, pl.levelm);
    return 0;
}

// Level-0 pl-pl detection by sort and sweep (k_sas.hip): same list as
// detect_pp, without the N^2 bit matrix.
int Ensemble::detect_pp_sas(double dt)
{
    hipStream_t st = stream;
    if (hipError_t e = sas_candidates(st, sas, pl, sy, npl, nsys, nfrozen ? d_active : nullptr, dt, d_total, d_scan))
        return hipfail(e, "sort-and-sweep candidates");
    int64_t *ht = (int64_t *)h_small;
    HIPCHK(hipMemcpyAsync(ht, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int64_t ncand = npl ? ht[0] : 0;
    int64_t nsel = 0;
    if (hipError_t e = sas_narrow(st, sas, pl, npl, ncand, dt, &nsel)) return hipfail(e, "sort-and-sweep narrow phase");
    npp = nsel;
    if (int rc = ensure_list(pp, npp, cap_pp)) return rc;
    if (hipError_t e = sas_fill(st, sas, pp, pl, npl, npp)) return hipfail(e, "sort-and-sweep list");
    launch_sys_flags(st, pp, sy.pl_off, sy.enc_pp, nsys);
    int32_t *hf = (int32_t *)(h_small + 1024);
    HIPCHK(hipMemcpyAsync(hf, sy.enc_pp, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hf + nsys, sy.err, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    h_enc_pp.assign(hf, hf + nsys);
    h_err.assign(hf + nsys, hf + 2 * nsys);
    return 0;
}

// Level-0 pl-tp detection for the systems in `run`.
int Ensemble::detect_pt(double dt, const std::vector<int32_t> &run)
{
    hipStream_t st = stream;
    npt = 0;
    std::vector<int4> tiles;
    for (int s = 0; s < nsys; s++) {
        if (!run[s]) continue;
        int m = h_tp_off[s + 1] - h_tp_off[s];
        for (int tb = 0; tb < (m + 63) / 64; tb++) tiles.push_back(make_int4(s, tb, 0, 0));
    }
    if (tiles.empty()) {
        HIPCHK(hipMemsetAsync(sy.enc_pt, 0, nsys * sizeof(int32_t), st));
        return 0;
    }
    if (!pt.words) {
        HIPCHK(dalloc(&pt.words, pt_words));
        HIPCHK(dalloc(&pt.wpre, pt_words));
    }
    HIPCHK(hipMemcpyAsync(d_tiles_pt, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_pt_run, run.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice, st));
    launch_pt_tiles(st, pl, tp, sy, pt, d_tiles_pt, (int)tiles.size(), dt);
    launch_rows_count(st, pt, pl.sys, 0, sy.pl_off, d_pt_run, npl);
    launch_scan_i32_i64(st, pt.rowcnt, pt.rowstart, npl, d_total, d_scan);
    launch_sys_flags(st, pt, sy.pl_off, sy.enc_pt, nsys);
    int64_t *ht = (int64_t *)h_small;
    int32_t *hf = (int32_t *)(h_small + 1024);
    HIPCHK(hipMemcpyAsync(ht, d_total, sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(hf, sy.enc_pt, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    npt = npl ? ht[0] : 0;
    h_enc_pt.assign(hf, hf + nsys);
    if (int rc = ensure_list(pt, npt, cap_pt)) return rc;
    if (npt)
        launch_rows_fill(st, pt, pl.sys, 0, sy.pl_off, sy.tp_off, npl, pl.levelg, pl.levelm, tp.levelg, tp.levelm);
    return 0;
}

int Ensemble::build_adjacency()
{
    hipStream_t st = stream;
    if (npp && prm.encounter_check == 1) {
        if (hipError_t e = sas_adjacency(st, sas, pp, npl, npp)) return hipfail(e, "sort-and-sweep adjacency");
    } else if (npp && npp <= COLS_MAX && npl <= COLS_NPL && !nfrozen) {
        launch_col_small(st, pp, npl, (int)npp);
    } else if (npp) {
        launch_col_t(st, pp, sy.pl_off, sy.pl_off, d_cb_pl, nsys, nblk_pl, nfrozen ? d_active : nullptr, 1, 0);
        launch_scan_i32_i64(st, pp.colcnt, pp.colstart, npl, nullptr, d_scan);
        launch_col_t(st, pp, sy.pl_off, sy.pl_off, d_cb_pl, nsys, nblk_pl, nfrozen ? d_active : nullptr, 1, 1);
    }
    if (npt) {
        launch_col_t(st, pt, sy.pl_off, sy.tp_off, d_cb_tp, nsys, nblk_tp, d_pt_run, 0, 0);
        launch_scan_i32_i64(st, pt.colcnt, pt.colstart, ntp, nullptr, d_scan);
        launch_col_t(st, pt, sy.pl_off, sy.tp_off, d_cb_tp, nsys, nblk_tp, d_pt_run, 0, 1);
    }
    return 0;
}

// Queue the copy of the encounter flag to the host; wait_any waits for that
// copy alone, so work queued after post_any keeps the device busy meanwhile.
int Ensemble::post_any()
{
    int32_t *h = (int32_t *)(h_small + 32);
    HIPCHK(hipMemcpyAsync(h, d_any, sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipEventRecord(ev_any, stream));
    return 0;
}

int Ensemble::wait_any(int32_t &v)
{
    HIPCHK(hipEventSynchronize(ev_any));
    v = *(const int32_t *)(h_small + 32);
    return 0;
}

LevelArgs Ensemble::level_args() const
{
    LevelArgs la;
    la.pl = pl;
    la.tp = tp;
    la.pp = pp;
    la.pt = pt;
    la.sy = sy;
    la.nsys = nsys;
    la.has_pp = npp > 0;
    la.has_pt = npt > 0;
    return la;
}

// The single level-0 substep of symba_step_recur_system (symba_step.f90:
// 163-262 with ireci = 0) over the whole encounter lists.
int Ensemble::recur0()
{
    hipStream_t st = stream;
    const double dt = prm.dt, dth = 0.5 * dt;
    const LevelArgs la = level_args();
    if (st_.max_depth < 0) st_.max_depth = 0;
    st_.nsubsteps++;
    launch_clear_flags(st, sy, nsys, d_any);
    launch_list_check_pp(st, pp, pl, sy, npp, 1, dt, d_any);
    launch_list_check_pt(st, pt, pl, tp, sy, npt, 1, dt, prm.short_circuit, d_any);
    // the counts of level 1 and the encounter flag are sent to the host
    // before the kicks and the drift, which change neither, run: the host
    // reads them while the device works
    if (int rc = build_begin(1, nullptr)) return rc;
    launch_level_kicks(st, la, nullptr, npl, ntp, 0, dth);
    launch_drift(st, pl.rh, pl.vb, pl.status, pl.levelg, pl.sys, sy.GMcb, sy.err, sy.frozen, npl, dt, 0, 1, 0);
    launch_drift(st, tp.rh, tp.vb, tp.status, tp.levelg, tp.sys, sy.GMcb, sy.err, sy.frozen, ntp, dt, 0, 1, 1);
    int32_t lenc = 0;
    if (int rc = build_wait(lenc)) return rc;
    if (lenc) {
        bool handoff = false;
        if (int rc = build_end(1, nullptr, lvv[1], lvc[1], handoff)) return rc;
        tr_hand[1] += handoff;
        if (handoff) {
            if (int rc = run_subtree(1, lvv[1])) return rc;
        } else {
            if (int rc = recur_level(1)) return rc;
        }
    }
    // a merge inside the recursion changes the bodies and the list
    const LevelArgs lb = level_args();
    launch_level_kicks(st, lb, nullptr, npl, ntp, 0, dth);
    if (prm.lclose) launch_collision(st, pp, pt, pl, tp, sy, npp, npt, 0, dt, d_any);
    if (merge_replay)
        if (int rc = merge_point(0, dt)) return rc;
    launch_set_levels(st, pl, tp, pp, pt, npl, ntp, npp, npt, 0);
    return 0;
}

// Build the whole-device working set of level dnew from src (nullptr: the
// level-0 adjacency) into the buffer of depth dnew, in three parts.
// build_begin queues the counts, the arena sizes and a copy of them and of the
// encounter flag to the host; it depends only on the levels the check just
// set, so it runs before the kicks and the drift of the current substep.
// build_wait waits for that copy.  build_end (only when the flag is set)
// sizes the buffer, queues the fill and tells whether every system's set is
// small enough for the persistent workgroups (handoff).
int Ensemble::build_begin(int dnew, const LevelView *src)
{
    hipStream_t st = stream;
    const LevelArgs la = level_args();
    if (launch_build_count(st, la, src, npl, ntp, dnew, bw, d_any, d_arena_off)) {
        // the arena offsets came with the ranges
    } else if (nsys <= SCAN_SMALL_MAX) {
        ScanSmall sc{};
        sc.in[0] = bw.caps;
        sc.out[0] = d_arena_off;
        sc.n[0] = nsys;
        sc.total[0] = d_arena_off + nsys;
        sc.total2[0] = bw.tot + 4;
        launch_scan_small(st, sc, 1);
    } else {
        launch_scan_i32_i64(st, bw.caps, d_arena_off, nsys, bw.tot + 4, d_scan_sys);
        HIPCHK(hipMemcpyAsync(d_arena_off + nsys, bw.tot + 4, sizeof(int64_t), hipMemcpyDeviceToDevice, st));
    }
    // counts, arena total, flag and the per-system maxima in one copy
    HIPCHK(hipMemcpyAsync(h_small + 128, bw.tot, 10 * sizeof(int64_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ev_any, st));
    return 0;
}

int Ensemble::build_wait(int32_t &lenc)
{
    HIPCHK(hipEventSynchronize(ev_any));
    lenc = (int32_t)((const int64_t *)(h_small + 128))[5];
    return 0;
}

int Ensemble::build_end(int dnew, const LevelView *src, LevelView &v, LevelCounts &c, bool &handoff)
{
    hipStream_t st = stream;
    const LevelArgs la = level_args();
    const int64_t *h = (const int64_t *)(h_small + 128);
    c.nB = h[0];
    c.nE = h[1];
    c.nBt = h[2];
    c.nEt = h[3];
    arena_need = h[4];
    const int32_t *mx = (const int32_t *)(h + 6);
    const int32_t maxsz = mx[0];
    if (c.nE > INT32_MAX || c.nEt > INT32_MAX)
        return fail(SWG_ERR_SCRATCH, "recursion working set of level " + std::to_string(dnew) + " exceeds 2^31 entries");
    const int64_t bytes = level_bytes(c);
    if (bytes > lv_cap[dnew]) {
        if (lv_buf[dnew]) (void)hipFree(lv_buf[dnew]);
        lv_buf[dnew] = nullptr;
        lv_cap[dnew] = 0;
        const int64_t cap = bytes + bytes / 4 + 256;
        HIPCHK(hipMalloc(&lv_buf[dnew], (size_t)cap));
        lv_cap[dnew] = cap;
    }
    launch_build_fill(st, la, src, npl, ntp, dnew, bw, lv_buf[dnew], c);
    v = level_view(lv_buf[dnew], c);
    // Hand the rest of the recursion to one workgroup per system when every
    // system's set fits the LDS-resident kernel, or when enough systems have
    // work to fill the device with workgroups.  A single large set stays on
    // the whole device, where its levels run many times faster than in one
    // workgroup.  SWG_SUBTREE_MAX in the environment forces the size rule
    // alone (hand-off modes of the tests).
    const bool fits = subtree_lds_fits(mx[1], mx[2], mx[3], mx[4]);
    level_fits = fits;
    handoff = !merge_replay && subtree_max >= 0 && maxsz <= subtree_max &&
              (subtree_forced || fits || mx[5] >= SUBTREE_MIN_SYSTEMS);
    return 0;
}

// symba_step_recur_system (symba_step.f90:163-262) at level d >= 1 on the
// whole device, all systems in lockstep: a system with no pair at level d has
// an empty share of the working set, and every launch is a no-op for it.
int Ensemble::recur_level(int d)
{
    hipStream_t st = stream;
    const double dt = prm.dt;
    double p3 = 1.0;
    for (int q = 0; q < d; q++) p3 *= NTENC;
    const double dtl = dt / p3, dth = 0.5 * dtl;
    if (dtl / dt < DBL_EPSILON) return fail(SWG_ERR_DEPTH, "recursion step below double precision");
    if (d > st_.max_depth) st_.max_depth = d;
    // the level's working set; a merge rebuilds it in place, so the sizes
    // and the state arguments are read again after anything that can merge
    const LevelView &v = lvv[d];
    const LevelCounts &c = lvc[d];
    for (int j = 1; j <= NTENC; j++) {
        LevelArgs la = level_args();
        int nB = (int)c.nB, nBt = (int)c.nBt;
        st_.nsubsteps++;
        tr_sub[d]++;
        tr_B[d] += c.nB + c.nBt;
        tr_E[d] += c.nE + c.nEt;
        if (j == 1) launch_clear_flags(st, sy, nsys, d_any);  // later substeps: cleared by the level reset
        launch_level_check(st, la, v, c, d, dtl, prm.short_circuit, d_any);
        const bool room = d + 1 <= MAXLEV;
        if (int rc = room ? build_begin(d + 1, &v) : post_any()) return rc;
        // opening kicks: from the previous substep's closing kicks when
        // nothing between them can have changed the sums (no collision test)
        const bool kept = j > 1 && !prm.lclose;
        launch_level_kicks(st, la, &v, nB, nBt, d, dth, kept ? 2 : 0, d_kc);
        launch_level_drift(st, la, v, nB, nBt, d, dtl);
        int32_t lenc = 0;
        if (int rc = room ? build_wait(lenc) : wait_any(lenc)) return rc;
        if (lenc) {
            if (!room) return fail(SWG_ERR_DEPTH, "recursion deeper than the level stack");
            bool handoff = false;
            if (int rc = build_end(d + 1, &v, lvv[d + 1], lvc[d + 1], handoff)) return rc;
            tr_hand[d + 1] += handoff;
            if (handoff) {
                if (int rc = run_subtree(d + 1, lvv[d + 1])) return rc;
            } else {
                if (int rc = recur_level(d + 1)) return rc;
            }
            la = level_args();
            nB = (int)c.nB;
            nBt = (int)c.nBt;
        }
        launch_level_kicks(st, la, &v, nB, nBt, d, dth, (j < NTENC && !prm.lclose) ? 1 : 0, d_kc);
        if (prm.lclose) launch_level_collision(st, la, v, nB, nBt, dtl);
        if (merge_replay) {
            if (int rc = merge_point(d, dtl)) return rc;
            la = level_args();
        }
        launch_level_set_levels(st, la, v, c, d, j < NTENC ? d_any : nullptr);
    }
    return 0;
}

// Run the recursion of every system from level d0 down in one persistent
// workgroup each (k_subtree.hip), from the level-d0 set just built.
int Ensemble::run_subtree(int d0, const LevelView &v)
{
    hipStream_t st = stream;
    if (arena_need > arena_words) {
        if (d_arena) (void)hipFree(d_arena);
        d_arena = nullptr;
        arena_words = 0;
        const int64_t w = arena_need + arena_need / 4 + 64;
        HIPCHK(dalloc(&d_arena, (size_t)w));
        arena_words = w;
    }
    launch_subtree(st, level_args(), v, bw.rng, d0, prm.dt, prm.lclose, prm.short_circuit, d_arena, d_arena_off,
                   d_sub_stats, level_fits);
    HIPCHK(hipGetLastError());
    return 0;
}

// Energy and angular momentum of every system (swg_system_energy).
int Ensemble::energy(const double *cb_mass, const double *cb_radius, const double *pl_mass, double *out)
{
    hipStream_t st = stream;
    if (!en_ready) {
        std::vector<int4> tiles;
        std::vector<int32_t> toff;
        energy_tiles(nsys, h_pl_off.data(), tiles, toff);
        en_ntiles = (int)tiles.size();
        if (alloc_vec(en_rb, npl) || own(&en_mass, npl) || own(&en_cb, 2 * (size_t)nsys) ||
            own(&en_part, (size_t)en_ntiles) || own(&en_out, (size_t)nsys * SWG_ENERGY_NOUT) ||
            own(&en_tiles, (size_t)en_ntiles) || own(&en_toff, (size_t)nsys + 1))
            return SWG_ERR_HIP;
        if (en_ntiles)
            HIPCHK(copy_wait(stream, en_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(stream, en_toff, toff.data(), toff.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        en_ready = true;
    }
    if (npl) HIPCHK(hipMemcpyAsync(en_mass, pl_mass, npl * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(en_cb, cb_mass, nsys * sizeof(double), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(en_cb + nsys, cb_radius, nsys * sizeof(double), hipMemcpyHostToDevice, st));
    // swiftest_io_conservation_report (swiftest_io.f90:135-141): WHM holds vh
    if (prm.integrator == SWG_WHM) launch_vh2vb(st, pl, sy, nsys, npl);
    launch_energy(st, pl, sy, nsys, en_mass, en_cb, en_cb + nsys, prm.lclose, en_rb, en_tiles, en_ntiles, en_toff,
                  en_part, en_out);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, en_out, (size_t)nsys * SWG_ENERGY_NOUT * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// Discard checks after a step (swg_system_discard).
int Ensemble::discard(const swg_discard_params &dp, const double *cb_radius, int64_t *npl_dis, int64_t *ntp_dis)
{
    hipStream_t st = stream;
    if (!dis_ready) {
        std::vector<int4> tiles;
        discard_tiles(nsys, h_tp_off.data(), tiles);
        dis_ntiles = (int)tiles.size();
        if (own(&dis_pl_isperi, npl) || own(&dis_tp_isperi, ntp) || own(&dis_lfirst, nsys) || own(&dis_pl_atp, npl) ||
            own(&dis_pl_peri, npl) || own(&dis_tp_atp, ntp) || own(&dis_tp_peri, ntp) ||
            own(&dis_sys, (size_t)nsys * discard_sys_words()) || own(&dis_rcb, nsys) || own(&dis_gone, npl) ||
            own(&dis_count, 2) || own(&dis_tiles, (size_t)dis_ntiles))
            return SWG_ERR_HIP;
        // swiftest_util_setup_body (swiftest_util.f90:2537): isperi = 1; the
        // first pericentre call of SyMBA's massive bodies is flagged per system
        std::vector<int32_t> one(std::max(std::max(npl, ntp), nsys), 1);
        std::vector<double> zero(std::max(npl, ntp), 0.0);
        if (npl) {
            HIPCHK(copy_wait(stream, dis_pl_isperi, one.data(), npl * sizeof(int32_t), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(stream, dis_pl_atp, zero.data(), npl * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(stream, dis_pl_peri, zero.data(), npl * sizeof(double), hipMemcpyHostToDevice));
        }
        if (ntp) {
            HIPCHK(copy_wait(stream, dis_tp_isperi, one.data(), ntp * sizeof(int32_t), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(stream, dis_tp_atp, zero.data(), ntp * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(stream, dis_tp_peri, zero.data(), ntp * sizeof(double), hipMemcpyHostToDevice));
        }
        HIPCHK(copy_wait(stream, dis_lfirst, one.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice));
        if (dis_ntiles)
            HIPCHK(copy_wait(stream, dis_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice));
        dis_ready = true;
    }
    HIPCHK(hipMemcpyAsync(dis_rcb, cb_radius, nsys * sizeof(double), hipMemcpyHostToDevice, st));
    const bool symba = prm.integrator == SWG_SYMBA;
    // pl%nplenc: the pairs of each body in the last step's level-0 list
    const bool enc = symba && npp > 0;
    launch_discard(st, pl, tp, sy, nsys, npl, symba, dp, prm.lclose, prm.dt, dis_rcb, enc ? pp.rowcnt : nullptr,
                   enc ? pp.colcnt : nullptr, dis_pl_isperi, dis_pl_atp, dis_pl_peri, dis_tp_isperi, dis_tp_atp,
                   dis_tp_peri, dis_lfirst, dis_gone, dis_sys, dis_count, dis_tiles, dis_ntiles);
    HIPCHK(hipGetLastError());
    unsigned long long c[2];
    HIPCHK(hipMemcpyAsync(c, dis_count, sizeof c, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (npl_dis) *npl_dis = (int64_t)c[0];
    if (ntp_dis) *ntp_dis = (int64_t)c[1];
    return 0;
}

int Ensemble::check_errors()
{
    int32_t *hf = (int32_t *)(h_small + 1024);
    HIPCHK(hipMemcpyAsync(hf, sy.err, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    h_err.assign(hf, hf + nsys);
    return err_code();
}

int Ensemble::err_code()
{
    for (int s = 0; s < nsys; s++) {
        if (h_err[s] & 1) return fail(SWG_ERR_COLLISION, "collision detected in system " + std::to_string(s));
        if (h_err[s] & 2) return fail(SWG_ERR_DRIFT, "massive body drift failure in system " + std::to_string(s));
        if (h_err[s] & 4)
            return fail(SWG_ERR_SCRATCH, "recursion working sets exceeded their arena in system " + std::to_string(s));
        if (h_err[s] & 8) return fail(SWG_ERR_DEPTH, "recursion step below double precision in system " + std::to_string(s));
    }
    return 0;
}

// Phase trace for diagnosis (SWG_TRACE=1 in the environment): synchronizes
// the stream and prints the time since the previous mark to stderr.
static bool trace_on()
{
    static int on = -1;
    if (on < 0) {
        const char *e = getenv("SWG_TRACE");
        on = (e && e[0] == '1') ? 1 : 0;
    }
    return on == 1;
}

static void trace_mark(hipStream_t st, const char *what)
{
    static std::chrono::steady_clock::time_point last = std::chrono::steady_clock::now();
    if (!trace_on()) return;
    (void)hipStreamSynchronize(st);
    auto now = std::chrono::steady_clock::now();
    fprintf(stderr, "[swg] %-18s %10.3f ms\n", what, std::chrono::duration<double, std::milli>(now - last).count());
    last = now;
}

void Ensemble::trace_levels()
{
    fprintf(stderr, "[swg] host level substeps   avg_nB   avg_nE handoffs_in\n");
    for (int d = 1; d <= MAXLEV; d++) {
        if (tr_sub[d] || tr_hand[d])
            fprintf(stderr, "[swg] %10d %8lld %8.0f %8.0f %8lld\n", d, (long long)tr_sub[d],
                    tr_sub[d] ? (double)tr_B[d] / tr_sub[d] : 0.0, tr_sub[d] ? (double)tr_E[d] / tr_sub[d] : 0.0,
                    (long long)tr_hand[d]);
        tr_sub[d] = tr_B[d] = tr_E[d] = tr_hand[d] = 0;
    }
}

// The two halves of the step around the recursion are fixed sequences of
// launches whose arguments change only when the tiles or the encounter lists'
// buffers move, or with whether the lists are empty.  With graphs on, each
// (half, pl-pl list empty or not, pl-tp list empty or not) variant is
// captured once as a HIP graph and replayed; the capture holds kernel nodes
// only.  Launch timing (ktiming) records events between the launches and
// always runs the halves directly.
uint64_t Ensemble::graph_key() const
{
    const void *ptrs[] = {d_tiles_pl, d_tiles_tp, pp.i1,     pp.i2,      pp.col,    pp.colcnt, pp.colstart,
                          pp.rowcnt,  pp.rowstart, pp.status, pt.i1,      pt.i2,     pt.col,    pt.colcnt,
                          pt.colstart, pt.status, d_arena,   (const void *)(intptr_t)ntiles_pl,
                          (const void *)(intptr_t)ntiles_tp, (const void *)(intptr_t)tune.kick_pl_block,
                          (const void *)(intptr_t)tune.kick_tp_block, (const void *)(intptr_t)tune.pair_fast,
                          (const void *)(intptr_t)tune.kick_pl_split, (const void *)(intptr_t)tune.kick_pl_eval};
    uint64_t h = 1469598103934665603ull;
    for (const void *p : ptrs) {
        h ^= (uint64_t)(uintptr_t)p;
        h *= 1099511628211ull;
    }
    return h;
}

void Ensemble::drop_graphs()
{
    // a replay still queued on the stream must end before its executable
    // graph (and the kernel arguments it owns) is destroyed
    bool any = false;
    for (auto &g : graph_exec) any |= g != nullptr;
    if (any && stream) (void)hipStreamSynchronize(stream);
    for (auto &g : graph_exec)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

// Allocations, frees and synchronous copies of one context can invalidate a
// graph capture in progress on another context's stream.  Captures and the
// creation and destruction of systems take this lock, so they never overlap;
// a capture that is invalidated all the same (by an allocation inside another
// context's step) is dropped, and its launches, none of which ran, are issued
// directly instead.
static std::mutex g_capture_mu;

int Ensemble::run_phase(int half, const std::function<int()> &body)
{
    if (!use_graphs || ktiming || capture_failures >= 4) return body();
    const uint64_t key = graph_key();
    if (key != graph_key_cached) {
        drop_graphs();
        graph_key_cached = key;
    }
    hipGraphExec_t &ex = graph_exec[half * 4 + (npp > 0) * 2 + (npt > 0)];
    if (!ex) {
        std::lock_guard<std::mutex> lk(g_capture_mu);
        hipGraph_t g = nullptr;
        // relaxed: the capture holds kernel launches only
        HIPCHK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        int rc = body();
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e == hipSuccess) {
            e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            if (e != hipSuccess) ex = nullptr;
        }
        if (g) (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            capture_failures++;
            return body();
        }
    }
    HIPCHK(hipGraphLaunch(ex, stream));
    return 0;
}

// One SyMBA / HELIO step of every system.
int Ensemble::step_helio_symba()
{
    hipStream_t st = stream;
    trace_mark(st, "step start");
    const double dt = prm.dt, dth = 0.5 * dt;
    const bool symba = prm.integrator == SWG_SYMBA;
    st_.nplpl_enc = 0;
    st_.npltp_enc = 0;
    st_.max_depth = -1;
    st_.nsubsteps = 0;
    launch_reset_levels(st, pl, tp, npl, ntp);
    HIPCHK(hipMemsetAsync(d_sub_stats, 0, 4 * sizeof(int32_t), st));
    npp = npt = 0;
    bool any = false;
    if (symba) {
        if (int rc = prm.encounter_check == 1 ? detect_pp_sas(dt) : detect_pp(dt)) return rc;
        trace_mark(st, "detect pl-pl");
        if (int rc = err_code()) return rc;
        std::vector<int32_t> run(nsys, 0);
        bool anyrun = false;
        for (int s = 0; s < nsys; s++) {
            run[s] = (h_tp_off[s + 1] > h_tp_off[s]) && (h_pl_off[s + 1] > h_pl_off[s]) &&
                     !(prm.short_circuit && h_enc_pp[s]) && !h_frozen[s];
            anyrun |= run[s] != 0;
        }
        h_enc_pt.assign(nsys, 0);
        if (anyrun) {
            if (int rc = detect_pt(dt, run)) return rc;
        } else {
            HIPCHK(hipMemsetAsync(sy.enc_pt, 0, nsys * sizeof(int32_t), st));
        }
        for (int s = 0; s < nsys; s++) any |= h_enc_pp[s] || h_enc_pt[s];
        st_.nplpl_enc = npp;
        st_.npltp_enc = npt;
        if (any) build_adjacency();
        trace_mark(st, "detect pl-tp+adj");
    } else {
        HIPCHK(hipMemsetAsync(sy.enc_pp, 0, nsys * sizeof(int32_t), st));
        HIPCHK(hipMemsetAsync(sy.enc_pt, 0, nsys * sizeof(int32_t), st));
    }
    const int use_radius = symba || prm.lclose;
    auto kick_pl = [&](int lbeg) -> int {
        if (ktiming && ntiles_pl)
            if (int rc = record_kernel_event(ev_pl, nev_pl)) return rc;
        launch_kick_pl(st, tune, pl, sy, pp, d_tiles_pl, ntiles_pl, dth, lbeg, use_radius, npp > 0);
        if (ktiming && ntiles_pl)
            if (int rc = record_kernel_event(ev_pl, nev_pl)) return rc;
        return 0;
    };
    auto kick_tp = [&](int lbeg) -> int {
        if (ktiming && ntiles_tp)
            if (int rc = record_kernel_event(ev_tp, nev_tp)) return rc;
        launch_kick_tp(st, tune, tp, pl, sy, pt, d_tiles_tp, ntiles_tp, dth, lbeg, npt > 0);
        if (ktiming && ntiles_tp)
            if (int rc = record_kernel_event(ev_tp, nev_tp)) return rc;
        return 0;
    };
    // first half: pl then tp
    auto first_half = [&]() -> int {
        launch_sums_begin(st, pl, sy, nsys, 0);
        launch_pl_begin_apply(st, pl, sy, npl, dth, 0);
        if (int rc = kick_pl(1)) return rc;
        launch_drift(st, pl.rh, pl.vb, pl.status, pl.levelg, pl.sys, sy.GMcb, sy.err, sy.frozen, npl, dt, -1, symba,
                     0);
        launch_tp_pre(st, tp, sy, ntp, dth);
        if (int rc = kick_tp(1)) return rc;
        launch_drift(st, tp.rh, tp.vb, tp.status, tp.levelg, tp.sys, sy.GMcb, sy.err, sy.frozen, ntp, dt, -1, symba,
                     1);
        return 0;
    };
    auto second_half = [&]() -> int {
        if (int rc = kick_pl(0)) return rc;
        launch_sums_end(st, pl, sy, nsys);
        launch_pl_end_apply(st, pl, sy, npl, dth);
        if (int rc = kick_tp(0)) return rc;
        launch_tp_post(st, tp, sy, ntp, dth);
        launch_lfirst_update(st, sy, nsys);
        return 0;
    };
    if (int rc = run_phase(0, first_half)) return rc;
    trace_mark(st, "first half");
    if (any) {
        if (int rc = recur0()) return rc;
        trace_mark(st, "recursion");
        if (trace_on()) {
            trace_levels();
            subtree_profile_report(st);
        }
    }
    // second half
    if (int rc = run_phase(1, second_half)) return rc;
    trace_mark(st, "second half");
    HIPCHK(hipGetLastError());
    // a collision flagged in this step's recursion ends the run at the end of
    // the step, with the colliding pairs flagged in the lists
    if (any && prm.lclose) return check_errors();
    return 0;
}

int Ensemble::snapshot()
{
    hipStream_t st = stream;
    if (prm.integrator == SWG_WHM) return whm_snapshot();
    launch_snapshot(st, pl, tp, sy, nsys, npl, ntp);
    if (!prm.rhill_present) {
        std::vector<double> rh(3 * (size_t)npl), vh(3 * (size_t)npl), rhill(npl);
        if (download_aos(st, pl.rh, rh.data(), npl) || download_aos(st, pl.vh, vh.data(), npl)) return SWG_ERR_HIP;
        if (nfrozen) HIPCHK(hipMemcpyAsync(rhill.data(), pl.rhill, npl * sizeof(double), hipMemcpyDeviceToHost, st));
        for (int s = 0; s < nsys; s++) {
            if (h_frozen[s]) continue;  // a frozen system keeps its state, rhill included
            int a = h_pl_off[s], n = h_pl_off[s + 1] - a;
            host_set_rhill(h_GMcb[s], n, &rh[3 * (size_t)a], &vh[3 * (size_t)a], &h_Gm[a], &rhill[a]);
        }
        HIPCHK(hipMemcpyAsync(pl.rhill, rhill.data(), npl * sizeof(double), hipMemcpyHostToDevice, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return 0;
}

int Ensemble::step()
{
    if (!tiles_ok) return fail(SWG_ERR_HIP, "the tile arrays are missing after a failed allocation; rebuild them first");
    int rc = prm.integrator == SWG_WHM ? step_whm() : merge_on ? step_merge() : step_helio_symba();
    if (rc && rc != SWG_ERR_COLLISION) return rc;
    iloop++;  // a step that found a collision was completed
    return rc;
}

int Ensemble::run(int64_t nsteps)
{
    if (iloop == 0 && prm.istep_out > 0 && !initial_snapshot_done) {
        if (int rc = snapshot()) return rc;
        initial_snapshot_done = true;
    }
    for (int64_t k = 0; k < nsteps;) {
        // WHM systems of few massive bodies: up to the next snapshot (at
        // most 4096 steps) in one launch; otherwise one step at a time
        int64_t m = 1;
        if (prm.integrator == SWG_WHM && whm_fused) {
            m = std::min<int64_t>(nsteps - k, 4096);
            if (prm.istep_out > 0) m = std::min<int64_t>(m, prm.istep_out - iout);
            if (int rc = run_whm_fused((int)m)) return rc;
            iloop += m;
        } else if (int rc = step()) {
            return rc;
        }
        k += m;
        if (prm.istep_out > 0) {
            iout += m;
            if (iout == prm.istep_out) {
                iout = 0;
                if (int rc = snapshot()) return rc;
            }
        }
    }
    HIPCHK(hipStreamSynchronize(stream));
    return check_errors();
}

int Ensemble::get(double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh, double *tp_vb,
                  int32_t *pl_status, int32_t *tp_status, int32_t *pl_id)
{
    hipStream_t st = stream;
    HIPCHK(hipStreamSynchronize(st));
    if (pl_rh && download_aos(st, pl.rh, pl_rh, npl)) return SWG_ERR_HIP;
    if (pl_vh && download_aos(st, pl.vh, pl_vh, npl)) return SWG_ERR_HIP;
    if (pl_vb && download_aos(st, pl.vb, pl_vb, npl)) return SWG_ERR_HIP;
    if (tp_rh && download_aos(st, tp.rh, tp_rh, ntp)) return SWG_ERR_HIP;
    if (tp_vh && download_aos(st, tp.vh, tp_vh, ntp)) return SWG_ERR_HIP;
    if (tp_vb && download_aos(st, tp.vb, tp_vb, ntp)) return SWG_ERR_HIP;
    if (pl_status) HIPCHK(copy_wait(stream, pl_status, pl.status, npl * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (tp_status) HIPCHK(copy_wait(stream, tp_status, tp.status, ntp * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (pl_id) std::copy(h_pl_id.begin(), h_pl_id.end(), pl_id);
    return 0;
}

int Ensemble::get_misc(double *rhill, double *cb_vb, double *ptbeg, double *ptend, int32_t *lfirst)
{
    HIPCHK(hipStreamSynchronize(stream));
    if (rhill) HIPCHK(copy_wait(stream, rhill, pl.rhill, npl * sizeof(double), hipMemcpyDeviceToHost));
    if (cb_vb) HIPCHK(copy_wait(stream, cb_vb, sy.cbvb, 3 * nsys * sizeof(double), hipMemcpyDeviceToHost));
    if (ptbeg) HIPCHK(copy_wait(stream, ptbeg, sy.ptbeg, 3 * nsys * sizeof(double), hipMemcpyDeviceToHost));
    if (ptend) HIPCHK(copy_wait(stream, ptend, sy.ptend, 3 * nsys * sizeof(double), hipMemcpyDeviceToHost));
    if (lfirst) HIPCHK(copy_wait(stream, lfirst, sy.lfirst, nsys * sizeof(int32_t), hipMemcpyDeviceToHost));
    return 0;
}

int Ensemble::enc_list(int which, int64_t *n, int32_t *i1, int32_t *i2)
{
    ListDev &L = which == 0 ? pp : pt;
    int64_t m = which == 0 ? npp : npt;
    *n = m;
    if (!i1 || !m) return 0;
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(copy_wait(stream, i1, L.i1, m * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(copy_wait(stream, i2, L.i2, m * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < m; k++) {
        i1[k] += 1;
        i2[k] += 1;
    }
    return 0;
}

int Ensemble::enc_status(int which, int64_t *n, int32_t *status)
{
    ListDev &L = which == 0 ? pp : pt;
    int64_t m = which == 0 ? npp : npt;
    *n = m;
    if (!status || !m) return 0;
    std::vector<int8_t> h(m);
    HIPCHK(hipStreamSynchronize(stream));
    HIPCHK(copy_wait(stream, h.data(), L.status, m, hipMemcpyDeviceToHost));
    for (int64_t k = 0; k < m; k++) status[k] = h[k];
    return 0;
}

int Ensemble::record_kernel_event(std::vector<hipEvent_t> &pool, size_t &n)
{
    if (n == pool.size()) {
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        pool.push_back(e);
    }
    HIPCHK(hipEventRecord(pool[n], stream));
    n++;
    return 0;
}

// average duration of the recorded launches (start/stop event pairs)
int Ensemble::kernel_times(double *ms_tp, int64_t *n_tp, double *ms_pl, int64_t *n_pl)
{
    HIPCHK(hipStreamSynchronize(stream));
    auto avg = [&](std::vector<hipEvent_t> &pool, size_t n, double *ms, int64_t *cnt) -> int {
        double tot = 0.0;
        int64_t k = 0;
        for (size_t q = 0; q + 1 < n; q += 2) {
            float e = 0.f;
            HIPCHK(hipEventElapsedTime(&e, pool[q], pool[q + 1]));
            tot += e;
            k++;
        }
        if (ms) *ms = k ? tot / k : 0.0;
        if (cnt) *cnt = k;
        return 0;
    };
    if (int rc = avg(ev_tp, nev_tp, ms_tp, n_tp)) return rc;
    if (int rc = avg(ev_pl, nev_pl, ms_pl, n_pl)) return rc;
    return 0;
}

int Ensemble::time_kernels(int reps, double *ms_tp, double *ms_pl)
{
    if (!tiles_ok) return fail(SWG_ERR_HIP, "the tile arrays are missing after a failed allocation; rebuild them first");
    // Launch the two acceleration kernels on copies of the velocity state so
    // that the timing runs leave the system unchanged.
    hipStream_t st = stream;
    if (!save_pl.x && (alloc_vec(save_pl, npl) || alloc_vec(save_tp, ntp))) return SWG_ERR_HIP;
    for (int d = 0; d < 3; d++) {
        double *src = d == 0 ? pl.vb.x : d == 1 ? pl.vb.y : pl.vb.z;
        double *dst = d == 0 ? save_pl.x : d == 1 ? save_pl.y : save_pl.z;
        HIPCHK(hipMemcpyAsync(dst, src, npl * sizeof(double), hipMemcpyDeviceToDevice, st));
        src = d == 0 ? tp.vb.x : d == 1 ? tp.vb.y : tp.vb.z;
        dst = d == 0 ? save_tp.x : d == 1 ? save_tp.y : save_tp.z;
        HIPCHK(hipMemcpyAsync(dst, src, ntp * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    const int use_radius = prm.integrator == SWG_SYMBA || prm.lclose;
    float ms = 0.f;
    if (ms_tp) {
        launch_kick_tp(st, tune, tp, pl, sy, pt, d_tiles_tp, ntiles_tp, 0.0, 1, 0);  // warm-up
        HIPCHK(hipEventRecord(ev0, st));
        for (int r = 0; r < reps; r++) launch_kick_tp(st, tune, tp, pl, sy, pt, d_tiles_tp, ntiles_tp, 0.0, 1, 0);
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipEventSynchronize(ev1));
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        *ms_tp = ms / reps;
    }
    if (ms_pl) {
        launch_kick_pl(st, tune, pl, sy, pp, d_tiles_pl, ntiles_pl, 0.0, 1, use_radius, 0);  // warm-up
        HIPCHK(hipEventRecord(ev0, st));
        for (int r = 0; r < reps; r++)
            launch_kick_pl(st, tune, pl, sy, pp, d_tiles_pl, ntiles_pl, 0.0, 1, use_radius, 0);
        HIPCHK(hipEventRecord(ev1, st));
        HIPCHK(hipEventSynchronize(ev1));
        HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
        *ms_pl = ms / reps;
    }
    for (int d = 0; d < 3; d++) {
        double *dst = d == 0 ? pl.vb.x : d == 1 ? pl.vb.y : pl.vb.z;
        double *src = d == 0 ? save_pl.x : d == 1 ? save_pl.y : save_pl.z;
        HIPCHK(hipMemcpyAsync(dst, src, npl * sizeof(double), hipMemcpyDeviceToDevice, st));
        dst = d == 0 ? tp.vb.x : d == 1 ? tp.vb.y : tp.vb.z;
        src = d == 0 ? save_tp.x : d == 1 ? save_tp.y : save_tp.z;
        HIPCHK(hipMemcpyAsync(dst, src, ntp * sizeof(double), hipMemcpyDeviceToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipGetLastError());
    return 0;
}

// ---------------------------------------------------------------------------
// Slots: an ensemble whose systems are fixed-capacity slots that take new
// systems while the others keep running (continuous batching).
// ---------------------------------------------------------------------------

// Replace system s by a new one of npl_new <= capacity massive bodies and
// ntp_new <= capacity test particles (initial conditions as for create), run
// the reference's initialize sequence for it alone, and put it back into the
// steps.  The other systems are not touched.
int Ensemble::load(int s, int npl_new, int ntp_new, double GMcb_new, const double *pl_rh, const double *pl_vh,
                   const double *Gmass, const double *radius, const double *rhill_in, const double *tp_rh,
                   const double *tp_vh)
{
    if (prm.integrator == SWG_WHM) return fail(SWG_ERR_ARG, "slots hold HELIO or SyMBA systems");
    if (s < 0 || s >= nsys) return fail(SWG_ERR_ARG, "no such slot");
    const int p0 = h_pl_off[s], cap_pl = h_pl_off[s + 1] - p0;
    const int t0 = h_tp_off[s], cap_tp = h_tp_off[s + 1] - t0;
    if (npl_new < 0 || ntp_new < 0 || npl_new > cap_pl || ntp_new > cap_tp)
        return fail(SWG_ERR_ARG, "system of " + std::to_string(npl_new) + " + " + std::to_string(ntp_new) +
                                     " bodies does not fit slot " + std::to_string(s) + " (" + std::to_string(cap_pl) +
                                     " + " + std::to_string(cap_tp) + ")");
    if (!(GMcb_new > 0.0)) return fail(SWG_ERR_ARG, "central body GM must be positive");
    hipStream_t st = stream;
    HIPCHK(hipStreamSynchronize(st));
    // host side of the initialize sequence, as in create: rhill, then the
    // bodies sorted by mass, descending (helio_util.f90:26)
    std::vector<double> rh(3 * (size_t)cap_pl), vh(3 * (size_t)cap_pl, 0.0), gm(cap_pl, 0.0), rad(cap_pl, 0.0),
        rhl(cap_pl, 0.0), rhill(npl_new ? npl_new : 1);
    if (rhill_in && prm.rhill_present)
        std::copy(rhill_in, rhill_in + npl_new, rhill.begin());
    else
        host_set_rhill(GMcb_new, npl_new, pl_rh, pl_vh, Gmass, rhill.data());
    std::vector<int> idx(npl_new);
    std::iota(idx.begin(), idx.end(), 0);
    std::stable_sort(idx.begin(), idx.end(), [&](int u, int v) { return Gmass[u] > Gmass[v]; });
    std::vector<int32_t> status(cap_pl, INACTIVE), tstatus(cap_tp, INACTIVE);
    for (int q = 0; q < cap_pl; q++) {
        if (q < npl_new) {
            const int src = idx[q];
            for (int d = 0; d < 3; d++) {
                rh[3 * (size_t)q + d] = pl_rh[3 * (size_t)src + d];
                vh[3 * (size_t)q + d] = pl_vh[3 * (size_t)src + d];
            }
            gm[q] = Gmass[src];
            rad[q] = radius[src];
            rhl[q] = rhill[src];
            status[q] = ACTIVE;
            h_pl_id[p0 + q] = src + 1;
        } else {
            pad_position(q - npl_new, &rh[3 * (size_t)q]);
            h_pl_id[p0 + q] = 0;
        }
        h_Gm[p0 + q] = gm[q];
    }
    std::vector<double> trh(3 * (size_t)cap_tp), tvh(3 * (size_t)cap_tp, 0.0);
    for (int q = 0; q < cap_tp; q++) {
        if (q < ntp_new) {
            for (int d = 0; d < 3; d++) {
                trh[3 * (size_t)q + d] = tp_rh[3 * (size_t)q + d];
                tvh[3 * (size_t)q + d] = tp_vh[3 * (size_t)q + d];
            }
            tstatus[q] = ACTIVE;
        } else {
            double pr[3];
            pad_position(q - ntp_new, pr);
            trh[3 * (size_t)q] = -pr[0];
            trh[3 * (size_t)q + 1] = pr[1];
            trh[3 * (size_t)q + 2] = pr[2];
        }
    }
    // the slot's ranges of the SoA arrays
    auto put3 = [&](Vec3p v, int off, const std::vector<double> &aos, int n) -> int {
        if (!n) return 0;
        std::vector<double> c(3 * (size_t)n);
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) c[(size_t)d * n + i] = aos[3 * (size_t)i + d];
        HIPCHK(copy_wait(st, v.x + off, c.data(), n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(st, v.y + off, c.data() + n, n * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(st, v.z + off, c.data() + 2 * (size_t)n, n * sizeof(double), hipMemcpyHostToDevice));
        return 0;
    };
    // rbeg / rend of the padding are read as partner positions by the
    // particle kicks, so the whole range gets the positions
    for (Vec3p v : {pl.rh, pl.rbeg, pl.rend})
        if (int rc = put3(v, p0, rh, cap_pl)) return rc;
    if (int rc = put3(pl.vh, p0, vh, cap_pl)) return rc;
    if (int rc = put3(tp.rh, t0, trh, cap_tp)) return rc;
    if (int rc = put3(tp.vh, t0, tvh, cap_tp)) return rc;
    if (cap_pl) {
        HIPCHK(copy_wait(st, pl.Gm + p0, gm.data(), cap_pl * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(st, pl.R + p0, rad.data(), cap_pl * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(st, pl.rhill + p0, rhl.data(), cap_pl * sizeof(double), hipMemcpyHostToDevice));
        HIPCHK(copy_wait(st, pl.status + p0, status.data(), cap_pl * sizeof(int32_t), hipMemcpyHostToDevice));
    }
    if (cap_tp) HIPCHK(copy_wait(st, tp.status + t0, tstatus.data(), cap_tp * sizeof(int32_t), hipMemcpyHostToDevice));
    // per-system scalars as create leaves them
    const double z3[3] = {0.0, 0.0, 0.0};
    const int32_t one = 1, zero = 0;
    h_GMcb[s] = GMcb_new;
    HIPCHK(copy_wait(st, sy.GMcb + s, &GMcb_new, sizeof(double), hipMemcpyHostToDevice));
    HIPCHK(copy_wait(st, sy.ptbeg + 3 * s, z3, sizeof z3, hipMemcpyHostToDevice));
    HIPCHK(copy_wait(st, sy.ptend + 3 * s, z3, sizeof z3, hipMemcpyHostToDevice));
    HIPCHK(copy_wait(st, sy.lfirst + s, &one, sizeof one, hipMemcpyHostToDevice));
    for (int32_t *q : {sy.enc_pp, sy.enc_pt, sy.flag_pp, sy.flag_pt, sy.err, sy.frozen})
        HIPCHK(copy_wait(st, q + s, &zero, sizeof zero, hipMemcpyHostToDevice));
    HIPCHK(copy_wait(st, d_active + s, &one, sizeof one, hipMemcpyHostToDevice));
    if (dis_ready) {
        // pericentre-passage state of the discard checks starts afresh
        std::vector<int32_t> ones(std::max(std::max(cap_pl, cap_tp), 1), 1);
        std::vector<double> zd(std::max(std::max(cap_pl, cap_tp), 1), 0.0);
        if (cap_pl) {
            HIPCHK(copy_wait(st, dis_pl_isperi + p0, ones.data(), cap_pl * sizeof(int32_t), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(st, dis_pl_atp + p0, zd.data(), cap_pl * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(st, dis_pl_peri + p0, zd.data(), cap_pl * sizeof(double), hipMemcpyHostToDevice));
        }
        if (cap_tp) {
            HIPCHK(copy_wait(st, dis_tp_isperi + t0, ones.data(), cap_tp * sizeof(int32_t), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(st, dis_tp_atp + t0, zd.data(), cap_tp * sizeof(double), hipMemcpyHostToDevice));
            HIPCHK(copy_wait(st, dis_tp_peri + t0, zd.data(), cap_tp * sizeof(double), hipMemcpyHostToDevice));
        }
        HIPCHK(copy_wait(st, dis_lfirst + s, &one, sizeof one, hipMemcpyHostToDevice));
    }
    {
        std::vector<int32_t> tiny(std::max(cap_pl, 1), 0);
        int32_t m = npl_new;
        if (prm.integrator == SWG_SYMBA && prm.gmtiny > 0.0) {
            m = 0;
            for (int q = 0; q < npl_new; q++) {
                tiny[q] = gm[q] < prm.gmtiny;
                m += !tiny[q];
            }
        }
        if (cap_pl) HIPCHK(copy_wait(st, pl.tiny + p0, tiny.data(), cap_pl * sizeof(int32_t), hipMemcpyHostToDevice));
        // the padding sits after the system's own bodies: as partners of
        // the fully interacting rows they add exact zeros
        HIPCHK(copy_wait(st, sy.nplm + s, &m, sizeof m, hipMemcpyHostToDevice));
        if (m < npl_new) {
            any_tiny = true;
            tune.any_tiny = 1;
        }
    }
    // device part of the initialize sequence for this system alone
    launch_init_one(st, pl, tp, sy, s);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    h_nreal_pl[s] = npl_new;
    h_nreal_tp[s] = ntp_new;
    if (merge_on) {
        // the new system's masses (Gmass / GU) and ids
        for (int q = 0; q < cap_pl; q++) h_mass[p0 + q] = q < npl_new ? gm[q] / GU : 0.0;
        merge_reset_system(s);
    }
    if (h_frozen[s]) nfrozen--;
    h_frozen[s] = 0;
    return build_tiles();
}

// The state of the system slot s holds (its real bodies, in storage order).
int Ensemble::get_slot(int s, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh,
                       double *tp_vb, int32_t *pl_status, int32_t *tp_status, int32_t *pl_id)
{
    if (s < 0 || s >= nsys) return fail(SWG_ERR_ARG, "no such slot");
    hipStream_t st = stream;
    HIPCHK(hipStreamSynchronize(st));
    const int p0 = h_pl_off[s], n = h_nreal_pl[s], t0 = h_tp_off[s], m = h_nreal_tp[s];
    auto sub = [](Vec3p v, int off) { return Vec3p{v.x + off, v.y + off, v.z + off}; };
    if (pl_rh && n && download_aos(st, sub(pl.rh, p0), pl_rh, n)) return SWG_ERR_HIP;
    if (pl_vh && n && download_aos(st, sub(pl.vh, p0), pl_vh, n)) return SWG_ERR_HIP;
    if (pl_vb && n && download_aos(st, sub(pl.vb, p0), pl_vb, n)) return SWG_ERR_HIP;
    if (tp_rh && m && download_aos(st, sub(tp.rh, t0), tp_rh, m)) return SWG_ERR_HIP;
    if (tp_vh && m && download_aos(st, sub(tp.vh, t0), tp_vh, m)) return SWG_ERR_HIP;
    if (tp_vb && m && download_aos(st, sub(tp.vb, t0), tp_vb, m)) return SWG_ERR_HIP;
    if (pl_status && n) HIPCHK(copy_wait(st, pl_status, pl.status + p0, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (tp_status && m) HIPCHK(copy_wait(st, tp_status, tp.status + t0, m * sizeof(int32_t), hipMemcpyDeviceToHost));
    if (pl_id) std::copy(h_pl_id.begin() + p0, h_pl_id.begin() + p0 + n, pl_id);
    return 0;
}

// Accelerations of the two direct sums at the current positions, with the
// kernels and tiles the steps use: every massive body's full row with the
// radius test (no encounter-pair removal), every test particle's sum over
// the massive bodies at their current positions.  The kick kernels run with
// dt = 1 on zeroed velocities (vb = fma(a, 1, 0) = a), the velocities are
// read back and restored.
int Ensemble::accel(double *pl_acc, double *tp_acc)
{
    if (!tiles_ok) return fail(SWG_ERR_HIP, "the tile arrays are missing after a failed allocation; rebuild them first");
    hipStream_t st = stream;
    if (!save_pl.x && (alloc_vec(save_pl, npl) || alloc_vec(save_tp, ntp))) return SWG_ERR_HIP;
    const Vec3p vs[2] = {pl.vb, tp.vb}, sv[2] = {save_pl, save_tp};
    const int ns[2] = {npl, ntp};
    for (int k = 0; k < 2; k++) {
        if (!ns[k]) continue;
        const size_t b = ns[k] * sizeof(double);
        HIPCHK(hipMemcpyAsync(sv[k].x, vs[k].x, b, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(sv[k].y, vs[k].y, b, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(sv[k].z, vs[k].z, b, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemsetAsync(vs[k].x, 0, b, st));
        HIPCHK(hipMemsetAsync(vs[k].y, 0, b, st));
        HIPCHK(hipMemsetAsync(vs[k].z, 0, b, st));
    }
    const int use_radius = prm.integrator == SWG_SYMBA || prm.lclose;
    launch_kick_pl(st, tune, pl, sy, pp, d_tiles_pl, ntiles_pl, 1.0, 1, use_radius, 0);
    launch_kick_tp(st, tune, tp, pl, sy, pt, d_tiles_tp, ntiles_tp, 1.0, 1, 0);
    HIPCHK(hipGetLastError());
    if (pl_acc && npl && download_aos(st, pl.vb, pl_acc, npl)) return SWG_ERR_HIP;
    if (tp_acc && ntp && download_aos(st, tp.vb, tp_acc, ntp)) return SWG_ERR_HIP;
    for (int k = 0; k < 2; k++) {
        if (!ns[k]) continue;
        const size_t b = ns[k] * sizeof(double);
        HIPCHK(hipMemcpyAsync(vs[k].x, sv[k].x, b, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(vs[k].y, sv[k].y, b, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(vs[k].z, sv[k].z, b, hipMemcpyDeviceToDevice, st));
    }
    HIPCHK(hipStreamSynchronize(st));
    return 0;
}

// Time the level-0 pl-pl check of the current state (the configured broad
// phase plus the list it produces), with the host synchronizations it
// contains; the stored level-0 list is rebuilt at the next step anyway.
int Ensemble::time_encounters(int reps, double *ms, int64_t *npairs, int64_t *ncand)
{
    if (prm.integrator != SWG_SYMBA) return fail(SWG_ERR_ARG, "encounter checks run in SyMBA only");
    const double dt = prm.dt;
    launch_reset_levels(stream, pl, tp, npl, ntp);  // the step's start: levels and level-0 renc
    if (int rc = prm.encounter_check == 1 ? detect_pp_sas(dt) : detect_pp(dt)) return rc;  // warm-up
    HIPCHK(hipStreamSynchronize(stream));
    auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < reps; r++)
        if (int rc = prm.encounter_check == 1 ? detect_pp_sas(dt) : detect_pp(dt)) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    auto t1 = std::chrono::steady_clock::now();
    if (ms) *ms = std::chrono::duration<double, std::milli>(t1 - t0).count() / reps;
    if (npairs) *npairs = npp;
    if (ncand) {
        *ncand = -1;
        if (prm.encounter_check == 1) {
            HIPCHK(copy_wait(stream, ncand, d_total, sizeof(int64_t), hipMemcpyDeviceToHost));
        }
    }
    return 0;
}

// Leave systems out of the following steps (frozen[s] = 1) or put them back.
// A frozen system's state stays exactly as it is; the other systems' steps
// are unaffected.
int Ensemble::set_frozen(const int32_t *frozen)
{
    if (prm.integrator == SWG_WHM) return fail(SWG_ERR_ARG, "freezing systems is supported for HELIO and SyMBA");
    nfrozen = 0;
    std::vector<int32_t> act(nsys);
    for (int s = 0; s < nsys; s++) {
        h_frozen[s] = frozen[s] ? 1 : 0;
        act[s] = !h_frozen[s];
        nfrozen += h_frozen[s];
    }
    HIPCHK(hipMemcpyAsync(sy.frozen, h_frozen.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(d_active, act.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    HIPCHK(hipStreamSynchronize(stream));
    return build_tiles();
}

// Pick the workgroup size of the two pair kernels (and with it the body
// tiles) by timing each candidate on the current state.
int Ensemble::autotune(int reps, int32_t *pl_block, int32_t *tp_block)
{
    const int cand[3] = {64, 128, 256};
    if (tune.pair_fast) {
        // fast pair mode: 64-body massive tiles shared by 2 ... 16
        // wavefronts, and the particle tile size
        const int split[4] = {2, 4, 8, 16};
        double best_tp = 1e300, best_pl = 1e300;
        int btp = tune.kick_tp_block, bsp = tune.kick_pl_split;
        for (int k = 0; k < 4; k++) {
            tune.kick_pl_split = split[k];
            if (k < 3) tune.kick_tp_block = cand[k];
            if (int rc = build_tiles()) return rc;
            double tms = 0.0, pms = 0.0;
            if (int rc = time_kernels(reps, (ntp && k < 3) ? &tms : nullptr, npl ? &pms : nullptr)) return rc;
            if (ntp && k < 3 && tms < best_tp) {
                best_tp = tms;
                btp = cand[k];
            }
            if (npl && pms < best_pl) {
                best_pl = pms;
                bsp = split[k];
            }
        }
        tune.kick_tp_block = btp;
        tune.kick_pl_split = bsp;
        if (int rc = build_tiles()) return rc;
        if (pl_block) *pl_block = bsp;
        if (tp_block) *tp_block = btp;
        return 0;
    }
    // exact mode: the pl-tp tile size, and for the massive bodies one lane
    // per body with 64, 128 or 256-body tiles or 64-body tiles shared by
    // 2, 4 or 8 evaluating wavefronts (reported as -2, -4, -8)
    const int evals[3] = {2, 4, 8};
    double best_tp = 1e300, best_pl = 1e300;
    int btp = tune.kick_tp_block, bpl = tune.kick_pl_block, bev = tune.kick_pl_eval;
    for (int k = 0; k < 6; k++) {
        if (k < 3) {
            tune.kick_tp_block = cand[k];
            tune.kick_pl_block = cand[k];
            tune.kick_pl_eval = 0;
        } else {
            tune.kick_pl_eval = evals[k - 3];
        }
        if (int rc = build_tiles()) return rc;
        double tms = 0.0, pms = 0.0;
        if (int rc = time_kernels(reps, (ntp && k < 3) ? &tms : nullptr, npl ? &pms : nullptr)) return rc;
        if (ntp && k < 3 && tms < best_tp) {
            best_tp = tms;
            btp = cand[k];
        }
        if (npl && pms < best_pl) {
            best_pl = pms;
            bpl = tune.kick_pl_block;
            bev = tune.kick_pl_eval;
        }
    }
    tune.kick_tp_block = btp;
    tune.kick_pl_block = bpl;
    tune.kick_pl_eval = bev;
    if (int rc = build_tiles()) return rc;
    if (pl_block) *pl_block = bev ? -bev : bpl;
    if (tp_block) *tp_block = btp;
    return 0;
}

}  // namespace swg

// ===========================================================================
// C ABI
// ===========================================================================
using namespace swg;

struct swg_system {
    swg::Ensemble *e;
};

extern "C" {

const char *swg_last_error(void) { return g_err.c_str(); }

int swg_init(int device, swg_ctx **ctx)
{
    if (!ctx) return fail(SWG_ERR_ARG, "ctx is NULL");
    *ctx = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(SWG_ERR_ARG, "no such device");
    HIPCHK(hipSetDevice(device));
    if (int rc = swg::noalloc_setup(device)) return rc;
    swg_ctx *c = new swg_ctx();
    c->device = device;
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete c;
        return hipfail(e, "hipStreamCreate");
    }
    c->leaf = new swg::LeafState();
    c->pair_mode = SWG_PAIR_EXACT;
    *ctx = c;
    return 0;
}

int swg_finalize(swg_ctx *ctx)
{
    if (!ctx) return 0;
    (void)hipSetDevice(ctx->device);
    delete ctx->leaf;
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return 0;
}

int swg_system_create(swg_ctx *ctx, const swg_params *prm, int32_t nsys, const int32_t *npl, const int32_t *ntp,
                      const double *GMcb, const double *pl_rh, const double *pl_vh, const double *Gmass,
                      const double *radius, const double *rhill, const double *tp_rh, const double *tp_vh,
                      swg_system **sys)
{
    if (!ctx || !prm || !sys || !npl || !ntp || !GMcb) return fail(SWG_ERR_ARG, "NULL argument");
    *sys = nullptr;
    HIPCHK(hipSetDevice(ctx->device));
    std::lock_guard<std::mutex> lk(swg::g_capture_mu);
    swg::Ensemble *e = new swg::Ensemble();
    e->stream = ctx->stream;
    int rc = e->create(*prm, nsys, npl, ntp, GMcb, pl_rh, pl_vh, Gmass, radius, rhill, tp_rh, tp_vh);
    if (rc) {
        delete e;
        return rc;
    }
    *sys = new swg_system{e};
    return 0;
}

int swg_system_run(swg_system *sys, int64_t nsteps)
{
    if (!sys || nsteps < 0) return fail(SWG_ERR_ARG, "bad argument");
    return sys->e->run(nsteps);
}

int swg_system_step(swg_system *sys)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    int rc = sys->e->step();
    if (rc) return rc;
    return sys->e->check_errors();
}

int swg_system_get(swg_system *sys, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh,
                   double *tp_vb, int32_t *pl_status, int32_t *tp_status, int32_t *pl_id)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    return sys->e->get(pl_rh, pl_vh, pl_vb, tp_rh, tp_vh, tp_vb, pl_status, tp_status, pl_id);
}

int swg_system_get_misc(swg_system *sys, double *rhill, double *cb_vb, double *ptbeg, double *ptend, int32_t *lfirst)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    return sys->e->get_misc(rhill, cb_vb, ptbeg, ptend, lfirst);
}

int swg_system_stats(swg_system *sys, swg_stats *st)
{
    if (!sys || !st) return fail(SWG_ERR_ARG, "NULL argument");
    swg::Ensemble *e = sys->e;
    int32_t dev[4] = {0, 0, 0, 0};
    HIPCHK(hipStreamSynchronize(e->stream));
    HIPCHK(copy_wait(e->stream, dev, e->d_sub_stats, sizeof dev, hipMemcpyDeviceToHost));
    *st = e->st_;
    st->nsubsteps += dev[0];  // substeps run by the persistent recursion kernel
    if (dev[1] > st->max_depth) st->max_depth = dev[1];
    return 0;
}

int swg_system_enc_list(swg_system *sys, int32_t which, int64_t *n, int32_t *index1, int32_t *index2)
{
    if (!sys || !n || (which != 0 && which != 1)) return fail(SWG_ERR_ARG, "bad argument");
    return sys->e->enc_list(which, n, index1, index2);
}

int swg_system_enc_status(swg_system *sys, int32_t which, int64_t *n, int32_t *status)
{
    if (!sys || !n || (which != 0 && which != 1)) return fail(SWG_ERR_ARG, "bad argument");
    return sys->e->enc_status(which, n, status);
}

int swg_system_time_kernels(swg_system *sys, int32_t reps, double *ms_kick_tp, double *ms_kick_pl)
{
    if (!sys || reps <= 0) return fail(SWG_ERR_ARG, "bad argument");
    return sys->e->time_kernels(reps, ms_kick_tp, ms_kick_pl);
}

int swg_system_create_slots(swg_ctx *ctx, const swg_params *prm, int32_t nslots, const int32_t *npl_cap,
                            const int32_t *ntp_cap, swg_system **sys)
{
    if (!ctx || !prm || !sys || !npl_cap || !ntp_cap || nslots <= 0) return fail(SWG_ERR_ARG, "bad argument");
    if (prm->integrator == SWG_WHM) return fail(SWG_ERR_ARG, "slots hold HELIO or SyMBA systems");
    *sys = nullptr;
    int64_t np = 0, nt = 0;
    for (int s = 0; s < nslots; s++) {
        if (npl_cap[s] < 0 || ntp_cap[s] < 0) return fail(SWG_ERR_ARG, "negative slot capacity");
        np += npl_cap[s];
        nt += ntp_cap[s];
    }
    if (np > INT32_MAX || nt > INT32_MAX) return fail(SWG_ERR_ARG, "slot capacities too large");
    // every slot starts as padding only, with no real body
    std::vector<double> rh(3 * (size_t)std::max<int64_t>(np, 1)), trh(3 * (size_t)std::max<int64_t>(nt, 1));
    std::vector<double> vh(rh.size(), 0.0), tvh(trh.size(), 0.0), z(std::max<int64_t>(np, 1), 0.0);
    std::vector<double> GM(nslots, 1.0);
    for (int s = 0, p = 0, t = 0; s < nslots; s++) {
        for (int q = 0; q < npl_cap[s]; q++, p++) pad_position(q, &rh[3 * (size_t)p]);
        for (int q = 0; q < ntp_cap[s]; q++, t++) {
            pad_position(q, &trh[3 * (size_t)t]);
            trh[3 * (size_t)t] = -trh[3 * (size_t)t];
        }
    }
    swg_params p = *prm;
    p.rhill_present = 1;
    int rc = swg_system_create(ctx, &p, nslots, npl_cap, ntp_cap, GM.data(), rh.data(), vh.data(), z.data(), z.data(),
                               z.data(), trh.data(), tvh.data(), sys);
    if (rc) return rc;
    swg::Ensemble *e = (*sys)->e;
    e->slot_pool = true;
    e->prm.rhill_present = prm->rhill_present;
    std::vector<int32_t> inact(std::max<int64_t>(std::max(np, nt), 1), INACTIVE), fr(nslots, 1);
    for (int s = 0; s < nslots; s++) {
        e->h_nreal_pl[s] = 0;
        e->h_nreal_tp[s] = 0;
    }
    if (np) rc = hipfail_or_ok(copy_wait(e->stream, e->pl.status, inact.data(), np * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!rc && nt) rc = hipfail_or_ok(copy_wait(e->stream, e->tp.status, inact.data(), nt * sizeof(int32_t), hipMemcpyHostToDevice));
    if (!rc) rc = e->set_frozen(fr.data());
    if (rc) {
        swg_system_destroy(*sys);
        *sys = nullptr;
    }
    return rc;
}

int swg_system_load(swg_system *sys, int32_t slot, int32_t npl, int32_t ntp, double GMcb, const double *pl_rh,
                    const double *pl_vh, const double *Gmass, const double *radius, const double *rhill,
                    const double *tp_rh, const double *tp_vh)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    if ((npl && (!pl_rh || !pl_vh || !Gmass || !radius)) || (ntp && (!tp_rh || !tp_vh)))
        return fail(SWG_ERR_ARG, "NULL array");
    std::lock_guard<std::mutex> lk(swg::g_capture_mu);
    return sys->e->load(slot, npl, ntp, GMcb, pl_rh, pl_vh, Gmass, radius, rhill, tp_rh, tp_vh);
}

int swg_system_get_slot(swg_system *sys, int32_t slot, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh,
                        double *tp_vh, double *tp_vb, int32_t *pl_status, int32_t *tp_status, int32_t *pl_id)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    return sys->e->get_slot(slot, pl_rh, pl_vh, pl_vb, tp_rh, tp_vh, tp_vb, pl_status, tp_status, pl_id);
}

int swg_system_slot_counts(swg_system *sys, int32_t slot, int32_t *npl, int32_t *ntp)
{
    if (!sys || !npl || !ntp) return fail(SWG_ERR_ARG, "bad argument");
    swg::Ensemble *e = sys->e;
    if (slot < 0 || slot >= e->nsys) return fail(SWG_ERR_ARG, "no such slot");
    *npl = e->h_nreal_pl[slot];
    *ntp = e->h_nreal_tp[slot];
    return 0;
}

int swg_system_accel(swg_system *sys, double *pl_acc, double *tp_acc)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    if (sys->e->prm.integrator == SWG_WHM) return fail(SWG_ERR_ARG, "accelerations are evaluated for HELIO and SyMBA");
    return sys->e->accel(pl_acc, tp_acc);
}

int swg_system_set_pair_mode(swg_system *sys, int32_t mode)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    if (mode != SWG_PAIR_EXACT && mode != SWG_PAIR_FAST) return fail(SWG_ERR_ARG, "unknown pair mode");
    swg::Ensemble *e = sys->e;
    HIPCHK(hipStreamSynchronize(e->stream));
    e->prm.pair_mode = mode;
    e->tune.pair_fast = mode == SWG_PAIR_FAST;
    return e->build_tiles();
}

int swg_set_pair_mode(swg_ctx *ctx, int32_t mode)
{
    if (!ctx) return fail(SWG_ERR_ARG, "ctx is NULL");
    if (mode != SWG_PAIR_EXACT && mode != SWG_PAIR_FAST) return fail(SWG_ERR_ARG, "unknown pair mode");
    ctx->pair_mode = mode;
    return 0;
}

int swg_system_energy(swg_system *sys, const double *cb_mass, const double *cb_radius, const double *pl_mass,
                      double *out)
{
    if (!sys || !cb_mass || !cb_radius || !out || (!pl_mass && sys->e->npl))
        return fail(SWG_ERR_ARG, "NULL argument");
    for (int s = 0; s < sys->e->nsys; s++)
        if (!(cb_radius[s] > 0.0)) return fail(SWG_ERR_ARG, "central body radius must be positive");
    return sys->e->energy(cb_mass, cb_radius, pl_mass, out);
}

int swg_system_discard(swg_system *sys, const swg_discard_params *dp, const double *cb_radius,
                       int64_t *npl_discarded, int64_t *ntp_discarded)
{
    if (!sys || !dp || !cb_radius) return fail(SWG_ERR_ARG, "NULL argument");
    if (dp->qmin_coord != 0 && dp->qmin_coord != 1) return fail(SWG_ERR_ARG, "qmin_coord must be 0 (HELIO) or 1 (BARY)");
    return sys->e->discard(*dp, cb_radius, npl_discarded, ntp_discarded);
}

int swg_system_time_encounters(swg_system *sys, int32_t reps, double *ms, int64_t *npairs, int64_t *ncand)
{
    if (!sys || reps <= 0) return fail(SWG_ERR_ARG, "bad argument");
    return sys->e->time_encounters(reps, ms, npairs, ncand);
}

int swg_system_use_graphs(swg_system *sys, int32_t enable)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    sys->e->use_graphs = enable != 0;
    if (!enable) sys->e->drop_graphs();
    return 0;
}

int swg_system_set_merge(swg_system *sys, double GU, const double *pl_mass)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    HIPCHK(hipStreamSynchronize(sys->e->stream));
    return sys->e->set_merge(GU, pl_mass);
}

int swg_system_counts(swg_system *sys, int32_t *npl, int32_t *ntp, int64_t *nmerge)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    if (npl) *npl = sys->e->npl;
    if (ntp) *ntp = sys->e->ntp;
    if (nmerge) *nmerge = sys->e->nmerge;
    return 0;
}

int swg_system_pl_offsets(swg_system *sys, int32_t *pl_off)
{
    if (!sys || !pl_off) return fail(SWG_ERR_ARG, "NULL argument");
    std::copy(sys->e->h_pl_off.begin(), sys->e->h_pl_off.end(), pl_off);
    return 0;
}

int swg_system_get_bodies(swg_system *sys, int32_t *id, double *mass, double *Gmass, double *radius)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    return sys->e->get_bodies(id, mass, Gmass, radius);
}

int swg_system_set_frozen(swg_system *sys, const int32_t *frozen)
{
    if (!sys || !frozen) return fail(SWG_ERR_ARG, "NULL argument");
    return sys->e->set_frozen(frozen);
}

int swg_system_autotune(swg_system *sys, int32_t reps, int32_t *pl_block, int32_t *tp_block)
{
    if (!sys || reps <= 0) return fail(SWG_ERR_ARG, "bad argument");
    return sys->e->autotune(reps, pl_block, tp_block);
}

int swg_system_kernel_timing(swg_system *sys, int32_t enable)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    HIPCHK(hipStreamSynchronize(sys->e->stream));
    sys->e->ktiming = enable != 0;
    sys->e->nev_tp = sys->e->nev_pl = 0;
    return 0;
}

int swg_system_kernel_times(swg_system *sys, double *ms_kick_tp, int64_t *n_kick_tp, double *ms_kick_pl,
                            int64_t *n_kick_pl)
{
    if (!sys) return fail(SWG_ERR_ARG, "NULL system");
    return sys->e->kernel_times(ms_kick_tp, n_kick_tp, ms_kick_pl, n_kick_pl);
}

int swg_system_destroy(swg_system *sys)
{
    if (!sys) return 0;
    std::lock_guard<std::mutex> lk(swg::g_capture_mu);
    delete sys->e;
    delete sys;
    return 0;
}

}  // extern "C"
