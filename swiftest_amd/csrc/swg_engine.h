// swg_engine.h -- host-side objects of the library.
#pragma once
#include <functional>
#include <string>
#include <vector>

#include "../../include/swiftest_hip.h"
#include "swg_kernels.h"

namespace swg {

extern thread_local std::string g_err;
int fail(int code, const std::string &msg);
int hipfail(hipError_t e, const char *what);
void host_set_rhill(double GMcb, int n, const double *rh, const double *vh, const double *Gm, double *rhill);

// Device arrays of the WHM integrator (Jacobi coordinates, whm_module.f90).
struct WhmDev {
    Vec3p xj, vj, ah, tah;
    double *eta, *muj, *ir3h, *ir3j;
    int32_t *iflag;
};

// Padding of a slot beyond the system it holds: inactive massive bodies
// with zero mass, radius and Hill radius, and inactive test particles, far
// from everything (10^30 AU, each at its own place) and at rest.  A padding
// body adds exactly nothing to any sum (every term it enters has a zero
// mass factor, fma(0, d, s) = s), is never drifted, kicked or listed in an
// encounter, and sorts after the real bodies.
inline void pad_position(int k, double *r)
{
    r[0] = 1e30;
    r[1] = 1e27 * (double)(k + 1);
    r[2] = 0.0;
}

// A resident ensemble of independent systems.
struct Ensemble {
    hipStream_t stream = nullptr;
    swg_params prm{};
    Tuning tune{};
    int nsys = 0, npl = 0, ntp = 0;
    std::vector<int32_t> h_pl_off, h_tp_off, h_pl_id, h_enc_pp, h_enc_pt, h_err;
    // bodies of each system that are not slot padding (all of them unless
    // the ensemble is a slot pool, swg_system_create_slots)
    std::vector<int32_t> h_nreal_pl, h_nreal_tp;
    bool slot_pool = false;  // made by swg_system_create_slots: slots keep their ranges
    bool any_tiny = false;  // some system has semi-interacting bodies (GMTINY)
    std::vector<double> h_GMcb, h_Gm;

    PlDev pl{};
    TpDev tp{};
    SysDev sy{};
    ListDev pp{}, pt{};
    WhmDev whm{};
    int64_t pp_words = 0, pt_words = 0, npp = 0, npt = 0, cap_pp = 0, cap_pt = 0;

    Tile *d_tiles_pl = nullptr, *d_tiles_tp = nullptr;
    bool tiles_ok = false;  // every tile array built (build_tiles succeeded)
    // WHM whole-step kernel: tiles (system, first / end particle, writer)
    int4 *d_tiles_whm = nullptr;
    // first 64-column block of each system's pl-pl / pl-tp bit matrix
    int32_t *d_cb_pl = nullptr, *d_cb_tp = nullptr;
    int nblk_pl = 0, nblk_tp = 0;
    int ntiles_whm = 0;
    bool whm_fused = false;
    int4 *d_tiles_pp = nullptr, *d_tiles_pt = nullptr;
    int ntiles_pl = 0, ntiles_tp = 0, ntiles_pp = 0;
    int32_t *d_any = nullptr, *d_pt_run = nullptr;
    int64_t *d_total = nullptr;
    void *d_scan = nullptr;
    // recursion below level 0: whole-device working sets, one buffer per
    // depth, until every system's set is at most subtree_max bodies + entries
    // and either fits the LDS-resident workgroup or enough systems have work
    // to fill the device; then one persistent workgroup per system
    BuildScratch bw{};
    std::vector<void *> lv_buf;
    std::vector<int64_t> lv_cap;
    int32_t *d_sub_stats = nullptr, *d_arena = nullptr;
    int64_t *d_arena_off = nullptr;
    void *d_scan_sys = nullptr;
    int64_t arena_words = 0, arena_need = 0;
    int64_t subtree_max = 16384;  // SWG_SUBTREE_MAX; < 0: every level on the whole device
    bool subtree_forced = false;  // SWG_SUBTREE_MAX given: hand off on size alone
    bool level_fits = false;      // every set of the level just built fits the LDS-resident kernel
    double *d_kc = nullptr;       // shell-kick sums kept from one substep's closing kicks for the next's opening ones
    static constexpr int SUBTREE_MIN_SYSTEMS = 128;  // systems with work that fill the device
    LevelArgs level_args() const;
    int build_begin(int dnew, const LevelView *src);
    int build_wait(int32_t &lenc);
    int build_end(int dnew, const LevelView *src, LevelView &v, LevelCounts &c, bool &handoff);
    int recur_level(int d);
    // the working set of each level the recursion is in (rebuilt in place
    // after a merge changes the bodies and the list)
    std::vector<LevelView> lvv;
    std::vector<LevelCounts> lvc;
    int run_subtree(int d0, const LevelView &v);
    // host-driven recursion counters per level (substeps, bodies, entries,
    // hand-offs), reported with SWG_TRACE=1
    int64_t tr_sub[MAXLEV + 2] = {}, tr_B[MAXLEV + 2] = {}, tr_E[MAXLEV + 2] = {}, tr_hand[MAXLEV + 2] = {};
    void trace_levels();
    char *h_small = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev_any = nullptr;
    std::vector<void *> owned;

    swg_stats st_{};
    int64_t iloop = 0, iout = 0;
    bool initial_snapshot_done = false;

    // live timing of the two acceleration kernels: HIP events recorded on the
    // launch stream around every launch while enabled
    bool ktiming = false;
    std::vector<hipEvent_t> ev_tp, ev_pl;
    size_t nev_tp = 0, nev_pl = 0;
    int record_kernel_event(std::vector<hipEvent_t> &pool, size_t &n);
    int kernel_times(double *ms_tp, int64_t *n_tp, double *ms_pl, int64_t *n_pl);

    ~Ensemble();
    int create(const swg_params &p, int nsys, const int32_t *npl, const int32_t *ntp, const double *GMcb,
               const double *pl_rh, const double *pl_vh, const double *Gmass, const double *radius,
               const double *rhill, const double *tp_rh, const double *tp_vh);
    int run(int64_t nsteps);
    int step();
    int snapshot();
    int get(double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh, double *tp_vb,
            int32_t *pl_status, int32_t *tp_status, int32_t *pl_id);
    int get_misc(double *rhill, double *cb_vb, double *ptbeg, double *ptend, int32_t *lfirst);
    int enc_list(int which, int64_t *n, int32_t *i1, int32_t *i2);
    int enc_status(int which, int64_t *n, int32_t *status);
    int time_kernels(int reps, double *ms_tp, double *ms_pl);
    int accel(double *pl_acc, double *tp_acc);
    int load(int s, int npl_new, int ntp_new, double GMcb_new, const double *pl_rh, const double *pl_vh,
             const double *Gmass, const double *radius, const double *rhill, const double *tp_rh, const double *tp_vh);
    int get_slot(int s, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh, double *tp_vb,
                 int32_t *pl_status, int32_t *tp_status, int32_t *pl_id);
    int autotune(int reps, int32_t *pl_block, int32_t *tp_block);
    int set_frozen(const int32_t *frozen);
    int time_encounters(int reps, double *ms, int64_t *npairs, int64_t *ncand);
    int energy(const double *cb_mass, const double *cb_radius, const double *pl_mass, double *out);
    // energy scratch, allocated on the first energy call
    bool en_ready = false;
    Vec3p en_rb{};
    double *en_mass = nullptr, *en_cb = nullptr, *en_part = nullptr, *en_out = nullptr;
    int4 *en_tiles = nullptr;
    int32_t *en_toff = nullptr;
    int en_ntiles = 0;
    int discard(const swg_discard_params &dp, const double *cb_radius, int64_t *npl_dis, int64_t *ntp_dis);
    // discard-check state, allocated on the first discard call
    bool dis_ready = false;
    int32_t *dis_pl_isperi = nullptr, *dis_tp_isperi = nullptr, *dis_lfirst = nullptr;
    double *dis_pl_atp = nullptr, *dis_pl_peri = nullptr, *dis_tp_atp = nullptr, *dis_tp_peri = nullptr;
    double *dis_sys = nullptr, *dis_rcb = nullptr;
    uint8_t *dis_gone = nullptr;
    unsigned long long *dis_count = nullptr;
    int4 *dis_tiles = nullptr;
    int dis_ntiles = 0;
    // HIP graphs of the step's two halves (see run_phase)
    bool use_graphs = false;
    hipGraphExec_t graph_exec[8] = {};
    uint64_t graph_key_cached = 0;
    int capture_failures = 0;  // captures dropped (after 4, the halves always run directly)
    uint64_t graph_key() const;
    void drop_graphs();
    int run_phase(int half, const std::function<int()> &body);
    std::vector<int32_t> h_frozen;
    int nfrozen = 0;
    int32_t *d_active = nullptr;  // 1 - frozen, per system
    Vec3p save_pl{}, save_tp{};  // velocity copies kept across kernel timings
    int check_errors();
    int err_code();

    // collisions resolved by merging inside the step (merge.hip)
    bool merge_on = false;
    double GU = 0.0;
    std::vector<double> h_mass;
    std::vector<int32_t> h_maxid, h_nplm;  // per system: largest id handed out, fully interacting bodies
    int64_t nmerge = 0;
    // a step runs as usual; one that flags a collision is run again from a
    // copy of the state taken at its start, with every level on the whole
    // device and a stop at each substep's collision check (merge_replay)
    bool merge_replay = false;
    Vec3p mb_rh{}, mb_vh{}, mb_vb{};
    int32_t *mb_status = nullptr, *mb_sys = nullptr;  // pl status; per system lfirst, err
    double *mb_cb = nullptr;                           // per system cbvb, ptbeg, ptend
    int step_merge();
    int merge_save(bool restore);
    int set_merge(double GU, const double *pl_mass);
    void merge_reset_system(int s);
    int merge_point(int irec, double dt);
    int resolve_merges(int irec, double dt);
    int resolve_system(void *host_sys, const std::vector<int32_t> &status, double GMcb, int irec, double dt,
                       int32_t &maxid_s, int32_t &nplm_s, int64_t &ncol, double *cbvb);
    int upload_merged(const void *host_systems, const std::vector<char> &hit, const double *cbvb, int irec);
    int get_bodies(int32_t *id, double *mass, double *Gmass, double *radius);

    // internals
    template <class T> int own(T **p, size_t n);
    template <class T> int own2(T **p, size_t n);
    int alloc_vec(Vec3p &v, size_t n);
    int build_tiles();
    int ensure_list(ListDev &L, int64_t n, int64_t &cap);
    int detect_pp(double dt);
    int detect_pp_sas(double dt);
    SasWork sas;
    int detect_pt(double dt, const std::vector<int32_t> &run);
    int build_adjacency();
    int post_any();
    int wait_any(int32_t &v);
    int recur0();
    int step_helio_symba();
    int whm_setup();
    int step_whm();
    int run_whm_fused(int nsteps);
    int whm_snapshot();
};

// The encounter check that allocates nothing (k_leaf.hip): the device
// address of its region on `device`, looked up once with that device current.
int noalloc_setup(int device);

// Scratch of the leaf-operator entry points.
struct LeafState {
    std::vector<void *> bufs;
    std::vector<size_t> sizes;
    ListDev L{};
    int64_t nenc = 0, cap = 0;
    ~LeafState();
    void *get(int slot, size_t bytes, int *rc);
};

}  // namespace swg

// A library context: one device, one stream, leaf-operator scratch.
struct swg_ctx {
    int device;
    hipStream_t stream;
    swg::LeafState *leaf;
    int pair_mode;  // SWG_PAIR_EXACT / SWG_PAIR_FAST for the leaf pair sums
};

namespace swg {
int noalloc_encounter_plpl(swg_ctx *ctx, int32_t npl, const double *r, const double *v, const double *renc, double dt,
                           int64_t cap, int64_t *nenc, int32_t *index1, int32_t *index2, int32_t *lvdotr);
}
