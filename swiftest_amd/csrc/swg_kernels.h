// swg_kernels.h -- host-side launchers of the device kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/swiftest_hip.h"
#include "swg_level.h"
#include "swg_state.h"

namespace swg {

// A copy on the library's own stream, waited for.  The library never uses the
// legacy stream: a synchronous legacy-stream call fails while any stream of
// the process is capturing a graph, and other threads capture theirs.
inline hipError_t copy_wait(hipStream_t st, void *dst, const void *src, size_t n, hipMemcpyKind kind)
{
    const hipError_t e = hipMemcpyAsync(dst, src, n, kind, st);
    return e == hipSuccess ? hipStreamSynchronize(st) : e;
}

// A block of consecutive bodies of one system: bodies [b0, b1) of system s.
struct Tile {
    int32_t s, b0, b1, pad;
};

// Kernel-tuning knobs chosen per launch shape by the autotuner.
struct Tuning {
    int kick_pl_block = 256;   // 64, 128 or 256 threads per workgroup
    int kick_tp_block = 256;
    int pair_fast = 0;         // 1: pair factors from the reciprocal square root (SWG_PAIR_FAST)
    int kick_pl_split = 8;     // fast mode: wavefronts sharing one 64-body tile's rows (1, 2, 4, 8, 16)
    int kick_pl_eval = 0;      // exact mode: 0 one lane per body (k_kick_pl), else the evaluating
                               // wavefronts per 64-body tile of k_kick_pl_split (2, 4, 8)
    int any_tiny = 0;          // some system has semi-interacting bodies: k_kick_pl_tiny for every tile
};
// bodies per massive-body tile of the fast pair kernel
constexpr int KPL_FAST_TILE = 64;

// ---- step kernels (k_step.hip) ----
void launch_reset_levels(hipStream_t st, PlDev pl, TpDev tp, int npl, int ntp);
void launch_sums_begin(hipStream_t st, PlDev pl, SysDev sy, int nsys, int force_vh2vb);
void launch_pl_begin_apply(hipStream_t st, PlDev pl, SysDev sy, int npl, double dth, int force_vh2vb);
void launch_kick_pl(hipStream_t st, const Tuning &tu, PlDev pl, SysDev sy, ListDev pp, const Tile *tiles,
                    int ntiles, double dt, int lbeg, int use_radius, int do_pairs);
void launch_drift(hipStream_t st, Vec3p r, Vec3p v, int32_t *status, int32_t *levelg, const int32_t *sys,
                  const double *GMcb, int32_t *err, const int32_t *frozen, int n, double dt, int irec, int use_level,
                  int is_tp);
void launch_tp_pre(hipStream_t st, TpDev tp, SysDev sy, int ntp, double dth);
void launch_kick_tp(hipStream_t st, const Tuning &tu, TpDev tp, PlDev pl, SysDev sy, ListDev pt,
                    const Tile *tiles, int ntiles, double dt, int lbeg, int do_pairs);
void launch_sums_end(hipStream_t st, PlDev pl, SysDev sy, int nsys);
void launch_pl_end_apply(hipStream_t st, PlDev pl, SysDev sy, int npl, double dth);
void launch_tp_post(hipStream_t st, TpDev tp, SysDev sy, int ntp, double dth);
void launch_lfirst_update(hipStream_t st, SysDev sy, int nsys);
void launch_snapshot(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, int nsys, int npl, int ntp);
void launch_init_tp_vb(hipStream_t st, TpDev tp, SysDev sy, int ntp);
void launch_init_one(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, int s);

// ---- energy and angular momentum (k_energy.hip) ----
void energy_tiles(int nsys, const int32_t *pl_off, std::vector<int4> &tiles, std::vector<int32_t> &toff);
void launch_energy(hipStream_t st, PlDev pl, SysDev sy, int nsys, const double *mass, const double *cbm,
                   const double *cbR, int lclose, Vec3p rb, const int4 *tiles, int ntiles, const int32_t *toff,
                   double *part, double *out);
void launch_vh2vb(hipStream_t st, PlDev pl, SysDev sy, int nsys, int npl);

// ---- discard checks (k_discard.hip) ----
void discard_tiles(int nsys, const int32_t *tp_off, std::vector<int4> &tiles);
size_t discard_sys_words();
void launch_discard(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, int nsys, int npl, int symba,
                    const swg_discard_params &dp, int lclose, double dt, const double *Rcb,
                    const int32_t *nplenc_row, const int32_t *nplenc_col, int32_t *pl_isperi, double *pl_atp,
                    double *pl_peri, int32_t *tp_isperi, double *tp_atp, double *tp_peri, int32_t *lfirst_peri,
                    uint8_t *gone, double *sys, unsigned long long *count, const int4 *tiles, int ntiles);

// ---- encounter detection (k_enc.hip) ----
void launch_pp_tiles(hipStream_t st, PlDev pl, SysDev sy, ListDev pp, const int4 *tiles, int ntiles,
                     double dt);
void launch_pt_tiles(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, ListDev pt, const int4 *tiles,
                     int ntiles, double dt);
void launch_rows_count(hipStream_t st, ListDev L, const int32_t *row_sys, int upper, const int32_t *pl_off,
                       const int32_t *run, int nrows);
void launch_scan_i32_i64(hipStream_t st, const int32_t *in, int64_t *out, int n, int64_t *total,
                         void *scratch);
size_t scan_scratch_bytes(int n);
// Exclusive scans of up to four int32 arrays of at most SCAN_SMALL_MAX
// elements each in a single launch; totals go to total[k] and, when given,
// total2[k]; zero[0, nzero) is cleared first.
constexpr int SCAN_SMALL_MAX = 16384;
struct ScanSmall {
    const int32_t *in[4];
    int64_t *out[4];
    int n[4];
    int64_t *total[4], *total2[4];
    int32_t *zero;
    int nzero;
};
void launch_scan_small(hipStream_t st, const ScanSmall &a, int na);
void launch_rows_fill(hipStream_t st, ListDev L, const int32_t *row_sys, int upper, const int32_t *pl_off,
                      const int32_t *col_off, int nrows, int32_t *lev1g = nullptr, int32_t *lev1m = nullptr,
                      int32_t *lev2g = nullptr, int32_t *lev2m = nullptr);
void launch_sys_flags(hipStream_t st, ListDev L, const int32_t *pl_off, int32_t *flag, int nsys);
// index2-side adjacency (counts, then fill) of a list whose bit matrix is in
// L.words; cb_off: each system's first 64-column block (nsys + 1 entries)
void launch_col_t(hipStream_t st, ListDev L, const int32_t *row_off, const int32_t *col_off, const int32_t *cb_off,
                  int nsys, int nblocks, const int32_t *run, int upper, int fill);
// the same arrays for a short pl-pl list, in one workgroup (counts, scan and
// fill over all npl columns); only for npp <= COLS_MAX and npl <= COLS_NPL
constexpr int COLS_MAX = 1024, COLS_NPL = 8192;
void launch_col_small(hipStream_t st, ListDev L, int npl, int npp);

// ---- sort-and-sweep broad phase of the level-0 pl-pl check (k_sas.hip) ----
struct SasWork {
    int n = 0;  // bodies the per-body arrays hold
    uint64_t *lokey = nullptr, *slokey = nullptr;  // lower x-extents as sortable keys, and sorted
    double *srt = nullptr;   // positions, velocities and encounter radii in sorted order (7 x n)
    int8_t *stiny = nullptr; // semi-interacting flags in sorted order
    double *hi = nullptr, *slo = nullptr;
    int32_t *idx = nullptr, *sidx = nullptr, *rank = nullptr, *cnt = nullptr;
    int64_t *coff = nullptr, *nsel = nullptr;
    uint64_t *key = nullptr, *sel = nullptr, *skey = nullptr;
    uint8_t *flag = nullptr;
    int32_t *pid = nullptr, *spid = nullptr, *si2 = nullptr;
    // radix sort and compaction scratch
    void *rtk = nullptr, *rtv = nullptr, *rhist = nullptr, *roffs = nullptr, *rscan = nullptr, *rtotal = nullptr;
    size_t key_cap = 0, flag_cap = 0, sel_cap = 0, skey_cap = 0, pid_cap = 0, spid_cap = 0, si2_cap = 0,
           rtk_cap = 0, rtv_cap = 0, rhist_cap = 0, roffs_cap = 0, rscan_cap = 0, rtotal_cap = 0;
    ~SasWork();
};
// extents, per-system sort, candidate counts and their scan (total in *ncand_dev)
hipError_t sas_candidates(hipStream_t st, SasWork &w, PlDev pl, SysDev sy, int npl, int nsys,
                          const int32_t *active, double dt, int64_t *ncand_dev, void *scan_scratch);
// narrow phase and compaction; returns the number of pairs (synchronizes)
hipError_t sas_narrow(hipStream_t st, SasWork &w, PlDev pl, int npl, int64_t ncand, double dt, int64_t *nsel_host);
// the ordered list, its row counts and the level-0 marks (list capacity >= npp)
hipError_t sas_fill(hipStream_t st, SasWork &w, ListDev &pp, PlDev pl, int npl, int64_t npp);
// index2-side adjacency
hipError_t sas_adjacency(hipStream_t st, SasWork &w, ListDev &pp, int npl, int64_t npp);

// ---- SyMBA recursion, level 0 (k_enc.hip) ----
void launch_list_check_pp(hipStream_t st, ListDev pp, PlDev pl, SysDev sy, int64_t n, int irec, double dt,
                          int32_t *any);
void launch_list_check_pt(hipStream_t st, ListDev pt, PlDev pl, TpDev tp, SysDev sy, int64_t n, int irec,
                          double dt, int short_circuit, int32_t *any);
void launch_collision(hipStream_t st, ListDev pp, ListDev pt, PlDev pl, TpDev tp, SysDev sy, int64_t npp,
                      int64_t npt, int irec, double dt, int32_t *any);
void launch_set_levels(hipStream_t st, PlDev pl, TpDev tp, ListDev pp, ListDev pt, int npl, int ntp,
                       int64_t npp, int64_t npt, int ireci);
void launch_clear_flags(hipStream_t st, SysDev sy, int nsys, int32_t *any);

// ---- SyMBA recursion levels on the whole device (k_level.hip) ----
struct LevelCounts {
    int64_t nB = 0, nE = 0, nBt = 0, nEt = 0;  // pl bodies, pl entries, tp bodies, tp entries
};
struct LevelArgs {
    PlDev pl;
    TpDev tp;
    ListDev pp, pt;
    SysDev sy;
    int nsys, has_pp, has_pt;
};
// temporaries of one level build (sized for all bodies) and its per-system results
struct BuildScratch {
    int32_t *keepB, *cntE, *keepBt, *cntEt;
    int64_t *posB, *posE, *posBt, *posEt;
    int64_t *tot;    // [10]: nB, nE, nBt, nEt of the new level, the arena total, the encounter flag, then maxsz
    void *scan;      // scan_scratch_bytes(max(npl, ntp))
    int32_t *rng;    // [4*nsys] new level's per-system body ranges
    int32_t *caps;   // [nsys] arena words of each system's persistent workgroup
    int32_t *maxsz;  // [8] = (int32_t *)(tot + 6): largest per-system set and its parts, systems with a set
};
LevelView level_view(void *base, const LevelCounts &c);
int64_t level_bytes(const LevelCounts &c);
// src == nullptr: build level 1 from the level-0 adjacency of all npl / ntp bodies
// any (nullable): the encounter flag, copied into w.tot[5] with the counts.
// Returns true when it also scanned the arena sizes into arena_off (and
// w.tot[4]), which it does for up to 1024 systems; otherwise the caller does.
bool launch_build_count(hipStream_t st, const LevelArgs &la, const LevelView *src, int npl, int ntp, int dnew,
                        const BuildScratch &w, const int32_t *any, int64_t *arena_off);
void launch_build_fill(hipStream_t st, const LevelArgs &la, const LevelView *src, int npl, int ntp, int dnew,
                       const BuildScratch &w, void *dst_base, const LevelCounts &c);
void launch_level_check(hipStream_t st, const LevelArgs &la, const LevelView &v, const LevelCounts &c, int d,
                        double dtl, int short_circuit, int32_t *any);
// v == nullptr: the level-0 kicks over the full adjacency of nB / nBt bodies
// keep 0: evaluate the shell kicks; 1: evaluate and keep the pl bodies' sums
// in kc (8 words per body of v); 2: apply the pl sums kept in kc by the
// closing kicks of the previous substep at the same level (the tp kicks are
// evaluated in every mode)
void launch_level_kicks(hipStream_t st, const LevelArgs &la, const LevelView *v, int nB, int nBt, int d, double dth,
                        int keep = 0, double *kc = nullptr);
void launch_level_drift(hipStream_t st, const LevelArgs &la, const LevelView &v, int nB, int nBt, int d, double dtl);
void launch_level_collision(hipStream_t st, const LevelArgs &la, const LevelView &v, int nB, int nBt, double dtl);
// any (non-null): also clear the encounter flags for the next substep
void launch_level_set_levels(hipStream_t st, const LevelArgs &la, const LevelView &v, const LevelCounts &c, int d,
                             int32_t *any = nullptr);

// ---- SyMBA recursion from level d0 down, one workgroup per system (k_subtree.hip) ----
// whether every system's set of this size fits the LDS-resident subtree kernel
bool subtree_lds_fits(int nb, int ne, int nbt, int net);
void subtree_profile_report(hipStream_t st);
void launch_subtree(hipStream_t st, const LevelArgs &la, const LevelView &start, const int32_t *rng, int d0,
                    double dt, int lclose, int short_circuit, int32_t *arena, const int64_t *arena_off,
                    int32_t *stats, bool all_lds);

}  // namespace swg
