/*
 * swiftest_hip.h -- C ABI of the MI355X implementation of Swiftest's per-step
 * hot path.
 *
 * Two layers:
 *
 *  1. Leaf operators: drop-in replacements for the numeric leaves that the
 *     reference's step layer calls.  They take the reference's own host
 *     arrays (Fortran (3,n) real(DP) == C double[n][3]; masks as 4-byte
 *     logicals == int32, nonzero = .true.; pair lists 1-based int32) and
 *     keep the reference's in/out contract.  Each call copies its inputs to
 *     HBM, runs the kernel, and copies the outputs back.
 *
 *  2. Resident systems: the whole WHM / HELIO / SyMBA step (kicks, drifts,
 *     encounter detection, SyMBA recursion) runs on the device with the state
 *     kept in HBM across steps; the host only sees it at output time.  One
 *     handle holds an ensemble of independent systems (a single system is an
 *     ensemble of one).
 *
 * All functions return 0 on success or a negative SWG_ERR_* code; the message
 * of the last error on the calling thread is available from swg_last_error().
 * Nothing here throws.  A context owns one HIP stream; calls on one context
 * are stream-ordered and must come from one host thread at a time.
 */
#ifndef SWIFTEST_HIP_H
#define SWIFTEST_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SWG_OK 0
#define SWG_ERR_ARG (-1)       /* invalid argument */
#define SWG_ERR_HIP (-2)       /* HIP runtime error (no device, out of memory, ...) */
#define SWG_ERR_COLLISION (-3) /* a collision was detected; its resolution is host work */
#define SWG_ERR_DRIFT (-4)     /* a massive body failed the Danby drift */
#define SWG_ERR_DEPTH (-5)     /* recursion step below double precision (symba_step.f90:196-201) */
#define SWG_ERR_SCRATCH (-6)   /* the recursion's working-set arena was too small */

#define SWG_WHM 0
#define SWG_HELIO 1
#define SWG_SYMBA 2

/* Pair factor Gm / r^3 of the direct sums (the pl-pl rows, the pl-tp sums and
 * the encounter-pair corrections at level 0; the recursion's shell kicks
 * always use the exact form):
 *   SWG_PAIR_EXACT  IEEE sqrt and division in the reference's order: every
 *                   sum is bit-identical to a strict evaluation of the
 *                   reference's expressions (the default);
 *   SWG_PAIR_FAST   the hardware reciprocal square root with a second-order
 *                   correction: each factor within a few ulp of the IEEE
 *                   one, sums in the same order, about half the
 *                   instructions per pair.  Not bit-identical. */
#define SWG_PAIR_EXACT 0
#define SWG_PAIR_FAST 1

typedef struct swg_ctx swg_ctx;
typedef struct swg_system swg_system;

/* Run configuration: the reference parameters that change the step
 * (swiftest_io.f90:2285-2760).  short_circuit = 1 reproduces the
 * reference's `pl check .or. tp check` as gfortran -O evaluates it: the pl-tp
 * check is skipped in a system whose pl-pl check found an encounter
 * (symba_step.f90:38, :210). */
typedef struct swg_params {
    double dt;
    int32_t integrator;    /* SWG_WHM, SWG_HELIO or SWG_SYMBA */
    int32_t lclose;        /* CHK_CLOSE */
    int32_t short_circuit;
    int32_t istep_out;     /* snapshot cadence in steps, 0 = none */
    int32_t rhill_present; /* RHILL_PRESENT: keep the given rhill */
    int32_t encounter_check; /* ENCOUNTER_CHECK for pl-pl: 0 TRIANGULAR, 1 SORTSWEEP (same list) */
    int32_t pair_mode;     /* SWG_PAIR_EXACT or SWG_PAIR_FAST (no counterpart in the reference) */
    int32_t reserved;
    double gmtiny;         /* GMTINY (SyMBA): massive bodies with Gmass < gmtiny interact only with the
                              others (semi-interacting, swiftest_kick.f90:190-262, encounter_check.f90:
                              42-109); 0 = every body fully interacting */
} swg_params;

typedef struct swg_stats {
    int64_t nplpl_enc;  /* level-0 pl-pl encounters in the last step (all systems) */
    int64_t npltp_enc;  /* level-0 pl-tp encounters in the last step */
    int32_t max_depth;  /* deepest recursion level entered in the last step, -1 = none */
    int32_t reserved;
    int64_t nsubsteps;  /* recursion substeps in the last step */
    int64_t ncollision; /* collisions detected (cumulative) */
    int64_t ndrifterr;  /* drift failures (cumulative) */
} swg_stats;

int swg_init(int device, swg_ctx **ctx);
int swg_finalize(swg_ctx *ctx);
const char *swg_last_error(void);

/* ---- leaf operators ------------------------------------------------------ */

/* swiftest_kick_getacch_int_all_tri_rad_pl (src/swiftest/swiftest_kick.f90:165)
 * and, with radius == NULL, ..._tri_norad_pl (:274).  acc(3,npl) is accumulated
 * into. */
int swg_kick_getacch_int_all_tri_pl(swg_ctx *ctx, int32_t npl, int32_t nplm, const double *r,
                                    const double *Gmass, const double *radius, double *acc);

/* swiftest_kick_getacch_int_all_flat_rad_pl (swiftest_kick.f90:69) and, with
 * radius == NULL, ..._flat_norad_pl (:118); k_plpl is the 1-based (2,nplpl)
 * pair list. */
int swg_kick_getacch_int_all_flat_pl(swg_ctx *ctx, int32_t npl, int64_t nplpl, const int32_t *k_plpl,
                                     const double *r, const double *Gmass, const double *radius, double *acc);

/* swiftest_kick_getacch_int_all_tp (swiftest_kick.f90:374). */
int swg_kick_getacch_int_all_tp(swg_ctx *ctx, int32_t ntp, int32_t npl, const double *rtp, const double *rpl,
                                const double *GMpl, const int32_t *lmask, double *acc);

/* swiftest_drift_all (src/swiftest/swiftest_drift.f90:60), GR off. */
int swg_drift_all(swg_ctx *ctx, const double *mu, double *x, double *v, int32_t n, double dt,
                  const int32_t *lmask, int32_t *iflag);

/* encounter_check_all_plpl / _pltp, TRIANGULAR (src/encounter/encounter_check.f90:14, :112).
 * The list is computed on the device and kept there; *nenc receives its length.
 * swg_encounter_list_copy() then copies index1, index2 (1-based) and lvdotr
 * into caller arrays of length nenc (the caller allocates, as the Fortran
 * allocatable-out wrapper does). */
int swg_encounter_check_all_plpl(swg_ctx *ctx, int32_t npl, const double *r, const double *v,
                                 const double *renc, double dt, int64_t *nenc);
int swg_encounter_check_all_pltp(swg_ctx *ctx, int32_t npl, int32_t ntp, const double *rpl, const double *vpl,
                                 const double *rtp, const double *vtp, const double *renc, double dt,
                                 int64_t *nenc);
int swg_encounter_list_copy(swg_ctx *ctx, int32_t *index1, int32_t *index2, int32_t *lvdotr);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): swg_encounter_check_all_plpl followed by
 * swg_encounter_list_copy in one call that allocates nothing.  Every device
 * buffer of the check (the staged inputs, the bit matrix, row counts, scan
 * partials and the list) lives in one library-owned region of
 * SWG_NOALLOC_BYTES (6.5 MiB) per GPU, held until the process ends.  Limits:
 * npl <= SWG_NOALLOC_MAX_NPL and at most SWG_NOALLOC_MAX_PAIRS listed pairs,
 * else SWG_ERR_ARG.  cap is the length of the caller's index1 / index2 /
 * lvdotr arrays (any may be NULL); when the list is longer than cap, *nenc
 * receives its length and SWG_ERR_ARG is returned with nothing copied.  The
 * outputs are those of the allocating pair of calls.  Every call on one GPU
 * uses that GPU's region and leaves the last call's contents there: calls
 * that overlap on one GPU (from other threads or contexts) race. */
#define SWG_NOALLOC_MAX_NPL 2048
#define SWG_NOALLOC_MAX_PAIRS 524288
#define SWG_NOALLOC_BYTES 6812160
int swg_encounter_check_all_plpl_noalloc(swg_ctx *ctx, int32_t npl, const double *r, const double *v,
                                         const double *renc, double dt, int64_t cap, int64_t *nenc,
                                         int32_t *index1, int32_t *index2, int32_t *lvdotr);

/* ---- resident systems ---------------------------------------------------- */

/* Create nsys independent systems from initial conditions as the reference
 * reads them (heliocentric rh, vh; Gmass; radius; optional rhill), then run
 * the reference's initialize sequence (sort by mass or by 1/r^3, set_rhill,
 * vh->vb, ...).  Bodies of all systems are concatenated system by system;
 * npl[s] / ntp[s] give the counts. */
int swg_system_create(swg_ctx *ctx, const swg_params *prm, int32_t nsys, const int32_t *npl, const int32_t *ntp,
                      const double *GMcb, const double *pl_rh, const double *pl_vh, const double *Gmass,
                      const double *radius, const double *rhill, const double *tp_rh, const double *tp_vh,
                      swg_system **sys);
/* Driver loop: nsteps steps with a snapshot every istep_out steps
 * (swiftest_driver.f90:130-160). */
int swg_system_run(swg_system *sys, int64_t nsteps);
/* One step, no snapshot. */
int swg_system_step(swg_system *sys);
/* Copy the state to host arrays (any may be NULL).  pl_id gives, for each
 * stored body, its 1-based position among its own system's bodies in the
 * creation input. */
int swg_system_get(swg_system *sys, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh, double *tp_vh,
                   double *tp_vb, int32_t *pl_status, int32_t *tp_status, int32_t *pl_id);
int swg_system_get_misc(swg_system *sys, double *rhill, double *cb_vb, double *ptbeg, double *ptend,
                        int32_t *lfirst);
int swg_system_stats(swg_system *sys, swg_stats *st);
/* Level-0 encounter list of the last step: which = 0 pl-pl, 1 pl-tp;
 * 1-based body indices in storage order.  With index1 == NULL only *n is set. */
int swg_system_enc_list(swg_system *sys, int32_t which, int64_t *n, int32_t *index1, int32_t *index2);
/* Status of each pair of that list after the last step: 0 ACTIVE, -14
 * COLLIDED for a pair the collision check flagged (the encounter list's
 * status component, encounter_module.f90:30, set at collision_check.f90:131
 * and :238).  With status == NULL only *n is set.  A SyMBA step with CHK_CLOSE
 * in which a collision is detected is completed with the colliding pairs
 * flagged (they drop out of the list kicks and checks for the rest of the
 * step) and swg_system_run / swg_system_step then return SWG_ERR_COLLISION;
 * resolving it is host work. */
int swg_system_enc_status(swg_system *sys, int32_t which, int64_t *n, int32_t *status);
/* Time the dominant kernels: launch the pl-tp and pl-pl acceleration kernels
 * reps times on the current state (no state change) and report the average
 * duration of one launch in ms, measured with HIP events on the system's
 * stream. */
int swg_system_time_kernels(swg_system *sys, int32_t reps, double *ms_kick_tp, double *ms_kick_pl);
/* Slot pools for continuous batching (no counterpart in the reference, which
 * runs one system per process).  swg_system_create_slots makes an ensemble of
 * nslots empty slots with room for npl_cap[s] massive bodies and ntp_cap[s]
 * test particles each, all allocated up front; empty slots are frozen.
 * swg_system_load puts a new system (initial conditions as for
 * swg_system_create: heliocentric rh, vh, Gmass, radius, optional rhill) into
 * a slot that is free or whose system has finished, runs the initialize
 * sequence for that system alone and puts the slot into the following steps,
 * while the other slots' systems keep their state.  The rest of a slot's
 * capacity holds padding bodies (inactive, zero mass, far away) that add
 * exactly nothing to any sum, so a system's trajectory in a slot is its
 * trajectory alone, bit for bit.  swg_system_get_slot copies out the state
 * of the system a slot holds (its own bodies only, in storage order; pl_id as
 * in swg_system_get); freezing a slot (swg_system_set_frozen) stops it.
 * HELIO and SyMBA. */
int swg_system_create_slots(swg_ctx *ctx, const swg_params *prm, int32_t nslots, const int32_t *npl_cap,
                            const int32_t *ntp_cap, swg_system **sys);
int swg_system_load(swg_system *sys, int32_t slot, int32_t npl, int32_t ntp, double GMcb, const double *pl_rh,
                    const double *pl_vh, const double *Gmass, const double *radius, const double *rhill,
                    const double *tp_rh, const double *tp_vh);
int swg_system_get_slot(swg_system *sys, int32_t slot, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh,
                        double *tp_vh, double *tp_vb, int32_t *pl_status, int32_t *tp_status, int32_t *pl_id);
/* Bodies the system in a slot has now (a merge, swg_system_set_merge,
 * lowers npl): the sizes of swg_system_get_slot's arrays.  Part of the slot
 * interface above, which has no counterpart in the reference. */
int swg_system_slot_counts(swg_system *sys, int32_t slot, int32_t *npl, int32_t *ntp);
/* Accelerations of the two direct sums at the current positions, computed by
 * the kernels and tiles the steps use (with the system's pair_mode), into
 * host arrays pl_acc[npl][3] and tp_acc[ntp][3] (either may be NULL): each
 * massive body's full row over its system with the radius test
 * (swiftest_kick_getacch_int_all_tri_rad_pl, swiftest_kick.f90:165-271,
 * without the encounter-pair removal), each active test particle's sum over
 * its system's massive bodies (swiftest_kick_getacch_int_all_tp, :374-415).
 * The state is unchanged.  HELIO and SyMBA.  A measurement and test hook
 * with no counterpart in the reference. */
int swg_system_accel(swg_system *sys, double *pl_acc, double *tp_acc);
/* Switch a resident system's pair mode for the following steps (the tiles
 * are rebuilt; the autotuner's choices of the other mode are kept). */
int swg_system_set_pair_mode(swg_system *sys, int32_t mode);
/* Pair mode of the leaf operators called on ctx (SWG_PAIR_EXACT by default);
 * resident systems take theirs from swg_params.pair_mode. */
int swg_set_pair_mode(swg_ctx *ctx, int32_t mode);
/* Replay the step's two halves around the recursion (frame changes, kicks,
 * drifts) from captured HIP graphs instead of launching them one by one;
 * results are unchanged.  Off by default.  No counterpart in the reference. */
int swg_system_use_graphs(swg_system *sys, int32_t enable);
/* Time the level-0 pl-pl encounter check (encounter_check_all_plpl with the
 * configured ENCOUNTER_CHECK) on the current state: average ms per call over
 * reps calls, the number of pairs found and, for SORTSWEEP, the number of
 * broad-phase candidates (else -1).  SyMBA only. */
int swg_system_time_encounters(swg_system *sys, int32_t reps, double *ms, int64_t *npairs, int64_t *ncand);
/* Leave systems out of the following steps: frozen[s] = 1 keeps system s's
 * state exactly as it is while the others advance; 0 puts it back.  Lets a
 * batch of systems with different step counts run together, each stopping at
 * its own count.  HELIO and SyMBA only.  No counterpart in the reference
 * (one system per run there). */
int swg_system_set_frozen(swg_system *sys, const int32_t *frozen);

/* Energy and angular momentum of every system, as the reference's
 * conservation report computes them: swiftest_util_coord_rh2rb_pl
 * (src/swiftest/swiftest_util.f90:488-518), then
 * swiftest_util_get_energy_and_momentum_system (:1172-1291) with the
 * triangular potential (:1350-1400), rotation and oblateness off.  WHM
 * systems first get vb from vh (swiftest_util_coord_vh2vb_pl, :424-459), as
 * swiftest_io_conservation_report does (src/swiftest/swiftest_io.f90:
 * 135-141); HELIO and SyMBA use the barycentric velocities they hold.
 *   cb_mass, cb_radius  per system (the reference's cb%mass and cb%radius)
 *   pl_mass             per stored body (pl%mass; Gmass = GU * mass)
 *   out                 nsys * SWG_ENERGY_NOUT doubles, fields SWG_E_*
 * The binding energy be is formed only with CHK_CLOSE (prm.lclose).  Sums run
 * in tree order, so the results agree with the reference's serial sums to
 * rounding. */
#define SWG_E_KE_ORBIT 0
#define SWG_E_KE_SPIN 1
#define SWG_E_PE 2
#define SWG_E_BE 3
#define SWG_E_BE_CB 4
#define SWG_E_TE 5
#define SWG_E_L_ORBIT 6  /* 3 components */
#define SWG_E_L_SPIN 9   /* 3 components */
#define SWG_E_L_TOTAL 12 /* 3 components */
#define SWG_E_GMTOT 15
#define SWG_ENERGY_NOUT 16
int swg_system_energy(swg_system *sys, const double *cb_mass, const double *cb_radius, const double *pl_mass,
                      double *out);

/* Discard checks after a step, on the resident state (swiftest_discard_system,
 * src/swiftest/swiftest_discard.f90:13-56): for SyMBA the massive-body checks
 * (symba_discard_pl, src/symba/symba_discard.f90:342-387, without its energy
 * bookkeeping), then the test-particle checks (swiftest_discard_tp,
 * swiftest_discard.f90:80-337).  The thresholds are the reference's
 * parameters; a negative rmin/rmax/rmaxu/qmin turns that check off.  As in
 * the reference, pl%h2b and tp%h2b overwrite the massive bodies', the central
 * body's and the test particles' barycentric velocities when the
 * central-body or barycentric pericentre checks run, and the pericentre
 * checks keep their passage state between calls.  The close-approach check
 * runs with CHK_CLOSE (prm.lclose).
 *   cb_radius  per system (cb%radius, the floor of rmin for test particles)
 * Discarded bodies get their DISCARDED_* status in place (statuses via
 * swg_system_get).  Discarded test particles drop out of the following steps.
 * A system that lost massive bodies must be rebuilt by the host (the
 * reference's rearray): swg_system_get, then swg_system_create without them. */
typedef struct {
    double rmin, rmax, rmaxu;   /* CHK_RMIN, CHK_RMAX, CHK_EJECT */
    double qmin, qmin_alo, qmin_ahi; /* CHK_QMIN, CHK_QMIN_RANGE */
    int32_t qmin_coord;         /* CHK_QMIN_COORD: 0 HELIO, 1 BARY */
    int32_t reserved;
} swg_discard_params;
#define SWG_DISCARDED_RMAX (-1)
#define SWG_DISCARDED_RMIN (-2)
#define SWG_DISCARDED_RMAXU (-3)
#define SWG_DISCARDED_PERI (-4)
#define SWG_DISCARDED_PLR (-5)
int swg_system_discard(swg_system *sys, const swg_discard_params *dp, const double *cb_radius,
                       int64_t *npl_discarded, int64_t *ntp_discarded);

/* ---- collisions resolved by merging ---------------------------------------
 * collision_resolve_plpl (src/collision/collision_resolve.f90:521) with the
 * MERGE collision model: collision_generate_merge
 * (src/collision/collision_generate.f90:158), collision_resolve_mergeaddsub
 * (collision_resolve.f90:326) and swiftest_util_rearray_pl
 * (src/swiftest/swiftest_util.f90:1638), called where the recursion finds the
 * collision (src/symba/symba_step.f90:245).  After swg_system_set_merge a
 * SyMBA step stops at the substep whose collision check flagged a pair,
 * merges each colliding pair into one body (masses and volumes added, the
 * pair's barycentre, its orbital angular momentum put back as a velocity
 * torque), removes the impactors, appends the new bodies, re-sorts by mass,
 * refreshes every rhill, re-runs the encounter check at that level over the
 * full step (keeping the old entries' state where the same pair sits at the
 * same list position and at most as many entries as before), checks the new
 * list for further collisions, and continues the step on the new arrays.
 * Needs SyMBA systems without test particles (one, an ensemble or a slot
 * pool: each system's merges are its own; a shrunk slot keeps its range
 * and pads the rest, and a system loaded into a slot later starts its own
 * ids and masses, Gmass / GU), CHK_CLOSE on and the triangular encounter
 * check; rotation and the ENERGY torque
 * (collision_resolve.f90:624-641) are not modelled, and a body in two
 * colliding pairs of one check (a collision family) returns
 * SWG_ERR_COLLISION at that substep: the system is then left as it stood
 * there, inside the step, and is not to be stepped further.  GU: G in the
 * system's units; pl_mass: masses in input order (NULL: Gmass / GU; NULL
 * for a slot pool).  Body ids are 1..npl in input order; a merged
 * body takes the next unused id.  A step that flags a collision runs twice:
 * as usual, then again from a copy of its starting state with every level
 * on the whole device and a stop at each substep's collision check; other
 * steps run as usual after a device-to-device copy of the bodies' state. */
int swg_system_set_merge(swg_system *sys, double GU, const double *pl_mass);
/* Current body counts (they shrink with merges) and the merges so far. */
int swg_system_counts(swg_system *sys, int32_t *npl, int32_t *ntp, int64_t *nmerge);
/* Where each system's massive bodies start in the concatenated arrays
 * (nsys + 1 entries; they move when merges shrink earlier systems). */
int swg_system_pl_offsets(swg_system *sys, int32_t *pl_off);
/* Ids, masses (0 unless merging is on), Gmass and radii of the massive
 * bodies, in storage order (npl of each; NULL pointers are skipped; a slot
 * pool's padding has id 0 and zero mass). */
int swg_system_get_bodies(swg_system *sys, int32_t *id, double *mass, double *Gmass, double *radius);
/* Autotuner: time the pl-tp and pl-pl acceleration kernels at each
 * workgroup size (64, 128, 256 bodies per tile) on the current state (no state
 * change) and keep the fastest for the following steps.  Reports the choice.
 * With SWG_PAIR_FAST the massive-body tiles are 64 bodies and *pl_block
 * reports the number of wavefronts sharing each tile's rows (2 ... 16).
 * No counterpart in the reference (its loops are OpenMP-scheduled). */
int swg_system_autotune(swg_system *sys, int32_t reps, int32_t *pl_block, int32_t *tp_block);
/* Live timing: while enabled, every launch of the pl-tp and pl-pl
 * acceleration kernels during steps is bracketed by HIP events on the
 * system's stream (enabling resets the record).  swg_system_kernel_times()
 * returns the average launch duration in ms and the number of launches. */
int swg_system_kernel_timing(swg_system *sys, int32_t enable);
int swg_system_kernel_times(swg_system *sys, double *ms_kick_tp, int64_t *n_kick_tp, double *ms_kick_pl,
                            int64_t *n_kick_pl);
int swg_system_destroy(swg_system *sys);

#ifdef __cplusplus
}
#endif
#endif
