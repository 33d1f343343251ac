/*
 * swiftest_oracle.h -- CPU restatement of Swiftest's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (swiftest_amd/, include/)
 * may link, load or call this code.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg use it, and only as the parity oracle / CPU baseline.
 *
 * Every function restates the arithmetic of the Fortran reference
 * (carlislewishard/swiftest v2023.10.2) in plain C, in the reference's
 * operation order; each cites the reference file:line it follows.
 *
 * Floating-point rule (identical in the oracle and in the HIP kernels; this
 * file is compiled with -ffp-contract=off so the compiler adds nothing):
 *   (1) 3-vector squared norms and dot products are a chain of fused
 *       multiply-adds:  d = fma(z1,z2, fma(y1,y2, x1*x2)).
 *   (2) accumulations acc = acc +/- f*x inside pair sums (kick kernels,
 *       SyMBA shell kicks, encounter-pair removal/add-back, WHM ah0) and the
 *       velocity kick / linear drift updates (v += a*dt, r += p*dt) are one fma.
 *   (3) every other operation is rounded separately, as written.
 * The reference leaves contraction to its compiler (GCC -O3 -march=native
 * contracts these same a*b+c shapes); the rule above fixes one choice.
 * sqrt and division are IEEE (correctly rounded).  r2**(-1.5) in the SyMBA
 * shell kick is evaluated correctly rounded by a double-double sequence
 * (swo_pow_m1p5), identical on both sides; glibc's pow agrees with it to
 * within 1 ulp (checked in tests).
 *
 * Arrays follow the reference layout: vectors are double[n][3] (the Fortran
 * (3,n) column-major layout), masks are int32 (nonzero = .true.), pair index
 * lists are 1-based int32.
 */
#ifndef SWIFTEST_ORACLE_H
#define SWIFTEST_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes (src/globals/globals_module.f90:86-103) ---- */
#define SWO_ACTIVE 0
#define SWO_INACTIVE 1
#define SWO_DISCARDED_DRIFTERR (-7)
#define SWO_COLLIDED (-14)

/* ---- leaf kernels ---- */
void swo_kick_getacch_int_all_tri(int npl, int nplm, const double *r, const double *Gmass,
                                  const double *radius, double *acc);
void swo_kick_getacch_int_all_flat(int npl, int64_t nplpl, const int32_t *k_plpl, const double *r,
                                   const double *Gmass, const double *radius, double *acc);
void swo_kick_getacch_int_all_tp(int ntp, int npl, const double *rtp, const double *rpl,
                                 const double *GMpl, const int32_t *lmask, double *acc);

void swo_drift_one(double mu, double *rx, double *ry, double *rz, double *vx, double *vy, double *vz,
                   double dt, int32_t *iflag);
void swo_drift_all(const double *mu, double *x, double *v, int n, double dt, const int32_t *lmask,
                   int32_t *iflag, int lgr, double inv_c2);
void swo_drift_kepu_stumpff(double x, double *c0, double *c1, double *c2, double *c3);

void swo_encounter_check_one(double xr, double yr, double zr, double vxr, double vyr, double vzr,
                             double renc, double dt, int32_t *lencounter, int32_t *lvdotr);
/* Triangular pair lists.  Outputs are malloc'd (free with swo_free); indices 1-based. */
/* encounter_check_all_plplm, TRIANGULAR (encounter_check.f90:42-109): the
 * pairs of the first nplm (fully interacting) bodies among themselves and
 * with the other npl - nplm, listed index1 ascending, then index2 ascending. */
int64_t swo_encounter_check_all_plplm(int npl, int nplm, const double *r, const double *v, const double *renc,
                                      double dt, int32_t **index1, int32_t **index2, int32_t **lvdotr);
int64_t swo_encounter_check_all_plpl(int npl, const double *r, const double *v, const double *renc,
                                     double dt, int32_t **index1, int32_t **index2, int32_t **lvdotr);
int64_t swo_encounter_check_all_pltp(int npl, int ntp, const double *rpl, const double *vpl,
                                     const double *rtp, const double *vtp, const double *renc,
                                     double dt, int32_t **index1, int32_t **index2, int32_t **lvdotr);
void swo_free(void *p);

void swo_collision_check_one(double xr, double yr, double zr, double vxr, double vyr, double vzr,
                             double Gmtot, double rlim, double dt, int32_t lvdotr,
                             int32_t *lcollision, int32_t *lclosest);
void swo_orbel_xv2aeq(double mu, double rx, double ry, double rz, double vx, double vy, double vz,
                      double *a, double *e, double *q);
double swo_orbel_xv2a(double mu, double rx, double ry, double rz, double vx, double vy, double vz);
double swo_pow_m1p5(double x);
double swo_powi(double x, int n);

/* ---- whole-system integrators (WHM / HELIO / SyMBA) ---- */
#define SWO_WHM 0
#define SWO_HELIO 1
#define SWO_SYMBA 2

typedef struct swo_sys swo_sys;

typedef struct swo_params {
    double dt;
    int lclose;          /* CHK_CLOSE (radius overlap test in pl-pl gravity; collision checks) */
    int short_circuit;   /* 1: `a .or. b` on encounter checks skips b when a is true (gfortran -O) */
    int istep_out;       /* snapshot cadence; 0 = never */
    int rhill_present;   /* RHILL_PRESENT: keep input rhill, never refresh */
    int nthreads;        /* OpenMP threads for the pair loops (1 = serial) */
    double gmtiny;       /* GMTINY (SyMBA): bodies with Gmass < gmtiny interact only with the others */
} swo_params;

typedef struct swo_stats {
    int64_t nplpl_enc;     /* level-0 pl-pl encounters in the last step */
    int64_t npltp_enc;     /* level-0 pl-tp encounters in the last step */
    int32_t max_depth;     /* deepest recursion level entered in the last step */
    int64_t nsubsteps;     /* recursion substeps in the last step */
    int64_t ncollision;    /* collisions detected (cumulative) */
    int64_t ndrifterr;     /* drift failures (cumulative) */
} swo_stats;

/* Create a system from initial conditions as read from the reference's
 * input files (heliocentric rh, vh; Gmass; radius; optional rhill), then run
 * the reference's initialize sequence for the chosen integrator. */
swo_sys *swo_sys_new(int integrator, int npl, int ntp, double GMcb, double Rcb,
                     const double *pl_rh, const double *pl_vh, const double *Gmass,
                     const double *radius, const double *rhill_in,
                     const double *tp_rh, const double *tp_vh, const swo_params *prm);
void swo_sys_free(swo_sys *s);
/* Run nsteps steps of the driver loop (step + output cadence).  Returns 0 or
 * a negative code: -1 collision detected (merge not modelled), -2 fatal
 * drift error (WHM pl), -3 recursion depth guard. */
int swo_sys_run(swo_sys *s, int64_t nsteps);
/* One integrator step only (no snapshot). */
int swo_sys_step(swo_sys *s);
void swo_sys_get(const swo_sys *s, double *pl_rh, double *pl_vh, double *pl_vb, double *tp_rh,
                 double *tp_vh, double *tp_vb, int32_t *pl_status, int32_t *tp_status);
void swo_sys_get_misc(const swo_sys *s, double *rhill, double *cb_vb, double *ptbeg, double *ptend,
                      int32_t *lfirst);
void swo_sys_get_order(const swo_sys *s, int32_t *pl_id);
void swo_sys_stats(const swo_sys *s, swo_stats *st);
/* recursion levels of the last step: levelg (current) and levelm (deepest
 * reached) of every massive body and test particle, in storage order */
void swo_sys_get_levels(const swo_sys *s, int32_t *levelg, int32_t *levelm, int32_t *tlevelg, int32_t *tlevelm);
/* the shell factor of symba_kick_list_plpl / _pltp for shell radius^2 ri and
 * separation^2 r2; returns 0 when the pair is dropped (r2 < ri*RSHELL^2) */
int swo_shell_fac(double ri, double r2, double *fac);
int swo_sys_npl(const swo_sys *s);
int swo_sys_ntp(const swo_sys *s);
/* Encounter lists produced by the last step's level-0 checks (1-based). */
int64_t swo_sys_list(const swo_sys *s, int which, int32_t *index1, int32_t *index2, int32_t *lvdotr);
/* Status of each pair of those lists (0 ACTIVE, -14 COLLIDED), list order. */
int64_t swo_sys_list_status(const swo_sys *s, int which, int32_t *status);
/* What a detected collision does.  0 (default): the step stops at the
 * collision check that found it and swo_sys_run returns -1 at once.  1: the
 * colliding pairs are flagged COLLIDED, which takes them out of the list
 * kicks and list checks for the rest of the step (symba_kick.f90:169/281,
 * symba_encounter_check.f90:112/187), the step is finished, and -1 is
 * returned at its end: the state a caller gets when the collision is not
 * resolved within the step. */
void swo_sys_set_collision_mode(swo_sys *s, int finish_step);
int swo_sys_set_merge(swo_sys *s, double GU, const double *pl_mass);
void swo_sys_get_bodies(const swo_sys *s, int32_t *id, double *mass, double *Gmass, double *radius);
int64_t swo_sys_nmerge(const swo_sys *s);
int swo_merge_bodies(const double *m, const double *Gm, const double *R, const double *rh, const double *vb,
                     const double *cb_rb, const double *cb_vb, double GU, double *out);

/* ---- energy and angular momentum (output cadence) ---- */
/* Fields of the energy record, as swiftest_nbody_system stores them. */
#define SWO_E_KE_ORBIT 0
#define SWO_E_KE_SPIN 1
#define SWO_E_PE 2
#define SWO_E_BE 3
#define SWO_E_BE_CB 4
#define SWO_E_TE 5
#define SWO_E_L_ORBIT 6  /* 3 components */
#define SWO_E_L_SPIN 9   /* 3 components */
#define SWO_E_L_TOTAL 12 /* 3 components */
#define SWO_E_GMTOT 15
#define SWO_ENERGY_NOUT 16
/* swiftest_util_coord_vh2vb_pl (swiftest_util.f90:424-459): barycentric
 * velocities of the massive bodies and of the central body from vh. */
void swo_coord_vh2vb(int npl, double GMcb, const double *Gmass, const double *vh, double *vb, double *cb_vb);
/* ---- discard checks (after each step) ---- */
#define SWO_DISCARDED_RMAX (-1)
#define SWO_DISCARDED_RMIN (-2)
#define SWO_DISCARDED_RMAXU (-3)
#define SWO_DISCARDED_PERI (-4)
#define SWO_DISCARDED_PLR (-5)
/* The discard checks of one system after a step, in the reference's order:
 * for SyMBA the massive-body checks (symba_discard_pl, symba_discard.f90:
 * 342-387: rh2rb, then symba_discard_cb_pl :14-104 and symba_discard_peri_pl
 * :292-339), then the test-particle checks (swiftest_discard_tp,
 * swiftest_discard.f90:80-119: pl%h2b and tp%h2b, which overwrite pl_vb,
 * cb_vb and tp_vb, then swiftest_discard_cb_tp :122-191,
 * swiftest_discard_peri_tp :194-243 and swiftest_discard_pl_tp :246-288).
 *   prm = {rmin, rmax, rmaxu, qmin, qmin_alo, qmin_ahi, dt}; a negative
 *   rmin/rmax/rmaxu/qmin switches that check off.  qmin_coord: 0 HELIO, 1 BARY.
 *   isperi/atp/peri and *lfirst_peri are the pericentre-passage state the
 *   reference keeps between calls (isperi starts at 1, lfirst_peri at 1).
 *   nplenc: pl-pl encounters of each body in the step's level-0 list.
 * The total mass Gmtot is set_msys's (swiftest_util.f90:2107-2118), formed
 * on every call.  Statuses change in place; counts[0], counts[1] receive the
 * numbers of massive bodies and test particles newly discarded. */
void swo_discard(int integrator, int npl, int ntp, double GMcb, double Rcb, double *cb_vb, const double *pl_rh,
                 const double *pl_vh, double *pl_vb, const double *Gmass, const double *radius, const double *rhill,
                 int32_t *pl_status, const int32_t *nplenc, int32_t *pl_isperi, double *pl_atp, double *pl_peri,
                 int32_t *lfirst_peri, const double *tp_rh, const double *tp_vh, double *tp_vb, int32_t *tp_status,
                 int32_t *tp_isperi, double *tp_atp, double *tp_peri, const double *prm, int qmin_coord, int lclose,
                 int64_t *counts);
/* swo_discard on a whole-system state after its last step; prm6 = {rmin,
 * rmax, rmaxu, qmin, qmin_alo, qmin_ahi} (dt and lclose are the system's). */
int swo_sys_discard(swo_sys *s, const double *prm6, int qmin_coord, int64_t *counts);
/* Fortran norm2 of a 3-vector as gfortran evaluates it (scaled sum of squares). */
double swo_norm2(double x, double y, double z);
/* swiftest_util_coord_rh2rb_pl (swiftest_util.f90:488-518) followed by
 * swiftest_util_get_energy_and_momentum_system (:1172-1291) with the
 * triangular potential (:1350-1400); rotation and oblateness off.  The
 * velocities are the barycentric ones the integrator holds (vb, cb_vb).
 * rb (npl x 3, optional) receives the barycentric positions. */
void swo_energy(int npl, double GMcb, double mcb, double Rcb, const double *cb_vb, const double *rh,
                const double *vb, const double *Gmass, const double *mass, const double *radius,
                const int32_t *status, int lclose, double *out, double *rb);

#ifdef __cplusplus
}
#endif
#endif
