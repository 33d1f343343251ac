// merge.hip -- SyMBA collisions resolved by merging inside the step.
//
// The reference resolves a pl-pl collision where the recursion finds it:
// right after the collision check of a substep (symba_step.f90:241-247),
// collision_resolve_plpl (collision_resolve.f90:521-662) merges each colliding
// pair into one body (collision_generate_merge, collision_generate.f90:
// 158-259; collision_resolve_mergeaddsub, collision_resolve.f90:326-520) and
// rebuilds the massive-body arrays and the encounter list
// (swiftest_util_rearray_pl, swiftest_util.f90:1638-1811); the remaining
// substeps of every level then run on the new arrays.
//
// Here the device stops at the substep whose check flagged a collision; this
// host code does the bookkeeping of the reference on a host copy of the
// system, uploads the new arrays and list, and rebuilds the working sets of
// the levels the recursion is in, so that the step continues where it
// stopped.  Scope: the MERGE collision model without rotation, systems
// (one, or an ensemble) without test particles, and no ENERGY torque
// (collision_resolve.f90:624-641).  A body in two colliding pairs of one check (a collision family,
// collision_resolve_make_impactors_pl) is refused with SWG_ERR_COLLISION.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "swg_engine.h"
#include "swg_kernels.h"

namespace swg {

int fail(int code, const std::string &msg);
int hipfail(hipError_t e, const char *what);
void host_set_rhill(double GMcb, int n, const double *rh, const double *vh, const double *Gm, double *rhill);

#define MCHK(x)                                             \
    do {                                                    \
        hipError_t e_ = (x);                                \
        if (e_ != hipSuccess) return hipfail(e_, #x);       \
    } while (0)

namespace {

constexpr double PI_ = 3.141592653589793238462643383279502884197;
constexpr double THIRD_ = 0.333333333333333333333333333333333333333;
constexpr double EPS_ = 2.220446049250313e-16;  // epsilon(1.0_DP)

// Fortran norm2 of a 3-vector as gfortran evaluates it: a running scale
// (the largest |component| so far) and a sum of squares relative to it.
double norm2_3(const double *a)
{
    double scale = 1.0, ssq = 0.0;
    for (int k = 0; k < 3; k++) {
        const double x = std::fabs(a[k]);
        if (a[k] == 0.0) continue;
        if (x > scale) {
            const double q = scale / x;
            ssq = 1.0 + ssq * q * q;
            scale = x;
        } else {
            const double q = x / scale;
            ssq += q * q;
        }
    }
    return scale * std::sqrt(ssq);
}

// operator .unit. (operator_unit.f90): zero for a vector of norm <= tiny
void unit3(const double *a, double *u)
{
    const double m = norm2_3(a);
    for (int d = 0; d < 3; d++) u[d] = m > DBL_TINY ? a[d] / m : 0.0;
}

// operator .cross. (operator_cross.f90)
void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

double cube(double r) { return (r * r) * r; }  // r**3 as gfortran expands it

// collision_util_velocity_torque (collision_util.f90:1001-1029)
void velocity_torque(const double *dL, double mass, const double *r, double *v)
{
    double dLu[3], ru[3], w[3], c[3];
    unit3(dL, dLu);
    unit3(r, ru);
    const double rmag = norm2_3(r), vmag = norm2_3(v);
    for (int d = 0; d < 3; d++) w[d] = dL[d] / (rmag * mass);
    const double vapply = norm2_3(w);
    if (vapply / vmag > EPS_) {
        cross3(dLu, ru, c);
        for (int d = 0; d < 3; d++) v[d] = v[d] + vapply * c[d];
    }
}

struct NewBody {
    double rh[3], vh[3], vb[3], mass, Gm, R;
    int32_t id, levelg, levelm;
};

// The merged body of two impactors a and b (consolidate_impactors without
// children, collision_resolve.f90:15-177; set_coordinate_impactors,
// collision_util.f90:609-670; generate_merge; mergeaddsub's coordinates,
// collision_resolve.f90:356-380).  Returns 0 when a is the larger impactor
// (maxloc: the first of equal masses), 1 when b is.
int merge_two(const double m[2], const double Gm[2], const double Rin[2], const double rh[2][3],
              const double vbin[2][3], const double *cb_rb, const double *cb_vb, double GU, NewBody &nb)
{
    double R[2], rb[2][3], vb[2][3];
    for (int k = 0; k < 2; k++) {
        const double vol = ((4.0 / 3.0) * PI_) * cube(Rin[k]);
        for (int d = 0; d < 3; d++) {
            rb[k][d] = rh[k][d] + cb_rb[d];
            vb[k][d] = vbin[k][d];
        }
        R[k] = std::pow((3.0 * vol) / (4.0 * PI_), 1.0 / 3.0);
    }
    const double mtot = m[0] + m[1];
    // overlapping impactors are moved apart along their relative velocity
    const double rlim = R[0] + R[1];
    double rrel[3];
    for (int d = 0; d < 3; d++) rrel[d] = rb[1][d] - rb[0][d];
    const double rrel_mag = norm2_3(rrel);
    if (rrel_mag < rlim) {
        double vrel[3], vu[3];
        for (int d = 0; d < 3; d++) vrel[d] = vb[1][d] - vb[0][d];
        unit3(vrel, vu);
        const double vdotr = (vu[0] * rrel[0] + vu[1] * rrel[1]) + vu[2] * rrel[2];
        const double root = std::sqrt(((rlim * rlim) - (rrel_mag * rrel_mag)) + (vdotr * vdotr));
        const double t = vdotr - std::copysign(1.0, vdotr) * root;
        for (int d = 0; d < 3; d++) {
            double dr = (-t) * vu[d];
            dr = (1.0 + 2.0 * EPS_) * dr;
            rb[0][d] = rb[0][d] - (dr * m[1]) / mtot;
            rb[1][d] = rb[1][d] + (dr * m[0]) / mtot;
        }
    }
    double rbcom[3], vbcom[3];
    for (int d = 0; d < 3; d++) {
        rbcom[d] = (m[0] * rb[0][d] + m[1] * rb[1][d]) / mtot;
        vbcom[d] = (m[0] * vb[0][d] + m[1] * vb[1][d]) / mtot;
    }
    // orbital angular momentum of the pair about its barycentre: lost in the
    // merge and put back as a velocity torque on the merged body
    double Lk[2][3];
    for (int k = 0; k < 2; k++) {
        double rc[3], vc[3];
        for (int d = 0; d < 3; d++) {
            rc[d] = rb[k][d] - rbcom[d];
            vc[d] = vb[k][d] - vbcom[d];
        }
        Lk[k][0] = m[k] * (rc[1] * vc[2] - rc[2] * vc[1]);
        Lk[k][1] = m[k] * (rc[2] * vc[0] - rc[0] * vc[2]);
        Lk[k][2] = m[k] * (rc[0] * vc[1] - rc[1] * vc[0]);
    }
    double dL[3], v[3];
    for (int d = 0; d < 3; d++) {
        dL[d] = -(0.0 - (Lk[0][d] + Lk[1][d]));  // -(L_after - L_before), L_after = 0
        v[d] = vbcom[d];
    }
    nb.mass = m[0] + m[1];
    const double volume = ((4.0 / 3.0) * PI_) * (cube(R[0]) + cube(R[1]));
    velocity_torque(dL, nb.mass, rbcom, v);
    for (int d = 0; d < 3; d++) {
        nb.rh[d] = rbcom[d] - cb_rb[d];
        nb.vh[d] = v[d] - cb_vb[d];
        nb.vb[d] = v[d];
    }
    nb.Gm = GU * nb.mass;
    nb.R = std::pow((3.0 * volume) / (4.0 * PI_), THIRD_);
    return Gm[1] > Gm[0] ? 1 : 0;
}

// The host copy of one system that the merges work on.
struct HostSys {
    int n = 0;
    std::vector<double> rh, vh, vb, Gm, mass, R, rhill;
    std::vector<int32_t> id, levelg, levelm, isperi;
    std::vector<double> atp, peri;
    // pl-pl list: 0-based indices and ids
    std::vector<int32_t> i1, i2, id1, id2, lvdotr, status, level;
};

}  // namespace

int Ensemble::set_merge(double GU_, const double *pl_mass)
{
    if (prm.integrator != SWG_SYMBA) return fail(SWG_ERR_ARG, "merging collisions needs SyMBA");
    if (slot_pool && pl_mass) return fail(SWG_ERR_ARG, "slot pools take the masses as Gmass / GU");
    if (ntp != 0) return fail(SWG_ERR_ARG, "merging collisions needs a system without test particles");
    if (!prm.lclose) return fail(SWG_ERR_ARG, "merging collisions needs the collision checks (lclose)");
    if (prm.encounter_check != 0) return fail(SWG_ERR_ARG, "merging collisions needs the triangular encounter check");
    if (!(GU_ > 0.0)) return fail(SWG_ERR_ARG, "GU must be positive");
    GU = GU_;
    // a slot's padding (after its real bodies) has mass 0
    h_mass.assign(npl, 0.0);
    for (int s = 0; s < nsys; s++)
        for (int i = h_pl_off[s]; i < h_pl_off[s] + h_nreal_pl[s]; i++)
            h_mass[i] = pl_mass ? pl_mass[h_pl_off[s] + h_pl_id[i] - 1] : h_Gm[i] / GU;
    if (!mb_cb) {
        if (alloc_vec(mb_rh, npl) || alloc_vec(mb_vh, npl) || alloc_vec(mb_vb, npl) || own(&mb_status, npl) ||
            own(&mb_sys, 2 * (size_t)nsys) || own(&mb_cb, 9 * (size_t)nsys))
            return SWG_ERR_HIP;
    }
    h_maxid.assign(nsys, 0);
    h_nplm.assign(nsys, 0);
    for (int s = 0; s < nsys; s++) merge_reset_system(s);
    merge_on = true;
    // the replay of a step that merges changes the body count mid-step: no
    // captured launch sequences
    use_graphs = false;
    drop_graphs();
    return 0;
}

// Id and fully-interacting count bookkeeping of system s as it stands: ids
// 1..n in input order, the next merged body n + 1 (swg_system_load calls this
// for a slot's new system).
void Ensemble::merge_reset_system(int s)
{
    const int a = h_pl_off[s], n = h_nreal_pl[s];
    h_maxid[s] = n + h_nreal_tp[s];
    h_nplm[s] = n;
    if (prm.gmtiny > 0.0) {
        h_nplm[s] = 0;
        for (int i = a; i < a + n; i++) h_nplm[s] += !(h_Gm[i] < prm.gmtiny);
    }
}

// The state a step changes (bodies' positions, velocities and status, the
// central body's sums, lfirst, the error bits), copied aside at the start of
// the step (restore = false) or put back (restore = true).  Everything else a
// step uses is rebuilt at its start (levels, renc, lists, flags).
int Ensemble::merge_save(bool restore)
{
    hipStream_t st = stream;
    auto cp = [&](void *a, void *b, size_t bytes) -> int {
        if (bytes) MCHK(hipMemcpyAsync(restore ? a : b, restore ? b : a, bytes, hipMemcpyDeviceToDevice, st));
        return 0;
    };
    const size_t n = (size_t)npl * sizeof(double);
    for (auto pr : {std::make_pair(pl.rh, mb_rh), std::make_pair(pl.vh, mb_vh), std::make_pair(pl.vb, mb_vb)}) {
        if (int rc = cp(pr.first.x, pr.second.x, n)) return rc;
        if (int rc = cp(pr.first.y, pr.second.y, n)) return rc;
        if (int rc = cp(pr.first.z, pr.second.z, n)) return rc;
    }
    if (int rc = cp(pl.status, mb_status, (size_t)npl * sizeof(int32_t))) return rc;
    if (int rc = cp(sy.lfirst, mb_sys, nsys * sizeof(int32_t))) return rc;
    if (int rc = cp(sy.err, mb_sys + nsys, nsys * sizeof(int32_t))) return rc;
    if (int rc = cp(sy.cbvb, mb_cb, 3 * nsys * sizeof(double))) return rc;
    if (int rc = cp(sy.ptbeg, mb_cb + 3 * nsys, 3 * nsys * sizeof(double))) return rc;
    if (int rc = cp(sy.ptend, mb_cb + 6 * nsys, 3 * nsys * sizeof(double))) return rc;
    return 0;
}

// One step with merges: the step as usual (the persistent recursion kernels
// included); when it flags a collision, the step again from its start with
// every level on the whole device, resolving each collision where the
// recursion finds it.  The recursion is deterministic and gives the same
// bits in every scheduling mode, so the replay retraces the first pass up to
// the collision.
int Ensemble::step_merge()
{
    if (int rc = merge_save(false)) return rc;
    merge_replay = false;
    const int64_t ncol = st_.ncollision;
    int rc = step_helio_symba();
    if (rc != SWG_ERR_COLLISION) return rc;
    if (int r2 = merge_save(true)) return r2;
    st_.ncollision = ncol;
    merge_replay = true;
    rc = step_helio_symba();
    merge_replay = false;
    return rc;
}

// After the collision launch of a substep at level irec: resolve what it
// flagged.  sy.err bit 1 is the device's collision flag.
int Ensemble::merge_point(int irec, double dt)
{
    std::vector<int32_t> err(nsys);
    MCHK(hipMemcpyAsync(err.data(), sy.err, nsys * sizeof(int32_t), hipMemcpyDeviceToHost, stream));
    MCHK(hipStreamSynchronize(stream));
    bool any = false;
    for (int s = 0; s < nsys; s++)
        if (err[s] & 1) {
            any = true;
            st_.ncollision++;  // one count per system whose check found a collision
            err[s] &= ~1;
        }
    if (!any) return 0;
    if (int rc = resolve_merges(irec, dt)) return rc;
    MCHK(hipMemcpyAsync(sy.err, err.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice, stream));
    MCHK(hipStreamSynchronize(stream));
    return 0;
}

// The merges of one system (local indices) on its host copy S: the
// central body's barycentric state, then rounds of extract, merge, rearray
// and collision check until no pair collides.  status: the bodies' status
// before the merges; cbvb receives the central body's barycentric velocity.
int Ensemble::resolve_system(void *hs, const std::vector<int32_t> &status, double GMcb, int irec, double dt,
                             int32_t &maxid_s, int32_t &nplm_s, int64_t &ncol, double *cbvb)
{
    HostSys &S = *(HostSys *)hs;
    const int n0 = S.n;
    // pl%vb2vh and pl%rh2rb (collision_resolve.f90:557-558): the central
    // body's barycentric velocity (bodies summed last to first) and position
    double cbrb[3];
    cbvb[0] = cbvb[1] = cbvb[2] = 0.0;
    for (int i = n0 - 1; i >= 0; i--) {
        if (status[i] == INACTIVE) continue;
        for (int d = 0; d < 3; d++) cbvb[d] = cbvb[d] - (S.Gm[i] * S.vb[3 * (size_t)i + d]) / GMcb;
    }
    for (int i = 0; i < n0; i++)
        for (int d = 0; d < 3; d++) S.vh[3 * (size_t)i + d] = S.vb[3 * (size_t)i + d] - cbvb[d];
    {
        double Gmtot = GMcb, xs[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < n0; i++) {
            if (status[i] == INACTIVE) continue;
            Gmtot = Gmtot + S.Gm[i];
            for (int d = 0; d < 3; d++) xs[d] = xs[d] + S.Gm[i] * S.rh[3 * (size_t)i + d];
        }
        for (int d = 0; d < 3; d++) cbrb[d] = -xs[d] / Gmtot;
    }

    for (int loop = 1;; loop++) {
        if (loop > 1000) return fail(SWG_ERR_COLLISION, "a runaway collisional cascade (more than 1000 rounds of merges)");
        // the colliding entries leave the list, in list order
        // (collision_resolve_extract_plpl, collision_resolve.f90:180-250)
        std::vector<std::pair<int, int>> coll;
        std::vector<char> seen(S.n, 0);
        bool family = false;
        {
            size_t o = 0;
            for (size_t k = 0; k < S.i1.size(); k++) {
                if (S.status[k] == COLLIDED) {
                    const int a = S.i1[k], b = S.i2[k];
                    family |= seen[a] || seen[b];
                    seen[a] = seen[b] = 1;
                    coll.emplace_back(a, b);
                    continue;
                }
                S.i1[o] = S.i1[k];
                S.i2[o] = S.i2[k];
                S.lvdotr[o] = S.lvdotr[k];
                S.status[o] = S.status[k];
                S.level[o] = S.level[k];
                o++;
            }
            S.i1.resize(o);
            S.i2.resize(o);
            S.lvdotr.resize(o);
            S.status.resize(o);
            S.level.resize(o);
        }
        if (coll.empty()) break;
        if (family)
            return fail(SWG_ERR_COLLISION,
                        "a body in two colliding pairs of one check (a collision family) is not merged here");
        const size_t ne = S.i1.size();
        S.id1.resize(ne);
        S.id2.resize(ne);
        for (size_t k = 0; k < ne; k++) {
            S.id1[k] = S.id[S.i1[k]];
            S.id2[k] = S.id[S.i2[k]];
        }
        // each pair in turn: the merged body, and the smaller impactor's
        // entries renamed to the larger (collision_generate.f90:232-246)
        std::vector<NewBody> add;
        std::vector<char> keep(S.n, 1);
        for (const auto &pr : coll) {
            const int idx[2] = {pr.first, pr.second};
            double m[2], Gm[2], R[2], rh[2][3], vb[2][3];
            for (int k = 0; k < 2; k++) {
                m[k] = S.mass[idx[k]];
                Gm[k] = S.Gm[idx[k]];
                R[k] = S.R[idx[k]];
                for (int d = 0; d < 3; d++) {
                    rh[k][d] = S.rh[3 * (size_t)idx[k] + d];
                    vb[k][d] = S.vb[3 * (size_t)idx[k] + d];
                }
            }
            NewBody nb;
            const int ib = idx[merge_two(m, Gm, R, rh, vb, cbrb, cbvb, GU, nb)];
            for (size_t k = 0; k < ne; k++)
                for (int q = 0; q < 2; q++) {
                    const int b = idx[q];
                    if (b == ib) continue;
                    if (S.id1[k] == S.id[b]) {
                        S.id1[k] = S.id[ib];
                        S.i1[k] = b;
                    }
                    if (S.id2[k] == S.id[b]) {
                        S.id2[k] = S.id[ib];
                        S.i2[k] = b;
                    }
                    if (S.id1[k] == S.id2[k]) S.status[k] = INACTIVE;
                }
            maxid_s = maxid_s + 1;
            nb.id = maxid_s;
            nb.levelg = S.levelg[ib];
            nb.levelm = S.levelm[ib];
            add.push_back(nb);
            keep[idx[0]] = keep[idx[1]] = 0;
            nmerge++;
        }

        // ---- swiftest_util_rearray_pl ----
        // discards out (order kept), the new bodies appended, then the mass
        // sort, descending (ties keep their order)
        std::vector<int> src;
        for (int i = 0; i < S.n; i++)
            if (keep[i]) src.push_back(i);
        for (size_t q = 0; q < add.size(); q++) src.push_back(-1 - (int)q);
        const int n = (int)src.size();
        auto key = [&](int s) { return s >= 0 ? S.mass[s] : add[-1 - s].mass; };
        std::vector<int> ord(n);
        std::iota(ord.begin(), ord.end(), 0);
        std::stable_sort(ord.begin(), ord.end(), [&](int u, int v) { return key(src[u]) > key(src[v]); });
        HostSys T;
        T.n = n;
        T.rh.resize(3 * (size_t)n);
        T.vh.resize(3 * (size_t)n);
        T.vb.resize(3 * (size_t)n);
        T.Gm.resize(n);
        T.mass.resize(n);
        T.R.resize(n);
        T.id.resize(n);
        T.levelg.resize(n);
        T.levelm.resize(n);
        const bool dis = !S.isperi.empty();
        if (dis) {
            T.isperi.resize(n);
            T.atp.resize(n);
            T.peri.resize(n);
        }
        for (int q = 0; q < n; q++) {
            const int a = src[ord[q]];
            if (a >= 0) {
                for (int d = 0; d < 3; d++) {
                    T.rh[3 * (size_t)q + d] = S.rh[3 * (size_t)a + d];
                    T.vh[3 * (size_t)q + d] = S.vh[3 * (size_t)a + d];
                    T.vb[3 * (size_t)q + d] = S.vb[3 * (size_t)a + d];
                }
                T.Gm[q] = S.Gm[a];
                T.mass[q] = S.mass[a];
                T.R[q] = S.R[a];
                T.id[q] = S.id[a];
                T.levelg[q] = S.levelg[a];
                T.levelm[q] = S.levelm[a];
                if (dis) {
                    T.isperi[q] = S.isperi[a];
                    T.atp[q] = S.atp[a];
                    T.peri[q] = S.peri[a];
                }
            } else {
                const NewBody &b = add[-1 - a];
                for (int d = 0; d < 3; d++) {
                    T.rh[3 * (size_t)q + d] = b.rh[d];
                    T.vh[3 * (size_t)q + d] = b.vh[d];
                    T.vb[3 * (size_t)q + d] = b.vb[d];
                }
                T.Gm[q] = b.Gm;
                T.mass[q] = b.mass;
                T.R[q] = b.R;
                T.id[q] = b.id;
                T.levelg[q] = b.levelg;
                T.levelm[q] = b.levelm;
                if (dis) {
                    T.isperi[q] = 1;  // a new body (swiftest_util_setup_body)
                    T.atp[q] = 0.0;
                    T.peri[q] = 0.0;
                }
            }
        }
        T.rhill.resize(n);
        host_set_rhill(GMcb, n, T.rh.data(), T.vh.data(), T.Gm.data(), T.rhill.data());
        int nplm_new = n;
        if (prm.gmtiny > 0.0) {
            nplm_new = 0;
            for (int i = 0; i < n; i++) nplm_new += !(T.Gm[i] < prm.gmtiny);
        }
        // the encounter check at level irec over the full step
        // (symba_encounter_check_pl): the list is replaced only when it
        // finds an encounter; every body it lists moves to level irec
        std::vector<double> renc(n);
        const double rs = rshell_pow(irec);
        for (int i = 0; i < n; i++) renc[i] = (T.rhill[i] * RHSCALE) * rs;
        std::vector<int32_t> ni1, ni2, nlv;
        for (int i = 0; i < nplm_new; i++)
            for (int j = i + 1; j < n; j++) {
                const double *ri = &T.rh[3 * (size_t)i], *rj = &T.rh[3 * (size_t)j];
                const double *vi = &T.vb[3 * (size_t)i], *vj = &T.vb[3 * (size_t)j];
                bool lv;
                if (encounter_one(rj[0] - ri[0], rj[1] - ri[1], rj[2] - ri[2], vj[0] - vi[0], vj[1] - vi[1],
                                  vj[2] - vi[2], renc[i] + renc[j], prm.dt, lv)) {
                    ni1.push_back(i);
                    ni2.push_back(j);
                    nlv.push_back(lv);
                }
            }
        const size_t nold = ne;
        if (!ni1.empty()) {
            const size_t m = ni1.size();
            T.i1 = ni1;
            T.i2 = ni2;
            T.lvdotr = nlv;
            T.status.assign(m, ACTIVE);
            T.level.assign(m, irec);
            T.id1.resize(m);
            T.id2.resize(m);
            for (size_t k = 0; k < m; k++) {
                T.id1[k] = T.id[ni1[k]];
                T.id2[k] = T.id[ni2[k]];
                T.levelg[ni1[k]] = T.levelm[ni1[k]] = irec;
                T.levelg[ni2[k]] = T.levelm[ni2[k]] = irec;
            }
        } else {
            T.i1 = S.i1;
            T.i2 = S.i2;
            T.lvdotr = S.lvdotr;
            T.status = S.status;
            T.level = S.level;
            T.id1 = S.id1;
            T.id2 = S.id2;
        }
        // the old entries' state where the same pair sits at the same list
        // position; the list keeps at most as many entries as it had
        if (nold > 0) {
            const size_t nmin = std::min(T.i1.size(), nold);
            T.i1.resize(nmin);
            T.i2.resize(nmin);
            T.lvdotr.resize(nmin);
            T.status.resize(nmin);
            T.level.resize(nmin);
            bool gone = false;
            for (size_t k = 0; k < nmin; k++) {
                if ((T.id1[k] == S.id1[k] && T.id2[k] == S.id2[k]) || (T.id1[k] == S.id2[k] && T.id2[k] == S.id1[k])) {
                    T.lvdotr[k] = S.lvdotr[k];
                    T.status[k] = S.status[k];
                    T.level[k] = S.level[k];
                }
                int a = -1, b = -1;
                for (int q = 0; q < n && (a < 0 || b < 0); q++) {
                    if (a < 0 && T.id[q] == T.id1[k]) a = q;
                    if (b < 0 && T.id[q] == T.id2[k]) b = q;
                }
                T.i1[k] = a;
                T.i2[k] = b;
                gone |= a < 0 || b < 0;
            }
            if (gone) {
                size_t o = 0;
                for (size_t k = 0; k < nmin; k++) {
                    if (T.i1[k] < 0 || T.i2[k] < 0) continue;
                    T.i1[o] = T.i1[k];
                    T.i2[o] = T.i2[k];
                    T.lvdotr[o] = T.lvdotr[k];
                    T.status[o] = T.status[k];
                    T.level[o] = T.level[k];
                    o++;
                }
                T.i1.resize(o);
                T.i2.resize(o);
                T.lvdotr.resize(o);
                T.status.resize(o);
                T.level.resize(o);
            }
        }
        T.id1.clear();
        T.id2.clear();
        S = std::move(T);
        nplm_s = nplm_new;
        // a new body may itself be colliding (collision_resolve.f90:646-648)
        bool any = false;
        for (size_t k = 0; k < S.i1.size(); k++) {
            const int i = S.i1[k], j = S.i2[k];
            if (S.status[k] != ACTIVE || S.levelg[i] < irec) continue;
            const double *ri = &S.rh[3 * (size_t)i], *rj = &S.rh[3 * (size_t)j];
            const double *vi = &S.vb[3 * (size_t)i], *vj = &S.vb[3 * (size_t)j];
            if (collision_one(ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2], vi[0] - vj[0], vi[1] - vj[1],
                              vi[2] - vj[2], S.Gm[i] + S.Gm[j], S.R[i] + S.R[j], dt, S.lvdotr[k] != 0)) {
                S.status[k] = COLLIDED;
                any = true;
            }
        }
        if (any) ncol++;
        if (!any) break;
    }
    return 0;
}

// Download the ensemble, resolve the merges of every system with a
// colliding pair, and put the result back on the device.
int Ensemble::resolve_merges(int irec, double dt)
{
    hipStream_t st = stream;
    MCHK(hipStreamSynchronize(st));
    const int n0 = npl;
    std::vector<double> rh(3 * (size_t)n0), vh(3 * (size_t)n0), vb(3 * (size_t)n0), R(n0), rhill(n0);
    std::vector<int32_t> status(n0), lg(n0), lm(n0), isperi;
    std::vector<double> atp, peri;
    {
        // device SoA -> host AoS
        std::vector<double> x(n0), y(n0), z(n0);
        auto down3 = [&](Vec3p v, std::vector<double> &out) -> int {
            MCHK(hipMemcpyAsync(x.data(), v.x, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(y.data(), v.y, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(z.data(), v.z, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            MCHK(hipStreamSynchronize(st));
            for (int i = 0; i < n0; i++) {
                out[3 * (size_t)i] = x[i];
                out[3 * (size_t)i + 1] = y[i];
                out[3 * (size_t)i + 2] = z[i];
            }
            return 0;
        };
        if (n0) {
            if (int rc = down3(pl.rh, rh)) return rc;
            if (int rc = down3(pl.vh, vh)) return rc;
            if (int rc = down3(pl.vb, vb)) return rc;
            MCHK(hipMemcpyAsync(R.data(), pl.R, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(rhill.data(), pl.rhill, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(status.data(), pl.status, n0 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(lg.data(), pl.levelg, n0 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(lm.data(), pl.levelm, n0 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        }
        if (dis_ready && n0) {
            isperi.resize(n0);
            atp.resize(n0);
            peri.resize(n0);
            MCHK(hipMemcpyAsync(isperi.data(), dis_pl_isperi, n0 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(atp.data(), dis_pl_atp, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
            MCHK(hipMemcpyAsync(peri.data(), dis_pl_peri, n0 * sizeof(double), hipMemcpyDeviceToHost, st));
        }
    }
    const int64_t m0 = npp;
    std::vector<int32_t> gi1(m0), gi2(m0);
    std::vector<int8_t> glv(m0), gst(m0), glev(m0);
    if (m0) {
        MCHK(hipMemcpyAsync(gi1.data(), pp.i1, m0 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        MCHK(hipMemcpyAsync(gi2.data(), pp.i2, m0 * sizeof(int32_t), hipMemcpyDeviceToHost, st));
        MCHK(hipMemcpyAsync(glv.data(), pp.lvdotr, m0, hipMemcpyDeviceToHost, st));
        MCHK(hipMemcpyAsync(gst.data(), pp.status, m0, hipMemcpyDeviceToHost, st));
        MCHK(hipMemcpyAsync(glev.data(), pp.level, m0, hipMemcpyDeviceToHost, st));
    }
    std::vector<double> cbvb(3 * (size_t)nsys);
    MCHK(hipMemcpyAsync(cbvb.data(), sy.cbvb, 3 * nsys * sizeof(double), hipMemcpyDeviceToHost, st));
    MCHK(hipStreamSynchronize(st));

    // one host system per system; the list is ordered by index1, so each
    // system's entries are one run of it
    std::vector<HostSys> sys(nsys);
    std::vector<char> hit(nsys, 0);
    {
        int64_t k = 0;
        for (int s = 0; s < nsys; s++) {
            const int a = h_pl_off[s], n = h_nreal_pl[s];
            HostSys &S = sys[s];
            S.n = n;
            S.rh.assign(rh.begin() + 3 * (size_t)a, rh.begin() + 3 * (size_t)(a + n));
            S.vh.assign(vh.begin() + 3 * (size_t)a, vh.begin() + 3 * (size_t)(a + n));
            S.vb.assign(vb.begin() + 3 * (size_t)a, vb.begin() + 3 * (size_t)(a + n));
            S.R.assign(R.begin() + a, R.begin() + a + n);
            S.rhill.assign(rhill.begin() + a, rhill.begin() + a + n);
            S.Gm.assign(h_Gm.begin() + a, h_Gm.begin() + a + n);
            S.mass.assign(h_mass.begin() + a, h_mass.begin() + a + n);
            S.id.assign(h_pl_id.begin() + a, h_pl_id.begin() + a + n);
            S.levelg.assign(lg.begin() + a, lg.begin() + a + n);
            S.levelm.assign(lm.begin() + a, lm.begin() + a + n);
            if (!isperi.empty()) {
                S.isperi.assign(isperi.begin() + a, isperi.begin() + a + n);
                S.atp.assign(atp.begin() + a, atp.begin() + a + n);
                S.peri.assign(peri.begin() + a, peri.begin() + a + n);
            }
            for (; k < m0 && gi1[k] < a + n; k++) {
                S.i1.push_back(gi1[k] - a);
                S.i2.push_back(gi2[k] - a);
                S.lvdotr.push_back(glv[k]);
                S.status.push_back(gst[k]);
                S.level.push_back(glev[k]);
                hit[s] |= gst[k] == COLLIDED;
            }
        }
        if (k != m0) return fail(SWG_ERR_COLLISION, "encounter list out of system order at a merge");
    }
    for (int s = 0; s < nsys; s++) {
        if (!hit[s]) continue;
        const int a = h_pl_off[s];
        std::vector<int32_t> st_s(status.begin() + a, status.begin() + a + sys[s].n);
        if (int rc = resolve_system(&sys[s], st_s, h_GMcb[s], irec, dt, h_maxid[s], h_nplm[s], st_.ncollision,
                                    &cbvb[3 * (size_t)s]))
            return rc;
    }
    return upload_merged(sys.data(), hit, cbvb.data(), irec);
}


// Put the merged ensemble on the device: bodies, list and its adjacency, the
// per-system sizes, tiles; then the working sets of levels 1..irec.  The
// systems without merges keep their device state; the arrays are written
// whole because the later systems' bodies move down.
int Ensemble::upload_merged(const void *hs, const std::vector<char> &hit, const double *cbvb, int irec)
{
    const HostSys *sys = (const HostSys *)hs;
    hipStream_t st = stream;
    // a slot pool keeps every slot's range and pads the rest of a shrunk slot
    // as swg_system_load does; other ensembles close up
    std::vector<int32_t> off(nsys + 1, 0);
    if (slot_pool)
        off = h_pl_off;
    else
        for (int s = 0; s < nsys; s++) off[s + 1] = off[s] + sys[s].n;
    const int n = off[nsys];
    // the body and list structures only shrink in a merge; the list may grow
    // when every old entry of a system was a collision
    if (n > npl) return fail(SWG_ERR_COLLISION, "merged ensemble larger than before");
    std::vector<double> rh(3 * (size_t)n), vh(3 * (size_t)n), vb(3 * (size_t)n), Gm(n), mass(n), R(n), rhill(n);
    std::vector<int32_t> id(n, 0), lg(n, -1), lm(n, -1), sysid(n), tiny(std::max(n, 1), 0), stat(std::max(n, 1), INACTIVE),
        isperi;
    std::vector<double> atp, peri;
    const bool dis = dis_ready && n;
    if (dis) {
        isperi.assign(n, 1);
        atp.assign(n, 0.0);
        peri.assign(n, 0.0);
    }
    std::vector<int32_t> i1, i2;
    std::vector<int8_t> lv, ls, ll;
    for (int s = 0; s < nsys; s++) {
        const HostSys &S = sys[s];
        const int a = off[s];
        std::copy(S.rh.begin(), S.rh.end(), rh.begin() + 3 * (size_t)a);
        std::copy(S.vh.begin(), S.vh.end(), vh.begin() + 3 * (size_t)a);
        std::copy(S.vb.begin(), S.vb.end(), vb.begin() + 3 * (size_t)a);
        std::copy(S.Gm.begin(), S.Gm.end(), Gm.begin() + a);
        std::copy(S.mass.begin(), S.mass.end(), mass.begin() + a);
        std::copy(S.R.begin(), S.R.end(), R.begin() + a);
        std::copy(S.rhill.begin(), S.rhill.end(), rhill.begin() + a);
        std::copy(S.id.begin(), S.id.end(), id.begin() + a);
        std::copy(S.levelg.begin(), S.levelg.end(), lg.begin() + a);
        std::copy(S.levelm.begin(), S.levelm.end(), lm.begin() + a);
        for (int q = 0; q < off[s + 1] - a; q++) {
            sysid[a + q] = s;
            if (q < S.n) {
                tiny[a + q] = prm.gmtiny > 0.0 && S.Gm[q] < prm.gmtiny;
                stat[a + q] = ACTIVE;
            } else {
                pad_position(q - S.n, &rh[3 * (size_t)(a + q)]);  // at rest, massless, inactive
            }
        }
        if (dis) {
            std::copy(S.isperi.begin(), S.isperi.end(), isperi.begin() + a);
            std::copy(S.atp.begin(), S.atp.end(), atp.begin() + a);
            std::copy(S.peri.begin(), S.peri.end(), peri.begin() + a);
        }
        for (size_t k = 0; k < S.i1.size(); k++) {
            if (k && S.i1[k] < S.i1[k - 1]) return fail(SWG_ERR_COLLISION, "encounter list out of row order after a merge");
            if (S.i1[k] >= S.i2[k]) return fail(SWG_ERR_COLLISION, "encounter pair out of order after a merge");
            i1.push_back(S.i1[k] + a);
            i2.push_back(S.i2[k] + a);
            lv.push_back((int8_t)S.lvdotr[k]);
            ls.push_back((int8_t)S.status[k]);
            ll.push_back((int8_t)S.level[k]);
        }
    }
    npl = n;
    h_pl_off = off;
    for (int s = 0; s < nsys; s++) h_nreal_pl[s] = sys[s].n;
    h_Gm = Gm;
    h_mass = mass;
    h_pl_id = id;
    {
        std::vector<double> x(n), y(n), z(n);
        auto up3 = [&](Vec3p v, const std::vector<double> &in) -> int {
            for (int i = 0; i < n; i++) {
                x[i] = in[3 * (size_t)i];
                y[i] = in[3 * (size_t)i + 1];
                z[i] = in[3 * (size_t)i + 2];
            }
            if (n) {
                MCHK(hipMemcpyAsync(v.x, x.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
                MCHK(hipMemcpyAsync(v.y, y.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
                MCHK(hipMemcpyAsync(v.z, z.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
            }
            MCHK(hipStreamSynchronize(st));
            return 0;
        };
        if (int rc = up3(pl.rh, rh)) return rc;
        if (int rc = up3(pl.vh, vh)) return rc;
        if (int rc = up3(pl.vb, vb)) return rc;
        // the padding's rbeg / rend are read as partner positions (zero
        // mass, but a distinct place for each, as swg_system_load leaves them)
        for (Vec3p v : {pl.rbeg, pl.rend})
            for (int s = 0; s < nsys && slot_pool; s++) {
                const int a = off[s] + sys[s].n, np = off[s + 1] - a;
                for (int i = 0; i < np; i++) {
                    x[i] = rh[3 * (size_t)(a + i)];
                    y[i] = rh[3 * (size_t)(a + i) + 1];
                    z[i] = rh[3 * (size_t)(a + i) + 2];
                }
                if (np) {
                    MCHK(hipMemcpyAsync(v.x + a, x.data(), np * sizeof(double), hipMemcpyHostToDevice, st));
                    MCHK(hipMemcpyAsync(v.y + a, y.data(), np * sizeof(double), hipMemcpyHostToDevice, st));
                    MCHK(hipMemcpyAsync(v.z + a, z.data(), np * sizeof(double), hipMemcpyHostToDevice, st));
                    MCHK(hipStreamSynchronize(st));
                }
            }
    }
    if (n) {
        MCHK(hipMemcpyAsync(pl.Gm, Gm.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.R, R.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.rhill, rhill.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.status, stat.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.sys, sysid.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.levelg, lg.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.levelm, lm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pl.tiny, tiny.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    if (dis) {
        MCHK(hipMemcpyAsync(dis_pl_isperi, isperi.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(dis_pl_atp, atp.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(dis_pl_peri, peri.data(), n * sizeof(double), hipMemcpyHostToDevice, st));
    }
    MCHK(hipMemcpyAsync(sy.cbvb, cbvb, 3 * nsys * sizeof(double), hipMemcpyHostToDevice, st));
    MCHK(hipMemcpyAsync(sy.pl_off, h_pl_off.data(), (nsys + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    MCHK(hipMemcpyAsync(sy.nplm, h_nplm.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice, st));
    // bit-matrix layout of the next step's detection and its column blocks
    std::vector<int64_t> woff(nsys);
    std::vector<int32_t> nb(nsys), cb(nsys + 1, 0);
    {
        int64_t w = 0;
        for (int s = 0; s < nsys; s++) {
            const int ns = off[s + 1] - off[s];
            nb[s] = (ns + 63) / 64;
            woff[s] = w;
            w += (int64_t)ns * nb[s];
            cb[s + 1] = cb[s] + nb[s];
        }
        nblk_pl = cb[nsys];
        MCHK(hipMemcpyAsync(pp.woff, woff.data(), nsys * sizeof(int64_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.nb, nb.data(), nsys * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(d_cb_pl, cb.data(), (nsys + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipStreamSynchronize(st));
    }
    // the list, its rows (index1 side, grouped in list order) and columns
    const int64_t m = (int64_t)i1.size();
    if (int rc = ensure_list(pp, m, cap_pp)) return rc;
    npp = m;
    std::vector<int32_t> rowcnt(n, 0), colcnt(n, 0), col(std::max<int64_t>(m, 1));
    std::vector<int64_t> rowstart(n, 0), colstart(n, 0);
    for (int64_t k = 0; k < m; k++) {
        rowcnt[i1[k]]++;
        colcnt[i2[k]]++;
    }
    for (int i = 1; i < n; i++) {
        rowstart[i] = rowstart[i - 1] + rowcnt[i - 1];
        colstart[i] = colstart[i - 1] + colcnt[i - 1];
    }
    {
        std::vector<int64_t> fill(colstart);
        for (int64_t k = 0; k < m; k++) col[fill[i2[k]]++] = (int32_t)k;
    }
    if (m) {
        MCHK(hipMemcpyAsync(pp.i1, i1.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.i2, i2.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.lvdotr, lv.data(), m, hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.status, ls.data(), m, hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.level, ll.data(), m, hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.col, col.data(), m * sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    if (n) {
        MCHK(hipMemcpyAsync(pp.rowcnt, rowcnt.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.rowstart, rowstart.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.colcnt, colcnt.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(pp.colstart, colstart.data(), n * sizeof(int64_t), hipMemcpyHostToDevice, st));
    }
    // sy.enc_pp keeps the step's level-0 flags: a system took the encounter
    // path whatever the merges leave in its list (it also decides lfirst)
    MCHK(hipStreamSynchronize(st));
    any_tiny = false;
    for (int s = 0; s < nsys; s++) any_tiny |= h_nplm[s] < sys[s].n;
    tune.any_tiny = any_tiny;
    if (int rc = build_tiles()) return rc;
    if (en_ready) {
        // the energy tiles follow the body counts
        std::vector<int4> tiles;
        std::vector<int32_t> toff;
        energy_tiles(nsys, h_pl_off.data(), tiles, toff);
        if ((int)tiles.size() > en_ntiles) return fail(SWG_ERR_COLLISION, "energy tiles grew in a merge");
        en_ntiles = (int)tiles.size();
        if (en_ntiles)
            MCHK(hipMemcpyAsync(en_tiles, tiles.data(), tiles.size() * sizeof(int4), hipMemcpyHostToDevice, st));
        MCHK(hipMemcpyAsync(en_toff, toff.data(), toff.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
        MCHK(hipStreamSynchronize(st));
    }
    drop_graphs();
    (void)hit;
    // the working sets of the levels the recursion is in, from the new list
    for (int k = 1; k <= irec; k++) {
        const LevelView *src = k == 1 ? nullptr : &lvv[k - 1];
        if (int rc = build_begin(k, src)) return rc;
        int32_t lenc = 0;
        if (int rc = build_wait(lenc)) return rc;
        bool handoff = false;
        if (int rc = build_end(k, src, lvv[k], lvc[k], handoff)) return rc;
    }
    MCHK(hipStreamSynchronize(st));
    return 0;
}

int Ensemble::get_bodies(int32_t *id, double *mass, double *Gmass, double *radius)
{
    MCHK(hipStreamSynchronize(stream));
    if (id) std::copy(h_pl_id.begin(), h_pl_id.end(), id);
    if (mass) {
        if (h_mass.size() == (size_t)npl)
            std::copy(h_mass.begin(), h_mass.end(), mass);
        else
            std::fill(mass, mass + npl, 0.0);
    }
    if (Gmass) std::copy(h_Gm.begin(), h_Gm.end(), Gmass);
    if (radius && npl) MCHK(hipMemcpy(radius, pl.R, npl * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace swg
