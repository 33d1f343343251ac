// k_discard.hip -- the discard checks run after each step, on the resident
// state.
//
// The order is the reference's:
// - For SyMBA, the massive-body checks come first (symba_discard_pl,
//   symba_discard.f90:342-387). These are rh2rb, the central-body and escape
//   tests (symba_discard_cb_pl, :14-104) and the pericentre test
//   (symba_discard_peri_pl, :292-339).
// - Then the test-particle checks (swiftest_discard_tp,
//   swiftest_discard.f90:80-119). These are pl%h2b and tp%h2b, which
//   overwrite pl vb, cb vb and tp vb as the reference does, then the
//   central-body and escape tests (:122-191), the pericentre test (:194-243)
//   and the close approach to a massive body (:246-337).
//
// Massive bodies discarded in the call leave the reference's arrays
// (pl%rearray) before the test particles are checked. Here they stay in
// place with their new status and a "gone" mark, and the test-particle
// pass skips them. The host rebuilds a system that lost massive bodies.
//
// The barycentre sums are serial in body order, as in the reference (one
// workgroup per system; dedicated threads fold LDS-staged chunks). The
// per-body tests run one thread per body. The close-approach test walks a
// system's massive bodies in order through LDS tiles and stops at the
// first hit.
#include <vector>

#include "../../include/swiftest_hip.h"
#include "swg_kernels.h"

namespace swg {

constexpr int DS_T = 256, DS_CH = 1024;

// per-system scratch words (doubles): see DS_*
constexpr int DS_GMTOT = 0, DS_RB = 1, DS_RB2 = 4, DS_VB2 = 7, DS_H2B = 10, DS_NW = 12;

struct DisPrm {
    double rmin, rmax, rmaxu, qmin, alo, ahi, dt;
    int qmin_coord, lclose, symba, lcb, h2b;
};

struct DisState {
    int32_t *pl_isperi, *tp_isperi, *lfirst_peri;
    double *pl_atp, *pl_peri, *tp_atp, *tp_peri;
    uint8_t *gone;
    double *sys;                // DS_NW per system
    unsigned long long *count;  // [0] massive bodies, [1] test particles
    const int32_t *nplenc_row, *nplenc_col;  // level-0 pl-pl pairs per body (null: none)
};

// swiftest_util_peri (swiftest_util.f90:1565-1603) for one body
__device__ inline void peri_one(double m, double rx, double ry, double rz, double vx, double vy, double vz,
                                int32_t &isperi, double &atp, double &q)
{
    const double vdotr = dot3(rx, ry, rz, vx, vy, vz);
    if (isperi == -1) {
        if (vdotr >= 0.0) {
            isperi = 0;
            xv2aq(m, rx, ry, rz, vx, vy, vz, atp, q);
        }
    } else {
        isperi = vdotr > 0.0 ? -1 : 1;
    }
}

// swiftest_discard_pl_close (swiftest_discard.f90:291-337)
__device__ inline bool pl_close(double dx, double dy, double dz, double ux, double uy, double uz, double dt,
                                double r2crit)
{
    const double r2 = dot3(dx, dy, dz, dx, dy, dz);
    if (r2 <= r2crit) return true;
    const double vdotr = dot3(dx, dy, dz, ux, uy, uz);
    if (vdotr > 0.0) return false;
    const double v2 = dot3(ux, uy, uz, ux, uy, uz);
    const double tmin = -vdotr / v2;
    double r2min;
    if (tmin < dt)
        r2min = r2 - (vdotr * vdotr) / v2;
    else
        r2min = (r2 + (2.0 * vdotr) * dt) + v2 * (dt * dt);
    r2min = fmin(r2min, r2);
    return r2min <= r2crit;
}

// Serial sums over one system's massive bodies, in order, each on its own
// thread: the staged chunk is folded by threads 0..nq-1.
struct DsStage {
    double gm[DS_CH], rx[DS_CH], ry[DS_CH], rz[DS_CH], vx[DS_CH], vy[DS_CH], vz[DS_CH];
    uint8_t in[DS_CH];
};

// set_msys (swiftest_util.f90:2115) and, for SyMBA, pl%rh2rb (:505-516)
__global__ __launch_bounds__(DS_T) void k_dis_sums1(PlDev pl, SysDev sy, DisPrm p, DisState d)
{
    __shared__ DsStage S;
    __shared__ double res[8];
    const int s = blockIdx.x, tid = threadIdx.x;
    if (sy.frozen[s]) return;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    const double GMcb = sy.GMcb[s];
    double acc = tid == 1 ? GMcb : 0.0;  // 0: set_msys sum, 1: rh2rb Gmtot, 2-4: sum Gm*rh
    for (int c0 = p0; c0 < p1; c0 += DS_CH) {
        const int n = min(DS_CH, p1 - c0);
        for (int k = tid; k < n; k += DS_T) {
            const int i = c0 + k;
            S.gm[k] = pl.Gm[i];
            S.rx[k] = pl.rh.x[i];
            S.ry[k] = pl.rh.y[i];
            S.rz[k] = pl.rh.z[i];
            S.in[k] = pl.status[i] != INACTIVE;
        }
        __syncthreads();
        if (tid == 0 || tid == 1) {
            for (int k = 0; k < n; k++)
                if (S.in[k]) acc = acc + S.gm[k];
        } else if (tid <= 4) {
            const double *r = tid == 2 ? S.rx : tid == 3 ? S.ry : S.rz;
            for (int k = 0; k < n; k++)
                if (S.in[k]) acc = acc + S.gm[k] * r[k];
        }
        __syncthreads();
    }
    if (tid < 5) res[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        double *w = d.sys + (size_t)s * DS_NW;
        w[DS_GMTOT] = GMcb + res[0];
        for (int k = 0; k < 3; k++) w[DS_RB + k] = p1 > p0 ? -res[2 + k] / res[1] : 0.0;
    }
}

// symba_discard_cb_pl and symba_discard_peri_pl, one thread per body
__global__ __launch_bounds__(DS_T) void k_dis_pl(PlDev pl, SysDev sy, DisPrm p, DisState d, int npl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npl) return;
    const int s = pl.sys[i];
    if (sy.frozen[s]) return;
    const double *w = d.sys + (size_t)s * DS_NW;
    const double Gmtot = w[DS_GMTOT];
    const bool in = pl.status[i] != INACTIVE;
    const double rbx = in ? pl.rh.x[i] + w[DS_RB] : 0.0, rby = in ? pl.rh.y[i] + w[DS_RB + 1] : 0.0,
                 rbz = in ? pl.rh.z[i] + w[DS_RB + 2] : 0.0;
    const double vx = pl.vb.x[i], vy = pl.vb.y[i], vz = pl.vb.z[i];
    int32_t st = pl.status[i];
    bool dis = false;
    if (p.lcb && st == ACTIVE) {
        const double rx = pl.rh.x[i], ry = pl.rh.y[i], rz = pl.rh.z[i];
        const double rh2 = dot3(rx, ry, rz, rx, ry, rz);
        if (p.rmax >= 0.0 && rh2 > p.rmax * p.rmax) {
            st = DISCARDED_RMAX;
        } else if (p.rmin >= 0.0 && rh2 < p.rmin * p.rmin) {
            st = DISCARDED_RMIN;
        } else if (p.rmaxu >= 0.0) {
            const double rb2 = dot3(rbx, rby, rbz, rbx, rby, rbz);
            const double vb2 = dot3(vx, vy, vz, vx, vy, vz);
            const double energy = 0.5 * vb2 - Gmtot / sqrt(rb2);
            if (energy > 0.0 && rb2 > p.rmaxu * p.rmaxu) st = DISCARDED_RMAXU;
        }
        dis = st != ACTIVE;
    }
    if (p.qmin >= 0.0) {
        const bool first = d.lfirst_peri[s] != 0;
        int32_t isp = first ? 0 : d.pl_isperi[i];
        double atp = d.pl_atp[i], q = d.pl_peri[i];
        if (p.qmin_coord == 0)
            peri_one(sy.GMcb[s] + pl.Gm[i], pl.rh.x[i], pl.rh.y[i], pl.rh.z[i], pl.vh.x[i], pl.vh.y[i], pl.vh.z[i],
                     isp, atp, q);
        else
            peri_one(Gmtot, rbx, rby, rbz, vx, vy, vz, isp, atp, q);
        d.pl_isperi[i] = isp;
        d.pl_atp[i] = atp;
        d.pl_peri[i] = q;
        if (!first && st == ACTIVE && isp == 0) {
            const int nenc = d.nplenc_row ? d.nplenc_row[i] + d.nplenc_col[i] : 0;
            if (nenc == 0 && atp >= p.alo && atp <= p.ahi && q <= p.qmin) {
                st = DISCARDED_PERI;
                dis = true;
            }
        }
    }
    if (dis) {
        pl.status[i] = st;
        d.gone[i] = 1;
        atomicAdd(&d.count[0], 1ull);
    }
}

// pl%h2b over the bodies still in the arrays (swiftest_util.f90:228-265);
// clears the first-call flag of the pericentre test
__global__ __launch_bounds__(DS_T) void k_dis_sums2(PlDev pl, SysDev sy, DisPrm p, DisState d)
{
    __shared__ DsStage S;
    __shared__ double res[8];
    __shared__ int nleft;
    const int s = blockIdx.x, tid = threadIdx.x;
    if (sy.frozen[s]) return;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    double *w = d.sys + (size_t)s * DS_NW;
    if (tid == 0) {
        nleft = 0;
        if (p.symba && p.qmin >= 0.0 && p1 > p0) d.lfirst_peri[s] = 0;
    }
    const bool run = p.h2b && sy.tp_off[s + 1] > sy.tp_off[s];
    if (!run) {
        if (tid == 0) w[DS_H2B] = 0.0;
        return;
    }
    const double GMcb = sy.GMcb[s];
    double acc = tid == 0 ? GMcb : 0.0;  // 0: Gmtot, 1-3: sum Gm*rh, 4-6: sum Gm*vh
    int left = 0;
    for (int c0 = p0; c0 < p1; c0 += DS_CH) {
        const int n = min(DS_CH, p1 - c0);
        for (int k = tid; k < n; k += DS_T) {
            const int i = c0 + k;
            S.gm[k] = pl.Gm[i];
            S.rx[k] = pl.rh.x[i];
            S.ry[k] = pl.rh.y[i];
            S.rz[k] = pl.rh.z[i];
            S.vx[k] = pl.vh.x[i];
            S.vy[k] = pl.vh.y[i];
            S.vz[k] = pl.vh.z[i];
            S.in[k] = (d.gone[i] ? 0 : 1) | (pl.status[i] != INACTIVE ? 2 : 0);
            left += d.gone[i] ? 0 : 1;
        }
        __syncthreads();
        if (tid == 0) {
            for (int k = 0; k < n; k++)
                if (S.in[k] == 3) acc = acc + S.gm[k];
        } else if (tid <= 6) {
            const double *q = tid == 1 ? S.rx : tid == 2 ? S.ry : tid == 3 ? S.rz : tid == 4 ? S.vx : tid == 5 ? S.vy : S.vz;
            for (int k = 0; k < n; k++)
                if (S.in[k] == 3) acc = acc + S.gm[k] * q[k];
        }
        __syncthreads();
    }
    if (left) atomicAdd(&nleft, left);
    if (tid < 7) res[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        if (nleft > 0) {
            for (int k = 0; k < 3; k++) {
                w[DS_RB2 + k] = -res[1 + k] / res[0];
                w[DS_VB2 + k] = -res[4 + k] / res[0];
                sy.cbvb[3 * s + k] = w[DS_VB2 + k];
            }
        } else {  // no massive body left: cb%rb and cb%vb stay as they are
            for (int k = 0; k < 3; k++) {
                w[DS_RB2 + k] = p.symba && p1 > p0 ? w[DS_RB + k] : 0.0;
                w[DS_VB2 + k] = sy.cbvb[3 * s + k];
            }
        }
        w[DS_H2B] = 1.0;
    }
}

// pl vb = vh + cb vb for the bodies h2b covers
__global__ __launch_bounds__(DS_T) void k_dis_h2b_pl(PlDev pl, SysDev sy, DisState d, int npl)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npl) return;
    const int s = pl.sys[i];
    if (sy.frozen[s]) return;
    const double *w = d.sys + (size_t)s * DS_NW;
    if (w[DS_H2B] == 0.0 || d.gone[i] || pl.status[i] == INACTIVE) return;
    pl.vb.x[i] = pl.vh.x[i] + w[DS_VB2];
    pl.vb.y[i] = pl.vh.y[i] + w[DS_VB2 + 1];
    pl.vb.z[i] = pl.vh.z[i] + w[DS_VB2 + 2];
}

// The test-particle checks.  One workgroup per tile of up to DS_T test
// particles of one system; the system's massive bodies are walked in order
// through LDS for the Hill-sphere and close-approach tests.
__global__ __launch_bounds__(DS_T) void k_dis_tp(PlDev pl, TpDev tp, SysDev sy, DisPrm p, DisState d,
                                                 const int4 *tiles)
{
    __shared__ double jx[DS_T], jy[DS_T], jz[DS_T], ju[DS_T], jv[DS_T], jw[DS_T], jr[DS_T];
    __shared__ uint8_t jin[DS_T];
    const int4 tl = tiles[blockIdx.x];
    const int s = tl.x, t = tl.y + threadIdx.x, tid = threadIdx.x;
    if (sy.frozen[s]) return;
    const bool mine = t < tl.z;
    const double *w = d.sys + (size_t)s * DS_NW;
    const bool h2b = w[DS_H2B] != 0.0;
    int32_t st = mine ? tp.status[t] : INACTIVE;
    const double rx = mine ? tp.rh.x[t] : 0.0, ry = mine ? tp.rh.y[t] : 0.0, rz = mine ? tp.rh.z[t] : 0.0;
    const double ux = mine ? tp.vh.x[t] : 0.0, uy = mine ? tp.vh.y[t] : 0.0, uz = mine ? tp.vh.z[t] : 0.0;
    double bx = 0.0, by = 0.0, bz = 0.0, vx = 0.0, vy = 0.0, vz = 0.0;
    if (mine) {
        vx = tp.vb.x[t];
        vy = tp.vb.y[t];
        vz = tp.vb.z[t];
        if (h2b && st != INACTIVE) {  // tp%h2b (swiftest_util.f90:268-296)
            bx = rx + w[DS_RB2];
            by = ry + w[DS_RB2 + 1];
            bz = rz + w[DS_RB2 + 2];
            vx = ux + w[DS_VB2];
            vy = uy + w[DS_VB2 + 1];
            vz = uz + w[DS_VB2 + 2];
            tp.vb.x[t] = vx;
            tp.vb.y[t] = vy;
            tp.vb.z[t] = vz;
        }
    }
    const int32_t st0 = st;
    const double Gmtot = w[DS_GMTOT];
    if (p.lcb && st == ACTIVE) {  // swiftest_discard_cb_tp (swiftest_discard.f90:136-188)
        const double rh2 = dot3(rx, ry, rz, rx, ry, rz), Rcb = d.sys[(size_t)s * DS_NW + DS_H2B + 1];
        if (p.rmax >= 0.0 && rh2 > p.rmax * p.rmax) {
            st = DISCARDED_RMAX;
        } else if (p.rmin >= 0.0 && rh2 < fmax(p.rmin * p.rmin, Rcb * Rcb)) {
            st = DISCARDED_RMIN;
        } else if (p.rmaxu >= 0.0) {
            const double rb2 = dot3(bx, by, bz, bx, by, bz);
            const double vb2 = dot3(vx, vy, vz, vx, vy, vz);
            const double energy = 0.5 * vb2 - Gmtot / sqrt(rb2);
            if (energy > 0.0 && rb2 > p.rmaxu * p.rmaxu) st = DISCARDED_RMAXU;
        }
    }
    bool hill = false;  // this particle needs the Hill-sphere walk
    double atp = 0.0, q = 0.0;
    if (p.qmin >= 0.0 && mine) {  // swiftest_discard_peri_tp (:209-240)
        int32_t isp = d.tp_isperi[t];
        atp = d.tp_atp[t];
        q = d.tp_peri[t];
        if (p.qmin_coord == 0)
            peri_one(sy.GMcb[s], rx, ry, rz, ux, uy, uz, isp, atp, q);
        else
            peri_one(Gmtot, bx, by, bz, vx, vy, vz, isp, atp, q);
        d.tp_isperi[t] = isp;
        d.tp_atp[t] = atp;
        d.tp_peri[t] = q;
        hill = st == ACTIVE && isp == 0;
    }
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    if (__syncthreads_or(hill)) {
        bool ih = true;
        for (int c0 = p0; c0 < p1; c0 += DS_T) {
            const int j = c0 + tid;
            if (j < p1) {
                jx[tid] = pl.rh.x[j];
                jy[tid] = pl.rh.y[j];
                jz[tid] = pl.rh.z[j];
                jr[tid] = pl.rhill[j];
                jin[tid] = !d.gone[j];
            }
            __syncthreads();
            if (hill) {
                const int n = min(DS_T, p1 - c0);
                for (int k = 0; k < n; k++) {
                    if (!jin[k]) continue;
                    const double dx = rx - jx[k], dy = ry - jy[k], dz = rz - jz[k];
                    if (dot3(dx, dy, dz, dx, dy, dz) <= jr[k] * jr[k]) ih = false;
                }
            }
            __syncthreads();
        }
        if (hill && ih && atp >= p.alo && atp <= p.ahi && q <= p.qmin) st = DISCARDED_PERI;
    }
    bool look = p.lclose && st == ACTIVE;  // swiftest_discard_pl_tp (:260-285)
    for (int c0 = p0; __syncthreads_or(look) && c0 < p1; c0 += DS_T) {
        const int j = c0 + tid;
        if (j < p1) {
            jx[tid] = pl.rh.x[j];
            jy[tid] = pl.rh.y[j];
            jz[tid] = pl.rh.z[j];
            ju[tid] = pl.vh.x[j];
            jv[tid] = pl.vh.y[j];
            jw[tid] = pl.vh.z[j];
            jr[tid] = pl.R[j] * pl.R[j];
            jin[tid] = !d.gone[j];
        }
        __syncthreads();
        if (look) {
            const int n = min(DS_T, p1 - c0);
            for (int k = 0; k < n; k++) {
                if (!jin[k]) continue;
                if (pl_close(rx - jx[k], ry - jy[k], rz - jz[k], ux - ju[k], uy - jv[k], uz - jw[k], p.dt, jr[k])) {
                    st = DISCARDED_PLR;
                    look = false;
                    break;
                }
            }
        }
        __syncthreads();
    }
    if (mine && st != st0) {
        tp.status[t] = st;
        atomicAdd(&d.count[1], 1ull);
    }
}

__global__ void k_dis_prep(SysDev sy, DisState d, const double *Rcb, int nsys, int npl)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < 2) d.count[k] = 0ull;
    if (k < nsys) d.sys[(size_t)k * DS_NW + DS_H2B + 1] = Rcb[k];
    if (k < npl) d.gone[k] = 0;
}

void discard_tiles(int nsys, const int32_t *tp_off, std::vector<int4> &tiles)
{
    tiles.clear();
    for (int s = 0; s < nsys; s++)
        for (int t0 = tp_off[s]; t0 < tp_off[s + 1]; t0 += DS_T)
            tiles.push_back(make_int4(s, t0, std::min(t0 + DS_T, tp_off[s + 1]), 0));
}

size_t discard_sys_words() { return DS_NW; }

void launch_discard(hipStream_t st, PlDev pl, TpDev tp, SysDev sy, int nsys, int npl, int symba,
                    const swg_discard_params &dp, int lclose, double dt, const double *Rcb,
                    const int32_t *nplenc_row, const int32_t *nplenc_col, int32_t *pl_isperi, double *pl_atp,
                    double *pl_peri, int32_t *tp_isperi, double *tp_atp, double *tp_peri, int32_t *lfirst_peri,
                    uint8_t *gone, double *sys, unsigned long long *count, const int4 *tiles, int ntiles)
{
    DisPrm p;
    p.rmin = dp.rmin;
    p.rmax = dp.rmax;
    p.rmaxu = dp.rmaxu;
    p.qmin = dp.qmin;
    p.alo = dp.qmin_alo;
    p.ahi = dp.qmin_ahi;
    p.dt = dt;
    p.qmin_coord = dp.qmin_coord;
    p.lclose = lclose;
    p.symba = symba;
    p.lcb = dp.rmin >= 0.0 || dp.rmax >= 0.0 || dp.rmaxu >= 0.0;
    p.h2b = p.lcb || (dp.qmin >= 0.0 && dp.qmin_coord == 1);
    DisState d{pl_isperi, tp_isperi, lfirst_peri, pl_atp, pl_peri, tp_atp, tp_peri, gone, sys, count,
               nplenc_row, nplenc_col};
    const int n = std::max(std::max(nsys, npl), 2);
    k_dis_prep<<<(n + 255) / 256, 256, 0, st>>>(sy, d, Rcb, nsys, npl);
    k_dis_sums1<<<nsys, DS_T, 0, st>>>(pl, sy, p, d);
    if (symba && npl) k_dis_pl<<<(npl + DS_T - 1) / DS_T, DS_T, 0, st>>>(pl, sy, p, d, npl);
    k_dis_sums2<<<nsys, DS_T, 0, st>>>(pl, sy, p, d);
    if (npl) k_dis_h2b_pl<<<(npl + DS_T - 1) / DS_T, DS_T, 0, st>>>(pl, sy, d, npl);
    if (ntiles) k_dis_tp<<<ntiles, DS_T, 0, st>>>(pl, tp, sy, p, d, tiles);
}

}  // namespace swg
