// k_energy.hip -- energy and angular momentum of every system.
// Follows swiftest_util_get_energy_and_momentum_system (swiftest_util.f90:
// 1172-1291), after swiftest_util_coord_rh2rb_pl (:488-518) and with the
// triangular potential (:1350-1400). Rotation and oblateness are off.
//
// k_en_body runs one workgroup per system. It forms the barycentre offset
// from sum(Gm*rh) over the bodies that are not INACTIVE. Each body then gets
// its barycentric position, kept for the pair pass. The workgroup sums the
// kinetic terms, angular momenta, central-body potentials and binding
// energies.
//
// k_en_pairs forms the pair potential -Gm_i*m_j/|rb_i - rb_j| over i < j, in
// tiles of EN_TI rows by EN_TJ columns. A thread walks one row of a tile,
// with the columns staged through LDS. Each tile's sum goes to its own slot,
// and k_en_final adds a system's slots in a fixed order.
//
// The sums run in tree order rather than the reference's serial order. |r|
// is the square root of the dot product, not gfortran's scaled norm2. The
// results therefore agree with the reference to rounding, and each run
// gives the same bits.
#include <vector>

#include "../../include/swiftest_hip.h"
#include "swg_kernels.h"

namespace swg {

constexpr int EN_T = 256, EN_TI = EN_T, EN_TJ = 2048;

// workgroup sum of one value per thread in a fixed tree order; every thread
// gets the result
__device__ inline double en_block_sum(double v, double *sh)
{
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int o = EN_T / 2; o > 0; o >>= 1) {
        if (tid < o) sh[tid] = sh[tid] + sh[tid + o];
        __syncthreads();
    }
    const double r = sh[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(EN_T) void k_en_body(PlDev pl, SysDev sy, const double *mass, const double *cbm,
                                                  const double *cbR, int lclose, Vec3p rb, double *out)
{
    __shared__ double sh[EN_T];
    const int s = blockIdx.x, tid = threadIdx.x;
    const int p0 = sy.pl_off[s], p1 = sy.pl_off[s + 1];
    const double GMcb = sy.GMcb[s];
    // swiftest_util_coord_rh2rb_pl (:505-516); a system without bodies keeps
    // the central body at the origin
    double gm = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
    for (int i = p0 + tid; i < p1; i += EN_T) {
        if (pl.status[i] == INACTIVE) continue;
        const double g = pl.Gm[i];
        gm = gm + g;
        gx = gx + g * pl.rh.x[i];
        gy = gy + g * pl.rh.y[i];
        gz = gz + g * pl.rh.z[i];
    }
    gm = en_block_sum(gm, sh);
    gx = en_block_sum(gx, sh);
    gy = en_block_sum(gy, sh);
    gz = en_block_sum(gz, sh);
    const double Gmtot = GMcb + gm;
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if (p1 > p0) {
        cx = -gx / Gmtot;
        cy = -gy / Gmtot;
        cz = -gz / Gmtot;
    }
    // per-body terms (:1209-1223, :1361-1370, :1278-1282)
    double ke = 0.0, lx = 0.0, ly = 0.0, lz = 0.0, pecb = 0.0, be = 0.0;
    for (int i = p0 + tid; i < p1; i += EN_T) {
        if (pl.status[i] == INACTIVE) continue;
        const double x = pl.rh.x[i] + cx, y = pl.rh.y[i] + cy, z = pl.rh.z[i] + cz;
        rb.x[i] = x;
        rb.y[i] = y;
        rb.z[i] = z;
        const double vx = pl.vb.x[i], vy = pl.vb.y[i], vz = pl.vb.z[i], m = mass[i];
        lx = lx + m * (y * vz - z * vy);
        ly = ly + m * (z * vx - x * vz);
        lz = lz + m * (x * vy - y * vx);
        ke = ke + m * dot3(vx, vy, vz, vx, vy, vz);
        pecb = pecb + (-(GMcb * m) / sqrt(dot3(x, y, z, x, y, z)));
        if (lclose) be = be + (-(3.0 * pl.Gm[i] * m) / (5.0 * pl.R[i]));
    }
    ke = en_block_sum(ke, sh);
    lx = en_block_sum(lx, sh);
    ly = en_block_sum(ly, sh);
    lz = en_block_sum(lz, sh);
    pecb = en_block_sum(pecb, sh);
    be = en_block_sum(be, sh);
    if (tid == 0) {
        const double mc = cbm[s], *v = sy.cbvb + 3 * s;
        const double kecb = mc * dot3(v[0], v[1], v[2], v[0], v[1], v[2]);
        double *o = out + (size_t)s * SWG_ENERGY_NOUT;
        for (int k = 0; k < SWG_ENERGY_NOUT; k++) o[k] = 0.0;
        o[SWG_E_KE_ORBIT] = 0.5 * (kecb + ke);
        o[SWG_E_PE] = pecb;  // the pair sum is added in front of it by k_en_final
        o[SWG_E_BE] = be;
        o[SWG_E_BE_CB] = -(3.0 * GMcb * mc) / (5.0 * cbR[s]);
        o[SWG_E_L_ORBIT + 0] = mc * (cy * v[2] - cz * v[1]) + lx;
        o[SWG_E_L_ORBIT + 1] = mc * (cz * v[0] - cx * v[2]) + ly;
        o[SWG_E_L_ORBIT + 2] = mc * (cx * v[1] - cy * v[0]) + lz;
        o[SWG_E_GMTOT] = Gmtot;
    }
}

// One tile: rows [i0, i0 + EN_TI) and columns [j0, j0 + EN_TJ) of system s,
// cut to the system's bodies; only j > i counts.
__global__ __launch_bounds__(EN_T) void k_en_pairs(PlDev pl, const double *mass, Vec3p rb, const int4 *tiles,
                                                   const int32_t *pl_off, double *part)
{
    __shared__ double cx[EN_T], cy[EN_T], cz[EN_T], cm[EN_T];
    __shared__ double sh[EN_T];
    const int4 t = tiles[blockIdx.x];
    const int s = t.x, i0 = t.y, j0 = t.z;
    const int p1 = pl_off[s + 1];
    const int tid = threadIdx.x, i = i0 + tid;
    const bool row = i < p1 && pl.status[i] != INACTIVE;
    const double xi = row ? rb.x[i] : 0.0, yi = row ? rb.y[i] : 0.0, zi = row ? rb.z[i] : 0.0;
    const double gi = row ? pl.Gm[i] : 0.0;
    const int j1 = min(j0 + EN_TJ, p1);
    double acc = 0.0;
    for (int c0 = j0; c0 < j1; c0 += EN_T) {
        const int j = c0 + tid;
        const bool col = j < j1 && pl.status[j] != INACTIVE;
        cx[tid] = col ? rb.x[j] : 0.0;
        cy[tid] = col ? rb.y[j] : 0.0;
        cz[tid] = col ? rb.z[j] : 0.0;
        cm[tid] = col ? mass[j] : -1.0;  // -1 marks a column left out
        __syncthreads();
        if (row) {
            const int kb = max(0, i + 1 - c0), ke = min(EN_T, j1 - c0);
            for (int k = kb; k < ke; k++) {
                if (cm[k] < 0.0) continue;
                const double dx = xi - cx[k], dy = yi - cy[k], dz = zi - cz[k];
                acc = acc + (-(gi * cm[k]) / sqrt(dot3(dx, dy, dz, dx, dy, dz)));
            }
        }
        __syncthreads();
    }
    acc = en_block_sum(acc, sh);
    if (tid == 0) part[blockIdx.x] = acc;
}

__global__ __launch_bounds__(EN_T) void k_en_final(const double *part, const int32_t *toff, double *out)
{
    __shared__ double sh[EN_T];
    const int s = blockIdx.x, tid = threadIdx.x;
    double acc = 0.0;
    for (int k = toff[s] + tid; k < toff[s + 1]; k += EN_T) acc = acc + part[k];
    acc = en_block_sum(acc, sh);
    if (tid == 0) {
        double *o = out + (size_t)s * SWG_ENERGY_NOUT;
        o[SWG_E_PE] = acc + o[SWG_E_PE];
        o[SWG_E_TE] = ((o[SWG_E_KE_ORBIT] + o[SWG_E_KE_SPIN]) + o[SWG_E_PE]) + o[SWG_E_BE];
        for (int d = 0; d < 3; d++) o[SWG_E_L_TOTAL + d] = o[SWG_E_L_ORBIT + d] + o[SWG_E_L_SPIN + d];
    }
}

void energy_tiles(int nsys, const int32_t *pl_off, std::vector<int4> &tiles, std::vector<int32_t> &toff)
{
    tiles.clear();
    toff.assign(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        toff[s] = (int32_t)tiles.size();
        const int p0 = pl_off[s], p1 = pl_off[s + 1];
        for (int i0 = p0; i0 < p1; i0 += EN_TI)
            for (int j0 = p0; j0 < p1; j0 += EN_TJ)
                if (j0 + EN_TJ > i0 + 1) tiles.push_back(make_int4(s, i0, j0, 0));
    }
    toff[nsys] = (int32_t)tiles.size();
}

void launch_energy(hipStream_t st, PlDev pl, SysDev sy, int nsys, const double *mass, const double *cbm,
                   const double *cbR, int lclose, Vec3p rb, const int4 *tiles, int ntiles, const int32_t *toff,
                   double *part, double *out)
{
    if (nsys <= 0) return;
    k_en_body<<<nsys, EN_T, 0, st>>>(pl, sy, mass, cbm, cbR, lclose, rb, out);
    if (ntiles) k_en_pairs<<<ntiles, EN_T, 0, st>>>(pl, mass, rb, tiles, sy.pl_off, part);
    k_en_final<<<nsys, EN_T, 0, st>>>(part, toff, out);
}

}  // namespace swg
