// Kernels shared by the straight and curved paths: time-step reduction
// (getdelt_lag, Getrhslag.f90:20600-20713), mass-matrix solve with the
// SSP-RK blend (Getrhslag.f90:1804-1904), node motion (:1954-1956), and the
// cell volume / volume centre (getcellvol_lag, Limiterlag.f90:1999-2162)
// fused with the Q2 centre-node reset (Getrhslag.f90:1962-1971).
#include "lag2d_dev.h"
#include "lag2d_kernels.h"

namespace lag2d {

// Per-block minima of the CFL time (edge length / sound speed) and of the
// volume-change time V/|dV/dt| over the quads, with dV/dt the degree-0
// tau right-hand side before it is scaled by dt.  Block b covers part
// b % nbpr of run b / nbpr (one run: nbpr blocks over all cells).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_dt_partial(Dev d, const double *__restrict__ Uin, int nbpr)
{
    __shared__ double s_cfl[LAG2D_BLOCK], s_vol[LAG2D_BLOCK];
    const int nc = d.nc, nd = d.nd;
    const double eps = 1.0e-6, g1 = d.gamlg - 1.0;
    const int r = blockIdx.x / nbpr, part = blockIdx.x - nbpr * r;
    const int c0 = d.nrun > 1 ? d.coff[r] : 0, c1 = d.nrun > 1 ? d.coff[r + 1] : nc;
    double mcfl = 1e10, mvol = 1e10;
    for (int ie = c0 + part * blockDim.x + threadIdx.x; ie < c1; ie += nbpr * blockDim.x) {
        if (ie < d.ntria) continue;  // the reference's dt loops over quads only (Getrhslag.f90:20624-20698)
        double xp[9], yp[9];
        const int nv = d.ncurv == 0 ? 4 : 9;
        for (int k = 0; k < nv; ++k) {
            int p = d.ipq[k * nc + ie];
            xp[k] = d.xn[p];
            yp[k] = d.yn[p];
        }
        double dl[4];
#define LEN(a, b) sqrt((xp[a] - xp[b]) * (xp[a] - xp[b]) + (yp[a] - yp[b]) * (yp[a] - yp[b]))
        if (d.ncurv == 0) {
            dl[0] = LEN(0, 1); dl[1] = LEN(1, 2); dl[2] = LEN(2, 3); dl[3] = LEN(3, 0);
        } else {
            dl[0] = LEN(0, 4) + LEN(4, 1);
            dl[1] = LEN(1, 5) + LEN(5, 2);
            dl[2] = LEN(2, 6) + LEN(6, 3);
            dl[3] = LEN(3, 7) + LEN(7, 0);
        }
#undef LEN
        double tau = Uin[(0 * nd) * nc + ie], u = Uin[(1 * nd) * nc + ie];
        double v = Uin[(2 * nd) * nc + ie], e = Uin[(3 * nd) * nc + ie];
        double rhoct = 1.0 / tau;
        double pctr = dmax(eps, g1 * rhoct * (e - 0.5 * (u * u + v * v)));
        double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
        double dlen = dmin(dmin(dmin(dl[0], dl[1]), dl[2]), dl[3]);
        mcfl = dmin(mcfl, dlen / sdctr);
        mvol = dmin(mvol, d.G[2 * nc + ie] / fabs(d.R[ie]));
    }
    s_cfl[threadIdx.x] = mcfl;
    s_vol[threadIdx.x] = mvol;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_cfl[threadIdx.x] = dmin(s_cfl[threadIdx.x], s_cfl[threadIdx.x + w]);
            s_vol[threadIdx.x] = dmin(s_vol[threadIdx.x], s_vol[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        d.red[blockIdx.x] = s_cfl[0];
        d.red[gridDim.x + blockIdx.x] = s_vol[0];
    }
}

// One block per run: the run's dt (Getrhslag.f90:20705-20709, clamped to its
// end time, :1768-1770) and time.  Minima do not depend on the order.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_dt_final(Dev d, int itime, int nbpr)
{
    __shared__ double s_cfl[LAG2D_BLOCK], s_vol[LAG2D_BLOCK];
    const int r = blockIdx.x, nall = gridDim.x * nbpr;
    double mcfl = 1e10, mvol = 1e10;
    for (int b = threadIdx.x; b < nbpr; b += blockDim.x) {
        mcfl = dmin(mcfl, d.red[r * nbpr + b]);
        mvol = dmin(mvol, d.red[nall + r * nbpr + b]);
    }
    s_cfl[threadIdx.x] = mcfl;
    s_vol[threadIdx.x] = mvol;
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s_cfl[threadIdx.x] = dmin(s_cfl[threadIdx.x], s_cfl[threadIdx.x + w]);
            s_vol[threadIdx.x] = dmin(s_vol[threadIdx.x], s_vol[threadIdx.x + w]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && run_active(d, r)) {
        const double tend = d.nrun > 1 ? d.tendv[r] : d.tend;
        double dt = d.scal[2 * r], t = d.scal[2 * r + 1];
        if (d.ndt == 1) {
            if (itime == 1)
                dt = dmin(d.cfl * s_cfl[0], 0.1 * s_vol[0]);
            else
                dt = dmin(dmin(d.cfl * s_cfl[0], 0.1 * s_vol[0]), 1.01 * dt);
            if (t + dt > tend) dt = tend - t;
        }
        d.scal[2 * r] = dt;
        d.scal[2 * r + 1] = t + dt;
    }
}

// Start of a step of an ensemble: a run advances unless it has reached its
// end time (the driver's exit test, dgflo2d.f90:458).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_run_flags(Dev d)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.nrun) return;
    const double t = d.scal[2 * r + 1], tend = d.tendv[r];
    d.act[r] = (fabs(t - tend) / tend <= 1e-11) ? 0 : 1;
}

// Quotients by the shared-divisor division the kernels use (qdiv), for the
// library's self-test against the IEEE division.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_division_check(const double *a, const double *b, double *q, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    q[i] = qdiv(a[i], recip(b[i]));
}

// Mass-matrix solve and SSP-RK blend (Getrhslag.f90:1822-1904):
// U_s = alfa U^n + (1 - alfa)(U_s + M^-1 dt R), one thread per DOF (q, id)
// of a cell, cells fastest so that every access is coalesced.  For meshes
// that stay in the caches; k_update_cells below reads each cell's M^-1
// and residual from HBM once, for meshes that do not.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_update(Dev d, double alfa)
{
    const int nc = d.nc, nd = d.nd;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 4 * nd * nc) return;
    const int ie = t % nc, qi = t / nc, q = qi / nd, id = qi - nd * q;
    const int run = cell_run(d, ie);
    const int o = (q * nd + id) * nc + ie;
    if (!run_active(d, run)) {
        d.Us[o] = d.Un[o];
        return;
    }
    const double dt = d.scal[2 * run];
    // row id of the inverse mass matrix: degree 0 decoupled (1/mass), the
    // rest packed as the upper triangle row-wise
    double mrow[6];
    if (nd == 3) {
        if (id == 0) {
            mrow[0] = d.A[3 * nc + ie]; mrow[1] = 0.0; mrow[2] = 0.0;
        } else if (id == 1) {
            mrow[0] = 0.0; mrow[1] = d.A[0 * nc + ie]; mrow[2] = d.A[1 * nc + ie];
        } else {
            mrow[0] = 0.0; mrow[1] = d.A[1 * nc + ie]; mrow[2] = d.A[2 * nc + ie];
        }
    } else {
        const int off[6] = {0, 0, 5, 9, 12, 14};
        for (int k = 0; k < 6; ++k) {
            if (id == 0 || k == 0) {
                mrow[k] = (id == 0 && k == 0) ? d.A[15 * nc + ie] : 0.0;
            } else {
                const int a = id < k ? id : k, b = id < k ? k : id;
                mrow[k] = d.A[(off[a] + (b - a)) * nc + ie];
            }
        }
    }
    double s4 = 0.0;
    for (int k = 0; k < nd; ++k) s4 = s4 + mrow[k] * (dt * d.R[(q * nd + k) * nc + ie]);
    d.Us[o] = alfa * d.Un[o] + (1.0 - alfa) * (d.Us[o] + s4);
}


// The same update, one thread per (cell,
// variable), the four variables of a cell in adjacent lanes: a cell's M^-1
// and residual are read from HBM once for all its DOFs (the lanes of a
// cell share the lines), and each lane reads its variable's rows for 16
// consecutive cells of a wave.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_update_cells(Dev d, double alfa)
{
    const int nc = d.nc, nd = d.nd;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int ie = t >> 2, q = t & 3;
    if (ie >= nc) return;
    const int run = cell_run(d, ie);
    if (!run_active(d, run)) {
        for (int id = 0; id < nd; ++id) d.Us[(q * nd + id) * nc + ie] = d.Un[(q * nd + id) * nc + ie];
        return;
    }
    const double dt = d.scal[2 * run];
    // the inverse mass matrix: degree 0 decoupled (1/mass), the rest packed
    // as the upper triangle row-wise
    const int nm = nd == 3 ? 4 : 16;
    double a[16];
    for (int k = 0; k < nm; ++k) a[k] = d.A[k * nc + ie];
    {
        double r[6];
        for (int k = 0; k < nd; ++k) r[k] = d.R[(q * nd + k) * nc + ie];
        for (int id = 0; id < nd; ++id) {
            double mrow[6];
            if (nd == 3) {
                if (id == 0) {
                    mrow[0] = a[3]; mrow[1] = 0.0; mrow[2] = 0.0;
                } else if (id == 1) {
                    mrow[0] = 0.0; mrow[1] = a[0]; mrow[2] = a[1];
                } else {
                    mrow[0] = 0.0; mrow[1] = a[1]; mrow[2] = a[2];
                }
            } else {
                const int off[6] = {0, 0, 5, 9, 12, 14};
                for (int k = 0; k < 6; ++k) {
                    if (id == 0 || k == 0) {
                        mrow[k] = (id == 0 && k == 0) ? a[15] : 0.0;
                    } else {
                        const int lo = id < k ? id : k, hi = id < k ? k : id;
                        mrow[k] = a[off[lo] + (hi - lo)];
                    }
                }
            }
            double s4 = 0.0;
            for (int k = 0; k < nd; ++k) s4 = s4 + mrow[k] * (dt * r[k]);
            const int o = (q * nd + id) * nc + ie;
            d.Us[o] = alfa * d.Un[o] + (1.0 - alfa) * (d.Us[o] + s4);
        }
    }
}

// x_s = alfa x^n + (1 - alfa)(x_src + dt u*)  (Getrhslag.f90:1954-1956)
__global__ void __launch_bounds__(LAG2D_BLOCK) k_nodemove(Dev d, double alfa, const double *__restrict__ xsrc,
                                                          const double *__restrict__ ysrc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int run = node_run(d, p);
    if (!run_active(d, run)) {
        d.xs[p] = d.xn[p];
        d.ys[p] = d.yn[p];
        return;
    }
    const double dt = d.scal[2 * run];
    double x = alfa * d.xn[p] + (1.0 - alfa) * (xsrc[p] + dt * d.usx[p]);
    double y = alfa * d.yn[p] + (1.0 - alfa) * (ysrc[p] + dt * d.usy[p]);
    d.xs[p] = x;
    d.ys[p] = y;
}

__device__ __forceinline__ void shape_q2_d(double r, double s, double *sh, double *dr_, double *ds_)
{
    const double c10 = 1.0;
    sh[0] = 0.25 * (c10 - r) * r * s * (c10 - s);
    sh[1] = -0.25 * (c10 + r) * (c10 - s) * r * s;
    sh[2] = 0.25 * (c10 + r) * (c10 + s) * r * s;
    sh[3] = -0.25 * (c10 - r) * (c10 + s) * r * s;
    sh[4] = -0.5 * (c10 - r * r) * (c10 - s) * s;
    sh[5] = 0.5 * (c10 + r) * (c10 - s * s) * r;
    sh[6] = 0.5 * (c10 - r * r) * (c10 + s) * s;
    sh[7] = -0.5 * (c10 - r) * (c10 - s * s) * r;
    sh[8] = (c10 - r * r) * (c10 - s * s);
    dr_[0] = 0.25 * (2.0 * r - c10) * s * (s - 1.0);
    dr_[1] = 0.25 * (2.0 * r + c10) * s * (s - 1.0);
    dr_[2] = 0.25 * (2.0 * r + c10) * s * (s + 1.0);
    dr_[3] = 0.25 * (2.0 * r - c10) * s * (s + 1.0);
    dr_[4] = -r * s * (s - c10);
    dr_[5] = 0.5 * (2.0 * r + c10) * (c10 - s * s);
    dr_[6] = -r * s * (s + c10);
    dr_[7] = 0.5 * (2.0 * r - c10) * (1.0 - s * s);
    dr_[8] = -2.0 * r * (1.0 - s * s);
    ds_[0] = 0.25 * (2.0 * s - c10) * r * (r - 1.0);
    ds_[1] = 0.25 * (2.0 * s - c10) * r * (r + 1.0);
    ds_[2] = 0.25 * (2.0 * s + c10) * r * (r + 1.0);
    ds_[3] = 0.25 * (2.0 * s + c10) * r * (r - 1.0);
    ds_[4] = 0.5 * (1.0 - r * r) * (2.0 * s - c10);
    ds_[5] = -r * s * (r + c10);
    ds_[6] = 0.5 * (1.0 - r * r) * (2.0 * s + c10);
    ds_[7] = -r * s * (r - c10);
    ds_[8] = -2.0 * s * (1.0 - r * r);
}

// 6x6 Gauss rule of ruqope (dgflo2d.f90:2752-2763): permuted points and
// weights truncated to 16 digits, as the reference lists them.
static constexpr double PG6[6] = {0.2386191860831969086305017, -0.2386191860831969086305017,
                                           0.6612093864662645136613996, -0.6612093864662645136613996,
                                           -0.9324695142031520278123016, 0.9324695142031520278123016};
static constexpr double WG6[6] = {0.4679139345726910, 0.4679139345726910, 0.3607615730481386,
                                           0.3607615730481386, 0.1713244923791704, 0.1713244923791704};

// Cell volume and volume centre on the 6x6 Gauss rule.  With set_centre
// (curved cells) the Q2 centre node is first reset to -1/4 sum(corners) +
// 1/2 sum(mid-nodes) (Getrhslag.f90:1962-1971); it belongs to this cell
// only, so the reset and the volume fuse.  Two work splits with the same
// arithmetic: k_cellvol, 8 lanes per cell (the lanes evaluate the
// Jacobians, lane 0 sums them in point order), for meshes too small to fill
// the chip with a thread per cell; k_cellvol_tpc, a thread per cell whose
// unrolled point loop turns the shape functions into constants.
// The area-weighted rows of a cell from its radius-weighted volume, area
// and area moments.
__device__ __forceinline__ void rz_rows(const Dev &d, int ie, double vol, double area, double rca, double sca)
{
    const int nc = d.nc;
    d.G[4 * nc + ie] = rca / area;
    d.G[5 * nc + ie] = sca / area;
    d.G[10 * nc + ie] = vol / area;
    d.G[11 * nc + ie] = area;
    d.G[14 * nc + ie] = rca / area;
    d.G[15 * nc + ie] = sca / area;
}

// rutope 7-point area rule of the reference triangle (dgflo2d.f90:2234-2255)
static constexpr double PT7R_C[7] = {0.333333333333333, 0.059715871789770, 0.470142064105115,
                                              0.470142064105115, 0.797426985353087, 0.101286507323456,
                                              0.101286507323456};
static constexpr double PT7S_C[7] = {0.333333333333333, 0.470142064105115, 0.059715871789770,
                                              0.470142064105115, 0.101286507323456, 0.797426985353087,
                                              0.101286507323456};
static constexpr double WT7_C[7] = {0.225000000000000, 0.132394152788506, 0.132394152788506,
                                             0.132394152788506, 0.125939180544827, 0.125939180544827,
                                             0.125939180544827};

// With RZ (nrz=2) the volume is the radius-weighted one (rcoef = 1 - alfrz
// + alfrz*y at each point) and the area-weighted rows follow
// (Limiterlag.f90:2001-2035, :2140-2160): geoel(11) mean radius = volume /
// area, geoel(12) area, and the area centre in geoel(5:6) and geoel(15:16).
// The 9 Q2 nodes of quad ie: straight cells (ncurv=0) take the mid and
// centre nodes from the corners; curved ones read them, or with set_centre
// reset the centre (written back when write_centre).
__device__ __forceinline__ void cellvol_nodes(const Dev &d, double *__restrict__ x, double *__restrict__ y, int ie,
                                              int set_centre, bool write_centre, double *xp, double *yp)
{
    const int nc = d.nc;
    if (d.ncurv == 0) {
        for (int k = 0; k < 4; ++k) {
            int p = d.ipq[k * nc + ie];
            xp[k] = x[p];
            yp[k] = y[p];
        }
        xp[4] = 0.5 * (xp[0] + xp[1]); yp[4] = 0.5 * (yp[0] + yp[1]);
        xp[5] = 0.5 * (xp[1] + xp[2]); yp[5] = 0.5 * (yp[1] + yp[2]);
        xp[6] = 0.5 * (xp[2] + xp[3]); yp[6] = 0.5 * (yp[2] + yp[3]);
        xp[7] = 0.5 * (xp[3] + xp[0]); yp[7] = 0.5 * (yp[3] + yp[0]);
        xp[8] = 0.5 * (xp[4] + xp[6]); yp[8] = 0.5 * (yp[4] + yp[6]);
        return;
    }
    for (int k = 0; k < 8; ++k) {
        int p = d.ipq[k * nc + ie];
        xp[k] = x[p];
        yp[k] = y[p];
    }
    const int pc = d.ipq[8 * nc + ie];
    if (set_centre) {
        const double cx = xp[0] + xp[1] + xp[2] + xp[3], mx = xp[4] + xp[5] + xp[6] + xp[7];
        const double cy = yp[0] + yp[1] + yp[2] + yp[3], my = yp[4] + yp[5] + yp[6] + yp[7];
        xp[8] = -0.25 * cx + 0.5 * mx;
        yp[8] = -0.25 * cy + 0.5 * my;
        if (write_centre) {
            x[pc] = xp[8];
            y[pc] = yp[8];
        }
    } else {
        xp[8] = x[pc];
        yp[8] = y[pc];
    }
}

// Weighted Jacobian determinant of the Q2 map at 6x6 point g (and the
// mapped y there, for RZ).
__device__ __forceinline__ double cellvol_point(const double *xp, const double *yp, int g, double *yg)
{
    const int i = g / 6, j = g - 6 * i;
    const double r = PG6[i], s = PG6[j], wi = WG6[i] * WG6[j];
    double sh[9], dr_[9], ds_[9];
    shape_q2_d(r, s, sh, dr_, ds_);
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 9; ++k) {
        dxdr = dxdr + dr_[k] * xp[k];
        dxds = dxds + ds_[k] * xp[k];
        dydr = dydr + dr_[k] * yp[k];
        dyds = dyds + ds_[k] * yp[k];
    }
    if (yg) {
        double v = 0.0;
        for (int k = 0; k < 9; ++k) v = v + sh[k] * yp[k];
        *yg = v;
    }
    return wi * (dxdr * dyds - dydr * dxds);
}

// triangle: 7-point rule on the quadratic triangle map with chord
// midpoints as mid-nodes (Limiterlag.f90:1898-2036)
template <int RZ>
__device__ __forceinline__ void cellvol_tria(const Dev &d, const double *__restrict__ x,
                                             const double *__restrict__ y, int ie)
{
    const int nc = d.nc;
    double xp[6], yp[6];
    // straight triangles: chord-midpoint mid-nodes; curved (ncurv=1): the
    // mesh's six nodes (Limiterlag.f90:1904-1914)
    const int nvt = d.ncurv == 1 ? 6 : 3;
    for (int k = 0; k < nvt; ++k) {
        const int p = d.ipq[k * nc + ie];
        xp[k] = x[p];
        yp[k] = y[p];
    }
    if (nvt == 3) {
        xp[3] = 0.5 * (xp[0] + xp[1]); yp[3] = 0.5 * (yp[0] + yp[1]);
        xp[4] = 0.5 * (xp[1] + xp[2]); yp[4] = 0.5 * (yp[1] + yp[2]);
        xp[5] = 0.5 * (xp[0] + xp[2]); yp[5] = 0.5 * (yp[0] + yp[2]);
    }
    const double c10 = 1.0;
    double rcv = 0.0, scv = 0.0, vol = 0.0, area = 0.0, rca = 0.0, sca = 0.0;
    for (int g = 0; g < 7; ++g) {
        const double r = PT7R_C[g], s = PT7S_C[g], wi = WT7_C[g];
        const double dr_[6] = {c10 - 4.0 * (c10 - r - s), -1.0 + 4.0 * r, 0.0, 4.0 * (1.0 - 2.0 * r - s),
                               4.0 * s, -4.0 * s};
        const double ds_[6] = {c10 - 4.0 * (c10 - r - s), 0.0, -1.0 + 4.0 * s, -4.0 * r, 4.0 * r,
                               4.0 * (1.0 - r - 2.0 * s)};
        double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
        for (int k = 0; k < 6; ++k) {
            dxdr = dxdr + dr_[k] * xp[k];
            dxds = dxds + ds_[k] * xp[k];
            dydr = dydr + dr_[k] * yp[k];
            dyds = dyds + ds_[k] * yp[k];
        }
        const double djaco = 0.5 * wi * (dxdr * dyds - dydr * dxds);
        if (RZ) {
            const double sh[6] = {-(c10 - r - s) * (c10 - 2.0 * (c10 - r - s)), -r * (c10 - 2.0 * r),
                                  -s * (c10 - 2.0 * s), 4.0 * r * (c10 - r - s), 4.0 * r * s,
                                  4.0 * s * (c10 - r - s)};
            double yg = 0.0;
            for (int k = 0; k < 6; ++k) yg = yg + sh[k] * yp[k];
            const double rcoef = 1.0 - d.alfrz + d.alfrz * yg;
            vol = vol + djaco * rcoef;
            area = area + djaco;
            rca = rca + r * djaco;
            sca = sca + s * djaco;
        } else {
            vol = vol + djaco * 1.0;
            rcv = rcv + r * djaco * 1.0;
            scv = scv + s * djaco * 1.0;
        }
    }
    d.G[2 * nc + ie] = vol;
    if (RZ) {
        rz_rows(d, ie, vol, area, rca, sca);
    } else {
        d.G[4 * nc + ie] = rcv / vol;
        d.G[5 * nc + ie] = scv / vol;
    }
}

// Sums over the 36 points in point order (planar: rcoef = 1 - alfrz +
// alfrz*y = 1) and the rows they give.
template <int RZ>
struct CellvolSum {
    double vol = 0.0, rcv = 0.0, scv = 0.0, area = 0.0, rca = 0.0, sca = 0.0;
    __device__ __forceinline__ void add(const Dev &d, int g, double djaco, double yg)
    {
        const int i = g / 6, j = g - 6 * i;
        const double r = PG6[i], s = PG6[j];
        if (RZ) {
            const double rcoef = 1.0 - d.alfrz + d.alfrz * yg;
            vol = vol + djaco * rcoef;
            area = area + djaco;
            rca = rca + r * djaco;
            sca = sca + s * djaco;
        } else {
            vol = vol + djaco * 1.0;
            rcv = rcv + r * djaco * 1.0;
            scv = scv + s * djaco * 1.0;
        }
    }
    __device__ __forceinline__ void store(const Dev &d, int ie) const
    {
        const int nc = d.nc;
        d.G[2 * nc + ie] = vol;
        if (RZ) {
            rz_rows(d, ie, vol, area, rca, sca);
        } else {
            d.G[4 * nc + ie] = rcv / vol;
            d.G[5 * nc + ie] = scv / vol;
        }
    }
};

template <int RZ>
__global__ void __launch_bounds__(256) k_cellvol(Dev d, double *__restrict__ x, double *__restrict__ y,
                                                 int set_centre)
{
    __shared__ double sdj[CV_CELLS][36 + 1];
    __shared__ double syg[RZ ? CV_CELLS : 1][36 + 1];
    const int cl = threadIdx.x >> 3, lane = threadIdx.x & 7;
    const int ie = blockIdx.x * CV_CELLS + cl;
    const int nc = d.nc;
    if (ie < nc && ie >= d.ntria) {
        double xp[9], yp[9];
        cellvol_nodes(d, x, y, ie, set_centre, lane == 0, xp, yp);
        for (int g = lane; g < 36; g += 8) {
            double yg;
            sdj[cl][g] = cellvol_point(xp, yp, g, RZ ? &yg : nullptr);
            if (RZ) syg[RZ ? cl : 0][g] = yg;
        }
    }
    __syncthreads();
    if (ie >= nc || lane != 0) return;
    if (ie < d.ntria) {
        cellvol_tria<RZ>(d, x, y, ie);
        return;
    }
    CellvolSum<RZ> acc;
#pragma unroll
    for (int g = 0; g < 36; ++g) acc.add(d, g, sdj[cl][g], RZ ? syg[RZ ? cl : 0][g] : 0.0);
    acc.store(d, ie);
}

template <int RZ>
__global__ void __launch_bounds__(256) k_cellvol_tpc(Dev d, double *__restrict__ x, double *__restrict__ y,
                                                     int set_centre)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= d.nc) return;
    if (ie < d.ntria) {
        cellvol_tria<RZ>(d, x, y, ie);
        return;
    }
    double xp[9], yp[9];
    cellvol_nodes(d, x, y, ie, set_centre, true, xp, yp);
    CellvolSum<RZ> acc;
#pragma unroll
    for (int g = 0; g < 36; ++g) {
        double yg = 0.0;
        const double djaco = cellvol_point(xp, yp, g, RZ ? &yg : nullptr);
        acc.add(d, g, djaco, yg);
        // the nodes pass through an empty asm statement after each group of
        // points, so the compiler keeps no products of them from one group to
        // the next, which bounds the registers the unrolled loop holds
        if (g % CV_GROUP == CV_GROUP - 1)
            for (int k = 0; k < 9; ++k) asm volatile("" : "+v"(xp[k]), "+v"(yp[k]));
    }
    acc.store(d, ie);
}

template __global__ void k_cellvol<0>(Dev, double *__restrict__, double *__restrict__, int);
template __global__ void k_cellvol<1>(Dev, double *__restrict__, double *__restrict__, int);
template __global__ void k_cellvol_tpc<0>(Dev, double *__restrict__, double *__restrict__, int);
template __global__ void k_cellvol_tpc<1>(Dev, double *__restrict__, double *__restrict__, int);

// 5x5 Gauss rule of ruqope (dgflo2d.f90:2734-2743 ordering)
static constexpr double PG5I[5] = {-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                                            0.5384693101056830910363144, 0.9061798459386639927976269};
static constexpr double WG5I[5] = {0.2369268850561890875142640, 0.4786286704993664680412915,
                                            0.5688888888888888888888889, 0.4786286704993664680412915,
                                            0.2369268850561890875142640};

// Weighted Jacobian determinant of the Q2 map at (r, s).
__device__ __forceinline__ double djaco_q2(const double *xp, const double *yp, double r, double s, double wi)
{
    double sh[9], dr_[9], ds_[9];
    shape_q2_d(r, s, sh, dr_, ds_);
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 9; ++k) {
        dxdr = dxdr + dr_[k] * xp[k];
        dxds = dxds + ds_[k] * xp[k];
        dydr = dydr + dr_[k] * yp[k];
        dyds = dyds + ds_[k] * yp[k];
    }
    return wi * (dxdr * dyds - dydr * dxds);
}

// Driver-side initialisation of the curved P2 path, one thread per cell,
// with the reference's loops and summation orders (planar: the radius
// coefficient 1 - alfrz + alfrz y is exactly 1 and its products drop):
//   getgeoel_initial_lag (GetInitial.f90:1507-1693): mass, volume, mass and
//     volume centres on the 6x6 rule;
//   getgeoelho_initial_lag (:1917-2090): the second moments about both
//     centres;
//   getunkno_initial_lag (:478-534): the mass matrix of the degree >= 1
//     block on the 5x5 rule, every entry accumulated as written there, and
//     its inverse by getinvmat (dgflo2d.f90:2792-2908: ludeco without
//     pivoting, then lureso per unit column), stored as amatr(1:15) = the
//     inverse's upper triangle row by row, amatr(16) = 1 / mass.
// The state is piecewise constant per cell: in = [rho; u; v; p; blast
// energy; blast flag] x nc.  The DOFs are tau = (5x5 volume) / mass, u, v,
// E = p / ((gamma - 1) rho) + |u|^2 / 2 (E = energy / mass in a blast
// cell), higher DOFs 0, as the host initialisation sets them.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_init_q2(Dev d, const double *__restrict__ in, double *A)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    const double rho = in[ie];
    double xp[9], yp[9];
    for (int k = 0; k < 9; ++k) {
        const int p = d.ipq[k * nc + ie];
        xp[k] = d.xn[p];
        yp[k] = d.yn[p];
    }
    // ---- getgeoel_initial_lag
    double masel = 0.0, volel = 0.0, xcent = 0.0, ycent = 0.0, xcel = 0.0, ycel = 0.0;
    for (int g = 0; g < 36; ++g) {
        const int i = g / 6, j = g - 6 * i;
        const double r = PG6[i], s = PG6[j];
        const double djaco = djaco_q2(xp, yp, r, s, WG6[i] * WG6[j]);
        masel = masel + rho * djaco;
        volel = volel + djaco;
        xcent = xcent + rho * r * djaco;
        ycent = ycent + rho * s * djaco;
        xcel = xcel + djaco * r;
        ycel = ycel + djaco * s;
    }
    const double rc = xcent / masel, sc = ycent / masel;
    const double rci = xcel / volel, sci = ycel / volel;
    // ---- getgeoelho_initial_lag
    double bq4 = 0.0, bq5 = 0.0, bq6 = 0.0, bi4 = 0.0, bi5 = 0.0, bi6 = 0.0;
    for (int g = 0; g < 36; ++g) {
        const int i = g / 6, j = g - 6 * i;
        const double r = PG6[i], s = PG6[j];
        const double djaco = djaco_q2(xp, yp, r, s, WG6[i] * WG6[j]);
        const double b2 = (r - rc) / 1.0, b3 = (s - sc) / 1.0;
        bq4 = bq4 + 0.5 * rho * b2 * b2 * djaco;
        bq5 = bq5 + 0.5 * rho * b3 * b3 * djaco;
        bq6 = bq6 + rho * b2 * b3 * djaco;
        const double c2 = (r - rci) / 1.0, c3 = (s - sci) / 1.0;
        bi4 = bi4 + 0.5 * c2 * c2 * djaco;
        bi5 = bi5 + 0.5 * c3 * c3 * djaco;
        bi6 = bi6 + c2 * c3 * djaco;
    }
    const double g18 = bq4 / masel, g19 = bq5 / masel, g20 = bq6 / masel;
    // ---- getunkno_initial_lag: mass matrix a(i, j) at a[i + 5 j]
    double a[25], f0 = 0.0, vol25 = 0.0;
    for (int k = 0; k < 25; ++k) a[k] = 0.0;
    for (int g = 0; g < 25; ++g) {
        const int i = g / 5, j = g - 5 * i;
        const double r = PG5I[i], s = PG5I[j];
        const double djaco = djaco_q2(xp, yp, r, s, WG5I[i] * WG5I[j]);
        double b[5];
        b[0] = (r - rc) / 1.0;
        b[1] = (s - sc) / 1.0;
        b[2] = 0.5 * b[0] * b[0] - g18;
        b[3] = 0.5 * b[1] * b[1] - g19;
        b[4] = b[0] * b[1] - g20;
        f0 = f0 + rho * djaco;
        for (int mx = 0; mx < 5; ++mx)
            for (int my = 0; my < 5; ++my) a[mx + 5 * my] = a[mx + 5 * my] + rho * b[mx] * b[my] * djaco;
        vol25 = vol25 + djaco;
    }
    // ludeco (dgflo2d.f90:2826-2864)
    for (int k = 0; k < 5; ++k) {
        for (int m = 0; m < k; ++m) a[k + 5 * k] = a[k + 5 * k] - a[k + 5 * m] * a[m + 5 * k];
        const double adicv = 1.0 / a[k + 5 * k];
        for (int i = k + 1; i < 5; ++i) {
            for (int m = 0; m < k; ++m) {
                a[k + 5 * i] = a[k + 5 * i] - a[k + 5 * m] * a[m + 5 * i];
                a[i + 5 * k] = a[i + 5 * k] - a[i + 5 * m] * a[m + 5 * k];
            }
            a[i + 5 * k] = a[i + 5 * k] * adicv;
        }
    }
    // lureso for each unit column (dgflo2d.f90:2872-2908); x5(row, col)
    double x5[25];
    for (int col = 0; col < 5; ++col) {
        double *x = &x5[5 * col];
        for (int i = 0; i < 5; ++i) {
            double c = 0.0;
            for (int j = 0; j < i; ++j) c = c + a[i + 5 * j] * x[j];
            x[i] = (i == col ? 1.0 : 0.0) - c;
        }
        for (int i = 4; i >= 0; --i) {
            double c = 0.0;
            for (int j = i + 1; j < 5; ++j) c = c + a[i + 5 * j] * x[j];
            x[i] = (x[i] - c) / a[i + 5 * i];
        }
    }
    int k = 0;
    for (int r = 0; r < 5; ++r)
        for (int c = r; c < 5; ++c) A[(k++) * nc + ie] = x5[r + 5 * c];
    A[15 * nc + ie] = 1.0 / f0;
    // ---- geometry rows and DOFs
    for (int row = 0; row < 24; ++row) d.G[row * nc + ie] = 0.0;
    d.G[0 * nc + ie] = rc;
    d.G[1 * nc + ie] = sc;
    d.G[2 * nc + ie] = volel;
    d.G[3 * nc + ie] = masel;
    d.G[4 * nc + ie] = rci;
    d.G[5 * nc + ie] = sci;
    d.G[6 * nc + ie] = rci;
    d.G[7 * nc + ie] = sci;
    d.G[18 * nc + ie] = g18;
    d.G[19 * nc + ie] = g19;
    d.G[20 * nc + ie] = g20;
    d.G[21 * nc + ie] = bi4 / volel;
    d.G[22 * nc + ie] = bi5 / volel;
    d.G[23 * nc + ie] = bi6 / volel;
    const double u0 = in[nc + ie], v0 = in[2 * nc + ie], p0 = in[3 * nc + ie];
    const double g1 = d.gamlg - 1.0;
    double e0 = p0 / (g1 * rho) + 0.5 * (u0 * u0 + v0 * v0);
    if (in[5 * nc + ie] != 0.0) e0 = in[4 * nc + ie] / masel;
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) d.Un[(q * 6 + id) * nc + ie] = 0.0;
    d.Un[(0 * 6) * nc + ie] = vol25 / masel;
    d.Un[(1 * 6) * nc + ie] = u0;
    d.Un[(2 * 6) * nc + ie] = v0;
    d.Un[(3 * 6) * nc + ie] = e0;
}

// Point g of a straight P1 cell's rule, on the six-node triangle map (edge
// midpoints as the extra nodes, the 7-point rule, djaco = wi/2 |J|) or the
// nine-node quad map (the 6x6 rule for GEO, else 5x5; djaco = wi |J|):
// reference coordinates, weighted Jacobian and the point's y.
template <bool GEO>
__device__ __forceinline__ void p1_point(bool tri, const double *xp, const double *yp, int g, double &r, double &s,
                                         double &djaco, double &yg)
{
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    yg = 0.0;
    if (tri) {
        const double c10 = 1.0;
        r = PT7R_C[g];
        s = PT7S_C[g];
        const double t = c10 - r - s;
        const double sh[6] = {-t * (c10 - 2.0 * t), -r * (c10 - 2.0 * r), -s * (c10 - 2.0 * s),
                              4.0 * r * t,          4.0 * r * s,          4.0 * s * t};
        const double dr_[6] = {c10 - 4.0 * t, -1.0 + 4.0 * r, 0.0, 4.0 * (1.0 - 2.0 * r - s), 4.0 * s, -4.0 * s};
        const double ds_[6] = {c10 - 4.0 * t, 0.0, -1.0 + 4.0 * s, -4.0 * r, 4.0 * r, 4.0 * (1.0 - r - 2.0 * s)};
        for (int k = 0; k < 6; ++k) {
            dxdr = dxdr + dr_[k] * xp[k];
            dxds = dxds + ds_[k] * xp[k];
            dydr = dydr + dr_[k] * yp[k];
            dyds = dyds + ds_[k] * yp[k];
        }
        djaco = 0.5 * WT7_C[g] * (dxdr * dyds - dydr * dxds);
        for (int k = 0; k < 6; ++k) yg = yg + sh[k] * yp[k];
        return;
    }
    double wi;
    if (GEO) {
        const int i = g / 6, j = g - 6 * i;
        r = PG6[i];
        s = PG6[j];
        wi = WG6[i] * WG6[j];
    } else {
        const int i = g / 5, j = g - 5 * i;
        r = PG5I[i];
        s = PG5I[j];
        wi = WG5I[i] * WG5I[j];
    }
    double sh[9], dr_[9], ds_[9];
    shape_q2_d(r, s, sh, dr_, ds_);
    for (int k = 0; k < 9; ++k) {
        dxdr = dxdr + dr_[k] * xp[k];
        dxds = dxds + ds_[k] * xp[k];
        dydr = dydr + dr_[k] * yp[k];
        dyds = dyds + ds_[k] * yp[k];
    }
    djaco = wi * (dxdr * dyds - dydr * dxds);
    for (int k = 0; k < 9; ++k) yg = yg + sh[k] * yp[k];
}

// Driver-side initialisation of the straight P1 path (triangles first,
// then quads; planar, or area-weighted RZ with RZ = 1), one thread per cell,
// with the reference's loops, products and summation orders:
//   getgeoel_initial_lag (GetInitial.f90:1278-1697): mass, volume, mass and
//     volume centres (rows 1-4, 7-8) and, RZ, the area mass, area centre,
//     mean radius and area (rows 1-2, 4, 7-8, 11-12), on the 7-point
//     triangle and 6x6 quad rules (its row 5, the corners' y extent, is
//     not kept: the step keeps the volume centre there);
//   getunkno_initial_lag (:151-693; it overwrites the amatr that
//     getamatr_initial_lag, :697-1274, computes first with another product
//     order): the slope mass matrix on the 7-point and 5x5 rules (dr = ds =
//     1/2 on triangles, 1 on quads), amatr(1:3) its inverse by the 2x2
//     cofactors, amatr(4) = 1 / f0, RZ all four divided by the mean radius.
// The radius coefficient 1 - alfrz + alfrz y is evaluated as written (it is
// exactly 1 when planar).  The state is piecewise constant per cell: in =
// [rho; u; v; p; blast energy; blast flag] x nc.  The DOFs are tau =
// (5x5 or 7-point volume) / f0, u, v, E = p tau / (gamma - 1) + |u|^2 / 2
// (E = energy / mass in a blast cell), slopes 0.
template <int RZ>
__global__ void __launch_bounds__(LAG2D_BLOCK) k_init_p1(Dev d, const double *__restrict__ in, double *A)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    const bool tri = ie < d.ntria;
    const double rho = in[ie];
    double xp[9], yp[9];
    if (tri) {
        for (int k = 0; k < 3; ++k) {
            const int p = d.ipq[k * nc + ie];
            xp[k] = d.xn[p];
            yp[k] = d.yn[p];
        }
        xp[3] = 0.5 * (xp[0] + xp[1]); yp[3] = 0.5 * (yp[0] + yp[1]);
        xp[4] = 0.5 * (xp[1] + xp[2]); yp[4] = 0.5 * (yp[1] + yp[2]);
        xp[5] = 0.5 * (xp[0] + xp[2]); yp[5] = 0.5 * (yp[0] + yp[2]);
    } else {
        cellvol_nodes(d, d.xn, d.yn, ie, 0, false, xp, yp);
    }
    // ---- getgeoel_initial_lag
    double masel = 0.0, volel = 0.0, xcent = 0.0, ycent = 0.0, xcel = 0.0, ycel = 0.0;
    double maselrz = 0.0, xcerz = 0.0, ycerz = 0.0, xceli = 0.0, yceli = 0.0, areel = 0.0;
    const int ng = tri ? 7 : 36;
    for (int g = 0; g < ng; ++g) {
        double r, s, djaco, yg;
        p1_point<true>(tri, xp, yp, g, r, s, djaco, yg);
        const double rcoef = 1.0 - d.alfrz + d.alfrz * yg;
        masel = masel + rho * djaco * rcoef;
        volel = volel + djaco * rcoef;
        xcent = xcent + rho * r * djaco * rcoef;
        ycent = ycent + rho * s * djaco * rcoef;
        xcel = xcel + djaco * r * rcoef;
        ycel = ycel + djaco * s * rcoef;
        if (RZ) {
            maselrz = maselrz + rho * djaco;
            xcerz = xcerz + rho * r * djaco;
            ycerz = ycerz + rho * s * djaco;
            xceli = xceli + djaco * r;
            yceli = yceli + djaco * s;
            areel = areel + djaco;
        }
    }
    for (int row = 0; row < 24; ++row) d.G[row * nc + ie] = 0.0;
    double rc, sc, mass, radius = 1.0;
    if (RZ) {
        rc = xcerz / maselrz;
        sc = ycerz / maselrz;
        mass = maselrz;
        radius = volel / areel;
        d.G[6 * nc + ie] = xceli / areel;
        d.G[7 * nc + ie] = yceli / areel;
        d.G[10 * nc + ie] = radius;
        d.G[11 * nc + ie] = areel;
    } else {
        rc = xcent / masel;
        sc = ycent / masel;
        mass = masel;
        d.G[6 * nc + ie] = xcel / volel;
        d.G[7 * nc + ie] = ycel / volel;
    }
    d.G[0 * nc + ie] = rc;
    d.G[1 * nc + ie] = sc;
    d.G[2 * nc + ie] = volel;
    d.G[3 * nc + ie] = mass;
    // ---- getunkno_initial_lag: the slope mass matrix and 1 / f0
    const double dr = tri ? 0.5 : 1.0, ds = dr;
    double f0 = 0.0, m11 = 0.0, m12 = 0.0, m22 = 0.0, volm = 0.0;
    const int nm = tri ? 7 : 25;
    for (int g = 0; g < nm; ++g) {
        double r, s, djaco, yg;
        p1_point<false>(tri, xp, yp, g, r, s, djaco, yg);
        const double b2 = (r - rc) / dr, b3 = (s - sc) / ds;
        if (RZ) {
            f0 = f0 + rho * djaco;
            m11 = m11 + rho * b2 * b2 * djaco;
            m12 = m12 + rho * b2 * b3 * djaco;
            m22 = m22 + rho * b3 * b3 * djaco;
            volm = volm + djaco;
        } else {
            const double rcoef = 1.0 - d.alfrz + d.alfrz * yg;
            f0 = f0 + rho * djaco * rcoef;
            m11 = m11 + rho * b2 * b2 * djaco * rcoef;
            m12 = m12 + rho * b2 * b3 * djaco * rcoef;
            m22 = m22 + rho * b3 * b3 * djaco * rcoef;
            volm = volm + djaco * rcoef;
        }
    }
    const double det = m11 * m22 - m12 * m12;
    double a[4] = {m22 / det, -m12 / det, m11 / det, 1.0 / f0};
    for (int k = 0; k < 4; ++k) A[k * nc + ie] = RZ ? a[k] / radius : a[k];
    // ---- DOFs
    const double tau = volm / f0;
    const double u0 = in[nc + ie], v0 = in[2 * nc + ie], p0 = in[3 * nc + ie];
    const double g1 = d.gamlg - 1.0;
    double e0 = p0 * tau / g1 + 0.5 * (u0 * u0 + v0 * v0);
    if (in[5 * nc + ie] != 0.0) e0 = in[4 * nc + ie] / mass;
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) d.Un[(q * 3 + id) * nc + ie] = 0.0;
    d.Un[(0 * 3) * nc + ie] = tau;
    d.Un[(1 * 3) * nc + ie] = u0;
    d.Un[(2 * 3) * nc + ie] = v0;
    d.Un[(3 * 3) * nc + ie] = e0;
}

template __global__ void k_init_p1<0>(Dev, const double *__restrict__, double *);
template __global__ void k_init_p1<1>(Dev, const double *__restrict__, double *);

}  // namespace lag2d
