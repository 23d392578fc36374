// Kernels shared by the straight and curved paths: time-step reduction
// (getdelt_lag, Getrhslag.f90:20600-20713), mass-matrix solve with the
// SSP-RK blend (Getrhslag.f90:1804-1904), node motion and the Q2 centre-node
// reset (:1954-1971), and the cell volume / volume centre (getcellvol_lag,
// Limiterlag.f90:1999-2162).
#include "lag2d_dev.h"
#include "lag2d_kernels.h"

namespace lag2d {

// Per-block minima of the CFL time (edge length / sound speed) and of the
// volume-change time V/|dV/dt| over the quads, with dV/dt the degree-0
// tau right-hand side before it is scaled by dt.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_dt_partial(Dev d, const double *__restrict__ Uin, int nblk)
{
    __shared__ double s_cfl[LAG2D_BLOCK], s_vol[LAG2D_BLOCK];
    const int nc = d.nc, nd = d.nd;
    const double eps = 1.0e-6, g1 = d.gamlg - 1.0;
    double mcfl = 1e10, mvol = 1e10;
    for (int ie = blockIdx.x * blockDim.x + threadIdx.x; ie < nc; ie += gridDim.x * blockDim.x) {
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
        d.red[nblk + blockIdx.x] = s_vol[0];
    }
}

__global__ void __launch_bounds__(LAG2D_BLOCK) k_dt_final(Dev d, int itime, int nblk)
{
    __shared__ double s_cfl[LAG2D_BLOCK], s_vol[LAG2D_BLOCK];
    double mcfl = 1e10, mvol = 1e10;
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        mcfl = dmin(mcfl, d.red[b]);
        mvol = dmin(mvol, d.red[nblk + b]);
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
        double dt = d.scal[0], t = d.scal[1];
        if (d.ndt == 1) {
            if (itime == 1)
                dt = dmin(d.cfl * s_cfl[0], 0.1 * s_vol[0]);
            else
                dt = dmin(dmin(d.cfl * s_cfl[0], 0.1 * s_vol[0]), 1.01 * dt);
            if (t + dt > d.tend) dt = d.tend - t;
        }
        d.scal[0] = dt;
        d.scal[1] = t + dt;
    }
}

// Mass-matrix solve and SSP-RK blend of one cell (Getrhslag.f90:1822-1904):
// U_s = alfa U^n + (1 - alfa)(U_s + M^-1 dt R).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_update(Dev d, double alfa)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc, nd = d.nd;
    if (ie >= nc) return;
    const double dt = d.scal[0];
    double m[6][6];
    if (nd == 3) {
        m[0][0] = d.A[3 * nc + ie]; m[0][1] = 0.0; m[0][2] = 0.0;
        m[1][0] = 0.0; m[1][1] = d.A[0 * nc + ie]; m[1][2] = d.A[1 * nc + ie];
        m[2][0] = 0.0; m[2][1] = d.A[1 * nc + ie]; m[2][2] = d.A[2 * nc + ie];
    } else {
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 6; ++j) m[i][j] = 0.0;
        m[0][0] = d.A[15 * nc + ie];
        int k = 0;
        for (int i = 1; i < 6; ++i)
            for (int j = i; j < 6; ++j) {
                double a = d.A[k * nc + ie];
                m[i][j] = a;
                m[j][i] = a;
                ++k;
            }
    }
    double rs[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < nd; ++id) rs[q][id] = dt * d.R[(q * nd + id) * nc + ie];
    for (int id = 0; id < nd; ++id) {
        double s4[4] = {0.0, 0.0, 0.0, 0.0};
        for (int k = 0; k < nd; ++k)
            for (int q = 0; q < 4; ++q) s4[q] = s4[q] + m[id][k] * rs[q][k];
        for (int q = 0; q < 4; ++q) {
            const int o = (q * nd + id) * nc + ie;
            d.Us[o] = alfa * d.Un[o] + (1.0 - alfa) * (d.Us[o] + s4[q]);
        }
    }
}

// x_s = alfa x^n + (1 - alfa)(x_src + dt u*)  (Getrhslag.f90:1954-1956)
__global__ void __launch_bounds__(LAG2D_BLOCK) k_nodemove(Dev d, double alfa, const double *__restrict__ xsrc,
                                                          const double *__restrict__ ysrc)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const double dt = d.scal[0];
    double x = alfa * d.xn[p] + (1.0 - alfa) * (xsrc[p] + dt * d.usx[p]);
    double y = alfa * d.yn[p] + (1.0 - alfa) * (ysrc[p] + dt * d.usy[p]);
    d.xs[p] = x;
    d.ys[p] = y;
}

// Q2 centre node := -1/4 sum(corners) + 1/2 sum(mid-nodes) (Getrhslag.f90:1962-1971)
__global__ void __launch_bounds__(LAG2D_BLOCK) k_center_q2(Dev d)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    int ip[9];
    for (int k = 0; k < 9; ++k) ip[k] = d.ipq[k * nc + ie];
    double cx = d.xs[ip[0]] + d.xs[ip[1]] + d.xs[ip[2]] + d.xs[ip[3]];
    double mx = d.xs[ip[4]] + d.xs[ip[5]] + d.xs[ip[6]] + d.xs[ip[7]];
    double cy = d.ys[ip[0]] + d.ys[ip[1]] + d.ys[ip[2]] + d.ys[ip[3]];
    double my = d.ys[ip[4]] + d.ys[ip[5]] + d.ys[ip[6]] + d.ys[ip[7]];
    d.xs[ip[8]] = -0.25 * cx + 0.5 * mx;
    d.ys[ip[8]] = -0.25 * cy + 0.5 * my;
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
__constant__ static const double PG6[6] = {0.2386191860831969086305017, -0.2386191860831969086305017,
                                           0.6612093864662645136613996, -0.6612093864662645136613996,
                                           -0.9324695142031520278123016, 0.9324695142031520278123016};
__constant__ static const double WG6[6] = {0.4679139345726910, 0.4679139345726910, 0.3607615730481386,
                                           0.3607615730481386, 0.1713244923791704, 0.1713244923791704};

__global__ void __launch_bounds__(LAG2D_BLOCK) k_cellvol(Dev d, const double *__restrict__ x,
                                                         const double *__restrict__ y)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    double xp[9], yp[9];
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
    } else {
        for (int k = 0; k < 9; ++k) {
            int p = d.ipq[k * nc + ie];
            xp[k] = x[p];
            yp[k] = y[p];
        }
    }
    double rcv = 0.0, scv = 0.0, vol = 0.0;
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < 6; ++j) {
            double r = PG6[i], s = PG6[j], wi = WG6[i] * WG6[j];
            double sh[9], dr_[9], ds_[9];
            shape_q2_d(r, s, sh, dr_, ds_);
            double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
            for (int k = 0; k < 9; ++k) {
                dxdr = dxdr + dr_[k] * xp[k];
                dxds = dxds + ds_[k] * xp[k];
                dydr = dydr + dr_[k] * yp[k];
                dyds = dyds + ds_[k] * yp[k];
            }
            // planar: rcoef = 1 - alfrz + alfrz*y = 1
            double djaco = wi * (dxdr * dyds - dydr * dxds);
            vol = vol + djaco * 1.0;
            rcv = rcv + r * djaco * 1.0;
            scv = scv + s * djaco * 1.0;
        }
    d.G[2 * nc + ie] = vol;
    d.G[4 * nc + ie] = rcv / vol;
    d.G[5 * nc + ie] = scv / vol;
}

}  // namespace lag2d
