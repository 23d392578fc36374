// Curved-quad P2 sub-cell path (ncurv=1, nfint=4, npoly=2): Persson
// troubled-cell indicator, characteristic WENO on troubled cells, sub-cell
// nodal Riemann solver with sub-cell density correction, sub-cell face
// integral and 5x5 Gauss-Lobatto domain integral.
//
// Reference routines restated here (getrhsdg_lag_mcnew branch
// Getrhslagmc.f90:7556-7708):
//   getfnds_lagsmsif_hybrid       Getrhslaggauss.f90:3356-3598
//   getfnds_lagsms_simpsonhybrid  Getrhslaggauss.f90:9660-10017
//   getSIHO_Persson               Limiterlag.f90:12042-12236
//   weno_char_quad_orthgl         WENO.f90:5927-6120 and its helpers
//                                 (WENO.f90:5889-8098, LinearAlgebra.f90:89-)
//   getndvelo_lagsubg             Getrhslaggauss.f90:4348-4614
//   getriem_quadsubg2             Getrhslaggauss.f90:10409-10864
//   getrhosubg_averg / _daverg    Getrhslaggauss.f90:7067-7392, :10949-11109
//   rhsifacedg_lagsubgq2          Getrhslaggauss.f90:8501-8745
//   rhsdomndg_lagmc_curvquad      Getrhslag.f90:4716-5013
// Work split: one thread per cell (prep, corner, cell RHS), per node (nodal
// solve), per troubled cell (WENO, over a compacted list).
#include <vector>

#include "lag2d_dev.h"
#include "lag2d_host.h"
#include "lag2d_kernels.h"

namespace lag2d {

// sub-cell vertex -> local Q2 node (ipqsg)
__constant__ static const int IPQSG[4][4] = {{0, 4, 8, 7}, {4, 1, 5, 8}, {8, 5, 2, 6}, {7, 8, 6, 3}};
// signed face-normal slots (fnqsg): sign = orientation, |v|-1 = gesgq slot
__constant__ static const int FNQSG[4][8] = {{8, 1, 13, -9, -9, 12, 12, 16},
                                             {9, 13, 2, 3, 14, -10, -10, 9},
                                             {-11, 10, 10, 14, 4, 5, 15, -11},
                                             {16, -12, -12, 11, 11, 15, 6, 7}};
__constant__ static const double XV9[9] = {-1.0, 1.0, 1.0, -1.0, 0.0, 1.0, 0.0, -1.0, 0.0};
__constant__ static const double YV9[9] = {-1.0, -1.0, 1.0, 1.0, -1.0, 0.0, 1.0, 0.0, 0.0};
// Riemann corner weights times 1/4 [isg][ivsg][f] (Getrhslaggauss.f90:10575-10599)
__constant__ static double WFR[4][4][2];
// face-integral weights: momentum/energy {0,1} and volume times 1/4 (:8573-8621)
__constant__ static double WFM[4][4][2];
__constant__ static double WFV[4][4][2];
// Leibniz-expansion term tables in the order of Matinv225 (LinearAlgebra.f90:130-):
// 5x5: 120 permutations; 4x4 minors: 24 permutations; sign in the last entry
__constant__ static signed char PERM5[120][6];
__constant__ static signed char PERM4[24][5];
// 5x5 Gauss rule of ruqope (dgflo2d.f90:2734-2743 ordering)
__constant__ static const double PG5[5] = {-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                                           0.5384693101056830910363144, 0.9061798459386639927976269};
__constant__ static const double WG5[5] = {0.2369268850561890875142640, 0.4786286704993664680412915,
                                           0.5688888888888888888888889, 0.4786286704993664680412915,
                                           0.2369268850561890875142640};
// 5-point Gauss-Lobatto (ruqope_lobatto, dgflo2d.f90:2571-2583)
__constant__ static const double PL5[5] = {-1.0, -0.6546536707079771437983, 0.0, 0.6546536707079771437983, 1.0};
__constant__ static const double WL5[5] = {0.1, 0.544444444444444444444, 0.7111111111111111111111,
                                           0.544444444444444444444, 0.1};

__device__ __forceinline__ void shape9(double r, double s, double *sh, double *dr_, double *ds_)
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

__device__ __forceinline__ void jac9(const double *xp, const double *yp, double r, double s, double J[4])
{
    double sh[9], dr_[9], ds_[9];
    shape9(r, s, sh, dr_, ds_);
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 9; ++k) {
        dxdr = dxdr + dr_[k] * xp[k];
        dxds = dxds + ds_[k] * xp[k];
        dydr = dydr + dr_[k] * yp[k];
        dyds = dyds + ds_[k] * yp[k];
    }
    J[0] = dxdr; J[1] = dxds; J[2] = dydr; J[3] = dyds;
}

__device__ __forceinline__ void map9(const double *xp, const double *yp, double r, double s, double *x, double *y)
{
    double sh[9], dr_[9], ds_[9];
    shape9(r, s, sh, dr_, ds_);
    double a = 0.0, b = 0.0;
    for (int k = 0; k < 9; ++k) {
        a = a + sh[k] * xp[k];
        b = b + sh[k] * yp[k];
    }
    *x = a;
    *y = b;
}

__device__ __forceinline__ void basis6(double r, double s, double rc, double sc, double g0, double g1, double g2,
                                       double *b)
{
    b[0] = 1.0;
    b[1] = (r - rc) / 1.0;
    b[2] = (s - sc) / 1.0;
    b[3] = 0.5 * b[1] * b[1] - g0;
    b[4] = 0.5 * b[2] * b[2] - g1;
    b[5] = b[1] * b[2] - g2;
}

__device__ __forceinline__ void load9(const Dev &d, const double *x, const double *y, int ie, double *xp, double *yp)
{
    for (int k = 0; k < 9; ++k) {
        int p = d.ipq[k * d.nc + ie];
        xp[k] = x[p];
        yp[k] = y[p];
    }
}

// Initial density (getrhoig_quadcurv2 / getrhog_initial) for the
// piecewise-constant problems: Sod splits on the initial cell centre.
__device__ __forceinline__ double rho0(int ncase, double xc, double yc)
{
    if (ncase == 6) return (sqrt(xc * xc + (yc - 0.0) * (yc - 0.0)) <= 0.5) ? 1.0 : 0.125;
    return 1.0;
}

#define UD(buf, q, id, ie) (buf)[((q) * 6 + (id)) * nc + (ie)]

// ---------------------------------------------------------------- prep --
// Copies the stage DOFs into the limited buffer, evaluates the Persson
// indicator on tau and appends troubled cells to the WENO list.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2_prep(Dev d, const double *__restrict__ Uin)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    double u[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) {
            u[q][id] = UD(Uin, q, id, ie);
            UD(d.Ul, q, id, ie) = u[q][id];
        }
    double xp[9], yp[9];
    load9(d, d.xs, d.ys, ie, xp, yp);
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    const double g0 = d.G[18 * nc + ie], g1 = d.G[19 * nc + ie], g2 = d.G[20 * nc + ie];
    double v1 = 0.0, v2 = 0.0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            double r = PG5[i], s = PG5[j], wi = WG5[i] * WG5[j];
            double J[4];
            jac9(xp, yp, r, s, J);
            double djaco = wi * (J[0] * J[3] - J[2] * J[1]);
            double b[6];
            basis6(r, s, rc, sc, g0, g1, g2, b);
            double un = 0.0, u1 = 0.0;
            for (int id = 0; id < 6; ++id) un = un + u[0][id] * b[id];
            for (int id = 0; id < 3; ++id) u1 = u1 + u[0][id] * b[id];
            v2 = v2 + (u1 - un) * (u1 - un) * djaco;
            v1 = v1 + un * un * djaco;
        }
    const double thr = -3.0 * log10(4.0) - 4.0;
    const bool troubled = !(log10(v2 / v1) <= thr);
    d.flag[ie] = troubled ? 1 : 0;
    if (troubled) {
        int pos = atomicAdd(d.tcount, 1);
        d.tlist[pos] = ie;
    }
}

// ---------------------------------------------------------------- WENO --
__device__ __forceinline__ void rotf(const double m[4], double &a, double &b)
{
    double x = a * m[0] + b * m[1];
    double y = a * m[2] + b * m[3];
    a = x;
    b = y;
}

// Matinv225, n = 5: inverse = adjugate / det with the Leibniz terms summed in
// the reference's order (tables PERM5 / PERM4). a, b column-major.
__device__ void inv5(const double *a, double *b)
{
    double det = 0.0;
    for (int t = 0; t < 120; ++t) {
        const signed char *p = PERM5[t];
        double v = a[0 + 5 * p[0]];
        for (int k = 1; k < 5; ++k) v = v * a[k + 5 * p[k]];
        if (p[5] < 0) v = -v;
        det = (t == 0) ? v : det + v;
    }
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            int rows[4], cols[4], kk = 0, ll = 0;
            for (int x = 0; x < 5; ++x)
                if (x != i) rows[kk++] = x;
            for (int x = 0; x < 5; ++x)
                if (x != j) cols[ll++] = x;
            double cof = 0.0;
            for (int t = 0; t < 24; ++t) {
                const signed char *p = PERM4[t];
                double v = a[rows[0] + 5 * cols[p[0]]];
                for (int k = 1; k < 4; ++k) v = v * a[rows[k] + 5 * cols[p[k]]];
                if (p[4] < 0) v = -v;
                cof = (t == 0) ? v : cof + v;
            }
            if ((i + j) & 1) cof = -cof;
            b[j + 5 * i] = cof / det;
        }
}

struct CellPts {
    double djacoi[25], rhogi[25], xl[25], yl[25];
};

__device__ void cell_points(const Dev &d, int je, const double m[4], CellPts &P)
{
    const int nc = d.nc;
    double xi[9], yi[9], xl[9], yl[9];
    load9(d, d.x0, d.y0, je, xi, yi);
    load9(d, d.xs, d.ys, je, xl, yl);
    for (int k = 0; k < 9; ++k) rotf(m, xl[k], yl[k]);
    double xcr, ycr;
    map9(xi, yi, d.G[0 * nc + je], d.G[1 * nc + je], &xcr, &ycr);
    const double rho = rho0(d.ncase, xcr, ycr);
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            const int g = 5 * i + j;
            double r = PG5[i], s = PG5[j], wi = WG5[i] * WG5[j];
            double J[4];
            jac9(xi, yi, r, s, J);
            P.djacoi[g] = wi * (J[0] * J[3] - J[2] * J[1]);
            map9(xl, yl, r, s, &P.xl[g], &P.yl[g]);
            P.rhogi[g] = rho;
        }
}

// Physical mass centre of a cell (weno_char_quad_orthgl, WENO.f90:5983-6064)
__device__ void mass_centre(const Dev &d, int ie, double &xc, double &yc)
{
    const int nc = d.nc;
    double xp[9], yp[9], xi[9], yi[9];
    load9(d, d.xs, d.ys, ie, xp, yp);
    load9(d, d.x0, d.y0, ie, xi, yi);
    double xcr, ycr;
    map9(xi, yi, d.G[0 * nc + ie], d.G[1 * nc + ie], &xcr, &ycr);
    const double rho = rho0(d.ncase, xcr, ycr);
    double xm = 0.0, ym = 0.0;
    for (int i = 0; i < 5; ++i)
        for (int j = 0; j < 5; ++j) {
            double r = PG5[i], s = PG5[j], wi = WG5[i] * WG5[j];
            double J[4];
            jac9(xi, yi, r, s, J);
            double djacoi = wi * (J[0] * J[3] - J[2] * J[1]);
            double xg, yg;
            map9(xp, yp, r, s, &xg, &yg);
            xm = xm + rho * djacoi * xg;
            ym = ym + rho * djacoi * yg;
        }
    const double masel = d.G[3 * nc + ie];
    xc = xm / masel;
    yc = ym / masel;
}

__device__ void char_mats(double gam, double dnx, double dny, const double ul[4], const double ur[4], double q[4][4],
                          double qi[4][4])
{
    const double eps = 1.e-6;
    double rhoc1 = 1.0 / ul[0], u1 = ul[1], v1 = ul[2], e1 = ul[3];
    double p1 = dmax(eps, (gam - 1.0) * rhoc1 * (e1 - 0.5 * (u1 * u1 + v1 * v1)));
    double rhoc2 = 1.0 / ur[0], u2 = ur[1], v2 = ur[2], e2 = ur[3];
    double p2 = dmax(eps, (gam - 1.0) * rhoc2 * (e2 - 0.5 * (u2 * u2 + v2 * v2)));
    (void)p1;
    (void)p2;
    double rhof = sqrt(rhoc1 * rhoc2);
    double rhosq = sqrt(rhoc1) + sqrt(rhoc2);
    double uf = (sqrt(rhoc1) * u1 + sqrt(rhoc2) * u2) / rhosq;
    double vf = (sqrt(rhoc1) * v1 + sqrt(rhoc2) * v2) / rhosq;
    double h1 = dmax(eps, gam * e1 - (gam - 1.0) * 0.5 * (u1 * u1 + v1 * v1));
    double h2 = dmax(eps, gam * e2 - (gam - 1.0) * 0.5 * (u2 * u2 + v2 * v2));
    double hf = (sqrt(rhoc1) * h1 + sqrt(rhoc2) * h2) / rhosq;
    double sdf = sqrt(dmax(eps, (gam - 1.0) * (hf - 0.5 * (uf * uf + vf * vf))));
    double rhosd = rhof * sdf;
    double pf = sdf * sdf * rhof / gam;
    double gam1 = (gam - 1.0) / gam;
    double vnm = vf * dnx - uf * dny;
    qi[0][0] = -.5 / gam;
    qi[0][1] = .5 * dnx / rhosd - 0.5 * uf / pf * gam1;
    qi[0][2] = .5 * dny / rhosd - 0.5 * vf / pf * gam1;
    qi[0][3] = .5 / pf * gam1;
    qi[1][0] = 1.0 / gam;
    qi[1][1] = -dny * vnm / pf - uf / pf / gam;
    qi[1][2] = dnx * vnm / pf - vf / pf / gam;
    qi[1][3] = 1.0 / pf / gam;
    qi[2][0] = 0.0;
    qi[2][1] = dny / pf;
    qi[2][2] = -dnx / pf;
    qi[2][3] = 0.0;
    qi[3][0] = .5 / gam;
    qi[3][1] = .5 * dnx / rhosd + 0.5 * uf / pf * gam1;
    qi[3][2] = .5 * dny / rhosd + 0.5 * vf / pf * gam1;
    qi[3][3] = -.5 / pf * gam1;
    double vedn = uf * dnx + vf * dny;
    q[0][0] = -1.0;
    q[0][1] = gam - 1.0;
    q[0][2] = (gam - 1.0) * vnm;
    q[0][3] = 1.0;
    q[1][0] = rhosd * dnx;
    q[1][1] = 0.0;
    q[1][2] = pf * dny;
    q[1][3] = rhosd * dnx;
    q[2][0] = rhosd * dny;
    q[2][1] = 0.0;
    q[2][2] = -pf * dnx;
    q[2][3] = rhosd * dny;
    q[3][0] = pf + rhosd * vedn;
    q[3][1] = pf;
    q[3][2] = 0.0;
    q[3][3] = -pf + rhosd * vedn;
}

// One troubled cell: 9-cell stencil, local orthogonal Taylor bases,
// 1 + 4 Lagrange + 4 Hermite candidates, characteristic nonlinear weights
// per real face, back-projection (weno_stncl_quad_orthol3 and helpers).
__global__ void __launch_bounds__(64) k_q2_weno(Dev d, const double *__restrict__ Uin)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= *d.tcount) return;
    const int nc = d.nc;
    const int ie = d.tlist[t];
    // stencil topology
    int ielaj[9];
    ielaj[0] = ie;
    for (int f = 0; f < 4; ++f) ielaj[1 + f] = d.esqua[f * nc + ie];
    for (int k = 5; k < 9; ++k) ielaj[k] = -2;
    bool mismatch = false;
    for (int iv = 0; iv < 4; ++iv) {
        int jel[3] = {-1, -1, -1}, nv = 0;
        int p = d.ipq[iv * nc + ie];
        for (int k = d.esuv2[p]; k < d.esuv2[p + 1]; ++k) {
            int je = d.esuv1[k];
            if (je != ie && nv < 3) jel[nv++] = je;
        }
        int idxod = 1, idxiv[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            for (int f = 0; f < 4; ++f)
                if (jel[a] == ielaj[1 + f]) {
                    idxod = idxod * (f + 1);
                    idxiv[a] = 1;
                }
        for (int a = 0; a < 3; ++a) {
            if (idxiv[a]) continue;
            if (idxod == 2) ielaj[5] = jel[a];
            else if (idxod == 6) ielaj[6] = jel[a];
            else if (idxod == 12) ielaj[7] = jel[a];
            else if (idxod == 4) ielaj[8] = jel[a];
            else mismatch = true;
        }
    }
    if (mismatch) {
        atomicAdd(d.wenobad, 1);
        return;
    }
    for (int k = 0; k < 9; ++k)
        if (ielaj[k] < 0) return;

    // velocity-aligned frame of the troubled cell
    double m[4];
    {
        double a = UD(Uin, 1, 0, ie), b = UD(Uin, 2, 0, ie);
        if (sqrt(a * a + b * b) > 1.e-10) {
            m[0] = a / sqrt(a * a + b * b);
            m[1] = b / sqrt(a * a + b * b);
            m[2] = -b / sqrt(a * a + b * b);
            m[3] = a / sqrt(a * a + b * b);
        } else {
            m[0] = m[1] = m[2] = m[3] = 0.0;
        }
    }
    double unkrl[9][4][6];
    for (int s = 0; s < 9; ++s)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) unkrl[s][q][id] = UD(Uin, q, id, ielaj[s]);
    for (int s = 0; s < 9; ++s)
        for (int id = 0; id < 6; ++id) rotf(m, unkrl[s][1][id], unkrl[s][2][id]);

    double geopl[9][5], dxcj[9];
    CellPts P[9];
    for (int s = 0; s < 9; ++s) {
        int je = ielaj[s];
        double xcl, ycl;
        mass_centre(d, je, xcl, ycl);
        rotf(m, xcl, ycl);
        geopl[s][0] = xcl;
        geopl[s][1] = ycl;
        dxcj[s] = sqrt(d.G[2 * nc + je]);
        cell_points(d, je, m, P[s]);
        double v1 = 0.0, v2 = 0.0, v3 = 0.0;
        for (int g = 0; g < 25; ++g) {
            double b2 = (P[s].xl[g] - xcl) / dxcj[s];
            double b3 = (P[s].yl[g] - ycl) / dxcj[s];
            double rd = P[s].rhogi[g] * P[s].djacoi[g];
            v1 = v1 + rd * b2 * b2;
            v2 = v2 + rd * b3 * b3;
            v3 = v3 + rd * b2 * b3;
        }
        double masel = d.G[3 * nc + je];
        geopl[s][2] = 0.5 * (v1 / masel);
        geopl[s][3] = 0.5 * (v2 / masel);
        geopl[s][4] = v3 / masel;
    }

    double unkpl[9][4][6];
    double mtbbq[25], mtbbi[25];
    for (int s = 0; s < 9; ++s) {
        int je = ielaj[s];
        double rcj = d.G[0 * nc + je], scj = d.G[1 * nc + je];
        double g0 = d.G[18 * nc + je], g1 = d.G[19 * nc + je], g2 = d.G[20 * nc + je];
        double bb[25], bbq[25];
        for (int k = 0; k < 25; ++k) bb[k] = bbq[k] = 0.0;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                const int g = 5 * i + j;
                double bq[6], bl[6];
                basis6(PG5[i], PG5[j], rcj, scj, g0, g1, g2, bq);
                bl[0] = 1.0;
                bl[1] = (P[s].xl[g] - geopl[s][0]) / dxcj[s];
                bl[2] = (P[s].yl[g] - geopl[s][1]) / dxcj[s];
                bl[3] = 0.5 * bl[1] * bl[1] - geopl[s][2];
                bl[4] = 0.5 * bl[2] * bl[2] - geopl[s][3];
                bl[5] = bl[1] * bl[2] - geopl[s][4];
                double rd = P[s].rhogi[g] * P[s].djacoi[g];
                for (int a = 1; a < 6; ++a)
                    for (int b = 1; b < 6; ++b) {
                        bb[(a - 1) + 5 * (b - 1)] = bb[(a - 1) + 5 * (b - 1)] + rd * bq[a] * bl[b];
                        bbq[(a - 1) + 5 * (b - 1)] = bbq[(a - 1) + 5 * (b - 1)] + rd * bl[a] * bl[b];
                    }
            }
        double rhs[4][6];
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 6; ++a) rhs[q][a] = 0.0;
        for (int a = 1; a < 6; ++a)
            for (int b = 1; b < 6; ++b)
                for (int q = 0; q < 4; ++q) rhs[q][a] = rhs[q][a] + unkrl[s][q][b] * bb[(b - 1) + 5 * (a - 1)];
        double bbqi[25], bbi[25];
        inv5(bbq, bbqi);
        inv5(bb, bbi);
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 6; ++a) unkpl[s][q][a] = 0.0;
        for (int a = 1; a < 6; ++a)
            for (int b = 1; b < 6; ++b)
                for (int q = 0; q < 4; ++q) unkpl[s][q][a] = unkpl[s][q][a] + rhs[q][b] * bbqi[(b - 1) + 5 * (a - 1)];
        if (s == 0) {
            for (int k = 0; k < 25; ++k) {
                mtbbq[k] = bbq[k];
                mtbbi[k] = bbi[k];
            }
        }
        for (int q = 0; q < 4; ++q) unkpl[s][q][0] = unkrl[s][q][0];
    }

    const int ST1[4][5] = {{1, 4, 5, 8, 7}, {2, 1, 6, 5, 8}, {3, 2, 7, 6, 5}, {4, 3, 8, 7, 6}};
    const int ST2[4][3] = {{1, 5, 8}, {2, 5, 6}, {3, 7, 6}, {4, 7, 8}};
    double unksl[9][4][6];
    for (int s = 0; s < 9; ++s)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) unksl[s][q][id] = 0.0;
    for (int s = 0; s < 9; ++s)
        for (int q = 0; q < 4; ++q) unksl[s][q][0] = unkpl[0][q][0];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) unksl[0][q][id] = unkpl[0][q][id];
    const double dxcl = sqrt(d.G[2 * nc + ie]);
    const double xmcl = geopl[0][0], ymcl = geopl[0][1];
    int isten = 0;
    for (int is = 0; is < 4; ++is) {
        ++isten;
        double mat[25], rhsl[4][5];
        for (int k = 0; k < 25; ++k) mat[k] = 0.0;
        for (int e = 0; e < 5; ++e) {
            int sj = ST1[is][e];
            double masej = d.G[3 * nc + ielaj[sj]];
            for (int g = 0; g < 25; ++g) {
                double bl[6];
                bl[1] = (P[sj].xl[g] - xmcl) / dxcl;
                bl[2] = (P[sj].yl[g] - ymcl) / dxcl;
                bl[3] = 0.5 * bl[1] * bl[1] - geopl[0][2];
                bl[4] = 0.5 * bl[2] * bl[2] - geopl[0][3];
                bl[5] = bl[1] * bl[2] - geopl[0][4];
                double rd = P[sj].rhogi[g] * P[sj].djacoi[g];
                for (int k = 0; k < 5; ++k) mat[e + 5 * k] = mat[e + 5 * k] + rd * bl[k + 1] / masej;
            }
            for (int q = 0; q < 4; ++q) rhsl[q][e] = unkpl[sj][q][0] - unkpl[0][q][0];
        }
        double mi[25];
        inv5(mat, mi);
        for (int a = 1; a < 6; ++a)
            for (int e = 0; e < 5; ++e)
                for (int q = 0; q < 4; ++q) unksl[isten][q][a] = unksl[isten][q][a] + mi[(a - 1) + 5 * e] * rhsl[q][e];
    }
    for (int is = 0; is < 4; ++is) {
        ++isten;
        double mh[45], rh[4][9];
        for (int k = 0; k < 45; ++k) mh[k] = 0.0;
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 9; ++k) rh[q][k] = 0.0;
        for (int e = 0; e < 3; ++e) {
            int sj = ST2[is][e];
            double masej = d.G[3 * nc + ielaj[sj]];
            double dxj = dxcj[sj];
            double xclj = geopl[sj][0], yclj = geopl[sj][1];
            const int row = 3 + 2 * e;
            for (int g = 0; g < 25; ++g) {
                double bl[6], bj[6];
                bl[1] = (P[sj].xl[g] - xmcl) / dxcl;
                bl[2] = (P[sj].yl[g] - ymcl) / dxcl;
                bl[3] = 0.5 * bl[1] * bl[1] - geopl[0][2];
                bl[4] = 0.5 * bl[2] * bl[2] - geopl[0][3];
                bl[5] = bl[1] * bl[2] - geopl[0][4];
                double rd = P[sj].rhogi[g] * P[sj].djacoi[g];
                for (int k = 0; k < 5; ++k) mh[e + 9 * k] = mh[e + 9 * k] + rd * bl[k + 1] / masej;
                bj[1] = (P[sj].xl[g] - xclj) / dxj;
                bj[2] = (P[sj].yl[g] - yclj) / dxj;
                bj[3] = 0.5 * bj[1] * bj[1] - geopl[sj][2];
                bj[4] = 0.5 * bj[2] * bj[2] - geopl[sj][3];
                bj[5] = bj[1] * bj[2] - geopl[sj][4];
                for (int a = 1; a < 6; ++a) {
                    mh[row + 9 * (a - 1)] = mh[row + 9 * (a - 1)] + rd * bl[a] * bj[1] / masej;
                    mh[row + 1 + 9 * (a - 1)] = mh[row + 1 + 9 * (a - 1)] + rd * bl[a] * bj[2] / masej;
                }
                for (int a = 1; a < 6; ++a)
                    for (int q = 0; q < 4; ++q) {
                        rh[q][row] = rh[q][row] + rd * bj[a] * bj[1] * unkpl[sj][q][a] / masej;
                        rh[q][row + 1] = rh[q][row + 1] + rd * bj[a] * bj[2] * unkpl[sj][q][a] / masej;
                    }
            }
            for (int q = 0; q < 4; ++q) rh[q][e] = unkpl[sj][q][0] - unkpl[0][q][0];
        }
        double ms[25], rs[4][5];
        for (int a = 0; a < 5; ++a)
            for (int b = 0; b < 5; ++b) {
                double v = 0.0;
                for (int k = 0; k < 9; ++k) v = v + mh[k + 9 * a] * mh[k + 9 * b];
                ms[a + 5 * b] = v;
            }
        for (int q = 0; q < 4; ++q)
            for (int a = 0; a < 5; ++a) {
                double v = 0.0;
                for (int k = 0; k < 9; ++k) v = v + mh[k + 9 * a] * rh[q][k];
                rs[q][a] = v;
            }
        double mi[25];
        inv5(ms, mi);
        for (int a = 1; a < 6; ++a)
            for (int e = 0; e < 5; ++e)
                for (int q = 0; q < 4; ++q) unksl[isten][q][a] = unksl[isten][q][a] + mi[(a - 1) + 5 * e] * rs[q][e];
    }
    const int nst = isten + 1;

    double xp[9], yp[9];
    load9(d, d.xs, d.ys, ie, xp, yp);
    const double volel = d.G[2 * nc + ie];
    const double dxc = sqrt(volel), dyc = dxc;
    double unk_phy[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) unk_phy[q][id] = 0.0;
    double uf[4][4][6];
    int ifprj = 0;
    for (int f = 0; f < 4; ++f) {
        int je = d.esqua[f * nc + ie];
        if (je >= nc) continue;
        int a0 = f, a1 = (f + 1) & 3;
        double dtx = xp[a1] - xp[a0];
        double dty = yp[a1] - yp[a0];
        double dlgt = sqrt(dtx * dtx + dty * dty);
        double dnx = dty / dlgt, dny = -dtx / dlgt;
        rotf(m, dnx, dny);
        double ul[4], ur[4];
        for (int q = 0; q < 4; ++q) {
            ul[q] = UD(Uin, q, 0, ie);
            ur[q] = UD(Uin, q, 0, je);
        }
        rotf(m, ul[1], ul[2]);
        rotf(m, ur[1], ur[2]);
        double qm[4][4], qi[4][4];
        char_mats(d.gamlg, dnx, dny, ul, ur, qm, qi);
        double os1[9];
        for (int s = 0; s < nst; ++s) {
            double o[4];
            for (int k = 0; k < 4; ++k) {
                double c1 = 0.0, c2 = 0.0;
                for (int q = 0; q < 4; ++q) {
                    c1 = c1 + qi[k][q] * unksl[s][q][1];
                    c2 = c2 + qi[k][q] * unksl[s][q][2];
                }
                o[k] = (c1 / dxc) * (c1 / dxc) + (c2 / dyc) * (c2 / dyc);
            }
            os1[s] = dmax(dmax(dmax(o[0], o[1]), o[2]), o[3]);
        }
        double wl[9], wh[9], wt = 0.0;
        wl[0] = .2;
        for (int s = 1; s < nst; ++s) wl[s] = 0.8 / ((double)nst - 1.0);
        for (int s = 0; s < nst; ++s) {
            double tt = 1.e-6 + os1[s];
            wh[s] = wl[s] / (tt * tt);
        }
        for (int s = 0; s < nst; ++s) wt = wt + wh[s];
        double fch[4][6];
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) fch[q][id] = 0.0;
        for (int s = 0; s < nst; ++s) {
            double ws = wh[s] / wt;
            for (int k = 0; k < 4; ++k)
                for (int id = 0; id < 6; ++id) {
                    double c = 0.0;
                    for (int q = 0; q < 4; ++q) c = c + qi[k][q] * unksl[s][q][id];
                    fch[k][id] = fch[k][id] + ws * c;
                }
        }
        for (int k = 0; k < 4; ++k)
            for (int id = 0; id < 6; ++id) uf[ifprj][k][id] = 0.0;
        for (int q = 0; q < 4; ++q)
            for (int id = 1; id < 6; ++id)
                for (int k = 0; k < 4; ++k) uf[ifprj][k][id] = uf[ifprj][k][id] + qm[k][q] * fch[q][id];
        ++ifprj;
    }
    const double wf = 1.0 / (double)ifprj;
    for (int a = 0; a < ifprj; ++a)
        for (int q = 0; q < 4; ++q)
            for (int id = 1; id < 6; ++id) unk_phy[q][id] = unk_phy[q][id] + wf * uf[a][q][id];
    double rhsuk[4][5];
    for (int q = 0; q < 4; ++q)
        for (int a = 0; a < 5; ++a) rhsuk[q][a] = 0.0;
    for (int a = 1; a < 6; ++a)
        for (int b = 1; b < 6; ++b)
            for (int q = 0; q < 4; ++q) rhsuk[q][a - 1] = rhsuk[q][a - 1] + unk_phy[q][b] * mtbbq[(b - 1) + 5 * (a - 1)];
    double out[4][6];
    for (int q = 0; q < 4; ++q)
        for (int a = 0; a < 6; ++a) out[q][a] = 0.0;
    for (int a = 1; a < 6; ++a)
        for (int b = 1; b < 6; ++b)
            for (int q = 0; q < 4; ++q) out[q][a] = out[q][a] + rhsuk[q][b - 1] * mtbbi[(b - 1) + 5 * (a - 1)];
    const double i11 = m[3], i12 = -m[1], i21 = -m[2], i22 = m[0];
    for (int a = 1; a < 6; ++a) {
        double x = out[1][a] * i11 + out[2][a] * i12;
        double y = out[1][a] * i21 + out[2][a] * i22;
        out[1][a] = x;
        out[2][a] = y;
    }
    for (int q = 0; q < 4; ++q)
        for (int a = 1; a < 6; ++a) UD(d.Ul, q, a, ie) = out[q][a];
}

// ------------------------------------------------------------- corners --
// Face geometry (stored for the face integral), sub-cell density and mean
// specific volume, vertex states and the 32 corner-face Riemann records.
// Records: rec[(r*32 + cf)*nc + ie], cf = f + 2*(ivsg + 4*isg),
// r = 0: m = mu*A, 1-2: m u_c, 3-4: A sigma n.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2_corner(Dev d)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    double xp[9], yp[9];
    load9(d, d.xs, d.ys, ie, xp, yp);
    // ---- face geometry
    double gq[16][3];
    {
        const double pq[4][2] = {{0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
        for (int k = 0; k < 4; ++k) {
            double xv, yv;
            map9(xp, yp, pq[k][0], pq[k][1], &xv, &yv);
            double fx[3] = {xp[4 + k], xp[8], xv};
            double fy[3] = {yp[4 + k], yp[8], yv};
            const double r = -1.0;
            const double dd[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
            double dxdr = 0.0, dydr = 0.0;
            for (int i = 0; i < 3; ++i) {
                dxdr = dxdr + dd[i] * fx[i];
                dydr = dydr + dd[i] * fy[i];
            }
            double djaco = sqrt(dxdr * dxdr + dydr * dydr);
            double anx = -dydr, any = dxdr;
            gq[8 + k][0] = anx / sqrt(anx * anx + any * any);
            gq[8 + k][1] = any / sqrt(anx * anx + any * any);
            gq[8 + k][2] = djaco * 2.0;
        }
        const double pi[12][2] = {{-1, -1}, {1, -1}, {1, -1}, {1, 1}, {1, 1}, {-1, 1},
                                  {-1, 1},  {-1, -1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
        const double vn[12][2] = {{0, -1}, {0, -1}, {1, 0}, {1, 0}, {0, 1}, {0, 1},
                                  {-1, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
        const int slot[12] = {0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15};
        for (int ig = 0; ig < 12; ++ig) {
            double J[4];
            jac9(xp, yp, pi[ig][0], pi[ig][1], J);
            double c11 = J[3], c12 = -J[2], c21 = -J[1], c22 = J[0];
            double anx = c11 * vn[ig][0] + c12 * vn[ig][1];
            double any = c21 * vn[ig][0] + c22 * vn[ig][1];
            double w1 = anx * 2.0, w2 = any * 2.0;
            double farea = sqrt(w1 * w1 + w2 * w2);
            gq[slot[ig]][0] = anx * 2.0 / farea;
            gq[slot[ig]][1] = any * 2.0 / farea;
            gq[slot[ig]][2] = farea;
        }
        for (int j = 0; j < 16; ++j)
            for (int k = 0; k < 3; ++k) d.gq[(3 * j + k) * nc + ie] = gq[j][k];
    }
    double u[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) u[q][id] = UD(d.Ul, q, id, ie);
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    const double gg0 = d.G[18 * nc + ie], gg1 = d.G[19 * nc + ie], gg2 = d.G[20 * nc + ie];
    // ---- sub-cell density: initial mass / current area
    double rsub[4];
    {
        const double pc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                                  {-1.0, .5},   {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
        const int ipsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
        double xc[12], yc[12];
        for (int k = 0; k < 12; ++k) map9(xp, yp, pc[k][0], pc[k][1], &xc[k], &yc[k]);
        for (int isg = 0; isg < 4; ++isg) {
            double xs[9], ys[9];
            for (int k = 0; k < 4; ++k) {
                xs[k] = xp[IPQSG[isg][k]];
                ys[k] = yp[IPQSG[isg][k]];
                xs[4 + k] = xc[ipsc[isg][k]];
                ys[4 + k] = yc[ipsc[isg][k]];
            }
            xs[8] = -0.25 * (xs[0] + xs[1] + xs[2] + xs[3]) + 0.5 * (xs[4] + xs[5] + xs[6] + xs[7]);
            ys[8] = -0.25 * (ys[0] + ys[1] + ys[2] + ys[3]) + 0.5 * (ys[4] + ys[5] + ys[6] + ys[7]);
            double a = 0.0;
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j) {
                    double J[4];
                    jac9(xs, ys, PG5[i], PG5[j], J);
                    a = a + WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
                }
            rsub[isg] = 0.0 + (1.0 / a) * d.msub[isg * nc + ie];
        }
    }
    // ---- sub-cell mean specific volume (aflim(1) = 1 on this path)
    double uq[4];
    {
        const double sr[4] = {-1.0, 1.0, 1.0, -1.0}, ss[4] = {-1.0, -1.0, 1.0, 1.0};
        const double rhomc = u[0][0];
        for (int isg = 0; isg < 4; ++isg) {
            double masel = 0.0, volel = 0.0;
            for (int i = 0; i < 5; ++i)
                for (int j = 0; j < 5; ++j) {
                    double r = (PG5[i] + sr[isg]) / 2.0;
                    double s = (PG5[j] + ss[isg]) / 2.0;
                    double b[6];
                    basis6(r, s, rc, sc, gg0, gg1, gg2, b);
                    double t = 0.0;
                    for (int id = 0; id < 6; ++id) t = t + u[0][id] * b[id];
                    t = rhomc + 1.0 * (t - rhomc);
                    double J[4];
                    jac9(xp, yp, r, s, J);
                    double djaco = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
                    masel = masel + 1.0 / t * djaco;
                    volel = volel + djaco;
                }
            uq[isg] = volel / masel;
        }
    }
    // ---- vertex states of the P2 polynomial at the 9 nodes
    double un[9][4];
    for (int iv = 0; iv < 9; ++iv) {
        double b[6];
        basis6(XV9[iv], YV9[iv], rc, sc, gg0, gg1, gg2, b);
        for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
        for (int id = 0; id < 6; ++id)
            for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + u[q][id] * b[id];
    }
    const double rhomc = u[0][0], uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
    const double rhoct = 1.0 / rhomc;
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    for (int isg = 0; isg < 4; ++isg) {
        const double dcor = d.cdrho * (rsub[isg] - 1.0 / uq[isg]);
        for (int iv = 0; iv < 4; ++iv) {
            const int ln = IPQSG[isg][iv];
            const int p = d.ipq[ln * nc + ie];
            const double rhovt = 1.0 / un[ln][0];
            const double rhovsg = rhovt + dcor;
            const double uv = un[ln][1], vv = un[ln][2], ev = un[ln][3];
            const double pv = dmax(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)));
            const double dux = d.usx[p] - uv, duy = d.usy[p] - vv;
            const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
            for (int f = 0; f < 2; ++f) {
                int tt = FNQSG[isg][2 * iv + f];
                int sl = (tt > 0 ? tt : -tt) - 1;
                double sg = tt > 0 ? 1.0 : -1.0;
                double nx = sg * gq[sl][0], ny = sg * gq[sl][1];
                double area = gq[sl][2] * WFR[isg][iv][f];
                double deltu = fabs(dux * nx + duy * ny);
                double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
                double m11;
                if (a3 < 1.e12) {
                    m11 = mur * area * fabs(1.0);
                } else {
                    double nrm = sqrt(dux * dux + duy * duy);
                    m11 = mur * area * fabs(nx * (-dux / nrm) + ny * (-duy / nrm));
                }
                const int cf = f + 2 * (iv + 4 * isg);
                d.rec[(0 * 32 + cf) * nc + ie] = m11;
                d.rec[(1 * 32 + cf) * nc + ie] = m11 * uv + 0.0 * vv;
                d.rec[(2 * 32 + cf) * nc + ie] = 0.0 * uv + m11 * vv;
                d.rec[(3 * 32 + cf) * nc + ie] = -pv * area * nx + 0.0 * area * ny;
                d.rec[(4 * 32 + cf) * nc + ie] = 0.0 * area * nx + -pv * area * ny;
            }
        }
    }
}

// Nodal solve over the sub-cell corners: contributions gathered in the
// reference's accumulation order (cell, sub-cell, vertex, face); walls on
// all three face nodes (Getrhslaggauss.f90:4494-4545).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_node_q2(Dev d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int nc = d.nc;
    double a11 = 0.0, a21 = 0.0, a12 = 0.0, a22 = 0.0, mu1 = 0.0, mu2 = 0.0, sn1 = 0.0, sn2 = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k];
        const int ie = e >> 4, slot = e & 15;
        for (int f = 0; f < 2; ++f) {
            const int cf = f + 2 * slot;
            double m = d.rec[(0 * 32 + cf) * nc + ie];
            a11 = a11 + m;
            a21 = a21 + 0.0;
            a12 = a12 + 0.0;
            a22 = a22 + m;
            mu1 = mu1 + d.rec[(1 * 32 + cf) * nc + ie];
            mu2 = mu2 + d.rec[(2 * 32 + cf) * nc + ie];
            sn1 = sn1 + d.rec[(3 * 32 + cf) * nc + ie];
            sn2 = sn2 + d.rec[(4 * 32 + cf) * nc + ie];
        }
    }
    const uint8_t fl = d.nflag[p];
    double ux = d.usx[p], uy = d.usy[p];
    if (!(fl & 4)) {
        double det = a11 * a22 - a21 * a12;
        double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        double r1 = mu1 - sn1, r2 = mu2 - sn2;
        ux = i11 * r1 + i12 * r2;
        uy = i21 * r1 + i22 * r2;
    }
    if (fl & 1) uy = 0.0;
    if (fl & 2) ux = 0.0;
    d.usx[p] = ux;
    d.usy[p] = uy;
}

// Sub-cell face integral, Gauss-Lobatto domain integral, the (identity)
// limiter update and, with apply_update, the mass-matrix solve and RK blend.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_cellrhs_q2(Dev d, double alfa, int apply_update)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie == 0) *d.tcount = 0;  // WENO list of this stage consumed; reset for the next
    if (ie >= nc) return;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    double u[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) u[q][id] = UD(d.Ul, q, id, ie);
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    const double gg0 = d.G[18 * nc + ie], gg1 = d.G[19 * nc + ie], gg2 = d.G[20 * nc + ie];
    double r[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) r[q][id] = 0.0;
    // ---- face integral
    {
        double bq[9][6];
        for (int iv = 0; iv < 9; ++iv) basis6(XV9[iv], YV9[iv], rc, sc, gg0, gg1, gg2, bq[iv]);
        for (int isg = 0; isg < 4; ++isg) {
            for (int iv = 0; iv < 4; ++iv) {
                const int ln = IPQSG[isg][iv];
                const int p = d.ipq[ln * nc + ie];
                const double us1 = d.usx[p], us2 = d.usy[p];
                double nx[2], ny[2], ar[2], fx[2], fy[2];
                for (int f = 0; f < 2; ++f) {
                    int tt = FNQSG[isg][2 * iv + f];
                    int sl = (tt > 0 ? tt : -tt) - 1;
                    double sg = tt > 0 ? 1.0 : -1.0;
                    nx[f] = sg * d.gq[(3 * sl + 0) * nc + ie];
                    ny[f] = sg * d.gq[(3 * sl + 1) * nc + ie];
                    ar[f] = d.gq[(3 * sl + 2) * nc + ie] * WFV[isg][iv][f];
                    const int cf = f + 2 * (iv + 4 * isg);
                    double m11 = d.rec[(0 * 32 + cf) * nc + ie];
                    double m1 = d.rec[(1 * 32 + cf) * nc + ie], m2 = d.rec[(2 * 32 + cf) * nc + ie];
                    double s1 = d.rec[(3 * 32 + cf) * nc + ie], s2 = d.rec[(4 * 32 + cf) * nc + ie];
                    fx[f] = s1 + m11 * us1 + 0.0 * us2 - m1;
                    fy[f] = s2 + 0.0 * us1 + m11 * us2 - m2;
                }
                const double w1 = WFM[isg][iv][0], w2 = WFM[isg][iv][1];
                for (int id = 0; id < 6; ++id) {
                    const double b = bq[ln][id];
                    r[0][id] = r[0][id] + us1 * nx[0] * ar[0] * b + us2 * ny[0] * ar[0] * b + us1 * nx[1] * ar[1] * b +
                               us2 * ny[1] * ar[1] * b;
                    r[1][id] = r[1][id] + fx[0] * b * w1 + fx[1] * b * w2;
                    r[2][id] = r[2][id] + fy[0] * b * w1 + fy[1] * b * w2;
                    r[3][id] = r[3][id] + us1 * fx[0] * b * w1 + us2 * fy[0] * b * w1 + us1 * fx[1] * b * w2 +
                               us2 * fy[1] * b * w2;
                }
            }
        }
    }
    // ---- domain integral (5x5 Gauss-Lobatto, djaco = w, cofactor form)
    {
        double xp[9], yp[9];
        load9(d, d.xs, d.ys, ie, xp, yp);
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                const double rr = PL5[i], ss = PL5[j], wi = WL5[i] * WL5[j];
                double J[4];
                jac9(xp, yp, rr, ss, J);
                const double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
                const double dbdr[6] = {0.0, 1.0 / 1.0, 0.0, (rr - rc) / 1.0, 0.0, (ss - sc) / 1.0 / 1.0};
                const double dbds[6] = {0.0, 0.0, 1.0 / 1.0, 0.0, (ss - sc) / 1.0, (rr - rc) / 1.0 / 1.0};
                double b[6];
                basis6(rr, ss, rc, sc, gg0, gg1, gg2, b);
                double un[4] = {0.0, 0.0, 0.0, 0.0};
                for (int id = 0; id < 6; ++id)
                    for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
                const double rhoad = 1.0 / un[0];
                const double uu = un[1], vv = un[2];
                const double pres = dmax(eps, g1 * rhoad * (un[3] - 0.5 * (uu * uu + vv * vv)));
                for (int id = 0; id < 6; ++id) {
                    const double a = dyds * dbdr[id] + (-dydr) * dbds[id];
                    const double c = (-dxds) * dbdr[id] + dxdr * dbds[id];
                    const double f0 = a * uu + c * vv;
                    r[0][id] = r[0][id] - f0 * wi;
                    r[1][id] = r[1][id] - a * (-pres) * wi;
                    r[2][id] = r[2][id] - c * (-pres) * wi;
                    r[3][id] = r[3][id] - (a * uu + c * vv) * (-pres) * wi;
                }
            }
    }
    // ---- derivative update with aflim = 1, afvec = I (Getrhslagmc.f90:7695-7708)
    double ul[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) ul[q][id] = u[q][id];
    ul[0][1] = u[0][1] * 1.0;
    ul[0][2] = u[0][2] * 1.0;
    ul[1][1] = 1.0 * u[1][1] + 0.0 * u[2][1];
    ul[1][2] = 1.0 * u[1][2] + 0.0 * u[2][2];
    ul[2][1] = 0.0 * u[1][1] + 1.0 * u[2][1];
    ul[2][2] = 0.0 * u[1][2] + 1.0 * u[2][2];
    ul[3][1] = u[3][1] * 1.0;
    ul[3][2] = u[3][2] * 1.0;
    if (apply_update) {
        const double dt = d.scal[0];
        double m[6][6];
        for (int a = 0; a < 6; ++a)
            for (int b = 0; b < 6; ++b) m[a][b] = 0.0;
        m[0][0] = d.A[15 * nc + ie];
        int k = 0;
        for (int a = 1; a < 6; ++a)
            for (int b = a; b < 6; ++b) {
                double v = d.A[k * nc + ie];
                m[a][b] = v;
                m[b][a] = v;
                ++k;
            }
        double rs[4][6];
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) rs[q][id] = dt * r[q][id];
        for (int id = 0; id < 6; ++id) {
            double s4[4] = {0.0, 0.0, 0.0, 0.0};
            for (int kk = 0; kk < 6; ++kk)
                for (int q = 0; q < 4; ++q) s4[q] = s4[q] + m[id][kk] * rs[q][kk];
            for (int q = 0; q < 4; ++q)
                UD(d.Us, q, id, ie) = alfa * UD(d.Un, q, id, ie) + (1.0 - alfa) * (ul[q][id] + s4[q]);
        }
    } else {
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) {
                UD(d.Us, q, id, ie) = ul[q][id];
                UD(d.R, q, id, ie) = r[q][id];
            }
    }
}

// Initial sub-cell masses (getrhosubg_averg second half): time-invariant,
// computed once from the initial coordinates.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2_submass(Dev d)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    double xi[9], yi[9];
    load9(d, d.x0, d.y0, ie, xi, yi);
    const double pc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                              {-1.0, .5},   {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
    const int ipsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
    double xc[12], yc[12];
    for (int k = 0; k < 12; ++k) map9(xi, yi, pc[k][0], pc[k][1], &xc[k], &yc[k]);
    double xcc, ycc;
    map9(xi, yi, 0.0, 0.0, &xcc, &ycc);
    const double rho = (d.ncase == 6) ? rho0(6, xcc, ycc) : 1.0;
    for (int isg = 0; isg < 4; ++isg) {
        double xs[9], ys[9];
        for (int k = 0; k < 4; ++k) {
            xs[k] = xi[IPQSG[isg][k]];
            ys[k] = yi[IPQSG[isg][k]];
            xs[4 + k] = xc[ipsc[isg][k]];
            ys[4 + k] = yc[ipsc[isg][k]];
        }
        xs[8] = -0.25 * (xs[0] + xs[1] + xs[2] + xs[3]) + 0.5 * (xs[4] + xs[5] + xs[6] + xs[7]);
        ys[8] = -0.25 * (ys[0] + ys[1] + ys[2] + ys[3]) + 0.5 * (ys[4] + ys[5] + ys[6] + ys[7]);
        double ms = 0.0;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                double J[4];
                jac9(xs, ys, PG5[i], PG5[j], J);
                double djaco = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
                ms = ms + rho * 1.0 * djaco;
            }
        d.msub[isg * nc + ie] = ms;
    }
}

#undef UD

// ---- host-side table set-up ------------------------------------------
// Enumeration of the Leibniz terms in Matinv225's order: the column of the
// last row ascends slowest, then the row before, ...; the first row takes
// the remaining column.
static void perm_table(int m, std::vector<std::vector<signed char>> &out)
{
    out.clear();
    int p[5];
    bool used[5] = {false, false, false, false, false};
    // recursive lambda over levels m-1 .. 1
    struct Rec {
        static void go(int level, int m, int *p, bool *used, std::vector<std::vector<signed char>> &out)
        {
            if (level == 0) {
                int j = 0;
                while (used[j]) ++j;
                p[0] = j;
                int inv = 0;
                for (int x = 0; x < m; ++x)
                    for (int y = x + 1; y < m; ++y)
                        if (p[x] > p[y]) ++inv;
                std::vector<signed char> v(m + 1);
                for (int k = 0; k < m; ++k) v[k] = (signed char)p[k];
                v[m] = (inv & 1) ? -1 : 1;
                out.push_back(v);
                return;
            }
            for (int j = 0; j < m; ++j) {
                if (used[j]) continue;
                used[j] = true;
                p[level] = j;
                go(level - 1, m, p, used, out);
                used[j] = false;
            }
        }
    };
    Rec::go(m - 1, m, p, used, out);
}

static int q2_tables()
{
    static bool done = false;
    if (done) return 0;
    std::vector<std::vector<signed char>> t5, t4;
    perm_table(5, t5);
    perm_table(4, t4);
    signed char p5[120][6], p4[24][5];
    for (int i = 0; i < 120; ++i)
        for (int k = 0; k < 6; ++k) p5[i][k] = t5[i][k];
    for (int i = 0; i < 24; ++i)
        for (int k = 0; k < 5; ++k) p4[i][k] = t4[i][k];
    const double rt[4][4][2] = {
        {{4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 8.0 / 3.0}, {1.0, 1.0}, {8.0 / 3.0, 4.0 / 3.0}},
        {{8.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 8.0 / 3.0}, {1.0, 1.0}},
        {{1.0, 1.0}, {8.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 8.0 / 3.0}},
        {{4.0 / 3.0, 8.0 / 3.0}, {1.0, 1.0}, {8.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}}};
    const double mt[4][4][2] = {{{1, 1}, {1, 1}, {0, 0}, {1, 1}},
                                {{1, 1}, {1, 1}, {1, 1}, {0, 0}},
                                {{0, 0}, {1, 1}, {1, 1}, {1, 1}},
                                {{1, 1}, {0, 0}, {1, 1}, {1, 1}}};
    const double vt[4][4][2] = {
        {{4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 0.0}, {0.0, 0.0}, {0.0, 4.0 / 3.0}},
        {{0.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 0.0 / 3.0}, {0.0, 0.0}},
        {{0.0, 0.0}, {0.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 0.0 / 3.0}},
        {{4.0 / 3.0, 0.0 / 3.0}, {0.0, 0.0}, {0.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}}};
    double wfr[4][4][2], wfm[4][4][2], wfv[4][4][2];
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int f = 0; f < 2; ++f) {
                wfr[a][b][f] = 0.25 * rt[a][b][f];
                wfm[a][b][f] = mt[a][b][f];
                wfv[a][b][f] = 0.25 * vt[a][b][f];
            }
    if (hipMemcpyToSymbol(HIP_SYMBOL(PERM5), p5, sizeof(p5)) != hipSuccess) return LAG2D_EDEVICE;
    if (hipMemcpyToSymbol(HIP_SYMBOL(PERM4), p4, sizeof(p4)) != hipSuccess) return LAG2D_EDEVICE;
    if (hipMemcpyToSymbol(HIP_SYMBOL(WFR), wfr, sizeof(wfr)) != hipSuccess) return LAG2D_EDEVICE;
    if (hipMemcpyToSymbol(HIP_SYMBOL(WFM), wfm, sizeof(wfm)) != hipSuccess) return LAG2D_EDEVICE;
    if (hipMemcpyToSymbol(HIP_SYMBOL(WFV), wfv, sizeof(wfv)) != hipSuccess) return LAG2D_EDEVICE;
    done = true;
    return 0;
}

}  // namespace lag2d

using namespace lag2d;

static inline int nblk(int n) { return (n + LAG2D_BLOCK - 1) / LAG2D_BLOCK; }

int upload_extra(lag2d_ctx &x)
{
    if (x.cfg.ncurv != 1) return 0;
    int rc = q2_tables();
    if (rc) return rc;
    hipLaunchKernelGGL(k_q2_submass, dim3(nblk(x.nc)), dim3(LAG2D_BLOCK), 0, x.stream, x.d);
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}

int launch_rhs_q2(lag2d_ctx *x, Dev &ds, const double *Uin, double alfa, int apply_update)
{
    const int nc = x->nc, np = x->np;
    x->tic(KID_Q2_PREP);
    hipLaunchKernelGGL(k_q2_prep, dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, Uin);
    x->toc(KID_Q2_PREP);
    x->tic(KID_Q2_WENO);
    hipLaunchKernelGGL(k_q2_weno, dim3((nc + 63) / 64), dim3(64), 0, x->stream, ds, Uin);
    x->toc(KID_Q2_WENO);
    x->tic(KID_Q2_CORNER);
    hipLaunchKernelGGL(k_q2_corner, dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
    x->toc(KID_Q2_CORNER);
    x->tic(KID_NODE_Q2);
    hipLaunchKernelGGL(k_node_q2, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
    x->toc(KID_NODE_Q2);
    x->tic(KID_CELLRHS_Q2);
    hipLaunchKernelGGL(k_cellrhs_q2, dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, alfa, apply_update);
    x->toc(KID_CELLRHS_Q2);
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}
