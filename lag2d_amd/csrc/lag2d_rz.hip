// Area-weighted RZ path (nrz=2) on straight P1 cells, triangles and quads:
// cylindrical coordinates with r = y (the axis is the wall y = 0), where the
// equations are integrated over the cell area and the radius enters as the
// cell's mean radius geoel(11) and as source terms.  One thread per cell or
// node.  Reference routines restated here (Getrhslagrz.f90 unless noted), in
// the order of a stage of getrhsdg_lag_mcnew (Getrhslagmc.f90:6846-6971,
// :7280-7387):
//   barthlimit_lagsym_{tria,quad}          the planar limiter (lag2d_limit.h)
//   getgeoel_lagmc_rzawrev    :4590-4998   mass centre with the limited density
//   getamatr_lagmc_rzawrev    :3633-4221   its P1 mass matrix / geoel(11)
//   getndvelo_lagmc_rzaw      :822-1100    nodal solve
//   getriem_{tria,quad}rzaw   :1104-1669   corner Riemann coefficients
//   getbc_lagmaire_rzaw       GetBClag.f90:57-250   slip-wall forces
//   rhsifacedg_lagmc_{tria,quad}rzaw :1673-1947      face integrals
//   rhsdomndg_lagmc_{tria,quad}rzaw  :1951-2555      domain integrals
//   rhsdomnsrcdg_lagmc_{tria,quad}rzaw :2559-3096    source terms
//   rhsdomndg_lagmc_rzawdt    :7220-7342   motion of the basis with the
//   getgeoel_lagmc_rzawdt     :5613-5960   predicted mass centre, and
//   rhsdomndg_lagmc_{tria,quad}rzawdt :5002-5609
#include "lag2d_dev.h"
#include "lag2d_kernels.h"
#include "lag2d_limit.h"

namespace lag2d {

namespace {

#define UI(q, id) Uin[((q) * nd + (id)) * nc + ie]

// rutope 7-point rule of the reference triangle (dgflo2d.f90:2234-2255)
static constexpr double T7R[7] = {0.333333333333333, 0.059715871789770, 0.470142064105115, 0.470142064105115,
                                    0.797426985353087, 0.101286507323456, 0.101286507323456};
static constexpr double T7S[7] = {0.333333333333333, 0.470142064105115, 0.059715871789770, 0.470142064105115,
                                    0.101286507323456, 0.797426985353087, 0.101286507323456};
static constexpr double T7W[7] = {0.225000000000000, 0.132394152788506, 0.132394152788506, 0.132394152788506,
                                    0.125939180544827, 0.125939180544827, 0.125939180544827};
// ruqope 5x5 and 6x6 Gauss points in the reference's order (dgflo2d.f90:2646-2783)
static constexpr double G5P[5] = {-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                                    0.5384693101056830910363144, 0.9061798459386639927976269};
static constexpr double G5W[5] = {0.2369268850561890875142640, 0.4786286704993664680412915,
                                    0.5688888888888888888888889, 0.4786286704993664680412915,
                                    0.2369268850561890875142640};
static constexpr double G6P[6] = {0.2386191860831969086305017, -0.2386191860831969086305017,
                                    0.6612093864662645136613996, -0.6612093864662645136613996,
                                    -0.9324695142031520278123016, 0.9324695142031520278123016};
static constexpr double G6W[6] = {0.4679139345726910, 0.4679139345726910, 0.3607615730481386,
                                    0.3607615730481386, 0.1713244923791704, 0.1713244923791704};

// reference vertices: triangle (0,0), (1,0), (0,1); quad (-1,-1), (1,-1), (1,1), (-1,1)
template <int NV>
__device__ __forceinline__ double xvr(int iv)
{
    if (NV == 3) return iv == 1 ? 1.0 : 0.0;
    return (iv == 1 || iv == 2) ? 1.0 : -1.0;
}
template <int NV>
__device__ __forceinline__ double yvr(int iv)
{
    if (NV == 3) return iv == 2 ? 1.0 : 0.0;
    return iv >= 2 ? 1.0 : -1.0;
}

// Geometry points of the quadratic map with chord midpoints as mid-nodes:
// triangle nodes 1-3, midpoints 12, 23, 13; quad nodes 1-4, midpoints 12,
// 23, 34, 41 and the centre as the midpoint of 12 and 34.
template <int NV>
__device__ __forceinline__ void geo_points(const double *xp, const double *yp, double *gx, double *gy)
{
    for (int k = 0; k < NV; ++k) {
        gx[k] = xp[k];
        gy[k] = yp[k];
    }
    if (NV == 3) {
        gx[3] = 0.5 * (gx[0] + gx[1]); gy[3] = 0.5 * (gy[0] + gy[1]);
        gx[4] = 0.5 * (gx[1] + gx[2]); gy[4] = 0.5 * (gy[1] + gy[2]);
        gx[5] = 0.5 * (gx[0] + gx[2]); gy[5] = 0.5 * (gy[0] + gy[2]);
    } else {
        gx[4] = 0.5 * (gx[0] + gx[1]); gy[4] = 0.5 * (gy[0] + gy[1]);
        gx[5] = 0.5 * (gx[1] + gx[2]); gy[5] = 0.5 * (gy[1] + gy[2]);
        gx[6] = 0.5 * (gx[2] + gx[3]); gy[6] = 0.5 * (gy[2] + gy[3]);
        gx[7] = 0.5 * (gx[3] + gx[0]); gy[7] = 0.5 * (gy[3] + gy[0]);
        gx[8] = 0.5 * (gx[4] + gx[6]); gy[8] = 0.5 * (gy[4] + gy[6]);
    }
}

// Jacobian determinant of the quadratic map at (r, s), in the order of the
// geometry routines: dxdr*dyds - dydr*dxds (without the weight; the
// triangle's 0.5 is the caller's).
template <int NV>
__device__ __forceinline__ double quad_map_det(double r, double s, const double *gx, const double *gy)
{
    const double c10 = 1.0;
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    if (NV == 3) {
        const double dr_[6] = {c10 - 4.0 * (c10 - r - s), -1.0 + 4.0 * r, 0.0, 4.0 * (1.0 - 2.0 * r - s),
                               4.0 * s, -4.0 * s};
        const double ds_[6] = {c10 - 4.0 * (c10 - r - s), 0.0, -1.0 + 4.0 * s, -4.0 * r, 4.0 * r,
                               4.0 * (1.0 - r - 2.0 * s)};
        for (int k = 0; k < 6; ++k) {
            dxdr = dxdr + dr_[k] * gx[k];
            dxds = dxds + ds_[k] * gx[k];
            dydr = dydr + dr_[k] * gy[k];
            dyds = dyds + ds_[k] * gy[k];
        }
    } else {
        const double dr_[9] = {0.25 * (2.0 * r - c10) * s * (s - 1.0), 0.25 * (2.0 * r + c10) * s * (s - 1.0),
                               0.25 * (2.0 * r + c10) * s * (s + 1.0), 0.25 * (2.0 * r - c10) * s * (s + 1.0),
                               -r * s * (s - c10),                     0.5 * (2.0 * r + c10) * (c10 - s * s),
                               -r * s * (s + c10),                     0.5 * (2.0 * r - c10) * (1.0 - s * s),
                               -2.0 * r * (1.0 - s * s)};
        const double ds_[9] = {0.25 * (2.0 * s - c10) * r * (r - 1.0), 0.25 * (2.0 * s - c10) * r * (r + 1.0),
                               0.25 * (2.0 * s + c10) * r * (r + 1.0), 0.25 * (2.0 * s + c10) * r * (r - 1.0),
                               0.5 * (1.0 - r * r) * (2.0 * s - c10),  -r * s * (r + c10),
                               0.5 * (1.0 - r * r) * (2.0 * s + c10),  -r * s * (r - c10),
                               -2.0 * s * (1.0 - r * r)};
        for (int k = 0; k < 9; ++k) {
            dxdr = dxdr + dr_[k] * gx[k];
            dxds = dxds + ds_[k] * gx[k];
            dydr = dydr + dr_[k] * gy[k];
            dyds = dyds + ds_[k] * gy[k];
        }
    }
    return dxdr * dyds - dydr * dxds;
}

// Weighted area element of the straight (linear / bilinear) map at (r, s),
// in the order of the source-term routines: w * (dxdr*dyds - dxds*dydr)
// (triangle: 0.5 w * ...).  Also returns the four derivatives.
template <int NV>
__device__ __forceinline__ double lin_map(double r, double s, double w, const double *xp, const double *yp,
                                          double J[4])
{
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    if (NV == 3) {
        const double dspr[3] = {-1.0, 1.0, 0.0}, dsps[3] = {-1.0, 0.0, 1.0};
        for (int k = 0; k < 3; ++k) {
            dxdr = dxdr + dspr[k] * xp[k];
            dxds = dxds + dsps[k] * xp[k];
            dydr = dydr + dspr[k] * yp[k];
            dyds = dyds + dsps[k] * yp[k];
        }
    } else {
        const double rp = 1.0 + r, rm = 1.0 - r, sp = 1.0 + s, sm = 1.0 - s;
        const double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
        const double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
        for (int k = 0; k < 4; ++k) {
            dxdr = dxdr + dsr[k] * xp[k];
            dxds = dxds + dss[k] * xp[k];
            dydr = dydr + dsr[k] * yp[k];
            dyds = dyds + dss[k] * yp[k];
        }
    }
    J[0] = dxdr; J[1] = dxds; J[2] = dydr; J[3] = dyds;
    return NV == 3 ? 0.5 * w * (dxdr * dyds - dxds * dydr) : w * (dxdr * dyds - dxds * dydr);
}

// Quadrature points: triangles the 7-point rule, quads a tensor Gauss rule
// (n = 5 or 6) in the reference's point order.
template <int NV, int NG>
__device__ __forceinline__ void qpoint(int g, double &r, double &s, double &w)
{
    if (NV == 3) {
        r = T7R[g]; s = T7S[g]; w = T7W[g];
    } else if (NG == 5) {
        const int i = g / 5, j = g - 5 * i;
        r = G5P[i]; s = G5P[j]; w = G5W[i] * G5W[j];
    } else {
        const int i = g / 6, j = g - 6 * i;
        r = G6P[i]; s = G6P[j]; w = G6W[i] * G6W[j];
    }
}

// getgeoel_lagmc_rzawrev for one cell: five fixed-point passes for the mass
// centre geoel(1:2) with the limited density (the first with the mean), and
// the first-order centre geoel(15:16).  Then getamatr_lagmc_rzawrev: the P1
// mass matrix with the limited density about the new centre, inverted and
// divided by geoel(11).
template <int NV>
__device__ void rev_geometry(const Dev &d, const double *__restrict__ Uin, int ie, const double *xp, const double *yp,
                             double al0, double &rc, double &sc)
{
    const int nc = d.nc, nd = d.nd;
    constexpr int NGEO = NV == 3 ? 7 : 36, NMAS = NV == 3 ? 7 : 25;
    const double dr = NV == 3 ? 0.5 : 1.0, ds = dr;
    double gx[9], gy[9];
    geo_points<NV>(xp, yp, gx, gy);
    const double u00 = UI(0, 0), u01 = UI(0, 1), u02 = UI(0, 2);
    double xcel = 0.0, ycel = 0.0;
    for (int iloop = 0; iloop < 5; ++iloop) {
        double xcent = 0.0, ycent = 0.0, masel = 0.0;
        xcel = 0.0;
        ycel = 0.0;
        for (int g = 0; g < NGEO; ++g) {
            double r, s, wi;
            qpoint<NV, 6>(g, r, s, wi);
            const double det = quad_map_det<NV>(r, s, gx, gy);
            const double djaco = NV == 3 ? 0.5 * wi * det : wi * det;
            const double b1 = (r - rc) / dr, b2 = (s - sc) / ds;
            const double rhomc = u00;
            double rhomg = 0.0;
            rhomg = rhomg + u00 * 1.0;
            rhomg = rhomg + u01 * b1;
            rhomg = rhomg + u02 * b2;
            if (iloop == 0) rhomg = rhomc;
            rhomg = rhomc + al0 * (rhomg - rhomc);
            const double rhog = 1.0 / rhomg;
            masel = masel + 1.0 / rhomc * djaco;
            xcent = xcent + rhog * r * djaco;
            ycent = ycent + rhog * s * djaco;
            xcel = xcel + 1.0 / rhomc * r * djaco;
            ycel = ycel + 1.0 / rhomc * s * djaco;
        }
        rc = xcent / masel;
        sc = ycent / masel;
        xcel = xcel / masel;
        ycel = ycel / masel;
    }
    d.G[0 * nc + ie] = rc;
    d.G[1 * nc + ie] = sc;
    d.G[14 * nc + ie] = xcel;
    d.G[15 * nc + ie] = ycel;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
    for (int g = 0; g < NMAS; ++g) {
        double r, s, wi;
        qpoint<NV, 5>(g, r, s, wi);
        const double det = quad_map_det<NV>(r, s, gx, gy);
        const double djaco = NV == 3 ? 0.5 * wi * det : wi * det;
        const double xg = r, yg = s;
        const double b1 = (r - rc) / dr, b2 = (s - sc) / ds;
        const double rhomc = u00;
        double rhomg = 0.0;
        rhomg = rhomg + u00 * 1.0;
        rhomg = rhomg + u01 * b1;
        rhomg = rhomg + u02 * b2;
        rhomg = rhomc + al0 * (rhomg - rhomc);
        const double rhog = 1.0 / rhomg;
        f0 = f0 + rhog * djaco;
        f1 = f1 + rhog * (xg - rc) / dr * (xg - rc) / dr * djaco;
        f2 = f2 + rhog * (xg - rc) / dr * (yg - sc) / ds * djaco;
        f3 = f3 + rhog * (yg - sc) / ds * (yg - sc) / ds * djaco;
    }
    const double det = f1 * f3 - f2 * f2;
    const double g11 = d.G[10 * nc + ie];
    double *A = d.A;
    A[0 * nc + ie] = f3 / det / g11;
    A[1 * nc + ie] = -f2 / det / g11;
    A[2 * nc + ie] = f1 / det / g11;
    A[3 * nc + ie] = 1.0 / f0 / g11;
}

// Limiter, (with rev) the stage's mass centre and mass matrix, and the
// corner-face Riemann records of one cell, in the record layout of the
// planar path (isotropic corner matrices): m, m u_c (2), A sigma n (2).
template <int NV>
__device__ void limit_corner_rz(const Dev &d, const double *__restrict__ Uin, int ie, int rev)
{
    const int nc = d.nc, nd = d.nd;
    const double g1 = d.gamlg - 1.0;
    const double dr = NV == 3 ? 0.5 : 1.0, ds = dr;
    int ip[NV];
    double xp[NV], yp[NV];
    for (int k = 0; k < NV; ++k) {
        ip[k] = d.ipq[k * nc + ie];
        xp[k] = d.xs[ip[k]];
        yp[k] = d.ys[ip[k]];
    }
    double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    double bq[NV][3], un[NV][4];
    auto vertex_values = [&]() {
        for (int iv = 0; iv < NV; ++iv) {
            bq[iv][0] = 1.0;
            bq[iv][1] = (xvr<NV>(iv) - rc) / dr;
            bq[iv][2] = (yvr<NV>(iv) - sc) / ds;
            for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + UI(q, id) * bq[iv][id];
        }
    };
    vertex_values();
    double al[5], af[4];
    barth_limit<NV>(d, Uin, ie, ip, un, al, af);
    for (int q = 0; q < 5; ++q) d.lim[q * nc + ie] = al[q];
    for (int k = 0; k < 4; ++k) d.lim[(5 + k) * nc + ie] = af[k];
    if (rev) {
        rev_geometry<NV>(d, Uin, ie, xp, yp, al[0], rc, sc);
        vertex_values();
    }

    // getriem_{tria,quad}rzaw: eps 1e-8, corner areas L/2, impedance
    // rho c + rho slpdu * (0 |du.n|)
    const double eps = 1.0e-8;
    double gnx[NV], gny[NV], glen[NV];
    for (int e = 0; e < NV; ++e) {
        const int a = e, b = e + 1 == NV ? 0 : e + 1;
        const double anx = yp[b] - yp[a];
        const double any = -(xp[b] - xp[a]);
        gnx[e] = anx / sqrt(anx * anx + any * any);
        gny[e] = any / sqrt(anx * anx + any * any);
        glen[e] = sqrt(anx * anx + any * any);
    }
    const double uctr = UI(1, 0), vctr = UI(2, 0), ectr = UI(3, 0);
    const double rhoct = 1.0 / UI(0, 0);
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double dudr = af[0] * UI(1, 1) + af[2] * UI(2, 1);
    const double duds = af[0] * UI(1, 2) + af[2] * UI(2, 2);
    const double dvdr = af[1] * UI(1, 1) + af[3] * UI(2, 1);
    const double dvds = af[1] * UI(1, 2) + af[3] * UI(2, 2);
    double uvx[NV], vvx[NV], pvx[NV], aj0[NV], aj1[NV], aj2[NV];
    for (int iv = 0; iv < NV; ++iv) {
        const double rhovt = 1.0 / un[iv][0];
        const double pv =
            dmax(eps, g1 * rhovt * (un[iv][3] - 0.5 * (un[iv][1] * un[iv][1] + un[iv][2] * un[iv][2])));
        uvx[iv] = uctr + dudr * bq[iv][1] + duds * bq[iv][2];
        vvx[iv] = vctr + dvdr * bq[iv][1] + dvds * bq[iv][2];
        pvx[iv] = pctr + al[3] * (pv - pctr);
        const int p = ip[iv];
        const double acnx = d.usx[p] - uvx[iv], acny = d.usy[p] - vvx[iv];
        if (sqrt(acnx * acnx + acny * acny) < (double)1.e-11f) {
            aj0[iv] = (double)1.e-11f;
            aj1[iv] = (double)1.e-11f;
        } else {
            aj0[iv] = acnx / sqrt(acnx * acnx + acny * acny);
            aj1[iv] = acny / sqrt(acnx * acnx + acny * acny);
        }
        aj2[iv] = sqrt(acnx * acnx + acny * acny);
    }
    const double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    for (int iv = 0; iv < NV; ++iv) {
        const double a3 = aj2[iv] / sdctr;
        const int p = ip[iv];
        const double dux = d.usx[p] - uvx[iv], duy = d.usy[p] - vvx[iv];
        for (int f = 0; f < 2; ++f) {
            // corner faces: the edge entering vertex iv, then the edge leaving it
            const int e = f == 0 ? (iv == 0 ? NV - 1 : iv - 1) : iv;
            const double nx = gnx[e], ny = gny[e], area = 0.5 * glen[e];
            const double deltu = 0.0 * fabs(dux * nx + duy * ny);
            const double mur = rhoct * sdctr + rhoct * d.slpdu * deltu;
            const double vnau = (a3 < 1.e12) ? 1.0 : nx * aj0[iv] + ny * aj1[iv];
            const double m11 = mur * area * fabs(vnau);
            const int cf = f + 2 * iv;
            d.rec[(0 * 8 + cf) * nc + ie] = m11;
            d.rec[(1 * 8 + cf) * nc + ie] = m11 * uvx[iv] + 0.0 * vvx[iv];
            d.rec[(2 * 8 + cf) * nc + ie] = 0.0 * uvx[iv] + m11 * vvx[iv];
            d.rec[(3 * 8 + cf) * nc + ie] = -pvx[iv] * area * nx + 0.0 * area * ny;
            d.rec[(4 * 8 + cf) * nc + ie] = 0.0 * area * nx + -pvx[iv] * area * ny;
        }
    }
}

// Face, domain and source integrals, in-place limiting, the basis-motion
// term and (with apply_update) the mass-matrix solve and RK blend of one cell.
template <int NV>
__device__ void cellrhs_rz(const Dev &d, const double *__restrict__ Uin, double alfa, int apply_update, double epsaw,
                           int ie)
{
    const int nc = d.nc, nd = d.nd;
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    const double g1 = d.gamlg - 1.0, dr = NV == 3 ? 0.5 : 1.0, ds = dr;
    constexpr int NGD = NV == 3 ? 7 : 25;
    int ip[NV];
    double xp[NV], yp[NV], us[NV], vs[NV];
    for (int k = 0; k < NV; ++k) {
        ip[k] = d.ipq[k * nc + ie];
        xp[k] = d.xs[ip[k]];
        yp[k] = d.ys[ip[k]];
        us[k] = d.usx[ip[k]];
        vs[k] = d.usy[ip[k]];
    }
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    const double g11 = d.G[10 * nc + ie];
    double gnx[NV], gny[NV], glen[NV];
    for (int e = 0; e < NV; ++e) {
        const int a = e, b = e + 1 == NV ? 0 : e + 1;
        const double anx = yp[b] - yp[a];
        const double any = -(xp[b] - xp[a]);
        gnx[e] = anx / sqrt(anx * anx + any * any);
        gny[e] = any / sqrt(anx * anx + any * any);
        glen[e] = sqrt(anx * anx + any * any);
    }
    double bv[NV][3];
    for (int iv = 0; iv < NV; ++iv) {
        bv[iv][0] = 1.0;
        bv[iv][1] = (xvr<NV>(iv) - rc) / dr;
        bv[iv][2] = (yvr<NV>(iv) - sc) / ds;
    }
    // face integral with consistent corner weights (2 b_i + b_j)
    double r[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) r[q][id] = 0.0;
    for (int iv = 0; iv < NV; ++iv) {
        double f[2][2];
        for (int fa = 0; fa < 2; ++fa) {
            const int cf = fa + 2 * iv;
            const double m11 = d.rec[(0 * 8 + cf) * nc + ie];
            const double mu1 = d.rec[(1 * 8 + cf) * nc + ie], mu2 = d.rec[(2 * 8 + cf) * nc + ie];
            const double sn1 = d.rec[(3 * 8 + cf) * nc + ie], sn2 = d.rec[(4 * 8 + cf) * nc + ie];
            f[fa][0] = sn1 + m11 * us[iv] + 0.0 * vs[iv] - mu1;
            f[fa][1] = sn2 + 0.0 * us[iv] + m11 * vs[iv] - mu2;
        }
        const int ivm = iv == 0 ? NV - 1 : iv - 1, ivp = iv + 1 == NV ? 0 : iv + 1;
        const int e0 = ivm, e1 = iv;  // edges entering and leaving the vertex
        for (int id = 0; id < 3; ++id) {
            const double w0 = 2.0 * bv[iv][id] + bv[ivm][id];
            const double w1 = 2.0 * bv[iv][id] + bv[ivp][id];
            const double l10 = c16 * w0 * gnx[e0] * glen[e0], l20 = c16 * w0 * gny[e0] * glen[e0];
            const double l11 = c16 * w1 * gnx[e1] * glen[e1], l21 = c16 * w1 * gny[e1] * glen[e1];
            r[0][id] = r[0][id] + us[iv] * l10 + vs[iv] * l20 + us[iv] * l11 + vs[iv] * l21;
            r[1][id] = r[1][id] + f[0][0] * c13 * w0 + f[1][0] * c13 * w1;
            r[2][id] = r[2][id] + f[0][1] * c13 * w0 + f[1][1] * c13 * w1;
            r[3][id] = r[3][id] + us[iv] * f[0][0] * c13 * w0 + vs[iv] * f[0][1] * c13 * w0 +
                       us[iv] * f[1][0] * c13 * w1 + vs[iv] * f[1][1] * c13 * w1;
        }
    }
    const double al0 = d.lim[0 * nc + ie], al3 = d.lim[3 * nc + ie], al4 = d.lim[4 * nc + ie];
    const double af0 = d.lim[5 * nc + ie], af1 = d.lim[6 * nc + ie], af2 = d.lim[7 * nc + ie],
                 af3 = d.lim[8 * nc + ie];
    double u0[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) u0[q][id] = UI(q, id);
    const double eps = 1.0e-6;
    const double rhomc = u0[0][0], rhoct = 1.0 / rhomc;
    const double uctr = u0[1][0], vctr = u0[2][0], ectr = u0[3][0];
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double dudr = af0 * u0[1][1] + af2 * u0[2][1];
    const double duds = af0 * u0[1][2] + af2 * u0[2][2];
    const double dvdr = af1 * u0[1][1] + af3 * u0[2][1];
    const double dvds = af1 * u0[1][2] + af3 * u0[2][2];
    const double dbdr[3] = {0.0, 1.0 / dr, 0.0}, dbds[3] = {0.0, 0.0, 1.0 / ds};
    // domain integral -int grad(b) . F dA in the planar form (no radius)
    for (int g = 0; g < NGD; ++g) {
        double rr, ss, wi, J[4];
        qpoint<NV, 5>(g, rr, ss, wi);
        lin_map<NV>(rr, ss, wi, xp, yp, J);
        const double djaco = NV == 3 ? 0.5 * wi : wi;
        const double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
        double gd0[3], gd1[3];
        for (int id = 0; id < 3; ++id) {
            gd0[id] = dyds * dbdr[id] + (-dydr) * dbds[id];
            gd1[id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
        }
        const double b[3] = {1.0, (rr - rc) / dr, (ss - sc) / ds};
        double un[4] = {0.0, 0.0, 0.0, 0.0};
        for (int id = 0; id < 3; ++id)
            for (int q = 0; q < 4; ++q) un[q] = un[q] + u0[q][id] * b[id];
        const double rhoad = 1.0 / un[0];
        double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
        const double uadv = uctr + dudr * b[1] + duds * b[2];
        const double vadv = vctr + dvdr * b[1] + dvds * b[2];
        pres = pctr + al3 * (pres - pctr);
        for (int id = 0; id < 3; ++id) {
            const double fl0 = gd0[id] * uadv + gd1[id] * vadv;
            const double fl1 = gd0[id] * (-pres);
            const double fl2 = gd1[id] * (-pres);
            const double fl3 = (gd0[id] * uadv + gd1[id] * vadv) * (-pres);
            r[0][id] = r[0][id] - fl0 * djaco;
            r[1][id] = r[1][id] - fl1 * djaco;
            r[2][id] = r[2][id] - fl2 * djaco;
            r[3][id] = r[3][id] - fl3 * djaco;
        }
    }
    // radius weighting and the source terms int v b dA (tau) and
    // -int v p b dA (E) with the cell-average velocity
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) r[q][id] = g11 * r[q][id];
    for (int g = 0; g < NGD; ++g) {
        double rr, ss, wi, J[4];
        qpoint<NV, 5>(g, rr, ss, wi);
        const double djaco = lin_map<NV>(rr, ss, wi, xp, yp, J);
        const double b[3] = {1.0, (rr - rc) / dr, (ss - sc) / ds};
        double un[4] = {0.0, 0.0, 0.0, 0.0};
        for (int id = 0; id < 3; ++id)
            for (int q = 0; q < 4; ++q) un[q] = un[q] + u0[q][id] * b[id];
        const double rhoad = 1.0 / un[0];
        double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
        const double vadv = vctr;
        pres = pctr + al3 * (pres - pctr);
        for (int id = 0; id < 3; ++id) {
            r[0][id] = r[0][id] + vadv * b[id] * djaco;
            r[1][id] = r[1][id] + 0.0 * djaco;
            r[2][id] = r[2][id] + 0.0 * djaco;
            r[3][id] = r[3][id] + vadv * (-pres) * b[id] * djaco;
        }
    }
    // in-place limiting of the stage DOFs (Getrhslagmc.f90:7283-7296)
    double ul[4][3];
    for (int q = 0; q < 4; ++q) ul[q][0] = u0[q][0];
    ul[0][1] = u0[0][1] * al0;
    ul[0][2] = u0[0][2] * al0;
    ul[1][1] = dudr; ul[1][2] = duds;
    ul[2][1] = dvdr; ul[2][2] = dvds;
    ul[3][1] = u0[3][1] * al4;
    ul[3][2] = u0[3][2] * al4;

    // Basis motion.  The predicted mean specific volume
    // tau_a = tau + epsaw (M^-1 dt R)_tau, and nodes x + dt epsaw u*, give
    // the mass centre geoel(13:14) a time epsaw dt ahead (the reference
    // predicts twice, from the limited and the unlimited state, whose means
    // agree); d b / dt follows from it and the first-order centre geoel(15:16).
    const double dt = d.scal[0];
    const double A3 = d.A[3 * nc + ie];
    double taua;
    {
        const double m0[3] = {A3, 0.0, 0.0};
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s = s + m0[k] * (dt * r[0][k]);
        taua = ul[0][0] + s * epsaw;
    }
    double rcdt, scdt;
    {
        double xa[NV], ya[NV];
        for (int k = 0; k < NV; ++k) {
            xa[k] = xp[k] + dt * epsaw * us[k];
            ya[k] = yp[k] + dt * epsaw * vs[k];
        }
        double gx[9], gy[9];
        geo_points<NV>(xa, ya, gx, gy);
        constexpr int NGEO = NV == 3 ? 7 : 36;
        double xcent = 0.0, ycent = 0.0, masel = 0.0;
        for (int g = 0; g < NGEO; ++g) {
            double rr, ss, wi;
            qpoint<NV, 6>(g, rr, ss, wi);
            const double det = quad_map_det<NV>(rr, ss, gx, gy);
            const double djaco = NV == 3 ? 0.5 * wi * det : wi * det;
            const double rhog = 1.0 / taua;
            masel = masel + 1.0 / taua * djaco;
            xcent = xcent + rhog * rr * djaco;
            ycent = ycent + rhog * ss * djaco;
        }
        rcdt = xcent / masel;
        scdt = ycent / masel;
        d.G[12 * nc + ie] = rcdt;
        d.G[13 * nc + ie] = scdt;
    }
    {
        const double rc1 = d.G[14 * nc + ie], sc1 = d.G[15 * nc + ie];
        const double dbdt1 = (rc1 - rcdt) / dr / (epsaw * dt);
        const double dbdt2 = (sc1 - scdt) / ds / (epsaw * dt);
        // the limited state, with the limiter applied once more as the
        // reference's routine does
        const double pdudr = af0 * ul[1][1] + af2 * ul[2][1];
        const double pduds = af0 * ul[1][2] + af2 * ul[2][2];
        const double pdvdr = af1 * ul[1][1] + af3 * ul[2][1];
        const double pdvds = af1 * ul[1][2] + af3 * ul[2][2];
        const double drhomdr = ul[0][1] * al0, drhomds = ul[0][2] * al0;
        const double dedr = ul[3][1] * al4, deds = ul[3][2] * al4;
        const double dq[4][2] = {{drhomdr, drhomds}, {pdudr, pduds}, {pdvdr, pdvds}, {dedr, deds}};
        double raw[4][3];
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) raw[q][id] = 0.0;
        for (int g = 0; g < NGD; ++g) {
            double rr, ss, wi, J[4];
            qpoint<NV, 5>(g, rr, ss, wi);
            const double djaco = lin_map<NV>(rr, ss, wi, xp, yp, J);
            const double b[3] = {1.0, (rr - rc) / dr, (ss - sc) / ds};
            double u0a = 0.0;
            for (int id = 0; id < 3; ++id) u0a = u0a + ul[0][id] * b[id];
            const double rhoma = ul[0][0] + al0 * (u0a - ul[0][0]);
            const double rhoad = 1.0 / rhoma;
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) {
                    const double fl = rhoad * b[id] * dq[q][0] * dbdt1 + rhoad * b[id] * dq[q][1] * dbdt2;
                    raw[q][id] = raw[q][id] - fl * djaco;
                }
        }
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) r[q][id] = r[q][id] + g11 * raw[q][id];
    }

    if (apply_update) {
        double m[3][3];
        m[0][0] = A3; m[0][1] = 0.0; m[0][2] = 0.0;
        m[1][0] = 0.0; m[1][1] = d.A[0 * nc + ie]; m[1][2] = d.A[1 * nc + ie];
        m[2][0] = 0.0; m[2][1] = d.A[1 * nc + ie]; m[2][2] = d.A[2 * nc + ie];
        double rs[4][3];
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) rs[q][id] = dt * r[q][id];
        for (int id = 0; id < 3; ++id) {
            double s4[4] = {0.0, 0.0, 0.0, 0.0};
            for (int k = 0; k < 3; ++k)
                for (int q = 0; q < 4; ++q) s4[q] = s4[q] + m[id][k] * rs[q][k];
            for (int q = 0; q < 4; ++q)
                d.Us[(q * nd + id) * nc + ie] =
                    alfa * d.Un[(q * nd + id) * nc + ie] + (1.0 - alfa) * (ul[q][id] + s4[q]);
        }
    } else {
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) {
                d.Us[(q * nd + id) * nc + ie] = ul[q][id];
                d.R[(q * nd + id) * nc + ie] = r[q][id];
            }
    }
}

#undef UI

}  // namespace

__global__ void __launch_bounds__(LAG2D_BLOCK) k_limit_corner_rz(Dev d, const double *__restrict__ Uin, int rev)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= d.nc) return;
    if (ie < d.ntria)
        limit_corner_rz<3>(d, Uin, ie, rev);
    else
        limit_corner_rz<4>(d, Uin, ie, rev);
}

// Nodal solve of getndvelo_lagmc_rzaw with the slip-wall force of
// getbc_lagmaire_rzaw: at a node of wall faces (code 22) the force
// (l . M^-1 rhs) / (l . M^-1 l) l, l the sum of the faces' L/2 normals
// (0, -1) for subcode 221 and (-1, 0) for 222, is taken off the right-hand
// side; then the wall components are zeroed as on the planar path.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_node_rz(Dev d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int nc = d.nc, np = d.np;
    double a11 = 0.0, a21 = 0.0, a12 = 0.0, a22 = 0.0, mu1 = 0.0, mu2 = 0.0, sn1 = 0.0, sn2 = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k];
        const int ie = e >> 2, iv = e & 3;
        for (int f = 0; f < 2; ++f) {
            const int cf = f + 2 * iv;
            const double m = d.rec[(0 * 8 + cf) * nc + ie];
            a11 = a11 + m;
            a21 = a21 + 0.0;
            a12 = a12 + 0.0;
            a22 = a22 + m;
            mu1 = mu1 + d.rec[(1 * 8 + cf) * nc + ie];
            mu2 = mu2 + d.rec[(2 * 8 + cf) * nc + ie];
            sn1 = sn1 + d.rec[(3 * 8 + cf) * nc + ie];
            sn2 = sn2 + d.rec[(4 * 8 + cf) * nc + ie];
        }
    }
    const double det = a11 * a22 - a21 * a12;
    const double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
    double fp1 = 0.0, fp2 = 0.0;
    const uint8_t fl = d.nflag[p];
    if (fl & 8) {
        double lnx = 0.0, lny = 0.0;
        for (int j = 0; j < 2; ++j) {
            const int w = d.wall[j * np + p];
            if (w < 0) continue;
            const int o = w >> 2, typ = w & 3;
            const double anx = d.ys[o] - d.ys[p], any = -(d.xs[o] - d.xs[p]);
            const double L = sqrt(anx * anx + any * any);
            double bsnx = 0.0, bsny = 0.0;
            if (typ == 1) {
                bsnx = 0.0 * L * 0.5;
                bsny = -1.0 * L * 0.5;
            } else if (typ == 2) {
                bsnx = -1.0 * L * 0.5;
                bsny = 0.0 * L * 0.5;
            }
            lnx = lnx + bsnx;
            lny = lny + bsny;
        }
        const double coefl = (i11 * lnx + i12 * lny) * lnx + (i21 * lnx + i22 * lny) * lny;
        const double rhsx = mu1 - sn1 - 0.0, rhsy = mu2 - sn2 - 0.0;
        const double coefr = (i11 * rhsx + i12 * rhsy) * lnx + (i21 * rhsx + i22 * rhsy) * lny;
        fp1 = coefr / dmax(1.e-16, coefl) * lnx;
        fp2 = coefr / dmax(1.e-16, coefl) * lny;
    }
    fp1 = fp1 + 0.0;
    fp2 = fp2 + 0.0;
    const double r1 = mu1 - sn1 - fp1, r2 = mu2 - sn2 - fp2;
    double ux = i11 * r1 + i12 * r2;
    double uy = i21 * r1 + i22 * r2;
    if (fl & 1) uy = 0.0;
    if (fl & 2) ux = 0.0;
    d.usx[p] = ux;
    d.usy[p] = uy;
}

__global__ void __launch_bounds__(LAG2D_BLOCK) k_cellrhs_rz(Dev d, const double *__restrict__ Uin, double alfa,
                                                            int apply_update, double epsaw)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= d.nc) return;
    if (ie < d.ntria)
        cellrhs_rz<3>(d, Uin, alfa, apply_update, epsaw, ie);
    else
        cellrhs_rz<4>(d, Uin, alfa, apply_update, epsaw, ie);
}

// Start of a step whose dt the problem fixes (Getrhslag.f90:1663-1737).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_schedule(Dev d, double dt)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= d.nrun) return;
    if (run_active(d, r)) d.scal[2 * r] = dt;
}

}  // namespace lag2d
