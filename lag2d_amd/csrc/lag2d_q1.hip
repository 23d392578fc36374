// Straight-quad P1 path (ncurv=0, nfint=1, nlimi=6): Barth limiter with the
// symmetry-preserving velocity limiter, nodal Riemann solver, lumped face
// integral and 5x5 Gauss domain integral.  One thread per cell or per node.
//
// Reference routines restated here, in order of a stage
// (Getrhslagmc.f90:6810-6939, :7283-7296):
//   getfnds_lag_mc_hybrid        Getrhslagmc.f90:2117-2262
//   barthlimit_lag_vtxunknew     Getrhslagmctria.f90:2373-2640
//   barthlimit_lagsym_quad       Limiterlag.f90:4-404
//   barthlimit_sympre_quadlb     Limiterlag.f90:8127-8692
//   getndvelo_lag_mc_matrixsym   Limiterlag.f90:408-704
//   getriem_quad_matrixsym       Limiterlag.f90:708-1033
//   getriecoef_matrixnew         Limiterlag.f90:1037-1076
//   rhsifacedg_lag_mc_hybridquad2 Getrhslagmc.f90:14468-14618
//   rhsdomndg_lag_mc_quadsym     Limiterlag.f90:1108-1490
// and for the triangles of a hybrid mesh (cells 0..ntria-1):
//   barthlimit_lagsym_tria       Getrhslagmctria.f90:422-842
//   barthlimit_sympre_trialb     Getrhslagmctria.f90:2644-3258
//   getriem_tria_matrixsym       Getrhslagmctria.f90:847-1204
//   rhsifacedg_lag_mc_hybridtria Getrhslagmc.f90:14006-14160
//   rhsdomndg_lag_mc_triasym     Getrhslagmctria.f90:1208-1510
#include "lag2d_dev.h"
#include "lag2d_kernels.h"
#include "lag2d_limit.h"

namespace lag2d {

__device__ __forceinline__ double xvq(int iv) { return (iv == 1 || iv == 2) ? 1.0 : -1.0; }
__device__ __forceinline__ double yvq(int iv) { return (iv >= 2) ? 1.0 : -1.0; }

#define UI(q, id) Uin[((q) * nd + (id)) * nc + ie]

// ---- triangles (cells 0..ntria-1 of a hybrid mesh)
// Reference triangle r, s in [0, 1] with vertices (0,0), (1,0), (0,1); the P1
// basis is scaled by dr = ds = 0.5 (Getrhslagmctria.f90:464-470).
__device__ __forceinline__ double xvt(int iv) { return iv == 1 ? 1.0 : 0.0; }
__device__ __forceinline__ double yvt(int iv) { return iv == 2 ? 1.0 : 0.0; }
// corner faces of vertex iv: (edge before, edge after), edge e joins nodes e
// and e+1; and the far end of each of those edges
static constexpr int TFE[3][2] = {{2, 0}, {0, 1}, {1, 2}};
static constexpr int TNB[3][2] = {{2, 1}, {0, 2}, {1, 0}};
// rutope 7-point area rule (dgflo2d.f90:2234-2255), its literal values
static constexpr double PT7R[7] = {0.333333333333333, 0.059715871789770, 0.470142064105115,
                                            0.470142064105115, 0.797426985353087, 0.101286507323456,
                                            0.101286507323456};
static constexpr double PT7S[7] = {0.333333333333333, 0.470142064105115, 0.059715871789770,
                                            0.470142064105115, 0.101286507323456, 0.797426985353087,
                                            0.101286507323456};
static constexpr double WT7[7] = {0.225000000000000, 0.132394152788506, 0.132394152788506,
                                           0.132394152788506, 0.125939180544827, 0.125939180544827,
                                           0.125939180544827};

// Limiter coefficients and corner-face Riemann records of one triangle:
// barthlimit_lagsym_tria (Getrhslagmctria.f90:422-842) with the velocity
// limiter of barthlimit_sympre_trialb (:2644-3258) in the frame of the
// cell-average velocity, and getriem_tria_matrixsym (:847-1204), whose
// impedance uses |u* - u_v| (not its normal component) at both faces of a
// corner.  Records as for quads, corner faces cf = f + 2*iv, iv = 0..2.
__device__ void limit_corner_tri(const Dev &d, const double *__restrict__ Uin, int ie)
{
    const int nc = d.nc, nd = d.nd, np = d.np;
    const double g1 = d.gamlg - 1.0, dr = 0.5, ds = 0.5;
    int ip[3];
    double xp[3], yp[3];
    for (int k = 0; k < 3; ++k) {
        ip[k] = d.ipq[k * nc + ie];
        xp[k] = d.xs[ip[k]];
        yp[k] = d.ys[ip[k]];
    }
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    double bt[3][3];
    for (int iv = 0; iv < 3; ++iv) {
        bt[iv][0] = 1.0;
        bt[iv][1] = (xvt(iv) - rc) / dr;
        bt[iv][2] = (yvt(iv) - sc) / ds;
    }
    double un[3][4];
    for (int iv = 0; iv < 3; ++iv) {
        for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
        for (int id = 0; id < 3; ++id)
            for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + UI(q, id) * bt[iv][id];
    }
    const double tau = UI(0, 0), uctr = UI(1, 0), vctr = UI(2, 0), ectr = UI(3, 0);
    const double rhoct = 1.0 / tau;
    double al[5], af[4];
    barth_limit<3>(d, Uin, ie, ip, un, al, af);
    for (int q = 0; q < 5; ++q) d.lim[q * nc + ie] = al[q];
    for (int k = 0; k < 4; ++k) d.lim[(5 + k) * nc + ie] = af[k];

    const double eps = 1.0e-6;
    double gnx[3], gny[3], glen[3];
    for (int e = 0; e < 3; ++e) {
        int a = e, b = e == 2 ? 0 : e + 1;
        double anx = yp[b] - yp[a];
        double any = -(xp[b] - xp[a]);
        gnx[e] = anx / sqrt(anx * anx + any * any);
        gny[e] = any / sqrt(anx * anx + any * any);
        glen[e] = sqrt(anx * anx + any * any);
    }
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double dudr = af[0] * UI(1, 1) + af[2] * UI(2, 1);
    const double duds = af[0] * UI(1, 2) + af[2] * UI(2, 2);
    const double dvdr = af[1] * UI(1, 1) + af[3] * UI(2, 1);
    const double dvds = af[1] * UI(1, 2) + af[3] * UI(2, 2);
    const double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    for (int iv = 0; iv < 3; ++iv) {
        double rhovt = 1.0 / un[iv][0];
        double pv = dmax(eps, g1 * rhovt * (un[iv][3] - 0.5 * (un[iv][1] * un[iv][1] + un[iv][2] * un[iv][2])));
        double uvx = uctr + dudr * bt[iv][1] + duds * bt[iv][2];
        double vvx = vctr + dvdr * bt[iv][1] + dvds * bt[iv][2];
        double pvx = pctr + al[3] * (pv - pctr);
        int p = ip[iv];
        double acnx = d.usx[p] - uvx, acny = d.usy[p] - vvx;
        double aj0, aj1;
        if (sqrt(acnx * acnx + acny * acny) < (double)1.e-11f) {
            aj0 = (double)1.e-11f;
            aj1 = (double)1.e-11f;
        } else {
            aj0 = acnx / sqrt(acnx * acnx + acny * acny);
            aj1 = acny / sqrt(acnx * acnx + acny * acny);
        }
        const double a3 = sqrt(acnx * acnx + acny * acny) / sdctr;
        const double deltu = sqrt(acnx * acnx + acny * acny);
        const double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
        for (int f = 0; f < 2; ++f) {
            int e = TFE[iv][f];
            double nx = gnx[e], ny = gny[e];
            // corner area of the face, planar rcoef = 1: (2*1 + 1)/3*0.5*L
            double area = (2.0 * 1.0 + 1.0) / 3.0 * 0.5 * glen[e];
            double vnau = (a3 < 1.e12) ? 1.0 : nx * aj0 + ny * aj1;
            double m11 = mur * area * fabs(vnau);
            double mu1 = m11 * uvx + 0.0 * vvx;
            double mu2 = 0.0 * uvx + m11 * vvx;
            double sn1 = -pvx * area * nx + 0.0 * area * ny;
            double sn2 = 0.0 * area * nx + -pvx * area * ny;
            int cf = f + 2 * iv;
            d.rec[(0 * 8 + cf) * nc + ie] = m11;
            d.rec[(1 * 8 + cf) * nc + ie] = mu1;
            d.rec[(2 * 8 + cf) * nc + ie] = mu2;
            d.rec[(3 * 8 + cf) * nc + ie] = sn1;
            d.rec[(4 * 8 + cf) * nc + ie] = sn2;
        }
    }
}

// Face integral with consistent corner weights (rhsifacedg_lag_mc_hybridtria,
// Getrhslagmc.f90:14006-14160), domain integral on the 7-point rule
// (rhsdomndg_lag_mc_triasym, Getrhslagmctria.f90:1208-1510), in-place
// limiting and, with apply_update, the mass-matrix solve and RK blend of one
// triangle.
__device__ void cellrhs_tri(const Dev &d, const double *__restrict__ Uin, double alfa, int apply_update, int ie)
{
    const int nc = d.nc, nd = d.nd;
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    const double g1 = d.gamlg - 1.0, dr = 0.5, ds = 0.5;
    int ip[3];
    double xp[3], yp[3], us[3], vs[3];
    for (int k = 0; k < 3; ++k) {
        ip[k] = d.ipq[k * nc + ie];
        xp[k] = d.xs[ip[k]];
        yp[k] = d.ys[ip[k]];
        us[k] = d.usx[ip[k]];
        vs[k] = d.usy[ip[k]];
    }
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    double gnx[3], gny[3], glen[3];
    for (int e = 0; e < 3; ++e) {
        int a = e, b = e == 2 ? 0 : e + 1;
        double anx = yp[b] - yp[a];
        double any = -(xp[b] - xp[a]);
        gnx[e] = anx / sqrt(anx * anx + any * any);
        gny[e] = any / sqrt(anx * anx + any * any);
        glen[e] = sqrt(anx * anx + any * any);
    }
    double bt[3][3];
    for (int iv = 0; iv < 3; ++iv) {
        bt[iv][0] = 1.0;
        bt[iv][1] = (xvt(iv) - rc) / dr;
        bt[iv][2] = (yvt(iv) - sc) / ds;
    }
    double r[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) r[q][id] = 0.0;
    for (int iv = 0; iv < 3; ++iv) {
        double f[2][2];  // [face][dim]
        for (int fa = 0; fa < 2; ++fa) {
            int cf = fa + 2 * iv;
            double m11 = d.rec[(0 * 8 + cf) * nc + ie];
            double mu1 = d.rec[(1 * 8 + cf) * nc + ie], mu2 = d.rec[(2 * 8 + cf) * nc + ie];
            double sn1 = d.rec[(3 * 8 + cf) * nc + ie], sn2 = d.rec[(4 * 8 + cf) * nc + ie];
            f[fa][0] = sn1 + m11 * us[iv] + 0.0 * vs[iv] - mu1;
            f[fa][1] = sn2 + 0.0 * us[iv] + m11 * vs[iv] - mu2;
        }
        const int e0 = TFE[iv][0], e1 = TFE[iv][1], j0 = TNB[iv][0], j1 = TNB[iv][1];
        for (int id = 0; id < 3; ++id) {
            double w0 = 2.0 * bt[iv][id] + bt[j0][id];
            double w1 = 2.0 * bt[iv][id] + bt[j1][id];
            double l10 = c16 * w0 * gnx[e0] * glen[e0], l20 = c16 * w0 * gny[e0] * glen[e0];
            double l11 = c16 * w1 * gnx[e1] * glen[e1], l21 = c16 * w1 * gny[e1] * glen[e1];
            r[0][id] = r[0][id] + us[iv] * l10 + vs[iv] * l20 + us[iv] * l11 + vs[iv] * l21;
            r[1][id] = r[1][id] + f[0][0] * c13 * w0 + f[1][0] * c13 * w1;
            r[2][id] = r[2][id] + f[0][1] * c13 * w0 + f[1][1] * c13 * w1;
            r[3][id] = r[3][id] + us[iv] * f[0][0] * c13 * w0 + vs[iv] * f[0][1] * c13 * w0 +
                       us[iv] * f[1][0] * c13 * w1 + vs[iv] * f[1][1] * c13 * w1;
        }
    }
    double al0 = d.lim[0 * nc + ie], al3 = d.lim[3 * nc + ie], al4 = d.lim[4 * nc + ie];
    double af0 = d.lim[5 * nc + ie], af1 = d.lim[6 * nc + ie], af2 = d.lim[7 * nc + ie], af3 = d.lim[8 * nc + ie];
    double u0[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) u0[q][id] = UI(q, id);
    const double eps = 1.0e-6;
    const double rhoct = 1.0 / u0[0][0];
    const double uctr = u0[1][0], vctr = u0[2][0], ectr = u0[3][0];
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double dudr = af0 * u0[1][1] + af2 * u0[2][1];
    const double duds = af0 * u0[1][2] + af2 * u0[2][2];
    const double dvdr = af1 * u0[1][1] + af3 * u0[2][1];
    const double dvds = af1 * u0[1][2] + af3 * u0[2][2];
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    {
        const double dspr[3] = {-1.0, 1.0, 0.0}, dsps[3] = {-1.0, 0.0, 1.0};
        for (int k = 0; k < 3; ++k) {
            dxdr = dxdr + dspr[k] * xp[k];
            dxds = dxds + dsps[k] * xp[k];
            dydr = dydr + dspr[k] * yp[k];
            dyds = dyds + dsps[k] * yp[k];
        }
    }
    const double dbdr[3] = {0.0, 1.0 / dr, 0.0}, dbds[3] = {0.0, 0.0, 1.0 / ds};
    double gd0[3], gd1[3];
    for (int id = 0; id < 3; ++id) {
        gd0[id] = dyds * dbdr[id] + (-dydr) * dbds[id];
        gd1[id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
    }
    for (int g = 0; g < 7; ++g) {
        const double rr = PT7R[g], ss = PT7S[g], djaco = 0.5 * WT7[g];
        // planar: rcoef = 1 - alfrz + alfrz*y = 1
        const double rcoef = 1.0 - 0.0 + 0.0 * 0.0;
        double b[3] = {1.0, (rr - rc) / dr, (ss - sc) / ds};
        double un[4] = {0.0, 0.0, 0.0, 0.0};
        for (int id = 0; id < 3; ++id)
            for (int q = 0; q < 4; ++q) un[q] = un[q] + u0[q][id] * b[id];
        double rhoad = 1.0 / un[0];
        double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
        double uadv = uctr + dudr * b[1] + duds * b[2];
        double vadv = vctr + dvdr * b[1] + dvds * b[2];
        pres = pctr + al3 * (pres - pctr);
        for (int id = 0; id < 3; ++id) {
            double f0 = gd0[id] * uadv + gd1[id] * vadv;
            double f1 = gd0[id] * (-pres);
            double f2 = gd1[id] * (-pres);
            double f3 = (gd0[id] * uadv + gd1[id] * vadv) * (-pres);
            r[0][id] = r[0][id] - f0 * djaco * rcoef;
            r[1][id] = r[1][id] - f1 * djaco * rcoef;
            r[2][id] = r[2][id] - f2 * djaco * rcoef;
            r[3][id] = r[3][id] - f3 * djaco * rcoef;
        }
    }
    double ul[4][3];
    for (int q = 0; q < 4; ++q) ul[q][0] = u0[q][0];
    ul[0][1] = u0[0][1] * al0;
    ul[0][2] = u0[0][2] * al0;
    ul[1][1] = dudr; ul[1][2] = duds;
    ul[2][1] = dvdr; ul[2][2] = dvds;
    ul[3][1] = u0[3][1] * al4;
    ul[3][2] = u0[3][2] * al4;
    if (apply_update) {
        const double dt = d.scal[0];
        double m[3][3];
        m[0][0] = d.A[3 * nc + ie]; m[0][1] = 0.0; m[0][2] = 0.0;
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

// Cell averages (tau, u, v, p, E) and the principal-strain frame of
// barthlimit_sympre_quadlb Part 1 (Limiterlag.f90:8262-8387).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_cellavg_q1(Dev d, const double *__restrict__ Uin)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc, nd = d.nd;
    if (ie >= nc) return;
    const double g1 = d.gamlg - 1.0;
    double tau = UI(0, 0), u = UI(1, 0), v = UI(2, 0), e = UI(3, 0);
    double rhoct = 1.0 / tau;
    double pctr = dmax(EPS_SP6, g1 * rhoct * (e - 0.5 * (u * u + v * v)));
    d.cavg[0 * nc + ie] = tau;
    d.cavg[1 * nc + ie] = u;
    d.cavg[2 * nc + ie] = v;
    d.cavg[3 * nc + ie] = pctr;
    d.cavg[4 * nc + ie] = e;
    if (ie < d.ntria) {
        // triangles: frame of the cell-average velocity
        // (barthlimit_sympre_trialb, Getrhslagmctria.f90:2810-2822)
        double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
        if (sqrt(u * u + v * v) > 1.e-10) {
            m0 = u / sqrt(u * u + v * v);
            m1 = v / sqrt(u * u + v * v);
            m2 = -v / sqrt(u * u + v * v);
            m3 = u / sqrt(u * u + v * v);
        }
        d.frame[0 * nc + ie] = m0;
        d.frame[1 * nc + ie] = m1;
        d.frame[2 * nc + ie] = m2;
        d.frame[3 * nc + ie] = m3;
        return;
    }

    double xp[4], yp[4];
    for (int k = 0; k < 4; ++k) {
        int p = d.ipq[k * nc + ie];
        xp[k] = d.xs[p];
        yp[k] = d.ys[p];
    }
    double r = d.G[0 * nc + ie], s = d.G[1 * nc + ie];
    double dudr = UI(1, 1) / 1.0, duds = UI(1, 2) / 1.0;
    double dvdr = UI(2, 1) / 1.0, dvds = UI(2, 2) / 1.0;
    double rp = 1.0 + r, rm = 1.0 - r, sp = 1.0 + s, sm = 1.0 - s;
    double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
    double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 4; ++k) {
        dxdr = dxdr + dsr[k] * xp[k];
        dxds = dxds + dss[k] * xp[k];
        dydr = dydr + dsr[k] * yp[k];
        dyds = dyds + dss[k] * yp[k];
    }
    double jacom = dxdr * dyds - dydr * dxds;
    double drdx = dyds / jacom, drdy = -dxds / jacom;
    double dsdx = -dydr / jacom, dsdy = dxdr / jacom;
    double dudx = dudr * drdx + duds * dsdx;
    double dudy = dudr * drdy + duds * dsdy;
    double dvdx = dvdr * drdx + dvds * dsdx;
    double dvdy = dvdr * drdy + dvds * dsdy;
    double ma = dudx, mb = 0.5 * (dudy + dvdx), mc = mb, md = dvdy;
    double mapt = ma + md, mapd = ma * md - mb * mc;
    double l1 = 0.5 * mapt + sqrt(0.25 * (mapt * mapt) - mapd);
    double l2 = 0.5 * mapt - sqrt(0.25 * (mapt * mapt) - mapd);
    double m0 = 0.0, m1 = 0.0, m2 = 0.0, m3 = 0.0;
    if (fabs(mc) > 1.e-8) {
        double n1 = sqrt((l1 - md) * (l1 - md) + mc * mc);
        double n2 = sqrt((l2 - md) * (l2 - md) + mc * mc);
        if (mc > 0.0) {
            m0 = (l1 - md) / n1; m1 = mc / n1; m2 = (l2 - md) / n2; m3 = mc / n2;
        } else {
            m2 = (l1 - md) / n1; m3 = mc / n1; m0 = (l2 - md) / n2; m1 = mc / n2;
        }
    }
    d.frame[0 * nc + ie] = m0;
    d.frame[1 * nc + ie] = m1;
    d.frame[2 * nc + ie] = m2;
    d.frame[3 * nc + ie] = m3;
}

// Node max/min of the cell averages over the surrounding cells
// (barthlimit_lag_vtxunknew, Getrhslagmctria.f90:2522-2636), as a CSR gather.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_nodeminmax(Dev d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int nc = d.nc;
    double mx[5], mn[5];
    for (int q = 0; q < 5; ++q) { mx[q] = -1.e10; mn[q] = 1.e10; }
    for (int k = d.esuv2[p]; k < d.esuv2[p + 1]; ++k) {
        int je = d.esuv1[k];
        for (int q = 0; q < 5; ++q) {
            double v = d.cavg[q * nc + je];
            if (v > mx[q]) mx[q] = v;
            if (v < mn[q]) mn[q] = v;
        }
    }
    for (int q = 0; q < 5; ++q) {
        d.nmm[q * d.np + p] = mx[q];
        d.nmm[(5 + q) * d.np + p] = mn[q];
    }
}

// Limiter coefficients (aflim, afvec) and the corner-face Riemann records of
// one quad.  Records per corner face cf = f + 2*iv (face f=0 is the edge
// entering vertex iv, f=1 the edge leaving it): m = mu*A (the isotropic
// corner matrix), m*u_c (2), A*sigma*n (2).
// With FROZEN the limiter coefficients and the impedance's |du.n| are not
// computed: the kernel reads them from d.lim and d.dlt, where a recorded
// pass left them (the adjoint's frozen forward; quads only).  Without
// FROZEN, a non-null d.dlt records |du.n| of every corner face.
template <int FROZEN>
__global__ void __launch_bounds__(LAG2D_BLOCK) k_limit_corner_q1(Dev d, const double *__restrict__ Uin)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc, nd = d.nd, np = d.np;
    if (ie >= nc) return;
    if (!FROZEN && ie < d.ntria) {
        limit_corner_tri(d, Uin, ie);
        return;
    }
    const double g1 = d.gamlg - 1.0;
    int ip[4];
    double xp[4], yp[4];
    for (int k = 0; k < 4; ++k) {
        ip[k] = d.ipq[k * nc + ie];
        xp[k] = d.xs[ip[k]];
        yp[k] = d.ys[ip[k]];
    }
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    double bq[4][3];
    for (int iv = 0; iv < 4; ++iv) {
        bq[iv][0] = 1.0;
        bq[iv][1] = (xvq(iv) - rc) / 1.0;
        bq[iv][2] = (yvq(iv) - sc) / 1.0;
    }
    double un[4][4];  // [iv][q] vertex values of the unlimited P1 polynomial
    for (int iv = 0; iv < 4; ++iv) {
        for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
        for (int id = 0; id < 3; ++id)
            for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + UI(q, id) * bq[iv][id];
    }
    const double tau = UI(0, 0), uctr = UI(1, 0), vctr = UI(2, 0), ectr = UI(3, 0);
    const double rhoct = 1.0 / tau;
    // barthlimit_lagsym_quad Parts 2-3 and barthlimit_sympre_quadlb Parts 4, 6
    double al[5], af[4];
    if (FROZEN) {
        for (int q = 0; q < 5; ++q) al[q] = d.lim[q * nc + ie];
        for (int k = 0; k < 4; ++k) af[k] = d.lim[(5 + k) * nc + ie];
    } else {
        barth_limit<4>(d, Uin, ie, ip, un, al, af);
        for (int q = 0; q < 5; ++q) d.lim[q * nc + ie] = al[q];
        for (int k = 0; k < 4; ++k) d.lim[(5 + k) * nc + ie] = af[k];
    }

    // Corner Riemann coefficients (getriem_quad_matrixsym, ndens=1, nlimi=6)
    const double eps = 1.0e-6;
    double gnx[4], gny[4], glen[4];
    for (int e = 0; e < 4; ++e) {
        int a = e, b = (e + 1) & 3;
        double anx = yp[b] - yp[a];
        double any = -(xp[b] - xp[a]);
        gnx[e] = anx / sqrt(anx * anx + any * any);
        gny[e] = any / sqrt(anx * anx + any * any);
        glen[e] = sqrt(anx * anx + any * any);
    }
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double dudr = af[0] * UI(1, 1) + af[2] * UI(2, 1);
    const double duds = af[0] * UI(1, 2) + af[2] * UI(2, 2);
    const double dvdr = af[1] * UI(1, 1) + af[3] * UI(2, 1);
    const double dvds = af[1] * UI(1, 2) + af[3] * UI(2, 2);
    const double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    for (int iv = 0; iv < 4; ++iv) {
        double rhovt = 1.0 / un[iv][0];
        double pv = dmax(eps, g1 * rhovt * (un[iv][3] - 0.5 * (un[iv][1] * un[iv][1] + un[iv][2] * un[iv][2])));
        double uvx = uctr + dudr * bq[iv][1] + duds * bq[iv][2];
        double vvx = vctr + dvdr * bq[iv][1] + dvds * bq[iv][2];
        double pvx = pctr + al[3] * (pv - pctr);
        int p = ip[iv];
        double dux = d.usx[p] - uvx, duy = d.usy[p] - vvx;
        double a3 = sqrt(dux * dux + duy * duy) / sdctr;
        for (int f = 0; f < 2; ++f) {
            int e = f == 0 ? ((iv + 3) & 3) : iv;
            double nx = gnx[e], ny = gny[e];
            // corner area of the face, planar rcoef = 1: (2*1 + 1)/3*0.5*L
            double area = (2.0 * 1.0 + 1.0) / 3.0 * 0.5 * glen[e];
            double deltu;
            if (FROZEN) {
                deltu = d.dlt[(f + 2 * iv) * nc + ie];
            } else {
                deltu = fabs(dux * nx + duy * ny);
                if (d.dlt) d.dlt[(f + 2 * iv) * nc + ie] = deltu;
            }
            double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
            double m11;
            if (a3 < 1.e12) {
                m11 = mur * area * fabs(1.0);
            } else {
                double nrm = sqrt(dux * dux + duy * duy);
                m11 = mur * area * fabs(nx * (-dux / nrm) + ny * (-duy / nrm));
            }
            double mu1 = m11 * uvx + 0.0 * vvx;
            double mu2 = 0.0 * uvx + m11 * vvx;
            double sn1 = -pvx * area * nx + 0.0 * area * ny;
            double sn2 = 0.0 * area * nx + -pvx * area * ny;
            int cf = f + 2 * iv;
            d.rec[(0 * 8 + cf) * nc + ie] = m11;
            d.rec[(1 * 8 + cf) * nc + ie] = mu1;
            d.rec[(2 * 8 + cf) * nc + ie] = mu2;
            d.rec[(3 * 8 + cf) * nc + ie] = sn1;
            d.rec[(4 * 8 + cf) * nc + ie] = sn2;
        }
    }
}

// Nodal solve: gather the corner-face records of the cells around each node
// in the reference's accumulation order (cell id ascending, then face), the
// 2x2 solve u* = M^-1 (sum M u_c - sum A sigma n) and the wall projection
// (Limiterlag.f90:526-651).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_node_q1(Dev d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int nc = d.nc;
    double a11 = 0.0, a21 = 0.0, a12 = 0.0, a22 = 0.0, mu1 = 0.0, mu2 = 0.0, sn1 = 0.0, sn2 = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        int e = d.n2c[k];
        int ie = e >> 2, iv = e & 3;
        for (int f = 0; f < 2; ++f) {
            int cf = f + 2 * iv;
            double m = d.rec[(0 * 8 + cf) * nc + ie];
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

// Face integral (lumped corner weights), domain integral and the in-place
// limiting of the stage DOFs of one quad.  With apply_update the mass-matrix
// solve and the SSP-RK blend follow at once (stages after the first, where dt
// is already known); otherwise R is stored for the dt reduction.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_cellrhs_q1(Dev d, const double *__restrict__ Uin, double alfa,
                                                            int apply_update)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc, nd = d.nd;
    if (ie >= nc) return;
    if (ie < d.ntria) {
        cellrhs_tri(d, Uin, alfa, apply_update, ie);
        return;
    }
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    const double g1 = d.gamlg - 1.0;
    int ip[4];
    double xp[4], yp[4], us[4], vs[4];
    for (int k = 0; k < 4; ++k) {
        ip[k] = d.ipq[k * nc + ie];
        xp[k] = d.xs[ip[k]];
        yp[k] = d.ys[ip[k]];
        us[k] = d.usx[ip[k]];
        vs[k] = d.usy[ip[k]];
    }
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    double gnx[4], gny[4], glen[4];
    for (int e = 0; e < 4; ++e) {
        int a = e, b = (e + 1) & 3;
        double anx = yp[b] - yp[a];
        double any = -(xp[b] - xp[a]);
        gnx[e] = anx / sqrt(anx * anx + any * any);
        gny[e] = any / sqrt(anx * anx + any * any);
        glen[e] = sqrt(anx * anx + any * any);
    }
    double r[4][3];  // [q][id]
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) r[q][id] = 0.0;
    for (int iv = 0; iv < 4; ++iv) {
        double b[3] = {1.0, (xvq(iv) - rc) / 1.0, (yvq(iv) - sc) / 1.0};
        double f[2][2];  // [face][dim]
        for (int fa = 0; fa < 2; ++fa) {
            int cf = fa + 2 * iv;
            double m11 = d.rec[(0 * 8 + cf) * nc + ie];
            double mu1 = d.rec[(1 * 8 + cf) * nc + ie], mu2 = d.rec[(2 * 8 + cf) * nc + ie];
            double sn1 = d.rec[(3 * 8 + cf) * nc + ie], sn2 = d.rec[(4 * 8 + cf) * nc + ie];
            f[fa][0] = sn1 + m11 * us[iv] + 0.0 * vs[iv] - mu1;
            f[fa][1] = sn2 + 0.0 * us[iv] + m11 * vs[iv] - mu2;
        }
        int e0 = (iv + 3) & 3, e1 = iv;
        for (int id = 0; id < 3; ++id) {
            double w = 2.0 * b[id] + b[id];
            double l11 = c16 * w * gnx[e0] * glen[e0], l21 = c16 * w * gny[e0] * glen[e0];
            double l12 = c16 * w * gnx[e1] * glen[e1], l22 = c16 * w * gny[e1] * glen[e1];
            r[0][id] = r[0][id] + us[iv] * l11 + vs[iv] * l21 + us[iv] * l12 + vs[iv] * l22;
            r[1][id] = r[1][id] + f[0][0] * c13 * w + f[1][0] * c13 * w;
            r[2][id] = r[2][id] + f[0][1] * c13 * w + f[1][1] * c13 * w;
            r[3][id] = r[3][id] + us[iv] * f[0][0] * c13 * w + vs[iv] * f[0][1] * c13 * w +
                       us[iv] * f[1][0] * c13 * w + vs[iv] * f[1][1] * c13 * w;
        }
    }
    // domain integral
    double al0 = d.lim[0 * nc + ie], al3 = d.lim[3 * nc + ie], al4 = d.lim[4 * nc + ie];
    double af0 = d.lim[5 * nc + ie], af1 = d.lim[6 * nc + ie], af2 = d.lim[7 * nc + ie], af3 = d.lim[8 * nc + ie];
    double u0[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) u0[q][id] = UI(q, id);
    const double eps = 1.0e-6;
    const double rhoct = 1.0 / u0[0][0];
    const double uctr = u0[1][0], vctr = u0[2][0], ectr = u0[3][0];
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double dudr = af0 * u0[1][1] + af2 * u0[2][1];
    const double duds = af0 * u0[1][2] + af2 * u0[2][2];
    const double dvdr = af1 * u0[1][1] + af3 * u0[2][1];
    const double dvds = af1 * u0[1][2] + af3 * u0[2][2];
    const double pg[5] = {-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                          0.5384693101056830910363144, 0.9061798459386639927976269};
    const double wg[5] = {0.2369268850561890875142640, 0.4786286704993664680412915, 0.5688888888888888888888889,
                          0.4786286704993664680412915, 0.2369268850561890875142640};
    for (int gi = 0; gi < 5; ++gi) {
        for (int gj = 0; gj < 5; ++gj) {
            double rr = pg[gi], ss = pg[gj], wi = wg[gi] * wg[gj];
            double rp = 1.0 + rr, rm = 1.0 - rr, sp = 1.0 + ss, sm = 1.0 - ss;
            double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
            double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
            double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
            for (int k = 0; k < 4; ++k) {
                dxdr = dxdr + dsr[k] * xp[k];
                dxds = dxds + dss[k] * xp[k];
                dydr = dydr + dsr[k] * yp[k];
                dyds = dyds + dss[k] * yp[k];
            }
            // planar: rcoef = 1 - alfrz + alfrz*y = 1
            const double rcoef = 1.0 - 0.0 + 0.0 * 0.0;
            double gd0[3] = {dyds * 0.0 + (-dydr) * 0.0, dyds * (1.0 / 1.0) + (-dydr) * 0.0,
                             dyds * 0.0 + (-dydr) * (1.0 / 1.0)};
            double gd1[3] = {(-dxds) * 0.0 + dxdr * 0.0, (-dxds) * (1.0 / 1.0) + dxdr * 0.0,
                             (-dxds) * 0.0 + dxdr * (1.0 / 1.0)};
            double b[3] = {1.0, (rr - rc) / 1.0, (ss - sc) / 1.0};
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + u0[q][id] * b[id];
            double rhoad = 1.0 / un[0];
            double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
            double uadv = uctr + dudr * b[1] + duds * b[2];
            double vadv = vctr + dvdr * b[1] + dvds * b[2];
            pres = pctr + al3 * (pres - pctr);
            double s11 = -pres + 0.0, s12 = 0.0 + 0.0, s21 = 0.0 + 0.0, s22 = -pres + 0.0;
            for (int id = 0; id < 3; ++id) {
                double f0 = gd0[id] * uadv + gd1[id] * vadv;
                double f1 = gd0[id] * s11 + gd1[id] * s12;
                double f2 = gd0[id] * s21 + gd1[id] * s22;
                double f3 = gd0[id] * (s11 * uadv + s12 * vadv) + gd1[id] * (s21 * uadv + s22 * vadv);
                r[0][id] = r[0][id] - f0 * wi * rcoef;
                r[1][id] = r[1][id] - f1 * wi * rcoef;
                r[2][id] = r[2][id] - f2 * wi * rcoef;
                r[3][id] = r[3][id] - f3 * wi * rcoef;
            }
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
    if (apply_update) {
        const double dt = d.scal[0];
        double m[3][3];
        m[0][0] = d.A[3 * nc + ie]; m[0][1] = 0.0; m[0][2] = 0.0;
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

template __global__ void k_limit_corner_q1<0>(Dev, const double *__restrict__);
template __global__ void k_limit_corner_q1<1>(Dev, const double *__restrict__);

}  // namespace lag2d
