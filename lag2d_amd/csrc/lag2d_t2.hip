// Curved six-node triangles of hybrid meshes on the curved P2 sub-cell path
// (ncurv=1, nfint=4, npoly=2; cells 0 .. ntria-1, the quads after them).
//
// Reference routines restated here (getrhsdg_lag_mcnew branch
// Getrhslagmc.f90:7556-7708, triangle calls :7584-7585, :7667, :7678):
//   getfnds_lagsmsif_hybrid, triangle part        Getrhslaggauss.f90:3419-3496
//   getfnds_lagsms_simpsonhybrid, triangle part   Getrhslaggauss.f90:9726-9822
//   getriem_triasubg3                             Getrhslaggauss.f90:12201-12590
//   getrhosubgt_average2                          Getrhslaggauss.f90:12595-12940
//   rhsifacedg_lagsubgt3                          Getrhslaggauss.f90:8297-8496
//   rhsdomndg_lagmc_curvtria                      Getrhslag.f90:4449-4711
// A triangle's sub-triangle corner states go to the same records as a
// quad's (RC2, corner 4 isg + ivsg), so the node gather takes both; its 15
// face slots use rows 0..44 of the face-geometry array.  Work split: a
// thread per triangle (a hybrid mesh's triangles are a fraction of its
// cells, and each thread's sums run in the reference's order).
#include "lag2d_dev.h"
#include "lag2d_kernels.h"
#include "lag2d_q2geom.h"

namespace lag2d {

#define UD6(buf, q, id, ie) (buf)[((q) * 6 + (id)) * nc + (ie)]

// quadratic triangle shape functions and their reference derivatives
__device__ __forceinline__ void shape6(double r, double s, double sh[6], double dr_[6], double ds_[6])
{
    const double c10 = 1.0;
    sh[0] = -(c10 - r - s) * (c10 - 2.0 * (c10 - r - s));
    sh[1] = -r * (c10 - 2.0 * r);
    sh[2] = -s * (c10 - 2.0 * s);
    sh[3] = 4.0 * r * (c10 - r - s);
    sh[4] = 4.0 * r * s;
    sh[5] = 4.0 * s * (c10 - r - s);
    dr_[0] = c10 - 4.0 * (c10 - r - s);
    dr_[1] = -1.0 + 4.0 * r;
    dr_[2] = 0.0;
    dr_[3] = 4.0 * (1.0 - 2.0 * r - s);
    dr_[4] = 4.0 * s;
    dr_[5] = -4.0 * s;
    ds_[0] = c10 - 4.0 * (c10 - r - s);
    ds_[1] = 0.0;
    ds_[2] = -1.0 + 4.0 * s;
    ds_[3] = -4.0 * r;
    ds_[4] = 4.0 * r;
    ds_[5] = 4.0 * (1.0 - r - 2.0 * s);
}

__device__ __forceinline__ void map6(const double *xp, const double *yp, double r, double s, double &x, double &y)
{
    double sh[6], dr_[6], ds_[6];
    shape6(r, s, sh, dr_, ds_);
    double a = 0.0, b = 0.0;
    for (int k = 0; k < 6; ++k) {
        a = a + sh[k] * xp[k];
        b = b + sh[k] * yp[k];
    }
    x = a;
    y = b;
}

__device__ __forceinline__ void jac6(const double *xp, const double *yp, double r, double s, double J[4])
{
    double sh[6], dr_[6], ds_[6];
    shape6(r, s, sh, dr_, ds_);
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 6; ++k) {
        dxdr = dxdr + dr_[k] * xp[k];
        dxds = dxds + ds_[k] * xp[k];
        dydr = dydr + dr_[k] * yp[k];
        dyds = dyds + ds_[k] * yp[k];
    }
    J[0] = dxdr; J[1] = dxds; J[2] = dydr; J[3] = dyds;
}

// P2 Taylor basis of a triangle about its mass centre, dr = ds = 0.5
__device__ __forceinline__ void basis_t6(double r, double s, double rc, double sc, double g0, double g1, double g2,
                                         double *b)
{
    b[0] = 1.0;
    b[1] = (r - rc) / 0.5;
    b[2] = (s - sc) / 0.5;
    b[3] = 0.5 * b[1] * b[1] - g0;
    b[4] = 0.5 * b[2] * b[2] - g1;
    b[5] = b[1] * b[2] - g2;
}

__device__ __forceinline__ void load6(const Dev &d, const double *x, const double *y, int ie, double *xp, double *yp)
{
    for (int k = 0; k < 6; ++k) {
        const int p = d.ipq[k * d.nc + ie];
        xp[k] = x[p];
        yp[k] = y[p];
    }
}

// points of the sub-triangles' quadratic maps (postc, iptsc of
// getrhosubgt_average2, Getrhslaggauss.f90:12644-12660)
static constexpr double PCT[9][2] = {{.25, 0.}, {.75, 0.}, {.75, .25}, {.25, .75}, {0., .75},
                                     {0., .25}, {.25, .25}, {.5, .25}, {.25, .5}};
static constexpr int IPTSC[4][3] = {{0, 6, 5}, {7, 8, 6}, {1, 2, 7}, {8, 3, 4}};

// Initial sub-triangle masses (getrhosubgt_average2, second half):
// time-invariant, computed once from the initial coordinates.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_t2_submass(Dev d)
{
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= d.ntria) return;
    const int nc = d.nc;
    double xi[6], yi[6];
    load6(d, d.x0, d.y0, it, xi, yi);
    double xc[9], yc[9];
    for (int k = 0; k < 9; ++k) map6(xi, yi, PCT[k][0], PCT[k][1], xc[k], yc[k]);
    // getrhoig_triacurv2 (Getrhslag.f90:17622-17757): the uniform density
    // of the problems this path takes, or the per-cell initial density
    const double rho = d.rhoi ? d.rhoi[it] : 1.0;
    for (int isg = 0; isg < 4; ++isg) {
        double xs[6], ys[6];
        for (int k = 0; k < 3; ++k) {
            xs[k] = xi[IPTSG[isg][k]];
            ys[k] = yi[IPTSG[isg][k]];
            xs[3 + k] = xc[IPTSC[isg][k]];
            ys[3 + k] = yc[IPTSC[isg][k]];
        }
        double ms = 0.0;
        for (int g = 0; g < 7; ++g) {
            double J[4];
            jac6(xs, ys, PT7T_R[g], PT7T_S[g], J);
            const double djaco = 0.5 * WT7T[g] * (J[0] * J[3] - J[2] * J[1]);
            ms = ms + rho * 1.0 * djaco;
        }
        d.msub[isg * nc + it] = ms;
    }
}

// Face geometry, sub-triangle densities and corner states of a triangle.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_t2_corner(Dev d)
{
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= d.ntria) return;
    const int nc = d.nc;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-15;  // getriem_triasubg3 clamps at 1.0d-15
    double xp[6], yp[6];
    load6(d, d.xs, d.ys, it, xp, yp);
    double g[15][3];
    {
        // interior sub-faces between mid-nodes, at both ends
        const double pc[3][2] = {{0.25, 0.25}, {0.5, 0.25}, {0.25, 0.5}};
        const int ipift[3][3] = {{3, 5, 0}, {4, 3, 1}, {5, 4, 2}};
        double xv[3], yv[3];
        for (int k = 0; k < 3; ++k) map6(xp, yp, pc[k][0], pc[k][1], xv[k], yv[k]);
        for (int f = 0; f < 3; ++f) {
            const double fx[3] = {xp[ipift[f][0]], xp[ipift[f][1]], xv[ipift[f][2]]};
            const double fy[3] = {yp[ipift[f][0]], yp[ipift[f][1]], yv[ipift[f][2]]};
            for (int e = 0; e < 2; ++e) {
                const double r = e == 0 ? -1.0 : 1.0;
                const double dd[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
                double dxdr = 0.0, dydr = 0.0;
                for (int i = 0; i < 3; ++i) {
                    dxdr = dxdr + dd[i] * fx[i];
                    dydr = dydr + dd[i] * fy[i];
                }
                const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
                const double anx = -dydr, any = dxdr;
                const int sl = (e == 0 ? 6 : 12) + f;
                g[sl][0] = anx / sqrt(anx * anx + any * any);
                g[sl][1] = any / sqrt(anx * anx + any * any);
                g[sl][2] = djaco * 2.0;
            }
        }
        // boundary-edge points: cofactor of the map's Jacobian
        const double pt[9][2] = {{0, 0}, {1, 0}, {1, 0}, {0, 1}, {0, 1}, {0, 0}, {.5, 0}, {.5, .5}, {0, .5}};
        const int slot[9] = {0, 1, 2, 3, 4, 5, 9, 10, 11};
        for (int ig = 0; ig < 9; ++ig) {
            double J[4];
            jac6(xp, yp, pt[ig][0], pt[ig][1], J);
            const double c11 = J[3], c12 = -J[2], c21 = -J[1], c22 = J[0];
            double v1, v2, larea;
            if (ig == 0 || ig == 1 || ig == 6) {
                v1 = 0.0; v2 = -1.0; larea = 1.0;
            } else if (ig == 2 || ig == 3 || ig == 7) {
                v1 = sqrt(2.0) * 0.5; v2 = sqrt(2.0) * 0.5; larea = sqrt(2.0);
            } else {
                v1 = -1.0; v2 = 0.0; larea = 1.0;
            }
            const double anx = c11 * v1 + c12 * v2;
            const double any = c21 * v1 + c22 * v2;
            const double w1 = anx * larea, w2 = any * larea;
            const double farea = sqrt(w1 * w1 + w2 * w2);
            g[slot[ig]][0] = anx * larea / farea;
            g[slot[ig]][1] = any * larea / farea;
            g[slot[ig]][2] = farea;
        }
        for (int sl = 0; sl < 15; ++sl)
            for (int k = 0; k < 3; ++k) d.gq[(3 * sl + k) * nc + it] = g[sl][k];
    }
    double u[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) u[q][id] = UD6(d.Ui, q, id, it);
    const double rc = d.G[0 * nc + it], sc = d.G[1 * nc + it];
    const double gg0 = d.G[18 * nc + it], gg1 = d.G[19 * nc + it], gg2 = d.G[20 * nc + it];
    const double rhomc = u[0][0], uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
    const double rhoct = 1.0 / rhomc;
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    // sub-triangle densities: initial mass / current area, and the
    // mass-weighted mean of tau, both in point order
    double dens[4], tau[4];
    {
        double xc[9], yc[9];
        for (int k = 0; k < 9; ++k) map6(xp, yp, PCT[k][0], PCT[k][1], xc[k], yc[k]);
        for (int isg = 0; isg < 4; ++isg) {
            double xs[6], ys[6];
            for (int k = 0; k < 3; ++k) {
                xs[k] = xp[IPTSG[isg][k]];
                ys[k] = yp[IPTSG[isg][k]];
                xs[3 + k] = xc[IPTSC[isg][k]];
                ys[3 + k] = yc[IPTSC[isg][k]];
            }
            double a = 0.0, masel = 0.0, volel = 0.0;
            for (int gp = 0; gp < 7; ++gp) {
                const double r = PT7T_R[gp], s = PT7T_S[gp], wi = WT7T[gp];
                const double shs[3] = {1.0 - r - s, r, s};
                double rg = 0.0, sg = 0.0;
                for (int k = 0; k < 3; ++k) {
                    rg = rg + shs[k] * XV6[IPTSG[isg][k]];
                    sg = sg + shs[k] * YV6[IPTSG[isg][k]];
                }
                double b[6];
                basis_t6(rg, sg, rc, sc, gg0, gg1, gg2, b);
                double t = 0.0;
                for (int id = 0; id < 6; ++id) t = t + u[0][id] * b[id];
                t = rhomc + 1.0 * (t - rhomc);
                double J[4];
                jac6(xs, ys, r, s, J);
                const double djaco = 0.5 * wi * (J[0] * J[3] - J[2] * J[1]);
                a = a + djaco;
                masel = masel + 1.0 / t * djaco;
                volel = volel + djaco;
            }
            dens[isg] = 0.0 + (1.0 / a) * d.msub[isg * nc + it];
            tau[isg] = volel / masel;
        }
    }
    double un[6][4];
    for (int iv = 0; iv < 6; ++iv) {
        double b[6];
        basis_t6(XV6[iv], YV6[iv], rc, sc, gg0, gg1, gg2, b);
        for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
        for (int id = 0; id < 6; ++id)
            for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + u[q][id] * b[id];
    }
    for (int isg = 0; isg < 4; ++isg)
        for (int iv = 0; iv < 3; ++iv) {
            const int ln = IPTSG[isg][iv];
            const int p = d.ipq[ln * nc + it];
            const double rhovt = 1.0 / un[ln][0];
            const double rhovsg = rhovt + 1.0 * (dens[isg] - 1.0 / tau[isg]);
            const double uv = un[ln][1], vv = un[ln][2], ev = un[ln][3];
            const double pv = dmax(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)));
            const double dux = d.usx[p] - uv, duy = d.usy[p] - vv;
            const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
            const int k = 4 * isg + iv;
            d.rec[RC2(it, k, 2)] = uv;
            d.rec[RC2(it, k, 3)] = vv;
            d.rec[RC2(it, k, 4)] = pv;
            for (int f = 0; f < 2; ++f) {
                const int tt = FNTSG[isg][2 * iv + f];
                const int sl = (tt > 0 ? tt : -tt) - 1;
                const double sg = tt > 0 ? 1.0 : -1.0;
                const double nx = sg * g[sl][0], ny = sg * g[sl][1];
                const double area = g[sl][2] * WT2[isg][iv][f];
                // the triangle corners carry no shock term (:12505)
                const double deltu = 0.0 * fabs(dux * nx + duy * ny);
                const double mur = rhoct * sdctr + rhoct * d.slpdu * deltu;
                double m11;
                if (a3 < 1.e12) {
                    m11 = mur * area * fabs(1.0);
                } else {
                    const double nrm = sqrt(dux * dux + duy * duy);
                    m11 = mur * area * fabs(nx * (dux / nrm) + ny * (duy / nrm));
                }
                d.rec[RC2(it, k, f)] = m11;
            }
        }
}

// Face integral (momentum/energy weights 1, volume weights wfgst), the
// 7-point domain integral, the identity derivative update and, with
// apply_update, the mass-matrix solve and RK blend of a triangle.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_t2_cellrhs(Dev d, double alfa, int apply_update)
{
    const int it = blockIdx.x * blockDim.x + threadIdx.x;
    if (it >= d.ntria) return;
    const int nc = d.nc;
    const double g1 = d.gamlg - 1.0;
    const double rc = d.G[0 * nc + it], sc = d.G[1 * nc + it];
    const double gg0 = d.G[18 * nc + it], gg1 = d.G[19 * nc + it], gg2 = d.G[20 * nc + it];
    double r[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) r[q][id] = 0.0;
    for (int isg = 0; isg < 4; ++isg)
        for (int iv = 0; iv < 3; ++iv) {
            const int k = 4 * isg + iv;
            const int ln = IPTSG[isg][iv];
            const int p = d.ipq[ln * nc + it];
            const double us1 = d.usx[p], us2 = d.usy[p];
            const double uv = d.rec[RC2(it, k, 2)], vv = d.rec[RC2(it, k, 3)], pv = d.rec[RC2(it, k, 4)];
            double nx[2], ny[2], ar[2], fx[2], fy[2];
            for (int f = 0; f < 2; ++f) {
                double ga, m1, m2, s1, s2;
                tri_face_geom(d.gq, nc, it, isg, iv, f, nx[f], ny[f], ga);
                ar[f] = ga * WT2[isg][iv][f];
                const double m11 = d.rec[RC2(it, k, f)];
                corner_face_rec(m11, uv, vv, pv, nx[f], ny[f], ga * WT2[isg][iv][f], m1, m2, s1, s2);
                fx[f] = s1 + m11 * us1 + 0.0 * us2 - m1;
                fy[f] = s2 + 0.0 * us1 + m11 * us2 - m2;
            }
            double bq[6];
            basis_t6(XV6[ln], YV6[ln], rc, sc, gg0, gg1, gg2, bq);
            for (int id = 0; id < 6; ++id) {
                const double b = bq[id];
                r[0][id] = r[0][id] + us1 * nx[0] * ar[0] * b + us2 * ny[0] * ar[0] * b + us1 * nx[1] * ar[1] * b +
                           us2 * ny[1] * ar[1] * b;
                r[1][id] = r[1][id] + fx[0] * b * 1.0 + fx[1] * b * 1.0;
                r[2][id] = r[2][id] + fy[0] * b * 1.0 + fy[1] * b * 1.0;
                r[3][id] = r[3][id] + us1 * fx[0] * b * 1.0 + us2 * fy[0] * b * 1.0 + us1 * fx[1] * b * 1.0 +
                           us2 * fy[1] * b * 1.0;
            }
        }
    double ul[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) ul[q][id] = UD6(d.Ui, q, id, it);
    {
        double xp[6], yp[6];
        load6(d, d.xs, d.ys, it, xp, yp);
        const double dr = 0.5, ds = 0.5;
        for (int gp = 0; gp < 7; ++gp) {
            const double rr = PT7T_R[gp], ss = PT7T_S[gp], wi = WT7T[gp];
            double J[4];
            jac6(xp, yp, rr, ss, J);
            const double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
            const double djaco = 0.5 * wi;
            const double dbdr[6] = {0.0, 1.0 / dr, 0.0, (rr - rc) / (dr * dr), 0.0, (ss - sc) / dr / ds};
            const double dbds[6] = {0.0, 0.0, 1.0 / ds, 0.0, (ss - sc) / (ds * ds), (rr - rc) / dr / ds};
            double b[6];
            basis_t6(rr, ss, rc, sc, gg0, gg1, gg2, b);
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 6; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + ul[q][id] * b[id];
            const double rhoad = 1.0 / un[0];
            const double uu = un[1], vv = un[2];
            const double pres = g1 * rhoad * (un[3] - 0.5 * (uu * uu + vv * vv));  // not clamped here
            for (int id = 0; id < 6; ++id) {
                const double a = dyds * dbdr[id] + (-dydr) * dbds[id];
                const double c = (-dxds) * dbdr[id] + dxdr * dbds[id];
                const double f0 = a * uu + c * vv;
                r[0][id] = r[0][id] - f0 * djaco;
                r[1][id] = r[1][id] - a * (-pres) * djaco;
                r[2][id] = r[2][id] - c * (-pres) * djaco;
                r[3][id] = r[3][id] - (a * uu + c * vv) * (-pres) * djaco;
            }
        }
    }
    // derivative update with aflim = 1, afvec = I (Getrhslagmc.f90:7695-7708)
    for (int id = 1; id < 3; ++id) {
        const double u1 = ul[1][id], u2 = ul[2][id];
        ul[0][id] = ul[0][id] * 1.0;
        ul[1][id] = 1.0 * u1 + 0.0 * u2;
        ul[2][id] = 0.0 * u1 + 1.0 * u2;
        ul[3][id] = ul[3][id] * 1.0;
    }
    if (apply_update) {
        const double dt = d.scal[0];
        const int off[6] = {0, 0, 5, 9, 12, 14};
        for (int id = 0; id < 6; ++id) {
            double mrow[6];
            for (int kk = 0; kk < 6; ++kk) {
                if (id == 0 || kk == 0) {
                    mrow[kk] = (id == 0 && kk == 0) ? d.A[15 * nc + it] : 0.0;
                } else {
                    const int a = id < kk ? id : kk, b = id < kk ? kk : id;
                    mrow[kk] = d.A[(off[a] + (b - a)) * nc + it];
                }
            }
            for (int q = 0; q < 4; ++q) {
                double s4 = 0.0;
                for (int kk = 0; kk < 6; ++kk) s4 = s4 + mrow[kk] * (dt * r[q][kk]);
                UD6(d.Us, q, id, it) = alfa * UD6(d.Un, q, id, it) + (1.0 - alfa) * (ul[q][id] + s4);
            }
        }
    } else {
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) {
                UD6(d.Us, q, id, it) = ul[q][id];
                UD6(d.R, q, id, it) = r[q][id];
            }
    }
}

}  // namespace lag2d
