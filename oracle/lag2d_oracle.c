/* CPU restatement of the reference's Lagrangian DG step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY: the parity tests, the smoke check and the bench's
 * CPU-baseline leg compare the HIP path against this code; the product never
 * links or calls it.  Built with -ffp-contract=off so every product and sum is
 * rounded where the reference rounds it; expressions keep the reference's
 * left-to-right evaluation order.  Each function names the reference
 * subroutine (file:line) whose arithmetic it restates.
 *
 * Paths covered: planar Lagrangian DG with nriem=1, nlimi=6, ndens=1 and
 *   - ncurv=0, nfint=1, P1 on straight quads and triangles (hybrid meshes),
 *     Barth + symmetry-preserving velocity limiter, Getrhslagmc.f90:6810-7405.
 */
#include "lag2d_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "orc_util.h"

/* The 9 geometry points of a quad: Q2 nodes, or for straight quads the
 * corners plus chord midpoints (Limiterlag.f90:2008-2021). */
static void quad_points9(const orc_case *c, const double *x, int iq, double xp[2][9])
{
    const int *ip = c->ipqua + c->nvqua * iq;
    if (c->ncurv == 0) {
        for (int k = 0; k < 4; ++k) {
            xp[0][k] = X(x, 0, ip[k]);
            xp[1][k] = X(x, 1, ip[k]);
        }
        for (int d = 0; d < 2; ++d) {
            xp[d][4] = 0.5 * (xp[d][0] + xp[d][1]);
            xp[d][5] = 0.5 * (xp[d][1] + xp[d][2]);
            xp[d][6] = 0.5 * (xp[d][2] + xp[d][3]);
            xp[d][7] = 0.5 * (xp[d][3] + xp[d][0]);
            xp[d][8] = 0.5 * (xp[d][4] + xp[d][6]);
        }
    } else {
        for (int k = 0; k < 9; ++k) {
            xp[0][k] = X(x, 0, ip[k]);
            xp[1][k] = X(x, 1, ip[k]);
        }
    }
}

/* rutope, 7-point area rule of the reference triangle r, s in [0, 1]
 * (dgflo2d.f90:2234-2255), its literal values. */
static const double PT7[7][2] = {
    {0.333333333333333, 0.333333333333333}, {0.059715871789770, 0.470142064105115},
    {0.470142064105115, 0.059715871789770}, {0.470142064105115, 0.470142064105115},
    {0.797426985353087, 0.101286507323456}, {0.101286507323456, 0.797426985353087},
    {0.101286507323456, 0.101286507323456}};
static const double WT7[7] = {0.225000000000000, 0.132394152788506, 0.132394152788506, 0.132394152788506,
                              0.125939180544827, 0.125939180544827, 0.125939180544827};

/* Reference-triangle vertices (0,0), (1,0), (0,1) and the basis scaling
 * dr = ds = 0.5 of the P1 triangle basis (Getrhslagmctria.f90:464-470). */
static const double XVT[3] = {0.0, 1.0, 0.0};
static const double YVT[3] = {0.0, 0.0, 1.0};

/* ------------------------------------------------------------------------ */
/* getcellvol_lag (Limiterlag.f90:1869-2164): current volume geoel(3) and
 * volume centre geoel(5:6); triangles by the 7-point rule on the quadratic
 * triangle map with chord-midpoint mid-nodes (:1898-2036), quads by the 6x6
 * Gauss rule (:2037-2162). */
void orc_cellvol(orc_case *c, const double *x)
{
    const double c10 = 1.0;
    for (int it = 0; it < c->ntria; ++it) {
        /* straight triangles: chord-midpoint mid-nodes; curved (ncurv=1):
         * the six nodes of the mesh (:1904-1914) */
        const int nvt = c->ncurv == 1 ? 6 : 3;
        const int *ip = c->iptri + nvt * it;
        double xp[2][6];
        for (int k = 0; k < nvt; ++k) {
            xp[0][k] = X(x, 0, ip[k]);
            xp[1][k] = X(x, 1, ip[k]);
        }
        if (nvt == 3)
            for (int d = 0; d < 2; ++d) {
                xp[d][3] = 0.5 * (xp[d][0] + xp[d][1]);
                xp[d][4] = 0.5 * (xp[d][1] + xp[d][2]);
                xp[d][5] = 0.5 * (xp[d][0] + xp[d][2]);
            }
        double rcv = 0.0, scv = 0.0, volel = 0.0, areel = 0.0, rcvrz = 0.0, scvrz = 0.0;
        for (int g = 0; g < 7; ++g) {
            const double r = PT7[g][0], s = PT7[g][1], wi = WT7[g];
            double sh[6], dr_[6], ds_[6];
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
            double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
            for (int k = 0; k < 6; ++k) {
                dxdr = dxdr + dr_[k] * xp[0][k];
                dxds = dxds + ds_[k] * xp[0][k];
                dydr = dydr + dr_[k] * xp[1][k];
                dyds = dyds + ds_[k] * xp[1][k];
            }
            double yg = 0.0;
            for (int k = 0; k < 6; ++k) yg = yg + sh[k] * xp[1][k];
            const double rcoef = 1.0 - c->alfrz + c->alfrz * yg;
            const double djaco = 0.5 * wi * (dxdr * dyds - dydr * dxds);
            volel = volel + djaco * rcoef;
            rcv = rcv + r * djaco * rcoef;
            scv = scv + s * djaco * rcoef;
            if (c->nrz == 2) {
                areel = areel + djaco;
                rcvrz = rcvrz + r * djaco;
                scvrz = scvrz + s * djaco;
            }
        }
        GE(c, 2, it) = volel;
        GE(c, 4, it) = rcv / volel;
        GE(c, 5, it) = scv / volel;
        if (c->nrz == 2) {
            GE(c, 4, it) = rcvrz / areel;
            GE(c, 5, it) = scvrz / areel;
            GE(c, 10, it) = volel / areel;
            GE(c, 11, it) = areel;
            GE(c, 14, it) = rcvrz / areel;
            GE(c, 15, it) = scvrz / areel;
        }
    }
    double pos[72], w[36];
    ruqope2(6, pos, w);
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        double xp[2][9];
        quad_points9(c, x, iq, xp);
        double rcv = 0.0, scv = 0.0, volel = 0.0, areel = 0.0, rcvrz = 0.0, scvrz = 0.0;
        for (int g = 0; g < 36; ++g) {
            double r = pos[2 * g], s = pos[2 * g + 1], wi = w[g];
            double sh[9], dr_[9], ds_[9];
            shape_q2(r, s, sh, dr_, ds_);
            double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
            for (int k = 0; k < 9; ++k) {
                dxdr = dxdr + dr_[k] * xp[0][k];
                dxds = dxds + ds_[k] * xp[0][k];
                dydr = dydr + dr_[k] * xp[1][k];
                dyds = dyds + ds_[k] * xp[1][k];
            }
            double yg = 0.0;
            for (int k = 0; k < 9; ++k) yg = yg + sh[k] * xp[1][k];
            double rcoef = 1.0 - c->alfrz + c->alfrz * yg;
            double djaco = wi * (dxdr * dyds - dydr * dxds);
            volel = volel + djaco * rcoef;
            rcv = rcv + r * djaco * rcoef;
            scv = scv + s * djaco * rcoef;
            if (c->nrz == 2) {
                areel = areel + djaco;
                rcvrz = rcvrz + r * djaco;
                scvrz = scvrz + s * djaco;
            }
        }
        GE(c, 2, ie) = volel;
        GE(c, 4, ie) = rcv / volel;
        GE(c, 5, ie) = scv / volel;
        if (c->nrz == 2) {
            GE(c, 4, ie) = rcvrz / areel;
            GE(c, 5, ie) = scvrz / areel;
            GE(c, 10, ie) = volel / areel;
            GE(c, 11, ie) = areel;
            GE(c, 14, ie) = rcvrz / areel;
            GE(c, 15, ie) = scvrz / areel;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* getfnds_lag_mc_hybrid, quad part (Getrhslagmc.f90:2212-2260): unit normal
 * and length of the four straight edges. gq[(k + 3*edge) + 12*ie]. */
static void fnds_q1(const orc_case *c, const double *x, double *gq)
{
    static const int ea[4] = {0, 1, 2, 3}, eb[4] = {1, 2, 3, 0};
    for (int ie = 0; ie < c->nquad; ++ie) {
        const int *ip = c->ipqua + c->nvqua * ie;
        for (int e = 0; e < 4; ++e) {
            double anx = X(x, 1, ip[eb[e]]) - X(x, 1, ip[ea[e]]);
            double any = -(X(x, 0, ip[eb[e]]) - X(x, 0, ip[ea[e]]));
            double *g = gq + 3 * e + 12 * ie;
            g[0] = anx / sqrt(anx * anx + any * any);
            g[1] = any / sqrt(anx * anx + any * any);
            g[2] = sqrt(anx * anx + any * any);
        }
    }
}

/* getfnds_lag_mc_hybrid, triangle loop (Getrhslagmc.f90:2163-2188): edges
 * 1-2, 2-3, 3-1.  gt[(k + 3*edge) + 9*it]. */
static void fnds_tri(const orc_case *c, const double *x, double *gt)
{
    static const int ea[3] = {0, 1, 2}, eb[3] = {1, 2, 0};
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        for (int e = 0; e < 3; ++e) {
            double anx = X(x, 1, ip[eb[e]]) - X(x, 1, ip[ea[e]]);
            double any = -(X(x, 0, ip[eb[e]]) - X(x, 0, ip[ea[e]]));
            double *g = gt + 3 * e + 9 * it;
            g[0] = anx / sqrt(anx * anx + any * any);
            g[1] = any / sqrt(anx * anx + any * any);
            g[2] = sqrt(anx * anx + any * any);
        }
    }
}

/* barthlimit_lag_vtxunknew, quad loop (Getrhslagmctria.f90:2522-2636): per
 * node max/min over surrounding cells of (tau, u, v, p, E, div u).
 * unmax[q + 6*ip]. The divergence row uses dr, ds that the reference only
 * sets in its triangle loop; it feeds nothing on these paths and is left
 * unset here. */
static void vtx_minmax(const orc_case *c, double *unmax, double *unmin)
{
    for (int i = 0; i < 6 * c->npoin; ++i) {
        unmax[i] = -1.e10;
        unmin[i] = 1.e10;
    }
    /* triangle loop (Getrhslagmctria.f90:2421-2521) */
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        double rhom = U(c, c->unold, 0, 0, it);
        double uctr = U(c, c->unold, 0, 1, it);
        double vctr = U(c, c->unold, 0, 2, it);
        double ectr = U(c, c->unold, 0, 3, it);
        double rhoct = 1.0 / rhom;
        double pctr = dmax(EPS_SP6, (c->gamlg - 1.0) * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double unctr[5] = {U(c, c->unold, 0, 0, it), uctr, vctr, pctr, ectr};
        for (int iv = 0; iv < 3; ++iv) {
            int p = ip[iv];
            for (int q = 0; q < 5; ++q) {
                if (unctr[q] > unmax[q + 6 * p]) unmax[q + 6 * p] = unctr[q];
                if (unctr[q] < unmin[q + 6 * p]) unmin[q + 6 * p] = unctr[q];
            }
        }
    }
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double rhom = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie);
        double vctr = U(c, c->unold, 0, 2, ie);
        double ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhom;
        double pctr = dmax(EPS_SP6, (c->gamlg - 1.0) * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double unctr[5] = {U(c, c->unold, 0, 0, ie), uctr, vctr, pctr, ectr};
        for (int iv = 0; iv < 4; ++iv) {
            int p = ip[iv];
            for (int q = 0; q < 5; ++q) {
                if (unctr[q] > unmax[q + 6 * p]) unmax[q + 6 * p] = unctr[q];
                if (unctr[q] < unmin[q + 6 * p]) unmin[q + 6 * p] = unctr[q];
            }
        }
    }
}

/* barthfct (Getrhslagmc.f90:9171-9219) */
static double barthfct(double umax, double umin, double uctr, double dunk)
{
    const double coef = 0.7;
    double fiy;
    if (dunk > 1.e-8) {
        fiy = coef * (umax - uctr) / dunk;
        return dmin((fiy * fiy + 2.0 * fiy) / (fiy * fiy + fiy + 2.0), 1.0);
    } else if (dunk < -1.e-8) {
        fiy = coef * (umin - uctr) / dunk;
        return dmin((fiy * fiy + 2.0 * fiy) / (fiy * fiy + fiy + 2.0), 1.0);
    }
    return 1.0;
}

static const double XVQ[4] = {-1.0, 1.0, 1.0, -1.0};
static const double YVQ[4] = {-1.0, -1.0, 1.0, 1.0};

/* barthlimit_sympre_quadlb (Limiterlag.f90:8127-8692): velocity limiter in
 * the principal-strain frame; writes afvec. Parts 5 and "further technique"
 * only write geoel(10), which nothing on this path reads; they are omitted. */
static void sympre_quad(orc_case *c, const int *indbd, double *afvec)
{
    const int nq = c->nq;
    double *mapmt = (double *)calloc((size_t)4 * c->nquad, sizeof(double));
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double xp[2][4];
        for (int k = 0; k < 4; ++k) {
            xp[0][k] = X(c->cooro, 0, ip[k]);
            xp[1][k] = X(c->cooro, 1, ip[k]);
        }
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double dudr = U(c, c->unold, 1, 1, ie) / 1.0;
        double duds = U(c, c->unold, 2, 1, ie) / 1.0;
        double dvdr = U(c, c->unold, 1, 2, ie) / 1.0;
        double dvds = U(c, c->unold, 2, 2, ie) / 1.0;
        double r = rc, s = sc;
        double rp = 1.0 + r, rm = 1.0 - r, sp = 1.0 + s, sm = 1.0 - s;
        double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
        double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
        double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
        for (int k = 0; k < 4; ++k) {
            dxdr = dxdr + dsr[k] * xp[0][k];
            dxds = dxds + dss[k] * xp[0][k];
            dydr = dydr + dsr[k] * xp[1][k];
            dyds = dyds + dss[k] * xp[1][k];
        }
        double jacom = dxdr * dyds - dydr * dxds;
        double drdx = dyds / jacom, drdy = -dxds / jacom;
        double dsdx = -dydr / jacom, dsdy = dxdr / jacom;
        double dudx = dudr * drdx + duds * dsdx;
        double dudy = dudr * drdy + duds * dsdy;
        double dvdx = dvdr * drdx + dvds * dsdx;
        double dvdy = dvdr * drdy + dvds * dsdy;
        double matra = dudx, matrb = 0.5 * (dudy + dvdx), matrc = matrb, matrd = dvdy;
        double mapt = matra + matrd;
        double mapd = matra * matrd - matrb * matrc;
        double lamda1 = 0.5 * mapt + sqrt(0.25 * (mapt * mapt) - mapd);
        double lamda2 = 0.5 * mapt - sqrt(0.25 * (mapt * mapt) - mapd);
        double *m = mapmt + 4 * iq; /* m[0]=(1,1) m[1]=(1,2) m[2]=(2,1) m[3]=(2,2) */
        if (fabs(matrc) > 1.e-8) {
            double l1 = sqrt((lamda1 - matrd) * (lamda1 - matrd) + matrc * matrc);
            double l2 = sqrt((lamda2 - matrd) * (lamda2 - matrd) + matrc * matrc);
            if (matrc > 0.0) {
                m[0] = (lamda1 - matrd) / l1; m[1] = matrc / l1;
                m[2] = (lamda2 - matrd) / l2; m[3] = matrc / l2;
            } else {
                m[2] = (lamda1 - matrd) / l1; m[3] = matrc / l1;
                m[0] = (lamda2 - matrd) / l2; m[1] = matrc / l2;
            }
        } else {
            m[0] = m[1] = m[2] = m[3] = 0.0;
        }
    }
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        const double *m = mapmt + 4 * iq;
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double alfa[2][4];
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        double uc2 = uctr * m[0] + vctr * m[1];
        double uc3 = uctr * m[2] + vctr * m[3];
        for (int iv = 0; iv < 4; ++iv) {
            double bq[3] = {1.0, (XVQ[iv] - rc) / 1.0, (YVQ[iv] - sc) / 1.0};
            double uv = 0.0, vv = 0.0;
            for (int id = 0; id < c->ndegr && id < 3; ++id) {
                uv = uv + U(c, c->unold, id, 1, ie) * bq[id];
                vv = vv + U(c, c->unold, id, 2, ie) * bq[id];
            }
            double un2 = uv * m[0] + vv * m[1];
            double un3 = uv * m[2] + vv * m[3];
            double mx2 = uc2, mx3 = uc3, mn2 = uc2, mn3 = uc3;
            int p = ip[iv];
            for (int k = c->esuv2[p]; k < c->esuv2[p + 1]; ++k) {
                int je = c->esuv1[k];
                double ucj = U(c, c->unold, 0, 1, je), vcj = U(c, c->unold, 0, 2, je);
#define ACC(UU, VV)                                          \
    do {                                                     \
        double ul = (UU) * m[0] + (VV) * m[1];               \
        double vl = (UU) * m[2] + (VV) * m[3];               \
        mx2 = dmax(mx2, ul); mx3 = dmax(mx3, vl);            \
        mn2 = dmin(mn2, ul); mn3 = dmin(mn3, vl);            \
    } while (0)
                ACC(ucj, vcj);
                if (c->ncase != 1) {
                    if (indbd[p] == 3) {
                        ACC(ucj, -vcj);
                    } else if (indbd[p] == 2) {
                        ACC(-ucj, vcj);
                    } else if (indbd[p] == 10) {
                        ACC(ucj, -vcj);
                        ACC(-ucj, vcj);
                        ACC(-ucj, -vcj);
                    }
                }
#undef ACC
            }
            alfa[0][iv] = barthfct(mx2, mn2, uc2, un2 - uc2);
            alfa[1][iv] = barthfct(mx3, mn3, uc3, un3 - uc3);
        }
        double fixi = 1.e300, fiet = 1.e300;
        for (int iv = 0; iv < 4; ++iv) {
            fixi = dmin(fixi, alfa[0][iv]);
            fiet = dmin(fiet, alfa[1][iv]);
        }
        (void)nq;
        double xicx = m[0], xicy = m[1], etcx = m[2], etcy = m[3];
        double fi11 = xicx * xicx * fixi + etcx * etcx * fiet;
        double fi12 = xicx * xicy * fixi + etcx * etcy * fiet;
        double fi22 = xicy * xicy * fixi + etcy * etcy * fiet;
        double *a = afvec + 4 * ie; /* a[0]=(1,1) a[1]=(2,1) a[2]=(1,2) a[3]=(2,2) */
        a[0] = fi11;
        a[2] = fi12;
        a[1] = fi12;
        a[3] = fi22;
    }
    free(mapmt);
}

/* Node colouring of the symmetry-preserving velocity limiters
 * (barthlimit_sympre_quadlb, Limiterlag.f90:8212-8250; the triangle one,
 * barthlimit_sympre_trialb, Getrhslagmctria.f90:2711-2746, colours the same
 * way): 3 on a v=0 wall, 2 on a u=0 wall, 10 where they meet. */
static void sym_colour(const orc_case *c, int *indbd)
{
    if (c->ncase == 1) return;
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 22 && b[3] == 221) {
            indbd[b[0]] = 3; indbd[b[1]] = 3;
            if (c->ncurv == 1) indbd[b[4]] = 3;
        }
    }
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 22 && b[3] == 222) {
            int nn = c->ncurv == 1 ? 3 : 2;
            for (int k = 0; k < nn; ++k) {
                int p = (k < 2) ? b[k] : b[4];
                indbd[p] = (indbd[p] == 3) ? 10 : 2;
            }
        }
    }
}

/* barthlimit_sympre_trialb (Getrhslagmctria.f90:2644-3258): the triangle
 * velocity limiter works in the frame of the cell-average velocity
 * (:2810-2822), with the same neighbour max/min and symmetry mirrors as the
 * quad limiter; afvec = R^T diag(phi_xi, phi_eta) R (:3236-3255).  Its
 * divergence and Mach factors (:3164-3233) feed nothing and are omitted. */
static void sympre_tri(orc_case *c, const int *indbd, double *afvec)
{
    const double dr = 0.5, ds = 0.5;
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        const double ma = U(c, c->unold, 0, 1, it), mb = U(c, c->unold, 0, 2, it);
        double m[4];
        if (sqrt(ma * ma + mb * mb) > 1.e-10) {
            m[0] = ma / sqrt(ma * ma + mb * mb);
            m[1] = mb / sqrt(ma * ma + mb * mb);
            m[2] = -mb / sqrt(ma * ma + mb * mb);
            m[3] = ma / sqrt(ma * ma + mb * mb);
        } else {
            m[0] = m[1] = m[2] = m[3] = 0.0;
        }
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        const double uctr = U(c, c->unold, 0, 1, it), vctr = U(c, c->unold, 0, 2, it);
        const double uc2 = uctr * m[0] + vctr * m[1];
        const double uc3 = uctr * m[2] + vctr * m[3];
        double fixi = 1.e300, fiet = 1.e300;
        for (int iv = 0; iv < 3; ++iv) {
            double bt[3] = {1.0, (XVT[iv] - rc) / dr, (YVT[iv] - sc) / ds};
            double uv = 0.0, vv = 0.0;
            for (int id = 0; id < 3; ++id) {
                uv = uv + U(c, c->unold, id, 1, it) * bt[id];
                vv = vv + U(c, c->unold, id, 2, it) * bt[id];
            }
            const double un2 = uv * m[0] + vv * m[1];
            const double un3 = uv * m[2] + vv * m[3];
            double mx2 = uc2, mx3 = uc3, mn2 = uc2, mn3 = uc3;
            const int p = ip[iv];
            for (int k = c->esuv2[p]; k < c->esuv2[p + 1]; ++k) {
                const int je = c->esuv1[k];
                const double ucj = U(c, c->unold, 0, 1, je), vcj = U(c, c->unold, 0, 2, je);
#define ACC(UU, VV)                                          \
    do {                                                     \
        double ul = (UU) * m[0] + (VV) * m[1];               \
        double vl = (UU) * m[2] + (VV) * m[3];               \
        mx2 = dmax(mx2, ul); mx3 = dmax(mx3, vl);            \
        mn2 = dmin(mn2, ul); mn3 = dmin(mn3, vl);            \
    } while (0)
                ACC(ucj, vcj);
                if (c->ncase != 1) {
                    if (indbd[p] == 3) {
                        ACC(ucj, -vcj);
                    } else if (indbd[p] == 2) {
                        ACC(-ucj, vcj);
                    } else if (indbd[p] == 10) {
                        ACC(ucj, -vcj);
                        ACC(-ucj, vcj);
                        ACC(-ucj, -vcj);
                    }
                }
#undef ACC
            }
            fixi = dmin(fixi, barthfct(mx2, mn2, uc2, un2 - uc2));
            fiet = dmin(fiet, barthfct(mx3, mn3, uc3, un3 - uc3));
        }
        const double xicx = m[0], xicy = m[1], etcx = m[2], etcy = m[3];
        double *a = afvec + 4 * it; /* a[0]=(1,1) a[1]=(2,1) a[2]=(1,2) a[3]=(2,2) */
        a[0] = xicx * xicx * fixi + etcx * etcx * fiet;
        a[2] = xicx * xicy * fixi + etcx * etcy * fiet;
        a[1] = a[2];
        a[3] = xicy * xicy * fixi + etcy * etcy * fiet;
    }
}

/* barthlimit_lagsym_tria (Getrhslagmctria.f90:422-842) for ndens=1, P1:
 * aflim(1:4) from the vertex values against the node max/min, the velocity
 * limiter, then the total-energy coefficient aflim(5).  The boundary-node
 * overrides of ncase 1, 2, 3 and -7 (:539-582) are not restated; those
 * problems are rejected on hybrid meshes. */
static void barth_tri(orc_case *c, const double *unmax, const double *unmin, double *aflim, double *afvec)
{
    const double g1 = c->gamlg - 1.0, dr = 0.5, ds = 0.5;
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        double unk[4][3];
        for (int iv = 0; iv < 3; ++iv) {
            double bt[3] = {1.0, (XVT[iv] - rc) / dr, (YVT[iv] - sc) / ds};
            for (int q = 0; q < 4; ++q) unk[q][iv] = 0.0;
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) unk[q][iv] = unk[q][iv] + U(c, c->unold, id, q, it) * bt[id];
            double rhov = 1.0 / unk[0][iv];
            double uv = unk[1][iv], vv = unk[2][iv], ev = unk[3][iv];
            double pv = dmax(EPS_SP6, g1 * rhov * (ev - 0.5 * (uv * uv + vv * vv)));
            unk[0][iv] = 1.0 / rhov;
            unk[3][iv] = pv;
        }
        double rhom = U(c, c->unold, 0, 0, it);
        double uctr = U(c, c->unold, 0, 1, it), vctr = U(c, c->unold, 0, 2, it);
        double ectr = U(c, c->unold, 0, 3, it);
        double rhoct = 1.0 / rhom;
        double pctr = dmax(EPS_SP6, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double unctr[4] = {1.0 / rhoct, uctr, vctr, pctr};
        for (int q = 0; q < 4; ++q) {
            double a = 1.e300;
            for (int iv = 0; iv < 3; ++iv) {
                int p = ip[iv];
                a = dmin(a, barthfct(unmax[q + 6 * p], unmin[q + 6 * p], unctr[q], unk[q][iv] - unctr[q]));
            }
            aflim[q + 5 * it] = a;
        }
    }
    int *indbd = (int *)calloc((size_t)c->npoin, sizeof(int));
    sym_colour(c, indbd);
    sympre_tri(c, indbd, afvec);
    free(indbd);
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        double rhom = U(c, c->unold, 0, 0, it);
        double uctr = U(c, c->unold, 0, 1, it), vctr = U(c, c->unold, 0, 2, it);
        double ectr = U(c, c->unold, 0, 3, it);
        double rhoct = 1.0 / rhom;
        double pctr = dmax(EPS_SP6, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double a5 = 1.e300;
        for (int iv = 0; iv < 3; ++iv) {
            double bt[3] = {1.0, (XVT[iv] - rc) / dr, (YVT[iv] - sc) / ds};
            double u[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) u[q] = u[q] + U(c, c->unold, id, q, it) * bt[id];
            double rhov = 1.0 / u[0];
            double pv = dmax(EPS_SP6, g1 * rhov * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2])));
            double t = rhom + aflim[0 + 5 * it] * (1.0 / rhov - rhom);
            double pl = pctr + aflim[3 + 5 * it] * (pv - pctr);
            double e5 = pl / g1 * t + 0.5 * (u[1] * u[1] + u[2] * u[2]);
            int p = ip[iv];
            a5 = dmin(a5, barthfct(unmax[4 + 6 * p], unmin[4 + 6 * p], ectr, e5 - ectr));
        }
        aflim[4 + 5 * it] = a5;
    }
}

/* barthlimit_lagsym_quad (Limiterlag.f90:4-404) for ndens=1, P1.
 * aflim[q + 5*ie] rows tau,u,v,p,E; afvec column-major 2x2 per cell. */
static void barth_quad(orc_case *c, const double *unmax, const double *unmin,
                       double *aflim, double *afvec)
{
    const double g1 = c->gamlg - 1.0;
    int *indbd = (int *)calloc((size_t)c->npoin, sizeof(int));
    sym_colour(c, indbd);
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double unk[4][4];
        for (int iv = 0; iv < 4; ++iv) {
            double bq[3] = {1.0, (XVQ[iv] - rc) / 1.0, (YVQ[iv] - sc) / 1.0};
            for (int q = 0; q < 4; ++q) unk[q][iv] = 0.0;
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) unk[q][iv] = unk[q][iv] + U(c, c->unold, id, q, ie) * bq[id];
            double rhov = 1.0 / unk[0][iv];
            double uv = unk[1][iv], vv = unk[2][iv], ev = unk[3][iv];
            double pv = dmax(EPS_SP6, g1 * rhov * (ev - 0.5 * (uv * uv + vv * vv)));
            unk[0][iv] = 1.0 / rhov;
            unk[3][iv] = pv;
        }
        double rhom = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        double ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhom;
        double pctr = dmax(EPS_SP6, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double unctr[4] = {1.0 / rhoct, uctr, vctr, pctr};
        for (int q = 0; q < 4; ++q) {
            double a = 1.e300;
            for (int iv = 0; iv < 4; ++iv) {
                int p = ip[iv];
                double al = barthfct(unmax[q + 6 * p], unmin[q + 6 * p], unctr[q], unk[q][iv] - unctr[q]);
                a = dmin(a, al);
            }
            aflim[q + 5 * ie] = a;
        }
    }
    sympre_quad(c, indbd, afvec);
    /* Part 3: total-energy coefficient (Limiterlag.f90:253-381) */
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double rhom = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        double ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhom;
        double pctr = dmax(EPS_SP6, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double a5 = 1.e300;
        for (int iv = 0; iv < 4; ++iv) {
            double bq[3] = {1.0, (XVQ[iv] - rc) / 1.0, (YVQ[iv] - sc) / 1.0};
            double u[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) u[q] = u[q] + U(c, c->unold, id, q, ie) * bq[id];
            double rhov = 1.0 / u[0];
            double pv = dmax(EPS_SP6, g1 * rhov * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2])));
            double t = rhom + aflim[0 + 5 * ie] * (1.0 / rhov - rhom);
            double pl = pctr + aflim[3 + 5 * ie] * (pv - pctr);
            double e5 = pl / g1 * t + 0.5 * (u[1] * u[1] + u[2] * u[2]);
            int p = ip[iv];
            double al = barthfct(unmax[4 + 6 * p], unmin[4 + 6 * p], ectr, e5 - ectr);
            a5 = dmin(a5, al);
        }
        aflim[4 + 5 * ie] = a5;
    }
    free(indbd);
}

/* getriem_quad_matrixsym (Limiterlag.f90:708-1033) + getriecoef_matrixnew
 * (:1037-1076) and the nodal solve of getndvelo_lag_mc_matrixsym
 * (:408-704), straight P1 quads. fstarq[(d + 2*(f + 2*iv)) + 16*ie]. */
static void ndvelo_q1(orc_case *c, const double *gt, const double *gq, const double *aflim, const double *afvec,
                      double *fstart, double *fstarq)
{
    const int np = c->npoin;
    const double g1 = c->gamlg - 1.0;
    const double eps = 1.0e-6;
    double *mn = (double *)calloc((size_t)4 * np, sizeof(double)); /* munacn(2,2,np) col-major */
    double *mu = (double *)calloc((size_t)2 * np, sizeof(double));
    double *sn = (double *)calloc((size_t)2 * np, sizeof(double));
    double *mcl = (double *)malloc(sizeof(double) * 32 * (c->nquad + 1)); /* munaclq(2,2,2,4) */
    double *mul = (double *)malloc(sizeof(double) * 16 * (c->nquad + 1));
    double *snl = (double *)malloc(sizeof(double) * 16 * (c->nquad + 1));
    double *mct = (double *)malloc(sizeof(double) * 24 * (c->ntria + 1)); /* munaclt(2,2,2,3) */
    double *mut = (double *)malloc(sizeof(double) * 12 * (c->ntria + 1));
    double *snt = (double *)malloc(sizeof(double) * 12 * (c->ntria + 1));
    /* triangle corners first (getriem_tria_matrixsym, Getrhslagmctria.f90:847-1204,
     * called before the quad routine at Limiterlag.f90:535-536) */
    {
        static const int tfe[3][2] = {{2, 0}, {0, 1}, {1, 2}}; /* (edge before, edge after) */
        static const int tnb[3][2] = {{2, 1}, {0, 2}, {1, 0}}; /* the other end of that edge */
        const double dr = 0.5, ds = 0.5;
        for (int it = 0; it < c->ntria; ++it) {
            const int *ip = c->iptri + 3 * it;
            const double rc = GE(c, 0, it), sc = GE(c, 1, it);
            double bt[3][3];
            for (int iv = 0; iv < 3; ++iv) {
                bt[iv][0] = 1.0;
                bt[iv][1] = (XVT[iv] - rc) / dr;
                bt[iv][2] = (YVT[iv] - sc) / ds;
            }
            double rco[3];
            for (int k = 0; k < 3; ++k) rco[k] = 1.0 - c->alfrz + c->alfrz * X(c->cooro, 1, ip[k]);
            double vn[3][2][3];
            for (int iv = 0; iv < 3; ++iv)
                for (int f = 0; f < 2; ++f) {
                    const double *g = gt + 3 * tfe[iv][f] + 9 * it;
                    vn[iv][f][0] = g[0];
                    vn[iv][f][1] = g[1];
                    vn[iv][f][2] = (2.0 * rco[iv] + rco[tnb[iv][f]]) / 3.0 * 0.5 * g[2];
                }
            const double rhomc = U(c, c->unold, 0, 0, it);
            const double uctr = U(c, c->unold, 0, 1, it), vctr = U(c, c->unold, 0, 2, it);
            const double ectr = U(c, c->unold, 0, 3, it);
            const double rhoct = 1.0 / rhomc;
            const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
            const double *af = afvec + 4 * it;
            const double dudr = af[0] * U(c, c->unold, 1, 1, it) + af[2] * U(c, c->unold, 1, 2, it);
            const double duds = af[0] * U(c, c->unold, 2, 1, it) + af[2] * U(c, c->unold, 2, 2, it);
            const double dvdr = af[1] * U(c, c->unold, 1, 1, it) + af[3] * U(c, c->unold, 1, 2, it);
            const double dvds = af[1] * U(c, c->unold, 2, 1, it) + af[3] * U(c, c->unold, 2, 2, it);
            double uvx[3], vvx[3], pvx[3], aj[3][3];
            for (int iv = 0; iv < 3; ++iv) {
                double u[4] = {0.0, 0.0, 0.0, 0.0};
                for (int id = 0; id < 3; ++id)
                    for (int q = 0; q < 4; ++q) u[q] = u[q] + U(c, c->unold, id, q, it) * bt[iv][id];
                const double rhovt = 1.0 / u[0];
                const double pv = dmax(eps, g1 * rhovt * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2])));
                uvx[iv] = U(c, c->unold, 0, 1, it) + dudr * bt[iv][1] + duds * bt[iv][2];
                vvx[iv] = U(c, c->unold, 0, 2, it) + dvdr * bt[iv][1] + dvds * bt[iv][2];
                pvx[iv] = pctr + aflim[3 + 5 * it] * (pv - pctr);
                const double acnx = X(c->ustar, 0, ip[iv]) - uvx[iv];
                const double acny = X(c->ustar, 1, ip[iv]) - vvx[iv];
                if (sqrt(acnx * acnx + acny * acny) < (double)1.e-11f) {
                    aj[iv][0] = (double)1.e-11f;
                    aj[iv][1] = (double)1.e-11f;
                } else {
                    aj[iv][0] = acnx / sqrt(acnx * acnx + acny * acny);
                    aj[iv][1] = acny / sqrt(acnx * acnx + acny * acny);
                }
                aj[iv][2] = sqrt(acnx * acnx + acny * acny);
            }
            const double sdctr = sqrt(dmax(eps, c->gamlg * pctr / rhoct));
            for (int iv = 0; iv < 3; ++iv) aj[iv][2] = aj[iv][2] / sdctr;
            double murie[3];
            for (int iv = 0; iv < 3; ++iv) {
                const double dux = X(c->ustar, 0, ip[iv]) - uvx[iv];
                const double duy = X(c->ustar, 1, ip[iv]) - vvx[iv];
                const double deltu = sqrt(dux * dux + duy * duy);
                murie[iv] = rhoct * sdctr + c->cimpd * rhoct * c->slpdu * deltu;
            }
            for (int iv = 0; iv < 3; ++iv) {
                const int p = ip[iv];
                for (int f = 0; f < 2; ++f) {
                    const double area = vn[iv][f][2], nx = vn[iv][f][0], ny = vn[iv][f][1];
                    const double vnau = (aj[iv][2] < 1.e12) ? 1.0 : nx * aj[iv][0] + ny * aj[iv][1];
                    const double m11 = murie[iv] * area * fabs(vnau), m21 = 0.0, m12 = 0.0;
                    const double m22 = murie[iv] * area * fabs(vnau);
                    const double mu1 = m11 * uvx[iv] + m21 * vvx[iv];
                    const double mu2 = m21 * uvx[iv] + m22 * vvx[iv];
                    const double sxx = -pvx[iv], sxy = 0.0, syx = 0.0, syy = -pvx[iv];
                    const double sn1 = sxx * area * nx + sxy * area * ny;
                    const double sn2 = syx * area * nx + syy * area * ny;
                    mn[0 + 4 * p] = mn[0 + 4 * p] + m11;
                    mn[1 + 4 * p] = mn[1 + 4 * p] + m21;
                    mn[2 + 4 * p] = mn[2 + 4 * p] + m12;
                    mn[3 + 4 * p] = mn[3 + 4 * p] + m22;
                    mu[0 + 2 * p] = mu[0 + 2 * p] + mu1;
                    mu[1 + 2 * p] = mu[1 + 2 * p] + mu2;
                    sn[0 + 2 * p] = sn[0 + 2 * p] + sn1;
                    sn[1 + 2 * p] = sn[1 + 2 * p] + sn2;
                    const int o = f + 2 * iv;
                    mct[0 + 4 * o + 24 * it] = m11;
                    mct[1 + 4 * o + 24 * it] = m21;
                    mct[2 + 4 * o + 24 * it] = m12;
                    mct[3 + 4 * o + 24 * it] = m22;
                    mut[0 + 2 * o + 12 * it] = mu1;
                    mut[1 + 2 * o + 12 * it] = mu2;
                    snt[0 + 2 * o + 12 * it] = sn1;
                    snt[1 + 2 * o + 12 * it] = sn2;
                }
            }
        }
    }
    /* corner faces: (edge before, edge after) of each vertex */
    static const int fe[4][2] = {{3, 0}, {0, 1}, {1, 2}, {2, 3}};
    static const int nb[4][2] = {{3, 1}, {0, 2}, {1, 3}, {2, 0}};
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double rco[4];
        for (int k = 0; k < 4; ++k) rco[k] = 1.0 - c->alfrz + c->alfrz * X(c->cooro, 1, ip[k]);
        double vn[4][2][3];
        for (int iv = 0; iv < 4; ++iv)
            for (int f = 0; f < 2; ++f) {
                const double *g = gq + 3 * fe[iv][f] + 12 * iq;
                vn[iv][f][0] = g[0];
                vn[iv][f][1] = g[1];
                vn[iv][f][2] = (2.0 * rco[iv] + rco[nb[iv][f]]) / 3.0 * 0.5 * g[2];
            }
        double rhomc = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        double ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhomc;
        double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        const double *af = afvec + 4 * ie;
        double dudr = af[0] * U(c, c->unold, 1, 1, ie) + af[2] * U(c, c->unold, 1, 2, ie);
        double duds = af[0] * U(c, c->unold, 2, 1, ie) + af[2] * U(c, c->unold, 2, 2, ie);
        double dvdr = af[1] * U(c, c->unold, 1, 1, ie) + af[3] * U(c, c->unold, 1, 2, ie);
        double dvds = af[1] * U(c, c->unold, 2, 1, ie) + af[3] * U(c, c->unold, 2, 2, ie);
        double uvx[4], vvx[4], pvx[4], a3[4];
        for (int iv = 0; iv < 4; ++iv) {
            double bq[3] = {1.0, (XVQ[iv] - rc) / 1.0, (YVQ[iv] - sc) / 1.0};
            double u[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < c->ndegr; ++id)
                for (int q = 0; q < 4; ++q) u[q] = u[q] + U(c, c->unold, id, q, ie) * bq[id];
            double rhovt = 1.0 / u[0];
            double pv = dmax(eps, g1 * rhovt * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2])));
            uvx[iv] = U(c, c->unold, 0, 1, ie) + dudr * bq[1] + duds * bq[2];
            vvx[iv] = U(c, c->unold, 0, 2, ie) + dvdr * bq[1] + dvds * bq[2];
            pvx[iv] = pctr + aflim[3 + 5 * ie] * (pv - pctr);
            double ax = X(c->ustar, 0, ip[iv]) - uvx[iv];
            double ay = X(c->ustar, 1, ip[iv]) - vvx[iv];
            a3[iv] = sqrt(ax * ax + ay * ay);
        }
        double sdctr = sqrt(dmax(eps, c->gamlg * pctr / rhoct));
        for (int iv = 0; iv < 4; ++iv) {
            a3[iv] = a3[iv] / sdctr;
            int p = ip[iv];
            double dux = X(c->ustar, 0, p) - uvx[iv];
            double duy = X(c->ustar, 1, p) - vvx[iv];
            for (int f = 0; f < 2; ++f) {
                double deltu = ORC_FROZEN(fabs(dux * vn[iv][f][0] + duy * vn[iv][f][1]));
                double mur = rhoct * sdctr + c->cimpd * rhoct * c->slpdu * deltu;
                double area = vn[iv][f][2];
                /* getriecoef_matrixnew: aujmp(3) < 1e12 gives vnau = 1 */
                double m11, m22;
                if (a3[iv] < 1.e12) {
                    m11 = mur * area * fabs(1.0);
                    m22 = mur * area * fabs(1.0);
                } else {
                    double ax = dux, ay = duy, nrm = sqrt(ax * ax + ay * ay);
                    double vnau = vn[iv][f][0] * (-ax / nrm) + vn[iv][f][1] * (-ay / nrm);
                    m11 = mur * area * fabs(vnau);
                    m22 = m11;
                }
                double m21 = 0.0, m12 = 0.0;
                double mu1 = m11 * uvx[iv] + m21 * vvx[iv];
                double mu2 = m21 * uvx[iv] + m22 * vvx[iv];
                double sxx = -pvx[iv], sxy = 0.0, syx = 0.0, syy = -pvx[iv];
                double sn1 = sxx * area * vn[iv][f][0] + sxy * area * vn[iv][f][1];
                double sn2 = syx * area * vn[iv][f][0] + syy * area * vn[iv][f][1];
                mn[0 + 4 * p] = mn[0 + 4 * p] + m11;
                mn[1 + 4 * p] = mn[1 + 4 * p] + m21;
                mn[2 + 4 * p] = mn[2 + 4 * p] + m12;
                mn[3 + 4 * p] = mn[3 + 4 * p] + m22;
                mu[0 + 2 * p] = mu[0 + 2 * p] + mu1;
                mu[1 + 2 * p] = mu[1 + 2 * p] + mu2;
                sn[0 + 2 * p] = sn[0 + 2 * p] + sn1;
                sn[1 + 2 * p] = sn[1 + 2 * p] + sn2;
                int o = f + 2 * iv;
                mcl[0 + 4 * o + 32 * iq] = m11;
                mcl[1 + 4 * o + 32 * iq] = m21;
                mcl[2 + 4 * o + 32 * iq] = m12;
                mcl[3 + 4 * o + 32 * iq] = m22;
                mul[0 + 2 * o + 16 * iq] = mu1;
                mul[1 + 2 * o + 16 * iq] = mu2;
                snl[0 + 2 * o + 16 * iq] = sn1;
                snl[1 + 2 * o + 16 * iq] = sn2;
            }
        }
    }
    /* fixed walls (code 25) keep their nodal velocity (Limiterlag.f90:476-493) */
    int *indnd = (int *)calloc((size_t)np, sizeof(int));
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 25) {
            indnd[b[0]] = 1; indnd[b[1]] = 1;
            if (c->ncurv == 1) indnd[b[4]] = 1;
        }
    }
    for (int p = 0; p < np; ++p) {
        if (indnd[p]) continue;
        double a11 = mn[0 + 4 * p], a21 = mn[1 + 4 * p], a12 = mn[2 + 4 * p], a22 = mn[3 + 4 * p];
        double det = a11 * a22 - a21 * a12;
        double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        double r1 = mu[0 + 2 * p] - sn[0 + 2 * p];
        double r2 = mu[1 + 2 * p] - sn[1 + 2 * p];
        X(c->ustar, 0, p) = i11 * r1 + i12 * r2;
        X(c->ustar, 1, p) = i21 * r1 + i22 * r2;
    }
    /* wall projection (Limiterlag.f90:607-651) */
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 22) {
            if (b[3] == 221) { X(c->ustar, 1, b[0]) = 0.0; X(c->ustar, 1, b[1]) = 0.0; }
            else if (b[3] == 222) { X(c->ustar, 0, b[0]) = 0.0; X(c->ustar, 0, b[1]) = 0.0; }
        } else if (b[2] == 25) {
            X(c->ustar, 1, b[0]) = 0.0; X(c->ustar, 1, b[1]) = 0.0;
            X(c->ustar, 0, b[0]) = 0.0; X(c->ustar, 0, b[1]) = 0.0;
        }
    }
    /* corner forces (Limiterlag.f90:663-698), triangles then quads */
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        for (int iv = 0; iv < 3; ++iv)
            for (int f = 0; f < 2; ++f) {
                int o = f + 2 * iv;
                const double *m = mct + 4 * o + 24 * it;
                double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
                fstart[0 + 2 * o + 12 * it] = snt[0 + 2 * o + 12 * it] + m[0] * us1 + m[1] * us2 - mut[0 + 2 * o + 12 * it];
                fstart[1 + 2 * o + 12 * it] = snt[1 + 2 * o + 12 * it] + m[2] * us1 + m[3] * us2 - mut[1 + 2 * o + 12 * it];
            }
    }
    for (int ie = 0; ie < c->nquad; ++ie) {
        const int *ip = c->ipqua + c->nvqua * ie;
        for (int iv = 0; iv < 4; ++iv)
            for (int f = 0; f < 2; ++f) {
                int o = f + 2 * iv;
                const double *m = mcl + 4 * o + 32 * ie;
                double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
                fstarq[0 + 2 * o + 16 * ie] = snl[0 + 2 * o + 16 * ie] + m[0] * us1 + m[1] * us2 - mul[0 + 2 * o + 16 * ie];
                fstarq[1 + 2 * o + 16 * ie] = snl[1 + 2 * o + 16 * ie] + m[2] * us1 + m[3] * us2 - mul[1 + 2 * o + 16 * ie];
            }
    }
    free(indnd);
    free(mn); free(mu); free(sn); free(mcl); free(mul); free(snl); free(mct); free(mut); free(snt);
}

/* rhsifacedg_lag_mc_hybridquad2 (Getrhslagmc.f90:14468-14618): lumped P1
 * face integral per corner; overwrites rhsel. */
static void iface_q1(orc_case *c, const double *gq, const double *fstarq)
{
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    static const int fe[4][2] = {{3, 0}, {0, 1}, {1, 2}, {2, 3}};
    const int nd = c->ndegr;
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double bq[4][3];
        for (int iv = 0; iv < 4; ++iv) {
            bq[iv][0] = 1.0;
            bq[iv][1] = (XVQ[iv] - rc) / 1.0;
            bq[iv][2] = (YVQ[iv] - sc) / 1.0;
        }
        double ul[3] = {0, 0, 0}, pl1[3] = {0, 0, 0}, pl2[3] = {0, 0, 0}, el[3] = {0, 0, 0};
        for (int iv = 0; iv < 4; ++iv) {
            double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
            const double *g0 = gq + 3 * fe[iv][0] + 12 * iq;
            const double *g1 = gq + 3 * fe[iv][1] + 12 * iq;
            const double *f = fstarq + 16 * iq + 4 * iv; /* f[d + 2*face] */
            for (int id = 0; id < nd; ++id) {
                double w = 2.0 * bq[iv][id] + bq[iv][id];
                double l11 = c16 * w * g0[0] * g0[2], l21 = c16 * w * g0[1] * g0[2];
                double l12 = c16 * w * g1[0] * g1[2], l22 = c16 * w * g1[1] * g1[2];
                ul[id] = ul[id] + us1 * l11 + us2 * l21 + us1 * l12 + us2 * l22;
                pl1[id] = pl1[id] + f[0] * c13 * w + f[2] * c13 * w;
                pl2[id] = pl2[id] + f[1] * c13 * w + f[3] * c13 * w;
                el[id] = el[id] + us1 * f[0] * c13 * w + us2 * f[1] * c13 * w +
                         us1 * f[2] * c13 * w + us2 * f[3] * c13 * w;
            }
        }
        for (int id = 0; id < nd; ++id) {
            U(c, c->rhsel, id, 0, ie) = ul[id];
            U(c, c->rhsel, id, 1, ie) = pl1[id];
            U(c, c->rhsel, id, 2, ie) = pl2[id];
            U(c, c->rhsel, id, 3, ie) = el[id];
        }
    }
}

/* rhsifacedg_lag_mc_hybridtria (Getrhslagmc.f90:14006-14160): P1 triangle
 * face integral with consistent corner weights (2 b_i + b_j)/6 (u* n) and
 * (2 b_i + b_j)/3 (forces); overwrites rhsel. */
static void iface_tri(orc_case *c, const double *gt, const double *fstart)
{
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    static const int tfe[3][2] = {{2, 0}, {0, 1}, {1, 2}};
    static const int tnb[3][2] = {{2, 1}, {0, 2}, {1, 0}};
    const double dr = 0.5, ds = 0.5;
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        double bt[3][3];
        for (int iv = 0; iv < 3; ++iv) {
            bt[iv][0] = 1.0;
            bt[iv][1] = (XVT[iv] - rc) / dr;
            bt[iv][2] = (YVT[iv] - sc) / ds;
        }
        double ul[3] = {0, 0, 0}, pl1[3] = {0, 0, 0}, pl2[3] = {0, 0, 0}, el[3] = {0, 0, 0};
        for (int iv = 0; iv < 3; ++iv) {
            const double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
            const double *g0 = gt + 3 * tfe[iv][0] + 9 * it;
            const double *g1 = gt + 3 * tfe[iv][1] + 9 * it;
            const double *f = fstart + 12 * it + 4 * iv; /* f[d + 2*face] */
            for (int id = 0; id < 3; ++id) {
                const double w0 = 2.0 * bt[iv][id] + bt[tnb[iv][0]][id];
                const double w1 = 2.0 * bt[iv][id] + bt[tnb[iv][1]][id];
                const double l10 = c16 * w0 * g0[0] * g0[2], l20 = c16 * w0 * g0[1] * g0[2];
                const double l11 = c16 * w1 * g1[0] * g1[2], l21 = c16 * w1 * g1[1] * g1[2];
                ul[id] = ul[id] + us1 * l10 + us2 * l20 + us1 * l11 + us2 * l21;
                pl1[id] = pl1[id] + f[0] * c13 * w0 + f[2] * c13 * w1;
                pl2[id] = pl2[id] + f[1] * c13 * w0 + f[3] * c13 * w1;
                el[id] = el[id] + us1 * f[0] * c13 * w0 + us2 * f[1] * c13 * w0 + us1 * f[2] * c13 * w1 +
                         us2 * f[3] * c13 * w1;
            }
        }
        for (int id = 0; id < 3; ++id) {
            U(c, c->rhsel, id, 0, it) = ul[id];
            U(c, c->rhsel, id, 1, it) = pl1[id];
            U(c, c->rhsel, id, 2, it) = pl2[id];
            U(c, c->rhsel, id, 3, it) = el[id];
        }
    }
}

/* rhsdomndg_lag_mc_triasym (Getrhslagmctria.f90:1208-1510): -int grad(b).F
 * on the 7-point rule (ngausd = 7 in the control file) with the cofactor
 * form, djaco = w/2.  With rz set, its RZ twin rhsdomndg_lagmc_triarzaw
 * (Getrhslagrz.f90:1951-2239), which differs only in leaving out rcoef. */
static void domn_tri(orc_case *c, const double *aflim, const double *afvec, int rz)
{
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6, dr = 0.5, ds = 0.5;
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        double xp[2][3];
        for (int k = 0; k < 3; ++k) {
            xp[0][k] = X(c->cooro, 0, ip[k]);
            xp[1][k] = X(c->cooro, 1, ip[k]);
        }
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        const double dbdr[3] = {0.0, 1.0 / dr, 0.0}, dbds[3] = {0.0, 0.0, 1.0 / ds};
        const double rhomc = U(c, c->unold, 0, 0, it);
        const double uctr = U(c, c->unold, 0, 1, it), vctr = U(c, c->unold, 0, 2, it);
        const double ectr = U(c, c->unold, 0, 3, it);
        const double rhoct = 1.0 / rhomc;
        const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        const double *af = afvec + 4 * it;
        const double dudr = af[0] * U(c, c->unold, 1, 1, it) + af[2] * U(c, c->unold, 1, 2, it);
        const double duds = af[0] * U(c, c->unold, 2, 1, it) + af[2] * U(c, c->unold, 2, 2, it);
        const double dvdr = af[1] * U(c, c->unold, 1, 1, it) + af[3] * U(c, c->unold, 1, 2, it);
        const double dvds = af[1] * U(c, c->unold, 2, 1, it) + af[3] * U(c, c->unold, 2, 2, it);
        const double dspr[3] = {-1.0, 1.0, 0.0}, dsps[3] = {-1.0, 0.0, 1.0};
        double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
        for (int k = 0; k < 3; ++k) {
            dxdr = dxdr + dspr[k] * xp[0][k];
            dxds = dxds + dsps[k] * xp[0][k];
            dydr = dydr + dspr[k] * xp[1][k];
            dyds = dyds + dsps[k] * xp[1][k];
        }
        double gd[2][3];
        for (int id = 0; id < 3; ++id) {
            gd[0][id] = dyds * dbdr[id] + (-dydr) * dbds[id];
            gd[1][id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
        }
        for (int g = 0; g < 7; ++g) {
            const double r = PT7[g][0], s = PT7[g][1], wi = WT7[g];
            const double sh[3] = {1.0 - r - s, r, s};
            const double djaco = 0.5 * wi;
            double yg = 0.0;
            for (int k = 0; k < 3; ++k) yg = yg + sh[k] * xp[1][k];
            const double rcoef = 1.0 - c->alfrz + c->alfrz * yg;
            const double b[3] = {1.0, (r - rc) / dr, (s - sc) / ds};
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + U(c, c->unold, id, q, it) * b[id];
            const double rhoad = 1.0 / un[0];
            double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
            const double uadv = U(c, c->unold, 0, 1, it) + dudr * b[1] + duds * b[2];
            const double vadv = U(c, c->unold, 0, 2, it) + dvdr * b[1] + dvds * b[2];
            pres = pctr + aflim[3 + 5 * it] * (pres - pctr);
            double fl[3][4];
            for (int id = 0; id < 3; ++id) {
                fl[id][0] = gd[0][id] * uadv + gd[1][id] * vadv;
                fl[id][1] = gd[0][id] * (-pres);
                fl[id][2] = gd[1][id] * (-pres);
                fl[id][3] = (gd[0][id] * uadv + gd[1][id] * vadv) * (-pres);
            }
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q)
                    U(c, c->rhsel, id, q, it) = rz ? U(c, c->rhsel, id, q, it) - fl[id][q] * djaco
                                                   : U(c, c->rhsel, id, q, it) - fl[id][q] * djaco * rcoef;
        }
    }
}

/* rhsdomndg_lag_mc_quadsym (Limiterlag.f90:1108-1490): -int grad(b).F on the
 * 5x5 Gauss rule with the cofactor form (djaco = w); P1 terms only.  The
 * artificial stress of getAS_llnl is identically zero (Limiterlag.f90:15701). */
static void domn_q1(orc_case *c, const double *aflim, const double *afvec)
{
    double pos[50], w[25];
    ruqope2(5, pos, w);
    const double g1 = c->gamlg - 1.0;
    const double eps = 1.0e-6;
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double xp[2][4];
        for (int k = 0; k < 4; ++k) {
            xp[0][k] = X(c->cooro, 0, ip[k]);
            xp[1][k] = X(c->cooro, 1, ip[k]);
        }
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double rhomc = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        double ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhomc;
        double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        const double *af = afvec + 4 * ie;
        double dudr = af[0] * U(c, c->unold, 1, 1, ie) + af[2] * U(c, c->unold, 1, 2, ie);
        double duds = af[0] * U(c, c->unold, 2, 1, ie) + af[2] * U(c, c->unold, 2, 2, ie);
        double dvdr = af[1] * U(c, c->unold, 1, 1, ie) + af[3] * U(c, c->unold, 1, 2, ie);
        double dvds = af[1] * U(c, c->unold, 2, 1, ie) + af[3] * U(c, c->unold, 2, 2, ie);
        for (int g = 0; g < 25; ++g) {
            double r = pos[2 * g], s = pos[2 * g + 1], wi = w[g];
            double rp = 1.0 + r, rm = 1.0 - r, sp = 1.0 + s, sm = 1.0 - s;
            double sh[4] = {0.25 * rm * sm, 0.25 * rp * sm, 0.25 * rp * sp, 0.25 * rm * sp};
            double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
            double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
            double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
            for (int k = 0; k < 4; ++k) {
                dxdr = dxdr + dsr[k] * xp[0][k];
                dxds = dxds + dss[k] * xp[0][k];
                dydr = dydr + dsr[k] * xp[1][k];
                dyds = dyds + dss[k] * xp[1][k];
            }
            double yg = 0.0;
            for (int k = 0; k < 4; ++k) yg = yg + sh[k] * xp[1][k];
            double rcoef = 1.0 - c->alfrz + c->alfrz * yg;
            double djaco = wi;
            /* cofactor G . grad_ref(b) for b2 = r - rc, b3 = s - sc */
            double gd[2][3];
            double dbdr[3] = {0.0, 1.0 / 1.0, 0.0}, dbds[3] = {0.0, 0.0, 1.0 / 1.0};
            for (int id = 0; id < 3; ++id) {
                gd[0][id] = dyds * dbdr[id] + (-dydr) * dbds[id];
                gd[1][id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
            }
            double b[3] = {1.0, (r - rc) / 1.0, (s - sc) / 1.0};
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < c->ndegr; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + U(c, c->unold, id, q, ie) * b[id];
            double rhoma = un[0];
            double rhoad = 1.0 / rhoma;
            double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
            double uadv = U(c, c->unold, 0, 1, ie) + dudr * b[1] + duds * b[2];
            double vadv = U(c, c->unold, 0, 2, ie) + dvdr * b[1] + dvds * b[2];
            pres = pctr + aflim[3 + 5 * ie] * (pres - pctr);
            double s11 = -pres + 0.0, s12 = 0.0 + 0.0, s21 = 0.0 + 0.0, s22 = -pres + 0.0;
            double fl[3][4];
            for (int id = 0; id < 3; ++id) {
                fl[id][0] = gd[0][id] * uadv + gd[1][id] * vadv;
                fl[id][1] = gd[0][id] * s11 + gd[1][id] * s12;
                fl[id][2] = gd[0][id] * s21 + gd[1][id] * s22;
                fl[id][3] = gd[0][id] * (s11 * uadv + s12 * vadv) + gd[1][id] * (s21 * uadv + s22 * vadv);
            }
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q)
                    U(c, c->rhsel, id, q, ie) = U(c, c->rhsel, id, q, ie) - fl[id][q] * djaco * rcoef;
        }
    }
}

/* getrhsdg_lag_mcnew, ncurv=0 / nriem=1 / nfint=1 / nlimi=6 / nrz=0 branch
 * (Getrhslagmc.f90:6810-6939 and the derivative update :7283-7296). */
static int rhs_q1(orc_case *c)
{
    if (c->npoly != 1 || c->nrz != 0 || c->ndens != 1 || c->nlimi != 6) return 2;
    if (c->ntria > 0 && (c->ncase == 1 || c->ncase == 2 || c->ncase == 3 || c->ncase == -7)) return 2;
    double *gq = (double *)malloc(sizeof(double) * 12 * (c->nquad + 1));
    double *gt = (double *)malloc(sizeof(double) * 9 * (c->ntria + 1));
    double *unmax = (double *)malloc(sizeof(double) * 6 * c->npoin);
    double *unmin = (double *)malloc(sizeof(double) * 6 * c->npoin);
    double *fstarq = (double *)malloc(sizeof(double) * 16 * (c->nquad + 1));
    double *fstart = (double *)malloc(sizeof(double) * 12 * (c->ntria + 1));
    fnds_tri(c, c->cooro, gt);
    fnds_q1(c, c->cooro, gq);
    vtx_minmax(c, unmax, unmin);
    barth_tri(c, unmax, unmin, c->aflim, c->afvec);
    barth_quad(c, unmax, unmin, c->aflim, c->afvec);
    for (int i = 0; i < 5 * c->ncell; ++i) c->aflim[i] = ORC_FROZEN(c->aflim[i]);
    for (int i = 0; i < 4 * c->ncell; ++i) c->afvec[i] = ORC_FROZEN(c->afvec[i]);
    ndvelo_q1(c, gt, gq, c->aflim, c->afvec, fstart, fstarq);
    iface_tri(c, gt, fstart);
    iface_q1(c, gq, fstarq);
    domn_tri(c, c->aflim, c->afvec, 0);
    domn_q1(c, c->aflim, c->afvec);
    for (int ie = 0; ie < c->ncell; ++ie) {
        const double *al = c->aflim + 5 * ie, *af = c->afvec + 4 * ie;
        U(c, c->unold, 1, 0, ie) = U(c, c->unold, 1, 0, ie) * al[0];
        U(c, c->unold, 2, 0, ie) = U(c, c->unold, 2, 0, ie) * al[0];
        double dudr = af[0] * U(c, c->unold, 1, 1, ie) + af[2] * U(c, c->unold, 1, 2, ie);
        double duds = af[0] * U(c, c->unold, 2, 1, ie) + af[2] * U(c, c->unold, 2, 2, ie);
        double dvdr = af[1] * U(c, c->unold, 1, 1, ie) + af[3] * U(c, c->unold, 1, 2, ie);
        double dvds = af[1] * U(c, c->unold, 2, 1, ie) + af[3] * U(c, c->unold, 2, 2, ie);
        U(c, c->unold, 1, 1, ie) = dudr;
        U(c, c->unold, 2, 1, ie) = duds;
        U(c, c->unold, 1, 2, ie) = dvdr;
        U(c, c->unold, 2, 2, ie) = dvds;
        U(c, c->unold, 1, 3, ie) = U(c, c->unold, 1, 3, ie) * al[4];
        U(c, c->unold, 2, 3, ie) = U(c, c->unold, 2, 3, ie) * al[4];
    }
    free(gq); free(gt); free(unmax); free(unmin); free(fstarq); free(fstart);
    return 0;
}

/* ======================================================================== */
/* RZ area weighting (nrz = 2): the ncurv=0 / nfint=1 branch of
 * getrhsdg_lag_mcnew with nrz = 2 (Getrhslagmc.f90:6846-6971, :7280-7387). */

static void tri_points6(const orc_case *c, const double *x, int it, double xp[2][6])
{
    const int *ip = c->iptri + 3 * it;
    for (int k = 0; k < 3; ++k) {
        xp[0][k] = X(x, 0, ip[k]);
        xp[1][k] = X(x, 1, ip[k]);
    }
    for (int d = 0; d < 2; ++d) {
        xp[d][3] = 0.5 * (xp[d][0] + xp[d][1]);
        xp[d][4] = 0.5 * (xp[d][1] + xp[d][2]);
        xp[d][5] = 0.5 * (xp[d][0] + xp[d][2]);
    }
}

/* quadratic-triangle shape derivatives (Getrhslagrz.f90:4648-4666) */
static void tri_p2_deriv(double r, double s, double dr_[6], double ds_[6])
{
    const double c10 = 1.0;
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

static double jac_det(const double *dr_, const double *ds_, const double xp0[], const double xp1[], int n)
{
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < n; ++k) {
        dxdr = dxdr + dr_[k] * xp0[k];
        dxds = dxds + ds_[k] * xp0[k];
        dydr = dydr + dr_[k] * xp1[k];
        dyds = dyds + ds_[k] * xp1[k];
    }
    return dxdr * dyds - dydr * dxds;
}

/* bilinear map Jacobian of a straight quad in the source-term order
 * dxdr*dyds - dxds*dydr (Getrhslagrz.f90:2905) */
static void quad_bilinear_jac(const double xp[2][4], double r, double s, double J[4])
{
    const double rp = 1.0 + r, rm = 1.0 - r, sp = 1.0 + s, sm = 1.0 - s;
    const double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
    const double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 4; ++k) {
        dxdr = dxdr + dsr[k] * xp[0][k];
        dxds = dxds + dss[k] * xp[0][k];
        dydr = dydr + dsr[k] * xp[1][k];
        dyds = dyds + dss[k] * xp[1][k];
    }
    J[0] = dxdr; J[1] = dxds; J[2] = dydr; J[3] = dyds;
}

static void quad_corners(const orc_case *c, const double *x, int iq, double xp[2][4])
{
    const int *ip = c->ipqua + c->nvqua * iq;
    for (int k = 0; k < 4; ++k) {
        xp[0][k] = X(x, 0, ip[k]);
        xp[1][k] = X(x, 1, ip[k]);
    }
}

/* getgeoel_lagmc_rzawrev (Getrhslagrz.f90:4590-4998): the reference-space mass
 * centre geoel(1:2) with the limited density, five fixed-point passes, and
 * the first-order centre geoel(15:16). */
static void geoel_rev(orc_case *c, const double *aflim)
{
    for (int it = 0; it < c->ntria; ++it) {
        double xp[2][6];
        tri_points6(c, c->cooro, it, xp);
        for (int iloop = 0; iloop < 5; ++iloop) {
            double xcent = 0.0, ycent = 0.0, xcel = 0.0, ycel = 0.0, masel = 0.0;
            const double rc = GE(c, 0, it), sc = GE(c, 1, it), dr = 0.5, ds = 0.5;
            for (int g = 0; g < 7; ++g) {
                const double r = PT7[g][0], s = PT7[g][1], wi = WT7[g];
                double dr_[6], ds_[6];
                tri_p2_deriv(r, s, dr_, ds_);
                const double djaco = 0.5 * wi * jac_det(dr_, ds_, xp[0], xp[1], 6);
                const double bt[3] = {1.0, (r - rc) / dr, (s - sc) / ds};
                const double rhomc = U(c, c->unold, 0, 0, it);
                double rhomg = 0.0;
                for (int id = 0; id < 3; ++id) rhomg = rhomg + U(c, c->unold, id, 0, it) * bt[id];
                if (iloop == 0) rhomg = rhomc;
                rhomg = rhomc + aflim[0 + 5 * it] * (rhomg - rhomc);
                const double rhog = 1.0 / rhomg;
                masel = masel + 1.0 / rhomc * djaco;
                xcent = xcent + rhog * r * djaco;
                ycent = ycent + rhog * s * djaco;
                xcel = xcel + 1.0 / rhomc * r * djaco;
                ycel = ycel + 1.0 / rhomc * s * djaco;
            }
            GE(c, 0, it) = xcent / masel;
            GE(c, 1, it) = ycent / masel;
            GE(c, 14, it) = xcel / masel;
            GE(c, 15, it) = ycel / masel;
        }
    }
    double pos[72], w[36];
    ruqope2(6, pos, w);
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        double xp[2][9];
        quad_points9(c, c->cooro, iq, xp);
        for (int iloop = 0; iloop < 5; ++iloop) {
            double xcent = 0.0, ycent = 0.0, xcel = 0.0, ycel = 0.0, masel = 0.0;
            const double rc = GE(c, 0, ie), sc = GE(c, 1, ie), dr = 1.0, ds = 1.0;
            for (int g = 0; g < 36; ++g) {
                const double r = pos[2 * g], s = pos[2 * g + 1], wi = w[g];
                double dr_[9], ds_[9];
                shape_q2(r, s, NULL, dr_, ds_);
                const double djaco = wi * jac_det(dr_, ds_, xp[0], xp[1], 9);
                const double bq[3] = {1.0, (r - rc) / dr, (s - sc) / ds};
                const double rhomc = U(c, c->unold, 0, 0, ie);
                double rhomg = 0.0;
                for (int id = 0; id < 3; ++id) rhomg = rhomg + U(c, c->unold, id, 0, ie) * bq[id];
                if (iloop == 0) rhomg = rhomc;
                rhomg = rhomc + aflim[0 + 5 * ie] * (rhomg - rhomc);
                const double rhog = 1.0 / rhomg;
                masel = masel + 1.0 / rhomc * djaco;
                xcent = xcent + rhog * r * djaco;
                ycent = ycent + rhog * s * djaco;
                xcel = xcel + 1.0 / rhomc * r * djaco;
                ycel = ycel + 1.0 / rhomc * s * djaco;
            }
            GE(c, 0, ie) = xcent / masel;
            GE(c, 1, ie) = ycent / masel;
            GE(c, 14, ie) = xcel / masel;
            GE(c, 15, ie) = ycel / masel;
        }
    }
}

/* getamatr_lagmc_rzawrev (Getrhslagrz.f90:3633-4221), P1: the P1 mass matrix
 * with the limited density (triangles on the 7-point rule, quads on 5x5
 * Gauss), inverted and divided by the cell's mean radius geoel(11). */
static void amatr_rev(orc_case *c, const double *aflim)
{
    for (int it = 0; it < c->ntria; ++it) {
        double xp[2][6];
        tri_points6(c, c->cooro, it, xp);
        const double rc = GE(c, 0, it), sc = GE(c, 1, it), dr = 0.5, ds = 0.5;
        double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
        for (int g = 0; g < 7; ++g) {
            const double r = PT7[g][0], s = PT7[g][1], wi = WT7[g];
            double dr_[6], ds_[6];
            tri_p2_deriv(r, s, dr_, ds_);
            const double djaco = 0.5 * wi * jac_det(dr_, ds_, xp[0], xp[1], 6);
            const double xg = r, yg = s;
            const double bt[3] = {1.0, (r - rc) / dr, (s - sc) / ds};
            const double rhomc = U(c, c->unold, 0, 0, it);
            double rhomg = 0.0;
            for (int id = 0; id < 3; ++id) rhomg = rhomg + U(c, c->unold, id, 0, it) * bt[id];
            rhomg = rhomc + aflim[0 + 5 * it] * (rhomg - rhomc);
            const double rhog = 1.0 / rhomg;
            f0 = f0 + rhog * djaco;
            f1 = f1 + rhog * (xg - rc) / dr * (xg - rc) / dr * djaco;
            f2 = f2 + rhog * (xg - rc) / dr * (yg - sc) / ds * djaco;
            f3 = f3 + rhog * (yg - sc) / ds * (yg - sc) / ds * djaco;
        }
        const double det = f1 * f3 - f2 * f2;
        double *a = c->amatr + c->nmatr * it;
        a[0] = f3 / det;
        a[1] = -f2 / det;
        a[2] = f1 / det;
        a[3] = 1.0 / f0;
        for (int k = 0; k < c->nmatr; ++k) a[k] = a[k] / GE(c, 10, it);
    }
    double pos[50], w[25];
    ruqope2(5, pos, w);
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        double xp[2][9];
        quad_points9(c, c->cooro, iq, xp);
        const double rc = GE(c, 0, ie), sc = GE(c, 1, ie), dr = 1.0, ds = 1.0;
        double f0 = 0.0, f1 = 0.0, f2 = 0.0, f3 = 0.0;
        for (int g = 0; g < 25; ++g) {
            const double r = pos[2 * g], s = pos[2 * g + 1], wi = w[g];
            double dr_[9], ds_[9];
            shape_q2(r, s, NULL, dr_, ds_);
            const double djaco = wi * jac_det(dr_, ds_, xp[0], xp[1], 9);
            const double xg = r, yg = s;
            const double bq[3] = {1.0, (r - rc) / dr, (s - sc) / ds};
            const double rhomc = U(c, c->unold, 0, 0, ie);
            double rhomg = 0.0;
            for (int id = 0; id < 3; ++id) rhomg = rhomg + U(c, c->unold, id, 0, ie) * bq[id];
            rhomg = rhomc + aflim[0 + 5 * ie] * (rhomg - rhomc);
            const double rhog = 1.0 / rhomg;
            f0 = f0 + rhog * djaco;
            f1 = f1 + rhog * (xg - rc) / dr * (xg - rc) / dr * djaco;
            f2 = f2 + rhog * (xg - rc) / dr * (yg - sc) / ds * djaco;
            f3 = f3 + rhog * (yg - sc) / ds * (yg - sc) / ds * djaco;
        }
        const double det = f1 * f3 - f2 * f2;
        double *a = c->amatr + c->nmatr * ie;
        a[0] = f3 / det;
        a[1] = -f2 / det;
        a[2] = f1 / det;
        a[3] = 1.0 / f0;
        for (int k = 0; k < c->nmatr; ++k) a[k] = a[k] / GE(c, 10, ie);
    }
}

/* Corner Riemann coefficients of getriem_{tria,quad}rzaw (Getrhslagrz.f90:
 * 1104-1669): corner areas L/2 without the radius factor, eps 1e-8, and an
 * impedance without the shock term (its |du.n| is multiplied by 0). */
static void corner_rz(orc_case *c, int ie, int nv, const int *ip, const double bt[][3], const double *gedge,
                      const int (*fe)[2], const double *aflim, const double *afvec, double *mn, double *mu,
                      double *sn, double *mcl, double *mul, double *snl)
{
    const double eps = 1.0e-8, g1 = c->gamlg - 1.0;
    const double rhomc = U(c, c->unold, 0, 0, ie);
    const double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
    const double ectr = U(c, c->unold, 0, 3, ie);
    const double rhoct = 1.0 / rhomc;
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double *af = afvec + 4 * ie;
    const double dudr = af[0] * U(c, c->unold, 1, 1, ie) + af[2] * U(c, c->unold, 1, 2, ie);
    const double duds = af[0] * U(c, c->unold, 2, 1, ie) + af[2] * U(c, c->unold, 2, 2, ie);
    const double dvdr = af[1] * U(c, c->unold, 1, 1, ie) + af[3] * U(c, c->unold, 1, 2, ie);
    const double dvds = af[1] * U(c, c->unold, 2, 1, ie) + af[3] * U(c, c->unold, 2, 2, ie);
    double uvx[4], vvx[4], pvx[4], aj[4][3];
    for (int iv = 0; iv < nv; ++iv) {
        double u[4] = {0.0, 0.0, 0.0, 0.0};
        for (int id = 0; id < 3; ++id)
            for (int q = 0; q < 4; ++q) u[q] = u[q] + U(c, c->unold, id, q, ie) * bt[iv][id];
        const double rhovt = 1.0 / u[0];
        const double pv = dmax(eps, g1 * rhovt * (u[3] - 0.5 * (u[1] * u[1] + u[2] * u[2])));
        uvx[iv] = U(c, c->unold, 0, 1, ie) + dudr * bt[iv][1] + duds * bt[iv][2];
        vvx[iv] = U(c, c->unold, 0, 2, ie) + dvdr * bt[iv][1] + dvds * bt[iv][2];
        pvx[iv] = pctr + aflim[3 + 5 * ie] * (pv - pctr);
        const double acnx = X(c->ustar, 0, ip[iv]) - uvx[iv];
        const double acny = X(c->ustar, 1, ip[iv]) - vvx[iv];
        if (sqrt(acnx * acnx + acny * acny) < (double)1.e-11f) {
            aj[iv][0] = (double)1.e-11f;
            aj[iv][1] = (double)1.e-11f;
        } else {
            aj[iv][0] = acnx / sqrt(acnx * acnx + acny * acny);
            aj[iv][1] = acny / sqrt(acnx * acnx + acny * acny);
        }
        aj[iv][2] = sqrt(acnx * acnx + acny * acny);
    }
    const double sdctr = sqrt(dmax(eps, c->gamlg * pctr / rhoct));
    for (int iv = 0; iv < nv; ++iv) aj[iv][2] = aj[iv][2] / sdctr;
    for (int iv = 0; iv < nv; ++iv) {
        const int p = ip[iv];
        const double dux = X(c->ustar, 0, p) - uvx[iv];
        const double duy = X(c->ustar, 1, p) - vvx[iv];
        for (int f = 0; f < 2; ++f) {
            const double *g = gedge + 3 * fe[iv][f];
            const double nx = g[0], ny = g[1], area = 0.5 * g[2];
            const double deltu = 0.0 * fabs(dux * nx + duy * ny);
            const double mur = rhoct * sdctr + rhoct * c->slpdu * deltu;
            const double vnau = (aj[iv][2] < 1.e12) ? 1.0 : nx * aj[iv][0] + ny * aj[iv][1];
            const double m11 = mur * area * fabs(vnau), m21 = 0.0, m12 = 0.0;
            const double m22 = mur * area * fabs(vnau);
            const double mu1 = m11 * uvx[iv] + m21 * vvx[iv];
            const double mu2 = m21 * uvx[iv] + m22 * vvx[iv];
            const double sxx = -pvx[iv], sxy = 0.0, syx = 0.0, syy = -pvx[iv];
            const double sn1 = sxx * area * nx + sxy * area * ny;
            const double sn2 = syx * area * nx + syy * area * ny;
            mn[0 + 4 * p] = mn[0 + 4 * p] + m11;
            mn[1 + 4 * p] = mn[1 + 4 * p] + m21;
            mn[2 + 4 * p] = mn[2 + 4 * p] + m12;
            mn[3 + 4 * p] = mn[3 + 4 * p] + m22;
            mu[0 + 2 * p] = mu[0 + 2 * p] + mu1;
            mu[1 + 2 * p] = mu[1 + 2 * p] + mu2;
            sn[0 + 2 * p] = sn[0 + 2 * p] + sn1;
            sn[1 + 2 * p] = sn[1 + 2 * p] + sn2;
            const int o = f + 2 * iv;
            mcl[0 + 4 * o] = m11;
            mcl[1 + 4 * o] = m21;
            mcl[2 + 4 * o] = m12;
            mcl[3 + 4 * o] = m22;
            mul[0 + 2 * o] = mu1;
            mul[1 + 2 * o] = mu2;
            snl[0 + 2 * o] = sn1;
            snl[1 + 2 * o] = sn2;
        }
    }
}

/* getndvelo_lagmc_rzaw (Getrhslagrz.f90:822-1100) with the wall forces of
 * getbc_lagmaire_rzaw (GetBClag.f90:57-250; slip walls, code 22) taken off
 * the nodal right-hand side. */
static void ndvelo_rz(orc_case *c, const double *gt, const double *gq, const double *gf, const double *aflim,
                      const double *afvec, double *fstart, double *fstarq)
{
    const int np = c->npoin;
    double *mn = (double *)calloc((size_t)4 * np, sizeof(double));
    double *mu = (double *)calloc((size_t)2 * np, sizeof(double));
    double *sn = (double *)calloc((size_t)2 * np, sizeof(double));
    double *mct = (double *)malloc(sizeof(double) * 24 * (c->ntria + 1));
    double *mut = (double *)malloc(sizeof(double) * 12 * (c->ntria + 1));
    double *snt = (double *)malloc(sizeof(double) * 12 * (c->ntria + 1));
    double *mcl = (double *)malloc(sizeof(double) * 32 * (c->nquad + 1));
    double *mul = (double *)malloc(sizeof(double) * 16 * (c->nquad + 1));
    double *snl = (double *)malloc(sizeof(double) * 16 * (c->nquad + 1));
    static const int tfe[3][2] = {{2, 0}, {0, 1}, {1, 2}};
    static const int qfe[4][2] = {{3, 0}, {0, 1}, {1, 2}, {2, 3}};
    for (int it = 0; it < c->ntria; ++it) {
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        double bt[3][3];
        for (int iv = 0; iv < 3; ++iv) {
            bt[iv][0] = 1.0;
            bt[iv][1] = (XVT[iv] - rc) / 0.5;
            bt[iv][2] = (YVT[iv] - sc) / 0.5;
        }
        corner_rz(c, it, 3, c->iptri + 3 * it, bt, gt + 9 * it, tfe, aflim, afvec, mn, mu, sn, mct + 24 * it,
                  mut + 12 * it, snt + 12 * it);
    }
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        const double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double bq[4][3];
        for (int iv = 0; iv < 4; ++iv) {
            bq[iv][0] = 1.0;
            bq[iv][1] = (XVQ[iv] - rc) / 1.0;
            bq[iv][2] = (YVQ[iv] - sc) / 1.0;
        }
        corner_rz(c, ie, 4, c->ipqua + c->nvqua * iq, bq, gq + 12 * iq, qfe, aflim, afvec, mn, mu, sn,
                  mcl + 32 * iq, mul + 16 * iq, snl + 16 * iq);
    }
    /* getbc_lagmaire_rzaw: wall-normal area vectors per node, then the force
     * that cancels the normal velocity of the node's 2x2 solve */
    double *bsnv = (double *)calloc((size_t)2 * np, sizeof(double));
    double *fpres = (double *)calloc((size_t)2 * np, sizeof(double));
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] != 22) continue;
        const double L = gf[3 * f + 2];
        double bsnx = 0.0, bsny = 0.0;
        if (b[3] == 221) {
            bsnx = 0.0 * L * 0.5;
            bsny = -1.0 * L * 0.5;
        } else if (b[3] == 222) {
            bsnx = -1.0 * L * 0.5;
            bsny = 0.0 * L * 0.5;
        }
        bsnv[0 + 2 * b[0]] = bsnv[0 + 2 * b[0]] + bsnx;
        bsnv[1 + 2 * b[0]] = bsnv[1 + 2 * b[0]] + bsny;
        bsnv[0 + 2 * b[1]] = bsnv[0 + 2 * b[1]] + bsnx;
        bsnv[1 + 2 * b[1]] = bsnv[1 + 2 * b[1]] + bsny;
    }
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] != 22) continue;
        for (int k = 0; k < 2; ++k) {
            const int p = b[k];
            const double a11 = mn[0 + 4 * p], a21 = mn[1 + 4 * p], a12 = mn[2 + 4 * p], a22 = mn[3 + 4 * p];
            const double det = a11 * a22 - a21 * a12;
            const double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
            const double lnx = bsnv[0 + 2 * p], lny = bsnv[1 + 2 * p];
            const double coefl = (i11 * lnx + i12 * lny) * lnx + (i21 * lnx + i22 * lny) * lny;
            const double rhsx = mu[0 + 2 * p] - sn[0 + 2 * p] - 0.0;
            const double rhsy = mu[1 + 2 * p] - sn[1 + 2 * p] - 0.0;
            const double coefr = (i11 * rhsx + i12 * rhsy) * lnx + (i21 * rhsx + i22 * rhsy) * lny;
            fpres[0 + 2 * p] = coefr / dmax(1.e-16, coefl) * lnx;
            fpres[1 + 2 * p] = coefr / dmax(1.e-16, coefl) * lny;
        }
    }
    for (int i = 0; i < 2 * np; ++i) fpres[i] = fpres[i] + 0.0;
    for (int p = 0; p < np; ++p) {
        double a11 = mn[0 + 4 * p], a21 = mn[1 + 4 * p], a12 = mn[2 + 4 * p], a22 = mn[3 + 4 * p];
        double det = a11 * a22 - a21 * a12;
        double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        double r1 = mu[0 + 2 * p] - sn[0 + 2 * p] - fpres[0 + 2 * p];
        double r2 = mu[1 + 2 * p] - sn[1 + 2 * p] - fpres[1 + 2 * p];
        X(c->ustar, 0, p) = i11 * r1 + i12 * r2;
        X(c->ustar, 1, p) = i21 * r1 + i22 * r2;
    }
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 22) {
            if (b[3] == 221) { X(c->ustar, 1, b[0]) = 0.0; X(c->ustar, 1, b[1]) = 0.0; }
            else if (b[3] == 222) { X(c->ustar, 0, b[0]) = 0.0; X(c->ustar, 0, b[1]) = 0.0; }
        }
    }
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = c->iptri + 3 * it;
        for (int iv = 0; iv < 3; ++iv)
            for (int f = 0; f < 2; ++f) {
                int o = f + 2 * iv;
                const double *m = mct + 4 * o + 24 * it;
                double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
                fstart[0 + 2 * o + 12 * it] = snt[0 + 2 * o + 12 * it] + m[0] * us1 + m[1] * us2 - mut[0 + 2 * o + 12 * it];
                fstart[1 + 2 * o + 12 * it] = snt[1 + 2 * o + 12 * it] + m[2] * us1 + m[3] * us2 - mut[1 + 2 * o + 12 * it];
            }
    }
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        for (int iv = 0; iv < 4; ++iv)
            for (int f = 0; f < 2; ++f) {
                int o = f + 2 * iv;
                const double *m = mcl + 4 * o + 32 * iq;
                double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
                fstarq[0 + 2 * o + 16 * iq] = snl[0 + 2 * o + 16 * iq] + m[0] * us1 + m[1] * us2 - mul[0 + 2 * o + 16 * iq];
                fstarq[1 + 2 * o + 16 * iq] = snl[1 + 2 * o + 16 * iq] + m[2] * us1 + m[3] * us2 - mul[1 + 2 * o + 16 * iq];
            }
    }
    free(mn); free(mu); free(sn); free(mct); free(mut); free(snt); free(mcl); free(mul); free(snl);
    free(bsnv); free(fpres);
}

/* rhsifacedg_lagmc_quadrzaw (Getrhslagrz.f90:1673-1812): quad face integral
 * with consistent corner weights; overwrites rhsel. */
static void iface_quad_rz(orc_case *c, const double *gq, const double *fstarq)
{
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    static const int qfe[4][2] = {{3, 0}, {0, 1}, {1, 2}, {2, 3}};
    static const int qnb[4][2] = {{3, 1}, {0, 2}, {1, 3}, {2, 0}};
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        const double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double bq[4][3];
        for (int iv = 0; iv < 4; ++iv) {
            bq[iv][0] = 1.0;
            bq[iv][1] = (XVQ[iv] - rc) / 1.0;
            bq[iv][2] = (YVQ[iv] - sc) / 1.0;
        }
        double ul[3] = {0, 0, 0}, pl1[3] = {0, 0, 0}, pl2[3] = {0, 0, 0}, el[3] = {0, 0, 0};
        for (int iv = 0; iv < 4; ++iv) {
            const double us1 = X(c->ustar, 0, ip[iv]), us2 = X(c->ustar, 1, ip[iv]);
            const double *g0 = gq + 3 * qfe[iv][0] + 12 * iq;
            const double *g1 = gq + 3 * qfe[iv][1] + 12 * iq;
            const double *f = fstarq + 16 * iq + 4 * iv;
            for (int id = 0; id < 3; ++id) {
                const double w0 = 2.0 * bq[iv][id] + bq[qnb[iv][0]][id];
                const double w1 = 2.0 * bq[iv][id] + bq[qnb[iv][1]][id];
                const double l10 = c16 * w0 * g0[0] * g0[2], l20 = c16 * w0 * g0[1] * g0[2];
                const double l11 = c16 * w1 * g1[0] * g1[2], l21 = c16 * w1 * g1[1] * g1[2];
                ul[id] = ul[id] + us1 * l10 + us2 * l20 + us1 * l11 + us2 * l21;
                pl1[id] = pl1[id] + f[0] * c13 * w0 + f[2] * c13 * w1;
                pl2[id] = pl2[id] + f[1] * c13 * w0 + f[3] * c13 * w1;
                el[id] = el[id] + us1 * f[0] * c13 * w0 + us2 * f[1] * c13 * w0 + us1 * f[2] * c13 * w1 +
                         us2 * f[3] * c13 * w1;
            }
        }
        for (int id = 0; id < 3; ++id) {
            U(c, c->rhsel, id, 0, ie) = ul[id];
            U(c, c->rhsel, id, 1, ie) = pl1[id];
            U(c, c->rhsel, id, 2, ie) = pl2[id];
            U(c, c->rhsel, id, 3, ie) = el[id];
        }
    }
}

/* rhsdomndg_lagmc_quadrzaw (Getrhslagrz.f90:2243-2555): planar-form domain
 * integral on 5x5 Gauss, energy flux (grad b . u)(-p), no radius factor. */
static void domn_quad_rz(orc_case *c, const double *aflim, const double *afvec)
{
    double pos[50], w[25];
    ruqope2(5, pos, w);
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        double xp[2][4];
        quad_corners(c, c->cooro, iq, xp);
        const double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        const double rhomc = U(c, c->unold, 0, 0, ie);
        const double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        const double ectr = U(c, c->unold, 0, 3, ie);
        const double rhoct = 1.0 / rhomc;
        const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        const double *af = afvec + 4 * ie;
        const double dudr = af[0] * U(c, c->unold, 1, 1, ie) + af[2] * U(c, c->unold, 1, 2, ie);
        const double duds = af[0] * U(c, c->unold, 2, 1, ie) + af[2] * U(c, c->unold, 2, 2, ie);
        const double dvdr = af[1] * U(c, c->unold, 1, 1, ie) + af[3] * U(c, c->unold, 1, 2, ie);
        const double dvds = af[1] * U(c, c->unold, 2, 1, ie) + af[3] * U(c, c->unold, 2, 2, ie);
        for (int g = 0; g < 25; ++g) {
            const double r = pos[2 * g], s = pos[2 * g + 1], djaco = w[g];
            double J[4];
            quad_bilinear_jac(xp, r, s, J);
            const double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
            const double dbdr[3] = {0.0, 1.0 / 1.0, 0.0}, dbds[3] = {0.0, 0.0, 1.0 / 1.0};
            double gd[2][3];
            for (int id = 0; id < 3; ++id) {
                gd[0][id] = dyds * dbdr[id] + (-dydr) * dbds[id];
                gd[1][id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
            }
            const double b[3] = {1.0, (r - rc) / 1.0, (s - sc) / 1.0};
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + U(c, c->unold, id, q, ie) * b[id];
            const double rhoad = 1.0 / un[0];
            double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
            const double uadv = U(c, c->unold, 0, 1, ie) + dudr * b[1] + duds * b[2];
            const double vadv = U(c, c->unold, 0, 2, ie) + dvdr * b[1] + dvds * b[2];
            pres = pctr + aflim[3 + 5 * ie] * (pres - pctr);
            for (int id = 0; id < 3; ++id) {
                const double fl[4] = {gd[0][id] * uadv + gd[1][id] * vadv, gd[0][id] * (-pres), gd[1][id] * (-pres),
                                      (gd[0][id] * uadv + gd[1][id] * vadv) * (-pres)};
                for (int q = 0; q < 4; ++q) U(c, c->rhsel, id, q, ie) = U(c, c->rhsel, id, q, ie) - fl[q] * djaco;
            }
        }
    }
}

/* The RZ source terms (Getrhslagmc.f90:6964-6969 and
 * rhsdomnsrcdg_lagmc_{tria,quad}rzaw, Getrhslagrz.f90:2559-3096): R is
 * scaled by the mean radius geoel(11), then int v b dA is added to tau and
 * -int v p b dA to E with the cell-average velocity. */
static void source_rz(orc_case *c, const double *aflim)
{
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    for (int ie = 0; ie < c->ncell; ++ie)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < c->ndegr; ++id) U(c, c->rhsel, id, q, ie) = GE(c, 10, ie) * U(c, c->rhsel, id, q, ie);
    double pos[50], w[25];
    ruqope2(5, pos, w);
    for (int ie = 0; ie < c->ncell; ++ie) {
        const int tri = ie < c->ntria;
        const double dr = tri ? 0.5 : 1.0, ds = dr;
        const int ng = tri ? 7 : 25;
        double xt[2][3], xq[2][4];
        if (tri) {
            const int *ip = c->iptri + 3 * ie;
            for (int k = 0; k < 3; ++k) {
                xt[0][k] = X(c->cooro, 0, ip[k]);
                xt[1][k] = X(c->cooro, 1, ip[k]);
            }
        } else {
            quad_corners(c, c->cooro, ie - c->ntria, xq);
        }
        const double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        const double rhomc = U(c, c->unold, 0, 0, ie);
        const double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        const double ectr = U(c, c->unold, 0, 3, ie);
        const double rhoct = 1.0 / rhomc;
        const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        for (int g = 0; g < ng; ++g) {
            double r, s, djaco;
            if (tri) {
                r = PT7[g][0];
                s = PT7[g][1];
                const double dspr[3] = {-1.0, 1.0, 0.0}, dsps[3] = {-1.0, 0.0, 1.0};
                double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
                for (int k = 0; k < 3; ++k) {
                    dxdr = dxdr + dspr[k] * xt[0][k];
                    dxds = dxds + dsps[k] * xt[0][k];
                    dydr = dydr + dspr[k] * xt[1][k];
                    dyds = dyds + dsps[k] * xt[1][k];
                }
                djaco = 0.5 * WT7[g] * (dxdr * dyds - dxds * dydr);
            } else {
                r = pos[2 * g];
                s = pos[2 * g + 1];
                double J[4];
                quad_bilinear_jac(xq, r, s, J);
                djaco = w[g] * (J[0] * J[3] - J[1] * J[2]);
            }
            const double b[3] = {1.0, (r - rc) / dr, (s - sc) / ds};
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + U(c, c->unold, id, q, ie) * b[id];
            const double rhoad = 1.0 / un[0];
            double pres = g1 * rhoad * (un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]));
            const double vadv = U(c, c->unold, 0, 2, ie);
            pres = pctr + aflim[3 + 5 * ie] * (pres - pctr);
            for (int id = 0; id < 3; ++id) {
                const double fl[4] = {vadv * b[id], 0.0, 0.0, vadv * (-pres) * b[id]};
                for (int q = 0; q < 4; ++q) U(c, c->rhsel, id, q, ie) = U(c, c->rhsel, id, q, ie) + fl[q] * djaco;
            }
        }
    }
}

/* getgeoel_lagmc_rzawdt (Getrhslagrz.f90:5613-5960): mass centre geoel(13:14)
 * on the predicted coordinates xa with the predicted cell-average density. */
static void geoel_dt(orc_case *c, const double *xa, const double *ua)
{
    for (int it = 0; it < c->ntria; ++it) {
        double xp[2][6];
        tri_points6(c, xa, it, xp);
        double xcent = 0.0, ycent = 0.0, masel = 0.0;
        for (int g = 0; g < 7; ++g) {
            const double r = PT7[g][0], s = PT7[g][1], wi = WT7[g];
            double dr_[6], ds_[6];
            tri_p2_deriv(r, s, dr_, ds_);
            const double djaco = 0.5 * wi * jac_det(dr_, ds_, xp[0], xp[1], 6);
            const double rhomc = U(c, ua, 0, 0, it);
            const double rhog = 1.0 / rhomc;
            masel = masel + 1.0 / rhomc * djaco;
            xcent = xcent + rhog * r * djaco;
            ycent = ycent + rhog * s * djaco;
        }
        GE(c, 12, it) = xcent / masel;
        GE(c, 13, it) = ycent / masel;
    }
    double pos[72], w[36];
    ruqope2(6, pos, w);
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        double xp[2][9];
        quad_points9(c, xa, iq, xp);
        double xcent = 0.0, ycent = 0.0, masel = 0.0;
        for (int g = 0; g < 36; ++g) {
            const double r = pos[2 * g], s = pos[2 * g + 1], wi = w[g];
            double dr_[9], ds_[9];
            shape_q2(r, s, NULL, dr_, ds_);
            const double djaco = wi * jac_det(dr_, ds_, xp[0], xp[1], 9);
            const double rhomc = U(c, ua, 0, 0, ie);
            const double rhog = 1.0 / rhomc;
            masel = masel + 1.0 / rhomc * djaco;
            xcent = xcent + rhog * r * djaco;
            ycent = ycent + rhog * s * djaco;
        }
        GE(c, 12, ie) = xcent / masel;
        GE(c, 13, ie) = ycent / masel;
    }
}

/* rhsdomndg_lagmc_{tria,quad}rzawdt (Getrhslagrz.f90:5002-5612): the
 * contribution of the basis' motion with the mass centre, d b / d t from the
 * first-order centre geoel(15:16) and the predicted one geoel(13:14). */
static void domn_dt(orc_case *c, const double *un, const double *aflim, const double *afvec, double *rhsaw)
{
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    double pos[50], w[25];
    ruqope2(5, pos, w);
    for (int ie = 0; ie < c->ncell; ++ie) {
        const int tri = ie < c->ntria;
        const double dr = tri ? 0.5 : 1.0, ds = dr;
        const int ng = tri ? 7 : 25;
        double xt[2][3], xq[2][4];
        if (tri) {
            const int *ip = c->iptri + 3 * ie;
            for (int k = 0; k < 3; ++k) {
                xt[0][k] = X(c->cooro, 0, ip[k]);
                xt[1][k] = X(c->cooro, 1, ip[k]);
            }
        } else {
            quad_corners(c, c->cooro, ie - c->ntria, xq);
        }
        const double rc = GE(c, 14, ie), sc = GE(c, 15, ie);
        const double rc2 = GE(c, 0, ie), sc2 = GE(c, 1, ie);
        const double rcdt = GE(c, 12, ie), scdt = GE(c, 13, ie);
        const double dbdt1 = (rc - rcdt) / dr / (c->epsaw * c->dtfix);
        const double dbdt2 = (sc - scdt) / ds / (c->epsaw * c->dtfix);
        const double rhomc = U(c, un, 0, 0, ie);
        const double uctr = U(c, un, 0, 1, ie), vctr = U(c, un, 0, 2, ie), ectr = U(c, un, 0, 3, ie);
        const double rhoct = 1.0 / rhomc;
        const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        (void)pctr;
        const double *af = afvec + 4 * ie;
        const double dudr = af[0] * U(c, un, 1, 1, ie) + af[2] * U(c, un, 1, 2, ie);
        const double duds = af[0] * U(c, un, 2, 1, ie) + af[2] * U(c, un, 2, 2, ie);
        const double dvdr = af[1] * U(c, un, 1, 1, ie) + af[3] * U(c, un, 1, 2, ie);
        const double dvds = af[1] * U(c, un, 2, 1, ie) + af[3] * U(c, un, 2, 2, ie);
        const double drhomdr = U(c, un, 1, 0, ie) * aflim[0 + 5 * ie];
        const double drhomds = U(c, un, 2, 0, ie) * aflim[0 + 5 * ie];
        const double dedr = U(c, un, 1, 3, ie) * aflim[4 + 5 * ie];
        const double deds = U(c, un, 2, 3, ie) * aflim[4 + 5 * ie];
        for (int g = 0; g < ng; ++g) {
            double r, s, djaco;
            if (tri) {
                r = PT7[g][0];
                s = PT7[g][1];
                const double dspr[3] = {-1.0, 1.0, 0.0}, dsps[3] = {-1.0, 0.0, 1.0};
                double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
                for (int k = 0; k < 3; ++k) {
                    dxdr = dxdr + dspr[k] * xt[0][k];
                    dxds = dxds + dsps[k] * xt[0][k];
                    dydr = dydr + dspr[k] * xt[1][k];
                    dyds = dyds + dsps[k] * xt[1][k];
                }
                djaco = 0.5 * WT7[g] * (dxdr * dyds - dxds * dydr);
            } else {
                r = pos[2 * g];
                s = pos[2 * g + 1];
                double J[4];
                quad_bilinear_jac(xq, r, s, J);
                djaco = w[g] * (J[0] * J[3] - J[1] * J[2]);
            }
            const double b[3] = {1.0, (r - rc2) / dr, (s - sc2) / ds};
            double u0 = 0.0;
            for (int id = 0; id < 3; ++id) u0 = u0 + U(c, un, id, 0, ie) * b[id];
            double rhoma = u0;
            rhoma = rhomc + aflim[0 + 5 * ie] * (rhoma - rhomc);
            const double rhoad = 1.0 / rhoma;
            const double dq[4][2] = {{drhomdr, drhomds}, {dudr, duds}, {dvdr, dvds}, {dedr, deds}};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) {
                    const double fl = rhoad * b[id] * dq[q][0] * dbdt1 + rhoad * b[id] * dq[q][1] * dbdt2;
                    U(c, rhsaw, id, q, ie) = U(c, rhsaw, id, q, ie) - fl * djaco;
                }
        }
    }
}

/* getrhsdg_lag_mcnew, ncurv=0 / nfint=1 / nlimi=6 / nrz=2, P1, slip walls */
static int rhs_rz(orc_case *c)
{
    if (c->npoly != 1 || c->ndens != 1 || c->nlimi != 6) return 2;
    /* ncase 1, 2, 12: wall normals from the node positions (GetBClag.f90:
     * 84-110); 3, 4: pressure boundaries; -7: limiter overrides */
    if (c->ncase == 1 || c->ncase == 2 || c->ncase == 3 || c->ncase == 4 || c->ncase == 12 || c->ncase == -7)
        return 2;
    /* slip walls (22) and free boundaries (21, which adds no force unless
     * ncase is 3 or 4, GetBClag.f90:112-127) */
    for (int f = 0; f < c->nbfac; ++f)
        if (c->bface[5 * f + 2] != 22 && c->bface[5 * f + 2] != 21) return 2;
    const int nd = c->ndegr, nq = c->nq, nt = nd * nq * c->ncell;
    double *gq = (double *)malloc(sizeof(double) * 12 * (c->nquad + 1));
    double *gt = (double *)malloc(sizeof(double) * 9 * (c->ntria + 1));
    double *gf = (double *)malloc(sizeof(double) * 3 * (c->nbfac + 1));
    double *unmax = (double *)malloc(sizeof(double) * 6 * c->npoin);
    double *unmin = (double *)malloc(sizeof(double) * 6 * c->npoin);
    double *fstarq = (double *)malloc(sizeof(double) * 16 * (c->nquad + 1));
    double *fstart = (double *)malloc(sizeof(double) * 12 * (c->ntria + 1));
    double *unori = (double *)malloc(sizeof(double) * nt);
    double *unlim = (double *)malloc(sizeof(double) * nt);
    double *rhsor = (double *)malloc(sizeof(double) * nt);
    double *rhsaw = (double *)calloc((size_t)nt, sizeof(double));
    double *unkaw = (double *)malloc(sizeof(double) * nt);
    double *coora = (double *)malloc(sizeof(double) * 2 * c->npoin);
    fnds_tri(c, c->cooro, gt);
    fnds_q1(c, c->cooro, gq);
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        double anx = X(c->cooro, 1, b[1]) - X(c->cooro, 1, b[0]);
        double any = -(X(c->cooro, 0, b[1]) - X(c->cooro, 0, b[0]));
        gf[3 * f + 0] = anx / sqrt(anx * anx + any * any);
        gf[3 * f + 1] = any / sqrt(anx * anx + any * any);
        gf[3 * f + 2] = sqrt(anx * anx + any * any);
    }
    vtx_minmax(c, unmax, unmin);
    barth_tri(c, unmax, unmin, c->aflim, c->afvec);
    barth_quad(c, unmax, unmin, c->aflim, c->afvec);
    if (c->itime > 1 || c->rkstg > 1) {
        geoel_rev(c, c->aflim);
        amatr_rev(c, c->aflim);
    }
    ndvelo_rz(c, gt, gq, gf, c->aflim, c->afvec, fstart, fstarq);
    iface_tri(c, gt, fstart);
    iface_quad_rz(c, gq, fstarq);
    domn_tri(c, c->aflim, c->afvec, 1);
    domn_quad_rz(c, c->aflim, c->afvec);
    source_rz(c, c->aflim);
    memcpy(unori, c->unold, sizeof(double) * nt);
    for (int ie = 0; ie < c->ncell; ++ie) {
        const double *al = c->aflim + 5 * ie, *af = c->afvec + 4 * ie;
        U(c, c->unold, 1, 0, ie) = U(c, c->unold, 1, 0, ie) * al[0];
        U(c, c->unold, 2, 0, ie) = U(c, c->unold, 2, 0, ie) * al[0];
        double dudr = af[0] * U(c, c->unold, 1, 1, ie) + af[2] * U(c, c->unold, 1, 2, ie);
        double duds = af[0] * U(c, c->unold, 2, 1, ie) + af[2] * U(c, c->unold, 2, 2, ie);
        double dvdr = af[1] * U(c, c->unold, 1, 1, ie) + af[3] * U(c, c->unold, 1, 2, ie);
        double dvds = af[1] * U(c, c->unold, 2, 1, ie) + af[3] * U(c, c->unold, 2, 2, ie);
        U(c, c->unold, 1, 1, ie) = dudr;
        U(c, c->unold, 2, 1, ie) = duds;
        U(c, c->unold, 1, 2, ie) = dvdr;
        U(c, c->unold, 2, 2, ie) = dvds;
        U(c, c->unold, 1, 3, ie) = U(c, c->unold, 1, 3, ie) * al[4];
        U(c, c->unold, 2, 3, ie) = U(c, c->unold, 2, 3, ie) * al[4];
    }
    /* rhsdomndg_lagmc_rzawdt (Getrhslagrz.f90:7220-7342) on the limited state,
     * then the same prediction repeated from the unlimited state
     * (Getrhslagmc.f90:7317-7387), which leaves geoel(13:14) behind */
    for (int pass = 0; pass < 2; ++pass) {
        const double *base = pass == 0 ? c->unold : unori;
        memcpy(unlim, base, sizeof(double) * nt);
        memcpy(rhsor, c->rhsel, sizeof(double) * nt);
        for (int i = 0; i < nt; ++i) rhsor[i] = c->dtfix * rhsor[i];
        for (int ie = 0; ie < c->ncell; ++ie) {
            const double *a = c->amatr + c->nmatr * ie;
            const double m[3][3] = {{a[3], 0.0, 0.0}, {0.0, a[0], a[1]}, {0.0, a[1], a[2]}};
            for (int id = 0; id < 3; ++id) {
                double s4[4] = {0.0, 0.0, 0.0, 0.0};
                for (int k = 0; k < 3; ++k)
                    for (int q = 0; q < 4; ++q) s4[q] = s4[q] + m[id][k] * U(c, rhsor, k, q, ie);
                for (int q = 0; q < 4; ++q) U(c, unkaw, id, q, ie) = U(c, unlim, id, q, ie) + s4[q] * c->epsaw;
            }
        }
        for (int p = 0; p < c->npoin; ++p)
            for (int d = 0; d < 2; ++d)
                X(coora, d, p) = X(c->cooro, d, p) + c->dtfix * c->epsaw * X(c->ustar, d, p);
        geoel_dt(c, coora, unkaw);
        if (pass == 0) domn_dt(c, c->unold, c->aflim, c->afvec, rhsaw);
    }
    for (int ie = 0; ie < c->ncell; ++ie)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < nd; ++id)
                U(c, c->rhsel, id, q, ie) = U(c, c->rhsel, id, q, ie) + GE(c, 10, ie) * U(c, rhsaw, id, q, ie);
    free(gq); free(gt); free(gf); free(unmax); free(unmin); free(fstarq); free(fstart);
    free(unori); free(unlim); free(rhsor); free(rhsaw); free(unkaw); free(coora);
    return 0;
}

int orc_rhs(orc_case *c)
{
    if (c->ncurv == 0 && c->nfint == 1) return c->nrz == 2 ? rhs_rz(c) : rhs_q1(c);
    if (c->ncurv == 1 && (c->nfint == 4 || c->nfint == 5)) return orc_rhs_q2(c);
    return 2;
}

/* getdelt_lag (Getrhslag.f90:20600-20713): dt on stage 1 from the CFL
 * condition over quads (edge length / sound speed) and the volume change. */
static void delt(orc_case *c, const double *dvdt)
{
    const double eps = 1.0e-6, g1 = c->gamlg - 1.0;
    double dtcfl = 1e10, dtvol = 1e10;
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int *ip = c->ipqua + c->nvqua * iq;
        const int ie = c->ntria + iq;
        double xp[2][9];
        int nv = c->ncurv == 0 ? 4 : 9;
        for (int k = 0; k < nv; ++k) {
            xp[0][k] = X(c->coord, 0, ip[k]);
            xp[1][k] = X(c->coord, 1, ip[k]);
        }
        double dl[4];
#define LEN(a, b) sqrt((xp[0][a] - xp[0][b]) * (xp[0][a] - xp[0][b]) + (xp[1][a] - xp[1][b]) * (xp[1][a] - xp[1][b]))
        if (c->ncurv == 0) {
            dl[0] = LEN(0, 1); dl[1] = LEN(1, 2); dl[2] = LEN(2, 3); dl[3] = LEN(3, 0);
        } else {
            dl[0] = LEN(0, 4) + LEN(4, 1);
            dl[1] = LEN(1, 5) + LEN(5, 2);
            dl[2] = LEN(2, 6) + LEN(6, 3);
            dl[3] = LEN(3, 7) + LEN(7, 0);
        }
#undef LEN
        double rhomc = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie);
        double ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhomc;
        double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double sdctr = sqrt(dmax(eps, c->gamlg * pctr / rhoct));
        double dlen = 1.e300;
        for (int k = 0; k < 4; ++k) dlen = dmin(dlen, dl[k]);
        dtcfl = dmin(dtcfl, dlen / sdctr);
    }
    for (int iq = 0; iq < c->nquad; ++iq) {
        const int ie = c->ntria + iq;
        dtvol = dmin(dtvol, GE(c, 2, ie) / fabs(dvdt[ie]));
    }
    if (c->itime == 1)
        c->dtfix = dmin(c->cfl * dtcfl, 0.1 * dtvol);
    else
        c->dtfix = dmin(dmin(c->cfl * dtcfl, 0.1 * dtvol), 1.01 * c->dtfix);
}

/* getsolvdg_lag (Getrhslag.f90:1553-2035), ndens=1, planar, and the driver's
 * copy-back of stage state into U^n and x^n (dgflo2d.f90:408-422). */
/* The per-step time-step and area-weighting schedule of getsolvdg_lag
 * (Getrhslag.f90:1663-1737): epsaw and, with set_dt, dtfix as functions of
 * itime for the problems that fix them. */
void orc_schedule(orc_case *c, int set_dt)
{
    const int it = c->itime;
    double dt = c->dtfix;
    c->epsaw = 1.e-1;
    if (c->nrz >= 1 && c->ncase == 7) {
        if (it <= 2000) { dt = 1e-6; c->epsaw = 1.0; }
        else if (it <= 6000) { dt = 2.5e-6; c->epsaw = 4.e-1; }
        else if (it <= 10000) { dt = 5.e-6; c->epsaw = 2.e-1; }
        else { dt = 5.e-5; c->epsaw = 2.e-2; }
    } else if (c->nrz >= 1 && c->ncase == 11) {
        if (it <= 7500) { dt = 1.e-4; c->epsaw = 1.e-3; }
        else { dt = 2.5e-6; c->epsaw = 2.e-2; }
    } else if (c->ncase == -4) {
        dt = 2.e-6;
    } else if (c->ncase == 13) {
        if (it <= 7500) dt = 1.e-4;
        else if (it <= 17500) dt = 2.e-5;
        else dt = 1e-5;
    } else if (c->ncase == -14) {
        if (it <= 1000) { dt = 5.e-4; c->epsaw = 1.e-2; }
        else dt = 1.e-4;
    }
    if (set_dt) c->dtfix = dt;
}

int orc_step(orc_case *c)
{
    static const double rklag[2][3] = {{0.0, 0.5, 0.0}, {0.0, 0.75, 0.3333333333333333}};
    const int nd = c->ndegr, nq = c->nq;
    ORC_PHASE(ORC_PH_CELLVOL);
    orc_cellvol(c, c->coord);
    orc_schedule(c, 1);
    double *dvdt = (double *)malloc(sizeof(double) * c->ncell);
    double m[6][6];
    for (int istag = 0; istag < c->nstag; ++istag) {
        double alfa = rklag[c->nstag - 2][istag];
        c->rkstg = istag + 1;
        int rc = orc_rhs(c);
        if (rc) { free(dvdt); return rc; }
        ORC_PHASE(ORC_PH_DT);
        if (c->ndt == 1) {
            if (istag == 0) {
                for (int ie = 0; ie < c->ncell; ++ie) dvdt[ie] = U(c, c->rhsel, 0, 0, ie);
                delt(c, dvdt);
                if (c->tacru + c->dtfix > c->tend) c->dtfix = c->tend - c->tacru;
                c->dtfix = ORC_FROZEN(c->dtfix);
                c->tacru = c->tacru + c->dtfix;
            }
        } else if (istag == 0) {
            c->tacru = c->tacru + c->dtfix;
        }
        ORC_PHASE(ORC_PH_UPDATE);
        for (int i = 0; i < nd * nq * c->ncell; ++i) c->rhsel[i] = c->dtfix * c->rhsel[i];
        for (int ie = 0; ie < c->ncell; ++ie) {
            const double *a = c->amatr + c->nmatr * ie;
            if (c->npoly == 1) {
                m[0][0] = a[3]; m[0][1] = 0.0; m[0][2] = 0.0;
                m[1][0] = m[0][1]; m[1][1] = a[0]; m[1][2] = a[1];
                m[2][0] = m[0][2]; m[2][1] = a[1]; m[2][2] = a[2];
            } else {
                for (int i = 0; i < 6; ++i)
                    for (int j = 0; j < 6; ++j) m[i][j] = 0.0;
                m[0][0] = a[15];
                int k = 0;
                for (int i = 1; i < nd; ++i)
                    for (int j = i; j < nd; ++j) {
                        m[i][j] = a[k];
                        m[j][i] = a[k];
                        ++k;
                    }
            }
            for (int id = 0; id < nd; ++id) {
                double un[4] = {0.0, 0.0, 0.0, 0.0};
                for (int k = 0; k < nd; ++k)
                    for (int q = 0; q < nq; ++q) un[q] = un[q] + m[id][k] * U(c, c->rhsel, k, q, ie);
                for (int q = 0; q < nq; ++q)
                    U(c, c->unold, id, q, ie) = alfa * U(c, c->unkno, id, q, ie) +
                                                (1.0 - alfa) * (U(c, c->unold, id, q, ie) + un[q]);
            }
        }
        ORC_PHASE(ORC_PH_NODEMOVE);
        for (int p = 0; p < c->npoin; ++p)
            for (int d = 0; d < 2; ++d)
                X(c->cooro, d, p) = alfa * X(c->coord, d, p) +
                                    (1.0 - alfa) * (X(c->cooro, d, p) + c->dtfix * X(c->ustar, d, p));
        ORC_PHASE(ORC_PH_CELLVOL);
        if (c->ncurv == 1) {
            for (int ie = 0; ie < c->nquad; ++ie) {
                const int *ip = c->ipqua + c->nvqua * ie;
                for (int d = 0; d < 2; ++d) {
                    double x1 = X(c->cooro, d, ip[0]), x2 = X(c->cooro, d, ip[1]);
                    double x3 = X(c->cooro, d, ip[2]), x4 = X(c->cooro, d, ip[3]);
                    double x5 = X(c->cooro, d, ip[4]), x6 = X(c->cooro, d, ip[5]);
                    double x7 = X(c->cooro, d, ip[6]), x8 = X(c->cooro, d, ip[7]);
                    X(c->cooro, d, ip[8]) = -0.25 * (x1 + x2 + x3 + x4) + 0.5 * (x5 + x6 + x7 + x8);
                }
            }
        }
        orc_cellvol(c, c->cooro);
    }
    /* driver copy-back */
    memcpy(c->unkno, c->unold, sizeof(double) * nd * nq * c->ncell);
    memcpy(c->coord, c->cooro, sizeof(double) * 2 * c->npoin);
    free(dvdt);
    return 0;
}

