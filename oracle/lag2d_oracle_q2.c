/* CPU restatement of the curved P2 sub-cell path of the reference's
 * Lagrangian DG right-hand side (the oracle; TEST INFRASTRUCTURE ONLY).
 *
 * ncurv=1, nfint=4, npoly=2 or 1, nlimi=6, ndens=1, planar
 * (getrhsdg_lag_mcnew branch Getrhslagmc.f90:7556-7708):
 *   getfnds_lagsmsif_hybrid      Getrhslaggauss.f90:3356-3598
 *   getfnds_lagsms_simpsonhybrid Getrhslaggauss.f90:9660-10017
 *   getSIHO_Persson              Limiterlag.f90:12042-12236
 *   weno_char_quad_orthgl        WENO.f90:5927-6120 (+ :5889-8098, LinearAlgebra.f90)
 *   getndvelo_lagsubg            Getrhslaggauss.f90:4348-4614
 *   getriem_quadsubg2            Getrhslaggauss.f90:10409-10864
 *   getrhosubg_averg / _daverg   Getrhslaggauss.f90:7067-7392, :10949-11109
 *   rhsifacedg_lagsubgq2         Getrhslaggauss.f90:8501-8745
 *   rhsdomndg_lagmc_curvquad     Getrhslag.f90:4716-5013
 * Curved six-node triangles of hybrid meshes (cells 0 .. ntria-1):
 *   getfnds_lagsmsif_hybrid, getfnds_lagsms_simpsonhybrid (triangle parts)
 *   getriem_triasubg3             Getrhslaggauss.f90:12201-12590
 *   getrhosubgt_average2          Getrhslaggauss.f90:12595-12940
 *   rhsifacedg_lagsubgt3          Getrhslaggauss.f90:8297-8496
 *   rhsdomndg_lagmc_curvtria      Getrhslag.f90:4449-4711
 * barthlimit_lagcurv is called by the reference on this branch but its
 * result (node max/min) is not read for P2; it is omitted.  For P1 the
 * reference also runs barthlimit_lagcurv and barthlimit_lag_curvquadcb
 * (:7598-7604), whose outputs (aflim, afvec, node max/min and geoel(10)) it
 * overwrites before any use (:7607, :7650-7656): they are omitted too, and
 * P1 takes every cell to WENO (weno_cell_p1) in place of the Persson flag.
 *
 * Sums follow the reference's loop order so the restatement rounds where
 * the reference rounds; the one exception is the 5x5 inverse of the WENO
 * normal equations (Matinv225, LinearAlgebra.f90:130-), computed here by
 * Gauss-Jordan elimination instead of the reference's expanded cofactor
 * formula, so troubled-cell reconstructions agree to round-off.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "lag2d_oracle.h"
#include "orc_util.h"

/* sub-cell vertex -> Q2 node (ipqsg, Getrhslaggauss.f90:4436-4439) */
static const int IPQSG[4][4] = {{0, 4, 8, 7}, {4, 1, 5, 8}, {8, 5, 2, 6}, {7, 8, 6, 3}};
/* signed face-normal table fnqsg(8,4) (Getrhslaggauss.f90:10490-10506):
 * entry [isg][2*ivsg + f], sign gives orientation, |value| the gesgq slot */
static const int FNQSG[4][8] = {{8, 1, 13, -9, -9, 12, 12, 16},
                                {9, 13, 2, 3, 14, -10, -10, 9},
                                {-11, 10, 10, 14, 4, 5, 15, -11},
                                {16, -12, -12, 11, 11, 15, 6, 7}};
/* reference coordinates of the 9 Q2 nodes */
static const double XV9[9] = {-1.0, 1.0, 1.0, -1.0, 0.0, 1.0, 0.0, -1.0, 0.0};
static const double YV9[9] = {-1.0, -1.0, 1.0, 1.0, -1.0, 0.0, 1.0, 0.0, 0.0};

/* Riemann-solver corner weights wfgsq (second assignment,
 * Getrhslaggauss.f90:10575-10597) times 1/4: [isg][ivsg][f] */
static void wfg_riem(double w[4][4][2])
{
    static const double t[4][4][2] = {
        {{4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 8.0 / 3.0}, {1.0, 1.0}, {8.0 / 3.0, 4.0 / 3.0}},
        {{8.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 8.0 / 3.0}, {1.0, 1.0}},
        {{1.0, 1.0}, {8.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 8.0 / 3.0}},
        {{4.0 / 3.0, 8.0 / 3.0}, {1.0, 1.0}, {8.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}}};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int f = 0; f < 2; ++f) w[a][b][f] = 0.25 * t[a][b][f];
}

/* face-integral weights (Getrhslaggauss.f90:8573-8621): momentum/energy
 * wfgsq in {0,1}, volume wfgsqm times 1/4: [isg][ivsg][f] */
static void wfg_face(double wm[4][4][2], double wv[4][4][2])
{
    static const double m[4][4][2] = {{{1, 1}, {1, 1}, {0, 0}, {1, 1}},
                                      {{1, 1}, {1, 1}, {1, 1}, {0, 0}},
                                      {{0, 0}, {1, 1}, {1, 1}, {1, 1}},
                                      {{1, 1}, {0, 0}, {1, 1}, {1, 1}}};
    static const double v[4][4][2] = {
        {{4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 0.0}, {0.0, 0.0}, {0.0, 4.0 / 3.0}},
        {{0.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 0.0 / 3.0}, {0.0, 0.0}},
        {{0.0, 0.0}, {0.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}, {4.0 / 3.0, 0.0 / 3.0}},
        {{4.0 / 3.0, 0.0 / 3.0}, {0.0, 0.0}, {0.0 / 3.0, 4.0 / 3.0}, {4.0 / 6.0, 4.0 / 6.0}}};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int f = 0; f < 2; ++f) {
                wm[a][b][f] = m[a][b][f];
                wv[a][b][f] = 0.25 * v[a][b][f];
            }
}

/* the nine nodes of quad cell ie (cells ntria .. ncell-1 are the quads) */
#define QIP(c, ie) ((c)->ipqua + 9 * ((ie) - (c)->ntria))

static void q2_nodes(const orc_case *c, const double *x, int ie, double xp[2][9])
{
    const int *ip = QIP(c, ie);
    for (int k = 0; k < 9; ++k) {
        xp[0][k] = X(x, 0, ip[k]);
        xp[1][k] = X(x, 1, ip[k]);
    }
}

static void q2_map(const double xp[2][9], double r, double s, double *x, double *y)
{
    double sh[9], dr_[9], ds_[9];
    shape_q2(r, s, sh, dr_, ds_);
    double a = 0.0, b = 0.0;
    for (int k = 0; k < 9; ++k) {
        a = a + sh[k] * xp[0][k];
        b = b + sh[k] * xp[1][k];
    }
    *x = a;
    *y = b;
}

static void q2_jac(const double xp[2][9], double r, double s, double J[4], double *xg, double *yg)
{
    double sh[9], dr_[9], ds_[9];
    shape_q2(r, s, sh, dr_, ds_);
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 9; ++k) {
        dxdr = dxdr + dr_[k] * xp[0][k];
        dxds = dxds + ds_[k] * xp[0][k];
        dydr = dydr + dr_[k] * xp[1][k];
        dyds = dyds + ds_[k] * xp[1][k];
    }
    J[0] = dxdr; J[1] = dxds; J[2] = dydr; J[3] = dyds;
    if (xg) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 9; ++k) {
            a = a + sh[k] * xp[0][k];
            b = b + sh[k] * xp[1][k];
        }
        *xg = a;
        *yg = b;
    }
}

/* Initial density at a point (getrhoig_quadcurv2, Getrhslag.f90:17761-17948,
 * and getrhog_initial, GetInitial.f90:4-147): the piecewise-constant cases of
 * the benchmark problems. */
static double rho_initial(const orc_case *c, double xc, double yc)
{
    if (c->ncase == 6) return (sqrt(xc * xc + (yc - 0.0) * (yc - 0.0)) <= 0.5) ? 1.0 : 0.125;
    return 1.0;
}

/* ------------------------------------------------------------------------ */
/* Face geometry: gq[(k + 3*j) + 48*ie], j = gesgq slot 0..15 */
static void fnds_q2(const orc_case *c, const double *x, double *gq)
{
    static const double posqc[4][2] = {{0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
    static const double posiq[12][2] = {{-1, -1}, {1, -1}, {1, -1}, {1, 1}, {1, 1}, {-1, 1},
                                        {-1, 1}, {-1, -1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
    static const double vn0[12][2] = {{0, -1}, {0, -1}, {1, 0}, {1, 0}, {0, 1}, {0, 1},
                                      {-1, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
    static const int slot[12] = {0, 1, 2, 3, 4, 5, 6, 7, 12, 13, 14, 15};
    for (int ie = c->ntria; ie < c->ncell; ++ie) {
        double xp[2][9];
        q2_nodes(c, x, ie, xp);
        double *g = gq + 48 * ie;
        /* interior half-edges: mid-node -> centre (Getrhslaggauss.f90:3549-3596) */
        double xv[4], yv[4];
        for (int k = 0; k < 4; ++k) q2_map(xp, posqc[k][0], posqc[k][1], &xv[k], &yv[k]);
        for (int k = 0; k < 4; ++k) {
            double fx[3] = {xp[0][4 + k], xp[0][8], xv[k]};
            double fy[3] = {xp[1][4 + k], xp[1][8], yv[k]};
            double r = -1.0;
            double d[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
            double dxdr = 0.0, dydr = 0.0;
            for (int i = 0; i < 3; ++i) {
                dxdr = dxdr + d[i] * fx[i];
                dydr = dydr + d[i] * fy[i];
            }
            double djaco = sqrt(dxdr * dxdr + dydr * dydr);
            double anx = -dydr, any = dxdr;
            g[3 * (8 + k) + 0] = anx / sqrt(anx * anx + any * any);
            g[3 * (8 + k) + 1] = any / sqrt(anx * anx + any * any);
            g[3 * (8 + k) + 2] = djaco * 2.0;
        }
        /* boundary-edge points by the cofactor of the Q2 Jacobian (:9884-10002) */
        for (int ig = 0; ig < 12; ++ig) {
            double J[4];
            q2_jac(xp, posiq[ig][0], posiq[ig][1], J, NULL, NULL);
            double c11 = J[3], c12 = -J[2], c21 = -J[1], c22 = J[0];
            double v1 = vn0[ig][0], v2 = vn0[ig][1], larea = 2.0;
            double anx = c11 * v1 + c12 * v2;
            double any = c21 * v1 + c22 * v2;
            double w1 = anx * larea, w2 = any * larea;
            double farea = sqrt(w1 * w1 + w2 * w2);
            g[3 * slot[ig] + 0] = anx * larea / farea;
            g[3 * slot[ig] + 1] = any * larea / farea;
            g[3 * slot[ig] + 2] = farea;
        }
    }
}

/* P2 Taylor basis about the mass centre (Limiterlag.f90:1305-1313) */
static void basis_p2(double r, double s, double rc, double sc, const double *g19, double b[6])
{
    b[0] = 1.0;
    b[1] = (r - rc) / 1.0;
    b[2] = (s - sc) / 1.0;
    b[3] = 0.5 * b[1] * b[1] - g19[0];
    b[4] = 0.5 * b[2] * b[2] - g19[1];
    b[5] = b[1] * b[2] - g19[2];
}

/* getSIHO_Persson (Limiterlag.f90:12042-12236): troubled flag 0 / 100 */
static void persson(const orc_case *c, const double *x, double *flag)
{
    double pos[50], w[25];
    ruqope2(5, pos, w);
    const double thr = -3.0 * log10(4.0) - 4.0;
    for (int ie = 0; ie < c->ntria; ++ie) flag[ie] = 0.0;
    for (int ie = c->ntria; ie < c->ncell; ++ie) {
        double xp[2][9];
        q2_nodes(c, x, ie, xp);
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double v1 = 0.0, v2 = 0.0;
        for (int g = 0; g < 25; ++g) {
            double r = pos[2 * g], s = pos[2 * g + 1], wi = w[g];
            double J[4];
            q2_jac(xp, r, s, J, NULL, NULL);
            double djaco = wi * (J[0] * J[3] - J[2] * J[1]);
            double b[6];
            basis_p2(r, s, rc, sc, &GE(c, 18, ie), b);
            double un = 0.0, u1 = 0.0;
            for (int id = 0; id < 6; ++id) un = un + U(c, c->unold, id, 0, ie) * b[id];
            for (int id = 0; id < 3; ++id) u1 = u1 + U(c, c->unold, id, 0, ie) * b[id];
            v2 = v2 + (u1 - un) * (u1 - un) * djaco;
            v1 = v1 + un * un * djaco;
        }
        flag[ie] = (log10(v2 / v1) <= thr) ? 0.0 : 100.0;
    }
}

/* ---------------------------------------------------------------- WENO -- */
/* Minor determinant by the Leibniz expansion, summed in the order the
 * reference's Matinv225 (LinearAlgebra.f90:130-) lists its terms: the column
 * of the last row varies slowest, ascending, then that of the row before it,
 * and so on; each term is the product of the entries row by row, left to
 * right; terms are accumulated left to right.  a is column-major n x n,
 * rows/cols hold the m rows and columns of the minor in ascending order. */
static double minor_det(const double *a, int n, const int *rows, const int *cols, int m)
{
    int p[5], used[5] = {0, 0, 0, 0, 0};
    double acc = 0.0;
    int first = 1;
    /* iterative enumeration: level k = m-1 down to 1 chooses p[k] ascending */
    int level = m - 1;
    int next[5] = {0, 0, 0, 0, 0};
    while (level < m) {
        if (level == 0) {
            int j = 0;
            while (used[j]) ++j;
            p[0] = j;
            /* sign = parity of the permutation p */
            int inv = 0;
            for (int x = 0; x < m; ++x)
                for (int y = x + 1; y < m; ++y)
                    if (p[x] > p[y]) ++inv;
            double t = a[rows[0] + n * cols[p[0]]];
            for (int k = 1; k < m; ++k) t = t * a[rows[k] + n * cols[p[k]]];
            if (inv & 1) t = -t;
            if (first) {
                acc = t;
                first = 0;
            } else {
                acc = acc + t;
            }
            ++level;
            continue;
        }
        /* advance the choice at this level */
        if (next[level] > 0) used[p[level]] = 0;
        int j = next[level];
        while (j < m && used[j]) ++j;
        if (j >= m) {
            next[level] = 0;
            ++level;
            continue;
        }
        p[level] = j;
        used[j] = 1;
        next[level] = j + 1;
        --level;
    }
    return acc;
}

/* Matinv225 for n = 5 (LinearAlgebra.f90:130-): inverse = adjugate / det,
 * with the determinant and every cofactor expanded as in minor_det. */
static void inv_small(int n, const double *a, double *b)
{
    int all[5] = {0, 1, 2, 3, 4};
    double det = minor_det(a, n, all, all, n);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            /* cofactor C(i, j): remove row i and column j */
            int rows[4], cols[4], k = 0, l = 0;
            for (int x = 0; x < n; ++x)
                if (x != i) rows[k++] = x;
            for (int x = 0; x < n; ++x)
                if (x != j) cols[l++] = x;
            double cof = minor_det(a, n, rows, cols, n - 1);
            if ((i + j) & 1) cof = -cof;
            b[j + n * i] = cof / det; /* inverse(j, i) = C(i, j) / det */
        }
}

/* Matinv225 for n = 2 (LinearAlgebra.f90:141-148, :457-458): adjugate
 * over det; a and b column-major */
static void inv2(const double *a, double *b)
{
    const double det = a[0] * a[3] - a[2] * a[1];
    b[0] = a[3] / det;
    b[2] = -a[2] / det;
    b[1] = -a[1] / det;
    b[3] = a[0] / det;
}

/* rotate (a, b) by the 2x2 frame m (getvector_local, Limiterlag.f90:19009) */
static inline void rot(const double m[4], double *a, double *b)
{
    double x = *a * m[0] + *b * m[1];
    double y = *a * m[2] + *b * m[3];
    *a = x;
    *b = y;
}

/* inverse rotation (getvector_local2 with idtrf=1, WENO.f90:5889-5923) */
static inline void rotinv(const double m[4], double *a, double *b)
{
    double a11 = m[3], a12 = -m[1], a21 = -m[2], a22 = m[0];
    double x = *a * a11 + *b * a12;
    double y = *a * a21 + *b * a22;
    *a = x;
    *b = y;
}

typedef struct {
    const orc_case *c;
    const int *esqua;     /* face neighbours */
    const double *xc;     /* current coordinates */
    const double *x0;     /* initial coordinates */
    const double *geoph;  /* physical mass centres [2*ncell] */
    const double *unknr;  /* snapshot of the DOFs */
    double pos[50], w[25];
} weno_env;

/* One Gauss-point sweep over cell j: initial-map Jacobian weight, initial
 * density and the point in the local frame (WENO.f90:6822-6870 pattern). */
typedef struct {
    double djacoi[25], rhogi[25], xl[25], yl[25];
} cell_pts;

static void cell_points(const weno_env *E, int je, const double m[4], cell_pts *P)
{
    const orc_case *c = E->c;
    double xi[2][9], xl[2][9];
    q2_nodes(c, E->x0, je, xi);
    q2_nodes(c, E->xc, je, xl);
    for (int k = 0; k < 9; ++k) rot(m, &xl[0][k], &xl[1][k]);
    double xcr, ycr;
    q2_map(xi, GE(c, 0, je), GE(c, 1, je), &xcr, &ycr);
    for (int g = 0; g < 25; ++g) {
        double r = E->pos[2 * g], s = E->pos[2 * g + 1], wi = E->w[g];
        double J[4], xgi, ygi;
        q2_jac(xi, r, s, J, &xgi, &ygi);
        P->djacoi[g] = wi * (J[0] * J[3] - J[2] * J[1]);
        double sh[9], d1[9], d2[9];
        shape_q2(r, s, sh, d1, d2);
        double a = 0.0, b = 0.0;
        for (int k = 0; k < 9; ++k) {
            a = a + sh[k] * xl[0][k];
            b = b + sh[k] * xl[1][k];
        }
        P->xl[g] = a;
        P->yl[g] = b;
        P->rhogi[g] = c->rho0 ? c->rho0[je] : rho_initial(c, xcr, ycr);
    }
}

/* getmatrix_prj2 (Limiterlag.f90:16455-16562): characteristic projection */
static void char_matrices(double gam, double dnx, double dny, const double ul[4], const double ur[4],
                          double q[4][4], double qi[4][4])
{
    const double eps = 1.e-6;
    double rhom1 = ul[0], u1 = ul[1], v1 = ul[2], e1 = ul[3];
    double rhoc1 = 1.0 / rhom1;
    double p1 = dmax(eps, (gam - 1.0) * rhoc1 * (e1 - 0.5 * (u1 * u1 + v1 * v1)));
    double rhom2 = ur[0], u2 = ur[1], v2 = ur[2], e2 = ur[3];
    double rhoc2 = 1.0 / rhom2;
    double p2 = dmax(eps, (gam - 1.0) * rhoc2 * (e2 - 0.5 * (u2 * u2 + v2 * v2)));
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

/* weno_stncl_quad_orthol3 (WENO.f90:7541-8098) with getgeophy_local_Taylor
 * (:6750-7001), getunknp_local_Taylor (:7005-7215) and weno_cell_Taylor
 * (:7219-7537), nbsis = 1.  Returns 0 when the cell was reconstructed or
 * skipped as a boundary stencil, 1 when its vertex neighbourhood does not
 * match the structured pattern the reference expects (the reference stops
 * there, WENO.f90:7699; here the cell is skipped like a boundary stencil). */
/* The 9-cell stencil of cell ie in the reference's order (WENO.f90:7641-
 * 7708): the cell, its 4 face neighbours, then the 4 diagonal cells placed by
 * the product of the face indices they share a vertex with.  Returns 0 for
 * a full stencil, 2 for a boundary stencil (the reference leaves the cell
 * unchanged, :7706-7708), 1 when a vertex neighbourhood matches none of the
 * expected products (the reference stops there, :7699). */
static int stencil_cells(const orc_case *c, int ie, const int *esqua, int ielaj[9])
{
    ielaj[0] = ie;
    for (int f = 0; f < 4; ++f) ielaj[1 + f] = esqua[f + 4 * ie];
    for (int k = 5; k < 9; ++k) ielaj[k] = -2; /* unset */
    const int *ip = QIP(c, ie);
    int mismatch = 0;
    for (int iv = 0; iv < 4; ++iv) {
        int jel[3] = {-1, -1, -1}, nv = 0;
        int p = ip[iv];
        for (int k = c->esuv2[p]; k < c->esuv2[p + 1]; ++k) {
            int je = c->esuv1[k];
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
            else mismatch = 1;
        }
    }
    if (mismatch) return 1;
    for (int k = 0; k < 9; ++k)
        if (ielaj[k] < 0) return 2;
    return 0;
}

static int weno_cell(const weno_env *E, int ie, const double m[4], const int *esqua, double *unkno)
{
    const orc_case *c = E->c;
    const int nc = c->ncell;
    int ielaj[9];
    ielaj[0] = ie;
    for (int f = 0; f < 4; ++f) ielaj[1 + f] = esqua[f + 4 * ie];
    for (int k = 5; k < 9; ++k) ielaj[k] = -2; /* unset */
    const int *ip = QIP(c, ie);
    int mismatch = 0;
    for (int iv = 0; iv < 4; ++iv) {
        int jel[3] = {-1, -1, -1}, nv = 0;
        int p = ip[iv];
        for (int k = c->esuv2[p]; k < c->esuv2[p + 1]; ++k) {
            int je = c->esuv1[k];
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
            else mismatch = 1;
        }
    }
    if (mismatch) return 1;
    for (int k = 0; k < 9; ++k)
        if (ielaj[k] < 0) return 0; /* boundary stencil: unchanged */
    (void)nc;

    /* local (rotated) DOFs of the stencil cells */
    double unkrl[9][4][6];
    for (int s = 0; s < 9; ++s)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 6; ++id) unkrl[s][q][id] = U(c, E->unknr, id, q, ielaj[s]);
    for (int s = 0; s < 9; ++s)
        for (int id = 0; id < 6; ++id) rot(m, &unkrl[s][1][id], &unkrl[s][2][id]);

    /* getgeophy_local_Taylor: local centres and P2 offsets (nbsis=1) */
    double geopl[9][5];
    cell_pts P[9];
    double dxcj[9];
    for (int s = 0; s < 9; ++s) {
        int je = ielaj[s];
        double xcl = E->geoph[2 * je], ycl = E->geoph[2 * je + 1];
        rot(m, &xcl, &ycl);
        geopl[s][0] = xcl;
        geopl[s][1] = ycl;
        dxcj[s] = sqrt(GE(c, 2, je));
        cell_points(E, je, m, &P[s]);
        double v1 = 0.0, v2 = 0.0, v3 = 0.0;
        for (int g = 0; g < 25; ++g) {
            double b2 = (P[s].xl[g] - xcl) / dxcj[s];
            double b3 = (P[s].yl[g] - ycl) / dxcj[s];
            double rd = P[s].rhogi[g] * P[s].djacoi[g];
            v1 = v1 + rd * b2 * b2;
            v2 = v2 + rd * b3 * b3;
            v3 = v3 + rd * b2 * b3;
        }
        double masel = GE(c, 3, je);
        v1 = v1 / masel;
        v2 = v2 / masel;
        v3 = v3 / masel;
        geopl[s][2] = 0.5 * v1;
        geopl[s][3] = 0.5 * v2;
        geopl[s][4] = v3;
    }

    /* getunknp_local_Taylor: DOFs in each cell's local orthogonal Taylor basis */
    double unkpl[9][4][6];
    double mtbbq[25], mtbbi[25];
    for (int s = 0; s < 9; ++s) {
        int je = ielaj[s];
        double rcj = GE(c, 0, je), scj = GE(c, 1, je);
        double bb[25], bbq[25];
        for (int k = 0; k < 25; ++k) bb[k] = bbq[k] = 0.0;
        for (int g = 0; g < 25; ++g) {
            double r = E->pos[2 * g], sr = E->pos[2 * g + 1];
            double bq[6];
            basis_p2(r, sr, rcj, scj, &GE(c, 18, je), bq);
            double bl[6];
            bl[0] = 1.0;
            bl[1] = (P[s].xl[g] - geopl[s][0]) / dxcj[s];
            bl[2] = (P[s].yl[g] - geopl[s][1]) / dxcj[s];
            bl[3] = 0.5 * bl[1] * bl[1] - geopl[s][2];
            bl[4] = 0.5 * bl[2] * bl[2] - geopl[s][3];
            bl[5] = bl[1] * bl[2] - geopl[s][4];
            double rd = P[s].rhogi[g] * P[s].djacoi[g];
            for (int i = 1; i < 6; ++i)
                for (int j = 1; j < 6; ++j) {
                    bb[(i - 1) + 5 * (j - 1)] = bb[(i - 1) + 5 * (j - 1)] + rd * bq[i] * bl[j];
                    bbq[(i - 1) + 5 * (j - 1)] = bbq[(i - 1) + 5 * (j - 1)] + rd * bl[i] * bl[j];
                }
        }
        double rhs[4][6];
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 6; ++i) rhs[q][i] = 0.0;
        for (int i = 1; i < 6; ++i)
            for (int j = 1; j < 6; ++j)
                for (int q = 0; q < 4; ++q) rhs[q][i] = rhs[q][i] + unkrl[s][q][j] * bb[(j - 1) + 5 * (i - 1)];
        double bbqi[25], bbi[25];
        inv_small(5, bbq, bbqi);
        inv_small(5, bb, bbi);
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 6; ++i) unkpl[s][q][i] = 0.0;
        for (int i = 1; i < 6; ++i)
            for (int j = 1; j < 6; ++j)
                for (int q = 0; q < 4; ++q) unkpl[s][q][i] = unkpl[s][q][i] + rhs[q][j] * bbqi[(j - 1) + 5 * (i - 1)];
        if (s == 0) {
            memcpy(mtbbq, bbq, sizeof(bbq));
            memcpy(mtbbi, bbi, sizeof(bbi));
        }
        for (int q = 0; q < 4; ++q) unkpl[s][q][0] = unkrl[s][q][0];
    }

    /* candidate polynomials: 1 centre, 4 Lagrange, 4 Hermite */
    static const int ST1[4][5] = {{1, 4, 5, 8, 7}, {2, 1, 6, 5, 8}, {3, 2, 7, 6, 5}, {4, 3, 8, 7, 6}};
    static const int ST2[4][3] = {{1, 5, 8}, {2, 5, 6}, {3, 7, 6}, {4, 7, 8}};
    double unksl[9][4][6];
    memset(unksl, 0, sizeof(unksl));
    for (int s = 0; s < 9; ++s)
        for (int q = 0; q < 4; ++q) unksl[s][q][0] = unkpl[0][q][0];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 6; ++i) unksl[0][q][i] = unkpl[0][q][i];
    const double dxcl = sqrt(GE(c, 2, ie));
    const double xmcl = geopl[0][0], ymcl = geopl[0][1];
    int isten = 0;
    for (int is = 0; is < 4; ++is) {
        ++isten;
        double mat[25], rhsl[4][5];
        for (int k = 0; k < 25; ++k) mat[k] = 0.0;
        for (int e = 0; e < 5; ++e) {
            int sj = ST1[is][e];
            double masej = GE(c, 3, ielaj[sj]);
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
        inv_small(5, mat, mi);
        for (int i = 1; i < 6; ++i)
            for (int e = 0; e < 5; ++e)
                for (int q = 0; q < 4; ++q)
                    unksl[isten][q][i] = unksl[isten][q][i] + mi[(i - 1) + 5 * e] * rhsl[q][e];
    }
    for (int is = 0; is < 4; ++is) {
        ++isten;
        double mh[9 * 5], rh[4][9];
        for (int k = 0; k < 45; ++k) mh[k] = 0.0;
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < 9; ++k) rh[q][k] = 0.0;
        for (int e = 0; e < 3; ++e) {
            int sj = ST2[is][e];
            double masej = GE(c, 3, ielaj[sj]);
            double dxj = dxcj[sj];
            double xclj = geopl[sj][0], yclj = geopl[sj][1];
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
                int row = 3 + 2 * (e + 1) - 1 - 1;
                for (int i = 1; i < 6; ++i) {
                    mh[row + 9 * (i - 1)] = mh[row + 9 * (i - 1)] + rd * bl[i] * bj[1] / masej;
                    mh[row + 1 + 9 * (i - 1)] = mh[row + 1 + 9 * (i - 1)] + rd * bl[i] * bj[2] / masej;
                }
                for (int i = 1; i < 6; ++i)
                    for (int q = 0; q < 4; ++q) {
                        rh[q][row] = rh[q][row] + rd * bj[i] * bj[1] * unkpl[sj][q][i] / masej;
                        rh[q][row + 1] = rh[q][row + 1] + rd * bj[i] * bj[2] * unkpl[sj][q][i] / masej;
                    }
            }
            for (int q = 0; q < 4; ++q) rh[q][e] = unkpl[sj][q][0] - unkpl[0][q][0];
        }
        /* matrix_sym (LinearAlgebra.f90:89-126): normal equations */
        double ms[25], rs[4][5];
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                double a = 0.0;
                for (int k = 0; k < 9; ++k) a = a + mh[k + 9 * i] * mh[k + 9 * j];
                ms[i + 5 * j] = a;
            }
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 5; ++i) {
                double a = 0.0;
                for (int k = 0; k < 9; ++k) a = a + mh[k + 9 * i] * rh[q][k];
                rs[q][i] = a;
            }
        double mi[25];
        inv_small(5, ms, mi);
        for (int i = 1; i < 6; ++i)
            for (int e = 0; e < 5; ++e)
                for (int q = 0; q < 4; ++q)
                    unksl[isten][q][i] = unksl[isten][q][i] + mi[(i - 1) + 5 * e] * rs[q][e];
    }
    const int nst = isten + 1; /* 9 candidates */

    /* weno_cell_Taylor: characteristic WENO across each real face */
    double xp[2][9];
    q2_nodes(c, E->xc, ie, xp);
    const double volel = GE(c, 2, ie);
    const double dxc = sqrt(volel), dyc = dxc;
    double unkf_phy[4][4][6];
    memset(unkf_phy, 0, sizeof(unkf_phy));
    int ifprj = 0;
    static const int mapfe[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    for (int f = 0; f < 4; ++f) {
        int je = esqua[f + 4 * ie];
        if (je >= c->ncell) continue;
        double dtx = xp[0][mapfe[f][1]] - xp[0][mapfe[f][0]];
        double dty = xp[1][mapfe[f][1]] - xp[1][mapfe[f][0]];
        double dlgt = sqrt(dtx * dtx + dty * dty);
        double dnx = dty / dlgt, dny = -dtx / dlgt;
        rot(m, &dnx, &dny);
        double ul[4], ur[4];
        for (int q = 0; q < 4; ++q) {
            ul[q] = U(c, E->unknr, 0, q, ie);
            ur[q] = U(c, E->unknr, 0, q, je);
        }
        rot(m, &ul[1], &ul[2]);
        rot(m, &ur[1], &ur[2]);
        double qm[4][4], qi[4][4];
        char_matrices(c->gamlg, dnx, dny, ul, ur, qm, qi);
        double cha[9][4][6];
        memset(cha, 0, sizeof(cha));
        for (int s = 0; s < nst; ++s)
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < 6; ++id)
                    for (int k = 0; k < 4; ++k) cha[s][k][id] = cha[s][k][id] + qi[k][q] * unksl[s][q][id];
        double os1[9];
        for (int s = 0; s < nst; ++s) {
            double o[4];
            for (int q = 0; q < 4; ++q)
                o[q] = (cha[s][q][1] / dxc) * (cha[s][q][1] / dxc) + (cha[s][q][2] / dyc) * (cha[s][q][2] / dyc);
            os1[s] = dmax(dmax(dmax(o[0], o[1]), o[2]), o[3]);
        }
        double wl[9];
        wl[0] = .2;
        for (int s = 1; s < nst; ++s) wl[s] = 0.8 / ((double)nst - 1.0);
        const double epsil = 1.e-6;
        double wh[9], wt = 0.0;
        for (int s = 0; s < nst; ++s) {
            double t = epsil + os1[s];
            wh[s] = wl[s] / (t * t);
        }
        for (int s = 0; s < nst; ++s) wt = wt + wh[s];
        double fch[4][6];
        memset(fch, 0, sizeof(fch));
        for (int s = 0; s < nst; ++s)
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < 6; ++id) fch[q][id] = fch[q][id] + (wh[s] / wt) * cha[s][q][id];
        for (int q = 0; q < 4; ++q)
            for (int id = 1; id < 6; ++id)
                for (int k = 0; k < 4; ++k) unkf_phy[ifprj][k][id] = unkf_phy[ifprj][k][id] + qm[k][q] * fch[q][id];
        ++ifprj;
    }
    double unk_phy[4][6];
    memset(unk_phy, 0, sizeof(unk_phy));
    const double wf = 1.0 / (double)ifprj;
    for (int a = 0; a < ifprj; ++a)
        for (int q = 0; q < 4; ++q)
            for (int id = 1; id < 6; ++id) unk_phy[q][id] = unk_phy[q][id] + wf * unkf_phy[a][q][id];

    /* back to the cell's Taylor basis and the global frame */
    double rhsuk[4][5];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 5; ++i) rhsuk[q][i] = 0.0;
    for (int i = 1; i < 6; ++i)
        for (int j = 1; j < 6; ++j)
            for (int q = 0; q < 4; ++q) rhsuk[q][i - 1] = rhsuk[q][i - 1] + unk_phy[q][j] * mtbbq[(j - 1) + 5 * (i - 1)];
    double out[4][6];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 6; ++i) out[q][i] = 0.0;
    for (int i = 1; i < 6; ++i)
        for (int j = 1; j < 6; ++j)
            for (int q = 0; q < 4; ++q) out[q][i] = out[q][i] + rhsuk[q][j - 1] * mtbbi[(j - 1) + 5 * (i - 1)];
    for (int i = 1; i < 6; ++i) rotinv(m, &out[1][i], &out[2][i]);
    for (int q = 0; q < 4; ++q)
        for (int i = 1; i < 6; ++i) U(c, unkno, i, q, ie) = out[q][i];
    return 0;
}

/* weno_stncl_quad_orthol3 (WENO.f90:7541-8098) for npoly = 1, with
 * getgeophy_local_Taylor (:6750-7001), getunknp_local_Taylor (:7005-7215)
 * and weno_cell_Taylor (:7219-7537), nbsis = 1.  Three quirks of the
 * reference's P1 code decide its results and are reproduced:
 *  - the Lagrange stencils fit 3 equations in 2 unknowns (matlg is 3 x 2),
 *    and Matinv225 is called on matlg itself with n = 2, so its 2 x 2
 *    explicit-shape dummy sees the first four elements of matlg in memory
 *    (column-major): a = [matlg(1,1) matlg(3,1); matlg(2,1) matlg(1,2)]
 *    (WENO.f90:7855-7858); the update then uses the first 2 equations;
 *  - weno_cell_Taylor's npoly = 1 smoothness indicator is the P2 formula
 *    (WENO.f90:7375-7392): it reads unksf_cha(4:6, iq, is) of an array
 *    dimensioned (ndegr = 3, nq, nsten), i.e. the elements at the flat
 *    offsets 3..5 past (1, iq, is) -- the next variables' (or the next
 *    stencil's) leading entries, or zeros;
 *  - there are 5 candidates (centre + 4 Lagrange), weights 0.2 and 0.2. */
static int weno_cell_p1(const weno_env *E, int ie, const double m[4], const int *ielaj, double *unkno)
{
    const orc_case *c = E->c;
    double unkrl[9][4][3];
    for (int s = 0; s < 9; ++s)
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) unkrl[s][q][id] = U(c, E->unknr, id, q, ielaj[s]);
    for (int s = 0; s < 9; ++s)
        for (int id = 0; id < 3; ++id) rot(m, &unkrl[s][1][id], &unkrl[s][2][id]);

    /* getgeophy_local_Taylor: local centres, the P2 offsets (nbsis = 1) */
    double geopl[9][5];
    cell_pts P[9];
    double dxcj[9];
    for (int s = 0; s < 9; ++s) {
        int je = ielaj[s];
        double xcl = E->geoph[2 * je], ycl = E->geoph[2 * je + 1];
        rot(m, &xcl, &ycl);
        geopl[s][0] = xcl;
        geopl[s][1] = ycl;
        dxcj[s] = sqrt(GE(c, 2, je));
        cell_points(E, je, m, &P[s]);
        double v1 = 0.0, v2 = 0.0, v3 = 0.0;
        for (int g = 0; g < 25; ++g) {
            double b2 = (P[s].xl[g] - xcl) / dxcj[s];
            double b3 = (P[s].yl[g] - ycl) / dxcj[s];
            double rd = P[s].rhogi[g] * P[s].djacoi[g];
            v1 = v1 + rd * b2 * b2;
            v2 = v2 + rd * b3 * b3;
            v3 = v3 + rd * b2 * b3;
        }
        double masel = GE(c, 3, je);
        v1 = v1 / masel;
        v2 = v2 / masel;
        v3 = v3 / masel;
        geopl[s][2] = 0.5 * v1;
        geopl[s][3] = 0.5 * v2;
        geopl[s][4] = v3;
    }

    /* getunknp_local_Taylor with 2 x 2 matrices */
    double unkpl[9][4][3];
    double mtbbq[4], mtbbi[4];
    for (int s = 0; s < 9; ++s) {
        int je = ielaj[s];
        double rcj = GE(c, 0, je), scj = GE(c, 1, je);
        double bb[4] = {0.0, 0.0, 0.0, 0.0}, bbq[4] = {0.0, 0.0, 0.0, 0.0};
        for (int g = 0; g < 25; ++g) {
            double bq[3];
            bq[0] = 1.0;
            bq[1] = (E->pos[2 * g] - rcj) / 1.0;
            bq[2] = (E->pos[2 * g + 1] - scj) / 1.0;
            double bl[3];
            bl[0] = 1.0;
            bl[1] = (P[s].xl[g] - geopl[s][0]) / dxcj[s];
            bl[2] = (P[s].yl[g] - geopl[s][1]) / dxcj[s];
            double rd = P[s].rhogi[g] * P[s].djacoi[g];
            for (int i = 1; i < 3; ++i)
                for (int j = 1; j < 3; ++j) {
                    bb[(i - 1) + 2 * (j - 1)] = bb[(i - 1) + 2 * (j - 1)] + rd * bq[i] * bl[j];
                    bbq[(i - 1) + 2 * (j - 1)] = bbq[(i - 1) + 2 * (j - 1)] + rd * bl[i] * bl[j];
                }
        }
        double rhs[4][3];
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 3; ++i) rhs[q][i] = 0.0;
        for (int i = 1; i < 3; ++i)
            for (int j = 1; j < 3; ++j)
                for (int q = 0; q < 4; ++q) rhs[q][i] = rhs[q][i] + unkrl[s][q][j] * bb[(j - 1) + 2 * (i - 1)];
        double bbqi[4], bbi[4];
        inv2(bbq, bbqi);
        inv2(bb, bbi);
        for (int q = 0; q < 4; ++q)
            for (int i = 0; i < 3; ++i) unkpl[s][q][i] = 0.0;
        for (int i = 1; i < 3; ++i)
            for (int j = 1; j < 3; ++j)
                for (int q = 0; q < 4; ++q) unkpl[s][q][i] = unkpl[s][q][i] + rhs[q][j] * bbqi[(j - 1) + 2 * (i - 1)];
        if (s == 0) {
            memcpy(mtbbq, bbq, sizeof(bbq));
            memcpy(mtbbi, bbi, sizeof(bbi));
        }
        for (int q = 0; q < 4; ++q) unkpl[s][q][0] = unkrl[s][q][0];
    }

    /* candidates: the centre polynomial and 4 Lagrange stencils of 3 cells */
    static const int ST1[4][3] = {{1, 4, 5}, {2, 1, 6}, {3, 2, 7}, {4, 3, 8}};
    double unksl[5][4][3];
    memset(unksl, 0, sizeof(unksl));
    for (int s = 0; s < 5; ++s)
        for (int q = 0; q < 4; ++q) unksl[s][q][0] = unkpl[0][q][0];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 3; ++i) unksl[0][q][i] = unkpl[0][q][i];
    const double dxcl = sqrt(GE(c, 2, ie));
    const double xmcl = geopl[0][0], ymcl = geopl[0][1];
    for (int is = 0; is < 4; ++is) {
        const int isten = is + 1;
        double matlg[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}; /* matlg(3, 2), column-major */
        double rhsl[4][3];
        for (int e = 0; e < 3; ++e) {
            int sj = ST1[is][e];
            double masej = GE(c, 3, ielaj[sj]);
            for (int g = 0; g < 25; ++g) {
                double bl[3];
                bl[1] = (P[sj].xl[g] - xmcl) / dxcl;
                bl[2] = (P[sj].yl[g] - ymcl) / dxcl;
                double rd = P[sj].rhogi[g] * P[sj].djacoi[g];
                for (int k = 0; k < 2; ++k) matlg[e + 3 * k] = matlg[e + 3 * k] + rd * bl[k + 1] / masej;
            }
            for (int q = 0; q < 4; ++q) rhsl[q][e] = unkpl[sj][q][0] - unkpl[0][q][0];
        }
        /* Matinv225(matlg, matin, 2): the first four elements of matlg as a
         * column-major 2 x 2 */
        double mi[4];
        inv2(matlg, mi);
        for (int i = 1; i < 3; ++i)
            for (int e = 0; e < 2; ++e)
                for (int q = 0; q < 4; ++q)
                    unksl[isten][q][i] = unksl[isten][q][i] + mi[(i - 1) + 2 * e] * rhsl[q][e];
    }
    const int nst = 5;

    /* weno_cell_Taylor, npoly = 1 */
    double xp[2][9];
    q2_nodes(c, E->xc, ie, xp);
    const double volel = GE(c, 2, ie);
    const double dxc = sqrt(volel), dyc = dxc;
    const double ct1 = geopl[0][2], ct2 = geopl[0][3], ct3 = geopl[0][4];
    double unkf_phy[4][4][3];
    memset(unkf_phy, 0, sizeof(unkf_phy));
    int ifprj = 0;
    static const int mapfe[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    for (int f = 0; f < 4; ++f) {
        int je = E->esqua[f + 4 * ie];
        if (je >= c->ncell) continue;
        double dtx = xp[0][mapfe[f][1]] - xp[0][mapfe[f][0]];
        double dty = xp[1][mapfe[f][1]] - xp[1][mapfe[f][0]];
        double dlgt = sqrt(dtx * dtx + dty * dty);
        double dnx = dty / dlgt, dny = -dtx / dlgt;
        rot(m, &dnx, &dny);
        double ul[4], ur[4];
        for (int q = 0; q < 4; ++q) {
            ul[q] = U(c, E->unknr, 0, q, ie);
            ur[q] = U(c, E->unknr, 0, q, je);
        }
        rot(m, &ul[1], &ul[2]);
        rot(m, &ur[1], &ur[2]);
        double qm[4][4], qi[4][4];
        char_matrices(c->gamlg, dnx, dny, ul, ur, qm, qi);
        /* unksf_cha(ndegr = 3, nq = 4, nsten = 10), column-major, zeroed */
        double cha[120];
        memset(cha, 0, sizeof(cha));
#define CHA(id, k, s) cha[(id) + 3 * (k) + 12 * (s)]
        for (int s = 0; s < nst; ++s)
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < 3; ++id)
                    for (int k = 0; k < 4; ++k) CHA(id, k, s) = CHA(id, k, s) + qi[k][q] * unksl[s][q][id];
        double os1[5];
        /* x**4 is evaluated as ((x*x)*x)*x by the reference's build (the
         * squared form (x*x)*(x*x) differs in the last bit here) */
        const double dx2 = dxc * dxc, dy2 = dyc * dyc, dx4 = dx2 * dxc * dxc, dy4 = dy2 * dyc * dyc;
        const double dxy = dyc * dxc, dxy2 = dxy * dxy;
        for (int s = 0; s < nst; ++s) {
            double o[4];
            for (int q = 0; q < 4; ++q) {
                /* degrees 4..6 read past the first dimension (flat offsets) */
                const double u2 = CHA(1, q, s), u3 = CHA(2, q, s);
                const double u4 = cha[3 + 3 * q + 12 * s], u5 = cha[4 + 3 * q + 12 * s], u6 = cha[5 + 3 * q + 12 * s];
                double sm1 = 1.0 / dx2 * (u2 * u2 + 2.0 * ct1 * (u4 * u4) + 2.0 * ct2 * (u6 * u6) + 2.0 * ct3 * u4 * u6);
                double sm2 = 1.0 / dy2 * (u3 * u3 + 2.0 * ct2 * (u5 * u5) + 2.0 * ct1 * (u6 * u6) + 2.0 * ct3 * u5 * u6);
                double sm3 = 1.0 / dx4 * (u4 * u4);
                double sm4 = 1.0 / dy4 * (u5 * u5);
                double sm5 = 1.0 / dxy2 * (u6 * u6);
                o[q] = (sm1 + sm2) + (sm3 + sm4 + sm5) * volel;
            }
            os1[s] = dmax(dmax(dmax(o[0], o[1]), o[2]), o[3]);
        }
        double wl[5];
        wl[0] = .2;
        for (int s = 1; s < nst; ++s) wl[s] = 0.8 / ((double)nst - 1.0);
        const double epsil = 1.e-6;
        double wh[5], wt = 0.0;
        for (int s = 0; s < nst; ++s) {
            double t = epsil + os1[s];
            wh[s] = wl[s] / (t * t);
        }
        for (int s = 0; s < nst; ++s) wt = wt + wh[s];
        double fch[4][3];
        memset(fch, 0, sizeof(fch));
        for (int s = 0; s < nst; ++s)
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < 3; ++id) fch[q][id] = fch[q][id] + (wh[s] / wt) * CHA(id, q, s);
#undef CHA
        for (int q = 0; q < 4; ++q)
            for (int id = 1; id < 3; ++id)
                for (int k = 0; k < 4; ++k) unkf_phy[ifprj][k][id] = unkf_phy[ifprj][k][id] + qm[k][q] * fch[q][id];
        ++ifprj;
    }
    double unk_phy[4][3];
    memset(unk_phy, 0, sizeof(unk_phy));
    const double wf = 1.0 / (double)ifprj;
    for (int a = 0; a < ifprj; ++a)
        for (int q = 0; q < 4; ++q)
            for (int id = 1; id < 3; ++id) unk_phy[q][id] = unk_phy[q][id] + wf * unkf_phy[a][q][id];

    /* back to the cell's Taylor basis and the global frame */
    double rhsuk[4][2];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 2; ++i) rhsuk[q][i] = 0.0;
    for (int i = 1; i < 3; ++i)
        for (int j = 1; j < 3; ++j)
            for (int q = 0; q < 4; ++q) rhsuk[q][i - 1] = rhsuk[q][i - 1] + unk_phy[q][j] * mtbbq[(j - 1) + 2 * (i - 1)];
    double out[4][3];
    for (int q = 0; q < 4; ++q)
        for (int i = 0; i < 3; ++i) out[q][i] = 0.0;
    for (int i = 1; i < 3; ++i)
        for (int j = 1; j < 3; ++j)
            for (int q = 0; q < 4; ++q) out[q][i] = out[q][i] + rhsuk[q][j - 1] * mtbbi[(j - 1) + 2 * (i - 1)];
    for (int i = 1; i < 3; ++i) rotinv(m, &out[1][i], &out[2][i]);
    for (int q = 0; q < 4; ++q)
        for (int i = 1; i < 3; ++i) U(c, unkno, i, q, ie) = out[q][i];
    return 0;
}

/* weno_char_quad_orthgl (WENO.f90:5927-6120) */
static int weno(orc_case *c, const double *flag, const int *esqua)
{
    weno_env E;
    E.c = c;
    E.xc = c->cooro;
    E.x0 = c->coold;
    ruqope2(5, E.pos, E.w);
    const int nc = c->ncell, nd = c->ndegr;
    double *geoph = (double *)malloc(sizeof(double) * 2 * nc);
    double *mapmt = (double *)malloc(sizeof(double) * 4 * nc);
    ORC_PHASE(ORC_PH_WENO_MC);
    for (int ie = c->ntria; ie < nc; ++ie) {
        double xp[2][9], xi[2][9];
        q2_nodes(c, c->cooro, ie, xp);
        q2_nodes(c, c->coold, ie, xi);
        double xcr, ycr;
        q2_map(xi, GE(c, 0, ie), GE(c, 1, ie), &xcr, &ycr);
        double masel = GE(c, 3, ie);
        double xmc = 0.0, ymc = 0.0;
        for (int g = 0; g < 25; ++g) {
            double r = E.pos[2 * g], s = E.pos[2 * g + 1], wi = E.w[g];
            double J[4];
            q2_jac(xi, r, s, J, NULL, NULL);
            double djacoi = wi * (J[0] * J[3] - J[2] * J[1]);
            double xg, yg;
            q2_map(xp, r, s, &xg, &yg);
            double rho = c->rho0 ? c->rho0[ie] : rho_initial(c, xcr, ycr);
            xmc = xmc + rho * djacoi * xg;
            ymc = ymc + rho * djacoi * yg;
        }
        geoph[2 * ie] = xmc / masel;
        geoph[2 * ie + 1] = ymc / masel;
        double a = U(c, c->unold, 0, 1, ie), b = U(c, c->unold, 0, 2, ie);
        double *m = mapmt + 4 * ie;
        if (sqrt(a * a + b * b) > 1.e-10) {
            m[0] = a / sqrt(a * a + b * b);
            m[1] = b / sqrt(a * a + b * b);
            m[2] = -b / sqrt(a * a + b * b);
            m[3] = a / sqrt(a * a + b * b);
        } else {
            m[0] = m[1] = m[2] = m[3] = 0.0;
        }
    }
    double *unknr = (double *)malloc(sizeof(double) * nd * 4 * nc);
    memcpy(unknr, c->unold, sizeof(double) * nd * 4 * nc);
    E.geoph = geoph;
    E.unknr = unknr;
    int bad = 0;
    ORC_PHASE(ORC_PH_WENO);
    E.esqua = esqua;
    for (int ie = c->ntria; ie < nc; ++ie) {
        if (flag[ie] < 10.0) continue;
        /* On a hybrid mesh the reference reads the face-neighbour table by
         * the cell's global number, not its quad number (WENO.f90:7659), so
         * a troubled cell's stencil is undefined there: such cells are
         * reported (strict) or keep their DOFs. */
        if (c->ntria > 0) {
            ++bad;
            c->nweno_skipped++;
            continue;
        }
        /* A cell at rest has the zero frame (WENO.f90:6066-6081): every
         * stencil point maps to the origin, the normal equations vanish and
         * Matinv225 returns 0/0.  Strict mode keeps that arithmetic (the
         * reference's NaN); otherwise the cell keeps its DOFs, like a
         * boundary stencil. */
        const double *mz = mapmt + 4 * ie;
        if (!c->weno_strict && mz[0] == 0.0 && mz[1] == 0.0 && mz[2] == 0.0 && mz[3] == 0.0) {
            c->nweno_rest++;
            continue;
        }
        ORC_WENO_CELL();
        int r;
        if (c->npoly == 1) {
            int ielaj[9];
            r = stencil_cells(c, ie, esqua, ielaj);
            if (r == 0) weno_cell_p1(&E, ie, mapmt + 4 * ie, ielaj, c->unold);
            if (r == 2) r = 0;
        } else {
            r = weno_cell(&E, ie, mapmt + 4 * ie, esqua, c->unold);
        }
        if (r) {
            ++bad;
            c->nweno_skipped++;
        }
    }
    for (size_t i = 0; i < (size_t)nd * 4 * nc; ++i) ORC_FROZEN_INCREMENT(c->unold[i], unknr[i]);
    free(geoph);
    free(mapmt);
    free(unknr);
    return (bad && c->weno_strict) ? 6 : 0;
}

/* -------------------------------------------------- sub-cell densities -- */
/* getrhosubg_averg (Getrhslaggauss.f90:7067-7392): sub-cell density =
 * initial sub-cell mass / current sub-cell area */
static void subcell_density(const orc_case *c, int ie, double rsub[4])
{
    static const double posqc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                                        {-1.0, .5}, {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
    static const int ipqsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
    double pos[50], w[25];
    ruqope2(5, pos, w);
    double xp[2][9], xi[2][9];
    q2_nodes(c, c->cooro, ie, xp);
    q2_nodes(c, c->coold, ie, xi);
    double xc[2][12], xic[2][12];
    for (int k = 0; k < 12; ++k) {
        q2_map(xp, posqc[k][0], posqc[k][1], &xc[0][k], &xc[1][k]);
        ORC_PHASE(ORC_PH_HOISTED);
        q2_map(xi, posqc[k][0], posqc[k][1], &xic[0][k], &xic[1][k]);
        ORC_PHASE(ORC_PH_CORNER);
    }
    double f0[4];
    for (int isg = 0; isg < 4; ++isg) {
        double xs[2][9];
        for (int d = 0; d < 2; ++d) {
            for (int k = 0; k < 4; ++k) xs[d][k] = xp[d][IPQSG[isg][k]];
            for (int k = 0; k < 4; ++k) xs[d][4 + k] = xc[d][ipqsc[isg][k]];
            xs[d][8] = -0.25 * (xs[d][0] + xs[d][1] + xs[d][2] + xs[d][3]) + 0.5 * (xs[d][4] + xs[d][5] + xs[d][6] + xs[d][7]);
        }
        double a = 0.0;
        for (int g = 0; g < 25; ++g) {
            double J[4];
            q2_jac(xs, pos[2 * g], pos[2 * g + 1], J, NULL, NULL);
            a = a + w[g] * (J[0] * J[3] - J[2] * J[1]);
        }
        f0[isg] = 1.0 / a;
    }
    ORC_PHASE(ORC_PH_HOISTED);
    for (int isg = 0; isg < 4; ++isg) {
        double xs[2][9];
        for (int d = 0; d < 2; ++d) {
            for (int k = 0; k < 4; ++k) xs[d][k] = xi[d][IPQSG[isg][k]];
            for (int k = 0; k < 4; ++k) xs[d][4 + k] = xic[d][ipqsc[isg][k]];
            xs[d][8] = -0.25 * (xs[d][0] + xs[d][1] + xs[d][2] + xs[d][3]) + 0.5 * (xs[d][4] + xs[d][5] + xs[d][6] + xs[d][7]);
        }
        double ms = 0.0;
        for (int g = 0; g < 25; ++g) {
            double J[4], xg, yg;
            q2_jac(xs, pos[2 * g], pos[2 * g + 1], J, &xg, &yg);
            double djaco = w[g] * (J[0] * J[3] - J[2] * J[1]);
            double xcc, ycc;
            q2_map(xi, 0.0, 0.0, &xcc, &ycc);
            double rhoi = c->rho0 ? c->rho0[ie] : ((c->ncase == 6) ? rho_initial(c, xcc, ycc) : 1.0);
            ms = ms + rhoi * 1.0 * djaco;
        }
        ORC_PHASE(ORC_PH_CORNER);
        rsub[isg] = 0.0 + f0[isg] * ms;
        ORC_PHASE(ORC_PH_HOISTED);
    }
    ORC_PHASE(ORC_PH_CORNER);
}

/* getrhosubg_daverg (Getrhslaggauss.f90:10949-11109): mass-weighted mean of
 * tau over each quarter of the reference square, 1/tau as density. */
static void subcell_tau(const orc_case *c, int ie, double uq[4])
{
    double pos[50], w[25];
    ruqope2(5, pos, w);
    double xp[2][9];
    q2_nodes(c, c->cooro, ie, xp);
    double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
    double rhomc = U(c, c->unold, 0, 0, ie);
    static const double sr[4] = {-1.0, 1.0, 1.0, -1.0}, ss[4] = {-1.0, -1.0, 1.0, 1.0};
    for (int isg = 0; isg < 4; ++isg) {
        double masel = 0.0, volel = 0.0;
        for (int g = 0; g < 25; ++g) {
            double r = (pos[2 * g] + sr[isg]) / 2.0;
            double s = (pos[2 * g + 1] + ss[isg]) / 2.0;
            double b[6];
            basis_p2(r, s, rc, sc, &GE(c, 18, ie), b);
            double t = 0.0;
            for (int id = 0; id < c->ndegr; ++id) t = t + U(c, c->unold, id, 0, ie) * b[id];
            t = rhomc + 1.0 * (t - rhomc); /* nlimi=6 with aflim(1)=1 */
            double J[4];
            q2_jac(xp, r, s, J, NULL, NULL);
            double djaco = w[g] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            masel = masel + 1.0 / t * djaco;
            volel = volel + djaco;
        }
        uq[isg] = volel / masel;
    }
}


/* ------------------------------------------------------------------------ */
/* Curved six-node triangles (hybrid meshes): cells 0 .. ntria-1, nodes
 * iptri[k + 6 it], corners 0-2 counter-clockwise then the mid-nodes of
 * edges 1-2, 2-3, 3-1. */
#define TIP(c, it) ((c)->iptri + 6 * (it))

/* sub-cell vertex -> local node (iptsg, Getrhslaggauss.f90:12262-12265) */
static const int IPTSG[4][3] = {{0, 3, 5}, {3, 4, 5}, {3, 1, 4}, {5, 4, 2}};
/* signed face-normal slots (fntsg, :12267-12285): [isg][2 ivsg + f] */
static const int FNTSG[4][6] = {{6, 1, 10, -7, -13, 12}, {7, 14, 8, 15, 9, 13},
                                {-14, 10, 2, 3, 11, -8}, {12, -9, -15, 11, 4, 5}};
/* corner weights times 1/4 (wfgst, :12287-12306; the volume weights of the
 * face integral, wfgstm, Getrhslaggauss.f90:8349-8367, are the same):
 * [isg][ivsg][f] */
static const double WT2[4][3][2] = {
    {{0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}},
    {{0.25 * (8.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * (8.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * (8.0 / 3.0), 0.25 * (8.0 / 3.0)}},
    {{0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}},
    {{0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}}};
/* reference coordinates of the six nodes */
static const double XV6[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
static const double YV6[6] = {0.0, 0.0, 1.0, 0.0, 0.5, 0.5};

/* rutope's 7-point area rule (dgflo2d.f90:2234-2255), as the triangle
 * routines read it with ngausd = 7 */
static const double PT7Q[7][2] = {
    {0.333333333333333, 0.333333333333333}, {0.059715871789770, 0.470142064105115},
    {0.470142064105115, 0.059715871789770}, {0.470142064105115, 0.470142064105115},
    {0.797426985353087, 0.101286507323456}, {0.101286507323456, 0.797426985353087},
    {0.101286507323456, 0.101286507323456}};
static const double WT7Q[7] = {0.225000000000000, 0.132394152788506, 0.132394152788506, 0.132394152788506,
                               0.125939180544827, 0.125939180544827, 0.125939180544827};

/* quadratic triangle shape functions and reference derivatives
 * (Getrhslaggauss.f90:9501-9523) */
static void shape_t2(double r, double s, double sh[6], double dr_[6], double ds_[6])
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

static void t2_nodes(const orc_case *c, const double *x, int it, double xp[2][6])
{
    const int *ip = TIP(c, it);
    for (int k = 0; k < 6; ++k) {
        xp[0][k] = X(x, 0, ip[k]);
        xp[1][k] = X(x, 1, ip[k]);
    }
}

static void t2_map(const double xp[2][6], double r, double s, double *x, double *y)
{
    double sh[6], dr_[6], ds_[6];
    shape_t2(r, s, sh, dr_, ds_);
    double a = 0.0, b = 0.0;
    for (int k = 0; k < 6; ++k) {
        a = a + sh[k] * xp[0][k];
        b = b + sh[k] * xp[1][k];
    }
    *x = a;
    *y = b;
}

static void t2_jac(const double xp[2][6], double r, double s, double J[4])
{
    double sh[6], dr_[6], ds_[6];
    shape_t2(r, s, sh, dr_, ds_);
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 6; ++k) {
        dxdr = dxdr + dr_[k] * xp[0][k];
        dxds = dxds + ds_[k] * xp[0][k];
        dydr = dydr + dr_[k] * xp[1][k];
        dyds = dyds + ds_[k] * xp[1][k];
    }
    J[0] = dxdr; J[1] = dxds; J[2] = dydr; J[3] = dyds;
}

/* P2 Taylor basis of a triangle about its mass centre (dr = ds = 0.5) */
static void basis_t2(double r, double s, double rc, double sc, const double *g19, double b[6])
{
    b[0] = 1.0;
    b[1] = (r - rc) / 0.5;
    b[2] = (s - sc) / 0.5;
    b[3] = 0.5 * b[1] * b[1] - g19[0];
    b[4] = 0.5 * b[2] * b[2] - g19[1];
    b[5] = b[1] * b[2] - g19[2];
}

/* Triangle face geometry gt[(k + 3 slot) + 45 it], 15 slots: the interior
 * sub-faces between mid-nodes at both ends (getfnds_lagsmsif_hybrid,
 * Getrhslaggauss.f90:3419-3496, slots 6-8 and 12-14) and the boundary-edge
 * points by the cofactor of the map's Jacobian (getfnds_lagsms_simpsonhybrid,
 * :9726-9822, slots 0-5 and 9-11). */
static void fnds_t2(const orc_case *c, const double *x, double *gt)
{
    static const double postc[3][2] = {{0.25, 0.25}, {0.5, 0.25}, {0.25, 0.5}};
    static const int ipift[3][3] = {{3, 5, 0}, {4, 3, 1}, {5, 4, 2}};
    static const double posit[9][2] = {{0, 0}, {1, 0}, {1, 0}, {0, 1}, {0, 1}, {0, 0}, {.5, 0}, {.5, .5}, {0, .5}};
    static const int slot[9] = {0, 1, 2, 3, 4, 5, 9, 10, 11};
    for (int it = 0; it < c->ntria; ++it) {
        double xp[2][6];
        t2_nodes(c, x, it, xp);
        double *g = gt + 45 * it;
        double xv[3], yv[3];
        for (int k = 0; k < 3; ++k) t2_map(xp, postc[k][0], postc[k][1], &xv[k], &yv[k]);
        for (int f = 0; f < 3; ++f) {
            double fx[3] = {xp[0][ipift[f][0]], xp[0][ipift[f][1]], xv[ipift[f][2]]};
            double fy[3] = {xp[1][ipift[f][0]], xp[1][ipift[f][1]], yv[ipift[f][2]]};
            for (int e = 0; e < 2; ++e) {
                const double r = e == 0 ? -1.0 : 1.0;
                const double d[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
                double dxdr = 0.0, dydr = 0.0;
                for (int i = 0; i < 3; ++i) {
                    dxdr = dxdr + d[i] * fx[i];
                    dydr = dydr + d[i] * fy[i];
                }
                const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
                const double anx = -dydr, any = dxdr;
                const int sl = (e == 0 ? 6 : 12) + f;
                g[3 * sl + 0] = anx / sqrt(anx * anx + any * any);
                g[3 * sl + 1] = any / sqrt(anx * anx + any * any);
                g[3 * sl + 2] = djaco * 2.0;
            }
        }
        for (int ig = 0; ig < 9; ++ig) {
            double J[4];
            t2_jac(xp, posit[ig][0], posit[ig][1], J);
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
            g[3 * slot[ig] + 0] = anx * larea / farea;
            g[3 * slot[ig] + 1] = any * larea / farea;
            g[3 * slot[ig] + 2] = farea;
        }
    }
}

/* Initial density at a point of a curved triangle (getrhoig_triacurv2,
 * Getrhslag.f90:17622-17757): the Sedov problem's uniform density, or the
 * per-cell initial density when given. */
static double rho_initial_t2(const orc_case *c, int it)
{
    return c->rho0 ? c->rho0[it] : 1.0;
}

/* getrhosubgt_average2 (Getrhslaggauss.f90:12595-12940): per sub-triangle
 * the density initial mass / current area (dens) and the mass-weighted mean
 * of tau as a density (tau) */
static void subcell_t2(const orc_case *c, int it, double dens[4], double tau[4])
{
    static const double postc[9][2] = {{.25, 0.}, {.75, 0.}, {.75, .25}, {.25, .75}, {0., .75},
                                       {0., .25}, {.25, .25}, {.5, .25}, {.25, .5}};
    static const int iptsc[4][3] = {{0, 6, 5}, {7, 8, 6}, {1, 2, 7}, {8, 3, 4}};
    double xp[2][6], xi[2][6];
    t2_nodes(c, c->cooro, it, xp);
    t2_nodes(c, c->coold, it, xi);
    double xc[2][9], xic[2][9];
    for (int k = 0; k < 9; ++k) {
        t2_map(xp, postc[k][0], postc[k][1], &xc[0][k], &xc[1][k]);
        t2_map(xi, postc[k][0], postc[k][1], &xic[0][k], &xic[1][k]);
    }
    const double rc = GE(c, 0, it), sc = GE(c, 1, it);
    const double rhomc = U(c, c->unold, 0, 0, it);
    double f0[4];
    for (int isg = 0; isg < 4; ++isg) {
        double xs[2][6];
        for (int d = 0; d < 2; ++d) {
            for (int k = 0; k < 3; ++k) xs[d][k] = xp[d][IPTSG[isg][k]];
            for (int k = 0; k < 3; ++k) xs[d][3 + k] = xc[d][iptsc[isg][k]];
        }
        double a = 0.0, masel = 0.0, volel = 0.0;
        for (int g = 0; g < 7; ++g) {
            const double r = PT7Q[g][0], s = PT7Q[g][1], wi = WT7Q[g];
            const double shs[3] = {1.0 - r - s, r, s};
            double rg = 0.0, sg = 0.0;
            for (int k = 0; k < 3; ++k) {
                rg = rg + shs[k] * XV6[IPTSG[isg][k]];
                sg = sg + shs[k] * YV6[IPTSG[isg][k]];
            }
            double b[6];
            basis_t2(rg, sg, rc, sc, &GE(c, 18, it), b);
            double t = 0.0;
            for (int id = 0; id < c->ndegr; ++id) t = t + U(c, c->unold, id, 0, it) * b[id];
            t = rhomc + 1.0 * (t - rhomc); /* nlimi=6 with aflim(1) = 1 */
            double J[4];
            t2_jac(xs, r, s, J);
            const double djaco = 0.5 * wi * (J[0] * J[3] - J[2] * J[1]);
            a = a + djaco;
            masel = masel + 1.0 / t * djaco;
            volel = volel + djaco;
        }
        f0[isg] = 1.0 / a;
        tau[isg] = volel / masel;
    }
    for (int isg = 0; isg < 4; ++isg) {
        double xs[2][6];
        for (int d = 0; d < 2; ++d) {
            for (int k = 0; k < 3; ++k) xs[d][k] = xi[d][IPTSG[isg][k]];
            for (int k = 0; k < 3; ++k) xs[d][3 + k] = xic[d][iptsc[isg][k]];
        }
        double ms = 0.0;
        for (int g = 0; g < 7; ++g) {
            double J[4];
            t2_jac(xs, PT7Q[g][0], PT7Q[g][1], J);
            const double djaco = 0.5 * WT7Q[g] * (J[0] * J[3] - J[2] * J[1]);
            ms = ms + rho_initial_t2(c, it) * 1.0 * djaco;
        }
        dens[isg] = 0.0 + f0[isg] * ms;
    }
}

/* getriem_triasubg3 (Getrhslaggauss.f90:12201-12590): the triangle's
 * sub-cell corners, scattered into the node sums and recorded per corner
 * face (mcl / mul / snl at o = f + 2 (ivsg + 3 isg), 24 per triangle) */
static void corners_t2(orc_case *c, const double *gt, double *mn, double *mu, double *sn, double *mcl, double *mul,
                       double *snl)
{
    const double g1 = c->gamlg - 1.0, eps = 1.0e-15;
    const int nd = c->ndegr;
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = TIP(c, it);
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        double un[6][4];
        for (int iv = 0; iv < 6; ++iv) {
            double b[6];
            basis_t2(XV6[iv], YV6[iv], rc, sc, &GE(c, 18, it), b);
            for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
            for (int id = 0; id < nd; ++id)
                for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + U(c, c->unold, id, q, it) * b[id];
        }
        const double rhomc = U(c, c->unold, 0, 0, it);
        const double uctr = U(c, c->unold, 0, 1, it), vctr = U(c, c->unold, 0, 2, it), ectr = U(c, c->unold, 0, 3, it);
        const double rhoct = 1.0 / rhomc;
        const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double dens[4], tau[4];
        subcell_t2(c, it, dens, tau);
        const double sdctr = sqrt(dmax(eps, c->gamlg * pctr / rhoct));
        const double *g = gt + 45 * it;
        for (int isg = 0; isg < 4; ++isg) {
            double vn[3][2][3];
            for (int iv = 0; iv < 3; ++iv)
                for (int f = 0; f < 2; ++f) {
                    const int t = FNTSG[isg][2 * iv + f];
                    const int sl = abs(t) - 1;
                    const double sg = t > 0 ? 1.0 : -1.0;
                    vn[iv][f][0] = sg * g[3 * sl];
                    vn[iv][f][1] = sg * g[3 * sl + 1];
                    vn[iv][f][2] = g[3 * sl + 2] * WT2[isg][iv][f];
                }
            for (int iv = 0; iv < 3; ++iv) {
                const int ln = IPTSG[isg][iv];
                const int p = ip[ln];
                const double rhovt = 1.0 / un[ln][0];
                const double rhovsg = rhovt + 1.0 * (dens[isg] - 1.0 / tau[isg]);
                const double uv = un[ln][1], vv = un[ln][2], ev = un[ln][3];
                const double pv = dmax(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)));
                const double dux = X(c->ustar, 0, p) - uv, duy = X(c->ustar, 1, p) - vv;
                const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
                for (int f = 0; f < 2; ++f) {
                    const double deltu = 0.0 * fabs(dux * vn[iv][f][0] + duy * vn[iv][f][1]);
                    const double mur = rhoct * sdctr + rhoct * c->slpdu * deltu;
                    const double area = vn[iv][f][2];
                    double m11;
                    if (a3 < 1.e12) {
                        m11 = mur * area * fabs(1.0);
                    } else {
                        const double nrm = sqrt(dux * dux + duy * duy);
                        m11 = mur * area * fabs(vn[iv][f][0] * (dux / nrm) + vn[iv][f][1] * (duy / nrm));
                    }
                    const double mu1 = m11 * uv + 0.0 * vv, mu2 = 0.0 * uv + m11 * vv;
                    const double sn1 = -pv * area * vn[iv][f][0] + 0.0 * area * vn[iv][f][1];
                    const double sn2 = 0.0 * area * vn[iv][f][0] + -pv * area * vn[iv][f][1];
                    mn[0 + 4 * p] += m11;
                    mn[1 + 4 * p] += 0.0;
                    mn[2 + 4 * p] += 0.0;
                    mn[3 + 4 * p] += m11;
                    mu[2 * p] += mu1;
                    mu[2 * p + 1] += mu2;
                    sn[2 * p] += sn1;
                    sn[2 * p + 1] += sn2;
                    const int o = f + 2 * (iv + 3 * isg);
                    mcl[o + 24 * it] = m11;
                    mul[2 * o + 48 * it] = mu1;
                    mul[2 * o + 1 + 48 * it] = mu2;
                    snl[2 * o + 48 * it] = sn1;
                    snl[2 * o + 1 + 48 * it] = sn2;
                }
            }
        }
    }
}

/* rhsifacedg_lagsubgt3 (Getrhslaggauss.f90:8297-8496): momentum and energy
 * weights 1, volume weights wfgstm; overwrites rhsel */
static void iface_t2(orc_case *c, const double *gt, const double *fstart)
{
    const int nd = c->ndegr;
    for (int it = 0; it < c->ntria; ++it) {
        const int *ip = TIP(c, it);
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        double bt[6][6];
        for (int iv = 0; iv < 6; ++iv) basis_t2(XV6[iv], YV6[iv], rc, sc, &GE(c, 18, it), bt[iv]);
        double ul[6] = {0}, p1[6] = {0}, p2[6] = {0}, el[6] = {0};
        const double *g = gt + 45 * it;
        for (int isg = 0; isg < 4; ++isg) {
            double vn[3][2][3];
            for (int iv = 0; iv < 3; ++iv)
                for (int f = 0; f < 2; ++f) {
                    const int t = FNTSG[isg][2 * iv + f];
                    const int sl = abs(t) - 1;
                    const double sg = t > 0 ? 1.0 : -1.0;
                    vn[iv][f][0] = sg * g[3 * sl];
                    vn[iv][f][1] = sg * g[3 * sl + 1];
                    vn[iv][f][2] = g[3 * sl + 2] * WT2[isg][iv][f];
                }
            for (int iv = 0; iv < 3; ++iv) {
                const int ln = IPTSG[isg][iv];
                const int p = ip[ln];
                const double us1 = X(c->ustar, 0, p), us2 = X(c->ustar, 1, p);
                const int o0 = 0 + 2 * (iv + 3 * isg), o1 = 1 + 2 * (iv + 3 * isg);
                const double *f0 = fstart + 2 * o0 + 48 * it, *f1 = fstart + 2 * o1 + 48 * it;
                for (int id = 0; id < nd; ++id) {
                    const double b = bt[ln][id];
                    ul[id] = ul[id] + us1 * vn[iv][0][0] * vn[iv][0][2] * b + us2 * vn[iv][0][1] * vn[iv][0][2] * b +
                             us1 * vn[iv][1][0] * vn[iv][1][2] * b + us2 * vn[iv][1][1] * vn[iv][1][2] * b;
                    p1[id] = p1[id] + f0[0] * b * 1.0 + f1[0] * b * 1.0;
                    p2[id] = p2[id] + f0[1] * b * 1.0 + f1[1] * b * 1.0;
                    el[id] = el[id] + us1 * f0[0] * b * 1.0 + us2 * f0[1] * b * 1.0 + us1 * f1[0] * b * 1.0 +
                             us2 * f1[1] * b * 1.0;
                }
            }
        }
        for (int id = 0; id < nd; ++id) {
            U(c, c->rhsel, id, 0, it) = ul[id];
            U(c, c->rhsel, id, 1, it) = p1[id];
            U(c, c->rhsel, id, 2, it) = p2[id];
            U(c, c->rhsel, id, 3, it) = el[id];
        }
    }
}

/* rhsdomndg_lagmc_curvtria (Getrhslag.f90:4449-4711): 7-point rule,
 * djaco = w/2 with the cofactor form of the basis gradients; the pressure
 * is not clamped here */
static void domn_t2(orc_case *c)
{
    const double g1 = c->gamlg - 1.0;
    const int nd = c->ndegr;
    for (int it = 0; it < c->ntria; ++it) {
        double xp[2][6];
        t2_nodes(c, c->cooro, it, xp);
        const double rc = GE(c, 0, it), sc = GE(c, 1, it);
        for (int g = 0; g < 7; ++g) {
            const double r = PT7Q[g][0], s = PT7Q[g][1], wi = WT7Q[g];
            double J[4];
            t2_jac(xp, r, s, J);
            const double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
            const double djaco = 0.5 * wi;
            const double dr = 0.5, ds = 0.5;
            const double dbdr[6] = {0.0, 1.0 / dr, 0.0, (r - rc) / (dr * dr), 0.0, (s - sc) / dr / ds};
            const double dbds[6] = {0.0, 0.0, 1.0 / ds, 0.0, (s - sc) / (ds * ds), (r - rc) / dr / ds};
            double b[6];
            basis_t2(r, s, rc, sc, &GE(c, 18, it), b);
            double un[4] = {0, 0, 0, 0};
            for (int id = 0; id < nd; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + U(c, c->unold, id, q, it) * b[id];
            const double rhoad = 1.0 / un[0];
            const double u = un[1], v = un[2];
            const double pres = g1 * rhoad * (un[3] - 0.5 * (u * u + v * v));
            for (int id = 0; id < nd; ++id) {
                const double a = dyds * dbdr[id] + (-dydr) * dbds[id];
                const double cc = (-dxds) * dbdr[id] + dxdr * dbds[id];
                const double f0 = a * u + cc * v;
                const double f1 = a * (-pres);
                const double f2 = cc * (-pres);
                const double f3 = (a * u + cc * v) * (-pres);
                U(c, c->rhsel, id, 0, it) = U(c, c->rhsel, id, 0, it) - f0 * djaco;
                U(c, c->rhsel, id, 1, it) = U(c, c->rhsel, id, 1, it) - f1 * djaco;
                U(c, c->rhsel, id, 2, it) = U(c, c->rhsel, id, 2, it) - f2 * djaco;
                U(c, c->rhsel, id, 3, it) = U(c, c->rhsel, id, 3, it) - f3 * djaco;
            }
        }
    }
}

/* sub-cell nodal solver (getndvelo_lagsubg + getriem_quadsubg2) and the
 * corner forces fstar[(d + 2*(f + 2*(ivsg + 4*isg))) + 32*ie] */
static void ndvelo_q2(orc_case *c, const double *gq, double *fstar, const double *gt, double *fstart)
{
    const int np = c->npoin, nc = c->ncell, nd = c->ndegr;
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    double wr[4][4][2];
    wfg_riem(wr);
    double *mn = (double *)calloc((size_t)4 * np, sizeof(double));
    double *mu = (double *)calloc((size_t)2 * np, sizeof(double));
    double *sn = (double *)calloc((size_t)2 * np, sizeof(double));
    ORC_PHASE(ORC_PH_CORNER);
    double *mcl = (double *)malloc(sizeof(double) * 32 * nc);  /* m = mu*A per corner face */
    double *mul = (double *)malloc(sizeof(double) * 64 * nc);  /* m u_c */
    double *snl = (double *)malloc(sizeof(double) * 64 * nc);  /* A sigma n */
    /* triangles first, as the reference accumulates (Getrhslaggauss.f90:4476-4481) */
    const int nt = c->ntria;
    double *mclt = (double *)malloc(sizeof(double) * (24 * (size_t)nt + 1));
    double *mult = (double *)malloc(sizeof(double) * (48 * (size_t)nt + 1));
    double *snlt = (double *)malloc(sizeof(double) * (48 * (size_t)nt + 1));
    corners_t2(c, gt, mn, mu, sn, mclt, mult, snlt);
    for (int ie = c->ntria; ie < nc; ++ie) {
        const int *ip = QIP(c, ie);
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double un[9][4];
        for (int iv = 0; iv < 9; ++iv) {
            double b[6];
            basis_p2(XV9[iv], YV9[iv], rc, sc, &GE(c, 18, ie), b);
            for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
            for (int id = 0; id < nd; ++id)
                for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + U(c, c->unold, id, q, ie) * b[id];
        }
        double rhomc = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie), ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhomc;
        double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
        double rsub[4], uq[4];
        subcell_density(c, ie, rsub);
        subcell_tau(c, ie, uq);
        const double *g = gq + 48 * ie;
        double sdctr = sqrt(dmax(eps, c->gamlg * pctr / rhoct));
        for (int isg = 0; isg < 4; ++isg) {
            double vn[4][2][3];
            for (int iv = 0; iv < 4; ++iv)
                for (int f = 0; f < 2; ++f) {
                    int t = FNQSG[isg][2 * iv + f];
                    int sl = abs(t) - 1;
                    double sg = t > 0 ? 1.0 : -1.0;
                    vn[iv][f][0] = sg * g[3 * sl];
                    vn[iv][f][1] = sg * g[3 * sl + 1];
                    vn[iv][f][2] = g[3 * sl + 2] * wr[isg][iv][f];
                }
            for (int iv = 0; iv < 4; ++iv) {
                int ln = IPQSG[isg][iv];
                int p = ip[ln];
                double rhovt = 1.0 / un[ln][0];
                double rhovsg = rhovt + c->cdrho * (rsub[isg] - 1.0 / uq[isg]);
                double uv = un[ln][1], vv = un[ln][2], ev = un[ln][3];
                double pv = dmax(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)));
                double dux = X(c->ustar, 0, p) - uv, duy = X(c->ustar, 1, p) - vv;
                double a3 = sqrt(dux * dux + duy * duy) / sdctr;
                for (int f = 0; f < 2; ++f) {
                    double deltu = ORC_FROZEN(fabs(dux * vn[iv][f][0] + duy * vn[iv][f][1]));
                    double mur = rhoct * sdctr + c->cimpd * rhoct * c->slpdu * deltu;
                    double area = vn[iv][f][2];
                    double m11;
                    if (a3 < 1.e12) {
                        m11 = mur * area * fabs(1.0);
                    } else {
                        double nrm = sqrt(dux * dux + duy * duy);
                        m11 = mur * area * fabs(vn[iv][f][0] * (-dux / nrm) + vn[iv][f][1] * (-duy / nrm));
                    }
                    double mu1 = m11 * uv + 0.0 * vv, mu2 = 0.0 * uv + m11 * vv;
                    double sn1 = -pv * area * vn[iv][f][0] + 0.0 * area * vn[iv][f][1];
                    double sn2 = 0.0 * area * vn[iv][f][0] + -pv * area * vn[iv][f][1];
                    ORC_PHASE(ORC_PH_NODE);
                    mn[0 + 4 * p] += m11;
                    mn[1 + 4 * p] += 0.0;
                    mn[2 + 4 * p] += 0.0;
                    mn[3 + 4 * p] += m11;
                    mu[2 * p] += mu1;
                    mu[2 * p + 1] += mu2;
                    sn[2 * p] += sn1;
                    sn[2 * p + 1] += sn2;
                    ORC_PHASE(ORC_PH_CORNER);
                    int o = f + 2 * (iv + 4 * isg);
                    mcl[o + 32 * ie] = m11;
                    mul[2 * o + 64 * ie] = mu1;
                    mul[2 * o + 1 + 64 * ie] = mu2;
                    snl[2 * o + 64 * ie] = sn1;
                    snl[2 * o + 1 + 64 * ie] = sn2;
                }
            }
        }
    }
    ORC_PHASE(ORC_PH_NODE);
    int *indnd = (int *)calloc((size_t)np, sizeof(int));
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 25) indnd[b[0]] = indnd[b[1]] = indnd[b[4]] = 1;
    }
    for (int p = 0; p < np; ++p) {
        if (indnd[p]) continue;
        double a11 = mn[4 * p], a21 = mn[4 * p + 1], a12 = mn[4 * p + 2], a22 = mn[4 * p + 3];
        double det = a11 * a22 - a21 * a12;
        double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        double r1 = mu[2 * p] - sn[2 * p], r2 = mu[2 * p + 1] - sn[2 * p + 1];
        X(c->ustar, 0, p) = i11 * r1 + i12 * r2;
        X(c->ustar, 1, p) = i21 * r1 + i22 * r2;
    }
    /* wall projection on all three face nodes (Getrhslaggauss.f90:4515-4545) */
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        int nodes[3] = {b[0], b[1], b[4]};
        for (int k = 0; k < 3; ++k) {
            if (b[2] == 22) {
                if (b[3] == 221) X(c->ustar, 1, nodes[k]) = 0.0;
                else if (b[3] == 222) X(c->ustar, 0, nodes[k]) = 0.0;
            } else if (b[2] == 25) {
                X(c->ustar, 1, nodes[k]) = 0.0;
                X(c->ustar, 0, nodes[k]) = 0.0;
            }
        }
    }
    ORC_PHASE(ORC_PH_CELLRHS);
    for (int it = 0; it < nt; ++it) {
        const int *ip = TIP(c, it);
        for (int isg = 0; isg < 4; ++isg)
            for (int iv = 0; iv < 3; ++iv)
                for (int f = 0; f < 2; ++f) {
                    const int o = f + 2 * (iv + 3 * isg);
                    const int p = ip[IPTSG[isg][iv]];
                    const double m11 = mclt[o + 24 * it];
                    const double mu1 = mult[2 * o + 48 * it], mu2 = mult[2 * o + 1 + 48 * it];
                    const double sn1 = snlt[2 * o + 48 * it], sn2 = snlt[2 * o + 1 + 48 * it];
                    const double us1 = X(c->ustar, 0, p), us2 = X(c->ustar, 1, p);
                    fstart[2 * o + 48 * it] = sn1 + m11 * us1 + 0.0 * us2 - mu1;
                    fstart[2 * o + 1 + 48 * it] = sn2 + 0.0 * us1 + m11 * us2 - mu2;
                }
    }
    free(mclt); free(mult); free(snlt);
    for (int ie = c->ntria; ie < nc; ++ie) {
        const int *ip = QIP(c, ie);
        for (int isg = 0; isg < 4; ++isg)
            for (int iv = 0; iv < 4; ++iv)
                for (int f = 0; f < 2; ++f) {
                    int o = f + 2 * (iv + 4 * isg);
                    int p = ip[IPQSG[isg][iv]];
                    double m11 = mcl[o + 32 * ie];
                    double mu1 = mul[2 * o + 64 * ie], mu2 = mul[2 * o + 1 + 64 * ie];
                    double sn1 = snl[2 * o + 64 * ie], sn2 = snl[2 * o + 1 + 64 * ie];
                    double us1 = X(c->ustar, 0, p), us2 = X(c->ustar, 1, p);
                    fstar[2 * o + 64 * ie] = sn1 + m11 * us1 + 0.0 * us2 - mu1;
                    fstar[2 * o + 1 + 64 * ie] = sn2 + 0.0 * us1 + m11 * us2 - mu2;
                }
    }
    free(indnd);
    free(mn); free(mu); free(sn); free(mcl); free(mul); free(snl);
}

/* rhsifacedg_lagsubgq2 (Getrhslaggauss.f90:8501-8745): overwrites rhsel */
static void iface_q2(orc_case *c, const double *gq, const double *fstar)
{
    double wm[4][4][2], wv[4][4][2];
    wfg_face(wm, wv);
    const int nd = c->ndegr;
    for (int ie = c->ntria; ie < c->ncell; ++ie) {
        const int *ip = QIP(c, ie);
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double bq[9][6];
        for (int iv = 0; iv < 9; ++iv) basis_p2(XV9[iv], YV9[iv], rc, sc, &GE(c, 18, ie), bq[iv]);
        double ul[6] = {0}, p1[6] = {0}, p2[6] = {0}, el[6] = {0};
        const double *g = gq + 48 * ie;
        for (int isg = 0; isg < 4; ++isg) {
            double vn[4][2][3];
            for (int iv = 0; iv < 4; ++iv)
                for (int f = 0; f < 2; ++f) {
                    int t = FNQSG[isg][2 * iv + f];
                    int sl = abs(t) - 1;
                    double sg = t > 0 ? 1.0 : -1.0;
                    vn[iv][f][0] = sg * g[3 * sl];
                    vn[iv][f][1] = sg * g[3 * sl + 1];
                    vn[iv][f][2] = g[3 * sl + 2] * wv[isg][iv][f];
                }
            for (int iv = 0; iv < 4; ++iv) {
                int ln = IPQSG[isg][iv];
                int p = ip[ln];
                double us1 = X(c->ustar, 0, p), us2 = X(c->ustar, 1, p);
                int o0 = 0 + 2 * (iv + 4 * isg), o1 = 1 + 2 * (iv + 4 * isg);
                const double *f0 = fstar + 2 * o0 + 64 * ie, *f1 = fstar + 2 * o1 + 64 * ie;
                double w1 = wm[isg][iv][0], w2 = wm[isg][iv][1];
                for (int id = 0; id < nd; ++id) {
                    double b = bq[ln][id];
                    ul[id] = ul[id] + us1 * vn[iv][0][0] * vn[iv][0][2] * b + us2 * vn[iv][0][1] * vn[iv][0][2] * b +
                             us1 * vn[iv][1][0] * vn[iv][1][2] * b + us2 * vn[iv][1][1] * vn[iv][1][2] * b;
                    p1[id] = p1[id] + f0[0] * b * w1 + f1[0] * b * w2;
                    p2[id] = p2[id] + f0[1] * b * w1 + f1[1] * b * w2;
                    el[id] = el[id] + us1 * f0[0] * b * w1 + us2 * f0[1] * b * w1 + us1 * f1[0] * b * w2 +
                             us2 * f1[1] * b * w2;
                }
            }
        }
        for (int id = 0; id < nd; ++id) {
            U(c, c->rhsel, id, 0, ie) = ul[id];
            U(c, c->rhsel, id, 1, ie) = p1[id];
            U(c, c->rhsel, id, 2, ie) = p2[id];
            U(c, c->rhsel, id, 3, ie) = el[id];
        }
    }
}

/* rhsdomndg_lagmc_curvquad (Getrhslag.f90:4716-5013): 5x5 Gauss-Lobatto */
static void domn_q2(orc_case *c)
{
    const double pl[5] = {-1.0, -0.6546536707079771437983, 0.0, 0.6546536707079771437983, 1.0};
    const double wl[5] = {0.1, 0.544444444444444444444, 0.7111111111111111111111, 0.544444444444444444444, 0.1};
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    const int nd = c->ndegr;
    for (int ie = c->ntria; ie < c->ncell; ++ie) {
        double xp[2][9];
        q2_nodes(c, c->cooro, ie, xp);
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) {
                double r = pl[i], s = pl[j], wi = wl[i] * wl[j];
                double J[4];
                q2_jac(xp, r, s, J, NULL, NULL);
                double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
                double dbdr[6] = {0.0, 1.0 / 1.0, 0.0, (r - rc) / 1.0, 0.0, (s - sc) / 1.0 / 1.0};
                double dbds[6] = {0.0, 0.0, 1.0 / 1.0, 0.0, (s - sc) / 1.0, (r - rc) / 1.0 / 1.0};
                double g1d[6], g2d[6];
                for (int id = 0; id < 6; ++id) {
                    g1d[id] = dyds * dbdr[id] + (-dydr) * dbds[id];
                    g2d[id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
                }
                double b[6];
                basis_p2(r, s, rc, sc, &GE(c, 18, ie), b);
                double un[4] = {0, 0, 0, 0};
                for (int id = 0; id < nd; ++id)
                    for (int q = 0; q < 4; ++q) un[q] = un[q] + U(c, c->unold, id, q, ie) * b[id];
                double rhoad = 1.0 / un[0];
                double u = un[1], v = un[2];
                double pres = dmax(eps, g1 * rhoad * (un[3] - 0.5 * (u * u + v * v)));
                for (int id = 0; id < nd; ++id) {
                    double f0 = g1d[id] * u + g2d[id] * v;
                    double f1 = g1d[id] * (-pres);
                    double f2 = g2d[id] * (-pres);
                    double f3 = (g1d[id] * u + g2d[id] * v) * (-pres);
                    U(c, c->rhsel, id, 0, ie) = U(c, c->rhsel, id, 0, ie) - f0 * wi;
                    U(c, c->rhsel, id, 1, ie) = U(c, c->rhsel, id, 1, ie) - f1 * wi;
                    U(c, c->rhsel, id, 2, ie) = U(c, c->rhsel, id, 2, ie) - f2 * wi;
                    U(c, c->rhsel, id, 3, ie) = U(c, c->rhsel, id, 3, ie) - f3 * wi;
                }
            }
    }
}

/* ------------------------------------------------------------------------ */
/* nfint = 5: the sub-cell scheme with vertex and non-vertex Gauss points on
 * the cell faces (getrhsdg_lag_mcnew branch Getrhslagmc.f90:7712-7829).
 *   getriem_quadsubg_gauss    Getrhslaggauss.f90:16657-17157
 *   getndvelo_lagsubg_gaussh  Getrhslaggauss.f90:17161-17546
 *   getvelo_quadsubg_gs3      Getrhslaggauss.f90:19312-19990
 *   getvar_face_general       Getrhslaggauss.f90:18184-18368
 *   getvectf                  Getrhslaggauss.f90:16547-16590
 *   rhsifacedg_lagsubgq_gs    Getrhslaggauss.f90:17550-17939
 * The face geometry routines of this branch (getfnds_lagsubgc,
 * getfnds_lag_simpsonh3) compute the quad slots with the same expressions as
 * the nfint=4 ones (fnds_q2).  Persson's flag is cleared before
 * barthlimit_lag_curvquadcb_HO runs (:7763-7766), so that limiter skips
 * every cell (Limiterlag.f90:10276) and aflim / afvec are reset to the
 * identity (:7782-7785): the whole limiting block is omitted. */

/* Lobatto point of the face rule, +-sqrt(3/7) as the reference writes it */
#define G5A 0.6546536707079771437983
/* cell-local face Gauss points 12..19 (posiq 13..20, :19443-19455): face lf
 * holds points 12 + 2 lf + k, k = 0 at r = -G5A along corner lf -> lf+1 */
static const double PQG5[8][2] = {{-G5A, -1.0}, {G5A, -1.0}, {1.0, -G5A}, {1.0, G5A},
                                  {G5A, 1.0},  {-G5A, 1.0}, {-1.0, G5A}, {-1.0, -G5A}};
/* sub-cell of each face Gauss point (sglgq, :18217-18220) */
static const int SGG5[8] = {0, 1, 1, 2, 2, 3, 3, 0};

/* corner weights of the nfint=5 Riemann solver (wfgsq, :16777-16796) */
static const double WR5[4][4][2] = {{{1, 1}, {.5, 1}, {0, 0}, {1, .5}},
                                    {{1, .5}, {1, 1}, {.5, 1}, {0, 0}},
                                    {{0, 0}, {1, .5}, {1, 1}, {.5, 1}},
                                    {{.5, 1}, {0, 0}, {1, .5}, {1, 1}}};

/* face-integral weights of rhsifacedg_lagsubgq_gs (:17651-17700) from the
 * 5-point Lobatto weights w1, w3: momentum/energy wm (w3/1.0) and volume wv
 * (w3/2.0), both halved */
static void wfg_face5(double wm[4][4][2], double wv[4][4][2])
{
    const double w1 = 0.1, w3 = 0.7111111111111111111111;
    /* 0: w1, 1: w3, 2: w3 / k (k = 1 momentum, 2 volume), 3: zero */
    static const int t[4][4][2] = {{{0, 0}, {2, 1}, {3, 3}, {1, 2}},
                                   {{1, 2}, {0, 0}, {2, 1}, {3, 3}},
                                   {{3, 3}, {1, 2}, {0, 0}, {2, 1}},
                                   {{2, 1}, {3, 3}, {1, 2}, {0, 0}}};
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int f = 0; f < 2; ++f) {
                const int k = t[a][b][f];
                const double m = k == 0 ? w1 : k == 1 ? w3 : k == 2 ? w3 / 1.0 : 0.0;
                const double v = k == 0 ? w1 : k == 1 ? w3 : k == 2 ? w3 / 2.0 : 0.0;
                wm[a][b][f] = 0.5 * m;
                wv[a][b][f] = 0.5 * v;
            }
}

/* getvectf: unit normal (dy/dr, -dx/dr) and 2 |dx/dr| of the quadratic face
 * (corner a, corner b, mid-node) at the face Lobatto point k */
static void face_normal5(const double fx[3], const double fy[3], int k, double n[3])
{
    const double r = k == 0 ? -G5A : G5A;
    const double d[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
    double dxdr = 0.0, dydr = 0.0;
    for (int i = 0; i < 3; ++i) {
        dxdr = dxdr + d[i] * fx[i];
        dydr = dydr + d[i] * fy[i];
    }
    const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
    const double anx = dydr, any = -dxdr;
    n[0] = anx / sqrt(anx * anx + any * any);
    n[1] = any / sqrt(anx * anx + any * any);
    n[2] = djaco * 2.0;
}

/* getvar_face_general: the cell's state at its face Gauss point g (0..7):
 * v[0..3] the DG values, v[4] pressure, v[5] rho c of the cell average,
 * v[6] c of the cell average */
static void face_state5(const orc_case *c, int ie, int g, const double *drho, double v[7])
{
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    const double rhoct = 1.0 / U(c, c->unold, 0, 0, ie);
    const double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie), ectr = U(c, c->unold, 0, 3, ie);
    const double eintc = ectr - 0.5 * (uctr * uctr + vctr * vctr);
    const double pctr = dmax(eps, g1 * rhoct * eintc);
    const double sdctr = sqrt(c->gamlg * pctr / rhoct);
    double b[6];
    basis_p2(PQG5[g][0], PQG5[g][1], GE(c, 0, ie), GE(c, 1, ie), &GE(c, 18, ie), b);
    for (int q = 0; q < 4; ++q) v[q] = 0.0;
    for (int id = 0; id < c->ndegr; ++id)
        for (int q = 0; q < 4; ++q) v[q] = v[q] + U(c, c->unold, id, q, ie) * b[id];
    const double rhovt = 1.0 / v[0] + c->cdrho * drho[4 * ie + SGG5[g]];
    const double eintv = v[3] - 0.5 * (v[1] * v[1] + v[2] * v[2]);
    v[4] = dmax(eps, g1 * rhovt * eintv);
    v[5] = rhoct * sdctr;
    v[6] = sdctr;
}

/* one side of a face Gauss point's Riemann problem (getvelo_quadsubg_gs3
 * :19668-19701 with getriecoef_matrixnew, Limiterlag.f90:1037-1076):
 * out[0] = M (isotropic), out[1..2] = M u, out[3..4] = A sigma n */
static void face_riem5(const orc_case *c, const double v[7], const double uf[2], const double n[3], double out[5])
{
    const double eps11 = (double)1.e-11f;
    double a1 = uf[0] - v[1], a2 = uf[1] - v[2];
    const double acnx = a1, acny = a2;
    if (sqrt(acnx * acnx + acny * acny) < eps11) {
        a1 = eps11;
        a2 = eps11;
    } else {
        a1 = a1 / sqrt(acnx * acnx + acny * acny);
        a2 = a2 / sqrt(acnx * acnx + acny * acny);
    }
    const double a3 = sqrt(acnx * acnx + acny * acny) / v[6];
    const double delu = c->cimpd * fabs(a1 * n[0] + a2 * n[1]);
    const double mur = v[5] + v[5] / v[6] * c->slpdu * delu;
    const double area = n[2];
    double m;
    if (a3 < 1.0e12) m = mur * area * fabs(1.0);
    else m = mur * area * fabs(n[0] * a1 + n[1] * a2);
    const double pv = v[4];
    out[0] = m;
    out[1] = m * v[1] + 0.0 * v[2];
    out[2] = 0.0 * v[1] + m * v[2];
    out[3] = -pv * area * n[0] + 0.0 * area * n[1];
    out[4] = 0.0 * area * n[0] + -pv * area * n[1];
}

/* boundary condition of each cell face (0 interior; 1 v = 0, 2 u = 0,
 * 3 u = v = 0) and whether the boundary face's nodes run against the cell's
 * counter-clockwise order (its normal then points inwards) */
static void face_bc5(const orc_case *c, int *bc, int *rev)
{
    const int nc = c->ncell;
    int *mid = (int *)malloc(sizeof(int) * c->npoin);
    for (int p = 0; p < c->npoin; ++p) mid[p] = -1;
    for (int ie = 0; ie < nc; ++ie)
        for (int lf = 0; lf < 4; ++lf) {
            mid[QIP(c, ie)[4 + lf]] = 4 * ie + lf;
            bc[4 * ie + lf] = 0;
            rev[4 * ie + lf] = 0;
        }
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        const int k = mid[b[4]];
        if (k < 0) continue;
        int code = 0;
        if (b[2] == 22) code = b[3] == 221 ? 1 : b[3] == 222 ? 2 : 0;
        else if (b[2] == 25 || b[2] == 27) code = 3;
        bc[k] = code;
        rev[k] = b[0] != QIP(c, k / 4)[k % 4];
    }
    free(mid);
}

/* nodal solve (getndvelo_lagsubg_gaussh): corners as getriem_quadsubg_gauss
 * computes them, the 2x2 node solve and walls, then the face Gauss points.
 * Outputs: fstar (vertex corner forces, as ndvelo_q2), and per cell and
 * local Gauss point fg[(j + 6*g) + 48*ie]: j = 0,1 the force, 2,3 the
 * correction force, 4,5 the interpolated nodal velocity ufgpq */
static void ndvelo_q2g(orc_case *c, const double *gq, double *fstar, double *fg)
{
    const int np = c->npoin, nc = c->ncell, nd = c->ndegr;
    const double g1 = c->gamlg - 1.0, eps = 1.0e-6;
    double *mn = (double *)calloc((size_t)4 * np, sizeof(double));
    double *mu = (double *)calloc((size_t)2 * np, sizeof(double));
    double *sn = (double *)calloc((size_t)2 * np, sizeof(double));
    double *mcl = (double *)malloc(sizeof(double) * 32 * nc);
    double *mul = (double *)malloc(sizeof(double) * 64 * nc);
    double *snl = (double *)malloc(sizeof(double) * 64 * nc);
    double *drho = (double *)malloc(sizeof(double) * 4 * nc);
    for (int ie = 0; ie < nc; ++ie) {
        const int *ip = QIP(c, ie);
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double un[9][4];
        for (int iv = 0; iv < 9; ++iv) {
            double b[6];
            basis_p2(XV9[iv], YV9[iv], rc, sc, &GE(c, 18, ie), b);
            for (int q = 0; q < 4; ++q) un[iv][q] = 0.0;
            for (int id = 0; id < nd; ++id)
                for (int q = 0; q < 4; ++q) un[iv][q] = un[iv][q] + U(c, c->unold, id, q, ie) * b[id];
        }
        double rhomc = U(c, c->unold, 0, 0, ie);
        double uctr = U(c, c->unold, 0, 1, ie), vctr = U(c, c->unold, 0, 2, ie), ectr = U(c, c->unold, 0, 3, ie);
        double rhoct = 1.0 / rhomc;
        double eintc = ectr - 0.5 * (uctr * uctr + vctr * vctr);
        /* GetEOS (Getsolid.f90:319-397, nmatel = 1): no clamp on c^2 */
        double pctr = dmax(eps, g1 * rhoct * eintc);
        double sdctr = sqrt(c->gamlg * pctr / rhoct);
        double rsub[4], uq[4];
        subcell_density(c, ie, rsub);
        subcell_tau(c, ie, uq);
        const double *g = gq + 48 * ie;
        for (int isg = 0; isg < 4; ++isg) {
            drho[4 * ie + isg] = rsub[isg] - 1.0 / uq[isg];
            double vn[4][2][3];
            for (int iv = 0; iv < 4; ++iv)
                for (int f = 0; f < 2; ++f) {
                    int t = FNQSG[isg][2 * iv + f];
                    int sl = abs(t) - 1;
                    double sg = t > 0 ? 1.0 : -1.0;
                    vn[iv][f][0] = sg * g[3 * sl];
                    vn[iv][f][1] = sg * g[3 * sl + 1];
                    vn[iv][f][2] = g[3 * sl + 2] * WR5[isg][iv][f];
                }
            for (int iv = 0; iv < 4; ++iv) {
                int ln = IPQSG[isg][iv];
                int p = ip[ln];
                double rhovt = 1.0 / un[ln][0];
                double rhovsg = rhovt + c->cdrho * (rsub[isg] - 1.0 / uq[isg]);
                double uv = un[ln][1], vv = un[ln][2], ev = un[ln][3];
                double eintv = ev - 0.5 * (uv * uv + vv * vv);
                double pv = dmax(eps, g1 * rhovsg * eintv);
                double dux = X(c->ustar, 0, p) - uv, duy = X(c->ustar, 1, p) - vv;
                double a3 = sqrt(dux * dux + duy * duy) / sdctr;
                for (int f = 0; f < 2; ++f) {
                    double deltu = c->cimpd * fabs(dux * vn[iv][f][0] + duy * vn[iv][f][1]);
                    double mur = rhoct * sdctr + rhoct * c->slpdu * deltu;
                    double area = vn[iv][f][2];
                    double m11;
                    if (a3 < 1.e12) {
                        m11 = mur * area * fabs(1.0);
                    } else {
                        double nrm = sqrt(dux * dux + duy * duy);
                        m11 = mur * area * fabs(vn[iv][f][0] * (dux / nrm) + vn[iv][f][1] * (duy / nrm));
                    }
                    double mu1 = m11 * uv + 0.0 * vv, mu2 = 0.0 * uv + m11 * vv;
                    double sn1 = -pv * area * vn[iv][f][0] + 0.0 * area * vn[iv][f][1];
                    double sn2 = 0.0 * area * vn[iv][f][0] + -pv * area * vn[iv][f][1];
                    mn[0 + 4 * p] += m11;
                    mn[1 + 4 * p] += 0.0;
                    mn[2 + 4 * p] += 0.0;
                    mn[3 + 4 * p] += m11;
                    mu[2 * p] += mu1;
                    mu[2 * p + 1] += mu2;
                    sn[2 * p] += sn1;
                    sn[2 * p + 1] += sn2;
                    int o = f + 2 * (iv + 4 * isg);
                    mcl[o + 32 * ie] = m11;
                    mul[2 * o + 64 * ie] = mu1;
                    mul[2 * o + 1 + 64 * ie] = mu2;
                    snl[2 * o + 64 * ie] = sn1;
                    snl[2 * o + 1 + 64 * ie] = sn2;
                }
            }
        }
    }
    int *indnd = (int *)calloc((size_t)np, sizeof(int));
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        if (b[2] == 25) indnd[b[0]] = indnd[b[1]] = indnd[b[4]] = 1;
    }
    for (int p = 0; p < np; ++p) {
        if (indnd[p]) continue;
        double a11 = mn[4 * p], a21 = mn[4 * p + 1], a12 = mn[4 * p + 2], a22 = mn[4 * p + 3];
        double det = a11 * a22 - a21 * a12;
        double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        double r1 = mu[2 * p] - sn[2 * p], r2 = mu[2 * p + 1] - sn[2 * p + 1];
        X(c->ustar, 0, p) = i11 * r1 + i12 * r2;
        X(c->ustar, 1, p) = i21 * r1 + i22 * r2;
    }
    for (int f = 0; f < c->nbfac; ++f) {
        const int *b = c->bface + 5 * f;
        int nodes[3] = {b[0], b[1], b[4]};
        for (int k = 0; k < 3; ++k) {
            if (b[2] == 22) {
                if (b[3] == 221) X(c->ustar, 1, nodes[k]) = 0.0;
                else if (b[3] == 222) X(c->ustar, 0, nodes[k]) = 0.0;
            } else if (b[2] == 25) {
                X(c->ustar, 1, nodes[k]) = 0.0;
                X(c->ustar, 0, nodes[k]) = 0.0;
            }
        }
    }
    /* ufgpq: u* interpolated along each face at its two Lobatto points
     * (:17416-17437) */
    for (int ie = 0; ie < nc; ++ie) {
        const int *ip = QIP(c, ie);
        for (int lf = 0; lf < 4; ++lf) {
            const int pa = ip[lf], pb = ip[(lf + 1) % 4], pm = ip[4 + lf];
            for (int k = 0; k < 2; ++k) {
                const double r = k == 0 ? -G5A : G5A;
                const double s1 = -0.5 * (1.0 - r) * r, s2 = 0.5 * (1.0 + r) * r, s3 = (1.0 + r) * (1.0 - r);
                double *o = fg + 6 * (2 * lf + k) + 48 * ie;
                for (int d = 0; d < 2; ++d)
                    o[4 + d] = s1 * X(c->ustar, d, pa) + s2 * X(c->ustar, d, pb) + s3 * X(c->ustar, d, pm);
            }
        }
    }
    /* the face Gauss points' Riemann problems, one face at a time
     * (getvelo_quadsubg_gs3): the cell on each side, its state at the
     * point, the 2x2 solve and the velocity correction dufgq; then the
     * forces of each side (:17470-17527) */
    int *bc = (int *)malloc(sizeof(int) * 4 * nc), *rev = (int *)malloc(sizeof(int) * 4 * nc);
    face_bc5(c, bc, rev);
    int *nbr = (int *)malloc(sizeof(int) * 4 * nc);
    {
        int *mid = (int *)malloc(sizeof(int) * np);
        for (int p = 0; p < np; ++p) mid[p] = -1;
        for (int ie = 0; ie < nc; ++ie)
            for (int lf = 0; lf < 4; ++lf) {
                const int m = QIP(c, ie)[4 + lf];
                nbr[4 * ie + lf] = -1;
                if (mid[m] >= 0) {
                    nbr[4 * ie + lf] = mid[m];
                    nbr[mid[m]] = 4 * ie + lf;
                } else {
                    mid[m] = 4 * ie + lf;
                }
            }
        free(mid);
    }
    for (int ie = 0; ie < nc; ++ie) {
        const int *ip = QIP(c, ie);
        for (int lf = 0; lf < 4; ++lf) {
            const int other = nbr[4 * ie + lf];
            if (other >= 0 && other < 4 * ie + lf) continue; /* the face was done from the other side */
            double fx[3], fy[3];
            const int fn[3] = {ip[lf], ip[(lf + 1) % 4], ip[4 + lf]};
            for (int i = 0; i < 3; ++i) {
                fx[i] = X(c->cooro, 0, fn[i]);
                fy[i] = X(c->cooro, 1, fn[i]);
            }
            for (int k = 0; k < 2; ++k) {
                double n[3];
                face_normal5(fx, fy, k, n);
                const int gl = 2 * lf + k;
                const double *uf = fg + 6 * gl + 48 * ie + 4;
                double vl[7], rl[5];
                face_state5(c, ie, gl, drho, vl);
                if (other < 0) {
                    /* boundary face: this side only, its normal as the
                     * boundary face's nodes orient it */
                    double nb[3] = {rev[4 * ie + lf] ? -n[0] : n[0], rev[4 * ie + lf] ? -n[1] : n[1], n[2]};
                    face_riem5(c, vl, uf, nb, rl);
                    double a11 = 0.0 + rl[0], a22 = 0.0 + rl[0], a21 = 0.0 + 0.0, a12 = 0.0 + 0.0;
                    double m1 = 0.0 + rl[1], m2 = 0.0 + rl[2], s1 = 0.0 + rl[3], s2 = 0.0 + rl[4];
                    double det = a11 * a22 - a21 * a12;
                    double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
                    double r1 = m1 - s1 - 0.0, r2 = m2 - s2 - 0.0;
                    double du1 = -uf[0] + (i11 * r1 + i12 * r2);
                    double du2 = -uf[1] + (i21 * r1 + i22 * r2);
                    const int b = bc[4 * ie + lf];
                    if (b == 1) du2 = 0.0;
                    else if (b == 2) du1 = 0.0;
                    else if (b == 3) du1 = du2 = 0.0;
                    double *o = fg + 6 * gl + 48 * ie;
                    o[0] = rl[3] + rl[0] * uf[0] + 0.0 * uf[1] - rl[1];
                    o[1] = rl[4] + rl[0] * uf[1] + 0.0 * uf[0] - rl[2];
                    o[2] = rl[0] * du1 + 0.0 * du2;
                    o[3] = rl[0] * du2 + 0.0 * du1;
                } else {
                    const int je = other / 4, gr = 2 * (other % 4) + (1 - k);
                    const double *ufr = fg + 6 * gr + 48 * je + 4;
                    double vr[7], rr[5];
                    face_state5(c, je, gr, drho, vr);
                    const double nr[3] = {-n[0], -n[1], n[2]};
                    face_riem5(c, vl, ufr, n, rl);
                    face_riem5(c, vr, ufr, nr, rr);
                    double a11 = 0.0 + rl[0] + rr[0], a22 = 0.0 + rl[0] + rr[0];
                    double a21 = 0.0 + 0.0 + 0.0, a12 = 0.0 + 0.0 + 0.0;
                    double m1 = 0.0 + rl[1] + rr[1], m2 = 0.0 + rl[2] + rr[2];
                    double s1 = 0.0 + rl[3] + rr[3], s2 = 0.0 + rl[4] + rr[4];
                    double det = a11 * a22 - a21 * a12;
                    double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
                    double r1 = m1 - s1, r2 = m2 - s2;
                    const double *side[2] = {rl, rr};
                    const int cell[2] = {ie, je}, gg[2] = {gl, gr};
                    for (int sd = 0; sd < 2; ++sd) {
                        const double *ufs = fg + 6 * gg[sd] + 48 * cell[sd] + 4;
                        double du1 = -ufs[0] + (i11 * r1 + i12 * r2);
                        double du2 = -ufs[1] + (i21 * r1 + i22 * r2);
                        const double *rs = side[sd];
                        double *o = fg + 6 * gg[sd] + 48 * cell[sd];
                        o[0] = rs[3] + rs[0] * ufs[0] + 0.0 * ufs[1] - rs[1];
                        o[1] = rs[4] + rs[0] * ufs[1] + 0.0 * ufs[0] - rs[2];
                        o[2] = rs[0] * du1 + 0.0 * du2;
                        o[3] = rs[0] * du2 + 0.0 * du1;
                    }
                }
            }
        }
    }
    /* the centre nodes' u* is zeroed before the vertex forces (:17475) */
    for (int ie = 0; ie < nc; ++ie) {
        const int p = QIP(c, ie)[8];
        X(c->ustar, 0, p) = 0.0;
        X(c->ustar, 1, p) = 0.0;
    }
    for (int ie = 0; ie < nc; ++ie) {
        const int *ip = QIP(c, ie);
        for (int isg = 0; isg < 4; ++isg)
            for (int iv = 0; iv < 4; ++iv)
                for (int f = 0; f < 2; ++f) {
                    int o = f + 2 * (iv + 4 * isg);
                    int p = ip[IPQSG[isg][iv]];
                    double m11 = mcl[o + 32 * ie];
                    double mu1 = mul[2 * o + 64 * ie], mu2 = mul[2 * o + 1 + 64 * ie];
                    double sn1 = snl[2 * o + 64 * ie], sn2 = snl[2 * o + 1 + 64 * ie];
                    double us1 = X(c->ustar, 0, p), us2 = X(c->ustar, 1, p);
                    fstar[2 * o + 64 * ie] = sn1 + m11 * us1 + 0.0 * us2 - mu1;
                    fstar[2 * o + 1 + 64 * ie] = sn2 + 0.0 * us1 + m11 * us2 - mu2;
                }
    }
    free(bc); free(rev); free(nbr); free(indnd); free(drho);
    free(mn); free(mu); free(sn); free(mcl); free(mul); free(snl);
}

/* rhsifacedg_lagsubgq_gs (Getrhslaggauss.f90:17550-17939): vertex terms
 * with the Lobatto-weighted corner faces, then per face the two Gauss
 * points' forces and, in a second pass, their correction forces;
 * overwrites rhsel */
static void iface_q2g(orc_case *c, const double *gq, const double *fstar, const double *fg)
{
    double wm[4][4][2], wv[4][4][2];
    wfg_face5(wm, wv);
    const double wi = 0.544444444444444444444 / 2.0;
    const int nd = c->ndegr;
    for (int ie = 0; ie < c->ncell; ++ie) {
        const int *ip = QIP(c, ie);
        double rc = GE(c, 0, ie), sc = GE(c, 1, ie);
        double bq[9][6];
        for (int iv = 0; iv < 9; ++iv) basis_p2(XV9[iv], YV9[iv], rc, sc, &GE(c, 18, ie), bq[iv]);
        double ul[6] = {0}, p1[6] = {0}, p2[6] = {0}, el[6] = {0};
        const double *g = gq + 48 * ie;
        for (int isg = 0; isg < 4; ++isg) {
            double vn[4][2][3];
            for (int iv = 0; iv < 4; ++iv)
                for (int f = 0; f < 2; ++f) {
                    int t = FNQSG[isg][2 * iv + f];
                    int sl = abs(t) - 1;
                    double sg = t > 0 ? 1.0 : -1.0;
                    vn[iv][f][0] = sg * g[3 * sl];
                    vn[iv][f][1] = sg * g[3 * sl + 1];
                    vn[iv][f][2] = g[3 * sl + 2] * wv[isg][iv][f];
                }
            for (int iv = 0; iv < 4; ++iv) {
                int ln = IPQSG[isg][iv];
                int p = ip[ln];
                double us1 = X(c->ustar, 0, p), us2 = X(c->ustar, 1, p);
                int o0 = 0 + 2 * (iv + 4 * isg), o1 = 1 + 2 * (iv + 4 * isg);
                const double *f0 = fstar + 2 * o0 + 64 * ie, *f1 = fstar + 2 * o1 + 64 * ie;
                double w1 = wm[isg][iv][0], w2 = wm[isg][iv][1];
                for (int id = 0; id < nd; ++id) {
                    double b = bq[ln][id];
                    ul[id] = ul[id] + us1 * vn[iv][0][0] * vn[iv][0][2] * b + us2 * vn[iv][0][1] * vn[iv][0][2] * b +
                             us1 * vn[iv][1][0] * vn[iv][1][2] * b + us2 * vn[iv][1][1] * vn[iv][1][2] * b;
                    p1[id] = p1[id] + f0[0] * b * w1 + f1[0] * b * w2;
                    p2[id] = p2[id] + f0[1] * b * w1 + f1[1] * b * w2;
                    el[id] = el[id] + us1 * f0[0] * b * w1 + us2 * f0[1] * b * w1 + us1 * f1[0] * b * w2 +
                             us2 * f1[1] * b * w2;
                }
            }
        }
        for (int lf = 0; lf < 4; ++lf) {
            double fx[3], fy[3];
            const int fn[3] = {ip[lf], ip[(lf + 1) % 4], ip[4 + lf]};
            for (int i = 0; i < 3; ++i) {
                fx[i] = X(c->cooro, 0, fn[i]);
                fy[i] = X(c->cooro, 1, fn[i]);
            }
            double bg[2][6], n[2][3];
            for (int k = 0; k < 2; ++k) {
                const int gl = 2 * lf + k;
                basis_p2(PQG5[gl][0], PQG5[gl][1], rc, sc, &GE(c, 18, ie), bg[k]);
                face_normal5(fx, fy, k, n[k]);
            }
            for (int k = 0; k < 2; ++k) {
                const double *o = fg + 6 * (2 * lf + k) + 48 * ie;
                const double uf1 = o[4], uf2 = o[5];
                for (int id = 0; id < nd; ++id) {
                    const double b = bg[k][id];
                    ul[id] = ul[id] + uf1 * n[k][0] * n[k][2] * b * wi + uf2 * n[k][1] * n[k][2] * b * wi;
                    p1[id] = p1[id] + o[0] * b * wi;
                    p2[id] = p2[id] + o[1] * b * wi;
                    el[id] = el[id] + uf1 * o[0] * b * wi + uf2 * o[1] * b * wi;
                }
            }
            for (int k = 0; k < 2; ++k) {
                const double *o = fg + 6 * (2 * lf + k) + 48 * ie;
                const double uf1 = o[4], uf2 = o[5];
                for (int id = 0; id < nd; ++id) {
                    const double b = bg[k][id];
                    p1[id] = p1[id] + o[2] * b * wi;
                    p2[id] = p2[id] + o[3] * b * wi;
                    el[id] = el[id] + uf1 * o[2] * b * wi + uf2 * o[3] * b * wi;
                }
            }
        }
        for (int id = 0; id < nd; ++id) {
            U(c, c->rhsel, id, 0, ie) = ul[id];
            U(c, c->rhsel, id, 1, ie) = p1[id];
            U(c, c->rhsel, id, 2, ie) = p2[id];
            U(c, c->rhsel, id, 3, ie) = el[id];
        }
    }
}

/* the nfint=5 right-hand side: geometry, the Gauss-point nodal solver, the
 * face integral, the Lobatto domain integral (unlimited: geoel(10) = 0) and
 * the identity derivative update */
static int rhs_q2g(orc_case *c)
{
    const int nc = c->ncell;
    double *gq = (double *)malloc(sizeof(double) * 48 * nc);
    double *fstar = (double *)malloc(sizeof(double) * 64 * nc);
    double *fg = (double *)malloc(sizeof(double) * 48 * nc);
    fnds_q2(c, c->cooro, gq);
    for (int ie = 0; ie < nc; ++ie) {
        GE(c, 9, ie) = 0.0;
        if (c->troubled) c->troubled[ie] = 0.0;
        for (int q = 0; q < 5; ++q) c->aflim[q + 5 * ie] = 1.0;
        c->afvec[4 * ie + 0] = 1.0;
        c->afvec[4 * ie + 1] = 0.0;
        c->afvec[4 * ie + 2] = 0.0;
        c->afvec[4 * ie + 3] = 1.0;
    }
    ndvelo_q2g(c, gq, fstar, fg);
    iface_q2g(c, gq, fstar, fg);
    domn_q2(c);
    /* derivative update with aflim = 1, afvec = I (Getrhslagmc.f90:7809-7823) */
    for (int ie = 0; ie < nc; ++ie) {
        U(c, c->unold, 1, 0, ie) = U(c, c->unold, 1, 0, ie) * 1.0;
        U(c, c->unold, 2, 0, ie) = U(c, c->unold, 2, 0, ie) * 1.0;
        double dudr = 1.0 * U(c, c->unold, 1, 1, ie) + 0.0 * U(c, c->unold, 1, 2, ie);
        double duds = 1.0 * U(c, c->unold, 2, 1, ie) + 0.0 * U(c, c->unold, 2, 2, ie);
        double dvdr = 0.0 * U(c, c->unold, 1, 1, ie) + 1.0 * U(c, c->unold, 1, 2, ie);
        double dvds = 0.0 * U(c, c->unold, 2, 1, ie) + 1.0 * U(c, c->unold, 2, 2, ie);
        U(c, c->unold, 1, 1, ie) = dudr;
        U(c, c->unold, 2, 1, ie) = duds;
        U(c, c->unold, 1, 2, ie) = dvdr;
        U(c, c->unold, 2, 2, ie) = dvds;
        U(c, c->unold, 1, 3, ie) = U(c, c->unold, 1, 3, ie) * 1.0;
        U(c, c->unold, 2, 3, ie) = U(c, c->unold, 2, 3, ie) * 1.0;
    }
    free(gq); free(fstar); free(fg);
    return 0;
}

int orc_rhs_q2(orc_case *c)
{
    if (!((c->npoly == 2 && c->ndegr == 6) || (c->npoly == 1 && c->ndegr == 3)) || c->nrz != 0 ||
        c->ndens != 1 || c->nlimi != 6 || !c->esqua)
        return 2;
    if (c->nfint == 5) return (c->npoly == 2 && c->ntria == 0) ? rhs_q2g(c) : 2;
    if (c->ntria > 0 && (c->npoly != 2 || c->ncase != 7)) return 2;
    const int nc = c->ncell;
    double *gq = (double *)malloc(sizeof(double) * 48 * nc);
    double *flag = (double *)malloc(sizeof(double) * nc);
    double *fstar = (double *)malloc(sizeof(double) * 64 * nc);
    double *gt = (double *)malloc(sizeof(double) * (45 * (size_t)c->ntria + 1));
    double *fstart = (double *)malloc(sizeof(double) * (48 * (size_t)c->ntria + 1));
    ORC_PHASE(ORC_PH_CORNER);
    fnds_q2(c, c->cooro, gq);
    fnds_t2(c, c->cooro, gt);
    ORC_PHASE(ORC_PH_PREP);
    if (c->npoly == 2) {
        persson(c, c->cooro, flag);
    } else {
        /* P1: the Barth limiters run and are then discarded (aflim = 1,
         * afvec = I below, Getrhslagmc.f90:7598-7604, :7650-7656); every cell
         * is flagged troubled (geoel(10) = 100, :7607), so WENO reconstructs
         * every cell whose stencil is complete */
        for (int ie = 0; ie < nc; ++ie) flag[ie] = 100.0;
    }
    int rc = weno(c, flag, c->esqua);
    if (rc) {
        free(gq); free(flag); free(fstar); free(gt); free(fstart);
        return rc;
    }
    ORC_PHASE(ORC_PH_OTHER);
    for (int ie = 0; ie < nc; ++ie) {
        GE(c, 9, ie) = 0.0;
        if (c->troubled) c->troubled[ie] = flag[ie];
        for (int q = 0; q < 5; ++q) c->aflim[q + 5 * ie] = 1.0;
        c->afvec[4 * ie + 0] = 1.0;
        c->afvec[4 * ie + 1] = 0.0;
        c->afvec[4 * ie + 2] = 0.0;
        c->afvec[4 * ie + 3] = 1.0;
    }
    ndvelo_q2(c, gq, fstar, gt, fstart);
    ORC_PHASE(ORC_PH_CELLRHS);
    iface_t2(c, gt, fstart);
    iface_q2(c, gq, fstar);
    domn_t2(c);
    domn_q2(c);
    /* derivative update with aflim = 1, afvec = I (Getrhslagmc.f90:7695-7708) */
    for (int ie = 0; ie < nc; ++ie) {
        U(c, c->unold, 1, 0, ie) = U(c, c->unold, 1, 0, ie) * 1.0;
        U(c, c->unold, 2, 0, ie) = U(c, c->unold, 2, 0, ie) * 1.0;
        double dudr = 1.0 * U(c, c->unold, 1, 1, ie) + 0.0 * U(c, c->unold, 1, 2, ie);
        double duds = 1.0 * U(c, c->unold, 2, 1, ie) + 0.0 * U(c, c->unold, 2, 2, ie);
        double dvdr = 0.0 * U(c, c->unold, 1, 1, ie) + 1.0 * U(c, c->unold, 1, 2, ie);
        double dvds = 0.0 * U(c, c->unold, 2, 1, ie) + 1.0 * U(c, c->unold, 2, 2, ie);
        U(c, c->unold, 1, 1, ie) = dudr;
        U(c, c->unold, 2, 1, ie) = duds;
        U(c, c->unold, 1, 2, ie) = dvdr;
        U(c, c->unold, 2, 2, ie) = dvds;
        U(c, c->unold, 1, 3, ie) = U(c, c->unold, 1, 3, ie) * 1.0;
        U(c, c->unold, 2, 3, ie) = U(c, c->unold, 2, 3, ie) * 1.0;
    }
    free(gq); free(flag); free(fstar); free(gt); free(fstart);
    return 0;
}
