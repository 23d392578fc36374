// Reference-element tables and geometry helpers of the curved (Q2) path,
// shared by its forward kernels (lag2d_q2.hip) and their reverse
// (lag2d_adjq2.hip).
#ifndef LAG2D_Q2GEOM_H
#define LAG2D_Q2GEOM_H
#include "lag2d_dev.h"

namespace lag2d {

// Compact sub-cell corner states in tiles of 64 cells: entry r of corner
// k = 4 isg + iv of cell ie at ((ie / 64) 80 + 5 k + r) 64 + ie % 64, with
// r = 0, 1 the impedance times area m = mu A of the corner's faces f = 0, 1
// and r = 2, 3, 4 the corner's velocity u_c, v_c and pressure p_c.  The
// 64 cells of a wave read or write each entry as 512 contiguous bytes (the
// thread-per-cell kernels), and a node's corners in neighbouring cells
// share the tile's lines (the node gather).  The array holds
// Q2_REC_CELLS(nc) cells, Q2_REC doubles each.
#define RC2(ie, k, r) (((((size_t)(ie)) >> 6) * 80 + (size_t)(k) * 5 + (r)) * 64 + (((size_t)(ie)) & 63))
#define Q2_REC_CELLS(nc) ((((size_t)(nc)) + 63) / 64 * 64)

// sub-cell vertex -> local Q2 node (ipqsg)
static constexpr int IPQSG[4][4] = {{0, 4, 8, 7}, {4, 1, 5, 8}, {8, 5, 2, 6}, {7, 8, 6, 3}};
// signed face-normal slots (fnqsg): sign = orientation, |v|-1 = gesgq slot
static constexpr int FNQSG[4][8] = {{8, 1, 13, -9, -9, 12, 12, 16},
                                             {9, 13, 2, 3, 14, -10, -10, 9},
                                             {-11, 10, 10, 14, 4, 5, 15, -11},
                                             {16, -12, -12, 11, 11, 15, 6, 7}};
static constexpr double XV9[9] = {-1.0, 1.0, 1.0, -1.0, 0.0, 1.0, 0.0, -1.0, 0.0};
static constexpr double YV9[9] = {-1.0, -1.0, 1.0, 1.0, -1.0, 0.0, 1.0, 0.0, 0.0};
// Riemann corner weights times 1/4 [isg][ivsg][f] (Getrhslaggauss.f90:10575-10599)
static constexpr double WFR[4][4][2] = {
    {{0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * 1.0, 0.25 * 1.0},
     {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}},
    {{0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)},
     {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * 1.0, 0.25 * 1.0}},
    {{0.25 * 1.0, 0.25 * 1.0}, {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)},
     {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}},
    {{0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * 1.0, 0.25 * 1.0}, {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)},
     {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}}};
// face-integral weights: momentum/energy {0,1} and volume times 1/4 (:8573-8621)
static constexpr double WFM[4][4][2] = {{{1, 1}, {1, 1}, {0, 0}, {1, 1}},
                                                 {{1, 1}, {1, 1}, {1, 1}, {0, 0}},
                                                 {{0, 0}, {1, 1}, {1, 1}, {1, 1}},
                                                 {{1, 1}, {0, 0}, {1, 1}, {1, 1}}};
static constexpr double WFV[4][4][2] = {
    {{0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.0}, {0.0, 0.0}, {0.0, 0.25 * (4.0 / 3.0)}},
    {{0.0, 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.0}, {0.0, 0.0}},
    {{0.0, 0.0}, {0.0, 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.0}},
    {{0.25 * (4.0 / 3.0), 0.0}, {0.0, 0.0}, {0.0, 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}}};
// The nfint=5 branch (vertex and non-vertex face Gauss points):
// Riemann corner weights (getriem_quadsubg_gauss, Getrhslaggauss.f90:16777-16796)
static constexpr double WR5[4][4][2] = {{{1, 1}, {.5, 1}, {0, 0}, {1, .5}},
                                        {{1, .5}, {1, 1}, {.5, 1}, {0, 0}},
                                        {{0, 0}, {1, .5}, {1, 1}, {.5, 1}},
                                        {{.5, 1}, {0, 0}, {1, .5}, {1, 1}}};
// face-integral weights of rhsifacedg_lagsubgq_gs (:17651-17700): halved
// 5-point Lobatto weights, momentum/energy with w3/1 and volume with w3/2
// at the faces' ends on the cell edges
constexpr double LW1 = 0.1, LW2 = 0.544444444444444444444, LW3 = 0.7111111111111111111111;
static constexpr double WM5[4][4][2] = {
    {{0.5 * LW1, 0.5 * LW1}, {0.5 * (LW3 / 1.0), 0.5 * LW3}, {0.0, 0.0}, {0.5 * LW3, 0.5 * (LW3 / 1.0)}},
    {{0.5 * LW3, 0.5 * (LW3 / 1.0)}, {0.5 * LW1, 0.5 * LW1}, {0.5 * (LW3 / 1.0), 0.5 * LW3}, {0.0, 0.0}},
    {{0.0, 0.0}, {0.5 * LW3, 0.5 * (LW3 / 1.0)}, {0.5 * LW1, 0.5 * LW1}, {0.5 * (LW3 / 1.0), 0.5 * LW3}},
    {{0.5 * (LW3 / 1.0), 0.5 * LW3}, {0.0, 0.0}, {0.5 * LW3, 0.5 * (LW3 / 1.0)}, {0.5 * LW1, 0.5 * LW1}}};
static constexpr double WV5[4][4][2] = {
    {{0.5 * LW1, 0.5 * LW1}, {0.5 * (LW3 / 2.0), 0.5 * LW3}, {0.0, 0.0}, {0.5 * LW3, 0.5 * (LW3 / 2.0)}},
    {{0.5 * LW3, 0.5 * (LW3 / 2.0)}, {0.5 * LW1, 0.5 * LW1}, {0.5 * (LW3 / 2.0), 0.5 * LW3}, {0.0, 0.0}},
    {{0.0, 0.0}, {0.5 * LW3, 0.5 * (LW3 / 2.0)}, {0.5 * LW1, 0.5 * LW1}, {0.5 * (LW3 / 2.0), 0.5 * LW3}},
    {{0.5 * (LW3 / 2.0), 0.5 * LW3}, {0.0, 0.0}, {0.5 * LW3, 0.5 * (LW3 / 2.0)}, {0.5 * LW1, 0.5 * LW1}}};
// the weight tables of branch NF (4: nfint=4, 5: nfint=5)
template <int NF>
__device__ __forceinline__ double wfr(int isg, int iv, int f) { return NF == 5 ? WR5[isg][iv][f] : WFR[isg][iv][f]; }
template <int NF>
__device__ __forceinline__ double wfm(int isg, int iv, int f) { return NF == 5 ? WM5[isg][iv][f] : WFM[isg][iv][f]; }
template <int NF>
__device__ __forceinline__ double wfv(int isg, int iv, int f) { return NF == 5 ? WV5[isg][iv][f] : WFV[isg][iv][f]; }
// face Lobatto point +-sqrt(3/7) as the reference writes it, and the
// cell-local face Gauss points g = 2 lf + k of face lf (corner lf -> lf+1),
// k = 0 at r = -G5A (posiq 13..20, :19443-19455), with their sub-cells
// (sglgq, :18217-18220)
constexpr double G5A = 0.6546536707079771437983;
static constexpr double PQG5[8][2] = {{-G5A, -1.0}, {G5A, -1.0}, {1.0, -G5A}, {1.0, G5A},
                                      {G5A, 1.0},  {-G5A, 1.0}, {-1.0, G5A}, {-1.0, -G5A}};
static constexpr int SGG5[8] = {0, 1, 1, 2, 2, 3, 3, 0};
// per cell and face Gauss point the face kernel's outputs, fg[(6 g + j) nc + ie]:
// j = 0, 1 the force, 2, 3 the correction force, 4, 5 the face velocity
#define FG5(g, j) ((6 * (g) + (j)))

// getvectf (Getrhslaggauss.f90:16547-16590): unit normal (dy/dr, -dx/dr)
// and 2 |dx/dr| of the quadratic face (corner a, corner b, mid-node) at its
// Lobatto point k
__device__ __forceinline__ void face_normal5(const double fx[3], const double fy[3], int k, double &nx, double &ny,
                                             double &ga)
{
    const double r = k == 0 ? -G5A : G5A;
    const double dd[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
    double dxdr = 0.0, dydr = 0.0;
    for (int i = 0; i < 3; ++i) {
        dxdr = dxdr + dd[i] * fx[i];
        dydr = dydr + dd[i] * fy[i];
    }
    const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
    const double anx = dydr, any = -dxdr;
    nx = anx / sqrt(anx * anx + any * any);
    ny = any / sqrt(anx * anx + any * any);
    ga = djaco * 2.0;
}

// 5x5 Gauss rule of ruqope (dgflo2d.f90:2734-2743 ordering)
static constexpr double PG5[5] = {-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                                           0.5384693101056830910363144, 0.9061798459386639927976269};
static constexpr double WG5[5] = {0.2369268850561890875142640, 0.4786286704993664680412915,
                                           0.5688888888888888888888889, 0.4786286704993664680412915,
                                           0.2369268850561890875142640};
// 5-point Gauss-Lobatto (ruqope_lobatto, dgflo2d.f90:2571-2583)
static constexpr double PL5[5] = {-1.0, -0.6546536707079771437983, 0.0, 0.6546536707079771437983, 1.0};
static constexpr double WL5[5] = {0.1, 0.544444444444444444444, 0.7111111111111111111111,
                                           0.544444444444444444444, 0.1};

// Shape-function table of the fixed quadrature points (k_q2_shape_tables):
// derivatives at 25 Gauss points, 4 x 25 sub-cell points and 25 Lobatto
// points, then the shape values at the 25 Gauss points.
constexpr int SHP_POINTS = 150;
constexpr int SHP_STRIDE = 18;   // dN/dr[9], dN/ds[9]

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

// Jacobian of a Q2 map from tabulated shape-function derivatives (the table
// is built by k_q2_shape_tables with shape9, so the sums equal jac9's).
__device__ __forceinline__ void jac9_tab(const double *xp, const double *yp, const double *t, double J[4])
{
    double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
    for (int k = 0; k < 9; ++k) {
        dxdr = dxdr + t[k] * xp[k];
        dxds = dxds + t[9 + k] * xp[k];
        dydr = dydr + t[k] * yp[k];
        dyds = dyds + t[9 + k] * yp[k];
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

// Face f of sub-cell corner (isg, iv): its slot's outward unit normal
// (signed by the slot orientation) and length from the face-geometry rows
// gq[(3 slot + c) nc + ie] the corner solve stores.
__device__ __forceinline__ void corner_face_geom(const double *__restrict__ gq, int nc, int ie, int isg, int iv,
                                                 int f, double &nx, double &ny, double &ga)
{
    const int tt = FNQSG[isg][2 * iv + f];
    const int sl = (tt > 0 ? tt : -tt) - 1;
    const double sg = tt > 0 ? 1.0 : -1.0;
    nx = sg * gq[(3 * sl + 0) * nc + ie];
    ny = sg * gq[(3 * sl + 1) * nc + ie];
    ga = gq[(3 * sl + 2) * nc + ie];
}

// Curved six-node triangles (cells 0 .. ntria-1 of a hybrid mesh): nodes
// 0-2 the corners counter-clockwise, 3-5 the mid-nodes of edges 1-2, 2-3,
// 3-1; four sub-triangles (iptsg, Getrhslaggauss.f90:12262-12265), their
// signed face slots among the triangle's 15 (fntsg, :12267-12285) and the
// corner weights times 1/4 (wfgst, :12287-12306, also the volume weights of
// the face integral, :8349-8367)
static constexpr int IPTSG[4][3] = {{0, 3, 5}, {3, 4, 5}, {3, 1, 4}, {5, 4, 2}};
static constexpr int FNTSG[4][6] = {{6, 1, 10, -7, -13, 12}, {7, 14, 8, 15, 9, 13},
                                    {-14, 10, 2, 3, 11, -8}, {12, -9, -15, 11, 4, 5}};
static constexpr double WT2[4][3][2] = {
    {{0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}, {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)},
     {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}},
    {{0.25 * (8.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * (8.0 / 3.0), 0.25 * (8.0 / 3.0)},
     {0.25 * (8.0 / 3.0), 0.25 * (8.0 / 3.0)}},
    {{0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)}, {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)},
     {0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}},
    {{0.25 * (4.0 / 3.0), 0.25 * (8.0 / 3.0)}, {0.25 * (8.0 / 3.0), 0.25 * (4.0 / 3.0)},
     {0.25 * (4.0 / 6.0), 0.25 * (4.0 / 6.0)}}};
static constexpr double XV6[6] = {0.0, 1.0, 0.0, 0.5, 0.5, 0.0};
static constexpr double YV6[6] = {0.0, 0.0, 1.0, 0.0, 0.5, 0.5};

// rutope's 7-point area rule (dgflo2d.f90:2234-2255), literal values, as
// the curved-triangle routines read it with ngausd = 7
static constexpr double PT7T_R[7] = {0.333333333333333, 0.059715871789770, 0.470142064105115, 0.470142064105115,
                                     0.797426985353087, 0.101286507323456, 0.101286507323456};
static constexpr double PT7T_S[7] = {0.333333333333333, 0.470142064105115, 0.059715871789770, 0.470142064105115,
                                     0.101286507323456, 0.797426985353087, 0.101286507323456};
static constexpr double WT7T[7] = {0.225000000000000, 0.132394152788506, 0.132394152788506, 0.132394152788506,
                                   0.125939180544827, 0.125939180544827, 0.125939180544827};

__device__ __forceinline__ void tri_face_geom(const double *__restrict__ gq, int nc, int ie, int isg, int iv, int f,
                                              double &nx, double &ny, double &ga)
{
    const int tt = FNTSG[isg][2 * iv + f];
    const int sl = (tt > 0 ? tt : -tt) - 1;
    const double sg = tt > 0 ? 1.0 : -1.0;
    nx = sg * gq[(3 * sl + 0) * nc + ie];
    ny = sg * gq[(3 * sl + 1) * nc + ie];
    ga = gq[(3 * sl + 2) * nc + ie];
}

// The Riemann record of a corner face (getriem_quadsubg2,
// Getrhslaggauss.f90:10639-10858): m u_c and A sigma n (A = the face length
// times the corner weight), with the reference's expressions, from the
// compact corner state; the node solve and the face integral rebuild them
// exactly as the corner solve formed them.
__device__ __forceinline__ void corner_face_rec(double m11, double uv, double vv, double pv, double nx, double ny,
                                                double area, double &m1, double &m2, double &s1, double &s2)
{
    m1 = m11 * uv + 0.0 * vv;
    m2 = 0.0 * uv + m11 * vv;
    s1 = -pv * area * nx + 0.0 * area * ny;
    s2 = 0.0 * area * nx + -pv * area * ny;
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

}  // namespace lag2d
#endif
