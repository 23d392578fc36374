// Discrete adjoint of the curved (Q2 sub-cell) step: the reverse of one
// SSP-RK stage of getsolvdg_lag on the ncurv = 1, nfint = 4 branch of
// getrhsdg_lag_mcnew (Getrhslagmc.f90:7556-7708), P1 and P2, with the WENO
// increment, the corner impedances' |du.n|, the max(eps, .) clamps and dt
// frozen at the recorded forward.
//
// The frozen stage, whose exact derivative these kernels compute:
//   Ul    = Uin + w                                  frozen WENO increment
//   rec, gq = corner(x_in, Ul)     getriem_quadsubg2 (Getrhslaggauss.f90:10409-10864),
//                                  sub-cell density getrhosubg_averg (:7067-7392)
//   u*    = node(rec)              getndvelo_lagsubg (:4348-4614)
//   r     = face(u*, gq, rec)      rhsifacedg_lagsubgq2 (:8501-8745)
//         + domain(x_in, Ul)       rhsdomndg_lagmc_curvquad (Getrhslag.f90:4716-5013)
//   U_out = alfa U^n + (1 - alfa)(Ul + dt M^-1 r)             (Getrhslag.f90:1822-1904)
//   x_pre = alfa x^n + (1 - alfa)(x_in + dt u*)                (:1954-1956)
//   x_out = x_pre, each cell's centre node set to -1/4 (corners) + 1/2 (mid-nodes)  (:1962-1971)
//
// Reverse order: node motion and centre reset (k_adjc_move), update and
// cell RHS (k_adjc_cellrhs), nodal solve (k_adjc_node), corners
// (k_adjc_corner), node sums of the per-cell coordinate adjoints
// (k_adjc_gather).  Node sums run over the node -> sub-cell-corner lists in
// a fixed order: no atomics, and the gradient is the same on every run.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "lag2d_host.h"
#include "lag2d_q2geom.h"

namespace lag2d {

namespace {

// FIRST[isg][iv]: local node IPQSG[isg][iv] appears in no sub-cell before
// isg, so a node's (cell, sub-cell corner) list names each (cell, local
// node) pair once by its first sub-cell.
static constexpr uint8_t FIRST[4][4] = {{1, 1, 1, 1}, {0, 1, 1, 0}, {0, 0, 1, 1}, {0, 0, 0, 1}};

// The corner faces that use face slot sl (the inverse of FNQSG: corner face
// cf = 2 (ivsg + 4 isg) + f names slot |FNQSG[isg][2 ivsg + f]| - 1), in
// increasing cf, -1 padded.
constexpr int kSlotUses[16][4] = {{1, -1, -1, -1},  {10, -1, -1, -1}, {11, -1, -1, -1}, {20, -1, -1, -1},
                                  {21, -1, -1, -1}, {30, -1, -1, -1}, {31, -1, -1, -1}, {0, -1, -1, -1},
                                  {3, 4, 8, 15},    {13, 14, 17, 18}, {16, 23, 27, 28}, {5, 6, 25, 26},
                                  {2, 9, -1, -1},   {12, 19, -1, -1}, {22, 29, -1, -1}, {7, 24, -1, -1}};
constexpr int kFnqsg[4][8] = {{8, 1, 13, -9, -9, 12, 12, 16},
                              {9, 13, 2, 3, 14, -10, -10, 9},
                              {-11, 10, 10, 14, 4, 5, 15, -11},
                              {16, -12, -12, 11, 11, 15, 6, 7}};
constexpr bool slot_uses_consistent()
{
    for (int sl = 0; sl < 16; ++sl) {
        int n = 0, last = -1;
        for (int cf = 0; cf < 32; ++cf) {
            const int t = kFnqsg[cf / 8][cf % 8];
            if ((t > 0 ? t : -t) - 1 != sl) continue;
            if (n >= 4 || kSlotUses[sl][n] != cf || cf <= last) return false;
            last = cf;
            ++n;
        }
        for (int j = n; j < 4; ++j)
            if (kSlotUses[sl][j] != -1) return false;
    }
    return true;
}
static_assert(slot_uses_consistent(), "kSlotUses must invert FNQSG");
static constexpr int SLOT_USES[16][4] = {{1, -1, -1, -1},  {10, -1, -1, -1}, {11, -1, -1, -1}, {20, -1, -1, -1},
                                           {21, -1, -1, -1}, {30, -1, -1, -1}, {31, -1, -1, -1}, {0, -1, -1, -1},
                                           {3, 4, 8, 15},    {13, 14, 17, 18}, {16, 23, 27, 28}, {5, 6, 25, 26},
                                           {2, 9, -1, -1},   {12, 19, -1, -1}, {22, 29, -1, -1}, {7, 24, -1, -1}};

// Row-symmetric inverse mass matrix of a cell (the packing of amatr used by
// the update: degree 0 decoupled, the rest the upper triangle row-wise).
template <int ND>
__device__ __forceinline__ void mass_inverse(const Dev &d, int ie, double M[ND][ND])
{
    const int nc = d.nc;
    if constexpr (ND == 6) {
        const int off[6] = {0, 0, 5, 9, 12, 14};
        for (int i = 0; i < 6; ++i)
            for (int k = 0; k < 6; ++k) {
                if (i == 0 || k == 0) {
                    M[i][k] = (i == 0 && k == 0) ? d.A[15 * nc + ie] : 0.0;
                } else {
                    const int a = i < k ? i : k, b = i < k ? k : i;
                    M[i][k] = d.A[(off[a] + (b - a)) * nc + ie];
                }
            }
    } else {
        M[0][0] = d.A[3 * nc + ie]; M[0][1] = 0.0; M[0][2] = 0.0;
        M[1][0] = 0.0; M[1][1] = d.A[0 * nc + ie]; M[1][2] = d.A[1 * nc + ie];
        M[2][0] = 0.0; M[2][1] = d.A[1 * nc + ie]; M[2][2] = d.A[2 * nc + ie];
    }
}

// lam of a 2x2 Jacobian J = (dx/dr, dx/ds, dy/dr, dy/ds) from lam of its
// determinant, pushed to the 9 nodes through the shape derivatives t.
__device__ __forceinline__ void det_reverse(const double J[4], double ldet, const double *t, double *lx, double *ly)
{
    const double l0 = ldet * J[3], l3 = ldet * J[0], l2 = -ldet * J[1], l1 = -ldet * J[2];
    for (int n = 0; n < 9; ++n) {
        lx[n] += t[n] * l0 + t[9 + n] * l1;
        ly[n] += t[n] * l2 + t[9 + n] * l3;
    }
}

}  // namespace

// ---- node motion and centre reset: lam of x_pre from lam of x_out, then
// x^n, x_in and u* parts.  A centre node's x_pre is overwritten by the reset.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adjc_move(Dev d, AdjStageC a, double alfa)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = d.np, nc = d.nc;
    if (p >= np) return;
    const double dt = *a.dt;
    double lx = a.LX[p], ly = a.LX[np + p];
    bool centre = false;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k], ie = e >> 4, slot = e & 15, isg = slot >> 2, iv = slot & 3;
        if (!FIRST[isg][iv]) continue;
        const int ln = IPQSG[isg][iv];
        if (ln == 8) {
            centre = true;
            continue;
        }
        const int pc = d.ipq[8 * nc + ie];
        const double c = ln < 4 ? -0.25 : 0.5;
        lx += c * a.LX[pc];
        ly += c * a.LX[np + pc];
    }
    if (centre) lx = ly = 0.0;
    a.LXN[p] += alfa * lx;
    a.LXN[np + p] += alfa * ly;
    a.LXI[p] = (1.0 - alfa) * lx;
    a.LXI[np + p] = (1.0 - alfa) * ly;
    a.LUS[p] = (1.0 - alfa) * dt * lx;
    a.LUS[np + p] = (1.0 - alfa) * dt * ly;
}

// ---- update, face integral and Lobatto domain integral, thread per cell.
// Writes lam of Ul (update and domain parts) to LUI, of the cell's nodes
// (domain part) to CX, of u* per sub-cell corner to CU, and of the records
// and face slots per corner face to LREC and LUSE; adds alfa lam(U_out)
// to lam(U^n).
template <int ND>
__global__ void __launch_bounds__(256) k_adjc_cellrhs(Dev d, AdjStageC a, double alfa)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    const double dt = *a.dt, g1 = d.gamlg - 1.0, eps = 1.0e-6;
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    const double gg0 = d.G[18 * nc + ie], gg1 = d.G[19 * nc + ie], gg2 = d.G[20 * nc + ie];
    double lr[4][ND], lu[4][ND];
    {
        double M[ND][ND];
        mass_inverse<ND>(d, ie, M);
        double lo[4][ND];
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < ND; ++id) {
                const size_t o = (size_t)(q * ND + id) * nc + ie;
                lo[q][id] = a.LU[o];
                a.LUN[o] += alfa * lo[q][id];
                lu[q][id] = (1.0 - alfa) * lo[q][id];
            }
        for (int q = 0; q < 4; ++q)
            for (int k = 0; k < ND; ++k) {
                double s = 0.0;
                for (int id = 0; id < ND; ++id) s += M[id][k] * lo[q][id];
                lr[q][k] = (1.0 - alfa) * dt * s;
            }
    }
    // ---- face integral
#pragma unroll 1
    for (int k = 0; k < 16; ++k) {
        const int isg = k >> 2, iv = k & 3;
        const int ln = IPQSG[isg][iv];
        const int p = d.ipq[ln * nc + ie];
        const double us1 = a.usx[p], us2 = a.usy[p];
        double bq[6];
        basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, bq);
        double L[4];
        for (int q = 0; q < 4; ++q) {
            double s = 0.0;
            for (int id = 0; id < ND; ++id) s += lr[q][id] * bq[id];
            L[q] = s;
        }
        double nx[2], ny[2], ar[2], fx[2], fy[2], m11[2], sg[2], w[2], lfx[2], lfy[2];
        const double uv = a.rec[RC2(ie, k, 2)], vv = a.rec[RC2(ie, k, 3)], pv = a.rec[RC2(ie, k, 4)];
        for (int f = 0; f < 2; ++f) {
            const int tt = FNQSG[isg][2 * iv + f];
            sg[f] = tt > 0 ? 1.0 : -1.0;
            double ga, m1, m2, s1, s2;
            corner_face_geom(a.gq, nc, ie, isg, iv, f, nx[f], ny[f], ga);
            ar[f] = ga * WFV[isg][iv][f];
            m11[f] = a.rec[RC2(ie, k, f)];
            corner_face_rec(m11[f], uv, vv, pv, nx[f], ny[f], ga * WFR[isg][iv][f], m1, m2, s1, s2);
            fx[f] = s1 + m11[f] * us1 + 0.0 * us2 - m1;
            fy[f] = s2 + 0.0 * us1 + m11[f] * us2 - m2;
            w[f] = WFM[isg][iv][f];
            lfx[f] = (L[1] + L[3] * us1) * w[f];
            lfy[f] = (L[2] + L[3] * us2) * w[f];
        }
        const double lus1 = L[0] * (nx[0] * ar[0] + nx[1] * ar[1]) + L[3] * (fx[0] * w[0] + fx[1] * w[1]) +
                            lfx[0] * m11[0] + lfx[1] * m11[1];
        const double lus2 = L[0] * (ny[0] * ar[0] + ny[1] * ar[1]) + L[3] * (fy[0] * w[0] + fy[1] * w[1]) +
                            lfy[0] * m11[0] + lfy[1] * m11[1];
        a.CU[(2 * k + 0) * nc + ie] = lus1;
        a.CU[(2 * k + 1) * nc + ie] = lus2;
        for (int f = 0; f < 2; ++f) {
            const int cf = f + 2 * k;
            a.LREC[(size_t)(cf * 5 + 0) * nc + ie] = lfx[f] * us1 + lfy[f] * us2;
            a.LREC[(size_t)(cf * 5 + 1) * nc + ie] = -lfx[f];
            a.LREC[(size_t)(cf * 5 + 2) * nc + ie] = -lfy[f];
            a.LREC[(size_t)(cf * 5 + 3) * nc + ie] = lfx[f];
            a.LREC[(size_t)(cf * 5 + 4) * nc + ie] = lfy[f];
            a.LUSE[(size_t)(cf * 3 + 0) * nc + ie] = sg[f] * (L[0] * us1 * ar[f]);
            a.LUSE[(size_t)(cf * 3 + 1) * nc + ie] = sg[f] * (L[0] * us2 * ar[f]);
            a.LUSE[(size_t)(cf * 3 + 2) * nc + ie] = WFV[isg][iv][f] * (L[0] * (us1 * nx[f] + us2 * ny[f]));
        }
    }
    // ---- domain integral
    double u[4][ND];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) u[q][id] = a.Ul[(size_t)(q * ND + id) * nc + ie];
    double xp[9], yp[9], lxp[9], lyp[9];
    load9(d, a.xin, a.yin, ie, xp, yp);
    for (int n = 0; n < 9; ++n) lxp[n] = lyp[n] = 0.0;
#pragma unroll 1
    for (int k = 0; k < 25; ++k) {
        const int i = k / 5, j = k - 5 * i;
        const double rr = PL5[i], ss = PL5[j], wi = WL5[i] * WL5[j];
        const double *t = d.shp + (125 + k) * SHP_STRIDE;
        double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
        for (int n = 0; n < 9; ++n) {
            dxdr = dxdr + t[n] * xp[n];
            dxds = dxds + t[9 + n] * xp[n];
            dydr = dydr + t[n] * yp[n];
            dyds = dyds + t[9 + n] * yp[n];
        }
        const double dbdr[6] = {0.0, 1.0, 0.0, rr - rc, 0.0, ss - sc};
        const double dbds[6] = {0.0, 0.0, 1.0, 0.0, ss - sc, rr - rc};
        double b[6];
        basis6(rr, ss, rc, sc, gg0, gg1, gg2, b);
        double un[4] = {0.0, 0.0, 0.0, 0.0};
        for (int id = 0; id < ND; ++id)
            for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
        const double rhoad = 1.0 / un[0];
        const double uu = un[1], vv = un[2];
        const double e = un[3] - 0.5 * (uu * uu + vv * vv);
        const bool clamped = a.dcl[(size_t)k * nc + ie] != 0;
        const double pres = clamped ? eps : g1 * rhoad * e;
        double luu = 0.0, lvv = 0.0, lp = 0.0, ldxdr = 0.0, ldxds = 0.0, ldydr = 0.0, ldyds = 0.0;
        for (int id = 0; id < ND; ++id) {
            const double av = dyds * dbdr[id] + (-dydr) * dbds[id];
            const double cv = (-dxds) * dbdr[id] + dxdr * dbds[id];
            const double l0 = lr[0][id], l1 = lr[1][id], l2 = lr[2][id], l3 = lr[3][id];
            const double kk = (-l0 + l3 * pres) * wi;
            const double lav = kk * uu + l1 * pres * wi;
            const double lcv = kk * vv + l2 * pres * wi;
            luu += kk * av;
            lvv += kk * cv;
            lp += (l1 * av + l2 * cv + l3 * (av * uu + cv * vv)) * wi;
            ldyds += lav * dbdr[id];
            ldydr -= lav * dbds[id];
            ldxds -= lcv * dbdr[id];
            ldxdr += lcv * dbds[id];
        }
        double lrho = 0.0, le = 0.0;
        if (!clamped) {
            lrho = lp * g1 * e;
            le = lp * g1 * rhoad;
            luu -= le * uu;
            lvv -= le * vv;
        }
        const double lun[4] = {-lrho * rhoad * rhoad, luu, lvv, le};
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < ND; ++id) lu[q][id] += lun[q] * b[id];
        for (int n = 0; n < 9; ++n) {
            lxp[n] += t[n] * ldxdr + t[9 + n] * ldxds;
            lyp[n] += t[n] * ldydr + t[9 + n] * ldyds;
        }
    }
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) a.LUI[(size_t)(q * ND + id) * nc + ie] = lu[q][id];
    for (int n = 0; n < 9; ++n) {
        a.CX[(2 * n + 0) * nc + ie] = lxp[n];
        a.CX[(2 * n + 1) * nc + ie] = lyp[n];
    }
}

// ---- nodal solve: u* = (sum m u - sum A sigma n) / sum m per component,
// walls zeroing a component.  NG = (lam_ux / M, lam_uy / M, -(.) . u*):
// the derivative of u* with respect to each record of the node.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adjc_node(Dev d, AdjStageC a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = d.np, nc = d.nc;
    if (p >= np) return;
    double lx = a.LUS[p], ly = a.LUS[np + p], m = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k], ie = e >> 4, slot = e & 15;
        lx += a.CU[(2 * slot + 0) * nc + ie];
        ly += a.CU[(2 * slot + 1) * nc + ie];
        for (int f = 0; f < 2; ++f) m = m + a.rec[RC2(ie, slot, f)];
    }
    const uint8_t fl = d.nflag[p];
    if (fl & 2) lx = 0.0;
    if (fl & 1) ly = 0.0;
    if (fl & 4) lx = ly = 0.0;
    const double gx = (fl & 4) ? 0.0 : lx / m, gy = (fl & 4) ? 0.0 : ly / m;
    a.NG[p] = gx;
    a.NG[np + p] = gy;
    a.NG[2 * np + p] = -(gx * a.usx[p] + gy * a.usy[p]);
}

// ---- corners, one lane per (cell, sub-cell): lam of the records (face
// integral part plus the node solve's) pushed through the corner states,
// the sub-cell density and mean specific volume, the centre state and the
// face geometry to the cell's nodes and DOFs.  Lanes exchange the face
// slots' adjoints through LDS (slots are shared by sub-cells) and sum
// their node and DOF adjoints across the cell's 4 lanes.
constexpr int ADJC_CELLS = 64;

template <int ND>
__global__ void __launch_bounds__(256) k_adjc_corner(Dev d, AdjStageC a)
{
    __shared__ double stab[125 * SHP_STRIDE];
    __shared__ double suse[ADJC_CELLS][32 * 3 + 1];
    for (int k = threadIdx.x; k < 125 * SHP_STRIDE; k += 256) stab[k] = d.shp[k];
    const int cl = threadIdx.x >> 2, isg = threadIdx.x & 3;
    const int ie = blockIdx.x * ADJC_CELLS + cl;
    const int nc = d.nc, np = d.np;
    const bool active = ie < nc;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    double xp[9], yp[9], lxp[9], lyp[9], lu[4][ND];
    for (int n = 0; n < 9; ++n) lxp[n] = lyp[n] = xp[n] = yp[n] = 0.0;
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) lu[q][id] = 0.0;
    __syncthreads();  // shape table staged
    if (active) {
        load9(d, a.xin, a.yin, ie, xp, yp);
        const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
        const double gg0 = d.G[18 * nc + ie], gg1 = d.G[19 * nc + ie], gg2 = d.G[20 * nc + ie];
        double u[4][ND];
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < ND; ++id) u[q][id] = a.Ul[(size_t)(q * ND + id) * nc + ie];
        // ---- forward: sub-cell map, area, mass-weighted mean specific volume
        const double pc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                                  {-1.0, .5},   {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
        const int ipsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
        double xs[9], ys[9];
        for (int k = 0; k < 4; ++k) {
            xs[k] = xp[IPQSG[isg][k]];
            ys[k] = yp[IPQSG[isg][k]];
            const int c = ipsc[isg][k];
            map9(xp, yp, pc[c][0], pc[c][1], &xs[4 + k], &ys[4 + k]);
        }
        xs[8] = -0.25 * (xs[0] + xs[1] + xs[2] + xs[3]) + 0.5 * (xs[4] + xs[5] + xs[6] + xs[7]);
        ys[8] = -0.25 * (ys[0] + ys[1] + ys[2] + ys[3]) + 0.5 * (ys[4] + ys[5] + ys[6] + ys[7]);
        const double sr[4] = {-1.0, 1.0, 1.0, -1.0}, ss[4] = {-1.0, -1.0, 1.0, 1.0};
        const double rhomc = u[0][0];
        double area = 0.0, masel = 0.0, volel = 0.0;
        for (int g = 0; g < 25; ++g) {
            const int i = g / 5, j = g - 5 * i;
            double J[4];
            jac9_tab(xs, ys, &stab[g * SHP_STRIDE], J);
            area = area + WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
            const double r = (PG5[i] + sr[isg]) / 2.0, s = (PG5[j] + ss[isg]) / 2.0;
            double b[6];
            basis6(r, s, rc, sc, gg0, gg1, gg2, b);
            double t = 0.0;
            for (int id = 0; id < ND; ++id) t = t + u[0][id] * b[id];
            t = rhomc + 1.0 * (t - rhomc);
            jac9_tab(xp, yp, &stab[(25 + 25 * isg + g) * SHP_STRIDE], J);
            const double dj = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            masel = masel + 1.0 / t * dj;
            volel = volel + dj;
        }
        const double msub = d.msub[isg * nc + ie];
        const double rsub = 0.0 + (1.0 / area) * msub;
        const double uq = volel / masel;
        const double uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
        const double rhoct = 1.0 / rhomc;
        const double ec = ectr - 0.5 * (uctr * uctr + vctr * vctr);
        const bool cpc = a.ccl[ie] != 0, csc = a.ccl[(size_t)nc + ie] != 0;
        const double pctr = cpc ? eps : g1 * rhoct * ec;
        const double sarg = csc ? eps : d.gamlg * pctr / rhoct;
        const double sdctr = sqrt(sarg);
        const double dcor = d.cdrho * (rsub - 1.0 / uq);
        // ---- reverse through the corner records
        double lrhoct = 0.0, lsdctr = 0.0, ldcor = 0.0;
        for (int iv = 0; iv < 4; ++iv) {
            const int ln = IPQSG[isg][iv];
            double b[6];
            basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, b);
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < ND; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
            const int p = d.ipq[ln * nc + ie];
            const double rhovt = 1.0 / un[0];
            const double rhovsg = rhovt + dcor;
            const double uv = un[1], vv = un[2], ev = un[3];
            const double e = ev - 0.5 * (uv * uv + vv * vv);
            const bool cpv = a.ccl[(size_t)(2 + 4 * isg + iv) * nc + ie] != 0;
            const double pv = cpv ? eps : g1 * rhovsg * e;
            const double gx = a.NG[p], gy = a.NG[np + p], gm = a.NG[2 * np + p];
            double lpv = 0.0, luv = 0.0, lvv = 0.0;
            for (int f = 0; f < 2; ++f) {
                const int tt = FNQSG[isg][2 * iv + f];
                const int sl = (tt > 0 ? tt : -tt) - 1;
                const double sg = tt > 0 ? 1.0 : -1.0;
                const double nx = sg * a.gq[(3 * sl + 0) * nc + ie], ny = sg * a.gq[(3 * sl + 1) * nc + ie];
                const double ar = a.gq[(3 * sl + 2) * nc + ie] * WFR[isg][iv][f];
                const int cf = f + 2 * (iv + 4 * isg);
                const double deltu = a.dlt[(size_t)cf * nc + ie];
                const double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
                const double m11 = mur * ar;
                const double L0 = a.LREC[(size_t)(cf * 5 + 0) * nc + ie] + gm;
                const double L1 = a.LREC[(size_t)(cf * 5 + 1) * nc + ie] + gx;
                const double L2 = a.LREC[(size_t)(cf * 5 + 2) * nc + ie] + gy;
                const double L3 = a.LREC[(size_t)(cf * 5 + 3) * nc + ie] - gx;
                const double L4 = a.LREC[(size_t)(cf * 5 + 4) * nc + ie] - gy;
                const double lm11 = L0 + L1 * uv + L2 * vv;
                luv += L1 * m11;
                lvv += L2 * m11;
                lpv += -(L3 * ar * nx + L4 * ar * ny);
                const double lar = -pv * (L3 * nx + L4 * ny) + lm11 * mur;
                const double lmur = lm11 * ar;
                lrhoct += lmur * (sdctr + d.cimpd * d.slpdu * deltu);
                lsdctr += lmur * rhoct;
                const size_t uo = (size_t)(cf * 3) * nc + ie;
                suse[cl][3 * cf + 0] = sg * (-L3 * pv * ar) + a.LUSE[uo];
                suse[cl][3 * cf + 1] = sg * (-L4 * pv * ar) + a.LUSE[uo + nc];
                suse[cl][3 * cf + 2] = WFR[isg][iv][f] * lar + a.LUSE[uo + 2 * (size_t)nc];
            }
            double lrhovsg = 0.0, lev = 0.0;
            if (!cpv) {
                lrhovsg = lpv * g1 * e;
                lev = lpv * g1 * rhovsg;
                luv -= lev * uv;
                lvv -= lev * vv;
            }
            ldcor += lrhovsg;
            const double lun0 = -lrhovsg * rhovt * rhovt;
            for (int id = 0; id < ND; ++id) {
                lu[0][id] += lun0 * b[id];
                lu[1][id] += luv * b[id];
                lu[2][id] += lvv * b[id];
                lu[3][id] += lev * b[id];
            }
        }
        // ---- dcor = cdrho (rsub - 1/uq), the centre state
        const double lrsub = d.cdrho * ldcor;
        const double luq = d.cdrho * ldcor / (uq * uq);
        {
            const double ls = csc ? 0.0 : lsdctr / (2.0 * sdctr);
            const double lpctr = ls * d.gamlg / rhoct;
            lrhoct += -ls * d.gamlg * pctr / (rhoct * rhoct);
            if (!cpc) {
                lrhoct += lpctr * g1 * ec;
                const double le = lpctr * g1 * rhoct;
                lu[3][0] += le;
                lu[1][0] += -le * uctr;
                lu[2][0] += -le * vctr;
            }
            lu[0][0] += -lrhoct * rhoct * rhoct;
        }
        // ---- rsub = msub / area, uq = volel / masel, over the 25 points
        const double larea = -lrsub * msub / (area * area);
        const double lvolel = luq / masel, lmasel = -luq * volel / (masel * masel);
        double lxs[9], lys[9];
        for (int n = 0; n < 9; ++n) lxs[n] = lys[n] = 0.0;
        for (int g = 0; g < 25; ++g) {
            const int i = g / 5, j = g - 5 * i;
            const double r = (PG5[i] + sr[isg]) / 2.0, s = (PG5[j] + ss[isg]) / 2.0;
            double b[6];
            basis6(r, s, rc, sc, gg0, gg1, gg2, b);
            double t = 0.0;
            for (int id = 0; id < ND; ++id) t = t + u[0][id] * b[id];
            t = rhomc + 1.0 * (t - rhomc);
            double J[4];
            const double *tp = &stab[(25 + 25 * isg + g) * SHP_STRIDE];
            jac9_tab(xp, yp, tp, J);
            const double w = WG5[i] * WG5[j];
            const double dj = w * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            const double ldj = lmasel / t + lvolel;
            const double lt = -lmasel * dj / (t * t);
            for (int id = 0; id < ND; ++id) lu[0][id] += lt * b[id];
            det_reverse(J, ldj * w * 0.25, tp, lxp, lyp);
            const double *ts = &stab[g * SHP_STRIDE];
            jac9_tab(xs, ys, ts, J);
            det_reverse(J, larea * w, ts, lxs, lys);
        }
        for (int k = 0; k < 4; ++k) {
            lxs[k] += -0.25 * lxs[8];
            lys[k] += -0.25 * lys[8];
            lxs[4 + k] += 0.5 * lxs[8];
            lys[4 + k] += 0.5 * lys[8];
        }
        for (int k = 0; k < 4; ++k) {
            lxp[IPQSG[isg][k]] += lxs[k];
            lyp[IPQSG[isg][k]] += lys[k];
            const int c = ipsc[isg][k];
            double sh[9], dr_[9], ds_[9];
            shape9(pc[c][0], pc[c][1], sh, dr_, ds_);
            for (int n = 0; n < 9; ++n) {
                lxp[n] += sh[n] * lxs[4 + k];
                lyp[n] += sh[n] * lys[4 + k];
            }
        }
    }
    __syncthreads();  // every corner face's slot adjoint staged
    if (active) {
        // ---- face slots 4 isg .. 4 isg + 3
        for (int uo = 0; uo < 4; ++uo) {
            const int slot = 4 * isg + uo;
            double lg[3] = {0.0, 0.0, 0.0};
            for (int j = 0; j < 4; ++j) {
                const int cf = SLOT_USES[slot][j];
                if (cf < 0) break;
                for (int c = 0; c < 3; ++c) lg[c] += suse[cl][3 * cf + c];
            }
            if (slot >= 8 && slot < 12) {
                const int k = slot - 8;
                const double pq[4][2] = {{0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
                double sh[9], dr_[9], ds_[9];
                shape9(pq[k][0], pq[k][1], sh, dr_, ds_);
                double xv = 0.0, yv = 0.0;
                for (int n = 0; n < 9; ++n) {
                    xv = xv + sh[n] * xp[n];
                    yv = yv + sh[n] * yp[n];
                }
                const double fx[3] = {xp[4 + k], xp[8], xv};
                const double fy[3] = {yp[4 + k], yp[8], yv};
                const double dd[3] = {-1.5, -0.5, 2.0};
                double dxdr = 0.0, dydr = 0.0;
                for (int i = 0; i < 3; ++i) {
                    dxdr = dxdr + dd[i] * fx[i];
                    dydr = dydr + dd[i] * fy[i];
                }
                const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
                const double anx = -dydr, any = dxdr;
                const double nn = sqrt(anx * anx + any * any);
                const double gx = anx / nn, gy = any / nn;
                const double gdot = lg[0] * gx + lg[1] * gy;
                const double lanx = (lg[0] - gdot * gx) / nn, lany = (lg[1] - gdot * gy) / nn;
                const double ldj = 2.0 * lg[2];
                const double ldxdr = lany + ldj * dxdr / djaco;
                const double ldydr = -lanx + ldj * dydr / djaco;
                lxp[4 + k] += dd[0] * ldxdr;
                lyp[4 + k] += dd[0] * ldydr;
                lxp[8] += dd[1] * ldxdr;
                lyp[8] += dd[1] * ldydr;
                for (int n = 0; n < 9; ++n) {
                    lxp[n] += sh[n] * (dd[2] * ldxdr);
                    lyp[n] += sh[n] * (dd[2] * ldydr);
                }
            } else {
                const int ig = slot < 8 ? slot : slot - 4;
                const double pi[12][2] = {{-1, -1}, {1, -1}, {1, -1}, {1, 1}, {1, 1}, {-1, 1},
                                          {-1, 1},  {-1, -1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
                const double vn[12][2] = {{0, -1}, {0, -1}, {1, 0}, {1, 0}, {0, 1}, {0, 1},
                                          {-1, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
                double sh[9], dr_[9], ds_[9];
                shape9(pi[ig][0], pi[ig][1], sh, dr_, ds_);
                double J[4] = {0.0, 0.0, 0.0, 0.0};
                for (int n = 0; n < 9; ++n) {
                    J[0] = J[0] + dr_[n] * xp[n];
                    J[1] = J[1] + ds_[n] * xp[n];
                    J[2] = J[2] + dr_[n] * yp[n];
                    J[3] = J[3] + ds_[n] * yp[n];
                }
                const double anx = J[3] * vn[ig][0] + (-J[2]) * vn[ig][1];
                const double any = (-J[1]) * vn[ig][0] + J[0] * vn[ig][1];
                const double w1 = anx * 2.0, w2 = any * 2.0;
                const double F = sqrt(w1 * w1 + w2 * w2);
                const double gx = w1 / F, gy = w2 / F;
                const double gdot = lg[0] * gx + lg[1] * gy;
                const double lw1 = (lg[0] - gdot * gx) / F + lg[2] * gx;
                const double lw2 = (lg[1] - gdot * gy) / F + lg[2] * gy;
                const double lanx = 2.0 * lw1, lany = 2.0 * lw2;
                const double lJ3 = lanx * vn[ig][0], lJ2 = -lanx * vn[ig][1];
                const double lJ1 = -lany * vn[ig][0], lJ0 = lany * vn[ig][1];
                for (int n = 0; n < 9; ++n) {
                    lxp[n] += dr_[n] * lJ0 + ds_[n] * lJ1;
                    lyp[n] += dr_[n] * lJ2 + ds_[n] * lJ3;
                }
            }
        }
    }
    // ---- sums over the 4 lanes of the cell; lane 0 adds them to CX, LUI
    for (int n = 0; n < 9; ++n) {
        lxp[n] += __shfl_xor(lxp[n], 1);
        lxp[n] += __shfl_xor(lxp[n], 2);
        lyp[n] += __shfl_xor(lyp[n], 1);
        lyp[n] += __shfl_xor(lyp[n], 2);
    }
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) {
            lu[q][id] += __shfl_xor(lu[q][id], 1);
            lu[q][id] += __shfl_xor(lu[q][id], 2);
        }
    if (!active || isg != 0) return;
    for (int n = 0; n < 9; ++n) {
        a.CX[(2 * n + 0) * nc + ie] += lxp[n];
        a.CX[(2 * n + 1) * nc + ie] += lyp[n];
    }
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) a.LUI[(size_t)(q * ND + id) * nc + ie] += lu[q][id];
}

// ---- node sums of the cells' coordinate adjoints into lam(x_in)
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adjc_gather(Dev d, AdjStageC a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const int np = d.np, nc = d.nc;
    if (p >= np) return;
    double lx = 0.0, ly = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k], ie = e >> 4, slot = e & 15, isg = slot >> 2, iv = slot & 3;
        if (!FIRST[isg][iv]) continue;
        const int ln = IPQSG[isg][iv];
        lx += a.CX[(2 * ln + 0) * nc + ie];
        ly += a.CX[(2 * ln + 1) * nc + ie];
    }
    a.LXI[p] += lx;
    a.LXI[np + p] += ly;
}

// ---- end of a step: lam of the step's input = the first stage's input
// part plus the accumulated U^n / x^n parts; the accumulators are cleared.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adjc_step_end(Dev d, AdjStageC a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t nu = (size_t)4 * d.nd * d.nc;
    if ((size_t)t < nu) {
        a.LU[t] = a.LUN[t] + a.LUI[t];
        a.LUN[t] = 0.0;
    }
    if (t < 2 * d.np) {
        a.LX[t] = a.LXN[t] + a.LXI[t];
        a.LXN[t] = 0.0;
    }
}

}  // namespace lag2d

using namespace lag2d;

static inline int nblk(size_t n) { return (int)((n + LAG2D_BLOCK - 1) / LAG2D_BLOCK); }

int launch_adjq2_stage(lag2d_ctx *x, AdjStageC &a, double alfa)
{
    Dev &d = x->d;
    const int nc = x->nc, np = x->np;
    x->tic(KID_ADJ_NODES);
    hipLaunchKernelGGL(k_adjc_move, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, d, a, alfa);
    x->toc(KID_ADJ_NODES);
    x->tic(KID_ADJ_CELLRHS);
    if (x->nd == 3)
        hipLaunchKernelGGL(k_adjc_cellrhs<3>, dim3(nblk(nc)), dim3(256), 0, x->stream, d, a, alfa);
    else
        hipLaunchKernelGGL(k_adjc_cellrhs<6>, dim3(nblk(nc)), dim3(256), 0, x->stream, d, a, alfa);
    x->toc(KID_ADJ_CELLRHS);
    x->tic(KID_ADJ_NODES);
    hipLaunchKernelGGL(k_adjc_node, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, d, a);
    x->toc(KID_ADJ_NODES);
    x->tic(KID_ADJ_CORNER);
    const dim3 gc((nc + ADJC_CELLS - 1) / ADJC_CELLS);
    if (x->nd == 3)
        hipLaunchKernelGGL(k_adjc_corner<3>, gc, dim3(256), 0, x->stream, d, a);
    else
        hipLaunchKernelGGL(k_adjc_corner<6>, gc, dim3(256), 0, x->stream, d, a);
    x->toc(KID_ADJ_CORNER);
    x->tic(KID_ADJ_NODES);
    hipLaunchKernelGGL(k_adjc_gather, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, d, a);
    x->toc(KID_ADJ_NODES);
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}

int launch_adjq2_step_end(lag2d_ctx *x, AdjStageC &a)
{
    const size_t n = std::max((size_t)4 * x->nd * x->nc, (size_t)2 * x->np);
    hipLaunchKernelGGL(k_adjc_step_end, dim3(nblk(n)), dim3(LAG2D_BLOCK), 0, x->stream, x->d, a);
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}
