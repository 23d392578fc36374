// Discrete adjoint of one SSP-RK2 step of the straight P1 planar path
// (quads, slip walls) with the limiter coefficients and the time step frozen
// at their forward values.  No counterpart in the reference.  Each kernel
// reverses one forward kernel of a stage, recomputing the forward
// intermediates of its cell or node from the stage's stored inputs:
//
//   forward (lag2d_q1.hip, lag2d_common.hip)     reverse (here)
//   k_nodemove   x_out = a x^n + (1-a)(x_in + dt u*)      k_adj_nodemove
//   k_cellrhs_q1 face integral, domain integral,        k_adj_cellrhs
//                limiting, U_out = a U^n + (1-a)(L U_in + M^-1 dt R)
//   k_node_q1    u* = (sum m u_c - sum A s n) / sum m     k_adj_node
//   k_limit_corner_q1 corner records m, m u_c, A s n    k_adj_corner
//                                                        k_adj_gather (node sums)
//
// Cell kernels write per-corner contributions to node adjoints, which node
// kernels gather through the node->corner CSR: no atomics, deterministic.
// The reference routines differentiated are getriem_quad_matrixsym and
// getriecoef_matrixnew (Limiterlag.f90:708-1076), getndvelo_lag_mc_matrixsym
// (:408-704), rhsifacedg_lag_mc_hybridquad2 (Getrhslagmc.f90:14468-14618),
// rhsdomndg_lag_mc_quadsym (Limiterlag.f90:1108-1490) and the update and
// node motion of getsolvdg_lag (Getrhslag.f90:1804-1956).
#include "lag2d_dev.h"
#include "lag2d_kernels.h"

namespace lag2d {

namespace {

__device__ __forceinline__ double xvq_(int iv) { return (iv == 1 || iv == 2) ? 1.0 : -1.0; }
__device__ __forceinline__ double yvq_(int iv) { return (iv >= 2) ? 1.0 : -1.0; }

static constexpr double AG5P[5] = {-0.9061798459386639927976269, -0.5384693101056830910363144, 0.0,
                                     0.5384693101056830910363144, 0.9061798459386639927976269};
static constexpr double AG5W[5] = {0.2369268850561890875142640, 0.4786286704993664680412915,
                                     0.5688888888888888888888889, 0.4786286704993664680412915,
                                     0.2369268850561890875142640};

struct CellIn {
    int ip[4];
    double x[4], y[4];
    double U[4][3];
    double al[5], af[4];
    double rc, sc;
};

__device__ __forceinline__ void load_cell(const Dev &d, const AdjStage &a, int ie, CellIn &c)
{
    const int nc = d.nc;
    for (int k = 0; k < 4; ++k) {
        c.ip[k] = d.ipq[k * nc + ie];
        c.x[k] = a.xin[c.ip[k]];
        c.y[k] = a.yin[c.ip[k]];
    }
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) c.U[q][id] = a.Uin[(q * 3 + id) * nc + ie];
    for (int k = 0; k < 5; ++k) c.al[k] = a.lim[k * nc + ie];
    for (int k = 0; k < 4; ++k) c.af[k] = a.lim[(5 + k) * nc + ie];
    c.rc = d.G[0 * nc + ie];
    c.sc = d.G[1 * nc + ie];
}

// edge e joins corners e and e+1: a = (y_b - y_a, -(x_b - x_a)), |a|, a/|a|
struct Edges {
    double anx[4], any[4], L[4], nx[4], ny[4];
};

__device__ __forceinline__ void edges(const CellIn &c, Edges &g)
{
    for (int e = 0; e < 4; ++e) {
        const int a = e, b = (e + 1) & 3;
        g.anx[e] = c.y[b] - c.y[a];
        g.any[e] = -(c.x[b] - c.x[a]);
        g.L[e] = sqrt(g.anx[e] * g.anx[e] + g.any[e] * g.any[e]);
        g.nx[e] = g.anx[e] / g.L[e];
        g.ny[e] = g.any[e] / g.L[e];
    }
}

// reverse of the edge geometry: adjoints of (a, |a|, a/|a|) onto the corners
__device__ __forceinline__ void edges_rev(const Edges &g, const double *Lanx0, const double *Lany0, const double *LL,
                                          const double *Lnx, const double *Lny, double *Lx, double *Ly)
{
    for (int e = 0; e < 4; ++e) {
        const int a = e, b = (e + 1) & 3;
        const double L = g.L[e], L3 = L * L * L;
        const double anx = g.anx[e], any = g.any[e];
        const double Lanx = Lanx0[e] + Lnx[e] * (1.0 / L - anx * anx / L3) - Lny[e] * any * anx / L3 + LL[e] * anx / L;
        const double Lany = Lany0[e] - Lnx[e] * anx * any / L3 + Lny[e] * (1.0 / L - any * any / L3) + LL[e] * any / L;
        Ly[b] += Lanx;
        Ly[a] -= Lanx;
        Lx[b] -= Lany;
        Lx[a] += Lany;
    }
}

}  // namespace

// Node motion reversed: lam_u* of the stage = lam_u*(next stage's corner
// input) + (1-a) dt lam_x_out; lam_x^n += a lam_x_out; lam_x_in (direct
// part) = (1-a) lam_x_out.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adj_nodemove(Dev d, AdjStage a, double alfa)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int np = d.np;
    const double dt = *a.dt;
    for (int c = 0; c < 2; ++c) {
        const double lx = a.LX[c * np + p];
        a.LUS[c * np + p] = a.LUSN[c * np + p] + (1.0 - alfa) * dt * lx;
        a.LXN[c * np + p] = a.LXN[c * np + p] + alfa * lx;
        a.LXI[c * np + p] = (1.0 - alfa) * lx;
    }
}

// Update, limiting, face and domain integrals of one quad reversed.
// In: lam_U_out (LU).  Out: LUN += a lam_U_out; LUI = lam_U_in; per corner
// lam_x_in (CX), lam_u* (CU); lam of the corner records (LREC).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adj_cellrhs(Dev d, AdjStage a, double alfa)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    const double c13 = 0.3333333333333333, c16 = 0.1666666666666666;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    CellIn c;
    load_cell(d, a, ie, c);
    Edges g;
    edges(c, g);
    const double dt = *a.dt;
    double us[4], vs[4];
    for (int k = 0; k < 4; ++k) {
        us[k] = a.usx[c.ip[k]];
        vs[k] = a.usy[c.ip[k]];
    }
    // update: U_out = a U^n + (1-a)(ul + M dt R)
    double G[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) {
            const int o = (q * 3 + id) * nc + ie;
            const double l = a.LU[o];
            a.LUN[o] = a.LUN[o] + alfa * l;
            G[q][id] = (1.0 - alfa) * l;
        }
    const double A0 = d.A[0 * nc + ie], A1 = d.A[1 * nc + ie], A2 = d.A[2 * nc + ie], A3 = d.A[3 * nc + ie];
    double Lr[4][3];
    for (int q = 0; q < 4; ++q) {
        Lr[q][0] = dt * A3 * G[q][0];
        Lr[q][1] = dt * (A0 * G[q][1] + A1 * G[q][2]);
        Lr[q][2] = dt * (A1 * G[q][1] + A2 * G[q][2]);
    }
    double LU[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) LU[q][id] = 0.0;
    const double al0 = c.al[0], al3 = c.al[3], al4 = c.al[4];
    const double af0 = c.af[0], af1 = c.af[1], af2 = c.af[2], af3 = c.af[3];
    // limited DOFs ul
    LU[0][0] += G[0][0];
    LU[0][1] += G[0][1] * al0;
    LU[0][2] += G[0][2] * al0;
    LU[1][0] += G[1][0];
    LU[2][0] += G[2][0];
    LU[3][0] += G[3][0];
    LU[3][1] += G[3][1] * al4;
    LU[3][2] += G[3][2] * al4;
    double Ldudr = G[1][1], Lduds = G[1][2], Ldvdr = G[2][1], Ldvds = G[2][2];

    // face integral (lumped corner weights 2 b_i + b_i)
    double Lx[4] = {0, 0, 0, 0}, Ly[4] = {0, 0, 0, 0};
    double Lanx[4] = {0, 0, 0, 0}, Lany[4] = {0, 0, 0, 0};
    for (int iv = 0; iv < 4; ++iv) {
        const double b[3] = {1.0, (xvq_(iv) - c.rc) / 1.0, (yvq_(iv) - c.sc) / 1.0};
        const int e0 = (iv + 3) & 3, e1 = iv;
        double m11[2], f[2][2];
        for (int fa = 0; fa < 2; ++fa) {
            const int cf = fa + 2 * iv;
            m11[fa] = a.rec[(0 * 8 + cf) * nc + ie];
            const double mu1 = a.rec[(1 * 8 + cf) * nc + ie], mu2 = a.rec[(2 * 8 + cf) * nc + ie];
            const double sn1 = a.rec[(3 * 8 + cf) * nc + ie], sn2 = a.rec[(4 * 8 + cf) * nc + ie];
            f[fa][0] = sn1 + m11[fa] * us[iv] - mu1;
            f[fa][1] = sn2 + m11[fa] * vs[iv] - mu2;
        }
        double Lus = 0.0, Lvs = 0.0, Lf[2][2] = {{0, 0}, {0, 0}};
        for (int id = 0; id < 3; ++id) {
            const double w = 2.0 * b[id] + b[id];
            const double l0 = Lr[0][id] * c16 * w;
            Lus += l0 * (g.anx[e0] + g.anx[e1]);
            Lvs += l0 * (g.any[e0] + g.any[e1]);
            Lanx[e0] += l0 * us[iv];
            Lany[e0] += l0 * vs[iv];
            Lanx[e1] += l0 * us[iv];
            Lany[e1] += l0 * vs[iv];
            const double l1 = Lr[1][id] * c13 * w, l2 = Lr[2][id] * c13 * w, l3 = Lr[3][id] * c13 * w;
            Lf[0][0] += l1 + l3 * us[iv];
            Lf[1][0] += l1 + l3 * us[iv];
            Lf[0][1] += l2 + l3 * vs[iv];
            Lf[1][1] += l2 + l3 * vs[iv];
            Lus += l3 * (f[0][0] + f[1][0]);
            Lvs += l3 * (f[0][1] + f[1][1]);
        }
        for (int fa = 0; fa < 2; ++fa) {
            const int cf = fa + 2 * iv;
            a.LREC[(0 * 8 + cf) * nc + ie] = Lf[fa][0] * us[iv] + Lf[fa][1] * vs[iv];
            a.LREC[(1 * 8 + cf) * nc + ie] = -Lf[fa][0];
            a.LREC[(2 * 8 + cf) * nc + ie] = -Lf[fa][1];
            a.LREC[(3 * 8 + cf) * nc + ie] = Lf[fa][0];
            a.LREC[(4 * 8 + cf) * nc + ie] = Lf[fa][1];
            Lus += Lf[fa][0] * m11[fa];
            Lvs += Lf[fa][1] * m11[fa];
        }
        a.CU[(0 * 4 + iv) * nc + ie] = Lus;
        a.CU[(1 * 4 + iv) * nc + ie] = Lvs;
    }

    // domain integral -int grad(b) . F on the 5x5 Gauss rule
    const double tau = c.U[0][0], rhoct = 1.0 / tau;
    const double uctr = c.U[1][0], vctr = c.U[2][0], ectr = c.U[3][0];
    const double pin = g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr));
    const bool pcl = !(eps > pin);
    const double pctr = pcl ? pin : eps;
    const double dudr = af0 * c.U[1][1] + af2 * c.U[2][1];
    const double duds = af0 * c.U[1][2] + af2 * c.U[2][2];
    const double dvdr = af1 * c.U[1][1] + af3 * c.U[2][1];
    const double dvds = af1 * c.U[1][2] + af3 * c.U[2][2];
    double Lpctr = 0.0;
    for (int gi = 0; gi < 5; ++gi)
        for (int gj = 0; gj < 5; ++gj) {
            const double rr = AG5P[gi], ss = AG5P[gj], wi = AG5W[gi] * AG5W[gj];
            const double rp = 1.0 + rr, rm = 1.0 - rr, sp = 1.0 + ss, sm = 1.0 - ss;
            const double dsr[4] = {-0.25 * sm, 0.25 * sm, 0.25 * sp, -0.25 * sp};
            const double dss[4] = {-0.25 * rm, -0.25 * rp, 0.25 * rp, 0.25 * rm};
            double dxdr = 0.0, dxds = 0.0, dydr = 0.0, dyds = 0.0;
            for (int k = 0; k < 4; ++k) {
                dxdr += dsr[k] * c.x[k];
                dxds += dss[k] * c.x[k];
                dydr += dsr[k] * c.y[k];
                dyds += dss[k] * c.y[k];
            }
            const double gd0[3] = {0.0, dyds, -dydr}, gd1[3] = {0.0, -dxds, dxdr};
            const double b[3] = {1.0, rr - c.rc, ss - c.sc};
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < 3; ++id)
                for (int q = 0; q < 4; ++q) un[q] += c.U[q][id] * b[id];
            const double rhoad = 1.0 / un[0];
            const double ek = un[3] - 0.5 * (un[1] * un[1] + un[2] * un[2]);
            const double praw = g1 * rhoad * ek;
            const double pres = pctr + al3 * (praw - pctr);
            const double uadv = uctr + dudr * b[1] + duds * b[2];
            const double vadv = vctr + dvdr * b[1] + dvds * b[2];
            const double s = -pres;
            double Lgd0[3], Lgd1[3], Luadv = 0.0, Lvadv = 0.0, Ls = 0.0;
            for (int id = 0; id < 3; ++id) {
                const double Lf0 = -wi * Lr[0][id], Lf1 = -wi * Lr[1][id];
                const double Lf2 = -wi * Lr[2][id], Lf3 = -wi * Lr[3][id];
                Lgd0[id] = Lf0 * uadv + Lf1 * s + Lf3 * s * uadv;
                Lgd1[id] = Lf0 * vadv + Lf2 * s + Lf3 * s * vadv;
                Luadv += Lf0 * gd0[id] + Lf3 * gd0[id] * s;
                Lvadv += Lf0 * gd1[id] + Lf3 * gd1[id] * s;
                Ls += Lf1 * gd0[id] + Lf2 * gd1[id] + Lf3 * (gd0[id] * uadv + gd1[id] * vadv);
            }
            const double Lpres = -Ls;
            Lpctr += (1.0 - al3) * Lpres;
            const double Lpraw = al3 * Lpres;
            const double Lrhoad = Lpraw * g1 * ek;
            double Lun[4];
            Lun[3] = Lpraw * g1 * rhoad;
            Lun[1] = -Lpraw * g1 * rhoad * un[1];
            Lun[2] = -Lpraw * g1 * rhoad * un[2];
            Lun[0] = -Lrhoad * rhoad * rhoad;
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < 3; ++id) LU[q][id] += Lun[q] * b[id];
            LU[1][0] += Luadv;
            Ldudr += Luadv * b[1];
            Lduds += Luadv * b[2];
            LU[2][0] += Lvadv;
            Ldvdr += Lvadv * b[1];
            Ldvds += Lvadv * b[2];
            const double Ldyds = Lgd0[1], Ldydr = -Lgd0[2], Ldxds = -Lgd1[1], Ldxdr = Lgd1[2];
            for (int k = 0; k < 4; ++k) {
                Lx[k] += dsr[k] * Ldxdr + dss[k] * Ldxds;
                Ly[k] += dsr[k] * Ldydr + dss[k] * Ldyds;
            }
        }
    if (pcl) {
        const double Lrhoct = Lpctr * g1 * (ectr - 0.5 * (uctr * uctr + vctr * vctr));
        LU[3][0] += Lpctr * g1 * rhoct;
        LU[1][0] -= Lpctr * g1 * rhoct * uctr;
        LU[2][0] -= Lpctr * g1 * rhoct * vctr;
        LU[0][0] -= Lrhoct / (tau * tau);
    }
    LU[1][1] += af0 * Ldudr + af1 * Ldvdr;
    LU[2][1] += af2 * Ldudr + af3 * Ldvdr;
    LU[1][2] += af0 * Lduds + af1 * Ldvds;
    LU[2][2] += af2 * Lduds + af3 * Ldvds;
    const double zero[4] = {0, 0, 0, 0};
    edges_rev(g, Lanx, Lany, zero, zero, zero, Lx, Ly);
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) a.LUI[(q * 3 + id) * nc + ie] = LU[q][id];
    for (int k = 0; k < 4; ++k) {
        a.CX[(0 * 4 + k) * nc + ie] = Lx[k];
        a.CX[(1 * 4 + k) * nc + ie] = Ly[k];
    }
}

// Nodal solve reversed: lam_u* (this stage's node motion and the cells'
// face integrals) -> the node's coefficients gx = lam_ux / M, gy, and
// gM = -(lam_ux ux + lam_uy uy) / M, M = sum of the corner m; wall
// components carry no adjoint (they are set, not solved).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adj_node(Dev d, AdjStage a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int nc = d.nc, np = d.np;
    double lx = a.LUS[p], ly = a.LUS[np + p], M = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k];
        const int ie = e >> 2, iv = e & 3;
        lx += a.CU[(0 * 4 + iv) * nc + ie];
        ly += a.CU[(1 * 4 + iv) * nc + ie];
        M += a.rec[(0 * 8 + 2 * iv) * nc + ie] + a.rec[(0 * 8 + 2 * iv + 1) * nc + ie];
    }
    const uint8_t fl = d.nflag[p];
    if (fl & 1) ly = 0.0;
    if (fl & 2) lx = 0.0;
    const double ux = a.usx[p], uy = a.usy[p];
    a.NG[p] = lx / M;
    a.NG[np + p] = ly / M;
    a.NG[2 * np + p] = -(lx * ux + ly * uy) / M;
}

// Corner records reversed (getriem_quad_matrixsym with the limiter and the
// impedance's |du.n| frozen): lam of the records (the cell RHS's, plus the
// nodal solve's through the node coefficients) -> LUI +=, CX += (corner
// coordinates).
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adj_corner(Dev d, AdjStage a)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc, np = d.np;
    if (ie >= nc) return;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    CellIn c;
    load_cell(d, a, ie, c);
    Edges g;
    edges(c, g);
    double bq[4][3], un[4][4];
    for (int iv = 0; iv < 4; ++iv) {
        bq[iv][0] = 1.0;
        bq[iv][1] = (xvq_(iv) - c.rc) / 1.0;
        bq[iv][2] = (yvq_(iv) - c.sc) / 1.0;
        for (int q = 0; q < 4; ++q) {
            un[iv][q] = 0.0;
            for (int id = 0; id < 3; ++id) un[iv][q] += c.U[q][id] * bq[iv][id];
        }
    }
    const double tau = c.U[0][0], rhoct = 1.0 / tau;
    const double uctr = c.U[1][0], vctr = c.U[2][0], ectr = c.U[3][0];
    const double pin = g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr));
    const bool pcl = !(eps > pin);
    const double pctr = pcl ? pin : eps;
    const double af0 = c.af[0], af1 = c.af[1], af2 = c.af[2], af3 = c.af[3], al3 = c.al[3];
    const double dudr = af0 * c.U[1][1] + af2 * c.U[2][1];
    const double duds = af0 * c.U[1][2] + af2 * c.U[2][2];
    const double dvdr = af1 * c.U[1][1] + af3 * c.U[2][1];
    const double dvds = af1 * c.U[1][2] + af3 * c.U[2][2];
    const double sin_ = d.gamlg * pctr / rhoct;
    const bool scl = !(eps > sin_);
    const double sdctr = sqrt(scl ? sin_ : eps);
    double LU[4][3];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) LU[q][id] = 0.0;
    double Lrhoct = 0.0, Lsdctr = 0.0, Lpctr = 0.0, Luctr = 0.0, Lvctr = 0.0;
    double Ldudr = 0.0, Lduds = 0.0, Ldvdr = 0.0, Ldvds = 0.0;
    double Lnx[4] = {0, 0, 0, 0}, Lny[4] = {0, 0, 0, 0}, LL[4] = {0, 0, 0, 0};
    for (int iv = 0; iv < 4; ++iv) {
        const int p = c.ip[iv];
        const double rhovt = 1.0 / un[iv][0];
        const double ekv = un[iv][3] - 0.5 * (un[iv][1] * un[iv][1] + un[iv][2] * un[iv][2]);
        const double pvin = g1 * rhovt * ekv;
        const bool pvcl = !(eps > pvin);
        const double pv = pvcl ? pvin : eps;
        const double uvx = uctr + dudr * bq[iv][1] + duds * bq[iv][2];
        const double vvx = vctr + dvdr * bq[iv][1] + dvds * bq[iv][2];
        const double pvx = pctr + al3 * (pv - pctr);
        const double gx = a.NG[p], gy = a.NG[np + p], gM = a.NG[2 * np + p];
        double Luvx = 0.0, Lvvx = 0.0, Lpvx = 0.0;
        for (int f = 0; f < 2; ++f) {
            const int e = f == 0 ? ((iv + 3) & 3) : iv;
            const int cf = f + 2 * iv;
            const double nx = g.nx[e], ny = g.ny[e], area = (2.0 * 1.0 + 1.0) / 3.0 * 0.5 * g.L[e];
            // the impedance's |du.n| is frozen with the limiter
            const double deltu = a.dlt[cf * nc + ie];
            const double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
            const double m11 = mur * area;
            const double Lm = a.LREC[(0 * 8 + cf) * nc + ie] + gM;
            const double Lmu1 = a.LREC[(1 * 8 + cf) * nc + ie] + gx;
            const double Lmu2 = a.LREC[(2 * 8 + cf) * nc + ie] + gy;
            const double Lsn1 = a.LREC[(3 * 8 + cf) * nc + ie] - gx;
            const double Lsn2 = a.LREC[(4 * 8 + cf) * nc + ie] - gy;
            const double Lm11 = Lm + Lmu1 * uvx + Lmu2 * vvx;
            Luvx += Lmu1 * m11;
            Lvvx += Lmu2 * m11;
            Lpvx -= Lsn1 * area * nx + Lsn2 * area * ny;
            const double Larea = -(Lsn1 * pvx * nx + Lsn2 * pvx * ny) + Lm11 * mur;
            Lnx[e] -= Lsn1 * pvx * area;
            Lny[e] -= Lsn2 * pvx * area;
            const double Lmur = Lm11 * area;
            Lrhoct += Lmur * (sdctr + d.cimpd * d.slpdu * deltu);
            Lsdctr += Lmur * rhoct;
            LL[e] += 0.5 * Larea;
        }
        Lpctr += (1.0 - al3) * Lpvx;
        const double Lpv = al3 * Lpvx;
        double Lun[4] = {0.0, 0.0, 0.0, 0.0};
        if (pvcl) {
            const double Lrhovt = Lpv * g1 * ekv;
            Lun[3] = Lpv * g1 * rhovt;
            Lun[1] = -Lpv * g1 * rhovt * un[iv][1];
            Lun[2] = -Lpv * g1 * rhovt * un[iv][2];
            Lun[0] = -Lrhovt * rhovt * rhovt;
        }
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < 3; ++id) LU[q][id] += Lun[q] * bq[iv][id];
        Luctr += Luvx;
        Ldudr += Luvx * bq[iv][1];
        Lduds += Luvx * bq[iv][2];
        Lvctr += Lvvx;
        Ldvdr += Lvvx * bq[iv][1];
        Ldvds += Lvvx * bq[iv][2];
    }
    if (scl) {
        const double Lsin = Lsdctr / (2.0 * sdctr);
        Lpctr += Lsin * d.gamlg / rhoct;
        Lrhoct -= Lsin * d.gamlg * pctr / (rhoct * rhoct);
    }
    if (pcl) {
        Lrhoct += Lpctr * g1 * (ectr - 0.5 * (uctr * uctr + vctr * vctr));
        LU[3][0] += Lpctr * g1 * rhoct;
        Luctr -= Lpctr * g1 * rhoct * uctr;
        Lvctr -= Lpctr * g1 * rhoct * vctr;
    }
    LU[0][0] -= Lrhoct / (tau * tau);
    LU[1][0] += Luctr;
    LU[2][0] += Lvctr;
    LU[1][1] += af0 * Ldudr + af1 * Ldvdr;
    LU[2][1] += af2 * Ldudr + af3 * Ldvdr;
    LU[1][2] += af0 * Lduds + af1 * Ldvds;
    LU[2][2] += af2 * Lduds + af3 * Ldvds;
    double Lx[4] = {0, 0, 0, 0}, Ly[4] = {0, 0, 0, 0};
    const double zero[4] = {0, 0, 0, 0};
    edges_rev(g, zero, zero, LL, Lnx, Lny, Lx, Ly);
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 3; ++id) {
            const int o = (q * 3 + id) * nc + ie;
            a.LUI[o] = a.LUI[o] + LU[q][id];
        }
    for (int k = 0; k < 4; ++k) {
        a.CX[(0 * 4 + k) * nc + ie] += Lx[k];
        a.CX[(1 * 4 + k) * nc + ie] += Ly[k];
    }
}

// Node sums of the corner adjoints: lam_x_in = direct part + sum CX, the
// previous stage's lam_x_out (LX).  The previous nodal velocity enters the
// corner records only through the impedance's |du.n|, which is frozen, so
// the previous stage's lam_u*_next (LUSN) is zero.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adj_gather(Dev d, AdjStage a)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    const int nc = d.nc, np = d.np;
    double lx = a.LXI[p], ly = a.LXI[np + p];
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k];
        const int ie = e >> 2, iv = e & 3;
        lx += a.CX[(0 * 4 + iv) * nc + ie];
        ly += a.CX[(1 * 4 + iv) * nc + ie];
    }
    a.LX[p] = lx;
    a.LX[np + p] = ly;
    a.LUSN[p] = 0.0;
    a.LUSN[np + p] = 0.0;
}

// End of a reversed step: lam_U^n = LUN + lam_U_in(stage 0) (into LU),
// lam_x^n = LXN + lam_x_in(stage 0) (into LX); LUN, LXN cleared.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_adj_step_end(Dev d, AdjStage a)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int nu = 12 * d.nc, nx = 2 * d.np;
    if (t < nu) {
        a.LU[t] = a.LUN[t] + a.LUI[t];
        a.LUN[t] = 0.0;
    }
    if (t < nx) {
        a.LX[t] = a.LXN[t] + a.LX[t];
        a.LXN[t] = 0.0;
    }
}

// The time step of a replayed (frozen) step: dt from the recording.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_replay_dt(Dev d, const double *dt)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        d.scal[0] = *dt;
        d.scal[1] = d.scal[1] + *dt;
    }
}

}  // namespace lag2d
