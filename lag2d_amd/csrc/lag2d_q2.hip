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
#include <utility>

#include "lag2d_dev.h"
#include "lag2d_host.h"
#include "lag2d_kernels.h"
#include "lag2d_q2geom.h"

namespace lag2d {

#define UD(buf, q, id, ie) (buf)[((q) * 6 + (id)) * nc + (ie)]
// the same for kernels templated on the number of DOFs per variable ND
#define UDN(buf, q, id, ie) (buf)[((q) * ND + (id)) * nc + (ie)]

// The limited DOFs of cell ie: WENO's output where it limited the cell (the
// troubled flag it leaves set), the stage input everywhere else.
__device__ __forceinline__ const double *cell_dofs(const Dev &d, int ie)
{
    return d.flag[ie] ? d.Ul : d.Ui;
}

// ---------------------------------------------------------------- prep --
// Evaluates the Persson indicator on tau, sets each cell's troubled flag
// and appends troubled cells to the WENO list (the stage DOFs are not
// copied: WENO writes only the troubled cells' limited DOFs to Ul, and
// cell_dofs reads the other cells' from the stage input).  Two work splits
// with the same arithmetic: k_q2_prep, 8 lanes per cell evaluating the 25
// Gauss points and lane 0 summing them in point order, for meshes too small
// to fill the chip with a thread per cell; k_q2_prep_tpc, a thread per cell
// whose unrolled point loop turns the shape functions into constants.
constexpr int PREP_CELLS = 32;
constexpr int PREP_GROUP = 5;  // points scheduled together in k_q2_prep_tpc

// The cell's data the indicator reads.
struct PrepCell {
    double u0[6], xp[9], yp[9], rc, sc, g0, g1, g2;
    __device__ __forceinline__ PrepCell(const Dev &d, const double *__restrict__ Uin, int ie)
    {
        const int nc = d.nc;
        for (int id = 0; id < 6; ++id) u0[id] = UD(Uin, 0, id, ie);
        load9(d, d.xs, d.ys, ie, xp, yp);
        rc = d.G[0 * nc + ie];
        sc = d.G[1 * nc + ie];
        g0 = d.G[18 * nc + ie];
        g1 = d.G[19 * nc + ie];
        g2 = d.G[20 * nc + ie];
    }
    __device__ __forceinline__ void opaque()
    {
        for (int k = 0; k < 9; ++k) asm volatile("" : "+v"(xp[k]), "+v"(yp[k]));
        asm volatile("" : "+v"(rc), "+v"(sc));
    }
    // the two integrands at 5x5 point g: |u|^2 and |u - u_P1|^2 times the
    // weighted Jacobian
    __device__ __forceinline__ void point(int g, double &a, double &b) const
    {
        const int i = g / 5, j = g - 5 * i;
        const double r = PG5[i], s = PG5[j], wi = WG5[i] * WG5[j];
        double J[4];
        jac9(xp, yp, r, s, J);
        const double djaco = wi * (J[0] * J[3] - J[2] * J[1]);
        double bs[6];
        basis6(r, s, rc, sc, g0, g1, g2, bs);
        double un = 0.0, u1 = 0.0;
        for (int id = 0; id < 6; ++id) un = un + u0[id] * bs[id];
        for (int id = 0; id < 3; ++id) u1 = u1 + u0[id] * bs[id];
        a = un * un * djaco;
        b = (u1 - un) * (u1 - un) * djaco;
    }
};

__device__ __forceinline__ void prep_flag(const Dev &d, int ie, double v1, double v2)
{
    const double thr = -3.0 * log10(4.0) - 4.0;
    const bool troubled = !(log10(v2 / v1) <= thr) && run_active(d, cell_run(d, ie));
    d.flag[ie] = troubled ? 1 : 0;
    if (troubled) {
        int pos = atomicAdd(d.tcount, 1);
        d.tlist[pos] = ie;
    }
}

__global__ void __launch_bounds__(256) k_q2_prep(Dev d, const double *__restrict__ Uin)
{
    __shared__ double sv[PREP_CELLS][2 * 25 + 1];
    const int nc = d.nc, ie0 = blockIdx.x * PREP_CELLS;
    const int cl = threadIdx.x >> 3, lane = threadIdx.x & 7, ie = ie0 + cl;
    const bool quad = ie < nc && ie >= d.ntria;  // the indicator runs on quads (Limiterlag.f90:12104)
    if (quad) {
        const PrepCell c(d, Uin, ie);
        for (int g = lane; g < 25; g += 8) c.point(g, sv[cl][g], sv[cl][25 + g]);
    }
    __syncthreads();
    if (!quad || lane != 0) return;
    double v1 = 0.0, v2 = 0.0;
#pragma unroll
    for (int g = 0; g < 25; ++g) {
        v2 = v2 + sv[cl][25 + g];
        v1 = v1 + sv[cl][g];
    }
    prep_flag(d, ie, v1, v2);
}

__global__ void __launch_bounds__(256) k_q2_prep_tpc(Dev d, const double *__restrict__ Uin)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    if (ie >= d.nc || ie < d.ntria) return;
    PrepCell c(d, Uin, ie);
    double v1 = 0.0, v2 = 0.0;
#pragma unroll
    for (int g = 0; g < 25; ++g) {
        double a, b;
        c.point(g, a, b);
        v2 = v2 + b;
        v1 = v1 + a;
        // the node coordinates pass through an empty asm statement after each
        // group of points, so the compiler keeps no products of them from
        // one group to the next: left alone it shares them across all 25
        // points and needs 428 registers
        if (g % PREP_GROUP == PREP_GROUP - 1) c.opaque();
    }
    prep_flag(d, ie, v1, v2);
}

// ---------------------------------------------------------------- WENO --
__device__ __forceinline__ void rotf(const double m[4], double &a, double &b)
{
    double x = a * m[0] + b * m[1];
    double y = a * m[2] + b * m[3];
    a = x;
    b = y;
}

// Matinv225, n = 5, split for a workgroup: the inverse is adjugate / det,
// each Leibniz expansion summed in the reference's term order.  The term
// tables are generated at compile time: the column taken by the last row
// ascends slowest, then the row before, ...; the first row takes the
// remaining column; the sign is the parity of the permutation.  Matrices are
// column-major; the inverse holds cof5(a, i, j) / det5(a) at (j, i).
template <int M>
struct Perms {
    static constexpr int N = (M == 5) ? 120 : 24;
    signed char p[N][M];
    bool neg[N];
};

template <int M>
constexpr Perms<M> make_perms()
{
    Perms<M> T{};
    int q[M] = {};  // q = (p[M-1], ..., p[0]) in lexicographic order
    for (int k = 0; k < M; ++k) q[k] = k;
    for (int t = 0; t < Perms<M>::N; ++t) {
        for (int k = 0; k < M; ++k) T.p[t][k] = (signed char)q[M - 1 - k];
        int inv = 0;
        for (int x = 0; x < M; ++x)
            for (int y = x + 1; y < M; ++y)
                if (T.p[t][x] > T.p[t][y]) ++inv;
        T.neg[t] = (inv & 1) != 0;
        int i = M - 2;
        while (i >= 0 && q[i] >= q[i + 1]) --i;
        if (i < 0) break;
        int j = M - 1;
        while (q[j] <= q[i]) --j;
        int tmp = q[i];
        q[i] = q[j];
        q[j] = tmp;
        for (int a = i + 1, b = M - 1; a < b; ++a, --b) {
            tmp = q[a];
            q[a] = q[b];
            q[b] = tmp;
        }
    }
    return T;
}

constexpr Perms<5> kP5 = make_perms<5>();
constexpr Perms<4> kP4 = make_perms<4>();
static_assert(kP5.p[0][4] == 0 && kP5.p[0][3] == 1 && kP5.p[0][0] == 4 && !kP5.neg[0] &&
                  kP5.p[119][4] == 4 && kP5.p[119][0] == 0,
              "Leibniz term order");

template <int T>
__device__ __forceinline__ double det5_term(const double (&m)[25])
{
    double v = m[0 + 5 * kP5.p[T][0]];
    v = v * m[1 + 5 * kP5.p[T][1]];
    v = v * m[2 + 5 * kP5.p[T][2]];
    v = v * m[3 + 5 * kP5.p[T][3]];
    v = v * m[4 + 5 * kP5.p[T][4]];
    if constexpr (kP5.neg[T]) v = -v;
    return v;
}

template <int... T>
__device__ __forceinline__ double det5_sum(const double (&m)[25], std::integer_sequence<int, T...>)
{
    double det = 0.0;
    ((det = (T == 0) ? det5_term<T>(m) : det + det5_term<T>(m)), ...);
    return det;
}

template <int T>
__device__ __forceinline__ double cof4_term(const double (&m)[16])
{
    double v = m[0 + 4 * kP4.p[T][0]];
    v = v * m[1 + 4 * kP4.p[T][1]];
    v = v * m[2 + 4 * kP4.p[T][2]];
    v = v * m[3 + 4 * kP4.p[T][3]];
    if constexpr (kP4.neg[T]) v = -v;
    return v;
}

template <int... T>
__device__ __forceinline__ double det4_sum(const double (&m)[16], std::integer_sequence<int, T...>)
{
    double det = 0.0;
    ((det = (T == 0) ? cof4_term<T>(m) : det + cof4_term<T>(m)), ...);
    return det;
}

__device__ double det5(const double *a)
{
    double m[25];
    for (int k = 0; k < 25; ++k) m[k] = a[k];
    return det5_sum(m, std::make_integer_sequence<int, 120>{});
}

// signed cofactor (i, j): Leibniz expansion of the minor without row i and
// column j
__device__ double cof5(const double *a, int i, int j)
{
    double m[16];
    for (int c = 0; c < 4; ++c)
        for (int r = 0; r < 4; ++r) m[r + 4 * c] = a[(r < i ? r : r + 1) + 5 * (c < j ? c : c + 1)];
    const double cof = det4_sum(m, std::make_integer_sequence<int, 24>{});
    return ((i + j) & 1) ? -cof : cof;
}

// Characteristic matrices of a real face (Roe-averaged state of the two
// cells, face normal in the velocity-aligned frame), in two parts so that a
// workgroup can spread them: the face's averaged quantities, then one
// matrix entry per call.  P = dnx, dny, uf, vf, rhosd, pf, gam1, vnm, vedn.
__device__ __forceinline__ void char_face(double gam, double dnx, double dny, const double ul[4], const double ur[4], double P[9])
{
    const double eps = 1.e-6;
    double rhoc1 = 1.0 / ul[0], u1 = ul[1], v1 = ul[2], e1 = ul[3];
    double rhoc2 = 1.0 / ur[0], u2 = ur[1], v2 = ur[2], e2 = ur[3];
    double rhof = sqrt(rhoc1 * rhoc2);
    double rhosq = sqrt(rhoc1) + sqrt(rhoc2);
    double uf = (sqrt(rhoc1) * u1 + sqrt(rhoc2) * u2) / rhosq;
    double vf = (sqrt(rhoc1) * v1 + sqrt(rhoc2) * v2) / rhosq;
    double h1 = dmax(eps, gam * e1 - (gam - 1.0) * 0.5 * (u1 * u1 + v1 * v1));
    double h2 = dmax(eps, gam * e2 - (gam - 1.0) * 0.5 * (u2 * u2 + v2 * v2));
    double hf = (sqrt(rhoc1) * h1 + sqrt(rhoc2) * h2) / rhosq;
    double sdf = sqrt(dmax(eps, (gam - 1.0) * (hf - 0.5 * (uf * uf + vf * vf))));
    P[0] = dnx;
    P[1] = dny;
    P[2] = uf;
    P[3] = vf;
    P[4] = rhof * sdf;                  // rhosd
    P[5] = sdf * sdf * rhof / gam;      // pf
    P[6] = (gam - 1.0) / gam;           // gam1
    P[7] = vf * dnx - uf * dny;         // vnm
    P[8] = uf * dnx + vf * dny;         // vedn
}

// Entry e of the face's matrices: e < 16 the left eigenvectors qi[e/4][e%4],
// else the right eigenvectors q[(e-16)/4][(e-16)%4].
__device__ __forceinline__ double char_entry(double gam, const double P[9], int e)
{
    const double dnx = P[0], dny = P[1], uf = P[2], vf = P[3], rhosd = P[4], pf = P[5], gam1 = P[6],
                 vnm = P[7], vedn = P[8];
    switch (e) {
    case 0: return -.5 / gam;
    case 1: return .5 * dnx / rhosd - 0.5 * uf / pf * gam1;
    case 2: return .5 * dny / rhosd - 0.5 * vf / pf * gam1;
    case 3: return .5 / pf * gam1;
    case 4: return 1.0 / gam;
    case 5: return -dny * vnm / pf - uf / pf / gam;
    case 6: return dnx * vnm / pf - vf / pf / gam;
    case 7: return 1.0 / pf / gam;
    case 8: return 0.0;
    case 9: return dny / pf;
    case 10: return -dnx / pf;
    case 11: return 0.0;
    case 12: return .5 / gam;
    case 13: return .5 * dnx / rhosd + 0.5 * uf / pf * gam1;
    case 14: return .5 * dny / rhosd + 0.5 * vf / pf * gam1;
    case 15: return -.5 / pf * gam1;
    case 16: return -1.0;
    case 17: return gam - 1.0;
    case 18: return (gam - 1.0) * vnm;
    case 19: return 1.0;
    case 20: return rhosd * dnx;
    case 21: return 0.0;
    case 22: return pf * dny;
    case 23: return rhosd * dnx;
    case 24: return rhosd * dny;
    case 25: return 0.0;
    case 26: return -pf * dnx;
    case 27: return rhosd * dny;
    case 28: return pf + rhosd * vedn;
    case 29: return pf;
    case 30: return 0.0;
    default: return -pf + rhosd * vedn;
    }
}

// Compiler fence: the LDS loads written above it are issued together and
// waited for once.  Without it the scheduler places each LDS load next to
// its use, which exposes the full LDS latency per load at one wave per SIMD.
#define LDS_FENCE() asm volatile("" ::: "memory")

// One troubled cell per workgroup: 9-cell stencil, local orthogonal Taylor
// bases, 1 + 4 Lagrange + 4 Hermite candidates, characteristic nonlinear
// weights per real face, back-projection (weno_stncl_quad_orthol3 and
// helpers).  The work of a cell is split into phases of independent tasks
// (quadrature points, matrix entries, cofactors, stencil x variable), with
// the stencil staged in LDS; every sum keeps the reference's sequential
// order, so the result does not depend on the split.  Blocks stride over
// the compacted troubled-cell list.
constexpr int WENO_WG = 256;
constexpr int WENO_GRID = 512;  // two resident blocks per CU

struct WenoShared {
    double m[4];
    double shv[25][9];                                   // shape values at the Gauss points
    double cxs[9][9], cys[9][9];                         // stage nodes of the stencil cells
    double cxr[9][9], cyr[9][9];                         // the same in the velocity-aligned frame
    double gc[9][7];                                     // rc, sc, offsets 19-21, volume, mass
    double xl[9][25], yl[9][25], rd[9][25];  // rotated Gauss points, rho * |J| * w
    double xg[9][25], yg[9][25];             // unrotated Gauss points
    double fdat[4][10];                      // per real face: edge vector, left and right states
    double fpro[4][9];                       // per real face: averaged quantities (char_face)
    double bl[9][25][5];                     // basis about the cell's own mass centre
    double bl0[9][25][5];                    // basis about the troubled cell's mass centre
    double geopl[9][5], dxcj[9], masel[9];
    double unkrl[9][4][6], unkpl[9][4][6], unksl[9][4][6];
    double bb[9][25], bbq[9][25];
    double inv[10][25], det[10];             // bbq^-1 of the 9 cells, bb^-1 of the troubled cell
    double s1[9][5], mh2[9][5][2], rh2[9][4][2];
    double qm[4][4][4], qi[4][4][4];
    union {
        double bq[9][25][5];                 // reference-coordinate basis, until bb is built
        struct {
            double rhs[9][4][5];
            double ms[8][25], mi[8][25], mdet[8];  // 0-3 Lagrange, 4-7 Hermite normal matrices
            double rs[4][4][5];
            double ws[4][9];
            double fch[4][4][6], uf[4][4][6];
            double phy[4][6], rhsuk[4][5];
        } l;
    };
    int top[16];  // stencil topology of the cell (k_q2_weno_setup)
    int ielaj[9];
    int nface;
    int face[4], fnei[4];
};

static constexpr int ST1[4][5] = {{1, 4, 5, 8, 7}, {2, 1, 6, 5, 8}, {3, 2, 7, 6, 5}, {4, 3, 8, 7, 6}};
static constexpr int ST2[4][3] = {{1, 5, 8}, {2, 5, 6}, {3, 7, 6}, {4, 7, 8}};

// Value of lane J of this lane's quad (lanes 4i..4i+3), by DPP quad_perm;
// every lane of the quad must be active.
template <int J>
__device__ __forceinline__ double quad_bcast(double v)
{
    constexpr int ctrl = J | (J << 2) | (J << 4) | (J << 6);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)b, ctrl, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xf, 0xf, false);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// Row k of the Hermite least-squares matrix of stencil is, column a
// (rows 0-2: the cell-average constraints, 3-8: the two gradient
// constraints of each neighbour).
__device__ __forceinline__ double mh_entry(const WenoShared &S, int is, int k, int a)
{
    if (k < 3) return S.s1[ST2[is][k]][a];
    const int e = (k - 3) >> 1;
    return S.mh2[ST2[is][e]][a][(k - 3) & 1];
}

__global__ void __launch_bounds__(WENO_WG) k_q2_weno(Dev d, const double *__restrict__ Uin)
{
    __shared__ WenoShared S;
    static_assert(WENO_WG == 256, "task spreading assumes 4 waves");
    const int nc = d.nc;
    const int ntr = *d.tcount;
    if (blockIdx.x >= ntr) return;
    for (int k = threadIdx.x; k < 225; k += WENO_WG) (&S.shv[0][0])[k] = d.shv[k];
    for (int t = blockIdx.x; t < ntr; t += gridDim.x) {
        // The thread index passes through an empty asm statement, so the
        // compiler cannot hoist the phases' index arithmetic out of the
        // cell loop and keep it live in registers the phases need.
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        // task index of this thread: consecutive tasks go to different waves,
        // so phases with few tasks still use every SIMD
        const int tsp = ((tid & 63) << 2) | (tid >> 6);
        __syncthreads();  // the previous cell's staging is dead
        const int ie = d.tlist[t];
        // 0a: stencil topology (time-invariant, from k_q2_weno_setup) and the
        //     velocity-aligned frame of the troubled cell
        if (tid < 16) {
            S.top[tid] = d.wtop[tid * nc + ie];
        } else if (tid == 64) {
            const double a = UD(Uin, 1, 0, ie), b = UD(Uin, 2, 0, ie);
            if (sqrt(a * a + b * b) > 1.e-10) {
                S.m[0] = a / sqrt(a * a + b * b);
                S.m[1] = b / sqrt(a * a + b * b);
                S.m[2] = -b / sqrt(a * a + b * b);
                S.m[3] = a / sqrt(a * a + b * b);
            } else {
                S.m[0] = S.m[1] = S.m[2] = S.m[3] = 0.0;
            }
        }
        __syncthreads();
        {
            const int flags = S.top[9];
            if (tid == 0 && (flags & 2)) atomicAdd(d.wenobad, 1);
            // a cell at rest has the zero frame and the reference's normal
            // equations vanish (0/0); outside strict mode it keeps its DOFs
            const bool rest = S.m[0] == 0.0 && S.m[1] == 0.0 && S.m[2] == 0.0 && S.m[3] == 0.0 && !d.weno_strict;
            if ((flags & 1) || rest) {  // not limited: the later kernels read the stage input
                if (tid == 0) d.flag[ie] = 0;
                continue;
            }
        }
        const double m[4] = {S.m[0], S.m[1], S.m[2], S.m[3]};
        int ielaj[9];
        for (int k = 0; k < 9; ++k) ielaj[k] = S.top[k];
        if (tid < 9) {
            S.ielaj[tid] = ielaj[tid];
        } else if (tid == 64) {
            const int nf = S.top[10];
            S.nface = nf;
            for (int a = 0; a < 4; ++a) {
                const int f = S.top[11 + a];
                S.face[a] = f;
                S.fnei[a] = (a < nf) ? ielaj[1 + f] : 0;
            }
        }

        // 0b: everything the cell reads from HBM, one batch per thread:
        //     rho |J| w at the stencil's Gauss points (time-invariant), stage
        //     nodes, geometry rows, DOFs, and the real faces' edge vectors and
        //     neighbouring states
        {
            const int k = tsp;
            double rdv = 0.0;
            if (k < 225) {
                const int s = k / 25, g = k - 25 * s;
                rdv = d.wrd[g * nc + ielaj[s]];
            }
            if (k < 81) {
                const int s = k / 9, q = k - 9 * s;
                const int p = d.ipq[q * nc + ielaj[s]];
                double x = d.xs[p], y = d.ys[p];
                S.cxs[s][q] = x;
                S.cys[s][q] = y;
                rotf(m, x, y);
                S.cxr[s][q] = x;
                S.cyr[s][q] = y;
            } else if (k < 81 + 63) {
                const int u = k - 81, s = u / 7, r = u - 7 * s;
                const int row[7] = {0, 1, 18, 19, 20, 2, 3};
                S.gc[s][r] = d.G[row[r] * nc + ielaj[s]];
            } else if (k < 81 + 63 + 54) {
                const int u = k - 81 - 63, s = u / 6, id = u - 6 * s;
                const int je = ielaj[s];
                double v[4];
                for (int q = 0; q < 4; ++q) v[q] = UD(Uin, q, id, je);
                rotf(m, v[1], v[2]);
                for (int q = 0; q < 4; ++q) S.unkrl[s][q][id] = v[q];
            } else if (k < 81 + 63 + 54 + 4) {
                const int a = k - 81 - 63 - 54;
                if (a < S.top[10]) {
                    const int f = S.top[11 + a], je = ielaj[1 + f];
                    const int p0 = d.ipq[f * nc + ie], p1 = d.ipq[((f + 1) & 3) * nc + ie];
                    S.fdat[a][0] = d.xs[p1] - d.xs[p0];
                    S.fdat[a][1] = d.ys[p1] - d.ys[p0];
                    for (int q = 0; q < 4; ++q) {
                        S.fdat[a][2 + q] = UD(Uin, q, 0, ie);
                        S.fdat[a][6 + q] = UD(Uin, q, 0, je);
                    }
                }
            }
            if (k < 225) S.rd[k / 25][k - 25 * (k / 25)] = rdv;
        }
        __syncthreads();

        // 1: Gauss points of the stencil (unrotated and rotated) and the
        //    reference-coordinate basis there
        for (int k = tsp; k < 225; k += WENO_WG) {
            const int s = k / 25, g = k - 25 * s, i = g / 5, j = g - 5 * i;
            double sh[9], xs[9], ys[9], xr[9], yr[9];
            for (int q = 0; q < 9; ++q) {
                sh[q] = S.shv[g][q];
                xs[q] = S.cxs[s][q];
                ys[q] = S.cys[s][q];
                xr[q] = S.cxr[s][q];
                yr[q] = S.cyr[s][q];
            }
            const double g0 = S.gc[s][0], g1 = S.gc[s][1], g2 = S.gc[s][2], g3 = S.gc[s][3], g4 = S.gc[s][4];
            LDS_FENCE();
            double a = 0.0, b = 0.0, ar = 0.0, br = 0.0;
            for (int q = 0; q < 9; ++q) {
                a = a + sh[q] * xs[q];
                b = b + sh[q] * ys[q];
                ar = ar + sh[q] * xr[q];
                br = br + sh[q] * yr[q];
            }
            S.xg[s][g] = a;
            S.yg[s][g] = b;
            S.xl[s][g] = ar;
            S.yl[s][g] = br;
            double bas[6];
            basis6(PG5[i], PG5[j], g0, g1, g2, g3, g4, bas);
            for (int c = 0; c < 5; ++c) S.bq[s][g][c] = bas[c + 1];
        }
        __syncthreads();
        // 2: physical mass centres
        if (tsp < 9) {
            const int s = tsp;
            double rd[25], xg[25], yg[25];
            for (int g = 0; g < 25; ++g) {
                rd[g] = S.rd[s][g];
                xg[g] = S.xg[s][g];
                yg[g] = S.yg[s][g];
            }
            const double masel = S.gc[s][6], vol = S.gc[s][5];
            LDS_FENCE();
            double xm = 0.0, ym = 0.0;
            for (int g = 0; g < 25; ++g) {
                xm = xm + rd[g] * xg[g];
                ym = ym + rd[g] * yg[g];
            }
            double xc = xm / masel, yc = ym / masel;
            rotf(m, xc, yc);
            S.geopl[s][0] = xc;
            S.geopl[s][1] = yc;
            S.dxcj[s] = sqrt(vol);
            S.masel[s] = masel;
        }
        __syncthreads();
        // 3: linear physical basis about each cell's own centre
        for (int k = tsp; k < 225; k += WENO_WG) {
            const int s = k / 25, g = k - 25 * s;
            S.bl[s][g][0] = (S.xl[s][g] - S.geopl[s][0]) / S.dxcj[s];
            S.bl[s][g][1] = (S.yl[s][g] - S.geopl[s][1]) / S.dxcj[s];
        }
        __syncthreads();
        // 4: second moments (waves 0-2); the real faces' averaged states for
        //    the characteristic matrices (wave 3)
        const int t3 = (tid & 63) * 3 + (tid >> 6);  // task index over waves 0-2 (tid < 192)
        if (tid < 192) {
            if (t3 < 27) {
                const int s = t3 / 3, c = t3 - 3 * s;
                double x[25], y[25], rd[25];
                for (int g = 0; g < 25; ++g) {
                    const double b2 = S.bl[s][g][0], b3 = S.bl[s][g][1];
                    rd[g] = S.rd[s][g];
                    x[g] = (c == 1) ? b3 : b2;
                    y[g] = (c == 0) ? b2 : b3;
                }
                const double masel = S.masel[s];
                LDS_FENCE();
                double v = 0.0;
                for (int g = 0; g < 25; ++g) v = v + rd[g] * x[g] * y[g];
                S.geopl[s][2 + c] = (c < 2) ? 0.5 * (v / masel) : v / masel;
            }
        } else if (tid - 192 < S.nface) {
            const int a = tid - 192;
            const double dtx = S.fdat[a][0], dty = S.fdat[a][1];
            const double dlgt = sqrt(dtx * dtx + dty * dty);
            double dnx = dty / dlgt, dny = -dtx / dlgt;
            rotf(m, dnx, dny);
            double ul[4], ur[4], P[9];
            for (int q = 0; q < 4; ++q) {
                ul[q] = S.fdat[a][2 + q];
                ur[q] = S.fdat[a][6 + q];
            }
            rotf(m, ul[1], ul[2]);
            rotf(m, ur[1], ur[2]);
            char_face(d.gamlg, dnx, dny, ul, ur, P);
            for (int k = 0; k < 9; ++k) S.fpro[a][k] = P[k];
        }
        __syncthreads();
        // 5: quadratic terms; the basis about the troubled cell's centre
        for (int k = tsp; k < 225; k += WENO_WG) {
            const int s = k / 25, g = k - 25 * s;
            const double b1 = S.bl[s][g][0], b2 = S.bl[s][g][1];
            S.bl[s][g][2] = 0.5 * b1 * b1 - S.geopl[s][2];
            S.bl[s][g][3] = 0.5 * b2 * b2 - S.geopl[s][3];
            S.bl[s][g][4] = b1 * b2 - S.geopl[s][4];
            const double c1 = (S.xl[s][g] - S.geopl[0][0]) / S.dxcj[0];
            const double c2 = (S.yl[s][g] - S.geopl[0][1]) / S.dxcj[0];
            S.bl0[s][g][0] = c1;
            S.bl0[s][g][1] = c2;
            S.bl0[s][g][2] = 0.5 * c1 * c1 - S.geopl[0][2];
            S.bl0[s][g][3] = 0.5 * c2 * c2 - S.geopl[0][3];
            S.bl0[s][g][4] = c1 * c2 - S.geopl[0][4];
        }
        __syncthreads();
        // 6: projection matrices bb (reference -> physical basis) and bbq
        //    (physical mass matrix), entry e of both per task; then the
        //    stencil-matrix moments of the neighbours (quotients by the
        //    cell mass through its reciprocal, bit-identical to division)
        if (tsp < 225) {
            const int s = tsp / 25, e = tsp - 25 * s, b = e / 5, a = e - 5 * b;
            double v = 0.0, w = 0.0;
#pragma unroll
            for (int g0 = 0; g0 < 25; g0 += 13) {
                constexpr int NB = 13;
                double rd[NB], lb[NB], qa[NB], la[NB];
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    const int g = (g0 + u < 25) ? g0 + u : 24;
                    rd[u] = S.rd[s][g];
                    lb[u] = S.bl[s][g][b];
                    qa[u] = S.bq[s][g][a];
                    la[u] = S.bl[s][g][a];
                }
                LDS_FENCE();
#pragma unroll
                for (int u = 0; u < NB; ++u) {
                    if (g0 + u < 25) {
                        v = v + rd[u] * qa[u] * lb[u];
                        w = w + rd[u] * la[u] * lb[u];
                    }
                }
            }
            S.bb[s][e] = v;
            S.bbq[s][e] = w;
        }
        for (int k = tsp; k < 40 + 80; k += WENO_WG) {
            // s1: rd * bl0[c] / masel; mh2: rd * bl0[a] * bl[c] / masel
            int s, a, c;
            bool lin;
            if (k < 40) {
                s = 1 + k / 5;
                a = k - 5 * (k / 5);
                c = 0;
                lin = true;
            } else {
                const int u = k - 40, r = u - 10 * (u / 10);
                s = 1 + u / 10;
                a = r >> 1;
                c = r & 1;
                lin = false;
            }
            const Recip rm = recip(S.masel[s]);
            bool ok = rm.ok;
            double rd[25], b0[25], bc[25];
#pragma unroll
            for (int g = 0; g < 25; ++g) {
                rd[g] = S.rd[s][g];
                b0[g] = S.bl0[s][g][a];
                bc[g] = S.bl[s][g][c];
            }
            LDS_FENCE();
            double v = 0.0;
#pragma unroll
            for (int g = 0; g < 25; ++g) {
                double x = rd[g] * b0[g];
                if (!lin) x = x * bc[g];
                v = v + qdiv_term(x, rm, ok);
            }
            if (!ok) {
                v = 0.0;
                for (int g = 0; g < 25; ++g) {
                    double x = S.rd[s][g] * S.bl0[s][g][a];
                    if (!lin) x = x * S.bl[s][g][c];
                    v = v + x / S.masel[s];
                }
            }
            if (lin) S.s1[s][a] = v;
            else S.mh2[s][a][c] = v;
        }
        __syncthreads();
        // 7: inverses of bbq (9 cells) and bb (troubled cell): determinants
        //    and cofactors; rhs of the projections; stencil matrices.  The
        //    ten 120-term determinants run on wave 0 with the copies, the
        //    cofactors and sums on waves 1-3, so no wave executes both.
        if (tid < 64) {
            if (tid < 10) S.det[tid] = det5(tid < 9 ? S.bbq[tid] : S.bb[0]);
            for (int k = tid - 10; k >= 0 && k < 100; k += 54) {
                const int is = k / 25, e = k - 25 * is, kk = e / 5, r = e - 5 * kk;
                S.l.ms[is][r + 5 * kk] = S.s1[ST1[is][r]][kk];
            }
        } else {
            const int w = ((tid & 63) * 3) + (tid >> 6) - 1;  // 0..191 spread over waves 1-3
            for (int k = w; k < 250 + 180 + 100; k += 192) {
                if (k < 250) {
                    const int mi = k / 25, e = k - 25 * mi, i = e / 5, j = e - 5 * i;
                    S.inv[mi][j + 5 * i] = cof5(mi < 9 ? S.bbq[mi] : S.bb[0], i, j);
                } else if (k < 430) {
                    const int u = k - 250, s = u / 20, q = (u / 5) & 3, a = u - 5 * (u / 5);
                    double v = 0.0;
                    for (int b = 0; b < 5; ++b) v = v + S.unkrl[s][q][b + 1] * S.bb[s][b + 5 * a];
                    S.l.rhs[s][q][a] = v;
                } else {
                    const int u = k - 430, is = u / 25, e = u - 25 * is, b = e / 5, a = e - 5 * b;
                    double v = 0.0;
                    for (int kk = 0; kk < 9; ++kk) v = v + mh_entry(S, is, kk, a) * mh_entry(S, is, kk, b);
                    S.l.ms[4 + is][a + 5 * b] = v;
                }
            }
        }
        __syncthreads();
        // 8: divide by the determinants; determinants and cofactors of the
        //    stencil matrices (determinants on wave 0, the rest on waves 1-3)
        if (tid < 64) {
            if (tid < 8) S.l.mdet[tid] = det5(S.l.ms[tid]);
        } else {
            const int w = ((tid & 63) * 3) + (tid >> 6) - 1;
            for (int k = w; k < 250 + 200; k += 192) {
                if (k < 250) {
                    const int mi = k / 25, e = k - 25 * mi;
                    S.inv[mi][e] = S.inv[mi][e] / S.det[mi];
                } else {
                    const int u = k - 250, mi = u / 25, e = u - 25 * mi, i = e / 5, j = e - 5 * i;
                    S.l.mi[mi][j + 5 * i] = cof5(S.l.ms[mi], i, j);
                }
            }
        }
        __syncthreads();
        // 9: stencil DOFs in the physical bases; divide the stencil inverses
        for (int k = tsp; k < 180 + 36 + 200; k += WENO_WG) {
            if (k < 180) {
                const int s = k / 20, q = (k / 5) & 3, a = k - 5 * (k / 5);
                double v = 0.0;
                for (int b = 0; b < 5; ++b) v = v + S.l.rhs[s][q][b] * S.inv[s][b + 5 * a];
                S.unkpl[s][q][a + 1] = v;
            } else if (k < 216) {
                const int u = k - 180, s = u >> 2, q = u & 3;
                S.unkpl[s][q][0] = S.unkrl[s][q][0];
            } else {
                const int u = k - 216, mi = u / 25, e = u - 25 * mi;
                S.l.mi[mi][e] = S.l.mi[mi][e] / S.l.mdet[mi];
            }
        }
        __syncthreads();
        // 10: Hermite gradient moments rh2 (8 cells x 4 variables x 2
        //     directions; 125 terms each, in the order g, a): lane j of a
        //     quad computes the terms of the Gauss points g = j, j+4, ...,
        //     then every lane of the quad sums all 125 in order, taking them
        //     from the quad by DPP broadcast.
        {
            const int task = tid >> 2, j = tid & 3;
            const int s = 1 + (task >> 3), q = (task >> 1) & 3, c = task & 1;
            const Recip rm = recip(S.masel[s]);
            bool ok = rm.ok;
            double u[5];
            for (int a = 0; a < 5; ++a) u[a] = S.unkpl[s][q][a + 1];
            double t[7][5];
#pragma unroll
            for (int k = 0; k < 7; ++k) {
                const int g = (4 * k + j < 25) ? 4 * k + j : 24;
                double bl[5];
                for (int a = 0; a < 5; ++a) bl[a] = S.bl[s][g][a];
                const double rd = S.rd[s][g], bc = S.bl[s][g][c];
                LDS_FENCE();
#pragma unroll
                for (int a = 0; a < 5; ++a) t[k][a] = qdiv_term(rd * bl[a] * bc * u[a], rm, ok);
            }
            if (!ok) {
                for (int k = 0; k < 7; ++k) {
                    const int g = (4 * k + j < 25) ? 4 * k + j : 24;
                    for (int a = 0; a < 5; ++a)
                        t[k][a] = S.rd[s][g] * S.bl[s][g][a] * S.bl[s][g][c] * S.unkpl[s][q][a + 1] / S.masel[s];
                }
            }
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < 6; ++k) {
#pragma unroll
                for (int a = 0; a < 5; ++a) v = v + quad_bcast<0>(t[k][a]);
#pragma unroll
                for (int a = 0; a < 5; ++a) v = v + quad_bcast<1>(t[k][a]);
#pragma unroll
                for (int a = 0; a < 5; ++a) v = v + quad_bcast<2>(t[k][a]);
#pragma unroll
                for (int a = 0; a < 5; ++a) v = v + quad_bcast<3>(t[k][a]);
            }
#pragma unroll
            for (int a = 0; a < 5; ++a) v = v + quad_bcast<0>(t[6][a]);
            if (j == 0) S.rh2[s][q][c] = v;
        }
        //     Lagrange candidates; candidate averages
        for (int k = tsp; k < 80 + 36 + 20; k += WENO_WG) {
            if (k < 80) {
                const int is = k / 20, q = (k / 5) & 3, a = k - 5 * (k / 5);
                double v = 0.0;
                for (int e = 0; e < 5; ++e)
                    v = v + S.l.mi[is][a + 5 * e] * (S.unkpl[ST1[is][e]][q][0] - S.unkpl[0][q][0]);
                S.unksl[1 + is][q][a + 1] = v;
            } else if (k < 116) {
                const int u = k - 80, s = u >> 2, q = u & 3;
                S.unksl[s][q][0] = S.unkpl[0][q][0];
            } else {
                const int u = k - 116, q = u / 5, a = u - 5 * q;
                S.unksl[0][q][a + 1] = S.unkpl[0][q][a + 1];
            }
        }
        __syncthreads();
        // 11: Hermite normal-equation right-hand sides (waves 0-2); the left
        //     characteristic matrices of the real faces (wave 3)
        if (tid >= 192) {
            const int a = tid - 192;
            if (a < S.nface) {
                double P[9];
                for (int k = 0; k < 9; ++k) P[k] = S.fpro[a][k];
#pragma unroll
                for (int e = 0; e < 16; ++e) S.qi[a][e >> 2][e & 3] = char_entry(d.gamlg, P, e);
            }
        } else if (t3 < 80) {
            const int is = t3 / 20, q = (t3 / 5) & 3, a = t3 - 5 * (t3 / 5);
            double v = 0.0;
            for (int kk = 0; kk < 9; ++kk) {
                double rh;
                if (kk < 3) {
                    rh = S.unkpl[ST2[is][kk]][q][0] - S.unkpl[0][q][0];
                } else {
                    const int e = (kk - 3) >> 1;
                    rh = S.rh2[ST2[is][e]][q][(kk - 3) & 1];
                }
                v = v + mh_entry(S, is, kk, a) * rh;
            }
            S.l.rs[is][q][a] = v;
        }
        __syncthreads();
        // 12: Hermite candidates (waves 0-2); the right characteristic
        //     matrices (wave 3)
        if (tid >= 192) {
            const int a = tid - 192;
            if (a < S.nface) {
                double P[9];
                for (int k = 0; k < 9; ++k) P[k] = S.fpro[a][k];
#pragma unroll
                for (int e = 16; e < 32; ++e) S.qm[a][(e - 16) >> 2][e & 3] = char_entry(d.gamlg, P, e);
            }
        } else if (t3 < 80) {
            const int is = t3 / 20, q = (t3 / 5) & 3, a = t3 - 5 * (t3 / 5);
            double v = 0.0;
            for (int e = 0; e < 5; ++e) v = v + S.l.mi[4 + is][a + 5 * e] * S.l.rs[is][q][e];
            S.unksl[5 + is][q][a + 1] = v;
        }
        __syncthreads();
        // 13: smoothness indicator of each candidate per real face, then the
        //     nonlinear weights
        const int nst = 9;
        for (int k = tsp; k < S.nface * nst; k += WENO_WG) {
            const int a = k / nst, s = k - nst * a;
            const double dxc = S.dxcj[0], dyc = dxc;
            double o[4];
            for (int kk = 0; kk < 4; ++kk) {
                double c1 = 0.0, c2 = 0.0;
                for (int q = 0; q < 4; ++q) {
                    c1 = c1 + S.qi[a][kk][q] * S.unksl[s][q][1];
                    c2 = c2 + S.qi[a][kk][q] * S.unksl[s][q][2];
                }
                o[kk] = (c1 / dxc) * (c1 / dxc) + (c2 / dyc) * (c2 / dyc);
            }
            S.l.ws[a][s] = dmax(dmax(dmax(o[0], o[1]), o[2]), o[3]);
        }
        __syncthreads();
        if (tsp < S.nface) {
            const int a = tsp;
            double wl[9], wh[9], wt = 0.0;
            wl[0] = .2;
            for (int s = 1; s < nst; ++s) wl[s] = 0.8 / ((double)nst - 1.0);
            for (int s = 0; s < nst; ++s) {
                const double tt = 1.e-6 + S.l.ws[a][s];
                wh[s] = wl[s] / (tt * tt);
            }
            for (int s = 0; s < nst; ++s) wt = wt + wh[s];
            for (int s = 0; s < nst; ++s) S.l.ws[a][s] = wh[s] / wt;
        }
        __syncthreads();
        // 14: weighted characteristic polynomial per real face
        for (int k = tsp; k < S.nface * 24; k += WENO_WG) {
            const int a = k / 24, r = k - 24 * a, kk = r / 6, id = r - 6 * kk;
            double v = 0.0;
            for (int s = 0; s < nst; ++s) {
                double c = 0.0;
                for (int q = 0; q < 4; ++q) c = c + S.qi[a][kk][q] * S.unksl[s][q][id];
                v = v + S.l.ws[a][s] * c;
            }
            S.l.fch[a][kk][id] = v;
        }
        __syncthreads();
        // 15: back to conserved variables
        for (int k = tsp; k < S.nface * 20; k += WENO_WG) {
            const int a = k / 20, r = k - 20 * a, kk = r / 5, id = 1 + r - 5 * (r / 5);
            double v = 0.0;
            for (int q = 0; q < 4; ++q) v = v + S.qm[a][kk][q] * S.l.fch[a][q][id];
            S.l.uf[a][kk][id] = v;
        }
        __syncthreads();
        // 16: average over the real faces
        if (tsp < 20) {
            const int q = tsp / 5, id = 1 + tsp - 5 * (tsp / 5);
            const double wf = 1.0 / (double)S.nface;
            double v = 0.0;
            for (int a = 0; a < S.nface; ++a) v = v + wf * S.l.uf[a][q][id];
            S.l.phy[q][id] = v;
        }
        __syncthreads();
        // 17: project back onto the reference-coordinate basis
        if (tsp < 20) {
            const int q = tsp / 5, a = tsp - 5 * (tsp / 5);
            double v = 0.0;
            for (int b = 0; b < 5; ++b) v = v + S.l.phy[q][b + 1] * S.bbq[0][b + 5 * a];
            S.l.rhsuk[q][a] = v;
        }
        __syncthreads();
        if (tsp < 5) {
            const int a = tsp;
            double out[4];
            for (int q = 0; q < 4; ++q) {
                double v = 0.0;
                for (int b = 0; b < 5; ++b) v = v + S.l.rhsuk[q][b] * S.inv[9][b + 5 * a];
                out[q] = v;
            }
            const double i11 = m[3], i12 = -m[1], i21 = -m[2], i22 = m[0];
            const double x = out[1] * i11 + out[2] * i12;
            const double y = out[1] * i21 + out[2] * i22;
            out[1] = x;
            out[2] = y;
            for (int q = 0; q < 4; ++q) UD(d.Ul, q, a + 1, ie) = out[q];
            if (a == 0)
                for (int q = 0; q < 4; ++q) UD(d.Ul, q, 0, ie) = UD(Uin, q, 0, ie);
        }
    }
}

// ------------------------------------------------------------- corners --
// Face geometry (stored for the face integral and the node solve), sub-cell
// density and mean specific volume, and the 16 sub-cell corner states.
// Corner states: rec[RC2(ie, ivsg + 4*isg, r)], r = 0-1: m = mu*A of the
// corner's two faces, 2-3: u_c, 4: p_c; the node solve and the face
// integral rebuild the corner-face records m u_c and A sigma n from them
// and the face geometry (corner_face_rec).
// 16 lanes per cell: first lane k evaluates face slot k and a quarter of
// the 25 quadrature points of sub-cell k/4 (terms staged in LDS), then lane
// (isg, ivsg) sums its sub-cell's terms in point order and writes its
// sub-cell corner's state.
constexpr int CORNER_CPB = 16;

__global__ void __launch_bounds__(256) k_q2_corner(Dev d)
{
    __shared__ double sgq[CORNER_CPB][16 * 3 + 1];  // padded: cells of a wave in different banks
    // node coordinates of the cells and of their sub-cells (sub-cell isg at
    // 9 isg), shared by the lanes that read them; strides keep the 4 cells x
    // 4 sub-cells a wave reads at once in distinct banks
    __shared__ double sxp[CORNER_CPB][9], syp[CORNER_CPB][9];
    __shared__ double ssx[CORNER_CPB][36], ssy[CORNER_CPB][36];
    const int cl = threadIdx.x >> 4, lane = threadIdx.x & 15;
    const int ie = blockIdx.x * CORNER_CPB + cl;
    const int nc = d.nc;
    const bool active = ie < nc;
    const int isg = lane >> 2, part = lane & 3;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    if (active && lane < 9) {
        const int p = d.ipq[lane * nc + ie];
        sxp[cl][lane] = d.xs[p];
        syp[cl][lane] = d.ys[p];
    }
    __syncthreads();
    const double *xp = sxp[cl], *yp = syp[cl];
    double *xs = &ssx[cl][9 * isg], *ys = &ssy[cl][9 * isg];
    double rc = 0.0, sc = 0.0, gg0 = 0.0, gg1 = 0.0, gg2 = 0.0;
    double a = 0.0, masel = 0.0, volel = 0.0;  // sub-cell area, mass and volume
    double u0[6];
    if (active) {
        rc = d.G[0 * nc + ie];
        sc = d.G[1 * nc + ie];
        gg0 = d.G[18 * nc + ie];
        gg1 = d.G[19 * nc + ie];
        gg2 = d.G[20 * nc + ie];
        // ---- face slot `lane`: 8-11 interior half-edges, others boundary points
        double gx, gy, ga;
        if (lane >= 8 && lane < 12) {
            const int k = lane - 8;
            const double pq[4][2] = {{0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
            double xv, yv;
            map9(xp, yp, pq[k][0], pq[k][1], &xv, &yv);
            const double fx[3] = {xp[4 + k], xp[8], xv};
            const double fy[3] = {yp[4 + k], yp[8], yv};
            const double r = -1.0;
            const double dd[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
            double dxdr = 0.0, dydr = 0.0;
            for (int i = 0; i < 3; ++i) {
                dxdr = dxdr + dd[i] * fx[i];
                dydr = dydr + dd[i] * fy[i];
            }
            const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
            const double anx = -dydr, any = dxdr;
            gx = anx / sqrt(anx * anx + any * any);
            gy = any / sqrt(anx * anx + any * any);
            ga = djaco * 2.0;
        } else {
            const int ig = lane < 8 ? lane : lane - 4;
            const double pi[12][2] = {{-1, -1}, {1, -1}, {1, -1}, {1, 1}, {1, 1}, {-1, 1},
                                      {-1, 1},  {-1, -1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
            const double vn[12][2] = {{0, -1}, {0, -1}, {1, 0}, {1, 0}, {0, 1}, {0, 1},
                                      {-1, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
            double J[4];
            jac9(xp, yp, pi[ig][0], pi[ig][1], J);
            const double c11 = J[3], c12 = -J[2], c21 = -J[1], c22 = J[0];
            const double anx = c11 * vn[ig][0] + c12 * vn[ig][1];
            const double any = c21 * vn[ig][0] + c22 * vn[ig][1];
            const double w1 = anx * 2.0, w2 = any * 2.0;
            const double farea = sqrt(w1 * w1 + w2 * w2);
            gx = anx * 2.0 / farea;
            gy = any * 2.0 / farea;
            ga = farea;
        }
        sgq[cl][3 * lane + 0] = gx;
        sgq[cl][3 * lane + 1] = gy;
        sgq[cl][3 * lane + 2] = ga;
        d.gq[(3 * lane + 0) * nc + ie] = gx;
        d.gq[(3 * lane + 1) * nc + ie] = gy;
        d.gq[(3 * lane + 2) * nc + ie] = ga;
        // ---- quadrature terms of sub-cell isg: area of its Q2 map and the
        //      mass-weighted mean specific volume of the DG polynomial
        for (int id = 0; id < 6; ++id) u0[id] = UD(cell_dofs(d, ie), 0, id, ie);
        // nodes of sub-cell isg: lane `part` stores corner `part` and maps
        // mid-node `part`; the sub-cell's four lanes form a DPP quad, which
        // shares the mid-nodes for the centre node
        const double pc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                                  {-1.0, .5},   {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
        const int ipsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
        const int c = ipsc[isg][part];
        double xm, ym;
        map9(xp, yp, pc[c][0], pc[c][1], &xm, &ym);
        xs[part] = xp[IPQSG[isg][part]];
        ys[part] = yp[IPQSG[isg][part]];
        xs[4 + part] = xm;
        ys[4 + part] = ym;
        const double mx0 = quad_bcast<0>(xm), mx1 = quad_bcast<1>(xm), mx2 = quad_bcast<2>(xm),
                     mx3 = quad_bcast<3>(xm);
        const double my0 = quad_bcast<0>(ym), my1 = quad_bcast<1>(ym), my2 = quad_bcast<2>(ym),
                     my3 = quad_bcast<3>(ym);
        if (part == 0) {
            xs[8] = -0.25 * (xp[IPQSG[isg][0]] + xp[IPQSG[isg][1]] + xp[IPQSG[isg][2]] + xp[IPQSG[isg][3]]) +
                    0.5 * (mx0 + mx1 + mx2 + mx3);
            ys[8] = -0.25 * (yp[IPQSG[isg][0]] + yp[IPQSG[isg][1]] + yp[IPQSG[isg][2]] + yp[IPQSG[isg][3]]) +
                    0.5 * (my0 + my1 + my2 + my3);
        }
    }
    __syncthreads();  // sub-cell nodes staged, face slots shared
    if (!active) return;
    {
        const double sr[4] = {-1.0, 1.0, 1.0, -1.0}, ss[4] = {-1.0, -1.0, 1.0, 1.0};
        const double rhomc = u0[0];
        // terms of Gauss point g: the sub-cell map's area element, and the
        // parent map's volume element with the mass element 1/tau |J| w
        auto terms = [&](int g, double &ta, double &tm, double &tv) {
            const int i = g / 5, j = g - 5 * i;
            // the node offsets pass through an empty asm statement, so the
            // compiler reads the nodes from LDS here rather than holding
            // them in registers across the loop
            int oc = cl, os = 9 * isg;
            asm volatile("" : "+v"(oc), "+v"(os));
            const double *xs = &ssx[oc][os], *ys = &ssy[oc][os];
            const double *xp = sxp[oc], *yp = syp[oc];
            double J[4];
            jac9(xs, ys, PG5[i], PG5[j], J);
            ta = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
            const double r = (PG5[i] + sr[isg]) / 2.0;
            const double s = (PG5[j] + ss[isg]) / 2.0;
            double b[6];
            basis6(r, s, rc, sc, gg0, gg1, gg2, b);
            double t = 0.0;
            for (int id = 0; id < 6; ++id) t = t + u0[id] * b[id];
            t = rhomc + 1.0 * (t - rhomc);
            jac9(xp, yp, r, s, J);
            const double djaco = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            tm = 1.0 / t * djaco;
            tv = djaco;
        };
        // Lane `part` evaluates point 4k + part in round k (the last round
        // has point 24 only); the quad adds each round's terms in point
        // order, so every lane of the sub-cell ends with the sequential sums
        // over points 0..24.
#pragma unroll 1
        for (int k = 0; k < 7; ++k) {
            double ta, tm, tv;
            terms(k < 6 ? 4 * k + part : 24, ta, tm, tv);
            a = a + quad_bcast<0>(ta);
            masel = masel + quad_bcast<0>(tm);
            volel = volel + quad_bcast<0>(tv);
            if (k < 6) {
                a = a + quad_bcast<1>(ta);
                masel = masel + quad_bcast<1>(tm);
                volel = volel + quad_bcast<1>(tv);
                a = a + quad_bcast<2>(ta);
                masel = masel + quad_bcast<2>(tm);
                volel = volel + quad_bcast<2>(tv);
                a = a + quad_bcast<3>(ta);
                masel = masel + quad_bcast<3>(tm);
                volel = volel + quad_bcast<3>(tv);
            }
        }
    }
    // ---- sub-cell corner (isg, iv)
    const int iv = part;
    const double rsub = 0.0 + (1.0 / a) * d.msub[isg * nc + ie];
    const double uq = volel / masel;
    double u[4][6];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < 6; ++id) u[q][id] = UD(cell_dofs(d, ie), q, id, ie);
    const int ln = IPQSG[isg][iv];
    double un[4];
    {
        double b[6];
        basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, b);
        for (int q = 0; q < 4; ++q) un[q] = 0.0;
        for (int id = 0; id < 6; ++id)
            for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
    }
    const double rhomc = u[0][0], uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
    const double rhoct = 1.0 / rhomc;
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    const double sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    const double dcor = d.cdrho * (rsub - 1.0 / uq);
    const int p = d.ipq[ln * nc + ie];
    const double rhovt = 1.0 / un[0];
    const double rhovsg = rhovt + dcor;
    const double uv = un[1], vv = un[2], ev = un[3];
    const double pv = dmax(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)));
    const double dux = d.usx[p] - uv, duy = d.usy[p] - vv;
    const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
    d.rec[RC2(ie, iv + 4 * isg, 2)] = uv;
    d.rec[RC2(ie, iv + 4 * isg, 3)] = vv;
    d.rec[RC2(ie, iv + 4 * isg, 4)] = pv;
    for (int f = 0; f < 2; ++f) {
        const int tt = FNQSG[isg][2 * iv + f];
        const int sl = (tt > 0 ? tt : -tt) - 1;
        const double sg = tt > 0 ? 1.0 : -1.0;
        const double nx = sg * sgq[cl][3 * sl + 0], ny = sg * sgq[cl][3 * sl + 1];
        const double area = sgq[cl][3 * sl + 2] * WFR[isg][iv][f];
        const double deltu = fabs(dux * nx + duy * ny);
        const double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
        double m11;
        if (a3 < 1.e12) {
            m11 = mur * area * fabs(1.0);
        } else {
            const double nrm = sqrt(dux * dux + duy * duy);
            m11 = mur * area * fabs(nx * (-dux / nrm) + ny * (-duy / nrm));
        }
        d.rec[RC2(ie, iv + 4 * isg, f)] = m11;
    }
}

// The same corner kernel organised for throughput on large meshes: one lane
// per (cell, sub-cell), 64 cells per block.  Each lane sums its sub-cell's
// 25 quadrature terms in registers in point order, so all lanes of a wave
// evaluate the same point at the same time; the shape-function derivatives
// at the quadrature points are read from a table (built by
// k_q2_shape_tables with the same arithmetic) staged in LDS.  Face slots
// are shared through LDS.  Results are identical to k_q2_corner's.
// Template shape: CPB cells per workgroup (4 CPB threads) and whether the
// shape-function table is staged in LDS (STAGE) or read through the caches;
// the autotuner picks among them.
constexpr int CSC_CELLS = 64;

__global__ void __launch_bounds__(256) k_q2_shape_tables(double *tab)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= SHP_POINTS) return;
    const double sr[4] = {-1.0, 1.0, 1.0, -1.0}, ss[4] = {-1.0, -1.0, 1.0, 1.0};
    double r, s;
    if (k < 25) {
        r = PG5[k / 5];
        s = PG5[k % 5];
    } else if (k < 125) {
        const int isg = (k - 25) / 25, g = (k - 25) % 25;
        r = (PG5[g / 5] + sr[isg]) / 2.0;
        s = (PG5[g % 5] + ss[isg]) / 2.0;
    } else {
        r = PL5[(k - 125) / 5];
        s = PL5[(k - 125) % 5];
    }
    double sh[9], dr_[9], ds_[9];
    shape9(r, s, sh, dr_, ds_);
    for (int q = 0; q < 9; ++q) {
        tab[k * SHP_STRIDE + q] = dr_[q];
        tab[k * SHP_STRIDE + 9 + q] = ds_[q];
        if (k < 25) tab[SHP_POINTS * SHP_STRIDE + k * 9 + q] = sh[q];
    }
}


// Frozen switches of the adjoint's recorded forward (MODE 1 records them,
// MODE 2 replays them, MODE 0 is the plain step): a max(eps, v) clamp and
// the impedance's |du.n|.
template <int MODE>
__device__ __forceinline__ double clamp_eps(double eps, double v, uint8_t *tape, size_t k)
{
    if constexpr (MODE == 0) {
        return dmax(eps, v);
    } else if constexpr (MODE == 1) {
        const bool c = eps > v;
        tape[k] = c ? 1 : 0;
        return c ? eps : v;
    } else {
        return tape[k] ? eps : v;
    }
}

// NF = 5: the nfint=5 branch's corners (getriem_quadsubg_gauss,
// Getrhslaggauss.f90:16657-17157): its corner weights WR5, the sound speed
// of GetEOS without a clamp on c^2 (Getsolid.f90:319-397), the impedance
// rho c + rho slpdu (cimpd |du.n|), and the sub-cell density correction
// drhosgq stored for the face Gauss points.
template <int ND, int MODE = 0, int CPB = CSC_CELLS, int STAGE = 1, int NF = 4>
__global__ void __launch_bounds__(4 * CPB) k_q2_corner_sc(Dev d)
{
    static_assert(NF == 4 || MODE == 0, "the adjoint's recorded forward runs the nfint=4 branch");
    __shared__ double stab_lds[STAGE ? 125 * SHP_STRIDE : 1];
    __shared__ double sgq[CPB][16 * 3 + 1];
    if constexpr (STAGE)
        for (int k = threadIdx.x; k < 125 * SHP_STRIDE; k += 4 * CPB) stab_lds[k] = d.shp[k];
    const double *stab = STAGE ? stab_lds : d.shp;
    const int cl = threadIdx.x >> 2, isg = threadIdx.x & 3;
    const int ie = blockIdx.x * CPB + cl;
    const int nc = d.nc;
    const bool active = ie < nc && ie >= d.ntria;  // triangles: k_t2_corner
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    double xp[9], yp[9];
    double rc = 0.0, sc = 0.0, gg0 = 0.0, gg1 = 0.0, gg2 = 0.0;
    double a = 0.0, masel = 0.0, volel = 0.0;
    if (active) {
        load9(d, d.xs, d.ys, ie, xp, yp);
        rc = d.G[0 * nc + ie];
        sc = d.G[1 * nc + ie];
        gg0 = d.G[18 * nc + ie];
        gg1 = d.G[19 * nc + ie];
        gg2 = d.G[20 * nc + ie];
        // ---- face slots 4 isg .. 4 isg + 3: 8-11 interior half-edges, others boundary points
        for (int u = 0; u < 4; ++u) {
            const int slot = 4 * isg + u;
            double gx, gy, ga;
            if (slot >= 8 && slot < 12) {
                const int k = slot - 8;
                const double pq[4][2] = {{0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
                double xv, yv;
                map9(xp, yp, pq[k][0], pq[k][1], &xv, &yv);
                const double fx[3] = {xp[4 + k], xp[8], xv};
                const double fy[3] = {yp[4 + k], yp[8], yv};
                const double r = -1.0;
                const double dd[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
                double dxdr = 0.0, dydr = 0.0;
                for (int i = 0; i < 3; ++i) {
                    dxdr = dxdr + dd[i] * fx[i];
                    dydr = dydr + dd[i] * fy[i];
                }
                const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
                const double anx = -dydr, any = dxdr;
                gx = anx / sqrt(anx * anx + any * any);
                gy = any / sqrt(anx * anx + any * any);
                ga = djaco * 2.0;
            } else {
                const int ig = slot < 8 ? slot : slot - 4;
                const double pi[12][2] = {{-1, -1}, {1, -1}, {1, -1}, {1, 1}, {1, 1}, {-1, 1},
                                          {-1, 1},  {-1, -1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
                const double vn[12][2] = {{0, -1}, {0, -1}, {1, 0}, {1, 0}, {0, 1}, {0, 1},
                                          {-1, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
                double J[4];
                jac9(xp, yp, pi[ig][0], pi[ig][1], J);
                const double c11 = J[3], c12 = -J[2], c21 = -J[1], c22 = J[0];
                const double anx = c11 * vn[ig][0] + c12 * vn[ig][1];
                const double any = c21 * vn[ig][0] + c22 * vn[ig][1];
                const double w1 = anx * 2.0, w2 = any * 2.0;
                const double farea = sqrt(w1 * w1 + w2 * w2);
                gx = anx * 2.0 / farea;
                gy = any * 2.0 / farea;
                ga = farea;
            }
            sgq[cl][3 * slot + 0] = gx;
            sgq[cl][3 * slot + 1] = gy;
            sgq[cl][3 * slot + 2] = ga;
            d.gq[(3 * slot + 0) * nc + ie] = gx;
            d.gq[(3 * slot + 1) * nc + ie] = gy;
            d.gq[(3 * slot + 2) * nc + ie] = ga;
        }
    }
    __syncthreads();  // shape table staged, face slots shared
    if (!active) return;
    {
        // ---- sub-cell isg: area of its Q2 map and the mass-weighted mean
        //      specific volume of the DG polynomial, summed in point order
        double u0[ND];
        for (int id = 0; id < ND; ++id) u0[id] = UDN(cell_dofs(d, ie), 0, id, ie);
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
        const double rhomc = u0[0];
        for (int g = 0; g < 25; ++g) {
            const int i = g / 5, j = g - 5 * i;
            double J[4];
            jac9_tab(xs, ys, &stab[g * SHP_STRIDE], J);
            a = a + WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
            const double r = (PG5[i] + sr[isg]) / 2.0;
            const double s = (PG5[j] + ss[isg]) / 2.0;
            double b[6];
            basis6(r, s, rc, sc, gg0, gg1, gg2, b);
            double t = 0.0;
            for (int id = 0; id < ND; ++id) t = t + u0[id] * b[id];
            t = rhomc + 1.0 * (t - rhomc);
            jac9_tab(xp, yp, &stab[(25 + 25 * isg + g) * SHP_STRIDE], J);
            const double djaco = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            masel = masel + 1.0 / t * djaco;
            volel = volel + djaco;
        }
    }
    const double rsub = 0.0 + (1.0 / a) * d.msub[isg * nc + ie];
    const double uq = volel / masel;
    double u[4][ND];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) u[q][id] = UDN(cell_dofs(d, ie), q, id, ie);
    const double rhomc = u[0][0], uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
    const double rhoct = 1.0 / rhomc;
    // the centre's clamps are the same on the 4 lanes of a cell: lane 0 records
    uint8_t cdummy[2];
    uint8_t *ctape = (MODE == 1 && isg != 0) ? cdummy : d.ccl;
    const size_t co = (MODE == 1 && isg != 0) ? 0 : (size_t)ie;
    const size_t cs = (MODE == 1 && isg != 0) ? 1 : (size_t)nc;
    const double pctr = clamp_eps<MODE>(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)), ctape, co);
    double sdctr;
    if constexpr (NF == 5) {
        sdctr = sqrt(d.gamlg * pctr / rhoct);
        d.drho[isg * nc + ie] = rsub - 1.0 / uq;
    } else {
        sdctr = sqrt(clamp_eps<MODE>(eps, d.gamlg * pctr / rhoct, ctape, co + cs));
    }
    const double dcor = d.cdrho * (rsub - 1.0 / uq);
    for (int iv = 0; iv < 4; ++iv) {
        // ---- sub-cell corner (isg, iv)
        const int ln = IPQSG[isg][iv];
        double un[4];
        {
            double b[6];
            basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, b);
            for (int q = 0; q < 4; ++q) un[q] = 0.0;
            for (int id = 0; id < ND; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
        }
        const int p = d.ipq[ln * nc + ie];
        const double rhovt = 1.0 / un[0];
        const double rhovsg = rhovt + dcor;
        const double uv = un[1], vv = un[2], ev = un[3];
        const double pv =
            clamp_eps<MODE>(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)), d.ccl, (size_t)(2 + 4 * isg + iv) * nc + ie);
        const double dux = d.usx[p] - uv, duy = d.usy[p] - vv;
        const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
        d.rec[RC2(ie, iv + 4 * isg, 2)] = uv;
        d.rec[RC2(ie, iv + 4 * isg, 3)] = vv;
        d.rec[RC2(ie, iv + 4 * isg, 4)] = pv;
        for (int f = 0; f < 2; ++f) {
            const int tt = FNQSG[isg][2 * iv + f];
            const int sl = (tt > 0 ? tt : -tt) - 1;
            const double sg = tt > 0 ? 1.0 : -1.0;
            const double nx = sg * sgq[cl][3 * sl + 0], ny = sg * sgq[cl][3 * sl + 1];
            const double area = sgq[cl][3 * sl + 2] * wfr<NF>(isg, iv, f);
            double deltu;
            const size_t od = (size_t)(f + 2 * (iv + 4 * isg)) * nc + ie;
            if constexpr (MODE == 2) {
                deltu = d.dlt[od];
            } else {
                deltu = fabs(dux * nx + duy * ny);
                if constexpr (MODE == 1) d.dlt[od] = deltu;
            }
            double mur;
            if constexpr (NF == 5)
                mur = rhoct * sdctr + rhoct * d.slpdu * (d.cimpd * deltu);
            else
                mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
            double m11;
            if (a3 < 1.e12) {
                m11 = mur * area * fabs(1.0);
            } else {
                const double nrm = sqrt(dux * dux + duy * duy);
                m11 = mur * area * fabs(nx * (-dux / nrm) + ny * (-duy / nrm));
            }
            d.rec[RC2(ie, iv + 4 * isg, f)] = m11;
        }
    }
}

// The corner kernel with one wave per sub-cell: wave isg of a 256-thread
// workgroup evaluates sub-cell isg of the workgroup's 64 cells, a lane per
// cell.  The sub-cell index is uniform across the wave, so
//   - the face slots 4 isg .. 4 isg + 3 and the sub-cell's nodes come from
//     code specialised per sub-cell (corner_wu_faces<ISG>), where the shape
//     functions at the fixed reference points are constants;
//   - the quadrature loop reads the shape-function rows of its points as
//     wave-uniform scalar loads (the table in the constant address space),
//     which the FP64 instructions take as scalar operands;
//   - every per-corner and per-face index (IPQSG, FNQSG, WFR, the corner
//     state and tape offsets) is uniform, and each wave reads and writes
//     its 64 cells' entries as contiguous runs.
// Every sum runs in the order of k_q2_corner_sc: identical results.
typedef __attribute__((address_space(4))) const double kdouble;

__device__ __forceinline__ void jac9_k(const double *xp, const double *yp, const kdouble *t, double J[4])
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

constexpr int CWU_CELLS = 64;

// Face slots 4 ISG .. 4 ISG + 3 of the cell (8-11 interior half-edges, the
// others boundary points) into sgq and gq, and the nodes of sub-cell ISG.
// U0 .. U1 - 1: the slots (of the four) this wave evaluates.
template <int ISG, int U0 = 0, int U1 = 4>
__device__ __forceinline__ void corner_wu_faces(const Dev &d, int ie, int cl, const double *xp, const double *yp,
                                                double (*sgq)[CWU_CELLS], double *xs, double *ys)
{
    const int nc = d.nc;
#pragma unroll
    for (int u = U0; u < U1; ++u) {
        const int slot = 4 * ISG + u;
        double gx, gy, ga;
        if (slot >= 8 && slot < 12) {
            const int k = slot - 8;
            const double pq[4][2] = {{0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
            double xv, yv;
            map9(xp, yp, pq[k][0], pq[k][1], &xv, &yv);
            const double fx[3] = {xp[4 + k], xp[8], xv};
            const double fy[3] = {yp[4 + k], yp[8], yv};
            const double r = -1.0;
            const double dd[3] = {-0.5 + r, 0.5 + r, -2.0 * r};
            double dxdr = 0.0, dydr = 0.0;
            for (int i = 0; i < 3; ++i) {
                dxdr = dxdr + dd[i] * fx[i];
                dydr = dydr + dd[i] * fy[i];
            }
            const double djaco = sqrt(dxdr * dxdr + dydr * dydr);
            const double anx = -dydr, any = dxdr;
            gx = anx / sqrt(anx * anx + any * any);
            gy = any / sqrt(anx * anx + any * any);
            ga = djaco * 2.0;
        } else {
            const int ig = slot < 8 ? slot : slot - 4;
            const double pi[12][2] = {{-1, -1}, {1, -1}, {1, -1}, {1, 1}, {1, 1}, {-1, 1},
                                      {-1, 1},  {-1, -1}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
            const double vn[12][2] = {{0, -1}, {0, -1}, {1, 0}, {1, 0}, {0, 1}, {0, 1},
                                      {-1, 0}, {-1, 0}, {0, -1}, {1, 0}, {0, 1}, {-1, 0}};
            double J[4];
            jac9(xp, yp, pi[ig][0], pi[ig][1], J);
            const double c11 = J[3], c12 = -J[2], c21 = -J[1], c22 = J[0];
            const double anx = c11 * vn[ig][0] + c12 * vn[ig][1];
            const double any = c21 * vn[ig][0] + c22 * vn[ig][1];
            const double w1 = anx * 2.0, w2 = any * 2.0;
            const double farea = sqrt(w1 * w1 + w2 * w2);
            gx = anx * 2.0 / farea;
            gy = any * 2.0 / farea;
            ga = farea;
        }
        sgq[3 * slot + 0][cl] = gx;
        sgq[3 * slot + 1][cl] = gy;
        sgq[3 * slot + 2][cl] = ga;
        d.gq[(3 * slot + 0) * nc + ie] = gx;
        d.gq[(3 * slot + 1) * nc + ie] = gy;
        d.gq[(3 * slot + 2) * nc + ie] = ga;
    }
    const double pc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                              {-1.0, .5},   {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
    const int ipsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
    const int ipqsg[4][4] = {{0, 4, 8, 7}, {4, 1, 5, 8}, {8, 5, 2, 6}, {7, 8, 6, 3}};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        xs[k] = xp[ipqsg[ISG][k]];
        ys[k] = yp[ipqsg[ISG][k]];
        const int c = ipsc[ISG][k];
        map9(xp, yp, pc[c][0], pc[c][1], &xs[4 + k], &ys[4 + k]);
    }
    xs[8] = -0.25 * (xs[0] + xs[1] + xs[2] + xs[3]) + 0.5 * (xs[4] + xs[5] + xs[6] + xs[7]);
    ys[8] = -0.25 * (ys[0] + ys[1] + ys[2] + ys[3]) + 0.5 * (ys[4] + ys[5] + ys[6] + ys[7]);
}

// WPE: the fewest waves per SIMD the register allocation must allow (1 no
// limit, 204 VGPRs; 3 caps it at 168 VGPRs, without spills).
template <int ND, int MODE = 0, int WPE = 1>
__global__ void __launch_bounds__(4 * CWU_CELLS) __attribute__((amdgpu_waves_per_eu(WPE))) k_q2_corner_wu(Dev d)
{
    __shared__ double sgq[16 * 3][CWU_CELLS];
    const int cl = threadIdx.x & (CWU_CELLS - 1);
    const int isg = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ie = blockIdx.x * CWU_CELLS + cl;
    const int nc = d.nc;
    const bool active = ie < nc;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    double xp[9], yp[9], xs[9], ys[9];
    double rc = 0.0, sc = 0.0, gg0 = 0.0, gg1 = 0.0, gg2 = 0.0;
    if (active) {
        load9(d, d.xs, d.ys, ie, xp, yp);
        rc = d.G[0 * nc + ie];
        sc = d.G[1 * nc + ie];
        gg0 = d.G[18 * nc + ie];
        gg1 = d.G[19 * nc + ie];
        gg2 = d.G[20 * nc + ie];
        switch (isg) {
        case 0: corner_wu_faces<0>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 1: corner_wu_faces<1>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 2: corner_wu_faces<2>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        default: corner_wu_faces<3>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        }
    }
    __syncthreads();  // face slots shared
    if (!active) return;
    // ---- sub-cell isg: area of its Q2 map and the mass-weighted mean
    //      specific volume of the DG polynomial, summed in point order
    double a = 0.0, masel = 0.0, volel = 0.0;
    {
        double u0[ND];
        for (int id = 0; id < ND; ++id) u0[id] = UDN(cell_dofs(d, ie), 0, id, ie);
        const double sr = (isg == 1 || isg == 2) ? 1.0 : -1.0, ss = isg >= 2 ? 1.0 : -1.0;
        const double rhomc = u0[0];
        const kdouble *stab = (const kdouble *)d.shp;
#pragma unroll 1
        for (int g = 0; g < 25; ++g) {
            const int i = g / 5, j = g - 5 * i;
            double J[4];
            jac9_k(xs, ys, stab + g * SHP_STRIDE, J);
            a = a + WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
            const double r = (PG5[i] + sr) / 2.0;
            const double s = (PG5[j] + ss) / 2.0;
            double b[6];
            basis6(r, s, rc, sc, gg0, gg1, gg2, b);
            double t = 0.0;
            for (int id = 0; id < ND; ++id) t = t + u0[id] * b[id];
            t = rhomc + 1.0 * (t - rhomc);
            jac9_k(xp, yp, stab + (25 + 25 * isg + g) * SHP_STRIDE, J);
            const double djaco = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            masel = masel + 1.0 / t * djaco;
            volel = volel + djaco;
        }
    }
    const double rsub = 0.0 + (1.0 / a) * d.msub[isg * nc + ie];
    const double uq = volel / masel;
    double u[4][ND];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) u[q][id] = UDN(cell_dofs(d, ie), q, id, ie);
    const double rhomc = u[0][0], uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
    const double rhoct = 1.0 / rhomc;
    // the centre's clamps are the same for the 4 sub-cells: wave 0 records
    uint8_t cdummy[2];
    uint8_t *ctape = (MODE == 1 && isg != 0) ? cdummy : d.ccl;
    const size_t co = (MODE == 1 && isg != 0) ? 0 : (size_t)ie;
    const size_t cs = (MODE == 1 && isg != 0) ? 1 : (size_t)nc;
    const double pctr = clamp_eps<MODE>(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)), ctape, co);
    const double sdctr = sqrt(clamp_eps<MODE>(eps, d.gamlg * pctr / rhoct, ctape, co + cs));
    const double dcor = d.cdrho * (rsub - 1.0 / uq);
#pragma unroll 1
    for (int iv = 0; iv < 4; ++iv) {
        // ---- sub-cell corner (isg, iv)
        const int ln = IPQSG[isg][iv];
        double un[4];
        {
            double b[6];
            basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, b);
            for (int q = 0; q < 4; ++q) un[q] = 0.0;
            for (int id = 0; id < ND; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
        }
        const int p = d.ipq[ln * nc + ie];
        const double rhovt = 1.0 / un[0];
        const double rhovsg = rhovt + dcor;
        const double uv = un[1], vv = un[2], ev = un[3];
        const double pv =
            clamp_eps<MODE>(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)), d.ccl, (size_t)(2 + 4 * isg + iv) * nc + ie);
        const double dux = d.usx[p] - uv, duy = d.usy[p] - vv;
        const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
        d.rec[RC2(ie, iv + 4 * isg, 2)] = uv;
        d.rec[RC2(ie, iv + 4 * isg, 3)] = vv;
        d.rec[RC2(ie, iv + 4 * isg, 4)] = pv;
        for (int f = 0; f < 2; ++f) {
            const int tt = FNQSG[isg][2 * iv + f];
            const int sl = (tt > 0 ? tt : -tt) - 1;
            const double sg = tt > 0 ? 1.0 : -1.0;
            const double nx = sg * sgq[3 * sl + 0][cl], ny = sg * sgq[3 * sl + 1][cl];
            const double area = sgq[3 * sl + 2][cl] * WFR[isg][iv][f];
            double deltu;
            const size_t od = (size_t)(f + 2 * (iv + 4 * isg)) * nc + ie;
            if constexpr (MODE == 2) {
                deltu = d.dlt[od];
            } else {
                deltu = fabs(dux * nx + duy * ny);
                if constexpr (MODE == 1) d.dlt[od] = deltu;
            }
            const double mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
            double m11;
            if (a3 < 1.e12) {
                m11 = mur * area * fabs(1.0);
            } else {
                const double nrm = sqrt(dux * dux + duy * duy);
                m11 = mur * area * fabs(nx * (-dux / nrm) + ny * (-duy / nrm));
            }
            d.rec[RC2(ie, iv + 4 * isg, f)] = m11;
        }
    }
}

// The wave-per-sub-cell corner kernel with two waves per sub-cell, for
// meshes too small to fill the chip, where a wave's instruction latency is
// the kernel's duration.  Waves isg and isg + 4 share sub-cell isg of the
// workgroup's 64 cells: each evaluates two of its four face slots and two of
// its four corners; the first sums the quadrature terms of points
// 0 .. CW2_SPLIT - 1 in order, the second evaluates the terms of the other
// points into LDS, and the first then continues its sums with them, so
// every sum runs in point order as in k_q2_corner_wu.
constexpr int CW2_SPLIT = 13;

template <int ND, int NF = 4>
__global__ void __launch_bounds__(8 * CWU_CELLS) k_q2_corner_wu2(Dev d)
{
    __shared__ double sgq[16 * 3][CWU_CELLS];
    __shared__ double sterm[4][3][25 - CW2_SPLIT][CWU_CELLS];  // second wave's terms
    __shared__ double ssum[4][3][CWU_CELLS];                   // a, masel, volel
    const int cl = threadIdx.x & (CWU_CELLS - 1);
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int isg = w & 3, half = w >> 2;
    const int ie = blockIdx.x * CWU_CELLS + cl;
    const int nc = d.nc;
    const bool active = ie < nc;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    double xp[9], yp[9], xs[9], ys[9];
    double rc = 0.0, sc = 0.0, gg0 = 0.0, gg1 = 0.0, gg2 = 0.0;
    double a = 0.0, masel = 0.0, volel = 0.0;
    if (active) {
        load9(d, d.xs, d.ys, ie, xp, yp);
        rc = d.G[0 * nc + ie];
        sc = d.G[1 * nc + ie];
        gg0 = d.G[18 * nc + ie];
        gg1 = d.G[19 * nc + ie];
        gg2 = d.G[20 * nc + ie];
        switch (w) {
        case 0: corner_wu_faces<0, 0, 2>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 1: corner_wu_faces<1, 0, 2>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 2: corner_wu_faces<2, 0, 2>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 3: corner_wu_faces<3, 0, 2>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 4: corner_wu_faces<0, 2, 4>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 5: corner_wu_faces<1, 2, 4>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        case 6: corner_wu_faces<2, 2, 4>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        default: corner_wu_faces<3, 2, 4>(d, ie, cl, xp, yp, sgq, xs, ys); break;
        }
        // ---- quadrature terms of sub-cell isg at this wave's points
        double u0[ND];
        for (int id = 0; id < ND; ++id) u0[id] = UDN(cell_dofs(d, ie), 0, id, ie);
        const double sr = (isg == 1 || isg == 2) ? 1.0 : -1.0, ss = isg >= 2 ? 1.0 : -1.0;
        const double rhomc = u0[0];
        const kdouble *stab = (const kdouble *)d.shp;
        const int gb = half ? CW2_SPLIT : 0, ge = half ? 25 : CW2_SPLIT;
#pragma unroll 1
        for (int g = gb; g < ge; ++g) {
            const int i = g / 5, j = g - 5 * i;
            double J[4];
            jac9_k(xs, ys, stab + g * SHP_STRIDE, J);
            const double ta = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
            const double r = (PG5[i] + sr) / 2.0;
            const double s = (PG5[j] + ss) / 2.0;
            double b[6];
            basis6(r, s, rc, sc, gg0, gg1, gg2, b);
            double t = 0.0;
            for (int id = 0; id < ND; ++id) t = t + u0[id] * b[id];
            t = rhomc + 1.0 * (t - rhomc);
            jac9_k(xp, yp, stab + (25 + 25 * isg + g) * SHP_STRIDE, J);
            const double djaco = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]) * 0.25;
            const double tm = 1.0 / t * djaco;
            if (half == 0) {
                a = a + ta;
                masel = masel + tm;
                volel = volel + djaco;
            } else {
                sterm[isg][0][g - CW2_SPLIT][cl] = ta;
                sterm[isg][1][g - CW2_SPLIT][cl] = tm;
                sterm[isg][2][g - CW2_SPLIT][cl] = djaco;
            }
        }
    }
    __syncthreads();  // face slots and the second wave's terms shared
    if (half == 0 && active) {
        for (int g = CW2_SPLIT; g < 25; ++g) {
            a = a + sterm[isg][0][g - CW2_SPLIT][cl];
            masel = masel + sterm[isg][1][g - CW2_SPLIT][cl];
            volel = volel + sterm[isg][2][g - CW2_SPLIT][cl];
        }
        ssum[isg][0][cl] = a;
        ssum[isg][1][cl] = masel;
        ssum[isg][2][cl] = volel;
    }
    __syncthreads();  // the sub-cell sums shared
    if (!active) return;
    a = ssum[isg][0][cl];
    masel = ssum[isg][1][cl];
    volel = ssum[isg][2][cl];
    const double rsub = 0.0 + (1.0 / a) * d.msub[isg * nc + ie];
    const double uq = volel / masel;
    double u[4][ND];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) u[q][id] = UDN(cell_dofs(d, ie), q, id, ie);
    const double rhomc = u[0][0], uctr = u[1][0], vctr = u[2][0], ectr = u[3][0];
    const double rhoct = 1.0 / rhomc;
    const double pctr = dmax(eps, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    double sdctr;
    if constexpr (NF == 5) {
        sdctr = sqrt(d.gamlg * pctr / rhoct);
        if (half == 0) d.drho[isg * nc + ie] = rsub - 1.0 / uq;
    } else {
        sdctr = sqrt(dmax(eps, d.gamlg * pctr / rhoct));
    }
    const double dcor = d.cdrho * (rsub - 1.0 / uq);
#pragma unroll 1
    for (int iv = 2 * half; iv < 2 * half + 2; ++iv) {
        // ---- sub-cell corner (isg, iv)
        const int ln = IPQSG[isg][iv];
        double un[4];
        {
            double b[6];
            basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, b);
            for (int q = 0; q < 4; ++q) un[q] = 0.0;
            for (int id = 0; id < ND; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
        }
        const int p = d.ipq[ln * nc + ie];
        const double rhovt = 1.0 / un[0];
        const double rhovsg = rhovt + dcor;
        const double uv = un[1], vv = un[2], ev = un[3];
        const double pv = dmax(eps, g1 * rhovsg * (ev - 0.5 * (uv * uv + vv * vv)));
        const double dux = d.usx[p] - uv, duy = d.usy[p] - vv;
        const double a3 = sqrt(dux * dux + duy * duy) / sdctr;
        d.rec[RC2(ie, iv + 4 * isg, 2)] = uv;
        d.rec[RC2(ie, iv + 4 * isg, 3)] = vv;
        d.rec[RC2(ie, iv + 4 * isg, 4)] = pv;
        for (int f = 0; f < 2; ++f) {
            const int tt = FNQSG[isg][2 * iv + f];
            const int sl = (tt > 0 ? tt : -tt) - 1;
            const double sg = tt > 0 ? 1.0 : -1.0;
            const double nx = sg * sgq[3 * sl + 0][cl], ny = sg * sgq[3 * sl + 1][cl];
            const double area = sgq[3 * sl + 2][cl] * wfr<NF>(isg, iv, f);
            const double deltu = fabs(dux * nx + duy * ny);
            double mur;
            if constexpr (NF == 5)
                mur = rhoct * sdctr + rhoct * d.slpdu * (d.cimpd * deltu);
            else
                mur = rhoct * sdctr + d.cimpd * rhoct * d.slpdu * deltu;
            double m11;
            if (a3 < 1.e12) {
                m11 = mur * area * fabs(1.0);
            } else {
                const double nrm = sqrt(dux * dux + duy * duy);
                m11 = mur * area * fabs(nx * (-dux / nrm) + ny * (-duy / nrm));
            }
            d.rec[RC2(ie, iv + 4 * isg, f)] = m11;
        }
    }
}

// Nodal solve over the sub-cell corners: contributions gathered in the
// reference's accumulation order (cell, sub-cell, vertex, face); walls on
// all three face nodes (Getrhslaggauss.f90:4494-4545).
// NF = 5 (getndvelo_lagsubg_gaussh, Getrhslaggauss.f90:17161-17546): the
// corner weights WR5, and the cell-centre nodes' u* set to zero (:17475);
// their solve divides by a zero impedance sum, as the centre's corner faces
// carry zero weight.
template <int NF = 4>
__global__ void __launch_bounds__(LAG2D_BLOCK) k_node_q2(Dev d)
{
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= d.np) return;
    if (!run_active(d, node_run(d, p))) return;  // frozen run: keep u*
    const int nc = d.nc;
    double a11 = 0.0, a21 = 0.0, a12 = 0.0, a22 = 0.0, mu1 = 0.0, mu2 = 0.0, sn1 = 0.0, sn2 = 0.0;
    for (int k = d.n2c_off[p]; k < d.n2c_off[p + 1]; ++k) {
        const int e = d.n2c[k];
        const int ie = e >> 4, slot = e & 15, isg = slot >> 2, iv = slot & 3;
        const double uv = d.rec[RC2(ie, slot, 2)], vv = d.rec[RC2(ie, slot, 3)], pv = d.rec[RC2(ie, slot, 4)];
        for (int f = 0; f < 2; ++f) {
            double nx, ny, ga, m1, m2, s1, s2, w;
            if (ie < d.ntria) {  // curved triangle: its slots and weights
                tri_face_geom(d.gq, nc, ie, isg, iv, f, nx, ny, ga);
                w = WT2[isg][iv][f];
            } else {
                corner_face_geom(d.gq, nc, ie, isg, iv, f, nx, ny, ga);
                w = wfr<NF>(isg, iv, f);
            }
            const double m = d.rec[RC2(ie, slot, f)];
            corner_face_rec(m, uv, vv, pv, nx, ny, ga * w, m1, m2, s1, s2);
            a11 = a11 + m;
            a21 = a21 + 0.0;
            a12 = a12 + 0.0;
            a22 = a22 + m;
            mu1 = mu1 + m1;
            mu2 = mu2 + m2;
            sn1 = sn1 + s1;
            sn2 = sn2 + s2;
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
    if (NF == 5 && (fl & 16)) ux = uy = 0.0;
    d.usx[p] = ux;
    d.usy[p] = uy;
}

// Sub-cell face integral, Gauss-Lobatto domain integral, the (identity)
// limiter update and, with apply_update, the mass-matrix solve and RK blend.
// 8 cells per 256-thread block.  Staging: 128 corner tasks compute the
// corner-only factors of every face-integral term and the basis at the
// corner's node; 200 Lobatto-point tasks compute velocity, pressure and the
// cofactor-form basis gradients of every DOF.  Sums: wave q owns variable q
// of the 8 cells, lane (cell, id) accumulates its residual entry over the
// corners and points in the reference's sequential order (each term is the
// staged factor times the basis value, rounded as the reference rounds it),
// then applies the update.
constexpr int RHS_CPB = 8;
// LDS strides (32 banks of 4 bytes): in the sums a wave reads 8 cells x 6
// DOFs -- the 6 lanes of a cell read 6 consecutive doubles -- and in the
// staging 64 lanes write the values of 64 different (cell, point) pairs.
// The per-cell strides were chosen by measurement among layouts whose read
// patterns are equally conflict-free on paper: this set waits least on LDS
// (SQ_WAIT_INST_LDS 0.87M against 1.44M cycles for per-cell strides 425 /
// 102 / 193 / 49, 31.8 against 33.8 us per launch at 14,400 cells).
constexpr int RHS_CN = 16 * 12 + 2;  // corner factors
constexpr int RHS_CS = 50;           // per-cell inputs: x[9], y[9], Ul[24], rc, sc, g19, g20, g21
constexpr int RHS_BN = 16 * 6 + 4;   // basis at the corner's node
constexpr int RHS_PS = 17;           // point values: a[6], c[6], u, v, p, w (+1 pad)
constexpr int RHS_PT = 25 * RHS_PS + 1;

constexpr int RHS_FS = 16;  // face-point values: b[6], uf1 nx A, uf2 ny A, f1, f2, uf1 f1, uf2 f2, c1, c2, uf1 c1, uf2 c2

// The face Gauss points' terms of variable Q, DOF id, added to r in the
// order of k_cellrhs_q2_tpc: per face the forces of its two points, then
// their correction forces, each weighted by w2 / 2.  e: a cell's staged
// face-point values (RHS_FS per point).
template <int Q>
__device__ __forceinline__ double face_points_q2(double r, const double *e, int id)
{
    const double wi = LW2 / 2.0;
#pragma unroll
    for (int lf = 0; lf < 4; ++lf) {
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double *p = e + RHS_FS * (2 * lf + k);
            const double b = p[id];
            if (Q == 0) r = r + p[6] * b * wi + p[7] * b * wi;
            if (Q == 1) r = r + p[8] * b * wi;
            if (Q == 2) r = r + p[9] * b * wi;
            if (Q == 3) r = r + p[10] * b * wi + p[11] * b * wi;
        }
        if (Q == 0) continue;
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const double *p = e + RHS_FS * (2 * lf + k);
            const double b = p[id];
            if (Q == 1) r = r + p[12] * b * wi;
            if (Q == 2) r = r + p[13] * b * wi;
            if (Q == 3) r = r + p[14] * b * wi + p[15] * b * wi;
        }
    }
    return r;
}

// NF = 5: the corner faces weighted by WR5 / WV5 / WM5, and the 8 face
// Gauss points of each cell staged by threads 72-135 (basis at the point and
// the point's force, correction and flux factors from the face kernel's
// outputs), summed after the corners in the order of k_cellrhs_q2_tpc.

template <int NF>
__global__ void __launch_bounds__(256) k_cellrhs_q2(Dev d, double alfa, int apply_update)
{
    __shared__ double scn[RHS_CPB][RHS_CN];
    __shared__ double sbn[RHS_CPB][RHS_BN];
    __shared__ double spt[RHS_CPB][RHS_PT];
    __shared__ double sres[RHS_CPB][4][6];
    __shared__ double scl[RHS_CPB][RHS_CS];
    __shared__ double sfp[NF == 5 ? RHS_CPB : 1][NF == 5 ? 8 * RHS_FS : 1];
    const int tid = threadIdx.x;
    const int nc = d.nc;
    const int ie0 = xcd_block(blockIdx.x, gridDim.x) * RHS_CPB;
    if (blockIdx.x == 0 && tid == 0) {  // WENO list of this stage consumed
        *d.tsum += (unsigned long long)*d.tcount;
        *d.tcount = 0;
    }
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    // the cells' node coordinates, limited DOFs and basis geometry, loaded
    // once per cell for the 16 corner and 25 point tasks that read them
    for (int u = tid; u < RHS_CPB * 47; u += 256) {
        const int cl = u / 47, k = u - 47 * cl, ie = ie0 + cl;
        if (ie >= nc) continue;
        double v;
        if (k < 9)
            v = d.xs[d.ipq[k * nc + ie]];
        else if (k < 18)
            v = d.ys[d.ipq[(k - 9) * nc + ie]];
        else if (k < 42)
            v = cell_dofs(d, ie)[(k - 18) * nc + ie];
        else
            v = d.G[(k == 42 ? 0 : (k == 43 ? 1 : 18 + (k - 44))) * nc + ie];
        scl[cl][k] = v;
    }
    __syncthreads();
    // corners on threads 0-127, points on threads 128-255 and 0-71
    {
        const int cl = tid >> 4, corner = tid & 15, ie = ie0 + cl;
        if (tid < RHS_CPB * 16 && ie < nc) {
            const int isg = corner >> 2, iv = corner & 3;
            const int ln = IPQSG[isg][iv];
            const int p = d.ipq[ln * nc + ie];
            const double us1 = d.usx[p], us2 = d.usy[p];
            double nx[2], ny[2], ar[2], fx[2], fy[2];
            const double uv = d.rec[RC2(ie, corner, 2)], vv = d.rec[RC2(ie, corner, 3)];
            const double pv = d.rec[RC2(ie, corner, 4)];
            for (int f = 0; f < 2; ++f) {
                double ga, m1, m2, s1, s2;
                corner_face_geom(d.gq, nc, ie, isg, iv, f, nx[f], ny[f], ga);
                ar[f] = ga * wfv<NF>(isg, iv, f);
                const double m11 = d.rec[RC2(ie, corner, f)];
                corner_face_rec(m11, uv, vv, pv, nx[f], ny[f], ga * wfr<NF>(isg, iv, f), m1, m2, s1, s2);
                fx[f] = s1 + m11 * us1 + 0.0 * us2 - m1;
                fy[f] = s2 + 0.0 * us1 + m11 * us2 - m2;
            }
            double *c = &scn[cl][12 * corner];
            c[0] = us1 * nx[0] * ar[0];  // tau: (u*.n A) b
            c[1] = us2 * ny[0] * ar[0];
            c[2] = us1 * nx[1] * ar[1];
            c[3] = us2 * ny[1] * ar[1];
            c[4] = fx[0];  // momentum: f* b w
            c[5] = fx[1];
            c[6] = fy[0];
            c[7] = fy[1];
            c[8] = us1 * fx[0];  // energy: (u*.f*) b w
            c[9] = us2 * fy[0];
            c[10] = us1 * fx[1];
            c[11] = us2 * fy[1];
            double bb[6];
            const double *sc_ = scl[cl];
            basis6(XV9[ln], YV9[ln], sc_[42], sc_[43], sc_[44], sc_[45], sc_[46], bb);
            for (int id = 0; id < 6; ++id) sbn[cl][6 * corner + id] = bb[id];
        }
    }
    if constexpr (NF == 5) {
        const int u = tid - 72, cl = u >> 3, g = u & 7, ie = ie0 + cl;
        if (u >= 0 && u < 8 * RHS_CPB && ie < nc) {
            const int lf = g >> 1, k = g & 1;
            const int pn[3] = {d.ipq[lf * nc + ie], d.ipq[((lf + 1) & 3) * nc + ie], d.ipq[(4 + lf) * nc + ie]};
            const double fxn[3] = {d.xs[pn[0]], d.xs[pn[1]], d.xs[pn[2]]};
            const double fyn[3] = {d.ys[pn[0]], d.ys[pn[1]], d.ys[pn[2]]};
            double gnx, gny, gsa;
            face_normal5(fxn, fyn, k, gnx, gny, gsa);
            const double *sc_ = scl[cl];
            double *e = &sfp[NF == 5 ? cl : 0][RHS_FS * g];
            basis6(PQG5[g][0], PQG5[g][1], sc_[42], sc_[43], sc_[44], sc_[45], sc_[46], e);
            const double f1 = d.fg[FG5(g, 0) * nc + ie], f2 = d.fg[FG5(g, 1) * nc + ie];
            const double c1 = d.fg[FG5(g, 2) * nc + ie], c2 = d.fg[FG5(g, 3) * nc + ie];
            const double uf1 = d.fg[FG5(g, 4) * nc + ie], uf2 = d.fg[FG5(g, 5) * nc + ie];
            e[6] = uf1 * gnx * gsa;
            e[7] = uf2 * gny * gsa;
            e[8] = f1;
            e[9] = f2;
            e[10] = uf1 * f1;
            e[11] = uf2 * f2;
            e[12] = c1;
            e[13] = c2;
            e[14] = uf1 * c1;
            e[15] = uf2 * c2;
        }
    }
    for (int u = (tid + 256 - RHS_CPB * 16) & 255; u < RHS_CPB * 25; u += 256) {
        const int cl = u / 25, pt = u - 25 * cl, ie = ie0 + cl;
        if (ie >= nc) continue;
        const int i = pt / 5, j = pt - 5 * i;
        const double rr = PL5[i], ss = PL5[j], wi = WL5[i] * WL5[j];
        const double *sc_ = scl[cl];
        const double rc = sc_[42], sc = sc_[43];
        double J[4];
        jac9(sc_, sc_ + 9, rr, ss, J);
        double b[6];
        basis6(rr, ss, rc, sc, sc_[44], sc_[45], sc_[46], b);
        double un[4] = {0.0, 0.0, 0.0, 0.0};
        for (int id = 0; id < 6; ++id)
            for (int q = 0; q < 4; ++q) un[q] = un[q] + sc_[18 + 6 * q + id] * b[id];
        const double rhoad = 1.0 / un[0];
        const double uu = un[1], vv = un[2];
        const double pres = dmax(eps, g1 * rhoad * (un[3] - 0.5 * (uu * uu + vv * vv)));
        const double dxdr = J[0], dxds = J[1], dydr = J[2], dyds = J[3];
        const double dbdr[6] = {0.0, 1.0 / 1.0, 0.0, (rr - rc) / 1.0, 0.0, (ss - sc) / 1.0 / 1.0};
        const double dbds[6] = {0.0, 0.0, 1.0 / 1.0, 0.0, (ss - sc) / 1.0, (rr - rc) / 1.0 / 1.0};
        double *s = &spt[cl][RHS_PS * pt];
        for (int id = 0; id < 6; ++id) {
            s[id] = dyds * dbdr[id] + (-dydr) * dbds[id];
            s[6 + id] = (-dxds) * dbdr[id] + dxdr * dbds[id];
        }
        s[12] = uu;
        s[13] = vv;
        s[14] = pres;
        s[15] = wi;
    }
    __syncthreads();
    const int q = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
    const int l = tid & 63, cl = l / 6, id = l - 6 * cl, ie = ie0 + cl;
    const bool active = l < 6 * RHS_CPB && ie < nc;
    // operands of the update, loaded ahead of the sums
    double uqi = 0.0, uo = 0.0, un = 0.0, mrow[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (active) {
        const double *Ue = cell_dofs(d, ie);
        uqi = UD(Ue, q, id, ie);
        if ((id == 1 || id == 2) && (q == 1 || q == 2)) uo = UD(Ue, 3 - q, id, ie);
        if (apply_update) {
            un = UD(d.Un, q, id, ie);
            // row id of the unpacked inverse mass matrix (degree 0 decoupled;
            // upper triangle of the 5x5 block packed row-wise, then 1/mass)
            const int off[6] = {0, 0, 5, 9, 12, 14};
            for (int kk = 0; kk < 6; ++kk) {
                if (id == 0 || kk == 0) {
                    mrow[kk] = (id == 0 && kk == 0) ? d.A[15 * nc + ie] : 0.0;
                } else {
                    const int a = id < kk ? id : kk, b = id < kk ? kk : id;
                    mrow[kk] = d.A[(off[a] + (b - a)) * nc + ie];
                }
            }
        }
        double r = 0.0;
        const double *cn = scn[cl];
        const double *bn = &sbn[cl][id];
        const double *pt = spt[cl];
        // ---- face integral, corners in (isg, iv) order; then the domain
        //      integral (5x5 Gauss-Lobatto, djaco = w, cofactor form)
        if (q == 0) {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double b = bn[6 * k];
                r = r + cn[12 * k + 0] * b + cn[12 * k + 1] * b + cn[12 * k + 2] * b + cn[12 * k + 3] * b;
            }
            if constexpr (NF == 5)
                r = face_points_q2<0>(r, sfp[NF == 5 ? cl : 0], id);
#pragma unroll
            for (int k = 0; k < 25; ++k) {
                const double *s = &pt[RHS_PS * k];
                const double f0 = s[id] * s[12] + s[6 + id] * s[13];
                r = r - f0 * s[15];
            }
        } else if (q == 1 || q == 2) {
            const int o = (q == 1) ? 4 : 6;
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double b = bn[6 * k];
                const double w1 = wfm<NF>(k >> 2, k & 3, 0), w2 = wfm<NF>(k >> 2, k & 3, 1);
                r = r + cn[12 * k + o] * b * w1 + cn[12 * k + o + 1] * b * w2;
            }
            if constexpr (NF == 5) {
                if (q == 1)
                    r = face_points_q2<1>(r, sfp[NF == 5 ? cl : 0], id);
                else
                    r = face_points_q2<2>(r, sfp[NF == 5 ? cl : 0], id);
            }
            const int ac = (q == 1) ? 0 : 6;
#pragma unroll
            for (int k = 0; k < 25; ++k) {
                const double *s = &pt[RHS_PS * k];
                r = r - s[ac + id] * (-s[14]) * s[15];
            }
        } else {
#pragma unroll
            for (int k = 0; k < 16; ++k) {
                const double b = bn[6 * k];
                const double w1 = wfm<NF>(k >> 2, k & 3, 0), w2 = wfm<NF>(k >> 2, k & 3, 1);
                r = r + cn[12 * k + 8] * b * w1 + cn[12 * k + 9] * b * w1 + cn[12 * k + 10] * b * w2 +
                    cn[12 * k + 11] * b * w2;
            }
            if constexpr (NF == 5)
                r = face_points_q2<3>(r, sfp[NF == 5 ? cl : 0], id);
#pragma unroll
            for (int k = 0; k < 25; ++k) {
                const double *s = &pt[RHS_PS * k];
                r = r - (s[id] * s[12] + s[6 + id] * s[13]) * (-s[14]) * s[15];
            }
        }
        sres[cl][q][id] = r;
    }
    __syncthreads();
    if (!active) return;
    // ---- derivative update with aflim = 1, afvec = I (Getrhslagmc.f90:7695-7708)
    double ul = uqi;
    if (id == 1 || id == 2) {
        if (q == 0 || q == 3) ul = uqi * 1.0;
        else if (q == 1) ul = 1.0 * uqi + 0.0 * uo;
        else ul = 0.0 * uo + 1.0 * uqi;
    }
    if (apply_update) {
        const int run = cell_run(d, ie);
        if (!run_active(d, run)) {  // frozen run: keep U^n
            UD(d.Us, q, id, ie) = un;
            return;
        }
        const double dt = d.scal[2 * run];
        double s4 = 0.0;
        for (int kk = 0; kk < 6; ++kk) s4 = s4 + mrow[kk] * (dt * sres[cl][q][kk]);
        UD(d.Us, q, id, ie) = alfa * un + (1.0 - alfa) * (ul + s4);
    } else {
        UD(d.Us, q, id, ie) = ul;
        UD(d.R, q, id, ie) = sres[cl][q][id];
    }
}

// The cell RHS organised for throughput on large meshes: one thread per
// cell (coalesced structure-of-arrays loads, the residual in registers).
// All lanes of a wave evaluate the same Lobatto point at the same time, so
// the shape-function derivatives are read from the table as wave-uniform
// loads.  Every residual entry accumulates in the same order as in
// k_cellrhs_q2: identical results.
// Registers are capped for 2 waves per SIMD: uncapped, the P2 form takes
// 266 (one wave per SIMD), and the 44 bytes it then spills cost far less
// than the lost latency hiding.
// NF = 5 (rhsifacedg_lagsubgq_gs, Getrhslaggauss.f90:17550-17939): the
// corner faces weighted by WR5 (records) and WM5 / WV5 (integral), then per
// face its two Gauss points' forces and, in a second pass, their correction
// forces, from the face kernel's outputs.
template <int ND, int MODE = 0, int BS = 256, int NF = 4>
__global__ void __launch_bounds__(BS) __attribute__((amdgpu_waves_per_eu(2)))
k_cellrhs_q2_tpc(Dev d, double alfa, int apply_update)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie == 0) {  // WENO list of this stage consumed
        *d.tsum += (unsigned long long)*d.tcount;
        *d.tcount = 0;
    }
    if (ie >= nc || ie < d.ntria) return;  // triangles: k_t2_cellrhs
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    const double rc = d.G[0 * nc + ie], sc = d.G[1 * nc + ie];
    const double gg0 = d.G[18 * nc + ie], gg1 = d.G[19 * nc + ie], gg2 = d.G[20 * nc + ie];
    double r[4][ND];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) r[q][id] = 0.0;
    // ---- face integral, corners in (isg, iv) order
#pragma unroll 1
    for (int k = 0; k < 16; ++k) {
        const int isg = k >> 2, iv = k & 3;
        const int ln = IPQSG[isg][iv];
        const int p = d.ipq[ln * nc + ie];
        const double us1 = d.usx[p], us2 = d.usy[p];
        double nx[2], ny[2], ar[2], fx[2], fy[2];
        const double uv = d.rec[RC2(ie, k, 2)], vv = d.rec[RC2(ie, k, 3)], pv = d.rec[RC2(ie, k, 4)];
        for (int f = 0; f < 2; ++f) {
            double ga, m1, m2, s1, s2;
            corner_face_geom(d.gq, nc, ie, isg, iv, f, nx[f], ny[f], ga);
            ar[f] = ga * wfv<NF>(isg, iv, f);
            const double m11 = d.rec[RC2(ie, k, f)];
            corner_face_rec(m11, uv, vv, pv, nx[f], ny[f], ga * wfr<NF>(isg, iv, f), m1, m2, s1, s2);
            fx[f] = s1 + m11 * us1 + 0.0 * us2 - m1;
            fy[f] = s2 + 0.0 * us1 + m11 * us2 - m2;
        }
        const double w1 = wfm<NF>(isg, iv, 0), w2 = wfm<NF>(isg, iv, 1);
        double bq[6];
        basis6(XV9[ln], YV9[ln], rc, sc, gg0, gg1, gg2, bq);
        for (int id = 0; id < ND; ++id) {
            const double b = bq[id];
            r[0][id] = r[0][id] + us1 * nx[0] * ar[0] * b + us2 * ny[0] * ar[0] * b + us1 * nx[1] * ar[1] * b +
                       us2 * ny[1] * ar[1] * b;
            r[1][id] = r[1][id] + fx[0] * b * w1 + fx[1] * b * w2;
            r[2][id] = r[2][id] + fy[0] * b * w1 + fy[1] * b * w2;
            r[3][id] = r[3][id] + us1 * fx[0] * b * w1 + us2 * fy[0] * b * w1 + us1 * fx[1] * b * w2 +
                       us2 * fy[1] * b * w2;
        }
    }
    if constexpr (NF == 5) {
        // ---- face Gauss points: per face the forces of both points, then
        //      their correction forces, each weighted by w2/2
        const double wi = LW2 / 2.0;
#pragma unroll 1
        for (int lf = 0; lf < 4; ++lf) {
            const int pn[3] = {d.ipq[lf * nc + ie], d.ipq[((lf + 1) & 3) * nc + ie], d.ipq[(4 + lf) * nc + ie]};
            const double fxn[3] = {d.xs[pn[0]], d.xs[pn[1]], d.xs[pn[2]]};
            const double fyn[3] = {d.ys[pn[0]], d.ys[pn[1]], d.ys[pn[2]]};
            double bg[2][6], gnx[2], gny[2], gsa[2];
            for (int k = 0; k < 2; ++k) {
                const int g = 2 * lf + k;
                basis6(PQG5[g][0], PQG5[g][1], rc, sc, gg0, gg1, gg2, bg[k]);
                face_normal5(fxn, fyn, k, gnx[k], gny[k], gsa[k]);
            }
            for (int k = 0; k < 2; ++k) {
                const int g = 2 * lf + k;
                const double f1 = d.fg[FG5(g, 0) * nc + ie], f2 = d.fg[FG5(g, 1) * nc + ie];
                const double uf1 = d.fg[FG5(g, 4) * nc + ie], uf2 = d.fg[FG5(g, 5) * nc + ie];
                for (int id = 0; id < ND; ++id) {
                    const double b = bg[k][id];
                    r[0][id] = r[0][id] + uf1 * gnx[k] * gsa[k] * b * wi + uf2 * gny[k] * gsa[k] * b * wi;
                    r[1][id] = r[1][id] + f1 * b * wi;
                    r[2][id] = r[2][id] + f2 * b * wi;
                    r[3][id] = r[3][id] + uf1 * f1 * b * wi + uf2 * f2 * b * wi;
                }
            }
            for (int k = 0; k < 2; ++k) {
                const int g = 2 * lf + k;
                const double c1 = d.fg[FG5(g, 2) * nc + ie], c2 = d.fg[FG5(g, 3) * nc + ie];
                const double uf1 = d.fg[FG5(g, 4) * nc + ie], uf2 = d.fg[FG5(g, 5) * nc + ie];
                for (int id = 0; id < ND; ++id) {
                    const double b = bg[k][id];
                    r[1][id] = r[1][id] + c1 * b * wi;
                    r[2][id] = r[2][id] + c2 * b * wi;
                    r[3][id] = r[3][id] + uf1 * c1 * b * wi + uf2 * c2 * b * wi;
                }
            }
        }
    }
    // ---- domain integral (5x5 Gauss-Lobatto, djaco = w, cofactor form);
    //      the limited DOFs, read once, also feed the update below
    double ul[4][ND];
    for (int q = 0; q < 4; ++q)
        for (int id = 0; id < ND; ++id) ul[q][id] = UDN(cell_dofs(d, ie), q, id, ie);
    {
        const double(&u)[4][ND] = ul;
        double xp[9], yp[9];
        load9(d, d.xs, d.ys, ie, xp, yp);
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
            const double dbdr[6] = {0.0, 1.0 / 1.0, 0.0, (rr - rc) / 1.0, 0.0, (ss - sc) / 1.0 / 1.0};
            const double dbds[6] = {0.0, 0.0, 1.0 / 1.0, 0.0, (ss - sc) / 1.0, (rr - rc) / 1.0 / 1.0};
            double b[6];
            basis6(rr, ss, rc, sc, gg0, gg1, gg2, b);
            double un[4] = {0.0, 0.0, 0.0, 0.0};
            for (int id = 0; id < ND; ++id)
                for (int q = 0; q < 4; ++q) un[q] = un[q] + u[q][id] * b[id];
            const double rhoad = 1.0 / un[0];
            const double uu = un[1], vv = un[2];
            const double pres = clamp_eps<MODE>(eps, g1 * rhoad * (un[3] - 0.5 * (uu * uu + vv * vv)), d.dcl,
                                                (size_t)k * nc + ie);
            for (int id = 0; id < ND; ++id) {
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
    for (int id = 1; id < 3; ++id) {
        const double u1 = ul[1][id], u2 = ul[2][id];
        ul[0][id] = ul[0][id] * 1.0;
        ul[1][id] = 1.0 * u1 + 0.0 * u2;
        ul[2][id] = 0.0 * u1 + 1.0 * u2;
        ul[3][id] = ul[3][id] * 1.0;
    }
    if (apply_update) {
        const int run = cell_run(d, ie);
        if (!run_active(d, run)) {  // frozen run: keep U^n
            for (int q = 0; q < 4; ++q)
                for (int id = 0; id < ND; ++id) UDN(d.Us, q, id, ie) = UDN(d.Un, q, id, ie);
            return;
        }
        const double dt = d.scal[2 * run];
        for (int id = 0; id < ND; ++id) {
            // row id of the inverse mass matrix (Getrhslag.f90:1830-1881):
            // P2 the upper triangle of the 5x5 block packed row-wise, then
            // 1/mass; P1 a11, a12, a22, 1/mass; degree 0 decoupled
            double mrow[ND];
            if constexpr (ND == 6) {
                const int off[6] = {0, 0, 5, 9, 12, 14};
                for (int kk = 0; kk < 6; ++kk) {
                    if (id == 0 || kk == 0) {
                        mrow[kk] = (id == 0 && kk == 0) ? d.A[15 * nc + ie] : 0.0;
                    } else {
                        const int a = id < kk ? id : kk, b = id < kk ? kk : id;
                        mrow[kk] = d.A[(off[a] + (b - a)) * nc + ie];
                    }
                }
            } else {
                if (id == 0) {
                    mrow[0] = d.A[3 * nc + ie];
                    mrow[1] = 0.0;
                    mrow[2] = 0.0;
                } else {
                    mrow[0] = 0.0;
                    mrow[1] = d.A[(id == 1 ? 0 : 1) * nc + ie];
                    mrow[2] = d.A[(id == 1 ? 1 : 2) * nc + ie];
                }
            }
            for (int q = 0; q < 4; ++q) {
                double s4 = 0.0;
                for (int kk = 0; kk < ND; ++kk) s4 = s4 + mrow[kk] * (dt * r[q][kk]);
                UDN(d.Us, q, id, ie) = alfa * UDN(d.Un, q, id, ie) + (1.0 - alfa) * (ul[q][id] + s4);
            }
        }
    } else {
        for (int q = 0; q < 4; ++q)
            for (int id = 0; id < ND; ++id) {
                UDN(d.Us, q, id, ie) = ul[q][id];
                UDN(d.R, q, id, ie) = r[q][id];
            }
    }
}

// Initial sub-cell masses (getrhosubg_averg second half): time-invariant,
// computed once from the initial coordinates.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2_submass(Dev d)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc || ie < d.ntria) return;
    double xi[9], yi[9];
    load9(d, d.x0, d.y0, ie, xi, yi);
    const double pc[12][2] = {{-0.5, -1.0}, {0.5, -1.0}, {1.0, -.5}, {1.0, .5}, {0.5, 1.0}, {-0.5, 1.0},
                              {-1.0, .5},   {-1.0, -.5}, {0.0, -.5}, {0.5, 0.0}, {0.0, .5}, {-0.5, 0.0}};
    const int ipsc[4][4] = {{0, 8, 11, 7}, {1, 2, 9, 8}, {9, 3, 4, 10}, {11, 10, 5, 6}};
    double xc[12], yc[12];
    for (int k = 0; k < 12; ++k) map9(xi, yi, pc[k][0], pc[k][1], &xc[k], &yc[k]);
    double xcc, ycc;
    map9(xi, yi, 0.0, 0.0, &xcc, &ycc);
    const double rho = d.rhoi ? d.rhoi[ie] : ((d.ncase == 6) ? rho0(6, xcc, ycc) : 1.0);
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

// ------------------------------------------------ curved P1 WENO (npoly=1) --
// On the curved P1 branch the reference flags every cell troubled
// (Getrhslagmc.f90:7607) and WENO reconstructs every cell whose stencil is
// complete: weno_stncl_quad_orthol3 with npoly = 1 (WENO.f90:7541-8098),
// getgeophy_local_Taylor / getunknp_local_Taylor with 2 x 2 matrices and
// weno_cell_Taylor with 5 candidates.  Two quirks of that code decide its
// results and are kept: each Lagrange stencil fits 3 equations in 2
// unknowns and Matinv225 reads the first four elements of the 3 x 2 matrix
// in memory order as the 2 x 2 to invert; the npoly = 1 smoothness
// indicator is the P2 formula, reading the entries at flat offsets 3..5
// past (1, iq, is) of an array dimensioned (3, nq, nsten).
// One thread per cell; every sum in the reference's order.
constexpr int P1_ND = 3;
#define UD3(buf, q, id, ie) (buf)[((q) * P1_ND + (id)) * nc + (ie)]

// Matinv225 for n = 2 (LinearAlgebra.f90:141-148, :457-458), column-major
__device__ __forceinline__ void inv2(const double *a, double *b)
{
    const double det = a[0] * a[3] - a[2] * a[1];
    b[0] = a[3] / det;
    b[2] = -a[2] / det;
    b[1] = -a[1] / det;
    b[3] = a[0] / det;
}

static constexpr int ST1P1[4][3] = {{1, 4, 5}, {2, 1, 6}, {3, 2, 7}, {4, 3, 8}};

// One of four values by a face index 0..3 (selects, not an indexed array,
// so nothing goes to scratch memory).
__device__ __forceinline__ double sel4(int f, double a0, double a1, double a2, double a3)
{
    return f == 0 ? a0 : (f == 1 ? a1 : (f == 2 ? a2 : a3));
}

// Characteristic DOF (id, k) of candidate c: sum over the variables in
// order, from 0.
__device__ __forceinline__ double c1_cha(const double (&qi)[4][4], const double (&unksl)[5][4][3], int c, int k,
                                         int id)
{
    double v = 0.0;
    for (int q = 0; q < 4; ++q) v = v + qi[k][q] * unksl[c][q][id];
    return v;
}

// The kernel keeps only what the reconstruction reads: the troubled cell's
// frame geometry, projections and slopes; of the eight neighbours, their
// cell averages and Lagrange-stencil moments (their own slope projections,
// which the reference also forms, feed nothing for npoly = 1).  The loop
// over the neighbours is unrolled, so every array index is static and the
// state stays in registers.
__global__ void __launch_bounds__(LAG2D_BLOCK) k_c1_weno(Dev d, const double *__restrict__ Uin)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    const kdouble *shv = (const kdouble *)d.shv;  // wave-uniform rows: scalar loads
    bool limited = false;
    if (ie < nc) {
        const int flags = d.wtop[9 * nc + ie];
        if (flags & 2) atomicAdd(d.wenobad, 1);
        double m[4];
        {
            const double a = UD3(Uin, 1, 0, ie), b = UD3(Uin, 2, 0, ie);
            if (sqrt(a * a + b * b) > 1.e-10) {
                m[0] = a / sqrt(a * a + b * b);
                m[1] = b / sqrt(a * a + b * b);
                m[2] = -b / sqrt(a * a + b * b);
                m[3] = a / sqrt(a * a + b * b);
            } else {
                m[0] = m[1] = m[2] = m[3] = 0.0;
            }
        }
        // a cell at rest has the zero frame and the reference's normal
        // equations vanish (0/0); outside strict mode it keeps its DOFs
        const bool rest = m[0] == 0.0 && m[1] == 0.0 && m[2] == 0.0 && m[3] == 0.0 && !d.weno_strict;
        if ((flags & 1) || rest || !run_active(d, cell_run(d, ie))) {
            d.flag[ie] = 0;  // not limited: the later kernels read the stage input
        } else {
            limited = true;
            int ielaj[9];
            for (int k = 0; k < 9; ++k) ielaj[k] = d.wtop[k * nc + ie];
            // the stencil's cell averages in the frame (unkrl(1, :, s))
            double avg[9][4];
#pragma unroll
            for (int s = 0; s < 9; ++s) {
                for (int q = 0; q < 4; ++q) avg[s][q] = UD3(Uin, q, 0, ielaj[s]);
                rotf(m, avg[s][1], avg[s][2]);
            }
            double ct1, ct2, ct3, xmcl, ymcl, dxcl;
            double mtbbq[4], mtbbi[4], unkp0[4][3];
            {
                // ---- the troubled cell (stencil cell 0)
                const int je = ielaj[0];
                double u0r[4][3];
                for (int id = 0; id < 3; ++id) {
                    for (int q = 0; q < 4; ++q) u0r[q][id] = UD3(Uin, q, id, je);
                    rotf(m, u0r[1][id], u0r[2][id]);
                }
                double xn[9], yn[9], xr[9], yr[9];
                for (int k = 0; k < 9; ++k) {
                    const int p = d.ipq[k * nc + je];
                    xn[k] = d.xs[p];
                    yn[k] = d.ys[p];
                    xr[k] = xn[k];
                    yr[k] = yn[k];
                    rotf(m, xr[k], yr[k]);
                }
                const double rcj = d.G[0 * nc + je], scj = d.G[1 * nc + je];
                const double volj = d.G[2 * nc + je], masj = d.G[3 * nc + je];
                // physical mass centre on the stage map, rho |J| w of the
                // initial map (WENO.f90:5989-6061)
                double xmc = 0.0, ymc = 0.0;
                for (int g = 0; g < 25; ++g) {
                    const kdouble *sh = shv + 9 * g;
                    double xg = 0.0, yg = 0.0;
                    for (int k = 0; k < 9; ++k) {
                        xg = xg + sh[k] * xn[k];
                        yg = yg + sh[k] * yn[k];
                    }
                    const double rd = d.wrd[g * nc + je];
                    xmc = xmc + rd * xg;
                    ymc = ymc + rd * yg;
                }
                double xcl = xmc / masj, ycl = ymc / masj;
                rotf(m, xcl, ycl);
                const double dxcj = sqrt(volj);
                xmcl = xcl;
                ymcl = ycl;
                dxcl = dxcj;
                double v1 = 0.0, v2 = 0.0, v3 = 0.0;
                double bb[4] = {0.0, 0.0, 0.0, 0.0}, bbq[4] = {0.0, 0.0, 0.0, 0.0};
                for (int g = 0; g < 25; ++g) {
                    const int i = g / 5, j = g - 5 * i;
                    const kdouble *sh = shv + 9 * g;
                    double xl = 0.0, yl = 0.0;
                    for (int k = 0; k < 9; ++k) {
                        xl = xl + sh[k] * xr[k];
                        yl = yl + sh[k] * yr[k];
                    }
                    const double rd = d.wrd[g * nc + je];
                    const double b2 = (xl - xcl) / dxcj, b3 = (yl - ycl) / dxcj;
                    v1 = v1 + rd * b2 * b2;
                    v2 = v2 + rd * b3 * b3;
                    v3 = v3 + rd * b2 * b3;
                    const double bq[3] = {1.0, (PG5[i] - rcj) / 1.0, (PG5[j] - scj) / 1.0};
                    const double bl[3] = {1.0, b2, b3};
                    for (int a = 1; a < 3; ++a)
                        for (int c = 1; c < 3; ++c) {
                            bb[(a - 1) + 2 * (c - 1)] = bb[(a - 1) + 2 * (c - 1)] + rd * bq[a] * bl[c];
                            bbq[(a - 1) + 2 * (c - 1)] = bbq[(a - 1) + 2 * (c - 1)] + rd * bl[a] * bl[c];
                        }
                }
                ct1 = 0.5 * (v1 / masj);
                ct2 = 0.5 * (v2 / masj);
                ct3 = v3 / masj;
                double rhs[4][3];
                for (int q = 0; q < 4; ++q) rhs[q][0] = rhs[q][1] = rhs[q][2] = 0.0;
                for (int a = 1; a < 3; ++a)
                    for (int c = 1; c < 3; ++c)
                        for (int q = 0; q < 4; ++q) rhs[q][a] = rhs[q][a] + u0r[q][c] * bb[(c - 1) + 2 * (a - 1)];
                double bbqi[4];
                inv2(bbq, bbqi);
                for (int k = 0; k < 4; ++k) mtbbq[k] = bbq[k];
                inv2(bb, mtbbi);
                for (int q = 0; q < 4; ++q) unkp0[q][1] = unkp0[q][2] = 0.0;
                for (int a = 1; a < 3; ++a)
                    for (int c = 1; c < 3; ++c)
                        for (int q = 0; q < 4; ++q)
                            unkp0[q][a] = unkp0[q][a] + rhs[q][c] * bbqi[(c - 1) + 2 * (a - 1)];
                for (int q = 0; q < 4; ++q) unkp0[q][0] = u0r[q][0];
            }
            // ---- Lagrange-stencil moments of the neighbours in the troubled
            //      cell's basis
            double ml[9][2];
            ml[0][0] = ml[0][1] = 0.0;
#pragma unroll
            for (int s = 1; s < 9; ++s) {
                const int je = ielaj[s];
                double xr[9], yr[9];
                for (int k = 0; k < 9; ++k) {
                    const int p = d.ipq[k * nc + je];
                    xr[k] = d.xs[p];
                    yr[k] = d.ys[p];
                    rotf(m, xr[k], yr[k]);
                }
                const double masj = d.G[3 * nc + je];
                double m0 = 0.0, m1 = 0.0;
#pragma unroll 1
                for (int g = 0; g < 25; ++g) {
                    const kdouble *sh = shv + 9 * g;
                    double xl = 0.0, yl = 0.0;
                    for (int k = 0; k < 9; ++k) {
                        xl = xl + sh[k] * xr[k];
                        yl = yl + sh[k] * yr[k];
                    }
                    const double rd = d.wrd[g * nc + je];
                    m0 = m0 + rd * ((xl - xmcl) / dxcl) / masj;
                    m1 = m1 + rd * ((yl - ymcl) / dxcl) / masj;
                }
                ml[s][0] = m0;
                ml[s][1] = m1;
                // one neighbour at a time: its loads are not hoisted above the
                // previous neighbour's loop
                asm volatile("" ::: "memory");
                __builtin_amdgcn_sched_barrier(0);
            }
            // candidates: the centre polynomial and the 4 Lagrange stencils
            double unksl[5][4][3];
            for (int c = 0; c < 5; ++c)
                for (int q = 0; q < 4; ++q) {
                    unksl[c][q][0] = unkp0[q][0];
                    unksl[c][q][1] = unksl[c][q][2] = 0.0;
                }
            for (int q = 0; q < 4; ++q) {
                unksl[0][q][1] = unkp0[q][1];
                unksl[0][q][2] = unkp0[q][2];
            }
#pragma unroll
            for (int is = 0; is < 4; ++is) {
                const int s0 = ST1P1[is][0], s1 = ST1P1[is][1], s2 = ST1P1[is][2];
                // matlg(3, 2) read as a column-major 2 x 2 from its first four elements
                const double a[4] = {ml[s0][0], ml[s1][0], ml[s2][0], ml[s0][1]};
                double mi[4];
                inv2(a, mi);
                double rhsl[4][2];
                for (int q = 0; q < 4; ++q) {
                    rhsl[q][0] = avg[s0][q] - unkp0[q][0];
                    rhsl[q][1] = avg[s1][q] - unkp0[q][0];
                }
                for (int i = 1; i < 3; ++i)
                    for (int e = 0; e < 2; ++e)
                        for (int q = 0; q < 4; ++q)
                            unksl[1 + is][q][i] = unksl[1 + is][q][i] + mi[(i - 1) + 2 * e] * rhsl[q][e];
            }
            // weno_cell_Taylor, npoly = 1: characteristic weights per real face
            double xp[4], yp[4];
            for (int k = 0; k < 4; ++k) {
                const int p = d.ipq[k * nc + ie];
                xp[k] = d.xs[p];
                yp[k] = d.ys[p];
            }
            const double volel = d.G[2 * nc + ie];
            const double dxc = sqrt(volel), dyc = dxc;
            // x**4 as ((x*x)*x)*x, as the reference's build evaluates it
            const double dx2 = dxc * dxc, dy2 = dyc * dyc, dx4 = dx2 * dxc * dxc, dy4 = dy2 * dyc * dyc;
            const double dxy = dyc * dxc, dxy2 = dxy * dxy;
            const int nf = d.wtop[10 * nc + ie];
            // the faces' contributions are averaged in face order as they come
            const double wf = 1.0 / (double)nf;
            double uph[4][3];
            for (int q = 0; q < 4; ++q) uph[q][0] = uph[q][1] = uph[q][2] = 0.0;
#pragma unroll 1
            for (int a = 0; a < nf; ++a) {
                const int f = d.wtop[(11 + a) * nc + ie], f1 = (f + 1) & 3;
                const double dtx = sel4(f1, xp[0], xp[1], xp[2], xp[3]) - sel4(f, xp[0], xp[1], xp[2], xp[3]);
                const double dty = sel4(f1, yp[0], yp[1], yp[2], yp[3]) - sel4(f, yp[0], yp[1], yp[2], yp[3]);
                const double dlgt = sqrt(dtx * dtx + dty * dty);
                double dnx = dty / dlgt, dny = -dtx / dlgt;
                rotf(m, dnx, dny);
                double ul[4], ur[4];
                for (int q = 0; q < 4; ++q) {
                    ul[q] = avg[0][q];
                    ur[q] = sel4(f, avg[1][q], avg[2][q], avg[3][q], avg[4][q]);
                }
                double P[9];
                char_face(d.gamlg, dnx, dny, ul, ur, P);
                double qi[4][4];
                for (int k = 0; k < 4; ++k)
                    for (int q = 0; q < 4; ++q) qi[k][q] = char_entry(d.gamlg, P, 4 * k + q);
                // the candidates' characteristic DOFs (unksf_cha(3, 4, 10),
                // column-major, zero past the 5 candidates): the smoothness
                // indicator of candidate c reads its 12 entries and the
                // next candidate's first 3
                double wh[5], wt = 0.0;
#pragma unroll
                for (int c = 0; c < 5; ++c) {
                    double ch[15];
                    for (int k = 0; k < 4; ++k)
                        for (int id = 0; id < 3; ++id) ch[id + 3 * k] = c1_cha(qi, unksl, c, k, id);
                    for (int id = 0; id < 3; ++id) ch[12 + id] = (c < 4) ? c1_cha(qi, unksl, c + 1, 0, id) : 0.0;
                    double o[4];
                    for (int q = 0; q < 4; ++q) {
                        const double u2 = ch[1 + 3 * q], u3 = ch[2 + 3 * q];
                        const double u4 = ch[3 + 3 * q], u5 = ch[4 + 3 * q], u6 = ch[5 + 3 * q];
                        const double sm1 =
                            1.0 / dx2 * (u2 * u2 + 2.0 * ct1 * (u4 * u4) + 2.0 * ct2 * (u6 * u6) + 2.0 * ct3 * u4 * u6);
                        const double sm2 =
                            1.0 / dy2 * (u3 * u3 + 2.0 * ct2 * (u5 * u5) + 2.0 * ct1 * (u6 * u6) + 2.0 * ct3 * u5 * u6);
                        const double sm3 = 1.0 / dx4 * (u4 * u4);
                        const double sm4 = 1.0 / dy4 * (u5 * u5);
                        const double sm5 = 1.0 / dxy2 * (u6 * u6);
                        o[q] = (sm1 + sm2) + (sm3 + sm4 + sm5) * volel;
                    }
                    const double os1 = dmax(dmax(dmax(o[0], o[1]), o[2]), o[3]);
                    const double wl = (c == 0) ? .2 : 0.8 / (5.0 - 1.0);
                    const double t = 1.e-6 + os1;
                    wh[c] = wl / (t * t);
                }
                for (int c = 0; c < 5; ++c) wt = wt + wh[c];
                // the weighted sum recomputes the characteristic DOFs: qi
                // passes through an empty asm statement, so the compiler
                // does not keep all 60 of the first pass live instead
                for (int k = 0; k < 4; ++k)
                    for (int q = 0; q < 4; ++q) asm volatile("" : "+v"(qi[k][q]));
                double fch[4][3];
                for (int q = 0; q < 4; ++q) fch[q][0] = fch[q][1] = fch[q][2] = 0.0;
#pragma unroll
                for (int c = 0; c < 5; ++c)
                    for (int q = 0; q < 4; ++q)
                        for (int id = 0; id < 3; ++id)
                            fch[q][id] = fch[q][id] + (wh[c] / wt) * c1_cha(qi, unksl, c, q, id);
                for (int q = 0; q < 4; ++q)
                    for (int id = 1; id < 3; ++id) {
                        double ufp = 0.0;
                        for (int r = 0; r < 4; ++r) ufp = ufp + char_entry(d.gamlg, P, 16 + 4 * q + r) * fch[r][id];
                        uph[q][id] = uph[q][id] + wf * ufp;
                    }
            }
            // back to the cell's Taylor basis, then the global frame
            double rhsuk[4][2], out[4][3];
            for (int q = 0; q < 4; ++q) rhsuk[q][0] = rhsuk[q][1] = 0.0;
            for (int i = 1; i < 3; ++i)
                for (int j = 1; j < 3; ++j)
                    for (int q = 0; q < 4; ++q)
                        rhsuk[q][i - 1] = rhsuk[q][i - 1] + uph[q][j] * mtbbq[(j - 1) + 2 * (i - 1)];
            for (int q = 0; q < 4; ++q) out[q][1] = out[q][2] = 0.0;
            for (int i = 1; i < 3; ++i)
                for (int j = 1; j < 3; ++j)
                    for (int q = 0; q < 4; ++q) out[q][i] = out[q][i] + rhsuk[q][j - 1] * mtbbi[(j - 1) + 2 * (i - 1)];
            for (int i = 1; i < 3; ++i) {  // getvector_local2 with the inverse frame
                const double a = out[1][i], b = out[2][i];
                out[1][i] = a * m[3] + b * (-m[1]);
                out[2][i] = a * (-m[2]) + b * m[0];
            }
            for (int q = 0; q < 4; ++q) {
                UD3(d.Ul, q, 0, ie) = UD3(Uin, q, 0, ie);
                UD3(d.Ul, q, 1, ie) = out[q][1];
                UD3(d.Ul, q, 2, ie) = out[q][2];
            }
            d.flag[ie] = 1;
        }
    }
    // cells reconstructed this stage (the WENO work counter)
    const unsigned long long bal = __ballot(limited);
    if ((threadIdx.x & 63) == 0 && bal) atomicAdd(d.tcount, (int)__popcll(bal));
}
#undef UD3

#undef UD
#undef UDN

// The WENO reconstruction of the adjoint's recorded and replayed forward:
// its increment U_lim - U_in per cell is a constant of the frozen map.
// RECORD (after the plain WENO, which it leaves as it is, so the recorded
// forward is the plain step): per cell whether WENO limited it and, where
// it did, the limited and the input DOFs go to the tape.  REPLAY: U_lim =
// U_lim,rec + (U_in - U_in,rec) on the limited cells (U_lim,rec itself at
// the recorded input), U_in elsewhere, and every cell flagged so that the
// sub-cell kernels read U_lim.  Tape: d.wdu = [U_lim,rec | U_in,rec]
// [2][4 nd][nc], the limited flags after the clamp branches at
// d.dcl + 25 nc.  Thread per cell.
template <int ND, int RECORD>
__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2_weno_frozen(Dev d, const double *__restrict__ Uin)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc) return;
    uint8_t *limited = d.dcl + (size_t)25 * nc;
    const size_t nu = (size_t)4 * ND * nc;
    if constexpr (RECORD) {
        const bool lim = d.flag[ie] != 0;
        limited[ie] = lim ? 1 : 0;
        if (lim)
            for (int k = 0; k < 4 * ND; ++k) {
                const size_t o = (size_t)k * nc + ie;
                d.wdu[o] = d.Ul[o];
                d.wdu[nu + o] = Uin[o];
            }
    } else {
        const bool lim = limited[ie] != 0;
        for (int k = 0; k < 4 * ND; ++k) {
            const size_t o = (size_t)k * nc + ie;
            d.Ul[o] = lim ? d.wdu[o] + (Uin[o] - d.wdu[nu + o]) : Uin[o];
        }
        d.flag[ie] = 1;
    }
}

}  // namespace lag2d

using namespace lag2d;

static inline int nblk(int n) { return (n + LAG2D_BLOCK - 1) / LAG2D_BLOCK; }

// Time-invariant data of every cell's WENO stencil (the mesh connectivity
// and the initial geometry and density do not change), one thread per cell:
//   wtop[0..8]   stencil cells: the cell, its face neighbours (>= nc on a
//                boundary face) and the vertex neighbour opposite each corner,
//                found by the product of the face numbers it shares
//                (WENO.f90:7659-7700); -2 where none
//   wtop[9]      bit 0: skip (boundary stencil or unclassified vertex
//                neighbourhood), bit 1: a neighbourhood the reference's
//                builder stops on (WENO.f90:7699)
//   wtop[10]     number of real faces, wtop[11..14] their face numbers
//   wrd[g]       rho0 |J| w at Gauss point g of the initial map, rho0 the
//                initial density at the initial mass centre (WENO.f90:5989-6061)
__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2_weno_setup(Dev d, int *wtop, double *wrd)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int nc = d.nc;
    if (ie >= nc || ie < d.ntria) return;
    int ielaj[9];
    ielaj[0] = ie;
    for (int f = 0; f < 4; ++f) ielaj[1 + f] = d.esqua[f * nc + ie];
    for (int k = 5; k < 9; ++k) ielaj[k] = -2;
    int jel[4][3];
    for (int iv = 0; iv < 4; ++iv) {
        int nv = 0;
        for (int a = 0; a < 3; ++a) jel[iv][a] = -1;
        const int p = d.ipq[iv * nc + ie];
        for (int k = d.esuv2[p]; k < d.esuv2[p + 1]; ++k) {
            const int je = d.esuv1[k];
            if (je != ie && nv < 3) jel[iv][nv++] = je;
        }
    }
    bool mismatch = false;
    for (int iv = 0; iv < 4; ++iv) {
        int idxod = 1, idxiv[3] = {0, 0, 0};
        for (int a = 0; a < 3; ++a)
            for (int f = 0; f < 4; ++f)
                if (jel[iv][a] == ielaj[1 + f]) {
                    idxod = idxod * (f + 1);
                    idxiv[a] = 1;
                }
        for (int a = 0; a < 3; ++a) {
            if (idxiv[a]) continue;
            if (idxod == 2) ielaj[5] = jel[iv][a];
            else if (idxod == 6) ielaj[6] = jel[iv][a];
            else if (idxod == 12) ielaj[7] = jel[iv][a];
            else if (idxod == 4) ielaj[8] = jel[iv][a];
            else mismatch = true;
        }
    }
    int skip = mismatch ? 1 : 0;
    for (int k = 0; k < 9; ++k)
        if (ielaj[k] < 0) skip = 1;
    // on a hybrid mesh the reference reads its face-neighbour table by the
    // cell's global number (WENO.f90:7659), so no troubled cell has a
    // defined stencil: every one is skipped, and reported in strict mode
    if (d.ntria > 0) mismatch = true, skip = 1;
    int nf = 0, face[4] = {0, 0, 0, 0};
    for (int f = 0; f < 4; ++f)
        if (ielaj[1 + f] < nc) face[nf++] = f;
    for (int k = 0; k < 9; ++k) wtop[k * nc + ie] = ielaj[k];
    wtop[9 * nc + ie] = skip | (mismatch ? 2 : 0);
    wtop[10 * nc + ie] = nf;
    for (int a = 0; a < 4; ++a) wtop[(11 + a) * nc + ie] = face[a];
    wtop[15 * nc + ie] = 0;

    double xi[9], yi[9];
    for (int q = 0; q < 9; ++q) {
        const int p = d.ipq[q * nc + ie];
        xi[q] = d.x0[p];
        yi[q] = d.y0[p];
    }
    double xcr, ycr;
    map9(xi, yi, d.G[0 * nc + ie], d.G[1 * nc + ie], &xcr, &ycr);
    const double rho = d.rhoi ? d.rhoi[ie] : rho0(d.ncase, xcr, ycr);
    for (int g = 0; g < 25; ++g) {
        const int i = g / 5, j = g - 5 * i;
        double J[4];
        jac9(xi, yi, PG5[i], PG5[j], J);
        const double djacoi = WG5[i] * WG5[j] * (J[0] * J[3] - J[2] * J[1]);
        wrd[g * nc + ie] = rho * djacoi;
    }
}

int upload_extra(lag2d_ctx &x)
{
    if (x.cfg.ncurv != 1) return 0;
    int rc = 0;
    if (!x.d.shp) {
        double *t;
        if ((rc = x.alloc(&t, (size_t)SHP_POINTS * SHP_STRIDE + 25 * 9))) return rc;
        x.d.shp = t;
        x.d.shv = t + SHP_POINTS * SHP_STRIDE;
        hipLaunchKernelGGL(k_q2_shape_tables, dim3(1), dim3(256), 0, x.stream, t);
    }
    hipLaunchKernelGGL(k_q2_submass, dim3(nblk(x.nc)), dim3(LAG2D_BLOCK), 0, x.stream, x.d);
    if (x.ntria > 0)
        hipLaunchKernelGGL(k_t2_submass, dim3(nblk(x.ntria)), dim3(LAG2D_BLOCK), 0, x.stream, x.d);
    if (!x.d.wtop) {
        int *t;
        double *r;
        if ((rc = x.alloc(&t, (size_t)16 * x.nc))) return rc;
        if ((rc = x.alloc(&r, (size_t)25 * x.nc))) return rc;
        x.d.wtop = t;
        x.d.wrd = r;
    }
    hipLaunchKernelGGL(k_q2_weno_setup, dim3(nblk(x.nc)), dim3(LAG2D_BLOCK), 0, x.stream, x.d, (int *)x.d.wtop,
                       (double *)x.d.wrd);
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}

// Persson indicator work split: a thread per cell on meshes large enough to
// fill the chip with it, else 8 lanes per cell.
static void launch_prep(lag2d_ctx *x, Dev &ds, const double *Uin)
{
    const int nc = x->nc;
    if (x->geom_cells)
        hipLaunchKernelGGL(k_q2_prep_tpc, dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, Uin);
    else
        hipLaunchKernelGGL(k_q2_prep, dim3((nc + PREP_CELLS - 1) / PREP_CELLS), dim3(256), 0, x->stream, ds, Uin);
}

// Corner kernel variants (identical results), the autotuner's choices:
//   1    16 lanes per cell, 16 cells per workgroup (short latency, small meshes)
//   2-4  a lane per sub-cell, 64 / 32 / 16 cells per workgroup, the
//        shape-function table staged in LDS (throughput, large meshes)
//   5-6  a lane per sub-cell, 64 / 32 cells per workgroup, the table read
//        through the caches (no LDS staging)
//   7-8  a wave per sub-cell, 64 cells per workgroup (uniform sub-cell
//        index: specialised face code, the table as scalar operands), with
//        registers for 2 or 3 waves per SIMD
//   9    two waves per sub-cell, 64 cells per workgroup (half the
//        instructions per wave: short latency, small meshes)
static void launch_corner(lag2d_ctx *x, Dev &ds, int variant)
{
    const int nc = x->nc;
    auto g = [nc](int cpb) { return dim3((nc + cpb - 1) / cpb); };
    switch (variant) {
    case 2: hipLaunchKernelGGL((k_q2_corner_sc<6, 0, 64, 1>), g(64), dim3(256), 0, x->stream, ds); break;
    case 3: hipLaunchKernelGGL((k_q2_corner_sc<6, 0, 32, 1>), g(32), dim3(128), 0, x->stream, ds); break;
    case 4: hipLaunchKernelGGL((k_q2_corner_sc<6, 0, 16, 1>), g(16), dim3(64), 0, x->stream, ds); break;
    case 5: hipLaunchKernelGGL((k_q2_corner_sc<6, 0, 64, 0>), g(64), dim3(256), 0, x->stream, ds); break;
    case 6: hipLaunchKernelGGL((k_q2_corner_sc<6, 0, 32, 0>), g(32), dim3(128), 0, x->stream, ds); break;
    case 7: hipLaunchKernelGGL((k_q2_corner_wu<6, 0, 1>), g(CWU_CELLS), dim3(4 * CWU_CELLS), 0, x->stream, ds); break;
    case 8: hipLaunchKernelGGL((k_q2_corner_wu<6, 0, 3>), g(CWU_CELLS), dim3(4 * CWU_CELLS), 0, x->stream, ds); break;
    case 9: hipLaunchKernelGGL(k_q2_corner_wu2<6>, g(CWU_CELLS), dim3(8 * CWU_CELLS), 0, x->stream, ds); break;
    default: hipLaunchKernelGGL(k_q2_corner, g(CORNER_CPB), dim3(256), 0, x->stream, ds); break;
    }
}

// Cell-RHS variants (identical results):
//   1    8 cells per workgroup, corner and point values staged in LDS (small meshes)
//   2-4  a thread per cell, workgroups of 256 / 128 / 64 threads (large meshes)
static void launch_cellrhs(lag2d_ctx *x, Dev &ds, double alfa, int apply_update, int variant)
{
    const int nc = x->nc;
    auto g = [nc](int bs) { return dim3((nc + bs - 1) / bs); };
    switch (variant) {
    case 2: hipLaunchKernelGGL((k_cellrhs_q2_tpc<6, 0, 256>), g(256), dim3(256), 0, x->stream, ds, alfa, apply_update); break;
    case 3: hipLaunchKernelGGL((k_cellrhs_q2_tpc<6, 0, 128>), g(128), dim3(128), 0, x->stream, ds, alfa, apply_update); break;
    case 4: hipLaunchKernelGGL((k_cellrhs_q2_tpc<6, 0, 64>), g(64), dim3(64), 0, x->stream, ds, alfa, apply_update); break;
    default:
        hipLaunchKernelGGL(k_cellrhs_q2<4>, g(RHS_CPB), dim3(256), 0, x->stream, ds, alfa, apply_update);
        break;
    }
}

// Autotuning on the context's own mesh: every variant is timed four times,
// in turn (the first round warms them up; the fastest of the other three
// counts), and the fastest is kept.  All variants write the same values,
// so running each costs only time.  Returns the variant, or 0 on a HIP
// error.
template <class Launch>
static int tune_variants(lag2d_ctx *x, int nvar, float *ms, Launch launch)
{
    hipEvent_t a, b;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return 0;
    for (int v = 0; v < nvar; ++v) ms[v] = 1e30f;
    bool ok = true;
    for (int rep = 0; rep < 4 && ok; ++rep)
        for (int v = 1; v <= nvar && ok; ++v) {
            ok = hipEventRecord(a, x->stream) == hipSuccess;
            launch(v);
            ok = ok && hipEventRecord(b, x->stream) == hipSuccess && hipEventSynchronize(b) == hipSuccess;
            float t = 0.f;
            ok = ok && hipEventElapsedTime(&t, a, b) == hipSuccess;
            if (ok && rep > 0 && t < ms[v - 1]) ms[v - 1] = t;
        }
    hipEventDestroy(a);
    hipEventDestroy(b);
    if (!ok) return 0;
    int best = 1;
    for (int v = 2; v <= nvar; ++v)
        if (ms[v - 1] < ms[best - 1]) best = v;
    return best;
}

// LAG2D_CORNER_VARIANT in the environment at context creation fixes the
// choice instead.
static int tune_corner(lag2d_ctx *x, Dev &ds)
{
    const int v = tune_variants(x, LAG2D_CORNER_VARIANTS, x->corner_ms, [&](int k) { launch_corner(x, ds, k); });
    if (!v) return LAG2D_EDEVICE;
    x->corner_variant = v;
    return 0;
}

// The cell RHS reads the stage input (Ui, and Ul for troubled cells), the
// corner records and u*, and writes R and the stage state Us.  Timing it
// several times is harmless only where it does not read what it writes:
// without the fused RK update, and with an input other than Us.  Elsewhere
// (the stages after the first, where Ui == Us and the update is applied)
// the default variant runs untimed and tuning waits for a call that allows
// it.
static int tune_cellrhs(lag2d_ctx *x, Dev &ds, double alfa, int apply_update)
{
    if (apply_update != 0 || ds.Ui == ds.Us) return 0;
    const int v = tune_variants(x, LAG2D_RHS_VARIANTS, x->rhs_ms,
                                [&](int k) { launch_cellrhs(x, ds, alfa, apply_update, k); });
    if (!v) return LAG2D_EDEVICE;
    x->rhs_variant = v;
    return 0;
}

int launch_rhs_q2(lag2d_ctx *x, Dev &ds, const double *Uin, double alfa, int apply_update)
{
    const int nc = x->nc, np = x->np;
    ds.Ui = Uin;
    if (x->nd == 3) {  // curved P1: WENO on every cell, then the sub-cell scheme
        x->tic(KID_Q2_WENO);
        hipLaunchKernelGGL(k_c1_weno, dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, Uin);
        x->toc(KID_Q2_WENO);
        x->tic(KID_Q2_CORNER);
        hipLaunchKernelGGL((k_q2_corner_wu<3, 0, 3>), dim3((nc + CWU_CELLS - 1) / CWU_CELLS), dim3(4 * CWU_CELLS), 0, x->stream, ds);
        x->toc(KID_Q2_CORNER);
        x->tic(KID_NODE_Q2);
        hipLaunchKernelGGL(k_node_q2<4>, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
        x->toc(KID_NODE_Q2);
        x->tic(KID_CELLRHS_Q2);
        hipLaunchKernelGGL(k_cellrhs_q2_tpc<3>, dim3(nblk(nc)), dim3(256), 0, x->stream, ds, alfa, apply_update);
        x->toc(KID_CELLRHS_Q2);
        return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
    }
    if (x->cfg.nfint == 5) {
        // no limiting on this branch (the Persson flag is cleared before
        // the limiter reads it, Getrhslagmc.f90:7763-7766); Gauss-point
        // corners, the node solve, the face points, the face and domain
        // integrals
        x->tic(KID_Q2_CORNER);
        if (x->corner_variant == 9)  // two waves per sub-cell (small meshes)
            hipLaunchKernelGGL((k_q2_corner_wu2<6, 5>), dim3((nc + CWU_CELLS - 1) / CWU_CELLS), dim3(8 * CWU_CELLS),
                               0, x->stream, ds);
        else
            hipLaunchKernelGGL((k_q2_corner_sc<6, 0, 64, 1, 5>), dim3((nc + 63) / 64), dim3(256), 0, x->stream, ds);
        x->toc(KID_Q2_CORNER);
        x->tic(KID_NODE_Q2);
        hipLaunchKernelGGL(k_node_q2<5>, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
        x->toc(KID_NODE_Q2);
        x->tic(KID_Q2G_FACE);
        hipLaunchKernelGGL(k_q2g_face, dim3(nblk(nc), 8), dim3(LAG2D_BLOCK), 0, x->stream, ds);
        x->toc(KID_Q2G_FACE);
        x->tic(KID_CELLRHS_Q2);
        if (x->rhs_variant == 1)  // 8 cells per workgroup (small meshes)
            hipLaunchKernelGGL(k_cellrhs_q2<5>, dim3((nc + RHS_CPB - 1) / RHS_CPB), dim3(256), 0, x->stream, ds,
                               alfa, apply_update);
        else
            hipLaunchKernelGGL((k_cellrhs_q2_tpc<6, 0, 256, 5>), dim3(nblk(nc)), dim3(256), 0, x->stream, ds, alfa,
                               apply_update);
        x->toc(KID_CELLRHS_Q2);
        return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
    }
    x->tic(KID_Q2_PREP);
    launch_prep(x, ds, Uin);
    x->toc(KID_Q2_PREP);
    x->tic(KID_Q2_WENO);
    hipLaunchKernelGGL(k_q2_weno, dim3(nc < WENO_GRID ? nc : WENO_GRID), dim3(WENO_WG), 0, x->stream, ds, Uin);
    x->toc(KID_Q2_WENO);
    if (x->corner_variant == 0) {
        int rc = tune_corner(x, ds);
        if (rc) return rc;
    }
    x->tic(KID_Q2_CORNER);
    launch_corner(x, ds, x->corner_variant);
    if (x->ntria > 0)
        hipLaunchKernelGGL(k_t2_corner, dim3(nblk(x->ntria)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
    x->toc(KID_Q2_CORNER);
    x->tic(KID_NODE_Q2);
    hipLaunchKernelGGL(k_node_q2<4>, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
    x->toc(KID_NODE_Q2);
    if (x->rhs_variant == 0) {
        int rc = tune_cellrhs(x, ds, alfa, apply_update);
        if (rc) return rc;
    }
    x->tic(KID_CELLRHS_Q2);
    launch_cellrhs(x, ds, alfa, apply_update, x->rhs_variant ? x->rhs_variant : 1);
    if (x->ntria > 0)
        hipLaunchKernelGGL(k_t2_cellrhs, dim3(nblk(x->ntria)), dim3(LAG2D_BLOCK), 0, x->stream, ds, alfa,
                           apply_update);
    x->toc(KID_CELLRHS_Q2);
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}

// One stage right-hand side of the adjoint's recorded (record = 1) or
// replayed (record = 0) forward on the curved path: the plain WENO (record
// only) and its frozen increment, then the corner (recording or reading the
// |du.n| and clamp tapes in ds.dlt, ds.ccl, ds.dcl), the nodal solve and
// the cell RHS with the fused update.  The wave-per-sub-cell corner and
// thread-per-cell RHS variants run (their results equal the others').
template <int ND>
static int launch_rhs_q2_adj_nd(lag2d_ctx *x, Dev &ds, const double *Uin, double alfa, int apply_update, int record)
{
    const int nc = x->nc, np = x->np;
    ds.Ui = Uin;
    if (record) {
        if constexpr (ND == 3) {
            hipLaunchKernelGGL(k_c1_weno, dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, Uin);
        } else {
            launch_prep(x, ds, Uin);
            hipLaunchKernelGGL(k_q2_weno, dim3(nc < WENO_GRID ? nc : WENO_GRID), dim3(WENO_WG), 0, x->stream, ds, Uin);
        }
        hipLaunchKernelGGL((k_q2_weno_frozen<ND, 1>), dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, Uin);
        hipLaunchKernelGGL((k_q2_corner_wu<ND, 1, 3>), dim3((nc + CWU_CELLS - 1) / CWU_CELLS), dim3(4 * CWU_CELLS), 0, x->stream, ds);
        hipLaunchKernelGGL(k_node_q2<4>, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
        hipLaunchKernelGGL((k_cellrhs_q2_tpc<ND, 1>), dim3(nblk(nc)), dim3(256), 0, x->stream, ds, alfa, apply_update);
    } else {
        hipLaunchKernelGGL((k_q2_weno_frozen<ND, 0>), dim3(nblk(nc)), dim3(LAG2D_BLOCK), 0, x->stream, ds, Uin);
        hipLaunchKernelGGL((k_q2_corner_wu<ND, 2, 3>), dim3((nc + CWU_CELLS - 1) / CWU_CELLS), dim3(4 * CWU_CELLS), 0, x->stream, ds);
        hipLaunchKernelGGL(k_node_q2<4>, dim3(nblk(np)), dim3(LAG2D_BLOCK), 0, x->stream, ds);
        hipLaunchKernelGGL((k_cellrhs_q2_tpc<ND, 2>), dim3(nblk(nc)), dim3(256), 0, x->stream, ds, alfa, apply_update);
    }
    return hipGetLastError() == hipSuccess ? 0 : LAG2D_EDEVICE;
}

int launch_rhs_q2_adj(lag2d_ctx *x, Dev &ds, const double *Uin, double alfa, int apply_update, int record)
{
    return x->nd == 3 ? launch_rhs_q2_adj_nd<3>(x, ds, Uin, alfa, apply_update, record)
                      : launch_rhs_q2_adj_nd<6>(x, ds, Uin, alfa, apply_update, record);
}
