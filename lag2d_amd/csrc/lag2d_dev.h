// Device-side layout and helpers shared by the lag2d kernels (gfx950, FP64).
//
// HBM layout (structure of arrays, one entry per cell or node so that a
// wavefront of 64 consecutive cells/nodes reads 64 consecutive doubles):
//   U[(q*nd + id)*nc + ie]   DOF id of variable q (tau, u, v, E) of cell ie
//   A[k*nc + ie]             packed inverse mass matrix (amatr row k)
//   G[k*nc + ie]             geometry rows of geoel (row k, 0-based)
//   X[ip], Y[ip]             node coordinates
//   IPQ[iv*nc + ie]          cell nodes (0-based); a triangle uses iv = 0..2
// Arithmetic is FP64 throughout and is built without FMA contraction so the
// results round exactly where the reference's do.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define LAG2D_BLOCK 256

namespace lag2d {

__device__ __forceinline__ double dmax(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmin(double a, double b) { return a < b ? a : b; }

// Fortran default-real literal 1.e-6 widened to REAL(8), as the limiter
// routines write it (Limiterlag.f90:53, :8239; Getrhslagmctria.f90:2408).
static constexpr double EPS_SP6 = (double)1.e-6f;

__device__ __forceinline__ double barthfct(double umax, double umin, double uctr, double dunk)
{
    // Getrhslagmc.f90:9171-9219 (coef 0.7 on both branches of its time test)
    const double coef = 0.7;
    if (dunk > 1.e-8) {
        double fiy = coef * (umax - uctr) / dunk;
        return dmin((fiy * fiy + 2.0 * fiy) / (fiy * fiy + fiy + 2.0), 1.0);
    } else if (dunk < -1.e-8) {
        double fiy = coef * (umin - uctr) / dunk;
        return dmin((fiy * fiy + 2.0 * fiy) / (fiy * fiy + fiy + 2.0), 1.0);
    }
    return 1.0;
}

// Reference-quad vertex coordinates (corner order 1..4, CCW)
static constexpr double XVQ[4] = {-1.0, 1.0, 1.0, -1.0};
static constexpr double YVQ[4] = {-1.0, -1.0, 1.0, 1.0};

struct Dev {
    int nc, np, nb, nd, nvq, nmatr;
    int ntria;               // cells 0..ntria-1 are triangles (straight P1 path), the rest quads
    int nrz;                 // 0 planar, 2 area-weighted RZ (r = y)
    double alfrz;            // 1 for RZ: rcoef = 1 - alfrz + alfrz*y
    int nstag, ncase, ncurv, npoly, ndt;
    double gamlg, cfl, cimpd, cdrho, tend, slpdu;
    // mesh
    const int *ipq;          // [nvq][nc]
    const int *esuv1, *esuv2;
    const int *n2c;          // node -> corner CSR entries (ie*NCORN + corner), offsets n2c_off
    const int *n2c_off;
    const uint8_t *nflag;    // bit0: v=0 wall, bit1: u=0 wall, bit2: fixed wall (no solve),
                             // bit3: node of a slip-wall face (code 22)
    const int *wall;         // RZ: [2][np] slip-wall faces of a node, other node*4 + (1: 221, 2: 222), -1
    const uint8_t *ncol;     // symmetry colouring for the velocity limiter (0,2,3,10)
    // state
    double *xn, *yn;         // x^n
    double *xs, *ys;         // stage coordinates
    const double *x0, *y0;   // initial coordinates
    double *Un, *Us;         // U^n and stage state
    double *A;               // inverse mass matrix (updated per stage on the RZ path)
    double *G;               // geoel rows [24][nc]
    double *usx, *usy;       // nodal velocity
    double *R;               // right-hand side [(q*nd+id)*nc + ie]
    double *cavg;            // [5][nc]  tau, u, v, p, E cell averages
    double *frame;           // [4][nc]  principal-strain frame (row-major 2x2)
    double *nmm;             // [10][np] node max (0..4) and min (5..9)
    double *lim;             // [9][nc]  aflim(1..5), afvec (col-major 2x2)
    double *dlt;             // |du.n| of each corner face's impedance, recorded for the adjoint's
                             // frozen forward: [8][nc] straight P1 quads, [32][nc] curved (null otherwise)
    double *rec;             // corner-face records
    double *red;             // reduction scratch
    double *scal;            // device scalars: 0 dtfix, 1 tacru
    // curved P2 sub-cell path
    double *msub;            // [4][nc] initial sub-cell masses (time-invariant)
    double *Ul;              // limited DOFs of the cells WENO limits (flag set) [(q*nd+id)*nc+ie]
    const double *Ui;        // the stage's input state: the DOFs of every other cell
    double *gq;              // face geometry [48][nc]: slot j -> (nx, ny, |J|) rows 3j..3j+2
    uint8_t *flag;           // Persson troubled flag per cell
    int *tlist;              // compacted troubled-cell list
    int *tcount;             // its length (device counter)
    unsigned long long *tsum;  // troubled cells processed since the counter was last cleared
    const int *esqua;        // face neighbours [4][nc]; >= nc marks a boundary face
    int weno_strict;         // 1: report cells the reference's stencil builder stops on
    int *wenobad;            // count of such cells
    // ensemble of independent runs (nrun > 1): run r owns cells
    // [coff[r], coff[r+1]); scal[2r], scal[2r+1] are its dt and time
    int nrun;
    const int *runc;         // run of each cell
    const int *runn;         // run of each node
    const int *coff;         // cell offsets of the runs [nrun+1]
    const double *tendv;     // end time of each run
    int *act;                // 1: the run advances in this step
    // initial density per cell, for initial states other than the
    // reference's built-in problems (null: the density of ncase)
    const double *rhoi;
    // shape-function derivatives at the curved path's fixed quadrature points,
    // and shape values at the 25 Gauss points
    const double *shp, *shv;
    // time-invariant WENO stencil data per cell: topology [16][nc] (stencil
    // cells, flags, real faces) and rho |J| w at the 25 Gauss points [25][nc]
    const int *wtop;
    const double *wrd;
    // the curved path's frozen switches of the adjoint's recorded forward
    // (null otherwise): WENO's limited and input DOFs of the cells it limited
    // [2][4 nd][nc], and which pressure and sound-speed clamps max(eps, .)
    // chose eps, per cell: corner kernel [18][nc] (centre pressure, sound
    // speed, the 16 sub-cell vertex pressures), domain integral [25][nc]
    // (its points), followed by the limited flags [nc]; the impedances'
    // |du.n| live in dlt [32][nc]
    double *wdu;
    uint8_t *ccl, *dcl;
    // the nfint=5 branch (face Gauss points): nfint, the sub-cell density
    // corrections of the corner solve [4][nc], the face kernel's outputs
    // [48][nc] (FG5 in lag2d_q2geom.h), and each cell face's other side
    // [4][nc]: neighbour * 4 + its local face, or -(1 + bc + 4 rev) on a
    // boundary (bc 0 free, 1 v = 0, 2 u = 0, 3 u = v = 0; rev: the boundary
    // face's nodes run clockwise round the cell)
    int nfint;
    double *drho;
    double *fg;
    const int *fnb;
};

// Quotients a / b by one divisor b, bit-identical to the IEEE division.
// gfx950 divides FP64 with div_scale (operand scaling), rcp, two Newton
// steps on the reciprocal, q0 = a*y, the correction fma(fma(-b, q0, a), y,
// q0) (div_fmas) and div_fixup (special operands).  When both exponents lie
// in [-380, 380] (or a is zero, whose quotient is q0 = a*y, a signed zero)
// the scaling and the fix-up are identities, so the quotient is those
// three operations on the reciprocal, which a shared divisor computes once.
// Other operands take the division itself.
struct Recip {
    double b, y;
    bool ok;
};

__device__ __forceinline__ bool div_exp_ok(double x)
{
    const unsigned e = (unsigned)(__double_as_longlong(x) >> 52) & 0x7ffu;
    return e - (1023u - 380u) <= 760u;
}

__device__ __forceinline__ Recip recip(double b)
{
    double y = __builtin_amdgcn_rcp(b);
    double e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    e = __builtin_fma(-b, y, 1.0);
    y = __builtin_fma(y, e, y);
    return Recip{b, y, div_exp_ok(b)};
}

__device__ __forceinline__ double qdiv(double a, const Recip &r)
{
    const double q0 = a * r.y;
    double q = __builtin_fma(__builtin_fma(-r.b, q0, a), r.y, q0);
    if (a == 0.0) q = q0;
    if (!(r.ok && (div_exp_ok(a) || a == 0.0))) q = a / r.b;
    return q;
}

// The shortcut alone, for a term of a sum that starts at +0: clears ok when
// a needs the division itself (the caller then recomputes the sum with it).
// For a = -0 it returns +0 where the division gives -0; a sum started at +0
// cannot tell them apart.
__device__ __forceinline__ double qdiv_term(double a, const Recip &r, bool &ok)
{
    ok = ok & (div_exp_ok(a) | (a == 0.0));
    const double q0 = a * r.y;
    return __builtin_fma(__builtin_fma(-r.b, q0, a), r.y, q0);
}

// One stage of the discrete adjoint (straight P1 planar quads): the stage's
// forward inputs and outputs, and the adjoint buffers.  Cell arrays are
// [k*nc + ie], node arrays [c*np + p], corner arrays [(c*4 + iv)*nc + ie].
struct AdjStage {
    const double *Uin;           // stage input DOFs [12][nc]
    const double *xin, *yin;     // stage input coordinates
    const double *usx, *usy;     // this stage's nodal velocity
    const double *rec;           // this stage's corner records [40][nc]
    const double *lim;           // this stage's (frozen) limiter coefficients [9][nc]
    const double *dlt;           // this stage's (frozen) impedance |du.n| per corner face [8][nc]
    const double *dt;            // the step's dt (device scalar)
    double *LU, *LUN, *LUI;      // lam of U_out, of U^n (accumulated), of U_in
    double *LX, *LXN, *LXI;      // lam of x_out, of x^n (accumulated), direct part of x_in
    double *LUS, *LUSN;          // lam of this stage's u*, of the next stage's u*_prev
    double *CX, *CU;             // per-corner lam of x_in, u* [2][4][nc]
    double *LREC;                // lam of the corner records [40][nc]
    double *NG;                  // node coefficients of the reversed solve [3][np]
};

// One reversed stage of the curved (Q2 sub-cell) step: the stage's data
// kept by the recomputed segment, its frozen switches, and the adjoint
// buffers (lam = derivative of the objective with respect to that array).
struct AdjStageC {
    const double *Ul;            // limited stage DOFs [4 nd][nc]
    const double *xin, *yin;     // stage input coordinates
    const double *usx, *usy;     // this stage's nodal velocity
    const double *rec;           // this stage's sub-cell corner states (RC2 layout)
    const double *gq;            // this stage's face slots [48][nc]
    const double *dlt;           // frozen |du.n| [32][nc]
    const uint8_t *ccl, *dcl;    // frozen clamp branches [18][nc], [25][nc]
    const double *dt;            // the step's dt (device scalar)
    double *LU, *LUN, *LUI;      // lam of U_out, of U^n (accumulated), of U_in  [4 nd][nc]
    double *LX, *LXN, *LXI;      // lam of x_out, of x^n (accumulated), of x_in  [2][np]
    double *LUS;                 // lam of this stage's u* [2][np]
    double *NG;                  // node coefficients of the reversed solve [3][np]
    double *CU;                  // lam of u* per sub-cell corner [16][2][nc]
    double *CX;                  // lam of the cell's 9 nodes [9][2][nc]
    double *LREC;                // lam of the corner-face records [32][5][nc]
    double *LUSE;                // lam of the face slots per corner face [32][3][nc]
};

__device__ __forceinline__ int cell_run(const Dev &d, int ie) { return d.nrun > 1 ? d.runc[ie] : 0; }
__device__ __forceinline__ int node_run(const Dev &d, int p) { return d.nrun > 1 ? d.runn[p] : 0; }
// A run that has reached its end time is frozen: later steps leave its state unchanged.
__device__ __forceinline__ bool run_active(const Dev &d, int r) { return d.nrun <= 1 || d.act[r] != 0; }

// Logical block index for an XCD-aware order.  Workgroups are dispatched to
// the 8 XCDs round-robin (block b runs on XCD b % 8, each XCD with its own
// L2); this maps the blocks of one XCD to a contiguous range of logical
// blocks, so that neighbouring cells -- which share cache lines of every
// cell-indexed array -- are processed under the same L2.  A bijection on
// [0, nb).
__device__ __forceinline__ int xcd_block(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

}  // namespace lag2d
