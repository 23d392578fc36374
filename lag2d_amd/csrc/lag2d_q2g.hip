// Face Gauss points of the curved P2 sub-cell path's nfint=5 branch
// (getrhsdg_lag_mcnew, Getrhslagmc.f90:7712-7829): on every cell face the
// two interior points of the 5-point Gauss-Lobatto face rule carry a
// Riemann problem of their own between the two cells' DG states, solved for
// a velocity correction of the nodal velocity interpolated along the face.
//
// Reference routines restated here:
//   getndvelo_lagsubg_gaussh  Getrhslaggauss.f90:17416-17527 (face velocity,
//                             forces at the face points)
//   getvelo_quadsubg_gs3      Getrhslaggauss.f90:19312-19990
//   getvar_face_general       Getrhslaggauss.f90:18184-18368
//   getvectf                  Getrhslaggauss.f90:16547-16590
//   getriecoef_matrixnew      Limiterlag.f90:1037-1076
//
// Work split: a thread per (cell, face point), 8 per cell.  Each thread
// solves its point's two-sided problem and writes its own cell's forces;
// the neighbour's thread solves the same problem from the other side.  The
// two solves give the same bits: the shared sums have two terms (added in
// either order), the face normal from the other side is the exact negation
// (the face parameter runs backwards), and the interpolated velocity uses
// the same three products.  Each cell's outputs are written by its own
// threads only, so no atomics and no ordering between threads.
#include "lag2d_dev.h"
#include "lag2d_kernels.h"
#include "lag2d_q2geom.h"

namespace lag2d {

// getvar_face_general: a cell's DG values at its face Gauss point g (tau,
// u, v, E in v[0..3]), the pressure with the sub-cell density correction
// (v[4]), and rho c and c of the cell average (v[5], v[6]).
__device__ __forceinline__ void face_state5(const Dev &d, int ie, int g, double v[7])
{
    const int nc = d.nc;
    const double g1 = d.gamlg - 1.0, eps = 1.0e-6;
    const double *u = d.Ui;
    const double rhoct = 1.0 / u[(0 * 6 + 0) * nc + ie];
    const double uctr = u[(1 * 6 + 0) * nc + ie], vctr = u[(2 * 6 + 0) * nc + ie], ectr = u[(3 * 6 + 0) * nc + ie];
    const double eintc = ectr - 0.5 * (uctr * uctr + vctr * vctr);
    const double pctr = dmax(eps, g1 * rhoct * eintc);
    const double sdctr = sqrt(d.gamlg * pctr / rhoct);
    double b[6];
    basis6(PQG5[g][0], PQG5[g][1], d.G[0 * nc + ie], d.G[1 * nc + ie], d.G[18 * nc + ie], d.G[19 * nc + ie],
           d.G[20 * nc + ie], b);
    for (int q = 0; q < 4; ++q) v[q] = 0.0;
    for (int id = 0; id < 6; ++id)
        for (int q = 0; q < 4; ++q) v[q] = v[q] + u[(q * 6 + id) * nc + ie] * b[id];
    const double rhovt = 1.0 / v[0] + d.cdrho * d.drho[SGG5[g] * nc + ie];
    const double eintv = v[3] - 0.5 * (v[1] * v[1] + v[2] * v[2]);
    v[4] = dmax(eps, g1 * rhovt * eintv);
    v[5] = rhoct * sdctr;
    v[6] = sdctr;
}

// One side's Riemann coefficients at a face point (getvelo_quadsubg_gs3
// :19668-19701, getriecoef_matrixnew): M = mu A (isotropic), M u, A sigma n.
__device__ __forceinline__ void face_riem5(const Dev &d, const double v[7], double uf1, double uf2, double nx,
                                           double ny, double area, double &m, double &mu1, double &mu2,
                                           double &sn1, double &sn2)
{
    const double eps11 = (double)1.e-11f;  // the default-real literal 1.e-11
    double a1 = uf1 - v[1], a2 = uf2 - v[2];
    const double acnx = a1, acny = a2;
    const double nrm = sqrt(acnx * acnx + acny * acny);
    if (nrm < eps11) {
        a1 = eps11;
        a2 = eps11;
    } else {
        a1 = a1 / nrm;
        a2 = a2 / nrm;
    }
    const double a3 = nrm / v[6];
    const double delu = d.cimpd * fabs(a1 * nx + a2 * ny);
    const double mur = v[5] + v[5] / v[6] * d.slpdu * delu;
    if (a3 < 1.0e12) m = mur * area * fabs(1.0);
    else m = mur * area * fabs(nx * a1 + ny * a2);
    mu1 = m * v[1] + 0.0 * v[2];
    mu2 = 0.0 * v[1] + m * v[2];
    sn1 = -v[4] * area * nx + 0.0 * area * ny;
    sn2 = 0.0 * area * nx + -v[4] * area * ny;
}

__global__ void __launch_bounds__(LAG2D_BLOCK) k_q2g_face(Dev d)
{
    const int ie = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;  // face point 0..7
    const int nc = d.nc;
    if (ie >= nc) return;
    const int lf = g >> 1, k = g & 1;
    // the face (corner lf, corner lf + 1, mid-node) in this cell's order
    const int pa = d.ipq[lf * nc + ie], pb = d.ipq[((lf + 1) & 3) * nc + ie], pm = d.ipq[(4 + lf) * nc + ie];
    const double fx[3] = {d.xs[pa], d.xs[pb], d.xs[pm]}, fy[3] = {d.ys[pa], d.ys[pb], d.ys[pm]};
    double nx, ny, ga;
    face_normal5(fx, fy, k, nx, ny, ga);
    // u* interpolated along the face (:17416-17437)
    const double r = k == 0 ? -G5A : G5A;
    const double s1 = -0.5 * (1.0 - r) * r, s2 = 0.5 * (1.0 + r) * r, s3 = (1.0 + r) * (1.0 - r);
    const double uf1 = s1 * d.usx[pa] + s2 * d.usx[pb] + s3 * d.usx[pm];
    const double uf2 = s1 * d.usy[pa] + s2 * d.usy[pb] + s3 * d.usy[pm];
    double v[7];
    face_state5(d, ie, g, v);
    const int nb = d.fnb[lf * nc + ie];
    double m, mu1, mu2, sn1, sn2;
    double du1, du2;
    if (nb < 0) {
        // boundary face: this side alone, its normal as the boundary face's
        // nodes orient it; walls clear the correction's normal components
        const int code = -nb - 1, bc = code & 3;
        const double sg = (code & 4) ? -1.0 : 1.0;
        face_riem5(d, v, uf1, uf2, sg * nx, sg * ny, ga, m, mu1, mu2, sn1, sn2);
        const double a11 = 0.0 + m, a22 = 0.0 + m, a21 = 0.0 + 0.0, a12 = 0.0 + 0.0;
        const double det = a11 * a22 - a21 * a12;
        const double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        const double r1 = (0.0 + mu1) - (0.0 + sn1) - 0.0, r2 = (0.0 + mu2) - (0.0 + sn2) - 0.0;
        du1 = -uf1 + (i11 * r1 + i12 * r2);
        du2 = -uf2 + (i21 * r1 + i22 * r2);
        if (bc == 1) du2 = 0.0;
        else if (bc == 2) du1 = 0.0;
        else if (bc == 3) du1 = du2 = 0.0;
    } else {
        // interior face: the neighbour's state at the same point (its face
        // runs the other way, so the point is its k' = 1 - k)
        const int je = nb >> 2, lr = nb & 3;
        double w[7];
        face_state5(d, je, 2 * lr + (1 - k), w);
        double mr, mr1, mr2, sr1, sr2;
        face_riem5(d, v, uf1, uf2, nx, ny, ga, m, mu1, mu2, sn1, sn2);
        face_riem5(d, w, uf1, uf2, -nx, -ny, ga, mr, mr1, mr2, sr1, sr2);
        const double a11 = 0.0 + m + mr, a22 = 0.0 + m + mr, a21 = 0.0 + 0.0 + 0.0, a12 = 0.0 + 0.0 + 0.0;
        const double det = a11 * a22 - a21 * a12;
        const double i11 = a22 / det, i12 = -a12 / det, i21 = -a21 / det, i22 = a11 / det;
        const double r1 = (0.0 + mu1 + mr1) - (0.0 + sn1 + sr1), r2 = (0.0 + mu2 + mr2) - (0.0 + sn2 + sr2);
        du1 = -uf1 + (i11 * r1 + i12 * r2);
        du2 = -uf2 + (i21 * r1 + i22 * r2);
    }
    // this side's force and correction force (:17479-17527)
    d.fg[FG5(g, 0) * nc + ie] = sn1 + m * uf1 + 0.0 * uf2 - mu1;
    d.fg[FG5(g, 1) * nc + ie] = sn2 + m * uf2 + 0.0 * uf1 - mu2;
    d.fg[FG5(g, 2) * nc + ie] = m * du1 + 0.0 * du2;
    d.fg[FG5(g, 3) * nc + ie] = m * du2 + 0.0 * du1;
    d.fg[FG5(g, 4) * nc + ie] = uf1;
    d.fg[FG5(g, 5) * nc + ie] = uf2;
}

}  // namespace lag2d
