// Kernel declarations shared between the kernel sources and the host driver.
#pragma once
#include "lag2d_dev.h"

namespace lag2d {

// straight-quad P1 path
__global__ void k_cellavg_q1(Dev d, const double *__restrict__ Uin);
__global__ void k_nodeminmax(Dev d);
template <int FROZEN>
__global__ void k_limit_corner_q1(Dev d, const double *__restrict__ Uin);
__global__ void k_node_q1(Dev d);
__global__ void k_cellrhs_q1(Dev d, const double *__restrict__ Uin, double alfa, int apply_update);

// common to both paths
__global__ void k_dt_partial(Dev d, const double *__restrict__ Uin, int nbpr);
__global__ void k_dt_final(Dev d, int itime, int nbpr);
__global__ void k_run_flags(Dev d);
__global__ void k_division_check(const double *a, const double *b, double *q, int n);
__global__ void k_update(Dev d, double alfa);
__global__ void k_update_cells(Dev d, double alfa);
__global__ void k_nodemove(Dev d, double alfa, const double *__restrict__ xsrc, const double *__restrict__ ysrc);
template <int RZ>
__global__ void k_cellvol(Dev d, double *__restrict__ x, double *__restrict__ y, int set_centre);
template <int RZ>
__global__ void k_cellvol_tpc(Dev d, double *__restrict__ x, double *__restrict__ y, int set_centre);
__global__ void k_init_q2(Dev d, const double *__restrict__ in, double *A);
template <int RZ>
__global__ void k_init_p1(Dev d, const double *__restrict__ in, double *A);

// area-weighted RZ path (nrz=2), straight P1 cells
__global__ void k_limit_corner_rz(Dev d, const double *__restrict__ Uin, int rev);
__global__ void k_node_rz(Dev d);
__global__ void k_cellrhs_rz(Dev d, const double *__restrict__ Uin, double alfa, int apply_update, double epsaw);
__global__ void k_schedule(Dev d, double dt);

// curved P2 sub-cell path, nfint=5: the face Gauss points' Riemann problems
__global__ void k_q2g_face(Dev d);
// curved P2 sub-cell path, six-node triangles of hybrid meshes
__global__ void k_t2_submass(Dev d);
__global__ void k_t2_corner(Dev d);
__global__ void k_t2_cellrhs(Dev d, double alfa, int apply_update);

// discrete adjoint of the straight P1 planar step (quads)
__global__ void k_adj_nodemove(Dev d, AdjStage a, double alfa);
__global__ void k_adj_cellrhs(Dev d, AdjStage a, double alfa);
__global__ void k_adj_node(Dev d, AdjStage a);
__global__ void k_adj_corner(Dev d, AdjStage a);
__global__ void k_adj_gather(Dev d, AdjStage a);
__global__ void k_adj_step_end(Dev d, AdjStage a);
__global__ void k_replay_dt(Dev d, const double *dt);
constexpr int CV_CELLS = 32;  // cells per block of k_cellvol
constexpr int CV_GROUP = 6;    // points scheduled together in k_cellvol_tpc
constexpr int Q2_REC = 80;    // curved path: sub-cell corner-state doubles per cell (16 corners x 5)
constexpr int GEOM_CELLS_MIN = 1 << 16;  // cells from which cell volume and prep run a thread per cell

}  // namespace lag2d
