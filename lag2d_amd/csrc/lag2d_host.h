// Host-side context of the lag2d library.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include "../../include/lag2d.h"
#include "lag2d_dev.h"

constexpr int LAG2D_CORNER_VARIANTS = 9, LAG2D_RHS_VARIANTS = 4;

enum {
    KID_CELLAVG_Q1 = 0, KID_NODEMINMAX, KID_LIMIT_Q1, KID_NODE_Q1, KID_CELLRHS_Q1, KID_DT_PARTIAL,
    KID_DT_FINAL, KID_UPDATE, KID_NODEMOVE, KID_CELLVOL,
    KID_Q2_PREP, KID_Q2_WENO, KID_Q2_CORNER, KID_NODE_Q2, KID_CELLRHS_Q2, KID_RUN_FLAGS,
    KID_LIMIT_RZ, KID_NODE_RZ, KID_CELLRHS_RZ, KID_SCHEDULE,
    KID_ADJ_CELLRHS, KID_ADJ_CORNER, KID_ADJ_NODES, KID_LIMIT_FROZEN, KID_Q2G_FACE, KID_COUNT
};
extern const char *const lag2d_kernel_names[KID_COUNT];

struct lag2d_timer {
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
};

// Recording and buffers of the discrete adjoint (straight P1 planar quads,
// and the curved sub-cell path).
struct lag2d_adj {
    int nsteps = 0, every = 0, itime0 = 0, nckpt = 0;
    bool curved = false;
    size_t cap_steps = 0, cap_ckpt = 0, cap_seg = 0;  // allocated sizes
    double *lim_tape = nullptr;  // straight: [nsteps][2][17*nc] limiter coefficients and |du.n| of every
                                 // stage; curved: [nsteps][nstag][(8 nd + 32) nc] WENO's limited and
                                 // input DOFs and |du.n|
    uint8_t *clamp_tape = nullptr;  // curved: [nsteps][nstag][44 nc] clamp branches, limited flags
    double *dt_tape = nullptr;   // [nsteps]
    double *ck = nullptr;        // [nckpt][12*nc + 4*np] checkpoints: U, x, y, u*
    double *seg = nullptr;       // [every][seg_stride] stage data of a recomputed segment
    size_t seg_stride = 0;
    double *buf = nullptr;       // adjoint buffers
    const double *lam_x0 = nullptr;  // dJ/dx at the recording's start, after lag2d_adj_gradient [2][np]
};

// A captured step: its launches depend on the parity of the state-buffer
// swaps and, where the problem fixes dt, on the step's dt and epsaw.
struct lag2d_graph {
    int parity = 0, sched = 0;
    double dt = 0.0, epsaw = 0.0;
    hipGraphExec_t exec = nullptr;
};

struct lag2d_ctx {
    lag2d_config cfg{};
    int nc = 0, np = 0, nb = 0, nd = 0, nvq = 0, nmatr = 0;
    int ntria = 0;          // triangles first in the cell numbering (straight P1 path)
    int nrec = 40;          // corner-face record doubles per cell
    int nblk_dt = 1;        // blocks per run of the dt reduction
    int nrun = 1;           // runs of an ensemble
    // curved-path kernel variants (launch_corner, launch_cellrhs in
    // lag2d_q2.hip): 0 = not tuned yet; corner 1 = lanes per cell, 2-6 = a
    // lane per sub-cell (cells per workgroup, LDS staging), 7-8 = a wave per
    // sub-cell (registers for 2 or 3 waves per SIMD), 9 = two waves per
    // sub-cell; cell RHS 1 = 8
    // cells per workgroup, 2-4 = a thread per cell (workgroup size)
    int corner_variant = 0;
    float corner_ms[LAG2D_CORNER_VARIANTS] = {};  // tuning timings
    int rhs_variant = 0;
    float rhs_ms[LAG2D_RHS_VARIANTS] = {};
    bool update_cells = false;  // k_update_cells (large meshes) instead of k_update
    bool geom_cells = false;    // k_cellvol_tpc, k_q2_prep_tpc (large meshes) instead of 8 lanes per cell
    std::vector<double> tend_host;
    hipStream_t stream = nullptr;
    lag2d::Dev d{};
    lag2d_adj adj;
    int parity = 0;                   // state buffers swapped an odd number of times
    bool use_graphs = false;          // replay captured steps (lag2d_use_graphs)
    std::vector<lag2d_graph> graphs;
    double *A = nullptr;    // inverse mass matrix (SoA)
    double *x0 = nullptr;   // initial coordinates [2][np]
    std::vector<void *> allocs;
    std::vector<int> ipq_host;
    std::vector<double> geo_host;
    bool timing = false;
    lag2d_timer timers[KID_COUNT];
    double kernel_total_ms[KID_COUNT] = {};
    int kernel_calls[KID_COUNT] = {};

    ~lag2d_ctx();
    template <class T>
    int alloc(T **p, size_t n);
    void tic(int k);
    void toc(int k);
};

// Node -> corner CSR: for every node the (cell, corner) pairs in the order
// the reference accumulates their contributions.
int build_node_corners(lag2d_ctx &x, const std::vector<int> &ipq, const std::vector<int> &esuv1,
                       const std::vector<int> &esuv2, std::vector<int> &n2c, std::vector<int> &n2c_off);
// Extra per-path data computed once the state is uploaded (e.g. the
// time-invariant initial sub-cell masses of the curved path).
int upload_extra(lag2d_ctx &x);
int launch_rhs_q2(lag2d_ctx *x, lag2d::Dev &ds, const double *Uin, double alfa, int apply_update);
// The curved path's stage RHS in the adjoint's recorded (record = 1) or
// replayed (record = 0) forward: WENO increment, |du.n| and clamp branches
// written to or read from the tapes ds.wdu, ds.dlt, ds.ccl, ds.dcl.
int launch_rhs_q2_adj(lag2d_ctx *x, lag2d::Dev &ds, const double *Uin, double alfa, int apply_update, int record);
// The reverse of one curved-path stage (lag2d_adjq2.hip): a.LU, a.LX hold
// the adjoint of the stage's output; on return a.LUI, a.LXI hold that of its
// input and a.LUN, a.LXN have gained the step input's direct part.
int launch_adjq2_stage(lag2d_ctx *x, lag2d::AdjStageC &a, double alfa);
int launch_adjq2_step_end(lag2d_ctx *x, lag2d::AdjStageC &a);
