/* lag2d: MI355X (gfx950) implementation of the Lagrangian DG right-hand side
 * and SSP-RK step of liuspace/Lag2D (srccurv-lag), behind a plain C ABI.
 *
 * The reference has no library boundary of its own: its time loop calls
 *   getsolvdg_lag(...)        once per step   (dgflo2d.f90:372-373)
 * which calls
 *   getrhsdg_lag_mcnew(...)   once per stage  (Getrhslag.f90:1747-1750).
 * The entry points below replace those two subroutines.  A Fortran caller
 * binds them with ISO_C_BINDING (lag2d_amd/fortran/lag2d_c.f90; a drop-in
 * getsolvdg_lag is lag2d_amd/fortran/getsolvdg_lag_shim.f90); arrays are
 * the reference's own column-major arrays, connectivity 1-based, sizes and
 * switches those of module `constant` (dgflo2d.f90:4-57).
 *
 * Every entry returns 0 on success or one of the LAG2D_E* codes; the message
 * of the last failure on the calling thread is lag2d_last_error().
 */
#ifndef LAG2D_H
#define LAG2D_H
#ifdef __cplusplus
extern "C" {
#endif

enum {
    LAG2D_OK = 0,
    LAG2D_EINVAL = 1,        /* bad argument or size */
    LAG2D_EUNSUPPORTED = 2,  /* switch combination not implemented */
    LAG2D_EDEVICE = 3,       /* HIP runtime failure */
    LAG2D_ENONFINITE = 4,    /* non-finite state detected by lag2d_check */
    LAG2D_EVOLUME = 5,       /* non-positive cell or specific volume detected by lag2d_check */
    LAG2D_ETOPOLOGY = 6      /* WENO stencil topology the reference stops on (strict) */
};

/* Switches and driver scalars of module `constant` (dgflo2d.f90:4-57) that the
 * Lagrangian step reads.  dtfix and tacru are in/out like the module globals
 * getsolvdg_lag mutates (Getrhslag.f90:1663-1785, :20705-20709). */
typedef struct lag2d_config {
    int ndegr;    /* 3 (P1) or 6 (P2) */
    int npoly;    /* 1 or 2 */
    int ncurv;    /* 0 straight quads (4 nodes), 1 quadratic quads (9 nodes) */
    int nstag;    /* 2: SSP-RK2, 3: SSP-RK3 */
    int ndens;    /* 1: specific volume evolved as a DG variable */
    int nlimi;    /* 6 */
    int ncase;    /* test-problem id (only boundary-colouring rules read it) */
    int nrz;      /* 0 planar, 2 area-weighted RZ (r = y; straight P1 only) */
    int nfint;    /* 1 (ncurv=0); 4 (ncurv=1 sub-cell scheme) or 5 (ncurv=1 sub-cell scheme
                     with face Gauss points, P2, the 5-point Lobatto face rule) */
    int nriem;    /* 1 */
    int ndt;      /* 1: CFL-controlled dt, 0: fixed dtfix */
    double gamlg, cfl, cimpd, cdrho, tend;
    double dtfix, tacru;
    /* 1: report troubled cells whose vertex neighbourhood the reference's
     * WENO stencil builder cannot classify (it stops there, WENO.f90:7699)
     * as LAG2D_ETOPOLOGY; 0: treat them like boundary stencils (skip). */
    int strict;
} lag2d_config;

typedef struct lag2d_ctx lag2d_ctx;

/* Create a context and upload the mesh (Preprocess.f90 outputs):
 *   ipqua(nvqua, nquad)  quad nodes, 1-based, nvqua = 4 or 9
 *   bface(5, nbfac)      n1, n2, BC code, BC sub-code, mid node (1-based)
 *   esuv1(esuv2(npoin+1)), esuv2(npoin+1)  elements around points (CSR)
 * Quads only (ntria = 0; lag2d_create_hybrid takes triangles too). */
int lag2d_create(const lag2d_config *cfg, int nquad, int npoin, int nbfac,
                 const int *ipqua, const int *bface, const int *esuv1,
                 const int *esuv2, lag2d_ctx **ctx);

/* Hybrid mesh of triangles and quads (straight P1 path, ncurv=0, nfint=1):
 * the reference numbers triangles first, so cell ie < ntria is triangle ie
 * (iptri(3, ntria), counter-clockwise, 1-based) and cell ntria + q is quad q
 * (Preprocess.f90:168-198); esuv1 holds those cell ids.  Each cell type runs
 * the reference's routines for it (Getrhslagmc.f90:6810-6939: the triangle
 * limiter, Riemann coefficients, consistent-weight face integral and
 * 7-point domain integral of Getrhslagmctria.f90 / Getrhslagmc.f90:14006-14160).
 * The problems ncase = 1, 2, 3, -7, whose triangle limiter overrides
 * boundary nodes (Getrhslagmctria.f90:539-582), return LAG2D_EUNSUPPORTED.
 * On the curved P2 path (ncurv=1, nfint=4, npoly=2, ncase 7) the triangles
 * are six-node curved triangles, iptri(6, ntria): corners, then the
 * mid-nodes of edges 1-2, 2-3, 3-1 (Getrhslaggauss.f90:12201-12590,
 * :8297-8496, Getrhslag.f90:4449-4711); esuv2(npoin+1) = 6 ntria + 9 nquad.
 * With cfg->nrz = 2 the cells take the area-weighted RZ routines instead
 * (Getrhslagrz.f90, Getrhslagmc.f90:6876-6971, :7299-7392); lag2d_create
 * accepts nrz = 2 on quad meshes as well.  RZ boundaries are slip walls
 * (code 22) or free (21). */
int lag2d_create_hybrid(const lag2d_config *cfg, int ntria, int nquad, int npoin, int nbfac,
                        const int *iptri, const int *ipqua, const int *bface, const int *esuv1,
                        const int *esuv2, lag2d_ctx **ctx);

/* Ensemble of independent runs (batched mode; curved P2 path only).  No
 * counterpart in the reference, which runs one problem per process: the
 * runs' meshes are concatenated in run order into one mesh -- the cells,
 * nodes and boundary faces of run r follow those of run r-1, connectivity
 * holds global 1-based ids, esuv1/esuv2 the CSR of the whole mesh -- and
 * every kernel launch of a stage covers all runs.  Run r owns cells
 * [cell_off[r], cell_off[r+1]) and nodes [node_off[r], node_off[r+1]); a
 * cell may use only nodes of its own run.  Each run keeps its own dt
 * (CFL-limited over its own cells, as getdelt_lag does for one problem),
 * time tacru[r] and end time tend[r]; a run that has reached its end time
 * is left unchanged by later steps.  cfg holds the switches shared by all
 * runs; its tend, dtfix and tacru are not used.  Upload, download and check
 * work on the concatenated arrays. */
int lag2d_create_ensemble(const lag2d_config *cfg, int nrun, const int *cell_off,
                          const int *node_off, const double *tend, const double *dtfix,
                          const double *tacru, int nquad, int npoin, int nbfac,
                          const int *ipqua, const int *bface, const int *esuv1,
                          const int *esuv2, lag2d_ctx **ctx);

/* Upload the solution state (reference layouts, nsize >= ncell):
 *   coord(2,npoin) x^n, coold(2,npoin) x^0, geoel(24,ncell), amatr(nmatr,ncell),
 *   unkno(ndegr,4,ncell), ustar(2,npoin).  Recomputes geoel(3,5,6) from coord
 *   as getcellvol_lag does (Limiterlag.f90:1869). */
int lag2d_upload(lag2d_ctx *ctx, const double *coord, const double *coold,
                 const double *geoel, const double *amatr, const double *unkno,
                 const double *ustar);

/* Driver-side initialisation on the device, in place of lag2d_upload, for
 * the planar curved P2 path on quads (ncurv=1, npoly=2, nrz=0) and the
 * straight P1 path (ncurv=0, npoly=1; planar or nrz=2, triangles first).
 * Replaces the start-up sequence of dgflo2d.f90:170-282 for piecewise
 * constant initial states, with the reference's loops and summation orders:
 * getgeoel_initial_lag (GetInitial.f90:1278-1697: mass, volume, mass and
 * volume centres on the 6x6 quad and 7-point triangle rules; RZ the area
 * mass, area centre, mean radius and area), for P2 getgeoelho_initial_lag
 * (:1917-2096: second moments) and getunkno_initial_lag (:151-693: the
 * degree >= 1 mass matrix on the 5x5 and 7-point rules and its inverse, by
 * getinvmat's LU, dgflo2d.f90:2792-2908, for P2 and by the 2x2 cofactors
 * for P1, RZ divided by the mean radius).  In: coord(2,npoin) (also taken
 * as coold; ustar = 0), rho, u, v, p per cell (ncell each), and nblast
 * cells (1-based) whose total energy is energy[k] (E = energy / mass,
 * Sedov).  Sets geoel, amatr and unkno (tau = volume / mass on the mass
 * matrix's rule, u, v, E = p / ((gamma-1) rho) + |u|^2/2 for P2,
 * p tau / (gamma-1) + |u|^2/2 for P1, higher DOFs 0) and, for P2, the
 * initial density of the sub-cell masses; download them with
 * lag2d_download and lag2d_get_mass_matrix.  Synchronises. */
int lag2d_init_state(lag2d_ctx *ctx, const double *coord, const double *rho, const double *u,
                     const double *v, const double *p, int nblast, const int *blast,
                     const double *energy);

/* Initial density of every cell (ncell values), for initial states other
 * than the reference's built-in problems.  No counterpart in the reference:
 * its initial sub-cell masses (getrhosubg_averg, Getrhslaggauss.f90:7293-7374)
 * and WENO mass centres (WENO.f90:5989-6061) evaluate the initial density
 * of problem ncase (getrhoig_quadcurv2, Getrhslag.f90:17761-17948); after
 * this call they use the given values.  Curved P2 path; call after
 * lag2d_upload and before the first step. */
int lag2d_set_initial_density(lag2d_ctx *ctx, const double *rho);

/* One SSP-RK time step: getsolvdg_lag (Getrhslag.f90:1553-2035) followed by
 * the driver's copy of the stage state into unkno/coord
 * (dgflo2d.f90:408-422).  itime is the driver's 1-based step counter.
 * Asynchronous on the context's stream, except with cfg->strict on the
 * curved path: that step synchronises and returns LAG2D_ETOPOLOGY if any
 * troubled cell met a WENO neighbourhood the reference stops on (the flag is
 * sticky: later steps of the context keep failing). */
int lag2d_step(lag2d_ctx *ctx, int itime);

/* nsteps steps (itime0, itime0+1, ...), stopping after the step that reaches
 * tend, as the driver does (dgflo2d.f90:458-463); an ensemble stops when
 * every run has reached its end time.  Returns the number taken
 * in *taken (synchronises). */
int lag2d_run(lag2d_ctx *ctx, int itime0, int nsteps, int *taken);

/* One right-hand side: getrhsdg_lag_mcnew (Getrhslagmc.f90:6693-8731) on the
 * current state x^n / unkno, with ustar as the previous nodal velocity.
 * Writes rhsel(ndegr,4,ncell) and updates the resident ustar and the limited
 * DOFs (the reference limits its unkno argument in place,
 * Getrhslagmc.f90:7283-7296).  Synchronises.  Strict mode reports the WENO
 * topology stop as lag2d_step does. */
int lag2d_rhs(lag2d_ctx *ctx, double *rhsel);

/* Download the state (same layouts as lag2d_upload); NULL pointers skip.
 * geoel receives rows 3, 5, 6 (current volume and volume centre) updated and
 * the uploaded values elsewhere; on the RZ path (nrz=2) also rows 1, 2
 * (mass centre), 11, 12 (mean radius, area), 13, 14 (predicted mass centre)
 * and 15, 16 (area centre), which it updates every stage. */
int lag2d_download(lag2d_ctx *ctx, double *coord, double *geoel, double *unkno,
                   double *ustar);

/* Driver scalars after the last step (module constant dtfix, tacru); of
 * run 0 for an ensemble. */
int lag2d_get_time(lag2d_ctx *ctx, double *dtfix, double *tacru);

/* The resident inverse mass matrix amatr(nmatr, ncell), as lag2d_upload takes
 * it.  The RZ path recomputes it every stage (getamatr_lagmc_rzawrev,
 * Getrhslagrz.f90:3633-4221, writing the amatr argument of getsolvdg_lag,
 * Getrhslag.f90:1747); the planar paths leave it as uploaded. */
int lag2d_get_mass_matrix(lag2d_ctx *ctx, double *amatr);

/* dt and time of every run of an ensemble (arrays of nrun; one entry for a
 * single problem).  No counterpart in the reference. */
int lag2d_get_run_times(lag2d_ctx *ctx, double *dtfix, double *tacru);

/* Scan the resident state: a non-finite DOF returns LAG2D_ENONFINITE; a
 * non-positive cell volume (geoel(3), e.g. a tangled cell) or a non-positive
 * cell-average specific volume tau returns LAG2D_EVOLUME; the message names
 * the first such cell (1-based).  The reference has no such check and would
 * propagate them silently.  Synchronises. */
int lag2d_check(lag2d_ctx *ctx);

/* Wait for all work queued on the context's stream. */
int lag2d_sync(lag2d_ctx *ctx);

/* Kernel timing: enable > 0 starts timing every kernel launch with HIP
 * events on the context's stream (and clears the WENO cell counter), 0 stops;
 * enable < 0 reports the average device time (ms) of the dominant kernel
 * over the timed launches, and its name.  Returns LAG2D_EINVAL if no timed
 * run happened. */
int lag2d_kernel_time(lag2d_ctx *ctx, int enable, double *avg_ms, const char **name);

/* Per-kernel totals of the last timed run: kernel index k = 0, 1, ...;
 * returns LAG2D_EINVAL past the last kernel. */
int lag2d_kernel_stats(lag2d_ctx *ctx, int k, double *total_ms, int *calls, const char **name);

/* Troubled cells the WENO kernel processed since kernel timing was last
 * enabled (the unit of that kernel's work; 0 on the straight path).  No
 * counterpart in the reference. */
int lag2d_weno_cells(lag2d_ctx *ctx, long long *total);

/* Self-test of the library's division by a shared divisor: q[i] = a[i] / b[i]
 * computed on the device by the reciprocal-and-correction sequence the
 * kernels use where many quotients share one divisor; it must equal the IEEE
 * quotient bit for bit.  No context needed; synchronises.  No counterpart in
 * the reference. */
int lag2d_division_check(int n, const double *a, const double *b, double *q);

/* lag2d_division_check without any device allocation: no hipMalloc, hipFree
 * or allocating library call.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid allocation).  The device copies
 * of a, b and q live in a static device array of 3 * LAG2D_NOALLOC_MAX_N
 * doubles (1.5 MiB) that each GPU the library runs on holds until the
 * process ends; n > LAG2D_NOALLOC_MAX_N returns LAG2D_EINVAL before the
 * device is touched.  Results equal lag2d_division_check's.  Every call on a
 * GPU shares that GPU's array: calls that overlap on one GPU (from several
 * host threads) race and must be serialised by the caller.  Works on the
 * current HIP device; synchronises. */
#define LAG2D_NOALLOC_MAX_N 65536
int lag2d_division_check_noalloc(int n, const double *a, const double *b, double *q);

/* With enable != 0, lag2d_step and lag2d_run replay each step from a HIP
 * graph captured from the step's launches (one per buffer parity and fixed
 * dt and epsaw), cutting the host's per-step launch cost; the results are
 * the eager step's.  The GPU's own gaps between dependent kernels are not
 * shortened, so small meshes bound by them gain nothing.  The first step of a run, steps with kernel timing on and
 * curved-path steps before its kernel variants are tuned run eagerly.  No
 * counterpart in the reference. */
int lag2d_use_graphs(lag2d_ctx *ctx, int enable);

/* Discrete adjoint of the step.  No counterpart in the reference.  Two
 * paths: the straight P1 planar quads with SSP-RK2 (ncurv=0, nrz=0), whose
 * limiter coefficients are frozen, and the curved sub-cell path (ncurv=1,
 * nfint=4, P1 or P2, Q2 quads, SSP-RK2 or RK3), whose WENO increment
 * U_lim - U_in per cell is frozen.  Both also freeze the corner impedances'
 * |du.n| (the shock switch of the Riemann coefficients, Limiterlag.f90:987-996,
 * Getrhslaggauss.f90:10810-10811), the branches of the max(eps, .) pressure and
 * sound-speed clamps and the time step, so the gradient is that of the map
 * the recorded forward computes with them held fixed.
 *
 * lag2d_adj_record: nsteps steps (itime0, itime0+1, ...) from the resident
 * state, as lag2d_step, recording the frozen switches of every stage and
 * every step's dt, and checkpointing U, x and u* every `every` steps.  The
 * recorded forward is lag2d_step's, bit for bit.  Device memory, straight:
 * 34*ncell doubles per step, a checkpoint of 12*ncell + 4*npoin doubles per
 * `every` steps, the stage data of `every` steps (104*ncell + 10*npoin
 * doubles each); curved (nd = ndegr, ns = nstag): ns*((8 nd + 32)*ncell
 * doubles + 44*ncell bytes) per step, a checkpoint of 4 nd*ncell + 4*npoin
 * doubles per `every` steps, the stage data of `every` steps
 * (ns*((4 nd + 128)*ncell + 4*npoin) doubles each, ncell rounded up to a
 * multiple of 64 for the 80 corner-state doubles).
 *
 * lag2d_adj_replay: nsteps steps from the resident state with the recorded
 * switches and dt (upload a perturbed initial state first): the frozen
 * forward map whose derivative the adjoint computes.  On the curved path a
 * limited cell's DOFs are U_lim,rec + (U_in - U_in,rec), the recording's
 * limited DOFs at the recorded input, so a replay from the recorded initial
 * state reproduces the recording bit for bit.
 *
 * lag2d_adj_gradient: given lam_coord(2,npoin) = dJ/dx at the end of the
 * recorded steps, returns grad(ndegr,4,ncell) = dJ/dunkno at their start.
 * Each checkpointed segment is recomputed forward (stage data kept), then
 * reversed step by step.  Leaves the resident state unspecified: upload
 * before stepping again.
 *
 * lag2d_adj_gradient_coord: after lag2d_adj_gradient, grad_coord(2,npoin) =
 * dJ/dx at the start of the recorded steps (the initial node positions). */
int lag2d_adj_record(lag2d_ctx *ctx, int itime0, int nsteps, int every);
int lag2d_adj_replay(lag2d_ctx *ctx, int nsteps);
int lag2d_adj_gradient(lag2d_ctx *ctx, const double *lam_coord, double *grad);
int lag2d_adj_gradient_coord(lag2d_ctx *ctx, double *grad_coord);

void lag2d_destroy(lag2d_ctx *ctx);
const char *lag2d_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
