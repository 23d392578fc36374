/* CPU restatement of the reference's Lagrangian DG step (the oracle).
 *
 * TEST INFRASTRUCTURE ONLY.  This code is the checker the parity tests, the
 * smoke check and the bench's CPU-baseline leg compare the HIP path against;
 * the product never links or calls it.
 *
 * Layouts follow the reference's Fortran arrays (column-major), with 0-based
 * indices:
 *   unkno[id + ndegr*(iq + nq*ie)]      unkno(ndegr,nq,nsize)  dgflo2d.f90:121
 *   geoel[k + 24*ie]                    geoel(24,nsize)        dgflo2d.f90:119
 *   amatr[k + nmatr*ie]                 amatr(nmatr,ncell)     dgflo2d.f90:238-245
 *   coord[d + 2*ip]                     coord(2,npoin)         dgflo2d.f90:118
 *   iptri[iv + 3*it], ipqua[iv + nvqua*iq]   0-based nodes; cell it is
 *                          triangle it, cell ntria + iq is quad iq
 *   bface[k + 5*ifa]       nodes (k=0,1,4) 0-based; k=2 BC code, k=3 subcode
 *   esuv2[npoin+1] offsets, esuv1[] 0-based element ids   Preprocess.f90:720-766
 */
#ifndef LAG2D_ORACLE_H
#define LAG2D_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_case {
    /* sizes */
    int ncell, nquad, npoin, nbfac;
    int ndegr, nq, npoly, ncurv, nvqua, nmatr;
    /* switches (dgflo2d.ctrl, Preprocess.f90:28-55) */
    int nstag, ndens, nlimi, ncase, nrz, nfint, ndt;
    double gamlg, cfl, cimpd, cdrho, tend;
    double slpdu, alfrz;            /* dgflo2d.f90:147-156 */
    /* mutable driver scalars (module constant) */
    double dtfix, tacru;
    int itime;
    /* mesh: ntria triangles (cells 0..ntria-1) then the quads (cell ntria+q),
     * as the reference numbers them (Preprocess.f90:168-198) */
    int ntria;
    const int *iptri;               /* iptri[iv + 3*it], 0-based */
    const int *ipqua, *bface, *esuv1, *esuv2;
    /* state: coord = x^n, cooro = stage x, coold = x^0 */
    double *coord, *cooro, *coold;
    double *geoel, *amatr;
    double *unkno, *unold;          /* U^n and stage state, nsize cells */
    double *ustar;                  /* nodal velocity (2, npoin) */
    /* per-stage outputs kept for inspection */
    double *rhsel;                  /* (ndegr, nq, ncell) */
    double *aflim, *afvec;          /* (nq+1, ncell), (2,2,ncell) */
    /* curved P2 path */
    const int *esqua;               /* face neighbours [f + 4*ie]; >= ncell: boundary */
    int weno_strict;                /* 1: fail where the reference stops (WENO.f90:7699) */
    int nweno_skipped;              /* troubled cells skipped for their neighbourhood */
    int nweno_rest;                 /* troubled cells at rest (zero frame) left unchanged */
    double *troubled;               /* Persson flag of the last stage (0 / 100), or NULL */
    /* RZ area weighting (nrz=2): module scalars getsolvdg_lag sets per stage
     * and per step (Getrhslag.f90:1633-1692, :1745) */
    int rkstg;
    double epsaw;
    /* per-cell initial density in place of problem ncase's (the role of
     * getrhoig_quadcurv2, Getrhslag.f90:17761-17948, and getrhog_initial,
     * in the initial sub-cell masses and the WENO mass centres), or NULL */
    const double *rho0;
} orc_case;

/* getcellvol_lag (Limiterlag.f90:1869-2164), quads */
void orc_cellvol(orc_case *c, const double *x);
/* getrhsdg_lag_mcnew (Getrhslagmc.f90:6693-8731), one stage right-hand side
 * on the stage coordinates; limits c->unold in place like the reference. */
int orc_rhs(orc_case *c);
/* getsolvdg_lag (Getrhslag.f90:1553-2035), one full SSP-RK step followed by
 * the driver's copy-back (dgflo2d.f90:408-422). */
int orc_step(orc_case *c);
/* epsaw (and with set_dt, dtfix) of step c->itime, Getrhslag.f90:1663-1737 */
void orc_schedule(orc_case *c, int set_dt);
/* curved P2 sub-cell right-hand side (lag2d_oracle_q2.c) */
int orc_rhs_q2(orc_case *c);
/* Op counting (oracle/opcount/): the arithmetic after a phase marker is
 * attributed to that phase, named for the GPU kernel that does the same
 * work (WENO_MC: the physical mass centres the reference computes for every
 * cell, of which a troubled cell's stencil needs 9).  No-op in the ordinary
 * build. */
enum orc_phase {
    ORC_PH_OTHER = 0, ORC_PH_PREP, ORC_PH_WENO_MC, ORC_PH_WENO, ORC_PH_CORNER, ORC_PH_NODE,
    ORC_PH_CELLRHS, ORC_PH_UPDATE, ORC_PH_NODEMOVE, ORC_PH_CELLVOL, ORC_PH_DT,
    ORC_PH_HOISTED,  /* time-invariant work the GPU path does once: initial sub-cell masses */
    ORC_PH_COUNT
};
#ifndef ORC_PHASE
#define ORC_PHASE(k) ((void)0)
#endif
#ifndef ORC_WENO_CELL
#define ORC_WENO_CELL() ((void)0)
#endif
/* Tangent build (oracle/tangent/): marks the values the discrete adjoint
 * holds at their forward values -- the limiter coefficients, the corner
 * impedance's |du.n| and the time step -- whose derivative is dropped.  The
 * identity in the ordinary build. */
#ifndef ORC_FROZEN
#define ORC_FROZEN(x) (x)
#endif
/* Tangent build: the curved path's WENO output takes the derivative of its
 * input (the reconstruction's increment is held at its forward value).
 * Nothing in the ordinary build. */
#ifndef ORC_FROZEN_INCREMENT
#define ORC_FROZEN_INCREMENT(out, in) ((void)0)
#endif

#ifdef __cplusplus
}
#endif
#endif
