// Tangent build of oracle/lag2d_oracle.c (see dual.hpp) and its entry point.
#include "dual.hpp"

// the plain (double) layout of the case, under other function names
namespace plain {
#define orc_cellvol orc_cellvol_plain_
#define orc_rhs orc_rhs_plain_
#define orc_step orc_step_plain_
#define orc_schedule orc_schedule_plain_
#define orc_rhs_q2 orc_rhs_q2_plain_
#include "../lag2d_oracle.h"
#undef orc_cellvol
#undef orc_rhs
#undef orc_step
#undef orc_schedule
#undef orc_rhs_q2
}  // namespace plain
#undef LAG2D_ORACLE_H

#define double Dual
#include "../lag2d_oracle.c"
#undef double

// J = sum over nodes |x - xt|^2 after nsteps steps from the case's state,
// and its derivative along the direction dunkno of the initial DOFs, with
// the limiter coefficients (straight path) or WENO increments (curved
// path), the corner impedances' |du.n| and dt frozen.
// The case's arrays are read, not modified.
extern "C" int orc_tangent_J(const plain::orc_case *p, const real_t *dunkno, int nsteps, const real_t *xt,
                             real_t *J, real_t *dJ)
{
    orc_case c;
    memset(&c, 0, sizeof(c));
    c.ncell = p->ncell; c.nquad = p->nquad; c.npoin = p->npoin; c.nbfac = p->nbfac;
    c.ndegr = p->ndegr; c.nq = p->nq; c.npoly = p->npoly; c.ncurv = p->ncurv; c.nvqua = p->nvqua;
    c.nmatr = p->nmatr; c.nstag = p->nstag; c.ndens = p->ndens; c.nlimi = p->nlimi; c.ncase = p->ncase;
    c.nrz = p->nrz; c.nfint = p->nfint; c.ndt = p->ndt;
    c.gamlg = p->gamlg; c.cfl = p->cfl; c.cimpd = p->cimpd; c.cdrho = p->cdrho; c.tend = p->tend;
    c.slpdu = p->slpdu; c.alfrz = p->alfrz; c.dtfix = p->dtfix; c.tacru = p->tacru; c.itime = p->itime;
    c.ntria = p->ntria; c.iptri = p->iptri; c.ipqua = p->ipqua; c.bface = p->bface;
    c.esuv1 = p->esuv1; c.esuv2 = p->esuv2; c.esqua = p->esqua; c.weno_strict = p->weno_strict;
    c.rkstg = p->rkstg; c.epsaw = p->epsaw;
    const size_t nu = (size_t)c.ndegr * c.nq * c.ncell, nx = 2 * (size_t)c.npoin;
    auto dup = [](const real_t *src, size_t n) {
        Dual *a = (Dual *)malloc(sizeof(Dual) * (n ? n : 1));
        for (size_t i = 0; i < n; ++i) a[i] = Dual(src ? src[i] : 0.0);
        return a;
    };
    c.coord = dup(p->coord, nx);
    c.cooro = dup(p->coord, nx);
    c.coold = dup(p->coold, nx);
    c.geoel = dup(p->geoel, 24 * (size_t)c.ncell);
    c.amatr = dup(p->amatr, (size_t)c.nmatr * c.ncell);
    c.unkno = dup(p->unkno, nu);
    for (size_t i = 0; i < nu; ++i) c.unkno[i].d = dunkno[i];
    c.unold = dup(nullptr, nu);
    memcpy(c.unold, c.unkno, sizeof(Dual) * nu);
    c.ustar = dup(p->ustar, nx);
    c.rhsel = dup(nullptr, nu);
    c.aflim = dup(nullptr, 5 * (size_t)c.ncell);
    c.afvec = dup(nullptr, 4 * (size_t)c.ncell);
    c.troubled = nullptr;
    Dual *rho0 = p->rho0 ? dup(p->rho0, (size_t)c.ncell) : nullptr;
    c.rho0 = rho0;
    int rc = 0;
    for (int n = 0; n < nsteps && !rc; ++n) {
        c.itime += 1;
        rc = orc_step(&c);
    }
    Dual j = 0.0;
    for (size_t i = 0; i < nx; ++i) j = j + (c.coord[i] - Dual(xt[i])) * (c.coord[i] - Dual(xt[i]));
    *J = j.v;
    *dJ = j.d;
    free(c.coord); free(c.cooro); free(c.coold); free(c.geoel); free(c.amatr);
    free(c.unkno); free(c.unold); free(c.ustar); free(c.rhsel); free(c.aflim); free(c.afvec);
    free(rho0);
    return rc;
}
