// Barth limiter of the straight P1 path, shared by the planar and the RZ
// kernels (the reference calls the same limiter before either branch,
// Getrhslagmc.f90:6823-6837).
#pragma once
#include "lag2d_dev.h"

namespace lag2d {

// Limiter coefficients of one cell with NV vertices (3: triangle, 4: quad):
// barthlimit_lagsym_quad Parts 2-3 (Limiterlag.f90:4-404) and
// barthlimit_lagsym_tria (Getrhslagmctria.f90:422-842) for tau, u, v, p and
// E, with the velocity limited in the cell's frame d.frame
// (barthlimit_sympre_quadlb Parts 4 and 6, Limiterlag.f90:8127-8692;
// barthlimit_sympre_trialb, Getrhslagmctria.f90:2644-3258).  un[iv][q] is
// the unlimited P1 polynomial at vertex iv, ip the cell's nodes.  Returns
// al = aflim(1..5) and af = afvec, column-major (1,1), (2,1), (1,2), (2,2).
template <int NV>
__device__ __forceinline__ void barth_limit(const Dev &d, const double *__restrict__ Uin, int ie, const int *ip,
                                            const double (*un)[4], double al[5], double af[4])
{
    const int nc = d.nc, nd = d.nd, np = d.np;
    const double g1 = d.gamlg - 1.0;
    const double tau = Uin[(0 * nd) * nc + ie], uctr = Uin[(1 * nd) * nc + ie];
    const double vctr = Uin[(2 * nd) * nc + ie], ectr = Uin[(3 * nd) * nc + ie];
    const double rhoct = 1.0 / tau;
    const double pctr_sp = dmax(EPS_SP6, g1 * rhoct * (ectr - 0.5 * (uctr * uctr + vctr * vctr)));
    {
        double unctr[4] = {1.0 / rhoct, uctr, vctr, pctr_sp};
        double a[4] = {1.e300, 1.e300, 1.e300, 1.e300};
        for (int iv = 0; iv < NV; ++iv) {
            double rhov = 1.0 / un[iv][0];
            double uv = un[iv][1], vv = un[iv][2], ev = un[iv][3];
            double pv = dmax(EPS_SP6, g1 * rhov * (ev - 0.5 * (uv * uv + vv * vv)));
            double vals[4] = {1.0 / rhov, uv, vv, pv};
            int p = ip[iv];
            for (int q = 0; q < 4; ++q) {
                double f = barthfct(d.nmm[q * np + p], d.nmm[(5 + q) * np + p], unctr[q], vals[q] - unctr[q]);
                a[q] = dmin(a[q], f);
            }
        }
        for (int q = 0; q < 4; ++q) al[q] = a[q];
    }
    {
        const double m0 = d.frame[0 * nc + ie], m1 = d.frame[1 * nc + ie];
        const double m2 = d.frame[2 * nc + ie], m3 = d.frame[3 * nc + ie];
        double uc2 = uctr * m0 + vctr * m1, uc3 = uctr * m2 + vctr * m3;
        double fixi = 1.e300, fiet = 1.e300;
        for (int iv = 0; iv < NV; ++iv) {
            double un2 = un[iv][1] * m0 + un[iv][2] * m1;
            double un3 = un[iv][1] * m2 + un[iv][2] * m3;
            double mx2 = uc2, mx3 = uc3, mn2 = uc2, mn3 = uc3;
            int p = ip[iv];
            int col = d.ncase != 1 ? d.ncol[p] : 0;
            for (int k = d.esuv2[p]; k < d.esuv2[p + 1]; ++k) {
                int je = d.esuv1[k];
                double ucj = d.cavg[1 * nc + je], vcj = d.cavg[2 * nc + je];
#define ACC(UU, VV)                                            \
    {                                                          \
        double ul = (UU) * m0 + (VV) * m1;                     \
        double vl = (UU) * m2 + (VV) * m3;                     \
        mx2 = dmax(mx2, ul); mx3 = dmax(mx3, vl);              \
        mn2 = dmin(mn2, ul); mn3 = dmin(mn3, vl);              \
    }
                ACC(ucj, vcj);
                if (col == 3) {
                    ACC(ucj, -vcj);
                } else if (col == 2) {
                    ACC(-ucj, vcj);
                } else if (col == 10) {
                    ACC(ucj, -vcj);
                    ACC(-ucj, vcj);
                    ACC(-ucj, -vcj);
                }
#undef ACC
            }
            fixi = dmin(fixi, barthfct(mx2, mn2, uc2, un2 - uc2));
            fiet = dmin(fiet, barthfct(mx3, mn3, uc3, un3 - uc3));
        }
        const double fi11 = m0 * m0 * fixi + m2 * m2 * fiet;
        const double fi12 = m0 * m1 * fixi + m2 * m3 * fiet;
        const double fi22 = m1 * m1 * fixi + m3 * m3 * fiet;
        af[0] = fi11; af[1] = fi12; af[2] = fi12; af[3] = fi22;
    }
    {
        double a5 = 1.e300;
        for (int iv = 0; iv < NV; ++iv) {
            double rhov = 1.0 / un[iv][0];
            double uv = un[iv][1], vv = un[iv][2];
            double pv = dmax(EPS_SP6, g1 * rhov * (un[iv][3] - 0.5 * (uv * uv + vv * vv)));
            double t = tau + al[0] * (1.0 / rhov - tau);
            double pl = pctr_sp + al[3] * (pv - pctr_sp);
            double e5 = pl / g1 * t + 0.5 * (uv * uv + vv * vv);
            int p = ip[iv];
            a5 = dmin(a5, barthfct(d.nmm[4 * np + p], d.nmm[9 * np + p], ectr, e5 - ectr));
        }
        al[4] = a5;
    }
}

}  // namespace lag2d
