/* Shared helpers of the C oracle (test infrastructure only). */
#ifndef ORC_UTIL_H
#define ORC_UTIL_H
#include <math.h>
#include "lag2d_oracle.h"

#define U(c, a, id, iq, ie) (a)[(id) + (c)->ndegr * ((iq) + (c)->nq * (ie))]
#define GE(c, k, ie) (c)->geoel[(k) + 24 * (ie)]
#define X(a, d, ip) (a)[(d) + 2 * (ip)]

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
/* Fortran default-real literal 1.e-6 widened to REAL(8) */
static const double EPS_SP6 = (double)1.e-6f;

/* ------------------------------------------------------------------------ */
/* Quadrature: ruqope (dgflo2d.f90:2646-2783), 2-D tensor rules, ngaus = n*n */
static inline void ruqope2(int n, double *pos /* [2*n*n] */, double *w /* [n*n] */)
{
    double pg[6], wg[6];
    if (n == 5) {
        pg[2] = 0.0;
        pg[3] = 0.5384693101056830910363144;
        pg[1] = -0.5384693101056830910363144;
        pg[4] = 0.9061798459386639927976269;
        pg[0] = -0.9061798459386639927976269;
        wg[2] = 0.5688888888888888888888889;
        wg[3] = 0.4786286704993664680412915;
        wg[1] = 0.4786286704993664680412915;
        wg[4] = 0.2369268850561890875142640;
        wg[0] = 0.2369268850561890875142640;
    } else if (n == 6) {
        pg[0] = 0.2386191860831969086305017;
        pg[1] = -0.2386191860831969086305017;
        pg[2] = 0.6612093864662645136613996;
        pg[3] = -0.6612093864662645136613996;
        pg[4] = -0.9324695142031520278123016;
        pg[5] = 0.9324695142031520278123016;
        wg[0] = 0.4679139345726910;
        wg[1] = 0.4679139345726910;
        wg[2] = 0.3607615730481386;
        wg[3] = 0.3607615730481386;
        wg[4] = 0.1713244923791704;
        wg[5] = 0.1713244923791704;
    } else { /* n == 2 */
        pg[0] = -0.5773502691896257645091488;
        pg[1] = 0.5773502691896257645091488;
        wg[0] = 1.0;
        wg[1] = 1.0;
    }
    int g = 0;
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            w[g] = wg[i] * wg[j];
            pos[2 * g] = pg[i];
            pos[2 * g + 1] = pg[j];
            ++g;
        }
}

/* Q2 shape functions and reference derivatives (Limiterlag.f90:2053-2091) */
static inline void shape_q2(double r, double s, double *sh, double *dr_, double *ds_)
{
    const double c10 = 1.0;
    if (sh) {
        sh[0] = 0.25 * (c10 - r) * r * s * (c10 - s);
        sh[1] = -0.25 * (c10 + r) * (c10 - s) * r * s;
        sh[2] = 0.25 * (c10 + r) * (c10 + s) * r * s;
        sh[3] = -0.25 * (c10 - r) * (c10 + s) * r * s;
        sh[4] = -0.5 * (c10 - r * r) * (c10 - s) * s;
        sh[5] = 0.5 * (c10 + r) * (c10 - s * s) * r;
        sh[6] = 0.5 * (c10 - r * r) * (c10 + s) * s;
        sh[7] = -0.5 * (c10 - r) * (c10 - s * s) * r;
        sh[8] = (c10 - r * r) * (c10 - s * s);
    }
    dr_[0] = 0.25 * (2.0 * r - c10) * s * (s - 1.0);
    dr_[1] = 0.25 * (2.0 * r + c10) * s * (s - 1.0);
    dr_[2] = 0.25 * (2.0 * r + c10) * s * (s + 1.0);
    dr_[3] = 0.25 * (2.0 * r - c10) * s * (s + 1.0);
    dr_[4] = -r * s * (s - c10);
    dr_[5] = 0.5 * (2.0 * r + c10) * (c10 - s * s);
    dr_[6] = -r * s * (s + c10);
    dr_[7] = 0.5 * (2.0 * r - c10) * (1.0 - s * s);
    dr_[8] = -2.0 * r * (1.0 - s * s);
    ds_[0] = 0.25 * (2.0 * s - c10) * r * (r - 1.0);
    ds_[1] = 0.25 * (2.0 * s - c10) * r * (r + 1.0);
    ds_[2] = 0.25 * (2.0 * s + c10) * r * (r + 1.0);
    ds_[3] = 0.25 * (2.0 * s + c10) * r * (r - 1.0);
    ds_[4] = 0.5 * (1.0 - r * r) * (2.0 * s - c10);
    ds_[5] = -r * s * (r + c10);
    ds_[6] = 0.5 * (1.0 - r * r) * (2.0 * s + c10);
    ds_[7] = -r * s * (r - c10);
    ds_[8] = -2.0 * s * (1.0 - r * r);
}


#endif
