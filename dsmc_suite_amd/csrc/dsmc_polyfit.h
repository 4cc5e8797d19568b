/*
 * dsmc_polyfit.h - least-squares polynomial fits, written once for the device
 * fit kernels (one lane per grid cell) and for the host export of polyfit.
 *
 * The arithmetic follows src/common/polyfit.c:10-101 operation for operation,
 * because the fits must equal the reference's bit for bit: power sums and
 * right-hand side accumulated point by point, the normal matrix built from the
 * power sums, Gaussian elimination with partial pivoting (largest magnitude
 * below the diagonal, first on ties, a pivot under 1e-30 is singular), back
 * substitution, coefficients returned highest power first.
 */
#ifndef DSMC_POLYFIT_H
#define DSMC_POLYFIT_H

#include "dsmc_lane.h"

enum { DSMC_POLY_MAX_DEG = 6 };

/* Degree-`deg` least squares through (x[i], y[i]), i < n; out[0] is the
 * coefficient of x^deg.  Returns 0, or -1 for fewer points than coefficients,
 * a degree outside 0..6 (the reference has room for degree 6 at most) or a
 * singular normal matrix. */
DSMC_HD int lsq_poly(const double *x, const double *y, int n, int deg, double *out)
{
    const int m = deg + 1;
    if (deg < 0 || deg > DSMC_POLY_MAX_DEG || n < m) return -1;
    double psum[2 * DSMC_POLY_MAX_DEG + 1];          /* sum over points of x^k */
    double rhs[DSMC_POLY_MAX_DEG + 1];               /* sum over points of x^k y */
    for (int k = 0; k <= 2 * DSMC_POLY_MAX_DEG; k++) psum[k] = 0.0;
    for (int k = 0; k <= DSMC_POLY_MAX_DEG; k++) rhs[k] = 0.0;
    for (int i = 0; i < n; i++) {
        double t = 1.0;
        for (int k = 0; k <= 2 * deg; k++) { psum[k] += t; t *= x[i]; }
    }
    for (int i = 0; i < n; i++) {
        double t = 1.0;
        for (int k = 0; k < m; k++) { rhs[k] += t * y[i]; t *= x[i]; }
    }
    double sys[DSMC_POLY_MAX_DEG + 1][DSMC_POLY_MAX_DEG + 2];   /* [normal matrix | rhs] */
    for (int r = 0; r < m; r++) {
        for (int c = 0; c < m; c++) sys[r][c] = psum[r + c];
        sys[r][m] = rhs[r];
    }
    for (int d = 0; d < m; d++) {
        int piv = d;
        double big = fabs(sys[d][d]);
        for (int r = d + 1; r < m; r++)
            if (fabs(sys[r][d]) > big) { big = fabs(sys[r][d]); piv = r; }
        if (big < 1e-30) return -1;
        if (piv != d)
            for (int c = 0; c <= m; c++) { const double t = sys[d][c]; sys[d][c] = sys[piv][c]; sys[piv][c] = t; }
        for (int r = d + 1; r < m; r++) {
            const double f = sys[r][d] / sys[d][d];
            for (int c = d; c <= m; c++) sys[r][c] -= f * sys[d][c];
        }
    }
    double sol[DSMC_POLY_MAX_DEG + 1];               /* constant term first */
    for (int r = m - 1; r >= 0; r--) {
        sol[r] = sys[r][m];
        for (int c = r + 1; c < m; c++) sol[r] -= sys[r][c] * sol[c];
        sol[r] /= sys[r][r];
    }
    for (int k = 0; k < m; k++) out[k] = sol[deg - k];
    return 0;
}

/* polyfit.c:103-110: Horner's rule from the highest coefficient c[0]. */
DSMC_HD double horner(const double *c, int deg, double x)
{
    double v = c[0];
    for (int k = 1; k <= deg; k++) v = v * x + c[k];
    return v;
}

#endif /* DSMC_POLYFIT_H */
