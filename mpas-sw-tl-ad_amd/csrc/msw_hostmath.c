/* Host-side loops of C-library transcendentals for the Williamson initial
 * conditions (mpas_sw_amd/testcases.py): one libm call per element, exactly
 * the call the reference's compiled Fortran makes for each value (a single
 * sincos where it evaluates sin and cos of one argument), so results are
 * those calls' bits.  Built with gcc -O2 -fno-builtin and no fast-math, so the
 * compiler neither vectorises, splits nor folds the calls. */
#define _GNU_SOURCE
#include <math.h>

void msw_vsincos(long n, const double *x, double *s, double *c)
{
    for (long i = 0; i < n; ++i) sincos(x[i], &s[i], &c[i]);
}

void msw_vsin(long n, const double *x, double *out)
{
    for (long i = 0; i < n; ++i) out[i] = sin(x[i]);
}

void msw_vcos(long n, const double *x, double *out)
{
    for (long i = 0; i < n; ++i) out[i] = cos(x[i]);
}

void msw_vpow(long n, const double *x, double y, double *out)
{
    for (long i = 0; i < n; ++i) out[i] = pow(x[i], y);
}
