/*
 * flops.c -- LAWN-41 operation counts used for every GFLOP/s figure
 * (test infrastructure).  Restates src/flops.h: FMULS/FADDS for POTRF
 * (:86-87), POTRS (:92-93), GEMM (:44-45), SYRK (:52-53), TRMM/TRSM (:62-70)
 * and the real-precision FLOPS_D* sums (:215,231,249,272,282).
 */
#include "oracle.h"

double orc_flops_dpotrf(double n)
{
    double fmuls = n * (((1. / 6.) * n + 0.5) * n + (1. / 3.));
    double fadds = n * (((1. / 6.) * n) * n - (1. / 6.));
    return fmuls + fadds;
}

double orc_flops_dpotrs(double n, double nrhs)
{
    return nrhs * n * (n + 1.) + nrhs * n * (n - 1.);
}

double orc_flops_dgemm(double m, double n, double k)
{
    return 2. * m * n * k;
}

double orc_flops_dsyrk(double k, double n)
{
    return 2. * (0.5 * k * n * (n + 1.));
}

double orc_flops_dtrsm(int side, double m, double n)
{
    /* FMULS_TRSM == FMULS_TRMM and FADDS_TRSM is defined as FMULS_TRMM
     * (flops.h:69-70), so the count is twice the multiply count. */
    double mul = (side == ORC_LEFT) ? 0.5 * n * m * (m + 1.) : 0.5 * m * n * (n + 1.);
    return 2. * mul;
}
