/*
 * ref_shim.c -- exports the reference's own header-only arithmetic so the
 * oracle can be checked against it (test infrastructure).  Compiled by
 * oracle/Makefile directly against the DPLASMA sources where they lie:
 *   src/cores/random.h:19-41            Rnd64_jump (LCG jump-ahead)
 *   src/include/dplasma/constants.h     enum values used by flops.h
 *   src/flops.h:86-93,215-282           LAWN-41 operation counts
 * Nothing here restates those files; the functions only expose their macros
 * and inline function through a C ABI.  The built library lands in
 * oracle/_ref/ and is never part of the product.
 */
#include "dplasma/constants.h"
#include "random.h"
#include "flops.h"

unsigned long long ref_rnd64_jump(unsigned long long n, unsigned long long seed)
{
    return Rnd64_jump(n, seed);
}
double ref_flops_dpotrf(double n) { return FLOPS_DPOTRF(n); }
double ref_flops_dpotrs(double n, double nrhs) { return FLOPS_DPOTRS(n, nrhs); }
double ref_flops_dgemm(double m, double n, double k) { return FLOPS_DGEMM(m, n, k); }
double ref_flops_dsyrk(double k, double n) { return FLOPS_DSYRK(k, n); }
double ref_flops_dtrsm(int side, double m, double n) { return FLOPS_DTRSM(side, m, n); }
double ref_flops_spotrf(double n) { return FLOPS_SPOTRF(n); }
