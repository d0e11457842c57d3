/* ref_glue.c -- links the reference's Fortran Cullum-Willoughby routines,
 * compiled here with flang, to the two runtime routines they call, as the
 * reference's own build links them with gfortran (-lgfortran -llapack,
 * Compilation/makefile_LINUX_TOBY:24-28).
 *
 * gfortran treats RAND and SRAND (genran.f:20-22) as GNU intrinsics, i.e.
 * calls into libgfortran's _gfortran_rand / _gfortran_srand.  flang knows no
 * such intrinsics, so it calls external procedures rand_ / srand_; under
 * genran.f's IMPLICIT DOUBLE PRECISION (R) the external RAND is typed
 * DOUBLE PRECISION, where gfortran converts the intrinsic's REAL*4 result.
 * DLAMCH (bisec.f:66, inverr.f:67, inverm.f:49, sturmi.f:50) is LAPACK's.
 * The functions below only forward to the GNU runtime and to LAPACK as
 * installed in this image (numpy's and scipy's bundled libgfortran and
 * OpenBLAS); nothing is reimplemented.  Test infrastructure only. */
#include <stddef.h>

float _gfortran_rand(int *flag);
void _gfortran_srand(int *seed);
double scipy_dlamch_(const char *cmach, size_t len);

double rand_(void) { return (double)_gfortran_rand(NULL); }

void srand_(int *seed) { _gfortran_srand(seed); }

double dlamch_(const char *cmach, size_t len) { return scipy_dlamch_(cmach, len); }
