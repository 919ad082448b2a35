// Checker build of the reference's special functions -- TEST INFRASTRUCTURE ONLY.
//
// The reference's special_functions.cpp defines free functions that need only
// the C++ standard library (its top-level translation unit, poisson_solver.cpp,
// also pulls in MPI, which this image lacks, so nothing else of the reference is
// built).  This translation unit includes that one source file where it lies
// under the reference checkout and exports its functions with C linkage, so the
// oracle's restatement can be compared against the reference's own arithmetic.
// Output goes to oracle/_ref/ only.
#include <cmath>
#include <complex>
#include <iostream>
#include "special_functions.cpp"

extern "C" double ref_sine_int(double x) { return sine_int(x); }
extern "C" double ref_bessel_int_J0(double x) { return bessel_int_J0(x); }
extern "C" void ref_sine_int_v(const double* x, double* y, long n)
{
	for (long i = 0; i < n; ++i) y[i] = sine_int(x[i]);
}
extern "C" void ref_bessel_int_J0_v(const double* x, double* y, long n)
{
	for (long i = 0; i < n; ++i) y[i] = bessel_int_J0(x[i]);
}
