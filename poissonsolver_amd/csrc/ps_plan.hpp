// Host-side planning: radix splits, elements per thread, tile shapes and
// twiddle tables for the pass kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "ps_fft.hpp"

namespace psamd {

// A 1D transform plan on the host.  E is one of the compiled
// elements-per-thread values; every radix divides E; TP = L / E.
struct Plan1D {
    int L = 0, E = 0, TP = 0, nstage = 0;
    int radix[kMaxStage] = {0};
};

// Elements-per-thread values compiled into the library (PS_FOR_EACH_E).
const std::vector<int>& compiled_E();

// Choose E and radices for length L with E <= maxE: among plans with the
// fewest stages, E closest to prefE.  Returns false (and a reason) when L has
// a prime factor above 7 or no compiled E fits.
bool make_plan(int L, int maxE, Plan1D* out, std::string* why, int prefE = 16);

// Largest E allowed for a kernel variant: the 3-component fused z pass
// holds three pencils per thread, so it gets a smaller register budget.
int max_E_for(int precision_bytes, int ncomp);

// Workgroup shape of a strided pass (columns per tile) and of the x pass
// (row pairs per workgroup), bounded by kMaxThreads and an LDS budget per
// workgroup in bytes (clamped to [16 KB, 160 KB]).
int choose_B(const Plan1D& p, int elem_bytes, int ncol, int lds_bytes = 80 * 1024);
int choose_P(const Plan1D& p, int elem_bytes, int npairs, int lds_bytes = 80 * 1024);

// Twiddle table exp(-2 pi i j/L), j < L, computed in long double.
void twiddles(int L, std::vector<double>& re, std::vector<double>& im);

template <typename T> FftPlan<T> device_plan(const Plan1D& p, const cplx<T>* tw)
{
    FftPlan<T> d;
    d.tw = tw;
    d.L = p.L;
    d.TP = p.TP;
    d.nstage = p.nstage;
    for (int i = 0; i < kMaxStage; ++i) d.radix[i] = p.radix[i];
    return d;
}

}  // namespace psamd
