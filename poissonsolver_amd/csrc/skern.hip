// Instantiates the specialised pass kernels for one length PS_SL.
#include "ps_skernels_impl.hpp"

#ifndef PS_SL
#error "build with -DPS_SL=<transform length>"
#endif

namespace psamd {
template hipError_t s_launch_x<float, PS_SL>(const XArgs<float>&, int, int, hipStream_t);
template hipError_t s_launch_x<double, PS_SL>(const XArgs<double>&, int, int, hipStream_t);
template hipError_t s_launch_pencil<float, PS_SL>(const PencilArgs<float>&, int, int, int, hipStream_t);
template hipError_t s_launch_pencil<double, PS_SL>(const PencilArgs<double>&, int, int, int, hipStream_t);
#if PS_SL == 1024
// z pencils of 2048 as two residues of 1024 (FP32)
template hipError_t s_launch_zres<float, 1024>(const PencilArgs<float>&, const void*, hipStream_t);
#endif
}  // namespace psamd
