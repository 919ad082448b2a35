// Instantiates every pass kernel for one elements-per-thread value PS_E.
#include "ps_kernels_impl.hpp"

#ifndef PS_E
#error "build with -DPS_E=<elements per thread>"
#endif

namespace psamd {
template hipError_t launch_x_r2c_E<float, PS_E>(const XArgs<float>&, hipStream_t);
template hipError_t launch_x_r2c_E<double, PS_E>(const XArgs<double>&, hipStream_t);
template hipError_t launch_x_c2r_E<float, PS_E>(const XArgs<float>&, hipStream_t);
template hipError_t launch_x_c2r_E<double, PS_E>(const XArgs<double>&, hipStream_t);
template hipError_t launch_pencil_E<float, PS_E, 1>(const PencilArgs<float>&, hipStream_t);
template hipError_t launch_pencil_E<double, PS_E, 1>(const PencilArgs<double>&, hipStream_t);
template hipError_t launch_pencil_E<float, PS_E, 2>(const PencilArgs<float>&, hipStream_t);
template hipError_t launch_pencil_E<float, PS_E, 3>(const PencilArgs<float>&, hipStream_t);
template hipError_t launch_pencil_E<double, PS_E, 3>(const PencilArgs<double>&, hipStream_t);
}  // namespace psamd
