// Runtime dispatch from a plan's elements-per-thread value to the kernels
// instantiated for it in kern_e.hip.
#include "ps_launch.hpp"

namespace psamd {

template <typename T> hipError_t launch_x_r2c(int E, const XArgs<T>& a, hipStream_t s)
{
    switch (E) {
#define PS_CASE(e) case e: return launch_x_r2c_E<T, e>(a, s);
        PS_FOR_EACH_E(PS_CASE)
#undef PS_CASE
    default: return hipErrorInvalidValue;
    }
}

template <typename T> hipError_t launch_x_c2r(int E, const XArgs<T>& a, hipStream_t s)
{
    switch (E) {
#define PS_CASE(e) case e: return launch_x_c2r_E<T, e>(a, s);
        PS_FOR_EACH_E(PS_CASE)
#undef PS_CASE
    default: return hipErrorInvalidValue;
    }
}

template <typename T> hipError_t launch_pencil(int E, int NC, const PencilArgs<T>& a, hipStream_t s)
{
    if (NC == 1) {
        switch (E) {
#define PS_CASE(e) case e: return launch_pencil_E<T, e, 1>(a, s);
            PS_FOR_EACH_E(PS_CASE)
#undef PS_CASE
        default: return hipErrorInvalidValue;
        }
    }
    switch (E) {
#define PS_CASE(e) case e: return launch_pencil_E<T, e, 3>(a, s);
        PS_FOR_EACH_E(PS_CASE)
#undef PS_CASE
    default: return hipErrorInvalidValue;
    }
}

template hipError_t launch_x_r2c<float>(int, const XArgs<float>&, hipStream_t);
template hipError_t launch_x_r2c<double>(int, const XArgs<double>&, hipStream_t);
template hipError_t launch_x_c2r<float>(int, const XArgs<float>&, hipStream_t);
template hipError_t launch_x_c2r<double>(int, const XArgs<double>&, hipStream_t);
template hipError_t launch_pencil<float>(int, int, const PencilArgs<float>&, hipStream_t);
template hipError_t launch_pencil<double>(int, int, const PencilArgs<double>&, hipStream_t);

}  // namespace psamd
