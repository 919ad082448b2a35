// Runtime dispatch from a plan's elements-per-thread value to the kernels
// instantiated for it in kern_e.hip.
#include "ps_launch.hpp"
#include "ps_static.hpp"

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
    if (NC == 2) {   // one input, two outputs: single precision only
        if constexpr (sizeof(T) == 4) {
            switch (E) {
#define PS_CASE(e) case e: return launch_pencil_E<T, e, 2>(a, s);
                PS_FOR_EACH_E(PS_CASE)
#undef PS_CASE
            default: return hipErrorInvalidValue;
            }
        }
        return hipErrorInvalidValue;
    }
    switch (E) {
#define PS_CASE(e) case e: return launch_pencil_E<T, e, 3>(a, s);
        PS_FOR_EACH_E(PS_CASE)
#undef PS_CASE
    default: return hipErrorInvalidValue;
    }
}

bool static_policy(int L, int elem_bytes, int ncomp, int* E, int* B, int* P)
{
    switch (L) {
#define PS_CASE(l)                                                                                  \
    case l:                                                                                         \
        if (ncomp == 3) {                                                                           \
            if (elem_bytes == 4) {                                                                  \
                *E = SPolicy<float, l, 3>::E; *B = SPolicy<float, l, 3>::B; *P = SPolicy<float, l, 3>::P;   \
            } else {                                                                                \
                *E = SPolicy<double, l, 3>::E; *B = SPolicy<double, l, 3>::B; *P = SPolicy<double, l, 3>::P; \
            }                                                                                       \
        } else if (elem_bytes == 4) {                                                               \
            *E = SPolicy<float, l>::E; *B = SPolicy<float, l>::B; *P = SPolicy<float, l>::P;         \
        } else {                                                                                    \
            *E = SPolicy<double, l>::E; *B = SPolicy<double, l>::B; *P = SPolicy<double, l>::P;      \
        }                                                                                           \
        return true;
        PS_FOR_EACH_SL(PS_CASE)
#undef PS_CASE
    default: return false;
    }
}

template <typename T> hipError_t static_launch_x(int L, const XArgs<T>& a, int inverse, int half, hipStream_t s)
{
    switch (L) {
#define PS_CASE(l) case l: return s_launch_x<T, l>(a, inverse, half, s);
        PS_FOR_EACH_SL(PS_CASE)
#undef PS_CASE
    default: return hipErrorInvalidValue;
    }
}

template <typename T>
hipError_t static_launch_pencil(int L, const PencilArgs<T>& a, int nc, int mode, int half, hipStream_t s)
{
    switch (L) {
#define PS_CASE(l) case l: return s_launch_pencil<T, l>(a, nc, mode, half, s);
        PS_FOR_EACH_SL(PS_CASE)
#undef PS_CASE
    default: return hipErrorInvalidValue;
    }
}

template hipError_t static_launch_x<float>(int, const XArgs<float>&, int, int, hipStream_t);
template hipError_t static_launch_x<double>(int, const XArgs<double>&, int, int, hipStream_t);
template hipError_t static_launch_pencil<float>(int, const PencilArgs<float>&, int, int, int, hipStream_t);
template hipError_t static_launch_pencil<double>(int, const PencilArgs<double>&, int, int, int, hipStream_t);

template hipError_t launch_x_r2c<float>(int, const XArgs<float>&, hipStream_t);
template hipError_t launch_x_r2c<double>(int, const XArgs<double>&, hipStream_t);
template hipError_t launch_x_c2r<float>(int, const XArgs<float>&, hipStream_t);
template hipError_t launch_x_c2r<double>(int, const XArgs<double>&, hipStream_t);
template hipError_t launch_pencil<float>(int, int, const PencilArgs<float>&, hipStream_t);
template hipError_t launch_pencil<double>(int, int, const PencilArgs<double>&, hipStream_t);

}  // namespace psamd
