// Per-E launchers of the pass kernels and the list of compiled E values
// (elements per thread).  Each E is instantiated in its own translation unit
// (kern_e.hip built with -DPS_E=<E>) so the build runs them in parallel.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_kernels.hpp"

// E values and the radices each one allows (radix must divide E).
#define PS_FOR_EACH_E(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(12) X(14) X(15) X(16) X(20) X(21) X(30) X(32) X(35) X(42) X(70)

namespace psamd {

template <typename T, int E> hipError_t launch_x_r2c_E(const XArgs<T>& a, hipStream_t s);
template <typename T, int E> hipError_t launch_x_c2r_E(const XArgs<T>& a, hipStream_t s);
template <typename T, int E, int NC> hipError_t launch_pencil_E(const PencilArgs<T>& a, hipStream_t s);

// runtime dispatch on E (defined in ps_dispatch.cpp)
template <typename T> hipError_t launch_x_r2c(int E, const XArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_x_c2r(int E, const XArgs<T>& a, hipStream_t s);
template <typename T> hipError_t launch_pencil(int E, int NC, const PencilArgs<T>& a, hipStream_t s);

}  // namespace psamd
