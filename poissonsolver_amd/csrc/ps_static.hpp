// Lengths with compile-time specialised pass kernels, their tile policy and
// the host entry points (defined in skern.hip, one translation unit per L).
#pragma once
#include <hip/hip_runtime.h>

#include "ps_kernels.hpp"

#define PS_FOR_EACH_SL(X) X(64) X(128) X(256) X(512) X(1024) X(2048)

namespace psamd {

enum StaticMode : int { SM_FWD = 0, SM_INV = 1, SM_FUSED = 2 };

constexpr int s_lds_len(int L) { return L + (L >> 5) + 1; }

// pencils per workgroup of the pencil-per-wavefront z kernel with 32
// elements per thread (FP32 pencils of 2048)
constexpr int kZwaveB32 = 8;

// columns per strided tile: 128-byte row segments when 512 threads and an
// 80 KB LDS budget (two workgroups per CU) allow it
constexpr int s_pick_B(int TP, int L, int cb)
{
    int B = 128 / cb;
    while (B > 1 && (TP * B > 512 || B * s_lds_len(L) * cb > 80 * 1024)) B /= 2;
    return B;
}

// row pairs per x-pass workgroup
constexpr int s_pick_P(int TP, int L, int cb)
{
    int P = 8;
    while (P > 1 && (TP * P > 512 || P * s_lds_len(L) * cb > 40 * 1024)) P /= 2;
    while (TP * P < 64 && TP * P * 2 <= 512) P *= 2;
    return P;
}

// NC = 3 (the z pass of curl, div and reprojection) holds three pencils per
// thread, so it uses E = 8 to stay in registers.
template <typename T, int L, int NC = 1> struct SPolicy {
    static constexpr int E = (NC == 3) ? 8 : 16;
    static constexpr int TP = L / E;
    static constexpr int B = s_pick_B(TP, L, 2 * (int)sizeof(T));
    static constexpr int P = s_pick_P(TP, L, 2 * (int)sizeof(T));
};

// Host query: whether L has specialised kernels, and their E / B / P
// (ncomp = 3 gives the policy of the three-component z pass).
bool static_policy(int L, int elem_bytes, int ncomp, int* E, int* B, int* P);

template <typename T, int L> hipError_t s_launch_x(const XArgs<T>& a, int inverse, int half, hipStream_t s);
template <typename T, int L>
hipError_t s_launch_pencil(const PencilArgs<T>& a, int nc, int mode, int half, hipStream_t s);

// runtime dispatch on L (ps_dispatch.cpp)
template <typename T> hipError_t static_launch_x(int L, const XArgs<T>& a, int inverse, int half, hipStream_t s);
template <typename T>
hipError_t static_launch_pencil(int L, const PencilArgs<T>& a, int nc, int mode, int half, hipStream_t s);

}  // namespace psamd
