// Lengths with compile-time specialised pass kernels, their tile policy and
// the host entry points (defined in skern.hip, one translation unit per L).
#pragma once
#include <hip/hip_runtime.h>

#include "ps_kernels.hpp"
#include "ps_sfft.hpp"

// powers of two, and the doubled lengths of 48, 96, 160, 192, 320 and 384 cells
#define PS_FOR_EACH_SL(X) X(64) X(96) X(128) X(192) X(256) X(320) X(384) X(512) X(640) X(768) X(1024) X(2048) X(4096) X(8192)

namespace psamd {

// SM_INVFAN: the y^-1 pass of one loaded spectrum into up to three outputs
// (PencilArgs::ik_fan; FP32)
enum StaticMode : int { SM_FWD = 0, SM_INV = 1, SM_FUSED = 2, SM_INVFAN = 3 };

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

// Elements per thread: the first candidate whose radices (the divisors of E
// in 16, 8, 7, 5, 4, 3, 2) split L.  Powers of two take 16; lengths with a
// factor 3 or 5 take 12 or 20.  NC = 3 (the z pass of curl, div and
// reprojection) holds three pencils per thread, so it prefers 8 and, in
// FP64, 10 over 20.
constexpr bool s_plan_ok(int L, int E)
{
    return E <= L && L % E == 0 && L / E <= kMaxThreads && make_cplan(L, E).n >= 0;
}
constexpr int s_pick_E(int L, int NC, int elem_bytes)
{
    const int one[6] = {16, 12, 20, 24, 10, 8};
    const int three32[6] = {8, 12, 20, 10, 6, 16};
    const int three64[6] = {8, 12, 10, 6, 20, 16};
    const int* c = (NC != 3) ? one : (elem_bytes == 8 ? three64 : three32);
    for (int i = 0; i < 6; ++i)
        if (s_plan_ok(L, c[i])) return c[i];
    return -1;
}

template <typename T, int L, int NC = 1> struct SPolicy {
    static constexpr int E = s_pick_E(L, NC, (int)sizeof(T));
    static_assert(E > 0, "no specialised radix split for this length");
    static constexpr int TP = L / E;
    static constexpr int B = s_pick_B(TP, L, 2 * (int)sizeof(T));
    static constexpr int P = s_pick_P(TP, L, 2 * (int)sizeof(T));
};

// Host query: whether L has specialised kernels, and their E / B / P
// (ncomp = 3 gives the policy of the three-component z pass).
bool static_policy(int L, int elem_bytes, int ncomp, int* E, int* B, int* P);

template <typename T, int L> hipError_t s_launch_x(const XArgs<T>& a, int inverse, int half, hipStream_t s);
// fused z pass of zero-padded pencils of 2 LH as two LH-point residues
// (instantiated for float, LH = 1024); tw_half is the LH-point twiddle table
template <typename T, int LH> hipError_t s_launch_zres(const PencilArgs<T>& a, const void* tw_half, hipStream_t s);
template <typename T, int L>
hipError_t s_launch_pencil(const PencilArgs<T>& a, int nc, int mode, int half, hipStream_t s);

// runtime dispatch on L (ps_dispatch.cpp)
template <typename T> hipError_t static_launch_x(int L, const XArgs<T>& a, int inverse, int half, hipStream_t s);
template <typename T>
hipError_t static_launch_pencil(int L, const PencilArgs<T>& a, int nc, int mode, int half, hipStream_t s);

}  // namespace psamd
