// Host entry points of the mixed-precision (matrix-core) pass kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_kernels.hpp"

namespace psamd {

// whether the fused z pass of a length-L transform (half: unbounded z, the
// input is nz = L/2 planes and only nz planes are kept) with operator `op`
// has a matrix-core kernel
bool zmfma_supported(int L, int half, int op);

// fused z pass on the matrix cores (FP16 hi/lo split DFT-32 stages, FP32
// accumulation); `a` as for the FP32 fused z pass, one-component operators.
// persistent = 0: one tile per workgroup, 4 wavefronts per SIMD; 1: one
// persistent workgroup per CU that prefetches its next tile
hipError_t launch_zmfma(const PencilArgs<float>& a, int persistent, hipStream_t s);

// three-component operators (div, curl) of the fused z pass on the matrix
// cores: inputs are components 0..2 of `a.in`, outputs components of `a.out`
bool zmfma3_supported(int L, int half, int op);
hipError_t launch_zmfma3(const PencilArgs<float>& a, hipStream_t s);

// per-lane fp16 hi/lo fragments of the 32-point DFT matrix that the
// matrix-core kernels multiply by (binary16 bit patterns, kZmfmaFragHalves
// entries; PencilArgs::mfrag points at a device copy)
constexpr int kZmfmaFragHalves = 64 * 8 * 8;
void zmfma_fragments(unsigned short* out);
// the same for the 16-point DFT (transforms of 512): kZmfma16FragHalves
// entries, [lane][rh, rl, ih, il][8 halves]
constexpr int kZmfma16FragHalves = 64 * 4 * 8;
void zmfma_fragments16(unsigned short* out);
// round-to-nearest-even binary16 bit pattern of a finite double below 65504
// in magnitude, and its value
unsigned short f16_bits(double v);
double f16_value(unsigned short h);

}  // namespace psamd
