// Host entry points of the mixed-precision (matrix-core) pass kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_kernels.hpp"

namespace psamd {

// whether the fused z pass of a length-L transform (half: unbounded z, the
// input is nz = L/2 planes and only nz planes are kept) with operator `op`
// has a matrix-core kernel
bool zmfma_supported(int L, int half, int op);

// fused z pass on the matrix cores (FP16 hi/lo split DFT stages, FP32
// accumulation); `a` as for the FP32 fused z pass, one-component operators.
// variant (tuning field mfma): transforms of 512, 1024 and 2048 run the
// persistent register-pencil kernels for 1 (when the arguments allow it:
// kx-contiguous rows, the octant in the [ky][kx][kz] layout); for 2,
// transforms of 512 and 1024 run the one-tile-per-workgroup kernels and
// transforms of 2048 8-pencil workgroups; for 3, transforms of 512 and 1024
// the one-tile kernels and of 2048 4-pencil workgroups
hipError_t launch_zmfma(const PencilArgs<float>& a, int variant, hipStream_t s);
// whether a context's matrix-core z pass of length L with this tuning
// variant reads the Gk octant in the [ky][kx][kz] layout
inline bool zmfma_reg_layout(int L, int variant) { return (L == 512 || L == 1024 || L == 2048) && variant == 1; }

// several z-pass outputs of one scalar from one forward transform (the
// potential form): mask bit 0 A (Gk), bit 1 d/dz A (i kzh Gk), bit 2 the
// Laplacian (-|k|^2 Gk); output o (the o-th set bit) is written to component
// o of a.out (stride a.oa.comp), the input is component 0 of a.in.  At least
// two outputs, variant 1, transforms of 512 (one pencil per wavefront) and
// 2048 (one pencil, both residues, per wavefront)
bool zmfma_fan_ok(const PencilArgs<float>& a, int mask, int variant);
hipError_t launch_zmfma_fan(const PencilArgs<float>& a, int mask, hipStream_t s);

// three-component operators (div, curl) of the fused z pass on the matrix
// cores: inputs are components 0..2 of `a.in`, outputs components of `a.out`;
// variant 1 the register-pencil kernel when the arguments allow it (as
// above), else one tile of 4 pencils per workgroup
bool zmfma3_supported(int L, int half, int op);
hipError_t launch_zmfma3(const PencilArgs<float>& a, int variant, hipStream_t s);

// per-lane fp16 hi/lo fragments of the 32-point DFT matrix that the
// matrix-core kernels multiply by (binary16 bit patterns, kZmfmaFragHalves
// entries; PencilArgs::mfrag points at a device copy)
constexpr int kZmfmaFragHalves = 64 * 8 * 8;
void zmfma_fragments(unsigned short* out);
// the same for the 16-point DFT (transforms of 512): kZmfma16FragHalves
// entries, [lane][rh, rl, ih, il][8 halves]
constexpr int kZmfma16FragHalves = 64 * 4 * 8;
void zmfma_fragments16(unsigned short* out);
// 16 x 16 tables W32^f(i, k) in that layout: kind 0 the 16-point DFT
// (f = 2ik), 1 P (f = k (2i + 1)), 2 Q (f = i (2k + 1)); the register-pencil
// kernel reads kinds 0, 1, 2 back to back (3 kZmfma16FragHalves entries)
void zmfma_fragments16k(unsigned short* out, int kind);
// round-to-nearest-even binary16 bit pattern of a finite double below 65504
// in magnitude, and its value
unsigned short f16_bits(double v);
double f16_value(unsigned short h);

}  // namespace psamd
