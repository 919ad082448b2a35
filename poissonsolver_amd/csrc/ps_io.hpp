// Strided field blocks at the push/pull boundary: the one-array form of the
// reference's Fortran interface (poisson_solver_push.f:8, poisson_solver_pull.f:8),
// an array real(ncomp, :, :, :) with the component index fastest and ghost
// cells around the physical cells.  Element (c, i, j, k) of such a block is
//   base[c*s0 + i*s1 + j*s2 + k*s3]   (base already offset to cell (0, 0, 0)).
#pragma once
#include <hip/hip_runtime.h>

namespace psamd {

struct Strided {
    long s0, s1, s2, s3;
};

// dst[c][k][j][i] (solve precision, elem_bytes 4 or 8) = src(c, i, j, k), c < ncomp
hipError_t launch_gather(const double* src, Strided st, void* dst, int elem_bytes, int ncomp, int nx, int ny,
                         int nz, hipStream_t s);

// dst(c, i, j, k) = src[c][k][j][i] for the components whose bit is set in cmask
hipError_t launch_scatter(const void* src, int elem_bytes, double* dst, Strided st, int cmask, int ncomp, int nx,
                          int ny, int nz, hipStream_t s);

// dst[i] = src[i], i < n, converting between double and float (elem bytes
// 8 or 4 each): the host-pointer push / pull stage FP64 data on the device
// and convert it there
hipError_t launch_cast(const void* src, int src_bytes, void* dst, int dst_bytes, long n, hipStream_t s);

// dst[i] = f * src[i], i < n (elem bytes 8 or 4; src == dst allowed)
hipError_t launch_scale(const void* src, void* dst, int elem_bytes, long n, double f, hipStream_t s);

// out[c][j][i] = in[c][i][j] for i < rows, j < cols: complex elements of
// elem_bytes (8 or 16), row strides in_ld / out_ld, component strides
// in_comp / out_comp (all in elements)
hipError_t launch_transpose_c(const void* in, void* out, int elem_bytes, int rows, int cols, long in_ld,
                              long out_ld, long in_comp, long out_comp, int ncomp, hipStream_t s);

}  // namespace psamd
