// Gather / scatter between strided field blocks and the solver's planar
// fields (ps_io.hpp).  Memory-bound copies with a conversion; one thread per
// element, the x index fastest across a wavefront.
#include "ps_io.hpp"

namespace psamd {

namespace {

template <typename T>
__global__ void k_gather(const double* src, Strided st, T* dst, int ncomp, int nx, int ny, int nz)
{
    const long n3 = (long)nx * ny * nz;
    const long n = n3 * ncomp;
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int c = (int)(id / n3);
        long r = id - (long)c * n3;
        const int i = (int)(r % nx);
        r /= nx;
        const int j = (int)(r % ny);
        const int k = (int)(r / ny);
        dst[id] = (T)src[c * st.s0 + i * st.s1 + j * st.s2 + k * st.s3];
    }
}

template <typename T>
__global__ void k_scatter(const T* src, double* dst, Strided st, int cmask, int ncomp, int nx, int ny, int nz)
{
    const long n3 = (long)nx * ny * nz;
    const long n = n3 * ncomp;
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int c = (int)(id / n3);
        if (!((cmask >> c) & 1)) continue;
        long r = id - (long)c * n3;
        const int i = (int)(r % nx);
        r /= nx;
        const int j = (int)(r % ny);
        const int k = (int)(r / ny);
        dst[c * st.s0 + i * st.s1 + j * st.s2 + k * st.s3] = (double)src[id];
    }
}

unsigned grid_for(long n)
{
    long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

hipError_t launch_gather(const double* src, Strided st, void* dst, int elem_bytes, int ncomp, int nx, int ny,
                         int nz, hipStream_t s)
{
    const unsigned g = grid_for((long)nx * ny * nz * ncomp);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_gather<double>, dim3(g), dim3(256), 0, s, src, st, (double*)dst, ncomp, nx, ny, nz);
    else
        hipLaunchKernelGGL(k_gather<float>, dim3(g), dim3(256), 0, s, src, st, (float*)dst, ncomp, nx, ny, nz);
    return hipGetLastError();
}

hipError_t launch_scatter(const void* src, int elem_bytes, double* dst, Strided st, int cmask, int ncomp, int nx,
                          int ny, int nz, hipStream_t s)
{
    const unsigned g = grid_for((long)nx * ny * nz * ncomp);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_scatter<double>, dim3(g), dim3(256), 0, s, (const double*)src, dst, st, cmask, ncomp,
                           nx, ny, nz);
    else
        hipLaunchKernelGGL(k_scatter<float>, dim3(g), dim3(256), 0, s, (const float*)src, dst, st, cmask, ncomp, nx,
                           ny, nz);
    return hipGetLastError();
}

}  // namespace psamd
