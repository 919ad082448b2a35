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

namespace {
// four elements per thread: 32-byte double and 16-byte float accesses
template <typename S, typename D> __global__ void __launch_bounds__(256) k_cast(const S* src, D* dst, long n)
{
    const long stride = 4L * gridDim.x * blockDim.x;
    for (long i = 4L * (blockIdx.x * (long)blockDim.x + threadIdx.x); i < n; i += stride) {
        if (i + 4 <= n) {
#pragma unroll
            for (int e = 0; e < 4; ++e) dst[i + e] = (D)src[i + e];
        } else {
            for (long e = i; e < n; ++e) dst[e] = (D)src[e];
        }
    }
}
}  // namespace

hipError_t launch_cast(const void* src, int src_bytes, void* dst, int dst_bytes, long n, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const unsigned g = grid_for((n + 3) / 4);
    if (src_bytes == 8 && dst_bytes == 4)
        hipLaunchKernelGGL((k_cast<double, float>), dim3(g), dim3(256), 0, s, (const double*)src, (float*)dst, n);
    else if (src_bytes == 4 && dst_bytes == 8)
        hipLaunchKernelGGL((k_cast<float, double>), dim3(g), dim3(256), 0, s, (const float*)src, (double*)dst, n);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

namespace {
template <typename T> __global__ void __launch_bounds__(256) k_scale(const T* src, T* dst, long n, T f)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = f * src[i];
}
}  // namespace

hipError_t launch_scale(const void* src, void* dst, int elem_bytes, long n, double f, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const unsigned g = grid_for(n);
    if (elem_bytes == 8)
        hipLaunchKernelGGL(k_scale<double>, dim3(g), dim3(256), 0, s, (const double*)src, (double*)dst, n, f);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(g), dim3(256), 0, s, (const float*)src, (float*)dst, n, (float)f);
    return hipGetLastError();
}

namespace {
template <typename V>
__global__ void __launch_bounds__(256) k_transpose_c(const V* __restrict__ in, V* __restrict__ out, int rows,
                                                    int cols, long in_ld, long out_ld, long in_comp,
                                                    long out_comp)
{
    __shared__ V tile[32][33];
    const V* ip = in + blockIdx.z * in_comp;
    V* op = out + blockIdx.z * out_comp;
    const int j0 = blockIdx.x * 32, i0 = blockIdx.y * 32;
    const int tx = threadIdx.x, ty = threadIdx.y;
    for (int r = ty; r < 32; r += 8) {
        const int i = i0 + r, j = j0 + tx;
        if (i < rows && j < cols) tile[r][tx] = ip[(long)i * in_ld + j];
    }
    __syncthreads();
    for (int r = ty; r < 32; r += 8) {
        const int j = j0 + r, i = i0 + tx;
        if (j < cols && i < rows) op[(long)j * out_ld + i] = tile[tx][r];
    }
}
}  // namespace

hipError_t launch_transpose_c(const void* in, void* out, int elem_bytes, int rows, int cols, long in_ld,
                              long out_ld, long in_comp, long out_comp, int ncomp, hipStream_t s)
{
    if (rows <= 0 || cols <= 0 || ncomp <= 0) return hipSuccess;
    const dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32), (unsigned)ncomp);
    const dim3 block(32, 8);
    if (elem_bytes == 16)
        hipLaunchKernelGGL(k_transpose_c<double2>, grid, block, 0, s, reinterpret_cast<const double2*>(in),
                           reinterpret_cast<double2*>(out), rows, cols, in_ld, out_ld, in_comp, out_comp);
    else
        hipLaunchKernelGGL(k_transpose_c<float2>, grid, block, 0, s, reinterpret_cast<const float2*>(in),
                           reinterpret_cast<float2*>(out), rows, cols, in_ld, out_ld, in_comp, out_comp);
    return hipGetLastError();
}

}  // namespace psamd
