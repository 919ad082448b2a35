// Green's kernel construction, FP64 on the device (setup only).
//
// The reference samples the kernel on the whole doubled grid, fft-shifts it
// and runs three FFT passes (greens3d.cpp:167-332).  The sampled kernel is even
// in every coordinate, so its DFT is real and even and equals a DCT-I of the
// octant n_d in [0, M_d/2]; only that octant is sampled and transformed here
// (each DCT-I runs as a length-M_d FFT of the even extension).  The result
// Ghat * zeta / (Mx My Mz) is stored as a real octant in the solve precision.
#include "ps_greens.hpp"

#include <cmath>
#include <vector>

#include "ps_launch.hpp"
#include "ps_plan.hpp"
#include "ps_special.hpp"

namespace psamd {

namespace {

constexpr double kPi = 3.1415926535897932;          // greens3d.cpp:16
constexpr double kGamma = 0.5772156649015329;       // greens2d.cpp:16

// UUU kernel Si(r/s)/(2 pi^2 r) dx dy dz, s = h/pi; 1/(2 pi^2 s) dx dy dz for
// r <= dx/4 (greens3d.cpp:198-224), at offsets (i dx, j dy, k dz).
__global__ void k_sample_uuu(double* g, int Ox, int Oy, int Oz, double dx, double dy, double dz,
                             double sigma)
{
    const long n = (long)Ox * Oy * Oz;
    const double c = 0.5 / (kPi * kPi);
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int i = (int)(id % Ox);
        const int j = (int)((id / Ox) % Oy);
        const int k = (int)(id / ((long)Ox * Oy));
        const double x = i * dx, y = j * dy, z = k * dz;
        const double r = sqrt(x * x + y * y + z * z);
        double v;
        if (r > 0.25 * dx) v = c * ps_sine_int(r / sigma) / r * dx * dy * dz;
        else v = c / sigma * dx * dy * dz;
        g[id] = v;
    }
}

// 2D kernel -(1/2pi)(Ji0(r/s) + ln(2s) - gamma) dx dy (greens2d.cpp:183-201)
__global__ void k_sample_2d(double* g, int Ox, int Oy, double dx, double dy, double sigma, double c2)
{
    const long n = (long)Ox * Oy;
    const double c = 0.5 / kPi;
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int i = (int)(id % Ox);
        const int j = (int)(id / Ox);
        const double x = i * dx, y = j * dy;
        const double r = sqrt(x * x + y * y);
        g[id] = -c * (ps_bessel_int_J0(r / sigma) + c2) * dx * dy;
    }
}

// 1D kernel -(s/pi)(rho Si(rho) + cos rho) dz, rho = |z|/s, s = dz/pi
// (greens2d.cpp:329-338, the kernel the PPU branch of greens3d.cpp:357-370 builds)
__global__ void k_sample_1d(double* g, int Oz, double dz)
{
    const double sigma = dz / kPi, c1 = sigma / kPi;
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < Oz; k += gridDim.x * blockDim.x) {
        const double rho = fabs(k * dz) / sigma;
        g[k] = -c1 * (ps_sine_int(rho) * rho + cos(rho)) * dz;
    }
}

__device__ inline double kap(int m, int M, double dk)
{
    const int mm = (2 * m < M) ? m : m - M;
    return 2.0 * kPi * (double)mm * dk;
}

// Ghat * zeta / N on the octant; the periodic branches use -1/(ik.ik)
// (greens3d.cpp:396, :461, :514) with their special columns/planes.
template <typename T>
__global__ void k_finalize(T* out, const double* gsp, int mode, const double* g2d, const double* g1d, int Ox,
                           int Oy, int Oz, int Mx, int My, int Mz, double dkx, double dky, double dkz,
                           double zarg, int reg, double invN, int layout)
{
    const long n = (long)Ox * Oy * Oz;
    for (long id = blockIdx.x * (long)blockDim.x + threadIdx.x; id < n; id += (long)gridDim.x * blockDim.x) {
        const int i = (int)(id % Ox);
        const int j = (int)((id / Ox) % Oy);
        const int k = (int)(id / ((long)Ox * Oy));
        const double kx = kap(i, Mx, dkx), ky = kap(j, My, dky), kz = kap(k, Mz, dkz);
        const double k2 = kx * kx + ky * ky + kz * kz;
        double g;
        if (mode == 0) {            // UUU: transformed real-space kernel
            g = gsp[id];
        } else if (mode == 1) {     // PPU: 1D kernel on the kx = ky = 0 column
            g = (i == 0 && j == 0) ? g1d[k] : 1.0 / k2;
        } else if (mode == 2) {     // UUP: 2D kernel on the kz = 0 plane
            g = (k == 0) ? g2d[(long)j * Ox + i] : 1.0 / k2;
        } else if (mode == 4) {     // 2D UU: the 2D kernel transform on the (kx, kz) plane (Oy = 1)
            g = gsp[id];
        } else if (mode == 5) {     // 2D UP: 1D kernel along x on the kz = 0 row
            g = (k == 0) ? g1d[i] : 1.0 / k2;
        } else {                    // PPP: zero mode removed
            g = (i == 0 && j == 0 && k == 0) ? 0.0 : 1.0 / k2;
        }
        double z = 1.0;
        if (reg > 0) {
            // zeta_m = e^-a sum_{n < m/2} a^n/n!  (greens3d.cpp:123-159)
            const double a = zarg * k2;
            double sum = 0.0, term = 1.0;
            for (int q = 0; q < reg / 2; ++q) {
                sum += term;
                term = term * a / (double)(q + 1);
            }
            z = sum * exp(-a);
        }
        // layout 0: [ky][kz][kx]; layout 1: [ky][kx][kz].  Either way a z-pass
        // tile (fixed ky) reads one contiguous slab.
        const long o = layout ? ((long)j * Ox + i) * Oz + k : ((long)j * Oz + k) * Ox + i;
        out[o] = (T)(g * z * invN);
    }
}

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) (void)hipFree(p);
    }
    hipError_t alloc(size_t bytes) { return hipMalloc(&p, bytes); }
};

#define PS_HIP(x)                                                                 \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            if (err) *err = std::string(#x) + ": " + hipGetErrorString(e_);       \
            return e_ == hipErrorOutOfMemory ? kGreensENOMEM : kGreensEHIP;       \
        }                                                                         \
    } while (0)

// DCT-I in place along dimension d of a real octant with dims O[3] (x fastest)
static int dct_pass(double* g, const int O[3], int d, int M, hipStream_t s, std::string* err)
{
    Plan1D p;
    std::string why;
    if (!make_plan(M, max_E_for(8, 1), &p, &why)) {
        if (err) *err = why;
        return kGreensEUNSUPPORTED;
    }
    std::vector<double> re, im;
    twiddles(M, re, im);
    std::vector<double> tw(2 * M);
    for (int j = 0; j < M; ++j) {
        tw[2 * j] = re[j];
        tw[2 * j + 1] = im[j];
    }
    DevBuf dtw;
    PS_HIP(dtw.alloc(tw.size() * sizeof(double)));
    PS_HIP(hipMemcpyAsync(dtw.p, tw.data(), tw.size() * sizeof(double), hipMemcpyHostToDevice, s));

    PencilArgs<double> a{};
    a.plan = device_plan<double>(p, reinterpret_cast<const double2*>(dtw.p));
    a.in = g;
    a.out = g;
    const long sx = 1, sy = O[0], sz = (long)O[0] * O[1];
    long pos_s, col_s, slab_s;
    int ncol, nslab;
    if (d == 0) {
        pos_s = sx; col_s = sy; slab_s = sz; ncol = O[1]; nslab = O[2];
    } else if (d == 1) {
        pos_s = sy; col_s = sx; slab_s = sz; ncol = O[0]; nslab = O[2];
    } else {
        pos_s = sz; col_s = sx; slab_s = sy; ncol = O[0]; nslab = O[1];
    }
    a.ia = a.oa = addr_plain(0, slab_s, pos_s, col_s);
    a.xcd_remap = 1;
    a.n_in = M;
    a.n_out = M / 2 + 1;
    a.ncol = ncol;
    a.B = choose_B(p, 8, ncol);
    if ((size_t)a.B * lds_len(M) * sizeof(double2) > 160 * 1024) {
        // one FP64 pencil of the kernel transform is staged whole in LDS
        if (err) *err = "kernel transform length " + std::to_string(M) + " exceeds the FP64 pencil held in LDS";
        return kGreensEUNSUPPORTED;
    }
    a.ncoltile = (ncol + a.B - 1) / a.B;
    a.nslab = nslab;
    a.ncomp = 1;
    a.load_mode = LOAD_REAL_EVEN;
    a.store_mode = STORE_REAL;
    a.do_fwd = 1;
    a.do_inv = 0;
    a.op = OP_NONE;
    PS_HIP(launch_pencil<double>(p.E, 1, a, s));
    PS_HIP(hipStreamSynchronize(s));   // dtw is freed on return
    return 0;
}

static unsigned grid_for(long n)
{
    long b = (n + 255) / 256;
    if (b > 65536) b = 65536;
    return (unsigned)(b < 1 ? 1 : b);
}

}  // namespace

// 2D grids (dim2) are held as (nx, 1, ny) with the y direction on the third
// axis; their branches (greens2d.cpp:157-429) map onto the modes as
//   UU (0,1,0) -> 4  Ji0 kernel sampled on the (x, z) quadrant, DCT-I along x and z
//   PU (1,1,0) -> 1  1D Si kernel along z on the kx = 0 column (:307-366)
//   UP (0,1,1) -> 5  1D Si kernel along x on the kz = 0 row; the reference
//                    stores every sample at a stale index j (:400), the
//                    intended index i is used
//   PP (1,1,1) -> 3  1/k^2, Ghat(0) = 0 (:268-303)
template <typename T>
int build_greens(const int ncell[3], const int bc[3], const double dx[3], int reg, T* out, int layout,
                 hipStream_t s, std::string* err, int dim2)
{
    int M[3], O[3];
    for (int d = 0; d < 3; ++d) {
        M[d] = (bc[d] == 0) ? 2 * ncell[d] : ncell[d];
        O[d] = M[d] / 2 + 1;
    }
    const double h = std::max(std::max(dx[0], dx[1]), dx[2]);    // greens3d.cpp:55
    const long n = (long)O[0] * O[1] * O[2];
    int mode;
    if (dim2) {
        if (ncell[1] != 1 || bc[1] != 1) {
            if (err) *err = "2D kernel: the grid must be held as (nx, 1, ny) with a periodic middle axis";
            return kGreensEBC;
        }
        mode = (bc[0] == 0) ? (bc[2] == 0 ? 4 : 5) : (bc[2] == 0 ? 1 : 3);
    } else if (bc[0] == 0 && bc[1] == 0 && bc[2] == 0) mode = 0;
    else if (bc[0] == 1 && bc[1] == 1 && bc[2] == 0) mode = 1;
    else if (bc[0] == 0 && bc[1] == 0 && bc[2] == 1) mode = 2;
    else if (bc[0] == 1 && bc[1] == 1 && bc[2] == 1) mode = 3;
    else {
        if (err) *err = "unsupported boundary condition tuple";
        return kGreensEBC;
    }

    DevBuf g3, g2, g1;
    if (mode == 0) {
        PS_HIP(g3.alloc(n * sizeof(double)));
        hipLaunchKernelGGL(k_sample_uuu, dim3(grid_for(n)), dim3(256), 0, s, (double*)g3.p, O[0], O[1], O[2],
                           dx[0], dx[1], dx[2], h / kPi);
        PS_HIP(hipGetLastError());
        for (int d = 0; d < 3; ++d) {
            int r = dct_pass((double*)g3.p, O, d, M[d], s, err);
            if (r) return r;
        }
    } else if (mode == 1) {
        PS_HIP(g1.alloc(O[2] * sizeof(double)));
        hipLaunchKernelGGL(k_sample_1d, dim3(grid_for(O[2])), dim3(256), 0, s, (double*)g1.p, O[2], dx[2]);
        PS_HIP(hipGetLastError());
        const int O1[3] = {O[2], 1, 1};
        int r = dct_pass((double*)g1.p, O1, 0, M[2], s, err);
        if (r) return r;
    } else if (mode == 4) {
        // greens2d.cpp:157-264 on (x, z): sigma = max(dx, dy)/pi, C2 = ln(2 sigma) - gamma
        PS_HIP(g3.alloc(n * sizeof(double)));
        const double sigma2 = h / kPi;
        const double c2 = std::log(2.0 * sigma2) - kGamma;
        hipLaunchKernelGGL(k_sample_2d, dim3(grid_for(n)), dim3(256), 0, s, (double*)g3.p, O[0], O[2], dx[0],
                           dx[2], sigma2, c2);
        PS_HIP(hipGetLastError());
        int r = dct_pass((double*)g3.p, O, 0, M[0], s, err);
        if (r) return r;
        r = dct_pass((double*)g3.p, O, 2, M[2], s, err);
        if (r) return r;
    } else if (mode == 5) {
        PS_HIP(g1.alloc(O[0] * sizeof(double)));
        hipLaunchKernelGGL(k_sample_1d, dim3(grid_for(O[0])), dim3(256), 0, s, (double*)g1.p, O[0], dx[0]);
        PS_HIP(hipGetLastError());
        const int O1[3] = {O[0], 1, 1};
        int r = dct_pass((double*)g1.p, O1, 0, M[0], s, err);
        if (r) return r;
    } else if (mode == 2) {
        // greens3d.cpp:413-425 calls greens2d with (nx, ny), dx[2] = 0, so h = max(dx, dy)
        const long n2 = (long)O[0] * O[1];
        PS_HIP(g2.alloc(n2 * sizeof(double)));
        const double h2 = std::max(dx[0], dx[1]);
        const double sigma2 = h2 / kPi;
        const double c2 = std::log(2.0 * sigma2) - kGamma;
        hipLaunchKernelGGL(k_sample_2d, dim3(grid_for(n2)), dim3(256), 0, s, (double*)g2.p, O[0], O[1], dx[0],
                           dx[1], sigma2, c2);
        PS_HIP(hipGetLastError());
        const int O2[3] = {O[0], O[1], 1};
        int r = dct_pass((double*)g2.p, O2, 0, M[0], s, err);
        if (r) return r;
        r = dct_pass((double*)g2.p, O2, 1, M[1], s, err);
        if (r) return r;
    }
    int regm = reg;
    if (regm > 0 && regm % 2) regm += 1;                            // greens3d.cpp:126
    if (regm < 0) regm = 0;
    const double sigma_z = 2.0 * h;                                 // greens3d.cpp:123
    const double zarg = 0.5 * sigma_z * sigma_z;
    const double invN = 1.0 / ((double)M[0] * (double)M[1] * (double)M[2]);
    hipLaunchKernelGGL((k_finalize<T>), dim3(grid_for(n)), dim3(256), 0, s, out, (const double*)g3.p, mode,
                       (const double*)g2.p, (const double*)g1.p, O[0], O[1], O[2], M[0], M[1], M[2],
                       1.0 / ((double)M[0] * dx[0]), 1.0 / ((double)M[1] * dx[1]), 1.0 / ((double)M[2] * dx[2]),
                       zarg, regm, invN, layout);
    PS_HIP(hipGetLastError());
    PS_HIP(hipStreamSynchronize(s));
    return 0;
}

template int build_greens<float>(const int[3], const int[3], const double[3], int, float*, int, hipStream_t,
                                 std::string*, int);
template int build_greens<double>(const int[3], const int[3], const double[3], int, double*, int, hipStream_t,
                                  std::string*, int);

// Si and Ji0 evaluated by the device code, for tests of the special functions.
__global__ void k_eval_special(const double* x, double* y, long n, int which)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        y[i] = which == 0 ? ps_sine_int(x[i]) : ps_bessel_int_J0(x[i]);
}

int eval_special_device(const double* hx, double* hy, long n, int which, std::string* err)
{
    DevBuf dxb, dyb;
    hipStream_t s = nullptr;
    PS_HIP(dxb.alloc(n * sizeof(double)));
    PS_HIP(dyb.alloc(n * sizeof(double)));
    PS_HIP(hipMemcpy(dxb.p, hx, n * sizeof(double), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_eval_special, dim3(grid_for(n)), dim3(256), 0, s, (const double*)dxb.p, (double*)dyb.p,
                       n, which);
    PS_HIP(hipGetLastError());
    PS_HIP(hipMemcpy(hy, dyb.p, n * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace psamd
