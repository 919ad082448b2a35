// Kernel bodies and per-E launchers.  Included only by kern_e.hip, which is
// compiled once per elements-per-thread value E.
#pragma once
#include "ps_kernels.hpp"
#include "ps_launch.hpp"

namespace psamd {

constexpr double kPi = 3.1415926535897932;   // greens3d.cpp:16

// Workgroups are dispatched round-robin over the 8 XCDs.  Give each XCD a
// contiguous range of work ids so that neighbouring tiles, which share cache
// lines at tile edges and rows of the Green's octant, meet in one L2.
__device__ __forceinline__ long xcd_remap(long lin, long nblk)
{
    const long per = nblk >> 3;
    if (lin < (per << 3)) return (lin & 7) * per + (lin >> 3);
    return lin;
}

// Wavenumber of index m of a length-M transform (greens3d.cpp:75-118): the
// index maps to m - M when 2m >= M, so the Nyquist wavenumber is negative.
__device__ __forceinline__ double kappa_full(int m, int M, double dk)
{
    const int mm = (2 * m < M) ? m : m - M;
    return 2.0 * kPi * (double)mm * dk;
}

// ---------------------------------------------------------------------------
// strided pencil pass (y passes, fused z pass, DCT-I of the kernel octant)
// ---------------------------------------------------------------------------
template <typename T, int E, int NC>
__global__ void __launch_bounds__(kMaxThreads) k_pencil(PencilArgs<T> a)
{
    using C = cplx<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);

    const int B = a.B, TP = a.plan.TP, L = a.plan.L;
    const int tid = threadIdx.x;
    const int p = tid % B, t = tid / B;
    const long nblk = gridDim.x;
    const long lin = xcd_remap(blockIdx.x, nblk);
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab * a.ncoltile);
    const int col = ct * B + p;
    const bool cv = col < a.ncol;
    C* my = lds + p;

    C v[NC][E];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int cc = (NC == 1) ? comp : c;
        const long base = cc * a.in_comp + slab * a.in_slab + col * a.in_col;
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int pos = t + TP * u;
            C x = C{0, 0};
            if (cv && pos < a.n_in) {
                if (a.load_mode == LOAD_COMPLEX) {
                    x = reinterpret_cast<const C*>(a.in)[base + pos * a.in_pos];
                } else {
                    const int n = (pos <= L - pos) ? pos : L - pos;
                    x.x = reinterpret_cast<const T*>(a.in)[base + n * a.in_pos];
                }
            }
            v[c][u] = x;
        }
    }
    if (a.do_fwd) {
#pragma unroll
        for (int c = 0; c < NC; ++c) fft_pencil<E, -1, T>(v[c], a.plan, my, B, t);
    }

    const int op = a.op;
    const int nout = (NC == 1 || op == OP_DIV) ? 1 : 3;
    // per-tile wavenumbers: x from the column, y from the slab (Hermitian
    // part kh: Nyquist zeroed; kn: Nyquist only)
    const int kx = col, ky = slab;
    T kxh = 0, kyh = 0, kxn = 0, kyn = 0;
    double kxf = 0, kyf = 0;
    const T* grow = nullptr;
    if (op != OP_NONE) {
        kxf = kappa_full(kx, a.Mx, a.dk[0]);
        kyf = kappa_full(ky, a.My, a.dk[1]);
        const bool nyx = (2 * kx == a.Mx), nyy = (2 * ky == a.My);
        kxh = nyx ? T(0) : (T)kxf;
        kxn = nyx ? (T)kxf : T(0);
        kyh = nyy ? T(0) : (T)kyf;
        kyn = nyy ? (T)kyf : T(0);
        const int kyo = (ky <= a.My - ky) ? ky : a.My - ky;
        if (cv && a.G) grow = a.G + kyo * a.g_ky + kx;
    }

    for (int o = 0; o < nout; ++o) {
        C w[E];
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int kz = t + TP * u;
            if (op == OP_NONE) {
                w[u] = v[0][u];
                continue;
            }
            const int kzo = (kz <= a.Mz - kz) ? kz : a.Mz - kz;
            const double kzf = kappa_full(kz, a.Mz, a.dk[2]);
            const bool nyz = (2 * kz == a.Mz);
            const T kzh = nyz ? T(0) : (T)kzf;
            const T kzn = nyz ? (T)kzf : T(0);
            const T g = grow ? grow[kzo * a.g_kz] : T(0);
            C r;
            if (op == OP_PLAIN) {
                r = cscale(v[0][u], g);
            } else if (op == OP_GRAD) {
                const T kd = (comp == 0) ? kxh : (comp == 1) ? kyh : kzh;
                const C s = v[0][u];
                r = C{-kd * g * s.y, kd * g * s.x};
            } else if (op == OP_DIV || op == OP_CURL) {
                const int vi = (NC == 3) ? 1 : 0, wi = (NC == 3) ? 2 : 0;
                const C sx = v[0][u], sy = v[vi][u], sz = v[wi][u];
                C q;
                if (op == OP_DIV) {
                    q = C{kxh * sx.x + kyh * sy.x + kzh * sz.x, kxh * sx.y + kyh * sy.y + kzh * sz.y};
                } else if (o == 0) {
                    q = C{kyh * sz.x - kzh * sy.x, kyh * sz.y - kzh * sy.y};
                } else if (o == 1) {
                    q = C{kzh * sx.x - kxh * sz.x, kzh * sx.y - kxh * sz.y};
                } else {
                    q = C{kxh * sy.x - kyh * sx.x, kxh * sy.y - kyh * sx.y};
                }
                r = C{-g * q.y, g * q.x};   // g * i * q
            } else {
                // filtered (and reprojected) rhs: zeta(|k|) / N
                double zf = a.invN;
                if (a.reg > 0) {
                    const double arg = a.zarg * (kxf * kxf + kyf * kyf + kzf * kzf);
                    double sum = 0.0, term = 1.0;
                    for (int n = 0; n < a.reg / 2; ++n) {
                        sum += term;
                        term = term * arg / (double)(n + 1);
                    }
                    zf *= sum * exp(-arg);
                }
                const T z = (T)zf;
                const int oc = (NC == 1) ? 0 : o;
                r = cscale(v[oc][u], z);
                if (NC == 3 && op == OP_RHSF_PROJ) {
                    const int vi = (NC == 3) ? 1 : 0, wi = (NC == 3) ? 2 : 0;
                    const C sx = v[0][u], sy = v[vi][u], sz = v[wi][u];
                    const C dh = C{kxh * sx.x + kyh * sy.x + kzh * sz.x, kxh * sx.y + kyh * sy.y + kzh * sz.y};
                    const C dn = C{kxn * sx.x + kyn * sy.x + kzn * sz.x, kxn * sx.y + kyn * sy.y + kzn * sz.y};
                    const T ah = (o == 0) ? kxh : (o == 1) ? kyh : kzh;
                    const T an = (o == 0) ? kxn : (o == 1) ? kyn : kzn;
                    r = C{r.x - g * (ah * dh.x + an * dn.x), r.y - g * (ah * dh.y + an * dn.y)};
                }
            }
            w[u] = r;
        }
        if (a.do_inv) fft_pencil<E, +1, T>(w, a.plan, my, B, t);

        const int oc = (NC == 1) ? comp : o;
        const long obase = oc * a.out_comp + slab * a.out_slab + col * a.out_col;
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int pos = t + TP * u;
            if (cv && pos < a.n_out) {
                if (a.store_mode == STORE_COMPLEX)
                    reinterpret_cast<C*>(a.out)[obase + pos * a.out_pos] = w[u];
                else
                    reinterpret_cast<T*>(a.out)[obase + pos * a.out_pos] = w[u].x;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// x pass, forward: two real rows -> one complex pencil -> two half spectra
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ void __launch_bounds__(kMaxThreads) k_x_r2c(XArgs<T> a)
{
    using C = cplx<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int TP = a.plan.TP, L = a.plan.L;
    const int pp = threadIdx.x / TP, t = threadIdx.x - (threadIdx.x / TP) * TP;
    const long pair = (long)blockIdx.x * a.P + pp;
    const int comp = blockIdx.y;
    const long ra = 2 * pair, rb = ra + 1;
    const bool va = ra < a.nrows, vb = rb < a.nrows;
    const T* in = reinterpret_cast<const T*>(a.in) + comp * a.r_comp;

    C v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int pos = t + TP * u;
        C x = C{0, 0};
        if (pos < a.nvalid) {
            if (va) x.x = in[ra * a.r_row + pos];
            if (vb) x.y = in[rb * a.r_row + pos];
        }
        v[u] = x;
    }
    C* my = lds + pp * lds_len(L);
    fft_pencil<E, -1, T>(v, a.plan, my, 1, t);

    // Z = Xa + i Xb:  Xa[k] = (Z[k] + conj Z[L-k])/2,  Xb[k] = (Z[k] - conj Z[L-k])/(2i)
#pragma unroll
    for (int u = 0; u < E; ++u) my[lds_pos(t + TP * u)] = v[u];
    __syncthreads();
    C* out = reinterpret_cast<C*>(a.out) + comp * a.c_comp;
    const T half = T(0.5);
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int pos = t + TP * u;
        if (pos < a.H) {
            const int mp = (pos == 0) ? 0 : L - pos;
            const C zk = v[u], zm = my[lds_pos(mp)];
            if (va) out[ra * a.c_row + pos] = C{half * (zk.x + zm.x), half * (zk.y - zm.y)};
            if (vb) out[rb * a.c_row + pos] = C{half * (zk.y + zm.y), half * (zm.x - zk.x)};
        }
    }
}

// ---------------------------------------------------------------------------
// x pass, inverse: two half spectra -> one complex pencil -> two real rows.
// DC and Nyquist bins contribute their real parts only, which is the real
// part of the full complex inverse the reference takes (pull.cpp:101-131).
// ---------------------------------------------------------------------------
template <typename T, int E>
__global__ void __launch_bounds__(kMaxThreads) k_x_c2r(XArgs<T> a)
{
    using C = cplx<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int TP = a.plan.TP, L = a.plan.L;
    const int pp = threadIdx.x / TP, t = threadIdx.x - (threadIdx.x / TP) * TP;
    const long pair = (long)blockIdx.x * a.P + pp;
    const int comp = blockIdx.y;
    const long ra = 2 * pair, rb = ra + 1;
    const bool va = ra < a.nrows, vb = rb < a.nrows;
    const C* in = reinterpret_cast<const C*>(a.in) + comp * a.c_comp;

    C v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int pos = t + TP * u;
        C za = C{0, 0}, zb = C{0, 0};
        if (pos < a.H) {
            if (va) za = in[ra * a.c_row + pos];
            if (vb) zb = in[rb * a.c_row + pos];
            if (pos == 0 || 2 * pos == L) {
                za.y = 0;
                zb.y = 0;
            }
        } else {
            const int m = L - pos;
            if (va) za = in[ra * a.c_row + m];
            if (vb) zb = in[rb * a.c_row + m];
            za.y = -za.y;
            zb.y = -zb.y;
        }
        v[u] = C{za.x - zb.y, za.y + zb.x};
    }
    C* my = lds + pp * lds_len(L);
    fft_pencil<E, +1, T>(v, a.plan, my, 1, t);
    T* out = reinterpret_cast<T*>(a.out) + comp * a.r_comp;
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int pos = t + TP * u;
        if (pos < a.nvalid) {
            if (va) out[ra * a.r_row + pos] = v[u].x;
            if (vb) out[rb * a.r_row + pos] = v[u].y;
        }
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename K> static hipError_t allow_lds(K kernel, size_t bytes)
{
    if (bytes <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel),
                               hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

template <typename T, int E> hipError_t launch_x_r2c_E(const XArgs<T>& a, hipStream_t s)
{
    const size_t shm = (size_t)a.P * lds_len(a.plan.L) * sizeof(cplx<T>);
    hipError_t e = allow_lds(k_x_r2c<T, E>, shm);
    if (e != hipSuccess) return e;
    dim3 grid((unsigned)((a.npairs + a.P - 1) / a.P), (unsigned)a.ncomp);
    hipLaunchKernelGGL((k_x_r2c<T, E>), grid, dim3(a.plan.TP * a.P), shm, s, a);
    return hipGetLastError();
}

template <typename T, int E> hipError_t launch_x_c2r_E(const XArgs<T>& a, hipStream_t s)
{
    const size_t shm = (size_t)a.P * lds_len(a.plan.L) * sizeof(cplx<T>);
    hipError_t e = allow_lds(k_x_c2r<T, E>, shm);
    if (e != hipSuccess) return e;
    dim3 grid((unsigned)((a.npairs + a.P - 1) / a.P), (unsigned)a.ncomp);
    hipLaunchKernelGGL((k_x_c2r<T, E>), grid, dim3(a.plan.TP * a.P), shm, s, a);
    return hipGetLastError();
}

template <typename T, int E, int NC> hipError_t launch_pencil_E(const PencilArgs<T>& a, hipStream_t s)
{
    const size_t shm = (size_t)a.B * lds_len(a.plan.L) * sizeof(cplx<T>);
    hipError_t e = allow_lds(k_pencil<T, E, NC>, shm);
    if (e != hipSuccess) return e;
    const long ntile = (long)a.ncoltile * a.nslab;
    const long nwork = ntile * (NC == 1 ? a.ncomp : 1);
    hipLaunchKernelGGL((k_pencil<T, E, NC>), dim3((unsigned)nwork), dim3(a.plan.TP * a.B), shm, s, a);
    return hipGetLastError();
}

}  // namespace psamd
