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
// slab order 0, M/2, 1, M-1, 2, M-2, ... : the two slabs that share a row of
// the kernel octant are processed back to back (same XCD, L2-resident row)
__device__ __forceinline__ int paired_slab(int s, int M)
{
    if (s == 0) return 0;
    if ((M & 1) == 0) {
        if (s == 1) return M / 2;
        s -= 2;
    } else {
        s -= 1;
    }
    const int j = s / 2 + 1;
    return (s & 1) ? M - j : j;
}

__device__ __forceinline__ long boff(int i, int sh, long lo, long hi)
{
    return (long)(i >> sh) * hi + (long)(i & ((1 << sh) - 1)) * lo;
}

__device__ __forceinline__ double kappa_full(int m, int M, double dk)
{
    const int mm = (2 * m < M) ? m : m - M;
    return 2.0 * kPi * (double)mm * dk;
}


// Per-tile wavenumber data of the z pass: x from the column, y from the slab;
// kh is the Hermitian part of i*k (Nyquist zeroed), kn the Nyquist-only part.
template <typename T> struct TileK {
    T kxh, kyh, kxn, kyn;
    double kxf, kyf;
    const T* grow;
};

template <typename T>
__device__ __forceinline__ TileK<T> tile_k(const PencilArgs<T>& a, int col, int slab, bool cv)
{
    TileK<T> k{T(0), T(0), T(0), T(0), 0.0, 0.0, nullptr};
    if (a.op == OP_NONE) return k;
    const int kx = col + a.col0, ky = slab;
    k.kxf = kappa_full(kx, a.Mx, a.dk[0]);
    k.kyf = kappa_full(ky, a.My, a.dk[1]);
    const bool nyx = (2 * kx == a.Mx), nyy = (2 * ky == a.My);
    k.kxh = nyx ? T(0) : (T)k.kxf;
    k.kxn = nyx ? (T)k.kxf : T(0);
    k.kyh = nyy ? T(0) : (T)k.kyf;
    k.kyn = nyy ? (T)k.kyf : T(0);
    const int kyo = (ky <= a.My - ky) ? ky : a.My - ky;
    if (cv && a.G) k.grow = a.G + kyo * a.g_ky + kx * a.g_kx;
    return k;
}

// ik_in mode `mode` (PencilArgs::ik_in) on one thread's loaded values v[u]
// (positions t + TP u)
template <typename T, int E>
__device__ __forceinline__ void apply_ik_mode(cplx<T> (&v)[E], const PencilArgs<T>& a, int mode, int col, int t,
                                              int TP)
{
    if (mode == 0) return;
    if (mode == 1) {
        const int kx = col + a.col0;
        const T k = (2 * kx == a.Mx) ? T(0) : (T)kappa_full(kx, a.Mx, a.dk[0]);
#pragma unroll
        for (int u = 0; u < E; ++u) v[u] = cplx<T>{-k * v[u].y, k * v[u].x};
    } else {
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int ky = t + TP * u;
            const T k = (2 * ky == a.My) ? T(0) : (T)kappa_full(ky, a.My, a.dk[1]);
            v[u] = cplx<T>{-k * v[u].y, k * v[u].x};
        }
    }
}

template <typename T, int E>
__device__ __forceinline__ void apply_ik_in(cplx<T> (&v)[E], const PencilArgs<T>& a, int col, int t, int TP)
{
    apply_ik_mode<T, E>(v, a, a.ik_in, col, t, TP);
}

// Signed wavenumber of vector component c under map `kmap` (KMap), from the
// axis wavenumbers kx, ky, kz; 0 for a component the map does not use.
template <typename T>
__device__ __forceinline__ T kcomp(int kmap, int c, T kx, T ky, T kz)
{
    if (kmap == KMAP_3D) return (c == 0) ? kx : (c == 1) ? ky : kz;
    if (kmap == KMAP_2D_CURL) return (c == 0) ? kz : (c == 1) ? -kx : T(0);
    if (kmap == KMAP_Z) return (c == 0) ? kz : T(0);
    return (c == 0) ? kx : (c == 1) ? kz : T(0);
}

// k-space operator of the fused z pass on one thread's E spectrum values
// (solve3d.cpp:329-396 in the Hermitian-part form of the half spectrum).
// Output o of the NC inputs v goes to w; positions are kz = t + TP u.
template <typename T, int E, int NC>
__device__ __forceinline__ void kspace_op(cplx<T> (&w)[E], const cplx<T> (&v)[NC][E], const PencilArgs<T>& a,
                                          const TileK<T>& tk, int o, int comp, int t, int TP)
{
    using C = cplx<T>;
    const int op = a.op;
    const int vi = (NC == 3) ? 1 : 0, wi = (NC == 3) ? 2 : 0;
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int kz = t + TP * u;
        if (op == OP_NONE) {
            w[u] = v[0][u];
            continue;
        }
        const int kzo = (kz <= a.Mz - kz) ? kz : a.Mz - kz;
        const T g = tk.grow ? tk.grow[kzo * a.g_kz] : T(0);
        if (op == OP_PLAIN) {
            w[u] = cscale(v[0][u], g);
            continue;
        }
        const double kzf = kappa_full(kz, a.Mz, a.dk[2]);
        const bool nyz = (2 * kz == a.Mz);
        const T kzh = nyz ? T(0) : (T)kzf;
        const T kzn = nyz ? (T)kzf : T(0);
        C r;
        if (op == OP_GRAD) {
            const T kd = kcomp<T>(a.kmap, comp, tk.kxh, tk.kyh, kzh);
            const C s = v[0][u];
            r = C{-kd * g * s.y, kd * g * s.x};
        } else if (op == OP_LAPL) {
            const T m = (T)(-(tk.kxf * tk.kxf + tk.kyf * tk.kyf + kzf * kzf)) * g;
            r = cscale(v[0][u], m);
        } else if (op == OP_DIV || op == OP_CURL) {
            const C sx = v[0][u], sy = v[vi][u], sz = v[wi][u];
            C q;
            if (op == OP_DIV) {
                const T k0 = kcomp<T>(a.kmap, 0, tk.kxh, tk.kyh, kzh);
                const T k1 = kcomp<T>(a.kmap, 1, tk.kxh, tk.kyh, kzh);
                const T k2 = kcomp<T>(a.kmap, 2, tk.kxh, tk.kyh, kzh);
                if (a.kmap == KMAP_2D_DIV)   // the third component is not loaded data
                    q = C{k0 * sx.x + k1 * sy.x, k0 * sx.y + k1 * sy.y};
                else
                    q = C{k0 * sx.x + k1 * sy.x + k2 * sz.x, k0 * sx.y + k1 * sy.y + k2 * sz.y};
            } else if (o == 0) {
                q = C{tk.kyh * sz.x - kzh * sy.x, tk.kyh * sz.y - kzh * sy.y};
            } else if (o == 1) {
                q = C{kzh * sx.x - tk.kxh * sz.x, kzh * sx.y - tk.kxh * sz.y};
            } else {
                q = C{tk.kxh * sy.x - tk.kyh * sx.x, tk.kxh * sy.y - tk.kyh * sx.y};
            }
            r = C{-g * q.y, g * q.x};   // g * i * q
        } else {
            // filtered (and reprojected) rhs: zeta(|k|) / N
            double zf = a.invN;
            if (a.reg > 0) {
                const double arg = a.zarg * (tk.kxf * tk.kxf + tk.kyf * tk.kyf + kzf * kzf);
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
                const C sx = v[0][u], sy = v[vi][u], sz = v[wi][u];
                const C dh = C{tk.kxh * sx.x + tk.kyh * sy.x + kzh * sz.x, tk.kxh * sx.y + tk.kyh * sy.y + kzh * sz.y};
                const C dn = C{tk.kxn * sx.x + tk.kyn * sy.x + kzn * sz.x, tk.kxn * sx.y + tk.kyn * sy.y + kzn * sz.y};
                const T ah = (o == 0) ? tk.kxh : (o == 1) ? tk.kyh : kzh;
                const T an = (o == 0) ? tk.kxn : (o == 1) ? tk.kyn : kzn;
                r = C{r.x - g * (ah * dh.x + an * dn.x), r.y - g * (ah * dh.y + an * dn.y)};
            }
        }
        w[u] = r;
    }
}

// One-component operators (plain, grad, filtered rhs) in place on v.  The
// spectrum index of v[u] is kzm (t + TP u) + kza (kzm 2: one residue of a
// transform split into even and odd frequencies).
template <typename T, int E>
__device__ __forceinline__ void kspace_op_inplace(cplx<T> (&v)[E], const PencilArgs<T>& a, const TileK<T>& tk,
                                                  int comp, int t, int TP, int kzm = 1, int kza = 0)
{
    using C = cplx<T>;
    const int op = a.op;
    if (op == OP_NONE) return;
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int kz = kzm * (t + TP * u) + kza;
        const int kzo = (kz <= a.Mz - kz) ? kz : a.Mz - kz;
        const T g = tk.grow ? tk.grow[kzo * a.g_kz] : T(0);
        if (op == OP_PLAIN) {
            v[u] = cscale(v[u], g);
        } else if (op == OP_LAPL) {
            const double kzf = kappa_full(kz, a.Mz, a.dk[2]);
            v[u] = cscale(v[u], (T)(-(tk.kxf * tk.kxf + tk.kyf * tk.kyf + kzf * kzf)) * g);
        } else if (op == OP_GRAD) {
            const T kzh = (2 * kz == a.Mz) ? T(0) : (T)kappa_full(kz, a.Mz, a.dk[2]);
            const T kd = kcomp<T>(a.kmap, comp, tk.kxh, tk.kyh, kzh);
            const C s = v[u];
            v[u] = C{-kd * g * s.y, kd * g * s.x};
        } else {
            const double kzf = kappa_full(kz, a.Mz, a.dk[2]);
            double zf = a.invN;
            if (a.reg > 0) {
                const double arg = a.zarg * (tk.kxf * tk.kxf + tk.kyf * tk.kyf + kzf * kzf);
                double sum = 0.0, term = 1.0;
                for (int n = 0; n < a.reg / 2; ++n) {
                    sum += term;
                    term = term * arg / (double)(n + 1);
                }
                zf *= sum * exp(-arg);
            }
            v[u] = cscale(v[u], (T)zf);
        }
    }
}

// ---------------------------------------------------------------------------
// strided pencil pass (y passes, fused z pass, DCT-I of the kernel octant)
// NC = 1: one component in, one out (launch component = blockIdx range);
// NC = 2: one component in, two out (a 2D grad or curl of one scalar: the
//         scalar is loaded and transformed once, output o is component o);
// NC = 3: three components in, one (div) or three out.
// ---------------------------------------------------------------------------
template <typename T, int E, int NC>
__global__ void __launch_bounds__(kMaxThreads) k_pencil(PencilArgs<T> a)
{
    constexpr int NV = (NC == 3) ? 3 : 1;   // input pencils held
    using C = cplx<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);

    const int B = a.B, TP = a.plan.TP, L = a.plan.L;
    const int tid = threadIdx.x;
    const int p = tid % B, t = tid / B;
    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col = ct * B + p;
    const bool cv = col < a.ncol;
    C* my = lds + p;

    C v[NV][E];
#pragma unroll
    for (int c = 0; c < NV; ++c) {
        const int cc = (NC == 3) ? c : comp;
        const long base = cc * a.ia.comp + col * a.ia.col + boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi);
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int pos = t + TP * u;
            C x = C{0, 0};
            if (cv && pos < a.n_in) {
                if (a.load_mode == LOAD_COMPLEX) {
                    x = reinterpret_cast<const C*>(a.in)[base + boff(pos, a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi)];
                } else {
                    const int n = (pos <= L - pos) ? pos : L - pos;
                    x.x = reinterpret_cast<const T*>(a.in)[base + boff(n, a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi)];
                }
            }
            v[c][u] = x;
        }
    }
    if (a.do_fwd) {
#pragma unroll
        for (int c = 0; c < NV; ++c) fft_pencil<E, -1, T>(v[c], a.plan, my, B, t);
    } else if constexpr (NC == 1) {
        apply_ik_in<T, E>(v[0], a, col, t, TP);
    }

    const int op = a.op;
    const int nout = (NC == 3) ? (op == OP_DIV ? 1 : 3) : NC;
    const TileK<T> tk = tile_k(a, col, slab, cv);

    // o is a compile-time index in each unrolled copy, so v[o] stays in registers
#pragma unroll
    for (int o = 0; o < NC; ++o) {
        if (o >= nout) break;
        C w[E];
        kspace_op<T, E, NV>(w, v, a, tk, o, (NC == 2) ? o : comp, t, TP);
        if (a.do_inv) fft_pencil<E, +1, T>(w, a.plan, my, B, t);

        const int oc = (NC == 1) ? comp : o;
        const long obase = oc * a.oa.comp + col * a.oa.col + boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi);
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int pos = t + TP * u;
            if (cv && pos < a.n_out) {
                if (a.store_mode == STORE_COMPLEX)
                    reinterpret_cast<C*>(a.out)[obase + boff(pos, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi)] = w[u];
                else
                    reinterpret_cast<T*>(a.out)[obase + boff(pos, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi)] = w[u].x;
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
    const long nwork = ntile * (NC == 3 ? 1 : a.ncomp);
    hipLaunchKernelGGL((k_pencil<T, E, NC>), dim3((unsigned)nwork), dim3(a.plan.TP * a.B), shm, s, a);
    return hipGetLastError();
}

}  // namespace psamd
