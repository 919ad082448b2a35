// Pass kernels specialised at compile time for one transform length L
// (instantiated by skern.hip, once per L).  Same data layouts, arguments and
// results as the runtime-plan kernels of ps_kernels_impl.hpp.
#pragma once
#include "ps_kernels_impl.hpp"
#include "ps_sfft.hpp"
#include "ps_static.hpp"

namespace psamd {

// ---------------------------------------------------------------------------
// x pass, forward (two real rows per complex pencil).  HALF: rows hold L/2
// values followed by zero padding (unbounded x).
// ---------------------------------------------------------------------------
// Waves per SIMD that the x kernels' LDS allows (at most 4), the occupancy
// bound handed to the compiler: it then keeps registers it would otherwise
// spill to reach waves that cannot be resident (FP64 pencils of 4096 and 8192
// hold 67 and 135 KB of LDS per row pair).
template <typename T, int L> constexpr int s_x_minw()
{
    using Pol = SPolicy<T, L>;
    const int lds = Pol::P * s_lds_len(L) * 2 * (int)sizeof(T);
    const int waves = ((160 * 1024) / lds) * ((Pol::TP * Pol::P + 63) / 64) / 4;
    return waves < 1 ? 1 : (waves > 4 ? 4 : waves);
}

// The inverse x kernel of FP64 pencils of 8192 reads its stage twiddles from
// the table: held in registers they take about 150 VGPRs besides the pencil's
// 64 and the kernel spilled (x^-1 of a 4096^2 2D solve 0.446 -> 0.335 ms).  The
// forward kernel, bounded to the occupancy its LDS allows, keeps them in
// registers (table twiddles measured 0.145 -> 0.163 ms there).
template <typename T, int L> constexpr bool s_xi_twreg() { return !(sizeof(T) == 8 && L >= 8192); }

template <typename T, int L, int HALF>
__global__ void __launch_bounds__(kMaxThreads, (s_x_minw<T, L>())) k_x_r2c_s(XArgs<T> a)
{
    using C = cplx<T>;
    using Pol = SPolicy<T, L>;
    constexpr int E = Pol::E, TP = Pol::TP, P = Pol::P;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int pp = threadIdx.x / TP, t = threadIdx.x % TP;
    const long pair = (long)blockIdx.x * P + pp;
    const int comp = blockIdx.y;
    const long ra = 2 * pair, rb = ra + 1;
    const bool va = ra < a.nrows, vb = rb < a.nrows;
    const T* in = reinterpret_cast<const T*>(a.in) + comp * a.r_comp;

    STw<L, E, T> W;
    load_stwiddles<L, E, T>(W, t, a.plan.tw);
    C v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int pos = t + TP * u;
        C x = C{0, 0};
        if (!(HALF && u >= E / 2) && (HALF || pos < a.nvalid)) {
            if (va) x.x = in[ra * a.r_row + pos];
            if (vb) x.y = in[rb * a.r_row + pos];
        }
        v[u] = x;
    }
    C* my = lds + pp * lds_len(L);
    constexpr int WS = (64 % TP == 0) ? 1 : 0;   // each pencil inside one wavefront
    sfft<L, E, 1, -1, T, WS>(v, W, my, t);

    if constexpr (TP % 32 == 0) {
        C* w0 = my + lds_pos(t);
#pragma unroll
        for (int u = 0; u < E; ++u) w0[u * (TP + TP / 32)] = v[u];
    } else {
#pragma unroll
        for (int u = 0; u < E; ++u) my[lds_pos(t + TP * u)] = v[u];
    }
    xsync<WS>();
    C* out = reinterpret_cast<C*>(a.out) + comp * a.c_comp;
    const T half = T(0.5);
#pragma unroll
    for (int u = 0; u <= E / 2; ++u) {
        const int pos = t + TP * u;
        if (u < E / 2 || pos == L / 2) {
            const int mp = (pos == 0) ? 0 : L - pos;
            const C zk = v[u], zm = my[lds_pos(mp)];
            if (va) out[ra * a.c_row + pos] = C{half * (zk.x + zm.x), half * (zk.y - zm.y)};
            if (vb) out[rb * a.c_row + pos] = C{half * (zk.y + zm.y), half * (zm.x - zk.x)};
        }
    }
}

// ---------------------------------------------------------------------------
// x pass, inverse.  HALF: only the first L/2 outputs are stored (truncation).
// FAN: the workgroups of grid component a.dx_also make a second pass (a
// separate instantiation: the loop costs registers in the one-pass kernel).
// ---------------------------------------------------------------------------
template <typename T, int L, int HALF, int FAN = 0>
__global__ void __launch_bounds__(kMaxThreads) k_x_c2r_s(XArgs<T> a)
{
    using C = cplx<T>;
    using Pol = SPolicy<T, L>;
    constexpr int E = Pol::E, TP = Pol::TP, P = Pol::P;
    constexpr int H = L / 2 + 1;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);
    const int pp = threadIdx.x / TP, t = threadIdx.x % TP;
    const long pair = (long)blockIdx.x * P + pp;
    const int comp = blockIdx.y;
    const long ra = 2 * pair, rb = ra + 1;
    const bool va = ra < a.nrows, vb = rb < a.nrows;
    const C* in = reinterpret_cast<const C*>(a.in) + (a.remap ? a.src[comp] : comp) * a.c_comp;
    // the workgroups of grid component dx_also make a second pass over their
    // rows (re-read from the cache) for d/dx of them
    const int npass = (FAN && a.remap && comp == a.dx_also) ? 2 : 1;

    STw<L, E, T> W;
    if constexpr (s_xi_twreg<T, L>()) load_stwiddles<L, E, T>(W, t, a.plan.tw);
    C* my = lds + pp * lds_len(L);
    for (int pass = 0; pass < npass; ++pass) {
    const bool dx = pass == 1 || ((a.ikx >> comp) & 1);
    const int oc = !a.remap ? comp : (pass ? a.dx_slot : a.oslot[comp]);
    C v[E];
#pragma unroll
    for (int u = 0; u < E; ++u) {
        const int pos = t + TP * u;
        C za = C{0, 0}, zb = C{0, 0};
        if (pos < H) {
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
    if (dx) {
        // the derivative's spectrum i kh(n) X[n] (kh 0 at the Nyquist index)
        // and its Hermitian extension -i kh(m) conj(X[m]) for both rows: kh
        // is real and the same for the two rows packed in v, so v times i kk
        const T kst = (T)(2.0 * kPi * a.dk0);
#pragma unroll
        for (int u = 0; u < E; ++u) {
            const int pos = t + TP * u;
            const T kk = (pos < H) ? ((2 * pos == L) ? T(0) : kst * (T)pos) : -kst * (T)(L - pos);
            v[u] = C{-kk * v[u].y, kk * v[u].x};
        }
    }
    constexpr int WS = (64 % TP == 0) ? 1 : 0;
    if constexpr (s_xi_twreg<T, L>()) sfft<L, E, 1, +1, T, WS>(v, W, my, t);
    else sfft_g<L, E, 1, +1, T, WS>(v, a.plan.tw, my, t);
    T* out = reinterpret_cast<T*>(a.out) + oc * a.r_comp;
#pragma unroll
    for (int u = 0; u < (HALF ? E / 2 : E); ++u) {
        const int pos = t + TP * u;
        if (HALF || pos < a.nvalid) {
            if (va) out[ra * a.r_row + pos] = v[u].x;
            if (vb) out[rb * a.r_row + pos] = v[u].y;
        }
    }
    }
}

// ---------------------------------------------------------------------------
// strided pencil pass.  MODE: SM_FWD (y pass), SM_INV (y^-1), SM_FUSED (z pass).
// HIN: input holds L/2 positions followed by zero padding; HOUT: only the
// first L/2 positions are stored.
// ---------------------------------------------------------------------------
// MINW: occupancy target in waves per SIMD; 4 keeps a kernel within 128
// VGPRs (two 512-thread workgroups per CU), 2 allows 256.
template <typename T, int L, int NC, int MODE, int HIN, int HOUT, int TWREG, int MINW, int B>
__global__ void __launch_bounds__(kMaxThreads, MINW) k_pencil_s(PencilArgs<T> a)
{
    using C = cplx<T>;
    using Pol = SPolicy<T, L, NC>;
    constexpr int E = Pol::E, TP = Pol::TP;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);

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
    const bool full = (ct + 1) * B <= a.ncol;     // uniform: every column of the tile exists
    C* my = lds + p;

    STw<L, E, T> W;
    if constexpr (TWREG) load_stwiddles<L, E, T>(W, t, a.plan.tw);

    // position t + TP u sits at off(t) + u * ustride (TP is a multiple of the
    // position block, checked on the host)
    const long iu = (a.ia.pos_sh >= 30) ? TP * a.ia.pos_lo : (long)(TP >> a.ia.pos_sh) * a.ia.pos_hi;
    const long ou = (a.oa.pos_sh >= 30) ? TP * a.oa.pos_lo : (long)(TP >> a.oa.pos_sh) * a.oa.pos_hi;

    C v[NC][E];
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        const int cc = (NC == 1) ? comp : c;
        const C* ip = reinterpret_cast<const C*>(a.in) + cc * a.ia.comp + col * a.ia.col +
                      boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi) +
                      boff(t, a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi);
        if (full) {
#pragma unroll
            for (int u = 0; u < E; ++u) {
                if (HIN && u >= E / 2) v[c][u] = C{0, 0};
                else if (HIN || t + TP * u < a.n_in) v[c][u] = ip[u * iu];
                else v[c][u] = C{0, 0};
            }
        } else {
#pragma unroll
            for (int u = 0; u < E; ++u) {
                C x = C{0, 0};
                if (!(HIN && u >= E / 2) && cv && (HIN || t + TP * u < a.n_in)) x = ip[u * iu];
                v[c][u] = x;
            }
        }
    }
    if constexpr (MODE == SM_FWD || MODE == SM_FUSED) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            if constexpr (TWREG) sfft<L, E, B, -1, T>(v[c], W, my, t);
            else sfft_g<L, E, B, -1, T>(v[c], a.plan.tw, my, t);
        }
    } else if constexpr (NC == 1 && MODE == SM_INV) {
        apply_ik_in<T, E>(v[0], a, col, t, TP);
    }

    if constexpr (MODE == SM_INVFAN) {
        // one loaded spectrum, outputs for the set bits m of a.ik_fan (ik
        // mode m applied to a copy, then the inverse transform), written to
        // components fan_slot[m] of the output addressing
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            if (!((a.ik_fan >> m) & 1)) continue;
            C w[E];
#pragma unroll
            for (int u = 0; u < E; ++u) w[u] = v[0][u];
            apply_ik_mode<T, E>(w, a, m, col, t, TP);
            if constexpr (TWREG) sfft<L, E, B, +1, T>(w, W, my, t);
            else sfft_g<L, E, B, +1, T>(w, a.plan.tw, my, t);
            C* op_ = reinterpret_cast<C*>(a.out) + a.fan_slot[m] * a.oa.comp + col * a.oa.col +
                     boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi) +
                     boff(t, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi);
            if (cv) {
#pragma unroll
                for (int u = 0; u < (HOUT ? E / 2 : E); ++u)
                    if (HOUT || t + TP * u < a.n_out) op_[u * ou] = w[u];
            }
        }
        return;
    }

    const int nout = (NC == 1 || a.op == OP_DIV) ? 1 : 3;
    TileK<T> tk{};
    if constexpr (MODE == SM_FUSED) tk = tile_k(a, col, slab, cv);

    // o is a compile-time index in each unrolled copy, so v[o] stays in registers
#pragma unroll
    for (int o = 0; o < (NC == 1 ? 1 : 3); ++o) {
        if (o >= nout) break;
        C wbuf[NC == 1 ? 1 : E];
        // one-component operators work in place on v[0]
        C(&w)[E] = *reinterpret_cast<C(*)[E]>(NC == 1 ? &v[0][0] : &wbuf[0]);
        if constexpr (MODE == SM_FUSED) {
            if constexpr (NC == 1) kspace_op_inplace<T, E>(v[0], a, tk, comp, t, TP);
            else kspace_op<T, E, NC>(w, v, a, tk, o, comp, t, TP);
        }
        if constexpr (MODE != SM_FWD) {
            if constexpr (TWREG) sfft<L, E, B, +1, T>(w, W, my, t);
            else sfft_g<L, E, B, +1, T>(w, a.plan.tw, my, t);
        }

        const int oc = (NC == 1) ? comp : o;
        C* op_ = reinterpret_cast<C*>(a.out) + oc * a.oa.comp + col * a.oa.col +
                 boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi) + boff(t, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi);
        if (full) {
#pragma unroll
            for (int u = 0; u < (HOUT ? E / 2 : E); ++u)
                if (HOUT || t + TP * u < a.n_out) op_[u * ou] = w[u];
        } else if (cv) {
#pragma unroll
            for (int u = 0; u < (HOUT ? E / 2 : E); ++u)
                if (HOUT || t + TP * u < a.n_out) op_[u * ou] = w[u];
        }
    }
}

// ---------------------------------------------------------------------------
// Fused z pass with two thread mappings (one-component operators).
//   coalesced mapping  tid = t*B + p : global loads/stores of B adjacent columns
//   wavefront mapping  tid = p*TP + t: each pencil inside one wavefront
// The first forward stage runs in the coalesced mapping on the loaded data;
// its exchange (one workgroup barrier) hands each pencil to its wavefront;
// the remaining forward stages, the k-space operator and all inverse stages
// but the last exchange through the pencil's own LDS region with wavefront
// ordering only; the last exchange (second barrier) returns the coalesced
// mapping for the final inverse stage and the store.
// ---------------------------------------------------------------------------
// TWREG = 1: stage twiddles from the register array W; 0: from the table tw
template <int L, int E, int DIR, int S, int SEND, int TWREG, typename T>
__device__ __forceinline__ void wstages(cplx<T> (&v)[E], const STw<L, E, T>& W, const cplx<T>* __restrict__ tw,
                                        cplx<T>* reg, int t)
{
    if constexpr (S < SEND) {
        sstage<L, E, DIR, S, TWREG, T>(v, W, tw, t);
        if constexpr (S + 1 < SEND) sxchg<L, E, S, 1, T>(v, reg, t, reg, t);
        wstages<L, E, DIR, S + 1, SEND, TWREG, T>(v, W, tw, reg, t);
    }
}

// elements per thread of the pencil-per-wavefront z kernel: 16 (stage
// twiddles in registers); FP32 pencils of 2048 use 32 so that a pencil still
// fits one wavefront, with stage twiddles read from the table (58 twiddles
// would not fit in registers next to 32 points)
template <typename T, int L> constexpr int zwave_E()
{
    return (L > 1024 && sizeof(T) == 4) ? 32 : 16;
}

template <typename T, int L, int B, int HIN, int HOUT>
__global__ void __launch_bounds__(kMaxThreads, 2) k_zfused_w(PencilArgs<T> a)
{
    using C = cplx<T>;
    constexpr int E = zwave_E<T, L>(), TP = L / E;
    constexpr int TWR = (E == 16) ? 1 : 0;
    constexpr int NS = SFft<L, E>::cp.n;
    static_assert(NS >= 2, "two or more stages");
    static_assert(64 % TP == 0, "each pencil inside one wavefront");
    constexpr int LP = lds_len(L);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);

    const int tid = threadIdx.x;
    const int tl = tid / B, pl = tid % B;
    const int pc = tid / TP, tc = tid % TP;
    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col0 = ct * B;
    const bool full = col0 + B <= a.ncol;
    const bool cvl = col0 + pl < a.ncol, cvc = col0 + pc < a.ncol;

    STw<L, E, T> W;
    if constexpr (TWR) load_stwiddles<L, E, T>(W, tc, a.plan.tw);

    const long iu = (a.ia.pos_sh >= 30) ? TP * a.ia.pos_lo : (long)(TP >> a.ia.pos_sh) * a.ia.pos_hi;
    const long ou = (a.oa.pos_sh >= 30) ? TP * a.oa.pos_lo : (long)(TP >> a.oa.pos_sh) * a.oa.pos_hi;

    C v[E];
    {
        const C* ip = reinterpret_cast<const C*>(a.in) + comp * a.ia.comp + (col0 + pl) * a.ia.col +
                      boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi) +
                      boff(tl, a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi);
        const bool ok = full || cvl;
#pragma unroll
        for (int u = 0; u < E; ++u) {
            C x = C{0, 0};
            if (!(HIN && u >= E / 2) && ok && (HIN || tl + TP * u < a.n_in)) x = ip[u * iu];
            v[u] = x;
        }
    }
    sstage<L, E, -1, 0, TWR, T>(v, W, a.plan.tw, tl);
    sxchg<L, E, 0, 0, T>(v, lds + pl * LP, tl, lds + pc * LP, tc);
    C* reg = lds + pc * LP;
    wstages<L, E, -1, 1, NS, TWR, T>(v, W, a.plan.tw, reg, tc);

    const TileK<T> tk = tile_k(a, col0 + pc, slab, cvc);
    kspace_op_inplace<T, E>(v, a, tk, comp, tc, TP);

    wstages<L, E, +1, 0, NS - 1, TWR, T>(v, W, a.plan.tw, reg, tc);
    sxchg<L, E, NS - 2, 0, T>(v, reg, tc, lds + pl * LP, tl);
    sstage<L, E, +1, NS - 1, 0, T>(v, W, a.plan.tw, tl);

    if (full || cvl) {
        C* op_ = reinterpret_cast<C*>(a.out) + comp * a.oa.comp + (col0 + pl) * a.oa.col +
                 boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi) + boff(tl, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi);
#pragma unroll
        for (int u = 0; u < (HOUT ? E / 2 : E); ++u)
            if (HOUT || tl + TP * u < a.n_out) op_[u * ou] = v[u];
    }
}

// Fused z pass of a zero-padded pencil of 2 LH (unbounded z, nz = LH) as two
// LH-point transforms of the unpadded data: the even and odd frequencies are
//   X[2k + r] = FFT_LH(x[n] W_{2LH}^{r n})[k],  r = 0, 1,
// and the kept outputs (all n < LH) are
//   y[n] = sum_r W_{2LH}^{-r n} IFFT_LH(op(X[2k + r]))[n].
// Each residue runs the pencil-per-wavefront machinery at E = 16 (a pencil of
// LH = 1024 is one wavefront); the input and the sum stay in the coalesced
// mapping's registers.  a.plan.tw is the LH-point table; the residue twiddles
// W_{2LH}^n come from sincospi with an exact argument n / LH.
template <typename T> __device__ __forceinline__ void zres_sincospi(T x, T* s, T* c)
{
    if constexpr (sizeof(T) == 4) sincospif(x, s, c);
    else sincospi(x, s, c);
}

// cos / sin (2 pi u / 32), u < 16: the residue twiddles' constant factors
struct Trig32h {
    static constexpr double c[16] = {1.0, 9.80785280403230449126e-1, 9.23879532511286756128e-1, 8.31469612302545237079e-1, 7.07106781186547524401e-1, 5.55570233019602224743e-1, 3.82683432365089771728e-1, 1.95090322016128267848e-1, 0.0, -1.95090322016128267848e-1, -3.82683432365089771728e-1, -5.55570233019602224743e-1, -7.07106781186547524401e-1, -8.31469612302545237079e-1, -9.23879532511286756128e-1, -9.80785280403230449126e-1};
    static constexpr double s[16] = {0.0, 1.95090322016128267848e-1, 3.82683432365089771728e-1, 5.55570233019602224743e-1, 7.07106781186547524401e-1, 8.31469612302545237079e-1, 9.23879532511286756128e-1, 9.80785280403230449126e-1, 1.0, 9.80785280403230449126e-1, 9.23879532511286756128e-1, 8.31469612302545237079e-1, 7.07106781186547524401e-1, 5.55570233019602224743e-1, 3.82683432365089771728e-1, 1.95090322016128267848e-1};
};

// TWR = 1: stage twiddles in registers, 0: from the table.  The input held
// for the second residue and the sum take 64 registers beyond the one-pencil
// kernel, so the kernel runs at 2 waves per SIMD (at 4 it spills 450+ bytes
// per lane even when the input is read again instead of held).
template <typename T, int LH, int B, int TWR>
__global__ void __launch_bounds__(kMaxThreads, 2) k_zfused_r2(PencilArgs<T> a)
{
    using C = cplx<T>;
    constexpr int E = 16, TP = LH / E;
    constexpr int NS = SFft<LH, E>::cp.n;
    static_assert(NS >= 2, "two or more stages");
    static_assert(64 % TP == 0, "each pencil inside one wavefront");
    constexpr int LP = lds_len(LH);
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C* lds = reinterpret_cast<C*>(smem_raw);

    const int tid = threadIdx.x;
    const int tl = tid / B, pl = tid % B;
    const int pc = tid / TP, tc = tid % TP;
    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col0 = ct * B;
    const bool full = col0 + B <= a.ncol;
    const bool cvl = col0 + pl < a.ncol, cvc = col0 + pc < a.ncol;

    STw<LH, E, T> W;
    if constexpr (TWR) load_stwiddles<LH, E, T>(W, tc, a.plan.tw);

    const long iu = (a.ia.pos_sh >= 30) ? TP * a.ia.pos_lo : (long)(TP >> a.ia.pos_sh) * a.ia.pos_hi;
    const long ou = (a.oa.pos_sh >= 30) ? TP * a.oa.pos_lo : (long)(TP >> a.oa.pos_sh) * a.oa.pos_hi;

    C x[E], acc[E];
    {
        const C* ip = reinterpret_cast<const C*>(a.in) + comp * a.ia.comp + (col0 + pl) * a.ia.col +
                      boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi) +
                      boff(tl, a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi);
        const bool ok = full || cvl;
#pragma unroll
        for (int u = 0; u < E; ++u) x[u] = (ok && tl + TP * u < a.n_in) ? ip[u * iu] : C{0, 0};
    }
    const TileK<T> tk = tile_k(a, col0 + pc, slab, cvc);
    C* reg = lds + pc * LP;
    // W_{2LH}^n at n = tl + TP u is W_{2LH}^tl W_{2LH}^{TP u}, and TP u / 2LH = u / 32
    static_assert(2 * LH == 32 * TP, "residue twiddles are W_{2LH}^tl times 32nd roots");
    C wb;
    {
        T sn, cs;
        zres_sincospi<T>((T)tl / (T)LH, &sn, &cs);
        wb = C{cs, -sn};
    }
    // unrolled: the input x is dead once the second residue has read it, and
    // the sum is only the first residue's result until then (two arrays live)
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        C v[E];
#pragma unroll
        for (int u = 0; u < E; ++u)
            v[u] = r ? cmul(x[u], cmul(wb, C{(T)Trig32h::c[u], (T)-Trig32h::s[u]})) : x[u];
        if (r) __syncthreads();   // the previous residue's last reads of the coalesced regions are done
        sstage<LH, E, -1, 0, TWR, T>(v, W, a.plan.tw, tl);
        sxchg<LH, E, 0, 0, T>(v, lds + pl * LP, tl, reg, tc);
        wstages<LH, E, -1, 1, NS, TWR, T>(v, W, a.plan.tw, reg, tc);
        kspace_op_inplace<T, E>(v, a, tk, comp, tc, TP, 2, r);
        wstages<LH, E, +1, 0, NS - 1, TWR, T>(v, W, a.plan.tw, reg, tc);
        sxchg<LH, E, NS - 2, 0, T>(v, reg, tc, lds + pl * LP, tl);
        sstage<LH, E, +1, NS - 1, 0, T>(v, W, a.plan.tw, tl);   // coalesced slot: table twiddles
#pragma unroll
        for (int u = 0; u < E; ++u) {
            if (r) {
                const C w = cmul(wb, C{(T)Trig32h::c[u], (T)-Trig32h::s[u]});
                acc[u] = cadd(acc[u], cmul(v[u], C{w.x, -w.y}));   // W_{2LH}^{-n}
            } else {
                acc[u] = v[u];
            }
        }
    }

    if (full || cvl) {
        C* op_ = reinterpret_cast<C*>(a.out) + comp * a.oa.comp + (col0 + pl) * a.oa.col +
                 boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi) + boff(tl, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi);
#pragma unroll
        for (int u = 0; u < E; ++u)
            if (tl + TP * u < a.n_out) op_[u * ou] = acc[u];
    }
}

// Launch of k_zfused_r2 for pencils of 2 LH; tw_half is the LH-point table.
template <typename T, int LH> hipError_t s_launch_zres(const PencilArgs<T>& a, const void* tw_half, hipStream_t s)
{
    using Pol = SPolicy<T, LH>;
    PencilArgs<T> b = a;
    b.plan.tw = reinterpret_cast<const cplx<T>*>(tw_half);
    constexpr int TP = LH / 16;
    const size_t shm = (size_t)a.B * lds_len(LH) * sizeof(cplx<T>);
    const long nwork = (long)a.ncoltile * a.nslab * a.ncomp;
    hipError_t e;
#define PS_ZRES(BB, TW)                                                                                   \
    do {                                                                                                  \
        if ((e = allow_lds(k_zfused_r2<T, LH, BB, TW>, shm)) != hipSuccess) return e;                     \
        hipLaunchKernelGGL((k_zfused_r2<T, LH, BB, TW>), dim3((unsigned)nwork), dim3(TP * BB), shm, s, b); \
        return hipGetLastError();                                                                         \
    } while (0)
    if (a.B == Pol::B) {
        if (a.twreg) PS_ZRES(Pol::B, 1);
        PS_ZRES(Pol::B, 0);
    }
    if constexpr (Pol::B >= 2) {
        if (a.B == Pol::B / 2) {
            if (a.twreg) PS_ZRES(Pol::B / 2, 1);
            PS_ZRES(Pol::B / 2, 0);
        }
    }
#undef PS_ZRES
    return hipErrorInvalidValue;
}

template <typename T, int L, int B>
hipError_t s_launch_zwave_b(const PencilArgs<T>& a, int half, hipStream_t s)
{
    const size_t shm = (size_t)B * lds_len(L) * sizeof(cplx<T>);
    const long nwork = (long)a.ncoltile * a.nslab * a.ncomp;
    constexpr int TP = L / zwave_E<T, L>();
    hipError_t e;
    if (half) {
        if ((e = allow_lds(k_zfused_w<T, L, B, 1, 1>, shm)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_zfused_w<T, L, B, 1, 1>), dim3((unsigned)nwork), dim3(TP * B), shm, s, a);
    } else {
        if ((e = allow_lds(k_zfused_w<T, L, B, 0, 0>, shm)) != hipSuccess) return e;
        hipLaunchKernelGGL((k_zfused_w<T, L, B, 0, 0>), dim3((unsigned)nwork), dim3(TP * B), shm, s, a);
    }
    return hipGetLastError();
}

template <typename T, int L> hipError_t s_launch_zwave(const PencilArgs<T>& a, int half, hipStream_t s)
{
    using Pol = SPolicy<T, L>;
    constexpr int EW = zwave_E<T, L>();
    // a pencil must fit one wavefront with a power-of-two split (lengths
    // with a factor 3 or 5 have no pencil-per-wavefront form)
    if constexpr (L % EW == 0 && 64 % (L / EW) == 0 && s_plan_ok(L, EW)) {
        if constexpr (EW == 32) {
            // 2 wavefronts per SIMD (registers): one workgroup of 8 pencils per CU
            // (8 x 16.9 KB LDS) reads 64-byte row segments
            if constexpr (SFft<L, EW>::cp.n >= 2) {
                if (a.B == kZwaveB32) return s_launch_zwave_b<T, L, kZwaveB32>(a, half, s);
                if (a.B == kZwaveB32 / 2) return s_launch_zwave_b<T, L, kZwaveB32 / 2>(a, half, s);
            }
        } else if constexpr (SFft<L, EW>::cp.n >= 2) {
            if (a.B == Pol::B) return s_launch_zwave_b<T, L, Pol::B>(a, half, s);
            if constexpr (Pol::B >= 2) {
                if (a.B == Pol::B / 2) return s_launch_zwave_b<T, L, Pol::B / 2>(a, half, s);
            }
        }
    }
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
template <typename T, int L> hipError_t s_launch_x(const XArgs<T>& a, int inverse, int half, hipStream_t s)
{
    using Pol = SPolicy<T, L>;
    const size_t shm = (size_t)Pol::P * lds_len(L) * sizeof(cplx<T>);
    dim3 grid((unsigned)((a.npairs + Pol::P - 1) / Pol::P), (unsigned)a.ncomp), block(Pol::TP * Pol::P);
    hipError_t e;
    if (!inverse) {
        if (half) {
            if ((e = allow_lds(k_x_r2c_s<T, L, 1>, shm)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_x_r2c_s<T, L, 1>), grid, block, shm, s, a);
        } else {
            if ((e = allow_lds(k_x_r2c_s<T, L, 0>, shm)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_x_r2c_s<T, L, 0>), grid, block, shm, s, a);
        }
    } else if (a.remap && a.dx_also >= 0) {
        // the second pass for d/dx of one component's rows (FP32)
        if constexpr (sizeof(T) == 4) {
            if (half) {
                if ((e = allow_lds(k_x_c2r_s<T, L, 1, 1>, shm)) != hipSuccess) return e;
                hipLaunchKernelGGL((k_x_c2r_s<T, L, 1, 1>), grid, block, shm, s, a);
            } else {
                if ((e = allow_lds(k_x_c2r_s<T, L, 0, 1>, shm)) != hipSuccess) return e;
                hipLaunchKernelGGL((k_x_c2r_s<T, L, 0, 1>), grid, block, shm, s, a);
            }
        } else {
            return hipErrorInvalidValue;
        }
    } else {
        if (half) {
            if ((e = allow_lds(k_x_c2r_s<T, L, 1>, shm)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_x_c2r_s<T, L, 1>), grid, block, shm, s, a);
        } else {
            if ((e = allow_lds(k_x_c2r_s<T, L, 0>, shm)) != hipSuccess) return e;
            hipLaunchKernelGGL((k_x_c2r_s<T, L, 0>), grid, block, shm, s, a);
        }
    }
    return hipGetLastError();
}

template <typename T, int L, int NC, int MODE, int HIN, int HOUT, int TWREG, int MINW, int B>
hipError_t s_launch_pencil_b(const PencilArgs<T>& a, hipStream_t s)
{
    using Pol = SPolicy<T, L, NC>;
    if (a.B != B) return hipErrorInvalidValue;
    const size_t shm = (size_t)B * lds_len(L) * sizeof(cplx<T>);
    hipError_t e = allow_lds(k_pencil_s<T, L, NC, MODE, HIN, HOUT, TWREG, MINW, B>, shm);
    if (e != hipSuccess) return e;
    const long nwork = (long)a.ncoltile * a.nslab * (NC == 1 ? a.ncomp : 1);
    hipLaunchKernelGGL((k_pencil_s<T, L, NC, MODE, HIN, HOUT, TWREG, MINW, B>), dim3((unsigned)nwork),
                       dim3(Pol::TP * B), shm, s, a);
    return hipGetLastError();
}

// tile width: the policy's B, or half of it (more, smaller workgroups per CU)
template <typename T, int L, int NC, int MODE, int HIN, int HOUT, int TWREG, int MINW>
hipError_t s_launch_pencil_tw(const PencilArgs<T>& a, hipStream_t s)
{
    using Pol = SPolicy<T, L, NC>;
    if constexpr (Pol::B >= 2) {
        if (a.B == Pol::B / 2) return s_launch_pencil_b<T, L, NC, MODE, HIN, HOUT, TWREG, MINW, Pol::B / 2>(a, s);
    }
    if constexpr (Pol::B >= 4 && MODE == SM_FUSED) {
        if (a.B == Pol::B / 4) return s_launch_pencil_b<T, L, NC, MODE, HIN, HOUT, TWREG, MINW, Pol::B / 4>(a, s);
    }
    return s_launch_pencil_b<T, L, NC, MODE, HIN, HOUT, TWREG, MINW, Pol::B>(a, s);
}

template <typename T, int L, int NC, int MODE, int HIN, int HOUT>
hipError_t s_launch_pencil_one(const PencilArgs<T>& a, hipStream_t s)
{
    if constexpr (NC == 3) {
        if (a.twreg) return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 1, 2>(a, s);
        return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 0, 2>(a, s);
    } else if constexpr (MODE == SM_FUSED) {
        if (a.occ4) {
            if (a.twreg) return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 1, 4>(a, s);
            return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 0, 4>(a, s);
        }
        if (a.twreg) return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 1, 2>(a, s);
        return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 0, 2>(a, s);
    } else {
        if (a.twreg) return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 1, 4>(a, s);
        return s_launch_pencil_tw<T, L, NC, MODE, HIN, HOUT, 0, 4>(a, s);
    }
}

template <typename T, int L>
hipError_t s_launch_pencil(const PencilArgs<T>& a, int nc, int mode, int half, hipStream_t s)
{
    if (mode == SM_FWD) {
        if (nc != 1) return hipErrorInvalidValue;
        return half ? s_launch_pencil_one<T, L, 1, SM_FWD, 1, 0>(a, s) : s_launch_pencil_one<T, L, 1, SM_FWD, 0, 0>(a, s);
    }
    if (mode == SM_INV) {
        if (nc != 1) return hipErrorInvalidValue;
        return half ? s_launch_pencil_one<T, L, 1, SM_INV, 0, 1>(a, s) : s_launch_pencil_one<T, L, 1, SM_INV, 0, 0>(a, s);
    }
    if (mode == SM_INVFAN) {
        // FP32 only; stage twiddles from the table (the loaded spectrum and a
        // working copy take the registers)
        if constexpr (sizeof(T) == 4) {
            if (nc != 1 || a.ik_fan < 1 || a.ik_fan > 7) return hipErrorInvalidValue;
            return half ? s_launch_pencil_tw<T, L, 1, SM_INVFAN, 0, 1, 0, 2>(a, s)
                        : s_launch_pencil_tw<T, L, 1, SM_INVFAN, 0, 0, 0, 2>(a, s);
        }
        return hipErrorInvalidValue;
    }
    if (nc == 1 && a.zwave) return s_launch_zwave<T, L>(a, half, s);
    if (nc == 1)
        return half ? s_launch_pencil_one<T, L, 1, SM_FUSED, 1, 1>(a, s) : s_launch_pencil_one<T, L, 1, SM_FUSED, 0, 0>(a, s);
    return half ? s_launch_pencil_one<T, L, 3, SM_FUSED, 1, 1>(a, s) : s_launch_pencil_one<T, L, 3, SM_FUSED, 0, 0>(a, s);
}

}  // namespace psamd
