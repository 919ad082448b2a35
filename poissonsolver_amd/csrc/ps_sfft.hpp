// Compile-time specialised pencil FFT (the hot lengths).
//
// Same register/LDS scheme as ps_fft.hpp, but the length L, the elements per
// thread E, the radix sequence and the LDS position stride LS are template
// constants, so every stride, modulo and LDS offset folds to an immediate.
// The stage twiddles of a thread depend only on its pencil slot t; they are
// loaded once into registers (load_stwiddles) and reused by every transform
// the kernel performs (forward and inverse use conjugates of the same set).
// Inputs known to be zero (zero padding) and outputs never stored (truncation)
// are compile-time facts at the call site, so the first and last butterflies
// are pruned by constant folding and dead-code elimination.
#pragma once
#include "ps_fft.hpp"

namespace psamd {

// Orders this wavefront's LDS writes before its later LDS reads without a
// workgroup barrier (used when every pencil lives inside one wavefront).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int WS> __device__ __forceinline__ void xsync()
{
    if constexpr (WS) wave_sync();
    else __syncthreads();
}

struct CPlan {
    int n;
    int r[kMaxStage];
};

// Greedy radix split, identical to the host planner's for a given (L, E).
constexpr CPlan make_cplan(int L, int E)
{
    CPlan c{0, {0}};
    const int rads[7] = {16, 8, 7, 5, 4, 3, 2};
    int rem = L;
    while (rem > 1) {
        int pick = 0;
        for (int i = 0; i < 7; ++i) {
            if (E % rads[i] == 0 && rem % rads[i] == 0) {
                pick = rads[i];
                break;
            }
        }
        if (!pick || c.n >= kMaxStage) return CPlan{-1, {0}};
        c.r[c.n++] = pick;
        rem /= pick;
    }
    return c;
}

constexpr int cplan_ns(const CPlan& c, int s)
{
    int ns = 1;
    for (int i = 0; i < s; ++i) ns *= c.r[i];
    return ns;
}

// offset of stage s's twiddles in the per-thread twiddle array
constexpr int cplan_twoff(const CPlan& c, int E, int s)
{
    int off = 0;
    for (int i = 1; i < s; ++i) off += (c.r[i] - 1) * (E / c.r[i]);
    return off;
}

constexpr int cplan_ntw(const CPlan& c, int E) { return c.n > 1 ? cplan_twoff(c, E, c.n) : 0; }

template <int L, int E> struct SFft {
    static constexpr CPlan cp = make_cplan(L, E);
    static_assert(cp.n >= 0, "no radix split for this (L, E)");
    static constexpr int TP = L / E;
    static constexpr int NTW = cplan_ntw(cp, E) > 0 ? cplan_ntw(cp, E) : 1;
};

// per-thread stage twiddles of an (L, E) transform
template <int L, int E, typename T> using STw = cplx<T>[SFft<L, E>::NTW];

template <int L, int E, int S, typename T>
__device__ __forceinline__ void load_stw_from(cplx<T>* W, int t, const cplx<T>* __restrict__ tw)
{
    constexpr CPlan cp = SFft<L, E>::cp;
    if constexpr (S < cp.n) {
        constexpr int R = cp.r[S];
        constexpr int Ns = cplan_ns(cp, S);
        constexpr int NB = E / R;
        constexpr int TP = L / E;
        constexpr int tws = L / (Ns * R);
        constexpr int off = cplan_twoff(cp, E, S);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int k = (t + i * TP) % Ns;
#pragma unroll
            for (int m = 1; m < R; ++m) W[off + i * (R - 1) + m - 1] = tw[m * k * tws];
        }
        load_stw_from<L, E, S + 1, T>(W, t, tw);
    }
}

template <int L, int E, typename T>
__device__ __forceinline__ void load_stwiddles(STw<L, E, T>& W, int t, const cplx<T>* tw)
{
    load_stw_from<L, E, 1, T>(W, t, tw);
}

// TWREG = 1: stage twiddles come from the register array W (loaded once);
// TWREG = 0: they are read from the global table `tw` when a stage needs them,
// which frees the registers W would hold.
template <int L, int E, int LS, int DIR, int S, int TWREG, int WS, typename T>
__device__ __forceinline__ void sfft_from(cplx<T> (&v)[E], const STw<L, E, T>& W, const cplx<T>* __restrict__ tw,
                                          cplx<T>* lds, int t)
{
    using C = cplx<T>;
    constexpr CPlan cp = SFft<L, E>::cp;
    if constexpr (S < cp.n) {
        constexpr int R = cp.r[S];
        constexpr int Ns = cplan_ns(cp, S);
        constexpr int NB = E / R;
        constexpr int TP = L / E;
        constexpr int off = cplan_twoff(cp, E, S);
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            C a[R];
#pragma unroll
            for (int m = 0; m < R; ++m) a[m] = v[i + m * NB];
            if constexpr (S > 0) {
                const int k = (t + i * TP) % Ns;
#pragma unroll
                for (int m = 1; m < R; ++m) {
                    C w;
                    if constexpr (TWREG) w = W[off + i * (R - 1) + m - 1];
                    else w = tw[m * k * (L / (Ns * R))];
                    if constexpr (DIR > 0) w.y = -w.y;
                    a[m] = cmul(a[m], w);
                }
            }
            dft<R, DIR>(a);
#pragma unroll
            for (int q = 0; q < R; ++q) v[i + q * NB] = a[q];
        }
        if constexpr (S + 1 < cp.n) {
#pragma unroll
            for (int i = 0; i < NB; ++i) {
                const int b = t + i * TP;
                const int k = b % Ns;
                const int base = (b - k) * R + k;
                if constexpr (Ns % 32 == 0) {
                    // lds_pos(base + q Ns) = lds_pos(base) + q (Ns + Ns/32)
                    C* w0 = lds + lds_pos(base) * LS;
#pragma unroll
                    for (int q = 0; q < R; ++q) w0[q * (Ns + Ns / 32) * LS] = v[i + q * NB];
                } else {
#pragma unroll
                    for (int q = 0; q < R; ++q) lds[lds_pos(base + q * Ns) * LS] = v[i + q * NB];
                }
            }
            xsync<WS>();
            if constexpr (TP % 32 == 0) {
                // lds_pos(t + TP u) = lds_pos(t) + u (TP + TP/32): immediate offsets
                const C* r0 = lds + lds_pos(t) * LS;
#pragma unroll
                for (int u = 0; u < E; ++u) v[u] = r0[u * (TP + TP / 32) * LS];
            } else {
#pragma unroll
                for (int u = 0; u < E; ++u) v[u] = lds[lds_pos(t + TP * u) * LS];
            }
            xsync<WS>();
        }
        sfft_from<L, E, LS, DIR, S + 1, TWREG, WS, T>(v, W, tw, lds, t);
    }
}

// ---------------------------------------------------------------------------
// Stage-level pieces, for kernels that change the thread <-> pencil mapping
// between stages (sstage: butterflies of one stage; sxchg: its LDS exchange).
// ---------------------------------------------------------------------------
template <int L, int E, int DIR, int S, int TWREG, typename T>
__device__ __forceinline__ void sstage(cplx<T> (&v)[E], const STw<L, E, T>& W, const cplx<T>* __restrict__ tw,
                                       int t)
{
    using C = cplx<T>;
    constexpr CPlan cp = SFft<L, E>::cp;
    constexpr int R = cp.r[S];
    constexpr int Ns = cplan_ns(cp, S);
    constexpr int NB = E / R;
    constexpr int TP = L / E;
    constexpr int off = cplan_twoff(cp, E, S);
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        C a[R];
#pragma unroll
        for (int m = 0; m < R; ++m) a[m] = v[i + m * NB];
        if constexpr (S > 0) {
            const int k = (t + i * TP) % Ns;
#pragma unroll
            for (int m = 1; m < R; ++m) {
                C w;
                if constexpr (TWREG) w = W[off + i * (R - 1) + m - 1];
                else w = tw[m * k * (L / (Ns * R))];
                if constexpr (DIR > 0) w.y = -w.y;
                a[m] = cmul(a[m], w);
            }
        }
        dft<R, DIR>(a);
#pragma unroll
        for (int q = 0; q < R; ++q) v[i + q * NB] = a[q];
    }
}

// Write the outputs of stage S held by slot tw_ into the pencil at lw (unit
// position stride), synchronise (WS = 1: wavefront, 0: workgroup), then read
// positions tr + TP u of the pencil at lr.  With WS = 1 the caller guarantees
// that only this wavefront touches lw and lr.
template <int L, int E, int S, int WS, typename T>
__device__ __forceinline__ void sxchg(cplx<T> (&v)[E], cplx<T>* lw, int tw_, const cplx<T>* lr, int tr)
{
    using C = cplx<T>;
    constexpr CPlan cp = SFft<L, E>::cp;
    constexpr int R = cp.r[S];
    constexpr int Ns = cplan_ns(cp, S);
    constexpr int NB = E / R;
    constexpr int TP = L / E;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = tw_ + i * TP;
        const int k = b % Ns;
        const int base = (b - k) * R + k;
        if constexpr (Ns % 32 == 0) {
            C* w0 = lw + lds_pos(base);
#pragma unroll
            for (int q = 0; q < R; ++q) w0[q * (Ns + Ns / 32)] = v[i + q * NB];
        } else {
#pragma unroll
            for (int q = 0; q < R; ++q) lw[lds_pos(base + q * Ns)] = v[i + q * NB];
        }
    }
    xsync<WS>();
    if constexpr (TP % 32 == 0) {
        const C* r0 = lr + lds_pos(tr);
#pragma unroll
        for (int u = 0; u < E; ++u) v[u] = r0[u * (TP + TP / 32)];
    } else {
#pragma unroll
        for (int u = 0; u < E; ++u) v[u] = lr[lds_pos(tr + TP * u)];
    }
    if constexpr (WS) wave_sync();
}

// Transform one pencil held as v[u] = x[t + TP u]; on return v[u] = X[t + TP u].
// `lds` points at this pencil's first position; positions are LS elements apart.
// WS = 1: the pencil's TP threads are lanes of one wavefront, so exchanges
// need only wavefront ordering, not a workgroup barrier.
template <int L, int E, int LS, int DIR, typename T, int WS = 0>
__device__ __forceinline__ void sfft(cplx<T> (&v)[E], const STw<L, E, T>& W, cplx<T>* lds, int t)
{
    sfft_from<L, E, LS, DIR, 0, 1, WS, T>(v, W, nullptr, lds, t);
}

// same, twiddles read from the global table when needed
template <int L, int E, int LS, int DIR, typename T, int WS = 0>
__device__ __forceinline__ void sfft_g(cplx<T> (&v)[E], const cplx<T>* tw, cplx<T>* lds, int t)
{
    STw<L, E, T> dummy;
    sfft_from<L, E, LS, DIR, 0, 0, WS, T>(v, dummy, tw, lds, t);
}

}  // namespace psamd
