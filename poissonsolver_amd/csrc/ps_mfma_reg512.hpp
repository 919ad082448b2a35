// Mixed-precision z pass on the matrix cores, transform length L = 512 (Mz of
// a 256-cell unbounded z direction), register-pencil form.
//
// The four-step 512 = 32 x 16 transform, with n = b + 16 m (b < 16, m < 32)
// and k = k1 + 32 k2 (k1 < 32, k2 < 16):
//   X[k1 + 32 k2] = sum_b W16^(b k2) W512^(b k1) sum_{m<16} W32^(m k1) x[b + 16 m]
//   x[n2 + 16 n1] = sum_k1 W32^(-n1 k1) W512^(-n2 k1) sum_k2 W16^(-n2 k2) Z[k1 + 32 k2]
// (m >= 16 is padding; n1 < 16 is kept; the reference's solve3d.cpp:243-441
// at Mz = 512).  One wavefront owns one pencil through all four stages in
// registers, chained as in the kernel of 1024: stage 1 (data x table; 16
// inputs, even k1 by E16, odd by P), stage 2 (the 16-point DFT E16 x the
// twiddled stage-1 result), inverse stage 1 (the conjugated spectrum x E16),
// inverse stage 2 (E16 and Q x the twiddled result, 16 outputs): 8 products
// of 6 MFMAs, 48 per pencil, 8 complex values per lane.
//
// Tiles: 16 adjacent kx columns of one ky slab x 256 z rows, 32 KB, two LDS
// buffers in the 1024 kernel's staging layout; each wavefront stores and
// fetches two 1 KB blocks per tile, at the start and after stage 1.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_mfma_reg.hpp"

namespace psamd {

struct MfmaZR5 {
    static constexpr int L = 512, NZ = 256, P = 16, NW = 16;
    static constexpr int STG = NZ * P * 8;            // 32 KB per staging buffer
    static constexpr int TWB = 8 * 64 * 8;            // one twiddle table [slot][lane]
    static constexpr int FTB = 3 * 4 * 64 * 16;       // [E16, P, Q][rh, ih, rl, il][lane]
    static constexpr size_t lds_bytes() { return 2 * (size_t)STG + 2 * TWB + FTB; }
};

__device__ __forceinline__ float mr5_absmax4(const float2 (&u)[4])
{
    float m = __builtin_fabsf(u[0].x);
    asm("v_max3_f32 %0, |%1|, |%2|, %0\n\t"
        "v_max3_f32 %0, |%3|, |%4|, %0\n\t"
        "v_max3_f32 %0, |%5|, |%6|, %0\n\t"
        "v_max3_f32 %0, |%7|, |%7|, %0\n\t"
        "s_nop 1"
        : "+v"(m)
        : "v"(u[0].y), "v"(u[1].x), "v"(u[1].y), "v"(u[2].x), "v"(u[2].y), "v"(u[3].x), "v"(u[3].y));
    return m;
}

// Multiplier inputs of lane (j, g)'s 8 spectrum points kz = 2 j + tt + 32 (4 g + e)
// (tt = k1 & 1): octant index kz below Mz / 2 = 256 (g < 2) and 512 - kz from
// there (g >= 2), so the lane's offsets run up (g < 2) or down; kept raw as
// in the kernel of 1024.  The wavenumber index is kz, or kz - 512 for g >= 2.
template <int OPK>
__device__ __forceinline__ void mr5_kload(float (&gr)[2][4], const PencilArgs<float>& a, const MrTile& T, int col, int j,
                                          int g, int kbl)
{
    const TileK<float> tk = tile_k(a, col, T.slab, col < a.ncol);
    const int kb = 2 * j + 128 * g;
    if constexpr (mr_loads_g<OPK>()) {
        // kbl: kb, possibly an opaque copy that orders the loads
        const float* row = tk.grow ? tk.grow : a.G;
        const bool up = g < 2;
        const float* r = row + (up ? kbl : 512 - kbl);
        const int sgn = up ? 1 : -1;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) gr[tt][e] = r[sgn * (tt + 32 * e)];
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int tt = 0; tt < 2; ++tt)
                gr[tt][e] = (OPK == OP_NONE) ? 1.0f : mf_rhsf(a, tk, kb + tt + 32 * e);
    }
}

// multiplier of point (tt, e); kbf = kb (g < 2) or kb - 512 (g >= 2) as a
// float, so the wavenumber index is kbf plus a compile-time integer; the
// gradient's Nyquist point kz = 256 is tt = e = 0 on the lane with kb = 256
template <int OPK>
__device__ __forceinline__ float mr5_kmul(const MfKCoef& kc, const float (&gr)[2][4], int tt, int e, int kb, float kbf)
{
    const float v = gr[tt][e];
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL) {
        const float kzw = (kbf + (float)(tt + 32 * e)) * kc.kst;
        if constexpr (OPK == OP_GRAD) {
            const float kzh = (tt == 0 && e == 0 && kb == 256) ? 0.0f : kzw;
            return (kc.c0 + kc.c1 * kzh) * v;
        } else {
            return (kc.c0 - kzw * kzw) * v;
        }
    }
    return v;
}

// Transform front of one tile for wavefront w: slice 0 (the previous tile's
// blocks w and w + 16 of nxt stored, the next tile's fetched into them), the
// pencil's input rows from cur, stage 1, slice 1 and the W512^(b k1)
// twiddle; returns the rows' staging offsets sb, the input scale's inverse
// and the twiddled stage-1 result Y.
__device__ __forceinline__ void mr5_front(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                          unsigned char* __restrict__ nxt, const float2* __restrict__ tw1,
                                          const mf_h8* __restrict__ ft, const MrSliceOffs& so, char* obase, int nv,
                                          bool hasprev, const char* ibase, bool more, int w, int l, const MfT16& F16,
                                          unsigned (&sb)[4], float& inv0, mf_f4 (&yr)[2], mf_f4 (&yi)[2])
{
    const int g = l >> 4, j = l & 15;
    const int p = w;
    mr_slice(nxt, so, obase, nv, hasprev, ibase, more, w, 0);
    // staging offsets of rows z = j + 16 (4 g + e): block (j >> 3) + 8 g + 2 e,
    // row j & 7; n & 1 = (j >> 3) & 1, (n >> 1) & 1 = e & 1
#pragma unroll
    for (int e = 0; e < 4; ++e)
        sb[e] = (unsigned)(((j >> 3) + 8 * g + 2 * e) * 1024) + mr_stg_pos(j & 7, p >> 1, e & 1, (j >> 3) & 1) +
                (p & 1) * 8;
    // ---- input x[b + 16 m], b = j, m = 4 g + e ----
    float2 x[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) x[e] = *reinterpret_cast<const float2*>(cur + sb[e]);
    float s0;
    mr_scale(mr_wave_max(mr5_absmax4(x)), s0, inv0);
    // ---- stage 1 (data x table): Y[b][k1], rows b = 4 g + e, column k1 = 2 j + tk ----
    {
        const MfFrag X = mr_split4s<false>(x, s0, a.fzero);
        mr_mfma_a<true>(yr[0], yi[0], X, F16);
        mr_mfma_a<true>(yr[1], yi[1], X, mr_table(ft, 1, l));
    }
    mr_slice(nxt, so, obase, nv, hasprev, ibase, more, w, 1);
    // twiddle W512^(b k1), lane-major: slot 4 tk + e of lane (j, g) holds W512^((4 g + e)(2 j + tk))
    const float2* t1 = tw1 + l;
#pragma unroll
    for (int tk = 0; tk < 2; ++tk)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float2 v = cmul(float2{yr[tk][e], yi[tk][e]}, t1[(4 * tk + e) * 64]);
            yr[tk][e] = v.x;
            yi[tk][e] = v.y;
        }
}

// One tile T for wavefront w (pencil w): transform of column w of cur,
// written back in place; the previous tile's blocks w and w + 16 of nxt are
// stored and the next tile's fetched into them, one at the start and one
// after stage 1; the next tile's multiplier inputs are loaded after the last
// use of gm.  Nothing here waits for vector memory.
template <int OPK>
__device__ __forceinline__ void mr5_step(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                         unsigned char* __restrict__ nxt, const float2* __restrict__ tw1,
                                         const float2* __restrict__ tw2, const mf_h8* __restrict__ ft,
                                         const MrTile& T, const MrSliceOffs& so, char* obase, int nv, bool hasprev,
                                         const char* ibase, bool more, const MrTile& Tn, bool load_g,
                                         float (&gm)[2][4], int w, int l)
{
    constexpr bool mimag = mf_imag_multiplier<OPK>();
    const int g = l >> 4, j = l & 15;
    const int p = w, kb = 2 * j + 128 * g;
    unsigned sb[4];
    float inv0;
    mf_f4 yr[2], yi[2];
    // the 16-point DFT table (6 of the 8 products) held for the tile
    const MfT16 F16 = mr_table(ft, 0, l);
    mr5_front(a, cur, nxt, tw1, ft, so, obase, nv, hasprev, ibase, more, w, l, F16, sb, inv0, yr, yi);
    const float2* t2 = tw2 + l;
    mf_f4 ur[2], ui[2];
    float s2[2], inv2[2];
    MfKCoef kc{};
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL)
        kc = mf_kcoef<OPK>(a, tile_k(a, T.col0 + p, T.slab, T.col0 + p < a.ncol), T.comp);
    const float kbf = (float)(g < 2 ? kb : kb - 512);
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        // ---- stage 2 (E16 x data): X[k1 = 2 j + tk][k2 = 4 g + e] ----
        float2 v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = float2{yr[tk][e], yi[tk][e]};
        const MfFrag B = mr_split4u(v, a.fone);
        mf_f4 xr, xi;
        mr_mfma_b<true>(xr, xi, F16, B);
        float2 z[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m = mr5_kmul<OPK>(kc, gm, tk, e, kb, kbf);
            z[e] = mimag ? float2{-xi[e] * m, xr[e] * m} : float2{xr[e] * m, xi[e] * m};
        }
        if (tk == 1 && load_g) {
            const float2 zz[4] = {z[0], z[1], z[2], z[3]};
            mr5_kload<OPK>(gm, a, Tn, Tn.col0 + p, j, g, mr_after(kb, z, zz));
        }
        // ---- inverse stage 1 (conj Z x E16): U[k1][n2], rows k1 = 2 (4 g + e) + tk, column n2 = j ----
        mr_scale(mr_wave_max(mr5_absmax4(z)), s2[tk], inv2[tk]);
        {
            const MfFrag Z = mr_split4s<true>(z, s2[tk], a.fzero);
            mr_mfma_a<true>(ur[tk], ui[tk], Z, F16);
        }
        // T = W512^(n2 k1) V: slot 4 tk + e of lane (j, g) holds W512^(j (2 (4 g + e) + tk))
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float2 vv = cmul(float2{ur[tk][e], ui[tk][e]}, t2[(4 * tk + e) * 64]);
            ur[tk][e] = vv.x;
            ui[tk][e] = vv.y;
        }
    }
    // inverse stage 2 input scales from the bound (a 16-point DFT of inputs
    // below 2^10 and unit twiddles), as in the kernel of 1024
    const float smin = __builtin_fminf(s2[0], s2[1]);
    const float sa = 0.5f * smin * inv2[0], sb2 = 0.5f * smin * inv2[1];
    const float fo = 2.0f * inv0 * __builtin_fmaxf(inv2[0], inv2[1]);
    // ---- inverse stage 2 (table x data): x[n2 + 16 n1] = conj(sum_k1 W32^(n1 k1) T[k1][n2]), n1 < 16 ----
    {
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = float2{ur[0][e], ui[0][e]};
            v1[e] = float2{ur[1][e], ui[1][e]};
        }
        const MfFrag B0 = mr_split4s<false>(v0, sa, a.fzero), B1 = mr_split4s<false>(v1, sb2, a.fzero);
        mf_f4 dr, di;
        mr_mfma_b<true>(dr, di, F16, B0);
        mr_mfma_b<false>(dr, di, mr_table(ft, 2, l), B1);
        // rows z = j + 16 (4 g + e)
#pragma unroll
        for (int e = 0; e < 4; ++e)
            *reinterpret_cast<float2*>(cur + sb[e]) = float2{dr[e] * fo, -di[e] * fo};
    }
}

template <int OPK>
__global__ void __launch_bounds__(1024, 4) k_zmfma512r(PencilArgs<float> a)
{
    using M = MfmaZR5;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg0 = smem_raw;
    float2* tw1 = reinterpret_cast<float2*>(smem_raw + 2 * M::STG);
    float2* tw2 = reinterpret_cast<float2*>(smem_raw + 2 * M::STG + M::TWB);
    mf_h8* ftw = reinterpret_cast<mf_h8*>(smem_raw + 2 * M::STG + 2 * M::TWB);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l0 = tid & 63;
    const int nunit = mr_nunit(a);
    const int G = gridDim.x;

    if (tid < 512) {
        // slot 4 tk + e, lane 16 g + j: W512^((4 g + e)(2 j + tk)) and W512^(j (2 (4 g + e) + tk))
        const int slot = tid >> 6, j = tid & 15, g = (tid >> 4) & 3, tk = slot >> 2, e = slot & 3;
        tw1[tid] = a.plan.tw[((4 * g + e) * (2 * j + tk)) & (M::L - 1)];
        tw2[tid] = a.plan.tw[(j * (2 * (4 * g + e) + tk)) & (M::L - 1)];
    }
    if (tid < 3 * 4 * 64) {
        const int ln = tid & 63, f = (tid >> 6) & 3, tb = tid >> 8;
        const int fo = (f == 0) ? 0 : (f == 1) ? 2 : (f == 2) ? 1 : 3;
        ftw[(tb * 4 + fo) * 64 + ln] = reinterpret_cast<const mf_h8*>(a.mfrag16)[(tb * 64 + ln) * 4 + f];
    }
    const MrSliceOffs so = mr_slice_offs(a, w, l0);
    int u = blockIdx.x, m = 0;
    MrUnit U{};
    MrTile T{};
    const bool any = u < nunit && (U = mr_unit<M::P>(a, u, nunit), mr_tile(a, U, 0, T));
    float gm[2][4];
    if (any) {
        const char* ib = mr_tile_base(a.in, a.ia, T);
#pragma unroll
        for (int k = 0; k < 2; ++k)
            __builtin_amdgcn_global_load_lds(ib + so.iu[k] + so.ilane, (mr_lds_ptr)(stg0 + (w + 16 * k) * 1024), 16, 0, 0);
        mr5_kload<OPK>(gm, a, T, T.col0 + w, l0 & 15, l0 >> 4, 2 * (l0 & 15) + 128 * (l0 >> 4));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    MrTile Tp = T;
    int it = 0;
    for (; any; ++it) {
        int l = l0;
        asm volatile("" : "+v"(l));
        unsigned char* cur = stg0 + (it & 1) * M::STG;
        unsigned char* nxt = stg0 + ((it & 1) ^ 1) * M::STG;
        MrTile Tn{};
        MrUnit Un = U;
        int un = u, mn = 1;
        if (m == 1 || !mr_tile(a, U, 1, Tn)) {
            un = u + G;
            mn = 0;
        }
        const bool more = un < nunit && (mn == 1 || (Un = mr_unit<M::P>(a, un, nunit), mr_tile(a, Un, 0, Tn)));
        const bool load_g = more && !(mn == 1 && Tn.kyo == T.kyo);
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, Tp));
        const char* ibase = mr_tile_base(a.in, a.ia, Tn);
        const int nv = a.ncol - Tp.col0 - 2 * (l & 7);
        mr5_step<OPK>(a, cur, nxt, tw1, tw2, ftw, T, so, obase, nv, it > 0, ibase, more, Tn, load_g, gm, w, l);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        Tp = T;
        T = Tn;
        U = Un;
        u = un;
        m = mn;
        if (!more) break;
    }
    if (any) {
        unsigned char* last = stg0 + (it & 1) * M::STG;
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, Tp));
        const int nv = a.ncol - Tp.col0 - 2 * (l0 & 7);
#pragma unroll
        for (int k = 0; k < 2; ++k) mr_slice(last, so, obase, nv, true, nullptr, false, w, k);
    }
}


// ---- Several outputs from one forward transform (the potential form) ----
//
// MASK selects the z-pass outputs of one scalar's spectrum, in this order:
// bit 0 A = IFFT(Gk X), bit 1 d/dz A = IFFT(i kzh Gk X) (kzh: the Hermitian
// part, 0 at the Nyquist index), bit 2 the Laplacian IFFT(-|k|^2 Gk X).
// Output o (the o-th set bit) is component o of the output addressing.  The
// pencil is read and transformed forward once and the multiplier inputs are
// loaded once; each output then runs the k-space multiply and the two inverse
// stages from the held spectrum and is written to the tile buffer cur.  An
// output other than the last is stored from cur by the wavefronts' slices
// before the next output overwrites it; the last stays in cur and is stored
// during the next tile, as in the one-output kernel.

__device__ __forceinline__ void mr5_lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

constexpr int mr5_popc(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1); }

// one output of the held spectrum X (stage-2 results of k1 parities 0, 1):
// multiply by output OPK's multiplier, inverse stages 1 and 2, rows written
// to the pencil's column of cur (after a barrier when WBAR: the previous
// output's rows have then been read by every wavefront); LASTG: the last
// use of gm, after which the next tile's multiplier inputs are loaded
template <int OPK, bool WBAR, bool LASTG>
__device__ __forceinline__ void mr5_fan_out(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                            const float2* __restrict__ tw2, const mf_h8* __restrict__ ft,
                                            const MrTile& T, const MrTile& Tn, bool load_g, float (&gm)[2][4],
                                            const mf_f4 (&xr)[2], const mf_f4 (&xi)[2], float inv0,
                                            const unsigned (&sb)[4], int p, int l, const MfT16& F16)
{
    constexpr bool mimag = mf_imag_multiplier<OPK>();
    const int g = l >> 4, j = l & 15, kb = 2 * j + 128 * g;
    MfKCoef kc{};
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL)
        kc = mf_kcoef<OPK>(a, tile_k(a, T.col0 + p, T.slab, T.col0 + p < a.ncol), 0);
    const float kbf = (float)(g < 2 ? kb : kb - 512);
    const float2* t2 = tw2 + l;
    mf_f4 ur[2], ui[2];
    float s2[2], inv2[2];
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        float2 z[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m = mr5_kmul<OPK>(kc, gm, tk, e, kb, kbf);
            z[e] = mimag ? float2{-xi[tk][e] * m, xr[tk][e] * m} : float2{xr[tk][e] * m, xi[tk][e] * m};
        }
        if (LASTG && tk == 1 && load_g) {
            const float2 zz[4] = {z[0], z[1], z[2], z[3]};
            mr5_kload<OPK>(gm, a, Tn, Tn.col0 + p, j, g, mr_after(kb, z, zz));
        }
        mr_scale(mr_wave_max(mr5_absmax4(z)), s2[tk], inv2[tk]);
        {
            const MfFrag Z = mr_split4s<true>(z, s2[tk], a.fzero);
            mr_mfma_a<true>(ur[tk], ui[tk], Z, F16);
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float2 vv = cmul(float2{ur[tk][e], ui[tk][e]}, t2[(4 * tk + e) * 64]);
            ur[tk][e] = vv.x;
            ui[tk][e] = vv.y;
        }
    }
    const float smin = __builtin_fminf(s2[0], s2[1]);
    const float sa = 0.5f * smin * inv2[0], sb2 = 0.5f * smin * inv2[1];
    const float fo = 2.0f * inv0 * __builtin_fmaxf(inv2[0], inv2[1]);
    float2 v0[4], v1[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        v0[e] = float2{ur[0][e], ui[0][e]};
        v1[e] = float2{ur[1][e], ui[1][e]};
    }
    const MfFrag B0 = mr_split4s<false>(v0, sa, a.fzero), B1 = mr_split4s<false>(v1, sb2, a.fzero);
    mf_f4 dr, di;
    mr_mfma_b<true>(dr, di, F16, B0);
    mr_mfma_b<false>(dr, di, mr_table(ft, 2, l), B1);
    if constexpr (WBAR) mr5_lds_barrier();
#pragma unroll
    for (int e = 0; e < 4; ++e) *reinterpret_cast<float2*>(cur + sb[e]) = float2{dr[e] * fo, -di[e] * fo};
}

// after a non-last output: every row of it is in cur; wavefront w stores its
// blocks w and w + 16 to that output (ob: its tile's column 0)
__device__ __forceinline__ void mr5_fan_store(unsigned char* __restrict__ cur, const MrSliceOffs& so, char* ob,
                                              int nvt, int w)
{
    mr5_lds_barrier();
#pragma unroll
    for (int k = 0; k < 2; ++k) mr_slice(cur, so, ob, nvt, true, nullptr, false, w, k);
}

template <int MASK>
__device__ __forceinline__ void mr5_fan_step(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                             unsigned char* __restrict__ nxt, const float2* __restrict__ tw1,
                                             const float2* __restrict__ tw2, const mf_h8* __restrict__ ft,
                                             const MrTile& T, const MrSliceOffs& so, char* obase, int nv,
                                             bool hasprev, const char* ibase, bool more, const MrTile& Tn,
                                             bool load_g, float (&gm)[2][4], int w, int l, long ostride)
{
    constexpr int NO = mr5_popc(MASK);
    constexpr int OD = MASK & 1, OL = mr5_popc(MASK & 3);   // output index of d/dz and of the Laplacian
    unsigned sb[4];
    float inv0;
    mf_f4 yr[2], yi[2];
    const MfT16 F16 = mr_table(ft, 0, l);
    mr5_front(a, cur, nxt, tw1, ft, so, obase, nv, hasprev, ibase, more, w, l, F16, sb, inv0, yr, yi);
    // stage 2 of both k1 parities, held for every output
    mf_f4 xr[2], xi[2];
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        float2 v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = float2{yr[tk][e], yi[tk][e]};
        mr_mfma_b<true>(xr[tk], xi[tk], F16, mr_split4u(v, a.fone));
    }
    char* ob = const_cast<char*>(mr_tile_base(a.out, a.oa, T));
    const int nvt = a.ncol - T.col0 - 2 * (l & 7);
    if constexpr ((MASK & 1) != 0) {
        mr5_fan_out<OP_PLAIN, false, NO == 1>(a, cur, tw2, ft, T, Tn, load_g, gm, xr, xi, inv0, sb, w, l, F16);
        if constexpr (NO > 1) mr5_fan_store(cur, so, ob, nvt, w);
    }
    if constexpr ((MASK & 2) != 0) {
        mr5_fan_out<OP_GRAD, (OD > 0), (OD == NO - 1)>(a, cur, tw2, ft, T, Tn, load_g, gm, xr, xi, inv0, sb, w, l,
                                                       F16);
        if constexpr (OD < NO - 1) mr5_fan_store(cur, so, ob + OD * ostride, nvt, w);
    }
    if constexpr ((MASK & 4) != 0)
        mr5_fan_out<OP_LAPL, (OL > 0), true>(a, cur, tw2, ft, T, Tn, load_g, gm, xr, xi, inv0, sb, w, l, F16);
}

// the fan-out kernel: launched as the one-output kernel (persistent, slab
// pairs x 16-column tiles of the one input component); a.kmap = KMAP_Z so
// the gradient coefficient is that of k_z
template <int MASK>
__global__ void __launch_bounds__(1024, 4) k_zmfma512f(PencilArgs<float> a)
{
    using M = MfmaZR5;
    constexpr int NO = mr5_popc(MASK);
    // the multiplier inputs: the operator of the first output that loads Gk rows
    constexpr int OPG = (MASK & 1) ? OP_PLAIN : (MASK & 2) ? OP_GRAD : OP_LAPL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg0 = smem_raw;
    float2* tw1 = reinterpret_cast<float2*>(smem_raw + 2 * M::STG);
    float2* tw2 = reinterpret_cast<float2*>(smem_raw + 2 * M::STG + M::TWB);
    mf_h8* ftw = reinterpret_cast<mf_h8*>(smem_raw + 2 * M::STG + 2 * M::TWB);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l0 = tid & 63;
    const int nunit = mr_nunit(a);
    const int G = gridDim.x;
    const long ostride = a.oa.comp * 8;
    const long olast = (long)(NO - 1) * ostride;

    if (tid < 512) {
        const int slot = tid >> 6, j = tid & 15, g = (tid >> 4) & 3, tk = slot >> 2, e = slot & 3;
        tw1[tid] = a.plan.tw[((4 * g + e) * (2 * j + tk)) & (M::L - 1)];
        tw2[tid] = a.plan.tw[(j * (2 * (4 * g + e) + tk)) & (M::L - 1)];
    }
    if (tid < 3 * 4 * 64) {
        const int ln = tid & 63, f = (tid >> 6) & 3, tb = tid >> 8;
        const int fo = (f == 0) ? 0 : (f == 1) ? 2 : (f == 2) ? 1 : 3;
        ftw[(tb * 4 + fo) * 64 + ln] = reinterpret_cast<const mf_h8*>(a.mfrag16)[(tb * 64 + ln) * 4 + f];
    }
    const MrSliceOffs so = mr_slice_offs(a, w, l0);
    int u = blockIdx.x, m = 0;
    MrUnit U{};
    MrTile T{};
    const bool any = u < nunit && (U = mr_unit<M::P>(a, u, nunit), mr_tile(a, U, 0, T));
    float gm[2][4];
    if (any) {
        const char* ib = mr_tile_base(a.in, a.ia, T);
#pragma unroll
        for (int k = 0; k < 2; ++k)
            __builtin_amdgcn_global_load_lds(ib + so.iu[k] + so.ilane, (mr_lds_ptr)(stg0 + (w + 16 * k) * 1024), 16, 0, 0);
        mr5_kload<OPG>(gm, a, T, T.col0 + w, l0 & 15, l0 >> 4, 2 * (l0 & 15) + 128 * (l0 >> 4));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    MrTile Tp = T;
    int it = 0;
    for (; any; ++it) {
        int l = l0;
        asm volatile("" : "+v"(l));
        unsigned char* cur = stg0 + (it & 1) * M::STG;
        unsigned char* nxt = stg0 + ((it & 1) ^ 1) * M::STG;
        MrTile Tn{};
        MrUnit Un = U;
        int un = u, mn = 1;
        if (m == 1 || !mr_tile(a, U, 1, Tn)) {
            un = u + G;
            mn = 0;
        }
        const bool more = un < nunit && (mn == 1 || (Un = mr_unit<M::P>(a, un, nunit), mr_tile(a, Un, 0, Tn)));
        const bool load_g = more && !(mn == 1 && Tn.kyo == T.kyo);
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, Tp)) + olast;
        const char* ibase = mr_tile_base(a.in, a.ia, Tn);
        const int nv = a.ncol - Tp.col0 - 2 * (l & 7);
        mr5_fan_step<MASK>(a, cur, nxt, tw1, tw2, ftw, T, so, obase, nv, it > 0, ibase, more, Tn, load_g, gm, w, l,
                           ostride);
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        Tp = T;
        T = Tn;
        U = Un;
        u = un;
        m = mn;
        if (!more) break;
    }
    if (any) {
        unsigned char* last = stg0 + (it & 1) * M::STG;
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, Tp)) + olast;
        const int nv = a.ncol - Tp.col0 - 2 * (l0 & 7);
#pragma unroll
        for (int k = 0; k < 2; ++k) mr_slice(last, so, obase, nv, true, nullptr, false, w, k);
    }
}

}  // namespace psamd
