// Mixed-precision z pass on the matrix cores, transform length L = 2048
// (Mz of a 1024-cell unbounded z direction), register-pencil form.
//
// The zero-padded 2048-point transform of a pencil x (1024 points) is two
// 1024-point transforms, one per residue of the frequency mod 2:
//   X[2 k'] = DFT1024(x)[k'],   X[2 k' + 1] = DFT1024(W2048^n x)[k'],
// and the truncated inverse is y[n] = IDFT1024(Z0)[n] + W2048^-n IDFT1024(Z1)[n],
// n < 1024 (the reference's solve3d.cpp:243-441 at Mz = 2048).  Each residue
// is the four-stage transform of the register-pencil kernel of 1024 (same
// tables, twiddles and accumulator-as-operand chaining), run by a wavefront of
// its own: wavefronts w and w + 8 of a workgroup own residues 0 and 1 of
// pencil w, column w of an 8-column tile.  With 32 inputs, stage 1 splits the
// outputs by parity through the inputs (a = x[m] + x[m + 16] by E16 gives the
// even k1, d = x[m] - x[m + 16] by P the odd ones); with 32 outputs, inverse
// stage 2 keeps its even and odd products apart and forms rows n1 and n1 + 16
// as E + O and E - O.  The residue-1 wavefront multiplies its input by
// W2048^n and leaves its result in the tile buffer, where the residue-0
// wavefront multiplies it by W2048^-n and adds its own.
//
// Tiles: 8 adjacent kx columns of one ky slab (64-byte rows) x 1024 z rows,
// 64 KB in LDS; two buffers, stored and fetched in four slices per wavefront
// over the first half of each transform, as in the kernel of 1024.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_mfma_reg.hpp"

namespace psamd {

struct MfmaZR2 {
    static constexpr int L = 2048, NZ = 1024, P = 8, NW = 16;
    static constexpr int STG = NZ * P * 8;            // 64 KB per staging buffer
    static constexpr int TWB = 16 * 64 * 8;           // W1024 twiddles [slot][lane]
    static constexpr int TSB = 16 * 64 * 8;           // W2048^n of the lane's 16 positions [slot][lane]
    static constexpr int FTB = 3 * 4 * 64 * 16;       // [E16, P, Q][rh, ih, rl, il][lane]
    static constexpr size_t lds_bytes() { return 2 * (size_t)STG + TWB + TSB + FTB; }
};

// Staging layout: block n (1 KB, one direct-load instruction) holds rows
// 16 n .. 16 n + 15 in 16-byte slots; chunk c (columns 2 c, 2 c + 1) of row r
// sits in slot 4 ((r + rot) & 15) + (c ^ ((r >> 2) & 3)), rot = (n >> 3) & 1.
// A wavefront's column access (its 64 lanes read rows 2 j + q of blocks
// (j >> 3) + 8 g + ...) then covers 16 of the 16-byte positions of a 256-byte
// bank period, 4 lanes each: the fewest conflicts an 8-byte column allows.
__device__ __forceinline__ int mr2_swz(int r) { return (r >> 2) & 3; }
__device__ __forceinline__ unsigned mr2_slot(int r, int c, int rot) { return 4u * ((r + rot) & 15) + (c ^ mr2_swz(r)); }

// 64-bit byte offset of position z0 (a multiple of 16) and the 32-bit one of
// zr < 16 (plain positions or blocks of 2^sh >= 8 planes)
__device__ __forceinline__ long mr2_zoff_u(const Addr& d, int z0)
{
    return (d.pos_sh >= 30) ? (long)z0 * d.pos_lo * 8
                            : ((long)(z0 >> d.pos_sh) * d.pos_hi + (long)(z0 & ((1 << d.pos_sh) - 1)) * d.pos_lo) * 8;
}

// Per-lane offsets of the slices of wavefront w (its blocks n = w + 16 k all
// have rot = (w >> 3) & 1): the fetch lands lane l in slot l, so the lane
// loads the row and chunk that slot holds; the store writes chunk l & 3 of
// row l >> 2 from that chunk's slot.
struct Mr2SliceOffs {
    unsigned lds;      // LDS offset of the stored chunk within a block
    unsigned ilane;    // fetch: byte offset of the chunk the lane loads, within 16 rows
    unsigned olane;    // store: byte offset of chunk l & 3, within 16 rows
    int c;             // store: chunk index (columns 2 c, 2 c + 1)
};

__device__ __forceinline__ Mr2SliceOffs mr2_slice_offs(const PencilArgs<float>& a, int w, int l)
{
    Mr2SliceOffs o;
    const int rot = (w >> 3) & 1;
    const int r = l >> 2, c = l & 3;                   // store: row, chunk
    const int fr = ((l >> 2) - rot) & 15;              // fetch: the row slot l holds
    const int fc = (l & 3) ^ mr2_swz(fr);              // and its chunk
    o.c = c;
    o.lds = 16u * mr2_slot(r, c, rot);
    o.ilane = mr_zoff_l(a.ia, fr) + 16u * (unsigned)fc;
    o.olane = mr_zoff_l(a.oa, r) + 16u * (unsigned)c;
    return o;
}

// Slice k of wavefront w: store block n = w + 16 k of the previous tile (from
// buf, to obase: that tile's column 0, ncv valid columns) and fetch the next
// tile's block n (from ibase) into the same place.
__device__ __forceinline__ void mr2_slice(unsigned char* __restrict__ buf, const PencilArgs<float>& a,
                                          const Mr2SliceOffs& so, char* obase, int ncv, bool hasprev,
                                          const char* ibase, bool more, int w, int k)
{
    const int n = w + 16 * k;
    if (hasprev) {
        const float4 v = *reinterpret_cast<const float4*>(buf + n * 1024 + so.lds);
        char* o = obase + mr2_zoff_u(a.oa, 16 * n) + so.olane;
        const int nv = ncv - 2 * so.c;
        if (nv >= 2) *reinterpret_cast<float4*>(o) = v;
        else if (nv == 1) *reinterpret_cast<float2*>(o) = float2{v.x, v.y};
    }
    if (more)
        __builtin_amdgcn_global_load_lds(ibase + mr2_zoff_u(a.ia, 16 * n) + so.ilane, (mr_lds_ptr)(buf + n * 1024),
                                         16, 0, 0);
}

// Multiplier inputs of residue rs: lane (j, g)'s 16 spectrum points
// kz = 2 k' + rs, k' = k1 + 32 k2, k1 = 2 j + tt, k2 = 4 g + e + 16 h, i.e.
// kz = kb + 2 tt + 64 e + 1024 h with kb = 4 j + 256 g + rs; octant index kz
// (h = 0) or 2048 - kz = 1024 - (kz of h = 0) (h = 1), [ky][kx][kz] layout.
// Kept raw as for the kernel of 1024 (the operator's k-factor is applied
// where the multiplier is used).
template <int OPK>
__device__ __forceinline__ void mr2_kload(float (&gr)[2][4][2], const PencilArgs<float>& a, const MrTile& T, int col,
                                          int kb, int kbl)
{
    const TileK<float> tk = tile_k(a, col, T.slab, col < a.ncol);
    if constexpr (mr_loads_g<OPK>()) {
        // kbl: kb, possibly an opaque copy that orders the loads
        const float* row = tk.grow ? tk.grow : a.G;
        const float* r0 = row + kbl;
        const float* r1 = row + 1024 - kbl;
#pragma unroll
        for (int e = 0; e < 4; ++e)
#pragma unroll
            for (int tt = 0; tt < 2; ++tt) {
                gr[0][e][tt] = r0[2 * tt + 64 * e];
                gr[1][e][tt] = r1[-(2 * tt + 64 * e)];
            }
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int tt = 0; tt < 2; ++tt)
                    gr[h][e][tt] = (OPK == OP_NONE) ? 1.0f : mf_rhsf(a, tk, kb + 2 * tt + 64 * e + 1024 * h);
    }
}

// multiplier of point (tt, h, e): kz = kb + 2 tt + 64 e + 1024 h, taken as
// kz - Mz from Mz / 2 = 1024 on (the points of h = 1), formed from kbf = kb
// as for the kernel of 1024; the gradient's Nyquist point is h = 1,
// tt = e = 0 on the lane with kb = 0
template <int OPK>
__device__ __forceinline__ float mr2_kmul(const MfKCoef& kc, const float (&gr)[2][4][2], int tt, int h, int e, int kb,
                                          float kbf)
{
    const float v = gr[h][e][tt];
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL) {
        const float kzw = (kbf + (float)(2 * tt + 64 * e - 1024 * h)) * kc.kst;
        if constexpr (OPK == OP_GRAD) {
            const float kzh = (h == 1 && tt == 0 && e == 0 && kb == 0) ? 0.0f : kzw;
            return (kc.c0 + kc.c1 * kzh) * v;
        } else {
            return (kc.c0 - kzw * kzw) * v;
        }
    }
    return v;
}

// LDS-only workgroup barrier: this wavefront's LDS accesses complete, vector
// memory (the slices' stores and fetches) still in flight
__device__ __forceinline__ void mr2_lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// rows 2 j + TN + 32 (4 g + e + 16 h) of a pencil: y = y0 + W2048^-n y1 with
// this wavefront's own residue TN in mine and the other one's in the column
template <int TN>
__device__ __forceinline__ void mr2_combine(unsigned char* __restrict__ cur, unsigned sbt, const float2* __restrict__ tsl,
                                            const float2 (&mine)[2][4])
{
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            float2* q = reinterpret_cast<float2*>(cur + sbt + 2048 * e + 32768 * h);
            const float2 t = tsl[(8 * TN + 4 * h + e) * 64];
            const float2 other = *q;
            const float2 y0 = TN ? other : mine[h][e];
            const float2 y1 = TN ? mine[h][e] : other;
            *q = float2{y0.x + y1.x * t.x + y1.y * t.y, y0.y + y1.y * t.x - y1.x * t.y};
        }
}

// Transform front of one tile for wavefront w (pencil p = w & 7, residue
// rs = w >> 3): the pencil's input rows from cur (then a barrier: every
// wavefront has read its pencil before any result is written), slice 0,
// residue 1's W2048^n twist, the pencil's scale, stage 1, slice 1 and the
// W1024^(b k1) twiddle; returns the rows' staging offsets sb, the input
// scale's inverse and the twiddled stage-1 result Y.
template <bool H16>
__device__ __forceinline__ void mr2_front(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                          unsigned char* __restrict__ nxt, const float2* __restrict__ twm,
                                          const float2* __restrict__ tws, const mf_h8* __restrict__ ft,
                                          const Mr2SliceOffs& so, char* obase, int ncv, bool hasprev,
                                          const char* ibase, bool more, int w, int l, const MfT16& F16,
                                          unsigned (&sb)[2], float& inv0, mf_f4 (&yr)[2][2], mf_f4 (&yi)[2][2])
{
    const int g = l >> 4, j = l & 15;
    const int p = w & 7, rs = w >> 3;
    // staging offsets of rows z = 2 j + q + 32 (4 g + e + 16 h) (q = 0, 1):
    // block (j >> 3) + 8 g + 2 e + 32 h (rot = g & 1), row 2 (j & 7) + q
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = 2 * (j & 7) + q;
        sb[q] = (unsigned)(((j >> 3) + 8 * g) * 1024) + 16u * mr2_slot(r, p >> 1, g & 1) + (p & 1) * 8;
    }
    // ---- input x[b + 32 m], b = 2 j + tb, m = 4 g + e + 16 h ----
    float2 x[2][2][4];
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                x[tb][h][e] = *reinterpret_cast<const float2*>(cur + sb[tb] + 2048 * e + 32768 * h);
    // every wavefront has read its pencil before any residue result is written
    mr2_lds_barrier();
    mr2_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 0);
    const float2* twl = twm + l;
    const float2* tsl = tws + l;   // slot 8 tb + 4 h + e: W2048^(2 j + tb + 32 (4 g + e + 16 h))
    if (rs) {
#pragma unroll
        for (int tb = 0; tb < 2; ++tb)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e) x[tb][h][e] = cmul(x[tb][h][e], tsl[(8 * tb + 4 * h + e) * 64]);
    }
    // one scale for the pencil: |x[m] +- x[m + 16]| <= 2 max
    float s0;
    {
        const float m0 = mr_absmax8(x[0][0], x[0][1]), m1 = mr_absmax8(x[1][0], x[1][1]);
        mr_scale(2.0f * mr_wave_max(__builtin_fmaxf(m0, m1)), s0, inv0);
    }
    // ---- stage 1 (data x table): Y[b][k1] = sum_m W32^(m k1) x[b + 32 m] ----
    {
        const MfT16 Fe = H16 ? F16 : mr_table(ft, 0, l), Fo = mr_table(ft, 1, l);
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
            float2 xa[4], xd[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                xa[e] = float2{x[tb][0][e].x + x[tb][1][e].x, x[tb][0][e].y + x[tb][1][e].y};
                xd[e] = float2{x[tb][0][e].x - x[tb][1][e].x, x[tb][0][e].y - x[tb][1][e].y};
            }
            const MfFrag XA = mr_split4s<false>(xa, s0, a.fzero), XD = mr_split4s<false>(xd, s0, a.fzero);
            mr_mfma_a<true>(yr[tb][0], yi[tb][0], XA, Fe);
            mr_mfma_a<true>(yr[tb][1], yi[tb][1], XD, Fo);
        }
    }
    mr2_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 1);
    // twiddle W1024^(b k1), lane-major as in the kernel of 1024
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int tk = 0; tk < 2; ++tk)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float2 v = cmul(float2{yr[tb][tk][e], yi[tb][tk][e]}, twl[(8 * tk + 2 * e + tb) * 64]);
                yr[tb][tk][e] = v.x;
                yi[tb][tk][e] = v.y;
            }
}

// One tile T for wavefront w (pencil p = w & 7, residue rs = w >> 3), with the
// slices of the previous and next tiles over its first half and the next
// tile's multiplier inputs loaded after the last use of gm; the result of
// pencil p is left in column p of cur.  Nothing here waits for vector memory.
template <int OPK>
__device__ __forceinline__ void mr2_step(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                         unsigned char* __restrict__ nxt, const float2* __restrict__ twm,
                                         const float2* __restrict__ tws, const mf_h8* __restrict__ ft,
                                         const MrTile& T, const Mr2SliceOffs& so, char* obase, int ncv, bool hasprev,
                                         const char* ibase, bool more, const MrTile& Tn, bool load_g,
                                         float (&gm)[2][4][2], int w, int l)
{
    constexpr bool mimag = mf_imag_multiplier<OPK>();
    const int g = l >> 4, j = l & 15;
    const int p = w & 7, rs = w >> 3, kb = 4 * j + 256 * g + rs;
    unsigned sb[2];
    float inv0;
    mf_f4 yr[2][2], yi[2][2];
    // the 16-point DFT table (8 of the 16 products) held for the tile where
    // the registers allow
    constexpr bool hold16 = OPK == OP_PLAIN || OPK == OP_NONE;
    MfT16 F16{};
    if constexpr (hold16) F16 = mr_table(ft, 0, l);
    mr2_front<hold16>(a, cur, nxt, twm, tws, ft, so, obase, ncv, hasprev, ibase, more, w, l, F16, sb, inv0, yr, yi);
    const float2* twl = twm + l;
    const float2* tsl = tws + l;
    mf_f4 ur[2][2], ui[2][2];
    float s2[2], inv2[2];
    MfKCoef kc{};
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL)
        kc = mf_kcoef<OPK>(a, tile_k(a, T.col0 + p, T.slab, T.col0 + p < a.ncol), T.comp);
    const float kbf = (float)kb;
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        // ---- stage 2 (table x data) for column k1 = 2 j + tk ----
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = float2{yr[0][tk][e], yi[0][tk][e]};
            v1[e] = float2{yr[1][tk][e], yi[1][tk][e]};
        }
        const MfFrag B0 = mr_split4u(v0, a.fone), B1 = mr_split4u(v1, a.fone);
        mf_f4 er, ei, orr, oi;
        mr_mfma_b<true>(er, ei, hold16 ? F16 : mr_table(ft, 0, l), B0);
        mr_mfma_b<true>(orr, oi, mr_table(ft, 2, l), B1);
        // the k-space operator and inverse stage 1's butterflies in one step:
        // with X = E +- O at rows k2 and k2 + 16 and multipliers m0, m1 there,
        // A = X0 m0 + X1 m1 = E (m0 + m1) + O (m0 - m1) and
        // B = X0 m0 - X1 m1 = E (m0 - m1) + O (m0 + m1) (times i for an
        // imaginary multiplier)
        float2 za[4], zb[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m0 = mr2_kmul<OPK>(kc, gm, tk, 0, e, kb, kbf), m1 = mr2_kmul<OPK>(kc, gm, tk, 1, e, kb, kbf);
            const float mp = m0 + m1, mm = m0 - m1;
            const float ar = er[e] * mp + orr[e] * mm, ai = ei[e] * mp + oi[e] * mm;
            const float br = er[e] * mm + orr[e] * mp, bi = ei[e] * mm + oi[e] * mp;
            if constexpr (mimag) {
                za[e] = float2{-ai, ar};
                zb[e] = float2{-bi, br};
            } else {
                za[e] = float2{ar, ai};
                zb[e] = float2{br, bi};
            }
        }
        if (tk == 0) mr2_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 2);
        if (tk == 1 && load_g) mr2_kload<OPK>(gm, a, Tn, Tn.col0 + p, kb, mr_after(kb, za, zb));
        // ---- inverse stage 1 (conj Z x table), row k1 = 2 j + tk ----
        mr_scale(mr_wave_max(mr_absmax8(za, zb)), s2[tk], inv2[tk]);
        {
            const MfFrag XA = mr_split4s<true>(za, s2[tk], a.fzero), XB = mr_split4s<true>(zb, s2[tk], a.fzero);
            mr_mfma_a<true>(ur[tk][0], ui[tk][0], XA, hold16 ? F16 : mr_table(ft, 0, l));
            mr_mfma_a<true>(ur[tk][1], ui[tk][1], XB, mr_table(ft, 1, l));
        }
        if (tk == 0) mr2_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 3);
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float2 v = cmul(float2{ur[tk][tn][e], ui[tk][tn][e]}, twl[(8 * tn + 2 * e + tk) * 64]);
                ur[tk][tn][e] = v.x;
                ui[tk][tn][e] = v.y;
            }
    }
    // inverse stage 2 input scales from the bound, as in the kernel of 1024
    const float smin = __builtin_fminf(s2[0], s2[1]);
    const float sa = 0.5f * smin * inv2[0], sb2 = 0.5f * smin * inv2[1];
    const float fo = 2.0f * inv0 * __builtin_fmaxf(inv2[0], inv2[1]);
    // ---- inverse stage 2: y[32 n1 + n2] = conj(sum_k1 W32^(n1 k1) T[k1][n2]) for n1 < 32:
    // rows n1 = 4 g + e (E + O) and n1 + 16 (E - O) ----
    float2 o[2][2][4];   // [tn][h][e]: rows z = 2 j + tn + 32 (4 g + e + 16 h)
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = float2{ur[0][tn][e], ui[0][tn][e]};
            v1[e] = float2{ur[1][tn][e], ui[1][tn][e]};
        }
        const MfFrag B0 = mr_split4s<false>(v0, sa, a.fzero), B1 = mr_split4s<false>(v1, sb2, a.fzero);
        mf_f4 er, ei, orr, oi;
        mr_mfma_b<true>(er, ei, hold16 ? F16 : mr_table(ft, 0, l), B0);
        mr_mfma_b<true>(orr, oi, mr_table(ft, 2, l), B1);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            o[tn][0][e] = float2{(er[e] + orr[e]) * fo, -(ei[e] + oi[e]) * fo};
            o[tn][1][e] = float2{(er[e] - orr[e]) * fo, -(ei[e] - oi[e]) * fo};
        }
        if (tn != rs) {
            // the half of the rows the other residue's wavefront combines
            // (residue 1 hands over rows of even z, residue 0 those of odd z)
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    *reinterpret_cast<float2*>(cur + sb[tn] + 2048 * e + 32768 * h) = o[tn][h][e];
        }
    }
    mr2_lds_barrier();
    // y = y0 + W2048^-n y1 on the rows of z parity rs, the other residue's
    // half from the column
    if (rs) mr2_combine<1>(cur, sb[1], tsl, o[1]);
    else mr2_combine<0>(cur, sb[0], tsl, o[0]);
}

// Persistent kernel over units of two tiles (slabs ky, My - ky) of 8 columns,
// as the kernel of 1024.
template <int OPK>
__global__ void __launch_bounds__(1024, 4) k_zmfma2048r(PencilArgs<float> a)
{
    using M = MfmaZR2;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg0 = smem_raw;
    float2* twm = reinterpret_cast<float2*>(smem_raw + 2 * M::STG);
    float2* tws = reinterpret_cast<float2*>(smem_raw + 2 * M::STG + M::TWB);
    mf_h8* ftw = reinterpret_cast<mf_h8*>(smem_raw + 2 * M::STG + M::TWB + M::TSB);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l0 = tid & 63;
    const int nunit = mr_nunit(a);
    const int G = gridDim.x;

    {
        const int slot = tid >> 6, j = tid & 15, g = (tid >> 4) & 3;
        // W1024^(r s) = W2048^(2 r s), r = 2 j + t, s = 8 g + 2 e + u (slot 8 t + 2 e + u)
        const int r = 2 * j + (slot >> 3), s = 8 * g + (slot & 7);
        twm[tid] = a.plan.tw[(2 * r * s) & (M::L - 1)];
        // W2048^n, n = 2 j + tb + 32 (4 g + e + 16 h) (slot 8 tb + 4 h + e)
        const int n = 2 * j + (slot >> 3) + 32 * (4 * g + (slot & 3) + 16 * ((slot >> 2) & 1));
        tws[tid] = a.plan.tw[n];
    }
    if (tid < 3 * 4 * 64) {
        const int ln = tid & 63, f = (tid >> 6) & 3, tb = tid >> 8;
        const int fo = (f == 0) ? 0 : (f == 1) ? 2 : (f == 2) ? 1 : 3;
        ftw[(tb * 4 + fo) * 64 + ln] = reinterpret_cast<const mf_h8*>(a.mfrag16)[(tb * 64 + ln) * 4 + f];
    }
    const Mr2SliceOffs so = mr2_slice_offs(a, w, l0);
    int u = blockIdx.x, m = 0;
    MrUnit U{};
    MrTile T{};
    const bool any = u < nunit && (U = mr_unit<M::P>(a, u, nunit), mr_tile(a, U, 0, T));
    float gm[2][4][2];
    if (any) {
        const char* ib = mr_tile_base(a.in, a.ia, T);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int n = w + 16 * k;
            __builtin_amdgcn_global_load_lds(ib + mr2_zoff_u(a.ia, 16 * n) + so.ilane, (mr_lds_ptr)(stg0 + n * 1024),
                                             16, 0, 0);
        }
        const int kb0 = 4 * (l0 & 15) + 256 * (l0 >> 4) + (w >> 3);
        mr2_kload<OPK>(gm, a, T, T.col0 + (w & 7), kb0, kb0);
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
        const int ncv = a.ncol - Tp.col0;   // valid columns of the previous tile (may exceed 8)
        mr2_step<OPK>(a, cur, nxt, twm, tws, ftw, T, so, obase, ncv, it > 0, ibase, more, Tn, load_g, gm, w, l);
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
        const int ncv = a.ncol - Tp.col0;
#pragma unroll
        for (int k = 0; k < 4; ++k) mr2_slice(last, a, so, obase, ncv, true, nullptr, false, w, k);
    }
}


// ---- Several outputs from one forward transform (the potential form) ----
//
// MASK as for transforms of 512: bit 0 A (Gk), bit 1 d/dz A (i kzh Gk), bit 2
// the Laplacian (-|k|^2 Gk); output o (the o-th set bit) is component o of
// the output addressing.  An output other than the last is stored from the
// tile buffer by the wavefronts' slices (after a barrier) before the next
// one is written (after another); the last stays in the buffer and is stored
// during the next tile.

constexpr int mr2_popc(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1); }

// multiplier of point (tt, h, e) for fan-out output OPK: A Gk, d/dz A
// i kzh Gk (applied as kzh Gk times i), the Laplacian (c0 - kz^2) Gk with
// c0 = -(kx^2 + ky^2) of the pencil
template <int OPK>
__device__ __forceinline__ float mr2_fan_kmul(float c0, float kst, const float (&gr)[2][4][2], int tt, int h, int e,
                                              int kb, float kbf)
{
    const float v = gr[h][e][tt];
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL) {
        const float kzw = (kbf + (float)(2 * tt + 64 * e - 1024 * h)) * kst;
        if constexpr (OPK == OP_GRAD) return ((h == 1 && tt == 0 && e == 0 && kb == 0) ? 0.0f : kzw) * v;
        else return (c0 - kzw * kzw) * v;
    }
    return v;
}

// ---- The fan-out with one wavefront per pencil (both residues) ----
//
// Workgroups of 8 wavefronts (2 per SIMD, up to 256 VGPRs each) on the same
// 8-column tiles: wavefront w owns pencil w and its column of the tile, runs
// both residues' forward transforms from its one read of the pencil, holds
// both residues' stage-2 results and multiplier inputs for every output, and
// combines the residues in its own column (residue 0's rows are written
// there, then read back as residue 1's are added), so no residue hand-over
// barrier is needed.  The slices move blocks w + 8 k (k < 8) of a tile.
struct MfmaZR2P {
    static constexpr int NW = 8;
};

// slice k of wavefront w (block w + 8 k, rot k & 1) with the offsets so[k & 1]
__device__ __forceinline__ void mr2p_slice(unsigned char* __restrict__ buf, const PencilArgs<float>& a,
                                           const Mr2SliceOffs (&so)[2], char* obase, int ncv, bool hasprev,
                                           const char* ibase, bool more, int w, int k)
{
    mr2_slice(buf, a, so[k & 1], obase, ncv, hasprev, ibase, more, w + 8 * (k & 1), k >> 1);
}

// residue RS's forward transform of the pencil x (the FP16 split scales it
// by s0) to its stage-2 results E and O of both k1 parities; residue 1
// multiplies x by W2048^n first
template <int RS>
__device__ __forceinline__ void mr2p_forward(const PencilArgs<float>& a, const float2 (&x)[2][2][4], float s0,
                                             const float2* __restrict__ twl, const float2* __restrict__ tsl,
                                             const mf_h8* __restrict__ ft, const MfT16& F16, int l, mf_f4 (&er)[2],
                                             mf_f4 (&ei)[2], mf_f4 (&orr)[2], mf_f4 (&oi)[2])
{
    mf_f4 yr[2][2], yi[2][2];
    {
        const MfT16& Fe = F16;
        const MfT16 Fo = mr_table(ft, 1, l);
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
            float2 xa[4], xd[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                float2 x0 = x[tb][0][e], x1 = x[tb][1][e];
                if (RS) {
                    x0 = cmul(x0, tsl[(8 * tb + e) * 64]);
                    x1 = cmul(x1, tsl[(8 * tb + 4 + e) * 64]);
                }
                xa[e] = float2{x0.x + x1.x, x0.y + x1.y};
                xd[e] = float2{x0.x - x1.x, x0.y - x1.y};
            }
            const MfFrag XA = mr_split4s<false>(xa, s0, a.fzero), XD = mr_split4s<false>(xd, s0, a.fzero);
            mr_mfma_a<true>(yr[tb][0], yi[tb][0], XA, Fe);
            mr_mfma_a<true>(yr[tb][1], yi[tb][1], XD, Fo);
        }
    }
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int tk = 0; tk < 2; ++tk)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float2 v = cmul(float2{yr[tb][tk][e], yi[tb][tk][e]}, twl[(8 * tk + 2 * e + tb) * 64]);
                yr[tb][tk][e] = v.x;
                yi[tb][tk][e] = v.y;
            }
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = float2{yr[0][tk][e], yi[0][tk][e]};
            v1[e] = float2{yr[1][tk][e], yi[1][tk][e]};
        }
        mr_mfma_b<true>(er[tk], ei[tk], F16, mr_split4u(v0, a.fone));
        mr_mfma_b<true>(orr[tk], oi[tk], mr_table(ft, 2, l), mr_split4u(v1, a.fone));
    }
}

// one output of residue RS from its held stage-2 results: multiply, inverse
// stages 1 and 2; residue 0 writes its rows to the pencil's column of cur,
// residue 1 adds W2048^-n times its rows to them.  FIRST: slices 5 + RS
// (and 7 after residue 1) of the tile's store and fetch; LASTG: the last use
// of gr, after which the next tile's multiplier inputs are loaded into it.
template <int OPK, int RS, bool FIRST, bool LASTG>
__device__ __forceinline__ void mr2p_inverse(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                             unsigned char* __restrict__ nxt, const float2* __restrict__ twl,
                                             const float2* __restrict__ tsl, const mf_h8* __restrict__ ft,
                                             const MrTile& T, const Mr2SliceOffs (&so)[2], char* obase, int ncv,
                                             bool hasprev, const char* ibase, bool more, const MrTile& Tn,
                                             bool load_g, float (&gr)[2][4][2], const mf_f4 (&er)[2],
                                             const mf_f4 (&ei)[2], const mf_f4 (&orr)[2], const mf_f4 (&oi)[2],
                                             float inv0, float c0, float kst, const unsigned (&sb)[2], int w, int l,
                                             const MfT16& F16)
{
    constexpr bool mimag = mf_imag_multiplier<OPK>();
    const int g = l >> 4, j = l & 15;
    const int kb = 4 * j + 256 * g + RS;
    const float kbf = (float)kb;
    mf_f4 ur[2][2], ui[2][2];
    float s2[2], inv2[2];
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        float2 za[4], zb[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m0 = mr2_fan_kmul<OPK>(c0, kst, gr, tk, 0, e, kb, kbf);
            const float m1 = mr2_fan_kmul<OPK>(c0, kst, gr, tk, 1, e, kb, kbf);
            const float mp = m0 + m1, mm = m0 - m1;
            const float ar = er[tk][e] * mp + orr[tk][e] * mm, ai = ei[tk][e] * mp + oi[tk][e] * mm;
            const float br = er[tk][e] * mm + orr[tk][e] * mp, bi = ei[tk][e] * mm + oi[tk][e] * mp;
            if constexpr (mimag) {
                za[e] = float2{-ai, ar};
                zb[e] = float2{-bi, br};
            } else {
                za[e] = float2{ar, ai};
                zb[e] = float2{br, bi};
            }
        }
        if (LASTG && tk == 1 && load_g) mr2_kload<OPK>(gr, a, Tn, Tn.col0 + w, kb, mr_after(kb, za, zb));
        mr_scale(mr_wave_max(mr_absmax8(za, zb)), s2[tk], inv2[tk]);
        {
            const MfFrag XA = mr_split4s<true>(za, s2[tk], a.fzero), XB = mr_split4s<true>(zb, s2[tk], a.fzero);
            mr_mfma_a<true>(ur[tk][0], ui[tk][0], XA, F16);
            mr_mfma_a<true>(ur[tk][1], ui[tk][1], XB, mr_table(ft, 1, l));
        }
        if (FIRST && tk == 0) mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 5 + RS);
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float2 v = cmul(float2{ur[tk][tn][e], ui[tk][tn][e]}, twl[(8 * tn + 2 * e + tk) * 64]);
                ur[tk][tn][e] = v.x;
                ui[tk][tn][e] = v.y;
            }
    }
    const float smin = __builtin_fminf(s2[0], s2[1]);
    const float sa = 0.5f * smin * inv2[0], sb2 = 0.5f * smin * inv2[1];
    const float fo = 2.0f * inv0 * __builtin_fmaxf(inv2[0], inv2[1]);
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = float2{ur[0][tn][e], ui[0][tn][e]};
            v1[e] = float2{ur[1][tn][e], ui[1][tn][e]};
        }
        const MfFrag B0 = mr_split4s<false>(v0, sa, a.fzero), B1 = mr_split4s<false>(v1, sb2, a.fzero);
        mf_f4 dr, di, qr, qi;
        mr_mfma_b<true>(dr, di, F16, B0);
        mr_mfma_b<true>(qr, qi, mr_table(ft, 2, l), B1);
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                // row z = 2 j + tn + 32 (4 g + e + 16 h): E + O for h = 0, E - O for h = 1
                const float yr = h ? (dr[e] - qr[e]) * fo : (dr[e] + qr[e]) * fo;
                const float yi = h ? -(di[e] - qi[e]) * fo : -(di[e] + qi[e]) * fo;
                float2* q = reinterpret_cast<float2*>(cur + sb[tn] + 2048 * e + 32768 * h);
                if (RS == 0) {
                    *q = float2{yr, yi};
                } else {
                    const float2 t = tsl[(8 * tn + 4 * h + e) * 64];
                    const float2 y0 = *q;
                    *q = float2{y0.x + yr * t.x + yi * t.y, y0.y + yi * t.x - yr * t.y};
                }
            }
    }
}

// after a non-last output: every column of cur holds it; wavefront w stores
// its blocks w + 8 k to that output (ob: its tile's column 0)
__device__ __forceinline__ void mr2p_fan_store(unsigned char* __restrict__ cur, const PencilArgs<float>& a,
                                               const Mr2SliceOffs (&so)[2], char* ob, int ncvt, int w)
{
    mr2_lds_barrier();
#pragma unroll
    for (int k = 0; k < 8; ++k) mr2p_slice(cur, a, so, ob, ncvt, true, nullptr, false, w, k);
}

template <int OPK, bool WBAR, bool FIRST, bool LASTG>
__device__ __forceinline__ void mr2p_fan_out(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                             unsigned char* __restrict__ nxt, const float2* __restrict__ twl,
                                             const float2* __restrict__ tsl, const mf_h8* __restrict__ ft,
                                             const MrTile& T, const Mr2SliceOffs (&so)[2], char* obase, int ncv,
                                             bool hasprev, const char* ibase, bool more, const MrTile& Tn,
                                             bool load_g, float (&gm)[2][2][4][2], const mf_f4 (&er)[2][2],
                                             const mf_f4 (&ei)[2][2], const mf_f4 (&orr)[2][2],
                                             const mf_f4 (&oi)[2][2], float inv0, const unsigned (&sb)[2], int w,
                                             int l, const MfT16& F16)
{
    float c0 = 0.0f;
    if constexpr (OPK == OP_LAPL) c0 = mf_kcoef<OPK>(a, tile_k(a, T.col0 + w, T.slab, T.col0 + w < a.ncol), 0).c0;
    const float kst = (float)(2.0 * kPi * a.dk[2]);
    // the previous output's rows have been read from cur by every wavefront
    if constexpr (WBAR) mr2_lds_barrier();
    mr2p_inverse<OPK, 0, FIRST, LASTG>(a, cur, nxt, twl, tsl, ft, T, so, obase, ncv, hasprev, ibase, more, Tn,
                                       load_g, gm[0], er[0], ei[0], orr[0], oi[0], inv0, c0, kst, sb, w, l, F16);
    mr2p_inverse<OPK, 1, FIRST, LASTG>(a, cur, nxt, twl, tsl, ft, T, so, obase, ncv, hasprev, ibase, more, Tn,
                                       load_g, gm[1], er[1], ei[1], orr[1], oi[1], inv0, c0, kst, sb, w, l, F16);
    if (FIRST) mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 7);
}

template <int MASK>
__device__ __forceinline__ void mr2p_fan_step(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                              unsigned char* __restrict__ nxt, const float2* __restrict__ twm,
                                              const float2* __restrict__ tws, const mf_h8* __restrict__ ft,
                                              const MrTile& T, const Mr2SliceOffs (&so)[2], char* obase, int ncv,
                                              bool hasprev, const char* ibase, bool more, const MrTile& Tn,
                                              bool load_g, float (&gm)[2][2][4][2], int w, int l, long ostride)
{
    constexpr int NO = mr2_popc(MASK);
    constexpr int OD = MASK & 1, OL = mr2_popc(MASK & 3);
    const int g = l >> 4, j = l & 15;
    const float2* twl = twm + l;
    const float2* tsl = tws + l;
    // staging offsets of rows z = 2 j + q + 32 (4 g + e + 16 h), column w
    unsigned sb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = 2 * (j & 7) + q;
        sb[q] = (unsigned)(((j >> 3) + 8 * g) * 1024) + 16u * mr2_slot(r, w >> 1, g & 1) + (w & 1) * 8;
    }
    float2 x[2][2][4];
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                x[tb][h][e] = *reinterpret_cast<const float2*>(cur + sb[tb] + 2048 * e + 32768 * h);
    mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 0);
    float s0, inv0;
    {
        const float m0 = mr_absmax8(x[0][0], x[0][1]), m1 = mr_absmax8(x[1][0], x[1][1]);
        mr_scale(2.0f * mr_wave_max(__builtin_fmaxf(m0, m1)), s0, inv0);
    }
    mf_f4 er[2][2], ei[2][2], orr[2][2], oi[2][2];
    // the 16-point DFT table (8 of each residue's 16 products) held for the tile
    const MfT16 F16 = mr_table(ft, 0, l);
    mr2p_forward<0>(a, x, s0, twl, tsl, ft, F16, l, er[0], ei[0], orr[0], oi[0]);
    mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 1);
    mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 2);
    mr2p_forward<1>(a, x, s0, twl, tsl, ft, F16, l, er[1], ei[1], orr[1], oi[1]);
    mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 3);
    mr2p_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 4);
    char* ob = const_cast<char*>(mr_tile_base(a.out, a.oa, T));
    const int ncvt = a.ncol - T.col0;
    if constexpr ((MASK & 1) != 0) {
        mr2p_fan_out<OP_PLAIN, false, true, NO == 1>(a, cur, nxt, twl, tsl, ft, T, so, obase, ncv, hasprev, ibase,
                                                     more, Tn, load_g, gm, er, ei, orr, oi, inv0, sb, w, l, F16);
        if constexpr (NO > 1) mr2p_fan_store(cur, a, so, ob, ncvt, w);
    }
    if constexpr ((MASK & 2) != 0) {
        mr2p_fan_out<OP_GRAD, (OD > 0), (OD == 0), (OD == NO - 1)>(a, cur, nxt, twl, tsl, ft, T, so, obase, ncv,
                                                                   hasprev, ibase, more, Tn, load_g, gm, er, ei,
                                                                   orr, oi, inv0, sb, w, l, F16);
        if constexpr (OD < NO - 1) mr2p_fan_store(cur, a, so, ob + OD * ostride, ncvt, w);
    }
    if constexpr ((MASK & 4) != 0)
        mr2p_fan_out<OP_LAPL, (OL > 0), (OL == 0), true>(a, cur, nxt, twl, tsl, ft, T, so, obase, ncv, hasprev,
                                                         ibase, more, Tn, load_g, gm, er, ei, orr, oi, inv0, sb, w,
                                                         l, F16);
}

template <int MASK>
__global__ void __launch_bounds__(512, 2) k_zmfma2048p(PencilArgs<float> a)
{
    using M = MfmaZR2;
    constexpr int NW = MfmaZR2P::NW;
    constexpr int NO = mr2_popc(MASK);
    constexpr int OPG = (MASK & 1) ? OP_PLAIN : (MASK & 2) ? OP_GRAD : OP_LAPL;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg0 = smem_raw;
    float2* twm = reinterpret_cast<float2*>(smem_raw + 2 * M::STG);
    float2* tws = reinterpret_cast<float2*>(smem_raw + 2 * M::STG + M::TWB);
    mf_h8* ftw = reinterpret_cast<mf_h8*>(smem_raw + 2 * M::STG + M::TWB + M::TSB);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l0 = tid & 63;
    const int nunit = mr_nunit(a);
    const int G = gridDim.x;
    const long ostride = a.oa.comp * 8;
    const long olast = (long)(NO - 1) * ostride;

    for (int i = tid; i < 1024; i += 64 * NW) {
        const int slot = i >> 6, j = i & 15, g = (i >> 4) & 3;
        const int r = 2 * j + (slot >> 3), s = 8 * g + (slot & 7);
        twm[i] = a.plan.tw[(2 * r * s) & (M::L - 1)];
        const int n = 2 * j + (slot >> 3) + 32 * (4 * g + (slot & 3) + 16 * ((slot >> 2) & 1));
        tws[i] = a.plan.tw[n];
    }
    for (int i = tid; i < 3 * 4 * 64; i += 64 * NW) {
        const int ln = i & 63, f = (i >> 6) & 3, tb = i >> 8;
        const int fo = (f == 0) ? 0 : (f == 1) ? 2 : (f == 2) ? 1 : 3;
        ftw[(tb * 4 + fo) * 64 + ln] = reinterpret_cast<const mf_h8*>(a.mfrag16)[(tb * 64 + ln) * 4 + f];
    }
    const Mr2SliceOffs so[2] = {mr2_slice_offs(a, w, l0), mr2_slice_offs(a, w + 8, l0)};
    int u = blockIdx.x, m = 0;
    MrUnit U{};
    MrTile T{};
    const bool any = u < nunit && (U = mr_unit<M::P>(a, u, nunit), mr_tile(a, U, 0, T));
    float gm[2][2][4][2];
    if (any) {
        const char* ib = mr_tile_base(a.in, a.ia, T);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int n = w + 8 * k;
            __builtin_amdgcn_global_load_lds(ib + mr2_zoff_u(a.ia, 16 * n) + so[k & 1].ilane,
                                             (mr_lds_ptr)(stg0 + n * 1024), 16, 0, 0);
        }
        const int kb0 = 4 * (l0 & 15) + 256 * (l0 >> 4);
        mr2_kload<OPG>(gm[0], a, T, T.col0 + w, kb0, kb0);
        mr2_kload<OPG>(gm[1], a, T, T.col0 + w, kb0 + 1, kb0 + 1);
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
        const int ncv = a.ncol - Tp.col0;
        mr2p_fan_step<MASK>(a, cur, nxt, twm, tws, ftw, T, so, obase, ncv, it > 0, ibase, more, Tn, load_g, gm, w, l,
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
        const int ncv = a.ncol - Tp.col0;
#pragma unroll
        for (int k = 0; k < 8; ++k) mr2p_slice(last, a, so, obase, ncv, true, nullptr, false, w, k);
    }
}

}  // namespace psamd
