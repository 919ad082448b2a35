// Mixed-precision z pass on the matrix cores, transform length L = 1024,
// three-component operators (curl, div), register-pencil form.
//
// The curl of G * psi and the divergence of G * psi mix the three input
// components at each wavenumber (the reference's solve3d.cpp:355-408 at
// Mz = 1024):
//   curl  u_x = i G (ky psi_z - kz psi_y), u_y = i G (kz psi_x - kx psi_z),
//         u_z = i G (kx psi_y - ky psi_x)
//   div   d   = i G (kx psi_x + ky psi_y + kz psi_z)
// (Hermitian wavenumbers: the Nyquist index contributes no derivative).
//
// One wavefront owns one pencil column of an 8-column tile and all three of
// its components: workgroups of 8 wavefronts (2 per SIMD, up to 256 VGPRs).
// A slab's three component tiles (8 columns x 512 z rows, 32 KB each) pass
// through the double-buffered staging in turn: each is transformed forward
// by the stages of the one-component kernel of 1024 and its stage-2 results
// are held in registers (32 VGPRs per component).  After the third, each
// output forms its k-space combination from the held spectra (each
// component's input scale folded into its factor), runs the two inverse
// stages and is written to the pencil's column of the tile buffer; an output
// other than the last is stored from there by the wavefronts' slices before
// the next one is written, the last is stored during the next slab's first
// component tile.  Multiplier inputs (Gk) are loaded once per slab pair.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_mfma_reg.hpp"
#include "ps_mfma_reg2048.hpp"

namespace psamd {

struct MfmaZR3 {
    static constexpr int L = 1024, NZ = 512, P = 8, NW = 8;
    static constexpr int STG = NZ * P * 8;            // 32 KB per staging buffer
    static constexpr int TWB = 16 * 64 * 8;           // W1024 twiddles [slot][lane]
    static constexpr int FTB = 3 * 4 * 64 * 16;       // [E16, P, Q][rh, ih, rl, il][lane]
    static constexpr size_t lds_bytes() { return 2 * (size_t)STG + TWB + FTB; }
};

// slice k < 4 of wavefront w: block n = w + 8 k (16 rows x 8 columns), whose
// rotation is k & 1, through the 2048 kernel's slice with the offsets so[k & 1]
__device__ __forceinline__ void mr3_slice(unsigned char* __restrict__ buf, const PencilArgs<float>& a,
                                          const Mr2SliceOffs (&so)[2], char* obase, int ncv, bool hasprev,
                                          const char* ibase, bool more, int w, int k)
{
    mr2_slice(buf, a, so[k & 1], obase, ncv, hasprev, ibase, more, w + 8 * (k & 1), k >> 1);
}

// staging offsets of the pencil rows z = 2 j + q + 32 (4 g + e) (q = 0, 1) of
// column w: block (j >> 3) + 8 g + 2 e (rotation g & 1), row 2 (j & 7) + q;
// + 2048 e bytes for e
__device__ __forceinline__ void mr3_rows(unsigned (&sb)[2], int w, int l)
{
    const int g = l >> 4, j = l & 15;
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = 2 * (j & 7) + q;
        sb[q] = (unsigned)(((j >> 3) + 8 * g) * 1024) + 16u * mr2_slot(r, w >> 1, g & 1) + (w & 1) * 8;
    }
}

// One component tile: the pencil's rows from cur, stage 1, the W1024 twiddle,
// stage 2; returns E and O of both k1 parities (kz = 2 j + tk + 32 k2, X = E
// + O at k2 = 4 g + e, E - O at k2 + 16) and the input scale's inverse.  Its
// four slices (the previous output's blocks of nxt stored, the next tile's
// fetched into them) are spread over it.
template <bool H16>
__device__ __forceinline__ void mr3_forward(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                            unsigned char* __restrict__ nxt, const float2* __restrict__ twm,
                                            const mf_h8* __restrict__ ft, const Mr2SliceOffs (&so)[2],
                                            char* obase, int ncv, bool hasprev, const char* ibase, bool more, int w,
                                            int l, const MfT16& F16, mf_f4 (&er)[2], mf_f4 (&ei)[2], mf_f4 (&orr)[2],
                                            mf_f4 (&oi)[2], float& inv0)
{
    mr3_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 0);
    unsigned sb[2];
    mr3_rows(sb, w, l);
    float2 x[2][4];
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int e = 0; e < 4; ++e) x[tb][e] = *reinterpret_cast<const float2*>(cur + sb[tb] + 2048 * e);
    float s0;
    mr_scale(mr_wave_max(mr_absmax8(x[0], x[1])), s0, inv0);
    mr3_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 1);
    mf_f4 yr[2][2], yi[2][2];
    {
        const MfT16 Fe = H16 ? F16 : mr_table(ft, 0, l), Fo = mr_table(ft, 1, l);
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
            const MfFrag X = mr_split4s<false>(x[tb], s0, a.fzero);
            mr_mfma_a<true>(yr[tb][0], yi[tb][0], X, Fe);
            mr_mfma_a<true>(yr[tb][1], yi[tb][1], X, Fo);
        }
    }
    mr3_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 2);
    const float2* twl = twm + l;
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
        mr_mfma_b<true>(er[tk], ei[tk], H16 ? F16 : mr_table(ft, 0, l), mr_split4u(v0, a.fone));
        mr_mfma_b<true>(orr[tk], oi[tk], mr_table(ft, 2, l), mr_split4u(v1, a.fone));
    }
    mr3_slice(nxt, a, so, obase, ncv, hasprev, ibase, more, w, 3);
}

// the held spectra of the three components
struct Mr3Spec {
    mf_f4 er[3][2], ei[3][2], orr[3][2], oi[3][2];
    float inv0[3];
};

// Output O of operator OPK: the k-space combination of the held spectra at
// the lane's 16 points, times i G, the inverse stages, the rows written to
// the pencil's column of cur (after a barrier when WBAR: the previous output
// has been read from cur by every wavefront).  LASTG: the last use of gm,
// after which the next slab pair's multiplier inputs are loaded into it.
template <int OPK, int O, bool WBAR, bool LASTG, bool H16>
__device__ __forceinline__ void mr3_output(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                           const float2* __restrict__ twm, const mf_h8* __restrict__ ft,
                                           const MrTile& T, const MrTile& Tn, bool load_g, mr_f2a (&gm)[2][4],
                                           Mr3Spec& S, int w, int l, const MfT16& F16)
{
    const int g = l >> 4, j = l & 15, kb = 2 * j + 128 * g;
    const float kbf = (float)kb;
    // each output forms its factors and sums from the held values anew:
    // values shared between outputs would stay live across them
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int e = 0; e < 4; ++e) asm volatile("" : "+v"(gm[h][e]));
#pragma unroll
    for (int c = 0; c < 3; ++c)
#pragma unroll
        for (int tk = 0; tk < 2; ++tk)
            asm volatile("" : "+v"(S.er[c][tk]), "+v"(S.ei[c][tk]), "+v"(S.orr[c][tk]), "+v"(S.oi[c][tk]));
    const TileK<float> tk0 = tile_k(a, T.col0 + w, T.slab, T.col0 + w < a.ncol);
    const float kst = (float)(2.0 * kPi * a.dk[2]);
    const float2* twl = twm + l;
    // factors of the components: kx psi_x + ky psi_y + kz psi_z (div), or
    // the pair of component CA times its k minus component CB times its k (curl)
    constexpr int CA = (O == 0) ? 2 : (O == 1) ? 0 : 1;   // u_x: ky psi_z - kz psi_y, u_y: kz psi_x - kx psi_z,
    constexpr int CB = (O == 0) ? 1 : (O == 1) ? 2 : 0;   // u_z: kx psi_y - ky psi_x
    mf_f4 ur[2][2], ui[2][2];
    float s2[2], inv2[2];
#pragma unroll
    for (int tk = 0; tk < 2; ++tk) {
        float2 za[4], zb[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // points k2 = 4 g + e (h = 0) and k2 + 16 (h = 1) of column k1 = 2 j + tk
            float2 z0{0.0f, 0.0f}, z1{0.0f, 0.0f};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float gv = mr_kmul<OP_PLAIN>(MfKCoef{}, gm, tk, h, e, kb, kbf);
                const float kzw = (kbf + (float)(tk + 32 * e - 512 * h)) * kst;
                const float kz = (h == 1 && tk == 0 && e == 0 && kb == 0) ? 0.0f : kzw;
                const float kc[3] = {tk0.kxh, tk0.kyh, kz};
                float2 t{0.0f, 0.0f};
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    float f;
                    if constexpr (OPK == OP_DIV) {
                        f = kc[c];
                    } else {
                        // the factor of component c: k[(c + 1) % 3 ...] per the curl rows above
                        f = (c == CA) ? kc[(O + 1) % 3] : (c == CB) ? -kc[(O + 2) % 3] : 0.0f;
                        if (c != CA && c != CB) continue;
                    }
                    f *= gv * S.inv0[c];
                    const float xr = h ? S.er[c][tk][e] - S.orr[c][tk][e] : S.er[c][tk][e] + S.orr[c][tk][e];
                    const float xi = h ? S.ei[c][tk][e] - S.oi[c][tk][e] : S.ei[c][tk][e] + S.oi[c][tk][e];
                    t.x += f * xr;
                    t.y += f * xi;
                }
                if (h) z1 = t;
                else z0 = t;
            }
            // i (z0 +- z1): the inverse stage-1 butterfly inputs times i
            za[e] = float2{-(z0.y + z1.y), z0.x + z1.x};
            zb[e] = float2{-(z0.y - z1.y), z0.x - z1.x};
        }
        if (LASTG && tk == 1 && load_g) mr_kload<OP_PLAIN>(gm, a, Tn, Tn.col0 + w, j, g, mr_after(kb, za, zb));
        mr_scale(mr_wave_max(mr_absmax8(za, zb)), s2[tk], inv2[tk]);
        {
            const MfFrag XA = mr_split4s<true>(za, s2[tk], a.fzero), XB = mr_split4s<true>(zb, s2[tk], a.fzero);
            mr_mfma_a<true>(ur[tk][0], ui[tk][0], XA, H16 ? F16 : mr_table(ft, 0, l));
            mr_mfma_a<true>(ur[tk][1], ui[tk][1], XB, mr_table(ft, 1, l));
        }
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float2 v = cmul(float2{ur[tk][tn][e], ui[tk][tn][e]}, twl[(8 * tn + 2 * e + tk) * 64]);
                ur[tk][tn][e] = v.x;
                ui[tk][tn][e] = v.y;
            }
    }
    // inverse stage 2 input scales from the bound, as in the one-component
    // kernel; the input scales are in the combination, so fo has no inv0
    const float smin = __builtin_fminf(s2[0], s2[1]);
    const float sa = 0.5f * smin * inv2[0], sb2 = 0.5f * smin * inv2[1];
    const float fo = 2.0f * __builtin_fmaxf(inv2[0], inv2[1]);
    unsigned sb[2];
    mr3_rows(sb, w, l);
    if constexpr (WBAR) mr2_lds_barrier();
#pragma unroll
    for (int tn = 0; tn < 2; ++tn) {
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = float2{ur[0][tn][e], ui[0][tn][e]};
            v1[e] = float2{ur[1][tn][e], ui[1][tn][e]};
        }
        const MfFrag B0 = mr_split4s<false>(v0, sa, a.fzero), B1 = mr_split4s<false>(v1, sb2, a.fzero);
        mf_f4 dr, di;
        mr_mfma_b<true>(dr, di, H16 ? F16 : mr_table(ft, 0, l), B0);
        mr_mfma_b<false>(dr, di, mr_table(ft, 2, l), B1);
        // rows z = 32 (4 g + e) + 2 j + tn
#pragma unroll
        for (int e = 0; e < 4; ++e)
            *reinterpret_cast<float2*>(cur + sb[tn] + 2048 * e) = float2{dr[e] * fo, -di[e] * fo};
    }
}

// after a non-last output: every column of cur holds it; wavefront w stores
// its blocks w + 8 k (k < 4) to it (ob: the tile's column 0 in that output)
__device__ __forceinline__ void mr3_store(unsigned char* __restrict__ cur, const PencilArgs<float>& a,
                                          const Mr2SliceOffs (&so)[2], char* ob, int ncvt, int w)
{
    mr2_lds_barrier();
#pragma unroll
    for (int k = 0; k < 4; ++k) mr3_slice(cur, a, so, ob, ncvt, true, nullptr, false, w, k);
}

// Persistent kernel over units (slab pair, 8-column tile) of the one
// three-component input; iterations are the component tiles (slab, c) of
// the unit's slabs in turn.
template <int OPK>
__global__ void __launch_bounds__(512, 2) k_zmfma1024c(PencilArgs<float> a)
{
    using M = MfmaZR3;
    constexpr int L = M::L;
    constexpr int NOUT = (OPK == OP_DIV) ? 1 : 3;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg0 = smem_raw;
    float2* twm = reinterpret_cast<float2*>(smem_raw + 2 * M::STG);
    mf_h8* ftw = reinterpret_cast<mf_h8*>(smem_raw + 2 * M::STG + M::TWB);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l0 = tid & 63;
    const int nunit = mr_nunit(a);
    const int G = gridDim.x;
    const long ostride = a.oa.comp * 8;

    for (int i = tid; i < 1024; i += 64 * M::NW) {
        const int slot = i >> 6, j = i & 15, g = (i >> 4) & 3;
        const int r = 2 * j + (slot >> 3), s = 8 * g + (slot & 7);
        twm[i] = a.plan.tw[(r * s) & (L - 1)];
    }
    for (int i = tid; i < 3 * 4 * 64; i += 64 * M::NW) {
        const int ln = i & 63, f = (i >> 6) & 3, tb = i >> 8;
        const int fo = (f == 0) ? 0 : (f == 1) ? 2 : (f == 2) ? 1 : 3;
        ftw[(tb * 4 + fo) * 64 + ln] = reinterpret_cast<const mf_h8*>(a.mfrag16)[(tb * 64 + ln) * 4 + f];
    }
    const Mr2SliceOffs so[2] = {mr2_slice_offs(a, w, l0), mr2_slice_offs(a, w + 8, l0)};
    int u = blockIdx.x, m = 0, c = 0;
    MrUnit U{};
    MrTile T{};
    const bool any = u < nunit && (U = mr_unit<M::P>(a, u, nunit), mr_tile(a, U, 0, T));
    mr_f2a gm[2][4];
    if (any) {
        T.comp = 0;
        const char* ib = mr_tile_base(a.in, a.ia, T);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int n = w + 8 * k;
            __builtin_amdgcn_global_load_lds(ib + mr2_zoff_u(a.ia, 16 * n) + so[k & 1].ilane,
                                             (mr_lds_ptr)(stg0 + n * 1024), 16, 0, 0);
        }
        mr_kload<OP_PLAIN>(gm, a, T, T.col0 + w, l0 & 15, l0 >> 4, 2 * (l0 & 15) + 128 * (l0 >> 4));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    Mr3Spec S;
    MrTile Tp = T;    // the slab tile whose last output the next slices store
    bool hasout = false;
    int it = 0;
    for (; any; ++it) {
        int l = l0;
        asm volatile("" : "+v"(l));
        unsigned char* cur = stg0 + (it & 1) * M::STG;
        unsigned char* nxt = stg0 + ((it & 1) ^ 1) * M::STG;
        // the next iteration: the slab's next component, else the pair's
        // second slab, else the next unit's first
        MrTile Tn = T;
        MrUnit Un = U;
        int un = u, mn = m, cn = c + 1;
        bool more = true;
        if (cn == 3) {
            cn = 0;
            mn = 1;
            if (m == 1 || !mr_tile(a, U, 1, Tn)) {
                un = u + G;
                mn = 0;
                more = un < nunit && (Un = mr_unit<M::P>(a, un, nunit), mr_tile(a, Un, 0, Tn));
            }
        }
        Tn.comp = cn;
        // a new slab pair's multipliers, loaded after this slab's last use (c = 2)
        const bool load_g = more && c == 2 && !(mn == 1 && Tn.kyo == T.kyo);
        MrTile To = Tp;
        To.comp = NOUT - 1;
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, To));
        const char* ibase = mr_tile_base(a.in, a.ia, Tn);
        const int ncv = a.ncol - Tp.col0;
        const bool hasprev = hasout && c == 0;
        // the 16-point DFT table held for the iteration (div: the registers allow)
        constexpr bool H16 = OPK == OP_DIV;
        MfT16 F16{};
        if constexpr (H16) F16 = mr_table(ftw, 0, l);
        if (c == 0)
            mr3_forward<H16>(a, cur, nxt, twm, ftw, so, obase, ncv, hasprev, ibase, more, w, l, F16, S.er[0],
                             S.ei[0], S.orr[0], S.oi[0], S.inv0[0]);
        else if (c == 1)
            mr3_forward<H16>(a, cur, nxt, twm, ftw, so, obase, ncv, false, ibase, more, w, l, F16, S.er[1], S.ei[1],
                             S.orr[1], S.oi[1], S.inv0[1]);
        else {
            mr3_forward<H16>(a, cur, nxt, twm, ftw, so, obase, ncv, false, ibase, more, w, l, F16, S.er[2], S.ei[2],
                             S.orr[2], S.oi[2], S.inv0[2]);
            MrTile T0 = T;
            T0.comp = 0;
            char* ob = const_cast<char*>(mr_tile_base(a.out, a.oa, T0));
            const int ncvt = a.ncol - T.col0;
            if constexpr (OPK == OP_DIV) {
                mr3_output<OPK, 0, false, true, H16>(a, cur, twm, ftw, T, Tn, load_g, gm, S, w, l, F16);
            } else {
                mr3_output<OPK, 0, false, false, H16>(a, cur, twm, ftw, T, Tn, load_g, gm, S, w, l, F16);
                mr3_store(cur, a, so, ob, ncvt, w);
                mr3_output<OPK, 1, true, false, H16>(a, cur, twm, ftw, T, Tn, load_g, gm, S, w, l, F16);
                mr3_store(cur, a, so, ob + ostride, ncvt, w);
                mr3_output<OPK, 2, true, true, H16>(a, cur, twm, ftw, T, Tn, load_g, gm, S, w, l, F16);
            }
            Tp = T;
            hasout = true;
        }
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        T = Tn;
        U = Un;
        u = un;
        m = mn;
        c = cn;
        if (!more) break;
    }
    // the last slab's last output (in the buffer of the last iteration)
    if (any) {
        unsigned char* last = stg0 + (it & 1) * M::STG;
        MrTile To = Tp;
        To.comp = NOUT - 1;
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, To));
        const int ncv = a.ncol - Tp.col0;
#pragma unroll
        for (int k = 0; k < 4; ++k) mr3_slice(last, a, so, obase, ncv, true, nullptr, false, w, k);
    }
}

}  // namespace psamd
