// Mixed-precision z pass on the matrix cores (PS_MIXED), transform length
// L = 1024 = 32 x 32 (Mz of a 512-cell unbounded z direction).
//
// The fused z pass (pad nz -> Mz, FFT, k-space operator, IFFT, keep nz; the
// reference's solve3d.cpp:243-441) is computed as a four-step transform whose
// two 32-point DFT stages run as v_mfma_f32_16x16x32_f16 products:
//
//   X[b' + 32 q'] = sum_b W32^{b q'} W1024^{b b'} sum_{m<16} W32^{m b'} x[b + 32 m]
//   x[32 n1 + n2] = sum_b' W32^{-n1 b'} W1024^{-n2 b'} sum_q' W32^{-n2 q'} Z[b' + 32 q']
//
// Each DFT stage multiplies the 32 x 32 DFT matrix (in real 2x2-block form,
// split once into fp16 hi + lo) by 16 columns of fp16 hi + lo data:
//   D = A_hi B_hi + A_hi B_lo + A_lo B_hi   (fp32 accumulation)
// which keeps about 22 significant bits per product (the dropped A_lo B_lo
// term is 2^-22 relative).  fp16 has a narrow exponent range, so every stage
// input is scaled by a power of two that maps its largest magnitude below
// 2^10; a 32-point DFT then grows values by at most 32 sqrt(2) < 2^5.5, which
// keeps the next stage's fp16 inputs below 65504.  Scales are per DFT column
// (stages whose input lives in one column's four lanes) or per pencil (stages
// whose input comes from LDS after a workgroup barrier).  The twiddles
// W1024^{b b'} are applied in fp32 on the vector ALUs.
//
// Workgroup: P = 8 pencils (adjacent kx columns of one ky slab) x the whole z
// extent, 8 wavefronts.  An MFMA column is (pencil p = j & 7, butterfly
// b = 2 blk + (j >> 3)); wavefront w owns blocks blk = w and w + 8 in every
// stage.  Lane l = 16 q + j holds, in stage input and output alike, the
// points 4q..4q+3 and 16+4q..16+4q+3 of its column, so the forward stage-2
// output feeds the inverse stage-1 input in registers, with the k-space
// operator in between.  The two data exchanges go through one LDS array
// [p][row][col]; the inverse stage-1 output is written into the row its
// forward input was read from (that row belongs to the same column), so only
// two workgroup barriers separate the four stages.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_kernels_impl.hpp"

namespace psamd {

typedef _Float16 mf_h8 __attribute__((ext_vector_type(8)));
typedef float mf_f4 __attribute__((ext_vector_type(4)));

struct MfmaZ {
    static constexpr int L = 1024;
    static constexpr int R = 32;
    static constexpr int P = 8;          // pencils per workgroup
    static constexpr int NW = 8;         // wavefronts per workgroup
    // LDS strides (complex): within each 16-lane group one access (64-bit)
    // touches 16 distinct bank pairs in both exchange patterns, row-strided
    // (p PS + bsub) and row-contiguous (p PS + bsub RS), as PS = 2 and
    // RS = 1 (mod 16)
    static constexpr int RS = 33;
    static constexpr int PS = 32 * 33 + 2;
    static constexpr size_t lds_bytes()
    {
        return (size_t)P * PS * 8 + (size_t)L * 8 + 2 * P * 4;
    }
};

// cos(2 pi j / 32), j = 0..31
struct Cos32 {
    static constexpr double c[32] = {
        1.0, 0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
        0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173, 0.19509032201612826785,
        0.0, -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
        -0.70710678118654752440, -0.83146961230254523708, -0.92387953251128675613, -0.98078528040323044913,
        -1.0, -0.98078528040323044913, -0.92387953251128675613, -0.83146961230254523708,
        -0.70710678118654752440, -0.55557023301960222474, -0.38268343236508977173, -0.19509032201612826785,
        0.0, 0.19509032201612826785, 0.38268343236508977173, 0.55557023301960222474,
        0.70710678118654752440, 0.83146961230254523708, 0.92387953251128675613, 0.98078528040323044913};
};

// ---------------------------------------------------------------------------
// fp16 hi/lo split of scaled fp32 values:
//   hi = f16(x s),  lo = f16(x s - hi)   (x s - hi is exact before rounding)
// Written as ordinary arithmetic (the compiler forms v_cvt_pk_f16_f32 and
// v_fma_mix_f32 from it) rather than inline assembly: the compiler tracks
// the wait states between matrix-core instructions and the vector
// instructions that read their results or overwrite their accumulator
// inputs only for instructions it generates itself.
// ---------------------------------------------------------------------------
typedef _Float16 mf_h2 __attribute__((ext_vector_type(2)));

// z: a run-time zero (see PencilArgs::fzero)
__device__ __forceinline__ void mf_split_pair(float x, float y, float sx, float sy, float z, unsigned& hi,
                                              unsigned& lo)
{
    const mf_h2 h = mf_h2{(_Float16)__builtin_fmaf(x, sx, z), (_Float16)__builtin_fmaf(y, sy, z)};
    const mf_h2 r = mf_h2{(_Float16)__builtin_fmaf(x, sx, -(float)h.x), (_Float16)__builtin_fmaf(y, sy, -(float)h.y)};
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, r);
}

// one: a run-time 1 (see PencilArgs::fone)
__device__ __forceinline__ void mf_split_pair1(float x, float y, float one, unsigned& hi, unsigned& lo)
{
    const mf_h2 h = mf_h2{(_Float16)x, (_Float16)y};
    const mf_h2 r = mf_h2{(_Float16)__builtin_fmaf(x, one, -(float)h.x), (_Float16)__builtin_fmaf(y, one, -(float)h.y)};
    hi = __builtin_bit_cast(unsigned, h);
    lo = __builtin_bit_cast(unsigned, r);
}

struct MfFrag {
    mf_h8 hi, lo;
};

// B fragment of four complex points (re, im interleaved) scaled by s; the
// imaginary parts are scaled by sim (= -s for the conjugate)
__device__ __forceinline__ MfFrag mf_split4(const float2 (&v)[4], float s, float sim, float z)
{
    unsigned H[4], Lo[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) mf_split_pair(v[e].x, v[e].y, s, sim, z, H[e], Lo[e]);
    MfFrag f;
    f.hi = __builtin_bit_cast(mf_h8, H);
    f.lo = __builtin_bit_cast(mf_h8, Lo);
    return f;
}

// B fragment of four complex points already inside the fp16 range
__device__ __forceinline__ MfFrag mf_split4u(const float2 (&v)[4], float one)
{
    unsigned H[4], Lo[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) mf_split_pair1(v[e].x, v[e].y, one, H[e], Lo[e]);
    MfFrag f;
    f.hi = __builtin_bit_cast(mf_h8, H);
    f.lo = __builtin_bit_cast(mf_h8, Lo);
    return f;
}

// negate points 1 and 3 of a fragment: rows k + 16 of the 32-point DFT are
// rows k with the input sign (-1)^m, and m = 4q + e + 16c has the parity of e
__device__ __forceinline__ mf_h8 mf_neg_odd(mf_h8 x)
{
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    u4 u = __builtin_bit_cast(u4, x);
    u.y ^= 0x80008000u;
    u.w ^= 0x80008000u;
    return __builtin_bit_cast(mf_h8, u);
}

// A fragments of rows 0..15 of the 32-point forward DFT in real form, for K
// chunk c (points m = 16 c + 0..15, re/im interleaved):
//   Yr row k: ( cos t,  sin t),  Yi row k: (-sin t, cos t),  t = 2 pi k m / 32
// Table layout (built by zmfma_fragments): [lane][c][rh, rl, ih, il][8 halves].
struct MfDft32 {
    mf_h8 rh[2], rl[2], ih[2], il[2];
};

__device__ __forceinline__ MfDft32 mf_dft32_frags(const void* table, int lane)
{
    const mf_h8* t = reinterpret_cast<const mf_h8*>(table) + lane * 8;
    MfDft32 F;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        F.rh[c] = t[4 * c + 0];
        F.rl[c] = t[4 * c + 1];
        F.ih[c] = t[4 * c + 2];
        F.il[c] = t[4 * c + 3];
    }
    return F;
}

// four independent accumulator chains (rows 0..15 and 16..31 of Yr and Yi)
// over NCH K chunks, issued round-robin so no product waits on the previous
struct MfAcc {
    mf_f4 r0, i0, r1, i1;
};

template <int NCH>
__device__ __forceinline__ MfAcc mf_dft32_rows32(const MfDft32& F, const MfFrag (&B)[2], const MfFrag (&Bn)[2])
{
    MfAcc d;
    d.r0 = d.i0 = d.r1 = d.i1 = mf_f4{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].hi, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].hi, d.i0, 0, 0, 0);
        d.r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], Bn[c].hi, d.r1, 0, 0, 0);
        d.i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], Bn[c].hi, d.i1, 0, 0, 0);
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].lo, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].lo, d.i0, 0, 0, 0);
        d.r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], Bn[c].lo, d.r1, 0, 0, 0);
        d.i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], Bn[c].lo, d.i1, 0, 0, 0);
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl[c], B[c].hi, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il[c], B[c].hi, d.i0, 0, 0, 0);
        d.r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl[c], Bn[c].hi, d.r1, 0, 0, 0);
        d.i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il[c], Bn[c].hi, d.i1, 0, 0, 0);
    }
    return d;
}

// same product with one fragment set: rows 0..15 first, then the odd points
// are negated in place for rows 16..31 (no second copy of the operands)
template <int NCH>
__device__ __forceinline__ MfAcc mf_dft32_rows32i(const MfDft32& F, MfFrag (&B)[2])
{
    MfAcc d;
    d.r0 = d.i0 = d.r1 = d.i1 = mf_f4{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].hi, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].hi, d.i0, 0, 0, 0);
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].lo, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].lo, d.i0, 0, 0, 0);
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl[c], B[c].hi, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il[c], B[c].hi, d.i0, 0, 0, 0);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        B[c].hi = mf_neg_odd(B[c].hi);
        B[c].lo = mf_neg_odd(B[c].lo);
    }
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        d.r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].hi, d.r1, 0, 0, 0);
        d.i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].hi, d.i1, 0, 0, 0);
        d.r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].lo, d.r1, 0, 0, 0);
        d.i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].lo, d.i1, 0, 0, 0);
        d.r1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl[c], B[c].hi, d.r1, 0, 0, 0);
        d.i1 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il[c], B[c].hi, d.i1, 0, 0, 0);
    }
    return d;
}

// rows 0..15 only (the truncated last inverse stage): two chains
__device__ __forceinline__ MfAcc mf_dft32_rows16(const MfDft32& F, const MfFrag (&B)[2])
{
    MfAcc d;
    d.r0 = d.i0 = d.r1 = d.i1 = mf_f4{0, 0, 0, 0};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].hi, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].hi, d.i0, 0, 0, 0);
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh[c], B[c].lo, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih[c], B[c].lo, d.i0, 0, 0, 0);
        d.r0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl[c], B[c].hi, d.r0, 0, 0, 0);
        d.i0 = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il[c], B[c].hi, d.i0, 0, 0, 0);
    }
    return d;
}

// byte offset pieces of position pos = base + 32 u along a pencil whose
// addressing is blocked by 2^sh (sh <= 5) or plain (sh >= 30): linear in u
__device__ __forceinline__ long mf_pos_stride32(int sh, long lo, long hi)
{
    return (sh >= 30) ? 32 * lo : (long)(32 >> sh) * hi;
}

// power-of-two scale mapping magnitudes below 2^e (m < 2^e) to below 2^10
__device__ __forceinline__ void mf_scale(float m, float& s, float& inv)
{
    int e = 0;
    (void)frexpf(m, &e);
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    s = ldexpf(1.0f, 10 - e);
    inv = ldexpf(1.0f, e - 10);
}

__device__ __forceinline__ float mf_absmax(float2 v, float m)
{
    return fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
}

// largest magnitude over the four lanes j, j+16, j+32, j+48 of a column:
// lane ^ 16 by a swizzle inside each half-wave, lane ^ 32 by a permute
// (bp32 = 4 (lane ^ 32), precomputed)
__device__ __forceinline__ float mf_colmax(float m, int bp32)
{
    m = fmaxf(m, __builtin_bit_cast(float, __builtin_amdgcn_ds_swizzle(__builtin_bit_cast(int, m), 0x401F)));
    m = fmaxf(m, __builtin_bit_cast(float, __builtin_amdgcn_ds_bpermute(bp32, __builtin_bit_cast(int, m))));
    return m;
}

// k-space multiplier of one spectrum point (one-component operators):
// plain Gk; grad i k_d Gk (Hermitian part, Nyquist zeroed); filtered rhs
// zeta/N; none 1.
// Geometry of tile `lin` (component, ky slab, kx column tile) and this lane's
// pencil in it.
struct MfTile {
    int comp, slab, col;
    bool cv;
};

__device__ __forceinline__ MfTile mf_tile(const PencilArgs<float>& a, long lin, int p)
{
    const long ntile = (long)a.ncoltile * a.nslab;
    MfTile t;
    t.comp = (int)(lin / ntile);
    const long rem = lin - (long)t.comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    t.slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    t.col = ct * MfmaZ::P + p;
    t.cv = t.col < a.ncol;
    return t;
}

// The one-component multipliers from the loaded Gk value g, with c0, c1
// fixed per pencil, so that the per-point work is branch-free:
//   plain  g
//   grad   (c0 + c1 kzh) g, applied as i k_d Gk: c0 = kx or ky, or c1 = 1 for kz
//   Laplacian (OP_LAPL)  (c0 - kz^2) g, c0 = -(kx^2 + ky^2)
// (Hermitian parts kzh: the Nyquist k_d is 0; kz^2 keeps the Nyquist kz.)  A
// column beyond the grid reads Gk row 0 instead: every load is in bounds and
// unconditional, and the column's results are never stored.
struct MfKCoef {
    float c0, c1, kst;
    const float* grow;
};

template <int OPK>
__device__ __forceinline__ MfKCoef mf_kcoef(const PencilArgs<float>& a, const TileK<float>& tk, int comp)
{
    MfKCoef k{0.0f, 0.0f, (float)(2.0 * kPi * a.dk[2]), tk.grow ? tk.grow : a.G};
    if constexpr (OPK == OP_GRAD) {
        k.c0 = kcomp<float>(a.kmap, comp, tk.kxh, tk.kyh, 0.0f);
        k.c1 = kcomp<float>(a.kmap, comp, 0.0f, 0.0f, 1.0f);
    } else if constexpr (OPK == OP_LAPL) {
        k.c0 = -(float)(tk.kxf * tk.kxf + tk.kyf * tk.kyf);
    }
    return k;
}

// multiplier at spectrum index kz from the loaded Gk value g
template <int OPK>
__device__ __forceinline__ float mf_kval(const PencilArgs<float>& a, const MfKCoef& k, int kz, float g)
{
    if constexpr (OPK == OP_PLAIN) {
        return g;
    } else {
        const float kzw = (float)((2 * kz < a.Mz) ? kz : kz - a.Mz) * k.kst;
        if constexpr (OPK == OP_GRAD) {
            const float kzh = (2 * kz == a.Mz) ? 0.0f : kzw;
            return (k.c0 + k.c1 * kzh) * g;
        } else {
            return (k.c0 - kzw * kzw) * g;
        }
    }
}

// filtered-rhs multiplier zeta(|k|) / N at spectrum index kz
__device__ __forceinline__ float mf_rhsf(const PencilArgs<float>& a, const TileK<float>& tk, int kz)
{
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
    return (float)zf;
}

// k-space multipliers of the eight spectrum points kz = bp + 32 q' (q' = 4q + e
// and 16 + 4q + e) that lane group q holds for butterfly bp, for the
// one-component operators; computed before the DFT stages (with the input
// loads), so that only FP32 products sit between the matrix-core stages:
//   plain  Gk                        (real)
//   grad   k_d Gk, applied as i k_d Gk (Hermitian part: Nyquist k_d = 0)
//   filtered rhs  zeta / N           (real)
//   none   1
// goff0 / goff1: byte offsets of Gk rows kzo = kz (q' < 16) and kzo = Mz - kz
// (q' >= 16) for e = 0, each stepping by -+gstep bytes (32 rows) per e.
template <int OPK>
__device__ __forceinline__ void mf_kfactors(float (&f)[8], const PencilArgs<float>& a, const TileK<float>& tk,
                                            int comp, int bp, int q, int goff0, int goff1, int gstep)
{
    if constexpr (OPK == OP_PLAIN || OPK == OP_GRAD || OPK == OP_LAPL) {
        const MfKCoef k = mf_kcoef<OPK>(a, tk, comp);
        float g[8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const unsigned go = h ? (unsigned)goff1 - (unsigned)e * (unsigned)gstep
                                      : (unsigned)goff0 + (unsigned)e * (unsigned)gstep;
                g[4 * h + e] = *reinterpret_cast<const float*>(reinterpret_cast<const char*>(k.grow) + go);
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) f[4 * h + e] = mf_kval<OPK>(a, k, bp + 32 * (4 * q + e + 16 * h), g[4 * h + e]);
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                f[4 * h + e] = (OPK == OP_NONE) ? 1.0f : mf_rhsf(a, tk, bp + 32 * (4 * q + e + 16 * h));
    }
}

// whether the k-space multiplier is imaginary (i k_d Gk: grad)
template <int OPK> constexpr bool mf_imag_multiplier()
{
    return OPK == OP_GRAD;
}

// One tile per workgroup (4 wavefronts per SIMD, two workgroups per CU).
template <int OPK>
__global__ void __launch_bounds__(512, 4) k_zmfma1024(PencilArgs<float> a)
{
    using M = MfmaZ;
    constexpr int P = M::P, RS = M::RS, PSTR = M::PS, L = M::L;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* sd = reinterpret_cast<float2*>(smem_raw);
    float2* stw = sd + P * PSTR;
    unsigned* smax = reinterpret_cast<unsigned*>(stw + L);

    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63, q = l >> 4, j = l & 15;
    const int p = j & 7, bsub = j >> 3;
    const int bp32 = (l ^ 32) << 2;

    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col = ct * P + p;
    const bool cv = col < a.ncol;
    // columns past the grid repeat the last one: loads stay in bounds, their
    // results are never stored, and no pencil's data mixes with another's
    const int colr = cv ? col : a.ncol - 1;
    constexpr bool mimag = mf_imag_multiplier<OPK>();

    for (int i = tid; i < L; i += 512) stw[i] = a.plan.tw[i];
    if (tid < 2 * P) smax[tid] = 0u;
    const MfDft32 F = mf_dft32_frags(a.mfrag, l);

    // addresses: a uniform base per component plus 32-bit byte offsets (the
    // host checks that a component and the Gk octant are below 4 GB); the
    // position z = z0 + 32 u of a pencil sits at off(z0) + u * s32
    const char* ibase = reinterpret_cast<const char*>(a.in) + (long)comp * a.ia.comp * 8;
    char* obase = reinterpret_cast<char*>(a.out) + (long)comp * a.oa.comp * 8;
    const unsigned icol = (unsigned)(colr * a.ia.col + boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi)) * 8u;
    const unsigned ocol = (unsigned)(col * a.oa.col + boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi)) * 8u;
    const unsigned is32 = (unsigned)mf_pos_stride32(a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi) * 8u;
    const unsigned os32 = (unsigned)mf_pos_stride32(a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi) * 8u;
    float2* myp = sd + p * PSTR;

    // k-space multipliers of this lane's 16 spectrum points, computed with the
    // input loads so their latency overlaps and the DFT stages see only FP32
    // products
    const TileK<float> tk = tile_k(a, col, slab, cv);
    float gk[2][8];
    const int gkz = (int)a.g_kz;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int bp = 2 * (w + 8 * bi) + bsub;
        mf_kfactors<OPK>(gk[bi], a, tk, comp, bp, q, (bp + 128 * q) * gkz * 4, (M::L / 2 - bp - 128 * q) * gkz * 4,
                         32 * gkz * 4);
    }

    // ---- load x[b + 32 m] (m < 16; the rest is zero padding) ----
    float2 x[2][4];
    float mx = 0.0f;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int b = 2 * (w + 8 * bi) + bsub;
        const unsigned o0 = icol + (unsigned)boff(b, a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi) * 8u + (unsigned)(4 * q) * is32;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            x[bi][e] = *reinterpret_cast<const float2*>(ibase + (o0 + (unsigned)e * is32));
            mx = mf_absmax(x[bi][e], mx);
        }
    }
    __syncthreads();                       // twiddle table and max slots initialised
    atomicMax(&smax[p], __float_as_uint(mx));
    __syncthreads();
    // one scale per pencil: the forward stage-1 output (<= 16 sqrt 2 2^10)
    // is the stage-2 input without rescaling
    float s0, inv0;
    mf_scale(__uint_as_float(smax[p]), s0, inv0);

    // ---- forward stage 1: DFT over m, 32 outputs k1 per column (p, b) ----
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int b = 2 * (w + 8 * bi) + bsub;
        MfFrag B[2];
        B[0] = mf_split4(x[bi], s0, s0, a.fzero);
        const MfAcc d = mf_dft32_rows32i<1>(F, B);
        // twiddle W1024^{b k1}; stored transposed as row k1, column b
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k1 = 4 * q + e + 16 * h;
                const float2 y = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                myp[k1 * RS + b] = cmul(y, stw[(b * k1) & (L - 1)]);
            }
        }
    }
    __syncthreads();

    // ---- forward stage 2, k-space operator, inverse stage 1 (registers) ----
    float m3 = 0.0f;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int bp = 2 * (w + 8 * bi) + bsub;
        float2* row = myp + bp * RS;
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = row[4 * q + e];
            v1[e] = row[16 + 4 * q + e];
        }
        MfFrag B[2];
        B[0] = mf_split4u(v0, a.fone);
        B[1] = mf_split4u(v1, a.fone);
        MfAcc d = mf_dft32_rows32i<2>(F, B);
        // s0 X[bp + 32 q'], q' = 4q + e (r0/i0) and 16 + 4q + e (r1/i1);
        // the k-space operator keeps the factor s0
        float2 z0[4], z1[4];
        float mz = 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m0 = gk[bi][e], m1 = gk[bi][4 + e];
            if constexpr (mimag) {
                // i k_d Gk: purely imaginary
                z0[e] = float2{-d.i0[e] * m0, d.r0[e] * m0};
                z1[e] = float2{-d.i1[e] * m1, d.r1[e] * m1};
            } else {
                z0[e] = float2{d.r0[e] * m0, d.i0[e] * m0};
                z1[e] = float2{d.r1[e] * m1, d.i1[e] * m1};
            }
            mz = mf_absmax(z0[e], mf_absmax(z1[e], mz));
        }
        float s2, inv2;
        mf_scale(mf_colmax(mz, bp32), s2, inv2);
        // inverse: conj(F) z = conj(F conj(z))
        B[0] = mf_split4(z0, s2, -s2, a.fzero);
        B[1] = mf_split4(z1, s2, -s2, a.fzero);
        d = mf_dft32_rows32i<2>(F, B);
        // U[n2] = conj of the inverse DFT output (times s0 s2); apply
        // W1024^{-n2 bp} and undo both scales: conj(U W) inv0 inv2, written
        // in place as row bp, column n2
        const float f = inv0 * inv2;
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int n2 = 4 * q + e + 16 * h;
                const float2 u = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                const float2 t = cmul(u, stw[(n2 * bp) & (L - 1)]);
                const float2 r = float2{t.x * f, -t.y * f};
                row[n2] = r;
                m3 = mf_absmax(r, m3);
            }
        }
    }
    atomicMax(&smax[P + p], __float_as_uint(m3));
    __syncthreads();

    // ---- inverse stage 2: column (p, n2), inputs over b', outputs n1 < 16 ----
    float s3, inv3;
    mf_scale(__uint_as_float(smax[P + p]), s3, inv3);
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int n2 = 2 * (w + 8 * bi) + bsub;
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = myp[(4 * q + e) * RS + n2];
            v1[e] = myp[(16 + 4 * q + e) * RS + n2];
        }
        MfFrag B[2];
        B[0] = mf_split4(v0, s3, -s3, a.fzero);
        B[1] = mf_split4(v1, s3, -s3, a.fzero);
        const MfAcc d = mf_dft32_rows16(F, B);
        if (cv) {
            const unsigned o0 = ocol + (unsigned)boff(n2, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi) * 8u + (unsigned)(4 * q) * os32;
#pragma unroll
            for (int e = 0; e < 4; ++e)
                *reinterpret_cast<float2*>(obase + (o0 + (unsigned)e * os32)) = float2{d.r0[e] * inv3, -d.i0[e] * inv3};
        }
    }
}

// ---------------------------------------------------------------------------
// Transform length 2048 (Mz of a 1024-cell unbounded z direction): the
// zero-padded 2048-point transform of a pencil x (1024 points) splits into
// two 1024-point residues,
//   X[2k']     = sum_n W1024^{n k'} x[n]                 (residue 0)
//   X[2k' + 1] = sum_n W1024^{n k'} (W2048^n x[n])       (residue 1)
// and the truncated inverse into
//   y[n] = IDFT1024(Z_0)[n] + W2048^{-n} IDFT1024(Z_1)[n],   n < 1024,
// so each residue runs the four DFT-32 stages of the 1024-point kernel, now
// with 32 inputs in forward stage 1 and 32 outputs in inverse stage 2.  The
// pencil data is loaded once and held in registers for both residues; the
// residue-0 result is held until residue 1 adds to it.  One workgroup of 8
// pencils per CU (2 wavefronts per SIMD, up to 256 registers); offsets are
// 64-bit (a component of a 1024^3 grid's spectrum exceeds 4 GB).
// ---------------------------------------------------------------------------
// P pencils per workgroup, P wavefronts: an MFMA column is (pencil
// p = j % P, butterfly b = (16 / P) blk + j / P).  P = 8 holds the pencil in
// registers for both residues; P = 4 reloads it for residue 1 (fewer
// registers, three workgroups per CU).
template <int P_> struct MfmaZ2048 {
    static constexpr int L = 2048, LH = 1024;
    static constexpr int P = P_, NW = P_, NB = 16 / P_;
    static constexpr bool HOLDX = P_ == 8;
    static constexpr int RS = 33;
    static constexpr int PS = 32 * 33 + (P_ == 8 ? 2 : 4);   // 2 or 4 (mod 16): p PS + bsub distinct
    static constexpr size_t lds_bytes()
    {
        return (size_t)P * PS * 8 + 2 * (size_t)LH * 8 + 3 * P * 4;
    }
};

// k-space multipliers of residue r: spectrum points kz = 2 (bp + 32 q') + r,
// q' = 4q + e + 16h; Gk rows kzo = min(kz, 2048 - kz)
template <int OPK>
__device__ __forceinline__ void mf2_kfactors(float (&f)[8], const PencilArgs<float>& a, const TileK<float>& tk,
                                             const MfKCoef& k, int bp, int q, int r)
{
    if constexpr (OPK == OP_PLAIN || OPK == OP_GRAD || OPK == OP_LAPL) {
        float g[8];
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int kz = 2 * (bp + 32 * (4 * q + e + 16 * h)) + r;
                const int kzo = (kz <= 1024) ? kz : 2048 - kz;
                g[4 * h + e] = k.grow[(long)kzo * a.g_kz];
            }
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                f[4 * h + e] = mf_kval<OPK>(a, k, 2 * (bp + 32 * (4 * q + e + 16 * h)) + r, g[4 * h + e]);
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
                f[4 * h + e] = (OPK == OP_NONE) ? 1.0f : mf_rhsf(a, tk, 2 * (bp + 32 * (4 * q + e + 16 * h)) + r);
    }
}

template <int OPK, int P_>
__global__ void __launch_bounds__(64 * P_, P_ == 8 ? 2 : 3) k_zmfma2048(PencilArgs<float> a)
{
    using M = MfmaZ2048<P_>;
    constexpr int P = M::P, RS = M::RS, PSTR = M::PS, LH = M::LH, NB = M::NB, NW = M::NW;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* sd = reinterpret_cast<float2*>(smem_raw);
    float2* stw = sd + P * PSTR;            // W1024^j, j < 1024
    float2* stw2 = stw + LH;                 // W2048^n, n < 1024
    unsigned* smax = reinterpret_cast<unsigned*>(stw2 + LH);   // [input | inverse r=0 | inverse r=1][P]

    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63, q = l >> 4, j = l & 15;
    const int p = j % P, bsub = j / P;
    const int bp32 = (l ^ 32) << 2;

    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col = ct * P + p;
    const bool cv = col < a.ncol;
    const int colr = cv ? col : a.ncol - 1;
    constexpr bool mimag = mf_imag_multiplier<OPK>();

    // the plan's table holds W2048^j, j < 2048: W1024^j = W2048^{2j}
    for (int i = tid; i < LH; i += 64 * NW) {
        stw[i] = a.plan.tw[2 * i];
        stw2[i] = a.plan.tw[i];
    }
    if (tid < 3 * P) smax[tid] = 0u;
    const MfDft32 F = mf_dft32_frags(a.mfrag, l);

    const float2* ib = reinterpret_cast<const float2*>(a.in) + (long)comp * a.ia.comp + (long)colr * a.ia.col +
                       boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi);
    float2* ob = reinterpret_cast<float2*>(a.out) + (long)comp * a.oa.comp + (long)col * a.oa.col +
                 boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi);
    const long is32 = mf_pos_stride32(a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi);
    const long os32 = mf_pos_stride32(a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi);
    float2* myp = sd + p * PSTR;
    const TileK<float> tk = tile_k(a, col, slab, cv);
    const MfKCoef kc = mf_kcoef<OPK>(a, tk, comp);
    int bb[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) bb[bi] = NB * (w + NW * bi) + bsub;

    // ---- load x[b + 32 m], m < 32: the whole nonzero pencil ----
    float2 x[2][2][4];
    float mx = 0.0f;
    auto load_x = [&]() {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const float2* ip = ib + boff(bb[bi], a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi) + (long)(4 * q) * is32;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e) x[bi][h][e] = ip[(long)(e + 16 * h) * is32];
        }
    };
    load_x();
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e) mx = mf_absmax(x[bi][h][e], mx);
    __syncthreads();                       // tables and scale slots initialised
    atomicMax(&smax[p], __float_as_uint(mx));
    __syncthreads();
    // one input scale per pencil for both residues (|W2048^n| = 1)
    float s0, inv0;
    mf_scale(__uint_as_float(smax[p]), s0, inv0);

    float2 y0[2][8];                       // residue-0 result at n = n2 + 32 n1
#pragma unroll
    for (int r = 0; r < 2; ++r) {
        // residue 1 without the held pencil: read it again (the output is
        // stored only after both residues, so in-place passes read the input)
        if constexpr (!M::HOLDX)
            if (r == 1) load_x();
        float gk[2][8];
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) mf2_kfactors<OPK>(gk[bi], a, tk, kc, bb[bi], q, r);

        // ---- forward stage 1: DFT over m (32 inputs), 32 outputs k1 per column (p, b) ----
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int b = bb[bi];
            MfFrag B[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                float2 v[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float2 xv = x[bi][h][e];
                    // residue 1: x[n] W2048^n, n = b + 32 (4q + e + 16h)
                    v[e] = r ? cmul(xv, stw2[b + 32 * (4 * q + e + 16 * h)]) : xv;
                }
                B[h] = mf_split4(v, s0, s0, a.fzero);
            }
            const MfAcc d = mf_dft32_rows32i<2>(F, B);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int k1 = 4 * q + e + 16 * h;
                    const float2 yv = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                    myp[k1 * RS + b] = cmul(yv, stw[(b * k1) & (LH - 1)]);
                }
            }
        }
        __syncthreads();

        // ---- forward stage 2, k-space operator, inverse stage 1 (registers) ----
        float m3 = 0.0f;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int bp = bb[bi];
            float2* row = myp + bp * RS;
            float2 v0[4], v1[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v0[e] = row[4 * q + e];
                v1[e] = row[16 + 4 * q + e];
            }
            MfFrag B[2];
            B[0] = mf_split4u(v0, a.fone);
            B[1] = mf_split4u(v1, a.fone);
            MfAcc d = mf_dft32_rows32i<2>(F, B);
            float2 z0[4], z1[4];
            float mz = 0.0f;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float m0 = gk[bi][e], m1 = gk[bi][4 + e];
                if constexpr (mimag) {
                    z0[e] = float2{-d.i0[e] * m0, d.r0[e] * m0};
                    z1[e] = float2{-d.i1[e] * m1, d.r1[e] * m1};
                } else {
                    z0[e] = float2{d.r0[e] * m0, d.i0[e] * m0};
                    z1[e] = float2{d.r1[e] * m1, d.i1[e] * m1};
                }
                mz = mf_absmax(z0[e], mf_absmax(z1[e], mz));
            }
            float s2, inv2;
            mf_scale(mf_colmax(mz, bp32), s2, inv2);
            B[0] = mf_split4(z0, s2, -s2, a.fzero);
            B[1] = mf_split4(z1, s2, -s2, a.fzero);
            d = mf_dft32_rows32i<2>(F, B);
            const float f = inv0 * inv2;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int n2 = 4 * q + e + 16 * h;
                    const float2 u = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                    const float2 t = cmul(u, stw[(n2 * bp) & (LH - 1)]);
                    const float2 rr = float2{t.x * f, -t.y * f};
                    row[n2] = rr;
                    m3 = mf_absmax(rr, m3);
                }
            }
        }
        atomicMax(&smax[P + r * P + p], __float_as_uint(m3));
        __syncthreads();

        // ---- inverse stage 2: column (p, n2), inputs over b', 32 outputs n1 ----
        float s3, inv3;
        mf_scale(__uint_as_float(smax[P + r * P + p]), s3, inv3);
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int n2 = bb[bi];
            float2 v0[4], v1[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v0[e] = myp[(4 * q + e) * RS + n2];
                v1[e] = myp[(16 + 4 * q + e) * RS + n2];
            }
            MfFrag B[2];
            B[0] = mf_split4(v0, s3, -s3, a.fzero);
            B[1] = mf_split4(v1, s3, -s3, a.fzero);
            const MfAcc d = mf_dft32_rows32i<2>(F, B);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float2 yv = h ? float2{d.r1[e] * inv3, -d.i1[e] * inv3}
                                        : float2{d.r0[e] * inv3, -d.i0[e] * inv3};
                    if (r == 0) {
                        y0[bi][4 * h + e] = yv;
                    } else {
                        // y[n] = y0[n] + W2048^{-n} y1[n], n = n2 + 32 n1
                        const int n = n2 + 32 * (4 * q + e + 16 * h);
                        const float2 wv = stw2[n];
                        const float2 t = float2{yv.x * wv.x + yv.y * wv.y, yv.y * wv.x - yv.x * wv.y};
                        y0[bi][4 * h + e] = float2{y0[bi][4 * h + e].x + t.x, y0[bi][4 * h + e].y + t.y};
                    }
                }
            }
        }
        // the exchange buffer is rewritten by the next residue's stage 1
        if (r == 0) __syncthreads();
    }
    if (cv) {
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            float2* op_ = ob + boff(bb[bi], a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi) + (long)(4 * q) * os32;
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int e = 0; e < 4; ++e) op_[(long)(e + 16 * h) * os32] = y0[bi][4 * h + e];
        }
    }
}

// ---------------------------------------------------------------------------
// Transform length 512 = 32 x 16 (Mz of a 256-cell unbounded z direction):
//   X[k1 + 32 k2] = sum_b W16^{b k2} W512^{b k1} sum_{m<16} W32^{m k1} x[b + 16 m]
//   x[b + 16 m]   = sum_k1 W32^{-m k1} W512^{-b k1} sum_k2 W16^{-b k2} Z[k1 + 32 k2]
// (b < 16, k1 < 32, k2 < 16; inputs and kept outputs m < 16).  Forward stage
// 1 (16 inputs) and inverse stage 2 (16 outputs) are the 32-point stages of
// the 1024-point kernel; the two middle stages are 16-point DFTs of one K
// chunk each, with the k-space operator between them in registers.
// Workgroup: P = 16 pencils (adjacent kx columns: 128-byte rows per z
// position) x the whole z extent, 8 wavefronts; an MFMA column is one pencil
// (p = lane & 15).  Wavefront w owns b = w, w + 8 in the 32-point stages and
// k1 = w + 8 bi (bi < 4) in the 16-point ones.  Offsets are 64-bit.
// ---------------------------------------------------------------------------
struct MfmaZ512 {
    static constexpr int L = 512;
    static constexpr int P = 16, NW = 8;
    // LDS [p][k1 < 32][b < 16]: in every exchange the 16 lanes of a group
    // differ in p only, and PS = 1 (mod 16) puts their 64-bit accesses on 16
    // distinct bank pairs
    static constexpr int RS = 16;
    static constexpr int PS = 32 * 16 + 1;
    static constexpr size_t lds_bytes()
    {
        return (size_t)P * PS * 8 + (size_t)L * 8 + 2 * P * 4;
    }
};

// A fragments of the 16-point forward DFT in real form (one K chunk: points
// m = 0..15, re/im interleaved; rows 0..15):
//   Yr row k: ( cos t,  sin t),  Yi row k: (-sin t, cos t),  t = 2 pi k m / 16
// Table layout (built by zmfma_fragments16): [lane][rh, rl, ih, il][8 halves].
struct MfDft16 {
    mf_h8 rh, rl, ih, il;
};

__device__ __forceinline__ MfDft16 mf_dft16_frags(const void* table, int lane)
{
    const mf_h8* t = reinterpret_cast<const mf_h8*>(table) + lane * 4;
    return MfDft16{t[0], t[1], t[2], t[3]};
}

struct MfAcc16 {
    mf_f4 r, i;
};

// the 16-point DFT of 16 columns (lane group q holds points, and receives
// outputs, 4q..4q+3)
__device__ __forceinline__ MfAcc16 mf_dft16(const MfDft16& F, const MfFrag& B)
{
    MfAcc16 d;
    d.r = d.i = mf_f4{0, 0, 0, 0};
    d.r = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh, B.hi, d.r, 0, 0, 0);
    d.i = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih, B.hi, d.i, 0, 0, 0);
    d.r = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh, B.lo, d.r, 0, 0, 0);
    d.i = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih, B.lo, d.i, 0, 0, 0);
    d.r = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl, B.hi, d.r, 0, 0, 0);
    d.i = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il, B.hi, d.i, 0, 0, 0);
    return d;
}

// k-space multipliers of spectrum points kz = k1 + 32 (4q + e), e < 4;
// Gk rows kzo = min(kz, 512 - kz)
template <int OPK>
__device__ __forceinline__ void mf5_kfactors(float (&f)[4], const PencilArgs<float>& a, const TileK<float>& tk,
                                             const MfKCoef& k, int k1, int q)
{
    if constexpr (OPK == OP_PLAIN || OPK == OP_GRAD || OPK == OP_LAPL) {
        float g[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int kz = k1 + 32 * (4 * q + e);
            const int kzo = (kz <= 256) ? kz : 512 - kz;
            g[e] = k.grow[(long)kzo * a.g_kz];
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) f[e] = mf_kval<OPK>(a, k, k1 + 32 * (4 * q + e), g[e]);
    } else {
#pragma unroll
        for (int e = 0; e < 4; ++e) f[e] = (OPK == OP_NONE) ? 1.0f : mf_rhsf(a, tk, k1 + 32 * (4 * q + e));
    }
}

template <int OPK>
__global__ void __launch_bounds__(512, 4) k_zmfma512(PencilArgs<float> a)
{
    using M = MfmaZ512;
    constexpr int P = M::P, RS = M::RS, PSTR = M::PS, L = M::L;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* sd = reinterpret_cast<float2*>(smem_raw);
    float2* stw = sd + P * PSTR;            // W512^j, j < 512
    unsigned* smax = reinterpret_cast<unsigned*>(stw + L);   // [input | inverse stage-2 input][P]

    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63, q = l >> 4, p = l & 15;
    const int bp32 = (l ^ 32) << 2;

    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const long ntile = (long)a.ncoltile * a.nslab;
    const int comp = (int)(lin / ntile);
    const long rem = lin - (long)comp * ntile;
    const int slab0 = (int)(rem / a.ncoltile);
    const int ct = (int)(rem - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col = ct * P + p;
    const bool cv = col < a.ncol;
    // columns past the grid repeat the last one: loads stay in bounds, their
    // results are never stored, and no pencil's data mixes with another's
    const int colr = cv ? col : a.ncol - 1;
    constexpr bool mimag = mf_imag_multiplier<OPK>();

    for (int i = tid; i < L; i += 64 * M::NW) stw[i] = a.plan.tw[i];
    if (tid < 2 * P) smax[tid] = 0u;
    const MfDft32 F = mf_dft32_frags(a.mfrag, l);
    const MfDft16 F16 = mf_dft16_frags(a.mfrag16, l);

    const float2* ib = reinterpret_cast<const float2*>(a.in) + (long)comp * a.ia.comp + (long)colr * a.ia.col +
                       boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi);
    float2* ob = reinterpret_cast<float2*>(a.out) + (long)comp * a.oa.comp + (long)col * a.oa.col +
                 boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi);
    float2* myp = sd + p * PSTR;

    // k-space multipliers of the four 16-point columns k1 = w + 8 bi
    const TileK<float> tk = tile_k(a, col, slab, cv);
    const MfKCoef kc = mf_kcoef<OPK>(a, tk, comp);
    float gk[4][4];
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) mf5_kfactors<OPK>(gk[bi], a, tk, kc, w + 8 * bi, q);

    // ---- load x[b + 16 m], m = 4q + e < 16 (the rest is zero padding) ----
    float2 x[2][4];
    float mx = 0.0f;
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int b = w + 8 * bi;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            x[bi][e] = ib[boff(b + 16 * (4 * q + e), a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi)];
            mx = mf_absmax(x[bi][e], mx);
        }
    }
    __syncthreads();                       // twiddle table and max slots initialised
    atomicMax(&smax[p], __float_as_uint(mx));
    __syncthreads();
    // one scale per pencil: the forward stage-1 output (<= 16 sqrt 2 2^10)
    // is the stage-2 input without rescaling
    float s0, inv0;
    mf_scale(__uint_as_float(smax[p]), s0, inv0);

    // ---- forward stage 1: column (p, b), DFT-32 over m, 32 outputs k1 ----
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int b = w + 8 * bi;
        MfFrag B[2];
        B[0] = mf_split4(x[bi], s0, s0, a.fzero);
        const MfAcc d = mf_dft32_rows32i<1>(F, B);
        // twiddle W512^{b k1}; stored as row k1, column b
#pragma unroll
        for (int h = 0; h < 2; ++h) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int k1 = 4 * q + e + 16 * h;
                const float2 y = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                myp[k1 * RS + b] = cmul(y, stw[(b * k1) & (L - 1)]);
            }
        }
    }
    __syncthreads();

    // ---- forward stage 2 (DFT-16 over b), k-space operator, inverse stage 1
    //      (DFT-16 over k2), column (p, k1) ----
    float m3 = 0.0f;
#pragma unroll
    for (int bi = 0; bi < 4; ++bi) {
        const int k1 = w + 8 * bi;
        float2* row = myp + k1 * RS;
        float2 v[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) v[e] = row[4 * q + e];
        MfFrag B = mf_split4u(v, a.fone);
        MfAcc16 d = mf_dft16(F16, B);
        // s0 X[k1 + 32 k2], k2 = 4q + e; the k-space operator keeps the factor s0
        float2 z[4];
        float mz = 0.0f;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m0 = gk[bi][e];
            if constexpr (mimag) z[e] = float2{-d.i[e] * m0, d.r[e] * m0};
            else z[e] = float2{d.r[e] * m0, d.i[e] * m0};
            mz = mf_absmax(z[e], mz);
        }
        float s2, inv2;
        mf_scale(mf_colmax(mz, bp32), s2, inv2);
        // inverse: conj(F) z = conj(F conj(z))
        B = mf_split4(z, s2, -s2, a.fzero);
        d = mf_dft16(F16, B);
        // output b = 4q + e: conj(d W512^{b k1}) inv0 inv2, in place as row k1, column b
        const float f = inv0 * inv2;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int b = 4 * q + e;
            const float2 t = cmul(float2{d.r[e], d.i[e]}, stw[(b * k1) & (L - 1)]);
            const float2 r = float2{t.x * f, -t.y * f};
            row[b] = r;
            m3 = mf_absmax(r, m3);
        }
    }
    atomicMax(&smax[P + p], __float_as_uint(m3));
    __syncthreads();

    // ---- inverse stage 2: column (p, b), inputs over k1, outputs m < 16 ----
    float s3, inv3;
    mf_scale(__uint_as_float(smax[P + p]), s3, inv3);
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) {
        const int b = w + 8 * bi;
        float2 v0[4], v1[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            v0[e] = myp[(4 * q + e) * RS + b];
            v1[e] = myp[(16 + 4 * q + e) * RS + b];
        }
        MfFrag B[2];
        B[0] = mf_split4(v0, s3, -s3, a.fzero);
        B[1] = mf_split4(v1, s3, -s3, a.fzero);
        const MfAcc d = mf_dft32_rows16(F, B);
        if (cv) {
#pragma unroll
            for (int e = 0; e < 4; ++e)
                ob[boff(b + 16 * (4 * q + e), a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi)] =
                    float2{d.r0[e] * inv3, -d.i0[e] * inv3};
        }
    }
}

// ---------------------------------------------------------------------------
// Three-component operators (div, curl) of the fused z pass on the matrix
// cores, L = 1024.  Workgroup: P3 = 4 pencils, 4 wavefronts; an MFMA column
// is (pencil p = j & 3, butterfly b = 4 blk + (j >> 2)).  Each input
// component runs forward stages 1 and 2 through one LDS exchange buffer and
// its spectrum stays in registers (three per lane); each output's k-space
// combination then runs inverse stages 1 and 2 through the same buffer.
// Per-pencil scales per component (forward) and per output (inverse stage
// 2), per-column scales for inverse stage 1, as in the one-component kernel.
// ---------------------------------------------------------------------------
struct MfmaZ3 {
    static constexpr int L = 1024;
    static constexpr int P = 4;          // pencils per workgroup
    static constexpr int NW = 4;         // wavefronts per workgroup
    static constexpr int NB = 16 / P;    // butterflies per MFMA block
    static constexpr int RS = 33;        // LDS row stride: 1 (mod 16)
    static constexpr int PS = 32 * 33 + 4;   // LDS pencil stride: 4 (mod 16), so 4p + bsub is distinct
    static constexpr size_t lds_bytes()
    {
        return (size_t)P * PS * 8 + (size_t)L * 8 + 6 * P * 4;
    }
};

template <int OPK>
__global__ void __launch_bounds__(256, 2) k_zmfma1024v(PencilArgs<float> a)
{
    using M = MfmaZ3;
    constexpr int P = M::P, RS = M::RS, PSTR = M::PS, L = M::L, NB = M::NB, W = M::NW;
    constexpr int NOUT = (OPK == OP_DIV) ? 1 : 3;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    float2* sd = reinterpret_cast<float2*>(smem_raw);
    float2* stw = sd + P * PSTR;
    unsigned* smax = reinterpret_cast<unsigned*>(stw + L);   // [3][P] forward, [3][P] inverse

    const int tid = threadIdx.x;
    const int w = tid >> 6, l = tid & 63, q = l >> 4, j = l & 15;
    const int p = j % P, bsub = j / P;
    const int bp32 = (l ^ 32) << 2;

    const long nblk = gridDim.x;
    const long lin = a.xcd_remap ? xcd_remap(blockIdx.x, nblk) : (long)blockIdx.x;
    const int slab0 = (int)(lin / a.ncoltile);
    const int ct = (int)(lin - (long)slab0 * a.ncoltile);
    const int slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    const int col = ct * P + p;
    const bool cv = col < a.ncol;
    const int colr = cv ? col : a.ncol - 1;

    for (int i = tid; i < L; i += 64 * W) stw[i] = a.plan.tw[i];
    if (tid < 6 * P) smax[tid] = 0u;
    const MfDft32 F = mf_dft32_frags(a.mfrag, l);

    // uniform bases and 32-bit byte offsets (host-checked below 4 GB)
    const char* ibase = reinterpret_cast<const char*>(a.in);
    char* obase = reinterpret_cast<char*>(a.out);
    const unsigned icol = (unsigned)(colr * a.ia.col + boff(slab, a.ia.slab_sh, a.ia.slab_lo, a.ia.slab_hi)) * 8u;
    const unsigned ocol = (unsigned)(col * a.oa.col + boff(slab, a.oa.slab_sh, a.oa.slab_lo, a.oa.slab_hi)) * 8u;
    const unsigned is32 = (unsigned)mf_pos_stride32(a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi) * 8u;
    const unsigned os32 = (unsigned)mf_pos_stride32(a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi) * 8u;
    float2* myp = sd + p * PSTR;
    int bb[2];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi) bb[bi] = NB * (w + W * bi) + bsub;

    // Green's octant values and wavenumbers (Hermitian parts: Nyquist 0)
    const TileK<float> tk = tile_k(a, col, slab, cv);
    const int gkz = (int)a.g_kz;
    float gk[2][8];
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
        mf_kfactors<OP_PLAIN>(gk[bi], a, tk, 0, bb[bi], q, (bb[bi] + 128 * q) * gkz * 4,
                              (M::L / 2 - bb[bi] - 128 * q) * gkz * 4, 32 * gkz * 4);
    const float kzc = (float)(2.0 * kPi * a.dk[2]);

    // ---- forward: each input component through stages 1 and 2 ----
    float2 X[3][2][8];
    float xinv[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float2 x[2][4];
        float mx = 0.0f;
        const char* cb = ibase + (long)c * a.ia.comp * 8;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const unsigned o0 = icol + (unsigned)boff(bb[bi], a.ia.pos_sh, a.ia.pos_lo, a.ia.pos_hi) * 8u +
                                (unsigned)(4 * q) * is32;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                x[bi][e] = *reinterpret_cast<const float2*>(cb + (o0 + (unsigned)e * is32));
                mx = mf_absmax(x[bi][e], mx);
            }
        }
        __syncthreads();      // tables and slots ready (c = 0); previous stage-2 reads done (c > 0)
        atomicMax(&smax[c * P + p], __float_as_uint(mx));
        __syncthreads();
        float s0, inv0;
        mf_scale(__uint_as_float(smax[c * P + p]), s0, inv0);
        xinv[c] = inv0;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int b = bb[bi];
            MfFrag B[2];
            B[0] = mf_split4(x[bi], s0, s0, a.fzero);
            const MfAcc d = mf_dft32_rows32i<1>(F, B);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int k1 = 4 * q + e + 16 * h;
                    const float2 y = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                    myp[k1 * RS + b] = cmul(y, stw[(b * k1) & (L - 1)]);
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const float2* row = myp + bb[bi] * RS;
            float2 v0[4], v1[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v0[e] = row[4 * q + e];
                v1[e] = row[16 + 4 * q + e];
            }
            MfFrag B[2];
            B[0] = mf_split4u(v0, a.fone);
            B[1] = mf_split4u(v1, a.fone);
            const MfAcc d = mf_dft32_rows32i<2>(F, B);
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                X[c][bi][e] = float2{d.r0[e], d.i0[e]};
                X[c][bi][4 + e] = float2{d.r1[e], d.i1[e]};
            }
        }
    }

    // ---- each output: k-space combination, inverse stages 1 and 2 ----
#pragma unroll
    for (int o = 0; o < NOUT; ++o) {
        __syncthreads();      // previous output's stage-2 reads (or forward stage-2 reads) done
        float m3 = 0.0f;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int bp = bb[bi];
            float2 z[2][4];
            float mz = 0.0f;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int kz = bp + 32 * (4 * q + e + 16 * h);
                    const int kzs = (kz < M::L / 2) ? kz : kz - M::L;
                    const float kzh = (kz == M::L / 2) ? 0.0f : (float)kzs * kzc;
                    const float g = gk[bi][4 * h + e];
                    const float2 sx = X[0][bi][4 * h + e], sy = X[1][bi][4 * h + e], sz = X[2][bi][4 * h + e];
                    const float fx = xinv[0], fy = xinv[1], fz = xinv[2];
                    float2 t;   // k-weighted sum before g * i
                    if constexpr (OPK == OP_DIV) {
                        const float ax = tk.kxh * fx, ay = tk.kyh * fy, az = kzh * fz;
                        t = float2{ax * sx.x + ay * sy.x + az * sz.x, ax * sx.y + ay * sy.y + az * sz.y};
                    } else if (o == 0) {
                        const float ay = tk.kyh * fz, az = kzh * fy;
                        t = float2{ay * sz.x - az * sy.x, ay * sz.y - az * sy.y};
                    } else if (o == 1) {
                        const float az = kzh * fx, ax = tk.kxh * fz;
                        t = float2{az * sx.x - ax * sz.x, az * sx.y - ax * sz.y};
                    } else {
                        const float ax = tk.kxh * fy, ay = tk.kyh * fx;
                        t = float2{ax * sy.x - ay * sx.x, ax * sy.y - ay * sx.y};
                    }
                    z[h][e] = float2{-g * t.y, g * t.x};   // g * i * t
                    mz = mf_absmax(z[h][e], mz);
                }
            }
            float s2, inv2;
            mf_scale(mf_colmax(mz, bp32), s2, inv2);
            MfFrag B[2];
            B[0] = mf_split4(z[0], s2, -s2, a.fzero);
            B[1] = mf_split4(z[1], s2, -s2, a.fzero);
            const MfAcc d = mf_dft32_rows32i<2>(F, B);
            float2* row = myp + bp * RS;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int n2 = 4 * q + e + 16 * h;
                    const float2 u = h ? float2{d.r1[e], d.i1[e]} : float2{d.r0[e], d.i0[e]};
                    const float2 t = cmul(u, stw[(n2 * bp) & (L - 1)]);
                    const float2 r = float2{t.x * inv2, -t.y * inv2};
                    row[n2] = r;
                    m3 = mf_absmax(r, m3);
                }
            }
        }
        atomicMax(&smax[3 * P + o * P + p], __float_as_uint(m3));
        __syncthreads();
        float s3, inv3;
        mf_scale(__uint_as_float(smax[3 * P + o * P + p]), s3, inv3);
        char* ob = obase + (long)o * a.oa.comp * 8;
#pragma unroll
        for (int bi = 0; bi < 2; ++bi) {
            const int n2 = bb[bi];
            float2 v0[4], v1[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                v0[e] = myp[(4 * q + e) * RS + n2];
                v1[e] = myp[(16 + 4 * q + e) * RS + n2];
            }
            MfFrag B[2];
            B[0] = mf_split4(v0, s3, -s3, a.fzero);
            B[1] = mf_split4(v1, s3, -s3, a.fzero);
            const MfAcc d = mf_dft32_rows16(F, B);
            if (cv) {
                const unsigned o0 = ocol + (unsigned)boff(n2, a.oa.pos_sh, a.oa.pos_lo, a.oa.pos_hi) * 8u +
                                    (unsigned)(4 * q) * os32;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    *reinterpret_cast<float2*>(ob + (o0 + (unsigned)e * os32)) =
                        float2{d.r0[e] * inv3, -d.i0[e] * inv3};
            }
        }
    }
}

}  // namespace psamd
