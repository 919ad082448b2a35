// Mixed-precision z pass on the matrix cores, transform length L = 1024
// (Mz of a 512-cell unbounded z direction), register-pencil form.
//
// One wavefront owns one pencil (kx, ky) for the whole fused pass (pad nz ->
// Mz, FFT, k-space operator, IFFT, keep nz; the reference's
// solve3d.cpp:243-441).  The 1024-point transform is the four-step
//   X[k1 + 32 k2] = sum_b W32^(b k2) W1024^(b k1) sum_{m<16} W32^(m k1) x[b + 32 m]
//   x[32 n1 + n2] = sum_k1 W32^(-n1 k1) W1024^(-n2 k1) sum_k2 W32^(-n2 k2) Z[k1 + 32 k2]
// whose DFT stages are v_mfma_f32_16x16x32_f16 products with fp16 hi/lo split
// operands (A_hi B_hi + A_hi B_lo + A_lo B_hi, fp32 accumulation).  The
// pencil never leaves the wavefront's registers between stages: an MFMA
// accumulator D[i][j] (column j on lanes j mod 16, rows i in groups of four
// per lane group) is, element for element, both the B operand with K = i and
// the A operand of the transpose (rows j, K = i).  Stage 1 multiplies the data
// (A) by the DFT (B); stage 2 the DFT (A) by the twiddled stage-1 result (B);
// inverse stage 1 the conjugated spectrum (A) by the DFT (B); inverse stage 2
// the DFT (A) by the twiddled result (B).  No exchange between stages, so no
// workgroup barrier inside the transform.
//
// Every 32-point stage is split into two 16-point products:
//   stage 1: even k1 by W16^(m k1'), odd k1 by W32^(m (2 k1' + 1))
//   stage 2: even b by W16^(k2 b'), odd b by W32^(k2 (2 b' + 1)); rows k2 and
//            k2 + 16 are E + O and E - O
//   inverse 1: A = Z[k2] + Z[k2 + 16] gives the even n2 by W16^(n' k2),
//            B = Z[k2] - Z[k2 + 16] the odd n2 by W32^(k2 (2 n' + 1))
//   inverse 2: even k1 by W16^(n1 k1'), odd k1 by W32^(n1 (2 k1' + 1))
// (six MFMAs per 16 x 16 x 16 complex product; 96 per pencil), with three
// 16 x 16 fp16 hi/lo tables: E16 = W16^(ik), P = W32^(k (2i + 1)),
// Q = W32^(i (2k + 1)).  The data index layout follows: lane j of a stage-1
// operand holds b = 2 j + tb, a stage-1 output column k1 = 2 j + tk.
//
// Workgroup: 16 wavefronts (one per CU, persistent), tiles of 16 adjacent kx
// columns of one ky slab, so every HBM row access is a full 128-byte line.
// A tile (512 z rows x 128 B) is fetched by direct global->LDS loads into
// one of two LDS buffers while the other tile is transformed; each wavefront
// reads its pencil (one column) from the buffer, transforms it and writes
// the result back into the same column; the workgroup then stores the tile.
//
// Scales: fp16 spans 2^-24 .. 65504, so every split input is scaled by a
// power of two mapping its largest magnitude below 2^10 (per pencil: a
// wavefront-wide maximum); stage-2 inputs (<= 16 sqrt 2 2^10) need none.
#pragma once
#include <hip/hip_runtime.h>

#include "ps_mfma.hpp"

namespace psamd {

struct MfmaZR {
    static constexpr int L = 1024, NZ = 512, P = 16, NW = 16;
    static constexpr int STG = NZ * P * 8;            // 64 KB per staging buffer
    static constexpr int TWB = 16 * 64 * 8;           // twiddles [slot][lane] (complex)
    static constexpr int FTB = 3 * 4 * 64 * 16;       // [E16, P, Q][rh, ih, rl, il][lane]
    static constexpr size_t lds_bytes() { return 2 * (size_t)STG + TWB + FTB; }
};

typedef __attribute__((address_space(3))) void* mr_lds_ptr;

// Staging layout: 1 KB blocks n of 8 rows of 128 bytes (one direct load
// instruction each); row r of block n sits in row slot (r + rot) & 7 with
// rot = (n >> 1) & 1, and chunk c (pencils 2 c, 2 c + 1) in chunk slot
// (c ^ r ^ (n & 1)) & 7.  A wavefront's pencil column (rows 2 j + q of blocks
// (j >> 2) + 16 g + ...) then covers all 16 of the 16-byte positions of a
// 256-byte bank period with its p & 1 half, 4 lanes each: the fewest
// conflicts an 8-byte column allows.
__device__ __forceinline__ unsigned mr_stg_pos(int r, int c, int rot, int par)
{
    return (unsigned)(128 * ((r + rot) & 7) + 16 * ((c ^ r ^ par) & 7));
}

// Wavefront-wide maximum of non-negative floats (their bit patterns order as
// unsigned integers): DPP within rows of 16 lanes, then row broadcasts;
// returned uniform.
__device__ __forceinline__ float mr_wave_max(float mf)
{
    unsigned m = __float_as_uint(mf);
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0xB1, 0xF, 0xF, false));    // quad [1,0,3,2]
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x4E, 0xF, 0xF, false));    // quad [2,3,0,1]
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x141, 0xF, 0xF, false));   // row half mirror
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp(0, (int)m, 0x140, 0xF, 0xF, false));   // row mirror
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x142, 0xA, 0xF, false));   // row bcast 15
    m = max(m, (unsigned)__builtin_amdgcn_update_dpp((int)m, (int)m, 0x143, 0xC, 0xF, false));   // row bcast 31
    return __uint_as_float((unsigned)__builtin_amdgcn_readlane((int)m, 63));
}

// Power-of-two scale s with s m < 2^10 and its inverse, from the exponent
// bits of the uniform maximum m >= 0 (clamped to 2^+-100 as mf_scale).
__device__ __forceinline__ void mr_scale(float m, float& s, float& inv)
{
    int e = (int)((__float_as_uint(m) >> 23) & 255) - 126;   // m < 2^e for normal m
    e = e < -100 ? -100 : (e > 100 ? 100 : e);
    s = __uint_as_float((unsigned)(127 + 10 - e) << 23);
    inv = __uint_as_float((unsigned)(127 - 10 + e) << 23);
}

// Inline assembly here never writes a register the compiler has not just
// written itself: every output is tied to an input produced by compiler-visible
// code, so the wait states that protect an in-flight MFMA's operands from a
// later write (which the compiler does not count for inline assembly) are
// already in place.  A trailing nop gives the results the two wait states a
// DPP or MFMA read of a VGPR written by a VALU instruction needs.

// largest |re|, |im| of eight complex values: max3 with absolute-value
// modifiers, one instruction per two values, on no NaN-canonicalised copies
// of memory operands (a NaN is skipped by the maximum and propagates through
// the transform regardless)
__device__ __forceinline__ float mr_absmax8(const float2 (&u)[4], const float2 (&v)[4])
{
    float m = __builtin_fabsf(u[0].x);
    asm("v_max3_f32 %0, |%1|, |%2|, %0\n\t"
        "v_max3_f32 %0, |%3|, |%4|, %0\n\t"
        "v_max3_f32 %0, |%5|, |%6|, %0\n\t"
        "v_max3_f32 %0, |%7|, |%8|, %0\n\t"
        "v_max3_f32 %0, |%9|, |%10|, %0\n\t"
        "v_max3_f32 %0, |%11|, |%12|, %0\n\t"
        "v_max3_f32 %0, |%13|, |%14|, %0\n\t"
        "v_max3_f32 %0, |%15|, |%15|, %0\n\t"
        "s_nop 1"
        : "+v"(m)
        : "v"(u[0].y), "v"(u[1].x), "v"(u[1].y), "v"(u[2].x), "v"(u[2].y), "v"(u[3].x), "v"(u[3].y), "v"(v[0].x),
          "v"(v[0].y), "v"(v[1].x), "v"(v[1].y), "v"(v[2].x), "v"(v[2].y), "v"(v[3].x), "v"(v[3].y));
    return m;
}

// fp16 hi/lo split of four complex values scaled by s (the imaginary parts
// by -s for NEGY; z a run-time zero): hi = f16(x s) rounded to nearest, by
// compiler-visible mixed FMAs, and lo = f16(x s - hi) by two v_fma_mix per
// value written into the register of the real part (op_sel picks the half of
// hi that the remainder subtracts).  The inputs are consumed.
template <bool NEGY>
__device__ __forceinline__ MfFrag mr_split4s(float2 (&v)[4], float s, float z)
{
    const float sy = NEGY ? -s : s;
    unsigned H[4], X[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        H[e] = __builtin_bit_cast(unsigned, mf_h2{(_Float16)__builtin_fmaf(v[e].x, s, z), (_Float16)__builtin_fmaf(v[e].y, sy, z)});
        X[e] = __float_as_uint(v[e].x);
    }
    asm("v_fma_mixlo_f16 %0, %0, %8, -%4 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %0, %9, %13, -%4 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixlo_f16 %1, %1, %8, -%5 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %1, %10, %13, -%5 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixlo_f16 %2, %2, %8, -%6 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %2, %11, %13, -%6 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixlo_f16 %3, %3, %8, -%7 op_sel_hi:[0,0,1]\n\t"
        "v_fma_mixhi_f16 %3, %12, %13, -%7 op_sel:[0,0,1] op_sel_hi:[0,0,1]\n\t"
        "s_nop 1"
        : "+v"(X[0]), "+v"(X[1]), "+v"(X[2]), "+v"(X[3])
        : "v"(H[0]), "v"(H[1]), "v"(H[2]), "v"(H[3]), "v"(s), "v"(v[0].y), "v"(v[1].y), "v"(v[2].y), "v"(v[3].y),
          "v"(sy));
    MfFrag f;
    f.hi = __builtin_bit_cast(mf_h8, H);
    f.lo = __builtin_bit_cast(mf_h8, X);
    return f;
}

// fp16 hi/lo split of four complex values already below 65504: hi rounded
// toward zero (the remainder x - hi is exact in fp32 and keeps the lost bits)
__device__ __forceinline__ MfFrag mr_split4u(const float2 (&v)[4], float one)
{
    unsigned H[4], Lo[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const mf_h2 h = __builtin_bit_cast(mf_h2, __builtin_amdgcn_cvt_pkrtz(v[e].x, v[e].y));
        const mf_h2 r = mf_h2{(_Float16)__builtin_fmaf(v[e].x, one, -(float)h.x),
                              (_Float16)__builtin_fmaf(v[e].y, one, -(float)h.y)};
        H[e] = __builtin_bit_cast(unsigned, h);
        Lo[e] = __builtin_bit_cast(unsigned, r);
    }
    MfFrag f;
    f.hi = __builtin_bit_cast(mf_h8, H);
    f.lo = __builtin_bit_cast(mf_h8, Lo);
    return f;
}

// the four fragments of one 16 x 16 table (real-form rows Yr, Yi, hi and lo)
struct MfT16 {
    mf_h8 rh, ih, rl, il;
};

// table tb (0 E16, 1 P, 2 Q) from LDS; the offset is opaque to the compiler
// so each product loads its table where it is used rather than holding all
// three in registers across the tile loop
__device__ __forceinline__ MfT16 mr_table(const mf_h8* ft, int tb, int l)
{
    int o = tb * 4 * 64;
    asm volatile("" : "+s"(o));
    const mf_h8* q = ft + o + l;
    MfT16 f;
    f.rh = q[0];
    f.ih = q[64];
    f.rl = q[128];
    f.il = q[192];
    return f;
}

// data (A operand, rows on lanes) times a table (B operand): r, i += ...;
// FIRST starts the accumulation chain from zero
template <bool FIRST>
__device__ __forceinline__ void mr_mfma_a(mf_f4& r, mf_f4& i, const MfFrag& X, const MfT16& F)
{
    const mf_f4 zr = FIRST ? mf_f4{0, 0, 0, 0} : r, zi = FIRST ? mf_f4{0, 0, 0, 0} : i;
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(X.hi, F.rh, zr, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(X.hi, F.ih, zi, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(X.lo, F.rh, r, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(X.lo, F.ih, i, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(X.hi, F.rl, r, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(X.hi, F.il, i, 0, 0, 0);
}

// a table (A operand) times data (B operand, columns on lanes)
template <bool FIRST>
__device__ __forceinline__ void mr_mfma_b(mf_f4& r, mf_f4& i, const MfT16& F, const MfFrag& X)
{
    const mf_f4 zr = FIRST ? mf_f4{0, 0, 0, 0} : r, zi = FIRST ? mf_f4{0, 0, 0, 0} : i;
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh, X.hi, zr, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih, X.hi, zi, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rh, X.lo, r, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.ih, X.lo, i, 0, 0, 0);
    r = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.rl, X.hi, r, 0, 0, 0);
    i = __builtin_amdgcn_mfma_f32_16x16x32_f16(F.il, X.hi, i, 0, 0, 0);
}

// Work units: unit u is (component, slab pair, column tile) and holds the
// tiles of slabs 2 sp and 2 sp + 1 of the paired slab order (ky and My - ky,
// which share a row of the kernel octant), transformed back to back by one
// workgroup so that the second reuses the first's multipliers.
struct MrTile {
    int comp, slab, col0;
    int kyo;     // octant row min(ky, My - ky)
};

__device__ __forceinline__ int mr_nunit(const PencilArgs<float>& a)
{
    return a.ncomp * ((a.nslab + 1) >> 1) * a.ncoltile;
}

// unit u's component, slab pair and column tile (tiles of P columns)
struct MrUnit {
    int comp, sp, col0;
};

template <int P = 16>
__device__ __forceinline__ MrUnit mr_unit(const PencilArgs<float>& a, int u, int nunit)
{
    int lin = u;
    if (a.xcd_remap) {
        // the grid is a multiple of 8 and workgroup g runs on XCD g mod 8:
        // give each XCD a contiguous range of units
        const int per = nunit >> 3;
        if (lin < (per << 3)) lin = (lin & 7) * per + (lin >> 3);
    }
    const int nunit1 = ((a.nslab + 1) >> 1) * a.ncoltile;
    MrUnit v;
    v.comp = lin / nunit1;
    const int rem = lin - v.comp * nunit1;
    v.sp = rem / a.ncoltile;
    v.col0 = (rem - v.sp * a.ncoltile) * P;
    return v;
}

// tile m (0, 1) of a unit; false if the unit has no such tile (odd slab count)
__device__ __forceinline__ bool mr_tile(const PencilArgs<float>& a, const MrUnit& v, int m, MrTile& z)
{
    const int slab0 = 2 * v.sp + m;
    z.comp = v.comp;
    z.col0 = v.col0;
    z.slab = a.slab_pair ? paired_slab(slab0, a.nslab) : slab0;
    z.kyo = (z.slab <= a.nslab - z.slab) ? z.slab : a.nslab - z.slab;
    return slab0 < a.nslab;
}

// byte offset of tile T's column 0 in a component array (addressing d)
__device__ __forceinline__ const char* mr_tile_base(const void* base, const Addr& d, const MrTile& T)
{
    return reinterpret_cast<const char*>(base) +
           ((long)T.comp * d.comp + (long)T.col0 * d.col + boff(T.slab, d.slab_sh, d.slab_lo, d.slab_hi)) * 8;
}

// Position z = z0 + zr of a pencil, z0 a multiple of 8 (uniform), zr < 8 or
// z0 a multiple of 128 and zr < 128: plain positions, or blocks of 2^sh >= 8
// planes (2^sh <= 128 for the second form), split the byte offset into a
// uniform part of z0 and a per-lane part of zr.
__device__ __forceinline__ unsigned mr_zoff_u(const Addr& d, int z0)
{
    return (d.pos_sh >= 30) ? (unsigned)z0 * (unsigned)d.pos_lo * 8u
                            : ((unsigned)(z0 >> d.pos_sh) * (unsigned)d.pos_hi +
                               (unsigned)(z0 & ((1 << d.pos_sh) - 1)) * (unsigned)d.pos_lo) * 8u;
}
__device__ __forceinline__ unsigned mr_zoff_l(const Addr& d, int zr)
{
    return (d.pos_sh >= 30) ? (unsigned)zr * (unsigned)d.pos_lo * 8u
                            : ((unsigned)(zr >> d.pos_sh) * (unsigned)d.pos_hi +
                               (unsigned)(zr & ((1 << d.pos_sh) - 1)) * (unsigned)d.pos_lo) * 8u;
}

// An opaque copy of the lane offset v that depends on the eight complex
// values z0, z1: loads addressed through it cannot be scheduled before those
// values exist, so loads of the next tile's multiplier inputs into the
// registers of this tile's (once z0, z1, the last values formed from them,
// exist) need no register copies.
__device__ __forceinline__ int mr_after(int v, const float2 (&z0)[4], const float2 (&z1)[4])
{
    asm volatile(""
                 : "+v"(v)
                 : "v"(z0[0].x), "v"(z0[0].y), "v"(z0[1].x), "v"(z0[1].y), "v"(z0[2].x), "v"(z0[2].y), "v"(z0[3].x),
                   "v"(z0[3].y), "v"(z1[0].x), "v"(z1[0].y), "v"(z1[1].x), "v"(z1[1].y), "v"(z1[2].x), "v"(z1[2].y),
                   "v"(z1[3].x), "v"(z1[3].y));
    return v;
}

// Multiplier inputs of lane (j, g)'s 16 spectrum points kz = k1 + 32 k2,
// k1 = 2 j + tt, k2 = 4 g + e + 16 h, for pencil column col of tile T, kept as
// loaded: gr[h][e] holds the pair tt = 0, 1 (h = 0) or tt = 1, 0 (h = 1).  The
// octant is [ky][kx][kz] (g_kz = 1): kz < 512 for h = 0, and the octant index
// of h = 1 is Mz - kz = 512 - (kz of h = 0), so each pair is one 8-byte load
// at base + immediate.  The operators that read the octant keep the raw Gk
// values (the same for slabs ky and My - ky, one octant row) and apply their
// k-dependent factor where the multiplier is used; the others compute theirs
// here.
typedef float mr_f2a __attribute__((ext_vector_type(2), aligned(4)));

template <int OPK> constexpr bool mr_loads_g() { return OPK == OP_PLAIN || OPK == OP_GRAD || OPK == OP_LAPL; }

template <int OPK>
__device__ __forceinline__ void mr_kload(mr_f2a (&gr)[2][4], const PencilArgs<float>& a, const MrTile& T, int col, int j,
                                         int g, int kbl)
{
    const TileK<float> tk = tile_k(a, col, T.slab, col < a.ncol);
    const int kb = 2 * j + 128 * g;
    if constexpr (mr_loads_g<OPK>()) {
        // kbl: kb, possibly an opaque copy that orders the loads
        const char* row = reinterpret_cast<const char*>(tk.grow ? tk.grow : a.G);
        const char* r0 = row + kbl * 4;
        const char* r1 = row + (512 - kbl) * 4 - 4 - 3 * 128;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            gr[0][e] = *reinterpret_cast<const mr_f2a*>(r0 + 128 * e);
            gr[1][e] = *reinterpret_cast<const mr_f2a*>(r1 + 128 * (3 - e));
        }
    } else {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
            for (int e = 0; e < 4; ++e)
#pragma unroll
                for (int tt = 0; tt < 2; ++tt) {
                    const int kz = kb + tt + 32 * e + 512 * h;
                    gr[h][e][h ? 1 - tt : tt] = (OPK == OP_NONE) ? 1.0f : mf_rhsf(a, tk, kz);
                }
    }
}

// multiplier of point (tt, h, e) from gr (kc: the tile's coefficients).  The
// point's wavenumber index is kz = kb + tt + 32 e + 512 h, taken as kz - Mz
// from Mz / 2 on (Mz = 1024: exactly the points of h = 1), so with kbf = kb
// as a float it is kbf plus a compile-time integer: (float)(kz or kz - Mz)
// in one addition, exact, as the gradient and Laplacian forms of the other
// kernels compute it.  The gradient has no derivative at the Nyquist point
// kz = Mz / 2 (h = 1, tt = e = 0 on the lane with kb = 0).
template <int OPK>
__device__ __forceinline__ float mr_kmul(const MfKCoef& kc, const mr_f2a (&gr)[2][4], int tt, int h, int e, int kb,
                                         float kbf)
{
    const float v = gr[h][e][h ? 1 - tt : tt];
    if constexpr (OPK == OP_GRAD || OPK == OP_LAPL) {
        const float kzw = (kbf + (float)(tt + 32 * e - 512 * h)) * kc.kst;
        if constexpr (OPK == OP_GRAD) {
            const float kzh = (h == 1 && tt == 0 && e == 0 && kb == 0) ? 0.0f : kzw;
            return (kc.c0 + kc.c1 * kzh) * v;
        } else {
            return (kc.c0 - kzw * kzw) * v;
        }
    }
    return v;
}

// Per-lane and per-wavefront offsets of the slices, fixed for the launch:
// slice k of wavefront w is block n = w + 16 k of a staging buffer (rows 8n ..
// 8n + 7).  For the store, lane l reads chunk l & 7 of row l >> 3 from its
// slot; the fetch lands lane l in slot l, so the lane loads the row and chunk
// that slot holds.
struct MrSliceOffs {
    unsigned lds;       // LDS offset of lane l's chunk in block 0 of its wavefront's slices (+ n 1024)
    unsigned ilane;     // fetch: byte offset of the chunk the lane loads (the staging permutation)
    unsigned olane;     // store: byte offset of the chunk the lane stores
    unsigned iu[4], ou[4];   // uniform byte offsets of rows 8n, fetch and store addressing
};

__device__ __forceinline__ MrSliceOffs mr_slice_offs(const PencilArgs<float>& a, int w, int l)
{
    MrSliceOffs o;
    // blocks n = w + 16 k: n & 1 = w & 1, (n >> 1) & 1 = (w >> 1) & 1
    const int par = w & 1, rot = (w >> 1) & 1;
    const int r = l >> 3, c = l & 7;                        // store: row, chunk
    const int fr = ((l >> 3) - rot) & 7;                    // fetch: the row slot l holds
    const int fc = ((l & 7) ^ fr ^ par) & 7;                // and its chunk
    o.lds = mr_stg_pos(r, c, rot, par);
    o.ilane = mr_zoff_l(a.ia, fr) + 16u * (unsigned)fc;
    o.olane = mr_zoff_l(a.oa, r) + 16u * (unsigned)c;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        o.iu[k] = mr_zoff_u(a.ia, 8 * (w + 16 * k));
        o.ou[k] = mr_zoff_u(a.oa, 8 * (w + 16 * k));
    }
    return o;
}

// Slice k: store the previous tile's output rows of the wavefront's block
// (to obase: that tile's column 0; nv valid columns from the lane's chunk),
// then fetch the next tile's rows (from ibase) into the same block; the
// wavefront alone reads and overwrites its blocks.  Four slices spread over a
// tile's transform keep the stores and the fetch from reaching the memory
// system as one burst.  Each lane handles the chunk whose LDS slot it reads,
// so the fetch and the store use the same (permuted) chunk index.
__device__ __forceinline__ void mr_slice(unsigned char* __restrict__ buf, const MrSliceOffs& so, char* obase, int nv,
                                         bool hasprev, const char* ibase, bool more, int w, int k)
{
    const int n = w + 16 * k;
    if (hasprev) {
        const float4 v = *reinterpret_cast<const float4*>(buf + n * 1024 + so.lds);
        char* o = obase + so.ou[k] + so.olane;
        if (nv >= 2) *reinterpret_cast<float4*>(o) = v;
        else if (nv == 1) *reinterpret_cast<float2*>(o) = float2{v.x, v.y};
    }
    if (more) __builtin_amdgcn_global_load_lds(ibase + so.iu[k] + so.ilane, (mr_lds_ptr)(buf + n * 1024), 16, 0, 0);
}

// One tile T for wavefront w: the transform of column w of cur with the
// multiplier inputs gm (already in registers), written back in place, and
// spread over its first half the four slices that store the previous tile
// from nxt and fetch the next one into it; after the last use of gm, the next
// tile Tn's multiplier inputs are loaded into it (load_g; the second tile of
// a slab pair keeps the first's).  Nothing here waits for vector memory: the loads are
// waited for at the end of the tile, so the multiplier loads never wait for a
// store (loads and stores share one counter).  cur, nxt, the twiddle matrix
// and the tables are disjoint LDS regions; as __restrict__ parameters of one
// function the compiler can tell that this tile's LDS reads of cur do not
// depend on the fetch into nxt.
template <int OPK>
__device__ __forceinline__ void mr_step(const PencilArgs<float>& a, unsigned char* __restrict__ cur,
                                        unsigned char* __restrict__ nxt, const float2* __restrict__ twm,
                                        const mf_h8* __restrict__ ft, const MrTile& T, const MrSliceOffs& so,
                                        char* obase, int nv, bool hasprev, const char* ibase, bool more,
                                        const MrTile& Tn, bool load_g, mr_f2a (&gm)[2][4], int w, int l)
{
    constexpr bool mimag = mf_imag_multiplier<OPK>();
    const int g = l >> 4, j = l & 15;
    const int p = w, kb = 2 * j + 128 * g;
    mr_slice(nxt, so, obase, nv, hasprev, ibase, more, w, 0);
    // the 16-point DFT table (8 of the 16 products) held for the tile where
    // the registers allow (the filtered rhs's multipliers take them)
    constexpr bool hold16 = OPK != OP_RHSF;
    MfT16 F16{};
    if constexpr (hold16) F16 = mr_table(ft, 0, l);
    // staging offsets of rows z = 2 j + q + 32 (4 g + e) (q = 0, 1; 128 g + 32 e
    // is a whole number of blocks): block (j >> 2) + 16 g + 4 e, row
    // 2 (j & 3) + q, n & 1 = (j >> 2) & 1, (n >> 1) & 1 = (j >> 3) & 1
    unsigned sb[2];
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int r = 2 * (j & 3) + q;
        sb[q] = (unsigned)(((j >> 2) + 16 * g) * 1024) + mr_stg_pos(r, p >> 1, (j >> 3) & 1, (j >> 2) & 1) + (p & 1) * 8;
    }
    // ---- input x[b + 32 m], b = 2 j + tb, m = 4 g + e (m >= 16 is padding) ----
    float2 x[2][4];
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
        for (int e = 0; e < 4; ++e) x[tb][e] = *reinterpret_cast<const float2*>(cur + sb[tb] + e * 4096);
    float s0, inv0;
    mr_scale(mr_wave_max(mr_absmax8(x[0], x[1])), s0, inv0);

    // ---- stage 1 (data x table): Y[b][k1] = sum_m W32^(m k1) x[b + 32 m] ----
    // yr/yi[tb][tk]: rows b = 2 (4 g + e) + tb, column k1 = 2 j + tk
    mf_f4 yr[2][2], yi[2][2];
    {
        const MfT16 Fe = hold16 ? F16 : mr_table(ft, 0, l), Fo = mr_table(ft, 1, l);
#pragma unroll
        for (int tb = 0; tb < 2; ++tb) {
            const MfFrag X = mr_split4s<false>(x[tb], s0, a.fzero);
            mr_mfma_a<true>(yr[tb][0], yi[tb][0], X, Fe);
            mr_mfma_a<true>(yr[tb][1], yi[tb][1], X, Fo);
        }
    }
    mr_slice(nxt, so, obase, nv, hasprev, ibase, more, w, 1);
    // twiddle W1024^(b k1): matrix row k1 = 2 j + tk, column b = 8 g + 2 e + tb
    // twiddles lane-major: slot 8 t + 2 e + u of lane (j, g) holds W1024^((2 j + t)(8 g + 2 e + u))
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
    // ur/ui[tk][tn]: rows k1 = 2 (4 g + e) + tk, column n2 = 2 j + tn; tile tk
    // carries the factor s0 s2[tk]
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
        // X[kz] (kz = k1 + 32 k2, k2 = 4 g + e and + 16) and the k-space operator
        // the k-space operator and inverse stage 1's butterflies in one step:
        // with X = E +- O at rows k2 and k2 + 16 and multipliers m0, m1 there,
        // A = X0 m0 + X1 m1 = E (m0 + m1) + O (m0 - m1) and
        // B = X0 m0 - X1 m1 = E (m0 - m1) + O (m0 + m1) (times i for an
        // imaginary multiplier)
        float2 za[4], zb[4];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const float m0 = mr_kmul<OPK>(kc, gm, tk, 0, e, kb, kbf), m1 = mr_kmul<OPK>(kc, gm, tk, 1, e, kb, kbf);
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
        if (tk == 0) mr_slice(nxt, so, obase, nv, hasprev, ibase, more, w, 2);
        // the next tile's multipliers, once this tile's are used
        if (tk == 1 && load_g) mr_kload<OPK>(gm, a, Tn, Tn.col0 + p, j, g, mr_after(kb, za, zb));
        // ---- inverse stage 1 (conj Z x table), row k1 = 2 j + tk ----
        mr_scale(mr_wave_max(mr_absmax8(za, zb)), s2[tk], inv2[tk]);
        {
            const MfFrag XA = mr_split4s<true>(za, s2[tk], a.fzero), XB = mr_split4s<true>(zb, s2[tk], a.fzero);
            mr_mfma_a<true>(ur[tk][0], ui[tk][0], XA, hold16 ? F16 : mr_table(ft, 0, l));
            mr_mfma_a<true>(ur[tk][1], ui[tk][1], XB, mr_table(ft, 1, l));
        }
        if (tk == 0) mr_slice(nxt, so, obase, nv, hasprev, ibase, more, w, 3);
        // T = W1024^(n2 k1) V (V = conj of the inverse transform's partial)
#pragma unroll
        for (int tn = 0; tn < 2; ++tn)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float2 v = cmul(float2{ur[tk][tn][e], ui[tk][tn][e]}, twl[(8 * tn + 2 * e + tk) * 64]);
                ur[tk][tn][e] = v.x;
                ui[tk][tn][e] = v.y;
            }
    }
    // Inverse stage 2 input scale from the bound, not a maximum: the split
    // inverse stage-1 inputs are below 2^10, a 16-point DFT and the unit
    // twiddles keep |T| below 16 sqrt 2 2^10 < 2^15, so tile tk, which carries
    // s0 s2[tk], is split with c[tk] = smin / (2 s2[tk]) (smin the smaller
    // s2): both tiles then carry s0 smin / 2 and stay below 2^14.5
    const float smin = __builtin_fminf(s2[0], s2[1]);
    const float sa = 0.5f * smin * inv2[0], sb2 = 0.5f * smin * inv2[1];
    const float fo = 2.0f * inv0 * __builtin_fmaxf(inv2[0], inv2[1]);
    // ---- inverse stage 2 (table x data): x[32 n1 + n2] = conj(sum_k1 W32^(n1 k1) T[k1][n2]), n1 < 16 ----
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
        mr_mfma_b<true>(dr, di, hold16 ? F16 : mr_table(ft, 0, l), B0);
        mr_mfma_b<false>(dr, di, mr_table(ft, 2, l), B1);
        // rows z = 32 (4 g + e) + 2 j + tn
#pragma unroll
        for (int e = 0; e < 4; ++e)
            *reinterpret_cast<float2*>(cur + sb[tn] + e * 4096) = float2{dr[e] * fo, -di[e] * fo};
    }
}

// Persistent kernel: workgroup g handles tiles g, g + G, g + 2G, ... (G = the
// grid); the first tile is fetched before the loop, each later one during
// the tile before it.  Every wavefront of a workgroup runs the same
// iterations, so the barriers are uniform.
template <int OPK>
__global__ void __launch_bounds__(1024, 4) k_zmfma1024r(PencilArgs<float> a)
{
    using M = MfmaZR;
    constexpr int L = M::L;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    unsigned char* stg0 = smem_raw;
    float2* twm = reinterpret_cast<float2*>(smem_raw + 2 * M::STG);
    mf_h8* ftw = reinterpret_cast<mf_h8*>(smem_raw + 2 * M::STG + M::TWB);

    const int tid = threadIdx.x;
    const int w = __builtin_amdgcn_readfirstlane(tid >> 6), l0 = tid & 63;
    const int nunit = mr_nunit(a);
    const int G = gridDim.x;

    // twiddles W1024^(r s) lane-major (1024 threads, one entry each): entry
    // (slot 8 t + 2 e + u, lane 16 g + j) is r = 2 j + t, s = 8 g + 2 e + u, so
    // every twiddle read of a wavefront is one linear 512-byte access
    {
        const int slot = tid >> 6, j = tid & 15, g = (tid >> 4) & 3;
        const int r = 2 * j + (slot >> 3), s = 8 * g + (slot & 7);
        twm[tid] = a.plan.tw[(r * s) & (L - 1)];
    }
    if (tid < 3 * 4 * 64) {
        // tables E16, P, Q: global [table][lane][rh, rl, ih, il] -> LDS [table][rh, ih, rl, il][lane]
        const int ln = tid & 63, f = (tid >> 6) & 3, tb = tid >> 8;
        const int fo = (f == 0) ? 0 : (f == 1) ? 2 : (f == 2) ? 1 : 3;
        ftw[(tb * 4 + fo) * 64 + ln] = reinterpret_cast<const mf_h8*>(a.mfrag16)[(tb * 64 + ln) * 4 + f];
    }
    // iterations: tiles (u, 0), (u, 1) of units u = g, g + G, ...
    const MrSliceOffs so = mr_slice_offs(a, w, l0);
    int u = blockIdx.x, m = 0;
    MrUnit U{};
    MrTile T{};
    const bool any = u < nunit && (U = mr_unit(a, u, nunit), mr_tile(a, U, 0, T));
    mr_f2a gm[2][4];   // multiplier inputs of the tile (kept for the second tile of a pair)
    if (any) {
        const char* ib = mr_tile_base(a.in, a.ia, T);
#pragma unroll
        for (int k = 0; k < 4; ++k)
            __builtin_amdgcn_global_load_lds(ib + so.iu[k] + so.ilane, (mr_lds_ptr)(stg0 + (w + 16 * k) * 1024), 16, 0, 0);
        mr_kload<OPK>(gm, a, T, T.col0 + w, l0 & 15, l0 >> 4, 2 * (l0 & 15) + 128 * (l0 >> 4));
    }
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();

    MrTile Tp = T;
    int it = 0;
    for (; any; ++it) {
        // lane indices recomputed per tile (an opaque copy): lane-dependent
        // offsets are not held in registers across the loop
        int l = l0;
        asm volatile("" : "+v"(l));
        unsigned char* cur = stg0 + (it & 1) * M::STG;
        unsigned char* nxt = stg0 + ((it & 1) ^ 1) * M::STG;
        // the next tile: the pair's second, else the next unit's first
        MrTile Tn{};
        MrUnit Un = U;
        int un = u, mn = 1;
        if (m == 1 || !mr_tile(a, U, 1, Tn)) {
            un = u + G;
            mn = 0;
        }
        const bool more = un < nunit && (mn == 1 || (Un = mr_unit(a, un, nunit), mr_tile(a, Un, 0, Tn)));
        // the second tile of a slab pair keeps the first one's multipliers
        const bool load_g = more && !(mn == 1 && Tn.kyo == T.kyo);
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, Tp));
        const char* ibase = mr_tile_base(a.in, a.ia, Tn);
        const int nv = a.ncol - Tp.col0 - 2 * (l & 7);   // valid columns of this lane's chunk of Tp
        mr_step<OPK>(a, cur, nxt, twm, ftw, T, so, obase, nv, it > 0, ibase, more, Tn, load_g, gm, w, l);
        // this wavefront's fetch of the next tile and the previous tile's
        // stores are done; every wavefront's output is in cur
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        Tp = T;
        T = Tn;
        U = Un;
        u = un;
        m = mn;
        if (!more) break;
    }
    // the last tile's output (in the buffer of the last iteration)
    if (any) {
        unsigned char* last = stg0 + (it & 1) * M::STG;
        char* obase = const_cast<char*>(mr_tile_base(a.out, a.oa, Tp));
        const int nv = a.ncol - Tp.col0 - 2 * (l0 & 7);
#pragma unroll
        for (int k = 0; k < 4; ++k) mr_slice(last, so, obase, nv, true, nullptr, false, w, k);
    }
}

}  // namespace psamd
