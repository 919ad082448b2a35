// Pass kernels of the 3D solve (declarations shared by host and device code).
//
// Data layout in HBM (x fastest, as the reference's ijk = (k ny + j) nx + i):
//   real fields        [c][nz][ny][nx]           T
//   A1 (after x pass)  [c][nz][ny][Hp]           complex T, kx in [0, H)
//   A2 (after y pass)  [c][nz][My][Hp]           complex T
//   Gk octant          [Mz/2+1][My/2+1][H]       real T  = Ghat * zeta / (Mx My Mz)
// The z pass reads nz planes of A2, transforms a zero-padded Mz pencil in
// registers, applies the k-space operator and writes nz planes back, so the
// z-doubled spectrum never touches HBM.
#pragma once
#include "ps_fft.hpp"

namespace psamd {

enum PencilOp : int {
    OP_NONE = 0,      // copy (pure transform pass)
    OP_PLAIN = 1,     // l_c = Gk r_c
    OP_GRAD = 2,      // l_d = i k_d Gk r_d           (component-wise)
    OP_DIV = 3,       // l_0 = Gk i k . r
    OP_CURL = 4,      // l = Gk i k x r
    OP_RHSF = 5,      // r_c * zeta / N               (filtered rhs, no reprojection)
    OP_RHSF_PROJ = 6, // r_d*zeta/N - Gk (kh_d kh.r + kn_d kn.r)  (filtered + reprojected)
    OP_POTGRAD = 7,   // pipeline only: one scalar r, outputs the quantities whose bits are set in
                      //   kmap (bit 0 Gk r, bits 1..3 i k_d Gk r, bit 4 -|k|^2 Gk r); its z pass
                      //   runs OP_PLAIN, OP_GRAD (KMAP_Z) and OP_LAPL
    OP_LAPL = 8,      // l = -|k|^2 Gk r               (|k| with the Nyquist wavenumbers kept)
};

// Wavenumber axis of each vector component for OP_GRAD and OP_DIV.  A 2D
// solve runs as a 3D one of shape (nx, 1, ny): its y direction is the pass
// axis z, and its operators pick their axes by this map.
enum KMap : int {
    KMAP_3D = 0,       // component c <-> axis c
    KMAP_2D_GRAD = 1,  // component 0 <-> k_x, 1 <-> k_z (both read the same scalar)
    KMAP_2D_CURL = 2,  // component 0 <-> +k_z, 1 <-> -k_x (curl of a scalar along z)
    KMAP_2D_DIV = 3,   // component 0 <-> k_x, 1 <-> k_z, component 2 unused
    KMAP_Z = 4,        // component 0 <-> k_z (the z derivative of a scalar)
};

enum LoadMode : int { LOAD_COMPLEX = 0, LOAD_REAL_EVEN = 1 };
enum StoreMode : int { STORE_COMPLEX = 0, STORE_REAL = 1 };

// Element addressing of a strided pass: element (c, slab, pos, col) sits at
//   c*comp + col*col + off(slab) + off(pos),
//   off(i) = (i >> sh)*hi + (i & (2^sh - 1))*lo,
// i.e. a plain stride (sh = 30, lo = stride) or a layout blocked by 2^sh.
struct Addr {
    long comp, col, slab_lo, slab_hi, pos_lo, pos_hi;
    int slab_sh, pos_sh;
};

inline Addr addr_plain(long comp, long slab, long pos, long col)
{
    return Addr{comp, col, slab, 0, pos, 0, 30, 30};
}

// Strided pencil pass: pencils along `pos`, tiles of B columns.
template <typename T> struct PencilArgs {
    FftPlan<T> plan;
    const void* in;
    void* out;
    Addr ia, oa;
    int n_in;       // input positions present (the rest are zero padding)
    int n_out;      // output positions stored
    int ncol, B, ncoltile, nslab, ncomp;
    int load_mode, store_mode;
    int do_fwd, do_inv;
    int op;
    int xcd_remap;   // 1: contiguous work-id ranges per XCD
    int twreg;       // specialised kernels: 1 = stage twiddles held in registers
    int occ4;        // specialised fused kernel: target 4 waves/SIMD (else 2)
    int zwave;       // specialised fused kernel: pencil-per-wavefront variant (k_zfused_w)
    int slab_pair;   // 1: visit slabs as 0, M/2, 1, M-1, 2, M-2, ... (z pass: ky and My-ky share a row of Gk)
    int kmap;        // KMap of the operator's vector components
    // k-space operator (z pass): slab = ky, col = kx, pos = kz
    const T* G;
    long g_ky, g_kx, g_kz;
    int Mx, My, Mz;
    double dk[3];    // 1/(M_d dx_d)
    double zarg;     // 0.5 (2h)^2
    int reg;         // regularisation order (even), 0 = none
    double invN;     // 1/(Mx My Mz)
    int col0;        // kx of the launch's column 0 (0: the launch covers every column)
    // y^-1 pass of one component: multiply the loaded spectrum by i k (its
    // Hermitian part, 0 at the Nyquist index) before the transform; 1: k = kx
    // of the column (Mx, dk[0]), 2: k = ky of the position (My, dk[1]); 0: none
    int ik_in;
    // y^-1 pass of one spectrum into several outputs (SM_INVFAN): bit m set
    // for an output with ik_in mode m (0 none, 1 i kx, 2 i ky), written to
    // component fan_slot[m] of the output addressing
    int ik_fan, fan_slot[3];
    // matrix-core z pass: per-lane fp16 hi/lo fragments of the 32-point DFT
    // matrix (64 lanes x 8 fragments x 8 halves), built once at setup
    const void* mfrag;
    const void* mfrag16;   // the same for the 16-point DFT (64 lanes x 4 fragments x 8 halves)
    // 1 and 0 supplied at run time: the compiler cannot fold them, so the
    // FP16 split forms each conversion as one mixed-precision FMA
    float fone, fzero;
};

// x pass: two real rows (a, b) per complex pencil.
template <typename T> struct XArgs {
    FftPlan<T> plan;
    const void* in;
    void* out;
    long r_comp, r_row;   // real rows
    long c_comp, c_row;   // complex rows
    int nrows;            // rows per component (ny * nz)
    int nvalid;           // nx
    int H;                // L/2 + 1
    int P;                // pairs per workgroup
    int npairs;
    int ncomp;
    // x^-1 pass (specialised kernel) of the potential form: with remap,
    // grid component c reads input component src[c] and writes output
    // component oslot[c], times i kx (the Hermitian part; Mx, dk0 = 1/(Mx
    // dx)) where bit c of ikx is set; the workgroups of grid component
    // dx_also (-1: none) also write i kx times their rows to dx_slot
    int remap, src[5], oslot[5], ikx, dx_also, dx_slot, Mx;
    double dk0;
};

}  // namespace psamd
