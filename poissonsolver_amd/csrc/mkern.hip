// Mixed-precision (matrix-core) pass kernels: instantiation and launchers.
#include "ps_mfma.hpp"
#include "ps_mfma_launch.hpp"
#include "ps_mfma_reg.hpp"
#include "ps_mfma_reg2048.hpp"
#include "ps_mfma_reg512.hpp"
#include "ps_mfma_reg3.hpp"

#include <algorithm>
#include <cmath>

namespace psamd {

unsigned short f16_bits(double v)
{
    const unsigned short sign = std::signbit(v) ? 0x8000 : 0;
    const double a = std::fabs(v);
    if (a == 0.0) return sign;
    int e = 0;
    (void)std::frexp(a, &e);   // a = m 2^e, m in [0.5, 1)
    int E = e - 1;             // a = 1.f 2^E
    if (E >= -14) {
        double r = std::nearbyint(std::ldexp(a, 10 - E));   // in [1024, 2048]
        if (r >= 2048.0) {
            r = 1024.0;
            E += 1;
        }
        return (unsigned short)(sign | ((E + 15) << 10) | ((int)r - 1024));
    }
    // subnormal: k 2^-24 (k = 1024 is the smallest normal, same bits)
    return (unsigned short)(sign | (int)std::nearbyint(std::ldexp(a, 24)));
}

double f16_value(unsigned short h)
{
    const int E = (h >> 10) & 31, f = h & 1023;
    const double m = (E == 0) ? std::ldexp((double)f, -24) : std::ldexp((double)(1024 + f), E - 25);
    return (h & 0x8000) ? -m : m;
}

void zmfma_fragments(unsigned short* out)
{
    for (int lane = 0; lane < 64; ++lane) {
        const int i = lane & 15, qa = lane >> 4;
        for (int c = 0; c < 2; ++c) {
            for (int e8 = 0; e8 < 8; ++e8) {
                const int m = 4 * qa + (e8 >> 1) + 16 * c;
                const int jj = (i * m) & 31;
                const double cs = Cos32::c[jj];
                const double sn = Cos32::c[(jj + 24) & 31];   // sin(2 pi jj / 32)
                const double vr = (e8 & 1) ? sn : cs;
                const double vi = (e8 & 1) ? cs : -sn;
                const unsigned short rh = f16_bits(vr), ih = f16_bits(vi);
                unsigned short* o = out + (lane * 8 + 4 * c) * 8 + e8;
                o[0] = rh;
                o[8] = f16_bits(vr - f16_value(rh));
                o[16] = ih;
                o[24] = f16_bits(vi - f16_value(ih));
            }
        }
    }
}

// 16 x 16 tables W32^f(i, k) in the layout of zmfma_fragments (row i = lane &
// 15, K points k = 4 (lane >> 4) + e8 / 2): kind 0 the 16-point DFT
// W16^(ik) (f = 2ik), kind 1 P (f = k (2i + 1)), kind 2 Q (f = i (2k + 1))
void zmfma_fragments16k(unsigned short* out, int kind)
{
    for (int lane = 0; lane < 64; ++lane) {
        const int i = lane & 15, qa = lane >> 4;
        for (int e8 = 0; e8 < 8; ++e8) {
            const int m = 4 * qa + (e8 >> 1);
            const int f = (kind == 0) ? 2 * i * m : (kind == 1) ? m * (2 * i + 1) : i * (2 * m + 1);
            const int jj = f & 31;
            const double cs = Cos32::c[jj];
            const double sn = Cos32::c[(jj + 24) & 31];
            const double vr = (e8 & 1) ? sn : cs;
            const double vi = (e8 & 1) ? cs : -sn;
            const unsigned short rh = f16_bits(vr), ih = f16_bits(vi);
            unsigned short* o = out + lane * 32 + e8;
            o[0] = rh;
            o[8] = f16_bits(vr - f16_value(rh));
            o[16] = ih;
            o[24] = f16_bits(vi - f16_value(ih));
        }
    }
}

void zmfma_fragments16(unsigned short* out)
{
    zmfma_fragments16k(out, 0);
}

static bool zmfma_reg_ok(const PencilArgs<float>& a);
static int device_cus();

template <int OPK> static hipError_t launch_zmfma512_op(PencilArgs<float> a, int variant, hipStream_t s)
{
    if (variant == 1 && zmfma_reg_ok(a)) {
        // register-pencil kernel: 16-column tiles of 256 rows, one workgroup
        // of 16 wavefronts per CU at a time, two per CU over the launch (a
        // tile is short: fewer workgroups start fewer table fills; 256^3
        // PPU z pass 0.070 -> 0.064 ms against four per CU)
        const size_t shm = MfmaZR5::lds_bytes();
        hipError_t e = allow_lds(k_zmfma512r<OPK>, shm);
        if (e != hipSuccess) return e;
        a.ncoltile = (a.ncol + MfmaZR5::P - 1) / MfmaZR5::P;
        if ((long)a.ncoltile * a.nslab * a.ncomp >= (1L << 31)) return hipErrorInvalidValue;
        const long nunit = (long)a.ncoltile * ((a.nslab + 1) / 2) * a.ncomp;
        const long ncu = std::max(device_cus(), 8);
        const long nwg = std::min<long>(nunit, 2 * ncu);
        hipLaunchKernelGGL((k_zmfma512r<OPK>), dim3((unsigned)nwg), dim3(64 * MfmaZR5::NW), shm, s, a);
        return hipGetLastError();
    }
    using M = MfmaZ512;
    const size_t shm = M::lds_bytes();
    hipError_t e = allow_lds(k_zmfma512<OPK>, shm);
    if (e != hipSuccess) return e;
    a.ncoltile = (a.ncol + M::P - 1) / M::P;
    const long nwork = (long)a.ncoltile * a.nslab * a.ncomp;
    hipLaunchKernelGGL((k_zmfma512<OPK>), dim3((unsigned)nwork), dim3(64 * M::NW), shm, s, a);
    return hipGetLastError();
}

// whether the register-pencil kernel can run these arguments: kx-contiguous
// rows (16-column tiles of 128 bytes), the octant in the [ky][kx][kz] layout
// and the E16/P/Q tables after the DFT-16 one
static bool zmfma_reg_ok(const PencilArgs<float>& a)
{
    auto zok = [](const Addr& d) { return d.pos_sh >= 30 || (d.pos_sh >= 3 && d.pos_sh <= 7); };
    return a.ia.col == 1 && a.oa.col == 1 && a.g_kz == 1 && a.mfrag16 && a.ncol > 0 && zok(a.ia) && zok(a.oa);
}

static int device_cus()
{
    int dev = 0, ncu = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return ncu;
}

template <int OPK> static hipError_t launch_zmfma_op(PencilArgs<float> a, int variant, hipStream_t s)
{
    hipError_t e;
    if (variant == 1 && zmfma_reg_ok(a)) {
        // persistent register-pencil kernel: 16-column tiles, one workgroup of
        // 16 wavefronts per CU; four workgroups per CU over the launch so that
        // CUs finishing early take the remaining ones
        const size_t shm = MfmaZR::lds_bytes();
        if ((e = allow_lds(k_zmfma1024r<OPK>, shm)) != hipSuccess) return e;
        a.ncoltile = (a.ncol + MfmaZR::P - 1) / MfmaZR::P;
        const long nunit = (long)a.ncoltile * ((a.nslab + 1) / 2) * a.ncomp;   // slab pairs x column tiles
        if ((long)a.ncoltile * a.nslab * a.ncomp >= (1L << 31)) return hipErrorInvalidValue;
        const long ncu = std::max(device_cus(), 8);
        const long nwg = std::min<long>(nunit, 4 * ncu);
        hipLaunchKernelGGL((k_zmfma1024r<OPK>), dim3((unsigned)nwg), dim3(64 * MfmaZR::NW), shm, s, a);
        return hipGetLastError();
    }
    const size_t shm = MfmaZ::lds_bytes();
    a.ncoltile = (a.ncol + MfmaZ::P - 1) / MfmaZ::P;
    const long nwork = (long)a.ncoltile * a.nslab * a.ncomp;
    if ((e = allow_lds(k_zmfma1024<OPK>, shm)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_zmfma1024<OPK>), dim3((unsigned)nwork), dim3(64 * MfmaZ::NW), shm, s, a);
    return hipGetLastError();
}

template <int OPK> static hipError_t launch_zmfma3_op(PencilArgs<float> a, hipStream_t s)
{
    const size_t shm = MfmaZ3::lds_bytes();
    hipError_t e = allow_lds(k_zmfma1024v<OPK>, shm);
    if (e != hipSuccess) return e;
    a.ncoltile = (a.ncol + MfmaZ3::P - 1) / MfmaZ3::P;
    const long nwork = (long)a.ncoltile * a.nslab;
    hipLaunchKernelGGL((k_zmfma1024v<OPK>), dim3((unsigned)nwork), dim3(64 * MfmaZ3::NW), shm, s, a);
    return hipGetLastError();
}

bool zmfma3_supported(int L, int half, int op)
{
    return L == MfmaZ3::L && half && (op == OP_DIV || op == OP_CURL);
}

// register-pencil form: workgroups of 8 wavefronts (one pencil column and
// its three components each, 2 per SIMD), one per CU at a time, four per CU
// over the launch
template <int OPK> static hipError_t launch_zmfma3r_op(PencilArgs<float> a, hipStream_t s)
{
    const size_t shm = MfmaZR3::lds_bytes();
    hipError_t e = allow_lds(k_zmfma1024c<OPK>, shm);
    if (e != hipSuccess) return e;
    a.ncomp = 1;
    a.ncoltile = (a.ncol + MfmaZR3::P - 1) / MfmaZR3::P;
    if ((long)a.ncoltile * a.nslab >= (1L << 31)) return hipErrorInvalidValue;
    const long nunit = (long)a.ncoltile * ((a.nslab + 1) / 2);
    const long ncu = std::max(device_cus(), 8);
    const long nwg = std::min<long>(nunit, 4 * ncu);
    hipLaunchKernelGGL((k_zmfma1024c<OPK>), dim3((unsigned)nwg), dim3(64 * MfmaZR3::NW), shm, s, a);
    return hipGetLastError();
}

hipError_t launch_zmfma3(const PencilArgs<float>& a, int variant, hipStream_t s)
{
    if (a.plan.L != MfmaZ3::L || a.n_in != MfmaZ3::L / 2 || a.n_out != MfmaZ3::L / 2) return hipErrorInvalidValue;
    if (variant == 1 && zmfma_reg_ok(a)) {
        if (a.op == OP_DIV) return launch_zmfma3r_op<OP_DIV>(a, s);
        if (a.op == OP_CURL) return launch_zmfma3r_op<OP_CURL>(a, s);
        return hipErrorInvalidValue;
    }
    if (a.op == OP_DIV) return launch_zmfma3_op<OP_DIV>(a, s);
    if (a.op == OP_CURL) return launch_zmfma3_op<OP_CURL>(a, s);
    return hipErrorInvalidValue;
}

bool zmfma_supported(int L, int half, int op)
{
    return (L == 512 || L == MfmaZ::L || L == 2048) && half &&
           (op == OP_NONE || op == OP_PLAIN || op == OP_GRAD || op == OP_RHSF || op == OP_LAPL);
}

template <int OPK, int P> static hipError_t launch_zmfma2048_p(PencilArgs<float> a, hipStream_t s)
{
    using M = MfmaZ2048<P>;
    const size_t shm = M::lds_bytes();
    hipError_t e = allow_lds(k_zmfma2048<OPK, P>, shm);
    if (e != hipSuccess) return e;
    a.ncoltile = (a.ncol + M::P - 1) / M::P;
    const long nwork = (long)a.ncoltile * a.nslab * a.ncomp;
    hipLaunchKernelGGL((k_zmfma2048<OPK, P>), dim3((unsigned)nwork), dim3(64 * M::NW), shm, s, a);
    return hipGetLastError();
}

// the fan-out of 2048: workgroups of 8 wavefronts (one pencil each, both
// residues; 2 per SIMD), one per CU at a time, four per CU over the launch
template <int MASK> static hipError_t launch_zmfma2048_fanp_m(PencilArgs<float> a, hipStream_t s)
{
    const size_t shm = MfmaZR2::lds_bytes();
    hipError_t e = allow_lds(k_zmfma2048p<MASK>, shm);
    if (e != hipSuccess) return e;
    a.kmap = KMAP_Z;
    a.ncoltile = (a.ncol + MfmaZR2::P - 1) / MfmaZR2::P;
    if ((long)a.ncoltile * a.nslab >= (1L << 31)) return hipErrorInvalidValue;
    const long nunit = (long)a.ncoltile * ((a.nslab + 1) / 2);
    const long ncu = std::max(device_cus(), 8);
    const long nwg = std::min<long>(nunit, 4 * ncu);
    hipLaunchKernelGGL((k_zmfma2048p<MASK>), dim3((unsigned)nwg), dim3(64 * MfmaZR2P::NW), shm, s, a);
    return hipGetLastError();
}

// tuning mfma 1: the register-pencil kernel when the arguments allow it (as
// for transforms of 1024), else 4-pencil workgroups that reload the pencil for
// residue 1; 2: 8-pencil workgroups that hold it; 3: the 4-pencil ones
template <int OPK> static hipError_t launch_zmfma2048_op(PencilArgs<float> a, int variant, hipStream_t s)
{
    if (variant == 1 && zmfma_reg_ok(a)) {
        // one workgroup of 16 wavefronts (two per pencil) per CU at a time,
        // four per CU over the launch, as for transforms of 1024
        const size_t shm = MfmaZR2::lds_bytes();
        hipError_t e = allow_lds(k_zmfma2048r<OPK>, shm);
        if (e != hipSuccess) return e;
        a.ncoltile = (a.ncol + MfmaZR2::P - 1) / MfmaZR2::P;
        if ((long)a.ncoltile * a.nslab * a.ncomp >= (1L << 31)) return hipErrorInvalidValue;
        const long nunit = (long)a.ncoltile * ((a.nslab + 1) / 2) * a.ncomp;
        const long ncu = std::max(device_cus(), 8);
        const long nwg = std::min<long>(nunit, 4 * ncu);
        hipLaunchKernelGGL((k_zmfma2048r<OPK>), dim3((unsigned)nwg), dim3(64 * MfmaZR2::NW), shm, s, a);
        return hipGetLastError();
    }
    return variant == 2 ? launch_zmfma2048_p<OPK, 8>(a, s) : launch_zmfma2048_p<OPK, 4>(a, s);
}

template <int MASK> static hipError_t launch_zmfma512_fan_m(PencilArgs<float> a, hipStream_t s)
{
    const size_t shm = MfmaZR5::lds_bytes();
    hipError_t e = allow_lds(k_zmfma512f<MASK>, shm);
    if (e != hipSuccess) return e;
    a.kmap = KMAP_Z;
    a.ncoltile = (a.ncol + MfmaZR5::P - 1) / MfmaZR5::P;
    if ((long)a.ncoltile * a.nslab >= (1L << 31)) return hipErrorInvalidValue;
    const long nunit = (long)a.ncoltile * ((a.nslab + 1) / 2);
    const long ncu = std::max(device_cus(), 8);
    const long nwg = std::min<long>(nunit, 2 * ncu);
    hipLaunchKernelGGL((k_zmfma512f<MASK>), dim3((unsigned)nwg), dim3(64 * MfmaZR5::NW), shm, s, a);
    return hipGetLastError();
}

bool zmfma_fan_ok(const PencilArgs<float>& a, int mask, int variant)
{
    const bool len = (a.plan.L == 512 || a.plan.L == 2048) && a.n_in == a.plan.L / 2 && a.n_out == a.plan.L / 2;
    return variant == 1 && len && a.ncomp == 1 && a.mfrag16 && zmfma_reg_ok(a) &&
           (mask == 3 || mask == 5 || mask == 6 || mask == 7);
}

hipError_t launch_zmfma_fan(const PencilArgs<float>& a, int mask, hipStream_t s)
{
    if (!zmfma_fan_ok(a, mask, 1)) return hipErrorInvalidValue;
    if (a.plan.L == 2048) {
        switch (mask) {
        case 3: return launch_zmfma2048_fanp_m<3>(a, s);
        case 5: return launch_zmfma2048_fanp_m<5>(a, s);
        case 6: return launch_zmfma2048_fanp_m<6>(a, s);
        default: return launch_zmfma2048_fanp_m<7>(a, s);
        }
    }
    switch (mask) {
    case 3: return launch_zmfma512_fan_m<3>(a, s);
    case 5: return launch_zmfma512_fan_m<5>(a, s);
    case 6: return launch_zmfma512_fan_m<6>(a, s);
    default: return launch_zmfma512_fan_m<7>(a, s);
    }
}

hipError_t launch_zmfma(const PencilArgs<float>& a, int variant, hipStream_t s)
{
    if (a.plan.L == 512) {
        if (a.n_in != 256 || a.n_out != 256 || !a.mfrag16) return hipErrorInvalidValue;
        switch (a.op) {
        case OP_NONE: return launch_zmfma512_op<OP_NONE>(a, variant, s);
        case OP_PLAIN: return launch_zmfma512_op<OP_PLAIN>(a, variant, s);
        case OP_GRAD: return launch_zmfma512_op<OP_GRAD>(a, variant, s);
        case OP_RHSF: return launch_zmfma512_op<OP_RHSF>(a, variant, s);
        case OP_LAPL: return launch_zmfma512_op<OP_LAPL>(a, variant, s);
        default: return hipErrorInvalidValue;
        }
    }
    if (a.plan.L == 2048) {
        // two 1024-point residues per pencil
        if (a.n_in != 1024 || a.n_out != 1024) return hipErrorInvalidValue;
        switch (a.op) {
        case OP_NONE: return launch_zmfma2048_op<OP_NONE>(a, variant, s);
        case OP_PLAIN: return launch_zmfma2048_op<OP_PLAIN>(a, variant, s);
        case OP_GRAD: return launch_zmfma2048_op<OP_GRAD>(a, variant, s);
        case OP_RHSF: return launch_zmfma2048_op<OP_RHSF>(a, variant, s);
        case OP_LAPL: return launch_zmfma2048_op<OP_LAPL>(a, variant, s);
        default: return hipErrorInvalidValue;
        }
    }
    if (a.plan.L != MfmaZ::L || a.n_in != MfmaZ::L / 2 || a.n_out != MfmaZ::L / 2) return hipErrorInvalidValue;
    switch (a.op) {
    case OP_NONE: return launch_zmfma_op<OP_NONE>(a, variant, s);
    case OP_PLAIN: return launch_zmfma_op<OP_PLAIN>(a, variant, s);
    case OP_GRAD: return launch_zmfma_op<OP_GRAD>(a, variant, s);
    case OP_RHSF: return launch_zmfma_op<OP_RHSF>(a, variant, s);
    case OP_LAPL: return launch_zmfma_op<OP_LAPL>(a, variant, s);
    default: return hipErrorInvalidValue;
    }
}

}  // namespace psamd
