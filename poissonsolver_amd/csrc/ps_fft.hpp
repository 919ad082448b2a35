// Register-resident batched pencil FFT engine for gfx950.
//
// A pencil of length L is owned by TP = L/E threads; thread t holds the E
// elements at positions t + TP*u (u = 0..E-1) in registers.  Each Stockham
// stage of radix R (R divides E) does E/R butterflies per thread on register
// data; between stages the pencil is exchanged through LDS (one write at the
// stage's output positions, one read back at positions t + TP*u).  The first
// stage therefore reads straight from the registers filled by a global load,
// and after the last stage every thread again holds positions t + TP*u in
// natural order, ready for a coalesced global store or a pointwise operation.
//
// Conventions follow the reference's Glassman FFT (pencil/fft.cpp,
// pencil/ifft.cpp): forward exp(-2 pi i nk/L), inverse exp(+2 pi i nk/L),
// inverse unnormalised.
#pragma once
#include <hip/hip_runtime.h>

namespace psamd {

template <typename T> struct cplx_of;
template <> struct cplx_of<float> { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename T> using cplx = typename cplx_of<T>::type;

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b)
{
    return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) { return C{a.x * s, a.y * s}; }
// multiply by -i (DIR = -1, forward) or +i (DIR = +1, inverse)
template <int DIR, typename C> __device__ __forceinline__ C crot(C a)
{
    if constexpr (DIR < 0) return C{a.y, -a.x};
    else return C{-a.y, a.x};
}

constexpr int kMaxStage = 14;
constexpr int kMaxThreads = 512;   // workgroup size bound of every pass kernel

// Runtime part of a 1D transform plan; E (elements per thread) is a template
// parameter of every kernel that uses it.
template <typename T> struct FftPlan {
    const cplx<T>* tw;   // tw[j] = exp(-2 pi i j / L), j < L
    int L;               // transform length
    int TP;              // threads per pencil = L / E
    int nstage;
    int radix[kMaxStage];
};

// ---------------------------------------------------------------------------
// small DFTs on registers: y_q = sum_m a_m w^(m q), w = exp(DIR 2 pi i / R)
// ---------------------------------------------------------------------------
template <int P> struct Trig;
template <> struct Trig<3> {
    static constexpr double c[3] = {1.0, -0.5, -0.5};
    static constexpr double s[3] = {0.0, 0.86602540378443864676, -0.86602540378443864676};
};
template <> struct Trig<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494742410, -0.80901699437494742410,
                                    -0.80901699437494742410, 0.30901699437494742410};
    static constexpr double s[5] = {0.0, 0.95105651629515357212, 0.58778525229247312917,
                                    -0.58778525229247312917, -0.95105651629515357212};
};
template <> struct Trig<7> {
    static constexpr double c[7] = {1.0, 0.62348980185873353053, -0.22252093395631440429,
                                    -0.90096886790241912624, -0.90096886790241912624,
                                    -0.22252093395631440429, 0.62348980185873353053};
    static constexpr double s[7] = {0.0, 0.78183148246802980871, 0.97492791218182360702,
                                    0.43388373911755812048, -0.43388373911755812048,
                                    -0.97492791218182360702, -0.78183148246802980871};
};
// cos/sin(2 pi j/16), j = 0..15
struct Trig16 {
    static constexpr double c[16] = {1.0, 0.92387953251128675613, 0.70710678118654752440,
                                     0.38268343236508977173, 0.0, -0.38268343236508977173,
                                     -0.70710678118654752440, -0.92387953251128675613, -1.0,
                                     -0.92387953251128675613, -0.70710678118654752440,
                                     -0.38268343236508977173, 0.0, 0.38268343236508977173,
                                     0.70710678118654752440, 0.92387953251128675613};
    static constexpr double s[16] = {0.0, 0.38268343236508977173, 0.70710678118654752440,
                                     0.92387953251128675613, 1.0, 0.92387953251128675613,
                                     0.70710678118654752440, 0.38268343236508977173, 0.0,
                                     -0.38268343236508977173, -0.70710678118654752440,
                                     -0.92387953251128675613, -1.0, -0.92387953251128675613,
                                     -0.70710678118654752440, -0.38268343236508977173};
};

// multiply by w16^j = exp(DIR 2 pi i j/16)
template <int DIR, int J, typename C> __device__ __forceinline__ C w16(C a)
{
    using T = decltype(a.x);
    constexpr int j = J & 15;
    if constexpr (j == 0) return a;
    else if constexpr (j == 4) return crot<DIR>(a);   // * DIR*i
    else if constexpr (j == 8) return C{-a.x, -a.y};
    else if constexpr (j == 12) return crot<-DIR>(a);  // * -DIR*i
    else {
        const T c = (T)Trig16::c[j];
        const T s = (T)(DIR * Trig16::s[j]);
        return C{a.x * c - a.y * s, a.x * s + a.y * c};
    }
}

template <int DIR, typename C> __device__ __forceinline__ void dft2(C& a0, C& a1)
{
    C t = a0;
    a0 = cadd(t, a1);
    a1 = csub(t, a1);
}

template <int DIR, typename C> __device__ __forceinline__ void dft4(C& a0, C& a1, C& a2, C& a3)
{
    C t0 = cadd(a0, a2), t1 = csub(a0, a2);
    C t2 = cadd(a1, a3), t3 = crot<DIR>(csub(a1, a3));
    a0 = cadd(t0, t2);
    a2 = csub(t0, t2);
    a1 = cadd(t1, t3);
    a3 = csub(t1, t3);
}

// odd prime P by pairing m and P-m
template <int P, int DIR, typename C> __device__ __forceinline__ void dft_odd(C* a)
{
    using T = decltype(a[0].x);
    constexpr int H = (P - 1) / 2;
    C sp[H], sm[H];
#pragma unroll
    for (int m = 1; m <= H; ++m) {
        sp[m - 1] = cadd(a[m], a[P - m]);
        sm[m - 1] = csub(a[m], a[P - m]);
    }
    C y[P];
    C y0 = a[0];
#pragma unroll
    for (int m = 0; m < H; ++m) y0 = cadd(y0, sp[m]);
    y[0] = y0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
        C A = a[0], Bv = C{0, 0};
#pragma unroll
        for (int m = 1; m <= H; ++m) {
            const T c = (T)Trig<P>::c[(m * k) % P];
            const T s = (T)Trig<P>::s[(m * k) % P];
            A = C{A.x + c * sp[m - 1].x, A.y + c * sp[m - 1].y};
            Bv = C{Bv.x + s * sm[m - 1].x, Bv.y + s * sm[m - 1].y};
        }
        // w^(mk) = cos + DIR i sin: y_k = A + DIR*i*B, y_{P-k} = A - DIR*i*B
        C iB = (DIR < 0) ? C{Bv.y, -Bv.x} : C{-Bv.y, Bv.x};
        y[k] = cadd(A, iB);
        y[P - k] = csub(A, iB);
    }
#pragma unroll
    for (int q = 0; q < P; ++q) a[q] = y[q];
}

template <int R, int DIR, typename C> __device__ __forceinline__ void dft(C* a)
{
    if constexpr (R == 1) {
    } else if constexpr (R == 2) {
        dft2<DIR>(a[0], a[1]);
    } else if constexpr (R == 4) {
        dft4<DIR>(a[0], a[1], a[2], a[3]);
    } else if constexpr (R == 3 || R == 5 || R == 7) {
        dft_odd<R, DIR>(a);
    } else if constexpr (R == 8) {
        C e0 = a[0], e1 = a[2], e2 = a[4], e3 = a[6];
        C o0 = a[1], o1 = a[3], o2 = a[5], o3 = a[7];
        dft4<DIR>(e0, e1, e2, e3);
        dft4<DIR>(o0, o1, o2, o3);
        o1 = w16<DIR, 2>(o1);
        o2 = w16<DIR, 4>(o2);
        o3 = w16<DIR, 6>(o3);
        a[0] = cadd(e0, o0); a[4] = csub(e0, o0);
        a[1] = cadd(e1, o1); a[5] = csub(e1, o1);
        a[2] = cadd(e2, o2); a[6] = csub(e2, o2);
        a[3] = cadd(e3, o3); a[7] = csub(e3, o3);
    } else if constexpr (R == 16) {
        // n = 4 n1 + n2, k = k1 + 4 k2
#pragma unroll
        for (int n2 = 0; n2 < 4; ++n2) dft4<DIR>(a[n2], a[n2 + 4], a[n2 + 8], a[n2 + 12]);
        // a[n2 + 4 k1] now holds F_{n2}[k1]; twiddle by w16^(n2 k1)
        a[5] = w16<DIR, 1>(a[5]);
        a[9] = w16<DIR, 2>(a[9]);
        a[13] = w16<DIR, 3>(a[13]);
        a[6] = w16<DIR, 2>(a[6]);
        a[10] = w16<DIR, 4>(a[10]);
        a[14] = w16<DIR, 6>(a[14]);
        a[7] = w16<DIR, 3>(a[7]);
        a[11] = w16<DIR, 6>(a[11]);
        a[15] = w16<DIR, 9>(a[15]);
        C y[16];
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) {
            C b0 = a[4 * k1 + 0], b1 = a[4 * k1 + 1], b2 = a[4 * k1 + 2], b3 = a[4 * k1 + 3];
            dft4<DIR>(b0, b1, b2, b3);
            y[k1] = b0;
            y[k1 + 4] = b1;
            y[k1 + 8] = b2;
            y[k1 + 12] = b3;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q) a[q] = y[q];
    } else {
        static_assert(R == 0, "unsupported radix");
    }
}

// ---------------------------------------------------------------------------
// Stockham stage on register data
// ---------------------------------------------------------------------------
// LDS index of pencil position pos: one pad element every 32 spreads the
// stride-R Stockham writes over the banks.
__device__ __forceinline__ int lds_pos(int pos) { return pos + (pos >> 5); }
__host__ __device__ constexpr int lds_len(int L) { return L + (L >> 5) + 1; }

template <int E, int R, int DIR, typename T>
__device__ __forceinline__ void radix_stage(cplx<T> (&v)[E], int t, int TP, int Ns, int tws,
                                            const cplx<T>* __restrict__ tw)
{
    using C = cplx<T>;
    constexpr int NB = E / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        C a[R];
#pragma unroll
        for (int m = 0; m < R; ++m) a[m] = v[i + m * NB];
        if (Ns > 1) {
            const int b = t + i * TP;
            const int k = b % Ns;
#pragma unroll
            for (int m = 1; m < R; ++m) {
                C w = tw[m * k * tws];
                if constexpr (DIR > 0) w.y = -w.y;
                a[m] = cmul(a[m], w);
            }
        }
        dft<R, DIR>(a);
#pragma unroll
        for (int q = 0; q < R; ++q) v[i + q * NB] = a[q];
    }
}

// Write the stage output (positions (b-k)R + k + q Ns) to LDS, read back the
// next stage's input positions t + TP u.  `ls` is the LDS element stride of
// one pencil position, `lds` points at this thread's pencil.
template <int E, int R, typename T>
__device__ __forceinline__ void stage_exchange(cplx<T> (&v)[E], cplx<T>* lds, int ls, int t, int TP,
                                               int Ns)
{
    constexpr int NB = E / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int b = t + i * TP;
        const int k = b % Ns;
        const int base = (b - k) * R + k;
#pragma unroll
        for (int q = 0; q < R; ++q) lds[lds_pos(base + q * Ns) * ls] = v[i + q * NB];
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < E; ++u) v[u] = lds[lds_pos(t + TP * u) * ls];
    __syncthreads();
}

template <int E, int R, int DIR, typename T>
__device__ __forceinline__ void stage_step(cplx<T> (&v)[E], const FftPlan<T>& p, cplx<T>* lds, int ls,
                                           int t, int Ns, bool last)
{
    if constexpr (E % R == 0) {
        radix_stage<E, R, DIR, T>(v, t, p.TP, Ns, p.L / (Ns * R), p.tw);
        if (!last) stage_exchange<E, R, T>(v, lds, ls, t, p.TP, Ns);
    }
}

// Full transform of one pencil.  Every thread of the workgroup must call it
// the same number of times (it contains __syncthreads()).
template <int E, int DIR, typename T>
__device__ __forceinline__ void fft_pencil(cplx<T> (&v)[E], const FftPlan<T>& p, cplx<T>* lds, int ls,
                                           int t)
{
    int Ns = 1;
    for (int s = 0; s < p.nstage; ++s) {
        const int R = p.radix[s];
        const bool last = (s + 1 == p.nstage);
        switch (R) {
        case 2: stage_step<E, 2, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        case 3: stage_step<E, 3, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        case 4: stage_step<E, 4, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        case 5: stage_step<E, 5, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        case 7: stage_step<E, 7, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        case 8: stage_step<E, 8, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        case 16: stage_step<E, 16, DIR, T>(v, p, lds, ls, t, Ns, last); break;
        default: break;
        }
        Ns *= R;
    }
}

}  // namespace psamd
