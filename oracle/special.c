/*
 * Oracle special functions -- TEST INFRASTRUCTURE ONLY.
 *
 * Plain-C restatement of the two special functions the reference evaluates
 * while building its Green's kernels.  Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load this library, and only as the
 * checker; the product path never links it.
 *
 *   oracle_sine_int(x)      Si(x) by the Numerical Recipes `cisi` scheme
 *                           (modified-Lentz continued fraction for |x| > 2,
 *                           paired power series below), eps = 1e-10,
 *                           maxit = 1000, fpmin = 1e-30, fpmax = 1e30.
 *                           Follows c++/source/poisson_solver/special_functions.cpp:238-352
 *                           (the convergence measure |Re(del-1)|+|Im(del-1)| of :286
 *                           and the loop bounds of :277 and :319 included).
 *   oracle_bessel_int_J0(x) Ji0(x) = int_0^x (1 - J0(t))/t dt by Luke (1975),
 *                           Table 9.3: Chebyshev series in T_2k(x/8) for x < 8
 *                           and a complex asymptotic Chebyshev series in
 *                           T_n(10/x - 1) for x >= 8.
 *                           Follows special_functions.cpp:115-229.
 *
 * Pinned by: tests/golden/special_ref.npz (outputs of the reference's own
 * special_functions.cpp, compiled by oracle/Makefile into oracle/_ref),
 * scipy.special.sici and mpmath quadrature of (1-J0(t))/t.
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>

/* Luke (1975) Table 9.3, x < 8: Ji0(x) = sum_k A[k] T_2k(x/8). */
static const double LUKE_A[18] = {
    1.35105091918187636388e+00, 8.3791030734868376979e-01,
   -3.5047963978529462711e-01,  1.2777415867753198659e-01,
   -2.981035698255560990e-02,   4.55219841169387328e-03,
   -4.8408621967185359e-04,     3.780202859916883e-05,
   -2.25886908506771e-06,       1.0664609068423e-07,
   -4.08005443149e-09,          1.2909996251e-10,
   -3.43577839e-12,             7.799552e-14,
   -1.52842e-15,                2.612e-17,
   -3.9e-19,                    1.0e-20 };

/* Luke (1975) Table 9.3, x >= 8: complex coefficients of T_n(10/x - 1). */
static const double LUKE_C_RE[39] = {
    9.5360150809738558095e-01, -5.860838853872331670e-02,
   -1.020283575659856676e-02,   1.96012704043622581e-03,
   -9.574977697756219e-05,     -3.570479477043714e-05,
    1.169677960430223e-05,     -1.64386246452682e-06,
   -7.415845751760e-08,         1.1434387527717e-07,
   -3.600903214141e-08,         6.01257386446e-09,
    1.9124656215e-10,          -5.4892028385e-10,
    2.2740445656e-10,          -5.671490865e-11,
    6.07510983e-12,             2.52060520e-12,
   -1.91255246e-12,             7.4056501e-13,
   -1.8950214e-13,              2.021389e-14,
    1.103617e-14,              -8.89993e-15,
    3.88558e-15,               -1.19200e-15,
    2.1456e-16,                 2.915e-17,
   -4.877e-17,                  2.737e-17,
   -1.080e-17,                  3.08e-18,
   -4.2e-19,                   -2.0e-19,
    2.0e-19,                   -1.1e-19,
    4.0e-20,                   -1.0e-20,
    0.0 };
static const double LUKE_C_IM[39] = {
   -1.3917925930200001236e-01, -1.2902065726135067062e-01,
    1.103004348109535741e-02,   5.1817180856880364e-04,
   -3.0928210173975681e-04,     4.647098443047525e-05,
   -8.198845340928e-08,        -1.91888381006925e-06,
    5.7813667761104e-07,       -8.448997773317e-08,
   -5.25612161520e-09,          7.63257790924e-09,
   -2.68643963177e-09,          5.4279949860e-10,
   -1.744365343e-11,           -3.975692920e-11,
    2.069683990e-11,           -6.39623674e-12,
    1.16359235e-12,             6.759603e-14,
   -1.6557337e-13,              8.425597e-14,
   -2.824474e-14,               6.07698e-15,
   -3.171e-17,                 -7.7237e-16,
    4.8022e-16,                -1.9502e-16,
    5.671e-17,                 -8.62e-18,
   -2.69e-18,                   3.09e-18,
   -1.67e-18,                   6.6e-19,
   -1.9e-19,                    3.0e-20,
    1.0e-20,                   -1.0e-20,
    1.0e-20 };

double oracle_sine_int(double x)
{
    const double pi = 3.1415926535897932;
    const double eps = 1.0e-10, fpmin = 1.0e-30, fpmax = 1.0e30, tmin = 2.0;
    const int maxit = 1000;
    double t = fabs(x), ans;

    if (t == 0.0) return 0.0;
    if (t > tmin) {
        /* continued fraction, special_functions.cpp:268-297 */
        double complex b = CMPLX(1.0, t);
        double complex c = CMPLX(fpmax, 0.0);
        double complex d = 1.0 / b;
        double complex h = d;
        double err = 1.0;
        int i = 2;
        while (i < maxit && err > eps) {
            double a = -(double)((i - 1) * (i - 1));
            b = b + 2.0;
            d = 1.0 / (a * d + b);
            c = b + a / c;
            double complex del = c * d;
            h = h * del;
            err = fabs(creal(del - 1.0)) + fabs(cimag(del - 1.0));
            ++i;
        }
        h = CMPLX(cos(t), -sin(t)) * h;
        ans = 0.5 * pi + cimag(h);
    } else if (t < sqrt(fpmin)) {
        ans = t;
    } else {
        /* paired series for Ci and Si, special_functions.cpp:299-344 */
        double sum = 0.0, sums = 0.0, sumc = 0.0, sign = 1.0, fact = 1.0;
        double err = 1.0;
        int odd = 1, k = 1;
        while (k <= maxit && err > eps) {
            fact = fact * t / k;
            double term = fact / k;
            sum = sum + sign * term;
            err = term / fabs(sum);
            if (odd) { sign = -sign; sums = sum; sum = sumc; }
            else     { sumc = sum; sum = sums; }
            odd = !odd;
            ++k;
        }
        ans = sums;
    }
    return x < 0.0 ? -ans : ans;
}

double oracle_bessel_int_J0(double x)
{
    const double pi = 3.1415926535897932;
    const double gamma = 0.5772156649015329;
    if (x < 8.0) {
        /* three-term recurrence for T_n(x/8); even orders carry A[n/2] */
        double x8 = 0.125 * x, t0 = 1.0, t1 = x8, ans = LUKE_A[0];
        for (int n = 2; n < 36; ++n) {
            double t2 = 2 * x8 * t1 - t0;
            if (n % 2 == 0) ans += LUKE_A[n / 2] * t2;
            t0 = t1;
            t1 = t2;
        }
        return ans;
    }
    double u = 2.0 * 5.0 / x - 1.0;
    double t0 = 1.0, t1 = u;
    double complex s = CMPLX(LUKE_C_RE[0], LUKE_C_IM[0]) * t0;
    s += CMPLX(LUKE_C_RE[1], LUKE_C_IM[1]) * t1;
    for (int n = 2; n < 39; ++n) {
        double t2 = 2.0 * u * t1 - t0;
        t0 = t1;
        t1 = t2;
        s += CMPLX(LUKE_C_RE[n], LUKE_C_IM[n]) * t2;
    }
    double fac = cos(x + 0.25 * pi) * creal(s) - sin(x + 0.25 * pi) * cimag(s);
    return sqrt(2.0 / (pi * x)) / x * fac + gamma + log(0.5 * x);
}

/* Vectorised entry points for ctypes. */
void oracle_sine_int_v(const double* x, double* y, size_t n)
{
    for (size_t i = 0; i < n; ++i) y[i] = oracle_sine_int(x[i]);
}

void oracle_bessel_int_J0_v(const double* x, double* y, size_t n)
{
    for (size_t i = 0; i < n; ++i) y[i] = oracle_bessel_int_J0(x[i]);
}
