// Special functions of the regularised Green's kernels, FP64, host and device.
//
//   ps_sine_int(x)       Si(x) by the Numerical Recipes `cisi` scheme with the
//                        reference's constants (eps 1e-10, maxit 1000,
//                        fpmin 1e-30, fpmax 1e30, branch at |x| = 2) and its
//                        convergence measure |Re(del-1)| + |Im(del-1)|
//                        (c++/source/poisson_solver/special_functions.cpp:238-352).
//                        The truncation error of eps = 1e-10 is part of the
//                        reference's kernel, so it is kept for FP64 parity.
//   ps_bessel_int_J0(x)  Ji0(x) = int_0^x (1 - J0(t))/t dt, Chebyshev
//                        approximations of Luke (1975) Table 9.3
//                        (special_functions.cpp:115-229).
#pragma once
#include <hip/hip_runtime.h>

namespace psamd {

struct zd {
    double x, y;
};
__host__ __device__ inline zd zadd(zd a, zd b) { return {a.x + b.x, a.y + b.y}; }
__host__ __device__ inline zd zmul(zd a, zd b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__host__ __device__ inline zd zscale(zd a, double s) { return {a.x * s, a.y * s}; }
// a / b by Smith's algorithm (no intermediate overflow for |b| up to 1e30)
__host__ __device__ inline zd zdiv(zd a, zd b)
{
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return {(a.x + a.y * r) / d, (a.y - a.x * r) / d};
    }
    const double r = b.x / b.y, d = b.x * r + b.y;
    return {(a.x * r + a.y) / d, (a.y * r - a.x) / d};
}

__host__ __device__ inline double ps_sine_int(double x)
{
    const double pi = 3.1415926535897932;
    const double eps = 1.0e-10, fpmin = 1.0e-30, fpmax = 1.0e30, tmin = 2.0;
    const int maxit = 1000;
    const double t = fabs(x);
    double ans;
    if (t == 0.0) return 0.0;
    if (t > tmin) {
        // modified Lentz evaluation of the continued fraction for E1(i t)
        zd b = {1.0, t};
        zd c = {fpmax, 0.0};
        zd d = zdiv(zd{1.0, 0.0}, b);
        zd h = d;
        double err = 1.0;
        for (int i = 2; i < maxit && err > eps; ++i) {
            const double a = -(double)((i - 1) * (i - 1));
            b.x += 2.0;
            d = zdiv(zd{1.0, 0.0}, zadd(zscale(d, a), b));
            c = zadd(b, zdiv(zd{a, 0.0}, c));
            const zd del = zmul(c, d);
            h = zmul(h, del);
            err = fabs(del.x - 1.0) + fabs(del.y);
        }
        h = zmul(zd{cos(t), -sin(t)}, h);
        ans = 0.5 * pi + h.y;
    } else if (t < sqrt(fpmin)) {
        ans = t;
    } else {
        // power series of Si and Ci summed together; the odd terms are Si's
        double sum = 0.0, sums = 0.0, sumc = 0.0, sign = 1.0, fact = 1.0, err = 1.0;
        bool odd = true;
        for (int k = 1; k <= maxit && err > eps; ++k) {
            fact = fact * t / k;
            const double term = fact / k;
            sum = sum + sign * term;
            err = term / fabs(sum);
            if (odd) {
                sign = -sign;
                sums = sum;
                sum = sumc;
            } else {
                sumc = sum;
                sum = sums;
            }
            odd = !odd;
        }
        ans = sums;
    }
    return x < 0.0 ? -ans : ans;
}

__host__ __device__ inline double ps_bessel_int_J0(double x)
{
    // Luke (1975) Table 9.3.  x < 8: sum_k a_k T_2k(x/8).
    const double a[18] = {1.35105091918187636388,  0.83791030734868376979,  -0.35047963978529462711,
                          0.12777415867753198659,  -0.02981035698255560990, 0.00455219841169387328,
                          -0.00048408621967185359, 0.00003780202859916883,  -0.00000225886908506771,
                          0.00000010664609068423,  -0.00000000408005443149, 0.00000000012909996251,
                          -0.00000000000343577839, 0.00000000000007799552,  -0.00000000000000152842,
                          0.00000000000000002612,  -0.00000000000000000039, 0.00000000000000000001};
    const double pi = 3.1415926535897932;
    const double gamma = 0.5772156649015329;
    if (x < 8.0) {
        const double u = 0.125 * x;
        double t0 = 1.0, t1 = u, s = a[0];
        for (int n = 2; n < 36; ++n) {
            const double t2 = 2 * u * t1 - t0;
            if ((n & 1) == 0) s += a[n / 2] * t2;
            t0 = t1;
            t1 = t2;
        }
        return s;
    }
    // x >= 8: complex series in T_n(10/x - 1)
    const double cr[39] = {
        0.95360150809738558095,  -0.05860838853872331670, -0.01020283575659856676, 0.00196012704043622581,
        -0.00009574977697756219, -0.00003570479477043714, 0.00001169677960430223,  -0.00000164386246452682,
        -0.00000007415845751760, 0.00000011434387527717,  -0.00000003600903214141, 0.00000000601257386446,
        0.00000000019124656215,  -0.00000000054892028385, 0.00000000022740445656,  -0.00000000005671490865,
        0.00000000000607510983,  0.00000000000252060520,  -0.00000000000191255246, 0.00000000000074056501,
        -0.00000000000018950214, 0.00000000000002021389,  0.00000000000001103617,  -0.00000000000000889993,
        0.00000000000000388558,  -0.00000000000000119200, 0.00000000000000021456,  0.00000000000000002915,
        -0.00000000000000004877, 0.00000000000000002737,  -0.00000000000000001080, 0.00000000000000000308,
        -0.00000000000000000042, -0.00000000000000000020, 0.00000000000000000020,  -0.00000000000000000011,
        0.00000000000000000004,  -0.00000000000000000001, 0.00000000000000000000};
    const double ci[39] = {
        -0.13917925930200001236, -0.12902065726135067062, 0.01103004348109535741,  0.00051817180856880364,
        -0.00030928210173975681, 0.00004647098443047525,  -0.00000008198845340928, -0.00000191888381006925,
        0.00000057813667761104,  -0.00000008448997773317, -0.00000000525612161520, 0.00000000763257790924,
        -0.00000000268643963177, 0.00000000054279949860,  -0.00000000001744365343, -0.00000000003975692920,
        0.00000000002069683990,  -0.00000000000639623674, 0.00000000000116359235,  0.00000000000006759603,
        -0.00000000000016557337, 0.00000000000008425597,  -0.00000000000002824474, 0.00000000000000607698,
        -0.00000000000000003171, -0.00000000000000077237, 0.00000000000000048022,  -0.00000000000000019502,
        0.00000000000000005671,  -0.00000000000000000862, -0.00000000000000000269, 0.00000000000000000309,
        -0.00000000000000000167, 0.00000000000000000066,  -0.00000000000000000019, 0.00000000000000000003,
        0.00000000000000000001,  -0.00000000000000000001, 0.00000000000000000001};
    const double u = 2.0 * 5.0 / x - 1.0;
    double t0 = 1.0, t1 = u;
    zd s = {cr[0] * t0, ci[0] * t0};
    s = zadd(s, zd{cr[1] * t1, ci[1] * t1});
    for (int n = 2; n < 39; ++n) {
        const double t2 = 2.0 * u * t1 - t0;
        t0 = t1;
        t1 = t2;
        s = zadd(s, zd{cr[n] * t2, ci[n] * t2});
    }
    const double fac = cos(x + 0.25 * pi) * s.x - sin(x + 0.25 * pi) * s.y;
    return sqrt(2.0 / (pi * x)) / x * fac + gamma + log(0.5 * x);
}

}  // namespace psamd
