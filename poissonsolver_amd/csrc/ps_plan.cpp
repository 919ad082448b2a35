#include "ps_plan.hpp"

#include <cmath>

#include "ps_launch.hpp"

namespace psamd {

const std::vector<int>& compiled_E()
{
    static const std::vector<int> es = {
#define PS_E_ITEM(e) e,
        PS_FOR_EACH_E(PS_E_ITEM)
#undef PS_E_ITEM
    };
    return es;
}

static const int kRadices[] = {16, 8, 7, 5, 4, 3, 2};
bool make_plan(int L, int maxE, Plan1D* out, std::string* why, int prefE)
{
    if (prefE < 1) prefE = 16;
    if (L < 1) {
        if (why) *why = "transform length must be positive";
        return false;
    }
    std::vector<int> primes;
    int r = L;
    for (int p = 2; p <= 7; ++p) {
        if (r % p == 0) {
            primes.push_back(p);
            while (r % p == 0) r /= p;
        }
    }
    if (r != 1) {
        if (why) *why = "transform length " + std::to_string(L) + " has a prime factor above 7";
        return false;
    }
    bool found = false;
    Plan1D best;
    for (int E : compiled_E()) {
        if (E > maxE || L % E != 0 || L / E > kMaxThreads) continue;
        bool ok = true;
        for (int p : primes) ok = ok && (E % p == 0);
        if (!ok) continue;
        Plan1D c;
        c.L = L;
        c.E = E;
        c.TP = L / E;
        int rem = L;
        while (rem > 1 && c.nstage < kMaxStage) {
            int pick = 0;
            for (int rr : kRadices) {
                if (E % rr == 0 && rem % rr == 0) {
                    pick = rr;
                    break;
                }
            }
            if (!pick) break;
            c.radix[c.nstage++] = pick;
            rem /= pick;
        }
        if (rem != 1) continue;
        // fewer stages first (one LDS exchange each), then E closest to prefE
        auto score = [prefE](const Plan1D& q) { return q.nstage * 100 + std::abs(q.E - prefE); };
        if (!found || score(c) < score(best)) {
            best = c;
            found = true;
        }
    }
    if (!found) {
        if (why) *why = "no compiled FFT plan for length " + std::to_string(L);
        return false;
    }
    *out = best;
    return true;
}

int max_E_for(int precision_bytes, int ncomp)
{
    if (ncomp >= 3) return 10;
    (void)precision_bytes;
    return 70;
}

static int clamp_lds(int bytes) { return bytes < 16384 ? 16384 : (bytes > 160 * 1024 ? 160 * 1024 : bytes); }

int choose_B(const Plan1D& p, int elem_bytes, int ncol, int lds_bytes)
{
    const int budget = clamp_lds(lds_bytes);
    const int cbytes = 2 * elem_bytes;
    const int target = (elem_bytes == 4) ? 16 : 8;   // 128-byte row segments
    int B = target;
    while (B > 1 && (p.TP * B > kMaxThreads || B * lds_len(p.L) * cbytes > budget)) B /= 2;
    while (B > 1 && B / 2 >= ncol) B /= 2;
    return B < 1 ? 1 : B;
}

int choose_P(const Plan1D& p, int elem_bytes, int npairs, int lds_bytes)
{
    const int budget = clamp_lds(lds_bytes);
    const int cbytes = 2 * elem_bytes;
    int P = 8;
    while (P > 1 && (p.TP * P > kMaxThreads || P * lds_len(p.L) * cbytes > budget / 2)) P /= 2;
    while (P > 1 && P / 2 >= npairs) P /= 2;
    // at least one wavefront per workgroup when the pencils are short
    while (p.TP * P < 64 && P * 2 <= npairs && p.TP * P * 2 <= kMaxThreads &&
           2 * P * lds_len(p.L) * cbytes <= budget)
        P *= 2;
    return P < 1 ? 1 : P;
}

void twiddles(int L, std::vector<double>& re, std::vector<double>& im)
{
    re.resize(L);
    im.resize(L);
    const long double two_pi = 6.283185307179586476925286766559005768L;
    for (int j = 0; j < L; ++j) {
        const long double a = two_pi * (long double)j / (long double)L;
        re[j] = (double)cosl(a);
        im[j] = (double)(-sinl(a));
    }
}

}  // namespace psamd
