// Host-side paths of the library built with AddressSanitizer and
// UndefinedBehaviorSanitizer (host code only; the device objects are the
// ordinary build): argument checks, error texts, tuning, planning, the
// C++ class's print-and-return behaviour, the VTK writer and batch creation.
// Runs without a GPU (device calls return PS_EHIP).  Prints "SANITIZE_OK".
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "poisson_solver.hpp"
#include "ps_capi.h"

static int fails = 0;
#define EXPECT(c)                                                              \
    do {                                                                       \
        if (!(c)) {                                                            \
            std::printf("FAILED %s:%d %s\n", __FILE__, __LINE__, #c);          \
            ++fails;                                                           \
        }                                                                      \
    } while (0)

int main(int argc, char** argv)
{
    const char* dir = argc > 1 ? argv[1] : ".";
    ps_ctx* c = nullptr;
    EXPECT(ps_create(&c) == PS_OK);
    int n[3] = {8, 6, 4}, bc[3] = {0, 0, 0};
    double dx[3] = {0.1, 0.1, 0.1};
    int bad_bc[3] = {0, 1, 0};
    EXPECT(ps_setup3d(c, n, bad_bc, dx, 0) == PS_EBC);
    EXPECT(std::strstr(ps_last_error(c), "z-direction") != nullptr);
    int bad_n[3] = {8, 0, 4};
    EXPECT(ps_setup3d(c, bad_n, bc, dx, 0) == PS_EINVAL);
    EXPECT(ps_setup3d(c, n, bc, dx, -3) == PS_EINVAL);
    const int st = ps_setup3d(c, n, bc, dx, 2);
    EXPECT(st == PS_EHIP || st == PS_OK);
    EXPECT(ps_solve3d(c) == PS_ESTATE || st == PS_OK);
    EXPECT(ps_set_precision(c, 9) == PS_EINVAL);
    EXPECT(ps_set_device_memory_limit(c, 1u << 20) == PS_OK);
    unsigned long long used = 1;
    EXPECT(ps_device_memory_used(c, &used) == PS_OK);
    EXPECT(ps_solve3d_potential(c, 0) == PS_ESTATE || ps_solve3d_potential(c, 0) == PS_EINVAL);
    EXPECT(ps_solve3d_noalloc(c, nullptr, 1, nullptr) != PS_OK);
    ps_tuning t;
    EXPECT(ps_tuning_default(&t) == PS_OK);
    t.mfma = 99;
    EXPECT(ps_set_tuning(c, &t) == PS_EINVAL);
    EXPECT(ps_set_tuning(c, nullptr) == PS_OK);
    for (int s = -1; s <= 7; ++s) EXPECT(ps_strerror(s) != nullptr);
    EXPECT(ps_destroy(c) == PS_OK);
    EXPECT(ps_destroy(nullptr) == PS_EINVAL);

    // planning of every length up to 4096
    for (int L = 1; L <= 4096; ++L) {
        int E = 0, tp = 0, ns = 0, rad[14];
        const int r = ps_plan_describe(L, 1, 8, &E, &tp, &ns, rad);
        if (r == PS_OK) {
            long prod = 1;
            for (int k = 0; k < ns; ++k) prod *= rad[k];
            EXPECT(prod == L && E * tp == L);
        }
        (void)ps_plan_describe(L, 3, 4, &E, &tp, &ns, rad);
    }

    // the reference's class: void methods, message and return
    {
        poisson_solver ps;
        ps.setup3d(n, bad_bc, dx, 0);
        EXPECT(ps.last_status() == PS_EBC);
        ps.solve3d();
        EXPECT(ps.last_status() == PS_ESTATE);
        std::vector<double> a(8 * 6 * 4, 1.0);
        ps.push(a.data(), nullptr, nullptr);
        ps.pull(a.data(), nullptr, nullptr);
        std::vector<class_partition> parts;
        partition_setup(0, n, bc, dx, true, true, true);
    }

    // VTK writer (host only)
    {
        const int nc[3] = {3, 2, 2};
        const double org[3] = {-1, -1, -1}, h[3] = {0.5, 0.5, 0.5};
        std::vector<double> f(12);
        for (int i = 0; i < 12; ++i) f[i] = i * 0.25;
        const char* names[1] = {"field"};
        const double* comps[3] = {f.data(), nullptr, f.data()};
        EXPECT(ps_write_vti(dir, "san", nc, org, h, 1, names, comps, 0) == PS_OK);
        EXPECT(ps_write_vti(dir, "san_raw", nc, org, h, 1, names, comps, 1) == PS_OK);
        EXPECT(ps_write_vti(nullptr, "x", nc, org, h, 1, names, comps, 0) != PS_OK);
    }

    // batch creation without a device fails cleanly
    {
        ps_batch* b = nullptr;
        const int ncell[6] = {8, 8, 8, 6, 6, 6}, bcs[6] = {0, 0, 0, 0, 0, 0}, reg[2] = {2, 2};
        const double dxs[6] = {0.1, 0.1, 0.1, 0.1, 0.1, 0.1};
        const int r = ps_batch_create(&b, 2, ncell, bcs, dxs, reg, PS_FP32);
        if (r == PS_OK) ps_batch_destroy(b);
        EXPECT(ps_batch_destroy(nullptr) == PS_EINVAL);
        EXPECT(ps_batch_solve(nullptr) == PS_EINVAL);
    }
    if (fails) return 1;
    std::printf("SANITIZE_OK\n");
    return 0;
}
