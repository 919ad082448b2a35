// Drop-in check of the C++ API: the call sequence of the reference's
// c++/test/unbounded3d/sources/main.cpp:116-247 (setup3d, set_return_curl,
// partition, push, solve3d, pull) on the compact vortex ring of
// :34-44 / :150-211, with the rel-L2 error against the analytic velocity of
// :252-314.  Written against include/poisson_solver.hpp only.
#include <cmath>
#include <cstdio>
#include <vector>

#include "poisson_solver.hpp"

int main(int argc, char** argv)
{
    const double c = 10.0, r0 = 0.5;
    const double xmin[3] = {-0.5, -1.0, -1.0}, xmax[3] = {0.5, 1.0, 1.0};
    int bounds[3] = {0, 0, 0};
    int n[3] = {32, 64, 64};
    if (argc > 1) {
        const int s = atoi(argv[1]);
        n[0] = s; n[1] = 2 * s; n[2] = 2 * s;
    }
    double dx[3];
    for (int d = 0; d < 3; ++d) dx[d] = (xmax[d] - xmin[d]) / n[d];

    poisson_solver poisson;
    poisson.setup3d(n, bounds, dx);
    if (poisson.last_status() != PS_OK) return 2;
    poisson.set_return_curl(true);

    const int nx = poisson.partition[0].ncell[0], ny = poisson.partition[0].ncell[1], nz = poisson.partition[0].ncell[2];
    const size_t N = (size_t)nx * ny * nz;
    std::vector<double> Ax(N), Ay(N), Az(N), Bx(N, 0.0), By(N, 0.0), Bz(N, 0.0);
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t ijk = ((size_t)k * ny + j) * nx + i;
                const double x = xmin[0] + (i + 0.5) * dx[0], y = xmin[1] + (j + 0.5) * dx[1],
                             z = xmin[2] + (k + 0.5) * dx[2];
                const double rho = std::sqrt(z * z + y * y), phi = std::sqrt((rho - r0) * (rho - r0) + x * x);
                const double th = std::atan2(z, y);
                if (phi < r0) {
                    const double q = 2.0 * r0 * rho - rho * rho - x * x;
                    const double x2 = x * x, R = rho;
                    const double p = 4 * c * c * std::pow(r0, 4) * x2 * R * R - 16 * std::pow(r0, 4) * std::pow(R, 4) +
                                     32 * std::pow(r0, 3) * std::pow(R, 5) - 24 * r0 * r0 * std::pow(R, 6) +
                                     8 * r0 * std::pow(R, 7) - 4 * std::pow(R, 6) * x2 - 6 * std::pow(R, 4) * x2 * x2 -
                                     4 * R * R * std::pow(x, 6) - 8 * c * std::pow(r0, 5) * std::pow(R, 3) +
                                     8 * c * std::pow(r0, 4) * std::pow(R, 4) - 6 * c * std::pow(r0, 3) * std::pow(R, 5) +
                                     4 * c * c * std::pow(r0, 6) * R * R - 8 * c * c * std::pow(r0, 5) * std::pow(R, 3) +
                                     4 * c * c * std::pow(r0, 4) * std::pow(R, 4) + 2 * c * r0 * r0 * std::pow(R, 6) +
                                     32 * std::pow(r0, 3) * std::pow(R, 3) * x2 - 48 * r0 * r0 * std::pow(R, 4) * x2 -
                                     24 * r0 * r0 * R * R * x2 * x2 + 24 * r0 * std::pow(R, 5) * x2 +
                                     24 * r0 * std::pow(R, 3) * x2 * x2 + 8 * r0 * R * std::pow(x, 6) +
                                     2 * c * std::pow(r0, 3) * R * x2 * x2 + 2 * c * r0 * r0 * R * R * x2 * x2 -
                                     4 * c * std::pow(r0, 3) * std::pow(R, 3) * x2 + 4 * c * r0 * r0 * std::pow(R, 4) * x2 -
                                     std::pow(R, 8) - std::pow(x, 8);
                    const double Bmag = -std::exp(-c * r0 * r0 / q) * p * std::pow(q, -4) / (R * R);
                    By[ijk] = -std::sin(th) * Bmag;
                    Bz[ijk] = std::cos(th) * Bmag;
                }
            }
    poisson.push(Bx.data(), By.data(), Bz.data());
    poisson.solve3d();
    poisson.pull(Ax.data(), Ay.data(), Az.data(), Bx.data(), By.data(), Bz.data());
    if (poisson.last_status() != PS_OK) return 3;

    double err = 0.0, nrm = 0.0;
    for (int k = 0; k < nz; ++k)
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                const size_t ijk = ((size_t)k * ny + j) * nx + i;
                const double x = xmin[0] + (i + 0.5) * dx[0], y = xmin[1] + (j + 0.5) * dx[1],
                             z = xmin[2] + (k + 0.5) * dx[2];
                const double rho = std::sqrt(z * z + y * y), phi = std::sqrt((rho - r0) * (rho - r0) + x * x);
                const double th = std::atan2(z, y);
                double sx = 0, sy = 0, sz = 0;
                if (phi < r0) {
                    const double q = 2.0 * r0 * rho - rho * rho - x * x;
                    const double e = std::exp(-c * r0 * r0 / q);
                    const double Amag = 2.0 * c * r0 * r0 * x * e / (q * q);
                    sx = e * (4 * r0 * r0 * rho * rho - 4 * r0 * std::pow(rho, 3) - 4 * r0 * rho * x * x +
                              std::pow(rho, 4) + 2 * rho * rho * x * x + std::pow(x, 4) + 2 * c * std::pow(r0, 3) * rho -
                              2 * c * r0 * r0 * rho * rho) / (q * q) / rho;
                    sy = std::cos(th) * Amag;
                    sz = std::sin(th) * Amag;
                }
                err += (Ax[ijk] - sx) * (Ax[ijk] - sx) + (Ay[ijk] - sy) * (Ay[ijk] - sy) + (Az[ijk] - sz) * (Az[ijk] - sz);
                nrm += sx * sx + sy * sy + sz * sz;
            }
    std::printf("Error: %17.6e %17.6e\n", dx[0], std::sqrt(err / nrm));
    return std::sqrt(err / nrm) < 5e-3 ? 0 : 1;
}
