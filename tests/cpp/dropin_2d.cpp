// Drop-in check of the C++ API for the 2D solver: the call sequences of the
// reference's c++/test/unbounded2d (setup2d, set_return_grad, push Bx, solve2d,
// pull Ax Ay; main.cpp:116-207), periodic2d (plain solve, pull of lhs and
// rhs; main.cpp:118-193) and mixed2d (x periodic, y unbounded, grad;
// main.cpp:122-220) on their own sources and grids, with the rel-L2 error
// against their analytic fields.  The unbounded case also writes the VTK
// files the reference programs write.  Written against
// include/poisson_solver.hpp only.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "poisson_solver.hpp"

namespace {

const double pi = std::acos(-1.0);

struct Case {
    const char* name;
    int ncell[2], bc[2];
    double xmin[2], xmax[2];
};

double run(const Case& cs, const char* vtk_dir)
{
    double dx[2];
    for (int d = 0; d < 2; ++d) dx[d] = (cs.xmax[d] - cs.xmin[d]) / cs.ncell[d];
    int n[2] = {cs.ncell[0], cs.ncell[1]}, bc[2] = {cs.bc[0], cs.bc[1]};
    poisson_solver poisson;
    poisson.setup2d(n, bc, dx);
    if (poisson.last_status() != PS_OK) return -1.0;
    const bool periodic = !std::strcmp(cs.name, "periodic2d");
    poisson.set_return_grad(!periodic);
    const int nx = poisson.partition[0].ncell[0], ny = poisson.partition[0].ncell[1];
    if (poisson.partition[0].ncell[2] != 1) return -1.0;
    const size_t N = (size_t)nx * ny;
    std::vector<double> Ax(N, 0.0), Ay(N, 0.0), Bx(N, 0.0), Sx(N, 0.0), Sy(N, 0.0);
    const double c = 10.0;
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            const size_t ij = (size_t)j * nx + i;
            const double x = cs.xmin[0] + (i + 0.5) * dx[0], y = cs.xmin[1] + (j + 0.5) * dx[1];
            if (!std::strcmp(cs.name, "unbounded2d")) {           // unbounded2d main.cpp:152-170, :230-249
                const double r2 = x * x + y * y;
                if (r2 < 1.0) {
                    const double d = 1.0 - r2, e = std::exp(-c / d);
                    Bx[ij] = 4.0 * c * e * (1.0 - x * x * x * x - y * y * y * y - 2.0 * x * x * y * y - c * r2) /
                             (d * d * d * d);
                    Sx[ij] = -2.0 * c * x * e / (d * d);
                    Sy[ij] = -2.0 * c * y * e / (d * d);
                }
            } else if (periodic) {                                // periodic2d main.cpp:154, :212
                Bx[ij] = 8.0 * pi * pi * std::sin(2.0 * pi * x) * std::sin(2.0 * pi * y);
                Sx[ij] = std::sin(2.0 * pi * x) * std::sin(2.0 * pi * y);
            } else if (std::fabs(y) < 1.0) {                      // mixed2d main.cpp:159-177, :239-247
                const double q = (y - 1.0) * (y + 1.0), e = std::exp(c * y * y / q);
                const double y2 = y * y, y4 = y2 * y2, y6 = y4 * y2, y8 = y4 * y4;
                Bx[ij] = e * std::sin(pi * x) *
                         (pi * pi * y8 - 4 * pi * pi * y6 + 6 * pi * pi * y4 - 6 * c * y4 - 4 * pi * pi * y2 -
                          4 * c * c * y2 + 4 * c * y2 + pi * pi + 2 * c) /
                         (q * q * q * q);
                Sx[ij] = pi * e * std::cos(pi * x);
                Sy[ij] = -2.0 * c * y * e * std::sin(pi * x) / (q * q);
            }
        }
    poisson.push(Bx.data(), NULL, NULL);
    poisson.solve2d();
    if (periodic) poisson.pull(Ax.data(), NULL, NULL, Bx.data(), NULL, NULL);
    else poisson.pull(Ax.data(), Ay.data(), NULL);
    if (poisson.last_status() != PS_OK) return -1.0;
    double err = 0.0, nrm = 0.0;
    for (size_t ij = 0; ij < N; ++ij) {
        err += (Ax[ij] - Sx[ij]) * (Ax[ij] - Sx[ij]);
        nrm += Sx[ij] * Sx[ij];
        if (!periodic) {
            err += (Ay[ij] - Sy[ij]) * (Ay[ij] - Sy[ij]);
            nrm += Sy[ij] * Sy[ij];
        }
    }
    if (vtk_dir) {
        const int n3[3] = {nx, ny, 1};
        const double org[3] = {cs.xmin[0], cs.xmin[1], 0.0}, sp[3] = {dx[0], dx[1], 0.0};
        const char* names[2] = {"B", "A"};
        const double* comps[6] = {Bx.data(), NULL, NULL, Ax.data(), Ay.data(), NULL};
        if (ps_write_vti(vtk_dir, "mesh", n3, org, sp, 2, names, comps, 0) != PS_OK) return -1.0;
    }
    return std::sqrt(err / nrm);
}

}  // namespace

int main(int argc, char** argv)
{
    const char* vtk_dir = argc > 1 ? argv[1] : NULL;
    const Case cases[3] = {
        {"unbounded2d", {128, 128}, {0, 0}, {-1.0, -1.0}, {1.0, 1.0}},
        {"periodic2d", {128, 128}, {1, 1}, {0.0, 0.0}, {1.0, 1.0}},
        {"mixed2d", {128, 256}, {1, 0}, {-1.0, -2.0}, {1.0, 2.0}},
    };
    int bad = 0;
    for (const Case& cs : cases) {
        const double e = run(cs, !std::strcmp(cs.name, "unbounded2d") ? vtk_dir : NULL);
        std::printf("%s Error: %17.6e\n", cs.name, e);
        if (!(e >= 0.0 && e < 1e-12)) bad = 1;
    }
    return bad;
}
