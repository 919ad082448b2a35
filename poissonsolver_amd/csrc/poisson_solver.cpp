// The reference's C++ class on top of the C ABI.  Error behaviour follows the
// reference: a message on std::cerr and a plain return (solve3d.cpp:57-60,
// greens3d.cpp:522-541); last_status() exposes the status code.
#include "../../include/poisson_solver.hpp"

#include <iostream>

bool poisson_solver::report(int status, const char* where)
{
    status_ = status;
    if (status == PS_OK) return true;
    std::cerr << " [poisson_solver." << where << "]: " << ps_strerror(status) << ": " << ps_last_error(ctx_)
              << std::endl;
    return false;
}

poisson_solver::poisson_solver(void) { report(ps_create(&ctx_), "constructor"); }

poisson_solver::~poisson_solver(void)
{
    if (ctx_) ps_destroy(ctx_);
    ctx_ = NULL;
}

void poisson_solver::set_precision(int precision)
{
    precision_ = precision;
    report(ps_set_precision(ctx_, precision), "set_precision");
}

void poisson_solver::setup2d(int ncell[2], int bc[2], double dx[2], int reg_order)
{
    if (!report(ps_setup2d(ctx_, ncell, bc, dx, reg_order), "setup2d")) return;
    refresh_partition();
}

void poisson_solver::setup3d(int ncell[3], int bc[3], double dx[3], int reg_order)
{
    if (!report(ps_setup3d(ctx_, ncell, bc, dx, reg_order), "setup3d")) return;
    refresh_partition();
}

// partition[rank] of the reference (setup3d.cpp:46, setup2d.cpp:46): one rank
// holding the whole grid; (nx, ny, 1) and dx (dx, dy, 0) after setup2d
void poisson_solver::refresh_partition()
{
    partition.clear();
    const int n = ps_partition_count(ctx_);
    for (int r = 0; r < n; ++r) {
        class_partition p;
        ps_partition(ctx_, r, p.ncell, p.icell, p.dx);
        partition.push_back(p);
    }
}

void poisson_solver::set_return_grad(bool t) { report(ps_set_return_grad(ctx_, t), "set_return_grad"); }
void poisson_solver::set_return_div(bool t) { report(ps_set_return_div(ctx_, t), "set_return_div"); }
void poisson_solver::set_return_curl(bool t) { report(ps_set_return_curl(ctx_, t), "set_return_curl"); }
void poisson_solver::set_reprojection(bool t) { report(ps_set_reprojection(ctx_, t), "set_reprojection"); }

void poisson_solver::solve2d() { report(ps_solve2d(ctx_), "solve2d"); }
void poisson_solver::solve3d() { report(ps_solve3d(ctx_), "solve3d"); }

void poisson_solver::push(double* bx, double* by, double* bz) { report(ps_push(ctx_, bx, by, bz), "push"); }

void poisson_solver::pull(double* ax, double* ay, double* az, double* bx, double* by, double* bz)
{
    report(ps_pull(ctx_, ax, ay, az, bx, by, bz), "pull");
}

void poisson_solver::push_device(const void* bx, const void* by, const void* bz)
{
    report(ps_push_device(ctx_, bx, by, bz), "push_device");
}

void poisson_solver::pull_device(void* ax, void* ay, void* az, void* bx, void* by, void* bz)
{
    report(ps_pull_device(ctx_, ax, ay, az, bx, by, bz), "pull_device");
}

std::vector<class_partition> partition_setup(int pencil_dir, int domain_ncell[3], int bound_cond[3], double dx[3],
                                             bool extendX, bool extendY, bool extendZ)
{
    std::vector<class_partition> all;
    if (pencil_dir != 0 && pencil_dir != 1) {
        std::cerr << " [poisson_solver.partition_setup3d]: Error." << std::endl;   // partition_setup.cpp:61-65
        return all;
    }
    const bool ext[3] = {extendX, extendY, extendZ};
    class_partition p;
    for (int d = 0; d < 3; ++d) {
        p.ncell[d] = (bound_cond[d] == 0 && ext[d]) ? 2 * domain_ncell[d] : domain_ncell[d];
        p.icell[d] = 0;
        p.dx[d] = dx[d];
    }
    all.push_back(p);
    return all;
}
