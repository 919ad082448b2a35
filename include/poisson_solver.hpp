// poisson_solver.hpp -- drop-in for c++/source/poisson_solver.hpp:39-124.
//
// Same public interface as the reference class: void methods, raw host
// double arrays (x fastest), NULL meaning "component absent", errors printed
// to std::cerr followed by a return.  The solve runs on the current HIP
// device through the C ABI of ps_capi.h.  Extensions (not in the reference):
// set_precision, push_device/pull_device, last_status, handle.
#ifndef PS_POISSON_SOLVER_HPP
#define PS_POISSON_SOLVER_HPP

#include <cstddef>
#include <vector>

#include "class_partition.hpp"
#include "ps_capi.h"

class poisson_solver
{
private:
    ps_ctx* ctx_ = NULL;
    int status_ = PS_OK;
    int precision_ = PS_FP64;
    bool report(int status, const char* where);
    void refresh_partition();

public:
    std::vector<class_partition> partition;

    poisson_solver(void);
    ~poisson_solver(void);
    poisson_solver(const poisson_solver&) = delete;
    poisson_solver& operator=(const poisson_solver&) = delete;

    void setup2d(int[2], int[2], double[2], int = 0);
    void setup3d(int[3], int[3], double[3], int = 0);

    void set_return_grad(bool);
    void set_return_div(bool);
    void set_return_curl(bool);
    void set_reprojection(bool);

    void solve2d();
    void solve3d();

    void push(double*, double*, double*);
    void pull(double*, double*, double*, double* = NULL, double* = NULL, double* = NULL);

    // extensions
    void set_precision(int precision);        // PS_FP64 (default), PS_FP32, PS_MIXED
    void push_device(const void*, const void*, const void*);
    void pull_device(void*, void*, void*, void* = NULL, void* = NULL, void* = NULL);
    int last_status() const { return status_; }
    ps_ctx* handle() { return ctx_; }
};

#endif
