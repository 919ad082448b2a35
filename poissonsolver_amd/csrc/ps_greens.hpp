// Green's kernel construction (FP64 on the device).
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "../../include/ps_capi.h"

namespace psamd {

constexpr int kGreensEBC = PS_EBC;
constexpr int kGreensENOMEM = PS_ENOMEM;
constexpr int kGreensEHIP = PS_EHIP;
constexpr int kGreensEUNSUPPORTED = PS_EUNSUPPORTED;

// Ghat * zeta / (Mx My Mz) on the octant kx <= Mx/2, ky <= My/2, kz <= Mz/2,
// stored [ky][kz][kx] (layout 0) or [ky][kx][kz] (layout 1) in `out` (device
// memory of that many T).  Returns a PS_* status.
// dim2 = 1: the 2D kernel of greens2d.cpp for a 2D grid (nx, ny) held as
// ncell = (nx, 1, ny), bc = (bc_x, 1, bc_y), dx = (dx, dx, dy).
template <typename T>
int build_greens(const int ncell[3], const int bc[3], const double dx[3], int reg, T* out, int layout,
                 hipStream_t s, std::string* err, int dim2 = 0);

// Device Si (which = 0) or Ji0 (which = 1) at n host arguments.
int eval_special_device(const double* x, double* y, long n, int which, std::string* err);

}  // namespace psamd
