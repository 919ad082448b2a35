"""MI355X-native solve path of the Green's-function Poisson solver
(lap(A) = -B on unbounded, mixed and periodic 3D domains).

The product is the HIP library libpoissonsolver_amd.so behind the C ABI of
include/ps_capi.h; `poisson_solver` mirrors the reference's C++ class.
"""
from ._capi import PS_FP32, PS_FP64, PS_MIXED  # noqa: F401
from .solver import batch, class_partition, eval_special, plan_describe, poisson_solver, write_vti  # noqa: F401
