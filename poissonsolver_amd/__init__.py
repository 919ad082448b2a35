"""MI355X-native solve path of the Green's-function Poisson solver
(lap(A) = -B on unbounded, mixed and periodic 3D domains).

The product is the HIP library libpoissonsolver_amd.so behind the C ABI of
include/ps_capi.h; `poisson_solver` mirrors the reference's C++ class.
"""
from ._capi import (PS_EBC, PS_EHIP, PS_EINVAL, PS_ENOMEM, PS_ESTATE, PS_EUNSUPPORTED, PS_FP32,  # noqa: F401
                    PS_FP64, PS_MIXED, PS_OK)
from .solver import batch, class_partition, eval_special, plan_describe, poisson_solver, write_vti  # noqa: F401
