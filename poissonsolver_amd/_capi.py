"""ctypes binding of the C ABI (include/ps_capi.h) of libpoissonsolver_amd.so.

The library is the in-tree build (python -m poissonsolver_amd.build).  There is
no fallback: if the shared library is missing, importing raises.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpoissonsolver_amd.so")

PS_OK, PS_EINVAL, PS_EBC, PS_ESTATE, PS_ENOMEM, PS_EHIP, PS_EUNSUPPORTED = range(7)
PS_FP64, PS_FP32, PS_MIXED = 0, 1, 2
PS_POT_A, PS_POT_DX, PS_POT_DY, PS_POT_DZ, PS_POT_LAP = 1, 2, 4, 8, 16

_c_int_p = ctypes.POINTER(ctypes.c_int)
_c_dbl_p = ctypes.POINTER(ctypes.c_double)
_c_long_p = ctypes.POINTER(ctypes.c_long)

TUNING_FIELDS = ("zblk", "tile_x", "tile_y", "tile_z", "radix_e", "lds_kb", "xcd", "specialised",
                 "specialised3", "twreg_z", "twreg_y", "occ4_z", "zwave", "halfb_z", "halfb_y", "mfma",
                 "t2d", "zfan")


class ps_tuning(ctypes.Structure):
    """struct ps_tuning of ps_capi.h."""
    _fields_ = [(f, ctypes.c_int) for f in TUNING_FIELDS]

    def as_dict(self):
        return {f: getattr(self, f) for f in TUNING_FIELDS}


_tun_p = ctypes.POINTER(ps_tuning)

# name -> (restype, argtypes)
SIGNATURES = {
    "ps_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p)]),
    "ps_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_set_precision": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_setup3d": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_dbl_p, ctypes.c_int]),
    "ps_setup2d": (ctypes.c_int, [ctypes.c_void_p, _c_int_p, _c_int_p, _c_dbl_p, ctypes.c_int]),
    "ps_set_return_grad": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_set_return_div": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_set_return_curl": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_set_reprojection": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_push": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 3),
    "ps_solve3d": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_solve2d": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_pull": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6),
    "ps_partition_count": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_partition": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p]),
    "ps_strerror": (ctypes.c_char_p, [ctypes.c_int]),
    "ps_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
    "ps_real_bytes": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_push_device": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 3),
    "ps_pull_device": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6),
    "ps_stream": (ctypes.c_void_p, [ctypes.c_void_p]),
    "ps_synchronize": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_profile_solve": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p]),
    "ps_plan_describe": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _c_int_p, _c_int_p,
                                        _c_int_p, _c_int_p]),
    "ps_eval_special": (ctypes.c_int, [_c_dbl_p, _c_dbl_p, ctypes.c_long, ctypes.c_int]),
    "ps_solve3d_async": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_solve3d_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    "ps_solve3d_potential": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_pull_potential": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5),
    "ps_pull_potential_device": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5),
    "ps_profile_potential": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _c_dbl_p]),
    "ps_set_device_memory_limit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_ulonglong]),
    "ps_device_memory_used": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ulonglong)]),
    "ps_solve2d_async": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_is_2d": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_push_strided": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_long_p, _c_long_p]),
    "ps_pull_strided": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, _c_long_p, ctypes.c_void_p,
                                       ctypes.c_int, _c_long_p, _c_long_p]),
    "ps_smooth3d": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_double]),
    "ps_write_vti": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _c_int_p, _c_dbl_p, _c_dbl_p, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]),
    "ps_zpass_variant": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_potential_zlaunches": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_potential_ylaunches": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_potential_youtputs": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_dft32_fragments": (ctypes.c_int, [ctypes.POINTER(ctypes.c_ushort), ctypes.c_int]),
    "ps_dft16_fragments": (ctypes.c_int, [ctypes.POINTER(ctypes.c_ushort), ctypes.c_int]),
    "ps_dft16_table": (ctypes.c_int, [ctypes.POINTER(ctypes.c_ushort), ctypes.c_int, ctypes.c_int]),
    "ps_tuning_default": (ctypes.c_int, [_tun_p]),
    "ps_set_tuning": (ctypes.c_int, [ctypes.c_void_p, _tun_p]),
    "ps_get_tuning": (ctypes.c_int, [ctypes.c_void_p, _tun_p]),
    "ps_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _tun_p, _c_dbl_p]),
    "ps_batch_create": (ctypes.c_int, [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, _c_int_p, _c_int_p, _c_dbl_p,
                                       _c_int_p, ctypes.c_int]),
    "ps_batch_destroy": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_batch_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int]),
    "ps_batch_set_ops": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_int] * 4),
    "ps_batch_context": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    "ps_batch_push": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3),
    "ps_batch_push_device": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 3),
    "ps_batch_solve": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_batch_pull": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_void_p] * 6),
    "ps_batch_profile": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, _c_dbl_p]),
    "ps_batch_grouped": (ctypes.c_int, [ctypes.c_void_p]),
    "ps_batch_last_error": (ctypes.c_char_p, [ctypes.c_void_p]),
}

_LIB = None


def lib():
    """Load the shared library once; raise if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libpoissonsolver_amd.so is not built: run python -m poissonsolver_amd.build")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def strerror(status: int) -> str:
    return lib().ps_strerror(status).decode()
