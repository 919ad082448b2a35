"""Writes tests/golden/special_ref.npz: arguments and the outputs of the
reference's own sine_int and bessel_int_J0 (c++/source/poisson_solver/
special_functions.cpp:115-352), taken from oracle/_ref/libref_special.so, which
oracle/Makefile compiles from the reference checkout.  Run in a container that
has the checkout:  make -C oracle && python tests/golden/make_special_golden.py
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def main():
    lib = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libref_special.so"))
    for name in ("ref_sine_int_v", "ref_bessel_int_J0_v"):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    rng = np.random.default_rng(20240601)
    # Si: both branches (t <= 2 series, t > 2 continued fraction), the branch
    # point, tiny and huge arguments, negative arguments, and the lattice
    # values rho = pi*n (dx = h) that the kernel samples hit exactly.
    si_x = np.concatenate([
        np.linspace(0.0, 12.0, 1201),
        np.nextafter(2.0, [0.0, 4.0]), [2.0],
        np.exp(rng.uniform(np.log(1e-17), np.log(6000.0), 1500)),
        -np.exp(rng.uniform(np.log(1e-3), np.log(50.0), 200)),
        np.pi * np.sqrt(np.arange(1, 400, dtype=np.float64)),
    ])
    ji_x = np.concatenate([
        np.linspace(0.0, 16.0, 1601),
        np.nextafter(8.0, [0.0, 10.0]), [8.0],
        np.exp(rng.uniform(np.log(1e-6), np.log(5000.0), 1500)),
        np.pi * np.sqrt(np.arange(1, 400, dtype=np.float64)),
    ])
    si_y = np.empty_like(si_x)
    ji_y = np.empty_like(ji_x)
    lib.ref_sine_int_v(si_x.ctypes.data, si_y.ctypes.data, si_x.size)
    lib.ref_bessel_int_J0_v(ji_x.ctypes.data, ji_y.ctypes.data, ji_x.size)
    out = os.path.join(HERE, "special_ref.npz")
    np.savez_compressed(out, si_x=si_x, si_y=si_y, ji0_x=ji_x, ji0_y=ji_y)
    print("wrote", out, si_x.size, ji_x.size)


if __name__ == "__main__":
    main()
