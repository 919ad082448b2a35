"""Python mirror of the reference's `poisson_solver` class (poisson_solver.hpp:39-124).

Same method names and argument meaning: ``setup3d(ncell, bc, dx, reg_order=0)``
and ``setup2d(ncell, bc, dx, reg_order=0)``, ``set_return_grad/div/curl(bool)``,
``set_reprojection(bool)``, ``push(bx, by, bz)`` with None for an absent
component, ``solve3d()`` / ``solve2d()``, ``pull(ax, ay, az, bx, by, bz)``
filling caller-owned float64 arrays, and the public ``partition`` list.  Arrays
hold nx*ny*nz doubles with x fastest (numpy shape (nz, ny, nx); (ny, nx) in 2D).
Like the reference, a failing call prints a message to stderr and returns;
``last_status`` holds the status code.  The solve runs on the GPU through the
C ABI; there is no CPU path.
"""
from __future__ import annotations

import ctypes
import sys
from dataclasses import dataclass

import numpy as np

from . import _capi


@dataclass
class class_partition:
    """class_partition.hpp:29-43."""
    ncell: list
    icell: list
    dx: list


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


class poisson_solver:
    def __init__(self, precision: int = _capi.PS_FP64):
        self._lib = _capi.lib()
        self._ctx = ctypes.c_void_p()
        self.last_status = self._lib.ps_create(ctypes.byref(self._ctx))
        self.partition = []
        self._n = None
        if self.last_status == _capi.PS_OK and precision != _capi.PS_FP64:
            self.set_precision(precision)

    # -- lifecycle ---------------------------------------------------------
    def close(self):
        if getattr(self, "_ctx", None) and self._ctx.value:
            self._lib.ps_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _report(self, status, where):
        self.last_status = status
        if status != _capi.PS_OK:
            msg = self._lib.ps_last_error(self._ctx).decode() if self._ctx.value else ""
            print(" [poisson_solver.%s]: %s: %s" % (where, _capi.strerror(status), msg), file=sys.stderr)
            return False
        return True

    # -- setup (setup3d.cpp:8-71) ------------------------------------------
    def set_precision(self, precision: int):
        return self._report(self._lib.ps_set_precision(self._ctx, int(precision)), "set_precision")

    def setup3d(self, ncell, bound_cond, dx, reg_order: int = 0):
        n = (ctypes.c_int * 3)(*[int(v) for v in ncell])
        b = (ctypes.c_int * 3)(*[int(v) for v in bound_cond])
        d = (ctypes.c_double * 3)(*[float(v) for v in dx])
        if not self._report(self._lib.ps_setup3d(self._ctx, n, b, d, int(reg_order)), "setup3d"):
            return
        self._n = tuple(int(v) for v in ncell)
        self._refresh_partition()

    def _refresh_partition(self):
        self.partition = []
        for r in range(self._lib.ps_partition_count(self._ctx)):
            nc = (ctypes.c_int * 3)()
            ic = (ctypes.c_int * 3)()
            dd = (ctypes.c_double * 3)()
            self._lib.ps_partition(self._ctx, r, nc, ic, dd)
            self.partition.append(class_partition(list(nc), list(ic), list(dd)))

    def setup2d(self, ncell, bound_cond, dx, reg_order: int = 0):
        """setup2d.cpp:8-72 (kernels of greens2d.cpp:8-441)."""
        n = (ctypes.c_int * 2)(*[int(v) for v in ncell])
        b = (ctypes.c_int * 2)(*[int(v) for v in bound_cond])
        d = (ctypes.c_double * 2)(*[float(v) for v in dx])
        if not self._report(self._lib.ps_setup2d(self._ctx, n, b, d, int(reg_order)), "setup2d"):
            return
        self._n = (int(ncell[0]), int(ncell[1]), 1)
        self._refresh_partition()

    # -- options (settings.cpp:9-30) ---------------------------------------
    def set_return_grad(self, toggle: bool):
        self._report(self._lib.ps_set_return_grad(self._ctx, int(bool(toggle))), "set_return_grad")

    def set_return_div(self, toggle: bool):
        self._report(self._lib.ps_set_return_div(self._ctx, int(bool(toggle))), "set_return_div")

    def set_return_curl(self, toggle: bool):
        self._report(self._lib.ps_set_return_curl(self._ctx, int(bool(toggle))), "set_return_curl")

    def set_reprojection(self, toggle: bool):
        self._report(self._lib.ps_set_reprojection(self._ctx, int(bool(toggle))), "set_reprojection")

    # -- data --------------------------------------------------------------
    def _host(self, a, writable):
        if a is None:
            return None
        if not isinstance(a, np.ndarray) or a.dtype != np.float64 or not a.flags.c_contiguous:
            raise TypeError("arrays must be C-contiguous float64 numpy arrays")
        if writable and not a.flags.writeable:
            raise TypeError("output arrays must be writable")
        if self._n is not None and a.size != self._n[0] * self._n[1] * self._n[2]:
            raise ValueError("array size %d does not match the grid %s" % (a.size, self._n))
        return a

    def push(self, bx=None, by=None, bz=None):
        """push.cpp:9-397."""
        arrs = [self._host(a, False) for a in (bx, by, bz)]
        self._report(self._lib.ps_push(self._ctx, *[_ptr(a) for a in arrs]), "push")

    def solve3d(self):
        """solve3d.cpp:8-635."""
        self._report(self._lib.ps_solve3d(self._ctx), "solve3d")

    def solve2d(self):
        """solve2d.cpp:8-475."""
        self._report(self._lib.ps_solve2d(self._ctx), "solve2d")

    def pull(self, ax=None, ay=None, az=None, bx=None, by=None, bz=None):
        """pull.cpp:9-469."""
        arrs = [self._host(a, True) for a in (ax, ay, az, bx, by, bz)]
        self._report(self._lib.ps_pull(self._ctx, *[_ptr(a) for a in arrs]), "pull")

    # -- the Fortran interface's one-array forms ---------------------------
    @staticmethod
    def _block(arr, first, writable):
        """A numpy array of shape (nz', ny', nx', ncomp) (C order: the memory
        layout of Fortran real(ncomp, nx', ny', nz')) or (ny', nx', ncomp) in
        2D, holding the physical cells from index `first` (x, y[, z])."""
        if not isinstance(arr, np.ndarray) or arr.dtype != np.float64:
            raise TypeError("blocks must be float64 numpy arrays")
        if writable and not arr.flags.writeable:
            raise TypeError("output blocks must be writable")
        if arr.ndim not in (3, 4):
            raise ValueError("blocks have 4 axes (3 in 2D)")
        st = [v // 8 for v in arr.strides]
        if any(v % 8 for v in arr.strides) or min(st) < 1:
            raise ValueError("blocks need positive element strides")
        if arr.ndim == 4:
            stride = (ctypes.c_long * 4)(st[3], st[2], st[1], st[0])
        else:   # (y, x, c): the z stride is never stepped (k = 0)
            stride = (ctypes.c_long * 4)(st[2], st[1], st[0], st[0] * arr.shape[0])
        f = list(first) + [0] * (3 - len(first))
        return arr.shape[-1], stride, (ctypes.c_long * 3)(*[int(v) for v in f])

    def push_strided(self, block, first=(0, 0, 0)):
        """poisson_solver_push_1pointer3d (poisson_solver_push.f:8): the
        components of `block` (1: X, 2: X, Y, 3: X, Y, Z) from cell `first` on."""
        nc, st, f = self._block(block, first, False)
        self._report(self._lib.ps_push_strided(self._ctx, ctypes.c_void_p(block.ctypes.data), nc, st, f),
                     "push_strided")

    def pull_strided(self, lhs_block, rhs_block=None, first=(0, 0, 0)):
        """poisson_solver_pull_1pointer3d (poisson_solver_pull.f:8)."""
        nl, sl, f = self._block(lhs_block, first, True)
        if rhs_block is not None:
            nr, sr, _ = self._block(rhs_block, first, True)
            rb = ctypes.c_void_p(rhs_block.ctypes.data)
        else:
            nr, sr, rb = 0, None, None
        self._report(self._lib.ps_pull_strided(self._ctx, ctypes.c_void_p(lhs_block.ctypes.data), nl, sl, rb, nr,
                                               sr, f), "pull_strided")

    def smooth3d(self, h: float):
        """poisson_solver_smooth3d (poisson_solver_smooth3d.f:8): Gaussian
        smoothing exp(-h^2 k^2 / 2) of every held rhs component."""
        self._report(self._lib.ps_smooth3d(self._ctx, float(h)), "smooth3d")

    # -- extensions --------------------------------------------------------
    def push_device(self, bx=None, by=None, bz=None):
        """Device pointers (int addresses, e.g. torch tensor.data_ptr()) of the
        solve precision's real type."""
        p = [None if v is None else ctypes.c_void_p(int(v)) for v in (bx, by, bz)]
        self._report(self._lib.ps_push_device(self._ctx, *p), "push_device")

    def pull_device(self, ax=None, ay=None, az=None, bx=None, by=None, bz=None):
        p = [None if v is None else ctypes.c_void_p(int(v)) for v in (ax, ay, az, bx, by, bz)]
        self._report(self._lib.ps_pull_device(self._ctx, *p), "pull_device")

    def stream(self) -> int:
        return self._lib.ps_stream(self._ctx) or 0

    def solve3d_async(self):
        """Enqueue solve3d on the context's stream and return."""
        self._report(self._lib.ps_solve3d_async(self._ctx), "solve3d_async")

    POT_NAMES = ("A", "dx", "dy", "dz", "lap")

    def solve3d_potential(self, outputs=("A", "dx", "dy", "dz")):
        """A, grad A and / or the Laplacian of A from the held scalar X in one
        pipeline (ps_solve3d_potential); outputs from POT_NAMES."""
        mask = 0
        for name in outputs:
            mask |= 1 << self.POT_NAMES.index(name)
        self._report(self._lib.ps_solve3d_potential(self._ctx, mask), "solve3d_potential")

    def potential_zlaunches(self, outputs=("A", "dx", "dy", "dz")) -> int:
        """z-pass launches a potential solve of these outputs makes (1: its
        z outputs from one forward transform); negative before setup3d."""
        mask = 0
        for name in outputs:
            mask |= 1 << self.POT_NAMES.index(name)
        return self._lib.ps_potential_zlaunches(self._ctx, mask)

    def potential_ylaunches(self, outputs=("A", "dx", "dy", "dz")) -> int:
        """y^-1 launches a potential solve of these outputs makes (A, dA/dx
        and dA/dy may share one); negative before setup3d."""
        mask = 0
        for name in outputs:
            mask |= 1 << self.POT_NAMES.index(name)
        return self._lib.ps_potential_ylaunches(self._ctx, mask)

    def potential_youtputs(self, outputs=("A", "dx", "dy", "dz")) -> int:
        """y^-1 outputs a potential solve of these outputs writes (dA/dx may be
        taken from A's in the x^-1 pass); negative before setup3d."""
        mask = 0
        for name in outputs:
            mask |= 1 << self.POT_NAMES.index(name)
        return self._lib.ps_potential_youtputs(self._ctx, mask)

    def pull_potential(self, A=None, dx=None, dy=None, dz=None, lap=None):
        arrs = [self._host(a, True) for a in (A, dx, dy, dz, lap)]
        self._report(self._lib.ps_pull_potential(self._ctx, *[_ptr(a) for a in arrs]), "pull_potential")

    def profile_potential(self, outputs=("A", "dx", "dy", "dz"), iters: int = 1):
        mask = 0
        for name in outputs:
            mask |= 1 << self.POT_NAMES.index(name)
        ms = (ctypes.c_double * 6)()
        if not self._report(self._lib.ps_profile_potential(self._ctx, mask, int(iters), ms), "profile_potential"):
            return None
        return dict(zip(("x", "y", "z", "y_inv", "x_inv", "solve"), list(ms)))

    def set_device_memory_limit(self, nbytes: int):
        """Device bytes this context may hold (0: no limit)."""
        return self._report(self._lib.ps_set_device_memory_limit(self._ctx, int(nbytes)), "set_device_memory_limit")

    def device_memory_used(self) -> int:
        v = ctypes.c_ulonglong()
        self._report(self._lib.ps_device_memory_used(self._ctx, ctypes.byref(v)), "device_memory_used")
        return int(v.value)

    def synchronize(self):
        self._report(self._lib.ps_synchronize(self._ctx), "synchronize")

    ZPASS_VARIANTS = {0: "runtime-plan", 1: "specialised", 2: "pencil-per-wavefront", 3: "matrix-core",
                      4: "two-residue"}

    def zpass_variant(self) -> str:
        """Kernel family of the one-component z pass chosen by setup3d."""
        v = self._lib.ps_zpass_variant(self._ctx)
        return self.ZPASS_VARIANTS.get(v, "not set up")

    def tuning(self) -> dict:
        """Tuning the last setup3d used (before setup: the one it will use)."""
        t = _capi.ps_tuning()
        self._report(self._lib.ps_get_tuning(self._ctx, ctypes.byref(t)), "get_tuning")
        return t.as_dict()

    def set_tuning(self, tuning=None, **fields):
        """Tuning for the next setup3d: a dict and/or keyword fields over the
        defaults; set_tuning(None) returns to the defaults."""
        if tuning is None and not fields:
            return self._report(self._lib.ps_set_tuning(self._ctx, None), "set_tuning")
        t = _capi.ps_tuning()
        self._lib.ps_tuning_default(ctypes.byref(t))
        for k, v in dict(tuning or {}, **fields).items():
            setattr(t, k, int(v))
        return self._report(self._lib.ps_set_tuning(self._ctx, ctypes.byref(t)), "set_tuning")

    def autotune(self, iters: int = 5):
        """Pick the fastest tuning for this grid and options; pushed data is
        discarded.  Returns (tuning dict, ms per solve) or None."""
        t = _capi.ps_tuning()
        ms = ctypes.c_double()
        if not self._report(self._lib.ps_autotune(self._ctx, int(iters), ctypes.byref(t), ctypes.byref(ms)),
                            "autotune"):
            return None
        return t.as_dict(), ms.value

    def profile_solve(self, iters: int = 1):
        """Per-pass average times (ms) measured with HIP events on the
        context's stream: dict x, y, z, y_inv, x_inv, solve."""
        ms = (ctypes.c_double * 6)()
        if not self._report(self._lib.ps_profile_solve(self._ctx, int(iters), ms), "profile_solve"):
            return None
        return dict(zip(("x", "y", "z", "y_inv", "x_inv", "solve"), list(ms)))


class batch:
    """Independent solves of uneven size replayed as one hipGraph launch
    (ps_batch_* of ps_capi.h).  problems: list of (ncell, bc, dx, reg_order)."""

    def __init__(self, problems, precision: int = _capi.PS_FP32):
        self._lib = _capi.lib()
        self._b = ctypes.c_void_p()
        n = len(problems)
        cell = (ctypes.c_int * (3 * n))(*[int(v) for p in problems for v in p[0]])
        bc = (ctypes.c_int * (3 * n))(*[int(v) for p in problems for v in p[1]])
        dx = (ctypes.c_double * (3 * n))(*[float(v) for p in problems for v in p[2]])
        reg = (ctypes.c_int * n)(*[int(p[3]) for p in problems])
        self.shapes = [tuple(int(v) for v in p[0]) for p in problems]
        st = self._lib.ps_batch_create(ctypes.byref(self._b), n, cell, bc, dx, reg, int(precision))
        if st != _capi.PS_OK:
            raise RuntimeError("ps_batch_create: " + _capi.strerror(st))

    def _check(self, st, where):
        if st != _capi.PS_OK:
            raise RuntimeError("%s: %s: %s" % (where, _capi.strerror(st), self._lib.ps_batch_last_error(self._b).decode()))

    def close(self):
        if getattr(self, "_b", None) and self._b.value:
            self._lib.ps_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_ops(self, grad=False, div=False, curl=False, reproject=False):
        self._check(self._lib.ps_batch_set_ops(self._b, int(grad), int(div), int(curl), int(reproject)), "set_ops")

    def autotune(self, iters: int = 3):
        self._check(self._lib.ps_batch_autotune(self._b, int(iters)), "autotune")

    def tuning(self, prob: int) -> dict:
        c = ctypes.c_void_p()
        self._check(self._lib.ps_batch_context(self._b, int(prob), ctypes.byref(c)), "context")
        t = _capi.ps_tuning()
        self._lib.ps_get_tuning(c, ctypes.byref(t))
        return t.as_dict()

    def push(self, prob: int, bx=None, by=None, bz=None):
        arrs = [None if a is None else np.ascontiguousarray(a, np.float64) for a in (bx, by, bz)]
        self._check(self._lib.ps_batch_push(self._b, int(prob), *[_ptr(a) for a in arrs]), "push")

    def push_device(self, prob: int, bx=None, by=None, bz=None):
        p = [None if v is None else ctypes.c_void_p(int(v)) for v in (bx, by, bz)]
        self._check(self._lib.ps_batch_push_device(self._b, int(prob), *p), "push_device")

    def solve(self):
        self._check(self._lib.ps_batch_solve(self._b), "solve")

    def pull(self, prob: int, ax=None, ay=None, az=None, bx=None, by=None, bz=None):
        for a in (ax, ay, az, bx, by, bz):
            if a is not None and (a.dtype != np.float64 or not a.flags.c_contiguous or not a.flags.writeable):
                raise TypeError("output arrays must be writable C-contiguous float64")
        self._check(self._lib.ps_batch_pull(self._b, int(prob), *[_ptr(a) for a in (ax, ay, az, bx, by, bz)]),
                    "pull")

    def profile(self, iters: int = 10) -> float:
        """Average ms of one graph launch over `iters` launches."""
        ms = ctypes.c_double()
        self._check(self._lib.ps_batch_profile(self._b, int(iters), ctypes.byref(ms)), "profile")
        return ms.value

    def grouped(self) -> int:
        """Problems the last solve ran inside group solves (one solve per grid)."""
        return int(self._lib.ps_batch_grouped(self._b))


def write_vti(directory, stem, ncell, origin, dx, fields, binary=False):
    """VTK output of the reference's test programs
    (c++/test/unbounded3d/sources/main.cpp:329-455): fields maps a name to a
    tuple of up to 3 float64 arrays of nx*ny*nz cells (x fastest; None =
    zeros).  Writes <directory>/<stem>_P00.vti and <directory>/<stem>.pvti."""
    names = list(fields)
    keep = []
    ptrs = []
    for nm in names:
        comps = list(fields[nm]) + [None] * (3 - len(fields[nm]))
        for c in comps:
            if c is None:
                ptrs.append(None)
                continue
            a = np.ascontiguousarray(c, dtype=np.float64)
            if a.size != int(ncell[0]) * int(ncell[1]) * int(ncell[2]):
                raise ValueError("field %s: %d cells, grid has %d" % (nm, a.size, int(np.prod(ncell))))
            keep.append(a)
            ptrs.append(a.ctypes.data)
    cn = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    cp = (ctypes.c_void_p * max(1, len(ptrs)))(*ptrs)
    st = _capi.lib().ps_write_vti(str(directory).encode(), str(stem).encode(),
                                  (ctypes.c_int * 3)(*[int(v) for v in ncell]),
                                  (ctypes.c_double * 3)(*[float(v) for v in origin]),
                                  (ctypes.c_double * 3)(*[float(v) for v in dx]), len(names), cn, cp, int(bool(binary)))
    if st != _capi.PS_OK:
        raise OSError("ps_write_vti: " + _capi.strerror(st))


def plan_describe(L: int, ncomp: int = 1, precision_bytes: int = 8):
    """(E, threads per pencil, radices) the library picks for length L, or None."""
    E, tp, ns = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    rad = (ctypes.c_int * 14)()
    st = _capi.lib().ps_plan_describe(int(L), int(ncomp), int(precision_bytes), ctypes.byref(E),
                                      ctypes.byref(tp), ctypes.byref(ns), rad)
    if st != _capi.PS_OK:
        return None
    return E.value, tp.value, [rad[i] for i in range(ns.value)]


def eval_special(x, which: str):
    """Device Si ('si') or Ji0 ('ji0') at float64 arguments x."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty_like(x)
    st = _capi.lib().ps_eval_special(x.ctypes.data_as(_capi._c_dbl_p), y.ctypes.data_as(_capi._c_dbl_p),
                                     x.size, 0 if which == "si" else 1)
    if st != _capi.PS_OK:
        raise RuntimeError("ps_eval_special: " + _capi.strerror(st))
    return y
