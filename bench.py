"""Benchmark of the 3D solve path on one MI355X.

Default workload: 512^3 fully unbounded scalar Poisson solve (bc 0,0,0; domain
doubled to 1024^3 through the R2C x pass and the pruned z pass), regularisation
order 2, Gaussian source (s = 0.1 on [-1,1]^3), FP32 storage, in the
mixed-precision form (z pass on the matrix cores: FP16 hi/lo DFT-32 stages with
FP32 accumulation); the plain FP32 path is timed beside it on the same input
and both are checked against the FP64 GPU solve.  One step = one `solve3d`
with the source already resident in HBM (pushed before the timed region).
Prints one JSON line.

  python bench.py [--steps K] [--warmup W] [--prec mixed|fp32|fp64]
                  [--n 512] [--op plain|curl] [--no-cpu] [--no-check] [--no-compare]
  python bench.py --workload ppu256|vortex512|large1024   (configurations 2, 3, 5)
  python bench.py --workload batch [--nprob 64] [--autotune]
      64 independent unbounded scalar problems, n from {32,...,192} (seed 4),
      sources = sums of 1-4 Gaussian blobs; one step = one hipGraph launch
      solving the whole batch.
  python bench.py --workload uu2d [--n 4096]
      2D solver: the reference's unbounded2d case (bump source on [-1,1]^2,
      grad) at n^2, FP64 with FP32 beside it; one step = one solve2d.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "solves/s + HBM GB/s, 512^3 unbounded: MFMA mixed-prec vs FP32 FFT @relL2≤1e-6"
HBM_PEAK_GBS = 8000.0


def gaussian_source(n, s=0.1):
    h = 2.0 / n
    x = -1.0 + (np.arange(n) + 0.5) * h
    r2 = (x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2)
    return np.exp(-r2 / (2 * s * s)) / ((2 * np.pi) ** 1.5 * s ** 3), h


def alg_bytes(n, cin, cout, eb):
    """Algorithmic HBM bytes of the pruned 5-pass pipeline (R2C on x, z never
    stored doubled, real octant of the kernel): per pass P1..P5."""
    nx = ny = nz = n
    Mx, My, Mz = 2 * n, 2 * n, 2 * n
    H = Mx // 2 + 1
    er, ec, eg = eb, 2 * eb, eb
    P1 = cin * (nx * ny * nz * er + H * ny * nz * ec)
    P2 = cin * (H * ny * nz + H * My * nz) * ec
    P3 = cin * H * My * nz * ec + (Mx // 2 + 1) * (My // 2 + 1) * (Mz // 2 + 1) * eg + cout * H * My * nz * ec
    P4 = cout * (H * My * nz + H * ny * nz) * ec
    P5 = cout * (H * ny * nz * ec + nx * ny * nz * er)
    return dict(x=P1, y=P2, z=P3, y_inv=P4, x_inv=P5, solve=P1 + P2 + P3 + P4 + P5)


def alg_bytes_grid(ncell, bc, cin, cout, eb):
    """The pruned 5-pass byte model for any boundary tuple (M_d = 2 n_d for an
    unbounded direction, n_d for a periodic one): cin spectra through the x,
    y and z passes (the z pass reads each input once), cout outputs back."""
    nx, ny, nz = ncell
    M = [2 * n if b == 0 else n for n, b in zip(ncell, bc)]
    H = M[0] // 2 + 1
    er, ec, eg = eb, 2 * eb, eb
    P1 = cin * (nx * ny * nz * er + H * ny * nz * ec)
    P2 = cin * (H * ny * nz + H * M[1] * nz) * ec
    P3 = cin * H * M[1] * nz * ec + (M[0] // 2 + 1) * (M[1] // 2 + 1) * (M[2] // 2 + 1) * eg + cout * H * M[1] * nz * ec
    P4 = cout * (H * M[1] * nz + H * ny * nz) * ec
    P5 = cout * (H * ny * nz * ec + nx * ny * nz * er)
    return dict(x=P1, y=P2, z=P3, y_inv=P4, x_inv=P5, solve=P1 + P2 + P3 + P4 + P5)


def alg_bytes_potential(ncell, bc, outputs, eb, zlaunches=None, ylaunches=None, youtputs=None):
    """Byte model of the potential form (ps_solve3d_potential): one scalar
    through the x and y passes; one z transform per distinct z multiplier (A,
    which d/dx and d/dy share, d/dz and the Laplacian), each reading the one
    spectrum and Gk once and writing its own, or, when the z outputs run from
    one forward transform (zlaunches 1), the spectrum and Gk read once and
    every output written; the y^-1 and x^-1 passes once per output (d/dx and
    d/dy read A's z output, times i k on load), the y^-1 reads once per
    launch when outputs share one (ylaunches) and writes youtputs spectra
    (d/dx A may be taken from A's by the x^-1 pass, which then reads A's
    rows once for both when A is wanted too)."""
    one = alg_bytes_grid(ncell, bc, 1, 1, eb)
    back = alg_bytes_grid(ncell, bc, 1, len(outputs), eb)
    nzq = int(any(k in outputs for k in ("A", "dx", "dy"))) + int("dz" in outputs) + int("lap" in outputs)
    if zlaunches == 1 and nzq > 1:
        two = alg_bytes_grid(ncell, bc, 1, 2, eb)
        z = one["z"] + (nzq - 1) * (two["z"] - one["z"])
    else:
        z = nzq * one["z"]
    y_inv = back["y_inv"]
    M = [2 * n if b == 0 else n for n, b in zip(ncell, bc)]
    H = M[0] // 2 + 1
    if ylaunches is not None and 0 < ylaunches < len(outputs):
        y_inv -= (len(outputs) - ylaunches) * H * M[1] * ncell[2] * 2 * eb
    x_inv = back["x_inv"]
    if youtputs is not None and 0 < youtputs < len(outputs):
        y_inv -= (len(outputs) - youtputs) * H * ncell[1] * ncell[2] * 2 * eb
        if "A" in outputs and "dx" in outputs:
            x_inv -= H * ncell[1] * ncell[2] * 2 * eb
    d = dict(x=one["x"], y=one["y"], z=z, y_inv=y_inv, x_inv=x_inv)
    d["solve"] = sum(d.values())
    return d


def cpu_baseline(n_target, cin, op, bc=(0, 0, 0)):
    """The oracle (numpy FP64 restatement of the reference, one core) timed on
    a bounded sample: push+solve+pull at 256^3 (15-20 s a solve on one core;
    repeated while the total stays under 10 s), scaled to n_target by
    n^3 log n.  The reference itself cannot be built here (it needs MPI)."""
    from oracle import oracle as O
    ns = min(256, n_target)
    b, h = gaussian_source(ns)
    s = O.OracleSolver()
    s.setup3d((ns,) * 3, tuple(bc), (h,) * 3, 2)
    s.set_return_curl(op == "curl")
    reps, t = 0, 0.0
    while reps < 3 and t < 10.0:
        t0 = time.perf_counter()
        s.push(*([b] * cin))
        s.solve3d()
        out = [np.empty_like(b) for _ in range(3)]
        s.pull(*out)
        t += time.perf_counter() - t0
        reps += 1
    per = t / reps
    scale = (n_target / ns) ** 3 * math.log2(2 * n_target) / math.log2(2 * ns)
    return {"value": 1.0 / (per * scale), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "oracle (numpy FP64 C2C restatement) push+solve3d+pull at %d^3 bc %s reg2, "
                      "%d reps, %.2f s/solve, scaled to %d^3 by n^3 log n (x%.0f)"
                      % (ns, "".join(str(v) for v in bc), reps, per, n_target, scale)}


CALIB_BYTES = 1 << 30


def is_zpass_kernel(name, nc):
    """The fused z pass: k_zmfma1024 / k_zmfma1024p (matrix-core, one
    component), k_zfused_w (pencil per wavefront, one component) or
    k_pencil_s<T, L, NC, MODE, ...> with MODE 2 and the given NC."""
    if nc == 1 and ("k_zmfma1024" in name or "k_zfused_w<" in name or "k_zfused_r2<" in name):
        return True
    if "k_pencil_s<" not in name:
        return False
    args = [x.strip() for x in name.split("k_pencil_s<", 1)[1].split(">", 1)[0].split(",")]
    return len(args) >= 4 and args[2] == str(nc) and args[3] == "2"


def calib_copy():
    """A plain 1 GiB device-to-device copy: the reference kernel for scaling
    the FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import torch
    src = torch.ones(CALIB_BYTES // 4, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    torch.cuda.synchronize()
    dst.copy_(src)
    torch.cuda.synchronize()
    del src, dst


def measure_traffic(argv_child, kernel_key, timeout=240, take_last=0):
    """HBM bytes per dispatch of the dominant kernel from two rocprofv3 --pmc
    runs (FETCH_SIZE, then WRITE_SIZE) of this script in a child process, each
    scaled by what the counter reads for the 1 GiB calibration copy.
    take_last > 0: average only the last take_last matching dispatches (the
    timed solves; set-up kernels of the same name run first).
    Returns (traffic_bytes, detail) or (None, reason)."""
    import csv
    import shutil
    import subprocess
    import tempfile
    exe = shutil.which("rocprofv3")
    if not exe:
        return None, "rocprofv3 not found"
    out = {}
    env = dict(os.environ, TMPDIR="/tmp")
    for ctr in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp(prefix="ps_pmc_", dir="/tmp")
        cmd = [exe, "--pmc", ctr, "--output-format", "csv", "-d", d, "-o", "p", "--",
               sys.executable, os.path.abspath(__file__)] + argv_child
        try:
            subprocess.run(cmd, cwd="/tmp", env=env, stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           timeout=timeout, check=False)
        except subprocess.TimeoutExpired:
            return None, "rocprofv3 %s pass timed out" % ctr
        rows = []
        for root, _, files in os.walk(d):
            for f in files:
                if f.endswith("counter_collection.csv"):
                    rows += list(csv.DictReader(open(os.path.join(root, f))))
        shutil.rmtree(d, ignore_errors=True)
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r["Kernel_Name"]]
        krows = [r for r in rows if kernel_key(r["Kernel_Name"])]
        if take_last and krows and "Dispatch_Id" in krows[0]:
            krows = sorted(krows, key=lambda r: int(r["Dispatch_Id"]))[-take_last:]
        kern = [float(r["Counter_Value"]) for r in krows]
        if not copy or not kern:
            return None, "%s: counters missing (copy %d, kernel %d rows)" % (ctr, len(copy), len(kern))
        factor = (CALIB_BYTES / 1024.0) / max(copy)      # KiB the counter should read / KiB it read
        out[ctr] = dict(kib=sum(kern) / len(kern), factor=factor)
    fetch = out["FETCH_SIZE"]["kib"] * out["FETCH_SIZE"]["factor"] * 1024
    write = out["WRITE_SIZE"]["kib"] * out["WRITE_SIZE"]["factor"] * 1024
    return fetch + write, dict(read_bytes=fetch, write_bytes=write,
                               fetch_factor=out["FETCH_SIZE"]["factor"], write_factor=out["WRITE_SIZE"]["factor"])


BATCH_SIZES = (32, 48, 64, 96, 128, 160, 192)


def batch_problems(nprob, seed=4):
    """Uneven batch: cubic grids on [-1,1]^3 with n drawn from BATCH_SIZES, each
    source a sum of 1-4 Gaussian blobs (centres in [-0.5,0.5]^3, widths in
    [0.05,0.15], amplitudes in [-1,1])."""
    rng = np.random.default_rng(seed)
    probs, srcs = [], []
    for _ in range(nprob):
        n = int(rng.choice(BATCH_SIZES))
        h = 2.0 / n
        x = -1.0 + (np.arange(n) + 0.5) * h
        k = int(rng.integers(1, 5))
        src = np.zeros((n, n, n))
        for _ in range(k):
            c = rng.uniform(-0.5, 0.5, 3)
            w = rng.uniform(0.05, 0.15)
            amp = rng.uniform(-1.0, 1.0)
            gz = np.exp(-(x - c[2]) ** 2 / (2 * w * w))
            gy = np.exp(-(x - c[1]) ** 2 / (2 * w * w))
            gx = np.exp(-(x - c[0]) ** 2 / (2 * w * w))
            src += amp * gz[:, None, None] * gy[None, :, None] * gx[None, None, :]
        probs.append(((n, n, n), (0, 0, 0), (h, h, h), 2))
        srcs.append(src)
    return probs, srcs


def separable_blobs(n, lo, hi, centres, widths, amps, nz_first=True):
    """sum_b amp_b exp(-|x - c_b|^2 / (2 w_b^2)) on an n^3 cell-centred grid of
    [lo, hi]^3, built plane by plane as rank-B products (x fastest)."""
    h = (hi - lo) / n
    x = lo + (np.arange(n) + 0.5) * h
    gx = np.exp(-(x[None, :] - centres[:, 0:1]) ** 2 / (2 * widths[:, None] ** 2))
    gy = np.exp(-(x[None, :] - centres[:, 1:2]) ** 2 / (2 * widths[:, None] ** 2))
    gz = np.exp(-(x[None, :] - centres[:, 2:3]) ** 2 / (2 * widths[:, None] ** 2))
    out = np.zeros((n, n, n))
    cut = 1e-18 * float(np.max(np.abs(amps)))      # below FP64 round-off of the largest term
    for k in range(n):
        wk = amps * gz[:, k]
        live = np.abs(wk) > cut
        if live.any():
            out[k] = (gy[live].T * wk[live]) @ gx[live]
    return out, h


def large_blobs(n, seed=5, nb=256):
    """Superposition of nb Gaussian blobs on an n^3 cell-centred grid of
    [-1,1]^3: widths in [0.01,0.03], amplitudes N(0,1), centres in
    [-0.8,0.8]^3 (seeded)."""
    rng = np.random.default_rng(seed)
    cen = rng.uniform(-0.8, 0.8, (nb, 3))
    wid = rng.uniform(0.01, 0.03, nb)
    amp = rng.standard_normal(nb)
    return separable_blobs(n, -1.0, 1.0, cen, wid, amp)


def vortex_ring_vorticity(n, seed=3, R=0.4, sc=0.05, nseg=2048):
    """Vorticity of a thin vortex ring on [-1,1]^3: circulation 1, Gaussian core
    sc, centreline R(t) = R (1 + 0.01 cos(5 t + phi)) in the z = 0 plane with a
    seeded phase phi.  Each of nseg centreline segments contributes a 3D
    Gaussian of width sc times its tangent vector, so the sum is a tube with a
    Gaussian cross-section (2D integral 1) directed along the centreline."""
    rng = np.random.default_rng(seed)
    phi = rng.uniform(0, 2 * np.pi)
    t = (np.arange(nseg) + 0.5) * 2 * np.pi / nseg
    r = R * (1 + 0.01 * np.cos(5 * t + phi))
    dr = -R * 0.01 * 5 * np.sin(5 * t + phi)
    X = np.stack([r * np.cos(t), r * np.sin(t), np.zeros_like(t)], 1)
    dX = np.stack([dr * np.cos(t) - r * np.sin(t), dr * np.sin(t) + r * np.cos(t), np.zeros_like(t)], 1)
    dX *= 2 * np.pi / nseg                                   # tangent * dt
    norm = 1.0 / ((2 * np.pi) ** 1.5 * sc ** 3)
    w = np.full(nseg, sc)
    comps = []
    for c in range(2):
        f, h = separable_blobs(n, -1.0, 1.0, X, w, norm * dX[:, c])
        comps.append(f)
    comps.append(np.zeros_like(comps[0]))
    return comps, h


def ppu_source(n, seed=2):
    """x, y periodic on [0,1)^2 (h = 1/n), z in [-1,1] (cell-centred, 2/n):
    sin(2 pi x) sin(4 pi y) exp(-z^2 / (2 0.05^2)) plus a seeded random-phase
    field in x, y (spectral amplitude exp(-|k|^2 / (2 16^2)), (0,0) mode zero,
    unit RMS) times the same z envelope."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    x = np.arange(n) * h
    hz = 2.0 / n
    z = -1.0 + (np.arange(n) + 0.5) * hz
    env = np.exp(-z ** 2 / (2 * 0.05 ** 2))
    k = np.fft.fftfreq(n, d=1.0 / n)
    kk = k[None, :] ** 2 + k[:, None] ** 2
    amp = np.exp(-kk / (2 * 16.0 ** 2))
    amp[0, 0] = 0.0
    spec = amp * np.exp(2j * np.pi * rng.uniform(size=(n, n)))
    field = np.real(np.fft.ifft2(spec))
    field /= np.sqrt(np.mean(field ** 2))
    plane = np.sin(2 * np.pi * x)[None, :] * np.sin(4 * np.pi * x)[:, None] + field   # [y][x]
    return env[:, None, None] * plane[None, :, :], (h, h, hz)


def _profile(s, steps, warmup):
    """warmup untimed solves, then steps timed ones, all of the pushed source
    (profile_solve does not apply the filter a solve3d leaves on the held
    rhs, so every step repeats the same work)."""
    if warmup > 0:
        s.profile_solve(warmup)
    s.synchronize()
    return s.profile_solve(steps)


def _rel(a, b):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    den = sum(float(np.sum(y ** 2)) for y in b)
    return math.sqrt(num / den)


def _rel_to(a, b, scale):
    num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(a, b))
    return math.sqrt(num / sum(float(np.sum(y ** 2)) for y in scale))


def bench_config(a):
    """Named workloads of the other configurations (one JSON line each):
      ppu256     x,y periodic / z unbounded 256^3, reg 6: A and grad A
      vortex512  512^3 unbounded vortex ring: streamfunction (3 components),
                 velocity by spectral curl, div of the streamfunction
      large1024  1024^3 unbounded, reg 10, 256 blobs: A, grad A and the
                 Laplacian of A (FP32 and mixed only; FP64 check on every 4th
                 point per direction)
    For ppu256 and large1024 the outputs come two ways: the reference API's
    component-wise solves (A plain; grad of the scalar pushed three times),
    and the potential extension (one pipeline from the one scalar).  `value`
    is the rate of the potential form.  Each solve times its passes with HIP
    events.  --pmc-child (for counter runs): the one precision --prec, no FP64
    check, a progress line per solve and no JSON line."""
    import poissonsolver_amd as P
    codes = {"fp32": P.PS_FP32, "mixed": P.PS_MIXED, "fp64": P.PS_FP64}
    precs = ["mixed", "fp32"] if a.prec == "mixed" and not a.pmc_child else [a.prec]
    pot = None          # outputs of the potential form
    if a.workload == "ppu256":
        n = a.n if a.n != 512 else 256
        src, dx = ppu_source(n)
        ncell, bc, reg = (n, n, n), (1, 1, 0), 6
        solves = [("A", "plain", [src]), ("gradA", "grad", [src, src, src])]
        cin_out = [(1, 1), (3, 3)]
        pot = ("A", "dx", "dy", "dz")
        desc = "%d^3 x,y periodic z unbounded (doubled in z only), reg 6, sin*sin*gauss + random-phase source: A and grad A" % n
    elif a.workload == "vortex512":
        n = a.n
        comps, h = vortex_ring_vorticity(n)
        ncell, bc, reg, dx = (n, n, n), (0, 0, 0), 2, (h, h, h)
        solves = [("psi", "plain", comps), ("u", "curl", comps), ("divpsi", "div", comps)]
        cin_out = [(3, 3), (3, 3), (3, 1)]
        desc = "%d^3 fully unbounded vortex ring (R 0.4, core 0.05, mode-5 1%%): streamfunction, velocity by curl, div psi" % n
    else:
        n = a.n if a.n != 512 else 1024
        src, h = large_blobs(n)
        ncell, bc, reg, dx = (n, n, n), (0, 0, 0), 10, (h, h, h)
        solves = [("A", "plain", [src]), ("gradA", "grad", [src, src, src])]
        cin_out = [(1, 1), (3, 3)]
        pot = ("A", "dx", "dy", "dz", "lap")
        desc = "%d^3 fully unbounded, reg 10, 256 Gaussian blobs: A, grad A and the Laplacian of A" % n
    sub = (slice(None, None, 4),) * 3 if a.workload == "large1024" else (slice(None),) * 3
    shape = (ncell[2], ncell[1], ncell[0])
    results = {}
    outs = {}
    for prec in precs + ([] if a.pmc_child else ["fp64"]):
        per = {}
        for name, op, rhs in solves:
            s = P.poisson_solver(codes[prec])
            s.setup3d(ncell, bc, dx, reg)
            s.set_return_grad(op == "grad")
            s.set_return_div(op == "div")
            s.set_return_curl(op == "curl")
            s.push(*rhs)
            if prec != "fp64":
                prof = _profile(s, a.steps, a.warmup)
                per[name] = {"ms": prof["solve"], "pass_ms": {k: prof[k] for k in ("x", "y", "z", "y_inv", "x_inv")},
                             "z_kernel": s.zpass_variant()}
            else:
                s.solve3d()
            if s.last_status:
                raise RuntimeError("solve %s %s failed: status %d" % (prec, name, s.last_status))
            if a.pmc_child:
                print("solved", prec, name, flush=True)
                s.close()
                continue
            nout = 1 if op == "div" else len(rhs) if op != "plain" else len(rhs)
            lhs = [np.empty(shape) for _ in range(nout)]
            s.pull(*lhs)
            outs[(prec, name)] = [x[sub].copy() for x in lhs]
            del lhs
            s.close()
        if pot is not None:
            # the potential form: one pipeline from the one scalar.  The FP64
            # reference runs the Laplacian alone, and at 1024^3 takes A and grad
            # A from the component-wise FP64 solves (a four-output FP64 solve
            # there would not fit in HBM next to its work buffers)
            big = ncell[0] * ncell[1] * ncell[2] >= 512 ** 3 * 4
            if prec != "fp64":
                # with the Laplacian, also the four outputs of the component-wise solves
                groups = [pot] + ([tuple(k for k in pot if k != "lap")] if "lap" in pot else [])
            else:
                groups = ([] if big else [tuple(k for k in pot if k != "lap")]) + ([("lap",)] if "lap" in pot else [])
            got = {}
            if prec == "fp64" and big:
                for k, x in zip(("A", "dx", "dy", "dz"), outs[("fp64", "A")] + outs[("fp64", "gradA")]):
                    got[k] = x
            for grp in groups:
                s = P.poisson_solver(codes[prec])
                s.setup3d(ncell, bc, dx, reg)
                s.push(src)
                if prec != "fp64":
                    if a.warmup > 0:
                        s.profile_potential(grp, a.warmup)
                    s.synchronize()
                    prof = s.profile_potential(grp, a.steps)
                    per["potential" if grp == pot else "potential_agrad"] = {
                        "ms": prof["solve"], "outputs": list(grp),
                        "pass_ms": {k: prof[k] for k in ("x", "y", "z", "y_inv", "x_inv")},
                        "z_kernel": s.zpass_variant(), "z_launches": s.potential_zlaunches(grp),
                        "y_inv_launches": s.potential_ylaunches(grp),
                        "y_inv_outputs": s.potential_youtputs(grp)}
                else:
                    s.solve3d_potential(grp)
                if s.last_status:
                    raise RuntimeError("potential solve %s %s failed: status %d" % (prec, grp, s.last_status))
                if a.pmc_child:
                    print("solved", prec, "potential", grp, flush=True)
                    s.close()
                    continue
                arrs = {k: (np.empty(shape) if k in grp else None) for k in P.poisson_solver.POT_NAMES}
                s.pull_potential(*[arrs[k] for k in P.poisson_solver.POT_NAMES])
                for k in grp:
                    if k not in got:
                        got[k] = arrs[k][sub].copy()
                del arrs
                s.close()
            if not a.pmc_child:
                outs[(prec, "potential")] = [got[k] for k in pot]
        results[prec] = per
    if a.pmc_child:
        return 0
    line_prec = precs[0]
    comp_ms = sum(results[line_prec][nm]["ms"] for nm, _, _ in solves)
    step_ms = results[line_prec]["potential"]["ms"] if pot is not None else comp_ms
    eb = 8 if line_prec == "fp64" else 4
    out = {
        "metric": METRIC, "value": 1000.0 / step_ms, "unit": "steps/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": line_prec, "data": "synthetic",
        "config": {"workload": a.workload + ": " + desc, "ncell": list(ncell), "bc": list(bc), "reg_order": reg,
                   "step": ["potential: " + ", ".join(pot)] if pot is not None else [nm for nm, _, _ in solves]},
        "solves": {}, "rel_l2_vs_fp64_gpu": {},
    }
    for prec in precs:
        for name, op, _ in solves:
            r = results[prec][name]
            key = "%s/%s" % (prec, name)
            out["solves"][key] = r
            if name == "divpsi":
                # div psi is ~0: its error is measured against the velocity scale
                out["rel_l2_vs_fp64_gpu"][key + " (vs |u|)"] = _rel_to(outs[(prec, name)], outs[("fp64", name)],
                                                                     outs[("fp64", "u")])
            else:
                out["rel_l2_vs_fp64_gpu"][key] = _rel(outs[(prec, name)], outs[("fp64", name)])
        if pot is not None:
            key = "%s/potential" % prec
            out["solves"][key] = results[prec]["potential"]
            if "potential_agrad" in results[prec]:
                out["solves"][key + "_agrad"] = results[prec]["potential_agrad"]
            for k, x, y in zip(pot, outs[(prec, "potential")], outs[("fp64", "potential")]):
                out["rel_l2_vs_fp64_gpu"]["%s/%s" % (key, k)] = _rel([x], [y])
    if pot is not None:
        if not (ncell[0] * ncell[1] * ncell[2] >= 512 ** 3 * 4):
            # the potential form against the component-wise solves (FP64: the same numbers)
            ref = outs[("fp64", "A")] + outs[("fp64", "gradA")]
            out["rel_l2_fp64_potential_vs_componentwise"] = _rel(outs[("fp64", "potential")][:4], ref)
        out["componentwise_ms"] = comp_ms
        agrad = results[line_prec].get("potential_agrad")
        # the same four outputs (A, grad A) as the component-wise solves
        out["speedup_potential_vs_componentwise"] = comp_ms / (agrad["ms"] if agrad else step_ms)
        zl = results[line_prec]["potential"].get("z_launches")
        nzq_pot = int(any(k in pot for k in ("A", "dx", "dy"))) + int("dz" in pot) + int("lap" in pot)
        B = alg_bytes_potential(ncell, bc, pot, eb, zl, results[line_prec]["potential"].get("y_inv_launches"),
                                results[line_prec]["potential"].get("y_inv_outputs"))
        Bc = sum(alg_bytes_grid(ncell, bc, ci, co, eb)["solve"] for ci, co in cin_out)
        pp = results[line_prec]["potential"]["pass_ms"]
        out["alg_bytes_per_step"] = B["solve"]
        if agrad:
            out["alg_bytes_agrad"] = alg_bytes_potential(ncell, bc, agrad["outputs"], eb, agrad.get("z_launches"),
                                                         agrad.get("y_inv_launches"),
                                                         agrad.get("y_inv_outputs"))["solve"]
        out["alg_bytes_componentwise"] = Bc
        out["step_hbm_gbs"] = B["solve"] / (step_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": B["z"] / (pp["z"] * 1e-3) / 1e9, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": B["z"] / (pp["z"] * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                           "kernel": "z pass of the potential form (%s kernel: one spectrum in, %d quantities "
                                     "out through %d inverse z transforms in %d launch(es))"
                                     % (results[line_prec]["potential"]["z_kernel"], len(pot), nzq_pot, zl),
                           "alg_bytes_per_launch": B["z"], "avg_ms": pp["z"]}
    if len(precs) == 2:
        f32 = sum(v["ms"] for k, v in results["fp32"].items() if pot is None or k == "potential")
        out["speedup_vs_fp32"] = f32 / step_ms
    if a.workload == "vortex512":
        # div psi relative to the velocity magnitude scale (L2 ratio, FP64)
        dv = outs[("fp64", "divpsi")]
        uu = outs[("fp64", "u")]
        out["div_psi_over_u_l2"] = math.sqrt(float(np.sum(dv[0] ** 2)) / sum(float(np.sum(x ** 2)) for x in uu))
    print(json.dumps(out), flush=True)
    return 0


def bump2d(n):
    """Source and analytic grad of the reference's unbounded2d test
    (c++/test/unbounded2d/sources/main.cpp:143-170, :230-249; c = 10, r0 = 1)
    at cell centres of [-1,1]^2, arrays (ny, nx)."""
    h = 2.0 / n
    x = -1.0 + (np.arange(n) + 0.5) * h
    X, Y = x[None, :], x[:, None]
    r2 = X * X + Y * Y
    inside = r2 < 1.0
    d = np.where(inside, 1.0 - r2, 1.0)
    e = np.where(inside, np.exp(-10.0 / d), 0.0)
    B = np.where(inside, 40.0 * e * (1.0 - X ** 4 - Y ** 4 - 2 * X ** 2 * Y ** 2 - 10.0 * r2) / d ** 4, 0.0)
    Ax = np.where(inside, -20.0 * X * e / d ** 2, 0.0)
    Ay = np.where(inside, -20.0 * Y * e / d ** 2, 0.0)
    return B, Ax, Ay, h


def alg_bytes_2d(n, cin, cout, eb):
    """Algorithmic HBM bytes of the 2D pipeline (n^2 unbounded): R2C x pass,
    fused y pass (pad, FFT, Ghat*zeta*op, IFFT, truncate; real quadrant of the
    kernel), C2R x pass.  cin spectra in, cout out."""
    H = n + 1
    er, ec = eb, 2 * eb
    P1 = cin * (n * n * er + H * n * ec)
    P3 = cin * H * n * ec + H * (n + 1) * er + cout * H * n * ec
    P5 = cout * (H * n * ec + n * n * er)
    return dict(x=P1, z=P3, x_inv=P5, solve=P1 + P3 + P5)


def cpu_baseline_2d(n_target):
    """The oracle (numpy FP64) 2D grad solve timed at 1024^2 and scaled to
    n_target by n^2 log n."""
    from oracle import oracle as O
    ns = 1024
    B, _, _, h = bump2d(ns)
    reps, t = 0, 0.0
    while reps < 3 and t < 20.0:
        t0 = time.perf_counter()
        O.solve2d((ns, ns), (0, 0), (h, h), 0, [B], op="grad")
        t += time.perf_counter() - t0
        reps += 1
    per = t / reps
    scale = (n_target / ns) ** 2 * math.log2(2 * n_target) / math.log2(2 * ns)
    return {"value": 1.0 / (per * scale), "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": "oracle (numpy FP64 C2C restatement of solve2d) setup2d+push+solve2d+pull, grad, at "
                      "%d^2 UU reg0, %d reps, %.2f s/solve, scaled to %d^2 by n^2 log n (x%.0f)"
                      % (ns, reps, per, n_target, scale)}


def bench_2d(a):
    import poissonsolver_amd as P
    n = a.n if a.n != 512 else 4096
    B, Ax, Ay, h = bump2d(n)
    codes = {"fp32": P.PS_FP32, "mixed": P.PS_MIXED, "fp64": P.PS_FP64}
    precs = ["fp64", "fp32"] if not a.pmc_child else [a.prec]
    res = {}
    for prec in precs:
        s = P.poisson_solver(codes[prec])
        s.setup2d((n, n), (0, 0), (h, h), 0)
        s.set_return_grad(True)
        s.push(B)
        if s.last_status:
            return 1
        if a.warmup > 0:
            s.profile_solve(a.warmup)
        s.synchronize()
        t0 = time.perf_counter()
        prof = s.profile_solve(a.steps)
        wall = time.perf_counter() - t0
        gx, gy = np.empty_like(B), np.empty_like(B)
        s.pull(gx, gy)
        err = _rel([gx, gy], [Ax, Ay])
        res[prec] = (prof, wall, err, s.zpass_variant())
        s.close()
    prec = precs[0]
    prof, wall, err, zvar = res[prec]
    eb = 8 if prec == "fp64" else 4
    Bb = alg_bytes_2d(n, 1, 2, eb)
    achieved = Bb["z"] / (prof["z"] * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": 1000.0 / prof["solve"], "unit": "solves/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": prof["solve"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": {"fp64": "f64", "fp32": "f32", "mixed": "f32"}[prec], "data": "synthetic",
        "config": {"workload": "uu2d: %d^2 fully unbounded 2D solve (solve2d), FFT domain doubled to %d^2, reg 0, "
                               "the reference's unbounded2d bump source, grad (2 components)" % (n, 2 * n),
                   "ncell": [n, n], "bc": [0, 0], "reg_order": 0, "op": "grad", "precision": prec},
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "fused y pass (%s kernel: pad, FFT, Ghat*op for 2 outputs, IFFT, truncate)" % zvar,
            "alg_bytes_per_launch": Bb["z"], "avg_ms": prof["z"]},
        "solve_hbm_gbs": Bb["solve"] / (prof["solve"] * 1e-3) / 1e9,
        "alg_bytes_per_solve": Bb["solve"],
        "pass_ms": {k: prof[k] for k in ("x", "z", "x_inv")},
        "pass_gbs": {k: Bb[k] / (prof[k] * 1e-3) / 1e9 for k in ("x", "z", "x_inv")},
        "rel_l2_vs_analytic_grad": err, "wall_s_timed": wall,
    }
    if prof["y"] > 1e-3:
        # long pencils: the x pass output transposed so that y pencils are
        # contiguous (a layout copy, outside the algorithmic byte model)
        out["pass_ms"]["transpose"] = prof["y"]
    if len(precs) == 2:
        p32, _, e32, zv32 = res["fp32"]
        out["fp32_path"] = {"value": 1000.0 / p32["solve"], "ms_per_step": p32["solve"], "z_kernel": zv32,
                            "pass_ms": {k: p32[k] for k in ("x", "y", "z", "x_inv")},
                            "rel_l2_vs_analytic_grad": e32}
    if not (a.no_traffic or a.pmc_child):
        child = ["--pmc-child", "--workload", "uu2d", "--no-traffic", "--no-cpu", "--steps", "2", "--warmup", "1",
                 "--n", str(n), "--prec", prec]
        traffic, detail = measure_traffic(child, lambda name: "k_pencil" in name or "k_zfused_w<" in name,
                                          take_last=3)
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_detail"] = detail
    if not (a.no_cpu or a.pmc_child):
        out["cpu_baseline"] = cpu_baseline_2d(n)
    if not a.pmc_child:
        print(json.dumps(out), flush=True)
    return 0


def bench_batch(a):
    import poissonsolver_amd as P

    probs, srcs = batch_problems(a.nprob)
    prec = {"fp32": P.PS_FP32, "mixed": P.PS_MIXED, "fp64": P.PS_FP64}[a.prec]
    t0 = time.perf_counter()
    b = P.batch(probs, prec)
    setup_s = time.perf_counter() - t0
    tune_s = 0.0
    if a.autotune:
        t0 = time.perf_counter()
        b.autotune(iters=3)
        tune_s = time.perf_counter() - t0
    for i, src in enumerate(srcs):
        b.push(i, src)
    for _ in range(max(a.warmup, 2)):        # eager pass, then capture + launch
        b.solve()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        b.solve()                            # one graph launch + wait
    wall = time.perf_counter() - t0
    ms = 1000.0 * wall / a.steps
    gpu_ms = b.profile(a.steps)              # graph launches only, HIP events
    eb = 8 if a.prec == "fp64" else 4
    alg = sum(alg_bytes(p[0][0], 1, 1, eb)["solve"] for p in probs)
    out = {
        "metric": METRIC,
        "value": a.nprob * 1000.0 / ms,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": max(a.warmup, 2),
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if a.prec == "fp64" else "f32",
        "data": "synthetic",
        "config": {
            "workload": "batch of %d independent fully unbounded scalar solves, n^3 with n drawn (seed 4) from "
                        "%s, reg order 2, Gaussian-blob sources, %s, one hipGraph launch per step"
                        % (a.nprob, list(BATCH_SIZES), a.prec),
            "sizes": [p[0][0] for p in probs], "autotune": bool(a.autotune), "precision": a.prec,
        },
        "gpu_ms_per_batch": gpu_ms,
        "alg_bytes_per_batch": alg,
        "batch_hbm_gbs": alg / (gpu_ms * 1e-3) / 1e9,
        "roofline": {"bound": "hbm", "achieved": alg / (gpu_ms * 1e-3) / 1e9, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": alg / (gpu_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "whole batch: one hipGraph launch of every problem's passes (grouped solves)",
                     "alg_bytes_per_launch": alg, "avg_ms": gpu_ms},
        "setup_s": setup_s, "autotune_s": tune_s,
    }
    if not a.no_check:
        errs = []
        for i, (p, src) in enumerate(zip(probs, srcs)):
            n = p[0]
            got = np.empty((n[2], n[1], n[0]))
            b.pull(i, got)
            if a.prec == "fp64":
                continue
            r = P.poisson_solver(P.PS_FP64)
            r.setup3d(*p)
            r.push(src)
            r.solve3d()
            ref = np.empty_like(got)
            r.pull(ref)
            r.close()
            errs.append(float(np.sqrt(np.sum((got - ref) ** 2) / np.sum(ref ** 2))))
        if errs:
            out["rel_l2_vs_fp64_gpu"] = {"max": max(errs), "median": float(np.median(errs)),
                                         "per_problem": errs}
    b.close()
    print(json.dumps(out), flush=True)
    return 0


def trace_gaps(path, per_launch, launches):
    """Idle time inside and between graph launches from a rocprofv3
    kernel_trace.csv: the last `launches` x `per_launch` dispatches, grouped
    in start order into launches; for each, the span from the first kernel's
    start to the last kernel's end against the union of the kernels' busy
    intervals, and the gap to the next launch."""
    import csv
    rows = list(csv.DictReader(open(path)))
    ks = sorted(((int(r["Start_Timestamp"]), int(r["End_Timestamp"]), r["Kernel_Name"]) for r in rows
                 if "copyBuffer" not in r["Kernel_Name"] and "fillBuffer" not in r["Kernel_Name"]))
    ks = ks[-per_launch * launches:]
    spans, busys, gaps = [], [], []
    for i in range(launches):
        grp = ks[i * per_launch:(i + 1) * per_launch]
        lo, hi = grp[0][0], max(e for _, e, _ in grp)
        busy, cur_s, cur_e = 0, None, None
        for st, en, _ in grp:
            if cur_e is None or st > cur_e:
                if cur_e is not None:
                    busy += cur_e - cur_s
                cur_s, cur_e = st, en
            else:
                cur_e = max(cur_e, en)
        busy += cur_e - cur_s
        spans.append(hi - lo)
        busys.append(busy)
        if i + 1 < launches:
            gaps.append(ks[(i + 1) * per_launch][0] - hi)
    span = sum(spans) / len(spans) / 1e6
    busy = sum(busys) / len(busys) / 1e6
    gap = (sum(gaps) / len(gaps) / 1e6) if gaps else 0.0
    out = {"launches": launches, "kernels_per_launch": per_launch, "span_ms": span, "busy_ms": busy,
           "idle_in_launch_ms": span - busy, "gap_between_launches_ms": gap,
           "idle_fraction": (span - busy + gap) / (span + gap)}
    print(json.dumps(out))
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--prec", default="mixed", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--no-compare", action="store_true", help="mixed: do not time the FP32 path beside it")
    ap.add_argument("--op", default="plain", choices=["plain", "curl"])
    ap.add_argument("--bc", default="0,0,0", choices=["0,0,0", "1,1,0", "0,0,1", "1,1,1"],
                    help="boundary tuple of the default workload (0 unbounded, 1 periodic)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-host", action="store_true", help="skip the host-pointer API end-to-end timing")
    ap.add_argument("--calib-copy", action="store_true",
                    help="only run a 1 GiB device-to-device copy (HBM counter calibration) and exit")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 FETCH_SIZE/WRITE_SIZE passes that fill roofline.traffic")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--workload", default="single",
                    choices=["single", "batch", "ppu256", "vortex512", "large1024", "uu2d"])
    ap.add_argument("--nprob", type=int, default=64)
    ap.add_argument("--autotune", action="store_true", help="batch: autotune each distinct size first")
    ap.add_argument("--trace-gaps", default=None, metavar="KERNEL_TRACE_CSV",
                    help="only summarise idle time of graph launches in a rocprofv3 kernel trace")
    ap.add_argument("--per-launch", type=int, default=35, help="--trace-gaps: kernels per graph launch")
    ap.add_argument("--launches", type=int, default=10, help="--trace-gaps: last launches to summarise")
    a = ap.parse_args(argv)
    if a.trace_gaps:
        return trace_gaps(a.trace_gaps, a.per_launch, a.launches)
    if a.gpus != 1:
        print("bench.py: this build runs on one GPU only; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        return 2

    if a.calib_copy or a.pmc_child:
        calib_copy()
        if a.calib_copy:
            return 0
    if a.workload == "batch":
        return bench_batch(a)
    if a.workload == "uu2d":
        return bench_2d(a)
    if a.workload in ("ppu256", "vortex512", "large1024"):
        return bench_config(a)

    import poissonsolver_amd as P

    n = a.n
    bc = tuple(int(v) for v in a.bc.split(","))
    uuu = bc == (0, 0, 0)
    b, h = gaussian_source(n)
    cin = 3 if a.op == "curl" else 1
    cout = 3 if a.op == "curl" else 1
    srcs = [b, 0.5 * b, -0.25 * b][:cin] if a.op == "curl" else [b]
    codes = {"fp32": P.PS_FP32, "mixed": P.PS_MIXED, "fp64": P.PS_FP64}

    def run(prec):
        """Time K solves of one precision (inputs resident in HBM); returns
        (per-pass ms, wall s, solver, z kernel family)."""
        s = P.poisson_solver(codes[prec])
        s.setup3d((n,) * 3, bc, (h,) * 3, 2)
        s.set_return_curl(a.op == "curl")
        s.push(*srcs)
        if s.last_status:
            return None, 0.0, s, ""
        if a.warmup > 0:
            s.profile_solve(a.warmup)            # untimed; solves the pushed source
        s.synchronize()
        t0 = time.perf_counter()
        prof = s.profile_solve(a.steps)          # K solves, HIP events on the solver stream
        s.synchronize()
        return prof, time.perf_counter() - t0, s, s.zpass_variant()

    ref = None

    def rel_to_fp64(s):
        nonlocal ref
        lhs = [np.empty_like(b) for _ in range(cout)]
        s.pull(*lhs)
        if ref is None:
            r = P.poisson_solver(P.PS_FP64)
            r.setup3d((n,) * 3, bc, (h,) * 3, 2)
            r.set_return_curl(a.op == "curl")
            r.push(*srcs)
            r.solve3d()
            ref = [np.empty_like(b) for _ in range(cout)]
            r.pull(*ref)
            r.close()
        num = sum(float(np.sum((x - y) ** 2)) for x, y in zip(lhs, ref))
        den = sum(float(np.sum(y ** 2)) for y in ref)
        return math.sqrt(num / den), lhs

    prof, wall, s, zvar = run(a.prec)
    if prof is None:
        return 1
    ms = prof["solve"]
    eb = 8 if a.prec == "fp64" else 4
    B = alg_bytes_grid((n, n, n), bc, cin, cout, eb)
    dom = "z"
    achieved = B[dom] / (prof[dom] * 1e-3) / 1e9
    dtype = {"fp64": "f64", "fp32": "f32",
             "mixed": "mixed: z pass f16 hi/lo MFMA DFT-32 stages, f32 accumulation; f32 elsewhere"}[a.prec]
    out = {
        "metric": METRIC,
        "value": 1000.0 / ms,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {
            "workload": ("%d^3 fully unbounded %s Poisson solve, FFT domain doubled to %d^3, reg order 2, "
                         "Gaussian source, %s" % (n, "vector curl" if a.op == "curl" else "scalar", 2 * n, a.prec))
                        if uuu else
                        ("%d^3 %s %s Poisson solve (bc %s: FFT lengths %s), reg order 2, Gaussian source, %s"
                         % (n, {(1, 1, 0): "x,y periodic z unbounded", (0, 0, 1): "x,y unbounded z periodic",
                                (1, 1, 1): "fully periodic"}[bc], "vector curl" if a.op == "curl" else "scalar",
                            a.bc, "x".join(str(2 * n if v == 0 else n) for v in bc), a.prec)),
            "ncell": [n, n, n], "bc": list(bc), "reg_order": 2, "op": a.op, "precision": a.prec,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None,
            "kernel": "z pass (%s kernel, %d component%s: pad, FFT, Ghat*zeta*op, IFFT, truncate)"
                      % (zvar, cin if a.op == "curl" else 1, "s" if a.op == "curl" else ""),
            "alg_bytes_per_launch": B[dom], "avg_ms": prof[dom],
        },
        "solve_hbm_gbs": B["solve"] / (ms * 1e-3) / 1e9,
        "solve_roofline_frac": B["solve"] / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "alg_bytes_per_solve": B["solve"],
        "pass_ms": {k: prof[k] for k in ("x", "y", "z", "y_inv", "x_inv")},
        "pass_gbs": {k: B[k] / (prof[k] * 1e-3) / 1e9 for k in ("x", "y", "z", "y_inv", "x_inv")},
        "wall_s_timed": wall,
    }
    if not (a.no_traffic or a.pmc_child):
        child = ["--pmc-child", "--no-traffic", "--no-cpu", "--no-check", "--no-compare", "--steps", "2",
                 "--warmup", "1", "--n", str(n), "--prec", a.prec, "--op", a.op, "--bc", a.bc]
        nc = 3 if a.op == "curl" else 1
        traffic, detail = measure_traffic(child, lambda name: is_zpass_kernel(name, nc))
        out["roofline"]["traffic"] = traffic
        out["roofline"]["traffic_detail"] = detail
    if not a.no_check:
        from scipy.special import erf
        if a.prec != "fp64":
            err, lhs = rel_to_fp64(s)
            out["rel_l2_vs_fp64_gpu"] = err
        else:
            lhs = [np.empty_like(b) for _ in range(cout)]
            s.pull(*lhs)
        if a.op == "plain" and uuu:
            x = -1.0 + (np.arange(n) + 0.5) * h
            r2 = x[:, None, None] ** 2 + x[None, :, None] ** 2 + x[None, None, :] ** 2
            rr = np.sqrt(r2)
            se = math.sqrt(0.1 ** 2 + (2 * h) ** 2)
            A = erf(rr / (math.sqrt(2) * se)) / (4 * np.pi * rr)
            out["rel_l2_vs_analytic"] = float(np.sqrt(np.sum((lhs[0] - A) ** 2) / np.sum(A ** 2)))
    if not (a.pmc_child or a.no_host):
        # the host-pointer API end to end (not `value`): push from host FP64
        # arrays, solve, pull to host FP64 arrays; PCIe transfers and the
        # device-side FP64 <-> FP32 conversion included
        lh = [np.empty_like(b) for _ in range(cout)]
        s.push(*srcs)
        s.solve3d()
        s.pull(*lh)
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            s.push(*srcs)
            s.solve3d()
            s.pull(*lh)
        host_ms = 1000.0 * (time.perf_counter() - t0) / reps
        out["host_api"] = {"value": 1000.0 / host_ms, "unit": "solves/s", "ms_per_solve": host_ms,
                           "bytes_over_pcie": (cin + cout) * n ** 3 * 8,
                           "what": "push(host FP64) + solve3d + pull(host FP64), %d repeats" % reps}
    s.close()
    # the plain FP32 GPU FFT path, timed beside the mixed form on the same input
    if a.prec == "mixed" and not (a.no_compare or a.pmc_child):
        p32, w32, s32, zv32 = run("fp32")
        if p32 is not None:
            f = {"value": 1000.0 / p32["solve"], "unit": "solves/s", "ms_per_step": p32["solve"],
                 "z_kernel": zv32, "pass_ms": {k: p32[k] for k in ("x", "y", "z", "y_inv", "x_inv")},
                 "z_achieved_gbs": B[dom] / (p32[dom] * 1e-3) / 1e9}
            if not a.no_check:
                f["rel_l2_vs_fp64_gpu"] = rel_to_fp64(s32)[0]
            out["fp32_path"] = f
            out["speedup_vs_fp32"] = p32["solve"] / ms
            out["z_speedup_vs_fp32"] = p32[dom] / prof[dom]
        s32.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(n, cin, a.op, bc)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
