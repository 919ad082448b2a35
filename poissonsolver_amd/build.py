"""Builds poissonsolver_amd/libpoissonsolver_amd.so for gfx950 with hipcc.

Every translation unit is compiled in parallel (one kern_e.hip object per
elements-per-thread value), then linked into one shared library in-tree.
Usage:  python -m poissonsolver_amd.build  [-j N] [--force]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libpoissonsolver_amd.so")
ARCH = os.environ.get("PS_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
    # no packed FP32 instructions: hipcc -O3 would otherwise form v_pk_* ops
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # lets x + 0.0 fold away where zero padding feeds the first butterflies
    "-fno-signed-zeros",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def e_values():
    txt = open(os.path.join(CSRC, "ps_launch.hpp")).read()
    m = re.search(r"#define PS_FOR_EACH_E\(X\)(.*)", txt)
    return [int(v) for v in re.findall(r"X\((\d+)\)", m.group(1))]


def sl_values():
    txt = open(os.path.join(CSRC, "ps_static.hpp")).read()
    m = re.search(r"#define PS_FOR_EACH_SL\(X\)(.*)", txt)
    return [int(v) for v in re.findall(r"X\((\d+)\)", m.group(1))]


def sources():
    units = []
    for e in e_values():
        units.append(("kern_e.hip", ["-DPS_E=%d" % e], "kern_e%d.o" % e))
    for L in sl_values():
        units.append(("skern.hip", ["-DPS_SL=%d" % L], "skern_%d.o" % L))
    units.append(("mkern.hip", [], "mkern.o"))
    for f in ("ps_greens.hip", "ps_io.hip", "ps_dispatch.cpp", "ps_plan.cpp", "ps_capi.cpp", "ps_batch.cpp", "poisson_solver.cpp",
              "ps_vtk.cpp"):
        units.append((f, [], os.path.splitext(f)[0] + ".o"))
    return units


_INC = re.compile(r'^\s*#\s*include\s+"([^"]+)"', re.M)


def _deps_mtime(path, seen=None):
    """Newest modification time of a source and the project headers it
    includes (quoted includes, followed recursively)."""
    seen = set() if seen is None else seen
    path = os.path.normpath(path)
    if path in seen or not os.path.exists(path):
        return 0.0
    seen.add(path)
    newest = os.path.getmtime(path)
    for inc in _INC.findall(open(path).read()):
        newest = max(newest, _deps_mtime(os.path.join(os.path.dirname(path), inc), seen))
    return newest


def _compile(unit, force, newest):
    src, defs, obj = unit
    out = os.path.join(BUILD, obj)
    newest = max(newest, _deps_mtime(os.path.join(CSRC, src)))
    if not force and os.path.exists(out) and os.path.getmtime(out) >= newest:
        return out, ""
    lang = [] if src.endswith(".hip") else ["-x", "hip"]
    cmd = [hipcc()] + COMMON + defs + lang + ["-c", os.path.join(CSRC, src), "-o", out]
    p = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "is not a recognized feature for this target" not in l)
    if p.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), log))
    return out, log


def build(jobs=None, force=False, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    newest = os.path.getmtime(os.path.abspath(__file__))   # compile flags live here
    units = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    objs = []
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for out, log in ex.map(lambda u: _compile(u, force, newest), units):
            objs.append(out)
            if verbose and log.strip():
                print(log)
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", LIB + ".tmp"] + objs
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), p.stdout + p.stderr))
    os.replace(LIB + ".tmp", LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.j, a.force, a.v))


if __name__ == "__main__":
    sys.exit(main())
