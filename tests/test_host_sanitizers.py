"""The library's host code built with AddressSanitizer and
UndefinedBehaviorSanitizer (host compilation only: each -fsanitize= follows
-Xarch_host), linked with the ordinary device objects of the in-tree build,
and run on a driver of the host paths (tests/cpp/host_sanitize.cpp) on the
CPU.  Any sanitizer report fails the run."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "poissonsolver_amd", "csrc")
BUILD = os.path.join(ROOT, "poissonsolver_amd", "_build")
HOST_UNITS = ["ps_capi.cpp", "ps_batch.cpp", "poisson_solver.cpp", "ps_plan.cpp", "ps_vtk.cpp"]
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=all", "-Xarch_host", "-fno-omit-frame-pointer"]


def test_host_code_under_asan_and_ubsan(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc"
    clangxx = "/opt/rocm/llvm/bin/clang++"
    if not (os.path.exists(hipcc) and os.path.exists(clangxx)):
        pytest.skip("ROCm compilers not available")
    dev_objs = [os.path.join(BUILD, f) for f in sorted(os.listdir(BUILD))] if os.path.isdir(BUILD) else []
    dev_objs = [o for o in dev_objs if o.endswith(".o") and os.path.basename(o)[:-2] + ".cpp" not in HOST_UNITS]
    if not any(os.path.basename(o).startswith("mkern") for o in dev_objs):
        pytest.skip("library not built (python -m poissonsolver_amd.build)")
    from poissonsolver_amd.build import COMMON
    objs = []
    for u in HOST_UNITS:
        o = str(tmp_path / (u + ".o"))
        cmd = [hipcc] + COMMON + ["-O1", "-g"] + SAN + ["-x", "hip", "-c", os.path.join(CSRC, u), "-o", o]
        p = subprocess.run(cmd, capture_output=True, text=True)
        assert p.returncode == 0, p.stderr[-3000:]
        objs.append(o)
    drv = str(tmp_path / "driver.o")
    p = subprocess.run([clangxx, "-std=c++17", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                        "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                        "-D__HIP_PLATFORM_AMD__", "-c", os.path.join(ROOT, "tests", "cpp", "host_sanitize.cpp"),
                        "-o", drv], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    exe = str(tmp_path / "host_sanitize")
    p = subprocess.run([clangxx, "-fsanitize=address,undefined", "-o", exe, drv] + objs + dev_objs +
                       ["-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"], capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    p = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=300)
    out = p.stdout + p.stderr
    assert p.returncode == 0 and "SANITIZE_OK" in p.stdout, out[-4000:]
    assert "runtime error" not in out and "AddressSanitizer" not in out, out[-4000:]
