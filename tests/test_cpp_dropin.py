"""The C++ drop-in header compiles and links against the library like the
reference's test programs (CPU); the program runs and meets the analytic
error of the reference's unbounded3d test (GPU)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_unbounded3d.cpp")
LIBDIR = os.path.join(ROOT, "poissonsolver_amd")
EXE = os.path.join("/tmp", "ps_dropin_unbounded3d")


def build_exe():
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIBDIR, "libpoissonsolver_amd.so")):
        pytest.skip("library not built")
    cmd = [cxx, "-std=c++11", "-O2", "-I", os.path.join(ROOT, "include"), SRC, "-o", EXE,
           "-L", LIBDIR, "-lpoissonsolver_amd", "-Wl,-rpath," + LIBDIR]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return EXE


def test_dropin_header_compiles_and_links():
    build_exe()


@pytest.mark.gpu
def test_dropin_program_runs_on_gpu():
    exe = build_exe()
    p = subprocess.run([exe, "32"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, (p.stdout, p.stderr)
    assert "Error:" in p.stdout
