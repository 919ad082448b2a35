"""The C++ drop-in header compiles and links against the library like the
reference's test programs (CPU); the programs run and meet the analytic
errors of the reference's unbounded3d, unbounded2d, periodic2d and mixed2d
tests (GPU)."""
import os
import shutil
import subprocess
import tempfile
import xml.etree.ElementTree as ET

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "poissonsolver_amd")
PROGRAMS = {"3d": "dropin_unbounded3d.cpp", "2d": "dropin_2d.cpp"}


def build_exe(which):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("g++ not available")
    if not os.path.exists(os.path.join(LIBDIR, "libpoissonsolver_amd.so")):
        pytest.skip("library not built")
    src = os.path.join(ROOT, "tests", "cpp", PROGRAMS[which])
    exe = os.path.join(tempfile.gettempdir(), "ps_dropin_" + which)
    cmd = [cxx, "-std=c++11", "-O2", "-I", os.path.join(ROOT, "include"), src, "-o", exe,
           "-L", LIBDIR, "-lpoissonsolver_amd", "-Wl,-rpath," + LIBDIR]
    p = subprocess.run(cmd, capture_output=True, text=True)
    assert p.returncode == 0, p.stderr
    return exe


@pytest.mark.parametrize("which", sorted(PROGRAMS))
def test_dropin_header_compiles_and_links(which):
    build_exe(which)


@pytest.mark.gpu
def test_dropin_program_runs_on_gpu():
    exe = build_exe("3d")
    p = subprocess.run([exe, "32"], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, (p.stdout, p.stderr)
    assert "Error:" in p.stdout


@pytest.mark.gpu
def test_dropin_2d_programs_run_on_gpu(tmp_path):
    """unbounded2d, periodic2d and mixed2d at their own sizes, each below 1e-12
    against the analytic field; the unbounded case writes mesh_P00.vti and
    mesh.pvti."""
    exe = build_exe("2d")
    p = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=120)
    print(p.stdout, p.stderr)
    assert p.returncode == 0, (p.stdout, p.stderr)
    assert p.stdout.count("Error:") == 3
    root = ET.parse(tmp_path / "mesh_P00.vti").getroot()
    assert root.find("ImageData").get("WholeExtent").split() == ["0", "128", "0", "128", "0", "1"]
    assert (tmp_path / "mesh.pvti").exists()
