"""The Fortran boundary: include/phlx_iface.F90 binds entry points that
include/phlx.h declares (CPU: compiled with amdflang, names checked), and a
Fortran time_loop-shaped caller (phlegethon_amd/fortran/phlx_demo.F90)
builds against libphlx.so and runs on the GPU (mass conserved, PASS)."""
import os
import re
import shutil
import subprocess

import pytest

from phlegethon_amd.runtime import LIB_PATH

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
IFACE = os.path.join(ROOT, "include", "phlx_iface.F90")
DEMO = os.path.join(ROOT, "phlegethon_amd", "fortran", "phlx_demo.F90")


def _flang():
    for c in ("/opt/rocm/llvm/bin/amdflang", shutil.which("amdflang") or ""):
        if c and os.path.exists(c):
            return c
    return None


def test_iface_names_are_declared():
    hdr = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "phlx.h")).read(),
                 flags=re.S)
    declared = set(re.findall(r"\b(phlx_[a-z0-9_]+)\s*\(", hdr))
    bound = set(re.findall(r'bind\(c,\s*name="(phlx_[a-z0-9_]+)"\)', open(IFACE).read()))
    assert bound and bound <= declared, sorted(bound - declared)


def test_iface_compiles(tmp_path):
    fc = _flang()
    if fc is None:
        pytest.skip("amdflang not available")
    r = subprocess.run([fc, "-c", IFACE, "-o", str(tmp_path / "i.o")], cwd=tmp_path,
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_fortran_demo_runs(tmp_path):
    fc = _flang()
    if fc is None:
        pytest.skip("amdflang not available")
    libdir = os.path.dirname(LIB_PATH)
    exe = tmp_path / "phlx_demo"
    r = subprocess.run([fc, IFACE, DEMO, "-o", str(exe), f"-L{libdir}", "-lphlx",
                        f"-Wl,-rpath,{libdir}"], cwd=tmp_path, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    print("running the Fortran caller", flush=True)
    r = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert "PASS" in r.stdout
