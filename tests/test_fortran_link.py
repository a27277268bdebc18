"""The Fortran boundary: include/specfwat_hip.f90 (the iso_c_binding module a
SpecFWAT maintainer adds) compiled with amdflang, a program using it linked
against libspecfwat_hip.so, and a host-only entry point called through it
(tests/fortran/link_check.f90; no GPU needed)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or ("/opt/rocm/bin/amdflang" if os.path.exists("/opt/rocm/bin/amdflang") else None)


@pytest.mark.skipif(FLANG is None, reason="amdflang is not installed")
def test_fortran_module_links_and_calls_the_library(tmp_path):
    from specfwat_amd import _lib
    libdir = os.path.dirname(_lib.LIB_PATH)
    run = lambda cmd: subprocess.run(cmd, cwd=tmp_path, capture_output=True, text=True, timeout=300)
    r = run([FLANG, "-c", os.path.join(ROOT, "include", "specfwat_hip.f90"), "-o", "module.o"])
    assert r.returncode == 0, r.stderr
    r = run([FLANG, os.path.join(ROOT, "tests", "fortran", "link_check.f90"), "module.o", f"-L{libdir}",
             "-lspecfwat_hip", f"-Wl,-rpath,{libdir}", "-o", "link_check"])
    assert r.returncode == 0, r.stderr
    r = run([str(tmp_path / "link_check")])
    assert r.returncode == 0 and r.stdout.strip() == "OK", (r.returncode, r.stdout, r.stderr)
