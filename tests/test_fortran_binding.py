"""The Fortran side of the drop-in, on the CPU.

* The iso_c_binding interface module links against libchanflow.so and a
  Fortran program reaches chan_version and chan_create's parameter checks,
  which read every field of chan_params (a layout mismatch between the
  Fortran derived type and the C struct shows as a wrong message).
* With the reference present, the replacement modules transform, pass1,
  pass1_linearized, pass2 and solve compile against the reference's kept
  modules runparms, gridparms, data and diff_int.
"""
import os
import shutil
import subprocess

import pytest

from helpers import PKG_DIR, chanflow

FSRC = os.path.join(PKG_DIR, "fsrc")
FLANG = "/opt/rocm/lib/llvm/bin/flang"
REF = "/root/reference/ParallelFortran-Code"

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")


def _make(*args):
    return subprocess.run(["make", "-C", FSRC, *args], capture_output=True, text=True, timeout=300)


def test_fortran_abi_check():
    chanflow().load_library()  # the library must be built
    r = _make("B=_build/abi")
    assert r.returncode == 0, r.stdout + r.stderr
    out = subprocess.run([os.path.join(FSRC, "_build", "abi", "abi_check")], capture_output=True, text=True,
                         timeout=60)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert "chanflow" in lines[0] and "gfx950" in lines[0]
    assert "nx=31 rc=1" in lines[1] and "even" in lines[1]
    assert "nz2=50 rc=1" in lines[2] and "2^a 3^b" in lines[2]
    assert "re=-1 rc=1" in lines[3] and "re/lx/lz" in lines[3]
    assert "bl_grid=1 rc=1" in lines[4] and "bl_grid" in lines[4]


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present")
def test_replacement_modules_compile_against_reference():
    r = _make("replace", f"REF={REF}", "B=_build/ref")
    assert r.returncode == 0, r.stdout + r.stderr
    for m in ("transform", "pass1", "pass1_linearized", "pass2", "solve"):
        assert os.path.exists(os.path.join(FSRC, "_build", "ref", f"{m}.o"))
    shutil.rmtree(os.path.join(FSRC, "_build", "ref"), ignore_errors=True)
