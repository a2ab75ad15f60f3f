"""The Fortran side of the drop-in, on the CPU.

* The iso_c_binding interface module links against libchanflow.so and a
  Fortran program reaches chan_version and chan_create's parameter checks,
  which read every field of chan_params (a layout mismatch between the
  Fortran derived type and the C struct shows as a wrong message).
* With the reference present, the replacement modules transform, pass1,
  pass1_linearized, pass2 and solve compile against the reference's kept
  modules runparms, gridparms, data and diff_int.
* With the reference and MPICH present, the reference's own driver
  (PFC/main_mpi.f90 with setup, read_write, init_vel, init_jet_buffer,
  pass_data) compiles from the reference's sources in the order of their USE
  statements, with the replacement modules in between, and links against
  libchanflow.so and MPI with no unresolved symbol; run on the host (no GPU)
  it reads its flow.dat, allocates its arrays and stops with the library's
  own message at the first device call.
"""
import os
import resource
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
def test_replacement_modules_compile_against_reference(tmp_path):
    r = _make("replace", f"REF={REF}", f"B={tmp_path}")
    assert r.returncode == 0, r.stdout + r.stderr
    for m in ("transform", "pass1", "pass1_linearized", "pass2", "solve"):
        assert os.path.exists(os.path.join(tmp_path, f"{m}.o"))


MPICH = "/opt/conda"


@pytest.mark.skipif(not (os.path.isdir(REF) and os.path.exists(os.path.join(MPICH, "include", "mpif.h"))),
                    reason="reference sources or MPICH not present")
def test_reference_driver_builds_and_links(tmp_path):
    chanflow().load_library()
    b = tmp_path / "ref"
    r = _make("refdriver", f"REF={REF}", f"B={b}", f"MPIDIR={MPICH}")
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    exe = b / "mpiflow"
    assert exe.exists()
    # nothing the kept driver calls is left unresolved (FFTW is gone, the library and MPI resolve the rest)
    ldd = subprocess.run(["ldd", str(exe)], capture_output=True, text=True, timeout=60)
    assert "libchanflow.so" in ldd.stdout and "not found" not in ldd.stdout, ldd.stdout
    nm = subprocess.run(["nm", "-u", str(exe)], capture_output=True, text=True, timeout=60).stdout
    assert "dfftw_" not in nm
    if not os.path.exists("/dev/kfd"):
        # on a host without a GPU the program gets as far as the first device call
        import numpy as np
        from helpers import write_flow_dat
        cf = chanflow()
        d1 = np.zeros((3, 66, 17, 32), np.complex128)
        cf.savevel_write(str(tmp_path / "savevel.out"), d1, 2 * np.pi, np.pi, 2500.0, 0.01, 0.0)
        write_flow_dat(tmp_path / "flow.dat", 2500.0, 2.0, 2.0, 32, 65, 32, 48, 48, tend=1, output_step=1,
                       save_step=10, l2_step=10)

        def big_stack():
            resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))
        run = subprocess.run([str(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=120,
                             preexec_fn=big_stack)
        assert run.returncode != 0
        assert "Nx =   32   Nx2 =   48" in run.stdout and "DATA ARRAYS ALLOCATED" in run.stdout, run.stdout
        assert "chan_create" in run.stdout + run.stderr
