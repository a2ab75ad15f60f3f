"""Velocity checkpoints in the reference's savevel.out / newvel.out format
(PFC/read_write.f90:21-292), host side (no GPU).

The format is pinned three ways: the file's bytes are parsed here
independently of the library (4-byte record markers, the header record,
one slice record per x-mode in data1(1:nz, x, 1:ny+1, 1:3) order); a file
written by a Fortran program through the Fortran runtime's own unformatted
I/O is read by the library; and a file written by the library is read back
by that Fortran program.
"""
import os
import struct
import subprocess

import numpy as np
import pytest

from helpers import ROOT, chanflow

FLANG = "/opt/rocm/lib/llvm/bin/flang"
NX, NY, NZ = 8, 5, 6


def _field(nx, ny, nz, seed=0):
    rng = np.random.default_rng(seed)
    shape = (3, ny + 1, nx // 2 + 1, nz)
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def _records(path):
    raw = open(path, "rb").read()
    out, i = [], 0
    while i < len(raw):
        (n,) = struct.unpack_from("<i", raw, i)
        body = raw[i + 4:i + 4 + n]
        (m,) = struct.unpack_from("<i", raw, i + 4 + n)
        assert m == n
        out.append(body)
        i += 8 + n
    return out


def _reference_field(nx, ny, nz):
    """The deterministic field of tests/fortran/ckpt_io.f90 (1-based indices)."""
    d1 = np.empty((3, ny + 1, nx // 2 + 1, nz), np.complex128)
    for c in range(1, 4):
        for j in range(1, ny + 2):
            for x in range(1, nx // 2 + 2):
                for k in range(1, nz + 1):
                    d1[c - 1, j - 1, x - 1, k - 1] = complex(1000 * c + 100 * j + 10 * x + k, -0.5 * (k + x + j + c))
    return d1


@pytest.fixture(scope="module")
def work(tmp_path_factory):
    return str(tmp_path_factory.mktemp("checkpoint"))


def test_roundtrip_and_byte_layout(work):
    cf = chanflow()
    d1 = _field(NX, NY, NZ)
    path = os.path.join(work, "rt.out")
    cf.savevel_write(path, d1, lx=6.5, lz=3.25, re=2500.0, dt=0.01, t_total=3.5)
    hdr, back = cf.savevel_read(path)
    assert hdr == dict(nx=NX, ny=NY, nz=NZ, lx=6.5, lz=3.25, re=2500.0, dt=0.01, t_total=3.5)
    assert np.array_equal(back, d1)
    recs = _records(path)
    assert len(recs) == 1 + NX // 2 + 1
    assert struct.unpack("<3i5d", recs[0]) == (NX, NY, NZ, 6.5, 3.25, 2500.0, 0.01, 3.5)
    for x in range(NX // 2 + 1):
        s = np.frombuffer(recs[1 + x], np.complex128).reshape(3, NY + 1, NZ)  # slice(nz, ny+1, 3)
        assert np.array_equal(s, d1[:, :, x, :])


def test_read_errors(work):
    cf = chanflow()
    path = os.path.join(work, "short.out")
    cf.savevel_write(path, _field(NX, NY, NZ), 1.0, 1.0, 100.0, 0.1, 0.0)
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-10])
    with pytest.raises(cf.ChanError, match="record|end of file"):
        cf.savevel_read(path)
    lib = cf.load_library()
    good = os.path.join(work, "good.out")
    cf.savevel_write(good, _field(NX, NY, NZ), 1.0, 1.0, 100.0, 0.1, 0.0)
    out = np.empty((3, NY + 1, NX // 2 + 1, NZ + 2), np.complex128)
    assert lib.chan_savevel_read(os.fsencode(good), NX, NY, NZ + 2, cf._ptr(out)) != 0
    assert "INCONSISTENT PARAMETERS" in lib.chan_io_last_error().decode()
    with pytest.raises(cf.ChanError, match="cannot open"):
        cf.savevel_read(os.path.join(work, "missing.out"))


@pytest.fixture(scope="module")
def fortran_prog(work):
    if not os.path.exists(FLANG):
        pytest.skip("flang not installed")
    exe = os.path.join(work, "ckpt_io")
    src = os.path.join(ROOT, "tests", "fortran", "ckpt_io.f90")
    r = subprocess.run([FLANG, "-O1", "-J", work, src, "-o", exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def test_fortran_written_file_is_read(work, fortran_prog):
    cf = chanflow()
    path = os.path.join(work, "from_fortran.out")
    r = subprocess.run([fortran_prog, "write", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    hdr, d1 = cf.savevel_read(path)
    assert (hdr["nx"], hdr["ny"], hdr["nz"]) == (NX, NY, NZ)
    assert (hdr["lx"], hdr["lz"], hdr["re"], hdr["dt"], hdr["t_total"]) == (
        6.283185307179586, 3.141592653589793, 2500.0, 0.01, 1.25)
    assert np.array_equal(d1, _reference_field(NX, NY, NZ))


def test_library_written_file_is_read_by_fortran(work, fortran_prog):
    cf = chanflow()
    d1 = _field(NX, NY, NZ, seed=3)
    path = os.path.join(work, "from_lib.out")
    cf.savevel_write(path, d1, 2.0, 1.0, 1234.5, 0.02, 7.0)
    r = subprocess.run([fortran_prog, "read", path], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = r.stdout.split("\n")
    assert [int(v) for v in lines[0].split()] == [NX, NY, NZ]
    assert np.allclose([float(v) for v in lines[1].split()], [2.0, 1.0, 1234.5, 0.02, 7.0], rtol=1e-15)
    w = np.arange(1, d1.size + 1, dtype=np.float64)
    flat = d1.reshape(-1)  # Fortran order data1(nz, nxh, ny+1, 3) == C order [3][ny+1][nxh][nz]
    s1, s2 = (float(v) for v in lines[2].split())
    assert abs(s1 - np.dot(w, flat.real)) <= 1e-12 * np.abs(w * flat.real).sum()
    assert abs(s2 - np.dot(w, flat.imag)) <= 1e-12 * np.abs(w * flat.imag).sum()
