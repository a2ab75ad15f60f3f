"""VTK output in output_hdf_dfsd's format (PFC/read_write.f90:296-500,
xskip = yskip = zskip = 1), host side: the file is rebuilt here byte by
byte from the reference's write statements (stream access, big-endian,
CHARACTER(30) header variables with an integer field written into
columns 12-26 (3I5), 15-17 (I3) or 12-20 (I9)) and compared with the
library's output."""
import os
import struct

import numpy as np

from helpers import chanflow



def _c30(text, a, b, fmt_vals):
    s = list(text.ljust(30))
    f = "".join(fmt_vals)
    s[a - 1:b] = list(f)
    return ("".join(s)).encode() + b"\n"


def _expected(nx, ny, nz, lx, lz, yh, phys):
    be = lambda vals: b"".join(struct.pack(">f", float(np.float32(v))) for v in vals)
    out = b"# vtk DataFile Version 2.0\n"
    out += b"(This is header line) vtk rectilinear grid with binary data\n"
    out += b"BINARY\n"
    out += b"DATASET RECTILINEAR_GRID\n"
    out += _c30("DIMENSIONS", 12, 26, [f"{nx:5d}", f"{ny + 1:5d}", f"{nz:5d}"])
    out += _c30("X_COORDINATES     float", 15, 17, [f"{nx:3d}"])
    out += be([lx * x / nx for x in range(nx)]) + b"\n"
    out += _c30("Y_COORDINATES     float", 15, 17, [f"{ny + 1:3d}"])
    out += be(yh) + b"\n"
    out += _c30("Z_COORDINATES     float", 15, 17, [f"{nz:3d}"])
    out += be([lz * z / nz for z in range(nz)]) + b"\n"
    out += _c30("POINT_DATA ", 12, 20, [f"{nx * (ny + 1) * nz:9d}"])
    out += b"VECTORS flowvec float\n"
    vals = []
    for z in range(nz):
        for j in range(ny + 1):
            for x in range(nx):
                for c in range(3):
                    vals.append(phys[c, j, z, x])
    out += be(vals) + b"\n"
    return out


def test_vtk_bytes_match_reference_format(tmp_path):
    cf = chanflow()
    nx, ny, nz = 8, 6, 4
    rng = np.random.default_rng(0)
    phys = rng.standard_normal((3, ny + 1, nz, nx))
    yh = np.linspace(-1, 1, ny + 1)
    path = os.path.join(tmp_path, "t_001.vtk")
    cf.vtk_write(path, phys, 2.5, 1.25, yh)
    got = open(path, "rb").read()
    assert got == _expected(nx, ny, nz, 2.5, 1.25, yh, phys)
    assert got.startswith(b"# vtk DataFile Version 2.0\n")
    assert b"DIMENSIONS     8    7    4    \n" in got
