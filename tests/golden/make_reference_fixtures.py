"""Extract numeric fixtures from the reference's own data files.

Reads (as text, nothing is executed) PFC/makegrid.m -- the grid dump written
by output_grids (PFC/setup.f90:1331-1359) for the 128x200x64, Lx=6pi, Lz=2pi
run -- and the initial-dt line of the run log PFC/diagnostics.txt:98, and
writes tests/golden/reference_fixtures.npz.  Run in a checkout that has the
reference mounted at /root/reference.
"""
import os
import re

import numpy as np

REF = "/root/reference/ParallelFortran-Code"
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    txt = open(os.path.join(REF, "makegrid.m")).read()

    def block(name):
        m = re.search(r"\b" + name + r"\s*=\s*\[(.*?)\];", txt, re.S)
        return np.array([float(v) for v in m.group(1).split()])

    diag = open(os.path.join(REF, "diagnostics.txt")).read().splitlines()
    m = re.search(r"INITIAL DT =\s*([0-9.E+-]+)", diag[97])
    dt0 = float(m.group(1))
    np.savez(os.path.join(HERE, "reference_fixtures.npz"),
             makegrid_y=block("y"), makegrid_yh=block("yh"), makegrid_x=block("x"), makegrid_z=block("z"),
             makegrid_case=np.array([128, 200, 64, 6.0, 3.0]),   # nx, ny, nz, lx/pi, lx/lz (PFC/diagnostics.txt:7-10)
             diagnostics_dt0=np.array(dt0))


if __name__ == "__main__":
    main()
