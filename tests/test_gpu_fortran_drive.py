"""The Fortran host path on the GPU: fsrc/drive_steps (built by build() next
to the ABI check program) creates the 32 x 65 x 32 channel through the
iso_c_binding module, builds the operators on the host side of the library
(chan_build_operators, stretch 1.5), and runs the reference's timestep loop
(boundary, then three nonlinear / rhs_solve substeps per step) on data1 in
the reference's Fortran layout.  Its data1 after 5 steps matches the
oracle's RK3 steps from the same start to relative L2 <= 1e-10 per
component."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import PKG_DIR, grid_cfg0, init_cfg0, rel_l2
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu

DRIVE = os.path.join(PKG_DIR, "fsrc", "_build", "abi", "drive_steps")


@pytest.mark.skipif(not os.path.exists(DRIVE), reason="drive_steps not built (needs flang)")
def test_fortran_drive_five_steps(tmp_path):
    g = grid_cfg0(O)
    U0 = init_cfg0(O, g, seed=31, amp=0.05)
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    # data1(nz, nx/2+1, ny+1, 3) column-major is the C-order array [3][ny+1][nx/2+1][nz]
    np.ascontiguousarray(U0, dtype=np.complex128).tofile(fin)
    dt, nsteps = 0.01, 5
    print("drive_steps: 32x65x32, 5 steps, dt 0.01", flush=True)
    sys.stdout.flush()
    r = subprocess.run([DRIVE, str(fin), str(fout), str(nsteps), repr(dt)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.startswith("steps=5"), r.stdout
    d1 = np.fromfile(fout, dtype=np.complex128).reshape(U0.shape)

    st = O.initial_state(g, U0, dt=dt)
    for s in range(1, nsteps + 1):
        O.rk3_step(g, st, s)
    assert rel_l2(d1, U0) > 1e-6          # the flow moved
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10, c
