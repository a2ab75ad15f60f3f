"""The reference's own driver on the GPU.

fsrc/_build/ref/mpiflow is the reference's main program (PFC/main_mpi.f90)
with its kept modules (setup, read_write, init_vel, init_jet_buffer,
pass_data and the module-state files runparms, gridparms, data, diff_int),
compiled from the reference's sources by `make -C fsrc refdriver` (build()
does this where the reference is present) and linked with the replacement
modules transform, pass1, pass1_linearized, pass2, solve and libchanflow.so.

Here it reads a flow.dat and a savevel.out written by the test, runs the
whole program (read_parms, define_coordinates, define_parameters,
define_initial_field with remove_initial_divergence and
compute_initial_pressure through the per-mode solve routines, the timestep
loop, save_velocity) and its newvel.out after 10 steps matches the oracle's
10 RK3 steps from the same file: relative L2 <= 1e-10 per component.  The
run uses the reference's own defaults (grid stretching 2.75, initial dt
0.25 lx/nx held fixed), which the unpatched driver hard-codes."""
import os
import resource
import subprocess

import pytest

from helpers import PKG_DIR, chanflow, init_cfg0, rel_l2, write_flow_dat
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(PKG_DIR, "fsrc", "_build", "ref", "mpiflow")


def _unlimited_stack():
    # nonlinear keeps plane-sized automatic arrays on the stack (PFC/pass1.f90:29-33)
    resource.setrlimit(resource.RLIMIT_STACK, (resource.RLIM_INFINITY, resource.RLIM_INFINITY))


def run_driver(workdir, timeout=300):
    return subprocess.run([DRIVER], cwd=workdir, capture_output=True, text=True, timeout=timeout,
                          preexec_fn=_unlimited_stack)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="reference driver not built (needs the reference "
                    "sources, flang and MPICH at build time)")
def test_reference_driver_ten_steps(tmp_path):
    cf = chanflow()
    g = O.make_grid(32, 65, 32, 48, 48, 2.0, 2.0, 2500.0, stretch=2.75, fix_mass=1)
    dt = O.default_dt(g)
    U0 = init_cfg0(O, g, seed=41, amp=1e-2)
    cf.savevel_write(str(tmp_path / "savevel.out"), U0, g.lx, g.lz, g.re, dt, 0.0)
    write_flow_dat(tmp_path / "flow.dat", g.re, 2.0, 2.0, g.nx, g.ny, g.nz, g.nx2, g.nz2, tend=10,
                   output_step=10, save_step=100000, l2_step=100000, read_vel=1, save_vel=1)
    print("reference driver: 32x65x32, 10 steps", flush=True)
    r = run_driver(tmp_path)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "Normal Termination" in r.stdout, r.stdout[-3000:]
    assert "WTIME" in r.stdout
    hdr, d1 = cf.savevel_read(str(tmp_path / "newvel.out"))
    assert (hdr["nx"], hdr["ny"], hdr["nz"]) == (g.nx, g.ny, g.nz)

    st = O.initial_state(g, U0, dt=dt)
    for s in range(1, 11):
        O.rk3_step(g, st, s)
    assert abs(hdr["t_total"] - st.t_total) <= 1e-13
    assert rel_l2(d1, U0) > 1e-6            # the flow moved
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10, c
