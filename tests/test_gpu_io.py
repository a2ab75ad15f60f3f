"""Device checkpoints through the C-ABI: SAVE_VELOCITY of the device state
and read_velocity into a fresh context (PFC/read_write.f90:21-292)."""

import numpy as np
import pytest

from helpers import grid_cfg0, init_cfg0, make_channel
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu


def test_save_and_read_velocity(tmp_path):
    g = grid_cfg0(O)
    a = make_channel(g)
    U0 = init_cfg0(O, g, seed=2, amp=1e-2)
    a.set_state(data1=U0)
    a.set_scalars(dt=0.01, t_total=0.0, p_grad=-2.0 / g.re)
    a.initialize_flow()
    a.step(3, step0=1)
    path = str(tmp_path / "newvel.out")
    a.save_velocity(path)
    b = make_channel(g)
    b.read_velocity(path)
    assert np.array_equal(b.get_data1(), a.get_data1())
    sa, sb = a.get_scalars(), b.get_scalars()
    assert sb["t_total"] == sa["t_total"]
    # a run with another box length is refused
    c = make_channel(O.make_grid(32, 65, 32, 48, 48, 2.5, 2.0, 2500.0, stretch=1.5))
    with pytest.raises(Exception, match="INCONSISTENT PARAMETERS"):
        c.read_velocity(path)
    for ch in (a, b, c):
        ch.close()


def test_write_vtk_from_device(tmp_path):
    import struct
    g = O.make_grid(16, 17, 12, 24, 18, 2.0, 2.0, 1000.0, stretch=1.5)
    ch = make_channel(g)
    U = O.remove_initial_divergence(g, init_cfg0(O, g, seed=4, amp=1e-1))
    ch.set_state(data1=U)
    path = str(tmp_path / "t_000.vtk")
    ch.write_vtk(path, g.yh)
    raw = open(path, "rb").read()
    i = raw.index(b"VECTORS flowvec float\n") + len(b"VECTORS flowvec float\n")
    n = g.nx * (g.ny + 1) * g.nz * 3
    got = np.array(struct.unpack(f">{n}f", raw[i:i + 4 * n])).reshape(g.nz, g.ny + 1, g.nx, 3)
    # expected: u, w at the planes, v averaged to yh, small transform
    F = U.copy()
    F[1, 0] = U[1, 0]
    F[1, 1:g.ny] = 0.5 * (U[1, 0:g.ny - 1] + U[1, 1:g.ny])
    F[1, g.ny] = U[1, g.ny - 1]
    phys = O.small_c2r(g, F.reshape(-1, g.nxh, g.nz)).reshape(3, g.ny + 1, g.nz, g.nx)
    exp = np.transpose(phys, (2, 1, 3, 0)).astype(np.float32)
    assert np.abs(got - exp).max() <= 2.0 ** -22 * max(1.0, np.abs(exp).max())
    ch.close()
