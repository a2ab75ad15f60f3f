"""Device checkpoints through the C-ABI: SAVE_VELOCITY of the device state
and read_velocity into a fresh context (PFC/read_write.f90:21-292)."""
import os

import numpy as np
import pytest

from helpers import ROOT, grid_cfg0, init_cfg0, make_channel
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu
WORK = os.path.join(ROOT, "gpurun_out", "test_gpu_io")


def test_save_and_read_velocity():
    os.makedirs(WORK, exist_ok=True)
    g = grid_cfg0(O)
    a = make_channel(g)
    U0 = init_cfg0(O, g, seed=2, amp=1e-2)
    a.set_state(data1=U0)
    a.set_scalars(dt=0.01, t_total=0.0, p_grad=-2.0 / g.re)
    a.initialize_flow()
    a.step(3, step0=1)
    path = os.path.join(WORK, "newvel.out")
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
