"""Device diagnostics through the C-ABI against the oracle: l2_norm and the
wall drag, per-mode kinetic energy and per-plane velocity moments.
Tolerance: rel 1e-12 for l2 and the energies, rel 1e-10 with an absolute
floor of 1e-12 of the row scale for the moments (reductions in another order)."""
import numpy as np
import pytest

from helpers import grid_cfg0, init_cfg0, make_channel, rel_l2
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu


def test_diagnostics_match_oracle():
    g = grid_cfg0(O)
    ch = make_channel(g)
    U0 = init_cfg0(O, g, seed=9, amp=5e-2)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.01, t_total=0.0, p_grad=-2.0 / g.re)
    ch.initialize_flow()
    st = O.initial_state(g, U0, dt=0.01)
    ch.step(2, step0=1)
    for s in (1, 2):
        O.rk3_step(g, st, s)
    U = ch.get_data1()
    l2, drag = ch.l2_norm()
    assert abs(l2 - O.l2_norm(g, U)) <= 1e-12 * O.l2_norm(g, U)
    assert abs(drag - (st.d_um_dy[0] - st.d_um_dy[-1])) <= 1e-10 * abs(st.d_um_dy[0])
    assert rel_l2(ch.mode_energy(), O.mode_energy(g, U)) <= 1e-12
    M = ch.plane_moments()
    R = O.plane_moments(g, U)
    # rows of one kind (means, 2nd, 3rd, 4th moments, fluxes) share a scale;
    # the mean of v is zero up to round-off, so an absolute floor is used
    for k0 in range(0, 15, 3):
        scale = np.abs(R[k0:k0 + 3]).max()
        np.testing.assert_allclose(M[k0:k0 + 3], R[k0:k0 + 3], rtol=1e-10, atol=1e-12 * scale)
    ch.close()
