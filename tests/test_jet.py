"""The wall jet (Gaussian jet-in-crossflow with its mean removed) and its
coupling to the step.

CPU pins of the restatement (oracle/chan_oracle.py: jet_profile_spec,
jet_amplitude): the wall profile round-trips through the small transforms to
the Gaussian minus its mean (sign: blowing into the channel at either wall),
with zero net mass flux and the peak at the jet centre; the substep time of
the amplitude sums rkc + rkd to one full step at rk = 3.
GPU (-m gpu): 3 RK3 steps with a pulsed bottom jet and wall-shear feedback
(ctrlu) against the oracle, relative L2 <= 1e-10, divergence <= 1e-12."""
import math

import numpy as np
import pytest

from helpers import grid_cfg0, init_cfg0, make_channel, rel_l2
from oracle import chan_oracle as O


def _gauss(g, wall, A, r, xj, zj):
    x = np.arange(g.nx) * g.lx / g.nx
    z = np.arange(g.nz) * g.lz / g.nz
    f = np.zeros((g.nz, g.nx))
    for a in (-1, 0, 1):            # periodic images, summed where they matter
        for b in (-1, 0, 1):
            dx = x[None, :] - xj + a * g.lx
            dz = z[:, None] - zj + b * g.lz
            f = np.maximum(f, np.exp(-(dx * dx + dz * dz) / (r * r)))
    return (1.0 if wall == 0 else -1.0) * A * f


@pytest.mark.parametrize("wall", [0, 1])
def test_jet_profile_round_trip_and_mass(wall):
    g = grid_cfg0(O)
    A, r, xj, zj = 0.2, 0.3, 1.0, 0.5
    C = O.jet_profile_spec(g, wall, A, r, xj, zj)
    assert C[0, 0] == 0.0                                       # no net mass flux
    f = O.small_c2r(g, C)
    ref = _gauss(g, wall, A, r, xj, zj)
    np.testing.assert_allclose(f, ref - ref.mean(), rtol=0, atol=1e-14)
    iz, ix = np.unravel_index(np.argmax((1 if wall == 0 else -1) * f), f.shape)
    assert abs(ix * g.lx / g.nx - xj) <= g.lx / g.nx and abs(iz * g.lz / g.nz - zj) <= g.lz / g.nz


def test_jet_amplitude_substep_times():
    st = O.State(U=None, R=None, P=None, dt=0.01, p_grad=0.0, d_um_dy=None)
    st.t_total = 2.0
    s3 = sum(O.RKC[r] + O.RKD[r] for r in range(3))
    assert abs(s3 - 1.0) < 1e-7                                  # the three substeps make one step
    f, fr = 0.3, 0.7
    assert abs(O.jet_amplitude(st, 3, f, fr) - (1 + f * math.sin(2 * math.pi * fr * (2.0 + 0.01 * s3)))) < 1e-15
    assert O.jet_amplitude(st, 1, 0.0, fr) == 1.0


@pytest.mark.gpu
def test_gpu_jet_with_feedback_steps():
    g = grid_cfg0(O, ctrlu=0.5)
    A, r, xj, zj, forcing, freq = 0.2, 0.3, 1.0, 0.5, 0.25, 0.8
    ch = make_channel(g)
    U0 = init_cfg0(O, g, seed=17, amp=1e-2)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.01)
    ch.jet_configure(0, A, r, xj, zj, forcing, freq)
    ch.initialize_flow()
    st = O.initial_state(g, U0, dt=0.01)
    ch.step(3, step0=1)
    jet = (0, O.jet_profile_spec(g, 0, A, r, xj, zj), forcing, freq)
    for s in (1, 2, 3):
        O.rk3_step(g, st, s, jet=jet)
    d1 = ch.get_data1()
    assert rel_l2(d1, st.U) < 1e-10
    assert ch.divergence_max() <= 1e-12
    # the jet is there: v at the first interior face follows the wall blowing
    assert np.abs(d1[1, 0]).max() > 1e-3
    ch.close()


@pytest.mark.gpu
def test_gpu_jet_control_config_64x97x64():
    """The jet-control run's own settings: 64x97x64 (96x96 dealiased, box
    2 pi x 2 x pi, stretching 2.75), Re_c = 2800, a steady Gaussian jet of
    radius 0.1 and peak wall-normal velocity 0.2 at the lower wall, and the
    wall-shear feedback law of boundary (PFC/read_write.f90:1465-1532) with
    gain ctrlu = ctrlw = 0.5; 3 RK3 steps at dt 0.002 from a perturbed
    laminar profile against the oracle: relative L2 <= 1e-10, divergence
    <= 1e-12."""
    g = O.make_grid(64, 97, 64, 96, 96, 2.0, 2.0, 2800.0, stretch=2.75, fix_mass=1, ctrlu=0.5, ctrlw=0.5)
    A, r, xj, zj = 0.2, 0.1, 0.25 * g.lx, 0.5 * g.lz
    dt = 0.002
    s = __import__("helpers").synth()
    U0 = s.laminar_perturbed(g.nx, g.ny, g.nz, g.yh, g.y, seed=23, amp=1e-2, kmax=4.0)
    ch = make_channel(g)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=dt)
    ch.jet_configure(0, A, r, xj, zj, 0.0, 0.0)
    ch.initialize_flow()
    st = O.initial_state(g, U0, dt=dt)
    ch.step(3, step0=1)
    jet = (0, O.jet_profile_spec(g, 0, A, r, xj, zj), 0.0, 0.0)
    for k in (1, 2, 3):
        O.rk3_step(g, st, k, jet=jet)
    d1 = ch.get_data1()
    assert rel_l2(d1, st.U) < 1e-10
    assert ch.divergence_max() <= 1e-12
    vwall = O.small_c2r(g, d1[1, 0][None])[0]     # wall-normal velocity on the lower wall
    assert 0.19 < vwall.max() <= 0.2              # the jet's peak (less the uniform suction)
    ch.close()
