"""The configurations' own run lengths on the GPU, held to the invariants the
scheme fixes (and, where the flow is still near laminar, to the oracle):

  * configuration 0 (32x65x32, Re_c 2500, dt 0.01): its 100 steps, against
    the oracle's 100 steps;
  * configuration 1 (the headline 192x129x192 turbulent channel, dt 0.005):
    its 500 steps;
  * configuration 3 (the 48-member batched ensemble): its 200 steps;
  * configuration 4 (the 64x65x64 learned-closure LES, dt 0.01): its 2000
    steps.

The invariants: a finite state, discrete divergence <= 1e-12 (PFC/pass2.f90:
380-401), t_total = steps x dt, and the bulk mass flux held by the
constant-mass-flux pressure gradient (PFC/pass2.f90:356-375).  That gradient
cancels the explicit part of each substep's flux change, not the implicit
part's change in wall shear (the identity pinned in test_oracle_rhs_pins), so
the flux drifts slowly by the scheme's design: relative change <= 1e-5 over
the run (2e-6 for the 2000 LES steps; without the gradient it changes by
percents)."""
import os

import numpy as np
import pytest

from helpers import chanflow, grid_cfg0, init_cfg0, make_channel, rel_l2, synth
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu


def _invariants(ch, q0, t_expect, what):
    d1 = ch.get_data1()
    assert np.all(np.isfinite(d1.real)) and np.all(np.isfinite(d1.imag)), what
    assert ch.divergence_max() <= 1e-12, what
    assert abs(ch.mass_flux() - q0) <= 1e-5 * abs(q0), (what, ch.mass_flux(), q0)
    assert abs(ch.get_scalars()["t_total"] - t_expect) <= 1e-9 * t_expect, what
    return d1


def test_config0_hundred_steps_against_oracle():
    g = grid_cfg0(O)
    U0 = init_cfg0(O, g, seed=7, amp=1e-3)
    ch = make_channel(g)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.01, t_total=0.0, p_grad=-2.0 / g.re)
    ch.initialize_flow()
    q0 = ch.mass_flux()
    print("config 0: 100 steps", flush=True)
    ch.step(100, step0=1)
    d1 = _invariants(ch, q0, 1.0, "config 0")
    ch.close()
    st = O.initial_state(g, U0, dt=0.01)
    for s in range(1, 101):
        O.rk3_step(g, st, s)
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10, c


def test_config1_five_hundred_steps():
    g = O.make_grid(192, 129, 192, 288, 288, 4.0, 3.0, 3300.0, stretch=2.75, fix_mass=1)
    U0 = synth().turbulent_like(g.nx, g.ny, g.nz, g.yh, g.y, seed=0, rms=0.1, re_tau=180.0)
    ch = make_channel(g)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.005, t_total=0.0, p_grad=-2.0 / g.re)
    ch.initialize_flow()
    q0 = ch.mass_flux()
    print("config 1: 500 steps at 192x129x192", flush=True)
    ch.step(500, step0=1)
    _invariants(ch, q0, 2.5, "config 1")
    ch.close()


def test_config3_two_hundred_steps():
    cf = chanflow()
    mem, q0 = [], []
    for d in synth().ensemble_draw(48, 3):
        g = O.make_grid(d["nx"], d["ny"], d["nz"], d["nx2"], d["nz2"], 2.0, 2.0, d["re"], stretch=1.5, fix_mass=1)
        U0 = synth().laminar_perturbed(d["nx"], d["ny"], d["nz"], g.yh, g.y, seed=d["seed"], amp=1e-2, kmax=4.0)
        ch = make_channel(g)
        ch.set_state(data1=U0)
        ch.set_scalars(dt=0.005, t_total=0.0, p_grad=-2.0 / d["re"])
        ch.initialize_flow()
        mem.append(ch)
        q0.append(ch.mass_flux())
    ens = cf.Ensemble(mem)
    print("config 3: 200 steps of the 48-member batched ensemble", flush=True)
    ens.step(200)
    for i, ch in enumerate(mem):
        _invariants(ch, q0[i], 1.0, f"member {i}")
    ens.close()
    for ch in mem:
        ch.close()


def test_config4_two_thousand_les_steps():
    cf = chanflow()
    g = O.make_grid(64, 65, 64, 96, 96, 4.0, 3.0, 3300.0, stretch=2.75, fix_mass=1)
    U0 = synth().turbulent_like(g.nx, g.ny, g.nz, g.yh, g.y, seed=9, rms=0.1, re_tau=180.0)
    pkg = os.path.dirname(cf.__file__)
    net = cf.Mlp(np.load(os.path.join(pkg, "les_net_64x65x64.npy")))
    ch = make_channel(g)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.01, t_total=0.0, p_grad=-2.0 / g.re)
    ch.les_configure("learned", net=net, beta=1.0, every_substep=False)
    ch.initialize_flow()
    q0 = ch.mass_flux()
    ens = cf.Ensemble([ch], nstreams=1)   # the step as a captured graph, as the LES runs
    print("config 4: 2000 LES steps at 64x65x64", flush=True)
    ens.step(2000)
    _invariants(ch, q0, 20.0, "config 4")
    ens.close()
    ch.close()
    net.close()
