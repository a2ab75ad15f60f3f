"""Batched ensemble and launch-shape tuning through the C-ABI.

* Every replay of the captured timestep graph is checked against twins of
  the members stepped one by one with chan_step, with host synchronisation
  between the replays (the graph holds only kernels; bitwise equality is
  expected, 1e-14 is allowed): the stream-parallel graph on 5 members, and
  the batched launches (one launch per phase and transform length over all
  members) on the 48 members of the seeded ensemble of uneven sizes.
* Members are also checked against the oracle after the replays.
* chan_autotune leaves the state unchanged and its choice gives the same
  step to round-off.
"""
import numpy as np
import pytest

from helpers import make_channel, rel_l2, synth
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu

SHAPES = [  # nx, ny, nz, nx2, nz2, Lx/pi, Lx/Lz, Re
    (32, 49, 32, 48, 48, 2.0, 2.0, 2100.0),
    (48, 65, 32, 72, 48, 2.0, 2.0, 2600.0),
    (64, 49, 48, 96, 72, 2.0, 2.0, 3100.0),
    (32, 97, 64, 48, 96, 2.0, 2.0, 3400.0),
    (96, 49, 32, 144, 48, 2.0, 2.0, 2900.0),
]


def _member(shape, seed):
    nx, ny, nz, nx2, nz2, lxp, asp, re = shape
    g = O.make_grid(nx, ny, nz, nx2, nz2, lxp, asp, re, stretch=1.5, fix_mass=1)
    U0 = synth().laminar_perturbed(nx, ny, nz, g.yh, g.y, seed=seed, amp=1e-2, kmax=4.0)
    ch = make_channel(g)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.005, t_total=0.0, p_grad=-2.0 / re)
    ch.initialize_flow()
    return g, U0, ch


def test_ensemble_replays_match_eager():
    cf = __import__("helpers").chanflow()
    mem = [_member(s, 20 + i) for i, s in enumerate(SHAPES)]
    twins = [_member(s, 20 + i) for i, s in enumerate(SHAPES)]
    ens = cf.Ensemble([m[2] for m in mem], nstreams=3)   # members' own kernels on 3 streams
    for step in range(1, 4):
        ens.step(1)
        for (g, _, a), (_, _, b) in zip(mem, twins):
            b.step(1, step0=step)
            da, db = a.get_state(), b.get_state()
            for x, y in zip(da, db):
                assert rel_l2(x, y) <= 1e-14
            assert a.get_scalars()["t_total"] == b.get_scalars()["t_total"]
    # member 1 against the oracle after the three replays
    g, U0, ch = mem[1]
    st = O.initial_state(g, U0, dt=0.005)
    for s in (1, 2, 3):
        O.rk3_step(g, st, s)
    d1 = ch.get_data1()
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10
    assert ch.divergence_max() <= 1e-12
    ens.step(2)  # more replays after the checks above
    ens.close()
    for _, _, ch in mem + twins:
        ch.close()


def test_autotune_restores_state_and_keeps_results():
    g, U0, a = _member(SHAPES[2], 31)
    _, _, b = _member(SHAPES[2], 31)
    before = a.get_state()
    ms = a.autotune(steps=1)
    assert ms > 0.0
    after = a.get_state()
    for x, y in zip(before, after):
        assert np.array_equal(x, y)
    t = a.get_tuning()
    assert t["solve_lanes"] in (1, 3) and t["solve_bl"] in (4, 8, 16)
    a.step(2, step0=1)
    b.step(2, step0=1)
    for x, y in zip(a.get_state(), b.get_state()):
        assert rel_l2(x, y) <= 1e-12
    with pytest.raises(Exception):
        a.set_tuning("solve_bl", 5)
    a.close(); b.close()


def _draw_member(d, rows=0, seg8=False):
    g = O.make_grid(d["nx"], d["ny"], d["nz"], d["nx2"], d["nz2"], 2.0, 2.0, d["re"], stretch=1.5, fix_mass=1)
    U0 = synth().laminar_perturbed(d["nx"], d["ny"], d["nz"], g.yh, g.y, seed=d["seed"], amp=1e-2, kmax=4.0)
    ch = make_channel(g)
    if rows:
        ch.set_tuning("xpass_rows", rows)
    if seg8:
        ch.set_tuning("solve_seg", 3)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.005, t_total=0.0, p_grad=-2.0 / d["re"])
    ch.initialize_flow()
    return g, U0, ch


def test_batched_ensemble_48_members():
    """The seeded 48-member ensemble (Nx, Nz in {32, 48, 64, 96}, Ny in {49, 65,
    97}, Re_c in [2000, 3500]) advanced by the batched launches: two replays,
    each checked against the 48 twins stepped one by one (bitwise equality
    expected, 1e-14 allowed), then the two smallest members against the
    oracle's RK3 steps (relative L2 <= 1e-10, divergence <= 1e-12).  The
    batched x-pass takes two z-rows per wavefront at every length of the draw
    (nx2/2 <= 72), a single channel's only at nx2/2 <= 24, so members and
    twins are asked for two rows (tuning key 12; the initial projection runs a
    nonlinear pass too); the batched solve takes 8 modes per wavefront at
    ny = 65 (8 segments of 9 rows), so those twins are asked for that layout
    (tuning key 11 = 3): the same launch shapes, the same arithmetic."""
    cf = __import__("helpers").chanflow()
    draws = synth().ensemble_draw(48, seed=3)
    assert all(d["nx2"] // 2 <= 72 for d in draws)
    mem = [_draw_member(d, rows=2) for d in draws]
    assert {d["ny"] for d in draws} == {49, 65, 97}
    twins = [_draw_member(d, rows=2, seg8=d["ny"] == 65) for d in draws]
    ens = cf.Ensemble([m[2] for m in mem])            # batched launches
    for step in (1, 2):
        print(f"batched ensemble replay {step}", flush=True)
        ens.step(1)
        for (g, _, a), (_, _, b) in zip(mem, twins):
            b.step(1, step0=step)
            for x, y in zip(a.get_state(), b.get_state()):
                assert rel_l2(x, y) <= 1e-14, (g.nx, g.ny, g.nz)
            assert a.get_scalars()["t_total"] == b.get_scalars()["t_total"]
    small = sorted(range(len(draws)), key=lambda i: draws[i]["nx"] * draws[i]["ny"] * draws[i]["nz"])[:2]
    for i in small:
        g, U0, ch = mem[i]
        st = O.initial_state(g, U0, dt=0.005)
        for s in (1, 2):
            O.rk3_step(g, st, s)
        d1 = ch.get_data1()
        for c in range(3):
            assert rel_l2(d1[c], st.U[c]) <= 1e-10, (i, c)
        assert ch.divergence_max() <= 1e-12
    ens.close()
    for _, _, ch in mem + twins:
        ch.close()


def test_batched_ensemble_autotune_keeps_state_and_results():
    """chan_ensemble_autotune on 12 members of the seeded draw: every member's
    state is the same after the tuning as before it (bit for bit), and the
    tuned ensemble's next two steps equal those of an untuned twin ensemble
    to round-off (the plans differ in the x-pass plane ranges, which leave
    the arithmetic as it is, and may differ in the tile shapes: rows per
    x-pass wavefront, modes per solve wavefront)."""
    cf = __import__("helpers").chanflow()
    draws = synth().ensemble_draw(12, seed=5)
    mem = [_draw_member(d) for d in draws]
    twins = [_draw_member(d) for d in draws]
    ens = cf.Ensemble([m[2] for m in mem])
    ref = cf.Ensemble([m[2] for m in twins])
    before = [m[2].get_state() for m in mem]
    ms = ens.autotune(steps=2)
    assert ms > 0.0
    for (_, _, ch), st in zip(mem, before):
        for x, y in zip(ch.get_state(), st):
            assert np.array_equal(x, y)
    ens.step(2)
    ref.step(2)
    for (g, _, a), (_, _, b) in zip(mem, twins):
        for x, y in zip(a.get_state(), b.get_state()):
            # round-off of the O(1) velocity carried into the O(1e-4) pressure
            assert rel_l2(x, y) <= 1e-10, (g.nx, g.ny, g.nz)
    streams = cf.Ensemble([m[2] for m in mem[:2]], nstreams=1)
    with pytest.raises(cf.ChanError):
        streams.autotune(steps=1)
    streams.close()
    ens.close()
    ref.close()
    for _, _, ch in mem + twins:
        ch.close()


def test_batched_ensemble_autotune_failure_restores_state():
    """chan_ensemble_autotune when every plan's trial trips the CFL check (dt
    far above the limit): it reports the failure, and every member's state
    and scalars are what they were before the tuning."""
    cf = __import__("helpers").chanflow()
    draws = synth().ensemble_draw(3, seed=11)
    mem = [_draw_member(d) for d in draws]
    for _, _, ch in mem:
        ch.set_scalars(dt=5.0)
    ens = cf.Ensemble([m[2] for m in mem])
    before = [m[2].get_state() for m in mem]
    sc_before = [m[2].get_scalars() for m in mem]
    with pytest.raises(cf.ChanError, match="failed the CFL check"):
        ens.autotune(steps=1)
    for (_, _, ch), st, sc in zip(mem, before, sc_before):
        for x, y in zip(ch.get_state(), st):
            assert np.array_equal(x, y)
        after = ch.get_scalars()
        assert after["dt"] == sc["dt"] and after["t_total"] == sc["t_total"]
    ens.close()
    for _, _, ch in mem:
        ch.close()


def test_batched_ensemble_rejects_ineligible_member():
    cf = __import__("helpers").chanflow()
    g, _, a = _member(SHAPES[0], 40)
    _, _, b = _member(SHAPES[1], 41)
    b.set_tuning("solve_seg", 0)
    with pytest.raises(cf.ChanError, match="member 1 cannot join the batched launches"):
        cf.Ensemble([a, b])
    a.close(); b.close()
