"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

Floating-point tolerances (FP64 throughout):
  * single transforms / solves / one nonlinear or rhs_solve call: relative
    L2 <= 1e-12 (both sides are round-off different restatements of the same
    arithmetic);
  * 10 RK3 steps of config 0: relative L2 of u, v, w and of the mean profile
    <= 1e-10; discrete divergence <= 1e-12.
"""
import math

import numpy as np
import pytest

from helpers import chanflow, grid_cfg0, init_cfg0, make_channel, rel_l2
from oracle import chan_oracle as O

pytestmark = pytest.mark.gpu

TOL1 = 1e-12


@pytest.fixture(scope="module")
def g0():
    return grid_cfg0(O)


def _rand_spec(rng, g, npl):
    C = rng.standard_normal((npl, g.nxh, g.nz)) + 1j * rng.standard_normal((npl, g.nxh, g.nz))
    return C


def test_version_and_create(g0):
    cf = chanflow()
    assert "gfx950" in cf.version()
    ch = make_channel(g0)
    ch.close()


@pytest.mark.parametrize("dims", [(32, 48, 32, 48), (16, 24, 16, 24), (32, 32, 32, 32), (24, 36, 12, 18)])
def test_xz_transforms(dims):
    nx, nx2, nz, nz2 = dims
    g = O.make_grid(nx, 9, nz, nx2, nz2, 2.0, 2.0, 1000.0, stretch=2.75)
    ch = make_channel(g)
    rng = np.random.default_rng(1)
    C = _rand_spec(rng, g, 5)
    got = ch.xz_c2r(C)
    ref = O.xz_c2r(g, C)
    assert rel_l2(got, ref) < TOL1
    f = rng.standard_normal((5, g.nz2, g.nx2))
    assert rel_l2(ch.xz_r2c(f), O.xz_r2c(g, f)) < TOL1
    fx, fy, fz = (rng.standard_normal((3, g.nz2, g.nx2)) for _ in range(3))
    assert rel_l2(ch.div_fluxes(fx, fy, fz), O.div_fluxes(g, fx, fy, fz)) < TOL1
    Cs = _rand_spec(rng, g, 4)
    assert rel_l2(ch.small_c2r(Cs), O.small_c2r(g, Cs)) < TOL1
    fs = rng.standard_normal((4, g.nz, g.nx))
    assert rel_l2(ch.small_r2c(fs), O.small_r2c(g, fs)) < TOL1
    ch.close()


def test_solvers(g0):
    ch = make_channel(g0)
    rng = np.random.default_rng(2)
    ns = 37
    n = g0.ny + 1
    f = rng.standard_normal((ns, n)) + 1j * rng.standard_normal((ns, n))
    alpha = 1.0 + rng.random(ns) * 3.0
    top = rng.standard_normal(ns) + 1j * rng.standard_normal(ns)
    bot = rng.standard_normal(ns) + 1j * rng.standard_normal(ns)
    diff = 3.7e-4
    got = ch.solve_helmholtz(f, alpha, diff, g0.d2h, top, bot)
    ref = O.solve_helmholtz(g0, n, f.T, alpha, diff, g0.d2h, top, bot).T
    assert rel_l2(got, ref) < TOL1
    fv = f[:, :g0.ny]
    got = ch.solve_helmholtz(fv, alpha, diff, g0.d2n, top, bot)
    ref = O.solve_helmholtz(g0, g0.ny, fv.T, alpha, diff, g0.d2n, top, bot).T
    assert rel_l2(got, ref) < TOL1
    p = rng.standard_normal((ns, n)) + 1j * rng.standard_normal((ns, n))
    al = -(0.5 + rng.random(ns) * 40.0)
    got = ch.solve_p(p, al)
    ref = O.solve_p(g0, p.T, al).T
    assert rel_l2(got, ref) < 1e-11
    ch.close()


def test_strip_divergence_modes(g0):
    """The per-mode strip_divergence and subtract_pressure_gradient of the
    kept initialisation code (PFC/solve.f90:14-77; one (kx,kz) column per
    reference call, here 48 columns per call, the x- and z-Nyquist columns
    among them) against the oracle's whole-field routines at the same modes;
    the mean mode, whose Neumann system is singular and which the reference's
    callers never pass, is rejected."""
    cf = chanflow()
    rng = np.random.default_rng(21)
    ch = make_channel(g0)
    nyp = g0.ny + 1
    cu = rng.standard_normal((3, nyp, g0.nxh, g0.nz)) + 1j * rng.standard_normal((3, nyp, g0.nxh, g0.nz))
    coef = 0.37
    ref_u, ref_p = O.strip_divergence(g0, cu, coef)
    modes = [(0, 1), (0, g0.nz // 2), (g0.nxh - 1, 0), (g0.nxh - 1, g0.nz - 1), (3, 5), (1, g0.nz // 2)]
    modes += [(int(a), int(b)) for a, b in zip(rng.integers(1, g0.nxh, 42), rng.integers(0, g0.nz, 42))]
    cols = np.stack([cu[:, :, ix, iz] for ix, iz in modes])
    p = [ix + 1 for ix, _ in modes]
    q = [iz + 1 for _, iz in modes]
    got_u, got_p = ch.strip_divergence_modes(cols, coef, p, q)
    for k, (ix, iz) in enumerate(modes):
        assert rel_l2(got_u[k], ref_u[:, :, ix, iz]) <= TOL1, (ix, iz)
        assert rel_l2(got_p[k], ref_p[:, ix, iz]) <= TOL1, (ix, iz)
    cpf = rng.standard_normal((nyp, g0.nxh, g0.nz)) + 1j * rng.standard_normal((nyp, g0.nxh, g0.nz))
    ref_g = O.subtract_pressure_gradient(g0, cu, cpf, coef)
    got_g = ch.subtract_pressure_gradient_modes(cols, np.stack([cpf[:, ix, iz] for ix, iz in modes]), coef, p, q)
    for k, (ix, iz) in enumerate(modes):
        assert rel_l2(got_g[k], ref_g[:, :, ix, iz]) <= TOL1, (ix, iz)
    with pytest.raises(cf.ChanError, match="mean mode"):
        ch.strip_divergence_modes(cols[:1], coef, [1], [1])
    with pytest.raises(cf.ChanError, match="out of range"):
        ch.strip_divergence_modes(cols[:1], coef, [g0.nxh + 1], [1])
    ch.close()


def test_strip_divergence(g0):
    ch = make_channel(g0)
    U = init_cfg0(O, g0, seed=3, amp=1e-2)
    got, cp = ch.strip_divergence(U, 0.37)
    ref, cpr = O.strip_divergence(g0, U, 0.37)
    assert rel_l2(got, ref) < TOL1
    assert rel_l2(cp, O.to_fortran_data4(cpr)) < 1e-11
    ch.close()


def test_nonlinear_single(g0):
    ch = make_channel(g0)
    U = O.remove_initial_divergence(g0, init_cfg0(O, g0, seed=4, amp=1e-2))
    ch.set_state(data1=U)
    dt0 = 0.01
    dt, cfl = ch.nonlinear(1, 1, dt=dt0)
    got = ch.get_data1()
    ref, dtr, cflr = O.nonlinear(g0, U, step=1, rk_step=1, dt=dt0, return_cfl=True)
    assert rel_l2(got, ref) < TOL1
    assert dt == dtr
    assert abs(cfl - cflr) <= 1e-14 * abs(cflr)
    ch.close()


@pytest.mark.parametrize("nch", [2, 5])
def test_nonlinear_plane_ranges(g0, nch):
    """The nonlinear pass run over plane ranges in turn (tuning key nl_chunks)
    computes the same term as the whole-plane kernels, bit for bit."""
    U = init_cfg0(O, g0, seed=21, amp=1e-2)
    out = []
    for n in (1, nch):
        ch = make_channel(g0)
        ch.set_tuning("nl_chunks", n)
        ch.set_state(data1=U)
        ch.nonlinear(1, 1, dt=0.01)
        out.append(ch.get_data1())
        ch.close()
    ref, _ = O.nonlinear(g0, U, step=1, rk_step=1, dt=0.01)
    assert rel_l2(out[1], ref) < TOL1
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("chunk", [2, 3, 7, 64])
def test_nonlinear_xpass_chunks(g0, chunk):
    """The x-pass with its planes split into ranges of at most `chunk` planes
    (tuning key xpass_chunk; 0, the default, picks the ranges from the
    occupancy) computes the same term, bit for bit, CFL included."""
    U = init_cfg0(O, g0, seed=22, amp=1e-2)
    out = []
    for c in (0, chunk):
        ch = make_channel(g0)
        ch.set_tuning("xpass_chunk", c)
        assert ch.get_tuning()["xpass_chunk"] == c
        ch.set_state(data1=U)
        dt, cfl = ch.nonlinear(1, 1, dt=0.01)
        out.append((ch.get_data1(), cfl))
        ch.close()
    ref, _, cflr = O.nonlinear(g0, U, step=1, rk_step=1, dt=0.01, return_cfl=True)
    assert rel_l2(out[1][0], ref) < TOL1
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1] == out[1][1] and abs(out[1][1] - cflr) <= 1e-14 * abs(cflr)


@pytest.mark.parametrize("key,value", [("xpass_wave", 0), ("xpass_prefetch", 0), ("xpass_threads", 256),
                                       ("zfwd_v2", 0), ("zinv_kx", 4), ("zinv_kx", 16), ("zfwd_kx", 4)])
@pytest.mark.parametrize("dims", [(32, 65, 32, 48, 48), (192, 9, 32, 288, 48)])
def test_nonlinear_tuning_variants(dims, key, value):
    """Every launch-shape variant that chan_set_tuning / chan_autotune can
    select for the nonlinear pass (the workgroup x-pass, the x-pass without
    its LDS row prefetch, a smaller x-pass workgroup, the z-forward with A and
    B together, z-inverse and z-forward workgroups of 4 or 16 columns), at
    nx2/2 = 24 and at the headline's 144: the same term as the default
    kernels to 1e-12 and the oracle's to 1e-12, CFL included."""
    nx, ny, nz, nx2, nz2 = dims
    g = O.make_grid(nx, ny, nz, nx2, nz2, 2.0, 1.0, 2500.0, stretch=1.5, fix_mass=1)
    s = __import__("helpers").synth()
    U = O.remove_initial_divergence(g, s.laminar_perturbed(nx, ny, nz, g.yh, g.y, seed=13, amp=5e-2, kmax=6.0))
    out = []
    for v in (None, value):
        ch = make_channel(g)
        if v is not None:
            ch.set_tuning(key, v)
        ch.set_state(data1=U)
        dt, cfl = ch.nonlinear(1, 1, dt=0.002)
        out.append((ch.get_data1(), cfl))
        ch.close()
    ref, _, cflr = O.nonlinear(g, U, step=1, rk_step=1, dt=0.002, return_cfl=True)
    assert rel_l2(out[1][0], out[0][0]) <= 1e-12
    assert rel_l2(out[1][0], ref) < TOL1
    assert abs(out[1][1] - cflr) <= 1e-14 * abs(cflr)


def test_nonlinear_dt_flag(g0):
    g = grid_cfg0(O, dt_flag=1)
    ch = make_channel(g)
    U = O.remove_initial_divergence(g, init_cfg0(O, g, seed=5, amp=1e-2))
    ch.set_state(data1=U)
    dt, cfl = ch.nonlinear(1, 1, dt=0.02)
    _, dtr, cflr = O.nonlinear(g, U, step=1, rk_step=1, dt=0.02, return_cfl=True)
    assert abs(dt - dtr) <= 1e-14 * dtr
    ch.close()


@pytest.mark.parametrize("rk", [1, 2, 3])
def test_rhs_solve_single(g0, rk):
    ch = make_channel(g0)
    rng = np.random.default_rng(10 + rk)
    st = O.initial_state(g0, init_cfg0(O, g0, seed=6, amp=1e-2), dt=0.01)
    NL, _ = O.nonlinear(g0, st.U, step=1, rk_step=rk, dt=0.01)
    st.U = NL
    st.R = st.R + 1e-3 * (rng.standard_normal(st.R.shape) + 1j * rng.standard_normal(st.R.shape))
    st.R[1, g0.ny] = 0.0
    ch.set_state(data1=NL, data2=O.to_fortran_data2(st.R), data4=O.to_fortran_data4(st.P))
    ch.set_scalars(dt=st.dt, p_grad=st.p_grad, d_um_dy=st.d_um_dy)
    pg = ch.rhs_solve(1, rk)
    ref = O.rhs_solve(g0, st.copy(), rk)
    d1, d2, d4 = ch.get_state()
    assert rel_l2(d1, ref.U) < TOL1
    assert rel_l2(d2, O.to_fortran_data2(ref.R)) < TOL1
    assert rel_l2(d4, O.to_fortran_data4(ref.P)) < 1e-11
    assert abs(pg - ref.p_grad) <= 1e-12 * abs(ref.p_grad)
    sc = ch.get_scalars()
    assert rel_l2(sc["d_um_dy"], ref.d_um_dy) < TOL1
    ch.close()


@pytest.mark.parametrize("variant", [(3, 8, 1), (3, 8, 0), (3, 16, 0), (1, 8, 0)])
def test_rhs_solve_variants(g0, variant):
    """Every sweep-based solve kernel variant (lanes per mode, row block, LDS
    ring; the segmented solve switched off) against the oracle for the three
    substeps; the ring and the register-blocked loads of the 3-lane kernel run
    the same recurrences (the compiler may contract multiply-adds differently
    in the two) and agree to round-off."""
    lanes, bl, dma = variant
    rng = np.random.default_rng(40)
    st = O.initial_state(g0, init_cfg0(O, g0, seed=9, amp=1e-2), dt=0.01)
    for rk in (1, 2, 3):
        NL, _ = O.nonlinear(g0, st.U, step=1, rk_step=rk, dt=0.01)
        s0 = st.copy()
        s0.U = NL
        s0.R = s0.R + 1e-3 * (rng.standard_normal(s0.R.shape) + 1j * rng.standard_normal(s0.R.shape))
        s0.R[1, g0.ny] = 0.0
        outs = []
        for d in ((dma, 1 - dma) if lanes == 3 else (dma,)):
            ch = make_channel(g0)
            ch.set_tuning("solve_seg", 0)
            ch.set_tuning("solve_lanes", lanes)
            ch.set_tuning("solve_bl", bl)
            ch.set_tuning("solve_dma", d)
            ch.set_state(data1=NL, data2=O.to_fortran_data2(s0.R), data4=O.to_fortran_data4(s0.P))
            ch.set_scalars(dt=s0.dt, p_grad=s0.p_grad, d_um_dy=s0.d_um_dy)
            ch.rhs_solve(1, rk)
            outs.append(ch.get_state())
            ch.close()
        ref = O.rhs_solve(g0, s0.copy(), rk)
        d1, d2, d4 = outs[0]
        assert rel_l2(d1, ref.U) < TOL1, rk
        assert rel_l2(d2, O.to_fortran_data2(ref.R)) < TOL1, rk
        assert rel_l2(d4, O.to_fortran_data4(ref.P)) < 1e-11, rk
        for x, y in zip(outs[0], outs[-1]):
            assert rel_l2(x, y) <= 1e-14, rk


@pytest.mark.parametrize("ny,layout", [(33, 1), (65, 1), (97, 1), (129, 1), (33, 2), (97, 2), (129, 2),
                                       (33, 3), (49, 3), (65, 3), (97, 3)])
def test_rhs_solve_segmented(ny, layout):
    """The segmented column solve (partitioned tridiagonal solves, columns in
    registers) against the oracle for the three substeps, with a perturbed RK
    register and wall data on both walls, and against the sweep-based 3-lane
    solve to 1e-12.  Layout 1, 4 modes per wavefront: segments of 5 rows
    (ny = 33: 7 segments, idle lanes past them), 6 (65: 11), 7 (97: 14) and
    9 (129: 15).  Layout 2, 2 modes per wavefront: 5-row segments, 7 of 32
    at ny = 33, 20 at 97 (3 rows in the last), 26 at 129.  Layout 3, 8 modes
    per wavefront: 5-row segments at ny = 33 (7 of 8), 9-row segments at 49
    (6) and 65 (all 8, 3 rows in the last); 98 rows do not fit 8 segments, so
    ny = 97 takes layout 1."""
    # 9 x 16 = 144 modes: 72 two-mode groups, 64 in the XCD-grouped prefix of
    # the grid and 8 past it
    g = O.make_grid(16, ny, 16, 24, 24, 2.0, 2.0, 2500.0, stretch=1.5, fix_mass=1, dt_flag=0)
    rng = np.random.default_rng(ny)
    s = __import__("helpers").synth()
    U0 = s.laminar_perturbed(g.nx, g.ny, g.nz, g.yh, g.y, seed=ny, amp=1e-2, kmax=4.0)
    st = O.initial_state(g, U0, dt=0.01)
    cbc_top = 1e-3 * (rng.standard_normal((3, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.nxh, g.nz)))
    cbc_bot = 1e-3 * (rng.standard_normal((3, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.nxh, g.nz)))
    cbc_top[:, 0, 0] = cbc_top[:, 0, 0].real
    cbc_bot[:, 0, 0] = cbc_bot[:, 0, 0].real
    for rk in (1, 2, 3):
        NL, _ = O.nonlinear(g, st.U, step=1, rk_step=rk, dt=0.01)
        s0 = st.copy()
        s0.U = NL
        s0.R = s0.R + 1e-3 * (rng.standard_normal(s0.R.shape) + 1j * rng.standard_normal(s0.R.shape))
        s0.R[1, g.ny] = 0.0
        s0.cbc_top, s0.cbc_bot = cbc_top, cbc_bot
        outs = []
        for seg in (layout, 0):
            ch = make_channel(g)
            ch.set_tuning("solve_seg", seg)
            ch.set_state(data1=NL, data2=O.to_fortran_data2(s0.R), data4=O.to_fortran_data4(s0.P))
            ch.set_scalars(dt=s0.dt, p_grad=s0.p_grad, d_um_dy=s0.d_um_dy)
            ch.set_wall_bc(cbc_top, cbc_bot)
            ch.rhs_solve(1, rk)
            outs.append((ch.get_state(), ch.get_scalars()))
            ch.close()
        ref = O.rhs_solve(g, s0.copy(), rk)
        (d1, d2, d4), sc = outs[0]
        assert rel_l2(d1, ref.U) < TOL1, rk
        assert rel_l2(d2, O.to_fortran_data2(ref.R)) < TOL1, rk
        assert rel_l2(d4, O.to_fortran_data4(ref.P)) < 1e-11, rk
        assert rel_l2(sc["d_um_dy"], ref.d_um_dy) < TOL1, rk
        assert abs(sc["p_grad"] - ref.p_grad) <= 1e-12 * abs(ref.p_grad), rk
        for x, y in zip(outs[0][0], outs[1][0]):
            assert rel_l2(x, y) <= 1e-12, rk


def test_initialize_flow(g0):
    ch = make_channel(g0)
    U0 = init_cfg0(O, g0, seed=7)
    ch.set_state(data1=U0)
    ch.initialize_flow()
    st = O.initial_state(g0, U0, dt=0.01)
    d1, d2, d4 = ch.get_state()
    assert rel_l2(d1, st.U) < TOL1
    assert rel_l2(d2, O.to_fortran_data2(st.R)) < TOL1
    assert rel_l2(d4, O.to_fortran_data4(st.P)) < 1e-11
    assert rel_l2(ch.get_scalars()["d_um_dy"], st.d_um_dy) < TOL1
    ch.close()


def test_ten_steps_config0(g0):
    """Config 0 parity: 10 RK3 steps, rel L2 <= 1e-10, divergence <= 1e-12."""
    ch = make_channel(g0)
    U0 = init_cfg0(O, g0, seed=0)
    dt = 0.01
    ch.set_state(data1=U0)
    ch.set_scalars(dt=dt, p_grad=-2.0 / g0.re)
    ch.initialize_flow()
    st = O.initial_state(g0, U0, dt=dt)
    ch.step(10, step0=1)
    for s in range(1, 11):
        O.rk3_step(g0, st, s)
    d1 = ch.get_data1()
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10, c
    assert rel_l2(d1[0, :, 0, 0].real, st.U[0, :, 0, 0].real) <= 1e-10
    assert ch.divergence_max() <= 1e-12
    sc = ch.get_scalars()
    assert abs(sc["t_total"] - st.t_total) <= 1e-14
    assert abs(sc["p_grad"] - st.p_grad) <= 1e-9 * abs(st.p_grad)
    ch.close()


def test_ten_steps_reference_defaults(g0):
    """Config 0 at the reference's own defaults rather than the patched ones:
    grid stretching 2.75 (PFC/setup.f90:545) and the initial dt rule
    0.25 lx/nx (PFC/setup.f90:1377, dt 0.0491 here) held fixed (dt_flag 0);
    10 RK3 steps, rel L2 <= 1e-10, divergence <= 1e-12."""
    g = O.make_grid(g0.nx, g0.ny, g0.nz, g0.nx2, g0.nz2, 2.0, 2.0, g0.re, stretch=2.75, fix_mass=1)
    dt = O.default_dt(g)
    ch = make_channel(g)
    U0 = init_cfg0(O, g, seed=5)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=dt, p_grad=-2.0 / g.re)
    ch.initialize_flow()
    st = O.initial_state(g, U0, dt=dt)
    ch.step(10, step0=1)
    for s in range(1, 11):
        O.rk3_step(g, st, s)
    d1 = ch.get_data1()
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10, c
    assert ch.divergence_max() <= 1e-12
    assert abs(ch.get_scalars()["t_total"] - st.t_total) <= 1e-13
    ch.close()


def test_step_api_matches_phase_calls(g0):
    """chan_step == the driver's boundary/nonlinear/rhs_solve call sequence."""
    U0 = init_cfg0(O, g0, seed=8, amp=1e-2)
    a = make_channel(g0)
    b = make_channel(g0)
    for ch in (a, b):
        ch.set_state(data1=U0)
        ch.set_scalars(dt=0.01)
        ch.initialize_flow()
    a.step(2, step0=1)
    for s in (1, 2):
        b.boundary()
        for rk in (1, 2, 3):
            b.nonlinear(s, rk)
            b.rhs_solve(s, rk)
    assert rel_l2(a.get_data1(), b.get_data1()) < 1e-14
    a.close(); b.close()


def test_feedback_control(g0):
    """boundary() with nonzero gains (PFC/read_write.f90:1465-1532) inside the step."""
    g = grid_cfg0(O, ctrlu=0.5, ctrlw=0.3)
    ch = make_channel(g)
    U0 = init_cfg0(O, g, seed=9, amp=1e-2)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.01)
    ch.initialize_flow()
    st = O.initial_state(g, U0, dt=0.01)
    ch.step(3, step0=1)
    for s in (1, 2, 3):
        O.rk3_step(g, st, s)
    assert rel_l2(ch.get_data1(), st.U) < 1e-11
    ch.close()


@pytest.mark.parametrize("dims", [(32, 33, 32, 48, 48), (32, 33, 32, 32, 32), (16, 33, 24, 24, 36)])
def test_feedback_control_grids(dims):
    """The feedback law of boundary() (PFC/read_write.f90:1465-1532) on
    dealiased grids, where the library applies it to the Fourier
    coefficients (the physical-space round trip there keeps a spectrum except
    for the kx = 0 column's Hermitian part), and on an undealiased grid, where
    it goes through the physical wall-adjacent planes: 3 steps against the
    oracle's physical-space law, both gains on, at 1e-11."""
    nx, ny, nz, nx2, nz2 = dims
    g = O.make_grid(nx, ny, nz, nx2, nz2, 2.0, 2.0, 2500.0, stretch=1.5, fix_mass=1, dt_flag=0,
                    ctrlu=0.4, ctrlw=0.7)
    s = __import__("helpers").synth()
    U0 = s.laminar_perturbed(nx, ny, nz, g.yh, g.y, seed=31, amp=1e-2, kmax=4.0)
    ch = make_channel(g)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.01)
    ch.initialize_flow()
    st = O.initial_state(g, U0, dt=0.01)
    ch.step(3, step0=1)
    for k in (1, 2, 3):
        O.rk3_step(g, st, k)
    assert rel_l2(ch.get_data1(), st.U) < 1e-11
    ch.close()


def test_linearized_matvec():
    """Arnoldi matvec compute_time_deriv on a small grid."""
    g = O.make_grid(16, 33, 16, 24, 24, 2.0, 2.0, 2800.0, stretch=2.75, fix_mass=0)
    ch = make_channel(g)
    base = O.laminar_baseflow(g)
    ch.set_baseflow(base)
    dt = O.linearized_dt(g)
    ch.set_scalars(dt=dt)
    rng = np.random.default_rng(11)
    v = rng.standard_normal((3, g.ny + 1, g.nz, g.nx))
    w, T = ch.linearized_matvec(v, 3)
    wr, Tr = O.compute_time_deriv(g, v, base, dt, 3)
    assert rel_l2(w, wr) < 1e-11
    assert T == Tr
    ch.close()


def test_cfl_abort(g0):
    """A too-large dt with dt_flag == 0 is reported as an error (PFC/pass1.f90:235-244)."""
    cf = chanflow()
    ch = make_channel(g0)
    ch.set_state(data1=init_cfg0(O, g0, seed=1))
    with pytest.raises(cf.ChanError, match="CFL IS TOO BIG"):
        ch.nonlinear(1, 1, dt=1.0)
    ch.close()


@pytest.mark.parametrize("dims", [(16, 17, 16, 24, 24), (32, 17, 16, 48, 24), (64, 17, 32, 96, 48),
                                  (96, 17, 16, 144, 24), (128, 9, 16, 192, 24), (192, 9, 16, 288, 24),
                                  (384, 9, 8, 576, 12)])
def test_nonlinear_x_lengths(dims):
    """One nonlinear pass at x lengths that take each x-pass variant: the
    workgroup kernel with the runtime-planned transform (nx2/2 = 12), the
    wave-per-row kernel with one, two and three points per lane
    (nx2/2 = 24 ... 144), and the workgroup kernel with a compile-time
    length (nx2/2 = 288)."""
    nx, ny, nz, nx2, nz2 = dims
    g = O.make_grid(nx, ny, nz, nx2, nz2, 2.0, 2.0, 2500.0, stretch=1.5, fix_mass=1)
    ch = make_channel(g)
    s = __import__("helpers").synth()
    U = O.remove_initial_divergence(g, s.laminar_perturbed(nx, ny, nz, g.yh, g.y, seed=11, amp=5e-2, kmax=6.0))
    ch.set_state(data1=U)
    dt, cfl = ch.nonlinear(1, 1, dt=0.002)
    ref, dtr, cflr = O.nonlinear(g, U, step=1, rk_step=1, dt=0.002, return_cfl=True)
    assert rel_l2(ch.get_data1(), ref) < TOL1
    assert abs(cfl - cflr) <= 1e-14 * abs(cflr)
    ch.close()


@pytest.mark.parametrize("lin", [False, True])
@pytest.mark.parametrize("dims,rows", [((16, 17, 16, 24, 24), 4), ((32, 17, 16, 48, 24), 2),
                                       ((32, 17, 16, 48, 24), 4), ((32, 17, 12, 48, 18), 4),
                                       ((48, 17, 16, 72, 24), 2), ((64, 17, 32, 96, 48), 2),
                                       ((96, 17, 16, 144, 24), 2), ((32, 17, 16, 48, 24), 1)])
def test_nonlinear_xpass_rows(dims, rows, lin):
    """The wavefront x-pass with several z-rows per wavefront (tuning key 12):
    one nonlinear pass, and its linearized form, at nx2/2 = 16 ... 72 with 1,
    2 and 4 rows against the oracle, CFL included, and against the pass with
    one row per wavefront to round-off (the arithmetic per point is the same;
    only the grouping of the sequences differs).  nz2 = 18 with 4 rows asked
    for falls back to 2 (18 is no multiple of 4)."""
    nx, ny, nz, nx2, nz2 = dims
    g = O.make_grid(nx, ny, nz, nx2, nz2, 2.0, 1.0, 2800.0, stretch=2.75, fix_mass=0)
    s = __import__("helpers").synth()
    base = O.laminar_baseflow(g) if lin else None
    amp = 1e-3 if lin else 5e-2
    U = O.remove_initial_divergence(g, s.laminar_perturbed(nx, ny, nz, g.yh, g.y, seed=14, amp=amp, kmax=6.0))
    out = []
    for r in (1, rows):
        ch = make_channel(g)
        ch.set_tuning("xpass_rows", r)
        assert ch.get_tuning()["xpass_rows"] == r
        if lin:
            ch.set_baseflow(base)
        ch.set_state(data1=U)
        dt, cfl = ch.nonlinear(1, 1, dt=0.002, linearized=lin)
        out.append((ch.get_data1(), cfl))
        ch.close()
    if lin:
        ref, _, cflr = O.nonlinear_linearized(g, U, base, step=1, rk_step=1, dt=0.002, return_cfl=True)
    else:
        ref, _, cflr = O.nonlinear(g, U, step=1, rk_step=1, dt=0.002, return_cfl=True)
    assert rel_l2(out[1][0], out[0][0]) <= 1e-14
    assert rel_l2(out[1][0], ref) < TOL1
    assert abs(out[1][1] - cflr) <= 1e-14 * abs(cflr)


@pytest.mark.parametrize("dims", [(16, 17, 16, 24, 24), (32, 17, 16, 48, 24), (64, 97, 64, 96, 96),
                                  (96, 17, 16, 144, 24), (192, 9, 16, 288, 24)])
def test_nonlinear_linearized_x_lengths(dims):
    """One linearized nonlinear pass (PFC/pass1_linearized.f90:135-509,
    products U.u' + u'.U about the laminar base flow, CFL with the base flow
    included) through each x-pass variant's linearized instance: the workgroup
    kernel with the runtime-planned transform (nx2/2 = 12) and the
    wave-per-row kernel with one, two and three points per lane (nx2/2 = 24,
    48, 72, 144; 64x97x64 is configuration 2's shape, nx2/2 = 144 the
    headline's x length)."""
    nx, ny, nz, nx2, nz2 = dims
    g = O.make_grid(nx, ny, nz, nx2, nz2, 2.0, 1.0, 2800.0, stretch=2.75, fix_mass=0)
    ch = make_channel(g)
    base = O.laminar_baseflow(g)
    ch.set_baseflow(base)
    s = __import__("helpers").synth()
    U = O.remove_initial_divergence(g, s.laminar_perturbed(nx, ny, nz, g.yh, g.y, seed=12, amp=1e-3, kmax=6.0))
    ch.set_state(data1=U)
    dt, cfl = ch.nonlinear(1, 1, dt=0.002, linearized=True)
    ref, dtr, cflr = O.nonlinear_linearized(g, U, base, step=1, rk_step=1, dt=0.002, return_cfl=True)
    assert rel_l2(ch.get_data1(), ref) < TOL1
    assert dt == dtr
    assert abs(cfl - cflr) <= 1e-14 * abs(cflr)
    ch.close()


@pytest.fixture(scope="module")
def g1():
    """Headline shape: Re_c 3300, box 4pi x 2 x 4pi/3, 192x129x192, 288x288, stretch 2.75."""
    return O.make_grid(192, 129, 192, 288, 288, 4.0, 3.0, 3300.0, stretch=2.75, fix_mass=1)


def _headline_field(g):
    s = __import__("helpers").synth()
    return s.turbulent_like(g.nx, g.ny, g.nz, g.yh, g.y, seed=3, rms=0.1, re_tau=180.0)


def test_headline_shape_nonlinear(g1):
    ch = make_channel(g1)
    U = _headline_field(g1)
    ch.set_state(data1=U)
    dt, cfl = ch.nonlinear(1, 1, dt=0.005)
    ref, dtr, cflr = O.nonlinear(g1, U, step=1, rk_step=1, dt=0.005, return_cfl=True)
    assert rel_l2(ch.get_data1(), ref) < TOL1
    assert abs(cfl - cflr) <= 1e-14 * abs(cflr)
    ch.close()


def test_headline_shape_one_step(g1):
    """Initialisation plus one RK3 step at the headline shape: rel L2 <= 1e-11
    for u, v, w, the pressure and the next RK register; divergence <= 1e-12."""
    ch = make_channel(g1)
    U0 = _headline_field(g1)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.005, p_grad=-2.0 / g1.re)
    ch.initialize_flow()
    st = O.initial_state(g1, U0, dt=0.005)
    ch.step(1, step0=1)
    O.rk3_step(g1, st, 1)
    d1, d2, d4 = ch.get_state()
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-11, c
    assert rel_l2(d2, O.to_fortran_data2(st.R)) <= 1e-11
    assert rel_l2(d4, O.to_fortran_data4(st.P)) <= 1e-10
    assert ch.divergence_max() <= 1e-12
    assert abs(ch.get_scalars()["p_grad"] - st.p_grad) <= 1e-10 * abs(st.p_grad)
    ch.close()


def test_headline_shape_ten_steps(g1):
    """The headline workload's parity criterion: 10 RK3 steps of the
    192x129x192 channel (PFC/main_mpi.f90:70-87) from the same seeded field
    as the oracle: relative L2 <= 1e-10 for u, v, w and the mean profile,
    discrete divergence <= 1e-12 (PFC/pass2.f90:380-401)."""
    ch = make_channel(g1)
    U0 = _headline_field(g1)
    ch.set_state(data1=U0)
    ch.set_scalars(dt=0.005, p_grad=-2.0 / g1.re)
    ch.initialize_flow()
    ch.step(10, step0=1)
    d1 = ch.get_data1()
    div = ch.divergence_max()
    ch.close()
    st = O.initial_state(g1, U0, dt=0.005)
    for s in range(1, 11):
        print(f"oracle step {s}/10 at 192x129x192", flush=True)
        O.rk3_step(g1, st, s)
    for c in range(3):
        assert rel_l2(d1[c], st.U[c]) <= 1e-10, c
    assert rel_l2(d1[0, :, 0, 0].real, st.U[0, :, 0, 0].real) <= 1e-10
    assert rel_l2(d1, U0) > 1e-4            # the flow moved
    assert div <= 1e-12
