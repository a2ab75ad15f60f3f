"""Pins of the CPU oracle (CPU only): reference fixtures, closed forms,
invariants and convergence orders.  Each check is chosen so that a plausible
slip in the restatement (dropped term, wrong sign, index or operand) fails it.
"""
import math
import os

import numpy as np
import pytest

from oracle import chan_oracle as O
from helpers import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = np.load(os.path.join(HERE, "golden", "reference_fixtures.npz"))


# --------------------------------------------------------------------------- grid
def test_grid_matches_reference_makegrid():
    """PFC/makegrid.m (written with f16.8 by PFC/setup.f90:1331-1359)."""
    nx, ny, nz, lxp, asp = FIX["makegrid_case"]
    g = O.make_grid(int(nx), int(ny), int(nz), 192, 96, lxp, asp, 5000.0, stretch=2.75)
    assert np.abs(g.y - FIX["makegrid_y"]).max() < 6e-9
    assert np.abs(g.yh - FIX["makegrid_yh"]).max() < 6e-9
    x = g.lx * np.arange(int(nx) + 1) / nx
    z = g.lz * np.arange(int(nz) + 1) / nz
    assert np.abs(x - FIX["makegrid_x"]).max() < 6e-9
    assert np.abs(z - FIX["makegrid_z"]).max() < 6e-9


def test_dt_rule_matches_reference_log():
    """INITIAL DT of PFC/diagnostics.txt:98 == 0.25*lx/nx (PFC/setup.f90:1377)."""
    g = O.make_grid(128, 200, 64, 192, 96, 6.0, 3.0, 5000.0)
    assert O.default_dt(g) == float(FIX["diagnostics_dt0"])


def test_rk_constants_are_float32_literals():
    """PFC/pass2.f90:14-18: single-precision literal quotients."""
    assert O.RKA[1] == -0.28333333134651184
    assert O.RKB[0] == 0.5333333611488342
    assert O.RKC[0] == 0.3020833432674408
    assert O.RKD[0] == 0.23125000298023224
    # they still form the SMR91 scheme up to float rounding
    assert abs(sum(O.RKB) + O.RKA[1] + O.RKA[2] - 1.0) < 1e-6
    assert abs(sum(O.RKC) + sum(O.RKD) - 1.0) < 1e-6


@pytest.mark.parametrize("stretch", [1.5, 2.75])
def test_y_operators_exactness(stretch):
    g = O.make_grid(8, 33, 8, 12, 12, 2.0, 2.0, 100.0, stretch=stretch)
    ny = g.ny
    # second-difference stencils annihilate linear profiles (all rows, incl. the DM-closed ones)
    lin = 0.3 - 1.7 * g.yh
    assert np.abs(O.d2_dy(lin[:, None], g.d2h)[1:ny, 0]).max() < 1e-10
    linv = 0.3 - 1.7 * g.y
    assert np.abs(O.d2_dy(linv[:, None], g.d2n)[1:ny - 1, 0]).max() < 1e-10
    # finite-volume (telescoping) form: sum_j hh_j D2H f_j = flux(top face) - flux(bottom face)
    f = np.cos(1.3 * g.yh) + g.yh ** 3
    d2 = O.d2_dy(f[:, None], g.d2h)[:, 0]
    js = np.arange(2, ny - 1)
    lhs = np.sum(g.hh[js] * d2[js])
    rhs = g.dn[ny - 2] * (f[ny - 1] - f[ny - 2]) - g.dn[1] * (f[2] - f[1])
    assert abs(lhs - rhs) < 1e-12 * np.abs(d2).max()
    # consistency: the stencils converge to the second derivative under refinement
    errs = []
    for n in (33, 65, 129):
        gg = O.make_grid(8, n, 8, 12, 12, 2.0, 2.0, 100.0, stretch=stretch)
        ff = np.cos(1.3 * gg.yh)
        dd = O.d2_dy(ff[:, None], gg.d2h)[:, 0]
        errs.append(np.abs(dd[2:n - 1] + 1.69 * np.cos(1.3 * gg.yh[2:n - 1])).max())
    assert errs[0] > 1.8 * errs[1] > 1.8 * 1.8 * errs[2], errs
    # DM: one-sided first derivative at the walls, exact for quadratics
    dm = g.dm
    q = 1.0 + 0.5 * g.yh + 2.0 * g.yh ** 2
    fb = dm[0, 0] * q[0] + dm[1, 0] * q[1] + dm[2, 0] * q[2]
    ft = dm[0, 1] * q[ny - 2] + dm[1, 1] * q[ny - 1] + dm[2, 1] * q[ny]
    assert abs(fb - (0.5 + 4.0 * g.yh[0])) < 1e-8
    assert abs(ft - (0.5 + 4.0 * g.yh[ny])) < 1e-8
    # quadratures: channel height and exactness for linear profiles
    assert abs(g.dyh.sum() - 2.0) < 1e-14
    assert abs(g.dy.sum() - 2.0) < 1e-14
    assert abs(np.sum(g.dyh * g.yh)) < 1e-14
    # iu weights form a partition of unity on the planes where they are used (2..ny-1)
    assert np.abs(g.iu[:, 1:ny - 1].sum(axis=0) - 1.0).max() < 1e-13


def test_wavenumbers():
    g = O.make_grid(8, 9, 8, 12, 12, 2.0, 2.0, 100.0)
    assert np.allclose(g.kx, np.arange(5) * 2 * np.pi / g.lx)
    assert np.allclose(g.kz, np.fft.fftfreq(8, 1.0 / 8) * 2 * np.pi / g.lz * np.where(np.arange(8) == 4, -1, 1))


# --------------------------------------------------------------------------- transforms
def _grid_t(nx=8, nx2=12, nz=8, nz2=12):
    return O.make_grid(nx, 9, nz, nx2, nz2, 2.0, 2.0, 100.0)


def _phys_coords(g):
    x = g.lx * np.arange(g.nx2) / g.nx2
    z = g.lz * np.arange(g.nz2) / g.nz2
    return x[None, :], z[:, None]


@pytest.mark.parametrize("dims", [(8, 12, 8, 12), (8, 8, 8, 8), (6, 12, 10, 15)])
def test_c2r_single_mode_closed_form(dims):
    """Physical field of one stored coefficient (FFTW conventions,
    PFC/transform.f90:339-409): 2 Re(c e^{i(kx x + kz z)}) for an interior
    mode; the x-Nyquist column is halved when nx2 > nx, the z-Nyquist entry is
    split into +-nz/2 halves when nz2 > nz."""
    nx, nx2, nz, nz2 = dims
    if nz2 % 2 and False:
        pass
    g = _grid_t(nx, nx2, nz, nz2)
    x, z = _phys_coords(g)
    rng = np.random.default_rng(0)
    c = complex(rng.standard_normal(), rng.standard_normal())
    for (ix, iz) in [(1, 2), (1, nz - 2), (0, 3), (g.nxh - 1, 1), (2, nz // 2), (0, 0)]:
        C = np.zeros((g.nxh, nz), complex)
        C[ix, iz] = c
        got = O.xz_c2r(g, C)
        kx, kz = g.kx[ix], g.kz[iz]
        e = np.exp(1j * (kx * x + kz * z))
        xnyq = (ix == g.nxh - 1)
        znyq = (iz == nz // 2)
        if ix == 0:
            if iz == 0:
                ref = np.full_like(got, c.real)
            elif znyq:
                # kx = 0, kz-Nyquist: (c e^{i kz z} + conj(c) e^{-i kz z})/2 = Re(c e^{ikz z})
                ref = (c * e).real if nz2 > nz else (c.real * np.cos(kz * z) + 0 * x)
            else:
                # c2r keeps the real part of the z-transform of the kx=0 column
                ref = (c * e).real
        else:
            fac = 2.0
            if xnyq and nx2 > nx:
                fac = 1.0
            if xnyq and nx2 == nx:
                # true Nyquist of the c2r: only Re of the z-transformed bin
                ref = ((c * np.exp(1j * kz * z)).real * np.cos(kx * x))
                assert np.abs(got - ref).max() < 1e-12
                continue
            if znyq and nz2 > nz:
                # half at +nz/2, half conj at -nz/2 (in z) for kx > 0
                ref = fac * 0.5 * ((c * np.exp(1j * (kx * x + kz * z))).real
                                   + (np.conj(c) * np.exp(1j * (kx * x - kz * z))).real)
            else:
                ref = fac * (c * e).real
        assert np.abs(got - ref).max() < 1e-12, (ix, iz)


@pytest.mark.parametrize("dims", [(8, 12, 8, 12), (16, 24, 8, 12), (8, 12, 8, 8), (8, 8, 8, 12)])
def test_round_trip_identity(dims):
    """xz_r2c(xz_c2r(C)) == C on the stored modes when the kx = 0 column is
    Hermitian in kz (the fold/double of :327-333 undo the split/halve of
    :349-357)."""
    nx, nx2, nz, nz2 = dims
    g = _grid_t(nx, nx2, nz, nz2)
    rng = np.random.default_rng(1)
    C = rng.standard_normal((3, g.nxh, nz)) + 1j * rng.standard_normal((3, g.nxh, nz))
    for zz in range(1, nz // 2):
        C[:, 0, nz - zz] = np.conj(C[:, 0, zz])
    C[:, 0, 0] = C[:, 0, 0].real
    C[:, 0, nz // 2] = C[:, 0, nz // 2].real
    if nx2 == nx or nz2 == nz:   # the x-Nyquist column without z padding: see below
        C[:, g.nxh - 1] = 0.0
    if nz2 == nz:   # undealiased z: the Nyquist row is treated separately below
        C[:, :, nz // 2] = 0.0
    back = O.xz_r2c(g, O.xz_c2r(g, C))
    assert rel_l2(back, C) < 1e-14


def test_undealiased_nyquist_conventions():
    """Without padding (nz2 == nz, nx2 == nx) the fold of PFC/transform.f90:327-328
    and the doubling of :331-333 return twice the real part of a z-Nyquist
    coefficient and twice the x-Nyquist coefficient."""
    g = _grid_t(8, 8, 8, 8)
    c = 0.7 - 0.4j
    C = np.zeros((g.nxh, g.nz), complex)
    C[1, g.nz // 2] = c
    back = O.xz_r2c(g, O.xz_c2r(g, C))
    ref = np.zeros_like(C)
    ref[1, g.nz // 2] = 2 * c.real
    assert np.abs(back - ref).max() < 1e-14
    C = np.zeros((g.nxh, g.nz), complex)
    C[g.nxh - 1, 0] = 0.3
    back = O.xz_r2c(g, O.xz_c2r(g, C))
    ref = np.zeros_like(C)
    ref[g.nxh - 1, 0] = 0.6
    assert np.abs(back - ref).max() < 1e-14
    # the c2r halving of the x-Nyquist column sits inside the nz2 > nz branch
    # (PFC/transform.f90:345-357), so with nx2 > nx but nz2 == nz it is skipped
    # while r2c still doubles it
    g = _grid_t(8, 12, 8, 8)
    C = np.zeros((g.nxh, g.nz), complex)
    C[g.nxh - 1, 1] = c
    back = O.xz_r2c(g, O.xz_c2r(g, C))
    assert abs(back[g.nxh - 1, 1] - 2 * c) < 1e-14


def test_divergence_of_fluxes_closed_form():
    """div = ikx F(fx) + F(fy) + ikz F(fz) (PFC/transform.f90:573-690) for
    band-limited fields whose derivatives are known analytically."""
    g = _grid_t(8, 12, 8, 12)
    x, z = _phys_coords(g)
    a1, b1 = g.kx[1], g.kz[2]
    a2, b2 = g.kx[2], g.kz[g.nz - 1]
    fx = np.cos(a1 * x + b1 * z) + 0.3 * np.sin(a2 * x)
    fy = 0.7 * np.cos(a2 * x + b2 * z)
    fz = np.sin(a1 * x - b2 * z) + 0.2 * np.cos(b1 * z)
    dfx = -a1 * np.sin(a1 * x + b1 * z) + 0.3 * a2 * np.cos(a2 * x)
    dfz = b2 * -np.cos(a1 * x - b2 * z) - 0.2 * b1 * np.sin(b1 * z)
    got = O.div_fluxes(g, fx, fy, fz)
    ref = O.xz_r2c(g, dfx + fy + dfz)
    assert np.abs(got - ref).max() < 1e-13


def test_small_transforms():
    """small_fft_* (PFC/transform.f90:950-1142): plain nx x nz DFTs scaled by
    1/(nx nz) forward; round trip and a single-mode closed form."""
    g = _grid_t(8, 12, 8, 12)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((2, g.nz, g.nx))
    assert rel_l2(O.small_c2r(g, O.small_r2c(g, f)), f) < 1e-14
    C = np.zeros((g.nxh, g.nz), complex)
    C[1, 3] = 0.4 - 0.1j
    xs = g.lx * np.arange(g.nx)[None, :] / g.nx
    zs = g.lz * np.arange(g.nz)[:, None] / g.nz
    ref = 2.0 * ((0.4 - 0.1j) * np.exp(1j * (g.kx[1] * xs + g.kz[3] * zs))).real
    assert np.abs(O.small_c2r(g, C) - ref).max() < 1e-13


# --------------------------------------------------------------------------- solvers
def _dense_helmholtz(g, n, alpha, diff, d2, top, bot, f):
    A = np.zeros((n, n))
    A[0, 0] = 1.0
    A[n - 1, n - 1] = 1.0
    for i in range(1, n - 1):
        A[i, i - 1] = -diff * d2[0, i]
        A[i, i] = -diff * d2[1, i] + alpha
        A[i, i + 1] = -diff * d2[2, i]
    rhs = f.copy()
    rhs[0] = bot
    rhs[n - 1] = top
    return np.linalg.solve(A, rhs)


def test_helmholtz_vs_dense():
    g = O.make_grid(8, 33, 8, 12, 12, 2.0, 2.0, 100.0)
    rng = np.random.default_rng(3)
    for n, d2 in ((g.ny + 1, g.d2h), (g.ny, g.d2n)):
        f = rng.standard_normal(n) + 1j * rng.standard_normal(n)
        got = O.solve_helmholtz(g, n, f, 1.7, 2e-3, d2, 0.3 + 0.1j, -0.2j)
        ref = _dense_helmholtz(g, n, 1.7, 2e-3, d2, 0.3 + 0.1j, -0.2j, f)
        assert rel_l2(got, ref) < 1e-13


def test_projection_postcondition_and_idempotence():
    """strip_divergence (PFC/solve.f90:14-50) leaves a discretely
    divergence-free field, and projecting twice changes nothing."""
    g = O.make_grid(8, 33, 8, 12, 12, 2.0, 2.0, 100.0)
    rng = np.random.default_rng(4)
    U = rng.standard_normal((3, g.ny + 1, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.ny + 1, g.nxh, g.nz))
    U[:, 0] = 0.0
    U[0, g.ny] = 0.0
    U[2, g.ny] = 0.0
    U[1, g.ny - 1:] = 0.0
    P1, cp = O.strip_divergence(g, U, 0.3)
    d = O.divergence(g, P1)
    d[:, 0, 0] = 0.0
    assert np.abs(d).max() < 1e-12 * np.abs(U).max() * 1e3
    P2, cp2 = O.strip_divergence(g, P1, 0.3)
    assert rel_l2(P2, P1) < 1e-13
    # the Neumann wall condition on the correction (PFC/solve.f90:335-336)
    dm = g.dm
    wall = dm[0, 0] * cp[0] + dm[1, 0] * cp[1] + dm[2, 0] * cp[2]
    assert np.abs(wall[g.k2 != 0]).max() < 1e-10 * np.abs(cp).max()


# --------------------------------------------------------------------------- nonlinear term
def _nl_bruteforce(g, U):
    """Independent restatement of pass1::nonlinear as the reference's plane
    loop (PFC/pass1.f90:46-99): explicit sliding window over reference plane
    numbers, in-place overwrite of a copy of data1."""
    ny = g.ny
    data1 = U.copy()
    half = 0.5

    def phys(k, i):
        return O.xz_c2r(g, data1[i, k - 1])

    u = np.stack([phys(1, i) for i in range(3)])
    up = np.stack([phys(2, i) for i in range(3)])
    yflux = np.zeros_like(u)
    yflux[0] = u[1] * u[0]
    temp = half * (u[1] + up[1])
    yflux[1] = temp * temp
    yflux[2] = u[1] * u[2]
    u = up.copy()
    data1[:, 0] = 0.0

    def wall_parallel(k):
        if k < ny:
            fpu = u[1] * (half * (u[0] + up[0]))
            fpw = u[1] * (half * (u[2] + up[2]))
        else:
            fpu = u[1] * up[0]
            fpw = u[1] * up[2]
        yflux[0] = g.dh[k - 1] * (fpu - yflux[0])
        yflux[2] = g.dh[k - 1] * (fpw - yflux[2])
        data1[0, k - 1] = O.div_fluxes(g, u[0] * u[0], yflux[0], u[2] * u[0])
        data1[2, k - 1] = O.div_fluxes(g, u[0] * u[2], yflux[2], u[2] * u[2])
        yflux[0] = fpu
        yflux[2] = fpw

    def wall_normal(k):
        t = half * (u[1] + up[1])
        fp = t * t
        yflux[1] = g.dn[k - 1] * (fp - yflux[1])
        i1, i2 = g.iu[0, k - 1], g.iu[1, k - 1]
        data1[1, k - 1] = O.div_fluxes(g, u[1] * (i1 * u[0] + i2 * up[0]), yflux[1],
                                       u[1] * (i1 * u[2] + i2 * up[2]))
        yflux[1] = fp

    for k in range(2, ny):
        up = np.stack([phys(k + 1, i) for i in range(3)])
        wall_parallel(k)
        wall_normal(k)
        u = up.copy()
    up = up.copy()
    up[0] = O.xz_c2r(g, data1[0, ny])
    up[2] = O.xz_c2r(g, data1[2, ny])
    up[1] = u[1]
    wall_parallel(ny)
    data1[1, ny - 1] = 0.0
    data1[:, ny] = 0.0
    return data1


def test_nonlinear_matches_plane_loop():
    g = O.make_grid(8, 17, 8, 12, 12, 2.0, 2.0, 100.0, stretch=2.0)
    rng = np.random.default_rng(5)
    U = 0.1 * (rng.standard_normal((3, g.ny + 1, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.ny + 1, g.nxh, g.nz)))
    U[:, :, 0] = U[:, :, 0].real
    NL, _ = O.nonlinear(g, U, step=-1, rk_step=3)
    assert rel_l2(NL, _nl_bruteforce(g, U)) < 1e-14


def _analytic_nl(g):
    """Velocity u = a + b cos(t), v = c cos(t), w = d sin(t), t = alpha x + gamma z
    and its exact conservative nonlinear term on the staggered points."""
    al, ga = g.kx[1], g.kz[1]
    x = g.lx * np.arange(g.nx2)[None, None, :] / g.nx2
    z = g.lz * np.arange(g.nz2)[None, :, None] / g.nz2
    t = al * x + ga * z
    def prof(y):
        a = 1 - y ** 2; da = -2 * y
        b = 0.3 * (1 - y ** 2) ** 2; db = 0.3 * 2 * (1 - y ** 2) * (-2 * y)
        c = 0.2 * (1 - y ** 2) ** 2; dc = 0.2 * 2 * (1 - y ** 2) * (-2 * y)
        d = 0.1 * (1 - y ** 2); dd = 0.1 * (-2 * y)
        return [v[:, None, None] for v in (a, b, c, d, da, db, dc, dd)]
    def fields(y):
        a, b, c, d, *_ = prof(y)
        return a + b * np.cos(t), c * np.cos(t), d * np.sin(t)
    def nl(y):
        a, b, c, d, da, db, dc, dd = prof(y)
        ct, st = np.cos(t), np.sin(t)
        u, v, w = a + b * ct, c * ct, d * st
        ux, uy, uz = -al * b * st, da + db * ct, -ga * b * st
        vx, vy, vz = -al * c * st, dc * ct, -ga * c * st
        wx, wy, wz = al * d * ct, dd * st, ga * d * ct
        nu = 2 * u * ux + (uy * v + u * vy) + (uz * w + u * wz)
        nv = (vx * u + v * ux) + 2 * v * vy + (vz * w + v * wz)
        nw = (wx * u + w * ux) + (wy * v + w * vy) + 2 * w * wz
        return nu, nv, nw
    return fields, nl


def test_nonlinear_second_order_in_y():
    """The discrete NL term converges to the exact conservative form at
    second order under y refinement (pins fluxes, interpolation, y-differences
    and the boundary planes)."""
    errs = []
    for ny in (33, 65, 129):
        g = O.make_grid(8, ny, 8, 12, 12, 2.0, 2.0, 100.0, stretch=1.5)
        fields, nl = _analytic_nl(g)
        uh, _, wh = fields(g.yh)
        _, vy_, _ = fields(g.y)
        U = np.zeros((3, ny + 1, g.nxh, g.nz), complex)
        U[0] = O.xz_r2c(g, uh)
        U[2] = O.xz_r2c(g, wh)
        U[1, :ny] = O.xz_r2c(g, vy_)
        NL, _ = O.nonlinear(g, U, step=-1, rk_step=3)
        nu, _, nw = nl(g.yh)
        _, nv, _ = nl(g.y)
        ex = np.zeros_like(U)
        ex[0] = O.xz_r2c(g, nu)
        ex[2] = O.xz_r2c(g, nw)
        ex[1, :ny] = O.xz_r2c(g, nv)
        sl = slice(2, ny - 2)  # interior planes
        e = max(np.abs(NL[c, sl] - ex[c, sl]).max() for c in range(3))
        errs.append(e)
    r1, r2 = errs[0] / errs[1], errs[1] / errs[2]
    assert r1 > 3.0 and r2 > 3.0, errs


def test_linearized_is_the_bilinear_cross_term():
    """N(U + eps u) = N(U) + eps N_lin(u; U) + eps^2 N(u) exactly, with the
    linearized term of PFC/pass1_linearized.f90 about the physical U."""
    g = O.make_grid(8, 17, 8, 12, 12, 2.0, 2.0, 100.0, stretch=2.0)
    rng = np.random.default_rng(6)
    def rnd():
        V = rng.standard_normal((3, g.ny + 1, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.ny + 1, g.nxh, g.nz))
        V[:, :, 0] = V[:, :, 0].real
        return V
    Ub, u = rnd(), rnd()
    base = O.xz_c2r(g, Ub)
    eps = 0.37
    N1, _ = O.nonlinear(g, Ub + eps * u, step=-1, rk_step=3)
    N0, _ = O.nonlinear(g, Ub, step=-1, rk_step=3)
    Nu, _ = O.nonlinear(g, u, step=-1, rk_step=3)
    Nl, _ = O.nonlinear_linearized(g, u, base, step=-1, rk_step=3)
    assert rel_l2(N1, N0 + eps * Nl + eps ** 2 * Nu) < 1e-13


# --------------------------------------------------------------------------- time stepper
def _seeded(g, seed, amp):
    from helpers import synth
    return synth().laminar_perturbed(g.nx, g.ny, g.nz, g.yh, g.y, seed=seed, amp=amp, kmax=3.0)


def test_rk3_cn_second_order_in_time():
    """Self-convergence of the RK3/CN + projection step (PFC/pass2.f90):
    halving dt reduces the difference by ~4 (second order)."""
    g = O.make_grid(8, 17, 8, 12, 12, 2.0, 2.0, 50.0, stretch=1.5, fix_mass=1)
    U0 = _seeded(g, 7, 5e-2)
    T = 0.32
    sols = []
    for dt in (0.04, 0.02, 0.01, 0.005):
        st = O.initial_state(g, U0, dt=dt)
        for s in range(int(round(T / dt))):
            O.rk3_step(g, st, s + 1)
        sols.append(st.U)
    e = [rel_l2(sols[i], sols[i + 1]) for i in range(3)]
    assert e[0] / e[1] > 3.3 and e[1] / e[2] > 3.5, e


def test_invariants_over_steps():
    """Discrete divergence stays at round-off and the mass flux is held by
    the mean-mode pressure gradient (PFC/pass2.f90:356-375)."""
    g = O.make_grid(16, 33, 16, 24, 24, 2.0, 2.0, 2500.0, stretch=1.5, fix_mass=1)
    st = O.initial_state(g, _seeded(g, 8, 1e-2), dt=0.01)
    q0 = O.mass_flux(g, st.U)
    for s in range(1, 11):
        O.rk3_step(g, st, s)
        assert O.max_divergence(g, st.U) < 1e-12
    assert abs(O.mass_flux(g, st.U) - q0) < 1e-9
    # laminar Poiseuille: p_grad = -2/Re holds the profile (PFC/setup.f90:1384)
    assert abs(st.p_grad + 2.0 / g.re) < 1e-3 * 2.0 / g.re


def test_wall_control_law_fixed_point():
    """boundary (PFC/read_write.f90:1483): the bottom-wall u satisfies
    ctrl*u_w = DM(:,1).u - 2 (Robin law) on the physical wall plane."""
    g = O.make_grid(8, 17, 8, 12, 12, 2.0, 2.0, 100.0, stretch=2.0, ctrlu=0.5)
    st = O.initial_state(g, _seeded(g, 9, 1e-2), dt=0.01)
    O.boundary(g, st)
    uw = O.xz_c2r(g, st.cbc_bot[0])
    u2 = O.xz_c2r(g, st.U[0, 1]); u3 = O.xz_c2r(g, st.U[0, 2])
    dm = g.dm
    lhs = g.ctrlu * uw
    rhs = dm[0, 0] * uw + dm[1, 0] * u2 + dm[2, 0] * u3 - 2.0
    assert np.abs(lhs - rhs).max() < 1e-12


# --------------------------------------------------------------------------- Arnoldi / OS known answer
OS_FIXTURE = os.path.join(os.path.dirname(__file__), "golden", "os_block_eigs.npz")


def test_orr_sommerfeld_known_answer():
    """The leading Ritz value of the linearized propagator about the laminar
    profile at Re = 7500 in the (kx = 1, kz = 0) block (written by
    tests/golden/make_os_fixture.py from the oracle) is the least-stable
    Orr-Sommerfeld mode: phase speed c_r = 0.24989154, the value the reference
    hard-codes (PFC/setup.f90:1387), and the textbook growth rate
    alpha c_i = 0.00223497, both to the second-order accuracy of the
    97-point stretched grid."""
    f = np.load(OS_FIXTURE)
    cr, ci = f["cr"], f["ci"]
    i = int(np.argmax(cr))
    assert abs(abs(ci[i]) - 0.24989154) < 2e-3
    assert abs(cr[i] - 0.00223497) < 0.1 * 0.00223497
    # the eigenvalue map of the fixture is eig_map of its Ritz values
    cr2, ci2 = O.eig_map(f["lam"], float(f["t_total"]))
    np.testing.assert_allclose(cr2, cr, rtol=0, atol=1e-15)


@pytest.mark.slow
@pytest.mark.skipif(os.environ.get("CHANFLOW_SLOW_TESTS") != "1", reason="about 80 s; set CHANFLOW_SLOW_TESTS=1")
def test_orr_sommerfeld_fixture_reproducible():
    import importlib.util
    spec = importlib.util.spec_from_file_location("mk_os", os.path.join(os.path.dirname(__file__), "golden",
                                                                        "make_os_fixture.py"))
    mk = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mk)
    g = mk.grid()
    base = O.laminar_baseflow(g)
    dt = O.linearized_dt(g)
    tend = O.linearized_tend(dt)
    lam = mk.leading_eigs(g, lambda v: O.compute_time_deriv(g, v, base, dt, tend)[0])
    np.testing.assert_allclose(lam, np.load(OS_FIXTURE)["lam"], rtol=1e-9)


# ---------------------------------------------------------------------------
# diagnostics (l2_norm, per-mode energy, plane moments)
# ---------------------------------------------------------------------------
def _diag_grid():
    return O.make_grid(16, 17, 12, 24, 18, 2.0, 2.0, 1000.0, stretch=1.5, fix_mass=1)


def test_mode_energy_parseval():
    """Sum of the per-mode energies (kx > 0 counted twice) equals the energy
    integrated in physical space on the nx x nz grid (no Nyquist content)."""
    g = _diag_grid()
    rng = np.random.default_rng(5)
    U = rng.standard_normal((3, g.ny + 1, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.ny + 1, g.nxh, g.nz))
    U[:, :, -1, :] = 0.0                       # x-Nyquist
    U[:, :, :, g.nz // 2] = 0.0                # z-Nyquist
    U[:, :, 0, :] = 0.5 * (U[:, :, 0, :] + np.conj(U[:, :, 0, (-np.arange(g.nz)) % g.nz]))  # Hermitian kx = 0
    ke = O.mode_energy(g, U)
    w = np.full(g.nxh, 2.0)
    w[0] = 1.0
    spectral = float(np.sum(w[:, None] * ke))
    phys = O.small_c2r(g, U.reshape(-1, g.nxh, g.nz)).reshape(3, g.ny + 1, g.nz, g.nx)
    t = phys[0] ** 2 + phys[2] ** 2
    vi = np.empty_like(phys[1])
    vi[0] = phys[1, 0]
    vi[1:g.ny] = 0.5 * (phys[1, 1:g.ny] + phys[1, 0:g.ny - 1])
    vi[g.ny] = phys[1, g.ny - 1]
    t = t + vi ** 2
    physical = float(np.sum(g.dyh[:, None, None] * t) / (g.nx * g.nz))
    assert abs(spectral - physical) <= 1e-12 * physical


def test_plane_moments_closed_form():
    """u = a cos x + m, v = b cos x, w = c sin 2x on every plane: variance
    a^2/2, third moment 0, fourth 3a^4/8, <u'v'> = ab/2."""
    g = _diag_grid()
    a, b, c, mval = 0.7, -1.3, 0.4, 0.25
    U = np.zeros((3, g.ny + 1, g.nxh, g.nz), np.complex128)
    U[0, :, 0, 0] = mval
    U[0, :, 1, 0] = a / 2
    U[1, :, 1, 0] = b / 2
    U[2, :, 2, 0] = -0.5j * c
    M = O.plane_moments(g, U)
    np.testing.assert_allclose(M[0], mval, rtol=1e-14)
    np.testing.assert_allclose(M[1:3], 0.0, atol=1e-15)
    np.testing.assert_allclose(M[3], a * a / 2, rtol=1e-13)
    np.testing.assert_allclose(M[4], b * b / 2, rtol=1e-13)
    np.testing.assert_allclose(M[5], c * c / 2, rtol=1e-13)
    np.testing.assert_allclose(M[6:9], 0.0, atol=1e-14)
    np.testing.assert_allclose(M[9], 3 * a ** 4 / 8, rtol=1e-13)
    np.testing.assert_allclose(M[10], 3 * b ** 4 / 8, rtol=1e-13)
    np.testing.assert_allclose(M[12], a * b / 2, rtol=1e-13)
    np.testing.assert_allclose(M[13], b * b / 2, rtol=1e-13)
    np.testing.assert_allclose(M[14], 0.0, atol=1e-15)


def test_l2_norm_closed_form():
    """u = 1 - y(j)^2 + a cos x on every plane, v = w = 0:
    l2 = sqrt(a^2/2 * sum_j dy_j / 2)."""
    g = _diag_grid()
    a = 0.3
    yext = np.concatenate([g.y, g.y[-1:]])
    U = np.zeros((3, g.ny + 1, g.nxh, g.nz), np.complex128)
    U[0, :, 0, 0] = 1.0 - yext ** 2
    U[0, :, 1, 0] = a / 2
    l2 = O.l2_norm(g, U)
    assert abs(l2 - math.sqrt(a * a / 2 * float(np.sum(g.dy)) / 2)) <= 1e-14


@pytest.mark.parametrize("dims", [(32, 32, 48, 48), (48, 32, 72, 48), (32, 16, 36, 18), (64, 64, 64, 64)])
def test_physical_round_trip_is_the_kx0_hermitian_projection(dims):
    """xz_r2c(xz_c2r(C)) on a dealiased grid (nx2 > nx, nz2 > nz) keeps every
    coefficient except those of the kx = 0 column, of which it keeps the
    Hermitian part: C(0, kz) -> (C(0, kz) + conj C(0, -kz)) / 2, the real part
    at kz = 0, the z-Nyquist coefficient as it is.  The device applies the
    (affine, pointwise) wall feedback law to Fourier coefficients with this
    projection instead of through the physical planes; on an undealiased grid
    (the last case) the round trip is not this projection, and the device
    keeps the physical path there."""
    nx, nz, nx2, nz2 = dims
    g = O.make_grid(nx, 17, nz, nx2, nz2, 2.0, 2.0, 2800.0, stretch=2.75)
    rng = np.random.default_rng(3)
    C = rng.standard_normal((3, g.nxh, g.nz)) + 1j * rng.standard_normal((3, g.nxh, g.nz))
    R = O.xz_r2c(g, O.xz_c2r(g, C))
    P = C.copy()
    h = nz // 2
    P[:, 0, 0] = C[:, 0, 0].real
    for iz in range(1, nz):
        if iz != h:
            P[:, 0, iz] = 0.5 * (C[:, 0, iz] + np.conj(C[:, 0, nz - iz]))
    err = np.abs(R - P).max()
    if nx2 > nx and nz2 > nz:
        assert err <= 1e-14
    else:
        assert err > 1e-3
