"""Pins of the oracle's rhs_solve sequencing (CPU only) that come from the
mathematics of the scheme rather than from the oracle's own formulas, and a
mutation test showing that plausible slips in those lines fail them.

1. Pressure update (Brown-Cortez-Minion rotational form,
   PFC/pass2.f90:307-309: p += alpha*phi - vis_imp*D2H phi).  Started on an
   eigenmode of the semi-discrete Stokes operator (the reference's own
   staggered y-operators, continuous in time, computed here by a dense
   generalized eigensolve independent of the time stepper) with its exact
   pressure, the rotational update keeps the pressure on the eigenmode's
   pressure vector to round-off at every dt, and its amplitude tends to
   p_hat exp(lambda (t - dt/6)) (the last substep's centre) at order > 1.4.
   The standard incremental update p += phi leaves that vector (the pressure
   boundary layer of non-rotational pressure-correction schemes).

2. Mean-mode pressure gradient (compute_p_grad, PFC/pass2.f90:356-375, with
   d_um_dy lagged from the end of the previous substep, :84-89).  For the
   Stokes mean mode the discrete flux budget of one substep is, exactly,
       Q_new - Q_old = d_r dt/Re (S(u_new) - S(u_old)),
   S(u) = sum_j dyh_j (D2H u)_j over the interior rows (= the wall-shear
   difference by telescoping): the explicit part c_r of the viscous flux and
   the lagged p_grad cancel, only the implicit part's change in wall shear is
   left.  Taking the wall shear from any other field breaks the identity.
"""
import os
import sys
import types

import numpy as np
import pytest
import scipy.linalg as sla

from oracle import chan_oracle as O

ORACLE_SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "chan_oracle.py")


# --------------------------------------------------------------------------- semi-discrete Stokes eigenmodes
def _tri(d, n):
    """Interior block (rows 1..n of the 0-based column, walls zero) of a
    three-point y operator d[3, :]."""
    A = np.zeros((n, n))
    for i in range(n):
        j = i + 1
        A[i, i] = d[1, j]
        if i > 0:
            A[i, i - 1] = d[0, j]
        if i < n - 1:
            A[i, i + 1] = d[2, j]
    return A


def stokes_eigenmode(g, ix, iz):
    """Least-damped eigenmode with a pressure of the semi-discrete Stokes
    problem of Fourier mode (ix, iz):
        du/dt = (D2H - k^2) u/Re - ikx p,  dw/dt = (D2H - k^2) w/Re - ikz p,
        dv/dt = (D2N - k^2) v/Re - dn (p_{j+1} - p_j),
        ikx u + dh (v_j - v_{j-1}) + ikz w = 0,
    u, w on the cell centres 1..ny-1 and v on the faces 1..ny-2 (0-based;
    walls zero), p on 1..ny-1.  Returns (lambda, u, v, w, p)."""
    ny, re = g.ny, g.re
    k2 = g.kx[ix] ** 2 + g.kz[iz] ** 2
    ikx, ikz = g.ikx[ix], g.ikz[iz]
    nu, nv, npp = ny - 1, ny - 2, ny - 1
    N = 2 * nu + nv
    L = np.zeros((N, N), complex)
    Lu = (_tri(g.d2h, nu) - k2 * np.eye(nu)) / re
    L[:nu, :nu] = Lu
    L[nu:nu + nv, nu:nu + nv] = (_tri(g.d2n, nv) - k2 * np.eye(nv)) / re
    L[nu + nv:, nu + nv:] = Lu
    G = np.zeros((N, npp), complex)
    D = np.zeros((npp, N), complex)
    for i in range(nu):
        G[i, i] = ikx
        G[nu + nv + i, i] = ikz
        D[i, i] = ikx
        D[i, nu + nv + i] = ikz
    for i in range(nv):
        j = i + 1
        G[nu + i, j] += g.dn[j]
        G[nu + i, j - 1] -= g.dn[j]
    for i in range(npp):
        j = i + 1
        if j <= ny - 2:
            D[i, nu + j - 1] += g.dh[j]
        if j >= 2:
            D[i, nu + j - 2] -= g.dh[j]
    A = np.block([[L, -G], [D, np.zeros((npp, npp))]])
    B = np.block([[np.eye(N), np.zeros((N, npp))], [np.zeros((npp, N + npp))]])
    lam, V = sla.eig(A, B)
    ok = np.isfinite(lam) & (np.abs(lam) < 1e6)
    lam, V = lam[ok], V[:, ok]
    share = np.linalg.norm(V[N:], axis=0) / np.linalg.norm(V, axis=0)
    cand = np.where(share > 1e-3)[0]
    i = cand[np.argmax(lam.real[cand])]
    v = V[:, i]
    return lam[i], v[:nu], v[nu:nu + nv], v[nu + nv:N], v[N:]


def pressure_update_residuals(Om, ny=33, re=20.0, stretch=1.5, ix=1, iz=0, steps=(8, 16, 32, 64)):
    """Run the Stokes eigenmode (nonlinear term zero) through Om.rhs_solve
    over one e-folding time.  Returns (collinearity residuals of P with the
    mode's pressure, amplitude errors against exp(lambda (T - dt/6)))."""
    g = Om.make_grid(8, ny, 8, 12, 12, 2.0, 2.0, re, stretch=stretch, fix_mass=0)
    lam, u, v, w, p = stokes_eigenmode(g, ix, iz)
    T = 1.0 / abs(lam.real)
    U = np.zeros((3, ny + 1, g.nxh, g.nz), complex)
    U[0, 1:ny, ix, iz] = u
    U[1, 1:ny - 1, ix, iz] = v
    U[2, 1:ny, ix, iz] = w
    P0 = np.zeros((ny + 1, g.nxh, g.nz), complex)
    P0[1:ny, ix, iz] = p
    col, amp = [], []
    for n in steps:
        dt = T / n
        st = Om.State(U=np.zeros_like(U), R=U.copy(), P=P0.copy(), dt=dt, p_grad=0.0, d_um_dy=np.zeros(ny))
        for _ in range(n):
            for rk in (1, 2, 3):
                st.U = np.zeros_like(U)      # Stokes: no nonlinear term
                Om.rhs_solve(g, st, rk)
        Pn = st.P[1:ny, ix, iz]
        c = np.vdot(p, Pn) / np.vdot(p, p)
        col.append(float(np.linalg.norm(Pn - c * p) / np.linalg.norm(Pn)))
        amp.append(float(abs(c - np.exp(lam * (T - dt / 6.0))) / abs(np.exp(lam * T))))
    return col, amp


def flux_budget_residual(Om, ny=33, re=20.0, dt=0.05, steps=3):
    """Largest |Q_new - Q_old - d_r dt/Re (S(u_new) - S(u_old))| over the
    substeps of a Stokes mean-flow transient under constant mass flux."""
    g = Om.make_grid(8, ny, 8, 12, 12, 2.0, 2.0, re, stretch=1.5, fix_mass=1)
    y = g.yh
    U = np.zeros((3, ny + 1, g.nxh, g.nz), complex)
    U[0, :, 0, 0] = (1 - y ** 2) + 0.3 * np.sin(np.pi * y) * (1 - y ** 2) + 0.2 * (1 - y ** 2) ** 3
    U[0, 0, 0, 0] = U[0, ny, 0, 0] = 0.0

    def S(u):
        return float(np.sum(g.dyh[1:ny] * O.d2_dy(u, g.d2h)[1:ny].real))

    st = Om.State(U=np.zeros_like(U), R=U.copy(), P=np.zeros((ny + 1, g.nxh, g.nz), complex), dt=dt,
                  p_grad=0.0, d_um_dy=Om.d_um_dy_of(g, U[0, :, 0, 0]))
    uold = U[0, :, 0, 0].copy()
    worst = 0.0
    for _ in range(steps):
        for rk in (1, 2, 3):
            st.U = np.zeros_like(U)
            Om.rhs_solve(g, st, rk)
            unew = st.U[0, :, 0, 0].copy()
            dq = float(np.sum(g.dyh * unew.real)) - float(np.sum(g.dyh * uold.real))
            pred = O.RKD[rk - 1] * dt / re * (S(unew) - S(uold))
            worst = max(worst, abs(dq - pred))
            uold = unew
    return worst


# --------------------------------------------------------------------------- the pins
@pytest.mark.parametrize("ny,re,stretch,ix,iz", [(33, 20.0, 1.5, 1, 0), (33, 20.0, 1.5, 1, 1), (17, 100.0, 2.0, 2, 1)])
def test_pressure_update_keeps_stokes_mode_pressure(ny, re, stretch, ix, iz):
    col, amp = pressure_update_residuals(O, ny, re, stretch, ix, iz, steps=(8, 16, 32))
    assert max(col) < 1e-11, col
    orders = [np.log2(amp[i] / amp[i + 1]) for i in range(len(amp) - 1)]
    assert min(orders) > 1.4, (amp, orders)


def test_stokes_eigenmode_is_an_eigenmode():
    """The dense eigensolve itself: the mode is divergence free with the
    oracle's divergence, and lambda u = L u - G p holds with the oracle's
    y-operators and pressure gradient."""
    g = O.make_grid(8, 33, 8, 12, 12, 2.0, 2.0, 20.0, stretch=1.5, fix_mass=0)
    ix, iz = 1, 1
    lam, u, v, w, p = stokes_eigenmode(g, ix, iz)
    ny = g.ny
    U = np.zeros((3, ny + 1, g.nxh, g.nz), complex)
    U[0, 1:ny, ix, iz] = u
    U[1, 1:ny - 1, ix, iz] = v
    U[2, 1:ny, ix, iz] = w
    P = np.zeros((ny + 1, g.nxh, g.nz), complex)
    P[1:ny, ix, iz] = p
    assert np.abs(O.divergence(g, U)).max() < 1e-12 * np.abs(U).max()
    k2 = g.k2
    lap = np.zeros_like(U)
    lap[0] = (O.d2_dy(U[0], g.d2h) - k2 * U[0]) / g.re
    lap[1, :ny] = (O.d2_dy(U[1, :ny], g.d2n) - k2 * U[1, :ny]) / g.re
    lap[2] = (O.d2_dy(U[2], g.d2h) - k2 * U[2]) / g.re
    rhs = O.subtract_pressure_gradient(g, lap, P, 1.0)
    res = lam * U - rhs
    assert np.abs(res[0, 1:ny]).max() < 1e-11 * abs(lam) * np.abs(U).max()
    assert np.abs(res[1, 1:ny - 1]).max() < 1e-11 * abs(lam) * np.abs(U).max()
    assert np.abs(res[2, 1:ny]).max() < 1e-11 * abs(lam) * np.abs(U).max()


def test_mean_mode_flux_budget_identity():
    assert flux_budget_residual(O) < 1e-13


def test_flux_budget_telescoping():
    """S(u) equals the wall-shear difference d_um_dy(ny) - d_um_dy(1) that
    compute_p_grad uses (the identity behind the flux budget)."""
    g = O.make_grid(8, 33, 8, 12, 12, 2.0, 2.0, 20.0, stretch=1.5, fix_mass=1)
    ny = g.ny
    u = np.cos(1.1 * g.yh) + 0.4 * g.yh ** 3 + 0.0j
    s = float(np.sum(g.dyh[1:ny] * O.d2_dy(u, g.d2h)[1:ny].real))
    d = O.d_um_dy_of(g, u)
    assert abs(s - (d[ny - 1] - d[0])) < 1e-12


# --------------------------------------------------------------------------- mutation test
def _mutant(old, new):
    src = open(ORACLE_SRC).read()
    assert src.count(old) == 1, old
    name = "chan_oracle_mutant"
    m = types.ModuleType(name)
    m.__file__ = ORACLE_SRC
    sys.modules[name] = m  # the dataclasses of the module look themselves up there
    try:
        exec(compile(src.replace(old, new), ORACLE_SRC + " (mutant)", "exec"), m.__dict__)
    finally:
        del sys.modules[name]
    return m


_BCM = "cp_new = cp + alpha2[None] * ch - vis_imp * d2_dy(ch, g.d2h)"
_PGRAD = "p_grad = -tmpv + 0.5 * (st.d_um_dy[ny - 1] - st.d_um_dy[0]) / re"
SLIPS = {
    # pressure update, PFC/pass2.f90:307-309
    "pressure update without the rotational terms": ("pressure", _BCM, "cp_new = cp + ch"),
    "pressure update without alpha": ("pressure", _BCM, "cp_new = cp + ch - vis_imp * d2_dy(ch, g.d2h)"),
    "pressure update with the viscous sign flipped": (
        "pressure", _BCM, "cp_new = cp + alpha2[None] * ch + vis_imp * d2_dy(ch, g.d2h)"),
    "pressure update with D2N for D2H": ("pressure", _BCM, "cp_new = cp + alpha2[None] * ch - vis_imp * "
                                         "np.concatenate([d2_dy(ch[:-1], g.d2n), ch[-1:] * 0.0])"),
    # compute_p_grad, PFC/pass2.f90:84-89,356-375
    "p_grad from the current right-hand side": (
        "flux", _PGRAD, "_d = d_um_dy_of(g, cv[0, :, 0, 0]); p_grad = -tmpv + 0.5 * (_d[ny - 1] - _d[0]) / re"),
    "p_grad from the RK register": (
        "flux", _PGRAD, "_d = d_um_dy_of(g, st.R[0, :, 0, 0]); p_grad = -tmpv + 0.5 * (_d[ny - 1] - _d[0]) / re"),
    "p_grad without the factor 1/2": (
        "flux", _PGRAD, "p_grad = -tmpv + 1.0 * (st.d_um_dy[ny - 1] - st.d_um_dy[0]) / re"),
    "d_um_dy never updated": ("flux", "st.d_um_dy = d_um_dy_of(g, cv[0, :, 0, 0])", "pass"),
    "p_grad with rkb for the pressure coefficient": (
        "flux", "cvm[0] = cvm[0] - math.cos(g.beta) * rk_press * p_grad",
        "cvm[0] = cvm[0] - math.cos(g.beta) * coef * p_grad"),
}


@pytest.mark.parametrize("name", sorted(SLIPS))
def test_slip_is_caught(name):
    """Each named slip, applied to an in-memory copy of the oracle, fails the
    pin aimed at its line while the unmodified oracle passes it."""
    kind, old, new = SLIPS[name]
    M = _mutant(old, new)
    if kind == "pressure":
        try:
            col, amp = pressure_update_residuals(M, steps=(8, 16))
        except (FloatingPointError, np.linalg.LinAlgError):
            return
        bad = (not np.all(np.isfinite(col))) or max(col) > 1e-9
        assert bad, (name, col)
    else:
        r = flux_budget_residual(M)
        assert not (r < 1e-11), (name, r)
