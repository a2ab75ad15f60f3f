"""Pins of the channel-grid closure restatement (oracle/les_oracle.py, the
part the device's LES module mirrors): closed forms for the closure inputs
(linear shear, a single Fourier mode), the Smagorinsky stress of a shear, the
Germano identity vanishing for x-z uniform fields and its invariance under
scaling of the velocity, the network inputs' invariants, and second-order
convergence of the subgrid-stress divergence that the nonlinear term adds
(which fixes where each tau_ij sits on the staggered grid)."""
import math

import numpy as np
import pytest

from helpers import grid_cfg0, init_cfg0
from oracle import chan_oracle as O
from oracle import les_oracle as L


@pytest.fixture(scope="module")
def g0():
    return grid_cfg0(O)


def _shear(g, c):
    U = np.zeros((3, g.ny + 1, g.nxh, g.nz), dtype=np.complex128)
    U[0, :, 0, 0] = c * g.yh
    return U


def test_fields_linear_shear_and_smagorinsky(g0):
    c = 0.7
    Ph = L.les_fields(g0, _shear(g0, c))
    k = slice(1, g0.ny)
    np.testing.assert_allclose(Ph[1, k], c, rtol=0, atol=1e-13)          # du/dy
    for f in (0, 2, 3, 4, 5, 6, 7, 8):
        assert np.abs(Ph[f, k]).max() <= 1e-13, f
    np.testing.assert_allclose(Ph[9, k], c * g0.yh[k][:, None, None] + 0 * Ph[9, k], rtol=0, atol=1e-13)
    assert np.abs(Ph[:, 0]).max() == 0.0 and np.abs(Ph[:, g0.ny]).max() == 0.0
    cs2 = np.full(g0.ny + 1, 0.01)
    tau = L.tau_smagorinsky(g0, Ph, cs2)
    d2 = L.les_delta2(g0)
    # S_12 = c/2, |S| = |c|: tau_12 = -2 C D^2 |c| c/2
    np.testing.assert_allclose(tau[1, k], (-0.01 * d2[k] * abs(c) * c)[:, None, None] + 0 * tau[1, k],
                               rtol=1e-13, atol=0)
    for t in (0, 2, 3, 4, 5):
        assert np.abs(tau[t, k]).max() <= 1e-16, t


def test_fields_single_fourier_mode(g0):
    """u = a cos(kx x) on one plane: du/dx = -a kx sin(kx x), dw/dz = 0."""
    a, k = 0.3, 5
    U = np.zeros((3, g0.ny + 1, g0.nxh, g0.nz), dtype=np.complex128)
    U[0, k, 1, 0] = 0.5 * a
    Ph = L.les_fields(g0, U)
    x = np.arange(g0.nx2) * g0.lx / g0.nx2
    kx = 2 * math.pi / g0.lx
    np.testing.assert_allclose(Ph[9, k], np.broadcast_to(a * np.cos(kx * x), Ph[9, k].shape), atol=1e-14)
    np.testing.assert_allclose(Ph[0, k], np.broadcast_to(-a * kx * np.sin(kx * x), Ph[0, k].shape), atol=1e-13)
    assert np.abs(Ph[8, k]).max() <= 1e-15


def test_dynamic_coefficient_invariants(g0):
    # x-z uniform field: every test-filtered quantity equals itself, L = 0
    assert np.all(L.dynamic_coefficient(g0, _shear(g0, 0.4)) == 0.0)
    U = init_cfg0(O, g0, seed=3, amp=0.05)
    c1 = L.dynamic_coefficient(g0, U)
    assert np.all(np.isfinite(c1)) and np.all(c1 >= 0.0)
    assert c1[0] == 0.0 and c1[g0.ny] == 0.0
    assert c1.max() > 0.0
    # L and M both scale with the square of the velocity
    c2 = L.dynamic_coefficient(g0, 3.0 * U)
    np.testing.assert_allclose(c2, c1, rtol=1e-10, atol=1e-14)


def test_features_invariants(g0):
    U = init_cfg0(O, g0, seed=4, amp=0.05)
    X1, s1 = L.features(g0, L.les_fields(g0, U))
    X2, s2 = L.features(g0, L.les_fields(g0, 2.0 * U))
    assert X1.shape == ((g0.ny - 1) * g0.nz2 * g0.nx2, 10) and X1.dtype == np.float32
    np.testing.assert_allclose(X1, X2, rtol=1e-6, atol=1e-7)
    np.testing.assert_allclose(s2, 4.0 * s1, rtol=1e-12)
    n = np.sqrt((X1[:, :9].astype(np.float64) ** 2).sum(axis=1))
    assert np.all((np.abs(n - 1.0) < 1e-6) | (n == 0.0))
    assert X1[:, 9].min() >= 0.0 and X1[:, 9].max() <= 1.0


def test_tau_gate_picks_smagorinsky_for_backscatter(g0):
    U = init_cfg0(O, g0, seed=5, amp=0.05)
    Ph = L.les_fields(g0, U)
    cs2 = L.dynamic_coefficient(g0, U)
    X, sc = L.features(g0, Ph)
    n = X.shape[0]
    ts = L.tau_smagorinsky(g0, Ph, cs2)
    # learned stress = -3 x Smagorinsky (strong backscatter) everywhere: all gated
    Y = np.zeros((n, 6), np.float32)
    ok = sc > 0
    for t in range(6):
        Y[ok, t] = (-3.0 * ts[t, 1:g0.ny].reshape(-1)[ok] / sc[ok]).astype(np.float32)
    tau, nfb = L.tau_learned(g0, Ph, Y, sc, cs2, beta=1.0)
    active = (cs2[1:g0.ny] > 0).sum() * g0.nz2 * g0.nx2
    assert nfb >= active - np.count_nonzero(~ok)
    np.testing.assert_allclose(tau[:, 1:g0.ny][:, cs2[1:g0.ny] > 0], ts[:, 1:g0.ny][:, cs2[1:g0.ny] > 0])
    # learned stress = Smagorinsky's: nothing gated, tau unchanged
    for t in range(6):
        Y[ok, t] = (ts[t, 1:g0.ny].reshape(-1)[ok] / sc[ok]).astype(np.float32)
    tau2, nfb2 = L.tau_learned(g0, Ph, Y, sc, cs2, beta=1.0)
    assert nfb2 == 0
    np.testing.assert_allclose(tau2, ts, rtol=1e-6, atol=1e-12)


def _tau_field(g, kxn=1):
    """tau_ij = a_ij phi(yh) cos(kx x), phi = (1 - y^2)^2, zero on the walls."""
    x = np.arange(g.nx2) * g.lx / g.nx2
    cx = np.cos(2 * math.pi * kxn * x / g.lx)
    phi = (1.0 - g.yh ** 2) ** 2
    amp = np.array([0.3, 0.5, -0.2, 0.4, 0.7, -0.6])
    tau = amp[:, None, None, None] * phi[None, :, None, None] * cx[None, None, None, :] * np.ones((1, 1, g.nz2, 1))
    tau[:, 0] = 0.0
    tau[:, g.ny] = 0.0
    return tau, amp


@pytest.mark.parametrize("comp", [0, 1])
def test_tau_divergence_second_order(comp):
    """NL(U = 0, tau) -> div(tau) at second order in y (spectral in x, z):
    u: d/dx tau_11 + d/dy tau_12 at yh; v: d/dx tau_12 + d/dy tau_22 at y."""
    errs = []
    for ny in (33, 65, 129):
        g = O.make_grid(8, ny, 8, 12, 12, 2.0, 2.0, 1000.0, stretch=1.5)
        tau, a = _tau_field(g)
        NL, _ = O.nonlinear(g, np.zeros((3, g.ny + 1, g.nxh, g.nz), dtype=np.complex128), rk_step=2, tau=tau)
        ph = O.xz_c2r(g, NL[comp])
        x = np.arange(g.nx2) * g.lx / g.nx2
        kx = 2 * math.pi / g.lx
        if comp == 0:
            ks = np.arange(2, g.ny - 1)
            yv = g.yh[ks]
            ex = (-a[0] * kx * (1 - yv ** 2) ** 2)[:, None] * np.sin(kx * x)[None] + \
                 (a[1] * (-4 * yv * (1 - yv ** 2)))[:, None] * np.cos(kx * x)[None]
        else:
            ks = np.arange(2, g.ny - 2)
            yv = g.y[ks]
            ex = (-a[1] * kx * (1 - yv ** 2) ** 2)[:, None] * np.sin(kx * x)[None] + \
                 (a[3] * (-4 * yv * (1 - yv ** 2)))[:, None] * np.cos(kx * x)[None]
        errs.append(np.abs(ph[ks, 0, :] - ex).max())
    assert errs[0] / errs[1] > 3.0 and errs[1] / errs[2] > 3.0, errs
    assert errs[2] < 2e-3, errs


def test_test_filter_transfer_function(g0):
    """The trapezoidal test filter multiplies a Fourier mode exp(i(a x + b z))
    of the dealiased grid by (1 + cos a h_x)/2 (1 + cos b h_z)/2 and keeps a
    constant; a Nyquist mode in x is removed."""
    x = np.arange(g0.nx2) * g0.lx / g0.nx2
    z = np.arange(g0.nz2) * g0.lz / g0.nz2
    hx, hz = g0.lx / g0.nx2, g0.lz / g0.nz2
    for kx, kz in ((1, 0), (0, 2), (3, 5), (7, 1)):
        a, b = 2 * math.pi * kx / g0.lx, 2 * math.pi * kz / g0.lz
        f = np.cos(a * x[None, :] + b * z[:, None])[None]
        ref = 0.25 * (1 + math.cos(a * hx)) * (1 + math.cos(b * hz)) * f
        np.testing.assert_allclose(L.test_filter(f), ref, rtol=0, atol=1e-14)
    np.testing.assert_allclose(L.test_filter(np.full((1, g0.nz2, g0.nx2), 2.5)), 2.5, rtol=0, atol=1e-15)
    nyq = np.cos(math.pi * np.arange(g0.nx2))[None, None, :] * np.ones((1, g0.nz2, 1))
    assert np.abs(L.test_filter(nyq)).max() < 1e-15
