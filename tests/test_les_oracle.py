"""Pins of the closure's CPU restatement (oracle/les_oracle.py): the MLP
gradient against central finite differences in float64, GELU and its
derivative against their definitions, one Adam step against its closed
form, bfloat16 rounding on representable and tie values, the Smagorinsky
stress on linear fields."""
import math

import numpy as np
import pytest

from oracle import les_oracle as L


def test_gelu_definition_and_derivative():
    x = np.linspace(-6, 6, 241)
    g = L.gelu(x)
    phi = 0.5 * (1 + np.array([math.erf(v / math.sqrt(2)) for v in x]))
    np.testing.assert_allclose(g, x * phi, rtol=0, atol=1e-15)
    h = 1e-6
    fd = (L.gelu(x + h) - L.gelu(x - h)) / (2 * h)
    np.testing.assert_allclose(L.gelu_grad(x), fd, rtol=1e-7, atol=1e-9)
    assert L.gelu(0.0) == 0.0 and abs(L.gelu_grad(0.0) - 0.5) < 1e-16


def test_mlp_gradient_finite_differences():
    W, b = L.init_params(seed=3)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((24, 10))
    T = rng.standard_normal((24, 6))
    gW, gb = L.backward(W, b, X, T)
    h = 1e-6
    for l, (i, j) in [(0, (5, 3)), (1, (17, 100)), (2, (127, 0)), (3, (4, 77))]:
        Wp = [w.copy() for w in W]; Wp[l][i, j] += h
        Wm = [w.copy() for w in W]; Wm[l][i, j] -= h
        fd = (L.mse(L.forward(Wp, b, X), T) - L.mse(L.forward(Wm, b, X), T)) / (2 * h)
        assert abs(fd - gW[l][i, j]) <= 1e-7 * max(1.0, abs(fd)), (l, i, j)
    for l, i in [(0, 9), (1, 64), (3, 5)]:
        bp = [v.copy() for v in b]; bp[l][i] += h
        bm = [v.copy() for v in b]; bm[l][i] -= h
        fd = (L.mse(L.forward(W, bp, X), T) - L.mse(L.forward(W, bm, X), T)) / (2 * h)
        assert abs(fd - gb[l][i]) <= 1e-7 * max(1.0, abs(fd)), (l, i)


def test_adam_first_step_closed_form():
    """At t = 1 the bias-corrected step is lr * g/(|g| + eps), i.e. about lr*sign(g)."""
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -4.0, 1e-3])
    p1, m, v = L.adam_step(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(p1, p - 0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)


def test_bf16_rounding():
    vals = np.array([1.0, -2.5, 3.0e-3, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 65504.0], np.float32)
    r = L.to_bf16(vals)
    assert r[0] == 1.0 and r[1] == -2.5
    assert r[3] == 1.0                 # tie rounds to even (mantissa ...0)
    assert r[4] == 1.0 + 2 * 2.0 ** -7  # tie rounds up to even
    assert abs(r[2] - 3.0e-3) <= 3.0e-3 * 2.0 ** -8


def test_smagorinsky_linear_fields():
    """Pure shear du/dy = s: S_12 = s/2, |S| = |s|, tau_12 = -(cs D)^2 |s| s."""
    s, cs, D = 2.5, 0.17, 0.1
    G = np.zeros((1, 3, 3)); G[0, 0, 1] = s
    tau = L.smagorinsky_tau(G, np.array([D]), cs)
    np.testing.assert_allclose(tau[0], [0, -(cs * D) ** 2 * abs(s) * s, 0, 0, 0, 0], atol=1e-15)
    # plane strain du/dx = a, dv/dy = -a: |S| = 2|a|, tau_11 = -2 (cs D)^2 2|a| a
    a = 1.5
    G = np.zeros((1, 3, 3)); G[0, 0, 0] = a; G[0, 1, 1] = -a
    tau = L.smagorinsky_tau(G, np.array([D]), cs)
    np.testing.assert_allclose(tau[0], [-4 * (cs * D) ** 2 * abs(a) * a, 0, 0, 4 * (cs * D) ** 2 * abs(a) * a, 0, 0],
                               atol=1e-15)


# ---------------------------------------------------------------------------
# Filtered DNS (oracle/les_oracle.py: filter_spec, filter_dns)
# ---------------------------------------------------------------------------
def _filter_grids():
    from oracle import chan_oracle as O
    gf = O.make_grid(32, 33, 32, 48, 48, 2.0, 2.0, 2500.0, stretch=1.5)
    gc = O.make_grid(16, 17, 16, 24, 24, 2.0, 2.0, 2500.0, stretch=1.5)
    return O, gf, gc


def test_filter_grids_share_every_other_face():
    O, gf, gc = _filter_grids()
    assert np.allclose(gc.y, gf.y[::2], rtol=0, atol=1e-15)


def test_filter_keeps_divergence_zero():
    """The y restriction (height-weighted cell averages for u and w, shared
    faces for v) with a mode-wise x-z filter maps a divergence-free fine field
    onto a divergence-free coarse one: the coarse cell's flux balance is the
    sum of its two fine cells'."""
    O, gf, gc = _filter_grids()
    rng = np.random.default_rng(4)
    U = rng.standard_normal((3, gf.ny + 1, gf.nxh, gf.nz)) + 1j * rng.standard_normal((3, gf.ny + 1, gf.nxh, gf.nz))
    U = O.remove_initial_divergence(gf, U)
    assert O.max_divergence(gf, U) < 1e-10
    Uc = np.stack([L.filter_spec(gf, gc, U[0], 0), L.filter_spec(gf, gc, U[1], 2), L.filter_spec(gf, gc, U[2], 0)])
    assert np.abs(Uc).max() > 0.1
    assert O.max_divergence(gc, Uc) < 1e-10
    # unfiltered random data has divergence of order one: the check has teeth
    V = rng.standard_normal(U.shape) + 0j
    Vc = np.stack([L.filter_spec(gf, gc, V[0], 0), L.filter_spec(gf, gc, V[1], 2), L.filter_spec(gf, gc, V[2], 0)])
    assert O.max_divergence(gc, Vc) > 1.0


def test_filter_transfer_and_constants():
    """A single Fourier mode constant in y comes out multiplied by the box
    transfer (1 + 2 cos(kx dx))/3 (1 + 2 cos(kz dz))/3 of the fine grid; a
    constant passes unchanged; the coarse Nyquist modes are dropped."""
    O, gf, gc = _filter_grids()
    C = np.zeros((gf.ny + 1, gf.nxh, gf.nz), np.complex128)
    ix, kz = 3, -2
    C[:, ix, kz % gf.nz] = 0.7 - 0.2j
    C[:, 0, 0] = 1.5
    out = L.filter_spec(gf, gc, C, 0)
    f = (1 + 2 * np.cos(2 * np.pi * ix / gf.nx)) / 3 * (1 + 2 * np.cos(2 * np.pi * kz / gf.nz)) / 3
    assert np.allclose(out[:, ix, kz % gc.nz], f * (0.7 - 0.2j), rtol=0, atol=1e-15)
    assert np.allclose(out[:, 0, 0], 1.5, rtol=0, atol=1e-15)
    C2 = np.zeros_like(C)
    C2[:, gc.nxh - 1, 1] = 1.0
    C2[:, 1, gc.nz // 2] = 1.0
    o2 = L.filter_spec(gf, gc, C2, 0)
    assert np.abs(o2).max() == 0.0


def test_filter_stress_vanishes_for_cellwise_constant_mean_flow():
    """tau = F(u u) - F(u) F(u) is zero for a mean flow that is constant over
    each coarse cell (equal on its two fine cells): nothing is unresolved."""
    O, gf, gc = _filter_grids()
    U = np.zeros((3, gf.ny + 1, gf.nxh, gf.nz), np.complex128)
    prof = np.zeros(gf.ny + 1)
    for K in range(1, gc.ny):
        prof[2 * K - 1] = prof[2 * K] = 1.0 - gc.yh[K] ** 2
    U[0, :, 0, 0] = prof
    _, tau = L.filter_dns(gf, gc, U)
    assert np.abs(tau).max() < 1e-14
