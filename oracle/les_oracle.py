"""CPU restatement of the learned subgrid closure and its Smagorinsky
fallback.  TEST INFRASTRUCTURE ONLY: imported by tests/, never by the
library or its bindings.

The reference has no LES closure (its subgrid flags and dynamic-model
arrays are vestiges: PFC/flow.dat:19-22, PFC/data.f90:14-15, calls commented
out at PFC/main_mpi.f90:10-13,74,77-78), so this file is pinned by
mathematics only: the MLP gradients by central finite differences in
float64, GELU by its definition through erf, Adam by its published update
(Kingma & Ba 2015, Algorithm 1), the Smagorinsky stress by closed forms on
linear velocity fields.

Network: 10 inputs (the 9 filtered velocity-gradient components du_i/dx_j
and y+) -> 128 -> 128 -> 128 -> 6 outputs (tau_11, tau_12, tau_13, tau_22,
tau_23, tau_33), GELU on the hidden layers, linear output.
"""
from __future__ import annotations

import math

import numpy as np

LAYERS = (10, 128, 128, 128, 6)

_erf = np.vectorize(math.erf)


def gelu(x):
    """x * Phi(x) = x/2 (1 + erf(x / sqrt 2))."""
    x = np.asarray(x, dtype=np.float64)
    return 0.5 * x * (1.0 + _erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    """d/dx [x Phi(x)] = Phi(x) + x phi(x)."""
    x = np.asarray(x, dtype=np.float64)
    return 0.5 * (1.0 + _erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2.0 * math.pi)


def init_params(seed=0, layers=LAYERS):
    """He-style uniform initialisation: W[l] (n_out, n_in), b[l] (n_out)."""
    rng = np.random.default_rng(seed)
    W, b = [], []
    for n_in, n_out in zip(layers[:-1], layers[1:]):
        lim = math.sqrt(6.0 / n_in)
        W.append(rng.uniform(-lim, lim, (n_out, n_in)))
        b.append(rng.uniform(-0.1, 0.1, n_out))
    return W, b


def forward(W, b, X, keep=False):
    """X (n, 10) -> Y (n, 6); with keep, also the pre-activations Z[l] and
    the layer inputs A[l] (A[0] = X)."""
    A = [np.asarray(X, dtype=np.float64)]
    Z = []
    h = A[0]
    L = len(W)
    for l in range(L):
        z = h @ W[l].T + b[l]
        Z.append(z)
        h = gelu(z) if l < L - 1 else z
        if l < L - 1:
            A.append(h)
    return (h, Z, A) if keep else h


def mse(Y, T):
    """mean over points and outputs of (Y - T)^2."""
    return float(np.mean((Y - T) ** 2))


def backward(W, b, X, T):
    """Gradients of mse(forward(X), T) w.r.t. W and b (exact, float64)."""
    Y, Z, A = forward(W, b, X, keep=True)
    n, m = Y.shape
    d = 2.0 * (Y - T) / (n * m)
    L = len(W)
    gW, gb = [None] * L, [None] * L
    for l in range(L - 1, -1, -1):
        gW[l] = d.T @ A[l]
        gb[l] = d.sum(axis=0)
        if l > 0:
            d = (d @ W[l]) * gelu_grad(Z[l - 1])
    return gW, gb


def adam_step(p, g, m, v, t, lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, Algorithm 1); t counts from 1.
    Returns (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def to_bf16(x):
    """Round float32 values to bfloat16 (round to nearest even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(a.shape)


def forward_bf16(W, b, X):
    """Inference as the MFMA kernel computes it: inputs, weights and hidden
    activations rounded to bfloat16, products accumulated in float32 (here
    float64, which bounds the float32 accumulation error), biases in float32."""
    h = to_bf16(X).astype(np.float64)
    L = len(W)
    for l in range(L):
        z = h @ to_bf16(W[l]).astype(np.float64).T + np.asarray(b[l], np.float32).astype(np.float64)
        if l < L - 1:
            h = to_bf16(gelu(z).astype(np.float32)).astype(np.float64)
        else:
            h = z
    return h


# ---------------------------------------------------------------------------
# Smagorinsky fallback
# ---------------------------------------------------------------------------
def smagorinsky_tau(grad, delta, cs):
    """Deviatoric Smagorinsky stress tau_ij = -2 (cs delta)^2 |S| S_ij with
    S = (G + G^T)/2, |S| = sqrt(2 S_ij S_ij); grad (n, 3, 3) with
    grad[:, i, j] = du_i/dx_j, delta (n,).  Returns (n, 6) in the order
    11, 12, 13, 22, 23, 33."""
    G = np.asarray(grad, dtype=np.float64)
    S = 0.5 * (G + np.swapaxes(G, 1, 2))
    Smag = np.sqrt(2.0 * np.einsum("nij,nij->n", S, S))
    nu = (cs * np.asarray(delta, dtype=np.float64)) ** 2 * Smag
    idx = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    return np.stack([-2.0 * nu * S[:, i, j] for i, j in idx], axis=1)


# ---------------------------------------------------------------------------
# The closure on the channel grid (the device's LES module)
#
# Fields at the cell centres yh of planes 1..ny-1 (zero on the wall planes):
# G[0..8] = du_i/dx_j (row-major, i = u, v, w; j = x, y, z), G[9..11] = u, v,
# w; d/dx and d/dz spectral, v at yh the mean of its faces, du/dy and dw/dy
# central over yh(k-1), yh(k+1), dv/dy the face difference dh(k) (v(k) -
# v(k-1)).  Pinned by closed forms (linear shear, single Fourier modes), by
# the Germano identity vanishing for x-z uniform fields and by second-order
# convergence of the subgrid-stress divergence (tests/test_les_closure.py).
# ---------------------------------------------------------------------------
SYM = ((0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2))


def _co():
    from oracle import chan_oracle as co
    return co


def les_spectral_fields(g, U):
    """The 12 spectral fields [12, ny+1, nxh, nz] of the closure."""
    ny = g.ny
    G = np.zeros((12, ny + 1, g.nxh, g.nz), dtype=np.complex128)
    k = np.arange(1, ny)
    ikx = 1j * g.kx[:, None]
    ikz = 1j * g.kz[None, :]
    u, w = U[0, k], U[2, k]
    vm, v0 = U[1, k - 1], U[1, k]
    vc = 0.5 * (vm + v0)
    dyc = (1.0 / (g.hn[k] + g.hn[k - 1]))[:, None, None]
    dhk = g.dh[k][:, None, None]
    G[0, k] = ikx * u
    G[1, k] = (U[0, k + 1] - U[0, k - 1]) * dyc
    G[2, k] = ikz * u
    G[3, k] = ikx * vc
    G[4, k] = dhk * (v0 - vm)
    G[5, k] = ikz * vc
    G[6, k] = ikx * w
    G[7, k] = (U[2, k + 1] - U[2, k - 1]) * dyc
    G[8, k] = ikz * w
    G[9, k], G[10, k], G[11, k] = u, vc, w
    return G


def les_fields(g, U):
    """Physical closure inputs [12, ny+1, nz2, nx2]."""
    return _co().xz_c2r(g, les_spectral_fields(g, U))


def strain(gf):
    """S_ij in the order 11, 12, 13, 22, 23, 33 and |S| = sqrt(2 S:S) from
    gradient fields gf[9, ...]."""
    s = np.stack([gf[0], 0.5 * (gf[1] + gf[3]), 0.5 * (gf[2] + gf[6]), gf[4], 0.5 * (gf[5] + gf[7]), gf[8]])
    ss = s[0] * s[0] + s[3] * s[3] + s[5] * s[5] + 2.0 * (s[1] * s[1] + s[2] * s[2] + s[4] * s[4])
    return s, np.sqrt(2.0 * ss)


def contract(a, b):
    """sum_ij a_ij b_ij of symmetric tensors stored as 6 components."""
    return a[0] * b[0] + a[3] * b[3] + a[5] * b[5] + 2.0 * (a[1] * b[1] + a[2] * b[2] + a[4] * b[4])


def les_delta2(g):
    """Filter width squared per plane, D = (dx dz (y(k)-y(k-1)))^(1/3)."""
    d2 = np.zeros(g.ny + 1)
    dx, dz = g.lx / g.nx, g.lz / g.nz
    for k in range(1, g.ny):
        d = np.cbrt(dx * dz * (g.y[k] - g.y[k - 1]))
        d2[k] = d * d
    return d2


def les_ydist(g):
    """Wall distance 1 - |yh| per plane (zero on the walls)."""
    yd = np.zeros(g.ny + 1)
    k = np.arange(1, g.ny)
    yd[k] = 1.0 - np.abs(0.5 * (g.y[k - 1] + g.y[k]))
    return yd


def test_filter(f):
    """Test filter of the dynamic procedure: the trapezoidal rule (1/4, 1/2,
    1/4) along x, then along z, on the periodic dealiased grid (last two axes
    z, x), in that order of operations."""
    fx = 0.25 * np.roll(f, 1, axis=-1) + 0.5 * f + 0.25 * np.roll(f, -1, axis=-1)
    return 0.25 * np.roll(fx, 1, axis=-2) + 0.5 * fx + 0.25 * np.roll(fx, -1, axis=-2)


def dynamic_coefficient(g, U):
    """Dynamic Smagorinsky C(y): Germano identity L = C M with
    L_ij = T(u_i u_j) - T(u_i) T(u_j), M_ij = 2 D^2 (T(|S| S_ij) - 4 |S^| S^_ij)
    (test-to-grid filter width ratio 2), T the trapezoidal test filter in x
    and z of the closure fields on the dealiased grid, plane least squares
    <L:M>/<M:M> (Lilly 1992), clipped at zero."""
    ny = g.ny
    Ph = les_fields(g, U)
    s, mag = strain(Ph[:9])
    u = Ph[9:12]
    P = np.zeros((12,) + Ph.shape[1:])
    for t, (i, j) in enumerate(SYM):
        P[t] = u[i] * u[j]
        P[6 + t] = mag * s[t]
    TP = test_filter(P)
    TG = test_filter(Ph)
    st, magt = strain(TG[:9])
    ut = TG[9:12]
    d2 = les_delta2(g)
    cs2 = np.zeros(ny + 1)
    for k in range(1, ny):
        L = np.stack([TP[t, k] - ut[i, k] * ut[j, k] for t, (i, j) in enumerate(SYM)])
        M = np.stack([2.0 * d2[k] * (TP[6 + t, k] - 4.0 * magt[k] * st[t, k]) for t in range(6)])
        lm = contract(L, M).sum()
        mm = contract(M, M).sum()
        c = lm / mm if mm > 0.0 else 0.0
        cs2[k] = c if c > 0.0 else 0.0
    return cs2


def tau_smagorinsky(g, Ph, cs2):
    """tau_ij = -2 C D^2 |S| S_ij at the interior planes, [6, ny+1, nz2, nx2]."""
    s, mag = strain(Ph[:9])
    d2 = les_delta2(g)
    tau = np.zeros((6,) + Ph.shape[1:])
    k = np.arange(1, g.ny)
    c = (-2.0 * cs2[k] * d2[k])[:, None, None] * mag[k]
    for t in range(6):
        tau[t, k] = c * s[t, k]
    return tau


def features(g, Ph):
    """Network inputs [(ny-1) nz2 nx2, 10] (float32) and the stress scale
    D^2 |G|^2 per point, points ordered ((k-1) nz2 + z) nx2 + x."""
    ny = g.ny
    gf = Ph[:9, 1:ny]
    n2 = (gf * gf).sum(axis=0)
    gn = np.sqrt(n2)
    inv = np.where(gn > 0.0, 1.0 / np.where(gn > 0.0, gn, 1.0), 0.0)
    X = np.empty(gf.shape[1:] + (10,), dtype=np.float32)
    for f in range(9):
        X[..., f] = (gf[f] * inv).astype(np.float32)
    X[..., 9] = np.float32(0.0)
    X[..., 9] = les_ydist(g)[1:ny, None, None].astype(np.float32)
    scale = les_delta2(g)[1:ny, None, None] * n2
    return X.reshape(-1, 10), scale.reshape(-1)


def tau_learned(g, Ph, Y, scale, cs2, beta=1.0):
    """Learned stress tau = scale * Y with the fallback gate: a point whose
    energy transfer -tau:S is below -beta times Smagorinsky's (beta >= 0), or
    is not finite, takes the Smagorinsky stress.  Returns (tau, fallback
    count)."""
    ny = g.ny
    s, mag = strain(Ph[:9])
    ts = tau_smagorinsky(g, Ph, cs2)
    sh = (ny - 1,) + Ph.shape[2:]
    tm = np.stack([(scale * Y[:, t].astype(np.float64)).reshape(sh) for t in range(6)])
    si = s[:, 1:ny]
    em = -contract(tm, si)
    es = -contract(ts[:, 1:ny], si)
    with np.errstate(invalid="ignore"):
        fb = ~np.isfinite(em)
        if beta >= 0.0:
            fb = fb | ~(em >= -beta * es)
    tau = np.zeros_like(ts)
    tau[:, 1:ny] = np.where(fb[None], ts[:, 1:ny], tm)
    return tau, int(fb.sum())


# ---------------------------------------------------------------------------
# Filtered DNS: restatement of the device filter (chan_les_filter_dns) for
# its GPU test.  Fine grid gf, coarse grid gc with the same box and every
# other fine face (gf.ny = 2 gc.ny - 1).  Pinned by the mathematics too:
# tests/test_les_oracle.py checks that the filter keeps the discrete
# divergence zero, passes a constant, and that tau vanishes for a field with
# no modes above the coarse cut-off and no y structure inside coarse cells.
# ---------------------------------------------------------------------------
def filter_spec(gf, gc, C, kind):
    """One fine spectral field C [ny_f+1, nxh_f, nz_f] -> coarse [ny_c+1,
    nxh_c, nz_c]: 3-point box in x, z (transfer (1 + 2 cos(k dx))/3 on the
    fine grid), cut-off without the coarse Nyquist modes, then in y
    kind 0: height-weighted average of fine cells 2K-1, 2K, walls copied;
    kind 1: the same, walls zero; kind 2: faces 2K (plane ny_c zero)."""
    nyf, nyc = gf.ny, gc.ny
    ix = np.arange(gc.nxh)
    iz = np.arange(gc.nz)
    kzi = np.where(iz <= gc.nz // 2, iz, iz - gc.nz)
    src = np.where(kzi >= 0, kzi, gf.nz + kzi)
    keep = (ix[:, None] < gc.nxh - 1) & (np.abs(kzi)[None, :] < gc.nz // 2)
    tx = (1.0 + 2.0 * np.cos(2.0 * math.pi * ix / gf.nx)) / 3.0
    tz = (1.0 + 2.0 * np.cos(2.0 * math.pi * kzi / gf.nz)) / 3.0
    f = np.where(keep, tx[:, None] * tz[None, :], 0.0)
    Cc = C[:, :gc.nxh, :][:, :, src] * f[None]
    out = np.zeros((nyc + 1, gc.nxh, gc.nz), np.complex128)
    if kind == 2:
        out[:nyc] = Cc[0:2 * nyc:2]
        return out
    hh = np.zeros(nyf + 1)
    hh[1:nyf] = gf.y[1:] - gf.y[:-1]
    K = np.arange(1, nyc)
    a, b = hh[2 * K - 1], hh[2 * K]
    out[1:nyc] = (a[:, None, None] * Cc[2 * K - 1] + b[:, None, None] * Cc[2 * K]) / (a + b)[:, None, None]
    if kind == 0:
        out[0] = Cc[0]
        out[nyc] = Cc[nyf]
    return out


def filter_dns(gf, gc, U):
    """Filtered velocity [3, ny_c+1, nxh_c, nz_c] and the exact subgrid stress
    tau_ij = F(u_i u_j) - F(u_i) F(u_j) [6, ny_c+1, nz2_c, nx2_c] of the fine
    velocity U (products of the cell-centre velocities, v averaged from its
    faces, on the fine dealiased grid; tau zero on the wall planes)."""
    co = _co()
    nyf, nyc = gf.ny, gc.ny
    Uc = np.zeros((3, nyc + 1, gc.nxh, gc.nz), np.complex128)
    Uc[0] = filter_spec(gf, gc, U[0], 0)
    Uc[1] = filter_spec(gf, gc, U[1], 2)
    Uc[2] = filter_spec(gf, gc, U[2], 0)
    vc = np.zeros_like(U[1])
    vc[1:nyf] = 0.5 * (U[1, 0:nyf - 1] + U[1, 1:nyf])
    spec = (U[0], vc, U[2])
    ph = [co.xz_c2r(gf, s) for s in spec]
    ub = [co.xz_c2r(gc, filter_spec(gf, gc, s, 1)) for s in spec]
    tau = np.zeros((6, nyc + 1, gc.nz2, gc.nx2))
    for t, (i, j) in enumerate(zip((0, 0, 0, 1, 1, 2), (0, 1, 2, 1, 2, 2))):
        P = ph[i] * ph[j]
        P[0] = 0.0
        P[nyf] = 0.0
        Fp = co.xz_c2r(gc, filter_spec(gf, gc, co.xz_r2c(gf, P), 1))
        tau[t] = Fp - ub[i] * ub[j]
        tau[t, 0] = 0.0
        tau[t, nyc] = 0.0
    return Uc, tau
