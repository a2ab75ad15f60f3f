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
