"""The closure network on the GPU against the CPU restatement.

Tolerances: bfloat16 matrix-core inference vs the restatement of the same
bfloat16 rounding: 2e-2 relative to the output scale (float32 accumulation
order, erff, and bf16 re-rounding of activations that sit on rounding
boundaries); float32 forward 1e-5 and gradient 2e-4 relative (float32 vs
float64); Adam training 5e-4 relative on the parameters after 3 steps.
"""
import numpy as np
import pytest

from helpers import chanflow
from oracle import les_oracle as L

pytestmark = pytest.mark.gpu


def _setup(seed=7, n=300):
    W, b = L.init_params(seed)
    W = [w.astype(np.float32).astype(np.float64) for w in W]
    b = [v.astype(np.float32).astype(np.float64) for v in b]
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((n, 10)).astype(np.float32).astype(np.float64)
    T = (0.3 * rng.standard_normal((n, 6))).astype(np.float32).astype(np.float64)
    net = chanflow().Mlp(chanflow().Mlp.flatten(W, b))
    return net, W, b, X, T


def test_param_roundtrip_and_count():
    net, W, b, X, T = _setup()
    assert net.nparams == 35206
    np.testing.assert_array_equal(net.get_params(), chanflow().Mlp.flatten(W, b))
    net.close()


@pytest.mark.parametrize("n", [1, 47, 1000])
def test_bf16_inference_matches_restatement(n):
    # a wavefront takes 48 points at a time: one point, one short of a batch,
    # and a count with a ragged last batch
    net, W, b, X, T = _setup(n=n)
    y = net.infer(X)
    ref = L.forward_bf16(W, b, X)
    scale = np.abs(ref).max()
    assert y.shape == (n, 6)
    assert np.abs(y - ref).max() <= 2e-2 * scale
    # and close to the exact network
    assert np.abs(y - L.forward(W, b, X)).max() <= 5e-2 * scale
    net.close()


def test_fp32_forward_and_gradient():
    net, W, b, X, T = _setup()
    y = net.forward(X)
    ref = L.forward(W, b, X)
    assert np.abs(y - ref).max() <= 1e-5 * np.abs(ref).max()
    g, loss = net.gradient(X, T)
    gW, gb = L.backward(W, b, X, T)
    gref = chanflow().Mlp.flatten(gW, gb).astype(np.float64)
    assert np.linalg.norm(g - gref) <= 2e-4 * np.linalg.norm(gref)
    assert abs(loss - L.mse(ref, T)) <= 1e-5 * L.mse(ref, T)
    net.close()


def test_adam_training_matches_restatement():
    net, W, b, X, T = _setup(n=256)
    loss = net.train(X, T, epochs=3, batch=256, lr=1e-3)
    p = chanflow().Mlp.flatten(W, b).astype(np.float64)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    sizes = [(w.shape, bb.shape) for w, bb in zip(W, b)]
    for t in range(1, 4):
        Wc, bc, o = [], [], 0
        for (ws, bs) in sizes:
            nw = ws[0] * ws[1]
            Wc.append(p[o:o + nw].reshape(ws)); o += nw
            bc.append(p[o:o + bs[0]]); o += bs[0]
        gW, gb = L.backward(Wc, bc, X, T)
        g = chanflow().Mlp.flatten(gW, gb).astype(np.float64)
        p, m, v = L.adam_step(p, g, m, v, t, lr=1e-3)
    got = net.get_params().astype(np.float64)
    assert np.linalg.norm(got - p) <= 5e-4 * np.linalg.norm(p - chanflow().Mlp.flatten(W, b))
    assert loss > 0
    net.close()


def test_training_reduces_loss_on_a_smooth_map():
    """Fit the Smagorinsky stress of random gradients (a smooth map): the loss
    after training is far below the initial loss."""
    rng = np.random.default_rng(1)
    n = 4096
    G = rng.standard_normal((n, 3, 3))
    yplus = rng.uniform(0, 1, n)
    X = np.concatenate([G.reshape(n, 9), yplus[:, None]], axis=1)
    T = L.smagorinsky_tau(G, 0.5 + yplus, 0.5)
    net = chanflow().Mlp(chanflow().Mlp.flatten(*L.init_params(2)))
    _, l0 = net.gradient(X, T)
    net.train(X, T, epochs=40, batch=512, lr=2e-3)
    _, l1 = net.gradient(X, T)
    assert l1 < 0.1 * l0
    net.close()
