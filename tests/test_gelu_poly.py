"""The branch-free GELU of the bf16 inference kernel (gelu_half in
csrc/k_mlp.hip, which takes y = x/2 from the hidden layers' halved weights and
biases), evaluated in float32 exactly as the kernel writes it from the
constants in that source file, against GELU(x) = x (1 + erf(x / sqrt 2)) / 2.
Each fused multiply-add is emulated as an exact float64 product and sum
rounded once to float32 (a float32 product is exact in float64).  This pins
the committed constants and the claim in the kernel's comment: absolute error
<= 6.3e-5 for |x| <= 8 and relative error < 1e-5 beyond."""
import math
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "3d-navier-stokes-thesis-paper_amd", "csrc", "k_mlp.hip")


def _kernel_constants():
    body = open(SRC).read().split("float gelu_half(float y) {", 1)[1].split("\n}", 1)[0]
    first = re.search(r"float e = ([-+0-9.e]+)f;", body).group(1)
    rest = re.findall(r"e = fmaf\(e, w, ([-+0-9.e]+)f\);", body)
    return [np.float32(float(v)) for v in [first] + rest]


def _fma32(a, b, c):
    f64 = lambda v: np.asarray(v, np.float32).astype(np.float64)
    return (f64(a) * f64(b) + f64(c)).astype(np.float32)


def _eval(c, x):
    x = np.asarray(x, np.float32)
    y = (np.float32(0.5) * x).astype(np.float32)   # exact
    w = np.clip((np.abs(y) * np.float32(1.0 / 2.25)).astype(np.float32), np.float32(0.0), np.float32(1.0))
    e = np.full_like(w, c[0])
    for k in c[1:]:
        e = _fma32(e, w, k)
    return _fma32(np.abs(y), e, y)


def _gelu_kernel(x):
    return _eval(_kernel_constants(), x)


def _gelu(x):
    return 0.5 * x * (1.0 + np.array([math.erf(v / math.sqrt(2.0)) for v in x]))


def test_constants_are_a_degree_9_polynomial():
    assert len(_kernel_constants()) == 10


def test_absolute_error_up_to_8():
    x = np.linspace(-8.0, 8.0, 160001)
    err = np.abs(_gelu_kernel(x).astype(np.float64) - _gelu(x))
    assert err.max() <= 6.3e-5, err.max()


def test_relative_error_beyond_8_and_limits():
    x = np.concatenate([np.linspace(8.0, 200.0, 2001), -np.linspace(8.0, 200.0, 2001)])
    g, ref = _gelu_kernel(x).astype(np.float64), _gelu(x)
    big = x > 0
    assert np.all(np.abs(g[big] - ref[big]) <= 1e-5 * ref[big])
    assert np.all(np.abs(g[~big]) <= 1e-5 * np.abs(x[~big]))   # GELU -> 0 from below
    assert _gelu_kernel(np.array([0.0]))[0] == 0.0


def test_a_wrong_constant_fails():
    """A slip in one constant (a sign) breaks the accuracy bound."""
    c = _kernel_constants()
    x = np.linspace(-8.0, 8.0, 20001).astype(np.float32)
    bad = list(c)
    bad[3] = -bad[3]
    g = _eval(bad, x)
    assert np.abs(g - _gelu(x.astype(np.float64))).max() > 1e-3


def test_halved_form_is_the_polynomial_on_x():
    """gelu_half(x/2) with the doubled coefficients equals, value for value,
    x/2 + |x| e(min(|x|/4.5, 1)) with the halved ones: halving the argument
    and doubling the coefficients only move every rounding by a power of two."""
    c = _kernel_constants()
    half = [np.float32(v) * np.float32(0.5) for v in c]
    x = np.concatenate([np.linspace(-8.0, 8.0, 40001), np.geomspace(1e-6, 300.0, 2001),
                        -np.geomspace(1e-6, 300.0, 2001)]).astype(np.float32)
    w = np.clip((np.abs(x) * np.float32(1.0 / 4.5)).astype(np.float32), np.float32(0.0), np.float32(1.0))
    e = np.full_like(w, half[0])
    for k in half[1:]:
        e = _fma32(e, w, k)
    direct = _fma32(np.abs(x), e, (np.float32(0.5) * x).astype(np.float32))
    assert np.array_equal(_gelu_kernel(x), direct)
