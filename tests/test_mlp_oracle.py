"""The MLP oracle's gradients against central finite differences (FP64)."""
import numpy as np

from oracle import mlp_oracle as mo


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.standard_normal(mo.NPARAMS) * 0.1
    X = rng.standard_normal((7, mo.IN))
    T = rng.standard_normal((7, mo.OUT))
    w = rng.uniform(0.5, 1.5, 7)
    L, g = mo.loss_and_grad(p, X, T, w)
    idx = rng.choice(mo.NPARAMS, 60, replace=False)
    # include every bias block and the output layer
    idx = np.concatenate([idx, [48 * 128, 48 * 128 + 128 + 128 * 128 + 5, mo.NPARAMS - 1, mo.NPARAMS - 3]])
    h = 1e-6
    for i in idx:
        e = np.zeros_like(p)
        e[i] = h
        fd = (mo.loss_and_grad(p + e, X, T, w)[0] - mo.loss_and_grad(p - e, X, T, w)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    x = np.array([1.0, 1.00390625, 1.0078125, 1.01171875, -3.14159])
    r = mo.bf16(x)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7: rounds to even (1.0)
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125 and r[3] == 1.015625
    assert abs(r[4] + 3.140625) < 1e-12


def test_gelu_is_the_tanh_form():
    """The activation is GELU x Phi(x) in its tanh approximation: within
    1e-3 of the erf form (the approximation's known error), exactly x/2 at
    0, and odd part x/2 (g(x) - g(-x) = x)."""
    from scipy.special import erf

    from oracle import mlp_oracle as mo
    x = np.linspace(-8, 8, 4001)
    exact = 0.5 * x * (1.0 + erf(x / np.sqrt(2.0)))
    assert np.max(np.abs(mo.gelu(x) - exact)) < 1e-3
    assert np.allclose(mo.gelu(x) - mo.gelu(-x), x, rtol=0, atol=1e-12)
    # a hard-sigmoid or a wrong constant would fail the closed form
    assert abs(mo.gelu(np.array([1.0]))[0] - 0.8411919906) < 1e-9
