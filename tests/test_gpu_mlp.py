"""MLP kernels (BF16 MFMA) against the FP64 oracle: forward, gradients, and
that AdamW training on the GPU reduces the loss."""
import numpy as np
import pytest

from imodel_amd.mlp import Mlp
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def test_forward_matches_oracle():
    say("mlp forward")
    m = Mlp(seed=3)
    p = m.params().astype(np.float64)
    X = np.random.default_rng(1).standard_normal((1000, 48)).astype(np.float32)
    Y = m.forward(X)
    Yo, _ = mo.forward(p, X.astype(np.float64), emulate_bf16=True)
    # BF16 operands, FP32 accumulation: relative L2 error at BF16 level
    err = np.linalg.norm(Y - Yo) / np.linalg.norm(Yo)
    assert err < 1e-2, err


@pytest.mark.parametrize("n", [1000, 4096 + 13])
def test_gradient_matches_oracle(n):
    say("mlp gradient", n)
    m = Mlp(seed=5)
    p = m.params().astype(np.float64)
    rng = np.random.default_rng(n)
    X = rng.standard_normal((n, 48)).astype(np.float32)
    T = rng.standard_normal((n, 2)).astype(np.float32)
    w = rng.uniform(0.5, 1.5, n).astype(np.float32) / n
    loss, g = m.train_step(X, T, w, apply=False)
    Lo, go = mo.loss_and_grad(p, X.astype(np.float64), T.astype(np.float64), w.astype(np.float64),
                              emulate_bf16=True)
    assert abs(loss / Lo - 1) < 1e-2
    # every parameter block within BF16-level relative error
    o = 0
    for name, s in mo.SHAPES:
        k = int(np.prod(s))
        gb, gob = g[o:o + k], go[o:o + k]
        assert np.linalg.norm(gb - gob) <= 3e-2 * np.linalg.norm(gob) + 1e-12, name
        o += k
    # the parameters did not move (apply=False)
    assert np.array_equal(m.params(), p.astype(np.float32))


def test_training_reduces_loss():
    say("mlp training")
    m = Mlp(seed=7)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((8192, 48)).astype(np.float32)
    T = np.stack([np.sin(X[:, 0]) + 0.5 * X[:, 1] * X[:, 2], 0.3 * X[:, 3] ** 2 - X[:, 4]], 1).astype(np.float32)
    w = np.full(len(X), 1.0 / len(X), dtype=np.float32)
    losses = [m.train_step(X, T, w, lr=3e-3)[0] for _ in range(200)]
    assert losses[-1] < 0.1 * losses[0], (losses[0], losses[-1])


def test_init_restarts_optimisation():
    """init(seed) gives Mlp(seed)'s parameters and a cleared AdamW state:
    training from it repeats a fresh handle's training bit for bit."""
    say("mlp init")
    rng = np.random.default_rng(3)
    X = rng.standard_normal((2048, 48)).astype(np.float32)
    T = rng.standard_normal((2048, 2)).astype(np.float32)
    a = Mlp(seed=4)
    for _ in range(10):
        a.train_step(X, T, lr=1e-3)
    a.init(9)
    b = Mlp(seed=9)
    assert np.array_equal(a.params(), b.params())
    la = [a.train_step(X, T, lr=1e-3)[0] for _ in range(5)]
    lb = [b.train_step(X, T, lr=1e-3)[0] for _ in range(5)]
    assert la == lb
    assert np.array_equal(a.params(), b.params())


def test_forward_large_grid_rows_unchanged():
    """Above 32768 tiles the inference grid doubles; every row's result is
    independent of the grid, so it equals the same rows evaluated alone."""
    n = 32768 * 32 + 77
    say("mlp forward large grid", n)
    m = Mlp(seed=6)
    rng = np.random.default_rng(5)
    X = rng.standard_normal((n, 48)).astype(np.float32)
    Y = m.forward(X)
    for a in (0, n // 2 + 3, n - 1000):
        np.testing.assert_array_equal(Y[a:a + 1000], m.forward(X[a:a + 1000]))


def test_gradient_input_mask_matches_oracle():
    """gmask=2: T[:, 1] is dL/dy1 itself. The oracle gets the same gradient
    from the pseudo-target t1 = y1 - g / (2 w) of its squared loss."""
    say("mlp gradient input mask")
    n = 2048 + 5
    m = Mlp(seed=11)
    p = m.params().astype(np.float64)
    rng = np.random.default_rng(7)
    X = rng.standard_normal((n, 48)).astype(np.float32)
    T = rng.standard_normal((n, 2)).astype(np.float32)
    T[:, 1] = (rng.standard_normal(n) / n).astype(np.float32)
    w = np.full(n, 0.5, dtype=np.float32)
    loss, g = m.train_step(X, T, w, apply=False, gmask=2)
    Xd = X.astype(np.float64)
    Yo, _ = mo.forward(p, Xd, emulate_bf16=True)
    To = T.astype(np.float64)
    To[:, 1] = Yo[:, 1] - T[:, 1].astype(np.float64) / (2 * 0.5)
    Lo, go = mo.loss_and_grad(p, Xd, To, w.astype(np.float64), emulate_bf16=True)
    Lo0 = float(np.sum(0.5 * (Yo[:, 0] - To[:, 0]) ** 2))
    assert abs(loss / Lo0 - 1) < 1e-2, (loss, Lo0)
    assert Lo > Lo0  # the pseudo part is in the oracle's loss only
    o = 0
    for name, s in mo.SHAPES:
        k = int(np.prod(s))
        gb, gob = g[o:o + k], go[o:o + k]
        assert np.linalg.norm(gb - gob) <= 3e-2 * np.linalg.norm(gob) + 1e-12, name
        o += k
