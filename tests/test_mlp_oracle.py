"""Pins the MLP oracle (oracle/mlp_oracle.py): analytic gradient against
central finite differences, Adam against its closed form for one step, and
a hand-computed forward pass. CPU only."""
import numpy as np

import mlp_oracle as mo


def setup(seed=0, nin=5, h=8, nout=2, B=7):
    rng = np.random.default_rng(seed)
    P = h * nin + h + h * h + h + nout * h + nout
    p = rng.normal(0, 0.5, P)
    X = rng.normal(0, 1, (B, nin))
    T = rng.normal(0, 1, (B, nout))
    return p, X, T, (nin, h, nout)


def test_gradient_matches_finite_differences():
    p, X, T, sh = setup()
    g = mo.grad(p, X, T, *sh)
    idx = np.arange(p.size)
    fd = mo.fd_grad(p, X, T, *sh, idx)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-9)


def test_forward_by_hand():
    nin, h, nout = 1, 1, 1
    # W1=2, b1=0.5, W2=-1, b2=0, W3=3, b3=1
    p = np.array([2.0, 0.5, -1.0, 0.0, 3.0, 1.0])
    X = np.array([[0.25]])
    Y, H1, H2 = mo.forward(p, X, nin, h, nout)
    h1 = np.tanh(2 * 0.25 + 0.5)
    h2 = np.tanh(-h1)
    assert Y[0, 0] == 3 * h2 + 1


def test_adam_first_step_is_sign_times_lr():
    """With m = v = 0 and t = 1 the bias-corrected step is lr * g / (|g| + eps)."""
    g = np.array([0.3, -2.0, 1e-3])
    p = np.zeros(3)
    p1, m, v = mo.adam(p, g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(p1, -0.01 * g / (np.abs(g) + 1e-8), rtol=1e-12)
