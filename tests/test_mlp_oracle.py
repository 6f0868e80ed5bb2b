"""Pins the CPU network oracle (oracle/mlp.py) that checks the library's HIP
network kernels: analytic gradients against central finite differences,
the one-layer closed form, Adam's first step, and a planted mistake that
the finite-difference check must catch."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp as M  # noqa: E402


def data(widths, n, seed):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(n, widths[0]))
    T = rng.normal(size=(n, widths[-1]))
    Mk = (rng.random((n, widths[-1])) < 0.7).astype(np.float64)
    p = rng.normal(size=M.nparams(widths)) * 0.5
    return X, T, Mk, p


def fd_grad(widths, p, X, T, Mk, h=1e-6):
    g = np.zeros_like(p)
    for k in range(len(p)):
        e = np.zeros_like(p)
        e[k] = h
        g[k] = (M.loss(widths, p + e, X, T, Mk) - M.loss(widths, p - e, X, T, Mk)) / (2 * h)
    return g


@pytest.mark.parametrize("widths", [[5, 3], [5, 7, 3], [4, 6, 5, 2], [3, 4, 4, 4, 3]])
def test_gradient_matches_finite_differences(widths):
    X, T, Mk, p = data(widths, 23, sum(widths))
    _, g = M.loss_grad(widths, p, X, T, Mk)
    gfd = fd_grad(widths, p, X, T, Mk)
    assert np.abs(g - gfd).max() <= 1e-6 * max(1.0, np.abs(gfd).max())


def test_one_layer_closed_form():
    widths = [4, 3]
    X, T, Mk, p = data(widths, 17, 1)
    W, b = M.split(widths, p)[0]
    Y = X @ W.T + b
    R = (Y - T) * Mk * 2.0 / Mk.sum()
    want = np.concatenate([(R.T @ X).ravel(), R.sum(axis=0)])
    L, g = M.loss_grad(widths, p, X, T, Mk)
    assert np.allclose(g, want, rtol=1e-12, atol=1e-14)
    assert np.isclose(L, ((Y - T) ** 2 * Mk).sum() / Mk.sum())


def test_empty_mask_divides_by_one():
    widths = [3, 2]
    X, T, _, p = data(widths, 5, 2)
    Z = np.zeros((5, 2))
    L, g = M.loss_grad(widths, p, X, T, Z)
    assert L == 0.0 and not g.any()


def test_adam_first_step_is_lr_sign():
    rng = np.random.default_rng(3)
    g = rng.normal(size=50)
    p = rng.normal(size=50)
    opt = M.Adam(50)
    p1 = opt.step(p, g, 1e-3)
    assert np.allclose(p1 - p, -1e-3 * np.sign(g), rtol=1e-4)


def test_planted_mistake_is_caught(monkeypatch):
    """A backward pass that forgets the ReLU mask must fail the check."""
    widths = [5, 7, 3]
    X, T, Mk, p = data(widths, 23, 9)
    real = M.loss_grad

    def wrong(widths, p, X, T, Mk):
        p = np.asarray(p, np.float64)
        acts = M.forward(widths, p, X)
        layers = M.split(widths, p)
        den = Mk.sum()
        dY = 2.0 * Mk * (acts[-1] - T) / den
        grads = []
        for l in range(len(layers) - 1, -1, -1):
            grads.append((dY.T @ acts[l], dY.sum(axis=0)))
            if l > 0:
                dY = dY @ layers[l][0]  # mask dropped
        out = []
        for dW, db in reversed(grads):
            out += [dW.ravel(), db]
        return 0.0, np.concatenate(out)

    monkeypatch.setattr(M, "loss_grad", wrong)
    _, g = M.loss_grad(widths, p, X, T, Mk)
    gfd = fd_grad(widths, p, X, T, Mk)
    assert np.abs(g - gfd).max() > 1e-3
    monkeypatch.setattr(M, "loss_grad", real)
