"""The FP64 surrogate oracle (oracle/mlp_oracle.py) pinned by the
mathematics: analytic gradient vs central finite differences, SiLU' vs its
closed form, and a one-layer case against the normal-equation gradient."""
import numpy as np

from oracle import mlp_oracle as mo

W = [4, 16, 16, 3]


def _params(widths, seed=0):
    rng = np.random.default_rng(seed)
    n = sum(a * b + b for a, b in zip(widths[:-1], widths[1:]))
    return rng.normal(0, 0.5, n)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(W)
    x = rng.normal(size=(9, 4)); t = rng.normal(size=(9, 3))
    _, g = mo.loss_grad(W, p, x, t)
    h = 1e-6
    for k in rng.choice(p.size, 60, replace=False):
        e = np.zeros_like(p); e[k] = h
        fd = (mo.loss_grad(W, p + e, x, t)[0] - mo.loss_grad(W, p - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[k]) <= 1e-7 + 1e-5 * abs(fd), (k, fd, g[k])


def test_silu_derivative_closed_form():
    z = np.linspace(-8, 8, 101)
    h = 1e-6
    assert np.allclose(mo.dsilu(z), (mo.silu(z + h) - mo.silu(z - h)) / (2 * h), atol=1e-8)
    assert mo.silu(0.0) == 0.0 and mo.dsilu(0.0) == 0.5


def test_linear_layer_gradient_is_normal_equation():
    rng = np.random.default_rng(2)
    w = [3, 2]
    p = rng.normal(size=3 * 2 + 2)
    x = rng.normal(size=(5, 3)); t = rng.normal(size=(5, 2))
    loss, g = mo.loss_grad(w, p, x, t)
    Wm, b = mo.unpack(w, p)[0]
    r = x @ Wm + b - t
    assert np.isclose(loss, (r ** 2).mean())
    assert np.allclose(g[:6].reshape(3, 2), 2 * x.T @ r / r.size)
    assert np.allclose(g[6:], 2 * r.sum(0) / r.size)
