"""The surrogate's BF16 MFMA kernels (mfa_brush_amd/csrc/mfa_mlp.hip) on the
device, checked against the FP64 oracle (oracle/mlp_oracle.py).

Tolerances: operands are rounded to BF16 (8-bit significand, relative
rounding 2^-9) before the FP32-accumulated products, so a forward pass of a
4-layer, 128-wide SiLU network agrees with FP64 to ~1e-2 relative; gradients
are compared by relative norm and direction."""
import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

WIDTHS = [4, 128, 128, 128, 128, 2]


def _mlp(widths=WIDTHS, seed=3):
    from mfa_brush_amd.surrogate import Mlp
    return Mlp(widths, seed=seed)


@pytest.mark.parametrize("shape", [(32, 32, 8), (70, 45, 19), (128, 128, 128), (5, 130, 300), (64, 64, 16),
                                   (200, 150, 70), (129, 65, 33)])
def test_mfma_tile_layout_exact_on_integers(shape):
    """Small integers are exact in BF16 and their products sum exactly in
    FP32, so the MFMA tile must reproduce A @ B bit for bit; a wrong lane
    layout (rows/columns/k swapped) fails."""
    from mfa_brush_amd.surrogate import gemm_check
    M, N, K = shape
    rng = np.random.default_rng(M * 7 + N)
    a = rng.integers(-8, 9, size=(M, K)).astype(np.float32)
    b = rng.integers(-8, 9, size=(K, N)).astype(np.float32)
    c = gemm_check(a, b)
    assert np.array_equal(c, (a.astype(np.int64) @ b.astype(np.int64)).astype(np.float32))


def test_forward_matches_fp64_oracle():
    rng = np.random.default_rng(5)
    x = rng.normal(size=(300, 4)).astype(np.float32)
    with _mlp() as m:
        p = m.params()
        y = m.forward(x)
    yo = mo.forward(WIDTHS, p, x)
    err = np.abs(y - yo).max() / np.abs(yo).max()
    assert err < 2e-2, err


def test_gradient_matches_fp64_oracle():
    rng = np.random.default_rng(6)
    x = rng.normal(size=(256, 4)).astype(np.float32)
    t = rng.normal(size=(256, 2)).astype(np.float32)
    with _mlp() as m:
        p = m.params()
        loss, g = m.gradient(x, t)
    lo, go = mo.loss_grad(WIDTHS, p, x, t)
    assert loss == pytest.approx(lo, rel=2e-2)
    rel = np.linalg.norm(g - go) / np.linalg.norm(go)
    cos = g @ go / (np.linalg.norm(g) * np.linalg.norm(go))
    assert rel < 5e-2 and cos > 0.998, (rel, cos)


def test_training_fits_a_smooth_map():
    """Adam on the device fits y = (sin(x0 + x1 z), exp(-z^2) x2) on 4 inputs:
    the loss falls by two orders of magnitude and the fit generalises."""
    rng = np.random.default_rng(8)

    def target(x):
        return np.stack([np.sin(x[:, 0] + x[:, 1] * x[:, 3]), np.exp(-x[:, 3] ** 2) * x[:, 2]], 1)

    x = rng.uniform(-1, 1, size=(8192, 4)).astype(np.float32)
    xt = rng.uniform(-1, 1, size=(1024, 4)).astype(np.float32)
    t = target(x).astype(np.float32)
    with _mlp(seed=11) as m:
        y0 = m.forward(xt)
        hist = m.train(x, t, steps=1500, batch=512, lr=2e-3, seed=3)
        y1 = m.forward(xt)
    e0 = ((y0 - target(xt)) ** 2).mean()
    e1 = ((y1 - target(xt)) ** 2).mean()
    assert np.all(np.isfinite(hist)) and hist[-50:].mean() < hist[:50].mean() / 30
    assert e1 < e0 / 100 and e1 < 2e-3, (e0, e1)


def test_graph_training_equals_eager():
    """Training by replaying one captured step gives the bits of launching
    every step's kernels: parameters and loss history after each of two
    train calls, with forward passes (eager work) between them."""
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, size=(3000, 4)).astype(np.float32)
    t = np.stack([np.sin(x[:, 0] * x[:, 3]), x[:, 1] * x[:, 2]], 1).astype(np.float32)
    runs = []
    for graph in (True, False):
        out = []
        with _mlp(seed=4) as m:
            m.set_graph(graph)
            for call in range(2):
                h = m.train(x, t, steps=120, batch=256, lr=2e-3, seed=5 + call)
                y = m.forward(x[:64])
                out.append((h, m.params(), y))
        runs.append(out)
    for (hg, pg, yg), (he, pe, ye) in zip(*runs):
        assert np.array_equal(hg, he) and np.array_equal(pg, pe) and np.array_equal(yg, ye)
    assert runs[0][1][0][-10:].mean() < runs[0][0][0][:10].mean()


def test_lr_schedule_matches_stepwise_calls():
    """A train call with the cosine learning-rate schedule equals one-step
    calls at the rates the schedule gives, lr (f + (1 - f)(1 + cos(pi s /
    (N - 1))) / 2) at step s, with Adam's moments and bias-correction count
    carried across the calls (one training row, so every minibatch is the
    same row)."""
    from mfa_brush_amd.surrogate import Mlp
    widths = [4, 32, 32, 2]
    rng = np.random.default_rng(3)
    x = rng.normal(size=(1, 4)).astype(np.float32)
    t = rng.normal(size=(1, 2)).astype(np.float32)
    n, lr, f = 12, 3e-3, 0.1
    with Mlp(widths, seed=4) as a, Mlp(widths, seed=4) as b, Mlp(widths, seed=4) as c:
        p0 = a.params()
        a.train(x, t, steps=n, batch=1, lr=lr, lr_final=f)
        for s in range(n):
            b.train(x, t, steps=1, batch=1, lr=float(lr * (f + (1 - f) * 0.5 * (1 + np.cos(np.pi * s / (n - 1))))))
        c.train(x, t, steps=n, batch=1, lr=lr)            # constant rate: a different end point
        pa, pb, pc = a.params(), b.params(), c.params()
    assert np.abs(pa - p0).max() > 5 * lr * f
    assert np.abs(pa - pb).max() < 1e-6, np.abs(pa - pb).max()
    assert np.abs(pa - pc).max() > 1e-4
