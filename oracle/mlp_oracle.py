"""TEST INFRASTRUCTURE ONLY -- FP64 restatement of the profile surrogate
(include/mfa_mlp.h), the test reference for the BF16 MFMA kernels of
mfa_brush_amd/csrc/mfa_mlp.hip.  The surrogate has no counterpart in the
reference, so this is pinned by the mathematics instead: the analytic
gradient against central finite differences and the SiLU derivative against
its closed form (tests/test_mlp_oracle.py).

Network: x -> [Linear -> SiLU]^(L-1) -> Linear -> y; flat parameters
W_0 (n_0 x n_1, row-major), b_0, W_1, b_1, ...; loss = mean squared error."""
import numpy as np


def unpack(widths, p):
    p = np.asarray(p, np.float64)
    out, o = [], 0
    for a, b in zip(widths[:-1], widths[1:]):
        W = p[o:o + a * b].reshape(a, b); o += a * b
        bb = p[o:o + b]; o += b
        out.append((W, bb))
    assert o == p.size
    return out


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(widths, p, x):
    h = np.asarray(x, np.float64)
    layers = unpack(widths, p)
    for i, (W, b) in enumerate(layers):
        z = h @ W + b
        h = silu(z) if i < len(layers) - 1 else z
    return h


def loss_grad(widths, p, x, t):
    """MSE over all outputs and its gradient w.r.t. the flat parameters."""
    layers = unpack(widths, p)
    hs, zs = [np.asarray(x, np.float64)], []
    for i, (W, b) in enumerate(layers):
        z = hs[-1] @ W + b
        zs.append(z)
        hs.append(silu(z) if i < len(layers) - 1 else z)
    t = np.asarray(t, np.float64)
    n = t.size
    loss = float(((hs[-1] - t) ** 2).sum() / n)
    dz = 2.0 * (hs[-1] - t) / n
    grads = []
    for i in range(len(layers) - 1, -1, -1):
        W, _ = layers[i]
        grads.append((hs[i].T @ dz, dz.sum(0)))
        if i > 0:
            dz = (dz @ W.T) * dsilu(zs[i - 1])
    g = []
    for dW, db in reversed(grads):
        g.append(dW.ravel()); g.append(db)
    return loss, np.concatenate(g)
