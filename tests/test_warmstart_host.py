"""Host-side pieces of the surrogate warm start (no GPU): the general channel
generator's geometry, the time-to-equilibrium metric and the feature map."""
import numpy as np
from scipy.spatial import cKDTree

from mfa_brush_amd import configs
from mfa_brush_amd import warmstart as ws


def _min_dist(p, r0):
    L, lz = p.boundary[0], p.boundary[2]
    pts = np.c_[r0[:, 0] % L, r0[:, 1] % L, r0[:, 2] + lz]
    d, _ = cKDTree(pts, boxsize=[L, L, 4 * lz]).query(pts, k=2)
    return d[:, 1].min()


def test_plain_channel_geometry():
    p, t, r0, v = configs.channel_general(6, 2.0, 8, 20.0, skin=0.3)
    L = p.boundary[0]
    assert p.n_part == int(round(0.8 * L * L * 20.0))
    assert (t[:p.n_chain * 8:8] == 1).all() and (t[p.n_chain * 8:] == 3).all()
    assert _min_dist(p, r0) >= 0.85 - 1e-9
    rb = r0[:p.n_chain * 8].reshape(p.n_chain, 8, 3)
    db = np.diff(rb, axis=1)
    db[..., :2] -= L * np.round(db[..., :2] / L)
    assert np.allclose(np.linalg.norm(db, axis=2), configs.BOND_Z)
    assert 0.8 <= r0[:, 2].min() and r0[:, 2].max() <= 19.2


def test_warm_channel_follows_requested_heights():
    p, t, r0, v = configs.channel_general(6, 2.0, 8, 20.0, skin=0.3, brush_z=lambda m: 0.6 * m,
                                          solvent_weight=lambda z: 3.0 if z < 10 else 1.0)
    n = p.n_chain * 8
    bottom = r0[:n // 2].reshape(-1, 8, 3)
    assert np.allclose(bottom[:, :, 2].mean(0), 1.2 + 0.6 * np.arange(8), atol=1e-9)
    sol = r0[n:, 2]
    p2, t2, r2, v2 = configs.channel_general(6, 2.0, 8, 20.0, skin=0.3, brush_z=lambda m: 0.6 * m)
    sol2 = r2[n:, 2]
    # excluded volume near the brushes damps it, but the weight shifts solvent down
    assert (sol < 10).sum() / (sol >= 10).sum() > 1.15 * (sol2 < 10).sum() / (sol2 >= 10).sum()


def test_time_to_equilibrium_metric():
    ref = np.ones((2, 64))
    blocks = [ref * 1.3, ref * 1.1, ref * 1.01, ref * 1.03, ref * 1.01] + [ref * 1.005] * 6
    k, err = ws.time_to_equilibrium(blocks, ref, tol=0.02, consecutive=5)
    assert k == 5 and abs(err[0] - 0.3) < 1e-12
    assert ws.time_to_equilibrium(blocks[:6], ref)[0] is None


def test_features_are_normalised():
    s = ws.Spec(0.275, 20, 1.0)
    f = ws.features(s, np.array([0.0, ws.LZ / 2, ws.LZ]))
    assert np.allclose(f[:, :3], 0.0) and np.allclose(f[:, 3], [-1, 0, 1])
