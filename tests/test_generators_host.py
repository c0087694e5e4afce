"""The reference's generators restated on the host (gen_brush.f90 modes 1-6,
gen_droplet.f90 case 1, conf_default.f90 labels and velocities): geometry
each mode fixes, checked on the generated arrays, and a short oracle run
from a generated start."""
import math

import numpy as np
import pytest

from mfa_brush_amd import configs, generators as G

R_CHAIN, SQRT3 = 1.5, 1.73205080756


def _params(n_mon=8, n_chain=32, n_chain_d=300, n_mon_d=4, box=(16.0, 16.0, 20.0)):
    p = configs.channel_params(box, n_mon, n_chain, n_chain_d, thermostat=True, rc_dpd=1.0, force_switch_on=True)
    return p.replace(n_mon_d=n_mon_d)


def _steps(r, L):
    d = np.diff(r, axis=0)
    d[:, :2] -= L[:2] * np.round(d[:, :2] / L[:2])
    return d


@pytest.mark.parametrize("mode", [1, 2, 3, 4, 5, 6])
def test_gen_brush_geometry(mode):
    p = _params(n_chain=32 if mode != 6 else 16)
    L = np.array(p.boundary)
    r = G.gen_brush(p, mode, np.random.default_rng(mode), z_head=1.2)
    chains = r.reshape(p.n_chain, p.n_mon, 3)
    top = np.zeros(p.n_chain, bool)
    if mode in (1, 3, 5):
        top[:p.n_chain // 2] = True                      # first half on the top wall
    assert np.allclose(chains[top, 0, 2], L[2] - 1.2) and np.allclose(chains[~top, 0, 2], 1.2)
    dz = 0.4 if mode == 3 else 0.5
    for c in range(p.n_chain):
        d = _steps(chains[c], L)
        assert np.allclose(d[:, 2], -dz if top[c] else dz)
        assert np.abs(d[:, :2]).max() <= R_CHAIN / SQRT3 + 1e-12
    assert (r[:, :2] > 0).all() and (r[:, 0] < L[0]).all() and (r[:, 1] < L[1]).all()
    heads = chains[:, 0, :2]
    if mode in (3, 4):
        for w in (top, ~top):
            h = heads[w]
            if len(h) < 2:
                continue
            d2 = ((h[:, None, :] - h[None, :, :]) ** 2).sum(-1)
            np.fill_diagonal(d2, 9.0)
            assert d2.min() > 1.0                        # no two grafting points within sigma
    if mode in (5, 6):
        walls = 2 if mode == 5 else 1
        a = math.sqrt(walls * L[0] * L[1] / p.n_chain)
        n = int(round(L[0] / a))
        per_wall = heads.reshape(walls, n, n, 2)           # column by column, rows inner
        for w in range(walls):
            gx, gy = np.meshgrid(a / 2 + a * np.arange(n), a / 2 + a * np.arange(n), indexing="ij")
            assert np.allclose(per_wall[w, :, :, 0], gx) and np.allclose(per_wall[w, :, :, 1], gy)


def test_gen_brush_lattice_refuses_mismatched_counts():
    with pytest.raises(ValueError, match="n_chain"):
        G.gen_brush(_params(n_chain=30), 5, np.random.default_rng(0))


def test_gen_droplet_bonds_and_heights():
    p = _params()
    r = G.gen_droplet(p, np.random.default_rng(3), z_skin=1.0)
    L = np.array(p.boundary)
    ch = r.reshape(p.n_chain_d, p.n_mon_d, 3)
    for c in range(p.n_chain_d):
        d = np.diff(ch[c], axis=0)
        d -= L * np.round(d / L)
        assert np.allclose(np.linalg.norm(d, axis=1), 0.95 * R_CHAIN)
    assert (r[:, 2] > 1.0 - 1e-12).all() and (r[:, 2] <= L[2] - 1.0 + 1e-12).all()
    assert (r >= 0).all() and (r < L).all()


def test_conf_default_labels_velocities_and_a_stable_start(oracle_mod):
    """Labels (1 head, 2 brush, 3 melt), zero head velocities, Maxwell-
    Boltzmann spread, and 40 oracle steps with the force switch-on from the
    ordered channel brush plus random melt stay finite and inside the walls."""
    p = _params(n_mon=8, n_chain=32, n_chain_d=200, n_mon_d=4)
    t, r0, v = G.conf_default(p, 5, seed=2)
    heads = np.arange(p.n_chain) * p.n_mon
    assert (t[heads] == 1).all() and (np.delete(t[:p.n_chain * p.n_mon], heads) == 2).all()
    assert (t[p.n_chain * p.n_mon:] == 3).all()
    assert not v[heads].any() and abs(v[p.n_chain:].std() - 1.0) < 0.1
    p = p.replace(r_2_min_init=0.64, r_2_min_time=1.0)
    r1, v1, a1, f1, e, _ = oracle_mod.Oracle(p).md_steps(r0, v, np.zeros_like(r0), t, 0, 40, noise=1)
    assert np.isfinite(r1).all() and np.isfinite(v1).all()
    assert (r1[:, 2] > 0).all() and (r1[:, 2] < p.boundary[2]).all()
