"""Chain observables of the oracle (observation.f90:182-290) pinned by closed
forms: a straight chain's radius of gyration, invariance under periodic
refolding and translation, the anisotropy of chains along z and along x,
and the momentum of a uniform velocity field."""
from types import SimpleNamespace

import numpy as np
import pytest

from oracle.oracle import chain_observables


def _system(n_chain, n_mon, n_chain_d, n_mon_d, L=(10.0, 12.0, 20.0)):
    return SimpleNamespace(n_chain=n_chain, n_mon=n_mon, n_chain_d=n_chain_d, n_mon_d=n_mon_d,
                           boundary=L, mass_type=[1.0, 1.0, 2.0, 1.0])


def _straight(n_ch, n, b, axis, start, L):
    r = np.zeros((n_ch * n, 3))
    for c in range(n_ch):
        for j in range(n):
            r[c * n + j] = start + np.array([0.3 * c, 0.7 * c, 0.5 * c])
            r[c * n + j, axis] += j * b
    return r


def test_straight_chain_closed_form():
    n, b = 16, 0.97
    p = _system(3, n, 0, 0, L=(40.0, 40.0, 40.0))
    r = _straight(3, n, b, 0, np.array([2.0, 3.0, 4.0]), p.boundary)
    t = np.full(len(r), 2, np.int32)
    gb, gm, an, _ = chain_observables(p, r, np.zeros_like(r), t)
    assert gb[0] == pytest.approx(b * b * (n * n - 1) / 12.0, rel=1e-13)
    assert abs(gb[1]) < 1e-28 and abs(gb[2]) < 1e-28 and np.all(gm == 0.0)   # sum * (1/n) rounding only
    assert an[0] == pytest.approx(-0.5, rel=1e-13)


def test_refolding_and_translation_invariant():
    rng = np.random.default_rng(3)
    n_ch, n = 5, 12
    p = _system(n_ch, n, 4, 6)
    steps = rng.normal(0, 0.5, (n_ch * n + 24, 3))
    r = np.zeros_like(steps)
    for c in range(n_ch):                        # random walks, unfolded
        r[c * n:(c + 1) * n] = np.cumsum(steps[c * n:(c + 1) * n], axis=0) + rng.uniform(0, 10, 3)
    base = n_ch * n
    for c in range(4):
        r[base + 6 * c:base + 6 * (c + 1)] = np.cumsum(steps[base + 6 * c:base + 6 * (c + 1)], axis=0) + 5.0
    r[:, 2] = np.abs(r[:, 2]) % 19.0 + 0.5       # z stays inside the walls, no jumps > Lz/2 per bond
    t = np.full(len(r), 3, np.int32)
    L = np.array(p.boundary)
    ref = chain_observables(p, r, np.zeros_like(r), t)
    folded = r.copy()
    folded[:, :2] = np.mod(folded[:, :2], L[:2])  # into the box: the chains now cross the x, y boundaries
    shifted = folded.copy()
    shifted[:, :2] = np.mod(shifted[:, :2] + np.array([3.3, -7.1]), L[:2])
    for rr in (folded, shifted):
        out = chain_observables(p, rr, np.zeros_like(rr), t)
        for a, b in zip(out[:3], ref[:3]):
            assert np.allclose(a, b, rtol=1e-11, atol=1e-12)


def test_anisotropy_along_z():
    n = 10
    p = _system(4, n, 0, 0, L=(30.0, 30.0, 30.0))
    r = _straight(4, n, 1.0, 2, np.array([1.0, 1.0, 1.0]), p.boundary)
    _, _, an, _ = chain_observables(p, r, np.zeros_like(r), np.full(len(r), 2, np.int32))
    assert an[0] == pytest.approx(1.0, rel=1e-13)


def test_momentum_uniform_velocity():
    p = _system(2, 3, 5, 1)
    n = 2 * 3 + 5
    t = np.array([1, 2, 2, 1, 2, 2, 3, 3, 3, 3, 3], np.int32)
    v = np.tile([0.5, -1.0, 2.0], (n, 1))
    _, _, _, mom = chain_observables(p, np.zeros((n, 3)), v, t)
    mass = 6 * 1.0 + 5 * 2.0
    assert np.allclose(mom, mass * np.array([0.5, -1.0, 2.0]), rtol=0, atol=1e-14)
