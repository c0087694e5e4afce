"""The reference-run statistics fixture (tests/golden/ref_dpd_stats.npz) and
the analysis code that made it (tests/dpd_stats.py), on the CPU: the
reference's thermostat holds kT = 1 in both halves of its sampling window,
its y/z velocities are Gaussian, each density profile integrates to its
species' bead count, and the analysis gives the closed forms on a synthetic
state."""
import os

import numpy as np

import dpd_stats

HERE = os.path.dirname(os.path.abspath(__file__))


def test_reference_fixture_holds_kt_and_counts():
    z = np.load(os.path.join(HERE, "golden", "ref_dpd_stats.npz"))
    p, t, r0, v = dpd_stats.system()
    dz = p.boundary[2] / dpd_stats.N_BINS
    area = p.boundary[0] * p.boundary[1]
    for h in ("a", "b", "c", "d"):
        assert abs(float(z[h + "_kt"]) - 1.0) < 0.01
        assert abs(float(z[h + "_var_yz"]) - 1.0) < 0.01
        assert abs(float(z[h + "_kurt_yz"])) < 0.02
        n_sol = int((t == 3).sum())
        assert np.isclose(z[h + "_rho"][1].sum() * dz * area, n_sol)
        assert np.isclose(z[h + "_rho"][0].sum() * dz * area, t.size - n_sol)
        assert np.isclose(z[h + "_sol_count"].sum(), n_sol)


def test_stats_closed_forms():
    p, t, r0, _ = dpd_stats.system()
    rng = np.random.default_rng(2)
    st = dpd_stats.Stats(p, t)
    n = t.size
    v = rng.normal(0.0, np.sqrt(2.0), (n, 3))        # kT 2 (unit masses)
    v[t == 1] = 0.0
    v[:, 0] += 0.5                                    # uniform flow along x
    st.add(r0, v)
    r = st.result()
    free = t != 1
    assert np.isclose(r["kt"], (v[free] ** 2).sum() / (3 * free.sum()))
    assert abs(r["var_yz"] - 2.0) < 0.1 and abs(r["kurt_yz"]) < 0.2
    sol = t == 3
    assert np.allclose(r["vx"][r["sol_count"] > 0],
                       [v[sol & (np.clip((r0[:, 2] / p.boundary[2] * dpd_stats.N_BINS).astype(int), 0,
                                         dpd_stats.N_BINS - 1) == b), 0].mean()
                        for b in np.nonzero(r["sol_count"] > 0)[0]])
