"""The Philox DPD thermostat against the reference's own ziggurat-driven
runs, statistically: the same channel (tests/dpd_stats.py: C1 size, DPD with
the LJ cutoff as range, bending, Poiseuille body force 0.1) from the same
initial state, 10,000 steps then 800 samples every 20 steps, on the GPU
(both paths) and in the reference's step routines (fixture
tests/golden/ref_dpd_stats.npz, written by tests/golden/make_ref_dpd_stats.py:
two reference runs with different ziggurat seeds, whose difference gives
the spread between runs).

Bars: kinetic temperature and solvent y/z velocity variance within 1 % of
the reference's (a noise amplitude off by 1 % moves kT by 2 %); excess
kurtosis of those components below 0.02 (Maxwell-Boltzmann); brush and
solvent density profiles within twice the L1 distance between the two
reference runs; mean solvent flow within three standard errors of the
two-run reference mean, the per-run standard deviation estimated as the
two runs' difference over sqrt(2).  The flow and the profiles move with
the slow arrangement of the collapsed brush in the channel's middle, so
the two halves of one run agree more closely than two runs do."""
import os

import numpy as np
import pytest

import dpd_stats
from mfa_brush_amd.sim import Simulation

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


class GpuStepper:
    def __init__(self, sim):
        self.s = sim

    def md_steps(self, n):
        self.s.md_steps(n, graph=True)

    def state(self):
        st = self.s.get_state()
        return st["r0"], st["v"]


def _mean_flow(h):
    return float((h["vx"] * h["sol_count"]).sum() / h["sol_count"].sum())


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_dpd_statistics_match_reference_runs(precision):
    z = np.load(os.path.join(HERE, "golden", "ref_dpd_stats.npz"))
    half = {h: {k[2:]: z[k] for k in z.files if k.startswith(h + "_")} for h in "abcd"}
    r1, r2 = dpd_stats.merge(half["a"], half["b"]), dpd_stats.merge(half["c"], half["d"])
    ref = dpd_stats.merge(r1, r2)
    p, t, r0, v = dpd_stats.system()
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        a, b = dpd_stats.run(GpuStepper(s), p, t, r0, v)
        assert s.flags()["nonfinite"] == 0
    g = dpd_stats.merge(a, b)
    rho_runs = dpd_stats.l1(r1["rho"], r2["rho"])
    flow_sd = abs(_mean_flow(r1) - _mean_flow(r2)) / np.sqrt(2.0)
    flow_se = flow_sd * np.sqrt(1.0 + 0.5)         # one GPU run against the mean of two reference runs
    print("%s: kT %.4f (ref %.4f), var %.4f (ref %.4f), kurt %.4f (ref %.4f), rho L1 %.4f (ref runs %.4f), "
          "flow %.4f (ref %.4f, standard error %.4f)"
          % (precision, g["kt"], ref["kt"], g["var_yz"], ref["var_yz"], g["kurt_yz"], ref["kurt_yz"],
             dpd_stats.l1(g["rho"], ref["rho"]), rho_runs, _mean_flow(g), _mean_flow(ref), flow_se))
    assert abs(g["kt"] - ref["kt"]) < 0.01 * ref["kt"]
    assert abs(g["var_yz"] - ref["var_yz"]) < 0.01 * ref["var_yz"]
    assert abs(g["kurt_yz"]) < 0.02 and abs(ref["kurt_yz"]) < 0.02
    assert dpd_stats.l1(g["rho"], ref["rho"]) < 2.0 * rho_runs
    assert abs(_mean_flow(g) - _mean_flow(ref)) < 3.0 * flow_se
