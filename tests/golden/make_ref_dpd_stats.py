"""Writes tests/golden/ref_dpd_stats.npz: the DPD-thermostat statistics
(tests/dpd_stats.py) of the reference's own step routines -- fluid_fluid
with dpd_forces drawing ziggurat normals, intra_molec, fluid_wall,
bending, constant_force, the Verlet halves (oracle/_ref, compiled from the
reference sources by oracle/ref.mk; OpenMP) -- on the test channel, from the
same initial state the GPU test uses.  Two runs with different ziggurat
seeds (halves a, b and c, d), so the spread between runs is known as well
as the spread within one.  Run in the container that holds the reference
checkout: python tests/golden/make_ref_dpd_stats.py"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, os.path.dirname(HERE))

import dpd_stats  # noqa: E402
from oracle.oracle import RefLib  # noqa: E402


class RefStepper:
    def __init__(self, p, t, r0, v, seed):
        self.ref = RefLib(p, omp=True, seed=seed)
        self.ref.set_state(r0, v, np.zeros_like(r0), t)
        self.t0 = 0

    def md_steps(self, n):
        self.ref.md_steps(n, self.t0)
        self.t0 += n

    def state(self):
        st = self.ref.get_state()
        return st["r0"], st["v"]


def main():
    p, t, r0, v = dpd_stats.system()
    out = {}
    for names, seed in (("ab", 4711), ("cd", 4712)):
        t0 = time.time()
        stepper = RefStepper(p, t, r0, v, seed=seed)
        a, b = dpd_stats.run(stepper, p, t, r0, v)
        for name, h in zip(names, (a, b)):
            for k, x in h.items():
                out["%s_%s" % (name, k)] = np.asarray(x)
        out["threads"] = np.array(stepper.ref.threads())
        print("seed %d: kT %.4f / %.4f, var_yz %.4f / %.4f, kurt %.4f / %.4f, rho halves L1 %.4f, %.0f s" %
              (seed, a["kt"], b["kt"], a["var_yz"], b["var_yz"], a["kurt_yz"], b["kurt_yz"],
               dpd_stats.l1(a["rho"], b["rho"]), time.time() - t0), flush=True)
    out["steps"] = np.array([dpd_stats.EQUIL, dpd_stats.SAMPLES, dpd_stats.EVERY])
    np.savez(os.path.join(HERE, "ref_dpd_stats.npz"), **out)


if __name__ == "__main__":
    main()
