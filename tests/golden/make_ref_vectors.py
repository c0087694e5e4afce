"""Writes tests/golden/ref_vectors.npz with the reference's own routines.

Needs oracle/_ref (built by `make -f oracle/ref.mk` from the reference
sources).  The configuration is tests/test_oracle_vs_reference.ref_system
(n_side=6: 922 beads); the vectors are what the reference's fluid_fluid
(LJ + dissipative DPD, sig=0), intra_molec, bending, fluid_wall and 10 steps
of the md_main loop produce on it.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))
sys.path.insert(0, os.path.join(HERE, ".."))

from oracle.oracle import RefLib  # noqa: E402
from test_oracle_vs_reference import ref_system  # noqa: E402


def main():
    n_side = 6
    p, t, r0, v = ref_system(n_side=n_side)
    ref = RefLib(p)
    ref.set_state(r0, v, np.zeros_like(r0), t)
    out = dict(n_side=np.array(n_side), r0=r0, a_type=t)
    out["f_fluid_fluid"], e = ref.call(ref.FLUID_FLUID)
    out["e_fluid_fluid"] = np.array(e)
    out["f_intra_molec"], e = ref.call(ref.INTRA_MOLEC)
    out["e_intra_molec"] = np.array(e)
    out["f_bending"], _ = ref.call(ref.BENDING)
    out["f_fluid_wall"], _ = ref.call(ref.FLUID_WALL)
    ref.md_steps(10, 3)
    st = ref.get_state()
    out["r0_10"], out["v_10"] = st["r0"], st["v"]
    np.savez_compressed(os.path.join(HERE, "ref_vectors.npz"), **out)
    print("wrote ref_vectors.npz (%d beads)" % p.n_part)


if __name__ == "__main__":
    main()
