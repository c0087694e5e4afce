"""Writes tests/golden/try_fixture.npz from the reference's own sample run.

Run in a container that has the reference checkout; the committed .npz is
what the tests read (the reference itself never travels).  Contents are
data the reference ships under try/: the conf_old configuration (6000 beads,
box 30x20x40), its system_input / mfa_input text, and the brute-force
energy the reference printed for it (try/log:152).
"""
import os
import sys

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", ".."))

from mfa_brush_amd.params import Flags, read_conf_old, read_mfa_input, read_system_input  # noqa: E402

# flags of the 2016 build that produced try/ (try/log:5-35,129)
TRY_FLAGS = Flags(solvent=3, brush_in_good_sv=True, dpd_cut_off=2.24, bending=True, poiseuille=True)


def main():
    si_txt = open(os.path.join(REF, "try", "system_input")).read()
    mi_txt = open(os.path.join(REF, "try", "mfa_input")).read()
    si = read_system_input(si_txt, TRY_FLAGS)
    n = si.n_mon * si.n_chain + si.n_mon_d * si.n_chain_d
    a_type, r0, v, a = read_conf_old(open(os.path.join(REF, "try", "conf_old")).read(), n)
    log = open(os.path.join(REF, "try", "log")).read().splitlines()
    vff = None
    for ln in log:
        if "[check_fluid_fluid] V=" in ln:
            vff = float(ln.split("=")[1])
    assert vff is not None
    # fort.41 (i_time, V_ff, V_bond, V_wall every 5 steps) first rows
    f41 = np.loadtxt(os.path.join(REF, "try", "fort.41"), max_rows=20)
    np.savez_compressed(os.path.join(HERE, "try_fixture.npz"), a_type=a_type, r0=r0, v=v, a=a,
                        system_input=np.array(si_txt), mfa_input=np.array(mi_txt),
                        v_ff_check=np.array(vff), fort41=f41)
    print("wrote try_fixture.npz: n=%d V_ff=%r" % (n, vff))


if __name__ == "__main__":
    main()
