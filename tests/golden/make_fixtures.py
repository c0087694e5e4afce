"""Writes tests/golden/try_fixture.npz from the reference's own sample run,
and tests/golden/ref_inputs.npz from its shipped input sets.

Run in a container that has the reference checkout; the committed .npz files
are what the tests read (the reference itself never travels).  Contents are
data the reference ships:
* try/: the conf_old configuration (6000 beads, box 30x20x40), its
  system_input / mfa_input text, the brute-force energy the reference
  printed for it (try/log:152) and the first rows of fort.41;
* inputs/Filled_Channel and inputs/Fully_Flexible_Brush_Channel (the
  channel set-ups, SYSTEM 0) and inputs/Clog (a droplet in the channel,
  SYSTEM 1): conf_old positions, velocities and types, system_input /
  mfa_input text, and the hot-path flags of their control_simulation.h.
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


# hot-path flags of the input sets' control_simulation.h (ORIENTATION,
# which Filled_Channel and Clog also define, is outside the path).  Clog is
# SYSTEM 1, which the force and integration routines treat as SYSTEM 0 (its
# droplet reshaping, verlet_positions.f90:270-274, runs only while i_time <=
# n_relax, and its start time is past that); its THERMOSTAT 1 (Langevin)
# adds no thermostat force in the reference's current fluid_fluid.f90, whose
# lgv_forces call is commented out (:455-458), so it runs with the
# thermostat off.
INPUT_FLAGS = {
    "Filled_Channel": dict(solvent=3, brush_in_good_sv=True, bending=True, thermostat=True, dpd_cut_off=2.24,
                           force_switch_on=True, poiseuille=False),
    "Fully_Flexible_Brush_Channel": dict(solvent=2, brush_in_good_sv=False, bending=False, thermostat=True,
                                         dpd_cut_off=2.24, force_switch_on=True, poiseuille=False),
    "Clog": dict(solvent=3, brush_in_good_sv=True, bending=True, thermostat=False, dpd_cut_off=2.24,
                 force_switch_on=True, poiseuille=True),
}


def inputs():
    import json
    out = {}
    for name, fl in INPUT_FLAGS.items():
        d = os.path.join(REF, "inputs", name)
        si_txt = open(os.path.join(d, "system_input")).read()
        mi_txt = open(os.path.join(d, "mfa_input")).read()
        si = read_system_input(si_txt, Flags(**fl))
        n = si.n_mon * si.n_chain + si.n_mon_d * si.n_chain_d
        a_type, r0, v, _ = read_conf_old(open(os.path.join(d, "conf_old")).read(), n)
        out[name + "/system_input"] = np.array(si_txt)
        out[name + "/mfa_input"] = np.array(mi_txt)
        out[name + "/flags"] = np.array(json.dumps(fl))
        out[name + "/a_type"] = a_type.astype(np.int8)
        out[name + "/r0"] = r0
        out[name + "/v"] = v
        print("input set %s: n=%d" % (name, n))
    np.savez_compressed(os.path.join(HERE, "ref_inputs.npz"), **out)


def main():
    inputs()
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
