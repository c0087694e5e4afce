"""NVE energy drift of the channel with the thermostat off.

Total energy E = kinetic + LJ (fluid_fluid) + FENE (intra_molec) + walls
(fluid_wall) + bending; heads are frozen (verlet_velocities.f90:113-121) and
act as a static external field, so E is conserved by the exact dynamics.
Reports (E(t) - E0) / |E0| every `every` steps and the slope of a linear fit
per 1000 steps, for the mixed and FP64 paths.

usage: python -m mfa_brush_amd.nve report.json [steps] [every] [config]
       config: c2 (998,784 beads, default) or c1 (4096 beads)
"""
from __future__ import annotations

import json
import sys
import time

import numpy as np

from . import configs
from .params import Flags
from .sim import Simulation


def total_energy(e):
    return e["t_fluid"] + e["v_fluid_fluid"] + e["v_intra_molec"] + e["v_fluid_wall"] + e["v_bend"] + e["v_bend_melt"]


def drift(p, t, r0, v, precision, steps, every):
    out = []
    t0 = time.perf_counter()
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        for k in range(steps // every + 1):     # E0 after the first full step
            if k:
                s.md_steps(every, graph=True)
            else:
                s.md_steps(1)
            out.append(total_energy(s.energies()))
        fl = s.flags()
    e = np.array(out)
    rel = (e - e[0]) / abs(e[0])
    x = np.arange(len(e)) * every
    slope = float(np.polyfit(x, rel, 1)[0] * 1000) if len(e) > 2 else 0.0
    return dict(precision=precision, steps=steps, every=every, e0=float(e[0]), rel=rel.tolist(),
                max_abs_rel=float(np.abs(rel).max()), slope_per_1000_steps=slope,
                seconds=time.perf_counter() - t0, flags=fl)


def main(argv=None):
    argv = sys.argv[1:] if argv is None else argv
    path = argv[0] if argv else "nve_report.json"
    steps = int(argv[1]) if len(argv) > 1 else 10000
    every = int(argv[2]) if len(argv) > 2 else 100
    cfg = argv[3] if len(argv) > 3 else "c2"
    mk = configs.channel_c2 if cfg == "c2" else configs.channel_c1
    p, t, r0, v = mk(thermostat=False, rc_dpd=None, skin=0.4)
    p = p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": False}), friction=0.0, sig=0.0)
    rep = {"config": cfg, "n_part": p.n_part, "dt": p.dt, "runs": []}
    for prec in ("mixed", "fp64"):
        r = drift(p, t, r0, v, prec, steps, every)
        print(prec, "max |dE/E0| %.3e, slope %.3e per 1000 steps, %.1f s" %
              (r["max_abs_rel"], r["slope_per_1000_steps"], r["seconds"]), flush=True)
        rep["runs"].append(r)
    with open(path, "w") as f:
        json.dump(rep, f)
    print("report:", path, flush=True)


if __name__ == "__main__":
    main()
