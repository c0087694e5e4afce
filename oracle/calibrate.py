"""CPU-baseline calibration (test infrastructure; runs where oracle/_ref is
built, i.e. next to the reference checkout).

Times the reference's own per-step routines (oracle/_ref/libmfa_ref_omp.so:
the reference's Fortran sources with the default control_simulation.h, its
OpenMP fluid_fluid and serial integrator, called step by step through
oracle/ref_harness.f90) against the oracle's OpenMP restatement
(liboracle_omp.so) on identical inputs and thread counts, and prints
particle-steps/s of both and their ratio.  Each measurement runs in a child
process with OMP_NUM_THREADS set before the OpenMP runtime starts.

    python oracle/calibrate.py [threads ...]      # default: 1 8
"""
import json
import os
import subprocess
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (label, n_side (None: the 4096-bead channel), steps)
CASES = (("4096-bead channel", None, 200), ("33k-bead channel", 36, 40))


def _child(kind, n_side, steps):
    sys.path.insert(0, ROOT)
    import numpy as np
    from oracle import oracle
    from tests.test_oracle_vs_reference import ref_system
    p, t, r0, v = ref_system(n_side)
    p = p.replace(sig=(2.0 * 1.0 * 4.5 / p.dt) ** 0.5)          # the thermostat's noise on
    a0 = np.zeros_like(r0)
    if kind == "reference":
        ref = oracle.RefLib(p, omp=True)
        ref.set_state(r0, v, a0, t)
        ref.md_steps(2, 0)                                       # first-touch and cell set-up
        t0 = time.perf_counter()
        ref.md_steps(steps, 2)
        dt = time.perf_counter() - t0
        threads = ref.threads()
    else:
        o = oracle.Oracle(p, omp=True)
        r1, v1, a1, *_ = o.md_steps(r0, v, a0, t, 0, 2, noise=1)
        t0 = time.perf_counter()
        o.md_steps(r1, v1, a1, t, 2, steps, noise=1)
        dt = time.perf_counter() - t0
        threads = o.lib.orc_num_threads()
    print(json.dumps({"kind": kind, "n_part": p.n_part, "steps": steps, "seconds": dt, "threads": threads,
                      "particle_steps_per_s": p.n_part * steps / dt}))


def main(argv):
    if len(argv) > 1 and argv[1] == "--child":
        _child(argv[2], None if argv[3] == "none" else int(argv[3]), int(argv[4]))
        return
    threads = [int(x) for x in argv[1:]] or [1, 8]
    rows = []
    for label, n_side, steps in CASES:
        for th in threads:
            res = {}
            for kind in ("reference", "port"):
                env = dict(os.environ, OMP_NUM_THREADS=str(th))
                out = subprocess.run([sys.executable, os.path.abspath(__file__), "--child", kind,
                                      "none" if n_side is None else str(n_side), str(steps)],
                                     env=env, capture_output=True, text=True, check=True)
                res[kind] = json.loads(out.stdout.strip().splitlines()[-1])
            row = {"case": label, "n_part": res["port"]["n_part"], "threads": th,
                   "reference": res["reference"]["particle_steps_per_s"],
                   "port": res["port"]["particle_steps_per_s"]}
            row["port_over_reference"] = row["port"] / row["reference"]
            rows.append(row)
            print(json.dumps(row), flush=True)
    print(json.dumps({"rows": rows}))


if __name__ == "__main__":
    main(sys.argv)
