"""Throughput of the per-step MD path on one MI355X.

Workload: the 998,784-bead brush channel (box 102*sqrt2 x 102*sqrt2 x 60,
2 x 102^2 grafted chains of 32 at grafting density 0.5, 332,928 solvent,
LJ rc 2.5 / brush-solvent WCA, DPD gamma 4.5 rc 1, kT 1, dt 0.005, 9-3
walls), synthetic initial state (seeded), DPD thermostat on, cells rebuilt
on the skin criterion (skin 0.4 by default, --skin).  A step is one full md_main iteration:
integrate, cell maintenance, LJ+DPD pairs, bonds, walls, second kick.

value = beads x steps / seconds of the mixed-precision path (FP32 pair
arithmetic, FP64 integration and sums): after 100 warm-up steps, 5 timed runs
of --steps steps each, one after the other, and the median run is the value
(every run's rate is in "runs"); the plain FP64 path is timed the same way and
reported beside it.  The pair kernel's average launch duration comes from HIP
events on the library's stream over the timed steps.

time_to_equilibrium: a bounded run of the learned warm start (training
profiles from 64 small batched channels, surrogate training, then a held-out
25k-bead channel run for 100 blocks of 1000 steps from the lattice start, the
reference's generator start and the warm start side by side), blocks to the
10 / 5 / 2 % profile bands for the three starts.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = 79          # LJ 41 + DPD 38 per interacting pair (thermostat on)
PEAK_FP32 = 157.3e12         # MI355X vector FP32
PEAK_FP64 = 78.6e12          # MI355X vector FP64
HBM_PEAK = 8.0e12
# The oracle's OpenMP restatement runs this many times faster than the
# reference's own routines (flang -O3, OpenMP fluid_fluid, serial
# integrator) on identical inputs and thread counts: measured by
# oracle/calibrate.py on an 8-core Xeon host (33,178-bead channel, 8 threads:
# 1.08e6 vs 4.11e5 particle-steps/s; 1 thread: 1.48x; the 4096-bead channel:
# 2.59x at 8 threads, 1.42x at 1).  reference_equivalent_value divides the
# port's rate by it.
PORT_OVER_REFERENCE = 2.63
CALIBRATION = ("oracle/calibrate.py: port/reference = 2.63 at 8 threads (33k beads), 1.48 at 1 thread; the "
               "reference's serial integrator weighs more at higher thread counts, so the reference-"
               "equivalent figure is an upper bound on the reference's own rate")


def run_gpu(precision, steps, warmup, ps, states, nlist=-1, choice=None, repeats=1):
    """Warm-up steps (graph replay), then `repeats` timed runs of `steps`
    steps each, one after the other, each bracketed by a synchronisation:
    the steps are launched in order with a pair of HIP events around each
    pair-kernel launch on the library's stream (so the pair kernel's average
    launch duration comes from the same steps as the wall time; two event
    records per step cost about 2.5 % against graph replay, eight about
    15 %).  After them, outside the timed region: steps with events around
    every kernel group for the per-kernel breakdown, and steps replayed as
    graphs without events.  ps and states list the channels of one
    (possibly batched) context."""
    from mfa_brush_amd.sim import Simulation
    s = Simulation(ps if len(ps) > 1 else ps[0], precision=precision)
    if choice is not None:
        from mfa_brush_amd import autotune
        autotune.apply(s, choice, s.neighbor_list())
    else:
        s.set_neighbor_list(nlist)
    for i, (t, r0, v) in enumerate(states):
        s.set_state(r0, v, t, isys=i)
    s.md_steps(warmup, graph=True)
    s.set_timing(2)
    runs = []
    for _ in range(repeats):
        s.synchronize()
        t0 = time.perf_counter()
        s.md_steps(steps, graph=True)
        s.synchronize()
        runs.append(time.perf_counter() - t0)
    pair_ms, pair_n = s.kernel_time("pair")
    n_break = max(8, min(steps, 40))
    s.set_timing(1)
    s.md_steps(n_break)
    s.synchronize()
    tot = {k: s.kernel_time(k) for k in ("vv1", "rebuild", "pair", "bonded")}
    s.set_timing(0)
    s.prepare_graphs()
    s.synchronize()
    t0 = time.perf_counter()
    s.md_steps(steps, graph=True)
    s.synchronize()
    dt_plain = time.perf_counter() - t0
    npairs = sum(s.pair_count(i) for i in range(len(ps)))
    fl = s.flags()
    cap = s.neighbor_list()
    lists = None
    if cap > 0:                     # the lists of the last rebuild, summed over the channels
        per = [s.list_stats(i) for i in range(len(ps))]
        lists = {k: sum(x[k] for x in per) for k in per[0]}
    s.close()
    return dict(seconds=sorted(runs)[len(runs) // 2], runs=runs, seconds_plain=dt_plain, pair_ms=pair_ms,
                pair_n=pair_n, npairs=npairs, flags=fl, list_cap=cap, lists=lists,
                kernels={k: {"total_ms": a, "launches": b} for k, (a, b) in tot.items()})


def build_workload(name, skin, n_rep=256):
    """(params list, states list, description) of a bench workload."""
    from mfa_brush_amd import configs
    if name == "c2":
        p, t, r0, v = configs.channel_c2(thermostat=True, rc_dpd=1.0, skin=skin)
        return [p], [(t, r0, v)], ("brush channel 998,784 beads, box 144.25x144.25x60, 2x10404 chains N=32, "
                                   "332,928 solvent, LJ rc 2.5 + DPD rc 1, dt 0.005")
    if name == "c3":
        p, t, r0, v = configs.channel_melt(thermostat=True, rc_dpd=1.0, skin=skin)
        return [p], [(t, r0, v)], ("melt channel 1,200,000 beads, 18,750 chains N=64, box 216.93x216.93x30, "
                                   "FENE + cosine bending k_b 5, Poiseuille g 0.01 x, DPD rc 1, dt 0.005")
    if name == "c4":
        reps = configs.c4_batch(n_rep)
        ps = [r[0].replace(skin=skin) for r in reps]
        return ps, [(r[1], r[2], r[3]) for r in reps], (
            "%d brush-channel replicas, 2k-40k beads log-uniform, grafting density U[0.05,0.5], N in "
            "{8,16,32,64}, eps_bs U[0.5,1.5], Lz 30 (more where straight chains need it), DPD rc 1, one batched "
            "context" % n_rep)
    raise ValueError(name)


def traffic_from_profile():
    """HBM bytes per pair-kernel launch from the committed rocprofv3 --pmc
    summary (FETCH_SIZE + WRITE_SIZE, scaled by a plain copy's reading), or
    None when no such summary is present."""
    f = os.path.join(ROOT, "profiles", "k_pair32_c2_traffic.md")
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


# kernel groups the library times (per launch of the group): the first half
# kick once per md_steps call, the gated rebuild and the pair kernel every
# step, and the integrator group every step (bending when a channel bends,
# then k_integrate: the second half kick and the next step's first half)
KERNEL_LABELS = {"vv1": "vv1_once_per_call", "rebuild": "rebuild", "pair": "pair", "bonded": "integrate"}


def kernels_ms(kernels):
    return {KERNEL_LABELS.get(k, k): d["total_ms"] / max(d["launches"], 1) for k, d in kernels.items()}


def neighbor_list(res, ps):
    """Verlet-list figures of a run: capacity, overflows, and from the
    lists of the last rebuild the entries per bead and the entries a wave of
    64 slots walks (its longest lane's list, in groups of four)."""
    out = {"capacity": res["list_cap"], "overflowed_beads": res["flags"]["list_overflow"]}
    if res["lists"]:
        n = sum(p.n_part for p in ps)
        waves = sum((p.n_part + 63) // 64 for p in ps)
        out["entries_per_bead"] = res["lists"]["entries"] / n
        out["walked_per_wave"] = res["lists"]["wave_walk"] / waves
    return out


def calib_copy(mib):
    """A plain device-to-device copy of `mib` MiB (read + write of that many
    bytes), the reference reading for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print("calib copy: %d MiB x 3" % mib, file=sys.stderr, flush=True)


def cpu_baseline(ps, states, n_steps, max_channels=16):
    """The oracle's OpenMP build (same algorithm as the reference: linked
    cells, half shell, full-array force reduction, serial integrator) on
    this host's cores, on a bounded sample: n_steps steps of the first
    max_channels channels of the workload."""
    import numpy as np
    from oracle.oracle import Oracle
    beads, dt, threads = 0, 0.0, 1
    for p, (t, r0, v) in list(zip(ps, states))[:max_channels]:
        o = Oracle(p, omp=True)
        threads = o.lib.orc_num_threads()
        t0 = time.perf_counter()
        o.md_steps(r0, v, np.zeros_like(r0), t, 0, n_steps, noise=1)
        dt += time.perf_counter() - t0
        beads += p.n_part
    value = beads * n_steps / dt
    return dict(value=value, unit="particle-steps/s", cores=threads, kind="port",
                sample="%d steps of %d channel(s) of the workload, %d beads (%.1f s)" %
                       (n_steps, min(len(ps), max_channels), beads, dt),
                port_over_reference=PORT_OVER_REFERENCE,
                reference_equivalent_value=value / PORT_OVER_REFERENCE,
                calibration=CALIBRATION)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200, help="steps per timed run")
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--repeats", type=int, default=5,
                    help="timed runs of --steps steps, one after the other; value is the median run")
    ap.add_argument("--cpu-steps", type=int, default=80,
                    help="steps of the CPU baseline sample (10-20 s of CPU work at 1M beads on 16 cores)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tte", action="store_true",
                    help="skip the bounded time-to-equilibrium run of the learned warm start")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-list", action="store_true", help="pair kernels scan the cells every step")
    ap.add_argument("--skin", type=float, default=0.4,
                    help="Verlet skin (cells and lists); the pair set does not depend on it")
    ap.add_argument("--workload", choices=("c2", "c3", "c4"), default="c2",
                    help="c2: 1M-bead brush channel (the metric); c3: 1.2M-bead melt with bending and "
                         "Poiseuille flow; c4: 256 batched replicas")
    ap.add_argument("--autotune", action="store_true",
                    help="pick skin, list capacity, pair-kernel launch shape and DPD queue depth by timing short "
                         "runs first (coordinate descent, mfa_brush_amd/autotune.py)")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="run a plain device copy of this size first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: there is no multi-GPU path (a channel would shard into x-slabs "
              "with a halo exchange; replicas shard by replica)", file=sys.stderr)
        sys.exit(2)

    if args.calib_copy_mib > 0:
        calib_copy(args.calib_copy_mib)
    ps, states, desc = build_workload(args.workload, args.skin)
    n = sum(p.n_part for p in ps)
    print("workload %s: %d channel(s), %d beads" % (args.workload, len(ps), n), file=sys.stderr, flush=True)
    tuned = None
    if args.autotune:
        from mfa_brush_amd import autotune
        tuned = autotune.tune(ps, states, "mixed")
        print("autotune:", json.dumps(tuned), file=sys.stderr, flush=True)
        ps = [p.replace(skin=tuned["skin"]) for p in ps]
        args.skin = tuned["skin"]

    nlist = 0 if args.no_list else -1
    if args.repeats < 1:
        ap.error("--repeats must be at least 1")
    mixed = run_gpu("mixed", args.steps, args.warmup, ps, states, nlist, tuned, args.repeats)
    print("mixed: median %.4f s for %d steps (%d runs)" % (mixed["seconds"], args.steps, args.repeats),
          file=sys.stderr, flush=True)
    fp64 = None
    if not args.no_fp64:
        fp64 = run_gpu("fp64", args.steps, args.warmup, ps, states, nlist, None, args.repeats)
        print("fp64: median %.4f s for %d steps (%d runs)" % (fp64["seconds"], args.steps, args.repeats),
              file=sys.stderr, flush=True)

    value = n * args.steps / mixed["seconds"]
    pair_avg_s = mixed["pair_ms"] / max(mixed["pair_n"], 1) / 1e3
    achieved = FLOPS_PER_PAIR * mixed["npairs"] / pair_avg_s
    out = {
        "metric": "particle-steps/s @1M beads, mixed vs plain FP64 GPU; time-to-equilibrium",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "repeats": args.repeats,
        "runs": [n * args.steps / t for t in mixed["runs"]],
        "ms_per_step": 1e3 * mixed["seconds"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-pair/fp64-integrate",
        "data": "synthetic (seeded column-lattice start, Maxwell-Boltzmann velocities)",
        "config": {"workload": desc, "skin": args.skin, "channels": len(ps),
                   "n_beads": n, "interacting_pairs": mixed["npairs"]},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP32, "traffic": traffic_from_profile(),
                     "kernel": "k_pair32 (vector-ALU FP32 bound; the pair path uses no MFMA)",
                     "flops_per_launch": FLOPS_PER_PAIR * mixed["npairs"], "avg_launch_us": pair_avg_s * 1e6,
                     "pair_launches_timed": mixed["pair_n"]},
        "kernels_ms": kernels_ms(mixed["kernels"]),
        "graph_no_event_ms_per_step": 1e3 * mixed["seconds_plain"] / args.steps,
        "rebuilds": mixed["flags"]["rebuilds"],
        "neighbor_list": neighbor_list(mixed, ps),
    }
    if fp64 is not None:
        v64 = n * args.steps / fp64["seconds"]
        p64 = fp64["pair_ms"] / max(fp64["pair_n"], 1) / 1e3
        out["fp64"] = {"value": v64, "runs": [n * args.steps / t for t in fp64["runs"]],
                       "ms_per_step": 1e3 * fp64["seconds"] / args.steps,
                       "graph_no_event_ms_per_step": 1e3 * fp64["seconds_plain"] / args.steps,
                       "pair_avg_us": p64 * 1e6,
                       "pair_frac_fp64_peak": FLOPS_PER_PAIR * fp64["npairs"] / p64 / PEAK_FP64,
                       "kernels_ms": kernels_ms(fp64["kernels"])}
        out["mixed_over_fp64"] = value / v64
    if not args.no_tte:
        from mfa_brush_amd.warmstart import bounded_tte
        out["time_to_equilibrium"] = bounded_tte()
        print("time to equilibrium: %.1f s" % out["time_to_equilibrium"]["seconds"], file=sys.stderr, flush=True)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(ps, states, args.cpu_steps)
    if tuned is not None:
        out["autotune"] = tuned
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
