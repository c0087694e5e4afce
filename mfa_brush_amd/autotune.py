"""Autotuner for the runtime choices of the GPU path.

The knobs change speed, not the pair set (hence not the physics):

* skin: cell size, Verlet-list radius and rebuild frequency;
* neighbor_list: Verlet-list capacity per bead (-1 automatic, 0 none, or a
  multiple of the automatic capacity);
* pair_grid: launch shape of the mixed pair kernel (0 one block per tile of
  256 slots, > 0 a persistent launch of that many blocks over a work
  counter);
* dpd_queue: DPD pairs per bead deferred to the second pass (0..8).

For a workload (one channel or a batch) and its start state, each candidate
runs `steps` graph-replayed steps after `warmup` steps.  The knobs are tuned
one after the other (coordinate descent: each in turn over its candidates
with the others at their best so far), and the best setting is returned with
every timing.  Results can be cached in a JSON file keyed by a workload
signature, so a production run tunes once.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

from .sim import Simulation

SKINS = (0.2, 0.3, 0.4, 0.5)
PAIR_GRIDS = (0, 1536, 2048)
DPD_QUEUES = (8, 6, 4)
LIST_SCALES = (-1, 0.8, 1.25)      # -1: automatic; x: that multiple of the automatic capacity

DEFAULT = {"skin": 0.4, "neighbor_list": -1, "pair_grid": 0, "dpd_queue": 8}


def signature(params, precision):
    ps = params if isinstance(params, (list, tuple)) else [params]
    h = hashlib.sha1()
    for p in ps:
        h.update(repr((p.n_mon, p.n_chain, p.n_mon_d, p.n_chain_d, tuple(p.boundary), p.r_cut_max,
                       bool(p.flags.thermostat), p.flags.dpd_cut_off)).encode())
    h.update(precision.encode())
    return h.hexdigest()[:16]


def apply(s: Simulation, choice: dict, auto_cap: int | None = None):
    """Set a tuned choice on a Simulation (before its states are uploaded)."""
    nl = choice.get("neighbor_list", -1)
    if isinstance(nl, float) and nl > 0 and auto_cap:
        nl = max(8, int(round(nl * auto_cap / 8.0)) * 8)
    s.set_neighbor_list(int(nl) if nl == -1 or nl >= 0 else -1)
    if s.precision == "mixed":
        s.set_pair_grid(int(choice.get("pair_grid", 0)))
    s.set_dpd_queue(int(choice.get("dpd_queue", 8)))


def _time(params, states, precision, choice, steps, warmup):
    ps = [p.replace(skin=choice["skin"]) for p in (params if isinstance(params, (list, tuple)) else [params])]
    with Simulation(ps if len(ps) > 1 else ps[0], precision=precision) as s:
        auto_cap = s.neighbor_list()
        apply(s, choice, auto_cap)
        for i, (t, r0, v) in enumerate(states):
            s.set_state(r0, v, t, isys=i)
        s.md_steps(warmup, graph=True)
        s.prepare_graphs()
        s.synchronize()
        t0 = time.perf_counter()
        s.md_steps(steps, graph=True)
        s.synchronize()
        return (time.perf_counter() - t0) / steps


def tune(params, states, precision="mixed", skins=SKINS, grids=PAIR_GRIDS, queues=DPD_QUEUES,
         lists=LIST_SCALES, steps=40, warmup=10, cache=None):
    """states: list of (a_type, r0, v) per channel.  Returns the best
    {"skin", "neighbor_list", "pair_grid", "dpd_queue", "ms_per_step"} plus
    "timings" (every candidate timed, in order)."""
    key = signature(params, precision)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            return db[key]
    best = dict(DEFAULT)
    timings = []
    axes = [("skin", skins), ("pair_grid", grids if precision == "mixed" else (0,)), ("dpd_queue", queues),
            ("neighbor_list", lists)]
    best_ms = None
    for name, cands in axes:
        scored = []
        for v in cands:
            choice = dict(best, **{name: v})
            ms = 1e3 * _time(params, states, precision, choice, steps, warmup)
            timings.append(dict(choice, ms_per_step=ms))
            scored.append((ms, v))
        ms, v = min(scored, key=lambda x: x[0])
        best[name] = v
        best_ms = ms
    out = dict(best, ms_per_step=best_ms, timings=timings)
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = out
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return out
