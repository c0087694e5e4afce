"""Autotuner for the runtime choices of the GPU path.

The choices that change speed but not the pair set (and hence not the
physics) are the skin (cell size, Verlet-list radius and rebuild frequency)
and whether the pair kernels use the Verlet list.  For a workload (one
channel or a batch) and its start state, each candidate runs `steps` graph-
replayed steps after `warmup` steps and the fastest is returned with all
timings.  Results can be cached in a JSON file keyed by a workload
signature, so a production run tunes once.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

from .sim import Simulation

SKINS = (0.2, 0.3, 0.4, 0.5, 0.6)


def signature(params, precision):
    ps = params if isinstance(params, (list, tuple)) else [params]
    h = hashlib.sha1()
    for p in ps:
        h.update(repr((p.n_mon, p.n_chain, p.n_mon_d, p.n_chain_d, tuple(p.boundary), p.r_cut_max,
                       bool(p.flags.thermostat), p.flags.dpd_cut_off)).encode())
    h.update(precision.encode())
    return h.hexdigest()[:16]


def _time(params, states, precision, skin, nlist, steps, warmup):
    ps = [p.replace(skin=skin) for p in (params if isinstance(params, (list, tuple)) else [params])]
    with Simulation(ps if len(ps) > 1 else ps[0], precision=precision) as s:
        s.set_neighbor_list(nlist)
        for i, (t, r0, v) in enumerate(states):
            s.set_state(r0, v, t, isys=i)
        s.md_steps(warmup, graph=True)
        s.synchronize()
        t0 = time.perf_counter()
        s.md_steps(steps, graph=True)
        s.synchronize()
        return (time.perf_counter() - t0) / steps


def tune(params, states, precision="mixed", skins=SKINS, lists=(-1,), steps=40, warmup=10, cache=None):
    """states: list of (a_type, r0, v) per channel.  Returns
    {"skin", "neighbor_list", "ms_per_step", "timings"}."""
    key = signature(params, precision)
    if cache and os.path.exists(cache):
        with open(cache) as f:
            db = json.load(f)
        if key in db:
            return db[key]
    timings = []
    for sk in skins:
        for nl in lists:
            dt = _time(params, states, precision, sk, nl, steps, warmup)
            timings.append({"skin": sk, "neighbor_list": nl, "ms_per_step": 1e3 * dt})
    best = min(timings, key=lambda r: r["ms_per_step"])
    out = dict(best, timings=timings)
    if cache:
        db = {}
        if os.path.exists(cache):
            with open(cache) as f:
                db = json.load(f)
        db[key] = out
        with open(cache, "w") as f:
            json.dump(db, f, indent=1)
    return out
