"""Surrogate warm start for brush channels.

The surrogate (mfa_brush_amd.surrogate.Mlp, BF16 MFMA kernels) maps
(grafting density, chain length, brush-solvent epsilon, z) to the brush and
solvent number densities at height z.  It is trained on equilibrium profiles
that the batched GPU simulator produces for a set of channels, and its
prediction for a new channel seeds the initial configuration
(configs.channel_general: monomer heights from the predicted brush density,
solvent sites weighted by the predicted solvent density).  The plain lattice
start (straight chains) remains the fallback, and plain MD decides the profile
error: time to equilibrium is measured on the simulated profiles of each start.

Time to equilibrium: the first block of `block` steps after which the
block-averaged brush and solvent profiles (n_bins bins over [0, Lz], mirrored
about the mid-plane) stay within `tol` (L1 distance relative to the L1 norm of
the reference) for `consecutive` blocks.  The reference is the equilibrium
estimate from the second halves of the runs together (equal weight); the
largest L1 distance between two runs' second-half means is reported with it,
as the check that the runs reached the same state.  Three starts are
compared: the plain lattice start, the reference's own generator start
(gen_brush.f90 mode 5 and gen_droplet.f90 case 1, restated in
generators.py) and the warm start.  No counterpart in the reference code.
"""
from __future__ import annotations

import json
import math
import os
import sys
import time
from dataclasses import dataclass, asdict

import numpy as np

from . import configs, generators
from .sim import Simulation
from .surrogate import Mlp

LZ = 40.0
N_BINS = 64
WIDTHS = [4, 128, 128, 128, 128, 2]


@dataclass
class Spec:
    sigma_g: float       # grafting density (chains per sigma^2 per wall)
    n_mon: int           # beads per grafted chain (head included)
    eps_bs: float        # brush-solvent WCA epsilon
    side: float = 20.0   # box side in sigma (rounded to whole grafting cells)
    seed: int = 1

    def a_graft(self):
        return 1.0 / math.sqrt(self.sigma_g)

    def n_side(self):
        return max(2, int(round(self.side / self.a_graft())))


def _soft(built):
    """Both starts use the reference's force switch-on (FORCE_SWITCH_ON):
    pair distances below r_min are clamped to r_min, with r_min^2 ramping
    from 0.64 to 0 over the first 1 tau (200 steps)."""
    p, t, r0, v = built
    return p.replace(r_2_min_init=0.64, r_2_min_time=1.0), t, r0, v


def build(spec: Spec, mlp: Mlp | None = None, skin: float = 0.3):
    """Plain start (mlp None) or surrogate warm start of a channel."""
    kw = dict(eps_bs=spec.eps_bs, thermostat=True, rc_dpd=None, skin=skin, seed=spec.seed, force_switch_on=True)
    if mlp is None:
        return _soft(configs.channel_general(spec.n_side(), spec.a_graft(), spec.n_mon, LZ, **kw))
    zc, rb, rs = predict(mlp, spec)
    half = zc <= LZ / 2
    z0 = 1.2
    zb, wb = zc[half], np.maximum(rb[half], 0.0) * (zc[half] >= z0)
    cdf = np.cumsum(wb)
    if cdf[-1] <= 0:
        return _soft(configs.channel_general(spec.n_side(), spec.a_graft(), spec.n_mon, LZ, **kw))
    cdf = cdf / cdf[-1]
    n = spec.n_mon

    def brush_z(m):
        if m == 0:
            return 0.0
        q = (m - 0.5) / (n - 1)
        return float(np.interp(q, cdf, zb)) - z0

    ws = np.maximum(rs, 1e-3)

    def solvent_weight(z):
        return float(np.interp(z, zc, ws))

    return _soft(configs.channel_general(spec.n_side(), spec.a_graft(), spec.n_mon, LZ, brush_z=brush_z,
                                         solvent_weight=solvent_weight, **kw))


def build_generated(spec: Spec, skin: float = 0.3):
    """The reference's own start for the channel (generators.conf_default):
    the ordered two-wall brush of gen_brush mode 5 (heads on the square
    lattice of spacing a_graft, chains as random walks), solvent beads placed
    uniformly by gen_droplet case 1 without an overlap check, and
    Maxwell-Boltzmann velocities; same bead counts, types and force switch-on
    as the plain start."""
    p, _, _, _ = build(spec, None, skin)
    a_type, r0, v = generators.conf_default(p, brush_mode=5, seed=spec.seed)
    return p, a_type, r0, v


def _bins():
    dz = LZ / N_BINS
    return (np.arange(N_BINS) + 0.5) * dz, dz


def densities(p, dens, n_obs):
    """Number densities (brush, solvent) per bin, mirrored about the mid-plane."""
    _, dz = _bins()
    area = p.boundary[0] * p.boundary[1]
    rho = dens / (max(n_obs, 1) * area * dz)
    return 0.5 * (rho + rho[:, ::-1])


def features(spec: Spec, z):
    z = np.asarray(z, float)
    return np.stack([np.full_like(z, (spec.sigma_g - 0.275) / 0.225), np.full_like(z, (spec.n_mon - 20.0) / 12.0),
                     np.full_like(z, (spec.eps_bs - 1.0) / 0.5), 2.0 * z / LZ - 1.0], 1).astype(np.float32)


def predict(mlp: Mlp, spec: Spec):
    zc, _ = _bins()
    y = mlp.forward(features(spec, zc))
    return zc, y[:, 0], y[:, 1]


def _build_plain(args):
    spec, skin = args
    return build(spec, None, skin)


def build_many(specs, skin: float = 0.3, workers: int = 16):
    """Plain starts of many channels; from 64 channels on, built by a pool of
    fresh worker processes (spawned, so they share nothing with a process
    that already uses the GPU)."""
    if len(specs) < 64 or workers <= 1:
        return [build(s, None, skin) for s in specs]
    import multiprocessing as mp
    from concurrent.futures import ProcessPoolExecutor
    n = max(1, min(workers, os.cpu_count() or 1))
    out = []
    with ProcessPoolExecutor(max_workers=n, mp_context=mp.get_context("spawn")) as ex:
        for c0 in range(0, len(specs), 1024):
            part = specs[c0:c0 + 1024]
            out.extend(ex.map(_build_plain, [(s, skin) for s in part], chunksize=16))
            print("  built %d/%d training channels" % (len(out), len(specs)), flush=True)
    return out


# slots per batched context: the Verlet-list entries address slots below 2^23
MAX_SLOTS = 7_000_000


def _groups(built, max_slots: int = MAX_SLOTS):
    """Consecutive groups of channels whose 256-aligned slot ranges fit one
    context."""
    groups, cur, slots = [], [], 0
    for b in built:
        n = (b[0].n_part + 255) // 256 * 256
        if cur and slots + n > max_slots:
            groups.append(cur)
            cur, slots = [], 0
        cur.append(b)
        slots += n
    if cur:
        groups.append(cur)
    return groups


def equilibrium_profiles(specs, equil_steps: int, avg_steps: int, every: int = 100, skin: float = 0.3,
                         workers: int = 16, halves: list | None = None):
    """Run the channels batched (as many per context as its slot range
    holds); profiles averaged over the last avg_steps steps.  Returns the
    list of (brush, solvent) densities, the summed anomaly flags and the bead
    count.  With `halves` (a list), the two halves of the averaging window
    are averaged separately and their pairs appended to it; the profile
    returned is then their mean (equal sample counts)."""
    built = build_many(specs, skin, workers)
    n_beads = int(sum(b[0].n_part for b in built))
    print("  built %d training channels (%d beads)" % (len(built), n_beads), flush=True)
    out, flags = [], {}
    groups = _groups(built)
    for gi, group in enumerate(groups):
        with Simulation([b[0] for b in group]) as sim:
            for i, (p, t, r0, v) in enumerate(group):
                sim.set_state(r0, v, t, isys=i)
            done = 0
            while done < equil_steps:
                n = min(1000, equil_steps - done)
                sim.md_steps(n, graph=True)
                done += n
                if done % 5000 == 0 or done == equil_steps:
                    print("  batch %d/%d (%d channels): equilibrate %d/%d" % (gi + 1, len(groups), len(group), done,
                                                                            equil_steps), flush=True)
            n_avg = avg_steps // every
            parts = [n_avg // 2, n_avg - n_avg // 2] if halves is not None else [n_avg]
            got = []
            q = 0
            for part in parts:
                sim.profiles_reset(N_BINS)
                for _ in range(part):
                    sim.md_steps(every, graph=True)
                    sim.profiles_accumulate()
                    q += 1
                    if q * every % 5000 == 0:
                        print("  batch %d/%d: average %d/%d" % (gi + 1, len(groups), q * every, avg_steps),
                              flush=True)
                got.append([densities(p, sim.profiles(i)[1], sim.profiles(i)[0]) for i, (p, *_r) in enumerate(group)])
            if halves is None:
                out.extend(got[0])
            else:
                halves.extend(zip(got[0], got[1]))
                out.extend(0.5 * (a + b) for a, b in zip(got[0], got[1]))
            for k, v in sim.flags().items():
                flags[k] = flags.get(k, 0) + int(v)
    return out, flags, n_beads


def training_specs(n: int, seed: int = 11):
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(n):
        specs.append(Spec(sigma_g=float(rng.uniform(0.05, 0.5)), n_mon=int(rng.choice([8, 16, 32])),
                          eps_bs=float(rng.uniform(0.5, 1.5)), seed=100 + i))
    return specs


def train(specs, profiles, steps: int = 3000, seed: int = 5, batch: int = 512, lr_final: float = 0.02):
    """The surrogate trained on the channels' profiles: Adam at 2e-3, decayed
    by a cosine to lr_final of it over the steps (1: constant)."""
    zc, _ = _bins()
    X = np.concatenate([features(s, zc) for s in specs])
    T = np.concatenate([np.stack([pr[0], pr[1]], 1) for pr in profiles]).astype(np.float32)
    mlp = Mlp(WIDTHS, seed=seed)
    hist = mlp.train(X, T, steps=steps, batch=batch, lr=2e-3, seed=seed, lr_final=lr_final)
    return mlp, hist


def block_profiles_starts(starts, n_blocks: int, block: int = 1000, every: int = 50):
    """Block-averaged (brush, solvent) densities of several starts of one
    channel (each a (p, a_type, r0, v) tuple), run side by side in one batched
    context (each channel's trajectory is the one it has alone)."""
    out = [[] for _ in starts]
    t0 = time.perf_counter()
    with Simulation([b[0] for b in starts]) as sim:
        for i, (p, t, r0, v) in enumerate(starts):
            sim.set_state(r0, v, t, isys=i)
        for b in range(n_blocks):
            sim.profiles_reset(N_BINS)
            for q in range(block // every):
                sim.md_steps(every, graph=True)
                sim.profiles_accumulate()
            for i, (p, *_r) in enumerate(starts):
                n_obs, dens, _, _ = sim.profiles(i)
                out[i].append(densities(p, dens, n_obs))
        sim.synchronize()
        fl = sim.flags()
    return [np.array(o) for o in out], starts[0][0].n_part, time.perf_counter() - t0, fl


def l1_error(prof, ref):
    return float(np.abs(prof - ref).sum() / np.abs(ref).sum())


def time_to_equilibrium(blocks, ref, tol: float = 0.02, consecutive: int = 5):
    """Index (1-based, in blocks) of the first block after which `consecutive`
    blocks stay within tol of ref; None if never."""
    err = [l1_error(b, ref) for b in blocks]
    for i in range(len(err) - consecutive + 1):
        if all(e <= tol for e in err[i:i + consecutive]):
            return i + 1, err
    return None, err


START_NAMES = ("plain", "generated", "warm")


def bounded_tte(n_train: int = 64, equil: int = 20000, avg: int = 20000, n_blocks: int = 100, block: int = 1000,
                mlp_steps: int = 6000, held: Spec | None = None, seed: int = 21, batch: int = 512):
    """The whole warm-start method at a bounded size (tens of seconds on one
    GPU): n_train small channels (12 sigma boxes, N in {8, 16}) run batched
    for equil + avg steps to give training profiles, the surrogate trains on
    them, and a held-out channel runs n_blocks blocks of `block` steps from
    three starts side by side: the plain lattice start, the reference's
    generator start (build_generated) and the warm start.  Returns the
    first-block L1 errors, the surrogate's own L1 error and the blocks to the
    10 / 5 / 2 % bands (5 consecutive blocks) of each start, against the mean
    of the second halves of all three runs, with the warm start's speedup
    over each of the other two."""
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    specs = [Spec(sigma_g=float(rng.uniform(0.05, 0.5)), n_mon=int(rng.choice([8, 16])),
                  eps_bs=float(rng.uniform(0.5, 1.5)), side=12.0, seed=300 + i) for i in range(n_train)]
    profs, _, n_beads = equilibrium_profiles(specs, equil, avg)
    t_sim = time.perf_counter() - t0
    t_train = time.perf_counter()
    mlp, hist = train(specs, profs, steps=mlp_steps, batch=batch)
    t_train = time.perf_counter() - t_train
    held = held or Spec(0.25, 16, 1.0, side=28.0, seed=907)
    starts = [build(held, None), build_generated(held), build(held, mlp)]
    runs, n_part, secs, fl = block_profiles_starts(starts, n_blocks, block=block)
    late = [r[n_blocks // 2:].mean(0) for r in runs]
    ref = sum(late) / len(late)
    zc, rb, rs = predict(mlp, held)
    out = {"held_out": asdict(held), "n_part": n_part, "block_steps": block, "blocks": n_blocks,
           "train_channels": n_train, "train_beads": n_beads, "train_steps": equil + avg,
           "mlp_steps": mlp_steps, "mlp_batch": batch, "mlp_seconds": t_train,
           "mlp_loss_first_last": [float(hist[:20].mean()), float(hist[-20:].mean())],
           "surrogate_l1": l1_error(np.stack([rb, rs]), ref),
           "late_gap_l1": max(l1_error(a, b) for a in late for b in late),
           "nonfinite": int(fl["nonfinite"]), "seconds": time.perf_counter() - t0,
           "train_sim_seconds": t_sim, "held_out_seconds": secs}
    for name, r in zip(START_NAMES, runs):
        out["first_block_l1_" + name] = l1_error(r[0], ref)
        out["l1_every_10th_block_" + name] = [round(l1_error(b, ref), 4) for b in r[::10]]
    for tol in (0.10, 0.05, 0.02):
        tte = {name: time_to_equilibrium(r, ref, tol=tol)[0] for name, r in zip(START_NAMES, runs)}
        tw = tte["warm"]
        for name in ("plain", "generated"):
            tte["speedup_vs_" + name] = (tte[name] / tw) if (tte[name] and tw) else None
        out["tte_blocks_%d%%" % round(100 * tol)] = tte
    return out


def learning_curve(n_train: int = 1024, n_held: int = 128, equil: int = 30000, avg: int = 10000,
                   sizes=(16, 32, 64, 128, 256, 512, 1024), mlp_steps: int = 4000, batch: int = 512,
                   side: float = 12.0, seed: int = 41, profiles_path: str | None = None):
    """What the surrogate gains from its training data: n_train + n_held
    channels (box side `side`, grafting density U[0.05, 0.5], N in {8, 16},
    eps_bs U[0.5, 1.5]) run batched to equilibrium profiles; the surrogate is
    trained on the first n channels for each n in sizes and scored on the
    n_held held-out channels by the relative L1 distance of its predicted
    (brush, solvent) profiles from their simulated ones.  The score of the
    lattice start's own profile (the uniform solvent and straight chains a
    plain start places) is given beside it as the no-surrogate baseline.
    profiles_path (.npz): the simulated profiles are loaded from it when it
    exists (the same channels: same counts, seed and run lengths), else
    written to it, so the training can be repeated without the runs."""
    t0 = time.perf_counter()
    rng = np.random.default_rng(seed)
    specs = [Spec(sigma_g=float(rng.uniform(0.05, 0.5)), n_mon=int(rng.choice([8, 16])),
                  eps_bs=float(rng.uniform(0.5, 1.5)), side=side, seed=5000 + i) for i in range(n_train + n_held)]
    halves = []
    key = np.array([n_train, n_held, equil, avg, seed, side])
    if profiles_path and os.path.exists(profiles_path):
        with np.load(profiles_path) as z:
            assert np.array_equal(z["key"], key), "profiles file is for other channels"
            halves = list(zip(z["half_a"], z["half_b"]))
            profs = [0.5 * (a + b) for a, b in halves]
            flags, n_beads = {}, int(z["beads"])
    else:
        profs, flags, n_beads = equilibrium_profiles(specs, equil, avg, halves=halves)
        if profiles_path:
            np.savez(profiles_path, key=key, beads=n_beads, half_a=np.array([a for a, _ in halves]),
                     half_b=np.array([b for _, b in halves]))
    t_sim = time.perf_counter() - t0
    # noise of a simulated profile: with independent halves A and B of the
    # averaging window, their mean M lies about |A - B| / 2 = |A - M| from
    # the noiseless profile, so L1(A, M) estimates how far even an exact
    # surrogate would score from the held-out profiles
    noise = [l1_error(a, 0.5 * (a + b)) for a, b in halves[n_train:]]
    train_specs, held_specs = specs[:n_train], specs[n_train:]
    train_profs, held_profs = profs[:n_train], profs[n_train:]
    zc, dz = _bins()

    def lattice_profile(spec):
        # the plain start's own density profile: straight chains (one bead
        # per sigma from the wall) and uniform solvent filling the rest
        p, _, r0, _ = build(spec, None)
        area = p.boundary[0] * p.boundary[1]
        n_brush = p.n_chain * p.n_mon            # brush beads come first, as the device profiles count them
        h_b = np.histogram(r0[:n_brush, 2], bins=N_BINS, range=(0.0, p.boundary[2]))[0]
        h_s = np.histogram(r0[n_brush:, 2], bins=N_BINS, range=(0.0, p.boundary[2]))[0]
        rho = np.stack([h_b, h_s]) / (area * dz)
        return 0.5 * (rho + rho[:, ::-1])

    base = [l1_error(lattice_profile(s), pr) for s, pr in zip(held_specs, held_profs)]
    curve = []
    for n in sizes:
        if n > n_train:
            break
        t1 = time.perf_counter()
        mlp, hist = train(train_specs[:n], train_profs[:n], steps=mlp_steps, batch=batch)
        errs = []
        for s, pr in zip(held_specs, held_profs):
            _, rb, rs = predict(mlp, s)
            errs.append(l1_error(np.stack([rb, rs]), pr))
        errs = np.array(errs)
        curve.append({"n_train": n, "samples": n * N_BINS, "held_l1_mean": float(errs.mean()),
                      "held_l1_median": float(np.median(errs)), "held_l1_p90": float(np.quantile(errs, 0.9)),
                      "loss_last": float(hist[-20:].mean()), "seconds": time.perf_counter() - t1})
        print(json.dumps(curve[-1]), flush=True)
    return {"n_train": n_train, "n_held": n_held, "side": side, "equil_steps": equil, "avg_steps": avg,
            "beads": n_beads, "sim_seconds": t_sim, "flags": flags, "mlp_steps": mlp_steps, "mlp_batch": batch,
            "lattice_start_l1_mean": float(np.mean(base)), "lattice_start_l1_median": float(np.median(base)),
            "held_noise_l1_mean": float(np.mean(noise)), "held_noise_l1_median": float(np.median(noise)),
            "curve": curve, "seconds": time.perf_counter() - t0}


def main_curve(argv):
    """--curve report.json [n_train n_held equil avg mlp_steps [profiles.npz]]: learning_curve, written as JSON."""
    out_path = argv[0]
    d = [1024, 128, 30000, 10000, 4000]
    vals = [int(x) for x in argv[1:6]]
    vals = vals + d[len(vals):]
    prof = argv[6] if len(argv) > 6 else None
    sizes = tuple(n for n in (16, 32, 64, 128, 256, 512, 1024, 2048, 4096, 8192) if n <= vals[0])
    r = learning_curve(n_train=vals[0], n_held=vals[1], equil=vals[2], avg=vals[3], sizes=sizes, mlp_steps=vals[4],
                       profiles_path=prof)
    print(json.dumps({k: v for k, v in r.items() if k != "curve"}), flush=True)
    with open(out_path, "w") as f:
        json.dump(r, f, indent=1)


def main_tte(argv):
    """--tte report.json [n_train equil avg n_blocks side sigma_g n_mon eps_bs
    mlp_steps batch]: one bounded_tte run, written as JSON."""
    out_path = argv[0]
    vals = [float(x) for x in argv[1:]]
    d = [64, 20000, 20000, 100, 28.0, 0.25, 16, 1.0, 6000, 512]
    vals = vals + d[len(vals):]
    held = Spec(vals[5], int(vals[6]), vals[7], side=vals[4], seed=907)
    r = bounded_tte(n_train=int(vals[0]), equil=int(vals[1]), avg=int(vals[2]), n_blocks=int(vals[3]), held=held,
                    mlp_steps=int(vals[8]), batch=int(vals[9]))
    print(json.dumps(r), flush=True)
    with open(out_path, "w") as f:
        json.dump(r, f, indent=1)


def main(argv=None):
    """Train on batched runs, then compare plain and warm starts on held-out
    channels.  Arguments: report path, training channels, equilibration and
    averaging steps of the training runs, blocks of 1000 steps per held-out
    run, held-out box side (sigma), optional training-profile file (.npz:
    loaded if present, else written after the training runs), optional
    comma-separated held-out indices (0-2, default all).  Writes a JSON
    report."""
    argv = sys.argv[1:] if argv is None else argv
    if argv and argv[0] == "--tte":
        return main_tte(argv[1:])
    if argv and argv[0] == "--curve":
        return main_curve(argv[1:])
    out_path = argv[0] if argv else "warmstart_report.json"
    n_train = int(argv[1]) if len(argv) > 1 else 48
    equil, avg = (int(argv[2]), int(argv[3])) if len(argv) > 3 else (40000, 20000)
    n_blocks = int(argv[4]) if len(argv) > 4 else 100
    side = float(argv[5]) if len(argv) > 5 else 56.0
    prof_path = argv[6] if len(argv) > 6 and argv[6] != "-" else None
    which = [int(x) for x in argv[7].split(",") if x] if len(argv) > 7 else [0, 1, 2]
    rep = {"n_train": n_train, "equil_steps": equil, "avg_steps": avg, "n_bins": N_BINS, "lz": LZ}
    t0 = time.perf_counter()
    specs = training_specs(n_train)
    if prof_path and os.path.exists(prof_path):
        with np.load(prof_path) as z:
            profs = list(z["profiles"])
            assert len(profs) == n_train and int(z["equil"]) == equil and int(z["avg"]) == avg
        rep["train_profiles_from"] = prof_path
    else:
        profs, fl, nb = equilibrium_profiles(specs, equil, avg)
        rep["train_sim_seconds"] = time.perf_counter() - t0
        rep["train_beads"] = nb
        rep["train_flags"] = fl
        print("training profiles: %.1f s" % rep["train_sim_seconds"], flush=True)
        if prof_path:
            np.savez(prof_path, profiles=np.array(profs), equil=equil, avg=avg)
    t0 = time.perf_counter()
    mlp, hist = train(specs, profs)
    rep["train_mlp_seconds"] = time.perf_counter() - t0
    rep["loss_first_last"] = [float(hist[:20].mean()), float(hist[-20:].mean())]
    held = [Spec(0.30, 32, 1.0, side=side, seed=901), Spec(0.12, 16, 0.8, side=side, seed=902),
            Spec(0.45, 32, 1.3, side=side, seed=903)]
    rep["held_out"] = []
    for s in [held[i] for i in which]:
        starts = [build(s, None), build_generated(s), build(s, mlp)]
        runs, n_part, secs, fl = block_profiles_starts(starts, n_blocks)
        late = [r[n_blocks // 2:].mean(0) for r in runs]
        ref = sum(late) / len(late)
        zc, rb, rs = predict(mlp, s)
        r = {"spec": asdict(s), "n_part": n_part, "blocks": n_blocks, "block_steps": 1000,
             "surrogate_l1_vs_reference": l1_error(np.stack([rb, rs]), ref),
             "late_gap_l1": max(l1_error(a, b) for a in late for b in late), "seconds": secs, "flags": fl}
        for tol in (0.10, 0.05, 0.02):
            tte = {name: time_to_equilibrium(x, ref, tol=tol)[0] for name, x in zip(START_NAMES, runs)}
            for name in ("plain", "generated"):
                tte["speedup_vs_" + name] = (tte[name] / tte["warm"]) if (tte[name] and tte["warm"]) else None
            r["tte_blocks_%d%%" % round(100 * tol)] = tte
        for name, x in zip(START_NAMES, runs):
            r["err_" + name] = [l1_error(b, ref) for b in x]
        print(json.dumps({k: v for k, v in r.items() if not k.startswith("err_")}), flush=True)
        rep["held_out"].append(r)
    with open(out_path, "w") as f:
        json.dump(rep, f, indent=1)
    print("report:", out_path, flush=True)


if __name__ == "__main__":
    main()
