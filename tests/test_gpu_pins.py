"""GPU parity of the pair and bond virial, the DPD range beyond the LJ range, the force
carried through a rebuild, and the mixed path's force error at full size.

Bars: FP64 within 1e-10 of the oracle (relative to the largest diagonal
virial component, or to the largest force); mixed within 1e-5 (max) and
2e-6 (RMS) of the largest force, per workload."""
import numpy as np
import pytest

from mfa_brush_amd import configs
from mfa_brush_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def _perturbed(builder, sigma=0.05, seed=5, **kw):
    p, t, r0, v = builder(**kw)
    rng = np.random.default_rng(seed)
    r0 = r0 + rng.normal(0, sigma, r0.shape)
    r0[:, 0] %= p.boundary[0]
    r0[:, 1] %= p.boundary[1]
    return p, t, r0, v


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("thermostat", [False, True])
def test_pair_virial_vs_oracle(oracle_mod, precision, thermostat):
    """sum over pairs of r_ij (x) F_ij (LJ, then LJ + dissipative + random
    DPD at the same Philox noise) from the pair kernel against the oracle."""
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=thermostat, rc_dpd=1.0)
    with Simulation(p, precision=precision, t_idx0=17) as s:
        s.set_virial(True)
        s.set_state(r0, v, t)
        s.fluid_fluid()
        w = s.energies()["virial"]
        s.set_virial(False)
        s.fluid_fluid()
        assert not s.energies()["virial"].any()
    wo = oracle_mod.Oracle(p).pair_virial(r0, v, t, step=17, noise=1)
    scale = np.abs(wo[:3]).max()
    err = np.abs(w - wo).max() / scale
    print("virial %s thermostat=%s: %s vs %s, err %.2e" % (precision, thermostat, w, wo, err))
    assert err < (1e-10 if precision == "fp64" else 1e-5), err


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_dpd_beyond_lj_range(oracle_mod, precision):
    """Solvent-solvent LJ cut at 0.9 under a DPD cutoff of 1: the pairs
    between 0.9 and 1 take only the DPD term, with their own distance."""
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=True, rc_dpd=1.0)
    r2 = np.asarray(p.range_2).reshape(4, 4).T.copy()
    r2[2, 2] = 0.81
    col = lambda m: np.ascontiguousarray(m.T.reshape(-1))
    p = p.replace(range_2=col(r2), inv_range_2=col(np.where(r2 > 0, 1 / np.where(r2 > 0, r2, 1), 0)))
    with Simulation(p, precision=precision, t_idx0=3) as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f = s.get_state()["force"]
    fo, _ = oracle_mod.Oracle(p).fluid_fluid(r0, v, t, step=3, noise=1)
    err = np.linalg.norm(f - fo, axis=1).max() / np.linalg.norm(fo, axis=1).max()
    assert err < (1e-11 if precision == "fp64" else 1e-5), err


def _wclose(w, want, rtol):
    scale = np.abs(want[:3]).max()
    assert scale > 0.0
    err = np.abs(np.asarray(w) - np.asarray(want)).max() / scale
    assert err < rtol, (err, w, want)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bond_virial_vs_oracle(oracle_mod, precision):
    """The FENE bond part of the virial against the oracle's sum over bonds
    of r_ij (x) F_ij, and the whole virial (pairs + bonds) against the
    oracle's two parts: routine path (fluid_fluid, then intra_molec), fused
    steps (the pair kernel's bonded epilogue; forces evaluated at the
    positions the last step leaves), and the DPD-VV step (bonds in the
    bonded kernel).  The bond part is FP64 on both paths."""
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=False, rc_dpd=1.0)
    o = oracle_mod.Oracle(p)
    tol = 1e-10 if precision == "fp64" else 1e-5
    with Simulation(p, precision=precision) as s:
        s.set_virial(True)
        s.set_state(r0, v, t)
        s.fluid_fluid()
        s.intra_molec()
        e = s.energies()
        wbo = o.bond_virial(r0)
        _wclose(e["virial_bond"], wbo, 1e-11)
        _wclose(e["virial"], o.pair_virial(r0, v, t, noise=0) + wbo, tol)
        s.fluid_fluid()                                  # a pair evaluation clears the bond part
        assert not s.energies()["virial_bond"].any()
        s.md_steps(3)
        e = s.energies()
        r = s.get_state()["r0"]
        wbo = o.bond_virial(r)
        _wclose(e["virial_bond"], wbo, 1e-11)
        _wclose(e["virial"], o.pair_virial(r, v, t, noise=0) + wbo, tol)
        s.set_virial(False)
        s.md_steps(1)
        e = s.energies()
        assert not e["virial"].any() and not e["virial_bond"].any()
    p2, t2, r2, v2 = _dpd_vv_c1(1.0)
    with Simulation(p2, precision=precision) as s:
        s.set_virial(True)
        s.set_state(r2, v2, t2)
        s.md_steps(3)
        e = s.energies()
        r = s.get_state()["r0"]
    _wclose(e["virial_bond"], oracle_mod.Oracle(p2).bond_virial(r), 1e-11)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_force_follows_beads_through_rebuild(precision):
    """verlet_positions re-sorts the beads by cell (every step without a
    skin); the force read afterwards is still the last evaluation's force of
    each bead, as the reference's particle-indexed force array is."""
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.0)
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        s.md_steps(3)
        f0 = s.get_state()["force"]
        s.verlet_positions()
        f1 = s.get_state()["force"]
        assert s.flags()["rebuilds"] >= 4
    assert np.abs(f0).max() > 1.0
    assert np.array_equal(f0, f1)


@pytest.mark.parametrize("workload", ["c2", "c3"])
def test_full_size_force_error(oracle_mod, workload):
    """Thermostat off, full size (998,784-bead brush channel; 1.2M-bead melt
    with bending): both paths against the oracle's linked-cell forces.  The
    mixed path's error is stated per workload: max and RMS of
    |F - F_oracle| over beads, relative to the largest (RMS) oracle force."""
    if workload == "c2":
        p, t, r0, v = _perturbed(configs.channel_c2, sigma=0.03, thermostat=False, rc_dpd=1.0)
    else:
        p, t, r0, v = _perturbed(configs.channel_melt, sigma=0.03, thermostat=False, rc_dpd=1.0)
    p = p.replace(skin=0.4)
    fo, eo = oracle_mod.Oracle(p, omp=True).fluid_fluid(r0, v, t, noise=0)
    no = np.linalg.norm(fo, axis=1)
    for precision in ("fp64", "mixed"):
        with Simulation(p, precision=precision) as s:
            s.set_state(r0, v, t)
            s.fluid_fluid()
            f = s.get_state()["force"]
            e = s.energies()["v_fluid_fluid"]
        d = np.linalg.norm(f - fo, axis=1)
        err_max = d.max() / no.max()
        err_rms = np.sqrt((d ** 2).mean() / (no ** 2).mean())
        err_e = abs(e - eo) / abs(eo)
        print("%s %s: force err max %.2e rms %.2e, energy rel err %.2e (max|F| %.1f)"
              % (workload, precision, err_max, err_rms, err_e, no.max()))
        if precision == "fp64":
            assert err_max < 1e-11 and err_e < 1e-11
        else:
            assert err_max < 1e-5 and err_rms < 2e-6 and err_e < 1e-5


def test_batched_replicas_force_error(oracle_mod):
    """Thermostat off, the first 16 replicas of the 256-replica batch (2k-40k
    beads; grafting density, chain length and eps_bs drawn per replica) in
    one batched context: every replica's forces on both paths against the
    oracle's.  The mixed path's error is stated as the largest over the
    replicas of the max and RMS of |F - F_oracle| (relative to the
    replica's largest and RMS oracle force) and of the energy's relative
    error."""
    from mfa_brush_amd.params import Flags
    rng = np.random.default_rng(7)
    ps, states = [], []
    for i in range(16):
        p, t, r0, v = configs.c4_replica(i)
        p = p.replace(skin=0.4, flags=Flags(**{**p.flags.__dict__, "thermostat": False}))
        r0 = r0 + rng.normal(0, 0.03, r0.shape)
        r0[:, 0] %= p.boundary[0]
        r0[:, 1] %= p.boundary[1]
        ps.append(p)
        states.append((t, r0, v))
    ref = [oracle_mod.Oracle(p, omp=True).fluid_fluid(r0, v, t, noise=0) for p, (t, r0, v) in zip(ps, states)]
    for precision in ("fp64", "mixed"):
        worst = np.zeros(3)
        with Simulation(ps, precision=precision) as s:
            for i, (t, r0, v) in enumerate(states):
                s.set_state(r0, v, t, isys=i)
            s.fluid_fluid()
            for i, (fo, eo) in enumerate(ref):
                f = s.get_state(i)["force"]
                e = s.energies(i)["v_fluid_fluid"]
                no = np.linalg.norm(fo, axis=1)
                d = np.linalg.norm(f - fo, axis=1)
                worst = np.maximum(worst, [d.max() / no.max(), np.sqrt((d ** 2).mean() / (no ** 2).mean()),
                                           abs(e - eo) / abs(eo)])
        print("c4 (16 replicas, %d beads) %s: worst force err max %.2e rms %.2e, energy rel err %.2e"
              % (sum(p.n_part for p in ps), precision, *worst))
        if precision == "fp64":
            assert worst[0] < 1e-11 and worst[2] < 1e-11
        else:
            assert worst[0] < 1e-5 and worst[1] < 2e-6 and worst[2] < 1e-5


def test_pair_timing_same_bits():
    """Pair-kernel timing (events around each pair launch, eager launches)
    times every pair launch and leaves the trajectory's bits unchanged."""
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.4)
    outs = []
    for mode in (0, 2):
        with Simulation(p, precision="mixed") as s:
            s.set_state(r0, v, t)
            s.set_timing(mode)
            s.prepare_graphs()
            s.md_steps(45, graph=True)
            s.synchronize()
            ms, n = s.kernel_time("pair")
            outs.append(s.get_state())
        if mode == 2:
            assert n == 45 and ms > 0.0, (n, ms)
        else:
            assert n == 0
    for k in ("r0", "v", "a", "force"):
        assert np.array_equal(outs[0][k], outs[1][k]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cosine_bending_routines_and_fused(oracle_mod, precision):
    """Cosine bending k_b (1 - cos theta) of brush (bending) and melt
    (bending_melt) chains: the routine-level entry points against the
    oracle's forces and energies, and 10 fused steps against the oracle."""
    from mfa_brush_amd.params import Flags
    p = configs.channel_params((16.0, 16.0, 20.0), 16, 128, 256, thermostat=True, rc_dpd=1.0, bending=True,
                               k_bend=5.0, const_force=(0.01, 0.0, 0.0))
    p = p.replace(n_mon_d=8, r_2_min_init=0.0, r_2_min_time=0.0,
                  flags=Flags(**{**p.flags.__dict__, "bending_melt": True, "bend_cosine": True}))
    t, r0, v = configs.build_channel(p, 2.0, 8)
    rng = np.random.default_rng(4)
    r0 = r0 + rng.normal(0, 0.05, r0.shape)
    r0[:, :2] %= 16.0
    o = oracle_mod.Oracle(p)
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f0 = s.get_state()["force"]
        s.bending(1)
        f1 = s.get_state()["force"]
        s.bending_melt(1)
        f2 = s.get_state()["force"]
        e = s.energies()
    fb, eb = o.bending(r0, 0)
    fm, em = o.bending(r0, 1)
    for got, want in ((f1 - f0, fb), (f2 - f1, fm)):
        assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()
    assert e["v_bend"] == pytest.approx(eb, rel=1e-11) and e["v_bend_melt"] == pytest.approx(em, rel=1e-11)
    ro, vo, ao, fo, eo, _ = o.md_steps(r0, v, np.zeros_like(r0), t, 0, 10, noise=1)
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        s.md_steps(10, graph=True)
        st = s.get_state()
        e = s.energies()
    tol = 1e-9 if precision == "fp64" else 1e-4
    assert np.abs(st["r0"] - ro).max() <= tol * np.abs(ro).max()
    for k in ("v_bend", "v_bend_melt"):
        assert e[k] == pytest.approx(eo[k], rel=1e3 * tol), k


def _dpd_vv_c1(rc_dpd, skin=0.3):
    from mfa_brush_amd.params import Flags
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=True, rc_dpd=rc_dpd)
    heads = np.arange(p.n_chain) * p.n_mon
    v = v.copy()
    v[heads] = 0.0
    p = p.replace(skin=skin, flags=Flags(**{**p.flags.__dict__, "dpd_vv": True}))
    return p, t, r0, v


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("rc_dpd", [1.0, None])
def test_dpd_vv_vs_oracle(oracle_mod, precision, rc_dpd):
    """DPD-VV (new_dpd_fd.f90): 10 fused steps, graph-replayed, against the
    oracle's steps, which evaluate the dissipative forces again at the new
    velocities; the device adds F_D of the kick dt/2 a instead (the force is
    linear in the velocities).  The routine-level new_dpd_fd after
    verlet_velocities gives the fused step's state."""
    p, t, r0, v = _dpd_vv_c1(rc_dpd)
    o = oracle_mod.Oracle(p)
    a0 = np.zeros_like(r0)
    ro, vo, ao, fo, eo, _ = o.md_steps(r0, v, a0, t, 0, 10, noise=1)
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t, a0)
        s.md_steps(10, graph=True)
        st = s.get_state()
    tol = 1e-9 if precision == "fp64" else 1e-4
    for k, want in (("r0", ro), ("v", vo), ("a", ao), ("force", fo)):
        err = np.abs(st[k] - want).max() / np.abs(want).max()
        assert err <= (tol if k == "r0" else 1e3 * tol), (k, err)
    with Simulation(p, precision=precision) as s1, Simulation(p, precision=precision) as s2:
        s1.set_state(r0, v, t, a0)
        s2.set_state(r0, v, t, a0)
        for _ in range(3):
            s1.step_by_routines()
        s2.md_steps(3)
        a, b = s1.get_state(), s2.get_state()
    for k in ("r0", "v", "a", "force"):
        assert np.abs(a[k] - b[k]).max() <= 1e-12 * np.abs(b[k]).max(), k


def test_dpd_vv_thermostat_holds_kt():
    """DPD-VV keeps the DPD thermostat's temperature: kT = 1 within 5 % on C1
    with DPD on every LJ pair."""
    p, t, r0, v = _dpd_vv_c1(None, skin=0.4)
    with Simulation(p, precision="mixed") as s:
        s.set_state(r0, v, t)
        s.md_steps(400, graph=True)
        kes = []
        for _ in range(20):
            s.md_steps(20, graph=True)
            kes.append(s.energies()["t_fluid"])
        assert s.flags()["nonfinite"] == 0
    kt = 2.0 * np.mean(kes) / (3 * (p.n_part - p.n_chain))
    assert abs(kt - 1.0) < 0.05, kt


@pytest.mark.parametrize("blocks", [7, 64])
def test_persistent_pair_launch_same_bits(blocks):
    """The persistent launch of the mixed pair kernel (a few blocks taking
    tiles from a device counter, which they reset for the next launch) gives
    the bits of one block per tile, over graph-replayed steps with rebuilds,
    on one channel and on a batch of uneven replicas."""
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.3)
    reps = [configs.channel_scaled(n, n_mon=8, lz=14.0, seed=300 + n, skin=0.3) for n in (6, 9, 7)]
    for ps, states in (([p], [(t, r0, v)]), ([r[0] for r in reps], [(r[1], r[2], r[3]) for r in reps])):
        outs = []
        for grid in (0, blocks):
            with Simulation(ps, precision="mixed") as s:
                s.set_pair_grid(grid)
                for i, (tt, rr, vv) in enumerate(states):
                    s.set_state(rr, vv, tt, isys=i)
                s.md_steps(30, graph=True)
                outs.append([s.get_state(i) for i in range(len(ps))] + [s.energies(i)["v_fluid_fluid"] for i in range(len(ps))])
                assert s.flags()["rebuilds"] > 1
        for a, b in zip(outs[0], outs[1]):
            if isinstance(a, dict):
                for k in ("r0", "v", "a", "force"):
                    assert np.array_equal(a[k], b[k]), k
            else:
                assert a == b


def test_autotuner_choices_and_cache(tmp_path):
    """The autotuner times every candidate of each knob in turn, returns the
    fastest setting, caches it by workload signature, and the setting it
    returns runs the same pair set (same positions within the FP64 summation
    order of the deferred DPD parts)."""
    from mfa_brush_amd import autotune
    p, t, r0, v = _perturbed(configs.channel_c1, thermostat=True, rc_dpd=1.0)
    cache = str(tmp_path / "tune.json")
    r = autotune.tune([p], [(t, r0, v)], "mixed", skins=(0.3, 0.4), grids=(0, 64), queues=(8, 4),
                      lists=(-1, 1.25), steps=8, warmup=4, cache=cache)
    assert len(r["timings"]) == 8
    assert r["skin"] in (0.3, 0.4) and r["pair_grid"] in (0, 64) and r["dpd_queue"] in (8, 4)
    assert r["ms_per_step"] == min(x["ms_per_step"] for x in r["timings"][-2:])
    again = autotune.tune([p], [(t, r0, v)], "mixed", cache=cache)        # from the cache, no timing
    assert again == r
    outs = []
    for choice in (autotune.DEFAULT, r):
        with Simulation(p.replace(skin=choice["skin"]), precision="mixed") as s:
            autotune.apply(s, choice, s.neighbor_list())
            s.set_state(r0, v, t)
            s.md_steps(10, graph=True)
            outs.append(s.get_state()["r0"])
    assert np.abs(outs[0] - outs[1]).max() < 1e-9
