"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: the interacting-pair set is bit-exact (per-bead partner counts and
hashes, and the kernel's pair count); FP64 forces agree with the oracle to
max|dF| <= 1e-11 max|F| (summation order differs: full shell per bead on the
GPU, half shell + Newton-3 in the reference); energies to 1e-11 relative;
multi-step trajectories to 1e-9 over 10 steps; the mixed path's forces to
<= 1e-5 of max|F| with the identical pair set.
"""
import math

import numpy as np
import pytest

from mfa_brush_amd import configs
from mfa_brush_amd.params import Flags
from mfa_brush_amd.sim import Simulation
from conftest import REF_INPUT_CASES

pytestmark = pytest.mark.gpu

FTOL = 1e-11


def _perturbed_c1(seed=5, **kw):
    p, t, r0, v = configs.channel_c1(**kw)
    rng = np.random.default_rng(seed)
    r0 = r0 + rng.normal(0, 0.05, r0.shape)
    r0[:, :2] %= 16.0
    return p, t, r0, v


def _close(a, b, tol=FTOL):
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    assert err <= tol, "max rel err %.3e > %.1e" % (err, tol)


def _oracle(oracle_mod, p):
    return oracle_mod.Oracle(p)


def test_try_fixture_golden_on_gpu(try_fixture, oracle_mod):
    p = try_fixture["p"]
    with Simulation(p) as s:
        s.set_state(try_fixture["r0"], try_fixture["v"], try_fixture["a_type"], try_fixture["a"])
        assert abs(s.check_fluid_fluid() - try_fixture["v_ff"]) < 2e-10
        pt = p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": False}))
    with Simulation(pt) as s:
        s.set_state(try_fixture["r0"], try_fixture["v"], try_fixture["a_type"], try_fixture["a"])
        s.fluid_fluid()
        e = s.energies()["v_fluid_fluid"]
        assert abs(e - try_fixture["v_ff"]) < 1e-9
        f = s.get_state()["force"]
        fo, _ = _oracle(oracle_mod, pt).fluid_fluid(try_fixture["r0"], try_fixture["v"], try_fixture["a_type"], noise=0)
        _close(f, fo)
        cnt, h = s.pair_stats()
        co, ho = _oracle(oracle_mod, pt).pair_stats(try_fixture["r0"], try_fixture["a_type"])
        assert np.array_equal(cnt, co) and np.array_equal(h, ho)
        assert s.pair_count() == co.sum() // 2


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_try_trajectory_follows_fort41(try_fixture, precision):
    """From try/conf_old, 6 fused calls of 5 steps: V_ff, V_bond and V_wall
    after each follow the reference's own try/fort.41 record (its ziggurat
    noise against our Philox: V_ff within 0.5 %, bond and wall within 1e-4,
    the oracle's bars in tests/test_oracle_pins.py)."""
    from test_oracle_pins import FORT41_TOL, fort41_errors
    p = try_fixture["p"]
    ens = []
    with Simulation(p, precision=precision) as s:
        s.set_state(try_fixture["r0"], try_fixture["v"], try_fixture["a_type"], try_fixture["a"])
        for _ in range(6):
            s.md_steps(5)
            ens.append(s.energies())
    for i, err in enumerate(fort41_errors(ens, try_fixture["fort41"][:6])):
        for k, tol in FORT41_TOL.items():
            assert err[k] < tol, (5 * (i + 1), k, err[k])


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("name", REF_INPUT_CASES)
def test_ref_input_sets_vs_oracle(oracle_mod, name, precision):
    """The reference's shipped channel configurations (inputs/<name>, with
    the flags of their control_simulation.h): the brute-force energy, the
    pair set (per-bead partner counts and hashes) and the pair forces with
    the DPD terms at the same Philox noise against the oracle (FP64 1e-11 of
    max|F|, mixed 1e-5), and 10 fused FP64 steps against the oracle's."""
    from conftest import load_ref_input
    p, t, r0, v = load_ref_input(name)
    p = p.replace(skin=0.3)
    o = _oracle(oracle_mod, p)
    step = 11
    with Simulation(p, precision=precision, t_idx0=step) as s:
        s.set_state(r0, v, t)
        assert s.check_fluid_fluid() == pytest.approx(o.check_fluid_fluid(r0, t), rel=1e-10)
        s.fluid_fluid()
        f = s.get_state()["force"]
        cnt, h = s.pair_stats()
    fo, eo = o.fluid_fluid(r0, v, t, step=step, noise=1)
    co, ho = o.pair_stats(r0, t)
    assert np.array_equal(cnt, co) and np.array_equal(h, ho)
    _close(f, fo, FTOL if precision == "fp64" else 1e-5)
    if precision == "fp64":
        with Simulation(p, t_idx0=step) as s:
            s.set_state(r0, v, t)
            s.md_steps(10)
            st = s.get_state()
        ro, vo, *_ = o.md_steps(r0, v, np.zeros_like(r0), t, step, 10, noise=1)
        _close(st["r0"], ro, 1e-9)
        _close(st["v"], vo, 1e-6)


@pytest.mark.parametrize("thermostat,noise,rc_dpd", [(False, 0, None), (True, 0, None), (True, 1, None),
                                                      (True, 1, 1.0)])
def test_pair_forces_fp64(oracle_mod, thermostat, noise, rc_dpd):
    p, t, r0, v = _perturbed_c1(thermostat=thermostat, rc_dpd=rc_dpd)
    if thermostat and noise == 0:
        p = p.replace(sig=0.0)                      # dissipative part only
    step = 17
    with Simulation(p, t_idx0=step) as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f = s.get_state()["force"]
        e = s.energies()["v_fluid_fluid"]
        np_gpu = s.pair_count()
    fo, eo = _oracle(oracle_mod, p).fluid_fluid(r0, v, t, step=step, noise=1)
    _close(f, fo)
    assert e == pytest.approx(eo, rel=1e-11)
    cnt, _ = _oracle(oracle_mod, p).pair_stats(r0, t)
    assert np_gpu == cnt.sum() // 2


@pytest.mark.parametrize("l_box,skin", [(1, 0.0), (2, 0.0), (1, 0.4), (2, 0.3)])
def test_cell_geometries_same_forces(oracle_mod, l_box, skin):
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=skin, flags=Flags(**{**p.flags.__dict__, "l_box": l_box}))
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f = s.get_state()["force"]
    fo, _ = _oracle(oracle_mod, p).fluid_fluid(r0, v, t, step=0, noise=1)
    _close(f, fo)


def _chain_system(bending_melt=False):
    """Brush (128 x 16) + melt chains (256 x 8) with bending and a body force."""
    p = configs.channel_params((16.0, 16.0, 20.0), 16, 128, 256, thermostat=False, bending=True, k_bend=5.0,
                               const_force=(0.01, 0.0, 0.002))
    p = p.replace(n_mon_d=8, flags=Flags(**{**p.flags.__dict__, "bending_melt": bending_melt}))
    t, r0, v = configs.build_channel(p, 2.0, 8)
    rng = np.random.default_rng(2)
    r0 = r0 + rng.normal(0, 0.04, r0.shape)
    r0[:, :2] %= 16.0
    return p, t, r0, v


def test_bonded_wall_body_force_routines(oracle_mod):
    p, t, r0, v = _chain_system(bending_melt=True)
    o = _oracle(oracle_mod, p)
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f_pair = s.get_state()["force"]
        s.fluid_wall(2)
        f1 = s.get_state()["force"]
        s.intra_molec()
        f2 = s.get_state()["force"]
        s.bending(1)
        f3 = s.get_state()["force"]
        s.bending_melt(1)
        f4 = s.get_state()["force"]
        s.constant_force(1)
        f5 = s.get_state()["force"]
        e = s.energies()
    fw, ew = o.fluid_wall(r0, t)
    _close(f1 - f_pair, fw, 1e-12)
    fb, eb, _ = o.intra_molec(r0)
    _close(f2 - f1, fb, 1e-11)
    fbe, ebe = o.bending(r0, 0)
    _close(f3 - f2, fbe, 1e-10)
    fbm, ebm = o.bending(r0, 1)
    _close(f4 - f3, fbm, 1e-10)
    assert np.abs(f5 - (f4 + o.constant_force(t))).max() <= 4e-16 * np.abs(f4).max()
    assert e["v_fluid_wall"] == pytest.approx(ew, rel=1e-12)
    assert e["v_intra_molec"] == pytest.approx(eb, rel=1e-12)
    assert e["v_bend"] == pytest.approx(ebe, rel=1e-11)
    assert e["v_bend_melt"] == pytest.approx(ebm, rel=1e-11)


def test_integrator_routines(oracle_mod):
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    o = _oracle(oracle_mod, p)
    rng = np.random.default_rng(9)
    a = rng.normal(0, 3.0, r0.shape)
    heads = np.arange(p.n_chain) * p.n_mon
    a[heads] = 0.0
    with Simulation(p) as s:
        s.set_state(r0, v, t, a)
        s.verlet_positions()
        st = s.get_state()
        s.fluid_fluid()
        f = s.get_state()["force"]
        s.verlet_velocities()
        st2 = s.get_state()
        ke = s.energies()["t_fluid"]
        fh = s.f_on_heads()
    r1, v1 = o.verlet_positions(r0, v, a)
    assert np.array_equal(st["r0"], r1)                     # same IEEE ops, no contraction
    assert np.array_equal(st["v"], v1)
    v2, a2, ke_o, fh_o = o.verlet_velocities(f, st["v"], t)
    assert np.array_equal(st2["v"], v2) and np.array_equal(st2["a"], a2)
    assert ke == pytest.approx(ke_o, rel=1e-13)
    assert np.array_equal(fh, fh_o)


@pytest.mark.parametrize("thermostat", [False, True])
@pytest.mark.parametrize("n_steps", [1, 10])
def test_fused_steps_vs_oracle(oracle_mod, thermostat, n_steps):
    p, t, r0, v = _chain_system(bending_melt=False)
    p = p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": thermostat}), r_2_min_init=0.0,
                  r_2_min_time=0.0)
    o = _oracle(oracle_mod, p)
    a0 = np.zeros_like(r0)
    ro, vo, ao, fo, eo, fho = o.md_steps(r0, v, a0, t, 5, n_steps, noise=1)
    for graph in (False, True):
        with Simulation(p, t_idx0=5) as s:
            s.set_state(r0, v, t, a0)
            s.md_steps(n_steps, graph=graph)
            st = s.get_state()
            e = s.energies()
            assert s.get_time() == 5 + n_steps
        tol = 1e-12 if n_steps == 1 else 1e-9
        _close(st["r0"], ro, tol)
        _close(st["v"], vo, 1e3 * tol)
        _close(st["force"], fo, 1e3 * tol)
        for k in ("v_fluid_fluid", "v_intra_molec", "v_fluid_wall", "v_bend", "t_fluid"):
            assert e[k] == pytest.approx(eo[k], rel=1e3 * tol, abs=1e-9), k


def test_routines_equal_fused_step(oracle_mod):
    p, t, r0, v = _chain_system(bending_melt=False)
    with Simulation(p) as s1, Simulation(p) as s2:
        s1.set_state(r0, v, t)
        s2.set_state(r0, v, t)
        for _ in range(3):
            s1.step_by_routines()
        s2.md_steps(3)
        a, b = s1.get_state(), s2.get_state()
        for k in ("r0", "v", "a"):
            _close(a[k], b[k], 1e-12)


def test_determinism_and_graph_replay():
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.4)
    outs = []
    for graph in (False, True, True):
        with Simulation(p) as s:
            s.set_state(r0, v, t)
            s.md_steps(25, graph=graph)
            outs.append(s.get_state())
    for o in outs[1:]:
        for k in ("r0", "v", "a", "force"):
            assert np.array_equal(o[k], outs[0][k]), k


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_neighbor_list_equals_cell_scan(precision):
    """The Verlet list keeps the cell-scan order and arithmetic, so list on,
    list off (cells scanned every step) and an overflowing list (beads past
    the capacity scan their cells) give the same bits over many steps with
    several rebuilds."""
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.3)
    outs = []
    for cap in (-1, 0, 24):
        with Simulation(p, precision=precision) as s:
            s.set_neighbor_list(cap)
            s.set_state(r0, v, t)
            s.md_steps(30, graph=True)
            outs.append((s.get_state(), s.energies(), s.flags(), s.neighbor_list()))
    assert outs[0][3] > 24 and outs[1][3] == 0 and outs[2][3] == 24
    assert outs[0][2]["rebuilds"] > 1 and outs[0][2]["list_overflow"] == 0
    assert outs[2][2]["list_overflow"] > 0
    for st, e, fl, _ in outs[1:]:
        for k in ("r0", "v", "a", "force"):
            assert np.array_equal(st[k], outs[0][0][k]), k
        assert e["v_fluid_fluid"] == outs[0][1]["v_fluid_fluid"]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_check_skin_reports_f_skin(precision):
    """check_skin after verlet_positions is f_skin of check_skin.f90:11-24:
    1 on the steps whose skin test fired (a bead more than skin/2 from its
    position at the last rebuild), else 0, and each such step rebuilt the
    cells once.  One solvent bead drifts 0.15 sigma a step while every other
    bead rests (no forces are evaluated, so nothing accelerates), so with
    skin 0.4 the test fires every second step."""
    p, t, r0, v = configs.channel_c1(thermostat=False)
    p = p.replace(skin=0.4)
    k = p.n_part - 1                                  # a solvent bead
    r0 = r0.copy()
    r0[k, 0] = 3.0                                    # far from the periodic wrap in x
    v = np.zeros_like(v)
    v[k, 0] = 0.15 / p.dt
    expect, d = [], 0.0
    for _ in range(6):
        d += 0.15
        fired = d * d > (0.5 * p.skin) ** 2
        expect.append(int(fired))
        d = 0.0 if fired else d
    assert expect == [0, 1, 0, 1, 0, 1]
    got, rebuilds = [], []
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        for _ in range(6):
            s.verlet_positions()
            got.append(s.check_skin())
            rebuilds.append(s.flags()["rebuilds"])
        assert abs(s.get_state()["r0"][k, 0] - (3.0 + 6 * 0.15)) < 1e-12
    assert got == expect
    # the first call also ran the rebuild pending after set_state
    assert list(np.diff(rebuilds)) == expect[1:]


def _list_count(p, t, r0, rel):
    """Host brute force: pairs (i, j != i) closer than the type pair's list
    radius (interaction range, or the DPD range when the DPD thermostat acts
    beyond it, plus the skin), the squared radius scaled by 1 + rel."""
    r = np.sqrt(np.maximum(np.asarray(p.range_2, float), 0.0)).reshape(4, 4)
    if p.flags.thermostat and p.flags.dpd_cut_off:
        r = np.maximum(r, math.sqrt(p.r_cut_dpd_2))
    rl2 = np.where(r > 0, (r + p.skin) ** 2, -1.0) * (1.0 + rel)
    L = np.asarray(p.boundary, float)
    ti = np.asarray(t) - 1
    n = 0
    for a in range(0, len(r0), 512):
        d = r0[a:a + 512, None, :] - r0[None, :, :]
        d[..., :2] -= L[:2] * np.rint(d[..., :2] / L[:2])
        d2 = (d * d).sum(-1)
        ok = d2 < rl2[ti[a:a + 512, None], ti[None, :]]
        ok[np.arange(ok.shape[0]), a + np.arange(ok.shape[0])] = False
        n += int(ok.sum())
    return n


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_list_stats_vs_host_count(precision):
    """The Verlet list of every bead holds its partners within the type
    pair's list radius: the entry total lies between host brute-force counts
    on the build's positions at the radius and at the radius the build
    widens it to (FP32: 2^-12 of the squared radius, FP64: 2^-40), the waves
    walk at least 1/64 of the entries and at most the entries plus three per
    wave, and no list overflows (C1)."""
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.3)
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        st = s.list_stats()
    widen = 2.0 ** -12 if precision == "mixed" else 2.0 ** -40
    lo = _list_count(p, t, r0, -1e-12 if precision == "fp64" else -1e-5)
    hi = _list_count(p, t, r0, widen + (1e-12 if precision == "fp64" else 1e-5))
    print(precision, st, lo, hi)
    assert lo <= st["entries"] <= hi
    n_waves = (len(t) + 63) // 64
    assert st["entries"] <= 64 * st["wave_walk"] and st["wave_walk"] <= st["entries"] + 3 * n_waves
    assert st["overflowed"] == 0
    with Simulation(p, precision=precision) as s:        # no list: the statistics are refused
        s.set_neighbor_list(0)
        s.set_state(r0, v, t)
        with pytest.raises(RuntimeError, match="no Verlet list"):
            s.list_stats()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("rc_dpd", [1.0, None])
def test_dpd_queue_depths_agree(oracle_mod, precision, rc_dpd):
    """The DPD part of a pair runs from the per-lane queue or, once the queue
    is full, in place.  Depths 8 (default) and 0 (all in place) differ only
    by the FP64 summation order of the two parts; with a 2-deep queue (both
    routes taken) the list and the cell scan agree bit for bit over steps
    with rebuilds; at depth 0 the FP64 path matches the oracle."""
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=rc_dpd)
    p = p.replace(skin=0.3)
    fs = []
    for depth in (8, 0):
        with Simulation(p, precision=precision, t_idx0=17) as s:
            s.set_dpd_queue(depth)
            s.set_state(r0, v, t)
            s.fluid_fluid()
            fs.append(s.get_state()["force"])
    _close(fs[1], fs[0], 1e-13)
    if precision == "fp64":
        fo, _ = _oracle(oracle_mod, p).fluid_fluid(r0, v, t, step=17, noise=1)
        _close(fs[1], fo)
    outs = []
    for cap in (-1, 0):
        with Simulation(p, precision=precision) as s:
            s.set_dpd_queue(2)
            s.set_neighbor_list(cap)
            s.set_state(r0, v, t)
            s.md_steps(20, graph=True)
            outs.append((s.get_state(), s.flags()))
    assert outs[0][1]["rebuilds"] > 1
    for k in ("r0", "v", "a", "force"):
        assert np.array_equal(outs[1][0][k], outs[0][0][k]), k


def test_batched_replicas_equal_single_runs():
    ps = []
    states = []
    for k, (n_side, eps_bs) in enumerate([(6, 1.0), (9, 0.7), (7, 1.3)]):   # >= 3 cells across at rc 2.5
        p, t, r0, v = configs.channel_scaled(n_side, n_mon=8, lz=14.0, eps_bs=eps_bs, seed=100 + k, skin=0.3)
        ps.append(p)
        states.append((t, r0, v))
    with Simulation(ps) as sb:
        for i, (t, r0, v) in enumerate(states):
            sb.set_state(r0, v, t, isys=i)
        sb.md_steps(20, graph=True)
        batched = [sb.get_state(i) for i in range(3)]
        eb = [sb.energies(i) for i in range(3)]
    for i, (t, r0, v) in enumerate(states):
        with Simulation(ps[i]) as s:
            s.set_state(r0, v, t)
            s.md_steps(20)
            single = s.get_state()
            es = s.energies()
        for k in ("r0", "v", "a"):
            assert np.array_equal(batched[i][k], single[k]), (i, k)
        assert eb[i]["v_fluid_fluid"] == es["v_fluid_fluid"]


@pytest.mark.parametrize("graph", [False, True])
def test_md_steps_calls_compose(graph):
    """md_steps(5) then md_steps(3) gives the bits of md_steps(8), forces
    included: the fused step writes the total force back on the last step of
    each call only, and the state read between the calls is that of step 5."""
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.md_steps(5, graph=graph)
        mid = s.get_state()
        s.md_steps(3, graph=graph)
        two = s.get_state()
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.md_steps(5, graph=graph)
        five = s.get_state()
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.md_steps(8, graph=graph)
        eight = s.get_state()
    for k in ("r0", "v", "a", "force"):
        assert np.array_equal(mid[k], five[k]), ("5", k)
        assert np.array_equal(two[k], eight[k]), ("5+3", k)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_bending_and_plain_channels_batched_equal_single_runs(precision):
    """A batch of a DPD-thermostatted brush channel, a bending channel without
    a thermostat and a Langevin channel.  The batch runs the DPD pair kernel
    and the fused bonded kernel with the bending code; alone, the plain
    channel runs the variant without it and the others the kernel without
    DPD.  Every channel gives the same bits batched and alone, forces
    included."""
    p0, t0, r00, v0 = configs.channel_scaled(7, n_mon=8, lz=14.0, seed=200, skin=0.3)
    p1, t1, r01, v1 = _chain_system(bending_melt=True)
    p2, t2, r02, v2 = _langevin_c1()
    ps, states = [p0, p1, p2], [(t0, r00, v0), (t1, r01, v1), (t2, r02, v2)]
    with Simulation(ps, precision=precision) as sb:
        for i, (t, r0, v) in enumerate(states):
            sb.set_state(r0, v, t, isys=i)
        sb.md_steps(12, graph=True)
        batched = [sb.get_state(i) for i in range(3)]
        eb = [sb.energies(i) for i in range(3)]
    for i, (t, r0, v) in enumerate(states):
        with Simulation(ps[i], precision=precision) as s:
            s.set_state(r0, v, t)
            s.md_steps(12, graph=True)
            single = s.get_state()
            es = s.energies()
        for k in ("r0", "v", "a", "force"):
            assert np.array_equal(batched[i][k], single[k]), (i, k)
        for k in ("v_fluid_fluid", "v_fluid_wall", "v_intra_molec", "v_bend", "v_bend_melt", "t_fluid"):
            assert eb[i][k] == es[k], (i, k, eb[i][k], es[k])


def test_mixed_precision_force_error(oracle_mod):
    p, t, r0, v = _perturbed_c1(thermostat=False, rc_dpd=None)
    p = p.replace(skin=0.3)
    with Simulation(p, precision="mixed") as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f = s.get_state()["force"]
        e = s.energies()["v_fluid_fluid"]
        npair = s.pair_count()
    fo, eo = _oracle(oracle_mod, p).fluid_fluid(r0, v, t, noise=0)
    cnt, _ = _oracle(oracle_mod, p).pair_stats(r0, t)
    assert npair == cnt.sum() // 2                      # FP64-decided pair set
    err = np.linalg.norm(f - fo, axis=1).max() / np.linalg.norm(fo, axis=1).max()
    assert err < 1e-5, err
    rms = np.sqrt(((f - fo) ** 2).sum(1).mean() / (fo ** 2).sum(1).mean())
    assert rms < 2e-6, rms
    assert e == pytest.approx(eo, rel=1e-5)


def _mean_kt(p, t, r0, v, precision, equil=400, blocks=20, per=20):
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t)
        s.md_steps(equil, graph=True)
        kes = []
        for _ in range(blocks):
            s.md_steps(per, graph=True)
            kes.append(s.energies()["t_fluid"])
        fl = s.flags()
    assert fl["nonfinite"] == 0
    return 2.0 * np.mean(kes) / (3 * (p.n_part - p.n_chain))


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_dpd_thermostat_holds_kt(precision):
    """Fluctuation-dissipation: with DPD on every LJ pair (the reference's
    default, no DPD_CUT_OFF) the channel holds kT = 1 within 5 %.  (The
    oracle, run the same way on the CPU, gives 1.015.)"""
    p, t, r0, v = configs.channel_c1(thermostat=True, rc_dpd=None)
    p = p.replace(skin=0.4)
    kT = _mean_kt(p, t, r0, v, precision)
    assert abs(kT - 1.0) < 0.05, kT


def test_mixed_precision_dpd_statistics():
    """DPD_CUT_OFF 1.0 with heads excluded thermostats few pairs, and this
    start keeps heating (the oracle reaches kT ~ 1.8 after 800 steps): the
    mixed path must follow the FP64 path (which matches the oracle step by
    step) statistically, within 5 % in mean kT."""
    p, t, r0, v = configs.channel_c1(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.4)
    k32 = _mean_kt(p, t, r0, v, "mixed")
    k64 = _mean_kt(p, t, r0, v, "fp64")
    assert abs(k32 / k64 - 1.0) < 0.05, (k32, k64)


def test_pair_set_scaled_c2(oracle_mod):
    """C2 geometry (n_side 30, ~86k beads): per-bead pair-set fingerprint."""
    p, t, r0, v = configs.channel_scaled(30)
    rng = np.random.default_rng(1)
    r0 = r0 + rng.normal(0, 0.05, r0.shape)
    r0[:, :2] %= p.boundary[0]
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        cnt, h = s.pair_stats()
        s.fluid_fluid()
        npair = s.pair_count()
    co, ho = _oracle(oracle_mod, p).pair_stats_cells(r0, t)
    assert np.array_equal(cnt, co) and np.array_equal(h, ho)
    assert npair == co.sum() // 2


def test_full_c2_pair_count_and_momentum(oracle_mod):
    """Full C2 (998,784 beads): the kernel's interacting-pair count equals the
    oracle's linked-cell count, and the pair forces sum to ~0.

    Both paths: each pair's two evaluations are exact negatives (in FP32 the
    separation is the difference of the two cell-frame coordinates minus the
    cell offset), and every pair force is added in FP64, so only the per-bead
    FP64 sums round."""
    p, t, r0, v = configs.channel_c2(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.4)
    co, _ = oracle_mod.Oracle(p).pair_stats_cells(r0, t)
    sums = {}
    for prec in ("fp64", "mixed"):
        with Simulation(p, precision=prec) as s:
            s.set_state(r0, v, t)
            s.fluid_fluid()
            npair = s.pair_count()
            f = s.get_state()["force"]
        assert npair == co.sum() // 2, prec
        sums[prec] = (np.abs(f.sum(0)).max(), np.abs(f).max())
    print("C2 |sum F|, max|F|:", sums)
    n = math.sqrt(p.n_part)
    assert sums["fp64"][0] < 1e-11 * sums["fp64"][1] * n, sums
    assert sums["mixed"][0] < 1e-11 * sums["mixed"][1] * n, sums


def test_full_c2_fused_steps_vs_oracle(oracle_mod):
    """Full C2 (998,784 beads): 16 fused FP64 steps, graph-replayed, against
    the oracle's linked-cell steps.  At this size the grid-stride rebuild
    kernels give blocks one or two tiles each, and the skin test triggers
    device-gated rebuilds after the initial one."""
    p, t, r0, v = configs.channel_c2(thermostat=True, rc_dpd=1.0)
    p = p.replace(skin=0.4)
    a0 = np.zeros_like(r0)
    n = 16
    ro, vo, ao, fo, eo, fho = oracle_mod.Oracle(p, omp=True).md_steps(r0, v, a0, t, 0, n, noise=1)
    with Simulation(p) as s:
        s.set_state(r0, v, t, a0)
        s.md_steps(n, graph=True)
        st = s.get_state()
        e = s.energies()
        rebuilds = s.flags()["rebuilds"]
    print("C2 rebuilds in %d steps (with the initial one): %d" % (n, rebuilds))
    assert rebuilds >= 2
    _close(st["r0"], ro, 1e-9)
    _close(st["v"], vo, 1e-6)
    _close(st["force"], fo, 1e-6)
    for k in ("v_fluid_fluid", "v_intra_molec", "v_fluid_wall", "t_fluid"):
        assert e[k] == pytest.approx(eo[k], rel=1e-6, abs=1e-9), k


def test_profiles_vs_oracle(oracle_mod):
    p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.profiles_reset(200)
        s.profiles_accumulate()
        s.md_steps(3)
        s.profiles_accumulate()
        st = s.get_state()
        nobs, dens, vp, vp2 = s.profiles()
    o = _oracle(oracle_mod, p)
    d0 = o.histo(r0, 200) + o.histo(st["r0"], 200)
    vp0a, vp20a = o.velocity_prof(r0, v, 200)
    vp0b, vp20b = o.velocity_prof(st["r0"], st["v"], 200)
    assert nobs == 2 and np.array_equal(dens, d0)
    assert np.abs(vp - (vp0a + vp0b)).max() < 1e-8
    assert np.abs(vp2 - (vp20a + vp20b)).max() < 1e-6


def test_error_paths():
    p, t, r0, v = configs.channel_c1()
    bad = p.replace(boundary=(4.0, 4.0, 20.0))              # < 3 cells in x,y
    with pytest.raises(Exception, match="too small"):
        Simulation(bad)
    with Simulation(p) as s:
        with pytest.raises(Exception, match="no state"):
            s.fluid_fluid()
        t2 = t.copy(); t2[0] = 7
        with pytest.raises(Exception, match="a_type"):
            s.set_state(r0, v, t2)
        s.set_state(r0, v, t)
        with pytest.raises(Exception, match="inter_type"):
            s.fluid_wall(1)


@pytest.mark.parametrize("cosine", [True, False])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_melt_channel_vs_oracle(oracle_mod, precision, cosine):
    """Melt without a brush (n_chain = 0): FENE + bending_melt (cosine and
    harmonic forms) + Poiseuille body force + DPD, 10 fused steps against
    the oracle."""
    p, t, r0, v = configs.channel_melt(n_cols=8, n_z=32, skin=0.3, bend_cosine=cosine)
    o = _oracle(oracle_mod, p)
    a0 = np.zeros_like(r0)
    ro, vo, ao, fo, eo, _ = o.md_steps(r0, v, a0, t, 0, 10, noise=1)
    with Simulation(p, precision=precision) as s:
        s.set_state(r0, v, t, a0)
        s.md_steps(10, graph=True)
        st = s.get_state()
        e = s.energies()
    tol = 1e-9 if precision == "fp64" else 1e-4
    _close(st["r0"], ro, tol)
    _close(st["v"], vo, 1e3 * tol)
    for k in ("v_fluid_fluid", "v_intra_molec", "v_bend_melt", "t_fluid"):
        assert e[k] == pytest.approx(eo[k], rel=1e3 * tol), k


def test_nve_drift_bounded():
    """Thermostat off, 2000 steps of C1: total energy drift stays below
    5e-4 of |E0| on both paths and the mixed path drifts like FP64."""
    from mfa_brush_amd.nve import drift
    p, t, r0, v = configs.channel_c1(thermostat=False, rc_dpd=None, skin=0.4)
    p = p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": False}), friction=0.0, sig=0.0)
    r64 = drift(p, t, r0, v, "fp64", 2000, 100)
    r32 = drift(p, t, r0, v, "mixed", 2000, 100)
    assert r64["max_abs_rel"] < 5e-4 and r32["max_abs_rel"] < 5e-4, (r64["max_abs_rel"], r32["max_abs_rel"])
    assert abs(r32["slope_per_1000_steps"] - r64["slope_per_1000_steps"]) < 2e-5


def test_nve_drift_full_c2():
    """Thermostat off, 1e4 steps of the full 998,784-bead brush channel
    (the NVE check of that workload): the total energy stays within 5e-4 of
    |E0| on both paths, and the mixed path's drift slope matches FP64's."""
    from mfa_brush_amd.nve import drift
    p, t, r0, v = configs.channel_c2(thermostat=False, rc_dpd=None, skin=0.4)
    p = p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": False}), friction=0.0, sig=0.0)
    r64 = drift(p, t, r0, v, "fp64", 10000, 500)
    r32 = drift(p, t, r0, v, "mixed", 10000, 500)
    print("C2 NVE: fp64 max %.2e slope %.2e; mixed max %.2e slope %.2e per 1000 steps"
          % (r64["max_abs_rel"], r64["slope_per_1000_steps"], r32["max_abs_rel"], r32["slope_per_1000_steps"]))
    assert r64["flags"]["nonfinite"] == 0 and r32["flags"]["nonfinite"] == 0
    assert r64["max_abs_rel"] < 5e-4 and r32["max_abs_rel"] < 5e-4, (r64["max_abs_rel"], r32["max_abs_rel"])
    assert abs(r32["slope_per_1000_steps"] - r64["slope_per_1000_steps"]) < 1e-5


@pytest.mark.parametrize("system", ["brush_solvent", "brush_melt"])
def test_chain_observables_vs_oracle(system):
    """Radius of gyration per axis of the brush and melt chains (unfolded
    across the periodic images), anisotropy and total momentum from the
    device state match the oracle's restatement of observation.f90:182-290
    on the same downloaded state, after steps that move chains across the
    periodic boundaries."""
    from oracle.oracle import chain_observables
    if system == "brush_solvent":
        p, t, r0, v = _perturbed_c1(thermostat=True, rc_dpd=1.0)
    else:
        p, t, r0, v = _chain_system()
    with Simulation(p) as s:
        s.set_state(r0, v, t)
        s.md_steps(30, graph=True)
        st = s.get_state()
        got = s.chain_observables()
    gb, gm, an, mom = chain_observables(p, st["r0"], st["v"], t)
    assert np.allclose(got["rg2_brush"], gb, rtol=1e-12, atol=0)
    assert np.allclose(got["rg2_melt"], gm, rtol=1e-12, atol=1e-300)
    assert np.allclose(got["anisotropy"], an, rtol=1e-11, atol=1e-13, equal_nan=True)
    scale = np.abs(np.asarray(p.mass_type)[t - 1][:, None] * st["v"]).sum()
    assert np.all(np.abs(got["momentum"] - mom) <= 1e-13 * scale)


def _langevin_c1():
    p, t, r0, v = _perturbed_c1(thermostat=False, rc_dpd=None)
    return p.replace(flags=Flags(**{**p.flags.__dict__, "langevin": True})), t, r0, v


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_langevin_forces_vs_oracle(oracle_mod, precision):
    """Langevin bead forces (lgv_forces.f90:14-25) with the same Philox
    noise on the device and in the oracle: FP64 within the FP64 bar, mixed
    within 1e-5 of the largest force."""
    p, t, r0, v = _langevin_c1()
    with Simulation(p, precision=precision, t_idx0=17) as s:
        s.set_state(r0, v, t)
        s.fluid_fluid()
        f = s.get_state()["force"]
    fo, _ = _oracle(oracle_mod, p).fluid_fluid(r0, v, t, step=17, noise=1)
    if precision == "fp64":
        _close(f, fo)
    else:
        err = np.linalg.norm(f - fo, axis=1).max() / np.linalg.norm(fo, axis=1).max()
        assert err < 1e-5, err


def test_langevin_thermostat_holds_kt():
    """The Langevin thermostat (friction gamma, sig = sqrt(2 kT gamma / dt))
    holds kT = 1 within 5 % on the channel."""
    p, t, r0, v = _langevin_c1()
    kt = _mean_kt(p, t, r0, v, "mixed")
    assert abs(kt - 1.0) < 0.05, kt
