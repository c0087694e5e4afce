"""The CPU oracle against the reference's own routines (oracle/_ref: the
reference sources compiled with flang, default control_simulation.h, called
through oracle/ref_harness.f90) and against golden vectors those routines
wrote (tests/golden/ref_vectors.npz, tests/golden/make_ref_vectors.py).

The reference draws its DPD noise from the ziggurat stream, the oracle from
Philox, so the comparisons use the noise-free parts: sig = 0 (the random term
is then exactly 0 in both) with the dissipative term on or off."""
import os

import numpy as np
import pytest

from mfa_brush_amd import configs
from mfa_brush_amd.params import Flags

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden", "ref_vectors.npz")


def ref_system(n_side=None):
    """Default-header physics: SOLVENT 3 style cutoffs, DPD cutoff = LJ
    cutoff, BENDING, POISEUILLE, FORCE_SWITCH_ON with a zero ramp."""
    if n_side is None:
        p, t, r0, v = configs.channel_c1(thermostat=True, rc_dpd=None, bending=True, k_bend=5.0,
                                         const_force=(0.01, 0.0, 0.002))
    else:
        p, t, r0, v = configs.channel_scaled(n_side, n_mon=12, lz=16.0, thermostat=True, rc_dpd=None,
                                             bending=True, k_bend=5.0, const_force=(0.01, 0.0, 0.002))
    p = p.replace(flags=Flags(solvent=3, thermostat=True, dpd_cut_off=None, force_switch_on=True, bending=True,
                              poiseuille=True), r_2_min_init=0.0, r_2_min_time=0.0, sig=0.0)
    rng = np.random.default_rng(3)
    r0 = r0 + rng.normal(0, 0.05, r0.shape)
    r0[:, :2] %= p.boundary[0]
    return p, t, r0, v


@pytest.fixture(scope="module")
def ref_mod(oracle_mod):
    if not oracle_mod.ref_available():
        pytest.skip("oracle/_ref not built (needs the reference checkout)")
    return oracle_mod


def _no_dpd(p):
    return p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": False}), friction=0.0)


@pytest.mark.parametrize("dissipative", [False, True])
def test_fluid_fluid_bit_exact(ref_mod, dissipative):
    p, t, r0, v = ref_system()
    if not dissipative:
        p = _no_dpd(p)
    ref = ref_mod.RefLib(p)
    ref.set_state(r0, v, np.zeros_like(r0), t)
    fr, er = ref.call(ref.FLUID_FLUID)
    fo, eo = ref_mod.Oracle(p).fluid_fluid(r0, v, t, noise=0)
    assert np.array_equal(fr, fo) and er == eo


def test_bonded_wall_body_bit_exact(ref_mod):
    p, t, r0, v = ref_system()
    ref = ref_mod.RefLib(p)
    ref.set_state(r0, v, np.zeros_like(r0), t)
    o = ref_mod.Oracle(p)
    fr, er = ref.call(ref.INTRA_MOLEC)
    fo, eo, _ = o.intra_molec(r0)
    assert np.array_equal(fr, fo) and er == eo
    fr, er = ref.call(ref.BENDING)
    fo, eo = o.bending(r0, 0)
    assert np.array_equal(fr, fo) and er == eo
    fr, _ = ref.call(ref.CONSTANT_FORCE)
    assert np.array_equal(fr, o.constant_force(t))
    # 9-3 wall: x**9, x**10 may be evaluated in another multiplication order
    fr, er = ref.call(ref.FLUID_WALL)
    fo, eo = o.fluid_wall(r0, t)
    assert np.abs(fr - fo).max() <= 4 * np.finfo(float).eps * np.abs(fr).max()
    assert eo == pytest.approx(er, rel=1e-14)
    _, er = ref.call(ref.CHECK_FLUID_FLUID)
    assert o.check_fluid_fluid(r0, t) == er


def test_md_steps_match_reference(ref_mod):
    p, t, r0, v = ref_system()
    ref = ref_mod.RefLib(p)
    ref.set_state(r0, v, np.zeros_like(r0), t)
    er = ref.md_steps(20, 3)
    st = ref.get_state()
    ro, vo, ao, fo, eo, _ = ref_mod.Oracle(p).md_steps(r0, v, np.zeros_like(r0), t, 3, 20, noise=0)
    assert np.abs(st["r0"] - ro).max() < 1e-12
    assert np.abs(st["v"] - vo).max() < 1e-11
    assert np.abs(st["a"] - ao).max() < 1e-9
    # the force array after the step, heads zeroed by verlet_velocities
    assert np.abs(st["force"] - fo).max() <= 1e-10 * np.abs(fo).max()
    heads = np.arange(p.n_chain) * p.n_mon
    assert not st["force"][heads].any() and not fo[heads].any()
    for k in ("v_fluid_fluid", "v_intra_molec", "v_fluid_wall", "v_bend", "t_fluid"):
        assert eo[k] == pytest.approx(er[k], rel=1e-12), k


def test_golden_vectors_from_reference(oracle_mod):
    """Vectors the reference's routines wrote (committed), checked without
    the reference build."""
    z = np.load(GOLD)
    p, t, r0, v = ref_system(n_side=int(z["n_side"]))
    assert np.array_equal(r0, z["r0"]) and np.array_equal(t, z["a_type"])
    o = oracle_mod.Oracle(p)
    f, e = o.fluid_fluid(r0, v, t, noise=0)
    assert np.array_equal(f, z["f_fluid_fluid"]) and e == float(z["e_fluid_fluid"])
    f, e, _ = o.intra_molec(r0)
    assert np.array_equal(f, z["f_intra_molec"]) and e == float(z["e_intra_molec"])
    f, e = o.bending(r0, 0)
    assert np.array_equal(f, z["f_bending"])
    f, e = o.fluid_wall(r0, t)
    assert np.abs(f - z["f_fluid_wall"]).max() <= 4 * np.finfo(float).eps * np.abs(f).max()
    ro, vo, *_ = o.md_steps(r0, v, np.zeros_like(r0), t, 3, 10, noise=0)
    assert np.abs(ro - z["r0_10"]).max() < 1e-12 and np.abs(vo - z["v_10"]).max() < 1e-11
