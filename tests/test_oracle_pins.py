"""Pins of the oracle branches that the reference build in oracle/_ref cannot
reach (its default control_simulation.h undefines DPD_CUT_OFF and has no
channel-mode virial) and of the melt-chain bending range.  Each is checked
against what the mathematics fixes: closed forms on two beads, the
virial's scaling and shear derivatives of the LJ energy, and the identity
sum_pairs r_ij (x) F_ij = sum_i r_i (x) F_i for a cluster that no pair
crosses the periodic boundary of.  A plausible slip (a dropped w factor, a
DPD term on a head pair, a stale 1/r, the melt range shifted by one chain,
a transposed virial component) fails one of them."""
import math

import numpy as np
import pytest

from mfa_brush_amd import configs
from mfa_brush_amd.params import Flags


def _two(p, dr, ti, tj, v=None):
    r0 = np.array([[6.0, 6.0, 8.0], [6.0, 6.0, 8.0]])
    r0[1] += np.asarray(dr, float)
    if v is None:
        v = np.zeros((2, 3))
    return r0, np.array([ti, tj], np.int32), np.asarray(v, float)


def _pair_params(rc_dpd=1.0, **kw):
    p = configs.channel_params((12.0, 12.0, 16.0), 1, 0, 2, rc_dpd=rc_dpd, **kw)
    return p


def _dpd_part(o, r0, t, v, step, noise):
    f_lj, _ = o.fluid_fluid_brute(r0, np.zeros_like(v), t, noise=0)
    f, _ = o.fluid_fluid_brute(r0, v, t, step=step, noise=noise)
    return f - f_lj, f_lj


# ------------------------------------------- DPD_CUT_OFF (rc_dpd = 1) ----
@pytest.mark.parametrize("ti,tj", [(3, 3), (2, 3), (2, 2)])
def test_dpd_cutoff_dissipative_closed_form(oracle_mod, ti, tj):
    """r = 0.8 < rc_dpd = 1 (fluid_fluid.f90:217-262, dpd_forces.f90:47-52):
    F_D on bead 0 = -gamma w^2 (dv . rhat) rhat with w = 1 - r/rc_dpd, and
    -F_D on bead 1; the LJ part is unchanged by the velocities."""
    p = _pair_params().replace(sig=0.0)                # dissipative part only
    o = oracle_mod.Oracle(p)
    d = np.array([0.48, -0.64, 0.0])                  # |d| = 0.8, off-axis
    r0, t, v = _two(p, d, ti, tj, v=[[0.3, -0.2, 0.1], [-0.5, 0.4, 0.25]])
    fd, _ = _dpd_part(o, r0, t, v, 5, 1)
    r = 0.8
    rhat = (r0[0] - r0[1]) / r
    w = 1.0 - r / 1.0
    want = -p.friction * w * w * (v[0] - v[1]).dot(rhat) * rhat
    assert np.allclose(fd[0], want, rtol=1e-12, atol=1e-15)
    assert np.array_equal(fd[1], -fd[0])


def test_dpd_cutoff_random_closed_form(oracle_mod):
    """Random part at r = 0.8: sig (1 - r/rc_dpd) g rhat, g the Philox normal
    of the unordered pair at this step; exactly antisymmetric."""
    lib = oracle_mod.load()
    p = _pair_params(gamma=0.0).replace(sig=2.5)
    o = oracle_mod.Oracle(p)
    r0, t, v = _two(p, [0.0, 0.0, -0.8], 3, 3)
    fr, _ = _dpd_part(o, r0, t, v, 23, 1)
    g = lib.orc_pair_gauss(p.seed, 23, 0, 1)
    want = 2.5 * (1.0 - 0.8) * g * np.array([0.0, 0.0, 1.0])
    assert np.allclose(fr[0], want, rtol=1e-12, atol=1e-15)
    assert np.array_equal(fr[1], -fr[0])
    assert abs(g) > 1e-3                               # the draw is not degenerate


@pytest.mark.parametrize("ti,tj", [(1, 3), (1, 2), (3, 1)])
def test_dpd_cutoff_excludes_heads(oracle_mod, ti, tj):
    """A pair with a grafted head (type 1) takes no DPD term under
    DPD_CUT_OFF (fluid_fluid.f90:223-236), only its LJ force."""
    p = _pair_params()
    o = oracle_mod.Oracle(p)
    r0, t, v = _two(p, [0.8, 0.0, 0.0], ti, tj, v=[[1.0, 0.5, 0.0], [-1.0, 0.0, 0.3]])
    fd, f_lj = _dpd_part(o, r0, t, v, 9, 1)
    assert not fd.any()
    assert np.abs(f_lj).max() > 1.0                    # the LJ part is there


@pytest.mark.parametrize("r", [1.05, 1.6, 2.4])
def test_dpd_cutoff_lj_only_beyond_rc_dpd(oracle_mod, r):
    """1 < r < 2.5: LJ only (the DPD range is the DPD cutoff, not the LJ one)."""
    p = _pair_params()
    o = oracle_mod.Oracle(p)
    r0, t, v = _two(p, [r, 0.0, 0.0], 3, 3, v=[[1.0, 0.5, 0.0], [-1.0, 0.0, 0.3]])
    fd, f_lj = _dpd_part(o, r0, t, v, 9, 1)
    assert not fd.any()
    s6 = r ** -6
    assert f_lj[0, 0] == pytest.approx(-24.0 * (2 * s6 * s6 - s6) / r, rel=1e-12)


def test_dpd_beyond_lj_range_uses_own_distance(oracle_mod):
    """A pair inside rc_dpd but outside its LJ range (solvent-solvent LJ cut
    at 0.9 here) takes the DPD term with its own 1/r: the force is the closed
    form along its own rhat, and no LJ force.  (The reference would reuse
    the previous pair's 1/r, fluid_fluid.f90:253-256; the restatement
    deliberately does not.)"""
    p = _pair_params()
    r2 = np.asarray(p.range_2).reshape(4, 4).T.copy()
    r2[2, 2] = 0.81
    col = lambda m: np.ascontiguousarray(m.T.reshape(-1))
    p = p.replace(range_2=col(r2), inv_range_2=col(np.where(r2 > 0, 1 / np.where(r2 > 0, r2, 1), 0)))
    o = oracle_mod.Oracle(p)
    lib = oracle_mod.load()
    d = np.array([0.0, 0.57, 0.76])                   # r = 0.95
    r0, t, v = _two(p, d, 3, 3, v=[[0.2, 0.7, -0.4], [0.0, -0.3, 0.6]])
    f, e = o.fluid_fluid_brute(r0, v, t, step=4, noise=1)
    r = 0.95
    rhat = (r0[0] - r0[1]) / r
    w = 1.0 - r
    g = lib.orc_pair_gauss(p.seed, 4, 0, 1)
    want = (p.sig * w * g - p.friction * w * w * (v[0] - v[1]).dot(rhat)) * rhat
    assert e == 0.0
    assert np.allclose(f[0], want, rtol=1e-12, atol=1e-15)
    assert np.allclose(f[1], -want, rtol=1e-12, atol=1e-15)


# ------------------------------------------------------- bending_melt ----
def _melt_system(oracle_mod, k_bend=5.0):
    """4 brush chains of 5 and 6 melt chains of 7 beads, crumpled."""
    p = configs.channel_params((12.0, 12.0, 14.0), 5, 4, 6, thermostat=False, bending=True, k_bend=k_bend)
    p = p.replace(n_mon_d=7, flags=Flags(**{**p.flags.__dict__, "bending_melt": True}))
    rng = np.random.default_rng(11)
    r0, t = [], []
    for c in range(4):
        base = np.array([1.5 + 3.0 * c, 2.0, 1.2])
        for m in range(5):
            r0.append(base + [0.2 * rng.standard_normal(), 0.2 * rng.standard_normal(), 0.95 * m])
            t.append(1 if m == 0 else 2)
    for c in range(6):
        x = np.array([2.0 + 1.7 * c, 7.0, 7.0])
        for m in range(7):
            r0.append(x.copy())
            t.append(3)
            step = rng.normal(0, 1, 3)
            x = x + 0.95 * step / np.linalg.norm(step)
    r0 = np.array(r0)
    r0[:, :2] %= 12.0
    return p, oracle_mod.Oracle(p), r0, np.array(t, np.int32)


def test_bending_melt_force_is_minus_gradient(oracle_mod):
    p, o, r0, t = _melt_system(oracle_mod)
    f, e = o.bending(r0, 1)
    assert np.abs(f[:20]).max() == 0.0                 # brush beads untouched by the melt term
    assert np.abs(f[20:]).max() > 0.1
    h = 1e-6
    for i in range(20, len(r0)):
        for d in range(3):
            rp = r0.copy(); rp[i, d] += h
            rm = r0.copy(); rm[i, d] -= h
            g = (o.bending(rp, 1)[1] - o.bending(rm, 1)[1]) / (2 * h)
            assert -g == pytest.approx(f[i, d], rel=2e-5, abs=2e-6)


def test_bending_melt_range_and_right_angle(oracle_mod):
    """Only the melt chains bend under bending_melt, each over its own beads:
    a melt chain of three beads with a right angle carries energy
    k (pi/2)^2 / 2, the end beads get force k (pi/2) / |b| perpendicular to
    their bond, the chain's forces sum to zero, straight chains get none.
    (The reference indexes the melt chain without its offset,
    bending_melt.f90:43; the restatement uses the chain's own beads.)"""
    p = configs.channel_params((12.0, 12.0, 14.0), 3, 2, 3, thermostat=False, bending=True, k_bend=5.0)
    p = p.replace(n_mon_d=3, flags=Flags(**{**p.flags.__dict__, "bending_melt": True}))
    o = oracle_mod.Oracle(p)
    brush = [[1, 1, 2.0], [1, 1, 3.0], [2, 1, 3.0], [5, 5, 10.0], [5, 5, 9.0], [5, 6, 9.0]]
    melt = [[3, 3, 6.0], [3, 3, 7.0], [3, 3, 8.0],          # straight
            [8, 8, 6.0], [8, 8, 7.0], [9, 8, 7.0],          # right angle
            [6, 9, 4.0], [6, 9, 5.0], [6, 9, 6.0]]          # straight
    r0 = np.array(brush + melt, float)
    f, e = o.bending(r0, 1)
    assert e == pytest.approx(0.5 * 5.0 * (math.pi / 2) ** 2, rel=1e-12)
    assert not f[:6].any() and np.abs(f[6:9]).max() < 1e-12 and np.abs(f[12:15]).max() < 1e-12
    fc = f[9:12]
    assert np.abs(fc.sum(0)).max() < 1e-12
    for end, nb in ((0, 1), (2, 1)):
        b = r0[9 + end] - r0[9 + nb]
        assert abs(fc[end].dot(b)) < 1e-12
        assert np.linalg.norm(fc[end]) == pytest.approx(5.0 * math.pi / 2 / np.linalg.norm(b), rel=1e-12)
    # the brush chains with their own right angles stay with bending(…, 0)
    fb, eb = o.bending(r0, 0)
    assert eb == pytest.approx(2 * 0.5 * 5.0 * (math.pi / 2) ** 2, rel=1e-12) and not fb[6:].any()


# ---------------------------------------------------------- pair virial ----
def _lj_fluid(n, box, y_slab=None, seed=3, dmin=0.88):
    rng = np.random.default_rng(seed)
    pts = []
    lo = np.array([0.0, 0.0, 1.0])
    hi = np.array([box[0], box[1], box[2] - 1.0])
    if y_slab is not None:
        lo[1], hi[1] = y_slab
    while len(pts) < n:
        q = rng.uniform(lo, hi)
        ok = True
        for r in pts:
            d = q - r
            d[:2] -= np.round(d[:2] / np.asarray(box[:2])) * np.asarray(box[:2])
            if d.dot(d) < dmin * dmin:
                ok = False
                break
        if ok:
            pts.append(q)
    t = np.where(np.arange(n) % 3 == 0, 2, 3).astype(np.int32)
    return np.array(pts), t


def _vparams(box, n, thermostat=False, rc_dpd=1.0):
    return configs.channel_params(box, 1, 0, n, thermostat=thermostat, rc_dpd=rc_dpd)


def test_virial_two_beads_closed_form(oracle_mod):
    p = _vparams((12.0, 12.0, 16.0), 2)
    o = oracle_mod.Oracle(p)
    d = np.array([-0.7, 0.5, 0.6])
    r0, t, v = _two(p, d, 3, 3)
    w = o.pair_virial(r0, v, t, noise=0)
    r01 = -d
    r2 = r01.dot(r01)
    s6 = r2 ** -3
    fmag = 24.0 * (2 * s6 * s6 - s6) / r2              # F_0 = fmag * r01
    want = [r01[a] * r01[b] * fmag for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))]
    assert np.allclose(w, want, rtol=1e-12)


def test_virial_two_beads_across_periodic_boundary(oracle_mod):
    """The separation is the minimum image: a pair split by the x boundary
    gives the virial of the same pair in the box interior."""
    p = _vparams((12.0, 12.0, 16.0), 2)
    o = oracle_mod.Oracle(p)
    r0 = np.array([[0.3, 5.0, 8.0], [11.5, 5.4, 8.3]])
    r1 = np.array([[6.3, 5.0, 8.0], [5.5, 5.4, 8.3]])
    t = np.array([3, 3], np.int32)
    w0 = o.pair_virial(r0, np.zeros((2, 3)), t, noise=0)
    w1 = o.pair_virial(r1, np.zeros((2, 3)), t, noise=0)
    assert np.abs(w0).max() > 0.1 and np.allclose(w0, w1, rtol=1e-12, atol=1e-14)


def test_virial_is_energy_derivative_under_strain(oracle_mod):
    """W_aa = -dV/d(lambda_a) under r_a -> lambda_a r_a (x and y with the box,
    z inside the walls), W_xz = -dV/dg under x -> x + g z, W_yz likewise,
    W_xy = -dV/dg under x -> x + g y for a fluid no pair of which wraps in y."""
    box = (11.0, 11.0, 11.0)
    r0, t = _lj_fluid(260, box, y_slab=(2.6, 8.4))
    p = _vparams(box, len(r0))
    o = oracle_mod.Oracle(p)
    v = np.zeros_like(r0)
    w = o.pair_virial(r0, v, t, noise=0)
    h = 1e-6

    def energy(r, bx):
        return oracle_mod.Oracle(p.replace(boundary=tuple(bx))).check_fluid_fluid(r, t)

    def d_scale(a):
        out = []
        for s in (1 + h, 1 - h):
            r = r0.copy(); r[:, a] *= s
            bx = list(box)
            if a < 2:
                bx[a] *= s
            out.append(energy(r, bx))
        return (out[0] - out[1]) / (2 * h)

    def d_shear(a, b):
        out = []
        for g in (h, -h):
            r = r0.copy(); r[:, a] += g * r0[:, b]
            out.append(energy(r, box))
        return (out[0] - out[1]) / (2 * h)

    want = [-d_scale(0), -d_scale(1), -d_scale(2), -d_shear(0, 1), -d_shear(0, 2), -d_shear(1, 2)]
    scale = np.abs(w[:3]).max()
    assert scale > 10.0
    assert np.allclose(w, want, rtol=0, atol=1e-6 * scale), (w, want)
    assert np.abs(w[3:]).max() > 1e-3 * scale          # off-diagonals are not trivially zero here


@pytest.mark.parametrize("noise", [0, 1])
def test_virial_cluster_identity_with_dpd(oracle_mod, noise):
    """sum_pairs r_ij (x) F_ij = sum_i r_i (x) F_i for a cluster that no pair
    crosses a periodic boundary of, with the DPD dissipative (and random)
    forces included: the pair force of the virial is the one the force
    routine adds."""
    box = (14.0, 14.0, 12.0)
    r0, t = _lj_fluid(300, (8.0, 8.0, 12.0), seed=8)
    r0[:, :2] += 3.0
    p = _vparams(box, len(r0), thermostat=True, rc_dpd=1.0)
    o = oracle_mod.Oracle(p)
    v = np.random.default_rng(2).normal(0, 1, r0.shape)
    f, _ = o.fluid_fluid_brute(r0, v, t, step=7, noise=noise)
    want = [(r0[:, a] * f[:, b]).sum() for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))]
    w = o.pair_virial(r0, v, t, step=7, noise=noise)
    f_lj, _ = o.fluid_fluid_brute(r0, np.zeros_like(v), t, noise=0)
    assert np.abs(f - f_lj).max() > 0.1                # DPD forces present
    assert np.allclose(w, want, rtol=0, atol=1e-10 * np.abs(want).max())


# ------------------------------------------------ cosine bending form ----
def _cosine(p):
    return p.replace(flags=Flags(**{**p.flags.__dict__, "bend_cosine": True}))


@pytest.mark.parametrize("melt", [0, 1])
def test_cosine_bending_is_minus_gradient(oracle_mod, melt):
    p, _, r0, t = _melt_system(oracle_mod)
    o = oracle_mod.Oracle(_cosine(p))
    f, e = o.bending(r0, melt)
    lo, hi = (20, len(r0)) if melt else (0, 20)
    assert np.abs(f[lo:hi]).max() > 0.1
    mask = np.ones(len(r0), bool); mask[lo:hi] = False
    assert not f[mask].any()
    h = 1e-6
    for i in range(lo, hi):
        for d in range(3):
            rp = r0.copy(); rp[i, d] += h
            rm = r0.copy(); rm[i, d] -= h
            g = (o.bending(rp, melt)[1] - o.bending(rm, melt)[1]) / (2 * h)
            assert -g == pytest.approx(f[i, d], rel=2e-5, abs=2e-6)
    assert np.abs(f.sum(0)).max() < 1e-12 * max(1.0, np.abs(f).max())


def test_cosine_bending_closed_forms(oracle_mod):
    """k (1 - cos theta): straight chains 0, a right angle k with end forces
    k / |b| perpendicular to their bonds, a full fold 2k with zero force
    (the extremum), and the harmonic form's small-angle limit: for theta
    -> 0, k (1 - cos theta) -> k theta^2 / 2, so both forms' forces agree to
    O(theta^2)."""
    p = configs.channel_params((12.0, 12.0, 14.0), 3, 3, 1, thermostat=False, bending=True, k_bend=5.0)
    oc = oracle_mod.Oracle(_cosine(p))
    th = 0.02
    r0 = np.array([[1, 1, 2.0], [1, 1, 3.0], [1, 1, 4.0],                       # straight
                   [5, 5, 10.0], [5, 5, 9.0], [6, 5, 9.0],                       # right angle
                   [8, 8, 3.0], [8, 8, 4.0], [8, 8, 3.0 + 1e-9],                 # folded back
                   [3, 9, 5.0]])
    f, e = oc.bending(r0, 0)
    want_e = 5.0 * 1.0 + 5.0 * 2.0
    assert e == pytest.approx(want_e, rel=1e-9)
    assert np.abs(f[0:3]).max() < 1e-12 and np.abs(f[6:9]).max() < 1e-6
    for end, nb in ((3, 4), (5, 4)):
        b = r0[end] - r0[nb]
        assert abs(f[end].dot(b)) < 1e-12
        assert np.linalg.norm(f[end]) == pytest.approx(5.0 / np.linalg.norm(b), rel=1e-12)
    # small angle: cosine vs harmonic (alpha_eq = 0)
    q = configs.channel_params((12.0, 12.0, 14.0), 3, 1, 1, thermostat=False, bending=True, k_bend=5.0)
    r1 = np.array([[5, 5, 3.0], [5, 5, 4.0], [5 + math.sin(th), 5, 4.0 + math.cos(th)], [2, 2, 7.0]])
    fc, ec = oracle_mod.Oracle(_cosine(q)).bending(r1, 0)
    fh, eh = oracle_mod.Oracle(q).bending(r1, 0)
    assert ec == pytest.approx(5.0 * (1 - math.cos(th)), rel=1e-9)
    assert ec == pytest.approx(eh, rel=th * th)
    assert np.abs(fc - fh).max() < th * th * np.abs(fh).max()


# ------------------------------------------------------------- DPD-VV ----
@pytest.mark.parametrize("rc_dpd", [1.0, None])
def test_dpd_vv_force_is_pair_force_at_new_velocities(oracle_mod, rc_dpd):
    """With DPD_VV the force an md step leaves behind is the total force with
    the DPD pair part evaluated at the step's final velocities (same
    positions, same noise): fluid_fluid at (r1, v1) plus walls, bonds and
    body force, heads zero, and a = F / m (new_dpd_fd.f90).  Without DPD_VV
    it is the force at the half-step velocities, which differs."""
    p, t, r0, v = configs.channel_c1(thermostat=True, rc_dpd=rc_dpd)
    rng = np.random.default_rng(3)
    r0 = r0 + rng.normal(0, 0.05, r0.shape)
    r0[:, :2] %= 16.0
    heads = np.arange(p.n_chain) * p.n_mon
    v = v.copy()
    v[heads] = 0.0                                   # heads start at their fixed velocity
    p = p.replace(flags=Flags(**{**p.flags.__dict__, "dpd_vv": True}), r_2_min_init=0.0, r_2_min_time=0.0)
    o = oracle_mod.Oracle(p)
    a0 = np.zeros_like(r0)
    step = 7
    r1, v1, a1, f1, _, _ = o.md_steps(r0, v, a0, t, step, 1, noise=1)
    fp, _ = o.fluid_fluid_brute(r1, v1, t, step=step, noise=1)
    fw, _ = o.fluid_wall(r1, t)
    fb, _, _ = o.intra_molec(r1)
    want = fp + fw + fb
    want[heads] = 0.0
    assert np.abs(f1 - want).max() <= 1e-11 * np.abs(want).max()
    mass, inv = p.masses(t)
    assert np.abs(a1 - f1 * inv[:, None]).max() <= 1e-14 * np.abs(a1).max()
    q = p.replace(flags=Flags(**{**p.flags.__dict__, "dpd_vv": False}))
    _, v2, _, f2, _, _ = oracle_mod.Oracle(q).md_steps(r0, v, a0, t, step, 1, noise=1)
    assert np.array_equal(v2, v1)                    # DPD-VV leaves the velocities alone
    assert np.abs(f2 - f1).max() > 1e-3 * np.abs(want).max()


# ------------------------------------------------ FENE bond virial ----
def test_bond_virial_is_energy_derivative_under_strain(oracle_mod):
    """Bond part of the virial: W_aa = -dV_bond/d(lambda_a) under r_a ->
    lambda_a r_a (x and y with the box), and the shears likewise, V_bond the
    oracle's FENE energy of brush and melt chains (some bonds crossing the
    periodic x boundary)."""
    p, o, r0, t = _melt_system(oracle_mod)
    r0 = r0.copy()
    r0[:, 0] = (r0[:, 0] + 10.0) % 12.0                 # melt chains now wrap in x
    w = o.bond_virial(r0)
    h = 1e-6

    def energy(r, bx):
        return oracle_mod.Oracle(p.replace(boundary=tuple(bx))).intra_molec(r)[1]

    def d_scale(a):
        out = []
        for s in (1 + h, 1 - h):
            r = r0.copy(); r[:, a] *= s
            bx = list(p.boundary)
            if a < 2:
                bx[a] *= s
            out.append(energy(r, bx))
        return (out[0] - out[1]) / (2 * h)

    def d_shear(a, b):
        out = []
        for g in (h, -h):
            r = r0.copy(); r[:, a] += g * r0[:, b]
            out.append(energy(r, p.boundary))
        return (out[0] - out[1]) / (2 * h)

    want = [-d_scale(0), -d_scale(1), -d_scale(2), None, -d_shear(0, 2), -d_shear(1, 2)]
    scale = np.abs(w[:3]).max()
    assert scale > 1.0
    for q in (0, 1, 2, 4, 5):                           # z is not periodic: the xz and yz shears are exact
        assert abs(w[q] - want[q]) < 1e-6 * scale, (q, w, want)


def test_bond_virial_cluster_identity(oracle_mod):
    """sum_bonds r_ij (x) F_ij = sum_i r_i (x) F_i for chains that cross no
    periodic boundary, F the oracle's FENE force: the bond virial uses the
    force intra_molec adds, all six components."""
    p, o, r0, t = _melt_system(oracle_mod)
    r0 = r0.copy()
    r0[:, :2] = r0[:, :2] * 0.5 + 3.0                     # squeeze into the box interior
    f, _, _ = o.intra_molec(r0)
    assert np.abs(f).max() > 1.0
    want = [(r0[:, a] * f[:, b]).sum() for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))]
    w = o.bond_virial(r0)
    assert np.allclose(w, want, rtol=0, atol=1e-11 * np.abs(want).max()), (w, want)
    assert abs(w[3]) > 1e-3 * np.abs(w[:3]).max()


def test_bond_virial_single_bond_closed_form(oracle_mod):
    """One FENE bond along d: W = -k d (x) d / (R0^2 (1 - |d|^2/R0^2)), the
    reference's force factor k / (1 - q) / R0^2 (intra_molec.f90:46) times the
    separation, k the table value k_chain (already times R0^2)."""
    p = configs.channel_params((12.0, 12.0, 16.0), 2, 1, 0, thermostat=False)
    o = oracle_mod.Oracle(p)
    d = np.array([0.6, -0.5, 0.4])
    r0 = np.array([[6.0, 6.0, 8.0], [6.0, 6.0, 8.0]]) + np.array([[0.0, 0.0, 0.0], d])
    q = d.dot(d) * p.inv_r_chain_2
    rd = p.k_chain / (1.0 - q) * p.inv_r_chain_2
    dij = -d                                              # r_0 - r_1
    want = [-rd * dij[a] * dij[b] for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))]
    assert np.allclose(o.bond_virial(r0), want, rtol=1e-12)


# The reference's own record of its try/ run (try/fort.41: step, V_ff, V_bond,
# V_wall every 5 steps, from try/conf_old) is followed for 30 steps.  Its
# DPD noise came from the ziggurat stream, ours from Philox, so the
# trajectories part slowly: V_ff within 0.5 %, the bond and wall energies
# within 1e-4 (measured: 0.18 %, 3e-5 and 2e-5 at step 30).
FORT41_TOL = {"v_fluid_fluid": 5e-3, "v_intra_molec": 1e-4, "v_fluid_wall": 1e-4}


def fort41_errors(energies, fort41):
    """Relative errors of (V_ff, V_bond, V_wall) against the fort.41 rows."""
    out = []
    for en, row in zip(energies, fort41):
        out.append({k: abs(en[k] - ref) / abs(ref)
                    for k, ref in zip(("v_fluid_fluid", "v_intra_molec", "v_fluid_wall"), row[1:4])})
    return out


def test_oracle_follows_try_fort41(try_fixture, oracle_mod):
    p = try_fixture["p"]
    o = oracle_mod.Oracle(p, omp=True)
    r0, v, a = try_fixture["r0"], try_fixture["v"], try_fixture["a"]
    t = try_fixture["a_type"]
    f41 = try_fixture["fort41"][:6]
    assert np.array_equal(f41[:, 0], [5, 10, 15, 20, 25, 30])
    ens = []
    for i in range(6):
        r0, v, a, _, en, _ = o.md_steps(r0, v, a, t, 5 * i, 5, noise=1)
        ens.append(en)
    for i, err in enumerate(fort41_errors(ens, f41)):
        for k, tol in FORT41_TOL.items():
            assert err[k] < tol, (5 * (i + 1), k, err[k])
