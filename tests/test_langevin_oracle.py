"""Langevin bead forces of the oracle (lgv_forces.f90:14-25), pinned by
their closed form: without noise the added force is exactly -gamma m v; with
noise, (F_lgv + gamma m v) / (sig sqrt(m)) is standard normal over beads
and axes and independent of the bead's DPD pair noise stream."""
import numpy as np
import pytest

from mfa_brush_amd import configs
from mfa_brush_amd.params import Flags
from oracle.oracle import Oracle


def _langevin(p):
    return p.replace(flags=Flags(**{**p.flags.__dict__, "thermostat": False, "langevin": True}))


def _forces(p, r0, v, t, step=11, noise=1):
    f, _ = Oracle(p).fluid_fluid(r0, v, t, step=step, noise=noise)
    return f


def test_friction_term_closed_form():
    p, t, r0, v = configs.channel_c1(thermostat=False, rc_dpd=None)
    plain = _forces(p, r0, v, t)
    pl = _langevin(p).replace(sig=0.0)
    lg = _forces(pl, r0, v, t)
    m = np.asarray(p.mass_type)[t - 1][:, None]
    assert np.array_equal(lg - plain, -(p.friction * m) * v) or \
        np.allclose(lg - plain, -(p.friction * m) * v, rtol=0, atol=1e-13 * np.abs(plain).max())


def test_noise_term_standard_normal():
    p, t, r0, v = configs.channel_c1(thermostat=False, rc_dpd=None)
    pl = _langevin(p)
    lg = _forces(pl, r0, v, t, noise=1)
    quiet = _forces(pl.replace(sig=0.0), r0, v, t, noise=1)
    m = np.asarray(p.mass_type)[t - 1][:, None]
    z = ((lg - quiet) / (pl.sig * np.sqrt(m))).ravel()
    n = z.size
    assert abs(z.mean()) < 4.0 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 6.0 * np.sqrt(2.0 / n)
    # a different step draws different noise
    lg2 = _forces(pl, r0, v, t, step=12, noise=1)
    assert np.corrcoef((lg - quiet).ravel(), (lg2 - quiet).ravel())[0, 1] < 0.1
