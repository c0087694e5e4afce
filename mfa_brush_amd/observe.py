"""Running observables and the reference's observable files.

Mirrors observation.f90:14-330 for the channel (WALL 2) case: each call to
Observer.observe() adds the current energies (totals as in
verlet_velocities.f90:132-178), the chain observables of the device state
(mean squared radius of gyration per axis of the brush and melt chains,
their anisotropy, the total momentum; Simulation.chain_observables) and, on
every fifth observation step after n_relax (observation.f90:295), returns
the lines the reference writes to fort.40-43 and fort.80 (WR_FORCES 0), in
its formats.
"""
from __future__ import annotations

import math

import numpy as np

from .params import wall_wall


def fortran_e(x: float, w: int = 18, d: int = 7) -> str:
    """Fortran Ew.d output: 0.ddddddd mantissa, signed two-digit exponent."""
    if x == 0.0 or not math.isfinite(x):
        body = ("0." + "0" * d + "E+00") if x == 0.0 else str(x)
        return body.rjust(w)
    m, e = abs(x), 0
    m_str = ""
    e = math.floor(math.log10(m)) + 1
    for _ in range(3):              # the rounded mantissa can reach 1.0: renormalise
        mant = round(m / 10.0 ** e, d)
        if mant >= 1.0:
            e += 1
            continue
        if mant < 0.1:
            e -= 1
            continue
        m_str = ("%.*f" % (d, mant))[1:] if mant < 1.0 else "%.*f" % (d, mant)
        break
    s = "%s0%sE%s%02d" % ("-" if x < 0 else "", m_str, "-" if e < 0 else "+", abs(e))
    return s.rjust(w)


class Observer:
    """Running sums of observation.f90 for one channel of a Simulation."""

    def __init__(self, sim, params, isys: int = 0, n_relax: int = 0):
        self.sim, self.p, self.isys, self.n_relax = sim, params, isys, n_relax
        self.count_obs = 0
        self.v_total_1 = self.t_total_1 = self.e_total_1 = 0.0
        self.tot_P = np.zeros(3)
        self.r_g2_mean = np.zeros((3, 3))          # brushes, melt, total (:274-276)
        self.r_par_per2 = np.zeros(2)

    def totals(self, e: dict):
        """v_total, t_total, e_total of verlet_velocities.f90:155-178 (WALL 2: t_wall = 0)."""
        f = self.p.flags
        v_total = wall_wall(self.p) + e["v_fluid_wall"] + e["v_fluid_fluid"] + e["v_intra_molec"]
        if f.bending:
            v_total = v_total + e["v_bend"]
        if f.bending_melt:
            v_total = v_total + e["v_bend_melt"]
        t_total = 0.0 + e["t_fluid"]
        return v_total, t_total, v_total + t_total

    def observe(self, i_time: int, r_time: float | None = None) -> dict:
        """Accumulate one observation; returns {file: line} on output steps."""
        e = self.sim.energies(self.isys)
        co = self.sim.chain_observables(self.isys)
        self.count_obs += 1
        v_total, t_total, e_total = self.totals(e)
        self.v_total_1 += v_total
        self.t_total_1 += t_total
        self.e_total_1 += e_total
        gb, gm = co["rg2_brush"], co["rg2_melt"]
        self.r_g2_mean[0] += gb
        self.r_g2_mean[1] += gm
        self.r_g2_mean[2] += (gb + gm) / 2.0
        self.r_par_per2 += co["anisotropy"]
        self.tot_P += co["momentum"]
        if (i_time - self.n_relax) % 5 != 0:
            return {}
        inv = 1.0 / float(self.count_obs)
        out = {
            "fort.40": "%7d" % i_time + "".join("%15.5f" % x for x in
                                                 [self.v_total_1 * inv, self.t_total_1 * inv, self.e_total_1 * inv,
                                                  *(self.tot_P * inv)]),
            "fort.41": "%7d" % i_time + "".join("%15.5f" % x for x in
                                                 [e["v_fluid_fluid"], e["v_intra_molec"], e["v_fluid_wall"]]),
            "fort.42": "%7d" % i_time + "".join("%15.7f" % x for x in
                                                 [*(self.r_g2_mean[0] * inv), self.r_g2_mean[0].sum() * inv,
                                                  *(self.r_g2_mean[1] * inv), self.r_g2_mean[1].sum() * inv]),
            "fort.43": "%7d" % i_time + "".join("%15.7f" % x for x in self.r_par_per2 * inv),
        }
        n_chain = self.p.n_chain
        if n_chain:
            fh = self.sim.f_on_heads(self.isys)
            half = n_chain // 2
            rt = r_time if r_time is not None else i_time * self.p.dt
            vals = [rt, *(-fh[:half].sum(axis=0)), *(-fh[half:].sum(axis=0))]
            out["fort.80"] = "".join(fortran_e(x) for x in vals)
        return out
