"""Run-time parameters of the per-step path and the tables it consumes.

The reference selects physics with compile-time macros (control_simulation.h)
and derives every kernel constant once in init_params.f90:393-887.  Here the
hot-path subset of those macros is a run-time :class:`Flags` record and
:func:`derive_tables` restates the derivation, including the reference's
single-precision literals widened to FP64 (``2.**(1./6.)``, ``2.2``, ``0.98``,
``0.4``, ``0.9``: init_params.f90:497,511,546,567,575-584,827,
init_force_switch_on.f90:14).

Readers for the reference's run-time inputs follow init_system.f90:14-138
(``system_input``), init_params.f90:139-245 (``mfa_input``) and
init_config.f90:103-173 (``conf_old``); :func:`write_conf_new` writes the
restart format of store_config.f90:14-63.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

N_TYPE = 4  # WALL == 2, no PARTICLE_4 (mfa_common.f90:51-57)

_F = np.float32


def f2d(x: float) -> float:
    """A single-precision Fortran literal as the FP64 value it widens to."""
    return float(_F(x))


# (2.**(1./6.)) folded in single precision, then widened (Q1)
TWO_SIXTH = float(_F(2.0) ** (_F(1.0) / _F(6.0)))


@dataclass
class Flags:
    """Hot-path subset of control_simulation.h as run-time switches.

    Field -> macro: solvent SOLVENT (:42), hydrophobia HYDROPHOBIA (:44),
    brush_in_good_sv BRUSH_IN_GOOD_SV (:46), bending BENDING (:47),
    bending_melt BENDING_MELT (:48), thermostat THERMOSTAT==0 (:59),
    dpd_cut_off DPD_CUT_OFF (:63, None = undefined), force_switch_on
    FORCE_SWITCH_ON (:73), poiseuille POISEUILLE (:80), l_box L_BOX (:105),
    langevin THERMOSTAT==1 with the Langevin bead forces of lgv_forces.f90
    (used when thermostat is off; the reference's current fluid_fluid.f90
    leaves its call commented out, :455-458), bend_cosine: the bending terms
    take the cosine form k_bend (1 - cos theta) instead of the reference's
    harmonic angle (bending.f90:97-100), dpd_vv DPD_VV (:65): the
    dissipative forces are evaluated again after verlet_velocities
    (new_dpd_fd.f90).
    WALL is fixed to 2 (implicit 9-3 walls) and SYMMETRY to 0 (channel).
    """
    solvent: int = 3
    hydrophobia: bool = False
    brush_in_good_sv: bool = False
    bending: bool = True
    bending_melt: bool = False
    thermostat: bool = True
    dpd_cut_off: Optional[float] = None
    force_switch_on: bool = True
    poiseuille: bool = True
    l_box: int = 1
    langevin: bool = False
    bend_cosine: bool = False
    dpd_vv: bool = False


@dataclass
class SystemInput:
    """Contents of ``system_input`` (init_system.f90:14-138)."""
    n_mon: int
    n_chain: int
    n_cell_w_x: int
    n_cell_w_y: int
    n_mon_d: int
    n_chain_d: int
    delta_sig: float = 1.0
    a_w: float = 1.0
    sigma_w: float = 1.0
    const_force: tuple = (0.0, 0.0, 0.0)


@dataclass
class MfaInput:
    """Contents of ``mfa_input`` (init_params.f90:139-245)."""
    s_time: int
    n_relax: int
    n_obser: int
    n_safe: int
    dt: float
    temp: float
    x_space: float
    y_space: float
    z_space_wall: float
    mass_type: List[float]
    epsil: List[List[float]]          # epsil[i][j], i,j = 0..3 (type-1)
    sigma_diag: List[float]
    friction: List[float]
    k_spring_wall: float
    k_bend: float
    alpha_eq: float
    iseed: int
    k_chain: float
    r_chain: float
    f_twall: tuple = (0, 0, 0)
    va_spring_twall: tuple = (0.0, 0.0, 0.0)


class _Records:
    """Fortran list-directed reading: an item read skips blank records and
    takes the leading tokens; a bare ``read(u,*)`` consumes one record."""

    def __init__(self, text: str):
        self.lines = text.splitlines()
        self.k = 0

    def skip(self, n: int = 1):
        self.k += n

    def items(self, n: int) -> List[str]:
        out: List[str] = []
        while len(out) < n:
            if self.k >= len(self.lines):
                raise ValueError("input ended while reading %d items" % n)
            toks = self.lines[self.k].replace(",", " ").split()
            self.k += 1
            out.extend(toks)
        return out[:n]

    def ints(self, n):
        return [int(t) for t in self.items(n)]

    def reals(self, n):
        return [float(t.replace("d", "e").replace("D", "E")) for t in self.items(n)]


def read_system_input(text: str, flags: Flags) -> SystemInput:
    r = _Records(text)
    r.skip(4)
    n_mon, n_chain, nx, ny = r.ints(4)
    r.skip()
    tk = r.items(5)
    n_mon_d, n_chain_d = int(tk[0]), int(tk[1])
    delta_sig = 1.0
    if flags.solvent in (2, 3):
        r.skip()
        delta_sig = r.reals(1)[0]
    r.skip()
    a_w, sigma_w = r.reals(2)
    const_force = (0.0, 0.0, 0.0)
    if flags.poiseuille:
        r.skip()
        const_force = tuple(r.reals(3))
    return SystemInput(n_mon, n_chain, nx, ny, n_mon_d, n_chain_d, delta_sig, a_w, sigma_w, const_force)


def read_mfa_input(text: str, flags: Flags) -> MfaInput:
    r = _Records(text)
    s_time, = r.ints(1)
    n_relax, = r.ints(1)
    n_obser, = r.ints(1)
    n_safe, = r.ints(1)
    dt, = r.reals(1)
    r.skip(2)
    temp, = r.reals(1)
    r.skip(3)
    x_space, = r.reals(1)
    y_space, = r.reals(1)
    z_space_wall, = r.reals(1)
    r.skip()
    mass, eps, sig, fric = [], [], [], []
    for _ in range(N_TYPE):
        mass.append(r.reals(1)[0])
        eps.append([r.reals(1)[0] for _ in range(N_TYPE)])
        sig.append(r.reals(1)[0])
        fric.append(r.reals(1)[0])
    k_spring_wall, = r.reals(1)
    k_bend = alpha_eq = 0.0
    if flags.bending:
        k_bend, = r.reals(1)
        alpha_eq, = r.reals(1)
    r.skip(2)
    iseed, = r.ints(1)
    r.skip()
    k_chain, = r.reals(1)
    r_chain, = r.reals(1)
    r.skip(4)
    f_twall = tuple(r.ints(3))
    va = tuple(r.reals(3))
    return MfaInput(s_time, n_relax, n_obser, n_safe, dt, temp, x_space, y_space, z_space_wall,
                    mass, eps, sig, fric, k_spring_wall, k_bend, alpha_eq, iseed, k_chain, r_chain,
                    f_twall, va)


@dataclass
class SimParams:
    """Everything the per-step kernels read for one channel (one replica).

    Tables are Fortran column-major n_type x n_type: ``tab[(ti-1) + 4*(tj-1)]``.
    """
    n_mon: int
    n_chain: int
    n_mon_d: int
    n_chain_d: int
    boundary: tuple
    epsil: np.ndarray
    sigma_2: np.ndarray
    range_2: np.ndarray
    e_shift: np.ndarray
    inv_range_2: np.ndarray
    r_cut_max: float
    r_cut_dpd_2: float
    sig: float
    friction: float
    r_2_min_init: float
    r_2_min_time: float
    sigma_wall: np.ndarray
    a_wall: np.ndarray
    a_w: float
    sigma_w: float
    k_chain: float
    inv_r_chain_2: float
    k_bend: float
    alpha_eq: float
    dt: float
    mass_type: np.ndarray
    const_force: tuple = (0.0, 0.0, 0.0)
    head_vel: tuple = (0.0, 0.0)
    seed: int = 0x5EED0001
    flags: Flags = field(default_factory=Flags)
    skin: float = 0.0
    temp: float = 1.0

    @property
    def n_part(self) -> int:
        return self.n_mon * self.n_chain + self.n_mon_d * self.n_chain_d

    @property
    def part_init_d(self) -> int:
        return self.n_mon * self.n_chain

    @property
    def dt_2(self) -> float:
        return self.dt ** 2

    @property
    def z_space_wall(self) -> float:
        return float(self.boundary[2])

    def replace(self, **kw) -> "SimParams":
        return dataclasses.replace(self, **kw)

    def masses(self, a_type: np.ndarray):
        mass = self.mass_type[np.asarray(a_type) - 1].astype(np.float64)
        inv = np.array([1.0 / m for m in mass])
        return mass, inv

    def r_2_min(self, r_time: float) -> float:
        """Force switch-on ramp, fluid_fluid.f90:35-42."""
        if not self.flags.force_switch_on:
            return 0.0
        if r_time < self.r_2_min_time:
            return (1.0 - r_time / self.r_2_min_time) * self.r_2_min_init
        return 0.0


def _t(ti: int, tj: int) -> int:
    return (ti - 1) + N_TYPE * (tj - 1)


def derive_tables(flags: Flags, si: SystemInput, mi: MfaInput, *,
                  s_time: Optional[int] = None, seed: Optional[int] = None,
                  skin: float = 0.0) -> SimParams:
    """init_params.f90:393-887 for WALL=2, SYSTEM=0, SYMMETRY=0, n_type=4."""
    nt = N_TYPE
    dt = mi.dt
    r_chain_2 = mi.r_chain ** 2
    inv_r_chain_2 = 1.0 / r_chain_2
    k_chain = mi.k_chain * r_chain_2                                   # :396
    sigma = np.zeros((nt, nt))
    for i in range(nt):
        sigma[i, i] = mi.sigma_diag[i]
    for i in range(nt - 1):                                            # :406-411 / :424-429
        for j in range(i + 1, nt):
            sigma[i, j] = (sigma[i, i] + sigma[j, j]) / 2
            sigma[j, i] = sigma[i, j]
    if flags.solvent in (0, 1):
        sigma[:, nt - 1] = 0.0                                         # :414-415
        sigma[nt - 1, :] = 0.0
    else:
        for (a, b) in ((1, 2), (2, 1), (0, 2), (2, 0)):                # :432-435
            sigma[a, b] = si.delta_sig
    sigma_2 = sigma ** 2                                               # :467
    epsil = 4.0 * np.array(mi.epsil, dtype=np.float64)                 # :468
    epsil[0, 0] = 0.0                                                  # :492
    ts = TWO_SIXTH
    range_2 = (ts * sigma) ** 2                                        # :497
    range_2[0, nt - 1] = (f2d(2.2) * sigma[0, nt - 1]) ** 2            # :511-512
    range_2[nt - 1, 0] = (f2d(2.2) * sigma[nt - 1, 0]) ** 2
    two_ts = float(_F(2.0) * _F(ts))
    if flags.solvent == 0:                                             # :546-551
        range_2 = (two_ts * sigma) ** 2
        range_2[:, nt - 1] = 0.0
        range_2[nt - 1, :] = 0.0
    if flags.solvent == 3:                                             # :567
        range_2 = (two_ts * sigma) ** 2
    if flags.hydrophobia:                                              # :575-578
        for (a, b) in ((0, 2), (1, 2), (2, 0), (2, 1)):
            range_2[a, b] = (ts * sigma[a, b]) ** 2
    if flags.brush_in_good_sv:                                         # :582-585
        for (a, b) in ((0, 1), (0, 0), (1, 1), (1, 0)):
            range_2[a, b] = (ts * sigma[a, b]) ** 2
    sigma_wall = np.array([(si.sigma_w + sigma[t, t]) / 2 for t in range(nt)])   # :593-596
    a_wall = np.full(nt, si.a_w)
    range_2[:, nt - 1] = 0.0                                           # :618-619
    range_2[nt - 1, :] = 0.0
    range_2[0, 0] = 0.0                                                # :719
    e_shift = np.zeros((nt, nt))
    for i in range(nt):                                                # :726-737
        for j in range(nt):
            if range_2[i, j] > 0.0:
                rd = (sigma[j, i] ** 2 / range_2[j, i]) ** 3
                e_shift[j, i] = rd * (rd - 1.0)
    r_cut_max = math.sqrt(range_2.max())                               # :792
    r_cut_dpd_2 = 0.0
    if flags.dpd_cut_off is not None:                                  # :800-803
        r_cut_dpd = f2d(flags.dpd_cut_off)
        r_cut_max = max(r_cut_dpd, r_cut_max)
        r_cut_dpd_2 = r_cut_dpd ** 2
    inv_range_2 = np.where(range_2 != 0.0, 1.0 / np.where(range_2 != 0.0, range_2, 1.0), 0.0)  # :812-821
    sig = math.sqrt(2.0 * mi.temp * mi.friction[0] / dt)               # :887
    maxsig = sigma.max()                                               # init_force_switch_on.f90:12-21
    r_2_min_init = (f2d(0.9) * maxsig) ** 2
    r_2_min_time = 40.0 if mi.n_relax * dt < 40.0 else mi.n_relax * dt
    s_t = mi.s_time if s_time is None else s_time
    if s_t != 1:                           # continuation run: no ramp (init_config.f90:287-302)
        r_2_min_init = r_2_min_time = 0.0
    boundary = (mi.x_space * si.n_cell_w_x, mi.y_space * si.n_cell_w_y, mi.z_space_wall)  # init_config.f90:11-19

    def col(m):  # Fortran (i,j) column-major flattening
        return np.ascontiguousarray(m.T.reshape(-1))

    return SimParams(
        n_mon=si.n_mon, n_chain=si.n_chain, n_mon_d=si.n_mon_d, n_chain_d=si.n_chain_d,
        boundary=boundary, epsil=col(epsil), sigma_2=col(sigma_2), range_2=col(range_2),
        e_shift=col(e_shift), inv_range_2=col(inv_range_2), r_cut_max=r_cut_max,
        r_cut_dpd_2=r_cut_dpd_2, sig=sig, friction=mi.friction[0],
        r_2_min_init=r_2_min_init, r_2_min_time=r_2_min_time,
        sigma_wall=sigma_wall, a_wall=a_wall, a_w=si.a_w, sigma_w=si.sigma_w,
        k_chain=k_chain, inv_r_chain_2=inv_r_chain_2, k_bend=mi.k_bend, alpha_eq=mi.alpha_eq,
        dt=dt, mass_type=np.array(mi.mass_type, dtype=np.float64),
        const_force=tuple(si.const_force), head_vel=tuple(mi.va_spring_twall[:2]),
        seed=int(mi.iseed if seed is None else seed) & 0xFFFFFFFFFFFFFFFF,
        flags=flags, skin=skin, temp=mi.temp)


def wall_wall(p: SimParams) -> float:
    """wall_wall(2), wall_wall.f90:45-47: the constant wall-wall energy."""
    q = p.sigma_w / p.z_space_wall
    return 2.0 * (abs(p.a_w) * q ** 9 - p.a_w * q ** 3)


def read_conf_old(text: str, n_part: int):
    """conf_old reader, init_config.f90:103-173: returns (a_type, r0, v, a) with
    r0/v/a shaped (n_part, 3) -- row i is reference column i+1."""
    lines = text.splitlines()
    k = 4                                   # header, (n_mon,n_chain), cells, j_order
    a_type = np.empty(n_part, dtype=np.int32)
    r0 = np.empty((n_part, 3))
    for i in range(n_part):
        t = lines[k + i].split()
        a_type[i] = int(t[0])
        r0[i] = [float(x) for x in t[1:4]]
    k += n_part
    v = np.array([[float(x) for x in lines[k + i].split()[:3]] for i in range(n_part)])
    k += n_part
    a = np.array([[float(x) for x in lines[k + i].split()[:3]] for i in range(n_part)])
    return a_type, r0, v, a


def write_conf_new(path: str, p: SimParams, a_type, r0, v, a, n_cell_w=(1, 1)) -> None:
    """Restart file in the layout of store_config.f90:14-63 (WALL=2: no
    wall-atom records; the trailing wall-state lines are zero)."""
    with open(path, "w") as fh:
        fh.write("1\n")
        fh.write("%10d%10d%10d%10d\n" % (p.n_mon, p.n_chain, p.n_mon_d, p.n_chain_d))
        fh.write("%10d%10d\n" % tuple(n_cell_w))
        fh.write("%3d\n" % 2)
        for t, r in zip(a_type, r0):
            fh.write("%3d%13.7f%13.7f%13.7f\n" % (t, r[0], r[1], r[2]))
        for row in v:
            fh.write("%15.7f%15.7f%15.7f\n" % tuple(row))
        for row in a:
            fh.write("%15.7f%15.7f%15.7f\n" % tuple(row))
