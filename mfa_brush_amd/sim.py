"""Host-side mirror of the reference's step-path subroutines over the C ABI.

:class:`Simulation` exposes the names md_main.f90 calls (verlet_positions,
fluid_fluid, fluid_wall, wall_wall, intra_molec, bending, bending_melt,
constant_force, verlet_velocities) plus the fused step, state transfer in
the reference's (3, n_part) layout and the observables.  Every call goes to
libmfa_hip.so; nothing here computes physics.
"""
from __future__ import annotations

import ctypes as C
from typing import List, Sequence

import numpy as np

from . import _lib
from .params import SimParams, wall_wall as _wall_wall_host


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def to_c(p: SimParams, t_idx0: int = 0) -> _lib.SystemParams:
    s = _lib.SystemParams()
    fl = p.flags
    s.n_mon, s.n_chain, s.n_mon_d, s.n_chain_d = p.n_mon, p.n_chain, p.n_mon_d, p.n_chain_d
    s.thermostat = 1 if fl.thermostat else (2 if getattr(fl, "langevin", False) else 0)
    s.dpd_cut_off = int(fl.dpd_cut_off is not None)
    s.force_switch_on = int(fl.force_switch_on)
    s.bending, s.bending_melt, s.poiseuille = int(fl.bending), int(fl.bending_melt), int(fl.poiseuille)
    s.l_box = fl.l_box
    s.bend_form = 1 if getattr(fl, "bend_cosine", False) else 0
    s.boundary = _lib._D3(*p.boundary)
    for name in ("epsil", "sigma_2", "range_2", "e_shift", "inv_range_2"):
        setattr(s, name, _lib._D16(*[float(x) for x in getattr(p, name)]))
    s.r_cut_max, s.r_cut_dpd_2, s.skin = p.r_cut_max, p.r_cut_dpd_2, p.skin
    s.sig, s.friction = p.sig, p.friction
    s.r_2_min_init, s.r_2_min_time = p.r_2_min_init, p.r_2_min_time
    s.sigma_wall = _lib._D4(*p.sigma_wall)
    s.a_wall = _lib._D4(*p.a_wall)
    s.a_w = p.a_w
    s.k_chain, s.inv_r_chain_2 = p.k_chain, p.inv_r_chain_2
    s.k_bend, s.alpha_eq = p.k_bend, p.alpha_eq
    s.dt = p.dt
    s.const_force = _lib._D3(*p.const_force)
    s.head_vel = _lib._D2(*p.head_vel)
    s.mass_type = _lib._D4(*p.mass_type)
    s.seed = p.seed
    s.t_idx0 = t_idx0
    s.dpd_vv = 1 if getattr(fl, "dpd_vv", False) else 0
    return s


class Simulation:
    """One or more channels (replicas) advanced together on the GPU."""

    def __init__(self, params: Sequence[SimParams] | SimParams, precision: str = "fp64", t_idx0: int = 0):
        if isinstance(params, SimParams):
            params = [params]
        self.params: List[SimParams] = list(params)
        self.lib = _lib.load()
        arr = (_lib.SystemParams * len(self.params))(*[to_c(p, t_idx0) for p in self.params])
        self._keep = arr
        prec = {"fp64": _lib.MFA_PREC_FP64, "mixed": _lib.MFA_PREC_MIXED}[precision]
        self.precision = precision
        h = C.c_void_p()
        _lib.check(self.lib.mfa_hip_create(arr, len(self.params), prec, C.byref(h)), "mfa_hip_create")
        self.h = h

    def close(self):
        if self.h:
            self.lib.mfa_hip_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc, what):
        _lib.check(rc, what)

    # ---- state ----------------------------------------------------------
    def set_state(self, r0, v, a_type, a=None, isys: int = 0):
        n = self.params[isys].n_part
        r0 = np.ascontiguousarray(r0, np.float64).reshape(n, 3)
        v = np.ascontiguousarray(v, np.float64).reshape(n, 3)
        a = None if a is None else np.ascontiguousarray(a, np.float64).reshape(n, 3)
        t = np.ascontiguousarray(a_type, np.int32).reshape(n)
        self._c(self.lib.mfa_hip_set_state(self.h, isys, _ptr(r0), _ptr(v), _ptr(a), _ptr(t)), "set_state")

    def get_state(self, isys: int = 0):
        n = self.params[isys].n_part
        out = {k: np.empty((n, 3)) for k in ("r0", "v", "a", "force")}
        self._c(self.lib.mfa_hip_get_state(self.h, isys, _ptr(out["r0"]), _ptr(out["v"]), _ptr(out["a"]),
                                           _ptr(out["force"])), "get_state")
        return out

    def set_time(self, t_idx: int):
        self._c(self.lib.mfa_hip_set_time(self.h, t_idx), "set_time")

    def get_time(self) -> int:
        t = C.c_int64()
        self._c(self.lib.mfa_hip_get_time(self.h, C.byref(t)), "get_time")
        return t.value

    # ---- step routines (md_main.f90:85-185) -----------------------------
    def verlet_positions(self):
        self._c(self.lib.mfa_hip_verlet_positions(self.h), "verlet_positions")

    def check_skin(self) -> int:
        """f_skin after the last verlet_positions (check_skin.f90:4-28): 1
        when its skin test fired, else 0; the rebuild it calls for is done."""
        f = C.c_int32()
        self._c(self.lib.mfa_hip_check_skin(self.h, C.byref(f)), "check_skin")
        return f.value

    def fluid_fluid(self):
        self._c(self.lib.mfa_hip_fluid_fluid(self.h), "fluid_fluid")

    def fluid_wall(self, inter_type: int = 2):
        self._c(self.lib.mfa_hip_fluid_wall(self.h, inter_type), "fluid_wall")

    def wall_wall(self, inter_type: int = 2) -> float:
        v = C.c_double()
        self._c(self.lib.mfa_hip_wall_wall(self.h, inter_type, self.params[0].sigma_w, C.byref(v)), "wall_wall")
        return v.value

    def intra_molec(self):
        self._c(self.lib.mfa_hip_intra_molec(self.h), "intra_molec")

    def bending(self, mode: int = 1):
        self._c(self.lib.mfa_hip_bending(self.h, mode), "bending")

    def bending_melt(self, mode: int = 1):
        self._c(self.lib.mfa_hip_bending_melt(self.h, mode), "bending_melt")

    def constant_force(self, case_pois: int = 1):
        self._c(self.lib.mfa_hip_constant_force(self.h, case_pois), "constant_force")

    def verlet_velocities(self):
        self._c(self.lib.mfa_hip_verlet_velocities(self.h), "verlet_velocities")

    def new_dpd_fd(self):
        """new_dpd_fd.f90 (DPD_VV): dissipative forces at the new velocities."""
        self._c(self.lib.mfa_hip_new_dpd_fd(self.h), "new_dpd_fd")

    def step_by_routines(self):
        """One md_main.f90:77-185 iteration through the single-routine entry points."""
        p = self.params[0]
        self.verlet_positions()
        self.fluid_fluid()
        self.fluid_wall(2)
        self.intra_molec()
        if p.flags.bending:
            self.bending(1)
        if p.flags.bending_melt:
            self.bending_melt(1)
        if p.flags.poiseuille:
            self.constant_force(1)
        self.verlet_velocities()
        if getattr(p.flags, "dpd_vv", False) and p.flags.thermostat:
            self.new_dpd_fd()

    def md_steps(self, n: int, graph: bool = False):
        self._c(self.lib.mfa_hip_md_steps(self.h, n, int(graph)), "md_steps")

    def prepare_graphs(self):
        """Capture the graphs md_steps(graph=True) replays in the current
        timing mode, without running a step."""
        self._c(self.lib.mfa_hip_prepare_graphs(self.h), "prepare_graphs")

    def synchronize(self):
        self._c(self.lib.mfa_hip_synchronize(self.h), "synchronize")

    # ---- observables ----------------------------------------------------
    def energies(self, isys: int = 0) -> dict:
        e = _lib.Energies()
        self._c(self.lib.mfa_hip_get_energies(self.h, isys, C.byref(e)), "get_energies")
        d = {k: getattr(e, k) for k, _ in _lib.Energies._fields_ if not k.startswith("virial")}
        d["virial"] = np.array(list(e.virial))
        d["virial_bond"] = np.array(list(e.virial_bond))
        return d

    def f_on_heads(self, isys: int = 0):
        n = max(self.params[isys].n_chain, 1)
        f = np.zeros((n, 3))
        self._c(self.lib.mfa_hip_get_f_on_heads(self.h, isys, _ptr(f)), "get_f_on_heads")
        return f

    def chain_observables(self, isys: int = 0) -> dict:
        """Gyration and momentum observables of observation.f90:182-290 from
        the device state: mean squared radius of gyration per axis of the
        brush and melt chains (r_g2), their anisotropy (r_par_per2) and the
        total momentum (tot_P before its running average)."""
        out = np.zeros(9)
        self._c(self.lib.mfa_hip_chain_observables(self.h, isys, _ptr(out)), "chain_observables")
        gb, gm = out[0:3], out[3:6]
        with np.errstate(invalid="ignore", divide="ignore"):
            aniso = np.array([(g[2] - 0.5 * (g[0] + g[1])) / (g[0] + g[1] + g[2]) for g in (gb, gm)])
        return dict(rg2_brush=gb, rg2_melt=gm, anisotropy=aniso, momentum=out[6:9])

    def list_stats(self, isys: int = 0) -> dict:
        """Verlet-list statistics of the last rebuild: list entries, the
        entries the pair kernel's waves walk (each wave the longest list of
        its 64 lanes, in groups of four) and overflowed beads."""
        out = np.zeros(3, np.int64)
        self._c(self.lib.mfa_hip_list_stats(self.h, isys, _ptr(out)), "list_stats")
        return dict(entries=int(out[0]), wave_walk=int(out[1]), overflowed=int(out[2]))

    def flags(self) -> dict:
        f = _lib.Flags()
        self._c(self.lib.mfa_hip_get_flags(self.h, C.byref(f)), "get_flags")
        return {k: getattr(f, k) for k, _ in _lib.Flags._fields_}

    def check_fluid_fluid(self, isys: int = 0) -> float:
        v = C.c_double()
        self._c(self.lib.mfa_hip_check_fluid_fluid(self.h, isys, C.byref(v)), "check_fluid_fluid")
        return v.value

    def pair_stats(self, isys: int = 0):
        n = self.params[isys].n_part
        cnt = np.zeros(n, np.int32)
        h = np.zeros(n, np.uint64)
        self._c(self.lib.mfa_hip_pair_stats(self.h, isys, _ptr(cnt), _ptr(h)), "pair_stats")
        return cnt, h

    def pair_count(self, isys: int = 0) -> int:
        n = C.c_int64()
        self._c(self.lib.mfa_hip_get_pair_count(self.h, isys, C.byref(n)), "get_pair_count")
        return n.value

    def cells(self, isys: int = 0):
        p = self.params[isys]
        nc = np.zeros(3, np.int32)
        self._c(self.lib.mfa_hip_get_cells(self.h, isys, _ptr(nc), None, None, None), "get_cells")
        nct = int(np.prod(nc))
        st = np.zeros(nct, np.int32); cn = np.zeros(nct, np.int32); beads = np.zeros(p.n_part, np.int32)
        self._c(self.lib.mfa_hip_get_cells(self.h, isys, _ptr(nc), _ptr(st), _ptr(cn), _ptr(beads)), "get_cells")
        return nc, st, cn, beads

    def profiles_reset(self, n_bins: int = 200):
        self._c(self.lib.mfa_hip_profiles_reset(self.h, n_bins), "profiles_reset")
        self._nb = n_bins

    def profiles_accumulate(self):
        self._c(self.lib.mfa_hip_profiles_accumulate(self.h), "profiles_accumulate")

    def profiles(self, isys: int = 0):
        nb = self._nb
        nobs = C.c_int32()
        dens = np.zeros((2, nb)); vp = np.zeros((2, 3, nb)); vp2 = np.zeros((2, 3, nb))
        self._c(self.lib.mfa_hip_profiles_get(self.h, isys, C.byref(nobs), _ptr(dens), _ptr(vp), _ptr(vp2)),
                "profiles_get")
        return nobs.value, dens, vp, vp2

    def set_virial(self, on: bool = True):
        self._c(self.lib.mfa_hip_set_virial(self.h, int(on)), "set_virial")

    def set_neighbor_list(self, capacity: int = -1):
        """Verlet-list capacity per bead: -1 automatic, 0 cell scan only."""
        self._c(self.lib.mfa_hip_set_neighbor_list(self.h, int(capacity)), "set_neighbor_list")

    def set_dpd_queue(self, depth: int = 8):
        """DPD pairs per bead whose DPD part is deferred to a second pass (0..8)."""
        self._c(self.lib.mfa_hip_set_dpd_queue(self.h, int(depth)), "set_dpd_queue")

    def set_pair_grid(self, blocks: int = 0):
        """Mixed pair kernel launch: 0 one block per tile, > 0 persistent blocks."""
        self._c(self.lib.mfa_hip_set_pair_grid(self.h, int(blocks)), "set_pair_grid")

    def neighbor_list(self) -> int:
        n = C.c_int32()
        self._c(self.lib.mfa_hip_get_neighbor_list(self.h, C.byref(n)), "get_neighbor_list")
        return n.value

    def set_timing(self, on=True):
        """0/False off; 1/True every kernel group, eager launches; 2 the pair
        kernel only, with graph replay (event records captured as nodes)."""
        self._c(self.lib.mfa_hip_set_timing(self.h, int(on)), "set_timing")

    def kernel_time(self, name: str):
        ms = C.c_double(); n = C.c_int64()
        self._c(self.lib.mfa_hip_get_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)), "get_kernel_time")
        return ms.value, n.value


def make_binning(bin_case: int, boundary, r_cut: float):
    """make_binning (cell_list.f90:129-158) through the C ABI."""
    lib = _lib.load()
    b = np.ascontiguousarray(boundary, np.float64)
    l = np.zeros(3); il = np.zeros(3); nc = np.zeros(3, np.int32)
    nct = C.c_int32(); nn = C.c_int32()
    _lib.check(lib.mfa_hip_make_binning(bin_case, _ptr(b), r_cut, _ptr(l), _ptr(il), _ptr(nc),
                                        C.byref(nct), C.byref(nn)), "make_binning")
    return nc, nct.value, nn.value, l, il


def wall_wall(p: SimParams) -> float:
    return _wall_wall_host(p)
