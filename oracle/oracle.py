"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.  Arrays follow the
reference layout: r0/v/a/force are (n_part, 3) C-contiguous FP64 numpy arrays,
i.e. Fortran (3, n_part) column-major; a_type is int32 in 1..4.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

_D3 = C.c_double * 3
_D16 = C.c_double * 16
_D4 = C.c_double * 4
_D2 = C.c_double * 2


class OrcParams(C.Structure):
    _fields_ = [
        ("n_part", C.c_int), ("n_mon", C.c_int), ("n_chain", C.c_int),
        ("n_mon_d", C.c_int), ("n_chain_d", C.c_int),
        ("thermostat", C.c_int), ("dpd_cut_off", C.c_int), ("force_switch_on", C.c_int),
        ("bending", C.c_int), ("bending_melt", C.c_int), ("poiseuille", C.c_int),
        ("l_box", C.c_int),
        ("boundary", _D3), ("inv_boundary", _D3),
        ("epsil", _D16), ("sigma_2", _D16), ("range_2", _D16), ("e_shift", _D16),
        ("inv_range_2", _D16),
        ("r_cut_max", C.c_double), ("r_cut_dpd_2", C.c_double),
        ("sig", C.c_double), ("friction", C.c_double),
        ("r_2_min_init", C.c_double), ("r_2_min_time", C.c_double),
        ("sigma_wall", _D4), ("a_wall", _D4), ("a_w", C.c_double), ("z_space_wall", C.c_double),
        ("k_chain", C.c_double), ("inv_r_chain_2", C.c_double),
        ("k_bend", C.c_double), ("alpha_eq", C.c_double),
        ("dt", C.c_double), ("dt_2", C.c_double),
        ("const_force", _D3), ("head_vel", _D2), ("mass_type", _D4),
        ("seed", C.c_uint64), ("bend_form", C.c_int), ("dpd_vv", C.c_int),
    ]


class OrcEnergies(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("v_fluid_fluid", "v_intra_molec", "v_fluid_wall", "v_bend", "v_bend_melt", "t_fluid")]


class OrcCells(C.Structure):
    _fields_ = [("n_cells", C.c_int * 3), ("n_cells_tot", C.c_int), ("n_nei", C.c_int),
                ("l_cell", _D3), ("inv_l_cell", _D3),
                ("part_in_cell", C.POINTER(C.c_int)), ("r_nei", C.POINTER(C.c_int)),
                ("l_nei", C.POINTER(C.c_int)), ("cell_neigh_ls", C.POINTER(C.c_int))]


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def load(omp: bool = False) -> C.CDLL:
    name = "liboracle_omp.so" if omp else "liboracle.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError("oracle not built: run `make -C oracle` (%s missing)" % path)
    lib = C.CDLL(path)
    vp, i, d, u64 = C.c_void_p, C.c_int, C.c_double, C.c_uint64
    lib.orc_philox4x32_10.argtypes = [vp, vp, vp]
    lib.orc_pair_gauss.argtypes = [u64, u64, C.c_int32, C.c_int32]
    lib.orc_pair_gauss.restype = C.c_float
    lib.orc_make_binning.argtypes = [i, vp, d, vp, vp, vp, vp, vp]
    lib.orc_get_cell.argtypes = [vp, vp, vp]
    lib.orc_get_cell.restype = i
    lib.orc_neigh_list.argtypes = [vp, i, vp, i, vp, vp, vp, vp, vp, vp]
    lib.orc_get_cell_neigh_list.argtypes = [i, vp, i, vp]
    lib.orc_r2_min.argtypes = [vp, d]
    lib.orc_r2_min.restype = d
    lib.orc_fluid_fluid.argtypes = [vp, vp, vp, vp, vp, vp, vp, i, i, d, u64, i, vp, vp]
    lib.orc_check_fluid_fluid.argtypes = [vp, vp, vp, d]
    lib.orc_check_fluid_fluid.restype = d
    lib.orc_pair_stats.argtypes = [vp, vp, vp, vp, vp]
    lib.orc_pair_stats_cells.argtypes = [vp, vp, vp, vp, vp]
    lib.orc_pair_stats_cells.restype = i
    lib.orc_fluid_fluid_brute.argtypes = [vp, vp, vp, vp, d, u64, i, vp, vp]
    lib.orc_pair_virial.argtypes = [vp, vp, vp, vp, d, u64, i, vp]
    lib.orc_bond_virial.argtypes = [vp, vp, vp]
    lib.orc_fluid_wall.argtypes = [vp, vp, vp, vp]
    lib.orc_fluid_wall.restype = d
    lib.orc_wall_wall.argtypes = [vp, d]
    lib.orc_wall_wall.restype = d
    lib.orc_intra_molec.argtypes = [vp, vp, vp, vp]
    lib.orc_intra_molec.restype = d
    lib.orc_bending.argtypes = [vp, vp, vp, i]
    lib.orc_bending.restype = d
    lib.orc_constant_force.argtypes = [vp, vp, vp, vp]
    lib.orc_verlet_positions.argtypes = [vp, vp, vp, vp, i, vp, vp, vp, vp, vp]
    lib.orc_verlet_velocities.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    lib.orc_verlet_velocities.restype = d
    lib.orc_histo.argtypes = [vp, vp, i, vp]
    lib.orc_velocity_prof.argtypes = [vp, vp, vp, i, vp, vp]
    lib.orc_cells_init.argtypes = [vp, vp, vp, d]
    lib.orc_cells_init.restype = i
    lib.orc_cells_free.argtypes = [vp]
    lib.orc_md_steps.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_longlong, i, i, vp, vp]
    lib.orc_num_threads.restype = i
    return lib


def make_params(p) -> OrcParams:
    """OrcParams from a mfa_brush_amd.params.SimParams."""
    o = OrcParams()
    fl = p.flags
    o.n_part = p.n_part
    o.n_mon, o.n_chain, o.n_mon_d, o.n_chain_d = p.n_mon, p.n_chain, p.n_mon_d, p.n_chain_d
    o.thermostat = 1 if fl.thermostat else (2 if getattr(fl, "langevin", False) else 0)
    o.dpd_cut_off = int(fl.dpd_cut_off is not None)
    o.force_switch_on = int(fl.force_switch_on)
    o.bending, o.bending_melt, o.poiseuille = int(fl.bending), int(fl.bending_melt), int(fl.poiseuille)
    o.l_box = fl.l_box
    o.boundary = _D3(*p.boundary)
    o.inv_boundary = _D3(*[1.0 / b for b in p.boundary])
    for name in ("epsil", "sigma_2", "range_2", "e_shift", "inv_range_2"):
        setattr(o, name, _D16(*[float(x) for x in getattr(p, name)]))
    o.r_cut_max, o.r_cut_dpd_2 = p.r_cut_max, p.r_cut_dpd_2
    o.sig, o.friction = p.sig, p.friction
    o.r_2_min_init, o.r_2_min_time = p.r_2_min_init, p.r_2_min_time
    o.sigma_wall = _D4(*p.sigma_wall)
    o.a_wall = _D4(*p.a_wall)
    o.a_w, o.z_space_wall = p.a_w, p.boundary[2]
    o.k_chain, o.inv_r_chain_2 = p.k_chain, p.inv_r_chain_2
    o.k_bend, o.alpha_eq = p.k_bend, p.alpha_eq
    o.dt, o.dt_2 = p.dt, p.dt ** 2
    o.const_force = _D3(*p.const_force)
    o.head_vel = _D2(*p.head_vel)
    o.mass_type = _D4(*p.mass_type)
    o.seed = p.seed
    o.bend_form = 1 if getattr(fl, "bend_cosine", False) else 0
    o.dpd_vv = 1 if getattr(fl, "dpd_vv", False) else 0
    return o


class Oracle:
    """Convenience front-end used by the tests (one channel)."""

    def __init__(self, p, omp: bool = False):
        self.lib = load(omp)
        self.p = p
        self.o = make_params(p)
        self.po = C.byref(self.o)

    # -- cell list --------------------------------------------------------
    def make_binning(self, r_cut=None):
        l = np.zeros(3); il = np.zeros(3); nc = np.zeros(3, np.int32)
        nct = C.c_int(); nn = C.c_int()
        rc = self.p.r_cut_max if r_cut is None else r_cut
        self.lib.orc_make_binning(self.p.flags.l_box, _p(np.array(self.p.boundary, float)), rc,
                                  _p(l), _p(il), _p(nc), C.byref(nct), C.byref(nn))
        return nc, nct.value, nn.value, l, il

    def cell_lists(self, r0, r_cut=None):
        nc, nct, nn, l, il = self.make_binning(r_cut)
        pic = np.zeros(nct + 1, np.int32)
        rn = np.zeros(self.p.n_part, np.int32)
        ln = np.zeros(self.p.n_part, np.int32)
        nob = C.c_int()
        self.lib.orc_neigh_list(_p(r0), self.p.n_part, _p(nc), nct, _p(np.array(self.p.boundary, float)),
                                _p(il), _p(pic), _p(rn), _p(ln), C.byref(nob))
        cnl = np.zeros(nct * nn, np.int32)
        self.lib.orc_get_cell_neigh_list(nct, _p(nc), nn, _p(cnl))
        return dict(n_cells=nc, n_cells_tot=nct, n_nei=nn, l_cell=l, inv_l_cell=il,
                    part_in_cell=pic, r_nei=rn, l_nei=ln, cell_neigh_ls=cnl, n_out=nob.value)

    # -- forces -----------------------------------------------------------
    def fluid_fluid(self, r0, v, a_type, step=0, r_time=0.0, noise=1, cells=None):
        r0 = np.ascontiguousarray(r0, float); v = np.ascontiguousarray(v, float)
        a_type = np.ascontiguousarray(a_type, np.int32)
        c = cells or self.cell_lists(r0)
        f = np.zeros((self.p.n_part, 3))
        e = C.c_double()
        self.lib.orc_fluid_fluid(self.po, _p(r0), _p(v), _p(a_type), _p(c["part_in_cell"]), _p(c["r_nei"]),
                                 _p(c["cell_neigh_ls"]), c["n_cells_tot"], c["n_nei"],
                                 self.lib.orc_r2_min(self.po, r_time), step, noise, _p(f), C.byref(e))
        return f, e.value

    def fluid_fluid_brute(self, r0, v, a_type, step=0, r_time=0.0, noise=1):
        f = np.zeros((self.p.n_part, 3)); e = C.c_double()
        self.lib.orc_fluid_fluid_brute(self.po, _p(np.ascontiguousarray(r0, float)),
                                       _p(np.ascontiguousarray(v, float)),
                                       _p(np.ascontiguousarray(a_type, np.int32)),
                                       self.lib.orc_r2_min(self.po, r_time), step, noise, _p(f), C.byref(e))
        return f, e.value

    def pair_virial(self, r0, v, a_type, step=0, r_time=0.0, noise=1):
        """sum over interacting pairs of r_ij (x) F_ij: xx yy zz xy xz yz."""
        w = np.zeros(6)
        self.lib.orc_pair_virial(self.po, _p(np.ascontiguousarray(r0, float)), _p(np.ascontiguousarray(v, float)),
                                 _p(np.ascontiguousarray(a_type, np.int32)), self.lib.orc_r2_min(self.po, r_time),
                                 step, noise, _p(w))
        return w

    def bond_virial(self, r0):
        """sum over FENE bonds of r_ij (x) F_ij: xx yy zz xy xz yz."""
        w = np.zeros(6)
        self.lib.orc_bond_virial(self.po, _p(np.ascontiguousarray(r0, float)), _p(w))
        return w

    def check_fluid_fluid(self, r0, a_type, r_2_min=0.0):
        return self.lib.orc_check_fluid_fluid(self.po, _p(np.ascontiguousarray(r0, float)),
                                              _p(np.ascontiguousarray(a_type, np.int32)), r_2_min)

    def pair_stats(self, r0, a_type):
        cnt = np.zeros(self.p.n_part, np.int32); h = np.zeros(self.p.n_part, np.uint64)
        self.lib.orc_pair_stats(self.po, _p(np.ascontiguousarray(r0, float)),
                                _p(np.ascontiguousarray(a_type, np.int32)), _p(cnt), _p(h))
        return cnt, h

    def pair_stats_cells(self, r0, a_type):
        cnt = np.zeros(self.p.n_part, np.int32); h = np.zeros(self.p.n_part, np.uint64)
        rc = self.lib.orc_pair_stats_cells(self.po, _p(np.ascontiguousarray(r0, float)),
                                           _p(np.ascontiguousarray(a_type, np.int32)), _p(cnt), _p(h))
        if rc != 0:
            raise MemoryError("oracle cells")
        return cnt, h

    def fluid_wall(self, r0, a_type):
        f = np.zeros((self.p.n_part, 3))
        e = self.lib.orc_fluid_wall(self.po, _p(np.ascontiguousarray(r0, float)),
                                    _p(np.ascontiguousarray(a_type, np.int32)), _p(f))
        return f, e

    def intra_molec(self, r0):
        f = np.zeros((self.p.n_part, 3)); ns = C.c_int()
        e = self.lib.orc_intra_molec(self.po, _p(np.ascontiguousarray(r0, float)), _p(f), C.byref(ns))
        return f, e, ns.value

    def bending(self, r0, melt=0):
        f = np.zeros((self.p.n_part, 3))
        e = self.lib.orc_bending(self.po, _p(np.ascontiguousarray(r0, float)), _p(f), melt)
        return f, e

    def constant_force(self, a_type):
        mass, _ = self.p.masses(a_type)
        f = np.zeros((self.p.n_part, 3))
        self.lib.orc_constant_force(self.po, _p(mass), _p(np.ascontiguousarray(a_type, np.int32)), _p(f))
        return f

    def verlet_positions(self, r0, v, a):
        r0 = np.array(r0, float); v = np.array(v, float)
        nc = np.ones(3, np.int32); il = np.ones(3)
        self.lib.orc_verlet_positions(self.po, _p(r0), _p(v), _p(np.ascontiguousarray(a, float)), 0,
                                      _p(nc), _p(il), None, None, None)
        return r0, v

    def verlet_velocities(self, force, v, a_type):
        mass, inv = self.p.masses(a_type)
        v = np.array(v, float); a = np.zeros_like(v)
        fh = np.zeros((max(self.p.n_chain, 1), 3))
        force = np.array(force, float)          # heads' force is zeroed in this copy
        t = self.lib.orc_verlet_velocities(self.po, _p(force), _p(inv), _p(mass), _p(v), _p(a), _p(fh))
        return v, a, t, fh

    def histo(self, r0, hist_dim=200):
        dz = np.zeros((2, hist_dim))
        self.lib.orc_histo(self.po, _p(np.ascontiguousarray(r0, float)), hist_dim, _p(dz))
        return dz

    def velocity_prof(self, r0, v, n_layers=200):
        vp = np.zeros((2, 3, n_layers)); vp2 = np.zeros((2, 3, n_layers))
        self.lib.orc_velocity_prof(self.po, _p(np.ascontiguousarray(r0, float)),
                                   _p(np.ascontiguousarray(v, float)), n_layers, _p(vp), _p(vp2))
        return vp, vp2

    def md_steps(self, r0, v, a, a_type, t_idx0, n_steps, noise=1):
        """md_main.f90:77-185 for n_steps; returns new (r0, v, a, force, energies)."""
        r0 = np.array(r0, float); v = np.array(v, float); a = np.array(a, float)
        a_type = np.ascontiguousarray(a_type, np.int32)
        f = np.zeros_like(r0)
        mass, inv = self.p.masses(a_type)
        ext = np.zeros_like(r0)
        if self.p.flags.poiseuille:
            ext[a_type != 1] = mass[a_type != 1, None] * np.array(self.p.const_force)[None, :]
        cells = OrcCells()
        rc = self.p.r_cut_max
        if self.lib.orc_cells_init(C.byref(cells), self.po, _p(r0), rc) != 0:
            raise MemoryError("oracle cells")
        e = OrcEnergies()
        fh = np.zeros((max(self.p.n_chain, 1), 3))
        try:
            self.lib.orc_md_steps(self.po, C.byref(cells), _p(r0), _p(v), _p(a), _p(f), _p(a_type),
                                  _p(mass), _p(inv), _p(ext), t_idx0, n_steps, noise, C.byref(e), _p(fh))
        finally:
            self.lib.orc_cells_free(C.byref(cells))
        en = {k: getattr(e, k) for k, _ in OrcEnergies._fields_}
        return r0, v, a, f, en, fh


# ---------------------------------------------------------------------------
# The reference's own routines (oracle/_ref, built by oracle/ref.mk from the
# reference sources with the harness oracle/ref_harness.f90).
# ---------------------------------------------------------------------------
REF_DIR = os.path.join(HERE, "_ref")


def chain_observables(p, r0, v, a_type):
    """Gyration and momentum part of observation.f90:182-290 (numpy, bead order).

    Chains are unfolded bead by bead against the previous, already unfolded
    bead (:182-237); per chain the centre of mass and the mean squared
    deviation per axis are summed in bead order and normalised as the
    reference does, brush times 1/n_mon, melt divided by n_mon_d (:247-271);
    the chain means are summed in chain order, brush times 1/n_chain, melt
    divided by n_chain_d (:256, :272); the anisotropy is
    (Rg2_z - (Rg2_x + Rg2_y)/2) / sum(Rg2) (:278-279); the momentum is
    sum(mass * v) in bead order (:286-290).  Returns (rg2_brush[3],
    rg2_melt[3], anisotropy[2], momentum[3])."""
    r0 = np.asarray(r0, float)
    L = np.array(p.boundary, float)
    h = L / 2.0

    def rg2(first, n_ch, n_m, brush):
        if n_ch == 0 or n_m == 0:
            return np.zeros(3)
        X = r0[first:first + n_ch * n_m].reshape(n_ch, n_m, 3).copy()
        for j in range(1, n_m):
            x, prev = X[:, j, :], X[:, j - 1, :]
            up = (x - prev) > h
            dn = ~up & ((prev - x) > h)
            X[:, j, :] = np.where(up, x - L, np.where(dn, x + L, x))
        inv = 1.0 / float(n_m)
        s = np.zeros((n_ch, 3))
        for j in range(n_m):
            s = s + X[:, j, :]
        cm = s * inv if brush else s / float(n_m)
        acc = np.zeros((n_ch, 3))
        for j in range(n_m):
            e = X[:, j, :] - cm
            acc = acc + e * e
        g = acc * inv if brush else acc / float(n_m)
        tot = np.zeros(3)
        for c in range(n_ch):
            tot = tot + g[c]
        return tot * (1.0 / float(n_ch)) if brush else tot / float(n_ch)

    gb = rg2(0, p.n_chain, p.n_mon, True)
    gm = rg2(p.n_chain * p.n_mon, p.n_chain_d, p.n_mon_d, False)
    with np.errstate(invalid="ignore", divide="ignore"):
        aniso = np.array([(g[2] - 0.5 * (g[0] + g[1])) / (g[0] + g[1] + g[2]) for g in (gb, gm)])
    m = np.asarray(p.mass_type, float)[np.asarray(a_type) - 1]
    mom = np.cumsum(m[:, None] * np.asarray(v, float), axis=0)[-1]
    return gb, gm, aniso, mom


def ref_available(omp: bool = False) -> bool:
    return os.path.exists(os.path.join(REF_DIR, "libmfa_ref_omp.so" if omp else "libmfa_ref.so"))


class RefLib:
    """One channel in the reference's module state; routines via ref_call."""
    FLUID_FLUID, FLUID_WALL, INTRA_MOLEC, BENDING, CONSTANT_FORCE, CHECK_FLUID_FLUID = 1, 2, 3, 4, 5, 6

    def __init__(self, p, omp: bool = False, seed: int = 4711):
        path = os.path.join(REF_DIR, "libmfa_ref_omp.so" if omp else "libmfa_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("reference build missing: make -f oracle/ref.mk (%s)" % path)
        lib = C.CDLL(path)
        vp, i, d = C.c_void_p, C.c_int, C.c_double
        lib.ref_setup.argtypes = [i, i, i, i, vp, vp, vp, vp, vp, vp, d, d, d, vp, vp, d, d, d, d, d, d, vp, vp, i]
        lib.ref_set_state.argtypes = [vp, vp, vp, vp]
        lib.ref_get_state.argtypes = [vp, vp, vp, vp]
        lib.ref_call.argtypes = [i, vp, vp]
        lib.ref_md_steps.argtypes = [i, i, vp]
        lib.ref_num_threads.restype = i
        self.lib, self.p = lib, p
        a = lambda x: np.ascontiguousarray(x, np.float64)
        self._keep = [a(p.boundary), a(p.epsil), a(p.sigma_2), a(p.range_2), a(p.e_shift), a(p.inv_range_2),
                      a(p.sigma_wall), a(p.a_wall), a(p.const_force), a(p.mass_type)]
        k = self._keep
        lib.ref_setup(p.n_mon, p.n_chain, p.n_mon_d, p.n_chain_d, _p(k[0]), _p(k[1]), _p(k[2]), _p(k[3]),
                      _p(k[4]), _p(k[5]), p.r_cut_max, p.sig, p.friction, _p(k[6]), _p(k[7]), p.a_w,
                      p.k_chain, p.inv_r_chain_2, p.k_bend, p.alpha_eq, p.dt, _p(k[8]), _p(k[9]), seed)

    def threads(self) -> int:
        return self.lib.ref_num_threads()

    def set_state(self, r0, v, a, a_type):
        self._st = [np.ascontiguousarray(r0, np.float64), np.ascontiguousarray(v, np.float64),
                    np.ascontiguousarray(a, np.float64), np.ascontiguousarray(a_type, np.int32)]
        self.lib.ref_set_state(*[_p(x) for x in self._st])

    def get_state(self):
        n = self.p.n_part
        out = [np.zeros((n, 3)) for _ in range(4)]
        self.lib.ref_get_state(*[_p(x) for x in out])
        return dict(zip(("r0", "v", "a", "force"), out))

    def call(self, which):
        f = np.zeros((self.p.n_part, 3)); e = C.c_double()
        self.lib.ref_call(which, _p(f), C.byref(e))
        return f, e.value

    def md_steps(self, n, t0=0):
        en = np.zeros(6)
        self.lib.ref_md_steps(n, t0, _p(en))
        return dict(zip(("v_fluid_fluid", "v_intra_molec", "v_fluid_wall", "v_bend", "v_bend_melt", "t_fluid"), en))
