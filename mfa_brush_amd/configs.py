"""Synthetic channel configurations and the named workloads.

The reference builds initial states with gen_brush.f90 (grafted chains on a
square lattice, mode 5 ordered, :362-441) and gen_droplet.f90 (melt/solvent
random walks, :18-69) and relaxes overlaps with the force switch-on ramp.
Here the initial state is overlap-free by construction: every bead sits on a
column site (columns on a square lattice, sites 0.97 sigma apart in z, close
to the FENE+LJ bond length), grafted chains are straight columns from their
heads at z=1.2 (bottom wall) and z=Lz-1.2 (top wall, lattice offset by half a
cell so the two brushes interleave), and solvent beads take randomly chosen
free column sites.  Velocities are Maxwell-Boltzmann at kT, heads at rest.

Workloads (sizes as stated, seeds fixed):
  * channel_c1:  box 16x16x20, 2x64 chains of 16 on 8x8 lattices (a=2),
    2048 solvent -> 4096 beads; LJ rc=2.5 (brush-brush, solvent-solvent,
    head-solvent), brush-solvent WCA, DPD gamma=4.5 rc_DPD=1, kT=1, dt=0.005.
  * channel_c2:  box 144.25x144.25x60 (102 sqrt2 square), 2x102^2 chains of
    32 (a=sqrt2, grafting density 0.5), 332,928 solvent -> 998,784 beads,
    same potentials.
  * melt_c3:     1.2M melt beads in 18,750 chains of 64, box 216.93^2 x 30,
    harmonic-angle bending k=5, Poiseuille g=0.01 x, DPD.
"""
from __future__ import annotations

import math
import warnings
from dataclasses import replace
from typing import Optional

import numpy as np

from .params import Flags, SimParams, TWO_SIXTH, f2d

BOND_Z = 0.97


def _tables(rc: np.ndarray, eps: np.ndarray, sig: np.ndarray):
    """Fortran-ordered tables from per-type-pair cutoffs (rc=0: off),
    epsilon and sigma, with the shift of init_params.f90:726-737."""
    range_2 = rc ** 2
    sigma_2 = sig ** 2
    epsil = 4.0 * eps
    e_shift = np.zeros((4, 4))
    for i in range(4):
        for j in range(4):
            if range_2[i, j] > 0.0:
                rd = (sigma_2[i, j] / range_2[i, j]) ** 3
                e_shift[i, j] = rd * (rd - 1.0)
    inv = np.where(range_2 > 0, 1.0 / np.where(range_2 > 0, range_2, 1.0), 0.0)
    col = lambda m: np.ascontiguousarray(m.T.reshape(-1))
    return col(epsil), col(sigma_2), col(range_2), col(e_shift), col(inv), float(math.sqrt(range_2.max()))


def channel_params(boundary, n_mon, n_chain, n_solvent, *, rc_lj=2.5, gamma=4.5, kT=1.0, dt=0.005,
                   rc_dpd: Optional[float] = 1.0, thermostat=True, a_w=1.0, sigma_w=1.0, seed=0x5EED0001,
                   eps_bs=1.0, skin=0.0, l_box=1, bending=False, k_bend=0.0, const_force=(0.0, 0.0, 0.0),
                   force_switch_on=False) -> SimParams:
    """Brush-in-channel parameters: type 1 heads, 2 brush, 3 solvent, 4 off.
    LJ cutoff rc_lj for brush-brush, solvent-solvent, head-brush and
    head-solvent; brush-solvent WCA (float 2^(1/6)) with epsilon eps_bs;
    head-head off (init_params.f90:719)."""
    rc = np.zeros((4, 4))
    eps = np.ones((4, 4))
    sig = np.ones((4, 4))
    wca = TWO_SIXTH
    for (a, b, r) in ((0, 1, rc_lj), (1, 1, rc_lj), (2, 2, rc_lj), (0, 2, rc_lj), (1, 2, wca)):
        rc[a, b] = rc[b, a] = r
    eps[1, 2] = eps[2, 1] = eps_bs
    eps[0, 0] = 0.0
    epsil, sigma_2, range_2, e_shift, inv, rcmax = _tables(rc, eps, sig)
    fl = Flags(solvent=3, thermostat=thermostat, dpd_cut_off=rc_dpd, force_switch_on=force_switch_on,
               poiseuille=any(c != 0.0 for c in const_force), bending=bending, l_box=l_box)
    r_cut_dpd_2 = 0.0
    if rc_dpd is not None:
        r = f2d(rc_dpd)
        r_cut_dpd_2 = r * r
        rcmax = max(rcmax, r)
    mass = np.array([1.0, 1.0, 1.0, 1.0])
    return SimParams(
        n_mon=n_mon, n_chain=n_chain, n_mon_d=1, n_chain_d=n_solvent, boundary=tuple(boundary),
        epsil=epsil, sigma_2=sigma_2, range_2=range_2, e_shift=e_shift, inv_range_2=inv,
        r_cut_max=rcmax, r_cut_dpd_2=r_cut_dpd_2, sig=math.sqrt(2.0 * kT * gamma / dt), friction=gamma,
        r_2_min_init=(f2d(0.9) * 1.0) ** 2, r_2_min_time=40.0,
        sigma_wall=np.array([(sigma_w + 1.0) / 2] * 4), a_wall=np.array([a_w] * 4), a_w=a_w, sigma_w=sigma_w,
        k_chain=30.0 * 1.5 ** 2, inv_r_chain_2=1.0 / 1.5 ** 2, k_bend=k_bend, alpha_eq=0.0, dt=dt,
        mass_type=mass, const_force=tuple(const_force), seed=seed, flags=fl, skin=skin, temp=kT)


def _columns(L, a, offset_frac, n_side):
    g = (np.arange(n_side) + offset_frac) * a
    xx, yy = np.meshgrid(g, g, indexing="ij")
    return np.stack([xx.ravel(), yy.ravel()], 1) % L


def build_channel(p: SimParams, a_graft: float, n_side: int, rng_seed: int = 12345, kT: float = 1.0):
    """Initial state (a_type, r0, v) for a two-brush channel in reference
    bead order: bottom chains, top chains (heads first), then solvent."""
    rng = np.random.default_rng(rng_seed)
    L, Ly, Lz = p.boundary
    assert abs(L - Ly) < 1e-9
    n_chain_half = n_side * n_side
    assert p.n_chain == 2 * n_chain_half
    bot = _columns(L, a_graft, 0.5, n_side)
    top = _columns(L, a_graft, 0.0, n_side)
    cols = [bot, top]
    if a_graft / 2 >= 0.95:   # extra solvent columns between brush columns
        cols.append((bot + np.array([a_graft / 2, 0.0])) % L)
        cols.append((bot + np.array([0.0, a_graft / 2])) % L)
    z_lo, z_hi = 1.2, Lz - 1.2
    n_site = int(math.floor((z_hi - z_lo) / BOND_Z)) + 1
    zs_up = z_lo + BOND_Z * np.arange(n_site)
    zs_dn = z_hi - BOND_Z * np.arange(n_site)
    n_mon = p.n_mon
    r0 = np.zeros((p.n_part, 3))
    a_type = np.zeros(p.n_part, np.int32)
    k = 0
    for side, (xy, zs) in enumerate(((bot, zs_up), (top, zs_dn))):
        for c in range(n_chain_half):
            r0[k:k + n_mon, 0] = xy[c, 0]
            r0[k:k + n_mon, 1] = xy[c, 1]
            r0[k:k + n_mon, 2] = zs[:n_mon]
            a_type[k] = 1
            a_type[k + 1:k + n_mon] = 2
            k += n_mon
    # free column sites (bottom-brush columns above the chains, top-brush
    # columns below them, solvent-only columns whole)
    site_runs = []
    for ci, xy in enumerate(cols):
        zs = zs_up if ci != 1 else zs_dn
        start = n_mon if ci < 2 else 0
        zz = zs[start:]
        zz = zz[(zz > 0.6) & (zz < Lz - 0.6)]
        for c in range(len(xy)):
            site_runs.append(np.stack([np.full(len(zz), xy[c, 0]), np.full(len(zz), xy[c, 1]), zz], 1))
    n_sol = p.n_chain_d * p.n_mon_d
    if p.n_mon_d == 1:
        free = np.concatenate(site_runs)
        if n_sol > len(free):
            raise ValueError("not enough free sites for %d solvent beads (have %d)" % (n_sol, len(free)))
        pick = rng.choice(len(free), size=n_sol, replace=False)
        r0[k:k + n_sol] = free[np.sort(pick)]
    else:
        # melt chains: runs of n_mon_d consecutive sites of one column
        runs = [(ri, q) for ri, run in enumerate(site_runs) for q in range(len(run) // p.n_mon_d)]
        if p.n_chain_d > len(runs):
            raise ValueError("not enough column runs for %d melt chains" % p.n_chain_d)
        pick = np.sort(rng.choice(len(runs), size=p.n_chain_d, replace=False))
        for c, idx in enumerate(pick):
            ri, q = runs[idx]
            r0[k + c * p.n_mon_d:k + (c + 1) * p.n_mon_d] = site_runs[ri][q * p.n_mon_d:(q + 1) * p.n_mon_d]
    a_type[k:k + n_sol] = 3
    # small jitter (keeps the minimum distance above 0.9) and wrap
    r0[:, :2] += rng.uniform(-0.02, 0.02, size=(p.n_part, 2))
    r0[:, 0] %= L
    r0[:, 1] %= Ly
    v = rng.normal(0.0, math.sqrt(kT), size=(p.n_part, 3))
    heads = np.arange(p.n_chain) * n_mon
    v[heads] = 0.0
    return a_type, r0, v


def channel_c1(**kw):
    p = channel_params((16.0, 16.0, 20.0), 16, 128, 2048, **kw)
    a_type, r0, v = build_channel(p, 2.0, 8)
    return p, a_type, r0, v


def channel_c2(**kw):
    a = math.sqrt(2.0)
    L = 102 * a
    p = channel_params((L, L, 60.0), 32, 2 * 102 * 102, 332928, **kw)
    a_type, r0, v = build_channel(p, a, 102)
    return p, a_type, r0, v


def channel_scaled(n_side: int, n_mon: int = 32, lz: float = 60.0, density: float = 0.8, **kw):
    """C2 geometry with n_side^2 chains per wall (C2 is n_side=102)."""
    a = math.sqrt(2.0)
    L = n_side * a
    n_brush = 2 * n_side * n_side * n_mon
    n_sol = int(round(density * L * L * lz)) - n_brush
    p = channel_params((L, L, lz), n_mon, 2 * n_side * n_side, n_sol, **kw)
    a_type, r0, v = build_channel(p, a, n_side)
    return p, a_type, r0, v


def _free_sites(L: float, lz: float, occupied: np.ndarray, spacing: float = 1.0, min_dist: float = 0.9,
                z_margin: float = 0.8) -> np.ndarray:
    """Sites of a cubic lattice (about `spacing` apart, fitted to the box)
    farther than min_dist (minimum image in x, y) from every occupied bead."""
    from scipy.spatial import cKDTree
    nxy = max(1, int(math.floor(L / spacing)))
    sx = L / nxy
    nz = max(1, int(math.floor((lz - 2 * z_margin) / spacing)) + 1)
    sz = (lz - 2 * z_margin) / max(1, nz - 1)
    g = (np.arange(nxy) + 0.5) * sx
    xx, yy, zz = np.meshgrid(g, g, z_margin + sz * np.arange(nz), indexing="ij")
    sites = np.stack([xx.ravel(), yy.ravel(), zz.ravel()], 1)
    if len(occupied) == 0:
        return sites
    box = [L, L, 4.0 * lz]          # periodic in x, y; z never wraps
    occ = np.c_[occupied[:, 0] % L, occupied[:, 1] % L, occupied[:, 2] + lz]
    tree = cKDTree(occ, boxsize=box)
    d, _ = tree.query(np.c_[sites[:, 0] % L, sites[:, 1] % L, sites[:, 2] + lz], k=1)
    return sites[d >= min_dist]


class _Hash:
    """Spatial hash of placed beads (periodic in x, y) for overlap checks."""

    def __init__(self, L, cell=1.0):
        self.L, self.n = L, max(1, int(L // cell))
        self.c = L / self.n
        self.d = {}

    def _key(self, r):
        return (int(r[0] // self.c) % self.n, int(r[1] // self.c) % self.n, int(math.floor(r[2] / self.c)))

    def add(self, r):
        self.d.setdefault(self._key(r), []).append(np.array(r, float))

    def nearest(self, r, skip=None, reach=1.0):
        """Distance to the nearest placed bead within `reach` (else reach)."""
        kx, ky, kz = self._key(r)
        best = reach * reach
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in self.d.get(((kx + dx) % self.n, (ky + dy) % self.n, kz + dz), ()):
                        if skip is not None and q is skip:
                            continue
                        e = r - q
                        e[:2] -= self.L * np.round(e[:2] / self.L)
                        best = min(best, e @ e)
        return math.sqrt(best)

    def clear_of(self, r, dmin, skip=None):
        kx, ky, kz = self._key(r)
        for dx in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dz in (-1, 0, 1):
                    for q in self.d.get(((kx + dx) % self.n, (ky + dy) % self.n, kz + dz), ()):
                        if skip is not None and q is skip:
                            continue
                        e = r - q
                        e[:2] -= self.L * np.round(e[:2] / self.L)
                        if e @ e < dmin * dmin:
                            return False
        return True


def _insert_solvent(L, lz, occupied, n_sol, rng, weight=None, dmin=0.85, z_margin=0.8, max_tries=400):
    """Random sequential insertion of n_sol beads in [0,L)^2 x [z_margin,
    lz - z_margin], each at least dmin (minimum image in x, y) from the
    occupied beads and from the beads inserted before it.  Heights are
    uniform, or drawn from weight(z) (inverse CDF on a 0.05 sigma grid).

    Candidates come in batches of 4096 and are taken in order: a candidate
    is kept when it is clear of the occupied beads, of the beads kept from
    earlier batches (k-d trees) and of the candidates kept before it in its
    own batch (only the few close pairs inside a batch are walked in order),
    which is the sequential insertion rule."""
    from scipy.spatial import cKDTree
    zs = np.arange(z_margin, lz - z_margin, 0.05)
    cdf = None
    if weight is not None:
        w = np.maximum(np.asarray([weight(z) for z in zs], float), 1e-9)
        cdf = np.cumsum(w)
        cdf /= cdf[-1]
    box = [L, L, 4 * lz]                 # periodic in x, y; z shifted by lz never wraps

    def lift(q):
        return np.c_[q[:, 0] % L, q[:, 1] % L, q[:, 2] + lz]

    tree = cKDTree(lift(occupied), boxsize=box) if len(occupied) else None
    out = np.zeros((n_sol, 3))
    k, tries = 0, 0
    kept_tree, kept_n = None, 0
    while k < n_sol:
        if tries > max_tries * max(n_sol, 1) // 4096 + 50:
            raise ValueError("solvent insertion stalled at %d of %d beads" % (k, n_sol))
        tries += 1
        m = 4096
        cand = np.c_[rng.uniform(0, L, m), rng.uniform(0, L, m),
                     rng.uniform(z_margin, lz - z_margin, m) if cdf is None else np.interp(rng.uniform(0, 1, m), cdf, zs)]
        if tree is not None:
            d, _ = tree.query(lift(cand), k=1)
            cand = cand[d >= dmin]
        if kept_n != k:
            kept_tree, kept_n = cKDTree(lift(out[:k]), boxsize=box), k
        if kept_tree is not None and len(cand):
            d, _ = kept_tree.query(lift(cand), k=1)
            cand = cand[d >= dmin]
        if not len(cand):
            continue
        # conflicts inside the batch: candidate j loses to an earlier kept i
        pairs = cKDTree(lift(cand), boxsize=box).query_pairs(dmin, output_type="ndarray")
        earlier = [[] for _ in range(len(cand))]
        for i, j in pairs:
            earlier[max(i, j)].append(min(i, j))
        keep = np.zeros(len(cand), bool)
        for j in range(len(cand)):
            keep[j] = not any(keep[i] for i in earlier[j])
        take = cand[keep][:n_sol - k]
        out[k:k + len(take)] = take
        k += len(take)
    return out


def channel_general(n_side: int, a_graft: float, n_mon: int, lz: float, density: float = 0.8,
                    brush_z=None, solvent_weight=None, seed: int = 1, **kw):
    """Two-brush channel of n_side^2 chains per wall at grafting spacing
    a_graft (grafting density 1/a_graft^2), chains of n_mon beads, solvent to
    the given total density.

    Plain start: straight chains at bond 0.97 along z (heads at z = 1.2 and
    lz - 1.2, the top lattice shifted by a_graft/2), solvent inserted at
    uniform random positions at least 0.85 sigma from every bead.  Warm
    start: brush_z(m) gives the height above its wall of monomer m (0 =
    head) and solvent_weight(z) the relative density of solvent at height z;
    monomers that rise less than a bond step sideways (random direction clear
    of placed beads) so the bond stays 0.97."""
    rng = np.random.default_rng(seed)
    L = n_side * a_graft
    n_chain = 2 * n_side * n_side
    n_brush = n_chain * n_mon
    n_sol = int(round(density * L * L * lz)) - n_brush
    if n_sol < 0:
        raise ValueError("brush alone exceeds the requested density")
    p = channel_params((L, L, lz), n_mon, n_chain, n_sol, **kw)
    g = (np.arange(n_side) + 0.5) * a_graft
    gx, gy = np.meshgrid(g, g, indexing="ij")
    base = np.stack([gx.ravel(), gy.ravel()], 1)
    r0 = np.zeros((p.n_part, 3))
    a_type = np.zeros(p.n_part, np.int32)
    k = 0
    crowded = 0
    hsh = _Hash(L) if brush_z is not None else None
    for side in (0, 1):
        heads = base if side == 0 else (base + a_graft / 2) % L
        for c in range(len(heads)):
            x, y = heads[c]
            zprev = 1.2
            prev = None
            for m in range(n_mon):
                h = 1.2 + (BOND_Z * m if brush_z is None else float(brush_z(m)))
                if m > 0 and brush_z is not None:
                    # rise towards the predicted height; the rest of the bond
                    # goes sideways, in a direction clear of placed beads
                    dz = min(max(h - zprev, 0.0), BOND_Z)
                    best, best_d = None, -1.0
                    for attempt in range(160):
                        # past 80 tries the rise grows towards a straight bond
                        dzt = dz if attempt < 80 else dz + (BOND_Z - dz) * (attempt - 79) / 80.0
                        lat = math.sqrt(max(BOND_Z * BOND_Z - dzt * dzt, 0.0))
                        ang = rng.uniform(0, 2 * math.pi)
                        cand = np.array([(x + lat * math.cos(ang)) % L, (y + lat * math.sin(ang)) % L, zprev + dzt])
                        zw = cand[2] if side == 0 else lz - cand[2]
                        dn = hsh.nearest(np.array([cand[0], cand[1], zw]), skip=prev)
                        if dn > best_d:
                            best, best_d = cand, dn
                        if dn >= 0.85:
                            break
                    if best_d < 0.85:
                        crowded += 1    # the least crowded of the tries
                    x, y, h = best
                elif m > 0:
                    h = zprev + BOND_Z
                zprev = h
                r0[k] = (x, y, h if side == 0 else lz - h)
                if hsh is not None:
                    hsh.add(r0[k])
                    prev = hsh.d[hsh._key(r0[k])][-1]
                a_type[k] = 1 if m == 0 else 2
                k += 1
    if crowded:
        warnings.warn("warm start: %d monomers closer than 0.85 sigma to another bead" % crowded)
    r0[k:] = _insert_solvent(L, lz, r0[:k], n_sol, rng, solvent_weight)
    a_type[k:] = 3
    v = rng.normal(0.0, 1.0, size=(p.n_part, 3))
    v[np.arange(n_chain) * n_mon] = 0.0
    return p, a_type, r0, v


def channel_melt(n_cols: int = 200, n_z: int = 30, n_mon: int = 64, lz: float = 30.0, density: float = 0.85,
                 k_bend: float = 5.0, g: float = 0.01, seed: int = 3, bend_cosine: bool = True, **kw):
    """Melt in a channel (no brush): n_cols^2 columns of n_z sites fill the
    box at `density`; chains of n_mon beads follow a serpentine path through
    the columns (up one column, down the next, rows of columns in
    alternating x direction), so every bond is either the in-column site
    spacing or the column spacing.  FENE + bending of the melt chains
    (bending_melt, k_bend): the cosine form k_bend (1 - cos theta) by
    default, the reference's harmonic angle with bend_cosine=False; a
    Poiseuille body force g along x on every bead.  The default is the
    1.2M-bead melt (18,750 chains of 64, L = 216.93, Lz = 30)."""
    n_sites = n_cols * n_cols * n_z
    if n_sites % n_mon:
        raise ValueError("columns x sites must be a multiple of the chain length")
    L = math.sqrt(n_sites / (density * lz))
    ac = L / n_cols
    zs = np.linspace(0.7, lz - 0.7, n_z)
    if min(ac, zs[1] - zs[0]) < 0.9 or max(ac, zs[1] - zs[0]) > 1.4:
        raise ValueError("site spacing out of range (%.3f, %.3f)" % (ac, zs[1] - zs[0]))
    rng = np.random.default_rng(seed)
    p = channel_params((L, L, lz), 0, 0, n_sites // n_mon, bending=False, k_bend=k_bend,
                       const_force=(g, 0.0, 0.0), **kw)
    p = p.replace(n_mon_d=n_mon, n_chain_d=n_sites // n_mon,
                  flags=Flags(**{**p.flags.__dict__, "bending_melt": k_bend > 0.0, "poiseuille": g != 0.0,
                                 "bend_cosine": bend_cosine}))
    r0 = np.zeros((n_sites, 3))
    k = 0
    col = 0
    for iy in range(n_cols):
        xs = range(n_cols) if iy % 2 == 0 else range(n_cols - 1, -1, -1)
        for ix in xs:
            zz = zs if col % 2 == 0 else zs[::-1]
            r0[k:k + n_z, 0] = (ix + 0.5) * ac
            r0[k:k + n_z, 1] = (iy + 0.5) * ac
            r0[k:k + n_z, 2] = zz
            k += n_z
            col += 1
    a_type = np.full(n_sites, 3, np.int32)
    v = rng.normal(0.0, 1.0, size=(n_sites, 3))
    return p, a_type, r0, v


def c4_replica(i: int):
    """Replica i of the batched workload: size log-uniform in [2k, 40k]
    beads, grafting density U[0.05, 0.5], N in {8, 16, 32, 64}, brush-solvent
    epsilon U[0.5, 1.5], DPD rc 1; per-replica seeds.  Lz is 30, raised
    where straight chains from both walls would not fit (N 32 and 64) or the
    two brushes alone would exceed 5/8 of the beads at density 0.8.  The
    crowded starts take the force switch-on ramp over the first 1 tau."""
    rng = np.random.default_rng(1000 + i)
    n_target = math.exp(rng.uniform(math.log(2000), math.log(40000)))
    sg = rng.uniform(0.05, 0.5)
    n_mon = int(rng.choice([8, 16, 32, 64]))
    eps = rng.uniform(0.5, 1.5)
    lz = max(30.0, 2.0 * sg * n_mon / 0.5, 2.0 * (1.2 + 0.97 * (n_mon - 1)) + 2.0)
    a = 1.0 / math.sqrt(sg)
    side = max(9.0, math.sqrt(n_target / (0.8 * lz)))
    n_side = max(2, int(round(side / a)))
    while n_side * a < 9.0:
        n_side += 1
    p, t, r0, v = channel_general(n_side, a, n_mon, lz, eps_bs=eps, thermostat=True,
                                  rc_dpd=1.0, seed=0x5EED0001 + i, force_switch_on=True)
    return p.replace(r_2_min_init=0.64, r_2_min_time=1.0), t, r0, v


def c4_batch(n_rep: int = 256, workers: int = 16):
    """The first n_rep replicas of the batched workload, built in parallel."""
    import os
    from concurrent.futures import ProcessPoolExecutor
    with ProcessPoolExecutor(max_workers=max(1, min(workers, os.cpu_count() or 1))) as ex:
        return list(ex.map(c4_replica, range(n_rep)))
