"""Initial configurations as the reference generates them (host side).

Restatements of gen_brush.f90 (modes 1-6), gen_droplet.f90 case 1 and the
labelling and velocity set-up of conf_default.f90, on numpy's generator
instead of the reference's R250 / ziggurat streams (the reference's own
sequences are compiler-dependent; the geometry and the distributions are
what these functions keep).  Arrays are (n_part, 3) in the reference bead
order: n_chain grafted chains of n_mon beads (head first; the first
n_chain / 2 chains on the top wall in the channel modes), then n_chain_d
melt / solvent chains of n_mon_d beads.
"""
from __future__ import annotations

import math

import numpy as np

from .params import SimParams

SQRT3 = 1.73205080756      # mfa_common.f90:315, the FP64 parameter


def z_head_default(sigma_11: float = 1.0, explicit_wall: bool = False) -> float:
    """Height of a grafted head above its wall, conf_default.f90:19-24."""
    return 2.0 ** (1.0 / 6.0) if explicit_wall else 1.2 * sigma_11


def z_step(solvent: int) -> float:
    """Monomer z step of the brush generators: 0.2 for SOLVENT 1/2, else 0.5
    (gen_brush.f90:58-70)."""
    return 0.2 if solvent in (1, 2) else 0.5


def _wrap_xy(x, L):
    """The reference's single wrap of a random-walk step (gen_brush.f90:50-55):
    >= L subtracts L, <= 0 adds L."""
    if x >= L:
        return x - L
    if x <= 0.0:
        return x + L
    return x


def gen_brush(p: SimParams, mode: int, rng: np.random.Generator, z_head: float = 1.2, r_chain: float = 1.5,
              solvent: int = 3, sigma_11: float = 1.0) -> np.ndarray:
    """Grafted chains, gen_brush.f90:
      1  channel: heads at random on both walls (first n_chain/2 on the top
         wall, z = Lz - z_head; the rest at z = z_head), chains as walks with
         x, y steps (2u-1) r_chain/sqrt3 and z steps -/+ 0.5 (0.2 for SOLVENT
         1, 2); a monomer outside 0 < z < Lz is drawn again (:16-97);
      2  droplet: as 1, every head on the bottom wall (:99-150);
      3  channel without overlapping grafting points: a head is drawn again
         until its x-y distance^2 to the earlier heads of its wall exceeds
         sigma_11^2; z steps -/+ 0.4 (0.2), no z check (:152-232);
      4  droplet without overlapping grafting points (:234-296; its overlap
         loop starts at the chain index left by the previous call, taken
         here as the first chain);
      5  ordered channel brush: heads on the square lattice
         a_br = sqrt(2 Lx Ly / n_chain), n_col = nint(Lx / a_br) columns by
         n_row = nint(Ly / a_br) rows per wall, column by column, the top
         wall first; steps as in mode 1 without the z check (:298-361);
      6  ordered droplet brush on the bottom wall, a_br = sqrt(Lx Ly /
         n_chain) (:363-420).
    Returns the brush part of r0 (n_chain * n_mon, 3)."""
    n_mon, n_chain = p.n_mon, p.n_chain
    Lx, Ly, Lz = p.boundary
    L = (Lx, Ly)
    out = np.zeros((n_mon * n_chain, 3))
    if n_chain == 0 or n_mon == 0:
        return out
    step = 0.4 if mode == 3 and solvent not in (1, 2) else z_step(solvent)
    top_half = n_chain // 2

    def chain(i0, head, sign, check_z):
        out[i0] = head
        for m in range(1, n_mon):
            while True:
                u = rng.uniform(0.0, 1.0, 3)
                q = out[i0 + m - 1].copy()
                for d in range(2):
                    q[d] = _wrap_xy(q[d] + (2.0 * u[d] - 1.0) * r_chain / SQRT3, L[d])
                q[2] = q[2] + sign * step
                if not check_z or (0.0 < q[2] < Lz):
                    break
            out[i0 + m] = q

    if mode in (1, 2, 3, 4):
        heads = []
        for c in range(n_chain):
            on_top = mode in (1, 3) and c < top_half
            zh = Lz - z_head if on_top else z_head
            first_of_wall = 0 if (on_top or mode in (2, 4)) else top_half
            while True:
                xy = rng.uniform(0.0, 1.0, 2) * np.array(L)
                if mode in (3, 4):
                    clash = any((xy[0] - h[0]) ** 2 + (xy[1] - h[1]) ** 2 <= sigma_11 ** 2
                                for h in heads[first_of_wall:c])
                    if clash:
                        continue
                break
            heads.append(xy)
            chain(c * n_mon, (xy[0], xy[1], zh), -1.0 if on_top else 1.0, check_z=mode in (1, 2, 4))
        return out
    if mode in (5, 6):
        walls = 2 if mode == 5 else 1
        a_br = math.sqrt(walls * Lx * Ly / n_chain)
        n_col, n_row = int(round(Lx / a_br)), int(round(Ly / a_br))
        if n_chain != walls * n_col * n_row:
            raise ValueError("gen_brush(%d): n_chain %d must equal %d x n_col x n_row = %d"
                             % (mode, n_chain, walls, walls * n_col * n_row))
        c = 0
        for w in range(walls):
            on_top = mode == 5 and w == 0
            for i_col in range(n_col):
                for i_row in range(n_row):
                    head = (a_br / 2 + a_br * i_col, a_br / 2 + a_br * i_row, Lz - z_head if on_top else z_head)
                    chain(c * n_mon, head, -1.0 if on_top else 1.0, check_z=mode == 6)
                    c += 1
        return out
    raise ValueError("gen_brush: mode must be 1..6")


def gen_droplet(p: SimParams, rng: np.random.Generator, z_skin: float = 1.0, r_chain: float = 1.5,
                max_tries: int = 1000) -> np.ndarray:
    """Melt / solvent chains, gen_droplet.f90 case 1 (:18-69): the first bead
    of a chain uniform in the box, moved by z_skin away from a wall it lies
    within z_skin of; each next bead 0.95 r_chain away in the direction of
    (2u - 1) / sqrt3, drawn again (at most max_tries times) while its z is
    <= z_skin or > Lz - z_skin; single periodic wrap in every direction.
    Returns the melt part of r0 (n_chain_d * n_mon_d, 3)."""
    n, nc = p.n_mon_d, p.n_chain_d
    B = np.array(p.boundary, float)
    out = np.zeros((n * nc, 3))
    for c in range(nc):
        i = c * n
        q = rng.uniform(0.0, 1.0, 3) * B
        if q[2] <= z_skin:
            q[2] += z_skin
        if q[2] >= B[2] - z_skin:
            q[2] -= z_skin
        out[i] = q
        for m in range(1, n):
            for t in range(max_tries + 1):
                if t == max_tries:
                    raise ValueError("gen_droplet: no configuration after %d tries" % max_tries)
                d = (2.0 * rng.uniform(0.0, 1.0, 3) - 1.0) / SQRT3
                d = 0.95 * r_chain * d / math.sqrt(d.dot(d))
                q = out[i + m - 1] + d
                if not (q[2] <= z_skin or q[2] > B[2] - z_skin):
                    break
            for k in range(3):
                if q[k] >= B[k]:
                    q[k] -= B[k]
                elif q[k] <= 0.0:
                    q[k] += B[k]
            out[i + m] = q
    return out


def labels(p: SimParams) -> np.ndarray:
    """a_type of conf_default.f90:40-66 (SYSTEM 0/1 without PARTICLE_4):
    1 for the first bead of a grafted chain, 2 for its other beads, 3 for
    the melt / solvent."""
    t = np.full(p.n_part, 3, np.int32)
    n_brush = p.n_mon * p.n_chain
    t[:n_brush] = 2
    if p.n_chain:
        t[0:n_brush:p.n_mon] = 1
    return t


def conf_default(p: SimParams, brush_mode: int, seed: int = 1, z_skin: float = 1.0, r_chain: float = 1.5,
                 solvent: int = 3, z_head: float | None = None):
    """conf_default.f90 for SYSTEM 0 (channel) / 1 (droplet) with implicit
    walls: labels, gen_brush(brush_mode), gen_droplet(1), Maxwell-Boltzmann
    velocities sqrt(T / m) N(0, 1) (:197-201) with the heads' velocities
    zeroed (:211-222).  Returns (a_type, r0, v)."""
    rng = np.random.default_rng(seed)
    zh = z_head_default() if z_head is None else z_head
    a_type = labels(p)
    r0 = np.zeros((p.n_part, 3))
    n_brush = p.n_mon * p.n_chain
    r0[:n_brush] = gen_brush(p, brush_mode, rng, z_head=zh, r_chain=r_chain, solvent=solvent)
    r0[n_brush:] = gen_droplet(p, rng, z_skin=z_skin, r_chain=r_chain)
    m = np.asarray(p.mass_type, float)[a_type - 1]
    v = np.sqrt(p.temp / m)[:, None] * rng.normal(0.0, 1.0, (p.n_part, 3))
    if p.n_chain:
        v[0:n_brush:p.n_mon] = 0.0
    return a_type, r0, v
