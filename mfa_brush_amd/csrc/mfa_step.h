// mfa_step.h -- per-bead kernels of the step: integrator halves, bonded
// forces, walls, body force, cell-list rebuild (counting sort), observables.
#pragma once
#include "mfa_kernels.h"
#include "mfa_pair.h"

namespace mfa {


// ---------------------------------------------------------------------------
// verlet_positions, verlet_positions.f90:150-209: r += dt v + 0.5 dt^2 a,
// single wrap in x,y, v += 0.5 dt a; plus the skin test of check_skin.f90:11-24
// (displacement since the last build > skin/2 -> rebuild).
// ---------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(BLK) void k_vv1(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B) {
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (threadIdx.x == 0 && P.skin_mode == 0) { *B.rebuild = 1; B.chan_rb[s] = 1; }
    // the fused step's k_bonded stores the total force only on the last step
    // of an md_steps call (the next pair kernel overwrites it); t_idx is
    // read here because k_bonded itself advances it
    if (blockIdx.x == 0 && threadIdx.x == 0) *B.store_frc = *B.t_idx >= *B.t_end ? 1 : 0;
    if (k - P.slot_base >= P.n_part) return;
    double4 r = CUR(B.pos)[k];
    double4 v = CUR(B.vel)[k];
    const double4 a = CUR(B.acc)[k];
    {
#pragma clang fp contract(off)
        const double h = 0.5 * P.dt_2;
        r.x = r.x + P.dt * v.x + h * a.x;
        r.y = r.y + P.dt * v.y + h * a.y;
        r.z = r.z + P.dt * v.z + h * a.z;
        if (r.x > P.L[0]) r.x = r.x - P.L[0]; else if (r.x < 0.0) r.x = r.x + P.L[0];
        if (r.y > P.L[1]) r.y = r.y - P.L[1]; else if (r.y < 0.0) r.y = r.y + P.L[1];
        const double hd = 0.5 * P.dt;
        v.x = v.x + hd * a.x;
        v.y = v.y + hd * a.y;
        v.z = v.z + hd * a.z;
    }
    CUR(B.pos)[k] = r;
    CUR(B.vel)[k] = v;
    if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) atomicAdd(&B.flags[0], 1);
    if (r.z < 0.0 || r.z > P.L[2]) atomicAdd(&B.flags[1], 1);
    if (P.skin_mode) {
        const double4 b = CUR(B.posb)[k];
        double dx = r.x - b.x, dy = r.y - b.y, dz = r.z - b.z;
        dx -= P.L[0] * rint(dx * P.invL[0]);
        dy -= P.L[1] * rint(dy * P.invL[1]);
        if (!(dx * dx + dy * dy + dz * dz <= P.skin_half_2)) { atomicOr(B.rebuild, 1); atomicOr(&B.chan_rb[s], 1); }
    }
    if (MIXED) {
        const int c = CUR(B.cell_of)[k] - P.cell_base;
        if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) return;
        const int nx = P.nc[0], ny = P.nc[1];
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        double rx = r.x - cx * P.l_cell[0], ry = r.y - cy * P.l_cell[1], rz = r.z - cz * P.l_cell[2];
        if (rx > 0.5 * P.L[0]) rx -= P.L[0]; else if (rx < -0.5 * P.L[0]) rx += P.L[0];
        if (ry > 0.5 * P.L[1]) ry -= P.L[1]; else if (ry < -0.5 * P.L[1]) ry += P.L[1];
        CUR(B.posf)[k] = make_float4((float)rx, (float)ry, (float)rz, mkwf(worig(r.w), wtype(r.w)));
        CUR(B.velf)[k] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
    }
}

// FP32 velocity copy refreshed from the FP64 velocities (mixed path): the
// routine-level pair evaluation may follow any routine, not only k_vv1.
__global__ __launch_bounds__(BLK) void k_velf_sync(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                                  Bufs B) {
    const int s = blk_sys[blockIdx.x];
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k - P.slot_base >= P.n_part) return;
    const double4 v = CUR(B.vel)[k];
    CUR(B.velf)[k] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
}

struct BendOut { double fp[3], fn[3], v; };

// Harmonic-angle bending centred on c, bending.f90:30-109 (including the
// small-angle fallback :59-92 and the cos >= 1 guard :93-98); with
// bend_form 1 the cosine form k_bend (1 - cos theta) over the same angle.
__device__ __forceinline__ BendOut bend_angle(const SysDev &P, const double4 &rm, const double4 &rc, const double4 &rpn) {
#pragma clang fp contract(off)
    BendOut o;
    double rp[3], rn[3];
    rp[0] = mic(rc.x - rm.x, P.L[0], P.invL[0]);
    rp[1] = mic(rc.y - rm.y, P.L[1], P.invL[1]);
    rp[2] = rc.z - rm.z;
    rn[0] = mic(rpn.x - rc.x, P.L[0], P.invL[0]);
    rn[1] = mic(rpn.y - rc.y, P.L[1], P.invL[1]);
    rn[2] = rpn.z - rc.z;
    double dot = 0.0, rp2 = 0.0, rn2 = 0.0;
    for (int d = 0; d < 3; ++d) { dot += rn[d] * rp[d]; rp2 += rp[d] * rp[d]; rn2 += rn[d] * rn[d]; }
    double cosa = dot / sqrt(rp2 * rn2);
    if (P.bend_form == 1) {
        // cosine form k (1 - cos theta): fp = k dcos/dr_prev, fn = k dcos/dr_next
        const double inv_pn = 1.0 / sqrt(rp2 * rn2);
        o.v = P.k_bend * (1.0 - cosa);
        for (int d = 0; d < 3; ++d) {
            o.fp[d] = P.k_bend * (cosa * rp[d] / rp2 - rn[d] * inv_pn);
            o.fn[d] = P.k_bend * (rp[d] * inv_pn - cosa * rn[d] / rn2);
        }
        return o;
    }
    double dn[3], dp[3], dn2 = 0.0, dp2 = 0.0;
    for (int d = 0; d < 3; ++d) { dn[d] = rp[d] * rn2 - rn[d] * dot; dp[d] = -rn[d] * rp2 + rp[d] * dot; }
    for (int d = 0; d < 3; ++d) { dp2 += dp[d] * dp[d]; dn2 += dn[d] * dn[d]; }
    const double small = (double)0.000001f;
    if (dn2 < small) {
        double s1 = sqrt(rp2), s2 = sqrt(rn2);
        for (int d = 0; d < 3; ++d) dn[d] = rp[d] / s1 - rn[d] / s2;
        double nrm = sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
        for (int d = 0; d < 3; ++d) dn[d] = nrm != 0.0 ? dn[d] / nrm / s2 : 0.0;
    } else {
        double sc = sqrt(dn2 * rn2);
        for (int d = 0; d < 3; ++d) dn[d] = dn[d] / sc;
    }
    if (dp2 < small) {
        double s1 = sqrt(rp2), s2 = sqrt(rn2);
        for (int d = 0; d < 3; ++d) dp[d] = rp[d] / s1 - rn[d] / s2;
        double nrm = sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
        for (int d = 0; d < 3; ++d) dp[d] = nrm != 0.0 ? dp[d] / nrm / s1 : 0.0;
    } else {
        double sc = sqrt(dp2 * rp2);
        for (int d = 0; d < 3; ++d) dp[d] = dp[d] / sc;
    }
    double dal = cosa >= 1.0 ? 0.0 : acos(cosa) - P.alpha_eq;
    double fm = P.k_bend * dal;
    o.v = 0.5 * fm * dal;
    for (int d = 0; d < 3; ++d) { o.fp[d] = fm * dp[d]; o.fn[d] = fm * dn[d]; }
    return o;
}

// M_NOBEND: a fused step of channels that all have bending and melt bending
// off; the bending code is compiled out (its registers cap the occupancy of
// the fused kernel), its energy partials are still written (zero).
// M_BENDF: the bending of the channels whose flags ask for it, added to the
// stored force (the fused step with the fused integrator, between the pair
// kernel and k_integrate).
enum Mode { M_WALL = 1, M_FENE = 2, M_BEND = 4, M_BENDM = 8, M_CONST = 16, M_VV2 = 32, M_FUSED = 64, M_NOBEND = 128,
            M_BENDF = 256 };

// Bending contributions to the bead at chain position m (0-based) of a chain
// of nm beads starting at global index g0, in the reference's loop order
// (centres ascending): +fn of centre m-1, -fp-fn of centre m, +fp of m+1.
__device__ __forceinline__ void bend_bead(const SysDev &P, const Bufs &B, size_t cur_off, int g0, int nm, int m,
                                          const double4 ri, double3 &F, double &ev) {
#pragma clang fp contract(off)
    // positions of the beads m-2 .. m+2 of the chain (absent ones are never used)
    double4 rm2 = ri, rm1 = ri, rp1 = ri, rp2 = ri;
    if (m >= 2) rm2 = bead_at(B, cur_off, g0 + m - 2);
    if (m >= 1) rm1 = bead_at(B, cur_off, g0 + m - 1);
    if (m + 1 < nm) rp1 = bead_at(B, cur_off, g0 + m + 1);
    if (m + 2 < nm) rp2 = bead_at(B, cur_off, g0 + m + 2);
    if (m - 1 >= 1 && m - 1 <= nm - 2) {
        BendOut o = bend_angle(P, rm2, rm1, ri);
        F.x += o.fn[0]; F.y += o.fn[1]; F.z += o.fn[2];
    }
    if (m >= 1 && m <= nm - 2) {
        BendOut o = bend_angle(P, rm1, ri, rp1);
        F.x = F.x - o.fp[0] - o.fn[0];
        F.y = F.y - o.fp[1] - o.fn[1];
        F.z = F.z - o.fp[2] - o.fn[2];
        ev += o.v;
    }
    if (m + 1 >= 1 && m + 1 <= nm - 2) {
        BendOut o = bend_angle(P, ri, rp1, rp2);
        F.x += o.fp[0]; F.y += o.fp[1]; F.z += o.fp[2];
    }
}

// Everything after the pair forces, per bead, in md_main.f90's order:
// fluid_wall(2) :132, intra_molec :142, bending(1) :144, bending_melt(1) :147,
// constant_force(1) :170, verlet_velocities :185.  MODE selects the terms
// (one term for the single-routine entry points, all for the fused step).
template <int MODE, bool MIXED>
__global__ __launch_bounds__(BLK) void k_bonded(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B,
                                               double *__restrict__ partial, double *__restrict__ f_heads, int vir) {
    __shared__ double red[BLK];
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    const bool active = (k - P.slot_base) < P.n_part;
    constexpr bool FUSED = (MODE & M_FUSED) != 0;
    // M_BENDF always stores (k_integrate reads the force), sums only on the last step
    const bool store_f = !FUSED || *B.store_frc != 0;
    const bool sums = (MODE & M_BENDF) ? *B.store_frc != 0 : store_f;
    double e_wall = 0.0, e_fene = 0.0, e_bend = 0.0, e_bendm = 0.0, ke = 0.0;
    double wb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (active) {
        // no contraction: the kernel's variants (with and without the
        // bending code) then round the walls, sums and kinetic energy alike,
        // as the oracle does
#pragma clang fp contract(off)
        const double4 ri = CUR(B.pos)[k];
        const int ti = wtype(ri.w);
        const int g = worig(ri.w);
        const int o = g - P.part_base;
        guard(B, ti >= 1 && ti <= 4, G_TYPE, k);
        guard(B, in_range(o, P.n_part), G_ORIG, k);
        double4 f4 = CUR(B.frc)[k];
        double3 F = make_double3(f4.x, f4.y, f4.z);
        const bool brush = o < P.pd;
        const int nm = brush ? P.n_mon : P.n_mon_d;
        const int ol = brush ? o : o - P.pd;
        const int m = nm > 0 ? ol % nm : 0;
        const int g0 = g - m;
        if (MODE & (M_WALL | M_FUSED)) wall_bead(P, ri, ti, F, e_wall);
        if (MODE & (M_FENE | M_FUSED)) {
            if (vir) fene_bead<true>(P, B, cur_off, ri, g, m, nm, F, e_fene, wb);
            else fene_bead(P, B, cur_off, ri, g, m, nm, F, e_fene);
        }
        constexpr bool FBEND = (FUSED && !(MODE & M_NOBEND)) || (MODE & M_BENDF) != 0;
        if ((MODE & M_BEND) || (FBEND && P.bending)) {
            if (brush) bend_bead(P, B, cur_off, g0, nm, m, ri, F, e_bend);
        }
        if ((MODE & M_BENDM) || (FBEND && P.bending_melt)) {
            if (!brush) bend_bead(P, B, cur_off, g0, nm, m, ri, F, e_bendm);
        }
        if ((MODE & M_CONST) || (FUSED && P.poiseuille)) {
            if (ti != 1) {
                const double ms = P.mass_type[(ti - 1) & 3];
                F.x += ms * P.const_force[0];
                F.y += ms * P.const_force[1];
                F.z += ms * P.const_force[2];
            }
        }
        if (MODE & (M_VV2 | M_FUSED)) {
            const bool head = brush && m == 0;
            if (head && guard(B, in_range(P.head_base + ol / P.n_mon, B.n_heads), G_HEAD, k)) {
                double *fh = f_heads + 3 * (size_t)(P.head_base + ol / P.n_mon);
                fh[0] = F.x; fh[1] = F.y; fh[2] = F.z;
            }
            double4 v = CUR(B.vel)[k];
            double4 a;
            const double im = P.inv_mass_type[(ti - 1) & 3];
            {
#pragma clang fp contract(off)
                a.x = F.x * im; a.y = F.y * im; a.z = F.z * im; a.w = 0.0;
                const double hd = 0.5 * P.dt;
                v.x = v.x + hd * a.x; v.y = v.y + hd * a.y; v.z = v.z + hd * a.z;
            }
            if (head) {
                if (P.poiseuille) { v.x = 0.0; v.y = 0.0; }
                else {
                    const double sg = (ol / P.n_mon) < P.n_chain / 2 ? 1.0 : -1.0;
                    v.x = sg * P.head_vel[0]; v.y = sg * P.head_vel[1];
                }
                v.z = 0.0;
                a.x = a.y = a.z = 0.0;
                F.x = F.y = F.z = 0.0;
            }
            ke = P.mass_type[(ti - 1) & 3] * (v.x * v.x + v.y * v.y + v.z * v.z);
            CUR(B.vel)[k] = v;
            CUR(B.acc)[k] = a;
            // (the fused step's next pair evaluation reads the FP32 copy that
            // k_vv1 writes after its half kick)
            if (MIXED && !FUSED) CUR(B.velf)[k] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
        }
        f4.x = F.x; f4.y = F.y; f4.z = F.z;
        if (store_f) CUR(B.frc)[k] = f4;
    }
    // the fused step sums its energies only on the step whose force it
    // stores (the last of an md_steps call), the only one a caller can read
    double *out = partial + (size_t)blockIdx.x * NPART;
    if (sums) {
        if (MODE & (M_WALL | M_FUSED)) { double t = block_sum(e_wall, red); if (threadIdx.x == 0) out[C_VWALL] = t; }
        if (MODE & (M_FENE | M_FUSED)) {
            double t = block_sum(e_fene, red);
            if (threadIdx.x == 0) out[C_VINTRA] = t;
            for (int q = 0; q < 6; ++q) {       // the FENE bond virial (zero when off)
                double bq = vir ? block_sum(wb[q], red) : 0.0;
                if (threadIdx.x == 0) out[C_WBXX + q] = bq;
            }
        }
        if ((MODE & (M_BEND | M_BENDF)) || FUSED) { double t = block_sum(e_bend, red); if (threadIdx.x == 0) out[C_VBEND] = t; }
        if ((MODE & (M_BENDM | M_BENDF)) || FUSED) { double t = block_sum(e_bendm, red); if (threadIdx.x == 0) out[C_VBENDM] = t; }
        if (MODE & (M_VV2 | M_FUSED)) {
            double t = block_sum(ke, red);
            if (threadIdx.x == 0) out[C_KE] = t;
        }
    }
    if ((MODE & (M_VV2 | M_FUSED)) && blockIdx.x == 0 && threadIdx.x == 0) *B.t_idx += 1;
}

// ---------------------------------------------------------------------------
// DPD-VV (new_dpd_fd.f90, md_main.f90:187-198): after the second half kick
// the dissipative forces are evaluated again at the new velocities and
// replace those of the half-step velocities in the force and acceleration
// carried into the next step.  The dissipative force is linear in the
// velocity differences, so F_D(v_new) - F_D(v_half) = F_D(v_new - v_half),
// and v_new - v_half is the kick just applied, dt/2 a (zero for the heads,
// whose velocities are fixed).  k_dpd_vv sums that difference per bead over
// the bead's DPD pairs (the pair set and weights of the pair kernel's DPD
// part: with DPD_CUT_OFF the non-head pairs within rc_dpd, else the pairs
// within the LJ range, with the switch-on clamp) into dfd; k_dpd_vv_apply
// adds it to the acceleration and the force.  `all`: every DPD channel (the
// routine-level call), else only the channels with dpd_vv set.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool dpd_vv_on(const SysDev &P, int all) {
    return P.thermostat == 1 && (all || P.dpd_vv);
}

__global__ __launch_bounds__(BLK) void k_dpd_vv(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B,
                                               const int2 *__restrict__ crange, double4 *__restrict__ dfd, int all) {
#pragma clang fp contract(off)
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k - P.slot_base >= P.n_part || !dpd_vv_on(P, all)) return;
    const double4 *__restrict__ pos = CUR(B.pos);
    const double4 *__restrict__ acc = CUR(B.acc);
    // the step's pair evaluation ran at the time index before the kick advanced it
    double r_2_min = 0.0;
    if (P.force_switch_on) {
        const double r_time = (double)(*B.t_idx - 1) * P.dt;
        if (r_time < P.r_2_min_time) r_2_min = (1.0 - r_time / P.r_2_min_time) * P.r_2_min_init;
    }
    const double hd = 0.5 * P.dt;
    const double4 pi = pos[k];
    const int ti = wtype(pi.w);
    const double4 ai = acc[k];
    const double dvix = hd * ai.x, dviy = hd * ai.y, dviz = hd * ai.z;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    auto visit = [&](int m) {
        const double4 pj = pos[m];
        const int tj = wtype(pj.w);
        const int tt = ((ti - 1) & 3) + 4 * ((tj - 1) & 3);
        const double dx = mic(pi.x - pj.x, P.L[0], P.invL[0]);
        const double dy = mic(pi.y - pj.y, P.L[1], P.invL[1]);
        const double dz = pi.z - pj.z;
        double r_2 = r2_exact(dx, dy, dz);
        const bool lj = r_2 < P.range_2[tt];
        const bool dpd = P.dpd_cut_off ? (ti != 1 && tj != 1 && r_2 < P.r_cut_dpd_2) : lj;
        if (!dpd) return;
        if (lj && P.force_switch_on && r_2 < r_2_min) r_2 = r_2_min;
        const double inv_sqrt = sqrt(1.0 / r_2);
        const double w = P.dpd_cut_off ? 1.0 - sqrt(r_2 / P.r_cut_dpd_2) : 1.0 - sqrt(r_2 * P.inv_range_2[tt]);
        const double4 aj = acc[m];
        const double rvx = dx * inv_sqrt, rvy = dy * inv_sqrt, rvz = dz * inv_sqrt;
        const double rdv = (dvix - hd * aj.x) * rvx + (dviy - hd * aj.y) * rvy + (dviz - hd * aj.z) * rvz;
        const double sd = -1.0 * P.friction * (w * w) * rdv;
        fx += sd * rvx; fy += sd * rvy; fz += sd * rvz;
    };
    const int nn = B.nl_cap > 0 ? B.nl_n[k] : -1;
    if (nn >= 0 && nn <= B.nl_cap) {
        for (int q = 0; q < nn; ++q) {
            const unsigned m = (unsigned)B.nl[nl_index(k, q, (size_t)B.n_slots)] >> NL_SLOT_SHIFT;
            if (guard(B, m < (unsigned)B.n_slots, G_SLOT, k)) visit((int)m);
        }
    } else {
        int c = CUR(B.cell_of)[k] - P.cell_base;
        if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) c = 0;
        const int nx = P.nc[0], ny = P.nc[1], nz = P.nc[2];
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        for (int dz = -P.S; dz <= P.S; ++dz) {
            const int z2 = cz + dz;
            if (z2 < 0 || z2 >= nz) continue;
            for (int dy = -P.S; dy <= P.S; ++dy) {
                int y2 = cy + dy;
                y2 += (y2 < 0) ? ny : 0;
                y2 -= (y2 >= ny) ? ny : 0;
                const int row = P.cell_base + nx * (y2 + ny * z2);
                for (int dx = -P.S; dx <= P.S; ++dx) {
                    int x2 = cx + dx;
                    x2 += (x2 < 0) ? nx : 0;
                    x2 -= (x2 >= nx) ? nx : 0;
                    int2 rg = crange[row + x2];
                    if (!guard(B, rg.x >= 0 && rg.x <= rg.y && rg.y <= B.n_slots, G_SLOT, k)) rg = make_int2(0, 0);
                    for (int m = rg.x; m < rg.y; ++m)
                        if (m != k) visit(m);
                }
            }
        }
    }
    dfd[k] = make_double4(fx, fy, fz, 0.0);
}

// fused: the fused step, whose force is stored only on a call's last step
__global__ __launch_bounds__(BLK) void k_dpd_vv_apply(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                                     Bufs B, const double4 *__restrict__ dfd, int all, int fused) {
#pragma clang fp contract(off)
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k - P.slot_base >= P.n_part || !dpd_vv_on(P, all)) return;
    const double4 ri = CUR(B.pos)[k];
    const int ti = wtype(ri.w);
    const int o = worig(ri.w) - P.part_base;
    if (o < P.pd && P.n_mon > 0 && o % P.n_mon == 0) return;     // heads keep force and acceleration zero
    const double4 d = dfd[k];
    const double im = P.inv_mass_type[(ti - 1) & 3];
    double4 a = CUR(B.acc)[k];
    a.x = a.x + d.x * im; a.y = a.y + d.y * im; a.z = a.z + d.z * im;
    CUR(B.acc)[k] = a;
    if (!fused || *B.store_frc) {
        double4 f = CUR(B.frc)[k];
        f.x = f.x + d.x; f.y = f.y + d.y; f.z = f.z + d.z;
        CUR(B.frc)[k] = f;
    }
}

// ---------------------------------------------------------------------------
// Fused integrator of the fused step: the second half kick of step n
// (verlet_velocities.f90:18-153, after constant_force(1)) and, unless step n
// is the last of the md_steps call, the first half of step n + 1
// (verlet_positions.f90:150-209 with the skin test of check_skin.f90:11-24).
// The force of step n -- pair, walls, bonds (and bending) -- is complete in
// frc when it runs; every bead touches only its own state, so positions are
// updated in place.  On the call's last step it stores the force and the
// acceleration and sums the kinetic energy (the pair kernel published that
// decision in store_frc); otherwise the acceleration lives only in registers
// between the two halves.
// ---------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(BLK) void k_integrate(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                                  Bufs B, double *__restrict__ partial, double *__restrict__ f_heads) {
    __shared__ double red[BLK];
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    const bool active = (k - P.slot_base) < P.n_part;
    const bool last = *B.store_frc != 0;
    if (!last && threadIdx.x == 0 && P.skin_mode == 0) { *B.rebuild = 1; B.chan_rb[s] = 1; }
    double ke = 0.0;
    if (active) {
#pragma clang fp contract(off)
        double4 r = CUR(B.pos)[k];
        const int ti = wtype(r.w);
        const int o = worig(r.w) - P.part_base;
        const bool brush = o < P.pd;
        const int ol = brush ? o : o - P.pd;
        const bool head = brush && P.n_mon > 0 && ol % P.n_mon == 0;
        double4 f4 = CUR(B.frc)[k];
        double3 F = make_double3(f4.x, f4.y, f4.z);
        if (P.poiseuille && ti != 1) {             // constant_force(1), md_main.f90:170
            const double ms = P.mass_type[(ti - 1) & 3];
            F.x += ms * P.const_force[0];
            F.y += ms * P.const_force[1];
            F.z += ms * P.const_force[2];
        }
        if (head && guard(B, in_range(P.head_base + ol / P.n_mon, B.n_heads), G_HEAD, k)) {
            double *fh = f_heads + 3 * (size_t)(P.head_base + ol / P.n_mon);
            fh[0] = F.x; fh[1] = F.y; fh[2] = F.z;
        }
        double4 v = CUR(B.vel)[k];
        double4 a;
        const double im = P.inv_mass_type[(ti - 1) & 3];
        const double hd = 0.5 * P.dt;
        a.x = F.x * im; a.y = F.y * im; a.z = F.z * im; a.w = 0.0;
        v.x = v.x + hd * a.x; v.y = v.y + hd * a.y; v.z = v.z + hd * a.z;
        if (head) {
            if (P.poiseuille) { v.x = 0.0; v.y = 0.0; }
            else {
                const double sg = (ol / P.n_mon) < P.n_chain / 2 ? 1.0 : -1.0;
                v.x = sg * P.head_vel[0]; v.y = sg * P.head_vel[1];
            }
            v.z = 0.0;
            a.x = a.y = a.z = 0.0;
            F.x = F.y = F.z = 0.0;
        }
        if (last) {
            ke = P.mass_type[(ti - 1) & 3] * (v.x * v.x + v.y * v.y + v.z * v.z);
            CUR(B.vel)[k] = v;
            CUR(B.acc)[k] = a;
            f4.x = F.x; f4.y = F.y; f4.z = F.z;
            CUR(B.frc)[k] = f4;
        } else {
            // the next step's first half (k_vv1)
            r.x = r.x + P.dt * v.x + (0.5 * P.dt_2) * a.x;
            r.y = r.y + P.dt * v.y + (0.5 * P.dt_2) * a.y;
            r.z = r.z + P.dt * v.z + (0.5 * P.dt_2) * a.z;
            if (r.x > P.L[0]) r.x = r.x - P.L[0]; else if (r.x < 0.0) r.x = r.x + P.L[0];
            if (r.y > P.L[1]) r.y = r.y - P.L[1]; else if (r.y < 0.0) r.y = r.y + P.L[1];
            v.x = v.x + hd * a.x; v.y = v.y + hd * a.y; v.z = v.z + hd * a.z;
            CUR(B.pos)[k] = r;
            CUR(B.vel)[k] = v;
            if (!isfinite(r.x) || !isfinite(r.y) || !isfinite(r.z)) atomicAdd(&B.flags[0], 1);
            if (r.z < 0.0 || r.z > P.L[2]) atomicAdd(&B.flags[1], 1);
            if (P.skin_mode) {
                const double4 b = CUR(B.posb)[k];
                double dx = r.x - b.x, dy = r.y - b.y, dz = r.z - b.z;
                dx -= P.L[0] * rint(dx * P.invL[0]);
                dy -= P.L[1] * rint(dy * P.invL[1]);
                if (!(dx * dx + dy * dy + dz * dz <= P.skin_half_2)) { atomicOr(B.rebuild, 1); atomicOr(&B.chan_rb[s], 1); }
            }
            if (MIXED) {
                const int c = CUR(B.cell_of)[k] - P.cell_base;
                if (guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) {
                    const int nx = P.nc[0], ny = P.nc[1];
                    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
                    double rx = r.x - cx * P.l_cell[0], ry = r.y - cy * P.l_cell[1], rz = r.z - cz * P.l_cell[2];
                    if (rx > 0.5 * P.L[0]) rx -= P.L[0]; else if (rx < -0.5 * P.L[0]) rx += P.L[0];
                    if (ry > 0.5 * P.L[1]) ry -= P.L[1]; else if (ry < -0.5 * P.L[1]) ry += P.L[1];
                    CUR(B.posf)[k] = make_float4((float)rx, (float)ry, (float)rz, mkwf(worig(r.w), wtype(r.w)));
                    CUR(B.velf)[k] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
                }
            }
        }
    }
    if (last) {
        double t = block_sum(ke, red);
        if (threadIdx.x == 0) partial[(size_t)blockIdx.x * NPART + C_KE] = t;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *B.t_idx += 1;
}

// Time index of the last of the n steps an md_steps call enqueues (the
// current index plus n - 1); read by k_vv1 at every fused step.
__global__ void k_set_t_end(Bufs B, int n) { *B.t_end = *B.t_idx + n - 1; }

// Per-channel reduction of the block partials in fixed order.
__global__ __launch_bounds__(BLK) void k_reduce(const SysDev *__restrict__ sys, const double *__restrict__ partial,
                                               double *__restrict__ energies) {
    __shared__ double red[BLK];
    const SysDev &P = sys[blockIdx.x];
    for (int col = 0; col < C_NCOL; ++col) {
        double acc = 0.0;
        for (int b = threadIdx.x; b < P.n_blocks; b += BLK) acc += partial[(size_t)(P.block_base + b) * NPART + col];
        double t = block_sum(acc, red);
        if (threadIdx.x == 0) energies[blockIdx.x * NPART + col] = t;
    }
}

// ---------------------------------------------------------------------------
// Cell-list rebuild: counting sort of the beads by cell (replaces the
// per-bead relinking update_part_cell, cell_list.f90:215-268, and
// neigh_list, :56-106).  Every kernel is gated on *rebuild.
// ---------------------------------------------------------------------------
// Request a rebuild of channel s (after its state was uploaded).
__global__ __launch_bounds__(BLK) void k_mark_rebuild(Bufs B, int s) {
    *B.rebuild = 1;
    B.chan_rb[s] = 1;
}

// Clear the rebuild requests of all n_sys channels.
// Clears the rebuild gates once the gated rebuild has run.  fired (may be
// null): set to whether the gates were up, i.e. whether the skin test of the
// step fired (f_skin after check_skin.f90:11-24).
__global__ __launch_bounds__(BLK) void k_clear_rebuild(Bufs B, int n_sys, int *fired) {
    for (int s = threadIdx.x; s < n_sys; s += BLK) B.chan_rb[s] = 0;
    if (threadIdx.x == 0) {
        if (fired) *fired = *B.rebuild != 0 ? 1 : 0;
        *B.rebuild = 0;
    }
}

__global__ __launch_bounds__(BLK) void k_bin(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B,
                                            int *__restrict__ cell_count, int *__restrict__ bin_cell,
                                            int *__restrict__ bin_rank, int *__restrict__ bin_key, int n_slots) {
    if (!*B.rebuild) return;
    const int lane = threadIdx.x & 63;
    const unsigned long long below = (1ull << lane) - 1ull;
    for (int k0 = blockIdx.x * BLK; k0 < n_slots; k0 += gridDim.x * BLK) {
        const int k = k0 + threadIdx.x;
        const int si = blk_sys[k0 / BLK];
        const SysDev &P = sys[si];
        int c = -1;
        if (k - P.slot_base < P.n_part && B.chan_rb[si]) {   // only rebuilding channels
            MFA_CUR(B, si);
            const double4 r = CUR(B.pos)[k];
            const int cc = cell_of_pos(P, r.x, r.y, r.z);
            if (guard(B, in_range(cc, B.n_cells), G_BIN, k)) {
                c = cc;
                bin_cell[k] = c;
                bin_key[k] = worig(r.w);
            }
        }
        // rank within the cell (arbitrary order; k_rank orders the cell):
        // the beads stay sorted by cell between rebuilds, so a wave spans a
        // few cells, and one atomic per distinct cell replaces one per bead
        unsigned long long todo = __ballot(c >= 0);
        while (todo) {
            const int leader = __ffsll((long long)todo) - 1;
            const int cl = __shfl(c, leader);
            const unsigned long long peers = __ballot(c == cl);
            int base = 0;
            if (lane == leader) base = atomicAdd(&cell_count[cl], __popcll(peers));
            base = __shfl(base, leader);
            if (c == cl) bin_rank[k] = base + __popcll(peers & below);
            todo &= ~peers;
        }
    }
}

// exclusive scan, 1024 cells per block (4 per thread)
__global__ __launch_bounds__(BLK) void k_scan1(const int *__restrict__ rebuild, const int *__restrict__ cnt, int n,
                                              int *__restrict__ pre, int *__restrict__ bsum) {
    if (!*rebuild) return;
    __shared__ int sh[BLK];
    const int base = blockIdx.x * 4 * BLK + 4 * threadIdx.x;
    int v[4], t = 0;
    for (int i = 0; i < 4; ++i) { v[i] = (base + i < n) ? cnt[base + i] : 0; t += v[i]; }
    sh[threadIdx.x] = t;
    __syncthreads();
    for (int off = 1; off < BLK; off <<= 1) {
        int x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
        __syncthreads();
        sh[threadIdx.x] += x;
        __syncthreads();
    }
    int ex = sh[threadIdx.x] - t;
    for (int i = 0; i < 4; ++i) { if (base + i < n) pre[base + i] = ex; ex += v[i]; }
    if (threadIdx.x == BLK - 1) bsum[blockIdx.x] = sh[BLK - 1];
}

__global__ __launch_bounds__(BLK) void k_scan2(const int *__restrict__ rebuild, int *__restrict__ bsum, int nb) {
    if (!*rebuild) return;
    __shared__ int sh[BLK];
    int carry = 0;
    for (int base = 0; base < nb; base += BLK) {
        const int i = base + threadIdx.x;
        const int t = i < nb ? bsum[i] : 0;
        sh[threadIdx.x] = t;
        __syncthreads();
        for (int off = 1; off < BLK; off <<= 1) {
            int x = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
            __syncthreads();
            sh[threadIdx.x] += x;
            __syncthreads();
        }
        if (i < nb) bsum[i] = carry + sh[threadIdx.x] - t;
        carry += sh[BLK - 1];
        __syncthreads();
    }
}

// Cell ranges of the rebuilding channels: the global exclusive prefix counts
// only their beads (k_bin skips the others), so a channel's ranges are prefix
// differences from its first cell plus its slot base; the other channels keep
// their ranges.
__global__ __launch_bounds__(BLK) void k_scan3(const SysDev *__restrict__ sys, const int *__restrict__ rebuild,
                                              const int *__restrict__ chan_rb, int *__restrict__ cnt,
                                              const int *__restrict__ pre, const int *__restrict__ bsum,
                                              const int *__restrict__ cell_sys, int2 *__restrict__ crange, int n) {
    if (!*rebuild) return;
    for (int c = blockIdx.x * BLK + threadIdx.x; c < n; c += gridDim.x * BLK) {
        const int s = cell_sys[c];
        if (!chan_rb[s]) continue;
        const SysDev &P = sys[s];
        const int cb = P.cell_base;
        const int st = (pre[c] + bsum[c / (4 * BLK)]) - (pre[cb] + bsum[cb / (4 * BLK)]) + P.slot_base;
        crange[c] = make_int2(st, st + cnt[c]);
        cnt[c] = 0;
    }
}

__global__ __launch_bounds__(BLK) void k_scatter(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                                Bufs B, const int *__restrict__ bin_cell,
                                                const int *__restrict__ bin_rank, const int *__restrict__ bin_key,
                                                const int2 *__restrict__ crange, int *__restrict__ key_tmp,
                                                int n_slots) {
    if (!*B.rebuild) return;
    for (int k = blockIdx.x * BLK + threadIdx.x; k < n_slots; k += gridDim.x * BLK) {
        const int si = blk_sys[k / BLK];
        const SysDev &P = sys[si];
        if (k - P.slot_base >= P.n_part || !B.chan_rb[si]) continue;
        const int c = bin_cell[k];
        if (!guard(B, in_range(c, B.n_cells), G_BIN, k)) continue;
        const int m = crange[c].x + bin_rank[k];
        if (!guard(B, in_range(m, n_slots), G_RANGE, k)) continue;
        key_tmp[m] = bin_key[k];
    }
}

// Order each cell by reference bead index (the atomic ranks of k_bin are
// arbitrary): every bead counts the beads of its cell with a smaller index,
// from the cell's keys that k_scatter laid out contiguously, and takes that
// slot.
__global__ __launch_bounds__(BLK) void k_rank(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B,
                                             const int *__restrict__ bin_cell, const int2 *__restrict__ crange,
                                             const int *__restrict__ bin_key, const int *__restrict__ key_tmp,
                                             int *__restrict__ perm, int n_slots) {
    if (!*B.rebuild) return;
    for (int k = blockIdx.x * BLK + threadIdx.x; k < n_slots; k += gridDim.x * BLK) {
        const int si = blk_sys[k / BLK];
        const SysDev &P = sys[si];
        if (k - P.slot_base >= P.n_part || !B.chan_rb[si]) continue;
        const int c = bin_cell[k];
        if (!guard(B, in_range(c, B.n_cells), G_BIN, k)) continue;
        const int2 rg = crange[c];
        if (!guard(B, rg.x >= 0 && rg.x <= rg.y && rg.y <= n_slots, G_RANGE, k)) continue;
        const int key = bin_key[k];
        int rank = 0;
        for (int m = rg.x; m < rg.y; ++m) rank += key_tmp[m] < key ? 1 : 0;
        if (guard(B, rank < rg.y - rg.x, G_RANGE, k)) perm[rg.x + rank] = k;
    }
}

// Permutes the rebuilding channels into the other half.  The last block to
// finish (every block has read the parities by then) flips those channels'
// parities and counts the rebuild; `done` is its ticket counter, left at 0.
// keep_frc: the total force moves with its bead (the state a caller reads
// after verlet_positions holds the last step's force, as the reference's
// particle-indexed force array does); the fused step passes 0, since its pair
// kernel overwrites the force right after the rebuild.
template <bool MIXED>
__global__ __launch_bounds__(BLK) void k_permute(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B,
                                                const int *__restrict__ perm, const int *__restrict__ bin_cell,
                                                int n_slots, int *__restrict__ done, int *__restrict__ rebuild_count,
                                                int n_sys, int keep_frc) {
    if (!*B.rebuild) return;
    for (int m = blockIdx.x * BLK + threadIdx.x; m < n_slots; m += gridDim.x * BLK) {
        const int si = blk_sys[m / BLK];
        const SysDev &P = sys[si];
        if (m - P.slot_base >= P.n_part || !B.chan_rb[si]) continue;
        MFA_CUR(B, si);
        const int old = perm[m];
        if (!guard(B, in_range(old, n_slots), G_PERM, m)) continue;
        const double4 r = CUR(B.pos)[old];
        if (!guard(B, in_range(worig(r.w), B.n_beads), G_ORIG, m)) continue;
        const double4 v = CUR(B.vel)[old];
        NXT(B.pos)[m] = r;
        NXT(B.vel)[m] = v;
        NXT(B.acc)[m] = CUR(B.acc)[old];
        if (keep_frc) NXT(B.frc)[m] = CUR(B.frc)[old];
        NXT(B.posb)[m] = r;                            // skin reference of this rebuild
        const int cg = bin_cell[old];
        NXT(B.cell_of)[m] = cg;
        B.where[worig(r.w)] = m;
        if (MIXED) {
            const int c = cg - P.cell_base;
            if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, m)) continue;
            const int nx = P.nc[0], ny = P.nc[1];
            const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
            double rx = r.x - cx * P.l_cell[0], ry = r.y - cy * P.l_cell[1], rz = r.z - cz * P.l_cell[2];
            if (rx > 0.5 * P.L[0]) rx -= P.L[0]; else if (rx < -0.5 * P.L[0]) rx += P.L[0];
            if (ry > 0.5 * P.L[1]) ry -= P.L[1]; else if (ry < -0.5 * P.L[1]) ry += P.L[1];
            NXT(B.posf)[m] = make_float4((float)rx, (float)ry, (float)rz, mkwf(worig(r.w), wtype(r.w)));
            NXT(B.velf)[m] = make_float4((float)v.x, (float)v.y, (float)v.z, 0.f);
        }
    }
    // (no fence: the last block reads nothing the others wrote, and every
    // block's parity reads were consumed before its ticket)
    __shared__ int last;
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(done, 1) == (int)gridDim.x - 1;
    __syncthreads();
    if (last && threadIdx.x == 0) {
        for (int s = 0; s < n_sys; ++s)
            if (B.chan_rb[s]) B.parity[s] ^= 1;
        *rebuild_count += 1;
        *done = 0;
    }
}

// ---------------------------------------------------------------------------
// Host <-> device in reference order
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(BLK) void k_upload(const SysDev *__restrict__ sys, int s, Bufs B, const double *__restrict__ r0,
                                               const double *__restrict__ v, const double *__restrict__ a,
                                               const int *__restrict__ type) {
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= P.n_part) return;
    const int k = P.slot_base + i;
    const int g = P.part_base + i;
    CUR(B.pos)[k] = make_double4(r0[3 * i], r0[3 * i + 1], r0[3 * i + 2], mkw(g, type[i]));
    CUR(B.vel)[k] = make_double4(v[3 * i], v[3 * i + 1], v[3 * i + 2], 0.0);
    CUR(B.acc)[k] = a ? make_double4(a[3 * i], a[3 * i + 1], a[3 * i + 2], 0.0) : make_double4(0.0, 0.0, 0.0, 0.0);
    CUR(B.frc)[k] = make_double4(0.0, 0.0, 0.0, 0.0);
    CUR(B.cell_of)[k] = -1;
    B.where[g] = k;
}

template <int WHICH>
__global__ __launch_bounds__(BLK) void k_download(const SysDev *__restrict__ sys, int s, Bufs B, double *__restrict__ out) {
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i >= P.n_part) return;
    const int k = B.where[P.part_base + i];
    if (!guard(B, in_range(k, B.n_slots), G_WHERE, i)) return;
    const double4 *src = WHICH == 0 ? CUR(B.pos) : WHICH == 1 ? CUR(B.vel) : WHICH == 2 ? CUR(B.acc) : CUR(B.frc);
    const double4 q = src[k];
    out[3 * i] = q.x; out[3 * i + 1] = q.y; out[3 * i + 2] = q.z;
}

// ---------------------------------------------------------------------------
// check_fluid_fluid, util.f90:284-330: O(N^2) brute-force LJ energy (and the
// pair-set fingerprint), one thread per bead i over all j != i.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(BLK) void k_brute(const SysDev *__restrict__ sys, int s, Bufs B, double r_2_min,
                                              double *__restrict__ e_out, int *__restrict__ cnt_out,
                                              unsigned long long *__restrict__ hash_out) {
#pragma clang fp contract(off)
    MFA_CUR(B, s);
    __shared__ double4 tile[BLK];
    const SysDev &P = sys[s];
    const int i = blockIdx.x * BLK + threadIdx.x;
    const bool act = i < P.n_part;
    double4 ri = act ? bead_at(B, cur_off, P.part_base + i) : make_double4(0, 0, 0, 0);
    const int ti = act ? wtype(ri.w) : 1;
    double e = 0.0;
    int cnt = 0;
    unsigned long long h = 0ull;
    for (int j0 = 0; j0 < P.n_part; j0 += BLK) {
        const int jj = j0 + threadIdx.x;
        tile[threadIdx.x] = jj < P.n_part ? bead_at(B, cur_off, P.part_base + jj) : make_double4(0, 0, 0, 0);
        __syncthreads();
        const int jn = min(BLK, P.n_part - j0);
        if (act) {
            for (int q = 0; q < jn; ++q) {
                const int j = j0 + q;
                if (j == i) continue;
                const double4 rj = tile[q];
                const int tj = wtype(rj.w);
                const int tt = ((ti - 1) & 3) + 4 * ((tj - 1) & 3);
                double dx = mic(ri.x - rj.x, P.L[0], P.invL[0]);
                double dy = mic(ri.y - rj.y, P.L[1], P.invL[1]);
                double dz = ri.z - rj.z;
                double r_2 = r2_exact(dx, dy, dz);
                if (r_2 < P.range_2[tt]) {
                    ++cnt;
                    h += mix64((unsigned long long)j);
                    if (j > i) {
                        if (r_2 < r_2_min) r_2 = r_2_min;
                        double qd = P.sigma_2[tt] / r_2;
                        double r_6 = qd * qd * qd;
                        double r_12 = r_6 * r_6;
                        e += P.epsil[tt] * ((r_12 - r_6) - P.e_shift[tt]);
                    }
                }
            }
        }
        __syncthreads();
    }
    if (act) { e_out[i] = e; cnt_out[i] = cnt; hash_out[i] = h; }
}

// ---------------------------------------------------------------------------
// Profiles: histo(2) functions.f90:157-213 and velocity_prof(2) :319-365.
// Per observation the sums go through 64-bit fixed point (order-independent,
// exact), then one thread per bin adds them to the FP64 running sums.
// ---------------------------------------------------------------------------
constexpr double FX_V = 4294967296.0;      // 2^32 for v
constexpr double FX_V2 = 4294967296.0;     // 2^32 for v^2

__global__ __launch_bounds__(BLK) void k_prof_bin(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B,
                                                 int nb, unsigned int *__restrict__ cnt,
                                                 unsigned long long *__restrict__ vs,
                                                 unsigned long long *__restrict__ v2s) {
    const int s = blk_sys[blockIdx.x];
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    const int k = blockIdx.x * BLK + threadIdx.x;
    if (k - P.slot_base >= P.n_part) return;
    const double4 r = CUR(B.pos)[k];
    const double4 v = CUR(B.vel)[k];
    const int o = worig(r.w) - P.part_base;
    const int sp = o < P.pd ? 0 : 1;
    const double inv_step = 1.0 / (P.L[2] / (double)nb);
    int z = (int)(r.z * inv_step) + 1;
    z = z < 1 ? 1 : (z > nb ? nb : z);
    const size_t base = (size_t)s * 2 * nb + (size_t)sp * nb + (z - 1);
    atomicAdd(&cnt[base], 1u);
    const double vv[3] = {v.x, v.y, v.z};
    for (int d = 0; d < 3; ++d) {
        const size_t o3 = ((size_t)s * 2 * 3 + (size_t)sp * 3 + d) * nb + (z - 1);
        atomicAdd(&vs[o3], (unsigned long long)(long long)llrint(vv[d] * FX_V));
        atomicAdd(&v2s[o3], (unsigned long long)(long long)llrint(vv[d] * vv[d] * FX_V2));
    }
}

// ---------------------------------------------------------------------------
// Chain observables of channel s (observation.f90:182-290).
// ---------------------------------------------------------------------------
// Radius of gyration per axis of every brush chain (n_mon beads, head first)
// and melt chain (n_mon_d beads).  A chain is unfolded bead by bead against
// its previous, already unfolded bead: a jump of more than half the box
// along an axis moves the bead by one box length (:182-237).  Then the
// centre of mass and the mean squared deviation per axis, summed in bead
// order and normalised as the reference does (brush: times 1/n_mon; melt:
// divided by n_mon_d) (:247-271).  Per block, in a fixed order: part[8 b +
// 0..2] the sum over its brush chains, part[8 b + 3..5] over its melt chains.
__device__ __forceinline__ void unfold_step(double *x, const double *prev, const double *L, const double *h) {
#pragma clang fp contract(off)
    for (int d = 0; d < 3; ++d) {
        if (x[d] - prev[d] > h[d]) x[d] = x[d] - L[d];
        else if (prev[d] - x[d] > h[d]) x[d] = x[d] + L[d];
    }
}

__global__ __launch_bounds__(BLK) void k_chain_rg2(const SysDev *__restrict__ sys, int s, Bufs B,
                                                  double *__restrict__ part) {
#pragma clang fp contract(off)
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    __shared__ double red[BLK];
    const int c = blockIdx.x * BLK + threadIdx.x;
    const int nb = P.n_chain, nt = P.n_chain + P.n_chain_d;
    const bool brush = c < nb;
    double g[3] = {0.0, 0.0, 0.0};
    const int len = brush ? P.n_mon : P.n_mon_d;
    if (c < nt && len > 0) {
        const int first = P.part_base + (brush ? c * P.n_mon : P.pd + (c - nb) * P.n_mon_d);
        const double L[3] = {P.L[0], P.L[1], P.L[2]};
        const double h[3] = {P.L[0] / 2.0, P.L[1] / 2.0, P.L[2] / 2.0};
        const double inv_len = 1.0 / (double)len;
        auto bead = [&](int j, double *x) {
            int k = B.where[first + j];
            if (!guard(B, in_range(k, B.n_slots), G_WHERE, first + j)) k = 0;
            const double4 r = CUR(B.pos)[k];
            x[0] = r.x; x[1] = r.y; x[2] = r.z;
        };
        double sum[3] = {0.0, 0.0, 0.0}, prev[3] = {0.0, 0.0, 0.0}, x[3];
        for (int j = 0; j < len; ++j) {
            bead(j, x);
            if (j > 0) unfold_step(x, prev, L, h);
            for (int d = 0; d < 3; ++d) { sum[d] = sum[d] + x[d]; prev[d] = x[d]; }
        }
        double cm[3], acc[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; ++d) cm[d] = brush ? sum[d] * inv_len : sum[d] / (double)len;
        for (int j = 0; j < len; ++j) {
            bead(j, x);
            if (j > 0) unfold_step(x, prev, L, h);
            for (int d = 0; d < 3; ++d) {
                const double e = x[d] - cm[d];
                acc[d] = acc[d] + e * e;
                prev[d] = x[d];
            }
        }
        for (int d = 0; d < 3; ++d) g[d] = brush ? acc[d] * inv_len : acc[d] / (double)len;
    }
    for (int q = 0; q < 6; ++q) {
        const bool mine = c < nt && ((q < 3) == brush);
        const double t = block_sum(mine ? g[q % 3] : 0.0, red);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8 + q] = t;
    }
}

// Total momentum of channel s, sum of mass * v over its beads
// (observation.f90:286-290): per block of slots, part[8 b + 0..2].
__global__ __launch_bounds__(BLK) void k_momentum(const SysDev *__restrict__ sys, int s, Bufs B,
                                                 double *__restrict__ part) {
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    __shared__ double red[BLK];
    const int k = P.slot_base + blockIdx.x * BLK + threadIdx.x;
    double m[3] = {0.0, 0.0, 0.0};
    if (k - P.slot_base < P.n_part) {
        const double4 r = CUR(B.pos)[k], v = CUR(B.vel)[k];
        const double ms = P.mass_type[(wtype(r.w) - 1) & 3];
        m[0] = ms * v.x; m[1] = ms * v.y; m[2] = ms * v.z;
    }
    for (int d = 0; d < 3; ++d) {
        const double t = block_sum(m[d], red);
        if (threadIdx.x == 0) part[(size_t)blockIdx.x * 8 + d] = t;
    }
}

// Verlet-list statistics of channel s (a diagnostic, outside the step): per
// block of slots, part[3 b + 0] the list entries of its beads, part[3 b + 1]
// the entries its waves walk (a wave walks the longest list of its lanes, in
// groups of four) and part[3 b + 2] its beads whose list overflowed (they scan
// their cells instead).
__global__ __launch_bounds__(BLK) void k_list_stats(const SysDev *__restrict__ sys, int s, Bufs B,
                                                   double *__restrict__ part) {
    const SysDev &P = sys[s];
    __shared__ int nn[BLK];
    const int i = blockIdx.x * BLK + threadIdx.x;
    nn[threadIdx.x] = i < P.n_part ? B.nl_n[P.slot_base + i] : -1;
    __syncthreads();
    if (threadIdx.x != 0) return;
    double ent = 0.0, walk = 0.0, over = 0.0;
    for (int w = 0; w < BLK; w += 64) {
        int mx = 0;
        for (int l = 0; l < 64; ++l) {
            const int n = nn[w + l];
            if (n < 0) continue;
            if (n > B.nl_cap) { over += 1.0; continue; }
            ent += (double)n;
            mx = max(mx, n);
        }
        walk += (double)((mx + 3) / 4 * 4);
    }
    part[3 * (size_t)blockIdx.x] = ent;
    part[3 * (size_t)blockIdx.x + 1] = walk;
    part[3 * (size_t)blockIdx.x + 2] = over;
}

__global__ __launch_bounds__(BLK) void k_prof_fold(int n_cnt, int n_v, unsigned int *__restrict__ cnt,
                                                  unsigned long long *__restrict__ vs,
                                                  unsigned long long *__restrict__ v2s, double *__restrict__ dens,
                                                  double *__restrict__ vp, double *__restrict__ vp2) {
    const int i = blockIdx.x * BLK + threadIdx.x;
    if (i < n_cnt) { dens[i] += (double)cnt[i]; cnt[i] = 0u; }
    if (i < n_v) {
        vp[i] += (double)(long long)vs[i] / FX_V;
        vp2[i] += (double)(long long)v2s[i] / FX_V2;
        vs[i] = 0ull; v2s[i] = 0ull;
    }
}

}  // namespace mfa
