// mfa_pair.h -- non-bonded pair kernels: truncated-shifted LJ + DPD.
//
// One thread per bead (sorted slot), full neighbour shell, no Newton-3
// write to the partner: every unordered pair is evaluated by both of its
// beads.  The two evaluations are exact negatives of each other (the minimum
// image and every product below are odd in delta_r, and the DPD normal is
// keyed on the unordered pair), so no atomics are needed and each bead's
// force is a fixed-order register sum, hence deterministic.
//
// Arithmetic follows fluid_fluid.f90:97-154 and dpd_forces.f90:45-115.
#pragma once
#include "mfa_kernels.h"

namespace mfa {

struct PairAcc {
    double fx, fy, fz, e;
    double w[6];
    int np;
};

// One candidate j for bead i in FP64.  Returns nothing; accumulates into a.
template <bool DPD, bool VIR>
__device__ __forceinline__ void pair64(const SysDev &P, const double *tab, double xi, double yi, double zi,
                                       int ti, int oi, double vxi, double vyi, double vzi,
                                       double4 pj, const double4 *__restrict__ vel, int m,
                                       double r_2_min, uint64_t t_idx, PairAcc &a) {
    const int tj = wtype(pj.w);
    const int tt = ((ti - 1) & 3) + 4 * ((tj - 1) & 3);
    double dx = mic(xi - pj.x, P.L[0], P.invL[0]);
    double dy = mic(yi - pj.y, P.L[1], P.invL[1]);
    double dz = zi - pj.z;
    double r_2 = r2_exact(dx, dy, dz);
    const double r_cut2 = tab[32 + tt];
    bool lj = r_2 < r_cut2;
    double inv_sqrt = 0.0;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (lj) {
        a.np += 1;
        if (P.force_switch_on && r_2 < r_2_min) r_2 = r_2_min;
        double inv_r_2 = 1.0 / r_2;
        inv_sqrt = sqrt(inv_r_2);
        double l_eps = tab[tt];
        double r_61 = tab[16 + tt] * inv_r_2;
        double r_6 = r_61 * r_61 * r_61;
        double r_12 = r_6 * r_6;
        double pot = (r_12 - r_6) - tab[48 + tt];
        a.e += l_eps * pot;
        double rd = l_eps * (-12.0 * r_12 + 6.0 * r_6) * inv_r_2;
        fx = -rd * dx; fy = -rd * dy; fz = -rd * dz;
    }
    if (DPD) {
        bool do_dpd;
        if (P.dpd_cut_off) {
            do_dpd = ti != 1 && tj != 1 && r_2 < P.r_cut_dpd_2;
        } else {
            do_dpd = lj;
        }
        if (do_dpd) {
            double w;
            if (P.dpd_cut_off) {
                if (!lj) inv_sqrt = sqrt(1.0 / r_2);
                w = 1.0 - sqrt(r_2 / P.r_cut_dpd_2);
            } else {
                w = 1.0 - sqrt(r_2 * tab[64 + tt]);
            }
            const int oj = worig(pj.w);
            double g = (double)pair_gauss(P.seed, t_idx, oi - P.part_base, oj - P.part_base);
            double rvx = dx * inv_sqrt, rvy = dy * inv_sqrt, rvz = dz * inv_sqrt;
            double sr = P.sig * w * g;
            double4 vj = vel[m];
            double rdv = (vxi - vj.x) * rvx + (vyi - vj.y) * rvy + (vzi - vj.z) * rvz;
            double sd = -1.0 * P.friction * (w * w) * rdv;
            fx += sr * rvx + sd * rvx;
            fy += sr * rvy + sd * rvy;
            fz += sr * rvz + sd * rvz;
        }
    }
    a.fx += fx; a.fy += fy; a.fz += fz;
    if (VIR) {
        a.w[0] += dx * fx; a.w[1] += dy * fy; a.w[2] += dz * fz;
        a.w[3] += dx * fy; a.w[4] += dx * fz; a.w[5] += dy * fz;
    }
}

__device__ __forceinline__ double block_sum(double v, double *sh) {
    // fixed-order tree over the block (deterministic)
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int s = BLK / 2; s > 0; s >>= 1) {
        if (t < s) sh[t] += sh[t + s];
        __syncthreads();
    }
    double r = sh[0];
    __syncthreads();
    return r;
}

__device__ __forceinline__ void load_tables(const SysDev &P, double *tab) {
    const int t = threadIdx.x;
    if (t < 16) {
        tab[t] = P.epsil[t];
        tab[16 + t] = P.sigma_2[t];
        tab[32 + t] = P.range_2[t];
        tab[48 + t] = P.e_shift[t];
        tab[64 + t] = P.inv_range_2[t];
    }
}

// ---------------------------------------------------------------------------
// FP64 pair kernel.  Neighbour cells are visited as x-strips: for each
// (dy, dz) the cells cx-S..cx+S of a row are contiguous slots unless the row
// wraps, so one strip is one or two slot ranges.
// ---------------------------------------------------------------------------
template <bool DPD, bool VIR>
__global__ __launch_bounds__(BLK) void k_pair64(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                               Bufs B, const int2 *__restrict__ crange,
                                               double *__restrict__ partial, int assign) {
    MFA_CUR(B);
    const double4 *__restrict__ pos = CUR(B.pos);
    const double4 *__restrict__ vel = CUR(B.vel);
    const int *__restrict__ cell_of = CUR(B.cell_of);
    double4 *__restrict__ frc = B.frc;
    const long long *t_dev = B.t_idx;
    int *rebuild = B.rebuild;
    __shared__ double tab[80];
    __shared__ double red[BLK];
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    const SysDev &P = sys[s];
    load_tables(P, tab);
    if (blockIdx.x == 0 && threadIdx.x == 0) *rebuild = 0;
    if ((int)blockIdx.x == P.block_base && threadIdx.x == 0) B.chan_rb[s] = 0;
    __syncthreads();
    const int k = blockIdx.x * BLK + threadIdx.x;
    const bool active = (k - P.slot_base) < P.n_part;
    const long long t_idx = *t_dev;
    double r_2_min = 0.0;
    if (P.force_switch_on) {
        double r_time = (double)t_idx * P.dt;
        if (r_time < P.r_2_min_time) r_2_min = (1.0 - r_time / P.r_2_min_time) * P.r_2_min_init;
    }
    PairAcc a;
    a.fx = a.fy = a.fz = a.e = 0.0;
    for (int q = 0; q < 6; ++q) a.w[q] = 0.0;
    a.np = 0;
    if (active) {
        const double4 pi = pos[k];
        const int ti = wtype(pi.w), oi = worig(pi.w);
        guard(B, ti >= 1 && ti <= 4, G_TYPE, k);
        double vxi = 0.0, vyi = 0.0, vzi = 0.0;
        if (DPD) { double4 v = vel[k]; vxi = v.x; vyi = v.y; vzi = v.z; }
        int c = cell_of[k] - P.cell_base;
        if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) c = 0;
        const int nx = P.nc[0], ny = P.nc[1], nz = P.nc[2];
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        const int S = P.S;
        for (int dz = -S; dz <= S; ++dz) {
            const int z2 = cz + dz;
            if (z2 < 0 || z2 >= nz) continue;
            for (int dy = -S; dy <= S; ++dy) {
                int y2 = cy + dy;
                y2 += (y2 < 0) ? ny : 0;
                y2 -= (y2 >= ny) ? ny : 0;
                const int row = P.cell_base + nx * (y2 + ny * z2);
                int xa = cx - S, xb = cx + S;
                int seg_a0, seg_a1, seg_b0 = 0, seg_b1 = -1;
                if (xa >= 0 && xb < nx) { seg_a0 = xa; seg_a1 = xb; }
                else if (xa < 0) { seg_a0 = 0; seg_a1 = xb; seg_b0 = nx + xa; seg_b1 = nx - 1; }
                else { seg_a0 = xa; seg_a1 = nx - 1; seg_b0 = 0; seg_b1 = xb - nx; }
                for (int seg = 0; seg < 2; ++seg) {
                    const int c0 = seg == 0 ? seg_a0 : seg_b0, c1 = seg == 0 ? seg_a1 : seg_b1;
                    if (c1 < c0) continue;
                    const int m0 = crange[row + c0].x, m1 = crange[row + c1].y;
                    if (!guard(B, m0 >= 0 && m0 <= m1 && m1 <= B.n_slots, G_SLOT, k)) continue;
                    for (int m = m0; m < m1; ++m) {
                        if (m == k) continue;
                        pair64<DPD, VIR>(P, tab, pi.x, pi.y, pi.z, ti, oi, vxi, vyi, vzi, pos[m], vel, m,
                                         r_2_min, (uint64_t)t_idx, a);
                    }
                }
            }
        }
        double4 f = assign ? make_double4(0.0, 0.0, 0.0, 0.0) : frc[k];
        f.x += a.fx; f.y += a.fy; f.z += a.fz;
        frc[k] = f;
    }
    double *out = partial + (size_t)blockIdx.x * NPART;
    double e = block_sum(a.e, red);
    double np = block_sum((double)a.np, red);
    if (threadIdx.x == 0) { out[C_VFF] = e; out[C_NPAIR] = np; }
    if (VIR) {
        for (int q = 0; q < 6; ++q) {
            double wq = block_sum(a.w[q], red);
            if (threadIdx.x == 0) out[C_WXX + q] = wq;
        }
    }
}

// ---------------------------------------------------------------------------
// Mixed-precision pair kernel: positions relative to the build-time cell
// origin in FP32 (posf), FP32 pair arithmetic, per-cell FP32 partial sums
// flushed into FP64 accumulators.  Pairs whose FP32 r^2 lies within a
// relative band of 2^-14 around the cutoff are decided again in FP64 with the
// reference arithmetic, so the interacting-pair set is the FP64 one.
// ---------------------------------------------------------------------------
template <bool DPD, bool VIR>
__global__ __launch_bounds__(BLK) void k_pair32(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                               Bufs B, const int2 *__restrict__ crange,
                                               double *__restrict__ partial, int assign) {
    MFA_CUR(B);
    const double4 *__restrict__ pos = CUR(B.pos);
    const float4 *__restrict__ posf = CUR(B.posf);
    const float4 *__restrict__ velf = CUR(B.velf);
    const int *__restrict__ cell_of = CUR(B.cell_of);
    double4 *__restrict__ frc = B.frc;
    const long long *t_dev = B.t_idx;
    int *rebuild = B.rebuild;
    __shared__ float tf[96];
    __shared__ double red[BLK];
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    const SysDev &P = sys[s];
    {
        const int t = threadIdx.x;
        if (t < 16) {
            tf[t] = P.epsil_f[t];
            tf[16 + t] = P.sigma_2f[t];
            tf[32 + t] = P.range_2f_lo[t];
            tf[48 + t] = P.range_2f_hi[t];
            tf[64 + t] = P.e_shift_f[t];
            tf[80 + t] = P.inv_rc_f[t];
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *rebuild = 0;
    if ((int)blockIdx.x == P.block_base && threadIdx.x == 0) B.chan_rb[s] = 0;
    __syncthreads();
    const int k = blockIdx.x * BLK + threadIdx.x;
    const bool active = (k - P.slot_base) < P.n_part;
    const long long t_idx = *t_dev;
    float r_2_min = 0.0f;
    double r_2_min_d = 0.0;
    if (P.force_switch_on) {
        double r_time = (double)t_idx * P.dt;
        if (r_time < P.r_2_min_time) r_2_min_d = (1.0 - r_time / P.r_2_min_time) * P.r_2_min_init;
        r_2_min = (float)r_2_min_d;
    }
    const float sigf = (float)P.sig, gamf = (float)P.friction;
    const float rdpd2 = (float)P.r_cut_dpd_2;
    const float inv_rdpd = (float)(1.0 / sqrt(P.r_cut_dpd_2 > 0.0 ? P.r_cut_dpd_2 : 1.0));
    double fx = 0.0, fy = 0.0, fz = 0.0, e = 0.0;
    double w[6] = {0, 0, 0, 0, 0, 0};
    int np = 0;
    if (active) {
        const float4 pf = posf[k];
        const int ti = fwtype(pf.w), oi = fworig(pf.w);
        guard(B, ti >= 1 && ti <= 4, G_TYPE, k);
        float4 vi = make_float4(0.f, 0.f, 0.f, 0.f);
        if (DPD) vi = velf[k];
        int c = cell_of[k] - P.cell_base;
        if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) c = 0;
        const int nx = P.nc[0], ny = P.nc[1], nz = P.nc[2];
        const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
        const int S = P.S;
        const float lx = (float)P.l_cell[0], ly = (float)P.l_cell[1], lz = (float)P.l_cell[2];
        for (int dz = -S; dz <= S; ++dz) {
            const int z2 = cz + dz;
            if (z2 < 0 || z2 >= nz) continue;
            const float qz = pf.z - (float)dz * lz;
            for (int dy = -S; dy <= S; ++dy) {
                int y2 = cy + dy;
                y2 += (y2 < 0) ? ny : 0;
                y2 -= (y2 >= ny) ? ny : 0;
                const float qy = pf.y - (float)dy * ly;
                const int row = P.cell_base + nx * (y2 + ny * z2);
                for (int dx = -S; dx <= S; ++dx) {
                    int x2 = cx + dx;
                    x2 += (x2 < 0) ? nx : 0;
                    x2 -= (x2 >= nx) ? nx : 0;
                    const float qx = pf.x - (float)dx * lx;
                    int2 rg = crange[row + x2];
                    if (!guard(B, rg.x >= 0 && rg.x <= rg.y && rg.y <= B.n_slots, G_SLOT, k)) rg = make_int2(0, 0);
                    float cfx = 0.f, cfy = 0.f, cfz = 0.f, ce = 0.f;
                    float cw[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                    for (int m = rg.x; m < rg.y; ++m) {
                        if (m == k) continue;
                        const float4 pj = posf[m];
                        const float ddx = qx - pj.x, ddy = qy - pj.y, ddz = qz - pj.z;
                        float r2 = ddx * ddx + ddy * ddy + ddz * ddz;
                        const int tj = fwtype(pj.w);
                        const int tt = ((ti - 1) & 3) + 4 * ((tj - 1) & 3);
                        const bool maybe_dpd = DPD && P.dpd_cut_off && r2 < rdpd2 * 1.0001f;
                        if (!(r2 < tf[48 + tt]) && !maybe_dpd) continue;
                        bool lj = r2 < tf[32 + tt];
                        if (!lj && r2 < tf[48 + tt]) {
                            // borderline: decide with the FP64 reference arithmetic
                            const double4 a = pos[k], b = pos[m];
                            double ex = mic(a.x - b.x, P.L[0], P.invL[0]);
                            double ey = mic(a.y - b.y, P.L[1], P.invL[1]);
                            double ez = a.z - b.z;
                            lj = r2_exact(ex, ey, ez) < P.range_2[tt];
                        }
                        float inv_sqrt = 0.f;
                        float pfx = 0.f, pfy = 0.f, pfz = 0.f;
                        if (lj) {
                            np += 1;
                            if (P.force_switch_on && r2 < r_2_min) r2 = r_2_min;
                            const float inv_r_2 = __builtin_amdgcn_rcpf(r2);
                            inv_sqrt = __builtin_amdgcn_rsqf(r2);
                            const float l_eps = tf[tt];
                            const float r_61 = tf[16 + tt] * inv_r_2;
                            const float r_6 = r_61 * r_61 * r_61;
                            const float r_12 = r_6 * r_6;
                            ce += l_eps * ((r_12 - r_6) - tf[64 + tt]);
                            const float rd = l_eps * (6.0f * r_6 - 12.0f * r_12) * inv_r_2;
                            pfx = -rd * ddx; pfy = -rd * ddy; pfz = -rd * ddz;
                        }
                        if (DPD) {
                            bool do_dpd;
                            float wgt;
                            if (P.dpd_cut_off) {
                                do_dpd = ti != 1 && tj != 1 && r2 < rdpd2;
                                if (do_dpd && !lj) inv_sqrt = __builtin_amdgcn_rsqf(r2);
                                wgt = 1.0f - r2 * inv_sqrt * inv_rdpd;
                            } else {
                                do_dpd = lj;
                                wgt = 1.0f - r2 * inv_sqrt * tf[80 + tt];
                            }
                            if (do_dpd) {
                                const float g = pair_gauss(P.seed, (uint64_t)t_idx, oi - P.part_base, fworig(pj.w) - P.part_base);
                                const float rvx = ddx * inv_sqrt, rvy = ddy * inv_sqrt, rvz = ddz * inv_sqrt;
                                const float4 vj = velf[m];
                                const float rdv = (vi.x - vj.x) * rvx + (vi.y - vj.y) * rvy + (vi.z - vj.z) * rvz;
                                const float sc = sigf * wgt * g - gamf * (wgt * wgt) * rdv;
                                pfx += sc * rvx; pfy += sc * rvy; pfz += sc * rvz;
                            }
                        }
                        cfx += pfx; cfy += pfy; cfz += pfz;
                        if (VIR) {
                            cw[0] += ddx * pfx; cw[1] += ddy * pfy; cw[2] += ddz * pfz;
                            cw[3] += ddx * pfy; cw[4] += ddx * pfz; cw[5] += ddy * pfz;
                        }
                    }
                    fx += (double)cfx; fy += (double)cfy; fz += (double)cfz; e += (double)ce;
                    if (VIR) for (int q = 0; q < 6; ++q) w[q] += (double)cw[q];
                }
            }
        }
        double4 f = assign ? make_double4(0.0, 0.0, 0.0, 0.0) : frc[k];
        f.x += fx; f.y += fy; f.z += fz;
        frc[k] = f;
    }
    double *out = partial + (size_t)blockIdx.x * NPART;
    double es = block_sum(e, red);
    double nps = block_sum((double)np, red);
    if (threadIdx.x == 0) { out[C_VFF] = es; out[C_NPAIR] = nps; }
    if (VIR) {
        for (int q = 0; q < 6; ++q) {
            double wq = block_sum(w[q], red);
            if (threadIdx.x == 0) out[C_WXX + q] = wq;
        }
    }
}

}  // namespace mfa
