// mfa_pair.h -- non-bonded pair kernels: truncated-shifted LJ + DPD.
//
// One thread per bead (sorted slot), full neighbour shell, no Newton-3
// write to the partner: every unordered pair is evaluated by both of its
// beads.  The two evaluations are exact negatives of each other (the minimum
// image, the FP32 cell-frame difference below and every product are odd in
// delta_r, and the DPD normal is keyed on the unordered pair), so no atomics
// are needed and each bead's force is a fixed-order register sum, hence
// deterministic, and the pair forces conserve momentum up to the rounding of
// those sums.
//
// Each pair is evaluated in two parts.  The LJ part runs while the bead walks
// its partners.  A pair within the DPD range is queued in LDS, and its DPD
// part (Philox normal, dissipative and random force) runs after the walk,
// with every lane busy.  Only a few partners per bead are within the DPD
// range (about 3 of 30 at a 1 sigma DPD cutoff), so evaluating the DPD part
// in place would run the Philox rounds at nearly every step of the walk for
// a few lanes of the wave.  Both parts are added into the FP64 accumulators
// separately, in a fixed order: the LJ parts in walk order, then the queued
// DPD parts in walk order.  A pair found when the queue is full has its DPD
// part evaluated in place; the rule depends only on the walk order, so the
// list and the cell scan give the same bits.
//
// Partners come from a Verlet list built when the bead's channel rebuilds its
// cells (k_nlist): every bead of the 27 (or 125) neighbour cells closer than
// the pair type's interaction range plus the skin.  The skin test (every
// bead within skin/2 of its position at that rebuild) keeps the list
// complete until the next rebuild, so the interacting pairs are exactly the
// ones a cell scan would find.  The list keeps the cell-scan order (dz, dy,
// dx, slot), and a bead whose list overflowed scans the cells instead, in the
// same order with the same arithmetic, so both routes give the same bits.
//
// Arithmetic follows fluid_fluid.f90:97-154 and dpd_forces.f90:45-115.
#pragma once
#include <type_traits>

#include "mfa_bonded.h"
#include "mfa_kernels.h"

namespace mfa {

struct PairAcc {
    double fx, fy, fz, e;
    double w[6];
    int np;
};

// list entry: partner slot << NL_SLOT_SHIFT | (partner type - 1) << 7 |
// cell-offset code, the code being (dx+S) + (2S+1)((dy+S) + (2S+1)(dz+S)) for
// the partner's cell relative to the bead's cell (slots below 2^23)
constexpr int NL_SLOT_SHIFT = 9;
constexpr int NL_TYPE_SHIFT = 7;
constexpr int NL_CODE_MASK = 0x7f;
__device__ __forceinline__ int nl_entry(int slot, int type, int code) {
    return (int)(((unsigned)slot << NL_SLOT_SHIFT) | ((unsigned)((type - 1) & 3) << NL_TYPE_SHIFT) | (unsigned)code);
}
constexpr int NL_UNROLL = 4;      // partner loads issued together (one 16-byte list load)
constexpr int NL_STAGE = 512;     // partner beads a list-build block stages in LDS per row of cells
constexpr int DPD_Q = 8;          // LDS queue of deferred DPD pairs per lane
constexpr int DPD_UNROLL = 2;     // deferred pairs whose partner loads are issued together
// Entry q of slot k: groups of four entries, slots interleaved per group, so
// a wave reading group q/4 of its 64 slots reads one contiguous 1 KB span and
// a slot's four consecutive entries share one 16-byte word.
__device__ __forceinline__ size_t nl_index(int k, int q, size_t stride) {
    return ((size_t)(q >> 2) * stride + (size_t)k) * 4 + (size_t)(q & 3);
}
// One 16-byte group of list entries, loaded with the non-temporal (streaming)
// hint: a launch reads the list once, while the partner coordinates it
// points to are gathered tens of times, so the list should not displace
// them from the caches (C2: pair step 173 -> 165 us, whole step -4 %).
__device__ __forceinline__ int4 nl_group(const int *__restrict__ nl, size_t idx) {
    typedef int v4i __attribute__((ext_vector_type(4)));
    const v4i g = __builtin_nontemporal_load((const v4i *)(nl + idx));
    return make_int4(g.x, g.y, g.z, g.w);
}
// Slot i of a per-slot array through an unsigned 32-bit byte offset, so the
// load takes the uniform base in scalar registers and one 32-bit lane offset
// (no 64-bit address add per gather).  Lists hold slots below 2^23 (the host
// refuses more), so i * 32 stays below 2^32.
template <typename T>
__device__ __forceinline__ T ld_slot(const T *__restrict__ base, unsigned i) {
    return *(const T *)((const char *)base + i * (unsigned)sizeof(T));
}
__device__ __forceinline__ int nl_code(int dx, int dy, int dz, int S) {
    const int w = 2 * S + 1;
    return (dx + S) + w * ((dy + S) + w * (dz + S));
}

// Slot range (lo, count <= cap) of the partner cells that a block of beads
// in cells c_lo .. c_hi (channel-relative, sorted) has in row offset (dy, dz):
// without a periodic wrap in y the shifted cells are c_lo + o .. c_hi + o,
// o = (dy + ny dz) nx, taken with S more cells on each side, one contiguous
// slot range; with a wrap, the cells x_lo - S .. x_hi + S of the first row (to
// its end if the block runs into the next row).  Cells outside the range are
// read from memory by the caller.  Block-uniform.
__device__ __forceinline__ int2 stage_range(const SysDev &P, const Bufs &B, const int2 *__restrict__ crange,
                                            int c_lo, int c_hi, int dy, int dz, int cap) {
    const int nx = P.nc[0], ny = P.nc[1], nz = P.nc[2], S = P.S;
    const int x_lo = c_lo % nx, y_lo = (c_lo / nx) % ny, z_lo = c_lo / (nx * ny);
    const int x_hi = c_hi % nx, yz_lo = c_lo / nx, yz_hi = c_hi / nx;
    int ca = 0, cb = -1;
    const int zz = z_lo + dz;
    const int ya = y_lo + dy, yb = (c_hi / nx) % ny + dy;
    if (zz >= 0 && zz < nz) {
        if (ya >= 0 && yb < ny && yb >= ya) {
            const int o = (dy + ny * dz) * nx;
            ca = max(0, c_lo + o - S);
            cb = min(P.n_cells_tot - 1, c_hi + o + S);
        } else {
            const int yy = ya < 0 ? ya + ny : (ya >= ny ? ya - ny : ya);
            const int rowc = nx * (yy + ny * zz);
            ca = rowc + max(0, x_lo - S);
            cb = rowc + ((yz_hi == yz_lo) ? min(nx - 1, x_hi + S) : nx - 1);
        }
    }
    if (cb >= ca) {
        const int a = crange[P.cell_base + ca].x, b = crange[P.cell_base + cb].y;
        if (a >= 0 && b <= B.n_slots && b > a) return make_int2(a, min(b - a, cap));
    }
    return make_int2(0, 0);
}

// Channel scalars the FP64 pair arithmetic reads per candidate, copied out of
// the channel record once per kernel: read through the record, the compiler
// reloads them (scalar loads and waits) after every possible store.
struct Pair64Ctx {
    double L0, L1, iL0, iL1, rdpd2, sig, fric, r_2_min;
    uint64_t seed, t_idx;
    int part_base, dpd_cut, fso;
};

__device__ __forceinline__ Pair64Ctx pair64_ctx(const SysDev &P, long long t_idx, double r_2_min) {
    Pair64Ctx C;
    C.L0 = P.L[0]; C.L1 = P.L[1]; C.iL0 = P.invL[0]; C.iL1 = P.invL[1];
    C.rdpd2 = P.r_cut_dpd_2; C.sig = P.sig; C.fric = P.friction; C.r_2_min = r_2_min;
    C.seed = P.seed; C.t_idx = (uint64_t)t_idx;
    C.part_base = P.part_base; C.dpd_cut = P.dpd_cut_off; C.fso = P.force_switch_on;
    // a batch may mix thermostats: a channel whose thermostat is not DPD takes
    // the DPD_CUT_OFF rule with an empty range, so no pair is flagged for a
    // DPD part (no test per candidate)
    if (P.thermostat != 1) { C.dpd_cut = 1; C.rdpd2 = -1.0; }
    return C;
}

// LJ part of candidate pj for bead i in FP64 (fluid_fluid.f90:104-135), and
// whether the pair takes a DPD term: 0 none, 1 within the LJ range, 2 beyond
// it (DPD_CUT_OFF only).
template <bool DPD, bool VIR>
__device__ __forceinline__ int pair64_lj(const Pair64Ctx &C, const double *tab, double xi, double yi, double zi,
                                         int ti, double4 pj, bool calc_e, PairAcc &a) {
#pragma clang fp contract(off)
    const int tj = wtype(pj.w);
    const int tt = ((ti - 1) & 3) + 4 * ((tj - 1) & 3);
    const double dx = mic(xi - pj.x, C.L0, C.iL0);
    const double dy = mic(yi - pj.y, C.L1, C.iL1);
    const double dz = zi - pj.z;
    double r_2 = r2_exact(dx, dy, dz);
    const bool lj = r_2 < tab[32 + tt];
    if (lj) {
        if (C.fso && r_2 < C.r_2_min) r_2 = C.r_2_min;
        const double inv_r_2 = 1.0 / r_2;
        const double l_eps = tab[tt];
        const double r_61 = tab[16 + tt] * inv_r_2;
        const double r_6 = r_61 * r_61 * r_61;
        const double r_12 = r_6 * r_6;
        if (calc_e) {
            a.np += 1;
            a.e += l_eps * ((r_12 - r_6) - tab[48 + tt]);
        }
        const double rd = l_eps * (-12.0 * r_12 + 6.0 * r_6) * inv_r_2;
        const double fx = -rd * dx, fy = -rd * dy, fz = -rd * dz;
        a.fx += fx; a.fy += fy; a.fz += fz;
        if (VIR) {
            a.w[0] += dx * fx; a.w[1] += dy * fy; a.w[2] += dz * fz;
            a.w[3] += dx * fy; a.w[4] += dx * fz; a.w[5] += dy * fz;
        }
    }
    if (!DPD) return 0;
    const bool do_dpd = C.dpd_cut ? (ti != 1 && tj != 1 && r_2 < C.rdpd2) : lj;
    return do_dpd ? (lj ? 1 : 2) : 0;
}

// DPD part of a pair that pair64_lj flagged (dpd_forces.f90:45-115).  lj: the
// pair is within the LJ range, where the switch-on clamp of r^2 applies as in
// the LJ part.
template <bool VIR>
__device__ __forceinline__ void pair64_dpd(const Pair64Ctx &C, const double *tab, double xi, double yi, double zi,
                                           int ti, int oi, double vxi, double vyi, double vzi, double4 pj,
                                           double4 vj, bool lj, PairAcc &a) {
#pragma clang fp contract(off)
    const double dx = mic(xi - pj.x, C.L0, C.iL0);
    const double dy = mic(yi - pj.y, C.L1, C.iL1);
    const double dz = zi - pj.z;
    double r_2 = r2_exact(dx, dy, dz);
    if (lj && C.fso && r_2 < C.r_2_min) r_2 = C.r_2_min;
    const double inv_sqrt = sqrt(1.0 / r_2);
    double w;
    if (C.dpd_cut) {
        w = 1.0 - sqrt(r_2 / C.rdpd2);
    } else {
        const int tt = ((ti - 1) & 3) + 4 * ((wtype(pj.w) - 1) & 3);
        w = 1.0 - sqrt(r_2 * tab[64 + tt]);
    }
    const double g = (double)pair_gauss(C.seed, C.t_idx, oi - C.part_base, worig(pj.w) - C.part_base);
    const double rvx = dx * inv_sqrt, rvy = dy * inv_sqrt, rvz = dz * inv_sqrt;
    const double sr = C.sig * w * g;
    const double rdv = (vxi - vj.x) * rvx + (vyi - vj.y) * rvy + (vzi - vj.z) * rvz;
    const double sd = -1.0 * C.fric * (w * w) * rdv;
    const double fx = sr * rvx + sd * rvx, fy = sr * rvy + sd * rvy, fz = sr * rvz + sd * rvz;
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
// Verlet-list build (after a rebuild, for the channels that rebuilt).  A
// partner is listed when its distance is below (range + skin) per type pair
// (rl2; < 0 for pairs that never interact).  The distance only has to select
// a superset of the pairs that can interact before the next rebuild, so it
// is taken from the cell-frame coordinates with the cell offset: FP32 on the
// mixed path (16-byte partner loads, threshold widened by 2^-12), FP64
// otherwise.
// ---------------------------------------------------------------------------
template <bool MIXED>
__global__ __launch_bounds__(BLK) void k_nlist(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                              Bufs B, const int2 *__restrict__ crange) {
    if (!*B.rebuild) return;
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    if (!B.chan_rb[s]) return;
    MFA_CUR(B, s);
    const SysDev &P = sys[s];
    using Bead = typename std::conditional<MIXED, float4, double4>::type;
    // list radius^2 per (bead type row, partner type bits 0..7); the FP32
    // threshold is widened by 2^-12 and the FP64 one by 2^-40, so rounding
    // can only add partners
    __shared__ double rl2d[32];
    __shared__ float rl2f[32];
    // the block's staged partners: one row of cells at a time
    __shared__ alignas(16) unsigned char stage_raw[(NL_STAGE + NL_UNROLL) * sizeof(Bead)];
    Bead *const stage = (Bead *)stage_raw;
    if (threadIdx.x < 32) {
        const int tr = threadIdx.x >> 3, tb = threadIdx.x & 7;
        const double r = P.rl2[tr + 4 * ((tb - 1) & 3)];
        rl2d[threadIdx.x] = r * (1.0 + 0x1p-40);
        rl2f[threadIdx.x] = (float)(r * (1.0 + 1.0 / 4096.0));
    }
    const Bead *__restrict__ src;
    if constexpr (MIXED) src = CUR(B.posf); else src = CUR(B.pos);
    const int *__restrict__ cell_of = CUR(B.cell_of);
    const int nx = P.nc[0], ny = P.nc[1], nz = P.nc[2];
    const int S = P.S;
    const int cap = B.nl_cap;
    const int k = blockIdx.x * BLK + threadIdx.x;
    const bool act = k - P.slot_base < P.n_part;
    // the block's first and last cells (slots are sorted by cell)
    const int k_last = min((int)(blockIdx.x * BLK + BLK - 1), P.slot_base + P.n_part - 1);
    int c_lo = cell_of[blockIdx.x * BLK] - P.cell_base, c_hi = cell_of[k_last] - P.cell_base;
    if (!in_range(c_lo, P.n_cells_tot) || !in_range(c_hi, P.n_cells_tot)) c_lo = c_hi = 0;
    int c = 0;
    if (act) {
        c = cell_of[k] - P.cell_base;
        if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) c = 0;
    }
    const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
    double4 pi = make_double4(0.0, 0.0, 0.0, 0.0);
    float4 pfi = make_float4(0.f, 0.f, 0.f, 0.f);
    int ti = 1;
    if (act) {
        if constexpr (MIXED) { pfi = src[k]; ti = fwtype(pfi.w); }
        else { pi = src[k]; ti = wtype(pi.w); }
    }
    const int trow = ((ti - 1) & 3) * 8;
    // entries are collected four at a time in registers (b0..b3, oldest
    // first) and written as one 16-byte group: a complete chunk per store
    int4 *gp = (int4 *)B.nl + k;                // group n / 4 of slot k
    const size_t group_step = (size_t)B.n_slots;
    int b0 = 0, b1 = 0, b2 = 0, b3 = 0;
    int n = 0;
    for (int dz = -S; dz <= S; ++dz) {
        for (int dy = -S; dy <= S; ++dy) {
            // stage the partner cells of the block's cells for this (dy, dz)
            const int2 st = stage_range(P, B, crange, c_lo, c_hi, dy, dz, NL_STAGE);
            const int lo = st.x, cnt = st.y;
            __syncthreads();                 // the previous row's stage is no longer read
            for (int i = threadIdx.x; i < cnt; i += BLK) stage[i] = src[lo + i];
            __syncthreads();
            const int z2 = cz + dz;
            if (!act || z2 < 0 || z2 >= nz) continue;
            int y2 = cy + dy;
            double sy = 0.0;                 // periodic image of the partner row
            if (y2 < 0) { y2 += ny; sy = -P.L[1]; } else if (y2 >= ny) { y2 -= ny; sy = P.L[1]; }
            const int row = P.cell_base + nx * (y2 + ny * z2);
            for (int dx = -S; dx <= S; ++dx) {
                int x2 = cx + dx;
                double sx = 0.0;
                if (x2 < 0) { x2 += nx; sx = -P.L[0]; } else if (x2 >= nx) { x2 -= nx; sx = P.L[0]; }
                int2 rg = crange[row + x2];
                if (!guard(B, rg.x >= 0 && rg.x <= rg.y && rg.y <= B.n_slots, G_SLOT, k)) rg = make_int2(0, 0);
                const int ecode = nl_code(dx, dy, dz, S);
                // the bead's position moved into the partner cell's frame
                // (mixed: minus the cell offset; FP64: minus the periodic image)
                const float qfx = pfi.x - (float)dx * (float)P.l_cell[0];
                const float qfy = pfi.y - (float)dy * (float)P.l_cell[1];
                const float qfz = pfi.z - (float)dz * (float)P.l_cell[2];
                const double qx = pi.x - sx, qy = pi.y - sy;
                // One group of four candidates from m0.  TAIL: the group may
                // run past the cell's end (only the cell's last group does);
                // SELF: the cell is the bead's own, the only one holding slot
                // k (boxes have at least 2S+1 cells per periodic axis), so
                // both tests leave the other groups' per-candidate work.
                auto group = [&](auto load, int m0, auto tail, auto self) {
                    bool take[NL_UNROLL];
                    Bead pj[NL_UNROLL];
#pragma unroll
                    for (int u = 0; u < NL_UNROLL; ++u) pj[u] = load(m0 + u);
#pragma unroll
                    for (int u = 0; u < NL_UNROLL; ++u) {
                        if constexpr (MIXED) {
                            const float ex = qfx - pj[u].x, ey = qfy - pj[u].y, ez = qfz - pj[u].z;
                            take[u] = __builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)) <
                                      rl2f[trow + (int)(__float_as_uint(pj[u].w) & 7u)];
                        } else {
                            const double ex = qx - pj[u].x, ey = qy - pj[u].y, ez = pi.z - pj[u].z;
                            take[u] = __builtin_fma(ez, ez, __builtin_fma(ey, ey, ex * ex)) <
                                      rl2d[trow + wtype(pj[u].w)];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < NL_UNROLL; ++u) {
                        const int m = m0 + u;
                        bool ok = take[u];
                        if constexpr (decltype(tail)::value) ok = ok && m < rg.y;
                        if constexpr (decltype(self)::value) ok = ok && m != k;
                        if (ok) {
                            int tj;
                            if constexpr (MIXED) tj = fwtype(pj[u].w); else tj = wtype(pj[u].w);
                            b0 = b1; b1 = b2; b2 = b3; b3 = nl_entry(m, tj, ecode);
                            if ((n & 3) == 3) {
                                if (n < cap) *gp = make_int4(b0, b1, b2, b3);
                                gp += group_step;
                            }
                            ++n;
                        }
                    }
                };
                // partners from the stage when the cell lies inside it, else
                // from memory (padded buffer); four loads in flight, loads past
                // the cell's end are not used
                auto scan = [&](auto load, auto self) {
                    int m0 = rg.x;
                    for (; m0 + NL_UNROLL <= rg.y; m0 += NL_UNROLL)
                        group(load, m0, std::false_type{}, self);
                    if (m0 < rg.y) group(load, m0, std::true_type{}, self);
                };
                const bool own = dx == 0 && dy == 0 && dz == 0;
                if (rg.x >= lo && rg.y <= lo + cnt) {
                    auto ld = [&](int m) { return stage[m - lo]; };
                    if (own) scan(ld, std::true_type{}); else scan(ld, std::false_type{});
                } else {
                    auto ld = [&](int m) { return src[m]; };
                    if (own) scan(ld, std::true_type{}); else scan(ld, std::false_type{});
                }
            }
        }
    }
    if (!act) return;
    // the last, partial group (the pair kernels read only its first n % 4)
    const int r = n & 3;
    if (r != 0 && n <= cap) *gp = r == 1 ? make_int4(b3, 0, 0, 0) : r == 2 ? make_int4(b2, b3, 0, 0)
                                                                  : make_int4(b1, b2, b3, 0);
    B.nl_n[k] = n;
    if (n > cap) atomicAdd(&B.flags[4], 1);
}

// The fused step's bonded epilogue of the pair kernels (mode 2): the 9-3
// walls and the FENE bonds of bead k (FP64 position ri, global index g, type
// ti) added to its pair force f in md_main.f90's order (fluid_wall :132,
// intra_molec :142), energies to e_wall and e_fene.
template <bool VIR>
__device__ __forceinline__ void bonded_epilogue(const SysDev &P, const Bufs &B, size_t cur_off, const double4 &ri,
                                                int ti, double4 &f, double &e_wall, double &e_fene, double *wb) {
    const int g = worig(ri.w);
    const int o = g - P.part_base;
    const bool brush = o < P.pd;
    const int nm = brush ? P.n_mon : P.n_mon_d;
    const int ol = brush ? o : o - P.pd;
    const int m = nm > 0 ? ol % nm : 0;
    double3 F = make_double3(f.x, f.y, f.z);
    wall_bead(P, ri, ti, F, e_wall);
    fene_bead<VIR>(P, B, cur_off, ri, g, m, nm, F, e_fene, wb);
    f.x = F.x; f.y = F.y; f.z = F.z;
}

// Energy partials of a pair-kernel block: the pair energy and count, the
// pair virial (zero when off), the FENE bond virial (the epilogue's in mode 2,
// else zero: the bonded kernel of the step writes it) and, in mode 2, the
// walls and bonds the epilogue added (bending columns zeroed; the bending
// kernel, if any, writes them).
template <bool VIR>
__device__ __forceinline__ void pair_partials(double *out, const PairAcc &a, int mode, double e_wall, double e_fene,
                                              const double *wb, double *red) {
    double es = block_sum(a.e, red);
    double nps = block_sum((double)a.np, red);
    if (threadIdx.x == 0) { out[C_VFF] = es; out[C_NPAIR] = nps; }
    for (int q = 0; q < 6; ++q) {       // zero when the virial is off
        double wq = VIR ? block_sum(a.w[q], red) : 0.0;
        double bq = (VIR && mode == 2) ? block_sum(wb[q], red) : 0.0;
        if (threadIdx.x == 0) { out[C_WXX + q] = wq; out[C_WBXX + q] = bq; }
    }
    if (mode == 2) {
        double ew = block_sum(e_wall, red);
        double ef = block_sum(e_fene, red);
        if (threadIdx.x == 0) { out[C_VWALL] = ew; out[C_VINTRA] = ef; out[C_VBEND] = 0.0; out[C_VBENDM] = 0.0; }
    }
}

// ---------------------------------------------------------------------------
// FP64 pair kernel.
// ---------------------------------------------------------------------------
template <bool DPD, bool VIR>
__global__ __launch_bounds__(BLK) void k_pair64(const SysDev *__restrict__ sys, const int *__restrict__ blk_sys,
                                               Bufs B, const int2 *__restrict__ crange,
                                               double *__restrict__ partial, int fused) {
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[blockIdx.x]);
    MFA_CUR(B, s);
    const double4 *__restrict__ pos = CUR(B.pos);
    const double4 *__restrict__ vel = CUR(B.vel);
    const int *__restrict__ cell_of = CUR(B.cell_of);
    double4 *__restrict__ frc = CUR(B.frc);
    const long long *t_dev = B.t_idx;
    int *rebuild = B.rebuild;
    __shared__ double tab[80];
    __shared__ double red[BLK];
    __shared__ unsigned dq_buf[DPD_Q * BLK];   // deferred DPD pairs: slot | lj << 31, [j * BLK + lane]
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
    const Pair64Ctx C = pair64_ctx(P, t_idx, r_2_min);
    // the fused step (mode 1, 2) sums the energy and counts the pairs only on
    // the last step of an md_steps call, the one whose force it stores and
    // whose energies a caller can read; block 0 publishes that decision for
    // the step's later kernels
    const bool calc_e = fused == 0 || *B.t_idx >= *B.t_end;
    if (fused != 0 && blockIdx.x == 0 && threadIdx.x == 0) *B.store_frc = calc_e ? 1 : 0;
    PairAcc a;
    a.fx = a.fy = a.fz = a.e = 0.0;
    for (int q = 0; q < 6; ++q) a.w[q] = 0.0;
    a.np = 0;
    double e_wall = 0.0, e_fene = 0.0;
    double wb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (active) {
        const double4 pi = pos[k];
        const int ti = wtype(pi.w), oi = worig(pi.w);
        guard(B, ti >= 1 && ti <= 4, G_TYPE, k);
        double vxi = 0.0, vyi = 0.0, vzi = 0.0;
        if (DPD) { double4 v = vel[k]; vxi = v.x; vyi = v.y; vzi = v.z; }
        unsigned *const myq = dq_buf + threadIdx.x;
        const int dcap = B.dpd_queue;
        int dq = 0;
        // LJ part of candidate m (position pj); a DPD pair is queued, or
        // evaluated in place when the queue is full
        auto visit = [&](int m, double4 pj) {
            const int f = pair64_lj<DPD, VIR>(C, tab, pi.x, pi.y, pi.z, ti, pj, calc_e, a);
            if (DPD && f) {
                if (dq < dcap) {
                    myq[dq * BLK] = (unsigned)m | (f == 1 ? 0x80000000u : 0u);
                    ++dq;
                } else {
                    pair64_dpd<VIR>(C, tab, pi.x, pi.y, pi.z, ti, oi, vxi, vyi, vzi, pj, vel[m], f == 1, a);
                }
            }
        };
        const int nn = B.nl_cap > 0 ? B.nl_n[k] : -1;
        if (nn >= 0 && nn <= B.nl_cap) {
            const unsigned nsl = (unsigned)B.n_slots;
            for (int q0 = 0; q0 < nn; q0 += NL_UNROLL) {
                unsigned mm[NL_UNROLL];
                double4 pj[NL_UNROLL];
                const int4 e4 = nl_group(B.nl, nl_index(k, q0, (size_t)B.n_slots));
                // unsigned slots: one compare is the range check (as in k_pair32)
                mm[0] = (unsigned)e4.x >> NL_SLOT_SHIFT; mm[1] = (unsigned)e4.y >> NL_SLOT_SHIFT;
                mm[2] = (unsigned)e4.z >> NL_SLOT_SHIFT; mm[3] = (unsigned)e4.w >> NL_SLOT_SHIFT;
#pragma unroll
                for (int u = 0; u < NL_UNROLL; ++u) pj[u] = ld_slot(pos, min(mm[u], nsl - 1u));
#pragma unroll
                for (int u = 0; u < NL_UNROLL; ++u) {
                    if (q0 + u < nn && guard(B, mm[u] < nsl, G_SLOT, k)) visit((int)mm[u], pj[u]);
                }
            }
        } else {
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
                    for (int dx = -S; dx <= S; ++dx) {
                        int x2 = cx + dx;
                        x2 += (x2 < 0) ? nx : 0;
                        x2 -= (x2 >= nx) ? nx : 0;
                        int2 rg = crange[row + x2];
                        if (!guard(B, rg.x >= 0 && rg.x <= rg.y && rg.y <= B.n_slots, G_SLOT, k)) rg = make_int2(0, 0);
                        for (int m = rg.x; m < rg.y; ++m) {
                            if (m == k) continue;
                            visit(m, pos[m]);
                        }
                    }
                }
            }
        }
        // the queued DPD parts, in walk order
        if (DPD) {
            for (int j0 = 0; j0 < dq; j0 += DPD_UNROLL) {
                unsigned e[DPD_UNROLL];
                double4 pj[DPD_UNROLL], vj[DPD_UNROLL];
#pragma unroll
                for (int u = 0; u < DPD_UNROLL; ++u) e[u] = myq[min(j0 + u, dq - 1) * BLK];
#pragma unroll
                for (int u = 0; u < DPD_UNROLL; ++u) {
                    const int m = (int)(e[u] & 0x7fffffffu);
                    pj[u] = pos[m];
                    vj[u] = vel[m];
                }
#pragma unroll
                for (int u = 0; u < DPD_UNROLL; ++u) {
                    if (j0 + u < dq)
                        pair64_dpd<VIR>(C, tab, pi.x, pi.y, pi.z, ti, oi, vxi, vyi, vzi, pj[u], vj[u],
                                        (e[u] >> 31) != 0, a);
                }
            }
        }
        double4 f = make_double4(a.fx, a.fy, a.fz, 0.0);
        if (P.thermostat == 2) lgv_bead(P, (uint64_t)t_idx, oi - P.part_base, ti, vel[k], f);
        if (fused == 2) bonded_epilogue<VIR>(P, B, cur_off, pi, ti, f, e_wall, e_fene, wb);
        frc[k] = f;
    }
    if (!calc_e) return;
    pair_partials<VIR>(partial + (size_t)blockIdx.x * NPART, a, fused, e_wall, e_fene, wb, red);
}

// ---------------------------------------------------------------------------
// Mixed-precision pair kernel: positions relative to the build-time cell
// origin in FP32 (posf), FP32 pair arithmetic, each pair force added into
// FP64 accumulators.  Pairs whose FP32 r^2 lies within a relative band of
// 2^-14 around the cutoff are decided again in FP64 with the reference
// arithmetic, so the interacting-pair set is the FP64 one.
// ---------------------------------------------------------------------------
struct Pair32Ctx {
    float r_2_min, rdpd2, inv_rdpd, sigf, gamf;
    uint64_t t_idx, seed;
    int part_base, dpd_cut, fso;      // channel scalars, copied once (see Pair64Ctx)
};

// Separation of candidate pj from bead pf in FP32 cell-frame coordinates, o
// being the partner cell's offset (d * l_cell, exact for |d| <= 2): the
// difference of the two cell-frame coordinates first, so the partner's
// evaluation computes exactly the negated vector.  r^2 is written with
// explicit fused multiply-adds (and contraction is off elsewhere in the pair
// arithmetic), so every evaluation of a pair -- list walk, cell scan, queued
// DPD part -- rounds the same way.
struct Sep32 { float x, y, z, r2; };
__device__ __forceinline__ Sep32 sep32(float4 pf, float4 pj, float ox, float oy, float oz) {
#pragma clang fp contract(off)
    Sep32 d;
    d.x = (pf.x - pj.x) - ox;
    d.y = (pf.y - pj.y) - oy;
    d.z = (pf.z - pj.z) - oz;
    d.r2 = __builtin_fmaf(d.z, d.z, __builtin_fmaf(d.y, d.y, d.x * d.x));
    return d;
}

// Per-channel pair tables of the mixed kernel in LDS, one row of 8 entries
// per bead type and one entry per partner type bits (the low 3 bits of a
// posf identity word, types 1..4), so a partner's entry is
// T[field * 32 + (((ti - 1) & 3) * 8 | (w & 7))]: one index per candidate
// for every field.
enum T32Field { F_REACH = 0, F_LO, F_HI, F_EPS, F_SIG2, F_ESH, F_DPD, F_IRC, F_NF };

// LJ part of candidate pj (slot m) for bead k, and whether the pair takes a
// DPD term, with in_lj whether it lies within the LJ range (a DPD pair
// beyond it: DPD_CUT_OFF only).  Two flags rather than one three-valued
// result, so both stay lane masks.  ix: the candidate's table index.  The
// early exit compares r^2 with the reach, max(LJ range widened by 2^-14, DPD
// range widened by 1e-4).
template <bool DPD, bool VIR>
__device__ __forceinline__ bool pair32_lj(const SysDev &P, const float *T, const Pair32Ctx &X,
                                          const double4 *__restrict__ pos, int k, int ti, float4 pf,
                                          float ox, float oy, float oz, float4 pj, int m, int ix, bool calc_e,
                                          PairAcc &a, bool &in_lj) {
#pragma clang fp contract(off)
    const Sep32 d = sep32(pf, pj, ox, oy, oz);
    float r2 = d.r2;
    in_lj = false;
    if (!(r2 < T[F_REACH * 32 + ix])) return false;
    bool lj = r2 < T[F_LO * 32 + ix];
    if (!lj && r2 < T[F_HI * 32 + ix]) {
        // borderline: decide with the FP64 reference arithmetic
        const int tt = ((ti - 1) & 3) + 4 * (((ix & 7) - 1) & 3);
        const double4 pa = pos[k], pb = pos[m];
        double ex = mic(pa.x - pb.x, P.L[0], P.invL[0]);
        double ey = mic(pa.y - pb.y, P.L[1], P.invL[1]);
        double ez = pa.z - pb.z;
        lj = r2_exact(ex, ey, ez) < P.range_2[tt];
    }
    if (lj) {
        if (X.fso && r2 < X.r_2_min) r2 = X.r_2_min;
        const float inv_r_2 = __builtin_amdgcn_rcpf(r2);
        const float l_eps = T[F_EPS * 32 + ix];
        const float r_61 = T[F_SIG2 * 32 + ix] * inv_r_2;
        const float r_6 = r_61 * r_61 * r_61;
        const float r_12 = r_6 * r_6;
        if (calc_e) {
            a.np += 1;
            a.e += (double)(l_eps * ((r_12 - r_6) - T[F_ESH * 32 + ix]));
        }
        const float rd = l_eps * __builtin_fmaf(6.0f, r_6, -12.0f * r_12) * inv_r_2;
        const float fx = -rd * d.x, fy = -rd * d.y, fz = -rd * d.z;
        a.fx += (double)fx; a.fy += (double)fy; a.fz += (double)fz;
        if (VIR) {
            a.w[0] += (double)(d.x * fx); a.w[1] += (double)(d.y * fy); a.w[2] += (double)(d.z * fz);
            a.w[3] += (double)(d.x * fy); a.w[4] += (double)(d.x * fz); a.w[5] += (double)(d.y * fz);
        }
    }
    in_lj = lj;
    if (!DPD) return false;
    // DPD_CUT_OFF: the F_DPD entry is the DPD range^2 for non-head pairs and
    // -1 for pairs with a head (or a channel without the DPD thermostat)
    return X.dpd_cut ? (r2 < T[F_DPD * 32 + ix]) : lj;
}

// DPD part of a pair that pair32_lj flagged (dpd_forces.f90:45-115).  lj: the
// pair is within the LJ range, where the switch-on clamp of r^2 applies as in
// the LJ part.
template <bool VIR>
__device__ __forceinline__ void pair32_dpd(const float *T, const Pair32Ctx &X, int oi, float4 pf, float4 vi,
                                           float ox, float oy, float oz, float4 pj, float4 vj, bool lj, int ix,
                                           PairAcc &a) {
#pragma clang fp contract(off)
    const Sep32 d = sep32(pf, pj, ox, oy, oz);
    float r2 = d.r2;
    if (lj && X.fso && r2 < X.r_2_min) r2 = X.r_2_min;
    const float inv_sqrt = __builtin_amdgcn_rsqf(r2);
    float wgt;
    if (X.dpd_cut) {
        wgt = 1.0f - r2 * inv_sqrt * X.inv_rdpd;
    } else {
        wgt = 1.0f - r2 * inv_sqrt * T[F_IRC * 32 + ix];
    }
    const float g = pair_gauss(X.seed, X.t_idx, oi - X.part_base, fworig(pj.w) - X.part_base);
    const float rvx = d.x * inv_sqrt, rvy = d.y * inv_sqrt, rvz = d.z * inv_sqrt;
    const float rdv = __builtin_fmaf(vi.z - vj.z, rvz, __builtin_fmaf(vi.y - vj.y, rvy, (vi.x - vj.x) * rvx));
    const float sc = __builtin_fmaf(X.sigf * wgt, g, -(X.gamf * (wgt * wgt) * rdv));
    const float fx = sc * rvx, fy = sc * rvy, fz = sc * rvz;
    a.fx += (double)fx; a.fy += (double)fy; a.fz += (double)fz;
    if (VIR) {
        a.w[0] += (double)(d.x * fx); a.w[1] += (double)(d.y * fy); a.w[2] += (double)(d.z * fz);
        a.w[3] += (double)(d.x * fy); a.w[4] += (double)(d.x * fz); a.w[5] += (double)(d.y * fz);
    }
}

// One tile of 256 slots (the tile-th block's slots) of the mixed pair kernel.
template <bool DPD, bool VIR>
__device__ __forceinline__ void pair32_tile(const int tile, const SysDev *__restrict__ sys,
                                            const int *__restrict__ blk_sys, const Bufs &B,
                                            const int2 *__restrict__ crange, double *__restrict__ partial,
                                            int fused) {
    const int s = __builtin_amdgcn_readfirstlane(blk_sys[tile]);
    MFA_CUR(B, s);
    const double4 *__restrict__ pos = CUR(B.pos);
    const float4 *__restrict__ posf = CUR(B.posf);
    const float4 *__restrict__ velf = CUR(B.velf);
    const int *__restrict__ cell_of = CUR(B.cell_of);
    double4 *__restrict__ frc = CUR(B.frc);
    const long long *t_dev = B.t_idx;
    int *rebuild = B.rebuild;
    __shared__ float T[F_NF * 32];
    __shared__ float4 off[128];         // cell offset (dx lx, dy ly, dz lz) per list code
    __shared__ double red[BLK];
    __shared__ int dq_m[DPD_Q * BLK];              // deferred DPD pairs: partner slot, [j * BLK + lane]
    __shared__ unsigned short dq_c[DPD_Q * BLK];   // and its cell-offset code | lj << 7
    const SysDev &P = sys[s];
    const int S = P.S;
    const float lx = (float)P.l_cell[0], ly = (float)P.l_cell[1], lz = (float)P.l_cell[2];
    Pair32Ctx X;
    X.t_idx = (uint64_t)*t_dev;
    X.r_2_min = 0.0f;
    if (P.force_switch_on) {
        const double r_time = (double)X.t_idx * P.dt;
        double r_2_min_d = 0.0;
        if (r_time < P.r_2_min_time) r_2_min_d = (1.0 - r_time / P.r_2_min_time) * P.r_2_min_init;
        X.r_2_min = (float)r_2_min_d;
    }
    X.sigf = (float)P.sig; X.gamf = (float)P.friction;
    X.seed = P.seed; X.part_base = P.part_base; X.dpd_cut = P.dpd_cut_off; X.fso = P.force_switch_on;
    X.rdpd2 = (float)P.r_cut_dpd_2;
    X.inv_rdpd = (float)(1.0 / sqrt(P.r_cut_dpd_2 > 0.0 ? P.r_cut_dpd_2 : 1.0));
    // a channel whose thermostat is not DPD flags no pair for a DPD part (as
    // in pair64_ctx: the DPD_CUT_OFF rule with an empty range)
    if (P.thermostat != 1) { X.dpd_cut = 1; X.rdpd2 = -1.0f; }
    {
        const int t = threadIdx.x;
        if (t < 32) {
            // entry (bead type row, partner type bits): type bits outside 1..4
            // (padding) never reach
            const int row = t >> 3, tb = t & 7;
            const bool valid = tb >= 1 && tb <= 4;
            const int tt = row + 4 * ((tb - 1) & 3);
            const bool head_pair = row == 0 || tb == 1;
            const float hi = P.range_2f_hi[tt];
            const float dpd2 = (X.dpd_cut && !head_pair) ? X.rdpd2 : -1.0f;
            const float reach = (DPD && X.dpd_cut && !head_pair) ? fmaxf(hi, X.rdpd2 * 1.0001f) : hi;
            T[F_REACH * 32 + t] = valid ? reach : -1.0f;
            T[F_LO * 32 + t] = valid ? P.range_2f_lo[tt] : -1.0f;
            T[F_HI * 32 + t] = valid ? hi : -1.0f;
            T[F_EPS * 32 + t] = P.epsil_f[tt];
            T[F_SIG2 * 32 + t] = P.sigma_2f[tt];
            T[F_ESH * 32 + t] = P.e_shift_f[tt];
            T[F_DPD * 32 + t] = valid ? dpd2 : -1.0f;
            T[F_IRC * 32 + t] = P.inv_rc_f[tt];
        }
        if (t < 128) {
            const int w = 2 * S + 1;
            const int dx = t % w - S, dy = (t / w) % w - S, dz = t / (w * w) - S;
            off[t] = make_float4((float)dx * lx, (float)dy * ly, (float)dz * lz, 0.f);
        }
    }
    if (tile == 0 && threadIdx.x == 0) *rebuild = 0;
    if (tile == P.block_base && threadIdx.x == 0) B.chan_rb[s] = 0;
    // energy and pair count only on a fused call's last step (k_pair64)
    const bool calc_e = fused == 0 || *B.t_idx >= *B.t_end;
    if (fused != 0 && tile == 0 && threadIdx.x == 0) *B.store_frc = calc_e ? 1 : 0;
    __syncthreads();
    const int k = tile * BLK + threadIdx.x;
    const bool active = (k - P.slot_base) < P.n_part;
    PairAcc a;
    a.fx = a.fy = a.fz = a.e = 0.0;
    for (int q = 0; q < 6; ++q) a.w[q] = 0.0;
    a.np = 0;
    double e_wall = 0.0, e_fene = 0.0;
    double wb[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (active) {
        const float4 pf = posf[k];
        const int ti = fwtype(pf.w), oi = fworig(pf.w);
        guard(B, ti >= 1 && ti <= 4, G_TYPE, k);
        const int trow8 = ((ti - 1) & 3) * 8;
        float4 vi = make_float4(0.f, 0.f, 0.f, 0.f);
        if (DPD) vi = velf[k];
        int *const myq = dq_m + threadIdx.x;
        unsigned short *const myc = dq_c + threadIdx.x;
        const int dcap = B.dpd_queue;
        int dq = 0;
        // LJ part of candidate m (position pj, cell-offset code, table index
        // ix); a DPD pair is queued, or evaluated in place when the queue is
        // full (the partner's identity word reloaded with its position)
        auto visit = [&](int m, int code, float4 pj, int ix) {
            const float4 o = off[code];
            bool in_lj;
            const bool f = pair32_lj<DPD, VIR>(P, T, X, pos, k, ti, pf, o.x, o.y, o.z, pj, m, ix, calc_e, a, in_lj);
            if (DPD && f) {
                if (dq < dcap) {
                    myq[dq * BLK] = m;
                    myc[dq * BLK] = (unsigned short)(code | (in_lj ? 0x80 : 0));
                    ++dq;
                } else {
                    pair32_dpd<VIR>(T, X, oi, pf, vi, o.x, o.y, o.z, posf[m], velf[m], in_lj, ix, a);
                }
            }
        };
        const int nn = B.nl_cap > 0 ? B.nl_n[k] : -1;
        if (nn >= 0 && nn <= B.nl_cap) {
            // list entries are slots below n_slots: k_nlist takes them from
            // checked cell ranges and writes the unused entries of a bead's
            // partial last group as zeros, so every entry gathered is in range
            // one 16-byte group of four entries, their four gathers in
            // flight together; TAIL: the list's last, partial group.  The
            // partner's type comes with the entry, so only its three cell-
            // frame coordinates are gathered (12 of the 16 bytes).  The next
            // group's entries are loaded after this group's gathers are
            // issued, so their memory latency overlaps this group's pair
            // arithmetic (the group past a bead's last one, clamped to the
            // row's last group, is loaded and not used).
            const int trow8p1 = trow8 + 1;
            const int q_last = B.nl_cap - NL_UNROLL;
            int4 e4 = nl_group(B.nl, nl_index(k, 0, (size_t)B.n_slots));
            auto group = [&](int q0, auto tail) {
                const int ent[NL_UNROLL] = {e4.x, e4.y, e4.z, e4.w};
                unsigned mm[NL_UNROLL];
                float3 pj[NL_UNROLL];
#pragma unroll
                for (int u = 0; u < NL_UNROLL; ++u) {
                    mm[u] = (unsigned)ent[u] >> NL_SLOT_SHIFT;
                    pj[u] = *(const float3 *)((const char *)posf + mm[u] * (unsigned)sizeof(float4));
                }
                if (!decltype(tail)::value)
                    e4 = nl_group(B.nl, nl_index(k, min(q0 + NL_UNROLL, q_last), (size_t)B.n_slots));
#pragma unroll
                for (int u = 0; u < NL_UNROLL; ++u) {
                    if (!decltype(tail)::value || q0 + u < nn)
                        visit((int)mm[u], ent[u] & NL_CODE_MASK, make_float4(pj[u].x, pj[u].y, pj[u].z, 0.f),
                              trow8p1 + ((ent[u] >> NL_TYPE_SHIFT) & 3));
                }
            };
            int q0 = 0;
            for (; q0 + NL_UNROLL <= nn; q0 += NL_UNROLL) group(q0, std::false_type{});
            if (q0 < nn) group(q0, std::true_type{});
        } else {
            int c = cell_of[k] - P.cell_base;
            if (!guard(B, in_range(c, P.n_cells_tot), G_CELL, k)) c = 0;
            const int nx = P.nc[0], ny = P.nc[1], nz = P.nc[2];
            const int cx = c % nx, cy = (c / nx) % ny, cz = c / (nx * ny);
            for (int dz = -S; dz <= S; ++dz) {
                const int z2 = cz + dz;
                if (z2 < 0 || z2 >= nz) continue;
                for (int dy = -S; dy <= S; ++dy) {
                    int y2 = cy + dy;
                    y2 += (y2 < 0) ? ny : 0;
                    y2 -= (y2 >= ny) ? ny : 0;
                    const int row = P.cell_base + nx * (y2 + ny * z2);
                    for (int dx = -S; dx <= S; ++dx) {
                        int x2 = cx + dx;
                        x2 += (x2 < 0) ? nx : 0;
                        x2 -= (x2 >= nx) ? nx : 0;
                        int2 rg = crange[row + x2];
                        if (!guard(B, rg.x >= 0 && rg.x <= rg.y && rg.y <= B.n_slots, G_SLOT, k)) rg = make_int2(0, 0);
                        const int code = nl_code(dx, dy, dz, S);
                        for (int m = rg.x; m < rg.y; ++m) {
                            if (m == k) continue;
                            const float4 pj = posf[m];
                            visit(m, code, pj, trow8 | (int)(__float_as_uint(pj.w) & 7u));
                        }
                    }
                }
            }
        }
        // the queued DPD parts, in walk order
        if (DPD) {
            for (int j0 = 0; j0 < dq; j0 += DPD_UNROLL) {
                int mm[DPD_UNROLL], cc[DPD_UNROLL];
                float4 pj[DPD_UNROLL], vj[DPD_UNROLL];
#pragma unroll
                for (int u = 0; u < DPD_UNROLL; ++u) {
                    const int j = min(j0 + u, dq - 1);
                    mm[u] = myq[j * BLK];
                    cc[u] = myc[j * BLK];
                }
#pragma unroll
                for (int u = 0; u < DPD_UNROLL; ++u) {
                    pj[u] = posf[mm[u]];
                    vj[u] = velf[mm[u]];
                }
#pragma unroll
                for (int u = 0; u < DPD_UNROLL; ++u) {
                    if (j0 + u < dq) {
                        const float4 o = off[cc[u] & 0x7f];
                        const int ix = trow8 | (int)(__float_as_uint(pj[u].w) & 7u);
                        pair32_dpd<VIR>(T, X, oi, pf, vi, o.x, o.y, o.z, pj[u], vj[u], (cc[u] & 0x80) != 0, ix, a);
                    }
                }
            }
        }
        double4 f = make_double4(a.fx, a.fy, a.fz, 0.0);
        if (P.thermostat == 2) lgv_bead(P, X.t_idx, oi - P.part_base, ti, CUR(B.vel)[k], f);
        if (fused == 2) bonded_epilogue<VIR>(P, B, cur_off, pos[k], ti, f, e_wall, e_fene, wb);
        frc[k] = f;
    }
    if (!calc_e) return;
    pair_partials<VIR>(partial + (size_t)tile * NPART, a, fused, e_wall, e_fene, wb, red);
}

// Six waves per SIMD: caps the kernel at 80 VGPRs (91 without the hint, five
// waves) for a 12-byte spill; measured 5-9 % faster on the pair step at 0.3-5M
// beads.  Seven waves (72 VGPRs, 40-byte spill) is 60 % slower, and five waves
// forced on k_pair64 (96 VGPRs) is slower than its natural four.
//
// PERSIST false: one tile per block.  PERSIST true: a persistent launch: the
// blocks take tiles from the work counter tiles[0] until it passes n_tiles
// (every block's last take does, so every block leaves the loop); the last
// block to leave resets tiles[0] and tiles[1] (its ticket counter) to 0 for
// the next launch.
template <bool DPD, bool VIR, bool PERSIST>
__global__ __launch_bounds__(BLK) __attribute__((amdgpu_waves_per_eu(PERSIST ? 5 : 6, 8))) void k_pair32(
    const SysDev *__restrict__ sys, const int *__restrict__ blk_sys, Bufs B, const int2 *__restrict__ crange,
    double *__restrict__ partial, int fused, int *tiles, int n_tiles) {
    if constexpr (!PERSIST) {
        pair32_tile<DPD, VIR>(blockIdx.x, sys, blk_sys, B, crange, partial, fused);
        return;
    }
    __shared__ int next;
    for (;;) {
        // every thread is past its previous tile (and its reads of the LDS
        // tables) once it reaches this barrier
        __syncthreads();
        if (threadIdx.x == 0) next = atomicAdd(&tiles[0], 1);
        __syncthreads();
        const int tile = next;
        if (tile >= n_tiles) break;
        pair32_tile<DPD, VIR>(tile, sys, blk_sys, B, crange, partial, fused);
    }
    if (threadIdx.x == 0 && atomicAdd(&tiles[1], 1) == (int)gridDim.x - 1) {
        tiles[0] = 0;
        tiles[1] = 0;
    }
}

}  // namespace mfa
