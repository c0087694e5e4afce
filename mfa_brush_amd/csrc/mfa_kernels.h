// mfa_kernels.h -- device-side data and arithmetic shared by the gfx950 kernels.
//
// Layout in HBM (one context, possibly many channels):
//   slots   : every channel owns a 256-aligned slot range; within it the
//             beads are kept sorted by cell (counting sort), ascending
//             reference index inside a cell.  Padding slots are never in a
//             cell.
//   pos     : double4 (x, y, z, w) with w carrying the bits of
//             (global bead index << 3 | type), so one 32-byte load gives a
//             bead's position, identity and type.
//   vel/acc/frc : double4 (x, y, z, unused).
//   posf/velf   : float4 copies for the mixed path; posf holds the position
//             relative to the origin of the bead's build-time cell.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mfa {

constexpr int BLK = 256;          // threads per block and slot-range alignment
constexpr int NPART = 24;         // partial-sum columns per block
// energies, kinetic energy, pair virial (C_W*), pair count, FENE bond virial (C_WB*)
enum Col { C_VFF = 0, C_VWALL, C_VINTRA, C_VBEND, C_VBENDM, C_KE,
           C_WXX, C_WYY, C_WZZ, C_WXY, C_WXZ, C_WYZ, C_NPAIR,
           C_WBXX, C_WBYY, C_WBZZ, C_WBXY, C_WBXZ, C_WBYZ, C_NCOL };
static_assert(C_NCOL <= NPART, "partial-sum columns");

struct SysDev {
    double L[3], invL[3];
    double l_cell[3], inv_l_cell[3];
    double epsil[16], sigma_2[16], range_2[16], e_shift[16], inv_range_2[16];
    float  epsil_f[16], sigma_2f[16], range_2f_lo[16], range_2f_hi[16], e_shift_f[16], inv_rc_f[16];
    double r_cut_dpd_2, sig, friction, r_2_min_init, r_2_min_time;
    double sigma_wall[4], a_wall[4], a_w;
    double k_chain, inv_r_chain_2, k_bend, alpha_eq;
    double dt, dt_2;
    double const_force[3], head_vel[2];
    double mass_type[4], inv_mass_type[4];
    double skin_half_2;
    double rl2[16];               // Verlet-list radius^2 per type pair (range + skin), < 0: never listed
    uint64_t seed;
    int nc[3], n_cells_tot, cell_base;
    int slot_base, n_part, part_base;
    int n_mon, n_chain, n_mon_d, n_chain_d, pd, head_base;
    int thermostat, dpd_cut_off, force_switch_on, bending, bending_melt, poiseuille, bend_form, dpd_vv;
    int S, skin_mode;
    int block_base, n_blocks;
};

// Double-buffered bead arrays: a rebuild permutes a channel into the other
// half and the permute kernel's last block flips that channel's `parity`, so no copy-back
// pass is needed, channels that do not rebuild are not touched, and the host
// never has to know whether the device decided to rebuild.  Each array is one
// allocation of 2 * n_slots entries; channel s's current half starts at
// parity[s] * n_slots.  The halves are reached by offset arithmetic on a single
// pointer, never by choosing between pointer fields of this struct at run
// time: the AMDGPU backend of this toolchain left the pointer register
// undefined on one branch of a run-time choice between pointers indexed
// dynamically out of the by-value kernel-argument struct.
struct Bufs {
    double4 *pos, *vel, *acc, *posb;
    float4 *posf, *velf;
    int *cell_of;
    double4 *frc;
    int *where;
    int *parity;        // per channel: which half holds the current layout
    int *rebuild;       // any channel rebuilds this step (gates the rebuild kernels)
    int *chan_rb;       // per channel: this channel's cells are rebuilt this step
    int *flags;         // [0..3] anomaly counters, [4] list overflows, [6] offending index, [7] first guard code
    long long *t_idx;
    long long *t_end;   // md_steps: time index of the call's last step
    int *store_frc;     // fused step: k_bonded stores the total force (set by k_vv1)
    int *nl;            // Verlet list: entry q of slot k at nl[((q / 4) * n_slots + k) * 4 + q % 4]
    int *nl_n;          // list length per slot (> nl_cap: overflowed, the cells are scanned)
    int nl_cap;         // entries per slot; 0: no list, the pair kernels scan the cells
    int dpd_queue;      // DPD pairs a pair-kernel lane defers (0..DPD_Q); the rest are evaluated in place
    int n_slots, n_beads, n_cells, n_heads;   // buffer extents, for the guards
};

#define MFA_CUR(B, s)                                                     \
    const int par = (B).parity[s] & 1;                                    \
    const size_t cur_off = (size_t)par * (size_t)(B).n_slots;             \
    const size_t nxt_off = (size_t)(par ^ 1) * (size_t)(B).n_slots;       \
    (void)cur_off; (void)nxt_off
#define CUR(arr) ((arr) + cur_off)
#define NXT(arr) ((arr) + nxt_off)

// Device-side guards.  Every data-dependent index (cell ids, cell ranges,
// permutations, bead identities, chain neighbours) is checked against the
// extent of the buffer it addresses before it is used; a bad index records
// the first failure (code, index) in flags[7], flags[6] and the access is
// skipped, so corrupted data is reported by the host instead of faulting
// the device.
enum GuardCode { G_CELL = 1, G_RANGE, G_PERM, G_ORIG, G_WHERE, G_TYPE, G_HEAD, G_BIN, G_SLOT };

__device__ __forceinline__ bool guard(const Bufs &B, bool ok, int code, long long idx) {
    if (!ok) {
        if (atomicCAS(&B.flags[7], 0, code) == 0) B.flags[6] = (int)idx;
    }
    return ok;
}
__device__ __forceinline__ bool in_range(long long v, long long n) { return v >= 0 && v < n; }

__device__ __forceinline__ uint64_t wbits(double w) { return (uint64_t)__double_as_longlong(w); }
__device__ __forceinline__ int wtype(double w) { return (int)(wbits(w) & 7ull); }
__device__ __forceinline__ int worig(double w) { return (int)(wbits(w) >> 3); }
__device__ __forceinline__ double mkw(int orig, int type) {
    return __longlong_as_double((long long)(((uint64_t)orig << 3) | (uint64_t)type));
}
__device__ __forceinline__ int fwtype(float w) { return (int)(__float_as_uint(w) & 7u); }
__device__ __forceinline__ int fworig(float w) { return (int)(__float_as_uint(w) >> 3); }
__device__ __forceinline__ float mkwf(int orig, int type) { return __uint_as_float(((uint32_t)orig << 3) | (uint32_t)type); }

// ---------------------------------------------------------------------------
// IEEE-exact pieces.  The pair decision and the integrator are written with
// FP contraction off so they compute what the reference's x86-64 build
// computes (no FMA): minimum image d - L*int(2 d/L) (fluid_fluid.f90:111-112),
// r^2 = (dx^2+dy^2)+dz^2 (:117).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double mic(double d, double L, double invL) {
#pragma clang fp contract(off)
    return d - L * (double)(int)(2.0 * d * invL);
}
__device__ __forceinline__ double r2_exact(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// Philox-4x32-10 (Salmon et al., SC'11).  Each 32x32 product is taken as one
// 64-bit product (one v_mad_u64_u32) rather than separate high and low
// halves (two quarter-rate multiplies).
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * (uint64_t)c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * (uint64_t)c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0;
        uint32_t n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

// Box-Muller normal from two 32-bit words with IEEE +,-,*,/,sqrt only (FP32,
// correctly rounded divide and square root, no contraction), so the value is
// the same bits on the host and on the device.
//
// The FP32 divide and square root are taken from FP64 and rounded once: the
// HIP __fsqrt_rn of this toolchain is the native v_sqrt_f32 (not correctly
// rounded) unless OCML_BASIC_ROUNDED_OPERATIONS is defined.  Rounding a
// correctly rounded (or few-ulp) FP64 quotient or root to FP32 gives the
// correctly rounded FP32 result: for a root, 4x - (2k+1)^2 is a nonzero odd
// integer, so sqrt(x) lies ~2^24 FP64 ulps from any FP32 midpoint; for a
// quotient, 53 >= 2*24+1 makes the double rounding innocuous.
__device__ __forceinline__ float div_rn_f32(float a, float b) { return (float)((double)a / (double)b); }
__device__ __forceinline__ float sqrt_rn_f32(float a) { return (float)sqrt((double)a); }

__device__ __forceinline__ float gauss_from_bits(uint32_t x0, uint32_t x1) {
#pragma clang fp contract(off)
    float u1 = (float)((x0 >> 8) + 1u) * 5.9604644775390625e-08f;
    uint32_t b = __float_as_uint(u1);
    int e = (int)(b >> 23) - 127;
    float m = __uint_as_float((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float s = div_rn_f32(m - 1.0f, m + 1.0f);
    float s2 = s * s;
    float p = 0.222222222f;
    p = p * s2 + 0.285714286f;
    p = p * s2 + 0.4f;
    p = p * s2 + 0.666666667f;
    p = p * s2 + 2.0f;
    float lnu = (float)e * 0.693147182f + s * p;
    float rad = sqrt_rn_f32(-2.0f * lnu);
    uint32_t q = x1 >> 8;
    uint32_t k = (q + (1u << 21)) >> 22;
    float f = (float)((int32_t)q - (int32_t)(k << 22)) * 2.38418579e-07f;
    float a = f * 1.57079633f;
    float a2 = a * a;
    float c = 2.48015873e-05f;
    c = c * a2 - 1.38888889e-03f;
    c = c * a2 + 4.16666667e-02f;
    c = c * a2 - 0.5f;
    c = c * a2 + 1.0f;
    float sn = 2.75573192e-06f;
    sn = sn * a2 - 1.98412698e-04f;
    sn = sn * a2 + 8.33333333e-03f;
    sn = sn * a2 - 1.66666667e-01f;
    sn = sn * a2 + 1.0f;
    sn = sn * a;
    uint32_t kk = k & 3u;
    float cs = kk == 0u ? c : (kk == 1u ? -sn : (kk == 2u ? -c : sn));
    return rad * cs;
}

// The DPD noise of the unordered pair {i, j} at time index t: one normal,
// identical for (i,j) and (j,i), so the random pair force is exactly
// antisymmetric although each bead evaluates the pair itself.
__device__ __forceinline__ float pair_gauss(uint64_t seed, uint64_t t, int i, int j) {
    uint32_t c[4];
    c[0] = (uint32_t)(i < j ? i : j);
    c[1] = (uint32_t)(i < j ? j : i);
    c[2] = (uint32_t)t;
    c[3] = (uint32_t)(t >> 32);
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return gauss_from_bits(c[0], c[1]);
}

// Langevin noise of bead i along axis d at time index t (lgv_forces.f90:14-20,
// one normal per (i, d, t)).  The counter's second word carries d with the
// top bit set, so it never equals a pair counter (second word: a bead index).
__device__ __forceinline__ float lgv_gauss(uint64_t seed, uint64_t t, int i, int d) {
    uint32_t c[4];
    c[0] = (uint32_t)i;
    c[1] = 0x80000000u | (uint32_t)d;
    c[2] = (uint32_t)t;
    c[3] = (uint32_t)(t >> 32);
    philox4x32_10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    return gauss_from_bits(c[0], c[1]);
}

// Langevin force on bead i (local index) of type ti, lgv_forces.f90:14-25:
// (sig g_d) sqrt(m) - (gamma m) v_d, added to f.
__device__ __forceinline__ void lgv_bead(const SysDev &P, uint64_t t, int i, int ti, double4 v, double4 &f) {
#pragma clang fp contract(off)
    const double m = P.mass_type[(ti - 1) & 3];
    const double sm = sqrt(m);
    const double vv[3] = {v.x, v.y, v.z};
    double add[3];
    for (int d = 0; d < 3; ++d) {
        double vd = P.sig * (double)lgv_gauss(P.seed, t, i, d) * sm;
        vd = vd - P.friction * m * vv[d];
        add[d] = vd;
    }
    f.x = f.x + add[0];
    f.y = f.y + add[1];
    f.z = f.z + add[2];
}

__device__ __forceinline__ double ipow(double x, int n) {
    double r = 1.0;
    while (n) { if (n & 1) r *= x; x *= x; n >>= 1; }
    return r;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// Cell of a position (get_cell, cell_list.f90:186-213), clamped into the grid.
__device__ __forceinline__ int cell_of_pos(const SysDev &P, double x, double y, double z) {
    int cx = clampi((int)(x * P.inv_l_cell[0]), 0, P.nc[0] - 1);
    int cy = clampi((int)(y * P.inv_l_cell[1]), 0, P.nc[1] - 1);
    int cz = clampi((int)(z * P.inv_l_cell[2]), 0, P.nc[2] - 1);
    return P.cell_base + cx + P.nc[0] * (cy + P.nc[1] * cz);
}

}  // namespace mfa
