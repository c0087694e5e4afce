// mfa_bonded.h -- per-bead bonded and wall terms shared by the bonded kernel
// and the fused step's pair-kernel epilogue.
#pragma once
#include "mfa_kernels.h"

namespace mfa {

__device__ __forceinline__ double4 bead_at(const Bufs &B, size_t cur_off, int gorig) {
    if (!guard(B, in_range(gorig, B.n_beads), G_ORIG, gorig)) return make_double4(0.0, 0.0, 1.0, 0.0);
    const int k = B.where[gorig];
    if (!guard(B, in_range(k, B.n_slots), G_WHERE, gorig)) return make_double4(0.0, 0.0, 1.0, 0.0);
    return CUR(B.pos)[k];
}

// ---------------------------------------------------------------------------
// Bonded terms
// ---------------------------------------------------------------------------
struct BondOut { double fx, fy, fz, pot, dx, dy, dz; };

// FENE bond (a, b=a+1), intra_molec.f90:21-58: returns force_loc (the force
// the reference adds to b and subtracts from a) and -0.5 ln(1-q).
__device__ __forceinline__ BondOut fene(const SysDev &P, const double4 &ra, const double4 &rb, int *stretched) {
#pragma clang fp contract(off)
    BondOut o;
    double dx = mic(ra.x - rb.x, P.L[0], P.invL[0]);
    double dy = mic(ra.y - rb.y, P.L[1], P.invL[1]);
    double dz = ra.z - rb.z;
    double r_2 = r2_exact(dx, dy, dz);
    double q = r_2 * P.inv_r_chain_2;
    const double clamp = (double)0.98f;
    if (q > clamp) { *stretched = 1; q = clamp; }
    o.pot = -0.5 * log(1.0 - q);
    double rd = P.k_chain / (1.0 - q) * P.inv_r_chain_2;
    o.fx = rd * dx; o.fy = rd * dy; o.fz = rd * dz;
    o.dx = dx; o.dy = dy; o.dz = dz;
    return o;
}

// Virial r_ab (x) F_ab of a FENE bond, added to w (xx yy zz xy xz yz): r_ab
// the minimum-image separation a - b and F_ab = -force_loc the force on a.
// Both beads of a bond add the same value (the bond seen from b negates both
// vectors), so a sum over beads counts every bond twice.
__device__ __forceinline__ void bond_virial(const BondOut &b, double *w) {
#pragma clang fp contract(off)
    w[0] -= b.fx * b.dx; w[1] -= b.fy * b.dy; w[2] -= b.fz * b.dz;
    w[3] -= b.fx * b.dy; w[4] -= b.fx * b.dz; w[5] -= b.fy * b.dz;
}

// fluid_wall(2), fluid_wall.f90:93-121, on bead ri of type ti: the 9-3 force
// of both walls added to F.z and the energy (with the global a_w in the
// attractive term, Q9) to e_wall.
__device__ __forceinline__ void wall_bead(const SysDev &P, const double4 &ri, int ti, double3 &F, double &e_wall) {
#pragma clang fp contract(off)
    const double sw = P.sigma_wall[(ti - 1) & 3], aw = P.a_wall[(ti - 1) & 3];
    double inv_z = 1.0 / ri.z;
    double rd = sw * inv_z;
    e_wall += fabs(aw) * ipow(rd, 9) - P.a_w * ipow(rd, 3);
    F.z += 9.0 * fabs(aw) * ipow(sw, 9) * ipow(inv_z, 10);
    F.z -= 3.0 * aw * ipow(sw, 3) * ipow(inv_z, 4);
    inv_z = 1.0 / (P.L[2] - ri.z);
    rd = sw * inv_z;
    e_wall += fabs(aw) * ipow(rd, 9) - P.a_w * ipow(rd, 3);
    F.z -= 9.0 * fabs(aw) * ipow(sw, 9) * ipow(inv_z, 10);
    F.z += 3.0 * aw * ipow(sw, 3) * ipow(inv_z, 4);
}

// intra_molec, intra_molec.f90:16-125, on the bead at chain position m of a
// chain of nm beads (global index g): the bond to the previous bead then the
// bond to the next one; the bond's energy goes to the bead that starts it.
// VIR: each bond's virial added to wb (twice per bond over all beads).
template <bool VIR = false>
__device__ __forceinline__ void fene_bead(const SysDev &P, const Bufs &B, size_t cur_off, const double4 &ri, int g,
                                          int m, int nm, double3 &F, double &e_fene, double *wb = nullptr) {
#pragma clang fp contract(off)
    int str = 0;
    if (m > 0) {
        BondOut b = fene(P, bead_at(B, cur_off, g - 1), ri, &str);
        F.x += b.fx; F.y += b.fy; F.z += b.fz;
        if (VIR) bond_virial(b, wb);
    }
    if (m < nm - 1) {
        BondOut b = fene(P, ri, bead_at(B, cur_off, g + 1), &str);
        F.x -= b.fx; F.y -= b.fy; F.z -= b.fz;
        e_fene += b.pot;
        if (VIR) bond_virial(b, wb);
    }
    if (str) atomicAdd(&B.flags[2], 1);
}

}  // namespace mfa
