/*
 * mfa_oracle.c -- CPU FP64 restatement of the mfa_brush per-step MD path.
 *
 * TEST INFRASTRUCTURE ONLY (the test reference, never the thing measured or
 * shipped).  Loaded by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg; the HIP product never calls into it.
 *
 * Every routine names the reference file:line it restates.  The arithmetic
 * follows the reference expression by expression and is compiled with
 * -ffp-contract=off, so the interacting-pair decision (r^2 < range_2) is the
 * same IEEE computation as the reference's (no FMA on the decision path).
 *
 * Parity pins: try/conf_old brute-force V_ff golden value, finite
 * differences of every potential, Newton-3 / momentum invariants, the
 * Random123 Philox known-answer vectors, brute force vs cell list
 * (tests/test_oracle.py); the real Fortran reference built into oracle/_ref
 * (thermostat-off and dissipative-only forces, bonds, walls, bending,
 * integrator; tests/test_oracle_vs_reference.py and the golden vectors it
 * wrote); closed forms of the DPD_CUT_OFF branch, of the cosine bending and
 * of the pair and bond virial, the bending_melt range, the virial as the
 * strain derivative of the energy, and the reference's own try/fort.41
 * energy record over 30 steps (tests/test_oracle_pins.py); the reference's
 * shipped input sets, cell list vs brute force (tests/test_ref_inputs_host.py).
 */
#include "mfa_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define T2(ti, tj) (((ti) - 1) + ORC_NT * ((tj) - 1))

/* (double)(float)x: the reference writes single-precision literals that the
 * compiler widens to FP64 (e.g. 0.98 in intra_molec.f90:43). */
#define F2D(x) ((double)(float)(x))

/* ------------------------------------------------------------------------
 * Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11).  Replaces the per-thread ziggurat stream of
 * par_ziggurat.f90:155-199 / ziggurat.f90:110-147 with a counter-based
 * stream keyed on (seed) and counted by (i, j, step).
 * ---------------------------------------------------------------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0;
        uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* Box-Muller normal from two 32-bit words using only IEEE +,-,*,/,sqrt in
 * FP32 (no libm), so the HIP kernel reproduces it bit for bit:
 *   u1 = ((x0>>8)+1) 2^-24 in (0,1],  ln u1 = e ln2 + 2 atanh((m-1)/(m+1)),
 *   angle 2 pi u2 reduced to k pi/2 + a, |a| <= pi/4, Taylor cos/sin. */
static float orc_gauss_from_bits(uint32_t x0, uint32_t x1) {
    float u1 = (float)((x0 >> 8) + 1u) * 5.9604644775390625e-08f;
    uint32_t b = f2u(u1);
    int e = (int)(b >> 23) - 127;
    float m = u2f((b & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float s = (m - 1.0f) / (m + 1.0f);
    float s2 = s * s;
    float p = 0.222222222f;
    p = p * s2 + 0.285714286f;
    p = p * s2 + 0.4f;
    p = p * s2 + 0.666666667f;
    p = p * s2 + 2.0f;
    float lnu = (float)e * 0.693147182f + s * p;
    float rad = sqrtf(-2.0f * lnu);
    uint32_t q = x1 >> 8;                       /* 24-bit u2 = q 2^-24 */
    uint32_t k = (q + (1u << 21)) >> 22;        /* round(4 u2) in 0..4 */
    float f = (float)((int32_t)q - (int32_t)(k << 22)) * 2.38418579e-07f; /* 4u2-k */
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
    float cs;
    switch (k & 3u) {
        case 0: cs = c; break;
        case 1: cs = -sn; break;
        case 2: cs = -c; break;
        default: cs = sn; break;
    }
    return rad * cs;
}

float orc_pair_gauss(uint64_t seed, uint64_t step, int32_t i, int32_t j) {
    uint32_t ctr[4], key[2], out[4];
    ctr[0] = (uint32_t)(i < j ? i : j);
    ctr[1] = (uint32_t)(i < j ? j : i);
    ctr[2] = (uint32_t)step;
    ctr[3] = (uint32_t)(step >> 32);
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    orc_philox4x32_10(ctr, key, out);
    return orc_gauss_from_bits(out[0], out[1]);
}

/* Langevin noise of bead i along axis d at time index step: one normal per
 * (i, d, step).  The counter's second word carries d with the top bit set,
 * so it never equals a pair counter, whose second word is a bead index. */
float orc_lgv_gauss(uint64_t seed, uint64_t step, int32_t i, int32_t d) {
    uint32_t ctr[4], key[2], out[4];
    ctr[0] = (uint32_t)i;
    ctr[1] = 0x80000000u | (uint32_t)d;
    ctr[2] = (uint32_t)step;
    ctr[3] = (uint32_t)(step >> 32);
    key[0] = (uint32_t)seed;
    key[1] = (uint32_t)(seed >> 32);
    orc_philox4x32_10(ctr, key, out);
    return orc_gauss_from_bits(out[0], out[1]);
}

/* lgv_forces.f90:14-25 for bead i: (sig g_d) sqrt(m) - (gamma m) v_d added to
 * its force.  The reference applies it per particle inside the pair loop
 * (fluid_fluid_old.f90:233-236; commented out in fluid_fluid.f90:455-458). */
static void lgv_bead(const orc_params *p, const double *v, const int *a_type, int i, uint64_t step, int noise,
                     double *force) {
    const double m = p->mass_type[(a_type[i] - 1) & 3];
    const double sm = sqrt(m);
    for (int d = 0; d < 3; ++d) {
        const double g = noise ? (double)orc_lgv_gauss(p->seed, step, i, d) : 0.0;
        double vd = p->sig * g * sm;
        vd = vd - p->friction * m * v[3 * i + d];
        force[3 * i + d] = force[3 * i + d] + vd;
    }
}

/* ------------------------------------------------------------------------
 * Cell list (cell_list.f90)
 * ---------------------------------------------------------------------- */

/* make_binning, cell_list.f90:129-158 */
void orc_make_binning(int bin_case, const double boundary[3], double r_cut,
                      double l_cell[3], double inv_l_cell[3], int n_cells[3],
                      int *n_cells_tot, int *n_nei) {
    for (int d = 0; d < 3; ++d) {
        if (bin_case == 2) n_cells[d] = (int)(boundary[d] / (r_cut / 2.0));
        else n_cells[d] = (int)(boundary[d] / r_cut);
    }
    *n_nei = bin_case == 2 ? (125 - 1) / 2 : (27 - 1) / 2;
    int tot = 1;
    for (int d = 0; d < 3; ++d) {
        l_cell[d] = boundary[d] / (double)(float)n_cells[d];   /* float(n_cells) */
        inv_l_cell[d] = 1.0 / l_cell[d];
        tot *= n_cells[d];
    }
    *n_cells_tot = tot;
}

/* get_cell, cell_list.f90:186-213 (1-based, x fastest) */
int orc_get_cell(const double r[3], const int n_cells[3], const double inv_l_cell[3]) {
    int pf = 1, p = 1;
    for (int d = 0; d < 3; ++d) {
        p += pf * (int)(r[d] * inv_l_cell[d]);
        pf *= n_cells[d];
    }
    return p;
}

/* Cell of a bead, clamped into the grid.  The reference's get_cell has no
 * clamp: a bead exactly at r = L (or outside the box) maps to a neighbouring
 * row or out of range (cell_list.f90:200, Q13).  Clamping keeps every bead
 * in the cell that contains its periodic image. */
static int orc_get_cell_clamped(const double r[3], const int n_cells[3], const double inv_l_cell[3]) {
    int pf = 1, p = 1;
    for (int d = 0; d < 3; ++d) {
        int c = (int)(r[d] * inv_l_cell[d]);
        if (c < 0) c = 0;
        if (c > n_cells[d] - 1) c = n_cells[d] - 1;
        p += pf * c;
        pf *= n_cells[d];
    }
    return p;
}

/* neigh_list + check_bound, cell_list.f90:56-127: head insertion */
void orc_neigh_list(const double *r0, int n_part, const int n_cells[3], int n_cells_tot,
                    const double boundary[3], const double inv_l_cell[3],
                    int *part_in_cell, int *r_nei, int *l_nei, int *n_out_of_box) {
    int nob = 0;
    for (int c = 0; c <= n_cells_tot; ++c) part_in_cell[c] = 0;
    for (int i = 0; i < n_part; ++i) { r_nei[i] = 0; l_nei[i] = 0; }
    for (int ip = 1; ip <= n_part; ++ip) {
        const double *r = r0 + 3 * (ip - 1);
        for (int d = 0; d < 3; ++d)
            if (r[d] < 0.0 || r[d] > boundary[d]) { ++nob; break; }
        int p = orc_get_cell_clamped(r, n_cells, inv_l_cell);
        if (part_in_cell[p] == 0) {
            part_in_cell[p] = ip;
        } else {
            int jp = part_in_cell[p];
            part_in_cell[p] = ip;
            r_nei[ip - 1] = jp;
            l_nei[ip - 1] = 0;
            l_nei[jp - 1] = ip;
        }
    }
    part_in_cell[0] = 0;
    if (n_out_of_box) *n_out_of_box = nob;
}

/* update_part_cell, cell_list.f90:215-268 (doubly linked relink) */
void orc_update_part_cell(const double r[3], int ip, const int n_cells[3],
                          const double inv_l_cell[3], int old_cell,
                          int *part_in_cell, int *r_nei, int *l_nei) {
    int new_cell = orc_get_cell_clamped(r, n_cells, inv_l_cell);
    if (new_cell == old_cell) return;
    if (part_in_cell[old_cell] == ip) {
        int jp = r_nei[ip - 1];
        part_in_cell[old_cell] = jp;
        if (jp != 0) l_nei[jp - 1] = 0;
    } else {
        int jp = r_nei[ip - 1];
        r_nei[l_nei[ip - 1] - 1] = jp;
        if (jp != 0) l_nei[jp - 1] = l_nei[ip - 1];
    }
    if (part_in_cell[new_cell] == 0) {
        part_in_cell[new_cell] = ip;
        l_nei[ip - 1] = 0;
        r_nei[ip - 1] = 0;
    } else {
        int jp = part_in_cell[new_cell];
        part_in_cell[new_cell] = ip;
        r_nei[ip - 1] = jp;
        l_nei[ip - 1] = 0;
        l_nei[jp - 1] = ip;
    }
}

/* get_cell_neigh_list / make_cell_list2, cell_list.f90:270-433, n_dim = 3,
 * SYMMETRY = 0 (periodic x,y; out-of-range z -> phantom cell 0). */
void orc_get_cell_neigh_list(int n_cells_tot, const int n_cells[3], int n_nei,
                             int *cell_neigh_ls) {
    const int nx = n_cells[0], ny = n_cells[1], nz = n_cells[2];
    for (int in_cell = 1; in_cell <= n_cells_tot; ++in_cell) {
        int z_cell = (in_cell - 1) / (nx * ny) + 1;
        int y_cell = (in_cell - nx * ny * (z_cell - 1) - 1) / nx + 1;
        int x_cell = in_cell - nx * ny * (z_cell - 1) - nx * (y_cell - 1);
        int il = 0;
        if (n_nei == 13) {
            for (int i = 2; i <= 3; ++i)
                for (int j = 1; j <= 3; ++j) {
                    if (i == 2 && j <= 1) continue;
                    for (int k = 1; k <= 3; ++k) {
                        if (i == 2 && j == 2 && k >= 3) continue;
                        int jc = in_cell + i - 2 + (j - 2) * nx + (k - 2) * ny * nx;
                        if (jc == in_cell) continue;
                        if (x_cell == 1 && i == 1) jc += nx;
                        if (x_cell == nx && i == 3) jc -= nx;
                        if (y_cell == 1 && j == 1) jc += nx * ny;
                        if (y_cell == ny && j == 3) jc -= nx * ny;
                        if (z_cell + k - 2 < 1) jc = 0;
                        if (z_cell + k - 2 > nz) jc = 0;
                        cell_neigh_ls[(in_cell - 1) + (size_t)n_cells_tot * il] = jc;
                        ++il;
                    }
                }
        } else { /* 62 */
            for (int i = 2; i <= 4; ++i)
                for (int j = 0; j <= 4; ++j) {
                    if (i == 2 && j <= 1) continue;
                    for (int k = 0; k <= 4; ++k) {
                        if (i == 2 && j == 2 && k >= 3) continue;
                        int jc = in_cell + i - 2 + (j - 2) * nx + (k - 2) * ny * nx;
                        if (jc == in_cell) continue;
                        if (x_cell + i - 2 > nx) jc -= nx;
                        if (y_cell + j - 2 < 1) jc += nx * ny;
                        if (y_cell + j - 2 > ny) jc -= nx * ny;
                        if (z_cell + k - 2 < 1) jc = 0;
                        if (z_cell + k - 2 > nz) jc = 0;
                        cell_neigh_ls[(in_cell - 1) + (size_t)n_cells_tot * il] = jc;
                        ++il;
                    }
                }
        }
    }
}

/* ------------------------------------------------------------------------
 * Pair forces: fluid_fluid.f90:65-475 with dpd_forces.f90:1-129
 * ---------------------------------------------------------------------- */

/* Force switch-on ramp, fluid_fluid.f90:35-42 */
double orc_r2_min(const orc_params *p, double r_time) {
    if (!p->force_switch_on) return 0.0;
    if (r_time < p->r_2_min_time) return (1.0 - r_time / p->r_2_min_time) * p->r_2_min_init;
    return 0.0;
}

/* minimum image in x,y: delta - L*int(2*delta*inv_L), fluid_fluid.f90:111-112 */
static inline double mic(double d, double L, double invL) {
    return d - L * (double)(int)(2.0 * d * invL);
}

/* DPD pair term, dpd_forces.f90:45-115.  fi/fj receive +vec / -vec. */
static inline void dpd_pair(const orc_params *p, const double dr[3], double r_2, double inv_sqrt_r_2,
                            int ti, int tj, const double *vi, const double *vj, double g,
                            double *fi, double *fj) {
    double w;
    if (p->dpd_cut_off) w = 1.0 - sqrt(r_2 / p->r_cut_dpd_2);
    else w = 1.0 - sqrt(r_2 * p->inv_range_2[T2(ti, tj)]);
    double w_r = w, w_d = w * w;
    double rv[3], vec[3];
    for (int d = 0; d < 3; ++d) rv[d] = dr[d] * inv_sqrt_r_2;
    for (int d = 0; d < 3; ++d) vec[d] = p->sig * w_r * g * rv[d];
    for (int d = 0; d < 3; ++d) { fi[d] += vec[d]; fj[d] -= vec[d]; }
    double dv[3];
    for (int d = 0; d < 3; ++d) dv[d] = vi[d] - vj[d];
    double rd = dv[0] * rv[0] + dv[1] * rv[1] + dv[2] * rv[2];
    for (int d = 0; d < 3; ++d) vec[d] = -1.0 * p->friction * w_d * rd * rv[d];
    for (int d = 0; d < 3; ++d) { fi[d] += vec[d]; fj[d] -= vec[d]; }
}

/* One candidate pair (i,j) (0-based), fluid_fluid.f90:97-261.  LJ force on
 * i goes to f_ipart, on j to force[j]; DPD goes to force[i], force[j]. */
static inline void pair_body(const orc_params *p, const double *r0, const double *v, const int *a_type,
                             int i, int j, double r_2_min, uint64_t step, int noise,
                             double *f_ipart, double *force, double *v_ff) {
    int ti = a_type[i], tj = a_type[j];
    double r_cut2 = p->range_2[T2(ti, tj)];
    double dr[3];
    dr[0] = r0[3 * i + 0] - r0[3 * j + 0];
    dr[1] = r0[3 * i + 1] - r0[3 * j + 1];
    dr[2] = r0[3 * i + 2] - r0[3 * j + 2];
    dr[0] = mic(dr[0], p->boundary[0], p->inv_boundary[0]);
    dr[1] = mic(dr[1], p->boundary[1], p->inv_boundary[1]);
    double r_2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    double inv_sqrt_r_2 = 0.0;
    int lj = r_2 < r_cut2;
    if (lj) {
        if (p->force_switch_on && r_2 < r_2_min) r_2 = r_2_min;
        double inv_r_2 = 1.0 / r_2;
        inv_sqrt_r_2 = sqrt(1.0 / r_2);
        double l_eps = p->epsil[T2(ti, tj)];
        double r_61 = p->sigma_2[T2(ti, tj)] * inv_r_2;
        double r_6 = r_61 * r_61 * r_61;
        double r_12 = r_6 * r_6;
        double pot = (r_12 - r_6) - p->e_shift[T2(ti, tj)];
        *v_ff += l_eps * pot;
        double rdm = l_eps * (-12.0 * r_12 + 6.0 * r_6) * inv_r_2;
        double fl[3] = {rdm * dr[0], rdm * dr[1], rdm * dr[2]};
        f_ipart[0] -= fl[0]; f_ipart[1] -= fl[1]; f_ipart[2] -= fl[2];
        force[3 * j + 0] += fl[0]; force[3 * j + 1] += fl[1]; force[3 * j + 2] += fl[2];
        if (p->thermostat == 1 && !p->dpd_cut_off) {
            double g = noise ? (double)orc_pair_gauss(p->seed, step, i, j) : 0.0;
            dpd_pair(p, dr, r_2, inv_sqrt_r_2, ti, tj, v + 3 * i, v + 3 * j, g,
                     force + 3 * i, force + 3 * j);
        }
    }
    if (p->thermostat == 1 && p->dpd_cut_off) {
        if (ti == 1 || tj == 1) return;                 /* heads excluded, :223-236 */
        if (r_2 < p->r_cut_dpd_2) {
            /* the reference reuses the last 1/r when the LJ branch did not run
             * (stale inv_sqrt_r_2); here it is this pair's own 1/r */
            if (!lj) inv_sqrt_r_2 = sqrt(1.0 / r_2);
            double g = noise ? (double)orc_pair_gauss(p->seed, step, i, j) : 0.0;
            dpd_pair(p, dr, r_2, inv_sqrt_r_2, ti, tj, v + 3 * i, v + 3 * j, g,
                     force + 3 * i, force + 3 * j);
        }
    }
}

static void ff_cell(const orc_params *p, const double *r0, const double *v, const int *a_type,
                    const int *part_in_cell, const int *r_nei, const int *cell_neigh_ls,
                    int n_cells_tot, int n_nei, double r_2_min, uint64_t step, int noise,
                    int i_cell, double *force, double *v_ff) {
    int ip = part_in_cell[i_cell];
    while (ip != 0) {
        int i = ip - 1;
        double f_ipart[3] = {0.0, 0.0, 0.0};
        for (int jd = 0; jd < n_nei; ++jd) {
            int j_cell = cell_neigh_ls[(i_cell - 1) + (size_t)n_cells_tot * jd];
            int jp = part_in_cell[j_cell];
            while (jp != 0) {
                pair_body(p, r0, v, a_type, i, jp - 1, r_2_min, step, noise, f_ipart, force, v_ff);
                jp = r_nei[jp - 1];
            }
        }
        int jp = part_in_cell[i_cell];
        while (jp != 0) {
            if (ip < jp)
                pair_body(p, r0, v, a_type, i, jp - 1, r_2_min, step, noise, f_ipart, force, v_ff);
            jp = r_nei[jp - 1];
        }
        force[3 * i + 0] += f_ipart[0];
        force[3 * i + 1] += f_ipart[1];
        force[3 * i + 2] += f_ipart[2];
        ip = r_nei[ip - 1];
    }
}

/* fluid_fluid (BIN_TYPE 2) -- adds into force; OpenMP over cells with a
 * full-array reduction, as fluid_fluid.f90:65-70. */
void orc_fluid_fluid(const orc_params *p, const double *r0, const double *v, const int *a_type,
                     const int *part_in_cell, const int *r_nei, const int *cell_neigh_ls,
                     int n_cells_tot, int n_nei, double r_2_min, uint64_t step,
                     int noise, double *force, double *v_fluid_fluid) {
    double v_ff = 0.0;
#ifdef _OPENMP
    const size_t n3 = 3 * (size_t)p->n_part;
    if (omp_get_max_threads() > 1) {
#pragma omp parallel
        {
            double *fl = (double *)calloc(n3, sizeof(double));
            double vloc = 0.0;
#pragma omp for schedule(static, 10)
            for (int c = 1; c <= n_cells_tot; ++c)
                ff_cell(p, r0, v, a_type, part_in_cell, r_nei, cell_neigh_ls, n_cells_tot, n_nei,
                        r_2_min, step, noise, c, fl, &vloc);
#pragma omp critical
            {
                for (size_t k = 0; k < n3; ++k) force[k] += fl[k];
                v_ff += vloc;
            }
            free(fl);
        }
        *v_fluid_fluid = v_ff;
        if (p->thermostat == 2)
            for (int i = 0; i < p->n_part; ++i) lgv_bead(p, v, a_type, i, step, noise, force);
        return;
    }
#endif
    for (int c = 1; c <= n_cells_tot; ++c)
        ff_cell(p, r0, v, a_type, part_in_cell, r_nei, cell_neigh_ls, n_cells_tot, n_nei,
                r_2_min, step, noise, c, force, &v_ff);
    *v_fluid_fluid = v_ff;
    if (p->thermostat == 2)
        for (int i = 0; i < p->n_part; ++i) lgv_bead(p, v, a_type, i, step, noise, force);
}

/* check_fluid_fluid, util.f90:284-330: O(N^2) brute-force LJ energy with the
 * same MIC, cutoff test and switch-on clamp; sigma_2/r_2 by division. */
double orc_check_fluid_fluid(const orc_params *p, const double *r0, const int *a_type, double r_2_min) {
    double vff = 0.0;
    const int n = p->n_part;
    for (int i = 0; i < n - 1; ++i) {
        int ti = a_type[i];
        for (int j = i + 1; j < n; ++j) {
            int tj = a_type[j];
            double dr[3];
            for (int d = 0; d < 3; ++d) dr[d] = r0[3 * i + d] - r0[3 * j + d];
            for (int d = 0; d < 2; ++d)
                dr[d] = dr[d] - p->boundary[d] * (double)(int)(2.0 * dr[d] * p->inv_boundary[d]);
            double r_2 = 0.0;
            for (int d = 0; d < 3; ++d) r_2 = r_2 + dr[d] * dr[d];
            if (r_2 < p->range_2[T2(ti, tj)]) {
                if (r_2 < r_2_min) r_2 = r_2_min;
                double q = p->sigma_2[T2(ti, tj)] / r_2;
                double r_6 = q * q * q;
                double r_12 = r_6 * r_6;
                double pot = (r_12 - r_6) - p->e_shift[T2(ti, tj)];
                vff = vff + p->epsil[T2(ti, tj)] * pot;
            }
        }
    }
    return vff;
}

/* 64-bit mix of a partner index (splitmix64 finaliser) for pair-set hashes. */
static inline uint64_t mix64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

/* Brute-force interacting-pair set: per bead the number of partners with
 * r^2 < range_2 and the wrapping sum of mix64(partner).  Bit-exact pair-set
 * fingerprint (same decision arithmetic as fluid_fluid.f90:104-121). */
void orc_pair_stats(const orc_params *p, const double *r0, const int *a_type,
                    int32_t *count, uint64_t *hash) {
    const int n = p->n_part;
    for (int i = 0; i < n; ++i) { count[i] = 0; hash[i] = 0; }
    for (int i = 0; i < n - 1; ++i) {
        for (int j = i + 1; j < n; ++j) {
            double dr[3];
            for (int d = 0; d < 3; ++d) dr[d] = r0[3 * i + d] - r0[3 * j + d];
            dr[0] = mic(dr[0], p->boundary[0], p->inv_boundary[0]);
            dr[1] = mic(dr[1], p->boundary[1], p->inv_boundary[1]);
            double r_2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            if (r_2 < p->range_2[T2(a_type[i], a_type[j])]) {
                count[i]++; count[j]++;
                hash[i] += mix64((uint64_t)j);
                hash[j] += mix64((uint64_t)i);
            }
        }
    }
}

/* Same fingerprint through the linked-cell half-shell traversal of
 * fluid_fluid.f90:71-442 (O(N), for full-size configurations). */
int orc_pair_stats_cells(const orc_params *p, const double *r0, const int *a_type,
                         int32_t *count, uint64_t *hash) {
    orc_cells c;
    if (orc_cells_init(&c, p, r0, p->r_cut_max) != 0) { orc_cells_free(&c); return -1; }
    const int n = p->n_part;
    for (int i = 0; i < n; ++i) { count[i] = 0; hash[i] = 0; }
    for (int ic = 1; ic <= c.n_cells_tot; ++ic) {
        for (int ip = c.part_in_cell[ic]; ip != 0; ip = c.r_nei[ip - 1]) {
            for (int jd = 0; jd <= c.n_nei; ++jd) {
                int jc = jd < c.n_nei ? c.cell_neigh_ls[(ic - 1) + (size_t)c.n_cells_tot * jd] : ic;
                for (int jp = c.part_in_cell[jc]; jp != 0; jp = c.r_nei[jp - 1]) {
                    if (jd == c.n_nei && !(ip < jp)) continue;
                    int i = ip - 1, j = jp - 1;
                    double dr[3];
                    for (int d = 0; d < 3; ++d) dr[d] = r0[3 * i + d] - r0[3 * j + d];
                    dr[0] = mic(dr[0], p->boundary[0], p->inv_boundary[0]);
                    dr[1] = mic(dr[1], p->boundary[1], p->inv_boundary[1]);
                    double r_2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                    if (r_2 < p->range_2[T2(a_type[i], a_type[j])]) {
                        count[i]++; count[j]++;
                        hash[i] += mix64((uint64_t)j);
                        hash[j] += mix64((uint64_t)i);
                    }
                }
            }
        }
    }
    orc_cells_free(&c);
    return 0;
}

/* All-pairs fluid_fluid (same pair body, i<j order): the brute-force force
 * oracle used to check the cell-list traversal. */
void orc_fluid_fluid_brute(const orc_params *p, const double *r0, const double *v, const int *a_type,
                           double r_2_min, uint64_t step, int noise, double *force, double *v_ff) {
    double e = 0.0;
    for (int i = 0; i < p->n_part; ++i) {
        double fi[3] = {0.0, 0.0, 0.0};
        for (int j = i + 1; j < p->n_part; ++j)
            pair_body(p, r0, v, a_type, i, j, r_2_min, step, noise, fi, force, &e);
        force[3 * i + 0] += fi[0];
        force[3 * i + 1] += fi[1];
        force[3 * i + 2] += fi[2];
    }
    *v_ff = e;
    if (p->thermostat == 2)
        for (int i = 0; i < p->n_part; ++i) lgv_bead(p, v, a_type, i, step, noise, force);
}

/* Pair virial W = sum over interacting pairs i<j of r_ij (x) F_ij, with r_ij
 * the minimum-image separation r_i - r_j and F_ij the total pair force on i
 * (LJ, dissipative and random parts), components xx yy zz xy xz yz.  The
 * reference accumulates a pressure tensor only in its bulk build
 * (fluid_fluid.f90:156-166, dpd_forces.f90:118-127) and none in channel mode;
 * this is the channel-mode virial of the device path.  Each pair goes through
 * pair_body, so the pair force is exactly the one orc_fluid_fluid_brute adds:
 * the LJ part lands in f_i, the DPD part in the scratch rows i and j, which
 * are read back and cleared after every pair. */
void orc_pair_virial(const orc_params *p, const double *r0, const double *v, const int *a_type,
                     double r_2_min, uint64_t step, int noise, double W[6]) {
    double *tmp = (double *)calloc(3 * (size_t)p->n_part, sizeof(double));
    double e = 0.0;
    for (int q = 0; q < 6; ++q) W[q] = 0.0;
    for (int i = 0; i < p->n_part; ++i) {
        for (int j = i + 1; j < p->n_part; ++j) {
            double fi[3] = {0.0, 0.0, 0.0};
            pair_body(p, r0, v, a_type, i, j, r_2_min, step, noise, fi, tmp, &e);
            double f[3], dr[3];
            for (int d = 0; d < 3; ++d) {
                f[d] = fi[d] + tmp[3 * i + d];
                tmp[3 * i + d] = 0.0;
                tmp[3 * j + d] = 0.0;
                dr[d] = r0[3 * i + d] - r0[3 * j + d];
            }
            if (f[0] == 0.0 && f[1] == 0.0 && f[2] == 0.0) continue;
            dr[0] = mic(dr[0], p->boundary[0], p->inv_boundary[0]);
            dr[1] = mic(dr[1], p->boundary[1], p->inv_boundary[1]);
            W[0] += dr[0] * f[0]; W[1] += dr[1] * f[1]; W[2] += dr[2] * f[2];
            W[3] += dr[0] * f[1]; W[4] += dr[0] * f[2]; W[5] += dr[1] * f[2];
        }
    }
    free(tmp);
}

/* ------------------------------------------------------------------------
 * Walls, bonds, bending, body force
 * ---------------------------------------------------------------------- */

static inline double ipow(double x, int n) {
    double r = 1.0;
    while (n) { if (n & 1) r *= x; x *= x; n >>= 1; }
    return r;
}

/* fluid_wall(2), fluid_wall.f90:77-123 (NO_WARNS: no teleport).  The energy
 * uses the global a_w in the attractive term, the force a_wall(type) (Q9). */
double orc_fluid_wall(const orc_params *p, const double *r0, const int *a_type, double *force) {
    double vfw = 0.0;
    for (int i = 0; i < p->n_part; ++i) {
        int t = a_type[i] - 1;
        double sw = p->sigma_wall[t], aw = p->a_wall[t];
        double inv_z = 1.0 / r0[3 * i + 2];
        double rd = sw * inv_z;
        vfw = vfw + fabs(aw) * ipow(rd, 9) - p->a_w * ipow(rd, 3);
        force[3 * i + 2] += 9.0 * fabs(aw) * ipow(sw, 9) * ipow(inv_z, 10);
        force[3 * i + 2] -= 3.0 * aw * ipow(sw, 3) * ipow(inv_z, 4);
        inv_z = 1.0 / (p->z_space_wall - r0[3 * i + 2]);
        rd = sw * inv_z;
        vfw = vfw + fabs(aw) * ipow(rd, 9) - p->a_w * ipow(rd, 3);
        force[3 * i + 2] -= 9.0 * fabs(aw) * ipow(sw, 9) * ipow(inv_z, 10);
        force[3 * i + 2] += 3.0 * aw * ipow(sw, 3) * ipow(inv_z, 4);
    }
    return vfw;
}

/* wall_wall(2), wall_wall.f90:45-47: a constant */
double orc_wall_wall(const orc_params *p, double sigma_w) {
    double q = sigma_w / p->z_space_wall;
    return 2.0 * (fabs(p->a_w) * ipow(q, 9) - p->a_w * ipow(q, 3));
}

static double fene_chain_set(const orc_params *p, const double *r0, double *force, int first,
                             int n_mon, int n_chain, int *n_str, double *W) {
    double v = 0.0;
    const double clamp = F2D(0.98);
    for (int c = 0; c < n_chain; ++c)
        for (int m = 0; m < n_mon - 1; ++m) {
            int i = first + c * n_mon + m, j = i + 1;
            double dr[3];
            for (int d = 0; d < 3; ++d) dr[d] = r0[3 * i + d] - r0[3 * j + d];
            dr[0] = mic(dr[0], p->boundary[0], p->inv_boundary[0]);
            dr[1] = mic(dr[1], p->boundary[1], p->inv_boundary[1]);
            double r_2 = 0.0;
            for (int d = 0; d < 3; ++d) r_2 = r_2 + dr[d] * dr[d];
            double q = r_2 * p->inv_r_chain_2;
            if (q > clamp) { if (n_str) ++*n_str; q = clamp; }
            v = v + (-0.5 * log(1.0 - q));
            double rd = p->k_chain / (1.0 - q) * p->inv_r_chain_2;
            double fl[3];
            for (int d = 0; d < 3; ++d) {
                fl[d] = rd * dr[d];
                force[3 * i + d] -= fl[d];
                force[3 * j + d] += fl[d];
            }
            if (W) {    /* r_ij (x) F_ij, F_ij = -fl the force on i (intra_molec.f90:115-120) */
                W[0] -= fl[0] * dr[0]; W[1] -= fl[1] * dr[1]; W[2] -= fl[2] * dr[2];
                W[3] -= fl[0] * dr[1]; W[4] -= fl[0] * dr[2]; W[5] -= fl[1] * dr[2];
            }
        }
    return p->k_chain * v;
}

/* intra_molec, intra_molec.f90:12-129,266: FENE brush + melt bonds */
double orc_intra_molec(const orc_params *p, const double *r0, double *force, int *n_stretched) {
    if (n_stretched) *n_stretched = 0;
    double vb = fene_chain_set(p, r0, force, 0, p->n_mon, p->n_chain, n_stretched, NULL);
    double vm = fene_chain_set(p, r0, force, p->n_mon * p->n_chain, p->n_mon_d, p->n_chain_d, n_stretched, NULL);
    return vb + vm;
}

/* FENE bond virial W = sum over the bonds of brush and melt chains of
 * r_ij (x) F_ij (xx yy zz xy xz yz), r_ij = r_i - r_j the minimum image in x
 * and y and F_ij the FENE force on i, as intra_molec.f90:115-120 accumulates
 * it for the melt (its brush loop, :61-66, adds the opposite sign). */
void orc_bond_virial(const orc_params *p, const double *r0, double W[6]) {
    double *tmp = (double *)calloc(3 * (size_t)(p->n_part > 0 ? p->n_part : 1), sizeof(double));
    for (int q = 0; q < 6; ++q) W[q] = 0.0;
    fene_chain_set(p, r0, tmp, 0, p->n_mon, p->n_chain, NULL, W);
    fene_chain_set(p, r0, tmp, p->n_mon * p->n_chain, p->n_mon_d, p->n_chain_d, NULL, W);
    free(tmp);
}

/* bending(1), bending.f90:19-121 (melt=0) / bending_melt(1),
 * bending_melt.f90:22-120 (melt=1; the missing `beg+` of :43 is restored).
 * bend_form 1 replaces the harmonic angle term by k_bend (1 - cos theta) over
 * the same angles (no reference counterpart; the melt workload's form). */
double orc_bending(const orc_params *p, const double *r0, double *force, int melt) {
    int first = melt ? p->n_mon * p->n_chain : 0;
    int n_mon = melt ? p->n_mon_d : p->n_mon;
    int n_chain = melt ? p->n_chain_d : p->n_chain;
    const double small = F2D(0.000001);
    double vb = 0.0;
    for (int l = 0; l < n_chain; ++l) {
        int b = first + l * n_mon;     /* 0-based index of bead 1 of the chain */
        double rn[3], rp[3];
        for (int d = 0; d < 3; ++d) rn[d] = r0[3 * (b + 1) + d] - r0[3 * b + d];
        rn[0] = mic(rn[0], p->boundary[0], p->inv_boundary[0]);
        rn[1] = mic(rn[1], p->boundary[1], p->inv_boundary[1]);
        for (int i = 2; i <= n_mon - 1; ++i) {
            for (int d = 0; d < 3; ++d) rp[d] = rn[d];
            int ic = b + i - 1;          /* centre bead (reference l*n_mon+i) */
            for (int d = 0; d < 3; ++d) rn[d] = r0[3 * (ic + 1) + d] - r0[3 * ic + d];
            rn[0] = mic(rn[0], p->boundary[0], p->inv_boundary[0]);
            rn[1] = mic(rn[1], p->boundary[1], p->inv_boundary[1]);
            double dot = 0.0, rp2 = 0.0, rn2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                dot = dot + rn[d] * rp[d];
                rp2 = rp2 + rp[d] * rp[d];
                rn2 = rn2 + rn[d] * rn[d];
            }
            double cosa = dot / sqrt(rp2 * rn2);
            double dn[3], dp[3], dn2 = 0.0, dp2 = 0.0;
            for (int d = 0; d < 3; ++d) {
                dn[d] = rp[d] * rn2 - rn[d] * dot;
                dp[d] = -rn[d] * rp2 + rp[d] * dot;
            }
            for (int d = 0; d < 3; ++d) { dp2 = dp2 + dp[d] * dp[d]; dn2 = dn2 + dn[d] * dn[d]; }
            if (dn2 < small) {
                double s1 = sqrt(rp2), s2 = sqrt(rn2);
                for (int d = 0; d < 3; ++d) dn[d] = rp[d] / s1 - rn[d] / s2;
                double nrm = sqrt(dn[0] * dn[0] + dn[1] * dn[1] + dn[2] * dn[2]);
                if (nrm != 0.0) { for (int d = 0; d < 3; ++d) dn[d] = dn[d] / nrm / s2; }
                else { for (int d = 0; d < 3; ++d) dn[d] = dn[d] * 0.0; }
            } else {
                double s = sqrt(dn2 * rn2);
                for (int d = 0; d < 3; ++d) dn[d] = dn[d] / s;
            }
            if (dp2 < small) {
                double s1 = sqrt(rp2), s2 = sqrt(rn2);
                for (int d = 0; d < 3; ++d) dp[d] = rp[d] / s1 - rn[d] / s2;
                double nrm = sqrt(dp[0] * dp[0] + dp[1] * dp[1] + dp[2] * dp[2]);
                if (nrm != 0.0) { for (int d = 0; d < 3; ++d) dp[d] = dp[d] / nrm / s1; }
                else { for (int d = 0; d < 3; ++d) dp[d] = dp[d] * 0.0; }
            } else {
                double s = sqrt(dp2 * rp2);
                for (int d = 0; d < 3; ++d) dp[d] = dp[d] / s;
            }
            if (p->bend_form == 1) {
                /* cosine form V = k (1 - cos theta), theta the angle between
                 * consecutive bonds (0 straight): F_prev = k dcos/dr_prev =
                 * k (cos rp/|rp|^2 - rn/(|rp||rn|)), F_next = k (rp/(|rp||rn|)
                 * - cos rn/|rn|^2), the centre takes minus their sum */
                double inv_pn = 1.0 / sqrt(rp2 * rn2);
                vb = vb + p->k_bend * (1.0 - cosa);
                for (int d = 0; d < 3; ++d) {
                    double fp = p->k_bend * (cosa * rp[d] / rp2 - rn[d] * inv_pn);
                    double fn = p->k_bend * (rp[d] * inv_pn - cosa * rn[d] / rn2);
                    force[3 * (ic - 1) + d] += fp;
                    force[3 * ic + d] = force[3 * ic + d] - fp - fn;
                    force[3 * (ic + 1) + d] += fn;
                }
                continue;
            }
            double dalpha = cosa >= 1.0 ? 0.0 : acos(cosa) - p->alpha_eq;
            double fmod = p->k_bend * dalpha;
            vb = vb + 0.5 * fmod * dalpha;
            for (int d = 0; d < 3; ++d) {
                double fp = fmod * dp[d], fn = fmod * dn[d];
                force[3 * (ic - 1) + d] += fp;
                force[3 * ic + d] = force[3 * ic + d] - fp - fn;
                force[3 * (ic + 1) + d] += fn;
            }
        }
    }
    return vb;
}

/* constant_force(1), constant_force.f90:17-21 with ext_force set up as
 * init_config.f90:370-378: m*g on every bead except heads (type 1). */
void orc_constant_force(const orc_params *p, const double *mass, const int *a_type, double *force) {
    for (int i = 0; i < p->n_part; ++i) {
        if (a_type[i] == 1) continue;
        for (int d = 0; d < 3; ++d) force[3 * i + d] += mass[i] * p->const_force[d];
    }
}

/* ------------------------------------------------------------------------
 * Integrator
 * ---------------------------------------------------------------------- */

/* verlet_positions, verlet_positions.f90:150-209 (BIN_TYPE 2 relink) */
void orc_verlet_positions(const orc_params *p, double *r0, double *v, const double *a,
                          int relink, const int n_cells[3], const double inv_l_cell[3],
                          int *part_in_cell, int *r_nei, int *l_nei) {
    for (int i = 0; i < p->n_part; ++i) {
        double *r = r0 + 3 * i;
        int old_cell = relink ? orc_get_cell_clamped(r, n_cells, inv_l_cell) : 0;
        for (int d = 0; d < 3; ++d) r[d] = r[d] + p->dt * v[3 * i + d] + 0.5 * p->dt_2 * a[3 * i + d];
        for (int d = 0; d < 2; ++d) {
            if (r[d] > p->boundary[d]) r[d] = r[d] - p->boundary[d];
            else if (r[d] < 0.0) r[d] = r[d] + p->boundary[d];
        }
        if (relink)
            orc_update_part_cell(r, i + 1, n_cells, inv_l_cell, old_cell, part_in_cell, r_nei, l_nei);
        for (int d = 0; d < 3; ++d) v[3 * i + d] = v[3 * i + d] + 0.5 * p->dt * a[3 * i + d];
    }
}

/* verlet_velocities, verlet_velocities.f90:14-153: f_on_heads snapshot,
 * a = F/m, second half-kick, heads frozen (velocity, acceleration and force
 * of every head zeroed, :95-119), kinetic energy t_fluid. */
double orc_verlet_velocities(const orc_params *p, double *force_in, const double *inv_mass,
                             const double *mass, double *v, double *a, double *f_on_heads) {
    const int n_head_span = p->n_mon * p->n_chain;
    if (f_on_heads)
        for (int c = 0; c < p->n_chain; ++c)
            for (int d = 0; d < 3; ++d) f_on_heads[3 * c + d] = force_in[3 * (c * p->n_mon) + d];
    (void)n_head_span;
    for (int i = 0; i < p->n_part; ++i)
        for (int d = 0; d < 3; ++d) {
            a[3 * i + d] = force_in[3 * i + d] * inv_mass[i];
            v[3 * i + d] = v[3 * i + d] + 0.5 * p->dt * a[3 * i + d];
        }
    for (int c = 0; c < p->n_chain; ++c) {
        int h = c * p->n_mon;
        if (p->poiseuille) {
            v[3 * h + 0] = 0.0; v[3 * h + 1] = 0.0;
        } else {
            double s = c < p->n_chain / 2 ? 1.0 : -1.0;
            v[3 * h + 0] = s * p->head_vel[0]; v[3 * h + 1] = s * p->head_vel[1];
        }
        v[3 * h + 2] = 0.0;
        for (int d = 0; d < 3; ++d) {
            force_in[3 * h + d] = 0.0;
            a[3 * h + d] = 0.0;
        }
    }
    double t = 0.0;
    for (int i = 0; i < p->n_part; ++i) {
        double td = mass[i] * (v[3 * i] * v[3 * i] + v[3 * i + 1] * v[3 * i + 1] + v[3 * i + 2] * v[3 * i + 2]);
        t = t + td;
    }
    return 0.5 * t;
}

/* ------------------------------------------------------------------------
 * Observables: histo(2) functions.f90:157-213, velocity_prof(2) :319-365
 * dens_z(hist_dim,3), v_prof(n_layers,3,3) column-major like the reference.
 * ---------------------------------------------------------------------- */
static inline int zbin(double z, double inv_step, int n) {
    int k = (int)(z * inv_step) + 1;
    if (k < 1) k = 1;
    if (k > n) k = n;
    return k - 1;
}

void orc_histo(const orc_params *p, const double *r0, int hist_dim, double *dens_z) {
    double inv_step = 1.0 / (p->boundary[2] / (double)hist_dim);
    int pd = p->n_mon * p->n_chain, pe = pd + p->n_mon_d * p->n_chain_d;
    for (int i = 0; i < pd; ++i) dens_z[zbin(r0[3 * i + 2], inv_step, hist_dim)] += 1.0;
    for (int i = pd; i < pe; ++i) dens_z[hist_dim + zbin(r0[3 * i + 2], inv_step, hist_dim)] += 1.0;
}

void orc_velocity_prof(const orc_params *p, const double *r0, const double *v, int n_layers,
                       double *v_prof, double *v_prof_2) {
    double dz = p->boundary[2] / (double)n_layers;
    double inv_step = 1.0 / dz;
    int pd = p->n_mon * p->n_chain, pe = pd + p->n_mon_d * p->n_chain_d;
    for (int i = 0; i < pe; ++i) {
        int s = i < pd ? 0 : 1;
        int k = zbin(r0[3 * i + 2], inv_step, n_layers);
        for (int d = 0; d < 3; ++d) {
            size_t o = k + (size_t)n_layers * (d + 3 * s);
            v_prof[o] += v[3 * i + d];
            v_prof_2[o] += v[3 * i + d] * v[3 * i + d];
        }
    }
}

/* ------------------------------------------------------------------------
 * One md_main step, md_main.f90:77-185 (BIN_TYPE 2, WALL 2, POISEUILLE)
 * ---------------------------------------------------------------------- */
int orc_cells_init(orc_cells *c, const orc_params *p, const double *r0, double r_cut) {
    memset(c, 0, sizeof(*c));
    orc_make_binning(p->l_box, p->boundary, r_cut, c->l_cell, c->inv_l_cell, c->n_cells,
                     &c->n_cells_tot, &c->n_nei);
    c->part_in_cell = (int *)calloc((size_t)c->n_cells_tot + 1, sizeof(int));
    c->r_nei = (int *)calloc((size_t)p->n_part, sizeof(int));
    c->l_nei = (int *)calloc((size_t)p->n_part, sizeof(int));
    c->cell_neigh_ls = (int *)calloc((size_t)c->n_cells_tot * c->n_nei, sizeof(int));
    if (!c->part_in_cell || !c->r_nei || !c->l_nei || !c->cell_neigh_ls) return -1;
    int nob = 0;
    orc_neigh_list(r0, p->n_part, c->n_cells, c->n_cells_tot, p->boundary, c->inv_l_cell,
                   c->part_in_cell, c->r_nei, c->l_nei, &nob);
    orc_get_cell_neigh_list(c->n_cells_tot, c->n_cells, c->n_nei, c->cell_neigh_ls);
    return 0;
}

void orc_cells_free(orc_cells *c) {
    free(c->part_in_cell); free(c->r_nei); free(c->l_nei); free(c->cell_neigh_ls);
    memset(c, 0, sizeof(*c));
}

/* Dissipative part of one pair (i, j), as pair_body decides and weighs it:
 * with DPD_CUT_OFF the non-head pairs within rc_dpd (their own 1/r), else the
 * pairs within the LJ range (with the switch-on clamp of r^2). */
static inline void dpd_d_pair(const orc_params *p, const double *r0, const double *v, const int *a_type,
                              int i, int j, double r_2_min, double *force_d) {
    int ti = a_type[i], tj = a_type[j];
    double dr[3];
    for (int d = 0; d < 3; ++d) dr[d] = r0[3 * i + d] - r0[3 * j + d];
    dr[0] = mic(dr[0], p->boundary[0], p->inv_boundary[0]);
    dr[1] = mic(dr[1], p->boundary[1], p->inv_boundary[1]);
    double r_2 = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    int lj = r_2 < p->range_2[T2(ti, tj)];
    double w;
    if (p->dpd_cut_off) {
        if (ti == 1 || tj == 1 || !(r_2 < p->r_cut_dpd_2)) return;
        if (lj && p->force_switch_on && r_2 < r_2_min) r_2 = r_2_min;
        w = 1.0 - sqrt(r_2 / p->r_cut_dpd_2);
    } else {
        if (!lj) return;
        if (p->force_switch_on && r_2 < r_2_min) r_2 = r_2_min;
        w = 1.0 - sqrt(r_2 * p->inv_range_2[T2(ti, tj)]);
    }
    double inv = sqrt(1.0 / r_2), rv[3];
    for (int d = 0; d < 3; ++d) rv[d] = dr[d] * inv;
    double rd = (v[3 * i] - v[3 * j]) * rv[0] + (v[3 * i + 1] - v[3 * j + 1]) * rv[1] +
                (v[3 * i + 2] - v[3 * j + 2]) * rv[2];
    for (int d = 0; d < 3; ++d) {
        double f = -1.0 * p->friction * (w * w) * rd * rv[d];
        force_d[3 * i + d] += f;
        force_d[3 * j + d] -= f;
    }
}

/* new_dpd_fd.f90:19-115: the dissipative forces over every pair of the cell
 * list's half shell (each unordered pair once). */
void orc_dpd_dissipative(const orc_params *p, const orc_cells *c, const double *r0, const double *v,
                         const int *a_type, double r_2_min, double *force_d) {
    for (int ic = 1; ic <= c->n_cells_tot; ++ic) {
        for (int ip = c->part_in_cell[ic]; ip != 0; ip = c->r_nei[ip - 1]) {
            for (int jd = 0; jd < c->n_nei; ++jd) {
                int jc = c->cell_neigh_ls[(ic - 1) + (size_t)c->n_cells_tot * jd];
                for (int jp = c->part_in_cell[jc]; jp != 0; jp = c->r_nei[jp - 1])
                    dpd_d_pair(p, r0, v, a_type, ip - 1, jp - 1, r_2_min, force_d);
            }
            for (int jp = c->part_in_cell[ic]; jp != 0; jp = c->r_nei[jp - 1])
                if (ip < jp) dpd_d_pair(p, r0, v, a_type, ip - 1, jp - 1, r_2_min, force_d);
        }
    }
}

void orc_md_steps(const orc_params *p, orc_cells *c, double *r0, double *v, double *a,
                  double *force, const int *a_type, const double *mass, const double *inv_mass,
                  const double *ext_force, long long t_idx0, int n_steps, int noise,
                  orc_energies *e, double *f_on_heads) {
    const size_t n3 = 3 * (size_t)p->n_part;
    for (int s = 0; s < n_steps; ++s) {
        long long t_idx = t_idx0 + s;
        double r_time = (double)t_idx * p->dt;
        orc_verlet_positions(p, r0, v, a, 1, c->n_cells, c->inv_l_cell, c->part_in_cell, c->r_nei, c->l_nei);
        memset(force, 0, n3 * sizeof(double));
        double r_2_min = orc_r2_min(p, r_time);
        orc_fluid_fluid(p, r0, v, a_type, c->part_in_cell, c->r_nei, c->cell_neigh_ls,
                        c->n_cells_tot, c->n_nei, r_2_min, (uint64_t)t_idx, noise, force, &e->v_fluid_fluid);
        e->v_fluid_wall = orc_fluid_wall(p, r0, a_type, force);
        e->v_intra_molec = orc_intra_molec(p, r0, force, NULL);
        e->v_bend = p->bending ? orc_bending(p, r0, force, 0) : 0.0;
        e->v_bend_melt = p->bending_melt ? orc_bending(p, r0, force, 1) : 0.0;
        if (ext_force)
            for (size_t k = 0; k < n3; ++k) force[k] += ext_force[k];
        double *v_half = NULL;
        if (p->dpd_vv && p->thermostat == 1) {
            v_half = (double *)malloc(n3 * sizeof(double));
            memcpy(v_half, v, n3 * sizeof(double));
        }
        e->t_fluid = orc_verlet_velocities(p, force, inv_mass, mass, v, a, f_on_heads);
        if (v_half) {
            /* new_dpd_fd.f90 (md_main.f90:187-198): the dissipative part of
             * the force is evaluated again at the new velocities and replaces
             * the one of the half-step velocities; the accelerations follow.
             * (The reference's routine also subtracts a force_d its pair
             * routine never fills and applies a further half kick; the
             * published DPD-VV scheme, Besold et al. PRE 62 R7611 (2000),
             * only replaces F_D, as here.) */
            double *fd = (double *)calloc(2 * n3, sizeof(double));
            orc_dpd_dissipative(p, c, r0, v_half, a_type, r_2_min, fd);
            orc_dpd_dissipative(p, c, r0, v, a_type, r_2_min, fd + n3);
            for (int i = 0; i < p->n_part; ++i) {
                const int head = i < p->n_mon * p->n_chain && i % p->n_mon == 0;
                for (int d = 0; d < 3; ++d) {
                    size_t k = 3 * (size_t)i + d;
                    if (head) continue;            /* heads keep force and a zero */
                    force[k] = force[k] + (fd[n3 + k] - fd[k]);
                    a[k] = force[k] * inv_mass[i];
                }
            }
            free(fd);
            free(v_half);
        }
    }
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
