/*
 * mfa_oracle.h -- CPU FP64 restatement of the mfa_brush per-step MD path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the test reference for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (mfa_brush_amd/, include/) never links or calls it.
 *
 * Conventions follow the reference's `module commons` (mfa_common.f90):
 *   - r0, v, a, force: (3, n_part) column-major FP64, i.e. x,y,z of one bead
 *     contiguous: r0[3*i + d] for 0-based bead i (reference index i+1);
 *   - a_type[i] in 1..4 (1 = brush head, 2 = brush, 3 = melt/solvent,
 *     4 = implicit-wall type, never interacting);
 *   - n_type x n_type tables are Fortran column-major: tab[(ti-1) + 4*(tj-1)];
 *   - cell lists are 1-based with phantom cell 0 and 0-terminated links.
 */
#ifndef MFA_ORACLE_H
#define MFA_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NT 4

typedef struct {
    int n_part;            /* n_mon_tot */
    int n_mon, n_chain;    /* grafted chains: beads 1..n_mon*n_chain, head first */
    int n_mon_d, n_chain_d;/* melt / solvent chains after the brush */
    int thermostat;        /* 1: DPD pair forces (THERMOSTAT==0); 2: Langevin bead forces
                              (THERMOSTAT==1, lgv_forces.f90) */
    int dpd_cut_off;       /* 1: DPD_CUT_OFF defined (own cutoff, heads excluded) */
    int force_switch_on;   /* FORCE_SWITCH_ON defined */
    int bending;           /* BENDING (brush chains) */
    int bending_melt;      /* BENDING_MELT (melt chains, index bug Q12 fixed) */
    int poiseuille;        /* POISEUILLE: heads get v=0 (else +-head_vel in x,y) */
    int l_box;             /* 1: l_cell >= r_cut (13 cells); 2: r_cut/2 (62 cells) */
    double boundary[3];
    double inv_boundary[3];
    double epsil[16];      /* 4*epsilon, epsil(1,1)=0 (init_params.f90:468,492) */
    double sigma_2[16];
    double range_2[16];    /* cutoff^2; 0 disables the pair */
    double e_shift[16];
    double inv_range_2[16];
    double r_cut_max;
    double r_cut_dpd_2;
    double sig;            /* sqrt(2 kT gamma / dt) (init_params.f90:887) */
    double friction;       /* gamma = friction(1) */
    double r_2_min_init, r_2_min_time;
    double sigma_wall[4], a_wall[4];
    double a_w, z_space_wall;
    double k_chain;        /* k * R0^2 (init_params.f90:396) */
    double inv_r_chain_2;
    double k_bend, alpha_eq;
    double dt, dt_2;
    double const_force[3];
    double head_vel[2];
    double mass_type[4];
    uint64_t seed;         /* Philox key for the DPD noise */
    int bend_form;         /* 0: harmonic angle, bending.f90:97-100; 1: cosine k_bend (1 - cos theta) */
    int dpd_vv;            /* DPD_VV: dissipative forces re-evaluated after verlet_velocities (new_dpd_fd.f90) */
} orc_params;

/* ---- RNG: Philox-4x32-10 (Salmon et al., SC'11) and the pair normal ---- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_pair_gauss(uint64_t seed, uint64_t step, int32_t i, int32_t j);
float orc_lgv_gauss(uint64_t seed, uint64_t step, int32_t i, int32_t d);

/* ---- cell list (cell_list.f90) ---- */
void orc_make_binning(int bin_case, const double boundary[3], double r_cut,
                      double l_cell[3], double inv_l_cell[3], int n_cells[3],
                      int *n_cells_tot, int *n_nei);
int  orc_get_cell(const double r[3], const int n_cells[3], const double inv_l_cell[3]);
void orc_neigh_list(const double *r0, int n_part, const int n_cells[3], int n_cells_tot,
                    const double boundary[3], const double inv_l_cell[3],
                    int *part_in_cell, int *r_nei, int *l_nei, int *n_out_of_box);
void orc_update_part_cell(const double r[3], int i_part, const int n_cells[3],
                          const double inv_l_cell[3], int old_cell,
                          int *part_in_cell, int *r_nei, int *l_nei);
void orc_get_cell_neigh_list(int n_cells_tot, const int n_cells[3], int n_nei,
                             int *cell_neigh_ls);

/* ---- forces ---- */
double orc_r2_min(const orc_params *p, double r_time);
void orc_fluid_fluid(const orc_params *p, const double *r0, const double *v, const int *a_type,
                     const int *part_in_cell, const int *r_nei, const int *cell_neigh_ls,
                     int n_cells_tot, int n_nei, double r_2_min, uint64_t step,
                     int noise, double *force, double *v_fluid_fluid);
double orc_check_fluid_fluid(const orc_params *p, const double *r0, const int *a_type, double r_2_min);
void orc_pair_stats(const orc_params *p, const double *r0, const int *a_type,
                    int32_t *count, uint64_t *hash);
int  orc_pair_stats_cells(const orc_params *p, const double *r0, const int *a_type,
                          int32_t *count, uint64_t *hash);
void orc_fluid_fluid_brute(const orc_params *p, const double *r0, const double *v, const int *a_type,
                           double r_2_min, uint64_t step, int noise, double *force, double *v_ff);
/* pair virial sum_{i<j} r_ij (x) F_ij (xx yy zz xy xz yz), all pair-force parts */
void orc_pair_virial(const orc_params *p, const double *r0, const double *v, const int *a_type,
                     double r_2_min, uint64_t step, int noise, double W[6]);
double orc_fluid_wall(const orc_params *p, const double *r0, const int *a_type, double *force);
double orc_wall_wall(const orc_params *p, double sigma_w);
double orc_intra_molec(const orc_params *p, const double *r0, double *force, int *n_stretched);
void orc_bond_virial(const orc_params *p, const double *r0, double W[6]);
double orc_bending(const orc_params *p, const double *r0, double *force, int melt);
void orc_constant_force(const orc_params *p, const double *mass, const int *a_type, double *force);

/* ---- integrator ---- */
void orc_verlet_positions(const orc_params *p, double *r0, double *v, const double *a,
                          int relink, const int n_cells[3], const double inv_l_cell[3],
                          int *part_in_cell, int *r_nei, int *l_nei);
double orc_verlet_velocities(const orc_params *p, double *force_in, const double *inv_mass,
                             const double *mass, double *v, double *a, double *f_on_heads);

/* ---- observables (functions.f90 histo / velocity_prof mode 2) ---- */
void orc_histo(const orc_params *p, const double *r0, int hist_dim, double *dens_z);
void orc_velocity_prof(const orc_params *p, const double *r0, const double *v, int n_layers,
                       double *v_prof, double *v_prof_2);

/* ---- one md_main step (md_main.f90:77-215, BIN_TYPE 2) ---- */
typedef struct {
    double v_fluid_fluid, v_intra_molec, v_fluid_wall, v_bend, v_bend_melt, t_fluid;
} orc_energies;

typedef struct {
    int n_cells[3];
    int n_cells_tot, n_nei;
    double l_cell[3], inv_l_cell[3];
    int *part_in_cell;   /* 0..n_cells_tot */
    int *r_nei, *l_nei;
    int *cell_neigh_ls;  /* (n_cells_tot, n_nei) column-major */
} orc_cells;

int  orc_cells_init(orc_cells *c, const orc_params *p, const double *r0, double r_cut);
void orc_cells_free(orc_cells *c);
void orc_md_steps(const orc_params *p, orc_cells *c, double *r0, double *v, double *a,
                  double *force, const int *a_type, const double *mass, const double *inv_mass,
                  const double *ext_force, long long t_idx0, int n_steps, int noise,
                  orc_energies *e, double *f_on_heads);
/* dissipative DPD forces of every DPD pair at velocities v (the pair set and
 * weights of fluid_fluid's DPD part at this r_2_min), added into force_d */
void orc_dpd_dissipative(const orc_params *p, const orc_cells *c, const double *r0, const double *v,
                         const int *a_type, double r_2_min, double *force_d);
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
