/*
 * mfa_hip.h -- C ABI of the MI355X (gfx950) per-step MD path of mfa_brush.
 *
 * Drop-in boundary: each entry point replaces one Fortran subroutine that
 * md_main.f90 calls on the step path (citations are file:line in the
 * reference).  Plain pointers and sizes only; arrays use the reference's
 * `module commons` layout (mfa_common.f90:98-324):
 *   r0, v, a, force  real(8) (3, n_part) column-major  -> double[3*n_part]
 *   a_type           integer (n_part), values 1..4      -> int32_t[n_part]
 *   n_type tables    real(8) (4, 4) column-major         -> double[16]
 * Bead order is the reference's: n_chain grafted chains of n_mon beads
 * (head first), then n_chain_d melt/solvent chains of n_mon_d beads.
 *
 * Every call returns 0 on success or a negative MFA_E_* code; the text of
 * the last error is in mfa_hip_last_error().  A context is used from one
 * host thread (the reference calls the step routines from its master
 * thread, outside any OpenMP region).  One context holds one or more
 * independent channels ("systems", i.e. replicas) advanced together.
 */
#ifndef MFA_HIP_H
#define MFA_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MFA_E_ARG     (-1)   /* invalid argument / unsupported configuration */
#define MFA_E_HIP     (-2)   /* HIP runtime error */
#define MFA_E_STATE   (-3)   /* call out of order (e.g. no state uploaded) */
#define MFA_E_NOMEM   (-4)

#define MFA_PREC_FP64  0     /* all pair arithmetic in FP64 */
#define MFA_PREC_MIXED 1     /* FP32 pair arithmetic on cell-relative positions, FP64 integration and sums */

/* Parameters of one channel.  Tables are the ones init_params.f90:393-887
 * derives; the flags are the hot-path subset of control_simulation.h. */
typedef struct mfa_system_params {
    int32_t n_mon, n_chain;       /* grafted chains (init_system.f90:33) */
    int32_t n_mon_d, n_chain_d;   /* melt / solvent chains (init_system.f90:36) */
    int32_t thermostat;           /* 1: DPD pair forces (THERMOSTAT==0, control_simulation.h:59);
                                     2: Langevin bead forces (THERMOSTAT==1, lgv_forces.f90) */
    int32_t dpd_cut_off;          /* 1: DPD_CUT_OFF defined (control_simulation.h:63) */
    int32_t force_switch_on;      /* FORCE_SWITCH_ON (control_simulation.h:73) */
    int32_t bending;              /* BENDING (control_simulation.h:47) */
    int32_t bending_melt;         /* BENDING_MELT (control_simulation.h:48) */
    int32_t poiseuille;           /* POISEUILLE (control_simulation.h:80) */
    int32_t l_box;                /* L_BOX 1 or 2 (control_simulation.h:105) */
    int32_t bend_form;            /* bending terms: 0 the reference's harmonic angle 0.5 k_bend (theta - alpha_eq)^2
                                     (bending.f90:97-100); 1 cosine k_bend (1 - cos theta) (no counterpart) */
    double boundary[3];           /* box; z = z_space_wall (init_config.f90:17-19) */
    double epsil[16];             /* 4*epsilon, epsil(1,1)=0 */
    double sigma_2[16];
    double range_2[16];           /* cutoff^2, 0 disables the pair */
    double e_shift[16];
    double inv_range_2[16];
    double r_cut_max;
    double r_cut_dpd_2;           /* used when dpd_cut_off */
    double skin;                  /* 0: cells of r_cut_max rebuilt every step (BIN_TYPE 2);
                                     >0: cells of r_cut_max+skin rebuilt when a bead moved
                                     more than skin/2 (check_skin.f90:11-24) */
    double sig;                   /* sqrt(2 kT gamma/dt) (init_params.f90:887) */
    double friction;              /* gamma = friction(1) (dpd_forces.f90:105) */
    double r_2_min_init, r_2_min_time;   /* force switch-on ramp (init_force_switch_on.f90:14-21) */
    double sigma_wall[4], a_wall[4];     /* init_params.f90:593-596 */
    double a_w;
    double k_chain;               /* k * R0^2 (init_params.f90:396) */
    double inv_r_chain_2;
    double k_bend, alpha_eq;
    double dt;
    double const_force[3];        /* Poiseuille body force g (init_config.f90:370-378) */
    double head_vel[2];           /* va_spring_twall(1:2) for heads when !poiseuille */
    double mass_type[4];
    uint64_t seed;                /* Philox key of the DPD noise */
    int64_t t_idx0;               /* time index of the next step: r_time = t_idx*dt */
    int32_t dpd_vv;               /* DPD_VV (control_simulation.h:65): md_steps re-evaluates the dissipative
                                     forces after the second half kick (new_dpd_fd.f90) */
    int32_t reserved1;
} mfa_system_params;

typedef struct mfa_energies {
    double v_fluid_fluid;   /* fluid_fluid.f90:137 */
    double v_fluid_wall;    /* fluid_wall.f90:98,116 */
    double v_intra_molec;   /* intra_molec.f90:72,129,266 */
    double v_bend;          /* bending.f90:100 */
    double v_bend_melt;     /* bending_melt.f90 */
    double t_fluid;         /* 0.5 sum m v^2, verlet_velocities.f90:132-153 */
    double virial[6];       /* virial W = sum over interacting pairs and FENE bonds of r_ij (x) F_ij, xx yy zz
                               xy xz yz, r_ij the minimum-image separation i - j and F_ij the force on i from
                               j: for pairs the whole pair force, LJ, DPD dissipative and DPD random parts (the
                               reference's bulk-mode press_tensor, dpd_forces.f90:118-127, keeps only the
                               dissipative DPD part; it has none in channel mode); for bonds the FENE force
                               (intra_molec.f90:61-66,115-120, whose brush loop adds the opposite sign).  The
                               pair part is that of the last fluid_fluid (or fused step), the bond part that
                               of the intra_molec after it (or of the fused step); walls and bending add
                               none.  Zero while mfa_hip_set_virial is off */
    double virial_bond[6];  /* the FENE bond part of virial */
} mfa_energies;

typedef struct mfa_flags {
    int32_t nonfinite;       /* beads with NaN/Inf position */
    int32_t out_of_box_z;    /* beads with z outside [0, Lz] (fluid_wall.f90:86,106 warnings) */
    int32_t fene_stretched;  /* bonds clamped at 0.98 R0^2 (intra_molec.f90:39-43) */
    int32_t rebuilds;        /* cell-list rebuilds since creation */
    int32_t list_overflow;   /* beads whose Verlet list exceeded the capacity (they scanned the cells) */
} mfa_flags;

typedef struct mfa_ctx mfa_ctx;

/* ---- lifetime and state -------------------------------------------------
 * mfa_hip_create allocates device buffers for n_sys channels (replaces the
 * allocation of init_system.f90:196-287 on the device side). */
int  mfa_hip_create(const mfa_system_params *sys, int32_t n_sys, int32_t precision, mfa_ctx **out);
int  mfa_hip_destroy(mfa_ctx *ctx);
/* host -> device copy of one channel's state (conf_old / gen_* output,
 * init_config.f90:164-173).  a may be NULL (zero).  Builds the cell list
 * (make_binning + neigh_list + get_cell_neigh_list, init_config.f90:261-272). */
int  mfa_hip_set_state(mfa_ctx *ctx, int32_t isys, const double *r0, const double *v,
                       const double *a, const int32_t *a_type);
/* device -> host, reference bead order; any pointer may be NULL.
 * Used at the host-consumer sites observation / store_config / obser_out
 * (md_main.f90:202-217). */
int  mfa_hip_get_state(mfa_ctx *ctx, int32_t isys, double *r0, double *v, double *a, double *force);
int  mfa_hip_set_time(mfa_ctx *ctx, int64_t t_idx);   /* r_time = t_idx*dt (md_main.f90:79) */
int  mfa_hip_get_time(mfa_ctx *ctx, int64_t *t_idx);

/* ---- step routines (drop-in for md_main.f90:85-185) -------------------- */
int  mfa_hip_verlet_positions(mfa_ctx *ctx);       /* verlet_positions.f90:150-209 + cell relink/rebuild */
int  mfa_hip_check_skin(mfa_ctx *ctx, int32_t *f_skin); /* check_skin.f90:4-28: f_skin of the last
                                                      verlet_positions, 1 when its skin test fired (some bead
                                                      more than skin/2 from its position at the last rebuild),
                                                      else 0.  The skin test and the rebuild it calls for
                                                      (binning.f90 / my_binning.f90, md_main.f90:97-102) run
                                                      inside verlet_positions, so the rebuild is done by then;
                                                      the rebuild count is in mfa_flags */
int  mfa_hip_fluid_fluid(mfa_ctx *ctx);            /* fluid_fluid.f90:1-512 + dpd_forces.f90; also
                                                      performs md_main.f90:107 force(:,:)=0 */
int  mfa_hip_fluid_wall(mfa_ctx *ctx, int32_t inter_type);   /* fluid_wall.f90:77-123 (inter_type 2) */
int  mfa_hip_wall_wall(mfa_ctx *ctx, int32_t inter_type, double sigma_w, double *v_wall_wall); /* wall_wall.f90:45-47 */
int  mfa_hip_intra_molec(mfa_ctx *ctx);            /* intra_molec.f90:1-268 */
int  mfa_hip_bending(mfa_ctx *ctx, int32_t mode);  /* bending.f90:19-121 (mode 1) */
int  mfa_hip_bending_melt(mfa_ctx *ctx, int32_t mode);   /* bending_melt.f90:22-120 (mode 1) */
int  mfa_hip_constant_force(mfa_ctx *ctx, int32_t case_pois); /* constant_force.f90:17-21 (case 1) */
int  mfa_hip_verlet_velocities(mfa_ctx *ctx);      /* verlet_velocities.f90:18-153; ends the step */
/* new_dpd_fd.f90:1-216 (md_main.f90:187-198, DPD_VV), right after
 * verlet_velocities: the dissipative forces of every DPD channel are
 * evaluated again at the new velocities and replace those of the half-step
 * velocities in the force and the acceleration the next step starts from.
 * The velocities and the kinetic energy are unchanged.  (The reference's
 * routine also subtracts a force_d that its pair routine never fills and
 * adds a further half kick; this follows the published DPD-VV scheme,
 * Besold et al., PRE 62, R7611 (2000), which only replaces F_D.)  Heads keep
 * zero force; a head's velocity change is taken as zero (heads are frozen). */
int  mfa_hip_new_dpd_fd(mfa_ctx *ctx);

/* Fused steps: n_steps of md_main.f90:77-185 in 4 kernels per step
 * (integrate+skin check, [gated cell rebuild], pair, bonded+wall+kick).
 * use_graph=1 replays captured graphs of 8 steps and of 1 step.  The force
 * that mfa_hip_get_state returns afterwards is the total force of the last
 * step, and the energies and pair count that mfa_hip_get_energies and
 * mfa_hip_get_pair_count return are the last step's (the steps before it
 * neither write the force back nor sum the energies). */
int  mfa_hip_md_steps(mfa_ctx *ctx, int32_t n_steps, int32_t use_graph);
/* Captures the graphs md_steps(use_graph=1) replays, without running a step
 * (a timed region then pays no capture); a no-op while timing. */
int  mfa_hip_prepare_graphs(mfa_ctx *ctx);
int  mfa_hip_synchronize(mfa_ctx *ctx);

/* ---- observables and checks ------------------------------------------- */
int  mfa_hip_get_energies(mfa_ctx *ctx, int32_t isys, mfa_energies *out);
int  mfa_hip_get_f_on_heads(mfa_ctx *ctx, int32_t isys, double *f_on_heads /* (3, n_chain) */);
int  mfa_hip_get_flags(mfa_ctx *ctx, mfa_flags *out);
/* check_fluid_fluid, util.f90:284-330: O(N^2) brute-force LJ energy */
int  mfa_hip_check_fluid_fluid(mfa_ctx *ctx, int32_t isys, double *v_fluid_fluid);
/* check_fluid_fluid without any allocation.  No counterpart in the reference:
 * an extra entry point for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Same result bits as mfa_hip_check_fluid_fluid on the same positions, but
 * on caller-owned device arrays and without a context:
 *   d_r0      device, (3, n_part) FP64 in the commons layout
 *   d_a_type  device, n_part int32 in 1..4
 *   r_2_min   switch-on clamp of r^2 (0 when FORCE_SWITCH_ON is off)
 *   d_v_ff    device, one double: receives V_ff
 *   stream    a hipStream_t, or NULL for the null stream
 * Only sys->boundary, epsil, sigma_2, range_2 and e_shift are read.  The
 * per-bead energies go to a static device array of MFA_NOALLOC_MAX_PART
 * doubles (8 MiB, held by each GPU the library runs on until the process
 * ends); n_part beyond it is rejected with MFA_E_ARG.  Every call on one GPU
 * shares that array: calls that overlap on one GPU (other streams, other
 * threads) race.  The call is asynchronous on `stream`. */
#define MFA_NOALLOC_MAX_PART (1 << 20)
int  mfa_hip_check_fluid_fluid_noalloc(const mfa_system_params *sys, const double *d_r0, const int32_t *d_a_type,
                                       int32_t n_part, double r_2_min, double *d_v_ff, void *stream);
/* interacting-pair fingerprint per bead (reference order): number of partners
 * with r^2 < range_2 and the wrapping sum of mix64(partner index) */
int  mfa_hip_pair_stats(mfa_ctx *ctx, int32_t isys, int32_t *count, uint64_t *hash);
/* histo(2) functions.f90:157-213 and velocity_prof(2) :319-365, accumulated
 * on the device: dens_z (n_bins,2) brush/melt counts; v_prof, v_prof_2
 * (n_bins,3,2) sums of v and v^2 per bin, species brush/melt. */
int  mfa_hip_profiles_reset(mfa_ctx *ctx, int32_t n_bins);
int  mfa_hip_profiles_accumulate(mfa_ctx *ctx);
int  mfa_hip_profiles_get(mfa_ctx *ctx, int32_t isys, int32_t *n_obs, double *dens_z,
                          double *v_prof, double *v_prof_2);
/* make_binning, cell_list.f90:129-158 (host): cells for a cutoff */
int  mfa_hip_make_binning(int32_t bin_case, const double *boundary, double r_cut, double *l_cell,
                          double *inv_l_cell, int32_t *n_cells, int32_t *n_cells_tot, int32_t *n_nei);
/* the device cell list of one channel: cell_start/cell_count (n_cells_tot)
 * and the bead (reference index, 0-based) in each cell slot (n_part) */
int  mfa_hip_get_cells(mfa_ctx *ctx, int32_t isys, int32_t *n_cells, int32_t *cell_start,
                       int32_t *cell_count, int32_t *cell_beads);

/* ---- timing ------------------------------------------------------------
 * Mode 1: each fused-step kernel group is bracketed by hipEvents on the
 * context's stream: "vv1" (the first half kick, once per md_steps call when
 * the step uses the fused integrator), "rebuild", "pair", "bonded" (bending
 * if any, then the integrator; the bonded kernel on the routine path) and
 * "dpd_vv".  Mode 2: only the pair kernel is.  Mode 0: off.  While timing,
 * md_steps launches eagerly.  Totals are in milliseconds. */
int  mfa_hip_set_timing(mfa_ctx *ctx, int32_t on);
int  mfa_hip_get_kernel_time(mfa_ctx *ctx, const char *kernel, double *total_ms, int64_t *launches);
int  mfa_hip_set_virial(mfa_ctx *ctx, int32_t on);   /* accumulate the FP64 pair and bond virial (default off);
                                                        waits for the context's stream and drops the captured
                                                        graphs */

/* Verlet-list capacity (entries per bead) of the pair kernels: -1 automatic
 * (default), 0 no list (the pair kernels scan the 27 or 125 neighbour cells
 * every step), > 0 explicit (rounded up to a multiple of 8).  The list holds the partners within the
 * interaction range plus the skin; it is rebuilt with the cells, so the pair
 * set and the forces are the same bits with or without it.  No counterpart in
 * the reference (its BIN_TYPE 0/1 Verlet lists are host data structures,
 * check_skin.f90 / binning.f90); a tuning knob of this implementation. */
int  mfa_hip_set_neighbor_list(mfa_ctx *ctx, int32_t capacity);
int  mfa_hip_get_neighbor_list(mfa_ctx *ctx, int32_t *capacity);
int  mfa_hip_get_pair_count(mfa_ctx *ctx, int32_t isys, int64_t *n_pairs);  /* interacting pairs, last evaluation */

/* Chain observables of channel isys from the device state (the gyration and
 * momentum part of observation.f90:182-290): out[0..2] the brush chains'
 * mean squared radius of gyration per axis (r_g2(1,:)), out[3..5] the melt
 * chains' (r_g2(2,:)), out[6..8] the total momentum sum(mass * v) (tot_P
 * before its running average).  Chains are unfolded across the periodic
 * images as the reference does.  The anisotropy r_par_per2 follows from
 * out[0..5] on the host. */
int  mfa_hip_chain_observables(mfa_ctx *ctx, int32_t isys, double *out);

/* Verlet-list statistics of channel isys (no reference counterpart; a
 * diagnostic of the pair kernels' work, outside the step): out[0] the list
 * entries of its beads, out[1] the entries its waves walk (a wave of 64
 * slots walks the longest list of its lanes, in groups of four), out[2] its
 * beads whose list overflowed the capacity (they scan their cells).  Lists
 * are those of the last rebuild; fails when the list is off (capacity 0). */
int  mfa_hip_list_stats(mfa_ctx *ctx, int32_t isys, int64_t *out);

/* DPD pairs each bead defers (0..8, default 8).  The pair kernels evaluate
 * the LJ part of every pair while walking the partners and queue the DPD part
 * (Philox normal, random and dissipative forces) of up to `depth` pairs per
 * bead for a second pass with all lanes busy; pairs past the queue have their
 * DPD part evaluated in place.  The forces are the same for every depth up
 * to the FP64 summation order of the two parts.  No counterpart in the
 * reference; a tuning and test knob of this implementation. */
int  mfa_hip_set_dpd_queue(mfa_ctx *ctx, int32_t depth);

/* Launch shape of the mixed pair kernel: 0 (default) one block per tile of
 * 256 slots; > 0 a persistent launch of that many blocks that take the tiles
 * from a device work counter (fewer, longer-lived blocks; the order of work
 * inside a tile and the results are the same bits).  No counterpart in the
 * reference; a tuning knob (mfa_brush_amd/autotune.py). */
int  mfa_hip_set_pair_grid(mfa_ctx *ctx, int32_t blocks);

const char *mfa_hip_last_error(void);
int  mfa_hip_version(void);

#ifdef __cplusplus
}
#endif
#endif
