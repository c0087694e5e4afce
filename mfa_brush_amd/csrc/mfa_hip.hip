// mfa_hip.hip -- C ABI (include/mfa_hip.h) and host orchestration of the
// gfx950 per-step path.  One HIP stream per context; every kernel of a step
// is enqueued on it, and the cell-list rebuild is decided on the device (a
// flag written by the integrator and read by the gated rebuild kernels), so
// a step needs no host round trip and can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mfa_hip.h"
#include "mfa_kernels.h"
#include "mfa_pair.h"
#include "mfa_step.h"

using namespace mfa;

static thread_local std::string g_err;

static int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess)                                                              \
            return fail(MFA_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_));        \
    } while (0)

namespace {

// grid of the rebuild kernels (grid-stride): eight blocks per CU keep more
// tiles' loads and atomics in flight than four (rebuild 566 -> 541 us per
// step on 256 replicas, 3.5M beads) and cost nothing measurable when idle
constexpr int GATED_GRID = 2048;
constexpr int GRAPH_STEPS = 8;       // steps per replayed graph (plus a one-step graph for the rest)

struct TimedLaunch {
    hipEvent_t a, b;
    int kernel;
};

enum KId { K_VV1 = 0, K_REBUILD, K_PAIR, K_BONDED, K_DPDVV, K_NK };
const char *kNames[K_NK] = {"vv1", "rebuild", "pair", "bonded", "dpd_vv"};

}  // namespace

struct mfa_ctx {
    int n_sys = 0, prec = 0;
    std::vector<mfa_system_params> hp;
    std::vector<SysDev> hs;
    int S_tot = 0, N_tot = 0, C_tot = 0, n_blocks = 0, H_tot = 0;
    hipStream_t st = nullptr;
    SysDev *d_sys = nullptr;
    int *d_blk_sys = nullptr, *d_cell_sys = nullptr;
    Bufs B{};
    int *d_cell_count = nullptr, *d_pre = nullptr, *d_bsum = nullptr;
    int2 *d_crange = nullptr;
    int *d_bin_cell = nullptr, *d_bin_rank = nullptr, *d_perm = nullptr, *d_key_tmp = nullptr, *d_bin_key = nullptr;
    double *d_partial = nullptr, *d_energies = nullptr, *d_fheads = nullptr;
    double4 *d_dfd = nullptr;      // DPD-VV: change of each bead's dissipative force (slot order)
    int *d_tiles = nullptr;        // persistent mixed pair kernel: work counter, ticket counter
    int pair_grid = 0;             // 0: one block per tile; > 0: persistent launch of this many blocks
    int *d_rebuild_count = nullptr, *d_permute_done = nullptr;
    int *d_skin_fired = nullptr;   // whether the last verlet_positions' skin test fired (f_skin)
    double *d_obs = nullptr;       // block partials of the chain observables (2 x channel blocks x 8)
    double *d_stage = nullptr;
    int *d_istage = nullptr;
    char *h_stage = nullptr;      // pinned staging buffer for every host<->device copy
    // profiles
    int prof_nb = 0, prof_nobs = 0;
    unsigned int *d_pcnt = nullptr;
    unsigned long long *d_pvs = nullptr, *d_pv2s = nullptr;
    double *d_pdens = nullptr, *d_pvp = nullptr, *d_pvp2 = nullptr;
    bool have_state = false;
    std::vector<char> uploaded;
    bool virial = false;
    int nl_request = -1;          // Verlet-list capacity: -1 automatic, 0 off, > 0 entries per bead
    // timing: bit k of timing_mask set when kernel group k is bracketed by events
    int timing_mask = 0;
    std::vector<hipEvent_t> ev_pool;
    std::vector<TimedLaunch> pending;
    double k_ms[K_NK] = {};
    long long k_n[K_NK] = {};
    // graph
    hipGraphExec_t gexec = nullptr;      // one step
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec_k = nullptr;    // GRAPH_STEPS steps
    hipGraph_t graph_k = nullptr;
};

static hipEvent_t take_event(mfa_ctx *c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

struct Timer {
    mfa_ctx *c;
    int kid;
    hipEvent_t a = nullptr;
    Timer(mfa_ctx *c_, int k) : c(c_), kid(k) {
        if (c->timing_mask & (1 << kid)) { a = take_event(c); if (a) (void)hipEventRecord(a, c->st); }
    }
    ~Timer() {
        if (a) {
            hipEvent_t b = take_event(c);
            if (b) { (void)hipEventRecord(b, c->st); c->pending.push_back({a, b, kid}); }
        }
    }
};

static void add_elapsed(mfa_ctx *c, const TimedLaunch &t) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, t.a, t.b) == hipSuccess) {
        c->k_ms[t.kernel] += ms;
        c->k_n[t.kernel] += 1;
    }
}

static int collect_timing(mfa_ctx *c) {
    if (c->pending.empty()) return 0;
    HIPCHK(hipStreamSynchronize(c->st));
    for (auto &t : c->pending) {
        add_elapsed(c, t);
        c->ev_pool.push_back(t.a);
        c->ev_pool.push_back(t.b);
    }
    c->pending.clear();
    return 0;
}

template <class T>
static int dalloc(T **p, size_t n) {
    if (n == 0) n = 1;
    HIPCHK(hipMalloc((void **)p, n * sizeof(T)));
    HIPCHK(hipMemset(*p, 0, n * sizeof(T)));
    return 0;
}

// Host<->device transfers.  The stream is synchronised first, so an error of
// an earlier kernel is reported by that synchronisation; the data then moves
// in chunks between device memory and the context's pinned staging buffer
// (DMA only ever touches pinned host memory), and the CPU copies between the
// staging buffer and the caller's memory.
constexpr size_t STAGE_BYTES = 16u << 20;

static int sync_for(mfa_ctx *c, const char *what) {
    hipError_t e = hipStreamSynchronize(c->st);
    if (e != hipSuccess) return fail(MFA_E_HIP, std::string("device work before ") + what + ": " + hipGetErrorString(e));
    return 0;
}

static int to_host(mfa_ctx *c, void *dst, const void *src, size_t bytes, const char *what) {
    if (int rc = sync_for(c, what)) return rc;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        hipError_t e = hipMemcpyAsync(c->h_stage, (const char *)src + off, n, hipMemcpyDeviceToHost, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) return fail(MFA_E_HIP, std::string("device-to-host copy of ") + what + ": " + hipGetErrorString(e));
        memcpy((char *)dst + off, c->h_stage, n);
    }
    return 0;
}

static int to_device(mfa_ctx *c, void *dst, const void *src, size_t bytes, const char *what) {
    if (int rc = sync_for(c, what)) return rc;
    for (size_t off = 0; off < bytes; off += STAGE_BYTES) {
        const size_t n = std::min(STAGE_BYTES, bytes - off);
        memcpy(c->h_stage, (const char *)src + off, n);
        hipError_t e = hipMemcpyAsync((char *)dst + off, c->h_stage, n, hipMemcpyHostToDevice, c->st);
        if (e == hipSuccess) e = hipStreamSynchronize(c->st);
        if (e != hipSuccess) return fail(MFA_E_HIP, std::string("host-to-device copy of ") + what + ": " + hipGetErrorString(e));
    }
    return 0;
}

static void free_all(mfa_ctx *c) {
    auto F = [](void *p) { if (p) (void)hipFree(p); };
    F(c->B.pos); F(c->B.vel); F(c->B.acc); F(c->B.posb); F(c->B.posf); F(c->B.velf); F(c->B.cell_of);
    F(c->B.frc); F(c->B.where); F(c->B.parity); F(c->B.rebuild); F(c->B.chan_rb); F(c->B.flags); F(c->B.t_idx); F(c->B.t_end); F(c->B.store_frc);
    F(c->B.nl); F(c->B.nl_n);
    F(c->d_sys); F(c->d_blk_sys); F(c->d_cell_sys); F(c->d_cell_count); F(c->d_pre); F(c->d_bsum);
    F(c->d_crange); F(c->d_bin_cell); F(c->d_bin_rank); F(c->d_perm); F(c->d_key_tmp); F(c->d_bin_key); F(c->d_partial);
    F(c->d_energies); F(c->d_fheads); F(c->d_rebuild_count); F(c->d_skin_fired); F(c->d_permute_done); F(c->d_obs); F(c->d_stage); F(c->d_istage);
    F(c->d_pcnt); F(c->d_pvs); F(c->d_pv2s); F(c->d_pdens); F(c->d_pvp); F(c->d_pvp2); F(c->d_dfd); F(c->d_tiles);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    for (auto e : c->ev_pool) (void)hipEventDestroy(e);
    for (auto &t : c->pending) { (void)hipEventDestroy(t.a); (void)hipEventDestroy(t.b); }
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    if (c->gexec_k) (void)hipGraphExecDestroy(c->gexec_k);
    if (c->graph_k) (void)hipGraphDestroy(c->graph_k);
    if (c->st) (void)hipStreamDestroy(c->st);
}

// make_binning, cell_list.f90:129-158
static void make_binning(int bin_case, const double *boundary, double r_cut, double *l_cell, double *inv_l_cell,
                         int *n_cells, int *n_cells_tot, int *n_nei) {
    for (int d = 0; d < 3; ++d)
        n_cells[d] = bin_case == 2 ? (int)(boundary[d] / (r_cut / 2.0)) : (int)(boundary[d] / r_cut);
    *n_nei = bin_case == 2 ? 62 : 13;
    int tot = 1;
    for (int d = 0; d < 3; ++d) {
        l_cell[d] = boundary[d] / (double)n_cells[d];
        inv_l_cell[d] = 1.0 / l_cell[d];
        tot *= n_cells[d];
    }
    *n_cells_tot = tot;
}

static int build_sysdev(const mfa_system_params &p, SysDev &d) {
    memset(&d, 0, sizeof(d));
    for (int k = 0; k < 3; ++k) {
        if (!(p.boundary[k] > 0.0)) return fail(MFA_E_ARG, "boundary must be positive");
        d.L[k] = p.boundary[k];
        d.invL[k] = 1.0 / p.boundary[k];
    }
    if (p.l_box != 1 && p.l_box != 2) return fail(MFA_E_ARG, "l_box must be 1 or 2");
    if (p.n_mon < 0 || p.n_chain < 0 || p.n_mon_d < 0 || p.n_chain_d < 0)
        return fail(MFA_E_ARG, "negative chain counts");
    if (p.n_chain > 0 && p.n_mon < 1) return fail(MFA_E_ARG, "n_mon must be >= 1 with grafted chains");
    if (p.n_chain_d > 0 && p.n_mon_d < 1) return fail(MFA_E_ARG, "n_mon_d must be >= 1 with melt chains");
    if (!(p.dt > 0.0)) return fail(MFA_E_ARG, "dt must be positive");
    if (p.skin < 0.0) return fail(MFA_E_ARG, "skin must be >= 0");
    const double rc = p.r_cut_max + p.skin;
    if (!(rc > 0.0)) return fail(MFA_E_ARG, "r_cut_max + skin must be positive");
    int nnei = 0, nct = 0;
    make_binning(p.l_box, p.boundary, rc, d.l_cell, d.inv_l_cell, d.nc, &nct, &nnei);
    d.S = p.l_box == 2 ? 2 : 1;
    // The half-shell stencil of cell_list.f90:345-402 repeats a cell when a
    // periodic direction has fewer than 2S+1 cells (the reference then
    // double counts pairs); such boxes are rejected.
    if (d.nc[0] < 2 * d.S + 1 || d.nc[1] < 2 * d.S + 1)
        return fail(MFA_E_ARG, "box too small for the cell stencil: need >= 2*L_BOX+1 cells in x and y");
    if (d.nc[2] < 1) return fail(MFA_E_ARG, "box shorter than one cell in z");
    d.n_cells_tot = nct;
    for (int t = 0; t < 16; ++t) {
        d.epsil[t] = p.epsil[t]; d.sigma_2[t] = p.sigma_2[t]; d.range_2[t] = p.range_2[t];
        d.e_shift[t] = p.e_shift[t]; d.inv_range_2[t] = p.inv_range_2[t];
        d.epsil_f[t] = (float)p.epsil[t]; d.sigma_2f[t] = (float)p.sigma_2[t];
        d.range_2f_lo[t] = (float)(p.range_2[t] * (1.0 - 1.0 / 16384.0));
        d.range_2f_hi[t] = (float)(p.range_2[t] * (1.0 + 1.0 / 16384.0));
        d.e_shift_f[t] = (float)p.e_shift[t];
        d.inv_rc_f[t] = (float)sqrt(p.inv_range_2[t]);
    }
    d.r_cut_dpd_2 = p.r_cut_dpd_2;
    if (p.dpd_cut_off && !(p.r_cut_dpd_2 > 0.0)) return fail(MFA_E_ARG, "dpd_cut_off needs r_cut_dpd_2 > 0");
    // the cells and the Verlet list are complete between rebuilds for pairs
    // within r_cut_max (every bead stays within skin/2 of its build position),
    // so a DPD range beyond it would lose pairs: init_params.f90:801 folds
    // the DPD cutoff into r_cut_max, and a caller must do the same
    if (p.thermostat == 1 && p.dpd_cut_off && sqrt(p.r_cut_dpd_2) > p.r_cut_max)
        return fail(MFA_E_ARG, "DPD cutoff exceeds r_cut_max (fold it into r_cut_max as init_params.f90:801 does)");
    d.sig = p.sig; d.friction = p.friction;
    d.r_2_min_init = p.r_2_min_init; d.r_2_min_time = p.r_2_min_time;
    for (int t = 0; t < 4; ++t) {
        d.sigma_wall[t] = p.sigma_wall[t]; d.a_wall[t] = p.a_wall[t];
        d.mass_type[t] = p.mass_type[t];
        d.inv_mass_type[t] = p.mass_type[t] != 0.0 ? 1.0 / p.mass_type[t] : 0.0;
    }
    d.a_w = p.a_w;
    d.k_chain = p.k_chain; d.inv_r_chain_2 = p.inv_r_chain_2; d.k_bend = p.k_bend; d.alpha_eq = p.alpha_eq;
    d.dt = p.dt; d.dt_2 = p.dt * p.dt;
    for (int k = 0; k < 3; ++k) d.const_force[k] = p.const_force[k];
    d.head_vel[0] = p.head_vel[0]; d.head_vel[1] = p.head_vel[1];
    d.skin_mode = p.skin > 0.0 ? 1 : 0;
    d.skin_half_2 = 0.25 * p.skin * p.skin;
    // Verlet-list radius per type pair: the LJ range, or the DPD range where
    // the thermostat acts beyond it (DPD_CUT_OFF), plus the skin
    for (int t = 0; t < 16; ++t) {
        double r = p.range_2[t] > 0.0 ? sqrt(p.range_2[t]) : 0.0;
        if (p.thermostat == 1 && p.dpd_cut_off) r = std::max(r, sqrt(p.r_cut_dpd_2));
        d.rl2[t] = r > 0.0 ? (r + p.skin) * (r + p.skin) : -1.0;
    }
    d.seed = p.seed;
    d.n_part = p.n_mon * p.n_chain + p.n_mon_d * p.n_chain_d;
    d.n_mon = p.n_mon; d.n_chain = p.n_chain; d.n_mon_d = p.n_mon_d; d.n_chain_d = p.n_chain_d;
    d.pd = p.n_mon * p.n_chain;
    d.thermostat = p.thermostat; d.dpd_cut_off = p.dpd_cut_off; d.force_switch_on = p.force_switch_on;
    d.bending = p.bending; d.bending_melt = p.bending_melt; d.poiseuille = p.poiseuille;
    if (p.bend_form != 0 && p.bend_form != 1) return fail(MFA_E_ARG, "bend_form must be 0 (harmonic) or 1 (cosine)");
    d.bend_form = p.bend_form;
    d.dpd_vv = p.dpd_vv != 0;
    if (d.n_part < 1) return fail(MFA_E_ARG, "channel without beads");
    return 0;
}

// kernels are launched by name (templated kernels parenthesised)
#define LAUNCH(kern, grid, st, ...) hipLaunchKernelGGL(kern, dim3(grid), dim3(BLK), 0, st, __VA_ARGS__)

// ---------------------------------------------------------------------------
// step pieces
// ---------------------------------------------------------------------------
// Verlet-list storage for the requested capacity (nl_request: -1 automatic,
// 0 off).  The automatic capacity covers twice the mean bead density of the
// densest channel over the largest list sphere, plus a margin; beads with
// longer lists scan the cells instead (counted in flags[4]).  Lists are used
// only when some channel has a skin (otherwise every step rebuilds).
static int alloc_list(mfa_ctx *c) {
    if (c->B.nl) { (void)hipFree(c->B.nl); c->B.nl = nullptr; }
    if (c->B.nl_n) { (void)hipFree(c->B.nl_n); c->B.nl_n = nullptr; }
    c->B.nl_cap = 0;
    int cap = c->nl_request;
    if (cap < 0) {
        bool any_skin = false;
        double need = 0.0;
        for (const SysDev &d : c->hs) {
            any_skin = any_skin || d.skin_mode;
            double rl2 = 0.0;
            for (int t = 0; t < 16; ++t) rl2 = std::max(rl2, d.rl2[t]);
            const double rho = d.n_part / (d.L[0] * d.L[1] * d.L[2]);
            need = std::max(need, 2.0 * rho * 4.18879020478639 * rl2 * sqrt(rl2));
        }
        cap = any_skin ? std::min(1024, ((int)ceil(need) + 16 + 7) / 8 * 8) : 0;
    }
    if (cap == 0) return 0;
    cap = (cap + 7) / 8 * 8;      // rows of whole 16-byte groups (the pair kernels load 4 entries at once)
    if (c->S_tot >= (1 << (32 - NL_SLOT_SHIFT))) {      // entries read back as unsigned
        if (c->nl_request > 0) return fail(MFA_E_ARG, "neighbour list: too many slots for the list encoding");
        return 0;
    }
    HIPCHK(hipMalloc((void **)&c->B.nl, (size_t)cap * c->S_tot * sizeof(int)));
    HIPCHK(hipMalloc((void **)&c->B.nl_n, (size_t)c->S_tot * sizeof(int)));
    // lengths above any capacity until a bead's list is built: the pair
    // kernels then scan the cells for it
    HIPCHK(hipMemset(c->B.nl_n, 0x7f, (size_t)c->S_tot * sizeof(int)));
    HIPCHK(hipDeviceSynchronize());
    c->B.nl_cap = cap;
    return 0;
}

// keep_frc: the permute carries the total force with the beads (every
// rebuild except the fused step's, whose pair kernel overwrites the force)
static void enqueue_rebuild(mfa_ctx *c, int keep_frc) {
    Timer t(c, K_REBUILD);
    const int g = std::min(GATED_GRID, std::max(1, c->n_blocks));
    const int gc = std::min(GATED_GRID, (c->C_tot + BLK - 1) / BLK);
    const int nsb = (c->C_tot + 4 * BLK - 1) / (4 * BLK);
    LAUNCH(k_bin, g, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_cell_count, c->d_bin_cell, c->d_bin_rank,
           c->d_bin_key, c->S_tot);
    LAUNCH(k_scan1, nsb, c->st, c->B.rebuild, c->d_cell_count, c->C_tot, c->d_pre, c->d_bsum);
    LAUNCH(k_scan2, 1, c->st, c->B.rebuild, c->d_bsum, nsb);
    LAUNCH(k_scan3, gc, c->st, c->d_sys, c->B.rebuild, (const int *)c->B.chan_rb, c->d_cell_count, c->d_pre,
           c->d_bsum, c->d_cell_sys, c->d_crange, c->C_tot);
    LAUNCH(k_scatter, g, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_bin_cell, c->d_bin_rank, c->d_bin_key,
           c->d_crange, c->d_key_tmp, c->S_tot);
    LAUNCH(k_rank, g, c->st, c->d_sys, c->d_blk_sys, c->B, (const int *)c->d_bin_cell, (const int2 *)c->d_crange,
           (const int *)c->d_bin_key, (const int *)c->d_key_tmp, c->d_perm, c->S_tot);
    if (c->prec == MFA_PREC_MIXED)
        LAUNCH((k_permute<true>), g, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_perm, c->d_bin_cell, c->S_tot,
               c->d_permute_done, c->d_rebuild_count, c->n_sys, keep_frc);
    else
        LAUNCH((k_permute<false>), g, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_perm, c->d_bin_cell, c->S_tot,
               c->d_permute_done, c->d_rebuild_count, c->n_sys, keep_frc);
    if (c->B.nl_cap > 0) {
        if (c->prec == MFA_PREC_MIXED)
            LAUNCH((k_nlist<true>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, (const int2 *)c->d_crange);
        else
            LAUNCH((k_nlist<false>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, (const int2 *)c->d_crange);
    }
}

static void enqueue_vv1(mfa_ctx *c) {
    Timer t(c, K_VV1);
    if (c->prec == MFA_PREC_MIXED) LAUNCH((k_vv1<true>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B);
    else LAUNCH((k_vv1<false>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B);
}

// fused: inside md_steps (energies and the pair count on the call's last step only)
static void enqueue_pair(mfa_ctx *c, int fused) {
    Timer t(c, K_PAIR);
    bool dpd = false;
    for (auto &p : c->hp) dpd = dpd || p.thermostat == 1;
    const int nb = c->n_blocks;
    if (c->prec == MFA_PREC_MIXED) {
        // persistent launch: pair_grid blocks take the nb tiles from a counter
        const bool pers = c->pair_grid > 0 && c->pair_grid < nb;
        const int g = pers ? c->pair_grid : nb;
        int *tl = pers ? c->d_tiles : nullptr;
#define PAIR32(D, V)                                                                                     \
    do {                                                                                                 \
        if (pers) LAUNCH((k_pair32<D, V, true>), g, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_crange,     \
                         c->d_partial, fused, tl, nb);                                                   \
        else LAUNCH((k_pair32<D, V, false>), g, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_crange,         \
                    c->d_partial, fused, tl, nb);                                                        \
    } while (0)
        if (dpd) {
            if (c->virial) PAIR32(true, true); else PAIR32(true, false);
        } else {
            if (c->virial) PAIR32(false, true); else PAIR32(false, false);
        }
#undef PAIR32
    } else {
        if (dpd) {
            if (c->virial) LAUNCH((k_pair64<true, true>), nb, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_crange, c->d_partial, fused);
            else LAUNCH((k_pair64<true, false>), nb, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_crange, c->d_partial, fused);
        } else {
            if (c->virial) LAUNCH((k_pair64<false, true>), nb, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_crange, c->d_partial, fused);
            else LAUNCH((k_pair64<false, false>), nb, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_crange, c->d_partial, fused);
        }
    }
}

template <int MODE>
static void enqueue_bonded(mfa_ctx *c) {
    Timer t(c, K_BONDED);
    if (c->prec == MFA_PREC_MIXED)
        LAUNCH((k_bonded<MODE, true>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_partial, c->d_fheads,
                   (int)c->virial);
    else
        LAUNCH((k_bonded<MODE, false>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_partial, c->d_fheads,
                   (int)c->virial);
}

// DPD-VV re-evaluation after the second half kick (new_dpd_fd.f90); all:
// every DPD channel, else those with dpd_vv set; fused: inside md_steps
static void enqueue_dpd_vv(mfa_ctx *c, int all, int fused) {
    Timer t(c, K_DPDVV);
    LAUNCH(k_dpd_vv, c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, (const int2 *)c->d_crange, c->d_dfd, all);
    LAUNCH(k_dpd_vv_apply, c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, (const double4 *)c->d_dfd, all, fused);
}

static bool any_bending(const mfa_ctx *c) {
    for (const SysDev &d : c->hs)
        if (d.bending || d.bending_melt) return true;
    return false;
}

// The fused step runs with the fused integrator (k_integrate: the second
// half kick of a step and the first half of the next one) unless a channel
// re-evaluates its dissipative forces between the two (DPD-VV).
static bool fused_integrator(const mfa_ctx *c) {
    for (const SysDev &d : c->hs)
        if (d.dpd_vv && d.thermostat == 1) return false;
    return true;
}

// One fused step.  With the fused integrator: [gated rebuild], pair forces
// with the walls and bonds in the pair kernel's epilogue, [bending],
// k_integrate (which also starts the next step, so md_steps enqueues k_vv1
// once, before its first step).  Otherwise: k_vv1, [gated rebuild], pair,
// k_bonded (walls, bonds, bending, body force, second kick), DPD-VV.
static void enqueue_step(mfa_ctx *c) {
    if (fused_integrator(c)) {
        enqueue_rebuild(c, 0);
        enqueue_pair(c, 2);
        Timer t(c, K_BONDED);
        if (any_bending(c)) {
            if (c->prec == MFA_PREC_MIXED)
                LAUNCH((k_bonded<M_BENDF, true>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_partial, c->d_fheads,
                   (int)c->virial);
            else
                LAUNCH((k_bonded<M_BENDF, false>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_partial, c->d_fheads,
                   (int)c->virial);
        }
        if (c->prec == MFA_PREC_MIXED)
            LAUNCH((k_integrate<true>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_partial, c->d_fheads);
        else
            LAUNCH((k_integrate<false>), c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->d_partial, c->d_fheads);
        return;
    }
    enqueue_vv1(c);
    enqueue_rebuild(c, 0);
    enqueue_pair(c, 1);
    if (any_bending(c)) enqueue_bonded<M_FUSED>(c);
    else enqueue_bonded<M_FUSED | M_NOBEND>(c);
    enqueue_dpd_vv(c, 0, 1);
}

// Reads the device guard word (one sync): a nonzero code means a kernel met
// an out-of-range index and skipped the access.
static const char *guard_name(int g) {
    switch (g) {
        case G_CELL: return "cell id"; case G_RANGE: return "cell slot range"; case G_PERM: return "permutation";
        case G_ORIG: return "bead index"; case G_WHERE: return "bead slot"; case G_TYPE: return "bead type";
        case G_HEAD: return "head index"; case G_BIN: return "binned cell"; case G_SLOT: return "neighbour range";
        default: return "unknown";
    }
}

static int check_guard(mfa_ctx *c) {
    int f[2] = {0, 0};
    if (int rc = to_host(c, f, c->B.flags + 6, 2 * sizeof(int), "guard flags")) return rc;
    if (f[1] != 0)
        return fail(MFA_E_STATE, std::string("device bounds guard: bad ") + guard_name(f[1]) + " at index " +
                                     std::to_string(f[0]) + " (state is corrupt; the access was skipped)");
    return 0;
}

static int check_ctx(mfa_ctx *c, bool need_state = true) {
    if (!c) return fail(MFA_E_ARG, "null context");
    if (need_state && !c->have_state) return fail(MFA_E_STATE, "no state uploaded (mfa_hip_set_state)");
    return 0;
}

static void invalidate_graph(mfa_ctx *c) {
    if (c->gexec) { (void)hipGraphExecDestroy(c->gexec); c->gexec = nullptr; }
    if (c->graph) { (void)hipGraphDestroy(c->graph); c->graph = nullptr; }
    if (c->gexec_k) { (void)hipGraphExecDestroy(c->gexec_k); c->gexec_k = nullptr; }
    if (c->graph_k) { (void)hipGraphDestroy(c->graph_k); c->graph_k = nullptr; }
}

// Captures `steps` consecutive steps into one executable graph.  Every step
// reads its time index and rebuild gates from device memory, so identical
// captured steps advance the state exactly as separate launches do.
static int capture_steps(mfa_ctx *c, int steps, hipGraph_t *g, hipGraphExec_t *ge) {
    HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
    for (int i = 0; i < steps; ++i) enqueue_step(c);
    hipError_t e = hipStreamEndCapture(c->st, g);
    if (e != hipSuccess) { invalidate_graph(c); return fail(MFA_E_HIP, std::string("capture: ") + hipGetErrorString(e)); }
    e = hipGraphInstantiate(ge, *g, nullptr, nullptr, 0);
    if (e != hipSuccess) { invalidate_graph(c); return fail(MFA_E_HIP, std::string("instantiate: ") + hipGetErrorString(e)); }
    return 0;
}

// ===========================================================================
// C ABI
// ===========================================================================
// ---------------------------------------------------------------------------
// Allocation-free check_fluid_fluid (mfa_hip_check_fluid_fluid_noalloc).  The
// per-bead energies live in `scratch`, a file-scope device array each GPU
// holds for the life of the process; host code reaches it only through the
// address hipGetSymbolAddress returns for the current device.
// ---------------------------------------------------------------------------
namespace {

constexpr int NOALLOC_MAX_PART = MFA_NOALLOC_MAX_PART;
constexpr int MAX_DEVICES = 64;

static __device__ __align__(256) double scratch[NOALLOC_MAX_PART];

struct BruteTables {
    double L0, L1, iL0, iL1;
    double epsil[16], sigma_2[16], range_2[16], e_shift[16];
};

// Per-bead part of util.f90:284-330 over caller arrays in the reference
// layout: e_i = sum over j > i of the LJ energy, j ascending, with the same
// expressions as k_brute, both without FP contraction (so the energies are
// the same bits, and the reference's x86-64 arithmetic).
__global__ __launch_bounds__(BLK) void k_brute_flat(BruteTables T, const double *__restrict__ r0,
                                                   const int *__restrict__ a_type, int n, double r_2_min,
                                                   double *__restrict__ e_out) {
#pragma clang fp contract(off)
    __shared__ double4 tile[BLK];
    const int i = blockIdx.x * BLK + threadIdx.x;
    const bool act = i < n;
    double4 ri = make_double4(0.0, 0.0, 0.0, 0.0);
    int ti = 1;
    if (act) {
        ri = make_double4(r0[3 * (size_t)i], r0[3 * (size_t)i + 1], r0[3 * (size_t)i + 2], 0.0);
        ti = a_type[i];
    }
    double e = 0.0;
    for (int j0 = 0; j0 < n; j0 += BLK) {
        const int jj = j0 + threadIdx.x;
        tile[threadIdx.x] = jj < n ? make_double4(r0[3 * (size_t)jj], r0[3 * (size_t)jj + 1], r0[3 * (size_t)jj + 2],
                                                  __longlong_as_double((long long)a_type[jj]))
                                   : make_double4(0.0, 0.0, 0.0, 0.0);
        __syncthreads();
        const int jn = min(BLK, n - j0);
        if (act) {
            for (int q = 0; q < jn; ++q) {
                const int j = j0 + q;
                if (j <= i) continue;
                const double4 rj = tile[q];
                const int tj = (int)__double_as_longlong(rj.w);
                const int tt = ((ti - 1) & 3) + 4 * ((tj - 1) & 3);
                const double dx = mic(ri.x - rj.x, T.L0, T.iL0);
                const double dy = mic(ri.y - rj.y, T.L1, T.iL1);
                const double dz = ri.z - rj.z;
                double r_2 = r2_exact(dx, dy, dz);
                if (r_2 < T.range_2[tt]) {
                    if (r_2 < r_2_min) r_2 = r_2_min;
                    const double qd = T.sigma_2[tt] / r_2;
                    const double r_6 = qd * qd * qd;
                    const double r_12 = r_6 * r_6;
                    e += T.epsil[tt] * ((r_12 - r_6) - T.e_shift[tt]);
                }
            }
        }
        __syncthreads();
    }
    if (act) e_out[i] = e;
}

// v = sum of e[0..n-1] in index order (one thread: the order of the host sum
// of mfa_hip_check_fluid_fluid, so the result is the same bits).
__global__ void k_sum_in_order(const double *__restrict__ e, int n, double *__restrict__ v) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += e[i];
    *v = s;
}

std::mutex g_scratch_mu;
double *g_scratch_ptr[MAX_DEVICES] = {};

// Device address of this device's `scratch`, looked up once per device with
// that device current.
int scratch_for_current_device(double **out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail(MFA_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= MAX_DEVICES) return fail(MFA_E_ARG, "device ordinal beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MFA_E_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace

extern "C" {

const char *mfa_hip_last_error(void) { return g_err.c_str(); }
int mfa_hip_version(void) { return 1; }

int mfa_hip_make_binning(int32_t bin_case, const double *boundary, double r_cut, double *l_cell, double *inv_l_cell,
                         int32_t *n_cells, int32_t *n_cells_tot, int32_t *n_nei) {
    if (!boundary || !l_cell || !inv_l_cell || !n_cells || !n_cells_tot || !n_nei || !(r_cut > 0.0))
        return fail(MFA_E_ARG, "make_binning: bad argument");
    if (bin_case != 1 && bin_case != 2) return fail(MFA_E_ARG, "make_binning: bin_case must be 1 or 2");
    make_binning(bin_case, boundary, r_cut, l_cell, inv_l_cell, n_cells, n_cells_tot, n_nei);
    return 0;
}

int mfa_hip_create(const mfa_system_params *sys, int32_t n_sys, int32_t precision, mfa_ctx **out) {
    if (!sys || n_sys < 1 || !out) return fail(MFA_E_ARG, "create: bad argument");
    if (precision != MFA_PREC_FP64 && precision != MFA_PREC_MIXED) return fail(MFA_E_ARG, "create: bad precision");
    *out = nullptr;
    int ndev = 0;
    hipError_t de = hipGetDeviceCount(&ndev);
    if (de != hipSuccess || ndev < 1) {
        const char *hv = getenv("HIP_VISIBLE_DEVICES"), *rv = getenv("ROCR_VISIBLE_DEVICES");
        return fail(MFA_E_HIP, std::string("no HIP device: hipGetDeviceCount -> ") + hipGetErrorString(de) +
                                   ", count " + std::to_string(ndev) + ", HIP_VISIBLE_DEVICES=" +
                                   (hv ? (*hv ? hv : "<empty>") : "<unset>") + ", ROCR_VISIBLE_DEVICES=" +
                                   (rv ? (*rv ? rv : "<empty>") : "<unset>"));
    }
    mfa_ctx *c = new mfa_ctx();
    c->n_sys = n_sys;
    c->prec = precision;
    c->hp.assign(sys, sys + n_sys);
    c->hs.resize(n_sys);
    c->uploaded.assign(n_sys, 0);
    long long slots = 0, parts = 0, cells = 0, heads = 0;
    for (int s = 0; s < n_sys; ++s) {
        int rc = build_sysdev(sys[s], c->hs[s]);
        if (rc) { delete c; return rc; }
        SysDev &d = c->hs[s];
        d.slot_base = (int)slots;
        d.part_base = (int)parts;
        d.cell_base = (int)cells;
        d.head_base = (int)heads;
        d.block_base = (int)(slots / BLK);
        d.n_blocks = (d.n_part + BLK - 1) / BLK;
        slots += (long long)d.n_blocks * BLK;
        parts += d.n_part;
        cells += d.n_cells_tot;
        heads += d.n_chain;
    }
    if (slots >= (1ll << 28) || cells >= (1ll << 30)) { delete c; return fail(MFA_E_ARG, "too many beads or cells"); }
    c->S_tot = (int)slots; c->N_tot = (int)parts; c->C_tot = (int)cells; c->H_tot = (int)heads;
    c->n_blocks = c->S_tot / BLK;
    c->B.n_slots = c->S_tot;
    c->B.dpd_queue = DPD_Q;
    c->B.n_beads = c->N_tot;
    c->B.n_cells = c->C_tot;
    c->B.n_heads = std::max(1, c->H_tot);
    std::vector<int> blk_sys(c->n_blocks), cell_sys(c->C_tot);
    for (int s = 0; s < n_sys; ++s) {
        for (int b = 0; b < c->hs[s].n_blocks; ++b) blk_sys[c->hs[s].block_base + b] = s;
        for (int q = 0; q < c->hs[s].n_cells_tot; ++q) cell_sys[c->hs[s].cell_base + q] = s;
    }
    int rc = 0;
#define ALLOC(p, n) do { rc = dalloc(&(p), (n)); if (rc) { free_all(c); delete c; return rc; } } while (0)
    if (hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking) != hipSuccess) {
        delete c; return fail(MFA_E_HIP, "stream creation failed");
    }
    if (hipHostMalloc((void **)&c->h_stage, STAGE_BYTES, hipHostMallocDefault) != hipSuccess) {
        c->h_stage = nullptr; free_all(c); delete c; return fail(MFA_E_HIP, "pinned staging allocation failed");
    }
    // two halves of S_tot slots each (current / next, selected by parity)
    // pos and posf carry NL_UNROLL - 1 zeroed entries past the upper half: the
    // list build reads whole groups of four partners past a cell's end
    ALLOC(c->B.pos, 2 * (size_t)c->S_tot + NL_UNROLL); ALLOC(c->B.vel, 2 * (size_t)c->S_tot);
    ALLOC(c->B.acc, 2 * (size_t)c->S_tot);
    ALLOC(c->B.posb, 2 * (size_t)c->S_tot); ALLOC(c->B.cell_of, 2 * (size_t)c->S_tot);
    if (precision == MFA_PREC_MIXED) {
        ALLOC(c->B.posf, 2 * (size_t)c->S_tot + NL_UNROLL); ALLOC(c->B.velf, 2 * (size_t)c->S_tot);
    }
    ALLOC(c->B.frc, 2 * (size_t)c->S_tot); ALLOC(c->B.where, c->N_tot); ALLOC(c->B.parity, n_sys); ALLOC(c->B.rebuild, 1);
    ALLOC(c->B.chan_rb, n_sys);
    ALLOC(c->B.flags, 8); ALLOC(c->B.t_idx, 1); ALLOC(c->B.t_end, 1); ALLOC(c->B.store_frc, 1);
    ALLOC(c->d_sys, n_sys); ALLOC(c->d_blk_sys, c->n_blocks); ALLOC(c->d_cell_sys, c->C_tot);
    ALLOC(c->d_cell_count, c->C_tot); ALLOC(c->d_pre, c->C_tot);
    ALLOC(c->d_bsum, (c->C_tot + 4 * BLK - 1) / (4 * BLK) + 1);
    ALLOC(c->d_crange, c->C_tot);
    ALLOC(c->d_bin_cell, c->S_tot); ALLOC(c->d_bin_rank, c->S_tot); ALLOC(c->d_perm, c->S_tot);
    ALLOC(c->d_key_tmp, c->S_tot); ALLOC(c->d_bin_key, c->S_tot);
    ALLOC(c->d_partial, (size_t)c->n_blocks * NPART); ALLOC(c->d_energies, (size_t)n_sys * NPART);
    ALLOC(c->d_fheads, 3 * (size_t)std::max(1, c->H_tot)); ALLOC(c->d_rebuild_count, 1); ALLOC(c->d_skin_fired, 1); ALLOC(c->d_permute_done, 1);
    {
        int mb = 1;
        for (const SysDev &d : c->hs) mb = std::max(mb, (d.n_part + BLK - 1) / BLK + 1);
        ALLOC(c->d_obs, 16 * (size_t)mb);
    }
    ALLOC(c->d_stage, 3 * (size_t)c->N_tot); ALLOC(c->d_istage, c->N_tot);
    ALLOC(c->d_dfd, c->S_tot);
    ALLOC(c->d_tiles, 2);
#undef ALLOC
    if ((rc = alloc_list(c))) { free_all(c); delete c; return rc; }
    // the zero fills of dalloc ran on the null stream: finish them before
    // anything is enqueued on the context's non-blocking stream
    if (hipDeviceSynchronize() != hipSuccess) { free_all(c); delete c; return fail(MFA_E_HIP, "allocation sync failed"); }
    long long t0 = sys[0].t_idx0;
    // until an md_steps call sets t_end, every fused step stores the force
    const long long t_end0 = LLONG_MIN;
    const int store1 = 1;
    if ((rc = to_device(c, c->d_sys, c->hs.data(), n_sys * sizeof(SysDev), "channel parameters")) ||
        (rc = to_device(c, c->B.t_end, &t_end0, sizeof(long long), "last-step index")) ||
        (rc = to_device(c, c->B.store_frc, &store1, sizeof(int), "force-store flag")) ||
        (rc = to_device(c, c->d_blk_sys, blk_sys.data(), c->n_blocks * sizeof(int), "block table")) ||
        (rc = to_device(c, c->d_cell_sys, cell_sys.data(), c->C_tot * sizeof(int), "cell table")) ||
        (rc = to_device(c, c->B.t_idx, &t0, sizeof(long long), "time index"))) {
        free_all(c); delete c; return rc;
    }
    *out = c;
    return 0;
}

int mfa_hip_destroy(mfa_ctx *c) {
    if (!c) return 0;
    (void)hipStreamSynchronize(c->st);
    free_all(c);
    delete c;
    return 0;
}

int mfa_hip_set_state(mfa_ctx *c, int32_t isys, const double *r0, const double *v, const double *a,
                      const int32_t *a_type) {
    if (int rc = check_ctx(c, false)) return rc;
    if (isys < 0 || isys >= c->n_sys || !r0 || !v || !a_type) return fail(MFA_E_ARG, "set_state: bad argument");
    const SysDev &d = c->hs[isys];
    const size_t n = d.n_part;
    for (size_t i = 0; i < n; ++i)
        if (a_type[i] < 1 || a_type[i] > 4) return fail(MFA_E_ARG, "set_state: a_type outside 1..4");
    // r0, v, a, a_type go through d_stage (3 n doubles) one array at a time
    double *dr = nullptr, *dv = nullptr, *da = nullptr;
    HIPCHK(hipMalloc((void **)&dr, 3 * n * sizeof(double)));
    HIPCHK(hipMalloc((void **)&dv, 3 * n * sizeof(double)));
    if (a) HIPCHK(hipMalloc((void **)&da, 3 * n * sizeof(double)));
    int rc = to_device(c, dr, r0, 3 * n * sizeof(double), "r0");
    if (!rc) rc = to_device(c, dv, v, 3 * n * sizeof(double), "v");
    if (!rc && a) rc = to_device(c, da, a, 3 * n * sizeof(double), "a");
    if (!rc) rc = to_device(c, c->d_istage, a_type, n * sizeof(int), "a_type");
    if (!rc) {
        LAUNCH(k_upload, (int)((n + BLK - 1) / BLK), c->st, c->d_sys, isys, c->B, (const double *)dr, (const double *)dv,
               (const double *)da, (const int *)c->d_istage);
        hipLaunchKernelGGL(k_mark_rebuild, dim3(1), dim3(1), 0, c->st, c->B, (int)isys);
        rc = sync_for(c, "state upload");
    }
    (void)hipFree(dr); (void)hipFree(dv); if (da) (void)hipFree(da);
    if (rc) return rc;
    if (int rg = check_guard(c)) return rg;
    c->uploaded[isys] = 1;
    c->have_state = std::all_of(c->uploaded.begin(), c->uploaded.end(), [](char u) { return u != 0; });
    // cell list of the new state (init_config.f90:261-272); pending until every
    // channel has a state, since the rebuild runs over all of them
    return 0;
}

static int ensure_cells(mfa_ctx *c) {
    // A rebuild is pending after set_state; run it (it is a no-op otherwise).
    enqueue_rebuild(c, 1);
    hipLaunchKernelGGL(k_clear_rebuild, dim3(1), dim3(BLK), 0, c->st, c->B, c->n_sys, (int *)nullptr);
    return 0;
}

int mfa_hip_get_state(mfa_ctx *c, int32_t isys, double *r0, double *v, double *a, double *force) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys) return fail(MFA_E_ARG, "get_state: bad channel");
    const SysDev &d = c->hs[isys];
    const size_t n = d.n_part;
    if (int rc = check_guard(c)) return rc;
    double *outs[4] = {r0, v, a, force};
    for (int w = 0; w < 4; ++w) {
        if (!outs[w]) continue;
        const int gd = (int)((n + BLK - 1) / BLK);
        if (w == 0) LAUNCH((k_download<0>), gd, c->st, c->d_sys, isys, c->B, c->d_stage);
        else if (w == 1) LAUNCH((k_download<1>), gd, c->st, c->d_sys, isys, c->B, c->d_stage);
        else if (w == 2) LAUNCH((k_download<2>), gd, c->st, c->d_sys, isys, c->B, c->d_stage);
        else LAUNCH((k_download<3>), gd, c->st, c->d_sys, isys, c->B, c->d_stage);
        HIPCHK(hipGetLastError());
        static const char *names[4] = {"r0", "v", "a", "force"};
        if (int rc = to_host(c, outs[w], c->d_stage, 3 * n * sizeof(double), names[w])) return rc;
    }
    return check_guard(c);
}

int mfa_hip_set_time(mfa_ctx *c, int64_t t_idx) {
    if (int rc = check_ctx(c, false)) return rc;
    long long t = t_idx;
    return to_device(c, c->B.t_idx, &t, sizeof(long long), "time index");
}

int mfa_hip_get_time(mfa_ctx *c, int64_t *t_idx) {
    if (int rc = check_ctx(c, false)) return rc;
    if (!t_idx) return fail(MFA_E_ARG, "get_time: null");
    long long t = 0;
    if (int rc = to_host(c, &t, c->B.t_idx, sizeof(long long), "time index")) return rc;
    *t_idx = t;
    return 0;
}

int mfa_hip_verlet_positions(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    ensure_cells(c);
    enqueue_vv1(c);
    enqueue_rebuild(c, 1);
    hipLaunchKernelGGL(k_clear_rebuild, dim3(1), dim3(BLK), 0, c->st, c->B, c->n_sys, c->d_skin_fired);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_check_skin(mfa_ctx *c, int32_t *f_skin) {
    if (int rc = check_ctx(c)) return rc;
    if (!f_skin) return fail(MFA_E_ARG, "check_skin: null");
    // f_skin of the last verlet_positions: 1 when its skin test fired (some
    // bead more than skin/2 from its position at the last rebuild), else 0.
    // verlet_positions has already rebuilt the cells and lists by then, which
    // is the binning the reference calls next (md_main.f90:97-102).
    int v = 0;
    if (int rc = to_host(c, &v, c->d_skin_fired, sizeof(int), "skin test flag")) return rc;
    *f_skin = v;
    return 0;
}

int mfa_hip_fluid_fluid(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    enqueue_rebuild(c, 1);    // pending rebuild after set_state / verlet_positions (gated)
    if (c->prec == MFA_PREC_MIXED) LAUNCH(k_velf_sync, c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B);
    enqueue_pair(c, 0);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_fluid_wall(mfa_ctx *c, int32_t inter_type) {
    if (int rc = check_ctx(c)) return rc;
    if (inter_type != 2) return fail(MFA_E_ARG, "fluid_wall: only inter_type 2 (implicit 9-3 walls) is supported");
    enqueue_bonded<M_WALL>(c);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_wall_wall(mfa_ctx *c, int32_t inter_type, double sigma_w, double *v_wall_wall) {
    if (int rc = check_ctx(c, false)) return rc;
    if (!v_wall_wall) return fail(MFA_E_ARG, "wall_wall: null");
    if (inter_type != 2) return fail(MFA_E_ARG, "wall_wall: only inter_type 2 is supported");
    const mfa_system_params &p = c->hp[0];
    const double q = sigma_w / p.boundary[2];
    *v_wall_wall = 2.0 * (fabs(p.a_w) * pow(q, 9) - p.a_w * pow(q, 3));
    return 0;
}

int mfa_hip_intra_molec(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    enqueue_bonded<M_FENE>(c);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_bending(mfa_ctx *c, int32_t mode) {
    if (int rc = check_ctx(c)) return rc;
    if (mode == 0) return 0;
    if (mode != 1) return fail(MFA_E_ARG, "bending: mode must be 0 or 1");
    enqueue_bonded<M_BEND>(c);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_bending_melt(mfa_ctx *c, int32_t mode) {
    if (int rc = check_ctx(c)) return rc;
    if (mode == 0) return 0;
    if (mode != 1) return fail(MFA_E_ARG, "bending_melt: mode must be 0 or 1");
    enqueue_bonded<M_BENDM>(c);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_constant_force(mfa_ctx *c, int32_t case_pois) {
    if (int rc = check_ctx(c)) return rc;
    if (case_pois != 1) return fail(MFA_E_ARG, "constant_force: only case 1 (Poiseuille body force) is supported");
    enqueue_bonded<M_CONST>(c);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_verlet_velocities(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    enqueue_bonded<M_VV2>(c);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_new_dpd_fd(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    enqueue_dpd_vv(c, 1, 0);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_md_steps(mfa_ctx *c, int32_t n_steps, int32_t use_graph) {
    if (int rc = check_ctx(c)) return rc;
    if (n_steps < 0) return fail(MFA_E_ARG, "md_steps: negative step count");
    ensure_cells(c);
    if (n_steps > 0) {
        hipLaunchKernelGGL(k_set_t_end, dim3(1), dim3(1), 0, c->st, c->B, (int)n_steps);
        if (fused_integrator(c)) enqueue_vv1(c);      // the first step's first half
    }
    if (!use_graph || c->timing_mask) {
        for (int i = 0; i < n_steps; ++i) enqueue_step(c);
        HIPCHK(hipGetLastError());
        return 0;
    }
    int i = 0;
    if (n_steps >= GRAPH_STEPS) {
        if (!c->gexec_k)
            if (int rc = capture_steps(c, GRAPH_STEPS, &c->graph_k, &c->gexec_k)) return rc;
        for (; i + GRAPH_STEPS <= n_steps; i += GRAPH_STEPS) HIPCHK(hipGraphLaunch(c->gexec_k, c->st));
    }
    if (i < n_steps && !c->gexec)
        if (int rc = capture_steps(c, 1, &c->graph, &c->gexec)) return rc;
    for (; i < n_steps; ++i) HIPCHK(hipGraphLaunch(c->gexec, c->st));
    return 0;
}

int mfa_hip_prepare_graphs(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    if (c->timing_mask) return 0;       // timed steps launch eagerly
    HIPCHK(hipStreamSynchronize(c->st));
    if (!c->gexec_k)
        if (int rc = capture_steps(c, GRAPH_STEPS, &c->graph_k, &c->gexec_k)) return rc;
    if (!c->gexec)
        if (int rc = capture_steps(c, 1, &c->graph, &c->gexec)) return rc;
    return 0;
}

int mfa_hip_synchronize(mfa_ctx *c) {
    if (int rc = check_ctx(c, false)) return rc;
    HIPCHK(hipStreamSynchronize(c->st));
    return check_guard(c);
}

int mfa_hip_get_energies(mfa_ctx *c, int32_t isys, mfa_energies *out) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !out) return fail(MFA_E_ARG, "get_energies: bad argument");
    hipLaunchKernelGGL(k_reduce, dim3(c->n_sys), dim3(BLK), 0, c->st, c->d_sys, (const double *)c->d_partial,
                       c->d_energies);
    HIPCHK(hipGetLastError());
    double e[NPART];
    if (int rc = to_host(c, e, c->d_energies + (size_t)isys * NPART, NPART * sizeof(double), "energies")) return rc;
    if (int rc = check_guard(c)) return rc;
    const mfa_system_params &p = c->hp[isys];
    out->v_fluid_fluid = 0.5 * e[C_VFF];
    out->v_fluid_wall = e[C_VWALL];
    out->v_intra_molec = p.k_chain * e[C_VINTRA];
    out->v_bend = e[C_VBEND];
    out->v_bend_melt = e[C_VBENDM];
    out->t_fluid = 0.5 * e[C_KE];
    for (int q = 0; q < 6; ++q) {
        out->virial_bond[q] = 0.5 * e[C_WBXX + q];
        out->virial[q] = 0.5 * (e[C_WXX + q] + e[C_WBXX + q]);
    }
    return 0;
}

int mfa_hip_get_pair_count(mfa_ctx *c, int32_t isys, int64_t *n_pairs) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !n_pairs) return fail(MFA_E_ARG, "get_pair_count: bad argument");
    hipLaunchKernelGGL(k_reduce, dim3(c->n_sys), dim3(BLK), 0, c->st, c->d_sys, (const double *)c->d_partial,
                       c->d_energies);
    double e[NPART];
    if (int rc = to_host(c, e, c->d_energies + (size_t)isys * NPART, NPART * sizeof(double), "pair count")) return rc;
    *n_pairs = (int64_t)llround(0.5 * e[C_NPAIR]);
    return 0;
}

int mfa_hip_get_f_on_heads(mfa_ctx *c, int32_t isys, double *f) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !f) return fail(MFA_E_ARG, "get_f_on_heads: bad argument");
    const SysDev &d = c->hs[isys];
    if (d.n_chain == 0) return 0;
    return to_host(c, f, c->d_fheads + 3 * (size_t)d.head_base, 3 * (size_t)d.n_chain * sizeof(double), "f_on_heads");
}

int mfa_hip_list_stats(mfa_ctx *c, int32_t isys, int64_t *out) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !out) return fail(MFA_E_ARG, "list_stats: bad argument");
    if (c->B.nl_cap <= 0) return fail(MFA_E_STATE, "list_stats: no Verlet list (capacity 0)");
    ensure_cells(c);                  // the rebuild pending after set_state builds the lists
    const SysDev &d = c->hs[isys];
    const int gm = std::max(1, (d.n_part + BLK - 1) / BLK);   // 3 gm partials fit in d_obs (16 per block)
    LAUNCH(k_list_stats, gm, c->st, c->d_sys, isys, c->B, c->d_obs);
    std::vector<double> h(3 * (size_t)gm);
    if (int rc = to_host(c, h.data(), c->d_obs, h.size() * sizeof(double), "list statistics")) return rc;
    double t[3] = {0.0, 0.0, 0.0};
    for (int b = 0; b < gm; ++b)
        for (int q = 0; q < 3; ++q) t[q] += h[3 * (size_t)b + q];
    for (int q = 0; q < 3; ++q) out[q] = (int64_t)t[q];
    return 0;
}

int mfa_hip_chain_observables(mfa_ctx *c, int32_t isys, double *out) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !out) return fail(MFA_E_ARG, "chain_observables: bad argument");
    const SysDev &d = c->hs[isys];
    const int nt = d.n_chain + d.n_chain_d;
    const int gb = std::max(1, (nt + BLK - 1) / BLK), gm = std::max(1, (d.n_part + BLK - 1) / BLK);
    // both partial sets fit: gb, gm <= blocks of the channel's slots
    LAUNCH(k_chain_rg2, gb, c->st, c->d_sys, isys, c->B, c->d_obs);
    LAUNCH(k_momentum, gm, c->st, c->d_sys, isys, c->B, c->d_obs + 8 * (size_t)gb);
    std::vector<double> h(8 * (size_t)(gb + gm));
    if (int rc = to_host(c, h.data(), c->d_obs, h.size() * sizeof(double), "chain observables")) return rc;
    double rb[3] = {0, 0, 0}, rm[3] = {0, 0, 0}, pm[3] = {0, 0, 0};
    for (int b = 0; b < gb; ++b)
        for (int q = 0; q < 3; ++q) { rb[q] += h[8 * (size_t)b + q]; rm[q] += h[8 * (size_t)b + 3 + q]; }
    for (int b = 0; b < gm; ++b)
        for (int q = 0; q < 3; ++q) pm[q] += h[8 * (size_t)(gb + b) + q];
    // observation.f90:256,272: mean over chains (brush times 1/n_chain, melt divided by n_chain_d)
    const double inv_nc = d.n_chain > 0 ? 1.0 / (double)d.n_chain : 0.0;
    for (int q = 0; q < 3; ++q) {
        out[q] = rb[q] * inv_nc;
        out[3 + q] = d.n_chain_d > 0 ? rm[q] / (double)d.n_chain_d : 0.0;
        out[6 + q] = pm[q];
    }
    return 0;
}

int mfa_hip_get_flags(mfa_ctx *c, mfa_flags *out) {
    if (int rc = check_ctx(c, false)) return rc;
    if (!out) return fail(MFA_E_ARG, "get_flags: null");
    int f[8], rb = 0;
    if (int rc = to_host(c, f, c->B.flags, 8 * sizeof(int), "flags")) return rc;
    if (int rc = to_host(c, &rb, c->d_rebuild_count, sizeof(int), "rebuild count")) return rc;
    out->nonfinite = f[0];
    out->out_of_box_z = f[1];
    out->fene_stretched = f[2];
    out->list_overflow = f[4];
    out->rebuilds = rb;
    return 0;
}

static int brute(mfa_ctx *c, int isys, double *e_sum, int32_t *count, uint64_t *hash) {
    const SysDev &d = c->hs[isys];
    const size_t n = d.n_part;
    double *de = nullptr; int *dc = nullptr; unsigned long long *dh = nullptr;
    HIPCHK(hipMalloc((void **)&de, n * sizeof(double)));
    HIPCHK(hipMalloc((void **)&dc, n * sizeof(int)));
    HIPCHK(hipMalloc((void **)&dh, n * sizeof(unsigned long long)));
    long long t = 0;
    if (int rc = to_host(c, &t, c->B.t_idx, sizeof(long long), "time index")) return rc;
    const mfa_system_params &p = c->hp[isys];
    double r_2_min = 0.0;
    if (p.force_switch_on) {
        double r_time = (double)t * p.dt;
        if (r_time < p.r_2_min_time) r_2_min = (1.0 - r_time / p.r_2_min_time) * p.r_2_min_init;
    }
    LAUNCH(k_brute, (int)((n + BLK - 1) / BLK), c->st, c->d_sys, isys, c->B, r_2_min, de, dc, dh);
    HIPCHK(hipGetLastError());
    std::vector<double> he(n);
    int rc = to_host(c, he.data(), de, n * sizeof(double), "brute-force energies");
    if (!rc && count) rc = to_host(c, count, dc, n * sizeof(int), "pair counts");
    if (!rc && hash) rc = to_host(c, hash, dh, n * sizeof(unsigned long long), "pair hashes");
    if (rc) { (void)hipFree(de); (void)hipFree(dc); (void)hipFree(dh); return rc; }
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s += he[i];
    if (e_sum) *e_sum = s;
    (void)hipFree(de); (void)hipFree(dc); (void)hipFree(dh);
    return check_guard(c);
}

int mfa_hip_check_fluid_fluid(mfa_ctx *c, int32_t isys, double *v) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !v) return fail(MFA_E_ARG, "check_fluid_fluid: bad argument");
    return brute(c, isys, v, nullptr, nullptr);
}

int mfa_hip_check_fluid_fluid_noalloc(const mfa_system_params *sys, const double *d_r0, const int32_t *d_a_type,
                                      int32_t n_part, double r_2_min, double *d_v_ff, void *stream) {
    if (!sys || !d_r0 || !d_a_type || !d_v_ff) return fail(MFA_E_ARG, "check_fluid_fluid_noalloc: null argument");
    if (n_part < 1) return fail(MFA_E_ARG, "check_fluid_fluid_noalloc: n_part must be >= 1");
    if (n_part > NOALLOC_MAX_PART)
        return fail(MFA_E_ARG, "check_fluid_fluid_noalloc: n_part " + std::to_string(n_part) +
                                   " exceeds the scratch capacity of " + std::to_string(NOALLOC_MAX_PART) + " beads");
    if (!(sys->boundary[0] > 0.0) || !(sys->boundary[1] > 0.0))
        return fail(MFA_E_ARG, "check_fluid_fluid_noalloc: boundary must be positive");
    double *scr = nullptr;
    if (int rc = scratch_for_current_device(&scr)) return rc;
    BruteTables T;
    T.L0 = sys->boundary[0]; T.L1 = sys->boundary[1];
    T.iL0 = 1.0 / sys->boundary[0]; T.iL1 = 1.0 / sys->boundary[1];
    for (int t = 0; t < 16; ++t) {
        T.epsil[t] = sys->epsil[t]; T.sigma_2[t] = sys->sigma_2[t];
        T.range_2[t] = sys->range_2[t]; T.e_shift[t] = sys->e_shift[t];
    }
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_brute_flat, dim3((n_part + BLK - 1) / BLK), dim3(BLK), 0, st, T, d_r0, (const int *)d_a_type,
                       (int)n_part, r_2_min, scr);
    hipLaunchKernelGGL(k_sum_in_order, dim3(1), dim3(64), 0, st, (const double *)scr, (int)n_part, d_v_ff);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_pair_stats(mfa_ctx *c, int32_t isys, int32_t *count, uint64_t *hash) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys || !count || !hash) return fail(MFA_E_ARG, "pair_stats: bad argument");
    return brute(c, isys, nullptr, count, hash);
}

int mfa_hip_get_cells(mfa_ctx *c, int32_t isys, int32_t *n_cells, int32_t *cell_start, int32_t *cell_count,
                      int32_t *cell_beads) {
    if (int rc = check_ctx(c)) return rc;
    if (isys < 0 || isys >= c->n_sys) return fail(MFA_E_ARG, "get_cells: bad channel");
    ensure_cells(c);
    const SysDev &d = c->hs[isys];
    if (n_cells) for (int k = 0; k < 3; ++k) n_cells[k] = d.nc[k];
    std::vector<int2> rg(d.n_cells_tot);
    if (int rc = to_host(c, rg.data(), c->d_crange + d.cell_base, d.n_cells_tot * sizeof(int2), "cell ranges")) return rc;
    int par = 0;
    if (int rc = to_host(c, &par, c->B.parity + isys, sizeof(int), "parity")) return rc;
    std::vector<double4> pos(d.n_part);
    const double4 *cur = c->B.pos + (size_t)(par & 1) * (size_t)c->B.n_slots;
    if (int rc = to_host(c, pos.data(), cur + d.slot_base, d.n_part * sizeof(double4), "positions")) return rc;
    for (int q = 0; q < d.n_cells_tot; ++q) {
        if (cell_start) cell_start[q] = rg[q].x - d.slot_base;
        if (cell_count) cell_count[q] = rg[q].y - rg[q].x;
    }
    if (cell_beads)
        for (int i = 0; i < d.n_part; ++i) {
            uint64_t w;
            memcpy(&w, &pos[i].w, 8);
            cell_beads[i] = (int)(w >> 3) - d.part_base;
        }
    return 0;
}

int mfa_hip_profiles_reset(mfa_ctx *c, int32_t n_bins) {
    if (int rc = check_ctx(c, false)) return rc;
    if (n_bins < 1 || n_bins > 100000) return fail(MFA_E_ARG, "profiles_reset: n_bins out of range");
    HIPCHK(hipStreamSynchronize(c->st));
    auto F = [](void *p) { if (p) (void)hipFree(p); };
    F(c->d_pcnt); F(c->d_pvs); F(c->d_pv2s); F(c->d_pdens); F(c->d_pvp); F(c->d_pvp2);
    c->d_pcnt = nullptr; c->d_pvs = c->d_pv2s = nullptr; c->d_pdens = c->d_pvp = c->d_pvp2 = nullptr;
    const size_t nc = (size_t)c->n_sys * 2 * n_bins, nv = nc * 3;
    if (int rc = dalloc(&c->d_pcnt, nc)) return rc;
    if (int rc = dalloc(&c->d_pvs, nv)) return rc;
    if (int rc = dalloc(&c->d_pv2s, nv)) return rc;
    if (int rc = dalloc(&c->d_pdens, nc)) return rc;
    if (int rc = dalloc(&c->d_pvp, nv)) return rc;
    if (int rc = dalloc(&c->d_pvp2, nv)) return rc;
    HIPCHK(hipDeviceSynchronize());
    c->prof_nb = n_bins;
    c->prof_nobs = 0;
    return 0;
}

int mfa_hip_profiles_accumulate(mfa_ctx *c) {
    if (int rc = check_ctx(c)) return rc;
    if (c->prof_nb < 1) return fail(MFA_E_STATE, "profiles_accumulate: call profiles_reset first");
    LAUNCH(k_prof_bin, c->n_blocks, c->st, c->d_sys, c->d_blk_sys, c->B, c->prof_nb, c->d_pcnt, c->d_pvs, c->d_pv2s);
    const int nc = c->n_sys * 2 * c->prof_nb, nv = nc * 3;
    LAUNCH(k_prof_fold, (nv + BLK - 1) / BLK, c->st, nc, nv, c->d_pcnt, c->d_pvs, c->d_pv2s, c->d_pdens, c->d_pvp,
           c->d_pvp2);
    HIPCHK(hipGetLastError());
    c->prof_nobs += 1;
    return 0;
}

int mfa_hip_profiles_get(mfa_ctx *c, int32_t isys, int32_t *n_obs, double *dens_z, double *v_prof, double *v_prof_2) {
    if (int rc = check_ctx(c, false)) return rc;
    if (isys < 0 || isys >= c->n_sys || c->prof_nb < 1) return fail(MFA_E_ARG, "profiles_get: bad argument");
    const size_t nb = c->prof_nb;
    if (n_obs) *n_obs = c->prof_nobs;
    // device layout: dens[s][sp][bin], v[s][sp][d][bin]; reference layout:
    // dens_z(bin, sp), v_prof(bin, d, sp)
    std::vector<double> hd(2 * nb), hv(6 * nb), hv2(6 * nb);
    int rc = to_host(c, hd.data(), c->d_pdens + (size_t)isys * 2 * nb, 2 * nb * sizeof(double), "density profile");
    if (!rc) rc = to_host(c, hv.data(), c->d_pvp + (size_t)isys * 6 * nb, 6 * nb * sizeof(double), "velocity profile");
    if (!rc) rc = to_host(c, hv2.data(), c->d_pvp2 + (size_t)isys * 6 * nb, 6 * nb * sizeof(double), "velocity^2 profile");
    if (rc) return rc;
    if (dens_z) memcpy(dens_z, hd.data(), 2 * nb * sizeof(double));
    for (int sp = 0; sp < 2; ++sp)
        for (int d = 0; d < 3; ++d)
            for (size_t b = 0; b < nb; ++b) {
                const size_t src = ((size_t)sp * 3 + d) * nb + b;
                const size_t dst = b + nb * (d + 3 * sp);
                if (v_prof) v_prof[dst] = hv[src];
                if (v_prof_2) v_prof_2[dst] = hv2[src];
            }
    return 0;
}

int mfa_hip_set_timing(mfa_ctx *c, int32_t on) {
    if (int rc = check_ctx(c, false)) return rc;
    if (on < 0 || on > 2) return fail(MFA_E_ARG, "set_timing: mode must be 0, 1 or 2");
    if (int rc = collect_timing(c)) return rc;
    c->timing_mask = on == 1 ? (1 << K_NK) - 1 : on == 2 ? (1 << K_PAIR) : 0;
    for (int k = 0; k < K_NK; ++k) { c->k_ms[k] = 0.0; c->k_n[k] = 0; }
    return 0;
}

int mfa_hip_get_kernel_time(mfa_ctx *c, const char *kernel, double *total_ms, int64_t *launches) {
    if (int rc = check_ctx(c, false)) return rc;
    if (!kernel || !total_ms || !launches) return fail(MFA_E_ARG, "get_kernel_time: null");
    if (int rc = collect_timing(c)) return rc;
    for (int k = 0; k < K_NK; ++k)
        if (!strcmp(kernel, kNames[k])) { *total_ms = c->k_ms[k]; *launches = c->k_n[k]; return 0; }
    return fail(MFA_E_ARG, std::string("unknown kernel ") + kernel);
}

int mfa_hip_set_neighbor_list(mfa_ctx *c, int32_t capacity) {
    if (int rc = check_ctx(c, false)) return rc;
    if (capacity < -1 || capacity > 4096) return fail(MFA_E_ARG, "set_neighbor_list: capacity must be -1, 0 or 1..4096");
    HIPCHK(hipStreamSynchronize(c->st));
    c->nl_request = capacity;
    invalidate_graph(c);
    if (int rc = alloc_list(c)) return rc;
    // the new list is built by the next rebuild: schedule one for every channel
    for (int s = 0; s < c->n_sys; ++s)
        hipLaunchKernelGGL(k_mark_rebuild, dim3(1), dim3(1), 0, c->st, c->B, s);
    HIPCHK(hipGetLastError());
    return 0;
}

int mfa_hip_get_neighbor_list(mfa_ctx *c, int32_t *capacity) {
    if (int rc = check_ctx(c, false)) return rc;
    if (!capacity) return fail(MFA_E_ARG, "get_neighbor_list: null");
    *capacity = c->B.nl_cap;
    return 0;
}

int mfa_hip_set_dpd_queue(mfa_ctx *c, int32_t depth) {
    if (int rc = check_ctx(c, false)) return rc;
    if (depth < 0 || depth > DPD_Q) return fail(MFA_E_ARG, "set_dpd_queue: depth must be 0..8");
    HIPCHK(hipStreamSynchronize(c->st));
    c->B.dpd_queue = depth;
    invalidate_graph(c);
    return 0;
}

int mfa_hip_set_pair_grid(mfa_ctx *c, int32_t blocks) {
    if (int rc = check_ctx(c, false)) return rc;
    if (blocks < 0) return fail(MFA_E_ARG, "set_pair_grid: blocks must be >= 0");
    HIPCHK(hipStreamSynchronize(c->st));
    c->pair_grid = blocks;
    invalidate_graph(c);
    return 0;
}

int mfa_hip_set_virial(mfa_ctx *c, int32_t on) {
    if (int rc = check_ctx(c, false)) return rc;
    HIPCHK(hipStreamSynchronize(c->st));      // no replay of the old graphs is in flight
    c->virial = on != 0;
    invalidate_graph(c);
    return 0;
}

}  // extern "C"
