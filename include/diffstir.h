/*
 * diffstir.h -- C ABI of the MI355X stirring simulator (libdiffstir_hip.so).
 *
 * The reference exposes one operator pair for this path: Taichi's custom-gradient
 * forward(s)/backward(s) (src/simulator/simulation_functions.py:341-365) over module-global
 * fields (src/simulator/initialization.py:123-165), recorded by ti.ad.Tape
 * (src/gradient_descent_module.py:103-107).  This library replaces that whole layer: a context
 * owns the device state of B independent instances (the reference has B = 1), advances them a
 * number of substeps at a time, records the stick force, and runs the checkpointed adjoint.
 *
 * Conventions
 *  - Every call returns 0 on success or a DS_E* code; ds_last_error() gives the message.
 *  - Host pointers (plain C arrays), except ds_substep_noalloc; the context owns all device memory.
 *  - Calls on one context are serialised on the context's HIP stream (its own, or the caller's
 *    after ds_set_stream); a context is used by one host thread at a time.
 *  - Forces: forces[b][s][3] for s = 0..max_steps-1, F(s) = m_stick/dt * sum_p dv[s,p]
 *    (simulation_functions.py:330-336, force_calculation_and_animation_module.py:16-19);
 *    row 0 is 0 because the stick starts with dv = 0.
 *  - Matrices are row-major 3x3; particle arrays are [n][3] / [n][9] float32.
 */
#ifndef DIFFSTIR_H
#define DIFFSTIR_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ds_ctx ds_ctx;

enum {
    DS_OK = 0,
    DS_EINVAL = 1,     /* bad argument */
    DS_EHIP = 2,       /* HIP runtime error */
    DS_ENOMEM = 3,     /* device allocation failed */
    DS_ESTATE = 4,     /* call out of order (e.g. backward before forward) */
    DS_ENONFINITE = 5, /* a particle left the grid or became non-finite */
};

enum { DS_MODEL_CARREAU = 0, DS_MODEL_CROSS = 1, DS_MODEL_HERSCHEL_BULKLEY = 2 };

/* Options of ds_set_option (no counterpart in the reference: they choose how the substep is
 * executed, never what it computes). */
enum {
    DS_OPT_KERNELS = 1,   /* DS_KERNELS_TILED (default) or DS_KERNELS_PER_PARTICLE; before ds_reset */
    DS_OPT_PARTICLES_PER_THREAD = 2,   /* tiled kernels: 1, 2 (default) or 4 particles per thread */
    DS_OPT_KERNEL_TIMING = 3,          /* time every N-th substep's kernels with HIP events (0 = off) */
    DS_OPT_GATHER_PARTICLES_PER_THREAD = 4,  /* tiled gather kernels (G2P, adjoint P2G): 1 (default), 2, 4 */
    DS_OPT_ADJOINT_P2G = 5,            /* DS_ADJ_P2G_FUSED (default) or DS_ADJ_P2G_SPLIT */
    DS_OPT_SCATTER_CHUNKS = 6,         /* DS_SCATTER_ROWS (default) or DS_SCATTER_PACKED */
    DS_OPT_GRAPH = 7,                  /* 1: a whole-rollout ds_forward (from step 0 to the end) and the
                                          ds_backward after it replay hipGraphs captured at their first
                                          call (no per-kernel timing inside them); 0 (default): launches */
    DS_OPT_OVERLAP = 8,                /* tiled adjoint: 1 (default) recomputes segment k-1 on a second
                                          stream while segment k's adjoint runs (a second segment buffer
                                          set, allocated at the first ds_backward); 0: one stream */
    DS_OPT_CHUNK_GRID = 9,             /* workgroups of each tiled chunk kernel launch (the chunk count lives
                                          on the device; chunks beyond the grid loop): 0 = 16384 (default),
                                          or 64..1048576 */
    DS_OPT_SORT_EVERY = 10,            /* substeps between bucket sorts: 0 (default) = the checkpoint interval,
                                          or a divisor of it.  Sorting more often keeps drifting particles
                                          inside their chunk's tile; the adjoint replays the forward's sorts.
                                          Reallocates the sort tables: takes effect at the next ds_reset */
    DS_OPT_REDUCE_PARTS = 11,          /* wavefronts per node block of the tiled grid reductions: 1, 2 or 4
                                          (default 1); with more, each takes every n-th round of covering
                                          chunk slots and the partial sums are added in a fixed order */
    DS_OPT_XGRAD_STREAM = 12,          /* tiled adjoint: 1 computes the adjoint P2G's position gradient
                                          (through G2P's gather) in its own kernel on a third stream,
                                          beside the adjoint G2P and grid reduction of the same step;
                                          0 (default): inside the adjoint P2G */
    DS_OPT_G2P_SPLIT = 13,             /* workgroups per gather chunk in the forward's G2P: 1 (default), 2
                                          or 4 (at most the gather particles per thread): with 2 particles
                                          per thread the adjoint P2G runs 512-particle chunks while the
                                          forward's G2P runs their two 256-particle halves */
    DS_OPT_TILE_WINDOW = 14,           /* diagnostic: a stencil uses its chunk's LDS tile when its base lies
                                          within this many cells of the tile origin per axis, 0..5 (default
                                          5, the largest the tile holds); others take the global fall-back
                                          grid.  Same results, to the FP32 order of the fall-back atomics */
    DS_OPT_COUNT_FALLBACK = 15,        /* diagnostic: 1 counts the stencils P2G sends to the fall-back grid
                                          (ds_fallback_count); 0 (default): no counting */
    DS_OPT_FUSED_STEP = 16,            /* tiled kernels with row chunks: 1 runs each forward G2P together
                                          with the next step's P2G in one kernel (k_g2p_p2g_tiled) wherever
                                          no sort falls between them and both steps are in the same
                                          ds_forward call (the adjoint's recompute keeps them apart);
                                          0 (default): separate G2P and P2G launches */
};
enum { DS_ADJ_P2G_FUSED = 0, DS_ADJ_P2G_SPLIT = 1 };
/* Scatter kernels (P2G, adjoint G2P): chunks of rows whose lanes have distinct stencil bases, or
 * the packed chunks of the gather kernels, where lanes sharing a base take successive passes. */
enum { DS_SCATTER_ROWS = 0, DS_SCATTER_PACKED = 1 };

/* Kernels timed by DS_OPT_KERNEL_TIMING (tiled family). */
enum {
    DS_K_P2G = 0,          /* k_p2g_tiled: P2G into staging slots */
    DS_K_GRID_REDUCE = 1,  /* k_grid_reduce: slots -> grid update -> velocity */
    DS_K_G2P = 2,          /* k_g2p_tiled */
    DS_K_G2P_ADJ = 3,      /* k_g2p_adj_tiled */
    DS_K_VBAR_REDUCE = 4,  /* k_vbar_reduce */
    DS_K_P2G_ADJ = 5,      /* k_p2g_adj_tiled: adjoint P2G gathers */
    DS_K_SORT = 6,         /* the whole bucket sort of one checkpoint */
    DS_K_P2G_ADJ_STRESS = 7, /* k_p2g_adj_stress: adjoint of the stress and of the FP64 SVD */
    DS_K_G2P_FACTOR = 8,   /* k_g2p_tiled of the recompute, with the full FP64 SVD */
    DS_K_P2G_ADJ_XGRAD = 9, /* k_p2g_adj_xgrad: the adjoint P2G's position gradient */
    DS_K_G2P_P2G = 10,     /* k_g2p_p2g_tiled: G2P fused with the next step's P2G */
    DS_NKERNELS = 11
};
enum { DS_KERNELS_TILED = 0, DS_KERNELS_PER_PARTICLE = 1 };

/* Scene constants shared by all instances of a context.
 * Reference: initialization.py:14-54 (constants), simulation_functions.py:104-225 (hard-coded
 * k = 20000, wall radius 0.0415 - dx/2, bound 3, g = -9.81). */
typedef struct ds_scene_desc {
    int32_t n_grid;        /* nodes per axis; dx = 1/n_grid */
    int32_t bound;         /* box-boundary width in nodes (initialization.py:15) */
    int32_t model;         /* DS_MODEL_* (simulation_functions.py:120,129,138) */
    int32_t k_from_param;  /* 1: bulk modulus = params[3] (Herschel-Bulkley only) */
    double dt;
    double gravity;        /* y acceleration added to liquid velocity (simulation_functions.py:152) */
    double k_bulk;         /* bulk modulus when k_from_param == 0 (reference: 20000) */
    double wall_r, wall_cx, wall_cz;   /* cylindrical slip wall (simulation_functions.py:205-210) */
    double m_liq, vol_liq, m_stick, m_cont;   /* per-particle mass / volume per species */
} ds_scene_desc;

/* Create a context for n_inst instances with the given particle counts per species, a
 * trajectory of max_steps frames (max_steps - 1 substeps, gradient_descent_module.py:105) and
 * an adjoint checkpoint every ckpt_every substeps.  Replaces the field allocation of
 * initialization.py:123-165 (which stores every step; here only checkpoints are kept). */
int ds_create(const ds_scene_desc* desc, int32_t n_inst, const int32_t* n_liquid,
              const int32_t* n_stick, const int32_t* n_container, int32_t max_steps,
              int32_t ckpt_every, ds_ctx** out);
void ds_destroy(ds_ctx* ctx);
/* Queue the context's work on `stream` (a hipStream_t) from the next call on; NULL: the context's
 * own stream (a blocking stream, so it is ordered after work on the null stream).  The library's
 * launches then follow what the caller queued on that stream (e.g. torch's current stream), and
 * calls that return host results wait for that stream.  Waits for the previous stream's work.  No
 * counterpart in the reference, whose kernels run on the Taichi runtime's own stream. */
int ds_set_stream(ds_ctx* ctx, void* stream);
/* Set an execution option (DS_OPT_*).  DS_OPT_KERNELS selects the kernel family: sorted
 * particles with LDS grid tiles (default) or one thread per particle with global atomics; it
 * takes effect at the next ds_reset. */
int ds_set_option(ds_ctx* ctx, int32_t option, int32_t value);
const char* ds_last_error(const ds_ctx* ctx);

/* Initial particle positions of one instance ([n][3] each; NULL allowed for an empty species).
 * DS_EINVAL (nothing stored) when a position is not finite or its stencil leaves the grid: the
 * base cell int(x/dx - 0.5) must lie in [0, n_grid - 3] on every axis.
 * Replaces initialize_objects' from_numpy uploads (simulation_functions.py:293-297). */
int ds_set_particles(ds_ctx* ctx, int32_t inst, const float* x_liquid, const float* x_stick,
                     const float* x_container);
/* Stick trajectory of one instance, [max_steps][3] (operator_kinematic_seq, initialization.py:151). */
int ds_set_trajectory(ds_ctx* ctx, int32_t inst, const float* traj);
/* Constitutive parameters, [n_inst][4] (constitutive_parameters, initialization.py:134).  New values
 * after a recorded rollout (as a new trajectory, ds_set_trajectory) make ds_backward return
 * DS_ESTATE until the next ds_reset: the adjoint recomputes its segments with the current values. */
int ds_set_params(ds_ctx* ctx, const double* params);

/* Step-0 state from the initial positions (v = 0, F = I, C = 0), forces zeroed, step counter
 * reset.  Replaces initialize_ti_field (simulation_functions.py:301-319). */
int ds_reset(ds_ctx* ctx);
/* Advance every instance nsteps substeps from the current step, recording forces and storing
 * a checkpoint every ckpt_every steps.  Replaces the loop of forward(s) calls
 * (simulation_functions.py:341-348, gradient_descent_module.py:105-106). */
int ds_forward(ds_ctx* ctx, int32_t nsteps);
/* Number of substeps taken since ds_reset. */
int ds_current_step(const ds_ctx* ctx, int32_t* step);
/* Stick forces recorded so far, [n_inst][max_steps][3] (total_operator_dv * m / dt,
 * force_calculation_and_animation_module.py:16-19). */
int ds_get_forces(ds_ctx* ctx, double* forces);
/* Adjoint of the whole recorded rollout: given dL/dforces [n_inst][max_steps][3], writes
 * dL/dparams [n_inst][4].  Recomputes each checkpoint segment.  Replaces backward(s)
 * (simulation_functions.py:351-365) as replayed by ti.ad.Tape (gradient_descent_module.py:103-107). */
int ds_backward(ds_ctx* ctx, const double* dl_dforces, double* dl_dparams);
/* Liquid state of one instance at the current step, in the order the particles were given.
 * Any output pointer may be NULL.  x/v: [n][3], C/F: [n][9]. */
int ds_read_liquid(ds_ctx* ctx, int32_t inst, float* x, float* v, float* C, float* F);
/* Stick state of one instance at the current step ([n][3], [n][3], [n][9]). */
int ds_read_stick(ds_ctx* ctx, int32_t inst, float* x, float* v, float* C);
/* Stencils P2G (forward and recompute) sent to the global fall-back grid since the last reset of the
 * counter, when DS_OPT_COUNT_FALLBACK is on; reset != 0 clears the counter afterwards.  A diagnostic
 * of the tiled kernels with no counterpart in the reference (which scatters every particle with
 * global atomics, simulation_functions.py:161-163). */
int ds_fallback_count(ds_ctx* ctx, int64_t* count, int32_t reset);
/* One substep of one instance on caller-owned device memory, allocating nothing.  No counterpart
 * in the reference: it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It is the per-particle form of
 * ds_forward (DS_KERNELS_PER_PARTICLE, one instance, one substep): P2G with global atomics into a
 * dense grid, the grid update, G2P.  Every pointer but desc is device memory:
 *   params      [4] doubles (constitutive_parameters)
 *   state_in    [24][P] floats, component-major: x 0-2, v 3-5, C 6-14 (row-major), F 15-23 for P
 *               slots = the liquid, stick and container runs in that order, each padded to a
 *               multiple of 64 (padding slots are ignored); stick and container F are not read
 *   state_out   [24][P], the state after the substep (may be state_in)
 *   traj2       [2][3] floats: the stick trajectory frames f and f+1
 *   force_out   [3] doubles: the substep's stick force m_stick/dt * sum_p dv (the row f+1 that
 *               ds_get_forces reports)
 *   status_out  [1]: 0, or 1 if a particle left the grid's interior or became non-finite
 * Limits: n_grid <= 64 and P <= 262144 slots (DS_EINVAL beyond them).  Its grid, tables and flag
 * live in one 9 MiB device array per GPU, held until the process ends; the call clears what must
 * start at zero.  Calls that overlap on one GPU race on that array.  The work is queued on
 * `stream` (a hipStream_t; NULL is the null stream) and the call returns without waiting.
 * Argument errors return DS_EINVAL with the message in ds_last_error(NULL). */
int ds_substep_noalloc(const ds_scene_desc* desc, int32_t n_liquid, int32_t n_stick, int32_t n_container,
                       const double* params, const float* state_in, float* state_out, const float* traj2,
                       double* force_out, uint32_t* status_out, void* stream);
/* Wait for all work queued on the context's stream. */
int ds_synchronize(ds_ctx* ctx);

/* Timing of the most recent ds_forward / ds_backward, measured with HIP events on the
 * context's stream: elapsed milliseconds and kernel launches. */
int ds_last_timing(const ds_ctx* ctx, double* fwd_ms, double* bwd_ms);

/* Per-kernel GPU time accumulated since the last reset by DS_OPT_KERNEL_TIMING sampling:
 * total_ms[DS_NKERNELS], launches[DS_NKERNELS]; reset != 0 clears the counters afterwards.  A
 * sample is a HIP event pair on the launch (hipExtLaunchKernel), taken after a short hold of the
 * stream so that the host has queued the substep; it includes the dispatch gap an event-recording
 * packet causes (about 4.6 us on MI355X against the kernel trace). */
int ds_kernel_timing(ds_ctx* ctx, double* total_ms, int64_t* launches, int32_t reset);

/* FP64 3x3 signed SVD on the device, for tests: F[n][9] -> U[n][9], S[n][3], V[n][9]
 * with F = U diag(S) V^T, det U = det V = +1 and S[2] carrying the sign of det F.
 * (The reference's svd(), simulation_functions.py:55-58, uses ti.svd in f32.) */
int ds_svd3(const double* F, int32_t n, double* U, double* S, double* V);
/* The singular values alone, as the forward substep computes them for J = S0 S1 S2
 * (simulation_functions.py:55-58 and :113): F[n][9] -> S[n][3], same signed convention as
 * ds_svd3, for tests. */
int ds_svd3_values(const double* F, int32_t n, double* S);
/* FP64 SVD backward on the device, for tests: given U, S, V and the adjoints gU, gS (diagonal,
 * [n][3]), gV, writes gF.  simulation_functions.py:61-87 (clamped 1/(s_j^2 - s_i^2)). */
int ds_svd3_backward(const double* U, const double* S, const double* V, const double* gU,
                     const double* gS, const double* gV, int32_t n, double* gF);

#ifdef __cplusplus
}
#endif

#endif /* DIFFSTIR_H */
