/*
 * pdpd_hip.h -- C ABI of the MI355X (gfx950) DPD engine library libpdpd_hip.so.
 *
 * The reference has no foreign-function boundary: its engines are C++ classes
 * behind DPDEngine (include/dpd/core/dpd_engine.hpp:10-216 in poets-dpd). This
 * ABI is the thin layer the repository's C++ DPDEngine subclasses
 * (include/pdpd/hip_dpd_engine.hpp) call; each entry point below names the
 * reference interface whose work it carries. Plain pointers and sizes only.
 *
 * Conventions
 *   - Every function returns PDPD_OK (0) or a negative PDPD_ERR_* code; the
 *     message of the last failure on the calling thread is pdpd_last_error().
 *   - A context owns one HIP device and one stream. It holds any number of
 *     resident worlds ("slots"); pdpd_run advances all of them together in one
 *     set of launches per step (the plain engine is a context with one world).
 *   - World data crosses the ABI in the reference's WorldState layout
 *     (dpd_state.hpp:141-264): FP64 x/v/f in bead-index order plus the polymer
 *     topology. On the device the state is FP32 SoA sorted by cell.
 *   - Calls on one context must not overlap; different contexts are independent.
 */
#ifndef PDPD_HIP_H
#define PDPD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 2: pdpd_world_desc gained cell_conservative / cell_dissipative;
 * 3: pdpd_server_config gained contexts;
 * 4: pdpd_request_result gained fail_flags (status 4: failed);
 * 5: pdpd_stats gained relayouts / relayout_s / relayout_max_s / relayout_max_beads,
 *    pdpd_server_stats gained relayouts / relayout_mean_s / relayout_max_s /
 *    max_resident_beads. */
#define PDPD_ABI_VERSION 5

enum {
    PDPD_OK = 0,
    PDPD_ERR_INVALID = -1,      /* bad argument or unknown slot / key            */
    PDPD_ERR_UNSUPPORTED = -2,  /* world the engine cannot run (see can_support)  */
    PDPD_ERR_HIP = -3,          /* HIP runtime failure                            */
    PDPD_ERR_CAPACITY = -4,     /* a kernel flagged an occupancy overflow         */
    PDPD_ERR_STATE = -5,        /* a kernel flagged a bead that left the box       */
    PDPD_ERR_NOMEM = -6         /* device or host allocation failed               */
};

/* Error flags a step can raise (pdpd_stats.flags). */
#define PDPD_FLAG_BAD_POSITION   0x1u   /* non-finite x, or a jump of more than one box */
#define PDPD_FLAG_SCAN_TIMEOUT   0x2u   /* the cell scan look-back ran out of spins      */
#define PDPD_FLAG_DENSE_ROW      0x4u   /* reserved: no longer raised (rows of any length are handled) */
/* Per-world failure flags, raised by a world's own beads, bonds and forces
 * (the reference's engines throw or exit on the same conditions): */
#define PDPD_FLAG_CELL_JUMP      0x8u   /* a bead moved more than one cell in a step
                                           (dpd_engine_avx2_half_merge_tbb.hpp:800-811)  */
#define PDPD_FLAG_BOND_STRETCH   0x10u  /* a Hookean or angle bond longer than the engine's
                                           maximum bond length, half the smallest box side
                                           (naive_dpd_engine_half_step.hpp:435-437)       */
#define PDPD_FLAG_FORCE_BOUND    0x20u  /* a force component not finite                     */
#define PDPD_FLAG_VELOCITY_BOUND 0x40u  /* a velocity component not finite                  */

/* Text of the last error on this thread ("" if none). */
const char *pdpd_last_error(void);
int pdpd_abi_version(void);

/* One world, in the reference's WorldState terms (dpd_state.hpp:141-264).
 * Arrays are read during pdpd_world_add only. */
typedef struct pdpd_world_desc {
    double box[3];                 /* WorldState::box, integer valued          */
    double dt;                     /* WorldState::dt                           */
    double lambda;                 /* WorldState::lambda (engine uses 0.5 form) */
    uint64_t seed;                 /* WorldState::seed                         */
    int64_t t;                     /* WorldState::t                            */

    uint32_t n_bead_types;
    const double *conservative;    /* n_bead_types^2, row-major (interactions[i].conservative) */
    const double *dissipative;     /* n_bead_types^2 (interactions[i].dissipative = gamma)     */
    const uint8_t *stationary;     /* n_bead_types or NULL (BeadType::stationary)             */

    uint32_t n_beads;
    const double *x, *v, *f;       /* 3*n_beads, xyz per bead, bead-index order */
    const uint8_t *bead_type;      /* Bead::bead_type      */
    const uint8_t *is_monomer;     /* Bead::is_monomer     */
    const uint32_t *polymer_id;    /* Bead::polymer_id     */
    const uint8_t *polymer_offset; /* Bead::polymer_offset */

    uint32_t n_polymer_types;      /* PolymerType (dpd_state.hpp:212-224)           */
    const uint32_t *pt_bond_start; /* n_polymer_types+1 -> bonds                    */
    const uint32_t *bond_head;     /* Bond::bead_offset_head                        */
    const uint32_t *bond_tail;     /* Bond::bead_offset_tail                        */
    const double *bond_kappa;      /* Bond::kappa                                    */
    const double *bond_r0;         /* Bond::r0                                       */
    const uint32_t *pt_angle_start;/* n_polymer_types+1 -> bond pairs               */
    const uint32_t *angle_head_bond;   /* BondPair::bond_offset_head                */
    const uint32_t *angle_tail_bond;   /* BondPair::bond_offset_tail                */
    const double *angle_kappa;     /* BondPair::kappa                                */
    const double *angle_theta0;    /* BondPair::theta0                               */

    uint32_t n_polymers;           /* Polymer (dpd_state.hpp:226-234)               */
    const uint32_t *polymer_type;  /* Polymer::polymer_type                          */
    const uint32_t *polymer_bead_start; /* n_polymers+1 -> polymer_bead_ids         */
    const uint32_t *polymer_bead_ids;   /* Polymer::bead_ids                         */

    /* Spatially varying DPD strengths (naive_dpd_engine.hpp:145-188): the
     * interactions evaluated at every cell, n_cells * n_bead_types^2 each, cell
     * index (x * Ly + y) * Lz + z, or NULL for constant strengths (then the
     * arrays above are used). A pair uses the mean of its two beads' cells
     * (naive_dpd_engine.hpp:452-468). */
    const double *cell_conservative;
    const double *cell_dissipative;
} pdpd_world_desc;

typedef struct pdpd_limits {
    uint32_t max_bead_types;      /* 16: the 4-bit type field of BeadHash (dpd_state.hpp:18-35) */
    uint32_t min_box;             /* 3: smallest side for a 27-cell stencil without aliasing     */
    uint32_t max_box;             /* largest side accepted                                       */
    uint32_t max_polymer_length;  /* 127: the 7-bit offset field of BeadHash                     */
    uint32_t max_beads;           /* beads one context may hold (26-bit slots in the pair lists)  */
    uint32_t max_bonds_per_bead;  /* bonds + angles touching one bead                            */
} pdpd_limits;

/* Capability / limit query (DPDEngine::CanSupport, dpd_engine.hpp:57-128). */
int pdpd_get_limits(pdpd_limits *out);

/* "" if the world can run, otherwise the reason, as CanSupport returns it.
 * The pointer stays valid until the next call on this thread. */
const char *pdpd_can_support(const pdpd_world_desc *w);

typedef struct pdpd_ctx pdpd_ctx;

/* Context = device + stream + resident worlds (DPDEngine instance state). */
int pdpd_ctx_create(int device, pdpd_ctx **out);
int pdpd_ctx_destroy(pdpd_ctx *ctx);

/* Attach a world (DPDEngine::Attach, dpd_engine.hpp:171-177): imports AoS FP64
 * into FP32 with the v -= f*dt/2 shift of dpd_engine_avx2_half_merge_tbb.hpp:524.
 * The upload takes effect at the next pdpd_run. */
int pdpd_world_add(pdpd_ctx *ctx, const pdpd_world_desc *w, int *slot_out);

/* Detach (Attach(nullptr)). Resident data of the slot is dropped. */
int pdpd_world_remove(pdpd_ctx *ctx, int slot);

/* Steps this world may advance in the next pdpd_run (default: all of them).
 * A world whose budget is spent keeps its state while the others advance. */
int pdpd_world_set_steps(pdpd_ctx *ctx, int slot, uint32_t steps);

/* Advance all resident worlds by nsteps (DPDEngine::Run(nSteps),
 * dpd_engine.hpp:179-185). Asynchronous: returns once the work is enqueued. */
int pdpd_run(pdpd_ctx *ctx, uint32_t nsteps);

/* Wait for enqueued work; reports kernel-raised flags as errors. */
int pdpd_sync(pdpd_ctx *ctx);

/* PDPD_FLAG_* bits raised by one world since it was added (waits for the
 * enqueued steps). pdpd_sync and the downloads fail with PDPD_ERR_STATE when a
 * resident world raises one, naming the world; the flags stay with the world,
 * so a caller running many worlds (the serving loop) can retire the failing
 * one and continue with the others, whose results do not depend on it. */
int pdpd_world_flags(pdpd_ctx *ctx, int slot, uint32_t *flags);

/* Synchronise and export one world in bead-index order (the end of Run:
 * dpd_engine_avx2_half_merge_tbb.hpp:537-557, v += f*dt/2 in FP64).
 * Any output pointer may be NULL. */
int pdpd_world_download(pdpd_ctx *ctx, int slot, double *x, double *v, double *f, int64_t *t);

/* Interval snapshots overlapped with compute (the Run(intervals) contract,
 * dpd_engine.hpp:187-215, with run_world.cpp:177-275's callback per interval).
 * pdpd_snapshot enqueues a device-side copy of every resident world's state as
 * it will be when the work enqueued so far has run, and returns at once; more
 * steps may be enqueued behind it. pdpd_world_snapshot_download waits for that
 * copy only (not for later steps), brings one world's share to the host on a
 * second stream and exports it as pdpd_world_download does, with the world's
 * t at the snapshot; worlds removed since remain downloadable, worlds added
 * since the last pdpd_run are not in it. One snapshot
 * is held per context (48 bytes of device memory per resident bead); a new
 * one replaces it. A bead that had left the box by the snapshot is reported
 * here as PDPD_ERR_STATE. */
int pdpd_snapshot(pdpd_ctx *ctx);
int pdpd_world_snapshot_download(pdpd_ctx *ctx, int slot, double *x, double *v, double *f, int64_t *t);

/* Raw FP32 engine state of one world in bead-index order: x, the shifted
 * velocity w = v - f*dt/2, f; and the device step counter. For parity tests. */
int pdpd_world_download_f32(pdpd_ctx *ctx, int slot, float *x, float *w, float *f, uint32_t *t);

/* Cell order of one world after the last step: order[k] = bead index at sorted
 * position k; cell_start[c] for c in [0, Lx*Ly*Lz] (world-local). */
int pdpd_world_download_cells(pdpd_ctx *ctx, int slot, int32_t *order, int32_t *cell_start);

/* Engine configuration (DPDEngine::GetConfigEntries / SetConfigEntry,
 * dpd_engine.hpp:150-169). Unknown keys fail with PDPD_ERR_INVALID.
 * Keys: "graph_steps" (steps per replayed hipGraph, even; 0 = direct launches),
 *       "force_grid" (pair-force grid-stride size; 0 = one workgroup per 64 beads). */
int pdpd_set_config(pdpd_ctx *ctx, const char *key, const char *value);
int pdpd_get_config(pdpd_ctx *ctx, const char *key, char *buf, size_t buflen);
/* Newline-separated "key=value\tdescription" list of all entries. */
int pdpd_list_config(pdpd_ctx *ctx, char *buf, size_t buflen);

/* Timings and kernel statistics (DPDEngine::GetTimings, dpd_engine.hpp:38-46,145). */
typedef struct pdpd_stats {
    double upload_s;              /* host->device import of attached worlds (timings_t::configure) */
    double run_s;                 /* wall time inside pdpd_run + pdpd_sync                        */
    uint64_t steps;               /* steps enqueued                                               */
    uint64_t bead_steps;          /* sum over worlds of beads x steps actually advanced           */
    uint32_t flags;               /* PDPD_FLAG_* raised since the last reset                      */
    uint32_t n_worlds;
    uint64_t n_beads;
    uint32_t n_force_blocks;      /* work items of the pair-force kernel (64 beads each)          */
    /* average kernel durations (ms) of the last profiled run, 0 if none; k_force includes the
     * bonded terms, k_stationary the velocity clear of stationary beads (0 without them) */
    double k_integrate_ms, k_scan_ms, k_scatter_ms, k_reorder_ms, k_force_ms, k_stationary_ms;
    uint64_t profiled_steps;
    uint64_t relayouts;           /* device layouts built (a world added or removed before a run) */
    double relayout_s;            /* host time spent building them (the device work is enqueued)  */
    double relayout_max_s;        /* the longest of them                                           */
    uint64_t relayout_max_beads;  /* resident beads of the layout that took relayout_max_s         */
} pdpd_stats;

int pdpd_get_stats(pdpd_ctx *ctx, pdpd_stats *out);

/* Run nsteps with direct launches and HIP events around every kernel on the
 * context's stream; fills the k_*_ms averages of pdpd_stats. Synchronous. */
int pdpd_run_profiled(pdpd_ctx *ctx, uint32_t nsteps);

/* Measurement hook (no reference counterpart; the reference times whole Run
 * calls, benchmark_engine.cpp:128-147): re-launches the pair-force kernel of
 * the last step reps times back to back on the context's stream, between two
 * HIP events, and returns the average duration of one launch. The kernel
 * rewrites the same forces from the same state, so the state is unchanged.
 * Needs at least one completed step. Synchronous. */
int pdpd_time_force(pdpd_ctx *ctx, uint32_t reps, double *avg_ms);

/* One evaluated pair of a force trace: the quantities the reference's force
 * logging records per bead pair (dpd_maths_core_simd.hpp:232-275,
 * dpd_maths_core_half_step_raw.hpp:190-214), from the home bead's side. */
typedef struct pdpd_pair_record {
    uint32_t home, other;        /* bead hashes (BeadHash, with the type bits)           */
    float dx[3];                 /* x_home - x_other, nearest image                       */
    float dr;                    /* |dx| (< 1)                                            */
    float dv[3];                 /* v_home - v_other, half-step velocities                */
    float diss_strength;         /* gamma_ij                                              */
    float inv_root_dt;           /* sqrt(24 kT / dt)                                      */
    float gammap;                /* (sqrt(gamma_ij) (1 - dr))^2                           */
    float rng;                   /* the pair's uniform draw u in [-0.5, 0.5)             */
    float con, diss, rng_scale, rand;   /* conservative, dissipative, sqrt(gammap) sqrt(24/dt), random */
    float f[3];                  /* the pair's force on the home bead, as the step used it */
} pdpd_pair_record;

/* One Hookean bond of a force trace (dpd_engine_avx2_half_merge_tbb.hpp:1244-1259). */
typedef struct pdpd_bond_record {
    uint32_t head, tail;         /* bead hashes                                          */
    double dx[3];                /* x_tail - x_head, nearest image                        */
    double dr;
    double f_head[3];            /* force on the head (the tail gets its negation)        */
} pdpd_bond_record;

/* Force trace of the step that produced a world's current state: every pair
 * evaluation of that step (one record per moving home bead and neighbour
 * within the cutoff, sorted by (home, other)) and every Hookean bond. The step
 * must have run on the current layout (no world added or removed, no autotune
 * since), else PDPD_ERR_STATE. *n_pairs / *n_bonds receive the counts; records
 * are written up to the capacities, and PDPD_ERR_INVALID is returned when a
 * capacity is too small (pass 0 to query the counts). A debugging aid, not a
 * fast path: one extra kernel and a download. Pairs with a stationary bead
 * use its stored velocity, which is 0 after the world's first step. */
int pdpd_trace_forces(pdpd_ctx *ctx, int slot, pdpd_pair_record *pairs, uint32_t pair_capacity,
                      uint32_t *n_pairs, pdpd_bond_record *bonds, uint32_t bond_capacity, uint32_t *n_bonds);

/* FP64 observables of one world, reduced on the device from its FP32 state
 * (no download of the state): the momentum and sum v.v the reference checks
 * (test_ordered_mesh.hpp:85-93; stats_dpd_engine_half_step.hpp:539-572 reports
 * per-step velocity statistics), and the conservative virial. v is the
 * exported velocity w + f dt/2; pairs are the unordered pairs of beads closer
 * than the cutoff (nearest image, FP64 distance of the FP32 positions), each
 * contributing r_ij . F^C_ij = a_ij (1 - r) r with the world's FP32 a_ij. The
 * reference has no pressure; the one here is the textbook virial form with
 * m = kT = 1. */
typedef struct pdpd_observables {
    int32_t slot;                /* world slot                                           */
    uint32_t n_beads;
    int64_t t;                   /* WorldState::t of the state the sums are taken from   */
    double momentum[3];          /* sum_i v_i                                            */
    double kinetic;              /* sum_i v_i . v_i                                      */
    double virial;               /* sum over pairs with r < 1 of a_ij (1 - r) r          */
    uint64_t n_pairs;            /* unordered pairs with r < 1                           */
    double temperature;          /* kinetic / (3 n_beads)                                */
    double pressure;             /* (n_beads temperature + virial / 3) / volume          */
} pdpd_observables;

/* Observables of every resident world (layout order, at most `capacity`; *n_out
 * receives the count). Waits for the enqueued steps. */
int pdpd_get_observables(pdpd_ctx *ctx, pdpd_observables *out, uint32_t capacity, uint32_t *n_out);
/* Observables of one world. */
int pdpd_world_observables(pdpd_ctx *ctx, int slot, pdpd_observables *out);

/* pdpd_get_observables without any allocation. No counterpart in the
 * reference; for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The
 * partial sums of the bead blocks go to a file-scope device array of 8192
 * blocks x 8 doubles (512 KiB, held by every GPU the library runs on until the
 * process ends); the results go to out_dev, device memory the caller passes,
 * 8 doubles per resident world in layout order: momentum x, y, z, sum v.v,
 * virial, pairs, beads, 0 (the pdpd_observables fields before the derived
 * ones). Enqueued on the context's stream, not waited for. Fails with
 * PDPD_ERR_CAPACITY when the resident worlds span more than 8192 bead blocks
 * (2^21 beads), PDPD_ERR_STATE when the layout must be rebuilt first (a world
 * was added or removed since the last step), PDPD_ERR_INVALID when out_worlds
 * is below the resident world count. Calls on one GPU share the array: calls
 * that overlap (other contexts' streams, other threads) race. */
int pdpd_observables_noalloc(pdpd_ctx *ctx, double *out_dev, uint32_t out_worlds);

/* pdpd_snapshot plus the FP64 observables (as pdpd_get_observables) of every
 * world in that snapshot, reduced on the device behind the enqueued steps;
 * returns without waiting, and steps may be enqueued behind it.
 * pdpd_world_snapshot_observables waits for those sums only and copies one
 * world's (a few dozen bytes, no copy of its state), or fails with
 * PDPD_ERR_STATE, naming the flags, when that world had raised a failure
 * flag by then. For per-interval statistics of resident worlds (the serving
 * loop's observables callback). No counterpart in the reference, whose
 * statistics engine (stats_dpd_engine_half_step.hpp:539-572) reduces on the
 * host after every step. */
int pdpd_snapshot_observables(pdpd_ctx *ctx);
int pdpd_world_snapshot_observables(pdpd_ctx *ctx, int slot, pdpd_observables *out);

/* Pick the launch shape (pair-force grid, graph length) for the resident worlds by timing
 * candidates for steps_per_trial steps each (state is saved and restored, so
 * the worlds do not advance); the choice stays in the configuration. */
int pdpd_autotune(pdpd_ctx *ctx, uint32_t steps_per_trial, char *choice, size_t choice_len);

/* The same over n contexts of one device that run side by side (an ensemble
 * spread over several contexts, whose streams overlap): every trial enqueues
 * the steps of all of them and is timed from the first context's start to the
 * last one's end; every context takes the chosen shape. No counterpart in the
 * reference, which tunes nothing (its optimisable.hpp:5-46 only sketches a
 * tuning interface). */
int pdpd_autotune_group(pdpd_ctx **ctxs, uint32_t n, uint32_t steps_per_trial, char *choice, size_t choice_len);

/* ---- serving: continuous batching over one context (no counterpart in the
 * reference, whose engines run one world per Run) ----
 * Requests queue in arrival order. At each interval boundary the loop retires
 * finished requests, admits arrived ones while the resident set fits the HBM
 * budget, and advances every resident request by one interval in shared
 * launches; requests with fewer steps left stop early via their step budget. */
typedef struct pdpd_server pdpd_server;

typedef struct pdpd_server_config {
    int device;
    uint64_t hbm_budget_bytes;    /* admission budget for resident worlds, 0 = unlimited */
    uint32_t interval_steps;      /* steps between admissions / callbacks (e.g. 500)     */
    uint32_t max_resident;        /* 0 = unlimited                                       */
    double time_scale;            /* arrival times are multiplied by this; 0 = all at once */
    uint32_t contexts;            /* resident requests are spread over this many contexts
                                     (streams whose launches overlap); 0 or 1 = one     */
} pdpd_server_config;

/* Called after every interval for every resident request with its exported
 * state (bead-index order). Return 0 to cancel the request. */
typedef int (*pdpd_interval_fn)(void *user, uint64_t request, uint32_t steps_done, int64_t t, uint32_t n_beads,
                                const double *x, const double *v, const double *f);

typedef struct pdpd_request_result {
    uint64_t id;
    uint32_t n_beads;
    uint32_t steps;
    double arrival_s, admit_s, done_s, latency_s;   /* seconds from the start of pdpd_server_run */
    int status;                                     /* 0 done, 1 queued, 2 resident, 3 cancelled,
                                                       4 failed (its world raised fail_flags)       */
    uint32_t fail_flags;                            /* PDPD_FLAG_* of a failed request            */
} pdpd_request_result;

typedef struct pdpd_server_stats {
    uint64_t requests, completed, intervals, max_resident;
    double makespan_s, gpu_busy_s;                  /* gpu_busy_s: time with a request resident */
    double bead_steps, bead_steps_per_s;             /* sum of beads x steps over the makespan */
    double p50_latency_s, p99_latency_s, mean_latency_s, max_latency_s;
    uint64_t failed;                                 /* requests retired because their world failed */
    /* layouts rebuilt on admission or retirement (host time; the moves run on the device) */
    uint64_t relayouts, max_resident_beads;
    double relayout_mean_s, relayout_max_s;
} pdpd_server_stats;

int pdpd_server_create(const pdpd_server_config *cfg, pdpd_server **out);
int pdpd_server_destroy(pdpd_server *s);
int pdpd_server_set_config(pdpd_server *s, const char *key, const char *value);
int pdpd_server_set_callback(pdpd_server *s, pdpd_interval_fn fn, void *user);
/* Per-interval observables instead of (or beside) the state: fn(user,
 * request, steps_done, observables) after every interval of every resident
 * request, the FP64 sums reduced on the device from the interval's snapshot
 * (pdpd_snapshot_observables), so a request's state is copied to the host
 * only when it completes or when a state callback is also set. Returning 0
 * cancels the request, as for the state callback. NULL removes it. */
typedef int (*pdpd_observables_fn)(void *user, uint64_t request, uint32_t steps_done, const pdpd_observables *obs);
int pdpd_server_set_observables_callback(pdpd_server *s, pdpd_observables_fn fn, void *user);
/* Queue a request; the world is copied. */
int pdpd_server_submit(pdpd_server *s, const pdpd_world_desc *w, double arrival_s, uint32_t steps, uint64_t *id_out);
/* Serve until every queued request has completed (arrivals against the wall clock). */
int pdpd_server_run(pdpd_server *s);
int pdpd_server_result(pdpd_server *s, uint64_t id, pdpd_request_result *out, double *x, double *v, double *f);
int pdpd_server_get_stats(pdpd_server *s, pdpd_server_stats *out);

/* ---- test hooks (no counterpart in the reference) ---- */

/* hash_rng_sym / default_hash (hash.hpp:81-88, dpd_maths_core.hpp:55-79) on the
 * device for n triples; outputs may be NULL. */
int pdpd_debug_rng(int device, uint32_t n, const uint64_t *t_hash, const uint32_t *a,
                   const uint32_t *b, uint32_t *u32_out, float *f32_out);

#ifdef __cplusplus
}
#endif
#endif
