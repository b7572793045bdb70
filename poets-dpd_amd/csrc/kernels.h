// kernels.h -- launch interface of the step kernels (kernels.hip).
#pragma once

#include "device_types.h"
#include "pdpd_hip.h"

namespace pdpd {

// Everything one step needs to launch.
struct StepLaunch {
    DeviceState S;
    int cur;                 // buffer holding the state at step start (the sort writes cur ^ 1)
    int cpar;                // counter copy this step reads (the next step's is cpar ^ 1)
    int integrate;           // run k_integrate (the first step of a chain)
    int fuse;                // the pair-force kernel integrates the next step (all but a chain's last step)
    const NextBuffers *next; // device table of 4 entries, [cur * 2 + cpar]
    uint32_t n_beads;
    uint32_t n_cells;
    uint32_t n_worlds;
    uint32_t n_bead_blocks;
    uint32_t n_force_blocks;
    uint32_t n_stationary;   // beads of stationary types
    uint32_t n_terms;        // bonded terms (Hookean bonds and angles) of all worlds
    uint32_t force_grid;     // 0 = one workgroup per force block, else persistent grid size
    uint32_t force_variant;  // pair-force kernel variant (loads in flight / occupancy), same results
    uint32_t *tile_sum;      // scan_tiles(n_cells) entries: tile sums of the three-phase scan
    uint32_t lookback_tiles; // single-pass look-back scan up to this many tiles, three-phase above
    uint32_t spatial;        // some world has per-cell strengths (the pair-force kernel with the cell lookup)
    uint32_t force_kernel;   // 0: per-slot blocks (k_force), 1: cell bricks staged in LDS (k_force_brick)
    uint32_t brick_shape;    // brick shape index (brick_shape_dims)
    uint32_t n_bricks;
    const BrickRec *bricks;
};

constexpr uint32_t kForceVariants = 8;

// Events recorded between the kernels of one step (profiling runs only):
// [0] integrate [1] scan [2] scatter [3] reorder [4] force (+ bonded terms) [5] stationary [6]
struct KernelEvents {
    hipEvent_t ev[7];
};

uint32_t scan_tiles(uint32_t n_cells);
hipError_t launch_step(const StepLaunch &L, hipStream_t st, KernelEvents *ev);
// The pair-force launch of the step L describes (reads and writes buffer
// L.cur ^ 1; the bonded terms were evaluated during the step's scatter).
hipError_t launch_force(const StepLaunch &L, hipStream_t st);
// The pair-force launch over cell bricks (force_brick.hip); same inputs, outputs
// and bits as launch_force's per-slot kernel.
hipError_t launch_force_brick(const StepLaunch &L, hipStream_t st);
uint32_t brick_shape_count();
void brick_shape_dims(uint32_t shape, uint32_t *bx, uint32_t *by, uint32_t *bz);
// The pair-force launch over slabs of cells one cell thick in z
// (force_slab.hip); same inputs, outputs and bits. The brick shapes from
// kSlabShape0 on name its shapes.
hipError_t launch_force_slab(const StepLaunch &L, uint32_t shape, hipStream_t st);
// Pair forces with one wave per stencil row, nine per workgroup (force_rows.hip):
// eight candidate loads per lane in flight, 512-pair queues.
hipError_t launch_force_rows(const StepLaunch &L, hipStream_t st);
uint32_t slab_shape_count();
void slab_shape_dims(uint32_t shape, uint32_t *bx, uint32_t *by);
hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st);
// Relayout: one segment per world of the new layout, in order (see
// k_move_worlds); blocks [first_block, next segment's first_block) serve it,
// 256 entries each of max(n, n_list, n_terms).
struct MoveSeg {
    uint32_t first_block, world, fresh, n;
    uint32_t src_off, dst_off;
    uint32_t n_list, src_list, dst_list;
    uint32_t n_terms, src_term, dst_term;
};
struct MoveBuffers {
    const float4 *sp, *sv, *sf;
    float4 *dp, *dv, *df;
    const uint32_t *ssob, *sstart, *slist;
    uint32_t *dsob, *dstart, *dlist;
    const TermRec *sterms;
    TermRec *dterms;
};
hipError_t launch_move_worlds(const MoveSeg *segs, uint32_t n_segs, uint32_t n_blocks, const MoveBuffers &B,
                              hipStream_t st);
// Per-world flags carried to a new layout (nw[i] = old[map[i]], 0 where map[i] is ~0).
hipError_t launch_remap_u32(uint32_t *nw, const uint32_t *old, const uint32_t *map, uint32_t n, hipStream_t st);
// Pair records of world `world` (n_beads beads) from buffer buf, step counter t; appended at *count.
hipError_t launch_trace_pairs(const DeviceState &S, int buf, uint32_t world, uint32_t n_beads, uint32_t t,
                              pdpd_pair_record *out, uint32_t cap, uint32_t *count, hipStream_t st);
// Per-world FP64 observables of buffer buf (momentum, sum v.v, conservative
// virial, pair and bead counts; kObsValues doubles per world) into out, via
// one partial per bead block in partial (n_bead_blocks x kObsValues doubles).
// first: n_worlds + 1 first bead-block indices.
constexpr int kObsValues = 8;
hipError_t launch_observables(const DeviceState &S, int buf, uint32_t n_bead_blocks, const uint32_t *first,
                              uint32_t n_worlds, double *partial, double *out, hipStream_t st);
hipError_t launch_debug_rng(uint32_t n, const uint64_t *th, const uint32_t *a, const uint32_t *b, uint32_t *u32,
                            float *f32, hipStream_t st);

}  // namespace pdpd
