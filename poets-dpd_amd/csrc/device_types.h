// device_types.h -- layouts shared by the gfx950 kernels and the host side of
// libpdpd_hip.so. All device state lives in HBM as SoA-of-float4, sorted by cell.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdpd {

constexpr int kMaxTypes = 16;          // 4-bit type field of the bead hash
constexpr int kBeadThreads = 256;      // bead-parallel kernels
constexpr int kForceBeads = 64;        // beads per pair-force workgroup (one per lane of a wave)
constexpr int kForcePlanes = 3;        // pair waves per pair-force workgroup: one per x-plane of the stencil
constexpr int kForceThreads = kForceBeads * (kForcePlanes + 1);   // + one wave for the bonded terms
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 2;
constexpr int kScanBlock = kScanThreads * kScanPerThread;   // cells per scan tile
constexpr uint32_t kMaxLayoutBeads = 1u << 26;              // pair-list entries carry 26-bit slots

// Per-world parameters, read by every kernel through the world index.
struct WorldDev {
    int32_t L[3];          // box in cells
    float Lf[3];
    float dt, hdt;         // dt and dt/2
    float rng_scale;       // sqrt(24 kT / dt), kT = 1
    uint32_t bead_off;     // first slot (= first global bead index) of this world
    uint32_t n_beads;
    uint32_t cell_off;     // first global cell id
    uint32_t n_cells;
    uint32_t seed_lo, seed_hi;
    uint32_t n_types;
    uint32_t stationary;   // bit k set: bead type k is stationary (never moves, v = f = 0 after a step)
    uint32_t spatial;      // 1: strengths per cell in cell_par (a and sg below unused)
    uint32_t cell_par_off; // first entry of this world's per-cell strengths in cell_par
    // failure bounds: |v| <= max_v and |f| <= max_f per component (FLT_MAX:
    // finite), bonds shorter than max_bond (half the smallest side, the
    // nearest image's limit)
    float max_v, max_f;
    double max_bond;
    float a[kMaxTypes * kMaxTypes];    // conservative a_ij
    float sg[kMaxTypes * kMaxTypes];   // sqrt(gamma_ij)
};

// Mutable per-world counters (kept apart from WorldDev so the table stays read-only).
struct WorldCtr {
    uint32_t t;            // step counter used for the RNG (WorldState::t truncated to u32)
    uint32_t remaining;    // steps this world may still advance in the current run
};

// One bead-parallel block: [begin, end) slots of one world.
struct BeadBlock {
    int32_t world;
    uint32_t begin, end;
    uint32_t pad;
};

// One pair-force brick: a box of cells of one world, [x0, x0 + bx) x [y0, y0 +
// by) x [z0, z0 + bz) (clipped at the world's far faces). origin packs x0 | y0
// << 10 | z0 << 20 (box sides are at most 1024 cells); dims packs bx | by << 8 |
// bz << 16.
struct BrickRec {
    int32_t world;
    uint32_t origin;
    uint32_t dims;
    uint32_t pad;
};

// Bonded terms: one record per Hookean bond or angle, global bead indices.
//   kind 0 (Hookean): a = head, b = tail, p0 = r0
//   kind 1 (angle):   a = head, b = middle, c = tail, p0 = cos(theta0), p1 = sin(theta0)
// kind holds the term kind in bits 0-7 and the world (index into the world
// table) in bits 8-31. Bonded terms are evaluated in FP64 (during k_scatter), once per
// term per step, into term_frc.
struct TermRec { uint32_t a, b, c, kind; double kappa, p0, p1; };
constexpr uint32_t kTermBond = 0, kTermAngle = 1;
constexpr uint32_t kTermWorldShift = 8;
// term_frc: 6 doubles per term. Hookean: the tail's force (the head gets its
// negation). Angle: the head's force, then the tail's (the middle gets
// -(head + tail)).
constexpr uint32_t kTermFrcStride = 6;

// Each bead's terms are bead_terms[term_start[bead], term_start[bead + 1]),
// (role << 29) | record index, in the order the reference's per-polymer loop
// applies them (the polymer's bonds, then its bond pairs).
constexpr uint32_t kRoleShift = 29;
constexpr uint32_t kRoleBondHead = 0, kRoleBondTail = 1, kRoleAngleHead = 2, kRoleAngleTail = 3,
                   kRoleAngleMid = 4;

// A bead of a stationary type (cleared after each step its world advances).
struct StationaryBead { uint32_t bead; int32_t world; };

// Device buffers of one layout.
struct DeviceState {
    float4 *pos[2];        // x, y, z, hash bits
    float4 *vel[2];        // w = v - f dt/2 (x,y,z), global bead index bits
    float4 *frc[2];        // fx, fy, fz, 0
    uint32_t *key;         // cell id per slot after the drift | inactive bit
    uint32_t *perm;        // sorted position -> old slot | inactive bit
    uint32_t *sorted_idx;  // sorted position -> global bead index
    uint32_t *cell_count;  // n_cells_total: counts; after the scan, the scatter's arrival cursors
    uint32_t *cell_start;  // n_cells_total + 1
    unsigned long long *scan_tiles;   // look-back state per scan tile
    uint32_t *scan_ticket;
    uint32_t *slot_of_bead;    // bonded (and stationary) beads: slot in the buffer a step starts from
    uint32_t *flags;       // context-wide flags (PDPD_FLAG_SCAN_TIMEOUT)
    uint32_t *wflags;      // per world: PDPD_FLAG_* raised by its beads, bonds and forces
    WorldDev *worlds;
    WorldCtr *ctr;
    BeadBlock *bead_blocks;    // kBeadThreads beads per block
    BeadBlock *force_blocks;   // kForceBeads beads per block
    uint32_t *term_start;      // n_beads + 1, by global bead index
    uint32_t *bead_terms;
    TermRec *terms;
    double *term_frc;          // kTermFrcStride per term, written during k_scatter each step
    StationaryBead *stationary_beads;
    // Spatially varying strengths: per world cell (local cell id) and ordered
    // type pair (home type * n_types + other type), (a, gamma) in FP32. A pair
    // uses the mean of its two cells' values.
    const float2 *cell_par;
};

// The buffers the pair-force kernel writes when it integrates the next step
// of a chain: for a step sorting from buffer c with counter copy cp, buffer c
// and counter copy cp ^ 1. One entry per (c, cp), in device memory.
struct NextBuffers {
    float4 *pos, *vel, *frc;
    uint32_t *key, *cell_count, *flags, *wflags;
    unsigned long long *scan_tiles;
    uint32_t *scan_ticket;
    const WorldCtr *ctr;
    uint32_t n_tiles;
};

}  // namespace pdpd
