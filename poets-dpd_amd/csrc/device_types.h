// device_types.h -- layouts shared by the gfx950 kernels and the host side of
// libpdpd_hip.so. All device state lives in HBM as SoA-of-float4, sorted by cell.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pdpd {

constexpr int kMaxTypes = 16;          // 4-bit type field of the bead hash
constexpr int kForceThreads = 256;     // pair-force workgroup (4 wave64)
constexpr int kMaxHaloCells = 512;     // (Bx+2)(By+2)(Bz+2) of one brick
constexpr int kMaxInnerCells = 256;    // Bx*By*Bz of one brick
constexpr int kTileBeads = 1280;       // beads of one brick's halo staged in LDS (u16 indices)
constexpr int kPairList = 16;          // accepted-pair list per lane in the pair-force kernel
constexpr int kScanThreads = 256;
constexpr int kScanPerThread = 8;
constexpr int kScanBlock = kScanThreads * kScanPerThread;   // cells per scan block
constexpr int kBeadThreads = 256;

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
    int32_t brick[3];      // brick shape of the pair-force work items
    uint32_t seed_lo, seed_hi;
    uint32_t n_types;
    uint32_t pad0;
    float a[kMaxTypes * kMaxTypes];    // conservative a_ij
    float sg[kMaxTypes * kMaxTypes];   // sqrt(gamma_ij)
};

// Mutable per-world counters (kept apart from WorldDev so the table stays read-only).
struct WorldCtr {
    uint32_t t;            // step counter used for the RNG (WorldState::t truncated to u32)
    uint32_t remaining;    // steps this world may still advance in the current run
};

// One pair-force work item: a brick of cells of one world.
struct BrickItem {
    int32_t world;
    int32_t bx, by, bz;    // brick origin in cells
};

// One bead-parallel block: [begin, end) slots of one world.
struct BeadBlock {
    int32_t world;
    uint32_t begin, end;
    uint32_t pad;
};

// Bonded topology, flattened per polymer in the reference's order
// (bonds of the polymer type, then its bond pairs).
struct PolyItem {
    int32_t world;
    uint32_t bond_begin, bond_end;
    uint32_t angle_begin, angle_end;
    uint32_t pad[3];
};
struct BondRec { uint32_t head, tail; float kappa, r0; };               // global bead indices
struct AngleRec { uint32_t head, mid, tail, pad; float kappa, cos0, sin0, pad1; };

// Device buffers of one layout.
struct DeviceState {
    float4 *pos[2];        // x, y, z, hash bits
    float4 *vel[2];        // w = v - f dt/2 (x,y,z), global bead index bits
    float4 *frc[2];        // fx, fy, fz, 0
    uint32_t *key;         // cell id per slot after the drift
    uint32_t *rank;        // arrival rank inside the cell | inactive bit
    uint32_t *perm;        // sorted position -> old slot | inactive bit
    uint32_t *cell_count;  // n_cells_total
    uint32_t *cell_start;  // n_cells_total + 1
    uint32_t *block_sums;  // scan partials
    uint32_t *slot_of_bead;
    uint32_t *flags;
    WorldDev *worlds;
    WorldCtr *ctr;
    BrickItem *bricks;
    BeadBlock *bead_blocks;
    PolyItem *polys;
    BondRec *bonds;
    AngleRec *angles;
};

}  // namespace pdpd
