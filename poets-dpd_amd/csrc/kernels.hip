// kernels.hip -- gfx950 kernels of one DPD time step.
//
// Step (one set of launches advances every resident world):
//   k_integrate   finish-kick + drift + half-kick + wrap, cell key, cell counts
//                 (dpd_maths_core_half_step_raw.hpp:42-72,114-124 fused as
//                 dpd_engine_avx2_half_merge_tbb.hpp:780-800 does); only the
//                 first step of a chain: later steps are integrated by the
//                 pair-force kernel of the step before them
//   k_scan        single-pass exclusive scan of the cell counts (look-back)
//   k_scatter     arrival position of every bead in its cell; advances the
//                 per-world step counters; extra workgroups evaluate the bonded
//                 terms (one thread per Hookean bond / angle, FP64; terms of
//                 naive_dpd_engine_half_merge.hpp:638-690 /
//                 dpd_maths_core_half_step.hpp:178-231), once per term
//   k_reorder     stable gather into the other buffer, bead-index order in a cell
//   k_force       per workgroup 64 consecutive beads: three waves for the pair
//                 forces, one per x-plane of the 27-cell stencil, neighbours read
//                 straight from the cell-sorted arrays; register accumulation, no
//                 atomics (dpd_maths_core_half_step_raw.hpp:129-216; RNG of
//                 hash.hpp:61-88 + dpd_maths_core.hpp:55-79); the home beads'
//                 bonded terms summed in FP64 in the per-polymer order of
//                 naive_dpd_engine_half_merge.hpp:376-385 (by a fourth wave or
//                 after the middle plane's pairs); within a chain, then the
//                 next step's integration of its beads
//   k_stationary  (only with stationary bead types) v = 0 after the pair forces
// The final half-kick v += f dt/2 of a step is folded into the next step's
// integration (the engine stores w = v - f dt/2).
//
// Arithmetic is FP32, built with -ffp-contract=off: every fused multiply-add is
// an explicit fmaf, sqrtf and '/' are IEEE correctly rounded. The CPU oracle
// performs the same operations in the same order.
#include "device_types.h"
#include "kernels.h"
#include "pdpd_hip.h"
#include "step_device.h"

#include <algorithm>

namespace pdpd {

__global__ void k_debug_rng(uint32_t n, const uint64_t *__restrict__ th, const uint32_t *__restrict__ a,
                            const uint32_t *__restrict__ b, uint32_t *__restrict__ u32, float *__restrict__ f32)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t t = th[i];
    uint32_t r = rng_sym((uint32_t)t, (uint32_t)(t >> 32), a[i], b[i]);
    u32[i] = r;
    f32[i] = rng_uniform(r);
}

__device__ __forceinline__ void reset_scan_state(const DeviceState &S, uint32_t n_tiles, int tid, int nthreads)
{
    for (uint32_t i = (uint32_t)tid; i < n_tiles; i += (uint32_t)nthreads) S.scan_tiles[i] = 0ull;
    if (tid == 0) *S.scan_ticket = 0u;
}

// The first step of a chain of steps: integration in place in buffer cur from
// the forces of the previous step (later steps of a chain are integrated by
// the pair-force kernel of the step before them).
__global__ __launch_bounds__(kBeadThreads) void k_integrate(DeviceState S, int cur, uint32_t n_tiles, uint32_t n_blocks)
{
    if (blockIdx.x == 0) reset_scan_state(S, n_tiles, threadIdx.x, blockDim.x);   // for this step's scan
    if (blockIdx.x >= n_blocks) return;
    const BeadBlock blk = S.bead_blocks[blockIdx.x];
    const uint32_t s = blk.begin + threadIdx.x;
    if (s >= blk.end) return;
    const WorldDev *__restrict__ W = &S.worlds[blk.world];
    const bool active = S.ctr[blk.world].remaining > 0;   // S.ctr points at this step's counters
    float4 p = S.pos[cur][s];
    uint32_t fl = 0;
    if (active) {
        float4 v = S.vel[cur][s];
        fl = drift_bead(W, p, v, S.frc[cur][s]);
        S.pos[cur][s] = p;
        S.vel[cur][s] = v;
    }
    // lanes past the block's end have returned
    count_bead(&S.wflags[blk.world], fl, S.cell_count, S.key, W, s, p, active, true);
}

// --------------------------------------------------------------- scan ----

// Single-pass exclusive scan of the cell counts with decoupled look-back: tiles
// are taken in ticket order, so every tile a block waits on belongs to a block
// that has already started and publishes its aggregate without waiting. The
// spin is bounded; running out raises PDPD_FLAG_SCAN_TIMEOUT.
__global__ __launch_bounds__(kScanThreads) void k_scan(uint32_t *__restrict__ count, uint32_t n,
                                                      uint32_t *__restrict__ cell_start,
                                                      unsigned long long *__restrict__ tiles,
                                                      uint32_t *__restrict__ ticket, uint32_t n_tiles,
                                                      uint32_t *__restrict__ flags)
{
    __shared__ uint32_t s_wsum[kScanThreads / 64];
    __shared__ uint32_t s_tile, s_prefix;
    if (threadIdx.x == 0) s_tile = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint32_t tile = s_tile;
    const uint32_t base = tile * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        const uint32_t i = base + k;
        v[k] = (i < n) ? count[i] : 0u;
        acc += v[k];
    }
    uint32_t agg;
    uint32_t ex = block_scan<kScanThreads / 64>(acc, s_wsum, &agg);
    if (threadIdx.x < 64) {
        // look-back by the first wave: lane l reads tile (base - l); the window
        // moves back 64 tiles at a time until it holds an inclusive prefix (tile
        // -1 counts as inclusive 0)
        const int lane = threadIdx.x;
        uint32_t prefix = 0;
        if (tile == 0) {
            if (lane == 0) __hip_atomic_store(&tiles[0], kTileIncl | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            if (lane == 0) __hip_atomic_store(&tiles[tile], kTileAgg | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int base = (int)tile - 1;
            uint32_t spins = 0;
            for (;;) {
                const int j = base - lane;
                const unsigned long long st =
                    j >= 0 ? __hip_atomic_load(&tiles[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) : kTileIncl;
                const unsigned long long status = st & 0xffffffff00000000ull;
                const unsigned long long incl = __ballot(status == kTileIncl);
                const unsigned long long none = __ballot(status == 0ull);
                const int first = incl ? __ffsll((long long)incl) - 1 : 64;
                const unsigned long long upto = first >= 63 ? ~0ull : ((2ull << first) - 1ull);
                if (none & upto) {                          // a tile in the window has not published yet
                    if (++spins > kScanSpinLimit) {
                        if (lane == 0) atomicOr(flags, PDPD_FLAG_SCAN_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                uint32_t val = lane <= first ? (uint32_t)st : 0u;
#pragma unroll
                for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, 64);
                prefix += val;
                if (first < 64) break;
                base -= 64;
            }
            if (lane == 0)
                __hip_atomic_store(&tiles[tile], kTileIncl | (prefix + agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        if (lane == 0) {
            s_prefix = prefix;
            if (tile == n_tiles - 1) cell_start[n] = prefix + agg;
        }
    }
    __syncthreads();
    ex += s_prefix;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        const uint32_t i = base + k;
        if (i < n) {
            cell_start[i] = ex;
            count[i] = ex;          // the scatter's arrival cursor (k_reorder clears it)
        }
        ex += v[k];
    }
}

// Large layouts (more tiles than StepLaunch::lookback_tiles): reduce each tile,
// scan the tile sums in one workgroup, then rescan each tile from its offset.
// Three launches, no inter-workgroup waiting.
constexpr int kTopThreads = 1024, kTopPerThread = 4;

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ count, uint32_t n,
                                                             uint32_t *__restrict__ tile_sum)
{
    __shared__ uint32_t s_wsum[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        const uint32_t i = base + k;
        acc += (i < n) ? count[i] : 0u;
    }
    uint32_t tot;
    (void)block_scan<kScanThreads / 64>(acc, s_wsum, &tot);
    if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(kTopThreads) void k_scan_top(uint32_t *__restrict__ tile_sum, uint32_t n_tiles,
                                                         uint32_t *__restrict__ cell_start, uint32_t n)
{
    __shared__ uint32_t s_wsum[kTopThreads / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < n_tiles; base += kTopThreads * kTopPerThread) {   // uniform loop
        const uint32_t i0 = base + threadIdx.x * kTopPerThread;
        uint32_t v[kTopPerThread], acc = 0;
#pragma unroll
        for (int k = 0; k < kTopPerThread; k++) {
            v[k] = (i0 + k < n_tiles) ? tile_sum[i0 + k] : 0u;
            acc += v[k];
        }
        uint32_t tot;
        uint32_t ex = block_scan<kTopThreads / 64>(acc, s_wsum, &tot) + carry;
#pragma unroll
        for (int k = 0; k < kTopPerThread; k++) {
            if (i0 + k < n_tiles) tile_sum[i0 + k] = ex;
            ex += v[k];
        }
        carry += tot;
    }
    if (threadIdx.x == 0) cell_start[n] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t *__restrict__ count, uint32_t n,
                                                           uint32_t *__restrict__ cell_start,
                                                           const uint32_t *__restrict__ tile_off)
{
    __shared__ uint32_t s_wsum[kScanThreads / 64];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread], acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        const uint32_t i = base + k;
        v[k] = (i < n) ? count[i] : 0u;
        acc += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_scan<kScanThreads / 64>(acc, s_wsum, &tot) + tile_off[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        const uint32_t i = base + k;
        if (i < n) {
            cell_start[i] = ex;
            count[i] = ex;          // the scatter's arrival cursor (k_reorder clears it)
        }
        ex += v[k];
    }
}

// ------------------------------------------------------- sort scatter ----

// Stable order inside a cell: beads ranked by global bead index, so the order
// (and every force sum) is independent of arrival order and of the batch.
// slot_of_bead is read only for bonded beads (the bonded terms) and stationary
// ones; monomers (BeadHash bit 27) have no bonds, so their entries are not
// written unless stationary beads are resident.
__global__ __launch_bounds__(kBeadThreads) void k_reorder(DeviceState S, int cur, uint32_t n, uint32_t n_cells,
                                                          int all_slots)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n_cells) S.cell_count[j] = 0u;          // ready for the next step's counts
    if (j >= n) return;
    const int nxt = cur ^ 1;
    const uint32_t e = S.perm[j];
    const uint32_t s = e & ~kInactive;
    const uint32_t k = S.key[s] & ~kInactive;
    const uint32_t a = S.cell_start[k], b = S.cell_start[k + 1];
    const uint32_t me = S.sorted_idx[j];
    uint32_t rank = 0;
    for (uint32_t q = a; q < b; q++) rank += S.sorted_idx[q] < me ? 1u : 0u;
    const uint32_t dst = a + rank;
    const float4 p = S.pos[cur][s];
    S.pos[nxt][dst] = p;
    S.vel[nxt][dst] = S.vel[cur][s];
    if (e & kInactive) S.frc[nxt][dst] = S.frc[cur][s];
    if (all_slots || !((fbits(p.w) >> 27) & 1u)) S.slot_of_bead[me] = dst;
}

// -------------------------------------------------------------- bonds ----

__device__ __forceinline__ double min_image(double d, double len)
{
    if (d > 0.5 * len) d -= len;
    else if (d < -0.5 * len) d += len;
    return d;
}

// One bonded term per thread, in FP64 from the FP32 positions (differences of
// floats are exact in double). FP64 keeps the theta0 != 0 angle force, which
// divides by sin(phi), stable near straight chains. Hookean:
// naive_dpd_engine_half_merge.hpp:638-654 (stored: the tail's force
// dx * kappa (r0 - r) / r; the head gets its negation); angle:
// dpd_maths_core_half_step.hpp:178-231 with the prefactor floor of
// naive_dpd_engine.hpp:681-682, the argument clamped at 0 (stored: head, then
// tail; the middle gets -(head + tail)). Terms of frozen worlds are skipped.
// Reads the drifted positions in buffer cur through slot_of_bead, which maps
// each bonded bead to its slot in that buffer (the previous step's sort, or
// the layout an import or relayout produced).
__device__ __forceinline__ void eval_term(const DeviceState &S, int cur, uint32_t k)
{
    const TermRec R = S.terms[k];
    const uint32_t w = R.kind >> kTermWorldShift;
    if (S.ctr[w].remaining == 0) return;
    const float4 *__restrict__ pos = S.pos[cur];
    const double Lx = S.worlds[w].Lf[0], Ly = S.worlds[w].Lf[1], Lz = S.worlds[w].Lf[2];
    double *__restrict__ out = S.term_frc + (size_t)k * kTermFrcStride;
    const float4 pa = pos[S.slot_of_bead[R.a]], pb = pos[S.slot_of_bead[R.b]];
    if ((R.kind & 0xffu) == kTermBond) {
        const double dx = min_image((double)pb.x - (double)pa.x, Lx);
        const double dy = min_image((double)pb.y - (double)pa.y, Ly);
        const double dz = min_image((double)pb.z - (double)pa.z, Lz);
        const double r = sqrt(fma(dz, dz, fma(dy, dy, dx * dx)));
        if (!(r < S.worlds[w].max_bond)) atomicOr(&S.wflags[w], PDPD_FLAG_BOND_STRETCH);
        const double hf = R.kappa * (R.p0 - r);
        const double sc = hf / r;
        out[0] = dx * sc;
        out[1] = dy * sc;
        out[2] = dz * sc;
        return;
    }
    const float4 pc = pos[S.slot_of_bead[R.c]];
    const double d0x = min_image((double)pb.x - (double)pa.x, Lx), d0y = min_image((double)pb.y - (double)pa.y, Ly),
                 d0z = min_image((double)pb.z - (double)pa.z, Lz);
    const double d1x = min_image((double)pc.x - (double)pb.x, Lx), d1y = min_image((double)pc.y - (double)pb.y, Ly),
                 d1z = min_image((double)pc.z - (double)pb.z, Lz);
    const double r01 = sqrt(fma(d0z, d0z, fma(d0y, d0y, d0x * d0x)));
    const double r12 = sqrt(fma(d1z, d1z, fma(d1y, d1y, d1x * d1x)));
    if (!(r01 < S.worlds[w].max_bond && r12 < S.worlds[w].max_bond)) atomicOr(&S.wflags[w], PDPD_FLAG_BOND_STRETCH);
    const double mag = r01 * r12;
    double hx = 0, hy = 0, hz = 0, tx = 0, ty = 0, tz = 0;
    if (mag > 0.0001) {
        const double b1 = r01 * r01, b2 = r12 * r12;
        const double dot = fma(d0z, d1z, fma(d0y, d1y, d0x * d1x));
        const double o1 = dot / b1, o2 = dot / b2;
        const double cos2 = o1 * o2;
        double fm = R.kappa / mag;
        if (fabs(dot) > 0.000001 && R.p1 > 0.0) {
            const double pref = fmax(sqrt(fmax(1.0 / cos2 - 1.0, 0.0)), 0.000001);
            fm = fm * (R.p0 - R.p1 * (1.0 / pref));
        }
        hx = (d0x * o1 - d1x) * fm; hy = (d0y * o1 - d1y) * fm; hz = (d0z * o1 - d1z) * fm;
        tx = (d0x - d1x * o2) * fm; ty = (d0y - d1y * o2) * fm; tz = (d0z - d1z * o2) * fm;
    }
    out[0] = hx; out[1] = hy; out[2] = hz;
    out[3] = tx; out[4] = ty; out[5] = tz;
}

// Sort scatter: the sorted position of every bead's slot (n_scatter_blocks
// workgroups), then the bonded terms (the remaining workgroups). Also advances
// the step counters: this step reads S.ctr; the next one reads ctr_next, which
// nothing of this step reads before the pair-force kernel.
__global__ __launch_bounds__(kBeadThreads) void k_scatter(DeviceState S, int cur, uint32_t n,
                                                          uint32_t n_scatter_blocks, uint32_t n_terms,
                                                          WorldCtr *__restrict__ ctr_next, uint32_t n_worlds)
{
    {
        const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i < n_worlds) {
            const WorldCtr c = S.ctr[i];
            WorldCtr o = c;
            if (c.remaining > 0) { o.t = c.t + 1; o.remaining = c.remaining - 1; }
            ctr_next[i] = o;
        }
    }
    if (blockIdx.x >= n_scatter_blocks) {
        const uint32_t k = (blockIdx.x - n_scatter_blocks) * blockDim.x + threadIdx.x;
        if (k < n_terms) eval_term(S, cur, k);
        return;
    }
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    // arrival position in the cell: each run of equal keys in consecutive
    // lanes reserves its positions with one atomic on the cell's cursor
    // (cell_count holds the cell's start after the scan)
    const uint32_t k = S.key[s];
    const uint32_t key = k & ~kInactive;
    int start, next;
    bool head;
    key_runs(key, true, start, next, head);                // lanes past the end have returned
    uint32_t base = 0;
    if (head) base = atomicAdd(&S.cell_count[key], (uint32_t)(next - (int)(threadIdx.x & 63)));
    const uint32_t j = __shfl(base, start, 64) + (uint32_t)((int)(threadIdx.x & 63) - start);
    S.perm[j] = s | (k & kInactive);
    S.sorted_idx[j] = fbits(S.vel[cur][s].w);
}

// --------------------------------------------------------- pair force ----

// Shift codes of a candidate's periodic image, 2 bits per axis: 0 none,
// 1 = +L, 2 = -L. Stored with the candidate's slot in the per-lane list.
__device__ __forceinline__ float shift_of(uint32_t code, float L)
{
    return code == 1u ? L : (code == 2u ? -L : 0.0f);
}

constexpr int kQueue = 256;       // accepted pairs buffered per pair wave before they are evaluated (at most)
constexpr int kLaneList = 16;     // accepted pairs buffered per home bead (queue positions, 8 bits)

// Slots are below 2^26, so the byte offset of a float4 fits 32 bits: the loads
// take a scalar base and a 32-bit vector offset.
__device__ __forceinline__ const float4 &at_slot(const float4 *__restrict__ p, uint32_t i)
{
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(p) + (i << 4));
}

__device__ __forceinline__ float3 load_xyz(const float4 *__restrict__ p, uint32_t i)
{
    const float *q = reinterpret_cast<const float *>(reinterpret_cast<const char *>(p) + (i << 4));
    return make_float3(q[0], q[1], q[2]);
}

// The bonded terms of the bead in slot s, summed in FP64 into s_bond /
// s_nterms (lane = the bead's lane in its workgroup).
__device__ __forceinline__ void bonded_terms(const DeviceState &S, const float4 *__restrict__ vel, uint32_t s,
                                             bool moving, int lane, double (*s_bond)[kForceBeads],
                                             uint32_t *s_nterms)
{
    double bx = 0.0, by = 0.0, bz = 0.0;
    const uint32_t nt = moving ? bonded_sum(S, fbits(vel[s].w), bx, by, bz) : 0u;
    s_bond[0][lane] = bx;
    s_bond[1][lane] = by;
    s_bond[2][lane] = bz;
    s_nterms[lane] = nt;
}

// Pair forces. A workgroup takes 64 consecutive slots of one world (one home
// bead per lane) and has one wave per x-plane of the 27-cell stencil. A wave
// scans the three rows (dy = -1, 0, +1) of its plane in turn; per row the
// candidates come in z order -1, 0, +1 (an image cell below 0, the contiguous
// run of cells, an image cell above L-1) and, within a cell, in slot order.
// Accepted candidates of all 64 home beads go to one queue per wave; when the
// queue fills, all lanes evaluate queued pairs side by side (64 per round,
// whatever bead they belong to) and store each pair's unit vector and force
// magnitude; each home bead then accumulates its own pairs in acceptance
// order. So every force is summed in one fixed order: per plane one running
// sum over its rows' candidates in order, planes as ((p0 + p1) + p2), independent of
// batch, launch shape and queue flushes. kScanWidth (candidate loads in flight
// per lane) and the occupancy target trade registers for memory-level
// parallelism; every variant gives bit-identical forces.
// kBondWave: a fourth wave sums the bonded terms beside the pair waves;
// otherwise the middle plane's wave sums them after its pairs (three waves per
// workgroup, more workgroups resident). Same sums either way.
// kSpatial: worlds flagged spatial take a pair's strengths as the mean of the
// values at the two beads' cells (naive_dpd_engine.hpp:452-468): a = (a_h +
// a_o) / 2, sqrt(gamma) = sqrtf((g_h + g_o) / 2), index home type * n_types +
// other type; other worlds use the constant tables as without it.
template <int kScanWidth, int kWaves, bool kBondWave, bool kSpatial, int kQ = kQueue>
__global__ __launch_bounds__(kBondWave ? kForceThreads : kForceBeads * kForcePlanes)
__attribute__((amdgpu_waves_per_eu(kWaves))) void k_force(DeviceState S, int nxt, uint32_t n_blocks,
                                                           const NextBuffers *next)
{
    // a batch accepts at most 64 * kScanWidth pairs, which an emptied queue must hold
    static_assert(kQ >= kForceBeads * kScanWidth && kQ <= kQueue, "queue positions are 8-bit");
    static_assert(kLaneList >= kScanWidth, "a lane lists up to kScanWidth pairs per batch");
    constexpr int kThreads = kBondWave ? kForceThreads : kForceBeads * kForcePlanes;
    constexpr int kPairLanes = kForceBeads * kForcePlanes;
    // per pair wave: queued pairs (candidate slot | image code bits 26..31, and
    // home lane) and their evaluated (unit vector, force magnitude)
    __shared__ uint32_t s_qslot[kForcePlanes][kQ];
    __shared__ uint8_t s_qhome[kForcePlanes][kQ];
    __shared__ float4 s_qres[kForcePlanes][kQ];
    // per home bead and plane: its queued pairs in acceptance order
    __shared__ uint8_t s_list[kLaneList][kPairLanes];
    __shared__ float4 s_hp[kForceBeads], s_hv[kForceBeads];
    __shared__ uint32_t s_world;
    __shared__ float s_part[kForcePlanes - 1][3][kForceBeads];
    __shared__ double s_bond[3][kForceBeads];
    __shared__ uint32_t s_nterms[kForceBeads];
    __shared__ float s_a[kMaxTypes * kMaxTypes], s_sg[kMaxTypes * kMaxTypes];

    const int tid = threadIdx.x;
    const int plane = tid / kForceBeads;            // wave-uniform
    const int lane = tid % kForceBeads;
    const unsigned long long below = (1ull << lane) - 1ull;
    const float4 *__restrict__ pos = S.pos[nxt];
    const float4 *__restrict__ vel = S.vel[nxt];
    float4 *__restrict__ frc = S.frc[nxt];
    const uint32_t *__restrict__ cs = S.cell_start;
    if (next && blockIdx.x == 0) {                       // for the next step's scan
        const NextBuffers B = next_buffers(next);
        for (uint32_t i = (uint32_t)tid; i < B.n_tiles; i += kThreads) B.scan_tiles[i] = 0ull;
        if (tid == 0) *B.scan_ticket = 0u;
    }

    for (uint32_t li = blockIdx.x; li < n_blocks; li += gridDim.x) {
        const BeadBlock blk = S.force_blocks[xcd_chunk(li, n_blocks)];
        if (S.ctr[blk.world].remaining == 0) continue;       // uniform: frozen world (never in a chain)
        const WorldDev *__restrict__ W = &S.worlds[blk.world];
        {   // the world's n_types^2 strengths (rows of kMaxTypes)
            const int nt = (int)W->n_types;
            for (int i = tid; i < nt * nt; i += kThreads) {
                const int k = (i / nt) * kMaxTypes + i % nt;
                s_a[k] = W->a[k];
                s_sg[k] = W->sg[k];
            }
        }
        const uint32_t s = blk.begin + lane;
        const bool live = s < blk.end;
        const uint32_t sl = live ? s : blk.begin;            // a valid slot for every lane's loads
        const float4 hp = pos[sl];
        const uint32_t hh = fbits(hp.w);
        // a stationary home bead ends the step with f = 0 (naive_dpd_engine_half_merge.hpp:387-390);
        // its pairs with moving beads are evaluated from the moving side
        const bool moving = live && !((W->stationary >> (hh >> 28)) & 1u);
        if (plane == 0) {
            s_hp[lane] = hp;
            s_hv[lane] = vel[sl];
            if (lane == 0) s_world = blk.world;
        }
        __syncthreads();
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        if (kBondWave && plane == kForcePlanes) {
            bonded_terms(S, vel, s, moving, lane, s_bond, s_nterms);
        } else {
            const WorldCtr wc = S.ctr[blk.world];
            const uint64_t th = step_hash(wc.t, W->seed_lo, W->seed_hi);
            const uint32_t th_lo = (uint32_t)th, th_hi = (uint32_t)(th >> 32);
            const float scale = W->rng_scale;
            const int L0 = W->L[0], L1 = W->L[1], L2 = W->L[2];
            const float Lx = W->Lf[0], Ly = W->Lf[1], Lz = W->Lf[2];
            const bool spatial = kSpatial && W->spatial;
            const int cx = cell_coord(hp.x, L0), cy = cell_coord(hp.y, L1), cz = cell_coord(hp.z, L2);

            // the plane's three rows; per row a stream of at most two slot ranges,
            // [ra, ra + na) then [rb, rb + nb)
            int ox = cx + plane - 1;
            uint32_t codex = 0;
            if (ox < 0) { ox += L0; codex = 2u; } else if (ox >= L0) { ox -= L0; codex = 1u; }
            const float sx = shift_of(codex, Lx);
            const bool pre = cz == 0, post = cz == L2 - 1;
            const int zi = pre ? L2 - 1 : 0;                 // the image cell, if any
            const float szA = pre ? -Lz : 0.0f, szB = post ? Lz : 0.0f;
            const uint32_t zcA = pre ? (2u << 30) : 0u, zcB = post ? (1u << 30) : 0u;
            // Cells wholly beyond the cutoff are skipped: the gap from the home
            // bead to a neighbour cell is the distance to the shared face, edge
            // or corner (fractions in the home cell are exact in FP32). A cell is
            // skipped only when its squared gap exceeds 1 by a margin far above
            // the rounding of the pair test, so no candidate the test would
            // accept is lost and the forces are unchanged.
            constexpr float kReach = 1.0f + 1e-5f;
            const float fxr = hp.x - (float)cx, fyr = hp.y - (float)cy, fzr = hp.z - (float)cz;
            const float gx = plane == 0 ? fxr : (plane == 2 ? 1.0f - fxr : 0.0f);
            const float gx2 = gx * gx, glo = fzr * fzr, ghi = (1.0f - fzr) * (1.0f - fzr);
            uint32_t ra[3], na[3], rb[3], nb[3], cxy[3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                int oy = cy + r - 1;
                uint32_t codey = 0;
                if (oy < 0) { oy += L1; codey = 2u; } else if (oy >= L1) { oy -= L1; codey = 1u; }
                cxy[r] = (codex << 26) | (codey << 28);
                const float gy = r == 0 ? fyr : (r == 2 ? 1.0f - fyr : 0.0f);
                const float g = fmaf(gy, gy, gx2);
                const bool rowon = moving && g < kReach, loon = rowon && g + glo < kReach,
                           hion = rowon && g + ghi < kReach;
                const int m0 = pre ? 0 : (loon ? cz - 1 : cz);    // the contiguous run of cells
                const int m1 = post ? cz : (hion ? cz + 1 : cz);
                const bool img = pre ? loon : (post && hion);    // the image cell taken
                const uint32_t row = W->cell_off + (uint32_t)((ox * L1 + oy) * L2);
                const uint32_t mb = cs[row + m0];
                const uint32_t me = rowon ? cs[row + m1 + 1] : mb;
                const uint32_t ib = img ? cs[row + zi] : 0u;
                const uint32_t ie = img ? cs[row + zi + 1] : 0u;
                ra[r] = pre ? ib : mb;
                na[r] = pre ? ie - ib : me - mb;
                rb[r] = pre ? mb : ib;
                nb[r] = pre ? me - mb : ie - ib;
            }

            uint32_t *__restrict__ qslot = s_qslot[plane];
            uint8_t *__restrict__ qhome = s_qhome[plane];
            float4 *__restrict__ qres = s_qres[plane];
            float px = 0.0f, py = 0.0f, pz = 0.0f;  // this plane's sum
            uint32_t listed = 0;                // this lane's queued pairs (its list length)
            uint32_t queued = 0;                // queue fill, the same in every lane

            // Evaluate every queued pair (dpd_maths_core_half_step_raw.hpp:129-216
            // with the terms grouped as a wr + sgw (scale u - sgw rdotv)), then add
            // each lane's own pairs to its plane sum in acceptance order.
            auto flush = [&]() {
                wave_sync();
                for (uint32_t i = (uint32_t)lane; i < queued; i += kForceBeads) {
                    const uint32_t e = qslot[i];
                    const uint32_t hl = qhome[i];
                    const float4 op = at_slot(pos, e & kSlotMask), ov = at_slot(vel, e & kSlotMask);
                    const float4 p = s_hp[hl], v = s_hv[hl];
                    const float ox_ = op.x + shift_of((e >> 26) & 3u, Lx);
                    const float oy_ = op.y + shift_of((e >> 28) & 3u, Ly);
                    const float oz_ = op.z + shift_of((e >> 30) & 3u, Lz);
                    const float dx = p.x - ox_, dy = p.y - oy_, dz = p.z - oz_;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float rr = sqrtf(r2);
                    const float inv_r = 1.0f / rr;
                    const float wr = 1.0f - rr;
                    const float ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
                    const float dvx = v.x - ov.x, dvy = v.y - ov.y, dvz = v.z - ov.z;
                    const float rdotv = fmaf(ez, dvz, fmaf(ey, dvy, ex * dvx));
                    const uint32_t ph = fbits(p.w), oh = fbits(op.w);
                    float a, sg;
                    if (kSpatial && spatial) {
                        const uint32_t nt = W->n_types, ti = (ph >> 28) * nt + (oh >> 28);
                        const uint32_t hc = (uint32_t)((cell_coord(p.x, L0) * L1 + cell_coord(p.y, L1)) * L2 +
                                                       cell_coord(p.z, L2));
                        const uint32_t oc = (uint32_t)((cell_coord(op.x, L0) * L1 + cell_coord(op.y, L1)) * L2 +
                                                       cell_coord(op.z, L2));
                        const float2 ph_ = S.cell_par[W->cell_par_off + hc * nt * nt + ti];
                        const float2 po_ = S.cell_par[W->cell_par_off + oc * nt * nt + ti];
                        a = (ph_.x + po_.x) * 0.5f;
                        sg = sqrtf((ph_.y + po_.y) * 0.5f);
                    } else {
                        const uint32_t ti = (ph >> 28) * kMaxTypes + (oh >> 28);
                        a = s_a[ti];
                        sg = s_sg[ti];
                    }
                    const float sgw = sg * wr;
                    const float u = rng_uniform(rng_sym(th_lo, th_hi, ph, oh));
                    const float F = fmaf(a, wr, sgw * fmaf(-sgw, rdotv, scale * u));
                    qres[i] = make_float4(ex, ey, ez, F);
                }
                wave_sync();
                for (uint32_t c = 0; c < listed; c++) {
                    const float4 q = qres[s_list[c][tid]];
                    px = fmaf(q.x, q.w, px);
                    py = fmaf(q.y, q.w, py);
                    pz = fmaf(q.z, q.w, pz);
                }
                queued = 0;
                listed = 0;
                wave_sync();
            };

            // Rows in turn, kScanWidth candidates per lane at a time. The queue is
            // flushed when the next batch would overflow it or a lane's list, and
            // at the end.
#pragma unroll
            for (int r = 0; r < 3; r++) {
                const uint32_t len = na[r] + nb[r];
                const uint32_t wlen = wave_max(len);
                const uint32_t rbm = rb[r] - na[r];              // slot of stream position q >= na: rbm + q
                const float sy = shift_of((cxy[r] >> 28) & 3u, Ly);
                for (uint32_t k = 0; k < wlen; k += kScanWidth) {
                    uint32_t slot[kScanWidth];
                    float3 pc[kScanWidth];
#pragma unroll
                    for (int j = 0; j < kScanWidth; j++) {
                        const uint32_t q = k + j;
                        // beyond the row's stream: the home slot, never accepted (r2 = 0)
                        const uint32_t inrow = q < na[r] ? ra[r] + q : rbm + q;
                        slot[j] = q < len ? inrow : sl;
                        pc[j] = load_xyz(pos, slot[j]);
                    }
                    bool ok[kScanWidth];
                    unsigned long long bal[kScanWidth];
                    uint32_t total = 0;
#pragma unroll
                    for (int j = 0; j < kScanWidth; j++) {
                        const float ozs = k + j < na[r] ? szA : szB;
                        const float dx = hp.x - (pc[j].x + sx), dy = hp.y - (pc[j].y + sy),
                                    dz = hp.z - (pc[j].z + ozs);
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        ok[j] = r2 < 1.0f && r2 >= kMinR2;
                        bal[j] = __builtin_amdgcn_ballot_w64(ok[j]);
                        total += (uint32_t)__popcll(bal[j]);
                    }
                    if (total == 0) continue;
                    // room for this batch: a lane may list up to kScanWidth more pairs
                    if (queued + total > (uint32_t)kQ ||
                        __builtin_amdgcn_ballot_w64(listed + kScanWidth > (uint32_t)kLaneList))
                        flush();
                    uint32_t at = queued;
#pragma unroll
                    for (int j = 0; j < kScanWidth; j++) {
                        if (ok[j]) {
                            const uint32_t qp = at + (uint32_t)__popcll(bal[j] & below);
                            qslot[qp] = slot[j] | cxy[r] | (k + j < na[r] ? zcA : zcB);
                            qhome[qp] = (uint8_t)lane;
                            s_list[listed++][tid] = (uint8_t)qp;
                        }
                        at += (uint32_t)__popcll(bal[j]);
                    }
                    queued = at;
                }
            }
            if (queued) flush();
            fx = px;
            fy = py;
            fz = pz;
        }
        if (!kBondWave && plane == 1) bonded_terms(S, vel, s, moving, lane, s_bond, s_nterms);
        // combine the three planes in order: (p0 + p1) + p2
        if (plane > 0 && plane < kForcePlanes) {
            s_part[plane - 1][0][lane] = fx;
            s_part[plane - 1][1][lane] = fy;
            s_part[plane - 1][2][lane] = fz;
        }
        __syncthreads();
        if (plane == 0 && live) {
            fx = (fx + s_part[0][0][lane]) + s_part[1][0][lane];
            fy = (fy + s_part[0][1][lane]) + s_part[1][1][lane];
            fz = (fz + s_part[0][2][lane]) + s_part[1][2][lane];
            if (s_nterms[lane]) {      // pair force + bonded terms, rounded once
                fx = (float)((double)fx + s_bond[0][lane]);
                fy = (float)((double)fy + s_bond[1][lane]);
                fz = (float)((double)fz + s_bond[2][lane]);
            }
            frc[s] = make_float4(fx, fy, fz, 0.0f);
            if (const uint32_t ff = force_flags(W, fx, fy, fz)) atomicOr(&S.wflags[blk.world], ff);
        }
        __syncthreads();
        if (next && plane == 0) {          // the whole wave, converged; f as stored above by this lane
            const NextBuffers B = next_buffers(next);          // (after its compiler barrier: reloaded here)
            const uint32_t s2 = blk.begin + lane;
            advance_next(B, &S.worlds[s_world], s_world, s2, live, s_hp[lane], s_hv[lane],
                         live ? frc[s2] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        }
    }
}

// Force trace (debugging): every pair the pair-force kernel evaluates for one
// world, recomputed from the buffers the step left (x(t+dt) and the half-step
// velocities the step used) by one thread per home bead over its 27 cells, with
// the acceptance test and the arithmetic of k_force, plus the per-term
// quantities the reference logs. Records are appended in any order.
__global__ __launch_bounds__(kBeadThreads) void k_trace_pairs(DeviceState S, int buf, uint32_t w, uint32_t t,
                                                              pdpd_pair_record *__restrict__ out, uint32_t cap,
                                                              uint32_t *__restrict__ count)
{
    const WorldDev *__restrict__ W = &S.worlds[w];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= W->n_beads) return;
    const uint32_t s = W->bead_off + i;
    const float4 *__restrict__ pos = S.pos[buf];
    const float4 *__restrict__ vel = S.vel[buf];
    const float4 hp = pos[s], hv = vel[s];
    const uint32_t hh = fbits(hp.w);
    if ((W->stationary >> (hh >> 28)) & 1u) return;
    const uint64_t th = step_hash(t, W->seed_lo, W->seed_hi);
    const uint32_t th_lo = (uint32_t)th, th_hi = (uint32_t)(th >> 32);
    const float scale = W->rng_scale;
    const int L[3] = {W->L[0], W->L[1], W->L[2]};
    const int c[3] = {cell_coord(hp.x, L[0]), cell_coord(hp.y, L[1]), cell_coord(hp.z, L[2])};
    for (int ox = -1; ox <= 1; ox++)
        for (int oy = -1; oy <= 1; oy++)
            for (int oz = -1; oz <= 1; oz++) {
                const int o[3] = {ox, oy, oz};
                int n[3];
                float sh[3];
                for (int d = 0; d < 3; d++) {
                    n[d] = c[d] + o[d];
                    uint32_t code = 0;
                    if (n[d] < 0) { n[d] += L[d]; code = 2u; } else if (n[d] >= L[d]) { n[d] -= L[d]; code = 1u; }
                    sh[d] = shift_of(code, W->Lf[d]);
                }
                const uint32_t cell = W->cell_off + (uint32_t)((n[0] * L[1] + n[1]) * L[2] + n[2]);
                for (uint32_t j = S.cell_start[cell]; j < S.cell_start[cell + 1]; j++) {
                    const float4 op = pos[j];
                    const float dx = hp.x - (op.x + sh[0]), dy = hp.y - (op.y + sh[1]), dz = hp.z - (op.z + sh[2]);
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    if (!(r2 < 1.0f && r2 >= kMinR2)) continue;
                    const float4 ov = vel[j];
                    const uint32_t oh = fbits(op.w);
                    const float rr = sqrtf(r2);
                    const float inv_r = 1.0f / rr;
                    const float wr = 1.0f - rr;
                    const float ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
                    const float dvx = hv.x - ov.x, dvy = hv.y - ov.y, dvz = hv.z - ov.z;
                    const float rdotv = fmaf(ez, dvz, fmaf(ey, dvy, ex * dvx));
                    float a, sg, gamma;
                    if (W->spatial) {
                        const uint32_t nt = W->n_types, ti = (hh >> 28) * nt + (oh >> 28);
                        const uint32_t hc = (uint32_t)((c[0] * L[1] + c[1]) * L[2] + c[2]);
                        const uint32_t oc = (uint32_t)((n[0] * L[1] + n[1]) * L[2] + n[2]);
                        const float2 ph_ = S.cell_par[W->cell_par_off + hc * nt * nt + ti];
                        const float2 po_ = S.cell_par[W->cell_par_off + oc * nt * nt + ti];
                        a = (ph_.x + po_.x) * 0.5f;
                        gamma = (ph_.y + po_.y) * 0.5f;
                        sg = sqrtf(gamma);
                    } else {
                        const uint32_t ti = (hh >> 28) * kMaxTypes + (oh >> 28);
                        a = W->a[ti];
                        sg = W->sg[ti];
                        gamma = sg * sg;
                    }
                    const float sgw = sg * wr;
                    const float u = rng_uniform(rng_sym(th_lo, th_hi, hh, oh));
                    const float F = fmaf(a, wr, sgw * fmaf(-sgw, rdotv, scale * u));
                    pdpd_pair_record r;
                    r.home = hh;
                    r.other = oh;
                    r.dx[0] = dx; r.dx[1] = dy; r.dx[2] = dz;
                    r.dr = rr;
                    r.dv[0] = dvx; r.dv[1] = dvy; r.dv[2] = dvz;
                    r.diss_strength = gamma;
                    r.inv_root_dt = scale;
                    r.gammap = sgw * sgw;
                    r.rng = u;
                    r.con = a * wr;
                    r.diss = -(sgw * sgw) * rdotv;
                    r.rng_scale = sgw * scale;
                    r.rand = r.rng_scale * u;
                    r.f[0] = ex * F; r.f[1] = ey * F; r.f[2] = ez * F;
                    const uint32_t k = atomicAdd(count, 1u);
                    if (k < cap) out[k] = r;
                }
            }
}

// After the pair forces have read them: beads of stationary types end the step
// with v = 0 (naive_dpd_engine_half_merge.hpp:387-390). Runs only when such
// beads are resident.
__global__ __launch_bounds__(kBeadThreads) void k_stationary(DeviceState S, int nxt, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const StationaryBead e = S.stationary_beads[i];
    if (S.ctr[e.world].remaining == 0) return;       // S.ctr points at this step's counters
    const uint32_t slot = S.slot_of_bead[e.bead];
    const float4 v = S.vel[nxt][slot];
    S.vel[nxt][slot] = make_float4(0.0f, 0.0f, 0.0f, v.w);
}

// -------------------------------------------------------- observables ----

// Per-world sums in FP64 from the FP32 state (the reference's checks:
// momentum sum_i v_i and sum_i v_i.v_i of test_ordered_mesh.hpp:85-93, the
// per-step velocity statistics of stats_dpd_engine_half_step.hpp:539-572),
// plus the conservative virial sum over unordered pairs closer than the
// cutoff of r_ij . F^C_ij = a_ij (1 - r) r. v is the exported velocity
// w + f dt/2 (FP64, as the export computes it); r is the nearest-image
// distance of the FP32 positions in FP64 (exact differences), a_ij the world's
// FP32 strength. Each bead block (256 beads of one world) writes its partial
// sums to partial[block]; the world's blocks are then reduced in a fixed tree,
// so the result depends only on the state and the block layout.
constexpr int kObsThreads = kBeadThreads;

__device__ __forceinline__ void block_sum_f64(double *v, double (*s_red)[kObsThreads], int tid)
{
#pragma unroll
    for (int k = 0; k < kObsValues; k++) s_red[k][tid] = v[k];
    __syncthreads();
    for (int o = kObsThreads / 2; o > 0; o >>= 1) {
        if (tid < o) {
#pragma unroll
            for (int k = 0; k < kObsValues; k++) s_red[k][tid] += s_red[k][tid + o];
        }
        __syncthreads();
    }
#pragma unroll
    for (int k = 0; k < kObsValues; k++) v[k] = s_red[k][0];
    __syncthreads();
}

__global__ __launch_bounds__(kObsThreads) void k_obs_partial(DeviceState S, int buf, double *__restrict__ partial)
{
    __shared__ double s_red[kObsValues][kObsThreads];
    const BeadBlock blk = S.bead_blocks[blockIdx.x];
    const WorldDev *__restrict__ W = &S.worlds[blk.world];
    const int tid = threadIdx.x;
    const uint32_t s = blk.begin + (uint32_t)tid;
    double acc[kObsValues] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (s < blk.end) {
        const float4 *__restrict__ pos = S.pos[buf];
        const float4 *__restrict__ vel = S.vel[buf];
        const float4 hp = pos[s], hv = vel[s], hf = S.frc[buf][s];
        const double hdt = (double)W->dt * 0.5;
        const double vx = (double)hv.x + (double)hf.x * hdt, vy = (double)hv.y + (double)hf.y * hdt,
                     vz = (double)hv.z + (double)hf.z * hdt;
        acc[0] = vx;
        acc[1] = vy;
        acc[2] = vz;
        acc[3] = vx * vx + vy * vy + vz * vz;
        acc[6] = 1.0;
        const uint32_t me = fbits(hv.w), ht = fbits(hp.w) >> 28;
        const int L[3] = {W->L[0], W->L[1], W->L[2]};
        const int c[3] = {cell_coord(hp.x, L[0]), cell_coord(hp.y, L[1]), cell_coord(hp.z, L[2])};
        for (int ox = -1; ox <= 1; ox++)
            for (int oy = -1; oy <= 1; oy++)
                for (int oz = -1; oz <= 1; oz++) {
                    const int o[3] = {ox, oy, oz};
                    int n[3];
                    double sh[3];
                    for (int d = 0; d < 3; d++) {
                        n[d] = c[d] + o[d];
                        sh[d] = 0.0;
                        if (n[d] < 0) { n[d] += L[d]; sh[d] = -(double)L[d]; }
                        else if (n[d] >= L[d]) { n[d] -= L[d]; sh[d] = (double)L[d]; }
                    }
                    const uint32_t cell = W->cell_off + (uint32_t)((n[0] * L[1] + n[1]) * L[2] + n[2]);
                    for (uint32_t j = S.cell_start[cell]; j < S.cell_start[cell + 1]; j++) {
                        if (fbits(vel[j].w) <= me) continue;            // each unordered pair once
                        const float4 op = pos[j];
                        const double dx = (double)hp.x - ((double)op.x + sh[0]);
                        const double dy = (double)hp.y - ((double)op.y + sh[1]);
                        const double dz = (double)hp.z - ((double)op.z + sh[2]);
                        const double r = sqrt(dx * dx + dy * dy + dz * dz);
                        if (!(r < 1.0)) continue;
                        const double a = (double)W->a[ht * kMaxTypes + (fbits(op.w) >> 28)];
                        acc[4] += a * (1.0 - r) * r;
                        acc[5] += 1.0;
                    }
                }
    }
    block_sum_f64(acc, s_red, tid);
    if (tid < kObsValues) partial[(size_t)blockIdx.x * kObsValues + tid] = acc[tid];
}

// One workgroup per world: its blocks [first[w], first[w + 1]) summed in a
// fixed order into out[w * kObsValues ...].
__global__ __launch_bounds__(kObsThreads) void k_obs_final(const double *__restrict__ partial,
                                                           const uint32_t *__restrict__ first,
                                                           double *__restrict__ out)
{
    __shared__ double s_red[kObsValues][kObsThreads];
    const uint32_t w = blockIdx.x, b0 = first[w], b1 = first[w + 1];
    double acc[kObsValues] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t b = b0 + threadIdx.x; b < b1; b += kObsThreads)
#pragma unroll
        for (int k = 0; k < kObsValues; k++) acc[k] += partial[(size_t)b * kObsValues + k];
    block_sum_f64(acc, s_red, threadIdx.x);
    if (threadIdx.x < kObsValues) out[(size_t)w * kObsValues + threadIdx.x] = acc[threadIdx.x];
}

// ------------------------------------------------------------ utility ----

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// A new layout built from the previous one: one segment per world, blocks
// dealt to segments by their first_block. A resident world's beads (state in
// its sorted order, bead index in v.w rebased), bead-slot map entries and
// term-list starts move from src_off to dst_off, its term-list entries and
// term records move with their term indices and bead indices shifted and the
// world's new index; differences are taken modulo 2^32. An uploaded world's
// state and tables come from the host: only its bead-slot map (the identity:
// uploaded in bead order) is set here.
__global__ __launch_bounds__(256) void k_move_worlds(const MoveSeg *__restrict__ segs, uint32_t n_segs,
                                                     MoveBuffers B)
{
    __shared__ uint32_t s_seg;
    if (threadIdx.x == 0) {
        uint32_t lo = 0, hi = n_segs;                 // last segment with first_block <= blockIdx.x
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (segs[mid].first_block <= blockIdx.x) lo = mid;
            else hi = mid;
        }
        s_seg = lo;
    }
    __syncthreads();
    const MoveSeg g = segs[s_seg];
    const uint32_t i = (blockIdx.x - g.first_block) * blockDim.x + threadIdx.x;
    if (g.fresh) {
        if (i < g.n) B.dsob[g.dst_off + i] = g.dst_off + i;
        return;
    }
    const uint32_t d_bead = g.dst_off - g.src_off;
    if (i < g.n) {
        const uint32_t s = g.src_off + i, d = g.dst_off + i;
        B.dp[d] = B.sp[s];
        float4 v = B.sv[s];
        v.w = __uint_as_float(__float_as_uint(v.w) + d_bead);
        B.dv[d] = v;
        B.df[d] = B.sf[s];
        B.dsob[d] = B.ssob[s] + d_bead;
        B.dstart[d] = B.sstart[s] + (g.dst_list - g.src_list);
    }
    if (i < g.n_list) {
        constexpr uint32_t kIndexMask = (1u << kRoleShift) - 1u;
        const uint32_t e = B.slist[g.src_list + i];
        B.dlist[g.dst_list + i] = (e & ~kIndexMask) | (((e & kIndexMask) + (g.dst_term - g.src_term)) & kIndexMask);
    }
    if (i < g.n_terms) {
        TermRec r = B.sterms[g.src_term + i];
        r.a += d_bead; r.b += d_bead; r.c += d_bead;
        r.kind = (r.kind & 0xffu) | (g.world << kTermWorldShift);
        B.dterms[g.dst_term + i] = r;
    }
}

// Per-world flags carried to a new layout: nw[i] = old[map[i]], 0 for map ~0.
__global__ void k_remap_u32(uint32_t *__restrict__ nw, const uint32_t *__restrict__ old,
                            const uint32_t *__restrict__ map, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) nw[i] = map[i] == ~0u ? 0u : old[map[i]];
}

// ------------------------------------------------------------ launchers ----

static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t scan_tiles(uint32_t n_cells) { return std::max<uint32_t>(1, ceil_div(n_cells, kScanBlock)); }

hipError_t launch_step(const StepLaunch &L, hipStream_t st, KernelEvents *ev)
{
    const DeviceState &S = L.S;
    const int cur = L.cur;
    DeviceState Sp = S;
    Sp.ctr = S.ctr + L.cpar * L.n_worlds;       // counters read this step
    auto rec = [&](int k) { if (ev) (void)hipEventRecord(ev->ev[k], st); };
    const uint32_t nt = scan_tiles(L.n_cells);

    rec(0);
    if (L.integrate)
        k_integrate<<<std::max<uint32_t>(1, L.n_bead_blocks), kBeadThreads, 0, st>>>(Sp, cur, nt, L.n_bead_blocks);
    rec(1);
    if (nt <= L.lookback_tiles) {
        k_scan<<<nt, kScanThreads, 0, st>>>(S.cell_count, L.n_cells, S.cell_start, S.scan_tiles, S.scan_ticket, nt,
                                             S.flags);
    } else {
        k_scan_reduce<<<nt, kScanThreads, 0, st>>>(S.cell_count, L.n_cells, L.tile_sum);
        k_scan_top<<<1, kTopThreads, 0, st>>>(L.tile_sum, nt, S.cell_start, L.n_cells);
        k_scan_down<<<nt, kScanThreads, 0, st>>>(S.cell_count, L.n_cells, S.cell_start, L.tile_sum);
    }
    rec(2);
    const uint32_t gb = ceil_div(L.n_beads, kBeadThreads);
    const uint32_t sg = std::max(gb + ceil_div(L.n_terms, kBeadThreads), ceil_div(L.n_worlds, kBeadThreads));
    k_scatter<<<std::max<uint32_t>(1, sg), kBeadThreads, 0, st>>>(Sp, cur, L.n_beads, gb, L.n_terms,
                                                                   S.ctr + (L.cpar ^ 1) * L.n_worlds, L.n_worlds);
    rec(3);
    k_reorder<<<std::max<uint32_t>(1, ceil_div(std::max(L.n_beads, L.n_cells), kBeadThreads)), kBeadThreads, 0, st>>>(
        S, cur, L.n_beads, L.n_cells, L.n_stationary > 0 ? 1 : 0);
    rec(4);
    const hipError_t e = launch_force(L, st);
    if (e != hipSuccess) return e;
    rec(5);
    if (L.n_stationary)
        k_stationary<<<ceil_div(L.n_stationary, kBeadThreads), kBeadThreads, 0, st>>>(Sp, cur ^ 1, L.n_stationary);
    rec(6);
    return hipGetLastError();
}

hipError_t launch_force(const StepLaunch &L, hipStream_t st)
{
    // cell bricks staged in LDS, unless a world has per-cell strengths (the
    // per-slot kernel carries that lookup)
    if (L.force_kernel == 1 && !L.spatial && L.bricks) return launch_force_brick(L, st);
    DeviceState Sp = L.S;
    Sp.ctr = L.S.ctr + L.cpar * L.n_worlds;      // counters read by that step
    const int nxt = L.cur ^ 1;
    const NextBuffers *ns = L.fuse ? L.next + (L.cur * 2 + L.cpar) : nullptr;
    const uint32_t fg = L.force_grid ? std::min<uint32_t>(L.force_grid, L.n_force_blocks) : L.n_force_blocks;
    constexpr uint32_t t3 = kForceBeads * kForcePlanes;
    if (fg && L.spatial) {
        // any world with spatially varying strengths: one shape, with the per-cell lookup
        k_force<3, 4, false, true><<<fg, t3, 0, st>>>(Sp, nxt, L.n_force_blocks, ns);
    } else if (fg) {
        switch (L.force_variant) {
        case 1: k_force<4, 4, true, false><<<fg, kForceThreads, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        case 2: k_force<2, 5, false, false><<<fg, t3, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        case 3: k_force<3, 5, true, false><<<fg, kForceThreads, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        // smaller queues, fewer registers per lane, more waves per SIMD (the
        // queue and register budget of variant 4 give 6 waves per SIMD against 5)
        case 4: k_force<2, 6, false, false, 128><<<fg, t3, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        case 5: k_force<3, 6, false, false, 192><<<fg, t3, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        case 6: k_force<2, 6, true, false, 128><<<fg, kForceThreads, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        // one wave per stencil row (nine per workgroup): shorter dependent chains
        // per wave, for worlds too small to give each SIMD several waves
        case 7: return launch_force_rows(L, st);
        default: k_force<3, 5, false, false><<<fg, t3, 0, st>>>(Sp, nxt, L.n_force_blocks, ns); break;
        }
    }
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_fill_u32<<<ceil_div(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_trace_pairs(const DeviceState &S, int buf, uint32_t world, uint32_t n_beads, uint32_t t,
                              pdpd_pair_record *out, uint32_t cap, uint32_t *count, hipStream_t st)
{
    if (n_beads == 0) return hipSuccess;
    k_trace_pairs<<<ceil_div(n_beads, kBeadThreads), kBeadThreads, 0, st>>>(S, buf, world, t, out, cap, count);
    return hipGetLastError();
}

hipError_t launch_move_worlds(const MoveSeg *segs, uint32_t n_segs, uint32_t n_blocks, const MoveBuffers &B,
                              hipStream_t st)
{
    if (n_segs == 0 || n_blocks == 0) return hipSuccess;
    k_move_worlds<<<n_blocks, 256, 0, st>>>(segs, n_segs, B);
    return hipGetLastError();
}

hipError_t launch_remap_u32(uint32_t *nw, const uint32_t *old, const uint32_t *map, uint32_t n, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_remap_u32<<<ceil_div(n, 256), 256, 0, st>>>(nw, old, map, n);
    return hipGetLastError();
}

hipError_t launch_observables(const DeviceState &S, int buf, uint32_t n_bead_blocks, const uint32_t *first,
                              uint32_t n_worlds, double *partial, double *out, hipStream_t st)
{
    if (n_worlds == 0) return hipSuccess;
    if (n_bead_blocks) k_obs_partial<<<n_bead_blocks, kObsThreads, 0, st>>>(S, buf, partial);
    k_obs_final<<<n_worlds, kObsThreads, 0, st>>>(partial, first, out);
    return hipGetLastError();
}

hipError_t launch_debug_rng(uint32_t n, const uint64_t *th, const uint32_t *a, const uint32_t *b, uint32_t *u32,
                            float *f32, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_debug_rng<<<ceil_div(n, 256), 256, 0, st>>>(n, th, a, b, u32, f32);
    return hipGetLastError();
}

}  // namespace pdpd
