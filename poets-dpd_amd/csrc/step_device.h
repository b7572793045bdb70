// step_device.h -- device helpers shared by the step kernels (kernels.hip,
// force_brick.hip): the counter-based RNG, block scans, cell coordinates, the
// integrate-and-count of one bead, wave synchronisation and the XCD mapping.
#pragma once

#include "device_types.h"
#include "pdpd_hip.h"

#include <type_traits>

namespace pdpd {

constexpr float kMinR2 = 1e-9f * 1e-9f;           // MIN_DISTANCE_CUTOFF_SQR
constexpr float kTwoM32 = 0.00000000023283064365386962890625f;  // 2^-32
constexpr uint32_t kInactive = 0x80000000u;
constexpr uint32_t kSlotMask = kMaxLayoutBeads - 1;
constexpr unsigned long long kTileAgg = 1ull << 32, kTileIncl = 2ull << 32;
constexpr uint32_t kScanSpinLimit = 1u << 24;
constexpr uint32_t kXcds = 8;     // MI355X: 8 accelerator complex dies, each with its own L2

// ---------------------------------------------------------------- RNG ----

__device__ __forceinline__ uint64_t mix64_m2(uint64_t x)
{
    const uint64_t C = 0xd392d2a7ull;
    x ^= x >> 32;
    x *= C;
    x ^= x >> 32;
    x *= C;
    x ^= x >> 32;
    return x;
}

__device__ __forceinline__ uint64_t step_hash(uint32_t t, uint32_t seed_lo, uint32_t seed_hi)
{
    const uint64_t seed = ((uint64_t)seed_hi << 32) | seed_lo;
    return mix64_m2(seed + mix64_m2((uint64_t)t));
}

// hash_rng_sym(t_hash, a, b): low word of mix64_m2((min | max << 32) ^ t_hash),
// carried in 32-bit halves: each 64x32 product is mul_lo + mul_hi + mul_lo.
__device__ __forceinline__ uint32_t rng_sym(uint32_t th_lo, uint32_t th_hi, uint32_t a, uint32_t b)
{
    const uint32_t C = 0xd392d2a7u;
    uint32_t lo = min(a, b) ^ th_lo;
    uint32_t hi = max(a, b) ^ th_hi;
    lo ^= hi;
    uint32_t l1 = lo * C;
    uint32_t h1 = __umulhi(lo, C) + hi * C;
    l1 ^= h1;
    uint32_t l2 = l1 * C;
    uint32_t h2 = __umulhi(l1, C) + h1 * C;
    return l2 ^ h2;
}

__device__ __forceinline__ float rng_uniform(uint32_t r)
{
    return (float)(int32_t)r * kTwoM32;
}

// ------------------------------------------------------------ helpers ----

__device__ __forceinline__ uint32_t fbits(float x) { return __float_as_uint(x); }

// The pair test kMinR2 <= r2 < 1 for r2 >= 0 or NaN (a sum of squares), on the
// bit patterns: non-negative floats order like their bits, so one subtraction
// and one unsigned compare select the same set as the two float compares (NaN
// and +inf fall outside).
__device__ __forceinline__ bool in_cutoff(float r2)
{
    const uint32_t lo = __float_as_uint(kMinR2), span = __float_as_uint(1.0f) - lo;
    return __float_as_uint(r2) - lo < span;
}

// Exclusive scan of one value per thread across a block of NW wave64s.
// Returns the exclusive prefix; *total receives the block sum.
template <int NW>
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *s_wsum, uint32_t *total)
{
    const int lane = threadIdx.x & 63;
    const int wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) s_wsum[wid] = x;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (int k = 0; k < NW; k++) {
        uint32_t s = s_wsum[k];
        if (k < wid) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

__device__ __forceinline__ int cell_coord(float x, int L)
{
    int c = (int)floorf(x);
    c = c < 0 ? 0 : c;
    return c >= L ? L - 1 : c;
}

// ---------------------------------------------------------- integrate ----

// One bead's finish-kick, drift, half-kick and wrap from its force f:
// v(t) = w + f dt/2, v(t+dt/2) = v(t) + f dt/2, x(t+dt) = x + dt v(t+dt/2),
// then w = v(t+dt/2). A bead of a stationary type keeps x; the pair force sees
// v(t) (the imported v on the first step, 0 after) --
// naive_dpd_engine_half_merge.hpp:327-331. Returns the failure flags the move
// raises: PDPD_FLAG_CELL_JUMP for a move of more than one cell on an axis
// (the periodic wrap aside; dpd_engine_avx2_half_merge_tbb.hpp:800-811),
// PDPD_FLAG_VELOCITY_BOUND for a velocity component beyond the world's bound
// (by default: not finite).
__device__ __forceinline__ uint32_t drift_bead(const WorldDev *__restrict__ W, float4 &p, float4 &v, const float4 f)
{
    uint32_t fl = 0;
    const float dt = W->dt, hdt = W->hdt;
    const float fv[3] = {f.x, f.y, f.z};
    float x[3] = {p.x, p.y, p.z};
    float wv[3] = {v.x, v.y, v.z};
    if ((W->stationary >> (fbits(p.w) >> 28)) & 1u) {
#pragma unroll
        for (int d = 0; d < 3; d++) wv[d] = fmaf(fv[d], hdt, wv[d]);
    } else {
#pragma unroll
        for (int d = 0; d < 3; d++) {
            const float Lf = W->Lf[d];
            float vv = fmaf(fv[d], hdt, wv[d]);     // v(t)       = w + f dt/2
            float vh = fmaf(fv[d], hdt, vv);        // v(t+dt/2)  = v + f dt/2
            float xn = fmaf(vh, dt, x[d]);          // x(t+dt)    = x + dt v(t+dt/2)
            xn += (xn < 0.0f) ? Lf : 0.0f;
            xn -= (xn >= Lf) ? Lf : 0.0f;
            const int jump = abs(cell_coord(xn, W->L[d]) - cell_coord(x[d], W->L[d]));
            fl |= (jump > 1 && jump < W->L[d] - 1) ? PDPD_FLAG_CELL_JUMP : 0u;
            fl |= fabsf(vh) <= W->max_v ? 0u : PDPD_FLAG_VELOCITY_BOUND;
            x[d] = xn;
            wv[d] = vh;
        }
    }
    p = make_float4(x[0], x[1], x[2], p.w);
    v = make_float4(wv[0], wv[1], wv[2], v.w);
    return fl;
}

// Runs of equal keys in consecutive lanes of one wave (called converged; the
// valid lanes are a prefix and hold consecutive slots in the last step's cell
// order, so beads of one cell come in runs): for each valid lane, the first
// lane of its run (start) and the run's length from this lane on.
__device__ __forceinline__ void key_runs(uint32_t key, bool valid, int &start, int &next, bool &head)
{
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(key, 1, 64);
    head = valid && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(head);
    const unsigned long long lanes = __ballot(valid);
    const unsigned long long upto = (2ull << lane) - 1ull;          // lanes 0..lane
    start = 63 - __clzll((long long)(heads & upto));                 // first lane of this run
    const unsigned long long above = heads & ~upto;
    next = above ? __ffsll((long long)above) - 1 : (int)__popcll(lanes);
}

// The cell key of the bead in slot s at position p (with kInactive for a
// bead that does not move this step), and its count in cell_count: one
// atomic per run of equal keys, whose result is not waited for. Raises
// PDPD_FLAG_BAD_POSITION for a position outside the box, with the flags
// `more` of its move, in its world's flag word wflag. Called by all lanes of a
// wave, converged.
__device__ __forceinline__ void count_bead(uint32_t *__restrict__ wflag, uint32_t more,
                                           uint32_t *__restrict__ cell_count, uint32_t *__restrict__ key_out,
                                           const WorldDev *__restrict__ W, uint32_t s, const float4 p, bool active,
                                           bool valid)
{
    const float x[3] = {p.x, p.y, p.z};
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; d++) bad |= !(x[d] >= 0.0f && x[d] < W->Lf[d]);
    const uint32_t fl = more | (bad ? PDPD_FLAG_BAD_POSITION : 0u);
    if (valid && fl) atomicOr(wflag, fl);
    const int c0 = cell_coord(x[0], W->L[0]), c1 = cell_coord(x[1], W->L[1]), c2 = cell_coord(x[2], W->L[2]);
    const uint32_t key = valid ? W->cell_off + (uint32_t)((c0 * W->L[1] + c1) * W->L[2] + c2) : 0xffffffffu;
    int start, next;
    bool head;
    key_runs(key, valid, start, next, head);
    if (head) atomicAdd(&cell_count[key], (uint32_t)(next - (int)(threadIdx.x & 63)));
    if (valid) key_out[s] = key | (active ? 0u : kInactive);
}

// Workgroups are dealt to the XCDs round-robin (workgroup i runs on XCD i mod 8).
// Map logical index i to a work item so that each XCD takes one contiguous
// share of the (cell-ordered) work list and its L2 holds one region of the
// world instead of all of it. A bijection on [0, n); with a grid-stride loop
// whose stride is a multiple of 8, a workgroup stays on its XCD's share.
__device__ __forceinline__ uint32_t xcd_chunk(uint32_t i, uint32_t n)
{
    const uint32_t q = n / kXcds, r = n % kXcds;
    const uint32_t x = i % kXcds, k = i / kXcds;
    return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Maximum over the wave's active lanes (the device library's DPP reduction:
// register moves between lanes, no LDS round trips).
__device__ __forceinline__ uint32_t wave_max(uint32_t v) { return __ockl_wfred_max_u32(v); }

// Orders this wave's LDS writes before its later LDS reads by other lanes.
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The bonded terms of the bead with global index b (evaluated during
// k_scatter, kTermFrcStride doubles per term), summed in FP64 in the
// per-polymer order of naive_dpd_engine_half_merge.hpp:376-385 (its bonds, then
// its angles). Returns the number of terms.
__device__ __forceinline__ uint32_t bonded_sum(const DeviceState &S, uint32_t b, double &bx, double &by, double &bz)
{
    const uint32_t t0 = S.term_start[b], t1 = S.term_start[b + 1];
    for (uint32_t t = t0; t < t1; t++) {
        const uint32_t e = S.bead_terms[t];
        const uint32_t role = e >> kRoleShift;
        const double *__restrict__ F = S.term_frc + (size_t)(e & ((1u << kRoleShift) - 1u)) * kTermFrcStride;
        double cx, cy, cz;
        if (role == kRoleBondHead) { cx = -F[0]; cy = -F[1]; cz = -F[2]; }
        else if (role == kRoleBondTail || role == kRoleAngleHead) { cx = F[0]; cy = F[1]; cz = F[2]; }
        else if (role == kRoleAngleTail) { cx = F[3]; cy = F[4]; cz = F[5]; }
        else { cx = -(F[0] + F[3]); cy = -(F[1] + F[4]); cz = -(F[2] + F[5]); }
        bx += cx; by += cy; bz += cz;
    }
    return t1 - t0;
}

// The next step's integration, done by the pair-force kernel for every step of
// a chain but the last. From the position, half-step velocity and the force
// just computed, each bead's drifted state goes into buffer nxt ^ 1 (the
// step's source buffer, free after the sort) at the same slot, with its cell
// key and count: what k_integrate would compute at the start of the
// next step, so the next step starts at its scan. Chains are only formed over
// steps in which every resident world steps (no step budget runs out inside
// a chain but at its end), so every bead is active in both steps.
// The kernel reads the next step's buffers through an entry of a table in
// device memory (null: not chained), after its pair loop, so that their
// addresses take no scalar registers during the loop.

// One bead of world `world` (all lanes of the wave call it, the valid ones a prefix).
__device__ __forceinline__ void advance_next(const NextBuffers &B, const WorldDev *__restrict__ W, uint32_t world,
                                             uint32_t s, bool valid, float4 p, float4 v, const float4 f)
{
    uint32_t fl = 0;
    if (valid) {
        // k_stationary zeroes the velocity of a stationary bead after this kernel
        if ((W->stationary >> (fbits(p.w) >> 28)) & 1u) v = make_float4(0.0f, 0.0f, 0.0f, v.w);
        fl = drift_bead(W, p, v, f);
        B.pos[s] = p;
        B.vel[s] = v;
    }
    count_bead(&B.wflags[world], fl, B.cell_count, B.key, W, s, p, true, valid);
}

// PDPD_FLAG_FORCE_BOUND for a force component beyond the world's bound (by
// default: not finite).
__device__ __forceinline__ uint32_t force_flags(const WorldDev *__restrict__ W, float fx, float fy, float fz)
{
    const float m = W->max_f;
    return (fabsf(fx) <= m && fabsf(fy) <= m && fabsf(fz) <= m) ? 0u : PDPD_FLAG_FORCE_BOUND;
}

// The table entry, read after a compiler barrier so the loads stay here.
__device__ __forceinline__ NextBuffers next_buffers(const NextBuffers *b)
{
    __asm__ volatile("" ::: "memory");
    return *b;
}

}  // namespace pdpd
