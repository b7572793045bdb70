// kernels.hip -- gfx950 kernels of one DPD time step.
//
// Step (one set of launches advances every resident world):
//   k_integrate   finish-kick + drift + half-kick + wrap, cell key, arrival rank
//                 (dpd_maths_core_half_step_raw.hpp:42-72,114-124 fused as
//                 dpd_engine_avx2_half_merge_tbb.hpp:780-800 does)
//   k_scan        single-pass exclusive scan of the cell counts (look-back)
//   k_scatter     sorted position of every bead's slot
//   k_reorder     stable gather into the other buffer, bead-index order in a cell
//   k_force       pair forces: per workgroup 64 consecutive beads x 3 waves, one
//                 wave per x-plane of the 27-cell stencil; neighbours are read
//                 straight from the cell-sorted arrays (contiguous z-rows); each
//                 lane buffers the candidates inside the cutoff and then
//                 evaluates them; register accumulation, no atomics
//                 (dpd_maths_core_half_step_raw.hpp:129-216; RNG of
//                 hash.hpp:61-88 + dpd_maths_core.hpp:55-79)
//   k_bonds       Hookean + angle terms per bonded bead, in the per-polymer
//                 order of naive_dpd_engine_half_merge.hpp:376-385 (terms of
//                 :638-690 / dpd_maths_core_half_step.hpp:178-231), then the
//                 per-world step counter
// The final half-kick v += f dt/2 of a step is folded into the next k_integrate
// (the engine stores w = v - f dt/2).
//
// Arithmetic is FP32, built with -ffp-contract=off: every fused multiply-add is
// an explicit fmaf, sqrtf and '/' are IEEE correctly rounded. The CPU oracle
// performs the same operations in the same order.
#include "device_types.h"
#include "kernels.h"
#include "pdpd_hip.h"

#include <algorithm>

namespace pdpd {

constexpr float kMinR2 = 1e-9f * 1e-9f;           // MIN_DISTANCE_CUTOFF_SQR
constexpr float kTwoM32 = 0.00000000023283064365386962890625f;  // 2^-32
constexpr uint32_t kInactive = 0x80000000u;
constexpr uint32_t kSlotMask = kMaxLayoutBeads - 1;
constexpr unsigned long long kTileAgg = 1ull << 32, kTileIncl = 2ull << 32;
constexpr uint32_t kScanSpinLimit = 1u << 24;

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

// ------------------------------------------------------------ helpers ----

__device__ __forceinline__ uint32_t fbits(float x) { return __float_as_uint(x); }

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

__global__ __launch_bounds__(kBeadThreads) void k_integrate(DeviceState S, int cur, uint32_t n_tiles, uint32_t n_blocks)
{
    if (blockIdx.x == 0) {       // reset the scan's look-back state for this step
        for (uint32_t i = threadIdx.x; i < n_tiles; i += blockDim.x) S.scan_tiles[i] = 0ull;
        if (threadIdx.x == 0) *S.scan_ticket = 0u;
    }
    if (blockIdx.x >= n_blocks) return;
    const BeadBlock blk = S.bead_blocks[blockIdx.x];
    const uint32_t s = blk.begin + threadIdx.x;
    if (s >= blk.end) return;
    const WorldDev *__restrict__ W = &S.worlds[blk.world];
    const bool active = S.ctr[blk.world].remaining > 0;   // S.ctr points at this step's counters

    float4 p = S.pos[cur][s];
    float x[3] = {p.x, p.y, p.z};
    if (active) {
        const float4 f = S.frc[cur][s];
        float4 v = S.vel[cur][s];
        const float dt = W->dt, hdt = W->hdt;
        const float fv[3] = {f.x, f.y, f.z};
        float wv[3] = {v.x, v.y, v.z};
        if ((W->stationary >> (fbits(p.w) >> 28)) & 1u) {
            // stationary type: x stays; the pair force sees v(t) (the imported v on
            // the first step, 0 after) -- naive_dpd_engine_half_merge.hpp:327-331
#pragma unroll
            for (int d = 0; d < 3; d++) wv[d] = fmaf(fv[d], hdt, wv[d]);
            S.vel[cur][s] = make_float4(wv[0], wv[1], wv[2], v.w);
        } else {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                const float Lf = W->Lf[d];
                float vv = fmaf(fv[d], hdt, wv[d]);     // v(t)       = w + f dt/2
                float vh = fmaf(fv[d], hdt, vv);        // v(t+dt/2)  = v + f dt/2
                float xn = fmaf(vh, dt, x[d]);          // x(t+dt)    = x + dt v(t+dt/2)
                xn += (xn < 0.0f) ? Lf : 0.0f;
                xn -= (xn >= Lf) ? Lf : 0.0f;
                x[d] = xn;
                wv[d] = vh;
            }
            S.pos[cur][s] = make_float4(x[0], x[1], x[2], p.w);
            S.vel[cur][s] = make_float4(wv[0], wv[1], wv[2], v.w);
        }
    }
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; d++) bad |= !(x[d] >= 0.0f && x[d] < W->Lf[d]);
    if (bad) atomicOr(S.flags, PDPD_FLAG_BAD_POSITION);
    const int c0 = cell_coord(x[0], W->L[0]), c1 = cell_coord(x[1], W->L[1]), c2 = cell_coord(x[2], W->L[2]);
    const uint32_t key = W->cell_off + (uint32_t)((c0 * W->L[1] + c1) * W->L[2] + c2);
    const uint32_t r = atomicAdd(&S.cell_count[key], 1u);
    S.key[s] = key;
    S.rank[s] = r | (active ? 0u : kInactive);
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
    if (threadIdx.x == 0) {
        uint32_t prefix = 0;
        if (tile == 0) {
            __hip_atomic_store(&tiles[0], kTileIncl | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&tiles[tile], kTileAgg | agg, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int j = (int)tile - 1;
            uint32_t spins = 0;
            while (j >= 0) {
                const unsigned long long st = __hip_atomic_load(&tiles[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                const unsigned long long status = st & 0xffffffff00000000ull;
                if (status == 0ull) {
                    if (++spins > kScanSpinLimit) {
                        atomicOr(flags, PDPD_FLAG_SCAN_TIMEOUT);
                        break;
                    }
                    __builtin_amdgcn_s_sleep(1);
                    continue;
                }
                prefix += (uint32_t)st;
                if (status == kTileIncl) break;
                j--;
            }
            __hip_atomic_store(&tiles[tile], kTileIncl | (prefix + agg), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_prefix = prefix;
        if (tile == n_tiles - 1) cell_start[n] = prefix + agg;
    }
    __syncthreads();
    ex += s_prefix;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        const uint32_t i = base + k;
        if (i < n) {
            cell_start[i] = ex;
            count[i] = 0u;          // ready for the next step's arrivals
        }
        ex += v[k];
    }
}

// ------------------------------------------------------- sort scatter ----

__global__ __launch_bounds__(kBeadThreads) void k_scatter(DeviceState S, int cur, uint32_t n)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = S.rank[s];
    const uint32_t j = S.cell_start[S.key[s]] + (r & ~kInactive);
    S.perm[j] = s | (r & kInactive);
    S.sorted_idx[j] = fbits(S.vel[cur][s].w);
}

// Stable order inside a cell: beads ranked by global bead index, so the order
// (and every force sum) is independent of arrival order and of the batch.
__global__ __launch_bounds__(kBeadThreads) void k_reorder(DeviceState S, int cur, uint32_t n)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int nxt = cur ^ 1;
    const uint32_t e = S.perm[j];
    const uint32_t s = e & ~kInactive;
    const uint32_t k = S.key[s];
    const uint32_t a = S.cell_start[k], b = S.cell_start[k + 1];
    const uint32_t me = S.sorted_idx[j];
    uint32_t rank = 0;
    for (uint32_t q = a; q < b; q++) rank += S.sorted_idx[q] < me ? 1u : 0u;
    const uint32_t dst = a + rank;
    S.pos[nxt][dst] = S.pos[cur][s];
    S.vel[nxt][dst] = S.vel[cur][s];
    if (e & kInactive) S.frc[nxt][dst] = S.frc[cur][s];
    S.slot_of_bead[me] = dst;
}

// --------------------------------------------------------- pair force ----

// Shift codes of a candidate's periodic image, 2 bits per axis: 0 none,
// 1 = +L, 2 = -L. Stored with the candidate's slot in the per-lane list.
__device__ __forceinline__ float shift_of(uint32_t code, float L)
{
    return code == 1u ? L : (code == 2u ? -L : 0.0f);
}

__global__ __launch_bounds__(kForceThreads) void k_force(DeviceState S, int nxt, uint32_t n_blocks)
{
    __shared__ uint32_t s_list[kPairList * kForceThreads];
    __shared__ float s_part[kForcePlanes - 1][3][kForceBeads];

    const int tid = threadIdx.x;
    const int plane = tid / kForceBeads;            // wave-uniform
    const int lane = tid % kForceBeads;
    const float4 *__restrict__ pos = S.pos[nxt];
    const float4 *__restrict__ vel = S.vel[nxt];
    const uint32_t *__restrict__ cs = S.cell_start;

    for (uint32_t bi = blockIdx.x; bi < n_blocks; bi += gridDim.x) {
        const BeadBlock blk = S.force_blocks[bi];
        if (S.ctr[blk.world].remaining == 0) continue;       // uniform: frozen world
        const WorldDev *__restrict__ W = &S.worlds[blk.world];
        const uint32_t s = blk.begin + lane;
        const bool live = s < blk.end;
        float fx = 0.0f, fy = 0.0f, fz = 0.0f;
        // a stationary home bead ends the step with f = 0 (naive_dpd_engine_half_merge.hpp:387-390);
        // its pairs with moving beads are evaluated from the moving side
        if (live && !((W->stationary >> (fbits(pos[s].w) >> 28)) & 1u)) {
            const WorldCtr wc = S.ctr[blk.world];
            const uint64_t th = step_hash(wc.t, W->seed_lo, W->seed_hi);
            const uint32_t th_lo = (uint32_t)th, th_hi = (uint32_t)(th >> 32);
            const float scale = W->rng_scale;
            const int L0 = W->L[0], L1 = W->L[1], L2 = W->L[2];
            const float4 hp = pos[s];
            const float4 hv = vel[s];
            const uint32_t hh = fbits(hp.w);
            const uint32_t trow = (hh >> 28) * kMaxTypes;
            const int cx = cell_coord(hp.x, L0), cy = cell_coord(hp.y, L1), cz = cell_coord(hp.z, L2);
            uint32_t cnt = 0;
            // one accepted pair: dpd_maths_core_half_step_raw.hpp:129-216 with the
            // terms grouped as a wr + sgw (scale u - sgw rdotv)
            auto pair = [&](uint32_t e, const float4 &op, const float4 &ov) {
                const float ox = op.x + shift_of((e >> 26) & 3u, W->Lf[0]);
                const float oy = op.y + shift_of((e >> 28) & 3u, W->Lf[1]);
                const float oz = op.z + shift_of((e >> 30) & 3u, W->Lf[2]);
                const float dx = hp.x - ox, dy = hp.y - oy, dz = hp.z - oz;
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float rr = sqrtf(r2);
                const float inv_r = 1.0f / rr;
                const float wr = 1.0f - rr;
                const float ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
                const float dvx = hv.x - ov.x, dvy = hv.y - ov.y, dvz = hv.z - ov.z;
                const float rdotv = fmaf(ez, dvz, fmaf(ey, dvy, ex * dvx));
                const uint32_t oh = fbits(op.w);
                const uint32_t ti = trow + (oh >> 28);
                const float sgw = W->sg[ti] * wr;
                const float u = rng_uniform(rng_sym(th_lo, th_hi, hh, oh));
                const float F = fmaf(W->a[ti], wr, sgw * fmaf(-sgw, rdotv, scale * u));
                fx = fmaf(ex, F, fx);
                fy = fmaf(ey, F, fy);
                fz = fmaf(ez, F, fz);
            };
            // evaluate the buffered pairs in order, loads of two pairs in flight
            auto flush = [&]() {
                uint32_t q = 0;
                for (; q + 2 <= cnt; q += 2) {
                    const uint32_t e0 = s_list[q * kForceThreads + tid];
                    const uint32_t e1 = s_list[(q + 1) * kForceThreads + tid];
                    const float4 op0 = pos[e0 & kSlotMask], ov0 = vel[e0 & kSlotMask];
                    const float4 op1 = pos[e1 & kSlotMask], ov1 = vel[e1 & kSlotMask];
                    pair(e0, op0, ov0);
                    pair(e1, op1, ov1);
                }
                if (q < cnt) {
                    const uint32_t e = s_list[q * kForceThreads + tid];
                    pair(e, pos[e & kSlotMask], vel[e & kSlotMask]);
                }
                cnt = 0;
            };
            auto test = [&](const float4 &op, uint32_t m, bool valid, uint32_t code, float sx, float sy, float sz) {
                const float dx = hp.x - (op.x + sx), dy = hp.y - (op.y + sy), dz = hp.z - (op.z + sz);
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (valid && r2 < 1.0f && r2 >= kMinR2 && m != s) s_list[(cnt++) * kForceThreads + tid] = m | code;
            };
            // candidates of cells [m0, m1) with one periodic image, four loads in
            // flight; the list is flushed between groups so no loaded candidate is
            // live across a flush
            auto scan = [&](uint32_t m0, uint32_t m1, uint32_t code, float sx, float sy, float sz) {
                for (uint32_t m = m0; m < m1; m += 4) {
                    if (cnt > kPairList - 4) flush();
                    const uint32_t last = m1 - 1;
                    const float4 p0 = pos[m];
                    const float4 p1 = pos[min(m + 1, last)];
                    const float4 p2 = pos[min(m + 2, last)];
                    const float4 p3 = pos[min(m + 3, last)];
                    test(p0, m, true, code, sx, sy, sz);
                    test(p1, m + 1, m + 1 < m1, code, sx, sy, sz);
                    test(p2, m + 2, m + 2 < m1, code, sx, sy, sz);
                    test(p3, m + 3, m + 3 < m1, code, sx, sy, sz);
                }
            };
            int ox = cx + plane - 1;
            uint32_t codex = 0;
            if (ox < 0) { ox += L0; codex = 2u; } else if (ox >= L0) { ox -= L0; codex = 1u; }
            const float sx = shift_of(codex, W->Lf[0]);
#pragma unroll 1
            for (int dyi = -1; dyi <= 1; dyi++) {
                int oy = cy + dyi;
                uint32_t codey = 0;
                if (oy < 0) { oy += L1; codey = 2u; } else if (oy >= L1) { oy -= L1; codey = 1u; }
                const float sy = shift_of(codey, W->Lf[1]);
                const uint32_t row = W->cell_off + (uint32_t)((ox * L1 + oy) * L2);
                const uint32_t cxy = (codex << 26) | (codey << 28);
                // z order -1, 0, +1: an image below 0 first, the contiguous run, an image above L-1 last
                if (cz == 0) scan(cs[row + L2 - 1], cs[row + L2], cxy | (2u << 30), sx, sy, -W->Lf[2]);
                const int zlo = cz > 0 ? cz - 1 : 0, zhi = cz < L2 - 1 ? cz + 1 : L2 - 1;
                scan(cs[row + zlo], cs[row + zhi + 1], cxy, sx, sy, 0.0f);
                if (cz == L2 - 1) scan(cs[row], cs[row + 1], cxy | (1u << 30), sx, sy, W->Lf[2]);
            }
            flush();
        }
        // combine the three planes in order: (p0 + p1) + p2
        if (plane > 0) {
            s_part[plane - 1][0][lane] = fx;
            s_part[plane - 1][1][lane] = fy;
            s_part[plane - 1][2][lane] = fz;
        }
        __syncthreads();
        if (plane == 0 && live) {
            fx = (fx + s_part[0][0][lane]) + s_part[1][0][lane];
            fy = (fy + s_part[0][1][lane]) + s_part[1][1][lane];
            fz = (fz + s_part[0][2][lane]) + s_part[1][2][lane];
            S.frc[nxt][s] = make_float4(fx, fy, fz, 0.0f);
        }
        __syncthreads();
    }
}

// -------------------------------------------------------------- bonds ----

__device__ __forceinline__ double min_image(double d, double len)
{
    if (d > 0.5 * len) d -= len;
    else if (d < -0.5 * len) d += len;
    return d;
}

// Contribution of one bonded term to one of its beads, in FP64 from the FP32
// positions (differences of floats are exact in double). FP64 keeps the
// theta0 != 0 angle force, which divides by sin(phi), stable near straight
// chains. Hookean: naive_dpd_engine_half_merge.hpp:638-654; angle:
// dpd_maths_core_half_step.hpp:178-231 with the prefactor floor of
// naive_dpd_engine.hpp:681-682 (the argument clamped at 0).
__device__ void bond_term(const DeviceState &S, const float4 *__restrict__ pos, const WorldDev *__restrict__ W,
                          uint32_t code, double &cx, double &cy, double &cz)
{
    const uint32_t role = code >> kRoleShift;
    const uint32_t idx = code & ((1u << kRoleShift) - 1u);
    const double Lx = W->Lf[0], Ly = W->Lf[1], Lz = W->Lf[2];
    if (role <= kRoleBondTail) {
        const BondRec b = S.bonds[idx];
        const float4 xh = pos[S.slot_of_bead[b.head]], xt = pos[S.slot_of_bead[b.tail]];
        const double dx = min_image((double)xt.x - (double)xh.x, Lx);
        const double dy = min_image((double)xt.y - (double)xh.y, Ly);
        const double dz = min_image((double)xt.z - (double)xh.z, Lz);
        const double r = sqrt(fma(dz, dz, fma(dy, dy, dx * dx)));
        const double hf = b.kappa * (b.r0 - r);
        const double sc = hf / r;
        const double sgn = role == kRoleBondHead ? -1.0 : 1.0;   // head -= f, tail += f
        cx = sgn * (dx * sc);
        cy = sgn * (dy * sc);
        cz = sgn * (dz * sc);
        return;
    }
    const AngleRec a = S.angles[idx];
    const float4 xh = pos[S.slot_of_bead[a.head]], xm = pos[S.slot_of_bead[a.mid]], xt = pos[S.slot_of_bead[a.tail]];
    const double d0x = min_image((double)xm.x - (double)xh.x, Lx), d0y = min_image((double)xm.y - (double)xh.y, Ly),
                 d0z = min_image((double)xm.z - (double)xh.z, Lz);
    const double d1x = min_image((double)xt.x - (double)xm.x, Lx), d1y = min_image((double)xt.y - (double)xm.y, Ly),
                 d1z = min_image((double)xt.z - (double)xm.z, Lz);
    const double r01 = sqrt(fma(d0z, d0z, fma(d0y, d0y, d0x * d0x)));
    const double r12 = sqrt(fma(d1z, d1z, fma(d1y, d1y, d1x * d1x)));
    const double mag = r01 * r12;
    double hx = 0, hy = 0, hz = 0, tx = 0, ty = 0, tz = 0;
    if (mag > 0.0001) {
        const double b1 = r01 * r01, b2 = r12 * r12;
        const double dot = fma(d0z, d1z, fma(d0y, d1y, d0x * d1x));
        const double o1 = dot / b1, o2 = dot / b2;
        const double cos2 = o1 * o2;
        double fm = a.kappa / mag;
        if (fabs(dot) > 0.000001 && a.sin0 > 0.0) {
            const double pref = fmax(sqrt(fmax(1.0 / cos2 - 1.0, 0.0)), 0.000001);
            fm = fm * (a.cos0 - a.sin0 * (1.0 / pref));
        }
        hx = (d0x * o1 - d1x) * fm; hy = (d0y * o1 - d1y) * fm; hz = (d0z * o1 - d1z) * fm;
        tx = (d0x - d1x * o2) * fm; ty = (d0y - d1y * o2) * fm; tz = (d0z - d1z * o2) * fm;
    }
    if (role == kRoleAngleHead) { cx = hx; cy = hy; cz = hz; }
    else if (role == kRoleAngleTail) { cx = tx; cy = ty; cz = tz; }
    else { cx = -(hx + tx); cy = -(hy + ty); cz = -(hz + tz); }
}

// Bonds per group of whole bonded beads: one lane per term computes its
// contribution (all loads of a group in flight at once), then one lane per bead
// adds its terms to the pair force in the per-polymer loop's order (bonds, then
// bond pairs; within a bond pair head, tail, middle) in FP64 and rounds once.
__global__ __launch_bounds__(kBondThreads) void k_bonds(DeviceState S, int nxt, int par, uint32_t n_blocks,
                                                       uint32_t n_worlds)
{
    __shared__ double s_c[3][kBondThreads];
    const uint32_t tid = threadIdx.x;
    const WorldCtr *__restrict__ ctr_in = S.ctr + par * n_worlds;
    if (blockIdx.x < n_blocks) {
        const BondBlock bk = S.bond_blocks[blockIdx.x];
        if (ctr_in[bk.world].remaining > 0) {                  // uniform per block
            const WorldDev *__restrict__ W = &S.worlds[bk.world];
            const float4 *__restrict__ pos = S.pos[nxt];
            const uint32_t t = bk.term_begin + tid;
            if (t < bk.term_end) {
                double cx, cy, cz;
                bond_term(S, pos, W, S.bead_terms[t], cx, cy, cz);
                s_c[0][tid] = cx;
                s_c[1][tid] = cy;
                s_c[2][tid] = cz;
            }
            __syncthreads();
            const uint32_t b = bk.bead_begin + tid;
            if (b < bk.bead_end) {
                const BondedBead bb = S.bonded[b];
                const uint32_t me = S.slot_of_bead[bb.bead];
                float4 f = S.frc[nxt][me];
                double ax = f.x, ay = f.y, az = f.z;
                for (uint32_t k = bb.t0 - bk.term_begin; k < bb.t1 - bk.term_begin; k++) {
                    ax += s_c[0][k];
                    ay += s_c[1][k];
                    az += s_c[2][k];
                }
                f.x = (float)ax;
                f.y = (float)ay;
                f.z = (float)az;
                if ((S.worlds[bk.world].stationary >> (fbits(pos[me].w) >> 28)) & 1u) f = make_float4(0, 0, 0, 0);
                S.frc[nxt][me] = f;
            }
        }
    }
    // step counters: read ctr[par], write ctr[par^1] (nothing reads that copy this step)
    const uint32_t i = blockIdx.x * blockDim.x + tid;
    if (i < n_worlds) {
        const WorldCtr c = S.ctr[par * n_worlds + i];
        WorldCtr o = c;
        if (c.remaining > 0) { o.t = c.t + 1; o.remaining = c.remaining - 1; }
        S.ctr[(par ^ 1) * n_worlds + i] = o;
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

// ------------------------------------------------------------ utility ----

__global__ void k_fill_u32(uint32_t *__restrict__ p, uint32_t n, uint32_t v)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_rebase_index(float4 *__restrict__ vel, uint32_t n, uint32_t delta)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float4 v = vel[i];
        v.w = __uint_as_float(__float_as_uint(v.w) + delta);
        vel[i] = v;
    }
}

// ------------------------------------------------------------ launchers ----

static inline uint32_t ceil_div(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

uint32_t scan_tiles(uint32_t n_cells) { return std::max<uint32_t>(1, ceil_div(n_cells, kScanBlock)); }

hipError_t launch_step(const StepLaunch &L, hipStream_t st, KernelEvents *ev)
{
    const DeviceState &S = L.S;
    const int cur = L.cur;
    const int par = L.cur;   // counters ping-pong with the buffers
    DeviceState Sp = S;
    Sp.ctr = S.ctr + par * L.n_worlds;       // counters read this step
    auto rec = [&](int k) { if (ev) (void)hipEventRecord(ev->ev[k], st); };
    const uint32_t nt = scan_tiles(L.n_cells);

    rec(0);
    k_integrate<<<std::max<uint32_t>(1, L.n_bead_blocks), kBeadThreads, 0, st>>>(Sp, cur, nt, L.n_bead_blocks);
    rec(1);
    k_scan<<<nt, kScanThreads, 0, st>>>(S.cell_count, L.n_cells, S.cell_start, S.scan_tiles, S.scan_ticket, nt,
                                         S.flags);
    rec(2);
    const uint32_t gb = ceil_div(L.n_beads, kBeadThreads);
    k_scatter<<<gb, kBeadThreads, 0, st>>>(S, cur, L.n_beads);
    rec(3);
    k_reorder<<<gb, kBeadThreads, 0, st>>>(S, cur, L.n_beads);
    rec(4);
    const uint32_t fg = L.force_grid ? std::min<uint32_t>(L.force_grid, L.n_force_blocks) : L.n_force_blocks;
    if (fg) k_force<<<fg, kForceThreads, 0, st>>>(Sp, cur ^ 1, L.n_force_blocks);
    rec(5);
    if (L.n_stationary)
        k_stationary<<<ceil_div(L.n_stationary, kBeadThreads), kBeadThreads, 0, st>>>(Sp, cur ^ 1, L.n_stationary);
    const uint32_t bg = std::max<uint32_t>(L.n_bond_blocks, ceil_div(L.n_worlds, kBondThreads));
    k_bonds<<<bg, kBondThreads, 0, st>>>(S, cur ^ 1, par, L.n_bond_blocks, L.n_worlds);
    rec(6);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_fill_u32<<<ceil_div(n, 256), 256, 0, st>>>(p, n, v);
    return hipGetLastError();
}

hipError_t launch_rebase_index(float4 *vel, uint32_t n, uint32_t delta, hipStream_t st)
{
    if (n == 0) return hipSuccess;
    k_rebase_index<<<ceil_div(n, 256), 256, 0, st>>>(vel, n, delta);
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
