// kernels.hip -- gfx950 kernels of one DPD time step.
//
// Step (one set of launches advances every resident world):
//   k_integrate   finish-kick + drift + half-kick + wrap, cell key, arrival rank
//                 (dpd_maths_core_half_step_raw.hpp:42-72,114-124 fused as
//                 dpd_engine_avx2_half_merge_tbb.hpp:780-800 does)
//   k_scan_*      exclusive prefix sum of cell counts -> cell starts (counting sort)
//   k_scatter     sorted position of every bead's slot
//   k_reorder     stable gather into the other buffer, bead-index order in a cell
//   k_force       pair forces: a brick of cells + its halo staged in LDS, one lane
//                 per home bead over the 27-cell stencil, register accumulation,
//                 no atomics (dpd_maths_core_half_step_raw.hpp:129-216; the RNG of
//                 hash.hpp:61-88 + dpd_maths_core.hpp:55-79)
//   k_bonds       Hookean + angle bonds per polymer, then the per-world step
//                 counter (naive_dpd_engine_half_merge.hpp:638-690,
//                 dpd_maths_core_half_step.hpp:178-231)
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

// Exclusive scan of one value per thread across a 256-thread block (4 wave64).
// Returns the exclusive prefix; *total receives the block sum.
__device__ __forceinline__ uint32_t block_scan_256(uint32_t v, uint32_t *s_wsum, uint32_t *total)
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
    for (int k = 0; k < 4; k++) {
        uint32_t s = s_wsum[k];
        if (k < wid) base += s;
        tot += s;
    }
    __syncthreads();
    *total = tot;
    return base + x - v;
}

// ---------------------------------------------------------- integrate ----

__global__ __launch_bounds__(kBeadThreads) void k_integrate(DeviceState S, int cur)
{
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
    int c[3];
    bool bad = false;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const float Lf = W->Lf[d];
        bad |= !(x[d] >= 0.0f && x[d] < Lf);
        int cc = (int)floorf(x[d]);
        cc = cc < 0 ? 0 : cc;
        cc = cc >= W->L[d] ? W->L[d] - 1 : cc;
        c[d] = cc;
    }
    if (bad) atomicOr(S.flags, PDPD_FLAG_BAD_POSITION);
    const uint32_t key = W->cell_off + (uint32_t)((c[0] * W->L[1] + c[1]) * W->L[2] + c[2]);
    const uint32_t r = atomicAdd(&S.cell_count[key], 1u);
    S.key[s] = key;
    S.rank[s] = r | (active ? 0u : kInactive);
}

// --------------------------------------------------------------- scan ----

__global__ __launch_bounds__(kScanThreads) void k_scan_reduce(const uint32_t *__restrict__ count, uint32_t n,
                                                             uint32_t *__restrict__ block_sums)
{
    __shared__ uint32_t s_wsum[4];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        uint32_t i = base + k;
        acc += (i < n) ? count[i] : 0u;
    }
    uint32_t tot;
    block_scan_256(acc, s_wsum, &tot);
    if (threadIdx.x == 0) block_sums[blockIdx.x] = tot;
}

// One block: exclusive scan of the block sums in place; writes the grand total
// to cell_start[n_cells].
__global__ __launch_bounds__(kScanThreads) void k_scan_top(uint32_t *__restrict__ block_sums, uint32_t nb,
                                                          uint32_t *__restrict__ cell_start, uint32_t n_cells)
{
    __shared__ uint32_t s_wsum[4];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nb; base += kScanThreads) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = (i < nb) ? block_sums[i] : 0u;
        uint32_t tot;
        uint32_t ex = block_scan_256(v, s_wsum, &tot);
        if (i < nb) block_sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) cell_start[n_cells] = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_down(uint32_t *__restrict__ count, uint32_t n,
                                                           const uint32_t *__restrict__ block_sums,
                                                           uint32_t *__restrict__ cell_start)
{
    __shared__ uint32_t s_wsum[4];
    const uint32_t base = blockIdx.x * kScanBlock + threadIdx.x * kScanPerThread;
    uint32_t v[kScanPerThread];
    uint32_t acc = 0;
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        uint32_t i = base + k;
        v[k] = (i < n) ? count[i] : 0u;
        acc += v[k];
    }
    uint32_t tot;
    uint32_t ex = block_scan_256(acc, s_wsum, &tot) + block_sums[blockIdx.x];
#pragma unroll
    for (int k = 0; k < kScanPerThread; k++) {
        uint32_t i = base + k;
        if (i < n) {
            cell_start[i] = ex;
            count[i] = 0u;          // ready for the next step's arrivals
        }
        ex += v[k];
    }
}

// ------------------------------------------------------- sort scatter ----

__global__ __launch_bounds__(kBeadThreads) void k_scatter(DeviceState S, uint32_t n)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t r = S.rank[s];
    S.perm[S.cell_start[S.key[s]] + (r & ~kInactive)] = s | (r & kInactive);
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
    const float4 v = S.vel[cur][s];
    const uint32_t me = fbits(v.w);
    uint32_t rank = 0;
    for (uint32_t q = a; q < b; q++) {
        const uint32_t s2 = S.perm[q] & ~kInactive;
        rank += fbits(S.vel[cur][s2].w) < me ? 1u : 0u;
    }
    const uint32_t dst = a + rank;
    S.pos[nxt][dst] = S.pos[cur][s];
    S.vel[nxt][dst] = v;
    if (e & kInactive) S.frc[nxt][dst] = S.frc[cur][s];
    S.slot_of_bead[me] = dst;
}

// --------------------------------------------------------- pair force ----

__global__ __launch_bounds__(kForceThreads) void k_force(DeviceState S, int nxt, uint32_t n_items)
{
    __shared__ float4 s_pos[kTileBeads];
    __shared__ float4 s_vel[kTileBeads];
    __shared__ uint32_t s_cstart[kMaxHaloCells + 1];
    __shared__ uint32_t s_cglob[kMaxHaloCells];
    __shared__ uint32_t s_hstart[kMaxInnerCells + 1];
    __shared__ float s_a[kMaxTypes * kMaxTypes];
    __shared__ float s_sg[kMaxTypes * kMaxTypes];
    __shared__ uint32_t s_wsum[4];
    __shared__ uint16_t s_list[kPairList * kForceThreads];

    const int tid = threadIdx.x;
    const WorldCtr *__restrict__ ctr = S.ctr;   // this step's counters
    int cur_world = -1;

    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        const BrickItem it = S.bricks[item];
        if (ctr[it.world].remaining == 0) continue;
        const WorldDev *__restrict__ W = &S.worlds[it.world];
        const int L0 = W->L[0], L1 = W->L[1], L2 = W->L[2];
        const int n0 = min(W->brick[0], L0 - it.bx);
        const int n1 = min(W->brick[1], L1 - it.by);
        const int n2 = min(W->brick[2], L2 - it.bz);
        const int h1 = n1 + 2, h2 = n2 + 2;
        const int nh = (n0 + 2) * h1 * h2;
        const int ni = n0 * n1 * n2;

        __syncthreads();    // previous item is done with LDS
        if (it.world != cur_world) {
            s_a[tid] = W->a[tid];
            s_sg[tid] = W->sg[tid];
            cur_world = it.world;
        }

        // 1. halo cell table: global start and count of every halo cell
        uint32_t cnt[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int hc = tid * 2 + q;
            cnt[q] = 0;
            if (hc < nh) {
                const int iz = hc % h2, iy = (hc / h2) % h1, ix = hc / (h1 * h2);
                int gx = it.bx - 1 + ix, gy = it.by - 1 + iy, gz = it.bz - 1 + iz;
                gx += gx < 0 ? L0 : 0; gx -= gx >= L0 ? L0 : 0;
                gy += gy < 0 ? L1 : 0; gy -= gy >= L1 ? L1 : 0;
                gz += gz < 0 ? L2 : 0; gz -= gz >= L2 ? L2 : 0;
                const uint32_t c = W->cell_off + (uint32_t)((gx * L1 + gy) * L2 + gz);
                const uint32_t st = S.cell_start[c];
                s_cglob[hc] = st;
                cnt[q] = S.cell_start[c + 1] - st;
            }
        }
        uint32_t total;
        const uint32_t ex = block_scan_256(cnt[0] + cnt[1], s_wsum, &total);
        if (tid * 2 < nh) s_cstart[tid * 2] = ex;
        if (tid * 2 + 1 < nh) s_cstart[tid * 2 + 1] = ex + cnt[0];
        if (tid == 0) s_cstart[nh] = total;
        __syncthreads();

        // inner (home) cells: prefix of their counts
        uint32_t hcnt = 0;
        int my_hc = 0;
        if (tid < ni) {
            const int jz = tid % n2, jy = (tid / n2) % n1, jx = tid / (n1 * n2);
            my_hc = ((jx + 1) * h1 + (jy + 1)) * h2 + (jz + 1);
            hcnt = s_cstart[my_hc + 1] - s_cstart[my_hc];
        }
        uint32_t n_home;
        const uint32_t hex = block_scan_256(hcnt, s_wsum, &n_home);
        if (tid < ni) s_hstart[tid] = hex;
        if (tid == 0) s_hstart[ni] = n_home;

        if (total > (uint32_t)kTileBeads) {           // uniform: skip, host reports
            if (tid == 0) atomicOr(S.flags, PDPD_FLAG_TILE_OVERFLOW);
            continue;
        }

        // 2. stage halo beads (positions shifted by the periodic image)
        for (uint32_t j = tid; j < total; j += kForceThreads) {
            int lo = 0, hi = nh;                       // last hc with s_cstart[hc] <= j
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_cstart[mid] <= j) lo = mid; else hi = mid;
            }
            const int hc = lo;
            const uint32_t g = s_cglob[hc] + (j - s_cstart[hc]);
            const int iz = hc % h2, iy = (hc / h2) % h1, ix = hc / (h1 * h2);
            const int gx = it.bx - 1 + ix, gy = it.by - 1 + iy, gz = it.bz - 1 + iz;
            const float sx = gx < 0 ? -W->Lf[0] : (gx >= L0 ? W->Lf[0] : 0.0f);
            const float sy = gy < 0 ? -W->Lf[1] : (gy >= L1 ? W->Lf[1] : 0.0f);
            const float sz = gz < 0 ? -W->Lf[2] : (gz >= L2 ? W->Lf[2] : 0.0f);
            float4 p = S.pos[nxt][g];
            p.x += sx; p.y += sy; p.z += sz;
            s_pos[j] = p;
            s_vel[j] = S.vel[nxt][g];
        }
        __syncthreads();

        // 3. one lane per home bead, 27 cells as 9 contiguous z-rows of 3 cells.
        //    Candidates within the cutoff are first collected into a per-lane
        //    list in LDS, then evaluated: the expensive pair body runs once per
        //    accepted pair instead of once per candidate of any lane in the wave.
        //    The list keeps candidate order, so the sum order is unchanged.
        const WorldCtr wc = ctr[it.world];
        const uint64_t th = step_hash(wc.t, W->seed_lo, W->seed_hi);
        const uint32_t th_lo = (uint32_t)th, th_hi = (uint32_t)(th >> 32);
        const float scale = W->rng_scale;
        for (uint32_t k = tid; k < n_home; k += kForceThreads) {
            int lo = 0, hi = ni;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (s_hstart[mid] <= k) lo = mid; else hi = mid;
            }
            const int ic = lo;
            const int jz = ic % n2, jy = (ic / n2) % n1, jx = ic / (n1 * n2);
            const int hc = ((jx + 1) * h1 + (jy + 1)) * h2 + (jz + 1);
            const uint32_t off = k - s_hstart[ic];
            const uint32_t li = s_cstart[hc] + off;
            const float4 hp = s_pos[li];
            const float4 hv = s_vel[li];
            const uint32_t hh = fbits(hp.w);
            const uint32_t trow = (hh >> 28) * kMaxTypes;
            float fx = 0.0f, fy = 0.0f, fz = 0.0f;
            uint32_t cnt = 0;
            auto flush = [&]() {
                for (uint32_t q = 0; q < cnt; q++) {
                    const uint32_t m = s_list[q * kForceThreads + tid];
                    const float4 op = s_pos[m];
                    const float4 ov = s_vel[m];
                    const float dx = hp.x - op.x, dy = hp.y - op.y, dz = hp.z - op.z;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float rr = sqrtf(r2);
                    const float inv_r = 1.0f / rr;
                    const float wr = 1.0f - rr;
                    const float ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
                    const float dvx = hv.x - ov.x, dvy = hv.y - ov.y, dvz = hv.z - ov.z;
                    const float rdotv = fmaf(ez, dvz, fmaf(ey, dvy, ex * dvx));
                    const uint32_t oh = fbits(op.w);
                    const uint32_t ti = trow + (oh >> 28);
                    const float sgw = s_sg[ti] * wr;
                    const float u = rng_uniform(rng_sym(th_lo, th_hi, hh, oh));
                    const float F = fmaf(s_a[ti], wr, sgw * fmaf(-sgw, rdotv, scale * u));
                    fx = fmaf(ex, F, fx);
                    fy = fmaf(ey, F, fy);
                    fz = fmaf(ez, F, fz);
                }
                cnt = 0;
            };
#pragma unroll 1
            for (int r = 0; r < 9; r++) {
                const int ddx = r / 3 - 1, ddy = r % 3 - 1;
                const int rb = hc + ddx * h1 * h2 + ddy * h2;
                const uint32_t m0 = s_cstart[rb - 1], m1 = s_cstart[rb + 2];
                for (uint32_t m = m0; m < m1; m++) {
                    const float4 op = s_pos[m];
                    const float dx = hp.x - op.x, dy = hp.y - op.y, dz = hp.z - op.z;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    if (r2 < 1.0f && r2 >= kMinR2 && m != li) {
                        s_list[cnt * kForceThreads + tid] = (uint16_t)m;
                        if (++cnt == kPairList) flush();
                    }
                }
            }
            flush();
            S.frc[nxt][s_cglob[hc] + off] = make_float4(fx, fy, fz, 0.0f);
        }
    }
}

// -------------------------------------------------------------- bonds ----

__device__ __forceinline__ float min_image(float d, float len)
{
    if (d > 0.5f * len) d -= len;
    else if (d < -0.5f * len) d += len;
    return d;
}

__global__ __launch_bounds__(kBeadThreads) void k_bonds(DeviceState S, int nxt, int par, uint32_t n_polys,
                                                       uint32_t n_worlds)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const WorldCtr *__restrict__ ctr_in = S.ctr + par * n_worlds;
    if (p < n_polys) {
        const PolyItem it = S.polys[p];
        if (ctr_in[it.world].remaining > 0) {
            const WorldDev *__restrict__ W = &S.worlds[it.world];
            const float Lx = W->Lf[0], Ly = W->Lf[1], Lz = W->Lf[2];
            float4 *__restrict__ frc = S.frc[nxt];
            const float4 *__restrict__ pos = S.pos[nxt];
            for (uint32_t bi = it.bond_begin; bi < it.bond_end; bi++) {
                const BondRec b = S.bonds[bi];
                const uint32_t sh = S.slot_of_bead[b.head], st = S.slot_of_bead[b.tail];
                const float4 xh = pos[sh], xt = pos[st];
                const float dx = min_image(xt.x - xh.x, Lx);
                const float dy = min_image(xt.y - xh.y, Ly);
                const float dz = min_image(xt.z - xh.z, Lz);
                const float r = sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx)));
                const float hf = b.kappa * (b.r0 - r);
                const float s = hf / r;
                float4 fh = frc[sh];
                fh.x -= dx * s; fh.y -= dy * s; fh.z -= dz * s;
                frc[sh] = fh;
                float4 ft = frc[st];
                ft.x += dx * s; ft.y += dy * s; ft.z += dz * s;
                frc[st] = ft;
            }
            for (uint32_t ai = it.angle_begin; ai < it.angle_end; ai++) {
                const AngleRec a = S.angles[ai];
                const uint32_t sh = S.slot_of_bead[a.head], sm = S.slot_of_bead[a.mid], st = S.slot_of_bead[a.tail];
                const float4 xh = pos[sh], xm = pos[sm], xt = pos[st];
                const float d0x = min_image(xm.x - xh.x, Lx), d0y = min_image(xm.y - xh.y, Ly),
                            d0z = min_image(xm.z - xh.z, Lz);
                const float d1x = min_image(xt.x - xm.x, Lx), d1y = min_image(xt.y - xm.y, Ly),
                            d1z = min_image(xt.z - xm.z, Lz);
                const float r01 = sqrtf(fmaf(d0z, d0z, fmaf(d0y, d0y, d0x * d0x)));
                const float r12 = sqrtf(fmaf(d1z, d1z, fmaf(d1y, d1y, d1x * d1x)));
                const float mag = r01 * r12;
                float hx = 0.f, hy = 0.f, hz = 0.f, tx = 0.f, ty = 0.f, tz = 0.f, mx = 0.f, my = 0.f, mz = 0.f;
                if (mag > 0.0001f) {
                    const float b1 = r01 * r01, b2 = r12 * r12;
                    const float dot = fmaf(d0z, d1z, fmaf(d0y, d1y, d0x * d1x));
                    const float o1 = dot / b1, o2 = dot / b2;
                    const float cos2 = o1 * o2;
                    float fm = a.kappa / mag;
                    if (fabsf(dot) > 0.000001f && a.sin0 > 0.0f) {
                        const float ip = 1.0f / sqrtf(1.0f / cos2 - 1.0f);
                        fm = fm * (a.cos0 - a.sin0 * ip);
                    }
                    hx = (d0x * o1 - d1x) * fm; hy = (d0y * o1 - d1y) * fm; hz = (d0z * o1 - d1z) * fm;
                    tx = (d0x - d1x * o2) * fm; ty = (d0y - d1y * o2) * fm; tz = (d0z - d1z * o2) * fm;
                    mx = -(hx + tx); my = -(hy + ty); mz = -(hz + tz);
                }
                float4 f;
                f = frc[sh]; f.x += hx; f.y += hy; f.z += hz; frc[sh] = f;
                f = frc[st]; f.x += tx; f.y += ty; f.z += tz; frc[st] = f;
                f = frc[sm]; f.x += mx; f.y += my; f.z += mz; frc[sm] = f;
            }
        }
    }
    // step counters: read ctr[par], write ctr[par^1] (nothing reads that copy this step)
    if (p < n_worlds) {
        const WorldCtr c = S.ctr[par * n_worlds + p];
        WorldCtr o = c;
        if (c.remaining > 0) { o.t = c.t + 1; o.remaining = c.remaining - 1; }
        S.ctr[(par ^ 1) * n_worlds + p] = o;
    }
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

hipError_t launch_step(const StepLaunch &L, hipStream_t st, KernelEvents *ev)
{
    const DeviceState &S = L.S;
    const int cur = L.cur;
    const int par = L.cur;   // counters ping-pong with the buffers
    DeviceState Sp = S;
    Sp.ctr = S.ctr + par * L.n_worlds;       // counters read this step
    auto rec = [&](int k) { if (ev) (void)hipEventRecord(ev->ev[k], st); };

    rec(0);
    k_integrate<<<L.n_bead_blocks, kBeadThreads, 0, st>>>(Sp, cur);
    rec(1);
    const uint32_t nb = ceil_div(L.n_cells, kScanBlock);
    k_scan_reduce<<<nb, kScanThreads, 0, st>>>(S.cell_count, L.n_cells, S.block_sums);
    k_scan_top<<<1, kScanThreads, 0, st>>>(S.block_sums, nb, S.cell_start, L.n_cells);
    k_scan_down<<<nb, kScanThreads, 0, st>>>(S.cell_count, L.n_cells, S.block_sums, S.cell_start);
    rec(2);
    const uint32_t gb = ceil_div(L.n_beads, kBeadThreads);
    k_scatter<<<gb, kBeadThreads, 0, st>>>(S, L.n_beads);
    rec(3);
    k_reorder<<<gb, kBeadThreads, 0, st>>>(S, cur, L.n_beads);
    rec(4);
    const uint32_t fg = L.force_blocks ? std::min<uint32_t>(L.force_blocks, L.n_bricks) : L.n_bricks;
    if (L.n_bricks) k_force<<<fg, kForceThreads, 0, st>>>(Sp, cur ^ 1, L.n_bricks);
    rec(5);
    const uint32_t pw = std::max<uint32_t>(L.n_polys, L.n_worlds);
    k_bonds<<<ceil_div(pw, kBeadThreads), kBeadThreads, 0, st>>>(S, cur ^ 1, par, L.n_polys, L.n_worlds);
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
