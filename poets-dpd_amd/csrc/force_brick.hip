// force_brick.hip -- pair forces by cell bricks staged in LDS.
//
// A workgroup takes one brick of cells of one world (BX x BY x BZ cells,
// clipped at the world's far faces) and stages the brick and its one-cell
// halo -- (bx+2)(by+2)(bz+2) cells, periodic images included -- into LDS with
// coalesced loads: positions with the image shift already added (x + s, the
// FP32 addition the pair test would make), hash bits, and the half-step
// velocities. Every candidate the brick's home beads test then comes from
// LDS.
//
// Work is one task per (moving home bead, stencil x-plane): three rows of the
// plane (dy = -1, 0, +1), each row one contiguous LDS range, the z cells
// z-1, z, z+1 of the halo row (cells wholly beyond the cutoff skipped, as in
// the per-slot kernel). Tasks are sorted by candidate count with a counting
// sort in LDS, so the 64 lanes of a wave test streams of nearly equal
// length. Accepted candidates go to a per-wave queue; when it fills, all lanes
// evaluate queued pairs side by side (dpd_maths_core_half_step_raw.hpp:129-216,
// RNG of hash.hpp:61-88 + dpd_maths_core.hpp:55-79) and each lane then adds its
// own pairs to its plane sum in acceptance order.
//
// Summation order per bead (the oracle's): one running sum per stencil
// plane over its rows' candidates in row, z-cell and slot order; planes
// (p0 + p1) + p2; then the FP64 bonded sum, rounded once. Which lane or wave
// evaluates a pair and when the queue flushes change nothing in it, so forces
// are bit-identical to the per-slot kernel's and the oracle's.
//
// A brick whose halo or home beads exceed the LDS capacity (dense regions)
// takes a slower path in the same workgroup: its home cells one at a time, up
// to one bead per thread, and each stencil row staged in chunks that fit, the
// row sums carried across chunks. Any density is handled.
#include "device_types.h"
#include "kernels.h"
#include "pdpd_hip.h"
#include "step_device.h"

namespace pdpd {

namespace {

constexpr int kBins = 64;          // counting-sort bins of task lengths (longer tasks share the last)
constexpr int kList = 16;          // accepted pairs listed per lane between flushes (queue positions, 8 bits)
constexpr int kUnroll = 4;         // candidates per lane per iteration

// Halo cell (i, j, k), k fastest, of a halo of dims hy x hz per plane.
__device__ __forceinline__ int halo_cell(int i, int j, int k, int hy, int hz) { return (i * hy + j) * hz + k; }

// Image shift of a halo coordinate g = origin - 1 + i of a world side L.
__device__ __forceinline__ float image_shift(int g, int L, float Lf)
{
    return g < 0 ? -Lf : (g >= L ? Lf : 0.0f);
}

__device__ __forceinline__ int wrap(int g, int L) { return g < 0 ? g + L : (g >= L ? g - L : g); }

}  // namespace

template <int BX, int BY, int BZ, int NW, int HCAP, int HOMECAP, int QCAP>
__global__ __launch_bounds__(NW * 64) void k_force_brick(DeviceState S, int nxt, uint32_t n_bricks,
                                                         const BrickRec *__restrict__ bricks,
                                                         const NextBuffers *next)
{
    constexpr int NT = NW * 64;
    constexpr int HX = BX + 2, HY = BY + 2, HZ = BZ + 2;
    constexpr int HC = HX * HY * HZ;                 // halo cells (strides fixed by the shape; clipped cells are empty)
    constexpr int NHC = BX * BY * BZ;                // home cells
    constexpr int CPT = (HC + NT - 1) / NT;          // halo cells per thread
    constexpr int HPT = (NHC + NT - 1) / NT;         // home cells per thread
    constexpr int kDense = NT;                       // dense path: home beads per pass (one per thread)
    static_assert(QCAP >= 64 && QCAP <= 256, "queue positions are 8-bit; an emptied queue holds one candidate per lane");
    static_assert(kList >= kUnroll, "a lane lists up to kUnroll pairs per batch");
    static_assert(HCAP < 65536 && 3 * HOMECAP < 65536, "16-bit LDS indices");
    static_assert(HCAP >= kDense + 64, "the dense path stages home beads and a chunk of candidates");

    __shared__ float4 s_pos[HCAP];                   // x, y, z (+ image shift), hash bits
    __shared__ float4 s_vel[HCAP];                   // half-step velocity, bead index bits
    __shared__ uint32_t s_cstart[HC + 1];            // LDS start of each halo cell
    __shared__ uint32_t s_gstart[HC];                // global slot of each halo cell's first bead
    __shared__ uint16_t s_hidx[HOMECAP];             // home bead -> LDS index
    __shared__ uint16_t s_hcell[HOMECAP];            // home bead -> halo cell
    __shared__ uint16_t s_task[3 * HOMECAP];         // tasks sorted by length, longest first
    __shared__ uint32_t s_bin[kBins];
    __shared__ float s_part[3][3][HOMECAP];          // plane sums [plane][axis][home bead]
    __shared__ uint32_t s_q[NW][QCAP];               // queued pairs: candidate | home << 16 (LDS indices)
    __shared__ float4 s_qres[NW][QCAP];              // their (unit vector, force magnitude)
    __shared__ uint8_t s_list[kList][NT];            // per lane its queued pairs (queue positions) in order
    __shared__ float s_a[kMaxTypes * kMaxTypes], s_sg[kMaxTypes * kMaxTypes];
    __shared__ uint32_t s_wsum[NW];
    __shared__ uint32_t s_misc[4];                   // -, -, tasks, next chunk

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    const float4 *__restrict__ gpos = S.pos[nxt];
    const float4 *__restrict__ gvel = S.vel[nxt];
    float4 *__restrict__ gfrc = S.frc[nxt];
    const uint32_t *__restrict__ cs = S.cell_start;
    uint32_t *__restrict__ qs = s_q[wave];
    float4 *__restrict__ qr = s_qres[wave];

    if (next && blockIdx.x == 0) {                  // for the next step's scan
        const NextBuffers B = next_buffers(next);
        for (uint32_t i = (uint32_t)tid; i < B.n_tiles; i += NT) B.scan_tiles[i] = 0ull;
        if (tid == 0) *B.scan_ticket = 0u;
    }

    for (uint32_t li = blockIdx.x; li < n_bricks; li += gridDim.x) {
        const BrickRec br = bricks[xcd_chunk(li, n_bricks)];
        if (S.ctr[br.world].remaining == 0) continue;       // uniform: frozen world (never in a chain)
        const WorldDev *__restrict__ W = &S.worlds[br.world];
        const int x0 = (int)(br.origin & 1023u), y0 = (int)((br.origin >> 10) & 1023u), z0 = (int)(br.origin >> 20);
        const int bx = (int)(br.dims & 255u), by = (int)((br.dims >> 8) & 255u), bz = (int)((br.dims >> 16) & 255u);
        const int L0 = W->L[0], L1 = W->L[1], L2 = W->L[2];
        const float Lx = W->Lf[0], Ly = W->Lf[1], Lz = W->Lf[2];
        {   // the world's n_types^2 strengths (rows of kMaxTypes)
            const int nt = (int)W->n_types;
            for (int i = tid; i < nt * nt; i += NT) {
                const int k = (i / nt) * kMaxTypes + i % nt;
                s_a[k] = W->a[k];
                s_sg[k] = W->sg[k];
            }
        }
        const WorldCtr wc = S.ctr[br.world];
        const uint64_t th = step_hash(wc.t, W->seed_lo, W->seed_hi);
        const uint32_t th_lo = (uint32_t)th, th_hi = (uint32_t)(th >> 32);
        const float scale = W->rng_scale;

        // ---- halo cells (thread c: cells c * CPT ...): global starts, counts, LDS starts
        uint32_t cnt[CPT];
        uint32_t mysum = 0;
#pragma unroll
        for (int q = 0; q < CPT; q++) {
            const int c = tid * CPT + q;
            const int k = c % HZ, j = (c / HZ) % HY, i = c / (HZ * HY);
            cnt[q] = 0;
            if (c < HC && i < bx + 2 && j < by + 2 && k < bz + 2) {
                const int gx = wrap(x0 - 1 + i, L0), gy = wrap(y0 - 1 + j, L1), gz = wrap(z0 - 1 + k, L2);
                const uint32_t id = W->cell_off + (uint32_t)((gx * L1 + gy) * L2 + gz);
                const uint32_t a = cs[id], b = cs[id + 1];
                s_gstart[c] = a;
                cnt[q] = b - a;
            }
            mysum += cnt[q];
        }
        uint32_t total;
        const uint32_t ex0 = block_scan<NW>(mysum, s_wsum, &total);
        {
            uint32_t ex = ex0;
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = tid * CPT + q;
                if (c < HC) s_cstart[c] = ex;
                ex += cnt[q];
            }
        }
        if (tid == 0) s_cstart[HC] = total;
        __syncthreads();
        // home beads: home cells (1..bx, 1..by, 1..bz) in halo order
        uint32_t hcnt[HPT];
        uint32_t hsum = 0;
#pragma unroll
        for (int q = 0; q < HPT; q++) {
            const int e = tid * HPT + q;
            const int k = e % BZ, j = (e / BZ) % BY, i = e / (BZ * BY);
            hcnt[q] = 0;
            if (e < NHC && i < bx && j < by && k < bz) {
                const int c = halo_cell(i + 1, j + 1, k + 1, HY, HZ);
                hcnt[q] = s_cstart[c + 1] - s_cstart[c];
            }
            hsum += hcnt[q];
        }
        uint32_t n_home;
        const uint32_t hex0 = block_scan<NW>(hsum, s_wsum, &n_home);
        const bool dense = total > (uint32_t)HCAP || n_home > (uint32_t)HOMECAP;   // uniform

        auto eval_pair = [&](uint32_t e) -> float4 {
            const uint32_t c = e & 0xffffu, hidx = e >> 16;
            const float4 p = s_pos[hidx], v = s_vel[hidx], op = s_pos[c], ov = s_vel[c];
            const float dx = p.x - op.x, dy = p.y - op.y, dz = p.z - op.z;
            const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
            const float rr = sqrtf(r2);
            const float inv_r = 1.0f / rr;
            const float wr = 1.0f - rr;
            const float ux = dx * inv_r, uy = dy * inv_r, uz = dz * inv_r;
            const float dvx = v.x - ov.x, dvy = v.y - ov.y, dvz = v.z - ov.z;
            const float rdotv = fmaf(uz, dvz, fmaf(uy, dvy, ux * dvx));
            const uint32_t ph = fbits(p.w), oh = fbits(op.w);
            const uint32_t ti = (ph >> 28) * kMaxTypes + (oh >> 28);
            const float a = s_a[ti], sg = s_sg[ti];
            const float sgw = sg * wr;
            const float u = rng_uniform(rng_sym(th_lo, th_hi, ph, oh));
            const float F = fmaf(a, wr, sgw * fmaf(-sgw, rdotv, scale * u));
            return make_float4(ux, uy, uz, F);
        };

        // Per-lane stream state of the current task: its plane sum and its
        // queued pairs (s_list). The queue fill is the same in every lane of a wave.
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        uint32_t listed = 0, queued = 0;
        auto flush = [&]() {
            wave_sync();
            for (uint32_t i = (uint32_t)lane; i < queued; i += 64) qr[i] = eval_pair(qs[i]);
            wave_sync();
            for (uint32_t c = 0; c < listed; c++) {
                const float4 q = qr[s_list[c][tid]];
                px = fmaf(q.x, q.w, px);
                py = fmaf(q.y, q.w, py);
                pz = fmaf(q.z, q.w, pz);
            }
            queued = 0;
            listed = 0;
            wave_sync();
        };
        // One lane's candidate stream: up to three LDS ranges (its rows) in
        // order, tested against home bead hidx at (hpx, hpy, hpz). Called by
        // the whole wave; an inactive lane passes empty ranges.
        auto stream = [&](uint32_t hidx, float hpx, float hpy, float hpz, uint32_t a0, uint32_t n0, uint32_t a1,
                          uint32_t n1, uint32_t a2, uint32_t n2) {
            const uint32_t n01 = n0 + n1, len = n01 + n2;
            const uint32_t b1 = a1 - n0, b2 = a2 - n01;     // candidate q of row 1 / 2 at q + b1 / q + b2
            const uint32_t hkey = hidx << 16;
            const uint32_t wlen = __builtin_amdgcn_readfirstlane(wave_max(len));
            for (uint32_t k = 0; k < wlen; k += kUnroll) {
                uint32_t idx[kUnroll];
                float4 pc[kUnroll];
#pragma unroll
                for (int j = 0; j < kUnroll; j++) {
                    const uint32_t q = k + j;
                    const uint32_t at = q + (q < n0 ? a0 : (q < n01 ? b1 : b2));
                    idx[j] = q < len ? at : hidx;          // beyond the stream: the home bead, never accepted
                    pc[j] = s_pos[idx[j]];
                }
                unsigned long long bal[kUnroll];
                uint32_t tot = 0;
#pragma unroll
                for (int j = 0; j < kUnroll; j++) {
                    const float dx = hpx - pc[j].x, dy = hpy - pc[j].y, dz = hpz - pc[j].z;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    bal[j] = __builtin_amdgcn_ballot_w64(r2 < 1.0f && r2 >= kMinR2);
                    tot += (uint32_t)__popcll(bal[j]);
                }
                if (tot == 0) continue;
                if (__builtin_amdgcn_ballot_w64(listed + kUnroll > (uint32_t)kList)) flush();
#pragma unroll
                for (int j = 0; j < kUnroll; j++) {
                    const uint32_t nj = (uint32_t)__popcll(bal[j]);
                    if (nj == 0) continue;
                    if (queued + nj > (uint32_t)QCAP) flush();
                    if ((bal[j] >> lane) & 1ull) {
                        const uint32_t qp = queued + (uint32_t)__popcll(bal[j] & below);
                        qs[qp] = idx[j] | hkey;
                        s_list[listed++][tid] = (uint8_t)qp;
                    }
                    queued += nj;
                }
            }
        };
        // The force of one home bead from its plane sums (+ bonded terms),
        // stored at its slot; in a chain, its integration for the next step.
        // Called by every lane of a wave, the valid lanes a prefix.
        auto finish_bead = [&](bool valid, uint32_t slot, float4 p, float4 v, bool moving, float fx, float fy,
                               float fz) {
            if (valid) {
                if (!moving) {
                    fx = fy = fz = 0.0f;
                } else if (!((fbits(p.w) >> 27) & 1u)) {       // a polymer bead (monomers have no bonded terms)
                    double bx_ = 0.0, by_ = 0.0, bz_ = 0.0;
                    if (bonded_sum(S, fbits(v.w), bx_, by_, bz_)) {     // pair force + bonded terms, rounded once
                        fx = (float)((double)fx + bx_);
                        fy = (float)((double)fy + by_);
                        fz = (float)((double)fz + bz_);
                    }
                }
                gfrc[slot] = make_float4(fx, fy, fz, 0.0f);
                if (const uint32_t ff = force_flags(W, fx, fy, fz)) atomicOr(&S.wflags[br.world], ff);
            }
            if (next) {
                const NextBuffers B = next_buffers(next);
                advance_next(B, W, (uint32_t)br.world, slot, valid, p, v, make_float4(fx, fy, fz, 0.0f));
            }
        };

        if (!dense) {
            // ---- stage the halo: each thread copies its cells' beads, image shifts added
#pragma unroll
            for (int q = 0; q < CPT; q++) {
                const int c = tid * CPT + q;
                if (c < HC && cnt[q]) {
                    const int k = c % HZ, j = (c / HZ) % HY, i = c / (HZ * HY);
                    const float sx = image_shift(x0 - 1 + i, L0, Lx), sy = image_shift(y0 - 1 + j, L1, Ly),
                                sz = image_shift(z0 - 1 + k, L2, Lz);
                    const uint32_t g = s_gstart[c], d = s_cstart[c];
                    for (uint32_t r = 0; r < cnt[q]; r++) {
                        float4 p = gpos[g + r];
                        const float4 v = gvel[g + r];
                        p.x = p.x + sx;
                        p.y = p.y + sy;
                        p.z = p.z + sz;
                        s_pos[d + r] = p;
                        s_vel[d + r] = v;
                    }
                }
            }
            // home bead tables
            {
                uint32_t hex = hex0;
#pragma unroll
                for (int q = 0; q < HPT; q++) {
                    const int e = tid * HPT + q;
                    if (e < NHC && hcnt[q]) {
                        const int k = e % BZ, j = (e / BZ) % BY, i = e / (BZ * BY);
                        const int c = halo_cell(i + 1, j + 1, k + 1, HY, HZ);
                        const uint32_t first = s_cstart[c];
                        for (uint32_t r = 0; r < hcnt[q]; r++) {
                            s_hidx[hex + r] = (uint16_t)(first + r);
                            s_hcell[hex + r] = (uint16_t)c;
                        }
                    }
                    hex += hcnt[q];
                }
            }
            for (int i = tid; i < kBins; i += NT) s_bin[i] = 0u;
            if (tid == 0) s_misc[3] = 0u;
            for (uint32_t i = (uint32_t)tid; i < 9u * HOMECAP; i += NT) (&s_part[0][0][0])[i] = 0.0f;
            __syncthreads();

            // ---- task ranges (home bead h, plane): cells of the halo row
            // (hi + plane - 1, hj + r - 1) from z-cell hk - 1 to hk + 1, with
            // cells wholly beyond the cutoff skipped. The gap from the home
            // bead to a neighbour cell is the distance to the shared face,
            // edge or corner (fractions in the home cell are exact in FP32); a
            // cell is skipped only when its squared gap exceeds 1 by a margin
            // far above the rounding of the pair test, so no candidate the
            // test would accept is lost. Counted into length bins.
            auto task_rows = [&](uint32_t t, uint32_t &hidx, float4 &hp, uint32_t (&ra)[3], uint32_t (&rn)[3]) {
                const uint32_t h = t / 3u;
                const int plane = (int)(t - 3u * h);
                const int c = s_hcell[h];
                const int hk = c % HZ, hj = (c / HZ) % HY, hi = c / (HZ * HY);
                hidx = s_hidx[h];
                hp = s_pos[hidx];
                const bool moving = !((W->stationary >> (fbits(hp.w) >> 28)) & 1u);
                constexpr float kReach = 1.0f + 1e-5f;
                const float fxr = hp.x - (float)(x0 - 1 + hi), fyr = hp.y - (float)(y0 - 1 + hj),
                            fzr = hp.z - (float)(z0 - 1 + hk);
                const float gx = plane == 0 ? fxr : (plane == 2 ? 1.0f - fxr : 0.0f);
                const float gx2 = gx * gx, glo = fzr * fzr, ghi = (1.0f - fzr) * (1.0f - fzr);
#pragma unroll
                for (int r = 0; r < 3; r++) {
                    const float gy = r == 0 ? fyr : (r == 2 ? 1.0f - fyr : 0.0f);
                    const float g = fmaf(gy, gy, gx2);
                    const bool rowon = moving && g < kReach, loon = rowon && g + glo < kReach,
                               hion = rowon && g + ghi < kReach;
                    const int base = c + (plane - 1) * (HY * HZ) + (r - 1) * HZ;
                    const uint32_t b = s_cstart[base + (loon ? -1 : 0)];
                    const uint32_t e = s_cstart[base + (hion ? 2 : 1)];
                    ra[r] = b;
                    rn[r] = rowon ? e - b : 0u;
                }
            };
            const uint32_t n_tasks_all = 3u * n_home;
            for (uint32_t t = (uint32_t)tid; t < n_tasks_all; t += NT) {
                uint32_t hidx, ra[3], rn[3];
                float4 hp;
                task_rows(t, hidx, hp, ra, rn);
                const uint32_t len = rn[0] + rn[1] + rn[2];
                if (len) atomicAdd(&s_bin[min(len, (uint32_t)kBins - 1u)], 1u);
            }
            __syncthreads();
            if (tid < 64) {       // exclusive scan of the bins, longest first
                const int b = kBins - 1 - tid;
                uint32_t v = tid < kBins ? s_bin[b] : 0u, x = v;
#pragma unroll
                for (int o = 1; o < 64; o <<= 1) {
                    const uint32_t y = __shfl_up(x, o, 64);
                    if (lane >= o) x += y;
                }
                if (tid < kBins) s_bin[b] = x - v;
                if (tid == 63) s_misc[2] = x;
            }
            __syncthreads();
            for (uint32_t t = (uint32_t)tid; t < n_tasks_all; t += NT) {
                uint32_t hidx, ra[3], rn[3];
                float4 hp;
                task_rows(t, hidx, hp, ra, rn);
                const uint32_t len = rn[0] + rn[1] + rn[2];
                if (len) s_task[atomicAdd(&s_bin[min(len, (uint32_t)kBins - 1u)], 1u)] = (uint16_t)t;
            }
            __syncthreads();

            // ---- chunks of 64 tasks, longest first, taken by the waves in turn
            const uint32_t n_tasks = s_misc[2];
            const uint32_t n_chunks = (n_tasks + 63u) / 64u;
            for (;;) {
                uint32_t ch = 0;
                if (lane == 0) ch = atomicAdd(&s_misc[3], 1u);
                ch = __builtin_amdgcn_readfirstlane(__shfl(ch, 0, 64));
                if (ch >= n_chunks) break;
                const uint32_t ti = ch * 64u + (uint32_t)lane;
                const bool act = ti < n_tasks;
                const uint32_t t = act ? s_task[ti] : 0u;
                uint32_t hidx, ra[3], rn[3];
                float4 hp;
                task_rows(t, hidx, hp, ra, rn);
                if (!act) rn[0] = rn[1] = rn[2] = 0u;
                px = py = pz = 0.0f;
                stream(hidx, hp.x, hp.y, hp.z, ra[0], rn[0], ra[1], rn[1], ra[2], rn[2]);
                if (queued) flush();
                if (act) {
                    const uint32_t h = t / 3u, plane = t - 3u * h;
                    s_part[plane][0][h] = px;
                    s_part[plane][1][h] = py;
                    s_part[plane][2][h] = pz;
                }
            }
            __syncthreads();

            // ---- forces of the home beads (planes (p0 + p1) + p2), next-step integration
            for (uint32_t base = 0; base < n_home; base += NT) {
                const uint32_t h = base + (uint32_t)tid;
                const bool valid = h < n_home;
                const uint32_t hh = valid ? h : 0u;
                const uint32_t hidx = s_hidx[hh];
                const int c = s_hcell[hh];
                const uint32_t slot = s_gstart[c] + (hidx - s_cstart[c]);
                const float4 p = s_pos[hidx], v = s_vel[hidx];
                const bool moving = !((W->stationary >> (fbits(p.w) >> 28)) & 1u);
                const float fx = (s_part[0][0][hh] + s_part[1][0][hh]) + s_part[2][0][hh];
                const float fy = (s_part[0][1][hh] + s_part[1][1][hh]) + s_part[2][1][hh];
                const float fz = (s_part[0][2][hh] + s_part[1][2][hh]) + s_part[2][2][hh];
                finish_bead(valid, slot, p, v, moving, fx, fy, fz);
            }
            __syncthreads();
        } else {
            // ---- dense brick: home cell by home cell, up to NT home beads at
            // a time in LDS [0, NT), each stencil row staged in chunks at [NT, HCAP)
            constexpr uint32_t kChunk = (uint32_t)(HCAP - kDense);
            const int nhc = bx * by * bz;
            for (int e = 0; e < nhc; e++) {
                const int ek = e % bz, ej = (e / bz) % by, ei = e / (bz * by);
                const int cx = x0 + ei, cy = y0 + ej, cz = z0 + ek;
                const uint32_t cid = W->cell_off + (uint32_t)((cx * L1 + cy) * L2 + cz);
                const uint32_t g0 = cs[cid], g1 = cs[cid + 1];
                for (uint32_t gb = g0; gb < g1; gb += kDense) {
                    const uint32_t slot = gb + (uint32_t)tid;
                    const bool valid = slot < g1;
                    const float4 hp = valid ? gpos[slot] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    const float4 hv = valid ? gvel[slot] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                    __syncthreads();                    // the previous pass's LDS reads are done
                    if (tid < kDense) { s_pos[tid] = hp; s_vel[tid] = hv; }
                    const bool moving = valid && !((W->stationary >> (fbits(hp.w) >> 28)) & 1u);
                    float sx_[3], sy_[3], sz_[3];
                    for (int plane = 0; plane < 3; plane++) {
                        px = py = pz = 0.0f;
                        const int gxr = cx + plane - 1, gx = wrap(gxr, L0);
                        const float shx = image_shift(gxr, L0, Lx);
                        for (int r = 0; r < 3; r++) {
                            const int gyr = cy + r - 1, gy = wrap(gyr, L1);
                            const float shy = image_shift(gyr, L1, Ly);
                            const uint32_t rowc = W->cell_off + (uint32_t)((gx * L1 + gy) * L2);
                            // the row's three cells z-1, z, z+1 (images shifted)
                            uint32_t cb[3], cn[3];
                            float shz[3];
#pragma unroll
                            for (int d = 0; d < 3; d++) {
                                const int gzr = cz + d - 1;
                                shz[d] = image_shift(gzr, L2, Lz);
                                const uint32_t id = rowc + (uint32_t)wrap(gzr, L2);
                                cb[d] = cs[id];
                                cn[d] = cs[id + 1] - cb[d];
                            }
                            const uint32_t rlen = cn[0] + cn[1] + cn[2];
                            for (uint32_t off = 0; off < rlen; off += kChunk) {
                                const uint32_t m = min(kChunk, rlen - off);
                                __syncthreads();                // LDS candidates of the last chunk are done
                                for (uint32_t q = (uint32_t)tid; q < m; q += NT) {
                                    const uint32_t rq = off + q;
                                    const int d = rq < cn[0] ? 0 : (rq < cn[0] + cn[1] ? 1 : 2);
                                    const uint32_t g = cb[d] + (rq - (d == 0 ? 0u : (d == 1 ? cn[0] : cn[0] + cn[1])));
                                    float4 p = gpos[g];
                                    p.x = p.x + shx;
                                    p.y = p.y + shy;
                                    p.z = p.z + shz[d];
                                    s_pos[kDense + q] = p;
                                    s_vel[kDense + q] = gvel[g];
                                }
                                __syncthreads();
                                const uint32_t n = moving ? m : 0u;
                                const uint32_t a = (uint32_t)kDense;
                                stream((uint32_t)tid, hp.x, hp.y, hp.z, a, r == 0 ? n : 0u, a, r == 1 ? n : 0u, a,
                                       r == 2 ? n : 0u);
                                if (queued) flush();
                            }
                        }
                        sx_[plane] = px;
                        sy_[plane] = py;
                        sz_[plane] = pz;
                    }
                    finish_bead(valid, slot, hp, hv, moving, (sx_[0] + sx_[1]) + sx_[2], (sy_[0] + sy_[1]) + sy_[2],
                                (sz_[0] + sz_[1]) + sz_[2]);
                }
            }
            __syncthreads();
        }
    }
}

// Brick shapes (cells per axis), waves per workgroup and LDS capacities. The
// capacities hold a brick's halo and home beads at density 3 with a wide
// margin; a brick beyond them takes the dense path.
#define PDPD_BRICK_SHAPES(X)                      \
    X(0, 4, 4, 4, 4, 768, 288, 192)               \
    X(1, 2, 4, 4, 4, 544, 160, 192)               \
    X(2, 4, 4, 2, 4, 544, 160, 192)               \
    X(3, 2, 2, 4, 2, 384, 96, 128)                \
    X(4, 4, 4, 4, 8, 768, 288, 192)               \
    X(5, 4, 4, 8, 8, 1280, 512, 192)

// Shapes 0-5 are bricks (this kernel); shapes from kSlabShape0 on are the
// slabs of force_slab.hip, one cell thick in z.
constexpr uint32_t kSlabShape0 = 6;

uint32_t brick_shape_count() { return kSlabShape0 + slab_shape_count(); }

void brick_shape_dims(uint32_t shape, uint32_t *bx, uint32_t *by, uint32_t *bz)
{
    if (shape >= kSlabShape0) {
        slab_shape_dims(shape - kSlabShape0, bx, by);
        *bz = 1;
        return;
    }
    switch (shape) {
#define X(id, a, b, c, nw, hcap, homecap, qcap) case id: *bx = a; *by = b; *bz = c; return;
        PDPD_BRICK_SHAPES(X)
#undef X
    default: *bx = *by = *bz = 4; return;
    }
}

hipError_t launch_force_brick(const StepLaunch &L, hipStream_t st)
{
    if (L.brick_shape >= kSlabShape0) return launch_force_slab(L, L.brick_shape - kSlabShape0, st);
    DeviceState Sp = L.S;
    Sp.ctr = L.S.ctr + L.cpar * L.n_worlds;      // counters read by that step
    const int nxt = L.cur ^ 1;
    const NextBuffers *ns = L.fuse ? L.next + (L.cur * 2 + L.cpar) : nullptr;
    const uint32_t n = L.n_bricks;
    if (!n) return hipSuccess;
    const uint32_t grid = L.force_grid ? std::min<uint32_t>(L.force_grid, n) : n;
    switch (L.brick_shape) {
#define X(id, a, b, c, nw, hcap, homecap, qcap)                                                               \
    case id:                                                                                                  \
        k_force_brick<a, b, c, nw, hcap, homecap, qcap><<<grid, nw * 64, 0, st>>>(Sp, nxt, n, L.bricks, ns);  \
        break;
        PDPD_BRICK_SHAPES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace pdpd
