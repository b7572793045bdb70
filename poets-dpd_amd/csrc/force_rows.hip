// force_rows.hip -- pair forces with one wave per stencil row, nine per workgroup.
//
// The per-slot pair-force kernel (k_force) gives each x-plane of the 27-cell
// stencil to one wave, which scans the plane's three rows in turn: a chain of
// some twenty dependent rounds of candidate loads per wave. When a world of
// 10^4-10^5 beads gives each SIMD only one or two waves, that chain is the
// kernel's duration. Here each (plane, row) of a block of 64 home beads has a
// wave of its own: the nine waves scan their rows and evaluate their accepted
// pairs side by side (k_force's acceptance test, queue and pair arithmetic),
// then one wave per plane adds its rows' pairs to the plane sum, row by row in
// acceptance order. The sums are k_force's (per plane one running sum over its
// rows' candidates in order, planes as ((p0 + p1) + p2), bonded terms in FP64
// rounded once), so the forces are bit-identical.
//
// A row whose accepted pairs do not fit its queue or a home bead's list is
// finished in further rounds: each round a row wave whose previous chunk has
// been added scans on; a plane wave adds chunks in row order and a row's chunk
// only once the rows before it are complete, so the order holds for any
// density. Worlds with per-cell strengths take k_force's spatial form.
#include "device_types.h"
#include "kernels.h"
#include "pdpd_hip.h"
#include "step_device.h"

#include <algorithm>

namespace pdpd {

namespace {

constexpr int kRowWaves = 9;                          // wave w: plane w / 3, row w % 3
constexpr int kRowThreads = kRowWaves * kForceBeads;
constexpr int kRowList = 16;                          // listed pairs per home bead and row, per chunk

// Image shift codes of a candidate, 2 bits per axis: 0 none, 1 = +L, 2 = -L.
__device__ __forceinline__ float image_shift(uint32_t code, float L)
{
    return code == 1u ? L : (code == 2u ? -L : 0.0f);
}

// Slots are below 2^26: the byte offset of a float4 fits 32 bits.
__device__ __forceinline__ const float4 &slot_at(const float4 *__restrict__ p, uint32_t i)
{
    return *reinterpret_cast<const float4 *>(reinterpret_cast<const char *>(p) + (i << 4));
}

// The bonded terms of the bead in slot s, summed in FP64 into s_bond /
// s_nterms (lane = the bead's lane in its block).
__device__ __forceinline__ void bonded_terms_of(const DeviceState &S, const float4 *__restrict__ vel, uint32_t s,
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

template <int kW, int kRQ>
__global__ __launch_bounds__(kRowThreads) void k_force_rows(DeviceState S, int nxt, uint32_t n_blocks,
                                                            const NextBuffers *next)
{
    static_assert(kRQ >= kForceBeads * kW && kRQ <= 65536, "a chunk takes at least one batch");
    using Pos = typename std::conditional<(kRQ > 256), uint16_t, uint8_t>::type;   // queue positions
    static_assert(kRowList >= kW, "a lane lists up to kW pairs per batch");
    // per row wave: queued pairs (candidate slot | image codes in bits 26..31,
    // home lane), their evaluated (unit vector, force magnitude), and per home
    // bead its queued pairs in acceptance order
    __shared__ uint32_t s_qslot[kRowWaves][kRQ];
    __shared__ uint8_t s_qhome[kRowWaves][kRQ];
    __shared__ float4 s_qres[kRowWaves][kRQ];
    __shared__ Pos s_list[kRowList][kRowThreads];
    __shared__ uint8_t s_nlist[kRowWaves][kForceBeads];
    // per row wave: a chunk waits to be added; the row is scanned to its end
    __shared__ uint32_t s_pending[kRowWaves], s_complete[kRowWaves];
    __shared__ uint32_t s_plane_done[kForcePlanes];
    __shared__ float4 s_hp[kForceBeads], s_hv[kForceBeads];
    __shared__ uint32_t s_world;
    __shared__ float s_part[kForcePlanes - 1][3][kForceBeads];
    __shared__ double s_bond[3][kForceBeads];
    __shared__ uint32_t s_nterms[kForceBeads];
    __shared__ float s_a[kMaxTypes * kMaxTypes], s_sg[kMaxTypes * kMaxTypes];

    const int tid = threadIdx.x;
    const int w = tid / kForceBeads;                 // wave-uniform
    const int plane = w / 3, row = w % 3;
    const int lane = tid % kForceBeads;
    const unsigned long long below = (1ull << lane) - 1ull;
    const float4 *__restrict__ pos = S.pos[nxt];
    const float4 *__restrict__ vel = S.vel[nxt];
    float4 *__restrict__ frc = S.frc[nxt];
    const uint32_t *__restrict__ cs = S.cell_start;
    if (next && blockIdx.x == 0) {                       // for the next step's scan
        const NextBuffers B = next_buffers(next);
        for (uint32_t i = (uint32_t)tid; i < B.n_tiles; i += kRowThreads) B.scan_tiles[i] = 0ull;
        if (tid == 0) *B.scan_ticket = 0u;
    }

    for (uint32_t li = blockIdx.x; li < n_blocks; li += gridDim.x) {
        const BeadBlock blk = S.force_blocks[xcd_chunk(li, n_blocks)];
        if (S.ctr[blk.world].remaining == 0) continue;       // uniform: frozen world (never in a chain)
        const WorldDev *__restrict__ W = &S.worlds[blk.world];
        {   // the world's n_types^2 strengths (rows of kMaxTypes)
            const int nt = (int)W->n_types;
            for (int i = tid; i < nt * nt; i += kRowThreads) {
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
        if (w == 0) {
            s_hp[lane] = hp;
            s_hv[lane] = vel[sl];
            if (lane == 0) s_world = blk.world;
        }
        if (lane == 0) {
            s_pending[w] = 0u;
            s_complete[w] = 0u;
        }
        __syncthreads();
        // the bonded terms of the home beads, by the middle row of the middle plane
        if (w == 4) bonded_terms_of(S, vel, s, moving, lane, s_bond, s_nterms);

        const WorldCtr wc = S.ctr[blk.world];
        const uint64_t th = step_hash(wc.t, W->seed_lo, W->seed_hi);
        const uint32_t th_lo = (uint32_t)th, th_hi = (uint32_t)(th >> 32);
        const float scale = W->rng_scale;
        const int L0 = W->L[0], L1 = W->L[1], L2 = W->L[2];
        const float Lx = W->Lf[0], Ly = W->Lf[1], Lz = W->Lf[2];
        const int cx = cell_coord(hp.x, L0), cy = cell_coord(hp.y, L1), cz = cell_coord(hp.z, L2);

        // This wave's row of its plane: a stream of at most two slot ranges,
        // [ra, ra + na) then [rb, rb + nb), as k_force builds it (cells wholly
        // beyond the cutoff skipped with the same margin).
        int ox = cx + plane - 1;
        uint32_t codex = 0;
        if (ox < 0) { ox += L0; codex = 2u; } else if (ox >= L0) { ox -= L0; codex = 1u; }
        const float sx = image_shift(codex, Lx);
        const bool pre = cz == 0, post = cz == L2 - 1;
        const int zi = pre ? L2 - 1 : 0;
        const float szA = pre ? -Lz : 0.0f, szB = post ? Lz : 0.0f;
        const uint32_t zcA = pre ? (2u << 30) : 0u, zcB = post ? (1u << 30) : 0u;
        constexpr float kReach = 1.0f + 1e-5f;
        const float fxr = hp.x - (float)cx, fyr = hp.y - (float)cy, fzr = hp.z - (float)cz;
        const float gx = plane == 0 ? fxr : (plane == 2 ? 1.0f - fxr : 0.0f);
        const float gx2 = gx * gx, glo = fzr * fzr, ghi = (1.0f - fzr) * (1.0f - fzr);
        int oy = cy + row - 1;
        uint32_t codey = 0;
        if (oy < 0) { oy += L1; codey = 2u; } else if (oy >= L1) { oy -= L1; codey = 1u; }
        const uint32_t cxy = (codex << 26) | (codey << 28);
        const float sy = image_shift(codey, Ly);
        const float gy = row == 0 ? fyr : (row == 2 ? 1.0f - fyr : 0.0f);
        const float g = fmaf(gy, gy, gx2);
        const bool rowon = moving && g < kReach, loon = rowon && g + glo < kReach, hion = rowon && g + ghi < kReach;
        const int m0 = pre ? 0 : (loon ? cz - 1 : cz);
        const int m1 = post ? cz : (hion ? cz + 1 : cz);
        const bool img = pre ? loon : (post && hion);
        const uint32_t rbase = W->cell_off + (uint32_t)((ox * L1 + oy) * L2);
        const uint32_t mb = cs[rbase + m0];
        const uint32_t me = rowon ? cs[rbase + m1 + 1] : mb;
        const uint32_t ib = img ? cs[rbase + zi] : 0u;
        const uint32_t ie = img ? cs[rbase + zi + 1] : 0u;
        const uint32_t ra = pre ? ib : mb, na = pre ? ie - ib : me - mb;
        const uint32_t rb = pre ? mb : ib, nb = pre ? me - mb : ie - ib;
        const uint32_t len = na + nb;
        const uint32_t wlen = wave_max(len);
        const uint32_t rbm = rb - na;                         // slot of stream position q >= na: rbm + q

        uint32_t *__restrict__ qslot = s_qslot[w];
        uint8_t *__restrict__ qhome = s_qhome[w];
        float4 *__restrict__ qres = s_qres[w];
        uint32_t k = 0;                                       // next stream position to scan
        float px = 0.0f, py = 0.0f, pz = 0.0f;                // plane sum (the plane's row-0 wave)
        int next_row = 0;                                     // plane wave: the next row to add

        for (;;) {
            // 1. Unless its last chunk still waits to be added, the wave scans on
            // from k until the row ends, the queue is full or a home bead's list
            // is, then evaluates the chunk's pairs
            // (dpd_maths_core_half_step_raw.hpp:129-216, terms grouped as
            // a wr + sgw (scale u - sgw rdotv)).
            if (!s_pending[w] && !s_complete[w]) {
                uint32_t queued = 0, listed = 0;
                for (; k < wlen; k += kW) {
                    uint32_t slot[kW];
                    float3 pc[kW];
#pragma unroll
                    for (int j = 0; j < kW; j++) {
                        const uint32_t q = k + j;
                        // beyond the row's stream: the home slot, never accepted (r2 = 0)
                        const uint32_t inrow = q < na ? ra + q : rbm + q;
                        slot[j] = q < len ? inrow : sl;
                        const float4 &c = slot_at(pos, slot[j]);
                        pc[j] = make_float3(c.x, c.y, c.z);
                    }
                    bool ok[kW];
                    unsigned long long bal[kW];
                    uint32_t total = 0;
#pragma unroll
                    for (int j = 0; j < kW; j++) {
                        const float ozs = k + j < na ? szA : szB;
                        const float dx = hp.x - (pc[j].x + sx), dy = hp.y - (pc[j].y + sy),
                                    dz = hp.z - (pc[j].z + ozs);
                        const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        ok[j] = in_cutoff(r2);
                        bal[j] = __builtin_amdgcn_ballot_w64(ok[j]);
                        total += (uint32_t)__popcll(bal[j]);
                    }
                    if (total == 0) continue;
                    if (queued + total > (uint32_t)kRQ || __builtin_amdgcn_ballot_w64(listed + kW > (uint32_t)kRowList))
                        break;                                 // the rest in a later chunk
                    uint32_t at = queued;
#pragma unroll
                    for (int j = 0; j < kW; j++) {
                        if (ok[j]) {
                            const uint32_t qp = at + (uint32_t)__popcll(bal[j] & below);
                            qslot[qp] = slot[j] | cxy | (k + j < na ? zcA : zcB);
                            qhome[qp] = (uint8_t)lane;
                            s_list[listed++][tid] = (Pos)qp;
                        }
                        at += (uint32_t)__popcll(bal[j]);
                    }
                    queued = at;
                }
                wave_sync();
                for (uint32_t i = (uint32_t)lane; i < queued; i += kForceBeads) {
                    const uint32_t e = qslot[i];
                    const uint32_t hl = qhome[i];
                    const float4 op = slot_at(pos, e & kSlotMask), ov = slot_at(vel, e & kSlotMask);
                    const float4 p = s_hp[hl], v = s_hv[hl];
                    const float ox_ = op.x + image_shift((e >> 26) & 3u, Lx);
                    const float oy_ = op.y + image_shift((e >> 28) & 3u, Ly);
                    const float oz_ = op.z + image_shift((e >> 30) & 3u, Lz);
                    const float dx = p.x - ox_, dy = p.y - oy_, dz = p.z - oz_;
                    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float rr = sqrtf(r2);
                    const float inv_r = 1.0f / rr;
                    const float wr = 1.0f - rr;
                    const float ex = dx * inv_r, ey = dy * inv_r, ez = dz * inv_r;
                    const float dvx = v.x - ov.x, dvy = v.y - ov.y, dvz = v.z - ov.z;
                    const float rdotv = fmaf(ez, dvz, fmaf(ey, dvy, ex * dvx));
                    const uint32_t ph = fbits(p.w), oh = fbits(op.w);
                    const uint32_t ti = (ph >> 28) * kMaxTypes + (oh >> 28);
                    const float a = s_a[ti], sg = s_sg[ti];
                    const float sgw = sg * wr;
                    const float u = rng_uniform(rng_sym(th_lo, th_hi, ph, oh));
                    const float F = fmaf(a, wr, sgw * fmaf(-sgw, rdotv, scale * u));
                    qres[i] = make_float4(ex, ey, ez, F);
                }
                s_nlist[w][lane] = (uint8_t)listed;
                if (lane == 0) {
                    s_pending[w] = 1u;
                    s_complete[w] = k >= wlen ? 1u : 0u;
                }
            }
            __syncthreads();
            // 2. Each plane's row-0 wave adds the waiting chunks of its rows in
            // order, up to the first row not yet scanned to its end.
            if (row == 0) {
                while (next_row < 3) {
                    const int rw = w + next_row;
                    if (!s_pending[rw]) break;
                    const uint32_t n = s_nlist[rw][lane];
                    const Pos *lst = &s_list[0][rw * kForceBeads + lane];
                    for (uint32_t c = 0; c < n; c++) {
                        const float4 q = s_qres[rw][lst[c * kRowThreads]];
                        px = fmaf(q.x, q.w, px);
                        py = fmaf(q.y, q.w, py);
                        pz = fmaf(q.z, q.w, pz);
                    }
                    wave_sync();
                    const bool complete = s_complete[rw] != 0u;
                    if (lane == 0) s_pending[rw] = 0u;
                    if (!complete) break;
                    next_row++;
                }
                if (lane == 0) s_plane_done[plane] = next_row == 3 ? 1u : 0u;
            }
            __syncthreads();
            if (s_plane_done[0] && s_plane_done[1] && s_plane_done[2]) break;   // uniform
        }

        // combine the three planes in order: (p0 + p1) + p2
        if (row == 0 && plane > 0) {
            s_part[plane - 1][0][lane] = px;
            s_part[plane - 1][1][lane] = py;
            s_part[plane - 1][2][lane] = pz;
        }
        __syncthreads();
        if (w == 0 && live) {
            float fx = (px + s_part[0][0][lane]) + s_part[1][0][lane];
            float fy = (py + s_part[0][1][lane]) + s_part[1][1][lane];
            float fz = (pz + s_part[0][2][lane]) + s_part[1][2][lane];
            if (s_nterms[lane]) {      // pair force + bonded terms, rounded once
                fx = (float)((double)fx + s_bond[0][lane]);
                fy = (float)((double)fy + s_bond[1][lane]);
                fz = (float)((double)fz + s_bond[2][lane]);
            }
            frc[s] = make_float4(fx, fy, fz, 0.0f);
            if (const uint32_t ff = force_flags(W, fx, fy, fz)) atomicOr(&S.wflags[blk.world], ff);
        }
        __syncthreads();
        if (next && w == 0) {              // the whole wave, converged; f as stored above by this lane
            const NextBuffers B = next_buffers(next);
            const uint32_t s2 = blk.begin + lane;
            advance_next(B, &S.worlds[s_world], s_world, s2, live, s_hp[lane], s_hv[lane],
                         live ? frc[s2] : make_float4(0.0f, 0.0f, 0.0f, 0.0f));
        }
    }
}

}  // namespace

hipError_t launch_force_rows(const StepLaunch &L, hipStream_t st)
{
    DeviceState Sp = L.S;
    Sp.ctr = L.S.ctr + L.cpar * L.n_worlds;      // counters read by that step
    const int nxt = L.cur ^ 1;
    const NextBuffers *ns = L.fuse ? L.next + (L.cur * 2 + L.cpar) : nullptr;
    const uint32_t fg = L.force_grid ? std::min<uint32_t>(L.force_grid, L.n_force_blocks) : L.n_force_blocks;
    if (!fg) return hipSuccess;
    k_force_rows<8, 512><<<fg, kRowThreads, 0, st>>>(Sp, nxt, L.n_force_blocks, ns);
    return hipGetLastError();
}

}  // namespace pdpd
