// capi.hip -- host side of libpdpd_hip.so: contexts, world import/export, device
// layout, step launches (direct or hipGraph replay), profiling and autotuning.
// Implements the C ABI declared in include/pdpd_hip.h.
#include "pdpd_hip.h"
#include "device_types.h"
#include "kernels.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

using namespace pdpd;

namespace {

thread_local std::string g_err;
thread_local std::string g_support;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                           \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return fail(PDPD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

inline float u2f(uint32_t u)
{
    float f;
    memcpy(&f, &u, 4);
    return f;
}

inline uint32_t f2u(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

constexpr uint32_t kMaxBox = 1024;
constexpr uint32_t kMaxPolymerLength = 127;

// Brick shapes tried by the heuristic and the autotuner, largest first.
const int kBrickCandidates[][3] = {
    {4, 4, 8}, {4, 4, 4}, {2, 4, 4}, {2, 2, 4}, {2, 2, 2}, {1, 2, 2}, {1, 1, 2}, {1, 1, 1},
};

uint32_t bead_hash(uint32_t type, bool monomer, uint32_t polymer_id, uint32_t offset)
{
    // dpd_state.hpp:52-60
    uint32_t base = (offset << 20) | polymer_id;
    base |= uint32_t(monomer ? 1u : 0u) << 27;
    base |= type << 28;
    return base;
}

struct WorldHost {
    int slot = -1;
    bool resident = false;        // lives in the current device layout
    int L[3] = {0, 0, 0};
    double dt = 0, lambda = 0.5;
    uint64_t seed = 0;
    int64_t t = 0;
    uint32_t budget = UINT32_MAX; // steps allowed in the next run
    uint32_t n = 0;
    uint32_t n_types = 0;
    float a[kMaxTypes * kMaxTypes] = {};
    float sg[kMaxTypes * kMaxTypes] = {};
    std::vector<float4> up_pos, up_vel, up_frc;   // pending upload (bead order)
    std::vector<uint32_t> cell_occupancy;          // at import, for brick feasibility
    std::vector<BondRec> bonds;                    // local bead indices
    std::vector<AngleRec> angles;
    struct Poly { uint32_t b0, b1, a0, a1; };
    std::vector<Poly> polys;
    // layout
    uint32_t bead_off = 0, cell_off = 0;
    int brick[3] = {4, 4, 4};
    uint32_t n_cells() const { return (uint32_t)L[0] * L[1] * L[2]; }
};

template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

template <class T>
hipError_t dev_alloc(DevBuf<T> &b, size_t n)
{
    b.release();
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)&b.p, n * sizeof(T));
    if (e == hipSuccess) b.n = n;
    return e;
}

}  // namespace

void pdpd_internal_set_error(const char *msg) { g_err = msg ? msg : ""; }

struct pdpd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::map<int, WorldHost> worlds;
    int next_slot = 0;
    bool dirty = true;
    uint64_t layout_version = 0;

    // device layout
    DevBuf<float4> pos[2], vel[2], frc[2];
    DevBuf<uint32_t> key, rank, perm, cell_count, cell_start, block_sums, slot_of_bead, flags;
    DevBuf<WorldDev> wdev;
    DevBuf<WorldCtr> ctr;
    DevBuf<BrickItem> bricks;
    DevBuf<BeadBlock> blocks;
    DevBuf<PolyItem> polys;
    DevBuf<BondRec> bonds;
    DevBuf<AngleRec> angles;
    int cur = 0;
    uint32_t n_beads = 0, n_cells = 0, n_worlds = 0, n_bricks = 0, n_blocks = 0, n_polys = 0;
    std::vector<int> order;               // slots in layout order
    std::vector<WorldCtr> ctr_host;

    // configuration
    std::string brick_cfg = "auto";
    uint32_t graph_steps = 20;
    uint32_t persistent_blocks = 0;
    int brick_used[3] = {0, 0, 0};

    // graph cache
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_version = ~0ull;
    int graph_cur = -1;
    uint32_t graph_len = 0;

    pdpd_stats stats{};

    ~pdpd_ctx()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (int b = 0; b < 2; b++) { pos[b].release(); vel[b].release(); frc[b].release(); }
        key.release(); rank.release(); perm.release(); cell_count.release(); cell_start.release();
        block_sums.release(); slot_of_bead.release(); flags.release(); wdev.release(); ctr.release();
        bricks.release(); blocks.release(); polys.release(); bonds.release(); angles.release();
        if (stream) (void)hipStreamDestroy(stream);
    }

    DeviceState state() const
    {
        DeviceState S{};
        for (int b = 0; b < 2; b++) { S.pos[b] = pos[b].p; S.vel[b] = vel[b].p; S.frc[b] = frc[b].p; }
        S.key = key.p; S.rank = rank.p; S.perm = perm.p;
        S.cell_count = cell_count.p; S.cell_start = cell_start.p; S.block_sums = block_sums.p;
        S.slot_of_bead = slot_of_bead.p; S.flags = flags.p;
        S.worlds = wdev.p; S.ctr = ctr.p; S.bricks = bricks.p; S.bead_blocks = blocks.p;
        S.polys = polys.p; S.bonds = bonds.p; S.angles = angles.p;
        return S;
    }

    StepLaunch step_launch(int c) const
    {
        StepLaunch L{};
        L.S = state();
        L.cur = c;
        L.n_beads = n_beads; L.n_cells = n_cells; L.n_worlds = n_worlds;
        L.n_bead_blocks = n_blocks; L.n_bricks = n_bricks; L.n_polys = n_polys;
        L.force_blocks = persistent_blocks;
        return L;
    }
};

namespace {

// Largest halo count over the bricks of a world for a brick shape, from the
// occupancy recorded at import (or downloaded for resident worlds).
uint32_t max_halo_count(const WorldHost &w, const int *B)
{
    const int L0 = w.L[0], L1 = w.L[1], L2 = w.L[2];
    if (w.cell_occupancy.empty()) return 0;
    uint32_t best = 0;
    for (int bx = 0; bx < L0; bx += B[0])
    for (int by = 0; by < L1; by += B[1])
    for (int bz = 0; bz < L2; bz += B[2]) {
        int n0 = std::min(B[0], L0 - bx), n1 = std::min(B[1], L1 - by), n2 = std::min(B[2], L2 - bz);
        uint32_t s = 0;
        for (int ix = -1; ix <= n0; ix++)
        for (int iy = -1; iy <= n1; iy++)
        for (int iz = -1; iz <= n2; iz++) {
            int gx = (bx + ix + L0) % L0, gy = (by + iy + L1) % L1, gz = (bz + iz + L2) % L2;
            s += w.cell_occupancy[(size_t)(gx * L1 + gy) * L2 + gz];
        }
        best = std::max(best, s);
    }
    return best;
}

bool brick_fits_lds(const int *B)
{
    return (B[0] + 2) * (B[1] + 2) * (B[2] + 2) <= kMaxHaloCells && B[0] * B[1] * B[2] <= kMaxInnerCells;
}

uint64_t count_bricks(const WorldHost &w, const int *B)
{
    return (uint64_t)((w.L[0] + B[0] - 1) / B[0]) * ((w.L[1] + B[1] - 1) / B[1]) * ((w.L[2] + B[2] - 1) / B[2]);
}

bool parse_brick(const std::string &s, int *B)
{
    int a, b, c;
    char x1, x2;
    if (sscanf(s.c_str(), "%d%c%d%c%d", &a, &x1, &b, &x2, &c) != 5) return false;
    if (x1 != 'x' || x2 != 'x' || a < 1 || b < 1 || c < 1) return false;
    B[0] = a; B[1] = b; B[2] = c;
    return brick_fits_lds(B);
}

// Choose one brick shape for the layout: an explicit configuration, or the
// largest candidate whose halo fits the tile with 25% headroom in every world
// and that still gives at least 1024 work items.
void choose_bricks(pdpd_ctx *ctx)
{
    int B[3];
    if (ctx->brick_cfg != "auto" && parse_brick(ctx->brick_cfg, B)) {
        for (auto &kv : ctx->worlds) std::copy(B, B + 3, kv.second.brick);
        std::copy(B, B + 3, ctx->brick_used);
        return;
    }
    const int *pick = nullptr;
    for (const auto &cand : kBrickCandidates) {
        bool ok = brick_fits_lds(cand);
        uint64_t items = 0;
        for (auto &kv : ctx->worlds) {
            if (!ok) break;
            uint32_t h = max_halo_count(kv.second, cand);
            ok = h * 5 / 4 + 16 <= (uint32_t)kTileBeads;
            items += count_bricks(kv.second, cand);
        }
        if (!ok) continue;
        pick = cand;
        if (items >= 1024) break;
    }
    static const int one[3] = {1, 1, 1};
    if (!pick) pick = one;
    for (auto &kv : ctx->worlds) std::copy(pick, pick + 3, kv.second.brick);
    std::copy(pick, pick + 3, ctx->brick_used);
}

int build_brick_table(pdpd_ctx *ctx)
{
    std::vector<BrickItem> items;
    for (size_t wi = 0; wi < ctx->order.size(); wi++) {
        const WorldHost &w = ctx->worlds.at(ctx->order[wi]);
        for (int bx = 0; bx < w.L[0]; bx += w.brick[0])
        for (int by = 0; by < w.L[1]; by += w.brick[1])
        for (int bz = 0; bz < w.L[2]; bz += w.brick[2]) items.push_back({(int32_t)wi, bx, by, bz});
    }
    // interleave worlds so a persistent grid sees mixed sizes evenly
    ctx->n_bricks = (uint32_t)items.size();
    if (ctx->bricks.n < items.size()) HIP_TRY(dev_alloc(ctx->bricks, items.size()));
    if (!items.empty())
        HIP_TRY(hipMemcpy(ctx->bricks.p, items.data(), items.size() * sizeof(BrickItem), hipMemcpyHostToDevice));
    // world table carries the brick shape
    std::vector<WorldDev> wd(ctx->order.size());
    HIP_TRY(hipMemcpy(wd.data(), ctx->wdev.p, wd.size() * sizeof(WorldDev), hipMemcpyDeviceToHost));
    for (size_t wi = 0; wi < ctx->order.size(); wi++) {
        const WorldHost &w = ctx->worlds.at(ctx->order[wi]);
        for (int d = 0; d < 3; d++) wd[wi].brick[d] = w.brick[d];
    }
    HIP_TRY(hipMemcpy(ctx->wdev.p, wd.data(), wd.size() * sizeof(WorldDev), hipMemcpyHostToDevice));
    ctx->stats.n_bricks = ctx->n_bricks;
    std::copy(ctx->brick_used, ctx->brick_used + 3, ctx->stats.brick);
    ctx->graph_version = ~0ull;
    return PDPD_OK;
}

// Rebuild the device layout: resident worlds are moved to their new offsets,
// pending worlds are uploaded (unsorted; the next step sorts them).
int relayout(pdpd_ctx *ctx)
{
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<int> order;
    for (auto &kv : ctx->worlds) order.push_back(kv.first);

    uint64_t nb = 0, nc = 0;
    std::vector<uint32_t> new_bead_off(order.size()), new_cell_off(order.size());
    for (size_t i = 0; i < order.size(); i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        new_bead_off[i] = (uint32_t)nb;
        new_cell_off[i] = (uint32_t)nc;
        nb += w.n;
        nc += w.n_cells();
    }
    if (nb >= 0x7fffffffull || nc >= 0x7fffffffull) return fail(PDPD_ERR_INVALID, "layout exceeds 2^31 beads or cells");

    // new bead buffers
    DevBuf<float4> npos[2], nvel[2], nfrc[2];
    for (int b = 0; b < 2; b++) {
        HIP_TRY(dev_alloc(npos[b], nb));
        HIP_TRY(dev_alloc(nvel[b], nb));
        HIP_TRY(dev_alloc(nfrc[b], nb));
    }
    for (size_t i = 0; i < order.size(); i++) {
        WorldHost &w = ctx->worlds.at(order[i]);
        const uint32_t off = new_bead_off[i];
        if (w.n == 0) continue;
        if (w.resident) {
            const int c = ctx->cur;
            HIP_TRY(hipMemcpyAsync(npos[0].p + off, ctx->pos[c].p + w.bead_off, w.n * sizeof(float4),
                                   hipMemcpyDeviceToDevice, ctx->stream));
            HIP_TRY(hipMemcpyAsync(nvel[0].p + off, ctx->vel[c].p + w.bead_off, w.n * sizeof(float4),
                                   hipMemcpyDeviceToDevice, ctx->stream));
            HIP_TRY(hipMemcpyAsync(nfrc[0].p + off, ctx->frc[c].p + w.bead_off, w.n * sizeof(float4),
                                   hipMemcpyDeviceToDevice, ctx->stream));
            HIP_TRY(launch_rebase_index(nvel[0].p + off, w.n, off - w.bead_off, ctx->stream));
        } else {
            for (uint32_t k = 0; k < w.n; k++) w.up_vel[k].w = u2f(off + k);
            HIP_TRY(hipMemcpyAsync(npos[0].p + off, w.up_pos.data(), w.n * sizeof(float4), hipMemcpyHostToDevice,
                                   ctx->stream));
            HIP_TRY(hipMemcpyAsync(nvel[0].p + off, w.up_vel.data(), w.n * sizeof(float4), hipMemcpyHostToDevice,
                                   ctx->stream));
            HIP_TRY(hipMemcpyAsync(nfrc[0].p + off, w.up_frc.data(), w.n * sizeof(float4), hipMemcpyHostToDevice,
                                   ctx->stream));
        }
    }
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    for (int b = 0; b < 2; b++) {
        ctx->pos[b].release(); ctx->vel[b].release(); ctx->frc[b].release();
        ctx->pos[b] = npos[b]; ctx->vel[b] = nvel[b]; ctx->frc[b] = nfrc[b];
        npos[b].p = nullptr; nvel[b].p = nullptr; nfrc[b].p = nullptr;
    }
    ctx->cur = 0;
    for (size_t i = 0; i < order.size(); i++) {
        WorldHost &w = ctx->worlds.at(order[i]);
        w.bead_off = new_bead_off[i];
        w.cell_off = new_cell_off[i];
        if (!w.resident) {
            w.up_pos.clear(); w.up_pos.shrink_to_fit();
            w.up_vel.clear(); w.up_vel.shrink_to_fit();
            w.up_frc.clear(); w.up_frc.shrink_to_fit();
            w.resident = true;
        }
    }
    ctx->order = order;
    ctx->n_beads = (uint32_t)nb;
    ctx->n_cells = (uint32_t)nc;
    ctx->n_worlds = (uint32_t)order.size();

    HIP_TRY(dev_alloc(ctx->key, nb));
    HIP_TRY(dev_alloc(ctx->rank, nb));
    HIP_TRY(dev_alloc(ctx->perm, nb));
    HIP_TRY(dev_alloc(ctx->slot_of_bead, nb));
    HIP_TRY(dev_alloc(ctx->cell_count, nc));
    HIP_TRY(dev_alloc(ctx->cell_start, nc + 1));
    HIP_TRY(dev_alloc(ctx->block_sums, (nc + kScanBlock - 1) / kScanBlock + 1));
    if (!ctx->flags.p) {
        HIP_TRY(dev_alloc(ctx->flags, 1));
        HIP_TRY(launch_fill_u32(ctx->flags.p, 1, 0, ctx->stream));
    }
    HIP_TRY(launch_fill_u32(ctx->cell_count.p, (uint32_t)nc, 0, ctx->stream));

    // world table
    std::vector<WorldDev> wd(order.size());
    for (size_t i = 0; i < order.size(); i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        WorldDev &d = wd[i];
        memset(&d, 0, sizeof(d));
        for (int k = 0; k < 3; k++) { d.L[k] = w.L[k]; d.Lf[k] = (float)w.L[k]; d.brick[k] = w.brick[k]; }
        d.dt = (float)w.dt;
        d.hdt = 0.5f * d.dt;
        d.rng_scale = (float)std::sqrt(24.0 / (double)d.dt);
        d.bead_off = w.bead_off; d.n_beads = w.n;
        d.cell_off = w.cell_off; d.n_cells = w.n_cells();
        d.seed_lo = (uint32_t)w.seed; d.seed_hi = (uint32_t)(w.seed >> 32);
        d.n_types = w.n_types;
        memcpy(d.a, w.a, sizeof(d.a));
        memcpy(d.sg, w.sg, sizeof(d.sg));
    }
    HIP_TRY(dev_alloc(ctx->wdev, wd.size()));
    if (!wd.empty()) HIP_TRY(hipMemcpy(ctx->wdev.p, wd.data(), wd.size() * sizeof(WorldDev), hipMemcpyHostToDevice));
    HIP_TRY(dev_alloc(ctx->ctr, 2 * std::max<size_t>(1, order.size())));
    ctx->ctr_host.assign(2 * std::max<size_t>(1, order.size()), WorldCtr{0, 0});

    // bead blocks
    std::vector<BeadBlock> blocks;
    for (size_t i = 0; i < order.size(); i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        for (uint32_t b = 0; b < w.n; b += kBeadThreads)
            blocks.push_back({(int32_t)i, w.bead_off + b, w.bead_off + std::min(w.n, b + kBeadThreads), 0});
    }
    ctx->n_blocks = (uint32_t)blocks.size();
    HIP_TRY(dev_alloc(ctx->blocks, blocks.size()));
    if (!blocks.empty())
        HIP_TRY(hipMemcpy(ctx->blocks.p, blocks.data(), blocks.size() * sizeof(BeadBlock), hipMemcpyHostToDevice));

    // bonded topology with global bead indices
    std::vector<PolyItem> pitems;
    std::vector<BondRec> brec;
    std::vector<AngleRec> arec;
    for (size_t i = 0; i < order.size(); i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        for (const auto &p : w.polys) {
            PolyItem it{};
            it.world = (int32_t)i;
            it.bond_begin = (uint32_t)brec.size();
            for (uint32_t k = p.b0; k < p.b1; k++) {
                BondRec r = w.bonds[k];
                r.head += w.bead_off; r.tail += w.bead_off;
                brec.push_back(r);
            }
            it.bond_end = (uint32_t)brec.size();
            it.angle_begin = (uint32_t)arec.size();
            for (uint32_t k = p.a0; k < p.a1; k++) {
                AngleRec r = w.angles[k];
                r.head += w.bead_off; r.mid += w.bead_off; r.tail += w.bead_off;
                arec.push_back(r);
            }
            it.angle_end = (uint32_t)arec.size();
            pitems.push_back(it);
        }
    }
    ctx->n_polys = (uint32_t)pitems.size();
    HIP_TRY(dev_alloc(ctx->polys, pitems.size()));
    HIP_TRY(dev_alloc(ctx->bonds, brec.size()));
    HIP_TRY(dev_alloc(ctx->angles, arec.size()));
    if (!pitems.empty())
        HIP_TRY(hipMemcpy(ctx->polys.p, pitems.data(), pitems.size() * sizeof(PolyItem), hipMemcpyHostToDevice));
    if (!brec.empty())
        HIP_TRY(hipMemcpy(ctx->bonds.p, brec.data(), brec.size() * sizeof(BondRec), hipMemcpyHostToDevice));
    if (!arec.empty())
        HIP_TRY(hipMemcpy(ctx->angles.p, arec.data(), arec.size() * sizeof(AngleRec), hipMemcpyHostToDevice));

    choose_bricks(ctx);
    int rc = build_brick_table(ctx);
    if (rc) return rc;

    ctx->stats.n_worlds = ctx->n_worlds;
    ctx->stats.n_beads = ctx->n_beads;
    ctx->layout_version++;
    ctx->graph_version = ~0ull;
    ctx->dirty = false;
    return PDPD_OK;
}

int ensure_layout(pdpd_ctx *ctx)
{
    if (!ctx->dirty) return PDPD_OK;
    return relayout(ctx);
}

// Upload this run's step counters into the copy the next step reads.
int load_counters(pdpd_ctx *ctx, uint32_t nsteps, uint64_t *bead_steps)
{
    const uint32_t W = ctx->n_worlds;
    uint64_t bs = 0;
    for (uint32_t i = 0; i < W; i++) {
        const WorldHost &w = ctx->worlds.at(ctx->order[i]);
        WorldCtr c;
        c.t = (uint32_t)w.t;
        c.remaining = std::min(w.budget, nsteps);
        ctx->ctr_host[ctx->cur * W + i] = c;
        bs += (uint64_t)w.n * c.remaining;
    }
    if (W) HIP_TRY(hipMemcpyAsync(ctx->ctr.p + ctx->cur * W, ctx->ctr_host.data() + ctx->cur * W,
                                  W * sizeof(WorldCtr), hipMemcpyHostToDevice, ctx->stream));
    if (bead_steps) *bead_steps = bs;
    return PDPD_OK;
}

void finish_counters(pdpd_ctx *ctx, uint32_t nsteps)
{
    for (auto &kv : ctx->worlds) {
        WorldHost &w = kv.second;
        w.t += std::min(w.budget, nsteps);
        w.budget = UINT32_MAX;
    }
}

int launch_direct(pdpd_ctx *ctx, uint32_t nsteps, KernelEvents *ev)
{
    for (uint32_t s = 0; s < nsteps; s++) {
        HIP_TRY(launch_step(ctx->step_launch(ctx->cur), ctx->stream, ev));
        ctx->cur ^= 1;
    }
    return PDPD_OK;
}

int ensure_graph(pdpd_ctx *ctx, uint32_t len)
{
    if (ctx->graph_exec && ctx->graph_version == ctx->layout_version && ctx->graph_cur == ctx->cur &&
        ctx->graph_len == len)
        return PDPD_OK;
    if (ctx->graph_exec) {
        HIP_TRY(hipGraphExecDestroy(ctx->graph_exec));
        ctx->graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    int c = ctx->cur;
    for (uint32_t s = 0; s < len; s++) {
        hipError_t e = launch_step(ctx->step_launch(c), ctx->stream, nullptr);
        if (e != hipSuccess) {
            hipGraph_t junk;
            (void)hipStreamEndCapture(ctx->stream, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            return fail(PDPD_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        }
        c ^= 1;
    }
    HIP_TRY(hipStreamEndCapture(ctx->stream, &g));
    hipError_t e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(PDPD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    ctx->graph_version = ctx->layout_version;
    ctx->graph_cur = ctx->cur;
    ctx->graph_len = len;
    return PDPD_OK;
}

int check_flags(pdpd_ctx *ctx)
{
    uint32_t f = 0;
    HIP_TRY(hipMemcpy(&f, ctx->flags.p, sizeof(f), hipMemcpyDeviceToHost));
    if (!f) return PDPD_OK;
    ctx->stats.flags |= f;
    HIP_TRY(launch_fill_u32(ctx->flags.p, 1, 0, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    if (f & PDPD_FLAG_TILE_OVERFLOW)
        return fail(PDPD_ERR_CAPACITY,
                    "a brick's halo held more beads than the LDS tile (" + std::to_string(kTileBeads) +
                        "); set a smaller \"brick\"");
    return fail(PDPD_ERR_STATE, "a bead left the box (non-finite position or a jump of more than one box length)");
}

WorldHost *find_world(pdpd_ctx *ctx, int slot)
{
    auto it = ctx->worlds.find(slot);
    return it == ctx->worlds.end() ? nullptr : &it->second;
}

std::string support_reason(const pdpd_world_desc *w)
{
    if (!w) return "No world.";
    for (int d = 0; d < 3; d++) {
        if (std::round(w->box[d]) != w->box[d]) return "World bounds must be integer.";
        if (w->box[d] < 3) return "All dimensions must be at least 3.";
        if (w->box[d] > kMaxBox) return "Dimension exceeds " + std::to_string(kMaxBox) + ".";
    }
    if (!(w->dt > 0)) return "dt must be positive.";
    if (w->n_bead_types == 0 || w->n_bead_types > (uint32_t)kMaxTypes)
        return "Engine supports 1 to " + std::to_string(kMaxTypes) + " bead types.";
    if (w->stationary) {
        for (uint32_t i = 0; i < w->n_bead_types; i++)
            if (w->stationary[i]) return "Engine does not support stationary bead types.";
    }
    for (uint32_t p = 0; p < w->n_polymers; p++) {
        uint32_t len = w->polymer_bead_start[p + 1] - w->polymer_bead_start[p];
        if (len > kMaxPolymerLength) return "Polymer longer than 127 beads.";
    }
    return "";
}

}  // namespace

// ============================================================== C ABI ====

extern "C" {

const char *pdpd_last_error(void) { return g_err.c_str(); }

int pdpd_abi_version(void) { return PDPD_ABI_VERSION; }

int pdpd_get_limits(pdpd_limits *out)
{
    if (!out) return fail(PDPD_ERR_INVALID, "null output");
    out->max_bead_types = kMaxTypes;
    out->min_box = 3;
    out->max_box = kMaxBox;
    out->max_polymer_length = kMaxPolymerLength;
    out->tile_capacity = kTileBeads;
    out->max_bonds_per_bead = 0xffffffffu;
    return PDPD_OK;
}

const char *pdpd_can_support(const pdpd_world_desc *w)
{
    g_support = support_reason(w);
    return g_support.c_str();
}

int pdpd_ctx_create(int device, pdpd_ctx **out)
{
    if (!out) return fail(PDPD_ERR_INVALID, "null output");
    int n = 0;
    HIP_TRY(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(PDPD_ERR_INVALID, "no HIP device " + std::to_string(device));
    HIP_TRY(hipSetDevice(device));
    std::unique_ptr<pdpd_ctx> ctx(new pdpd_ctx());
    ctx->device = device;
    HIP_TRY(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    *out = ctx.release();
    return PDPD_OK;
}

int pdpd_ctx_destroy(pdpd_ctx *ctx)
{
    if (!ctx) return PDPD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return PDPD_OK;
}

int pdpd_world_add(pdpd_ctx *ctx, const pdpd_world_desc *d, int *slot_out)
{
    if (!ctx || !d) return fail(PDPD_ERR_INVALID, "null argument");
    std::string why = support_reason(d);
    if (!why.empty()) return fail(PDPD_ERR_UNSUPPORTED, why);
    double t0 = now_s();
    WorldHost w;
    for (int k = 0; k < 3; k++) w.L[k] = (int)d->box[k];
    w.dt = d->dt;
    w.lambda = d->lambda;
    w.seed = d->seed;
    w.t = d->t;
    w.n = d->n_beads;
    w.n_types = d->n_bead_types;
    const uint32_t nt = d->n_bead_types;
    for (uint32_t i = 0; i < nt; i++)
        for (uint32_t j = 0; j < nt; j++) {
            w.a[i * kMaxTypes + j] = (float)d->conservative[i * nt + j];
            w.sg[i * kMaxTypes + j] = (float)std::sqrt(d->dissipative[i * nt + j]);
        }
    // import: FP64 AoS -> FP32, w = v - f dt/2 (dpd_engine_avx2_half_merge_tbb.hpp:515-524)
    const float dt = (float)d->dt, hdt = 0.5f * dt;
    w.up_pos.resize(w.n);
    w.up_vel.resize(w.n);
    w.up_frc.resize(w.n);
    w.cell_occupancy.assign(w.n_cells(), 0);
    for (uint32_t i = 0; i < w.n; i++) {
        float x[3], v[3], f[3];
        int c[3];
        for (int k = 0; k < 3; k++) {
            const float Lf = (float)w.L[k];
            x[k] = (float)d->x[3 * i + k];
            if (!(x[k] >= 0.0f)) return fail(PDPD_ERR_INVALID, "bead " + std::to_string(i) + " outside the box");
            if (x[k] >= Lf) x[k] = std::nextafter(Lf, 0.0f);   // FP32 rounding of x < L up to L
            v[k] = (float)d->v[3 * i + k];
            f[k] = (float)d->f[3 * i + k];
            v[k] = std::fmaf(f[k], -hdt, v[k]);
            c[k] = std::min((int)std::floor(x[k]), w.L[k] - 1);
        }
        if (d->x[3 * i] >= w.L[0] || d->x[3 * i + 1] >= w.L[1] || d->x[3 * i + 2] >= w.L[2])
            return fail(PDPD_ERR_INVALID, "bead " + std::to_string(i) + " outside the box");
        const uint32_t h = bead_hash(d->bead_type[i], d->is_monomer[i] != 0, d->polymer_id[i], d->polymer_offset[i]);
        w.up_pos[i] = make_float4(x[0], x[1], x[2], u2f(h));
        w.up_vel[i] = make_float4(v[0], v[1], v[2], 0.0f);
        w.up_frc[i] = make_float4(f[0], f[1], f[2], 0.0f);
        w.cell_occupancy[(size_t)(c[0] * w.L[1] + c[1]) * w.L[2] + c[2]]++;
    }
    // topology in the reference's order: per polymer, its type's bonds then bond pairs
    for (uint32_t p = 0; p < d->n_polymers; p++) {
        const uint32_t pt = d->polymer_type[p];
        if (pt >= d->n_polymer_types) return fail(PDPD_ERR_INVALID, "polymer type out of range");
        const uint32_t *ids = d->polymer_bead_ids + d->polymer_bead_start[p];
        const uint32_t len = d->polymer_bead_start[p + 1] - d->polymer_bead_start[p];
        WorldHost::Poly poly{(uint32_t)w.bonds.size(), 0, 0, 0};
        for (uint32_t b = d->pt_bond_start[pt]; b < d->pt_bond_start[pt + 1]; b++) {
            if (d->bond_head[b] >= len || d->bond_tail[b] >= len) return fail(PDPD_ERR_INVALID, "bond offset out of range");
            w.bonds.push_back({ids[d->bond_head[b]], ids[d->bond_tail[b]], (float)d->bond_kappa[b], (float)d->bond_r0[b]});
        }
        poly.b1 = (uint32_t)w.bonds.size();
        poly.a0 = (uint32_t)w.angles.size();
        for (uint32_t a = d->pt_angle_start[pt]; a < d->pt_angle_start[pt + 1]; a++) {
            const uint32_t hb = d->pt_bond_start[pt] + d->angle_head_bond[a];
            const uint32_t tb = d->pt_bond_start[pt] + d->angle_tail_bond[a];
            if (hb >= d->pt_bond_start[pt + 1] || tb >= d->pt_bond_start[pt + 1])
                return fail(PDPD_ERR_INVALID, "bond pair offset out of range");
            AngleRec r{};
            r.head = ids[d->bond_head[hb]];
            r.mid = ids[d->bond_tail[hb]];
            r.tail = ids[d->bond_tail[tb]];
            r.kappa = (float)d->angle_kappa[a];
            r.cos0 = (float)std::cos(d->angle_theta0[a]);
            r.sin0 = (float)std::sin(d->angle_theta0[a]);
            w.angles.push_back(r);
        }
        poly.a1 = (uint32_t)w.angles.size();
        if (poly.b1 > poly.b0 || poly.a1 > poly.a0) w.polys.push_back(poly);
    }
    for (const auto &b : w.bonds)
        if (b.head >= w.n || b.tail >= w.n) return fail(PDPD_ERR_INVALID, "bond bead id out of range");
    for (const auto &a : w.angles)
        if (a.head >= w.n || a.mid >= w.n || a.tail >= w.n) return fail(PDPD_ERR_INVALID, "angle bead id out of range");

    int slot = ctx->next_slot++;
    w.slot = slot;
    ctx->worlds.emplace(slot, std::move(w));
    ctx->dirty = true;
    ctx->stats.upload_s += now_s() - t0;
    if (slot_out) *slot_out = slot;
    return PDPD_OK;
}

int pdpd_world_remove(pdpd_ctx *ctx, int slot)
{
    if (!ctx || !find_world(ctx, slot)) return fail(PDPD_ERR_INVALID, "unknown slot");
    ctx->worlds.erase(slot);
    ctx->dirty = true;
    return PDPD_OK;
}

int pdpd_world_set_steps(pdpd_ctx *ctx, int slot, uint32_t steps)
{
    WorldHost *w = ctx ? find_world(ctx, slot) : nullptr;
    if (!w) return fail(PDPD_ERR_INVALID, "unknown slot");
    w->budget = steps;
    return PDPD_OK;
}

int pdpd_run(pdpd_ctx *ctx, uint32_t nsteps)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    double t0 = now_s();
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    if (nsteps == 0 || ctx->n_worlds == 0) {
        finish_counters(ctx, 0);
        return PDPD_OK;
    }
    uint64_t bs = 0;
    rc = load_counters(ctx, nsteps, &bs);
    if (rc) return rc;
    uint32_t left = nsteps;
    const uint32_t G = ctx->graph_steps & ~1u;     // even: buffers return to the same parity
    if (G >= 2 && left >= G) {
        rc = ensure_graph(ctx, G);
        if (rc) return rc;
        while (left >= G) {
            HIP_TRY(hipGraphLaunch(ctx->graph_exec, ctx->stream));
            left -= G;
        }
    }
    rc = launch_direct(ctx, left, nullptr);
    if (rc) return rc;
    finish_counters(ctx, nsteps);
    ctx->stats.steps += nsteps;
    ctx->stats.bead_steps += bs;
    ctx->stats.run_s += now_s() - t0;
    return PDPD_OK;
}

int pdpd_sync(pdpd_ctx *ctx)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    double t0 = now_s();
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    ctx->stats.run_s += now_s() - t0;
    if (ctx->flags.p) return check_flags(ctx);
    return PDPD_OK;
}

int pdpd_world_download(pdpd_ctx *ctx, int slot, double *x, double *v, double *f, int64_t *t)
{
    WorldHost *w = ctx ? find_world(ctx, slot) : nullptr;
    if (!w) return fail(PDPD_ERR_INVALID, "unknown slot");
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    rc = pdpd_sync(ctx);
    if (rc) return rc;
    std::vector<float4> P(w->n), V(w->n), F(w->n);
    const int c = ctx->cur;
    if (w->n) {
        HIP_TRY(hipMemcpy(P.data(), ctx->pos[c].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(V.data(), ctx->vel[c].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(F.data(), ctx->frc[c].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
    }
    const double hdt = (double)(float)w->dt * 0.5;
    for (uint32_t s = 0; s < w->n; s++) {
        const uint32_t i = f2u(V[s].w) - w->bead_off;
        if (i >= w->n) return fail(PDPD_ERR_STATE, "corrupt bead index in device state");
        const double px[3] = {P[s].x, P[s].y, P[s].z};
        const double fx[3] = {F[s].x, F[s].y, F[s].z};
        const double wx[3] = {V[s].x, V[s].y, V[s].z};
        for (int k = 0; k < 3; k++) {
            if (x) x[3 * i + k] = px[k];
            if (f) f[3 * i + k] = fx[k];
            if (v) v[3 * i + k] = wx[k] + fx[k] * hdt;   // post-correct, avx2 engine :554
        }
    }
    if (t) *t = w->t;
    return PDPD_OK;
}

int pdpd_world_download_f32(pdpd_ctx *ctx, int slot, float *x, float *wout, float *f, uint32_t *t)
{
    WorldHost *w = ctx ? find_world(ctx, slot) : nullptr;
    if (!w) return fail(PDPD_ERR_INVALID, "unknown slot");
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    rc = pdpd_sync(ctx);
    if (rc) return rc;
    std::vector<float4> P(w->n), V(w->n), F(w->n);
    const int c = ctx->cur;
    if (w->n) {
        HIP_TRY(hipMemcpy(P.data(), ctx->pos[c].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(V.data(), ctx->vel[c].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(F.data(), ctx->frc[c].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
    }
    for (uint32_t s = 0; s < w->n; s++) {
        const uint32_t i = f2u(V[s].w) - w->bead_off;
        if (i >= w->n) return fail(PDPD_ERR_STATE, "corrupt bead index in device state");
        if (x) { x[3 * i] = P[s].x; x[3 * i + 1] = P[s].y; x[3 * i + 2] = P[s].z; }
        if (wout) { wout[3 * i] = V[s].x; wout[3 * i + 1] = V[s].y; wout[3 * i + 2] = V[s].z; }
        if (f) { f[3 * i] = F[s].x; f[3 * i + 1] = F[s].y; f[3 * i + 2] = F[s].z; }
    }
    if (t) *t = (uint32_t)w->t;
    return PDPD_OK;
}

int pdpd_world_download_cells(pdpd_ctx *ctx, int slot, int32_t *order, int32_t *cell_start)
{
    WorldHost *w = ctx ? find_world(ctx, slot) : nullptr;
    if (!w) return fail(PDPD_ERR_INVALID, "unknown slot");
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    rc = pdpd_sync(ctx);
    if (rc) return rc;
    if (order && w->n) {
        std::vector<float4> V(w->n);
        HIP_TRY(hipMemcpy(V.data(), ctx->vel[ctx->cur].p + w->bead_off, w->n * sizeof(float4), hipMemcpyDeviceToHost));
        for (uint32_t s = 0; s < w->n; s++) order[s] = (int32_t)(f2u(V[s].w) - w->bead_off);
    }
    if (cell_start) {
        const uint32_t nc = w->n_cells();
        std::vector<uint32_t> cs(nc + 1);
        HIP_TRY(hipMemcpy(cs.data(), ctx->cell_start.p + w->cell_off, (nc + 1) * sizeof(uint32_t),
                          hipMemcpyDeviceToHost));
        for (uint32_t c = 0; c <= nc; c++) cell_start[c] = (int32_t)(cs[c] - w->bead_off);
    }
    return PDPD_OK;
}

int pdpd_set_config(pdpd_ctx *ctx, const char *key, const char *value)
{
    if (!ctx || !key || !value) return fail(PDPD_ERR_INVALID, "null argument");
    std::string k(key), v(value);
    if (k == "brick") {
        int B[3];
        if (v != "auto" && !parse_brick(v, B))
            return fail(PDPD_ERR_INVALID, "brick must be \"auto\" or BXxBYxBZ with (BX+2)(BY+2)(BZ+2) <= " +
                                              std::to_string(kMaxHaloCells));
        ctx->brick_cfg = v;
        if (!ctx->dirty && ctx->n_worlds) {
            choose_bricks(ctx);
            return build_brick_table(ctx);
        }
        return PDPD_OK;
    }
    if (k == "graph_steps") {
        char *end = nullptr;
        long n = strtol(v.c_str(), &end, 10);
        if (!end || *end || n < 0 || n > 100000) return fail(PDPD_ERR_INVALID, "graph_steps must be an integer >= 0");
        ctx->graph_steps = (uint32_t)n;
        return PDPD_OK;
    }
    if (k == "persistent_blocks") {
        char *end = nullptr;
        long n = strtol(v.c_str(), &end, 10);
        if (!end || *end || n < 0 || n > (1 << 24)) return fail(PDPD_ERR_INVALID, "persistent_blocks must be >= 0");
        ctx->persistent_blocks = (uint32_t)n;
        ctx->graph_version = ~0ull;
        return PDPD_OK;
    }
    return fail(PDPD_ERR_INVALID, "Unknown engine config name '" + k + "'");
}

int pdpd_get_config(pdpd_ctx *ctx, const char *key, char *buf, size_t buflen)
{
    if (!ctx || !key || !buf || !buflen) return fail(PDPD_ERR_INVALID, "null argument");
    std::string k(key), v;
    if (k == "brick") v = ctx->brick_cfg;
    else if (k == "graph_steps") v = std::to_string(ctx->graph_steps);
    else if (k == "persistent_blocks") v = std::to_string(ctx->persistent_blocks);
    else return fail(PDPD_ERR_INVALID, "Unknown engine config name '" + k + "'");
    snprintf(buf, buflen, "%s", v.c_str());
    return PDPD_OK;
}

int pdpd_list_config(pdpd_ctx *ctx, char *buf, size_t buflen)
{
    if (!ctx || !buf || !buflen) return fail(PDPD_ERR_INVALID, "null argument");
    std::string s;
    s += "brick=" + ctx->brick_cfg + "\tcells per pair-force workgroup, \"auto\" or BXxBYxBZ\n";
    s += "graph_steps=" + std::to_string(ctx->graph_steps) + "\tsteps per captured hipGraph (even; 0 = direct launches)\n";
    s += "persistent_blocks=" + std::to_string(ctx->persistent_blocks) +
         "\tpair-force grid size for a persistent launch (0 = one workgroup per brick)\n";
    snprintf(buf, buflen, "%s", s.c_str());
    return PDPD_OK;
}

int pdpd_get_stats(pdpd_ctx *ctx, pdpd_stats *out)
{
    if (!ctx || !out) return fail(PDPD_ERR_INVALID, "null argument");
    *out = ctx->stats;
    return PDPD_OK;
}

int pdpd_run_profiled(pdpd_ctx *ctx, uint32_t nsteps)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    if (nsteps == 0 || ctx->n_worlds == 0) return PDPD_OK;
    uint64_t bs = 0;
    rc = load_counters(ctx, nsteps, &bs);
    if (rc) return rc;
    KernelEvents ev;
    for (int k = 0; k < 7; k++) HIP_TRY(hipEventCreate(&ev.ev[k]));
    double acc[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t s = 0; s < nsteps; s++) {
        HIP_TRY(launch_step(ctx->step_launch(ctx->cur), ctx->stream, &ev));
        ctx->cur ^= 1;
        HIP_TRY(hipEventSynchronize(ev.ev[6]));
        for (int k = 0; k < 6; k++) {
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev.ev[k], ev.ev[k + 1]));
            acc[k] += ms;
        }
    }
    for (int k = 0; k < 7; k++) (void)hipEventDestroy(ev.ev[k]);
    finish_counters(ctx, nsteps);
    ctx->stats.steps += nsteps;
    ctx->stats.bead_steps += bs;
    ctx->stats.k_integrate_ms = acc[0] / nsteps;
    ctx->stats.k_scan_ms = acc[1] / nsteps;
    ctx->stats.k_scatter_ms = acc[2] / nsteps;
    ctx->stats.k_reorder_ms = acc[3] / nsteps;
    ctx->stats.k_force_ms = acc[4] / nsteps;
    ctx->stats.k_bonds_ms = acc[5] / nsteps;
    ctx->stats.profiled_steps = nsteps;
    return check_flags(ctx);
}

int pdpd_autotune(pdpd_ctx *ctx, uint32_t steps_per_trial, char *choice, size_t choice_len)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    if (ctx->n_worlds == 0) return fail(PDPD_ERR_INVALID, "no worlds to tune on");
    if (steps_per_trial < 2) steps_per_trial = 2;
    steps_per_trial &= ~1u;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    // save the state: trials run on a copy and the worlds do not advance
    const size_t nb = ctx->n_beads;
    DevBuf<float4> sp, sv, sf;
    HIP_TRY(dev_alloc(sp, nb));
    HIP_TRY(dev_alloc(sv, nb));
    HIP_TRY(dev_alloc(sf, nb));
    const int c0 = ctx->cur;
    HIP_TRY(hipMemcpy(sp.p, ctx->pos[c0].p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(sv.p, ctx->vel[c0].p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(sf.p, ctx->frc[c0].p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
    std::vector<int64_t> saved_t;
    for (int s : ctx->order) saved_t.push_back(ctx->worlds.at(s).t);

    const std::string saved_cfg = ctx->brick_cfg;
    const uint32_t saved_persist = ctx->persistent_blocks;
    std::string best;
    uint32_t best_persist = 0;
    double best_ms = 1e30;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    const uint32_t persist_opts[2] = {0, 256u * 8u};
    for (const auto &cand : kBrickCandidates) {
        if (!brick_fits_lds(cand)) continue;
        bool ok = true;
        for (auto &kv : ctx->worlds) {
            uint32_t h = max_halo_count(kv.second, cand);
            if (h * 5 / 4 + 16 > (uint32_t)kTileBeads) ok = false;
        }
        if (!ok) continue;
        char name[32];
        snprintf(name, sizeof(name), "%dx%dx%d", cand[0], cand[1], cand[2]);
        for (uint32_t po : persist_opts) {
            ctx->brick_cfg = name;
            ctx->persistent_blocks = po;
            choose_bricks(ctx);
            rc = build_brick_table(ctx);
            if (rc) return rc;
            for (int s = 0; s < 2; s++) {   // restore state for each trial
                HIP_TRY(hipMemcpy(ctx->pos[c0].p, sp.p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
                HIP_TRY(hipMemcpy(ctx->vel[c0].p, sv.p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
                HIP_TRY(hipMemcpy(ctx->frc[c0].p, sf.p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
                ctx->cur = c0;
                rc = load_counters(ctx, steps_per_trial, nullptr);
                if (rc) return rc;
                if (s == 0) {     // warm-up
                    rc = launch_direct(ctx, 2, nullptr);
                    if (rc) return rc;
                    HIP_TRY(hipStreamSynchronize(ctx->stream));
                    continue;
                }
                HIP_TRY(hipEventRecord(e0, ctx->stream));
                rc = launch_direct(ctx, steps_per_trial, nullptr);
                if (rc) return rc;
                HIP_TRY(hipEventRecord(e1, ctx->stream));
                HIP_TRY(hipEventSynchronize(e1));
            }
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            uint32_t f = 0;
            HIP_TRY(hipMemcpy(&f, ctx->flags.p, sizeof(f), hipMemcpyDeviceToHost));
            if (f) {      // candidate overflowed: discard it
                HIP_TRY(launch_fill_u32(ctx->flags.p, 1, 0, ctx->stream));
                HIP_TRY(hipStreamSynchronize(ctx->stream));
                continue;
            }
            if (ms < best_ms) { best_ms = ms; best = name; best_persist = po; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // restore
    HIP_TRY(hipMemcpy(ctx->pos[c0].p, sp.p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(ctx->vel[c0].p, sv.p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
    HIP_TRY(hipMemcpy(ctx->frc[c0].p, sf.p, nb * sizeof(float4), hipMemcpyDeviceToDevice));
    ctx->cur = c0;
    for (size_t i = 0; i < ctx->order.size(); i++) ctx->worlds.at(ctx->order[i]).t = saved_t[i];
    sp.release(); sv.release(); sf.release();
    if (best.empty()) {
        ctx->brick_cfg = saved_cfg;
        ctx->persistent_blocks = saved_persist;
        choose_bricks(ctx);
        rc = build_brick_table(ctx);
        if (rc) return rc;
        return fail(PDPD_ERR_CAPACITY, "no brick shape fits the worlds");
    }
    ctx->brick_cfg = best;
    ctx->persistent_blocks = best_persist;
    choose_bricks(ctx);
    rc = build_brick_table(ctx);
    if (rc) return rc;
    if (choice && choice_len)
        snprintf(choice, choice_len, "brick=%s persistent_blocks=%u ms_per_step=%.5f", best.c_str(), best_persist,
                 best_ms / steps_per_trial);
    return PDPD_OK;
}

int pdpd_debug_rng(int device, uint32_t n, const uint64_t *t_hash, const uint32_t *a, const uint32_t *b,
                   uint32_t *u32_out, float *f32_out)
{
    if (n == 0) return PDPD_OK;
    if (!t_hash || !a || !b) return fail(PDPD_ERR_INVALID, "null input");
    HIP_TRY(hipSetDevice(device));
    DevBuf<uint64_t> dt;
    DevBuf<uint32_t> da, db, du;
    DevBuf<float> df;
    HIP_TRY(dev_alloc(dt, n));
    HIP_TRY(dev_alloc(da, n));
    HIP_TRY(dev_alloc(db, n));
    HIP_TRY(dev_alloc(du, n));
    HIP_TRY(dev_alloc(df, n));
    HIP_TRY(hipMemcpy(dt.p, t_hash, n * 8, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(da.p, a, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(db.p, b, n * 4, hipMemcpyHostToDevice));
    HIP_TRY(launch_debug_rng(n, dt.p, da.p, db.p, du.p, df.p, nullptr));
    HIP_TRY(hipDeviceSynchronize());
    if (u32_out) HIP_TRY(hipMemcpy(u32_out, du.p, n * 4, hipMemcpyDeviceToHost));
    if (f32_out) HIP_TRY(hipMemcpy(f32_out, df.p, n * 4, hipMemcpyDeviceToHost));
    dt.release(); da.release(); db.release(); du.release(); df.release();
    return PDPD_OK;
}

}  // extern "C"
