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
#include <cfloat>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
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
// Chained steps (the pair-force kernel integrating the next step) measured
// faster at 98k beads (46.1 against 47.5 us per step) and slower at 6.3M
// (1.77 against 1.74 ms): on by default up to this many resident beads.
constexpr uint32_t kFuseAutoBeads = 1u << 20;
// force_variant=auto: variant 4 (smaller pair queue, fewer registers, 6 waves
// per SIMD) from kVariantLargeBeads resident beads (measured: 128^3 world 1.19
// against 1.34 ms per pair-force launch, 256-world ensemble over 12 contexts of
// 0.5M beads +8 %; 32^3 and 16^3 worlds 1 us slower), variant 7 (a wave per
// stencil row) below kVariantSmallBeads (16^3 world 14.3 against 19.3 us; at
// 20^3 23.5 against 19.9, at 32^3 2.5 times variant 0's time), variant 0 between
constexpr uint32_t kVariantAuto = UINT32_MAX;
constexpr uint32_t kVariantLargeBeads = 1u << 18;
constexpr uint32_t kVariantSmallBeads = 1u << 14;
constexpr uint64_t kBandMinCells = 64 * 64;   // y-z plane size from which force blocks are listed in bands
constexpr uint32_t kBandRows = 16;            // default band height (y rows)
constexpr uint32_t kMaxPolymerLength = 127;

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
    uint32_t stationary = 0;      // bit per stationary bead type
    float a[kMaxTypes * kMaxTypes] = {};
    float sg[kMaxTypes * kMaxTypes] = {};
    std::vector<float2> cell_par;                  // per cell and type pair (a, gamma), or empty: constant strengths
    std::vector<float4> up_pos, up_vel, up_frc;   // pending upload (bead order)
    std::vector<TermRec> bonds, angles;            // local bead indices, reference order
    std::vector<uint32_t> stationary_beads;        // local indices of beads of stationary types
    // per bead (local index) its terms, in the order the per-polymer loop
    // applies them (its bonds, then its angles): role << kRoleShift | term
    // index within bonds then angles; built once, when the world is added
    std::vector<uint32_t> term_start_local, term_list_local;
    uint32_t bead_off = 0, cell_off = 0;          // layout
    uint32_t term_base = 0;                        // index of the world's first bonded term (bonds, then angles)
    uint32_t list_base = 0;                        // index of the world's first per-bead term-list entry
    int64_t t_at_layout = 0;                       // t when the current layout was built
    uint32_t flags = 0;                            // PDPD_FLAG_* raised by this world so far
    // work lists in local offsets, kept between layouts (see local_force_blocks, build_bricks)
    std::vector<BeadBlock> fb_local;
    uint32_t fb_rows = UINT32_MAX;
    std::vector<BrickRec> bricks_local;
    uint32_t bricks_shape = UINT32_MAX;
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

template <class T>
hipError_t dev_upload(DevBuf<T> &b, const std::vector<T> &v)
{
    hipError_t e = dev_alloc(b, v.size());
    if (e != hipSuccess || v.empty()) return e;
    return hipMemcpy(b.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice);
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
    uint64_t steps_on_layout = 0;          // steps run since the layout (and its cell lists) was built

    // device layout
    DevBuf<float4> pos[2], vel[2], frc[2];
    DevBuf<uint32_t> key, perm, sorted_idx, cell_count, cell_start, slot_of_bead, flags, scan_ticket;
    DevBuf<uint32_t> wflags;              // per world in layout order: PDPD_FLAG_* raised on the device
    // the previous layout's copies of the tables a relayout rebuilds on the
    // device from the current ones (then swapped in)
    DevBuf<uint32_t> slot_of_bead_alt, wflags_alt, term_start_alt, bead_terms_alt;
    DevBuf<TermRec> terms_alt;
    DevBuf<uint32_t> remap;               // relayout: new world index -> old
    DevBuf<MoveSeg> move_segs;            // relayout: one segment per world (k_move_worlds)
    // buffers replaced while the stream may still read them, freed once it is idle
    std::vector<void *> graveyard;
    std::vector<void *> host_graveyard;   // pinned arenas replaced while their copies may run
    // pinned staging of relayouts' host-to-device uploads: a ring of arenas,
    // each reused once the copies of the relayout that last filled it are done
    // (its event), so a relayout does not wait for the steps enqueued before it
    static constexpr int kStageArenas = 4;
    struct StageArena {
        char *p = nullptr;
        size_t cap = 0;
        hipEvent_t done = nullptr;
        bool pending = false;
    } arenas[kStageArenas];
    int stage_cur = 0;                    // arena of the relayout being built
    char *stage = nullptr;                // = arenas[stage_cur].p while staging
    size_t stage_cap = 0, stage_used = 0;
    uint32_t reserved_beads = 0;          // reserve_beads setting
    DevBuf<unsigned long long> scan_tiles;
    DevBuf<uint32_t> tile_sum;
    DevBuf<WorldDev> wdev;
    DevBuf<WorldCtr> ctr;
    DevBuf<BeadBlock> bblocks, fblocks;
    DevBuf<uint32_t> term_start, bead_terms;
    DevBuf<TermRec> terms;
    DevBuf<double> term_frc;
    DevBuf<StationaryBead> stationary;
    DevBuf<float2> cell_par;
    bool spatial = false;                 // some resident world has per-cell strengths
    DevBuf<NextBuffers> next;             // [buffer * 2 + counter copy]: what a chained step's force kernel writes
    DevBuf<BrickRec> bricks;              // pair-force bricks of the resident worlds
    uint32_t n_bricks = 0;
    // observables: first bead block of every world (n_worlds + 1), per-block
    // partial sums and per-world results (kObsValues doubles each)
    DevBuf<uint32_t> obs_first;
    DevBuf<double> obs_partial, obs_out;

    // interval snapshot: pos, vel, frc of every resident world (and the flags
    // word) copied on the device behind the enqueued steps; a world's share
    // goes to pinned host memory on a second stream when it is downloaded
    struct SnapWorld {
        uint32_t index;                  // position in the layout (its flag word)
        uint32_t bead_off, n;
        int64_t t;
        double dt;
    };
    hipStream_t copy_stream = nullptr;
    hipEvent_t snap_event = nullptr;
    DevBuf<float4> snap;             // [pos | vel | frc], snap_beads each
    DevBuf<uint32_t> snap_flags;
    // FP64 observables of the snapshot's state (pdpd_snapshot_observables):
    // per-block partials and per-world sums on the device, their event, and
    // pinned host room for one world's sums and flags
    DevBuf<double> snap_obs_partial, snap_obs;
    hipEvent_t snap_obs_event = nullptr;
    bool snap_has_obs = false;
    double *snap_obs_host = nullptr;
    uint32_t snap_beads = 0;
    std::map<int, SnapWorld> snap_worlds;
    float4 *snap_host = nullptr;
    size_t snap_host_n = 0;
    int cur = 0;                          // buffer holding the state after the enqueued steps
    int cpar = 0;                         // counter copy the next step reads
    uint32_t all_active = 0;              // the run's first steps in which every world steps (chains stay inside)
    uint32_t n_beads = 0, n_cells = 0, n_worlds = 0, n_bblocks = 0, n_fblocks = 0, n_stationary = 0, n_terms = 0;
    std::vector<int> order;               // slots in layout order
    std::vector<WorldCtr> ctr_host;

    // configuration
    uint32_t graph_steps = 20;
    uint32_t force_grid = 0;
    uint32_t force_variant = kVariantAuto;   // pair-force kernel variant, or kVariantAuto (by size)
    uint32_t lookback_tiles = 256;
    uint32_t fuse_mode = 2;               // chained steps: 0 off, 1 on, 2 on up to kFuseAutoBeads resident beads
    uint32_t band_rows = kBandRows;       // force-block order of large worlds (0 = layout order)
    uint32_t force_kernel = 0;            // 0: per-slot blocks, 1: cell bricks staged in LDS (autotuned)
    uint32_t brick_shape = 0;             // brick shape index (brick_shape_dims)

    bool fuse() const { return fuse_mode == 1 || (fuse_mode == 2 && n_beads <= kFuseAutoBeads); }
    uint32_t variant() const
    {
        if (force_variant != kVariantAuto) return force_variant;
        return n_beads >= kVariantLargeBeads ? 4u : (n_beads < kVariantSmallBeads ? 7u : 0u);
    }

    // graph cache
    hipGraphExec_t graph_exec = nullptr;
    uint64_t graph_version = ~0ull;
    int graph_cur = -1, graph_cpar = -1;
    uint32_t graph_len = 0;

    pdpd_stats stats{};

    ~pdpd_ctx()
    {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (int b = 0; b < 2; b++) { pos[b].release(); vel[b].release(); frc[b].release(); }
        key.release(); perm.release(); sorted_idx.release(); cell_count.release();
        cell_start.release(); slot_of_bead.release(); flags.release(); scan_ticket.release(); scan_tiles.release();
        wflags.release();
        slot_of_bead_alt.release(); wflags_alt.release(); term_start_alt.release(); bead_terms_alt.release();
        terms_alt.release(); remap.release(); move_segs.release();
        for (void *q : graveyard) (void)hipFree(q);
        for (void *q : host_graveyard) (void)hipHostFree(q);
        for (StageArena &a : arenas) {
            if (a.p) (void)hipHostFree(a.p);
            if (a.done) (void)hipEventDestroy(a.done);
        }
        tile_sum.release();
        wdev.release(); ctr.release(); bblocks.release(); fblocks.release(); term_start.release();
        bead_terms.release(); terms.release(); term_frc.release(); stationary.release();
        cell_par.release();
        next.release();
        bricks.release();
        obs_first.release(); obs_partial.release(); obs_out.release();
        snap.release(); snap_flags.release();
        snap_obs_partial.release(); snap_obs.release();
        if (snap_obs_event) (void)hipEventDestroy(snap_obs_event);
        if (snap_obs_host) (void)hipHostFree(snap_obs_host);
        if (snap_host) (void)hipHostFree(snap_host);
        if (snap_event) (void)hipEventDestroy(snap_event);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
        if (stream) (void)hipStreamDestroy(stream);
    }

    DeviceState state() const
    {
        DeviceState S{};
        for (int b = 0; b < 2; b++) { S.pos[b] = pos[b].p; S.vel[b] = vel[b].p; S.frc[b] = frc[b].p; }
        S.key = key.p; S.perm = perm.p; S.sorted_idx = sorted_idx.p;
        S.cell_count = cell_count.p; S.cell_start = cell_start.p;
        S.scan_tiles = scan_tiles.p; S.scan_ticket = scan_ticket.p;
        S.slot_of_bead = slot_of_bead.p; S.flags = flags.p; S.wflags = wflags.p;
        S.worlds = wdev.p; S.ctr = ctr.p; S.bead_blocks = bblocks.p; S.force_blocks = fblocks.p;
        S.term_start = term_start.p; S.bead_terms = bead_terms.p; S.terms = terms.p; S.term_frc = term_frc.p;
        S.stationary_beads = stationary.p;
        S.cell_par = cell_par.p;
        return S;
    }

    StepLaunch step_launch(int c, int cp, bool integrate, bool fuse) const
    {
        StepLaunch L{};
        L.S = state();
        L.cur = c;
        L.cpar = cp;
        L.integrate = integrate ? 1 : 0;
        L.fuse = fuse ? 1 : 0;
        L.next = next.p;
        L.n_beads = n_beads; L.n_cells = n_cells; L.n_worlds = n_worlds;
        L.n_bead_blocks = n_bblocks; L.n_force_blocks = n_fblocks;
        L.n_stationary = n_stationary;
        L.n_terms = n_terms;
        L.force_grid = force_grid;
        L.force_variant = variant();
        L.tile_sum = tile_sum.p;
        L.lookback_tiles = lookback_tiles;
        L.spatial = spatial ? 1u : 0u;
        L.force_kernel = force_kernel;
        L.brick_shape = brick_shape;
        L.n_bricks = n_bricks;
        L.bricks = bricks.p;
        return L;
    }
};

namespace {

// Work lists over all resident worlds, in layout (world, then cell) order. The
// pair-force kernel deals contiguous shares of this list to the XCDs.
// A world whose y-z plane holds at least kBandMinCells cells has its force
// blocks listed in bands of band_rows y-rows instead: band by band, each band
// plane by plane (x), within a plane in slot order. An XCD then works through
// a few planes of one band at a time, so the stencil neighbours its
// workgroups read stay in its L2, instead of streaming whole planes. A
// block's cell is estimated from its middle slot at the world's mean density
// (slots are in cell order). The order changes which XCD evaluates a block and
// when, never a block's beads or any sum, so forces are bit-identical.
// The pair-force work items of one world (kForceBeads slots each) in local
// slot offsets, in band order when band_rows is set; kept with the world and
// rebuilt only when the band setting changes.
const std::vector<BeadBlock> &local_force_blocks(WorldHost &w, uint32_t band_rows)
{
    if (w.fb_rows == band_rows) return w.fb_local;
    std::vector<BeadBlock> &fb = w.fb_local;
    fb.clear();
    for (uint32_t b = 0; b < w.n; b += kForceBeads) fb.push_back({0, b, std::min(w.n, b + kForceBeads), 0});
    const uint64_t plane = (uint64_t)w.L[1] * (uint64_t)w.L[2];
    if (!(band_rows == 0 || plane < kBandMinCells || (uint32_t)w.L[1] <= band_rows || w.n == 0)) {
        const double cells_per_slot = (double)w.n_cells() / (double)w.n;
        auto band_plane = [&](const BeadBlock &b) {
            const uint64_t mid = ((uint64_t)b.begin + (uint64_t)b.end) / 2;
            const uint64_t c = std::min<uint64_t>((uint64_t)((double)mid * cells_per_slot), w.n_cells() - 1);
            const uint64_t y = (c / (uint64_t)w.L[2]) % (uint64_t)w.L[1];
            return std::make_pair(y / band_rows, c / plane);
        };
        std::stable_sort(fb.begin(), fb.end(),
                         [&](const BeadBlock &a, const BeadBlock &b) { return band_plane(a) < band_plane(b); });
    }
    w.fb_rows = band_rows;
    return fb;
}

// Work lists of the worlds in `order` placed at bead offsets off[i]: bead
// blocks (kBeadThreads slots: integrate, observables) and pair-force blocks.
void build_blocks(pdpd_ctx *ctx, const std::vector<int> &order, const std::vector<uint32_t> &off,
                  std::vector<BeadBlock> &bb, std::vector<BeadBlock> &fb)
{
    bb.clear();
    fb.clear();
    for (size_t i = 0; i < order.size(); i++) {
        WorldHost &w = ctx->worlds.at(order[i]);
        for (uint32_t b = 0; b < w.n; b += kBeadThreads)
            bb.push_back({(int32_t)i, off[i] + b, off[i] + std::min(w.n, b + kBeadThreads), 0});
        for (const BeadBlock &l : local_force_blocks(w, ctx->band_rows))
            fb.push_back({(int32_t)i, off[i] + l.begin, off[i] + l.end, 0});
    }
}

// Pair-force bricks of every resident world: boxes of the shape's cells,
// clipped at the world's far faces, listed world by world in x, y, z order
// (the pair-force kernel deals contiguous shares of the list to the XCDs, so
// an XCD works through neighbouring bricks whose halos overlap in its L2).
// Bricks name cells by world and coordinates, so a world's list is kept with
// it and only its world index changes between layouts.
void build_bricks(pdpd_ctx *ctx, const std::vector<int> &order, std::vector<BrickRec> &out)
{
    uint32_t bx, by, bz;
    brick_shape_dims(ctx->brick_shape, &bx, &by, &bz);
    out.clear();
    for (size_t i = 0; i < order.size(); i++) {
        WorldHost &w = ctx->worlds.at(order[i]);
        if (w.n == 0) continue;
        if (w.bricks_shape != ctx->brick_shape) {
            w.bricks_local.clear();
            for (int x = 0; x < w.L[0]; x += (int)bx)
                for (int y = 0; y < w.L[1]; y += (int)by)
                    for (int z = 0; z < w.L[2]; z += (int)bz) {
                        BrickRec r{};
                        r.origin = (uint32_t)x | ((uint32_t)y << 10) | ((uint32_t)z << 20);
                        const uint32_t dx = std::min<uint32_t>(bx, (uint32_t)(w.L[0] - x));
                        const uint32_t dy = std::min<uint32_t>(by, (uint32_t)(w.L[1] - y));
                        const uint32_t dz = std::min<uint32_t>(bz, (uint32_t)(w.L[2] - z));
                        r.dims = dx | (dy << 8) | (dz << 16);
                        w.bricks_local.push_back(r);
                    }
            w.bricks_shape = ctx->brick_shape;
        }
        for (BrickRec r : w.bricks_local) {
            r.world = (int32_t)i;
            out.push_back(r);
        }
    }
}

int upload_bricks(pdpd_ctx *ctx)
{
    std::vector<BrickRec> br;
    build_bricks(ctx, ctx->order, br);
    ctx->n_bricks = (uint32_t)br.size();
    HIP_TRY(dev_upload(ctx->bricks, br));
    return PDPD_OK;
}

// Move the device's per-world flags into the worlds (the layout's order names
// them; worlds removed since are skipped) and clear them. Returns the union.
int harvest_world_flags(pdpd_ctx *ctx, uint32_t *any)
{
    *any = 0;
    if (!ctx->wflags.p || ctx->order.empty()) return PDPD_OK;
    std::vector<uint32_t> f(ctx->order.size());
    HIP_TRY(hipMemcpyAsync(f.data(), ctx->wflags.p, f.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    uint32_t u = 0;
    for (size_t i = 0; i < f.size(); i++) {
        if (!f[i]) continue;
        u |= f[i];
        auto it = ctx->worlds.find(ctx->order[i]);
        if (it != ctx->worlds.end()) it->second.flags |= f[i];
    }
    if (u) HIP_TRY(launch_fill_u32(ctx->wflags.p, (uint32_t)f.size(), 0, ctx->stream));
    *any = u;
    return PDPD_OK;
}

std::string flag_text(uint32_t f)
{
    std::string s;
    auto add = [&](uint32_t bit, const char *what) {
        if (f & bit) s += (s.empty() ? "" : "; ") + std::string(what);
    };
    add(PDPD_FLAG_BAD_POSITION, "a bead left the box (non-finite position or a jump of more than one box length)");
    add(PDPD_FLAG_CELL_JUMP, "a bead moved more than one cell in a time-step");
    add(PDPD_FLAG_BOND_STRETCH, "a bond is longer than half the smallest box side (it has snapped)");
    add(PDPD_FLAG_FORCE_BOUND, "a force is not finite");
    add(PDPD_FLAG_VELOCITY_BOUND, "a velocity is not finite");
    return s;
}

// Device buffer with room for n entries; contents are not kept. A replaced
// buffer may still be read by enqueued work: it is freed when the stream is
// next idle (free_graveyard), so growing never waits for the device.
template <class T>
int reserve(pdpd_ctx *ctx, DevBuf<T> &b, size_t n)
{
    if (n == 0) n = 1;
    if (b.p && b.n >= n) return PDPD_OK;
    const size_t cap = std::max(n, b.n + b.n / 2);
    T *p = nullptr;
    const hipError_t e = hipMalloc((void **)&p, cap * sizeof(T));
    if (e != hipSuccess) return fail(PDPD_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    if (b.p) ctx->graveyard.push_back(b.p);
    b.p = p;
    b.n = cap;
    return PDPD_OK;
}

// After the stream is idle: free the buffers replaced while it ran.
void free_graveyard(pdpd_ctx *ctx)
{
    for (void *q : ctx->graveyard) (void)hipFree(q);
    ctx->graveyard.clear();
    for (void *q : ctx->host_graveyard) (void)hipHostFree(q);
    ctx->host_graveyard.clear();
}

// Pinned staging for one relayout's uploads (total bytes known in advance):
// waits only for the previous relayout's copies, which precede everything
// enqueued since.
int stage_begin(pdpd_ctx *ctx, size_t total)
{
    total += 4096;
    // the first arena, from the one after the last used, whose copies are done
    // and which is large enough; else the first done one (grown); else wait
    // for the next in turn
    int pick = -1, idle = -1;
    for (int k = 1; k <= pdpd_ctx::kStageArenas && pick < 0; k++) {
        const int i = (ctx->stage_cur + k) % pdpd_ctx::kStageArenas;
        pdpd_ctx::StageArena &a = ctx->arenas[i];
        if (!a.done) HIP_TRY(hipEventCreateWithFlags(&a.done, hipEventDisableTiming));
        if (a.pending) {
            const hipError_t q = hipEventQuery(a.done);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return fail(PDPD_ERR_HIP, std::string("hipEventQuery: ") + hipGetErrorString(q));
            a.pending = false;
        }
        if (a.cap >= total) pick = i;
        else if (idle < 0) idle = i;
    }
    if (pick < 0) pick = idle;
    if (pick < 0) {
        pick = (ctx->stage_cur + 1) % pdpd_ctx::kStageArenas;
        HIP_TRY(hipEventSynchronize(ctx->arenas[pick].done));
        ctx->arenas[pick].pending = false;
    }
    pdpd_ctx::StageArena &a = ctx->arenas[pick];
    if (a.cap < total) {
        const size_t cap = std::max(total, a.cap + a.cap / 2);
        if (a.p) ctx->host_graveyard.push_back(a.p);   // freeing pinned memory may wait for the device
        a.p = nullptr;
        a.cap = 0;
        HIP_TRY(hipHostMalloc((void **)&a.p, cap, hipHostMallocDefault));
        a.cap = cap;
    }
    ctx->stage_cur = pick;
    ctx->stage = a.p;
    ctx->stage_cap = a.cap;
    ctx->stage_used = 0;
    return PDPD_OK;
}

int stage_put(pdpd_ctx *ctx, void *dst, const void *src, size_t bytes)
{
    if (!bytes) return PDPD_OK;
    if (ctx->stage_used + bytes > ctx->stage_cap) return fail(PDPD_ERR_STATE, "relayout staging overflow");
    char *h = ctx->stage + ctx->stage_used;
    memcpy(h, src, bytes);
    HIP_TRY(hipMemcpyAsync(dst, h, bytes, hipMemcpyHostToDevice, ctx->stream));
    ctx->stage_used += (bytes + 255) & ~(size_t)255;
    return PDPD_OK;
}

int stage_end(pdpd_ctx *ctx)
{
    pdpd_ctx::StageArena &a = ctx->arenas[ctx->stage_cur];
    HIP_TRY(hipEventRecord(a.done, ctx->stream));
    a.pending = true;
    ctx->stage = nullptr;
    ctx->stage_cap = 0;
    return PDPD_OK;
}

// Capacity for layouts of up to n beads (n cells, n / 2 bonded terms, n
// term-list entries, 256 worlds) and the pinned staging their tables and one
// large upload need, so that building them allocates nothing. Only before any
// layout exists: the buffers' contents are not kept.
int reserve_capacity(pdpd_ctx *ctx, size_t n)
{
    if (n == 0) return PDPD_OK;
    int rc;
    for (int b = 0; b < 2; b++)
        if ((rc = reserve(ctx, ctx->pos[b], n)) || (rc = reserve(ctx, ctx->vel[b], n)) ||
            (rc = reserve(ctx, ctx->frc[b], n)))
            return rc;
    const size_t W = 256, nc = n, nt = n / 2 + 1;
    if ((rc = reserve(ctx, ctx->slot_of_bead, n)) || (rc = reserve(ctx, ctx->slot_of_bead_alt, n)) ||
        (rc = reserve(ctx, ctx->term_start, n + 1)) || (rc = reserve(ctx, ctx->term_start_alt, n + 1)) ||
        (rc = reserve(ctx, ctx->bead_terms, n)) || (rc = reserve(ctx, ctx->bead_terms_alt, n)) ||
        (rc = reserve(ctx, ctx->terms, nt)) || (rc = reserve(ctx, ctx->terms_alt, nt)) ||
        (rc = reserve(ctx, ctx->term_frc, nt * kTermFrcStride)) || (rc = reserve(ctx, ctx->key, n)) ||
        (rc = reserve(ctx, ctx->perm, n)) || (rc = reserve(ctx, ctx->sorted_idx, n)) ||
        (rc = reserve(ctx, ctx->cell_count, nc)) || (rc = reserve(ctx, ctx->cell_start, nc + 1)) ||
        (rc = reserve(ctx, ctx->scan_tiles, scan_tiles((uint32_t)nc))) ||
        (rc = reserve(ctx, ctx->tile_sum, scan_tiles((uint32_t)nc))) || (rc = reserve(ctx, ctx->wflags, W)) ||
        (rc = reserve(ctx, ctx->wflags_alt, W)) || (rc = reserve(ctx, ctx->remap, W)) ||
        (rc = reserve(ctx, ctx->move_segs, W)) || (rc = reserve(ctx, ctx->wdev, W)) ||
        (rc = reserve(ctx, ctx->ctr, 2 * W)) || (rc = reserve(ctx, ctx->bblocks, n / kBeadThreads + W)) ||
        (rc = reserve(ctx, ctx->fblocks, n / kForceBeads + W)) || (rc = reserve(ctx, ctx->bricks, nc / 8 + W)) ||
        (rc = reserve(ctx, ctx->obs_first, W + 1)))
        return rc;
    // staging, per arena: the tables above for n beads plus an upload of n / 8 beads
    const size_t stage = n / kBeadThreads * sizeof(BeadBlock) + (n / kForceBeads) * sizeof(BeadBlock) +
                         (nc / 8) * sizeof(BrickRec) + W * (sizeof(WorldDev) + sizeof(MoveSeg) + 16) +
                         (n / 8) * (3 * sizeof(float4) + 32) + (1u << 20);
    for (int k = 0; k < pdpd_ctx::kStageArenas; k++) {
        pdpd_ctx::StageArena &a = ctx->arenas[k];
        if (a.cap >= stage) continue;
        if (a.p) HIP_TRY(hipHostFree(a.p));      // no layout yet: no copy is pending
        a.p = nullptr;
        a.cap = 0;
        HIP_TRY(hipHostMalloc((void **)&a.p, stage, hipHostMallocDefault));
        a.cap = stage;
    }
    return PDPD_OK;
}

template <class T>
size_t staged_bytes(const std::vector<T> &v)
{
    return (v.size() * sizeof(T) + 255) & ~(size_t)255;
}

// Rebuild the device layout: resident worlds are moved to their new offsets,
// pending worlds are uploaded (unsorted; the next step sorts them). All of it
// is enqueued on the context's stream behind the work already there -- no
// wait for the device: buffers keep their capacity (a replaced one is freed
// once the stream is idle), resident worlds and their bonded-term tables,
// bead-slot map and flags are moved on the device, and host tables go up from
// pinned staging. The host work is per world and per work item, plus the
// upload of new worlds.
int relayout(pdpd_ctx *ctx)
{
    const double t_start = now_s();
    HIP_TRY(hipSetDevice(ctx->device));
    if ((!ctx->graveyard.empty() || !ctx->host_graveyard.empty()) && hipStreamQuery(ctx->stream) == hipSuccess)
        free_graveyard(ctx);
    std::vector<int> order;
    for (auto &kv : ctx->worlds) order.push_back(kv.first);
    const size_t W = order.size();
    if (W >= (1u << (32 - kTermWorldShift))) return fail(PDPD_ERR_INVALID, "too many worlds in one context");

    uint64_t nb = 0, nc = 0, ntl = 0, nterm = 0;
    std::vector<uint32_t> new_bead_off(W), new_cell_off(W), new_list_base(W), new_term_base(W);
    for (size_t i = 0; i < W; i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        new_bead_off[i] = (uint32_t)nb;
        new_cell_off[i] = (uint32_t)nc;
        new_list_base[i] = (uint32_t)ntl;
        new_term_base[i] = (uint32_t)nterm;
        nb += w.n;
        nc += w.n_cells();
        ntl += w.term_list_local.size();
        nterm += w.bonds.size() + w.angles.size();
    }
    if (nb >= kMaxLayoutBeads) return fail(PDPD_ERR_INVALID, "resident worlds exceed 2^26 beads in one context");
    if (nc >= 0x7fffffffull) return fail(PDPD_ERR_INVALID, "resident worlds exceed 2^31 cells in one context");
    if (nterm >= (1u << kRoleShift)) return fail(PDPD_ERR_INVALID, "too many bond terms");
    std::map<int, uint32_t> old_index;
    for (uint32_t i = 0; i < ctx->order.size(); i++) old_index[ctx->order[i]] = i;

    // ---- host tables of the new layout (small: per world and per work item)
    std::vector<WorldDev> wd(W);
    std::vector<float2> cpar;
    for (size_t i = 0; i < W; i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        WorldDev &d = wd[i];
        memset(&d, 0, sizeof(d));
        for (int k = 0; k < 3; k++) { d.L[k] = w.L[k]; d.Lf[k] = (float)w.L[k]; }
        d.dt = (float)w.dt;
        d.hdt = 0.5f * d.dt;
        d.rng_scale = (float)std::sqrt(24.0 / (double)d.dt);
        d.bead_off = new_bead_off[i]; d.n_beads = w.n;
        d.cell_off = new_cell_off[i]; d.n_cells = w.n_cells();
        d.seed_lo = (uint32_t)w.seed; d.seed_hi = (uint32_t)(w.seed >> 32);
        d.n_types = w.n_types;
        d.stationary = w.stationary;
        // failure bounds: velocities and forces must stay finite (the bounds of
        // the reference's validator, dpd_state_validator.hpp:147-152, apply to
        // the states its tools read, not to its engines' runs); bonds must be
        // shorter than the nearest image resolves
        d.max_v = FLT_MAX;
        d.max_f = FLT_MAX;
        d.max_bond = 0.5 * (double)std::min(w.L[0], std::min(w.L[1], w.L[2]));
        memcpy(d.a, w.a, sizeof(d.a));
        memcpy(d.sg, w.sg, sizeof(d.sg));
        if (!w.cell_par.empty()) {
            if (cpar.size() + w.cell_par.size() > UINT32_MAX) return fail(PDPD_ERR_INVALID, "per-cell strength tables too large");
            d.spatial = 1;
            d.cell_par_off = (uint32_t)cpar.size();
            cpar.insert(cpar.end(), w.cell_par.begin(), w.cell_par.end());
        }
    }
    std::vector<BeadBlock> bb, fb;
    build_blocks(ctx, order, new_bead_off, bb, fb);
    std::vector<BrickRec> bricks;
    build_bricks(ctx, order, bricks);
    std::vector<uint32_t> first(W + 1, (uint32_t)bb.size());   // first bead block of each world
    for (size_t b = bb.size(); b-- > 0;) first[(size_t)bb[b].world] = (uint32_t)b;
    for (size_t i = W; i-- > 0;)
        if (first[i] > first[i + 1]) first[i] = first[i + 1];
    std::vector<StationaryBead> sb;
    for (size_t i = 0; i < W; i++)
        for (uint32_t b : ctx->worlds.at(order[i]).stationary_beads) sb.push_back({new_bead_off[i] + b, (int32_t)i});
    std::vector<uint32_t> remap(std::max<size_t>(1, W), ~0u);
    for (size_t i = 0; i < W; i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        auto it = old_index.find(order[i]);
        if (w.resident && it != old_index.end()) remap[i] = it->second;
    }
    const std::vector<uint32_t> tstart_end{(uint32_t)ntl};

    // ---- device buffers of the new layout (capacity kept across relayouts)
    const int src = ctx->cur, dst = ctx->cur ^ 1;
    int rc;
    if ((rc = reserve(ctx, ctx->pos[dst], nb)) || (rc = reserve(ctx, ctx->vel[dst], nb)) ||
        (rc = reserve(ctx, ctx->frc[dst], nb)) || (rc = reserve(ctx, ctx->slot_of_bead_alt, nb)) ||
        (rc = reserve(ctx, ctx->term_start_alt, nb + 1)) || (rc = reserve(ctx, ctx->bead_terms_alt, ntl)) ||
        (rc = reserve(ctx, ctx->terms_alt, nterm)) || (rc = reserve(ctx, ctx->wflags_alt, W)) ||
        (rc = reserve(ctx, ctx->remap, W)) || (rc = reserve(ctx, ctx->move_segs, W)))
        return rc;

    // ---- uploads: pending worlds' state and bonded-term tables, host tables
    size_t need = staged_bytes(wd) + staged_bytes(cpar) + staged_bytes(bb) + staged_bytes(fb) + staged_bytes(bricks) +
                  staged_bytes(first) + staged_bytes(sb) + staged_bytes(remap) + staged_bytes(tstart_end) +
                  W * sizeof(MoveSeg) + 4 * sizeof(NextBuffers) + 2048;
    for (size_t i = 0; i < W; i++) {
        const WorldHost &w = ctx->worlds.at(order[i]);
        if (w.resident && remap[i] != ~0u) continue;
        need += 3 * staged_bytes(w.up_pos) + 2 * ((w.n + 1) * 4 + 255) + staged_bytes(w.term_list_local) +
                (w.bonds.size() + w.angles.size()) * sizeof(TermRec) + 256;
    }
    if ((rc = stage_begin(ctx, need))) return rc;

    // resident worlds: state, slot map, bonded-term tables moved on the device
    // (one launch for all of them); pending worlds: uploaded with their tables
    // in place
    std::vector<TermRec> trec;
    std::vector<uint32_t> tst, tl;
    std::vector<MoveSeg> segs;
    uint32_t move_blocks = 0;
    for (size_t i = 0; i < W; i++) {
        WorldHost &w = ctx->worlds.at(order[i]);
        const uint32_t off = new_bead_off[i];
        const uint32_t n_terms_w = (uint32_t)(w.bonds.size() + w.angles.size());
        const uint32_t n_list_w = (uint32_t)w.term_list_local.size();
        const bool moved = w.resident && remap[i] != ~0u;
        MoveSeg g{};
        g.first_block = move_blocks;
        g.world = (uint32_t)i;
        g.fresh = moved ? 0u : 1u;
        g.n = w.n;
        g.src_off = w.bead_off; g.dst_off = off;
        g.n_list = moved ? n_list_w : 0u; g.src_list = w.list_base; g.dst_list = new_list_base[i];
        g.n_terms = moved ? n_terms_w : 0u; g.src_term = w.term_base; g.dst_term = new_term_base[i];
        const uint32_t span = std::max(g.n, std::max(g.n_list, g.n_terms));
        if (span) {
            segs.push_back(g);
            move_blocks += (span + 255) / 256;
        }
        if (moved) continue;
        if (w.n) {
            for (uint32_t k = 0; k < w.n; k++) w.up_vel[k].w = u2f(off + k);
            if ((rc = stage_put(ctx, ctx->pos[dst].p + off, w.up_pos.data(), w.n * sizeof(float4))) ||
                (rc = stage_put(ctx, ctx->vel[dst].p + off, w.up_vel.data(), w.n * sizeof(float4))) ||
                (rc = stage_put(ctx, ctx->frc[dst].p + off, w.up_frc.data(), w.n * sizeof(float4))))
                return rc;
        }
        tst.resize(w.n);
        for (uint32_t b = 0; b < w.n; b++)
            tst[b] = new_list_base[i] + (w.term_start_local.empty() ? 0u : w.term_start_local[b]);
        if ((rc = stage_put(ctx, ctx->term_start_alt.p + off, tst.data(), tst.size() * 4))) return rc;
        constexpr uint32_t kIndexMask = (1u << kRoleShift) - 1u;
        tl.resize(n_list_w);
        for (size_t k = 0; k < tl.size(); k++) {
            const uint32_t e = w.term_list_local[k];
            tl[k] = (e & ~kIndexMask) | ((e & kIndexMask) + new_term_base[i]);
        }
        if ((rc = stage_put(ctx, ctx->bead_terms_alt.p + new_list_base[i], tl.data(), tl.size() * 4))) return rc;
        trec.clear();
        for (const auto *list : {&w.bonds, &w.angles})
            for (const TermRec &r0 : *list) {
                TermRec r = r0;
                r.a += off; r.b += off; r.c += off;
                r.kind = (r0.kind & 0xffu) | ((uint32_t)i << kTermWorldShift);
                trec.push_back(r);
            }
        if ((rc = stage_put(ctx, ctx->terms_alt.p + new_term_base[i], trec.data(), trec.size() * sizeof(TermRec))))
            return rc;
    }
    if (!segs.empty()) {
        if ((rc = stage_put(ctx, ctx->move_segs.p, segs.data(), segs.size() * sizeof(MoveSeg)))) return rc;
        MoveBuffers mb;
        mb.sp = ctx->pos[src].p; mb.sv = ctx->vel[src].p; mb.sf = ctx->frc[src].p;
        mb.dp = ctx->pos[dst].p; mb.dv = ctx->vel[dst].p; mb.df = ctx->frc[dst].p;
        mb.ssob = ctx->slot_of_bead.p; mb.sstart = ctx->term_start.p; mb.slist = ctx->bead_terms.p;
        mb.dsob = ctx->slot_of_bead_alt.p; mb.dstart = ctx->term_start_alt.p; mb.dlist = ctx->bead_terms_alt.p;
        mb.sterms = ctx->terms.p; mb.dterms = ctx->terms_alt.p;
        HIP_TRY(launch_move_worlds(ctx->move_segs.p, (uint32_t)segs.size(), move_blocks, mb, ctx->stream));
    }
    if ((rc = stage_put(ctx, ctx->term_start_alt.p + nb, tstart_end.data(), 4))) return rc;
    // per-world flags follow their worlds (removed worlds' are dropped)
    if ((rc = stage_put(ctx, ctx->remap.p, remap.data(), remap.size() * 4))) return rc;
    if (ctx->wflags.p && !ctx->order.empty())
        HIP_TRY(launch_remap_u32(ctx->wflags_alt.p, ctx->wflags.p, ctx->remap.p, (uint32_t)W, ctx->stream));
    else
        HIP_TRY(launch_fill_u32(ctx->wflags_alt.p, (uint32_t)std::max<size_t>(1, W), 0, ctx->stream));
    std::swap(ctx->slot_of_bead, ctx->slot_of_bead_alt);
    std::swap(ctx->term_start, ctx->term_start_alt);
    std::swap(ctx->bead_terms, ctx->bead_terms_alt);
    std::swap(ctx->terms, ctx->terms_alt);
    std::swap(ctx->wflags, ctx->wflags_alt);
    if ((rc = reserve(ctx, ctx->pos[src], nb)) || (rc = reserve(ctx, ctx->vel[src], nb)) ||
        (rc = reserve(ctx, ctx->frc[src], nb)))
        return rc;
    ctx->cur = dst;
    ctx->cpar = 0;

    // ---- remaining per-bead, per-cell and per-world buffers
    if ((rc = reserve(ctx, ctx->key, nb)) || (rc = reserve(ctx, ctx->perm, nb)) ||
        (rc = reserve(ctx, ctx->sorted_idx, nb)) || (rc = reserve(ctx, ctx->cell_count, nc)) ||
        (rc = reserve(ctx, ctx->cell_start, nc + 1)) || (rc = reserve(ctx, ctx->scan_tiles, scan_tiles((uint32_t)nc))) ||
        (rc = reserve(ctx, ctx->tile_sum, scan_tiles((uint32_t)nc))) || (rc = reserve(ctx, ctx->wdev, W)) ||
        (rc = reserve(ctx, ctx->cell_par, cpar.size())) || (rc = reserve(ctx, ctx->ctr, 2 * std::max<size_t>(1, W))) ||
        (rc = reserve(ctx, ctx->bblocks, bb.size())) || (rc = reserve(ctx, ctx->fblocks, fb.size())) ||
        (rc = reserve(ctx, ctx->bricks, bricks.size())) || (rc = reserve(ctx, ctx->obs_first, first.size())) ||
        (rc = reserve(ctx, ctx->term_frc, std::max<size_t>(1, nterm) * kTermFrcStride)) ||
        (rc = reserve(ctx, ctx->stationary, sb.size())) || (rc = reserve(ctx, ctx->next, 4)))
        return rc;
    if (!ctx->flags.p) {
        HIP_TRY(dev_alloc(ctx->flags, 1));
        HIP_TRY(launch_fill_u32(ctx->flags.p, 1, 0, ctx->stream));
    }
    if (!ctx->scan_ticket.p) HIP_TRY(dev_alloc(ctx->scan_ticket, 1));
    HIP_TRY(launch_fill_u32(ctx->cell_count.p, (uint32_t)nc, 0, ctx->stream));
    if ((rc = stage_put(ctx, ctx->wdev.p, wd.data(), wd.size() * sizeof(WorldDev))) ||
        (rc = stage_put(ctx, ctx->cell_par.p, cpar.data(), cpar.size() * sizeof(float2))) ||
        (rc = stage_put(ctx, ctx->bblocks.p, bb.data(), bb.size() * sizeof(BeadBlock))) ||
        (rc = stage_put(ctx, ctx->fblocks.p, fb.data(), fb.size() * sizeof(BeadBlock))) ||
        (rc = stage_put(ctx, ctx->bricks.p, bricks.data(), bricks.size() * sizeof(BrickRec))) ||
        (rc = stage_put(ctx, ctx->obs_first.p, first.data(), first.size() * 4)) ||
        (rc = stage_put(ctx, ctx->stationary.p, sb.data(), sb.size() * sizeof(StationaryBead))))
        return rc;
    std::vector<NextBuffers> nbuf(4);
    for (int c = 0; c < 2; c++)
        for (int cp = 0; cp < 2; cp++) {
            NextBuffers &b = nbuf[c * 2 + cp];
            b.pos = ctx->pos[c].p; b.vel = ctx->vel[c].p; b.frc = ctx->frc[c].p;
            b.key = ctx->key.p; b.cell_count = ctx->cell_count.p; b.flags = ctx->flags.p; b.wflags = ctx->wflags.p;
            b.scan_tiles = ctx->scan_tiles.p; b.scan_ticket = ctx->scan_ticket.p;
            b.ctr = ctx->ctr.p + (cp ^ 1) * std::max<size_t>(1, W);
            b.n_tiles = scan_tiles((uint32_t)nc);
        }
    if ((rc = stage_put(ctx, ctx->next.p, nbuf.data(), nbuf.size() * sizeof(NextBuffers)))) return rc;
    if ((rc = stage_end(ctx))) return rc;

    // ---- host bookkeeping
    for (size_t i = 0; i < W; i++) {
        WorldHost &w = ctx->worlds.at(order[i]);
        w.bead_off = new_bead_off[i];
        w.cell_off = new_cell_off[i];
        w.list_base = new_list_base[i];
        w.term_base = new_term_base[i];
        w.t_at_layout = w.t;
        if (!w.resident) {
            std::vector<float4>().swap(w.up_pos);
            std::vector<float4>().swap(w.up_vel);
            std::vector<float4>().swap(w.up_frc);
            w.resident = true;
        }
    }
    ctx->order = order;
    ctx->n_beads = (uint32_t)nb;
    ctx->n_cells = (uint32_t)nc;
    ctx->n_worlds = (uint32_t)W;
    ctx->n_bblocks = (uint32_t)bb.size();
    ctx->n_fblocks = (uint32_t)fb.size();
    ctx->n_bricks = (uint32_t)bricks.size();
    ctx->n_stationary = (uint32_t)sb.size();
    ctx->n_terms = (uint32_t)nterm;
    ctx->spatial = !cpar.empty();
    ctx->ctr_host.assign(2 * std::max<size_t>(1, W), WorldCtr{0, 0});
    ctx->stats.n_worlds = ctx->n_worlds;
    ctx->stats.n_beads = ctx->n_beads;
    ctx->stats.n_force_blocks = ctx->n_fblocks;
    const double took = now_s() - t_start;
    ctx->stats.relayouts++;
    ctx->stats.relayout_s += took;
    if (took > ctx->stats.relayout_max_s) {
        ctx->stats.relayout_max_s = took;
        ctx->stats.relayout_max_beads = nb;
    }
    ctx->layout_version++;
    ctx->steps_on_layout = 0;
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
    ctx->all_active = nsteps;
    for (uint32_t i = 0; i < W; i++) {
        const WorldHost &w = ctx->worlds.at(ctx->order[i]);
        WorldCtr c;
        c.t = (uint32_t)w.t;
        c.remaining = std::min(w.budget, nsteps);
        ctx->all_active = std::min(ctx->all_active, c.remaining);
        ctx->ctr_host[ctx->cpar * W + i] = c;
        bs += (uint64_t)w.n * c.remaining;
    }
    if (W) HIP_TRY(hipMemcpyAsync(ctx->ctr.p + ctx->cpar * W, ctx->ctr_host.data() + ctx->cpar * W,
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

// A chain of len steps from buffer c: the first step integrates with
// k_integrate, the pair-force kernel of every other step's predecessor
// integrates it; every step sorts from buffer c into c ^ 1, so the chain ends
// in c ^ 1. The step counters alternate every step. With ev, each step is
// timed and waited for (profiling).
hipError_t launch_chain(const pdpd_ctx *ctx, int &c, int &cp, uint32_t len, KernelEvents *ev, double *acc)
{
    for (uint32_t i = 0; i < len; i++) {
        hipError_t e = launch_step(ctx->step_launch(c, cp, i == 0, i + 1 < len), ctx->stream, ev);
        if (e != hipSuccess) return e;
        cp ^= 1;
        if (ev) {
            e = hipEventSynchronize(ev->ev[6]);
            if (e != hipSuccess) return e;
            for (int k = 0; k < 6; k++) {
                float ms = 0;
                e = hipEventElapsedTime(&ms, ev->ev[k], ev->ev[k + 1]);
                if (e != hipSuccess) return e;
                acc[k] += ms;
            }
        }
    }
    if (len) c ^= 1;
    return hipSuccess;
}

// nsteps steps from the run's step `first` on: one chain while every world
// steps, then single steps (a step budget runs out in them).
int launch_direct(pdpd_ctx *ctx, uint32_t first, uint32_t nsteps)
{
    const uint32_t chained =
        ctx->fuse() && first < ctx->all_active ? std::min(nsteps, ctx->all_active - first) : 0;
    HIP_TRY(launch_chain(ctx, ctx->cur, ctx->cpar, chained, nullptr, nullptr));
    for (uint32_t k = chained; k < nsteps; k++) HIP_TRY(launch_chain(ctx, ctx->cur, ctx->cpar, 1, nullptr, nullptr));
    return PDPD_OK;
}

int ensure_graph(pdpd_ctx *ctx, uint32_t len)
{
    if (ctx->graph_exec && ctx->graph_version == ctx->layout_version && ctx->graph_cur == ctx->cur &&
        ctx->graph_cpar == ctx->cpar && ctx->graph_len == len)
        return PDPD_OK;
    if (ctx->graph_exec) {
        HIP_TRY(hipGraphExecDestroy(ctx->graph_exec));
        ctx->graph_exec = nullptr;
    }
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    // two chains of len / 2 steps (len is even), or len single steps: buffers
    // and counters end where they started, so the graph replays back to back
    int c = ctx->cur, cp = ctx->cpar;
    const bool f = ctx->fuse();
    for (uint32_t h = 0; h < (f ? 2u : len); h++) {
        hipError_t e = launch_chain(ctx, c, cp, f ? len / 2 : 1u, nullptr, nullptr);
        if (e != hipSuccess) {
            hipGraph_t junk = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &junk);
            if (junk) (void)hipGraphDestroy(junk);
            return fail(PDPD_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
        }
    }
    HIP_TRY(hipStreamEndCapture(ctx->stream, &g));
    hipError_t e = hipGraphInstantiate(&ctx->graph_exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(PDPD_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    ctx->graph_version = ctx->layout_version;
    ctx->graph_cur = ctx->cur;
    ctx->graph_cpar = ctx->cpar;
    ctx->graph_len = len;
    return PDPD_OK;
}

// Enqueue nsteps: whole graphs of graph_steps (even, so buffer parity returns)
// then single steps.
int enqueue_steps(pdpd_ctx *ctx, uint32_t nsteps)
{
    uint32_t done = 0;
    const uint32_t G = ctx->graph_steps & ~1u;
    // graphs only over steps in which every world steps (chains must not
    // span the end of a step budget)
    if (G >= 2 && ctx->all_active >= G) {
        int rc = ensure_graph(ctx, G);
        if (rc) return rc;
        while (ctx->all_active - done >= G) {
            HIP_TRY(hipGraphLaunch(ctx->graph_exec, ctx->stream));
            done += G;
        }
    }
    return launch_direct(ctx, done, nsteps - done);
}

int check_flags(pdpd_ctx *ctx)
{
    uint32_t f = 0, any = 0;
    HIP_TRY(hipMemcpy(&f, ctx->flags.p, sizeof(f), hipMemcpyDeviceToHost));
    if (int rc = harvest_world_flags(ctx, &any)) return rc;
    ctx->stats.flags |= f | any;
    if (f) {
        HIP_TRY(launch_fill_u32(ctx->flags.p, 1, 0, ctx->stream));
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        if (f & PDPD_FLAG_SCAN_TIMEOUT) return fail(PDPD_ERR_STATE, "cell scan look-back did not complete");
    }
    // every resident world that has raised a flag (now or before) fails the call, named
    std::string msg;
    for (int slot : ctx->order) {
        auto it = ctx->worlds.find(slot);
        if (it == ctx->worlds.end() || !it->second.flags) continue;
        msg += (msg.empty() ? "" : " | ") + std::string("world ") + std::to_string(slot) + ": " +
               flag_text(it->second.flags);
    }
    if (!msg.empty()) return fail(PDPD_ERR_STATE, msg);
    return PDPD_OK;
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
    for (uint32_t p = 0; p < w->n_polymers; p++) {
        uint32_t len = w->polymer_bead_start[p + 1] - w->polymer_bead_start[p];
        if (len > kMaxPolymerLength) return "Polymer longer than 127 beads.";
    }
    if (w->n_beads >= kMaxLayoutBeads) return "World has more than 2^26 beads.";
    return "";
}

// Download the three state arrays of a world (current buffers).
int fetch_world(pdpd_ctx *ctx, const WorldHost &w, std::vector<float4> &P, std::vector<float4> &V,
                std::vector<float4> &F)
{
    P.resize(w.n); V.resize(w.n); F.resize(w.n);
    if (!w.n) return PDPD_OK;
    const int c = ctx->cur;
    HIP_TRY(hipMemcpy(P.data(), ctx->pos[c].p + w.bead_off, w.n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(V.data(), ctx->vel[c].p + w.bead_off, w.n * sizeof(float4), hipMemcpyDeviceToHost));
    HIP_TRY(hipMemcpy(F.data(), ctx->frc[c].p + w.bead_off, w.n * sizeof(float4), hipMemcpyDeviceToHost));
    return PDPD_OK;
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
    out->max_beads = kMaxLayoutBeads - 1;
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
    if (d->stationary)
        for (uint32_t i = 0; i < nt; i++)
            if (d->stationary[i]) w.stationary |= 1u << i;
    for (uint32_t i = 0; i < nt; i++)
        for (uint32_t j = 0; j < nt; j++) {
            w.a[i * kMaxTypes + j] = (float)d->conservative[i * nt + j];
            w.sg[i * kMaxTypes + j] = (float)std::sqrt(d->dissipative[i * nt + j]);
        }
    if (d->cell_conservative || d->cell_dissipative) {
        // per-cell strengths, rounded to FP32 (the pair's mean is taken in FP32)
        if (!d->cell_conservative || !d->cell_dissipative)
            return fail(PDPD_ERR_INVALID, "cell_conservative and cell_dissipative must be given together");
        const size_t m = (size_t)w.n_cells() * nt * nt;
        w.cell_par.resize(m);
        for (size_t k = 0; k < m; k++) {
            const double a = d->cell_conservative[k], g = d->cell_dissipative[k];
            if (!std::isfinite(a) || !std::isfinite(g) || g < 0)
                return fail(PDPD_ERR_INVALID, "per-cell strengths must be finite, dissipative strengths non-negative (cell " +
                                                  std::to_string(k / ((size_t)nt * nt)) + ")");
            w.cell_par[k] = make_float2((float)a, (float)g);
        }
    }
    // import: FP64 AoS -> FP32, w = v - f dt/2 (dpd_engine_avx2_half_merge_tbb.hpp:515-524)
    const float dt = (float)d->dt, hdt = 0.5f * dt;
    w.up_pos.resize(w.n);
    w.up_vel.resize(w.n);
    w.up_frc.resize(w.n);
    for (uint32_t i = 0; i < w.n; i++) {
        float x[3], v[3], f[3];
        for (int k = 0; k < 3; k++) {
            if (!(d->x[3 * i + k] >= 0.0) || !(d->x[3 * i + k] < (double)w.L[k]))
                return fail(PDPD_ERR_INVALID, "bead " + std::to_string(i) + " outside the box");
            const float Lf = (float)w.L[k];
            x[k] = (float)d->x[3 * i + k];
            if (x[k] >= Lf) x[k] = std::nextafter(Lf, 0.0f);   // FP32 rounding of x < L up to L
            v[k] = (float)d->v[3 * i + k];
            f[k] = (float)d->f[3 * i + k];
            v[k] = std::fmaf(f[k], -hdt, v[k]);
        }
        if (d->bead_type[i] >= nt) return fail(PDPD_ERR_INVALID, "bead " + std::to_string(i) + ": bead type out of range");
        // BeadHash fields (dpd_state.hpp:18-35): 27-bit monomer id, or 20-bit polymer id + 7-bit offset
        if (d->is_monomer[i] ? d->polymer_id[i] >= (1u << 27)
                             : (d->polymer_id[i] >= (1u << 20) || d->polymer_offset[i] > kMaxPolymerLength))
            return fail(PDPD_ERR_INVALID, "bead " + std::to_string(i) + ": polymer id or offset does not fit the bead hash");
        const uint32_t h = bead_hash(d->bead_type[i], d->is_monomer[i] != 0, d->polymer_id[i], d->polymer_offset[i]);
        w.up_pos[i] = make_float4(x[0], x[1], x[2], u2f(h));
        if ((w.stationary >> d->bead_type[i]) & 1u) w.stationary_beads.push_back(i);
        w.up_vel[i] = make_float4(v[0], v[1], v[2], 0.0f);
        w.up_frc[i] = make_float4(f[0], f[1], f[2], 0.0f);
    }
    // topology in the reference's order: per polymer, its type's bonds then bond pairs
    for (uint32_t p = 0; p < d->n_polymers; p++) {
        const uint32_t pt = d->polymer_type[p];
        if (pt >= d->n_polymer_types) return fail(PDPD_ERR_INVALID, "polymer type out of range");
        const uint32_t *ids = d->polymer_bead_ids + d->polymer_bead_start[p];
        const uint32_t len = d->polymer_bead_start[p + 1] - d->polymer_bead_start[p];
        for (uint32_t b = d->pt_bond_start[pt]; b < d->pt_bond_start[pt + 1]; b++) {
            if (d->bond_head[b] >= len || d->bond_tail[b] >= len) return fail(PDPD_ERR_INVALID, "bond offset out of range");
            TermRec r{};
            r.a = ids[d->bond_head[b]];
            r.b = r.c = ids[d->bond_tail[b]];
            r.kind = kTermBond;
            r.kappa = d->bond_kappa[b];
            r.p0 = d->bond_r0[b];
            w.bonds.push_back(r);
        }
        for (uint32_t a = d->pt_angle_start[pt]; a < d->pt_angle_start[pt + 1]; a++) {
            const uint32_t hb = d->pt_bond_start[pt] + d->angle_head_bond[a];
            const uint32_t tb = d->pt_bond_start[pt] + d->angle_tail_bond[a];
            if (hb >= d->pt_bond_start[pt + 1] || tb >= d->pt_bond_start[pt + 1])
                return fail(PDPD_ERR_INVALID, "bond pair offset out of range");
            TermRec r{};
            r.a = ids[d->bond_head[hb]];        // head
            r.b = ids[d->bond_tail[hb]];        // middle
            r.c = ids[d->bond_tail[tb]];        // tail
            r.kind = kTermAngle;
            r.kappa = d->angle_kappa[a];
            r.p0 = std::cos(d->angle_theta0[a]);
            r.p1 = std::sin(d->angle_theta0[a]);
            w.angles.push_back(r);
        }
    }
    for (const auto &b : w.bonds)
        if (b.a >= w.n || b.b >= w.n) return fail(PDPD_ERR_INVALID, "bond bead id out of range");
    for (const auto &a : w.angles)
        if (a.a >= w.n || a.b >= w.n || a.c >= w.n) return fail(PDPD_ERR_INVALID, "angle bead id out of range");
    // a monomer (BeadHash monomer bit) has no bonds: the engine tracks the slots
    // of polymer beads only
    for (const auto &b : w.bonds)
        if (d->is_monomer[b.a] || d->is_monomer[b.b])
            return fail(PDPD_ERR_INVALID, "a bead in a bond is marked as a monomer");
    for (const auto &a : w.angles)
        if (d->is_monomer[a.a] || d->is_monomer[a.b] || d->is_monomer[a.c])
            return fail(PDPD_ERR_INVALID, "a bead in an angle bond is marked as a monomer");
    // per bead its terms in the order they apply (bonds, then angles), by counting
    if (!w.bonds.empty() || !w.angles.empty()) {
        auto roles = [&](auto fn) {
            uint32_t k = 0;
            for (const TermRec &r : w.bonds) { fn(r.a, kRoleBondHead, k); fn(r.b, kRoleBondTail, k); k++; }
            for (const TermRec &r : w.angles) {
                fn(r.a, kRoleAngleHead, k); fn(r.c, kRoleAngleTail, k); fn(r.b, kRoleAngleMid, k); k++;
            }
        };
        w.term_start_local.assign(w.n + 1, 0);
        roles([&](uint32_t bead, uint32_t, uint32_t) { w.term_start_local[bead + 1]++; });
        for (uint32_t b = 0; b < w.n; b++) w.term_start_local[b + 1] += w.term_start_local[b];
        w.term_list_local.resize(w.term_start_local[w.n]);
        std::vector<uint32_t> at(w.term_start_local.begin(), w.term_start_local.end() - 1);
        roles([&](uint32_t bead, uint32_t role, uint32_t k) { w.term_list_local[at[bead]++] = (role << kRoleShift) | k; });
    }

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
    rc = enqueue_steps(ctx, nsteps);
    if (rc) return rc;
    finish_counters(ctx, nsteps);
    ctx->stats.steps += nsteps;
    ctx->steps_on_layout += nsteps;
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
    free_graveyard(ctx);
    ctx->stats.run_s += now_s() - t0;
    if (ctx->flags.p) return check_flags(ctx);
    return PDPD_OK;
}

int pdpd_world_flags(pdpd_ctx *ctx, int slot, uint32_t *flags)
{
    WorldHost *w = ctx ? find_world(ctx, slot) : nullptr;
    if (!w || !flags) return fail(PDPD_ERR_INVALID, !w ? "unknown slot" : "null output");
    HIP_TRY(hipSetDevice(ctx->device));
    if (w->resident && !ctx->dirty) {
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        uint32_t any;
        if (int rc = harvest_world_flags(ctx, &any)) return rc;
    }
    *flags = w->flags;
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
    std::vector<float4> P, V, F;
    rc = fetch_world(ctx, *w, P, V, F);
    if (rc) return rc;
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

int pdpd_snapshot(pdpd_ctx *ctx)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    HIP_TRY(hipSetDevice(ctx->device));
    // the worlds of the current device layout; worlds added since (not yet on
    // the device) are not in it, and no relayout is forced here
    if (!ctx->copy_stream) HIP_TRY(hipStreamCreateWithFlags(&ctx->copy_stream, hipStreamNonBlocking));
    if (!ctx->snap_event) HIP_TRY(hipEventCreateWithFlags(&ctx->snap_event, hipEventDisableTiming));
    const size_t n = ctx->n_beads;
    if (ctx->snap.n < 3 * n || !ctx->snap.p) {
        // the stream may still read the old buffer: let it finish first
        if (ctx->snap.p) HIP_TRY(hipEventSynchronize(ctx->snap_event));
        HIP_TRY(dev_alloc(ctx->snap, 3 * n));
    }
    if (ctx->snap_flags.n < 1 + (size_t)ctx->n_worlds) {
        if (ctx->snap_flags.p) HIP_TRY(hipEventSynchronize(ctx->snap_event));
        HIP_TRY(dev_alloc(ctx->snap_flags, 1 + (size_t)ctx->n_worlds));
    }
    const int c = ctx->cur;        // the buffer the enqueued steps end in
    if (n) {
        HIP_TRY(hipMemcpyAsync(ctx->snap.p, ctx->pos[c].p, n * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->snap.p + n, ctx->vel[c].p, n * sizeof(float4), hipMemcpyDeviceToDevice,
                               ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->snap.p + 2 * n, ctx->frc[c].p, n * sizeof(float4), hipMemcpyDeviceToDevice,
                               ctx->stream));
    }
    if (ctx->flags.p)
        HIP_TRY(hipMemcpyAsync(ctx->snap_flags.p, ctx->flags.p, sizeof(uint32_t), hipMemcpyDeviceToDevice, ctx->stream));
    else
        HIP_TRY(launch_fill_u32(ctx->snap_flags.p, 1, 0, ctx->stream));
    if (ctx->n_worlds && ctx->wflags.p)
        HIP_TRY(hipMemcpyAsync(ctx->snap_flags.p + 1, ctx->wflags.p, ctx->n_worlds * sizeof(uint32_t),
                               hipMemcpyDeviceToDevice, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->snap_event, ctx->stream));
    ctx->snap_has_obs = false;
    ctx->snap_beads = (uint32_t)n;
    ctx->snap_worlds.clear();
    for (uint32_t i = 0; i < ctx->order.size(); i++) {      // worlds removed since the layout was built are skipped
        auto it = ctx->worlds.find(ctx->order[i]);
        if (it == ctx->worlds.end() || !it->second.resident) continue;
        const WorldHost &w = it->second;
        ctx->snap_worlds[ctx->order[i]] = {i, w.bead_off, w.n, w.t, w.dt};
    }
    return PDPD_OK;
}

int pdpd_world_snapshot_download(pdpd_ctx *ctx, int slot, double *x, double *v, double *f, int64_t *t)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    auto it = ctx->snap_worlds.find(slot);
    if (it == ctx->snap_worlds.end()) return fail(PDPD_ERR_INVALID, "no snapshot of this slot");
    const pdpd_ctx::SnapWorld sw = it->second;
    HIP_TRY(hipSetDevice(ctx->device));
    const size_t n = sw.n, N = ctx->snap_beads;
    if (ctx->snap_host_n < 3 * n + 1) {
        if (ctx->snap_host) HIP_TRY(hipHostFree(ctx->snap_host));
        ctx->snap_host = nullptr;
        ctx->snap_host_n = 0;
        HIP_TRY(hipHostMalloc((void **)&ctx->snap_host, (3 * n + 1) * sizeof(float4), hipHostMallocDefault));
        ctx->snap_host_n = 3 * n + 1;
    }
    HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->snap_event, 0));
    for (int a = 0; a < 3 && n; a++)
        HIP_TRY(hipMemcpyAsync(ctx->snap_host + a * n, ctx->snap.p + a * N + sw.bead_off, n * sizeof(float4),
                               hipMemcpyDeviceToHost, ctx->copy_stream));
    uint32_t *fh = reinterpret_cast<uint32_t *>(ctx->snap_host + 3 * n);      // [context-wide, this world]
    HIP_TRY(hipMemcpyAsync(fh, ctx->snap_flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->copy_stream));
    HIP_TRY(hipMemcpyAsync(fh + 1, ctx->snap_flags.p + 1 + sw.index, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->copy_stream));
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    const uint32_t gflags = fh[0], wflags = fh[1];
    if (gflags & PDPD_FLAG_SCAN_TIMEOUT) return fail(PDPD_ERR_STATE, "cell scan look-back did not complete");
    if (wflags) {
        if (WorldHost *w = find_world(ctx, slot)) w->flags |= wflags;
        return fail(PDPD_ERR_STATE, "world " + std::to_string(slot) + ": " + flag_text(wflags));
    }
    const float4 *P = ctx->snap_host, *V = ctx->snap_host + n, *F = ctx->snap_host + 2 * n;
    const double hdt = (double)(float)sw.dt * 0.5;
    for (uint32_t s = 0; s < n; s++) {
        const uint32_t i = f2u(V[s].w) - sw.bead_off;
        if (i >= n) return fail(PDPD_ERR_STATE, "corrupt bead index in device state");
        const double px[3] = {P[s].x, P[s].y, P[s].z};
        const double fx[3] = {F[s].x, F[s].y, F[s].z};
        const double wx[3] = {V[s].x, V[s].y, V[s].z};
        for (int k = 0; k < 3; k++) {
            if (x) x[3 * i + k] = px[k];
            if (f) f[3 * i + k] = fx[k];
            if (v) v[3 * i + k] = wx[k] + fx[k] * hdt;
        }
    }
    if (t) *t = sw.t;
    return PDPD_OK;
}

static void fill_observables(const WorldHost &w, uint32_t n, int64_t t, const double *r, pdpd_observables *o);

int pdpd_snapshot_observables(pdpd_ctx *ctx)
{
    int rc = pdpd_snapshot(ctx);
    if (rc) return rc;
    if (!ctx->snap_obs_event) HIP_TRY(hipEventCreateWithFlags(&ctx->snap_obs_event, hipEventDisableTiming));
    const size_t np = std::max<size_t>(1, ctx->n_bblocks) * kObsValues;
    const size_t no = std::max<size_t>(1, ctx->n_worlds) * kObsValues;
    if (ctx->snap_obs_partial.n < np || ctx->snap_obs.n < no) {
        // the stream may still use the old buffers: let their last use finish first
        HIP_TRY(hipEventSynchronize(ctx->snap_obs_event));
        if (ctx->snap_obs_partial.n < np) HIP_TRY(dev_alloc(ctx->snap_obs_partial, np));
        if (ctx->snap_obs.n < no) HIP_TRY(dev_alloc(ctx->snap_obs, no));
    }
    if (!ctx->snap_obs_host)
        HIP_TRY(hipHostMalloc((void **)&ctx->snap_obs_host, (kObsValues + 2) * sizeof(double), hipHostMallocDefault));
    // the same state the snapshot copied: the buffer the enqueued steps end in
    if (ctx->n_worlds)
        HIP_TRY(launch_observables(ctx->state(), ctx->cur, ctx->n_bblocks, ctx->obs_first.p, ctx->n_worlds,
                                   ctx->snap_obs_partial.p, ctx->snap_obs.p, ctx->stream));
    HIP_TRY(hipEventRecord(ctx->snap_obs_event, ctx->stream));
    ctx->snap_has_obs = true;
    return PDPD_OK;
}

int pdpd_world_snapshot_observables(pdpd_ctx *ctx, int slot, pdpd_observables *out)
{
    if (!ctx || !out) return fail(PDPD_ERR_INVALID, "null argument");
    if (!ctx->snap_has_obs) return fail(PDPD_ERR_STATE, "no snapshot with observables (pdpd_snapshot_observables)");
    auto it = ctx->snap_worlds.find(slot);
    if (it == ctx->snap_worlds.end()) return fail(PDPD_ERR_INVALID, "no snapshot of this slot");
    const pdpd_ctx::SnapWorld sw = it->second;
    WorldHost *w = find_world(ctx, slot);
    if (!w) return fail(PDPD_ERR_INVALID, "unknown slot");
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamWaitEvent(ctx->copy_stream, ctx->snap_obs_event, 0));
    double *h = ctx->snap_obs_host;
    HIP_TRY(hipMemcpyAsync(h, ctx->snap_obs.p + (size_t)sw.index * kObsValues, kObsValues * sizeof(double),
                           hipMemcpyDeviceToHost, ctx->copy_stream));
    uint32_t *fh = reinterpret_cast<uint32_t *>(h + kObsValues);      // [context-wide, this world]
    HIP_TRY(hipMemcpyAsync(fh, ctx->snap_flags.p, sizeof(uint32_t), hipMemcpyDeviceToHost, ctx->copy_stream));
    HIP_TRY(hipMemcpyAsync(fh + 1, ctx->snap_flags.p + 1 + sw.index, sizeof(uint32_t), hipMemcpyDeviceToHost,
                           ctx->copy_stream));
    HIP_TRY(hipStreamSynchronize(ctx->copy_stream));
    if (fh[0] & PDPD_FLAG_SCAN_TIMEOUT) return fail(PDPD_ERR_STATE, "cell scan look-back did not complete");
    if (fh[1]) {
        w->flags |= fh[1];
        return fail(PDPD_ERR_STATE, "world " + std::to_string(slot) + ": " + flag_text(fh[1]));
    }
    fill_observables(*w, sw.n, sw.t, h, out);     // the snapshot's bead count and time
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
    std::vector<float4> P, V, F;
    rc = fetch_world(ctx, *w, P, V, F);
    if (rc) return rc;
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

static int parse_u32(const std::string &v, uint32_t lo, uint32_t hi, uint32_t *out)
{
    char *end = nullptr;
    long long n = strtoll(v.c_str(), &end, 10);
    if (v.empty() || !end || *end || n < (long long)lo || n > (long long)hi) return -1;
    *out = (uint32_t)n;
    return 0;
}

// Re-lists the force blocks of the current layout in the order band_rows asks
// for (a layout not yet built takes it when it is).
static int rebuild_force_blocks(pdpd_ctx *ctx)
{
    if (ctx->dirty || ctx->n_fblocks == 0) return PDPD_OK;
    HIP_TRY(hipSetDevice(ctx->device));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    std::vector<uint32_t> off;
    for (int slot : ctx->order) off.push_back(ctx->worlds.at(slot).bead_off);
    std::vector<BeadBlock> bb, fb;
    build_blocks(ctx, ctx->order, off, bb, fb);
    if ((uint32_t)fb.size() != ctx->n_fblocks) return fail(PDPD_ERR_STATE, "force block count changed");
    HIP_TRY(hipMemcpy(ctx->fblocks.p, fb.data(), fb.size() * sizeof(BeadBlock), hipMemcpyHostToDevice));
    return PDPD_OK;
}

int pdpd_set_config(pdpd_ctx *ctx, const char *key, const char *value)
{
    if (!ctx || !key || !value) return fail(PDPD_ERR_INVALID, "null argument");
    std::string k(key), v(value);
    if (k == "graph_steps") {
        uint32_t n;
        if (parse_u32(v, 0, 100000, &n)) return fail(PDPD_ERR_INVALID, "graph_steps must be an integer in [0, 100000]");
        ctx->graph_steps = n;
        return PDPD_OK;
    }
    if (k == "force_grid") {
        uint32_t n;
        if (parse_u32(v, 0, 1u << 24, &n)) return fail(PDPD_ERR_INVALID, "force_grid must be an integer >= 0");
        ctx->force_grid = n;
        ctx->graph_version = ~0ull;
        return PDPD_OK;
    }
    if (k == "scan_lookback_tiles") {
        uint32_t n;
        if (parse_u32(v, 0, 1u << 20, &n)) return fail(PDPD_ERR_INVALID, "scan_lookback_tiles must be an integer >= 0");
        ctx->lookback_tiles = n;
        ctx->graph_version = ~0ull;
        return PDPD_OK;
    }
    if (k == "fused_integrate") {
        uint32_t n;
        if (parse_u32(v, 0, 2, &n)) return fail(PDPD_ERR_INVALID, "fused_integrate must be 0, 1 or 2");
        ctx->fuse_mode = n;
        ctx->graph_version = ~0ull;
        return PDPD_OK;
    }
    if (k == "force_band_rows") {
        uint32_t n;
        if (parse_u32(v, 0, 1u << 16, &n)) return fail(PDPD_ERR_INVALID, "force_band_rows must be an integer >= 0");
        if (n != ctx->band_rows) {
            ctx->band_rows = n;
            const int rc = rebuild_force_blocks(ctx);
            if (rc != PDPD_OK) return rc;
        }
        return PDPD_OK;
    }
    if (k == "reserve_beads") {
        uint32_t n;
        if (parse_u32(v, 0, kMaxLayoutBeads - 1, &n))
            return fail(PDPD_ERR_INVALID, "reserve_beads must be an integer in [0, 2^26)");
        if (!ctx->worlds.empty() || !ctx->order.empty())
            return fail(PDPD_ERR_STATE, "reserve_beads must be set before any world is added");
        HIP_TRY(hipSetDevice(ctx->device));
        if (int rc = reserve_capacity(ctx, n)) return rc;
        ctx->reserved_beads = std::max(ctx->reserved_beads, n);
        return PDPD_OK;
    }
    if (k == "force_kernel") {
        uint32_t n;
        if (parse_u32(v, 0, 1, &n)) return fail(PDPD_ERR_INVALID, "force_kernel must be 0 or 1");
        ctx->force_kernel = n;
        ctx->graph_version = ~0ull;
        return PDPD_OK;
    }
    if (k == "force_brick") {
        uint32_t n;
        if (parse_u32(v, 0, brick_shape_count() - 1, &n))
            return fail(PDPD_ERR_INVALID, "force_brick must be an integer in [0, " + std::to_string(brick_shape_count() - 1) + "]");
        if (n != ctx->brick_shape) {
            ctx->brick_shape = n;
            ctx->graph_version = ~0ull;
            if (!ctx->dirty) {     // re-list the bricks of the current layout
                HIP_TRY(hipSetDevice(ctx->device));
                HIP_TRY(hipStreamSynchronize(ctx->stream));
                if (int rc = upload_bricks(ctx)) return rc;
            }
        }
        return PDPD_OK;
    }
    if (k == "force_variant") {
        uint32_t n = kVariantAuto;
        if (v != "auto" && parse_u32(v, 0, kForceVariants - 1, &n))
            return fail(PDPD_ERR_INVALID, "force_variant must be 'auto' or an integer in [0, " + std::to_string(kForceVariants - 1) + "]");
        ctx->force_variant = n;
        ctx->graph_version = ~0ull;
        return PDPD_OK;
    }
    return fail(PDPD_ERR_INVALID, "Unknown engine config name '" + k + "'");
}

int pdpd_get_config(pdpd_ctx *ctx, const char *key, char *buf, size_t buflen)
{
    if (!ctx || !key || !buf || !buflen) return fail(PDPD_ERR_INVALID, "null argument");
    std::string k(key), v;
    if (k == "graph_steps") v = std::to_string(ctx->graph_steps);
    else if (k == "force_grid") v = std::to_string(ctx->force_grid);
    else if (k == "force_variant") v = ctx->force_variant == kVariantAuto ? "auto" : std::to_string(ctx->force_variant);
    else if (k == "scan_lookback_tiles") v = std::to_string(ctx->lookback_tiles);
    else if (k == "fused_integrate") v = std::to_string(ctx->fuse_mode);
    else if (k == "force_band_rows") v = std::to_string(ctx->band_rows);
    else if (k == "force_kernel") v = std::to_string(ctx->force_kernel);
    else if (k == "force_brick") v = std::to_string(ctx->brick_shape);
    else if (k == "force_brick_shapes") v = std::to_string(brick_shape_count());
    else if (k == "reserve_beads") v = std::to_string(ctx->reserved_beads);
    else return fail(PDPD_ERR_INVALID, "Unknown engine config name '" + k + "'");
    snprintf(buf, buflen, "%s", v.c_str());
    return PDPD_OK;
}

int pdpd_list_config(pdpd_ctx *ctx, char *buf, size_t buflen)
{
    if (!ctx || !buf || !buflen) return fail(PDPD_ERR_INVALID, "null argument");
    std::string s;
    s += "graph_steps=" + std::to_string(ctx->graph_steps) + "\tsteps per captured hipGraph (even; 0 = direct launches)\n";
    s += "force_grid=" + std::to_string(ctx->force_grid) +
         "\tpair-force grid for a persistent grid-stride launch (0 = one workgroup per 64 beads)\n";
    s += "scan_lookback_tiles=" + std::to_string(ctx->lookback_tiles) +
         "\tcell scan: single pass with look-back up to this many 512-cell tiles, three phases above\n";
    s += "force_variant=" + (ctx->force_variant == kVariantAuto ? std::string("auto") : std::to_string(ctx->force_variant)) +
         "\tpair-force kernel variant: 0 = 3 candidate loads in flight, bonded sums after the middle plane's pairs;"
         " 1 = 4 loads, bonded sums by a fourth wave; 2 = 2 loads, middle plane; 3 = 3 loads, fourth wave;"
         " 4 = 2 loads, middle plane, 128-pair queues, 6 waves per SIMD; 5 = 3 loads, 192-pair queues, 6 waves;"
         " 6 = 2 loads, fourth wave, 128-pair queues, 6 waves; 7 = one wave per stencil row (nine per workgroup),"
         " 8 loads, 512-pair queues; auto = 4 from " + std::to_string(kVariantLargeBeads) +
         " resident beads, 7 below " + std::to_string(kVariantSmallBeads) + ", else 0\n";
    s += "fused_integrate=" + std::to_string(ctx->fuse_mode) +
         "\tchained steps, the pair-force kernel integrating the next step: 0 = off, 1 = on, 2 = on up to " +
         std::to_string(kFuseAutoBeads) + " resident beads\n";
    s += "force_band_rows=" + std::to_string(ctx->band_rows) +
         "\tpair-force work order of worlds with at least " + std::to_string(kBandMinCells) +
         " cells per y-z plane: bands of this many y rows, plane by plane (0 = layout order)\n";
    s += "force_kernel=" + std::to_string(ctx->force_kernel) +
         "\tpair-force kernel: 0 = per-slot blocks of 64 beads, candidates gathered from HBM; 1 = cell bricks"
         " staged in LDS with their halo (worlds with per-cell strengths always use 0)\n";
    std::string shapes;
    for (uint32_t i = 0; i < brick_shape_count(); i++) {
        uint32_t a, b, c;
        brick_shape_dims(i, &a, &b, &c);
        shapes += (i ? ", " : "") + std::to_string(i) + " = " + std::to_string(a) + "x" + std::to_string(b) + "x" +
                  std::to_string(c);
    }
    s += "force_brick=" + std::to_string(ctx->brick_shape) + "\tbrick shape in cells (x, y, z): " + shapes + "\n";
    s += "reserve_beads=" + std::to_string(ctx->reserved_beads) +
         "\tdevice and pinned staging capacity set aside for this many resident beads (layouts up to it are built"
         " without allocating)\n";
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
    // chains as a replayed graph runs them, while every world steps
    const uint32_t chain = ctx->fuse() ? std::max<uint32_t>(1, (ctx->graph_steps & ~1u) / 2) : 1u;
    for (uint32_t s = 0; s < nsteps;) {
        const uint32_t len = s < ctx->all_active ? std::min(chain, ctx->all_active - s) : 1;
        HIP_TRY(launch_chain(ctx, ctx->cur, ctx->cpar, len, &ev, acc));
        s += len;
    }
    for (int k = 0; k < 7; k++) (void)hipEventDestroy(ev.ev[k]);
    finish_counters(ctx, nsteps);
    ctx->stats.steps += nsteps;
    ctx->steps_on_layout += nsteps;
    ctx->stats.bead_steps += bs;
    ctx->stats.k_integrate_ms = acc[0] / nsteps;
    ctx->stats.k_scan_ms = acc[1] / nsteps;
    ctx->stats.k_scatter_ms = acc[2] / nsteps;
    ctx->stats.k_reorder_ms = acc[3] / nsteps;
    ctx->stats.k_force_ms = acc[4] / nsteps;
    ctx->stats.k_stationary_ms = acc[5] / nsteps;
    ctx->stats.profiled_steps = nsteps;
    return check_flags(ctx);
}

int pdpd_time_force(pdpd_ctx *ctx, uint32_t reps, double *avg_ms)
{
    if (!ctx || !avg_ms) return fail(PDPD_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->dirty || ctx->n_worlds == 0 || ctx->steps_on_layout == 0)
        return fail(PDPD_ERR_STATE, "time_force needs a completed step on the current layout");
    reps = std::max<uint32_t>(1, reps);
    // the step that produced the current buffers (the last of a chain)
    const StepLaunch L = ctx->step_launch(ctx->cur ^ 1, ctx->cpar ^ 1, false, false);
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    HIP_TRY(launch_force(L, ctx->stream));                  // warm
    HIP_TRY(hipEventRecord(e0, ctx->stream));
    for (uint32_t r = 0; r < reps; r++) HIP_TRY(launch_force(L, ctx->stream));
    HIP_TRY(hipEventRecord(e1, ctx->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float ms = 0;
    HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *avg_ms = ms / reps;
    return check_flags(ctx);
}

int pdpd_trace_forces(pdpd_ctx *ctx, int slot, pdpd_pair_record *pairs, uint32_t pair_capacity,
                      uint32_t *n_pairs, pdpd_bond_record *bonds, uint32_t bond_capacity, uint32_t *n_bonds)
{
    WorldHost *w = ctx ? find_world(ctx, slot) : nullptr;
    if (!w) return fail(PDPD_ERR_INVALID, "unknown slot");
    if ((pair_capacity && !pairs) || (bond_capacity && !bonds)) return fail(PDPD_ERR_INVALID, "null record buffer");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = pdpd_sync(ctx);
    if (rc) return rc;
    if (ctx->dirty || ctx->steps_on_layout == 0 || !w->resident || w->t == w->t_at_layout)
        return fail(PDPD_ERR_STATE, "force trace needs a step of this world on the current layout");
    uint32_t world = 0;
    while (world < ctx->order.size() && ctx->order[world] != slot) world++;
    // pair records: appended on the device, then sorted by (home, other)
    DevBuf<pdpd_pair_record> drec;
    DevBuf<uint32_t> dcount;
    HIP_TRY(dev_alloc(drec, pair_capacity));
    HIP_TRY(dev_alloc(dcount, 1));
    HIP_TRY(launch_fill_u32(dcount.p, 1, 0, ctx->stream));
    HIP_TRY(launch_trace_pairs(ctx->state(), ctx->cur, world, w->n, (uint32_t)(w->t - 1), drec.p, pair_capacity,
                               dcount.p, ctx->stream));
    uint32_t count = 0;
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    HIP_TRY(hipMemcpy(&count, dcount.p, sizeof(count), hipMemcpyDeviceToHost));
    if (n_pairs) *n_pairs = count;
    const uint32_t got = std::min(count, pair_capacity);
    if (got) {
        HIP_TRY(hipMemcpy(pairs, drec.p, got * sizeof(pdpd_pair_record), hipMemcpyDeviceToHost));
        std::sort(pairs, pairs + got, [](const pdpd_pair_record &a, const pdpd_pair_record &b) {
            return a.home != b.home ? a.home < b.home : a.other < b.other;
        });
    }
    drec.release();
    dcount.release();
    // Hookean bonds: positions and hashes from the world's buffers, forces as the step evaluated them
    const uint32_t nbond = (uint32_t)w->bonds.size();
    if (n_bonds) *n_bonds = nbond;
    if (bond_capacity && nbond) {
        std::vector<float4> P, V, F;
        if ((rc = fetch_world(ctx, *w, P, V, F))) return rc;
        std::vector<uint32_t> slot_of(w->n);
        for (uint32_t s = 0; s < w->n; s++) {
            const uint32_t i = f2u(V[s].w) - w->bead_off;
            if (i >= w->n) return fail(PDPD_ERR_STATE, "corrupt bead index in device state");
            slot_of[i] = s;
        }
        const uint32_t nb = std::min(nbond, bond_capacity);
        std::vector<double> tf((size_t)nb * kTermFrcStride);
        HIP_TRY(hipMemcpy(tf.data(), ctx->term_frc.p + (size_t)w->term_base * kTermFrcStride,
                          tf.size() * sizeof(double), hipMemcpyDeviceToHost));
        auto min_image = [](double d, double len) { return d > 0.5 * len ? d - len : (d < -0.5 * len ? d + len : d); };
        for (uint32_t j = 0; j < nb; j++) {
            const float4 a = P[slot_of[w->bonds[j].a]], b = P[slot_of[w->bonds[j].b]];
            pdpd_bond_record &r = bonds[j];
            r.head = f2u(a.w);
            r.tail = f2u(b.w);
            const double da[3] = {a.x, a.y, a.z}, db[3] = {b.x, b.y, b.z};
            for (int d = 0; d < 3; d++) r.dx[d] = min_image(db[d] - da[d], (double)(float)w->L[d]);
            r.dr = std::sqrt(std::fma(r.dx[2], r.dx[2], std::fma(r.dx[1], r.dx[1], r.dx[0] * r.dx[0])));
            for (int d = 0; d < 3; d++) r.f_head[d] = -tf[(size_t)j * kTermFrcStride + d];
        }
    }
    if (count > pair_capacity || nbond > bond_capacity) {
        if (pair_capacity == 0 && bond_capacity == 0) return PDPD_OK;    // a count query
        return fail(PDPD_ERR_INVALID, "force trace: record capacity too small");
    }
    return PDPD_OK;
}

// Per-world observables of the current state into ctx->obs_out (device),
// stream-ordered. partial / out: device buffers of n_bblocks and n_worlds
// entries of kObsValues doubles.
static int enqueue_observables(pdpd_ctx *ctx, double *partial, double *out)
{
    HIP_TRY(launch_observables(ctx->state(), ctx->cur, ctx->n_bblocks, ctx->obs_first.p, ctx->n_worlds, partial, out,
                               ctx->stream));
    return PDPD_OK;
}

static void fill_observables(const WorldHost &w, uint32_t n, int64_t t, const double *r, pdpd_observables *o)
{
    memset(o, 0, sizeof(*o));
    o->slot = w.slot;
    o->n_beads = n;
    o->t = t;
    for (int k = 0; k < 3; k++) o->momentum[k] = r[k];
    o->kinetic = r[3];
    o->virial = r[4];
    o->n_pairs = (uint64_t)r[5];
    const double vol = (double)w.L[0] * (double)w.L[1] * (double)w.L[2];
    o->temperature = n ? r[3] / (3.0 * (double)n) : 0.0;
    o->pressure = ((double)n * o->temperature + r[4] / 3.0) / vol;
}

int pdpd_get_observables(pdpd_ctx *ctx, pdpd_observables *out, uint32_t capacity, uint32_t *n_out)
{
    if (!ctx || (capacity && !out)) return fail(PDPD_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    if (n_out) *n_out = ctx->n_worlds;
    if (ctx->n_worlds == 0) return PDPD_OK;
    if ((rc = pdpd_sync(ctx))) return rc;
    const size_t np = std::max<size_t>(1, ctx->n_bblocks) * kObsValues, no = (size_t)ctx->n_worlds * kObsValues;
    if (ctx->obs_partial.n < np) HIP_TRY(dev_alloc(ctx->obs_partial, np));
    if (ctx->obs_out.n < no) HIP_TRY(dev_alloc(ctx->obs_out, no));
    if ((rc = enqueue_observables(ctx, ctx->obs_partial.p, ctx->obs_out.p))) return rc;
    std::vector<double> r(no);
    HIP_TRY(hipMemcpyAsync(r.data(), ctx->obs_out.p, no * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(hipStreamSynchronize(ctx->stream));
    const uint32_t n = std::min(capacity, ctx->n_worlds);
    for (uint32_t i = 0; i < n; i++) {
        const WorldHost &w = ctx->worlds.at(ctx->order[i]);
        fill_observables(w, w.n, w.t, &r[(size_t)i * kObsValues], &out[i]);
    }
    return PDPD_OK;
}

// ---- observables without allocation: partial sums in a file-scope device array

// Bead blocks (256 beads each) whose partial sums the array holds: contexts of
// up to 2^21 resident beads (about 8192 x 256), 512 KiB per GPU.
constexpr uint32_t kObsStaticBlocks = 8192;
static __device__ __align__(256) double scratch[kObsStaticBlocks * kObsValues];

// The array's device address on `device` (current), looked up once per device.
static int scratch_address(int device, double **out)
{
    static std::mutex mtx;
    static double *addr[64] = {};
    if (device < 0 || device >= 64) return fail(PDPD_ERR_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(mtx);
    if (!addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(PDPD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        addr[device] = (double *)p;
    }
    *out = addr[device];
    return PDPD_OK;
}

int pdpd_observables_noalloc(pdpd_ctx *ctx, double *out_dev, uint32_t out_worlds)
{
    if (!ctx || !out_dev) return fail(PDPD_ERR_INVALID, "null argument");
    HIP_TRY(hipSetDevice(ctx->device));
    if (ctx->dirty)
        return fail(PDPD_ERR_STATE, "the device layout must be rebuilt first (run a step or call an allocating "
                                    "entry point after adding or removing worlds)");
    if (out_worlds < ctx->n_worlds)
        return fail(PDPD_ERR_INVALID, "out_dev holds " + std::to_string(out_worlds) + " worlds, " +
                                          std::to_string(ctx->n_worlds) + " are resident");
    if (ctx->n_bblocks > kObsStaticBlocks)
        return fail(PDPD_ERR_CAPACITY, "resident worlds need " + std::to_string(ctx->n_bblocks) +
                                           " bead blocks of partial sums; the static array holds " +
                                           std::to_string(kObsStaticBlocks) + " (" +
                                           std::to_string(kObsStaticBlocks * kBeadThreads) + " beads)");
    if (ctx->n_worlds == 0) return PDPD_OK;
    double *partial = nullptr;
    if (int rc = scratch_address(ctx->device, &partial)) return rc;
    return enqueue_observables(ctx, partial, out_dev);
}

int pdpd_world_observables(pdpd_ctx *ctx, int slot, pdpd_observables *out)
{
    if (!ctx || !out) return fail(PDPD_ERR_INVALID, "null argument");
    if (!find_world(ctx, slot)) return fail(PDPD_ERR_INVALID, "unknown slot");
    int rc = ensure_layout(ctx);
    if (rc) return rc;
    std::vector<pdpd_observables> all(ctx->n_worlds);
    uint32_t n = 0;
    if ((rc = pdpd_get_observables(ctx, all.data(), (uint32_t)all.size(), &n))) return rc;
    for (uint32_t i = 0; i < n; i++)
        if (all[i].slot == slot) {
            *out = all[i];
            return PDPD_OK;
        }
    return fail(PDPD_ERR_STATE, "world not in the device layout");
}

// Times launch shapes that do not change results (the pair-force kernel,
// its grid, variant or brick shape, the graph length and chained steps) on a
// saved copy of the state, keeps the fastest. A group of contexts on one
// device (an ensemble spread over several streams) is tuned as it runs: every
// trial enqueues the steps of all of them, and its time is from the first
// context's start to the last one's end; every context takes the same shape.
static int autotune_group(const std::vector<pdpd_ctx *> &all, uint32_t steps_per_trial, char *choice,
                          size_t choice_len)
{
    std::vector<pdpd_ctx *> cs;
    for (pdpd_ctx *c : all) {
        if (!c) return fail(PDPD_ERR_INVALID, "null context");
        if (c->device != all[0]->device) return fail(PDPD_ERR_INVALID, "contexts of a group must share a device");
        HIP_TRY(hipSetDevice(c->device));
        if (int rc = ensure_layout(c)) return rc;
        if (c->n_worlds) cs.push_back(c);
    }
    if (cs.empty()) return fail(PDPD_ERR_INVALID, "no worlds to tune on");
    HIP_TRY(hipSetDevice(cs[0]->device));
    steps_per_trial = std::max<uint32_t>(4, steps_per_trial) & ~1u;
    const size_t nc = cs.size();
    struct Saved {
        DevBuf<float4> sp, sv, sf;
        DevBuf<uint32_t> ss;
        int c0 = 0, cp0 = 0;
        hipEvent_t e0 = nullptr, e1 = nullptr;
    };
    std::vector<Saved> sv(nc);
    auto cleanup = [&]() {
        for (Saved &q : sv) {
            q.sp.release(); q.sv.release(); q.sf.release(); q.ss.release();
            if (q.e0) (void)hipEventDestroy(q.e0);
            if (q.e1) (void)hipEventDestroy(q.e1);
            q.e0 = q.e1 = nullptr;
        }
    };
    struct Guard { std::function<void()> f; ~Guard() { f(); } } guard{cleanup};
    for (size_t i = 0; i < nc; i++) {
        pdpd_ctx *ctx = cs[i];
        Saved &q = sv[i];
        HIP_TRY(hipStreamSynchronize(ctx->stream));
        const size_t nb = ctx->n_beads;
        HIP_TRY(dev_alloc(q.sp, nb));
        HIP_TRY(dev_alloc(q.sv, nb));
        HIP_TRY(dev_alloc(q.sf, nb));
        HIP_TRY(dev_alloc(q.ss, nb));
        q.c0 = ctx->cur;
        q.cp0 = ctx->cpar;
        // saved and restored on the engine's stream: the stream is non-blocking,
        // so a plain hipMemcpy would not be ordered with the trial steps
        HIP_TRY(hipMemcpyAsync(q.sp.p, ctx->pos[q.c0].p, nb * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(q.sv.p, ctx->vel[q.c0].p, nb * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(q.sf.p, ctx->frc[q.c0].p, nb * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(q.ss.p, ctx->slot_of_bead.p, nb * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               ctx->stream));
        HIP_TRY(hipEventCreate(&q.e0));
        HIP_TRY(hipEventCreate(&q.e1));
    }
    auto restore = [&](size_t i) -> int {
        pdpd_ctx *ctx = cs[i];
        const Saved &q = sv[i];
        const size_t nb = ctx->n_beads;
        HIP_TRY(hipMemcpyAsync(ctx->pos[q.c0].p, q.sp.p, nb * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->vel[q.c0].p, q.sv.p, nb * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->frc[q.c0].p, q.sf.p, nb * sizeof(float4), hipMemcpyDeviceToDevice, ctx->stream));
        HIP_TRY(hipMemcpyAsync(ctx->slot_of_bead.p, q.ss.p, nb * sizeof(uint32_t), hipMemcpyDeviceToDevice,
                               ctx->stream));
        ctx->cur = q.c0;
        ctx->cpar = q.cp0;
        return PDPD_OK;
    };
    int cus = 256;
    {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, cs[0]->device) == hipSuccess && prop.multiProcessorCount > 0)
            cus = prop.multiProcessorCount;
    }
    const uint32_t grids[] = {0u, (uint32_t)cus * 4, (uint32_t)cus * 8, (uint32_t)cus * 16};
    const uint32_t graphs[] = {10u, 20u, 50u};
    pdpd_ctx *c0 = cs[0];
    uint32_t best_grid = c0->force_grid, best_graph = c0->graph_steps, best_var = c0->force_variant;
    uint32_t best_kernel = c0->force_kernel, best_shape = c0->brick_shape;
    const uint32_t fuse0 = c0->fuse() ? 1u : 0u;
    uint32_t best_fuse = fuse0;
    double best_ms = 1e30;
    auto set_shape = [&](pdpd_ctx *ctx, uint32_t kernel, uint32_t shape) -> int {
        ctx->force_kernel = kernel;
        if (shape != ctx->brick_shape) {
            HIP_TRY(hipStreamSynchronize(ctx->stream));
            ctx->brick_shape = shape;
            if (int r = upload_bricks(ctx)) return r;
        }
        return PDPD_OK;
    };
    uint32_t cur_kernel = c0->force_kernel, cur_shape = c0->brick_shape;
    auto trial = [&](uint32_t g, uint32_t var, uint32_t gs, uint32_t fz) -> int {
        const uint32_t steps = std::max(steps_per_trial, gs);
        for (pdpd_ctx *ctx : cs) {
            if (int r = set_shape(ctx, cur_kernel, cur_shape)) return r;
            ctx->force_grid = g;
            ctx->force_variant = var;
            ctx->graph_steps = gs;
            ctx->fuse_mode = fz;
            ctx->graph_version = ~0ull;
        }
        double per = 0;
        for (int pass = 0; pass < 2; pass++) {   // warm-up (graph capture), then timed
            for (size_t i = 0; i < nc; i++) {
                int r;
                if ((r = restore(i))) return r;
                if ((r = load_counters(cs[i], steps, nullptr))) return r;
            }
            for (size_t i = 0; i < nc; i++) {    // every context's restore done before any trial step
                HIP_TRY(hipStreamSynchronize(cs[i]->stream));
            }
            for (size_t i = 0; i < nc; i++) {
                int r;
                if (pass) HIP_TRY(hipEventRecord(sv[i].e0, cs[i]->stream));
                if ((r = enqueue_steps(cs[i], steps))) return r;
                if (pass) HIP_TRY(hipEventRecord(sv[i].e1, cs[i]->stream));
            }
            for (size_t i = 0; i < nc; i++) HIP_TRY(hipStreamSynchronize(cs[i]->stream));
            if (pass) {
                float end = 0;
                for (size_t i = 0; i < nc; i++) {
                    float ms = 0;
                    HIP_TRY(hipEventElapsedTime(&ms, sv[0].e0, sv[i].e1));
                    end = std::max(end, ms);
                }
                per = end / steps;
            }
        }
        if (per < best_ms) {
            best_ms = per; best_grid = g; best_var = var; best_graph = gs; best_fuse = fz;
            best_kernel = cur_kernel; best_shape = cur_shape;
        }
        return PDPD_OK;
    };
    // pair-force kernels: the per-slot kernel over grid x variant, the brick
    // and slab kernels over their shapes (one workgroup per brick), at 20
    // steps per graph; then the graph length with chained steps off and on
    bool spatial = false;
    for (pdpd_ctx *ctx : cs) spatial |= ctx->spatial;
    int rc;
    cur_kernel = 0;
    for (uint32_t g : grids)
        for (uint32_t var = 0; var < kForceVariants; var++)
            if ((rc = trial(g, var, 20u, fuse0))) return rc;
    if (!spatial) {
        cur_kernel = 1;
        for (uint32_t sh = 0; sh < brick_shape_count(); sh++) {
            cur_shape = sh;
            if ((rc = trial(0u, c0->force_variant, 20u, fuse0))) return rc;
        }
    }
    cur_kernel = best_kernel;
    cur_shape = best_shape;
    const uint32_t g_best = best_grid, v_best = best_var;
    for (uint32_t gs : graphs)
        for (uint32_t fz = 0; fz < 2; fz++)
            if ((gs != 20u || fz != fuse0) && (rc = trial(g_best, v_best, gs, fz))) return rc;
    for (size_t i = 0; i < nc; i++)
        if ((rc = restore(i))) return rc;
    for (pdpd_ctx *ctx : cs) HIP_TRY(hipStreamSynchronize(ctx->stream));
    cleanup();
    for (pdpd_ctx *ctx : cs) {
        ctx->steps_on_layout = 0;      // the cell lists are the last trial's, not the restored state's
        if ((rc = set_shape(ctx, best_kernel, best_shape))) return rc;
        ctx->force_grid = best_grid;
        ctx->force_variant = best_var;
        ctx->graph_steps = best_graph;
        ctx->fuse_mode = best_fuse;
        ctx->graph_version = ~0ull;
        if ((rc = check_flags(ctx))) return rc;
    }
    // empty contexts of the group take the shape too
    for (pdpd_ctx *ctx : all) {
        if (ctx->n_worlds) continue;
        ctx->force_kernel = best_kernel;
        if (ctx->brick_shape != best_shape) {
            ctx->brick_shape = best_shape;
            if (!ctx->dirty && (rc = upload_bricks(ctx))) return rc;
        }
        ctx->force_grid = best_grid;
        ctx->force_variant = best_var;
        ctx->graph_steps = best_graph;
        ctx->fuse_mode = best_fuse;
        ctx->graph_version = ~0ull;
    }
    if (choice && choice_len) {
        const std::string var = best_var == kVariantAuto ? std::string("auto") : std::to_string(best_var);
        snprintf(choice, choice_len,
                 "force_kernel=%u force_brick=%u force_grid=%u force_variant=%s graph_steps=%u fused_integrate=%u "
                 "ms_per_step=%.5f contexts=%zu",
                 best_kernel, best_shape, best_grid, var.c_str(), best_graph, best_fuse, best_ms, all.size());
    }
    return PDPD_OK;
}

int pdpd_autotune(pdpd_ctx *ctx, uint32_t steps_per_trial, char *choice, size_t choice_len)
{
    if (!ctx) return fail(PDPD_ERR_INVALID, "null context");
    return autotune_group({ctx}, steps_per_trial, choice, choice_len);
}

int pdpd_autotune_group(pdpd_ctx **ctxs, uint32_t n, uint32_t steps_per_trial, char *choice, size_t choice_len)
{
    if (!ctxs || n == 0) return fail(PDPD_ERR_INVALID, "no contexts");
    return autotune_group(std::vector<pdpd_ctx *>(ctxs, ctxs + n), steps_per_trial, choice, choice_len);
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
