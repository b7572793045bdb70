// hip_dpd_engine.hpp -- DPDEngine subclasses that run on an MI355X through the
// C ABI of libpdpd_hip.so (include/pdpd_hip.h).
//
//   HipDPDEngine        plain engine: one WorldState, the reference's
//                       CanSupport / Attach / Run(n) / Run(intervals) surface
//                       (include/dpd/core/dpd_engine.hpp:57-215 in poets-dpd)
//   HipBatchDPDEngine   many WorldStates of uneven size advanced together: one
//                       launch set per step in each of a few contexts whose
//                       streams overlap (the batched pattern; no counterpart in
//                       the reference)
//
// With a force logger installed (ForceLogging::set_logger), HipDPDEngine runs
// one step at a time and logs what the reference's AVX2 half-merge engine logs
// (dpd_engine_avx2_half_merge_tbb.hpp:675-760,1244-1259;
// dpd_maths_core_simd.hpp:232-275): per bead b_hash, x, v, f before the step
// and v_next, f_next after it; per evaluated pair dx, dr, dv, the dissipative
// strength, 1/sqrt(dt) scale, gammap, the random draw, the conservative,
// dissipative and random terms, the random scale and the pair force; per
// Hookean bond its dx, dr and force, both ways.
//
// The header compiles against this repository's restatement of the engine
// interface (pdpd/dpd_engine.hpp) or, with PDPD_HOST_REFERENCE defined and the
// reference's include/ on the path, against poets-dpd's own
// dpd/core/dpd_engine.hpp: it touches only the WorldState fields both share.
#pragma once

#if defined(PDPD_HOST_REFERENCE)
#include "dpd/core/dpd_engine.hpp"
#include "dpd/core/logging.hpp"
#else
#include "pdpd/dpd_engine.hpp"
#include "pdpd/force_logging.hpp"
#endif
#include "pdpd_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <unordered_map>
#include <vector>

namespace pdpd_host {

inline void check(int rc, const char *what)
{
    if (rc != PDPD_OK) throw std::runtime_error(std::string(what) + ": " + pdpd_last_error());
}

// WorldState -> pdpd_world_desc. Owns the flattened arrays the descriptor
// points into; valid while this object lives. DPD strengths that are
// expressions of the cell position are evaluated at every cell with the
// bindings of naive_dpd_engine.hpp:158-183 (x, y, z the cell coordinates, ux,
// uy, uz = (coordinate + 0.5) / cells along the axis), unless cell_tables is
// false (the support check, which needs only the constant part).
struct WorldDescBuffers {
    pdpd_world_desc desc{};
    std::vector<double> cons, diss, x, v, f, bond_kappa, bond_r0, angle_kappa, angle_theta0;
    std::vector<double> cell_cons, cell_diss;
    std::vector<uint8_t> stationary, bead_type, is_monomer, polymer_offset;
    std::vector<uint32_t> polymer_id, pt_bond_start, bond_head, bond_tail, pt_angle_start, angle_head_bond,
        angle_tail_bond, polymer_type, polymer_bead_start, polymer_bead_ids;

    explicit WorldDescBuffers(const WorldState &s, bool cell_tables = true)
    {
        const size_t nt = s.bead_types.size();
        bool spatial = false;
        for (size_t i = 0; i < s.interactions.size(); i++) {
            const auto &ii = s.interactions[i];
            spatial |= !ii.conservative.is_constant() || !ii.dissipative.is_constant();
            cons.push_back(ii.conservative.is_constant() ? (double)ii.conservative : 0.0);
            diss.push_back(ii.dissipative.is_constant() ? (double)ii.dissipative : 0.0);
        }
        bool box_ok = true;
        for (int k = 0; k < 3; k++) box_ok &= s.box[k] >= 1 && s.box[k] <= 4096 && s.box[k] == (double)(int)s.box[k];
        if (spatial && cell_tables && box_ok && s.interactions.size() == nt * nt) {
            const int L[3] = {(int)s.box[0], (int)s.box[1], (int)s.box[2]};
            const size_t m = nt * nt;
            cell_cons.resize((size_t)L[0] * L[1] * L[2] * m);
            cell_diss.resize(cell_cons.size());
            std::unordered_map<std::string, double> b;
            size_t c = 0;
            for (int cx = 0; cx < L[0]; cx++)
                for (int cy = 0; cy < L[1]; cy++)
                    for (int cz = 0; cz < L[2]; cz++, c++) {
                        b["x"] = cx; b["y"] = cy; b["z"] = cz;
                        b["ux"] = (cx + 0.5) / L[0];
                        b["uy"] = (cy + 0.5) / L[1];
                        b["uz"] = (cz + 0.5) / L[2];
                        for (size_t k = 0; k < m; k++) {
                            auto pc = s.interactions[k].conservative;     // (evaluate is not const in every Parameter)
                            auto pd = s.interactions[k].dissipative;
                            cell_cons[c * m + k] = pc.evaluate(b);
                            cell_diss[c * m + k] = pd.evaluate(b);
                        }
                    }
        }
        bool any_stationary = false;
        for (const auto &bt : s.bead_types) {
            stationary.push_back(bt.stationary ? 1 : 0);
            any_stationary |= bt.stationary;
        }
        const size_t n = s.beads.size();
        x.resize(3 * n); v.resize(3 * n); f.resize(3 * n);
        bead_type.resize(n); is_monomer.resize(n); polymer_offset.resize(n); polymer_id.resize(n);
        for (size_t i = 0; i < n; i++) {
            const auto &b = s.beads[i];
            for (int d = 0; d < 3; d++) { x[3 * i + d] = b.x[d]; v[3 * i + d] = b.v[d]; f[3 * i + d] = b.f[d]; }
            bead_type[i] = b.bead_type;
            is_monomer[i] = b.is_monomer ? 1 : 0;
            polymer_offset[i] = b.polymer_offset;
            polymer_id[i] = b.polymer_id;
        }
        pt_bond_start.push_back(0);
        pt_angle_start.push_back(0);
        for (const auto &pt : s.polymer_types) {
            for (const auto &b : pt.bonds) {
                bond_head.push_back(b.bead_offset_head);
                bond_tail.push_back(b.bead_offset_tail);
                bond_kappa.push_back((double)b.kappa);
                bond_r0.push_back((double)b.r0);
            }
            pt_bond_start.push_back((uint32_t)bond_head.size());
            for (const auto &bp : pt.bond_pairs) {
                angle_head_bond.push_back(bp.bond_offset_head);
                angle_tail_bond.push_back(bp.bond_offset_tail);
                angle_kappa.push_back((double)bp.kappa);
                angle_theta0.push_back((double)bp.theta0);
            }
            pt_angle_start.push_back((uint32_t)angle_head_bond.size());
        }
        polymer_bead_start.push_back(0);
        for (const auto &p : s.polymers) {
            polymer_type.push_back(p.polymer_type);
            for (auto id : p.bead_ids) polymer_bead_ids.push_back(id);
            polymer_bead_start.push_back((uint32_t)polymer_bead_ids.size());
        }
        pdpd_world_desc &d = desc;
        for (int k = 0; k < 3; k++) d.box[k] = s.box[k];
        d.dt = s.dt;
        d.lambda = s.lambda;
        d.seed = s.seed;
        d.t = (int64_t)s.t;
        d.n_bead_types = (uint32_t)nt;
        d.conservative = cons.data();
        d.dissipative = diss.data();
        d.stationary = any_stationary ? stationary.data() : nullptr;
        d.n_beads = (uint32_t)n;
        d.x = x.data(); d.v = v.data(); d.f = f.data();
        d.bead_type = bead_type.data();
        d.is_monomer = is_monomer.data();
        d.polymer_id = polymer_id.data();
        d.polymer_offset = polymer_offset.data();
        d.n_polymer_types = (uint32_t)s.polymer_types.size();
        d.pt_bond_start = pt_bond_start.data();
        d.bond_head = bond_head.data(); d.bond_tail = bond_tail.data();
        d.bond_kappa = bond_kappa.data(); d.bond_r0 = bond_r0.data();
        d.pt_angle_start = pt_angle_start.data();
        d.angle_head_bond = angle_head_bond.data(); d.angle_tail_bond = angle_tail_bond.data();
        d.angle_kappa = angle_kappa.data(); d.angle_theta0 = angle_theta0.data();
        d.n_polymers = (uint32_t)s.polymers.size();
        d.polymer_type = polymer_type.data();
        d.polymer_bead_start = polymer_bead_start.data();
        d.polymer_bead_ids = polymer_bead_ids.data();
        d.cell_conservative = cell_cons.empty() ? nullptr : cell_cons.data();
        d.cell_dissipative = cell_diss.empty() ? nullptr : cell_diss.data();
    }
};

// Host copy of what the engine last wrote into a WorldState, to notice edits
// made between runs (a caller may change velocities without re-attaching).
struct Snapshot {
    std::vector<double> xvf;
    long t = 0;
    void take(const WorldState &s)
    {
        xvf.resize(9 * s.beads.size());
        for (size_t i = 0; i < s.beads.size(); i++)
            for (int d = 0; d < 3; d++) {
                xvf[9 * i + d] = s.beads[i].x[d];
                xvf[9 * i + 3 + d] = s.beads[i].v[d];
                xvf[9 * i + 6 + d] = s.beads[i].f[d];
            }
        t = s.t;
    }
    bool same(const WorldState &s) const
    {
        if (xvf.size() != 9 * s.beads.size() || t != s.t) return false;
        for (size_t i = 0; i < s.beads.size(); i++)
            for (int d = 0; d < 3; d++)
                if (xvf[9 * i + d] != s.beads[i].x[d] || xvf[9 * i + 3 + d] != s.beads[i].v[d] ||
                    xvf[9 * i + 6 + d] != s.beads[i].f[d])
                    return false;
        return true;
    }
};

// Device context shared by the engines (one per engine instance).
class Context {
public:
    explicit Context(int device) : m_device(device) {}
    ~Context() { if (m_ctx) pdpd_ctx_destroy(m_ctx); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;

    pdpd_ctx *get()
    {
        if (!m_ctx) check(pdpd_ctx_create(m_device, &m_ctx), "pdpd_ctx_create");
        return m_ctx;
    }
    bool created() const { return m_ctx != nullptr; }

private:
    int m_device;
    pdpd_ctx *m_ctx = nullptr;
};

// The world's state (after all enqueued work), or with from_snapshot its last
// pdpd_snapshot (later steps may still run), into s.
inline void download_into(pdpd_ctx *ctx, int slot, WorldState &s, bool from_snapshot = false)
{
    const size_t n = s.beads.size();
    std::vector<double> x(3 * n), v(3 * n), f(3 * n);
    int64_t t = 0;
    if (from_snapshot)
        check(pdpd_world_snapshot_download(ctx, slot, x.data(), v.data(), f.data(), &t), "pdpd_world_snapshot_download");
    else
        check(pdpd_world_download(ctx, slot, x.data(), v.data(), f.data(), &t), "pdpd_world_download");
    for (size_t i = 0; i < n; i++) {
        auto &b = s.beads[i];
        for (int d = 0; d < 3; d++) { b.x[d] = x[3 * i + d]; b.v[d] = v[3 * i + d]; b.f[d] = f[3 * i + d]; }
    }
    s.t = (long)t;
}

}  // namespace pdpd_host

class HipDPDEngine : public DPDEngine {
protected:
    bool CanSupportHookeanBonds() const override { return true; }
    bool CanSupportAngleBonds() const override { return true; }
    bool CanSupportStationaryBeadTypes() const override { return true; }
    bool CanSupportSpatialDPDParameters() const override { return true; }

public:
    explicit HipDPDEngine(int device = 0, unsigned graph_steps = 20) : m_ctx(device), m_graph_steps(graph_steps) {}

    ~HipDPDEngine() override
    {
        if (m_ctx.created() && m_slot >= 0) pdpd_world_remove(m_ctx.get(), m_slot);
    }

    std::string CanSupport(const WorldState *s) const override
    {
        std::string r = DPDEngine::CanSupport(s);
        if (!r.empty()) return r;
        pdpd_host::WorldDescBuffers b(*s, false);
        return pdpd_can_support(&b.desc);
    }

    // Bonds use the nearest image, which resolves a bond only while it is
    // shorter than half the smallest side of the box: just below that for the
    // attached world, and for the smallest box the engine supports (3) when
    // none is attached. A longer bond raises PDPD_FLAG_BOND_STRETCH and the
    // run fails, as the reference's engines throw when a bond snaps
    // (naive_dpd_engine_half_step.hpp:435-437).
    double GetMaxBondLength() const override
    {
        double side = 3.0;
        if (m_state) side = std::min(m_state->box[0], std::min(m_state->box[1], m_state->box[2]));
        return std::nextafter(0.5 * side, 0.0);
    }

    void PrepareResources() override { m_ctx.get(); }

    bool GetTimings(timings_t &t) override
    {
        if (!m_ctx.created()) return false;
        pdpd_stats st;
        pdpd_host::check(pdpd_get_stats(m_ctx.get(), &st), "pdpd_get_stats");
        t = timings_t{};
        t.configure = st.upload_s;
        t.execute_to_first_bead = st.run_s;
        t.execute_to_last_bead = st.run_s;
        return true;
    }

    void GetConfigEntries(std::vector<engine_config_entry_t> &config) override
    {
        config.clear();
        std::vector<char> buf(4096);
        pdpd_host::check(pdpd_list_config(ctx(), buf.data(), buf.size()), "pdpd_list_config");
        std::string all(buf.data());
        size_t pos = 0;
        while (pos < all.size()) {
            size_t eol = all.find('\n', pos);
            std::string line = all.substr(pos, eol == std::string::npos ? std::string::npos : eol - pos);
            pos = eol == std::string::npos ? all.size() : eol + 1;
            size_t eq = line.find('='), tab = line.find('\t');
            if (eq == std::string::npos || tab == std::string::npos) continue;
            config.push_back({line.substr(0, eq), line.substr(tab + 1), line.substr(eq + 1, tab - eq - 1)});
        }
    }

    void SetConfigEntry(const std::string &key, const std::string &value) override
    {
        if (pdpd_set_config(ctx(), key.c_str(), value.c_str()) != PDPD_OK)
            throw std::runtime_error(pdpd_last_error());
    }

    void Attach(WorldState *state) override
    {
        if (m_slot >= 0) {
            pdpd_host::check(pdpd_world_remove(ctx(), m_slot), "pdpd_world_remove");
            m_slot = -1;
        }
        m_state = state;
        if (!state) return;
        std::string why = CanSupport(state);
        if (!why.empty()) throw std::runtime_error("Can't support this world : " + why);
        import_state();
    }

    void Run(unsigned nSteps) override
    {
        require_attached();
        if (!m_snap.same(*m_state)) reimport();
        if (ForceLogging::logger()) {
            for (unsigned i = 0; i < nSteps; i++) logged_step();
            return;
        }
        pdpd_host::check(pdpd_run(ctx(), nSteps), "pdpd_run");
        export_state();
    }

    // Interval run with the next interval already on the device while the
    // previous one's state comes to the host and its callback runs (the
    // contract of dpd_engine.hpp:187-194 allows execution to continue during
    // the callback): a device-side snapshot is enqueued after each interval and
    // the next interval behind it, so the copy to the host and the callback
    // overlap the next interval's steps.
    unsigned Run(int interval_count, unsigned interval_size, std::function<bool()> interval_callback) override
    {
        require_attached();
        if (interval_count == 0) return 0;
        if (!m_snap.same(*m_state)) reimport();
        unsigned done = 0;
        if (ForceLogging::logger()) {     // logged: one step at a time, no interval in flight
            for (;;) {
                for (unsigned i = 0; i < interval_size; i++) logged_step();
                done += interval_size;
                const bool cont = interval_callback();
                if (interval_count > 0) --interval_count;
                if (!cont || interval_count == 0) return done;
                if (!m_snap.same(*m_state)) reimport();
            }
        }
        pdpd_host::check(pdpd_run(ctx(), interval_size), "pdpd_run");
        for (;;) {
            pdpd_host::check(pdpd_snapshot(ctx()), "pdpd_snapshot");
            done += interval_size;
            const bool more = interval_count < 0 || interval_count > 1;
            if (more) pdpd_host::check(pdpd_run(ctx(), interval_size), "pdpd_run");
            export_state(true);
            const bool cont = interval_callback();
            if (interval_count > 0) --interval_count;
            if (more && !m_snap.same(*m_state)) {
                // the callback edited the world: drop the speculative interval
                pdpd_host::check(pdpd_sync(ctx()), "pdpd_sync");
                reimport();
                if (cont) pdpd_host::check(pdpd_run(ctx(), interval_size), "pdpd_run");
            }
            if (!cont || !more) {
                if (more && cont == false) m_device_ahead = true;
                break;
            }
        }
        if (m_device_ahead) {
            pdpd_host::check(pdpd_sync(ctx()), "pdpd_sync");
            reimport();
            m_device_ahead = false;
        }
        return done;
    }

protected:
    pdpd_ctx *ctx() { return m_ctx.get(); }

    void require_attached() const
    {
        if (!m_state || m_slot < 0) throw std::runtime_error("HipDPDEngine: no world attached");
    }

    void import_state()
    {
        pdpd_host::WorldDescBuffers b(*m_state);
        pdpd_ctx *c = ctx();
        pdpd_host::check(pdpd_set_config(c, "graph_steps", std::to_string(m_graph_steps).c_str()), "pdpd_set_config");
        pdpd_host::check(pdpd_world_add(c, &b.desc, &m_slot), "pdpd_world_add");
        m_snap.take(*m_state);
    }

    void reimport()
    {
        if (m_slot >= 0) pdpd_host::check(pdpd_world_remove(ctx(), m_slot), "pdpd_world_remove");
        m_slot = -1;
        import_state();
    }

    void export_state(bool from_snapshot = false)
    {
        pdpd_host::download_into(ctx(), m_slot, *m_state, from_snapshot);
        m_snap.take(*m_state);
    }

    // One step with the force logger: the state before, the step's pair and
    // bond evaluations (pdpd_trace_forces), the state after.
    void logged_step()
    {
        ForceLogging *log = ForceLogging::logger();
        log->SetTime(m_state->t);
        for (const auto &b : m_state->beads) {
            const BeadHash h = b.get_hash_code();
            const double hv = (double)h.hash;
            const double x[3] = {b.x[0], b.x[1], b.x[2]}, v[3] = {b.v[0], b.v[1], b.v[2]},
                         f[3] = {b.f[0], b.f[1], b.f[2]};
            log->LogBeadProperty(h, "b_hash", 1, &hv);
            log->LogBeadProperty(h, "x", 3, x);
            log->LogBeadProperty(h, "v", 3, v);
            log->LogBeadProperty(h, "f", 3, f);
        }
        pdpd_host::check(pdpd_run(ctx(), 1), "pdpd_run");
        uint32_t np = 0, nb = 0;
        pdpd_host::check(pdpd_trace_forces(ctx(), m_slot, nullptr, 0, &np, nullptr, 0, &nb), "pdpd_trace_forces");
        std::vector<pdpd_pair_record> pairs(np);
        std::vector<pdpd_bond_record> bonds(nb);
        pdpd_host::check(pdpd_trace_forces(ctx(), m_slot, pairs.data(), np, &np, bonds.data(), nb, &nb),
                         "pdpd_trace_forces");
        for (const pdpd_pair_record &r : pairs) {
            const BeadHash h{r.home}, o{r.other};
            const double dx[3] = {r.dx[0], r.dx[1], r.dx[2]}, dv[3] = {r.dv[0], r.dv[1], r.dv[2]},
                         ff[3] = {r.f[0], r.f[1], r.f[2]};
            const double one[8] = {r.dr, r.diss_strength, r.inv_root_dt, r.gammap, r.rng, r.con, r.diss, r.rng_scale};
            const double rnd = r.rand;
            log->LogBeadPairProperty(h, o, "dx", 3, dx);
            log->LogBeadPairProperty(h, o, "dr", 1, &one[0]);
            log->LogBeadPairProperty(h, o, "dv", 3, dv);
            log->LogBeadPairProperty(h, o, "dpd-diss-strength", 1, &one[1]);
            log->LogBeadPairProperty(h, o, "dpd-invrootdt", 1, &one[2]);
            log->LogBeadPairProperty(h, o, "dpd-gammap", 1, &one[3]);
            log->LogBeadPairProperty(h, o, "dpd-rng", 1, &one[4]);
            log->LogBeadPairProperty(h, o, "dpd-con", 1, &one[5]);
            log->LogBeadPairProperty(h, o, "dpd-diss", 1, &one[6]);
            log->LogBeadPairProperty(h, o, "dpd-rng-scale", 1, &one[7]);
            log->LogBeadPairProperty(h, o, "dpd-rand", 1, &rnd);
            log->LogBeadPairProperty(h, o, "f_next_dpd", 3, ff);
        }
        for (const pdpd_bond_record &r : bonds) {
            const BeadHash h{r.head}, t{r.tail};
            double dx[3] = {r.dx[0], r.dx[1], r.dx[2]}, ff[3] = {r.f_head[0], r.f_head[1], r.f_head[2]};
            log->LogBeadPairProperty(h, t, "f_next_hookean_dx", 3, dx);
            log->LogBeadPairProperty(h, t, "f_next_hookean_dr", 1, &r.dr);
            log->LogBeadPairProperty(h, t, "f_next_hookean", 3, ff);
            for (int d = 0; d < 3; d++) { dx[d] = -dx[d]; ff[d] = -ff[d]; }
            log->LogBeadPairProperty(t, h, "f_next_hookean", 3, ff);
            log->LogBeadPairProperty(t, h, "f_next_hookean_dx", 3, dx);
            log->LogBeadPairProperty(t, h, "f_next_hookean_dr", 1, &r.dr);
        }
        export_state();
        for (const auto &b : m_state->beads) {
            const double v[3] = {b.v[0], b.v[1], b.v[2]}, f[3] = {b.f[0], b.f[1], b.f[2]};
            log->LogBeadProperty(b.get_hash_code(), "v_next", 3, v);
            log->LogBeadProperty(b.get_hash_code(), "f_next", 3, f);
        }
        log->Flush();
    }

    pdpd_host::Context m_ctx;
    unsigned m_graph_steps;
    WorldState *m_state = nullptr;
    int m_slot = -1;
    pdpd_host::Snapshot m_snap;
    bool m_device_ahead = false;
};

// Many independent worlds advanced together: every step is one launch set over
// all of them. Worlds may be attached and detached between runs; each keeps its
// own t, seed and parameters, and its results equal a solo run bit for bit.
// Many worlds advanced together. The worlds are spread over `parts` device
// contexts of one GPU, each with its own stream (a new world goes to the part
// holding the fewest beads); Enqueue launches every part's steps, so the parts'
// launches overlap on the device. Each world's results are those of the world
// run alone, bit for bit, whatever part it lands in. SuggestedParts gives the
// part count measured best for an ensemble's total bead count.
class HipBatchDPDEngine {
public:
    explicit HipBatchDPDEngine(int device = 0, unsigned parts = 1)
    {
        for (unsigned i = 0; i < std::max(1u, parts); i++) m_parts.emplace_back(new pdpd_host::Context(device));
        m_load.assign(m_parts.size(), 0);
    }

    // 256 mixed worlds, 6.0M beads: 1 part 3.62e9 bead-steps/s, 4: 3.81e9,
    // 8: 3.94e9, 12: 4.10e9, 16: 3.58e9.
    static unsigned SuggestedParts(uint64_t total_beads)
    {
        return total_beads >= 4000000u ? 12u : (total_beads >= 1000000u ? 4u : 1u);
    }

    std::string CanSupport(const WorldState *s) const
    {
        HipDPDEngine probe;
        return probe.CanSupport(s);
    }

    // Attach one more world; returns its slot.
    int Attach(WorldState *s)
    {
        std::string why = CanSupport(s);
        if (!why.empty()) throw std::runtime_error("Can't support this world : " + why);
        pdpd_host::WorldDescBuffers b(*s);
        const size_t part = (size_t)(std::min_element(m_load.begin(), m_load.end()) - m_load.begin());
        int cslot = -1;
        pdpd_host::check(pdpd_world_add(m_parts[part]->get(), &b.desc, &cslot), "pdpd_world_add");
        m_load[part] += s->beads.size();
        m_worlds.push_back({m_next_slot, part, cslot, s});
        return m_next_slot++;
    }

    void Detach(int slot)
    {
        for (size_t i = 0; i < m_worlds.size(); i++)
            if (m_worlds[i].slot == slot) {
                const Entry &e = m_worlds[i];
                pdpd_host::check(pdpd_world_remove(m_parts[e.part]->get(), e.cslot), "pdpd_world_remove");
                m_load[e.part] -= e.state->beads.size();
                m_worlds.erase(m_worlds.begin() + i);
                return;
            }
        throw std::runtime_error("HipBatchDPDEngine: unknown slot");
    }

    // Step budget of one world for the next Run/Enqueue (default: all steps).
    void SetSteps(int slot, unsigned steps)
    {
        const Entry &e = find(slot);
        pdpd_host::check(pdpd_world_set_steps(m_parts[e.part]->get(), e.cslot, steps), "set_steps");
    }

    // Advance every attached world and synchronise all WorldStates.
    void Run(unsigned nSteps)
    {
        Enqueue(nSteps);
        for (auto &w : m_worlds) pdpd_host::download_into(m_parts[w.part]->get(), w.cslot, *w.state);
    }

    // Asynchronous advance of every part; Sync(slot) brings one world's WorldState up to date.
    void Enqueue(unsigned nSteps)
    {
        for (size_t p = 0; p < m_parts.size(); p++)
            if (has_worlds(p)) pdpd_host::check(pdpd_run(m_parts[p]->get(), nSteps), "pdpd_run");
    }
    void Sync(int slot)
    {
        const Entry &e = find(slot);
        pdpd_host::download_into(m_parts[e.part]->get(), e.cslot, *e.state);
    }

    // Let the autotuner pick the launch shape (pair-force kernel, grid,
    // variant or brick shape, steps per graph, chained steps), timed with all
    // parts running side by side; every part takes it.
    std::string Autotune(unsigned steps_per_trial)
    {
        std::vector<pdpd_ctx *> cs;
        for (auto &c : m_parts) cs.push_back(c->get());
        char buf[256] = {0};
        pdpd_host::check(pdpd_autotune_group(cs.data(), (uint32_t)cs.size(), steps_per_trial, buf, sizeof(buf)),
                         "pdpd_autotune_group");
        return buf;
    }

    void SetConfigEntry(const std::string &key, const std::string &value)
    {
        for (auto &c : m_parts)
            if (pdpd_set_config(c->get(), key.c_str(), value.c_str()) != PDPD_OK) throw std::runtime_error(pdpd_last_error());
    }

    // Statistics of one part (part 0 by default).
    pdpd_stats Stats(unsigned part = 0)
    {
        pdpd_stats st;
        pdpd_host::check(pdpd_get_stats(m_parts.at(part)->get(), &st), "pdpd_get_stats");
        return st;
    }

    size_t size() const { return m_worlds.size(); }
    size_t parts() const { return m_parts.size(); }

private:
    struct Entry { int slot; size_t part; int cslot; WorldState *state; };
    const Entry &find(int slot) const
    {
        for (const auto &w : m_worlds)
            if (w.slot == slot) return w;
        throw std::runtime_error("HipBatchDPDEngine: unknown slot");
    }
    bool has_worlds(size_t p) const
    {
        for (const auto &w : m_worlds)
            if (w.part == p) return true;
        return false;
    }
    std::vector<std::unique_ptr<pdpd_host::Context>> m_parts;
    std::vector<uint64_t> m_load;
    std::vector<Entry> m_worlds;
    int m_next_slot = 0;
};
