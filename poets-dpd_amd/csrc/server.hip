// server.hip -- continuous-batching serving loop over one pdpd context.
//
// Requests (a world + a step count + an arrival time) queue in arrival order.
// At every interval boundary the loop retires finished requests, admits queued
// ones that have arrived while the resident set fits the HBM budget, and runs
// the next interval for all resident requests in one launch set per step; a
// request with fewer steps left than the interval stops early through its step
// budget. Results and per-request latencies (completion - arrival) are kept;
// an optional callback sees each request's state after every interval.
// Host code only; no counterpart in the reference (its engines run one world).
#include "pdpd_hip.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

// Sets the text pdpd_last_error() returns (capi.hip).
void pdpd_internal_set_error(const char *msg);

namespace {

int sfail(int code, const std::string &m)
{
    pdpd_internal_set_error(m.c_str());
    return code;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Deep copy of a pdpd_world_desc.
struct OwnedDesc {
    pdpd_world_desc d{};
    std::vector<double> cons, diss, x, v, f, bk, br, ak, at0, ccons, cdiss;
    std::vector<uint8_t> stat, btype, mono, poff;
    std::vector<uint32_t> pid, ptbs, bh, bt, ptas, ahb, atb, ptype, pbs, pbi;

    explicit OwnedDesc(const pdpd_world_desc &s)
    {
        d = s;
        const size_t nt = s.n_bead_types, n = s.n_beads, npt = s.n_polymer_types, np = s.n_polymers;
        auto cp = [](auto &vec, const auto *src, size_t cnt) {
            if (src && cnt) vec.assign(src, src + cnt);
            else vec.clear();
            return vec.empty() ? nullptr : vec.data();
        };
        d.conservative = cp(cons, s.conservative, nt * nt);
        d.dissipative = cp(diss, s.dissipative, nt * nt);
        d.stationary = s.stationary ? cp(stat, s.stationary, nt) : nullptr;
        d.x = cp(x, s.x, 3 * n);
        d.v = cp(v, s.v, 3 * n);
        d.f = cp(f, s.f, 3 * n);
        d.bead_type = cp(btype, s.bead_type, n);
        d.is_monomer = cp(mono, s.is_monomer, n);
        d.polymer_id = cp(pid, s.polymer_id, n);
        d.polymer_offset = cp(poff, s.polymer_offset, n);
        d.pt_bond_start = cp(ptbs, s.pt_bond_start, npt + 1);
        const size_t nb = npt ? s.pt_bond_start[npt] : 0;
        d.bond_head = cp(bh, s.bond_head, nb);
        d.bond_tail = cp(bt, s.bond_tail, nb);
        d.bond_kappa = cp(bk, s.bond_kappa, nb);
        d.bond_r0 = cp(br, s.bond_r0, nb);
        d.pt_angle_start = cp(ptas, s.pt_angle_start, npt + 1);
        const size_t na = npt ? s.pt_angle_start[npt] : 0;
        d.angle_head_bond = cp(ahb, s.angle_head_bond, na);
        d.angle_tail_bond = cp(atb, s.angle_tail_bond, na);
        d.angle_kappa = cp(ak, s.angle_kappa, na);
        d.angle_theta0 = cp(at0, s.angle_theta0, na);
        d.polymer_type = cp(ptype, s.polymer_type, np);
        d.polymer_bead_start = cp(pbs, s.polymer_bead_start, np + 1);
        d.polymer_bead_ids = cp(pbi, s.polymer_bead_ids, np ? s.polymer_bead_start[np] : 0);
        const size_t ncp = (size_t)s.box[0] * (size_t)s.box[1] * (size_t)s.box[2] * nt * nt;
        d.cell_conservative = cp(ccons, s.cell_conservative, ncp);
        d.cell_dissipative = cp(cdiss, s.cell_dissipative, ncp);
    }
};

struct Request {
    uint64_t id;
    std::unique_ptr<OwnedDesc> desc;
    double arrival;         // seconds after the start of the run (scaled)
    uint32_t steps;
    uint32_t part = 0;      // the context the request is resident in
    uint32_t done = 0;
    uint32_t inflight = 0;  // steps enqueued whose results are not yet processed
    int slot = -1;
    uint64_t bytes = 0;
    pdpd_request_result res{};
    std::vector<double> x, v, f;
};

// Device bytes one resident world needs: double-buffered pos/vel/frc (96 B),
// sort keys, permutation, sorted indices and bead->slot map (16 B) per bead, the cell
// counts and starts (8 B) per cell, and per-cell strengths (8 B per cell and
// type pair) if it has them.
uint64_t device_bytes(const pdpd_world_desc &d)
{
    const uint64_t cells = (uint64_t)d.box[0] * (uint64_t)d.box[1] * (uint64_t)d.box[2];
    const uint64_t par = d.cell_conservative ? 8ull * cells * d.n_bead_types * d.n_bead_types : 0;
    return 112ull * d.n_beads + 8ull * cells + par + 4096;
}

}  // namespace

struct pdpd_server {
    pdpd_server_config cfg{};
    std::vector<pdpd_ctx *> ctxs;      // resident requests are spread over these (one stream each)
    std::vector<uint64_t> part_beads;  // resident beads per context
    std::map<uint64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request *> waiting;
    std::vector<Request *> resident;
    uint64_t next_id = 0;
    pdpd_interval_fn cb = nullptr;
    void *cb_user = nullptr;
    pdpd_observables_fn ocb = nullptr;     // per-interval observables, no state download
    void *ocb_user = nullptr;
    pdpd_server_stats stats{};

    ~pdpd_server()
    {
        for (pdpd_ctx *c : ctxs) pdpd_ctx_destroy(c);
    }
};

extern "C" {

int pdpd_server_create(const pdpd_server_config *cfg, pdpd_server **out)
{
    if (!cfg || !out) return sfail(PDPD_ERR_INVALID, "null argument");
    if (cfg->interval_steps == 0) return sfail(PDPD_ERR_INVALID, "interval_steps must be positive");
    std::unique_ptr<pdpd_server> s(new pdpd_server());
    s->cfg = *cfg;
    const uint32_t parts = std::max<uint32_t>(1, cfg->contexts);
    if (parts > 64) return sfail(PDPD_ERR_INVALID, "contexts must be at most 64");
    for (uint32_t i = 0; i < parts; i++) {
        pdpd_ctx *c = nullptr;
        int rc = pdpd_ctx_create(cfg->device, &c);
        if (rc) return sfail(rc, pdpd_last_error());
        s->ctxs.push_back(c);
    }
    s->part_beads.assign(parts, 0);
    if (cfg->hbm_budget_bytes) {
        // set the budget's worth of beads aside in every context (a context may
        // hold up to twice its share), so admissions never allocate
        const uint64_t total = cfg->hbm_budget_bytes / 112ull;
        const uint64_t each = std::min<uint64_t>(std::min<uint64_t>(total, parts > 1 ? 2 * total / parts : total),
                                                 (1ull << 26) - 1);
        for (pdpd_ctx *c : s->ctxs) {
            int rc = pdpd_set_config(c, "reserve_beads", std::to_string(each).c_str());
            if (rc) return sfail(rc, pdpd_last_error());
        }
    }
    *out = s.release();
    return PDPD_OK;
}

int pdpd_server_destroy(pdpd_server *s)
{
    delete s;
    return PDPD_OK;
}

int pdpd_server_set_config(pdpd_server *s, const char *key, const char *value)
{
    if (!s) return sfail(PDPD_ERR_INVALID, "null server");
    for (pdpd_ctx *c : s->ctxs) {
        int rc = pdpd_set_config(c, key, value);
        if (rc) return sfail(rc, pdpd_last_error());
    }
    return PDPD_OK;
}

int pdpd_server_set_callback(pdpd_server *s, pdpd_interval_fn fn, void *user)
{
    if (!s) return sfail(PDPD_ERR_INVALID, "null server");
    s->cb = fn;
    s->cb_user = user;
    return PDPD_OK;
}

int pdpd_server_set_observables_callback(pdpd_server *s, pdpd_observables_fn fn, void *user)
{
    if (!s) return sfail(PDPD_ERR_INVALID, "null server");
    s->ocb = fn;
    s->ocb_user = user;
    return PDPD_OK;
}

int pdpd_server_submit(pdpd_server *s, const pdpd_world_desc *w, double arrival_s, uint32_t steps, uint64_t *id_out)
{
    if (!s || !w) return sfail(PDPD_ERR_INVALID, "null argument");
    const char *why = pdpd_can_support(w);
    if (why[0]) return sfail(PDPD_ERR_UNSUPPORTED, why);
    if (w->n_beads == 0) return sfail(PDPD_ERR_INVALID, "request has no beads");
    std::unique_ptr<Request> r(new Request());
    r->id = s->next_id++;
    r->desc.reset(new OwnedDesc(*w));
    r->arrival = arrival_s;
    r->steps = steps;
    r->bytes = device_bytes(*w);
    if (s->cfg.hbm_budget_bytes && r->bytes > s->cfg.hbm_budget_bytes)
        return sfail(PDPD_ERR_CAPACITY, "request larger than the HBM budget");
    r->res.id = r->id;
    r->res.n_beads = w->n_beads;
    r->res.steps = steps;
    r->res.arrival_s = arrival_s;
    r->res.status = 1;   // queued
    // keep the queue in arrival order (stable for equal times)
    Request *raw = r.get();
    auto it = std::upper_bound(s->waiting.begin(), s->waiting.end(), arrival_s,
                               [](double a, const Request *q) { return a < q->arrival; });
    s->waiting.insert(it, raw);
    s->reqs.emplace(raw->id, std::move(r));
    if (id_out) *id_out = raw->id;
    return PDPD_OK;
}

int pdpd_server_run(pdpd_server *s)
{
    if (!s) return sfail(PDPD_ERR_INVALID, "null server");
    const double scale = s->cfg.time_scale;
    const uint64_t budget = s->cfg.hbm_budget_bytes;
    const uint32_t K = s->cfg.interval_steps;
    const double t0 = now_s();
    uint64_t used = 0;
    for (Request *r : s->resident) used += r->bytes;
    double idle = 0;      // time with no request resident
    auto rel = [&]() { return now_s() - t0; };
    auto fail_rc = [&](int rc) { return sfail(rc, pdpd_last_error()); };

    // One interval is computed while the previous one's results are handled:
    // after an interval is enqueued, the previous interval's snapshot is
    // downloaded (finished and callback requests) and its callbacks run, then
    // a snapshot of the new interval is enqueued behind it. Requests finishing
    // with an interval are frozen (step budget 0) in the next; a request a
    // callback cancels computes one interval more, which is discarded.
    std::vector<std::pair<Request *, uint32_t>> pending;    // the snapshotted interval: request, steps advanced
    bool have_pending = false;
    auto process = [&]() -> int {
        for (auto &pa : pending) {
            Request *r = pa.first;
            const uint32_t adv = pa.second;
            r->done += adv;
            r->inflight -= adv;
            s->stats.bead_steps += (double)r->res.n_beads * adv;
            const bool finished = r->done >= r->steps;
            bool cancel = false;
            uint32_t failed = 0;
            // the request's own world failed: retire it alone; the others do
            // not depend on it and run on
            auto world_failed = [&]() -> bool {
                uint32_t wf = 0;
                if (pdpd_world_flags(s->ctxs[r->part], r->slot, &wf) == PDPD_OK && wf) failed = wf;
                return failed != 0;
            };
            if (s->ocb) {      // the interval's observables, reduced on the device
                pdpd_observables o;
                int rc = pdpd_world_snapshot_observables(s->ctxs[r->part], r->slot, &o);
                if (rc == PDPD_ERR_STATE) {
                    if (!world_failed()) return rc;
                } else if (rc) {
                    return rc;
                } else {
                    cancel = s->ocb(s->ocb_user, r->id, r->done, &o) == 0;
                }
            }
            if ((s->cb || finished) && !failed) {
                const size_t n = r->res.n_beads;
                r->x.resize(3 * n); r->v.resize(3 * n); r->f.resize(3 * n);
                int64_t t = 0;
                int rc = pdpd_world_snapshot_download(s->ctxs[r->part], r->slot, r->x.data(), r->v.data(),
                                                      r->f.data(), &t);
                if (rc == PDPD_ERR_STATE) {
                    if (!world_failed()) return rc;
                } else if (rc) {
                    return rc;
                }
                if (s->cb && !failed)
                    cancel |= s->cb(s->cb_user, r->id, r->done, t, (uint32_t)n, r->x.data(), r->v.data(),
                                    r->f.data()) == 0;
            }
            if (failed) {
                r->x.clear(); r->v.clear(); r->f.clear();
                r->res.fail_flags = failed;
                s->stats.failed++;
            }
            if (finished || cancel || failed) {
                int rc = pdpd_world_remove(s->ctxs[r->part], r->slot);
                if (rc) return rc;
                s->part_beads[r->part] -= r->res.n_beads;
                r->slot = -1;
                r->inflight = 0;
                r->res.done_s = rel();
                r->res.latency_s = r->res.done_s - r->arrival * scale;
                r->res.status = failed ? 4 : (finished ? 0 : 3);   // failed / done / cancelled
                used -= r->bytes;
                if (!failed) s->stats.completed++;
            }
        }
        std::vector<Request *> keep;
        for (Request *r : s->resident)
            if (r->slot >= 0) keep.push_back(r);
        s->resident.swap(keep);
        pending.clear();
        return PDPD_OK;
    };

    while (!s->waiting.empty() || !s->resident.empty()) {
        // admit in arrival order while the head has arrived and fits
        const double now = rel();
        while (!s->waiting.empty()) {
            Request *r = s->waiting.front();
            if (r->arrival * scale > now) break;
            if (budget && used + r->bytes > budget && !s->resident.empty()) break;
            if (s->cfg.max_resident && s->resident.size() >= s->cfg.max_resident) break;
            const uint32_t part = (uint32_t)(std::min_element(s->part_beads.begin(), s->part_beads.end()) -
                                             s->part_beads.begin());
            int rc = pdpd_world_add(s->ctxs[part], &r->desc->d, &r->slot);
            if (rc) return fail_rc(rc);
            r->part = part;
            s->part_beads[part] += r->res.n_beads;
            r->res.admit_s = now;
            r->res.status = 2;  // resident
            used += r->bytes;
            s->resident.push_back(r);
            s->waiting.pop_front();
        }
        if (s->resident.empty()) {
            const double next = s->waiting.front()->arrival * scale;
            const double wait = next - rel();
            if (wait > 0) {
                std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.05)));
                idle += std::min(wait, 0.05);
            }
            continue;
        }
        s->stats.max_resident = std::max<uint64_t>(s->stats.max_resident, s->resident.size());
        {
            uint64_t rb = 0;
            for (uint64_t pb : s->part_beads) rb += pb;
            s->stats.max_resident_beads = std::max(s->stats.max_resident_beads, rb);
        }
        uint32_t k = 0;
        for (Request *r : s->resident) k = std::max(k, std::min(K, r->steps - r->done - r->inflight));
        std::vector<std::pair<Request *, uint32_t>> next;
        if (k) {
            for (Request *r : s->resident) {
                const uint32_t adv = std::min(k, r->steps - r->done - r->inflight);
                int rc = pdpd_world_set_steps(s->ctxs[r->part], r->slot, adv);
                if (rc) return fail_rc(rc);
                if (adv) next.push_back({r, adv});
            }
            for (uint32_t p = 0; p < s->ctxs.size(); p++) {
                if (!s->part_beads[p]) continue;
                int rc = pdpd_run(s->ctxs[p], k);
                if (rc) return fail_rc(rc);
            }
            for (auto &pa : next) pa.first->inflight += pa.second;
            s->stats.intervals++;
        }
        if (have_pending) {
            int rc = process();
            if (rc) return fail_rc(rc);
            have_pending = false;
        }
        if (k) {
            for (auto &pa : next)
                if (pa.first->slot >= 0) pending.push_back(pa);
            // cancelled requests left the context; their interval is discarded
            for (uint32_t p = 0; p < s->ctxs.size(); p++) {
                if (!s->part_beads[p]) continue;
                int rc = s->ocb ? pdpd_snapshot_observables(s->ctxs[p]) : pdpd_snapshot(s->ctxs[p]);
                if (rc) return fail_rc(rc);
            }
            have_pending = !pending.empty();
        }
    }
    for (pdpd_ctx *c : s->ctxs) {
        int rc = pdpd_sync(c);
        if (rc) return fail_rc(rc);
    }
    s->stats.makespan_s = rel();
    {
        double sum = 0;
        s->stats.relayouts = 0;
        s->stats.relayout_max_s = 0;
        for (pdpd_ctx *c : s->ctxs) {
            pdpd_stats cs;
            if (pdpd_get_stats(c, &cs) != PDPD_OK) continue;
            s->stats.relayouts += cs.relayouts;
            sum += cs.relayout_s;
            s->stats.relayout_max_s = std::max(s->stats.relayout_max_s, cs.relayout_max_s);
        }
        s->stats.relayout_mean_s = s->stats.relayouts ? sum / (double)s->stats.relayouts : 0.0;
    }
    s->stats.gpu_busy_s += s->stats.makespan_s - idle;
    s->stats.requests = s->reqs.size();
    std::vector<double> lat;
    for (auto &kv : s->reqs)
        if (kv.second->res.status == 0) lat.push_back(kv.second->res.latency_s);

    std::sort(lat.begin(), lat.end());
    auto pct = [&](double p) {
        if (lat.empty()) return 0.0;
        const double pos = p * (lat.size() - 1);
        const size_t lo = (size_t)std::floor(pos), hi = std::min(lat.size() - 1, lo + 1);
        return lat[lo] + (lat[hi] - lat[lo]) * (pos - lo);
    };
    s->stats.p50_latency_s = pct(0.50);
    s->stats.p99_latency_s = pct(0.99);
    double sum = 0;
    for (double l : lat) sum += l;
    s->stats.mean_latency_s = lat.empty() ? 0 : sum / lat.size();
    s->stats.max_latency_s = lat.empty() ? 0 : lat.back();
    s->stats.bead_steps_per_s = s->stats.makespan_s > 0 ? s->stats.bead_steps / s->stats.makespan_s : 0;
    return PDPD_OK;
}

int pdpd_server_result(pdpd_server *s, uint64_t id, pdpd_request_result *out, double *x, double *v, double *f)
{
    if (!s) return sfail(PDPD_ERR_INVALID, "null server");
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return sfail(PDPD_ERR_INVALID, "unknown request");
    Request *r = it->second.get();
    if (out) *out = r->res;
    const size_t n = r->res.n_beads;
    if (r->res.status == 0 || r->res.status == 3) {
        if (x && !r->x.empty()) memcpy(x, r->x.data(), 3 * n * sizeof(double));
        if (v && !r->v.empty()) memcpy(v, r->v.data(), 3 * n * sizeof(double));
        if (f && !r->f.empty()) memcpy(f, r->f.data(), 3 * n * sizeof(double));
    }
    return PDPD_OK;
}

int pdpd_server_get_stats(pdpd_server *s, pdpd_server_stats *out)
{
    if (!s || !out) return sfail(PDPD_ERR_INVALID, "null argument");
    *out = s->stats;
    return PDPD_OK;
}

}  // extern "C"
