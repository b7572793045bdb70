// CPU-emulated runs of the engine through its C ABI (AddressSanitizer + UBSan build).
#include "pdpd_hip.h"
#include "pdpd/hip_dpd_engine.hpp"
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>
#include <cmath>
#include <cstring>

struct W {
    double box[3];
    std::vector<double> cons, diss, x, v, f, bk, br, ak, at, ccons, cdiss;
    std::vector<uint8_t> stat, bt, mono, poff;
    std::vector<uint32_t> pid, ptbs, bh, btl, ptas, ahb, atb, ptype, pbs, pbi;
    pdpd_world_desc d{};
};

static void make(W &w, int Lx, int Ly, int Lz, double bonded, bool wall, double theta0, uint64_t seed,
                 int density = 3)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> U(0, 1);
    std::normal_distribution<double> N(0, 1);
    w.box[0] = Lx; w.box[1] = Ly; w.box[2] = Lz;
    const int n = density * Lx * Ly * Lz;
    const int clen = 6;
    const int nch = (int)(bonded * n / clen);
    const int nt = wall ? 4 : 3;
    w.cons.assign(nt * nt, 25); w.diss.assign(nt * nt, 4.5); w.stat.assign(nt, 0);
    if (wall) w.stat[3] = 1;
    // polymer types: 0 monomer W, 1 chain, 2 wall monomer
    w.ptbs = {0, 0, 5, 5}; w.ptas = {0, 0, 4, 4};
    for (int i = 0; i < 5; i++) { w.bh.push_back(i); w.btl.push_back(i + 1); w.bk.push_back(128); w.br.push_back(0.5); }
    for (int i = 0; i < 4; i++) { w.ahb.push_back(i); w.atb.push_back(i + 1); w.ak.push_back(20); w.at.push_back(theta0); }
    int b = 0, p = 0;
    w.pbs.push_back(0);
    for (int c = 0; c < nch; c++) {
        double s[3] = {U(rng) * Lx, U(rng) * Ly, U(rng) * Lz};
        for (int k = 0; k < clen; k++) {
            for (int d = 0; d < 3; d++) {
                double L = w.box[d];
                double xx = std::fmod(s[d] + 0.4 * k * (d == 0 ? 1 : 0.3) + L, L);
                if (xx >= L) xx = 0;
                w.x.push_back(xx); w.v.push_back(N(rng)); w.f.push_back(0);
            }
            w.bt.push_back(k < 2 ? 1 : 2); w.mono.push_back(0); w.pid.push_back(p); w.poff.push_back(k);
            w.pbi.push_back(b++);
        }
        w.ptype.push_back(1); p++; w.pbs.push_back(b);
    }
    while (b < n) {
        double xx = U(rng) * Lx, yy = U(rng) * Ly, zz = U(rng) * Lz;
        bool st = wall && xx < 1.0;
        w.x.push_back(xx); w.x.push_back(yy); w.x.push_back(zz);
        for (int d = 0; d < 3; d++) { w.v.push_back(st ? 0 : N(rng)); w.f.push_back(0); }
        w.bt.push_back(st ? 3 : 0); w.mono.push_back(1); w.pid.push_back(p); w.poff.push_back(0);
        w.pbi.push_back(b++);
        w.ptype.push_back(st ? 2 : 0); p++; w.pbs.push_back(b);
    }
    // polymer type 2 (wall monomer) has no bonds: fix ptbs/ptas to 3 types
    w.ptbs = {0, 0, 5, 5}; w.ptas = {0, 0, 4, 4};
    pdpd_world_desc &d = w.d;
    for (int k = 0; k < 3; k++) d.box[k] = w.box[k];
    d.dt = 0.01; d.lambda = 0.5; d.seed = seed; d.t = 0;
    d.n_bead_types = nt; d.conservative = w.cons.data(); d.dissipative = w.diss.data(); d.stationary = w.stat.data();
    d.n_beads = n; d.x = w.x.data(); d.v = w.v.data(); d.f = w.f.data();
    d.bead_type = w.bt.data(); d.is_monomer = w.mono.data(); d.polymer_id = w.pid.data(); d.polymer_offset = w.poff.data();
    d.n_polymer_types = 3; d.pt_bond_start = w.ptbs.data(); d.bond_head = w.bh.data(); d.bond_tail = w.btl.data();
    d.bond_kappa = w.bk.data(); d.bond_r0 = w.br.data(); d.pt_angle_start = w.ptas.data();
    d.angle_head_bond = w.ahb.data(); d.angle_tail_bond = w.atb.data(); d.angle_kappa = w.ak.data(); d.angle_theta0 = w.at.data();
    d.n_polymers = p; d.polymer_type = w.ptype.data(); d.polymer_bead_start = w.pbs.data(); d.polymer_bead_ids = w.pbi.data();
}

// The same world as a WorldState (for the C++ engine classes).
static WorldState to_state(const W &w)
{
    const pdpd_world_desc &d = w.d;
    WorldState s;
    s.box = {d.box[0], d.box[1], d.box[2]};
    s.dt = d.dt; s.lambda = d.lambda; s.seed = d.seed; s.t = d.t;
    const unsigned nt = d.n_bead_types;
    for (unsigned i = 0; i < nt; i++) {
        BeadType bt;
        bt.name = "T" + std::to_string(i); bt.id = i; bt.stationary = d.stationary && d.stationary[i];
        s.bead_types.push_back(bt);
    }
    s.interactions.resize(nt * nt);
    for (unsigned k = 0; k < nt * nt; k++) {
        s.interactions[k].conservative = d.conservative[k];
        s.interactions[k].dissipative = d.dissipative[k];
    }
    for (unsigned t = 0; t < d.n_polymer_types; t++) {
        PolymerType pt;
        pt.name = "P" + std::to_string(t); pt.polymer_id = t;
        for (unsigned b = d.pt_bond_start[t]; b < d.pt_bond_start[t + 1]; b++) {
            Bond bd; bd.kappa = d.bond_kappa[b]; bd.r0 = d.bond_r0[b];
            bd.bead_offset_head = d.bond_head[b]; bd.bead_offset_tail = d.bond_tail[b];
            pt.bonds.push_back(bd);
        }
        for (unsigned a = d.pt_angle_start[t]; a < d.pt_angle_start[t + 1]; a++) {
            BondPair bp; bp.kappa = d.angle_kappa[a]; bp.theta0 = d.angle_theta0[a];
            bp.bond_offset_head = d.angle_head_bond[a]; bp.bond_offset_tail = d.angle_tail_bond[a];
            pt.bond_pairs.push_back(bp);
        }
        s.polymer_types.push_back(pt);
    }
    s.beads.resize(d.n_beads);
    for (unsigned p = 0; p < d.n_polymers; p++) {
        Polymer poly; poly.polymer_id = p; poly.polymer_type = d.polymer_type[p];
        for (unsigned k = d.polymer_bead_start[p]; k < d.polymer_bead_start[p + 1]; k++) {
            const unsigned i = d.polymer_bead_ids[k];
            poly.bead_ids.push_back(i);
            if (s.polymer_types[poly.polymer_type].bead_types.size() <= k - d.polymer_bead_start[p])
                s.polymer_types[poly.polymer_type].bead_types.push_back(d.bead_type[i]);
            Bead &b = s.beads[i];
            b.bead_id = i; b.polymer_id = d.polymer_id[i]; b.polymer_offset = d.polymer_offset[i];
            b.bead_type = d.bead_type[i]; b.polymer_type = (uint8_t)poly.polymer_type; b.is_monomer = d.is_monomer[i];
            for (int c = 0; c < 3; c++) { b.x[c] = d.x[3 * i + c]; b.v[c] = d.v[3 * i + c]; b.f[c] = d.f[3 * i + c]; }
        }
        s.polymers.push_back(poly);
    }
    return s;
}

#define CHECK(e) do { int rc_ = (e); if (rc_) { fprintf(stderr, "%s -> %d: %s\n", #e, rc_, pdpd_last_error()); exit(1); } } while (0)

int main(int argc, char **argv)
{
    struct Case { const char *name; int L[3]; double bonded; bool wall; double th; int steps; const char *cfg; };
    Case cases[] = {
        {"water4", {4, 4, 4}, 0.0, false, 0.0, 1, nullptr},
        {"water3", {3, 3, 3}, 0.0, false, 0.0, 2, nullptr},
        {"poly_6x7x8", {6, 7, 8}, 0.2, false, 0.0, 3, nullptr},
        {"wall_bent", {6, 6, 6}, 0.2, true, 1.0, 3, nullptr},
        {"poly_graph", {5, 6, 7}, 0.2, false, 0.0, 4, "graph_steps=2"},
        {"poly_grid", {6, 7, 8}, 0.1, false, 0.0, 2, "force_grid=5"},
        {"poly_v1", {6, 7, 8}, 0.1, false, 0.0, 2, "force_variant=1"},
        {"wall_v2", {6, 5, 7}, 0.2, true, 1.0, 2, "force_variant=2"},
        {"wall_v3", {6, 5, 7}, 0.2, true, 1.0, 2, "force_variant=3"},
        {"poly_v4", {6, 7, 8}, 0.1, false, 0.0, 2, "force_variant=4"},
        {"wall_v5", {6, 5, 7}, 0.2, true, 1.0, 2, "force_variant=5"},
        {"wall_v6", {6, 5, 7}, 0.2, true, 1.0, 2, "force_variant=6,force_grid=4"},
        {"poly_v7", {6, 7, 8}, 0.2, false, 0.0, 3, "force_variant=7"},             // a wave per stencil row
        {"wall_v7", {6, 5, 7}, 0.2, true, 1.0, 2, "force_variant=7,force_grid=4"},
        {"dense_v7", {3, 4, 3}, 0.0, false, 0.0, 2, "force_variant=7"},           // rows in several rounds
        {"poly_scan3", {9, 8, 7}, 0.1, false, 0.0, 3, "scan_lookback_tiles=0"},
        {"dense", {3, 4, 3}, 0.0, false, 0.0, 2, nullptr},
        {"dense_brick3", {4, 4, 3}, 0.0, false, 0.0, 2, "force_kernel=1,force_brick=3"},    // every brick beyond LDS
        {"poly_brick1", {6, 7, 8}, 0.2, false, 0.0, 2, "force_kernel=1,force_brick=1"},
        {"wall_brick2", {6, 5, 7}, 0.2, true, 1.0, 2, "force_kernel=1,force_brick=2"},
        {"poly_brick5", {9, 8, 10}, 0.1, false, 0.0, 2, "force_kernel=1,force_brick=5"},
        {"poly_brick_grid", {9, 8, 10}, 0.1, false, 0.0, 3, "force_kernel=1,force_grid=3"},
        {"dense_slab6", {4, 4, 3}, 0.0, false, 0.0, 2, "force_kernel=1,force_brick=6"},     // every slab beyond LDS
        {"poly_slab7", {6, 7, 8}, 0.2, false, 0.0, 2, "force_kernel=1,force_brick=7"},
        {"wall_slab8", {6, 5, 7}, 0.2, true, 1.0, 2, "force_kernel=1,force_brick=8"},
        {"poly_slab9", {19, 6, 5}, 0.1, false, 0.0, 3, "force_kernel=1,force_brick=9"},
        {"poly_slab_grid", {9, 8, 10}, 0.1, false, 0.0, 3, "force_kernel=1,force_brick=6,force_grid=3"},
        {"poly_per_slot", {6, 7, 8}, 0.1, false, 0.0, 2, "force_kernel=0"},
        {"spatial", {5, 4, 6}, 0.2, true, 0.5, 2, nullptr},      // strengths per cell, batched with a constant world
    };
    for (auto &c : cases) {
        fprintf(stderr, "case %s\n", c.name);
        W w;
        make(w, c.L[0], c.L[1], c.L[2], c.bonded, c.wall, c.th, 7, strncmp(c.name, "dense", 5) ? 3 : 24);
        if (!strcmp(c.name, "spatial")) {
            const int nt = (int)w.d.n_bead_types, nc = c.L[0] * c.L[1] * c.L[2];
            for (int cell = 0; cell < nc; cell++)
                for (int k = 0; k < nt * nt; k++) {
                    w.ccons.push_back(20.0 + cell % 7 + k);
                    w.cdiss.push_back(3.0 + 0.1 * (cell % 5));
                }
            w.d.cell_conservative = w.ccons.data();
            w.d.cell_dissipative = w.cdiss.data();
        }
        const char *why = pdpd_can_support(&w.d);
        if (why && *why) { fprintf(stderr, "unsupported: %s\n", why); return 1; }
        pdpd_ctx *ctx;
        CHECK(pdpd_ctx_create(0, &ctx));
        for (const char *p = c.cfg; p && *p;) {     // key=value[,key=value...]
            char k[64], v[64];
            int used = 0;
            if (sscanf(p, "%63[^=]=%63[^,]%n", k, v, &used) != 2) { fprintf(stderr, "bad config\n"); return 1; }
            CHECK(pdpd_set_config(ctx, k, v));
            p += used;
            if (*p == ',') p++;
        }
        int slot;
        CHECK(pdpd_world_add(ctx, &w.d, &slot));
        if (!strcmp(c.name, "poly_6x7x8")) {      // batch: two more worlds
            W w2, w3;
            make(w2, 4, 5, 6, 0.1, false, 0.0, 9);
            make(w3, 7, 3, 5, 0.0, false, 0.0, 11);
            int s2, s3;
            CHECK(pdpd_world_add(ctx, &w2.d, &s2));
            CHECK(pdpd_world_add(ctx, &w3.d, &s3));
            CHECK(pdpd_world_set_steps(ctx, s3, 1));
            CHECK(pdpd_run(ctx, c.steps));
            CHECK(pdpd_sync(ctx));
            W w4;                                 // a world joins, one leaves: resident worlds move
            make(w4, 5, 4, 6, 0.2, true, 0.5, 13);
            int s4;
            CHECK(pdpd_world_add(ctx, &w4.d, &s4));
            CHECK(pdpd_run(ctx, 1));
            CHECK(pdpd_world_remove(ctx, s2));
            CHECK(pdpd_run(ctx, 1));
            CHECK(pdpd_sync(ctx));
        } else if (!strcmp(c.name, "spatial")) {
            W w2;
            make(w2, 4, 5, 6, 0.1, false, 0.0, 9);
            int s2;
            CHECK(pdpd_world_add(ctx, &w2.d, &s2));
            CHECK(pdpd_run(ctx, c.steps));
            CHECK(pdpd_sync(ctx));
        } else {
            CHECK(pdpd_run(ctx, c.steps));
            CHECK(pdpd_sync(ctx));
        }
        {   // force trace of the last step: counts, then the records
            uint32_t np = 0, nbd = 0;
            CHECK(pdpd_trace_forces(ctx, slot, nullptr, 0, &np, nullptr, 0, &nbd));
            std::vector<pdpd_pair_record> pr(np);
            std::vector<pdpd_bond_record> br(nbd);
            CHECK(pdpd_trace_forces(ctx, slot, pr.data(), np, &np, br.data(), nbd, &nbd));
            if (np == 0 || (c.bonded > 0 && nbd == 0)) { fprintf(stderr, "empty trace\n"); return 1; }
            fprintf(stderr, "  trace %u pairs %u bonds\n", np, nbd);
        }
        std::vector<double> x(3 * w.d.n_beads), v(x.size()), f(x.size());
        int64_t t;
        CHECK(pdpd_world_download(ctx, slot, x.data(), v.data(), f.data(), &t));
        {   // a snapshot taken before one more step holds the state before it
            std::vector<double> sx(x.size()), sv(x.size()), sf(x.size());
            int64_t st;
            CHECK(pdpd_snapshot(ctx));
            CHECK(pdpd_run(ctx, 1));
            CHECK(pdpd_world_snapshot_download(ctx, slot, sx.data(), sv.data(), sf.data(), &st));
            if (st != t || sx != x || sv != v || sf != f) { fprintf(stderr, "snapshot differs\n"); return 1; }
            CHECK(pdpd_world_download(ctx, slot, x.data(), v.data(), f.data(), &t));
        }
        double fs = 0;
        for (double q : f) fs += std::fabs(q);
        fprintf(stderr, "  ok t=%lld sum|f|=%g\n", (long long)t, fs);
        CHECK(pdpd_ctx_destroy(ctx));
    }
    {   // the C++ batched engine over three contexts against each world run alone
        fprintf(stderr, "case batch_engine\n");
        W a, b, c, e;
        make(a, 4, 5, 6, 0.2, false, 0.0, 21);
        make(b, 5, 4, 4, 0.1, true, 0.5, 22);
        make(c, 6, 5, 5, 0.0, false, 0.0, 23);
        make(e, 4, 4, 4, 0.2, false, 0.0, 24);
        W *ws[4] = {&a, &b, &c, &e};
        std::vector<WorldState> st, solo;
        for (W *w : ws) { st.push_back(to_state(*w)); solo.push_back(to_state(*w)); }
        HipBatchDPDEngine batch(0, 3);
        std::vector<int> slots;
        for (auto &s : st) slots.push_back(batch.Attach(&s));
        batch.SetSteps(slots[1], 2);
        batch.Run(5);
        for (size_t i = 0; i < solo.size(); i++) {
            HipDPDEngine one;
            one.Attach(&solo[i]);
            one.Run(i == 1 ? 2 : 5);
            for (size_t k = 0; k < solo[i].beads.size(); k++)
                for (int q = 0; q < 3; q++)
                    if (solo[i].beads[k].x[q] != st[i].beads[k].x[q] || solo[i].beads[k].v[q] != st[i].beads[k].v[q] ||
                        solo[i].beads[k].f[q] != st[i].beads[k].f[q]) {
                        fprintf(stderr, "batch engine: world %zu bead %zu differs from its solo run\n", i, k);
                        return 1;
                    }
            if (solo[i].t != st[i].t) { fprintf(stderr, "batch engine: t differs\n"); return 1; }
            one.Attach(nullptr);
        }
        fprintf(stderr, "  ok %zu worlds over %zu contexts\n", batch.size(), batch.parts());
    }
    {   // worlds retired and admitted between runs: resident worlds (bonded ones
        // too) move to new offsets on the device and go on as if run alone
        fprintf(stderr, "case relayout\n");
        W a, b, c, e;
        make(a, 5, 4, 6, 0.2, false, 0.0, 31);
        make(b, 6, 5, 5, 0.2, true, 0.5, 32);
        make(c, 4, 4, 5, 0.0, false, 0.0, 33);
        make(e, 4, 5, 4, 0.2, false, 0.0, 34);
        pdpd_ctx *ctx;
        CHECK(pdpd_ctx_create(0, &ctx));
        int sa, sb, sc, se;
        CHECK(pdpd_world_add(ctx, &a.d, &sa));
        CHECK(pdpd_world_add(ctx, &b.d, &sb));
        CHECK(pdpd_world_add(ctx, &c.d, &sc));
        CHECK(pdpd_run(ctx, 3));
        CHECK(pdpd_world_remove(ctx, sa));
        CHECK(pdpd_world_add(ctx, &e.d, &se));
        CHECK(pdpd_run(ctx, 2));
        CHECK(pdpd_world_remove(ctx, se));
        CHECK(pdpd_run(ctx, 1));
        CHECK(pdpd_sync(ctx));
        pdpd_stats stt;
        CHECK(pdpd_get_stats(ctx, &stt));
        if (stt.relayouts != 3) { fprintf(stderr, "relayouts %llu, expected 3\n", (unsigned long long)stt.relayouts); return 1; }
        const struct { W *w; int slot; } kept[] = {{&b, sb}, {&c, sc}};
        for (const auto &k : kept) {
            const size_t n3 = 3 * k.w->d.n_beads;
            std::vector<double> x(n3), v(n3), f(n3), x1(n3), v1(n3), f1(n3);
            int64_t t, t1;
            CHECK(pdpd_world_download(ctx, k.slot, x.data(), v.data(), f.data(), &t));
            pdpd_ctx *one;
            CHECK(pdpd_ctx_create(0, &one));
            int s1;
            CHECK(pdpd_world_add(one, &k.w->d, &s1));
            CHECK(pdpd_run(one, 6));
            CHECK(pdpd_world_download(one, s1, x1.data(), v1.data(), f1.data(), &t1));
            CHECK(pdpd_ctx_destroy(one));
            if (t != 6 || t1 != 6 || x != x1 || v != v1 || f != f1) {
                fprintf(stderr, "relayout: world %d differs from its solo run\n", k.slot);
                return 1;
            }
        }
        CHECK(pdpd_ctx_destroy(ctx));
        fprintf(stderr, "  ok %llu relayouts\n", (unsigned long long)stt.relayouts);
    }
    fprintf(stderr, "all cases done\n");
    return 0;
}
