// CPU-emulated runs of the engine through its C ABI (AddressSanitizer + UBSan build).
#include "pdpd_hip.h"
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
        {"poly_scan3", {9, 8, 7}, 0.1, false, 0.0, 3, "scan_lookback_tiles=0"},
        {"dense", {3, 4, 3}, 0.0, false, 0.0, 2, nullptr},
        {"spatial", {5, 4, 6}, 0.2, true, 0.5, 2, nullptr},      // strengths per cell, batched with a constant world
    };
    for (auto &c : cases) {
        fprintf(stderr, "case %s\n", c.name);
        W w;
        make(w, c.L[0], c.L[1], c.L[2], c.bonded, c.wall, c.th, 7, strcmp(c.name, "dense") ? 3 : 24);
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
        if (c.cfg) {
            char k[64], v[64];
            sscanf(c.cfg, "%63[^=]=%63s", k, v);
            CHECK(pdpd_set_config(ctx, k, v));
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
    fprintf(stderr, "all cases done\n");
    return 0;
}
