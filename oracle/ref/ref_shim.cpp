// ref_shim.cpp -- C ABI around the reference's own engines, compiled from the
// headers where they lie under /root/reference/include (oracle/Makefile, target
// "ref"). Output goes to oracle/_ref/ only. TEST INFRASTRUCTURE: used to pin the
// oracle restatement, to write tests/golden fixtures and (its -O3 -ffast-math
// build) to time the CPU baseline. The built library travels with the tree to
// the GPU host, where only tests and the benchmark's CPU-baseline leg load it;
// the engine library never links it.
//
// It builds a reference WorldState from the same pdpd_world_desc the HIP engine
// takes, runs one of the reference engines for n steps, and copies x/v/f back.
#include "dpd/core/dpd_engine.hpp"
#include "dpd/core/hash.hpp"
#include "dpd/core/dpd_state_validator.hpp"
#include "dpd/maths/dpd_maths_core.hpp"
#include "dpd/engines/naive/naive_dpd_engine.hpp"
#include "dpd/engines/naive/naive_dpd_engine_half_merge.hpp"
#include "dpd/engines/naive/naive_dpd_engine_half_step.hpp"
#include "dpd/core/logging.hpp"
#include "dpd/core/logging_impl.hpp"
#include "dpd/core/dpd_state_io.hpp"

#include "pdpd_hip.h"

#include <chrono>
#include <cstring>
#include <memory>
#include <string>

static WorldState make_state(const pdpd_world_desc *d)
{
    WorldState s;
    for (int k = 0; k < 3; k++) s.box[k] = d->box[k];
    s.dt = d->dt;
    s.lambda = d->lambda;
    s.seed = d->seed;
    s.t = d->t;
    const unsigned nt = d->n_bead_types;
    for (unsigned i = 0; i < nt; i++) {
        BeadType bt;
        bt.name = "B" + std::to_string(i);
        bt.r = 0.5;
        bt.id = i;
        bt.stationary = d->stationary ? d->stationary[i] != 0 : false;
        s.bead_types.push_back(bt);
    }
    for (unsigned i = 0; i < nt * nt; i++) {
        InteractionStrength is;
        is.conservative = d->conservative[i];
        is.dissipative = d->dissipative[i];
        s.interactions.push_back(is);
    }
    std::vector<int> pt_first_polymer(d->n_polymer_types, -1);
    for (unsigned p = 0; p < d->n_polymers; p++)
        if (pt_first_polymer[d->polymer_type[p]] < 0) pt_first_polymer[d->polymer_type[p]] = (int)p;
    for (unsigned pt = 0; pt < d->n_polymer_types; pt++) {
        PolymerType T;
        T.name = "P" + std::to_string(pt);
        T.polymer_id = pt;
        int p0 = pt_first_polymer[pt];
        if (p0 >= 0) {
            for (unsigned k = d->polymer_bead_start[p0]; k < d->polymer_bead_start[p0 + 1]; k++)
                T.bead_types.push_back(d->bead_type[d->polymer_bead_ids[k]]);
        }
        for (unsigned b = d->pt_bond_start[pt]; b < d->pt_bond_start[pt + 1]; b++) {
            Bond B;
            B.kappa = d->bond_kappa[b];
            B.r0 = d->bond_r0[b];
            B.bead_offset_head = d->bond_head[b];
            B.bead_offset_tail = d->bond_tail[b];
            T.bonds.push_back(B);
        }
        for (unsigned a = d->pt_angle_start[pt]; a < d->pt_angle_start[pt + 1]; a++) {
            BondPair P;
            P.kappa = d->angle_kappa[a];
            P.theta0 = d->angle_theta0[a];
            P.bond_offset_head = d->angle_head_bond[a];
            P.bond_offset_tail = d->angle_tail_bond[a];
            T.bond_pairs.push_back(P);
        }
        s.polymer_types.push_back(T);
    }
    std::vector<uint32_t> ptype_of_bead(d->n_beads, 0);
    for (unsigned p = 0; p < d->n_polymers; p++) {
        Polymer P;
        P.polymer_id = p;
        P.polymer_type = d->polymer_type[p];
        for (unsigned k = d->polymer_bead_start[p]; k < d->polymer_bead_start[p + 1]; k++) {
            P.bead_ids.push_back(d->polymer_bead_ids[k]);
            ptype_of_bead[d->polymer_bead_ids[k]] = d->polymer_type[p];
        }
        s.polymers.push_back(P);
    }
    for (unsigned i = 0; i < d->n_beads; i++) {
        Bead b;
        b.bead_id = i;
        b.polymer_id = d->polymer_id[i];
        b.polymer_offset = d->polymer_offset[i];
        b.bead_type = d->bead_type[i];
        b.polymer_type = (uint8_t)ptype_of_bead[i];
        b.is_monomer = d->is_monomer[i] != 0;
        for (int k = 0; k < 3; k++) {
            b.x[k] = d->x[3 * i + k];
            b.v[k] = d->v[3 * i + k];
            b.f[k] = d->f[3 * i + k];
        }
        s.beads.push_back(b);
    }
    return s;
}

static std::string g_err;

#include "../../tests/cpp/force_log_sample.inc"

extern "C" {

const char *ref_last_error(void) { return g_err.c_str(); }

uint64_t ref_get_t_hash(uint32_t t, uint64_t seed) { return get_t_hash(t, seed); }

uint32_t ref_hash_rng_sym(uint64_t t_hash, uint32_t a, uint32_t b) { return hash_rng_sym(t_hash, a, b); }

float ref_default_hash(uint64_t t_hash, uint32_t a, uint32_t b) { return dpd_maths_core::default_hash(t_hash, a, b); }

uint32_t ref_bead_hash(uint32_t bead_type, int is_monomer, uint32_t polymer_id, uint32_t polymer_offset)
{
    return BeadHash::construct(bead_type, is_monomer != 0, polymer_id, polymer_offset).hash;
}

// Run(nsteps) is timed alone (Attach excluded, as benchmark_engine.cpp:128-147 does).
// engine: "naive" (NaiveDPDEngine<false>), "naive_core" (NaiveDPDEngine<true>),
// "half_merge" (NaiveDPDEngineHalfMerge), "half_step" (NaiveDPDEngineHalfStep).
// Returns 0, 1 if CanSupport rejects the world (reason in ref_last_error), -1 on exception.
int ref_run(const pdpd_world_desc *d, const char *engine, int nsteps, double *x, double *v, double *f,
            int64_t *t_out, double *run_seconds)
{
    try {
        WorldState s = make_state(d);
        std::unique_ptr<DPDEngine> e;
        std::string name(engine);
        if (name == "naive") e.reset(new NaiveDPDEngine<false>());
        else if (name == "naive_core") e.reset(new NaiveDPDEngine<true>());
        else if (name == "half_merge") e.reset(new NaiveDPDEngineHalfMerge());
        else if (name == "half_step") e.reset(new NaiveDPDEngineHalfStep());
        else { g_err = "unknown engine " + name; return -1; }
        std::string why = e->CanSupport(&s);
        if (!why.empty()) { g_err = why; return 1; }
        e->Attach(&s);
        auto t0 = std::chrono::steady_clock::now();
        if (nsteps > 0) e->Run((unsigned)nsteps);
        auto t1 = std::chrono::steady_clock::now();
        if (run_seconds) *run_seconds = std::chrono::duration<double>(t1 - t0).count();
        for (unsigned i = 0; i < d->n_beads; i++) {
            for (int k = 0; k < 3; k++) {
                if (x) x[3 * i + k] = s.beads[i].x[k];
                if (v) v[3 * i + k] = s.beads[i].v[k];
                if (f) f[3 * i + k] = s.beads[i].f[k];
            }
        }
        if (t_out) *t_out = s.t;
        e->Attach(nullptr);
        return 0;
    } catch (std::exception &ex) {
        g_err = ex.what();
        return -1;
    } catch (std::string &sx) {
        g_err = sx;
        return -1;
    }
}

// A run of a world read from a state file by the reference's own reader
// (dpd_state_io.hpp), so parameters given as expressions reach the engine as
// the reference parses them. x, v, f: 3 * beads, bead-index order.
int ref_run_file(const char *path, const char *engine, int nsteps, double *x, double *v, double *f, int64_t *t_out)
{
    try {
        WorldState s = read_world_state(std::string(path));
        std::unique_ptr<DPDEngine> e;
        std::string name(engine);
        if (name == "naive") e.reset(new NaiveDPDEngine<false>());
        else if (name == "naive_core") e.reset(new NaiveDPDEngine<true>());
        else { g_err = "unknown engine " + name; return -1; }
        std::string why = e->CanSupport(&s);
        if (!why.empty()) { g_err = why; return 1; }
        e->Attach(&s);
        if (nsteps > 0) e->Run((unsigned)nsteps);
        for (size_t i = 0; i < s.beads.size(); i++)
            for (int k = 0; k < 3; k++) {
                if (x) x[3 * i + k] = s.beads[i].x[k];
                if (v) v[3 * i + k] = s.beads[i].v[k];
                if (f) f[3 * i + k] = s.beads[i].f[k];
            }
        if (t_out) *t_out = s.t;
        e->Attach(nullptr);
        return 0;
    } catch (std::exception &ex) {
        g_err = ex.what();
        return -1;
    } catch (std::string &sx) {
        g_err = sx;
        return -1;
    }
}

// The same run with the reference's FileLogger writing every logged force
// property to log_path (the force-log format of logging_impl.hpp).
int ref_run_logged(const pdpd_world_desc *d, const char *engine, int nsteps, const char *log_path)
{
    FileLogger *log = new FileLogger(log_path);
    ForceLogging::set_logger(log);
    const int rc = ref_run(d, engine, nsteps, nullptr, nullptr, nullptr, nullptr, nullptr);
    ForceLogging::set_logger(nullptr);
    delete log;
    return rc;
}

// tests/cpp/force_log_sample.inc through the reference's FileLogger.
int ref_force_log_sample(const char *path)
{
    FileLogger log(path);
    force_log_sample(log);
    return 0;
}

}  // extern "C"
