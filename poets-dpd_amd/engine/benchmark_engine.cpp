// benchmark_engine -- bead-steps/s of a registered engine on a uniform world,
// with the output and timing rule of poets-dpd's src/benchmark_engine.cpp:
// Attach is timed apart from Run, steps grow by 3/2 until the rate settles, and
// each line is "engine,mode,volume,beads,steps,attach_s,run_s,bead_steps_per_s".
//
//   benchmark_engine <engine> uniform-<n> [start_steps]
//
// The world is n^3 cells at density 3 with one bead type, placed on the R3
// quasi-random sequence (alpha = 1/g, 1/g^2, 1/g^3 with g the plastic number
// 1.22074408460575947536) as the reference's make_uniform does
// (src/benchmark_engine.cpp:32-76); interactions are the builder defaults
// a = 1, gamma = 0.1 that the reference's world ends up with.
#include "pdpd/hip_dpd_engine.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

extern "C" int pdpd_engines_registered(void);

static double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static WorldState make_uniform(int dim)
{
    WorldState s;
    s.box = {double(dim), double(dim), double(dim)};
    s.dt = 0.01;
    s.lambda = 0.5;
    s.t = 0;
    s.seed = 1;
    BeadType w;
    w.name = "W";
    w.r = 0.5;
    w.id = 0;
    s.bead_types.push_back(w);
    InteractionStrength is;
    is.conservative = 1.0;
    is.dissipative = 0.1;
    s.interactions.push_back(is);
    PolymerType pt;
    pt.name = "M";
    pt.bead_types = {0};
    pt.polymer_id = 0;
    s.polymer_types.push_back(pt);
    const double g = 1.22074408460575947536;
    const double inc[3] = {1.0 / g, 1.0 / (g * g), 1.0 / (g * g * g)};
    double x[3] = {0.5, 0.5, 0.5};
    const int n = 3 * dim * dim * dim;
    for (int i = 0; i < n; i++) {
        Bead b;
        b.bead_id = i;
        b.polymer_id = i;
        b.polymer_offset = 0;
        b.bead_type = 0;
        b.polymer_type = 0;
        b.is_monomer = true;
        for (int d = 0; d < 3; d++) b.x[d] = x[d] * dim;
        s.beads.push_back(b);
        Polymer p;
        p.bead_ids = {(uint32_t)i};
        p.polymer_id = i;
        p.polymer_type = 0;
        s.polymers.push_back(p);
        for (int d = 0; d < 3; d++) {
            x[d] += inc[d];
            if (x[d] >= 1) x[d] -= 1;
        }
    }
    return s;
}

int main(int argc, const char *argv[])
{
    pdpd_engines_registered();
    if (argc < 3 || std::string(argv[2]).rfind("uniform-", 0) != 0) {
        fprintf(stderr, "benchmark_engine : engine-name uniform-n [start_steps]\n  engine names:\n");
        for (const auto &s : DPDEngineFactory::ListFactories()) fprintf(stderr, "    %s\n", s.c_str());
        return 1;
    }
    const std::string engine_name = argv[1], mode = argv[2];
    const int dim = std::stoi(mode.substr(8));
    int start = argc > 3 ? std::stoi(argv[3]) : 1;
    WorldState state = make_uniform(dim);
    auto engine = DPDEngineFactory::create(engine_name);
    const std::string why = engine->CanSupport(&state);
    if (!why.empty()) {
        fprintf(stderr, "engine cannot run this world: %s\n", why.c_str());
        return 1;
    }
    double prev = 0;
    for (int todo = start; todo < 1000000; todo = std::max(2, todo * 3 / 2)) {
        const double t0 = now();
        engine->Attach(&state);
        const double t1 = now();
        engine->Run(todo);
        const double t2 = now();
        fprintf(stdout, "%s,%s,%d,%d,%d,%g,%g,%g\n", engine_name.c_str(), mode.c_str(), dim * dim * dim,
                (int)state.beads.size(), todo, t1 - t0, t2 - t1, state.beads.size() / (t2 - t1) * todo);
        fflush(stdout);
        const double rate = (t2 - t0) / todo;
        if (prev > 0 && (t2 - t0) > 2 && std::fabs((rate - prev) / prev) < 0.05) break;
        prev = rate;
    }
    return 0;
}
