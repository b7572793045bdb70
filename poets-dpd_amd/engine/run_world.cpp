// run_world -- advance a WorldState file with a registered engine, writing the
// state after intervals and one progress line per interval. Command line,
// outputs and progress line are those of poets-dpd's src/run_world.cpp:
//
//   run_world <engine> <src-file> <output-base-name> [interval_count=1000]
//             [state_interval_size=1] [snapshot_interval_size] [--text-state]
//
// The engine runs intervals of min(state, snapshot) steps; every
// state_interval_size steps the state is written as
// "<base>.<t, 9 digits>.state.gz" in v2binary (src/run_world.cpp:201-209).
// --text-state (not in the reference) writes those files as text v0 with 17
// significant digits instead, which keeps every bit of the doubles. The
// reference's visual snapshots (--vtk-snapshot, --povray-*, --solvent-free-snapshot)
// are not produced by this tool and are rejected. The progress line on stdout is
// "engine,src,volume,beads,steps_done,attach_s,interval_s,
// bead_steps_per_s_so_far,bead_steps_per_s_interval" (src/run_world.cpp:181-185).
// The world is checked with validate_world_state(state, GetMaxBondLength())
// before it is attached, as src/run_world.cpp:155-157 does. States are written
// on host threads while the engine runs the next interval. With PDPD_LOG=path
// in the environment every step's forces are logged to path in the reference's
// force-log format (src/run_world.cpp:120-125; the engine then runs one step
// at a time).
#include "pdpd/hip_dpd_engine.hpp"
#include "pdpd/world_state_io.hpp"
#include "pdpd/world_state_validate.hpp"

#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <memory>
#include <queue>
#include <string>
#include <vector>

extern "C" int pdpd_engines_registered(void);

static double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

static int usage()
{
    fprintf(stderr, "run_world : engine-name src-file output-base-name interval_count state_interval_size "
                    "[snapshot_interval_size] [--text-state]\n");
    fprintf(stderr, "   --text-state : write the states as text v0 with 17 significant digits instead of v2binary.\n");
    fprintf(stderr, "  engine names:\n");
    for (const auto &n : DPDEngineFactory::ListFactories()) fprintf(stderr, "    %s\n", n.c_str());
    fprintf(stderr, "  env:\n     PDPD_LOG=log-path : do full force logging to given file.\n");
    fprintf(stderr, "    Either snapshot_interval_size must divide state_interval_size, or vice-versa\n");
    return 1;
}

int main(int argc, const char *argv[])
{
    pdpd_engines_registered();
    bool text_state = false;
    std::vector<std::string> args;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        if (a == "--text-state") {
            text_state = true;
        } else if (a == "--vtk-snapshot" || a == "--povray-snapshot" || a == "--povray-render" ||
                   a == "--solvent-free-snapshot" || a == "--gzip-snapshot") {
            fprintf(stderr, "run_world: %s is not supported (this tool writes state files only)\n", a.c_str());
            return 1;
        } else {
            args.push_back(a);
        }
    }
    if (args.size() < 3) return usage();
    std::unique_ptr<FileLogger> logger;
    if (const char *dst = getenv("PDPD_LOG")) {
        logger.reset(new FileLogger(dst));
        ForceLogging::set_logger(logger.get());
        fprintf(stderr, "Logging to %s\n", dst);
    }
    struct Unset {
        ~Unset() { ForceLogging::set_logger(nullptr); }
    } unset;
    try {
        const std::string engine_name = args[0], src = args[1], base = args[2];
        const int interval_count = args.size() > 3 ? std::stoi(args[3]) : 1000;
        const int state_interval = args.size() > 4 ? std::stoi(args[4]) : 1;
        const int snapshot_interval = args.size() > 5 ? std::stoi(args[5]) : state_interval;
        if (state_interval <= 0 || snapshot_interval <= 0) {
            fprintf(stderr, "interval sizes must be positive.\n");
            return 1;
        }
        const int interval_size = std::min(state_interval, snapshot_interval);
        fprintf(stderr, "interval_count=%d, interval_size=%d, state_interval_size=%d, snapshot_interval_size=%d\n",
                interval_count, interval_size, state_interval, snapshot_interval);
        if (state_interval % snapshot_interval != 0 && snapshot_interval % state_interval != 0) {
            fprintf(stderr, "state_interval_size and snapshot_interval_size are not compatible.\n");
            return 1;
        }

        WorldState state = pdpd_io::read_world_state(src);
        std::shared_ptr<DPDEngine> engine = DPDEngineFactory::create(engine_name);
        fprintf(stderr, "Validating with max_r=%f....\n", engine->GetMaxBondLength());
        pdpd_io::validate_world_state(state, engine->GetMaxBondLength());
        fprintf(stderr, "  ok\n");
        const std::string why = engine->CanSupport(&state);
        if (!why.empty()) {
            fprintf(stderr, "Engine can't support world : %s\n", why.c_str());
            return 1;
        }
        const int volume = int(state.box[0] * state.box[1] * state.box[2]);
        const double t0 = now();
        engine->Attach(&state);
        double t1 = now();
        const double attach_s = t1 - t0;
        const int state_modulus = state_interval / interval_size;
        unsigned done = 0, slice = 0;
        std::queue<std::future<void>> writes;
        engine->Run(interval_count, (unsigned)interval_size, [&]() {
            ++slice;
            const double t2 = now();
            done += interval_size;
            const double n = (double)state.beads.size();
            fprintf(stdout, "%s,%s,%d,%d,%u,%g,%g,%g,%g\n", engine_name.c_str(), src.c_str(), volume, (int)n, done,
                    attach_s, t2 - t1, n / (t2 - t0) * done, n / (t2 - t1) * interval_size);
            fflush(stdout);
            t1 = t2;
            if (slice % state_modulus == 0) {
                char suffix[64];
                snprintf(suffix, sizeof(suffix), ".%09d.state.gz", (int)state.t);
                auto snap = std::make_shared<WorldState>(state);
                const std::string file = base + suffix;
                while (writes.size() > 4) {
                    writes.front().get();
                    writes.pop();
                }
                writes.push(std::async(std::launch::async, [snap, file, text_state]() {
                    pdpd_io::write_world_state(file, *snap, !text_state, text_state);
                }));
            }
            return true;
        });
        while (!writes.empty()) {
            writes.front().get();
            writes.pop();
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "run_world: exception: %s\n", e.what());
        return 1;
    }
    return 0;
}
