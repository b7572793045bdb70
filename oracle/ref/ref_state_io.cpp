// ref_state_io.cpp -- the reference's own WorldState reader and writer
// (dpd_state_io.hpp, cvec3.hpp), compiled from its headers where they lie: reads
// a state file in any of its formats and writes it back. TEST INFRASTRUCTURE
// ONLY: the tests feed it files written by this repository and compare the
// bytes both ways.
//
//   ref_state_io <in> <out> [text|v2|v1]
//
// text and v2 are the reference's write_world_state(..., false / true). The
// reference never writes v1binary (its writer has the v1 header commented out),
// so v1 takes its v2 header, renames it, and writes the records with the
// reference's own write_polymer_binary in v1 mode.
#include "dpd/core/dpd_state_io.hpp"

#include <cstdio>
#include <cstring>
#include <sstream>

int main(int argc, const char *argv[])
{
    if (argc != 3 && argc != 4) {
        fprintf(stderr, "usage: ref_state_io <in> <out> [text|v2|v1]\n");
        return 2;
    }
    const std::string mode = argc > 3 ? argv[3] : "text";
    try {
        WorldState s = read_world_state(std::string(argv[1]));
        if (mode == "text" || mode == "v2") {
            write_world_state(std::string(argv[2]), s, mode == "v2");
        } else if (mode == "v1") {
            std::ostringstream head;
            write_world_state(head, s, true);
            std::string h = head.str();
            const std::string mark = "BinaryData\n";
            h = h.substr(0, h.find(mark) + mark.size());
            h.replace(h.find("v2binary"), 8, "v1binary");
            with_optional_gzip_ostream(argv[2], [&](std::ostream &dst) {
                dst << h;
                char nul = 0;
                dst.write(&nul, 1);
                BinaryIOWriteContext ctxt{dst};
                ctxt.use_v2_binary = false;
                for (const Polymer &p : s.polymers) write_polymer_binary(p, s, ctxt);
            });
        } else {
            fprintf(stderr, "ref_state_io: unknown mode %s\n", mode.c_str());
            return 2;
        }
    } catch (const std::exception &e) {
        fprintf(stderr, "ref_state_io: %s\n", e.what());
        return 1;
    }
    return 0;
}
