// pdpd_state -- read a WorldState file in any format (text v0, v1binary,
// v2binary; ".gz" / ".bz2" compressed) and write it back, e.g. to convert between
// formats, compress, or re-check a state file.
//
//   pdpd_state <in> <out> [--binary] [--precise] [--validate]
//     --binary    write v2binary (default: text v0)
//     --precise   text numbers with 17 significant digits (default 6, as the reference)
//     --validate  run the reference tools' WorldState checks first
#include "pdpd/world_state_io.hpp"
#include "pdpd/world_state_validate.hpp"

#include <cstdio>
#include <cstring>

int main(int argc, const char *argv[])
{
    if (argc < 3) {
        fprintf(stderr, "usage: pdpd_state <in> <out> [--binary] [--precise] [--validate]\n");
        return 1;
    }
    bool binary = false, precise = false, check = false;
    for (int i = 3; i < argc; i++) {
        if (!strcmp(argv[i], "--binary")) binary = true;
        else if (!strcmp(argv[i], "--precise")) precise = true;
        else if (!strcmp(argv[i], "--validate")) check = true;
        else {
            fprintf(stderr, "pdpd_state: unknown option %s\n", argv[i]);
            return 1;
        }
    }
    try {
        WorldState s = pdpd_io::read_world_state(std::string(argv[1]));
        if (check) pdpd_io::validate_world_state(s);
        pdpd_io::write_world_state(std::string(argv[2]), s, binary, precise);
    } catch (const std::exception &e) {
        fprintf(stderr, "pdpd_state: %s\n", e.what());
        return 1;
    }
    return 0;
}
