// Writes tests/cpp/force_log_sample.inc's calls through this repository's
// FileLogger (include/pdpd/force_logging.hpp) to argv[1]. Built and run by
// tests/test_force_log.py.
#include "pdpd/force_logging.hpp"

#include "force_log_sample.inc"

int main(int argc, char **argv)
{
    if (argc != 2) return 2;
    FileLogger log(argv[1]);
    force_log_sample(log);
    return 0;
}
