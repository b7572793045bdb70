// ref_engine_tests.cpp -- runs the reference's own engine test suite against the
// HIP engine. Built by oracle/Makefile ("ref" target) from the reference's
// headers where they lie (include/dpd/tests/*.hpp, naive_dpd_engine.hpp) plus
// this repository's include/pdpd/hip_dpd_engine.hpp compiled with
// PDPD_HOST_REFERENCE, i.e. against the reference's own DPDEngine / WorldState.
// That build is the drop-in check: the engine header compiles and runs inside
// the reference's type system unchanged. TEST INFRASTRUCTURE ONLY.
//
// Modes (TAP output on stdout, exit status 1 if any test fails):
//   scenarios [prefix]     test_runner.hpp on every registered scenario
//                          (the role of src/test/test_engine.cpp:33-96)
//   differential [prefix]  test_differential.hpp against NaiveDPDEngine<false>
//                          (the role of src/test/test_engine_diff.cpp:29-103)
#include "dpd/tests/test_ordered_mesh.hpp"
#include "dpd/tests/test_bonded_bead_angle.hpp"
#include "dpd/tests/test_bonded_bead_triple.hpp"
#include "dpd/tests/test_bonded_bead_pair.hpp"
#include "dpd/tests/test_colliding_bead_pair_diss.hpp"
#include "dpd/tests/test_colliding_bead_pair.hpp"
#include "dpd/tests/test_isolated_moving_bead.hpp"
#include "dpd/tests/test_multiple_moving_beads.hpp"
#include "dpd/engines/naive/naive_dpd_engine.hpp"
#include "dpd/core/logging_impl.hpp"
#include "dpd/tests/test_runner.hpp"
#include "dpd/tests/test_differential.hpp"

#define PDPD_HOST_REFERENCE 1
#include "pdpd/hip_dpd_engine.hpp"

#include <cstdio>
#include <string>

int main(int argc, const char *argv[])
{
    std::string mode = argc > 1 ? argv[1] : "scenarios";
    std::string prefix = argc > 2 ? argv[2] : "";
    if (mode != "scenarios" && mode != "differential") {
        fprintf(stderr, "usage: ref_engine_tests scenarios|differential [prefix]\n");
        return 2;
    }
    TestOrderedMesh::register_tests();
    TestBondedBeadAngle::register_tests();
    TestBondedBeadTriple::register_tests();
    TestBondedBeadPair::register_tests();
    TestCollidingBeadPair::register_tests();
    TestCollidingBeadPairDiss::register_tests();
    TestIsolatedMovingBead::register_tests();
    TestMultipleMovingBeads::register_tests();

    fprintf(stdout, "TAP version 13\n");
    int num = 1, failed = 0;
    for (const auto &tt : TestBase::get_test_factories_ordered_by_complexity()) {
        if (!prefix.empty() && std::get<0>(tt).substr(0, prefix.size()) != prefix) continue;
        auto test = std::get<2>(tt)();
        HipDPDEngine dut;
        std::pair<bool, std::string> r;
        if (mode == "scenarios") {
            r = test_runner(*test, dut);
        } else {
            NaiveDPDEngine<false> ref;
            r = test_differential(*test, ref, dut);
        }
        fprintf(stdout, "%s %d - %s : %s\n", r.first ? "ok" : "not ok", num, test->name().c_str(), r.second.c_str());
        fflush(stdout);
        failed += r.first ? 0 : 1;
        num++;
    }
    fprintf(stdout, "1..%d\n", num - 1);
    return failed > 0;
}
