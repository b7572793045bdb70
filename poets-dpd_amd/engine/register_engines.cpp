// register_engines.cpp -- registers the MI355X engines in DPDEngineFactory the
// way every reference engine registers itself (a static initialiser per engine,
// e.g. src/engines/naive_dpd_engine.cpp:3-6 in poets-dpd), so tools that create
// engines by name find them.
//   hip_dpd_engine          graph-captured steps (20 per graph)
//   hip_dpd_engine_direct   one launch set per step, no graph
#include "pdpd/hip_dpd_engine.hpp"

namespace {

const bool registered_graph = DPDEngineFactory::RegisterFactory(
    "hip_dpd_engine", []() { return std::make_shared<HipDPDEngine>(0, 20); });

const bool registered_direct = DPDEngineFactory::RegisterFactory(
    "hip_dpd_engine_direct", []() { return std::make_shared<HipDPDEngine>(0, 0); });

}  // namespace

// Referenced by tools so the linker keeps this object's initialisers.
extern "C" int pdpd_engines_registered(void) { return registered_graph && registered_direct; }
