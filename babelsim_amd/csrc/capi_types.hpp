// Opaque handle types of the C ABI (include/bsim.h) and the conversions of
// its configuration structs, shared by the translation units that implement
// the entry points.
#pragma once

#include "../../include/bsim.h"
#include "piso.hpp"
#include "simple.hpp"

struct bsim_mesh {
    bsim::MeshDesc d;
};
struct bsim_ctx {
    std::unique_ptr<bsim::Context> c;
};

namespace bsim {
// bsim_simple_config -> SimpleCfg with the solution dictionary's checks
// (src/physics/simple/main.cpp:19-44; LinearSolverConfig::validate).
SimpleCfg toSimpleCfg(const bsim_simple_config& cfg);
// bsim_piso_config -> PisoCfg (src/physics/piso/main.cpp:83-105).
PisoCfg toPisoCfg(const bsim_piso_config& cfg);
// EquationPerformanceCounters of one persistent system (petsc_backend.cpp:236-286 accounting).
bsim_equation_counters equationCounters(const DeviceSystem& system, const char* identity);
}  // namespace bsim
