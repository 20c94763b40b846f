// internal.hpp -- interface between the solver (solver.hip) and the served
// form (server.hip) inside libminiaero_amd; not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>

#include "../../include/miniaero.h"

namespace ma {

struct PlannedCase;

// Host-side plan of one case (layout and the exact device bytes it needs).
ma_status plan_case(const ma_mesh_data* md, const ma_options* opt, PlannedCase** out, size_t* bytes);
void destroy_plan(PlannedCase* pc);
// Builds the planned case inside [region, region + region_bytes) on its own
// stream (no device allocation); the plan's arrays are consumed.
ma_status create_from_plan(PlannedCase* pc, void* region, size_t region_bytes, ma_solver** out);
// ma_solver_step without the argument checks.
ma_status solver_advance(ma_solver* s, int nsteps);
hipStream_t solver_stream(const ma_solver* s);

}  // namespace ma
