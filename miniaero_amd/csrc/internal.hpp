// internal.hpp -- interface between the solver (solver.hip) and the served
// form (server.hip) inside libminiaero_amd; not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>

#include "../../include/miniaero.h"

namespace ma {

struct PlannedCase;
struct LayoutPlan;

// Host-side plan of one case (layout and the exact device bytes it needs).
ma_status plan_case(const ma_mesh_data* md, const ma_options* opt, PlannedCase** out, size_t* bytes);
// The same in two parts: the device layout of a mesh (depends on the mesh and
// the tile shape only, so it can serve every case of that mesh; *cells and
// *host_bytes describe it), then a case's plan on that layout with its own
// options and exact device bytes.
ma_status plan_layout_shared(const ma_mesh_data* md, const ma_options* opt, std::shared_ptr<const LayoutPlan>* out,
                             long long* cells, size_t* host_bytes);
ma_status plan_case_with(const std::shared_ptr<const LayoutPlan>& lp, const ma_options* opt, PlannedCase** out,
                         size_t* bytes);
void destroy_plan(PlannedCase* pc);
// Builds the planned case inside [region, region + region_bytes) on its own
// stream (no device allocation); the plan is only read.
ma_status create_from_plan(PlannedCase* pc, void* region, size_t region_bytes, ma_solver** out);
// Which of the case's two state buffers holds sol_n (0 or 1).
int solver_cur(const ma_solver* s);
// Records that `steps` steps ran outside the solver's own launches (in a
// Multi): flips the state buffer for an odd count, adds to the step count.
void solver_set_progress(ma_solver* s, int steps);

// Several single-case solvers of one scheme (second order, viscous, limiter,
// precision, Prandtl) advanced by the same launches on `stream`.
struct Multi;
ma_status multi_create(const ma_options& scheme, int capacity, hipStream_t stream, Multi** out);
// Sets the cases of the following steps (uploaded on the stream, after the
// steps already queued).  Case k's state buffer for sol_n at a step of
// parity p is parity[k] ^ p.
ma_status multi_set(Multi* M, ma_solver* const* cases, const int* parity, int n);
// One RK4 step of every case (asynchronous on the stream).
ma_status multi_step(Multi* M, int parity);
void multi_destroy(Multi* M);

}  // namespace ma
