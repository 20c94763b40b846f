/*
 * miniaero.h -- C ABI of the MI355X-native miniAero explicit RK4 residual path.
 *
 * The library (miniaero_amd/libminiaero_amd.so) replaces the data-parallel hot
 * path that TimeSolverExplicitRK4<Device>::Solve drives in the reference
 * (kokkos/TimeSolverExplicitRK4.h:160-539): face fluxes (Roe_Flux.h,
 * Viscous_Flux.h), the face-to-cell residual (Flux.h), Green-Gauss gradients
 * (GreenGauss.h), stencil min/max and limiters (StencilLimiter.h,
 * VenkatLimiter.h, VanAlbadaLimiter.h), boundary-face fluxes (Extrapolate_BC.h,
 * Tangent_BC.h, Inflow_BC.h, NoSlip_BC.h) and the RK4 stage update.  Its host
 * side keeps the reference's inputs: the nine-value miniaero.inp options
 * (Options.h:47-101) and the MeshData face/cell lists that
 * Parallel3DMesh::fillMeshData produces (MeshData.h:42-58, Faces.h:42-71,
 * Cells.h:41-76).
 *
 * Conventions: plain pointers and sizes, no C++ types.  Every entry point
 * returns ma_status (0 = success); on failure ma_last_error() describes the
 * cause (thread-local).  Host arrays are row-major, as the reference's
 * LayoutRight views: a (n x 3) array holds x0 y0 z0 x1 y1 z1 ...
 * A solver handle is not thread-safe; distinct handles are.
 */
#ifndef MINIAERO_H
#define MINIAERO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef int ma_status;
enum {
  MA_OK = 0,
  MA_ERR_INVALID = 1,      /* bad argument or inconsistent mesh */
  MA_ERR_UNSUPPORTED = 2,  /* valid input outside what this build handles (e.g. ghosts) */
  MA_ERR_NOMEM = 3,        /* host or device allocation failed */
  MA_ERR_DEVICE = 4,       /* HIP runtime error, or no GPU */
  MA_ERR_NONFINITE = 5     /* NaN/Inf found by ma_solver_check_finite */
};

/* Boundary-condition kinds: the names of MeshData::boundary_faces
 * (Parallel3DMesh.h:382-396).  Their numeric order is the order in which the
 * reference applies the boundary fluxes (TimeSolverExplicitRK4.h:428-475). */
enum { MA_BC_EXTRAPOLATE = 1, MA_BC_TANGENT = 2, MA_BC_NOSLIP = 3, MA_BC_INFLOW = 4 };

/* Side order of the generated box's boundary groups (Parallel3DMesh.h:382-396). */
enum { MA_SIDE_BOTTOM = 0, MA_SIDE_TOP, MA_SIDE_FRONT, MA_SIDE_BACK, MA_SIDE_RIGHT, MA_SIDE_LEFT };

enum { MA_LIMITER_VENKAT = 0, MA_LIMITER_VANALBADA = 1 };
enum { MA_IC_REFERENCE = 0, MA_IC_SOD = 1, MA_IC_PERTURBED_INFLOW = 2 };
/* Arithmetic of fluxes, gradients and limiters.  The conserved state, the RK4
 * accumulation and the residual sums are FP64 in both. */
enum { MA_PRECISION_FP64 = 0, MA_PRECISION_FP32 = 1 };

/* Options.h:47-57 plus config-mode extensions.  A zeroed extension block
 * reproduces the reference exactly. */
typedef struct ma_options {
  int problem_type;          /* 0 Sod, 1 viscous flat plate, 2 inviscid ramp */
  double lx, ly, lz, angle;  /* domain size; ramp angle in degrees */
  int nx, ny, nz;            /* cells per direction */
  int ntimesteps;
  double dt;
  int output_results;
  int output_frequency;
  int second_order_space;
  int viscous;
  /* ---- extensions (no reference counterpart) ---- */
  int bc_kind[6];            /* per MA_SIDE_*: 0 = reference assignment, else MA_BC_* */
  double y_stretch;          /* geometric y-spacing ratio away from y=0; 0 or 1 = uniform */
  int limiter;               /* MA_LIMITER_* */
  int ic_kind;               /* MA_IC_* */
  double ic[5];              /* MA_IC_SOD: rhoL, pL, rhoR, pR, x0 */
  int use_inflow;            /* 1: inflow[] replaces the hard-coded inflow state */
  double inflow[5];          /* conserved inflow / freestream state */
  double prandtl;            /* 0 = 0.71 (GasModel.h:103) */
  double cfl;                /* > 0: dt = cfl * min_c vol_c / (|u| S_x + |v| S_y + |w| S_z + c sum_f |n_f|
                              * + viscous term), S_d = sum_f |n_f,d| over the cell's faces, per step */
  double perturb;            /* MA_IC_PERTURBED_INFLOW: relative amplitude */
  unsigned long long seed;   /* MA_IC_PERTURBED_INFLOW: generator key */
  int precision;             /* MA_PRECISION_* */
  /* Flux tiles: cells per workgroup as an i x j x k brick of a lattice mesh
   * (extents >= 1, product <= 256); all 0 = 4 x 8 x 8, or 4 x 4 x 4 when the
   * run has fewer cells than 256 per compute unit; tile_shape[0] < 0 =
   * 256 consecutive cells of the caller's order.  Meshes that are not a
   * lattice always use consecutive cells.  Results do not depend on it. */
  int tile_shape[3];
  int autotune;              /* 1: time candidate tile shapes and workgroups_per_cu at create, keep the fastest */
  int workgroups_per_cu;     /* flux kernel: resident workgroups per CU walking the tiles
                              * (0 = one workgroup per tile); results do not depend on it */
} ma_options;

/* One Faces<Device> (Faces.h:42-71), as host arrays. */
typedef struct ma_faces {
  int nfaces;
  const double* coordinates;    /* nfaces x 3, face centroid */
  const double* face_normal;    /* nfaces x 3, area-weighted normal (left -> right) */
  const double* face_tangent;   /* nfaces x 3 */
  const double* face_binormal;  /* nfaces x 3 */
  const int* face_cell_conn;    /* nfaces x 2: left cell, right cell (-1 on boundary faces) */
  const int* cell_flux_index;   /* nfaces x 2, local face slot in each cell; may be NULL */
  const int* permute_vector;    /* nfaces, processing order; NULL = stable sort by left cell */
} ma_faces;

/* One std::pair<std::string, Faces<Device>> of MeshData::boundary_faces. */
typedef struct ma_bc_group {
  int kind;                     /* MA_BC_* (the reference's string name) */
  ma_faces faces;
} ma_bc_group;

/* MeshData<Device> (MeshData.h:42-58) + Cells<Device> (Cells.h:41-76), one rank. */
typedef struct ma_mesh_data {
  int num_owned_cells;
  int num_ghosts;                   /* must be 0: one GPU, no halo */
  const double* cell_coordinates;   /* num_owned_cells x 3 */
  const double* cell_volumes;       /* num_owned_cells */
  ma_faces internal_faces;
  int num_boundary_groups;
  const ma_bc_group* boundary_faces;
} ma_mesh_data;

typedef struct ma_mesh ma_mesh;
typedef struct ma_solver ma_solver;

/* Human-readable description of the last failure on this thread. */
const char* ma_last_error(void);

/* Reads the nine values of miniaero.inp (Options::read_options_file,
 * Options.h:73-101) from `path` into *opt; extension fields are zeroed.
 * Unlike the reference, a missing or short file is an error. */
ma_status ma_options_read(const char* path, ma_options* opt);

/* ---- mesh (replaces Parallel3DMesh + fillMeshData, Parallel3DMesh.h:56-449,
 *      Parallel3DMesh.C:36-303, MeshProcessor.C:39-229, Face.C:37-98,
 *      ElementTopoHexa8.C:134-148) ---- */

/* Generates the reference's one-rank box mesh in closed form: identical
 * cells, faces, face order (including the std::random_shuffle of internal
 * faces, which re-seeds the C library rand() with 1) and boundary groups. */
ma_status ma_mesh_generate(const ma_options* opt, ma_mesh** out);
/* Fills a view of the generated mesh; pointers stay valid until destroy. */
ma_status ma_mesh_view(const ma_mesh* mesh, ma_mesh_data* view);
void ma_mesh_destroy(ma_mesh* mesh);

/* Pre-flight check of a mesh, host only (no GPU needed): validates the
 * arrays and builds the device layout ma_solver_create would use.  stats[0..2]
 * = lattice dimensions found from the face connectivity (0 when the mesh is
 * not a lattice), [3] = flux tiles, [4] = most internal faces in one tile,
 * [5] = tile face entries in total, [6] = 1 when every face's tangent and
 * binormal follow from its area vector as in Face.C (the kernels then rebuild
 * the face frame instead of reading it).  No reference counterpart. */
ma_status ma_mesh_plan(const ma_mesh_data* mesh, long long* stats);

/* ---- solver (replaces TimeSolverExplicitRK4<Device>, TimeSolverExplicitRK4.h:160-539) ---- */

/* TimeSolverExplicitRK4(MeshData&, const Options&) (TimeSolverExplicitRK4.h:164,179-196):
 * copies the mesh to the GPU and builds the per-cell face lists.  Picks the
 * current HIP device. */
ma_status ma_solver_create(const ma_mesh_data* mesh, const ma_options* opt, ma_solver** out);
/* Initial condition + copy (TimeSolverExplicitRK4.h:324-336), time = 0. */
ma_status ma_solver_initialize(ma_solver* s);
/* Replaces the state (num_owned_cells x 5 conserved values); residual = 0. */
ma_status ma_solver_set_state(ma_solver* s, const double* U);
/* Runs nsteps RK4 steps (the body of the time loop, TimeSolverExplicitRK4.h:340-490).
 * Asynchronous on the solver's stream. */
ma_status ma_solver_step(ma_solver* s, int nsteps);
/* Blocks until the solver's stream is idle. */
ma_status ma_solver_synchronize(ma_solver* s);
/* Replays of a captured HIP graph: with steps_per_graph > 0 (an even count),
 * ma_solver_step runs whole groups of steps_per_graph steps as one graph
 * launch (captured on first use) and the remainder as plain launches; 0 (the
 * default) launches every kernel separately.  Same results either way.
 * Graphs are captured and replayed only on the thread that created the
 * solver: a step call from another thread that would use one fails with
 * MA_ERR_UNSUPPORTED and runs nothing.  No reference counterpart. */
ma_status ma_solver_set_graph(ma_solver* s, int steps_per_graph);
/* Persistent steps: with enabled = 1, ma_solver_step runs all its steps, for
 * every case of the run, in one cooperative launch whose workgroups stay
 * resident and pass a grid-wide barrier between the phases that are separate
 * kernels otherwise (the time step, the gradients, the boundary faces, the
 * flux tiles of each stage); same results as the per-kernel launches.  The
 * call then waits for the device to be idle (every stream of the process)
 * before the launch, and for the launch.  Turns graphs off (ma_solver_set_graph
 * with a count > 0 turns this off).  MA_ERR_UNSUPPORTED when the device has
 * no cooperative launch; MA_ERR_DEVICE when a barrier wait ran over its 10 s
 * limit.  No reference counterpart: the reference launches each kernel of
 * the stage loop separately (TimeSolverExplicitRK4.h:352-486). */
ma_status ma_solver_set_persistent(ma_solver* s, int enabled);
/* Launch shape in use: tile shape (-1 -1 -1 for consecutive cells),
 * resident workgroups per CU (0 = one per tile) and, after autotuning, the
 * measured ms per step per million cells of that choice (else 0).  Any
 * argument may be NULL.  No reference counterpart. */
ma_status ma_solver_tuning(ma_solver* s, int* tile_shape, int* workgroups_per_cu, double* tuned_ms);
/* Copies the current solution (sol_n) to host, num_owned_cells x 5. */
ma_status ma_solver_get_state(ma_solver* s, double* U);
/* Residual of the last stage, num_owned_cells x 5. */
ma_status ma_solver_get_residual(ma_solver* s, double* R);
/* Gradients (n x 5 x 3) and limiters (n x 5) of the last stage; either may be NULL.
   Inviscid second order does not keep them while stepping: the call computes
   them from the last stage's primitives (same values) on the solver's stream. */
ma_status ma_solver_get_gradients(ma_solver* s, double* grad, double* lim);
/* Simulated time and the dt of the last step. */
ma_status ma_solver_get_time(ma_solver* s, double* time, double* last_dt);
/* L2 norms of the residual of the last RK stage, per case and conserved
 * variable: norms[5*k + v] = sqrt(sum over case k's cells of R_v^2), with
 * R = dt / vol * (sum of face fluxes) as in ma_solver_get_residual.  Summed on
 * the device in a fixed order (repeatable bit for bit).  No reference
 * counterpart (a convergence monitor). */
ma_status ma_solver_residual_norms(ma_solver* s, double* norms);
/* The same norms, bit for bit, for callers that cannot allocate device
 * memory or block: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.
 * `norms` is a DEVICE array of 5 * ncases doubles; the call queues two
 * kernels on the solver's stream and returns without waiting (the values are
 * there once the stream reaches them).  It allocates nothing: the per-tile
 * partial sums go to a static device array of at most 262,144 tiles (10 MiB
 * per GPU, held until the process ends); a solver with more flux tiles gets
 * MA_ERR_INVALID.  That array is shared by every call on the GPU: calls that
 * overlap on one GPU (other streams or threads) race, calls queued in order
 * on one stream do not. */
ma_status ma_solver_residual_norms_noalloc(ma_solver* s, double* norms);
/* Sets *nonfinite to the number of non-finite values in the state. */
ma_status ma_solver_check_finite(ma_solver* s, long long* nonfinite);
/* The HIP stream (hipStream_t) all of the solver's work is queued on. */
ma_status ma_solver_stream(ma_solver* s, void** stream);
void ma_solver_destroy(ma_solver* s);

/* ---- batched form: independent cases stepped together ---- */

/* One run over ncases independent meshes (the reference runs one mesh per
 * process, Main.C:96-141; no counterpart).  Case k has its own mesh, initial
 * state, inflow state and dt or CFL number; all cases must share
 * second_order_space, viscous, limiter, precision and prandtl.  Every step
 * advances every case in the same launches; the result of each case is bit
 * for bit that of ma_solver_create on its mesh and options alone.
 * ma_solver_* calls apply to the whole batch: states, residuals and
 * gradients are the cases' cells in order, ma_solver_get_time reports case 0. */
ma_status ma_batch_create(int ncases, const ma_mesh_data* const* meshes, const ma_options* opts,
                          ma_solver** out);
/* Number of cases and, if case_cell_offsets is not NULL, the ncases + 1
 * offsets of each case's first cell in the batch's cell numbering. */
ma_status ma_batch_sizes(ma_solver* s, int* ncases, int* case_cell_offsets);
/* ma_mesh_plan over the concatenated layout ma_batch_create would build
 * (lattice dimensions are 0 for more than one case).  No reference counterpart. */
ma_status ma_batch_plan(int ncases, const ma_mesh_data* const* meshes, long long* stats);
/* State of case k alone (its num_owned_cells x 5 conserved values). */
ma_status ma_batch_get_case_state(ma_solver* s, int k, double* U);
/* Time and last dt of case k (dt differs between CFL-driven cases). */
ma_status ma_batch_get_case_time(ma_solver* s, int k, double* time, double* last_dt);

/* ---- served form: a request loop over independent cases ---- */

/* A server keeps admitted cases' state on the device between steps and runs
 * them together while new requests arrive (continuous batching): the cases
 * of one scheme (order, viscous, limiter, precision, Prandtl) are advanced
 * by the same launches, one RK4 step at a time; a case joins at the next step
 * after its admission and leaves after its last one.  Requests are planned
 * on host threads (mesh generated from the options as Parallel3DMesh does,
 * layout -- reused for a mesh seen before -- and exact device size) and
 * admitted in submission order when a slot is free and their device bytes
 * fit in the server's region of HBM, which is allocated once at creation.  Each case runs opts->ntimesteps RK4 steps from its initial
 * state; its result equals a ma_solver_create run of the same case alone,
 * bit for bit.  No reference counterpart (the reference runs one case per
 * process, Main.C:96-141). */
typedef struct ma_server ma_server;

typedef struct {
  long long hbm_budget_bytes; /* device region for admitted cases; 0 = 200e9, capped by free HBM */
  int max_active;             /* cases on the device at once; 0 = 16, at most 64 */
  int setup_threads;          /* host threads planning requests; 0 = 4 */
  int steps_in_flight;        /* RK4 steps queued ahead per scheme group; 0 = 2 */
  long long plan_cache_bytes; /* host bytes of device layouts kept for reuse by requests of the same
                               * mesh (box, cells, stretch, boundary kinds, tile shape); 0 = 16 GiB,
                               * < 0 = none */
} ma_server_config;

typedef struct {
  long long id;
  int status;              /* ma_status of the case; MA_OK when it ran to the end */
  int steps;               /* steps run */
  long long cells;
  long long device_bytes;  /* bytes of the case in the server's region */
  double t_submit, t_admit, t_finish;  /* seconds on the server's clock (ma_server_now) */
} ma_case_result;

ma_status ma_server_create(const ma_server_config* cfg, ma_server** out);
/* Queues a case; *id identifies it in results.  Never blocks on the device. */
ma_status ma_server_submit(ma_server* srv, const ma_options* opts, long long* id);
/* Moves up to max finished cases (in completion order) into out, waiting up
 * to timeout_s seconds for the first; returns the count (-1 on error). */
int ma_server_poll(ma_server* srv, ma_case_result* out, int max, double timeout_s);
/* Final state of a finished case (cells x 5 conserved values, the generated
 * mesh's cell order); kept on the host until ma_server_release.  A case whose
 * final state holds non-finite values finishes with MA_ERR_NONFINITE; its
 * state is still copied out and the call returns MA_ERR_NONFINITE. */
ma_status ma_server_get_state(ma_server* srv, long long id, double* U);
ma_status ma_server_release(ma_server* srv, long long id);
/* stats[0] submitted, [1] waiting (planning or for admission), [2] active,
 * [3] finished, [4] bytes of the region in use, [5] region size. */
ma_status ma_server_stats(ma_server* srv, long long* stats);
/* Seconds since the server was created. */
double ma_server_now(ma_server* srv);
/* Requests whose layout came from the plan cache (*hits) and those that
 * generated and planned their mesh (*misses). */
ma_status ma_server_plan_cache(ma_server* srv, long long* hits, long long* misses);
/* Cancels cases not yet admitted, waits for the active ones to finish their
 * launched steps, frees everything. */
void ma_server_destroy(ma_server* srv);

/* ---- measurement ---- */

/* Per-kernel timing: runs nsteps with a HIP event pair around every launch on
 * the solver's stream and reports, for each kernel name, the number of
 * launches and the mean duration in milliseconds.  Returns the number of
 * kernel names written (at most max). names[i] points to static strings. */
int ma_solver_profile(ma_solver* s, int nsteps, const char** names, int* launches,
                      double* mean_ms, int max);
/* Sizes: counts[0] cells, [1] internal faces, [2] boundary faces, [3] device bytes held. */
ma_status ma_solver_sizes(ma_solver* s, long long* counts);
/* Streaming-copy ceiling of the current GPU: copies `bytes` bytes `reps` times
 * with a plain copy kernel and reports the mean time per copy (it reads and
 * writes `bytes` each).  Used to calibrate HBM counters and roofline numbers;
 * no reference counterpart. */
ma_status ma_measure_copy(long long bytes, int reps, double* ms_per_copy);

#ifdef __cplusplus
}
#endif
#endif
