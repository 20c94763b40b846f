/*
 * oracle.h -- C interface of the CPU oracle for the miniAero explicit RK4 path.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a fixed-order, single-threaded CPU
 * restatement of the reference (Mantevo/miniAero, kokkos/) used as the checker
 * for the HIP product path.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it.
 *
 * The structs below are the oracle's own POD layout.  They intentionally carry
 * the same fields as the product's public options so that one Python dict can
 * drive both, but nothing here is shared with the product's headers.
 */
#ifndef MINIAERO_ORACLE_H
#define MINIAERO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Options.h:47-57 (9 input values) plus the config-mode extensions.  The zero
 * value of every extension field reproduces the reference behaviour. */
typedef struct orc_options {
  int problem_type;          /* 0 Sod, 1 flat plate, 2 ramp */
  double lx, ly, lz, angle;
  int nx, ny, nz;
  int ntimesteps;
  double dt;
  int second_order_space;
  int viscous;
  /* ---- extensions ---- */
  int bc_kind[6];            /* per side (bottom, top, front, back, right, left);
                                0 = reference assignment, 1 Extrapolate,
                                2 Tangent, 3 NoSlip, 4 Inflow */
  double y_stretch;          /* geometric ratio of y spacing away from y=0; 0 or 1 = uniform */
  int limiter;               /* 0 Venkat (reference), 1 Van Albada */
  int ic_kind;               /* 0 reference IC, 1 Sod-type, 2 inflow state + seeded perturbation */
  double ic[5];              /* ic_kind 1: rhoL, pL, rhoR, pR, x0 */
  int use_inflow;            /* 1: inflow[] replaces the reference inflow state */
  double inflow[5];          /* conserved inflow state */
  double prandtl;            /* 0 = 0.71 (GasModel.h:103) */
  double cfl;                /* > 0: dt = cfl * min_c vol_c / sum_f(|u.n| + c|n|) each step */
  double perturb;            /* ic_kind 2: relative amplitude */
  unsigned long long seed;   /* ic_kind 2: counter-based generator key */
} orc_options;

typedef struct orc_mesh orc_mesh;
typedef struct orc_solver orc_solver;

/* Literal restatement of Parallel3DMesh::fillMeshData for one rank. */
orc_mesh* orc_mesh_new(const orc_options* o);
void orc_mesh_free(orc_mesh* m);
/* counts[0] = ncells, counts[1] = n internal faces, counts[2] = n groups,
 * counts[3 + g] = faces in group g.  Returns the number of ints written. */
int orc_mesh_counts(const orc_mesh* m, int* counts, int max);
void orc_mesh_cells(const orc_mesh* m, double* xyz, double* vol);
/* Node coordinates ((nx+1)(ny+1)(nz+1) x 3), node (i, j, k) at row
 * (i (ny+1) + j)(nz+1) + k (Parallel3DMesh.h:453-460, one rank). */
void orc_mesh_nodes(const orc_mesh* m, double* xyz);
/* group = -1 for internal faces.  Any pointer may be NULL. kind gets the BC kind
 * (1..4) of the group, 0 for internal faces. */
void orc_mesh_faces(const orc_mesh* m, int group, double* xf, double* a, double* t,
                    double* b, int* conn, int* fidx, int* perm, int* kind);

/* TimeSolverExplicitRK4 restatement.  The solver keeps a pointer to the mesh. */
orc_solver* orc_solver_new(const orc_mesh* m, const orc_options* o);
void orc_solver_free(orc_solver* s);
void orc_solver_init(orc_solver* s);                 /* IC + copy (TimeSolverExplicitRK4.h:324-336) */
void orc_solver_set_state(orc_solver* s, const double* U);   /* (ncells x 5) row-major */
void orc_solver_step(orc_solver* s, int nsteps);
void orc_solver_get_state(const orc_solver* s, double* U);
void orc_solver_get_residual(const orc_solver* s, double* R);
/* Gradient (ncells x 5 x 3) and limiter (ncells x 5) of the last stage. */
void orc_solver_get_gradients(const orc_solver* s, double* grad, double* lim);
double orc_solver_time(const orc_solver* s);
double orc_solver_last_dt(const orc_solver* s);

/* Scalar pieces, exposed for unit checks against closed forms. */
void orc_primitives(const double* U, double* V);
void orc_roe_flux(const double* VL, const double* VR, const double* n, const double* t,
                  const double* b, double* F);
void orc_viscous_flux(const double* grad15, const double* V, const double* a, double prandtl,
                      double* F);
double orc_venkat(double dumax, double dumin, double du, double deltax3);
double orc_vanalbada(double dumax, double dumin, double du);
void orc_hex_volume_centroid(const double* ex, const double* ey, const double* ez,
                             double* vol, double* centroid);

#ifdef __cplusplus
}
#endif
#endif
