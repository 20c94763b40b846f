// miniaero.hpp -- C++ host facade over the C ABI (include/miniaero.h).
//
// Mirrors the reference's host-side classes so a caller of the reference can
// switch with the same sequence of calls (Main.C:96-141):
//
//   miniaero::Options opts;             opts.read_options_file();        // Options.h:73
//   miniaero::Parallel3DMesh mesh(nx, ny, nz, lx, ly, lz, problem_type, angle);
//   miniaero::MeshData mesh_data;        mesh.fillMeshData(mesh_data);   // Parallel3DMesh.h:173
//   miniaero::TimeSolverExplicitRK4 solver(mesh_data, opts);             // TimeSolverExplicitRK4.h:164
//   solver.Solve();                                                      // TimeSolverExplicitRK4.h:166
//
// The facade owns no numerics: every step runs in the HIP kernels of
// libminiaero_amd.so.  Failures throw miniaero::Error (the reference has no
// error reporting; it prints and continues).
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/miniaero.h"

namespace miniaero {

struct Error : std::runtime_error {
  ma_status status;
  Error(ma_status s, const std::string& what) : std::runtime_error(what), status(s) {}
};

inline void check(ma_status s) {
  if (s != MA_OK) throw Error(s, std::string("miniaero: ") + ma_last_error());
}

// Options.h:47-101 with the constructor defaults of Options.h:59-70; the
// extension fields of ma_options start zeroed (reference behaviour).
struct Options : ma_options {
  Options() {
    std::memset(static_cast<ma_options*>(this), 0, sizeof(ma_options));
    problem_type = 0;
    nx = ny = nz = 10;
    ntimesteps = 1;
    dt = 5e-8;
    output_results = 0;
    output_frequency = 10;
    second_order_space = 0;
  }
  void read_options_file(const char* path = "miniaero.inp") {
    ma_options o;
    check(ma_options_read(path, &o));
    // keep the extension block, replace the nine reference values
    problem_type = o.problem_type;
    lx = o.lx, ly = o.ly, lz = o.lz, angle = o.angle;
    nx = o.nx, ny = o.ny, nz = o.nz;
    ntimesteps = o.ntimesteps;
    dt = o.dt;
    output_results = o.output_results;
    output_frequency = o.output_frequency;
    second_order_space = o.second_order_space;
    viscous = o.viscous;
  }
};

// MeshData<Device> for one rank: a view (ma_mesh_data) plus, for generated
// meshes, ownership of the arrays it points to.
class MeshData {
 public:
  MeshData() { std::memset(&view_, 0, sizeof(view_)); }
  ~MeshData() { reset(); }
  MeshData(const MeshData&) = delete;
  MeshData& operator=(const MeshData&) = delete;

  // Adopts caller-owned arrays (they must outlive the solver's creation).
  void set_view(const ma_mesh_data& v) {
    reset();
    view_ = v;
  }
  const ma_mesh_data& view() const { return view_; }
  int num_owned_cells() const { return view_.num_owned_cells; }
  int num_ghosts() const { return view_.num_ghosts; }
  const double* cell_coordinates() const { return view_.cell_coordinates; }

 private:
  friend class Parallel3DMesh;
  void adopt(ma_mesh* m) {
    reset();
    owned_ = m;
    check(ma_mesh_view(m, &view_));
  }
  void reset() {
    if (owned_) ma_mesh_destroy(owned_);
    owned_ = nullptr;
    std::memset(&view_, 0, sizeof(view_));
  }
  ma_mesh* owned_ = nullptr;
  ma_mesh_data view_;
};

// Parallel3DMesh(nx, ny, nz, lx, ly, lz, problem_type, angle), one rank
// (Parallel3DMesh.h:56-59).  `ext` supplies the config-mode extensions
// (y_stretch, bc_kind) when given.
class Parallel3DMesh {
 public:
  Parallel3DMesh(int nx, int ny, int nz, double lx, double ly, double lz, int problem_type,
                 double angle = 0.0, const ma_options* ext = nullptr) {
    if (ext) static_cast<ma_options&>(opt_) = *ext;
    opt_.nx = nx, opt_.ny = ny, opt_.nz = nz;
    opt_.lx = lx, opt_.ly = ly, opt_.lz = lz;
    opt_.problem_type = problem_type;
    opt_.angle = angle;
  }
  void fillMeshData(MeshData& md) {
    ma_mesh* m = nullptr;
    check(ma_mesh_generate(&opt_, &m));
    md.adopt(m);
  }

 private:
  Options opt_;
};

// TimeSolverExplicitRK4<Device>(MeshData&, const Options&) + Solve()
// (TimeSolverExplicitRK4.h:160-539).
class TimeSolverExplicitRK4 {
 public:
  TimeSolverExplicitRK4(MeshData& mesh, const ma_options& options) : mesh_(mesh), opt_(options) {
    check(ma_solver_create(&mesh.view(), &opt_, &s_));
  }
  ~TimeSolverExplicitRK4() { ma_solver_destroy(s_); }
  TimeSolverExplicitRK4(const TimeSolverExplicitRK4&) = delete;
  TimeSolverExplicitRK4& operator=(const TimeSolverExplicitRK4&) = delete;

  // Initial condition, ntimesteps RK4 steps with the reference's progress
  // line every output_frequency steps, the device run time, and
  // results.<rank> when output_results is set.
  void Solve(int rank = 0) {
    const auto t0 = std::chrono::steady_clock::now();
    check(ma_solver_initialize(s_));
    const int freq = opt_.output_frequency > 0 ? opt_.output_frequency : opt_.ntimesteps + 1;
    int done = 0;
    while (done < opt_.ntimesteps) {
      const int n = std::min(freq - done % freq, opt_.ntimesteps - done);
      check(ma_solver_step(s_, n));
      done += n;
      if (done % freq == 0 && rank == 0) {
        double t = 0, dt = 0;
        check(ma_solver_get_time(s_, &t, &dt));
        std::fprintf(stdout, "\nTime Step #%i:  Time = %16.9e; dt = %16.9e\n", done, t, dt);
      }
    }
    check(ma_solver_synchronize(s_));
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rank == 0) std::fprintf(stdout, "\n ... Device Run time: %8.2f seconds ...\n", secs);
    if (opt_.output_results) write_results(rank);
  }

  void initialize() { check(ma_solver_initialize(s_)); }
  void step(int n) { check(ma_solver_step(s_, n)); }
  std::vector<double> state() {
    std::vector<double> U((size_t)mesh_.num_owned_cells() * 5);
    check(ma_solver_get_state(s_, U.data()));
    return U;
  }

  // results.<rank>: TimeSolverExplicitRK4.h:514-538, default ostream format.
  void write_results(int rank = 0) {
    const std::vector<double> U = state();
    std::stringstream fs;
    fs << "results." << rank;
    std::ofstream out(fs.str().c_str(), std::ios::out);
    const double* x = mesh_.cell_coordinates();
    for (int i = 0; i < mesh_.num_owned_cells(); ++i) {
      out << x[3 * i] << "\t" << x[3 * i + 1] << "\t" << x[3 * i + 2] << "\t";
      for (int k = 0; k < 5; ++k) out << U[(size_t)i * 5 + k] << "\t";
      out << std::endl;
    }
  }

  ma_solver* handle() { return s_; }

 private:
  MeshData& mesh_;
  ma_options opt_;
  ma_solver* s_ = nullptr;
};

// Independent cases advanced by the same launches (ma_batch_create); no
// reference counterpart (the reference runs one case per process).  Each
// case's result equals a TimeSolverExplicitRK4 run of it alone, bit for bit.
class BatchSolver {
 public:
  BatchSolver(const std::vector<MeshData*>& meshes, const std::vector<ma_options>& options) {
    if (meshes.size() != options.size() || meshes.empty()) throw Error(MA_ERR_INVALID, "BatchSolver: sizes");
    std::vector<const ma_mesh_data*> views;
    for (MeshData* m : meshes) {
      views.push_back(&m->view());
      cells_.push_back(m->num_owned_cells());
    }
    check(ma_batch_create((int)views.size(), views.data(), options.data(), &s_));
  }
  ~BatchSolver() { ma_solver_destroy(s_); }
  BatchSolver(const BatchSolver&) = delete;
  BatchSolver& operator=(const BatchSolver&) = delete;

  void initialize() { check(ma_solver_initialize(s_)); }
  void step(int n) { check(ma_solver_step(s_, n)); }
  void set_graph(int steps_per_graph) { check(ma_solver_set_graph(s_, steps_per_graph)); }
  void set_persistent(bool enabled) { check(ma_solver_set_persistent(s_, enabled ? 1 : 0)); }
  int cases() const { return (int)cells_.size(); }
  std::vector<double> case_state(int k) {
    std::vector<double> U((size_t)cells_.at(k) * 5);
    check(ma_batch_get_case_state(s_, k, U.data()));
    return U;
  }
  double case_time(int k) {
    double t = 0, dt = 0;
    check(ma_batch_get_case_time(s_, k, &t, &dt));
    return t;
  }
  ma_solver* handle() { return s_; }

 private:
  std::vector<int> cells_;
  ma_solver* s_ = nullptr;
};

// Request loop over independent cases (ma_server_*): submit options, poll
// finished cases, read their states.  No reference counterpart.
class Server {
 public:
  explicit Server(long long hbm_budget_bytes = 0, int max_active = 0, int setup_threads = 0,
                  int steps_in_flight = 0, long long plan_cache_bytes = 0) {
    ma_server_config cfg{hbm_budget_bytes, max_active, setup_threads, steps_in_flight, plan_cache_bytes};
    check(ma_server_create(&cfg, &s_));
  }
  ~Server() { ma_server_destroy(s_); }
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  long long submit(const ma_options& options) {
    long long id = 0;
    check(ma_server_submit(s_, &options, &id));
    return id;
  }
  std::vector<ma_case_result> poll(double timeout_s, int max_results = 64) {
    std::vector<ma_case_result> out((size_t)max_results);
    const int n = ma_server_poll(s_, out.data(), max_results, timeout_s);
    if (n < 0) throw Error(MA_ERR_INVALID, ma_last_error());
    out.resize((size_t)n);
    return out;
  }
  std::vector<double> state(const ma_case_result& r) {
    std::vector<double> U((size_t)r.cells * 5);
    check(ma_server_get_state(s_, r.id, U.data()));
    return U;
  }
  void release(long long id) { check(ma_server_release(s_, id)); }
  double now() { return ma_server_now(s_); }

 private:
  ma_server* s_ = nullptr;
};

}  // namespace miniaero
