// main.cpp -- the `miniaero` driver: the reference's main()/run() sequence
// (Main.C:53-151) over the MI355X library.  Reads miniaero.inp from the
// working directory, builds the mesh, runs the RK4 solve on the current GPU
// and writes results.0 when output_results is set.
#include <chrono>
#include <cstdio>

#include "miniaero.hpp"

int main(int argc, char** argv) {
  const char* input = argc > 1 ? argv[1] : "miniaero.inp";
  try {
    const auto t0 = std::chrono::steady_clock::now();
    miniaero::Options opts;
    opts.read_options_file(input);

    miniaero::Parallel3DMesh mesh(opts.nx, opts.ny, opts.nz, opts.lx, opts.ly, opts.lz,
                                  opts.problem_type, opts.angle);
    miniaero::MeshData mesh_data;
    mesh.fillMeshData(mesh_data);
    const double setup =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stdout, "\n ... Setup time: %8.2f seconds ...\n", setup);

    {
      miniaero::TimeSolverExplicitRK4 solver(mesh_data, opts);
      solver.Solve();
    }
    const double total =
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    std::fprintf(stdout, "\n ... Total elapsed time: %8.2f seconds ...\n", total);
  } catch (const miniaero::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 1;
  }
  return 0;
}
