// many_cases -- the batched and served forms through the C++ facade: every
// case's state equals a TimeSolverExplicitRK4 run of it alone, bit for bit.
// Exit status 0 when all agree, 1 on a mismatch, 2 on a library error.
#include <cstdio>
#include <cstring>
#include <map>
#include <vector>

#include "miniaero.hpp"

namespace {

miniaero::Options plate(int nx, int ny, int nz, int steps, double seed_like) {
  miniaero::Options o;
  o.problem_type = 1;
  o.lx = 1.0, o.ly = 0.05, o.lz = 0.025;
  o.nx = nx, o.ny = ny, o.nz = nz;
  o.ntimesteps = steps;
  o.dt = 0.0;
  o.cfl = 0.5;
  o.second_order_space = 1;
  o.viscous = 1;
  o.y_stretch = 1.05;
  const int bc[6] = {3, 2, 2, 2, 1, 4};
  for (int k = 0; k < 6; ++k) o.bc_kind[k] = bc[k];
  o.ic_kind = MA_IC_PERTURBED_INFLOW;
  o.perturb = 0.01;
  o.seed = (unsigned long long)(seed_like * 1000);
  return o;
}

std::vector<double> alone(const miniaero::Options& o) {
  miniaero::MeshData md;
  miniaero::Parallel3DMesh(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.problem_type, o.angle, &o).fillMeshData(md);
  miniaero::TimeSolverExplicitRK4 s(md, o);
  s.initialize();
  s.step(o.ntimesteps);
  return s.state();
}

bool same(const std::vector<double>& a, const std::vector<double>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
}

}  // namespace

int main() {
  try {
    const std::vector<miniaero::Options> cases = {plate(12, 8, 4, 7, 1), plate(20, 10, 6, 7, 2),
                                                  plate(9, 14, 5, 7, 3)};
    // batched
    std::vector<miniaero::MeshData> meshes(cases.size());
    std::vector<miniaero::MeshData*> mp;
    std::vector<ma_options> opts;
    for (size_t k = 0; k < cases.size(); ++k) {
      const miniaero::Options& o = cases[k];
      miniaero::Parallel3DMesh(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.problem_type, o.angle, &o)
          .fillMeshData(meshes[k]);
      mp.push_back(&meshes[k]);
      opts.push_back(o);
    }
    miniaero::BatchSolver b(mp, opts);
    b.initialize();
    b.step(7);
    int bad = 0;
    for (int k = 0; k < b.cases(); ++k) {
      const bool ok = same(b.case_state(k), alone(cases[k]));
      std::printf("batched case %d: %s\n", k, ok ? "bit-identical" : "DIFFERS");
      bad += !ok;
    }
    // served
    miniaero::Server srv(1ll << 30, 2, 2, 0);
    std::map<long long, const miniaero::Options*> ids;
    for (const auto& o : cases) ids[srv.submit(o)] = &o;
    size_t got = 0;
    for (int tries = 0; got < ids.size() && tries < 600; ++tries)
      for (const ma_case_result& r : srv.poll(0.1)) {
        ++got;
        const bool ok = r.status == MA_OK && same(srv.state(r), alone(*ids.at(r.id)));
        std::printf("served case %lld (%lld cells): %s\n", r.id, r.cells, ok ? "bit-identical" : "DIFFERS");
        bad += !ok;
        srv.release(r.id);
      }
    if (got != ids.size()) {
      std::printf("served: only %zu of %zu cases finished\n", got, ids.size());
      return 1;
    }
    std::printf(bad ? "FAILED\n" : "OK\n");
    return bad ? 1 : 0;
  } catch (const miniaero::Error& e) {
    std::fprintf(stderr, "%s\n", e.what());
    return 2;
  }
}
