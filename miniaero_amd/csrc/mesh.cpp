/*
 * mesh.cpp -- host-side generator of miniAero's one-rank box mesh.
 *
 * Replaces Parallel3DMesh::fillMeshData (Parallel3DMesh.h:173-449) and the
 * MeshProcessor free functions it calls (MeshProcessor.C:39-229).  The
 * reference finds each face by hashing its smallest node id and comparing
 * sorted node lists (std::list per face), which is O(N) with a large constant
 * and a lot of memory.  On a structured box the outcome is known in closed
 * form, so this generator writes it directly:
 *
 *   - cells are visited i, j, k with k fastest (Parallel3DMesh.C:80-96) and
 *     cell (i,j,k) creates local faces 1, 2 and 5 always, face 0 only when
 *     j == 0, face 3 only when i == 0 and face 4 only when k == 0 (the other
 *     three are created earlier by the lower-indexed neighbour); the creating
 *     cell is the left cell and its local face number is the flux index;
 *   - face nodes follow the local-face table of MeshProcessor.C:44 in the
 *     creating cell's node order, so centroid, area vector, tangent and
 *     binormal are evaluated with the operation order of Face.C:57-96;
 *   - internal faces keep creation order and are then permuted by
 *     std::random_shuffle from the C library's default rand() state
 *     (Parallel3DMesh.h:362), reproduced by a private srand(1)-seeded state
 *     so that generation is thread-safe;
 *   - the processing order (Faces.h:128-144) is a stable sort of the shuffled
 *     faces by left cell, done here as a counting sort;
 *   - boundary faces are grouped per side in cell order, in the group order
 *     of Parallel3DMesh.h:382-396 (bottom, top, front, back, right, left).
 *
 * Cell volume uses the 2x2x2 Gauss rule with the reference's truncated Gauss
 * points (ElementTopoHexa8.h:35-43, ElementTopoHexa8.C:62-148) and the
 * centroid is the node mean (MeshProcessor.C:151-171), so every array is bit
 * for bit what the reference builds.  Compile with -ffp-contract=off.
 *
 * Extensions: y_stretch (geometric spacing toward y = 0) and bc_kind (per-side
 * boundary kinds) change coordinates and group kinds only, never the topology
 * or the order.
 */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <new>
#include <string>
#include <vector>

#include "../../include/miniaero.h"
#include "common.hpp"

namespace ma {
std::string& last_error() {
  static thread_local std::string msg;
  return msg;
}
}  // namespace ma

namespace {

struct FaceSet {
  int n = 0;
  std::vector<double> xf, a, t, b;
  std::vector<int> conn, fidx, perm;
  void resize(int nf) {
    n = nf;
    xf.resize((size_t)nf * 3);
    a.resize((size_t)nf * 3);
    t.resize((size_t)nf * 3);
    b.resize((size_t)nf * 3);
    conn.resize((size_t)nf * 2);
    fidx.resize((size_t)nf * 2);
    perm.resize(nf);
  }
};

// Local face table of MeshProcessor.C:44 (hex8 node numbering of
// Parallel3DMesh.C:86-93): the same table is needed to reproduce the node
// order, hence the operation order, of every face's geometry.
const int kFaceNodes[6][4] = {{0, 1, 5, 4}, {1, 2, 6, 5}, {2, 3, 7, 6},
                              {3, 0, 4, 7}, {0, 3, 2, 1}, {4, 5, 6, 7}};

}  // namespace

struct ma_mesh {
  ma_options opt;
  int nx, ny, nz, ncells;
  std::vector<double> node;  // (nx+1)(ny+1)(nz+1) x 3
  std::vector<double> cell_xyz, cell_vol;
  FaceSet internal;
  FaceSet groups[6];
  ma_bc_group group_view[6];

  int node_id(int i, int j, int k) const { return i * ((ny + 1) * (nz + 1)) + j * (nz + 1) + k; }

  void cell_nodes(int i, int j, int k, int* ids) const {
    ids[0] = node_id(i, j, k);
    ids[1] = node_id(i + 1, j, k);
    ids[2] = node_id(i + 1, j + 1, k);
    ids[3] = node_id(i, j + 1, k);
    ids[4] = node_id(i, j, k + 1);
    ids[5] = node_id(i + 1, j, k + 1);
    ids[6] = node_id(i + 1, j + 1, k + 1);
    ids[7] = node_id(i, j + 1, k + 1);
  }

  // Geometry of local face q of cell c, written at slot f of `F`.
  void face_geometry(int c, int q, FaceSet& F, int f) const {
    const int i = c / (ny * nz), j = (c / nz) % ny, k = c % nz;
    int en[8];
    cell_nodes(i, j, k, en);
    double p[4][3];
    for (int m = 0; m < 4; ++m)
      for (int d = 0; d < 3; ++d) p[m][d] = node[(size_t)en[kFaceNodes[q][m]] * 3 + d];
    double* xf = &F.xf[(size_t)f * 3];
    double* a = &F.a[(size_t)f * 3];
    double* t = &F.t[(size_t)f * 3];
    double* b = &F.b[(size_t)f * 3];
    // centroid: running sum of the four nodes, then scale by 1/4
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (int m = 0; m < 4; ++m) {
      s0 += p[m][0];
      s1 += p[m][1];
      s2 += p[m][2];
    }
    const double quarter = 1.0 / 4;
    xf[0] = s0 * quarter;
    xf[1] = s1 * quarter;
    xf[2] = s2 * quarter;
    double e1[3], e2[3], e3[3];
    for (int d = 0; d < 3; ++d) {
      e1[d] = p[1][d] - p[0][d];
      e2[d] = p[2][d] - p[0][d];
      e3[d] = p[3][d] - p[0][d];
    }
    // area vector: mean of the cross products of the two triangles' edges
    double c1[3], c2[3];
    c1[0] = e1[1] * e2[2] - e2[1] * e1[2];
    c1[1] = -e1[0] * e2[2] + e2[0] * e1[2];
    c1[2] = e1[0] * e2[1] - e2[0] * e1[1];
    c2[0] = e2[1] * e3[2] - e3[1] * e2[2];
    c2[1] = -e2[0] * e3[2] + e3[0] * e2[2];
    c2[2] = e2[0] * e3[1] - e3[0] * e2[1];
    for (int d = 0; d < 3; ++d) a[d] = 0.5 * (c1[d] + c2[d]);
    // tangent: zero in the component after the first largest |a_i| (i1),
    // rotated in the plane of a_i1 and the component two after it (i3)
    const double m0 = std::fabs(a[0]), m1 = std::fabs(a[1]), m2 = std::fabs(a[2]);
    int i1 = 0;
    if (m1 > m0) i1 = 1;
    if (m2 > (i1 == 0 ? m0 : m1)) i1 = 2;
    const int i2 = (i1 + 1) % 3, i3 = (i1 + 2) % 3;
    const double den = std::sqrt(a[i1] * a[i1] + a[i3] * a[i3]);
    t[i2] = 0.0;
    t[i1] = a[i3] / den;
    t[i3] = -a[i1] / den;
    b[0] = a[1] * t[2] - t[1] * a[2];
    b[1] = -a[0] * t[2] + t[0] * a[2];
    b[2] = a[0] * t[1] - t[0] * a[1];
  }
};

namespace {

void node_coordinate(const ma_options& o, int gi, int gj, int gk, double c[3]) {
  // Parallel3DMesh.h:466-487 (one rank: global index == local index).
  const double PI = 3.14159265;
  c[0] = (double)gi / (o.nx) * o.lx;
  c[2] = (double)gk / (o.nz) * o.lz;
  const bool stretched = o.y_stretch > 0.0 && o.y_stretch != 1.0;
  double eta = 0.0;
  if (stretched) eta = (std::pow(o.y_stretch, gj) - 1.0) / (std::pow(o.y_stretch, o.ny) - 1.0);
  if (c[0] < o.lx / 2.0) {
    c[1] = stretched ? eta * o.ly : (double)gj / (o.ny) * o.ly;
  } else {
    const double y_ramp = (c[0] - (o.lx / 2.0)) * std::tan(o.angle * PI / 180.0);
    const double ly_scaled = o.ly - y_ramp;
    c[1] = stretched ? y_ramp + eta * ly_scaled : y_ramp + (double)gj * ly_scaled / (o.ny);
  }
}

void hex_volume_centroid(const double* ex, const double* ey, const double* ez, double* vol,
                         double* cen) {
  // 2x2x2 Gauss rule of ElementTopoHexa8.C:134-148; dN/dxi etc. from the
  // trilinear shape functions, evaluated at the truncated points +-g.
  const double g = 0.577350269189626;
  static const signed char sx[8] = {-1, 1, 1, -1, -1, 1, 1, -1};
  static const signed char sy[8] = {-1, -1, 1, 1, -1, -1, 1, 1};
  static const signed char sz[8] = {-1, -1, -1, -1, 1, 1, 1, 1};
  double volume = 0.0;
  for (int ig = 0; ig < 8; ++ig) {
    const double xi = sx[ig] * g, eta = sy[ig] * g, zeta = sz[ig] * g;
    double dxi[8], deta[8], dzeta[8];
    for (int n = 0; n < 8; ++n) {
      // N_n = 1/8 (1 + sx xi)(1 + sy eta)(1 + sz zeta); each derivative is
      // written as +-0.125 * (1 +- a) * (1 +- b), the reference's factor order.
      const double fe = sy[n] > 0 ? 1.0 + eta : 1.0 - eta;
      const double fz = sz[n] > 0 ? 1.0 + zeta : 1.0 - zeta;
      const double fx = sx[n] > 0 ? 1.0 + xi : 1.0 - xi;
      dxi[n] = (sx[n] > 0 ? 0.125 : -0.125) * fe * fz;
      deta[n] = (sy[n] > 0 ? 0.125 : -0.125) * fx * fz;
      dzeta[n] = (sz[n] > 0 ? 0.125 : -0.125) * fx * fe;
    }
    double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 8; ++n) {
      J[0] += dxi[n] * ex[n];
      J[1] += dxi[n] * ey[n];
      J[2] += dxi[n] * ez[n];
      J[3] += deta[n] * ex[n];
      J[4] += deta[n] * ey[n];
      J[5] += deta[n] * ez[n];
      J[6] += dzeta[n] * ex[n];
      J[7] += dzeta[n] * ey[n];
      J[8] += dzeta[n] * ez[n];
    }
    const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) + J[1] * (J[5] * J[6] - J[3] * J[8]) +
                       J[2] * (J[3] * J[7] - J[4] * J[6]);
    volume += det;
  }
  *vol = volume;
  double s0 = 0, s1 = 0, s2 = 0;
  for (int n = 0; n < 8; ++n) {
    s0 += ex[n];
    s1 += ey[n];
    s2 += ez[n];
  }
  cen[0] = s0 / 8;
  cen[1] = s1 / 8;
  cen[2] = s2 / 8;
}

}  // namespace

extern "C" ma_status ma_options_read(const char* path, ma_options* opt) {
  if (!path || !opt) return ma::fail(MA_ERR_INVALID, "ma_options_read: null argument");
  std::ifstream f(path);
  if (!f) return ma::fail(MA_ERR_INVALID, std::string("miniaero.inp not found: ") + path);
  ma_options o;
  std::memset(&o, 0, sizeof(o));
  // Options.h:91-99, the same nine whitespace-separated values.
  f >> o.problem_type;
  f >> o.lx >> o.ly >> o.lz >> o.angle;
  f >> o.nx >> o.ny >> o.nz;
  f >> o.ntimesteps;
  f >> o.dt;
  f >> o.output_results;
  f >> o.output_frequency;
  f >> o.second_order_space;
  f >> o.viscous;
  if (!f) return ma::fail(MA_ERR_INVALID, std::string("miniaero.inp is short or malformed: ") + path);
  *opt = o;
  return MA_OK;
}

extern "C" ma_status ma_mesh_generate(const ma_options* o, ma_mesh** out) {
  if (!o || !out) return ma::fail(MA_ERR_INVALID, "ma_mesh_generate: null argument");
  if (o->nx < 1 || o->ny < 1 || o->nz < 1)
    return ma::fail(MA_ERR_INVALID, "ma_mesh_generate: nx, ny, nz must be >= 1");
  if ((double)o->nx * o->ny * o->nz > 2.0e8)
    return ma::fail(MA_ERR_UNSUPPORTED, "ma_mesh_generate: more than 2e8 cells");
  for (int s = 0; s < 6; ++s)
    if (o->bc_kind[s] < 0 || o->bc_kind[s] > MA_BC_INFLOW)
      return ma::fail(MA_ERR_INVALID, "ma_mesh_generate: bc_kind out of range");
  ma_mesh* m = new (std::nothrow) ma_mesh();
  if (!m) return ma::fail(MA_ERR_NOMEM, "ma_mesh_generate: out of host memory");
  try {
    m->opt = *o;
    const int nx = o->nx, ny = o->ny, nz = o->nz;
    m->nx = nx;
    m->ny = ny;
    m->nz = nz;
    const int nc = nx * ny * nz;
    m->ncells = nc;
    m->node.resize((size_t)(nx + 1) * (ny + 1) * (nz + 1) * 3);
    for (int i = 0; i <= nx; ++i)
      for (int j = 0; j <= ny; ++j)
        for (int k = 0; k <= nz; ++k) node_coordinate(*o, i, j, k, &m->node[(size_t)m->node_id(i, j, k) * 3]);

    m->cell_xyz.resize((size_t)nc * 3);
    m->cell_vol.resize(nc);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const int c = (i * ny + j) * nz + k;
          int en[8];
          m->cell_nodes(i, j, k, en);
          double ex[8], ey[8], ez[8];
          for (int n = 0; n < 8; ++n) {
            ex[n] = m->node[(size_t)en[n] * 3];
            ey[n] = m->node[(size_t)en[n] * 3 + 1];
            ez[n] = m->node[(size_t)en[n] * 3 + 2];
          }
          hex_volume_centroid(ex, ey, ez, &m->cell_vol[c], &m->cell_xyz[(size_t)c * 3]);
        }

    // Internal faces in creation order: per cell, local face 1 (x+), 2 (y+),
    // 5 (z+) when the neighbour exists.  Key = cell * 8 + local face.
    std::vector<long long> keys;
    keys.reserve((size_t)nc * 3);
    for (int i = 0; i < nx; ++i)
      for (int j = 0; j < ny; ++j)
        for (int k = 0; k < nz; ++k) {
          const long long c = (i * ny + j) * nz + k;
          if (i + 1 < nx) keys.push_back(c * 8 + 1);
          if (j + 1 < ny) keys.push_back(c * 8 + 2);
          if (k + 1 < nz) keys.push_back(c * 8 + 5);
        }
    // The C library's rand() after srand(1), drawn from a private state of
    // the same generator (random_r over a 128-byte state, as rand() uses) so
    // that meshes can be generated on several threads at once; rand() is
    // random(), and random_shuffle(first, last) draws rand() % (i + 1)
    // exactly as the functor form does with this functor.
    struct random_data rd;
    std::memset(&rd, 0, sizeof rd);
    char rstate[128];
    std::memset(rstate, 0, sizeof rstate);
    if (initstate_r(1u, rstate, sizeof rstate, &rd) != 0 || srandom_r(1u, &rd) != 0) {
      delete m;
      return ma::fail(MA_ERR_INVALID, "mesh generator: random state initialisation failed");
    }
    auto draw = [&rd](long long n) {
      int32_t r = 0;
      random_r(&rd, &r);
      return (long long)(r % n);
    };
    std::random_shuffle(keys.begin(), keys.end(), draw);
    const int nf = (int)keys.size();
    FaceSet& F = m->internal;
    F.resize(nf);
    for (int f = 0; f < nf; ++f) {
      const int c = (int)(keys[f] >> 3), q = (int)(keys[f] & 7);
      int right = 0, fq = 0;
      if (q == 1) { right = c + ny * nz; fq = 3; }
      else if (q == 2) { right = c + nz; fq = 0; }
      else { right = c + 1; fq = 4; }
      F.conn[2 * f] = c;
      F.conn[2 * f + 1] = right;
      F.fidx[2 * f] = q;
      F.fidx[2 * f + 1] = fq;
      m->face_geometry(c, q, F, f);
    }
    // Stable counting sort by left cell (serial BinSort, Faces.h:128-144).
    {
      std::vector<int> start((size_t)nc + 1, 0);
      for (int f = 0; f < nf; ++f) start[F.conn[2 * f] + 1]++;
      for (int c = 0; c < nc; ++c) start[c + 1] += start[c];
      for (int f = 0; f < nf; ++f) F.perm[start[F.conn[2 * f]]++] = f;
    }

    // Boundary groups, each in cell order.
    const int pt = o->problem_type;
    const int ref_kind[6] = {pt == 1 ? MA_BC_NOSLIP : MA_BC_TANGENT,
                             pt == 1 ? MA_BC_EXTRAPOLATE : MA_BC_TANGENT,
                             MA_BC_TANGENT, MA_BC_TANGENT, MA_BC_EXTRAPOLATE,
                             pt == 0 ? MA_BC_EXTRAPOLATE : MA_BC_INFLOW};
    const int side_q[6] = {0, 2, 4, 5, 1, 3};
    for (int s = 0; s < 6; ++s) {
      std::vector<int> cells;
      for (int i = 0; i < nx; ++i)
        for (int j = 0; j < ny; ++j)
          for (int k = 0; k < nz; ++k) {
            const bool on = (s == 0 && j == 0) || (s == 1 && j == ny - 1) || (s == 2 && k == 0) ||
                            (s == 3 && k == nz - 1) || (s == 4 && i == nx - 1) || (s == 5 && i == 0);
            if (on) cells.push_back((i * ny + j) * nz + k);
          }
      FaceSet& G = m->groups[s];
      G.resize((int)cells.size());
      for (int f = 0; f < G.n; ++f) {
        G.conn[2 * f] = cells[f];
        G.conn[2 * f + 1] = -1;
        G.fidx[2 * f] = side_q[s];
        G.fidx[2 * f + 1] = -1;
        G.perm[f] = f;
        m->face_geometry(cells[f], side_q[s], G, f);
      }
      ma_bc_group& v = m->group_view[s];
      v.kind = o->bc_kind[s] ? o->bc_kind[s] : ref_kind[s];
    }
  } catch (const std::bad_alloc&) {
    delete m;
    return ma::fail(MA_ERR_NOMEM, "ma_mesh_generate: out of host memory");
  }
  *out = m;
  return MA_OK;
}

namespace {
void view_of(const FaceSet& F, ma_faces* v) {
  v->nfaces = F.n;
  v->coordinates = F.xf.data();
  v->face_normal = F.a.data();
  v->face_tangent = F.t.data();
  v->face_binormal = F.b.data();
  v->face_cell_conn = F.conn.data();
  v->cell_flux_index = F.fidx.data();
  v->permute_vector = F.perm.data();
}
}  // namespace

extern "C" ma_status ma_mesh_view(const ma_mesh* cm, ma_mesh_data* v) {
  if (!cm || !v) return ma::fail(MA_ERR_INVALID, "ma_mesh_view: null argument");
  ma_mesh* m = const_cast<ma_mesh*>(cm);
  v->num_owned_cells = m->ncells;
  v->num_ghosts = 0;
  v->cell_coordinates = m->cell_xyz.data();
  v->cell_volumes = m->cell_vol.data();
  view_of(m->internal, &v->internal_faces);
  for (int s = 0; s < 6; ++s) view_of(m->groups[s], &m->group_view[s].faces);
  v->num_boundary_groups = 6;
  v->boundary_faces = m->group_view;
  return MA_OK;
}

extern "C" void ma_mesh_destroy(ma_mesh* m) { delete m; }
