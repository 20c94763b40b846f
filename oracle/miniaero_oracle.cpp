/*
 * miniaero_oracle.cpp -- CPU oracle for the miniAero explicit RK4 residual path.
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  This file restates, in plain
 * single-threaded C++, what the reference computes when it runs serially:
 *
 *   mesh:   Parallel3DMesh.{h,C}, MeshProcessor.C, Face.C, ElementTopoHexa8.C
 *   solver: TimeSolverExplicitRK4.h, Flux.h, Roe_Flux.h, Viscous_Flux.h,
 *           GasModel.h, MathToolsDevice.h, GreenGauss.h, StencilLimiter.h,
 *           VenkatLimiter.h, VanAlbadaLimiter.h, Extrapolate_BC.h,
 *           Tangent_BC.h, Inflow_BC.h, NoSlip_BC.h, Initial_Conditions.h,
 *           Cells.h (zero_cell_flux), Faces.h (permutation)
 *
 * Fixed order.  The reference scatters with atomics (-DATOMICS_FLUX), so its
 * floating-point summation order is only defined for a one-thread run.  The
 * oracle uses that one-thread order: internal faces in the permuted order
 * (a stable sort of the shuffled faces by left cell, which is what the
 * serial Kokkos BinSort with sort_within_bins produces, Faces.h:128-144),
 * then the boundary groups, in the order each reference kernel walks them.
 * Build with -ffp-contract=off: the reference host build has no FMA.
 *
 * Pinning (what makes this oracle trustworthy, each checked in tests/):
 *   - mesh arrays: bit-exact against oracle/_ref, a build of the reference's
 *     own MeshProcessor.C / Face.C / ElementTopoHexa8.C sources;
 *   - end to end: the reference's six gold result sets (tests/golden/);
 *   - scalar kernels: closed forms (Roe consistency F(V,V) = physical flux,
 *     rotation invariance, supersonic upwinding; viscous Couette stress;
 *     Venkat/Van Albada closed forms; Green-Gauss exact for linear fields;
 *     uniform flow preserved exactly by the full RK4 step).
 * Config-mode extensions (stretching, per-side BCs, Van Albada in the stencil
 * limiter, CFL dt, custom ICs, Prandtl) have no reference counterpart and are
 * pinned by closed forms and invariants: tests/test_oracle_extensions.py
 * (stretched nodes, Sod-type and perturbed ICs, the generator against the
 * published SplitMix64 sequence, the Mach/Reynolds freestream) and
 * tests/test_oracle.py (CFL time step, Van Albada).
 */
#include "oracle.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cfloat>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <vector>

namespace {

const double kRgas = 287.05;   // GasModel.h:39
const double kGamma = 1.4;     // GasModel.h:51

enum { BC_EXTRAPOLATE = 1, BC_TANGENT = 2, BC_NOSLIP = 3, BC_INFLOW = 4 };

struct Faces {
  int n = 0;
  std::vector<double> xf, a, t, b;  // n x 3 each
  std::vector<int> conn, fidx;      // n x 2 each
  std::vector<int> perm;            // n
  std::vector<std::array<int, 4>> nodes;
};

}  // namespace

struct orc_mesh {
  orc_options opt;
  int nx, ny, nz, ncells;
  std::vector<double> node_xyz;
  std::vector<int> elem_nodes;
  std::vector<double> cell_xyz, cell_vol;
  Faces internal;
  std::vector<Faces> groups;
  std::vector<int> group_kind;
};

namespace {

/* ------------------------------------------------------------------ mesh -- */

// Parallel3DMesh.h:466-487 (getNodeCoordinate), one rank so global == local.
// The y_stretch extension replaces j/ny by a geometric fraction.
void node_coordinate(const orc_options& o, int gi, int gj, int gk, double c[3]) {
  const double PI = 3.14159265;
  c[0] = (double)gi / (o.nx) * o.lx;
  c[2] = (double)gk / (o.nz) * o.lz;
  const bool stretched = o.y_stretch > 0.0 && o.y_stretch != 1.0;
  double eta = 0.0;
  if (stretched) eta = (std::pow(o.y_stretch, gj) - 1.0) / (std::pow(o.y_stretch, o.ny) - 1.0);
  if (c[0] < o.lx / 2.0) {
    c[1] = stretched ? eta * o.ly : (double)gj / (o.ny) * o.ly;
  } else {
    double y_ramp = (c[0] - (o.lx / 2.0)) * std::tan(o.angle * PI / 180.0);
    double ly_scaled = o.ly - y_ramp;
    c[1] = stretched ? y_ramp + eta * ly_scaled : y_ramp + (double)gj * ly_scaled / (o.ny);
  }
}

// Parallel3DMesh.h:453-460 for one rank.
inline int node_id(const orc_mesh& m, int i, int j, int k) {
  return i * ((m.ny + 1) * (m.nz + 1)) + j * (m.nz + 1) + k;
}

// Face::Face geometry, Face.C:37-98 with MathTools.h:41-88.
void face_geometry(const std::vector<double>& X, const std::array<int, 4>& nd, double* xf,
                   double* a, double* t, double* b) {
  double p[4][3];
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 3; ++j) p[i][j] = X[nd[i] * 3 + j];
  xf[0] = xf[1] = xf[2] = 0.0;
  for (int i = 0; i < 4; ++i) {
    xf[0] += p[i][0];
    xf[1] += p[i][1];
    xf[2] += p[i][2];
  }
  const double s = 1.0 / 4;
  xf[0] *= s;
  xf[1] *= s;
  xf[2] *= s;
  const double v1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
  const double v2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
  const double v3[3] = {p[3][0] - p[0][0], p[3][1] - p[0][1], p[3][2] - p[0][2]};
  auto cross = [](const double* x, const double* y, double* z) {
    z[0] = x[1] * y[2] - y[1] * x[2];
    z[1] = -x[0] * y[2] + y[0] * x[2];
    z[2] = x[0] * y[1] - y[0] * x[1];
  };
  double n1[3], n2[3];
  cross(v1, v2, n1);
  cross(v2, v3, n2);
  for (int j = 0; j < 3; ++j) a[j] = 0.5 * (n1[j] + n2[j]);
  const double aa[3] = {std::fabs(a[0]), std::fabs(a[1]), std::fabs(a[2])};
  int i1 = (int)(std::max_element(aa, aa + 3) - aa);
  int i2 = i1 + 1, i3 = i1 + 2;
  i2 = (i2 > 2) ? i2 - 3 : i2;
  i3 = (i3 > 2) ? i3 - 3 : i3;
  const double denom = std::sqrt(a[i1] * a[i1] + a[i3] * a[i3]);
  t[i2] = 0.0;
  t[i1] = a[i3] / denom;
  t[i3] = -a[i1] / denom;
  cross(a, t, b);
}

}  // namespace

extern "C" void orc_hex_volume_centroid(const double* ex, const double* ey, const double* ez,
                                        double* vol, double* centroid) {
  // ElementTopoHexa8.C:62-148 with the truncated Gauss points of ElementTopoHexa8.h:35-43.
  const double g = 0.577350269189626;
  const double xi_[8] = {-g, g, g, -g, -g, g, g, -g};
  const double eta_[8] = {-g, -g, g, g, -g, -g, g, g};
  const double zeta_[8] = {-g, -g, -g, -g, g, g, g, g};
  double volume = 0.0;
  for (int ig = 0; ig < 8; ++ig) {
    const double xi = xi_[ig], eta = eta_[ig], zeta = zeta_[ig];
    double dx[8], de[8], dz[8];
    dx[0] = -0.125 * (1.0 - eta) * (1.0 - zeta);
    dx[1] = 0.125 * (1.0 - eta) * (1.0 - zeta);
    dx[2] = 0.125 * (1.0 + eta) * (1.0 - zeta);
    dx[3] = -0.125 * (1.0 + eta) * (1.0 - zeta);
    dx[4] = -0.125 * (1.0 - eta) * (1.0 + zeta);
    dx[5] = 0.125 * (1.0 - eta) * (1.0 + zeta);
    dx[6] = 0.125 * (1.0 + eta) * (1.0 + zeta);
    dx[7] = -0.125 * (1.0 + eta) * (1.0 + zeta);
    de[0] = -0.125 * (1.0 - xi) * (1.0 - zeta);
    de[1] = -0.125 * (1.0 + xi) * (1.0 - zeta);
    de[2] = 0.125 * (1.0 + xi) * (1.0 - zeta);
    de[3] = 0.125 * (1.0 - xi) * (1.0 - zeta);
    de[4] = -0.125 * (1.0 - xi) * (1.0 + zeta);
    de[5] = -0.125 * (1.0 + xi) * (1.0 + zeta);
    de[6] = 0.125 * (1.0 + xi) * (1.0 + zeta);
    de[7] = 0.125 * (1.0 - xi) * (1.0 + zeta);
    dz[0] = -0.125 * (1.0 - xi) * (1.0 - eta);
    dz[1] = -0.125 * (1.0 + xi) * (1.0 - eta);
    dz[2] = -0.125 * (1.0 + xi) * (1.0 + eta);
    dz[3] = -0.125 * (1.0 - xi) * (1.0 + eta);
    dz[4] = 0.125 * (1.0 - xi) * (1.0 - eta);
    dz[5] = 0.125 * (1.0 + xi) * (1.0 - eta);
    dz[6] = 0.125 * (1.0 + xi) * (1.0 + eta);
    dz[7] = 0.125 * (1.0 - xi) * (1.0 + eta);
    auto dot = [](const double* u, const double* v) {
      double r = 0;
      for (int i = 0; i < 8; ++i) r += u[i] * v[i];
      return r;
    };
    double J[9];
    J[0] = dot(dx, ex); J[1] = dot(dx, ey); J[2] = dot(dx, ez);
    J[3] = dot(de, ex); J[4] = dot(de, ey); J[5] = dot(de, ez);
    J[6] = dot(dz, ex); J[7] = dot(dz, ey); J[8] = dot(dz, ez);
    const double detJ = J[0] * (J[4] * J[8] - J[5] * J[7]) + J[1] * (J[5] * J[6] - J[3] * J[8]) +
                        J[2] * (J[3] * J[7] - J[4] * J[6]);
    volume += detJ;
  }
  *vol = volume;
  // MeshProcessor.C:151-171
  double sx = 0, sy = 0, sz = 0;
  for (int j = 0; j < 8; ++j) {
    sx += ex[j];
    sy += ey[j];
    sz += ez[j];
  }
  centroid[0] = sx / 8;
  centroid[1] = sy / 8;
  centroid[2] = sz / 8;
}

extern "C" orc_mesh* orc_mesh_new(const orc_options* o) {
  orc_mesh* m = new orc_mesh();
  m->opt = *o;
  m->nx = o->nx;
  m->ny = o->ny;
  m->nz = o->nz;
  const int nx = m->nx, ny = m->ny, nz = m->nz;
  m->ncells = nx * ny * nz;
  // Parallel3DMesh.C:178-197
  const int nnodes = (nx + 1) * (ny + 1) * (nz + 1);
  m->node_xyz.resize((size_t)nnodes * 3);
  for (int i = 0; i < nx + 1; ++i)
    for (int j = 0; j < ny + 1; ++j)
      for (int k = 0; k < nz + 1; ++k) node_coordinate(*o, i, j, k, &m->node_xyz[(size_t)node_id(*m, i, j, k) * 3]);
  // Parallel3DMesh.C:80-96, k fastest
  m->elem_nodes.reserve((size_t)m->ncells * 8);
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        const int ids[8] = {node_id(*m, i, j, k),         node_id(*m, i + 1, j, k),
                            node_id(*m, i + 1, j + 1, k), node_id(*m, i, j + 1, k),
                            node_id(*m, i, j, k + 1),     node_id(*m, i + 1, j, k + 1),
                            node_id(*m, i + 1, j + 1, k + 1), node_id(*m, i, j + 1, k + 1)};
        m->elem_nodes.insert(m->elem_nodes.end(), ids, ids + 8);
      }
  // create_faces, MeshProcessor.C:39-128: faces are found through the list
  // of faces already seen at their smallest node id.
  static const int fno[24] = {0, 1, 5, 4, 1, 2, 6, 5, 2, 3, 7, 6, 3, 0, 4, 7, 0, 3, 2, 1, 4, 5, 6, 7};
  struct FD {
    int minus = -1, plus = -1, fminus = -1, fplus = -1;
    std::array<int, 4> nodes, sorted;
  };
  std::vector<FD> fd;
  std::vector<std::vector<int>> at_node(nnodes);
  for (int ie = 0; ie < m->ncells; ++ie) {
    const int* en = &m->elem_nodes[(size_t)ie * 8];
    for (int ifa = 0; ifa < 6; ++ifa) {
      std::array<int, 4> nd;
      for (int q = 0; q < 4; ++q) nd[q] = en[fno[ifa * 4 + q]];
      std::array<int, 4> srt = nd;
      std::sort(srt.begin(), srt.end());
      bool found = false;
      for (int f : at_node[srt[0]]) {
        if (fd[f].sorted == srt) {
          fd[f].plus = ie;
          fd[f].fplus = ifa;
          found = true;
          break;
        }
      }
      if (found) continue;
      at_node[srt[0]].push_back((int)fd.size());
      FD x;
      x.minus = ie;
      x.fminus = ifa;
      x.nodes = nd;
      x.sorted = srt;
      fd.push_back(x);
    }
  }
  // ElementTopoHexa8 volume and node-mean centroid, MeshProcessor.C:130-171
  m->cell_xyz.resize((size_t)m->ncells * 3);
  m->cell_vol.resize(m->ncells);
  for (int ie = 0; ie < m->ncells; ++ie) {
    double ex[8], ey[8], ez[8];
    for (int j = 0; j < 8; ++j) {
      const int n = m->elem_nodes[(size_t)ie * 8 + j];
      ex[j] = m->node_xyz[(size_t)n * 3];
      ey[j] = m->node_xyz[(size_t)n * 3 + 1];
      ez[j] = m->node_xyz[(size_t)n * 3 + 2];
    }
    orc_hex_volume_centroid(ex, ey, ez, &m->cell_vol[ie], &m->cell_xyz[(size_t)ie * 3]);
  }
  // extract_BC_faces (MeshProcessor.C:189-203): stable partition on elem2 == -1.
  std::vector<int> internal_ids, bc_ids;
  for (int f = 0; f < (int)fd.size(); ++f) (fd[f].plus == -1 ? bc_ids : internal_ids).push_back(f);
  // random_shuffle of the internal faces (Parallel3DMesh.h:362) from the
  // C library's default rand() state.
  std::srand(1);
  std::random_shuffle(internal_ids.begin(), internal_ids.end());

  auto fill = [&](Faces& F, const std::vector<int>& ids) {
    F.n = (int)ids.size();
    F.xf.resize((size_t)F.n * 3);
    F.a.resize((size_t)F.n * 3);
    F.t.resize((size_t)F.n * 3);
    F.b.resize((size_t)F.n * 3);
    F.conn.resize((size_t)F.n * 2);
    F.fidx.resize((size_t)F.n * 2);
    F.nodes.resize(F.n);
    for (int i = 0; i < F.n; ++i) {
      const FD& x = fd[ids[i]];
      F.conn[2 * i] = x.minus;
      F.conn[2 * i + 1] = x.plus;
      F.fidx[2 * i] = x.fminus;
      F.fidx[2 * i + 1] = x.fplus;
      F.nodes[i] = x.nodes;
      face_geometry(m->node_xyz, x.nodes, &F.xf[3 * i], &F.a[3 * i], &F.t[3 * i], &F.b[3 * i]);
    }
    // Faces.h:128-144: BinSort by left cell; serial binning + in-bin
    // insertion sort is a stable sort on the left cell index.
    F.perm.resize(F.n);
    std::iota(F.perm.begin(), F.perm.end(), 0);
    std::stable_sort(F.perm.begin(), F.perm.end(),
                     [&](int p, int q) { return F.conn[2 * p] < F.conn[2 * q]; });
  };
  fill(m->internal, internal_ids);

  // Boundary groups by node membership (Parallel3DMesh.C:199-245,
  // MeshProcessor.C:205-229), pushed in the order of Parallel3DMesh.h:382-396:
  // bottom, top, front, back, right, left.
  auto on_side = [&](int side, int n) {
    const int i = n / ((ny + 1) * (nz + 1));
    const int j = (n / (nz + 1)) % (ny + 1);
    const int k = n % (nz + 1);
    switch (side) {
      case 0: return j == 0;
      case 1: return j == ny;
      case 2: return k == 0;
      case 3: return k == nz;
      case 4: return i == nx;
      default: return i == 0;
    }
  };
  const int pt = o->problem_type;
  const int ref_kind[6] = {pt == 1 ? BC_NOSLIP : BC_TANGENT, pt == 1 ? BC_EXTRAPOLATE : BC_TANGENT,
                           BC_TANGENT, BC_TANGENT, BC_EXTRAPOLATE,
                           pt == 0 ? BC_EXTRAPOLATE : BC_INFLOW};
  for (int side = 0; side < 6; ++side) {
    std::vector<int> ids;
    for (int f : bc_ids) {
      bool in = true;
      for (int q = 0; q < 4; ++q) in = in && on_side(side, fd[f].nodes[q]);
      if (in) ids.push_back(f);
    }
    Faces F;
    fill(F, ids);
    m->groups.push_back(F);
    m->group_kind.push_back(o->bc_kind[side] ? o->bc_kind[side] : ref_kind[side]);
  }
  return m;
}

extern "C" void orc_mesh_free(orc_mesh* m) { delete m; }

extern "C" int orc_mesh_counts(const orc_mesh* m, int* c, int max) {
  int w = 0;
  auto put = [&](int v) { if (w < max) c[w++] = v; };
  put(m->ncells);
  put(m->internal.n);
  put((int)m->groups.size());
  for (const auto& g : m->groups) put(g.n);
  return w;
}

extern "C" void orc_mesh_nodes(const orc_mesh* m, double* xyz) {
  std::copy(m->node_xyz.begin(), m->node_xyz.end(), xyz);
}

extern "C" void orc_mesh_cells(const orc_mesh* m, double* xyz, double* vol) {
  if (xyz) std::copy(m->cell_xyz.begin(), m->cell_xyz.end(), xyz);
  if (vol) std::copy(m->cell_vol.begin(), m->cell_vol.end(), vol);
}

extern "C" void orc_mesh_faces(const orc_mesh* m, int group, double* xf, double* a, double* t,
                               double* b, int* conn, int* fidx, int* perm, int* kind) {
  const Faces& F = group < 0 ? m->internal : m->groups[group];
  if (xf) std::copy(F.xf.begin(), F.xf.end(), xf);
  if (a) std::copy(F.a.begin(), F.a.end(), a);
  if (t) std::copy(F.t.begin(), F.t.end(), t);
  if (b) std::copy(F.b.begin(), F.b.end(), b);
  if (conn) std::copy(F.conn.begin(), F.conn.end(), conn);
  if (fidx) std::copy(F.fidx.begin(), F.fidx.end(), fidx);
  if (perm) std::copy(F.perm.begin(), F.perm.end(), perm);
  if (kind) *kind = group < 0 ? 0 : m->group_kind[group];
}

/* ----------------------------------------------------------- physics -- */

extern "C" void orc_primitives(const double* U, double* V) {
  // GasModel.h:69-90
  double gamma = 1.4;
  double Rgas = 287.05;
  double r = U[0];
  double ri = 1.0 / r;
  double u = U[1] * ri;
  double v = U[2] * ri;
  double w = U[3] * ri;
  double k = 0.5 * (u * u + v * v + w * w);
  double e = U[4] * ri - k;
  double T = e * (gamma - 1.0) / Rgas;
  V[0] = r;
  V[1] = u;
  V[2] = v;
  V[3] = w;
  V[4] = T;
}

namespace {
inline double pressure(const double* V) { return V[0] * kRgas * V[4]; }  // GasModel.h:37-44
inline double enthalpy(const double* V) { return 1004.0 * V[4]; }        // GasModel.h:62-67
inline double dmin(double a, double b) { return (b < a) ? b : a; }       // std::min
inline double dmax(double a, double b) { return (a < b) ? b : a; }       // std::max

// MathToolsDevice.h:62-70 with alpha = 1, beta = 0 on a zeroed y.
void matvec5(const double* A, const double* x, double* y) {
  for (int i = 0; i < 5; ++i) {
    y[i] *= 0.0;
    for (int j = 0; j < 5; ++j) y[i] += 1.0 * A[5 * i + j] * x[j];
  }
}
}  // namespace

extern "C" void orc_roe_flux(const double* pl, const double* pr, const double* fn, const double* ft,
                             const double* fb, double* flux) {
  // Roe_Flux.h:50-265
  const double efix_u = 0.1, efix_c = 0.1, gm1 = 0.4;
  const double rl = pl[0], ul = pl[1], vl = pl[2], wl = pl[3];
  const double p_l = pressure(pl), h_l = enthalpy(pl);
  const double ht_l = h_l + 0.5 * (ul * ul + vl * vl + wl * wl);
  const double mf_l = rl * (fn[0] * ul + fn[1] * vl + fn[2] * wl);
  const double rr = pr[0], ur = pr[1], vr = pr[2], wr = pr[3];
  const double p_r = pressure(pr), h_r = enthalpy(pr);
  const double ht_r = h_r + 0.5 * (ur * ur + vr * vr + wr * wr);
  const double mf_r = rr * (fn[0] * ur + fn[1] * vr + fn[2] * wr);
  const double psum = p_l + p_r;
  flux[0] = 0.5 * (mf_l + mf_r);
  flux[1] = 0.5 * (mf_l * ul + mf_r * ur + fn[0] * psum);
  flux[2] = 0.5 * (mf_l * vl + mf_r * vr + fn[1] * psum);
  flux[3] = 0.5 * (mf_l * wl + mf_r * wr + fn[2] * psum);
  flux[4] = 0.5 * (mf_l * ht_l + mf_r * ht_r);
  const double nn = std::sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const double tn = std::sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
  const double bn = std::sqrt(fb[0] * fb[0] + fb[1] * fb[1] + fb[2] * fb[2]);
  const double nu[3] = {fn[0] / nn, fn[1] / nn, fn[2] / nn};
  const double tu[3] = {ft[0] / tn, ft[1] / tn, ft[2] / tn};
  const double bu[3] = {fb[0] / bn, fb[1] / bn, fb[2] / bn};
  const double denom = 1.0 / (std::sqrt(rl) + std::sqrt(rr));
  const double alpha = std::sqrt(rl) * denom;
  const double beta = 1.0 - alpha;
  const double u = alpha * ul + beta * ur;
  const double v = alpha * vl + beta * vr;
  const double w = alpha * wl + beta * wr;
  const double h = alpha * h_l + beta * h_r +
                   0.5 * alpha * beta *
                       ((ur - ul) * (ur - ul) + (vr - vl) * (vr - vl) + (wr - wl) * (wr - wl));
  const double c = std::sqrt(gm1 * h);
  const double vn = u * nu[0] + v * nu[1] + w * nu[2];
  const double vt = u * tu[0] + v * tu[1] + w * tu[2];
  const double vb = u * bu[0] + v * bu[1] + w * bu[2];
  const double ke = 0.5 * (u * u + v * v + w * w);
  const double ci2 = 1.0 / (c * c);
  const double hci2 = 0.5 * ci2;
  double dq[5];
  dq[0] = rr - rl;
  dq[1] = rr * ur - rl * ul;
  dq[2] = rr * vr - rl * vl;
  dq[3] = rr * wr - rl * wl;
  dq[4] = (rr * ht_r - p_r) - (rl * ht_l - p_l);
  const double cbar = c * nn;
  const double ubar = u * fn[0] + v * fn[1] + w * fn[2];
  const double cfl = std::fabs(ubar) + cbar;
  const double eig1 = ubar + cbar, eig2 = ubar - cbar, eig3 = ubar;
  double a1 = std::fabs(eig1), a2 = std::fabs(eig2), a3 = std::fabs(eig3);
  const double epuc = efix_u * cfl, epcc = efix_c * cfl;
  if (a1 < epcc) a1 = 0.5 * (eig1 * eig1 + epcc * epcc) / epcc;
  if (a2 < epcc) a2 = 0.5 * (eig2 * eig2 + epcc * epcc) / epcc;
  if (a3 < epuc) a3 = 0.5 * (eig3 * eig3 + epuc * epuc) / epuc;
  double eigp[5] = {0.5 * (eig1 + a1), 0.5 * (eig2 + a2), 0.5 * (eig3 + a3), 0.0, 0.0};
  eigp[3] = eigp[4] = eigp[2];
  double eigm[5] = {0.5 * (eig1 - a1), 0.5 * (eig2 - a2), 0.5 * (eig3 - a3), 0.0, 0.0};
  eigm[3] = eigm[4] = eigm[2];
  double ldq[5] = {0, 0, 0, 0, 0}, rlldq[5] = {0, 0, 0, 0, 0};
  double M[25];
  M[0] = gm1 * (ke - h) + c * (c - vn);
  M[1] = c * nu[0] - gm1 * u;
  M[2] = c * nu[1] - gm1 * v;
  M[3] = c * nu[2] - gm1 * w;
  M[4] = gm1;
  M[5] = gm1 * (ke - h) + c * (c + vn);
  M[6] = -c * nu[0] - gm1 * u;
  M[7] = -c * nu[1] - gm1 * v;
  M[8] = -c * nu[2] - gm1 * w;
  M[9] = gm1;
  M[10] = ke - h;
  M[11] = -u;
  M[12] = -v;
  M[13] = -w;
  M[14] = 1.0;
  M[15] = -vt;
  M[16] = tu[0];
  M[17] = tu[1];
  M[18] = tu[2];
  M[19] = 0.0;
  M[20] = -vb;
  M[21] = bu[0];
  M[22] = bu[1];
  M[23] = bu[2];
  M[24] = 0.0;
  matvec5(M, dq, ldq);
  for (int j = 0; j < 5; ++j) ldq[j] = (eigp[j] - eigm[j]) * ldq[j];
  M[0] = hci2;
  M[1] = hci2;
  M[2] = -gm1 * ci2;
  M[3] = 0.0;
  M[4] = 0.0;
  M[5] = (u + nu[0] * c) * hci2;
  M[6] = (u - nu[0] * c) * hci2;
  M[7] = -gm1 * u * ci2;
  M[8] = tu[0];
  M[9] = bu[0];
  M[10] = (v + nu[1] * c) * hci2;
  M[11] = (v - nu[1] * c) * hci2;
  M[12] = -gm1 * v * ci2;
  M[13] = tu[1];
  M[14] = bu[1];
  M[15] = (w + nu[2] * c) * hci2;
  M[16] = (w - nu[2] * c) * hci2;
  M[17] = -gm1 * w * ci2;
  M[18] = tu[2];
  M[19] = bu[2];
  M[20] = (h + ke + c * vn) * hci2;
  M[21] = (h + ke - c * vn) * hci2;
  M[22] = (c * c - gm1 * (h + ke)) * ci2;
  M[23] = vt;
  M[24] = vb;
  matvec5(M, ldq, rlldq);
  for (int i = 0; i < 5; ++i) flux[i] -= 0.5 * rlldq[i];
}

extern "C" void orc_viscous_flux(const double* g15, const double* V, const double* a, double prandtl,
                                 double* vflux) {
  // Viscous_Flux.h:66-98 with GasModel.h:92-106 (Pr a parameter, 0.71 in the reference).
  const double (*g)[3] = reinterpret_cast<const double (*)[3]>(g15);
  const double T = V[4];
  const double viscosity = 1.458e-6 * T * std::sqrt(T) / (T + 110.4);
  const double Pr = prandtl > 0.0 ? prandtl : 0.71;
  const double kcond = viscosity * 1006.0 / Pr;
  double div = 0;
  for (int i = 0; i < 5; ++i) vflux[i] = 0.0;
  for (int i = 0; i < 3; ++i) div += g[i + 1][i];
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) {
      const double delta = (i == j) ? 1 : 0;
      const double S = 0.5 * (g[i + 1][j] + g[j + 1][i]);
      const double tij = S - div * delta / 3.;
      vflux[1 + i] += (2 * viscosity * tij) * a[j];
      vflux[4] += (2 * viscosity * tij) * V[i + 1] * a[j];
    }
    vflux[4] += kcond * g[4][i] * a[i];
  }
}

extern "C" double orc_venkat(double dumax, double dumin, double du, double deltax3) {
  // VenkatLimiter.h:45-73
  double beta = 1;
  double eps2 = deltax3 * beta * beta * beta;
  double phi = 0, num = 0, den = 0;
  if (du > 1e-40) {
    num = (dumax * dumax + eps2) * du + 2 * du * du * dumax;
    den = du * (dumax * dumax + 2 * du * du + dumax * du + eps2);
    phi = num / den;
  } else if (du < -1e-40) {
    num = (dumin * dumin + eps2) * du + 2 * du * du * dumin;
    den = du * (dumin * dumin + 2 * du * du + dumin * du + eps2);
    phi = num / den;
  } else {
    phi = 1;
  }
  return phi;
}

extern "C" double orc_vanalbada(double dumax, double dumin, double du) {
  // VanAlbadaLimiter.h:45-65
  double y = 2;
  if (du > DBL_EPSILON)
    y = dumax / du;
  else if (du < -DBL_EPSILON)
    y = dumin / du;
  double phi = 1;
  if (y < 2) {
    phi = (4 * y - y * y) / (y * y - 4 * y + 8);
    phi = dmax(phi, 0.0);
    phi = dmin(phi, 1.0);
  }
  return phi;
}

/* ------------------------------------------------------------ solver -- */

namespace {
// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): output number i for seed s
// is mix64(s + i * 0x9E3779B97F4A7C15).
uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
// Counter-based uniform draw in [-1, 1) for (seed, cell, k): SplitMix64
// output number 8 cell + k + 1, top 53 bits.
double unif(uint64_t seed, long long cell, int k) {
  uint64_t z = mix64(seed + 0x9E3779B97F4A7C15ull * (uint64_t)(cell * 8 + k + 1));
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}
}  // namespace

struct orc_solver {
  const orc_mesh* m;
  orc_options o;
  int nc;
  double time = 0.0, last_dt = 0.0;
  double inflow[5];
  std::vector<double> sol_n, sol_np1, sol_temp, res;
  std::vector<double> cell_flux, cell_grad, grad, lim, smin, smax, stored_lim;
};

extern "C" orc_solver* orc_solver_new(const orc_mesh* m, const orc_options* o) {
  orc_solver* s = new orc_solver();
  s->m = m;
  s->o = *o;
  s->nc = m->ncells;
  // TimeSolverExplicitRK4.h:218-223
  const double ref_inflow[5] = {0.5805, 503.96, 0.0, 0.0, 343750.0};
  for (int i = 0; i < 5; ++i) s->inflow[i] = o->use_inflow ? o->inflow[i] : ref_inflow[i];
  const size_t n5 = (size_t)s->nc * 5;
  s->sol_n.assign(n5, 0.0);
  s->sol_np1.assign(n5, 0.0);
  s->sol_temp.assign(n5, 0.0);
  s->res.assign(n5, 0.0);
  s->cell_flux.assign(n5, 0.0);
  s->cell_grad.assign(n5 * 3, 0.0);
  s->grad.assign(n5 * 3, 0.0);
  s->lim.assign(n5, 0.0);
  s->smin.assign(n5, 0.0);
  s->smax.assign(n5, 0.0);
  s->stored_lim.assign(n5, 0.0);
  return s;
}

extern "C" void orc_solver_free(orc_solver* s) { delete s; }

extern "C" void orc_solver_init(orc_solver* s) {
  const orc_options& o = s->o;
  const double Rgas = 287.05, gamma = 1.4;
  for (int i = 0; i < s->nc; ++i) {
    double* U = &s->sol_n[(size_t)i * 5];
    const double x = s->m->cell_xyz[(size_t)i * 3];
    if (o.ic_kind == 1) {
      // Sod-type extension, same arithmetic form as initialize_sod3d.
      const double Cv = Rgas / (gamma - 1.0);
      const bool left = x < o.ic[4];
      const double rho = left ? o.ic[0] : o.ic[2];
      const double p = left ? o.ic[1] : o.ic[3];
      const double T = p / (rho * Rgas);
      U[0] = rho;
      U[1] = U[2] = U[3] = 0.0;
      U[4] = rho * (Cv * T);
    } else if (o.ic_kind == 2) {
      double V[5];
      orc_primitives(s->inflow, V);
      const double speed = std::sqrt(V[1] * V[1] + V[2] * V[2] + V[3] * V[3]);
      const double p = V[0] * Rgas * V[4];
      const double rho = V[0] * (1.0 + o.perturb * unif(o.seed, i, 0));
      const double u = V[1] + o.perturb * speed * unif(o.seed, i, 1);
      const double v = V[2] + o.perturb * speed * unif(o.seed, i, 2);
      const double w = V[3] + o.perturb * speed * unif(o.seed, i, 3);
      const double T = p / (rho * Rgas);
      const double Cv = Rgas / (gamma - 1.0);
      U[0] = rho;
      U[1] = rho * u;
      U[2] = rho * v;
      U[3] = rho * w;
      U[4] = rho * (Cv * T) + 0.5 * rho * (u * u + v * v + w * w);
    } else if (o.problem_type == 0) {
      // initialize_sod3d, Initial_Conditions.h:38-88, midx = lx/2 (TimeSolverExplicitRK4.h:217)
      const double midx = o.lx / 2.0;
      const double Cv = Rgas / (gamma - 1.0);
      const double P1 = 68947.57, T1 = 288.889, P2 = 6894.757, T2 = 231.11;
      const double d1 = P1 / (Rgas * T1), e1 = d1 * (Cv * T1);
      const double d2 = P2 / (Rgas * T2), e2 = d2 * (Cv * T2);
      U[0] = x < midx ? d1 : d2;
      U[1] = U[2] = U[3] = 0.0;
      U[4] = x < midx ? e1 : e2;
    } else {
      // initialize_constant, Initial_Conditions.h:95-133
      for (int c = 0; c < 5; ++c) U[c] = s->inflow[c];
    }
  }
  s->sol_np1 = s->sol_n;
  s->sol_temp = s->sol_n;
  std::fill(s->res.begin(), s->res.end(), 0.0);
  s->time = 0.0;
}

extern "C" void orc_solver_set_state(orc_solver* s, const double* U) {
  std::copy(U, U + (size_t)s->nc * 5, s->sol_n.begin());
  s->sol_np1 = s->sol_n;
  s->sol_temp = s->sol_n;
  std::fill(s->res.begin(), s->res.end(), 0.0);
}

namespace {

void prim_of(const std::vector<double>& U, int c, double* V) { orc_primitives(&U[(size_t)c * 5], V); }

// GreenGauss::compute_gradients, GreenGauss.h:51-137 (internal), 142-222 (boundary), 228-270 (sum)
void green_gauss(orc_solver* s) {
  const orc_mesh* m = s->m;
  const Faces& F = m->internal;
  for (int ii = 0; ii < F.n; ++ii) {
    const int i = F.perm[ii];
    const int L = F.conn[2 * i], R = F.conn[2 * i + 1];
    double pl[5], pr[5];
    prim_of(s->sol_temp, L, pl);
    prim_of(s->sol_temp, R, pr);
    const double vl = m->cell_vol[L], vr = m->cell_vol[R];
    for (int d = 0; d < 3; ++d) {
      const double fnorm = F.a[3 * i + d];
      for (int c = 0; c < 5; ++c) {
        const double g = 0.5 * (pl[c] + pr[c]) * fnorm;
        s->cell_grad[((size_t)L * 5 + c) * 3 + d] += g / vl;
        s->cell_grad[((size_t)R * 5 + c) * 3 + d] += -g / vr;
      }
    }
  }
  for (const Faces& G : m->groups) {
    for (int i = 0; i < G.n; ++i) {
      const int c0 = G.conn[2 * i];
      double p[5];
      prim_of(s->sol_temp, c0, p);
      for (int c = 0; c < 5; ++c)
        for (int d = 0; d < 3; ++d) {
          const double g = p[c] * G.a[3 * i + d];
          s->cell_grad[((size_t)c0 * 5 + c) * 3 + d] += g / m->cell_vol[c0];
        }
    }
  }
  for (size_t q = 0; q < s->grad.size(); ++q) {
    s->grad[q] = 0;
    s->grad[q] += s->cell_grad[q];
  }
}

// StencilLimiter::compute_min_max, StencilLimiter.h:56-282, 532-559
void min_max(orc_solver* s) {
  const orc_mesh* m = s->m;
  std::fill(s->smin.begin(), s->smin.end(), 1.0e300);
  std::fill(s->smax.begin(), s->smax.end(), -1.0e300);
  auto push = [&](int cell, const double* fmin, const double* fmax) {
    for (int c = 0; c < 5; ++c) {
      double& a = s->smin[(size_t)cell * 5 + c];
      double& b = s->smax[(size_t)cell * 5 + c];
      a = dmin(a, fmin[c]);
      b = dmax(b, fmax[c]);
    }
  };
  const Faces& F = m->internal;
  for (int ii = 0; ii < F.n; ++ii) {
    const int i = F.perm[ii];
    const int L = F.conn[2 * i], R = F.conn[2 * i + 1];
    double pl[5], pr[5], fmin[5], fmax[5];
    prim_of(s->sol_temp, L, pl);
    prim_of(s->sol_temp, R, pr);
    for (int c = 0; c < 5; ++c) {
      fmin[c] = pr[c] < pl[c] ? pr[c] : pl[c];
      fmax[c] = pr[c] > pl[c] ? pr[c] : pl[c];
    }
    push(L, fmin, fmax);
    push(R, fmin, fmax);
  }
  for (const Faces& G : m->groups)
    for (int ii = 0; ii < G.n; ++ii) {
      const int i = G.perm[ii];
      const int L = G.conn[2 * i];
      double p[5];
      prim_of(s->sol_temp, L, p);
      push(L, p, p);
    }
}

// StencilLimiter::compute_limiter, StencilLimiter.h:288-350, 356-500, 591-616.
// The limiter extension swaps VenkatLimiter::limit for VanAlbadaLimiter::limit.
void limiter(orc_solver* s) {
  const orc_mesh* m = s->m;
  std::fill(s->stored_lim.begin(), s->stored_lim.end(), 1.0);
  auto one_side = [&](int cell, const double* xf, double* out) {
    double p[5];
    prim_of(s->sol_temp, cell, p);
    double disp[3], dist = 0;
    for (int d = 0; d < 3; ++d) {
      disp[d] = xf[d] - m->cell_xyz[(size_t)cell * 3 + d];
      dist += disp[d] * disp[d];
    }
    for (int c = 0; c < 5; ++c) {
      double du = 0;
      for (int d = 0; d < 3; ++d) du += disp[d] * s->grad[((size_t)cell * 5 + c) * 3 + d];
      const double dumax = s->smax[(size_t)cell * 5 + c] - p[c];
      const double dumin = s->smin[(size_t)cell * 5 + c] - p[c];
      out[c] = s->o.limiter == 1 ? orc_vanalbada(dumax, dumin, du) : orc_venkat(dumax, dumin, du, dist);
    }
  };
  auto push = [&](int cell, const double* l) {
    for (int c = 0; c < 5; ++c) {
      double& a = s->stored_lim[(size_t)cell * 5 + c];
      a = dmin(a, l[c]);
    }
  };
  const Faces& F = m->internal;
  for (int ii = 0; ii < F.n; ++ii) {
    const int i = F.perm[ii];
    const int L = F.conn[2 * i], R = F.conn[2 * i + 1];
    double ll[5], lr[5];
    one_side(L, &F.xf[3 * i], ll);
    one_side(R, &F.xf[3 * i], lr);
    push(L, ll);
    push(R, lr);
  }
  for (const Faces& G : m->groups)
    for (int ii = 0; ii < G.n; ++ii) {
      const int i = G.perm[ii];
      double ll[5];
      one_side(G.conn[2 * i], &G.xf[3 * i], ll);
      push(G.conn[2 * i], ll);
    }
  s->lim = s->stored_lim;
}

// compute_face_flux, Flux.h:89-180 (ATOMICS_FLUX branch)
void face_flux(orc_solver* s) {
  const orc_mesh* m = s->m;
  const Faces& F = m->internal;
  const bool so = s->o.second_order_space != 0, visc = s->o.viscous != 0;
  for (int ii = 0; ii < F.n; ++ii) {
    const int i = F.perm[ii];
    const int L = F.conn[2 * i], R = F.conn[2 * i + 1];
    double flux[5], pl[5], pr[5];
    prim_of(s->sol_temp, L, pl);
    prim_of(s->sol_temp, R, pr);
    if (so) {
      for (int c = 0; c < 5; ++c) {
        double gl = 0, gr = 0;
        for (int d = 0; d < 3; ++d) {
          gl += (F.xf[3 * i + d] - m->cell_xyz[(size_t)L * 3 + d]) * s->grad[((size_t)L * 5 + c) * 3 + d];
          gr += (F.xf[3 * i + d] - m->cell_xyz[(size_t)R * 3 + d]) * s->grad[((size_t)R * 5 + c) * 3 + d];
        }
        pl[c] += gl * s->lim[(size_t)L * 5 + c];
        pr[c] += gr * s->lim[(size_t)R * 5 + c];
      }
    }
    orc_roe_flux(pl, pr, &F.a[3 * i], &F.t[3 * i], &F.b[3 * i], flux);
    if (visc) {
      double pf[5], gf[15], vf[5];
      for (int c = 0; c < 5; ++c) {
        pf[c] = 0.5 * (pl[c] + pr[c]);
        for (int d = 0; d < 3; ++d)
          gf[c * 3 + d] = 0.5 * (s->grad[((size_t)L * 5 + c) * 3 + d] + s->grad[((size_t)R * 5 + c) * 3 + d]);
      }
      orc_viscous_flux(gf, pf, &F.a[3 * i], s->o.prandtl, vf);
      for (int c = 0; c < 5; ++c) flux[c] -= vf[c];
    }
    for (int c = 0; c < 5; ++c) {
      s->cell_flux[(size_t)L * 5 + c] += -flux[c];
      s->cell_flux[(size_t)R * 5 + c] += flux[c];
    }
  }
}

// Boundary fluxes of one group: Extrapolate_BC.h:69-101, Tangent_BC.h:71-122,
// NoSlip_BC.h:76-156, Inflow_BC.h:72-110.
void bc_flux(orc_solver* s, const Faces& G, int kind) {
  const orc_mesh* m = s->m;
  for (int i = 0; i < G.n; ++i) {
    const int c0 = G.conn[2 * i];
    const double* n = &G.a[3 * i];
    double pl[5], pr[5], flux[5], vflux[5] = {0, 0, 0, 0, 0};
    prim_of(s->sol_temp, c0, pl);
    if (kind == BC_EXTRAPOLATE) {
      orc_roe_flux(pl, pl, n, &G.t[3 * i], &G.b[3 * i], flux);
    } else if (kind == BC_INFLOW) {
      orc_primitives(s->inflow, pr);
      orc_roe_flux(pl, pr, n, &G.t[3 * i], &G.b[3 * i], flux);
    } else {
      double an = 0;
      for (int d = 0; d < 3; ++d) an += n[d] * n[d];
      an = std::sqrt(an);
      double ub = 0.0;
      ub += pl[1] * n[0] / an;
      ub += pl[2] * n[1] / an;
      ub += pl[3] * n[2] / an;
      pr[0] = pl[0];
      pr[1] = pl[1] - 2 * ub * n[0] / an;
      pr[2] = pl[2] - 2 * ub * n[1] / an;
      pr[3] = pl[3] - 2 * ub * n[2] / an;
      pr[4] = pl[4];
      orc_roe_flux(pl, pr, n, &G.t[3 * i], &G.b[3 * i], flux);
      if (kind == BC_NOSLIP) {
        double pf[5] = {pl[0], 0.0, 0.0, 0.0, pl[4]}, gf[15], un[3], dist = 0;
        for (int d = 0; d < 3; ++d) {
          dist += std::pow(G.xf[3 * i + d] - m->cell_xyz[(size_t)c0 * 3 + d], 2);
          un[d] = n[d] / an;
        }
        const double inv = 1.0 / std::sqrt(dist);
        for (int c = 0; c < 5; ++c)
          for (int d = 0; d < 3; ++d) gf[c * 3 + d] = (pf[c] - pl[c]) * un[d] * inv;
        orc_viscous_flux(gf, pf, n, s->o.prandtl, vflux);
      }
    }
    for (int c = 0; c < 5; ++c) s->cell_flux[(size_t)c0 * 5 + c] += -flux[c] + vflux[c];
  }
}

// Stable dt for the CFL extension (no reference counterpart):
//   lc = |u| S_x + |v| S_y + |w| S_z + c N1,  lv = N2,
//   dt_c = vol / (lc + 4 kappa lv / vol),  kappa = max(4/3, gamma/Pr) mu / rho
// (kappa = 0 for inviscid runs; the convective plus viscous spectral radius
// of Blazek, Computational Fluid Dynamics, sec. 6.1.4, with sum_f |u.n_f|
// taken as sum_d |u_d| S_d, equal to it for faces normal to an axis and an
// upper bound otherwise), dt = cfl * min_c dt_c.  S_d = sum_f |n_f,d|,
// N1 = sum_f |n_f|, N2 = sum_f |n_f|^2 over the cell's faces in the
// Green-Gauss order.
double stable_dt(orc_solver* s) {
  const orc_mesh* m = s->m;
  std::vector<double> geo((size_t)s->nc * 5, 0.0);
  auto add = [&](int c, const double* n) {
    double* g = &geo[(size_t)c * 5];
    const double n2 = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    g[0] = g[0] + std::fabs(n[0]);
    g[1] = g[1] + std::fabs(n[1]);
    g[2] = g[2] + std::fabs(n[2]);
    g[3] = g[3] + std::sqrt(n2);
    g[4] = g[4] + n2;
  };
  const Faces& F = m->internal;
  for (int ii = 0; ii < F.n; ++ii) {
    const int i = F.perm[ii];
    add(F.conn[2 * i], &F.a[3 * i]);
    add(F.conn[2 * i + 1], &F.a[3 * i]);
  }
  for (const Faces& G : m->groups)
    for (int i = 0; i < G.n; ++i) add(G.conn[2 * i], &G.a[3 * i]);
  std::vector<double> V((size_t)s->nc * 5), lc((size_t)s->nc), lv((size_t)s->nc);
  for (int c = 0; c < s->nc; ++c) {
    double* p = &V[(size_t)c * 5];
    prim_of(s->sol_n, c, p);
    const double* g = &geo[(size_t)c * 5];
    const double snd = std::sqrt(1.4 * (p[0] * kRgas * p[4]) / p[0]);
    lc[c] = std::fabs(p[1]) * g[0] + std::fabs(p[2]) * g[1] + std::fabs(p[3]) * g[2] + snd * g[3];
    lv[c] = g[4];
  }
  const double Pr = s->o.prandtl > 0.0 ? s->o.prandtl : 0.71;
  const double kfac = std::max(4.0 / 3.0, 1.4 / Pr);
  double dtmin = 1e300;
  for (int c = 0; c < s->nc; ++c) {
    const double* p = &V[(size_t)c * 5];
    const double vol = m->cell_vol[c];
    double lam = lc[c];
    if (s->o.viscous) {
      const double mu = 1.458e-6 * p[4] * std::sqrt(p[4]) / (p[4] + 110.4);
      const double kappa = kfac * mu / p[0];
      lam = lam + 4.0 * kappa * lv[c] / vol;
    }
    const double d = vol / lam;
    dtmin = d < dtmin ? d : dtmin;
  }
  return s->o.cfl * dtmin;
}

}  // namespace

extern "C" void orc_solver_step(orc_solver* s, int nsteps) {
  // TimeSolverExplicitRK4.h:340-490
  const double alpha[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
  const double beta[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const orc_mesh* m = s->m;
  const int nc = s->nc;
  const bool so = s->o.second_order_space != 0, visc = s->o.viscous != 0;
  for (int it = 0; it < nsteps; ++it) {
    const double dt = s->o.cfl > 0.0 ? stable_dt(s) : s->o.dt;
    s->last_dt = dt;
    s->time += dt;
    for (int irk = 0; irk < 4; ++irk) {
      for (size_t q = 0; q < (size_t)nc * 5; ++q) s->sol_temp[q] = s->sol_n[q] + alpha[irk] * s->res[q];
      std::fill(s->cell_flux.begin(), s->cell_flux.end(), 0.0);
      std::fill(s->cell_grad.begin(), s->cell_grad.end(), 0.0);
      if (so || visc) green_gauss(s);
      if (so) {
        min_max(s);
        limiter(s);
      }
      face_flux(s);
      // Boundary kinds in the order of TimeSolverExplicitRK4.h:428-475:
      // Extrapolate, Tangent, NoSlip, Inflow; groups in boundary_faces order.
      for (int kind = BC_EXTRAPOLATE; kind <= BC_INFLOW; ++kind)
        for (size_t g = 0; g < m->groups.size(); ++g)
          if (m->group_kind[g] == kind) bc_flux(s, m->groups[g], kind);
      // apply_cell_flux, Flux.h:210-228
      for (int c = 0; c < nc; ++c)
        for (int k = 0; k < 5; ++k) {
          double r = 0.0;
          r = r + dt / m->cell_vol[c] * s->cell_flux[(size_t)c * 5 + k];
          s->res[(size_t)c * 5 + k] = r;
        }
      for (size_t q = 0; q < (size_t)nc * 5; ++q) s->sol_np1[q] = s->sol_np1[q] + beta[irk] * s->res[q];
    }
    s->sol_n = s->sol_np1;
  }
}

extern "C" void orc_solver_get_state(const orc_solver* s, double* U) {
  std::copy(s->sol_n.begin(), s->sol_n.end(), U);
}
extern "C" void orc_solver_get_residual(const orc_solver* s, double* R) {
  std::copy(s->res.begin(), s->res.end(), R);
}
extern "C" void orc_solver_get_gradients(const orc_solver* s, double* grad, double* lim) {
  if (grad) std::copy(s->grad.begin(), s->grad.end(), grad);
  if (lim) std::copy(s->lim.begin(), s->lim.end(), lim);
}
extern "C" double orc_solver_time(const orc_solver* s) { return s->time; }
extern "C" double orc_solver_last_dt(const orc_solver* s) { return s->last_dt; }
