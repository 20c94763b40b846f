/*
 * ref_mesh_driver.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * Thin C driver around the reference's own mesh-processing sources, compiled
 * where they lie (MeshProcessor.C, Face.C, ElementTopoHexa8.C, ElementTopo.C,
 * Cell.C under the reference's kokkos/ directory; see oracle/Makefile).  Those
 * files need no Kokkos or MPI, so the face creation, face geometry, cell
 * volume/centroid and boundary-face grouping below are the reference's code,
 * not a restatement.  What this driver supplies itself, because it lives in
 * Parallel3DMesh.{h,C} (which include Kokkos), is the one-rank node/element
 * loop (Parallel3DMesh.C:80-96, 178-245; Parallel3DMesh.h:453-487) and the
 * call sequence of Parallel3DMesh::fillMeshData (Parallel3DMesh.h:173-362).
 * Output lands in oracle/_ref/libref_mesh.so and is used only to pin the
 * oracle's mesh restatement bit for bit.
 */
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <set>
#include <vector>

#include "Cell.h"
#include "Face.h"
#include "MeshProcessor.h"

struct ref_mesh {
  int ncells;
  std::vector<Cell> cells;
  std::vector<Face> internal;
  std::vector<std::vector<Face>> groups;  // bottom, top, front, back, right, left
};

extern "C" {

ref_mesh* ref_mesh_new(int nx, int ny, int nz, double lx, double ly, double lz, double angle) {
  auto nid = [&](int i, int j, int k) { return i * ((ny + 1) * (nz + 1)) + j * (nz + 1) + k; };
  std::vector<int> conn;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j)
      for (int k = 0; k < nz; ++k) {
        conn.push_back(nid(i, j, k));
        conn.push_back(nid(i + 1, j, k));
        conn.push_back(nid(i + 1, j + 1, k));
        conn.push_back(nid(i, j + 1, k));
        conn.push_back(nid(i, j, k + 1));
        conn.push_back(nid(i + 1, j, k + 1));
        conn.push_back(nid(i + 1, j + 1, k + 1));
        conn.push_back(nid(i, j + 1, k + 1));
      }
  std::vector<double> xyz;
  const double PI = 3.14159265;
  for (int i = 0; i < nx + 1; ++i)
    for (int j = 0; j < ny + 1; ++j)
      for (int k = 0; k < nz + 1; ++k) {
        double c[3];
        c[0] = (double)i / (nx) * lx;
        c[2] = (double)k / (nz) * lz;
        if (c[0] < lx / 2.0) {
          c[1] = (double)j / (ny) * ly;
        } else {
          double y_ramp = (c[0] - (lx / 2.0)) * tan(angle * PI / 180.0);
          double ly_scaled = ly - y_ramp;
          c[1] = y_ramp + (double)j * ly_scaled / (ny);
        }
        xyz.push_back(c[0]);
        xyz.push_back(c[1]);
        xyz.push_back(c[2]);
      }
  ref_mesh* m = new ref_mesh();
  const int ne = nx * ny * nz, nn = (nx + 1) * (ny + 1) * (nz + 1);
  m->ncells = ne;
  m->cells.resize(ne);
  std::vector<Face> faces;
  create_faces(conn, faces, xyz, ne, nn);
  compute_cell_volumes(m->cells, conn, xyz, ne, 0);
  compute_cell_centroid(m->cells, conn, xyz, ne);
  delete_ghosted_faces(faces, ne);
  std::vector<Face> bc;
  extract_BC_faces(faces, bc);
  std::set<int> sides[6];
  for (int i = 0; i < nx + 1; ++i)
    for (int k = 0; k < nz + 1; ++k) {
      sides[0].insert(nid(i, 0, k));
      sides[1].insert(nid(i, ny, k));
    }
  for (int i = 0; i < nx + 1; ++i)
    for (int j = 0; j < ny + 1; ++j) {
      sides[2].insert(nid(i, j, 0));
      sides[3].insert(nid(i, j, nz));
    }
  for (int j = 0; j < ny + 1; ++j)
    for (int k = 0; k < nz + 1; ++k) {
      sides[4].insert(nid(nx, j, k));
      sides[5].insert(nid(0, j, k));
    }
  m->groups.resize(6);
  for (int s = 0; s < 6; ++s) organize_BC_faces(bc, m->groups[s], sides[s]);
  std::srand(1);
  std::random_shuffle(faces.begin(), faces.end());
  m->internal.swap(faces);
  return m;
}

void ref_mesh_free(ref_mesh* m) { delete m; }

int ref_mesh_counts(const ref_mesh* m, int* c) {
  c[0] = m->ncells;
  c[1] = (int)m->internal.size();
  c[2] = 6;
  for (int g = 0; g < 6; ++g) c[3 + g] = (int)m->groups[g].size();
  return 9;
}

void ref_mesh_cells(ref_mesh* m, double* xyz, double* vol) {
  for (int i = 0; i < m->ncells; ++i) {
    vol[i] = m->cells[i].GetVolume();
    for (int d = 0; d < 3; ++d) xyz[3 * i + d] = m->cells[i].GetCoords()[d];
  }
}

void ref_mesh_faces(const ref_mesh* m, int group, double* xf, double* a, double* t, double* b,
                    int* conn, int* fidx) {
  const std::vector<Face>& F = group < 0 ? m->internal : m->groups[group];
  for (size_t i = 0; i < F.size(); ++i) {
    double av[3], tv[3], bv[3];
    F[i].GetAreaVecAndTangentVec(av, tv, bv);
    for (int d = 0; d < 3; ++d) {
      xf[3 * i + d] = F[i].GetCoords()[d];
      a[3 * i + d] = av[d];
      t[3 * i + d] = tv[d];
      b[3 * i + d] = bv[d];
    }
    conn[2 * i] = F[i].GetElem1();
    conn[2 * i + 1] = F[i].GetElem2();
    fidx[2 * i] = F[i].GetElem1_FluxIndex();
    fidx[2 * i + 1] = F[i].GetElem2_FluxIndex();
  }
}

}  // extern "C"
