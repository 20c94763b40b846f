// solver.hip -- MI355X (gfx950) kernels and C ABI of the miniAero RK4 path.
//
// Replaces the kernel sequence that TimeSolverExplicitRK4::Solve launches per
// RK stage (TimeSolverExplicitRK4.h:352-486): update, zero_cell_flux,
// GreenGauss (3 kernels + per-group launches), StencilLimiter min/max and
// limiter (6 kernels + per-group launches), compute_face_flux, one launch per
// boundary group and kind, apply_cell_flux and the RK accumulate -- about 37
// launches and fences per second-order stage -- with three cell-centric
// kernels and no atomics:
//
//   k_prim   U_s = U_n + alpha_k R;  V = primitives(U_s)           (per cell, stage 0 when
//            V does not hold them; otherwise every stage's V comes from
//            the previous stage's k_flux)
//   k_grad   Green-Gauss gradient, stencil min/max and the limiter  (per cell)
//            of every cell, from V of the cell and its face neighbours;
//            inviscid second order: also the cell's extrapolated values
//            at its faces, which k_flux then reads instead of V, grad, phi
//   k_bc     boundary-face contributions (Roe of the ghost state    (per boundary face)
//            [+ no-slip viscous flux]); when k_grad runs, its extra
//            workgroups do this in the same launch
//   k_flux   per tile of 256 cells: every internal face touching    (per cell tile)
//            the tile evaluated once into LDS (Roe [+ reconstruction]
//            [+ viscous]), then per cell: sum of face and boundary
//            contributions, residual R = dt/V sum, U_acc += beta_k R,
//            and the next stage's V from U_n + alpha_k+1 R into the other
//            V buffer (stage 3: the next step's, from the new U; R itself
//            is stored on stage 3 only)
//
// Determinism and parity.  The reference scatters face results into cells
// with atomics; with one thread the additions land in a fixed order: internal
// faces in the permuted order (stable by left cell), then boundary faces.  At
// setup every cell gets its face list in exactly that order (slot table), and
// each kernel sums its cell's contributions in slot order.  A face between two
// tiles is evaluated by both tiles -- bitwise the same value, since it is a
// pure function of the same inputs -- so no workgroup ever waits on another
// and the result does not depend on scheduling.  With -ffp-contract=off the FP64
// results equal those of a one-thread CPU run of the reference.
//
// Layout in HBM (structure of arrays, one plane per component, so a
// wavefront's 64 lanes read 64 consecutive doubles; the gradient records,
// gathered per face by the flux kernel, are one 16-byte-aligned block per
// cell instead).  Cells are renumbered into 4 x 8 x 8 bricks of the mesh
// lattice (one brick per flux tile) when the mesh is a lattice; results go
// back to the caller in the caller's order.
//   cells   vol[nc], U[2][5][nc] (U_n / U_acc ping-pong), R[5][nc],
//           V[2][5][nc] (stage primitives, even/odd stages),
//           rec[nc][20] (gradient 15, limiter 5 per cell);
//           int32 slot[6][nc] (face*2 + side, reference order),
//           nbr[6][nc] (the cell across each slot), tile slot codes [6][nc];
//           xc[3][nc] for the initial condition only
//   faces   n[3][nf] (area vector; an FP32 copy in the FP32 form),
//           disp[6][nf] (x_f - x_c of the left and right cell, arithmetic
//           type), left[nf], right[nf], kind[nf] (int8), wall[nf]; tangent
//           and binormal are rebuilt from n in the kernels, and stored as
//           frame[6][nf] only for meshes whose tangents do not follow the
//           reference construction; internal faces in processing (left-cell)
//           order, then the boundary groups
//   tiles   ptr[ntiles+1], faces[] (internal faces touching the tile) with
//           lr[] (their left and right cells), cptr[ntiles+1], case[ntiles]
// Entry point declarations and the reference interfaces they replace are in
// include/miniaero.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/miniaero.h"
#include "common.hpp"
#include "internal.hpp"
#include "physics.hpp"

namespace {

constexpr int kMaxSlots = 6;  // hex8 cells (ElementTopoHexa8); checked at setup
constexpr int kBlock = 256;
constexpr int kTile = 256;    // cells per workgroup of the flux kernel
constexpr int kXcds = 8;      // MI355X dies; workgroups are dealt to them round-robin
// Tile-face cell entries of the face-value scheme (inviscid second order):
// cell index in the low kCellBits bits (a run has at most 2^28 cells), the
// face's slot in that cell's slot list above them; other schemes store the
// plain cell index.
constexpr int kCellBits = 28;
__device__ __forceinline__ int cell_of(int x) { return x & ((1 << kCellBits) - 1); }
__device__ __forceinline__ int slot_of(int x) { return x >> kCellBits; }

struct DevMesh {
  int nc;    // owned cells
  int nif;   // internal faces
  int nf;    // all faces
  const double* __restrict__ vol;
  const int* __restrict__ slot;    // kMaxSlots planes of nc; entry = face*2 + side, -1 empty
  const double* __restrict__ fn;   // 3 planes of nf: area vector
  const float* __restrict__ fn32;  // the same in FP32 (FP32-arithmetic form only)
  const double* __restrict__ ffr;  // 6 planes of nf: tangent, binormal as given (or null)
  int derive_frame;                // 1: rebuild tangent and binormal from fn (ffr unused)
  // 6 planes of nf in the arithmetic type: T(x_f - x_left) (planes 0-2) and
  // T(x_f - x_right) (3-5, internal faces), from FP64 centroids
  const void* __restrict__ fdx;
  const int* __restrict__ fl;
  const int* __restrict__ fr;
  const signed char* __restrict__ fkind;
  const double* __restrict__ fwall;  // 1 / |x_f - x_c| of boundary faces
  // CFL runs: 5 planes of nc, each cell's face geometry S_x, S_y, S_z, N1, N2
  // (see dt_tile); null otherwise
  const double* __restrict__ dtg;
  // Geometry records of the faces in the arithmetic type, kGeoStride values
  // each: area vector, T(x_f - x_left), T(x_f - x_right) (zero for boundary
  // faces), the same values as the fn/fn32 and fdx planes; faces with the
  // same values share a record (TileMap::fgid, sgeo)
  const void* __restrict__ gtab;
  // kMaxSlots planes of nc: (record << 1) | side of each slot's face, -1 for
  // an empty slot (the gradient kernel's slot list)
  const int* __restrict__ sgeo;
  // 1: every face's area vector has at most one nonzero component and every
  // cell's 1 / vol is finite (box meshes; see grad_cell)
  int axis;
  // 1: no cell has an empty slot (every cell of a hexahedral mesh has six faces)
  int full;
};

struct Inflow {
  double U[5];
};

// Tiles of the flux kernel: tile b holds at most kTile cells, a brick of the
// mesh lattice when there is one, and the internal faces
// faces[ptr[b] .. ptr[b+1]) that touch them (ascending).
struct TileMap {
  int ntiles;
  int per_xcd;                    // tiles per XCD, ceil(ntiles / kXcds)
  int max_faces;                  // largest ptr[b+1] - ptr[b]
  const int* __restrict__ ptr;    // ntiles + 1
  const int* __restrict__ faces;  // device face indices
  const int2* __restrict__ lr;    // parallel to faces: the face's left and right cells (+ slots, see kCellBits)
  const int* __restrict__ slot;   // kMaxSlots planes of nc, see k_flux
  const int* __restrict__ cptr;   // ntiles + 1: cells of tile b are [cptr[b], cptr[b+1])
  const int* __restrict__ tcase;  // ntiles: the case (independent mesh) tile b belongs to
  const int* __restrict__ fgid;   // parallel to faces: the face's record in DevMesh::gtab
};

template <class S, class T>
__device__ __forceinline__ void load5(const S* __restrict__ a, int n, int i, T* v) {
#pragma unroll
  for (int k = 0; k < 5; ++k) v[k] = T(a[(size_t)k * n + i]);
}

template <class S, class T>
__device__ __forceinline__ void load3(const S* __restrict__ a, int n, int i, T* v) {
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = T(a[(size_t)k * n + i]);
}

// Gradient record: the gradient (15 values, component k and direction d at
// 3k + d) and the limiter (5 values) of a cell stored together, kRec values
// per cell, cell after cell.  The flux kernel gathers both for the two cells
// of every face; as one 16-byte-aligned block per cell a gather moves only the
// cell's own bytes, where 20 separate planes would each bring in a cache
// line mostly of other cells (a tile's neighbours in a brick lie every 8th or
// 64th cell).
constexpr int kRec = 20;

template <class T>
__device__ __forceinline__ void load_rec(const T* __restrict__ rec, int c, T* v) {
  if constexpr (sizeof(T) == 8) {
    const double2* p = reinterpret_cast<const double2*>(rec + (size_t)c * kRec);
#pragma unroll
    for (int i = 0; i < kRec / 2; ++i) {
      const double2 x = p[i];
      v[2 * i] = x.x;
      v[2 * i + 1] = x.y;
    }
  } else {
    const float4* p = reinterpret_cast<const float4*>(rec + (size_t)c * kRec);
#pragma unroll
    for (int i = 0; i < kRec / 4; ++i) {
      const float4 x = p[i];
      v[4 * i] = x.x;
      v[4 * i + 1] = x.y;
      v[4 * i + 2] = x.z;
      v[4 * i + 3] = x.w;
    }
  }
}

// Shared memory that stores the gradient records of a workgroup of kBlock
// cells: a row of kRec + 1 values per thread (the pad spreads the rows over
// the LDS banks).
template <class T>
constexpr int kStageRow = kRec + 1;

// Each thread puts its cell's record (out) in its row, then every wavefront
// writes its 64 cells' records -- one contiguous block in rec, from its first
// cell c0 on -- 16 bytes per lane per store, so the stores are whole cache
// lines instead of each lane writing its own record.  Every thread of the
// workgroup calls it (it holds a barrier); cells at or beyond nc are not
// written.
template <class T>
__device__ __forceinline__ void store_records(T* __restrict__ rec, T* stage, const T* out, int c0, int nc) {
  T* row = stage + threadIdx.x * kStageRow<T>;
#pragma unroll
  for (int j = 0; j < kRec; ++j) row[j] = out[j];
  __syncthreads();
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int first = c0 + w * 64;
  const int nvalid = min(64, nc - first);
  constexpr int E = 16 / sizeof(T);  // values per 16-byte chunk; kRec is a multiple of it
  const T* wst = stage + w * 64 * kStageRow<T>;
  for (int e = lane * E; e < nvalid * kRec; e += 64 * E) {
    const T* src = wst + (e / kRec) * kStageRow<T> + e % kRec;
    T* dst = rec + (size_t)first * kRec + e;
    if constexpr (sizeof(T) == 8)
      *reinterpret_cast<double2*>(dst) = make_double2(src[0], src[1]);
    else
      *reinterpret_cast<float4*>(dst) = make_float4(src[0], src[1], src[2], src[3]);
  }
}

// Shared face-geometry record: area vector, x_f - x_left, x_f - x_right
// (kGeoVals values), padded to whole 16-byte chunks (kGeoStride values: 10
// in FP64, 12 in FP32) so that one record is a few vector loads.
constexpr int kGeoVals = 9;
template <class T>
constexpr int kGeoStride = (kGeoVals * (int)sizeof(T) + 15) / 16 * 16 / (int)sizeof(T);

// Record g of the shared geometry table: the area vector, and with kDisp
// both displacements.
template <class T, bool kDisp>
__device__ __forceinline__ void load_geo(const DevMesh& m, int g, T* a, T* dl, T* dr) {
  if constexpr (sizeof(T) == 8 && !kDisp) {
    const double2* p = reinterpret_cast<const double2*>(m.gtab) + (size_t)g * (kGeoStride<T> / 2);
    const double2 x0 = p[0];
    a[0] = x0.x;
    a[1] = x0.y;
    a[2] = reinterpret_cast<const double*>(p)[2];
  } else if constexpr (sizeof(T) == 8) {
    const double2* p = reinterpret_cast<const double2*>(m.gtab) + (size_t)g * (kGeoStride<T> / 2);
    const double2 x0 = p[0], x1 = p[1];
    a[0] = x0.x;
    a[1] = x0.y;
    a[2] = x1.x;
    {
      const double2 x2 = p[2], x3 = p[3], x4 = p[4];
      dl[0] = x1.y;
      dl[1] = x2.x;
      dl[2] = x2.y;
      dr[0] = x3.x;
      dr[1] = x3.y;
      dr[2] = x4.x;
    }
  } else if constexpr (!kDisp) {  // three values: no register for a fourth
    const float* p = static_cast<const float*>(m.gtab) + (size_t)g * kGeoStride<T>;
#pragma unroll
    for (int k = 0; k < 3; ++k) a[k] = p[k];
  } else {
    const float4* p = reinterpret_cast<const float4*>(m.gtab) + (size_t)g * (kGeoStride<T> / 4);
    const float4 x0 = p[0];
    a[0] = x0.x;
    a[1] = x0.y;
    a[2] = x0.z;
    {
      const float4 x1 = p[1], x2 = p[2];
      dl[0] = x0.w;
      dl[1] = x1.x;
      dl[2] = x1.y;
      dr[0] = x1.z;
      dr[1] = x1.w;
      dr[2] = x2.x;
    }
  }
}

// Displacement x_f - x_c of side `side` of record g.
template <class T>
__device__ __forceinline__ void load_geo_disp(const DevMesh& m, int g, int side, T* d) {
  const T* p = static_cast<const T*>(m.gtab) + (size_t)g * kGeoStride<T> + 3 + 3 * side;
#pragma unroll
  for (int k = 0; k < 3; ++k) d[k] = p[k];
}

// Area vectors in the arithmetic type.
template <class T>
__device__ __forceinline__ const T* area(const DevMesh& m);
template <>
__device__ __forceinline__ const double* area<double>(const DevMesh& m) {
  return m.fn;
}
template <>
__device__ __forceinline__ const float* area<float>(const DevMesh& m) {
  return m.fn32;
}

// Face frame of face f from its area vector a (already loaded): rebuilt from
// a, or read from the stored frame planes.
template <class T>
__device__ __forceinline__ ma::FaceFrameT<T> frame_of(const DevMesh& m, int f, const T* a) {
  if (m.derive_frame) return ma::frame_from_area_t<T>(a);
  const double* p = m.ffr;
  const size_t nf = m.nf;
  T t[3], b[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    t[d] = T(p[d * nf + f]);
    b[d] = T(p[(3 + d) * nf + f]);
  }
  return ma::make_frame_t<T>(a, t, b);
}

template <class T>
__device__ __forceinline__ ma::FaceFrameT<T> load_frame(const DevMesh& m, int f) {
  T a[3];
  load3(area<T>(m), m.nf, f, a);
  return frame_of<T>(m, f, a);
}

// ---------------------------------------------------------------- kernels --

// Initial condition: initialize_sod3d / initialize_constant
// (Initial_Conditions.h:38-133) and the Sod-type / perturbed-inflow
// extensions, written to U_n.
struct IcParams {
  int kind;          // MA_IC_*
  int problem_type;
  double midx;       // reference Sod split, lx / 2
  double ic[5];
  double inflow[5];
  double perturb;
  unsigned long long seed;
  int cell0;         // first cell of the case in the caller's union numbering
};

// SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): output number i for seed
// s is mix64(s + i * 0x9E3779B97F4A7C15).
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Counter-based uniform draw in [-1, 1) keyed by (seed, cell, k): SplitMix64
// output number 8 cell + k + 1, top 53 bits.
__device__ __forceinline__ double unif(unsigned long long seed, long long cell, int k) {
  unsigned long long z = mix64(seed + 0x9E3779B97F4A7C15ull * (unsigned long long)(cell * 8 + k + 1));
  return (double)(z >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// xc: cell centroids (3 planes of nc); org: device cell -> the caller's index.
__global__ __launch_bounds__(kBlock) void k_init(DevMesh m, const double* __restrict__ xc,
                                                 const int* __restrict__ org, const IcParams* __restrict__ cases,
                                                 const int* __restrict__ cell_case, double* __restrict__ U) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= m.nc) return;
  const IcParams& p = cases[cell_case[c]];
  const double Rgas = 287.05, gamma = 1.4;
  const double x = xc[c];
  double u[5];
  if (p.kind == MA_IC_SOD) {
    const double Cv = Rgas / (gamma - 1.0);
    const bool left = x < p.ic[4];
    const double rho = left ? p.ic[0] : p.ic[2];
    const double pr = left ? p.ic[1] : p.ic[3];
    const double T = pr / (rho * Rgas);
    u[0] = rho;
    u[1] = u[2] = u[3] = 0.0;
    u[4] = rho * (Cv * T);
  } else if (p.kind == MA_IC_PERTURBED_INFLOW) {
    double V[5];
    ma::primitives(p.inflow, V);
    const double speed = sqrt(V[1] * V[1] + V[2] * V[2] + V[3] * V[3]);
    const double pr = V[0] * Rgas * V[4];
    const int oc = org[c] - p.cell0;  // the draw is keyed by the case's own cell index
    const double rho = V[0] * (1.0 + p.perturb * unif(p.seed, oc, 0));
    const double uu = V[1] + p.perturb * speed * unif(p.seed, oc, 1);
    const double vv = V[2] + p.perturb * speed * unif(p.seed, oc, 2);
    const double ww = V[3] + p.perturb * speed * unif(p.seed, oc, 3);
    const double T = pr / (rho * Rgas);
    const double Cv = Rgas / (gamma - 1.0);
    u[0] = rho;
    u[1] = rho * uu;
    u[2] = rho * vv;
    u[3] = rho * ww;
    u[4] = rho * (Cv * T) + 0.5 * rho * (uu * uu + vv * vv + ww * ww);
  } else if (p.problem_type == 0) {
    const double Cv = Rgas / (gamma - 1.0);
    const double P1 = 68947.57, T1 = 288.889, P2 = 6894.757, T2 = 231.11;
    const double d1 = P1 / (Rgas * T1), e1 = d1 * (Cv * T1);
    const double d2 = P2 / (Rgas * T2), e2 = d2 * (Cv * T2);
    u[0] = x < p.midx ? d1 : d2;
    u[1] = u[2] = u[3] = 0.0;
    u[4] = x < p.midx ? e1 : e2;
  } else {
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = p.inflow[k];
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) U[(size_t)k * m.nc + c] = u[k];
}

// Stage state and its primitives: update<Device> with alpha_k
// (TimeSolverExplicitRK4.h:106-128, 355-357) fused with ComputePrimitives.
// The stage state is FP64; V is stored in the arithmetic type T.
template <class T>
__device__ __forceinline__ void prim_cell(int nc, const double* __restrict__ Un, const double* __restrict__ R,
                                          double alpha, T* __restrict__ V, int c) {
  double u[5], v[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) u[k] = Un[(size_t)k * nc + c] + alpha * R[(size_t)k * nc + c];
  ma::primitives(u, v);
#pragma unroll
  for (int k = 0; k < 5; ++k) V[(size_t)k * nc + c] = T(v[k]);
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_prim(int nc, const double* __restrict__ Un,
                                                 const double* __restrict__ R, double alpha,
                                                 T* __restrict__ V) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < nc) prim_cell(nc, Un, R, alpha, V, c);
}


// Green-Gauss gradient (GreenGauss.h:51-270), then, for second order, the
// stencil min/max (StencilLimiter.h:56-282) and the limiter
// (StencilLimiter.h:288-500) of one cell, all from the cell's slot list.
// nbr: kMaxSlots planes of nc, the cell across each slot's internal face (the
// cell itself for a boundary face or an empty slot).  A boundary face (side
// 0, neighbour = the cell) goes through the internal-face formulas with the
// same bits as the reference's boundary terms -- 0.5 * (v + v) == v and
// min(v, v) == v exactly -- so the slot loop has no branch; all six slots'
// area vectors and neighbour values are loaded before the first is used (one
// round of gathers per wavefront instead of one per face, measured 8-16%
// faster than the per-face form), and empty slots are masked: an empty slot
// adds +-0 to the gradient sums, which leaves them unchanged (a sum starts at
// +0 and is -0 only when both terms are, so it is never -0), and leaves the
// min/max and the limiter as they were.
// kFaceVals (inviscid second order): each slot's reconstructed value -- the
// extrapolation v + phi ((x_f - x_c) . grad) of Flux.h:107-130, the same
// expression in the same order as the flux kernel's otherwise -- is stored
// in fv[kMaxSlots * 5][nc], so the flux kernel gathers 5 values per face
// side instead of 25; the gradient record is then stored (to rec, kRec
// values per cell) only when rec is non-null, by the getter's call.
// Otherwise the record is left in out[kRec] for the caller, which stores a
// wavefront's records together (store_records).  First order has zero
// limiters.
// kAxis: 0 the general Green-Gauss loop, 1 the axis-aligned one (the mesh
// has DevMesh::axis set), 2 whichever m.axis selects at run time.
template <bool kSecondOrder, bool kFaceVals, class Limiter, class T, int kAxis = 0>
__device__ __forceinline__ void grad_cell(const DevMesh& m, const int* __restrict__ nbr, const T* __restrict__ V,
                                          T* __restrict__ rec, T* __restrict__ fv, int c, T* out) {
  const int nc = m.nc;
  // slot entries: (geometry record << 1) | side, -1 empty
  int ent[kMaxSlots], nb[kMaxSlots];
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    ent[s] = m.sgeo[(size_t)s * nc + c];
    nb[s] = nbr[(size_t)s * nc + c];
  }
  T vc[5];
  load5(V, nc, c, vc);
  const T vol = T(m.vol[c]);
  const T rvol = T(1) / vol;

  T g[5][3];
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) g[k][d] = T(0.0);
  T mn[5], mx[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    mn[k] = T(1.0e300);
    mx[k] = -T(1.0e300);
  }

  // kAxis 1 is launched only on meshes without empty slots (DevMesh::full),
  // so there every slot is a face and the Green-Gauss loop's masks fold away
  // (FP64 viscous kernel 168 -> 158 registers, no spills; folding the
  // limiter loop's as well made it spill 416 bytes per lane)
  constexpr bool kFull = kAxis == 1;
  auto rec_of = [&](int e) { return kFull ? e >> 1 : (e < 0 ? 0 : e >> 1); };
  T an[kMaxSlots][3], vnn[kMaxSlots][5];  // empty slots read record 0 and the cell itself
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    load_geo<T, false>(m, rec_of(ent[s]), an[s], nullptr, nullptr);
    load5(V, nc, nb[s], vnn[s]);
  }
  // FP32 face-value form: the limiter's first displacement is loaded here and
  // each later one a slot ahead, while the previous slot's limiter arithmetic
  // runs (512-case batched Sod workload: gradient kernel -3%; the other forms
  // lose occupancy to the extra registers and load each slot's displacement
  // when it is used)
  constexpr bool kPipe = kFaceVals && sizeof(T) == 4;
  T dnext[kPipe ? 3 : 1];
  if constexpr (kPipe) load_geo_disp(m, rec_of(ent[0]), ent[0] & 1, dnext);
  // Axis-aligned faces (every face of a box mesh): for a zero component n_d
  // of the area vector the Green-Gauss term gg = 0.5 (v_L + v_R) n_d is +-0
  // or NaN, and so are gg / vol and gg * (1 / vol) (the volume is nonzero and
  // its reciprocal finite).  Adding +-0 leaves a sum unchanged (it starts at
  // +0 and is never -0), so on a mesh whose area vectors have at most one
  // nonzero component (DevMesh::axis, checked on the host with the volumes)
  // a slot divides only that component: the same results with a third of
  // the divisions.  FP64 adds gg itself for the other two, with the side's
  // sign in the area vector (-(a b) = a (-b) and (-x) / vol = -(x / vol)
  // exactly); FP32 multiplies them by 1 / vol, the form that keeps the
  // multi-case kernel within its registers.  The multi-case and persistent
  // kernels (kAxis 2) have both loops and take this one in the FP32
  // face-value form only: in FP64 both loops together spill.  The FP32
  // viscous form multiplies by the reciprocal throughout, so there is no
  // division to save.
  constexpr bool kAxisForm = sizeof(T) == 4 ? kFaceVals : kAxis == 1;
  const bool axis = kAxisForm && (kAxis == 1 || (kAxis == 2 && m.axis));
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    const bool ok = kFull || ent[s] >= 0;
    const int e = ok ? ent[s] : 0;
    const int side = e & 1;
    const T* n = an[s];
    const T* vn = vnn[s];
    if (axis && sizeof(T) == 8) {  // uniform: a property of the mesh
      T sn[3];
      bool nz[3];
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        sn[d] = side ? -n[d] : n[d];
        nz[d] = n[d] != T(0);
      }
      // the nonzero component, exactly: the other two are +-0 (when all
      // three are zero its sign may differ, and a +-0 term changes no sum)
      const T snd = (sn[0] + sn[1]) + sn[2];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const T vl = side ? vn[k] : vc[k];
        const T vr = side ? vc[k] : vn[k];
        const T base = ok ? T(0.5) * (vl + vr) : T(0);
        const T q = (base * snd) / vol;
#pragma unroll
        for (int d = 0; d < 3; ++d) g[k][d] += nz[d] ? q : base * sn[d];
      }
    } else if (axis) {
      const int dn = n[1] != T(0) ? 1 : (n[2] != T(0) ? 2 : 0);  // the nonzero component (0 if none)
      // n[dn], exactly (a sum instead of a select by dn, which the compiler
      // would turn into an indexed load from memory)
      const T nd = (n[0] + n[1]) + n[2];
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const T vl = side ? vn[k] : vc[k];
        const T vr = side ? vc[k] : vn[k];
        const T base = T(0.5) * (vl + vr);
        const T gn = ok ? base * nd : T(0);
        const T q = side ? -gn / vol : gn / vol;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const T gd = ok ? base * n[d] : T(0);
          g[k][d] += d == dn ? q : (side ? -gd * rvol : gd * rvol);
        }
      }
    } else {
#pragma unroll
      for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < 5; ++k) {
          const T vl = side ? vn[k] : vc[k];
          const T vr = side ? vc[k] : vn[k];
          const T gg = ok ? T(0.5) * (vl + vr) * n[d] : T(0);
          // FP32 viscous form: one reciprocal per cell (k_grad -6%; in the
          // face-value scheme it measured 2% slower, so that keeps the
          // division); FP64: the reference's division per face
          if (sizeof(T) == 4 && !kFaceVals)
            g[k][d] += (side ? -gg : gg) * rvol;
          else
            g[k][d] += side ? -gg / vol : gg / vol;
        }
    }
    if (kSecondOrder) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const T vl = side ? vn[k] : vc[k];
        const T vr = side ? vc[k] : vn[k];
        const T fmin = vr < vl ? vr : vl;
        const T fmax = vr > vl ? vr : vl;
        mn[k] = ok ? ma::smin(mn[k], fmin) : mn[k];
        mx[k] = ok ? ma::smax(mx[k], fmax) : mx[k];
      }
    }
  }
  // green_gauss_gradient_sum: gradient = 0; gradient += slot (GreenGauss.h:244-269)
#pragma unroll
  for (int k = 0; k < 5; ++k)
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      T x = 0;
      x += g[k][d];
      g[k][d] = x;
      if constexpr (kFaceVals) {
        if (rec) rec[(size_t)c * kRec + k * 3 + d] = x;  // the getter's call only
      } else {
        out[k * 3 + d] = x;
      }
    }
  if (!kSecondOrder) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      if constexpr (!kFaceVals) out[15 + k] = T(0);
    }
    return;
  }

  T lim[5];
  T dus[kFaceVals ? kMaxSlots : 1][5];  // the extrapolation increments, kept for the face values
#pragma unroll
  for (int k = 0; k < 5; ++k) lim[k] = T(1.0);
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    const bool ok = ent[s] >= 0;
    const int e = ok ? ent[s] : 0;
    T disp[3], dist = 0;
    if constexpr (kPipe) {
#pragma unroll
      for (int d = 0; d < 3; ++d) disp[d] = dnext[d];
      if (s + 1 < kMaxSlots) load_geo_disp(m, rec_of(ent[s + 1]), ent[s + 1] & 1, dnext);
#pragma unroll
      for (int d = 0; d < 3; ++d) dist += disp[d] * disp[d];
    } else {
      load_geo_disp(m, e >> 1, e & 1, disp);
#pragma unroll
      for (int d = 0; d < 3; ++d) dist += disp[d] * disp[d];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      T du = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) du += disp[d] * g[k][d];
      if (kFaceVals) dus[kFaceVals ? s : 0][k] = du;
      const T dumax = mx[k] - vc[k];
      const T dumin = mn[k] - vc[k];
      const T l = Limiter::limit(dumax, dumin, du, dist);
      lim[k] = ok ? ma::smin(lim[k], l) : lim[k];
    }
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    if constexpr (kFaceVals) {
      if (rec) rec[(size_t)c * kRec + 15 + k] = lim[k];
    } else {
      out[15 + k] = lim[k];
    }
  }
  if (kFaceVals) {
#pragma unroll
    for (int s = 0; s < kMaxSlots; ++s) {
      const int e = ent[s];
      if (!kFull && e < 0) break;
#pragma unroll
      for (int k = 0; k < 5; ++k) fv[(size_t)(s * 5 + k) * nc + c] = vc[k] + dus[kFaceVals ? s : 0][k] * lim[k];
    }
  }
}

// Waves per SIMD the FP64 second-order gradient kernels are register-budgeted
// for: the compiler's own choice can land just over the 168 registers that
// allow 3 (then 2 waves run); other forms no constraint (the FP32 viscous
// kernel runs 5 waves at 84 registers; a 6-wave budget spills and was 14%
// slower on the 512x256x64 plate).  The axis-aligned FP64 viscous kernel
// fits 3 waves in 158 registers with its empty-slot masks folded; before
// that it spilled 12 bytes per lane and was still 10% faster than a 2-wave
// build without spills (0.222 against 0.245 ms on the 256x128x64 plate).
template <bool kSecondOrder, class T>
constexpr int kGradWaves = sizeof(T) == 8 && kSecondOrder ? 3 : 1;

// Flux of internal face f between cells L and Rr: compute_face_flux's body
// (Flux.h:89-160) without the scatter -- primitives of both cells, the
// second-order extrapolation v += phi * ((x_f - x_c) . grad) with the cell
// limiters (Flux.h:107-130; x_f - x_c from the displacement planes, side 0
// the left cell, 1 the right), Roe, minus the viscous flux from
// face-averaged gradients.  Every input is loaded before the first is used.
// Inviscid second order takes both sides' extrapolated values from fv (slot
// sL of cell L, slot sR of cell Rr), stored by the gradient kernel with the
// same arithmetic.  The result is a pure function of (f, V, grad, phi), so
// every workgroup that evaluates face f gets the same bits.
template <bool kSecondOrder, class Inv, class Visc, class T>
__device__ __forceinline__ void interior_flux(const DevMesh& m, const T* __restrict__ V,
                                              const T* __restrict__ rec, const T* __restrict__ fv, int f,
                                              int g, int L, int sL, int Rr, int sR, const Inv& inv,
                                              const Visc& visc, T* F) {
  constexpr bool kFaceVals = kSecondOrder && !Visc::isViscous;
  constexpr bool kRecon = kSecondOrder && !kFaceVals;  // extrapolated here
  constexpr bool kGrad = Visc::isViscous;
  const int nc = m.nc;
  T vl[5], vr[5], a[3];
  T gl[kGrad ? 15 : 1], gr[kGrad ? 15 : 1], pl[5], pr[5], dl[3], dr[3];
  if (kFaceVals) {
    load5(fv + (size_t)sL * 5 * nc, nc, L, vl);
    load5(fv + (size_t)sR * 5 * nc, nc, Rr, vr);
  } else {
    load5(V, nc, L, vl);
    load5(V, nc, Rr, vr);
  }
  load_geo<T, kRecon>(m, g, a, dl, dr);  // the face's geometry record
  if (kGrad) {  // (kRecon implies kGrad)
    T ra[kRec], rb[kRec];
    load_rec(rec, L, ra);
    load_rec(rec, Rr, rb);
#pragma unroll
    for (int j = 0; j < 15; ++j) {
      gl[j] = ra[j];
      gr[j] = rb[j];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      pl[k] = ra[15 + k];
      pr[k] = rb[15 + k];
    }
  }
  const ma::FaceFrameT<T> fr = frame_of<T>(m, f, a);
  if (kRecon) {
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      T tl = 0, tr = 0;
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        tl += dl[d] * gl[k * 3 + d];
        tr += dr[d] * gr[k * 3 + d];
      }
      vl[k] += tl * pl[k];
      vr[k] += tr * pr[k];
    }
  }
  ma::inviscid_flux(inv, vl, vr, F, fr);
  if (Visc::isViscous) {
    T vf[5], gf[5][3], fv[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      vf[k] = T(0.5) * (vl[k] + vr[k]);
#pragma unroll
      for (int d = 0; d < 3; ++d) gf[k][d] = T(0.5) * (gl[k * 3 + d] + gr[k * 3 + d]);
    }
    visc.compute_flux(gf, vf, fr.n, fv);
#pragma unroll
    for (int k = 0; k < 5; ++k) F[k] -= fv[k];
  }
}

// Boundary faces, one thread each: the contribution each adds to its cell's
// flux sum, from the boundary policy of the face's kind (BCs::Extrapolate,
// Tangent, NoSlip, Inflow; the reference's functors Extrapolate_BC.h:69-101,
// Tangent_BC.h:71-122, NoSlip_BC.h:76-156, Inflow_BC.h:72-110) with the
// inviscid policy; the no-slip wall's viscous part is always the Newtonian
// flux, even in inviscid runs (TimeSolverExplicitRK4.h:457-460).  Written to
// bc[5][nbf].
template <class Inv, class BCs, class T>
__device__ __forceinline__ void bc_face(const DevMesh& m, const T* __restrict__ V, const Inv& inv,
                                        double prandtl, const Inflow* __restrict__ inflows,
                                        const int* __restrict__ bc_case, T* __restrict__ bc, int i) {
  const int nbf = m.nf - m.nif;
  const int f = m.nif + i;
  const int c = m.fl[f];
  const int kind = m.fkind[f];
  T vc[5], out[5];
  load5(V, m.nc, c, vc);
  const ma::FaceFrameT<T> fr = load_frame<T>(m, f);
  const ma::NewtonianViscousFlux wallv{prandtl};
  switch (kind) {
    case MA_BC_EXTRAPOLATE:
      typename BCs::Extrapolate{}(vc, fr, nullptr, T(0), inv, wallv, out);
      break;
    case MA_BC_TANGENT:
      typename BCs::Tangent{}(vc, fr, nullptr, T(0), inv, wallv, out);
      break;
    case MA_BC_INFLOW:
      typename BCs::Inflow{}(vc, fr, inflows[bc_case[i]].U, T(0), inv, wallv, out);
      break;
    default:  // MA_BC_NOSLIP (kinds are checked at setup)
      typename BCs::NoSlip{}(vc, fr, nullptr, T(m.fwall[f]), inv, wallv, out);
      break;
  }
#pragma unroll
  for (int k = 0; k < 5; ++k) bc[(size_t)k * nbf + i] = out[k];
}

template <class Inv, class T>
__global__ __launch_bounds__(kBlock) void k_bc(DevMesh m, const T* __restrict__ V, Inv inv, double prandtl,
                                               const Inflow* __restrict__ inflows,
                                               const int* __restrict__ bc_case, T* __restrict__ bc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < m.nf - m.nif) bc_face<Inv, ma::ReferenceBoundaries>(m, V, inv, prandtl, inflows, bc_case, bc, i);
}

// Workgroups past the per_xcd * kXcds cell blocks take the boundary faces,
// k_bc's work: it reads the same stage primitives, so one launch does both.
template <bool kSecondOrder, bool kFaceVals, class Limiter, class T, int kAxis, class Inv>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kGradWaves<kSecondOrder, T>)))
void k_grad(DevMesh m, const int* __restrict__ nbr, const T* __restrict__ V, T* __restrict__ rec,
            T* __restrict__ fv, int per_xcd, Inv inv, double prandtl, const Inflow* __restrict__ inflows,
            const int* __restrict__ bc_case, T* __restrict__ bc) {
  __shared__ T stage[kFaceVals ? 1 : kBlock * kStageRow<T>];
  const int ncb = per_xcd * kXcds;
  if ((int)blockIdx.x >= ncb) {  // a boundary-face block (uniform over the workgroup)
    const int i = ((int)blockIdx.x - ncb) * kBlock + threadIdx.x;
    if (i < m.nf - m.nif) bc_face<Inv, ma::ReferenceBoundaries>(m, V, inv, prandtl, inflows, bc_case, bc, i);
    return;
  }
  // XCD-aware: each XCD takes a contiguous run of blocks (see k_flux)
  const int b = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  const int c = b * kBlock + threadIdx.x;
  T out[kRec];
  if (c < m.nc) grad_cell<kSecondOrder, kFaceVals, Limiter, T, kAxis>(m, nbr, V, rec, fv, c, out);
  if constexpr (!kFaceVals) {
    if (b * kBlock < m.nc) store_records(rec, stage, out, b * kBlock, m.nc);  // uniform over the workgroup
  }
}

// Residual and RK accumulate of one tile of at most kTile cells:
// compute_face_flux's atomic scatter (Flux.h:162-170), apply_cell_flux
// (Flux.h:188-229) and update<Device> with beta_k (TimeSolverExplicitRK4.h:483-485).
// Phase 1: the workgroup evaluates every internal face touching its cells
// once into LDS (faces between two tiles are evaluated by both).  Phase 2:
// every cell adds its face fluxes from LDS and its boundary contributions in
// the reference's one-thread order -- internal faces in processing order,
// then boundary faces by kind (Extrapolate, Tangent, NoSlip, Inflow) and group.
// The sums, the residual and the accumulator are FP64 whatever T is.
// Slot entries: (index << 3) | code, code 0 = left side, 1 = right side of an
// internal face (index = LDS slot), 4 + kind - 1 = boundary face of that kind
// (index = boundary-face number).
template <bool kSecondOrder, class Inv, class Visc, class T>
// Un, Vnext, alpha_next, last: with Vnext set, the cell's primitives for the
// next stage, primitives(base + alpha_next R) -- update<Device> and
// ComputePrimitives of that stage, the same arithmetic as k_prim -- are
// written to Vnext, base being Un on stages 0-2 and, on the last stage, the
// new state just written to Uacc_out (the next step's stage 0, alpha 0).
// R itself is stored on the last stage only.
__device__ __forceinline__ void flux_tile(const DevMesh& m, const TileMap& t, int tile, T* lds,
                                          const T* __restrict__ V, const T* __restrict__ rec,
                                          const T* __restrict__ fv, const T* __restrict__ bc,
                                          const double* Uacc_in, double* Uacc_out, double* __restrict__ R,
                                          const double* __restrict__ dt_dev, double beta, const Inv& inv,
                                          const Visc& visc, const double* __restrict__ Un,
                                          T* __restrict__ Vnext, double alpha_next, int last) {
  const int f0 = t.ptr[tile];
  const int nfl = t.ptr[tile + 1] - f0;
  const int mf = t.max_faces;
  // the next face's indices are fetched while the current face is evaluated
  int f = 0;
  int2 lr = make_int2(0, 0);
  int g = 0;
  if (threadIdx.x < nfl) {
    f = t.faces[f0 + threadIdx.x];
    lr = t.lr[f0 + threadIdx.x];
    g = t.fgid[f0 + threadIdx.x];
  }
  for (int i = threadIdx.x; i < nfl; i += kTile) {
    const int fc = f, gc = g;
    const int2 lrc = lr;
    if (i + kTile < nfl) {
      f = t.faces[f0 + i + kTile];
      lr = t.lr[f0 + i + kTile];
      g = t.fgid[f0 + i + kTile];
    }
    T F[5];
    // slots are packed into the cell entries only for the face-value scheme
    constexpr bool kFaceVals = kSecondOrder && !Visc::isViscous;
    interior_flux<kSecondOrder>(m, V, rec, fv, fc, gc, kFaceVals ? cell_of(lrc.x) : lrc.x, slot_of(lrc.x),
                                kFaceVals ? cell_of(lrc.y) : lrc.y, slot_of(lrc.y), inv, visc, F);
#pragma unroll
    for (int k = 0; k < 5; ++k) lds[k * mf + i] = F[k];
  }
  __syncthreads();
  const int c = t.cptr[tile] + threadIdx.x;
  if (c >= t.cptr[tile + 1]) return;  // (the caller's barrier is reached by every thread)
  const int nc = m.nc;
  const int nbf = m.nf - m.nif;
  // The cell's accumulator and base state are loaded before anything is
  // stored: Uacc_in is Uacc_out or a different array, never a shifted view,
  // so reading all five planes first reads the same values, and the loads
  // need not wait behind the stores.
  double uin[5], un[5];
  load5(Uacc_in, nc, c, uin);
  if (Vnext && !last) load5(Un, nc, c, un);
  const double dt = dt_dev[t.tcase[tile]];
  const double vol = m.vol[c];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int ent[kMaxSlots];
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) ent[s] = t.slot[(size_t)s * nc + c];
  bool any_bc = false;
#pragma unroll
  for (int s = 0; s < kMaxSlots; ++s) {
    const int e = ent[s];
    if (e < 0) break;
    const int code = e & 7, j = e >> 3;
    if (code < 2) {
#pragma unroll
      for (int k = 0; k < 5; ++k) {
        const double x = double(lds[k * mf + j]);
        acc[k] += code ? x : -x;
      }
    } else {
      any_bc = true;
    }
  }
  if (any_bc) {
#pragma unroll 1
    for (int code = 4; code < 8; ++code)
#pragma unroll 1
      for (int s = 0; s < kMaxSlots; ++s) {
        const int e = ent[s];
        if (e < 0) break;
        if ((e & 7) != code) continue;
        const int j = e >> 3;
#pragma unroll
        for (int k = 0; k < 5; ++k) acc[k] += double(bc[(size_t)k * nbf + j]);
      }
  }
  double rr[5], unew[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    double r = 0.0;
    r = r + dt / vol * acc[k];
    rr[k] = r;
    unew[k] = uin[k] + beta * r;
    Uacc_out[(size_t)k * nc + c] = unew[k];
  }
  if (Vnext) {
    double u[5], v[5];
#pragma unroll
    for (int k = 0; k < 5; ++k) u[k] = (last ? unew[k] : un[k]) + alpha_next * rr[k];
    ma::primitives(u, v);
#pragma unroll
    for (int k = 0; k < 5; ++k) Vnext[(size_t)k * nc + c] = T(v[k]);
  }
  if (last) {
#pragma unroll
    for (int k = 0; k < 5; ++k) R[(size_t)k * nc + c] = rr[k];
  }
}

// Waves per SIMD the flux kernels are register-budgeted for: FP64 3 (the
// compiler's own choice lands one register over the 168 that allow 3, at 2
// waves, measured 10% slower); FP32 5 (the viscous multi-case kernel would
// otherwise take 98 registers, 4 waves; the others use fewer than 5 waves'
// budget anyway).
template <class T>
constexpr int kFluxWaves = sizeof(T) == 8 ? 3 : 5;

template <bool kSecondOrder, class Inv, class Visc, class T>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(kFluxWaves<T>)))
void k_flux(DevMesh m, TileMap t, const T* __restrict__ V,
            const T* __restrict__ rec,  // gradient records, or the face values (inviscid second order)
            const T* __restrict__ bc,
            const double* Uacc_in,  // may alias Uacc_out
            double* Uacc_out, double* __restrict__ R,
            const double* __restrict__ dt_dev, double beta,
            Inv inv, Visc visc, const double* __restrict__ Un, T* Vnext,
            double alpha_next, int last) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);  // 5 planes of t.max_faces
  // XCD-aware order: the workgroups an XCD receives (every kXcds-th) take a
  // contiguous run of tiles, so neighbouring tiles share that XCD's L2.  With
  // fewer workgroups than tiles (gridDim.x a multiple of kXcds) each one
  // walks its XCD's run with stride gridDim.x / kXcds, which bounds how many
  // tiles an XCD has in flight.
  const int xcd = blockIdx.x % kXcds, stride = gridDim.x / kXcds;
  for (int j = blockIdx.x / kXcds; j < t.per_xcd; j += stride) {
    const int tile = xcd * t.per_xcd + j;
    if (tile >= t.ntiles) break;  // uniform over the workgroup
    // the face-value scheme reads no gradients: its face values come in rec
    flux_tile<kSecondOrder, Inv, Visc, T>(m, t, tile, lds, V, rec, rec, bc, Uacc_in, Uacc_out, R, dt_dev, beta,
                                          inv, visc, Un, Vnext, alpha_next, last);
    __syncthreads();  // the LDS is reused by the next tile
  }
}

// Stable time step for the CFL extension, per cell:
//   lc = |u| S_x + |v| S_y + |w| S_z + c N1,  lv = N2,
//   dt_c = vol / (lc + 4 kappa lv / vol),  kappa = max(4/3, gamma/Pr) mu / rho
// (kappa = 0 for inviscid runs: convective plus viscous spectral radius),
// with the cell's face geometry summed once at setup over its slot list:
// S_d = sum_f |n_f,d|, N1 = sum_f |n_f|, N2 = sum_f |n_f|^2.  For faces
// normal to an axis (every box mesh, stretched or not) |u| S_x + |v| S_y +
// |w| S_z is sum_f |u.n_f| exactly; for others it bounds it from above (a
// smaller, still stable step).  Reduced to the minimum per block.
// One workgroup of kTile threads per tile (a tile never spans two cases);
// red: kTile / 64 doubles of shared memory.
__device__ __forceinline__ void dt_tile(const DevMesh& m, const TileMap& t, int tile,
                                        const double* __restrict__ Un, int viscous, double prandtl,
                                        double* __restrict__ partial, double* red) {
  const int c = t.cptr[tile] + threadIdx.x;
  double dtc = 1e300;
  if (c < t.cptr[tile + 1]) {
    const int nc = m.nc;
    double u[5], v[5], gm[5];
    load5(Un, nc, c, u);
    load5(m.dtg, nc, c, gm);
    ma::primitives(u, v);
    const double snd = sqrt(1.4 * (v[0] * ma::kRgas * v[4]) / v[0]);
    const double lc = fabs(v[1]) * gm[0] + fabs(v[2]) * gm[1] + fabs(v[3]) * gm[2] + snd * gm[3];
    const double vol = m.vol[c];
    double lam = lc;
    if (viscous) {
      const double kfac = ma::smax(4.0 / 3.0, 1.4 / prandtl);
      const double mu = 1.458e-6 * v[4] * sqrt(v[4]) / (v[4] + 110.4);
      const double kappa = kfac * mu / v[0];
      lam = lam + 4.0 * kappa * gm[4] / vol;
    }
    dtc = vol / lam;
  }
  for (int off = 32; off > 0; off >>= 1) dtc = ma::smin(dtc, __shfl_down(dtc, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = dtc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = red[0];
    for (int w = 1; w < kTile / 64; ++w) x = ma::smin(x, red[w]);
    partial[tile] = x;
  }
}

__global__ __launch_bounds__(kTile) void k_dt_partial(DevMesh m, TileMap t,
                                                      const double* __restrict__ Un, int viscous,
                                                      double prandtl, double* __restrict__ partial) {
  __shared__ double red[kTile / 64];
  dt_tile(m, t, blockIdx.x, Un, viscous, prandtl, partial, red);
}

// One workgroup per case: dt = cfl * (minimum over the case's tiles) when the
// case runs at a CFL number (cfl > 0), else its fixed dt; then time += dt
// (TimeSolverExplicitRK4.h:343).
// One workgroup of kBlock threads per case; red: kBlock / 64 doubles of
// shared memory.
__device__ __forceinline__ void dt_case(const double* __restrict__ partial, const int* __restrict__ case_tiles,
                                        const double* __restrict__ cfl, double* __restrict__ dt_dev,
                                        double* __restrict__ time_dev, int k, double* red) {
  const double cf = cfl[k];
  double x = 1e300;
  if (cf > 0.0) {
    // eight loads in flight per thread (a case can have tens of thousands of
    // tiles); the minimum of positive finite bounds does not depend on the
    // order it is taken in, and the order is fixed in any case
    const int hi = case_tiles[k + 1], n = blockDim.x;
    int i = case_tiles[k] + threadIdx.x;
    for (; i + 7 * n < hi; i += 8 * n) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = partial[i + u * n];
#pragma unroll
      for (int u = 0; u < 8; ++u) x = ma::smin(x, v[u]);
    }
    for (; i < hi; i += n) x = ma::smin(x, partial[i]);
  }
  for (int off = 32; off > 0; off >>= 1) x = ma::smin(x, __shfl_down(x, off, 64));
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = x;
  __syncthreads();
  if (threadIdx.x == 0) {
    double y = red[0];
    for (int w = 1; w < kBlock / 64; ++w) y = ma::smin(y, red[w]);
    const double dt = cf > 0.0 ? cf * y : dt_dev[k];
    dt_dev[k] = dt;
    time_dev[k] = time_dev[k] + dt;
  }
}

__global__ __launch_bounds__(kBlock) void k_dt_final(const double* __restrict__ partial,
                                                     const int* __restrict__ case_tiles,
                                                     const double* __restrict__ cfl,
                                                     double* __restrict__ dt_dev,
                                                     double* __restrict__ time_dev) {
  __shared__ double red[kBlock / 64];
  dt_case(partial, case_tiles, cfl, dt_dev, time_dev, blockIdx.x, red);
}

// Residual norms: per tile, the five sums of R_k^2 over its cells (wavefront
// shuffles, then the tile's waves in order); per case, the tiles' sums in
// tile order and the square root.  Fixed order, so repeatable bit for bit.
__global__ __launch_bounds__(kTile) void k_norm_partial(TileMap t, int nc, const double* __restrict__ R,
                                                        double* __restrict__ partial) {
  __shared__ double red[5][kTile / 64];
  const int tile = blockIdx.x;
  const int c = t.cptr[tile] + threadIdx.x;
  const bool in = c < t.cptr[tile + 1];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    const double r = in ? R[(size_t)k * nc + c] : 0.0;
    double x = r * r;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_down(x, off, 64);
    if (lane == 0) red[k][wave] = x;
  }
  __syncthreads();
  if (threadIdx.x < 5) {
    double y = 0.0;
    for (int w = 0; w < kTile / 64; ++w) y += red[threadIdx.x][w];
    partial[(size_t)tile * 5 + threadIdx.x] = y;
  }
}

__global__ void k_norm_final(const double* __restrict__ partial, const int* __restrict__ case_tiles,
                             int ncases, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ncases * 5) return;
  const int k = i / 5, v = i % 5;
  double y = 0.0;
  for (int b = case_tiles[k]; b < case_tiles[k + 1]; ++b) y += partial[(size_t)b * 5 + v];
  out[i] = sqrt(y);
}

__global__ __launch_bounds__(kBlock) void k_count_nonfinite(int n, const double* __restrict__ a,
                                                            unsigned long long* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool bad = i < n && !isfinite(a[i]);
  const unsigned long long mask = __ballot(bad);
  if ((threadIdx.x & 63) == 0 && mask) atomicAdd(out, (unsigned long long)__popcll(mask));
}

// Plain streaming copy, the bandwidth ceiling the kernels above are compared
// with: one 16-byte load and store per thread over a grid covering the array
// (on MI355X this form reaches ~6.3 TB/s read + write, a grid-stride loop
// over 16 blocks per CU ~4.5 TB/s).
__global__ __launch_bounds__(kBlock) void k_copy(long long n2, const double2* __restrict__ a,
                                                 double2* __restrict__ b) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n2) b[i] = a[i];
}

// ------------------------------------------------- multi-case launches --
// Several independent single-case runs advanced by the same launches (the
// served form's continuous batching): each workgroup finds its case from a
// prefix table and runs the single-case body on that case's own arrays.

struct CaseView {
  DevMesh m;
  TileMap t;
  double* U[2];
  double* R;
  void *V[2], *grad, *bc;  // arithmetic type of the launch; V[stage & 1] holds the stage's primitives;
                           // grad: gradient records
  void* fv;                      // extrapolated face values (inviscid second order)
  double *dt, *time, *partial;
  const double* cfl;
  const int* case_tiles;  // {0, ntiles}
  const Inflow* inflow;
  const int* bc_case;
  const int* nbr;  // neighbour-cell planes (grad_cell)
  int parity;  // U[parity ^ step parity] holds the case's sol_n
  int prim0;   // 1: V[0] does not hold the primitives of the case's sol_n yet
};

// Prefix table: start[k] = first global block of case k, start[n] = total.
__device__ __forceinline__ int find_case(const int* __restrict__ start, int n, int g) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Stage 0, for the cases whose V[0] is not current (new in the group; the
// flux kernel of every stage writes the next stage's primitives, the last
// stage the next step's).
template <class T>
__global__ __launch_bounds__(kBlock) void k_prim_multi(const CaseView* __restrict__ cv, const int* __restrict__ start,
                                                       int n, double alpha, int parity) {
  const int g = blockIdx.x;
  const int k = find_case(start, n, g);
  const CaseView& v = cv[k];
  if (!v.prim0) return;
  const int c = (g - start[k]) * kBlock + threadIdx.x;
  if (c < v.m.nc) prim_cell(v.m.nc, v.U[v.parity ^ parity], v.R, alpha, static_cast<T*>(v.V[0]), c);
}

template <bool kSecondOrder, bool kFaceVals, class Limiter, class T>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kGradWaves<kSecondOrder, T>)))
void k_grad_multi(const CaseView* __restrict__ cv, const int* __restrict__ start,
                                                       int n, int per_xcd, int vsel, int store_grad) {
  const int g = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (g >= start[n]) return;
  const int k = find_case(start, n, g);
  const CaseView& v = cv[k];
  const int c0 = (g - start[k]) * kBlock, c = c0 + threadIdx.x;
  T out[kRec];
  if (c < v.m.nc)
    grad_cell<kSecondOrder, kFaceVals, Limiter, T, 2>(v.m, v.nbr, static_cast<const T*>(v.V[vsel]),
                                                   store_grad ? static_cast<T*>(v.grad) : nullptr,
                                                   static_cast<T*>(v.fv), c, out);
  if constexpr (!kFaceVals) {
    __shared__ T stage[kBlock * kStageRow<T>];
    if (store_grad) store_records(static_cast<T*>(v.grad), stage, out, c0, v.m.nc);  // uniform over the workgroup
  }
}

template <class Inv, class T>
__global__ __launch_bounds__(kBlock) void k_bc_multi(const CaseView* __restrict__ cv, const int* __restrict__ start,
                                                     int n, Inv inv, double prandtl, int vsel) {
  const int g = blockIdx.x;
  const int k = find_case(start, n, g);
  const CaseView& v = cv[k];
  const int i = (g - start[k]) * kBlock + threadIdx.x;
  if (i < v.m.nf - v.m.nif)
    bc_face<Inv, ma::ReferenceBoundaries>(v.m, static_cast<const T*>(v.V[vsel]), inv, prandtl, v.inflow,
                                          v.bc_case, static_cast<T*>(v.bc), i);
}

template <bool kSecondOrder, class Inv, class Visc, class T>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(kFluxWaves<T>)))
void k_flux_multi(const CaseView* __restrict__ cv, const int* __restrict__ start,
                  int n, int per_xcd, int parity, int irk, double beta,
                  double alpha_next, Inv inv, Visc visc) {
  extern __shared__ __align__(16) unsigned char lds_raw[];
  const int g = (blockIdx.x % kXcds) * per_xcd + blockIdx.x / kXcds;
  if (g >= start[n]) return;  // whole workgroup, before any barrier
  const int k = find_case(start, n, g);
  const CaseView& v = cv[k];
  const int cur = v.parity ^ parity;
  double* Uacc = v.U[1 - cur];
  flux_tile<kSecondOrder, Inv, Visc, T>(v.m, v.t, g - start[k], reinterpret_cast<T*>(lds_raw),
                                   static_cast<const T*>(v.V[irk & 1]), static_cast<const T*>(v.grad),
                                   static_cast<const T*>(v.fv), static_cast<const T*>(v.bc),
                                   irk == 0 ? v.U[cur] : Uacc, Uacc, v.R, v.dt, beta, inv, visc, v.U[cur],
                                   static_cast<T*>(v.V[1 - (irk & 1)]), alpha_next, irk == 3);
}

__global__ __launch_bounds__(kTile) void k_dt_partial_multi(const CaseView* __restrict__ cv,
                                                            const int* __restrict__ start, int n, int parity,
                                                            int viscous, double prandtl) {
  __shared__ double red[kTile / 64];
  const int g = blockIdx.x;
  const int k = find_case(start, n, g);
  const CaseView& v = cv[k];
  if (!v.m.dtg) return;  // a fixed-dt case: no geometry, its partial minima are not read
  dt_tile(v.m, v.t, g - start[k], v.U[v.parity ^ parity], viscous, prandtl, v.partial, red);
}

__global__ __launch_bounds__(kBlock) void k_dt_final_multi(const CaseView* __restrict__ cv) {
  __shared__ double red[kBlock / 64];
  const CaseView& v = cv[blockIdx.x];
  dt_case(v.partial, v.case_tiles, v.cfl, v.dt, v.time, 0, red);
}

// ------------------------------------------------------ persistent steps --
// nsteps whole RK4 steps of every case of a run in one cooperative launch:
// the workgroups stay resident and walk each phase's tiles or cell blocks
// (the same device functions, in the same order, as the per-kernel launches),
// with a grid-wide barrier between phases where the launches had a kernel
// boundary.  Same bits as the per-kernel launches.

struct PersistArgs {
  DevMesh m;
  TileMap t;
  const int* nbr;
  void* V[2];  // stage primitives: V (stages 0, 2), V2 (stages 1, 3)
  void* grad;  // gradient records
  void* fv;    // face values (inviscid second order)
  void* bc;
  double* U[2];
  double* R;
  double* dt;
  double* time;
  const double* cfl;
  double* partial;
  const int* case_tiles;
  const Inflow* inflow;
  const int* bc_case;
  unsigned* gsync;  // [0] arrivals, [1] generation, [2] wait exceeded; zero at launch
  long long wait_ticks;  // longest barrier wait, in wall-clock ticks
  int ncases, nsteps, cur0, prim0, any_cfl, viscous, per_blocks;
  double prandtl;
};

// Grid-wide barrier: every workgroup of the launch arrives once per call.
// The arrivals are counted in gsync[0]; the last one resets the count and
// advances the generation gsync[1], for which the others wait.  The fences
// make each workgroup's earlier writes visible to the whole device (every
// XCD's L2) before its arrival, and the new generation's observers see them.
// The wait is bounded: after wait_ticks of the wall clock gsync[2] is set,
// and a workgroup that finds it set leaves the kernel at this or any later
// barrier, so the launch ends.  Returns false when the workgroup must leave.
// Every thread of the workgroup calls it (it holds barriers).
__device__ __forceinline__ bool grid_sync(unsigned* gsync, unsigned& gen, long long wait_ticks, int* stop) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned want = gen + 1;
    int leave = 0;
    __threadfence();
    if (atomicAdd(&gsync[0], 1u) == gridDim.x - 1) {
      atomicExch(&gsync[0], 0u);
      __threadfence();
      atomicExch(&gsync[1], want);
    } else {
      const long long t0 = wall_clock64();
      while (__hip_atomic_load(&gsync[1], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) != want) {
        if (wall_clock64() - t0 > wait_ticks) {
          atomicOr(&gsync[2], 1u);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __threadfence();
    if (__hip_atomic_load(&gsync[2], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) leave = 1;
    gen = want;
    *stop = leave;
  }
  __syncthreads();
  return *stop == 0;
}

// Register budget of k_steps: the gradient and flux phases share one
// allocation, which in FP64 does not fit the 3-wave budget of the separate
// kernels without spilling, so FP64 runs 2 waves per SIMD (1 for the viscous
// second-order form, whose spills then go to the accumulation registers) with
// no scratch: the launch is refused if the kernel uses any, since scratch
// allocations can hold fewer waves resident than the occupancy the grid is
// sized for.
template <bool kViscousSO, class T>
constexpr int kStepsWaves = sizeof(T) == 8 ? (kViscousSO ? 1 : 2) : 1;

template <bool SO, class Inv, class Visc, class Lim, class T>
__global__ __launch_bounds__(kTile) __attribute__((amdgpu_waves_per_eu(kStepsWaves<SO && Visc::isViscous, T>)))
void k_steps(PersistArgs a, Inv inv, Visc visc) {
  // TimeSolverExplicitRK4.h:187-195: alpha = {0, 1/2, 1/2, 1}, beta = {1/6,
  // 1/3, 1/3, 1/6}, selected rather than indexed (an indexed array would sit
  // in scratch)
  auto beta = [](int k) { return k == 0 || k == 3 ? 1.0 / 6.0 : 1.0 / 3.0; };
  auto alpha_next = [](int k) { return k == 2 ? 1.0 : k == 3 ? 0.0 : 1.0 / 2.0; };
  constexpr bool FV = SO && !Visc::isViscous;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  T* lds = reinterpret_cast<T*>(lds_raw);  // flux face results, or the gradient records' staging
  __shared__ double red[kTile / 64];
  __shared__ int stop;
  const DevMesh& m = a.m;
  const int G = gridDim.x;
  // XCD-aware walks as in k_flux: the workgroups of one XCD (every kXcds-th)
  // take that XCD's contiguous run of tiles or blocks
  const int xcd = blockIdx.x % kXcds, j0 = blockIdx.x / kXcds, xstride = G / kXcds;
  const int nbf = m.nf - m.nif;
  unsigned gen = 0;
  for (int it = 0; it < a.nsteps; ++it) {
    const int cur = a.cur0 ^ (it & 1);
    double* Un = a.U[cur];
    double* Uacc = a.U[1 - cur];
    if (a.any_cfl) {
      for (int tile = blockIdx.x; tile < a.t.ntiles; tile += G) {
        dt_tile(m, a.t, tile, Un, a.viscous, a.prandtl, a.partial, red);
        __syncthreads();
      }
      if (!grid_sync(a.gsync, gen, a.wait_ticks, &stop)) return;
    }
    for (int k = blockIdx.x; k < a.ncases; k += G) {
      dt_case(a.partial, a.case_tiles, a.cfl, a.dt, a.time, k, red);
      __syncthreads();
    }
    if (!grid_sync(a.gsync, gen, a.wait_ticks, &stop)) return;
    if (it == 0 && a.prim0) {
      for (int c = blockIdx.x * kTile + threadIdx.x; c < m.nc; c += G * kTile)
        prim_cell(m.nc, Un, a.R, 0.0, static_cast<T*>(a.V[0]), c);
      if (!grid_sync(a.gsync, gen, a.wait_ticks, &stop)) return;
    }
    for (int irk = 0; irk < 4; ++irk) {
      const T* Vs = static_cast<const T*>(a.V[irk & 1]);
      T* Vn = static_cast<T*>(a.V[1 - (irk & 1)]);
      if (SO || Visc::isViscous) {
        T* rec = FV ? nullptr : static_cast<T*>(a.grad);
        for (int j = j0; j < a.per_blocks; j += xstride) {
          const int b = xcd * a.per_blocks + j;
          if (b * kBlock >= m.nc) break;  // uniform over the workgroup
          const int c = b * kBlock + threadIdx.x;
          T out[kRec];
          if (c < m.nc)
            grad_cell<SO, FV, Lim, T, 2>(m, a.nbr, Vs, rec, static_cast<T*>(a.fv), c, out);
          if constexpr (!FV) {
            store_records(rec, lds, out, b * kBlock, m.nc);
            __syncthreads();  // the staging rows are reused by the next block
          }
        }
        if (!grid_sync(a.gsync, gen, a.wait_ticks, &stop)) return;
      }
      if (nbf > 0) {
        for (int i = blockIdx.x * kTile + threadIdx.x; i < nbf; i += G * kTile)
          bc_face<Inv, ma::ReferenceBoundaries>(m, Vs, inv, visc.prandtl, a.inflow, a.bc_case,
                                                static_cast<T*>(a.bc), i);
        if (!grid_sync(a.gsync, gen, a.wait_ticks, &stop)) return;
      }
      const T* frec = static_cast<const T*>(FV ? a.fv : a.grad);
      for (int j = j0; j < a.t.per_xcd; j += xstride) {
        const int tile = xcd * a.t.per_xcd + j;
        if (tile >= a.t.ntiles) break;  // uniform over the workgroup
        flux_tile<SO, Inv, Visc, T>(m, a.t, tile, lds, Vs, frec, frec, static_cast<const T*>(a.bc),
                                    irk == 0 ? Un : Uacc, Uacc, a.R, a.dt, beta(irk), inv, visc, Un, Vn,
                                    alpha_next(irk), irk == 3 ? 1 : 0);
        __syncthreads();  // the LDS is reused by the next tile
      }
      if (!grid_sync(a.gsync, gen, a.wait_ticks, &stop)) return;
    }
  }
}

// ------------------------------------------------------------- host side --

#define MA_HIP(call)                                                                      \
  do {                                                                                    \
    hipError_t err_ = (call);                                                             \
    if (err_ != hipSuccess)                                                               \
      return ma::fail(MA_ERR_DEVICE, std::string(#call) + ": " + hipGetErrorString(err_)); \
  } while (0)

// Cells per flux tile: a brick of b[0] x b[1] x b[2] lattice cells (i, j, k
// extents, product <= kTile), or kTile consecutive cells of the caller's order
// when all three are zero.
struct TileShape {
  int b[3];
};
constexpr TileShape kDefaultShape = {{4, 8, 8}};
// The default for a run with fewer cells than one default tile per CU: with
// so few tiles a step's time is each tile's latency, and a quarter of the
// faces per 256-thread workgroup shortens it (64x16x16 first-order Sod,
// 16k cells: 77 -> 53 us per step as one graph, 81 -> 58 with a launch per
// kernel; 2x4x4 and 2x2x4 tiles 56 us, 2x2x2 93 us)
constexpr TileShape kSmallRunShape = {{4, 4, 4}};

bool valid_shape(const TileShape& t) {
  if (t.b[0] == 0 && t.b[1] == 0 && t.b[2] == 0) return true;
  return t.b[0] >= 1 && t.b[1] >= 1 && t.b[2] >= 1 && t.b[0] * t.b[1] * t.b[2] <= kTile;
}

struct Timing {
  bool on = false;
  struct Rec {
    const char* name;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
};

}  // namespace

struct ma_solver {
  ma_options opt;
  int device = 0;
  hipStream_t stream = nullptr;
  DevMesh m{};
  TileMap t{};
  std::vector<int> dev_of;  // caller's cell index -> device cell index
  std::vector<void*> allocs;
  size_t bytes = 0;
  double *U[2] = {nullptr, nullptr}, *R = nullptr;
  // Arithmetic-type buffers (double, or float for the FP32 form): primitives,
  // gradients, limiters, boundary contributions.
  void *V = nullptr, *V2 = nullptr, *bc = nullptr;
  void* grad = nullptr;  // gradient records (kRec values per cell: gradient, limiter)
  void* fv = nullptr;  // kMaxSlots x 5 planes: extrapolated face values (inviscid second order)
  bool fp32 = false;
  // Cases (independent meshes stepped together; a plain run is one case).
  int ncases = 1;
  std::vector<ma_options> case_opt;
  std::vector<int> case_cell_off;  // ncases + 1: cells of case k in the caller's union order
  bool any_cfl = false;
  double *dt_dev = nullptr, *time_dev = nullptr, *cfl_dev = nullptr;  // per case
  double* partial = nullptr;                                          // per tile
  int *cell_case = nullptr, *bc_case = nullptr, *case_tiles = nullptr;
  IcParams* ic_dev = nullptr;
  Inflow* inflow_dev = nullptr;
  unsigned long long* counter = nullptr;
  unsigned* gsync = nullptr;  // grid barrier of the persistent launch (k_steps)
  bool persistent = false;    // ma_solver_step runs its steps as one k_steps launch
  double *norm_partial = nullptr, *norm_out = nullptr;  // per tile x 5, per case x 5
  const double* xc = nullptr;  // cell centroids (3 planes), for the initial condition
  const int* org = nullptr;    // device cell -> caller's cell index
  const int* nbr = nullptr;    // kMaxSlots planes of nc: the cell across each slot (grad_cell)
  int cur = 0;  // U[cur] holds sol_n
  // V holds primitives(U[cur]) (written by the last stage of the previous
  // step, single- or multi-case); false after initialize and set_state
  bool v0_valid = false;
  // inviscid second order: the last stage's gradients and limiters are not in
  // grad/phi (the flux kernel reads face values); its primitives are in V2,
  // from which the getter computes them
  bool grads_stale = false;
  long long steps_done = 0;
  // HIP graphs of graph_steps whole RK4 steps (an even count, so U[cur] is
  // sol_n again after a replay), one per starting value of cur.
  int graph_steps = 0;
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  // thread that created the solver: the only one that captures or replays its graphs
  std::thread::id owner = std::this_thread::get_id();
  TileShape shape = kDefaultShape;  // flux tile shape in use
  int wg_per_cu = 0;                 // resident workgroups per CU of k_flux (0: full grid)
  int cus = 256;                     // compute units of the device
  bool all_lattice = true;
  double tuned_ms = 0.0;  // autotuner: measured ms per step per million cells of the chosen shape
  Timing timing;

  template <class T>
  ma_status alloc(T** p, size_t n) {
    return alloc_bytes((void**)p, n * sizeof(T));
  }
  // Device memory: from hipMalloc, or carved from a caller-provided region
  // (arena_base, 256-byte aligned), or only counted (counting: the pointers
  // are placeholders and nothing is copied).
  bool counting = false;
  char* arena_base = nullptr;
  size_t arena_size = 0, arena_used = 0;

  ma_status alloc_bytes(void** p, size_t nbytes) {
    *p = nullptr;
    if (nbytes == 0) nbytes = 8;
    if (counting || arena_base) {
      const size_t at = (arena_used + 255) & ~size_t(255);
      if (!counting && at + nbytes > arena_size)
        return ma::fail(MA_ERR_NOMEM, "device region too small for the case");
      *p = counting ? reinterpret_cast<void*>(at + 256) : arena_base + at;
      arena_used = at + nbytes;
      bytes += nbytes;
      return MA_OK;
    }
    hipError_t e = hipMalloc(p, nbytes);
    if (e != hipSuccess)
      return ma::fail(MA_ERR_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    allocs.push_back(*p);
    bytes += nbytes;
    return MA_OK;
  }

  // Launch wrapper: records a HIP event pair around the launch when profiling.
  template <class F>
  void launch(const char* name, F&& f) {
    if (timing.on) {
      Timing::Rec r{name, nullptr, nullptr};
      (void)hipEventCreate(&r.a);
      (void)hipEventCreate(&r.b);
      (void)hipEventRecord(r.a, stream);
      f();
      (void)hipEventRecord(r.b, stream);
      timing.recs.push_back(r);
    } else {
      f();
    }
  }
};

namespace {

inline int nblocks(int n) { return (n + kBlock - 1) / kBlock; }

ma_status check_faces(const ma_faces& F, int nc, bool internal, const char* what) {
  if (F.nfaces < 0) return ma::fail(MA_ERR_INVALID, std::string(what) + ": negative nfaces");
  if (F.nfaces == 0) return MA_OK;
  if (!F.coordinates || !F.face_normal || !F.face_tangent || !F.face_binormal || !F.face_cell_conn)
    return ma::fail(MA_ERR_INVALID, std::string(what) + ": null face array");
  for (int i = 0; i < F.nfaces; ++i) {
    const int L = F.face_cell_conn[2 * i], R = F.face_cell_conn[2 * i + 1];
    if (L < 0 || L >= nc) return ma::fail(MA_ERR_INVALID, std::string(what) + ": left cell out of range");
    if (internal && (R < 0 || R >= nc || R == L))
      return ma::fail(MA_ERR_INVALID, std::string(what) + ": right cell out of range");
  }
  if (F.permute_vector) {
    std::vector<char> seen(F.nfaces, 0);
    for (int i = 0; i < F.nfaces; ++i) {
      const int p = F.permute_vector[i];
      if (p < 0 || p >= F.nfaces || seen[p])
        return ma::fail(MA_ERR_INVALID, std::string(what) + ": permute_vector is not a permutation");
      seen[p] = 1;
    }
  }
  return MA_OK;
}

// Processing order of a face list: the given permutation, or a stable sort by
// left cell (what Faces.h:128-144 builds).
std::vector<int> processing_order(const ma_faces& F, int nc) {
  std::vector<int> perm(F.nfaces);
  if (F.permute_vector) {
    std::copy(F.permute_vector, F.permute_vector + F.nfaces, perm.begin());
    return perm;
  }
  std::vector<int> start((size_t)nc + 1, 0);
  for (int f = 0; f < F.nfaces; ++f) start[F.face_cell_conn[2 * f] + 1]++;
  for (int c = 0; c < nc; ++c) start[c + 1] += start[c];
  for (int f = 0; f < F.nfaces; ++f) perm[start[F.face_cell_conn[2 * f]]++] = f;
  return perm;
}

template <class T>
ma_status upload(ma_solver* s, T** dst, const std::vector<T>& h) {
  ma_status st = s->alloc(dst, h.size());
  if (st || s->counting) return st;
  MA_HIP(hipMemcpyAsync(*dst, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  return MA_OK;
}

// Lattice coordinates of a structured hex mesh from its face connectivity:
// every internal face links its two cells along the axis of the largest
// component of its normal; on a structured mesh those links form one chain
// per grid line, and a cell's position along its chain is its index on that
// axis.  Returns false when the mesh is not a full i x j x k lattice.
bool lattice_coords(const ma_mesh_data* md, std::vector<int> ijk[3], int dims[3]) {
  const int nc = md->num_owned_cells;
  const ma_faces& I = md->internal_faces;
  std::vector<int> succ[3], haspred[3];
  for (int a = 0; a < 3; ++a) {
    succ[a].assign(nc, -1);
    haspred[a].assign(nc, 0);
  }
  for (int i = 0; i < I.nfaces; ++i) {
    const double* n = &I.face_normal[3 * i];
    int a = 0;
    if (std::fabs(n[1]) > std::fabs(n[a])) a = 1;
    if (std::fabs(n[2]) > std::fabs(n[a])) a = 2;
    int from = I.face_cell_conn[2 * i], to = I.face_cell_conn[2 * i + 1];
    if (n[a] < 0) std::swap(from, to);
    if (succ[a][from] != -1 || haspred[a][to]) return false;
    succ[a][from] = to;
    haspred[a][to] = 1;
  }
  for (int a = 0; a < 3; ++a) {
    ijk[a].assign(nc, -1);
    dims[a] = 0;
    for (int c = 0; c < nc; ++c) {
      if (haspred[a][c]) continue;
      int k = 0;
      for (int x = c; x != -1; x = succ[a][x]) {
        if (ijk[a][x] != -1) return false;
        ijk[a][x] = k++;
      }
      dims[a] = std::max(dims[a], k);
    }
    for (int c = 0; c < nc; ++c)
      if (ijk[a][c] < 0) return false;  // a cycle
  }
  if ((long long)dims[0] * dims[1] * dims[2] != nc) return false;
  std::vector<char> seen(nc, 0);
  for (int c = 0; c < nc; ++c) {
    const long long id = ((long long)ijk[0][c] * dims[1] + ijk[1][c]) * dims[2] + ijk[2][c];
    if (id >= nc || seen[id]) return false;
    seen[id] = 1;
  }
  return true;
}

// Device cell order: bricks of the tile shape's lattice cells (cells inside a
// brick lexicographic, last axis fastest), so a tile is a compact block whose
// neighbours are mostly inside it.  Without a
// lattice, or with the all-zero shape, the original order is kept.  Fills
// dev_of (original -> device) and the cell range of every tile.
void cell_order(const ma_mesh_data* md, const TileShape& shape, std::vector<int>& dev_of,
                std::vector<int>& tile_cptr, int dims[3]) {
  const int nc = md->num_owned_cells;
  const int* kBrick = shape.b;
  std::vector<int> ijk[3];
  dev_of.resize(nc);
  tile_cptr.assign(1, 0);
  const bool bricks = kBrick[0] > 0;
  const bool lattice = lattice_coords(md, ijk, dims);
  if (!lattice) dims[0] = dims[1] = dims[2] = 0;
  if (!lattice || !bricks) {
    for (int c = 0; c < nc; ++c) dev_of[c] = c;
    for (int b = kTile; b < nc; b += kTile) tile_cptr.push_back(b);
    tile_cptr.push_back(nc);
    return;
  }
  std::vector<int> at((size_t)nc);
  for (int c = 0; c < nc; ++c) at[((size_t)ijk[0][c] * dims[1] + ijk[1][c]) * dims[2] + ijk[2][c]] = c;
  // Brick order: the brick grid is cut into kXcds slabs along its longest
  // axis (the kernels hand each XCD a contiguous run of tiles), and inside a
  // slab the axis of smallest extent runs fastest, so a brick's neighbours on
  // all three axes are close to it in dispatch order and still in that XCD's
  // L2 when it runs.
  int nb[3];
  for (int a = 0; a < 3; ++a) nb[a] = (dims[a] + kBrick[a] - 1) / kBrick[a];
  int sa = 0;
  for (int a = 1; a < 3; ++a)
    if (nb[a] > nb[sa]) sa = a;
  int ext[3] = {nb[0], nb[1], nb[2]};
  ext[sa] = (nb[sa] + kXcds - 1) / kXcds;
  int ax[3] = {0, 1, 2};  // slowest .. fastest inside a slab
  std::stable_sort(ax, ax + 3, [&](int a, int b) { return ext[a] > ext[b]; });
  int d = 0;
  for (int sl = 0; sl < kXcds; ++sl) {
    const int lo = (int)((long long)sl * nb[sa] / kXcds), hi = (int)((long long)(sl + 1) * nb[sa] / kXcds);
    if (lo >= hi) continue;
    int len[3], base[3];
    for (int a = 0; a < 3; ++a) {
      len[a] = a == sa ? hi - lo : nb[a];
      base[a] = a == sa ? lo : 0;
    }
    int q[3];
    for (q[0] = 0; q[0] < len[ax[0]]; ++q[0])
      for (q[1] = 0; q[1] < len[ax[1]]; ++q[1])
        for (q[2] = 0; q[2] < len[ax[2]]; ++q[2]) {
          int bb[3];
          for (int r = 0; r < 3; ++r) bb[ax[r]] = base[ax[r]] + q[r];
          const int bi = bb[0] * kBrick[0], bj = bb[1] * kBrick[1], bk = bb[2] * kBrick[2];
          for (int i = bi; i < std::min(bi + kBrick[0], dims[0]); ++i)
            for (int j = bj; j < std::min(bj + kBrick[1], dims[1]); ++j)
              for (int k = bk; k < std::min(bk + kBrick[2], dims[2]); ++k)
                dev_of[at[((size_t)i * dims[1] + j) * dims[2] + k]] = d++;
          tile_cptr.push_back(d);
        }
  }
}

// Everything the device needs, computed on the host (no GPU involved).
struct Layout {
  // frees every array of `other` (used to release merged parts)
  void swap_out(Layout& other) { std::swap(*this, other); }
  int nc = 0, nif = 0, nf = 0, nbf = 0, ntiles = 0, max_faces = 1;
  int dims[3] = {0, 0, 0};  // mesh lattice, zeros when there is none
  bool derive_frame = true;  // every face's t and b follow from its area vector
  std::vector<int> dev_of, org_of, tile_cptr, tptr, tfaces, slot, tslot, fl, fr;
  std::vector<double> fn, ffr, fx, fwall, vol, xc;
  std::vector<signed char> fkind;
  // geometry records of the faces (see dedupe_geometry): the record of each
  // face, and kGeoVals FP64 values per record
  std::vector<int> fgeo;
  std::vector<double> geo;
  // every area vector has at most one nonzero component and every cell
  // volume has a finite reciprocal in FP32 and FP64 (DevMesh::axis)
  bool axis = false;
};

// Every index a kernel will use, checked on the host before any upload.
ma_status validate_layout(const Layout& L) {
  const int nc = L.nc, nif = L.nif, nf = L.nf, nbf = L.nbf, ntiles = L.ntiles;
  const std::vector<int>& dev_of = L.dev_of;
  const std::vector<int>& tile_cptr = L.tile_cptr;
  const std::vector<int>& tptr = L.tptr;
  const std::vector<int>& tfaces = L.tfaces;
  const std::vector<int>& fl = L.fl;
  const std::vector<int>& fr = L.fr;
  const std::vector<int>& slot = L.slot;
  const std::vector<int>& tslot = L.tslot;
  if ((int)tile_cptr.size() != ntiles + 1 || (int)tptr.size() != ntiles + 1 || (int)dev_of.size() != nc)
    return ma::fail(MA_ERR_INVALID, "layout: sizes");
  if ((size_t)L.max_faces * 5 * sizeof(double) > 65536)
    return ma::fail(MA_ERR_UNSUPPORTED, "too many faces per cell tile for the flux kernel's LDS");
  std::vector<int> tile_of(nc, -1);
  for (int b = 0; b < ntiles; ++b)
    for (int d = tile_cptr[b]; d < tile_cptr[b + 1] && d < nc; ++d) tile_of[d] = b;
  for (int b = 0; b < ntiles; ++b)
    if (tptr[b + 1] - tptr[b] > L.max_faces) return ma::fail(MA_ERR_INVALID, "layout: tile faces");
  {
  std::vector<char> hit(nc, 0);
  for (int c = 0; c < nc; ++c) {
    if (dev_of[c] < 0 || dev_of[c] >= nc || hit[dev_of[c]]) return ma::fail(MA_ERR_INVALID, "layout: cell order");
    hit[dev_of[c]] = 1;
  }
  if (tile_cptr.front() != 0 || tile_cptr.back() != nc) return ma::fail(MA_ERR_INVALID, "layout: tile cells");
  {
    const size_t ng = L.geo.size() / kGeoVals;
    if ((int)L.fgeo.size() != nf || L.geo.size() % kGeoVals || (nf > 0 && ng == 0))
      return ma::fail(MA_ERR_INVALID, "layout: geometry records");
    for (int x : L.fgeo)
      if (x < 0 || (size_t)x >= ng) return ma::fail(MA_ERR_INVALID, "layout: geometry record index");
  }
  for (int b = 0; b < ntiles; ++b) {
    const int n = tile_cptr[b + 1] - tile_cptr[b];
    if (n < 1 || n > kTile) return ma::fail(MA_ERR_INVALID, "layout: tile size");
    for (int q = tptr[b]; q < tptr[b + 1]; ++q)
      if (tfaces[q] < 0 || tfaces[q] >= nif) return ma::fail(MA_ERR_INVALID, "layout: tile face");
  }
  for (int f = 0; f < nf; ++f)
    if (fl[f] < 0 || fl[f] >= nc || (f < nif && (fr[f] < 0 || fr[f] >= nc)))
      return ma::fail(MA_ERR_INVALID, "layout: face cells");
  for (int c = 0; c < nc; ++c) {
    const int b = tile_of[c];
    if (b < 0) return ma::fail(MA_ERR_INVALID, "layout: cell outside every tile");
    const int nloc = tptr[b + 1] - tptr[b];
    for (int q = 0; q < kMaxSlots; ++q) {
      const int e = tslot[(size_t)q * nc + c], g = slot[(size_t)q * nc + c];
      if ((e < 0) != (g < 0)) return ma::fail(MA_ERR_INVALID, "layout: slot tables");
      if (e < 0) continue;
      if ((g >> 1) >= nf) return ma::fail(MA_ERR_INVALID, "layout: slot face");
      if ((e & 7) < 2 ? (e >> 3) >= nloc || tfaces[tptr[b] + (e >> 3)] != (g >> 1)
                      : (e >> 3) >= nbf)
        return ma::fail(MA_ERR_INVALID, "layout: tile slot");
    }
  }
  }
  return MA_OK;
}

// Concatenation of independent cases' layouts into one: cells, internal faces,
// boundary faces and tiles of case k follow those of case k-1; tiles never
// span two cases.  Internal faces of all cases come before all boundary faces.
void merge_layouts(std::vector<Layout>& parts, Layout& M, std::vector<int>& tile_case,
                   std::vector<int>& case_tiles, std::vector<int>& bc_case, std::vector<int>& cell_case) {
  const int n = (int)parts.size();
  std::vector<int> co(n + 1, 0), io(n + 1, 0), bo(n + 1, 0), to(n + 1, 0), qo(n + 1, 0);
  M.derive_frame = true;
  M.max_faces = 1;
  for (int k = 0; k < n; ++k) {
    const Layout& L = parts[k];
    co[k + 1] = co[k] + L.nc;
    io[k + 1] = io[k] + L.nif;
    bo[k + 1] = bo[k] + L.nbf;
    to[k + 1] = to[k] + L.ntiles;
    qo[k + 1] = qo[k] + L.tptr[L.ntiles];
    M.derive_frame = M.derive_frame && L.derive_frame;
    M.max_faces = std::max(M.max_faces, L.max_faces);
  }
  M.nc = co[n];
  M.nif = io[n];
  M.nbf = bo[n];
  M.nf = M.nif + M.nbf;
  M.ntiles = to[n];
  const int nc = M.nc, nf = M.nf;
  M.dev_of.assign(nc, 0);
  M.org_of.assign(nc, 0);
  M.tile_cptr.assign(M.ntiles + 1, 0);
  M.tptr.assign(M.ntiles + 1, 0);
  M.tfaces.assign(std::max(1, qo[n]), 0);
  M.slot.assign((size_t)kMaxSlots * nc, -1);
  M.tslot.assign((size_t)kMaxSlots * nc, -1);
  M.fl.assign(nf, 0);
  M.fr.assign(nf, -1);
  M.fkind.assign(nf, 0);
  M.fwall.assign(nf, 0.0);
  M.fn.assign((size_t)3 * nf, 0.0);
  if (!M.derive_frame) M.ffr.assign((size_t)6 * nf, 0.0);
  M.fx.assign((size_t)3 * nf, 0.0);
  M.vol.assign(nc, 0.0);
  M.xc.assign((size_t)3 * nc, 0.0);
  M.fgeo.assign(M.nf, 0);
  M.geo.clear();
  M.axis = true;
  for (int k = 0; k < n; ++k) M.axis = M.axis && parts[k].axis;
  tile_case.assign(M.ntiles, 0);
  case_tiles.assign(n + 1, 0);
  bc_case.assign(std::max(1, M.nbf), 0);
  cell_case.assign(nc, 0);
  for (int k = 0; k < n; ++k) {
    Layout& L = parts[k];
    const int lnc = L.nc, lnif = L.nif, lnf = L.nf;
    auto gf = [&](int f) { return f < lnif ? io[k] + f : M.nif + bo[k] + (f - lnif); };
    for (int c = 0; c < lnc; ++c) {
      M.dev_of[co[k] + c] = co[k] + L.dev_of[c];
      M.org_of[co[k] + c] = co[k] + L.org_of[c];
      M.vol[co[k] + c] = L.vol[c];
      cell_case[co[k] + c] = k;
      for (int d = 0; d < 3; ++d) M.xc[(size_t)d * nc + co[k] + c] = L.xc[(size_t)d * lnc + c];
      for (int q = 0; q < kMaxSlots; ++q) {
        const int e = L.slot[(size_t)q * lnc + c];
        if (e >= 0) M.slot[(size_t)q * nc + co[k] + c] = gf(e >> 1) * 2 + (e & 1);
        const int te = L.tslot[(size_t)q * lnc + c];
        if (te >= 0)
          M.tslot[(size_t)q * nc + co[k] + c] = (te & 7) < 2 ? te : (((te >> 3) + bo[k]) << 3) | (te & 7);
      }
    }
    for (int b = 0; b < L.ntiles; ++b) {
      M.tile_cptr[to[k] + b + 1] = co[k] + L.tile_cptr[b + 1];
      M.tptr[to[k] + b + 1] = qo[k] + L.tptr[b + 1];
      tile_case[to[k] + b] = k;
    }
    case_tiles[k + 1] = to[k + 1];
    for (int q = 0; q < L.tptr[L.ntiles]; ++q) M.tfaces[qo[k] + q] = io[k] + L.tfaces[q];
    for (int f = 0; f < lnf; ++f) {
      const int g = gf(f);
      M.fl[g] = co[k] + L.fl[f];
      M.fr[g] = f < lnif ? co[k] + L.fr[f] : -1;
      M.fkind[g] = L.fkind[f];
      M.fwall[g] = L.fwall[f];
      for (int d = 0; d < 3; ++d) {
        M.fn[(size_t)d * nf + g] = L.fn[(size_t)d * lnf + f];
        M.fx[(size_t)d * nf + g] = L.fx[(size_t)d * lnf + f];
      }
      if (!M.derive_frame) {
        if (L.derive_frame) {  // this part stored no frames: they follow from the area vector
          const double a[3] = {L.fn[f], L.fn[(size_t)lnf + f], L.fn[(size_t)2 * lnf + f]};
          double t[3], b[3];
          ma::tangent_binormal(a, t, b);
          for (int d = 0; d < 3; ++d) {
            M.ffr[(size_t)d * nf + g] = t[d];
            M.ffr[(size_t)(3 + d) * nf + g] = b[d];
          }
        } else {
          for (int d = 0; d < 6; ++d) M.ffr[(size_t)d * nf + g] = L.ffr[(size_t)d * lnf + f];
        }
      }
      if (f >= lnif) bc_case[bo[k] + (f - lnif)] = k;
    }
    {  // this part's geometry records follow the earlier parts'
      const int r0 = (int)(M.geo.size() / kGeoVals);
      for (int f = 0; f < lnf; ++f) M.fgeo[gf(f)] = r0 + L.fgeo[f];
      M.geo.insert(M.geo.end(), L.geo.begin(), L.geo.end());
    }
    Layout().swap_out(L);  // release the part as soon as it is merged
  }
}

// Geometry records of the faces, which the flux and gradient kernels read:
// each face's area vector and both displacements x_f - x_left, x_f - x_right
// (FP64, the expressions the displacement planes use; the right one zero for
// a boundary face), with faces whose nine values have the same bits sharing
// one record.  Box meshes repeat few
// records over many faces (the 256 x 128 x 64 stretched plate: 7,534 records
// for 6.2M internal faces; a uniform box: 3), so the kernel reads a 4-byte
// record index per face (the gradient kernel per slot) instead of up to 9
// values, and the records stay in L2.
// When records repeat fewer than kGeoReuse times on average, every face gets
// its own record (record f for face f).
constexpr int kGeoReuse = 8;

struct GeoKeyHash {
  size_t operator()(const std::array<unsigned long long, kGeoVals>& k) const {
    unsigned long long h = 1469598103934665603ull;
    for (unsigned long long x : k) {
      h ^= x;
      h *= 1099511628211ull;
      h ^= h >> 29;
    }
    return (size_t)h;
  }
};

void dedupe_geometry(Layout& L) {
  const int nf = L.nf, nc = L.nc;
  L.axis = true;
  for (int f = 0; f < nf && L.axis; ++f)
    L.axis = int(L.fn[f] != 0.0) + int(L.fn[(size_t)nf + f] != 0.0) + int(L.fn[(size_t)2 * nf + f] != 0.0) <= 1;
  for (int c = 0; c < nc && L.axis; ++c)
    L.axis = std::fabs(L.vol[c]) >= (double)std::numeric_limits<float>::min() &&
             std::fabs(L.vol[c]) <= (double)std::numeric_limits<float>::max();
  auto record = [&](int f, double* r) {
    for (int d = 0; d < 3; ++d) {
      r[d] = L.fn[(size_t)d * nf + f];
      r[3 + d] = L.fx[(size_t)d * nf + f] - L.xc[(size_t)d * nc + L.fl[f]];
      r[6 + d] = L.fr[f] >= 0 ? L.fx[(size_t)d * nf + f] - L.xc[(size_t)d * nc + L.fr[f]] : 0.0;
    }
  };
  L.fgeo.assign((size_t)nf, 0);
  L.geo.clear();
  const size_t cap = std::min<size_t>((size_t)nf / kGeoReuse, size_t(1) << 20);
  std::unordered_map<std::array<unsigned long long, kGeoVals>, int, GeoKeyHash> ids;
  ids.reserve(std::min<size_t>(cap, 1 << 16) * 2);
  bool shared = nf >= kGeoReuse;
  for (int f = 0; shared && f < nf; ++f) {
    double r[kGeoVals];
    record(f, r);
    std::array<unsigned long long, kGeoVals> key;
    std::memcpy(key.data(), r, sizeof r);
    const auto it = ids.emplace(key, (int)ids.size());
    if (it.second) {
      if (ids.size() > cap) shared = false;  // too few repeats
      L.geo.insert(L.geo.end(), r, r + kGeoVals);
    }
    L.fgeo[f] = it.first->second;
  }
  if (shared) return;
  L.geo.assign((size_t)kGeoVals * nf, 0.0);
  for (int f = 0; f < nf; ++f) {
    record(f, &L.geo[(size_t)kGeoVals * f]);
    L.fgeo[f] = f;
  }
}

ma_status plan_layout(const ma_mesh_data* md, const TileShape& shape, Layout& L) {
  const int nc = md->num_owned_cells;
  L.nc = nc;
  const ma_faces& I = md->internal_faces;
  ma_status st = check_faces(I, nc, true, "internal_faces");
  if (st) return st;
  long long nbf = 0;
  for (int g = 0; g < md->num_boundary_groups; ++g) {
    const ma_bc_group& G = md->boundary_faces[g];
    if (G.kind < MA_BC_EXTRAPOLATE || G.kind > MA_BC_INFLOW)
      return ma::fail(MA_ERR_INVALID, "boundary_faces: unknown kind");
    st = check_faces(G.faces, nc, false, "boundary_faces");
    if (st) return st;
    nbf += G.faces.nfaces;
  }
  if ((long long)I.nfaces + nbf > (1ll << 28))
    return ma::fail(MA_ERR_UNSUPPORTED, "more than 2^28 faces");
  const int nif = I.nfaces, nf = nif + (int)nbf;
  L.nif = nif;
  L.nf = nf;
  L.nbf = (int)nbf;

  // Cell order on the device, and its inverse.
  std::vector<int>& dev_of = L.dev_of;
  std::vector<int>& tile_cptr = L.tile_cptr;
  cell_order(md, shape, dev_of, tile_cptr, L.dims);
  std::vector<int>& org_of = L.org_of;
  org_of.assign(nc, 0);
  for (int c = 0; c < nc; ++c) org_of[dev_of[c]] = c;
  const int ntiles = (int)tile_cptr.size() - 1;

  // Summation order: internal faces in processing order (rank[i] = position
  // of face i in it).  Device face order: internal faces sorted by the device
  // index of their left cell (stable in processing order), then the boundary
  // groups in order.  Memory order does not change any sum.
  const std::vector<int> perm = processing_order(I, nc);
  std::vector<int> dface(nif);  // device slot -> original internal face
  {
    std::vector<int> start((size_t)nc + 1, 0);
    for (int i = 0; i < nif; ++i) start[dev_of[I.face_cell_conn[2 * i]] + 1]++;
    for (int c = 0; c < nc; ++c) start[c + 1] += start[c];
    for (int ii = 0; ii < nif; ++ii) {
      const int i = perm[ii];
      dface[start[dev_of[I.face_cell_conn[2 * i]]]++] = i;
    }
  }
  std::vector<int> devf_of(nif);  // original internal face -> device slot
  for (int g = 0; g < nif; ++g) devf_of[dface[g]] = g;

  std::vector<double>& fn = L.fn;
  std::vector<double>& ffr = L.ffr;
  std::vector<double>& fx = L.fx;
  std::vector<double>& fwall = L.fwall;
  std::vector<int>& fl = L.fl;
  std::vector<int>& fr = L.fr;
  std::vector<signed char>& fkind = L.fkind;
  fn.assign((size_t)3 * nf, 0.0);
  fx.assign((size_t)3 * nf, 0.0);
  fwall.assign(nf, 0.0);
  fl.assign(nf, 0);
  fr.assign(nf, 0);
  fkind.assign(nf, 0);
  // Area vector and centroid of every face; whether every stored tangent
  // and binormal follow from the area vector as Face.C builds them (then the
  // kernels rebuild the frame and no frame array is stored).
  auto derivable = [](const ma_faces& F) {  // in the caller's order (sequential reads)
    for (int i = 0; i < F.nfaces; ++i) {
      double t[3], b[3];
      ma::tangent_binormal(&F.face_normal[3 * i], t, b);
      for (int d = 0; d < 3; ++d)
        if (t[d] != F.face_tangent[3 * i + d] || b[d] != F.face_binormal[3 * i + d] ||
            std::signbit(t[d]) != std::signbit(F.face_tangent[3 * i + d]) ||
            std::signbit(b[d]) != std::signbit(F.face_binormal[3 * i + d]))
          return false;
    }
    return true;
  };
  L.derive_frame = derivable(I);
  for (int g = 0; g < md->num_boundary_groups && L.derive_frame; ++g)
    L.derive_frame = derivable(md->boundary_faces[g].faces);
  auto put = [&](int f, const ma_faces& F, int i) {
    for (int d = 0; d < 3; ++d) {
      fn[(size_t)d * nf + f] = F.face_normal[3 * i + d];
      fx[(size_t)d * nf + f] = F.coordinates[3 * i + d];
    }
  };
  auto put_frame = [&](int f, const ma_faces& F, int i) {  // the caller's tangent and binormal
    for (int d = 0; d < 3; ++d) {
      ffr[(size_t)d * nf + f] = F.face_tangent[3 * i + d];
      ffr[(size_t)(3 + d) * nf + f] = F.face_binormal[3 * i + d];
    }
  };
  for (int g = 0; g < nif; ++g) {
    const int i = dface[g];
    put(g, I, i);
    fl[g] = dev_of[I.face_cell_conn[2 * i]];
    fr[g] = dev_of[I.face_cell_conn[2 * i + 1]];
  }
  {
    int f = nif;
    for (int g = 0; g < md->num_boundary_groups; ++g) {
      const ma_bc_group& G = md->boundary_faces[g];
      for (int i = 0; i < G.faces.nfaces; ++i, ++f) {
        put(f, G.faces, i);
        const int c = G.faces.face_cell_conn[2 * i];
        fl[f] = dev_of[c];
        fr[f] = -1;
        fkind[f] = (signed char)G.kind;
        // NoSlip_BC.h:124-129: 1 / sqrt(sum_d pow(x_f - x_c, 2))
        double dist = 0;
        for (int d = 0; d < 3; ++d)
          dist += std::pow(G.faces.coordinates[3 * i + d] - md->cell_coordinates[3 * c + d], 2);
        fwall[f] = 1.0 / std::sqrt(dist);
      }
    }
  }
  if (!L.derive_frame) {  // stored tangents and binormals, the caller's vectors
    ffr.assign((size_t)6 * nf, 0.0);
    for (int g = 0; g < nif; ++g) put_frame(g, I, dface[g]);
    int f = nif;
    for (int g = 0; g < md->num_boundary_groups; ++g) {
      const ma_bc_group& G = md->boundary_faces[g];
      for (int i = 0; i < G.faces.nfaces; ++i, ++f) put_frame(f, G.faces, i);
    }
  }

  // Per-cell face lists in summation order: internal faces by processing
  // rank, then boundary faces in group order.  Entry = device face * 2 + side.
  std::vector<int>& slot = L.slot;
  slot.assign((size_t)kMaxSlots * nc, -1);
  std::vector<unsigned char> used(nc, 0);
  auto add_slot = [&](int dc, int entry) -> bool {
    if (used[dc] >= kMaxSlots) return false;
    slot[(size_t)used[dc] * nc + dc] = entry;
    used[dc]++;
    return true;
  };
  for (int ii = 0; ii < nif; ++ii) {
    const int i = perm[ii], g = devf_of[i];
    if (!add_slot(fl[g], g * 2) || !add_slot(fr[g], g * 2 + 1))
      return ma::fail(MA_ERR_UNSUPPORTED, "a cell has more than 6 faces (hex8 meshes only)");
  }
  for (int f = nif; f < nf; ++f)
    if (!add_slot(fl[f], f * 2))
      return ma::fail(MA_ERR_UNSUPPORTED, "a cell has more than 6 faces (hex8 meshes only)");

  // Tiles of the flux kernel.  Every internal face is listed in the tile of
  // its left cell and, when different, in the tile of its right cell; faces
  // are appended in ascending device order, so each list is sorted.
  std::vector<int> tile_of(nc);
  for (int b = 0; b < ntiles; ++b)
    for (int d = tile_cptr[b]; d < tile_cptr[b + 1]; ++d) tile_of[d] = b;
  std::vector<int>& tptr = L.tptr;
  tptr.assign((size_t)ntiles + 1, 0);
  for (int g = 0; g < nif; ++g) {
    const int tl = tile_of[fl[g]], tr = tile_of[fr[g]];
    tptr[tl + 1]++;
    if (tr != tl) tptr[tr + 1]++;
  }
  for (int b = 0; b < ntiles; ++b) tptr[b + 1] += tptr[b];
  std::vector<int>& tfaces = L.tfaces;
  tfaces.assign((size_t)std::max(1, tptr[ntiles]), 0);
  std::vector<int> locl(nif), locr(nif);
  {
    std::vector<int> fill(tptr.begin(), tptr.end() - 1);
    for (int g = 0; g < nif; ++g) {
      const int tl = tile_of[fl[g]], tr = tile_of[fr[g]];
      locl[g] = fill[tl] - tptr[tl];
      tfaces[fill[tl]++] = g;
      if (tr != tl) {
        locr[g] = fill[tr] - tptr[tr];
        tfaces[fill[tr]++] = g;
      } else {
        locr[g] = locl[g];
      }
    }
  }
  int max_faces = 1;
  for (int b = 0; b < ntiles; ++b) max_faces = std::max(max_faces, tptr[b + 1] - tptr[b]);
  if ((size_t)max_faces * 5 * sizeof(double) > 65536)
    return ma::fail(MA_ERR_UNSUPPORTED, "too many faces per cell tile for the flux kernel's LDS");
  std::vector<int>& tslot = L.tslot;
  tslot.assign((size_t)kMaxSlots * nc, -1);
  for (int c = 0; c < nc; ++c)
    for (int q = 0; q < kMaxSlots; ++q) {
      const int e = slot[(size_t)q * nc + c];
      if (e < 0) break;
      const int g = e >> 1, side = e & 1;
      tslot[(size_t)q * nc + c] =
          g < nif ? (((side ? locr[g] : locl[g]) << 3) | side) : (((g - nif) << 3) | (3 + fkind[g]));
    }

  L.vol.assign(nc, 0.0);
  L.xc.assign((size_t)3 * nc, 0.0);
  for (int c = 0; c < nc; ++c) {
    const int d = dev_of[c];
    L.vol[d] = md->cell_volumes[c];
    for (int k = 0; k < 3; ++k) L.xc[(size_t)k * nc + d] = md->cell_coordinates[3 * c + k];
  }
  L.ntiles = ntiles;
  L.max_faces = max_faces;
  dedupe_geometry(L);
  return MA_OK;
}

double effective_prandtl(const ma_options& o) { return o.prandtl > 0.0 ? o.prandtl : 0.71; }

// Inflow state of a case: the explicit one, or TimeSolverExplicitRK4.h:218-223.
Inflow case_inflow(const ma_options& o) {
  const double ref_inflow[5] = {0.5805, 503.96, 0.0, 0.0, 343750.0};
  Inflow in{};
  for (int k = 0; k < 5; ++k) in.U[k] = o.use_inflow ? o.inflow[k] : ref_inflow[k];
  return in;
}

ma_status check_options(const ma_options* opt) {
  if (opt->cfl <= 0.0 && !(opt->dt > 0.0)) return ma::fail(MA_ERR_INVALID, "dt must be positive");
  if (opt->limiter != MA_LIMITER_VENKAT && opt->limiter != MA_LIMITER_VANALBADA)
    return ma::fail(MA_ERR_INVALID, "unknown limiter");
  if (opt->precision != MA_PRECISION_FP64 && opt->precision != MA_PRECISION_FP32)
    return ma::fail(MA_ERR_INVALID, "unknown precision");
  return MA_OK;
}

ma_status check_mesh(const ma_mesh_data* md, const char* who) {
  if (!md) return ma::fail(MA_ERR_INVALID, std::string(who) + ": null mesh");
  if (md->num_ghosts != 0)
    return ma::fail(MA_ERR_UNSUPPORTED, "ghost cells (multi-rank meshes) are not supported on one GPU");
  if (md->num_owned_cells <= 0 || !md->cell_coordinates || !md->cell_volumes)
    return ma::fail(MA_ERR_INVALID, std::string(who) + ": empty or null cell arrays");
  if (md->num_boundary_groups < 0 || (md->num_boundary_groups > 0 && !md->boundary_faces))
    return ma::fail(MA_ERR_INVALID, std::string(who) + ": bad boundary group list");
  return MA_OK;
}

// Uploads the 6 displacement planes T(x_f - x_left), T(x_f - x_right) of
// every face (right planes zero for boundary faces).
template <class T>
ma_status upload_disp(ma_solver* s, const Layout& L, void** out) {
  const int nf = L.nf, nc = L.nc;
  std::vector<T> h;
  if (!s->counting) {
    h.assign((size_t)6 * nf, T(0));
    for (int f = 0; f < nf; ++f)
      for (int side = 0; side < 2; ++side) {
        const int c = side ? L.fr[f] : L.fl[f];
        if (c < 0) continue;
        for (int d = 0; d < 3; ++d)
          h[(size_t)(side * 3 + d) * nf + f] = T(L.fx[(size_t)d * nf + f] - L.xc[(size_t)d * nc + c]);
      }
  } else {
    h.resize(1);
  }
  T* p = nullptr;
  ma_status st = s->counting ? s->alloc(&p, (size_t)6 * nf) : upload(s, &p, h);
  *out = p;
  return st;
}

// Uploads the shared geometry records in the arithmetic type, kGeoStride
// values each (T of the FP64 values, as the fn32 and displacement planes
// round them).
template <class T>
ma_status upload_geo(ma_solver* s, const Layout& L, void** out) {
  const size_t ng = L.geo.size() / kGeoVals;
  std::vector<T> h;
  if (!s->counting) {
    h.assign(ng * kGeoStride<T>, T(0));
    for (size_t r = 0; r < ng; ++r)
      for (int v = 0; v < kGeoVals; ++v) h[r * kGeoStride<T> + v] = T(L.geo[r * kGeoVals + v]);
  } else {
    h.resize(1);
  }
  T* p = nullptr;
  ma_status st = s->counting ? s->alloc(&p, ng * kGeoStride<T>) : upload(s, &p, h);
  *out = p;
  return st;
}

// Layout of a run over ncases meshes (concatenated when there are several),
// checked, with the case of every tile, boundary face and device cell.
struct RunPlan {
  Layout L;
  bool all_lattice = true;  // every case's mesh is an i x j x k lattice
  std::vector<int> tile_case, case_tiles, bc_case, cell_case, case_cell_off;
};

ma_status plan_run(int ncases, const ma_mesh_data* const* mds, const TileShape& shape, RunPlan& P) {
  if (!valid_shape(shape)) return ma::fail(MA_ERR_INVALID, "tile shape: extents >= 1 with product <= 256, or all 0");
  P.case_cell_off.assign(ncases + 1, 0);
  long long total_cells = 0, total_faces = 0;
  for (int k = 0; k < ncases; ++k) {
    ma_status st = check_mesh(mds[k], "plan");
    if (st) return st;
    total_cells += mds[k]->num_owned_cells;
    total_faces += mds[k]->internal_faces.nfaces;
    for (int g = 0; g < mds[k]->num_boundary_groups; ++g) total_faces += mds[k]->boundary_faces[g].faces.nfaces;
    if (total_cells > (1ll << 28) || total_faces > (1ll << 28))
      return ma::fail(MA_ERR_UNSUPPORTED, "more than 2^28 cells or faces in one run");
    P.case_cell_off[k + 1] = (int)total_cells;
  }
  ma_status st = MA_OK;
  Layout& L = P.L;
  if (ncases == 1) {
    if ((st = plan_layout(mds[0], shape, L))) return st;
    P.all_lattice = L.dims[0] > 0;
    P.tile_case.assign(L.ntiles, 0);
    P.case_tiles = {0, L.ntiles};
    P.bc_case.assign(std::max(1, L.nbf), 0);
    P.cell_case.assign(L.nc, 0);
  } else {
    // cases are planned independently on up to 16 host threads
    std::vector<Layout> parts(ncases);
    std::vector<ma_status> status(ncases, MA_OK);
    std::vector<std::string> message(ncases);
    std::atomic<int> next{0};
    auto worker = [&] {
      for (int k = next++; k < ncases; k = next++) {
        try {
          status[k] = plan_layout(mds[k], shape, parts[k]);
          if (status[k]) message[k] = ma::last_error();
        } catch (const std::bad_alloc&) {
          status[k] = MA_ERR_NOMEM;
          message[k] = "plan: out of host memory";
        }
      }
    };
    const int nthreads = std::max(1, std::min({ncases, 16, (int)std::thread::hardware_concurrency()}));
    std::vector<std::thread> pool;
    for (int i = 1; i < nthreads; ++i) pool.emplace_back(worker);
    worker();
    for (auto& th : pool) th.join();
    for (int k = 0; k < ncases; ++k) {
      if (status[k]) return ma::fail(status[k], "case " + std::to_string(k) + ": " + message[k]);
      P.all_lattice = P.all_lattice && parts[k].dims[0] > 0;
    }
    merge_layouts(parts, L, P.tile_case, P.case_tiles, P.bc_case, P.cell_case);
  }
  return validate_layout(L);
}

void plan_stats(const Layout& L, long long* stats) {
  stats[0] = L.dims[0];
  stats[1] = L.dims[1];
  stats[2] = L.dims[2];
  stats[3] = L.ntiles;
  stats[4] = L.max_faces;
  stats[5] = L.tptr.empty() ? 0 : L.tptr.back();
  stats[6] = L.derive_frame ? 1 : 0;
}

// Device arrays of one run from its plan: the layout, the per-case
// parameters and the work buffers.  With s->counting set nothing is
// allocated or copied; s->arena_used is then the device bytes the run needs.
// The plan is only read, so one plan can build many runs (the served form
// keeps plans of the mesh shapes it has seen).
ma_status build_planned(ma_solver* s, const RunPlan& P) {
  ma_status st = MA_OK;
  const int ncases = s->ncases;
  s->all_lattice = P.all_lattice;
  const Layout& L = P.L;
  const std::vector<int>& tile_case = P.tile_case;
  const std::vector<int>& case_tiles = P.case_tiles;
  const std::vector<int>& bc_case = P.bc_case;
  const std::vector<int>& cell_case = P.cell_case;
  const int nc = L.nc, nif = L.nif, nf = L.nf;
  const long long nbf = L.nbf;
  if (!s->counting) {
    s->case_cell_off = P.case_cell_off;
    s->dev_of = L.dev_of;
  }
  DevMesh& m = s->m;
  m.nc = nc;
  m.nif = nif;
  m.nf = nf;
  double *dvol, *dxc, *dfn, *dffr = nullptr, *dfw;
  int *dslot, *dfl, *dfr, *dorg;
  signed char* dkind;
  if ((st = upload(s, &dvol, L.vol)) || (st = upload(s, &dxc, L.xc)) || (st = upload(s, &dslot, L.slot)) ||
      (st = upload(s, &dfn, L.fn)) || (!L.derive_frame && (st = upload(s, &dffr, L.ffr))) ||
      (st = upload(s, &dorg, L.org_of)) ||
      (st = upload(s, &dfl, L.fl)) || (st = upload(s, &dfr, L.fr)) ||
      (st = upload(s, &dkind, L.fkind)) || (st = upload(s, &dfw, L.fwall)))
    return st;
  m.vol = dvol;
  s->xc = dxc;
  m.slot = dslot;
  s->org = dorg;
  m.fn = dfn;
  m.fn32 = nullptr;
  if (s->fp32) {  // the area vector in FP32 for the FP32-arithmetic kernels
    float* dfn32 = nullptr;
    if (s->counting) {
      st = s->alloc(&dfn32, (size_t)3 * nf);
    } else {
      std::vector<float> h(L.fn.begin(), L.fn.end());
      st = upload(s, &dfn32, h);
    }
    if (st) return st;
    m.fn32 = dfn32;
  }
  m.ffr = L.derive_frame ? nullptr : dffr;
  m.derive_frame = L.derive_frame ? 1 : 0;
  // displacement planes T(x_f - x_cell) from the FP64 centroids
  void* dfdx = nullptr;
  if ((st = s->fp32 ? upload_disp<float>(s, L, &dfdx) : upload_disp<double>(s, L, &dfdx))) return st;
  m.fdx = dfdx;
  m.fl = dfl;
  m.fr = dfr;
  m.fkind = dkind;
  m.fwall = dfw;
  // neighbour planes: the cell across each slot's internal face, else the cell itself
  {
    int* dnbr = nullptr;
    if (s->counting) {
      st = s->alloc(&dnbr, (size_t)kMaxSlots * nc);
    } else {
      std::vector<int> h((size_t)kMaxSlots * nc);
      for (int q = 0; q < kMaxSlots; ++q)
        for (int c = 0; c < nc; ++c) {
          const int e = L.slot[(size_t)q * nc + c], f = e >> 1;
          h[(size_t)q * nc + c] = (e >= 0 && f < nif) ? ((e & 1) ? L.fl[f] : L.fr[f]) : c;
        }
      st = upload(s, &dnbr, h);
    }
    if (st) return st;
    s->nbr = dnbr;
  }

  int *dtptr, *dtfaces, *dtslot, *dtcptr, *dtcase;
  if ((st = upload(s, &dtptr, L.tptr)) || (st = upload(s, &dtfaces, L.tfaces)) || (st = upload(s, &dtslot, L.tslot)) ||
      (st = upload(s, &dtcptr, L.tile_cptr)) || (st = upload(s, &dtcase, tile_case)))
    return st;
  s->t.ntiles = L.ntiles;
  s->t.per_xcd = (L.ntiles + kXcds - 1) / kXcds;
  s->t.max_faces = L.max_faces;
  s->t.ptr = dtptr;
  s->t.faces = dtfaces;
  {
    int2* dlr = nullptr;
    const size_t ntf = L.tfaces.size();
    if (s->counting) {
      st = s->alloc(&dlr, ntf);
    } else {
      // slot of each internal face in its left and right cell's slot list,
      // packed above the cell index for the face-value scheme (inviscid
      // second order) only
      const bool face_vals = s->opt.second_order_space && !s->opt.viscous;
      std::vector<int> fs((size_t)2 * nif, 0);
      for (int q = 0; face_vals && q < kMaxSlots; ++q)
        for (int c = 0; c < nc; ++c) {
          const int e = L.slot[(size_t)q * nc + c];
          if (e >= 0 && (e >> 1) < nif) fs[(size_t)(e & 1) * nif + (e >> 1)] = q;
        }
      // (tfaces is padded to one entry when there are none; padding stays 0)
      std::vector<int2> h(ntf, make_int2(0, 0));
      for (size_t i = 0; i < (size_t)L.tptr[L.ntiles]; ++i) {
        const int f = L.tfaces[i];
        h[i] = make_int2(L.fl[f] | (fs[f] << kCellBits), L.fr[f] | (fs[(size_t)nif + f] << kCellBits));
      }
      st = upload(s, &dlr, h);
    }
    if (st) return st;
    s->t.lr = dlr;
  }
  s->t.slot = dtslot;
  s->t.cptr = dtcptr;
  s->t.tcase = dtcase;
  // face geometry records, and each tile face's record index
  {
    void* dg = nullptr;
    if ((st = s->fp32 ? upload_geo<float>(s, L, &dg) : upload_geo<double>(s, L, &dg))) return st;
    int* dfg = nullptr;
    const size_t ntf = L.tfaces.size();
    if (s->counting) {
      st = s->alloc(&dfg, ntf);
    } else {
      std::vector<int> h(ntf, 0);
      for (size_t i = 0; i < (size_t)L.tptr[L.ntiles]; ++i) h[i] = L.fgeo[L.tfaces[i]];
      st = upload(s, &dfg, h);
    }
    if (st) return st;
    m.gtab = dg;
    s->t.fgid = dfg;
    m.axis = L.axis ? 1 : 0;
    m.full = std::all_of(L.slot.begin(), L.slot.end(), [](int e) { return e >= 0; }) ? 1 : 0;
    int* dsg = nullptr;
    if (s->counting) {
      st = s->alloc(&dsg, (size_t)kMaxSlots * nc);
    } else {
      std::vector<int> h((size_t)kMaxSlots * nc, -1);
      for (size_t i = 0; i < h.size(); ++i)
        if (L.slot[i] >= 0) h[i] = (L.fgeo[L.slot[i] >> 1] << 1) | (L.slot[i] & 1);
      st = upload(s, &dsg, h);
    }
    if (st) return st;
    m.sgeo = dsg;
  }

  // per-case parameters
  std::vector<double> dt0(ncases), cfl(ncases), zero(ncases, 0.0);
  std::vector<IcParams> ic(ncases);
  std::vector<Inflow> inflow(ncases);
  s->any_cfl = false;
  for (int k = 0; k < ncases; ++k) {
    const ma_options& o = s->case_opt[k];
    dt0[k] = o.dt;
    cfl[k] = o.cfl > 0.0 ? o.cfl : 0.0;
    s->any_cfl = s->any_cfl || o.cfl > 0.0;
    inflow[k] = case_inflow(o);
    IcParams& p = ic[k];
    p = IcParams{};
    p.kind = o.ic_kind;
    p.problem_type = o.problem_type;
    p.midx = o.lx / 2.0;
    for (int q = 0; q < 5; ++q) {
      p.ic[q] = o.ic[q];
      p.inflow[q] = inflow[k].U[q];
    }
    p.perturb = o.perturb;
    p.seed = o.seed;
    p.cell0 = (s->counting ? P.case_cell_off : s->case_cell_off)[k];
  }
  // CFL runs: each cell's face geometry for the time-step bound (dt_tile),
  // summed over its slot list in slot order
  s->m.dtg = nullptr;
  if (s->any_cfl) {
    std::vector<double> g;
    if (!s->counting) {
      g.assign((size_t)5 * nc, 0.0);
      for (int c = 0; c < nc; ++c) {
        double sx = 0.0, sy = 0.0, sz = 0.0, n1 = 0.0, n2 = 0.0;
        for (int q = 0; q < kMaxSlots; ++q) {
          const int e = L.slot[(size_t)q * nc + c];
          if (e < 0) break;
          const int f = e >> 1;
          const double a0 = L.fn[f], a1 = L.fn[(size_t)nf + f], a2 = L.fn[(size_t)2 * nf + f];
          const double q2 = a0 * a0 + a1 * a1 + a2 * a2;
          sx = sx + std::fabs(a0);
          sy = sy + std::fabs(a1);
          sz = sz + std::fabs(a2);
          n1 = n1 + std::sqrt(q2);
          n2 = n2 + q2;
        }
        g[c] = sx;
        g[(size_t)nc + c] = sy;
        g[(size_t)2 * nc + c] = sz;
        g[(size_t)3 * nc + c] = n1;
        g[(size_t)4 * nc + c] = n2;
      }
    } else {
      g.resize(1);
    }
    double* d = nullptr;
    if ((st = s->counting ? s->alloc(&d, (size_t)5 * nc) : upload(s, &d, g))) return st;
    s->m.dtg = d;
  }
  if ((st = upload(s, &s->dt_dev, dt0)) || (st = upload(s, &s->cfl_dev, cfl)) ||
      (st = upload(s, &s->time_dev, zero)) || (st = upload(s, &s->ic_dev, ic)) ||
      (st = upload(s, &s->inflow_dev, inflow)) || (st = upload(s, &s->cell_case, cell_case)) ||
      (st = upload(s, &s->bc_case, bc_case)) || (st = upload(s, &s->case_tiles, case_tiles)))
    return st;

  const size_t n5 = (size_t)5 * nc;
  const size_t tsize = s->fp32 ? sizeof(float) : sizeof(double);
  if ((st = s->alloc(&s->U[0], n5)) || (st = s->alloc(&s->U[1], n5)) || (st = s->alloc(&s->R, n5)) ||
      (st = s->alloc_bytes(&s->V, n5 * tsize)) || (st = s->alloc_bytes(&s->V2, n5 * tsize)) ||
      (st = s->alloc_bytes(&s->grad, (size_t)kRec * nc * tsize)) ||
      (s->opt.second_order_space && !s->opt.viscous &&
       (st = s->alloc_bytes(&s->fv, (size_t)kMaxSlots * n5 * tsize))) ||
      (st = s->alloc_bytes(&s->bc, (size_t)5 * std::max<long long>(nbf, 1) * tsize)) ||
      (st = s->alloc(&s->partial, (size_t)std::max(1, s->t.ntiles))) || (st = s->alloc(&s->counter, 1)) ||
      (st = s->alloc(&s->gsync, 4)) ||
      (st = s->alloc(&s->norm_partial, (size_t)5 * std::max(1, s->t.ntiles))) ||
      (st = s->alloc(&s->norm_out, (size_t)5 * ncases)))
    return st;
  if (s->counting) return MA_OK;
  MA_HIP(hipMemsetAsync(s->R, 0, n5 * sizeof(double), s->stream));
  MA_HIP(hipMemsetAsync(s->grad, 0, (size_t)kRec * nc * tsize, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  return MA_OK;
}

ma_status build(ma_solver* s, int ncases, const ma_mesh_data* const* mds) {
  RunPlan P;
  ma_status st = plan_run(ncases, mds, s->shape, P);
  if (st) return st;
  return build_planned(s, P);
}

// Workgroups per XCD for a kernel with `per` tiles or blocks per XCD.
inline int resident(const ma_solver* s, int per) {
  if (s->wg_per_cu <= 0) return per;
  return std::max(1, std::min(per, s->wg_per_cu * std::max(1, s->cus / kXcds)));
}

template <bool SO, class Inv, class Visc, class Lim, class T>
void run_stage(ma_solver* s, int irk) {
  // TimeSolverExplicitRK4.h:187-195
  constexpr double alpha[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
  constexpr double beta[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const DevMesh& m = s->m;
  const int nb = nblocks(m.nc);
  double* Un = s->U[s->cur];
  double* Uacc = s->U[1 - s->cur];
  // the stage's primitives: V for stages 0 and 2, V2 for 1 and 3; stage 0's
  // come from k_prim, the others from the previous stage's flux kernel
  // (stage 3 writes the next step's stage-0 primitives into V, so k_prim
  // runs only when V does not hold them yet)
  T* Vs = static_cast<T*>((irk & 1) ? s->V2 : s->V);
  T* Vn = static_cast<T*>((irk & 1) ? s->V : s->V2);
  if (irk == 0 && !s->v0_valid) {
    s->launch("k_prim", [&] {
      hipLaunchKernelGGL(k_prim<T>, dim3(nb), dim3(kBlock), 0, s->stream, m.nc, Un, s->R, alpha[0], Vs);
    });
  }
  // inviscid second order: extrapolated face values for the flux kernel;
  // gradients and limiters are then not stored (the getter computes the last
  // stage's from V2, see fill_gradients)
  constexpr bool FV = SO && !Visc::isViscous;
  Visc visc{effective_prandtl(s->opt)};
  const Inv inv{};
  const int nbf = m.nf - m.nif;
  if (SO || Visc::isViscous) {
    // the gradient kernel; its workgroups past the cell blocks compute the
    // boundary faces (k_bc's work, timed with it)
    s->launch(SO ? "k_grad_limiter" : "k_grad", [&] {
      const int per = (nb + kXcds - 1) / kXcds;
      const bool store = !FV;
      const dim3 grid(per * kXcds + (nbf > 0 ? nblocks(nbf) : 0));
      // the axis-aligned Green-Gauss loop when the mesh has DevMesh::axis
      // (every form that divides by the volume)
      constexpr bool kDiv = sizeof(T) == 8 || FV;
      if (kDiv && m.axis && m.full)
        hipLaunchKernelGGL((k_grad<SO, FV, Lim, T, kDiv ? 1 : 0, Inv>), grid, dim3(kBlock), 0, s->stream, m,
                           s->nbr, static_cast<const T*>(Vs), store ? static_cast<T*>(s->grad) : nullptr,
                           static_cast<T*>(s->fv), per, inv, visc.prandtl, s->inflow_dev, s->bc_case,
                           static_cast<T*>(s->bc));
      else
        hipLaunchKernelGGL((k_grad<SO, FV, Lim, T, 0, Inv>), grid, dim3(kBlock), 0, s->stream, m, s->nbr,
                           static_cast<const T*>(Vs), store ? static_cast<T*>(s->grad) : nullptr,
                           static_cast<T*>(s->fv), per, inv, visc.prandtl, s->inflow_dev, s->bc_case,
                           static_cast<T*>(s->bc));
    });
  }
  // Stage 0 accumulates from sol_n (sol_np1 == sol_n at the start of a step).
  const double* acc_in = irk == 0 ? Un : Uacc;
  if (nbf > 0 && !(SO || Visc::isViscous)) {
    s->launch("k_bc", [&] {
      hipLaunchKernelGGL((k_bc<Inv, T>), dim3(nblocks(nbf)), dim3(kBlock), 0, s->stream, m,
                         static_cast<const T*>(Vs), inv, visc.prandtl, s->inflow_dev, s->bc_case,
                         static_cast<T*>(s->bc));
    });
  }
  const size_t lds = (size_t)s->t.max_faces * 5 * sizeof(T);
  s->launch("k_flux", [&] {
    hipLaunchKernelGGL((k_flux<SO, Inv, Visc, T>), dim3(resident(s, s->t.per_xcd) * kXcds), dim3(kTile), lds,
                       s->stream, m, s->t, static_cast<const T*>(Vs), static_cast<const T*>(FV ? s->fv : s->grad),
                       static_cast<const T*>(s->bc), acc_in, Uacc, s->R,
                       s->dt_dev, beta[irk], inv, visc, Un, Vn, irk < 3 ? alpha[irk + 1] : alpha[0],
                       irk == 3 ? 1 : 0);
  });
  if (irk == 3) s->v0_valid = true;  // V now holds the next step's stage-0 primitives
}

template <bool SO, class Inv, class Visc, class Lim, class T>
void run_steps(ma_solver* s, int nsteps) {
  for (int it = 0; it < nsteps; ++it) {
    // per-case dt (CFL minimum or fixed) and time += dt
    if (s->any_cfl) {
      s->launch("k_dt_partial", [&] {
        hipLaunchKernelGGL(k_dt_partial, dim3(s->t.ntiles), dim3(kTile), 0, s->stream, s->m, s->t,
                           s->U[s->cur], s->opt.viscous, effective_prandtl(s->opt), s->partial);
      });
    }
    s->launch("k_dt_final", [&] {
      hipLaunchKernelGGL(k_dt_final, dim3(s->ncases), dim3(kBlock), 0, s->stream, s->partial, s->case_tiles,
                         s->cfl_dev, s->dt_dev, s->time_dev);
    });
    for (int irk = 0; irk < 4; ++irk) run_stage<SO, Inv, Visc, Lim, T>(s, irk);
    s->cur = 1 - s->cur;  // U_acc after the last stage is the new sol_n
    ++s->steps_done;
  }
}

// The policies of a run: inviscid flux (the reference's Roe_Flux.h; the only
// one this library instantiates), viscous flux and limiter selected from the
// options, as TimeSolverExplicitRK4 instantiates compute_face_flux and the
// limiter (TimeSolverExplicitRK4.h:388-486, StencilLimiter.h:455).
using Inviscid = ma::RoeFlux;

template <bool SO, class Visc, class T>
void dispatch_lim(ma_solver* s, int n) {
  if (s->opt.limiter == MA_LIMITER_VANALBADA)
    run_steps<SO, Inviscid, Visc, ma::VanAlbadaLimiter, T>(s, n);
  else
    run_steps<SO, Inviscid, Visc, ma::VenkatLimiter, T>(s, n);
}

template <class T>
void dispatch_t(ma_solver* s, int n) {
  const bool so = s->opt.second_order_space != 0, visc = s->opt.viscous != 0;
  if (so && visc) dispatch_lim<true, ma::NewtonianViscousFlux, T>(s, n);
  else if (so) dispatch_lim<true, ma::NoViscousFlux, T>(s, n);
  else if (visc) dispatch_lim<false, ma::NewtonianViscousFlux, T>(s, n);
  else dispatch_lim<false, ma::NoViscousFlux, T>(s, n);
}

void dispatch(ma_solver* s, int n) {
  if (s->fp32)
    dispatch_t<float>(s, n);
  else
    dispatch_t<double>(s, n);
}

// nsteps steps as one cooperative launch of k_steps: as many workgroups as
// the device keeps resident at once (no more than the largest phase has
// tiles or cell blocks), all of them walking every phase.  Waits for the
// device to be idle before the launch and for the launch itself, since a
// barrier wait that ran over its limit is reported here.
template <bool SO, class Inv, class Visc, class Lim, class T>
ma_status run_persistent(ma_solver* s, int nsteps) {
  if (nsteps == 0) return MA_OK;
  constexpr bool FV = SO && !Visc::isViscous;
  const auto kern = k_steps<SO, Inv, Visc, Lim, T>;
  const size_t lds = std::max((size_t)s->t.max_faces * 5 * sizeof(T),
                              FV ? (size_t)0 : (size_t)kBlock * kStageRow<T> * sizeof(T));
  int coop = 0, per_cu = 0, wall_khz = 0;
  MA_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, s->device));
  if (!coop) return ma::fail(MA_ERR_UNSUPPORTED, "persistent steps: the device has no cooperative launch");
  MA_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, s->device));
  hipFuncAttributes fa{};
  MA_HIP(hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(kern)));
  if (fa.localSizeBytes > 0)
    return ma::fail(MA_ERR_UNSUPPORTED, "persistent steps: the kernel uses scratch memory, which can keep "
                                        "workgroups of a cooperative launch from being resident together");
  MA_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kTile, lds));
  const int nc = s->m.nc;
  // one workgroup per cell block at most: smaller flux tiles (4x4x4 on small
  // runs) are walked several per workgroup rather than widening the grid
  // barriers (16k-cell Sod: 263 us per step with a workgroup per tile)
  const int most = std::max(nblocks(nc), std::max(nblocks(s->m.nf - s->m.nif), s->ncases));
  int grid = std::min(per_cu * s->cus, (most + kXcds - 1) / kXcds * kXcds);
  grid -= grid % kXcds;
  if (grid < kXcds || wall_khz <= 0)
    return ma::fail(MA_ERR_UNSUPPORTED, "persistent steps: too few resident workgroups");
  PersistArgs a{};
  a.m = s->m;
  a.t = s->t;
  a.nbr = s->nbr;
  a.V[0] = s->V;
  a.V[1] = s->V2;
  a.grad = s->grad;
  a.fv = s->fv;
  a.bc = s->bc;
  a.U[0] = s->U[0];
  a.U[1] = s->U[1];
  a.R = s->R;
  a.dt = s->dt_dev;
  a.time = s->time_dev;
  a.cfl = s->cfl_dev;
  a.partial = s->partial;
  a.case_tiles = s->case_tiles;
  a.inflow = s->inflow_dev;
  a.bc_case = s->bc_case;
  a.gsync = s->gsync;
  a.wait_ticks = (long long)wall_khz * 1000 * 10;  // 10 s
  a.ncases = s->ncases;
  a.nsteps = nsteps;
  a.cur0 = s->cur;
  a.prim0 = s->v0_valid ? 0 : 1;
  a.any_cfl = s->any_cfl ? 1 : 0;
  a.viscous = s->opt.viscous;
  a.per_blocks = (nblocks(nc) + kXcds - 1) / kXcds;
  a.prandtl = effective_prandtl(s->opt);
  Visc visc{a.prandtl};
  Inv inv{};
  void* args[] = {&a, &inv, &visc};
  // The grid is sized to every workgroup slot of the device, so no other
  // kernel of the process may hold slots while it is placed: wait for the
  // whole device (other solvers' streams included) first.
  MA_HIP(hipDeviceSynchronize());
  MA_HIP(hipMemsetAsync(s->gsync, 0, 4 * sizeof(unsigned), s->stream));
  MA_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3(grid), dim3(kTile), args, lds,
                                    s->stream));
  unsigned exceeded = 0;
  MA_HIP(hipMemcpyAsync(&exceeded, s->gsync + 2, sizeof(exceeded), hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  if (exceeded) return ma::fail(MA_ERR_DEVICE, "persistent steps: a grid barrier wait ran over its limit");
  s->cur ^= nsteps & 1;  // U_acc after the last stage is the new sol_n, every step
  s->steps_done += nsteps;
  s->v0_valid = true;
  return MA_OK;
}

template <bool SO, class Visc, class T>
ma_status persist_lim(ma_solver* s, int n) {
  if (s->opt.limiter == MA_LIMITER_VANALBADA) return run_persistent<SO, Inviscid, Visc, ma::VanAlbadaLimiter, T>(s, n);
  return run_persistent<SO, Inviscid, Visc, ma::VenkatLimiter, T>(s, n);
}

template <class T>
ma_status persist_t(ma_solver* s, int n) {
  const bool so = s->opt.second_order_space != 0, visc = s->opt.viscous != 0;
  if (so && visc) return persist_lim<true, ma::NewtonianViscousFlux, T>(s, n);
  if (so) return persist_lim<true, ma::NoViscousFlux, T>(s, n);
  if (visc) return persist_lim<false, ma::NewtonianViscousFlux, T>(s, n);
  return persist_lim<false, ma::NoViscousFlux, T>(s, n);
}

ma_status persist(ma_solver* s, int n) { return s->fp32 ? persist_t<float>(s, n) : persist_t<double>(s, n); }

void drop_graphs(ma_solver* s) {
  for (auto& g : s->graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
}

// Captures graph_steps steps starting from the current cur (no launch runs).
ma_status capture_graph(ma_solver* s) {
  const int cur = s->cur;
  const long long done = s->steps_done;
  hipGraph_t g = nullptr;
  MA_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
  dispatch(s, s->graph_steps);
  const hipError_t e = hipStreamEndCapture(s->stream, &g);
  s->cur = cur;  // nothing ran
  s->steps_done = done;
  if (e != hipSuccess || !g)
    return ma::fail(MA_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ei != hipSuccess) return ma::fail(MA_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
  s->graph[cur] = x;
  return MA_OK;
}

// The stage-0 primitives of U[cur] into V, when V does not hold them.
ma_status ensure_stage0_prims(ma_solver* s) {
  if (s->v0_valid) return MA_OK;
  const int nc = s->m.nc;
  const double* Un = s->U[s->cur];
  if (s->fp32)
    hipLaunchKernelGGL(k_prim<float>, dim3(nblocks(nc)), dim3(kBlock), 0, s->stream, nc, Un, s->R, 0.0,
                       static_cast<float*>(s->V));
  else
    hipLaunchKernelGGL(k_prim<double>, dim3(nblocks(nc)), dim3(kBlock), 0, s->stream, nc, Un, s->R, 0.0,
                       static_cast<double*>(s->V));
  MA_HIP(hipGetLastError());
  s->v0_valid = true;
  return MA_OK;
}

// nsteps steps: whole graphs while at least graph_steps remain, the rest as
// plain launches.  A graph starts from valid stage-0 primitives (made here
// when needed), so it never contains k_prim and leaves them valid.
ma_status advance(ma_solver* s, int nsteps) {
  if (s->graph_steps > 0 && nsteps >= s->graph_steps && std::this_thread::get_id() != s->owner)
    return ma::fail(MA_ERR_UNSUPPORTED,
                    "graph steps are captured and replayed only on the thread that created the solver; "
                    "call ma_solver_set_graph(s, 0) to step from another thread");
  if (nsteps > 0 && s->opt.second_order_space && !s->opt.viscous) s->grads_stale = true;
  if (s->persistent) return persist(s, nsteps);
  if (s->graph_steps > 0) {
    while (nsteps >= s->graph_steps) {
      ma_status st0 = ensure_stage0_prims(s);
      if (st0) return st0;
      if (!s->graph[s->cur]) {
        ma_status st = capture_graph(s);
        if (st) return st;
      }
      MA_HIP(hipGraphLaunch(s->graph[s->cur], s->stream));
      s->steps_done += s->graph_steps;
      nsteps -= s->graph_steps;
    }
  }
  dispatch(s, nsteps);
  return MA_OK;
}

template <class S>
ma_status download_aos(ma_solver* s, const S* dsrc, int ncomp, double* host) {
  const int nc = s->m.nc;
  std::vector<S> tmp((size_t)ncomp * nc);
  MA_HIP(hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * sizeof(S), hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  for (int c = 0; c < nc; ++c) {
    const int d = s->dev_of[c];
    for (int k = 0; k < ncomp; ++k) host[(size_t)c * ncomp + k] = tmp[(size_t)k * nc + d];
  }
  return MA_OK;
}

// Gradient records (kRec values per device cell) to the caller's order:
// grad (n x 5 x 3) and lim (n x 5), either may be null.
template <class S>
ma_status download_records(ma_solver* s, const S* dsrc, double* grad, double* lim) {
  const int nc = s->m.nc;
  std::vector<S> tmp((size_t)kRec * nc);
  MA_HIP(hipMemcpyAsync(tmp.data(), dsrc, tmp.size() * sizeof(S), hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  for (int c = 0; c < nc; ++c) {
    const S* r = &tmp[(size_t)s->dev_of[c] * kRec];
    if (grad)
      for (int j = 0; j < 15; ++j) grad[(size_t)c * 15 + j] = r[j];
    if (lim)
      for (int k = 0; k < 5; ++k) lim[(size_t)c * 5 + k] = r[15 + k];
  }
  return MA_OK;
}

}  // namespace

extern "C" const char* ma_last_error(void) { return ma::last_error().c_str(); }

extern "C" ma_status ma_batch_plan(int ncases, const ma_mesh_data* const* meshes, long long* stats) {
  if (!meshes || !stats || ncases < 1) return ma::fail(MA_ERR_INVALID, "plan: bad argument");
  RunPlan P;
  ma_status st = MA_OK;
  try {
    st = plan_run(ncases, meshes, kDefaultShape, P);
  } catch (const std::bad_alloc&) {
    st = ma::fail(MA_ERR_NOMEM, "plan: out of host memory");
  }
  if (st) return st;
  plan_stats(P.L, stats);
  return MA_OK;
}

extern "C" ma_status ma_mesh_plan(const ma_mesh_data* md, long long* stats) {
  if (!md) return ma::fail(MA_ERR_INVALID, "ma_mesh_plan: null argument");
  return ma_batch_plan(1, &md, stats);
}

namespace {

// Tile shape requested by the options: zeros = the default brick, a negative
// first extent = consecutive cells.
TileShape shape_of(const ma_options& o) {
  if (o.tile_shape[0] < 0) return TileShape{{0, 0, 0}};
  if (o.tile_shape[0] == 0 && o.tile_shape[1] == 0 && o.tile_shape[2] == 0) return kDefaultShape;
  return TileShape{{o.tile_shape[0], o.tile_shape[1], o.tile_shape[2]}};
}

// Launch-shape choices that do not change results: the flux tile shape and
// how many workgroups per CU the gradient and flux kernels keep resident
// (0 = one workgroup per tile or cell block).
struct Tuning {
  TileShape shape;
  int wg_per_cu;
};

ma_status create_run(int ncases, const ma_mesh_data* const* mds, const ma_options* opts, ma_solver** out,
                     const Tuning* force = nullptr);

// Times kTimed steps of a freshly built run (after kWarm untimed ones) from
// its initial state; ms per step per million cells.
ma_status time_candidate(int n, const ma_mesh_data* const* mds, const ma_options* so, const Tuning& c,
                         long long cells, double* v, bool* lattice) {
  constexpr int kWarm = 2, kTimed = 8;
  ma_solver* t = nullptr;
  ma_status st = create_run(n, mds, so, &t, &c);
  if (st) return st;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  bool ok = ma_solver_initialize(t) == MA_OK && hipEventCreate(&e0) == hipSuccess &&
            hipEventCreate(&e1) == hipSuccess;
  if (ok) {
    dispatch(t, kWarm);
    ok = hipEventRecord(e0, t->stream) == hipSuccess;
    dispatch(t, kTimed);
    ok = ok && hipEventRecord(e1, t->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
         hipEventElapsedTime(&ms, e0, e1) == hipSuccess;
  }
  *lattice = t->all_lattice;
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  ma_solver_destroy(t);
  if (!ok) return ma::fail(MA_ERR_DEVICE, "autotune: timing run failed");
  *v = ms / kTimed / (cells * 1e-6);
  return MA_OK;
}

// Autotuner: builds a sample of the cases (every stride-th case, at least
// 2^21 cells, or all of them; spread over the batch so that it sees the
// batch's mix of case sizes) with every candidate tile shape, then the best
// shape with several resident-workgroup counts, times a few steps of each
// from the initial state with HIP events, and returns the fastest choice and
// its time in ms per step per million cells.  Results do not depend on these choices
// (every cell sums its faces in the same order whatever the tiles and
// however the workgroups walk them), only the speed does.
ma_status autotune(int ncases, const ma_mesh_data* const* mds, const ma_options* opts, Tuning* best,
                   double* best_ms) {
  static const TileShape cand[] = {{{4, 8, 8}}, {{8, 8, 4}}, {{8, 4, 8}}, {{4, 4, 16}}, {{16, 4, 4}},
                                   {{2, 8, 16}}, {{8, 2, 16}}, {{4, 16, 4}}, {{2, 16, 8}}, {{4, 4, 8}},
                                   {{4, 4, 4}}, {{0, 0, 0}}};
  static const int wgs[] = {1, 2, 3, 4};
  long long total = 0;
  for (int k = 0; k < ncases; ++k) total += mds[k]->num_owned_cells;
  const int stride = (int)std::max<long long>(1, std::min<long long>(ncases, total >> 21));
  std::vector<const ma_mesh_data*> smd;
  std::vector<ma_options> so;
  long long cells = 0;
  for (int k = 0; k < ncases && cells < (1ll << 21); k += stride) {
    smd.push_back(mds[k]);
    so.push_back(opts[k]);
    cells += mds[k]->num_owned_cells;
  }
  const int n = (int)smd.size();
  for (auto& o : so) o.autotune = 0;
  *best_ms = 0.0;
  *best = Tuning{kDefaultShape, 0};
  double best_v = 1e300;
  auto consider = [&](const Tuning& c, bool* lattice) -> ma_status {
    double v = 0.0;
    ma_status st = time_candidate(n, smd.data(), so.data(), c, cells, &v, lattice);
    if (st == MA_ERR_UNSUPPORTED) return MA_OK;  // e.g. a tile with too many faces for LDS
    if (st) return st;
    if (v < best_v) {
      best_v = v;
      *best = c;
      *best_ms = v;
    }
    return MA_OK;
  };
  for (const TileShape& c : cand) {
    bool lattice = true;
    ma_status st = consider(Tuning{c, 0}, &lattice);
    if (st) return st;
    if (!lattice) break;  // without a lattice every shape is the same layout
  }
  const TileShape shape = best->shape;
  for (int w : wgs) {
    bool lattice = true;
    ma_status st = consider(Tuning{shape, w}, &lattice);
    if (st) return st;
  }
  return MA_OK;
}

// Shared body of ma_solver_create and ma_batch_create; `force` fixes the
// tile shape (no autotuning).
ma_status create_run(int ncases, const ma_mesh_data* const* mds, const ma_options* opts, ma_solver** out,
                     const Tuning* force) {
  const ma_options& o0 = opts[0];
  for (int k = 0; k < ncases; ++k) {
    ma_status st = check_mesh(mds[k], "create");
    if (st || (st = check_options(&opts[k]))) return st;
    const ma_options& o = opts[k];
    if (o.second_order_space != o0.second_order_space || o.viscous != o0.viscous ||
        o.limiter != o0.limiter || o.precision != o0.precision ||
        effective_prandtl(o) != effective_prandtl(o0))
      return ma::fail(MA_ERR_INVALID,
                      "batched cases must share second_order_space, viscous, limiter, precision and prandtl");
  }
  Tuning tune = force ? *force : Tuning{shape_of(o0), o0.workgroups_per_cu};
  if (!valid_shape(tune.shape)) return ma::fail(MA_ERR_INVALID, "tile shape: extents >= 1 with product <= 256");
  if (tune.wg_per_cu < 0 || tune.wg_per_cu > 64)
    return ma::fail(MA_ERR_INVALID, "workgroups_per_cu must be in 0..64");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
    return ma::fail(MA_ERR_DEVICE, "no HIP device available");
  if (!force && !o0.autotune && o0.tile_shape[0] == 0 && o0.tile_shape[1] == 0 && o0.tile_shape[2] == 0) {
    long long cells = 0;
    for (int k = 0; k < ncases; ++k) cells += mds[k]->num_owned_cells;
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess &&
        cells < (long long)cus * kTile)
      tune.shape = kSmallRunShape;
  }
  double tuned_ms = 0.0;
  if (!force && o0.autotune) {
    ma_status st = MA_OK;
    try {
      st = autotune(ncases, mds, opts, &tune, &tuned_ms);
    } catch (const std::bad_alloc&) {
      st = ma::fail(MA_ERR_NOMEM, "autotune: out of host memory");
    }
    if (st) return st;
  }
  ma_solver* s = new (std::nothrow) ma_solver();
  if (!s) return ma::fail(MA_ERR_NOMEM, "create: out of host memory");
  ma_status st = MA_OK;
  try {
    s->shape = tune.shape;
    s->wg_per_cu = tune.wg_per_cu;
    s->tuned_ms = tuned_ms;
    s->opt = o0;
    s->ncases = ncases;
    s->case_opt.assign(opts, opts + ncases);
    s->fp32 = o0.precision == MA_PRECISION_FP32;
    (void)hipGetDevice(&s->device);
    (void)hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
      st = ma::fail(MA_ERR_DEVICE, "hipStreamCreate failed");
    else
      st = build(s, ncases, mds);
  } catch (const std::bad_alloc&) {
    st = ma::fail(MA_ERR_NOMEM, "create: out of host memory");
  }
  if (st) {
    ma_solver_destroy(s);
    return st;
  }
  *out = s;
  return MA_OK;
}

}  // namespace

extern "C" ma_status ma_solver_create(const ma_mesh_data* md, const ma_options* opt, ma_solver** out) {
  if (!md || !opt || !out) return ma::fail(MA_ERR_INVALID, "ma_solver_create: null argument");
  *out = nullptr;
  return create_run(1, &md, opt, out);
}

extern "C" ma_status ma_batch_create(int ncases, const ma_mesh_data* const* meshes, const ma_options* opts,
                                     ma_solver** out) {
  if (!meshes || !opts || !out || ncases < 1) return ma::fail(MA_ERR_INVALID, "ma_batch_create: bad argument");
  *out = nullptr;
  return create_run(ncases, meshes, opts, out);
}

extern "C" ma_status ma_solver_initialize(ma_solver* s) {
  if (!s) return ma::fail(MA_ERR_INVALID, "null solver");
  s->cur = 0;
  hipLaunchKernelGGL(k_init, dim3(nblocks(s->m.nc)), dim3(kBlock), 0, s->stream, s->m, s->xc, s->org, s->ic_dev,
                     s->cell_case, s->U[0]);
  MA_HIP(hipGetLastError());
  const size_t n5 = (size_t)5 * s->m.nc;
  MA_HIP(hipMemsetAsync(s->R, 0, n5 * sizeof(double), s->stream));
  MA_HIP(hipMemsetAsync(s->time_dev, 0, s->ncases * sizeof(double), s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  s->steps_done = 0;
  s->v0_valid = false;
  return MA_OK;
}

extern "C" ma_status ma_solver_set_state(ma_solver* s, const double* U) {
  if (!s || !U) return ma::fail(MA_ERR_INVALID, "null argument");
  const int nc = s->m.nc;
  std::vector<double> tmp((size_t)5 * nc);
  for (int c = 0; c < nc; ++c)
    for (int k = 0; k < 5; ++k) tmp[(size_t)k * nc + s->dev_of[c]] = U[(size_t)c * 5 + k];
  s->cur = 0;
  MA_HIP(hipMemcpyAsync(s->U[0], tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice, s->stream));
  MA_HIP(hipMemsetAsync(s->R, 0, tmp.size() * sizeof(double), s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  s->v0_valid = false;
  return MA_OK;
}

extern "C" ma_status ma_solver_step(ma_solver* s, int nsteps) {
  if (!s) return ma::fail(MA_ERR_INVALID, "null solver");
  if (nsteps < 0) return ma::fail(MA_ERR_INVALID, "negative step count");
  ma_status st = advance(s, nsteps);
  if (st) return st;
  MA_HIP(hipGetLastError());
  return MA_OK;
}

extern "C" ma_status ma_solver_tuning(ma_solver* s, int* tile_shape, int* workgroups_per_cu, double* tuned_ms) {
  if (!s) return ma::fail(MA_ERR_INVALID, "ma_solver_tuning: null solver");
  const bool consecutive = s->shape.b[0] == 0;
  if (tile_shape)
    for (int k = 0; k < 3; ++k) tile_shape[k] = consecutive ? -1 : s->shape.b[k];
  if (workgroups_per_cu) *workgroups_per_cu = s->wg_per_cu;
  if (tuned_ms) *tuned_ms = s->tuned_ms;
  return MA_OK;
}

extern "C" ma_status ma_solver_set_graph(ma_solver* s, int steps_per_graph) {
  if (!s || steps_per_graph < 0 || (steps_per_graph & 1))
    return ma::fail(MA_ERR_INVALID, "ma_solver_set_graph: steps_per_graph must be an even count >= 0");
  MA_HIP(hipStreamSynchronize(s->stream));
  drop_graphs(s);
  s->graph_steps = steps_per_graph;
  if (steps_per_graph > 0) s->persistent = false;
  return MA_OK;
}

extern "C" ma_status ma_solver_set_persistent(ma_solver* s, int enabled) {
  if (!s || enabled < 0 || enabled > 1)
    return ma::fail(MA_ERR_INVALID, "ma_solver_set_persistent: enabled must be 0 or 1");
  MA_HIP(hipStreamSynchronize(s->stream));
  if (enabled) {
    drop_graphs(s);
    s->graph_steps = 0;
  }
  s->persistent = enabled != 0;
  return MA_OK;
}

extern "C" ma_status ma_solver_synchronize(ma_solver* s) {
  if (!s) return ma::fail(MA_ERR_INVALID, "null solver");
  MA_HIP(hipStreamSynchronize(s->stream));
  return MA_OK;
}

extern "C" ma_status ma_solver_get_state(ma_solver* s, double* U) {
  if (!s || !U) return ma::fail(MA_ERR_INVALID, "null argument");
  return download_aos(s, s->U[s->cur], 5, U);
}

extern "C" ma_status ma_solver_get_residual(ma_solver* s, double* R) {
  if (!s || !R) return ma::fail(MA_ERR_INVALID, "null argument");
  return download_aos(s, s->R, 5, R);
}

namespace {

template <class Lim, class T>
void stage3_gradients(ma_solver* s) {
  const int nb = nblocks(s->m.nc);
  const int per = (nb + kXcds - 1) / kXcds;
  // no boundary-face blocks: the grid is the cell blocks alone
  hipLaunchKernelGGL((k_grad<true, true, Lim, T, 0, Inviscid>), dim3(per * kXcds), dim3(kBlock), 0, s->stream,
                     s->m, s->nbr, static_cast<const T*>(s->V2), static_cast<T*>(s->grad), static_cast<T*>(s->fv),
                     per, Inviscid{}, 0.0, nullptr, nullptr, nullptr);
}

// Inviscid second order stores no gradients or limiters while stepping.  The
// last stage's primitives stay in V2 until the next step starts, so the same
// kernel run on them gives the values that stage computed (and rewrites the
// face values with the same bits).
ma_status fill_gradients(ma_solver* s) {
  if (!s->grads_stale) return MA_OK;
  const bool va = s->opt.limiter == MA_LIMITER_VANALBADA;
  if (s->fp32) {
    if (va) stage3_gradients<ma::VanAlbadaLimiter, float>(s);
    else stage3_gradients<ma::VenkatLimiter, float>(s);
  } else {
    if (va) stage3_gradients<ma::VanAlbadaLimiter, double>(s);
    else stage3_gradients<ma::VenkatLimiter, double>(s);
  }
  MA_HIP(hipGetLastError());
  s->grads_stale = false;
  return MA_OK;
}

}  // namespace

extern "C" ma_status ma_solver_get_gradients(ma_solver* s, double* grad, double* lim) {
  if (!s) return ma::fail(MA_ERR_INVALID, "null solver");
  ma_status st = fill_gradients(s);
  if (st) return st;
  return s->fp32 ? download_records(s, static_cast<const float*>(s->grad), grad, lim)
                 : download_records(s, static_cast<const double*>(s->grad), grad, lim);
}

namespace {

ma_status case_time(ma_solver* s, int k, double* time, double* last_dt) {
  double t = 0.0, d = 0.0;
  MA_HIP(hipMemcpyAsync(&t, s->time_dev + k, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipMemcpyAsync(&d, s->dt_dev + k, sizeof(double), hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  if (time) *time = t;
  if (last_dt) *last_dt = s->steps_done ? d : 0.0;
  return MA_OK;
}

}  // namespace

extern "C" ma_status ma_solver_get_time(ma_solver* s, double* time, double* last_dt) {
  if (!s) return ma::fail(MA_ERR_INVALID, "null solver");
  return case_time(s, 0, time, last_dt);
}

extern "C" ma_status ma_batch_get_case_time(ma_solver* s, int k, double* time, double* last_dt) {
  if (!s || k < 0 || k >= s->ncases) return ma::fail(MA_ERR_INVALID, "ma_batch_get_case_time: bad argument");
  return case_time(s, k, time, last_dt);
}

extern "C" ma_status ma_batch_get_case_state(ma_solver* s, int k, double* U) {
  if (!s || !U || k < 0 || k >= s->ncases) return ma::fail(MA_ERR_INVALID, "ma_batch_get_case_state: bad argument");
  const int nc = s->m.nc, c0 = s->case_cell_off[k], c1 = s->case_cell_off[k + 1];
  // the case's device cells are one contiguous range
  int lo = nc, hi = -1;
  for (int c = c0; c < c1; ++c) {
    lo = std::min(lo, s->dev_of[c]);
    hi = std::max(hi, s->dev_of[c]);
  }
  const int n = hi - lo + 1;
  std::vector<double> tmp((size_t)5 * n);
  for (int q = 0; q < 5; ++q)
    MA_HIP(hipMemcpyAsync(tmp.data() + (size_t)q * n, s->U[s->cur] + (size_t)q * nc + lo, n * sizeof(double),
                          hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  for (int c = c0; c < c1; ++c)
    for (int q = 0; q < 5; ++q) U[(size_t)(c - c0) * 5 + q] = tmp[(size_t)q * n + s->dev_of[c] - lo];
  return MA_OK;
}

extern "C" ma_status ma_batch_sizes(ma_solver* s, int* ncases, int* case_cell_offsets) {
  if (!s || !ncases) return ma::fail(MA_ERR_INVALID, "ma_batch_sizes: null argument");
  *ncases = s->ncases;
  if (case_cell_offsets)
    for (int k = 0; k <= s->ncases; ++k) case_cell_offsets[k] = s->case_cell_off[k];
  return MA_OK;
}

extern "C" ma_status ma_solver_residual_norms(ma_solver* s, double* norms) {
  if (!s || !norms) return ma::fail(MA_ERR_INVALID, "null argument");
  if (s->t.ntiles > 0) {
    hipLaunchKernelGGL(k_norm_partial, dim3(s->t.ntiles), dim3(kTile), 0, s->stream, s->t, s->m.nc, s->R,
                       s->norm_partial);
    hipLaunchKernelGGL(k_norm_final, dim3((5 * s->ncases + 255) / 256), dim3(256), 0, s->stream, s->norm_partial,
                       s->case_tiles, s->ncases, s->norm_out);
    MA_HIP(hipGetLastError());
  }
  MA_HIP(hipMemcpyAsync(norms, s->norm_out, (size_t)5 * s->ncases * sizeof(double), hipMemcpyDeviceToHost,
                        s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  return MA_OK;
}

// ------------------------------------------- residual norms, no allocation --
// ma_solver_residual_norms_noalloc keeps its per-tile partial sums in this
// file-scope array instead of the solver's own buffer.  Every GPU the library
// runs on holds its own copy until the process ends.  In host code the name
// `scratch` is only a stand-in for the symbol, not a device address: the
// address on each device is looked up once per device with
// hipGetSymbolAddress and kept in g_scratch, and only that pointer reaches
// kernels.  All calls on one GPU share its copy, so calls that overlap on one
// GPU (different streams, or threads) race; calls queued on one stream do not.
// Nothing in it needs to start at zero: k_norm_partial writes all 5 sums of
// every tile before k_norm_final reads them, on the same stream.
constexpr int kScratchTiles = 1 << 18;  // 5 doubles per tile: 10 MiB per GPU
constexpr int kScratchDevices = 64;
static __device__ __align__(256) double scratch[(size_t)5 * kScratchTiles];

static std::mutex g_scratch_mu;
static double* g_scratch[kScratchDevices] = {};

// Device address of `scratch` on `device`: looked up on the first call for
// that device, with that device current (and the caller's device restored).
static ma_status scratch_on(int device, double** out) {
  if (device < 0 || device >= kScratchDevices)
    return ma::fail(MA_ERR_UNSUPPORTED, "residual norms without allocation: device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    int prev = -1;
    MA_HIP(hipGetDevice(&prev));
    if (prev != device) MA_HIP(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (prev != device) (void)hipSetDevice(prev);
    if (e != hipSuccess || !p)
      return ma::fail(MA_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return MA_OK;
}

extern "C" ma_status ma_solver_residual_norms_noalloc(ma_solver* s, double* norms) {
  if (!s || !norms) return ma::fail(MA_ERR_INVALID, "null argument");
  if (s->t.ntiles > kScratchTiles)
    return ma::fail(MA_ERR_INVALID, "ma_solver_residual_norms_noalloc: " + std::to_string(s->t.ntiles) +
                                        " flux tiles, more than the " + std::to_string(kScratchTiles) +
                                        " its static scratch holds; use ma_solver_residual_norms");
  double* part = nullptr;
  if (ma_status st = scratch_on(s->device, &part)) return st;
  if (s->t.ntiles > 0)
    hipLaunchKernelGGL(k_norm_partial, dim3(s->t.ntiles), dim3(kTile), 0, s->stream, s->t, s->m.nc, s->R, part);
  hipLaunchKernelGGL(k_norm_final, dim3((5 * s->ncases + 255) / 256), dim3(256), 0, s->stream, part,
                     s->case_tiles, s->ncases, norms);
  MA_HIP(hipGetLastError());
  return MA_OK;
}

extern "C" ma_status ma_solver_check_finite(ma_solver* s, long long* nonfinite) {
  if (!s || !nonfinite) return ma::fail(MA_ERR_INVALID, "null argument");
  const int n = 5 * s->m.nc;
  MA_HIP(hipMemsetAsync(s->counter, 0, sizeof(unsigned long long), s->stream));
  hipLaunchKernelGGL(k_count_nonfinite, dim3(nblocks(n)), dim3(kBlock), 0, s->stream, n, s->U[s->cur],
                     s->counter);
  MA_HIP(hipGetLastError());
  unsigned long long h = 0;
  MA_HIP(hipMemcpyAsync(&h, s->counter, sizeof(h), hipMemcpyDeviceToHost, s->stream));
  MA_HIP(hipStreamSynchronize(s->stream));
  *nonfinite = (long long)h;
  return h ? ma::fail(MA_ERR_NONFINITE, "non-finite values in the solution") : MA_OK;
}

extern "C" ma_status ma_solver_stream(ma_solver* s, void** stream) {
  if (!s || !stream) return ma::fail(MA_ERR_INVALID, "null argument");
  *stream = (void*)s->stream;
  return MA_OK;
}

extern "C" ma_status ma_solver_sizes(ma_solver* s, long long* counts) {
  if (!s || !counts) return ma::fail(MA_ERR_INVALID, "null argument");
  counts[0] = s->m.nc;
  counts[1] = s->m.nif;
  counts[2] = s->m.nf - s->m.nif;
  counts[3] = (long long)s->bytes;
  return MA_OK;
}

extern "C" int ma_solver_profile(ma_solver* s, int nsteps, const char** names, int* launches,
                                 double* mean_ms, int max) {
  if (!s || nsteps < 0) {
    ma::fail(MA_ERR_INVALID, "ma_solver_profile: bad argument");
    return -1;
  }
  (void)hipStreamSynchronize(s->stream);
  s->timing.on = true;
  s->timing.recs.clear();
  dispatch(s, nsteps);
  s->timing.on = false;
  if (hipStreamSynchronize(s->stream) != hipSuccess) {
    ma::fail(MA_ERR_DEVICE, "ma_solver_profile: stream error");
    return -1;
  }
  std::vector<std::string> order;
  std::map<std::string, std::pair<int, double>> agg;
  for (auto& r : s->timing.recs) {
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, r.a, r.b);
    if (!agg.count(r.name)) order.push_back(r.name);
    agg[r.name].first += 1;
    agg[r.name].second += ms;
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  s->timing.recs.clear();
  static const char* kNames[] = {"k_prim", "k_grad", "k_grad_limiter", "k_bc", "k_flux",
                                 "k_dt_partial", "k_dt_final"};
  int w = 0;
  for (auto& n : order) {
    if (w >= max) break;
    const char* stable = nullptr;
    for (const char* k : kNames)
      if (n == k) stable = k;
    names[w] = stable;
    launches[w] = agg[n].first;
    mean_ms[w] = agg[n].second / agg[n].first;
    ++w;
  }
  return w;
}

extern "C" ma_status ma_measure_copy(long long bytes, int reps, double* ms_per_copy) {
  if (bytes < 16 || reps < 1 || !ms_per_copy) return ma::fail(MA_ERR_INVALID, "ma_measure_copy: bad argument");
  const long long n2 = bytes / 16;
  const long long blocks = (n2 + kBlock - 1) / kBlock;
  if (blocks > 0x7fffffffll) return ma::fail(MA_ERR_INVALID, "ma_measure_copy: too many bytes");
  const unsigned grid = (unsigned)blocks;
  double2 *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  ma_status rc = MA_OK;
  float ms = 0.f;
  if (hipMalloc((void**)&a, n2 * 16) != hipSuccess || hipMalloc((void**)&b, n2 * 16) != hipSuccess) {
    rc = ma::fail(MA_ERR_NOMEM, "ma_measure_copy: hipMalloc failed");
  } else if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess ||
             hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = ma::fail(MA_ERR_DEVICE, "ma_measure_copy: stream/event creation failed");
  } else {
    (void)hipMemsetAsync(a, 0, n2 * 16, st);
    hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, n2, a, b);  // warm-up
    (void)hipEventRecord(e0, st);
    for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy, dim3(grid), dim3(kBlock), 0, st, n2, a, b);
    (void)hipEventRecord(e1, st);
    if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
      rc = ma::fail(MA_ERR_DEVICE, "ma_measure_copy: kernel failed");
    *ms_per_copy = ms / reps;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

extern "C" void ma_solver_destroy(ma_solver* s) {
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  drop_graphs(s);
  for (void* p : s->allocs) (void)hipFree(p);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

// ------------------------------------------- library-internal interface --
// Used by the served form (server.hip): a single case planned on the host,
// its exact device size known before any device memory is taken, then built
// inside a region the caller hands over.

namespace ma {

// The device layout of one mesh, shareable between the runs of that mesh.
struct LayoutPlan {
  RunPlan P;
  TileShape shape;
  size_t host_bytes = 0;
};

struct PlannedCase {
  std::shared_ptr<const LayoutPlan> lp;
  ma_options opt;
  size_t bytes = 0;
};

namespace {
template <class V>
size_t vbytes(const V& v) {
  return v.size() * sizeof(typename V::value_type);
}
}  // namespace

ma_status plan_layout_shared(const ma_mesh_data* md, const ma_options* opt, std::shared_ptr<const LayoutPlan>* out,
                             long long* cells, size_t* host_bytes) {
  ma_status st = check_mesh(md, "serve");
  if (st || (st = check_options(opt))) return st;
  const TileShape shape = shape_of(*opt);
  if (!valid_shape(shape)) return ma::fail(MA_ERR_INVALID, "tile shape: extents >= 1 with product <= 256");
  try {
    std::shared_ptr<LayoutPlan> lp = std::make_shared<LayoutPlan>();
    lp->shape = shape;
    if ((st = plan_run(1, &md, shape, lp->P))) return st;
    const Layout& L = lp->P.L;
    lp->host_bytes = vbytes(L.dev_of) + vbytes(L.org_of) + vbytes(L.tile_cptr) + vbytes(L.tptr) +
                     vbytes(L.tfaces) + vbytes(L.slot) + vbytes(L.tslot) + vbytes(L.fl) + vbytes(L.fr) +
                     vbytes(L.fn) + vbytes(L.ffr) + vbytes(L.fx) + vbytes(L.fwall) + vbytes(L.vol) +
                     vbytes(L.xc) + vbytes(L.fkind) + vbytes(L.fgeo) + vbytes(L.geo) +
                     vbytes(lp->P.tile_case) + vbytes(lp->P.case_tiles) +
                     vbytes(lp->P.bc_case) + vbytes(lp->P.cell_case) + vbytes(lp->P.case_cell_off);
    *cells = L.nc;
    *host_bytes = lp->host_bytes;
    *out = std::move(lp);
  } catch (const std::bad_alloc&) {
    return ma::fail(MA_ERR_NOMEM, "serve: out of host memory");
  }
  return MA_OK;
}

ma_status plan_case_with(const std::shared_ptr<const LayoutPlan>& lp, const ma_options* opt, PlannedCase** out,
                         size_t* bytes) {
  *out = nullptr;
  ma_status st = check_options(opt);
  if (st) return st;
  PlannedCase* pc = new (std::nothrow) PlannedCase();
  if (!pc) return ma::fail(MA_ERR_NOMEM, "serve: out of host memory");
  try {
    pc->lp = lp;
    pc->opt = *opt;
    ma_solver count;  // sizes only: no device memory, no stream
    count.counting = true;
    count.opt = *opt;
    count.ncases = 1;
    count.case_opt.assign(1, *opt);
    count.fp32 = opt->precision == MA_PRECISION_FP32;
    count.shape = lp->shape;
    st = build_planned(&count, lp->P);
    pc->bytes = count.arena_used;
  } catch (const std::bad_alloc&) {
    st = ma::fail(MA_ERR_NOMEM, "serve: out of host memory");
  }
  if (st) {
    delete pc;
    return st;
  }
  *out = pc;
  *bytes = pc->bytes;
  return MA_OK;
}

ma_status plan_case(const ma_mesh_data* md, const ma_options* opt, PlannedCase** out, size_t* bytes) {
  *out = nullptr;
  std::shared_ptr<const LayoutPlan> lp;
  long long cells = 0;
  size_t hb = 0;
  ma_status st = plan_layout_shared(md, opt, &lp, &cells, &hb);
  return st ? st : plan_case_with(lp, opt, out, bytes);
}

void destroy_plan(PlannedCase* pc) { delete pc; }

ma_status create_from_plan(PlannedCase* pc, void* region, size_t region_bytes, ma_solver** out) {
  *out = nullptr;
  ma_solver* s = new (std::nothrow) ma_solver();
  if (!s) return ma::fail(MA_ERR_NOMEM, "serve: out of host memory");
  ma_status st = MA_OK;
  try {
    s->shape = pc->lp->shape;
    s->wg_per_cu = pc->opt.workgroups_per_cu;
    s->opt = pc->opt;
    s->ncases = 1;
    s->case_opt.assign(1, pc->opt);
    s->fp32 = pc->opt.precision == MA_PRECISION_FP32;
    s->arena_base = static_cast<char*>(region);
    s->arena_size = region_bytes;
    (void)hipGetDevice(&s->device);
    (void)hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, s->device);
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
      st = ma::fail(MA_ERR_DEVICE, "hipStreamCreate failed");
    else
      st = build_planned(s, pc->lp->P);
  } catch (const std::bad_alloc&) {
    st = ma::fail(MA_ERR_NOMEM, "serve: out of host memory");
  }
  if (st) {
    ma_solver_destroy(s);
    return st;
  }
  *out = s;
  return MA_OK;
}

}  // namespace ma

// Multi-case launches (continuous batching of the served form).
namespace ma {

struct Multi {
  ma_options scheme{};
  bool fp32 = false, any_cfl = false;
  int capacity = 0, n = 0;
  hipStream_t stream = nullptr;
  CaseView* d_views = nullptr;  // capacity
  int* d_start = nullptr;       // 3 x (capacity + 1): cell blocks, boundary-face blocks, tiles
  // pinned staging, two sets used in turn so that a new table can be staged
  // while the previous upload is still queued behind running steps
  CaseView* h_views[2] = {nullptr, nullptr};
  int* h_start[2] = {nullptr, nullptr};
  hipEvent_t uploaded[2] = {nullptr, nullptr};
  int slot = 0;
  int totals[3] = {0, 0, 0};
  int max_faces = 1;
  std::vector<ma_solver*> cases;  // the group's solvers, in table order
  bool any_prim0 = false;         // some case needs its stage-0 primitives computed
};

ma_status multi_create(const ma_options& scheme, int capacity, hipStream_t stream, Multi** out) {
  *out = nullptr;
  Multi* M = new (std::nothrow) Multi();
  if (!M) return ma::fail(MA_ERR_NOMEM, "multi: out of host memory");
  M->scheme = scheme;
  M->fp32 = scheme.precision == MA_PRECISION_FP32;
  M->capacity = capacity;
  M->stream = stream;
  const size_t nstart = (size_t)3 * (capacity + 1);
  bool ok = hipMalloc((void**)&M->d_views, capacity * sizeof(CaseView)) == hipSuccess &&
            hipMalloc((void**)&M->d_start, nstart * sizeof(int)) == hipSuccess;
  for (int b = 0; b < 2 && ok; ++b)
    ok = hipHostMalloc((void**)&M->h_views[b], capacity * sizeof(CaseView), hipHostMallocDefault) == hipSuccess &&
         hipHostMalloc((void**)&M->h_start[b], nstart * sizeof(int), hipHostMallocDefault) == hipSuccess &&
         hipEventCreateWithFlags(&M->uploaded[b], hipEventDisableTiming) == hipSuccess;
  if (!ok) {
    multi_destroy(M);
    return ma::fail(MA_ERR_NOMEM, "multi: device or pinned allocation failed");
  }
  *out = M;
  return MA_OK;
}

void multi_destroy(Multi* M) {
  if (!M) return;
  for (int b = 0; b < 2; ++b) {
    if (M->uploaded[b]) {
      (void)hipEventSynchronize(M->uploaded[b]);
      (void)hipEventDestroy(M->uploaded[b]);
    }
    if (M->h_views[b]) (void)hipHostFree(M->h_views[b]);
    if (M->h_start[b]) (void)hipHostFree(M->h_start[b]);
  }
  if (M->d_views) (void)hipFree(M->d_views);
  if (M->d_start) (void)hipFree(M->d_start);
  delete M;
}

ma_status multi_set(Multi* M, ma_solver* const* cases, const int* parity, int n) {
  if (n < 0 || n > M->capacity) return ma::fail(MA_ERR_INVALID, "multi: too many cases");
  const int b = M->slot;
  MA_HIP(hipEventSynchronize(M->uploaded[b]));  // this staging set's last upload has run
  int* st = M->h_start[b];
  const int cap1 = M->capacity + 1;
  st[0] = st[cap1] = st[2 * cap1] = 0;
  M->any_cfl = false;
  M->max_faces = 1;
  M->any_prim0 = false;
  M->cases.assign(cases, cases + n);
  for (int k = 0; k < n; ++k) {
    ma_solver* s = cases[k];
    if (s->ncases != 1 || s->fp32 != M->fp32) return ma::fail(MA_ERR_INVALID, "multi: incompatible case");
    CaseView& v = M->h_views[b][k];
    v.m = s->m;
    v.t = s->t;
    v.U[0] = s->U[0];
    v.U[1] = s->U[1];
    v.R = s->R;
    v.V[0] = s->V;
    v.V[1] = s->V2;
    v.grad = s->grad;
    v.fv = s->fv;
    v.bc = s->bc;
    v.dt = s->dt_dev;
    v.time = s->time_dev;
    v.partial = s->partial;
    v.cfl = s->cfl_dev;
    v.case_tiles = s->case_tiles;
    v.inflow = s->inflow_dev;
    v.bc_case = s->bc_case;
    v.nbr = s->nbr;
    v.parity = parity[k];
    v.prim0 = s->v0_valid ? 0 : 1;
    M->any_prim0 = M->any_prim0 || !s->v0_valid;
    st[k + 1] = st[k] + nblocks(s->m.nc);
    st[cap1 + k + 1] = st[cap1 + k] + nblocks(s->m.nf - s->m.nif);
    st[2 * cap1 + k + 1] = st[2 * cap1 + k] + s->t.ntiles;
    M->any_cfl = M->any_cfl || s->any_cfl;
    M->max_faces = std::max(M->max_faces, s->t.max_faces);
  }
  M->n = n;
  M->totals[0] = st[n];
  M->totals[1] = st[cap1 + n];
  M->totals[2] = st[2 * cap1 + n];
  if (n > 0) {
    MA_HIP(hipMemcpyAsync(M->d_views, M->h_views[b], n * sizeof(CaseView), hipMemcpyHostToDevice, M->stream));
    MA_HIP(hipMemcpyAsync(M->d_start, st, (size_t)3 * cap1 * sizeof(int), hipMemcpyHostToDevice, M->stream));
  }
  MA_HIP(hipEventRecord(M->uploaded[b], M->stream));
  M->slot = 1 - b;
  return MA_OK;
}

}  // namespace ma

namespace {

template <bool SO, class Inv, class Visc, class Lim, class T>
void multi_step_t(ma::Multi* M, int parity) {
  constexpr double alpha[4] = {0.0, 1.0 / 2.0, 1.0 / 2.0, 1.0};
  constexpr double beta[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};
  const int n = M->n, cap1 = M->capacity + 1;
  const int* sc = M->d_start;
  const int* sb = M->d_start + cap1;
  const int* stl = M->d_start + 2 * cap1;
  const double prandtl = effective_prandtl(M->scheme);
  if (M->any_cfl)
    hipLaunchKernelGGL(k_dt_partial_multi, dim3(M->totals[2]), dim3(kTile), 0, M->stream, M->d_views, stl, n,
                       parity, M->scheme.viscous, prandtl);
  hipLaunchKernelGGL(k_dt_final_multi, dim3(n), dim3(kBlock), 0, M->stream, M->d_views);
  Visc visc{prandtl};
  const Inv inv{};
  const int per_cells = (M->totals[0] + kXcds - 1) / kXcds;
  const int per_tiles = (M->totals[2] + kXcds - 1) / kXcds;
  const size_t lds = (size_t)M->max_faces * 5 * sizeof(T);
  if (M->any_prim0)
    hipLaunchKernelGGL(k_prim_multi<T>, dim3(M->totals[0]), dim3(kBlock), 0, M->stream, M->d_views, sc, n,
                       alpha[0], parity);
  for (int irk = 0; irk < 4; ++irk) {
    if (SO || Visc::isViscous)
      hipLaunchKernelGGL((k_grad_multi<SO, SO && !Visc::isViscous, Lim, T>), dim3(per_cells * kXcds),
                         dim3(kBlock), 0, M->stream, M->d_views, sc, n, per_cells, irk & 1,
                         Visc::isViscous ? 1 : 0);
    if (M->totals[1] > 0)
      hipLaunchKernelGGL((k_bc_multi<Inv, T>), dim3(M->totals[1]), dim3(kBlock), 0, M->stream, M->d_views, sb, n,
                         inv, prandtl, irk & 1);
    hipLaunchKernelGGL((k_flux_multi<SO, Inv, Visc, T>), dim3(per_tiles * kXcds), dim3(kTile), lds, M->stream,
                       M->d_views, stl, n, per_tiles, parity, irk, beta[irk], irk < 3 ? alpha[irk + 1] : 0.0,
                       inv, visc);
  }
}

template <bool SO, class Visc, class T>
void multi_lim(ma::Multi* M, int parity) {
  if (M->scheme.limiter == MA_LIMITER_VANALBADA)
    multi_step_t<SO, Inviscid, Visc, ma::VanAlbadaLimiter, T>(M, parity);
  else
    multi_step_t<SO, Inviscid, Visc, ma::VenkatLimiter, T>(M, parity);
}

template <class T>
void multi_t(ma::Multi* M, int parity) {
  const bool so = M->scheme.second_order_space != 0, visc = M->scheme.viscous != 0;
  if (so && visc) multi_lim<true, ma::NewtonianViscousFlux, T>(M, parity);
  else if (so) multi_lim<true, ma::NoViscousFlux, T>(M, parity);
  else if (visc) multi_lim<false, ma::NewtonianViscousFlux, T>(M, parity);
  else multi_lim<false, ma::NoViscousFlux, T>(M, parity);
}

}  // namespace

namespace ma {

ma_status multi_step(Multi* M, int parity) {
  if (M->n == 0) return MA_OK;
  if (M->fp32)
    multi_t<float>(M, parity);
  else
    multi_t<double>(M, parity);
  MA_HIP(hipGetLastError());
  // The last stage wrote every case's next stage-0 primitives (stream order).
  M->any_prim0 = false;
  for (ma_solver* s : M->cases) s->v0_valid = true;
  return MA_OK;
}

void solver_set_progress(ma_solver* s, int steps) {
  s->cur ^= (steps & 1);
  s->steps_done += steps;
  s->v0_valid = true;  // the multi-case flux kernel's last stage wrote the stage-0 primitives
  if (steps > 0 && s->opt.second_order_space && !s->opt.viscous) s->grads_stale = true;
}

int solver_cur(const ma_solver* s) { return s->cur; }

}  // namespace ma
