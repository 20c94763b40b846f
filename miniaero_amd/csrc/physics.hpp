// physics.hpp -- pointwise physics of the miniAero RK4 path as HIP policy structs.
//
// The structs keep the reference's functor interfaces so kernels are templated
// on them the way compute_face_flux<Device, second_order, Inviscid, Viscous>
// is (Flux.h:52-54):
//   RoeFlux::compute_flux(VL, VR, F, n, t, b)          Roe_Flux.h:50-54
//   NewtonianViscousFlux / NoViscousFlux (isViscous)    Viscous_Flux.h:39-99
//   VenkatLimiter::limit(dumax, dumin, du, deltax3)     VenkatLimiter.h:45
//   VanAlbadaLimiter::limit(dumax, dumin, du)           VanAlbadaLimiter.h:45
//   ExtrapolateBC, TangentBC, InflowBC, NoSlipBC        Extrapolate_BC.h:45-102,
//                                                       Tangent_BC.h:47-123,
//                                                       Inflow_BC.h:45-111,
//                                                       NoSlip_BC.h:47-157
//
// Everything is templated on the arithmetic type T.  With T = double every
// expression keeps the reference's operation order (left-to-right
// evaluation, the same association of sums and products, literals exact in
// double); built with -ffp-contract=off and IEEE division/sqrt, the FP64
// results are the same bits a serial CPU run of the reference produces.
// T = float is the mixed-precision form: the same formulas in FP32, used for
// flux, gradient and limiter arithmetic while state and sums stay FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

#define MA_HD __host__ __device__ __forceinline__

namespace ma {

constexpr double kRgas = 287.05;  // GasModel.h:39

// ComputePrimitives, GasModel.h:69-90.  (gamma - 1.0) is evaluated at run
// time in the reference; in double it is 0.3999999999999999, not 0.4.
template <class T>
MA_HD void primitives(const T* U, T* V) {
  T gamma = T(1.4);
  T r = U[0];
  T ri = T(1.0) / r;
  T u = U[1] * ri;
  T v = U[2] * ri;
  T w = U[3] * ri;
  T k = T(0.5) * (u * u + v * v + w * w);
  T e = U[4] * ri - k;
  T temp = e * (gamma - T(1.0)) / T(kRgas);
  V[0] = r;
  V[1] = u;
  V[2] = v;
  V[3] = w;
  V[4] = temp;
}

template <class T>
MA_HD T pressure(const T* V) { return V[0] * T(kRgas) * V[4]; }  // GasModel.h:37-44
template <class T>
MA_HD T enthalpy(const T* V) { return T(1004.0) * V[4]; }        // GasModel.h:62-67

// std::min / std::max semantics (NaN in the second argument is ignored).
template <class T>
MA_HD T smin(T a, T b) { return (b < a) ? b : a; }
template <class T>
MA_HD T smax(T a, T b) { return (a < b) ? b : a; }

// Geometry-only part of the Roe flux: the area vector, its norm and the unit
// normal, tangent and binormal (Roe_Flux.h:100-118).  It depends on the face
// alone, so the solver computes it once per face at setup, with the same
// operations, instead of once per flux evaluation.
// A face as the flux policies see it: the Faces<Device> vectors as stored
// (area-weighted normal n, tangent t, binormal b; Faces.h:42-71) and the
// unit vectors Roe_Flux.h:100-118 makes of them.  Unused members cost
// nothing: the kernels build the frame inline.
template <class T>
struct FaceFrameT {
  T n[3];   // area-weighted normal
  T t[3];   // tangent as stored
  T b[3];   // binormal as stored (n x t, length |n|)
  T nn;     // |n|
  T nu[3], tu[3], bu[3];
};
using FaceFrame = FaceFrameT<double>;

template <class T>
MA_HD FaceFrameT<T> make_frame_t(const T* fn, const T* ft, const T* fb) {
  FaceFrameT<T> f;
  const T nn = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const T tn = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
  const T bn = sqrt(fb[0] * fb[0] + fb[1] * fb[1] + fb[2] * fb[2]);
  for (int d = 0; d < 3; ++d) {
    f.n[d] = fn[d];
    f.t[d] = ft[d];
    f.b[d] = fb[d];
    f.nu[d] = fn[d] / nn;
    f.tu[d] = ft[d] / tn;
    f.bu[d] = fb[d] / bn;
  }
  f.nn = nn;
  return f;
}

MA_HD FaceFrame make_frame(const double* fn, const double* ft, const double* fb) {
  return make_frame_t<double>(fn, ft, fb);
}

// The tangent and binormal the reference attaches to a face are functions of
// its area vector a alone (Face.C:80-96): i1 = first index of the largest
// |a_i|, t[i1+1] = 0, t[i1] = a[i1+2]/den, t[i1+2] = -a[i1]/den with
// den = sqrt(a[i1]^2 + a[i1+2]^2) (indices mod 3), b = a x t.  When a mesh's
// t and b are exactly these, the frame can be rebuilt from a on the fly with
// the same operations, which is cheaper than reading it.  In FP64 this gives
// the reference's bits; the FP32-arithmetic form builds it in FP32.
//
// The three cases share one form: with i3 = i1 + 2 (mod 3), p = a[i1] and
// q = a[i3], den = sqrt(p^2 + q^2), t[i1] = q / den, t[i3] = -p / den.  The
// operands are selected first and the square root and the two divisions
// computed once (a face-parallel wavefront mixes orientations, and a
// branch per case would run all three).
template <class T>
MA_HD void tangent_binormal_t(const T* a, T* t, T* b) {
  const T m0 = fabs(a[0]), m1 = fabs(a[1]), m2 = fabs(a[2]);
  int i1 = 0;
  if (m1 > m0) i1 = 1;
  if (m2 > (i1 == 0 ? m0 : m1)) i1 = 2;
  const T p = i1 == 0 ? a[0] : (i1 == 1 ? a[1] : a[2]);
  const T q = i1 == 0 ? a[2] : (i1 == 1 ? a[0] : a[1]);
  const T den = sqrt(p * p + q * q);
  const T tp = q / den, tq = -p / den;
  t[0] = i1 == 0 ? tp : (i1 == 1 ? tq : T(0.0));
  t[1] = i1 == 0 ? T(0.0) : (i1 == 1 ? tp : tq);
  t[2] = i1 == 0 ? tq : (i1 == 1 ? T(0.0) : tp);
  b[0] = a[1] * t[2] - t[1] * a[2];
  b[1] = -a[0] * t[2] + t[0] * a[2];
  b[2] = a[0] * t[1] - t[0] * a[1];
}

MA_HD void tangent_binormal(const double* a, double* t, double* b) { tangent_binormal_t<double>(a, t, b); }

// An area vector along an axis -- two components +-0, the third of magnitude
// m in [2^-500, 2^500] (FP32: [2^-60, 2^60]) -- makes every rounding of the
// general path exact: sqrt(RN(x^2)) = |x| in binary floating point with
// rounding to nearest (no overflow or underflow), x / |x| = +-1, +-0 / y =
// +-0 for y > 0, and t has unit length.  Its frame is therefore built with
// selects (axis_frame_t), the same bits as tangent_binormal_t + make_frame_t,
// signed zeros included (b keeps the general formula).  Box meshes have only
// such faces.
template <class T>
MA_HD bool is_axis_area(const T* a) {
  const bool z0 = a[0] == T(0), z1 = a[1] == T(0), z2 = a[2] == T(0);
  const T m = fabs(!z0 ? a[0] : (!z1 ? a[1] : a[2]));
  const T lo = T(sizeof(T) == 8 ? 0x1p-500 : 0x1p-60), hi = T(sizeof(T) == 8 ? 0x1p500 : 0x1p60);
  return int(z0) + int(z1) + int(z2) == 2 && m >= lo && m <= hi;
}

template <class T>
MA_HD FaceFrameT<T> axis_frame_t(const T* a) {
  const bool z0 = a[0] == T(0), z1 = a[1] == T(0);
  const int i1 = !z0 ? 0 : (!z1 ? 1 : 2);  // tangent_binormal_t's i1: the largest |a_i|
  const T p = i1 == 0 ? a[0] : (i1 == 1 ? a[1] : a[2]);
  const T q = i1 == 0 ? a[2] : (i1 == 1 ? a[0] : a[1]);  // a[i1 + 2], +-0: q / |p| = q
  const T tq = copysign(T(1), -p);                       // -p / |p|
  FaceFrameT<T> f;
  f.t[0] = i1 == 0 ? q : (i1 == 1 ? tq : T(0.0));
  f.t[1] = i1 == 0 ? T(0.0) : (i1 == 1 ? q : tq);
  f.t[2] = i1 == 0 ? tq : (i1 == 1 ? T(0.0) : q);
  f.b[0] = a[1] * f.t[2] - f.t[1] * a[2];
  f.b[1] = -a[0] * f.t[2] + f.t[0] * a[2];
  f.b[2] = a[0] * f.t[1] - f.t[0] * a[1];
  for (int d = 0; d < 3; ++d) {
    f.n[d] = a[d];
    f.nu[d] = a[d] == T(0) ? a[d] : copysign(T(1), a[d]);      // a / |a|
    f.tu[d] = f.t[d];                                           // t / 1
    f.bu[d] = f.b[d] == T(0) ? f.b[d] : copysign(T(1), f.b[d]);  // b / |b|, |b| = |p|
  }
  f.nn = fabs(p);
  return f;
}

template <class T>
MA_HD FaceFrameT<T> frame_from_area_t(const T* a) {
  if (is_axis_area<T>(a)) return axis_frame_t<T>(a);
  T t[3], b[3];
  tangent_binormal_t<T>(a, t, b);
  return make_frame_t<T>(a, t, b);
}

MA_HD FaceFrame frame_from_area(const double* a) { return frame_from_area_t<double>(a); }

// ---- inviscid flux policies ----
// The reference templates compute_face_flux and every boundary functor on an
// inviscid policy with
//   void compute_flux(const double* V_L, const double* V_R, double* F,
//                     const double* n, const double* t, const double* b) const
// (Roe_Flux.h:50-54): primitives of both sides, the face's stored normal,
// tangent and binormal.  The kernels here take any policy of that shape,
// templated on the arithmetic type.  A policy may also offer
//   void flux_frame(const T* V_L, const T* V_R, T* F, const FaceFrameT<T>&) const
// which receives the unit frame already built (RoeFlux does: the kernels
// rebuild the frame once per face and the policy would otherwise normalise
// it again); inviscid_flux() calls whichever the policy has.
template <class Inv, class T, class = void>
struct takes_frame : std::false_type {};
template <class Inv, class T>
struct takes_frame<Inv, T,
                   decltype(std::declval<const Inv&>().flux_frame(
                                std::declval<const T*>(), std::declval<const T*>(), std::declval<T*>(),
                                std::declval<const FaceFrameT<T>&>()),
                            void())> : std::true_type {};

template <class Inv, class T>
MA_HD void inviscid_flux(const Inv& inv, const T* vl, const T* vr, T* F, const FaceFrameT<T>& fr) {
  if constexpr (takes_frame<Inv, T>::value)
    inv.flux_frame(vl, vr, F, fr);
  else
    inv.compute_flux(vl, vr, F, fr.n, fr.t, fr.b);
}

struct RoeFlux {
  // Roe_Flux.h:50-54 signature: builds the face frame (Roe_Flux.h:100-118),
  // then the flux; the same bits as flux_frame on make_frame_t(n, t, b).
  template <class T>
  MA_HD void compute_flux(const T* pl, const T* pr, T* flux, const T* fn, const T* ft, const T* fb) const {
    flux_frame(pl, pr, flux, make_frame_t<T>(fn, ft, fb));
  }

  // Roe flux with the eigenvalue fix of Roe_Flux.h:157-186.  The 5x5 left and
  // right eigenvector products are unrolled, each row summed in column order
  // from zero as MathTools::MatVec5 does (MathToolsDevice.h:62-70).
  template <class T>
  MA_HD void flux_frame(const T* pl, const T* pr, T* flux, const FaceFrameT<T>& fr) const {
    const T* fn = fr.n;
    const T efix = T(0.1);
    const T gm1 = T(0.4);
    const T half = T(0.5);
    const T rl = pl[0], ul = pl[1], vl = pl[2], wl = pl[3];
    const T p_l = pressure(pl), h_l = enthalpy(pl);
    const T ht_l = h_l + half * (ul * ul + vl * vl + wl * wl);
    const T mf_l = rl * (fn[0] * ul + fn[1] * vl + fn[2] * wl);
    const T rr = pr[0], ur = pr[1], vr = pr[2], wr = pr[3];
    const T p_r = pressure(pr), h_r = enthalpy(pr);
    const T ht_r = h_r + half * (ur * ur + vr * vr + wr * wr);
    const T mf_r = rr * (fn[0] * ur + fn[1] * vr + fn[2] * wr);
    const T psum = p_l + p_r;
    T f0 = half * (mf_l + mf_r);
    T f1 = half * (mf_l * ul + mf_r * ur + fn[0] * psum);
    T f2 = half * (mf_l * vl + mf_r * vr + fn[1] * psum);
    T f3 = half * (mf_l * wl + mf_r * wr + fn[2] * psum);
    T f4 = half * (mf_l * ht_l + mf_r * ht_r);

    const T nn = fr.nn;
    const T n0 = fr.nu[0], n1 = fr.nu[1], n2 = fr.nu[2];
    const T t0 = fr.tu[0], t1 = fr.tu[1], t2 = fr.tu[2];
    const T b0 = fr.bu[0], b1 = fr.bu[1], b2 = fr.bu[2];

    const T srl = sqrt(rl);
    const T denom = T(1.0) / (srl + sqrt(rr));
    const T alpha = srl * denom;
    const T beta = T(1.0) - alpha;
    const T u = alpha * ul + beta * ur;
    const T v = alpha * vl + beta * vr;
    const T w = alpha * wl + beta * wr;
    const T h = alpha * h_l + beta * h_r +
                half * alpha * beta *
                    ((ur - ul) * (ur - ul) + (vr - vl) * (vr - vl) + (wr - wl) * (wr - wl));
    const T c = sqrt(gm1 * h);
    const T vn = u * n0 + v * n1 + w * n2;
    const T vt = u * t0 + v * t1 + w * t2;
    const T vb = u * b0 + v * b1 + w * b2;
    const T ke = half * (u * u + v * v + w * w);
    const T ci2 = T(1.0) / (c * c);
    const T hci2 = half * ci2;

    const T dq0 = rr - rl;
    const T dq1 = rr * ur - rl * ul;
    const T dq2 = rr * vr - rl * vl;
    const T dq3 = rr * wr - rl * wl;
    const T dq4 = (rr * ht_r - p_r) - (rl * ht_l - p_l);

    const T cbar = c * nn;
    const T ubar = u * fn[0] + v * fn[1] + w * fn[2];
    const T cfl = fabs(ubar) + cbar;
    const T eig1 = ubar + cbar, eig2 = ubar - cbar, eig3 = ubar;
    T a1 = fabs(eig1), a2 = fabs(eig2), a3 = fabs(eig3);
    const T ep = efix * cfl;
    if (a1 < ep) a1 = half * (eig1 * eig1 + ep * ep) / ep;
    if (a2 < ep) a2 = half * (eig2 * eig2 + ep * ep) / ep;
    if (a3 < ep) a3 = half * (eig3 * eig3 + ep * ep) / ep;
    // |lambda| as (eig+)-(eig-) with eig+- = (eig +- |eig|)/2, as the reference forms it
    const T s1 = half * (eig1 + a1) - half * (eig1 - a1);
    const T s2 = half * (eig2 + a2) - half * (eig2 - a2);
    const T s3 = half * (eig3 + a3) - half * (eig3 - a3);

    const T kh = ke - h;
    const T zero = T(0.0), one = T(1.0);
    T x;
    // left eigenvectors times the conserved jump, scaled by |lambda|
    x = zero;
    x += (gm1 * kh + c * (c - vn)) * dq0;
    x += (c * n0 - gm1 * u) * dq1;
    x += (c * n1 - gm1 * v) * dq2;
    x += (c * n2 - gm1 * w) * dq3;
    x += gm1 * dq4;
    const T l0 = s1 * x;
    x = zero;
    x += (gm1 * kh + c * (c + vn)) * dq0;
    x += (-c * n0 - gm1 * u) * dq1;
    x += (-c * n1 - gm1 * v) * dq2;
    x += (-c * n2 - gm1 * w) * dq3;
    x += gm1 * dq4;
    const T l1 = s2 * x;
    x = zero;
    x += kh * dq0;
    x += -u * dq1;
    x += -v * dq2;
    x += -w * dq3;
    x += one * dq4;
    const T l2 = s3 * x;
    x = zero;
    x += -vt * dq0;
    x += t0 * dq1;
    x += t1 * dq2;
    x += t2 * dq3;
    x += zero * dq4;
    const T l3 = s3 * x;
    x = zero;
    x += -vb * dq0;
    x += b0 * dq1;
    x += b1 * dq2;
    x += b2 * dq3;
    x += zero * dq4;
    const T l4 = s3 * x;

    // right eigenvectors
    const T hk = h + ke;
    x = zero;
    x += hci2 * l0;
    x += hci2 * l1;
    x += -gm1 * ci2 * l2;
    x += zero * l3;
    x += zero * l4;
    f0 -= half * x;
    x = zero;
    x += (u + n0 * c) * hci2 * l0;
    x += (u - n0 * c) * hci2 * l1;
    x += -gm1 * u * ci2 * l2;
    x += t0 * l3;
    x += b0 * l4;
    f1 -= half * x;
    x = zero;
    x += (v + n1 * c) * hci2 * l0;
    x += (v - n1 * c) * hci2 * l1;
    x += -gm1 * v * ci2 * l2;
    x += t1 * l3;
    x += b1 * l4;
    f2 -= half * x;
    x = zero;
    x += (w + n2 * c) * hci2 * l0;
    x += (w - n2 * c) * hci2 * l1;
    x += -gm1 * w * ci2 * l2;
    x += t2 * l3;
    x += b2 * l4;
    f3 -= half * x;
    x = zero;
    x += (hk + c * vn) * hci2 * l0;
    x += (hk - c * vn) * hci2 * l1;
    x += (c * c - gm1 * hk) * ci2 * l2;
    x += vt * l3;
    x += vb * l4;
    f4 -= half * x;
    flux[0] = f0;
    flux[1] = f1;
    flux[2] = f2;
    flux[3] = f3;
    flux[4] = f4;
  }
};

// Thermal conductivity mu * Cp / Pr, Cp = 1006 (GasModel.h:100-106).
struct NewtonianViscousFlux {
  enum { isViscous = true };
  double prandtl;
  template <class T>
  MA_HD void compute_flux(const T g[5][3], const T* V, const T* a, T* vflux) const {
    // Viscous_Flux.h:66-98, Sutherland viscosity GasModel.h:92-98.
    const T temp = V[4];
    const T mu = T(1.458e-6) * temp * sqrt(temp) / (temp + T(110.4));
    const T kc = mu * T(1006.0) / T(prandtl);
    T div = 0;
    for (int i = 0; i < 5; ++i) vflux[i] = T(0.0);
    for (int i = 0; i < 3; ++i) div += g[i + 1][i];
    // div * delta / 3 with the Kronecker delta: div / 3 on the diagonal; off
    // it (div * 0) / 3, which is div * 0 itself (+-0 / 3 = +-0, NaN / 3 =
    // NaN), so the division is left out there
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const T S = T(0.5) * (g[i + 1][j] + g[j + 1][i]);
        const T tij = S - (i == j ? div / T(3.) : div * T(0));
        vflux[1 + i] += (T(2) * mu * tij) * a[j];
        vflux[4] += (T(2) * mu * tij) * V[i + 1] * a[j];
      }
      vflux[4] += kc * g[4][i] * a[i];
    }
  }
};

struct NoViscousFlux {
  enum { isViscous = false };
  double prandtl;
  template <class T>
  MA_HD void compute_flux(const T[5][3], const T*, const T*, T*) const {}
};

// ---- boundary policies ----
// The reference applies each boundary group with a functor
// compute_<kind>BC_flux<Device, Inviscid[, Viscous]> whose operator()(i)
// evaluates face i (Extrapolate_BC.h:60-101, Tangent_BC.h:62-122,
// Inflow_BC.h:61-110, NoSlip_BC.h:65-156).  Here each kind is a policy whose
// operator() evaluates one face from what the kernel loaded for it -- the
// primitives V of the cell inside, the face frame, the conserved inflow
// state (Inflow) and 1/|x_f - x_c| (NoSlip) -- with the inviscid policy and
// the wall's viscous policy as arguments, and writes the face's contribution
// to its cell's flux sum (the operand of the reference's atomic_add).

// Tangent_BC.h:86-102 / NoSlip_BC.h:96-112: the mirror state, normal
// velocity reflected; |n| of the frame is the reference's area_norm bit for bit.
template <class T>
MA_HD void mirror_state(const T* vc, const FaceFrameT<T>& fr, T* vr) {
  const T* n = fr.n;
  const T an = fr.nn;
  T ub = 0.0;
  ub += vc[1] * n[0] / an;
  ub += vc[2] * n[1] / an;
  ub += vc[3] * n[2] / an;
  vr[0] = vc[0];
  vr[1] = vc[1] - T(2) * ub * n[0] / an;
  vr[2] = vc[2] - T(2) * ub * n[1] / an;
  vr[3] = vc[3] - T(2) * ub * n[2] / an;
  vr[4] = vc[4];
}

struct ExtrapolateBC {  // compute_extrapolateBC_flux: the cell's own state on both sides
  template <class T, class Inv, class Wall>
  MA_HD void operator()(const T* vc, const FaceFrameT<T>& fr, const double*, T, const Inv& inv, const Wall&,
                        T* out) const {
    T F[5];
    inviscid_flux(inv, vc, vc, F, fr);
    for (int k = 0; k < 5; ++k) out[k] = -F[k];
  }
};

struct TangentBC {  // compute_tangentBC_flux: slip wall, mirror state outside
  template <class T, class Inv, class Wall>
  MA_HD void operator()(const T* vc, const FaceFrameT<T>& fr, const double*, T, const Inv& inv, const Wall&,
                        T* out) const {
    T vr[5], F[5];
    mirror_state(vc, fr, vr);
    inviscid_flux(inv, vc, vr, F, fr);
    for (int k = 0; k < 5; ++k) out[k] = -F[k];
  }
};

struct InflowBC {  // compute_inflowBC_flux: the inflow state outside
  template <class T, class Inv, class Wall>
  MA_HD void operator()(const T* vc, const FaceFrameT<T>& fr, const double* state, T, const Inv& inv,
                        const Wall&, T* out) const {
    double vi[5];
    primitives(state, vi);
    T vr[5], F[5];
    for (int k = 0; k < 5; ++k) vr[k] = T(vi[k]);
    inviscid_flux(inv, vc, vr, F, fr);
    for (int k = 0; k < 5; ++k) out[k] = -F[k];
  }
};

// compute_NoSlipBC_flux: the mirror state's inviscid flux, plus the viscous
// flux of the one-sided gradient from the cell to the wall state
// (rho, 0, 0, 0, T) over the centroid-to-face distance (NoSlip_BC.h:114-147).
struct NoSlipBC {
  template <class T, class Inv, class Wall>
  MA_HD void operator()(const T* vc, const FaceFrameT<T>& fr, const double*, T wall_inv, const Inv& inv,
                        const Wall& wallv, T* out) const {
    T vr[5], F[5];
    mirror_state(vc, fr, vr);
    inviscid_flux(inv, vc, vr, F, fr);
    const T vf[5] = {vc[0], T(0.0), T(0.0), T(0.0), vc[4]};
    T gf[5][3], fv[5];
    for (int k = 0; k < 5; ++k)
      for (int d = 0; d < 3; ++d) gf[k][d] = (vf[k] - vc[k]) * fr.nu[d] * wall_inv;
    wallv.compute_flux(gf, vf, fr.n, fv);
    for (int k = 0; k < 5; ++k) out[k] = -F[k] + fv[k];
  }
};

// The boundary policy of each kind the mesh's groups name (MA_BC_*).  A
// caller's own boundary treatment is a struct of the same shape in place of
// one of these.
struct ReferenceBoundaries {
  using Extrapolate = ExtrapolateBC;
  using Tangent = TangentBC;
  using NoSlip = NoSlipBC;
  using Inflow = InflowBC;
};

struct VenkatLimiter {
  enum { kind = 0 };
  // VenkatLimiter.h:45-73; deltax3 is the squared face distance
  // (StencilLimiter.h:426-455).  The reference's two branches apply one
  // formula to dumax (du > 1e-40) or dumin (du < -1e-40); the bound is
  // selected first and the formula evaluated once, with the same operations,
  // so a wavefront whose lanes take both branches does one division, not two.
  template <class T>
  static MA_HD T limit(T dumax, T dumin, T du, T deltax3) {
    T beta = 1;
    T eps2 = deltax3 * beta * beta * beta;
    const T d = du > T(0) ? dumax : dumin;
    T num = (d * d + eps2) * du + T(2) * du * du * d;
    T den = du * (d * d + T(2) * du * du + d * du + eps2);
    const T phi = num / den;
    return (du > T(1e-40) || du < -T(1e-40)) ? phi : T(1);
  }
};

struct VanAlbadaLimiter {
  enum { kind = 1 };
  // VanAlbadaLimiter.h:45-65 (defined but never called by the reference).
  // As in VenkatLimiter, the bound is selected before the one division.
  template <class T>
  static MA_HD T limit(T dumax, T dumin, T du, T /*deltax3*/) {
    const T eps = T(2.220446049250313080847e-16);  // DBL_EPSILON
    const T d = du > eps ? dumax : dumin;
    const T q = d / du;
    const T y = (du > eps || du < -eps) ? q : T(2);
    T phi = 1;
    if (y < T(2)) {
      phi = (T(4) * y - y * y) / (y * y - T(4) * y + T(8));
      phi = smax(phi, T(0.0));
      phi = smin(phi, T(1.0));
    }
    return phi;
  }
};

}  // namespace ma
