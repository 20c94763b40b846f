// physics.hpp -- pointwise physics of the miniAero RK4 path as HIP policy structs.
//
// The structs keep the reference's functor interfaces so kernels are templated
// on them the way compute_face_flux<Device, second_order, Inviscid, Viscous>
// is (Flux.h:52-54):
//   RoeFlux::compute_flux(VL, VR, F, n, t, b)          Roe_Flux.h:50-54
//   NewtonianViscousFlux / NoViscousFlux (isViscous)    Viscous_Flux.h:39-99
//   VenkatLimiter::limit(dumax, dumin, du, deltax3)     VenkatLimiter.h:45
//   VanAlbadaLimiter::limit(dumax, dumin, du)           VanAlbadaLimiter.h:45
//
// Every expression keeps the reference's operation order (left-to-right
// evaluation, the same association of sums and products).  Built with
// -ffp-contract=off and IEEE division/sqrt, the FP64 results are then the same
// bits a serial CPU run of the reference produces, which is what lets the
// parity tests demand bit equality with the oracle.
#pragma once

#include <hip/hip_runtime.h>

#define MA_HD __host__ __device__ __forceinline__

namespace ma {

constexpr double kRgas = 287.05;  // GasModel.h:39

// ComputePrimitives, GasModel.h:69-90.  (gamma - 1.0) is evaluated at run
// time in the reference; it is 0.3999999999999999, not 0.4.
MA_HD void primitives(const double* U, double* V) {
  double gamma = 1.4;
  double r = U[0];
  double ri = 1.0 / r;
  double u = U[1] * ri;
  double v = U[2] * ri;
  double w = U[3] * ri;
  double k = 0.5 * (u * u + v * v + w * w);
  double e = U[4] * ri - k;
  double T = e * (gamma - 1.0) / kRgas;
  V[0] = r;
  V[1] = u;
  V[2] = v;
  V[3] = w;
  V[4] = T;
}

MA_HD double pressure(const double* V) { return V[0] * kRgas * V[4]; }  // GasModel.h:37-44
MA_HD double enthalpy(const double* V) { return 1004.0 * V[4]; }        // GasModel.h:62-67

// std::min / std::max semantics (NaN in the second argument is ignored).
MA_HD double smin(double a, double b) { return (b < a) ? b : a; }
MA_HD double smax(double a, double b) { return (a < b) ? b : a; }

// Geometry-only part of the Roe flux: the area vector, its norm and the unit
// normal, tangent and binormal (Roe_Flux.h:100-118).  It depends on the face
// alone, so the solver computes it once per face at setup, with the same
// operations, instead of once per flux evaluation.
struct FaceFrame {
  double n[3];   // area-weighted normal
  double nn;     // |n|
  double nu[3], tu[3], bu[3];
};

MA_HD FaceFrame make_frame(const double* fn, const double* ft, const double* fb) {
  FaceFrame f;
  const double nn = sqrt(fn[0] * fn[0] + fn[1] * fn[1] + fn[2] * fn[2]);
  const double tn = sqrt(ft[0] * ft[0] + ft[1] * ft[1] + ft[2] * ft[2]);
  const double bn = sqrt(fb[0] * fb[0] + fb[1] * fb[1] + fb[2] * fb[2]);
  for (int d = 0; d < 3; ++d) {
    f.n[d] = fn[d];
    f.nu[d] = fn[d] / nn;
    f.tu[d] = ft[d] / tn;
    f.bu[d] = fb[d] / bn;
  }
  f.nn = nn;
  return f;
}

struct RoeFlux {
  // Roe_Flux.h:50-54 signature: builds the face frame, then the flux.
  MA_HD void compute_flux(const double* pl, const double* pr, double* flux, const double* fn,
                          const double* ft, const double* fb) const {
    const FaceFrame fr = make_frame(fn, ft, fb);
    flux_frame(pl, pr, flux, fr);
  }

  // Roe flux with the eigenvalue fix of Roe_Flux.h:157-186.  The 5x5 left and
  // right eigenvector products are unrolled, each row summed in column order
  // from zero as MathTools::MatVec5 does (MathToolsDevice.h:62-70).
  MA_HD void flux_frame(const double* pl, const double* pr, double* flux,
                        const FaceFrame& fr) const {
    const double* fn = fr.n;
    const double efix = 0.1;
    const double gm1 = 0.4;
    const double rl = pl[0], ul = pl[1], vl = pl[2], wl = pl[3];
    const double p_l = pressure(pl), h_l = enthalpy(pl);
    const double ht_l = h_l + 0.5 * (ul * ul + vl * vl + wl * wl);
    const double mf_l = rl * (fn[0] * ul + fn[1] * vl + fn[2] * wl);
    const double rr = pr[0], ur = pr[1], vr = pr[2], wr = pr[3];
    const double p_r = pressure(pr), h_r = enthalpy(pr);
    const double ht_r = h_r + 0.5 * (ur * ur + vr * vr + wr * wr);
    const double mf_r = rr * (fn[0] * ur + fn[1] * vr + fn[2] * wr);
    const double psum = p_l + p_r;
    double f0 = 0.5 * (mf_l + mf_r);
    double f1 = 0.5 * (mf_l * ul + mf_r * ur + fn[0] * psum);
    double f2 = 0.5 * (mf_l * vl + mf_r * vr + fn[1] * psum);
    double f3 = 0.5 * (mf_l * wl + mf_r * wr + fn[2] * psum);
    double f4 = 0.5 * (mf_l * ht_l + mf_r * ht_r);

    const double nn = fr.nn;
    const double n0 = fr.nu[0], n1 = fr.nu[1], n2 = fr.nu[2];
    const double t0 = fr.tu[0], t1 = fr.tu[1], t2 = fr.tu[2];
    const double b0 = fr.bu[0], b1 = fr.bu[1], b2 = fr.bu[2];

    const double srl = sqrt(rl);
    const double denom = 1.0 / (srl + sqrt(rr));
    const double alpha = srl * denom;
    const double beta = 1.0 - alpha;
    const double u = alpha * ul + beta * ur;
    const double v = alpha * vl + beta * vr;
    const double w = alpha * wl + beta * wr;
    const double h = alpha * h_l + beta * h_r +
                     0.5 * alpha * beta *
                         ((ur - ul) * (ur - ul) + (vr - vl) * (vr - vl) + (wr - wl) * (wr - wl));
    const double c = sqrt(gm1 * h);
    const double vn = u * n0 + v * n1 + w * n2;
    const double vt = u * t0 + v * t1 + w * t2;
    const double vb = u * b0 + v * b1 + w * b2;
    const double ke = 0.5 * (u * u + v * v + w * w);
    const double ci2 = 1.0 / (c * c);
    const double hci2 = 0.5 * ci2;

    const double dq0 = rr - rl;
    const double dq1 = rr * ur - rl * ul;
    const double dq2 = rr * vr - rl * vl;
    const double dq3 = rr * wr - rl * wl;
    const double dq4 = (rr * ht_r - p_r) - (rl * ht_l - p_l);

    const double cbar = c * nn;
    const double ubar = u * fn[0] + v * fn[1] + w * fn[2];
    const double cfl = fabs(ubar) + cbar;
    const double eig1 = ubar + cbar, eig2 = ubar - cbar, eig3 = ubar;
    double a1 = fabs(eig1), a2 = fabs(eig2), a3 = fabs(eig3);
    const double ep = efix * cfl;
    if (a1 < ep) a1 = 0.5 * (eig1 * eig1 + ep * ep) / ep;
    if (a2 < ep) a2 = 0.5 * (eig2 * eig2 + ep * ep) / ep;
    if (a3 < ep) a3 = 0.5 * (eig3 * eig3 + ep * ep) / ep;
    // |lambda| as (eig+)-(eig-) with eig+- = (eig +- |eig|)/2, as the reference forms it
    const double s1 = 0.5 * (eig1 + a1) - 0.5 * (eig1 - a1);
    const double s2 = 0.5 * (eig2 + a2) - 0.5 * (eig2 - a2);
    const double s3 = 0.5 * (eig3 + a3) - 0.5 * (eig3 - a3);

    const double kh = ke - h;
    double x;
    // left eigenvectors times the conserved jump, scaled by |lambda|
    x = 0.0;
    x += (gm1 * kh + c * (c - vn)) * dq0;
    x += (c * n0 - gm1 * u) * dq1;
    x += (c * n1 - gm1 * v) * dq2;
    x += (c * n2 - gm1 * w) * dq3;
    x += gm1 * dq4;
    const double l0 = s1 * x;
    x = 0.0;
    x += (gm1 * kh + c * (c + vn)) * dq0;
    x += (-c * n0 - gm1 * u) * dq1;
    x += (-c * n1 - gm1 * v) * dq2;
    x += (-c * n2 - gm1 * w) * dq3;
    x += gm1 * dq4;
    const double l1 = s2 * x;
    x = 0.0;
    x += kh * dq0;
    x += -u * dq1;
    x += -v * dq2;
    x += -w * dq3;
    x += 1.0 * dq4;
    const double l2 = s3 * x;
    x = 0.0;
    x += -vt * dq0;
    x += t0 * dq1;
    x += t1 * dq2;
    x += t2 * dq3;
    x += 0.0 * dq4;
    const double l3 = s3 * x;
    x = 0.0;
    x += -vb * dq0;
    x += b0 * dq1;
    x += b1 * dq2;
    x += b2 * dq3;
    x += 0.0 * dq4;
    const double l4 = s3 * x;

    // right eigenvectors
    const double hk = h + ke;
    x = 0.0;
    x += hci2 * l0;
    x += hci2 * l1;
    x += -gm1 * ci2 * l2;
    x += 0.0 * l3;
    x += 0.0 * l4;
    f0 -= 0.5 * x;
    x = 0.0;
    x += (u + n0 * c) * hci2 * l0;
    x += (u - n0 * c) * hci2 * l1;
    x += -gm1 * u * ci2 * l2;
    x += t0 * l3;
    x += b0 * l4;
    f1 -= 0.5 * x;
    x = 0.0;
    x += (v + n1 * c) * hci2 * l0;
    x += (v - n1 * c) * hci2 * l1;
    x += -gm1 * v * ci2 * l2;
    x += t1 * l3;
    x += b1 * l4;
    f2 -= 0.5 * x;
    x = 0.0;
    x += (w + n2 * c) * hci2 * l0;
    x += (w - n2 * c) * hci2 * l1;
    x += -gm1 * w * ci2 * l2;
    x += t2 * l3;
    x += b2 * l4;
    f3 -= 0.5 * x;
    x = 0.0;
    x += (hk + c * vn) * hci2 * l0;
    x += (hk - c * vn) * hci2 * l1;
    x += (c * c - gm1 * hk) * ci2 * l2;
    x += vt * l3;
    x += vb * l4;
    f4 -= 0.5 * x;
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
  MA_HD void compute_flux(const double g[5][3], const double* V, const double* a,
                          double* vflux) const {
    // Viscous_Flux.h:66-98, Sutherland viscosity GasModel.h:92-98.
    const double T = V[4];
    const double mu = 1.458e-6 * T * sqrt(T) / (T + 110.4);
    const double kc = mu * 1006.0 / prandtl;
    double div = 0;
    for (int i = 0; i < 5; ++i) vflux[i] = 0.0;
    for (int i = 0; i < 3; ++i) div += g[i + 1][i];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double delta = (i == j) ? 1 : 0;
        const double S = 0.5 * (g[i + 1][j] + g[j + 1][i]);
        const double tij = S - div * delta / 3.;
        vflux[1 + i] += (2 * mu * tij) * a[j];
        vflux[4] += (2 * mu * tij) * V[i + 1] * a[j];
      }
      vflux[4] += kc * g[4][i] * a[i];
    }
  }
};

struct NoViscousFlux {
  enum { isViscous = false };
  double prandtl;
  MA_HD void compute_flux(const double[5][3], const double*, const double*, double*) const {}
};

struct VenkatLimiter {
  enum { kind = 0 };
  // VenkatLimiter.h:45-73; deltax3 is the squared face distance (StencilLimiter.h:426-455).
  static MA_HD double limit(double dumax, double dumin, double du, double deltax3) {
    double beta = 1;
    double eps2 = deltax3 * beta * beta * beta;
    if (du > 1e-40) {
      double num = (dumax * dumax + eps2) * du + 2 * du * du * dumax;
      double den = du * (dumax * dumax + 2 * du * du + dumax * du + eps2);
      return num / den;
    } else if (du < -1e-40) {
      double num = (dumin * dumin + eps2) * du + 2 * du * du * dumin;
      double den = du * (dumin * dumin + 2 * du * du + dumin * du + eps2);
      return num / den;
    }
    return 1;
  }
};

struct VanAlbadaLimiter {
  enum { kind = 1 };
  // VanAlbadaLimiter.h:45-65 (defined but never called by the reference).
  static MA_HD double limit(double dumax, double dumin, double du, double /*deltax3*/) {
    const double eps = 2.220446049250313080847e-16;  // DBL_EPSILON
    double y = 2;
    if (du > eps)
      y = dumax / du;
    else if (du < -eps)
      y = dumin / du;
    double phi = 1;
    if (y < 2) {
      phi = (4 * y - y * y) / (y * y - 4 * y + 8);
      phi = smax(phi, 0.0);
      phi = smin(phi, 1.0);
    }
    return phi;
  }
};

}  // namespace ma
