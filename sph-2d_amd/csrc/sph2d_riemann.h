// sph2d_riemann.h -- per-pair star states of the HLLC and MUSCL solver variants,
// written once for a real type T: double in the FP64 form, float in the
// FP32-mixed form.
//
// Both variants depend on which particle of a pair is "left": the reference
// evaluates every pair once with pair_i < pair_j in local index order
// (find_particle_pair.f90:64-123) and feeds (pair_i, pair_j) as (left, right).
// Callers therefore order the two particles by local index before calling, and
// e is the unit vector from the left particle to the right one.
#pragma once
#include <hip/hip_runtime.h>

namespace sph2d {

__device__ __forceinline__ double r_pow(double a, double b) { return pow(a, b); }
__device__ __forceinline__ float r_pow(float a, float b) { return powf(a, b); }
__device__ __forceinline__ double r_sqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float r_sqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ double r_abs(double a) { return fabs(a); }
__device__ __forceinline__ float r_abs(float a) { return fabsf(a); }

// Limited MUSCL slopes, function_module.f90:34-80 (Toro's limiter with
// beta_lim = 1.3; when the limited slope overshoots half the jump, the side
// with the larger one-sided slope takes half the jump and the other none).
template <typename T>
__device__ __forceinline__ void muscl_limiter(T di, T dj, T dij, T& li, T& lj) {
  const T beta_lim = T(1.3);
  T lim = T(0);
  if (dj > T(0)) {
    const T a = beta_lim * di < dj ? beta_lim * di : dj;
    const T b = di < beta_lim * dj ? di : beta_lim * dj;
    if (a > lim) lim = a;
    if (b > lim) lim = b;
    if (lim > T(0.5) * dij && r_abs(lim) > T(0)) {
      const bool take_i = di > dj;
      li = dij * (take_i ? T(0.5) : T(0));
      lj = dij * (take_i ? T(0) : T(0.5));
    } else {
      li = lim;
      lj = lim;
    }
  } else {
    const T a = beta_lim * di > dj ? beta_lim * di : dj;
    const T b = di > beta_lim * dj ? di : beta_lim * dj;
    if (a < lim) lim = a;
    if (b < lim) lim = b;
    if (lim < T(0.5) * dij && r_abs(lim) > T(0)) {
      const bool take_i = di < dj;
      li = dij * (take_i ? T(0.5) : T(0));
      lj = dij * (take_i ? T(0) : T(0.5));
    } else {
      li = lim;
      lj = lim;
    }
  }
}

// HLLC star state, Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90:362-483:
// velocities rotated into the pair frame (normal, tangential), the acoustic
// U*, P*, wave speeds S = U -/+ c Q (Q = 1 unless P* exceeds that side's
// pressure, then the estimate from rho* = rho (P*/P)^(1/7)), the fan sampled at
// x/t = dx/dt, the chosen state rotated back. Returns v* and P*.
template <typename T>
__device__ __forceinline__ void hllc_star(T vlx, T vly, T pl, T rhol, T cl, T vrx, T vry, T pr, T rhor,
                                          T cr, T ex, T ey, T x_over_t, T& vsx, T& vsy, T& ps) {
  const T sn = ey, cs = ex;
  const T L1 = vlx * cs + vly * sn, L2 = -(vlx * sn) + vly * cs;
  const T R1 = vrx * cs + vry * sn, R2 = -(vrx * sn) + vry * cs;
  const T zsum = rhol * cl + rhor * cr;
  const T Us = (L1 * cl * rhol + R1 * cr * rhor + pl - pr) / zsum;
  const T Ps = (pl * cl * rhol + pr * cr * rhor - cl * rhol * cr * rhor * (R1 - L1)) / zsum;
  const T ex7 = T(0.142857142f);                 // REAL(4) literal of the exponent
  T QL = T(1), QR = T(1);
  if (!(Ps <= pl)) {
    const T rs = rhol * r_pow(Ps / pl, ex7);
    const T t = ((rhol / rs) * (pl - Ps)) / ((rhol - rs) * (cl * cl));
    QL = t >= T(0) ? r_sqrt(t) : T(1);
  }
  if (!(Ps <= pr)) {
    const T rs = rhor * r_pow(Ps / pr, ex7);
    const T t = ((rhor / rs) * (pr - Ps)) / ((rhor - rs) * (cr * cr));
    QR = t >= T(0) ? r_sqrt(t) : T(1);
  }
  const T SL = L1 - cl * QL;
  const T SR = R1 + cr * QR;
  T H1, H2;
  if (SL > x_over_t) {
    H1 = L1; H2 = L2;
  } else if (Us >= x_over_t && x_over_t >= SL) {
    const T q = rhol * ((SL - L1) / (SL - Us));
    H1 = (Us * q) / q; H2 = (L2 * q) / q;
  } else if (SR >= x_over_t && x_over_t >= Us) {
    const T q = rhor * ((SR - R1) / (SR - Us));
    H1 = (Us * q) / q; H2 = (R2 * q) / q;
  } else {
    H1 = R1; H2 = R2;
  }
  vsx = H1 * cs - H2 * sn;
  vsy = H1 * sn + H2 * cs;
  ps = Ps;
}

}  // namespace sph2d
