// sph2d_adjoint.hip -- reverse of one FP64 force evaluation (vector-Jacobian product).
//
// The forward evaluation (single_step_compute.f90:62-171 with
// Define_freesurface_and_Normailized_derivates.f90 and
// Solver_Normalized_Modified_Linear_Riemann_Delta_SPS.f90 /
// Solver_Normalized_Linear_Riemann_Delta_SPS.f90) maps the subdomain state
// (x, y, u, v, rho of every local particle, ghosts included) to drho/dt,
// dv/dt and the XSPH average velocity. Given adjoint weights of those outputs
// per fluid particle, the kernels below return d(sum of weighted outputs) /
// d(state) per fluid particle, the ghosts' shares added to the particles they
// copy. What stays fixed, as in any discrete adjoint of this step: the
// neighbour set (the acceptance masks of the last evaluation), the free-surface
// and near-surface flags, the branches of the limiter (vx < 0, max, min), the
// particle masses and the local order (the SPS term reads particle_mass at the
// local index, Solver...SPS.f90:450-492). The identity renormalisation of the reference keeps the
// normalised kernel gradient equal to the raw one, so no matrix inverse enters.
//
// Structure (all gather form, no float atomics, fixed summation order):
//   k_adj_weights   per local particle: its output weights (zero for ghosts)
//   k_adj_force     per local s, over neighbours o: d/d(state_s, grad rho_s, tau_s)
//                   of the force-pass pair terms summed by s (weights of s) and by
//                   o (weights of o); the state directions by reverse mode over the
//                   pair formulas, grad rho and tau (linear in the pair terms) in
//                   closed form
//   k_adj_tau       per local s: recompute the strain S_s, reverse tau(S_s, rho_s)
//   k_adj_sums      per local s, over neighbours o: reverse of the pass-A sums
//                   (S and grad rho) into d/d(state_s); result stored by gid
//   k_adj_corners   corner ghosts into the x-ghost they copy
//   k_adj_owner     every ghost into its fluid particle (same slot arithmetic as
//                   the ghost emission of Distribute_Period_Boundary_Particle.f90)
// Pair-level vector-Jacobian products with respect to the state come from a
// hand-written reverse mode of the FP64 force pass's pair formulas (one primal and
// one adjoint sweep per pair and side); the per-particle closures (the SPS stress of
// k_adj_tau, the walls) use forward-mode duals; the reverse sweep over the particle
// graph is explicit.
#include "sph2d_internal.h"
#include "sph2d_device.h"

namespace sph2d {

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

// Forward-mode dual number with N tangent directions.
template <int N>
struct Dual {
  double v;
  double d[N];
  __device__ __forceinline__ Dual() {}
  __device__ __forceinline__ Dual(double x) : v(x) {
#pragma unroll
    for (int k = 0; k < N; ++k) d[k] = 0.0;
  }
  friend __device__ __forceinline__ Dual operator+(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator-(const Dual& a) {
    Dual r;
    r.v = -a.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator*(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
  }
  friend __device__ __forceinline__ Dual operator/(const Dual& a, const Dual& b) {
    Dual r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
#pragma unroll
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
  }
};

__device__ __forceinline__ double val(double x) { return x; }
template <int N>
__device__ __forceinline__ double val(const Dual<N>& a) { return a.v; }
__device__ __forceinline__ double tsqrt(double x) { return sqrt(x); }
template <int N>
__device__ __forceinline__ Dual<N> tsqrt(const Dual<N>& a) {
  Dual<N> r;
  r.v = sqrt(a.v);
  const double g = r.v > 0.0 ? 0.5 / r.v : 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * g;
  return r;
}
__device__ __forceinline__ double texp(double x) { return exp(x); }
template <int N>
__device__ __forceinline__ Dual<N> texp(const Dual<N>& a) {
  Dual<N> r;
  r.v = exp(a.v);
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * r.v;
  return r;
}
__device__ __forceinline__ double tpow(double x, double e) { return pow(x, e); }
template <int N>
__device__ __forceinline__ Dual<N> tpow(const Dual<N>& a, double e) {
  Dual<N> r;
  r.v = pow(a.v, e);
  const double g = a.v != 0.0 ? e * r.v / a.v : 0.0;
#pragma unroll
  for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * g;
  return r;
}

// Kinds of particle as the evaluation sees them. FLUID (and periodic ghosts):
// the state is (x, y, u, v, rho), p and c from the Tait law, the mass fixed.
// TIED: wavemaker particles below the water height: the same, with mass rho V0
// (Wavemaker_Subroutine.f90:171-193 sets both from one rho). WALL: fixed ghosts
// whose (u, v, p) come from their neighbours (Assign_boundary_particles.f90:157-299):
// rho from the inverse Tait law with the REAL(4) exponent, c = c0 (rho/rho0)^3,
// mass rho V0. INERT: zero mass (walls with no support, wavemaker particles above
// the water), which adds nothing to any sum.
enum { KIND_FLUID = 0, KIND_TIED = 1, KIND_WALL = 2, KIND_INERT = 3 };

__device__ __forceinline__ int particle_kind(const Buffers& B, int s) {
  const int t = B.l_type[s];
  if (t == -3) return B.l_mass[s] != 0.0 ? KIND_WALL : KIND_INERT;
  if (t == 130) return B.l_mass[s] != 0.0 ? KIND_TIED : KIND_INERT;
  return KIND_FLUID;
}

// One side of a pair: the primary state (x, y, u, v, q with q = rho, or p for a
// wall), what follows from it (rho, p, c, mass), grad rho and the SPS stress.
template <class T>
struct Side {
  T x, y, u, v, q, rho, p, c, m, gx, gy, txx, txy, tyy;
};

// Adjoint weights of the summing particle's outputs: drho/dt (solver part),
// the delta-SPH correction (fluid only), acceleration (fluid only), XSPH, the
// shift sums (Shift_particle_position.f90:78-103) and grad rho.
struct PairLam {
  double d, c, ax, ay, gx, gy, shx, shy;
};
constexpr int kLam = 10;   // per local particle: d, c, ax, ay, gx, gy, shx, shy, grad rho x, y

// status_equation.f90:52-61 (Tait), as in the forward evaluation.
template <class T>
__device__ __forceinline__ void tait_t(const DevParams& P, const T& rho, T& p, T& c) {
  const T x = rho * (1.0 / P.rho0);
  const T x3 = x * x * x;
  p = P.eos_b * (x3 * x3 * x - 1.0);
  c = P.sqrt_c0sq * x3;
}

// rho, p, c and the mass from the primary state, per kind (k_assign_walls and the
// wavemaker placement compute them this way).
template <class T>
__device__ __forceinline__ void derive_side(const DevParams& P, Side<T>& a, int kind, double mass) {
  if (kind == KIND_WALL) {
    a.p = a.q;
    a.rho = P.rho0 * tpow(1.0 + 7.0 * a.p / (P.c0sq * P.rho0), (double)0.142857142f);
    const T x = a.rho * (1.0 / P.rho0);
    a.c = P.sqrt_c0sq * (x * x * x);
    a.m = a.rho * P.vol0;
  } else {
    a.rho = a.q;
    tait_t(P, a.rho, a.p, a.c);
    a.m = kind == KIND_TIED ? a.rho * P.vol0 : T(mass);
  }
}

// The primary state of local particle s as plain values.
__device__ __forceinline__ Side<double> load_side(const DevParams& P, const Buffers& B, int s, int& kind) {
  Side<double> a;
  kind = particle_kind(B, s);
  const double2 p = B.l_pos[s], v = B.l_vel[s];
  const double2 g = B.grad[s];
  a.x = p.x; a.y = p.y; a.u = v.x; a.v = v.y;
  a.q = kind == KIND_WALL ? B.l_pc[s].x : B.l_rv[s].x;
  a.gx = g.x; a.gy = g.y;
  a.txx = B.tau[3 * s]; a.txy = B.tau[3 * s + 1]; a.tyy = B.tau[3 * s + 2];
  derive_side(P, a, kind, B.l_mass[s]);
  return a;
}

// Reverse mode of the weighted force-pass pair terms of A summing over neighbour B
// (Solver_Normalized_*_Riemann_Delta_SPS.f90:300-583 as the FP64 force pass computes
// them, with the shift sum of Shift_particle_position.f90:78-103): the forward
// formulas, then their adjoints in reverse order. bA, bB receive d(J)/d(x, y, u, v, rho, p, c, m) of either side
// (what a side's primary state q moves is chained by side_q_bar), gaux d(J)/d(grad
// rho x, y, tau xx, xy, yy) of either side (linear, through the sums A + B). The branches of the forward (the limiter's cases, vx < 0, the
// near-surface flag) are taken as the forward takes them. One primal sweep and one
// adjoint sweep in FP64 registers: a fraction of the work and registers of 5-direction
// duals over the same formulas.
enum { RB_X = 0, RB_Y, RB_U, RB_V, RB_RHO, RB_P, RB_C, RB_M, RB_N };

template <int SOLVER>
__device__ __forceinline__ void force_pair_rev(const DevParams& P, const Side<double>& A, const Side<double>& B,
                                               double gmo, int nes, const PairLam& L, int do_shift, double* bA,
                                               double* bB, double* gaux) {
#pragma unroll
  for (int k = 0; k < RB_N; ++k) { bA[k] = 0.0; bB[k] = 0.0; }
  // forward
  const double dx = A.x - B.x, dy = A.y - B.y;
  const double r2 = dx * dx + dy * dy;
  const double r = sqrt(r2);
  const double inv_r = 1.0 / r;
  const double q = r * P.inv_h;
  const double eq = exp(-(q * q));
  const double w = (eq - P.e9) * P.kfactor;
  const double kq = (inv_r * P.inv_h) * P.kfactor;          // kt = eq (-2 q) kq
  const double kt = eq * (-2.0 * q) * kq;
  const double dwx = kt * dx, dwy = kt * dy;
  const double mo = B.m;
  const double Vo = mo / B.rho;
  const double ex = -(dx * inv_r), ey = -(dy * inv_r);
  const double UL = A.u * ex + A.v * ey;
  const double UR = B.u * ex + B.v * ey;
  const double dU = UL - UR;
  const double Pbar = (A.p + B.p) * 0.5;
  const double vbx = (A.u + B.u) * 0.5, vby = (A.v + B.v) * 0.5;
  const double cbar = (A.c + B.c) * 0.5;
  const double rbar = (A.rho + B.rho) * 0.5;
  double vsx, vsy, Ps;
  // modified linear: D = U* - U_bar; linear: zs, zo, 1 / (zs + zo), U*
  double D = 0.0, bR = 0.0, zs = 0.0, zo = 0.0, iz = 0.0, Us = 0.0, NP = 0.0, NU = 0.0;
  int bcase = 0;                                  // 0: bR = 0, 1: bR = 3 dU, 2: bR = c_bar
  if (SOLVER == SPH2D_SOLVER_LINEAR_RIEMANN) {
    zs = A.rho * A.c;
    zo = B.rho * B.c;
    iz = 1.0 / (zs + zo);
    NU = zs * UR + zo * UL + A.p - B.p;
    NP = zs * B.p + zo * A.p + zs * zo * dU;
    Us = NU * iz;
    vsx = Us * ex;
    vsy = Us * ey;
    Ps = NP * iz;
  } else {
    D = 0.5 * (A.p - B.p) / rbar * P.inv_c0;
    vsx = D * ex + vbx;
    vsy = D * ey + vby;
    if (dU > 0.0) { bR = 3.0 * dU; bcase = 1; }
    if (bR >= cbar) { bR = cbar; bcase = 2; }
    Ps = Pbar + 0.5 * bR * rbar * dU;
  }
  if (nes) { vsx = vbx; vsy = vby; Ps = Pbar; }
  const double ti = (A.u - vsx) * dwx + (A.v - vsy) * dwy;
  const double rij = 1.0 / (A.rho * B.rho);
  const double fp = mo * (-2.0 * Ps * rij);
  const double vijx = A.u - B.u, vijy = A.v - B.v;
  const double vx = vijx * dx + vijy * dy;
  const bool visc = vx < 0.0;
  const double fh2 = (P.fai * P.h) * (P.fai * P.h);
  double phi = 0.0, Pi = 0.0;
  if (visc) {
    phi = P.h * vx / (r2 + fh2);
    Pi = (-(P.alpha * cbar * phi) + P.beta * (phi * phi)) / rbar;
  }
  const double Rd = dx * dwx + dy * dwy;
  const double lf = Vo * 2.0 * P.nu * Rd / r2;
  const double k3 = P.delta_alpha * P.h * P.rho0;
  const double tmp = k3 * cbar * vx * Vo / (r2 * A.rho);
  const double sxx = A.txx + B.txx, sxy = A.txy + B.txy, syy = A.tyy + B.tyy;
  const double tp1 = sxx * dwx + sxy * dwy;
  const double tp1r = tp1 * rij;
  const double tp2p = tp1r + syy * dwy + sxy * dwx;
  const double xsphf = (SOLVER == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && P.epsilon_xsph != 0.0)
                           ? P.epsilon_xsph * P.xsph_f : 0.0;
  const double psi = (B.rho - A.rho) + 0.5 * ((A.gx + B.gx) * dx + (A.gy + B.gy) * dy);
  const double k2 = 2.0 * P.delta * P.h;
  const double cf = k2 * cbar * Rd / r2 * Vo;
  // d/d(grad rho x, y, tau xx, xy, yy): linear, through the sums A + B
  gaux[0] = -L.c * cf * 0.5 * dx;
  gaux[1] = -L.c * cf * 0.5 * dy;
  gaux[2] = gmo * rij * (L.ax * dwx + L.ay * (dwx * rij));
  gaux[3] = gmo * rij * (L.ax * dwy + L.ay * (dwy * rij + dwx));
  gaux[4] = gmo * rij * (L.ay * dwy);
  // reverse: adjoints of the intermediates
  double b_dx = 0.0, b_dy = 0.0, b_r2 = 0.0, b_w = 0.0, b_dwx = 0.0, b_dwy = 0.0, b_Vo = 0.0, b_cbar = 0.0;
  double b_rbar = 0.0, b_Rd = 0.0, b_vx = 0.0, b_vijx = 0.0, b_vijy = 0.0, b_rij = 0.0;
  if (do_shift && (L.shx != 0.0 || L.shy != 0.0)) {
    const double wr = w * P.inv_wdx;
    const double wr2 = wr * wr;
    const double sb = P.shift_beta * P.h2;
    const double f = sb * (1.0 + P.shift_R * (wr2 * wr2)) * Vo;
    const double b_f = -(L.shx * dwx + L.shy * dwy);
    b_dwx += -L.shx * f;
    b_dwy += -L.shy * f;
    b_Vo += b_f * sb * (1.0 + P.shift_R * (wr2 * wr2));
    const double b_wr2 = b_f * sb * Vo * P.shift_R * 2.0 * wr2;
    b_w += b_wr2 * 2.0 * wr * P.inv_wdx;
  }
  {                                               // delta-SPH: J -= L.c cf psi
    const double b_cf = -L.c * psi, b_psi = -L.c * cf;
    bB[RB_RHO] += b_psi;
    bA[RB_RHO] -= b_psi;
    b_dx += b_psi * 0.5 * (A.gx + B.gx);
    b_dy += b_psi * 0.5 * (A.gy + B.gy);
    b_cbar += b_cf * k2 * Rd / r2 * Vo;
    b_Rd += b_cf * k2 * cbar / r2 * Vo;
    b_r2 -= b_cf * cf / r2;
    b_Vo += b_cf * k2 * cbar * Rd / r2;
  }
  if (xsphf != 0.0) {                             // XSPH: J -= (L.gx vijx + L.gy vijy) f
    const double srho = A.rho + B.rho;
    const double f = xsphf * mo / srho * w;
    const double b_f = -(L.gx * vijx + L.gy * vijy);
    b_vijx += -L.gx * f;
    b_vijy += -L.gy * f;
    bB[RB_M] += b_f * xsphf / srho * w;
    const double b_s = -b_f * f / srho;
    bA[RB_RHO] += b_s;
    bB[RB_RHO] += b_s;
    b_w += b_f * xsphf * mo / srho;
  }
  const double b_drho = L.d, b_ax = L.ax, b_ay = L.ay;
  {                                               // SPS stress
    const double b_tp2 = b_ay * gmo;
    b_rij += b_tp2 * tp2p;
    const double b_tp2p = b_tp2 * rij;
    const double b_tp1r = b_tp2p + b_ax * gmo;
    b_dwy += b_tp2p * syy;
    b_dwx += b_tp2p * sxy;
    const double b_tp1 = b_tp1r * rij;
    b_rij += b_tp1r * tp1;
    b_dwx += b_tp1 * sxx;
    b_dwy += b_tp1 * sxy;
  }
  {                                               // delta-alpha term
    const double b_tmp = b_ax * dwx + b_ay * dwy;
    b_dwx += b_ax * tmp;
    b_dwy += b_ay * tmp;
    const double den = r2 * A.rho;
    b_cbar += b_tmp * k3 * vx * Vo / den;
    b_vx += b_tmp * k3 * cbar * Vo / den;
    b_Vo += b_tmp * k3 * cbar * vx / den;
    b_r2 -= b_tmp * tmp / r2;
    bA[RB_RHO] -= b_tmp * tmp / A.rho;
  }
  {                                               // laminar viscosity
    const double b_lf = b_ax * vijx + b_ay * vijy;
    b_vijx += b_ax * lf;
    b_vijy += b_ay * lf;
    b_Vo += b_lf * 2.0 * P.nu * Rd / r2;
    b_Rd += b_lf * Vo * 2.0 * P.nu / r2;
    b_r2 -= b_lf * lf / r2;
  }
  b_dx += b_Rd * dwx;
  b_dwx += b_Rd * dx;
  b_dy += b_Rd * dwy;
  b_dwy += b_Rd * dy;
  if (visc) {                                     // artificial viscosity
    const double g = b_ax * dwx + b_ay * dwy;
    bB[RB_M] += -Pi * g;
    const double b_Pi = -mo * g;
    b_dwx += -mo * Pi * b_ax;
    b_dwy += -mo * Pi * b_ay;
    b_cbar += b_Pi * (-P.alpha * phi) / rbar;
    const double b_phi = b_Pi * (-P.alpha * cbar + 2.0 * P.beta * phi) / rbar;
    b_rbar -= b_Pi * Pi / rbar;
    b_vx += b_phi * P.h / (r2 + fh2);
    b_r2 -= b_phi * phi / (r2 + fh2);
  }
  b_vijx += b_vx * dx;
  b_vijy += b_vx * dy;
  b_dx += b_vx * vijx;
  b_dy += b_vx * vijy;
  bA[RB_U] += b_vijx;
  bB[RB_U] -= b_vijx;
  bA[RB_V] += b_vijy;
  bB[RB_V] -= b_vijy;
  double b_Ps;
  {                                               // pressure term
    const double b_fp = b_ax * dwx + b_ay * dwy;
    b_dwx += b_ax * fp;
    b_dwy += b_ay * fp;
    bB[RB_M] += b_fp * (-2.0 * Ps * rij);
    b_Ps = b_fp * mo * (-2.0 * rij);
    b_rij += b_fp * mo * (-2.0 * Ps);
  }
  bA[RB_RHO] -= b_rij * rij / A.rho;
  bB[RB_RHO] -= b_rij * rij / B.rho;
  double b_vsx, b_vsy;
  {                                               // continuity
    bA[RB_RHO] += b_drho * 2.0 * Vo * ti;
    b_Vo += b_drho * 2.0 * A.rho * ti;
    const double b_ti = b_drho * 2.0 * A.rho * Vo;
    bA[RB_U] += b_ti * dwx;
    bA[RB_V] += b_ti * dwy;
    b_vsx = -b_ti * dwx;
    b_vsy = -b_ti * dwy;
    b_dwx += b_ti * (A.u - vsx);
    b_dwy += b_ti * (A.v - vsy);
  }
  double b_Pbar = 0.0, b_vbx = 0.0, b_vby = 0.0, b_dU = 0.0, b_ex = 0.0, b_ey = 0.0, b_UL = 0.0, b_UR = 0.0;
  if (nes) {
    b_vbx += b_vsx;
    b_vby += b_vsy;
    b_Pbar += b_Ps;
  } else if (SOLVER == SPH2D_SOLVER_LINEAR_RIEMANN) {
    const double b_Us = b_vsx * ex + b_vsy * ey;
    b_ex += b_vsx * Us;
    b_ey += b_vsy * Us;
    const double b_NP = b_Ps * iz, b_NU = b_Us * iz;
    const double b_iz = b_Ps * NP + b_Us * NU;
    double b_zs = -b_iz * iz * iz, b_zo = -b_iz * iz * iz;
    b_zs += b_NP * (B.p + zo * dU);
    bB[RB_P] += b_NP * zs;
    b_zo += b_NP * (A.p + zs * dU);
    bA[RB_P] += b_NP * zo;
    b_dU += b_NP * zs * zo;
    b_zs += b_NU * UR;
    b_UR += b_NU * zs;
    b_zo += b_NU * UL;
    b_UL += b_NU * zo;
    bA[RB_P] += b_NU;
    bB[RB_P] -= b_NU;
    bA[RB_RHO] += b_zs * A.c;
    bA[RB_C] += b_zs * A.rho;
    bB[RB_RHO] += b_zo * B.c;
    bB[RB_C] += b_zo * B.rho;
  } else {
    b_Pbar += b_Ps;
    const double b_bR = b_Ps * 0.5 * rbar * dU;
    b_rbar += b_Ps * 0.5 * bR * dU;
    b_dU += b_Ps * 0.5 * bR * rbar;
    if (bcase == 1) b_dU += 3.0 * b_bR;
    else if (bcase == 2) b_cbar += b_bR;
    const double b_D = b_vsx * ex + b_vsy * ey;
    b_ex += b_vsx * D;
    b_ey += b_vsy * D;
    b_vbx += b_vsx;
    b_vby += b_vsy;
    const double k = 0.5 / rbar * P.inv_c0;
    bA[RB_P] += b_D * k;
    bB[RB_P] -= b_D * k;
    b_rbar -= b_D * D / rbar;
  }
  bA[RB_P] += 0.5 * b_Pbar;
  bB[RB_P] += 0.5 * b_Pbar;
  bA[RB_U] += 0.5 * b_vbx;
  bB[RB_U] += 0.5 * b_vbx;
  bA[RB_V] += 0.5 * b_vby;
  bB[RB_V] += 0.5 * b_vby;
  bA[RB_C] += 0.5 * b_cbar;
  bB[RB_C] += 0.5 * b_cbar;
  bA[RB_RHO] += 0.5 * b_rbar;
  bB[RB_RHO] += 0.5 * b_rbar;
  b_UL += b_dU;
  b_UR -= b_dU;
  bA[RB_U] += b_UL * ex;
  bA[RB_V] += b_UL * ey;
  b_ex += b_UL * A.u + b_UR * B.u;
  b_ey += b_UL * A.v + b_UR * B.v;
  bB[RB_U] += b_UR * ex;
  bB[RB_V] += b_UR * ey;
  double b_inv_r = -(b_ex * dx + b_ey * dy);
  b_dx += -b_ex * inv_r;
  b_dy += -b_ey * inv_r;
  bB[RB_M] += b_Vo / B.rho;
  bB[RB_RHO] -= b_Vo * Vo / B.rho;
  // the kernel: dwx = kt dx, dwy = kt dy, kt = eq (-2 q) kq, kq = inv_r inv_h kfactor
  const double b_kt = b_dwx * dx + b_dwy * dy;
  b_dx += b_dwx * kt;
  b_dy += b_dwy * kt;
  double b_eq = b_kt * (-2.0 * q) * kq + b_w * P.kfactor;
  double b_q = b_kt * eq * (-2.0) * kq;
  b_inv_r += b_kt * eq * (-2.0 * q) * P.inv_h * P.kfactor;
  b_q += b_eq * eq * (-2.0 * q);
  double b_r = b_q * P.inv_h - b_inv_r * inv_r * inv_r;
  b_r2 += b_r * 0.5 / r;
  b_dx += b_r2 * 2.0 * dx;
  b_dy += b_r2 * 2.0 * dy;
  bA[RB_X] += b_dx;
  bB[RB_X] -= b_dx;
  bA[RB_Y] += b_dy;
  bB[RB_Y] -= b_dy;
}

// d(J)/dq of a side from d(J)/d(rho, p, c, m), through derive_side.
__device__ __forceinline__ double side_q_bar(const DevParams& P, const Side<double>& a, int kind, const double* b) {
  if (kind == KIND_WALL) {
    const double base = 1.0 + 7.0 * a.p / (P.c0sq * P.rho0);
    const double e = (double)0.142857142f;
    const double drho_dp = P.rho0 * (base != 0.0 ? e * pow(base, e) / base : 0.0) * (7.0 / (P.c0sq * P.rho0));
    const double x = a.rho * (1.0 / P.rho0);
    const double dc_drho = P.sqrt_c0sq * 3.0 * x * x * (1.0 / P.rho0);
    return b[RB_P] + (b[RB_RHO] + b[RB_C] * dc_drho + b[RB_M] * P.vol0) * drho_dp;
  }
  const double x = a.rho * (1.0 / P.rho0);
  const double x2 = x * x, x3 = x2 * x;
  const double dp = P.eos_b * 7.0 * x3 * x3 * (1.0 / P.rho0);
  const double dc = P.sqrt_c0sq * 3.0 * x2 * (1.0 / P.rho0);
  return b[RB_RHO] + b[RB_P] * dp + b[RB_C] * dc + (kind == KIND_TIED ? b[RB_M] * P.vol0 : 0.0);
}


// Reverse mode of the weighted pass-A terms of A with neighbour B: strain S_xx, S_xy,
// S_yy and grad rho (Solver...SPS.f90:144-182,508-526 with the identity
// renormalisation), weights g; bA, bB receive d(J)/d(x, y, u, v, rho, p, c, m) of
// either side.
__device__ __forceinline__ void sums_pair_rev(const DevParams& P, const Side<double>& A, const Side<double>& B,
                                              const double* g, double* bA, double* bB) {
#pragma unroll
  for (int k = 0; k < RB_N; ++k) { bA[k] = 0.0; bB[k] = 0.0; }
  const double dx = A.x - B.x, dy = A.y - B.y;
  const double r = sqrt(dx * dx + dy * dy);
  const double q = r * P.inv_h;
  const double eq = exp(-(q * q));
  const double kq = ((1.0 / r) * P.inv_h) * P.kfactor;
  const double kt = eq * (-2.0 * q) * kq;
  const double dwx = kt * dx, dwy = kt * dy;
  const double Vo = B.m / B.rho;
  const double vdx = B.u - A.u, vdy = B.v - A.v;
  const double drh = B.rho - A.rho;
  const double rd = drh * Vo;
  const double b_dwx = g[0] * vdx * Vo + g[1] * vdy * Vo + g[3] * rd;
  const double b_dwy = g[1] * vdx * Vo + g[2] * vdy * Vo + g[4] * rd;
  const double b_vdx = g[0] * dwx * Vo + g[1] * dwy * Vo;
  const double b_vdy = g[1] * dwx * Vo + g[2] * dwy * Vo;
  const double b_rd = g[3] * dwx + g[4] * dwy;
  const double b_Vo = g[0] * vdx * dwx + g[1] * (vdx * dwy + vdy * dwx) + g[2] * vdy * dwy + b_rd * drh;
  bB[RB_RHO] += b_rd * Vo - b_Vo * Vo / B.rho;
  bA[RB_RHO] -= b_rd * Vo;
  bB[RB_U] += b_vdx;
  bA[RB_U] -= b_vdx;
  bB[RB_V] += b_vdy;
  bA[RB_V] -= b_vdy;
  bB[RB_M] += b_Vo / B.rho;
  const double b_kt = b_dwx * dx + b_dwy * dy;
  double b_dx = b_dwx * kt, b_dy = b_dwy * kt;
  const double b_eq = b_kt * (-2.0 * q) * kq;
  double b_q = b_kt * eq * (-2.0) * kq + b_eq * eq * (-2.0 * q);
  const double b_invr = b_kt * eq * (-2.0 * q) * P.inv_h * P.kfactor;
  const double b_r = b_q * P.inv_h - b_invr / (r * r);
  b_dx += b_r * dx / r;
  b_dy += b_r * dy / r;
  bA[RB_X] += b_dx;
  bB[RB_X] -= b_dx;
  bA[RB_Y] += b_dy;
  bB[RB_Y] -= b_dy;
}

struct AdjBufs {
  const double* lam;   // gid mode: 5 per fluid gid (drho, ax, ay, avg_x, avg_y); slot mode: unused
  double* w;           // kLam per local particle (PairLam, then grad rho x, y)
  double* gs;          // 5 per local particle: d/d(x, y, u, v, q)
  double* ga;          // 5 per local particle: d/d(grad rho x, y) then d/d(tau xx, xy, yy); then
                       //   d/d(S xx, xy, yy, grad rho x, y) after k_adj_tau
  double* G;           // gid mode: 5 per gid (fluid and ghosts); slot mode: 5 per local particle
  int do_shift;        // the evaluation's shift sums carry weights
};

__global__ void __launch_bounds__(kBlock) k_adj_weights(DevParams P, Buffers B, AdjBufs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int g = B.name[s];
  const int fluid = B.l_type[s] == 2;
  double l[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  if (g >= 0 && g < B.counters[C_NFLUID])
    for (int k = 0; k < 5; ++k) l[k] = A.lam[5 * (size_t)g + k];
  const bool xsph = P.solver == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && P.epsilon_xsph != 0.0;
  double* w = A.w + kLam * (size_t)s;
  w[0] = l[0];
  w[1] = fluid ? l[0] : 0.0;
  w[2] = fluid ? l[1] : 0.0;
  w[3] = fluid ? l[2] : 0.0;
  w[4] = xsph ? l[3] : 0.0;
  w[5] = xsph ? l[4] : 0.0;
  w[6] = 0.0; w[7] = 0.0; w[8] = 0.0; w[9] = 0.0;
}

__device__ __forceinline__ PairLam pair_lam(const AdjBufs& A, int s) {
  const double* w = A.w + kLam * (size_t)s;
  return PairLam{w[0], w[1], w[2], w[3], w[4], w[5], w[6], w[7]};
}

// The two halves of the force pass's reverse for local s: SIDE 0, the terms s sums
// over its neighbours (weights of s; writes gs, ga); SIDE 1, the terms each neighbour
// o sums over s (weights of o; adds to gs, ga). Apart, each half fits 256 registers
// and runs 2 waves per SIMD where the two together would take one.
template <int SOLVER, int SIDE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(2)))
k_adj_force(DevParams P, Buffers B, AdjBufs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int ts = B.l_type[s];
  const double2 ps = B.l_pos[s];
  int ks;
  const Side<double> vs = load_side(P, B, s, ks);
  // d/d(x, y, u, v, rho, p, c, m) of s, then d/d(grad rho, tau) of s
  double bs[RB_N] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  double acc[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) {
    if (ts != 2 && B.l_type[o] != 2) return;
    int ko;
    const Side<double> vo = load_side(P, B, o, ko);
    double bx[RB_N], by[RB_N], a[5];
    if (SIDE == 0) {
      force_pair_rev<SOLVER>(P, vs, vo, B.g_mass[o], B.near_[s], pair_lam(A, s), A.do_shift, bx, by, a);
#pragma unroll
      for (int k = 0; k < RB_N; ++k) bs[k] = bs[k] + bx[k];
    } else {
      force_pair_rev<SOLVER>(P, vo, vs, B.g_mass[s], B.near_[o], pair_lam(A, o), A.do_shift, by, bx, a);
#pragma unroll
      for (int k = 0; k < RB_N; ++k) bs[k] = bs[k] + bx[k];
    }
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[5 + k] = acc[5 + k] + a[k];
  }, B.nb_list, B.nb_count, B.kstride);
  acc[0] = bs[RB_X];
  acc[1] = bs[RB_Y];
  acc[2] = bs[RB_U];
  acc[3] = bs[RB_V];
  acc[4] = side_q_bar(P, vs, ks, bs);
  double* gs = A.gs + 5 * (size_t)s;
  double* ga = A.ga + 5 * (size_t)s;
  if (SIDE == 0) {
    const double* w = A.w + kLam * (size_t)s;
    acc[5] = acc[5] + w[8];                       // grad rho is an output too (the shift uses it)
    acc[6] = acc[6] + w[9];
    for (int k = 0; k < 5; ++k) { gs[k] = acc[k]; ga[k] = acc[5 + k]; }
  } else {
    for (int k = 0; k < 5; ++k) { gs[k] = gs[k] + acc[k]; ga[k] = ga[k] + acc[5 + k]; }
  }
}

// tau = rho (2 nut S - k - b) etc. (Solver...SPS.f90:256-296) reversed: the tau
// weights become strain weights; the rho share goes to d/d(rho).
__global__ void __launch_bounds__(kBlock) k_adj_tau(DevParams P, Buffers B, AdjBufs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int ts = B.l_type[s];
  const double2 ps = B.l_pos[s];
  const double2 vs = B.l_vel[s];
  double* ga = A.ga + 5 * (size_t)s;
  const double gtx = ga[2], gty = ga[3], gtz = ga[4];
  double gS[3] = {0.0, 0.0, 0.0};
  if (ts == 2) {
    double Sxx = 0.0, Sxy = 0.0, Syy = 0.0;
    for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) {
      const double2 po = B.l_pos[o];
      const double dx = ps.x - po.x, dy = ps.y - po.y;
      const double r = sqrt(dx * dx + dy * dy);
      const double q = r * P.inv_h;
      const double kt = exp(-(q * q)) * (-2.0 * q) * ((1.0 / r) * P.inv_h) * P.kfactor;
      const double dwx = kt * dx, dwy = kt * dy;
      const double Vo = B.l_rv[o].y;
      const double2 vo = B.l_vel[o];
      const double vdx = vo.x - vs.x, vdy = vo.y - vs.y;
      Sxx = Sxx + vdx * dwx * Vo;
      Sxy = Sxy + (vdx * dwy + vdy * dwx) * Vo;
      Syy = Syy + vdy * dwy * Vo;
    }, B.nb_list, B.nb_count, B.kstride);
    Dual<4> xx(Sxx), xy(Sxy), yy(Syy), rho(B.l_rv[s].x);
    xx.d[0] = 1.0; xy.d[1] = 1.0; yy.d[2] = 1.0; rho.d[3] = 1.0;
    const Dual<4> abs2 = 2.0 * (xx * xx + yy * yy) + xy * xy;
    const double t0 = P.sps_c012 * P.sps_dl;
    const Dual<4> nut = (t0 * t0) * tsqrt(abs2);
    const Dual<4> kk = P.sps_c23 * nut * (xx + yy);
    const Dual<4> blin = P.sps_c23b * (P.sps_dl * P.sps_dl) * abs2;
    const Dual<4> J = gtx * ((2.0 * nut * xx - kk - blin) * rho) + gty * ((nut * xy) * rho) +
                      gtz * ((2.0 * nut * yy - kk - blin) * rho);
    gS[0] = J.d[0]; gS[1] = J.d[1]; gS[2] = J.d[2];
    A.gs[5 * (size_t)s + 4] += J.d[3];
  }
  const double ggx = ga[0], ggy = ga[1];
  ga[0] = gS[0]; ga[1] = gS[1]; ga[2] = gS[2]; ga[3] = ggx; ga[4] = ggy;
}

// Reverse of the pass-A sums (S and grad rho) into d/d(state_s). SLOT: the result
// per local particle in G; else per gid.
template <bool SLOT>
__global__ void __launch_bounds__(kBlock) k_adj_sums(DevParams P, Buffers B, AdjBufs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int ts = B.l_type[s];
  const double2 ps = B.l_pos[s];
  int ks;
  const Side<double> vs = load_side(P, B, s, ks);
  const double* gas = A.ga + 5 * (size_t)s;
  double bs[RB_N] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) {
    if (ts != 2 && B.l_type[o] != 2) return;
    int ko;
    const Side<double> vo = load_side(P, B, o, ko);
    double b1s[RB_N], b1o[RB_N], b2o[RB_N], b2s[RB_N];
    sums_pair_rev(P, vs, vo, gas, b1s, b1o);                       // s sums over o
    sums_pair_rev(P, vo, vs, A.ga + 5 * (size_t)o, b2o, b2s);       // o sums over s
#pragma unroll
    for (int k = 0; k < RB_N; ++k) bs[k] = bs[k] + (b1s[k] + b2s[k]);
  }, B.nb_list, B.nb_count, B.kstride);
  const double acc[5] = {bs[RB_X], bs[RB_Y], bs[RB_U], bs[RB_V], side_q_bar(P, vs, ks, bs)};
  const size_t dst = SLOT ? (size_t)s : (size_t)B.name[s];
  const double* gs = A.gs + 5 * (size_t)s;
  for (int k = 0; k < 5; ++k) A.G[5 * dst + k] = gs[k] + acc[k];
}

// Assign_boundary_particles.f90:157-299 reversed, gather form: each non-wall local
// particle j adds, over the walls w whose Shepard support holds it, the weights of
// (u_w, v_w, p_w) times the derivatives of the wall's interpolation with respect to
// (x_j, y_j, u_j, v_j, rho_j): p_w = sum_k F_k (p_k - g rho_k (y_w - y_k)),
// u_w = s_u sum_k F_k u_k, F_k = W_k / sum W, so dF_k/dx_j = dW_j/dx_j (delta_kj - F_k)
// / sum W and p_k = Tait(rho_k). Walls with fewer than 3 supports or p_w = 0 keep
// fixed values. A first pass stores each wall's sum W.
__device__ __forceinline__ double wall_wdw(const DevParams& P, double r, double& dwdr) {
  const double q = r / P.wall_h;
  const double e = exp(-(q * q));
  dwdr = e * (-2.0 * q) / P.wall_h * P.wall_kfactor;
  return (e - P.e9) * P.wall_kfactor;
}

__global__ void __launch_bounds__(kBlock) k_adj_wall_sum(DevParams P, Buffers B, double* wsum) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  if (B.l_type[s] != -3) return;
  const double2 xs = B.l_pos[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  const double rmax = 3 * P.wall_h, rmin = (double)1.0E-7f;
  double a = 0.0;
  int nt = 0;
  for (int d = 0; d < 3; ++d)
    for (int o = R.lo[d]; o < R.hi[d]; ++o) {
      if (B.l_type[o] == -3) continue;
      const double2 xo = B.l_pos[o];
      const double px = xs.x - xo.x, py = xs.y - xo.y;
      const double r = sqrt(px * px + py * py);
      if (r <= rmax && r >= rmin) {
        double dw;
        a = a + wall_wdw(P, r, dw) * 1.0;
        ++nt;
      }
    }
  wsum[s] = (nt >= 3 && B.l_pc[s].x != 0.0) ? a : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_adj_wall_push(DevParams P, Buffers B, AdjBufs A, const double* wsum) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= B.counters[C_NLOCAL]) return;
  if (B.l_type[j] == -3) return;
  const double2 xj = B.l_pos[j];
  // the wall particles lie outside [0, box_x] x [0, box_y] (wall_layout): a particle
  // farther than their support from every side of the box has none within it
  const double rsup = 3 * P.wall_h;
  if (xj.x > rsup && P.box_x - xj.x > rsup && xj.y > rsup && P.box_y - xj.y > rsup) return;
  const double2 vj = B.l_vel[j];
  const double rhoj = B.l_rv[j].x;
  const double pj = B.l_pc[j].x;
  const double x7 = rhoj / P.rho0, x6 = x7 * x7 * x7 * x7 * x7 * x7;
  const double dpdrho = P.eos_b * 7.0 * x6 / P.rho0;         // d Tait / d rho
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[j]);
  const double rmax = 3 * P.wall_h, rmin = (double)1.0E-7f;
  double g[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int d = 0; d < 3; ++d)
    for (int w = R.lo[d]; w < R.hi[d]; ++w) {
      if (B.l_type[w] != -3) continue;
      const double sw = wsum[w];
      if (sw == 0.0) continue;
      const double2 xw = B.l_pos[w];
      const double px = xw.x - xj.x, py = xw.y - xj.y;
      const double r = sqrt(px * px + py * py);
      if (!(r <= rmax && r >= rmin)) continue;
      double dwdr;
      const double W = wall_wdw(P, r, dwdr);
      const double F = 1.0 * ((1.0 / sw) * W);
      const int bt = B.g_btype[B.name[w]];
      const double su = (bt == -7 || bt == -8) ? -1.0 : 1.0, sv = (bt == -6 || bt == -7) ? -1.0 : 1.0;
      const double* gw = A.G + 5 * (size_t)w;
      const double lu = su * gw[2], lv = sv * gw[3], lp = gw[4];
      const double2 vw = B.l_vel[w];
      const double ub = su * vw.x, vb = sv * vw.y, pw = B.l_pc[w].x;
      const double qj = P.g != 0.0 ? pj - P.g * rhoj * (xw.y - xj.y) : pj;
      // dW_j / d(x_j, y_j) with r = |x_w - x_j|
      const double wx = -dwdr * px / r, wy = -dwdr * py / r;
      const double cw = (lp * (qj - pw) + lu * (vj.x - ub) + lv * (vj.y - vb)) / sw;
      g[0] += cw * wx;
      g[1] += cw * wy + (P.g != 0.0 ? lp * F * P.g * rhoj : 0.0);
      g[2] += lu * F;
      g[3] += lv * F;
      g[4] += lp * F * (dpdrho - (P.g != 0.0 ? P.g * (xw.y - xj.y) : 0.0));
    }
  double* gj = A.G + 5 * (size_t)j;
  for (int k = 0; k < 5; ++k) gj[k] = gj[k] + g[k];
}

__global__ void k_adj_zero(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// Corner ghosts (y-copies of x-ghosts, Distribute_Period_Boundary_Particle.f90:153-206)
// into the x-ghost they copy: x-ghost g owns the cnt_yx[g] slots from
// nxg + nyf + off_yx[g].
__global__ void k_adj_corners(Buffers B, AdjBufs A, int cap_ghost) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  const int nxg = B.counters[C_NXG], nyf = B.counters[C_NYF];
  if (g >= cap_ghost || g >= nxg) return;
  const int nc = B.cnt_yx[g];
  double* dst = A.G + 5 * (size_t)(N + g);
  for (int t = 0; t < nc; ++t) {
    const int slot = nxg + nyf + B.off_yx[g] + t;
    if (slot >= cap_ghost) break;                     // the emission clamps at the capacity too
    const double* src = A.G + 5 * (size_t)(N + slot);
    for (int k = 0; k < 5; ++k) dst[k] = dst[k] + src[k];
  }
}

// Every x- and y-ghost into the fluid particle it copies. The slots are found by
// the arithmetic of the ghost emission (block offsets gblk, ballot ranks within
// the block), so every fluid particle adds its own ghosts in a fixed order.
__global__ void __launch_bounds__(kBlock) k_adj_owner(DevParams P, Buffers B, AdjBufs A, int cap_fluid,
                                                      int cap_ghost) {
  __shared__ int sx[kBlock / 64], sy[kBlock / 64];
  const int b = blockIdx.x, nb = gridDim.x;
  const int nxg = B.counters[C_NXG];
  const int ox = B.gblk[b], oy = B.gblk[nb + b];
  const int i = b * kBlock + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  int l = 0, r = 0, yl = 0, yr = 0;
  if (i < N && i < cap_fluid && B.g_type[i] != 120) ghost_flags(P, B.g_pos[i], l, r, yl, yr);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint64_t bl = __ballot(l), br = __ballot(r), byl = __ballot(yl), byr = __ballot(yr);
  int px = __popcll(bl & lt) + __popcll(br & lt);
  int py = __popcll(byl & lt) + __popcll(byr & lt);
  if (lane == 0) { sx[w] = __popcll(bl) + __popcll(br); sy[w] = __popcll(byl) + __popcll(byr); }
  __syncthreads();
  for (int k = 0; k < w; ++k) { px += sx[k]; py += sy[k]; }
  if (i >= N || i >= cap_fluid) return;
  double acc[5];
  double* dst = A.G + 5 * (size_t)i;
  for (int k = 0; k < 5; ++k) acc[k] = dst[k];
  int slots[4], ns = 0;
  int k = ox + px;
  if (l) slots[ns++] = k++;
  if (r) slots[ns++] = k;
  k = nxg + oy + py;
  if (yl) slots[ns++] = k++;
  if (yr) slots[ns++] = k;
  for (int t = 0; t < ns; ++t) {
    if (slots[t] >= cap_ghost) continue;               // not emitted: the capacity clamps the emission
    const double* src = A.G + 5 * (size_t)(N + slots[t]);
    for (int q = 0; q < 5; ++q) acc[q] = acc[q] + src[q];
  }
  for (int q = 0; q < 5; ++q) dst[q] = acc[q];
}

// Weights of the device-array reverse: lam per fluid gid (drho/dt, ax, ay, avg x, y)
// from the caller's (ld, 2) columns; and the gradients back into them.
__global__ void k_vjp_pack(int n, int ld, const double* ldr, const double* lacc, const double* lavg, double* lam) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double* l = lam + 5 * (size_t)i;
  l[0] = ldr[i];
  l[1] = lacc[i];
  l[2] = lacc[(size_t)ld + i];
  l[3] = lavg ? lavg[i] : 0.0;
  l[4] = lavg ? lavg[(size_t)ld + i] : 0.0;
}

__global__ void k_vjp_unpack(int n, int ld, const double* G, double* gpos, double* gvel, double* grho) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double* g = G + 5 * (size_t)i;
  if (gpos) { gpos[i] = g[0]; gpos[(size_t)ld + i] = g[1]; }
  if (gvel) { gvel[i] = g[2]; gvel[(size_t)ld + i] = g[3]; }
  if (grho) grho[i] = g[4];
}

void launch_vjp_pack(hipStream_t st, int n, int ld, const double* ldr, const double* lacc, const double* lavg,
                     double* lam) {
  hipLaunchKernelGGL(k_vjp_pack, dim3(grid_for(n)), dim3(kBlock), 0, st, n, ld, ldr, lacc, lavg, lam);
}

void launch_vjp_unpack(hipStream_t st, int n, int ld, const double* G, double* gpos, double* gvel, double* grho) {
  hipLaunchKernelGGL(k_vjp_unpack, dim3(grid_for(n)), dim3(kBlock), 0, st, n, ld, G, gpos, gvel, grho);
}

static void launch_force_tau(hipStream_t st, const DevParams& P, Buffers& B, int g, const AdjBufs& A) {
  if (P.solver == SPH2D_SOLVER_LINEAR_RIEMANN) {
    hipLaunchKernelGGL((k_adj_force<SPH2D_SOLVER_LINEAR_RIEMANN, 0>), dim3(g), dim3(kBlock), 0, st, P, B, A);
    hipLaunchKernelGGL((k_adj_force<SPH2D_SOLVER_LINEAR_RIEMANN, 1>), dim3(g), dim3(kBlock), 0, st, P, B, A);
  } else {
    hipLaunchKernelGGL((k_adj_force<SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN, 0>), dim3(g), dim3(kBlock), 0, st, P, B, A);
    hipLaunchKernelGGL((k_adj_force<SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN, 1>), dim3(g), dim3(kBlock), 0, st, P, B, A);
  }
  hipLaunchKernelGGL(k_adj_tau, dim3(g), dim3(kBlock), 0, st, P, B, A);
}

void launch_eval_vjp(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, int cap_fluid, int cap_ghost,
                     int cap_global, const double* lam, double* w, double* gs, double* ga, double* G) {
  const AdjBufs A{lam, w, gs, ga, G, 0};
  const int g = grid_for(cap_local);
  const size_t nG = 5 * (size_t)cap_global;
  hipLaunchKernelGGL(k_adj_zero, dim3((unsigned)((nG + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, G, nG);
  hipLaunchKernelGGL(k_adj_weights, dim3(g), dim3(kBlock), 0, st, P, B, A);
  launch_force_tau(st, P, B, g, A);
  hipLaunchKernelGGL(k_adj_sums<false>, dim3(g), dim3(kBlock), 0, st, P, B, A);
  hipLaunchKernelGGL(k_adj_corners, dim3(grid_for(cap_ghost)), dim3(kBlock), 0, st, B, A, cap_ghost);
  hipLaunchKernelGGL(k_adj_owner, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, st, P, B, A, cap_fluid, cap_ghost);
}

// Slot form (the step reverse): weights w (kLam per local particle) already set by
// the caller; G receives d/d(x, y, u, v, q) per local particle with the walls'
// shares pushed into the particles they interpolate. wsum: one double per local
// particle.
void launch_eval_vjp_slots(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, int do_shift, double* w,
                           double* gs, double* ga, double* G, double* wsum) {
  const AdjBufs A{nullptr, w, gs, ga, G, do_shift};
  const int g = grid_for(cap_local);
  launch_force_tau(st, P, B, g, A);
  hipLaunchKernelGGL(k_adj_sums<true>, dim3(g), dim3(kBlock), 0, st, P, B, A);
  if (P.walls) {
    hipLaunchKernelGGL(k_adj_wall_sum, dim3(g), dim3(kBlock), 0, st, P, B, wsum);
    hipLaunchKernelGGL(k_adj_wall_push, dim3(g), dim3(kBlock), 0, st, P, B, A, wsum);
  }
}

}  // namespace sph2d
