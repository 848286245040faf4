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
//                   o (weights of o); state directions by 5-direction duals, grad rho
//                   and tau (linear in the pair terms) in closed form
//   k_adj_tau       per local s: recompute the strain S_s, reverse tau(S_s, rho_s)
//   k_adj_sums      per local s, over neighbours o: reverse of the pass-A sums
//                   (S and grad rho) into d/d(state_s); result stored by gid
//   k_adj_corners   corner ghosts into the x-ghost they copy
//   k_adj_owner     every ghost into its fluid particle (same slot arithmetic as
//                   the ghost emission of Distribute_Period_Boundary_Particle.f90)
// Pair-level Jacobians with respect to the state come from forward-mode duals over
// the same pair formulas as the FP64 force pass; the reverse sweep over the
// particle graph is explicit.
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

// One side of a pair: state, grad rho and the SPS stress of the particle.
template <class T>
struct Side {
  T x, y, u, v, rho, gx, gy, txx, txy, tyy;
};

// Adjoint weights of the summing particle's outputs: drho/dt (solver part),
// the delta-SPH correction (fluid only), acceleration (fluid only), XSPH.
struct PairLam {
  double d, c, ax, ay, gx, gy;
};

// status_equation.f90:52-61 (Tait), as in the forward evaluation.
template <class T>
__device__ __forceinline__ void tait_t(const DevParams& P, const T& rho, T& p, T& c) {
  const T x = rho * (1.0 / P.rho0);
  const T x3 = x * x * x;
  p = P.eos_b * (x3 * x3 * x - 1.0);
  c = P.sqrt_c0sq * x3;
}

// Weighted force-pass pair terms summed by particle S with neighbour O: the same
// formulas as eval_c in FP64 (Solver_Normalized_*_Riemann_Delta_SPS.f90:300-583).
// gaux receives d/d(grad rho x, y, tau xx, xy, yy) of either side: those enter the
// pair terms linearly, through the sums S + O, so both sides share one closed form.
template <class T, int SOLVER>
__device__ __forceinline__ T force_pair(const DevParams& P, const Side<T>& S, const Side<T>& O, double mo,
                                        double gmo, int nes, const PairLam& L, double* gaux) {
  const T dx = S.x - O.x, dy = S.y - O.y;
  const T r2 = dx * dx + dy * dy;
  const T r = tsqrt(r2);
  const T inv_r = 1.0 / r;
  const T q = r * P.inv_h;
  const T eq = texp(-(q * q));
  const T w = (eq - P.e9) * P.kfactor;
  const T kt = eq * (-2.0 * q) * (inv_r * P.inv_h) * P.kfactor;
  const T dwx = kt * dx, dwy = kt * dy;
  T ps, cs, po, co;
  tait_t(P, S.rho, ps, cs);
  tait_t(P, O.rho, po, co);
  const T Vo = mo / O.rho;
  const T ex = -(dx * inv_r), ey = -(dy * inv_r);
  const T UL = S.u * ex + S.v * ey;
  const T UR = O.u * ex + O.v * ey;
  const T dU = UL - UR;
  const T Pbar = (ps + po) * 0.5;
  const T vbx = (S.u + O.u) * 0.5, vby = (S.v + O.v) * 0.5;
  const T cbar = (cs + co) * 0.5;
  const T rbar = (S.rho + O.rho) * 0.5;
  T vsx, vsy, Ps;
  if (SOLVER == SPH2D_SOLVER_LINEAR_RIEMANN) {
    const T zs = S.rho * cs, zo = O.rho * co;
    const T inv_z = 1.0 / (zs + zo);
    const T Us = (zs * UR + zo * UL + ps - po) * inv_z;
    vsx = Us * ex;
    vsy = Us * ey;
    Ps = (zs * po + zo * ps + zs * zo * dU) * inv_z;
  } else {
    const T Ub = (UL + UR) * 0.5;
    const T Us = Ub + 0.5 * (ps - po) / rbar * P.inv_c0;
    vsx = (Us - Ub) * ex + vbx;
    vsy = (Us - Ub) * ey + vby;
    T bR = val(dU) > 0.0 ? 3.0 * dU : T(0.0);
    bR = val(bR) < val(cbar) ? bR : cbar;
    Ps = Pbar + 0.5 * bR * rbar * dU;
  }
  if (nes) { vsx = vbx; vsy = vby; Ps = Pbar; }
  const T ti = (S.u - vsx) * dwx + (S.v - vsy) * dwy;
  const T drho = 2.0 * S.rho * Vo * ti;
  const T rij = 1.0 / (S.rho * O.rho);
  const T fp = mo * (-2.0 * Ps * rij);
  T axp = fp * dwx, ayp = fp * dwy;
  const T vijx = S.u - O.u, vijy = S.v - O.v;
  const T vx = vijx * dx + vijy * dy;
  if (val(vx) < 0.0) {
    const double fh = P.fai * P.h;
    const T phi = P.h * vx / (r2 + fh * fh);
    const T Pi = (-(P.alpha * cbar * phi) + P.beta * (phi * phi)) / rbar;
    axp = axp + mo * (-Pi * dwx);
    ayp = ayp + mo * (-Pi * dwy);
  }
  const T Rd = dx * dwx + dy * dwy;
  const T lf = Vo * 2.0 * P.nu * Rd / r2;
  axp = axp + lf * vijx;
  ayp = ayp + lf * vijy;
  const T tmp = P.delta_alpha * P.h * cbar * vx / r2 * P.rho0 / S.rho * Vo;
  axp = axp + tmp * dwx;
  ayp = ayp + tmp * dwy;
  T tp = (S.txx + O.txx) * dwx + (S.txy + O.txy) * dwy;
  tp = tp * rij;
  axp = axp + gmo * tp;
  tp = tp + (S.tyy + O.tyy) * dwy + (S.txy + O.txy) * dwx;
  tp = tp * rij;
  ayp = ayp + gmo * tp;
  T J = L.d * drho + L.ax * axp + L.ay * ayp;
  if (SOLVER == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && P.epsilon_xsph != 0.0) {
    const T f = P.epsilon_xsph * 2.0 * mo / (S.rho + O.rho) * w;
    J = J - L.gx * (f * vijx) - L.gy * (f * vijy);
  }
  const T psi = (O.rho - S.rho) + 0.5 * ((S.gx + O.gx) * dx + (S.gy + O.gy) * dy);
  const T cf = 2.0 * P.delta * P.h * cbar * Rd / r2 * Vo;
  J = J - L.c * (cf * psi);
  {
    const double rv = val(rij), wx = val(dwx), wy = val(dwy), cv = val(cf);
    gaux[0] = -L.c * cv * 0.5 * val(dx);
    gaux[1] = -L.c * cv * 0.5 * val(dy);
    gaux[2] = gmo * rv * (L.ax * wx + L.ay * (wx * rv));
    gaux[3] = gmo * rv * (L.ax * wy + L.ay * (wy * rv + wx));
    gaux[4] = gmo * rv * (L.ay * wy);
  }
  return J;
}

// Weighted pass-A pair terms of S with neighbour O: strain S_xx, S_xy, S_yy and
// grad rho (Solver...SPS.f90:144-182,508-526 with the identity renormalisation).
template <class T>
__device__ __forceinline__ T sums_pair(const DevParams& P, const Side<T>& S, const Side<T>& O, double mo,
                                       const double* g) {
  const T dx = S.x - O.x, dy = S.y - O.y;
  const T r = tsqrt(dx * dx + dy * dy);
  const T q = r * P.inv_h;
  const T eq = texp(-(q * q));
  const T kt = eq * (-2.0 * q) * ((1.0 / r) * P.inv_h) * P.kfactor;
  const T dwx = kt * dx, dwy = kt * dy;
  const T Vo = mo / O.rho;
  const T vdx = O.u - S.u, vdy = O.v - S.v;
  const T rd = (O.rho - S.rho) * Vo;
  return g[0] * (vdx * dwx * Vo) + g[1] * ((vdx * dwy + vdy * dwx) * Vo) + g[2] * (vdy * dwy * Vo) +
         g[3] * (rd * dwx) + g[4] * (rd * dwy);
}

template <class T>
__device__ __forceinline__ Side<T> load_side(const Buffers& B, int s) {
  Side<T> a;
  const double2 p = B.l_pos[s], v = B.l_vel[s];
  const double2 g = B.grad[s];
  a.x = p.x; a.y = p.y; a.u = v.x; a.v = v.y; a.rho = B.l_rv[s].x;
  a.gx = g.x; a.gy = g.y;
  a.txx = B.tau[3 * s]; a.txy = B.tau[3 * s + 1]; a.tyy = B.tau[3 * s + 2];
  return a;
}

template <int N>
__device__ __forceinline__ void seed_side(Side<Dual<N>>& a, int first, int count) {
  Dual<N>* f[10] = {&a.x, &a.y, &a.u, &a.v, &a.rho, &a.gx, &a.gy, &a.txx, &a.txy, &a.tyy};
  for (int k = 0; k < count; ++k) f[k]->d[first + k] = 1.0;
}

struct AdjBufs {
  const double* lam;   // 5 per fluid gid: drho, ax, ay, avg_x, avg_y
  double* w;           // 6 per local particle: PairLam
  double* gs;          // 5 per local particle: d/d(x, y, u, v, rho)
  double* ga;          // 5 per local particle: d/d(grad rho x, y) then d/d(tau xx, xy, yy); then
                       //   d/d(S xx, xy, yy, grad rho x, y) after k_adj_tau
  double* G;           // 5 per gid (fluid and ghosts)
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
  double* w = A.w + 6 * (size_t)s;
  w[0] = l[0];
  w[1] = fluid ? l[0] : 0.0;
  w[2] = fluid ? l[1] : 0.0;
  w[3] = fluid ? l[2] : 0.0;
  w[4] = xsph ? l[3] : 0.0;
  w[5] = xsph ? l[4] : 0.0;
}

__device__ __forceinline__ PairLam pair_lam(const AdjBufs& A, int s) {
  const double* w = A.w + 6 * (size_t)s;
  return PairLam{w[0], w[1], w[2], w[3], w[4], w[5]};
}

template <int SOLVER>
__global__ void __launch_bounds__(kBlock) k_adj_force(DevParams P, Buffers B, AdjBufs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int ts = B.l_type[s];
  const double2 ps = B.l_pos[s];
  const Side<double> vs = load_side<double>(B, s);
  const PairLam Ls = pair_lam(A, s);
  const int nes = B.near_[s];
  const double ms = B.l_mass[s], gms = B.g_mass[s];
  double acc[10] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) {
    if (ts != 2 && B.l_type[o] != 2) return;
    const Side<double> vo = load_side<double>(B, o);
    Side<Dual<5>> S, O;
    S.x = vs.x; S.y = vs.y; S.u = vs.u; S.v = vs.v; S.rho = vs.rho;
    S.gx = vs.gx; S.gy = vs.gy; S.txx = vs.txx; S.txy = vs.txy; S.tyy = vs.tyy;
    O.x = vo.x; O.y = vo.y; O.u = vo.u; O.v = vo.v; O.rho = vo.rho;
    O.gx = vo.gx; O.gy = vo.gy; O.txx = vo.txx; O.txy = vo.txy; O.tyy = vo.tyy;
    seed_side(S, 0, 5);
    double a1[5], a2[5];
    // terms summed by s (s is the summing particle)
    const Dual<5> J1 = force_pair<Dual<5>, SOLVER>(P, S, O, B.l_mass[o], B.g_mass[o], nes, Ls, a1);
    // terms summed by o (s is its neighbour)
    const Dual<5> J2 = force_pair<Dual<5>, SOLVER>(P, O, S, ms, gms, B.near_[o], pair_lam(A, o), a2);
#pragma unroll
    for (int k = 0; k < 5; ++k) {
      acc[k] = acc[k] + (J1.d[k] + J2.d[k]);
      acc[5 + k] = acc[5 + k] + (a1[k] + a2[k]);
    }
  }, B.nb_list, B.nb_count, B.kstride);
  double* gs = A.gs + 5 * (size_t)s;
  double* ga = A.ga + 5 * (size_t)s;
  for (int k = 0; k < 5; ++k) { gs[k] = acc[k]; ga[k] = acc[5 + k]; }
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

__global__ void __launch_bounds__(kBlock) k_adj_sums(DevParams P, Buffers B, AdjBufs A) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int ts = B.l_type[s];
  const double2 ps = B.l_pos[s];
  const Side<double> vs = load_side<double>(B, s);
  const double ms = B.l_mass[s];
  const double* gas = A.ga + 5 * (size_t)s;
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) {
    if (ts != 2 && B.l_type[o] != 2) return;
    const Side<double> vo = load_side<double>(B, o);
    Side<Dual<5>> S, O;
    S.x = vs.x; S.y = vs.y; S.u = vs.u; S.v = vs.v; S.rho = vs.rho;
    O.x = vo.x; O.y = vo.y; O.u = vo.u; O.v = vo.v; O.rho = vo.rho;
    seed_side(S, 0, 5);
    const Dual<5> J1 = sums_pair<Dual<5>>(P, S, O, B.l_mass[o], gas);
    const Dual<5> J2 = sums_pair<Dual<5>>(P, O, S, ms, A.ga + 5 * (size_t)o);
#pragma unroll
    for (int k = 0; k < 5; ++k) acc[k] = acc[k] + (J1.d[k] + J2.d[k]);
  }, B.nb_list, B.nb_count, B.kstride);
  const int g = B.name[s];
  const double* gs = A.gs + 5 * (size_t)s;
  for (int k = 0; k < 5; ++k) A.G[5 * (size_t)g + k] = gs[k] + acc[k];
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

void launch_eval_vjp(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, int cap_fluid, int cap_ghost,
                     int cap_global, const double* lam, double* w, double* gs, double* ga, double* G) {
  const AdjBufs A{lam, w, gs, ga, G};
  const int g = grid_for(cap_local);
  const size_t nG = 5 * (size_t)cap_global;
  hipLaunchKernelGGL(k_adj_zero, dim3((unsigned)((nG + kBlock - 1) / kBlock)), dim3(kBlock), 0, st, G, nG);
  hipLaunchKernelGGL(k_adj_weights, dim3(g), dim3(kBlock), 0, st, P, B, A);
  if (P.solver == SPH2D_SOLVER_LINEAR_RIEMANN)
    hipLaunchKernelGGL(k_adj_force<SPH2D_SOLVER_LINEAR_RIEMANN>, dim3(g), dim3(kBlock), 0, st, P, B, A);
  else
    hipLaunchKernelGGL(k_adj_force<SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN>, dim3(g), dim3(kBlock), 0, st, P, B, A);
  hipLaunchKernelGGL(k_adj_tau, dim3(g), dim3(kBlock), 0, st, P, B, A);
  hipLaunchKernelGGL(k_adj_sums, dim3(g), dim3(kBlock), 0, st, P, B, A);
  hipLaunchKernelGGL(k_adj_corners, dim3(grid_for(cap_ghost)), dim3(kBlock), 0, st, B, A, cap_ghost);
  hipLaunchKernelGGL(k_adj_owner, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, st, P, B, A, cap_fluid, cap_ghost);
}

}  // namespace sph2d
