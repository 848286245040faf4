// sph2d_ms.hip -- the moving-square program of the reference
// (Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/, "MS" below) on the GPU:
// a tank of fixed ghost walls with a square body of fixed ghosts moved at a
// prescribed velocity, the linear EOS, no free surface, XSPH velocity inside the
// continuity equation, three Riemann solvers and the body force.
//
// Same gather form as the TGV evaluation: every particle sums over its own
// neighbours (pair i < j of the reference's local order is (left, right) of the
// star state; every other pair term is written from the summing particle's side,
// which the reference's j-side formulas reduce to), no pair list and no float
// atomics. The force evaluation is templated on the arithmetic type of the pair
// sums: double (the FP64 form) or float (the FP32-mixed form: neighbour data
// rounded to FP32 after FP64 differences of positions, velocities, densities and
// pressures, compensated sums), with positions, the EOS, the ghost assignment,
// the stages and the density MLS in FP64 in both forms.
#include "sph2d_internal.h"
#include "sph2d_device.h"
#include "sph2d_riemann.h"

namespace sph2d {

static inline int ms_grid(int n) { return (n + kBlock - 1) / kBlock; }

__device__ __forceinline__ double r_exp(double a) { return exp(a); }
__device__ __forceinline__ float r_exp(float a) { return expf(a); }

// Compensated (Kahan) sum in FP32, a plain sum in FP64.
template <typename T>
struct KSum {
  T s, c;
  __device__ __forceinline__ KSum() : s(T(0)), c(T(0)) {}
  __device__ __forceinline__ void add(T x) {
    if (sizeof(T) == 4) {
      const T y = x - c;
      const T t = s + y;
      c = (t - s) - y;
      s = t;
    } else {
      s = s + x;
    }
  }
  __device__ __forceinline__ T get() const { return s; }
};

// compute_kernel (MS Compute_Kernel.f90:251-262, model 4 = the improved Gaussian of
// support 3h) from the offset of the summing particle to its neighbour.
template <typename T>
struct MsKernel {
  T w, dwx, dwy;
};

template <typename T>
__device__ __forceinline__ MsKernel<T> ms_kernel(const DevParams& P, T dx, T dy, T r) {
  MsKernel<T> k;
  const T q = r * T(P.inv_h);
  const T eq = r_exp(-(q * q));
  k.w = (eq - T(P.e9)) * T(P.kfactor);
  const T t = eq * (T(-2.0) * q) * ((T(1) / r) * T(P.inv_h)) * T(P.kfactor);
  k.dwx = t * dx;
  k.dwy = t * dy;
  return k;
}

__device__ __forceinline__ bool ms_is_body(const Buffers& B, int s) { return B.g_init[B.name[s]] == -4; }

// ---------------------------------------------------------------------------
// MS Assign_boundary_particles.f90:611-760 (DistributeFixedBoundaryParticleOrNot = 1,
// ExternalBoundaryCondition = 0, MLS_order_Boundary = 1). One wavefront per local slot:
// the lanes test the candidates of the 3x3 cells in parallel (non-ghosts within 3 hbar,
// hbar = 1.2 smooth_length, no lower bound) and compact them into LDS in candidate
// order with what the sums need (W m/rho, p, rho, x, y, u, v, m); lane 0 sums them in
// that order (the Shepard weights Fai = W/sum W of the order-1 MLS, 3 or more
// candidates): P = sum p Fai + rho Fai g_ext . (x_i - x_j), v = sum v Fai, m = sum m Fai,
// g_ext = gravity (+ the body's acceleration for body particles). Walls: while
// 0 < t <= Shock_wave_release_time the characteristic correction of the normal velocity
// and pressure against (u, p) = (0, 0) for types -6 (y) and -7 (x), -8 keeping p = 0 as
// the reference's branches leave it; else the MLS values; then both components negated
// (no-slip). Body particles: the MLS pressure and the body's velocity. Then rho from
// the linear EOS and c = c0 (rho/rho0)^3 when m != 0, else the rest state.
// ---------------------------------------------------------------------------
constexpr int kMsWaves = 4, kMsList = 128;

__global__ void __launch_bounds__(64 * kMsWaves) k_ms_assign_walls(DevParams P, Buffers B) {
  __shared__ double sw[kMsWaves][kMsList], sp[kMsWaves][kMsList], sr[kMsWaves][kMsList];
  __shared__ double sx[kMsWaves][kMsList], sy[kMsWaves][kMsList], su[kMsWaves][kMsList];
  __shared__ double sv[kMsWaves][kMsList], sm[kMsWaves][kMsList];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * kMsWaves + wv;
  if (s >= B.counters[C_NLOCAL] || B.l_type[s] != -3) return;   // uniform across the wavefront
  const double2 xs = B.l_pos[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  const double rmax = 3.0 * P.wall_h;
  const uint64_t below = (1ull << lane) - 1ull;
  int cnt = 0;
  for (int d = 0; d < 3; ++d)
    for (int base = R.lo[d]; base < R.hi[d]; base += 64) {
      const int o = base + lane;
      bool ok = false;
      double w = 0.0;
      if (o < R.hi[d] && B.l_type[o] != -3) {
        const double2 xo = B.l_pos[o];
        const double px = xs.x - xo.x, py = xs.y - xo.y;
        const double r = sqrt(px * px + py * py);
        if (r <= rmax) {
          ok = true;
          const double q = r / P.wall_h;
          w = (exp(-(q * q)) - P.e9) * P.wall_kfactor;
        }
      }
      const uint64_t m = __ballot(ok);
      const int k = cnt + __popcll(m & below);
      if (ok && k < kMsList) {
        const double mo = B.l_mass[o], rho = B.l_rv[o].x;
        sw[wv][k] = w * mo / rho;
        sp[wv][k] = B.l_pc[o].x;
        sr[wv][k] = rho;
        const double2 xo = B.l_pos[o];
        sx[wv][k] = xo.x;
        sy[wv][k] = xo.y;
        const double2 vo = B.l_vel[o];
        su[wv][k] = vo.x;
        sv[wv][k] = vo.y;
        sm[wv][k] = mo;
      }
      cnt += __popcll(m);
    }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane != 0) return;
  const bool body = ms_is_body(B, s);
  const double gx = body ? 0.0 + P.cust_a : 0.0;               // Gravity_grad + Custom_accleration
  const double gy = body ? -P.g + 0.0 : -P.g;
  double iu = 0.0, iv = 0.0, ip = 0.0, m = 0.0;
  if (cnt > kMsList) {
    atomicOr(&B.counters[C_ERR], E_CELL);                      // more candidates than the list holds
  } else if (cnt >= 3) {
    double a = 0.0;
    for (int k = 0; k < cnt; ++k) a = a + sw[wv][k] * 1.0;
    const double ainv = 1.0 / a;
    for (int k = 0; k < cnt; ++k) {
      const double fai = 1.0 * (ainv * sw[wv][k]);
      const double dot = gx * (xs.x - sx[wv][k]) + gy * (xs.y - sy[wv][k]);
      ip = ip + sp[wv][k] * fai + sr[wv][k] * fai * dot;
      iu = iu + su[wv][k] * fai;
      iv = iv + sv[wv][k] * fai;
      m = m + sm[wv][k] * fai;
    }
  }
  double u, v, pr = 0.0;
  if (!body) {
    if (0.0 < P.t_eval && P.t_eval <= P.release_time) {
      const double irho = (ip - P.pb) / P.c0sq + P.rho0;
      const double xr = irho / P.rho0;
      const double tc = P.sqrt_c0sq * (xr * xr * xr);
      u = iu; v = iv;
      const int bt = B.g_btype[B.name[s]];
      if (bt == -6) {
        const double J2 = irho * tc * (iv - 0.0) + (ip - 0.0);
        const double J3 = -(irho * tc * (iv - 0.0)) + (ip - 0.0);
        v = 0.0 + (J2 - J3) / (2 * irho * tc);
        pr = 0.0 + 0.5 * (J2 + J3);
      } else if (bt == -7) {
        const double J2 = irho * tc * (iu - 0.0) + (ip - 0.0);
        const double J3 = -(irho * tc * (iu - 0.0)) + (ip - 0.0);
        u = 0.0 + (J2 - J3) / (2 * irho * tc);
        pr = 0.0 + 0.5 * (J2 + J3);
      }
    } else {
      u = iu; v = iv; pr = ip;
    }
    u = -u; v = -v;
  } else {
    pr = ip;
    u = P.cust_u; v = 0.0;
  }
  double rho, c;
  if (m != 0.0) {
    rho = (pr - P.pb) / P.c0sq + P.rho0;
    const double xr = rho / P.rho0;
    c = P.sqrt_c0sq * (xr * xr * xr);
  } else {
    u = 0.0; v = 0.0; pr = 0.0;
    rho = P.rho0; c = P.sqrt_c0sq;
  }
  B.l_vel[s] = make_double2(u, v);
  B.l_pc[s] = make_double2(pr, c);
  B.l_rv[s] = make_double2(rho, m / rho);
  B.l_mass[s] = m;
}

void launch_ms_assign_walls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_ms_assign_walls, dim3((cap_local + kMsWaves - 1) / kMsWaves), dim3(64 * kMsWaves), 0, s, P,
                     B);
}

// ---------------------------------------------------------------------------
// Pass A of an MS evaluation, every local particle: grad rho and grad v (and grad p
// for MUSCL-HLLC) with the kernel gradient (the identity renormalisation of MS
// Define_freesurface_and_Normailized_derivates.f90:165-172), the SPS strain and tau of
// the fluid (the block of MS Solver_Normalized_HLLC_Riemann_Delta_SPS.f90:95-235), and
// the XSPH velocity -epsilon sum m_j (v_i - v_j) / (rho_i + rho_j) W (every MS solver,
// e.g. MS/Solver_Normalized_MUSCL_HLLC_Riemann_Delta.f90:170-176). Records the accepted
// neighbours for the force pass.
// ---------------------------------------------------------------------------
template <typename T, bool MUSCL>
__device__ __forceinline__ void ms_a_body(const DevParams& P, Buffers& B, int blk_) {
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const double2 vs = B.l_vel[s];
  const double rhos = B.l_rv[s].x;
  const double press = B.l_pc[s].x;
  const int ts = B.l_type[s];
  const T eps = T(P.epsilon_xsph);
  KSum<T> grx, gry, gux, guy, gvx, gvy, gpx, gpy, sxx, sxy, syy, avx, avy;
  for_each_neighbour<kNbRecord>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    if (ts != 2 && B.l_type[o] != 2) return;     // no pair between two fixed ghosts
    const double2 po = B.l_pos[o];
    const T dx = T(ps.x - po.x), dy = T(ps.y - po.y);
    const T r = r_sqrt(dx * dx + dy * dy);
    const MsKernel<T> K = ms_kernel<T>(P, dx, dy, r);
    const double2 rvo = B.l_rv[o];
    const double2 vo = B.l_vel[o];
    const T Vo = T(rvo.y);
    const T vdx = T(vo.x - vs.x), vdy = T(vo.y - vs.y);       // -(v_i - v_j)
    const T rd = T(rvo.x - rhos);
    grx.add(rd * Vo * K.dwx);
    gry.add(rd * Vo * K.dwy);
    gux.add(vdx * Vo * K.dwx);
    guy.add(vdx * Vo * K.dwy);
    gvx.add(vdy * Vo * K.dwx);
    gvy.add(vdy * Vo * K.dwy);
    if (MUSCL) {
      const T pd = T(B.l_pc[o].x - press);
      gpx.add(pd * Vo * K.dwx);
      gpy.add(pd * Vo * K.dwy);
    }
    if (P.sps) {
      sxx.add(vdx * K.dwx * Vo);
      sxy.add((vdx * K.dwy + vdy * K.dwx) * Vo);
      syy.add(vdy * K.dwy * Vo);
    }
    const T mo = T(B.l_mass[o]);
    const T rsum = T(rhos + rvo.x);
    avx.add(-(eps * mo * (-vdx) / rsum * K.w));
    avy.add(-(eps * mo * (-vdy) / rsum * K.w));
  }, B.nb_list, B.nb_count, B.kstride);
  B.grad[s] = make_double2((double)grx.get(), (double)gry.get());
  B.avg[s] = make_double2((double)avx.get(), (double)avy.get());
  if (MUSCL) {
    B.mg[3 * s + 0] = make_double2((double)gpx.get(), (double)gpy.get());
    B.mg[3 * s + 1] = make_double2((double)gux.get(), (double)guy.get());
    B.mg[3 * s + 2] = make_double2((double)gvx.get(), (double)gvy.get());
  }
  double txx = 0.0, txy = 0.0, tyy = 0.0;
  if (ts == 2 && P.sps) {
    const double Sxx = (double)sxx.get(), Sxy = (double)sxy.get(), Syy = (double)syy.get();
    const double master = Sxx * Sxx + Syy * Syy + 0.0;
    const double last = Sxy * Sxy + 0.0 + 0.0;
    const double abs2 = 2 * master + last;
    const double t0 = P.sps_c012 * P.sps_dl;
    const double nut = (t0 * t0) * sqrt(abs2);
    const double kk = P.sps_c23 * nut * (Sxx + Syy);
    const double blin = P.sps_c23b * (P.sps_dl * P.sps_dl) * abs2;
    txx = (2 * nut * Sxx - kk - blin) * rhos;
    txy = (nut * Sxy) * rhos;
    tyy = (2 * nut * Syy - kk - blin) * rhos;
  }
  B.tau[3 * s + 0] = txx;
  B.tau[3 * s + 1] = txy;
  B.tau[3 * s + 2] = tyy;
  B.near_[s] = 0;                                  // Has_FreeSurface_Or_Not = 0
}

template <typename T, bool MUSCL>
__global__ void __launch_bounds__(kBlock) k_ms_a(DevParams P, Buffers B) {
  ms_a_body<T, MUSCL>(P, B, xcd_block());
}

// ---------------------------------------------------------------------------
// Force pass of an MS evaluation, fluid particles (the other particles' rates are
// never read). Per neighbour the star state of the pair in the reference's
// orientation (left = lower local index):
//  - MUSCL-HLLC (MS/Solver_Normalized_MUSCL_HLLC_Riemann_Delta.f90:200-330): p, u, v
//    reconstructed with the limited slopes, velocities in the pair frame, U* with each
//    side's impedance on its own velocity, the acoustic P*, rho* of the linear EOS,
//    Q factors, the HLLC fan sampled at the rotated mean velocity, the chosen state
//    rotated back, P* = c0^2 (rho_HLLC - rho0) + p_b;
//  - HLLC (MS/Solver_Normalized_HLLC_Riemann_Delta_SPS.f90:300-398): no
//    reconstruction, U* with the impedances crossed (the reference's form);
//  - modified linear (MS/Solver_Normalized_Modified_Linear_Riemann_Delta.f90:118-170),
//    which does not depend on the orientation;
// then the continuity with (v_i - v* + XSPH_i - XSPH_j), momentum, Monaghan viscosity,
// the Morris laplacian, the delta-alpha term, SPS (raw dW, particle_mass at the local
// index, temp_p carried from x into y), the delta-SPH density correction, the shifting
// sums, and on body-force evaluations the parts of Calculate_Body_Force.f90:51-142.
// Rate assembly of Single_step_compute.f90:253-283 (gravity as the external term).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void ms_hllc(T L1, T L2, T R1, T R2, T B1, T PL, T PR, T rL, T rR, T cL, T cR,
                                        bool muscl_form, const DevParams& P, T& H1, T& H2, T& Ps) {
  const T zsum = rL * cL + rR * cR;
  const T Us = muscl_form ? (L1 * cL * rL + R1 * cR * rR + PL - PR) / zsum
                          : (R1 * cL * rL + L1 * cR * rR + PL - PR) / zsum;
  const T Pst = (PR * cL * rL + PL * cR * rR - cL * rL * cR * rR * (R1 - L1)) / zsum;
  const T rst = (Pst - T(P.pb)) / T(P.c0sq) + T(P.rho0);
  T QL = T(1), QR = T(1);
  if (!(Pst <= PL)) QL = (rL > rst) ? r_sqrt((rst / rL) * (Pst - PL) / (rL - rst)) / cL : T(1);
  if (!(Pst <= PR)) QR = (rR > rst) ? r_sqrt((rst / rR) * (Pst - PR) / (rR - rst)) / cR : T(1);
  const T SL = L1 - cL * QL, SR = R1 + cR * QR, xot = B1;
  T rH;
  if (SL > xot) {
    rH = rL; H1 = L1; H2 = L2;
  } else if (Us >= xot && xot >= SL) {
    const T q = rL * ((SL - L1) / (SL - Us));
    rH = q; H1 = (Us * q) / rH; H2 = (L2 * q) / rH;
  } else if (SR >= xot && xot >= Us) {
    const T q = rR * ((SR - R1) / (SR - Us));
    rH = q; H1 = (Us * q) / rH; H2 = (R2 * q) / rH;
  } else {
    rH = rR; H1 = R1; H2 = R2;
  }
  Ps = T(P.c0sq) * (rH - T(P.rho0)) + T(P.pb);
}

template <typename T, int SOLVER>
__device__ __forceinline__ void ms_c_body(const DevParams& P, Buffers& B, int blk_) {
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const int ts = B.l_type[s];
  if (ts != 2) {
    B.drhodt[s] = 0.0;
    B.acc[s] = make_double2(0.0, 0.0);
    B.shift[s] = make_double2(0.0, 0.0);
    if (P.body_force) {
      B.bp[s] = make_double2(0.0, 0.0);
      B.bv[s] = make_double2(0.0, 0.0);
      B.near_body[s] = 0;
    }
    return;
  }
  constexpr bool kMuscl = SOLVER == SPH2D_SOLVER_MS_MUSCL_HLLC;
  constexpr bool kHllc = SOLVER == SPH2D_SOLVER_MS_MUSCL_HLLC || SOLVER == SPH2D_SOLVER_MS_HLLC;
  const double2 ps = B.l_pos[s];
  const double2 vs = B.l_vel[s];
  const double rhos_d = B.l_rv[s].x;
  const double2 pcs = B.l_pc[s];
  const double2 avs = B.avg[s];
  const double2 grs = B.grad[s];
  const T rhos = T(rhos_d), cs = T(pcs.y);
  const T txs = T(B.tau[3 * s]), tys = T(B.tau[3 * s + 1]), tzs = T(B.tau[3 * s + 2]);
  double2 gps = make_double2(0.0, 0.0), gus = gps, gvs = gps;
  if (kMuscl) { gps = B.mg[3 * s]; gus = B.mg[3 * s + 1]; gvs = B.mg[3 * s + 2]; }
  const T h = T(P.h), nu = T(P.nu);
  KSum<T> drho, corr, iax, iay, avx, avy, aix, aiy, spx, spy, lpx, lpy, shx, shy, bpx, bpy, bvx, bvy;
  int nb = 0;
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    const double2 po = B.l_pos[o];
    const T dx = T(ps.x - po.x), dy = T(ps.y - po.y);        // x_s - x_o
    const T r2 = dx * dx + dy * dy;
    const T r = r_sqrt(r2);
    const MsKernel<T> K = ms_kernel<T>(P, dx, dy, r);
    const double2 vo = B.l_vel[o];
    const double2 rvo = B.l_rv[o];
    const double2 pco = B.l_pc[o];
    const T rhoo = T(rvo.x), Vo = T(rvo.y), co = T(pco.y);
    const T mo = T(B.l_mass[o]);
    const T cbar = (cs + co) / T(2), rbar = (rhos + rhoo) / T(2);
    T vsx, vsy, Ps;
    if (kHllc) {
      const bool left = s < o;                                 // the pair (i, j) = (left, right)
      const double2 vi = left ? vs : vo, vj = left ? vo : vs;
      const double pi_ = left ? pcs.x : pco.x, pj_ = left ? pco.x : pcs.x;
      const T rL = left ? rhos : rhoo, rR = left ? rhoo : rhos;
      const T cL = left ? cs : co, cR = left ? co : cs;
      const T xjx = left ? -dx : dx, xjy = left ? -dy : dy;    // x_j - x_i
      const T ex = xjx / r, ey = xjy / r;                      // MUSCL: x_ij/|x_ij|; HLLC: -x_ij/|x_ij|
      T PL = T(pi_), PR = T(pj_);
      T Iu = T(vi.x), Iv = T(vi.y), Ju = T(vj.x), Jv = T(vj.y);
      if (kMuscl) {
        const double2* mgo = B.mg + 3 * o;
        const double2 gpo = mgo[0], guo = mgo[1], gvo = mgo[2];
        const double2 gpi = left ? gps : gpo, gpj = left ? gpo : gps;
        const double2 gui = left ? gus : guo, guj = left ? guo : gus;
        const double2 gvi = left ? gvs : gvo, gvj = left ? gvo : gvs;
        T li, lj;
        muscl_limiter<T>(T(0.5) * (T(gpi.x) * xjx + T(gpi.y) * xjy), -(T(0.5) * (T(gpj.x) * xjx + T(gpj.y) * xjy)),
                         T(pj_ - pi_), li, lj);
        PL = T(pi_) + li; PR = T(pj_) - lj;
        muscl_limiter<T>(T(0.5) * (T(gui.x) * xjx + T(gui.y) * xjy), -(T(0.5) * (T(guj.x) * xjx + T(guj.y) * xjy)),
                         T(vj.x - vi.x), li, lj);
        Iu = T(vi.x) + li; Ju = T(vj.x) - lj;
        muscl_limiter<T>(T(0.5) * (T(gvi.x) * xjx + T(gvi.y) * xjy), -(T(0.5) * (T(gvj.x) * xjx + T(gvj.y) * xjy)),
                         T(vj.y - vi.y), li, lj);
        Iv = T(vi.y) + li; Jv = T(vj.y) - lj;
      }
      const T vbx = (Iu + Ju) / T(2), vby = (Iv + Jv) / T(2);
      const T L1 = ex * Iu + ey * Iv, L2 = -ey * Iu + ex * Iv;   // Transfer_Matrix = [e1 e2; -e2 e1]
      const T R1 = ex * Ju + ey * Jv, R2 = -ey * Ju + ex * Jv;
      const T B1 = ex * vbx + ey * vby;
      T H1, H2;
      ms_hllc<T>(L1, L2, R1, R2, B1, PL, PR, rL, rR, cL, cR, kMuscl, P, H1, H2, Ps);
      vsx = ex * H1 + (-ey) * H2;                              // TRANSPOSE(Transfer_Matrix)
      vsy = ey * H1 + ex * H2;
    } else {
      const T ex = -dx / r, ey = -dy / r;                      // from s to o
      const T UL = T(vs.x) * ex + T(vs.y) * ey;
      const T UR = T(vo.x) * ex + T(vo.y) * ey;
      const T Ub = (UL + UR) / T(2);
      const T Pbar = T(pcs.x + pco.x) / T(2);
      const T vbx = T(vs.x + vo.x) / T(2), vby = T(vs.y + vo.y) / T(2);
      const T Us = Ub + T(0.5) * T(pcs.x - pco.x) / (rbar * T(P.sqrt_c0sq));
      vsx = (Us - Ub) * ex + vbx;
      vsy = (Us - Ub) * ey + vby;
      const T du = UL - UR;
      T bR = T(3) * (du > T(0) ? du : T(0));
      bR = bR < cbar ? bR : cbar;
      Ps = Pbar + T(0.5) * bR * rbar * du;
    }
    const double2 avo = B.avg[o];
    const T qx = T(vs.x) - vsx + T(avs.x) - T(avo.x);
    const T qy = T(vs.y) - vsy + T(avs.y) - T(avo.y);
    drho.add(T(2) * rhos * Vo * (qx * K.dwx + qy * K.dwy));
    const T rij = T(1) / (rhos * rhoo);
    iax.add(mo * (T(-2) * Ps * K.dwx * rij));
    iay.add(mo * (T(-2) * Ps * K.dwy * rij));
    const T vijx = T(vs.x - vo.x), vijy = T(vs.y - vo.y);
    const T vx = vijx * dx + vijy * dy;
    if (vx < T(0)) {
      const T fh = T(P.fai) * h;
      const T phi = h * vx / (r2 + fh * fh);
      const T Pi = (-(T(P.alpha) * cbar * phi) + T(P.beta) * (phi * phi)) / rbar;
      avx.add(mo * (-Pi * K.dwx));
      avy.add(mo * (-Pi * K.dwy));
    }
    const T Rd = dx * K.dwx + dy * K.dwy;
    lpx.add(Vo * T(2) * nu * Rd / r2 * vijx);
    lpy.add(Vo * T(2) * nu * Rd / r2 * vijy);
    const T tmp = T(P.delta_alpha) * h * cbar * vx / r2;
    aix.add(tmp * K.dwx * T(P.rho0) / rhos * Vo);
    aiy.add(tmp * K.dwy * T(P.rho0) / rhos * Vo);
    if (P.sps) {
      const T gmo = T(B.g_mass[o]);                            // particle_mass(j), j a local index
      const T txo = T(B.tau[3 * o]), tyo = T(B.tau[3 * o + 1]), tzo = T(B.tau[3 * o + 2]);
      T tp = T(0) + (txs + txo) * K.dwx;
      tp = tp + (tys + tyo) * K.dwy;
      spx.add(gmo * tp * rij);
      tp = tp + (tzs + tzo) * K.dwy + (tys + tyo) * K.dwx;
      spy.add(gmo * tp * rij);
    }
    {
      const double2 gro = B.grad[o];
      const T psi = T(rvo.x - rhos_d) - T(0.5) * (T(grs.x + gro.x) * (-dx) + T(grs.y + gro.y) * (-dy));
      const T ti = ((-dx) * K.dwx + (-dy) * K.dwy) / r2;
      corr.add(T(2) * T(P.delta) * h * cbar * psi * ti * Vo);
    }
    {
      const T wr = K.w * T(P.inv_wdx);
      const T f = T(P.shift_beta) * T(P.h2) * (T(1) + T(P.shift_R) * (wr * wr * wr * wr));
      shx.add(-(f * K.dwx * Vo));
      shy.add(-(f * K.dwy * Vo));
    }
    if (P.body_force && ms_is_body(B, o)) {                    // Calculate_Body_Force.f90:80-101
      nb = 1;
      const T psum = T(pcs.x + pco.x);
      bpx.add(-(Vo * psum * K.dwx));
      bpy.add(-(Vo * psum * K.dwy));
      bvx.add(rhos * Vo * T(2) * nu * Rd / r2 * vijx);
      bvy.add(rhos * Vo * T(2) * nu * Rd / r2 * vijy);
    }
  }, B.nb_list, B.nb_count, B.kstride);
  B.drhodt[s] = (double)drho.get() + (double)corr.get();
  const double ax = ((((double)iax.get() + 0.0) + (double)avx.get()) + (double)aix.get()) + (double)spx.get() +
                    (double)lpx.get();
  const double ay = ((((double)iay.get() + (0.0 - P.g)) + (double)avy.get()) + (double)aiy.get()) +
                    (double)spy.get() + (double)lpy.get();
  B.acc[s] = make_double2(ax, ay);
  B.shift[s] = make_double2((double)shx.get(), (double)shy.get());
  if (P.body_force) {
    B.bp[s] = make_double2((double)bpx.get(), (double)bpy.get());
    B.bv[s] = make_double2((double)bvx.get(), (double)bvy.get());
    B.near_body[s] = nb;
  }
}

template <typename T, int SOLVER>
__global__ void __launch_bounds__(kBlock) k_ms_c(DevParams P, Buffers B) {
  ms_c_body<T, SOLVER>(P, B, xcd_block());
}

template <typename T>
static void launch_ms_eval_t(hipStream_t s, const DevParams& P, Buffers& B, int g) {
  if (P.solver == SPH2D_SOLVER_MS_MUSCL_HLLC) {
    hipLaunchKernelGGL((k_ms_a<T, true>), dim3(g), dim3(kBlock), 0, s, P, B);
    hipLaunchKernelGGL((k_ms_c<T, SPH2D_SOLVER_MS_MUSCL_HLLC>), dim3(g), dim3(kBlock), 0, s, P, B);
  } else if (P.solver == SPH2D_SOLVER_MS_HLLC) {
    hipLaunchKernelGGL((k_ms_a<T, false>), dim3(g), dim3(kBlock), 0, s, P, B);
    hipLaunchKernelGGL((k_ms_c<T, SPH2D_SOLVER_MS_HLLC>), dim3(g), dim3(kBlock), 0, s, P, B);
  } else {
    hipLaunchKernelGGL((k_ms_a<T, false>), dim3(g), dim3(kBlock), 0, s, P, B);
    hipLaunchKernelGGL((k_ms_c<T, SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN>), dim3(g), dim3(kBlock), 0, s, P, B);
  }
}

void launch_ms_eval(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  const int g = ms_grid(cap_local);
  if (P.mixed) launch_ms_eval_t<float>(s, P, B, g);
  else launch_ms_eval_t<double>(s, P, B, g);
}

// Body force totals (Single_step_compute.f90:219-249): sum over the near-body fluid of
// V (pressure part), V (viscous part) and V (both + rho g); per-block partials in a
// fixed tree, then one block in block order, so the totals are deterministic.
__global__ void __launch_bounds__(kBlock) k_ms_body_part(DevParams P, Buffers B) {
  __shared__ double lds[6][kBlock];
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  double f[6] = {0, 0, 0, 0, 0, 0};
  if (s < B.counters[C_NLOCAL] && B.l_type[s] == 2 && B.near_body[s] == 1) {
    const double V = B.l_rv[s].y, rho = B.l_rv[s].x;
    const double2 bp = B.bp[s], bv = B.bv[s];
    const double ex = 0.0 + rho * 0.0, ey = 0.0 + rho * (-P.g);
    f[0] = V * bp.x; f[1] = V * bp.y; f[2] = V * bv.x; f[3] = V * bv.y;
    f[4] = V * (bp.x + bv.x + ex); f[5] = V * (bp.y + bv.y + ey);
  }
  for (int k = 0; k < 6; ++k) lds[k][threadIdx.x] = f[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 6; ++k) lds[k][threadIdx.x] = lds[k][threadIdx.x] + lds[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 6; ++k) B.bpart[6 * blockIdx.x + k] = lds[k][0];
}

// One wavefront per component: lane l sums the partials l, l + 64, ... in order,
// then the 64 lane sums fold pairwise (xor butterfly): a fixed order, so the
// result does not depend on the launch.
__global__ void __launch_bounds__(384) k_ms_body_final(Buffers B, int nb, double* out) {
  const int k = threadIdx.x >> 6, lane = threadIdx.x & 63;
  double a = 0.0;
  for (int b = lane; b < nb; b += 64) a = a + B.bpart[6 * b + k];
  for (int off = 32; off > 0; off >>= 1) a = a + __shfl_xor(a, off);
  if (lane == 0) out[k] = a;
}

void launch_ms_body_sum(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, double* out6) {
  const int g = ms_grid(cap_local);
  hipLaunchKernelGGL(k_ms_body_part, dim3(g), dim3(kBlock), 0, s, P, B);
  hipLaunchKernelGGL(k_ms_body_final, dim3(1), dim3(384), 0, s, B, g, out6);
}

// ---------------------------------------------------------------------------
// MS predictor-corrector (SPH_Time_Iteration.f90:313-347, Iteration_Method = 2), fluid
// particles, rates read by the current local index (the previous step's, or the first
// evaluation's on step 1):
//   predictor: before_* = state, temp_* = rates and velocity;
//     rho = rho_b + drho dt, x = x_b + dt (v + XSPH) + a dt^2 / 2, v = v_b + dt a
//   corrector: rho = rho_b + (drho + temp_drho) dt / 2,
//     x = x_b + dt (v + temp_v) / 2 + XSPH dt + a dt^2 / 2, v = v_b + dt (a + temp_a) / 2
// (the temp velocity lives in the first RK stage slab).
// ---------------------------------------------------------------------------
__global__ void k_ms_stage(DevParams P, Buffers B, int kind) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const double dt = P.dt;
  const int t = B.l_type[s];
  const double drho = B.drhodt[s];
  const double2 a = B.acc[s];
  const double2 avg = B.avg[s];
  double2 x = B.l_pos[s], v = B.l_vel[s];
  double rho = B.l_rv[s].x;
  if (kind == MS_PREDICT) {
    B.b_rho[s] = rho; B.b_vel[s] = v; B.b_pos[s] = x;
    B.t_drho[s] = drho; B.t_acc[s] = a; B.k_vel[s] = v;
    if (t != 2) return;
    rho = rho + drho * dt;
    x.x = x.x + dt * (v.x + avg.x) + 0.5 * a.x * (dt * dt);
    x.y = x.y + dt * (v.y + avg.y) + 0.5 * a.y * (dt * dt);
    v.x = v.x + dt * a.x;
    v.y = v.y + dt * a.y;
    B.l_pos[s] = x;
    B.l_pos32[s] = make_float2((float)x.x, (float)x.y);
    B.l_vel[s] = v;
    B.l_rv[s] = make_double2(rho, B.l_rv[s].y);
  } else {
    if (t != 2) return;
    const double2 bx = B.b_pos[s], bv = B.b_vel[s], ta = B.t_acc[s], tv = B.k_vel[s];
    rho = B.b_rho[s] + 0.5 * (drho + B.t_drho[s]) * dt;
    x.x = bx.x + 0.5 * dt * (v.x + tv.x) + avg.x * dt + 0.5 * a.x * (dt * dt);
    x.y = bx.y + 0.5 * dt * (v.y + tv.y) + avg.y * dt + 0.5 * a.y * (dt * dt);
    v.x = bv.x + 0.5 * dt * (a.x + ta.x);
    v.y = bv.y + 0.5 * dt * (a.y + ta.y);
    B.l_pos[s] = x;
    B.l_pos32[s] = make_float2((float)x.x, (float)x.y);
    B.l_vel[s] = v;
    B.l_rv[s] = make_double2(rho, B.l_rv[s].y);
  }
  const int cell = B.l_cell[s];
  const int cx = cell / P.chain_ny, cy = cell - cx * P.chain_ny;
  const double x0 = P.chain_ox + cx * P.chain_d, y0 = P.chain_oy + cy * P.chain_d;
  const double m = P.rebin_m;
  if (x.x < x0 - m || x.x > x0 + P.chain_d + m || x.y < y0 - m || x.y > y0 + P.chain_d + m)
    atomicOr(&B.counters[C_REBIN], 1);
}

void launch_ms_stage(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int kind) {
  hipLaunchKernelGGL(k_ms_stage, dim3(ms_grid(cap_local)), dim3(kBlock), 0, s, P, B, kind);
}

// ---------------------------------------------------------------------------
// MS Reinitialize_particle_density.f90:29-73 with RefreshDensity_MLS_Order = 3: per fluid
// particle, its candidates of the 3x3 cells within 3h (no lower bound: the particle
// itself counts), W V_j, the order-3 MLS of MS/SPH_Subroutine_Module.f90:2565-2792
// (basis (1, (x_j - x_i)/h, (y_j - y_i)/h), alpha^3 on the diagonal, the cofactor inverse
// of matrix_inversion_less_three, Fai = row 1 of A^-1 B; Shepard when det = 0);
// rho = sum rho_j Fai_j, a NaN keeping the old value; other particles keep theirs.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_ms_mls(DevParams P, Buffers B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const double rho_old = B.l_rv[s].x;
  if (B.l_type[s] != 2) { B.mls_rho[s] = rho_old; return; }
  const double2 xs = B.l_pos[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  double a00 = 0, a01 = 0, a02 = 0, a11 = 0, a12 = 0, a22 = 0, sw = 0;
  for (int d = 0; d < 3; ++d)
    for (int o = R.lo[d]; o < R.hi[d]; ++o) {
      const double2 xo = B.l_pos[o];
      const double px = xs.x - xo.x, py = xs.y - xo.y;
      const double r = sqrt(px * px + py * py);
      if (!(r <= P.three_h)) continue;
      const double q = r / P.h;
      const double w = (exp(-(q * q)) - P.e9) * P.kfactor * B.l_rv[o].y;
      const double p2 = (xo.x - xs.x) / P.h, p3 = (xo.y - xs.y) / P.h;
      a00 = a00 + 1.0 * w * 1.0; a01 = a01 + 1.0 * w * p2; a02 = a02 + 1.0 * w * p3;
      a11 = a11 + p2 * w * p2; a12 = a12 + p2 * w * p3; a22 = a22 + p3 * w * p3;
      sw = sw + w;
    }
  const double al3 = (P.mls_alpha * P.mls_alpha) * P.mls_alpha;
  const double a1 = a00 + al3, b1 = a01, c1 = a02;
  const double a2 = a01, b2 = a11 + al3, c2 = a12;
  const double a3 = a02, b3 = a12, c3 = a22 + al3;
  const double det = (a1 * (b2 * c3 - c2 * b3) - a2 * (b1 * c3 - c1 * b3) + a3 * (b1 * c2 - c1 * b2));
  double i0 = 0, i1 = 0, i2 = 0;
  const bool ok = det != 0;
  if (ok) {
    i0 = (b2 * c3 - c2 * b3) / det;
    i1 = (b3 * c1 - c3 * b1) / det;
    i2 = (b1 * c2 - c1 * b2) / det;
  } else {
    atomicAdd(&B.counters[C_MLSFB], 1);
  }
  double rho = 0.0;
  for (int d = 0; d < 3; ++d)
    for (int o = R.lo[d]; o < R.hi[d]; ++o) {
      const double2 xo = B.l_pos[o];
      const double px = xs.x - xo.x, py = xs.y - xo.y;
      const double r = sqrt(px * px + py * py);
      if (!(r <= P.three_h)) continue;
      const double q = r / P.h;
      const double w = (exp(-(q * q)) - P.e9) * P.kfactor * B.l_rv[o].y;
      const double p2 = (xo.x - xs.x) / P.h, p3 = (xo.y - xs.y) / P.h;
      const double fai = ok ? (i0 * (1.0 * w) + i1 * (p2 * w) + i2 * (p3 * w)) : (w / sw);
      rho = rho + B.l_rv[o].x * fai;
    }
  B.mls_rho[s] = isnan(rho) ? rho_old : rho;
}

void launch_ms_mls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_ms_mls, dim3(ms_grid(cap_local)), dim3(kBlock), 0, s, P, B);
}

// Body_Motion_Solver.f90:245-271: the body particles follow the change of the centre.
__global__ void k_ms_body_move(Buffers B, int n_total, double ddx, double ddy) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total || B.g_init[g] != -4) return;
  const double2 x = B.g_pos[g];
  B.g_pos[g] = make_double2(x.x + ddx, x.y + ddy);
}

void launch_ms_body_move(hipStream_t s, const DevParams& P, Buffers& B, int n_total, double ddx, double ddy) {
  (void)P;
  hipLaunchKernelGGL(k_ms_body_move, dim3(ms_grid(n_total)), dim3(kBlock), 0, s, B, n_total, ddx, ddy);
}

// MS Reorder_Particle_Index.f90:80-95: fluid outside [0, box_x] x [0, box_y] (lower bounds
// the REAL(4) -1.0E-10) becomes ill (120) and leaves the subdomain.
__global__ void k_ms_ill(DevParams P, Buffers B, int cap_fluid) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= cap_fluid || g >= B.counters[C_NFLUID] || B.g_type[g] != 2) return;
  const double2 x = B.g_pos[g];
  const double lo = (double)-1.0E-10f;
  if (x.x < lo || x.x > P.box_x || x.y < lo || x.y > P.box_y) B.g_type[g] = 120;
}

void launch_ms_ill(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid) {
  hipLaunchKernelGGL(k_ms_ill, dim3(ms_grid(cap_fluid)), dim3(kBlock), 0, s, P, B, cap_fluid);
}

// ---------------------------------------------------------------------------
// Host: the fixed ghosts of MS/Initial_Particle_From_First_Step.f90:178-620. Per side
// a run from a start point towards an end point: the first layer's count k comes from
// walking in steps of gdx until within gdx/2 of the end; a run places k - 1 points at
// the spacing |end - start| / (k - 1); k grows by 2 per layer outside (tank walls:
// bottom, right, top, left) and shrinks by 2 inside the square (bottom, right, top,
// left, layer i at (i - 1/2) gdy from its edges). Tank walls: -6 in the box's x range of
// the bottom and top, -8 in its y range of the sides, else -7; body: the first point of
// each run -7, the rest -6 (bottom, top) or -8 (sides).
// ---------------------------------------------------------------------------
static int ms_run(double x0, double y0, double x1, double y1, double sx, double sy, double gdx, int& k, bool first,
                  int dk, double* x, double* y, int at) {
  if (first) {
    int j = 0;
    for (;;) {
      ++j;
      const double px = x0 + sx * ((j - 1) * gdx), py = y0 + sy * ((j - 1) * gdx);
      if (std::sqrt((x1 - px) * (x1 - px) + (y1 - py) * (y1 - py)) <= gdx / 2.0) break;
    }
    k = j;
  }
  const double dist = std::sqrt((x1 - x0) * (x1 - x0) + (y1 - y0) * (y1 - y0)) / (k - 1);
  int n = 0;
  for (int j = 1; j <= k - 1; ++j, ++n)
    if (x) {
      x[at + n] = x0 + sx * ((j - 1) * dist);
      y[at + n] = y0 + sy * ((j - 1) * dist);
    }
  k = k + dk;
  return n;
}

int ms_ghost_layout(double bx, double by, double gdx, double gdy, int L, double cx, double cy, double len,
                    double* x, double* y, int* bt, int* init) {
  int n = 0, k = 0, n0 = 0;
  // b +/- (i-1) gdy +/- gdy/2, left to right as the reference writes it
  auto up = [&](double b, int i) { return b + (i - 1) * gdy + gdy / 2.0; };
  auto dn = [&](double b, int i) { return b - (i - 1) * gdy - gdy / 2.0; };
  n0 = n;
  for (int i = 1; i <= L; ++i) n += ms_run(dn(0, i), dn(0, i), up(bx, i), dn(0, i), 1, 0, gdx, k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= x[q] && x[q] <= bx) ? -6 : -7; init[q] = -3; }
  n0 = n;
  for (int i = 1; i <= L; ++i) n += ms_run(up(bx, i), dn(0, i), up(bx, i), up(by, i), 0, 1, gdx, k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= y[q] && y[q] <= by) ? -8 : -7; init[q] = -3; }
  n0 = n;
  for (int i = 1; i <= L; ++i) n += ms_run(up(bx, i), up(by, i), dn(0, i), up(by, i), -1, 0, gdx, k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= x[q] && x[q] <= bx) ? -6 : -7; init[q] = -3; }
  n0 = n;
  for (int i = 1; i <= L; ++i) n += ms_run(dn(0, i), up(by, i), dn(0, i), dn(0, i), 0, -1, gdx, k, i == 1, 2, x, y, n);
  if (x) for (int q = n0; q < n; ++q) { bt[q] = (0.0 <= y[q] && y[q] <= by) ? -8 : -7; init[q] = -3; }
  const double hl = 0.5 * len;
  const double sg[4][2] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
  for (int side = 0; side < 4; ++side) {
    n0 = n;
    for (int i = 1; i <= L; ++i) {
      const double lo_x = up(cx - hl, i), hi_x = dn(cx + hl, i), lo_y = up(cy - hl, i), hi_y = dn(cy + hl, i);
      double x0, y0, x1, y1;
      if (side == 0) { x0 = lo_x; y0 = lo_y; x1 = hi_x; y1 = lo_y; }
      else if (side == 1) { x0 = hi_x; y0 = lo_y; x1 = hi_x; y1 = hi_y; }
      else if (side == 2) { x0 = hi_x; y0 = hi_y; x1 = lo_x; y1 = hi_y; }
      else { x0 = lo_x; y0 = hi_y; x1 = lo_x; y1 = lo_y; }
      const int first = n;
      n += ms_run(x0, y0, x1, y1, sg[side][0], sg[side][1], gdx, k, i == 1, -2, x, y, n);
      if (x && first < n) bt[first] = -7;
      if (x)
        for (int q = first + 1; q < n; ++q) bt[q] = (side == 0 || side == 2) ? -6 : -8;
    }
    if (x) for (int q = n0; q < n; ++q) init[q] = -4;
  }
  return n;
}

}  // namespace sph2d
