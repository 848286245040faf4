// sph2d_eval64.hip -- one force evaluation of single_step_compute.f90 in FP64.
//
// Gather form: every particle sums over its own neighbours, so no pair list is
// stored and no float atomics are used. Each sum runs in a fixed order
// (neighbour columns cx-1, cx, cx+1, rows ascending, particles ascending), so the
// results are deterministic run to run.
//
// Pair terms are written for the summing particle s and its neighbour o with
// P = x_s - x_o and D = dW(P). The reference evaluates every pair once for
// i < j and scatters to both sides; its j-side formulas with x_ij = -P and
// dwdx = -D reduce to the i-side formulas written here (the two sign flips
// cancel in every term, and the Riemann star state v*, P* is the same vector
// from either side), so each term equals the reference's up to rounding.
// Built with -ffp-contract=off; reciprocals replace repeated divisions.
#include "sph2d_internal.h"
#include "sph2d_device.h"
#include "sph2d_riemann.h"

namespace sph2d {

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

// Kernel gradient and value from the neighbour offset (compute_kernel.f90:123-129).
struct KernelTerm {
  double w, dwx, dwy;
};

__device__ __forceinline__ KernelTerm kernel_term(const DevParams& P, double dx, double dy, double r,
                                                  double inv_r) {
  KernelTerm k;
  const double q = r * P.inv_h;
  const double eq = exp(-(q * q));
  k.w = (eq - P.e9) * P.kfactor;
  const double t = eq * (-2.0 * q) * (inv_r * P.inv_h) * P.kfactor;
  k.dwx = t * dx;
  k.dwy = t * dy;
  return k;
}

// ---------------------------------------------------------------------------
// pass A: volume fraction, sum dW, SPS strain -> tau, grad rho
// Define_freesurface_and_Normailized_derivates.f90:79-131,141-159 and
// Solver_Normalized_Modified_Linear_Riemann_Delta_SPS.f90:144-296,508-526.
// MUSCL also sums grad p, grad u, grad v the same way
// (Solver_Normalized_MUSCL_Linear_Riemann_Delta_SPS.f90:145-185).
// ---------------------------------------------------------------------------
// RENORM (a mode of this library): the dyadic tensor T = sum (x_o - x_s) (x) grad W V_o is
// formed and inverted (matrix_inversion_less_three; singular: the identity, counted),
// and the strain and grad rho use T^-1 grad W. T^-1 is not known inside the walk, so the
// walk sums the raw gradients' outer products and T^-1 is applied after it (the same
// sums in exact arithmetic, rounded differently).
template <bool MUSCL, bool RENORM = false>
__device__ __forceinline__ void k_eval_a_body(const DevParams& P, Buffers& B, int blk_, int nblk_) {
  __shared__ NbStage st;
  __shared__ float2 lds32[kStageCap];
  __shared__ double2 lds64[kStageCap];
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  stage_candidates(P, B.cell_start, B.l_cell, B.l_pos32, B.l_pos, n, st, lds32, lds64, blk_);
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const double2 vs = B.l_vel[s];
  const double rhos = B.l_rv[s].x;
  const double pres = B.l_pc[s].x;
  const int ts = B.l_type[s];
  double fra = 0.0, sdx = 0.0, sdy = 0.0, Sxx = 0.0, Sxy = 0.0, Syy = 0.0, grx = 0.0, gry = 0.0;
  double gpx = 0.0, gpy = 0.0, gux = 0.0, guy = 0.0, gvx = 0.0, gvy = 0.0;
  double t00 = 0.0, t01 = 0.0, t10 = 0.0, t11 = 0.0, gxx = 0.0, gxy = 0.0, gyx = 0.0, gyy = 0.0;
  for_each_neighbour<kNbRecord>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    if (ts != 2 && B.l_type[o] != 2) return;     // no pair between two non-fluid particles
    const double2 po = B.l_pos[o];
    const double dx = ps.x - po.x, dy = ps.y - po.y;
    const double r = sqrt(dx * dx + dy * dy);
    const KernelTerm K = kernel_term(P, dx, dy, r, 1.0 / r);
    const double2 rvo = B.l_rv[o];
    const double2 vo = B.l_vel[o];
    const double Vo = rvo.y;
    fra = fra + Vo * P.inv_sv;
    sdx = sdx + K.dwx * Vo;
    sdy = sdy + K.dwy * Vo;
    const double vdx = vo.x - vs.x, vdy = vo.y - vs.y;   // -(v_i - v_j)
    if (RENORM) {                                  // Define_freesurface...f90:141-159 (x_j - x_i = -dx)
      t00 = t00 + (-dx) * K.dwx * Vo;
      t01 = t01 + (-dx) * K.dwy * Vo;
      t10 = t10 + (-dy) * K.dwx * Vo;
      t11 = t11 + (-dy) * K.dwy * Vo;
      gxx = gxx + vdx * K.dwx * Vo;
      gxy = gxy + vdx * K.dwy * Vo;
      gyx = gyx + vdy * K.dwx * Vo;
      gyy = gyy + vdy * K.dwy * Vo;
    } else {                                       // the identity renormalisation
      Sxx = Sxx + vdx * K.dwx * Vo;
      Sxy = Sxy + (vdx * K.dwy + vdy * K.dwx) * Vo;
      Syy = Syy + vdy * K.dwy * Vo;
    }
    const double rd = (rvo.x - rhos) * Vo;
    grx = grx + rd * K.dwx;
    gry = gry + rd * K.dwy;
    if (MUSCL) {
      const double pd = (B.l_pc[o].x - pres) * Vo;
      gpx = gpx + pd * K.dwx;
      gpy = gpy + pd * K.dwy;
      const double ud = vdx * Vo, vd = vdy * Vo;
      gux = gux + ud * K.dwx;
      guy = guy + ud * K.dwy;
      gvx = gvx + vd * K.dwx;
      gvy = gvy + vd * K.dwy;
    }
  }, B.nb_list, B.nb_count, B.kstride, &st, lds32, lds64);
  B.fra_v[s] = fra;
  B.sdw[s] = make_double2(sdx, sdy);
  if (RENORM) {
    // matrix_inversion_less_three (function_module.f90:157-249) for fluid particles;
    // a singular tensor keeps the identity. i00 i01 / i10 i11 = T^-1 (rows, columns).
    double i00 = 1.0, i01 = 0.0, i10 = 0.0, i11 = 1.0;
    if (ts == 2) {
      const double det = (t00 * t11 - t01 * t10);
      if (det != 0) {
        i00 = t11 / det; i01 = -t01 / det; i10 = -t10 / det; i11 = t00 / det;
      } else {
        atomicAdd(&B.counters[C_NSING], 1);
      }
    }
    double* ti = B.tinv + 4 * (size_t)s;
    ti[0] = i00; ti[1] = i10; ti[2] = i01; ti[3] = i11;   // column-major, as T_inv(i, :, :)
    // strain and grad rho with T^-1 grad W (Solver...SPS.f90:144-182, :508-526)
    Sxx = i00 * gxx + i01 * gxy;
    Syy = i10 * gyx + i11 * gyy;
    Sxy = (i10 * gxx + i11 * gxy) + (i00 * gyx + i01 * gyy);
    const double rx = grx, ry = gry;
    grx = i00 * rx + i01 * ry;
    gry = i10 * rx + i11 * ry;
  }
  B.grad[s] = make_double2(grx, gry);
  if (MUSCL) {
    B.mg[3 * s + 0] = make_double2(gpx, gpy);
    B.mg[3 * s + 1] = make_double2(gux, guy);
    B.mg[3 * s + 2] = make_double2(gvx, gvy);
  }
  double txx = 0.0, txy = 0.0, tyy = 0.0;
  if (ts == 2 && P.sps) {                          // Solver...SPS.f90:256-296
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
  const int sf = (ts == 2 && fra <= P.fra_limit) ? 1 : 0;
  B.surf[s] = sf;
  B.near_[s] = 0;
  if (sf) atomicAdd(&B.counters[C_NSURF], 1);
}

template <bool MUSCL, bool RENORM>
__global__ void __launch_bounds__(kBlock) k_eval_a(DevParams P, Buffers B) {
  k_eval_a_body<MUSCL, RENORM>(P, B, xcd_block(), gridDim.x);
}

// Surface normal and geometric test (Define_freesurface...f90:295-376), surface particles only.
__device__ __forceinline__ void k_eval_b1_body(const DevParams& P, Buffers& B, int blk_, int nblk_) {
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  if (B.surf[s] != 1) { B.nrm[s] = make_double2(0.0, 0.0); return; }
  const double2 sd = B.sdw[s];
  double n1 = 1.0 * sd.x + 0.0 * sd.y;           // identity inverse: matmul(I, sum_dwdx)
  double n2 = 0.0 * sd.x + 1.0 * sd.y;
  if (P.renorm) {                                // matmul(T^-1, sum_dwdx)
    const double* ti = B.tinv + 4 * (size_t)s;
    n1 = ti[0] * sd.x + ti[2] * sd.y;
    n2 = ti[1] * sd.x + ti[3] * sd.y;
  }
  double len = 0.0;
  len = len + n1 * n1;
  len = len + n2 * n2;
  len = sqrt(len);
  bool keep = true;
  double nx = 0.0, ny = 0.0;
  if (isnan(len) || len == 0.0) {
    keep = false;
  } else {
    nx = -n1 / len;
    ny = -n2 / len;
    const double2 ps = B.l_pos[s];
    const double vx = ps.x + P.h * nx, vy = ps.y + P.h * ny;
    const double pi = 3.14159265358979324;
    // The reference tests every fluid candidate of the 3x3 cells; either test can only
    // fire within 2h (JT < h around x + h n, or JI < 1.414 h), so the accepted pair
    // partners (within 3h) hold every candidate that matters, and "any fires" does not
    // depend on the order.
    for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
      if (!keep || B.l_type[o] != 2) return;
      const double2 po = B.l_pos[o];
      const double pdx = po.x - ps.x, pdy = po.y - ps.y;
      const double JI = sqrt(pdx * pdx + pdy * pdy);
      const double Nd = nx * pdx + ny * pdy;
      const double qx = po.x - vx, qy = po.y - vy;
      const double JT = sqrt(qx * qx + qy * qy);
      if (JI >= P.surf_lim && JT < P.h) keep = false;
      else if (JI < P.surf_lim && acos(Nd / JI) <= (pi / 4.0)) keep = false;
    }, B.nb_list, B.nb_count, B.kstride);
  }
  B.nrm[s] = make_double2(nx, ny);
  if (!keep) {
    B.surf[s] = 0;
    atomicSub(&B.counters[C_NSURF], 1);
  }
}

__global__ void __launch_bounds__(kBlock) k_eval_b1(DevParams P, Buffers B) {
  k_eval_b1_body(P, B, xcd_block(), gridDim.x);
}

// Near-surface flag (Define_freesurface...f90:418-440): any pair partner is a surface particle.
__device__ __forceinline__ void k_eval_b2_body(const DevParams& P, Buffers& B, int blk_, int nblk_) {
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  if (B.counters[C_NSURF] == 0) return;            // near_ was zeroed by pass A
  const int ts = B.l_type[s];
  int nr = 0;
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, B.l_pos[s], B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    if (B.surf[o] == 1 && (ts == 2 || B.l_type[o] == 2)) nr = 1;
  }, B.nb_list, B.nb_count, B.kstride);
  B.near_[s] = nr;
}

__global__ void __launch_bounds__(kBlock) k_eval_b2(DevParams P, Buffers B) {
  k_eval_b2_body(P, B, xcd_block(), gridDim.x);
}

// ---------------------------------------------------------------------------
// pass C: Riemann continuity/momentum, artificial viscosity, laplacian,
// delta-SPH internal, SPS, XSPH, delta-SPH density correction, shift sums.
// Solver_Normalized_Modified_Linear_Riemann_Delta_SPS.f90:300-635,
// Solver_Normalized_Linear_Riemann_Delta_SPS.f90:362-371,
// Solver_Normalized_HLLC_Linear_Riemann_Delta_SPS.f90:362-483,
// Solver_Normalized_MUSCL_Linear_Riemann_Delta_SPS.f90:464-518,
// single_step_compute.f90:152-171, Shift_particle_position.f90:78-103.
// DEBUG keeps the five acceleration components apart and writes them.
// SOLVER is one of SPH2D_SOLVER_*: each variant is its own kernel.
// ---------------------------------------------------------------------------
template <bool DEBUG, int SOLVER>
__device__ __forceinline__ void eval_c(const DevParams& P, const Buffers& B, int do_shift, int blk_) {
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const double2 vs = B.l_vel[s];
  const double rhos = B.l_rv[s].x;
  const double2 pcs = B.l_pc[s];
  const int ts = B.l_type[s];
  const double txs = B.tau[3 * s], tys = B.tau[3 * s + 1], tzs = B.tau[3 * s + 2];
  const double2 grs = B.grad[s];
  const int nes = B.near_[s];
  const double inv_rhos = 1.0 / rhos;
  constexpr bool kMuscl = SOLVER == SPH2D_SOLVER_MUSCL;
  double2 gps = make_double2(0.0, 0.0), gus = gps, gvs = gps;
  if (kMuscl) { gps = B.mg[3 * s]; gus = B.mg[3 * s + 1]; gvs = B.mg[3 * s + 2]; }
  double drho = 0.0, corr = 0.0, ax = 0.0, ay = 0.0, shx = 0.0, shy = 0.0, agx = 0.0, agy = 0.0;
  double iax = 0.0, iay = 0.0, avx = 0.0, avy = 0.0, aix = 0.0, aiy = 0.0;
  double spx = 0.0, spy = 0.0, lpx = 0.0, lpy = 0.0;
  // the 2x2 renormalisation (a mode of this library): every normalised gradient below is
  // T^-1 grad W of the summing particle; the identity otherwise
  double i00 = 1.0, i01 = 0.0, i10 = 0.0, i11 = 1.0;
  if (P.renorm) {
    const double* ti = B.tinv + 4 * (size_t)s;
    i00 = ti[0]; i10 = ti[1]; i01 = ti[2]; i11 = ti[3];
  }
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    if (ts != 2 && B.l_type[o] != 2) return;
    const double2 po = B.l_pos[o];
    const double dx = ps.x - po.x, dy = ps.y - po.y;      // x_ij with i = s
    const double r2 = dx * dx + dy * dy;
    const double r = sqrt(r2);
    const double inv_r = 1.0 / r;
    const double inv_r2 = inv_r * inv_r;
    const KernelTerm K = kernel_term(P, dx, dy, r, inv_r);
    double ndx = K.dwx, ndy = K.dwy;                  // Normalized_dwdx of the summing particle
    if (P.renorm) {
      ndx = i00 * K.dwx + i01 * K.dwy;
      ndy = i10 * K.dwx + i11 * K.dwy;
    }
    const double2 vo = B.l_vel[o];
    const double2 rvo = B.l_rv[o];
    const double2 pco = B.l_pc[o];
    const double rhoo = rvo.x, Vo = rvo.y;
    // Riemann problem along e = -x_ij/r (Solver...:306-356)
    const double ex = -(dx * inv_r), ey = -(dy * inv_r);
    const double UL = vs.x * ex + vs.y * ey;
    const double UR = vo.x * ex + vo.y * ey;
    const double dU = UL - UR;
    const double Pbar = (pcs.x + pco.x) * 0.5;
    const double vbx = (vs.x + vo.x) * 0.5, vby = (vs.y + vo.y) * 0.5;
    const double cbar = (pcs.y + pco.y) * 0.5;
    const double rbar = (rhos + rhoo) * 0.5;
    const double inv_rbar = 1.0 / rbar;
    double vsx, vsy, Ps;
    double2 vown = vs;                                    // MUSCL: own reconstructed velocity
    if (SOLVER == SPH2D_SOLVER_HLLC || SOLVER == SPH2D_SOLVER_MUSCL) {
      // the reference's pair order: pair_i = lower local index, e from pair_i to pair_j
      const bool left = s < o;
      const double epx = left ? ex : -ex, epy = left ? ey : -ey;
      const double2 vi = left ? vs : vo, vj = left ? vo : vs;
      const double2 pci = left ? pcs : pco, pcj = left ? pco : pcs;
      if (SOLVER == SPH2D_SOLVER_HLLC) {
        const double rhoi = left ? rhos : rhoo, rhoj = left ? rhoo : rhos;
        hllc_star<double>(vi.x, vi.y, pci.x, rhoi, pci.y, vj.x, vj.y, pcj.x, rhoj, pcj.y, epx, epy,
                          P.x_over_t, vsx, vsy, Ps);
      } else {
        const double2* mgo = B.mg + 3 * o;
        const double2 gpo = mgo[0], guo = mgo[1], gvo = mgo[2];
        const double2 gpi = left ? gps : gpo, gpj = left ? gpo : gps;
        const double2 gui = left ? gus : guo, guj = left ? guo : gus;
        const double2 gvi = left ? gvs : gvo, gvj = left ? gvo : gvs;
        const double xrx = left ? -dx : dx, xry = left ? -dy : dy;   // x_j - x_i of the pair
        double li, lj;
        muscl_limiter<double>(0.5 * (gpi.x * xrx + gpi.y * xry), -(0.5 * (gpj.x * xrx + gpj.y * xry)),
                              pcj.x - pci.x, li, lj);
        const double PL = pci.x + li, PR = pcj.x - lj;
        muscl_limiter<double>(0.5 * (gui.x * xrx + gui.y * xry), -(0.5 * (guj.x * xrx + guj.y * xry)),
                              vj.x - vi.x, li, lj);
        const double Iu = vi.x + li, Ju = vj.x - lj;
        muscl_limiter<double>(0.5 * (gvi.x * xrx + gvi.y * xry), -(0.5 * (gvj.x * xrx + gvj.y * xry)),
                              vj.y - vi.y, li, lj);
        const double Iv = vi.y + li, Jv = vj.y - lj;
        const double UL = Iu * epx + Iv * epy;
        const double UR = Ju * epx + Jv * epy;
        const double Ub = (UL + UR) * 0.5;
        const double Us = Ub + 0.5 * (PL - PR) * inv_rbar * P.inv_c0;
        vsx = (Us - Ub) * epx + vbx;
        vsy = (Us - Ub) * epy + vby;
        const double dUp = UL - UR;
        double bR = 3.0 * (dUp > 0.0 ? dUp : 0.0);
        bR = bR < cbar ? bR : cbar;
        Ps = Pbar + 0.5 * bR * rbar * dUp;
        vown = left ? make_double2(Iu, Iv) : make_double2(Ju, Jv);
      }
    } else if (SOLVER == SPH2D_SOLVER_LINEAR_RIEMANN) {
      const double zs = rhos * pcs.y, zo = rhoo * pco.y;
      const double inv_z = 1.0 / (zs + zo);
      const double Us = (zs * UR + zo * UL + pcs.x - pco.x) * inv_z;
      vsx = Us * ex;
      vsy = Us * ey;
      Ps = (zs * pco.x + zo * pcs.x + zs * zo * dU) * inv_z;
    } else {
      const double Ub = (UL + UR) * 0.5;
      const double Us = Ub + 0.5 * (pcs.x - pco.x) * inv_rbar * P.inv_c0;
      vsx = (Us - Ub) * ex + vbx;
      vsy = (Us - Ub) * ey + vby;
      double bR = 3.0 * (dU > 0.0 ? dU : 0.0);
      bR = bR < cbar ? bR : cbar;
      Ps = Pbar + 0.5 * bR * rbar * dU;
    }
    if (nes) { vsx = vbx; vsy = vby; Ps = Pbar; }  // near the free surface: averages
    const double ti = (vown.x - vsx) * ndx + (vown.y - vsy) * ndy;
    drho = drho + 2 * rhos * Vo * ti;                        // :360-369
    const double rij = inv_rhos / rhoo;                      // 1/(rho_i rho_j)
    const double mo = B.l_mass[o];
    const double fp = mo * (-2 * Ps * rij);                  // :374-388
    double tix = fp * ndx, tiy = fp * ndy;
    const double vijx = vs.x - vo.x, vijy = vs.y - vo.y;
    const double vx = vijx * dx + vijy * dy;
    double tvx = 0.0, tvy = 0.0;
    if (vx < 0.0) {                                          // :392-417
      const double fh = P.fai * P.h;
      const double phi = P.h * vx / (r2 + fh * fh);
      const double Pi = (-(P.alpha * cbar * phi) + P.beta * (phi * phi)) * inv_rbar;
      tvx = mo * (-Pi * ndx);
      tvy = mo * (-Pi * ndy);
    }
    const double Rd = dx * ndx + dy * ndy;
    const double lf = Vo * 2 * P.nu * Rd * inv_r2;           // :432-436
    const double tlx = lf * vijx, tly = lf * vijy;
    const double tmp = P.delta_alpha * P.h * cbar * vx * inv_r2 * P.rho0 * inv_rhos * Vo;   // :442-446
    const double tax = tmp * ndx, tay = tmp * ndy;
    // SPS (:450-492): raw dwdx, particle_mass at the LOCAL index, temp_p carried from x to y
    const double gmo = B.g_mass[o];
    const double txo = B.tau[3 * o], tyo = B.tau[3 * o + 1], tzo = B.tau[3 * o + 2];
    double tp = (txs + txo) * K.dwx + (tys + tyo) * K.dwy;
    tp = tp * rij;
    const double tsx = gmo * tp;
    tp = tp + (tzs + tzo) * K.dwy + (tys + tyo) * K.dwx;
    tp = tp * rij;
    const double tsy = gmo * tp;
    if (DEBUG) {
      iax += tix; iay += tiy; avx += tvx; avy += tvy; aix += tax; aiy += tay;
      spx += tsx; spy += tsy; lpx += tlx; lpy += tly;
    } else {
      ax = ax + ((((tix + tvx) + tax) + tsx) + tlx);
      ay = ay + ((((tiy + tvy) + tay) + tsy) + tly);
    }
    if (SOLVER == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && P.epsilon_xsph != 0.0) {   // :495-500
      const double f = P.epsilon_xsph * P.xsph_f * mo / (rhos + rhoo) * K.w;
      agx = agx - f * vijx;
      agy = agy - f * vijy;
    }
    {  // delta-SPH density diffusion (:552-583) with x_j - x_i = -x_ij
      const double2 gro = B.grad[o];
      const double psi = (rhoo - rhos) + 0.5 * ((grs.x + gro.x) * dx + (grs.y + gro.y) * dy);
      corr = corr - 2 * P.delta * P.h * cbar * psi * Rd * inv_r2 * Vo;
    }
    if (do_shift) {                                           // Shift_particle_position.f90:78-103
      const double wr = K.w * P.inv_wdx;
      const double f = P.shift_beta * P.h2 * (1 + P.shift_R * (wr * wr * wr * wr)) * Vo;
      shx = shx - f * ndx;
      shy = shy - f * ndy;
    }
  }, B.nb_list, B.nb_count, B.kstride);
  if (ts == 2) drho = drho + corr;
  B.drhodt[s] = drho;
  if (DEBUG) {
    ax = iax + 0.0 + avx + aix + spx + lpx;
    ay = iay + (0.0 - P.g) + avy + aiy + spy + lpy;        // single_step_compute.f90:161-165
    double* d = B.dbg + 10 * s;
    d[0] = iax; d[1] = iay; d[2] = avx; d[3] = avy; d[4] = aix; d[5] = aiy;
    d[6] = spx; d[7] = spy; d[8] = lpx; d[9] = lpy;
  }
  else ay = ay - P.g;                                       // external acceleration (gravity, or 0)
  if (ts != 2) { ax = 0.0; ay = 0.0; }
  B.acc[s] = make_double2(ax, ay);
  if (P.epsilon_xsph != 0.0) B.avg[s] = make_double2(agx, agy);
  if (do_shift) B.shift[s] = make_double2(shx, shy);
}

// The compiler's own register allocation: asking for 3 waves per SIMD (168
// VGPRs) spills the force pass to scratch and costs more than it gains.
template <bool DEBUG, int SOLVER>
__global__ void __launch_bounds__(kBlock) k_eval_c(DevParams P, Buffers B, int do_shift) {
  eval_c<DEBUG, SOLVER>(P, B, do_shift, xcd_block());
}

template <bool DEBUG>
static void launch_c64(hipStream_t s, const DevParams& P, Buffers& B, int g) {
  switch (P.solver) {
    case SPH2D_SOLVER_LINEAR_RIEMANN:
      hipLaunchKernelGGL((k_eval_c<DEBUG, SPH2D_SOLVER_LINEAR_RIEMANN>), dim3(g), dim3(kBlock), 0, s, P, B, 1);
      break;
    case SPH2D_SOLVER_HLLC:
      hipLaunchKernelGGL((k_eval_c<DEBUG, SPH2D_SOLVER_HLLC>), dim3(g), dim3(kBlock), 0, s, P, B, 1);
      break;
    case SPH2D_SOLVER_MUSCL:
      hipLaunchKernelGGL((k_eval_c<DEBUG, SPH2D_SOLVER_MUSCL>), dim3(g), dim3(kBlock), 0, s, P, B, 1);
      break;
    default:
      hipLaunchKernelGGL((k_eval_c<DEBUG, SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN>), dim3(g), dim3(kBlock), 0, s,
                         P, B, 1);
  }
}

static void launch_a64(hipStream_t s, const DevParams& P, Buffers& B, int g) {
  if (P.solver == SPH2D_SOLVER_MUSCL) hipLaunchKernelGGL((k_eval_a<true, false>), dim3(g), dim3(kBlock), 0, s, P, B);
  else if (P.renorm) hipLaunchKernelGGL((k_eval_a<false, true>), dim3(g), dim3(kBlock), 0, s, P, B);
  else hipLaunchKernelGGL((k_eval_a<false, false>), dim3(g), dim3(kBlock), 0, s, P, B);
}

void launch_eval_part(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int part) {
  const int g = grid_for(cap_local);
  if (part == 0) {
    launch_fill_i32(s, B.counters + C_NSURF, 1, 0);
    if (P.mixed) launch_eval32_a(s, P, B, cap_local);
    else launch_a64(s, P, B, g);
  } else if (part == 1) {
    hipLaunchKernelGGL(k_eval_b1, dim3(g), dim3(kBlock), 0, s, P, B);
    hipLaunchKernelGGL(k_eval_b2, dim3(g), dim3(kBlock), 0, s, P, B);
  } else {
    if (P.mixed) launch_eval32_c(s, P, B, cap_local);
    else if (P.write_debug) launch_c64<true>(s, P, B, g);
    else launch_c64<false>(s, P, B, g);
  }
}

__global__ void k_clear_tau(Buffers B, int mixed) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  B.tau[3 * s] = 0.0; B.tau[3 * s + 1] = 0.0; B.tau[3 * s + 2] = 0.0;
  if (mixed) {
    const float4 r = B.m_r1[s];
    B.m_r1[s] = make_float4(0.f, 0.f, 0.f, r.w);   // .w: particle_mass at the local index
  }
}

void launch_clear_tau(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_clear_tau, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, B, P.mixed);
}

void launch_eval(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, hipEvent_t* ev) {
  const int g = grid_for(cap_local);
  launch_fill_i32(s, B.counters + C_NSURF, 1, 0);
  if (ev) (void)hipEventRecord(ev[0], s);
  if (P.mixed) launch_eval32_a(s, P, B, cap_local);
  else launch_a64(s, P, B, g);
  if (ev) (void)hipEventRecord(ev[1], s);
  hipLaunchKernelGGL(k_eval_b1, dim3(g), dim3(kBlock), 0, s, P, B);
  hipLaunchKernelGGL(k_eval_b2, dim3(g), dim3(kBlock), 0, s, P, B);
  if (ev) (void)hipEventRecord(ev[2], s);
  if (P.mixed) launch_eval32_c(s, P, B, cap_local);
  else if (P.write_debug) launch_c64<true>(s, P, B, g);
  else launch_c64<false>(s, P, B, g);
  if (ev) (void)hipEventRecord(ev[3], s);
}

// Refresh_domain.f90:320-335: the vorticity of the last evaluation's pairs,
// sum_j V_j (v_i - v_j) x Normalized_dwdx with the identity renormalisation; the
// pair's j-side term V_i (v_i - v_j) x dwdx_ij equals the same gather term seen
// from j. Written for fluid particles, in global order (Refresh's reduce).
__global__ void __launch_bounds__(kBlock) k_vorticity(DevParams P, Buffers B) {
  const int s = xcd_block() * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const int g = B.name[s];
  if (g >= B.counters[C_NFLUID]) return;           // ghosts are not refreshed
  const int ts = B.l_type[s];
  const double2 ps = B.l_pos[s];
  const double2 vs = B.l_vel[s];
  double wz = 0.0;
  double i00 = 1.0, i01 = 0.0, i10 = 0.0, i11 = 1.0;
  if (P.renorm) {
    const double* ti = B.tinv + 4 * (size_t)s;
    i00 = ti[0]; i10 = ti[1]; i01 = ti[2]; i11 = ti[3];
  }
  for_each_neighbour<kNbReplay>(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    if (ts != 2 && B.l_type[o] != 2) return;       // find_particle_pair.f90:76
    const double2 po = B.l_pos[o];
    const double pdx = ps.x - po.x, pdy = ps.y - po.y;
    const double r = sqrt(pdx * pdx + pdy * pdy);
    double w, dwx, dwy;
    kernel_w_dw(P, r, pdx, pdy, w, dwx, dwy);
    if (P.renorm) {                                 // Normalized_dwdx of the summing particle
      const double rx = dwx, ry = dwy;
      dwx = i00 * rx + i01 * ry;
      dwy = i10 * rx + i11 * ry;
    }
    const double2 vo = B.l_vel[o];
    const double vx = vs.x - vo.x, vy = vs.y - vo.y;
    wz = wz + B.l_rv[o].y * (vx * dwy - vy * dwx);   // Cross_Product(...)(3), function_module.f90:150
  }, B.nb_list, B.nb_count, B.kstride);
  B.g_vort[g] = wz;
}

void launch_vorticity(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_vorticity, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B);
}

void launch_eval_pass(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int pass) {
  const int g = grid_for(cap_local);
  if (pass == 0) {
    if (P.mixed) launch_eval32_c(s, P, B, cap_local);
    else launch_c64<false>(s, P, B, g);
  } else if (pass == 1) {
    if (P.mixed) launch_eval32_a(s, P, B, cap_local);
    else launch_a64(s, P, B, g);
  } else {
    launch_eval(s, P, B, cap_local);
  }
}

// Batched evaluation passes (see launch_batch_step): XCD-aware block order over
// the whole launch, then the instance of the block.
template <bool MUSCL>
__global__ void __launch_bounds__(kBlock) kb_eval_a(BatchArgs A) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_eval_a_body<MUSCL>(P, B, w.blk, w.nblk);
}

__global__ void __launch_bounds__(kBlock) kb_eval_b1(BatchArgs A) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_eval_b1_body(P, B, w.blk, w.nblk);
}

__global__ void __launch_bounds__(kBlock) kb_eval_b2(BatchArgs A) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_eval_b2_body(P, B, w.blk, w.nblk);
}

template <int SOLVER>
__global__ void __launch_bounds__(kBlock) kb_eval_c(BatchArgs A) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  const Buffers B = A.inst[w.i].B;
  eval_c<false, SOLVER>(P, B, 1, w.blk);
}

void launch_batch_eval(hipStream_t s, const BatchDev& D, int mixed, int solver) {
  const int g = D.total[BC_LOCAL];
  if (g <= 0) return;
  const BatchArgs A{D.inst, D.start[BC_LOCAL], D.n, D.xcd};
  if (mixed) launch_batch_eval32(s, D, 0, solver);
  else if (solver == SPH2D_SOLVER_MUSCL) hipLaunchKernelGGL(kb_eval_a<true>, dim3(g), dim3(kBlock), 0, s, A);
  else hipLaunchKernelGGL(kb_eval_a<false>, dim3(g), dim3(kBlock), 0, s, A);
  hipLaunchKernelGGL(kb_eval_b1, dim3(g), dim3(kBlock), 0, s, A);
  hipLaunchKernelGGL(kb_eval_b2, dim3(g), dim3(kBlock), 0, s, A);
  if (mixed) {
    launch_batch_eval32(s, D, 1, solver);
    return;
  }
  switch (solver) {
    case SPH2D_SOLVER_LINEAR_RIEMANN:
      hipLaunchKernelGGL(kb_eval_c<SPH2D_SOLVER_LINEAR_RIEMANN>, dim3(g), dim3(kBlock), 0, s, A);
      break;
    case SPH2D_SOLVER_HLLC:
      hipLaunchKernelGGL(kb_eval_c<SPH2D_SOLVER_HLLC>, dim3(g), dim3(kBlock), 0, s, A);
      break;
    case SPH2D_SOLVER_MUSCL:
      hipLaunchKernelGGL(kb_eval_c<SPH2D_SOLVER_MUSCL>, dim3(g), dim3(kBlock), 0, s, A);
      break;
    default:
      hipLaunchKernelGGL(kb_eval_c<SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN>, dim3(g), dim3(kBlock), 0, s, A);
  }
}

}  // namespace sph2d
