// sph2d_adjstep.hip -- reverse of whole predictor-corrector steps of a walled tank
// with a piston wavemaker (the discrete adjoint behind sph2d_forward/sph2d_backward).
//
// One forward step n (time_iteration.f90:541-657 with Wavemaker_Subroutine.f90:140-215
// before it and Shift_particle_position.f90 / Refresh_domain.f90 after it):
//   wavemaker update   x_wm = x_p - (layer-1) dx - dx/2, u_wm = u_p, rho_wm(H) below the
//                      water height H = max y of fluid within 3 dx of x_p, + dx/2;
//                      fluid behind the paddle mirrored, fluid within dx of it at u_p
//   predictor          v' = v0 + dt a_(n-1), x' = x0 + dt v', rho' = rho0 + dt d_(n-1)
//                      (the previous rates read by local slot, time_iteration.f90:542-555)
//   evaluation         (d_n, a_n, shift_n, grad rho_n) of the predicted state, walls
//                      assigned from it, end damping on a_n
//   corrector          v = v0 + dt/2 (a_n + a_(n-1)), x = x0 + dt/2 (v + v') + dt^2/2 a_n,
//                      rho = rho0 + dt/2 (d_n + d_(n-1))
//   shift              x += clamp(shift_n), rho += grad rho_n . clamp(shift_n), p = Tait(rho)
//                      for fluid not near the surface (else p = Tait(rho'))
// and, every few steps, the probe pressures (Sampling_subroutine.f90:21-177, MLS of
// order 6 with the hydrostatic correction) from the refreshed state.
//
// The reverse of step n, with lamG the adjoint of the refreshed fluid state (x, y, u,
// v, rho, p by gid) and lamR that of the rates d_n, a_n by local slot (from step n+1's
// predictor): the probe reverse adds to lamG; k_adj_stage1 reverses shift and corrector
// into weights on the evaluation's outputs; the evaluation's vector-Jacobian product
// (slot form) gives the adjoint of the predicted state; k_adj_stage2 reverses the
// predictor into lamG (step-start state) and lamR (previous rates); the wavemaker
// reverse takes the paddle's share (and the pull-back's) into (dL/dx_p, dL/du_p) of
// the step. What stays fixed: the neighbour sets, free-surface and near-surface
// flags, the limiter branches, the shift clamp's branch, the damping zone's membership,
// the pull-back branches, the water height's arg-max particle, which walls and
// wavemaker particles have mass, and the SPS term's particle_mass read by local index.
#include "sph2d_internal.h"
#include "sph2d_device.h"

namespace sph2d {

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

__global__ void k_adjs_zero(double* p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

void launch_adj_zero(hipStream_t s, double* p, size_t n) {
  if (n) hipLaunchKernelGGL(k_adjs_zero, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, s, p, n);
}

__device__ __forceinline__ double tait_dp(const DevParams& P, double rho) {
  const double x = rho / P.rho0;
  const double x2 = x * x, x6 = x2 * x2 * x2;
  return P.eos_b * 7.0 * x6 / P.rho0;
}

// End damping factor at (x, y) and its gradient (single_step_compute.f90:174-260 and the
// linear ramp); false outside the zone.
__device__ __forceinline__ bool damping_factor(const DevParams& P, double x, double y, double& f, double& fx,
                                               double& fy) {
  const double Lx = P.box_x, L = P.damp_len, d = P.water_depth;
  f = 1.0; fx = 0.0; fy = 0.0;
  if (!P.damping || !((Lx - x) <= L && x <= Lx)) return false;
  const double pi = 3.14159265358979324;
  if (P.damping == 1) {
    const double th = pi * (x + L - Lx) / L;
    f = 1 - 0.5 * P.damp_amp * (1 - cos(th));
    fx = -0.5 * P.damp_amp * sin(th) * (pi / L);
  } else if (P.damping == 2) {
    const double u = (Lx - x) / L;
    const double e = exp(pow(u, 3.5));
    f = (e - 1.0) / (P.e32 - 1.0) * (y / d);
    fx = e * 3.5 * pow(u, 2.5) * (-1.0 / L) / (P.e32 - 1.0) * (y / d);
    fy = (e - 1.0) / (P.e32 - 1.0) / d;
  } else if (P.damping == 4) {
    f = 1 - P.damp_amp * ((x + L - Lx) / L);
    fx = -P.damp_amp / L;
  } else {
    const double q = (x + L - Lx) / L;
    f = 1 - 6 * (q * q) * ((d - y) / d);
    fx = -12 * q / L * ((d - y) / d);
    fy = 6 * (q * q) / d;
  }
  return true;
}

// Reverse of shift and corrector for the fluid particle in slot s (gid g), and the
// evaluation's output weights of every local slot. scr keeps, per slot, what the
// predictor's reverse needs: lam v' (2), lam x0 (2), lam v0 (2), lam rho0, lam rho'
// (from p = Tait(rho') of an unshifted particle), lam a_(n-1) (2), lam d_(n-1).
constexpr int kScr = 11;

__global__ void __launch_bounds__(kBlock) k_adj_stage1(DevParams P, Buffers B, const double* lamG, const double* lamR,
                                                       double* W, double* scr, int shift_on) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int t = B.l_type[s];
  const int g = B.name[s];
  double* w = W + 10 * (size_t)s;
  const double* lr = lamR + 3 * (size_t)s;
  for (int k = 0; k < 10; ++k) w[k] = 0.0;
  w[0] = lr[0];                                  // the next step's predictor reads d_n by slot
  double* sc = scr + kScr * (size_t)s;
  for (int k = 0; k < kScr; ++k) sc[k] = 0.0;
  if (t != 2 || g < 0 || g >= B.counters[C_NFLUID]) return;
  const double dt = P.dt;
  const double* lg = lamG + 6 * (size_t)g;
  double lxf = lg[0], lyf = lg[1];
  const double luc = lg[2], lvc = lg[3];
  double lrf = lg[4];
  const double lpf = lg[5];
  const bool shifted = shift_on && B.near_[s] != 1;
  double lrp_x = 0.0;
  double lshx = 0.0, lshy = 0.0, lgrx = 0.0, lgry = 0.0;
  if (shifted) {
    const double rf = B.g_rho[g];                // refreshed density (after the shift)
    lrf = lrf + lpf * tait_dp(P, rf);
    const double2 raw = B.shift[s];
    const double amp = sqrt(raw.x * raw.x + raw.y * raw.y);
    double2 sh = raw;
    const bool clamp = amp > P.max_shift;
    if (clamp) { sh.x = raw.x / amp * P.max_shift; sh.y = raw.y / amp * P.max_shift; }
    const double2 gr = B.grad[s];
    const double lsx = lxf + lrf * gr.x, lsy = lyf + lrf * gr.y;
    lgrx = lrf * sh.x;
    lgry = lrf * sh.y;
    if (clamp) {
      const double ux = raw.x / amp, uy = raw.y / amp;
      const double dot = lsx * ux + lsy * uy;
      lshx = (P.max_shift / amp) * (lsx - dot * ux);
      lshy = (P.max_shift / amp) * (lsy - dot * uy);
    } else {
      lshx = lsx;
      lshy = lsy;
    }
  } else {
    lrp_x = lpf * tait_dp(P, B.l_rv[s].x);       // p = Tait(rho') for an unshifted particle
  }
  // x = x0 + dt/2 (v + v') + dt^2/2 a ; v = v0 + dt/2 (a + a') ; rho = rho0 + dt/2 (d + d')
  const double lvx = luc + 0.5 * dt * lxf, lvy = lvc + 0.5 * dt * lyf;
  const double lax = 0.5 * dt * lvx + 0.5 * dt * dt * lxf;
  const double lay = 0.5 * dt * lvy + 0.5 * dt * dt * lyf;
  const double ld = 0.5 * dt * lrf;
  const double2 xe = B.l_pos[s];
  double f, fx, fy;
  damping_factor(P, xe.x, xe.y, f, fx, fy);
  const double lax_t = lax + lr[1], lay_t = lay + lr[2];
  w[0] = ld + lr[0];
  w[1] = ld + lr[0];                             // the delta-SPH correction is part of d for fluid
  w[2] = f * lax_t;
  w[3] = f * lay_t;
  w[6] = lshx; w[7] = lshy;
  w[8] = lgrx; w[9] = lgry;
  sc[0] = 0.5 * dt * lxf; sc[1] = 0.5 * dt * lyf;          // lam v'
  sc[2] = lxf; sc[3] = lyf;                                // lam x0
  sc[4] = lvx; sc[5] = lvy;                                // lam v0
  sc[6] = lrf;                                             // lam rho0
  sc[7] = lrp_x;                                           // lam rho'
  sc[8] = 0.5 * dt * lvx; sc[9] = 0.5 * dt * lvy;          // lam a_(n-1)
  sc[10] = 0.5 * dt * lrf;                                 // lam d_(n-1)
}

// Reverse of the predictor (and of the damping factor's dependence on the evaluated
// position): G holds d/d(x', y', u', v', rho') per slot from the evaluation's reverse.
// Writes the step-start adjoint into lamG (p zero) and the previous rates' into lamRp.
__global__ void __launch_bounds__(kBlock) k_adj_stage2(DevParams P, Buffers B, const double* G, const double* scr,
                                                       const double* lamR, double* lamG, double* lamRp) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int t = B.l_type[s];
  const int g = B.name[s];
  double* lrp = lamRp + 3 * (size_t)s;
  lrp[0] = 0.0; lrp[1] = 0.0; lrp[2] = 0.0;
  if (g < 0 || g >= B.counters[C_NFLUID]) return;
  double* lg = lamG + 6 * (size_t)g;
  if (t != 2) {                                  // not moved by the stages: p = Tait(rho) again
    lg[4] = lg[4] + lg[5] * tait_dp(P, B.l_rv[s].x);
    lg[5] = 0.0;
    return;
  }
  const double dt = P.dt;
  const double* gs = G + 5 * (size_t)s;
  const double* sc = scr + kScr * (size_t)s;
  double lxp = gs[0], lyp = gs[1];
  const double2 xe = B.l_pos[s];
  double f, fx, fy;
  if (damping_factor(P, xe.x, xe.y, f, fx, fy) && f != 0.0) {
    const double* lr = lamR + 3 * (size_t)s;
    const double lax = 0.5 * dt * sc[4] + 0.5 * dt * dt * sc[2] + lr[1];
    const double lay = 0.5 * dt * sc[5] + 0.5 * dt * dt * sc[3] + lr[2];
    const double2 a = B.acc[s];
    const double la_raw = (lax * a.x + lay * a.y) / f;
    lxp = lxp + la_raw * fx;
    lyp = lyp + la_raw * fy;
  }
  double lvpx = sc[0] + gs[2], lvpy = sc[1] + gs[3];
  const double lrp_ = gs[4] + sc[7];
  // x' = x0 + dt v'
  const double lbx = sc[2] + lxp, lby = sc[3] + lyp;
  lvpx = lvpx + dt * lxp;
  lvpy = lvpy + dt * lyp;
  // v' = v0 + dt a' ; rho' = rho0 + dt d'
  lg[0] = lbx;
  lg[1] = lby;
  lg[2] = sc[4] + lvpx;
  lg[3] = sc[5] + lvpy;
  lg[4] = sc[6] + lrp_;
  lg[5] = 0.0;
  lrp[0] = sc[10] + dt * lrp_;
  lrp[1] = sc[8] + dt * lvpx;
  lrp[2] = sc[9] + dt * lvpy;
}

// Wavemaker_Subroutine.f90:196-215 branches of the fluid particle g at the step's start
// (before the update): bit 0 mirrored (x < x_p), bit 1 set to the paddle's velocity.
// Also the water height's arg-max: the smallest gid holding the maximum y within 3 dx.
__global__ void __launch_bounds__(kBlock) k_adj_wm_flags(DevParams P, Buffers B, double xp, int* flags,
                                                         const unsigned long long* hbits, int* argmax) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B.counters[C_NFLUID]) return;
  const double2 x = B.g_pos[g];
  int f = 0;
  if (B.g_type[g] != 120) {
    double xn = x.x;
    if (xn < xp) { xn = 2 * xp - xn; f |= 1; }
    if ((xn - xp) < P.dx) f |= 2;
  }
  flags[g] = f;
  const double hmax = __longlong_as_double(*hbits);
  if (hmax > 0.0 && fabs(xp - x.x) <= 3 * P.dx && x.y == hmax) atomicMin(argmax, g);
}

// Per gid, its share of (dL/dx_p, dL/du_p, dL/dH): wavemaker particles through their
// placement (G by slot of the evaluation's reverse), fluid through the pull-back; the
// fluid's start adjoint in lamG becomes that of its state before the update.
__global__ void __launch_bounds__(kBlock) k_adj_wm_share(DevParams P, Buffers B, const double* G, const int* flags,
                                                         const unsigned long long* hbits, double* lamG, double* c3,
                                                         int n_wm) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  if (g >= N + n_wm) return;
  double cx = 0.0, cu = 0.0, cH = 0.0;
  if (g < N) {
    double* lg = lamG + 6 * (size_t)g;
    const int f = flags[g];
    if (f & 2) { cu = lg[2]; lg[2] = 0.0; }
    if (f & 1) { cx = 2.0 * lg[0]; lg[0] = -lg[0]; }
  } else {
    const int s = B.l_of_g[g];
    if (s >= 0) {
      const double* gs = G + 5 * (size_t)s;
      cx = gs[0];
      cu = gs[2];
      if (B.l_mass[s] != 0.0) {                    // below the water height: rho(H)
        const double y = B.g_pos[g].y;
        const double hgt = __longlong_as_double(*hbits) + 0.5 * P.dx;   // the update's height
        const double e6 = (double)(1.0f / 6.0f);
        const double base = -6.0 * 9.81 * (y - hgt) / P.c0sq + 1.0;
        const double drdH = P.rho0 * e6 * pow(base, e6 - 1.0) * (6.0 * 9.81 / P.c0sq);
        cH = gs[4] * drdH;
      }
    }
  }
  c3[3 * (size_t)g + 0] = cx;
  c3[3 * (size_t)g + 1] = cu;
  c3[3 * (size_t)g + 2] = cH;
}

// Fixed-order sum of c3 over n gids: block partials (k_adj_wm_part, contiguous ranges
// summed in order by a fixed tree), then one block over the partials; dL/dH goes into
// the arg-max particle's y and (dL/dx_p, dL/du_p) of the step into out.
constexpr int kRedBlocks = 256;

__global__ void __launch_bounds__(256) k_adj_wm_part(const double* c3, int n, double* part) {
  __shared__ double sh[3][256];
  const int per = (n + kRedBlocks - 1) / kRedBlocks;
  const int lo = blockIdx.x * per, hi = min(n, lo + per);
  double a[3] = {0.0, 0.0, 0.0};
  for (int i = lo + threadIdx.x; i < hi; i += blockDim.x)
    for (int k = 0; k < 3; ++k) a[k] = a[k] + c3[3 * (size_t)i + k];
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = a[k];
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 3; ++k) part[3 * blockIdx.x + k] = sh[k][0];
}

__global__ void __launch_bounds__(kRedBlocks) k_adj_wm_reduce(const double* part, int n, const int* argmax,
                                                              double* lamG, double* out) {
  __shared__ double sh[3][kRedBlocks];
  for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = part[3 * threadIdx.x + k];
  __syncthreads();
  for (int off = blockDim.x / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off)
      for (int k = 0; k < 3; ++k) sh[k][threadIdx.x] = sh[k][threadIdx.x] + sh[k][threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sh[0][0];
    out[1] = sh[1][0];
    const int am = *argmax;
    if (am >= 0 && am < n) lamG[6 * (size_t)am + 1] = lamG[6 * (size_t)am + 1] + sh[2][0];
  }
}

// A snapshot copy in one launch: up to 10 arrays, in 4-byte words.
// Up to kAdjCopies copies of whole 4-byte words; 16 bytes a lane where both ends are
// 16-byte aligned (every allocation and image slot is), the last words one at a time.
__global__ void __launch_bounds__(256) k_adj_copy(AdjCopy C) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (int a = 0; a < C.n; ++a) {
    const uint32_t* src = static_cast<const uint32_t*>(C.src[a]);
    uint32_t* dst = static_cast<uint32_t*>(C.dst[a]);
    const size_t w = C.words[a];
    size_t done = 0;
    if ((((uintptr_t)src | (uintptr_t)dst) & 15) == 0) {
      const size_t n4 = w / 4;
      const uint4* s4 = reinterpret_cast<const uint4*>(src);
      uint4* d4 = reinterpret_cast<uint4*>(dst);
      for (size_t i = t0; i < n4; i += stride) d4[i] = s4[i];
      done = 4 * n4;
    }
    for (size_t i = done + t0; i < w; i += stride) dst[i] = src[i];
  }
}

void launch_adj_copy(hipStream_t st, const AdjCopy& C) {
  hipLaunchKernelGGL(k_adj_copy, dim3(1024), dim3(256), 0, st, C);
}

// 6x6 symmetric positive definite solve by Cholesky (the same factorisation as the
// forward's first-row solve), right-hand side rhs.
__device__ __forceinline__ void chol6_solve(const double* A, double al3, const double* rhs, double* y) {
  double L[21];
  const int diag[6] = {0, 6, 11, 15, 18, 20};
#define AU(m, l) A[diag[(m)] + (l) - (m)]
#define LL(i, j) L[(i) * ((i) + 1) / 2 + (j)]
  for (int i = 0; i < 6; ++i)
    for (int j = 0; j <= i; ++j) {
      double v = AU(j, i) + (i == j ? al3 : 0.0);
      for (int k = 0; k < j; ++k) v = v - LL(i, k) * LL(j, k);
      if (i == j) LL(i, i) = sqrt(v > 0.0 ? v : 1.0);
      else LL(i, j) = v / LL(j, j);
    }
  double z[6];
  for (int i = 0; i < 6; ++i) {
    double v = rhs[i];
    for (int k = 0; k < i; ++k) v = v - LL(i, k) * z[k];
    z[i] = v / LL(i, i);
  }
  for (int i = 5; i >= 0; --i) {
    double v = z[i];
    for (int k = i + 1; k < 6; ++k) v = v - LL(k, i) * y[k];
    y[i] = v / LL(i, i);
  }
#undef AU
#undef LL
}

// Reverse of one MLS probe pressure (k_probe_mls's second block) at (ex, ey) with
// weight lam: p = sum_j F_j q_j, q_j = p_j - g rho_j (ey - y_j), F_j = w_j (y . P_j) with
// y = A^-1 e1, A = sum w P P^T + al3 I (or Shepard w_j / sum w). With z = A^-1 b,
// b = sum w q P: dp = y . db - y^T dA z. One thread: the support is visited in the
// forward's order and each particle once.
__global__ void k_adj_probe(DevParams P, Buffers B, double ex, double ey, double gravity, const double* lam_ptr,
                            double* lamG) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const double lam = *lam_ptr;
  if (lam == 0.0) return;
  const int N = B.counters[C_NFLUID];
  double A[21], b[6];
  for (int q = 0; q < 21; ++q) A[q] = 0.0;
  for (int q = 0; q < 6; ++q) b[q] = 0.0;
  double sw = 0.0, swq = 0.0;
  int nt = 0;
  for_each_support(P, B, ex, ey, N, [&](int g, double2 pj, double w) {
    const double a = (pj.x - ex) / P.h, c = (pj.y - ey) / P.h;
    const double pv[6] = {1.0, a, c, a * a, a * c, c * c};
    const double qj = B.g_p[g] - (gravity != 0.0 ? gravity * B.g_rho[g] * (ey - pj.y) : 0.0);
    int q = 0;
    for (int m = 0; m < 6; ++m) {
      const double bm = pv[m] * w;
      for (int l = m; l < 6; ++l) { A[q] = A[q] + bm * pv[l]; ++q; }
      b[m] = b[m] + w * qj * pv[m];
    }
    sw = sw + w;
    swq = swq + w * qj;
    ++nt;
  });
  if (nt <= 3) return;
  const double al3 = (P.mls_alpha * P.mls_alpha) * P.mls_alpha;
  double y[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, z[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  bool mls = nt > 10;
  if (mls) mls = mls6_first_row(A, al3, y);
  if (mls) chol6_solve(A, al3, b, z);
  const double pshep = swq / sw;
  for_each_support(P, B, ex, ey, N, [&](int g, double2 pj, double w) {
    const double a = (pj.x - ex) / P.h, c = (pj.y - ey) / P.h;
    const double pv[6] = {1.0, a, c, a * a, a * c, c * c};
    const double qj = B.g_p[g] - (gravity != 0.0 ? gravity * B.g_rho[g] * (ey - pj.y) : 0.0);
    const double pdx = ex - pj.x, pdy = ey - pj.y;
    const double r = sqrt(pdx * pdx + pdy * pdy);
    double ww, dwx, dwy;
    kernel_w_dw(P, r, pdx, pdy, ww, dwx, dwy);
    const double wxj = -dwx, wyj = -dwy;         // dW/dx_j, dW/dy_j
    double F, gx, gy;
    if (mls) {
      double yp = 0.0, pz = 0.0;
      for (int m = 0; m < 6; ++m) { yp += y[m] * pv[m]; pz += pv[m] * z[m]; }
      F = w * yp;
      // dP/dx_j = (1/h) dP/da, dP/dy_j = (1/h) dP/dc
      const double da[6] = {0.0, 1.0, 0.0, 2.0 * a, c, 0.0};
      const double dc[6] = {0.0, 0.0, 1.0, 0.0, a, 2.0 * c};
      double yda = 0.0, zda = 0.0, ydc = 0.0, zdc = 0.0;
      for (int m = 0; m < 6; ++m) {
        yda += y[m] * da[m]; zda += z[m] * da[m];
        ydc += y[m] * dc[m]; zdc += z[m] * dc[m];
      }
      yda /= P.h; zda /= P.h; ydc /= P.h; zdc /= P.h;
      gx = qj * (wxj * yp + w * yda) - (wxj * yp * pz + w * (yda * pz + yp * zda));
      gy = qj * (wyj * yp + w * ydc) - (wyj * yp * pz + w * (ydc * pz + yp * zdc));
    } else {
      F = w / sw;
      gx = wxj * (qj - pshep) / sw;
      gy = wyj * (qj - pshep) / sw;
    }
    double* lg = lamG + 6 * (size_t)g;
    lg[0] += lam * gx;
    lg[1] += lam * (gy + (gravity != 0.0 ? F * gravity * B.g_rho[g] : 0.0));
    lg[4] += lam * (gravity != 0.0 ? -F * gravity * (ey - pj.y) : 0.0);
    lg[5] += lam * F;
  });
}

void launch_adj_stage1(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, const double* lamG,
                       const double* lamR, double* W, double* scr, int shift_on) {
  hipLaunchKernelGGL(k_adj_stage1, dim3(grid_for(cap_local)), dim3(kBlock), 0, st, P, B, lamG, lamR, W, scr, shift_on);
}

void launch_adj_stage2(hipStream_t st, const DevParams& P, Buffers& B, int cap_local, const double* G,
                       const double* scr, const double* lamR, double* lamG, double* lamRp) {
  hipLaunchKernelGGL(k_adj_stage2, dim3(grid_for(cap_local)), dim3(kBlock), 0, st, P, B, G, scr, lamR, lamG, lamRp);
}

void launch_adj_wm_flags(hipStream_t st, const DevParams& P, Buffers& B, int cap_fluid, double xp, int* flags,
                         const unsigned long long* hbits, int* argmax) {
  hipLaunchKernelGGL(k_adj_wm_flags, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, st, P, B, xp, flags, hbits, argmax);
}

void launch_adj_wm_reverse(hipStream_t st, const DevParams& P, Buffers& B, int n_total, int n_wm, const double* G,
                           const int* flags, const unsigned long long* hbits, double* lamG, double* c3,
                           double* part, const int* argmax, double* out2) {
  hipLaunchKernelGGL(k_adj_wm_share, dim3(grid_for(n_total)), dim3(kBlock), 0, st, P, B, G, flags, hbits, lamG, c3,
                     n_wm);
  hipLaunchKernelGGL(k_adj_wm_part, dim3(kRedBlocks), dim3(256), 0, st, c3, n_total, part);
  hipLaunchKernelGGL(k_adj_wm_reduce, dim3(1), dim3(kRedBlocks), 0, st, part, n_total, argmax, lamG, out2);
}

__global__ void k_adj_wm_init(unsigned long long* hbits, int* argmax) {
  if (threadIdx.x == 0) { *hbits = 0ull; *argmax = 0x7fffffff; }
}

void launch_adj_wm_init(hipStream_t st, unsigned long long* hbits, int* argmax) {
  hipLaunchKernelGGL(k_adj_wm_init, dim3(1), dim3(64), 0, st, hbits, argmax);
}

void launch_adj_probe(hipStream_t st, const DevParams& P, Buffers& B, double ex, double ey, double gravity,
                      const double* lam, double* lamG) {
  hipLaunchKernelGGL(k_adj_probe, dim3(1), dim3(64), 0, st, P, B, ex, ey, gravity, lam, lamG);
}

}  // namespace sph2d
