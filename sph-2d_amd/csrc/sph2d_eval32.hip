// sph2d_eval32.hip -- the FP32-mixed force evaluation.
//
// Same passes, order and pair terms as sph2d_eval64.hip, with the pair sums in
// FP32 and Kahan-compensated accumulation. What stays FP64: particle positions
// (the exact pair test 1e-7 <= r <= 3h in pass A, whose record of accepted
// neighbours the force pass replays), the integrator state and stage updates,
// the per-particle SPS closure and the Tait law: p and c of every particle are
// evaluated in FP64 from its FP64 density by the kernel that writes its records
// and stored rounded to FP32 (m_pc), so the pair sums never evaluate an EOS.
// Density enters the pair sums as delta rho = rho - rho0 (FP32), so rho_o - rho_s
// carries no cancellation. In the force pass the offset x_s - x_o comes from FP32
// offsets to each particle's cell centre (in y: the centre of its pair of cell
// rows) plus the whole-cell difference, which keeps it to ~1e-7 h; a neighbour
// costs 56 bytes (three float4 records and the (p, c) pair) there, 32 in pass A.
#include <cstring>
#include "sph2d_internal.h"
#include "sph2d_device.h"
#include "sph2d_riemann.h"

namespace sph2d {

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

// Kahan-compensated FP32 accumulator.
struct KSum {
  float s = 0.f, c = 0.f;
  __device__ __forceinline__ void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

struct DevParams32 {
  float inv_h, e9, kfactor, inv_sv, rho0, inv_c0, h, fai_h2, alpha, beta, two_nu;
  float delta_alpha_h, two_delta_h, shift_beta_h2, shift_R, inv_wdx, epsilon2, x_over_t;
  float cell_d, two_cd, inv_two_cd;
};

static DevParams32 make32(const DevParams& P) {
  DevParams32 Q;
  Q.inv_h = (float)P.inv_h;
  Q.e9 = (float)P.e9;
  Q.kfactor = (float)P.kfactor;
  Q.inv_sv = (float)P.inv_sv;
  Q.rho0 = (float)P.rho0;
  Q.inv_c0 = (float)P.inv_c0;
  Q.h = (float)P.h;
  Q.fai_h2 = (float)((P.fai * P.h) * (P.fai * P.h));
  Q.alpha = (float)P.alpha;
  Q.beta = (float)P.beta;
  Q.two_nu = (float)(2 * P.nu);
  Q.delta_alpha_h = (float)(P.delta_alpha * P.h);
  Q.two_delta_h = (float)(2 * P.delta * P.h);
  Q.shift_beta_h2 = (float)(P.shift_beta * P.h2);
  Q.shift_R = (float)P.shift_R;
  Q.inv_wdx = (float)P.inv_wdx;
  Q.epsilon2 = (float)(P.epsilon_xsph * P.xsph_f);
  Q.x_over_t = (float)P.x_over_t;
  Q.cell_d = (float)P.chain_d;
  Q.two_cd = (float)(2 * P.chain_d);
  Q.inv_two_cd = (float)(1.0 / (2 * P.chain_d));
  return Q;
}

struct KernelTerm32 {
  float w, dwx, dwy, t;   // grad W = t (dx, dy)
};

__device__ __forceinline__ KernelTerm32 kernel_term32(const DevParams32& Q, float dx, float dy, float r,
                                                      float inv_r) {
  KernelTerm32 k;
  const float q = r * Q.inv_h;
  const float eq = __expf(-(q * q));
  k.w = (eq - Q.e9) * Q.kfactor;
  const float t = eq * (-2.0f * q) * (inv_r * Q.inv_h) * Q.kfactor;
  k.t = t;
  k.dwx = t * dx;
  k.dwy = t * dy;
  return k;
}

// The next set bit of a 256-bit candidate mask (words m0..m3 of chunk c0), walked
// one word at a time: cur is the rest of word wi. Returns the candidate number, or
// -1 when the mask is spent; ascending order.
__device__ __forceinline__ int next_slot(uint64_t& cur, int& wi, uint64_t m1, uint64_t m2, uint64_t m3, int c0) {
  while (cur == 0) {
    if (wi >= 3) return -1;
    ++wi;
    cur = wi == 1 ? m1 : (wi == 2 ? m2 : m3);
  }
  const int b = __builtin_ctzll(cur);
  cur &= cur - 1;
  return c0 + 64 * wi + b;
}

// Pass A: the particle's own values, one survivor of the screen as loaded, and
// one survivor's terms.
struct A32Self {
  double2 ps;
  float2 vs;
  float drs, pres;
  bool fluid;
};

struct A32Rec {
  float4 r0;
  double2 pos;
  float p;
  int kk;
};

struct A32Terms {
  float fra, sdx, sdy, sxx, sxy, syy, grx, gry, gpx, gpy, gux, guy, gvx, gvy;
};

__device__ __forceinline__ A32Terms operator+(const A32Terms& a, const A32Terms& b) {
  return A32Terms{a.fra + b.fra, a.sdx + b.sdx, a.sdy + b.sdy, a.sxx + b.sxx, a.sxy + b.sxy,
                  a.syy + b.syy, a.grx + b.grx, a.gry + b.gry, a.gpx + b.gpx, a.gpy + b.gpy,
                  a.gux + b.gux, a.guy + b.guy, a.gvx + b.gvx, a.gvy + b.gvy};
}

// Pass A's sums, Kahan-compensated over the terms of one survivor or the FP32 sum
// of two (as the force pass's sums).
template <bool MUSCL>
struct A32Sums {
  KSum fra, sdx, sdy, Sxx, Sxy, Syy, grx, gry, gpx, gpy, gux, guy, gvx, gvy;
  __device__ __forceinline__ void add(const A32Terms& t) {
    fra.add(t.fra);
    sdx.add(t.sdx);
    sdy.add(t.sdy);
    Sxx.add(t.sxx);
    Sxy.add(t.sxy);
    Syy.add(t.syy);
    grx.add(t.grx);
    gry.add(t.gry);
    if (MUSCL) {
      gpx.add(t.gpx);
      gpy.add(t.gpy);
      gux.add(t.gux);
      guy.add(t.guy);
      gvx.add(t.gvx);
      gvy.add(t.gvy);
    }
  }
};

// One survivor of the FP32 screen: the exact FP64 pair test of find_particle_pair.f90
// :87 (acc), then the pass A terms of k_eval_a (sph2d_eval64.hip) in FP32. A survivor
// that fails the test, or a pair of two non-fluid particles, gets the volume zero,
// so every term is zero without a branch (r^2 is kept above 1e-30 so that a
// coincident survivor's terms stay finite).
template <bool MUSCL>
__device__ __forceinline__ A32Terms a32_terms(const DevParams& P, const DevParams32& Q, const A32Self& S,
                                              const A32Rec& N, bool& acc) {
  const double ddx = S.ps.x - N.pos.x, ddy = S.ps.y - N.pos.y;
  acc = is_pair_r2(P, ddx, ddy);
  const bool pair = acc && (S.fluid || N.r0.w >= 0.f);
  const float dx = (float)ddx, dy = (float)ddy;
  const float r2 = fmaxf(dx * dx + dy * dy, 1e-30f);
  const float inv_r = __builtin_amdgcn_rsqf(r2);
  const float r = r2 * inv_r;
  const KernelTerm32 K = kernel_term32(Q, dx, dy, r, inv_r);
  const float Vo = pair ? fabsf(N.r0.w) * __builtin_amdgcn_rcpf(Q.rho0 + N.r0.z) : 0.f;
  A32Terms T;
  T.fra = Vo * Q.inv_sv;
  T.sdx = K.dwx * Vo;
  T.sdy = K.dwy * Vo;
  const float vdx = N.r0.x - S.vs.x, vdy = N.r0.y - S.vs.y;
  T.sxx = vdx * K.dwx * Vo;
  T.sxy = (vdx * K.dwy + vdy * K.dwx) * Vo;
  T.syy = vdy * K.dwy * Vo;
  const float rd = (N.r0.z - S.drs) * Vo;
  T.grx = rd * K.dwx;
  T.gry = rd * K.dwy;
  T.gpx = T.gpy = T.gux = T.guy = T.gvx = T.gvy = 0.f;
  if (MUSCL) {
    const float pd = (N.p - S.pres) * Vo;
    T.gpx = pd * K.dwx;
    T.gpy = pd * K.dwy;
    const float ud = vdx * Vo, vd = vdy * Vo;
    T.gux = ud * K.dwx;
    T.guy = ud * K.dwy;
    T.gvx = vd * K.dwx;
    T.gvy = vd * K.dwy;
  }
  return T;
}

// pass A (mixed): see k_eval_a in sph2d_eval64.hip for the reference lines.
// CAP: candidate slots of the block's LDS stage (24 B each).
template <bool MUSCL, int CAP = kStageCap>
__device__ __forceinline__ void k_eval_a32_body(const DevParams& P, const DevParams32& Q, Buffers& B, int blk_, int nblk_) {
  __shared__ NbStage st;
  __shared__ float2 lds32[CAP];
  __shared__ double2 lds64[CAP];
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  stage_candidates(P, B.cell_start, B.l_cell, B.l_pos32, B.l_pos, n, st, lds32, lds64, blk_, CAP);
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const float4 r0s = B.m_r0[s];
  const float2 vs = make_float2(r0s.x, r0s.y);
  const float drs = r0s.z;
  const float pres = MUSCL ? B.m_pc[s].x : 0.f;   // FP64 Tait law, rounded to FP32
  const int ts = B.l_type[s];
  A32Sums<MUSCL> A;
  const __amdgpu_buffer_rsrc_t br0 = rec_rsrc(B.m_r0), bpos = rec_rsrc((const float4*)B.l_pos);
  const __amdgpu_buffer_rsrc_t bpc = rec_rsrc(reinterpret_cast<const float4*>(B.m_pc));
  // find_particle_pair.f90:64-123 written out (the shared neighbour loop records
  // the same slots): the FP32 screen of the 3x3 stencil, then the survivors in
  // ascending order, each with the exact FP64 test of :87. Two survivors' records
  // are in flight in turn (the next one's loads are issued before the current one's
  // test and sums, so their latency overlaps that work, and no loaded record is
  // copied between registers); the sums take the two survivors' terms together.
  // The accepted slots are recorded for the later passes (one byte each,
  // kNbOverflow when they do not fit).
  const float2 ps32 = B.l_pos32[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  const int n0 = R.hi[0] - R.lo[0], n1 = R.hi[1] - R.lo[1], n2 = R.hi[2] - R.lo[2];
  const int n01 = n0 + n1, total = n01 + n2;
  const int l0 = R.lo[0], l1 = R.lo[1] - n0, l2 = R.lo[2] - n01;
  const bool staged = st.use;
  const int so0 = staged ? l0 - st.lo[0] + st.base[0] : 0;
  const int so1 = staged ? l1 - st.lo[1] + st.base[1] : 0;
  const int so2 = staged ? l2 - st.lo[2] + st.base[2] : 0;
  const bool stored = total <= 256;
  uint32_t* const nbl = B.nb_list;
  const int lst = B.kstride;
  int nrec = 0;
  uint32_t word = 0;
  A32Self S;
  S.ps = ps;
  S.vs = vs;
  S.drs = drs;
  S.pres = pres;
  S.fluid = ts == 2;
#define SPH2D_A32_FETCH(R, KK)                                                              \
  do {                                                                                     \
    const int kk_ = (KK);                                                                  \
    const bool a_ = kk_ < n0, b_ = kk_ < n01;                                              \
    R.kk = kk_;                                                                            \
    const int o_ = kk_ + (a_ ? l0 : (b_ ? l1 : l2));                                       \
    R.r0 = buf_f4(br0, o_);                                                                \
    if (MUSCL) R.p = buf_f2(bpc, o_).x;                                                    \
    if (staged) R.pos = lds64[kk_ + (a_ ? so0 : (b_ ? so1 : so2))];                        \
    else R.pos = buf_d2(bpos, o_);                                                         \
  } while (0)
  // the accepted slot kk joins the record (ascending order)
#define SPH2D_A32_RECORD(KK)                                                                \
  do {                                                                                     \
    if (stored) {                                                                          \
      if (nrec < kNbListCap) {                                                             \
        word |= (uint32_t)(KK) << ((nrec & 3) << 3);                                       \
        if ((nrec & 3) == 3) {                                                             \
          nbl[(nrec >> 2) * lst + s] = word;                                               \
          word = 0;                                                                        \
        }                                                                                  \
      }                                                                                    \
      ++nrec;                                                                              \
    }                                                                                      \
  } while (0)
  for (int c0 = 0; c0 < total; c0 += 256) {
    const Mask256 sm = screen_chunk(P, B.l_pos32, s, ps32, R, n0, n1, n2, n01, total, l0, l1, l2, staged, so0, so1,
                                    so2, lds32, c0);
    uint64_t cur = sm.w0;
    int wi = 0;
    A32Rec ra, rb;
    int ka = next_slot(cur, wi, sm.w1, sm.w2, sm.w3, c0);
    if (ka >= 0) SPH2D_A32_FETCH(ra, ka);
    while (ka >= 0) {
      const int kb = next_slot(cur, wi, sm.w1, sm.w2, sm.w3, c0);
      if (kb >= 0) SPH2D_A32_FETCH(rb, kb);
      bool acc;
      const A32Terms ta = a32_terms<MUSCL>(P, Q, S, ra, acc);
      if (acc) SPH2D_A32_RECORD(ra.kk);
      if (kb < 0) {
        A.add(ta);
        break;
      }
      ka = next_slot(cur, wi, sm.w1, sm.w2, sm.w3, c0);
      if (ka >= 0) SPH2D_A32_FETCH(ra, ka);
      const A32Terms tb = a32_terms<MUSCL>(P, Q, S, rb, acc);
      if (acc) SPH2D_A32_RECORD(rb.kk);
      A.add(ta + tb);
    }
  }
#undef SPH2D_A32_FETCH
#undef SPH2D_A32_RECORD
  if (stored && (nrec & 3) && nrec < kNbListCap) nbl[(nrec >> 2) * lst + s] = word;
  B.nb_count[s] = nrec | ((!stored || nrec > kNbListCap) ? kNbOverflow : 0);
  const double f = (double)A.fra.s;
  B.fra_v[s] = f;
  B.sdw[s] = make_double2((double)A.sdx.s, (double)A.sdy.s);
  {
    const float4 r2s = B.m_r2[s];                  // keep the position offsets
    B.m_r2[s] = make_float4(A.grx.s, A.gry.s, r2s.z, r2s.w);
  }
  if (MUSCL) {
    B.m_mg[3 * s + 0] = make_float2(A.gpx.s, A.gpy.s);
    B.m_mg[3 * s + 1] = make_float2(A.gux.s, A.guy.s);
    B.m_mg[3 * s + 2] = make_float2(A.gvx.s, A.gvy.s);
    B.mg[3 * s + 0] = make_double2((double)A.gpx.s, (double)A.gpy.s);
    B.mg[3 * s + 1] = make_double2((double)A.gux.s, (double)A.guy.s);
    B.mg[3 * s + 2] = make_double2((double)A.gvx.s, (double)A.gvy.s);
  }
  double txx = 0.0, txy = 0.0, tyy = 0.0;
  if (ts == 2 && P.sps) {                          // SPS closure in FP64 (per particle)
    const double sxx = A.Sxx.s, sxy = A.Sxy.s, syy = A.Syy.s;
    const double rhos = (double)drs + P.rho0;
    const double master = sxx * sxx + syy * syy + 0.0;
    const double last = sxy * sxy + 0.0 + 0.0;
    const double abs2 = 2 * master + last;
    const double t0 = P.sps_c012 * P.sps_dl;
    const double nut = (t0 * t0) * sqrt(abs2);
    const double kk = P.sps_c23 * nut * (sxx + syy);
    const double blin = P.sps_c23b * (P.sps_dl * P.sps_dl) * abs2;
    txx = (2 * nut * sxx - kk - blin) * rhos;
    txy = (nut * sxy) * rhos;
    tyy = (2 * nut * syy - kk - blin) * rhos;
  }
  B.m_r1[s] = make_float4((float)txx, (float)txy, (float)tyy, B.m_r1[s].w);
  const int sf = (ts == 2 && f <= P.fra_limit) ? 1 : 0;
  B.surf[s] = sf;
  B.near_[s] = 0;
  if (sf) atomicAdd(&B.counters[C_NSURF], 1);
}

// Pass A waits on its loads more than on its arithmetic: without the MUSCL
// gradients it keeps 5 waves per SIMD (at most 96 registers, no spills), which hide
// that latency better than the 4 that 104 registers would leave; the MUSCL form
// needs its registers.
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
k_eval_a32(DevParams P, DevParams32 Q, Buffers B) {
  k_eval_a32_body<false>(P, Q, B, xcd_block(), gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_eval_a32_muscl(DevParams P, DevParams32 Q, Buffers B) {
  k_eval_a32_body<true>(P, Q, B, xcd_block(), gridDim.x);
}

// The particle's own values in the force pass.
struct C32Self {
  float2 vs, pcs, grs, gps, gus, gvs;
  float drs, rhos, inv_rhos, offx, offy;
  float4 taus;
  int s, ts, nes;
};

// One neighbour's contributions to the force pass's sums.
struct C32Terms {
  float drho, corr, ax, ay, shx, shy, agx, agy;
};

__device__ __forceinline__ C32Terms operator+(const C32Terms& a, const C32Terms& b) {
  return C32Terms{a.drho + b.drho, a.corr + b.corr, a.ax + b.ax, a.ay + b.ay,
                  a.shx + b.shx, a.shy + b.shy, a.agx + b.agx, a.agy + b.agy};
}

// The force pass's sums, Kahan-compensated. add() takes the contributions of one
// neighbour or the FP32 sum of two (the listed loop advances two neighbours at a
// time): the compensation then runs over neighbour pairs, and the one extra rounding
// of a two-term sum is of the order of each term's own FP32 rounding.
struct C32Sums {
  KSum drho, corr, ax, ay, shx, shy, agx, agy;
  __device__ __forceinline__ void add(const C32Terms& t, bool xsph, int do_shift) {
    drho.add(t.drho);
    corr.add(t.corr);
    ax.add(t.ax);
    ay.add(t.ay);
    if (xsph) { agx.add(t.agx); agy.add(t.agy); }
    if (do_shift) { shx.add(t.shx); shy.add(t.shy); }
  }
};

// A neighbour's records as the force pass loads them: r0 (u, v, delta rho, +-m),
// r1 (tau, m), r2 (grad rho, FP32 offsets), (p, c); its stencil column and slot.
struct C32Rec {
  float4 r0, r1, r2;
  float2 pc;
  int d, o;
};

// One recorded neighbour o (stencil column d) of the force pass: the pair terms of
// k_eval_c (sph2d_eval64.hip) in FP32. The offset x_s - x_o is the difference of the
// two particles' FP32 offsets to their cells' centres plus the whole-column
// difference (1 - d) cells in x; in y the offsets are taken to the centre of the pair
// of cell rows (rows 2k, 2k+1) that holds the particle, so the row difference aliases
// modulo two rows (8h) and the pair's |dy| <= 3h < 4h picks the one value, exactly.
//
// The modified linear Riemann solver's terms are written in the kernel gradient's
// scalar factor: with grad W = t (dx, dy), e = -(dx, dy) / r and v_ij.(dx, dy) = vx,
// U_L - U_R = -vx / r, (v_s - v*).grad W = t (vx / 2 + (U* - U_bar) r) (and t vx / 2
// near the surface, where v* = v_bar), (dx, dy).grad W / r^2 = t: the same terms as
// the reference's vector form, fewer operations, FP32 rounding in another order.
template <int SOLVER>
__device__ __forceinline__ C32Terms c32_terms(const DevParams32& Q, const Buffers& B, const C32Self& S,
                                              const C32Rec& N, int do_shift) {
  C32Terms T{0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  const float4 r0o = N.r0, r1o = N.r1, r2o = N.r2;
  const float2 pco = N.pc;
  const int o = N.o, d = N.d;
  // no pair between two non-fluid particles: such a pair's masses count as zero, so
  // every term below vanishes (no branch)
  const bool none = S.ts != 2 && r0o.w < 0.f;
  const float dx = (S.offx - r2o.z) + (float)(1 - d) * Q.cell_d;
  const float dyr = S.offy - r2o.w;
  const float dy = fmaf(-Q.two_cd, __builtin_rintf(dyr * Q.inv_two_cd), dyr);
  const float r2 = dx * dx + dy * dy;
  const float inv_r = __builtin_amdgcn_rsqf(r2);   // r2 >= 1e-14: no denormal scaling
  const float r = r2 * inv_r;
  const float inv_r2 = inv_r * inv_r;
  const float2 vs = S.vs, pcs = S.pcs;
  const float2 vo = make_float2(r0o.x, r0o.y);
  const float mo = none ? 0.f : fabsf(r0o.w);
  const float m1 = none ? 0.f : r1o.w;
  const float rhoo = Q.rho0 + r0o.z;
  const float inv_rhoo = __builtin_amdgcn_rcpf(rhoo);
  const float Vo = mo * inv_rhoo;
  const float vijx = vs.x - vo.x, vijy = vs.y - vo.y;
  const float vx = vijx * dx + vijy * dy;
  const float Pbar = (pcs.x + pco.x) * 0.5f;
  const float cbar = (pcs.y + pco.y) * 0.5f;
  const float rbar = Q.rho0 + (S.drs + r0o.z) * 0.5f;
  const float inv_rbar = __builtin_amdgcn_rcpf(rbar);    // v_rcp_f32, 1 ulp
  const float rij = S.inv_rhos * inv_rhoo;
  const KernelTerm32 K = kernel_term32(Q, dx, dy, r, inv_r);
  // artificial viscosity of approaching pairs (vx < 0); phi = 0 otherwise, which
  // adds zero: no branch
  const float phi = Q.h * (vx < 0.0f ? vx : 0.0f) * __builtin_amdgcn_rcpf(r2 + Q.fai_h2);
  const float Pi = (-(Q.alpha * cbar * phi) + Q.beta * (phi * phi)) * inv_rbar;
  const float dtmp = Q.delta_alpha_h * cbar * vx * inv_r2 * Q.rho0 * S.inv_rhos * Vo;
  float tx, ty;
  if (SOLVER == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN) {
    const float t = K.t;
    const float dU = -vx * inv_r;
    // near the surface v* = v_bar and P* = P_bar: U* - U_bar and the dissipation
    // are zero there (selects, no branch)
    const float dUs = S.nes ? 0.f : 0.5f * (pcs.x - pco.x) * inv_rbar * Q.inv_c0;   // U* - U_bar
    float bR = 3.0f * (dU > 0.0f ? dU : 0.0f);
    bR = bR < cbar ? bR : cbar;
    bR = S.nes ? 0.f : bR;
    const float Ps = Pbar + 0.5f * bR * rbar * dU;
    const float ti = t * (0.5f * vx + dUs * r);
    T.drho = 2.0f * S.rhos * Vo * ti;
    const float a = t * (mo * (-2.0f * Ps * rij) - mo * Pi + dtmp);
    tx = a * dx;
    ty = a * dy;
    if (Q.epsilon2 != 0.0f) {
      const float f = Q.epsilon2 * mo * __builtin_amdgcn_rcpf(S.rhos + rhoo) * K.w;
      T.agx = -f * vijx;
      T.agy = -f * vijy;
    }
  } else {
    const float ex = -(dx * inv_r), ey = -(dy * inv_r);
    const float UL = vs.x * ex + vs.y * ey;
    const float UR = vo.x * ex + vo.y * ey;
    const float vbx = (vs.x + vo.x) * 0.5f, vby = (vs.y + vo.y) * 0.5f;
    float vsx, vsy, Ps;
    float2 vown = vs;
    if (SOLVER == SPH2D_SOLVER_HLLC || SOLVER == SPH2D_SOLVER_MUSCL) {
      const bool left = S.s < o;                          // pair_i = lower local index
      const float epx = left ? ex : -ex, epy = left ? ey : -ey;
      const float2 vi = left ? vs : vo, vj = left ? vo : vs;
      const float2 pci = left ? pcs : pco, pcj = left ? pco : pcs;
      if (SOLVER == SPH2D_SOLVER_HLLC) {
        const float rhoi = left ? S.rhos : rhoo, rhoj = left ? rhoo : S.rhos;
        hllc_star<float>(vi.x, vi.y, pci.x, rhoi, pci.y, vj.x, vj.y, pcj.x, rhoj, pcj.y, epx, epy, Q.x_over_t,
                         vsx, vsy, Ps);
      } else {
        const float2* mgo = B.m_mg + 3 * o;
        const float2 gpo = mgo[0], guo = mgo[1], gvo = mgo[2];
        const float2 gpi = left ? S.gps : gpo, gpj = left ? gpo : S.gps;
        const float2 gui = left ? S.gus : guo, guj = left ? guo : S.gus;
        const float2 gvi = left ? S.gvs : gvo, gvj = left ? gvo : S.gvs;
        const float xrx = left ? -dx : dx, xry = left ? -dy : dy;
        float li, lj;
        muscl_limiter<float>(0.5f * (gpi.x * xrx + gpi.y * xry), -(0.5f * (gpj.x * xrx + gpj.y * xry)),
                             pcj.x - pci.x, li, lj);
        const float PL = pci.x + li, PR = pcj.x - lj;
        muscl_limiter<float>(0.5f * (gui.x * xrx + gui.y * xry), -(0.5f * (guj.x * xrx + guj.y * xry)),
                             vj.x - vi.x, li, lj);
        const float Iu = vi.x + li, Ju = vj.x - lj;
        muscl_limiter<float>(0.5f * (gvi.x * xrx + gvi.y * xry), -(0.5f * (gvj.x * xrx + gvj.y * xry)),
                             vj.y - vi.y, li, lj);
        const float Iv = vi.y + li, Jv = vj.y - lj;
        const float ULm = Iu * epx + Iv * epy;
        const float URm = Ju * epx + Jv * epy;
        const float dUs = 0.5f * (PL - PR) * inv_rbar * Q.inv_c0;   // U* - U_bar
        vsx = dUs * epx + vbx;
        vsy = dUs * epy + vby;
        const float dUp = ULm - URm;
        float bR = 3.0f * (dUp > 0.0f ? dUp : 0.0f);
        bR = bR < cbar ? bR : cbar;
        Ps = Pbar + 0.5f * bR * rbar * dUp;
        vown = left ? make_float2(Iu, Iv) : make_float2(Ju, Jv);
      }
    } else {                                           // SPH2D_SOLVER_LINEAR_RIEMANN
      const float dU = UL - UR;
      const float zs = S.rhos * pcs.y, zo = rhoo * pco.y;
      const float inv_z = __builtin_amdgcn_rcpf(zs + zo);
      const float Us = (zs * UR + zo * UL + (pcs.x - pco.x)) * inv_z;
      vsx = Us * ex;
      vsy = Us * ey;
      Ps = (zs * pco.x + zo * pcs.x + zs * zo * dU) * inv_z;
    }
    if (S.nes) { vsx = vbx; vsy = vby; Ps = Pbar; }
    const float ti = (vown.x - vsx) * K.dwx + (vown.y - vsy) * K.dwy;
    T.drho = 2.0f * S.rhos * Vo * ti;
    const float fp = mo * (-2.0f * Ps * rij);
    tx = fp * K.dwx + mo * (-Pi * K.dwx) + dtmp * K.dwx;
    ty = fp * K.dwy + mo * (-Pi * K.dwy) + dtmp * K.dwy;
  }
  const float4 taus = S.taus, tauo = r1o;
  float tp = (taus.x + tauo.x) * K.dwx + (taus.y + tauo.y) * K.dwy;
  tp = tp * rij;
  tx += m1 * tp;
  tp = tp + (taus.z + tauo.z) * K.dwy + (taus.y + tauo.y) * K.dwx;
  tp = tp * rij;
  ty += m1 * tp;
  const float lf = Vo * Q.two_nu * K.t;                 // (dx, dy).grad W / r^2 = t
  T.ax = tx + lf * vijx;
  T.ay = ty + lf * vijy;
  const float psi = (r0o.z - S.drs) + 0.5f * ((S.grs.x + r2o.x) * dx + (S.grs.y + r2o.y) * dy);
  T.corr = -(Q.two_delta_h * cbar * psi * K.t * Vo);
  if (do_shift) {
    const float wr = K.w * Q.inv_wdx;
    const float f = Q.shift_beta_h2 * (1.0f + Q.shift_R * (wr * wr * wr * wr)) * Vo;
    T.shx = -f * K.dwx;
    T.shy = -f * K.dwy;
  }
  return T;
}

// pass C (mixed): see k_eval_c in sph2d_eval64.hip for the reference lines.
// A block whose candidate union fits kRecCap copies the neighbour records r0 and r2
// to LDS once and reads them there; r1 (tau, read only by the SPS term) comes from
// L2, which keeps the block at 34 KB of LDS (4 blocks per CU instead of 3). Blocks
// whose union does not fit read every record from global memory.
constexpr int kRecCap = 1088;

// CAP: candidate slots of the block's LDS record stage (32 B each).
template <int SOLVER, int CAP = kRecCap>
__device__ __forceinline__ void k_eval_c32_body(const DevParams& P, const DevParams32& Q, Buffers& B, int do_shift, int blk_, int nblk_) {
  __shared__ NbStage st;
  __shared__ float4 rec0[CAP], rec2[CAP];
  const int blk = blk_;
  const int s = blk * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  plan_stage(P, B.cell_start, B.l_cell, n, CAP, st, blk);
  for (int i = threadIdx.x; i < st.total; i += blockDim.x) {
    const int o = stage_source(st, i);
    rec0[i] = B.m_r0[o];
    rec2[i] = B.m_r2[o];
  }
  __syncthreads();
  const bool staged = st.use;                      // uniform over the block
  if (s >= n) return;
  C32Self S;
  const float4 r0s = B.m_r0[s];
  const float4 r2s = B.m_r2[s];
  S.s = s;
  S.vs = make_float2(r0s.x, r0s.y);
  S.drs = r0s.z;
  S.taus = B.m_r1[s];
  S.pcs = B.m_pc[s];                               // FP64 Tait law, rounded to FP32
  S.ts = B.l_type[s];
  S.grs = make_float2(r2s.x, r2s.y);
  S.offx = r2s.z;
  S.offy = r2s.w;
  S.nes = B.near_[s];
  S.rhos = Q.rho0 + S.drs;
  S.inv_rhos = 1.0f / S.rhos;
  constexpr bool kMuscl = SOLVER == SPH2D_SOLVER_MUSCL;
  S.gps = make_float2(0.f, 0.f); S.gus = S.gps; S.gvs = S.gps;
  if (kMuscl) { S.gps = B.m_mg[3 * s]; S.gus = B.m_mg[3 * s + 1]; S.gvs = B.m_mg[3 * s + 2]; }
  C32Sums A;
  // Buffer-resource views of the records for blocks without an LDS copy: buffer and
  // LDS loads are different instructions, never merged into one flat load of a
  // selected address (flat loads of LDS data are slower than ds_read)
  const __amdgpu_buffer_rsrc_t br0 = rec_rsrc(B.m_r0), br1 = rec_rsrc(B.m_r1), br2 = rec_rsrc(B.m_r2);
  const __amdgpu_buffer_rsrc_t bpc = rec_rsrc(reinterpret_cast<const float4*>(B.m_pc));
  // The neighbours recorded by pass A (for_each_neighbour<kNbReplay>, written out:
  // a closure over this kernel's locals would be kept in scratch memory).
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  const int n0 = R.hi[0] - R.lo[0], n01 = n0 + R.hi[1] - R.lo[1], total = n01 + R.hi[2] - R.lo[2];
  const int l0 = R.lo[0], l1 = R.lo[1] - n0, l2 = R.lo[2] - n01;
  const int so0 = staged ? l0 - st.lo[0] + st.base[0] : 0;
  const int so1 = staged ? l1 - st.lo[1] + st.base[1] : 0;
  const int so2 = staged ? l2 - st.lo[2] + st.base[2] : 0;
  // pass A's record of the accepted neighbours (one byte per neighbour, in ascending
  // candidate order); a list longer than 256 candidates or more neighbours than the
  // record holds are screened and tested again, 256 candidates at a time.
  // The slots come one ahead of the pair they feed: each iteration takes the next
  // slot and issues its record loads before the current pair's arithmetic, so the
  // L2 and LDS latency of the loads overlaps the work of the pair before.
  const int cnt = B.nb_count[s];
  const bool listed = total <= 256 && !(cnt & kNbOverflow);
  const bool xsph = SOLVER == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && Q.epsilon2 != 0.0f;
#define SPH2D_C32_FETCH(R, KK)                                                             \
  do {                                                                                    \
    const int kk_ = (KK);                                                                 \
    const bool a_ = kk_ < n0, b_ = kk_ < n01;                                             \
    R.d = a_ ? 0 : (b_ ? 1 : 2);                                                          \
    R.o = kk_ + (a_ ? l0 : (b_ ? l1 : l2));                                               \
    const int lin_ = kk_ + (a_ ? so0 : (b_ ? so1 : so2));                                 \
    R.r1 = buf_f4(br1, R.o);                                                              \
    R.pc = buf_f2(bpc, R.o);                                                              \
    if (staged) { R.r0 = rec0[lin_]; R.r2 = rec2[lin_]; }                                 \
    else { R.r0 = buf_f4(br0, R.o); R.r2 = buf_f4(br2, R.o); }                            \
  } while (0)
  if (listed) {
    // two record sets in turn (the loop advances two neighbours), so the loads in
    // flight never need copying between registers
    const uint32_t* const nbl = B.nb_list;
    const int lst = B.kstride;
    uint32_t word = 0, word_nx = 0;                // the list word in use and the next one
    if (cnt > 0) word = nbl[s];
    if (cnt > 4) word_nx = nbl[lst + s];
    C32Rec ra, rb;
    if (cnt > 0) SPH2D_C32_FETCH(ra, (int)(word & 255u));
    for (int t = 0; t < cnt; t += 2) {
      if (t + 1 < cnt) SPH2D_C32_FETCH(rb, (int)((word >> (((t + 1) & 3) << 3)) & 255u));
      const C32Terms ta = c32_terms<SOLVER>(Q, B, S, ra, do_shift);
      if (t + 1 >= cnt) {
        A.add(ta, xsph, do_shift);
        break;
      }
      if (t + 2 < cnt) {
        if (((t + 2) & 3) == 0) {
          word = word_nx;
          if (t + 6 < cnt) word_nx = nbl[(((t + 2) >> 2) + 1) * lst + s];
        }
        SPH2D_C32_FETCH(ra, (int)((word >> (((t + 2) & 3) << 3)) & 255u));
      }
      A.add(ta + c32_terms<SOLVER>(Q, B, S, rb, do_shift), xsph, do_shift);
    }
  } else {
    C32Rec rn;
    for (int c0 = 0; c0 < total; c0 += 256) {
      const Mask256 mk = exact_chunk_mask(P, B.l_pos, B.l_pos32, s, B.l_pos[s], n0, n01, l0, l1, l2, total, c0);
      uint64_t cur = mk.w0;
      int wi = 0;
      int kk_n = next_slot(cur, wi, mk.w1, mk.w2, mk.w3, c0);
      if (kk_n >= 0) SPH2D_C32_FETCH(rn, kk_n);
      while (kk_n >= 0) {
        const C32Rec rc = rn;
        kk_n = next_slot(cur, wi, mk.w1, mk.w2, mk.w3, c0);
        if (kk_n >= 0) SPH2D_C32_FETCH(rn, kk_n);
        A.add(c32_terms<SOLVER>(Q, B, S, rc, do_shift), xsph, do_shift);
      }
    }
  }
#undef SPH2D_C32_FETCH
  const int ts = S.ts;
  double dr = (double)A.drho.s;
  if (ts == 2) dr = dr + (double)A.corr.s;
  B.drhodt[s] = dr;
  B.acc[s] = ts == 2 ? make_double2((double)A.ax.s, (double)A.ay.s - P.g) : make_double2(0.0, 0.0);
  if (P.epsilon_xsph != 0.0) B.avg[s] = make_double2((double)A.agx.s, (double)A.agy.s);
  if (do_shift) B.shift[s] = make_double2((double)A.shx.s, (double)A.shy.s);
}

template <int SOLVER>
__global__ void __launch_bounds__(kBlock) k_eval_c32(DevParams P, DevParams32 Q, Buffers B, int do_shift) {
  k_eval_c32_body<SOLVER>(P, Q, B, do_shift, xcd_block(), gridDim.x);
}

void launch_eval32_a(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  const int g = grid_for(cap_local);
  if (P.solver == SPH2D_SOLVER_MUSCL)
    hipLaunchKernelGGL(k_eval_a32_muscl, dim3(g), dim3(kBlock), 0, s, P, make32(P), B);
  else
    hipLaunchKernelGGL(k_eval_a32, dim3(g), dim3(kBlock), 0, s, P, make32(P), B);
}

void launch_eval32_c(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  const int g = grid_for(cap_local);
  const DevParams32 Q = make32(P);
  switch (P.solver) {
    case SPH2D_SOLVER_LINEAR_RIEMANN:
      hipLaunchKernelGGL(k_eval_c32<SPH2D_SOLVER_LINEAR_RIEMANN>, dim3(g), dim3(kBlock), 0, s, P, Q, B, 1);
      break;
    case SPH2D_SOLVER_HLLC:
      hipLaunchKernelGGL(k_eval_c32<SPH2D_SOLVER_HLLC>, dim3(g), dim3(kBlock), 0, s, P, Q, B, 1);
      break;
    case SPH2D_SOLVER_MUSCL:
      hipLaunchKernelGGL(k_eval_c32<SPH2D_SOLVER_MUSCL>, dim3(g), dim3(kBlock), 0, s, P, Q, B, 1);
      break;
    default:
      hipLaunchKernelGGL(k_eval_c32<SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN>, dim3(g), dim3(kBlock), 0, s, P, Q, B, 1);
  }
}

size_t q32_bytes() { return sizeof(DevParams32); }
void q32_fill(const DevParams& P, void* out) {
  const DevParams32 Q = make32(P);
  std::memcpy(out, &Q, sizeof Q);
}

// Batched mixed passes A and C (see launch_batch_step), in the tile shape SHAPE of
// kNbShape: BLK consecutive particles per block and the LDS stage capacities.
template <int SHAPE>
__global__ void __launch_bounds__(kNbShape[SHAPE].block) __attribute__((amdgpu_waves_per_eu(5)))
kb_eval_a32(BatchArgs A, const DevParams32* Qs) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  const DevParams32 Q = Qs[w.i];
  Buffers B = A.inst[w.i].B;
  k_eval_a32_body<false, kNbShape[SHAPE].cap_a>(P, Q, B, w.blk, w.nblk);
}

template <int SHAPE>
__global__ void __launch_bounds__(kNbShape[SHAPE].block) kb_eval_a32_muscl(BatchArgs A, const DevParams32* Qs) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  const DevParams32 Q = Qs[w.i];
  Buffers B = A.inst[w.i].B;
  k_eval_a32_body<true, kNbShape[SHAPE].cap_a>(P, Q, B, w.blk, w.nblk);
}

template <int SOLVER, int SHAPE>
__global__ void __launch_bounds__(kNbShape[SHAPE].block) kb_eval_c32(BatchArgs A, const DevParams32* Qs) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  const DevParams32 Q = Qs[w.i];
  Buffers B = A.inst[w.i].B;
  k_eval_c32_body<SOLVER, kNbShape[SHAPE].cap_c>(P, Q, B, 1, w.blk, w.nblk);
}

template <int SHAPE>
static void launch_batch_eval32_shape(hipStream_t s, const BatchDev& D, int pass, int solver) {
  const int g = D.total[BC_NBR];
  const dim3 t(kNbShape[SHAPE].block);
  const BatchArgs A{D.inst, D.start[BC_NBR], D.n, D.xcd};
  const auto* Qs = static_cast<const DevParams32*>(D.q32);
  if (pass == 0) {
    if (solver == SPH2D_SOLVER_MUSCL) hipLaunchKernelGGL((kb_eval_a32_muscl<SHAPE>), dim3(g), t, 0, s, A, Qs);
    else hipLaunchKernelGGL((kb_eval_a32<SHAPE>), dim3(g), t, 0, s, A, Qs);
    return;
  }
  switch (solver) {
    case SPH2D_SOLVER_LINEAR_RIEMANN:
      hipLaunchKernelGGL((kb_eval_c32<SPH2D_SOLVER_LINEAR_RIEMANN, SHAPE>), dim3(g), t, 0, s, A, Qs);
      break;
    case SPH2D_SOLVER_HLLC:
      hipLaunchKernelGGL((kb_eval_c32<SPH2D_SOLVER_HLLC, SHAPE>), dim3(g), t, 0, s, A, Qs);
      break;
    case SPH2D_SOLVER_MUSCL:
      hipLaunchKernelGGL((kb_eval_c32<SPH2D_SOLVER_MUSCL, SHAPE>), dim3(g), t, 0, s, A, Qs);
      break;
    default:
      hipLaunchKernelGGL((kb_eval_c32<SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN, SHAPE>), dim3(g), t, 0, s, A, Qs);
  }
}

void launch_batch_eval32(hipStream_t s, const BatchDev& D, int pass, int solver) {
  if (D.total[BC_NBR] <= 0) return;
  switch (D.shape) {
    case 1: launch_batch_eval32_shape<1>(s, D, pass, solver); break;
    case 2: launch_batch_eval32_shape<2>(s, D, pass, solver); break;
    case 3: launch_batch_eval32_shape<3>(s, D, pass, solver); break;
    default: launch_batch_eval32_shape<0>(s, D, pass, solver);
  }
}

}  // namespace sph2d
