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
// offsets to each particle's cell corner plus the whole-cell difference, which
// keeps it to ~1e-7 h; a neighbour costs 56 bytes (three float4 records and the
// (p, c) pair) there, 32 in pass A.
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
  float cell_d;
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
  return Q;
}

struct KernelTerm32 {
  float w, dwx, dwy;
};

__device__ __forceinline__ KernelTerm32 kernel_term32(const DevParams32& Q, float dx, float dy, float r,
                                                      float inv_r) {
  KernelTerm32 k;
  const float q = r * Q.inv_h;
  const float eq = __expf(-(q * q));
  k.w = (eq - Q.e9) * Q.kfactor;
  const float t = eq * (-2.0f * q) * (inv_r * Q.inv_h) * Q.kfactor;
  k.dwx = t * dx;
  k.dwy = t * dy;
  return k;
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
  KSum fra, sdx, sdy, Sxx, Sxy, Syy, grx, gry;
  KSum gpx, gpy, gux, guy, gvx, gvy;
  const __amdgpu_buffer_rsrc_t br0 = rec_rsrc(B.m_r0), bpos = rec_rsrc((const float4*)B.l_pos);
  const __amdgpu_buffer_rsrc_t bpc = rec_rsrc(reinterpret_cast<const float4*>(B.m_pc));
  // find_particle_pair.f90:64-123 written out (the shared neighbour loop records
  // the same slots): the FP32 screen of the 3x3 stencil, then the survivors in
  // ascending order, each with the exact FP64 test of :87; the next survivor's
  // loads are issued before the current survivor's test and sums, so their
  // latency overlaps that work. The accepted slots are recorded for the later
  // passes (one byte each, kNbOverflow when they do not fit).
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
  for (int c0 = 0; c0 < total; c0 += 256) {
    const Mask256 sm = screen_chunk(P, B.l_pos32, s, ps32, R, n0, n1, n2, n01, total, l0, l1, l2, staged, so0, so1,
                                    so2, lds32, c0);
    uint64_t m0 = sm.w0, m1 = sm.w1, m2 = sm.w2, m3 = sm.w3;
    int kk_n = -1;
    if ((m0 | m1 | m2 | m3) != 0) {
      const int w = m0 ? 0 : (m1 ? 1 : (m2 ? 2 : 3));
      const uint64_t mm = m0 ? m0 : (m1 ? m1 : (m2 ? m2 : m3));
      const int b = __builtin_ctzll(mm);
      const uint64_t bit = 1ull << b;
      if (w == 0) m0 ^= bit; else if (w == 1) m1 ^= bit; else if (w == 2) m2 ^= bit; else m3 ^= bit;
      kk_n = c0 + w * 64 + b;
    }
    float4 r0n = make_float4(0.f, 0.f, 0.f, 0.f);
    double2 pn = make_double2(0.0, 0.0);
    if (kk_n >= 0) {
      const int dn = kk_n < n0 ? 0 : (kk_n < n01 ? 1 : 2);
      const int on = dn == 0 ? l0 + kk_n : (dn == 1 ? l1 + kk_n : l2 + kk_n);
      r0n = buf_f4(br0, on);
      if (staged) pn = lds64[dn == 0 ? so0 + kk_n : (dn == 1 ? so1 + kk_n : so2 + kk_n)];
      else pn = buf_d2(bpos, on);
    }
    while (kk_n >= 0) {
      const int kk = kk_n;
      const int d = kk < n0 ? 0 : (kk < n01 ? 1 : 2);
      const int o = d == 0 ? l0 + kk : (d == 1 ? l1 + kk : l2 + kk);
      const float4 r0o = r0n;
      const double2 po = pn;
      kk_n = -1;
      if ((m0 | m1 | m2 | m3) != 0) {
        const int w = m0 ? 0 : (m1 ? 1 : (m2 ? 2 : 3));
        const uint64_t mm = m0 ? m0 : (m1 ? m1 : (m2 ? m2 : m3));
        const int b = __builtin_ctzll(mm);
        const uint64_t bit = 1ull << b;
        if (w == 0) m0 ^= bit; else if (w == 1) m1 ^= bit; else if (w == 2) m2 ^= bit; else m3 ^= bit;
        kk_n = c0 + w * 64 + b;
        const int dn = kk_n < n0 ? 0 : (kk_n < n01 ? 1 : 2);
        const int on = dn == 0 ? l0 + kk_n : (dn == 1 ? l1 + kk_n : l2 + kk_n);
        r0n = buf_f4(br0, on);
        if (staged) pn = lds64[dn == 0 ? so0 + kk_n : (dn == 1 ? so1 + kk_n : so2 + kk_n)];
        else pn = buf_d2(bpos, on);
      }
      if (!is_pair_r2(P, ps.x - po.x, ps.y - po.y)) continue;
      if (stored) {
        if (nrec < kNbListCap) {
          word |= (uint32_t)kk << ((nrec & 3) << 3);
          if ((nrec & 3) == 3) {
            nbl[(nrec >> 2) * lst + s] = word;
            word = 0;
          }
        }
        ++nrec;
      }
      if (ts != 2 && r0o.w < 0.f) continue;        // no pair between two non-fluid particles
      const float dx = (float)(ps.x - po.x), dy = (float)(ps.y - po.y);
      const float r2 = dx * dx + dy * dy;
      const float inv_r = __builtin_amdgcn_rsqf(r2);   // r2 >= 1e-14: no denormal scaling
      const float r = r2 * inv_r;
      const KernelTerm32 K = kernel_term32(Q, dx, dy, r, inv_r);
      const float2 rvo = make_float2(r0o.z, fabsf(r0o.w) * __builtin_amdgcn_rcpf(Q.rho0 + r0o.z));
      const float2 vo = make_float2(r0o.x, r0o.y);
      const float Vo = rvo.y;
      fra.add(Vo * Q.inv_sv);
      sdx.add(K.dwx * Vo);
      sdy.add(K.dwy * Vo);
      const float vdx = vo.x - vs.x, vdy = vo.y - vs.y;
      Sxx.add(vdx * K.dwx * Vo);
      Sxy.add((vdx * K.dwy + vdy * K.dwx) * Vo);
      Syy.add(vdy * K.dwy * Vo);
      const float rd = (rvo.x - drs) * Vo;
      grx.add(rd * K.dwx);
      gry.add(rd * K.dwy);
      if (MUSCL) {
        const float pd = (buf_f2(bpc, o).x - pres) * Vo;
        gpx.add(pd * K.dwx);
        gpy.add(pd * K.dwy);
        const float ud = vdx * Vo, vd = vdy * Vo;
        gux.add(ud * K.dwx);
        guy.add(ud * K.dwy);
        gvx.add(vd * K.dwx);
        gvy.add(vd * K.dwy);
      }
    }
  }
  if (stored && (nrec & 3) && nrec < kNbListCap) nbl[(nrec >> 2) * lst + s] = word;
  B.nb_count[s] = nrec | ((!stored || nrec > kNbListCap) ? kNbOverflow : 0);
  const double f = (double)fra.s;
  B.fra_v[s] = f;
  B.sdw[s] = make_double2((double)sdx.s, (double)sdy.s);
  {
    const float4 r2s = B.m_r2[s];                  // keep the position offsets
    B.m_r2[s] = make_float4(grx.s, gry.s, r2s.z, r2s.w);
  }
  if (MUSCL) {
    B.m_mg[3 * s + 0] = make_float2(gpx.s, gpy.s);
    B.m_mg[3 * s + 1] = make_float2(gux.s, guy.s);
    B.m_mg[3 * s + 2] = make_float2(gvx.s, gvy.s);
    B.mg[3 * s + 0] = make_double2((double)gpx.s, (double)gpy.s);
    B.mg[3 * s + 1] = make_double2((double)gux.s, (double)guy.s);
    B.mg[3 * s + 2] = make_double2((double)gvx.s, (double)gvy.s);
  }
  double txx = 0.0, txy = 0.0, tyy = 0.0;
  if (ts == 2 && P.sps) {                          // SPS closure in FP64 (per particle)
    const double sxx = Sxx.s, sxy = Sxy.s, syy = Syy.s;
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

template <bool MUSCL>
__global__ void __launch_bounds__(kBlock) k_eval_a32(DevParams P, DevParams32 Q, Buffers B) {
  k_eval_a32_body<MUSCL>(P, Q, B, xcd_block(), gridDim.x);
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
  const float4 r0s = B.m_r0[s];
  const float4 taus = B.m_r1[s];
  const float4 r2s = B.m_r2[s];
  const float2 vs = make_float2(r0s.x, r0s.y);
  const float2 rvs = make_float2(r0s.z, 0.f);
  const float2 pcs = B.m_pc[s];                    // FP64 Tait law, rounded to FP32
  const int ts = B.l_type[s];
  const float2 grs = make_float2(r2s.x, r2s.y);
  const int nes = B.near_[s];
  const float rhos = Q.rho0 + rvs.x;
  const float inv_rhos = 1.0f / rhos;
  // Cell rows of the stencil columns: a neighbour's row (hence its cell corner
  // relative to ours) follows from where its local index falls in its column.
  const int cell_s = B.l_cell[s];
  const int ny = P.chain_ny;
  const int cxs = cell_s / ny, cys = cell_s - cxs * ny;
  const int row0 = cys > 0 ? cys - 1 : 0;
  const int row1 = cys + 1 < ny ? cys + 1 : ny - 1;
  // first local index of rows row0+1 and row0+2 in stencil columns 0, 1, 2 (none: INT_MAX)
  const bool c0ok = cxs >= 1, c2ok = cxs + 1 < P.chain_nx;
  const bool r1ok = row0 + 1 <= row1, r2ok = row0 + 2 <= row1;
  const int* const cs = B.cell_start;
  const int rb10 = c0ok && r1ok ? cs[(cxs - 1) * ny + row0 + 1] : 0x7fffffff;
  const int rb11 = r1ok ? cs[cxs * ny + row0 + 1] : 0x7fffffff;
  const int rb12 = c2ok && r1ok ? cs[(cxs + 1) * ny + row0 + 1] : 0x7fffffff;
  const int rb20 = c0ok && r2ok ? cs[(cxs - 1) * ny + row0 + 2] : 0x7fffffff;
  const int rb21 = r2ok ? cs[cxs * ny + row0 + 2] : 0x7fffffff;
  const int rb22 = c2ok && r2ok ? cs[(cxs + 1) * ny + row0 + 2] : 0x7fffffff;
  constexpr bool kMuscl = SOLVER == SPH2D_SOLVER_MUSCL;
  float2 gps = make_float2(0.f, 0.f), gus = gps, gvs = gps;
  if (kMuscl) { gps = B.m_mg[3 * s]; gus = B.m_mg[3 * s + 1]; gvs = B.m_mg[3 * s + 2]; }
  KSum drho, corr, ax, ay, shx, shy, agx, agy;
  // Buffer-resource views of the records for blocks without an LDS copy: buffer and
  // LDS loads are different instructions, never merged into one flat load of a
  // selected address (flat loads of LDS data are slower than ds_read)
  const __amdgpu_buffer_rsrc_t br0 = rec_rsrc(B.m_r0), br1 = rec_rsrc(B.m_r1), br2 = rec_rsrc(B.m_r2);
  const __amdgpu_buffer_rsrc_t bpc = rec_rsrc(reinterpret_cast<const float4*>(B.m_pc));
  // The neighbours recorded by pass A (for_each_neighbour<kNbReplay>, written out:
  // a closure over this kernel's locals would be kept in scratch memory).
  const CellRanges R = cell_ranges(P, B.cell_start, cell_s);
  const int n0 = R.hi[0] - R.lo[0], n01 = n0 + R.hi[1] - R.lo[1], total = n01 + R.hi[2] - R.lo[2];
  const int l0 = R.lo[0], l1 = R.lo[1] - n0, l2 = R.lo[2] - n01;
  const int so0 = staged ? l0 - st.lo[0] + st.base[0] : 0;
  const int so1 = staged ? l1 - st.lo[1] + st.base[1] : 0;
  const int so2 = staged ? l2 - st.lo[2] + st.base[2] : 0;
  // pass A's record of the accepted neighbours (one byte per neighbour, in ascending
  // candidate order); a list longer than 256 candidates or more neighbours than the
  // record holds are screened and tested again, 256 candidates at a time
  const int cnt = B.nb_count[s];
  const bool listed = total <= 256 && !(cnt & kNbOverflow);
  const uint32_t* const nbl = B.nb_list;
  const int lst = B.kstride;
  for (int c0 = 0; c0 < total; c0 += 256) {
  uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  if (!listed) {
    const Mask256 mk = exact_chunk_mask(P, B.l_pos, B.l_pos32, s, B.l_pos[s], n0, n01, l0, l1, l2, total, c0);
    m0 = mk.w0; m1 = mk.w1; m2 = mk.w2; m3 = mk.w3;
  }
  // The slots come one ahead of the pair they feed: each iteration takes the
  // next slot and issues its record loads before the current pair's arithmetic,
  // so the L2 and LDS latency of the loads overlaps the work of the pair before.
  uint32_t word = 0, word_nx = 0;                  // the list word in use and the next one
  int t = 0;
  int kk_n = -1;
  if (listed) {
    if (cnt > 0) { word = nbl[s]; kk_n = (int)(word & 255u); }
    if (cnt > 4) word_nx = nbl[lst + s];
  } else if ((m0 | m1 | m2 | m3) != 0) {
    const int w = m0 ? 0 : (m1 ? 1 : (m2 ? 2 : 3));
    const uint64_t mm = m0 ? m0 : (m1 ? m1 : (m2 ? m2 : m3));
    const int b = __builtin_ctzll(mm);
    const uint64_t bit = 1ull << b;
    if (w == 0) m0 ^= bit; else if (w == 1) m1 ^= bit; else if (w == 2) m2 ^= bit; else m3 ^= bit;
    kk_n = c0 + w * 64 + b;
  }
  float4 r0n = make_float4(0.f, 0.f, 0.f, 0.f), r1n = r0n, r2n = r0n;
  float2 pcn = make_float2(0.f, 0.f);
  int dn = 0, on = 0;
  if (kk_n >= 0) {
    dn = kk_n < n0 ? 0 : (kk_n < n01 ? 1 : 2);
    on = dn == 0 ? l0 + kk_n : (dn == 1 ? l1 + kk_n : l2 + kk_n);
    const int lin = dn == 0 ? so0 + kk_n : (dn == 1 ? so1 + kk_n : so2 + kk_n);
    r1n = buf_f4(br1, on);
    pcn = buf_f2(bpc, on);
    if (staged) { r0n = rec0[lin]; r2n = rec2[lin]; }
    else { r0n = buf_f4(br0, on); r2n = buf_f4(br2, on); }
  }
  while (kk_n >= 0) {
    const int d = dn, o = on;
    const float4 r0o = r0n, r1o = r1n, r2o = r2n;
    const float2 pco = pcn;
    ++t;
    kk_n = -1;
    if (listed) {
      if (t < cnt) {
        if ((t & 3) == 0) {
          word = word_nx;
          if (t + 4 < cnt) word_nx = nbl[((t >> 2) + 1) * lst + s];
        }
        kk_n = (int)((word >> ((t & 3) << 3)) & 255u);
      }
    } else if ((m0 | m1 | m2 | m3) != 0) {
      const int w = m0 ? 0 : (m1 ? 1 : (m2 ? 2 : 3));
      const uint64_t mm = m0 ? m0 : (m1 ? m1 : (m2 ? m2 : m3));
      const int b = __builtin_ctzll(mm);
      const uint64_t bit = 1ull << b;
      if (w == 0) m0 ^= bit; else if (w == 1) m1 ^= bit; else if (w == 2) m2 ^= bit; else m3 ^= bit;
      kk_n = c0 + w * 64 + b;
    }
    if (kk_n >= 0) {
      dn = kk_n < n0 ? 0 : (kk_n < n01 ? 1 : 2);
      on = dn == 0 ? l0 + kk_n : (dn == 1 ? l1 + kk_n : l2 + kk_n);
      const int lin = dn == 0 ? so0 + kk_n : (dn == 1 ? so1 + kk_n : so2 + kk_n);
      r1n = buf_f4(br1, on);
      pcn = buf_f2(bpc, on);
      if (staged) { r0n = rec0[lin]; r2n = rec2[lin]; }
      else { r0n = buf_f4(br0, on); r2n = buf_f4(br2, on); }
    }
    if (ts != 2 && r0o.w < 0.f) continue;        // no pair between two non-fluid particles
    // bitwise blends, not selects: a select of captured values can become a
    // select of their addresses, which keeps the closure in scratch memory
    const int e0 = -(d == 0), e1 = -(d == 1), e2 = -(d == 2);
    const int b1 = (rb10 & e0) | (rb11 & e1) | (rb12 & e2);
    const int b2 = (rb20 & e0) | (rb21 & e1) | (rb22 & e2);
    const int dcy = cys - (row0 + (o >= b1) + (o >= b2));   // cy_s - cy_o
    const float dx = (r2s.z - r2o.z) + (float)(1 - d) * Q.cell_d;
    const float dy = (r2s.w - r2o.w) + (float)dcy * Q.cell_d;
    const float r2 = dx * dx + dy * dy;
    const float inv_r = __builtin_amdgcn_rsqf(r2);   // r2 >= 1e-14: no denormal scaling
    const float r = r2 * inv_r;
    const float inv_r2 = inv_r * inv_r;
    const KernelTerm32 K = kernel_term32(Q, dx, dy, r, inv_r);
    const float2 vo = make_float2(r0o.x, r0o.y);
    const float mo_x = fabsf(r0o.w);
    const float rhoo = Q.rho0 + r0o.z;
    const float2 rvo = make_float2(r0o.z, mo_x * __builtin_amdgcn_rcpf(rhoo));
    const float Vo = rvo.y;
    const float ex = -(dx * inv_r), ey = -(dy * inv_r);
    const float UL = vs.x * ex + vs.y * ey;
    const float UR = vo.x * ex + vo.y * ey;
    const float dU = UL - UR;
    const float Pbar = (pcs.x + pco.x) * 0.5f;
    const float vbx = (vs.x + vo.x) * 0.5f, vby = (vs.y + vo.y) * 0.5f;
    const float cbar = (pcs.y + pco.y) * 0.5f;
    const float rbar = Q.rho0 + (rvs.x + rvo.x) * 0.5f;
    const float inv_rbar = __builtin_amdgcn_rcpf(rbar);    // v_rcp_f32, 1 ulp
    float vsx, vsy, Ps;
    float2 vown = vs;
    if (SOLVER == SPH2D_SOLVER_HLLC || SOLVER == SPH2D_SOLVER_MUSCL) {
      const bool left = s < o;                          // pair_i = lower local index
      const float epx = left ? ex : -ex, epy = left ? ey : -ey;
      const float2 vi = left ? vs : vo, vj = left ? vo : vs;
      const float2 pci = left ? pcs : pco, pcj = left ? pco : pcs;
      if (SOLVER == SPH2D_SOLVER_HLLC) {
        const float rhoi = left ? rhos : rhoo, rhoj = left ? rhoo : rhos;
        hllc_star<float>(vi.x, vi.y, pci.x, rhoi, pci.y, vj.x, vj.y, pcj.x, rhoj, pcj.y, epx, epy, Q.x_over_t,
                         vsx, vsy, Ps);
      } else {
        const float2* mgo = B.m_mg + 3 * o;
        const float2 gpo = mgo[0], guo = mgo[1], gvo = mgo[2];
        const float2 gpi = left ? gps : gpo, gpj = left ? gpo : gps;
        const float2 gui = left ? gus : guo, guj = left ? guo : gus;
        const float2 gvi = left ? gvs : gvo, gvj = left ? gvo : gvs;
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
        const float UL = Iu * epx + Iv * epy;
        const float UR = Ju * epx + Jv * epy;
        const float dUs = 0.5f * (PL - PR) * inv_rbar * Q.inv_c0;   // U* - U_bar
        vsx = dUs * epx + vbx;
        vsy = dUs * epy + vby;
        const float dUp = UL - UR;
        float bR = 3.0f * (dUp > 0.0f ? dUp : 0.0f);
        bR = bR < cbar ? bR : cbar;
        Ps = Pbar + 0.5f * bR * rbar * dUp;
        vown = left ? make_float2(Iu, Iv) : make_float2(Ju, Jv);
      }
    } else if (SOLVER == SPH2D_SOLVER_LINEAR_RIEMANN) {
      const float zs = rhos * pcs.y, zo = rhoo * pco.y;
      const float inv_z = __builtin_amdgcn_rcpf(zs + zo);
      const float Us = (zs * UR + zo * UL + (pcs.x - pco.x)) * inv_z;
      vsx = Us * ex;
      vsy = Us * ey;
      Ps = (zs * pco.x + zo * pcs.x + zs * zo * dU) * inv_z;
    } else {
      const float Ub = (UL + UR) * 0.5f;
      const float dUs = 0.5f * (pcs.x - pco.x) * inv_rbar * Q.inv_c0;   // U* - U_bar
      vsx = dUs * ex + vbx;
      vsy = dUs * ey + vby;
      float bR = 3.0f * (dU > 0.0f ? dU : 0.0f);
      bR = bR < cbar ? bR : cbar;
      Ps = Pbar + 0.5f * bR * rbar * dU;
      (void)Ub;
    }
    if (nes) { vsx = vbx; vsy = vby; Ps = Pbar; }
    const float ti = (vown.x - vsx) * K.dwx + (vown.y - vsy) * K.dwy;
    drho.add(2.0f * rhos * Vo * ti);
    const float rij = inv_rhos * __builtin_amdgcn_rcpf(rhoo);
    const float2 mo = make_float2(mo_x, r1o.w);
    const float fp = mo.x * (-2.0f * Ps * rij);
    float tx = fp * K.dwx, ty = fp * K.dwy;
    const float vijx = vs.x - vo.x, vijy = vs.y - vo.y;
    const float vx = vijx * dx + vijy * dy;
    if (vx < 0.0f) {
      const float phi = Q.h * vx * __builtin_amdgcn_rcpf(r2 + Q.fai_h2);
      const float Pi = (-(Q.alpha * cbar * phi) + Q.beta * (phi * phi)) * inv_rbar;
      tx += mo.x * (-Pi * K.dwx);
      ty += mo.x * (-Pi * K.dwy);
    }
    const float Rd = dx * K.dwx + dy * K.dwy;
    const float tmp = Q.delta_alpha_h * cbar * vx * inv_r2 * Q.rho0 * inv_rhos * Vo;
    tx += tmp * K.dwx;
    ty += tmp * K.dwy;
    const float4 tauo = r1o;
    float tp = (taus.x + tauo.x) * K.dwx + (taus.y + tauo.y) * K.dwy;
    tp = tp * rij;
    tx += mo.y * tp;
    tp = tp + (taus.z + tauo.z) * K.dwy + (taus.y + tauo.y) * K.dwx;
    tp = tp * rij;
    ty += mo.y * tp;
    const float lf = Vo * Q.two_nu * Rd * inv_r2;
    tx += lf * vijx;
    ty += lf * vijy;
    ax.add(tx);
    ay.add(ty);
    if (SOLVER == SPH2D_SOLVER_MODIFIED_LINEAR_RIEMANN && Q.epsilon2 != 0.0f) {
      const float f = Q.epsilon2 * mo.x / (rhos + rhoo) * K.w;
      agx.add(-f * vijx);
      agy.add(-f * vijy);
    }
    const float2 gro = make_float2(r2o.x, r2o.y);
    const float psi = (rvo.x - rvs.x) + 0.5f * ((grs.x + gro.x) * dx + (grs.y + gro.y) * dy);
    corr.add(-(Q.two_delta_h * cbar * psi * Rd * inv_r2 * Vo));
    if (do_shift) {
      const float wr = K.w * Q.inv_wdx;
      const float f = Q.shift_beta_h2 * (1.0f + Q.shift_R * (wr * wr * wr * wr)) * Vo;
      shx.add(-f * K.dwx);
      shy.add(-f * K.dwy);
    }
  }
  }
  double dr = (double)drho.s;
  if (ts == 2) dr = dr + (double)corr.s;
  B.drhodt[s] = dr;
  B.acc[s] = ts == 2 ? make_double2((double)ax.s, (double)ay.s - P.g) : make_double2(0.0, 0.0);
  if (P.epsilon_xsph != 0.0) B.avg[s] = make_double2((double)agx.s, (double)agy.s);
  if (do_shift) B.shift[s] = make_double2((double)shx.s, (double)shy.s);
}

template <int SOLVER>
__global__ void __launch_bounds__(kBlock) k_eval_c32(DevParams P, DevParams32 Q, Buffers B, int do_shift) {
  k_eval_c32_body<SOLVER>(P, Q, B, do_shift, xcd_block(), gridDim.x);
}

void launch_eval32_a(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  const int g = grid_for(cap_local);
  if (P.solver == SPH2D_SOLVER_MUSCL)
    hipLaunchKernelGGL(k_eval_a32<true>, dim3(g), dim3(kBlock), 0, s, P, make32(P), B);
  else
    hipLaunchKernelGGL(k_eval_a32<false>, dim3(g), dim3(kBlock), 0, s, P, make32(P), B);
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
template <bool MUSCL, int SHAPE>
__global__ void __launch_bounds__(kNbShape[SHAPE].block) kb_eval_a32(BatchArgs A, const DevParams32* Qs) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  const DevParams32 Q = Qs[w.i];
  Buffers B = A.inst[w.i].B;
  k_eval_a32_body<MUSCL, kNbShape[SHAPE].cap_a>(P, Q, B, w.blk, w.nblk);
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
    if (solver == SPH2D_SOLVER_MUSCL) hipLaunchKernelGGL((kb_eval_a32<true, SHAPE>), dim3(g), t, 0, s, A, Qs);
    else hipLaunchKernelGGL((kb_eval_a32<false, SHAPE>), dim3(g), t, 0, s, A, Qs);
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
