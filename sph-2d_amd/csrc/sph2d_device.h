// sph2d_device.h -- device functions shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include "sph2d_internal.h"

namespace sph2d {

// compute_kernel.f90:123-133, improved Gaussian with support 3h, dim = 2.
// factor = 1/(PI4*h**2*(1-10*exp(-9.0))) and exp(-9.0) are REAL(4) constants
// folded on the host (kfactor, e9). The operation order is the reference's, so
// the only difference to the Fortran build is the last bit of exp().
__device__ __forceinline__ void kernel_w_dw(const DevParams& P, double r, double pdx, double pdy,
                                            double& w, double& dwx, double& dwy) {
  const double q = r / P.h;
  const double eq = exp(-(q * q));
  w = (eq - P.e9) * P.kfactor;
  const double t = eq * (-2.0 * q);
  const double rh = r * P.h;
  dwx = t * pdx / rh * P.kfactor;
  dwy = t * pdy / rh * P.kfactor;
}

// status_equation.f90:61-63: p = B*((rho/rho0)**7-1), c = sqrt(c0^2)*(rho/rho0)**3,
// with flang's left-to-right expansion of the integer powers.
// The moving-square program's law instead (Moving_Square_Or_Cylinder_Inside_a_Rectangular_Box/
// Status_Equation.f90:62-64, EquationOfState = 1): p = c0^2 (rho - rho0) + p_b.
__device__ __forceinline__ void tait(const DevParams& P, double rho, double& p, double& c) {
  const double x = rho / P.rho0;
  if (P.ms) p = P.c0sq * (rho - P.rho0) + P.pb;
  else p = P.eos_b * (x * x * x * x * x * x * x - 1.0);
  c = P.sqrt_c0sq * (x * x * x);
}

// FP32 record r0 of the mixed form: (u, v, rho - rho0, m), m negated if not fluid;
// and (p, c) of the FP64 Tait law (status_equation.f90:61-63) rounded to FP32.
__device__ __forceinline__ void write_r0(const DevParams& P, const Buffers& B, int s, double2 vel, double rho,
                                         double m, int type) {
  const float mf = (float)m;
  B.m_r0[s] = make_float4((float)vel.x, (float)vel.y, (float)(rho - P.rho0), type == 2 ? mf : -mf);
  double p, c;
  tait(P, rho, p, c);
  B.m_pc[s] = make_float2((float)p, (float)c);
}

// Offset of a position from the lower-left corner of its binned cell (the cell of
// Devide_block's binning: x_cell = chain_origin_x + cx * chain_dx).
// FP32 offset of a position to the centre of its cell in x (|offset| <= d/2) and to
// the centre of its pair of cell rows (rows 2k and 2k+1) in y (|offset| <= d): the
// mixed force pass recovers a pair's y offset modulo two rows (see c32_pair).
__device__ __forceinline__ float2 cell_offset(const DevParams& P, double2 x, int cell) {
  const int cx = cell / P.chain_ny, cy = cell - cx * P.chain_ny;
  return make_float2((float)(x.x - (P.chain_ox + (cx + 0.5) * P.chain_d)),
                     (float)(x.y - (P.chain_oy + ((cy & ~1) + 1) * P.chain_d)));
}

__device__ __forceinline__ double tait_p(const DevParams& P, double rho) {
  const double x = rho / P.rho0;
  return P.eos_b * (x * x * x * x * x * x * x - 1.0);
}

// Buffer resource over a float4 array (raw addressing, byte offsets < 2^31;
// word 3 = 0x00020000, the CDNA data-format setting for raw dword loads).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rec_rsrc(const float4* p) {
  return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, 0x7fffffff, 0x00020000);
}

__device__ __forceinline__ float4 buf_f4(__amdgpu_buffer_rsrc_t r, int idx) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 w = __builtin_amdgcn_raw_buffer_load_b128(r, idx * 16, 0, 0);
  return make_float4(__uint_as_float(w.x), __uint_as_float(w.y), __uint_as_float(w.z), __uint_as_float(w.w));
}

__device__ __forceinline__ float2 buf_f2(__amdgpu_buffer_rsrc_t r, int idx) {
  typedef unsigned int u2 __attribute__((ext_vector_type(2)));
  const u2 w = __builtin_amdgcn_raw_buffer_load_b64(r, idx * 8, 0, 0);
  return make_float2(__uint_as_float(w.x), __uint_as_float(w.y));
}

__device__ __forceinline__ double2 buf_d2(__amdgpu_buffer_rsrc_t r, int idx) {
  typedef unsigned int u4 __attribute__((ext_vector_type(4)));
  const u4 w = __builtin_amdgcn_raw_buffer_load_b128(r, idx * 16, 0, 0);
  return make_double2(__hiloint2double((int)w.y, (int)w.x), __hiloint2double((int)w.w, (int)w.z));
}

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup b to XCD b % 8),
// each XCD with its own L2. The logical block of workgroup b gives every XCD one
// contiguous run of blocks, so blocks whose neighbour columns overlap share an L2.
// The instance of batched block b and its block index within the instance.
struct BatchWhere {
  int i, blk, nblk;
};
__device__ __forceinline__ BatchWhere batch_where(const BatchArgs& A, int b) {
  int lo = 0, hi = A.n - 1;
  while (lo < hi) {                                  // last instance whose first block <= b
    const int mid = (lo + hi + 1) >> 1;
    if (A.start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return BatchWhere{lo, b - A.start[lo], A.start[lo + 1] - A.start[lo]};
}

__device__ __forceinline__ int xcd_block() {
  const int nb = gridDim.x, b = blockIdx.x;
  const int q = nb >> 3, r = nb & 7;
  const int x = b & 7, k = b >> 3;
  return (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + k;
}

// Row 1 of the inverse of the order-6 MLS moment matrix, y = (A + al3 I)^-1 e_1, with
// A symmetric positive semi-definite and given by its upper triangle, row-major over
// (m <= l): a00 a01 .. a05 a11 .. a55. The reference inverts with the full-pivot
// Gauss-Jordan BRINV (function_module.f90:568-639) and keeps row 1; a Cholesky
// solve in registers gives the same row with another rounding (relative ~ cond * eps).
// Returns false when a pivot is not positive.
__device__ __forceinline__ bool mls6_first_row(const double* A, double al3, double* y) {
  double L[21];
  bool ok = true;
  const int diag[6] = {0, 6, 11, 15, 18, 20};
#define AU(m, l) A[diag[(m)] + (l) - (m)]          // upper-triangle access, m <= l
#define LL(i, j) L[(i) * ((i) + 1) / 2 + (j)]      // lower-triangle access, j <= i
#pragma unroll
  for (int i = 0; i < 6; ++i) {
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      double v = AU(j, i) + (i == j ? al3 : 0.0);
#pragma unroll
      for (int k = 0; k < j; ++k) v = v - LL(i, k) * LL(j, k);
      if (i == j) {
        ok = ok && (v > 0.0);
        LL(i, i) = sqrt(v > 0.0 ? v : 1.0);
      } else {
        LL(i, j) = v / LL(j, j);
      }
    }
  }
  double z[6];
#pragma unroll
  for (int i = 0; i < 6; ++i) {
    double v = (i == 0) ? 1.0 : 0.0;
#pragma unroll
    for (int k = 0; k < i; ++k) v = v - LL(i, k) * z[k];
    z[i] = v / LL(i, i);
  }
#pragma unroll
  for (int i = 5; i >= 0; --i) {
    double v = z[i];
#pragma unroll
    for (int k = i + 1; k < 6; ++k) v = v - LL(k, i) * y[k];
    y[i] = v / LL(i, i);
  }
#undef AU
#undef LL
  return ok;
}

// Neighbour cells of a particle: 3 column ranges of the column-major cell table.
struct CellRanges {
  int lo[3], hi[3];
};

__device__ __forceinline__ CellRanges cell_ranges(const DevParams& P, const int* __restrict__ cell_start,
                                                  int cell) {
  CellRanges R;
  const int ny = P.chain_ny;
  const int cx = cell / ny, cy = cell - cx * ny;
  const int r0 = cy > 0 ? cy - 1 : 0;
  const int r1 = cy + 1 < ny ? cy + 1 : ny - 1;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const int col = cx + d - 1;
    if (col < 0 || col >= P.chain_nx) {
      R.lo[d] = 0; R.hi[d] = 0;
    } else {
      R.lo[d] = cell_start[col * ny + r0];
      R.hi[d] = cell_start[col * ny + r1 + 1];
    }
  }
  return R;
}

}  // namespace sph2d

namespace sph2d {

// Neighbour loop of the gather kernels. The 3 column ranges of the 3x3 cell
// stencil form one candidate list (about 207 candidates, 36 of them inside 3h).
// Phase 1 screens every candidate with FP32 positions against a bound that no
// pair within 3h can exceed and keeps a 256-bit mask of the survivors (about 40
// of 207); phase 2 applies the exact FP64 criterion of find_particle_pair.f90:87
// (1e-7 <= r <= 3h, r = sqrt(dx*dx+dy*dy)) to them and calls f(o, d, li) (d:
// the candidate's stencil column, 0..2 for cell columns cx-1..cx+1; li: its slot
// in a staged block's LDS copy, -1 when not staged) for the
// accepted ones, in ascending candidate order, so the lanes of a wave stay
// converged on the pair body. Lists longer than 256 run in chunks.
//
// Modes: kNbScan screens and tests; kNbRecord also records the exactly accepted
// candidates of a list of at most 256 (their slots, one byte each, 4 per word,
// word w of particle s at list[w * lstride + s], and their number in counts[s]);
// kNbReplay reads that record instead of screening and testing, so later passes of
// the same evaluation (same positions, same cell table) visit the same neighbours
// in the same order without reading positions for the test. A list longer than
// 256 candidates, or one with more than kNbListCap neighbours, is screened and
// tested again.
// find_particle_pair.f90:87 with r2 = dx*dx+dy*dy rounded as the reference rounds it
// (no contraction, whatever the translation unit's flags).
__device__ __forceinline__ bool is_pair_r2(const DevParams& P, double dx, double dy) {
  const double r2 = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
  return (r2 <= P.r2_max) && (r2 >= P.r2_min);
}

enum { kNbScan = 0, kNbRecord = 1, kNbReplay = 2 };

// Per-block staging of candidate positions in LDS (pass A). The particles of a
// block are consecutive in local order; when they all sit in one cell column,
// the union of their candidate lists is three contiguous local ranges (the
// stencil columns over rows first-1 .. last+1), which the block copies to LDS
// once -- FP32 and FP64 positions -- and every thread's own ranges are
// sub-ranges of them. A block that straddles two columns, or whose union
// exceeds the stage capacity (kStageCap slots unless the caller sizes its own
// stage), reads global memory as before.
constexpr int kStageCap = 1024;

struct NbStage {
  int use;
  int lo[3];      // first staged local index of each stencil column
  int base[3];    // its LDS slot
  int total;
};

// The block's candidate union (st.use = 0 when it cannot be staged within cap);
// ends with a barrier.
__device__ __forceinline__ void plan_stage(const DevParams& P, const int* __restrict__ cell_start,
                                           const int* __restrict__ l_cell, int n, int cap, NbStage& st, int blk) {
  if (threadIdx.x == 0) {
    int use = 0, t = 0;
    int lo[3] = {0, 0, 0}, cnt[3] = {0, 0, 0};
    const int b0 = blk * blockDim.x;
    if (b0 < n) {
      const int b1 = min(b0 + (int)blockDim.x, n) - 1;
      const int ny = P.chain_ny;
      const int c0 = l_cell[b0], c1 = l_cell[b1];
      const int cx = c0 / ny;
      if (c1 / ny == cx) {
        const int rlo = max(c0 - cx * ny - 1, 0), rhi = min(c1 - cx * ny + 1, ny - 1);
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          const int col = cx + d - 1;
          if (col >= 0 && col < P.chain_nx) {
            lo[d] = cell_start[col * ny + rlo];
            cnt[d] = cell_start[col * ny + rhi + 1] - lo[d];
          }
          t += cnt[d];
        }
        use = t <= cap;
      }
    }
    st.use = use;
    st.total = use ? t : 0;
#pragma unroll
    for (int d = 0; d < 3; ++d) st.lo[d] = lo[d];
    st.base[0] = 0;
    st.base[1] = cnt[0];
    st.base[2] = cnt[0] + cnt[1];
  }
  __syncthreads();
}

// Local index of LDS slot i of a planned stage.
__device__ __forceinline__ int stage_source(const NbStage& st, int i) {
  const int d = i < st.base[1] ? 0 : (i < st.base[2] ? 1 : 2);
  return st.lo[d] + (i - st.base[d]);
}

__device__ __forceinline__ void stage_candidates(const DevParams& P, const int* __restrict__ cell_start,
                                                 const int* __restrict__ l_cell, const float2* __restrict__ pos32,
                                                 const double2* __restrict__ pos, int n, NbStage& st,
                                                 float2* lds32, double2* lds64, int blk, int cap = kStageCap) {
  plan_stage(P, cell_start, l_cell, n, cap, st, blk);
  const int total = st.total;
  for (int i = threadIdx.x; i < total; i += blockDim.x) {
    const int o = stage_source(st, i);
    lds32[i] = pos32[o];
    lds64[i] = pos[o];
  }
  __syncthreads();
}

// Exact acceptance mask of candidates c0 .. c0+255 of a candidate list (screen,
// then the exact test on the survivors); for lists longer than the recorded 256.
struct Mask256 {
  uint64_t w0, w1, w2, w3;
};

__device__ __forceinline__ Mask256 exact_chunk_mask(const DevParams& P, const double2* __restrict__ pos,
                                                    const float2* __restrict__ pos32, int s, double2 ps,
                                                    int n0, int n01, int l0, int l1, int l2, int total, int c0) {
  const float2 ps32 = pos32[s];
  uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
#pragma unroll 1
  for (int w = 0; w < 4; ++w) {
    const int base = c0 + 64 * w;
    const int cn = min(64, total - base);
    uint64_t mw = 0;
    for (int k = 0; k < cn; ++k) {
      const int kk = base + k;
      const int o = kk < n0 ? l0 + kk : (kk < n01 ? l1 + kk : l2 + kk);
      const float2 q = pos32[o];
      const float fx = ps32.x - q.x, fy = ps32.y - q.y;
      bool ok = (fx * fx + fy * fy <= P.cut2_32) && (o != s);
      if (ok) {
        const double2 po = pos[o];
        ok = is_pair_r2(P, ps.x - po.x, ps.y - po.y);
      }
      mw |= (uint64_t)ok << k;
    }
    if (w == 0) m0 = mw; else if (w == 1) m1 = mw; else if (w == 2) m2 = mw; else m3 = mw;
  }
  return Mask256{m0, m1, m2, m3};
}

// (w3:w2:w1:w0) |= (hi:lo) << sh for sh in [0, 128]; bits beyond 256 are dropped.
__device__ __forceinline__ void or_shl128(uint64_t& w0, uint64_t& w1, uint64_t& w2, uint64_t& w3, uint64_t lo,
                                          uint64_t hi, int sh) {
  const int q = sh >> 6, r = sh & 63;
  const uint64_t a = lo << r;
  const uint64_t b = r ? ((hi << r) | (lo >> (64 - r))) : hi;
  const uint64_t c = r ? (hi >> (64 - r)) : 0ull;
  w0 |= q == 0 ? a : 0ull;
  w1 |= q == 1 ? a : (q == 0 ? b : 0ull);
  w2 |= q == 2 ? a : (q == 1 ? b : (q == 0 ? c : 0ull));
  w3 |= q == 2 ? b : (q == 1 ? c : 0ull);
}

// The flags of one wrapped fluid position: x-ghost to the right of the box (l),
// to the left (r), y-ghosts above (yl) and below (yr)
// (Distribute_Period_Boundary_Particle.f90:98-148, :153-206).
__device__ __forceinline__ void ghost_flags(const DevParams& P, double2 p, int& l, int& r, int& yl, int& yr) {
  const double r3h = P.three_h;
  l = (p.x <= (0.0 + r3h) && p.x >= 0.0);
  r = (p.x >= (P.box_x - r3h) && p.x <= P.box_x);
  yl = (p.y <= (0.0 + r3h) && p.y >= 0.0);
  yr = (p.y >= (P.box_y - r3h) && p.y <= P.box_y);
}

// The FP32 screen of candidates c0 .. c0+255 of a particle's list: bit k of the
// result is set when candidate c0 + k may lie within 3h (no pair within 3h fails
// it); the particle itself is never set.
__device__ __forceinline__ Mask256 screen_chunk(const DevParams& P, const float2* __restrict__ pos32, int s,
                                                float2 ps32, const CellRanges& R, int n0, int n1, int n2, int n01,
                                                int total, int l0, int l1, int l2, bool staged, int s0, int s1,
                                                int s2, const float2* lds32, int c0) {
uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
  if (staged && total <= 256 && n0 <= 128 && n1 <= 128 && n2 <= 128) {
    // Column by column from LDS: a column's slots are contiguous and the loop
    // index is the same in every lane, so a candidate costs one LDS read, the
    // distance test and one bit insert; the self pair is cleared afterwards.
#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
      const int nd = d == 0 ? n0 : (d == 1 ? n1 : n2);
      const float2* q = lds32 + (d == 0 ? s0 : (d == 1 ? s1 + n0 : s2 + n01));
      uint64_t lo = 0, hi = 0;
#pragma unroll 1
      for (int j = 0; j < 4; ++j) {
        uint32_t acc = 0;
        const int kend = min(nd - 32 * j, 32);
#pragma unroll 4
        for (int k = 0; k < kend; ++k) {
          const float2 v = q[32 * j + k];
          const float fx = ps32.x - v.x, fy = ps32.y - v.y;
          acc |= (uint32_t)(fx * fx + fy * fy <= P.cut2_32) << k;
        }
        if (j == 0) lo |= acc;
        else if (j == 1) lo |= (uint64_t)acc << 32;
        else if (j == 2) hi |= acc;
        else hi |= (uint64_t)acc << 32;
      }
      if (d == 1) {
        const int ks = s - R.lo[1];                 // the particle itself
        if (ks < 64) lo &= ~(1ull << ks); else hi &= ~(1ull << (ks - 64));
      }
      or_shl128(m0, m1, m2, m3, lo, hi, d == 0 ? 0 : (d == 1 ? n0 : n01));
    }
  } else {
#pragma unroll 1
    for (int w = 0; w < 4; ++w) {
      const int base = c0 + 64 * w;
      const int cn = min(64, total - base);
      uint64_t mw = 0;
      if (staged) {
#pragma unroll 2
        for (int k = 0; k < cn; ++k) {
          const int kk = base + k;
          const int o = kk < n0 ? l0 + kk : (kk < n01 ? l1 + kk : l2 + kk);
          const int li = kk < n0 ? s0 + kk : (kk < n01 ? s1 + kk : s2 + kk);
          const float2 q = lds32[li];
          const float fx = ps32.x - q.x, fy = ps32.y - q.y;
          const bool ok = (fx * fx + fy * fy <= P.cut2_32) && (o != s);
          mw |= (uint64_t)ok << k;
        }
      } else {
#pragma unroll 2
        for (int k = 0; k < cn; ++k) {
          const int kk = base + k;
          const int o = kk < n0 ? l0 + kk : (kk < n01 ? l1 + kk : l2 + kk);
          const float2 q = pos32[o];
          const float fx = ps32.x - q.x, fy = ps32.y - q.y;
          const bool ok = (fx * fx + fy * fy <= P.cut2_32) && (o != s);
          mw |= (uint64_t)ok << k;
        }
      }
      if (w == 0) m0 = mw; else if (w == 1) m1 = mw; else if (w == 2) m2 = mw; else m3 = mw;
    }
  }
  return Mask256{m0, m1, m2, m3};
}

template <int MODE = kNbScan, class F>
__device__ __forceinline__ void for_each_neighbour(const DevParams& P, const int* __restrict__ cell_start,
                                                   const double2* __restrict__ pos,
                                                   const float2* __restrict__ pos32, int s, double2 ps,
                                                   int cell, F&& f, uint32_t* __restrict__ list = nullptr,
                                                   int* __restrict__ counts = nullptr, int lstride = 0,
                                                   const NbStage* st = nullptr,
                                                   const float2* lds32 = nullptr, const double2* lds64 = nullptr) {
  const CellRanges R = cell_ranges(P, cell_start, cell);
  const int n0 = R.hi[0] - R.lo[0], n1 = R.hi[1] - R.lo[1], n2 = R.hi[2] - R.lo[2];
  const int n01 = n0 + n1, total = n01 + n2;
  const int l0 = R.lo[0], l1 = R.lo[1] - n0, l2 = R.lo[2] - n01;
  const bool stored = total <= 256;
  const bool staged = st != nullptr && st->use;     // uniform over the block
  // LDS slot of candidate kk: kk + s0/s1/s2 by stencil column
  const int s0 = staged ? l0 - st->lo[0] + st->base[0] : 0;
  const int s1 = staged ? l1 - st->lo[1] + st->base[1] : 0;
  const int s2 = staged ? l2 - st->lo[2] + st->base[2] : 0;
  if (MODE == kNbReplay && stored) {
    const int cnt = counts[s];
    if (!(cnt & kNbOverflow)) {
      // the recorded slots in ascending order: one byte per neighbour; the loop
      // index advances in every active lane together, so the word loads are uniform
      uint32_t word = 0;
#pragma unroll 1
      for (int t = 0; t < cnt; ++t) {
        if ((t & 3) == 0) word = list[(t >> 2) * lstride + s];
        const int kk = (int)((word >> ((t & 3) << 3)) & 255u);
        const int d = kk < n0 ? 0 : (kk < n01 ? 1 : 2);
        const int o = d == 0 ? l0 + kk : (d == 1 ? l1 + kk : l2 + kk);
        const int li = staged ? (d == 0 ? s0 + kk : (d == 1 ? s1 + kk : s2 + kk)) : -1;
        f(o, d, li);
      }
      return;
    }
  }
  const float2 ps32 = pos32[s];
  const __amdgpu_buffer_rsrc_t rpos = rec_rsrc(reinterpret_cast<const float4*>(pos));
  int nrec = 0;                                      // kNbRecord: accepted so far
  uint32_t word = 0;
  for (int c0 = 0; c0 < total; c0 += 256) {
    const Mask256 sm = screen_chunk(P, pos32, s, ps32, R, n0, n1, n2, n01, total, l0, l1, l2, staged, s0, s1, s2,
                                    lds32, c0);
    uint64_t m0 = sm.w0, m1 = sm.w1, m2 = sm.w2, m3 = sm.w3;
#pragma unroll 1
    while ((m0 | m1 | m2 | m3) != 0) {
      const int w = m0 ? 0 : (m1 ? 1 : (m2 ? 2 : 3));
      const uint64_t mm = m0 ? m0 : (m1 ? m1 : (m2 ? m2 : m3));
      const int b = __builtin_ctzll(mm);
      const uint64_t bit = 1ull << b;
      if (w == 0) m0 ^= bit; else if (w == 1) m1 ^= bit; else if (w == 2) m2 ^= bit; else m3 ^= bit;
      const int kk = c0 + w * 64 + b;
      const int d = kk < n0 ? 0 : (kk < n01 ? 1 : 2);     // stencil column of the candidate
      const int o = d == 0 ? l0 + kk : (d == 1 ? l1 + kk : l2 + kk);
      const int li = staged ? (d == 0 ? s0 + kk : (d == 1 ? s1 + kk : s2 + kk)) : -1;   // LDS slot
      double2 po;
      if (staged) po = lds64[li];
      else po = buf_d2(rpos, o);                   // a buffer load: never merged with the LDS one
      const bool accept = is_pair_r2(P, ps.x - po.x, ps.y - po.y);
      if (MODE == kNbRecord && accept && stored) {
        if (nrec < kNbListCap) {
          word |= (uint32_t)kk << ((nrec & 3) << 3);
          if ((nrec & 3) == 3) {
            list[(nrec >> 2) * lstride + s] = word;
            word = 0;
          }
        }
        ++nrec;
      }
      if (accept) f(o, d, li);                    // one call site: the body inlines once
    }
  }
  if (MODE == kNbRecord) {
    if (stored && (nrec & 3) && nrec < kNbListCap) list[(nrec >> 2) * lstride + s] = word;
    // a list longer than 256 candidates, or with more neighbours than recorded, is
    // screened again by every later pass
    counts[s] = nrec | ((!stored || nrec > kNbListCap) ? kNbOverflow : 0);
  }
}

// The support of a point for the global-domain MLS (Calculation_For_PostProceeding.f90
// :184-250, Sampling_subroutine.f90:60-95): the particles with gid < n_max (fluid,
// no periodic ghosts) of the point's 3x3 cells in the step-start binning whose
// refreshed position lies within [1e-7 (REAL(4)), 3h]; f(gid, position, W).
template <class F>
__device__ __forceinline__ void for_each_support(const DevParams& P, const Buffers& B, double ex, double ey,
                                                 int n_max, F&& f) {
  const int xn = (int)floor((ex - P.chain_ox) / P.chain_d + 1);
  const int yn = (int)floor((ey - P.chain_oy) / P.chain_d + 1);
  if (xn < 1 || xn > P.chain_nx || yn < 1 || yn > P.chain_ny) return;
  const CellRanges R = cell_ranges(P, B.cell_start, (xn - 1) * P.chain_ny + (yn - 1));
  const double rmin = (double)1.0e-7f, rmax = P.three_h;
  for (int d = 0; d < 3; ++d)
    for (int o = R.lo[d]; o < R.hi[d]; ++o) {
      const int g = B.name[o];
      if (g >= n_max) continue;
      const double2 pj = B.g_pos[g];
      const double pdx = ex - pj.x, pdy = ey - pj.y;
      const double r = sqrt(pdx * pdx + pdy * pdy);
      if (!(r <= rmax && r >= rmin)) continue;
      double w, dwx, dwy;
      kernel_w_dw(P, r, pdx, pdy, w, dwx, dwy);
      f(g, pj, w);
    }
}

}  // namespace sph2d
