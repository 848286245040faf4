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
__device__ __forceinline__ void tait(const DevParams& P, double rho, double& p, double& c) {
  const double x = rho / P.rho0;
  p = P.eos_b * (x * x * x * x * x * x * x - 1.0);
  c = P.sqrt_c0sq * (x * x * x);
}

// FP32 mirrors of the state one neighbour reads in the mixed form.
__device__ __forceinline__ void write_mirror(const DevParams& P, const Buffers& B, int s, double2 vel,
                                             double rho, double vol, double p, double c) {
  B.m_vel[s] = make_float2((float)vel.x, (float)vel.y);
  B.m_rv[s] = make_float2((float)(rho - P.rho0), (float)vol);
  B.m_pc[s] = make_float2((float)p, (float)c);
}

__device__ __forceinline__ double tait_p(const DevParams& P, double rho) {
  const double x = rho / P.rho0;
  return P.eos_b * (x * x * x * x * x * x * x - 1.0);
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
// (1e-7 <= r <= 3h, r = sqrt(dx*dx+dy*dy)) to them and calls f(o) for the
// accepted ones, in ascending candidate order, so the lanes of a wave stay converged on the
// expensive pair body. Lists longer than 256 run in several chunks.
// find_particle_pair.f90:87 with r2 = dx*dx+dy*dy rounded as the reference rounds it
// (no contraction, whatever the translation unit's flags).
__device__ __forceinline__ bool is_pair_r2(const DevParams& P, double dx, double dy) {
  const double r2 = __dadd_rn(__dmul_rn(dx, dx), __dmul_rn(dy, dy));
  return (r2 <= P.r2_max) && (r2 >= P.r2_min);
}

template <class F>
__device__ __forceinline__ void for_each_neighbour(const DevParams& P, const int* __restrict__ cell_start,
                                                   const double2* __restrict__ pos,
                                                   const float2* __restrict__ pos32, int s, double2 ps,
                                                   int cell, F&& f) {
  const float2 ps32 = pos32[s];
  const CellRanges R = cell_ranges(P, cell_start, cell);
  const int n0 = R.hi[0] - R.lo[0], n1 = R.hi[1] - R.lo[1], n2 = R.hi[2] - R.lo[2];
  const int n01 = n0 + n1, total = n01 + n2;
  const int l0 = R.lo[0], l1 = R.lo[1] - n0, l2 = R.lo[2] - n01;
  for (int c0 = 0; c0 < total; c0 += 256) {
    uint64_t m0 = 0, m1 = 0, m2 = 0, m3 = 0;
#pragma unroll 1
    for (int w = 0; w < 4; ++w) {
      const int base = c0 + 64 * w;
      const int cn = min(64, total - base);
      uint64_t mw = 0;
#pragma unroll 2
      for (int k = 0; k < cn; ++k) {
        const int kk = base + k;
        const int o = kk < n0 ? l0 + kk : (kk < n01 ? l1 + kk : l2 + kk);
        const float2 q = pos32[o];
        const float fx = ps32.x - q.x, fy = ps32.y - q.y;
        const bool ok = (fx * fx + fy * fy <= P.cut2_32) && (o != s);
        mw |= (uint64_t)ok << k;
      }
      if (w == 0) m0 = mw; else if (w == 1) m1 = mw; else if (w == 2) m2 = mw; else m3 = mw;
    }
#pragma unroll 1
    while ((m0 | m1 | m2 | m3) != 0) {
      const int w = m0 ? 0 : (m1 ? 1 : (m2 ? 2 : 3));
      const uint64_t mm = m0 ? m0 : (m1 ? m1 : (m2 ? m2 : m3));
      const int b = __builtin_ctzll(mm);
      const uint64_t bit = 1ull << b;
      if (w == 0) m0 ^= bit; else if (w == 1) m1 ^= bit; else if (w == 2) m2 ^= bit; else m3 ^= bit;
      const int kk = c0 + w * 64 + b;
      const int o = kk < n0 ? l0 + kk : (kk < n01 ? l1 + kk : l2 + kk);
      const double2 po = pos[o];
      if (is_pair_r2(P, ps.x - po.x, ps.y - po.y)) f(o);
    }
  }
}

}  // namespace sph2d
