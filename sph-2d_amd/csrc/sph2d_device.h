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
