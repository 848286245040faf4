// sph2d_fp64.hip -- plain FP64 form of the WCSPH step (gfx950).
//
// Gather form: every particle sums over its own neighbours, so no pair list is
// stored and no float atomics are used; each sum runs in a fixed order
// (neighbour columns cx-1, cx, cx+1, rows ascending, particles ascending) and
// the results are deterministic. Each pair term is evaluated in the reference's
// pair orientation (a = lower local index, b = higher), so a term equals the
// Fortran term up to the last bit of exp(); only the summation order differs.
// Built with -ffp-contract=off (no FMA contraction), like the reference build.
#include "sph2d_internal.h"
#include "sph2d_device.h"
#include <algorithm>
#include <cmath>
#include <cstdio>

namespace sph2d {

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

#define HIPCHK(x) do { hipError_t e_ = (x); if (e_ != hipSuccess) { \
  fprintf(stderr, "HIP error %s at %s:%d\n", hipGetErrorString(e_), __FILE__, __LINE__); } } while (0)

// ---------------------------------------------------------------------------
// utilities: fill, device-wide exclusive scan of int32 (3 launches, fixed grid)
// ---------------------------------------------------------------------------
__global__ void k_fill_i32(int* p, int n, int v) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void launch_fill_i32(hipStream_t s, int* p, int n, int v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n, v);
}

constexpr int kScanItems = 4;                    // items per thread
constexpr int kScanTile = kBlock * kScanItems;   // 1024 items per block

__device__ __forceinline__ int block_exclusive_scan(int v, int* lds, int& total) {
  // Hillis-Steele over 256 threads in LDS.
  const int t = threadIdx.x;
  lds[t] = v;
  __syncthreads();
  for (int off = 1; off < kBlock; off <<= 1) {
    int add = t >= off ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  total = lds[kBlock - 1];
  int incl = lds[t];
  __syncthreads();
  return incl - v;
}

__device__ __forceinline__ void k_scan_reduce_body(const int* __restrict__ in, int n, int* __restrict__ blocks, int blk_, int nblk_) {
  __shared__ int lds[kBlock];
  const int base = blk_ * kScanTile + threadIdx.x * kScanItems;
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) s += (base + k < n) ? in[base + k] : 0;
  int total;
  block_exclusive_scan(s, lds, total);
  if (threadIdx.x == 0) blocks[blk_] = total;
}

__global__ void k_scan_reduce(const int* __restrict__ in, int n, int* __restrict__ blocks) {
  k_scan_reduce_body(in, n, blocks, blockIdx.x, gridDim.x);
}

// One block: each thread scans a contiguous run of the block totals serially,
// then one block-wide scan of the run totals gives the runs' offsets.
__device__ __forceinline__ void k_scan_top_body(int* blocks, int nb, int* total_out, int* total_out2, int blk_, int nblk_) {
  __shared__ int lds[kBlock];
  const int per = (nb + kBlock - 1) / kBlock;
  const int b0 = threadIdx.x * per;
  const int b1 = min(b0 + per, nb);
  int run = 0;
  for (int i = b0; i < b1; ++i) run += blocks[i];
  int total;
  int ex = block_exclusive_scan(run, lds, total);
  for (int i = b0; i < b1; ++i) {
    const int v = blocks[i];
    blocks[i] = ex;
    ex += v;
  }
  if (threadIdx.x == 0) {
    if (total_out) *total_out = total;
    if (total_out2) *total_out2 = total;
  }
}

__global__ void k_scan_top(int* blocks, int nb, int* total_out, int* total_out2) {
  k_scan_top_body(blocks, nb, total_out, total_out2, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void k_scan_apply_body(const int* __restrict__ in, int n, const int* __restrict__ blocks, int* __restrict__ out, int blk_, int nblk_) {
  __shared__ int lds[kBlock];
  const int base = blk_ * kScanTile + threadIdx.x * kScanItems;
  int v[kScanItems];
  int s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) { v[k] = (base + k < n) ? in[base + k] : 0; s += v[k]; }
  int total;
  int ex = block_exclusive_scan(s, lds, total) + blocks[blk_];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = ex;
    ex += v[k];
  }
}

__global__ void k_scan_apply(const int* __restrict__ in, int n, const int* __restrict__ blocks,
                             int* __restrict__ out) {
  k_scan_apply_body(in, n, blocks, out, blockIdx.x, gridDim.x);
}

// out[i] = sum_{k<i} in[k] for i < n; the grand total goes to *total (and *total2).
static void exclusive_scan(hipStream_t s, const int* in, int n, int* out, int* blocks,
                           int* total, int* total2 = nullptr) {
  if (n <= 0) n = 1;
  int nb = (n + kScanTile - 1) / kScanTile;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kBlock), 0, s, in, n, blocks);
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kBlock), 0, s, blocks, nb, total, total2);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kBlock), 0, s, in, n, blocks, out);
}

void launch_exclusive_scan(hipStream_t s, const int* in, int n, int* out, int* blocks, int* total) {
  exclusive_scan(s, in, n, out, blocks, total);
}

int scan_block_count(int n) { return (n <= 0 ? 1 : (n + kScanTile - 1) / kScanTile); }

// ---------------------------------------------------------------------------
// Distribute_Initial_Particle.f90:204-249 (TGV lattice)
// ---------------------------------------------------------------------------
__global__ void k_init_tgv(DevParams P, Buffers B, int nx, int ny, double pi, double expo,
                           double vol0) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nx * ny) return;
  int i = k / nx + 1, j = k - (k / nx) * nx + 1;   // row i (y), column j (x), both 1-based
  double x = (j - 1) * P.dx + P.dx / 2.0;
  double y = (i - 1) * (P.box_y / ny) + (P.box_y / ny) / 2.0;
  B.g_pos[k] = make_double2(x, y);
  double u = P.U * sin(2 * pi * x) * cos(2 * pi * y);
  double v = -P.U * cos(2 * pi * x) * sin(2 * pi * y);
  B.g_vel[k] = make_double2(u, v);
  double pr = 0.25 * P.rho0 * (P.U * P.U) * (cos(4 * pi * x) + cos(4 * pi * y));
  double rho = P.rho0 * pow(1 + 7 * pr / (P.c0sq * P.rho0), expo);
  B.g_p[k] = pr;
  B.g_rho[k] = rho;
  B.g_mass[k] = vol0 * rho;
  B.g_type[k] = 2;
  B.g_c[k] = P.sqrt_c0sq;
  B.g_e[k] = P.c0sq / (7.0 - 1.0);
}

void launch_init_tgv(hipStream_t s, const DevParams& P, Buffers& B, int nx, int ny) {
  const double pi = 3.14159265358979324;            // information_module.f90:29
  const double expo = (double)0.142857142f;         // REAL(4) literal, Distribute_Initial_Particle.f90:222
  const double vol0 = P.dx * (P.box_y / ny);
  hipLaunchKernelGGL(k_init_tgv, dim3(grid_for(nx * ny)), dim3(kBlock), 0, s, P, B, nx, ny, pi, expo, vol0);
}

// ---------------------------------------------------------------------------
// Distribute_Period_Boundary_Particle.f90:56-206
// ---------------------------------------------------------------------------
// Exclusive scan in place of two arrays of nb ints that follow each other (block 0
// scans a[0..nb), block 1 a[nb..2nb)); totals to *t0 and *t1. Each of the 1024
// threads first sums a contiguous run (independent loads), then one block scan of
// the run totals gives the runs' offsets.
__device__ __forceinline__ void k_scan_top2_body(int* a, int nb, int* t0, int* t1, int blk_, int nblk_) {
  __shared__ int lds[1024];
  int* arr = a + (size_t)blk_ * nb;
  const int t = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = min(t * per, nb), b1 = min(b0 + per, nb);
  int run = 0;
  for (int c = b0; c < b1; c += 16) {            // 16 loads in flight at a time
    int v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = c + k < b1 ? arr[c + k] : 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) run += v[k];
  }
  lds[t] = run;
  __syncthreads();
  for (int off = 1; off < 1024; off <<= 1) {
    const int add = t >= off ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += add;
    __syncthreads();
  }
  int ex = lds[t] - run;
  for (int c = b0; c < b1; c += 16) {
    int v[16];
#pragma unroll
    for (int k = 0; k < 16; ++k) v[k] = c + k < b1 ? arr[c + k] : 0;
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      if (c + k < b1) arr[c + k] = ex;
      ex += v[k];
    }
  }
  if (t == 1023) *(blk_ ? t1 : t0) = lds[1023];
}

__global__ void __launch_bounds__(1024) k_scan_top2(int* a, int nb, int* t0, int* t1) {
  k_scan_top2_body(a, nb, t0, t1, blockIdx.x, gridDim.x);
}

// :56-90 wrap (written back only where it moves a particle), then per block of
// kBlock gids the number of x-ghosts (gblk[b]) and of y-ghosts of fluid
// particles (gblk[nb + b]).
__device__ __forceinline__ void k_wrap_count_body(const DevParams& P, Buffers& B, int cap_fluid, int blk_, int nblk_) {
  __shared__ int sx[kBlock / 64], sy[kBlock / 64];
  const int i = blk_ * blockDim.x + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  int l = 0, r = 0, yl = 0, yr = 0;
  if (i < N && i < cap_fluid && B.g_type[i] != 120) {
    const double2 p0 = B.g_pos[i];
    double2 p = p0;
    if (p.x > P.box_x) p.x = p.x - P.box_x;
    else if (p.x < 0.0) p.x = p.x + P.box_x;
    if (p.y > P.box_y) p.y = p.y - P.box_y;
    else if (p.y < 0.0) p.y = p.y + P.box_y;
    if (p.x != p0.x || p.y != p0.y) B.g_pos[i] = p;
    ghost_flags(P, p, l, r, yl, yr);
  }
  const int w = threadIdx.x >> 6;
  const int cx = __popcll(__ballot(l)) + __popcll(__ballot(r));
  const int cy = __popcll(__ballot(yl)) + __popcll(__ballot(yr));
  if ((threadIdx.x & 63) == 0) { sx[w] = cx; sy[w] = cy; }
  __syncthreads();
  if (threadIdx.x == 0) {
    int tx = 0, ty = 0;
#pragma unroll
    for (int k = 0; k < kBlock / 64; ++k) { tx += sx[k]; ty += sy[k]; }
    B.gblk[blk_] = tx;
    B.gblk[nblk_ + blk_] = ty;
  }
}

__global__ void __launch_bounds__(kBlock) k_wrap_count(DevParams P, Buffers B, int cap_fluid) {
  k_wrap_count_body(P, B, cap_fluid, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void copy_particle(Buffers& B, int dst, int src, double2 pos) {
  B.g_pos[dst] = pos;
  B.g_vel[dst] = B.g_vel[src];
  B.g_rho[dst] = B.g_rho[src];
  B.g_mass[dst] = B.g_mass[src];
  B.g_type[dst] = -101;
  B.g_p[dst] = B.g_p[src];
  B.g_c[dst] = B.g_c[src];
  B.g_e[dst] = B.g_e[src];
}

// The x-ghosts (:98-148, left copy before right copy for each gid) and the
// y-ghosts of fluid particles (:153-206, first part of the loop), each in gid
// order: a block's slots start at its scanned count (gblk) and each lane's
// after the ghosts of the lanes before it. Blocks without ghosts leave at once.
// Records each x-ghost's own y-ghost count for the corner pass.
__device__ __forceinline__ void k_ghost_emit_body(const DevParams& P, Buffers& B, int cap_fluid, int cap_ghost, int blk_, int nblk_) {
  __shared__ int sx[kBlock / 64], sy[kBlock / 64];
  const int b = blk_, nb = nblk_;
  const int nxg = B.counters[C_NXG], nyf = B.counters[C_NYF];
  const int ox = B.gblk[b], oy = B.gblk[nb + b];
  const int ex = (b + 1 < nb ? B.gblk[b + 1] : nxg) - ox;
  const int ey = (b + 1 < nb ? B.gblk[nb + b + 1] : nyf) - oy;
  if (ex == 0 && ey == 0) return;                  // uniform over the block
  const int i = b * kBlock + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  int l = 0, r = 0, yl = 0, yr = 0;
  double2 p = make_double2(0.0, 0.0);
  if (i < N && i < cap_fluid && B.g_type[i] != 120) {
    p = B.g_pos[i];
    ghost_flags(P, p, l, r, yl, yr);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const uint64_t lt = (1ull << lane) - 1ull;
  const uint64_t bl = __ballot(l), br = __ballot(r), byl = __ballot(yl), byr = __ballot(yr);
  int px = __popcll(bl & lt) + __popcll(br & lt);
  int py = __popcll(byl & lt) + __popcll(byr & lt);
  if (lane == 0) { sx[w] = __popcll(bl) + __popcll(br); sy[w] = __popcll(byl) + __popcll(byr); }
  __syncthreads();
  for (int k = 0; k < w; ++k) { px += sx[k]; py += sy[k]; }
  int k = ox + px;
  if (l) {
    if (k < cap_ghost) {
      copy_particle(B, N + k, i, make_double2(p.x + P.box_x, p.y));
      B.cnt_yx[k] = yl + yr;
      B.src_x[k] = i;
    } else {
      atomicOr(&B.counters[C_ERR], E_GHOST_CAP);
    }
    ++k;
  }
  if (r) {
    if (k < cap_ghost) {
      copy_particle(B, N + k, i, make_double2(p.x - P.box_x, p.y));
      B.cnt_yx[k] = yl + yr;
      B.src_x[k] = i;
    } else {
      atomicOr(&B.counters[C_ERR], E_GHOST_CAP);
    }
  }
  k = nxg + oy + py;
  if (yl) {
    if (k < cap_ghost) copy_particle(B, N + k, i, make_double2(p.x, p.y + P.box_y));
    else atomicOr(&B.counters[C_ERR], E_GHOST_CAP);
    ++k;
  }
  if (yr) {
    if (k < cap_ghost) copy_particle(B, N + k, i, make_double2(p.x, p.y - P.box_y));
    else atomicOr(&B.counters[C_ERR], E_GHOST_CAP);
  }
}

__global__ void __launch_bounds__(kBlock) k_ghost_emit(DevParams P, Buffers B, int cap_fluid, int cap_ghost) {
  k_ghost_emit_body(P, B, cap_fluid, cap_ghost, blockIdx.x, gridDim.x);
}

// y-direction ghosts of the x-ghosts (corners), in x-ghost order; sets the ghost count.
__device__ __forceinline__ void k_ghost_yx_body(const DevParams& P, Buffers& B, int cap_ghost, int blk_, int nblk_) {
  int g = blk_ * blockDim.x + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  const int nxg = B.counters[C_NXG], nyf = B.counters[C_NYF], nyx = B.counters[C_NYX];
  if (g == 0) {
    int ng = nxg + nyf + nyx;
    if (ng > cap_ghost) { atomicOr(&B.counters[C_ERR], E_GHOST_CAP); ng = cap_ghost; }
    B.counters[C_NGHOST] = ng;
  }
  if (g >= cap_ghost || g >= nxg || B.cnt_yx[g] == 0) return;
  const int src = N + g;
  const double2 p = B.g_pos[src];
  const double r3h = P.three_h;
  int k = nxg + nyf + B.off_yx[g];
  if (p.y <= (0.0 + r3h) && p.y >= 0.0) {
    if (k < cap_ghost) copy_particle(B, N + k, src, make_double2(p.x, p.y + P.box_y));
    else atomicOr(&B.counters[C_ERR], E_GHOST_CAP);
    ++k;
  }
  if (p.y >= (P.box_y - r3h) && p.y <= P.box_y) {
    if (k < cap_ghost) copy_particle(B, N + k, src, make_double2(p.x, p.y - P.box_y));
    else atomicOr(&B.counters[C_ERR], E_GHOST_CAP);
  }
}

__global__ void k_ghost_yx(DevParams P, Buffers B, int cap_ghost) {
  k_ghost_yx_body(P, B, cap_ghost, blockIdx.x, gridDim.x);
}

void launch_period_boundary(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, int cap_ghost) {
  const int nb = grid_for(cap_fluid);              // gblk holds 2 * nb ints
  hipLaunchKernelGGL(k_wrap_count, dim3(nb), dim3(kBlock), 0, s, P, B, cap_fluid);
  hipLaunchKernelGGL(k_scan_top2, dim3(2), dim3(1024), 0, s, B.gblk, nb, B.counters + C_NXG, B.counters + C_NYF);
  launch_fill_i32(s, B.cnt_yx, cap_ghost, 0);
  hipLaunchKernelGGL(k_ghost_emit, dim3(nb), dim3(kBlock), 0, s, P, B, cap_fluid, cap_ghost);
  exclusive_scan(s, B.cnt_yx, cap_ghost, B.off_yx, B.scan_blocks, B.counters + C_NYX);
  hipLaunchKernelGGL(k_ghost_yx, dim3(grid_for(cap_ghost)), dim3(kBlock), 0, s, P, B, cap_ghost);
}

// ---------------------------------------------------------------------------
// Devide_block_and_send_tasks.f90 on one rank
// ---------------------------------------------------------------------------
__global__ void k_analytic(DevParams P, Buffers B, int cap_fluid, double pi, double t, double Re) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= cap_fluid || i >= B.counters[C_NFLUID]) return;
  const double2 p = B.g_pos[i];
  double ua = P.U * sin(2 * pi * p.x) * cos(2 * pi * p.y) * exp(-8 * (pi * pi) / Re * t);
  double va = -P.U * cos(2 * pi * p.x) * sin(2 * pi * p.y) * exp(-8 * (pi * pi) / Re * t);
  double pa = 0.25 * P.rho0 * (P.U * P.U) * (cos(4 * pi * p.x) + cos(4 * pi * p.y)) *
              exp(-16 * (pi * pi) / Re * t);
  B.ana[4 * i + 0] = ua;
  B.ana[4 * i + 1] = va;
  B.ana[4 * i + 2] = pa;
  B.ana[4 * i + 3] = sqrt(ua * ua + va * va);
}

// :75-83 (y<0 -> ill) and :123-162 binning; key = column-major cell index.
__device__ __forceinline__ void k_cell_key_body(const DevParams& P, Buffers& B, int cap_global, int blk_, int nblk_) {
  const int g = blk_ * blockDim.x + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  const int ng = min(N + B.counters[C_NGHOST], cap_global);
  const bool valid = g < ng;                       // every lane stays for the wave operations below
  int key = -1;
  if (valid) {
    int t = B.g_type[g];
    const double2 p = B.g_pos[g];
    if (g < N && t != 120 && p.y < 0.0) { t = 120; B.g_type[g] = 120; }
    if (t != 120 && t != 0) {
      const int xn = (int)floor((p.x - P.chain_ox) / P.chain_d + 1);
      const int yn = (int)floor((p.y - P.chain_oy) / P.chain_d + 1);
      if (xn < 1 || xn > P.chain_nx || yn < 1 || yn > P.chain_ny) {
        B.g_type[g] = 120;
        t = 120;
      } else {
        key = (xn - 1) * P.chain_ny + (yn - 1);
      }
    }
    if (t == 120 && g < N) atomicAdd(&B.counters[C_NILL], 1);   // every fluid particle typed 120
    B.g_key[g] = key;
    B.l_of_g[g] = -1;                              // set again by k_cell_rank for kept particles
  }
  // Slots: runs of adjacent lanes with the same key (neighbouring gids share
  // cells) reserve their slots with one atomic from the run's first lane, all
  // runs of the wave at once. Slots only need to be distinct: the within-cell
  // order comes from the rank by gid in k_cell_rank.
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1);
  const bool cont = key >= 0 && lane > 0 && prev == key;
  const bool lead = key >= 0 && !cont;
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const uint64_t breaks = __ballot(!cont) & ~upto;  // run starts (or unkeyed lanes) above this lane
  const uint64_t leads = __ballot(lead) & upto;     // run starts at or below this lane
  const int next = breaks ? __builtin_ctzll(breaks) : 64;
  const int my_lead = leads ? 63 - __builtin_clzll(leads) : lane;
  int base = 0;
  if (lead) base = atomicAdd(&B.cell_count[key], next - lane);
  base = __shfl(base, my_lead);
  if (key >= 0) B.g_slot[g] = base + (lane - my_lead);
}

__global__ void k_cell_key(DevParams P, Buffers B, int cap_global) {
  k_cell_key_body(P, B, cap_global, blockIdx.x, gridDim.x);
}

__device__ __forceinline__ void k_cell_scatter_body(Buffers& B, int cap_global, int blk_, int nblk_) {
  int g = blk_ * blockDim.x + threadIdx.x;
  if (g >= cap_global) return;
  const int ng = B.counters[C_NFLUID] + B.counters[C_NGHOST];
  if (g >= ng) return;
  const int key = B.g_key[g];
  if (key < 0) return;
  B.cell_members[B.cell_start[key] + B.g_slot[g]] = g;
}

__global__ void k_cell_scatter(Buffers B, int cap_global) {
  k_cell_scatter_body(B, cap_global, blockIdx.x, gridDim.x);
}

// Rank inside the cell by gid gives the reference's within-cell order; then the
// global -> subdomain copy (:666-765) with EOS and volume (single_step_compute.f90:62-73,
// Assign_boundary_particles.f90:303-321).
// One wavefront per cell: each lane holds one member gid, its rank among the cell's
// members is the count of smaller gids (broadcast lane by lane), which gives the
// reference's within-cell order; then the global -> subdomain copy (:666-765) with
// EOS and volume (single_step_compute.f90:62-73, Assign_boundary_particles.f90:303-321)
// into the cell's contiguous slice of the local arrays.
__device__ __forceinline__ void gather_one(const DevParams& P, Buffers& B, int g, int l, int key) {
  const double2 pos = B.g_pos[g];
  B.l_pos[l] = pos;
  B.l_pos32[l] = make_float2((float)pos.x, (float)pos.y);
  const double2 vel = B.g_vel[g];
  B.l_vel[l] = vel;
  const double rho = B.g_rho[g];
  const double m = B.g_mass[g];
  const int t = B.g_type[g];
  B.l_mass[l] = m;
  B.l_type[l] = t;
  double p, c;
  if (t != 120) tait(P, rho, p, c);
  else { p = B.g_p[g]; c = B.g_c[g]; }
  B.l_pc[l] = make_double2(p, c);
  B.l_rv[l] = make_double2(rho, m / rho);
  if (P.mixed) {
    write_r0(P, B, l, vel, rho, m, t);
    B.m_r1[l] = make_float4(0.f, 0.f, 0.f, (float)B.g_mass[l]);   // particle_mass(j) with j local
    const float2 off = cell_offset(P, pos, key);
    B.m_r2[l] = make_float4(0.f, 0.f, off.x, off.y);
  }
}

// One wavefront per cell ranks the members by gid and writes the local order
// (name, cell, and the inverse map l_of_g); the state copy runs afterwards with
// one thread per local slot, so no lanes idle on the wide copy.
__device__ __forceinline__ void k_cell_rank_body(Buffers& B, int n_cells, int blk_, int nblk_) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int cell = blk_ * (kBlock / 64) + wave;
  if (cell >= n_cells) return;
  const int s0 = B.cell_start[cell];
  const int cnt = B.cell_start[cell + 1] - s0;
  if (cnt == 0) return;
  for (int base = 0; base < cnt; base += 64) {
    const int idx = base + lane;
    const int my = idx < cnt ? B.cell_members[s0 + idx] : 0x7fffffff;
    int rank = 0;
    for (int b2 = 0; b2 < cnt; b2 += 64) {
      const int v = (b2 + lane < cnt) ? B.cell_members[s0 + b2 + lane] : 0x7fffffff;
      const int m = min(64, cnt - b2);
      for (int k = 0; k < m; ++k) rank += (__builtin_amdgcn_readlane(v, k) < my);
    }
    if (idx < cnt) {
      const int l = s0 + rank;
      B.name[l] = my;
      B.l_cell[l] = cell;
      B.l_of_g[my] = l;
    }
  }
}

__global__ void __launch_bounds__(kBlock) k_cell_rank(Buffers B, int n_cells) {
  k_cell_rank_body(B, n_cells, blockIdx.x, gridDim.x);
}

__global__ void __launch_bounds__(kBlock) k_local_gather(DevParams P, Buffers B) {
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l >= B.counters[C_NLOCAL]) return;
  gather_one(P, B, B.name[l], l, B.l_cell[l]);
}

// Largest cell population (diagnostic, on demand): one block folds cell_start.
__global__ void k_max_cell(Buffers B, int n_cells) {
  __shared__ int lds[kBlock];
  int m = 0;
  for (int c = threadIdx.x; c < n_cells; c += kBlock) m = max(m, B.cell_start[c + 1] - B.cell_start[c]);
  lds[threadIdx.x] = m;
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) lds[threadIdx.x] = max(lds[threadIdx.x], lds[threadIdx.x + off]);
    __syncthreads();
  }
  if (threadIdx.x == 0) B.counters[C_MAXCELL] = lds[0];
}

void launch_max_cell(hipStream_t s, Buffers& B, int n_cells) {
  hipLaunchKernelGGL(k_max_cell, dim3(1), dim3(kBlock), 0, s, B, n_cells);
}

__global__ void __launch_bounds__(kBlock) k_gather_predict(DevParams P, Buffers B, int save_before, int lean);

void launch_devide_block(hipStream_t s, const DevParams& P, Buffers& B, int cap_global, int n_cells,
                         bool analytic, double t_nondim, int predict) {
  const double pi = 3.14159265358979324;
  if (analytic) {
    const double Re = P.U * P.box_x / P.nu;        // information_module.f90:60
    hipLaunchKernelGGL(k_analytic, dim3(grid_for(cap_global)), dim3(kBlock), 0, s, P, B, cap_global, pi,
                       t_nondim, Re);
  }
  launch_fill_i32(s, B.cell_count, n_cells, 0);
  hipLaunchKernelGGL(k_cell_key, dim3(grid_for(cap_global)), dim3(kBlock), 0, s, P, B, cap_global);
  exclusive_scan(s, B.cell_count, n_cells, B.cell_start, B.scan_blocks, B.counters + C_NLOCAL,
                 B.cell_start + n_cells);
  hipLaunchKernelGGL(k_cell_scatter, dim3(grid_for(cap_global)), dim3(kBlock), 0, s, B, cap_global);
  hipLaunchKernelGGL(k_cell_rank, dim3((n_cells + kBlock / 64 - 1) / (kBlock / 64)), dim3(kBlock), 0, s, B,
                     n_cells);
  if (predict) hipLaunchKernelGGL(k_gather_predict, dim3(grid_for(cap_global)), dim3(kBlock), 0, s, P, B,
                                  (int)(predict == 2), (int)(predict == 3));
  else hipLaunchKernelGGL(k_local_gather, dim3(grid_for(cap_global)), dim3(kBlock), 0, s, P, B);
}

// ---------------------------------------------------------------------------
// Stage updates of time_iteration.f90 (fluid particles only; whole-array saves)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void eos_volume(const DevParams& P, Buffers& B, int s, double rho) {
  double p, c;
  tait(P, rho, p, c);
  const double vol = B.l_mass[s] / rho;
  B.l_pc[s] = make_double2(p, c);
  B.l_rv[s] = make_double2(rho, vol);
  if (P.mixed) write_r0(P, B, s, B.l_vel[s], rho, B.l_mass[s], B.l_type[s]);
}

// Positions moved: the FP32 copy for the neighbour screen and, in the mixed
// form, the offsets to the binned cell's corner (grad rho is rewritten by the
// next pass A).
__device__ __forceinline__ void write_pos_mirrors(const DevParams& P, Buffers& B, int s, double2 x) {
  B.l_pos32[s] = make_float2((float)x.x, (float)x.y);
  if (P.mixed) {
    const float2 off = cell_offset(P, x, B.l_cell[s]);
    B.m_r2[s] = make_float4(0.f, 0.f, off.x, off.y);
  }
}

// time_iteration.f90:553-566: the predictor from the step-start state (bx, bv, brho)
// and the rates read by the current local index.
__device__ __forceinline__ void pc_predict(double dt, double drho, double2 a, double2 avg, double brho, double2 bv,
                                           double2 bx, double& rho, double2& v, double2& x) {
  rho = brho + drho * dt;
  v.x = bv.x + dt * a.x + avg.x;
  x.x = bx.x + dt * v.x;
  v.y = bv.y + dt * a.y + avg.y;
  x.y = bx.y + dt * v.y;
}

// :590-599: the corrector; (td, ta, tv) are the predictor's rates and velocity.
__device__ __forceinline__ void pc_correct(double dt, double drho, double2 a, double2 avg, double td, double2 ta,
                                           double2 tv, double brho, double2 bv, double2 bx, double& rho, double2& v,
                                           double2& x) {
  rho = brho + 0.5 * (drho + td) * dt;
  v.x = bv.x + 0.5 * dt * (a.x + ta.x) + avg.x;
  x.x = bx.x + 0.5 * dt * (v.x + tv.x) + 0.5 * dt * dt * a.x;
  v.y = bv.y + 0.5 * dt * (a.y + ta.y) + avg.y;
  x.y = bx.y + 0.5 * dt * (v.y + tv.y) + 0.5 * dt * dt * a.y;
}

// The 3x3 stencil of 4h cells holds every pair within 3h while each particle
// stays within h/2 of its binned cell (3h + h/2 + h/2 = 4h); flag a breach.
__device__ __forceinline__ void rebin_check(const DevParams& P, Buffers& B, int cell, double2 x) {
  const int cx = cell / P.chain_ny, cy = cell - cx * P.chain_ny;
  const double x0 = P.chain_ox + cx * P.chain_d, y0 = P.chain_oy + cy * P.chain_d;
  const double m = P.rebin_m;
  if (x.x < x0 - m || x.x > x0 + P.chain_d + m || x.y < y0 - m || x.y > y0 + P.chain_d + m)
    atomicOr(&B.counters[C_REBIN], 1);
}

__device__ __forceinline__ void k_stage_body(const DevParams& P, Buffers& B, int kind, int blk_, int nblk_) {
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const double dt = P.dt;
  const int t = B.l_type[s];
  const double drho = B.drhodt[s];
  const double2 a = B.acc[s];
  const double2 avg = (P.epsilon_xsph != 0.0) ? B.avg[s] : make_double2(0.0, 0.0);
  double rho = B.l_rv[s].x;
  double2 v = B.l_vel[s];
  double2 x = B.l_pos[s];
  if (kind == ST_SAVE_BEFORE) {
    B.b_rho[s] = rho; B.b_vel[s] = v; B.b_pos[s] = x;
    return;
  }
  if (kind == ST_EOS) {
    if (t != 120) {
      eos_volume(P, B, s, rho);
      write_pos_mirrors(P, B, s, x);
    }
    return;
  }
  if (kind == ST_PC_PREDICT || kind == ST_RK_HALF0) {      // before_* = state (whole arrays)
    B.b_rho[s] = rho; B.b_vel[s] = v; B.b_pos[s] = x;
  }
  const double brho = B.b_rho[s];
  const double2 bv = B.b_vel[s], bx = B.b_pos[s];
  if (kind == ST_PC_PREDICT) {                              // :553-566
    if (t == 2) {
      pc_predict(dt, drho, a, avg, brho, bv, bx, rho, v, x);
      B.l_vel[s] = v; B.l_pos[s] = x;
      write_pos_mirrors(P, B, s, x);
      eos_volume(P, B, s, rho);
    }
    B.t_drho[s] = drho; B.t_acc[s] = a;                     // :573-576 (temp velocity = l_vel)
  } else if (kind == ST_PC_CORRECT) {                       // :590-599
    if (t == 2) {
      pc_correct(dt, drho, a, avg, B.t_drho[s], B.t_acc[s], v, brho, bv, bx, rho, v, x);
      B.l_vel[s] = v; B.l_pos[s] = x;
      B.l_rv[s] = make_double2(rho, B.l_rv[s].y);           // p, c, volume keep the last evaluation's values
    }
  } else if (kind == ST_RK_HALF0 || kind == ST_RK_HALF1 || kind == ST_RK_FULL2) {
    const bool half = kind != ST_RK_FULL2;
    if (t == 2) {
      if (half) {
        rho = brho + drho * dt / 2;
        v.x = bv.x + 0.5 * dt * a.x + avg.x;
        x.x = bx.x + 0.5 * dt * v.x;
        v.y = bv.y + 0.5 * dt * a.y + avg.y;
        x.y = bx.y + 0.5 * dt * v.y;
      } else {
        rho = brho + drho * dt;
        v.x = bv.x + dt * a.x + avg.x;
        x.x = bx.x + dt * v.x;
        v.y = bv.y + dt * a.y + avg.y;
        x.y = bx.y + dt * v.y;
      }
      B.l_vel[s] = v; B.l_pos[s] = x;
      write_pos_mirrors(P, B, s, x);
      eos_volume(P, B, s, rho);
    }
    const int k = kind == ST_RK_HALF0 ? 0 : (kind == ST_RK_HALF1 ? 1 : 2);
    const size_t off = (size_t)k * (size_t)B.kstride;
    B.k_vel[off + s] = v; B.k_drho[off + s] = drho; B.k_acc[off + s] = a;
  } else if (kind == ST_RK_FINAL) {                         // :918-928
    if (t == 2) {
      const size_t st = (size_t)B.kstride;
      const double k1 = B.k_drho[s], k2 = B.k_drho[st + s], k3 = B.k_drho[2 * st + s];
      const double2 a1 = B.k_acc[s], a2 = B.k_acc[st + s], a3 = B.k_acc[2 * st + s];
      const double2 v1 = B.k_vel[s], v2 = B.k_vel[st + s], v3 = B.k_vel[2 * st + s];
      rho = brho + (k1 + 2 * k2 + 2 * k3 + drho) * dt / 6.0;
      v.x = bv.x + (a1.x + 2 * a2.x + 2 * a3.x + a.x) * dt / 6.0;
      x.x = bx.x + (v1.x + 2 * v2.x + 2 * v3.x + v.x) * dt / 6.0 + 0.5 * dt * dt * a.x;
      v.y = bv.y + (a1.y + 2 * a2.y + 2 * a3.y + a.y) * dt / 6.0;
      x.y = bx.y + (v1.y + 2 * v2.y + 2 * v3.y + v.y) * dt / 6.0 + 0.5 * dt * dt * a.y;
      B.l_vel[s] = v; B.l_pos[s] = x;
      B.l_rv[s] = make_double2(rho, B.l_rv[s].y);
    }
  }
  if (t == 2) rebin_check(P, B, B.l_cell[s], x);
}

__global__ void k_stage(DevParams P, Buffers B, int kind) {
  k_stage_body(P, B, kind, blockIdx.x, gridDim.x);
}

void launch_stage(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int kind) {
  hipLaunchKernelGGL(k_stage, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, kind);
}

// The global -> subdomain copy and the predictor in one pass (steps after the
// first, whose predictor uses the previous step's rates): the gathered state is
// the step-start state (kept as before_* when save_before: k_stage's corrector
// reads it; k_correct_refresh reads the global arrays instead), the local arrays
// receive the predicted state.
// Same arithmetic as k_local_gather followed by k_stage(ST_PC_PREDICT).
__device__ __forceinline__ void k_gather_predict_body(const DevParams& P, Buffers& B, int save_before, int lean, int blk_, int nblk_) {
  const int l = blk_ * blockDim.x + threadIdx.x;
  if (l >= B.counters[C_NLOCAL]) return;
  const int g = B.name[l], key = B.l_cell[l];
  const double2 pos = B.g_pos[g];
  const double2 vel = B.g_vel[g];
  const double rho0 = B.g_rho[g];
  const double m = B.g_mass[g];
  const int t = B.g_type[g];
  if (save_before) { B.b_rho[l] = rho0; B.b_vel[l] = vel; B.b_pos[l] = pos; }
  const double drho = B.drhodt[l];
  const double2 a = B.acc[l];
  B.t_drho[l] = drho; B.t_acc[l] = a;
  double rho = rho0;
  double2 v = vel, x = pos;
  if (t == 2) {
    const double2 avg = (P.epsilon_xsph != 0.0) ? B.avg[l] : make_double2(0.0, 0.0);
    pc_predict(P.dt, drho, a, avg, rho0, vel, pos, rho, v, x);
    rebin_check(P, B, key, x);
  }
  B.l_pos[l] = x;
  B.l_pos32[l] = make_float2((float)x.x, (float)x.y);
  B.l_type[l] = t;
  if (!lean) {                  // the FP32-mixed passes read the records below instead
    B.l_vel[l] = v;
    B.l_mass[l] = m;
    double p, c;
    if (t != 120) tait(P, rho, p, c);
    else { p = B.g_p[g]; c = B.g_c[g]; }
    B.l_pc[l] = make_double2(p, c);
    B.l_rv[l] = make_double2(rho, m / rho);
  }
  if (P.mixed) {
    write_r0(P, B, l, v, rho, m, t);
    B.m_r1[l] = make_float4(0.f, 0.f, 0.f, (float)B.g_mass[l]);   // particle_mass(j) with j local
    const float2 off = cell_offset(P, x, key);
    B.m_r2[l] = make_float4(0.f, 0.f, off.x, off.y);
  }
}

__global__ void __launch_bounds__(kBlock) k_gather_predict(DevParams P, Buffers B, int save_before, int lean) {
  k_gather_predict_body(P, B, save_before, lean, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------------------
// Reinitialize_particle_density.f90:66-182 + MLS_Interpolation_In_Subdomain_Domain.f90
// (order 6): A = sum_k w_k p_k p_k^T + alpha^3 I with p_k = (1, dx, dy, dx^2, dx dy, dy^2)/h
// powers, rho_i = sum_k rho_k Phi_k with Phi_k = w_k (A^-1 p_k)_1 = (A^-1 e_1) . (w_k p_k).
// The reference inverts A by full-pivot Gauss-Jordan (BRINV, function_module.f90:568-639)
// and keeps row 1; A is symmetric positive definite here (w_k >= 0 inside the support,
// alpha^3 = 1e-12 on the diagonal), so y = A^-1 e_1 comes from a Cholesky solve in
// registers: same row, different rounding (relative ~ cond(A) * eps). Fewer than 9
// neighbours fall back to the Shepard average, as the reference does (:284-295).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void k_mls_body(const DevParams& P, Buffers& B, int blk_, int nblk_) {
  // the block's candidate positions in LDS (as the evaluation's pass A stages them)
  __shared__ NbStage st;
  __shared__ float2 lds32[kStageCap];
  __shared__ double2 lds64[kStageCap];
  const int s = blk_ * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  stage_candidates(P, B.cell_start, B.l_cell, B.l_pos32, B.l_pos, n, st, lds32, lds64, blk_);
  if (s >= n) return;
  const int ts = B.l_type[s];
  if (ts != 2) { B.mls_rho[s] = 0.0; return; }     // :69 rho=0 for every non-fluid particle
  const double2 ps = B.l_pos[s];
  // upper triangle of A, row-major over (m <= l): a00 a01 .. a05 a11 .. a55
  double A[21];
#pragma unroll
  for (int k = 0; k < 21; ++k) A[k] = 0.0;
  double bv[6] = {0, 0, 0, 0, 0, 0};
  double sw = 0.0, srw = 0.0;
  int ne = 0;
  for_each_neighbour(P, B.cell_start, B.l_pos, B.l_pos32, s, ps, B.l_cell[s], [&](int o, int d, int li) __attribute__((always_inline)) {
    const double2 po = li >= 0 ? lds64[li] : B.l_pos[o];
    const double pdx = ps.x - po.x, pdy = ps.y - po.y;
    const double r = sqrt(pdx * pdx + pdy * pdy);
    double w, dwx, dwy;
    kernel_w_dw(P, r, pdx, pdy, w, dwx, dwy);
    const double a = (po.x - ps.x) * P.inv_h, b = (po.y - ps.y) * P.inv_h;
    const double pv[6] = {1.0, a, b, a * a, a * b, b * b};
    const double rho_o = B.l_rv[o].x;
    int k = 0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double bm = pv[m] * w;
#pragma unroll
      for (int l = m; l < 6; ++l) { A[k] = A[k] + bm * pv[l]; ++k; }
      bv[m] = bv[m] + rho_o * bm;
    }
    sw = sw + w;
    srw = srw + rho_o * w;
    ++ne;
  }, nullptr, nullptr, 0, &st, lds32, lds64);
  const double al3 = (P.mls_alpha * P.mls_alpha) * P.mls_alpha;
  double rho = 0.0;
  bool ok = ne > 8;
  if (ok) {
    double y[6];
    ok = mls6_first_row(A, al3, y);
#pragma unroll
    for (int l = 0; l < 6; ++l) rho = rho + y[l] * bv[l];
  }
  if (!ok) {
    rho = srw / sw;
    atomicAdd(&B.counters[C_MLSFB], 1);
  }
  B.mls_rho[s] = rho;
}

__global__ void __launch_bounds__(kBlock) k_mls(DevParams P, Buffers B) {
  k_mls_body(P, B, xcd_block(), gridDim.x);
}

void launch_mls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_mls, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B);
}

// MLS_Interpolation_In_Gobal_Domain.f90, order 6, at (ex, ey): y = row 1 of the
// moment inverse (mls = true) or Shepard weights (10 or fewer particles in the
// support, or no Cholesky factor); sw = sum of W. Returns the support count.
__device__ __forceinline__ int point_mls_weights(const DevParams& P, const Buffers& B, double ex, double ey,
                                                 int n_max, double* y, double& sw, bool& mls) {
  double A[21];
#pragma unroll
  for (int q = 0; q < 21; ++q) A[q] = 0.0;
  sw = 0.0;
  int nt = 0;
  for_each_support(P, B, ex, ey, n_max, [&](int g, double2 pj, double w) {
    (void)g;
    const double a = (pj.x - ex) / P.h, b = (pj.y - ey) / P.h;
    const double pv[6] = {1.0, a, b, a * a, a * b, b * b};
    int q = 0;
#pragma unroll
    for (int m = 0; m < 6; ++m) {
      const double bm = pv[m] * w;
#pragma unroll
      for (int l = m; l < 6; ++l) { A[q] = A[q] + bm * pv[l]; ++q; }
    }
    sw = sw + w;
    ++nt;
  });
#pragma unroll
  for (int m = 0; m < 6; ++m) y[m] = 0.0;
  mls = nt > 10;
  if (mls) mls = mls6_first_row(A, (P.mls_alpha * P.mls_alpha) * P.mls_alpha, y);
  return nt;
}

// Fai of one support particle: w (A^-1 p)_1 = sum_m y_m (p_m w), or w / sum W.
__device__ __forceinline__ double point_mls_fai(const DevParams& P, const double* y, double sw, bool mls, double ex,
                                                double ey, double2 pj, double w) {
  if (!mls) return w / sw;
  const double a = (pj.x - ex) / P.h, b = (pj.y - ey) / P.h;
  const double pv[6] = {1.0, a, b, a * a, a * b, b * b};
  double fai = 0.0;
#pragma unroll
  for (int m = 0; m < 6; ++m) fai = fai + y[m] * (pv[m] * w);
  return fai;
}

// ---------------------------------------------------------------------------
// Calculation_For_PostProceeding.f90:96-300 on the element nodes of
// Distribute_Grid_For_Postproceeding.f90:30-39 (node k = j + (i-1) nnx at
// ((j-1) edx, (i-1) edx): both coordinates use element_dx, as there), after the
// step's refresh. Candidates are the fluid particles (gid < N) of the node's 3x3
// cells in the step-start binning, as the reference's particle_in_chain (each
// particle is within h/2 of its binned cell, so the stencil holds every particle
// within 3h); positions, velocity, pressure and vorticity are the refreshed global
// ones. MLS_Interpolation_In_Gobal_Domain.f90, order 6: Fai_k = w_k (A^-1 p_k)_1,
// Shepard weights w_k / sum w when 10 or fewer particles are in the support or A
// has no Cholesky factor; at least 3 are needed. Level set from the nearest surface
// particle (:125-170). out: 9 blocks of n_nodes (u, v, p, vorticity_z, Fai,
// analytic u, v, p, |v| of Devide_block_and_send_tasks.f90:104-115).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_grid_mls(DevParams P, Buffers B, GridArgs G) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= G.n_nodes) return;
  const size_t nn = (size_t)G.n_nodes;
  const int N = B.counters[C_NFLUID];
  const double ex = (k % G.nnx) * G.edx, ey = (k / G.nnx) * G.edx;
  const double pi = 3.14159265358979324;
  const double au = P.U * sin(2 * pi * ex) * cos(2 * pi * ey) * exp(-8 * (pi * pi) / G.Re * G.t_nondim);
  const double av = -P.U * cos(2 * pi * ex) * sin(2 * pi * ey) * exp(-8 * (pi * pi) / G.Re * G.t_nondim);
  G.out[5 * nn + k] = au;
  G.out[6 * nn + k] = av;
  G.out[7 * nn + k] = 0.25 * P.rho0 * (P.U * P.U) * (cos(4 * pi * ex) + cos(4 * pi * ey)) *
                      exp(-16 * (pi * pi) / G.Re * G.t_nondim);
  G.out[8 * nn + k] = sqrt(au * au + av * av);
  double f = 1.0;
  if (G.n_surf > 0) {
    double best = 1000000000.0, dfx = 0.0, dfy = 0.0;
    int idx = 0;
    for (int q = 0; q < G.n_surf; ++q) {
      const double2 pj = B.g_pos[G.surf_gid[q]];
      const double pdx = ex - pj.x, pdy = ey - pj.y;
      const double d = sqrt(pdx * pdx + pdy * pdy);
      if (d <= best) { best = d; idx = q; dfx = -pdx; dfy = -pdy; }
    }
    const double2 nr = G.surf_nrm[idx];
    const double det = dfx * nr.x + dfy * nr.y;
    if (det <= -2 * P.h) f = -1.0;
    else if (det >= 2 * P.h) f = 1.0;
    else f = det / (2 * P.h);
  }
  G.out[4 * nn + k] = f;
  double su = 0.0, sv = 0.0, sp = 0.0, so = 0.0;
  if (f > -0.9) {
    double y[6], sw;
    bool mls;
    const int nt = point_mls_weights(P, B, ex, ey, N, y, sw, mls);
    if (nt >= 3)
      for_each_support(P, B, ex, ey, N, [&](int g, double2 pj, double w) {
        const double fai = point_mls_fai(P, y, sw, mls, ex, ey, pj, w);
        const double2 vj = B.g_vel[g];
        sp = sp + B.g_p[g] * fai;
        su = su + vj.x * fai;
        sv = sv + vj.y * fai;
        so = so + B.g_vort[g] * fai;
      });
  }
  G.out[0 * nn + k] = su;
  G.out[1 * nn + k] = sv;
  G.out[2 * nn + k] = sp;
  G.out[3 * nn + k] = so;
}

void launch_grid_mls(hipStream_t s, const DevParams& P, Buffers& B, const GridArgs& G) {
  hipLaunchKernelGGL(k_grid_mls, dim3(grid_for(G.n_nodes)), dim3(kBlock), 0, s, P, B, G);
}

// Sampling_subroutine.f90:21-177 (subroutine Sampling) at n sampling points:
// candidates are the fluid particles (gid < N); with more than 3 in the support,
// the average pressure over the support, the order-6 MLS pressure (minus
// g rho_j (y_s - y_j) Fai_j when gravity is on) and the MLS velocity; zeros
// otherwise. out: 4 blocks of n (average p, MLS p, u, v).
__global__ void __launch_bounds__(kBlock) k_probe_mls(DevParams P, Buffers B, ProbeArgs Q) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Q.n) return;
  const int N = B.counters[C_NFLUID];
  const double2 xs = Q.pos[k];
  double y[6], sw;
  bool mls;
  const int nt = point_mls_weights(P, B, xs.x, xs.y, N, y, sw, mls);
  double avg = 0.0, mp = 0.0, su = 0.0, sv = 0.0;
  if (nt > 3) {
    for_each_support(P, B, xs.x, xs.y, N, [&](int g, double2 pj, double w) {
      const double fai = point_mls_fai(P, y, sw, mls, xs.x, xs.y, pj, w);
      const double pj_p = B.g_p[g];
      if (Q.gravity != 0.0) mp = mp + pj_p * fai - Q.gravity * B.g_rho[g] * (xs.y - pj.y) * fai;
      else mp = mp + pj_p * fai;
      const double2 vj = B.g_vel[g];
      su = su + vj.x * fai;
      sv = sv + vj.y * fai;
      avg = avg + pj_p;
    });
    avg = avg / nt;
  }
  const size_t n = (size_t)Q.n;
  Q.out[k] = avg;
  Q.out[n + k] = mp;
  Q.out[2 * n + k] = su;
  Q.out[3 * n + k] = sv;
}

void launch_probe_mls(hipStream_t s, const DevParams& P, Buffers& B, const ProbeArgs& Q) {
  hipLaunchKernelGGL(k_probe_mls, dim3(grid_for(Q.n)), dim3(kBlock), 0, s, P, B, Q);
}

// Sampling_subroutine.f90:209-222: over the fluid particles, the wave front (largest
// x, from 0) and, for each wave probe, the largest y (from 0) of the particles within
// 0.55 (REAL(4)) dx of the probe's x. Maxima of non-negative doubles are taken on their bit
// patterns with integer atomics: exact and order-independent. out_bits[0] = front,
// out_bits[1 + j] = probe j (zeroed by the caller).
__global__ void __launch_bounds__(kBlock) k_wave_probes(DevParams P, Buffers B, const double* probe_x, int n_probes,
                                                        unsigned long long* out_bits) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int N = B.counters[C_NFLUID];
  double front = 0.0;
  double2 x = make_double2(0.0, 0.0);
  const bool ok = g < N;
  if (ok) {
    x = B.g_pos[g];
    if (x.x >= front) front = x.x;
  }
  // a wave-wide maximum first, then one atomic per wave
  unsigned long long fb = __double_as_longlong(front);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(fb, off, 64);
    fb = o > fb ? o : fb;
  }
  if ((threadIdx.x & 63) == 0 && fb != 0ull) atomicMax(&out_bits[0], fb);
  for (int j = 0; j < n_probes; ++j) {
    double e = 0.0;
    if (ok && fabs(probe_x[j] - x.x) <= (double)0.55f * P.dx && x.y > e) e = x.y;
    unsigned long long eb = __double_as_longlong(e);
    for (int off = 32; off > 0; off >>= 1) {
      const unsigned long long o = __shfl_xor(eb, off, 64);
      eb = o > eb ? o : eb;
    }
    if ((threadIdx.x & 63) == 0 && eb != 0ull) atomicMax(&out_bits[1 + j], eb);
  }
}

// single_step_compute.f90:174-260, the end damping zone, after the rate assembly of an
// evaluation: fluid particles with box_x - x <= L and x <= box_x (local positions of
// the evaluation) get their acceleration and XSPH velocity scaled by the zone's factor.
// The reference also scales subdomain_particle_velocity there; every stage update
// overwrites that velocity before it is read again (time_iteration.f90:559-565,
// :590-594 and the RK-variant stages), so only the rates are scaled here.
__global__ void __launch_bounds__(kBlock) k_end_damping(DevParams P, Buffers B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL] || B.l_type[s] != 2) return;
  const double2 x = B.l_pos[s];
  const double Lx = P.box_x, L = P.damp_len, d = P.water_depth;
  if (!((Lx - x.x) <= L && x.x <= Lx)) return;
  double f;
  if (P.damping == 1) {                                   // Cosin_Type (:181-201)
    const double pi = 3.14159265358979324;
    f = 1 - 0.5 * P.damp_amp * (1 - cos(pi * (x.x + L - Lx) / L));
  } else if (P.damping == 2) {                            // OpenFoam_Type (:203-228)
    f = (exp(pow((Lx - x.x) / L, 3.5)) - 1.0) / (P.e32 - 1.0) * (x.y / d);
  } else if (P.damping == 4) {                            // linear ramp (no reference form)
    f = 1 - P.damp_amp * ((x.x + L - Lx) / L);
  } else {                                                // DualSPH_Type (:230-255)
    const double q = (x.x + L - Lx) / L;
    f = 1 - 6 * (q * q) * ((d - x.y) / d);
  }
  const double2 a = B.acc[s];
  B.acc[s] = make_double2(a.x * f, a.y * f);
  if (P.epsilon_xsph != 0.0) {
    const double2 v = B.avg[s];
    B.avg[s] = make_double2(v.x * f, v.y * f);
  }
}

// Wavemaker_Subroutine.f90:196-215: each fluid particle (not typed 120) behind the
// paddle (x < x_p) is mirrored to x = 2 x_p - x; one within dx in front of it
// (x - x_p < dx, after the mirroring) takes the paddle's horizontal velocity.
__global__ void __launch_bounds__(kBlock) k_wavemaker_pull_back(DevParams P, Buffers B, double xp, double up) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B.counters[C_NFLUID] || B.g_type[g] == 120) return;
  double2 x = B.g_pos[g];
  if (x.x < xp) {
    x.x = 2 * xp - x.x;
    B.g_pos[g] = x;
  }
  if ((x.x - xp) < P.dx) {
    const double2 v = B.g_vel[g];
    B.g_vel[g] = make_double2(up, v.y);
  }
}

void launch_wavemaker_pull_back(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, double xp, double up) {
  hipLaunchKernelGGL(k_wavemaker_pull_back, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, s, P, B, xp, up);
}

// Wavemaker_Subroutine.f90:140-164: the largest fluid y within kernel_scale dx of
// the paddle (from 0), as a bit-pattern maximum (exact, order-independent).
__global__ void __launch_bounds__(kBlock) k_wavemaker_height(DevParams P, Buffers B, double xp,
                                                             unsigned long long* bits) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (g < B.counters[C_NFLUID]) {
    const double2 x = B.g_pos[g];
    if (fabs(xp - x.x) <= 3 * P.dx && x.y > e) e = x.y;
  }
  unsigned long long eb = __double_as_longlong(e);
  for (int off = 32; off > 0; off >>= 1) {
    const unsigned long long o = __shfl_xor(eb, off, 64);
    eb = o > eb ? o : eb;
  }
  if ((threadIdx.x & 63) == 0 && eb != 0ull) atomicMax(bits, eb);
}

// :171-193: wavemaker particle q (gid N + q, layer in g_btype) at x_p - (layer-1) dx
// - dx/2 with the paddle's velocity; below the height (+ dx/2): rho0 (-6.0 g (y - H)
// / c0^2 + 1.0)**(1.0/6.0) with REAL(4) literals and g = 9.81 (information_module.f90:32),
// p = c0^2 rho0 ((rho/rho0)**7.0 - 1)/7, mass rho V0; above it rho0, p 0, mass 0.
__global__ void k_wavemaker_place(DevParams P, Buffers B, int n_wm, double xp, double up,
                                  const unsigned long long* bits) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_wm) return;
  const int i = B.counters[C_NFLUID] + q;
  const double left = __longlong_as_double(*bits) + 0.5 * P.dx;
  double2 x = B.g_pos[i];
  x.x = xp - (B.g_btype[i] - 1) * P.dx - P.dx / 2.0;
  B.g_pos[i] = x;
  const double2 v = B.g_vel[i];
  B.g_vel[i] = make_double2(up, v.y);
  if (x.y > left) {
    B.g_rho[i] = P.rho0; B.g_p[i] = 0.0; B.g_mass[i] = 0.0;
  } else {
    const double rho = P.rho0 * pow(-6.0 * 9.81 * (x.y - left) / P.c0sq + 1.0, (double)(1.0f / 6.0f));
    B.g_rho[i] = rho;
    B.g_p[i] = P.c0sq * P.rho0 * (pow(rho / P.rho0, 7.0) - 1.0) / 7.0;
    B.g_mass[i] = rho * P.vol0;
  }
}

void launch_wavemaker_height(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, double xp,
                             unsigned long long* height_bits) {
  hipLaunchKernelGGL(k_wavemaker_height, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, s, P, B, xp, height_bits);
}

void launch_wavemaker_place(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, int n_wm, double xp,
                            double up, unsigned long long* height_bits) {
  hipLaunchKernelGGL(k_wavemaker_height, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, s, P, B, xp, height_bits);
  if (n_wm > 0)
    hipLaunchKernelGGL(k_wavemaker_place, dim3(grid_for(n_wm)), dim3(kBlock), 0, s, P, B, n_wm, xp, up, height_bits);
}

// ---------------------------------------------------------------------------
// Wave-theory wavemaker and its damping (relaxation) zone
// ---------------------------------------------------------------------------
__device__ __forceinline__ double inv_tait(const DevParams& P, double p) {
  return P.rho0 * pow(1 + 7 * p / (P.c0sq * P.rho0), (double)0.142857142f);
}

// Parallel_Taylor_Green_Vortices_2D_Release.f90:113-143 for gid g < n_total (fluid and
// wavemaker particles): analytic fields = the state, initial position = the position,
// fluid with x <= L in the zone with 1 - (exp((x/L)**3.5) - 1.0)/(exp(1.0) - 1.0)
// (exp(1.0) REAL(4)); L <= 0 leaves the zone empty.
__global__ void k_relax_init(DevParams P, Buffers B, RelaxArgs R, int n_total, double L) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n_total) return;
  const double2 x = B.g_pos[g];
  R.init_pos[g] = x;
  R.an_pos[g] = x;
  R.an_vel[g] = B.g_vel[g];
  R.an_rho[g] = B.g_rho[g];
  R.an_p[g] = B.g_p[g];
  int in = 0;
  double c = 0.0;
  if (g < B.counters[C_NFLUID] && L > 0.0 && x.x <= L) {
    in = 1;
    c = (1 - (exp(pow(x.x / L, 3.5)) - 1.0) / (P.e32 - 1.0));
  }
  R.in_zone[g] = in;
  R.coef[g] = c;
}

// Wavemaker_Subroutine.f90:268-276: fluid below y = 0 typed 120.
__global__ void k_wt_ill(Buffers B) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B.counters[C_NFLUID] || B.g_type[g] == 120) return;
  if (B.g_pos[g].y < 0.0) B.g_type[g] = 120;
}

// :300-346 (Stokes_2nd) and :354-396 (Airy): wavemaker particle q (gid N + q) from wave
// theory at its current position, around its initial position.
__global__ void k_wt_wavemaker(DevParams P, Buffers B, RelaxArgs R, WaveArgs W, int n_wm) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n_wm) return;
  const int i = B.counters[C_NFLUID] + q;
  const double pi = 3.14159265358979324;
  const double soft = W.soft, H = W.H, T = W.T, L = W.L, A = W.A, k = W.k, d = W.d, g = W.g;
  const double skd = W.skd, ckd = W.ckd;
  const double2 x0 = R.init_pos[i];
  const double2 xc = B.g_pos[i];
  const double ang = k * xc.x - W.om * W.t + W.phase;
  const double yo = xc.y;
  double el, u, v;
  double2 xn;
  if (W.wave_type == 1) {
    el = soft * (0.5 * H * cos(ang) + pi * H * H / (8 * L) * ckd / (skd * skd * skd) * (2 + cosh(2 * k * d)) * cos(2 * ang));
    u = pi * H / T * cosh(k * yo) * cos(ang) / skd + 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * cosh(2 * k * yo) * cos(2 * ang) / (skd * skd * skd * skd);
    v = pi * H / T * sinh(k * yo) * sin(ang) / skd + 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * sinh(2 * k * yo) * sin(2 * ang) / (skd * skd * skd * skd);
    xn.x = x0.x - soft * (A * cosh(k * yo) * sin(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * cosh(2 * k * yo) * sin(2 * ang) / (skd * skd * skd * skd));
    xn.y = x0.y + soft * (A * sinh(k * yo) * cos(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * sinh(2 * k * yo) * cos(2 * ang) / (skd * skd * skd * skd));
  } else {
    el = soft * 0.5 * H * cos(ang);
    u = pi * H / T * cosh(k * yo) * cos(ang) / skd;
    v = pi * H / T * sinh(k * yo) * sin(ang) / skd;
    xn.x = x0.x - soft * A * cosh(k * yo) * sin(ang) / skd;
    xn.y = x0.y + soft * A * sinh(k * yo) * cos(ang) / skd;
  }
  B.g_vel[i] = make_double2(soft * u, soft * v);
  B.g_pos[i] = xn;
  const double y = xn.y;
  if (y <= el + d) {
    double pr;
    if (W.wave_type == 1)
      pr = P.rho0 * g * (el + d - y) + soft * (P.rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd + 3.0 / 4.0 * P.rho0 * g * H * (pi * H / L) * (1.0 / sinh(2 * k * d)) * (cosh(2 * k * y) / (skd * skd) - (double)(1.0f / 3.0f)) * cos(2 * ang) - 1.0 / 4.0 * P.rho0 * g * H * (pi * H / L) * cosh(2 * k * y) / sinh(2 * k * d));
    else
      pr = P.rho0 * g * (d + el - y) + soft * (P.rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd);
    B.g_p[i] = pr;
    B.g_rho[i] = inv_tait(P, pr);
    B.g_mass[i] = P.rho0 * P.vol0;
  } else {
    B.g_p[i] = 0.0; B.g_rho[i] = P.rho0; B.g_mass[i] = 0.0;
  }
  B.g_c[i] = P.sqrt_c0sq;
}

// :617-700: the zone's analytic fields, from the previous analytic position.
__global__ void k_wt_relax_fields(DevParams P, Buffers B, RelaxArgs R, WaveArgs W) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= B.counters[C_NFLUID] || B.g_type[i] == 120 || R.in_zone[i] != 1) return;
  const double pi = 3.14159265358979324;
  const double soft = W.soft, H = W.H, T = W.T, L = W.L, A = W.A, k = W.k, d = W.d, g = W.g;
  const double skd = W.skd, ckd = W.ckd;
  const double2 x0 = R.init_pos[i];
  const double2 pc = R.an_pos[i];
  const double px = pc.x, py = pc.y;
  const double ang = k * px - W.om * W.t + W.phase;
  double el, u, v;
  double2 xn;
  if (W.wave_type == 1) {
    el = soft * (0.5 * H * cos(ang) + pi * H * H * (1.0 / tanh(k * d)) * (1 + 3.0 / (2 * (skd * skd * skd)) * cos(2 * ang)));
    u = pi * H / T * cosh(k * py) * cos(ang) / skd - 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * cosh(2 * k * py) * cos(2 * ang) / (skd * skd * skd * skd);
    v = pi * H / T * sinh(k * py) * sin(ang) / skd - 3.0 / 4.0 * (pi * pi) * H / T * (H / L) * sinh(2 * k * py) * sin(2 * ang) / (skd * skd * skd * skd);
    xn.x = x0.x - soft * (A * cosh(k * py) * sin(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * cosh(2 * k * py) * sin(2 * ang) / (skd * skd * skd * skd));
    xn.y = x0.y + soft * (A * sinh(k * py) * cos(ang) / skd + 3.0 * pi * (H * H) / (16 * L) * sinh(2 * k * py) * cos(2 * ang) / (skd * skd * skd * skd));
  } else {
    el = soft * 0.5 * H * cos(ang);
    u = pi * H / T * cosh(k * py) * cos(ang) / skd;
    v = pi * H / T * sinh(k * py) * sin(ang) / skd;
    xn.x = x0.x - soft * A * cosh(k * py) * sin(ang) / skd;
    xn.y = x0.y + soft * A * sinh(k * py) * cos(ang) / skd;
  }
  R.an_vel[i] = make_double2(soft * u, soft * v);
  R.an_pos[i] = xn;
  const double y = xn.y;
  if (y <= el + d) {
    double pr;
    if (W.wave_type == 1)
      pr = P.rho0 * g * (el + d - y) + soft * (P.rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd + 3.0 / 4.0 * P.rho0 * g * H * (pi * H / L) * (1.0 / sinh(2 * k * d)) * (cosh(2 * k * y) / (skd * skd) - (double)(1.0f / 3.0f)) * cos(2 * ang) - 1.0 / 4.0 * P.rho0 * g * H * (pi * H / L) * cosh(2 * k * y) / sinh(2 * k * d));
    else
      pr = P.rho0 * g * (el + d - y) + soft * (P.rho0 * g * H / 2.0 * cosh(k * y) * cos(ang) / ckd);
    R.an_p[i] = pr;
    R.an_rho[i] = inv_tait(P, pr);
  } else {
    R.an_p[i] = 0.0;
    R.an_rho[i] = P.rho0;
  }
}

// wave_maker_damping_zone.f90:61-95 on local slot s after the corrector: velocity,
// position and pressure blended toward the analytic fields, rho from the inverse Tait law.
__global__ void k_relax_blend(DevParams P, Buffers B, RelaxArgs R) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL]) return;
  const int g = B.name[s];
  if (g < 0 || g >= B.counters[C_NFLUID] || R.in_zone[g] != 1) return;
  const double f = R.coef[g];
  const double2 av = R.an_vel[g], ax = R.an_pos[g];
  double2 v = B.l_vel[s], x = B.l_pos[s];
  v.x = f * av.x + (1 - f) * v.x;
  v.y = f * av.y + (1 - f) * v.y;
  x.x = f * ax.x + (1 - f) * x.x;
  x.y = f * ax.y + (1 - f) * x.y;
  B.l_vel[s] = v;
  B.l_pos[s] = x;
  double2 pc = B.l_pc[s];
  pc.x = f * R.an_p[g] + (1 - f) * pc.x;
  B.l_pc[s] = pc;
  B.l_rv[s] = make_double2(inv_tait(P, pc.x), B.l_rv[s].y);
}

void launch_relax_init(hipStream_t s, const DevParams& P, Buffers& B, const RelaxArgs& R, int n_total,
                       double zone_length) {
  hipLaunchKernelGGL(k_relax_init, dim3(grid_for(n_total)), dim3(kBlock), 0, s, P, B, R, n_total, zone_length);
}

void launch_wave_theory(hipStream_t s, const DevParams& P, Buffers& B, const RelaxArgs& R, const WaveArgs& W,
                        int cap_fluid, int n_wm, int relax_on) {
  hipLaunchKernelGGL(k_wt_ill, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, s, B);
  if (!(W.t >= 0.0)) return;
  if (n_wm > 0) hipLaunchKernelGGL(k_wt_wavemaker, dim3(grid_for(n_wm)), dim3(kBlock), 0, s, P, B, R, W, n_wm);
  if (relax_on) hipLaunchKernelGGL(k_wt_relax_fields, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, s, P, B, R, W);
}

void launch_relax_blend(hipStream_t s, const DevParams& P, Buffers& B, const RelaxArgs& R, int cap_local) {
  hipLaunchKernelGGL(k_relax_blend, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, R);
}

// Assign_boundary_particles.f90:157-299 before an evaluation's passes: each fixed
// ghost takes the pressure and velocity of its non-ghost candidates (the 3x3 cells)
// within 3 hbar, hbar = 1.2 (h_i + h_j) / 2 with 1.2 a REAL(4) literal, by the
// order-1 MLS of MLS_Interpolation_In_Subdomain_Domain.f90 (Fai = (1 / sum W) W,
// with 3 or more particles, else nothing; with gravity each p_j is carried to the
// wall's height, p_j - g rho_j (y_i - y_j)); the velocity is mirrored by its boundary
// type (-6: v, -7: u and v, -8: u); then rho from the inverse Tait law (REAL(4)
// exponent), mass = rho V0 and c = c0 (rho/rho0)^3, or the rest state when p is 0.
__device__ __forceinline__ double wall_w(const DevParams& P, double r) {
  const double q = r / P.wall_h;
  return (exp(-(q * q)) - P.e9) * P.wall_kfactor;
}

// One wavefront per local slot (a wall's slot does work, the others leave at once):
// the lanes test the candidates of the 3x3 cells in parallel and compact the
// accepted ones into LDS in candidate order, with what the sums need; lane 0 then
// sums in that order, so the arithmetic is the sequential loop's. More than
// kWallList accepted candidates fall back to lane 0 walking the cells itself.
constexpr int kWallWaves = 4, kWallList = 128;

__device__ __forceinline__ void wall_finish(const DevParams& P, Buffers& B, int s, double u, double v, double pr) {
  const int bt = B.g_btype[B.name[s]];
  if (bt == -6) v = -v;
  else if (bt == -7) { u = -u; v = -v; }
  else if (bt == -8) u = -u;
  double rho, mass, c;
  if (pr != 0.0) {
    rho = P.rho0 * pow(1 + 7 * pr / (P.c0sq * P.rho0), (double)0.142857142f);
    mass = rho * P.vol0;
    const double x = rho / P.rho0;
    c = P.sqrt_c0sq * (x * x * x);
  } else {
    u = 0.0; v = 0.0; pr = 0.0;
    rho = P.rho0; mass = 0.0; c = P.sqrt_c0sq;
  }
  B.l_vel[s] = make_double2(u, v);
  B.l_pc[s] = make_double2(pr, c);
  B.l_rv[s] = make_double2(rho, mass / rho);
  B.l_mass[s] = mass;
  if (P.mixed) write_r0(P, B, s, make_double2(u, v), rho, mass, -3);
}

__global__ void __launch_bounds__(64 * kWallWaves) k_assign_walls(DevParams P, Buffers B) {
  __shared__ double sw[kWallWaves][kWallList], sp[kWallWaves][kWallList], sr[kWallWaves][kWallList];
  __shared__ double sy[kWallWaves][kWallList], su[kWallWaves][kWallList], sv[kWallWaves][kWallList];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int s = blockIdx.x * kWallWaves + wv;
  if (s >= B.counters[C_NLOCAL] || B.l_type[s] != -3) return;   // uniform across the wavefront
  const double2 xs = B.l_pos[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  const double rmax = 3 * P.wall_h, rmin = (double)1.0E-7f;
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
        if (r <= rmax && r >= rmin) { ok = true; w = wall_w(P, r); }
      }
      const uint64_t m = __ballot(ok);
      const int k = cnt + __popcll(m & below);
      if (ok && k < kWallList) {
        sw[wv][k] = w;
        sp[wv][k] = B.l_pc[o].x;
        sr[wv][k] = B.l_rv[o].x;
        sy[wv][k] = B.l_pos[o].y;
        const double2 vo = B.l_vel[o];
        su[wv][k] = vo.x;
        sv[wv][k] = vo.y;
      }
      cnt += __popcll(m);
    }
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  if (lane != 0) return;
  double u = 0.0, v = 0.0, pr = 0.0;
  if (cnt <= kWallList) {
    double a = 0.0;
    for (int k = 0; k < cnt; ++k) a = a + sw[wv][k] * 1.0;
    if (cnt >= 3) {
      const double ainv = 1.0 / a;
      for (int k = 0; k < cnt; ++k) {
        const double fai = 1.0 * (ainv * sw[wv][k]);
        if (P.g != 0.0) pr = pr + sp[wv][k] * fai - P.g * sr[wv][k] * (xs.y - sy[wv][k]) * fai;
        else pr = pr + sp[wv][k] * fai;
        u = u + su[wv][k] * fai;
        v = v + sv[wv][k] * fai;
      }
    }
  } else {                                               // a crowded support: the cells directly
    double a = 0.0;
    int nt = 0;
    for (int d = 0; d < 3; ++d)
      for (int o = R.lo[d]; o < R.hi[d]; ++o) {
        if (B.l_type[o] == -3) continue;
        const double2 xo = B.l_pos[o];
        const double px = xs.x - xo.x, py = xs.y - xo.y;
        const double r = sqrt(px * px + py * py);
        if (r <= rmax && r >= rmin) { a = a + wall_w(P, r) * 1.0; ++nt; }
      }
    if (nt >= 3) {
      const double ainv = 1.0 / a;
      for (int d = 0; d < 3; ++d)
        for (int o = R.lo[d]; o < R.hi[d]; ++o) {
          if (B.l_type[o] == -3) continue;
          const double2 xo = B.l_pos[o];
          const double px = xs.x - xo.x, py = xs.y - xo.y;
          const double r = sqrt(px * px + py * py);
          if (!(r <= rmax && r >= rmin)) continue;
          const double fai = 1.0 * (ainv * wall_w(P, r));
          const double2 vo = B.l_vel[o];
          if (P.g != 0.0) pr = pr + B.l_pc[o].x * fai - P.g * B.l_rv[o].x * (xs.y - xo.y) * fai;
          else pr = pr + B.l_pc[o].x * fai;
          u = u + vo.x * fai;
          v = v + vo.y * fai;
        }
    }
  }
  wall_finish(P, B, s, u, v, pr);
}

void launch_assign_walls(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_assign_walls, dim3((cap_local + kWallWaves - 1) / kWallWaves), dim3(64 * kWallWaves), 0, s,
                     P, B);
}

// Refresh_domain.f90:136-300 for the fixed ghosts: their assigned state goes back to
// the global arrays (the next gather reads it; particle_mass at a local index reads
// it in the SPS term).
__global__ void k_wall_refresh(DevParams P, Buffers B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= B.counters[C_NLOCAL] || B.l_type[s] != -3) return;
  const int g = B.name[s];
  B.g_vel[g] = B.l_vel[s];
  const double2 pc = B.l_pc[s];
  B.g_p[g] = pc.x;
  B.g_c[g] = pc.y;
  B.g_rho[g] = B.l_rv[s].x;
  B.g_mass[g] = B.l_mass[s];
}

void launch_wall_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_wall_refresh, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B);
}

static int wall_run(double x0, double y0, double x1, double y1, double sx, double sy, double step, double gdx,
                    int& k, bool first, double* x, double* y, int at) {
  if (first) {
    int j = 0;
    for (;;) {
      ++j;
      const double px = x0 + sx * ((j - 1) * step), py = y0 + sy * ((j - 1) * step);
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
  k = k + 2;
  return n;
}

// Layer i's runs: top from (Lx + o, Ly + o) leftwards, bottom from (-o, -o) rightwards,
// right from (Lx + o, -o) upwards, left from (-o, Ly + o) downwards, o = (i - 1) dy + dy/2,
// each without its end point; the first layer's count walks until within dx/2 of the end.
int wall_layout(double bx, double by, double gdx, double gdy, int layers, double* x, double* y, int* btype) {
  int n = 0, k = 0;
  for (int i = 1; i <= layers; ++i)
    n += wall_run(bx + (i - 1) * gdy + gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0,
                  by + (i - 1) * gdy + gdy / 2.0, -1, 0, gdx, gdx, k, i == 1, x, y, n);
  for (int i = 1; i <= layers; ++i)
    n += wall_run(-(i - 1) * gdy - gdy / 2.0, -(i - 1) * gdy - gdy / 2.0, bx + (i - 1) * gdy + gdy / 2.0,
                  -(i - 1) * gdy - gdy / 2.0, 1, 0, gdx, gdx, k, i == 1, x, y, n);
  for (int i = 1; i <= layers; ++i)
    n += wall_run(bx + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0, bx + (i - 1) * gdy + gdy / 2.0,
                  by + (i - 1) * gdy + gdy / 2.0, 0, 1, gdx, gdx, k, i == 1, x, y, n);
  for (int i = 1; i <= layers; ++i)
    n += wall_run(-(i - 1) * gdy - gdy / 2.0, by + (i - 1) * gdy + gdy / 2.0, -(i - 1) * gdy - gdy / 2.0,
                  -(i - 1) * gdy - gdy / 2.0, 0, -1, gdy, gdx, k, i == 1, x, y, n);
  if (x && btype)
    for (int q = 0; q < n; ++q) {
      const bool corner = x[q] < 0.0 || x[q] > bx;
      if (y[q] <= 0.0 || y[q] >= by) btype[q] = corner ? -7 : -6;
      else btype[q] = -8;
    }
  return n;
}

void launch_end_damping(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_end_damping, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B);
}

void launch_wave_probes(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, const double* probe_x,
                        int n_probes, unsigned long long* out_bits) {
  hipLaunchKernelGGL(k_wave_probes, dim3(grid_for(cap_fluid)), dim3(kBlock), 0, s, P, B, probe_x, n_probes,
                     out_bits);
}

// The density gradient of pass A: FP64 array, or the FP32-mixed record.
__device__ __forceinline__ double2 grad_rho(const DevParams& P, const Buffers& B, int s) {
  if (P.mixed) {
    const float4 r2 = B.m_r2[s];
    return make_double2((double)r2.x, (double)r2.y);
  }
  return B.grad[s];
}

// ---------------------------------------------------------------------------
// Shift_particle_position.f90:108-136 (after an optional MLS reinitialisation:
// mode 1 = take rho from mls_rho first, 2 = only commit mls_rho, 0 = shift only)
// ---------------------------------------------------------------------------
__global__ void k_shift_apply(DevParams P, Buffers B, int mode) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  double rho = B.l_rv[s].x;
  if (mode >= 1) {
    rho = B.mls_rho[s];
    B.l_rv[s] = make_double2(rho, B.l_rv[s].y);
  }
  if (mode == 2) return;
  if (B.l_type[s] != 2 || B.near_[s] == 1) return;
  double2 sh = B.shift[s];
  const double amp = sqrt(sh.x * sh.x + sh.y * sh.y);
  if (amp > P.max_shift) {
    sh.x = sh.x / amp * P.max_shift;
    sh.y = sh.y / amp * P.max_shift;
  }
  double2 x = B.l_pos[s];
  x.x = x.x + sh.x;
  x.y = x.y + sh.y;
  B.l_pos[s] = x;
  const double2 gr = grad_rho(P, B, s);
  if (P.ms) {
    rho = rho + (gr.x * sh.x + gr.y * sh.y);
  } else {
    rho = rho + gr.x * sh.x;
    rho = rho + gr.y * sh.y;
  }
  B.l_rv[s] = make_double2(rho, B.l_rv[s].y);
  B.l_pc[s] = make_double2(tait_p(P, rho), B.l_pc[s].y);
}

void launch_shift_apply(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  (void)P;
  hipLaunchKernelGGL(k_shift_apply, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, 0);
}

void launch_shift_apply_mode(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int mode) {
  hipLaunchKernelGGL(k_shift_apply, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, mode);
}

// Refresh_domain.f90:136-300 on one rank rebuilds the global arrays from the subdomain:
// a fluid particle left out of it (typed 120 by this step's binning) becomes a zero
// entry of type 0, which the next binning skips.
__device__ __forceinline__ void refresh_dropped(Buffers& B, int g) {
  B.g_pos[g] = make_double2(0.0, 0.0);
  B.g_vel[g] = make_double2(0.0, 0.0);
  B.g_rho[g] = 0.0; B.g_p[g] = 0.0; B.g_c[g] = 0.0; B.g_mass[g] = 0.0; B.g_e[g] = 0.0;
  B.g_type[g] = 0;
}

// Refresh_domain.f90:136-300 on one rank.
// Gather form: one thread per fluid gid reads its local slot (l_of_g), so the
// global arrays are written in order. Ghost gids are not written back: the
// next step regenerates them before anything reads them. Energy stays in the
// global arrays only: it is never updated (its rate is zeroed and never set).
__global__ void k_refresh(Buffers B) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= B.counters[C_NFLUID]) return;
  const int s = B.l_of_g[g];
  if (s < 0) { refresh_dropped(B, g); return; }
  B.g_pos[g] = B.l_pos[s];
  B.g_vel[g] = B.l_vel[s];
  const double2 rv = B.l_rv[s];
  const double2 pc = B.l_pc[s];
  B.g_rho[g] = rv.x;
  B.g_p[g] = pc.x;
  B.g_c[g] = pc.y;
  // mass, type and energy do not change inside a step (dedt = 0: single_step_compute.f90:113-126)
}

void launch_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  (void)P;
  hipLaunchKernelGGL(k_refresh, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, B);
}

// The end of a step in one pass: MLS density commit (mode >= 1), particle
// shifting (mode <= 1; Shift_particle_position.f90:105-136) and the refresh,
// computed for each fluid gid from its local slot. Same arithmetic as
// k_shift_apply followed by k_refresh; the local arrays are not updated (the
// next step rebuilds them from the global ones).
// Shift_particle_position.f90:108-136 for local slot s of a fluid particle that is
// not near the surface: the capped shift moves x, corrects rho by grad(rho).shift
// and p by the EOS.
__device__ __forceinline__ void shift_one(const DevParams& P, const Buffers& B, int s, double2& x, double& rho,
                                          double& p) {
  double2 sh = B.shift[s];
  const double amp = sqrt(sh.x * sh.x + sh.y * sh.y);
  if (amp > P.max_shift) {
    sh.x = sh.x / amp * P.max_shift;
    sh.y = sh.y / amp * P.max_shift;
  }
  x.x = x.x + sh.x;
  x.y = x.y + sh.y;
  const double2 gr = grad_rho(P, B, s);
  if (P.ms) {                      // MS Shift_particle_position.f90:117 (DOT_PRODUCT)
    rho = rho + (gr.x * sh.x + gr.y * sh.y);
  } else {
    rho = rho + gr.x * sh.x;
    rho = rho + gr.y * sh.y;
  }
  p = tait_p(P, rho);
}

__device__ __forceinline__ void k_shift_refresh_body(const DevParams& P, Buffers& B, int mode, int shift_on, int blk_, int nblk_) {
  const int g = blk_ * blockDim.x + threadIdx.x;
  if (g >= B.counters[C_NFLUID]) return;
  const int s = B.l_of_g[g];
  if (s < 0) { refresh_dropped(B, g); return; }
  double2 x = B.l_pos[s];
  const double2 rv = B.l_rv[s];
  double2 pc = B.l_pc[s];
  double rho = rv.x;
  if (mode >= 1) rho = B.mls_rho[s];
  if (shift_on && mode <= 1 && B.l_type[s] == 2 && B.near_[s] != 1) shift_one(P, B, s, x, rho, pc.x);
  B.g_pos[g] = x;
  B.g_vel[g] = B.l_vel[s];
  B.g_rho[g] = rho;
  B.g_p[g] = pc.x;
  B.g_c[g] = pc.y;
}

__global__ void k_shift_refresh(DevParams P, Buffers B, int mode, int shift_on) {
  k_shift_refresh_body(P, B, mode, shift_on, blockIdx.x, gridDim.x);
}

void launch_shift_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int mode, int shift_on) {
  hipLaunchKernelGGL(k_shift_refresh, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, mode, shift_on);
}

// The end of a predictor-corrector step without MLS in one pass: the corrector
// (time_iteration.f90:590-599), shifting and the refresh, for each fluid gid from
// its local slot. Same arithmetic as k_stage(ST_PC_CORRECT) followed by
// k_shift_refresh(mode 0); the local arrays keep the predicted state. The
// predictor's velocity is still in l_vel (the evaluation does not write it), and
// the global arrays of fluid gids still hold the step-start state.
__device__ __forceinline__ void k_correct_refresh_body(const DevParams& P, Buffers& B, int shift_on, int blk_, int nblk_) {
  const int g = blk_ * blockDim.x + threadIdx.x;
  if (g >= B.counters[C_NFLUID]) return;
  const int s = B.l_of_g[g];
  if (s < 0) { refresh_dropped(B, g); return; }
  const int t = B.l_type[s];
  // The step-start state is still this gid's global state (before_* of the gather).
  const double2 bx = B.g_pos[g], bv = B.g_vel[g];
  const double brho = B.g_rho[g];
  // The predicted state (velocity, and p, c of the evaluation) again, bit for bit as
  // the gather computed it; XSPH's average velocity of the previous evaluation is gone
  // by now, so with epsilon != 0 it is read back instead.
  double2 v, x;
  double rho;
  double2 pc;
  if (P.epsilon_xsph == 0.0) {
    double rp = brho;
    double2 xp = bx;
    v = bv;
    if (t == 2) pc_predict(P.dt, B.t_drho[s], B.t_acc[s], make_double2(0.0, 0.0), brho, bv, bx, rp, v, xp);
    tait(P, rp, pc.x, pc.y);
  } else {
    v = B.l_vel[s];
    pc = B.l_pc[s];
  }
  if (t == 2) {
    const double2 avg = (P.epsilon_xsph != 0.0) ? B.avg[s] : make_double2(0.0, 0.0);
    pc_correct(P.dt, B.drhodt[s], B.acc[s], avg, B.t_drho[s], B.t_acc[s], v, brho, bv, bx, rho, v, x);
    rebin_check(P, B, B.l_cell[s], x);
    if (shift_on && B.near_[s] != 1) shift_one(P, B, s, x, rho, pc.x);
  } else {
    x = bx;                                  // the stages move fluid particles only
    rho = brho;
  }
  B.g_pos[g] = x;
  B.g_vel[g] = v;
  B.g_rho[g] = rho;
  B.g_p[g] = pc.x;
  B.g_c[g] = pc.y;
}

__global__ void k_correct_refresh(DevParams P, Buffers B, int shift_on) {
  k_correct_refresh_body(P, B, shift_on, blockIdx.x, gridDim.x);
}

void launch_correct_refresh(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int shift_on) {
  hipLaunchKernelGGL(k_correct_refresh, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, shift_on);
}

// ---------------------------------------------------------------------------
// Sampling_subroutine.f90:340-461: block partials in fixed order, then one block.
// partial layout per block: [ke, vmax, (max, sum, sum2) x 4]
// ---------------------------------------------------------------------------
constexpr int kRed = 14;

__global__ void k_sample_partial(DevParams P, Buffers B, int cap_fluid) {
  __shared__ double lds[kBlock][kRed];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v[kRed];
#pragma unroll
  for (int k = 0; k < kRed; ++k) v[k] = 0.0;
  if (i < cap_fluid && i < B.counters[C_NFLUID]) {
    const double2 u = B.g_vel[i];
    const double vv = u.x * u.x + u.y * u.y;
    v[0] = 0.5 * B.g_mass[i] * vv;
    const double vm = sqrt(vv);
    v[1] = vm;
    const double p0 = P.rho0 * P.U * P.U;
    const double e[4] = {(u.x - B.ana[4 * i]) / P.U, (u.y - B.ana[4 * i + 1]) / P.U,
                         (B.g_p[i] - B.ana[4 * i + 2]) / p0, (vm - B.ana[4 * i + 3]) / P.U};
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const double ea = fabs(e[a]);
      v[2 + 3 * a] = ea; v[3 + 3 * a] = ea; v[4 + 3 * a] = ea * ea;
    }
  }
#pragma unroll
  for (int k = 0; k < kRed; ++k) lds[threadIdx.x][k] = v[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
#pragma unroll
      for (int k = 0; k < kRed; ++k) {
        const bool mx = (k == 1) || (k >= 2 && (k - 2) % 3 == 0);
        const double a = lds[threadIdx.x][k], b = lds[threadIdx.x + off][k];
        lds[threadIdx.x][k] = mx ? (a > b ? a : b) : a + b;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < kRed; ++k) B.red[blockIdx.x * kRed + k] = lds[0][k];
}

// One block: thread t folds partial blocks t, t+256, ... in order, then a fixed tree.
__global__ void k_sample_final(Buffers B, int nb, double* out) {
  __shared__ double lds[kBlock][kRed];
  double v[kRed];
#pragma unroll
  for (int k = 0; k < kRed; ++k) v[k] = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock)
#pragma unroll
    for (int k = 0; k < kRed; ++k) {
      const bool mx = (k == 1) || (k >= 2 && (k - 2) % 3 == 0);
      const double x = B.red[b * kRed + k];
      v[k] = mx ? (v[k] > x ? v[k] : x) : v[k] + x;
    }
#pragma unroll
  for (int k = 0; k < kRed; ++k) lds[threadIdx.x][k] = v[k];
  __syncthreads();
  for (int off = kBlock / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) {
#pragma unroll
      for (int k = 0; k < kRed; ++k) {
        const bool mx = (k == 1) || (k >= 2 && (k - 2) % 3 == 0);
        const double a = lds[threadIdx.x][k], c = lds[threadIdx.x + off][k];
        lds[threadIdx.x][k] = mx ? (a > c ? a : c) : a + c;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < kRed; ++k) out[k] = lds[0][k];
}

void launch_sample(hipStream_t s, const DevParams& P, Buffers& B, int cap_fluid, double* dev_out14) {
  const int nb = grid_for(cap_fluid);
  hipLaunchKernelGGL(k_sample_partial, dim3(nb), dim3(kBlock), 0, s, P, B, cap_fluid);
  hipLaunchKernelGGL(k_sample_final, dim3(1), dim3(kBlock), 0, s, B, nb, dev_out14);
}

// ---------------------------------------------------------------------------
// Pair set export (find_particle_pair.f90:64-123): pairs with a < b (local).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool is_pair(const DevParams& P, const Buffers& B, int s, int o,
                                        double2 ps, int ts) {
  const double2 po = B.l_pos[o];
  const double pdx = ps.x - po.x, pdy = ps.y - po.y;
  const double r = sqrt(pdx * pdx + pdy * pdy);
  if (!(r <= P.three_h && r >= P.r_min)) return false;
  return ts == 2 || B.l_type[o] == 2;
}

__global__ void k_pair_count(DevParams P, Buffers B) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const int ts = B.l_type[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  int cnt = 0;
  for (int d = 0; d < 3; ++d)
    for (int o = R.lo[d]; o < R.hi[d]; ++o)
      if (o > s && is_pair(P, B, s, o, ps, ts)) ++cnt;
  B.pair_count[s] = cnt;
}

__global__ void k_pair_write(DevParams P, Buffers B, int64_t capacity) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const int ts = B.l_type[s];
  const CellRanges R = cell_ranges(P, B.cell_start, B.l_cell[s]);
  int64_t k = B.pair_off[s];
  const int gs = B.name[s];
  for (int d = 0; d < 3; ++d)
    for (int o = R.lo[d]; o < R.hi[d]; ++o)
      if (o > s && is_pair(P, B, s, o, ps, ts)) {
        if (k < capacity) {
          const int go = B.name[o];
          B.pair_lo[k] = gs < go ? gs : go;
          B.pair_hi[k] = gs < go ? go : gs;
        }
        ++k;
      }
}

// find_particle_pair.f90:64-123 in the reference's own order: particle i's
// candidates are the 3x3 stencil of Distribute_Backgroud_Grid.f90:102-112 (self,
// left, right, upper left, up, upper right, lower left, down, lower right), each
// cell's particles in local order, at most 300 per cell and 300 in all
// (Assign_boundary_particles.f90:124-152); the pairs are the candidates j > i that
// pass the distance test. WRITE = false counts, true writes 1-based (i, j).
__constant__ int kStencilDx[9] = {0, -1, 1, -1, 0, 1, -1, 0, 1};
__constant__ int kStencilDy[9] = {0, 0, 0, 1, 1, 1, -1, -1, -1};

template <bool WRITE>
__global__ void k_ref_pairs(DevParams P, Buffers B, int64_t capacity) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  const int n = B.counters[C_NLOCAL];
  if (s >= n) return;
  const double2 ps = B.l_pos[s];
  const int ts = B.l_type[s];
  const int cell = B.l_cell[s];
  const int cx = cell / P.chain_ny, cy = cell - cx * P.chain_ny;
  int cnt = 0, cand = 0;
  int64_t k = WRITE ? (int64_t)B.pair_off[s] : 0;
  for (int d = 0; d < 9 && cand < 300; ++d) {
    const int x = cx + kStencilDx[d], y = cy + kStencilDy[d];
    if (x < 0 || x >= P.chain_nx || y < 0 || y >= P.chain_ny) continue;
    const int key = x * P.chain_ny + y;
    const int lo = B.cell_start[key];
    const int hi = min(B.cell_start[key + 1], lo + 300);
    for (int o = lo; o < hi && cand < 300; ++o, ++cand) {
      if (o > s && is_pair(P, B, s, o, ps, ts)) {
        if (WRITE && k < capacity) { B.pair_lo[k] = s + 1; B.pair_hi[k] = o + 1; }
        ++k;
        ++cnt;
      }
    }
  }
  if (!WRITE) B.pair_count[s] = cnt;
}

void launch_count_pair_list(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_ref_pairs<false>, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, (int64_t)0);
  launch_fill_i32(s, B.counters + C_NPAIR_LO, 1, 0);
  exclusive_scan(s, B.pair_count, cap_local, B.pair_off, B.scan_blocks, B.counters + C_NPAIR_LO);
}

void launch_pair_list(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int64_t capacity) {
  hipLaunchKernelGGL(k_ref_pairs<true>, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, capacity);
}

void launch_count_pairs(hipStream_t s, const DevParams& P, Buffers& B, int cap_local) {
  hipLaunchKernelGGL(k_pair_count, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B);
  launch_fill_i32(s, B.counters + C_NPAIR_LO, 1, 0);
  exclusive_scan(s, B.pair_count, cap_local, B.pair_off, B.scan_blocks, B.counters + C_NPAIR_LO);
}

void launch_pairs(hipStream_t s, const DevParams& P, Buffers& B, int cap_local, int64_t capacity,
                  int64_t* n_pairs) {
  (void)n_pairs;
  hipLaunchKernelGGL(k_pair_write, dim3(grid_for(cap_local)), dim3(kBlock), 0, s, P, B, capacity);
}

// ---------------------------------------------------------------------------
// array forms of compute_kernel / status_equation
// ---------------------------------------------------------------------------
__global__ void k_kernel_n(int n, const double* r, const double* dx, const double* h, double* w,
                           double* dwdx, double e9, float c4, double pik) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double hb = h[i], dist = r[i];
  const double q = dist / hb;
  const double factor = 1.0 / (pik * (hb * hb) * (double)c4);
  double wi = 0.0, d0 = 0.0, d1 = 0.0;
  if (dist <= 3.0 * hb) {
    const double eq = exp(-(q * q));
    wi = (eq - e9) * factor;
    const double t = eq * (-2.0 * q);
    d0 = t * dx[i] / (dist * hb) * factor;
    d1 = t * dx[n + i] / (dist * hb) * factor;
  }
  w[i] = wi;
  dwdx[i] = d0;
  dwdx[n + i] = d1;
}

void launch_kernel_n(hipStream_t s, int n, const double* r, const double* dx, const double* h,
                     double* w, double* dwdx, double e9) {
  const float c4 = 1.0f - 10.0f * expf(-9.0f);
  const double pik = (double)3.14159265358f;
  hipLaunchKernelGGL(k_kernel_n, dim3(grid_for(n)), dim3(kBlock), 0, s, n, r, dx, h, w, dwdx, e9, c4, pik);
}

__global__ void k_eos_n(int n, double rho0, double c0sq, const double* rho, const int* type,
                        double* p, double* c) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int t = type[i];
  if (t == 120) return;
  if (t == 1 || t == -1) {                          // air branch: c only (status_equation.f90:46-50)
    c[i] = sqrt(((double)1.4f - 1.0) * rho[i]);
    return;
  }
  const double b = c0sq * rho0 / 7.0;
  const double x = rho[i] / rho0;
  p[i] = b * (x * x * x * x * x * x * x - 1.0);
  c[i] = sqrt(c0sq) * (x * x * x);
}

void launch_eos_n(hipStream_t s, int n, double rho0, double c0sq, const double* rho,
                  const int* type, double* p, double* c) {
  hipLaunchKernelGGL(k_eos_n, dim3(grid_for(n)), dim3(kBlock), 0, s, n, rho0, c0sq, rho, type, p, c);
}

// ---------------------------------------------------------------------------
// Batched forms of the step's kernels: block b of a launch belongs to the instance
// found by batch_where; each runs the same body with that instance's parameters.
// ---------------------------------------------------------------------------
int batch_blocks(int cls, const Inst& I, int shape) {
  switch (cls) {
    case BC_FLUID: return grid_for(I.cap_fluid);
    case BC_TWO: return 2;
    case BC_ONE: return 1;
    case BC_FILL: return grid_for(std::max(std::max(I.cap_ghost, I.n_cells), 1));
    case BC_SCAN_G: return scan_block_count(I.cap_ghost);
    case BC_GHOST: return grid_for(I.cap_ghost);
    case BC_GLOBAL: return grid_for(I.cap_global);
    case BC_SCAN_C: return scan_block_count(I.n_cells);
    case BC_RANK: return (I.n_cells + kBlock / 64 - 1) / (kBlock / 64);
    case BC_LOCAL: return grid_for(I.cap_local);
    case BC_NBR: return (I.cap_local + kNbShape[shape].block - 1) / kNbShape[shape].block;
    default: return 0;
  }
}

// the zeroing of one step: corner-ghost counts, cell populations, ill and surface counters
__global__ void kb_step_fill(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const Inst& I = A.inst[w.i];
  const int k = w.blk * blockDim.x + threadIdx.x;
  if (k < I.cap_ghost) I.B.cnt_yx[k] = 0;
  if (k < I.n_cells) I.B.cell_count[k] = 0;
  if (k == 0) {
    I.B.counters[C_NILL] = 0;
    I.B.counters[C_NSURF] = 0;
    I.B.counters[C_MLSFB] = 0;
  }
}

__global__ void __launch_bounds__(kBlock) kb_wrap_count(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_wrap_count_body(P, B, A.inst[w.i].cap_fluid, w.blk, w.nblk);
}

__global__ void __launch_bounds__(1024) kb_scan_top2(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const Inst& I = A.inst[w.i];
  k_scan_top2_body(I.B.gblk, (I.cap_fluid + kBlock - 1) / kBlock, I.B.counters + C_NXG, I.B.counters + C_NYF, w.blk, w.nblk);
}

__global__ void __launch_bounds__(kBlock) kb_ghost_emit(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_ghost_emit_body(P, B, A.inst[w.i].cap_fluid, A.inst[w.i].cap_ghost, w.blk, w.nblk);
}

// which = 0: corner-ghost counts (cnt_yx -> off_yx, total C_NYX); 1: cell populations
// (cell_count -> cell_start, total C_NLOCAL and cell_start[n_cells]), as in the single path.
struct ScanArgs {
  const int* in;
  int n;
  int* out;
  int* blocks;
  int* total;
  int* total2;
};
__device__ __forceinline__ ScanArgs scan_args(const Inst& I, int which) {
  if (which == 0) return ScanArgs{I.B.cnt_yx, max(I.cap_ghost, 1), I.B.off_yx, I.B.scan_blocks, I.B.counters + C_NYX,
                                  nullptr};
  return ScanArgs{I.B.cell_count, max(I.n_cells, 1), I.B.cell_start, I.B.scan_blocks, I.B.counters + C_NLOCAL,
                  I.B.cell_start + I.n_cells};
}

__global__ void kb_scan_reduce(BatchArgs A, int which) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const ScanArgs a = scan_args(A.inst[w.i], which);
  k_scan_reduce_body(a.in, a.n, a.blocks, w.blk, w.nblk);
}

__global__ void kb_scan_top(BatchArgs A, int which) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const ScanArgs a = scan_args(A.inst[w.i], which);
  k_scan_top_body(a.blocks, (a.n + kScanTile - 1) / kScanTile, a.total, a.total2, 0, 1);
}

__global__ void kb_scan_apply(BatchArgs A, int which) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const ScanArgs a = scan_args(A.inst[w.i], which);
  k_scan_apply_body(a.in, a.n, a.blocks, a.out, w.blk, w.nblk);
}

__global__ void kb_ghost_yx(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_ghost_yx_body(P, B, A.inst[w.i].cap_ghost, w.blk, w.nblk);
}

__global__ void kb_cell_key(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_cell_key_body(P, B, A.inst[w.i].cap_global, w.blk, w.nblk);
}

__global__ void kb_cell_scatter(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  Buffers B = A.inst[w.i].B;
  k_cell_scatter_body(B, A.inst[w.i].cap_global, w.blk, w.nblk);
}

__global__ void __launch_bounds__(kBlock) kb_cell_rank(BatchArgs A) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  Buffers B = A.inst[w.i].B;
  k_cell_rank_body(B, A.inst[w.i].n_cells, w.blk, w.nblk);
}

__global__ void __launch_bounds__(kBlock) kb_gather_predict(BatchArgs A, int save_before, int lean) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_gather_predict_body(P, B, save_before, lean, w.blk, w.nblk);
}

__global__ void kb_stage(BatchArgs A, int kind) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_stage_body(P, B, kind, w.blk, w.nblk);
}

__global__ void __launch_bounds__(kBlock) kb_mls(BatchArgs A) {
  const BatchWhere w = batch_where(A, A.xcd ? xcd_block() : (int)blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_mls_body(P, B, w.blk, w.nblk);
}

__global__ void kb_shift_refresh(BatchArgs A, int mode, int shift_on) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_shift_refresh_body(P, B, mode, shift_on, w.blk, w.nblk);
}

__global__ void kb_correct_refresh(BatchArgs A, int shift_on) {
  const BatchWhere w = batch_where(A, blockIdx.x);
  const DevParams P = A.inst[w.i].P;
  Buffers B = A.inst[w.i].B;
  k_correct_refresh_body(P, B, shift_on, w.blk, w.nblk);
}

static BatchArgs bargs(const BatchDev& D, int cls) { return BatchArgs{D.inst, D.start[cls], D.n, D.xcd}; }

void launch_batch_step(hipStream_t s, const BatchDev& D, int mixed, int mls, int shift_on, int solver) {
  const int lean = mixed && !mls;
  auto go = [&](int cls, auto kernel, int threads, auto... args) {
    if (D.total[cls] > 0)
      hipLaunchKernelGGL(kernel, dim3(D.total[cls]), dim3(threads), 0, s, bargs(D, cls), args...);
  };
  go(BC_FILL, kb_step_fill, kBlock);
  // Distribute_Period_Boundary_Particle
  go(BC_FLUID, kb_wrap_count, kBlock);
  go(BC_TWO, kb_scan_top2, 1024);
  go(BC_FLUID, kb_ghost_emit, kBlock);
  go(BC_SCAN_G, kb_scan_reduce, kBlock, 0);
  go(BC_ONE, kb_scan_top, kBlock, 0);
  go(BC_SCAN_G, kb_scan_apply, kBlock, 0);
  go(BC_GHOST, kb_ghost_yx, kBlock);
  // Devide_block_and_send_tasks with the predictor
  go(BC_GLOBAL, kb_cell_key, kBlock);
  go(BC_SCAN_C, kb_scan_reduce, kBlock, 1);
  go(BC_ONE, kb_scan_top, kBlock, 1);
  go(BC_SCAN_C, kb_scan_apply, kBlock, 1);
  go(BC_GLOBAL, kb_cell_scatter, kBlock);
  go(BC_RANK, kb_cell_rank, kBlock);
  go(BC_GLOBAL, kb_gather_predict, kBlock, (int)mls, lean);
  // single_step_compute, then the corrector with shifting and the refresh
  launch_batch_eval(s, D, mixed, solver);
  if (!mls) {
    go(BC_LOCAL, kb_correct_refresh, kBlock, shift_on);
    return;
  }
  // an MLS step: corrector, Reinitialize_particle_density, then shifting (with the
  // MLS density) and the refresh
  go(BC_LOCAL, kb_stage, kBlock, (int)ST_PC_CORRECT);
  go(BC_LOCAL, kb_mls, kBlock);
  go(BC_LOCAL, kb_shift_refresh, kBlock, shift_on ? 1 : 2, shift_on);
}

}  // namespace sph2d
