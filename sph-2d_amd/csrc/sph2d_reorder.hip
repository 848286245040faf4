// sph2d_reorder.hip -- Reorder_Particle_Index (MS/Reorder_Particle_Index.f90:54-287)
// on device-resident arrays.
//
// The reference marks fluid particles outside the box as ill (:54-70), lists the
// particles of each kept type in index order (:84-115) and copies them back in the
// order fluid, wavemaker, fixed ghost, inlet, outlet (:157-287); every other type
// (ill, periodic, body, air) is dropped. That is a stable counting sort over five
// categories. Here: per-block category counts from wave ballots, one exclusive
// scan over the category-major (category, block) table -- which is exactly the
// stable partition's offsets -- a scatter of each kept particle's source index
// into `perm`, then one gather per column. Integer and copy work, HBM-bound.
#include "sph2d_internal.h"

#include <mutex>

namespace sph2d {

constexpr int kRoCats = 5;
constexpr int kRoWaves = kBlock / 64;

// Wave_Maker_Particle_Label 130, Fix_Ghost -3, Inlet -201, Outlet -202
// (MS/Information_Module.f90:61-67); 5 = dropped.
__device__ __forceinline__ int ro_category(int t) {
  switch (t) {
    case 2: return 0;
    case 130: return 1;
    case -3: return 2;
    case -201: return 3;
    case -202: return 4;
    default: return 5;
  }
}

// Category of particle i after the ill test (:56-64); writes the ill label back.
__global__ void __launch_bounds__(kBlock) k_ro_count(int n_total, int n_fluid, int* type, const double* x,
                                                     const double* y, double box_x, double box_y, double lo,
                                                     int* blk, int G) {
  __shared__ int wcnt[kRoWaves][kRoCats];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  int cat = 5;
  if (i < n_total) {
    int t = type[i];
    if (i < n_fluid && t == 2) {
      const double px = x[i], py = y[i];
      if (px <= lo || px > box_x || py <= lo || py > box_y) {
        t = 120;                                   // Ill_Particle_Label
        type[i] = t;
      }
    }
    cat = ro_category(t);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < kRoCats; ++c) {
    const unsigned long long m = __ballot(cat == c);
    if (lane == 0) wcnt[w][c] = __popcll(m);
  }
  __syncthreads();
  if (threadIdx.x < kRoCats) {
    int s = 0;
#pragma unroll
    for (int k = 0; k < kRoWaves; ++k) s += wcnt[k][threadIdx.x];
    blk[threadIdx.x * G + blockIdx.x] = s;
  }
}

// perm[new slot] = source index, in the order of the stable partition.
__global__ void __launch_bounds__(kBlock) k_ro_scatter(int n_total, const int* type, const int* off, int G,
                                                       int* perm) {
  __shared__ int wcnt[kRoWaves][kRoCats];
  const int i = blockIdx.x * kBlock + threadIdx.x;
  const int cat = i < n_total ? ro_category(type[i]) : 5;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const unsigned long long below = (1ull << lane) - 1ull;
  int rank = 0;
#pragma unroll
  for (int c = 0; c < kRoCats; ++c) {
    const unsigned long long m = __ballot(cat == c);
    if (lane == 0) wcnt[w][c] = __popcll(m);
    if (cat == c) rank = __popcll(m & below);
  }
  __syncthreads();
  if (cat < kRoCats) {
    for (int k = 0; k < w; ++k) rank += wcnt[k][cat];
    perm[off[cat * G + blockIdx.x] + rank] = i;
  }
}

// counts[c] = number kept in category c, counts[5] = total kept.
__global__ void k_ro_counts(const int* off, int G, const int* total, int* counts) {
  const int c = threadIdx.x;
  if (c < kRoCats) {
    const int a = off[c * G];
    const int b = c + 1 < kRoCats ? off[(c + 1) * G] : *total;
    counts[c] = b - a;
  }
  if (c == kRoCats) counts[kRoCats] = *total;
}

template <typename T>
__global__ void k_ro_gather(const int* perm, const int* total, const T* src, T* dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < *total) dst[k] = src[perm[k]];
}

void launch_reorder_partition(hipStream_t s, int n_total, int n_fluid, int* type, const double* x,
                              const double* y, double box_x, double box_y, int* blk, int* off,
                              int* scan_blocks, int* total, int* counts, int* perm) {
  const int G = (n_total + kBlock - 1) / kBlock;
  const double lo = (double)-1.0E-10f;            // REAL(4) literal (:58)
  hipLaunchKernelGGL(k_ro_count, dim3(G), dim3(kBlock), 0, s, n_total, n_fluid, type, x, y, box_x, box_y, lo,
                     blk, G);
  launch_exclusive_scan(s, blk, kRoCats * G, off, scan_blocks, total);
  hipLaunchKernelGGL(k_ro_scatter, dim3(G), dim3(kBlock), 0, s, n_total, type, off, G, perm);
  hipLaunchKernelGGL(k_ro_counts, dim3(1), dim3(64), 0, s, off, G, total, counts);
}

// ---- scratch of sph2d_reorder_particle_index_noalloc ----------------------------------
// One file-scope array per device (the HIP loader places a copy on every GPU the
// library runs on; it is held until the process ends) for inputs of at most
// kNoAllocMax particles: the category tables of k_ro_count/k_ro_scatter (5 per block),
// the scan tiles, the counters, the gather staging column and the permutation.
constexpr int kNoAllocMax = 1 << 21;
constexpr size_t kAlign = 256;
constexpr size_t align_up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr int kNoAllocBlocks = (kNoAllocMax + kBlock - 1) / kBlock;
constexpr size_t kNoAllocBytes = 2 * align_up(sizeof(int) * kRoCats * kNoAllocBlocks) +
                                 align_up(sizeof(int) * 1024) + align_up(sizeof(int) * 8) +
                                 align_up(sizeof(double) * kNoAllocMax) + align_up(sizeof(int) * kNoAllocMax);

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

int reorder_scratch(void** ptr, size_t* bytes, int* max_n) {
  constexpr int kMaxDevices = 64;
  static void* table[kMaxDevices] = {};
  static std::mutex lock;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices)
    return set_error(SPH2D_ERR_HIP, "hipGetDevice failed");
  {
    std::lock_guard<std::mutex> g(lock);
    if (!table[dev]) {
      void* p = nullptr;
      // host code holds only a stand-in for the array: its device address comes from
      // the runtime, per device, with that device current
      const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !p) return set_error(SPH2D_ERR_HIP, "hipGetSymbolAddress(scratch) failed");
      table[dev] = p;
    }
    *ptr = table[dev];
  }
  *bytes = kNoAllocBytes;
  *max_n = kNoAllocMax;
  return SPH2D_OK;
}

template <typename T>
__global__ void k_ro_copy(const int* total, const T* src, T* dst) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < *total) dst[k] = src[k];
}

void launch_reorder_copy_f64(hipStream_t s, int n_total, const int* total, const double* src, double* dst) {
  hipLaunchKernelGGL(k_ro_copy<double>, dim3((n_total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, total, src, dst);
}

void launch_reorder_copy_i32(hipStream_t s, int n_total, const int* total, const int* src, int* dst) {
  hipLaunchKernelGGL(k_ro_copy<int>, dim3((n_total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, total, src, dst);
}

void launch_reorder_gather_f64(hipStream_t s, int n_total, const int* perm, const int* total, const double* src,
                               double* dst) {
  hipLaunchKernelGGL(k_ro_gather<double>, dim3((n_total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, perm, total,
                     src, dst);
}

void launch_reorder_gather_i32(hipStream_t s, int n_total, const int* perm, const int* total, const int* src,
                               int* dst) {
  hipLaunchKernelGGL(k_ro_gather<int>, dim3((n_total + kBlock - 1) / kBlock), dim3(kBlock), 0, s, perm, total,
                     src, dst);
}

}  // namespace sph2d
