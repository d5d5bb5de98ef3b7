// Shared prepass of the fused forward and the adjoint: bounding box, 30-bit
// Morton keys, hipCUB radix sort of (key, index) and the packed FP64 record
// gather.  Kernels live in an anonymous namespace: each translation unit that
// includes this header gets its own copies.
#pragma once

#include <hipcub/hipcub.hpp>

#include "vpm_common.h"

namespace vpm {

namespace {

struct alignas(16) Src8 {
  double x, y, z, ax, ay, az, s, nu;
};

constexpr int kPreBlock = 256;
constexpr int kBboxBlocks = 512;

__global__ __launch_bounds__(kPreBlock) void bbox_partial_kernel(const double* __restrict__ pos,
                                                                 int64_t ld, int n,
                                                                 double* __restrict__ part) {
  __shared__ double red[6][kPreBlock / 64];
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = blockIdx.x * kPreBlock + threadIdx.x; i < n; i += gridDim.x * kPreBlock) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double v = pos[c * ld + i];
      mn[c] = fmin(mn[c], v);
      mx[c] = fmax(mx[c], v);
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mn[c] = wave_min(mn[c]);
    mx[c] = wave_max(mx[c]);
    if (lane == 0) {
      red[c][wave] = mn[c];
      red[3 + c][wave] = mx[c];
    }
  }
  __syncthreads();
  if (threadIdx.x < 6) {
    double v = red[threadIdx.x][0];
    for (int w = 1; w < kPreBlock / 64; ++w)
      v = threadIdx.x < 3 ? fmin(v, red[threadIdx.x][w]) : fmax(v, red[threadIdx.x][w]);
    part[blockIdx.x * 6 + threadIdx.x] = v;
  }
}

// One wave: each lane folds a strided share of the nb partial boxes, then
// the wave reduces (min/max are exact in any order).
__global__ __launch_bounds__(64) void bbox_final_kernel(const double* __restrict__ part, int nb,
                                                        double* __restrict__ bbox) {
  double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += 64) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      mn[c] = fmin(mn[c], part[b * 6 + c]);
      mx[c] = fmax(mx[c], part[b * 6 + 3 + c]);
    }
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    mn[c] = wave_min(mn[c]);
    mx[c] = wave_max(mx[c]);
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      bbox[c] = mn[c];
      bbox[3 + c] = mx[c];
    }
  }
}

__device__ inline uint32_t spread10(uint32_t v) {
  v &= 0x3ffu;
  v = (v | (v << 16)) & 0x030000FFu;
  v = (v | (v << 8)) & 0x0300F00Fu;
  v = (v | (v << 4)) & 0x030C30C3u;
  v = (v | (v << 2)) & 0x09249249u;
  return v;
}

__global__ __launch_bounds__(kPreBlock) void morton_kernel(const double* __restrict__ pos,
                                                           int64_t ld, int n,
                                                           const double* __restrict__ bbox,
                                                           uint32_t* __restrict__ keys,
                                                           int* __restrict__ idx) {
  const int i = blockIdx.x * kPreBlock + threadIdx.x;
  if (i >= n) return;
  uint32_t q[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double lo = bbox[c], ext = bbox[3 + c] - lo;
    const double f = ext > 0.0 ? (pos[c * ld + i] - lo) / ext : 0.0;
    int v = static_cast<int>(f * 1023.0);
    v = v < 0 ? 0 : (v > 1023 ? 1023 : v);
    q[c] = static_cast<uint32_t>(v);
  }
  keys[i] = spread10(q[0]) | (spread10(q[1]) << 1) | (spread10(q[2]) << 2);
  idx[i] = i;
}

__global__ __launch_bounds__(kPreBlock) void gather_kernel(const vpm_particles_t p, int n,
                                                           const int* __restrict__ perm,
                                                           Src8* __restrict__ out) {
  const int k = blockIdx.x * kPreBlock + threadIdx.x;
  if (k >= n) return;
  const int i = perm ? perm[k] : k;
  Src8 r;
  r.x = p.pos[i];
  r.y = p.pos[p.ld + i];
  r.z = p.pos[2 * p.ld + i];
  r.ax = p.str[i];
  r.ay = p.str[p.ld + i];
  r.az = p.str[2 * p.ld + i];
  r.s = p.sigma[i];
  r.nu = p.nu ? p.nu[i] : 0.0;
  out[k] = r;
}

__global__ void fill_kernel(double* __restrict__ p, int64_t n, double v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}


// Sorts particles by Morton key of their position.  Returns the permutation
// (sorted k -> caller index) in `perm`.  Scratch: bbox_part (kBboxBlocks*6),
// bbox (8), keys_in/keys_out/idx_in (n each) and cub_bytes of hipCUB storage.
inline int morton_sort(const vpm_particles_t* p, double* bbox_part, double* bbox, uint32_t* kin,
                       uint32_t* kout, int* iin, int* perm, void* cub, size_t cub_bytes,
                       hipStream_t st) {
  const int64_t n = p->n;
  const int ni = static_cast<int>(n);
  const int nbb = static_cast<int>(std::min<int64_t>(kBboxBlocks, ceil_div(n, kPreBlock)));
  hipLaunchKernelGGL(bbox_partial_kernel, dim3(nbb), dim3(kPreBlock), 0, st, p->pos, p->ld, ni,
                     bbox_part);
  VPM_LAUNCH_CHECK("bbox_partial_kernel");
  hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(64), 0, st, bbox_part, nbb, bbox);
  VPM_LAUNCH_CHECK("bbox_final_kernel");
  const int nb = static_cast<int>(ceil_div(n, kPreBlock));
  hipLaunchKernelGGL(morton_kernel, dim3(nb), dim3(kPreBlock), 0, st, p->pos, p->ld, ni, bbox,
                     kin, iin);
  VPM_LAUNCH_CHECK("morton_kernel");
  size_t cb = cub_bytes;
  VPM_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(cub, cb, kin, kout, iin, perm, ni, 0, 30, st));
  return VPM_OK;
}

inline int sort_temp_bytes(int64_t n, size_t* bytes) {
  *bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(
      nullptr, *bytes, static_cast<const uint32_t*>(nullptr), static_cast<uint32_t*>(nullptr),
      static_cast<const int*>(nullptr), static_cast<int*>(nullptr),
      static_cast<int>(n > 0 ? n : 1), 0, 30, hipStream_t(0));
  if (e != hipSuccess) {
    set_error("hipcub SortPairs size query failed: %s", hipGetErrorString(e));
    return VPM_ERR_LAUNCH;
  }
  return VPM_OK;
}

inline size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

}  // namespace vpm
