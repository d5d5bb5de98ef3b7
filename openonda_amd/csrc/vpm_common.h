// Internal helpers shared by the translation units of libvpm_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "vpm_hip.h"

namespace vpm {

constexpr double kOneOverFourPi = 0.079577471545947667884441881686257181;  // 1/(4 pi)
constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kEpsPair = 1e-8;  // ParticlesPhysicsCuda.py:6, ParticlesPhysics.py:4

// Thread-local error message (vpm_last_error()).
void set_error(const char* fmt, ...);
void clear_error();

inline int fail(int code, const char* what) {
  set_error("%s", what);
  return code;
}

#define VPM_HIP_TRY(expr)                                                       \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::vpm::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_),   \
                       __FILE__, __LINE__);                                     \
      return VPM_ERR_LAUNCH;                                                    \
    }                                                                           \
  } while (0)

// Check the launch that was just enqueued.
#define VPM_LAUNCH_CHECK(name)                                                  \
  do {                                                                          \
    hipError_t e_ = hipGetLastError();                                          \
    if (e_ != hipSuccess) {                                                     \
      ::vpm::set_error("launch of %s failed: %s", name, hipGetErrorString(e_)); \
      return VPM_ERR_LAUNCH;                                                    \
    }                                                                           \
  } while (0)

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

// Validates a particle set; `need_nu` requires the viscosity array.
int check_particles(const vpm_particles_t* p, bool need_nu, const char* who);
int check_points(const vpm_points_t* t, const char* who);
int check_out(const double* out, int64_t ld, int64_t n, const char* who);

inline int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

// ---- source splits -----------------------------------------------------
// When the target blocks of an all-pairs pass cannot fill the device, the
// sources are split among workgroups (grid.y) and FP64 partial sums per
// target are added in split order by a reduce kernel (deterministic).
constexpr int kMaxSplitBlocks = 4096;  // target blocks x splits

// Most workgroups in flight at once on the current device for `kern` (queried
// per call: kernels of one signature differ in register use).
template <typename Kern>
int resident_blocks(Kern kern, int block) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, block, 0) != hipSuccess ||
      per_cu <= 0)
    per_cu = 1;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      cus <= 0)
    cus = 1;
  return per_cu * cus;
}

// Splits actually used when `k` are asked for: whole tiles per split, none empty.
inline int effective_splits(int ntiles, int k) {
  const int per = (ntiles + k - 1) / k;
  return (ntiles + per - 1) / per;
}

// Number of source splits for `blocks` target blocks: equal-work workgroups
// run in rounds of `resident`, so take the smallest split count (<= kmax)
// whose launch fills its rounds best -- if that beats no split by 5%.
inline int choose_splits(int blocks, int ntiles, int kmax, int resident) {
  auto fill = [&](int k) {
    const int64_t b = static_cast<int64_t>(blocks) * effective_splits(ntiles, k);
    const int64_t rounds = (b + resident - 1) / resident;
    return static_cast<double>(b) / static_cast<double>(rounds * resident);
  };
  int best = 1;
  double best_fill = fill(1);
  for (int k = 2; k <= kmax; ++k) {
    const double f = fill(k);
    if (f > best_fill) {
      best = k;
      best_fill = f;
    }
  }
  return best_fill >= fill(1) + 0.05 ? effective_splits(ntiles, best) : 1;
}


// Largest split count launch code may pick for `blocks` target blocks over
// `ntiles` source tiles (1: never split).
inline int max_splits_for(int64_t blocks, int64_t ntiles) {
  if (blocks <= 0 || blocks >= kMaxSplitBlocks / 2) return 1;
  const int64_t k = ntiles < kMaxSplitBlocks / blocks ? ntiles : kMaxSplitBlocks / blocks;
  return k > 1 ? static_cast<int>(k) : 1;
}

// Wave64 butterfly reductions.
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ inline double wave_min(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmin(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline double wave_max(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

__device__ inline float wave_min(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fminf(v, __shfl_xor(v, off, 64));
  return v;
}
__device__ inline float wave_max(float v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v = fmaxf(v, __shfl_xor(v, off, 64));
  return v;
}

// Deterministic block sum (fixed shuffle tree + fixed LDS order).  All
// threads of the block must call it; the result is valid in thread 0.
template <int BLOCK>
__device__ inline double block_sum(double v, double* lds /* BLOCK/64 */) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  if (lane == 0) lds[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / 64; ++w) s += lds[w];
  }
  __syncthreads();
  return s;
}

// Sums `n` doubles in fixed order with one block and writes scale*sum.
__global__ void final_sum_kernel(const double* __restrict__ part, int64_t n, double scale,
                                 double* __restrict__ out);

}  // namespace vpm
