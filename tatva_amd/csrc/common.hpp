// Shared helpers: status/error plumbing, launch geometry, deterministic
// reductions.  Written for gfx950 (CDNA4): 64-wide wavefronts, 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/tatva_amd.h"

namespace tatva {

void set_error(const std::string &msg);

#define TATVA_TRY_HIP(expr)                                                     \
  do {                                                                          \
    hipError_t _e = (expr);                                                     \
    if (_e != hipSuccess) {                                                     \
      ::tatva::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));    \
      return TATVA_EHIP;                                                        \
    }                                                                           \
  } while (0)

#define TATVA_REQUIRE(cond, msg)                                                \
  do {                                                                          \
    if (!(cond)) {                                                              \
      ::tatva::set_error(msg);                                                  \
      return TATVA_EINVAL;                                                      \
    }                                                                           \
  } while (0)

#define TATVA_CHECK_LAUNCH()                                                    \
  do {                                                                          \
    hipError_t _e = hipGetLastError();                                          \
    if (_e != hipSuccess) {                                                     \
      ::tatva::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
      return TATVA_EHIP;                                                        \
    }                                                                           \
  } while (0)

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kNumXcd = 8;
// Upper bound on blocks of a grid-stride reduction kernel: bounds the partials
// buffers and fixes the summation tree for a given problem size.
constexpr int kMaxReduceBlocks = 2048;

inline int reduce_grid(int64_t n, int threads_per_item = 1) {
  int64_t items = n * threads_per_item;
  int64_t b = (items + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
  return (int)b;
}

inline int grid_for(int64_t n, int block = kBlock) {
  int64_t b = (n + block - 1) / block;
  return (int)(b < 1 ? 1 : b);
}

// Workgroups are dispatched round-robin over the 8 XCDs (block b -> XCD b % 8).
// Re-map so that each XCD walks one contiguous range of logical blocks: the
// neighbouring elements/rows it processes then share node data in its own L2.
__device__ __forceinline__ int64_t xcd_block(int b, int n) {
  const int xcd = b % kNumXcd;
  const int k = b / kNumXcd;
  const int q = n / kNumXcd;
  const int r = n % kNumXcd;
  return (int64_t)xcd * q + (xcd < r ? xcd : r) + k;
}

template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Deterministic block sum (fixed shuffle/LDS tree); result valid in thread 0.
template <class T>
__device__ __forceinline__ T block_sum(T v) {
  __shared__ T smem[kBlock / kWave];
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  T s = T(0);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += smem[w];
  }
  __syncthreads();
  return s;
}

// Returns true in every thread of the last block to arrive.  Callers write
// their block partials before calling; the last block then reads all partials
// in a fixed order, so the result does not depend on arrival order.
__device__ __forceinline__ bool last_block(unsigned int *counter) {
  __shared__ bool is_last;
  __threadfence();
  if (threadIdx.x == 0) {
    unsigned prev = atomicAdd(counter, 1u);
    is_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (is_last) __threadfence();
  return is_last;
}

// Sum of partials[0..n) by one block in a fixed order; result valid in thread 0.
template <class T>
__device__ __forceinline__ T block_sum_array(const T *partials, int n) {
  T v = T(0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[i];
  return block_sum(v);
}

}  // namespace tatva
