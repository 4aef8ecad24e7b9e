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

// Deterministic block sum whose result every thread receives.
template <class T>
__device__ __forceinline__ T block_sum_all(T v) {
  __shared__ T smem[kBlock / kWave + 1];
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (lane == 0) smem[wid] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    T s = T(0);
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) s += smem[w];
    smem[kBlock / kWave] = s;
  }
  __syncthreads();
  const T s = smem[kBlock / kWave];
  __syncthreads();
  return s;
}

// Sum of per-block partials written by an earlier kernel, in a fixed order;
// every thread receives the result.  Cross-block reductions go through a
// kernel boundary (which publishes the partials to every XCD) instead of an
// in-kernel device-scope fence and atomic: on CDNA4 with one L2 per XCD such a
// fence writes back the L2 and costs far more than re-reading a few KB of
// partials in each block.
template <class T>
__device__ __forceinline__ T sum_partials(const T *partials, int n) {
  T v = T(0);
  for (int i = threadIdx.x; i < n; i += blockDim.x) v += partials[i];
  return block_sum_all(v);
}

// Three partial sums at once, one barrier: every thread adds the four wave
// sums itself in the same order, so all threads of all blocks get the same
// bits.  Call at most once per kernel (its LDS is not re-armed).
__device__ __forceinline__ void sum_partials3(const double *__restrict__ a, int na,
                                              const double *__restrict__ b, int nb,
                                              const double *__restrict__ c, int nc, double &sa,
                                              double &sb, double &sc) {
  __shared__ double sm[3][kBlock / kWave];
  double va = 0.0, vb = 0.0, vc = 0.0;
  for (int i = threadIdx.x; i < na; i += blockDim.x) va += a[i];
  for (int i = threadIdx.x; i < nb; i += blockDim.x) vb += b[i];
  for (int i = threadIdx.x; i < nc; i += blockDim.x) vc += c[i];
  va = wave_sum(va);
  vb = wave_sum(vb);
  vc = wave_sum(vc);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (lane == 0) {
    sm[0][wid] = va;
    sm[1][wid] = vb;
    sm[2][wid] = vc;
  }
  __syncthreads();
  sa = sb = sc = 0.0;
#pragma unroll
  for (int w = 0; w < kBlock / kWave; ++w) {
    sa += sm[0][w];
    sb += sm[1][w];
    sc += sm[2][w];
  }
}

// Two block sums at once (one barrier); results valid in thread 0.
__device__ __forceinline__ void block_sum2(double &a, double &b) {
  __shared__ double sm[2][kBlock / kWave];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (lane == 0) {
    sm[0][wid] = a;
    sm[1][wid] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double x = 0.0, y = 0.0;
#pragma unroll
    for (int w = 0; w < kBlock / kWave; ++w) {
      x += sm[0][w];
      y += sm[1][w];
    }
    a = x;
    b = y;
  }
}

// out[0] = sum (subtract == 0) or out[0] -= sum of partials[0..n): one block.
template <class T>
__global__ __launch_bounds__(kBlock) void k_finalize_sum(const T *__restrict__ partials, int n,
                                                         T *__restrict__ out, int subtract) {
  const T s = sum_partials(partials, n);
  if (threadIdx.x == 0) out[0] = subtract ? out[0] - s : s;
}

}  // namespace tatva
