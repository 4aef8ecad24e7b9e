// VALU issue-rate calibration kernel.
//
// The FP32 tile kernels of the matrix-free operator are bound by VALU issue, so
// their roofline is a rate of wave64 VALU instructions.  This kernel measures
// that rate's ceiling on the device at hand with the same counter the tile
// kernels are profiled with (SQ_INSTS_VALU): every thread runs eight
// independent FP32 chains of one instruction form, 32 instructions per loop
// trip, and nothing else touches the vector ALU inside the loop (the trip count
// lives in scalar registers).  Per wave one launch issues 32 x iters of them
// plus a fixed prologue and epilogue of a few dozen instructions.

#include "common.hpp"

namespace tatva {

constexpr int kCalChains = 8;

// kOp 0: v = fma(v, a, b) with a, b uniform (VOP3 v_fma_f32, scalar operands);
// 1: v += a_t * b_t with per-thread a_t, b_t (VOP2 v_fmac_f32, vector operands);
// 2: v = v + b_t (VOP2 v_add_f32); 3: v = v * a_t (VOP2 v_mul_f32);
// 4: v = fma(v, a_t, b_t) (VOP3 v_fma_f32, vector operands only);
// 5: v += a * b_t with a uniform (VOP2 v_fmac_f32, one scalar operand).
template <int kOp>
__global__ __launch_bounds__(kBlock) void k_valu_chains(int32_t iters, float a, float b,
                                                        float *__restrict__ out) {
  const float at = a * (1.0f + 1e-7f * threadIdx.x);
  const float bt = b * (1.0f - 1e-7f * threadIdx.x);
  float v[kCalChains];
#pragma unroll
  for (int k = 0; k < kCalChains; ++k) v[k] = 1e-3f * threadIdx.x + k;
  for (int32_t i = 0; i < iters; ++i) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
#pragma unroll
      for (int k = 0; k < kCalChains; ++k) {
        if constexpr (kOp == 0) v[k] = fmaf(v[k], a, b);
        else if constexpr (kOp == 1) v[k] = fmaf(at, bt, v[k]);
        else if constexpr (kOp == 2) v[k] = v[k] + bt;
        else if constexpr (kOp == 3) v[k] = v[k] * at;
        else if constexpr (kOp == 4) v[k] = fmaf(v[k], at, bt);
        else v[k] = fmaf(a, bt, v[k]);
      }
    }
  }
  float s = 0.f;
#pragma unroll
  for (int k = 0; k < kCalChains; ++k) s += v[k];
  out[int64_t(blockIdx.x) * kBlock + threadIdx.x] = s;
}

}  // namespace tatva

using namespace tatva;

extern "C" int tatva_valu_calibration(int32_t op, int32_t blocks, int32_t iters, float *out,
                                      void *stream) {
  TATVA_REQUIRE(op >= 0 && op <= 5 && blocks > 0 && iters > 0 && out, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  // a < 1 and b small keep every chain bounded and away from subnormals for
  // iters up to 2^16 (fma with v as the product operand: fixed point b / (1 - a);
  // add and fmac: grow by 32 b per trip at most; mul: decays by a^32 per trip)
  const float a = 0.9999f, b = 1e-4f;
  if (op == 0) k_valu_chains<0><<<blocks, kBlock, 0, s>>>(iters, a, b, out);
  else if (op == 1) k_valu_chains<1><<<blocks, kBlock, 0, s>>>(iters, a, b, out);
  else if (op == 2) k_valu_chains<2><<<blocks, kBlock, 0, s>>>(iters, a, b, out);
  else if (op == 3) k_valu_chains<3><<<blocks, kBlock, 0, s>>>(iters, a, b, out);
  else if (op == 4) k_valu_chains<4><<<blocks, kBlock, 0, s>>>(iters, a, b, out);
  else k_valu_chains<5><<<blocks, kBlock, 0, s>>>(iters, a, b, out);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}
