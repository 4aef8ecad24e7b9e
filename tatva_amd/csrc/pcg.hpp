// Jacobi-PCG state shared by the single-solve and batched solvers.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "common.hpp"
#include "mesh.hpp"

namespace tatva {

// Device-side record of one solve; block 0 of the last kernel of an
// iteration writes it, the host reads it between launches.
struct PcgState {
  double rr;
  double bb;
  double tol2;
  int iter;
  int done;
  int status;
  int maxiter;
  // single-reduction CG: alpha_i, gamma_i and the iteration count written by
  // the step kernel of parity q = i mod 2 and read at parity q ^ 1
  double alpha[2];
  double gamma[2];
  int it[2];
  // FP32 matrix-free CG: the p and x update of iteration i, written by its
  // step kernel (slot i mod 2) and applied by the next operator launch
  CgDeferCoef defer[2];
};

// Vector kernels use at most this many blocks (4 per CU): the partials each
// block re-sums stay a few KB.
constexpr int kVecBlocks = 1024;

inline int vec_grid(int64_t n) {
  int64_t b = (n + kBlock - 1) / kBlock;
  if (b < 1) b = 1;
  if (b > kVecBlocks) b = kVecBlocks;
  return (int)b;
}

// Grid of a CG step kernel whose threads load `per_thread` elements before
// summing the shares of the previous step, and of the kernels that write
// those shares.  Every step block sums all of them, so the grid is no larger
// than the prefetch needs: enough blocks for the prefetched elements to cover
// n, at least one block per CU (fewer only when n gives fewer blocks of one
// element per thread), at most kVecBlocks.
constexpr int kStepMinBlocks = 256;

inline int step_grid(int64_t n, int per_thread) {
  const int64_t per_block = (int64_t)kBlock * per_thread;
  const int64_t lo = std::min<int64_t>(kStepMinBlocks, vec_grid(n));
  return (int)std::min<int64_t>(kVecBlocks, std::max(lo, (n + per_block - 1) / per_block));
}

// Matrix-free FP32 PCG set-up on stream s: st reset, dinv = 1/diag,
// r = b - H(u) x, p = D^-1 r, the shares of r.z (into part_rz1), r.r
// (part_rr1) and b.b over gv blocks, and the state (tol, maxiter, done when
// b == 0 or r already small).  lin: the density's cached linearisation at u
// (launch_linearize_f32), or NULL to linearise in the product; geo: an affine
// Hex8 mesh's geometry cache (launch_geometry_cache_f32), or NULL.  diag_map:
// the vectors are full-length and diag is the reduced diagonal (n_red entries)
// read through this dof map (dinv = 0 outside the system), or NULL: diag is n
// long.
int matfree_pcg_setup(tatva_mesh *m, const tatva_density *d, const float *u,
                      const int32_t *dof_map, int64_t n, const float *diag, const float *b,
                      const float *x, float *r, float *p, float *Ap, float *dinv,
                      double *part_rz1, double *part_rr1, double *part_bb, int gv, double tol,
                      int maxiter, PcgState *st, hipStream_t s, const float *lin = nullptr,
                      const float *geo = nullptr, const int32_t *diag_map = nullptr,
                      int64_t n_red = 0);

}  // namespace tatva
