// CSR SpMV, Jacobi-PCG (FP64 CSR and FP32 matrix-free) with hipGraph-captured
// iteration chunks, and the small vector kernels of the Newton loop.
//
// The single-solve PCG runs the single-reduction (Chronopoulos-Gear) form of
// Jacobi-preconditioned CG: with u = D^-1 r and w = A u,
//   gamma_i = r_i.u_i, delta_i = w_i.u_i, beta_i = gamma_i / gamma_{i-1},
//   alpha_i = gamma_i / (delta_i - beta_i gamma_i / alpha_{i-1}),
//   p_i = u_i + beta_i p_{i-1}, s_i = w_i + beta_i s_{i-1}  (s_i = A p_i),
//   x_{i+1} = x_i + alpha_i p_i, r_{i+1} = r_i - alpha_i s_i,
// which equals standard PCG in exact arithmetic and needs one cross-block sum
// per iteration.  One iteration i = two or three kernels:
//   KA  w = A u and the blocks' shares of w.u (written by the operator's last
//       kernel, or by k_dot_shares after the node-block SpMV)  -> part_delta
//   KB  every block re-sums part_delta and the shares of r.u, r.r of the
//       previous KB (fixed order) to get the scalars, tests convergence on
//       r_i, updates p, s, x, r, u and writes the shares of r.u and r.r of
//       r_{i+1}                                                 -> part_rz[q], part_rr[q]
// with q = i mod 2 (a kernel argument; chunks have an even length).  There is
// no in-kernel fence or atomic: cross-block sums cross a kernel boundary, and
// every block computes identical scalars, so the iteration is deterministic.
// Block 0 alone writes the state record; scalars a kernel reads are written
// by the other parity, so no block reads a value block 0 of the same kernel
// writes.  Each kernel returns at once when the state says done, so a chunk
// may overrun convergence harmlessly.  The batched solver (batch_kernels.hip)
// runs the same form for every slot of a launch.
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <mutex>

#include "common.hpp"
#include "mesh.hpp"
#include "pcg.hpp"

namespace tatva {

// Partial-sum buffers of one solve: pap, rz[2], rr[2], bb (kMaxReduceBlocks each).
struct PcgParts {
  double *pap;
  double *rz0, *rz1;
  double *rr0, *rr1;
  double *bb;
  double *rz(int q) const { return q ? rz1 : rz0; }
  double *rr(int q) const { return q ? rr1 : rr0; }
};

// ----------------------------------------------------------------- SpMV ----

// LPR lanes per row; rows processed in a grid-stride loop over row groups.
template <int LPR>
__device__ __forceinline__ double csr_row_dot(int64_t row, int lane, const int32_t *__restrict__ indptr,
                                              const int32_t *__restrict__ indices,
                                              const double *__restrict__ data,
                                              const double *__restrict__ x) {
  double s = 0.0;
  const int32_t k1 = indptr[row + 1];
  for (int32_t k = indptr[row] + lane; k < k1; k += LPR) s += data[k] * x[indices[k]];
#pragma unroll
  for (int off = LPR / 2; off > 0; off >>= 1) s += __shfl_xor(s, off, LPR);
  return s;
}

// y = A x, or y = b - A x when b != NULL.
template <int LPR>
__global__ __launch_bounds__(kBlock) void k_spmv(int64_t n, const int32_t *__restrict__ indptr,
                                                 const int32_t *__restrict__ indices,
                                                 const double *__restrict__ data,
                                                 const double *__restrict__ x,
                                                 const double *__restrict__ b, double *__restrict__ y) {
  constexpr int G = kBlock / LPR;
  const int lane = threadIdx.x % LPR;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t stride = (int64_t)gridDim.x * G;
  for (int64_t row = lb * G + threadIdx.x / LPR; row < n; row += stride) {
    const double s = csr_row_dot<LPR>(row, lane, indptr, indices, data, x);
    if (lane == 0) y[row] = b ? b[row] - s : s;
  }
}

// ------------------------------------------------------ node-block SpMV ----
//
// FEM patterns with B dofs per node have rows in groups of B (one node) that
// share one column list made of aligned B-wide blocks.  The CSR arrays stay as
// they are; a derived list holds one int32 block column per BxB block, so the
// operator reads 8 + 4/B^2 bytes per non-zero instead of 12.  For block row I:
//   base = indptr[B I], L = indptr[B I + 1] - base (row length, same for its
//   B rows), row c of the node starts at base + c L, and its j-th block
//   column is bcol[base / B^2 + j].
// 32 lanes per block row; lane j owns block j (B x B values, B x entries).
constexpr int kBLanes = 32;

template <int B>
__device__ __forceinline__ void bcsr_block_row(int64_t I, int lane, const int32_t *__restrict__ indptr,
                                               const int32_t *__restrict__ bcol,
                                               const double *__restrict__ data,
                                               const double *__restrict__ x, double acc[B]) {
  const int32_t base = indptr[B * I];
  const int32_t L = indptr[B * I + 1] - base;
  const int32_t nb = L / B;
  const int32_t k0 = base / (B * B);
#pragma unroll
  for (int c = 0; c < B; ++c) acc[c] = 0.0;
  for (int32_t j = lane; j < nb; j += kBLanes) {
    const int64_t m = bcol[k0 + j];
    double xv[B];
#pragma unroll
    for (int cm = 0; cm < B; ++cm) xv[cm] = x[m * B + cm];
#pragma unroll
    for (int c = 0; c < B; ++c) {
      const double *row = data + base + (int64_t)c * L + (int64_t)j * B;
#pragma unroll
      for (int cm = 0; cm < B; ++cm) acc[c] += row[cm] * xv[cm];
    }
  }
#pragma unroll
  for (int c = 0; c < B; ++c)
#pragma unroll
    for (int off = kBLanes / 2; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, kBLanes);
}

template <int B>
__global__ __launch_bounds__(kBlock) void k_bspmv(int64_t nbr, const int32_t *__restrict__ indptr,
                                                  const int32_t *__restrict__ bcol,
                                                  const double *__restrict__ data,
                                                  const double *__restrict__ x,
                                                  const double *__restrict__ b, double *__restrict__ y) {
  constexpr int G = kBlock / kBLanes;
  const int lane = threadIdx.x % kBLanes;
  const int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  const int64_t stride = (int64_t)gridDim.x * G;
  for (int64_t I = lb * G + threadIdx.x / kBLanes; I < nbr; I += stride) {
    double acc[B];
    bcsr_block_row<B>(I, lane, indptr, bcol, data, x, acc);
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < B; ++c) y[B * I + c] = b ? b[B * I + c] - acc[c] : acc[c];
    }
  }
}

// K1 (node-block view): Ap = A p.  The shares of p.Ap come from k_dot_shares:
// folding the dot into this kernel cost 7-9 us per launch at 64^3 Hex8 (every
// variant measured), more than a separate pass over p and Ap.  With rev the
// blocks take the node rows from the end: alternating the direction between
// iterations, each pass starts on the matrix lines the previous pass read
// last, which the 256 MB last-level cache still holds (64^3 Hex8, in the
// solve: 91 instead of 108 us per launch).
template <int B>
__global__ __launch_bounds__(kBlock) void k_pcg_bspmv(int64_t nbr, const int32_t *__restrict__ indptr,
                                                      const int32_t *__restrict__ bcol,
                                                      const double *__restrict__ data,
                                                      const double *__restrict__ p,
                                                      double *__restrict__ Ap,
                                                      const PcgState *__restrict__ st, int rev) {
  if (st->done) return;
  constexpr int G = kBlock / kBLanes;
  const int lane = threadIdx.x % kBLanes;
  int64_t lb = xcd_block(blockIdx.x, gridDim.x);
  if (rev) lb = (int64_t)gridDim.x - 1 - lb;
  const int64_t stride = (int64_t)gridDim.x * G;
  for (int64_t I = lb * G + threadIdx.x / kBLanes; I < nbr; I += stride) {
    double acc[B];
    bcsr_block_row<B>(I, lane, indptr, bcol, data, p, acc);
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < B; ++c) Ap[B * I + c] = acc[c];
    }
  }
}

// Shares of p.Ap for the CG step: block k sums its grid-stride range of
// products in a fixed order.
__global__ __launch_bounds__(kBlock) void k_dot_shares(int64_t n, const double *__restrict__ p,
                                                       const double *__restrict__ Ap,
                                                       double *__restrict__ part,
                                                       const PcgState *__restrict__ st) {
  if (st->done) return;
  const int64_t stride = (int64_t)gridDim.x * kBlock;
  double acc = 0.0;
#pragma unroll 4
  for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += stride) acc += p[i] * Ap[i];
  acc = block_sum(acc);
  if (threadIdx.x == 0) part[blockIdx.x] = acc;
}

// Derives the block column list; flags *bad if the pattern is not made of
// aligned BxB blocks shared by the B rows of each node.
template <int B>
__global__ __launch_bounds__(kBlock) void k_block_view(int64_t nbr, const int32_t *__restrict__ indptr,
                                                       const int32_t *__restrict__ indices,
                                                       int32_t *__restrict__ bcol,
                                                       int *__restrict__ bad) {
  const int64_t I = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (I >= nbr) return;
  const int32_t base = indptr[B * I];
  const int32_t L = indptr[B * I + 1] - base;
  if (L % B != 0 || base % (B * B) != 0) {
    *bad = 1;
    return;
  }
  for (int c = 1; c < B; ++c)
    if (indptr[B * I + c] != base + c * L || indptr[B * I + c + 1] != base + (c + 1) * L) {
      *bad = 1;
      return;
    }
  for (int32_t j = 0; j < L / B; ++j) {
    const int32_t m = indices[base + j * B];
    if (m % B != 0) {
      *bad = 1;
      return;
    }
    for (int c = 0; c < B; ++c)
      for (int cm = 0; cm < B; ++cm)
        if (indices[base + c * L + j * B + cm] != m + cm) {
          *bad = 1;
          return;
        }
    bcol[base / (B * B) + j] = m / B;
  }
}

// ------------------------------------------------------- sliced ELL view ----
//
// A CSR matrix whose rows do not form node blocks (a node with some of its
// dofs constrained, a scalar field) is re-laid for the PCG as sliced ELL:
// slices of kSellRows consecutive rows, slice s padded to the length J_s of
// its longest row, entry j of row kSellRows s + l at (soff[s] + j) kSellRows
// + l (padding: value 0, column = the row itself).  One wave streams a slice
// with one row per lane, so every load instruction reads kSellRows
// consecutive values and no lane waits on a row pointer between entries.
// Each row is summed in CSR order (ascending column), as k_spmv does.
constexpr int kSellRows = kWave;
constexpr int kSellSlicesPerBlock = kBlock / kSellRows;

// soff[s + 1] = J_s (longest row of slice s); soff[0] = 0.
__global__ __launch_bounds__(kBlock) void k_sell_len(int64_t n, int64_t nsl,
                                                    const int32_t *__restrict__ indptr,
                                                    int64_t *__restrict__ soff) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s == 0) soff[0] = 0;
  if (s >= nsl) return;
  const int64_t r0 = s * kSellRows;
  const int64_t r1 = r0 + kSellRows < n ? r0 + kSellRows : n;
  int32_t J = 0;
  for (int64_t i = r0; i < r1; ++i) J = max(J, indptr[i + 1] - indptr[i]);
  soff[s + 1] = J;
}

// In-place inclusive scan of a[0..m) by one block of kBlock threads (each
// thread scans one contiguous chunk, then adds the sum of the chunks before).
__global__ __launch_bounds__(kBlock) void k_scan_inplace(int64_t *__restrict__ a, int64_t m) {
  __shared__ int64_t tot[kBlock];
  const int64_t per = (m + kBlock - 1) / kBlock;
  const int64_t lo = threadIdx.x * per;
  const int64_t hi = lo + per < m ? lo + per : m;
  int64_t run = 0;
  for (int64_t i = lo; i < hi; ++i) {
    run += a[i];
    a[i] = run;
  }
  tot[threadIdx.x] = run;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int t = 1; t < kBlock; ++t) tot[t] += tot[t - 1];
  __syncthreads();
  const int64_t before = threadIdx.x > 0 ? tot[threadIdx.x - 1] : 0;
  for (int64_t i = lo; i < hi; ++i) a[i] += before;
}

// One wave per slice: lane l copies row kSellRows s + l into the slice.
__global__ __launch_bounds__(kBlock) void k_sell_fill(int64_t n, int64_t nsl,
                                                     const int32_t *__restrict__ indptr,
                                                     const int32_t *__restrict__ indices,
                                                     const double *__restrict__ data,
                                                     const int64_t *__restrict__ soff,
                                                     int32_t *__restrict__ cols,
                                                     double *__restrict__ vals) {
  const int l = threadIdx.x % kSellRows;
  const int64_t s = (int64_t)blockIdx.x * kSellSlicesPerBlock + threadIdx.x / kSellRows;
  if (s >= nsl) return;
  const int64_t i = s * kSellRows + l;
  const int32_t base = i < n ? indptr[i] : 0;
  const int32_t len = i < n ? indptr[i + 1] - base : 0;
  const int64_t o = soff[s];
  const int J = (int)(soff[s + 1] - o);
  const int32_t self = i < n ? (int32_t)i : 0;
  for (int j = 0; j < J; ++j) {
    const int64_t pos = (o + j) * kSellRows + l;
    cols[pos] = j < len ? indices[base + j] : self;
    vals[pos] = j < len ? data[base + j] : 0.0;
  }
}

// y = A x (b given: y = b - A x); with part: the block's share of y . x.
// One wave per slice; skipped when st says the solve is done.
__global__ __launch_bounds__(kBlock) void k_sell_spmv(int64_t n, int64_t nsl,
                                                     const int64_t *__restrict__ soff,
                                                     const int32_t *__restrict__ cols,
                                                     const double *__restrict__ vals,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ y,
                                                     double *__restrict__ part,
                                                     const PcgState *__restrict__ st, int rev) {
  if (st && st->done) return;
  const int l = threadIdx.x % kSellRows;
  int64_t bl = xcd_block(blockIdx.x, gridDim.x);
  if (rev) bl = (int64_t)gridDim.x - 1 - bl;  // slices from the end (see k_pcg_bspmv)
  const int64_t s = bl * kSellSlicesPerBlock + threadIdx.x / kSellRows;
  double dot = 0.0;
  if (s < nsl) {
    const int64_t i = s * kSellRows + l;
    const double xi = i < n ? x[i] : 0.0;  // for the dot, loaded before the row
    const int64_t o = soff[s];
    const int J = (int)(soff[s + 1] - o);
    const int32_t *__restrict__ c = cols + o * kSellRows + l;
    const double *__restrict__ v = vals + o * kSellRows + l;
    double a = 0.0;
#pragma unroll 8
    for (int j = 0; j < J; ++j) a += v[j * kSellRows] * x[c[j * kSellRows]];
    if (i < n) {
      const double yi = b ? b[i] - a : a;
      y[i] = yi;
      dot = yi * xi;
    }
  }
  if (part) {
    dot = block_sum(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = dot;
  }
}

static int64_t sell_slices(int64_t n) { return (n + kSellRows - 1) / kSellRows; }
static int sell_grid(int64_t n) {
  return (int)std::max<int64_t>(1, (sell_slices(n) + kSellSlicesPerBlock - 1) / kSellSlicesPerBlock);
}

// ------------------------------------------------------------ PCG steps ----

// K1 for a matrix-free operator: the block's share of p.Ap.
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_dot(int64_t n, const T *__restrict__ p,
                                                    const T *__restrict__ Ap,
                                                    double *__restrict__ part_pap,
                                                    const PcgState *__restrict__ st) {
  if (st->done) return;
  double dot = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    dot += (double)p[i] * (double)Ap[i];
  dot = block_sum(dot);
  if (threadIdx.x == 0) part_pap[blockIdx.x] = dot;
}

// K2: alpha, x and r updates, shares of r.z and r.r for parity q.
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_update(int64_t n, T *__restrict__ x, T *__restrict__ r,
                                                       const T *__restrict__ p,
                                                       const T *__restrict__ Ap,
                                                       const T *__restrict__ dinv,
                                                       const double *__restrict__ part_pap, int n_pap,
                                                       const double *__restrict__ part_rz_old,
                                                       double *__restrict__ part_rz_new,
                                                       double *__restrict__ part_rr_new,
                                                       PcgState *__restrict__ st) {
  if (st->done) return;
  const int iter = st->iter;
  const double pap = sum_partials(part_pap, n_pap);
  const double rz = sum_partials(part_rz_old, gridDim.x);
  if (!(pap > 0.0) || !isfinite(pap) || iter >= st->maxiter) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->status = (iter >= st->maxiter) ? TATVA_ENOTCONV : TATVA_EBREAKDOWN;
    }
    return;
  }
  const T alpha = T(rz / pap);
  double srz = 0.0, srr = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * Ap[i];
    r[i] = ri;
    srz += (double)ri * (double)(dinv[i] * ri);
    srr += (double)ri * (double)ri;
  }
  srz = block_sum(srz);
  srr = block_sum(srr);
  if (threadIdx.x == 0) {
    part_rz_new[blockIdx.x] = srz;
    part_rr_new[blockIdx.x] = srr;
  }
}

// K3: convergence test and p = D^-1 r + beta p.
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_pupdate(int64_t n, T *__restrict__ p,
                                                        const T *__restrict__ r,
                                                        const T *__restrict__ dinv,
                                                        const double *__restrict__ part_rz_old,
                                                        const double *__restrict__ part_rz_new,
                                                        const double *__restrict__ part_rr_new,
                                                        PcgState *__restrict__ st) {
  if (st->done) return;
  const double rz_old = sum_partials(part_rz_old, gridDim.x);
  const double rz_new = sum_partials(part_rz_new, gridDim.x);
  const double rr = sum_partials(part_rr_new, gridDim.x);
  const bool converged = rr <= st->tol2;
  const bool bad = !isfinite(rr) || !isfinite(rz_new);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->iter += 1;
    st->rr = rr;
    if (converged) {
      st->done = 1;
      st->status = TATVA_OK;
    } else if (bad) {
      st->done = 1;
      st->status = TATVA_EBREAKDOWN;
    }
  }
  if (converged || bad) return;
  const T beta = T(rz_new / rz_old);
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = dinv[i] * r[i] + beta * p[i];
}

// Vectors of the CG step: FP32 in 16-byte vectors of 4 (a clang vector type,
// which the non-temporal store builtin takes: 115 -> 94 us per step at n =
// 12.4 M), FP64 one element at a time (in 16-byte vectors of 2 it measured
// slower, 162 -> 168 us).
typedef float step_f4 __attribute__((ext_vector_type(4)));
template <class T>
struct StepVec;
template <>
struct StepVec<float> {
  using type = step_f4;
  static constexpr int V = 4;
};
template <>
struct StepVec<double> {
  using type = double;
  static constexpr int V = 1;
};

// Vectors per thread whose loads the CG step issues before its scalar sums.
template <class T>
constexpr int step_prefetch_vecs() {
  return sizeof(T) == 4 ? 2 : 4;
}
// The same in elements (the step grid is sized by it).
template <class T>
constexpr int step_prefetch() {
  return step_prefetch_vecs<T>() * StepVec<T>::V;
}

// KB of the single-reduction form (see the file comment).  u is D^-1 r.  In
// FP64, r is not stored: u_{i+1} = u_i - alpha D^-1 s_i, and r_{i+1} =
// u_{i+1} / D^-1 is formed for the shares of r.u and r.r (ten vector streams
// per element instead of twelve; iteration counts within 1 %).  In FP32 that
// recurrence took 6-8 % more iterations, so r is kept there (r unused and
// may be NULL in FP64).  In FP32 each thread moves 16-byte vectors of every
// stream (the vectors are 16-byte aligned: they sit at 256-byte offsets of a
// 256-byte aligned workspace); the last n % 4 elements go one per thread.
// kDefer (FP32): p and x are left to the next operator launch (their
// coefficients go to st->defer[q]; x and p are not touched and may be NULL),
// and u is written, not read: the new z goes to u while the operator's launch
// still reads the old one from the other buffer (seven streams instead of
// twelve on the step's critical path).
template <class T, bool kDefer = false>
__global__ __launch_bounds__(kBlock) void k_cg_step(int64_t n, T *__restrict__ x, T *__restrict__ r,
                                                    T *__restrict__ u, T *__restrict__ p,
                                                    T *__restrict__ s, const T *__restrict__ w,
                                                    const T *__restrict__ dinv,
                                                    const double *__restrict__ part_delta,
                                                    int n_delta,
                                                    const double *__restrict__ part_g_old,
                                                    const double *__restrict__ part_rr_old,
                                                    double *__restrict__ part_g_new,
                                                    double *__restrict__ part_rr_new, int q,
                                                    PcgState *__restrict__ st) {
  static_assert(!kDefer || sizeof(T) == 4, "the deferred update is FP32");
  if (st->done) return;
  using VT = typename StepVec<T>::type;
  constexpr int V = StepVec<T>::V;
  // (the deferred form loads four streams instead of seven: twice the vectors
  // in flight before the sums, 447.9 -> 445.7 us per 160^3 iteration)
  constexpr int kPre = kDefer ? 2 * step_prefetch_vecs<T>() : step_prefetch_vecs<T>();
  constexpr bool kStoreR = sizeof(T) == 4;
  union Pack {
    VT v;
    T e[V];
  };
  const int iter = st->it[q ^ 1];
  const int64_t nv = n / V;  // whole vectors
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t i0 = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const VT *u4 = reinterpret_cast<const VT *>(u), *p4 = reinterpret_cast<const VT *>(p);
  const VT *w4 = reinterpret_cast<const VT *>(w), *s4 = reinterpret_cast<const VT *>(s);
  const VT *x4 = reinterpret_cast<const VT *>(x), *r4 = reinterpret_cast<const VT *>(r);
  const VT *d4 = reinterpret_cast<const VT *>(dinv);
  // the first kPre vectors of this thread are loaded before the scalars are
  // summed, so their latency overlaps the partial sums
  Pack ru[kPre], rp[kPre], rw[kPre], rs[kPre], rx[kPre], rv[kPre], rd[kPre];
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t iv = i0 + k * stride;
    if (iv < nv) {
      if constexpr (!kDefer) {
        ru[k].v = u4[iv];
        rp[k].v = p4[iv];
        rx[k].v = x4[iv];
      }
      rw[k].v = w4[iv];
      rs[k].v = s4[iv];
      if constexpr (kStoreR) rv[k].v = r4[iv];
      rd[k].v = d4[iv];
    }
  }
  double gamma, rr, delta;
  sum_partials3(part_g_old, gridDim.x, part_rr_old, gridDim.x, part_delta, n_delta, gamma, rr,
                delta);
  const bool first = iter == 0;
  const double beta = first ? 0.0 : gamma / st->gamma[q ^ 1];
  const double denom = first ? delta : delta - beta * gamma / st->alpha[q ^ 1];
  const bool converged = rr <= st->tol2;
  const bool spent = iter >= st->maxiter;
  const bool bad = !(denom > 0.0) || !isfinite(denom) || !isfinite(gamma) || !isfinite(rr);
  if (converged || spent || bad) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->done = 1;
      st->rr = rr;
      st->iter = iter;
      st->status = converged ? TATVA_OK : spent ? TATVA_ENOTCONV : TATVA_EBREAKDOWN;
    }
    return;
  }
  const double alpha = gamma / denom;
  const T a = T(alpha), bt = T(beta);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->alpha[q] = alpha;
    st->gamma[q] = gamma;
    st->it[q] = iter + 1;
    st->iter = iter + 1;
    st->rr = rr;
    if constexpr (kDefer) st->defer[q] = CgDeferCoef{a, bt, first ? 1 : 2, 0};
  }
  double sg = 0.0, srr = 0.0;
  // one element: p and s hold no values before iteration 0 (selected away,
  // never scaled); returns the new u
  const auto step1 = [&](T ui, T pi, T wi, T si, T xi, T rvi, T di, T &pn, T &sn, T &xn, T &rn) {
    if constexpr (!kDefer) {
      pn = first ? ui : ui + bt * pi;
      xn = xi + a * pn;
    }
    sn = first ? wi : wi + bt * si;
    double ri;
    T un;
    if constexpr (kStoreR) {
      rn = rvi - a * sn;
      un = di * rn;
      ri = (double)rn;
    } else {
      un = ui - a * (di * sn);
      ri = (double)un / (double)di;
    }
    sg += ri * (double)un;
    srr += di != T(0) ? ri * ri : 0.0;  // D^-1 = 0: a dof outside the system
    return un;
  };
  // x, r, p, s are stored non-temporally (read again only by the next step;
  // measured 4 % faster on the 64^3 Hex8 solve), u through the cache (the next
  // operator gathers it)
  const auto stepv = [&](int64_t iv, const Pack &U, const Pack &P, const Pack &W, const Pack &S,
                         const Pack &X, const Pack &R, const Pack &D) {
    Pack pn, sn, xn, rn, un;
#pragma unroll
    for (int e = 0; e < V; ++e)
      un.e[e] = step1(kDefer ? T(0) : U.e[e], kDefer ? T(0) : P.e[e], W.e[e], S.e[e],
                      kDefer ? T(0) : X.e[e], kStoreR ? R.e[e] : T(0), D.e[e], pn.e[e], sn.e[e],
                      xn.e[e], rn.e[e]);
    if constexpr (!kDefer) {
      __builtin_nontemporal_store(pn.v, reinterpret_cast<VT *>(p) + iv);
      __builtin_nontemporal_store(xn.v, reinterpret_cast<VT *>(x) + iv);
    }
    __builtin_nontemporal_store(sn.v, reinterpret_cast<VT *>(s) + iv);
    if constexpr (kStoreR) __builtin_nontemporal_store(rn.v, reinterpret_cast<VT *>(r) + iv);
    reinterpret_cast<VT *>(u)[iv] = un.v;
  };
#pragma unroll
  for (int k = 0; k < kPre; ++k) {
    const int64_t iv = i0 + k * stride;
    if (iv < nv) stepv(iv, ru[k], rp[k], rw[k], rs[k], rx[k], rv[k], rd[k]);
  }
  for (int64_t iv = i0 + kPre * stride; iv < nv; iv += stride) {
    Pack U, P, W, S, X, R, D;
    if constexpr (!kDefer) {
      U.v = u4[iv];
      P.v = p4[iv];
      X.v = x4[iv];
    }
    W.v = w4[iv];
    S.v = s4[iv];
    if constexpr (kStoreR) R.v = r4[iv];
    D.v = d4[iv];
    stepv(iv, U, P, W, S, X, R, D);
  }
  // the last n % V elements, one per thread
  if (i0 < n - nv * V) {
    const int64_t i = nv * V + i0;
    T pn, sn, xn, rn;
    const T un = step1(kDefer ? T(0) : u[i], kDefer ? T(0) : p[i], w[i], s[i],
                       kDefer ? T(0) : x[i], kStoreR ? r[i] : T(0), dinv[i], pn, sn, xn, rn);
    if constexpr (!kDefer) {
      p[i] = pn;
      x[i] = xn;
    }
    s[i] = sn;
    if constexpr (kStoreR) r[i] = rn;
    u[i] = un;
  }
  block_sum2(sg, srr);
  if (threadIdx.x == 0) {
    part_g_new[blockIdx.x] = sg;
    part_rr_new[blockIdx.x] = srr;
  }
}

// r given: z = D^-1 r, p = z, shares of r.z (into the parity-1 slot, read by
// iteration 0 as "previous"), r.r and b.b.
template <class T>
__global__ __launch_bounds__(kBlock) void k_pcg_init(int64_t n, const T *__restrict__ r,
                                                     const T *__restrict__ dinv,
                                                     const T *__restrict__ b, T *__restrict__ p,
                                                     double *__restrict__ part_rz,
                                                     double *__restrict__ part_rr,
                                                     double *__restrict__ part_bb) {
  double rz = 0.0, rr = 0.0, bb = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const T ri = r[i], di = dinv[i];
    const T z = di * ri;
    p[i] = z;
    rz += (double)ri * (double)z;
    // D^-1 = 0 marks a dof outside the system (full-length vectors): its r is
    // not part of the residual
    rr += di != T(0) ? (double)ri * (double)ri : 0.0;
    bb += (double)b[i] * (double)b[i];
  }
  rz = block_sum(rz);
  rr = block_sum(rr);
  bb = block_sum(bb);
  if (threadIdx.x == 0) {
    part_rz[blockIdx.x] = rz;
    part_rr[blockIdx.x] = rr;
    part_bb[blockIdx.x] = bb;
  }
}

// One block: state from the initial sums (b == 0 leaves x as given, as the
// oracle does; an initial residual already below tol needs no iteration).
__global__ __launch_bounds__(kBlock) void k_pcg_init_state(const double *__restrict__ part_rr,
                                                           const double *__restrict__ part_bb,
                                                           int nparts, double tol, int maxiter,
                                                           PcgState *__restrict__ st) {
  const double rr = sum_partials(part_rr, nparts);
  const double bb = sum_partials(part_bb, nparts);
  if (threadIdx.x == 0) {
    const int bad = st->status;  // set by the diagonal kernel
    st->rr = rr;
    st->bb = bb;
    st->tol2 = tol * tol * bb;
    st->iter = 0;
    st->it[0] = 0;
    st->it[1] = 0;
    st->defer[0] = CgDeferCoef{0.0f, 0.0f, 0, 0};
    st->defer[1] = CgDeferCoef{0.0f, 0.0f, 0, 0};
    st->maxiter = maxiter;
    int done = 1, status = TATVA_OK;
    if (bad != TATVA_OK) {
      status = bad;
    } else if (bb == 0.0 || rr <= st->tol2) {
      status = TATVA_OK;
    } else if (maxiter <= 0) {
      status = TATVA_ENOTCONV;
    } else {
      done = 0;
    }
    st->status = status;
    st->done = done;
  }
}

// dinv[row] = 1 / A[row, row] by binary search of the sorted row; a missing or
// zero diagonal flags st->status.
__global__ __launch_bounds__(kBlock) void k_csr_diag_inv(int64_t n, const int32_t *__restrict__ indptr,
                                                         const int32_t *__restrict__ indices,
                                                         const double *__restrict__ data,
                                                         double *__restrict__ dinv,
                                                         PcgState *__restrict__ st) {
  const int64_t row = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= n) return;
  int32_t lo = indptr[row], hi = indptr[row + 1] - 1;
  double d = 0.0;
  while (lo <= hi) {
    const int32_t mid = (lo + hi) >> 1;
    const int32_t c = indices[mid];
    if (c == row) {
      d = data[mid];
      break;
    }
    if (c < row) lo = mid + 1;
    else hi = mid - 1;
  }
  if (d != 0.0 && isfinite(d)) {
    dinv[row] = 1.0 / d;
  } else {
    dinv[row] = 1.0;
    st->status = TATVA_EINVAL;
  }
}

// diag -> 1/diag for FP32 Jacobi; a zero diagonal flags st->status.
__global__ __launch_bounds__(kBlock) void k_inv_diag_f32(int64_t n, const float *__restrict__ diag,
                                                         float *__restrict__ dinv,
                                                         PcgState *__restrict__ st) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float d = diag[i];
  if (d != 0.0f && isfinite(d)) {
    dinv[i] = 1.0f / d;
  } else {
    dinv[i] = 1.0f;
    st->status = TATVA_EINVAL;
  }
}

// Full-length vectors of a reduced system of n_red entries (dof map injective,
// -1 outside the system; an index outside [0, n_red) is treated as -1, so a
// malformed map cannot reach past the caller's vectors): out[i] = in[map[i]] or
// 0; the inverse diagonal with 0 outside the system; the reduced vector back.
__global__ __launch_bounds__(kBlock) void k_expand_f32(int64_t n, const int32_t *__restrict__ map,
                                                       int64_t n_red,
                                                       const float *__restrict__ in,
                                                       float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t r = map[i];
  out[i] = (r >= 0 && r < n_red) ? in[r] : 0.0f;
}

__global__ __launch_bounds__(kBlock) void k_inv_diag_map_f32(int64_t n, const int32_t *__restrict__ map,
                                                             int64_t n_red,
                                                             const float *__restrict__ diag,
                                                             float *__restrict__ dinv,
                                                             PcgState *__restrict__ st) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t r = map[i];
  if (r < 0 || r >= n_red) {
    dinv[i] = 0.0f;
    return;
  }
  const float d = diag[r];
  if (d != 0.0f && isfinite(d)) {
    dinv[i] = 1.0f / d;
  } else {
    dinv[i] = 1.0f;
    st->status = TATVA_EINVAL;
  }
}

__global__ __launch_bounds__(kBlock) void k_compress_f32(int64_t n, const int32_t *__restrict__ map,
                                                         int64_t n_red,
                                                         const float *__restrict__ in,
                                                         float *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t r = map[i];
  if (r >= 0 && r < n_red) out[r] = in[i];
}

template <class T>
__global__ __launch_bounds__(kBlock) void k_sub(int64_t n, const T *__restrict__ b,
                                                const T *__restrict__ Ax, T *__restrict__ r) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - Ax[i];
}

__global__ void k_state_reset(PcgState *st) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    st->status = TATVA_OK;
    st->done = 0;
    st->iter = 0;
  }
}

// ----------------------------------------------------- device resources ----

// An instantiated chunk graph and the launch parameters it was captured with.
// Solver kernels read and write only workspace buffers and the operator's
// arrays, so a graph is reused whenever these pointers and sizes repeat
// (Newton iterations, load steps, benchmark steps).
struct GraphKey {
  const void *ws;
  const void *a, *b, *c, *d, *e;  // operator arrays (CSR / mesh / state)
  int64_t n, m, k;
  int i0, i1, i2, i3;
  double f0, f1, f2;  // density parameters baked into captured kernel arguments
  bool operator==(const GraphKey &o) const { return std::memcmp(this, &o, sizeof(*this)) == 0; }
};

struct CachedGraph {
  GraphKey key{};
  bool valid = false;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void reset() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
    valid = false;
  }
};

struct DeviceRes {
  std::mutex mu;
  std::atomic<bool> init{false};
  hipStream_t cap = nullptr;
  hipEvent_t ev_in = nullptr, ev_out = nullptr;
  hipEvent_t prof_a = nullptr, prof_b = nullptr;  // timing events around a sampled operator
  hipEvent_t prof_m = nullptr;  // after the sampled operator's first kernel
  PcgState *host_state = nullptr;
  CachedGraph csr, matfree;
};

static DeviceRes g_res[64];
static std::mutex g_res_init_mu;  // serialises the first use of a device's resources
static std::atomic<int> g_profiling{0};
static std::atomic<int> g_lin_cache{0};  // TATVA_OPT_MATFREE_LIN_CACHE: 0 never, 1 Hex8, 2 always
static std::atomic<int> g_geo_cache{1};  // TATVA_OPT_MATFREE_GEOM_CACHE: 0 never, 1 affine Hex8
static std::atomic<int> g_cg_defer{1};   // TATVA_OPT_MATFREE_DEFERRED_UPDATE: 0 off, 1 on
static std::atomic<int> g_full_vectors{1};  // TATVA_OPT_MATFREE_FULL_VECTORS: 0 off, 1 on

// Drops cached solver graphs that captured this mesh's arrays (called before
// the mesh's device memory is freed).
void forget_mesh_graphs(const tatva_mesh *m) {
  for (auto &r : g_res) {
    if (!r.init) continue;
    std::lock_guard<std::mutex> lock(r.mu);
    if (r.matfree.valid && r.matfree.key.a == (const void *)m) r.matfree.reset();
  }
}

static int get_res(int device, DeviceRes **out) {
  if (device < 0 || device >= 64) {
    set_error("device id out of range");
    return TATVA_EINVAL;
  }
  DeviceRes &r = g_res[device];
  std::lock_guard<std::mutex> lock(g_res_init_mu);
  if (!r.init.load()) {
    TATVA_TRY_HIP(hipStreamCreateWithFlags(&r.cap, hipStreamNonBlocking));
    TATVA_TRY_HIP(hipEventCreateWithFlags(&r.ev_in, hipEventDisableTiming));
    TATVA_TRY_HIP(hipEventCreateWithFlags(&r.ev_out, hipEventDisableTiming));
    TATVA_TRY_HIP(hipEventCreate(&r.prof_a));
    TATVA_TRY_HIP(hipEventCreate(&r.prof_b));
    TATVA_TRY_HIP(hipEventCreate(&r.prof_m));
    TATVA_TRY_HIP(hipHostMalloc((void **)&r.host_state, sizeof(PcgState), hipHostMallocDefault));
    r.init.store(true);
  }
  *out = &r;
  return TATVA_OK;
}

struct SellView {
  int64_t nsl = 0, steps = 0;
  const int64_t *soff = nullptr;
  const int32_t *cols = nullptr;
  const double *vals = nullptr;
};

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

// Lays the CSR matrix out as sliced ELL in the caller's workspace from byte
// offset `at` (slice offsets, values, columns); one host synchronisation reads
// the padded size, which must fit the workspace.
static int build_sell(int64_t n, const int32_t *indptr, const int32_t *indices,
                      const double *data, char *ws, size_t at, size_t ws_bytes, hipStream_t s,
                      SellView &V) {
  V.nsl = sell_slices(n);
  int64_t *soff = reinterpret_cast<int64_t *>(ws + at);
  const size_t vals_at = align256(at + (size_t)(V.nsl + 1) * sizeof(int64_t));
  TATVA_REQUIRE(vals_at <= ws_bytes, "PCG workspace too small for the sliced ELL copy");
  k_sell_len<<<grid_for(V.nsl), kBlock, 0, s>>>(n, V.nsl, indptr, soff);
  k_scan_inplace<<<1, kBlock, 0, s>>>(soff, V.nsl + 1);
  TATVA_CHECK_LAUNCH();
  TATVA_TRY_HIP(hipMemcpyAsync(&V.steps, soff + V.nsl, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  TATVA_TRY_HIP(hipStreamSynchronize(s));
  const size_t entries = (size_t)V.steps * kSellRows;
  TATVA_REQUIRE(vals_at + entries * (sizeof(double) + sizeof(int32_t)) <= ws_bytes,
                "PCG workspace too small for the sliced ELL copy");
  double *vals = reinterpret_cast<double *>(ws + vals_at);
  int32_t *cols = reinterpret_cast<int32_t *>(vals + entries);
  k_sell_fill<<<sell_grid(n), kBlock, 0, s>>>(n, V.nsl, indptr, indices, data, soff, cols, vals);
  TATVA_CHECK_LAUNCH();
  V.soff = soff;
  V.cols = cols;
  V.vals = vals;
  return TATVA_OK;
}

static int choose_lpr(int64_t n, int64_t nnz) {
  const double avg = n > 0 ? double(nnz) / double(n) : 1.0;
  if (avg <= 6) return 4;
  if (avg <= 14) return 8;
  if (avg <= 40) return 16;
  if (avg <= 96) return 32;
  return 64;
}

template <class F>
static int with_lpr(int lpr, F &&f) {
  switch (lpr) {
    case 4: return f(std::integral_constant<int, 4>{});
    case 8: return f(std::integral_constant<int, 8>{});
    case 16: return f(std::integral_constant<int, 16>{});
    case 32: return f(std::integral_constant<int, 32>{});
    default: return f(std::integral_constant<int, 64>{});
  }
}

static int spmv_grid(int64_t n, int lpr) {
  const int64_t rows_per_block = kBlock / lpr;
  int64_t b = (n + rows_per_block - 1) / rows_per_block;
  if (b > kMaxReduceBlocks) b = kMaxReduceBlocks;
  return (int)(b < 1 ? 1 : b);
}

// One block per rows_per_block rows, no cap: the hardware dispatcher balances
// the uneven rows (measured ~15 % faster than a grid-stride loop over 2048
// blocks on the 64^3 Hex8 stiffness).
static int spmv_grid_full(int64_t n, int lpr) {
  const int64_t rows_per_block = kBlock / lpr;
  const int64_t b = (n + rows_per_block - 1) / rows_per_block;
  return (int)(b < 1 ? 1 : b);
}

struct WsLayout {
  size_t vec[12];
  size_t parts;
  size_t state;
  size_t op_parts;  // per-block shares of the operator's dot (one per SpMV block)
  size_t total;
};


// nvec vectors of n elements of size esz, 6 partial buffers, the state record.
static WsLayout ws_layout(int64_t n, int nvec, size_t esz, int64_t n_op_parts = 0) {
  WsLayout L{};
  size_t o = 0;
  for (int i = 0; i < nvec; ++i) {
    L.vec[i] = o;
    o = align256(o + (size_t)n * esz);
  }
  L.parts = o;
  o = align256(o + 6 * kMaxReduceBlocks * sizeof(double));
  L.state = o;
  o = align256(o + sizeof(PcgState));
  L.op_parts = o;
  o = align256(o + (size_t)n_op_parts * sizeof(double));
  L.total = o;
  return L;
}

// Blocks' shares of a dot product (more than a CG step re-sums) reduced to
// kShareBlocks partials: block k sums the contiguous k-th part of p[0..n) in
// a fixed order, out[k] = its sum.
constexpr int kShareBlocks = 64;
__global__ __launch_bounds__(kBlock) void k_sum_shares(const double *__restrict__ p, int n,
                                                      double *__restrict__ out,
                                                      const PcgState *__restrict__ st) {
  if (st->done) return;
  const int per = (n + kShareBlocks - 1) / kShareBlocks;
  const int lo = blockIdx.x * per;
  const int hi = min(lo + per, n);
  double v = 0.0;
  for (int i = lo + threadIdx.x; i < hi; i += kBlock) v += p[i];
  v = block_sum(v);
  if (threadIdx.x == 0) out[blockIdx.x] = v;
}

// Typed views of a workspace; base is the workspace's first byte.
template <class T>
static T *ws_vec(char *base, const WsLayout &L, int i) {
  return reinterpret_cast<T *>(base + L.vec[i]);
}

static PcgState *ws_state(char *base, const WsLayout &L) {
  return reinterpret_cast<PcgState *>(base + L.state);
}

static PcgParts parts_at(char *w, const WsLayout &L) {
  double *b = (double *)(w + L.parts);
  return PcgParts{b, b + kMaxReduceBlocks, b + 2 * kMaxReduceBlocks, b + 3 * kMaxReduceBlocks,
                  b + 4 * kMaxReduceBlocks, b + 5 * kMaxReduceBlocks};
}

// Runs captured chunks until the device state reports done; fills stats.
// With profiling, the graph holds the chunk minus its first operator
// application, which `sampled_op` launches eagerly on the same stream between
// prof_a and prof_b; its duration is read after each chunk.
// With mid_sampled, sampled_op also records prof_m after its first kernel,
// whose duration (prof_a to prof_m) is summed into op_main_ms.
static int run_chunks(DeviceRes *res, hipGraphExec_t exec, PcgState *st, int maxiter, int chunk,
                      bool profiling, const std::function<int()> &sampled_op,
                      tatva_pcg_stats *stats,
                      const std::function<int()> &after_sample = std::function<int()>(),
                      bool mid_sampled = false) {
  const int max_launch = (maxiter + chunk - 1) / chunk + 1;
  double op_ms = 0.0, main_ms = 0.0;
  int samples = 0;
  bool launched = false;
  for (int l = 0; l <= max_launch + 1; ++l) {
    TATVA_TRY_HIP(hipMemcpyAsync(res->host_state, st, sizeof(PcgState), hipMemcpyDeviceToHost,
                                 res->cap));
    TATVA_TRY_HIP(hipStreamSynchronize(res->cap));
    if (launched && profiling) {
      float ms = 0.0f;
      TATVA_TRY_HIP(hipEventElapsedTime(&ms, res->prof_a, res->prof_b));
      op_ms += ms;
      if (mid_sampled) {
        TATVA_TRY_HIP(hipEventElapsedTime(&ms, res->prof_a, res->prof_m));
      }
      main_ms += ms;
      ++samples;
    }
    if (res->host_state->done) break;
    if (profiling) {
      TATVA_TRY_HIP(hipEventRecord(res->prof_a, res->cap));
      int rc = sampled_op();
      if (rc) return rc;
      TATVA_TRY_HIP(hipEventRecord(res->prof_b, res->cap));
      if (after_sample) {
        rc = after_sample();
        if (rc) return rc;
      }
    }
    TATVA_TRY_HIP(hipGraphLaunch(exec, res->cap));
    launched = true;
  }
  const PcgState &h = *res->host_state;
  if (stats) {
    stats->iterations = h.iter;
    stats->status = h.done ? h.status : TATVA_ENOTCONV;
    stats->b_norm = std::sqrt(h.bb);
    stats->rel_residual = h.bb > 0 ? std::sqrt(h.rr / h.bb) : 0.0;
    stats->op_ms = op_ms;
    stats->op_samples = samples;
    stats->reserved = 0;
    stats->op_main_ms = main_ms;
  }
  return TATVA_OK;
}

// Returns the chunk graph (even length) of `step(q)` calls, q = iteration
// parity, capturing it unless the cache already holds one for `key`.
static int chunk_graph(hipStream_t s, int chunk, const std::function<int(int, bool)> &step,
                       bool profiling, const GraphKey &key, CachedGraph &cache) {
  if (cache.valid && cache.key == key) return TATVA_OK;
  cache.reset();
  TATVA_TRY_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = TATVA_OK;
  for (int i = 0; i < chunk && rc == TATVA_OK; ++i) rc = step(i & 1, !(i == 0 && profiling));
  hipError_t ce = hipStreamEndCapture(s, &cache.graph);
  if (rc) {
    cache.reset();
    return rc;
  }
  if (ce != hipSuccess) {
    cache.reset();
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
    return TATVA_EHIP;
  }
  hipError_t ie = hipGraphInstantiate(&cache.exec, cache.graph, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    cache.reset();
    set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
    return TATVA_EHIP;
  }
  cache.key = key;
  cache.valid = true;
  return TATVA_OK;
}

}  // namespace tatva

using namespace tatva;

extern "C" int tatva_set_profiling(int enable) {
  g_profiling.store(enable ? 1 : 0);
  return TATVA_OK;
}

extern "C" int tatva_set_option(int32_t option, int32_t value) {
  switch (option) {
    case TATVA_OPT_MATFREE_LIN_CACHE:
      TATVA_REQUIRE(value >= 0 && value <= 2, "linearisation cache option: 0, 1 or 2");
      g_lin_cache.store(value);
      return TATVA_OK;
    case TATVA_OPT_MATFREE_GEOM_CACHE:
      TATVA_REQUIRE(value >= 0 && value <= 1, "geometry cache option: 0 or 1");
      g_geo_cache.store(value);
      return TATVA_OK;
    case TATVA_OPT_MATFREE_DEFERRED_UPDATE:
      TATVA_REQUIRE(value >= 0 && value <= 1, "deferred update option: 0 or 1");
      g_cg_defer.store(value);
      return TATVA_OK;
    case TATVA_OPT_MATFREE_FULL_VECTORS:
      TATVA_REQUIRE(value >= 0 && value <= 1, "full-length vectors option: 0 or 1");
      g_full_vectors.store(value);
      return TATVA_OK;
    default: set_error("unknown option"); return TATVA_EINVAL;
  }
}

extern "C" int tatva_spmv_csr(int64_t n, int64_t nnz, const int32_t *indptr,
                              const int32_t *indices, const double *data, const double *x,
                              double *y, void *stream) {
  TATVA_REQUIRE(n >= 0 && nnz >= 0 && indptr && indices && data && x && y, "invalid arguments");
  if (n == 0) return TATVA_OK;
  hipStream_t s = (hipStream_t)stream;
  // The average row length picks the lanes per row.
  return with_lpr(choose_lpr(n, nnz), [&](auto lc) -> int {
    constexpr int L = decltype(lc)::value;
    k_spmv<L><<<spmv_grid(n, L), kBlock, 0, s>>>(n, indptr, indices, data, x, nullptr, y);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
}

extern "C" int tatva_csr_block_view(int64_t n, int64_t nnz, const int32_t *indptr,
                                    const int32_t *indices, int32_t block_size,
                                    int32_t *block_cols, int32_t *ok, void *stream) {
  TATVA_REQUIRE(n > 0 && nnz >= 0 && indptr && indices && block_cols && ok, "invalid arguments");
  TATVA_REQUIRE(block_size == 2 || block_size == 3, "block_size must be 2 or 3");
  *ok = 0;
  if (n % block_size != 0 || nnz % (block_size * block_size) != 0) return TATVA_OK;
  hipStream_t s = (hipStream_t)stream;
  // setup call: a 4-byte flag, freed before returning
  int *bad = nullptr;
  TATVA_TRY_HIP(hipMalloc(&bad, sizeof(int)));
  int h_bad = 0;
  hipError_t e = hipMemsetAsync(bad, 0, sizeof(int), s);
  const int64_t nbr = n / block_size;
  if (e == hipSuccess) {
    if (block_size == 3)
      k_block_view<3><<<grid_for(nbr), kBlock, 0, s>>>(nbr, indptr, indices, block_cols, bad);
    else
      k_block_view<2><<<grid_for(nbr), kBlock, 0, s>>>(nbr, indptr, indices, block_cols, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&h_bad, bad, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(bad);
  if (e != hipSuccess) {
    set_error(std::string("block view: ") + hipGetErrorString(e));
    return TATVA_EHIP;
  }
  *ok = h_bad ? 0 : 1;
  return TATVA_OK;
}

extern "C" int tatva_spmv_bcsr(int64_t n, const int32_t *indptr, const int32_t *block_cols,
                               int32_t block_size, const double *data, const double *x, double *y,
                               void *stream) {
  TATVA_REQUIRE(n > 0 && indptr && block_cols && data && x && y, "invalid arguments");
  TATVA_REQUIRE((block_size == 2 || block_size == 3) && n % block_size == 0,
                "block_size must be 2 or 3 and divide n");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nbr = n / block_size;
  if (block_size == 3)
    k_bspmv<3><<<spmv_grid(nbr, kBLanes), kBlock, 0, s>>>(nbr, indptr, block_cols, data, x,
                                                         nullptr, y);
  else
    k_bspmv<2><<<spmv_grid(nbr, kBLanes), kBlock, 0, s>>>(nbr, indptr, block_cols, data, x,
                                                         nullptr, y);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

// workspace vectors: p, r, Ap, D^-1, x, b
// workspace: u, w, p, s, r, D^-1, x, b (FP64), partials, state
constexpr int kCsrVecs = 8;

// One operator-dot share per SpMV block when the SpMV grid is not capped: at
// most one block per 4 rows (64 lanes per row).
static int64_t csr_op_parts(int64_t n) { return (n + 3) / 4; }

extern "C" int64_t tatva_pcg_csr_workspace(int64_t n, const int32_t *host_indptr) {
  if (n <= 0) return -1;
  size_t total = ws_layout(n, kCsrVecs, sizeof(double), csr_op_parts(n)).total;
  if (host_indptr) {
    // the sliced ELL copy: slice offsets, then every slice padded to its longest row
    const int64_t nsl = sell_slices(n);
    int64_t steps = 0;
    for (int64_t sl = 0; sl < nsl; ++sl) {
      int64_t longest = 0;
      for (int64_t r = sl * kSellRows; r < std::min<int64_t>(n, (sl + 1) * kSellRows); ++r)
        longest = std::max<int64_t>(longest, (int64_t)host_indptr[r + 1] - host_indptr[r]);
      steps += longest;
    }
    total = align256(total + (size_t)(nsl + 1) * sizeof(int64_t));
    total += (size_t)steps * kSellRows * (sizeof(double) + sizeof(int32_t));
  }
  return (int64_t)total;
}

extern "C" int tatva_pcg_csr(int64_t n, int64_t nnz, const int32_t *indptr,
                             const int32_t *indices, const double *data,
                             const int32_t *block_cols, int32_t block_size, const double *b_in,
                             double *x_io, double tol, int32_t maxiter, int32_t chunk,
                             void *workspace, int64_t ws_bytes, tatva_pcg_stats *stats,
                             void *stream) {
  TATVA_REQUIRE(n > 0 && nnz > 0 && indptr && indices && data && b_in && x_io, "invalid arguments");
  TATVA_REQUIRE(chunk >= 1 && chunk <= 4096, "chunk must be in [1, 4096]");
  if (block_cols)
    TATVA_REQUIRE((block_size == 2 || block_size == 3) && n % block_size == 0,
                  "block view needs block_size 2 or 3 dividing n");
  TATVA_REQUIRE(tol >= 0.0, "tol must be non-negative");
  const WsLayout L = ws_layout(n, kCsrVecs, sizeof(double), csr_op_parts(n));
  TATVA_REQUIRE(workspace && ws_bytes >= (int64_t)L.total, "PCG workspace too small");
  TATVA_REQUIRE(((uintptr_t)workspace & 255) == 0, "PCG workspace must be 256-byte aligned");
  chunk += chunk & 1;  // parity ping-pong needs an even chunk
  int device = 0;
  TATVA_TRY_HIP(hipGetDevice(&device));
  DeviceRes *res = nullptr;
  int rc = get_res(device, &res);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(res->mu);
  char *wsb = (char *)workspace;
  double *u = ws_vec<double>(wsb, L, 0);
  double *w = ws_vec<double>(wsb, L, 1);
  double *p = ws_vec<double>(wsb, L, 2);
  double *sv = ws_vec<double>(wsb, L, 3);
  double *r = ws_vec<double>(wsb, L, 4);
  double *dinv = ws_vec<double>(wsb, L, 5);
  double *x = ws_vec<double>(wsb, L, 6);
  double *b = ws_vec<double>(wsb, L, 7);
  const PcgParts P = parts_at(wsb, L);
  PcgState *st = ws_state(wsb, L);
  hipStream_t caller = (hipStream_t)stream;
  hipStream_t s = res->cap;
  TATVA_TRY_HIP(hipEventRecord(res->ev_in, caller));
  TATVA_TRY_HIP(hipStreamWaitEvent(s, res->ev_in, 0));
  TATVA_TRY_HIP(hipMemcpyAsync(x, x_io, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipMemcpyAsync(b, b_in, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  const int gv = step_grid(n, step_prefetch<double>());
  const bool profiling = g_profiling.load() != 0;
  const int nblk = block_cols ? block_size : 1;
  // operator: node-block view when given, else a sliced ELL copy of the CSR
  SellView sv_ell;
  if (nblk == 1) {
    rc = build_sell(n, indptr, indices, data, wsb, L.total, (size_t)ws_bytes, s, sv_ell);
    if (rc) return rc;
  }
  const int g_op = nblk > 1 ? spmv_grid_full(n / nblk, kBLanes) : sell_grid(n);
  // more blocks than the partial buffers hold: the shares go to op_parts and
  // k_sum_shares reduces them to kShareBlocks partials in P.pap
  const bool big = g_op > kMaxReduceBlocks;
  TATVA_REQUIRE(g_op <= csr_op_parts(n), "SpMV grid exceeds the workspace shares");
  double *op_part = big ? reinterpret_cast<double *>(wsb + L.op_parts) : P.pap;
  // node-block operator: its shares come from k_dot_shares over g_dot blocks
  const int g_dot = step_grid(n, 4);
  const int n_delta = nblk > 1 ? g_dot : big ? kShareBlocks : g_op;
  auto apply_op = [&](const double *xin, const double *bin, double *yout) -> int {
    if (nblk == 3)
      k_bspmv<3><<<g_op, kBlock, 0, s>>>(n / 3, indptr, block_cols, data, xin, bin, yout);
    else if (nblk == 2)
      k_bspmv<2><<<g_op, kBlock, 0, s>>>(n / 2, indptr, block_cols, data, xin, bin, yout);
    else
      k_sell_spmv<<<g_op, kBlock, 0, s>>>(n, sv_ell.nsl, sv_ell.soff, sv_ell.cols, sv_ell.vals, xin,
                                          bin, yout, nullptr, nullptr, 0);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  };
  // KA: w = A u and the shares of w.u (node blocks: by k_dot_shares; sliced
  // ELL: fused, summed to kShareBlocks partials when big)
  // (rev: iteration parity; the matrix is streamed in alternating directions)
  auto spmv_dot = [&](int rev) -> int {
    if (nblk == 3)
      k_pcg_bspmv<3><<<g_op, kBlock, 0, s>>>(n / 3, indptr, block_cols, data, u, w, st, rev);
    else if (nblk == 2)
      k_pcg_bspmv<2><<<g_op, kBlock, 0, s>>>(n / 2, indptr, block_cols, data, u, w, st, rev);
    else
      k_sell_spmv<<<g_op, kBlock, 0, s>>>(n, sv_ell.nsl, sv_ell.soff, sv_ell.cols, sv_ell.vals, u,
                                          nullptr, w, op_part, st, rev);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  };
  const std::function<int()> sum_shares = [&]() -> int {
    if (nblk > 1)
      k_dot_shares<<<g_dot, kBlock, 0, s>>>(n, u, w, P.pap, st);
    else if (big)
      k_sum_shares<<<kShareBlocks, kBlock, 0, s>>>(op_part, g_op, P.pap, st);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  };
  const std::function<int(int)> apply_op_dot = [&](int q) -> int {
    const int e = spmv_dot(q);
    return e ? e : sum_shares();
  };
  // the eagerly timed operator of a profiled chunk is its iteration 0 (parity 0)
  const std::function<int()> sampled_spmv = [&]() -> int { return spmv_dot(0); };
  k_state_reset<<<1, 64, 0, s>>>(st);
  k_csr_diag_inv<<<grid_for(n), kBlock, 0, s>>>(n, indptr, indices, data, dinv, st);
  rc = apply_op(x, b, r);  // r = b - A x
  if (rc) return rc;
  // u = D^-1 r and the shares of r.u, r.r, b.b (parity 1: read by iteration 0)
  k_pcg_init<double><<<gv, kBlock, 0, s>>>(n, r, dinv, b, u, P.rz(1), P.rr(1), P.bb);
  k_pcg_init_state<<<1, kBlock, 0, s>>>(P.rr(1), P.bb, gv, tol, maxiter, st);
  TATVA_CHECK_LAUNCH();
  // with profiling, iteration 0 of a chunk has its SpMV (timed) and share sum
  // launched eagerly before the graph
  const std::function<int(int, bool)> step = [&](int q, bool with_op) -> int {
    if (with_op) {
      const int e = apply_op_dot(q);
      if (e) return e;
    }
    k_cg_step<double><<<gv, kBlock, 0, s>>>(n, x, nullptr, u, p, sv, w, dinv, P.pap, n_delta,
                                            P.rz(q ^ 1), P.rr(q ^ 1), P.rz(q), P.rr(q), q, st);
    return TATVA_OK;
  };
  GraphKey key{};
  key.ws = workspace;
  key.a = indptr;
  key.b = indices;
  key.c = data;
  key.d = block_cols ? (const void *)block_cols : (const void *)sv_ell.soff;
  key.e = sv_ell.vals;
  key.n = n;
  key.m = nnz;
  key.k = sv_ell.steps;
  key.i0 = nblk;
  key.i1 = chunk;
  key.i2 = profiling ? 1 : 0;
  rc = chunk_graph(s, chunk, step, profiling, key, res->csr);
  if (rc == TATVA_OK)
    rc = run_chunks(res, res->csr.exec, st, maxiter, chunk, profiling, sampled_spmv, stats,
                    sum_shares);
  if (rc == TATVA_OK)
    TATVA_TRY_HIP(hipMemcpyAsync(x_io, x, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipEventRecord(res->ev_out, s));
  TATVA_TRY_HIP(hipStreamWaitEvent(caller, res->ev_out, 0));
  return rc;
}

// workspace: reduced vectors p, r, Ap, D^-1, x, b, D (FP32), then the full FP32 state u
namespace tatva {

int matfree_pcg_setup(tatva_mesh *m, const tatva_density *d, const float *u,
                      const int32_t *dof_map, int64_t n, const float *diag, const float *b,
                      const float *x, float *r, float *p, float *Ap, float *dinv,
                      double *part_rz1, double *part_rr1, double *part_bb, int gv, double tol,
                      int maxiter, PcgState *st, hipStream_t s, const float *lin,
                      const float *geo, const int32_t *diag_map, int64_t n_red) {
  k_state_reset<<<1, 64, 0, s>>>(st);
  if (diag_map)
    k_inv_diag_map_f32<<<grid_for(n), kBlock, 0, s>>>(n, diag_map, n_red, diag, dinv, st);
  else
    k_inv_diag_f32<<<grid_for(n), kBlock, 0, s>>>(n, diag, dinv, st);
  // r = b - H x
  const int rc = launch_hvp_f32(m, d, u, x, dof_map, Ap, s, nullptr, lin, geo);
  if (rc) return rc;
  k_sub<float><<<grid_for(n), kBlock, 0, s>>>(n, b, Ap, r);
  k_pcg_init<float><<<gv, kBlock, 0, s>>>(n, r, dinv, b, p, part_rz1, part_rr1, part_bb);
  k_pcg_init_state<<<1, kBlock, 0, s>>>(part_rr1, part_bb, gv, tol, maxiter, st);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

}  // namespace tatva

// workspace: FP32 vectors z (= D^-1 r), w, p, s, r, D^-1, x, b, D, the second z
// buffer (reduced, or full-length under TATVA_OPT_MATFREE_FULL_VECTORS), the
// partials and the state, then the full FP32 state u, then the caches the
// options select for this mesh and density: the density's linearisation at u
// and the affine Hex8 geometry (each 0 bytes when not selected)
constexpr int kMatfreeVecs = 10;

// Length of the CG vectors the workspace is sized for: with full-length
// vectors the larger of n and the full dof count.
static int64_t matfree_vec_len(const tatva_mesh *m, const tatva_density *d, int64_t n) {
  return g_full_vectors.load() == 1 ? std::max<int64_t>(n, m->n_nodes * density_ncomp(m, d)) : n;
}

static size_t matfree_u_offset(int64_t n) { return ws_layout(n, kMatfreeVecs, sizeof(float)).total; }

static size_t matfree_lin_offset(const tatva_mesh *m, int64_t n) {
  return matfree_u_offset(n) + align256((size_t)(m->n_nodes * m->dim) * sizeof(float));
}

// Bytes of the two caches under the current options.
static size_t matfree_lin_bytes(const tatva_mesh *m, const tatva_density *d) {
  const int opt = g_lin_cache.load();
  const bool on = opt == 2 || (opt == 1 && m->kind == TATVA_HEX8);
  return on ? align256((size_t)lin_cache_floats(m, d) * sizeof(float)) : 0;
}
static size_t matfree_geo_bytes(const tatva_mesh *m) {
  return g_geo_cache.load() == 1 ? align256((size_t)geo_cache_floats(m) * sizeof(float)) : 0;
}

extern "C" int64_t tatva_pcg_matfree_workspace(tatva_mesh *m, const tatva_density *d, int64_t n) {
  if (!m || !d || n <= 0) return -1;
  const int64_t nv = matfree_vec_len(m, d, n);
  return (int64_t)(matfree_lin_offset(m, nv) + matfree_lin_bytes(m, d) + matfree_geo_bytes(m));
}

extern "C" int tatva_pcg_matfree_f32(tatva_mesh *m, const tatva_density *d, const float *u_in,
                                     const int32_t *dof_map, int64_t n, const float *diag_in,
                                     const float *b_in, float *x_io, double tol, int32_t maxiter,
                                     int32_t chunk, void *workspace, int64_t ws_bytes,
                                     tatva_pcg_stats *stats, void *stream) {
  TATVA_REQUIRE(m && d && u_in && diag_in && b_in && x_io && n > 0, "invalid arguments");
  TATVA_REQUIRE(chunk >= 1 && chunk <= 4096, "chunk must be in [1, 4096]");
  TATVA_REQUIRE(((uintptr_t)workspace & 255) == 0, "PCG workspace must be 256-byte aligned");
  TATVA_REQUIRE(workspace && ws_bytes >= tatva_pcg_matfree_workspace(m, d, n),
                "PCG workspace too small");
  const int64_t n_full = m->n_nodes * density_ncomp(m, d);
  // full-length vectors (option, on by default): with a dof map the CG runs on
  // every dof of the mesh, D^-1 = 0 outside the system, so that the products
  // index the vectors by node and no kernel of the iteration reads the map;
  // b, x and the diagonal go in and x comes out through the map once per solve
  const bool full = dof_map && g_full_vectors.load() == 1;
  const int64_t nv = full ? n_full : n;
  const int32_t *op_map = full ? nullptr : dof_map;
  const WsLayout L = ws_layout(nv, kMatfreeVecs, sizeof(float));
  chunk += chunk & 1;
  TATVA_TRY_HIP(hipSetDevice(m->device));
  DeviceRes *res = nullptr;
  int rc = get_res(m->device, &res);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(res->mu);
  char *wsb = (char *)workspace;
  float *z = ws_vec<float>(wsb, L, 0);
  float *w = ws_vec<float>(wsb, L, 1);
  float *p = ws_vec<float>(wsb, L, 2);
  float *sv = ws_vec<float>(wsb, L, 3);
  float *r = ws_vec<float>(wsb, L, 4);
  float *dinv = ws_vec<float>(wsb, L, 5);
  float *x = ws_vec<float>(wsb, L, 6);
  float *b = ws_vec<float>(wsb, L, 7);
  float *diag = ws_vec<float>(wsb, L, 8);
  float *z2 = ws_vec<float>(wsb, L, 9);
  float *u = reinterpret_cast<float *>(wsb + matfree_u_offset(nv));
  // the state is fixed during the solve: the density may be linearised once and
  // read back by every product (option; off by default: on the 160^3 Hex8 cube
  // the product reading 40 bytes per point ran at 515 us against 433 us for the
  // one recomputing them, its memory and VALU work not overlapping at 4 waves
  // per SIMD)
  const size_t lin_bytes = matfree_lin_bytes(m, d);
  float *lin = lin_bytes ? reinterpret_cast<float *>(wsb + matfree_lin_offset(m, nv)) : nullptr;
  // affine Hex8: the elements' J^-1 and det J are written once per solve and
  // read by every product instead of the coordinates (option, on by default: the
  // 64^3 linear elasticity product 22.1 -> 19.9 us, the 160^3 Neo-Hookean one
  // unchanged at 334 us)
  float *geo = matfree_geo_bytes(m)
                   ? reinterpret_cast<float *>(wsb + matfree_lin_offset(m, nv) + lin_bytes)
                   : nullptr;
  const PcgParts P = parts_at(wsb, L);
  PcgState *st = ws_state(wsb, L);
  hipStream_t caller = (hipStream_t)stream;
  hipStream_t s = res->cap;
  TATVA_TRY_HIP(hipEventRecord(res->ev_in, caller));
  TATVA_TRY_HIP(hipStreamWaitEvent(s, res->ev_in, 0));
  if (full) {
    k_expand_f32<<<grid_for(nv), kBlock, 0, s>>>(nv, dof_map, n, x_io, x);
    k_expand_f32<<<grid_for(nv), kBlock, 0, s>>>(nv, dof_map, n, b_in, b);
    TATVA_CHECK_LAUNCH();
  } else {
    TATVA_TRY_HIP(hipMemcpyAsync(x, x_io, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    TATVA_TRY_HIP(hipMemcpyAsync(b, b_in, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  }
  // (full-length: the reduced diagonal is copied as is and read through the map)
  TATVA_TRY_HIP(hipMemcpyAsync(diag, diag_in, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipMemcpyAsync(u, u_in, n_full * sizeof(float), hipMemcpyDeviceToDevice, s));
  const int gv = step_grid(nv, step_prefetch<float>());
  const bool profiling = g_profiling.load() != 0;
  // r = b - H x, z = D^-1 r, shares of r.z, r.r, b.b (w is scratch here)
  if (geo) {
    rc = launch_geometry_cache_f32(m, geo, s);
    if (rc) return rc;
  }
  if (lin) {
    rc = launch_linearize_f32(m, d, u, lin, s, geo);
    if (rc) return rc;
  }
  rc = matfree_pcg_setup(m, d, u, op_map, nv, diag, b, x, r, z, w, dinv, P.rz(1), P.rr(1), P.bb,
                         gv, tol, maxiter, st, s, lin, geo, full ? dof_map : nullptr, n);
  if (rc) return rc;
  const int g_op = hvp_gather_blocks(m);
  // deferred p and x (option, on by default): iteration i's z sits in zb[i % 2];
  // the step of iteration i writes z_{i+1} to the other buffer and leaves
  // p_i = z_i + beta p_{i-1}, x_{i+1} = x_i + alpha p_i to the tile kernel of
  // iteration i + 1, which reads z_i from that buffer before the step of
  // iteration i + 1 overwrites it.  The solve is the same bits either way.
  const bool defer = g_cg_defer.load() == 1;
  float *const zb[2] = {z, defer ? z2 : z};
  // KA: w = H z with the shares of z.w; the tiles run in alternating order
  // (iteration parity), as the CSR SpMV streams its matrix
  const auto op = [&](int q, hipEvent_t between) -> int {
    const CgDefer dx{zb[q ^ 1], p, x, &st->defer[q ^ 1], defer ? nv : 0};
    return launch_hvp_dot_f32(m, d, u, zb[q], op_map, w, P.pap, s, &st->done, q, between, lin,
                              geo, &dx);
  };
  // the sampled operator also times its tile kernel alone (prof_a to prof_m)
  const std::function<int()> sampled_op = [&]() -> int { return op(0, res->prof_m); };
  const std::function<int(int, bool)> step = [&](int q, bool with_op) -> int {
    if (with_op) {
      const int e = op(q, nullptr);
      if (e) return e;
    }
    if (defer)
      k_cg_step<float, true><<<gv, kBlock, 0, s>>>(nv, nullptr, r, zb[q ^ 1], nullptr, sv, w, dinv,
                                                   P.pap, g_op, P.rz(q ^ 1), P.rr(q ^ 1), P.rz(q),
                                                   P.rr(q), q, st);
    else
      k_cg_step<float><<<gv, kBlock, 0, s>>>(nv, x, r, z, p, sv, w, dinv, P.pap, g_op, P.rz(q ^ 1),
                                             P.rr(q ^ 1), P.rz(q), P.rr(q), q, st);
    return TATVA_OK;
  };
  GraphKey key{};
  key.ws = workspace;
  key.a = m;
  key.b = op_map;
  key.n = nv;
  key.m = n_full;
  key.i0 = d->kind;
  key.i1 = chunk;
  key.i2 = profiling ? 1 : 0;
  key.i3 = (lin ? 1 : 0) | (geo ? 2 : 0) | (defer ? 4 : 0);
  key.c = d->elem_params;
  key.f0 = d->p0;
  key.f1 = d->p1;
  key.f2 = d->p2;
  rc = chunk_graph(s, chunk, step, profiling, key, res->matfree);
  if (rc == TATVA_OK)
    rc = run_chunks(res, res->matfree.exec, st, maxiter, chunk, profiling, sampled_op, stats,
                    std::function<int()>(), true);
  if (rc == TATVA_OK) {
    if (full) {
      k_compress_f32<<<grid_for(nv), kBlock, 0, s>>>(nv, dof_map, n, x, x_io);
      TATVA_CHECK_LAUNCH();
    } else {
      TATVA_TRY_HIP(hipMemcpyAsync(x_io, x, n * sizeof(float), hipMemcpyDeviceToDevice, s));
    }
  }
  TATVA_TRY_HIP(hipEventRecord(res->ev_out, s));
  TATVA_TRY_HIP(hipStreamWaitEvent(caller, res->ev_out, 0));
  return rc;
}

// ------------------------------------- matrix-free PCG through a lift map ----
//
// Constraints that tie several full dofs to one reduced dof (Periodic) make the
// reduced operator P^T H P with P the linear part of the lift: the reduced
// vector is spread to full dofs through the lift's read map, H runs on the
// full dofs, and each reduced entry sums the rows that read it.

namespace tatva {

// out[j] = in[map[j]], 0 where map[j] < 0 (constraint values)
__global__ __launch_bounds__(kBlock) void k_lift_map_f32(int64_t n, const int32_t *__restrict__ map,
                                                         const float *__restrict__ in,
                                                         float *__restrict__ out,
                                                         const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int32_t r = map[i];
    out[i] = r >= 0 ? in[r] : 0.0f;
  }
}

// w[j] = sum of y[list[k]] over k in [ptr[j], ptr[j+1]) in ascending k; with
// part, the block's share of z.w (FP32 products summed in FP64).
__global__ __launch_bounds__(kBlock) void k_reduce_rows_dot_f32(int64_t n, const int32_t *__restrict__ ptr,
                                                                const int32_t *__restrict__ list,
                                                                const float *__restrict__ y,
                                                                const float *__restrict__ z,
                                                                float *__restrict__ w,
                                                                double *__restrict__ part,
                                                                const int *__restrict__ skip) {
  if (skip && *skip) return;
  double dot = 0.0;
  for (int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; j < n;
       j += (int64_t)gridDim.x * blockDim.x) {
    float sum = 0.0f;
    for (int32_t k = ptr[j]; k < ptr[j + 1]; ++k) sum += y[list[k]];
    w[j] = sum;
    dot += (double)z[j] * (double)sum;
  }
  if (part) {
    dot = block_sum(dot);
    if (threadIdx.x == 0) part[blockIdx.x] = dot;
  }
}

}  // namespace tatva

extern "C" int64_t tatva_pcg_matfree_mapped_workspace(tatva_mesh *m, int64_t n) {
  if (!m || n <= 0) return -1;
  const int64_t n_full = m->n_nodes * m->dim;
  return (int64_t)(matfree_u_offset(n) + 3 * align256((size_t)n_full * sizeof(float)));
}

extern "C" int tatva_pcg_matfree_f32_mapped(tatva_mesh *m, const tatva_density *d, const float *u_in,
                                            const int32_t *read_map, const int32_t *rows_ptr,
                                            const int32_t *rows_list, int64_t n,
                                            const float *diag_in, const float *b_in, float *x_io,
                                            double tol, int32_t maxiter, int32_t chunk,
                                            void *workspace, int64_t ws_bytes,
                                            tatva_pcg_stats *stats, void *stream) {
  TATVA_REQUIRE(m && d && u_in && read_map && rows_ptr && rows_list && diag_in && b_in && x_io &&
                    n > 0,
                "invalid arguments");
  TATVA_REQUIRE(chunk >= 1 && chunk <= 4096, "chunk must be in [1, 4096]");
  const WsLayout L = ws_layout(n, kMatfreeVecs, sizeof(float));
  TATVA_REQUIRE(((uintptr_t)workspace & 255) == 0, "PCG workspace must be 256-byte aligned");
  TATVA_REQUIRE(workspace && ws_bytes >= tatva_pcg_matfree_mapped_workspace(m, n),
                "PCG workspace too small");
  const int64_t n_full = m->n_nodes * density_ncomp(m, d);
  chunk += chunk & 1;
  TATVA_TRY_HIP(hipSetDevice(m->device));
  DeviceRes *res = nullptr;
  int rc = get_res(m->device, &res);
  if (rc) return rc;
  std::lock_guard<std::mutex> lock(res->mu);
  char *wsb = (char *)workspace;
  float *z = ws_vec<float>(wsb, L, 0);
  float *w = ws_vec<float>(wsb, L, 1);
  float *p = ws_vec<float>(wsb, L, 2);
  float *sv = ws_vec<float>(wsb, L, 3);
  float *r = ws_vec<float>(wsb, L, 4);
  float *dinv = ws_vec<float>(wsb, L, 5);
  float *x = ws_vec<float>(wsb, L, 6);
  float *b = ws_vec<float>(wsb, L, 7);
  float *diag = ws_vec<float>(wsb, L, 8);
  const size_t full_bytes = align256((size_t)m->n_nodes * m->dim * sizeof(float));
  float *u = reinterpret_cast<float *>(wsb + matfree_u_offset(n));
  float *v_full = reinterpret_cast<float *>(wsb + matfree_u_offset(n) + full_bytes);
  float *y_full = reinterpret_cast<float *>(wsb + matfree_u_offset(n) + 2 * full_bytes);
  const PcgParts P = parts_at(wsb, L);
  PcgState *st = ws_state(wsb, L);
  hipStream_t caller = (hipStream_t)stream;
  hipStream_t s = res->cap;
  TATVA_TRY_HIP(hipEventRecord(res->ev_in, caller));
  TATVA_TRY_HIP(hipStreamWaitEvent(s, res->ev_in, 0));
  TATVA_TRY_HIP(hipMemcpyAsync(x, x_io, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipMemcpyAsync(b, b_in, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipMemcpyAsync(diag, diag_in, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipMemcpyAsync(u, u_in, n_full * sizeof(float), hipMemcpyDeviceToDevice, s));
  const int gv = step_grid(n, step_prefetch<float>());
  const int g_red = vec_grid(n);
  const bool profiling = g_profiling.load() != 0;
  // y = P^T H(u) P v, with the shares of v.y when part is given
  const auto mapped_op = [&](const float *v, float *y, double *part, const int *skip) -> int {
    k_lift_map_f32<<<grid_for(n_full), kBlock, 0, s>>>(n_full, read_map, v, v_full, skip);
    const int e = launch_hvp_f32(m, d, u, v_full, nullptr, y_full, s, skip);
    if (e) return e;
    k_reduce_rows_dot_f32<<<g_red, kBlock, 0, s>>>(n, rows_ptr, rows_list, y_full, v, y, part, skip);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  };
  // r = b - H x, z = D^-1 r, shares of r.z, r.r, b.b (w is scratch here)
  k_state_reset<<<1, 64, 0, s>>>(st);
  k_inv_diag_f32<<<grid_for(n), kBlock, 0, s>>>(n, diag, dinv, st);
  rc = mapped_op(x, w, nullptr, nullptr);
  if (rc) return rc;
  k_sub<float><<<grid_for(n), kBlock, 0, s>>>(n, b, w, r);
  k_pcg_init<float><<<gv, kBlock, 0, s>>>(n, r, dinv, b, z, P.rz(1), P.rr(1), P.bb);
  k_pcg_init_state<<<1, kBlock, 0, s>>>(P.rr(1), P.bb, gv, tol, maxiter, st);
  TATVA_CHECK_LAUNCH();
  // KA: w = P^T H P z with the shares of z.w
  const std::function<int()> apply_op = [&]() -> int { return mapped_op(z, w, P.pap, &st->done); };
  const std::function<int(int, bool)> step = [&](int q, bool with_op) -> int {
    if (with_op) {
      const int e = apply_op();
      if (e) return e;
    }
    k_cg_step<float><<<gv, kBlock, 0, s>>>(n, x, r, z, p, sv, w, dinv, P.pap, g_red, P.rz(q ^ 1),
                                           P.rr(q ^ 1), P.rz(q), P.rr(q), q, st);
    return TATVA_OK;
  };
  GraphKey key{};
  key.ws = workspace;
  key.a = m;
  key.b = read_map;
  key.c = rows_ptr;
  key.d = rows_list;
  key.e = d->elem_params;
  key.n = n;
  key.m = n_full;
  key.i0 = d->kind;
  key.i1 = chunk;
  key.i2 = profiling ? 1 : 0;
  key.f0 = d->p0;
  key.f1 = d->p1;
  key.f2 = d->p2;
  rc = chunk_graph(s, chunk, step, profiling, key, res->matfree);
  if (rc == TATVA_OK)
    rc = run_chunks(res, res->matfree.exec, st, maxiter, chunk, profiling, apply_op, stats);
  if (rc == TATVA_OK)
    TATVA_TRY_HIP(hipMemcpyAsync(x_io, x, n * sizeof(float), hipMemcpyDeviceToDevice, s));
  TATVA_TRY_HIP(hipEventRecord(res->ev_out, s));
  TATVA_TRY_HIP(hipStreamWaitEvent(caller, res->ev_out, 0));
  return rc;
}

// ------------------------------------------------------- vector kernels ----

namespace tatva {

__global__ void k_scatter_free(int64_t n, const int32_t *__restrict__ idx, const double *__restrict__ in,
                               double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = in[i];
}

__global__ void k_gather_free(int64_t n, const int32_t *__restrict__ idx, const double *__restrict__ in,
                              double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[idx[i]];
}

__global__ void k_set_dofs(int64_t n, const int32_t *__restrict__ idx, const double *__restrict__ vals,
                           double value, double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[idx[i]] = vals ? vals[i] : value;
}

// out[d] = map[d] >= 0 ? in[map[d]] : 0 (linear part of a lift)
__global__ void k_lift_map(int64_t n, const int32_t *__restrict__ map, const double *__restrict__ in,
                           double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const int32_t r = map[i];
    out[i] = r >= 0 ? in[r] : 0.0;
  }
}

// out[j] = sum of in[list[k]] over k in [ptr[j], ptr[j+1]), ascending k
__global__ void k_reduce_rows(int64_t n, const int32_t *__restrict__ ptr,
                              const int32_t *__restrict__ list, const double *__restrict__ in,
                              double *__restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double s = 0.0;
  for (int32_t k = ptr[j]; k < ptr[j + 1]; ++k) s += in[list[k]];
  out[j] = s;
}

template <class T>
__global__ void k_axpy_free(int64_t n, const int32_t *__restrict__ idx, double alpha,
                            const T *__restrict__ du, double *__restrict__ u) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) u[idx[i]] += alpha * (double)du[i];
}

template <class A, class B>
__global__ void k_cast(int64_t n, double alpha, const A *__restrict__ x, B *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = B(alpha * (double)x[i]);
}

__global__ __launch_bounds__(kBlock) void k_dot_part(int64_t n, const double *__restrict__ a,
                                                     const double *__restrict__ b,
                                                     double *__restrict__ partials) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// rows[k] = row of entry k, col_colors[k] = colors[indices[k]]
// (compute_rows_cols, tatva/sparse/base.py:108-136).  One wavefront per row.
__global__ __launch_bounds__(kBlock) void k_rows_cols(int64_t n, const int32_t *__restrict__ indptr,
                                                      const int32_t *__restrict__ indices,
                                                      const int32_t *__restrict__ colors,
                                                      int32_t *__restrict__ rows,
                                                      int32_t *__restrict__ col_colors) {
  const int lane = threadIdx.x % kWave;
  const int64_t row = (int64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (row >= n) return;
  for (int32_t k = indptr[row] + lane; k < indptr[row + 1]; k += kWave) {
    rows[k] = (int32_t)row;
    col_colors[k] = colors[indices[k]];
  }
}

}  // namespace tatva

extern "C" int tatva_scatter_free(int64_t n, const int32_t *idx, const double *in, double *out,
                                  void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (idx && in && out)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_scatter_free<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, idx, in, out);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_set_dofs(int64_t n, const int32_t *idx, const double *vals, double value,
                              double *out, void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (idx && out)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_set_dofs<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, idx, vals, value, out);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_gather_free(int64_t n, const int32_t *idx, const double *in, double *out,
                                 void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (idx && in && out)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_gather_free<<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, idx, in, out);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_lift_map(int64_t n_full, const int32_t *map, const double *in_red,
                              double *out_full, void *stream) {
  TATVA_REQUIRE(n_full >= 0 && (n_full == 0 || (map && in_red && out_full)), "invalid arguments");
  if (n_full == 0) return TATVA_OK;
  k_lift_map<<<grid_for(n_full), kBlock, 0, (hipStream_t)stream>>>(n_full, map, in_red, out_full);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_reduce_rows(int64_t n_red, const int32_t *ptr, const int32_t *list,
                                 const double *in_full, double *out_red, void *stream) {
  TATVA_REQUIRE(n_red >= 0 && (n_red == 0 || (ptr && list && in_full && out_red)),
                "invalid arguments");
  if (n_red == 0) return TATVA_OK;
  k_reduce_rows<<<grid_for(n_red), kBlock, 0, (hipStream_t)stream>>>(n_red, ptr, list, in_full,
                                                                     out_red);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_axpy_free_f32(int64_t n, const int32_t *idx, double alpha, const float *du,
                                   double *u, void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (idx && du && u)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_axpy_free<float><<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, idx, alpha, du, u);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_axpy_free(int64_t n, const int32_t *idx, double alpha, const double *du,
                               double *u, void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (idx && du && u)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_axpy_free<double><<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, idx, alpha, du, u);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_cast_f64_f32(int64_t n, double alpha, const double *x, float *y, void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (x && y)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_cast<double, float><<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, alpha, x, y);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_cast_f32_f64(int64_t n, double alpha, const float *x, double *y, void *stream) {
  TATVA_REQUIRE(n >= 0 && (n == 0 || (x && y)), "invalid arguments");
  if (n == 0) return TATVA_OK;
  k_cast<float, double><<<grid_for(n), kBlock, 0, (hipStream_t)stream>>>(n, alpha, x, y);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int64_t tatva_dot_workspace(int64_t n) {
  if (n <= 0) return -1;
  return (int64_t)reduce_grid(n) * (int64_t)sizeof(double);
}

extern "C" int tatva_dot(int64_t n, const double *x, const double *y, double *out,
                         void *workspace, int64_t workspace_bytes, void *stream) {
  TATVA_REQUIRE(n > 0 && x && y && out, "invalid arguments");
  TATVA_REQUIRE(workspace && workspace_bytes >= tatva_dot_workspace(n), "dot workspace too small");
  hipStream_t s = (hipStream_t)stream;
  const int g = reduce_grid(n);
  double *partials = static_cast<double *>(workspace);
  k_dot_part<<<g, kBlock, 0, s>>>(n, x, y, partials);
  k_finalize_sum<double><<<1, kBlock, 0, s>>>(partials, g, out, 0);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

// Scratch of tatva_dot_noalloc: one partial sum per block of the reduction
// grid, which reduce_grid caps at kMaxReduceBlocks (16 KiB), on every device the
// library runs on, held until the process ends.  Host code never passes the
// array's name to a kernel or copy (in host code it is a stand-in, not a device
// address): each device's address is read once with hipGetSymbolAddress, with
// that device current, and used from then on.
static __device__ __align__(256) double scratch[kMaxReduceBlocks];
static double *g_scratch_dev[64];
static std::mutex g_scratch_mu;

static int scratch_on(int device, double **out) {
  TATVA_REQUIRE(device >= 0 && device < 64, "device id out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[device]) {
    void *p = nullptr;
    TATVA_TRY_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_dev[device] = static_cast<double *>(p);
  }
  *out = g_scratch_dev[device];
  return TATVA_OK;
}

extern "C" int tatva_dot_noalloc(int64_t n, const double *x, const double *y, double *out,
                                 void *stream) {
  TATVA_REQUIRE(n > 0 && x && y && out, "invalid arguments");
  const int g = reduce_grid(n);
  TATVA_REQUIRE(g <= kMaxReduceBlocks, "the reduction needs more partial sums than the "
                                       "no-allocation scratch holds");
  int device = 0;
  TATVA_TRY_HIP(hipGetDevice(&device));
  double *partials = nullptr;
  const int rc = scratch_on(device, &partials);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  // every block writes its partial before the sum reads them: nothing to clear
  k_dot_part<<<g, kBlock, 0, s>>>(n, x, y, partials);
  k_finalize_sum<double><<<1, kBlock, 0, s>>>(partials, g, out, 0);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

// ---------------------------------------------------------------- BiCGSTAB --
//
// Jacobi-preconditioned BiCGSTAB (H. A. van der Vorst, SIAM J. Sci. Stat.
// Comput. 13, 1992, right preconditioning) for the non-symmetric systems of the
// path: the reduced mass matrix E_free^T M P of an L2 projection under Periodic
// constraints.  Host-driven: the recurrence scalars are read back once per
// reduction (these solves are small and set-up only).

__global__ __launch_bounds__(kBlock) void k_lin3(int64_t n, double a, const double *__restrict__ x,
                                                 double b, const double *__restrict__ y, double c,
                                                 const double *__restrict__ z,
                                                 double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double v = a * x[i] + b * y[i];
  if (z) v += c * z[i];
  out[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_scale(int64_t n, const double *__restrict__ d,
                                                  const double *__restrict__ x,
                                                  double *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = d[i] * x[i];
}

// dinv[i] = 1 / A_ii (1 where the row has no diagonal entry or it is zero)
__global__ __launch_bounds__(kBlock) void k_diag_inv_plain(int64_t n, const int32_t *__restrict__ indptr,
                                                           const int32_t *__restrict__ indices,
                                                           const double *__restrict__ data,
                                                           double *__restrict__ dinv) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int32_t k = indptr[i]; k < indptr[i + 1]; ++k)
    if (indices[k] == i) d = data[k];
  dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
}

constexpr int kBicgVecs = 9;  // r, rh, p, v, s, t, ph, sh, dinv

extern "C" int64_t tatva_bicgstab_csr_workspace(int64_t n) {
  if (n <= 0) return -1;
  return (int64_t)(align256((size_t)kBicgVecs * n * sizeof(double)) +
                   align256((kMaxReduceBlocks + 8) * sizeof(double)));
}

extern "C" int tatva_bicgstab_csr(int64_t n, int64_t nnz, const int32_t *indptr,
                                  const int32_t *indices, const double *data, const double *b,
                                  double *x, double tol, int32_t maxiter, void *workspace,
                                  int64_t ws_bytes, tatva_pcg_stats *stats, void *stream) {
  TATVA_REQUIRE(n > 0 && nnz > 0 && indptr && indices && data && b && x && maxiter >= 0,
                "invalid arguments");
  TATVA_REQUIRE(tol >= 0.0, "tol must be non-negative");
  TATVA_REQUIRE(workspace && ws_bytes >= tatva_bicgstab_csr_workspace(n),
                "BiCGSTAB workspace too small");
  hipStream_t st = (hipStream_t)stream;
  double *w = static_cast<double *>(workspace);
  double *r = w, *rh = w + n, *p = w + 2 * n, *v = w + 3 * n, *sv = w + 4 * n, *t = w + 5 * n;
  double *ph = w + 6 * n, *sh = w + 7 * n, *dinv = w + 8 * n;
  double *parts = reinterpret_cast<double *>(static_cast<char *>(workspace) +
                                             align256((size_t)kBicgVecs * n * sizeof(double)));
  double *dots = parts + kMaxReduceBlocks;
  const int g = grid_for(n), gr = reduce_grid(n);
  const int lpr = choose_lpr(n, nnz);
  auto spmv = [&](const double *xin, const double *bin, double *yout) -> int {
    return with_lpr(lpr, [&](auto lc) -> int {
      constexpr int LL = decltype(lc)::value;
      k_spmv<LL><<<spmv_grid(n, LL), kBlock, 0, st>>>(n, indptr, indices, data, xin, bin, yout);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    });
  };
  // dots[slot] = a . b on the stream
  auto dot = [&](const double *a, const double *bb, int slot) {
    k_dot_part<<<gr, kBlock, 0, st>>>(n, a, bb, parts);
    k_finalize_sum<double><<<1, kBlock, 0, st>>>(parts, gr, dots + slot, 0);
  };
  double h[8];
  auto read = [&](int count) -> int {
    TATVA_CHECK_LAUNCH();
    TATVA_TRY_HIP(hipMemcpyAsync(h, dots, count * sizeof(double), hipMemcpyDeviceToHost, st));
    TATVA_TRY_HIP(hipStreamSynchronize(st));
    return TATVA_OK;
  };
  k_diag_inv_plain<<<g, kBlock, 0, st>>>(n, indptr, indices, data, dinv);
  int rc = spmv(x, b, r);  // r = b - A x
  if (rc) return rc;
  TATVA_TRY_HIP(hipMemcpyAsync(rh, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
  dot(b, b, 0);
  dot(r, r, 1);
  if ((rc = read(2))) return rc;
  const double bb = h[0], tol2 = tol * tol * bb;
  double rr = h[1], rho_old = 1.0, alpha = 1.0, omega = 1.0;
  int status = TATVA_ENOTCONV, it = 0;
  if (bb == 0.0 || rr <= tol2) status = TATVA_OK;
  while (status == TATVA_ENOTCONV && it < maxiter) {
    dot(rh, r, 0);
    if ((rc = read(1))) return rc;
    const double rho = h[0];
    if (!(std::fabs(rho) > 0.0) || !std::isfinite(rho)) {
      status = TATVA_EBREAKDOWN;
      break;
    }
    if (it == 0) {
      TATVA_TRY_HIP(hipMemcpyAsync(p, r, n * sizeof(double), hipMemcpyDeviceToDevice, st));
    } else {
      const double beta = (rho / rho_old) * (alpha / omega);
      k_lin3<<<g, kBlock, 0, st>>>(n, 1.0, r, beta, p, -beta * omega, v, p);
    }
    k_scale<<<g, kBlock, 0, st>>>(n, dinv, p, ph);
    if ((rc = spmv(ph, nullptr, v))) return rc;
    dot(rh, v, 0);
    if ((rc = read(1))) return rc;
    if (!(std::fabs(h[0]) > 0.0) || !std::isfinite(h[0])) {
      status = TATVA_EBREAKDOWN;
      break;
    }
    alpha = rho / h[0];
    k_lin3<<<g, kBlock, 0, st>>>(n, 1.0, r, -alpha, v, 0.0, nullptr, sv);
    dot(sv, sv, 0);
    if ((rc = read(1))) return rc;
    ++it;
    if (h[0] <= tol2) {  // x += alpha ph
      k_lin3<<<g, kBlock, 0, st>>>(n, 1.0, x, alpha, ph, 0.0, nullptr, x);
      rr = h[0];
      status = TATVA_OK;
      break;
    }
    k_scale<<<g, kBlock, 0, st>>>(n, dinv, sv, sh);
    if ((rc = spmv(sh, nullptr, t))) return rc;
    dot(t, sv, 0);
    dot(t, t, 1);
    if ((rc = read(2))) return rc;
    if (!(h[1] > 0.0) || !std::isfinite(h[0])) {
      status = TATVA_EBREAKDOWN;
      break;
    }
    omega = h[0] / h[1];
    k_lin3<<<g, kBlock, 0, st>>>(n, 1.0, x, alpha, ph, omega, sh, x);
    k_lin3<<<g, kBlock, 0, st>>>(n, 1.0, sv, -omega, t, 0.0, nullptr, r);
    dot(r, r, 0);
    if ((rc = read(1))) return rc;
    rr = h[0];
    if (rr <= tol2) status = TATVA_OK;
    else if (!(std::fabs(omega) > 0.0)) status = TATVA_EBREAKDOWN;
    rho_old = rho;
  }
  TATVA_CHECK_LAUNCH();
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->iterations = it;
    stats->status = status;
    stats->b_norm = std::sqrt(bb);
    stats->rel_residual = bb > 0 ? std::sqrt(rr / bb) : 0.0;
  }
  return TATVA_OK;
}

extern "C" int tatva_rows_cols(int64_t n, const int32_t *indptr, const int32_t *indices,
                               const int32_t *colors, int32_t *rows, int32_t *col_colors,
                               void *stream) {
  TATVA_REQUIRE(n >= 0 && indptr && indices && colors && rows && col_colors, "invalid arguments");
  if (n == 0) return TATVA_OK;
  const int64_t blocks = (n + (kBlock / kWave) - 1) / (kBlock / kWave);
  TATVA_REQUIRE(blocks < (int64_t(1) << 31), "too many rows");
  k_rows_cols<<<(int)blocks, kBlock, 0, (hipStream_t)stream>>>(n, indptr, indices, colors, rows,
                                                              col_colors);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}
