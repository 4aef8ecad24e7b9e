// Element-level kernels on gfx950: quadrature-point gather/eval/grad, weights,
// integration, energy, residual (node-gather assembly), Hessian-vector product,
// element stiffness and the colored-Hessian decompression into CSR.
//
// Threads map to elements (or element columns); FP64 geometry is recomputed
// from the gathered coordinates at every quadrature point (J, signed detJ,
// inv(J) in registers) instead of being stored, so a pass reads only the
// connectivity, the node coordinates and the nodal vectors.
#include <hip/hip_runtime.h>

#include <sched.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <new>
#include <thread>
#include <type_traits>
#include <vector>

#include "common.hpp"
#include "density.hpp"
#include "element.hpp"
#include "mesh.hpp"

namespace tatva {

template <class F>
static int with_kind(int kind, F &&f) {
  switch (kind) {
    case TATVA_TRI3: return f(std::integral_constant<int, TATVA_TRI3>{});
    case TATVA_TET4: return f(std::integral_constant<int, TATVA_TET4>{});
    case TATVA_HEX8: return f(std::integral_constant<int, TATVA_HEX8>{});
    default: set_error("unknown element kind"); return TATVA_EINVAL;
  }
}

template <int D, class T, class F>
static int with_density(const tatva_density *d, F &&f) {
  if (!d) {
    set_error("density descriptor is NULL");
    return TATVA_EINVAL;
  }
  switch (d->kind) {
    case TATVA_DENSITY_POISSON: return f(Poisson<D, T>{T(d->p0), d->elem_params});
    case TATVA_DENSITY_LINEAR_ELASTIC:
      return f(LinearElastic<D, T>{T(d->p0), T(d->p1), d->elem_params});
    case TATVA_DENSITY_NEO_HOOKEAN: return f(NeoHookean<D, T>{T(d->p0), T(d->p1), d->elem_params});
    default: set_error("unknown density kind"); return TATVA_EINVAL;
  }
}

int density_ncomp(const tatva_mesh *m, const tatva_density *d) {
  return d->kind == TATVA_DENSITY_POISSON ? 1 : m->dim;
}

static int set_device(const tatva_mesh *m) {
  TATVA_TRY_HIP(hipSetDevice(m->device));
  return TATVA_OK;
}

// ------------------------------------------------------------ weights/eval --

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_weights(int64_t n_elem, const double *__restrict__ coords,
                                                    const int32_t *__restrict__ conn, QuadRule quad,
                                                    double *__restrict__ out) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  for (int q = 0; q < quad.nq; ++q) out[e * quad.nq + q] = qp_detJ<KIND>(X, quad.xi[q]) * quad.w[q];
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_eval(int64_t n_elem, const int32_t *__restrict__ conn,
                                                 QuadRule quad, int nc, const double *__restrict__ u,
                                                 double *__restrict__ out) {
  using El = Elem<KIND>;
  constexpr int NNE = El::NNE;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
#pragma unroll
  for (int a = 0; a < NNE; ++a) nodes[a] = conn[e * NNE + a];
  for (int q = 0; q < quad.nq; ++q) {
    double N[NNE];
    El::N(quad.xi[q], N);
    for (int c = 0; c < nc; ++c) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NNE; ++a) s += N[a] * u[(int64_t)nodes[a] * nc + c];
      out[(e * quad.nq + q) * nc + c] = s;
    }
  }
}

template <int KIND>
__global__ __launch_bounds__(kBlock) void k_grad(int64_t n_elem, const double *__restrict__ coords,
                                                 const int32_t *__restrict__ conn, QuadRule quad,
                                                 int nc, const double *__restrict__ u,
                                                 double *__restrict__ out) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  for (int q = 0; q < quad.nq; ++q) {
    double B[D][NNE];
    qp_geometry<KIND>(X, quad.xi[q], B);
    for (int c = 0; c < nc; ++c) {
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NNE; ++a) s += B[j][a] * u[(int64_t)nodes[a] * nc + c];
        out[((e * quad.nq + q) * nc + c) * D + j] = s;
      }
    }
  }
}

// per_elem[e, c] = sum_q vals[e, q, c] * detJ * w_q, vals given at quadrature
// points or (nodal != 0) interpolated from nodal values.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_integrate_elem(int64_t n_elem, const double *__restrict__ coords,
                                                           const int32_t *__restrict__ conn,
                                                           QuadRule quad, int nc, int nodal,
                                                           const double *__restrict__ vals,
                                                           double *__restrict__ per_elem) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  for (int c = 0; c < nc; ++c) per_elem[e * nc + c] = 0.0;
  for (int q = 0; q < quad.nq; ++q) {
    const double w = qp_detJ<KIND>(X, quad.xi[q]) * quad.w[q];
    double N[NNE];
    El::N(quad.xi[q], N);
    for (int c = 0; c < nc; ++c) {
      double v;
      if (nodal) {
        v = 0.0;
#pragma unroll
        for (int a = 0; a < NNE; ++a) v += N[a] * vals[(int64_t)nodes[a] * nc + c];
      } else {
        v = vals[(e * quad.nq + q) * nc + c];
      }
      per_elem[e * nc + c] += v * w;
    }
  }
}

// partials[block] = block share of sum_e sum_q vals[e, q, c] * detJ * w_q for
// one component c (grid-stride over elements; no per-element buffer).
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_integrate_comp(int64_t n_elem, const double *__restrict__ coords,
                                                           const int32_t *__restrict__ conn,
                                                           QuadRule quad, int nc, int c, int nodal,
                                                           const double *__restrict__ vals,
                                                           double *__restrict__ partials) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE;
  double s = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_elem;
       e += (int64_t)gridDim.x * blockDim.x) {
    int32_t nodes[NNE];
    double X[NNE][D];
    load_element<KIND>(e, conn, coords, nodes, X);
    for (int q = 0; q < quad.nq; ++q) {
      const double w = qp_detJ<KIND>(X, quad.xi[q]) * quad.w[q];
      double v;
      if (nodal) {
        double N[NNE];
        El::N(quad.xi[q], N);
        v = 0.0;
#pragma unroll
        for (int a = 0; a < NNE; ++a) v += N[a] * vals[(int64_t)nodes[a] * nc + c];
      } else {
        v = vals[(e * quad.nq + q) * nc + c];
      }
      s += v * w;
    }
  }
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// partials[block] = sum over the block's grid-stride share of x[i*stride + offset];
// k_finalize_sum then adds the partials in a fixed order.
__global__ __launch_bounds__(kBlock) void k_strided_sum(int64_t n, const double *__restrict__ x,
                                                        int stride, int offset,
                                                        double *__restrict__ partials) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += x[i * stride + offset];
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// partials[block] = block share of sum_i a[i] * b[i]
__global__ __launch_bounds__(kBlock) void k_dot_partials(int64_t n, const double *__restrict__ a,
                                                         const double *__restrict__ b,
                                                         double *__restrict__ partials) {
  double s = 0.0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    s += a[i] * b[i];
  s = block_sum(s);
  if (threadIdx.x == 0) partials[blockIdx.x] = s;
}

// ----------------------------------------------------------- densities -----

template <int KIND, class Dens>
__global__ __launch_bounds__(kBlock) void k_density_eval(int64_t n_elem, int nq, Dens dens,
                                                         const double *__restrict__ grad,
                                                         const double *__restrict__ values,
                                                         double *__restrict__ psi) {
  constexpr int D = Elem<KIND>::D, NC = Dens::NC;
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_elem * nq) return;
  double G[NC][D];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int j = 0; j < D; ++j) G[c][j] = grad[(i * NC + c) * D + j];
  const double uq = values ? values[i] : 0.0;
  psi[i] = dens.at(i / nq).psi(G, uq);
}

// Sum over elements of sum_q psi(G_q) detJ w_q (+ Poisson source via u_q).
template <int KIND, class Dens>
__global__ __launch_bounds__(kBlock) void k_energy(int64_t n_elem, const double *__restrict__ coords,
                                                   const int32_t *__restrict__ conn, QuadRule quad,
                                                   Dens dens, const double *__restrict__ u,
                                                   double *__restrict__ partials) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
  double acc = 0.0;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < n_elem;
       e += (int64_t)gridDim.x * blockDim.x) {
    int32_t nodes[NNE];
    double X[NNE][D];
    load_element<KIND>(e, conn, coords, nodes, X);
    const Dens de = dens.at(e);
    double ue[NNE][NC];
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) ue[a][c] = u[(int64_t)nodes[a] * NC + c];
    double ee = 0.0;
    for (int q = 0; q < quad.nq; ++q) {
      double B[D][NNE];
      const double det = qp_geometry<KIND>(X, quad.xi[q], B);
      double G[NC][D];
#pragma unroll
      for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int j = 0; j < D; ++j) {
          double s = 0.0;
#pragma unroll
          for (int a = 0; a < NNE; ++a) s += B[j][a] * ue[a][c];
          G[c][j] = s;
        }
      double uq = 0.0;
      if constexpr (Dens::kSource) {
        double N[NNE];
        El::N(quad.xi[q], N);
#pragma unroll
        for (int a = 0; a < NNE; ++a) uq += N[a] * ue[a][0];
      }
      ee += de.psi(G, uq) * (det * quad.w[q]);
    }
    acc += ee;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Element stiffness, one thread per (element, column j = b*NC + k):
// column j of K_e is the element JVP of the residual with the unit tangent
// on local dof j.  Layout: Ke[e][row][col], row/col = a*NC + c.
template <int KIND, class Dens>
__global__ __launch_bounds__(kBlock) void k_elem_stiffness(int64_t n_elem, const double *__restrict__ coords,
                                                           const int32_t *__restrict__ conn,
                                                           QuadRule quad, Dens dens,
                                                           const double *__restrict__ u,
                                                           double *__restrict__ Ke) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC, ND = NNE * NC;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_elem * ND) return;
  const int64_t e = t / ND;
  const int col = (int)(t - e * ND);
  const int b = col / NC, k = col % NC;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  const Dens de = dens.at(e);
  double ue[NNE][NC], ke[NNE][NC];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      ue[a][c] = u[(int64_t)nodes[a] * NC + c];
      ke[a][c] = 0.0;
    }
  for (int q = 0; q < quad.nq; ++q) {
    double B[D][NNE];
    const double det = qp_geometry<KIND>(X, quad.xi[q], B);
    const double w = det * quad.w[q];
    double G[NC][D], dG[NC][D], dP[NC][D];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NNE; ++a) s += B[j][a] * ue[a][c];
        G[c][j] = s;
        double bj = 0.0;
#pragma unroll
        for (int a = 0; a < NNE; ++a) bj = (a == b) ? B[j][a] : bj;
        dG[c][j] = (c == k) ? bj : 0.0;
      }
    const auto L = de.linearize(G);
    de.dP(L, dG, dP);
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += dP[c][j] * B[j][a];
        ke[a][c] += w * s;
      }
  }
  double *dst = Ke + e * ND * ND + col;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) dst[(a * NC + c) * ND] = ke[a][c];
}

// Colored Hessian decompression, one wavefront per matrix row.
// The row's compressed Jacobian J[row, 0..n_colors) lives in LDS:
//   J[row, colors[s]] = sum over elements e containing the row's node (ascending
//   e) and element dofs s of K_e[row, s]   (the per-color JVP, sparse/base.py:257-270)
// then data[k] = J[row, colors[indices[k]]]  (sparse/base.py:173).
// Within one element all dof colors differ (distance-2 coloring), so lanes of
// one step never write the same LDS slot.
template <int NNE, int NC>
__global__ __launch_bounds__(kBlock) void k_colored_rows(
    int64_t n_rows, const int32_t *__restrict__ red_to_full, const int32_t *__restrict__ dof_map,
    const int32_t *__restrict__ colors, int n_colors, const int32_t *__restrict__ indptr,
    const int32_t *__restrict__ indices, const int32_t *__restrict__ inv_ptr,
    const int32_t *__restrict__ inv_list, const int32_t *__restrict__ conn,
    const double *__restrict__ Ke, double *__restrict__ data) {
  constexpr int ND = NNE * NC;
  extern __shared__ double lds[];
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  const int64_t row = xcd_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + wid;
  if (row >= n_rows) return;
  double *J = lds + (int64_t)wid * n_colors;
  for (int c = lane; c < n_colors; c += kWave) J[c] = 0.0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  const int64_t R = red_to_full ? (int64_t)red_to_full[row] : row;
  const int64_t node = R / NC;
  const int ci = (int)(R - node * NC);
  const int32_t b0 = inv_ptr[node], b1 = inv_ptr[node + 1];
  for (int32_t p = b0; p < b1; ++p) {
    const int32_t packed = inv_list[p];
    const int64_t e = packed >> 3;
    const int a = packed & 7;
    if (lane < ND) {
      const int bn = lane / NC, k = lane % NC;
      const int64_t S = (int64_t)conn[e * NNE + bn] * NC + k;
      const int32_t s = dof_map ? dof_map[S] : (int32_t)S;
      if (s >= 0) {
        const int c = colors[s];
        J[c] += Ke[(e * ND + a * NC + ci) * ND + lane];
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  const int32_t k0 = indptr[row], k1 = indptr[row + 1];
  for (int32_t k = k0 + lane; k < k1; k += kWave) data[k] = J[colors[indices[k]]];
}

// Reduced row r = sum of full rows row_list[row_ptr[r]..row_ptr[r+1]); columns
// through read_map.  serial: one lane accumulates an element's columns in turn
// (two element dofs may share a reduced column).
template <int NNE, int NC>
__global__ __launch_bounds__(kBlock) void k_colored_rows_mapped(
    int64_t n_rows, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ row_list,
    const int32_t *__restrict__ read_map, const int32_t *__restrict__ colors, int n_colors,
    const int32_t *__restrict__ indptr, const int32_t *__restrict__ indices,
    const int32_t *__restrict__ inv_ptr, const int32_t *__restrict__ inv_list,
    const int32_t *__restrict__ conn, const double *__restrict__ Ke, double *__restrict__ data,
    int serial) {
  constexpr int ND = NNE * NC;
  extern __shared__ double lds[];
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  const int64_t row = xcd_block(blockIdx.x, gridDim.x) * (kBlock / kWave) + wid;
  if (row >= n_rows) return;
  double *J = lds + (int64_t)wid * n_colors;
  for (int c = lane; c < n_colors; c += kWave) J[c] = 0.0;
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int32_t t = row_ptr[row]; t < row_ptr[row + 1]; ++t) {
    const int64_t R = row_list[t];
    const int64_t node = R / NC;
    const int ci = (int)(R - node * NC);
    const int32_t b0 = inv_ptr[node], b1 = inv_ptr[node + 1];
    for (int32_t p = b0; p < b1; ++p) {
      const int32_t packed = inv_list[p];
      const int64_t e = packed >> 3;
      const int a = packed & 7;
      const double *ke = Ke + (e * ND + a * NC + ci) * ND;
      if (serial) {
        if (lane == 0)
          for (int l = 0; l < ND; ++l) {
            const int32_t s = read_map[(int64_t)conn[e * NNE + l / NC] * NC + l % NC];
            if (s >= 0) J[colors[s]] += ke[l];
          }
      } else if (lane < ND) {
        const int32_t s = read_map[(int64_t)conn[e * NNE + lane / NC] * NC + lane % NC];
        if (s >= 0) J[colors[s]] += ke[lane];
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
  }
  const int32_t k0 = indptr[row], k1 = indptr[row + 1];
  for (int32_t k = k0 + lane; k < k1; k += kWave) data[k] = J[colors[indices[k]]];
}

// Element mass matrix of an NC-component field: Ke[(a,c),(b,c')] =
// delta_cc' sum_q N_a N_b detJ w (the Hessian of 1/2 |u|^2 integrated), in the
// layout of k_elem_stiffness.
template <int KIND, int NC>
__global__ __launch_bounds__(kBlock) void k_mass_stiffness(int64_t n_elem, const double *__restrict__ coords,
                                                           const int32_t *__restrict__ conn,
                                                           QuadRule quad, double *__restrict__ Ke) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, ND = NNE * NC;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  double M[NNE][NNE];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int b = 0; b < NNE; ++b) M[a][b] = 0.0;
  for (int q = 0; q < quad.nq; ++q) {
    const double w = qp_detJ<KIND>(X, quad.xi[q]) * quad.w[q];
    double N[NNE];
    El::N(quad.xi[q], N);
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int b = 0; b < NNE; ++b) M[a][b] += w * N[a] * N[b];
  }
  double *out = Ke + e * ND * ND;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int b = 0; b < NNE; ++b)
#pragma unroll
        for (int c2 = 0; c2 < NC; ++c2) out[(a * NC + c) * ND + b * NC + c2] = c == c2 ? M[a][b] : 0.0;
}

// out[i, c] = sum over the elements e of node i (local index a, ascending e)
// of sum_q N_a(xi_q) vals[e, q, c] detJ w_q.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_project_rhs(int64_t n_nodes, const double *__restrict__ coords,
                                                        const int32_t *__restrict__ conn,
                                                        const int32_t *__restrict__ inv_ptr,
                                                        const int32_t *__restrict__ inv_list,
                                                        QuadRule quad, int nc,
                                                        const double *__restrict__ vals,
                                                        double *__restrict__ out) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  for (int c = 0; c < nc; ++c) out[i * nc + c] = 0.0;
  for (int32_t p = inv_ptr[i]; p < inv_ptr[i + 1]; ++p) {
    const int64_t e = inv_list[p] >> 3;
    const int a = inv_list[p] & 7;
    int32_t nodes[NNE];
    double X[NNE][D];
    load_element<KIND>(e, conn, coords, nodes, X);
    for (int q = 0; q < quad.nq; ++q) {
      double N[NNE];
      El::N(quad.xi[q], N);
      const double w = qp_detJ<KIND>(X, quad.xi[q]) * quad.w[q] * N[a];
      for (int c = 0; c < nc; ++c) out[i * nc + c] += w * vals[(e * quad.nq + q) * nc + c];
    }
  }
}

// Operator.interpolate for 2-D elements (operator.py:399-468, mesh.py:295-390):
// the first element (lowest index) whose vertex polygon contains the point
// (bounding-box reject, points within 1e-8 of an edge count as inside, else
// ray-casting parity), one Newton step x(xi) = p from the first quadrature
// point, then the element interpolation of the nodal values.  Points in no
// element get NaN and are counted in *outside.
template <int KIND>
__global__ __launch_bounds__(kBlock) void k_interpolate(int64_t n_points, const double *__restrict__ pts,
                                                        PointGrid grid,
                                                        const int32_t *__restrict__ cell_ptr,
                                                        const int32_t *__restrict__ cell_elems,
                                                        const double *__restrict__ coords,
                                                        const int32_t *__restrict__ conn, QuadRule quad,
                                                        int nc, const double *__restrict__ nodal,
                                                        double *__restrict__ out,
                                                        int *__restrict__ outside) {
  using El = Elem<KIND>;
  constexpr int NNE = El::NNE;
  static_assert(El::D == 2, "polygon location is 2-D");
  const int64_t ip = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= n_points) return;
  const double px = pts[ip * 2], py = pts[ip * 2 + 1];
  int64_t found = -1;
  double X[NNE][2];
  int32_t nodes[NNE];
  // every element holding the point overlaps the point's cell, and the cell
  // lists its elements ascending: the first hit is the lowest element index,
  // as a scan over all elements would find
  const int64_t cell = (int64_t)grid_cell(py, grid.lo[1], grid.h[1], grid.n[1]) * grid.n[0] +
                       grid_cell(px, grid.lo[0], grid.h[0], grid.n[0]);
  const int32_t k1 = cell_ptr[cell + 1];
  for (int32_t k = cell_ptr[cell]; k < k1 && found < 0; ++k) {
    const int64_t e = cell_elems[k];
    double lo0 = 1e300, lo1 = 1e300, hi0 = -1e300, hi1 = -1e300;
#pragma unroll
    for (int a = 0; a < NNE; ++a) {
      nodes[a] = conn[e * NNE + a];
      X[a][0] = coords[(int64_t)nodes[a] * 2];
      X[a][1] = coords[(int64_t)nodes[a] * 2 + 1];
      lo0 = fmin(lo0, X[a][0]);
      hi0 = fmax(hi0, X[a][0]);
      lo1 = fmin(lo1, X[a][1]);
      hi1 = fmax(hi1, X[a][1]);
    }
    if (!(px >= lo0 && px <= hi0 && py >= lo1 && py <= hi1)) continue;
    bool boundary = false;
    int crossings = 0;
#pragma unroll
    for (int a = 0; a < NNE; ++a) {
      const double x1 = X[a][0], y1 = X[a][1];
      const double x2 = X[(a + 1) % NNE][0], y2 = X[(a + 1) % NNE][1];
      const double cross = (x2 - x1) * (py - y1) - (y2 - y1) * (px - x1);
      const bool on_seg = fmin(x1, x2) <= px && px <= fmax(x1, x2) && fmin(y1, y2) <= py &&
                          py <= fmax(y1, y2);
      if (fabs(cross) <= 1e-8 && on_seg) boundary = true;
      const bool ycond = (y1 <= py && y2 > py) || (y2 <= py && y1 > py);
      if (ycond) {
        const double xi = (x2 - x1) * (py - y1) / (y2 - y1) + x1;
        if (px < xi) ++crossings;
      }
    }
    if (boundary || (crossings & 1)) found = e;
  }
  if (found < 0) {
    for (int c = 0; c < nc; ++c) out[ip * nc + c] = __builtin_nan("");
    atomicAdd(outside, 1);
    return;
  }
  // one Newton step from the first quadrature point: J dxi = p - x(xi0)
  const double *xi0 = quad.xi[0];
  double N0[NNE], dn[2][NNE];
  El::N(xi0, N0);
  El::dN(xi0, dn);
  double x0[2] = {0.0, 0.0}, J[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      x0[i] += N0[a] * X[a][i];
#pragma unroll
      for (int k = 0; k < 2; ++k) J[i][k] += X[a][i] * dn[k][a];  // dx_i / dxi_k
    }
  double inv[2][2];
  det_inv(J, inv);
  const double r0 = px - x0[0], r1 = py - x0[1];
  double xi[3] = {xi0[0] + inv[0][0] * r0 + inv[0][1] * r1, xi0[1] + inv[1][0] * r0 + inv[1][1] * r1,
                  0.0};
  double N[NNE];
  El::N(xi, N);
  for (int c = 0; c < nc; ++c) {
    double v = 0.0;
#pragma unroll
    for (int a = 0; a < NNE; ++a) v += N[a] * nodal[(int64_t)nodes[a] * nc + c];
    out[ip * nc + c] = v;
  }
}

}  // namespace tatva

using namespace tatva;

// ------------------------------------------------------------ tiling -------

namespace {

struct HostTiles {
  int32_t n_tiles = 0, max_local = 0, max_colors = 0;
  std::vector<uint16_t> lnodes;     // n_tiles * kTile * nne
  std::vector<int16_t> color_off;   // n_tiles * (max_colors + 1)
  std::vector<int32_t> node_ptr;    // n_tiles + 1
  std::vector<int32_t> slot_ptr;    // n_nodes + 1
  std::vector<int32_t> slot_pos;    // slot -> node-major position (ascending tile per node)
  std::vector<int32_t> tnodes;      // global node of every slot
  std::vector<uint16_t> inc_end;    // (incidence lists only) n_slots
  std::vector<uint16_t> inc_list;   // (incidence lists only) n_tiles * kTile * nne
  std::vector<int32_t> elem;        // n_tiles * kTile: element of each tile slot, -1 padding
};

uint64_t spread3(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffff;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

// Host threads for the set-up passes: the CPUs this process may run on, at
// most 16, and no more than there are work items.
int host_threads(int64_t items) {
  int n = 1;
  cpu_set_t cs;
  if (sched_getaffinity(0, sizeof(cs), &cs) == 0) n = CPU_COUNT(&cs);
  n = std::min(n, 16);
  return (int)std::max<int64_t>(1, std::min<int64_t>(n, items));
}

// f(begin, end, worker) over [0, n) in `threads` contiguous ranges; false if
// a worker ran out of memory.
template <class F>
bool parallel_ranges(int64_t n, int threads, F &&f) {
  std::atomic<bool> ok{true};
  auto run = [&](int64_t b, int64_t e, int w) {
    try {
      f(b, e, w);
    } catch (const std::bad_alloc &) {
      ok = false;
    }
  };
  if (threads <= 1 || n < 2) {
    run(0, n, 0);
    return ok;
  }
  std::vector<std::thread> pool;
  for (int w = 0; w < threads; ++w) {
    const int64_t b = n * w / threads, e = n * (w + 1) / threads;
    if (b < e) pool.emplace_back(run, b, e, w);
  }
  for (auto &t : pool) t.join();
  return ok;
}

// Sorts (code, element) pairs: sorted ranges, then rounds of pairwise merges.
// The pairs are distinct, so the result equals one std::sort.
using MortonKey = std::pair<uint64_t, int32_t>;
bool sort_keys(std::vector<MortonKey> &v, int threads) {
  const int64_t n = (int64_t)v.size();
  if (threads <= 1 || n < (int64_t(1) << 16)) {
    std::sort(v.begin(), v.end());
    return true;
  }
  std::vector<int64_t> bnd((size_t)threads + 1);
  for (int i = 0; i <= threads; ++i) bnd[(size_t)i] = n * i / threads;
  if (!parallel_ranges(threads, threads, [&](int64_t b, int64_t e, int) {
        for (int64_t i = b; i < e; ++i) std::sort(v.begin() + bnd[(size_t)i], v.begin() + bnd[(size_t)i + 1]);
      }))
    return false;
  std::vector<MortonKey> tmp((size_t)n);
  std::vector<MortonKey> *src = &v, *dst = &tmp;
  for (int width = 1; width < threads; width *= 2) {
    const int64_t pairs = (threads + 2 * width - 1) / (2 * width);
    if (!parallel_ranges(pairs, host_threads(pairs), [&](int64_t b, int64_t e, int) {
          for (int64_t q = b; q < e; ++q) {
            const int lo = (int)(q * 2 * width);
            const int mid = std::min(threads, lo + width), hi = std::min(threads, lo + 2 * width);
            auto s0 = src->begin() + bnd[(size_t)lo], s1 = src->begin() + bnd[(size_t)mid];
            auto s2 = src->begin() + bnd[(size_t)hi];
            std::merge(s0, s1, s1, s2, dst->begin() + bnd[(size_t)lo]);
          }
        }))
      return false;
    std::swap(src, dst);
  }
  if (src != &v) v.swap(*src);
  return true;
}

// Z-order of element centroids, tiles of kTile consecutive elements in that
// order, local node numbering per tile (ascending global node), greedy
// element coloring inside each tile (elements of a color share no node),
// node -> slot lists.  Tiles are independent, so they are built by several
// host threads and concatenated in tile order (the result does not depend on
// the thread count).
int build_tiles(const double *coords, int dim, int64_t N, const int32_t *elements, int64_t E,
                int nne, bool incidence, HostTiles &T) {
  constexpr int kMaxColors = kMaxTileColors;  // the LDS copy of a tile's offsets holds this many
  const int64_t n_tiles = (E + kTile - 1) / kTile;
  if (n_tiles >= (int64_t(1) << 31)) {
    set_error("too many tiles");
    return TATVA_EINVAL;
  }
  const int th_e = host_threads(E / 65536 + 1);
  // centroids, their bounding box, Morton codes
  std::vector<double> cen((size_t)E * 3, 0.0);
  std::vector<std::array<double, 6>> box((size_t)th_e, {1e300, 1e300, 1e300, -1e300, -1e300, -1e300});
  bool ok = parallel_ranges(E, th_e, [&](int64_t b, int64_t e1, int w) {
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t e = b; e < e1; ++e)
      for (int j = 0; j < dim; ++j) {
        double s = 0.0;
        for (int a = 0; a < nne; ++a) s += coords[(int64_t)elements[e * nne + a] * dim + j];
        s /= nne;
        cen[(size_t)e * 3 + j] = s;
        lo[j] = std::min(lo[j], s);
        hi[j] = std::max(hi[j], s);
      }
    box[(size_t)w] = {lo[0], lo[1], lo[2], hi[0], hi[1], hi[2]};
  });
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (const auto &bx : box)
    for (int j = 0; j < 3; ++j) {
      lo[j] = std::min(lo[j], bx[(size_t)j]);
      hi[j] = std::max(hi[j], bx[(size_t)j + 3]);
    }
  std::vector<MortonKey> key((size_t)E);
  ok = ok && parallel_ranges(E, th_e, [&](int64_t b, int64_t e1, int) {
    for (int64_t e = b; e < e1; ++e) {
      uint64_t code = 0;
      for (int j = 0; j < dim; ++j) {
        const double span = hi[j] - lo[j];
        const double t = span > 0 ? (cen[(size_t)e * 3 + j] - lo[j]) / span : 0.0;
        const uint64_t qv = (uint64_t)std::min(2097151.0, std::max(0.0, t * 2097151.0));
        code |= spread3(qv) << j;
      }
      key[(size_t)e] = {code, (int32_t)e};
    }
  });
  ok = ok && sort_keys(key, th_e);
  if (!ok) {
    set_error("out of host memory while tiling the mesh");
    return TATVA_ENOMEM;
  }
  std::vector<double>().swap(cen);
  // per tile: local nodes, colors, color-ordered local connectivity
  T.n_tiles = (int32_t)n_tiles;
  T.lnodes.assign((size_t)n_tiles * kTile * nne, 0);
  T.elem.assign((size_t)n_tiles * kTile, -1);
  if (incidence) T.inc_list.assign((size_t)n_tiles * kTile * nne, 0);
  std::vector<std::vector<int32_t>> tloc((size_t)n_tiles);  // sorted local -> global node
  std::vector<std::vector<uint16_t>> tinc((size_t)n_tiles);  // incidence list ends per local node
  std::vector<int16_t> toff((size_t)n_tiles * (kMaxColors + 1), 0);
  std::vector<int32_t> tcol((size_t)n_tiles, 0);
  std::atomic<int> bad{0};  // 1: too many local nodes, 2: too many colors
  const int th_t = host_threads(n_tiles / 16 + 1);
  ok = parallel_ranges(n_tiles, th_t, [&](int64_t t0, int64_t t1, int) {
    std::vector<int32_t> mark((size_t)N, -1), lid((size_t)N);
    std::vector<uint64_t> used;
    std::vector<int> color(kTile);
    std::vector<std::array<uint16_t, 8>> ln(kTile);
    std::vector<int> ord(kTile);
    std::vector<int32_t> cntn;
    for (int64_t t = t0; t < t1 && !bad; ++t) {
      const int64_t b = t * kTile, cnt = std::min<int64_t>(kTile, E - b);
      std::vector<int32_t> &loc = tloc[(size_t)t];
      for (int64_t k = 0; k < cnt; ++k)
        for (int a = 0; a < nne; ++a) {
          const int32_t nd = elements[(int64_t)key[(size_t)(b + k)].second * nne + a];
          if (mark[(size_t)nd] != (int32_t)t) {
            mark[(size_t)nd] = (int32_t)t;
            loc.push_back(nd);
          }
        }
      std::sort(loc.begin(), loc.end());
      if (loc.size() > 65535) {
        bad = 1;
        return;
      }
      for (size_t i = 0; i < loc.size(); ++i) lid[(size_t)loc[i]] = (int32_t)i;
      used.assign(loc.size() * 2, 0);
      int ncol = 0;
      for (int64_t k = 0; k < cnt; ++k) {
        const int32_t e = key[(size_t)(b + k)].second;
        uint64_t m0 = 0, m1 = 0;
        for (int a = 0; a < nne; ++a) {
          const uint16_t l = (uint16_t)lid[(size_t)elements[(int64_t)e * nne + a]];
          ln[(size_t)k][a] = l;
          m0 |= used[2 * l];
          m1 |= used[2 * l + 1];
        }
        int c = -1;
        if (~m0) c = __builtin_ctzll(~m0);
        else if (~m1) c = 64 + __builtin_ctzll(~m1);
        if (c < 0) {
          bad = 2;
          return;
        }
        color[(size_t)k] = c;
        ncol = std::max(ncol, c + 1);
        for (int a = 0; a < nne; ++a) {
          const uint16_t l = ln[(size_t)k][a];
          used[2 * l + (c >> 6)] |= uint64_t(1) << (c & 63);
        }
      }
      tcol[(size_t)t] = ncol;
      // the tile's elements ordered by color, ties in tile order (counting sort)
      int16_t *off = &toff[(size_t)t * (kMaxColors + 1)];
      for (int64_t k = 0; k < cnt; ++k) off[color[(size_t)k] + 1]++;
      for (int c = 0; c < ncol; ++c) off[c + 1] += off[c];
      int16_t fillc[kMaxColors];
      for (int c = 0; c < ncol; ++c) fillc[c] = off[c];
      for (int64_t k = 0; k < cnt; ++k) ord[(size_t)fillc[color[(size_t)k]]++] = (int)k;
      for (int64_t k = 0; k < cnt; ++k) {
        const int src = ord[(size_t)k];
        for (int a = 0; a < nne; ++a) T.lnodes[(size_t)((b + k) * nne + a)] = ln[(size_t)src][a];
        T.elem[(size_t)(b + k)] = key[(size_t)(b + src)].second;
      }
      if (incidence) {
        // entries k * nne + a of each local node, ascending (a counting sort)
        cntn.assign(loc.size() + 1, 0);
        for (int64_t q = 0; q < cnt * nne; ++q) cntn[(size_t)T.lnodes[(size_t)(b * nne + q)] + 1]++;
        for (size_t j = 0; j < loc.size(); ++j) cntn[j + 1] += cntn[j];
        tinc[(size_t)t].assign(cntn.begin() + 1, cntn.end());
        for (int64_t q = 0; q < cnt * nne; ++q) {
          const uint16_t l = T.lnodes[(size_t)(b * nne + q)];
          T.inc_list[(size_t)(b * nne + cntn[l]++)] = (uint16_t)q;
        }
      }
    }
  });
  if (!ok) {
    set_error("out of host memory while tiling the mesh");
    return TATVA_ENOMEM;
  }
  if (bad == 1) {
    set_error("tile has too many local nodes");
    return TATVA_EINVAL;
  }
  if (bad == 2) {
    set_error("more than 128 element colors in a tile");
    return TATVA_EINVAL;
  }
  // concatenation in tile order
  T.node_ptr.assign((size_t)n_tiles + 1, 0);
  T.max_local = 0;
  T.max_colors = 0;
  for (int64_t t = 0; t < n_tiles; ++t) {
    const int64_t sz = (int64_t)tloc[(size_t)t].size();
    if ((int64_t)T.node_ptr[(size_t)t] + sz >= (int64_t(1) << 31)) {
      set_error("too many tile slots");
      return TATVA_EINVAL;
    }
    T.node_ptr[(size_t)t + 1] = T.node_ptr[(size_t)t] + (int32_t)sz;
    T.max_local = std::max<int32_t>(T.max_local, (int32_t)sz);
    T.max_colors = std::max(T.max_colors, tcol[(size_t)t]);
  }
  const size_t n_slots = (size_t)T.node_ptr.back();
  T.tnodes.resize(n_slots);
  if (incidence) T.inc_end.resize(n_slots);
  for (int64_t t = 0; t < n_tiles; ++t) {
    const size_t at = (size_t)T.node_ptr[(size_t)t];
    std::copy(tloc[(size_t)t].begin(), tloc[(size_t)t].end(), T.tnodes.begin() + at);
    if (incidence) std::copy(tinc[(size_t)t].begin(), tinc[(size_t)t].end(), T.inc_end.begin() + at);
  }
  std::vector<std::vector<int32_t>>().swap(tloc);
  std::vector<std::vector<uint16_t>>().swap(tinc);
  T.color_off.assign((size_t)n_tiles * (T.max_colors + 1), 0);
  for (int64_t t = 0; t < n_tiles; ++t) {
    const int16_t *off = &toff[(size_t)t * (kMaxColors + 1)];
    const int ncol = tcol[(size_t)t];
    for (int c = 0; c <= T.max_colors; ++c)
      T.color_off[(size_t)t * (T.max_colors + 1) + c] = off[std::min(c, ncol)];
  }
  // node -> slots (tiles visited in order, so each list ascends by tile)
  T.slot_ptr.assign((size_t)N + 1, 0);
  for (int32_t nd : T.tnodes) T.slot_ptr[(size_t)nd + 1]++;
  for (int64_t i = 0; i < N; ++i) T.slot_ptr[(size_t)i + 1] += T.slot_ptr[(size_t)i];
  T.slot_pos.assign(n_slots, 0);
  std::vector<int32_t> fill(T.slot_ptr.begin(), T.slot_ptr.end() - 1);
  for (size_t sl = 0; sl < n_slots; ++sl) T.slot_pos[sl] = fill[(size_t)T.tnodes[sl]]++;
  return TATVA_OK;
}

}  // namespace

// ---------------------------------------------------------------- C ABI ----

extern "C" int tatva_mesh_create(int device, int elem_kind, const double *coords,
                                 int64_t n_nodes, int dim, const int32_t *elements,
                                 int64_t n_elements, int nne, int nq, const double *quad_points,
                                 const double *quad_weights, tatva_mesh **out) {
  TATVA_REQUIRE(out != nullptr, "out is NULL");
  *out = nullptr;
  TATVA_REQUIRE(coords && elements, "coords/elements are NULL");
  TATVA_REQUIRE(n_nodes > 0, "Mesh must contain at least one node.");
  TATVA_REQUIRE(n_elements > 0, "Mesh must contain at least one element.");
  int exp_dim = 0, exp_nne = 0;
  switch (elem_kind) {
    case TATVA_TRI3: exp_dim = 2; exp_nne = 3; break;
    case TATVA_TET4: exp_dim = 3; exp_nne = 4; break;
    case TATVA_HEX8: exp_dim = 3; exp_nne = 8; break;
    default: set_error("unknown element kind"); return TATVA_EINVAL;
  }
  TATVA_REQUIRE(dim == exp_dim, "mesh dimension does not match the element");
  TATVA_REQUIRE(nne == exp_nne, "nodes per element do not match the element");
  TATVA_REQUIRE(nq >= 1 && nq <= kMaxQuad, "number of quadrature points must be in [1, 27]");
  TATVA_REQUIRE(quad_points && quad_weights, "quadrature is NULL");
  TATVA_REQUIRE(n_elements < (int64_t(1) << 28), "too many elements (packed incidence)");
  TATVA_REQUIRE(n_nodes < (int64_t(1) << 31), "too many nodes");
  const int64_t total = n_elements * nne;
  for (int64_t i = 0; i < total; ++i) {
    if (elements[i] < 0) {
      set_error("Mesh element connectivity contains negative node indices.");
      return TATVA_EINVAL;
    }
    if (elements[i] >= n_nodes) {
      set_error("Mesh element connectivity references nodes outside the mesh coordinates array.");
      return TATVA_EINVAL;
    }
  }
  TATVA_REQUIRE(total < (int64_t(1) << 31), "too many element-node incidences");
  // Node -> (element, local node) incidence by counting sort: ascending element
  // order per node, so every node gather sums in a fixed order.
  std::vector<int32_t> ptr((size_t)n_nodes + 1, 0), list((size_t)total);
  for (int64_t i = 0; i < total; ++i) ptr[(size_t)elements[i] + 1]++;
  int max_inc = 0;
  for (int64_t i = 0; i < n_nodes; ++i) {
    max_inc = std::max(max_inc, ptr[(size_t)i + 1]);
    ptr[(size_t)i + 1] += ptr[(size_t)i];
  }
  {
    std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
    for (int64_t e = 0; e < n_elements; ++e)
      for (int a = 0; a < nne; ++a) {
        const int32_t nd = elements[e * nne + a];
        list[(size_t)fill[(size_t)nd]++] = (int32_t)((e << 3) | a);
      }
  }
  tatva_mesh *m = new tatva_mesh();
  std::memset(m, 0, sizeof(*m));
  m->device = device;
  m->kind = elem_kind;
  m->dim = dim;
  m->nne = nne;
  m->n_elements = n_elements;
  m->n_nodes = n_nodes;
  m->max_incident = max_inc;
  m->quad.nq = nq;
  for (int q = 0; q < nq; ++q) {
    for (int j = 0; j < 3; ++j) m->quad.xi[q][j] = j < dim ? quad_points[q * dim + j] : 0.0;
    m->quad.w[q] = quad_weights[q];
  }
  set_quad_f32(m->quad);
  // Uploads go through a private non-blocking stream: mesh creation on a
  // helper thread (a server preparing requests) then does not wait for work
  // queued on the default stream.
  hipStream_t up = nullptr;
  auto fail = [&](hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    if (up) {
      (void)hipStreamSynchronize(up);
      (void)hipStreamDestroy(up);
    }
    tatva_mesh_destroy(m);
    return TATVA_EHIP;
  };
  HostTiles T;
  {
    const int rc = build_tiles(coords, dim, n_nodes, elements, n_elements, nne,
                               tile_gather_kind(elem_kind), T);
    if (rc) {
      delete m;
      return rc;
    }
  }
  m->n_slots = T.node_ptr.back();
  if (elem_kind == TATVA_HEX8) {
    bool affine = true;
    for (int64_t e = 0; e < n_elements && affine; ++e) {
      double X[8][3];
      for (int a = 0; a < 8; ++a)
        for (int j = 0; j < 3; ++j) X[a][j] = coords[(int64_t)elements[e * 8 + a] * 3 + j];
      affine = hex_is_affine_host(X);
    }
    m->hex_affine = affine ? 1 : 0;
  }
  hipError_t err = hipSetDevice(device);
  if (err != hipSuccess) return fail(err, "hipSetDevice");
  if ((err = hipStreamCreateWithFlags(&up, hipStreamNonBlocking)) != hipSuccess) {
    up = nullptr;
    return fail(err, "hipStreamCreateWithFlags");
  }
  // tile partial sums: n_slots x ncomp (ncomp <= dim) values of at most 8 bytes
  m->scratch_bytes = std::max<int64_t>(1, m->n_slots) * dim * (int64_t)sizeof(double);
  auto upload = [&](auto **dst, const auto &vec, const char *what) -> hipError_t {
    using V = typename std::remove_reference<decltype(vec)>::type::value_type;
    hipError_t e2 = hipMalloc((void **)dst, std::max<size_t>(1, vec.size()) * sizeof(V));
    if (e2 != hipSuccess) return e2;
    (void)what;
    return hipMemcpyAsync(*dst, vec.data(), vec.size() * sizeof(V), hipMemcpyHostToDevice, up);
  };
  if ((err = upload(&m->t_lnodes, T.lnodes, "tiles")) != hipSuccess) return fail(err, "tiles");
  if ((err = upload(&m->t_color_off, T.color_off, "tiles")) != hipSuccess) return fail(err, "tiles");
  if ((err = upload(&m->t_node_ptr, T.node_ptr, "tiles")) != hipSuccess) return fail(err, "tiles");
  if ((err = upload(&m->t_slot_ptr, T.slot_ptr, "tiles")) != hipSuccess) return fail(err, "tiles");
  if ((err = upload(&m->t_slot_pos, T.slot_pos, "tiles")) != hipSuccess) return fail(err, "tiles");
  if ((err = upload(&m->t_tnodes, T.tnodes, "tiles")) != hipSuccess) return fail(err, "tiles");
  if ((err = upload(&m->t_elem, T.elem, "tiles")) != hipSuccess) return fail(err, "tiles");
  if (tile_gather_kind(elem_kind)) {
    if ((err = upload(&m->t_inc_end, T.inc_end, "tiles")) != hipSuccess) return fail(err, "tiles");
    if ((err = upload(&m->t_inc_list, T.inc_list, "tiles")) != hipSuccess) return fail(err, "tiles");
  }
  m->tiles = TileSet{T.n_tiles, T.max_local, T.max_colors, 0, m->t_lnodes,
                     m->t_color_off, m->t_node_ptr, m->t_slot_ptr, m->t_slot_pos,
                     m->t_tnodes, m->t_inc_end, m->t_inc_list, m->t_elem};
  if ((err = hipMalloc(&m->coords, n_nodes * dim * sizeof(double))) != hipSuccess)
    return fail(err, "hipMalloc coords");
  if ((err = hipMalloc(&m->conn, total * sizeof(int32_t))) != hipSuccess)
    return fail(err, "hipMalloc conn");
  if ((err = hipMalloc(&m->inv_ptr, (n_nodes + 1) * sizeof(int32_t))) != hipSuccess)
    return fail(err, "hipMalloc inv_ptr");
  if ((err = hipMalloc(&m->inv_list, total * sizeof(int32_t))) != hipSuccess)
    return fail(err, "hipMalloc inv_list");
  if ((err = hipMalloc(&m->scratch, m->scratch_bytes)) != hipSuccess)
    return fail(err, "hipMalloc scratch");
  if ((err = hipMalloc(&m->partials, kMaxReduceBlocks * 4 * sizeof(double))) != hipSuccess)
    return fail(err, "hipMalloc partials");
  if ((err = hipMemcpyAsync(m->coords, coords, n_nodes * dim * sizeof(double), hipMemcpyHostToDevice, up)) != hipSuccess)
    return fail(err, "copy coords");
  if ((err = hipMemcpyAsync(m->conn, elements, total * sizeof(int32_t), hipMemcpyHostToDevice, up)) != hipSuccess)
    return fail(err, "copy conn");
  if ((err = hipMemcpyAsync(m->inv_ptr, ptr.data(), (n_nodes + 1) * sizeof(int32_t), hipMemcpyHostToDevice,
                            up)) != hipSuccess)
    return fail(err, "copy inv_ptr");
  if ((err = hipMemcpyAsync(m->inv_list, list.data(), total * sizeof(int32_t), hipMemcpyHostToDevice,
                            up)) != hipSuccess)
    return fail(err, "copy inv_list");
  // the host arrays above must outlive the copies
  if ((err = hipStreamSynchronize(up)) != hipSuccess) return fail(err, "mesh upload");
  (void)hipStreamDestroy(up);
  up = nullptr;
  *out = m;
  return TATVA_OK;
}

extern "C" int tatva_mesh_destroy(tatva_mesh *m) {
  if (!m) return TATVA_OK;
  (void)hipSetDevice(m->device);
  (void)hipDeviceSynchronize();
  forget_mesh_graphs(m);
  if (m->coords) (void)hipFree(m->coords);
  if (m->conn) (void)hipFree(m->conn);
  if (m->inv_ptr) (void)hipFree(m->inv_ptr);
  if (m->inv_list) (void)hipFree(m->inv_list);
  if (m->scratch) (void)hipFree(m->scratch);
  if (m->partials) (void)hipFree(m->partials);
  for (void *p : {(void *)m->t_lnodes, (void *)m->t_color_off,
                  (void *)m->t_node_ptr, (void *)m->t_slot_ptr, (void *)m->t_slot_pos,
                  (void *)m->t_tnodes, (void *)m->t_inc_end, (void *)m->t_inc_list,
                  (void *)m->t_elem,
                  (void *)m->loc_ptr, (void *)m->loc_elems})
    if (p) (void)hipFree(p);
  delete m;
  return TATVA_OK;
}

extern "C" int tatva_mesh_info(const tatva_mesh *m, int64_t *out, int32_t n) {
  TATVA_REQUIRE(m && out && n >= 0, "invalid arguments");
  const int64_t v[8] = {m->n_nodes,       m->n_elements,           m->tiles.n_tiles, m->n_slots,
                        m->tiles.max_local, m->tiles.max_colors, m->hex_affine,    m->quad.nq};
  for (int32_t i = 0; i < n && i < 8; ++i) out[i] = v[i];
  return TATVA_OK;
}

#define MESH_PRELUDE(m)                                   \
  TATVA_REQUIRE((m) != nullptr, "mesh is NULL");          \
  {                                                       \
    int _rc = set_device(m);                              \
    if (_rc) return _rc;                                  \
  }

extern "C" int tatva_weights(tatva_mesh *m, double *out, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(out, "out is NULL");
  hipStream_t s = (hipStream_t)stream;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    k_weights<K><<<grid_for(m->n_elements), kBlock, 0, s>>>(m->n_elements, m->coords, m->conn,
                                                             m->quad, out);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
}

extern "C" int tatva_eval(tatva_mesh *m, int nc, const double *u, double *out, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && out && nc >= 1, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    k_eval<K><<<grid_for(m->n_elements), kBlock, 0, s>>>(m->n_elements, m->conn, m->quad, nc, u, out);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
}

extern "C" int tatva_grad(tatva_mesh *m, int nc, const double *u, double *out, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && out && nc >= 1, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    k_grad<K><<<grid_for(m->n_elements), kBlock, 0, s>>>(m->n_elements, m->coords, m->conn,
                                                          m->quad, nc, u, out);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
}

extern "C" int tatva_integrate(tatva_mesh *m, int nc, int nodal, const double *values,
                               double *per_element, double *total, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(values && nc >= 1, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    const int g = reduce_grid(m->n_elements);
    if (per_element) {
      k_integrate_elem<K><<<grid_for(m->n_elements), kBlock, 0, s>>>(
          m->n_elements, m->coords, m->conn, m->quad, nc, nodal, values, per_element);
      if (total)
        for (int c = 0; c < nc; ++c) {
          k_strided_sum<<<g, kBlock, 0, s>>>(m->n_elements, per_element, nc, c, m->partials);
          k_finalize_sum<double><<<1, kBlock, 0, s>>>(m->partials, g, total + c, 0);
        }
    } else if (total) {
      for (int c = 0; c < nc; ++c) {
        k_integrate_comp<K><<<g, kBlock, 0, s>>>(m->n_elements, m->coords, m->conn, m->quad, nc,
                                                  c, nodal, values, m->partials);
        k_finalize_sum<double><<<1, kBlock, 0, s>>>(m->partials, g, total + c, 0);
      }
    }
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
}

extern "C" int tatva_density_eval(int dim, int64_t n_points, const tatva_density *d,
                                  const double *grad, const double *values, double *psi,
                                  void *stream) {
  TATVA_REQUIRE(grad && psi && n_points >= 0, "invalid arguments");
  TATVA_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  if (d && d->kind == TATVA_DENSITY_POISSON) TATVA_REQUIRE(values, "Poisson needs the values u_q");
  if (n_points == 0) return TATVA_OK;
  // per-element parameters: points come in runs of points_per_element per element
  int64_t ppe = 1;
  if (d && d->elem_params) {
    TATVA_REQUIRE(d->points_per_element >= 1 && n_points % d->points_per_element == 0,
                  "per-element parameters need points_per_element dividing n_points");
    ppe = d->points_per_element;
  }
  hipStream_t s = (hipStream_t)stream;
  const int kind = dim == 2 ? TATVA_TRI3 : TATVA_HEX8;
  return with_kind(kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, double>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      k_density_eval<K, Dens><<<grid_for(n_points), kBlock, 0, s>>>(n_points / ppe, (int)ppe, dens,
                                                                     grad, values, psi);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    });
  });
}

extern "C" int tatva_energy(tatva_mesh *m, const tatva_density *d, const double *u,
                            const double *f_ext, double *out, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && out, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  int rc = with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, double>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      const int g = reduce_grid(m->n_elements);
      k_energy<K, Dens><<<g, kBlock, 0, s>>>(m->n_elements, m->coords, m->conn, m->quad, dens, u,
                                             m->partials);
      k_finalize_sum<double><<<1, kBlock, 0, s>>>(m->partials, g, out, 0);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    });
  });
  if (rc) return rc;
  if (f_ext) {
    const int64_t n = m->n_nodes * density_ncomp(m, d);
    const int g = reduce_grid(n);
    k_dot_partials<<<g, kBlock, 0, s>>>(n, f_ext, u, m->partials);
    k_finalize_sum<double><<<1, kBlock, 0, s>>>(m->partials, g, out, 1);
    TATVA_CHECK_LAUNCH();
  }
  return TATVA_OK;
}

namespace tatva {

// Tile assembly + slot gather for one mode; out is reduced through dof_map.
template <class T, int MODE>
static int launch_assembly(tatva_mesh *m, const tatva_density *d, const T *u, const T *v,
                           const int32_t *dof_map, const double *f_ext, T *out, hipStream_t s,
                           const int *skip, double *part_dot = nullptr, int gather_blocks = 0,
                           int rev = 0, hipEvent_t between = nullptr, T *lin = nullptr,
                           const T *geo = nullptr, const CgDefer &dx = CgDefer{}) {
  T *partial = (T *)m->scratch;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, T>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      constexpr int NC = Dens::NC;
      if constexpr ((MODE == kModeLinearize || MODE == kModeHvpLin) && Dens::kLin == 0) {
        set_error("this density has no linearisation to cache");
        return TATVA_EINVAL;
      } else {
      const size_t lds = tile_lds_bytes(K, m->tiles.max_local, Elem<K>::D, Elem<K>::NNE, NC, sizeof(T), MODE);
      if (lds > kTileLdsMax) {
        set_error("mesh tiles need more LDS than a workgroup holds");
        return TATVA_EINVAL;
      }
      // element geometry: from a geometry cache (FP32 HVP / linearisation), once
      // per element (affine Hex8), or at every point
      constexpr bool kCacheable = sizeof(T) == 4 && (MODE == kModeHvp || MODE == kModeHvpLin ||
                                                     MODE == kModeLinearize);
      if (geo && !(kCacheable && K == TATVA_HEX8 && m->hex_affine)) {
        set_error("a geometry cache serves the FP32 HVP of affine Hex8 meshes only");
        return TATVA_EINVAL;
      }
      constexpr bool kDeferable = sizeof(T) == 4 && (MODE == kModeHvp || MODE == kModeHvpLin);
      if (dx.n != 0 && !kDeferable) {
        set_error("a deferred CG update rides on the FP32 HVP only");
        return TATVA_EINVAL;
      }
      const int grid = tile_grid(m->tiles.n_tiles, dx.n != 0);
      if (geo) {
        if constexpr (kCacheable)
          k_tile_assemble<K, Dens, T, MODE, kGeomCache><<<grid, kTile, lds, s>>>(
              m->coords, m->quad, dens, u, v, dof_map, m->tiles, partial, skip, rev, lin, geo, dx);
      } else if (K == TATVA_HEX8 && m->hex_affine) {
        k_tile_assemble<K, Dens, T, MODE, kGeomAffine><<<grid, kTile, lds, s>>>(
            m->coords, m->quad, dens, u, v, dof_map, m->tiles, partial, skip, rev, lin, nullptr,
            dx);
      } else {
        k_tile_assemble<K, Dens, T, MODE, kGeomGeneral><<<grid, kTile, lds, s>>>(
            m->coords, m->quad, dens, u, v, dof_map, m->tiles, partial, skip, rev, lin, nullptr,
            dx);
      }
      if constexpr (MODE == kModeLinearize) {
        TATVA_CHECK_LAUNCH();
        return TATVA_OK;  // the cache is the output: no gather
      }
      // optional event between the tile kernel and the gather (kernel timing)
      if (between) TATVA_TRY_HIP(hipEventRecord(between, s));
      if (part_dot)
        k_slot_gather_dot<NC, T><<<gather_blocks, kBlock, 0, s>>>(
            m->n_nodes, m->t_slot_ptr, partial, dof_map, v, out, part_dot, skip);
      else
        k_slot_gather<NC, T><<<grid_for(m->n_nodes), kBlock, 0, s>>>(
            m->n_nodes, m->t_slot_ptr, partial, f_ext, dof_map, out, skip);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
      }
    });
  });
}

int launch_hvp_f64(tatva_mesh *m, const tatva_density *d, const double *u, const double *v,
                   const int32_t *dof_map, double *y, hipStream_t s, const int *skip) {
  return launch_assembly<double, kModeHvp>(m, d, u, v, dof_map, nullptr, y, s, skip);
}

int launch_hvp_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                   const int32_t *dof_map, float *y, hipStream_t s, const int *skip,
                   const float *lin, const float *geo) {
  if (lin)
    return launch_assembly<float, kModeHvpLin>(m, d, u, v, dof_map, nullptr, y, s, skip, nullptr,
                                               0, 0, nullptr, const_cast<float *>(lin), geo);
  return launch_assembly<float, kModeHvp>(m, d, u, v, dof_map, nullptr, y, s, skip, nullptr, 0, 0,
                                          nullptr, nullptr, geo);
}

int64_t lin_cache_floats(const tatva_mesh *m, const tatva_density *d) {
  if (!m || !d || d->kind != TATVA_DENSITY_NEO_HOOKEAN) return 0;
  return (int64_t)m->tiles.n_tiles * kTile * m->quad.nq * (m->dim * m->dim + 1);
}

int launch_linearize_f32(tatva_mesh *m, const tatva_density *d, const float *u, float *lin,
                         hipStream_t s, const float *geo) {
  return launch_assembly<float, kModeLinearize>(m, d, u, nullptr, nullptr, nullptr, nullptr, s,
                                                nullptr, nullptr, 0, 0, nullptr, lin, geo);
}

// Writes the geometry cache of an affine Hex8 mesh (layout at kGeoFloats):
// one thread per tile slot, its element's nodes gathered from conn.
__global__ __launch_bounds__(kTile) void k_geometry_cache_hex8(const double *__restrict__ coords,
                                                               const int32_t *__restrict__ conn,
                                                               TileSet ts,
                                                               float *__restrict__ geo) {
  const int tile = blockIdx.x;
  const int64_t slot = (int64_t)tile * kTile + threadIdx.x;
  const int32_t e = ts.elem[slot];
  float *gp = geo + (int64_t)tile * kGeoFloats * kTile + threadIdx.x;
  if (e < 0) {  // padding slot: zeros, never read
#pragma unroll
    for (int k = 0; k < kGeoFloats; ++k) gp[k * kTile] = 0.0f;
    return;
  }
  double X[8][3];
#pragma unroll
  for (int a = 0; a < 8; ++a) {
    const int64_t g = conn[(int64_t)e * 8 + a];
#pragma unroll
    for (int k = 0; k < 3; ++k) X[a][k] = coords[g * 3 + k];
  }
  float iv[3][3];
  const double det = hex8_affine_inverse<float>(X, iv);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) gp[(i * 3 + j) * kTile] = iv[i][j];
  gp[9 * kTile] = __uint_as_float((uint32_t)__double2loint(det));
  gp[10 * kTile] = __uint_as_float((uint32_t)__double2hiint(det));
}

int64_t geo_cache_floats(const tatva_mesh *m) {
  if (!m || m->kind != TATVA_HEX8 || !m->hex_affine) return 0;
  return (int64_t)m->tiles.n_tiles * kTile * kGeoFloats;
}

int launch_geometry_cache_f32(tatva_mesh *m, float *geo, hipStream_t s) {
  if (geo_cache_floats(m) == 0) {
    set_error("a geometry cache needs an affine Hex8 mesh");
    return TATVA_EINVAL;
  }
  k_geometry_cache_hex8<<<m->tiles.n_tiles, kTile, 0, s>>>(m->coords, m->conn, m->tiles, geo);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

int hvp_gather_blocks(const tatva_mesh *m) {
  int64_t b = (m->n_nodes + kBlock - 1) / kBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, kMaxReduceBlocks));
}

int launch_hvp_dot_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                       const int32_t *dof_map, float *y, double *part_dot, hipStream_t s,
                       const int *skip, int rev, hipEvent_t between, const float *lin,
                       const float *geo, const CgDefer *dx) {
  const CgDefer none{};
  const CgDefer &defer = dx ? *dx : none;
  if (lin)
    return launch_assembly<float, kModeHvpLin>(m, d, u, v, dof_map, nullptr, y, s, skip, part_dot,
                                               hvp_gather_blocks(m), rev, between,
                                               const_cast<float *>(lin), geo, defer);
  return launch_assembly<float, kModeHvp>(m, d, u, v, dof_map, nullptr, y, s, skip, part_dot,
                                          hvp_gather_blocks(m), rev, between, nullptr, geo, defer);
}

}  // namespace tatva

extern "C" int tatva_residual(tatva_mesh *m, const tatva_density *d, const double *u,
                              const double *f_ext, const int32_t *dof_map, double *r,
                              void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && r, "invalid arguments");
  return launch_assembly<double, kModeResidual>(m, d, u, nullptr, dof_map, f_ext, r,
                                                (hipStream_t)stream, nullptr);
}

extern "C" int tatva_hvp(tatva_mesh *m, const tatva_density *d, const double *u, const double *v,
                         const int32_t *dof_map, double *y, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && v && y, "invalid arguments");
  return launch_hvp_f64(m, d, u, v, dof_map, y, (hipStream_t)stream);
}

extern "C" int tatva_hvp_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                             const int32_t *dof_map, float *y, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && v && y, "invalid arguments");
  return launch_hvp_f32(m, d, u, v, dof_map, y, (hipStream_t)stream);
}

extern "C" int tatva_hessian_diag(tatva_mesh *m, const tatva_density *d, const double *u,
                                  const int32_t *dof_map, double *diag, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && diag, "invalid arguments");
  return launch_assembly<double, kModeDiag>(m, d, u, nullptr, dof_map, nullptr, diag,
                                             (hipStream_t)stream, nullptr);
}

extern "C" int tatva_hessian_diag_f32(tatva_mesh *m, const tatva_density *d, const float *u,
                                      const int32_t *dof_map, float *diag, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && diag, "invalid arguments");
  return launch_assembly<float, kModeDiag>(m, d, u, nullptr, dof_map, nullptr, diag,
                                            (hipStream_t)stream, nullptr);
}

extern "C" int64_t tatva_colored_hessian_workspace(tatva_mesh *m, int nc) {
  if (!m || nc < 1) return -1;
  const int64_t nd = (int64_t)m->nne * nc;
  return m->n_elements * nd * nd * (int64_t)sizeof(double);
}

static constexpr int kMaxColors = 2048;

extern "C" int tatva_colored_hessian(tatva_mesh *m, const tatva_density *d, const double *u,
                                     const int32_t *dof_map, const int32_t *red_to_full,
                                     int64_t n_rows, const int32_t *colors, int32_t n_colors,
                                     const int32_t *indptr, const int32_t *indices, int64_t nnz,
                                     double *data, void *workspace, int64_t ws_bytes,
                                     void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && colors && indptr && indices && data, "invalid arguments");
  TATVA_REQUIRE(n_colors >= 1 && n_colors <= kMaxColors, "n_colors must be in [1, 2048]");
  TATVA_REQUIRE((dof_map == nullptr) == (red_to_full == nullptr),
                "dof_map and red_to_full must be given together");
  const int nc = density_ncomp(m, d);
  if (!dof_map) TATVA_REQUIRE(n_rows == m->n_nodes * nc, "pattern size does not match the mesh dofs");
  const int64_t need = tatva_colored_hessian_workspace(m, nc);
  TATVA_REQUIRE(workspace && ws_bytes >= need, "colored Hessian workspace too small");
  (void)nnz;
  hipStream_t s = (hipStream_t)stream;
  double *Ke = (double *)workspace;
  const int rows_per_block = kBlock / kWave;
  const size_t lds = (size_t)rows_per_block * n_colors * sizeof(double);
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, double>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      constexpr int NNE = Elem<K>::NNE, NC = Dens::NC;
      const int64_t nthreads = m->n_elements * NNE * NC;
      k_elem_stiffness<K, Dens><<<grid_for(nthreads), kBlock, 0, s>>>(
          m->n_elements, m->coords, m->conn, m->quad, dens, u, Ke);
      TATVA_CHECK_LAUNCH();
      const int64_t nblk = (n_rows + rows_per_block - 1) / rows_per_block;
      TATVA_REQUIRE(nblk < (int64_t(1) << 31), "too many rows");
      k_colored_rows<NNE, NC><<<(int)nblk, kBlock, lds, s>>>(
          n_rows, red_to_full, dof_map, colors, n_colors, indptr, indices, m->inv_ptr,
          m->inv_list, m->conn, Ke, data);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    });
  });
}

extern "C" int tatva_colored_hessian_mapped(tatva_mesh *m, const tatva_density *d, const double *u,
                                            const int32_t *read_map, const int32_t *row_ptr,
                                            const int32_t *row_list, int64_t n_rows,
                                            const int32_t *colors, int32_t n_colors,
                                            const int32_t *indptr, const int32_t *indices,
                                            int64_t nnz, int32_t flags, double *data,
                                            void *workspace, int64_t ws_bytes, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && read_map && row_ptr && row_list && colors && indptr && indices && data,
                "invalid arguments");
  TATVA_REQUIRE(n_colors >= 1 && n_colors <= kMaxColors, "n_colors must be in [1, 2048]");
  const int nc = density_ncomp(m, d);
  const int64_t need = tatva_colored_hessian_workspace(m, nc);
  TATVA_REQUIRE(workspace && ws_bytes >= need, "colored Hessian workspace too small");
  (void)nnz;
  hipStream_t s = (hipStream_t)stream;
  double *Ke = (double *)workspace;
  const int rows_per_block = kBlock / kWave;
  const size_t lds = (size_t)rows_per_block * n_colors * sizeof(double);
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, double>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      constexpr int NNE = Elem<K>::NNE, NC = Dens::NC;
      const int64_t nthreads = m->n_elements * NNE * NC;
      k_elem_stiffness<K, Dens><<<grid_for(nthreads), kBlock, 0, s>>>(
          m->n_elements, m->coords, m->conn, m->quad, dens, u, Ke);
      TATVA_CHECK_LAUNCH();
      if (n_rows == 0) return TATVA_OK;
      const int64_t nblk = (n_rows + rows_per_block - 1) / rows_per_block;
      TATVA_REQUIRE(nblk < (int64_t(1) << 31), "too many rows");
      k_colored_rows_mapped<NNE, NC><<<(int)nblk, kBlock, lds, s>>>(
          n_rows, row_ptr, row_list, read_map, colors, n_colors, indptr, indices, m->inv_ptr,
          m->inv_list, m->conn, Ke, data, flags & 1);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    });
  });
}

extern "C" int tatva_mass_matrix(tatva_mesh *m, int32_t ncomp, const int32_t *dof_map,
                                 const int32_t *red_to_full, int64_t n_rows,
                                 const int32_t *colors, int32_t n_colors, const int32_t *indptr,
                                 const int32_t *indices, int64_t nnz, double *data,
                                 void *workspace, int64_t ws_bytes, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(colors && indptr && indices && data, "invalid arguments");
  TATVA_REQUIRE(ncomp >= 1 && ncomp <= 3, "ncomp must be in [1, 3]");
  TATVA_REQUIRE(n_colors >= 1 && n_colors <= kMaxColors, "n_colors must be in [1, 2048]");
  TATVA_REQUIRE((dof_map == nullptr) == (red_to_full == nullptr),
                "dof_map and red_to_full must be given together");
  if (!dof_map) TATVA_REQUIRE(n_rows == m->n_nodes * ncomp, "pattern size does not match the mesh dofs");
  TATVA_REQUIRE(workspace && ws_bytes >= tatva_colored_hessian_workspace(m, ncomp),
                "mass matrix workspace too small");
  (void)nnz;
  hipStream_t s = (hipStream_t)stream;
  double *Ke = (double *)workspace;
  const int rows_per_block = kBlock / kWave;
  const size_t lds = (size_t)rows_per_block * n_colors * sizeof(double);
  const int64_t nblk = (n_rows + rows_per_block - 1) / rows_per_block;
  TATVA_REQUIRE(nblk < (int64_t(1) << 31), "too many rows");
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    constexpr int NNE = Elem<K>::NNE;
    auto run = [&](auto ncc) -> int {
      constexpr int NC = decltype(ncc)::value;
      k_mass_stiffness<K, NC><<<grid_for(m->n_elements), kBlock, 0, s>>>(m->n_elements, m->coords,
                                                                           m->conn, m->quad, Ke);
      if (n_rows > 0)
        k_colored_rows<NNE, NC><<<(int)nblk, kBlock, lds, s>>>(
            n_rows, red_to_full, dof_map, colors, n_colors, indptr, indices, m->inv_ptr,
            m->inv_list, m->conn, Ke, data);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    };
    if (ncomp == 1) return run(std::integral_constant<int, 1>{});
    if (ncomp == 2) return run(std::integral_constant<int, 2>{});
    return run(std::integral_constant<int, 3>{});
  });
}

extern "C" int tatva_mass_matrix_mapped(tatva_mesh *m, int32_t ncomp, const int32_t *read_map,
                                        const int32_t *row_ptr, const int32_t *row_list,
                                        int64_t n_rows, const int32_t *colors, int32_t n_colors,
                                        const int32_t *indptr, const int32_t *indices,
                                        int64_t nnz, int32_t flags, double *data,
                                        void *workspace, int64_t ws_bytes, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(read_map && row_ptr && row_list && colors && indptr && indices && data,
                "invalid arguments");
  TATVA_REQUIRE(ncomp >= 1 && ncomp <= 3, "ncomp must be in [1, 3]");
  TATVA_REQUIRE(n_colors >= 1 && n_colors <= kMaxColors, "n_colors must be in [1, 2048]");
  TATVA_REQUIRE(workspace && ws_bytes >= tatva_colored_hessian_workspace(m, ncomp),
                "mass matrix workspace too small");
  (void)nnz;
  hipStream_t s = (hipStream_t)stream;
  double *Ke = (double *)workspace;
  const int rows_per_block = kBlock / kWave;
  const size_t lds = (size_t)rows_per_block * n_colors * sizeof(double);
  const int64_t nblk = (n_rows + rows_per_block - 1) / rows_per_block;
  TATVA_REQUIRE(nblk < (int64_t(1) << 31), "too many rows");
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    constexpr int NNE = Elem<K>::NNE;
    auto run = [&](auto ncc) -> int {
      constexpr int NC = decltype(ncc)::value;
      k_mass_stiffness<K, NC><<<grid_for(m->n_elements), kBlock, 0, s>>>(m->n_elements, m->coords,
                                                                           m->conn, m->quad, Ke);
      if (n_rows > 0)
        k_colored_rows_mapped<NNE, NC><<<(int)nblk, kBlock, lds, s>>>(
            n_rows, row_ptr, row_list, read_map, colors, n_colors, indptr, indices, m->inv_ptr,
            m->inv_list, m->conn, Ke, data, flags & 1);
      TATVA_CHECK_LAUNCH();
      return TATVA_OK;
    };
    if (ncomp == 1) return run(std::integral_constant<int, 1>{});
    if (ncomp == 2) return run(std::integral_constant<int, 2>{});
    return run(std::integral_constant<int, 3>{});
  });
}

extern "C" int tatva_project_rhs(tatva_mesh *m, int32_t nc, const double *quad_values,
                                 double *out, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(quad_values && out && nc >= 1, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    k_project_rhs<K><<<grid_for(m->n_nodes), kBlock, 0, s>>>(m->n_nodes, m->coords, m->conn,
                                                            m->inv_ptr, m->inv_list, m->quad, nc,
                                                            quad_values, out);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
}

// Point-location grid of a 2-D mesh, built once from the device arrays: about
// two elements per cell, cells shaped after the bounding box of the nodes;
// each element is listed in every cell its bounding box overlaps, in
// ascending element order.
static int build_point_grid(tatva_mesh *m, hipStream_t s) {
  if (m->loc_ptr) return TATVA_OK;
  const int64_t N = m->n_nodes, E = m->n_elements;
  const int nne = m->nne;
  std::vector<double> X((size_t)N * 2);
  std::vector<int32_t> C((size_t)E * nne);
  hipError_t e = hipMemcpyAsync(X.data(), m->coords, X.size() * sizeof(double), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess)
    e = hipMemcpyAsync(C.data(), m->conn, C.size() * sizeof(int32_t), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error(std::string("point grid: ") + hipGetErrorString(e));
    return TATVA_EHIP;
  }
  double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
  for (int64_t i = 0; i < N; ++i)
    for (int j = 0; j < 2; ++j) {
      lo[j] = std::min(lo[j], X[(size_t)i * 2 + j]);
      hi[j] = std::max(hi[j], X[(size_t)i * 2 + j]);
    }
  const double w0 = hi[0] - lo[0], w1 = hi[1] - lo[1];
  const double cells = std::max<double>(1.0, 0.5 * (double)E);
  const double aspect = (w0 > 0.0 && w1 > 0.0) ? w0 / w1 : 1.0;
  PointGrid g{};
  g.n[0] = (int)std::min(8192.0, std::max(1.0, std::ceil(std::sqrt(cells * aspect))));
  g.n[1] = (int)std::min(8192.0, std::max(1.0, std::ceil(cells / g.n[0])));
  g.lo[0] = lo[0];
  g.lo[1] = lo[1];
  g.h[0] = w0 > 0.0 ? w0 / g.n[0] : 1.0;
  g.h[1] = w1 > 0.0 ? w1 / g.n[1] : 1.0;
  const int64_t n_cells = (int64_t)g.n[0] * g.n[1];
  std::vector<int64_t> ptr((size_t)n_cells + 1, 0);
  auto cell_range = [&](int64_t el, int c0[2], int c1[2]) {
    double a[2] = {1e300, 1e300}, b[2] = {-1e300, -1e300};
    for (int v = 0; v < nne; ++v)
      for (int j = 0; j < 2; ++j) {
        const double x = X[(size_t)C[(size_t)(el * nne + v)] * 2 + j];
        a[j] = std::min(a[j], x);
        b[j] = std::max(b[j], x);
      }
    for (int j = 0; j < 2; ++j) {
      c0[j] = grid_cell(a[j], g.lo[j], g.h[j], g.n[j]);
      c1[j] = grid_cell(b[j], g.lo[j], g.h[j], g.n[j]);
    }
  };
  for (int64_t el = 0; el < E; ++el) {
    int c0[2], c1[2];
    cell_range(el, c0, c1);
    for (int cy = c0[1]; cy <= c1[1]; ++cy)
      for (int cx = c0[0]; cx <= c1[0]; ++cx) ptr[(size_t)((int64_t)cy * g.n[0] + cx) + 1]++;
  }
  for (int64_t c = 0; c < n_cells; ++c) ptr[(size_t)c + 1] += ptr[(size_t)c];
  if (ptr.back() >= (int64_t(1) << 31)) {
    set_error("point grid: too many cell entries");
    return TATVA_EINVAL;
  }
  std::vector<int32_t> ptr32(ptr.begin(), ptr.end()), list((size_t)ptr.back());
  for (int64_t el = 0; el < E; ++el) {
    int c0[2], c1[2];
    cell_range(el, c0, c1);
    for (int cy = c0[1]; cy <= c1[1]; ++cy)
      for (int cx = c0[0]; cx <= c1[0]; ++cx)
        list[(size_t)ptr[(size_t)((int64_t)cy * g.n[0] + cx)]++] = (int32_t)el;
  }
  int32_t *d_ptr = nullptr, *d_list = nullptr;
  e = hipMalloc(&d_ptr, ptr32.size() * sizeof(int32_t));
  if (e == hipSuccess) e = hipMalloc(&d_list, std::max<size_t>(1, list.size()) * sizeof(int32_t));
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_ptr, ptr32.data(), ptr32.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
  if (e == hipSuccess && !list.empty())
    e = hipMemcpyAsync(d_list, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);  // the host arrays must outlive the copies
  if (e != hipSuccess) {
    if (d_ptr) (void)hipFree(d_ptr);
    if (d_list) (void)hipFree(d_list);
    set_error(std::string("point grid: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? TATVA_ENOMEM : TATVA_EHIP;
  }
  m->loc = g;
  m->loc_ptr = d_ptr;
  m->loc_elems = d_list;
  return TATVA_OK;
}

extern "C" int tatva_interpolate(tatva_mesh *m, int64_t n_points, const double *points, int32_t nc,
                                 const double *nodal, double *out, int32_t *n_outside,
                                 void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(n_points >= 0 && nc >= 1 && n_outside, "invalid arguments");
  TATVA_REQUIRE(n_points == 0 || (points && nodal && out), "invalid arguments");
  TATVA_REQUIRE(m->kind == TATVA_TRI3, "point location is implemented for 2-D (Tri3) meshes");
  *n_outside = 0;
  if (n_points == 0) return TATVA_OK;
  hipStream_t s = (hipStream_t)stream;
  {
    const int rc = build_point_grid(m, s);
    if (rc) return rc;
  }
  int *d_out = reinterpret_cast<int *>(m->partials);  // the mesh's reduction scratch
  hipError_t e = hipMemsetAsync(d_out, 0, sizeof(int), s);
  if (e == hipSuccess) {
    k_interpolate<TATVA_TRI3><<<grid_for(n_points), kBlock, 0, s>>>(
        n_points, points, m->loc, m->loc_ptr, m->loc_elems, m->coords, m->conn, m->quad, nc, nodal,
        out, d_out);
    e = hipGetLastError();
  }
  int h = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&h, d_out, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (e != hipSuccess) {
    set_error(std::string("interpolate: ") + hipGetErrorString(e));
    return TATVA_EHIP;
  }
  *n_outside = h;
  return TATVA_OK;
}
