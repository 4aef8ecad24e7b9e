// Element-level kernels on gfx950: quadrature-point gather/eval/grad, weights,
// integration, energy, residual (node-gather assembly), Hessian-vector product,
// element stiffness and the colored-Hessian decompression into CSR.
//
// Threads map to elements (or element columns); FP64 geometry is recomputed
// from the gathered coordinates at every quadrature point (J, signed detJ,
// inv(J) in registers) instead of being stored, so a pass reads only the
// connectivity, the node coordinates and the nodal vectors.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
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
    case TATVA_DENSITY_POISSON: return f(Poisson<D, T>{T(d->p0)});
    case TATVA_DENSITY_LINEAR_ELASTIC: return f(LinearElastic<D, T>{T(d->p0), T(d->p1)});
    case TATVA_DENSITY_NEO_HOOKEAN: return f(NeoHookean<D, T>{T(d->p0), T(d->p1)});
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
  psi[i] = dens.psi(G, uq);
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
      ee += dens.psi(G, uq) * (det * quad.w[q]);
    }
    acc += ee;
  }
  acc = block_sum(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// Element residual vectors r_e[a][c] = sum_q w (P(G) B)_ac - f w N_a.
template <int KIND, class Dens>
__global__ __launch_bounds__(kBlock) void k_elem_residual(int64_t n_elem, const double *__restrict__ coords,
                                                          const int32_t *__restrict__ conn,
                                                          QuadRule quad, Dens dens,
                                                          const double *__restrict__ u,
                                                          double *__restrict__ out) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  double ue[NNE][NC], re[NNE][NC];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      ue[a][c] = u[(int64_t)nodes[a] * NC + c];
      re[a][c] = 0.0;
    }
  for (int q = 0; q < quad.nq; ++q) {
    double B[D][NNE];
    const double det = qp_geometry<KIND>(X, quad.xi[q], B);
    const double w = det * quad.w[q];
    double G[NC][D], P[NC][D];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NNE; ++a) s += B[j][a] * ue[a][c];
        G[c][j] = s;
      }
    dens.P(G, P);
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < D; ++j) s += P[c][j] * B[j][a];
        re[a][c] += w * s;
      }
    if constexpr (Dens::kSource) {
      double N[NNE];
      El::N(quad.xi[q], N);
#pragma unroll
      for (int a = 0; a < NNE; ++a) re[a][0] -= w * dens.f * N[a];
    }
  }
  double *dst = out + e * NNE * NC;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) dst[a * NC + c] = re[a][c];
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
    const auto L = dens.linearize(G);
    dens.dP(L, dG, dP);
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

// Diagonal of the element stiffness, one thread per element.
template <int KIND, class Dens, class T>
__global__ __launch_bounds__(kBlock) void k_elem_diag(int64_t n_elem, const double *__restrict__ coords,
                                                      const int32_t *__restrict__ conn, QuadRule quad,
                                                      Dens dens, const T *__restrict__ u,
                                                      T *__restrict__ out) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  T ue[NNE][NC], de[NNE][NC];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      ue[a][c] = u[(int64_t)nodes[a] * NC + c];
      de[a][c] = T(0);
    }
  for (int q = 0; q < quad.nq; ++q) {
    double Bd[D][NNE];
    const double det = qp_geometry<KIND>(X, quad.xi[q], Bd);
    const T w = T(det * quad.w[q]);
    T B[D][NNE];
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
      for (int a = 0; a < NNE; ++a) B[j][a] = T(Bd[j][a]);
    T G[NC][D];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        T s = T(0);
#pragma unroll
        for (int a = 0; a < NNE; ++a) s += B[j][a] * ue[a][c];
        G[c][j] = s;
      }
    const auto L = dens.linearize(G);
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        T dG[NC][D], dP[NC][D];
#pragma unroll
        for (int cc = 0; cc < NC; ++cc)
#pragma unroll
          for (int j = 0; j < D; ++j) dG[cc][j] = (cc == c) ? B[j][a] : T(0);
        dens.dP(L, dG, dP);
        T s = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) s += dP[c][j] * B[j][a];
        de[a][c] += w * s;
      }
  }
  T *dst = out + e * NNE * NC;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) dst[a * NC + c] = de[a][c];
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

}  // namespace tatva

using namespace tatva;

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
  auto fail = [&](hipError_t e, const char *what) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    tatva_mesh_destroy(m);
    return TATVA_EHIP;
  };
  hipError_t err = hipSetDevice(device);
  if (err != hipSuccess) return fail(err, "hipSetDevice");
  m->scratch_bytes = n_elements * nne * dim * (int64_t)sizeof(double);
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
  if ((err = hipMemcpy(m->coords, coords, n_nodes * dim * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(err, "copy coords");
  if ((err = hipMemcpy(m->conn, elements, total * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(err, "copy conn");
  if ((err = hipMemcpy(m->inv_ptr, ptr.data(), (n_nodes + 1) * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(err, "copy inv_ptr");
  if ((err = hipMemcpy(m->inv_list, list.data(), total * sizeof(int32_t), hipMemcpyHostToDevice)) != hipSuccess)
    return fail(err, "copy inv_list");
  *out = m;
  return TATVA_OK;
}

extern "C" int tatva_mesh_destroy(tatva_mesh *m) {
  if (!m) return TATVA_OK;
  (void)hipSetDevice(m->device);
  (void)hipDeviceSynchronize();
  if (m->coords) (void)hipFree(m->coords);
  if (m->conn) (void)hipFree(m->conn);
  if (m->inv_ptr) (void)hipFree(m->inv_ptr);
  if (m->inv_list) (void)hipFree(m->inv_list);
  if (m->scratch) (void)hipFree(m->scratch);
  if (m->partials) (void)hipFree(m->partials);
  delete m;
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
  double *pe = per_element;
  if (!pe) {
    TATVA_REQUIRE(m->n_elements * nc * (int64_t)sizeof(double) <= m->scratch_bytes,
                  "too many components for the mesh scratch buffer");
    pe = (double *)m->scratch;
  }
  int rc = with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    k_integrate_elem<K><<<grid_for(m->n_elements), kBlock, 0, s>>>(
        m->n_elements, m->coords, m->conn, m->quad, nc, nodal, values, pe);
    TATVA_CHECK_LAUNCH();
    return TATVA_OK;
  });
  if (rc) return rc;
  if (total) {
    const int g = reduce_grid(m->n_elements);
    for (int c = 0; c < nc; ++c) {
      k_strided_sum<<<g, kBlock, 0, s>>>(m->n_elements, pe, nc, c, m->partials);
      k_finalize_sum<double><<<1, kBlock, 0, s>>>(m->partials, g, total + c, 0);
      TATVA_CHECK_LAUNCH();
    }
  }
  return TATVA_OK;
}

extern "C" int tatva_density_eval(int dim, int64_t n_points, const tatva_density *d,
                                  const double *grad, const double *values, double *psi,
                                  void *stream) {
  TATVA_REQUIRE(grad && psi && n_points >= 0, "invalid arguments");
  TATVA_REQUIRE(dim == 2 || dim == 3, "dim must be 2 or 3");
  if (d && d->kind == TATVA_DENSITY_POISSON) TATVA_REQUIRE(values, "Poisson needs the values u_q");
  if (n_points == 0) return TATVA_OK;
  hipStream_t s = (hipStream_t)stream;
  const int kind = dim == 2 ? TATVA_TRI3 : TATVA_HEX8;
  return with_kind(kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, double>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      k_density_eval<K, Dens><<<grid_for(n_points), kBlock, 0, s>>>(n_points, 1, dens, grad,
                                                                     values, psi);
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

template <int NNE, int NC, class T>
static int launch_gather(tatva_mesh *m, const T *elem, const double *f_ext, const int32_t *dof_map,
                         T *out, hipStream_t s, const int *skip = nullptr) {
  k_node_gather<NNE, NC, T><<<grid_for(m->n_nodes), kBlock, 0, s>>>(
      m->n_nodes, m->inv_ptr, m->inv_list, elem, f_ext, dof_map, out, skip);
  TATVA_CHECK_LAUNCH();
  return TATVA_OK;
}

extern "C" int tatva_residual(tatva_mesh *m, const tatva_density *d, const double *u,
                              const double *f_ext, const int32_t *dof_map, double *r,
                              void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && r, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  double *elem = (double *)m->scratch;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, double>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      k_elem_residual<K, Dens><<<grid_for(m->n_elements), kBlock, 0, s>>>(
          m->n_elements, m->coords, m->conn, m->quad, dens, u, elem);
      TATVA_CHECK_LAUNCH();
      return launch_gather<Elem<K>::NNE, Dens::NC, double>(m, elem, f_ext, dof_map, r, s);
    });
  });
}

namespace tatva {

template <class T>
static int launch_hvp_t(tatva_mesh *m, const tatva_density *d, const T *u, const T *v,
                        const int32_t *dof_map, T *y, hipStream_t s, const int *skip) {
  T *elem = (T *)m->scratch;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, T>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      k_elem_hvp<K, Dens, T><<<grid_for(m->n_elements), kBlock, 0, s>>>(
          m->n_elements, m->coords, m->conn, m->quad, dens, u, v, dof_map, elem, skip);
      TATVA_CHECK_LAUNCH();
      return launch_gather<Elem<K>::NNE, Dens::NC, T>(m, elem, nullptr, dof_map, y, s, skip);
    });
  });
}

int launch_hvp_f64(tatva_mesh *m, const tatva_density *d, const double *u, const double *v,
                   const int32_t *dof_map, double *y, hipStream_t s, const int *skip) {
  return launch_hvp_t<double>(m, d, u, v, dof_map, y, s, skip);
}

int launch_hvp_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                   const int32_t *dof_map, float *y, hipStream_t s, const int *skip) {
  return launch_hvp_t<float>(m, d, u, v, dof_map, y, s, skip);
}

template <class T>
static int launch_diag_t(tatva_mesh *m, const tatva_density *d, const T *u,
                         const int32_t *dof_map, T *diag, hipStream_t s) {
  T *elem = (T *)m->scratch;
  return with_kind(m->kind, [&](auto kc) -> int {
    constexpr int K = decltype(kc)::value;
    return with_density<Elem<K>::D, T>(d, [&](auto dens) -> int {
      using Dens = decltype(dens);
      k_elem_diag<K, Dens, T><<<grid_for(m->n_elements), kBlock, 0, s>>>(
          m->n_elements, m->coords, m->conn, m->quad, dens, u, elem);
      TATVA_CHECK_LAUNCH();
      return launch_gather<Elem<K>::NNE, Dens::NC, T>(m, elem, nullptr, dof_map, diag, s);
    });
  });
}

}  // namespace tatva

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
  return launch_diag_t<double>(m, d, u, dof_map, diag, (hipStream_t)stream);
}

extern "C" int tatva_hessian_diag_f32(tatva_mesh *m, const tatva_density *d, const float *u,
                                      const int32_t *dof_map, float *diag, void *stream) {
  MESH_PRELUDE(m);
  TATVA_REQUIRE(u && diag, "invalid arguments");
  return launch_diag_t<float>(m, d, u, dof_map, diag, (hipStream_t)stream);
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
