// Device-resident mesh handle and the element/node kernels shared by the
// operator entry points and the matrix-free solver.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "density.hpp"
#include "element.hpp"

struct tatva_mesh {
  int device;
  int kind;
  int dim;
  int nne;
  int64_t n_elements;
  int64_t n_nodes;
  tatva::QuadRule quad;
  double *coords;        // (N, dim) FP64
  int32_t *conn;         // (E, nne) int32, element-major as in tatva
  int32_t *inv_ptr;      // (N + 1) offsets into inv_list
  int32_t *inv_list;     // packed (element << 3 | local node), sorted by element per node
  void *scratch;         // E * nne * dim doubles: per-element vectors before the node gather
  int64_t scratch_bytes;
  double *partials;      // reduction partials, kMaxReduceBlocks * 4
  int max_incident;      // largest number of elements sharing a node
};

namespace tatva {

// Sum element vectors into nodes: out[map(node*NC+c)] = sum over the node's
// incident (e, a) in ascending e of elem[(e*NNE + a)*NC + c], minus f_ext.
// Every node has one writer and a fixed summation order: deterministic.
template <int NNE, int NC, class T>
__global__ __launch_bounds__(kBlock) void k_node_gather(
    int64_t n_nodes, const int32_t *__restrict__ inv_ptr, const int32_t *__restrict__ inv_list,
    const T *__restrict__ elem, const double *__restrict__ f_ext,
    const int32_t *__restrict__ dof_map, T *__restrict__ out, const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  T acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = T(0);
  const int32_t b = inv_ptr[i], en = inv_ptr[i + 1];
  for (int32_t k = b; k < en; ++k) {
    const int32_t p = inv_list[k];
    const int64_t e = p >> 3;
    const int a = p & 7;
    const T *src = elem + (e * NNE + a) * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += src[c];
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int64_t dof = i * NC + c;
    T v = acc[c];
    if (f_ext) v -= T(f_ext[dof]);
    if (dof_map) {
      const int32_t r = dof_map[dof];
      if (r >= 0) out[r] = v;
    } else {
      out[dof] = v;
    }
  }
}

// Element Hessian-vector product: y_e = sum_q w_q B^T dP(G(u_e); G(v_e)).
// v is read through dof_map (reduced vector, zero on constrained dofs).
template <int KIND, class Dens, class T>
__global__ __launch_bounds__(kBlock) void k_elem_hvp(
    int64_t n_elem, const double *__restrict__ coords, const int32_t *__restrict__ conn,
    QuadRule quad, Dens dens, const T *__restrict__ u, const T *__restrict__ v,
    const int32_t *__restrict__ dof_map, T *__restrict__ out, const int *__restrict__ skip) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
  if (skip && *skip) return;
  const int64_t e = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (e >= n_elem) return;
  int32_t nodes[NNE];
  double X[NNE][D];
  load_element<KIND>(e, conn, coords, nodes, X);
  T ue[NNE][NC], ve[NNE][NC], ye[NNE][NC];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = (int64_t)nodes[a] * NC + c;
      ue[a][c] = u[dof];
      if (dof_map) {
        const int32_t r = dof_map[dof];
        ve[a][c] = r >= 0 ? v[r] : T(0);
      } else {
        ve[a][c] = v[dof];
      }
      ye[a][c] = T(0);
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
    T G[NC][D], dG[NC][D], dP[NC][D];
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        T g = T(0), dg = T(0);
#pragma unroll
        for (int a = 0; a < NNE; ++a) {
          g += B[j][a] * ue[a][c];
          dg += B[j][a] * ve[a][c];
        }
        G[c][j] = g;
        dG[c][j] = dg;
      }
    const auto L = dens.linearize(G);
    dens.dP(L, dG, dP);
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        T s = T(0);
#pragma unroll
        for (int j = 0; j < D; ++j) s += dP[c][j] * B[j][a];
        ye[a][c] += w * s;
      }
  }
  T *dst = out + e * NNE * NC;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) dst[a * NC + c] = ye[a][c];
}

// Launch an element HVP + node gather for a density descriptor.  When skip is
// non-NULL and *skip != 0 at kernel start both kernels return immediately
// (used by graph-captured solver iterations after convergence).
int launch_hvp_f64(tatva_mesh *m, const tatva_density *d, const double *u, const double *v,
                   const int32_t *dof_map, double *y, hipStream_t s, const int *skip = nullptr);
int launch_hvp_f32(tatva_mesh *m, const tatva_density *d, const float *u, const float *v,
                   const int32_t *dof_map, float *y, hipStream_t s, const int *skip = nullptr);
int density_ncomp(const tatva_mesh *m, const tatva_density *d);

}  // namespace tatva
