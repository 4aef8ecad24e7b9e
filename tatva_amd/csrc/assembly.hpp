// Element residual / Hessian-vector product / Hessian diagonal, assembled into
// nodes through LDS tiles.
//
// The elements are processed in tiles of kTile (= one workgroup) taken from a
// Morton (Z-order) ordering of the element centroids, so a tile is spatially
// compact.  A workgroup computes its elements' vectors in registers and sums
// them into LDS accumulators of the tile's local nodes, one element color at
// a time (no two elements of a color share a node; colors are separated by
// barriers), then writes the tile's node sums to its slots of a partial buffer.
// A node gather adds, per node, the slots of the tiles containing it in tile
// order.  The result is deterministic with no float atomics, and HBM sees
// about (local nodes / elements) x ncomp values per element instead of
// nne x ncomp per element of a plain element-vector buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "density.hpp"
#include "element.hpp"

namespace tatva {

constexpr int kTile = 256;
constexpr int kMaxTileColors = 128;

// Device view of a mesh's tiling (built once at mesh creation).
struct TileSet {
  int32_t n_tiles;
  int32_t max_local;       // largest number of local nodes in a tile
  int32_t max_colors;      // largest number of element colors in a tile
  int32_t pad;
  const int32_t *elems;    // (n_tiles * kTile) element id per tile slot, -1 for padding
  const uint16_t *lnodes;  // (n_tiles * kTile * nne) local node index of each element node
  const int16_t *color_off;  // (n_tiles * (max_colors + 1)) start of each color in the tile
  const int32_t *node_ptr;   // (n_tiles + 1) first slot of each tile
  const int32_t *slot_ptr;   // (n_nodes + 1) node -> its slots
  const int32_t *slot_list;  // slots of each node, ascending tile
};

// ------------------------------------------------------ element kernels ----

// r_e[a][c] = sum_q w (P(G) B)_ac - f w N_a   (FP64)
template <int KIND, class Dens>
__device__ __forceinline__ void elem_residual(const double (&X)[Elem<KIND>::NNE][Elem<KIND>::D],
                                              const double (&ue)[Elem<KIND>::NNE][Dens::NC],
                                              const QuadRule &quad, const Dens &dens,
                                              double (&re)[Elem<KIND>::NNE][Dens::NC]) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) re[a][c] = 0.0;
  for (int q = 0; q < quad.nq; ++q) {
    double B[D][NNE];
    const double w = qp_geometry<KIND>(X, quad.xi[q], B) * quad.w[q];
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
}

// y_e = sum_q w B^T dP(G(u_e); G(v_e)); geometry in FP64, physics in T.
template <int KIND, class Dens, class T>
__device__ __forceinline__ void elem_hvp(const double (&X)[Elem<KIND>::NNE][Elem<KIND>::D],
                                         const T (&ue)[Elem<KIND>::NNE][Dens::NC],
                                         const T (&ve)[Elem<KIND>::NNE][Dens::NC],
                                         const QuadRule &quad, const Dens &dens,
                                         T (&ye)[Elem<KIND>::NNE][Dens::NC]) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) ye[a][c] = T(0);
  for (int q = 0; q < quad.nq; ++q) {
    double Bd[D][NNE];
    const T w = T(qp_geometry<KIND>(X, quad.xi[q], Bd) * quad.w[q]);
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
}

// diag of K_e: d_e[a][c] = sum_q w (dP(G; e_c B_a) B)_ac
template <int KIND, class Dens, class T>
__device__ __forceinline__ void elem_diag(const double (&X)[Elem<KIND>::NNE][Elem<KIND>::D],
                                          const T (&ue)[Elem<KIND>::NNE][Dens::NC],
                                          const QuadRule &quad, const Dens &dens,
                                          T (&de)[Elem<KIND>::NNE][Dens::NC]) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) de[a][c] = T(0);
  for (int q = 0; q < quad.nq; ++q) {
    double Bd[D][NNE];
    const T w = T(qp_geometry<KIND>(X, quad.xi[q], Bd) * quad.w[q]);
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
}

// ------------------------------------------------------ tile assembly ------

enum { kModeResidual = 0, kModeHvp = 1, kModeDiag = 2 };

// Body of one tile (one workgroup of kTile threads; acc = LDS of
// max_local * NC values).  u: full state; v (HVP): tangent read through
// dof_map (reduced, zero on constrained dofs).  Writes the tile's node sums to
// partial[(node_ptr[tile] + j) * NC + c].
template <int KIND, class Dens, class T, int MODE>
__device__ __forceinline__ void tile_assemble(int tile, const double *__restrict__ coords,
                                              const int32_t *__restrict__ conn,
                                              const QuadRule &quad, const Dens &dens,
                                              const T *__restrict__ u, const T *__restrict__ v,
                                              const int32_t *__restrict__ dof_map,
                                              const TileSet &ts, T *__restrict__ partial,
                                              T *__restrict__ acc) {
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
  const int32_t nb = ts.node_ptr[tile];
  const int nloc = ts.node_ptr[tile + 1] - nb;
  for (int i = threadIdx.x; i < nloc * NC; i += kTile) acc[i] = T(0);
  const int64_t slot = (int64_t)tile * kTile + threadIdx.x;
  const int32_t e = ts.elems[slot];
  T ye[NNE][NC];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) ye[a][c] = T(0);
  if (e >= 0) {
    int32_t nodes[NNE];
    double X[NNE][D];
    load_element<KIND>(e, conn, coords, nodes, X);
    T ue[NNE][NC];
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) ue[a][c] = u[(int64_t)nodes[a] * NC + c];
    if constexpr (MODE == kModeResidual) {
      elem_residual<KIND, Dens>(X, ue, quad, dens, ye);
    } else if constexpr (MODE == kModeHvp) {
      T ve[NNE][NC];
#pragma unroll
      for (int a = 0; a < NNE; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int64_t dof = (int64_t)nodes[a] * NC + c;
          if (dof_map) {
            const int32_t r = dof_map[dof];
            ve[a][c] = r >= 0 ? v[r] : T(0);
          } else {
            ve[a][c] = v[dof];
          }
        }
      elem_hvp<KIND, Dens, T>(X, ue, ve, quad, dens, ye);
    } else {
      elem_diag<KIND, Dens, T>(X, ue, quad, dens, ye);
    }
  }
  __syncthreads();
  const int16_t *off = ts.color_off + (int64_t)tile * (ts.max_colors + 1);
  const uint16_t *ln = ts.lnodes + slot * NNE;
  for (int c = 0; c < ts.max_colors; ++c) {
    const int lo = off[c], hi = off[c + 1];
    if (lo >= hi) break;  // colors are numbered 0..ncolors-1 without gaps
    if ((int)threadIdx.x >= lo && (int)threadIdx.x < hi) {
#pragma unroll
      for (int a = 0; a < NNE; ++a) {
        T *dst = acc + (int)ln[a] * NC;
#pragma unroll
        for (int k = 0; k < NC; ++k) dst[k] += ye[a][k];
      }
    }
    __syncthreads();
  }
  T *out = partial + (int64_t)nb * NC;
  for (int i = threadIdx.x; i < nloc * NC; i += kTile) out[i] = acc[i];
}

template <int KIND, class Dens, class T, int MODE>
__global__ __launch_bounds__(kTile) void k_tile_assemble(const double *__restrict__ coords,
                                                         const int32_t *__restrict__ conn,
                                                         QuadRule quad, Dens dens,
                                                         const T *__restrict__ u,
                                                         const T *__restrict__ v,
                                                         const int32_t *__restrict__ dof_map,
                                                         TileSet ts, T *__restrict__ partial,
                                                         const int *__restrict__ skip) {
  if (skip && *skip) return;
  extern __shared__ unsigned char lds_raw[];
  tile_assemble<KIND, Dens, T, MODE>((int)xcd_block(blockIdx.x, gridDim.x), coords, conn, quad,
                                     dens, u, v, dof_map, ts, partial,
                                     reinterpret_cast<T *>(lds_raw));
}

// Sum of node i's tile slots (ascending tile).
template <int NC, class T>
__device__ __forceinline__ void slot_sum(int64_t i, const int32_t *__restrict__ slot_ptr,
                                         const int32_t *__restrict__ slot_list,
                                         const T *__restrict__ partial, T (&acc)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = T(0);
  const int32_t b = slot_ptr[i], en = slot_ptr[i + 1];
  for (int32_t k = b; k < en; ++k) {
    const T *src = partial + (int64_t)slot_list[k] * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += src[c];
  }
}

// out[map(node*NC + c)] = sum of the node's tile slots (ascending tile) - f_ext.
template <int NC, class T>
__global__ __launch_bounds__(kBlock) void k_slot_gather(int64_t n_nodes, const int32_t *__restrict__ slot_ptr,
                                                        const int32_t *__restrict__ slot_list,
                                                        const T *__restrict__ partial,
                                                        const double *__restrict__ f_ext,
                                                        const int32_t *__restrict__ dof_map,
                                                        T *__restrict__ out,
                                                        const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  T acc[NC];
  slot_sum<NC, T>(i, slot_ptr, slot_list, partial, acc);
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const int64_t dof = i * NC + c;
    T val = acc[c];
    if (f_ext) val -= T(f_ext[dof]);
    if (dof_map) {
      const int32_t r = dof_map[dof];
      if (r >= 0) out[r] = val;
    } else {
      out[dof] = val;
    }
  }
}

}  // namespace tatva
