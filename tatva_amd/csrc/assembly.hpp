// Element residual / Hessian-vector product / Hessian diagonal, assembled into
// nodes through LDS tiles.
//
// The elements are processed in tiles of kTile (= one workgroup) taken from a
// Morton (Z-order) ordering of the element centroids, so a tile is spatially
// compact.  A workgroup loads its local nodes' data into LDS once, computes its
// elements' vectors in registers and sums them per local node, then writes the
// tile's node sums to its slots of a partial buffer.  The per-tile sum runs
// one element color at a time into LDS accumulators (no two elements of a
// color share a node; colors are separated by barriers) for Tri3 and Hex8
// (6-13 colors a tile); Tet4 tiles have ~30 colors, so there every element
// writes its vector to LDS and each local node adds its entries in ascending
// (slot, node) order from a per-tile incidence list (two barriers).
// A node gather adds, per node, the slots of the tiles containing it in tile
// order.  The result is deterministic with no float atomics, and HBM sees
// about (local nodes / elements) x ncomp values per element instead of
// nne x ncomp per element of a plain element-vector buffer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

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
  const uint16_t *lnodes;  // (n_tiles * kTile * nne) local node index of each element node
  const int16_t *color_off;  // (n_tiles * (max_colors + 1)) start of each color in the tile
  const int32_t *node_ptr;   // (n_tiles + 1) first slot of each tile
  const int32_t *slot_ptr;   // (n_nodes + 1) node -> its slots
  const int32_t *slot_pos;   // (n_slots) node-major position of each tile slot's sums
  const int32_t *tnodes;     // (n_slots) global node of each slot
  // Tet4 only: entries (slot * nne + a) of each local node, ascending, at
  // inc_list[tile * kTile * nne + ...]; local node j ends at inc_end[node_ptr + j]
  const uint16_t *inc_end;   // (n_slots)
  const uint16_t *inc_list;  // (n_tiles * kTile * nne)
  const int32_t *elem;       // (n_tiles * kTile) element of each tile slot (-1: padding)
};

// Element work of a tile kernel.  kModeLinearize stores the density's
// linearisation (Dens::kLin values per quadrature point, from the state u) in a
// cache and writes no output; kModeHvpLin is kModeHvp reading that cache instead
// of re-linearising at every call (same arithmetic, same result), for solvers
// whose state is fixed over many products.
enum { kModeResidual = 0, kModeHvp = 1, kModeDiag = 2, kModeLinearize = 3, kModeHvpLin = 4 };

// Cache layout: value k of quadrature point q of the element in tile slot t of
// tile b at lin[((b * nq + q) * kLin + k) * kTile + t] (a tile's threads read
// consecutive floats).

// Element geometry variants of the tile kernels.  kGeomGeneral: computed at
// every quadrature point from the node coordinates.  kGeomAffine: a Hex8
// parallelepiped, its Jacobian inverted once per element.  kGeomCache: an
// affine Hex8 mesh whose geometry is read from a cache written once per solve
// (FP32 HVP / linearisation only).
enum { kGeomGeneral = 0, kGeomAffine = 1, kGeomCache = 2 };

// Geometry cache layout (affine Hex8): value k of the element in tile slot t
// of tile b at geo[(b * kGeoFloats + k) * kTile + t] (a tile's threads read
// consecutive floats): the FP32 cast of the FP64 Jacobian inverse (9 values,
// row-major) and the FP64 det J as two 32-bit words, the values the kernels
// computing the geometry themselves use.
constexpr int kGeoFloats = 11;

// The p and x updates of an FP32 conjugate-gradient iteration, which its step
// kernel leaves to the next operator launch (nothing reads p or x in between):
// p = z + b p (p = z on the first iteration) and x = x + a p.  The step writes
// the coefficients; on: 0 no update pending, 1 first iteration, 2 otherwise.
struct CgDeferCoef {
  float a, b;
  int on, pad;
};
struct CgDefer {
  const float *z;  // the z the update belongs to (the previous iteration's)
  float *p, *x;
  const CgDeferCoef *coef;
  int64_t n;  // 0: the launch carries no update
};

typedef float cg_f4 __attribute__((ext_vector_type(4)));

// A tile kernel carrying a deferred update runs it in workgroups of its own,
// interleaved with the tiles' so that the streams spread over the whole launch
// and no tile waits for them: after every kDeferRows rows of kNumXcd tile
// workgroups comes one row of kNumXcd streaming workgroups (a row holds one
// workgroup per XCD, so the tiles keep their XCD mapping).
constexpr int kDeferRows = 8;
inline int tile_grid(int n_tiles, bool defer) {
  if (!defer) return n_tiles;
  const int per = kNumXcd * kDeferRows;
  return (n_tiles + per - 1) / per * (per + kNumXcd);
}

// Streaming workgroup sb of S: the contiguous range [sb nv / S, (sb + 1) nv / S)
// of the nv 16-byte vectors (16-byte aligned), kDeferUnroll vectors per thread
// in flight; workgroup 0 also takes the last n % 4 entries.  The expressions are
// those of the step kernel that would otherwise apply the update, so the
// iterates are the same bits.
constexpr int kDeferUnroll = 4;
__device__ __forceinline__ void cg_defer_stream(const CgDefer &dx, int64_t sb, int64_t S) {
  const CgDeferCoef c = *dx.coef;
  if (c.on == 0) return;
  const bool first = c.on == 1;
  const float a = c.a, b = c.b;
  const int64_t nv = dx.n / 4;
  const int64_t v0 = sb * nv / S, v1 = (sb + 1) * nv / S;
  const cg_f4 *z4 = reinterpret_cast<const cg_f4 *>(dx.z);
  cg_f4 *p4 = reinterpret_cast<cg_f4 *>(dx.p);
  cg_f4 *x4 = reinterpret_cast<cg_f4 *>(dx.x);
  constexpr int U = kDeferUnroll;
  for (int64_t base = v0 + threadIdx.x; base < v1; base += U * (int64_t)blockDim.x) {
    cg_f4 zv[U], pv[U], xv[U];
    // p is loaded on the first iteration too (and not used), so that no load
    // waits for another
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t iv = base + k * (int64_t)blockDim.x;
      if (iv < v1) {
        zv[k] = z4[iv];
        pv[k] = p4[iv];
        xv[k] = x4[iv];
      }
    }
#pragma unroll
    for (int k = 0; k < U; ++k) {
      const int64_t iv = base + k * (int64_t)blockDim.x;
      if (iv < v1) {
        cg_f4 pn, xn;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          pn[e] = first ? zv[k][e] : zv[k][e] + b * pv[k][e];
          xn[e] = xv[k][e] + a * pn[e];
        }
        __builtin_nontemporal_store(pn, p4 + iv);
        __builtin_nontemporal_store(xn, x4 + iv);
      }
    }
  }
  if (sb == 0 && threadIdx.x < dx.n - nv * 4) {
    const int64_t i = nv * 4 + threadIdx.x;
    const float zi = dx.z[i], pi = dx.p[i];
    const float pn = first ? zi : zi + b * pi;
    dx.p[i] = pn;
    dx.x[i] = dx.x[i] + a * pn;
  }
}

// Per-tile sum by the incidence lists instead of element colors: Tet4 (~30
// colors per tile) and Hex8 (8-15; 160^3 Neo-Hookean HVP 335.7 -> 301.7 us
// against the per-color sums, whose barriers and range tests it drops).
constexpr bool kHex8Incidence = true;
constexpr bool tile_gather_kind(int kind) {
  return kind == TATVA_TET4 || (kind == TATVA_HEX8 && kHex8Incidence);
}
template <int KIND>
constexpr bool tile_gather() {
  return tile_gather_kind(KIND);
}

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
// MODE kModeLinearize stores the linearisation at G(u_e) to lin (point q's
// values from lin + q * kLin * kTile, stride kTile) and leaves ye zero;
// kModeHvpLin reads it from there instead of linearising at G(u_e).
template <int KIND, class Dens, class T, int MODE = kModeHvp>
__device__ __forceinline__ void elem_hvp(const double (&X)[Elem<KIND>::NNE][Elem<KIND>::D],
                                         const T (&ue)[Elem<KIND>::NNE][Dens::NC],
                                         const T (&ve)[Elem<KIND>::NNE][Dens::NC],
                                         const QuadRule &quad, const Dens &dens,
                                         T (&ye)[Elem<KIND>::NNE][Dens::NC],
                                         T *lin = nullptr) {
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
          if constexpr (MODE != kModeHvpLin) g += B[j][a] * ue[a][c];
          if constexpr (MODE != kModeLinearize) dg += B[j][a] * ve[a][c];
        }
        G[c][j] = g;
        dG[c][j] = dg;
      }
    T *lq = (MODE == kModeLinearize || MODE == kModeHvpLin)
                ? lin + (int64_t)q * Dens::kLin * kTile : nullptr;
    if constexpr (MODE == kModeLinearize) {
      dens.store_lin(dens.linearize(G), lq, kTile);
      continue;
    }
    if constexpr (MODE == kModeHvpLin) {
      dens.dP(dens.load_lin(lq, kTile), dG, dP);
    } else {
      const auto L = dens.linearize(G);
      dens.dP(L, dG, dP);
    }
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


// ------------------------------------------------- Hex8 in modal form -----
//
// A trilinear Hex8 field f = sum_a N_a f_a equals (1/8) sum_m M_m psi_m(xi)
// with psi = {1, xi, eta, xi eta, zeta, xi zeta, eta zeta, xi eta zeta}
// (mode index b = i + 2j + 4k over the node sign pattern) and M = H f, H the
// 8x8 matrix of node sign products H[m][a] = prod_{d in m} s_d(a)
// (element/base.py:478-491,535); node values from modes use H^T.
// A natural derivative at a point then costs 3 FMAs instead of 8, and the
// element output sum_q B^T S_q is one butterfly of 7 accumulated modes.  An
// element whose bilinear coordinate modes are all zero (a parallelepiped,
// e.g. every cell of an axis-aligned structured grid; the butterfly along x
// first makes them exactly zero there) has a constant Jacobian, inverted once
// per element.  The results equal the nodal formulas up to rounding.

// Hex8 local node -> sign index i + 2j + 4k (sx, sy, sz > 0 -> i, j, k = 1).
__host__ __device__ constexpr int hex_bin(int a) {
  return a == 2 ? 3 : a == 3 ? 2 : a == 6 ? 7 : a == 7 ? 6 : a;
}

// In-place sign-product transform v <- H v (unnormalised), stages x, y, z.
template <class S>
__device__ __forceinline__ void hadamard8(S (&v)[8]) {
#pragma unroll
  for (int h = 1; h < 8; h <<= 1)
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (!(b & h)) {
        const S lo = v[b], hi = v[b | h];
        v[b] = hi + lo;
        v[b | h] = hi - lo;
      }
}

// Transpose of hadamard8 (node value from modes: lo - hi at the minus node,
// lo + hi at the plus node, per direction).
template <class S>
__device__ __forceinline__ void hadamard8_t(S (&v)[8]) {
#pragma unroll
  for (int h = 1; h < 8; h <<= 1)
#pragma unroll
    for (int b = 0; b < 8; ++b)
      if (!(b & h)) {
        const S lo = v[b], hi = v[b | h];
        v[b] = lo - hi;
        v[b | h] = lo + hi;
      }
}

// M[c][b] = (1/8) (H f_c)[b]
template <class S, int NC>
__device__ __forceinline__ void hex_modes(const S (&f)[8][NC], S (&M)[NC][8]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    S v[8];
#pragma unroll
    for (int a = 0; a < 8; ++a) v[hex_bin(a)] = f[a][c];
    hadamard8(v);
#pragma unroll
    for (int b = 0; b < 8; ++b) M[c][b] = v[b] * S(0.125);
  }
}

// d[k] = d f / d xi_k at (x, y, z) for a modal field M (scaled by 1/8)
template <class S>
__device__ __forceinline__ void hex_dmodes(const S (&M)[8], S x, S y, S z, S (&d)[3]) {
  d[0] = M[1] + M[3] * y + M[5] * z + M[7] * (y * z);
  d[1] = M[2] + M[3] * x + M[6] * z + M[7] * (x * z);
  d[2] = M[4] + M[5] * x + M[6] * y + M[7] * (x * y);
}

// The same with the point's products given (yz = y * z, xz = x * z, xy = x * y).
template <class S>
__device__ __forceinline__ void hex_dmodes(const S (&M)[8], S x, S y, S z, S yz, S xz, S xy,
                                           S (&d)[3]) {
  d[0] = M[1] + M[3] * y + M[5] * z + M[7] * yz;
  d[1] = M[2] + M[3] * x + M[6] * z + M[7] * xz;
  d[2] = M[4] + M[5] * x + M[6] * y + M[7] * xy;
}

// Affine Hex8: det J (FP64) and the FP32 cast of J^-1, from the coordinate
// modes (J rows = modes 1, 2, 4).  Plain multiplies and adds, no fused
// multiply-add contraction, so the geometry cache kernel and the tile kernels
// get the same bits.
template <class T>
__device__ __forceinline__ double hex8_affine_inverse(const double (&X)[8][3], T (&iv)[3][3]) {
#pragma clang fp contract(off)
  double XM[3][8];
  hex_modes<double, 3>(X, XM);
  double J[3][3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    J[0][j] = XM[j][1];
    J[1][j] = XM[j][2];
    J[2][j] = XM[j][4];
  }
  const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const double r = 1.0 / det;
  iv[0][0] = T(c00 * r);
  iv[1][0] = T(c01 * r);
  iv[2][0] = T(c02 * r);
  iv[0][1] = T((J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r);
  iv[1][1] = T((J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r);
  iv[2][1] = T((J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r);
  iv[0][2] = T((J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r);
  iv[1][2] = T((J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r);
  iv[2][2] = T((J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r);
  return det;
}

// Hex8 residual (MODE = kModeResidual: S = P(G), plus the Poisson source) or
// Hessian-vector product (S = dP(G(u); G(v))): ye = sum_q w_q B_q^T S_q.
// Geometry in FP64, physics in T.
// MODE kModeLinearize / kModeHvpLin: as elem_hvp (cache at lin, stride kTile).
// GEO (affine only): J^-1 and det J read from geo (this element's geometry
// cache, stride kTile) instead of computed from X.
template <class Dens, class T, int MODE, bool AFFINE, bool GEO = false>
__device__ __forceinline__ void hex8_element(const double (&X)[8][3], const T (&ue)[8][Dens::NC],
                                             const T (&ve)[8][Dens::NC], const QuadRule &quad,
                                             const Dens &dens, T (&ye)[8][Dens::NC],
                                             T *lin = nullptr, const T *geo = nullptr) {
  static_assert(!GEO || AFFINE, "the Hex8 geometry cache holds one inverse per element");
  constexpr int NC = Dens::NC;
  constexpr bool kUsesU = MODE != kModeHvpLin;
  constexpr bool kUsesV = MODE == kModeHvp || MODE == kModeHvpLin;
  // kModeHvpLin: the cached linearisation of point q + 1 is loaded while point q
  // is computed (a point's values are needed only after its tangent gradient)
  typename Dens::Lin Lnext{};
  if constexpr (MODE == kModeHvpLin) Lnext = dens.load_lin(lin, kTile);
  double XM[3][8];  // coordinate modes (general elements)
  if constexpr (!AFFINE) hex_modes<double, 3>(X, XM);
  T UM[NC][8], VM[NC][8];
  if constexpr (kUsesU) hex_modes<T, NC>(ue, UM);
  if constexpr (kUsesV) hex_modes<T, NC>(ve, VM);
  T A[NC][8];
#pragma unroll
  for (int c = 0; c < NC; ++c)
#pragma unroll
    for (int b = 0; b < 8; ++b) A[c][b] = T(0);
  // affine: the Jacobian inverse is the same at every point, held in T (the
  // values each point would cast from the FP64 inverse)
  T iv0[3][3] = {};
  double det0 = 0.0;
  if constexpr (GEO) {
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) iv0[i][j] = geo[(i * 3 + j) * kTile];
    const uint32_t lo = __float_as_uint(float(geo[9 * kTile]));
    const uint32_t hi = __float_as_uint(float(geo[10 * kTile]));
    det0 = __hiloint2double((int)hi, (int)lo);
  } else if constexpr (AFFINE) {
    det0 = hex8_affine_inverse<T>(X, iv0);
  }
  for (int q = 0; q < quad.nq; ++q) {
    const double x = quad.xi[q][0], y = quad.xi[q][1], z = quad.xi[q][2];
    // the point in T and its products: uniform values (FP32: from the kernel
    // argument, so they stay scalar operands)
    T xt, yt, zt, yz, xz, xy, xyz;
    if constexpr (sizeof(T) == 4) {
      xt = quad.xf[q][0], yt = quad.xf[q][1], zt = quad.xf[q][2];
      yz = quad.xf[q][3], xz = quad.xf[q][4], xy = quad.xf[q][5], xyz = quad.xf[q][6];
    } else {
      xt = T(x), yt = T(y), zt = T(z);
      yz = yt * zt, xz = xt * zt, xy = xt * yt, xyz = xt * yt * zt;
    }
    double inv[3][3], det;
    if constexpr (AFFINE) {
      det = det0;
    } else {
      double J[3][3];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double d[3];
        hex_dmodes<double>(XM[j], x, y, z, d);
        J[0][j] = d[0];
        J[1][j] = d[1];
        J[2][j] = d[2];
      }
      det = det_inv(J, inv);
    }
    const T w = T(det * quad.w[q]);
    T iv[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) iv[i][j] = AFFINE ? iv0[i][j] : T(inv[i][j]);
    // G[c][j] = sum_k H[c][k] inv[j][k], H = natural derivatives
    T G[NC][3], S[NC][3];
    if constexpr (kUsesU) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        T h[3];
        hex_dmodes<T>(UM[c], xt, yt, zt, yz, xz, xy, h);
#pragma unroll
        for (int j = 0; j < 3; ++j) G[c][j] = h[0] * iv[j][0] + h[1] * iv[j][1] + h[2] * iv[j][2];
      }
    }
    T *lq = (MODE == kModeLinearize || MODE == kModeHvpLin)
                ? lin + (int64_t)q * Dens::kLin * kTile : nullptr;
    if constexpr (MODE == kModeLinearize) {
      dens.store_lin(dens.linearize(G), lq, kTile);
      continue;
    }
    if constexpr (MODE == kModeResidual) {
      dens.P(G, S);
    } else {
      T dG[NC][3];
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        T h[3];
        hex_dmodes<T>(VM[c], xt, yt, zt, yz, xz, xy, h);
#pragma unroll
        for (int j = 0; j < 3; ++j) dG[c][j] = h[0] * iv[j][0] + h[1] * iv[j][1] + h[2] * iv[j][2];
      }
      // S = w dP(G; dG): the weight folded into the density's moduli
      if constexpr (MODE == kModeHvpLin) {
        const typename Dens::Lin Lq = Lnext;
        if (q + 1 < quad.nq) Lnext = dens.load_lin(lq + Dens::kLin * kTile, kTile);
        dens.dPw(Lq, dG, w, S);
      } else {
        const auto L = dens.linearize(G);
        dens.dPw(L, dG, w, S);
      }
    }
    // Q[c][k] = w sum_j P[c][j] inv[j][k] (residual) / sum_j S[c][j] inv[j][k],
    // accumulated into the derivative modes
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      T Q[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const T sk = S[c][0] * iv[0][k] + S[c][1] * iv[1][k] + S[c][2] * iv[2][k];
        Q[k] = MODE == kModeResidual ? w * sk : sk;
      }
      A[c][1] += Q[0];
      A[c][2] += Q[1];
      A[c][4] += Q[2];
      A[c][3] += Q[0] * yt + Q[1] * xt;
      A[c][5] += Q[0] * zt + Q[2] * xt;
      A[c][6] += Q[1] * zt + Q[2] * yt;
      A[c][7] += Q[0] * yz + Q[1] * xz + Q[2] * xy;
    }
    if constexpr (MODE == kModeResidual && Dens::kSource) {
      // - w f N_a with N_a = (1/8) sum_m phi_m(a) psi_m
      const T s = w * dens.f;
      A[0][0] -= s;
      A[0][1] -= s * xt;
      A[0][2] -= s * yt;
      A[0][3] -= s * xy;
      A[0][4] -= s * zt;
      A[0][5] -= s * xz;
      A[0][6] -= s * yz;
      A[0][7] -= s * xyz;
    }
  }
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    T v[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) v[b] = A[c][b];
    hadamard8_t(v);
#pragma unroll
    for (int a = 0; a < 8; ++a) ye[a][c] = v[hex_bin(a)] * T(0.125);
  }
}

// True when the element's bilinear coordinate modes are all exactly zero
// (computed with the same butterfly as hex_modes).
inline bool hex_is_affine_host(const double (&X)[8][3]) {
  for (int j = 0; j < 3; ++j) {
    double v[8];
    for (int a = 0; a < 8; ++a) v[hex_bin(a)] = X[a][j];
    for (int h = 1; h < 8; h <<= 1)
      for (int b = 0; b < 8; ++b)
        if (!(b & h)) {
          const double lo = v[b], hi = v[b | h];
          v[b] = hi + lo;
          v[b | h] = hi - lo;
        }
    if (v[3] != 0.0 || v[5] != 0.0 || v[6] != 0.0 || v[7] != 0.0) return false;
  }
  return true;
}

// ------------------------------------------------------ tile assembly ------

// LDS of one tile: per local node its coordinates (FP64), then the
// accumulator, the state and (HVP) the tangent in a T of tsize bytes.  With
// the incidence-list sum the element vectors (kTile x nne x NC) reuse that
// space, and the tile's incidence lists follow it.
constexpr size_t align16(size_t x) { return (x + 15) & ~size_t(15); }
constexpr size_t tile_stage_bytes(int max_local, int D, int NC, size_t tsize, int mode) {
  return (size_t)max_local *
         (D * sizeof(double) + ((mode == kModeHvp || mode == kModeHvpLin) ? 3 : 2) * NC * tsize);
}
// Values between two threads' element vectors in the incidence-list sum: nne*NC
// rounded up to an odd count, so that the threads of a wave writing their
// vectors hit distinct LDS banks (with 12 the lanes t and t + 8 share one).
constexpr int tile_elem_stride(int nne, int NC) { return nne * NC + ((nne * NC) % 2 == 0 ? 1 : 0); }
constexpr size_t tile_union_bytes(int max_local, int D, int nne, int NC, size_t tsize, int mode) {
  const size_t staged = tile_stage_bytes(max_local, D, NC, tsize, mode);
  const size_t elem = (size_t)kTile * tile_elem_stride(nne, NC) * tsize;
  return align16(elem > staged ? elem : staged);
}
// the color sum's copy of the tile's color offsets, after the staged nodes
constexpr size_t tile_offs_bytes() { return align16(sizeof(int16_t) * (kMaxTileColors + 1)); }
constexpr size_t tile_lds_bytes(int kind, int max_local, int D, int nne, int NC, size_t tsize,
                                int mode) {
  if (!tile_gather_kind(kind)) return align16(tile_stage_bytes(max_local, D, NC, tsize, mode)) + tile_offs_bytes();
  return tile_union_bytes(max_local, D, nne, NC, tsize, mode) + (size_t)kTile * nne * sizeof(uint16_t) +
         align16((size_t)max_local * sizeof(uint16_t));
}
constexpr size_t kTileLdsMax = 65536;  // LDS one workgroup may allocate

// Body of one tile (one workgroup of kTile threads; lds of tile_lds_bytes).
// GEOM (kGeomGeneral / kGeomAffine / kGeomCache): how the elements' geometry
// is taken; with kGeomCache the FP32 HVP and linearisation read it from geo
// instead of the node coordinates.  u: full state; v (HVP): tangent read through dof_map
// (reduced, zero on constrained dofs).  The tile's local nodes are loaded into
// LDS once (one thread per node), so each element reads its nodes from LDS
// instead of gathering them from global memory.  Writes the sums of local node
// j to partial[slot_pos[node_ptr[tile] + j] * NC + c] (node-major order).
template <int KIND, class Dens, class T, int MODE, int GEOM>
__device__ __forceinline__ void tile_assemble(int tile, const double *__restrict__ coords,
                                              const QuadRule &quad, const Dens &dens,
                                              const T *__restrict__ u, const T *__restrict__ v,
                                              const int32_t *__restrict__ dof_map,
                                              const TileSet &ts, T *__restrict__ partial,
                                              unsigned char *__restrict__ lds,
                                              T *__restrict__ lin = nullptr,
                                              const T *__restrict__ geo = nullptr) {
  constexpr bool kUsesU = MODE != kModeHvpLin;
  constexpr bool kUsesV = MODE == kModeHvp || MODE == kModeHvpLin;
  using El = Elem<KIND>;
  constexpr int D = El::D, NNE = El::NNE, NC = Dens::NC;
  // a geometry cache (affine Hex8 HVP / linearisation) replaces the coordinates
  constexpr bool kAffine = KIND == TATVA_HEX8 && GEOM != kGeomGeneral;
  constexpr bool kGeo = KIND == TATVA_HEX8 && GEOM == kGeomCache && sizeof(T) == 4 &&
                        (MODE == kModeHvp || MODE == kModeHvpLin || MODE == kModeLinearize);
  constexpr bool uses_x = !kGeo;
  const int32_t nb = ts.node_ptr[tile];
  const int nloc = ts.node_ptr[tile + 1] - nb;
  double *Xs = reinterpret_cast<double *>(lds);
  T *acc = reinterpret_cast<T *>(Xs + (size_t)ts.max_local * D);
  T *us = acc + (size_t)ts.max_local * NC;
  T *vs = us + (size_t)ts.max_local * NC;
  // Stage the local nodes, two per thread at a time (the second one is the
  // first again past the end: the same values stored twice).  Every load of
  // the pair is issued before the first wait: the global node ids, then the
  // coordinates, state and dof map entries, then the tangent through the map
  // (a constrained dof, r < 0, reads a valid dummy address and stores 0, so
  // that no load sits behind a branch).
  const T *dummy = reinterpret_cast<const T *>(coords);
  const auto stage = [&](auto mapped) {
    constexpr bool kMapped = decltype(mapped)::value;
    for (int j0 = threadIdx.x; j0 < nloc; j0 += 2 * kTile) {
      const int jj[2] = {j0, j0 + kTile < nloc ? j0 + kTile : j0};
      int64_t g[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) g[h] = ts.tnodes[nb + jj[h]];
      double xs[2][D];
      T uv[2][NC], vv[2][NC];
      int64_t r[2][NC];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        if constexpr (uses_x) {
#pragma unroll
          for (int k = 0; k < D; ++k) xs[h][k] = coords[g[h] * D + k];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          const int64_t dof = g[h] * NC + c;
          if constexpr (kUsesU) uv[h][c] = u[dof];
          if constexpr (kUsesV) r[h][c] = kMapped ? (int64_t)dof_map[dof] : dof;
        }
      }
      if constexpr (kUsesV) {
#pragma unroll
        for (int h = 0; h < 2; ++h)
#pragma unroll
          for (int c = 0; c < NC; ++c) {
            const T val = *(r[h][c] >= 0 ? v + r[h][c] : dummy);
            vv[h][c] = r[h][c] >= 0 ? val : T(0);
          }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int j = jj[h];
        if constexpr (uses_x) {
#pragma unroll
          for (int k = 0; k < D; ++k) Xs[j * D + k] = xs[h][k];
        }
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          acc[j * NC + c] = T(0);
          if constexpr (kUsesU) us[j * NC + c] = uv[h][c];
          if constexpr (kUsesV) vs[j * NC + c] = vv[h][c];
        }
      }
    }
  };
  if (kUsesV && dof_map) stage(std::true_type{});
  else stage(std::false_type{});
  // Tet4: the tile's incidence lists are loaded here and stored to LDS after
  // the element work, so their latency overlaps it (Hex8: loaded after the
  // element work, whose registers are all in use)
  using IncVec = std::conditional_t<NNE == 8, uint4, uint2>;  // NNE uint16 entries
  constexpr bool kIncEarly = NNE == 4;
  IncVec inc_pk{};
  uint32_t inc_e = 0;
  if constexpr (tile_gather<KIND>() && kIncEarly) {
    inc_pk = reinterpret_cast<const IncVec *>(ts.inc_list + (int64_t)tile * kTile * NNE)[threadIdx.x];
    if ((int)threadIdx.x < nloc) inc_e = ts.inc_end[nb + threadIdx.x];
  }
  const int64_t slot = (int64_t)tile * kTile + threadIdx.x;
  const int16_t *off = ts.color_off + (int64_t)tile * (ts.max_colors + 1);
  // the tile's elements fill its first off[max_colors] slots (the rest is padding)
  const bool real = (int)threadIdx.x < (int)off[ts.max_colors];
  // the color sum reads the offsets from LDS (copied with the nodes)
  int16_t *offs = reinterpret_cast<int16_t *>(
      lds + align16(tile_stage_bytes(ts.max_local, D, NC, sizeof(T), MODE)));
  if constexpr (!tile_gather<KIND>() && MODE != kModeLinearize) {
    for (int c = threadIdx.x; c <= ts.max_colors; c += kTile) offs[c] = off[c];
  }
  // the element's local node numbers (padding slots hold valid entries too):
  // one 16- / 8-byte load for Hex8 / Tet4
  int l[NNE];
  if constexpr (NNE == 8 || NNE == 4) {
    using Vec = std::conditional_t<NNE == 8, uint4, uint2>;
    const Vec pk = reinterpret_cast<const Vec *>(ts.lnodes)[slot];
    const uint32_t w[4] = {pk.x, pk.y, NNE == 8 ? ((const uint32_t *)&pk)[2] : 0u,
                           NNE == 8 ? ((const uint32_t *)&pk)[3] : 0u};
#pragma unroll
    for (int a = 0; a < NNE; ++a) l[a] = real ? (int)((w[a >> 1] >> (16 * (a & 1))) & 0xffffu) : 0;
  } else {
    const uint16_t *ln = ts.lnodes + slot * NNE;
#pragma unroll
    for (int a = 0; a < NNE; ++a) l[a] = real ? (int)ln[a] : 0;
  }
  T ye[NNE][NC];
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int c = 0; c < NC; ++c) ye[a][c] = T(0);
  // per-element material parameters: the density of this slot's element
  const Dens de = (dens.ep && real) ? dens.at(ts.elem[slot]) : dens;
  __syncthreads();
  if (real) {
    double X[NNE][D];
    T ue[NNE][NC], ve[NNE][NC];
#pragma unroll
    for (int a = 0; a < NNE; ++a) {
#pragma unroll
      for (int k = 0; k < D; ++k) X[a][k] = uses_x ? Xs[l[a] * D + k] : 0.0;
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        ue[a][c] = kUsesU ? us[l[a] * NC + c] : T(0);
        ve[a][c] = kUsesV ? vs[l[a] * NC + c] : T(0);
      }
    }
    // this thread's cached linearisation (kModeLinearize / kModeHvpLin)
    T *lp = lin ? lin + (int64_t)tile * quad.nq * Dens::kLin * kTile + threadIdx.x : nullptr;
    const T *gp = kGeo ? geo + (int64_t)tile * kGeoFloats * kTile + threadIdx.x : nullptr;
    if constexpr (KIND == TATVA_HEX8 && MODE == kModeResidual) {
      hex8_element<Dens, T, kModeResidual, kAffine>(X, ue, ue, quad, de, ye);
    } else if constexpr (MODE == kModeResidual) {
      elem_residual<KIND, Dens>(X, ue, quad, de, ye);
    } else if constexpr (MODE == kModeHvp || MODE == kModeHvpLin || MODE == kModeLinearize) {
      if constexpr (KIND == TATVA_HEX8)
        hex8_element<Dens, T, MODE, kAffine, kGeo>(X, ue, ve, quad, de, ye, lp, gp);
      else elem_hvp<KIND, Dens, T, MODE>(X, ue, ve, quad, de, ye, lp);
    } else {
      elem_diag<KIND, Dens, T>(X, ue, quad, de, ye);
    }
  }
  if constexpr (MODE == kModeLinearize) return;  // the cache is the output
  if constexpr (tile_gather<KIND>()) {
    if constexpr (!kIncEarly) {
      inc_pk = reinterpret_cast<const IncVec *>(ts.inc_list + (int64_t)tile * kTile * NNE)[threadIdx.x];
      if ((int)threadIdx.x < nloc) inc_e = ts.inc_end[nb + threadIdx.x];
    }
    __syncthreads();  // every element has read its nodes: the LDS is free
    T *ys = reinterpret_cast<T *>(lds);
    uint16_t *il = reinterpret_cast<uint16_t *>(
        lds + tile_union_bytes(ts.max_local, D, NNE, NC, sizeof(T), MODE));
    uint16_t *ie = il + kTile * NNE;
    constexpr int kYS = tile_elem_stride(NNE, NC);
#pragma unroll
    for (int a = 0; a < NNE; ++a)
#pragma unroll
      for (int c = 0; c < NC; ++c) ys[(int)threadIdx.x * kYS + a * NC + c] = ye[a][c];
    reinterpret_cast<IncVec *>(il)[threadIdx.x] = inc_pk;
    if ((int)threadIdx.x < nloc) ie[threadIdx.x] = (uint16_t)inc_e;
    for (int j = threadIdx.x + kTile; j < nloc; j += kTile) ie[j] = ts.inc_end[nb + j];
    __syncthreads();
    // one local node per thread, its components together: one read of each
    // incidence entry serves all NC sums (each still in ascending entry order)
    for (int j = threadIdx.x; j < nloc; j += kTile) {
      const int q0 = j > 0 ? (int)ie[j - 1] : 0;
      const int q1 = ie[j];
      T sum[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) sum[c] = T(0);
      for (int q = q0; q < q1; ++q) {
        const int e = il[q];  // entry slot * NNE + a
        const int base = (e / NNE) * kYS + (e % NNE) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) sum[c] += ys[base + c];
      }
      const int64_t pos = (int64_t)ts.slot_pos[nb + j] * NC;
#pragma unroll
      for (int c = 0; c < NC; ++c) partial[pos + c] = sum[c];
    }
    return;
  }
  for (int c = 0; c < ts.max_colors; ++c) {
    const int lo = offs[c], hi = offs[c + 1];
    if (lo >= hi) break;  // colors are numbered 0..ncolors-1 without gaps
    if ((int)threadIdx.x >= lo && (int)threadIdx.x < hi) {
#pragma unroll
      for (int a = 0; a < NNE; ++a) {
        T *dst = acc + l[a] * NC;
#pragma unroll
        for (int k = 0; k < NC; ++k) dst[k] += ye[a][k];
      }
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < nloc * NC; i += kTile) {
    const int j = i / NC;
    partial[(int64_t)ts.slot_pos[nb + j] * NC + (i - j * NC)] = acc[i];
  }
}

template <int KIND, class Dens, class T, int MODE, int GEOM>
__global__ __launch_bounds__(kTile) void k_tile_assemble(const double *__restrict__ coords,
                                                         QuadRule quad, Dens dens,
                                                         const T *__restrict__ u,
                                                         const T *__restrict__ v,
                                                         const int32_t *__restrict__ dof_map,
                                                         TileSet ts, T *__restrict__ partial,
                                                         const int *__restrict__ skip, int rev,
                                                         T *__restrict__ lin,
                                                         const T *__restrict__ geo, CgDefer dx) {
  if (skip && *skip) return;
  extern __shared__ unsigned char lds_raw[];
  // dx (FP32 HVP): a CG step's p and x streams, run by the launch's streaming
  // workgroups on the memory bandwidth the tiles leave idle (grid of tile_grid)
  int b = blockIdx.x;
  if constexpr (sizeof(T) == 4 && (MODE == kModeHvp || MODE == kModeHvpLin)) {
    if (dx.n) {
      constexpr int per = kNumXcd * kDeferRows;
      const int grp = b / (per + kNumXcd), j = b - grp * (per + kNumXcd);
      if (j >= per) {
        cg_defer_stream(dx, (int64_t)grp * kNumXcd + (j - per),
                        (int64_t)(gridDim.x / (per + kNumXcd)) * kNumXcd);
        return;
      }
      b = grp * per + j;
      if (b >= ts.n_tiles) return;
    }
  }
  // rev: tiles from the end (a solver alternates the order between iterations
  // so that each pass starts on the data the last-level cache kept from the
  // previous one; every tile writes its own slots, so the result is the same)
  int tile = (int)xcd_block(b, ts.n_tiles);
  if (rev) tile = ts.n_tiles - 1 - tile;
  tile_assemble<KIND, Dens, T, MODE, GEOM>(tile, coords, quad, dens, u, v, dof_map, ts, partial,
                                           lds_raw, lin, geo);
}

// Sum of node i's tile slots (ascending tile): the tiles wrote them to the
// node-major range [slot_ptr[i], slot_ptr[i + 1]) of the partial buffer.
template <int NC, class T>
__device__ __forceinline__ void slot_sum(int64_t i, const int32_t *__restrict__ slot_ptr,
                                         const T *__restrict__ partial, T (&acc)[NC]) {
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = T(0);
  const int32_t b = slot_ptr[i], en = slot_ptr[i + 1];
  for (int32_t k = b; k < en; ++k) {
    const T *src = partial + (int64_t)k * NC;
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] += src[c];
  }
}

// out[map(node*NC + c)] = sum of the node's tile slots (ascending tile) - f_ext.
template <int NC, class T>
__global__ __launch_bounds__(kBlock) void k_slot_gather(int64_t n_nodes, const int32_t *__restrict__ slot_ptr,
                                                        const T *__restrict__ partial,
                                                        const double *__restrict__ f_ext,
                                                        const int32_t *__restrict__ dof_map,
                                                        T *__restrict__ out,
                                                        const int *__restrict__ skip) {
  if (skip && *skip) return;
  const int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  T acc[NC];
  slot_sum<NC, T>(i, slot_ptr, partial, acc);
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

// As k_slot_gather without f_ext, grid-stride over nodes, plus the block's
// share of w . out (w and out reduced through dof_map): the PCG operator step
// y = H p with p.Hp in one pass.
template <int NC, class T>
__global__ __launch_bounds__(kBlock) void k_slot_gather_dot(int64_t n_nodes, const int32_t *__restrict__ slot_ptr,
                                                            const T *__restrict__ partial,
                                                            const int32_t *__restrict__ dof_map,
                                                            const T *__restrict__ w,
                                                            T *__restrict__ out,
                                                            double *__restrict__ part_dot,
                                                            const int *__restrict__ skip) {
  if (skip && *skip) return;
  double dot = 0.0;
  // two nodes per pass (i and i + stride), their slot lists walked side by side
  // so that the loads of both are in flight together; each node still sums its
  // slots in ascending tile order
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x; i < n_nodes;
       i += 2 * stride) {
    const int64_t i1 = i + stride;
    const bool has1 = i1 < n_nodes;
    const int32_t b0 = slot_ptr[i], e0 = slot_ptr[i + 1];
    const int32_t b1 = has1 ? slot_ptr[i1] : 0, e1 = has1 ? slot_ptr[i1 + 1] : 0;
    T acc0[NC], acc1[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc0[c] = acc1[c] = T(0);
    const int32_t n0 = e0 - b0, n1 = e1 - b1, nmax = n0 > n1 ? n0 : n1;
    for (int32_t k = 0; k < nmax; ++k) {
      if (k < n0) {
        const T *src = partial + (int64_t)(b0 + k) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc0[c] += src[c];
      }
      if (k < n1) {
        const T *src = partial + (int64_t)(b1 + k) * NC;
#pragma unroll
        for (int c = 0; c < NC; ++c) acc1[c] += src[c];
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      const int64_t dof = i * NC + c;
      const int64_t r = dof_map ? (int64_t)dof_map[dof] : dof;
      if (r >= 0) {
        out[r] = acc0[c];
        dot += (double)acc0[c] * (double)w[r];
      }
    }
    if (has1) {
#pragma unroll
      for (int c = 0; c < NC; ++c) {
        const int64_t dof = i1 * NC + c;
        const int64_t r = dof_map ? (int64_t)dof_map[dof] : dof;
        if (r >= 0) {
          out[r] = acc1[c];
          dot += (double)acc1[c] * (double)w[r];
        }
      }
    }
  }
  dot = block_sum(dot);
  if (threadIdx.x == 0) part_dot[blockIdx.x] = dot;
}

}  // namespace tatva
