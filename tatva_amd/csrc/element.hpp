// Reference-element math on the device: shape functions N(xi), their natural
// derivatives dN/dxi (d x nne), the Jacobian J = dNdr @ X (so J = (dx/dxi)^T),
// its signed determinant, inv(J) and dNdX = inv(J) @ dNdr.
// Follows tatva/element/base.py:90-115 (get_jacobian/gradient/interpolate),
// :245-265 (Tri3), :448-472 (Tetrahedron4), :475-568 (Hexahedron8).
#pragma once

#include <hip/hip_runtime.h>

#include "common.hpp"

namespace tatva {

constexpr int kMaxQuad = 27;

// Quadrature rule passed by value as a kernel argument (uniform -> SGPRs).
struct QuadRule {
  int nq;
  int pad;
  double xi[kMaxQuad][3];
  double w[kMaxQuad];
  // the points in FP32 and their products (x, y, z, yz, xz, xy, xyz), rounded
  // as a device cast of xi would round them: uniform kernel arguments that an
  // FP32 kernel uses as scalar operands (set by set_quad_f32)
  float xf[kMaxQuad][8];
};

inline void set_quad_f32(QuadRule &q) {
  for (int i = 0; i < q.nq; ++i) {
    const float x = (float)q.xi[i][0], y = (float)q.xi[i][1], z = (float)q.xi[i][2];
    const float v[8] = {x, y, z, y * z, x * z, x * y, x * y * z, 0.0f};
    for (int k = 0; k < 8; ++k) q.xf[i][k] = v[k];
  }
}

// Uniform 2-D grid of point-location cells: n[j] cells of width h[j] from
// lo[j] along axis j, cell (cx, cy) at index cy * n[0] + cx.
struct PointGrid {
  int n[2];
  double lo[2];
  double h[2];
};

// Cell coordinate of v along an axis.  Host and device evaluate the same
// correctly rounded operations, and the result does not decrease as v grows,
// so a point inside an element's bounding box [a, b] always falls in a cell
// between the cells of a and b.  Outside the grid (and NaN) it clamps.
__host__ __device__ inline int grid_cell(double v, double lo, double h, int n) {
  double t = (v - lo) / h;
  t = t >= 0.0 ? t : 0.0;
  t = t < double(n - 1) ? t : double(n - 1);
  return (int)floor(t);
}

template <int KIND>
struct Elem;

template <>
struct Elem<TATVA_TRI3> {
  static constexpr int D = 2, NNE = 3;
  template <class T>
  __device__ static void N(const double *xi, T n[NNE]) {
    n[0] = T(1.0 - xi[0] - xi[1]);
    n[1] = T(xi[0]);
    n[2] = T(xi[1]);
  }
  template <class T>
  __device__ static void dN(const double *, T dn[D][NNE]) {
    dn[0][0] = T(-1); dn[0][1] = T(1); dn[0][2] = T(0);
    dn[1][0] = T(-1); dn[1][1] = T(0); dn[1][2] = T(1);
  }
};

template <>
struct Elem<TATVA_TET4> {
  static constexpr int D = 3, NNE = 4;
  template <class T>
  __device__ static void N(const double *xi, T n[NNE]) {
    n[0] = T(1.0 - xi[0] - xi[1] - xi[2]);
    n[1] = T(xi[0]);
    n[2] = T(xi[1]);
    n[3] = T(xi[2]);
  }
  template <class T>
  __device__ static void dN(const double *, T dn[D][NNE]) {
#pragma unroll
    for (int i = 0; i < D; ++i) {
      dn[i][0] = T(-1);
#pragma unroll
      for (int a = 1; a < NNE; ++a) dn[i][a] = T(i == a - 1 ? 1 : 0);
    }
  }
};

template <>
struct Elem<TATVA_HEX8> {
  static constexpr int D = 3, NNE = 8;
  // Reference node signs (element/base.py:478-491).
  __device__ static constexpr int sx(int a) { return (a == 1 || a == 2 || a == 5 || a == 6) ? 1 : -1; }
  __device__ static constexpr int sy(int a) { return (a == 2 || a == 3 || a == 6 || a == 7) ? 1 : -1; }
  __device__ static constexpr int sz(int a) { return a >= 4 ? 1 : -1; }
  template <class T>
  __device__ static void N(const double *xi, T n[NNE]) {
#pragma unroll
    for (int a = 0; a < NNE; ++a)
      n[a] = T(0.125 * (1.0 + sx(a) * xi[0]) * (1.0 + sy(a) * xi[1]) * (1.0 + sz(a) * xi[2]));
  }
  template <class T>
  __device__ static void dN(const double *xi, T dn[D][NNE]) {
#pragma unroll
    for (int a = 0; a < NNE; ++a) {
      const double fx = 1.0 + sx(a) * xi[0];
      const double fy = 1.0 + sy(a) * xi[1];
      const double fz = 1.0 + sz(a) * xi[2];
      dn[0][a] = T(0.125 * sx(a) * fy * fz);
      dn[1][a] = T(0.125 * fx * sy(a) * fz);
      dn[2][a] = T(0.125 * fx * fy * sz(a));
    }
  }
};

// 1 / x: correctly rounded in FP64; in FP32 the hardware reciprocal (1 ulp),
// which the FP32 tangent of the matrix-free product uses.
template <class T>
__device__ __forceinline__ T trcp(T x);
template <>
__device__ __forceinline__ double trcp<double>(double x) { return 1.0 / x; }
template <>
__device__ __forceinline__ float trcp<float>(float x) { return __builtin_amdgcn_rcpf(x); }

// Signed det(J) and inv(J) by cofactors.
template <class T>
__device__ __forceinline__ T det_inv(const T J[2][2], T inv[2][2]) {
  const T det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
  const T r = trcp<T>(det);
  inv[0][0] = J[1][1] * r;
  inv[0][1] = -J[0][1] * r;
  inv[1][0] = -J[1][0] * r;
  inv[1][1] = J[0][0] * r;
  return det;
}

template <class T>
__device__ __forceinline__ T det_inv(const T J[3][3], T inv[3][3]) {
  const T c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
  const T c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
  const T c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
  const T det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
  const T r = trcp<T>(det);
  inv[0][0] = c00 * r;
  inv[1][0] = c01 * r;
  inv[2][0] = c02 * r;
  inv[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * r;
  inv[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * r;
  inv[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * r;
  inv[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * r;
  inv[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * r;
  inv[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * r;
  return det;
}

template <class T>
__device__ __forceinline__ T det_only(const T J[2][2]) {
  return J[0][0] * J[1][1] - J[0][1] * J[1][0];
}

template <class T>
__device__ __forceinline__ T det_only(const T J[3][3]) {
  return J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) +
         J[0][1] * (J[1][2] * J[2][0] - J[1][0] * J[2][2]) +
         J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
}

// Geometry of a simplex (Tri3 / Tet4: constant over the element), the general
// formula below with its constant dN/dxi folded in: J[i][j] = X[i+1][j] -
// X[0][j], B[i][a] = inv[i][a-1] for a >= 1, B[i][0] = -(inv[i][0] + ... +
// inv[i][D-1]) (the general form multiplies by the constants 1, -1 and 0,
// which strict FP semantics may not drop).
template <int D, int NNE>
__device__ __forceinline__ double simplex_geometry(const double X[NNE][D], double B[D][NNE]) {
  static_assert(NNE == D + 1, "simplex");
  double J[D][D], inv[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) J[i][j] = X[i + 1][j] - X[0][j];
  const double det = det_inv(J, inv);
#pragma unroll
  for (int i = 0; i < D; ++i) {
    double s = -inv[i][0];
#pragma unroll
    for (int k = 1; k < D; ++k) s -= inv[i][k];
    B[i][0] = s;
#pragma unroll
    for (int a = 1; a < NNE; ++a) B[i][a] = inv[i][a - 1];
  }
  return det;
}

// Geometry at one quadrature point: B = dNdX (D x NNE) and signed detJ, in FP64.
template <int KIND>
__device__ __forceinline__ double qp_geometry(const double X[Elem<KIND>::NNE][Elem<KIND>::D],
                                              const double *xi,
                                              double B[Elem<KIND>::D][Elem<KIND>::NNE]) {
  using E = Elem<KIND>;
  constexpr int D = E::D, NNE = E::NNE;
  if constexpr (KIND != TATVA_HEX8) return simplex_geometry<D, NNE>(X, B);
  double dn[D][NNE];
  E::dN(xi, dn);
  double J[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NNE; ++a) s += dn[i][a] * X[a][j];
      J[i][j] = s;
    }
  double inv[D][D];
  const double det = det_inv(J, inv);
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int a = 0; a < NNE; ++a) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < D; ++k) s += inv[i][k] * dn[k][a];
      B[i][a] = s;
    }
  return det;
}

template <int KIND>
__device__ __forceinline__ double qp_detJ(const double X[Elem<KIND>::NNE][Elem<KIND>::D],
                                          const double *xi) {
  using E = Elem<KIND>;
  constexpr int D = E::D, NNE = E::NNE;
  double dn[D][NNE];
  E::dN(xi, dn);
  double J[D][D];
#pragma unroll
  for (int i = 0; i < D; ++i)
#pragma unroll
    for (int j = 0; j < D; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < NNE; ++a) s += dn[i][a] * X[a][j];
      J[i][j] = s;
    }
  return det_only(J);
}

// Gather the element's node ids and coordinates.
template <int KIND>
__device__ __forceinline__ void load_element(int64_t e, const int32_t *__restrict__ conn,
                                             const double *__restrict__ coords,
                                             int32_t nodes[Elem<KIND>::NNE],
                                             double X[Elem<KIND>::NNE][Elem<KIND>::D]) {
  using E = Elem<KIND>;
  constexpr int D = E::D, NNE = E::NNE;
  const int32_t *c = conn + e * NNE;
  if constexpr (NNE == 8) {
    const int4 a = *reinterpret_cast<const int4 *>(c);
    const int4 b = *reinterpret_cast<const int4 *>(c + 4);
    nodes[0] = a.x; nodes[1] = a.y; nodes[2] = a.z; nodes[3] = a.w;
    nodes[4] = b.x; nodes[5] = b.y; nodes[6] = b.z; nodes[7] = b.w;
  } else if constexpr (NNE == 4) {
    const int4 a = *reinterpret_cast<const int4 *>(c);
    nodes[0] = a.x; nodes[1] = a.y; nodes[2] = a.z; nodes[3] = a.w;
  } else {
#pragma unroll
    for (int a = 0; a < NNE; ++a) nodes[a] = c[a];
  }
#pragma unroll
  for (int a = 0; a < NNE; ++a)
#pragma unroll
    for (int j = 0; j < D; ++j) X[a][j] = coords[(int64_t)nodes[a] * D + j];
}

}  // namespace tatva
