// Built-in energy densities with hand-derived first derivative P = dpsi/d(grad u)
// and its directional derivative dP = d^2psi/d(grad u)^2 [dG].
//   Poisson:  psi = |G|^2/2 - f u            (scalar field, NC = 1)
//   Linear elasticity (tests/test_sparse.py:20-38):
//             eps = sym(G), sigma = 2 mu eps + lambda tr(eps) I, psi = sigma:eps/2
//   Compressible Neo-Hookean (tests/test_sparse_tracer.py:103-115):
//             F = I + G, J = det F, psi = mu/2 (F:F - D - 2 ln J) + lambda/2 (ln J)^2
//             P  = mu (F - F^-T) + lambda ln J F^-T
//             dP = mu dG + (mu - lambda ln J) F^-T dG^T F^-T + lambda tr(F^-1 dG) F^-T
// G[c][j] = d u_c / d x_j.
// A density's parameters are uniform, or per element (ep != NULL): at(e) is the
// density of element e.
#pragma once

#include <hip/hip_runtime.h>

#include "element.hpp"

namespace tatva {

template <int D, class T>
struct Poisson {
  static constexpr int NC = 1;
  static constexpr bool kSource = true;
  T f;
  const double *ep = nullptr;  // per-element f (n_elements), or NULL
  __device__ __forceinline__ Poisson at(int64_t e) const {
    Poisson r = *this;
    if (ep && e >= 0) r.f = T(ep[e]);
    return r;
  }
  struct Lin {};
  static constexpr int kLin = 0;  // values of a cached linearisation per point
  __device__ __forceinline__ void store_lin(const Lin &, T *, int) const {}
  __device__ __forceinline__ Lin load_lin(const T *, int) const { return Lin{}; }
  __device__ __forceinline__ T psi(const T G[NC][D], T uq) const {
    T s = T(0);
#pragma unroll
    for (int j = 0; j < D; ++j) s += G[0][j] * G[0][j];
    return T(0.5) * s - f * uq;
  }
  __device__ __forceinline__ void P(const T G[NC][D], T out[NC][D]) const {
#pragma unroll
    for (int j = 0; j < D; ++j) out[0][j] = G[0][j];
  }
  __device__ __forceinline__ Lin linearize(const T (&)[NC][D]) const { return Lin{}; }
  __device__ __forceinline__ void dP(const Lin &, const T dG[NC][D], T out[NC][D]) const {
#pragma unroll
    for (int j = 0; j < D; ++j) out[0][j] = dG[0][j];
  }
  // w dP (the quadrature weight folded into the tangent)
  __device__ __forceinline__ void dPw(const Lin &, const T dG[NC][D], T w, T out[NC][D]) const {
#pragma unroll
    for (int j = 0; j < D; ++j) out[0][j] = w * dG[0][j];
  }
};

template <int D, class T>
struct LinearElastic {
  static constexpr int NC = D;
  static constexpr bool kSource = false;
  T mu, lam;
  const double *ep = nullptr;  // per-element (mu, lambda) pairs (n_elements x 2), or NULL
  __device__ __forceinline__ LinearElastic at(int64_t e) const {
    LinearElastic r = *this;
    if (ep && e >= 0) {
      r.mu = T(ep[2 * e]);
      r.lam = T(ep[2 * e + 1]);
    }
    return r;
  }
  struct Lin {};
  static constexpr int kLin = 0;
  __device__ __forceinline__ void store_lin(const Lin &, T *, int) const {}
  __device__ __forceinline__ Lin load_lin(const T *, int) const { return Lin{}; }
  __device__ __forceinline__ void stress(const T G[NC][D], T s[NC][D]) const {
    stress_with(mu, lam, G, s);
  }
  __device__ __forceinline__ static void stress_with(T m, T l, const T G[NC][D], T s[NC][D]) {
    T tr = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i) tr += G[i][i];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j)
        s[i][j] = i == j ? m * (G[i][j] + G[j][i]) + l * tr : m * (G[i][j] + G[j][i]);
  }
  __device__ __forceinline__ T psi(const T G[NC][D], T) const {
    T s[NC][D];
    stress(G, s);
    T e = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) e += s[i][j] * T(0.5) * (G[i][j] + G[j][i]);
    return T(0.5) * e;
  }
  __device__ __forceinline__ void P(const T G[NC][D], T out[NC][D]) const { stress(G, out); }
  __device__ __forceinline__ Lin linearize(const T (&)[NC][D]) const { return Lin{}; }
  __device__ __forceinline__ void dP(const Lin &, const T dG[NC][D], T out[NC][D]) const {
    stress(dG, out);
  }
  // w dP (the quadrature weight folded into the moduli)
  __device__ __forceinline__ void dPw(const Lin &, const T dG[NC][D], T w, T out[NC][D]) const {
    stress_with(w * mu, w * lam, dG, out);
  }
};

template <class T>
__device__ __forceinline__ T tlog(T x);
template <>
__device__ __forceinline__ double tlog<double>(double x) { return log(x); }
// FP32: the hardware log2 times ln 2 (the FP32 tangent's ln J; its absolute
// error of a few 1e-8 near J = 1 is below the FP32 product's rounding)
template <>
__device__ __forceinline__ float tlog<float>(float x) {
  return __builtin_amdgcn_logf(x) * 0.693147180559945309f;
}

template <int D, class T>
struct NeoHookean {
  static constexpr int NC = D;
  static constexpr bool kSource = false;
  T mu, lam;
  const double *ep = nullptr;  // per-element (mu, lambda) pairs (n_elements x 2), or NULL
  __device__ __forceinline__ NeoHookean at(int64_t e) const {
    NeoHookean r = *this;
    if (ep && e >= 0) {
      r.mu = T(ep[2 * e]);
      r.lam = T(ep[2 * e + 1]);
    }
    return r;
  }
  struct Lin {
    T Fi[D][D];  // F^-1
    T c1;        // mu - lambda ln J
    T lnJ;
  };
  __device__ __forceinline__ Lin linearize(const T (&G)[NC][D]) const {
    T F[D][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) F[i][j] = i == j ? G[i][j] + T(1) : G[i][j];
    Lin L;
    const T J = det_inv(F, L.Fi);
    L.lnJ = tlog<T>(J);
    L.c1 = mu - lam * L.lnJ;
    return L;
  }
  // dP needs F^-1 and c1 only: a cached linearisation holds those D*D + 1
  // values per point, value k of a point at dst[k * stride]
  static constexpr int kLin = D * D + 1;
  __device__ __forceinline__ void store_lin(const Lin &L, T *dst, int stride) const {
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) dst[(i * D + j) * stride] = L.Fi[i][j];
    dst[D * D * stride] = L.c1;
  }
  __device__ __forceinline__ Lin load_lin(const T *src, int stride) const {
    Lin L;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) L.Fi[i][j] = src[(i * D + j) * stride];
    L.c1 = src[D * D * stride];
    L.lnJ = T(0);  // not used by dP
    return L;
  }
  __device__ __forceinline__ T psi(const T G[NC][D], T) const {
    T F[D][D];
    T I1 = T(0);
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        F[i][j] = i == j ? G[i][j] + T(1) : G[i][j];
        I1 += F[i][j] * F[i][j];
      }
    const T lnJ = tlog<T>(det_only(F));
    return T(0.5) * mu * (I1 - T(D) - T(2) * lnJ) + T(0.5) * lam * lnJ * lnJ;
  }
  __device__ __forceinline__ void P(const T G[NC][D], T out[NC][D]) const {
    T Gc[NC][D];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) Gc[i][j] = G[i][j];
    const Lin L = linearize(Gc);
    const T b = lam * L.lnJ - mu;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j)
        out[i][j] = mu * (i == j ? G[i][j] + T(1) : G[i][j]) + b * L.Fi[j][i];
  }
  __device__ __forceinline__ void dP(const Lin &L, const T dG[NC][D], T out[NC][D]) const {
    dP_with(mu, L.c1, lam, L, dG, out);
  }
  // w dP (the quadrature weight folded into mu, mu - lambda ln J and lambda)
  __device__ __forceinline__ void dPw(const Lin &L, const T dG[NC][D], T w, T out[NC][D]) const {
    dP_with(w * mu, w * L.c1, w * lam, L, dG, out);
  }
  // m dG + c Z^T + l tr(W) F^-T with W = F^-1 dG, Z = W F^-1
  // ((F^-T dG^T F^-T)_ij = Z_ji, tr(F^-1 dG) = tr W)
  __device__ __forceinline__ static void dP_with(T m, T c, T l, const Lin &L, const T dG[NC][D],
                                                 T out[NC][D]) {
    T W[D][D];
#pragma unroll
    for (int j = 0; j < D; ++j)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        T s = T(0);
#pragma unroll
        for (int l = 0; l < D; ++l) s += L.Fi[j][l] * dG[l][k];
        W[j][k] = s;
      }
    T tr = T(0);
#pragma unroll
    for (int j = 0; j < D; ++j) tr += W[j][j];
    const T lt = l * tr;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        T z = T(0);
#pragma unroll
        for (int k = 0; k < D; ++k) z += W[j][k] * L.Fi[k][i];
        out[i][j] = m * dG[i][j] + c * z + lt * L.Fi[j][i];
      }
  }
};

}  // namespace tatva
