/* Oracle: plain-C FP64 restatement of the element residual, Hessian-vector
 * product and Hessian diagonal of oracle/fem.py (Problem.residual / hvp /
 * hessian_diag) for Tet4 and Hex8 with linear elasticity and compressible
 * Neo-Hookean, parallel over elements with OpenMP.  TEST INFRASTRUCTURE ONLY:
 * the timed CPU baseline of the benchmark and its own parity tests use it; the
 * product never links it.
 *
 * Element algebra as the reference (nodal form, no modal rewrite):
 *   J = dNdr @ X, detJ signed, dNdX = inv(J) @ dNdr (tatva/element/base.py:90-115)
 *   grad u [c][j] = sum_n dNdX[j][n] u[n][c]           (element/base.py:99-115)
 *   shape functions: Tetrahedron4 (element/base.py:448-472),
 *                    Hexahedron8 with the 1/8 factor (element/base.py:475-568)
 *   r_e[a][c] = sum_q detJ w_q (P(G) dNdX)[c][a]       (jax.grad of
 *                                                        tatva/operator.py:307-319)
 * Densities (user code of the reference tests):
 *   LE: sigma = mu (G + G^T) + lambda tr(G) I          (tests/test_sparse.py:20-38)
 *   NH: F = I + G, P = mu (F - F^-T) + lambda ln J F^-T (tests/test_sparse_tracer.py:103-115)
 *       dP = mu dG + (mu - lambda ln J) F^-T dG^T F^-T + lambda tr(F^-1 dG) F^-T
 * Scatter into nodes with atomic adds (the summation order varies between runs;
 * the results agree with oracle/fem.py to rounding). */
#include <math.h>
#include <stdint.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

enum { K_TET4 = 1, K_HEX8 = 2 };
enum { D_LE = 2, D_NH = 3 };
enum { M_RESIDUAL = 0, M_HVP = 1, M_DIAG = 2 };

static const double HEX_R[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                                   {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

/* dN[d][n] at xi */
static void shape_grad(int kind, const double *xi, double dN[3][8]) {
  if (kind == K_TET4) {
    static const double T[3][4] = {{-1, 1, 0, 0}, {-1, 0, 1, 0}, {-1, 0, 0, 1}};
    for (int d = 0; d < 3; ++d)
      for (int n = 0; n < 4; ++n) dN[d][n] = T[d][n];
    return;
  }
  for (int n = 0; n < 8; ++n) {
    const double fx = 1 + HEX_R[n][0] * xi[0], fy = 1 + HEX_R[n][1] * xi[1],
                 fz = 1 + HEX_R[n][2] * xi[2];
    dN[0][n] = 0.125 * HEX_R[n][0] * fy * fz;
    dN[1][n] = 0.125 * fx * HEX_R[n][1] * fz;
    dN[2][n] = 0.125 * fx * fy * HEX_R[n][2];
  }
}

static double inv3(const double A[3][3], double Ai[3][3]) {
  const double c00 = A[1][1] * A[2][2] - A[1][2] * A[2][1];
  const double c01 = A[1][2] * A[2][0] - A[1][0] * A[2][2];
  const double c02 = A[1][0] * A[2][1] - A[1][1] * A[2][0];
  const double det = A[0][0] * c00 + A[0][1] * c01 + A[0][2] * c02;
  const double id = 1.0 / det;
  Ai[0][0] = c00 * id;
  Ai[1][0] = c01 * id;
  Ai[2][0] = c02 * id;
  Ai[0][1] = (A[0][2] * A[2][1] - A[0][1] * A[2][2]) * id;
  Ai[1][1] = (A[0][0] * A[2][2] - A[0][2] * A[2][0]) * id;
  Ai[2][1] = (A[0][1] * A[2][0] - A[0][0] * A[2][1]) * id;
  Ai[0][2] = (A[0][1] * A[1][2] - A[0][2] * A[1][1]) * id;
  Ai[1][2] = (A[0][2] * A[1][0] - A[0][0] * A[1][2]) * id;
  Ai[2][2] = (A[0][0] * A[1][1] - A[0][1] * A[1][0]) * id;
  return det;
}

typedef struct {
  int density;
  double mu, lam;
  double Fi[3][3]; /* NH: F^-1 */
  double lnJ;
} Lin;

static void linearize(Lin *L, const double G[3][3]) {
  if (L->density != D_NH) return;
  double F[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) F[i][j] = G[i][j] + (i == j ? 1.0 : 0.0);
  L->lnJ = log(inv3(F, L->Fi));
}

static void stress(const Lin *L, const double G[3][3], double P[3][3]) {
  if (L->density == D_LE) {
    const double tr = G[0][0] + G[1][1] + G[2][2];
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        P[i][j] = L->mu * (G[i][j] + G[j][i]) + (i == j ? L->lam * tr : 0.0);
    return;
  }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      P[i][j] = L->mu * (G[i][j] + (i == j ? 1.0 : 0.0) - L->Fi[j][i]) +
                L->lam * L->lnJ * L->Fi[j][i];
}

static void dstress(const Lin *L, const double dG[3][3], double S[3][3]) {
  if (L->density == D_LE) {
    stress(L, dG, S);
    return;
  }
  double W[3][3]; /* F^-1 dG */
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) W[i][k] = L->Fi[i][0] * dG[0][k] + L->Fi[i][1] * dG[1][k] +
                                          L->Fi[i][2] * dG[2][k];
  const double tr = W[0][0] + W[1][1] + W[2][2];
  const double c1 = L->mu - L->lam * L->lnJ;
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      /* (F^-T dG^T F^-T)_ij = (F^-1 dG F^-1)_ji = sum_k W[j][k] Fi[k][i] */
      const double z = W[j][0] * L->Fi[0][i] + W[j][1] * L->Fi[1][i] + W[j][2] * L->Fi[2][i];
      S[i][j] = L->mu * dG[i][j] + c1 * z + L->lam * tr * L->Fi[j][i];
    }
}

/* out (n_nodes*3) = residual (mode 0), H(u) v (mode 1) or diag H(u) (mode 2).
 * Returns 0, or -1 for an unsupported kind or density. */
int oracle_assemble(int kind, int density, double mu, double lam, const double *X, int64_t n_nodes,
                    const int32_t *conn, int64_t n_elem, int nq, const double *qp,
                    const double *qw, const double *u, const double *v, int mode, double *out,
                    int threads) {
  if ((kind != K_TET4 && kind != K_HEX8) || (density != D_LE && density != D_NH)) return -1;
  const int nne = kind == K_TET4 ? 4 : 8;
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < n_nodes * 3; ++i) out[i] = 0.0;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < n_elem; ++e) {
    double Xe[8][3], ue[8][3], ve[8][3], re[8][3];
    for (int a = 0; a < nne; ++a) {
      const int64_t g = conn[e * nne + a];
      for (int k = 0; k < 3; ++k) {
        Xe[a][k] = X[g * 3 + k];
        ue[a][k] = u[g * 3 + k];
        ve[a][k] = v ? v[g * 3 + k] : 0.0;
        re[a][k] = 0.0;
      }
    }
    for (int q = 0; q < nq; ++q) {
      double dN[3][8], J[3][3], Ji[3][3], B[3][8];
      shape_grad(kind, qp + 3 * q, dN);
      for (int d = 0; d < 3; ++d)
        for (int k = 0; k < 3; ++k) {
          double s = 0.0;
          for (int n = 0; n < nne; ++n) s += dN[d][n] * Xe[n][k];
          J[d][k] = s;
        }
      const double w = inv3(J, Ji) * qw[q];
      for (int j = 0; j < 3; ++j)
        for (int n = 0; n < nne; ++n) B[j][n] = Ji[j][0] * dN[0][n] + Ji[j][1] * dN[1][n] +
                                                Ji[j][2] * dN[2][n];
      double G[3][3];
      for (int c = 0; c < 3; ++c)
        for (int j = 0; j < 3; ++j) {
          double s = 0.0;
          for (int n = 0; n < nne; ++n) s += B[j][n] * ue[n][c];
          G[c][j] = s;
        }
      Lin L = {density, mu, lam, {{0}}, 0.0};
      linearize(&L, G);
      if (mode == M_DIAG) {
        for (int a = 0; a < nne; ++a)
          for (int c = 0; c < 3; ++c) {
            double dG[3][3] = {{0}}, S[3][3];
            for (int j = 0; j < 3; ++j) dG[c][j] = B[j][a];
            dstress(&L, dG, S);
            re[a][c] += w * (S[c][0] * B[0][a] + S[c][1] * B[1][a] + S[c][2] * B[2][a]);
          }
        continue;
      }
      double S[3][3];
      if (mode == M_RESIDUAL) {
        stress(&L, G, S);
      } else {
        double dG[3][3];
        for (int c = 0; c < 3; ++c)
          for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int n = 0; n < nne; ++n) s += B[j][n] * ve[n][c];
            dG[c][j] = s;
          }
        dstress(&L, dG, S);
      }
      for (int a = 0; a < nne; ++a)
        for (int c = 0; c < 3; ++c)
          re[a][c] += w * (S[c][0] * B[0][a] + S[c][1] * B[1][a] + S[c][2] * B[2][a]);
    }
    for (int a = 0; a < nne; ++a) {
      const int64_t g = conn[e * nne + a];
      for (int c = 0; c < 3; ++c) {
#pragma omp atomic
        out[g * 3 + c] += re[a][c];
      }
    }
  }
  return 0;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}
