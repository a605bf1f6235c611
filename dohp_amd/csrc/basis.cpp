// Host-side 1-D rules and nodal bases (setup only).
//
// Gauss-Legendre / Gauss-Lobatto-Legendre nodes by Newton iteration on Legendre
// polynomials, Lagrange bases by the barycentric formula.  The reference computes
// the same tables with polylib (src/jacobi/impls/tensor/polylib.c:182-800) inside
// TensorBasisCreate (src/jacobi/impls/tensor/tensor.c:22-109) and TensorGetRule
// (src/jacobi/quadrature/impls/tensor/tensorquad.c:51-118).
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "common.h"

namespace dohp {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

// Legendre P_n(x) and P_n'(x) by the three-term recurrence.
static void legendre(int n, double x, double *P, double *dP)
{
  double p0 = 1.0, p1 = x;
  if (n == 0) { *P = 1.0; *dP = 0.0; return; }
  for (int k = 2; k <= n; k++) {
    const double p2 = ((2.0 * k - 1.0) * x * p1 - (k - 1.0) * p0) / k;
    p0 = p1;
    p1 = p2;
  }
  *P = p1;
  // P_n'(x) = n (x P_n - P_{n-1}) / (x^2 - 1), valid off the endpoints
  *dP = n * (x * p1 - p0) / (x * x - 1.0);
}

int gauss_legendre(int n, double *z, double *w)
{
  if (n < 1 || n > 64) return 1;
  if (n == 1) { z[0] = 0; w[0] = 2; return 0; }
  for (int i = 0; i < n; i++) {
    double x = -std::cos(M_PI * (i + 0.75) / (n + 0.5));
    for (int it = 0; it < 100; it++) {
      double P, dP;
      legendre(n, x, &P, &dP);
      const double dx = P / dP;
      x -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    double P, dP;
    legendre(n, x, &P, &dP);
    z[i] = x;
    w[i] = 2.0 / ((1.0 - x * x) * dP * dP);
  }
  return 0;
}

int gauss_lobatto(int n, double *z, double *w)
{
  if (n < 2 || n > 64) return 1;
  const int N = n - 1;  // interior nodes are roots of P_N'
  z[0] = -1.0;
  z[N] = 1.0;
  for (int i = 1; i < N; i++) {
    // initial guess: Chebyshev-Gauss-Lobatto point
    double x = -std::cos(M_PI * i / N);
    for (int it = 0; it < 100; it++) {
      // Newton on q(x) = P_N'(x); q'(x) = (2x P_N' - N(N+1) P_N) / (1 - x^2)
      double P, dP;
      legendre(N, x, &P, &dP);
      const double d2P = (2.0 * x * dP - N * (N + 1.0) * P) / (1.0 - x * x);
      const double dx = dP / d2P;
      x -= dx;
      if (std::fabs(dx) < 1e-16) break;
    }
    z[i] = x;
  }
  for (int i = 0; i <= N; i++) {
    double P, dP;
    if (i == 0 || i == N) P = (i == 0 && (N % 2)) ? -1.0 : 1.0;
    else legendre(N, z[i], &P, &dP);
    w[i] = 2.0 / (N * (N + 1.0) * P * P);
  }
  return 0;
}

void barycentric_weights(int n, const double *x, double *bw)
{
  for (int j = 0; j < n; j++) {
    double prod = 1.0;
    for (int k = 0; k < n; k++)
      if (k != j) prod *= (x[j] - x[k]);
    bw[j] = 1.0 / prod;
  }
}

void lagrange_matrix(int n, const double *x, int m, const double *xq, double *interp)
{
  double bw[kMaxN * 4];
  barycentric_weights(n, x, bw);
  for (int q = 0; q < m; q++) {
    int hit = -1;
    for (int j = 0; j < n; j++)
      if (std::fabs(xq[q] - x[j]) < 1e-14) hit = j;
    if (hit >= 0) {
      for (int j = 0; j < n; j++) interp[q * n + j] = j == hit ? 1.0 : 0.0;
      continue;
    }
    double den = 0.0;
    for (int j = 0; j < n; j++) den += bw[j] / (xq[q] - x[j]);
    for (int j = 0; j < n; j++) interp[q * n + j] = (bw[j] / (xq[q] - x[j])) / den;
  }
}

void collocation_derivative(int n, const double *x, double *D)
{
  double bw[kMaxN * 4];
  barycentric_weights(n, x, bw);
  for (int i = 0; i < n; i++) {
    double diag = 0.0;
    for (int j = 0; j < n; j++) {
      if (i == j) continue;
      D[i * n + j] = (bw[j] / bw[i]) / (x[i] - x[j]);
      diag -= D[i * n + j];
    }
    D[i * n + i] = diag;
  }
}

}  // namespace dohp

using namespace dohp;

extern "C" void dohpGetLastError(char *buf, size_t len)
{
  if (!buf || !len) return;
  std::strncpy(buf, g_last_error.c_str(), len - 1);
  buf[len - 1] = 0;
}

extern "C" dErr dohpTensorRule(int method, int family, int order, int *size, double *coord, double *weight)
{
  DOHP_CHECK(order >= 0 && order < 100, dohpERR_ARG, "rule order out of bounds");
  DOHP_CHECK(size && coord && weight, dohpERR_ARG, "null argument");
  int n;
  switch (method) {
  case dQUADRATURE_METHOD_SPARSE: {
    // Two points per sub-interval between the GLL nodes of the (n/2+1)-point rule.
    n = order > 0 ? order : 1;
    *size = n;
    if (n == 1) { coord[0] = 0; weight[0] = 2; return 0; }
    const int nel = n / 2;
    DOHP_CHECK(nel + 1 <= 64, dohpERR_ARG, "sparse rule too large");
    double iv[64], un[64];
    gauss_lobatto(nel + 1, iv, un);
    const double g = family == dGAUSS_LOBATTO ? 1.0 : 1.0 / std::sqrt(3.0);
    for (int i = 0; i < nel; i++) {
      const double h = iv[i + 1] - iv[i], mid = 0.5 * (iv[i] + iv[i + 1]);
      coord[2 * i] = mid - g * h / 2;
      coord[2 * i + 1] = mid + g * h / 2;
      weight[2 * i] = weight[2 * i + 1] = h / 2;
    }
    return 0;
  }
  case dQUADRATURE_METHOD_FAST: n = family == dGAUSS_LOBATTO ? 2 + order / 2 : 1 + order / 2; break;
  case dQUADRATURE_METHOD_SELF:
    DOHP_CHECK(family == dGAUSS_LOBATTO, dohpERR_ARG, "SELF rule needs the Lobatto family");
    n = 1 + order / 2;
    break;
  default: DOHP_CHECK(false, dohpERR_ARG, "unknown quadrature method");
  }
  *size = n;
  if (n == 1) { coord[0] = 0; weight[0] = 2; return 0; }
  const int rc = family == dGAUSS_LOBATTO ? gauss_lobatto(n, coord, weight) : gauss_legendre(n, coord, weight);
  DOHP_CHECK(rc == 0, dohpERR_ARG, "rule size unsupported");
  return 0;
}

extern "C" dErr dohpTensorBasis(int P, int Q, const double *rcoord, double *interp, double *deriv, double *node,
                                double *weight)
{
  DOHP_CHECK(P >= 2 && P <= kMaxN, dohpERR_ARG, "basis size P must be in [2,16]");
  DOHP_CHECK(Q >= 1 && Q <= 4 * kMaxN, dohpERR_ARG, "rule size Q must be in [1,64]");
  double z[kMaxN], w[kMaxN], Dc[kMaxN * kMaxN];
  gauss_lobatto(P, z, w);
  double B[4 * kMaxN * kMaxN];  // Q x P, on the stack (no allocation: used by the allocation-free apply)
  lagrange_matrix(P, z, Q, rcoord, B);
  collocation_derivative(P, z, Dc);
  for (int q = 0; q < Q; q++)
    for (int j = 0; j < P; j++) {
      double s = 0;
      for (int k = 0; k < P; k++) s += B[q * P + k] * Dc[k * P + j];
      if (deriv) deriv[q * P + j] = s;
      if (interp) interp[q * P + j] = B[q * P + j];
    }
  if (node) std::memcpy(node, z, P * sizeof(double));
  if (weight) std::memcpy(weight, w, P * sizeof(double));
  return 0;
}
