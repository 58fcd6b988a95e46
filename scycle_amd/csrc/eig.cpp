// Dense symmetric eigen-decomposition for the fast-diagonalization set-up:
// Householder reduction to tridiagonal form followed by the implicit-shift QL
// iteration (the EISPACK tred2/tql2 method as published by Martin, Reinsch and
// Wilkinson).  n is at most a few thousand and this runs once per operator.
#include <cmath>
#include <vector>

#include "fdiag.hpp"

namespace sc {

namespace {
double hypot2(double a, double b) { return std::hypot(a, b); }
}  // namespace

// a: n*n row-major symmetric; on return eigenvalues in w (ascending) and the
// eigenvectors as ROWS of vt (vt[i*n + k] = component k of vector i).
int sym_eig(int n, std::vector<double> a, std::vector<double>& w, std::vector<double>& vt) {
  std::vector<double> d(n), e(n);
  auto A = [&](int i, int j) -> double& { return a[(size_t)i * n + j]; };
  // --- Householder tridiagonalisation, accumulating the transformation in a
  for (int i = n - 1; i > 0; --i) {
    const int l = i - 1;
    double h = 0.0;
    if (l > 0) {
      double scale = 0.0;
      for (int k = 0; k <= l; ++k) scale += std::fabs(A(i, k));
      if (scale == 0.0) {
        e[i] = A(i, l);
      } else {
        for (int k = 0; k <= l; ++k) {
          A(i, k) /= scale;
          h += A(i, k) * A(i, k);
        }
        double f = A(i, l);
        double g = f >= 0.0 ? -std::sqrt(h) : std::sqrt(h);
        e[i] = scale * g;
        h -= f * g;
        A(i, l) = f - g;
        f = 0.0;
        for (int j = 0; j <= l; ++j) {
          A(j, i) = A(i, j) / h;
          g = 0.0;
          for (int k = 0; k <= j; ++k) g += A(j, k) * A(i, k);
          for (int k = j + 1; k <= l; ++k) g += A(k, j) * A(i, k);
          e[j] = g / h;
          f += e[j] * A(i, j);
        }
        const double hh = f / (h + h);
        for (int j = 0; j <= l; ++j) {
          f = A(i, j);
          e[j] = g = e[j] - hh * f;
          for (int k = 0; k <= j; ++k) A(j, k) -= (f * e[k] + g * A(i, k));
        }
      }
    } else {
      e[i] = A(i, l);
    }
    d[i] = h;
  }
  d[0] = 0.0;
  e[0] = 0.0;
  for (int i = 0; i < n; ++i) {
    const int l = i - 1;
    if (d[i] != 0.0) {
      for (int j = 0; j <= l; ++j) {
        double g = 0.0;
        for (int k = 0; k <= l; ++k) g += A(i, k) * A(k, j);
        for (int k = 0; k <= l; ++k) A(k, j) -= g * A(k, i);
      }
    }
    d[i] = A(i, i);
    A(i, i) = 1.0;
    for (int j = 0; j <= l; ++j) A(j, i) = A(i, j) = 0.0;
  }
  // eigenvector matrix Q has columns = vectors; keep it transposed so the QL
  // rotations touch two contiguous rows
  vt.assign((size_t)n * n, 0.0);
  for (int i = 0; i < n; ++i)
    for (int k = 0; k < n; ++k) vt[(size_t)i * n + k] = A(k, i);
  // --- implicit QL on the tridiagonal (d, e)
  for (int i = 1; i < n; ++i) e[i - 1] = e[i];
  e[n - 1] = 0.0;
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; ++l) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (iter++ == 200) return 1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot2(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0.0 ? std::fabs(r) : -std::fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; --i) {
          double f = s * e[i];
          const double b = c * e[i];
          e[i + 1] = (r = hypot2(f, g));
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          d[i + 1] = g + (p = s * r);
          g = c * r - b;
          double* z0 = &vt[(size_t)i * n];
          double* z1 = &vt[(size_t)(i + 1) * n];
          for (int k = 0; k < n; ++k) {
            f = z1[k];
            z1[k] = s * z0[k] + c * f;
            z0[k] = c * z0[k] - s * f;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  // sort ascending
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::vector<double> ws(n), vs((size_t)n * n);
  for (int i = 0; i < n; ++i) {
    int best = i;
    for (int j = i + 1; j < n; ++j)
      if (d[idx[j]] < d[idx[best]]) best = j;
    std::swap(idx[i], idx[best]);
  }
  for (int i = 0; i < n; ++i) {
    ws[i] = d[idx[i]];
    for (int k = 0; k < n; ++k) vs[(size_t)i * n + k] = vt[(size_t)idx[i] * n + k];
  }
  w.swap(ws);
  vt.swap(vs);
  return 0;
}

}  // namespace sc
