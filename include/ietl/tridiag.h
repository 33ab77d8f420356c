// ietl/tridiag.h -- symmetric tridiagonal eigensolver used by the T-matrix
// analysis.  It takes the place of the LAPACK dstev call the reference makes
// through src/ietl/ietl2lapack.h:44-99 (jobz 'N' and 'V'), so the engine
// needs no LAPACK on the host.
//
// Method: implicit QL iteration with Wilkinson shifts (Bowdler, Martin,
// Reinsch and Wilkinson, Numer. Math. 11 (1968) 293-306).  The T-matrix
// analysis only reads the last row of the eigenvector matrix (the Ritz error
// bounds |beta_n * z(n-1,i)|, tmatrix.h:164-183), so the rotations are applied
// to that row alone: O(n^2) work instead of O(n^3).
#ifndef ALPS_ED_IETL_TRIDIAG_H
#define ALPS_ED_IETL_TRIDIAG_H

#include <cmath>
#include <cstddef>
#include <limits>
#include <utility>
#include <vector>

namespace ietl {
namespace tridiag {

// sqrt(a^2 + b^2) without the cost of std::hypot on the common range; the
// scaled form guards against overflow and underflow outside it.
inline double pythag(double a, double b) {
  const double fa = std::fabs(a), fb = std::fabs(b);
  const double m = fa > fb ? fa : fb;
  if (m > 1e-150 && m < 1e150) return std::sqrt(a * a + b * b);
  if (m == 0.0) return 0.0;
  const double x = fa / m, y = fb / m;
  return m * std::sqrt(x * x + y * y);
}

// d[0..n): diagonal (overwritten by ascending eigenvalues).
// e[0..n-1): off-diagonal (not modified).
// zlast: if non-null, receives z(n-1, i) of the orthonormal eigenvectors,
//        in the order of the sorted eigenvalues.
// Returns 0, or l+1 when the l-th eigenvalue did not converge (dstev info).
inline int eig(int n, double* d, const double* e_in, double* zlast) {
  if (n <= 0) return 0;
  std::vector<double> e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  std::vector<double> z;
  if (zlast) z.assign(n, 0.0), z[n - 1] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    for (int iter = 0;; ++iter) {
      // look for a negligible off-diagonal element at or after l
      int m = l;
      while (m < n - 1 && std::fabs(e[m]) > eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) ++m;
      if (m == l) break;
      if (iter >= 60) return l + 1;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = pythag(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool split = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i];
        const double b = c * e[i];
        r = pythag(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // the matrix splits; restart the sweep
          d[i + 1] -= p;
          e[m] = 0.0;
          split = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (zlast) {
          const double t = z[i + 1];
          z[i + 1] = s * z[i] + c * t;
          z[i] = c * z[i] - s * t;
        }
      }
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  // insertion sort, ascending, carrying the eigenvector row along
  for (int i = 1; i < n; ++i) {
    double v = d[i];
    double zv = zlast ? z[i] : 0.0;
    int j = i - 1;
    while (j >= 0 && d[j] > v) {
      d[j + 1] = d[j];
      if (zlast) z[j + 1] = z[j];
      --j;
    }
    d[j + 1] = v;
    if (zlast) z[j + 1] = zv;
  }
  if (zlast)
    for (int i = 0; i < n; ++i) zlast[i] = z[i];
  return 0;
}

// Eigenvalues (ascending, into d) and all eigenvectors (column-major n x n
// into z) of the tridiagonal; same iteration, rotations applied to every row.
inline int eig_full(int n, double* d, const double* e_in, double* z) {
  if (n <= 0) return 0;
  std::vector<double> e(n, 0.0);
  for (int i = 0; i + 1 < n; ++i) e[i] = e_in[i];
  for (int i = 0; i < n * n; ++i) z[i] = 0.0;
  for (int i = 0; i < n; ++i) z[i + (std::size_t)i * n] = 1.0;
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    for (int iter = 0;; ++iter) {
      int m = l;
      while (m < n - 1 && std::fabs(e[m]) > eps * (std::fabs(d[m]) + std::fabs(d[m + 1]))) ++m;
      if (m == l) break;
      if (iter >= 60) return l + 1;
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = pythag(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool split = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i];
        const double b = c * e[i];
        r = pythag(f, g);
        e[i + 1] = r;
        if (r == 0.0) {
          d[i + 1] -= p;
          e[m] = 0.0;
          split = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        double* zi = z + (std::size_t)i * n;
        double* zi1 = z + (std::size_t)(i + 1) * n;
        for (int k = 0; k < n; ++k) {
          const double t = zi1[k];
          zi1[k] = s * zi[k] + c * t;
          zi[k] = c * zi[k] - s * t;
        }
      }
      if (split) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  for (int i = 1; i < n; ++i) {  // insertion sort carrying the columns
    int j = i;
    while (j > 0 && d[j - 1] > d[j]) {
      std::swap(d[j - 1], d[j]);
      for (int k = 0; k < n; ++k) std::swap(z[k + (std::size_t)(j - 1) * n], z[k + (std::size_t)j * n]);
      --j;
    }
  }
  return 0;
}

}  // namespace tridiag
}  // namespace ietl

#endif
