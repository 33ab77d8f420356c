// ed_tridiag.hpp -- host analysis of the Lanczos T matrix (no device code):
// the nlowest stopping rule on the ietl T-matrix analysis, an O(m) test of
// the lowest Ritz value that gates it, and Ritz coefficients by inverse
// iteration.  Host C++ only, so that it builds with any C++17 compiler.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#include <ietl/tmatrix.h>

namespace alps_ed {
namespace host {

struct TValues {
  std::vector<double> ev, err;
  std::vector<int> index;
  std::vector<double> all;  // every eigenvalue of T, ascending (index refers to it)
  // false for analyze_low: ev/err/index hold only the lowest entries, and
  // `all` only the eigenvalues those refer to (NaN elsewhere)
  bool complete = true;
};

inline TValues analyze_prefix(const std::vector<double>& a, const std::vector<double>& b, int m, bool fp32) {
  double amax = a[0], bmax = b[0];
  for (int i = 1; i < m; ++i) {
    amax = std::max(amax, a[i]);
    bmax = std::max(bmax, b[i]);
  }
  ietl::tmatrix_analysis an;
  an.run(a.data(), b.data(), m, amax, bmax, true);
  TValues t;
  t.all = std::move(an.eigval_all);
  const double dup = 64.0 * std::numeric_limits<float>::epsilon();
  for (size_t i = 0; i < an.eigval_noghost.size(); ++i) {
    const double lam = an.eigval_noghost[i];
    if (fp32 && !t.ev.empty() && lam - t.ev.back() <= dup * std::max(1.0, std::fabs(t.ev.back()))) {
      t.err.back() = std::min(t.err.back(), an.err_noghost[i]);
      continue;
    }
    t.ev.push_back(lam);
    t.err.push_back(an.err_noghost[i]);
    t.index.push_back(an.index_noghost[i]);
  }
  return t;
}

// Stopping tolerance (rtol = atol) of the nlowest rule: 100 eps, as
// lanczos_iteration_nlowest (iteration.h:110-113), for FP64 storage.  With
// FP32 storage the T-matrix error bounds stall near the FP32 noise level
// long before 100 eps; the solve stops at 1e-8 and reports FP64 Rayleigh
// quotients of the Ritz vectors, whose error is about the square of the
// residual (1e-14 and below on the tested chains and lattices).
inline double stop_tolerance(bool fp32) {
  return fp32 ? 1e-8 : 100. * std::numeric_limits<double>::epsilon();
}

// lanczos_iteration_nlowest::converged (iteration.h:116-130).
inline bool nlowest_converged(const TValues& t, int nev, bool fp32 = false) {
  const double tol = stop_tolerance(fp32);
  if ((int)t.ev.size() < nev) return false;
  for (int i = 0; i < nev; ++i)
    if (t.err[i] > std::max(tol, tol * std::fabs(t.ev[i]))) return false;
  return true;
}

// Cheap necessary-condition test for the nlowest stopping rule: the error
// bound |beta_{m-1} z_{m-1}| of the lowest eigenvalue of T[0..m).  O(m) per
// pass.
//
// The eigenvalue: Laguerre's iteration on p(x) = det(T - x I), started left
// of every eigenvalue (the Gershgorin bound, or hint_lo when the Sturm count
// confirms it; by interlacing the lowest value of a longer prefix only moves
// down, so the previous test's value bounds it from above).  For a
// polynomial with real roots, Laguerre's steps from the left of the smallest
// root increase monotonically towards it without passing it and converge
// cubically (a far start costs one or two steps, where Newton's would creep
// by 1/m of the distance).  G = p'/p = sum_i q_i'/q_i and H = G^2 - p''/p
// = -G' come with the LDL^T pivots q_i of the Sturm sequence and their
// first two derivatives.  A step that passes the root in rounding (Sturm
// count >= 1), or slow convergence (a nearly double lowest value), hands the
// bracket to bisection.
inline double lowest_ritz_error(const double* a, const double* b, int m, double* lambda_out,
                                double hint_lo = 0.0, double hint_hi = -1.0) {
  double lo = a[0], hi = a[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  const double tiny = std::numeric_limits<double>::min();
  auto count_below = [&](double x) {  // eigenvalues of T[0..m) below x
    int c = 0;
    double q = a[0] - x;
    if (q < 0) ++c;
    for (int i = 1; i < m; ++i) {
      if (std::fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
      q = a[i] - x - b[i - 1] * b[i - 1] / q;
      if (q < 0) ++c;
    }
    return c;
  };
  if (hint_lo < hint_hi && hint_lo >= lo && hint_hi <= hi && count_below(hint_lo) == 0 && count_below(hint_hi) >= 1) {
    lo = hint_lo;
    hi = hint_hi;
  }
  // Laguerre from lo: every iterate stays below the lowest eigenvalue.  Its
  // last step lands within rounding of the root (on either side), which two
  // Sturm counts at +-4 eps confirm; otherwise bisection goes on from the
  // bracket found so far.
  const double scale = 1e-3 * std::max(std::fabs(lo), std::fabs(hi));
  auto confirm = [&](double xc) {
    const double d = 4 * eps * std::max(std::fabs(xc), scale);
    if (count_below(xc - d) == 0 && count_below(xc + d) >= 1) {
      lo = std::max(lo, xc - d);
      hi = std::min(hi, xc + d);
      return true;
    }
    return false;
  };
  double x = lo;
  for (int it = 0; it < 24; ++it) {
    int c = 0;
    double q = a[0] - x, dq = -1.0, d2q = 0.0, g = 0.0, h = 0.0;
    if (q < 0) ++c;
    for (int i = 1; i <= m; ++i) {
      if (std::fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
      const double r = 1.0 / q, u = dq * r;
      g += u;                // q'/q
      h += u * u - d2q * r;  // -(q'/q)' = (q'/q)^2 - q''/q
      if (i == m) break;
      const double bb = b[i - 1] * b[i - 1] * r * r;  // b^2 / q^2
      d2q = bb * (d2q - 2.0 * dq * u);
      dq = -1.0 + bb * dq;
      q = a[i] - x - bb * q;
      if (q < 0) ++c;
    }
    if (c >= 1) {  // passed the root in rounding
      if (!confirm(x)) hi = std::min(hi, x);
      break;
    }
    lo = x;
    if (!(g < 0.0)) break;  // no usable slope: bisection decides
    const double n = (double)m;
    const double disc = std::sqrt(std::max(0.0, (n - 1.0) * (n * h - g * g)));
    const double step = -n / (g - disc);
    const double xn = x + step;
    if (!(xn > x)) {  // a step below one ulp: x is the root to rounding
      confirm(x);
      break;
    }
    if (xn >= hi) break;
    if (step <= 2 * eps * std::max(std::fabs(x), std::fabs(xn))) {
      confirm(xn);
      break;
    }
    x = xn;
  }
  for (int it = 0; it < 200 && hi - lo > 2 * eps * std::max(std::fabs(lo), std::fabs(hi)); ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) >= 1) hi = mid;
    else lo = mid;
  }
  const double lam = 0.5 * (lo + hi);
  if (lambda_out) *lambda_out = lam;
  // z_{m-1} = 1, then z_{i-1} = ((lam - a_i) z_i - b_i z_{i+1}) / b_{i-1}:
  // the backward three-term recurrence for the eigenvector, stable exactly
  // when the Ritz value has converged (the last components are small); for
  // an unconverged value it can underestimate the bound, which only sends
  // the test on to the full analysis.
  double z1 = 1.0, z2 = 0.0, nrm2 = 1.0, zlast = 1.0;
  for (int i = m - 1; i >= 1; --i) {
    const double zi = ((lam - a[i]) * z1 - (i + 1 < m ? b[i] * z2 : 0.0)) / b[i - 1];
    z2 = z1;
    z1 = zi;
    nrm2 += zi * zi;
    if (nrm2 > 1e200) {  // rescale: only the ratio zlast / |z| matters
      const double s = 1e-100;
      z1 *= s;
      z2 *= s;
      zlast *= s;
      nrm2 *= s * s;
    }
  }
  return std::fabs(b[m - 1]) * std::fabs(zlast) / std::sqrt(nrm2);
}

// Number of eigenvalues of T[0..m) below x (Sturm sequence of the LDL^T
// pivots).
inline int sturm_count_below(const double* a, const double* b, int m, double x) {
  const double tiny = std::numeric_limits<double>::min();
  int c = 0;
  double q = a[0] - x;
  if (q < 0) ++c;
  for (int i = 1; i < m; ++i) {
    if (std::fabs(q) < tiny) q = q < 0 ? -tiny : tiny;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++c;
  }
  return c;
}

// A necessary condition for nlowest_converged(analyze_prefix(a, b, m, fp32),
// nev, fp32), in O(m): the lowest eigenvalue of T[0..m) has an error bound
// within 4 times the stopping tolerance (the slack covers the difference
// between this bound's backward recurrence and the full analysis'
// eigenvectors), or another eigenvalue of T lies within
// 1e-5 max(|T|, 1) of it.  In the second case the full analysis may merge
// the two (the FP64 rule's thold, tmatrix.h:149-150, is about 1e-10 |T|; the
// FP32 near-copy window 64 eps_f32 max(|lambda|, 1) is below 1e-5 max(|T|,
// 1)) and give the group
// a smaller bound, so the full analysis has to decide.  When the test fails
// the stopping rule cannot hold, because the lowest distinct value of the
// analysis is then this eigenvalue with this bound.
inline bool may_be_converged(const double* a, const double* b, int m, bool fp32, double* lambda_out = nullptr,
                             double hint_lo = 0.0, double hint_hi = -1.0) {
  double lam = 0.0;
  const double err = lowest_ritz_error(a, b, m, &lam, hint_lo, hint_hi);
  if (lambda_out) *lambda_out = lam;
  const double tol = stop_tolerance(fp32);
  if (!(err > 4.0 * std::max(tol, tol * std::fabs(lam)))) return true;  // also NaN: the full analysis reports it
  double tn = 0.0;
  for (int i = 0; i < m; ++i)
    tn = std::max(tn, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0));
  return sturm_count_below(a, b, m, lam + 1e-5 * std::max({tn, std::fabs(lam), 1.0})) >= 2;
}

// Eigenvector of T[0..m) for the (already accurate) eigenvalue lambda by
// inverse iteration: Gaussian elimination with partial pivoting on the
// tridiagonal T - lambda I (zero pivots replaced by eps |T|), three solves
// from a fixed start, normalized.  z has m entries.
inline void tridiag_eigenvector(const double* a, const double* b, int m, double lambda, double* z) {
  const double eps = std::numeric_limits<double>::epsilon();
  double tnorm = 0.0;
  for (int i = 0; i < m; ++i)
    tnorm = std::max(tnorm, std::fabs(a[i]) + (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0));
  const double small = eps * std::max(tnorm, std::numeric_limits<double>::min());
  // LU of T - lambda I with row interchanges (the dgttrf scheme): U has the
  // diagonals dd, du, du2; dl holds the multipliers.
  std::vector<double> dd(m), dl(m, 0.0), du(m, 0.0), du2(m, 0.0);
  std::vector<char> swp(m, 0);
  for (int i = 0; i < m; ++i) {
    dd[i] = a[i] - lambda;
    if (i + 1 < m) dl[i] = du[i] = b[i];
  }
  auto guard = [&](double v) { return std::fabs(v) < small ? (v < 0 ? -small : small) : v; };
  for (int i = 0; i + 1 < m; ++i) {
    if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
      dd[i] = guard(dd[i]);
      const double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
    } else {
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < m) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      swp[i] = 1;
    }
  }
  dd[m - 1] = guard(dd[m - 1]);
  std::vector<double> x(m, 1.0);
  for (int it = 0; it < 3; ++it) {
    // forward: the row interchanges and multipliers
    for (int i = 0; i + 1 < m; ++i) {
      if (swp[i]) std::swap(x[i], x[i + 1]);
      x[i + 1] -= dl[i] * x[i];
    }
    // back substitution
    for (int i = m - 1; i >= 0; --i) {
      double v = x[i];
      if (i + 1 < m) v -= du[i] * x[i + 1];
      if (i + 2 < m) v -= du2[i] * x[i + 2];
      x[i] = v / dd[i];
    }
    double nrm = 0.0;
    for (double v : x) nrm += v * v;
    nrm = std::sqrt(nrm);
    for (double& v : x) v /= nrm;
  }
  for (int i = 0; i < m; ++i) z[i] = x[i];
}

// The lowest `want` entries of analyze_prefix(a, b, m, fp32) without
// diagonalizing T: the same rules (tmatrix.h:148-217 as restated in
// include/ietl/tmatrix.h, and the FP32 near-copy merge above) applied from
// the bottom of the spectrum up, with each eigenvalue that an entry needs
// found by Sturm bisection and each error bound |beta_{m-1} z_{m-1}| from
// the inverse-iteration eigenvector.  Per entry O(m) passes instead of the
// two O(m^2) QL diagonalizations, for stopping tests that read only the
// lowest few values.
//   - the extreme eigenvalues give thold and multol as in the full analysis;
//   - a group is the eigenvalues within thold of its first (a Sturm count);
//   - a one-member group is a ghost when T without its first row and column
//     has an eigenvalue within multol of it, not counting that matrix's
//     largest eigenvalue (the full analysis' scan stops short of it);
//   - with fp32, later values within 64 eps_f32 max(1, |last|) of the last
//     entry merge into it (the smaller error bound is kept), so one more
//     value is examined after the last entry.
// Equal to the full analysis up to the rounding of the eigenvalues (ulps),
// for eigenvalues that do not fall within rounding of a group or ghost
// boundary.
inline TValues analyze_low(const std::vector<double>& a, const std::vector<double>& b, int m, bool fp32, int want) {
  TValues t;
  t.complete = false;
  t.all.assign(m, std::numeric_limits<double>::quiet_NaN());
  if (m <= 0 || want <= 0) return t;
  double amax = a[0], bmax = b[0];
  for (int i = 1; i < m; ++i) {
    amax = std::max(amax, a[i]);
    bmax = std::max(bmax, b[i]);
  }
  double glo = a[0], ghi = a[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    glo = std::min(glo, a[i] - r);
    ghi = std::max(ghi, a[i] + r);
  }
  const double eps = std::numeric_limits<double>::epsilon();
  // eigenvalues of T below x, and of T2 = T without its first row and column
  auto count = [&](double x) { return sturm_count_below(a.data(), b.data(), m, x); };
  auto count2 = [&](double x) { return m > 1 ? sturm_count_below(a.data() + 1, b.data() + 1, m - 1, x) : 0; };
  // eigenvalue k (0-based) of T, given lo with count(lo) <= k
  auto kth = [&](int k, double lo) {
    double hi = ghi + eps * std::max(1.0, std::fabs(ghi));
    for (int it = 0; it < 300 && hi - lo > 2 * eps * std::max(std::fabs(lo), std::fabs(hi)); ++it) {
      const double mid = 0.5 * (lo + hi);
      if (mid <= lo || mid >= hi) break;
      if (count(mid) > k) hi = mid;
      else lo = mid;
    }
    return 0.5 * (lo + hi);
  };
  // the extreme eigenvalues by the Laguerre iteration of the convergence
  // test (the largest as the lowest of -T)
  double lam0 = 0.0, lamN = 0.0;
  lowest_ritz_error(a.data(), b.data(), m, &lam0);
  {
    std::vector<double> na(m);
    for (int i = 0; i < m; ++i) na[i] = -a[i];
    lowest_ritz_error(na.data(), b.data(), m, &lamN);
    lamN = -lamN;
  }
  const double error_tol = std::pow(eps, 2. / 3.);
  const double multol = std::max(amax, bmax) * 2 * eps * (10000 + m);
  const double thold = std::max(error_tol * std::max(lam0, lamN), 5 * multol);
  const double bn = b[m - 1];
  const double dup = 64.0 * std::numeric_limits<float>::epsilon();
  std::vector<double> z(m);
  int idx = 0;  // index of the next group's first eigenvalue
  double first = lam0;
  bool pending_merge_check = false;
  while (idx < m) {
    if (idx > 0) {
      // the next group's first value: beyond the previous group's end
      if (pending_merge_check) {
        // after `want` entries only a value that would merge matters
        const double lim = t.ev.back() + dup * std::max(1.0, std::fabs(t.ev.back()));
        if (count(lim) <= idx) return t;
      }
      first = kth(idx, first + thold);
    }
    t.all[idx] = first;
    const int end = std::max(idx + 1, std::min(m, count(first + thold)));  // members: eigenvalues <= first + thold
    const int g = end - idx;
    bool ghost = false;
    double err = 0.0;
    if (g == 1) {
      const int clo = count2(first - multol), chi = count2(first + multol);
      ghost = chi > clo && clo <= m - 3;
      if (!ghost) {
        tridiag_eigenvector(a.data(), b.data(), m, first, z.data());
        err = std::fabs(bn * z[m - 1]);
      }
    }
    if (!ghost) {
      if (fp32 && !t.ev.empty() && first - t.ev.back() <= dup * std::max(1.0, std::fabs(t.ev.back()))) {
        t.err.back() = std::min(t.err.back(), err);
      } else {
        if (pending_merge_check) return t;  // a new entry beyond `want`
        t.ev.push_back(first);
        t.err.push_back(err);
        t.index.push_back(idx);
      }
      if ((int)t.ev.size() >= want) {
        if (!fp32) return t;
        pending_merge_check = true;
      }
    }
    idx = end;
  }
  t.complete = true;
  return t;
}

// The analysis a stopping test and its results read: bottom-up (the lowest
// `nev` entries) for FP32 storage, whose tolerance (1e-8) is far above the
// differences between the two analyses' error bounds; the full analysis for
// FP64 storage, whose tolerance (100 eps) is not: next to a ghost the
// inverse-iteration bound of analyze_low can exceed the QL one (the route
// of the reference's LAPACK dstev) by 1e-12 |T|.
inline TValues stop_analysis(const std::vector<double>& a, const std::vector<double>& b, int m, bool fp32, int nev) {
  return fp32 ? analyze_low(a, b, m, true, nev) : analyze_prefix(a, b, m, false);
}

}  // namespace host
}  // namespace alps_ed
