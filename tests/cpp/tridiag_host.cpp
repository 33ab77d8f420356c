// Host checks of the T-matrix helpers (alps_amd/csrc/ed_tridiag.hpp), no
// device: the lowest eigenvalue of lowest_ritz_error against plain Sturm
// bisection, its error bound and tridiag_eigenvector against cyclic Jacobi
// eigenvectors of the dense T, may_be_converged as a necessary condition of
// the full nlowest rule on T matrices of real Lanczos runs, and the
// bottom-up analyze_low against the full analyze_prefix.
// Prints "name value" lines for tests/test_tridiag_host.py.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include "ed_tridiag.hpp"

using namespace alps_ed::host;

// eigenvalues of T[0..m) below x, and the lowest one by bisection to the
// last bit (the textbook method, independent of the code under test)
static int count_below(const std::vector<double>& a, const std::vector<double>& b, int m, double x) {
  int c = 0;
  double q = a[0] - x;
  if (q < 0) ++c;
  for (int i = 1; i < m; ++i) {
    if (q == 0.0) q = 1e-300;
    q = a[i] - x - b[i - 1] * b[i - 1] / q;
    if (q < 0) ++c;
  }
  return c;
}

static double lowest_by_bisection(const std::vector<double>& a, const std::vector<double>& b, int m) {
  double lo = 1e300, hi = -1e300;
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? std::fabs(b[i - 1]) : 0.0) + (i + 1 < m ? std::fabs(b[i]) : 0.0);
    lo = std::min(lo, a[i] - r);
    hi = std::max(hi, a[i] + r);
  }
  for (int it = 0; it < 5000; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(a, b, m, mid) >= 1) hi = mid;
    else lo = mid;
  }
  return hi;
}

// cyclic Jacobi on the dense symmetric matrix; returns the eigenvector of
// the lowest eigenvalue
static std::vector<double> jacobi_lowest(const std::vector<double>& a, const std::vector<double>& b, int m,
                                         double& lam) {
  std::vector<double> A(m * m, 0.0), V(m * m, 0.0);
  for (int i = 0; i < m; ++i) {
    A[i * m + i] = a[i];
    V[i * m + i] = 1.0;
    if (i + 1 < m) A[i * m + i + 1] = A[(i + 1) * m + i] = b[i];
  }
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) off += A[p * m + q] * A[p * m + q];
    if (off < 1e-60) break;
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = A[p * m + q];
        if (std::fabs(apq) < 1e-300) continue;
        const double th = (A[q * m + q] - A[p * m + p]) / (2 * apq);
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
        const double c = 1 / std::sqrt(t * t + 1), s = t * c;
        for (int k = 0; k < m; ++k) {
          const double akp = A[k * m + p], akq = A[k * m + q];
          A[k * m + p] = c * akp - s * akq;
          A[k * m + q] = s * akp + c * akq;
        }
        for (int k = 0; k < m; ++k) {
          const double apk = A[p * m + k], aqk = A[q * m + k];
          A[p * m + k] = c * apk - s * aqk;
          A[q * m + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < m; ++k) {
          const double vkp = V[k * m + p], vkq = V[k * m + q];
          V[k * m + p] = c * vkp - s * vkq;
          V[k * m + q] = s * vkp + c * vkq;
        }
      }
  }
  int j = 0;
  for (int i = 1; i < m; ++i)
    if (A[i * m + i] < A[j * m + j]) j = i;
  lam = A[j * m + j];
  std::vector<double> z(m);
  for (int k = 0; k < m; ++k) z[k] = V[k * m + j];
  return z;
}

// alpha/beta of a plain FP64 Lanczos run on a random symmetric sparse
// matrix (a random-bond XXZ chain's Sz=0 sector would do as well; any
// symmetric operator gives a T with the same structure)
static void lanczos_run(std::mt19937_64& rng, int n, int steps, std::vector<double>& al, std::vector<double>& be) {
  std::uniform_real_distribution<double> U(-1, 1);
  std::vector<std::vector<std::pair<int, double>>> rows(n);
  for (int i = 0; i < n; ++i) {
    rows[i].push_back({i, 2.0 * U(rng)});
    for (int k = 0; k < 3; ++k) {
      const int j = (int)(rng() % n);
      if (j == i) continue;
      const double v = U(rng);
      rows[i].push_back({j, v});
      rows[j].push_back({i, v});
    }
  }
  std::vector<double> v(n), vp(n, 0.0), w(n);
  double nrm = 0.0;
  for (double& x : v) {
    x = U(rng);
    nrm += x * x;
  }
  for (double& x : v) x /= std::sqrt(nrm);
  double beta = 0.0;
  for (int s = 0; s < steps; ++s) {
    for (int i = 0; i < n; ++i) {
      double acc = 0.0;
      for (auto& e : rows[i]) acc += e.second * v[e.first];
      w[i] = acc - beta * vp[i];
    }
    double alpha = 0.0;
    for (int i = 0; i < n; ++i) alpha += w[i] * v[i];
    nrm = 0.0;
    for (int i = 0; i < n; ++i) {
      w[i] -= alpha * v[i];
      nrm += w[i] * w[i];
    }
    beta = std::sqrt(nrm);
    al.push_back(alpha);
    be.push_back(beta);
    for (int i = 0; i < n; ++i) {
      vp[i] = v[i];
      v[i] = w[i] / beta;
    }
  }
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U(-1, 1), P(0.05, 1.5);
  double lam_rel = 0.0, err_over = 0.0, vec_bad = 0.0;
  long cases = 0, hint_bad = 0;
  for (int t = 0; t < 3000; ++t) {
    const int m = 1 + (int)(rng() % 120);
    const double scale = std::pow(10.0, (double)(rng() % 9) - 3.0);
    std::vector<double> a(m), b(m);
    for (int i = 0; i < m; ++i) {
      a[i] = scale * 3 * U(rng);
      b[i] = scale * P(rng);
    }
    switch (t % 5) {
      case 1:  // tiny couplings: nearly decoupled blocks, near-double lowest values
        for (int i = 0; i < m; i += 7) b[i] = scale * 1e-13 * P(rng);
        break;
      case 2:  // a repeated 2x2 block: the lowest value nearly doubled
        for (int i = 0; i < m; ++i) {
          a[i] = scale * (i % 2 ? 1.0 : -1.0);
          b[i] = (i % 2) ? scale * 1e-9 : scale * 0.7;
        }
        break;
      case 3:  // Wilkinson-like: |i - m/2| on the diagonal, unit couplings
        for (int i = 0; i < m; ++i) {
          a[i] = scale * std::fabs(i - 0.5 * (m - 1));
          b[i] = scale;
        }
        break;
      default:
        break;
    }
    double lam = 0.0;
    const double err = lowest_ritz_error(a.data(), b.data(), m, &lam);
    const double ref = lowest_by_bisection(a, b, m);
    double tn = 0.0;
    for (int i = 0; i < m; ++i) tn = std::max(tn, std::fabs(a[i]) + std::fabs(b[i]) + (i ? std::fabs(b[i - 1]) : 0.0));
    lam_rel = std::max(lam_rel, std::fabs(lam - ref) / tn);
    // the previous value of a longer prefix as a hint gives the same value
    if (m > 2) {
      double lam_h = 0.0;
      lowest_ritz_error(a.data(), b.data(), m, &lam_h, ref - 0.1 * scale, ref + 0.5 * scale);
      if (std::fabs(lam_h - ref) > 8e-16 * tn) ++hint_bad;
    }
    if (m <= 40) {
      double lj = 0.0;
      const std::vector<double> z = jacobi_lowest(a, b, m, lj);
      // the error bound |beta_{m-1} z_{m-1}| against Jacobi's vector, where
      // the lowest value is well separated (else the vector is not defined)
      const int below = count_below(a, b, m, lj + 1e-6 * tn);
      if (below == 1) {
        const double want = std::fabs(b[m - 1] * z[m - 1]);
        // the backward recurrence may underestimate an unconverged value's
        // bound (which only sends the gate on to the full analysis) but must
        // not overestimate a converged one
        err_over = std::max(err_over, (err - want) / std::max(want, 1e-13 * tn));
        std::vector<double> zi(m);
        tridiag_eigenvector(a.data(), b.data(), m, ref, zi.data());
        double dot = 0.0;
        for (int i = 0; i < m; ++i) dot += zi[i] * z[i];
        vec_bad = std::max(vec_bad, 1.0 - std::fabs(dot));
      }
    }
    ++cases;
  }
  // necessary condition: wherever the full rule holds, the gate passes
  long gate_checks = 0, gate_bad = 0, rule_true = 0, gate_true = 0;
  for (int run = 0; run < 40; ++run) {
    std::vector<double> al, be;
    lanczos_run(rng, 300 + 50 * (run % 5), 90, al, be);
    for (int m = 2; m <= (int)al.size(); ++m)
      for (int fp32 = 0; fp32 < 2; ++fp32) {
        const TValues tv = analyze_prefix(al, be, m, fp32 != 0);
        const bool rule = nlowest_converged(tv, 1 + run % 3, fp32 != 0);
        const bool gate = may_be_converged(al.data(), be.data(), m, fp32 != 0);
        ++gate_checks;
        rule_true += rule;
        gate_true += gate;
        if (rule && !gate) ++gate_bad;
      }
  }
  std::printf("cases %ld\nlam_rel %.3e\nhint_bad %ld\nerr_over %.3e\nvec_bad %.3e\n", cases, lam_rel, hint_bad,
              err_over, vec_bad);
  std::printf("gate_checks %ld\nrule_true %ld\ngate_true %ld\ngate_bad %ld\n", gate_checks, rule_true, gate_true,
              gate_bad);
  // the bottom-up analysis against the full one: the same lowest entries
  // (values to rounding, indices, bounds) and the same stopping decisions,
  // on Lanczos T matrices long enough to hold ghosts and on random ones
  long low_checks = 0, low_index_bad = 0, low_rule_bad32 = 0, low_rule_bad64 = 0, low_count_bad = 0, ghost_seen = 0;
  double low_ev = 0.0, low_err = 0.0;
  for (int run = 0; run < 60; ++run) {
    std::vector<double> al, be;
    if (run % 3 == 2) {
      const int m = 2 + (int)(rng() % 80);
      for (int i = 0; i < m; ++i) {
        al.push_back(3 * U(rng));
        be.push_back(P(rng));
      }
    } else {
      lanczos_run(rng, 120 + 40 * (run % 4), 150, al, be);
    }
    for (int m = 1; m <= (int)al.size(); m += (m < 20 ? 1 : 7))
      for (int fp32 = 0; fp32 < 2; ++fp32)
        for (int want = 1; want <= 4; ++want) {
          const TValues full = analyze_prefix(al, be, m, fp32 != 0);
          const TValues low = analyze_low(al, be, m, fp32 != 0, want);
          ++low_checks;
          double tn = 0.0;
          for (int i = 0; i < m; ++i) tn = std::max(tn, std::fabs(al[i]) + std::fabs(be[i]) + (i ? std::fabs(be[i - 1]) : 0.0));
          const size_t k = std::min<size_t>(want, full.ev.size());
          if (low.ev.size() < k || (low.complete && low.ev.size() != full.ev.size())) {
            ++low_count_bad;
            continue;
          }
          for (size_t i = 0; i < k; ++i) {
            low_ev = std::max(low_ev, std::fabs(low.ev[i] - full.ev[i]) / tn);
            low_err = std::max(low_err, std::fabs(low.err[i] - full.err[i]) / tn);
            if (low.index[i] != full.index[i]) ++low_index_bad;
          }
          if (nlowest_converged(low, want, fp32 != 0) != nlowest_converged(full, want, fp32 != 0))
            ++(fp32 ? low_rule_bad32 : low_rule_bad64);
          int ghosts = 0;
          for (size_t i = 0; i + 1 < full.index.size(); ++i) ghosts += full.index[i + 1] - full.index[i] > 1;
          ghost_seen += ghosts > 0;
        }
  }
  // FP64 decisions may differ next to a ghost (the bounds differ by up to
  // 1e-12 |T| there, the FP64 tolerance is 100 eps), which is why FP64
  // stopping tests keep the full analysis
  std::printf("low_checks %ld\nlow_count_bad %ld\nlow_index_bad %ld\nlow_rule_bad32 %ld\nlow_rule_bad64 %ld\n"
              "low_ev %.3e\nlow_err %.3e\nghost_seen %ld\n",
              low_checks, low_count_bad, low_index_bad, low_rule_bad32, low_rule_bad64, low_ev, low_err, ghost_seen);
  return 0;
}
