// Host-side small dense linear algebra for the Krylov solvers: the projected
// matrices are at most a few thousand wide, so they stay on the CPU.
#pragma once

#include <algorithm>
#include <cmath>
#include <limits>
#include <numeric>
#include <vector>

namespace qsi {
namespace hostla {

// All eigenpairs of the symmetric tridiagonal matrix with diagonal d (n) and
// off-diagonal e (n-1, e[i] couples i and i+1) by the implicit QL algorithm
// with Wilkinson-type shifts.  evals ascending.  If Z is non-null it receives
// the n x n eigenvector matrix, column-major (column j = eigenvector j).  If
// first_row is non-null (and Z is null) it receives only row 0 of that matrix
// — the first components needed by Gauss quadrature (FTLM) — in O(n^2).
inline bool tridiag_eig(std::vector<double> d, std::vector<double> e,
                        std::vector<double>& evals, std::vector<double>* Z,
                        std::vector<double>* first_row = nullptr) {
  const int n = static_cast<int>(d.size());
  evals.clear();
  if (n == 0) return true;
  e.resize(n, 0.0);
  e[n - 1] = 0.0;
  std::vector<double> z;     // full vectors, column-major
  std::vector<double> zrow;  // row 0 only
  if (Z) {
    z.assign(static_cast<size_t>(n) * n, 0.0);
    for (int i = 0; i < n; ++i) z[static_cast<size_t>(i) * n + i] = 1.0;
  } else if (first_row) {
    zrow.assign(n, 0.0);
    zrow[0] = 1.0;
  }
  const double eps = std::numeric_limits<double>::epsilon();
  for (int l = 0; l < n; ++l) {
    for (int iter = 0;; ++iter) {
      int m = l;
      for (; m < n - 1; ++m) {
        const double dd = std::fabs(d[m]) + std::fabs(d[m + 1]);
        if (std::fabs(e[m]) <= eps * dd) break;
      }
      if (m == l) break;
      if (iter >= 200) return false;
      // shift from the leading 2x2 block of the unreduced part
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double r = std::hypot(g, 1.0);
      g = d[m] - d[l] + e[l] / (g + std::copysign(r, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = m - 1; i >= l; --i) {
        const double f = s * e[i];
        const double b = c * e[i];
        r = std::hypot(f, g);
        e[i + 1] = r;
        if (r == 0.0) {  // split: the rotation chain ends here
          d[i + 1] -= p;
          e[m] = 0.0;
          deflated = true;
          break;
        }
        s = f / r;
        c = g / r;
        g = d[i + 1] - p;
        r = (d[i] - g) * s + 2.0 * c * b;
        p = s * r;
        d[i + 1] = g + p;
        g = c * r - b;
        if (Z) {
          double* zi = &z[static_cast<size_t>(i) * n];
          double* zi1 = &z[static_cast<size_t>(i + 1) * n];
          for (int k = 0; k < n; ++k) {
            const double t = zi1[k];
            zi1[k] = s * zi[k] + c * t;
            zi[k] = c * zi[k] - s * t;
          }
        } else if (first_row) {
          const double t = zrow[i + 1];
          zrow[i + 1] = s * zrow[i] + c * t;
          zrow[i] = c * zrow[i] - s * t;
        }
      }
      if (deflated) continue;
      d[l] -= p;
      e[l] = g;
      e[m] = 0.0;
    }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return d[a] < d[b]; });
  evals.resize(n);
  for (int j = 0; j < n; ++j) evals[j] = d[idx[j]];
  if (Z) {
    Z->assign(static_cast<size_t>(n) * n, 0.0);
    for (int j = 0; j < n; ++j)
      std::copy(z.begin() + static_cast<size_t>(idx[j]) * n,
                z.begin() + static_cast<size_t>(idx[j] + 1) * n,
                Z->begin() + static_cast<size_t>(j) * n);
  } else if (first_row) {
    first_row->resize(n);
    for (int j = 0; j < n; ++j) (*first_row)[j] = zrow[idx[j]];
  }
  return true;
}

// Number of eigenvalues of the tridiagonal (d, e) below x (Sturm count).
inline int sturm_count(const std::vector<double>& d, const std::vector<double>& e,
                       size_t n, double x) {
  int cnt = 0;
  double q = 1.0;
  for (size_t i = 0; i < n; ++i) {
    q = d[i] - x - (i > 0 ? e[i - 1] * e[i - 1] / q : 0.0);
    if (q == 0.0) q = -std::numeric_limits<double>::min();
    if (q < 0.0) ++cnt;
  }
  return cnt;
}

// Lowest eigenvalue of the leading n x n block of (d, e) by bisection.
inline double tridiag_lowest(const std::vector<double>& d,
                             const std::vector<double>& e, size_t n) {
  double lo = std::numeric_limits<double>::max(), hi = -lo;
  for (size_t i = 0; i < n; ++i) {
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) +
                     (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  for (int it = 0; it < 256; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (!(mid > lo && mid < hi)) break;
    if (sturm_count(d, e, n, mid) >= 1) hi = mid;
    else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Unit eigenvector of the tridiagonal (d, e) (leading n) for eigenvalue lam
// by inverse iteration with a row-pivoted LU of T - lam I, started from e_1.
// When lam belongs to a cluster of numerically equal eigenvalues (the ghost
// copies a long Lanczos run produces), every cluster direction is amplified
// alike, so the result is the normalised projection of e_1 onto the cluster:
// the copy carried by the first Lanczos vectors, i.e. the Ritz vector of the
// converged eigenvalue rather than one of its later ghosts.  Returned with a
// non-negative first component.
inline std::vector<double> tridiag_eigvec(const std::vector<double>& d,
                                          const std::vector<double>& e,
                                          size_t n, double lam) {
  std::vector<double> x(n, 0.0);
  x[0] = 1.0;
  if (n == 1) return x;
  double tn = 0.0;
  for (size_t i = 0; i < n; ++i)
    tn = std::max(tn, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) +
                          (i + 1 < n ? std::fabs(e[i]) : 0.0));
  const double tiny = std::max(tn, std::numeric_limits<double>::min()) * 1e-15;
  // rows: sub l[i] (i+1,i), diag a[i], super u[i] (i,i+1), second super w[i]
  std::vector<double> l(n - 1), a(n), u(n - 1), w(n, 0.0);
  std::vector<char> piv(n - 1, 0);
  // shift just below the cluster: all copies of lam are amplified alike and
  // the next distinct eigenvalue is suppressed by (shift / gap) per pass
  const double sigma = lam - std::max(1e-10 * tn, tiny);
  for (size_t i = 0; i < n; ++i) a[i] = d[i] - sigma;
  for (size_t i = 0; i + 1 < n; ++i) l[i] = u[i] = e[i];
  for (size_t i = 0; i + 1 < n; ++i) {
    if (std::fabs(a[i]) >= std::fabs(l[i])) {
      if (a[i] == 0.0) a[i] = tiny;
      const double f = l[i] / a[i];
      l[i] = f;
      a[i + 1] -= f * u[i];
    } else {
      const double f = a[i] / l[i];
      a[i] = l[i];
      l[i] = f;
      const double t = u[i];
      u[i] = a[i + 1];
      a[i + 1] = t - f * a[i + 1];
      if (i + 2 < n) {
        w[i] = u[i + 1];
        u[i + 1] = -f * u[i + 1];
      }
      piv[i] = 1;
    }
  }
  if (a[n - 1] == 0.0) a[n - 1] = tiny;
  auto iterate = [&](std::vector<double>& xv) {
    for (int pass = 0; pass < 4; ++pass) {
      std::vector<double> b = xv;
      for (size_t i = 0; i + 1 < n; ++i) {
        if (piv[i]) std::swap(b[i], b[i + 1]);
        b[i + 1] -= l[i] * b[i];
      }
      xv[n - 1] = b[n - 1] / a[n - 1];
      xv[n - 2] = (b[n - 2] - u[n - 2] * xv[n - 1]) / a[n - 2];
      for (size_t k = n - 2; k-- > 0;)
        xv[k] = (b[k] - u[k] * xv[k + 1] - w[k] * xv[k + 2]) / a[k];
      double s = 0.0;
      for (double v : xv) s += v * v;
      s = std::sqrt(s);
      if (xv[0] < 0) s = -s;
      for (auto& v : xv) v /= s;
    }
  };
  auto residual = [&](const std::vector<double>& xv) {
    double r = 0.0;
    for (size_t i = 0; i < n; ++i) {
      double t = (d[i] - lam) * xv[i];
      if (i > 0) t += e[i - 1] * xv[i - 1];
      if (i + 1 < n) t += e[i] * xv[i + 1];
      r = std::max(r, std::fabs(t));
    }
    return r;
  };
  iterate(x);
  // e_1 can be nearly orthogonal to the wanted vector (a localised
  // eigenvector); then restart from a flat vector
  if (!(residual(x) <= 1e-8 * std::max(tn, tiny))) {
    x.assign(n, 1.0 / std::sqrt(static_cast<double>(n)));
    iterate(x);
  }
  return x;
}

// Dense symmetric eigen-decomposition (n x n column-major) by cyclic Jacobi
// rotations; evals ascending, V column j = eigenvector j.  For the small
// projected matrices of the thick-restart solver (n <= 64).
inline void sym_eig(std::vector<double> A, int n, std::vector<double>& evals,
                    std::vector<double>& V) {
  V.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) V[static_cast<size_t>(i) * n + i] = 1.0;
  auto at = [&](int i, int j) -> double& { return A[static_cast<size_t>(j) * n + i]; };
  for (int sweep = 0; sweep < 64; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const double v = at(i, j) * at(i, j);
        tot += v;
        if (i != j) off += v;
      }
    if (off == 0.0 || off <= 1e-32 * tot) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = at(p, q);
        if (apq == 0.0) continue;
        const double th = (at(q, q) - at(p, p)) / (2.0 * apq);
        const double t = std::copysign(1.0, th) / (std::fabs(th) + std::hypot(th, 1.0));
        const double c = 1.0 / std::hypot(t, 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {
          const double kp = at(k, p), kq = at(k, q);
          at(k, p) = c * kp - s * kq;
          at(k, q) = s * kp + c * kq;
        }
        for (int k = 0; k < n; ++k) {
          const double pk = at(p, k), qk = at(q, k);
          at(p, k) = c * pk - s * qk;
          at(q, k) = s * pk + c * qk;
        }
        double* vp = &V[static_cast<size_t>(p) * n];
        double* vq = &V[static_cast<size_t>(q) * n];
        for (int k = 0; k < n; ++k) {
          const double a = vp[k], b = vq[k];
          vp[k] = c * a - s * b;
          vq[k] = s * a + c * b;
        }
      }
  }
  std::vector<int> idx(n);
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return at(a, a) < at(b, b); });
  evals.resize(n);
  std::vector<double> V2(static_cast<size_t>(n) * n);
  for (int j = 0; j < n; ++j) {
    evals[j] = at(idx[j], idx[j]);
    std::copy(V.begin() + static_cast<size_t>(idx[j]) * n,
              V.begin() + static_cast<size_t>(idx[j] + 1) * n,
              V2.begin() + static_cast<size_t>(j) * n);
  }
  V.swap(V2);
}

}  // namespace hostla
}  // namespace qsi
