// ORACLE — test infrastructure only.
//
// A CPU restatement (C++17 + OpenMP) of the reference's matrix-free ring
// exchange H·v and its in-house Lanczos.  It exists to check the HIP library:
// only tests/, __graft_entry__.smoke() and bench.py's CPU baseline may load
// it, and never as the thing measured or shipped.
//
// Each function cites the reference code it follows (paths relative to the
// reference root).  Pins (see tests/test_oracle_*.py):
//   pmr_construct / apply_state  - known answers of test/test_PMR_construct.cpp
//                                  and the compiled reference (oracle/_ref)
//   basis search                 - the compiled reference ZBasisInterp/ZBasisBST
//                                  (oracle/_ref) on the fixture bases
//   evaluate_add                 - dense numpy H built from first principles
//                                  (32 spins), symmetry <u,Hv> = <Hu,v>,
//                                  row-sum identity, golden E0 (48 spins)
//   sparse triplets              - nnz 15360 / 703680 and the flippable
//                                  histogram measured on the fixture bases
//   lanczos_eigval0              - golden E0 of test/test_correct.sh and the
//                                  test/test_lanczos.cpp mock scenario
//   random_unit                  - libstdc++ std::mt19937 +
//                                  std::normal_distribution<double>, as used by
//                                  the reference (same library, same stream)
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <sstream>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

namespace {

using u128 = unsigned __int128;

inline u128 mk(uint64_t lo, uint64_t hi) { return (static_cast<u128>(hi) << 64) | lo; }
inline uint64_t lo64(u128 x) { return static_cast<uint64_t>(x); }
inline uint64_t hi64(u128 x) { return static_cast<uint64_t>(x >> 64); }
inline int popc(u128 x) { return __builtin_popcountll(lo64(x)) + __builtin_popcountll(hi64(x)); }
inline u128 bit(int i) { return static_cast<u128>(1) << i; }

// One term, same layout as qsi_term_t in include/qsi_gpu.h.
struct Term {
  double coeff;
  uint64_t X[2];
  uint64_t Z[2];
  uint64_t up[2];
  uint64_t down[2];
  int32_t sign;
  int32_t slot;
};

struct Pmr {
  u128 X = 0, Z = 0, down = 0, up = 0;
  int sign = 1;
};

Pmr from_term(const Term& t) {
  Pmr p;
  p.X = mk(t.X[0], t.X[1]);
  p.Z = mk(t.Z[0], t.Z[1]);
  p.up = mk(t.up[0], t.up[1]);
  p.down = mk(t.down[0], t.down[1]);
  p.sign = t.sign;
  return p;
}

// Right-to-left construction of the permutation-matrix representation,
// following SymbolicPMROperator::_construct
// (components/operator/include/operator.hpp:383-466).
// Returns false on an unknown op character (bad_operator_spec).
bool construct(const char* ops, const int* ids, int n, Pmr& p) {
  p = Pmr();
  for (int k = n - 1; k >= 0; --k) {
    const int j = ids[k];
    const u128 b = bit(j);
    switch (ops[k]) {
      case 'x': case 'X':
        p.X ^= b;
        if (p.Z & b) p.sign = -p.sign;
        break;
      case 'z': case 'Z':
        p.Z ^= b;
        break;
      case '+':
        if (p.down & p.X & b) { p.sign = 0; break; }  // '+' after '+'
        if (p.Z & b) p.sign = -p.sign;
        p.X ^= b;
        if (p.X & b) p.down |= b; else p.up |= b;
        break;
      case '-':
        if (p.up & p.X & b) { p.sign = 0; break; }  // '-' after '-'
        if (p.Z & b) p.sign = -p.sign;
        p.X ^= b;
        if (p.X & b) p.up |= b; else p.down |= b;
        break;
      case 'p': case 'P':  // projector on up (S+ S-)
        if (p.up & p.X & b) { p.sign = 0; break; }
        if (p.X & b) p.up |= b; else p.down |= b;
        break;
      case 'q': case 'Q':  // projector on down (S- S+)
        if (p.down & p.X & b) { p.sign = 0; break; }
        if (p.Z & b) p.sign = -p.sign;
        p.X ^= b;
        if (p.X & b) p.down |= b; else p.up |= b;
        break;
      default:
        return false;
    }
  }
  return true;
}

// SymbolicPMROperator::applyState (operator.hpp:232-254): returns the signed
// amplitude and replaces s by the image state (unchanged when it vanishes).
inline int apply_state(const Pmr& p, u128& s) {
  const int ok = ((s & p.down) == 0) && ((s & p.up) == p.up);
  s = ok ? (s ^ p.X) : s;
  const int par = popc(~s & p.Z) & 1;
  return p.sign * (1 - 2 * par) * ok;
}

// ZBasisInterp (operator.hpp:197-204; operator.cpp:38-86): a sorted state
// vector and a map from the high word to the [first, last] index range.
struct Basis {
  std::vector<u128> states;
  std::unordered_map<uint64_t, std::pair<int64_t, int64_t>> bounds;

  void build_bounds() {  // find_bounds, operator.cpp:73-86
    bounds.clear();
    for (int64_t J = 0; J < static_cast<int64_t>(states.size()); ++J) {
      const uint64_t h = hi64(states[J]);
      auto it = bounds.find(h);
      if (it == bounds.end()) bounds[h] = {J, J};
      else it->second.second = J;
    }
  }

  // search, operator.cpp:38-70.  Returns 1 (found), 0 (missing), -1 when the
  // high word has no bucket (the reference's bounds.at throws there).
  int search(u128 s, int64_t& J) const {
    auto it = bounds.find(hi64(s));
    if (it == bounds.end()) return -1;
    int64_t left = it->second.first, right = it->second.second;
    const uint64_t key_lo = lo64(s);
    while (right - left > 32) {  // bisect on the low word
      const int64_t mid = (left + right) / 2;
      if (lo64(states[mid]) < key_lo) left = mid + 1;
      else right = mid;
    }
    for (J = left; J <= right; ++J)  // linear scan of the remaining window
      if (states[J] == s) return 1;
    return 0;
  }

  // ZBasisBST::search, operator.cpp:4-34 (bisection on the full 128 bits)
  int search_bst(u128 s, int64_t& J) const {
    int64_t left = 0, right = static_cast<int64_t>(states.size()) - 1;
    while (right - left > 32) {
      const int64_t mid = (left + right) / 2;
      if (states[mid] < s) left = mid + 1;
      else right = mid;
    }
    for (J = left; J <= right; ++J)
      if (J >= 0 && states[J] == s) return 1;
    return 0;
  }
};

// SymbolicOpSum::add_term (operator.hpp:489-495): X == 0 -> diagonal list.
struct OpSum {
  std::vector<std::pair<double, Pmr>> diag, off;
  OpSum(const Term* t, int n) {
    for (int k = 0; k < n; ++k) {
      Pmr p = from_term(t[k]);
      if (p.X == 0) diag.emplace_back(t[k].coeff, p);
      else off.emplace_back(t[k].coeff, p);
    }
  }
};

constexpr double kApplyTol = 1e-15;  // LazyOpSum::APPLY_TOL, operator_matrix.hpp:113

// LazyOpSum::evaluate_add = evaluate_add_off_diag_omp + evaluate_add_diagonal
// (components/operator/src/operator_matrix.cpp:28-82,108-112): push form,
// term-outer loop, static row schedule, atomic scatter.
// The pushes of rows [r0, r1) only (the full apply is r0 = 0, r1 = D): used
// to time the algorithm on a bounded sample of a very large basis.
int evaluate_add_rows(const Basis& B, const OpSum& H, const double* x, double* y,
                      int64_t r0, int64_t r1) {
  int missing_bucket = 0;
  for (const auto& [c, op] : H.off) {
#pragma omp parallel for schedule(static) reduction(| : missing_bucket)
    for (int64_t i = r0; i < r1; ++i) {
      u128 s = B.states[i];
      int sgn = apply_state(op, s);
      if (sgn != 0 && std::fabs(x[i]) > kApplyTol) {
        int64_t J = -1;
        const int f = B.search(s, J);
        if (f < 0) { missing_bucket = 1; continue; }
        if (f == 0) continue;  // reference adds c*x*0 (a no-op)
        const double dy = c * x[i] * sgn;
#pragma omp atomic
        y[J] += dy;
      }
    }
  }
  for (const auto& [c, op] : H.diag) {
#pragma omp parallel for schedule(static)
    for (int64_t i = r0; i < r1; ++i) {
      u128 s = B.states[i];
      y[i] += c * x[i] * static_cast<double>(apply_state(op, s));
    }
  }
  return missing_bucket ? -1 : 0;
}

int evaluate_add(const Basis& B, const OpSum& H, const double* x, double* y) {
  return evaluate_add_rows(B, H, x, y, 0, static_cast<int64_t>(B.states.size()));
}

// ---- small dense symmetric eigensolver (cyclic Jacobi) ----------------------
// Used for the Lanczos tridiagonal matrix, standing in for Eigen's
// SelfAdjointEigenSolver (components/linalg_routines/include/lapack_adapter.hpp:
// 37-104).  Eigenvalues ascending, column j of V = eigenvector j.
void jacobi_eig(std::vector<double> A, int n, std::vector<double>& w,
                std::vector<double>& V) {
  V.assign(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) V[i * n + i] = 1.0;
  auto a = [&](int i, int j) -> double& { return A[static_cast<size_t>(j) * n + i]; };
  for (int sweep = 0; sweep < 100; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        tot += a(i, j) * a(i, j);
        if (i != j) off += a(i, j) * a(i, j);
      }
    if (off <= 1e-30 * tot || off == 0.0) break;
    for (int p = 0; p < n - 1; ++p)
      for (int q = p + 1; q < n; ++q) {
        const double apq = a(p, q);
        if (apq == 0.0) continue;
        const double theta = (a(q, q) - a(p, p)) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) /
                         (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < n; ++k) {  // A <- A J (columns p, q)
          const double akp = a(k, p), akq = a(k, q);
          a(k, p) = c * akp - s * akq;
          a(k, q) = s * akp + c * akq;
        }
        for (int k = 0; k < n; ++k) {  // A <- J^T A (rows p, q)
          const double apk = a(p, k), aqk = a(q, k);
          a(p, k) = c * apk - s * aqk;
          a(q, k) = s * apk + c * aqk;
        }
        for (int k = 0; k < n; ++k) {
          const double vkp = V[static_cast<size_t>(p) * n + k], vkq = V[static_cast<size_t>(q) * n + k];
          V[static_cast<size_t>(p) * n + k] = c * vkp - s * vkq;
          V[static_cast<size_t>(q) * n + k] = s * vkp + c * vkq;
        }
      }
  }
  std::vector<int> idx(n);
  for (int i = 0; i < n; ++i) idx[i] = i;
  std::sort(idx.begin(), idx.end(), [&](int x, int y) { return a(x, x) < a(y, y); });
  w.resize(n);
  std::vector<double> V2(static_cast<size_t>(n) * n);
  for (int j = 0; j < n; ++j) {
    w[j] = a(idx[j], idx[j]);
    std::copy(V.begin() + static_cast<size_t>(idx[j]) * n,
              V.begin() + static_cast<size_t>(idx[j] + 1) * n,
              V2.begin() + static_cast<size_t>(j) * n);
  }
  V.swap(V2);
}

// Lowest eigenvalue of the symmetric tridiagonal (d, e) by Sturm-sequence
// bisection; the convergence test of the reference only needs this number.
double tridiag_lowest(const std::vector<double>& d, const std::vector<double>& e) {
  const size_t n = d.size();
  double lo = std::numeric_limits<double>::max(), hi = -lo;
  for (size_t i = 0; i < n; ++i) {  // Gershgorin interval
    const double r = (i > 0 ? std::fabs(e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(e[i]) : 0.0);
    lo = std::min(lo, d[i] - r);
    hi = std::max(hi, d[i] + r);
  }
  auto count_below = [&](double x) {
    int cnt = 0;
    double q = 1.0;
    for (size_t i = 0; i < n; ++i) {
      const double e2 = i > 0 ? e[i - 1] * e[i - 1] : 0.0;
      q = d[i] - x - (i > 0 ? e2 / q : 0.0);
      if (q == 0.0) q = -1e-300;
      if (q < 0) ++cnt;
    }
    return cnt;
  };
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid <= lo || mid >= hi) break;
    if (count_below(mid) >= 1) hi = mid; else lo = mid;
  }
  return 0.5 * (lo + hi);
}

// Eigenvector of the symmetric tridiagonal (d, e) for the eigenvalue lam by
// inverse iteration with a pivoted tridiagonal LU of (T - lam I), started from
// e_1: inside a cluster of ghost copies of lam this converges to the
// projection of e_1 on the cluster, the copy carried by the first Lanczos
// vectors.  First component made non-negative.
std::vector<double> tridiag_inverse_iteration(const std::vector<double>& d,
                                              const std::vector<double>& e,
                                              double lam) {
  const int n = static_cast<int>(d.size());
  std::vector<double> x(n, 0.0);
  x[0] = 1.0;
  if (n == 1) return x;
  double tnorm = 0.0;
  for (int i = 0; i < n; ++i)
    tnorm = std::max(tnorm, std::fabs(d[i]) + (i > 0 ? std::fabs(e[i - 1]) : 0.0) +
                                (i + 1 < n ? std::fabs(e[i]) : 0.0));
  const double tiny = std::max(tnorm, 1e-300) * 1e-15;
  std::vector<double> dl(n - 1), dd(n), du(n - 1), du2(std::max(n - 2, 0), 0.0);
  std::vector<int> swp(n - 1, 0);
  // shift just below the cluster so every ghost copy is amplified alike
  const double sigma = lam - std::max(1e-10 * tnorm, tiny);
  for (int i = 0; i < n; ++i) dd[i] = d[i] - sigma;
  for (int i = 0; i + 1 < n; ++i) { dl[i] = e[i]; du[i] = e[i]; }
  for (int i = 0; i + 1 < n; ++i) {
    if (std::fabs(dd[i]) >= std::fabs(dl[i])) {
      if (dd[i] == 0.0) dd[i] = tiny;
      const double f = dl[i] / dd[i];
      dl[i] = f;
      dd[i + 1] -= f * du[i];
    } else {  // swap rows i and i+1
      const double f = dd[i] / dl[i];
      dd[i] = dl[i];
      dl[i] = f;
      const double t = du[i];
      du[i] = dd[i + 1];
      dd[i + 1] = t - f * dd[i + 1];
      if (i + 2 < n) {
        du2[i] = du[i + 1];
        du[i + 1] = -f * du[i + 1];
      }
      swp[i] = 1;
    }
  }
  if (dd[n - 1] == 0.0) dd[n - 1] = tiny;
  auto solve_passes = [&](std::vector<double>& v) {
    for (int pass = 0; pass < 4; ++pass) {
      std::vector<double> b = v;
      for (int i = 0; i + 1 < n; ++i) {
        if (swp[i]) std::swap(b[i], b[i + 1]);
        b[i + 1] -= dl[i] * b[i];
      }
      v[n - 1] = b[n - 1] / dd[n - 1];
      v[n - 2] = (b[n - 2] - du[n - 2] * v[n - 1]) / dd[n - 2];
      for (int i = n - 3; i >= 0; --i)
        v[i] = (b[i] - du[i] * v[i + 1] - du2[i] * v[i + 2]) / dd[i];
      double s = 0.0;
      for (double t : v) s += t * t;
      s = std::sqrt(s);
      if (v[0] < 0) s = -s;
      for (auto& t : v) t /= s;
    }
  };
  solve_passes(x);
  double res = 0.0;
  for (int i = 0; i < n; ++i) {
    double t = (d[i] - lam) * x[i];
    if (i > 0) t += e[i - 1] * x[i - 1];
    if (i + 1 < n) t += e[i] * x[i + 1];
    res = std::max(res, std::fabs(t));
  }
  if (!(res <= 1e-8 * std::max(tnorm, tiny))) {  // e_1 missed a localised vector
    x.assign(n, 1.0 / std::sqrt(static_cast<double>(n)));
    solve_passes(x);
  }
  return x;
}

// tridiagonalise_one (lapack_adapter.hpp:94-104): lowest eigenpair of the
// tridiagonal matrix with diagonal D and off-diagonal E (first n-1 used).
// The reference uses a dense Eigen SelfAdjointEigenSolver; the eigenvalue is
// found here by bisection and the vector by inverse iteration.  The
// eigenvector sign is arbitrary in both.
void tridiag_one(const std::vector<double>& D, const std::vector<double>& E,
                 double& e0, std::vector<double>& vec) {
  e0 = tridiag_lowest(D, E);
  vec = tridiag_inverse_iteration(D, E, e0);
}

struct LzSettings {  // lanczos::Settings, components/linalg_routines/include/lanczos.hpp:9-20
  int64_t krylov_dim;
  double abs_tol;
  double rel_tol;
  int64_t min_iterations;
  int64_t max_iterations;
  uint32_t x0_seed;
  int calc_eigenvector;
  int verify_eigenvector;
};

struct LzResult {  // lanczos::Result, lanczos.hpp:22-28
  int eigval_converged;
  int eigvec_converged;
  int64_t n_iterations;
  double eigval_error;
  double eigvec_error;
};

// set_random_unit (lanczos.cpp:55-62)
void random_unit(std::vector<double>& v, uint32_t seed) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (auto& x : v) x = dist(rng);
  double s = 0.0;
  for (double x : v) s += x * x;
  const double nrm = std::sqrt(s);
  for (auto& x : v) x /= nrm;
}

double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}

template <typename Apply>
LzResult lanczos_iterate(Apply&& apply_add, std::vector<double>& v,
                         std::vector<double>& alphas, std::vector<double>& betas,
                         const LzSettings& st, const std::vector<double>* ritz,
                         std::vector<double>* eigvec, bool check_each_step) {
  // lanczos.cpp:67-177
  const size_t dim = v.size();
  random_unit(v, st.x0_seed);
  std::vector<double> u(dim, 0.0);
  alphas.clear();
  betas.clear();
  LzResult r{};
  double beta = 0.0;
  double eigval = std::numeric_limits<double>::max();
  int64_t it = 0;
  for (it = 0; it < st.max_iterations; ++it) {
    if (eigvec && ritz && it < static_cast<int64_t>(ritz->size())) {
      const double c = (*ritz)[it];
      for (size_t i = 0; i < dim; ++i) (*eigvec)[i] += c * v[i];
    }
    if (it > 0)
      for (auto& x : u) x *= -beta;
    apply_add(v.data(), u.data());
    const double alpha = dot(v, u);
    for (size_t i = 0; i < dim; ++i) u[i] += -alpha * v[i];
    beta = std::sqrt(dot(u, u));
    if (beta < st.abs_tol) break;
    std::swap(v, u);
    const double inv = 1 / beta;
    for (auto& x : v) x *= inv;
    alphas.push_back(alpha);
    betas.push_back(beta);
    if (it >= st.min_iterations && (check_each_step || ritz)) {
      // check_lanczos_convergence (lanczos.cpp:14-53)
      const double old = eigval;
      eigval = tridiag_lowest(alphas, betas);
      r.eigval_error = std::fabs(eigval - old);
      const double rel = r.eigval_error / std::max(std::fabs(eigval), 1e-10);
      r.eigval_converged = (r.eigval_error < st.abs_tol || rel < st.rel_tol);
      if (r.eigval_converged && ritz) {
        r.eigvec_converged = 1;
        break;
      }
    }
  }
  r.n_iterations = it;
  return r;
}

}  // namespace

// ============================================================================
// C ABI for ctypes
// ============================================================================
extern "C" {

int oracle_num_threads(void) { return omp_get_max_threads(); }
void oracle_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }

// masks out: X[2], Z[2], down[2], up[2]
int oracle_pmr_construct(const char* ops, const int* ids, int n, uint64_t* m,
                         int* sign) {
  Pmr p;
  if (!construct(ops, ids, n, p)) return -1;
  const u128 v[4] = {p.X, p.Z, p.down, p.up};
  for (int k = 0; k < 4; ++k) { m[2 * k] = lo64(v[k]); m[2 * k + 1] = hi64(v[k]); }
  *sign = p.sign;
  return 0;
}

// SymbolicPMROperator(const std::string&) -> parseSpec (operator.hpp:349-381):
// whitespace separated tokens "<digits><op>".
int oracle_pmr_from_spec(const char* spec, uint64_t* m, int* sign) {
  std::istringstream ss(spec);
  std::string tok;
  std::vector<char> ops;
  std::vector<int> ids;
  while (ss >> tok) {
    if (tok.size() < 2) return -1;
    size_t i = 0;
    int idx = 0;
    while (i < tok.size() - 1 && std::isdigit(static_cast<unsigned char>(tok[i]))) {
      idx = idx * 10 + (tok[i] - '0');
      ++i;
    }
    const char op = tok[i];
    if (std::string("xyzXYZ+-pqPQ").find(op) == std::string::npos) return -1;
    ops.push_back(op);
    ids.push_back(idx);
  }
  return oracle_pmr_construct(ops.data(), ids.data(), static_cast<int>(ops.size()), m, sign);
}

// state in/out [lo, hi]; returns the amplitude sign
int oracle_apply_state(const uint64_t* m, int sign, uint64_t* state) {
  Pmr p;
  p.X = mk(m[0], m[1]);
  p.Z = mk(m[2], m[3]);
  p.down = mk(m[4], m[5]);
  p.up = mk(m[6], m[7]);
  p.sign = sign;
  u128 s = mk(state[0], state[1]);
  const int r = apply_state(p, s);
  state[0] = lo64(s);
  state[1] = hi64(s);
  return r;
}

// operator*= (operator.hpp:312-330): a <- a * b (masks as above)
void oracle_pmr_mul(uint64_t* a, int* asign, const uint64_t* b, int bsign) {
  u128 X = mk(a[0], a[1]), Z = mk(a[2], a[3]), dn = mk(a[4], a[5]), up = mk(a[6], a[7]);
  const u128 bX = mk(b[0], b[1]), bZ = mk(b[2], b[3]), bdn = mk(b[4], b[5]), bup = mk(b[6], b[7]);
  // the product's sign starts from a's sign only: operator*= never reads
  // b.sign (operator.hpp:312-330)
  int s = *asign;
  s *= 1 - 2 * (popc(X & bZ) & 1);
  Z ^= bZ;
  X ^= bX;
  const u128 req_up = (dn & bX) | (up & ~bX);
  const u128 req_dn = (up & bX) | (dn & ~bX);
  up = bup | req_up;
  dn = bdn | req_dn;
  if ((up & dn) != 0) s = 0;
  const u128 v[4] = {X, Z, dn, up};
  for (int k = 0; k < 4; ++k) { a[2 * k] = lo64(v[k]); a[2 * k + 1] = hi64(v[k]); }
  *asign = s;
  (void)bsign;
}

void* oracle_basis_new(const uint64_t* lohi, int64_t D) {
  auto* b = new Basis();
  b->states.resize(static_cast<size_t>(D));
  for (int64_t i = 0; i < D; ++i) b->states[i] = mk(lohi[2 * i], lohi[2 * i + 1]);
  b->build_bounds();
  return b;
}

void oracle_basis_free(void* b) { delete static_cast<Basis*>(b); }

// n queries; rank[i] = index or -1 (missing) or -2 (high word without bucket)
void oracle_basis_search(const void* b, const uint64_t* q, int64_t n,
                         int64_t* rank, int bst) {
  const Basis& B = *static_cast<const Basis*>(b);
  for (int64_t i = 0; i < n; ++i) {
    int64_t J = -1;
    const u128 s = mk(q[2 * i], q[2 * i + 1]);
    const int f = bst ? B.search_bst(s, J) : B.search(s, J);
    rank[i] = f == 1 ? J : (f == 0 ? -1 : -2);
  }
}

// y += H x (LazyOpSum::evaluate_add).  Returns -1 if a connected state's high
// word had no bucket (where the reference would throw std::out_of_range).
int oracle_evaluate_add(const void* b, const Term* terms, int n_terms,
                        const double* x, double* y) {
  const Basis& B = *static_cast<const Basis*>(b);
  OpSum H(terms, n_terms);
  return evaluate_add(B, H, x, y);
}

// evaluate_add restricted to the pushes of rows [r0, r1).
int oracle_evaluate_add_rows(const void* b, const Term* terms, int n_terms,
                             const double* x, double* y, int64_t r0, int64_t r1) {
  const Basis& B = *static_cast<const Basis*>(b);
  OpSum H(terms, n_terms);
  const int64_t D = static_cast<int64_t>(B.states.size());
  return evaluate_add_rows(B, H, x, y, std::max<int64_t>(0, r0), std::min(r1, D));
}

// toSparseMatrix triplets (operator_matrix.hpp:79-105), in the reference's
// order: for each column j, the off-diagonal terms (J, j, c*sign) with
// |c*sign| > tol, then the diagonal (j, j, acc).  Two-call protocol: with
// rows == NULL returns the count only.
int64_t oracle_sparse_triplets(const void* b, const Term* terms, int n_terms,
                               double tol, int64_t cap, int64_t* rows,
                               int64_t* cols, double* vals) {
  const Basis& B = *static_cast<const Basis*>(b);
  OpSum H(terms, n_terms);
  const int64_t D = static_cast<int64_t>(B.states.size());
  int64_t k = 0;
  for (int64_t j = 0; j < D; ++j) {
    for (const auto& [c, op] : H.off) {
      u128 s = B.states[j];
      int sg = apply_state(op, s);
      int64_t J = j;
      if (sg != 0) {
        const int f = B.search(s, J);
        sg *= (f == 1 ? 1 : 0);
      }
      const double res = c * sg;
      if (std::fabs(res) > tol) {
        if (rows && k < cap) { rows[k] = J; cols[k] = j; vals[k] = res; }
        ++k;
      }
    }
    double acc = 0.0;
    for (const auto& [c, op] : H.diag) {
      u128 s = B.states[j];
      acc += c * apply_state(op, s);
    }
    if (rows && k < cap) { rows[k] = j; cols[k] = j; vals[k] = acc; }
    ++k;
  }
  return k;
}

// Raw draws of std::normal_distribution<double>(0, 1) over std::mt19937(seed)
// (the generator of test/test_lanczos.cpp:44-60 and lanczos.cpp:55-62).
void oracle_normal_stream(uint32_t seed, double* v, int64_t n) {
  std::mt19937 rng(seed);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (int64_t i = 0; i < n; ++i) v[i] = dist(rng);
}

void oracle_random_unit(uint32_t seed, double* v, int64_t n) {
  std::vector<double> w(static_cast<size_t>(n));
  random_unit(w, seed);
  std::memcpy(v, w.data(), sizeof(double) * static_cast<size_t>(n));
}

// lanczos::eigval0 (lanczos.cpp:184-227) on the oracle apply.  alphas/betas
// (optional, capacity cap) receive pass 1's recurrence coefficients.
int oracle_lanczos_eigval0(const void* b, const Term* terms, int n_terms,
                           const LzSettings* st, double* eigval, double* eigvec,
                           LzResult* res, double* alphas_out, double* betas_out,
                           int64_t cap, int64_t* n_ab) {
  const Basis& B = *static_cast<const Basis*>(b);
  OpSum H(terms, n_terms);
  const size_t D = B.states.size();
  if (D == 0) return -1;
  int bad = 0;
  auto apply_add = [&](const double* x, double* y) {
    if (evaluate_add(B, H, x, y) != 0) bad = 1;
  };
  std::vector<double> v(D), alphas, betas;
  LzResult r = lanczos_iterate(apply_add, v, alphas, betas, *st, nullptr,
                               nullptr, false);
  if (alphas_out && betas_out) {
    for (size_t i = 0; i < alphas.size() && static_cast<int64_t>(i) < cap; ++i) {
      alphas_out[i] = alphas[i];
      betas_out[i] = betas[i];
    }
  }
  if (n_ab) *n_ab = static_cast<int64_t>(alphas.size());
  std::vector<double> ritz;
  double e0 = 0.0;
  tridiag_one(alphas, betas, e0, ritz);
  *eigval = e0;
  if (st->calc_eigenvector) {
    std::vector<double> ev(D, 0.0);
    r = lanczos_iterate(apply_add, v, alphas, betas, *st, &ritz, &ev, true);
    std::memcpy(eigvec, ev.data(), sizeof(double) * D);
  }
  if (res) *res = r;
  return bad ? -2 : 0;
}

// M steps of the 3-term recurrence of lanczos_iterate (lanczos.cpp:111-146,
// no convergence test) from a given start vector (normalised here); the
// FTLM sample coefficients.  A beta below 1e-300 ends the recurrence (the
// remaining coefficients are left 0).
int oracle_lanczos_coeffs(const void* b, const Term* terms, int n_terms,
                          const double* v0, int M, double* alphas,
                          double* betas) {
  const Basis& B = *static_cast<const Basis*>(b);
  OpSum H(terms, n_terms);
  const size_t D = B.states.size();
  std::vector<double> v(v0, v0 + D), u(D, 0.0);
  const double n0 = std::sqrt(dot(v, v));
  for (auto& x : v) x /= n0;
  double beta = 0.0;
  for (int j = 0; j < M; ++j) {
    alphas[j] = betas[j] = 0.0;
  }
  for (int j = 0; j < M; ++j) {
    if (j > 0)
      for (auto& x : u) x *= -beta;
    if (evaluate_add(B, H, v.data(), u.data()) != 0) return -1;
    const double a = dot(v, u);
    for (size_t i = 0; i < D; ++i) u[i] -= a * v[i];
    beta = std::sqrt(dot(u, u));
    alphas[j] = a;
    betas[j] = beta;
    if (!(beta > 1e-300)) break;
    std::swap(v, u);
    for (auto& x : v) x /= beta;
  }
  return 0;
}

// Same recurrence driven by an arbitrary dense symmetric matrix (the
// MockHam scenario of test/test_lanczos.cpp:12-40,113-174).
int oracle_lanczos_dense(const double* M, int64_t n, const LzSettings* st,
                         double* eigval, double* eigvec, LzResult* res) {
  auto apply_add = [&](const double* x, double* y) {
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t j = 0; j < n; ++j) s += M[i * n + j] * x[j];
      y[i] += s;
    }
  };
  std::vector<double> v(static_cast<size_t>(n)), alphas, betas;
  LzResult r = lanczos_iterate(apply_add, v, alphas, betas, *st, nullptr,
                               nullptr, false);
  std::vector<double> ritz;
  double e0 = 0.0;
  tridiag_one(alphas, betas, e0, ritz);
  *eigval = e0;
  if (st->calc_eigenvector) {
    std::vector<double> ev(static_cast<size_t>(n), 0.0);
    r = lanczos_iterate(apply_add, v, alphas, betas, *st, &ritz, &ev, true);
    std::memcpy(eigvec, ev.data(), sizeof(double) * static_cast<size_t>(n));
  }
  if (res) *res = r;
  return 0;
}

// Symmetric eigen-decomposition by cyclic Jacobi (n x n column-major),
// exposed so tests can pin it against numpy.
void oracle_jacobi_eig(const double* A, int n, double* w, double* V) {
  std::vector<double> a(A, A + static_cast<size_t>(n) * n), ww, VV;
  jacobi_eig(a, n, ww, VV);
  std::memcpy(w, ww.data(), sizeof(double) * n);
  std::memcpy(V, VV.data(), sizeof(double) * static_cast<size_t>(n) * n);
}

double oracle_tridiag_lowest(const double* d, const double* e, int64_t n) {
  std::vector<double> D(d, d + n), E(e, e + std::max<int64_t>(n - 1, 0));
  return tridiag_lowest(D, E);
}

void oracle_tridiag_one(const double* d, const double* e, int64_t n,
                        double* e0, double* vec) {
  std::vector<double> D(d, d + n), E(e, e + std::max<int64_t>(n - 1, 0)), v;
  tridiag_one(D, E, *e0, v);
  std::memcpy(vec, v.data(), sizeof(double) * static_cast<size_t>(n));
}

}  // extern "C"
