// Krylov drivers (host control, GPU vectors) and the host basis enumerator.
//
//   qsi_lanczos_eigval0  lanczos::eigval0 / lanczos_iterate
//                        (components/linalg_routines/src/lanczos.cpp:14-235)
//   qsi_eigsh            thick-restart Lanczos standing in for
//                        Spectra::SymEigsSolver behind `mfsparse`
//                        (include/matrix_diag_bits.hpp:51-95)
//   qsi_enumerate_basis  gen_spinon_basis DFS
//                        (components/build_basis/src/pyro_tree.cpp:20-240)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <random>
#include <thread>
#include <vector>

#include "qsi_enum.hpp"
#include "qsi_hostla.hpp"
#include "qsi_internal.hpp"

namespace qsi {
namespace {

// RAII device buffer
struct DevBuf {
  void* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  hipError_t alloc(size_t bytes) { return hipMalloc(&p, std::max<size_t>(bytes, 8)); }
  double* d() const { return static_cast<double*>(p); }
};

#define QSI_TRY(expr)            \
  do {                           \
    const int _rc = (expr);      \
    if (_rc != QSI_OK) return _rc; \
  } while (0)

// The reference's start vector: std::mt19937(seed), N(0,1) entries, then
// divided by the 2-norm (lanczos.cpp:55-62, 80-81).
std::vector<double> random_unit(int64_t n, uint32_t seed) {
  std::vector<double> v(static_cast<size_t>(n));
  std::mt19937 rng(seed);
  std::normal_distribution<double> dist(0.0, 1.0);
  for (auto& x : v) x = dist(rng);
  double s = 0.0;
  for (double x : v) s += x * x;
  const double nrm = std::sqrt(s);
  for (auto& x : v) x /= nrm;
  return v;
}

// Step `it` of the 3-term recurrence on the device (lanczos.cpp:111-146).
// The vectors are kept unnormalised: the buffer v holds v_it / s_it and u
// holds v_{it-1} / s_{it-1}, where s_j = 1/beta_{j-1} sits on the device at
// ab[3(j-1) + 2] (s_0 = 1: the start vector is normalised in place).  The
// apply and the axpy read the buffers times their scales, which gives the
// values a separate normalising pass would have stored, so the coefficients
// have the same bits as with that pass, one pass fewer per step.
// On exit u holds beta_it v_{it+1}, unnormalised (its scale 1/beta_it is at
// ab[3 it + 2]); alpha_it -> ab[3 it], beta_it -> ab[3 it + 1]; the caller
// swaps v and u.
inline const double* lanczos_scale(const double* ab, int64_t j) {
  return j == 0 ? nullptr : ab + 3 * (j - 1) + 2;
}

int lanczos_step(const qsi_op_s* op, const double* v, double* u, int64_t it,
                 double* ab_base, double* partials, hipStream_t s) {
  const int64_t D = op->basis->dim;
  double* ab = ab_base + 3 * it;
  const double* vs = lanczos_scale(ab_base, it);
  // u = -beta_{it-1} v_{it-1} + H v_it;  alpha = <v_it, u>   (lanczos.cpp:117-126)
  QSI_TRY(launch_apply_scaled(op, v, u, it == 0 ? nullptr : u, -1.0,
                              it == 0 ? nullptr : ab - 2, vs,
                              it == 0 ? nullptr : lanczos_scale(ab_base, it - 1), partials, ab,
                              s));
  // u -= alpha v_it;  beta = ||u||, 1/beta                   (lanczos.cpp:129-139)
  QSI_TRY(launch_axpy_norm(ab, v, u, D, partials, ab + 1, s, vs));
  return QSI_OK;
}

struct ConvState {
  double eigval = std::numeric_limits<double>::max();
  double error = 0.0;
  bool converged = false;
};

// check_lanczos_convergence (lanczos.cpp:14-53) on the leading n entries.
void check_convergence(const std::vector<double>& a, const std::vector<double>& b,
                       size_t n, const qsi_lanczos_settings_t& st, ConvState& c) {
  const double old = c.eigval;
  c.eigval = hostla::tridiag_lowest(a, b, n);
  c.error = std::fabs(c.eigval - old);
  const double rel = c.error / std::max(std::fabs(c.eigval), 1e-10);
  c.converged = (c.error < st.abs_tol || rel < st.rel_tol);
}

}  // namespace

int64_t lanczos_coeffs_partials(const qsi_op_s* op) {
  return std::max<int64_t>(max_vec_partials(), apply_num_blocks(op));
}

int lanczos_coeffs_run(const qsi_op_s* op, double* v, double* u, double* partials, double* ab,
                       double* scal, int M, double* h_alpha, double* h_beta, int* n_steps,
                       hipStream_t s) {
  const int64_t D = op->basis->dim;
  double n2 = 0.0;
  QSI_TRY(launch_dot(v, v, D, partials, scal, s));
  QSI_CHECK_HIP(hipMemcpy(&n2, scal, sizeof(double), hipMemcpyDeviceToHost));
  QSI_REQUIRE(n2 > 0.0, "qsi_lanczos_coeffs: zero start vector");
  QSI_TRY(launch_scal(1.0 / std::sqrt(n2), v, D, s));
  std::fill(h_alpha, h_alpha + M, 0.0);
  std::fill(h_beta, h_beta + M, 0.0);
  constexpr int kChunk = 64;  // coefficients copied back every kChunk steps
  double hab[3 * kChunk];
  int done = 0;
  for (int it0 = 0; it0 < M; it0 += kChunk) {
    const int it1 = std::min(M, it0 + kChunk);
    for (int it = it0; it < it1; ++it) {
      QSI_TRY(lanczos_step(op, v, u, it, ab, partials, s));
      std::swap(v, u);
    }
    QSI_CHECK_HIP(hipMemcpy(hab, ab + 3 * it0, 3 * (it1 - it0) * sizeof(double),
                            hipMemcpyDeviceToHost));
    bool stop = false;
    for (int it = it0; it < it1; ++it) {
      h_alpha[it] = hab[3 * (it - it0)];
      h_beta[it] = hab[3 * (it - it0) + 1];
      done = it + 1;
      if (!(h_beta[it] > 1e-300)) {  // invariant subspace: the recurrence ends
        h_beta[it] = 0.0;
        stop = true;
        break;
      }
    }
    if (stop) break;
  }
  if (n_steps) *n_steps = done;
  return QSI_OK;
}

}  // namespace qsi

using namespace qsi;

extern "C" {

int qsi_lanczos_eigval0(qsi_op_t op, const qsi_lanczos_settings_t* st,
                        double* eigval, double* d_eigvec,
                        qsi_lanczos_result_t* result) {
  QSI_REQUIRE(op && st && eigval, "qsi_lanczos_eigval0: null argument");
  const int64_t D = op->basis->dim;
  QSI_REQUIRE(D > 0, "qsi_lanczos_eigval0: v called with size 0");
  QSI_REQUIRE(!st->calc_eigenvector || d_eigvec,
              "qsi_lanczos_eigval0: calc_eigenvector needs d_eigvec");
  QSI_REQUIRE(st->max_iterations >= 0 && st->min_iterations >= 0,
              "qsi_lanczos_eigval0: negative iteration bounds");
  const int64_t max_it = st->max_iterations;
  const size_t bytes = static_cast<size_t>(D) * sizeof(double);
  DevBuf bv, bu, bpart, bab;
  QSI_CHECK_HIP(bv.alloc(bytes));
  QSI_CHECK_HIP(bu.alloc(bytes));
  // partials of the fused apply+dot (one per apply workgroup) and of the
  // streaming reductions
  const int64_t n_part = std::max<int64_t>(max_vec_partials(), apply_num_blocks(op));
  QSI_CHECK_HIP(bpart.alloc(static_cast<size_t>(n_part) * sizeof(double)));
  const int64_t chunk = 64;
  // per step: alpha, beta, 1/beta
  QSI_CHECK_HIP(bab.alloc(static_cast<size_t>(3 * std::max<int64_t>(max_it, 1)) * sizeof(double)));
  hipStream_t s = nullptr;
  const std::vector<double> v0 = random_unit(D, st->x0_seed);

  // ---- pass 1: the recurrence; stops at max_iterations or beta < abs_tol
  // (lanczos.cpp:111-160; pass 1 never stops on convergence)
  std::vector<double> alphas, betas;
  std::vector<double> hab;
  int64_t n_loop = max_it;  // value of the loop counter at exit
  {
    double* v = bv.d();
    double* u = bu.d();
    QSI_CHECK_HIP(hipMemcpy(v, v0.data(), bytes, hipMemcpyHostToDevice));
    bool stop = false;
    for (int64_t it0 = 0; it0 < max_it && !stop; it0 += chunk) {
      const int64_t it1 = std::min(max_it, it0 + chunk);
      for (int64_t it = it0; it < it1; ++it) {
        QSI_TRY(lanczos_step(op, v, u, it, bab.d(), bpart.d(), s));
        std::swap(v, u);
      }
      hab.resize(static_cast<size_t>(3 * (it1 - it0)));
      QSI_CHECK_HIP(hipMemcpyAsync(hab.data(), bab.d() + 3 * it0,
                                   hab.size() * sizeof(double),
                                   hipMemcpyDeviceToHost, s));
      QSI_CHECK_HIP(hipStreamSynchronize(s));
      for (int64_t it = it0; it < it1; ++it) {
        const double a = hab[3 * (it - it0)], b = hab[3 * (it - it0) + 1];
        if (!(b >= st->abs_tol)) {  // beta < abs_tol (or NaN): break
          n_loop = it;
          stop = true;
          break;
        }
        alphas.push_back(a);
        betas.push_back(b);
      }
    }
  }
  const size_t n1 = alphas.size();
  qsi_lanczos_result_t res{};
  // tridiagonalise_one after pass 1 (lanczos.cpp:209-211)
  std::vector<double> ritz;
  if (n1 > 0) {
    *eigval = hostla::tridiag_lowest(alphas, betas, n1);
    ritz = hostla::tridiag_eigvec(alphas, betas, n1, *eigval);
  } else {
    // the Krylov space is one-dimensional: v0 is an eigenvector with the
    // Rayleigh quotient alpha_0 (the reference reads an empty vector here)
    *eigval = hab.empty() ? 0.0 : hab[0];
  }

  if (!st->calc_eigenvector) {
    // pass-1 Result: the last convergence check (lanczos.cpp:149-159)
    res.n_iterations = n_loop;
    const int64_t mi = st->min_iterations;
    if (n1 > 0 && static_cast<int64_t>(n1) - 1 >= mi) {
      ConvState c;
      if (static_cast<int64_t>(n1) - 2 >= mi) check_convergence(alphas, betas, n1 - 1, *st, c);
      check_convergence(alphas, betas, n1, *st, c);
      res.eigval_converged = c.converged;
      res.eigval_error = c.error;
    }
    if (result) *result = res;
    return QSI_OK;
  }

  // ---- pass 2: same start vector, accumulate eigvec += ritz[j] v_j until
  // the per-step check converges (lanczos.cpp:112-115,155-158).  The device
  // recurrence is deterministic, so pass 2 reproduces pass 1's coefficients
  // and its stopping iteration is found on the host first.
  ConvState c;
  int64_t stop_it = -1;  // iteration at which pass 2 breaks on convergence
  for (size_t k = 0; k < n1; ++k) {
    if (static_cast<int64_t>(k) >= st->min_iterations) {
      check_convergence(alphas, betas, k + 1, *st, c);
      if (c.converged) {
        stop_it = static_cast<int64_t>(k);
        break;
      }
    }
  }
  // last iteration whose top-of-loop accumulation runs
  const int64_t last_acc = stop_it >= 0 ? stop_it : std::min<int64_t>(n_loop, max_it - 1);
  {
    double* v = bv.d();
    double* u = bu.d();
    QSI_CHECK_HIP(hipMemcpy(v, v0.data(), bytes, hipMemcpyHostToDevice));
    QSI_CHECK_HIP(hipMemsetAsync(d_eigvec, 0, bytes, s));
    for (int64_t it = 0; it <= last_acc; ++it) {
      if (it < static_cast<int64_t>(ritz.size()))
        QSI_TRY(launch_axpy(ritz[it], v, d_eigvec, D, s, lanczos_scale(bab.d(), it)));
      if (it == last_acc) break;
      QSI_TRY(lanczos_step(op, v, u, it, bab.d(), bpart.d(), s));
      std::swap(v, u);
    }
    QSI_CHECK_HIP(hipStreamSynchronize(s));
  }
  res.eigval_converged = c.converged;
  res.eigvec_converged = stop_it >= 0;
  res.n_iterations = stop_it >= 0 ? stop_it : n_loop;
  res.eigval_error = c.error;
  if (st->verify_eigenvector) {
    // u = H eigvec - eigval * eigvec; error = ||u|| (lanczos.cpp:163-172)
    double* u = bu.d();
    QSI_TRY(launch_apply(op, d_eigvec, u, nullptr, 0.0, nullptr, nullptr, s));
    QSI_TRY(launch_axpy(-c.eigval, d_eigvec, u, D, s));
    DevBuf out;
    QSI_CHECK_HIP(out.alloc(2 * sizeof(double)));
    QSI_TRY(launch_dot(u, u, D, bpart.d(), out.d(), s));
    double nn = 0.0;
    QSI_CHECK_HIP(hipMemcpy(&nn, out.d(), sizeof(double), hipMemcpyDeviceToHost));
    res.eigvec_error = std::sqrt(nn);
  }
  if (result) *result = res;
  return QSI_OK;
}

int qsi_lanczos_coeffs(qsi_op_t op, const double* d_v0, uint64_t seed, int M,
                       double* h_alpha, double* h_beta, int* n_steps) {
  QSI_REQUIRE(op && h_alpha && h_beta, "qsi_lanczos_coeffs: null argument");
  QSI_REQUIRE(M >= 1, "qsi_lanczos_coeffs: M must be >= 1");
  const int64_t D = op->basis->dim;
  QSI_REQUIRE(D > 0, "qsi_lanczos_coeffs: empty basis");
  const size_t bytes = static_cast<size_t>(D) * sizeof(double);
  DevBuf bv, bu, bpart, bab, bscal;
  QSI_CHECK_HIP(bv.alloc(bytes));
  QSI_CHECK_HIP(bu.alloc(bytes));
  QSI_CHECK_HIP(bpart.alloc(static_cast<size_t>(lanczos_coeffs_partials(op)) * sizeof(double)));
  QSI_CHECK_HIP(bab.alloc(static_cast<size_t>(3 * M) * sizeof(double)));
  QSI_CHECK_HIP(bscal.alloc(2 * sizeof(double)));
  hipStream_t s = nullptr;
  if (d_v0) {
    QSI_CHECK_HIP(hipMemcpyAsync(bv.d(), d_v0, bytes, hipMemcpyDeviceToDevice, s));
  } else {
    QSI_TRY(launch_randn(bv.d(), D, seed, s));
  }
  return lanczos_coeffs_run(op, bv.d(), bu.d(), bpart.d(), bab.d(), bscal.d(), M, h_alpha,
                            h_beta, n_steps, s);
}

// One FTLM sample with operator expectations.  Pass 1 runs the 3-term
// recurrence from r; T_n = Y diag(e) Y^T gives, for each beta, the Lanczos
// coefficients of psi_beta = exp(-beta (H - e_shift) / 2) r,
//   c_beta = Y (exp(-beta (e - e_shift) / 2) * Y[0, :])^T;
// pass 2 runs the same recurrence again (same kernels, so the same vectors)
// and adds c_beta[j] v_j into Psi[:, beta]; then <psi|psi> and <psi|O|psi>
// per beta and operator come from device dots and one apply of O per
// column.
int qsi_ftlm_sample(qsi_op_t H, const qsi_op_t* ops, int n_ops, const double* d_r, int M,
                    const double* h_betas, int n_beta, double e_shift, double* e_used,
                    double* h_norm, double* h_expect, double* h_alpha, double* h_beta,
                    int* n_steps) {
  QSI_REQUIRE(H && d_r && h_betas && h_norm && h_alpha && h_beta,
              "qsi_ftlm_sample: null argument");
  QSI_REQUIRE(n_ops >= 0 && (n_ops == 0 || (ops && h_expect)),
              "qsi_ftlm_sample: operators need ops and h_expect");
  QSI_REQUIRE(M >= 1 && n_beta >= 1, "qsi_ftlm_sample: need M >= 1 and n_beta >= 1");
  const int64_t D = H->basis->dim;
  QSI_REQUIRE(D > 0, "qsi_ftlm_sample: empty basis");
  for (int l = 0; l < n_ops; ++l)
    QSI_REQUIRE(ops[l] && ops[l]->basis->dim == D,
                "qsi_ftlm_sample: every operator must act on the Hamiltonian's basis size");
  const size_t bytes = static_cast<size_t>(D) * sizeof(double);
  DevBuf bv, bu, bt, bpsi, bpart, bab, bscal, bc;
  QSI_CHECK_HIP(bv.alloc(bytes));
  QSI_CHECK_HIP(bu.alloc(bytes));
  QSI_CHECK_HIP(bpart.alloc(static_cast<size_t>(lanczos_coeffs_partials(H)) * sizeof(double)));
  QSI_CHECK_HIP(bab.alloc(static_cast<size_t>(3 * M) * sizeof(double)));
  QSI_CHECK_HIP(bscal.alloc(2 * sizeof(double)));
  hipStream_t s = nullptr;
  // pass 1: the coefficients
  QSI_CHECK_HIP(hipMemcpyAsync(bv.d(), d_r, bytes, hipMemcpyDeviceToDevice, s));
  int n = 0;
  QSI_TRY(lanczos_coeffs_run(H, bv.d(), bu.d(), bpart.d(), bab.d(), bscal.d(), M, h_alpha,
                             h_beta, &n, s));
  if (n_steps) *n_steps = n;
  // the Lanczos coefficients of psi_beta
  std::vector<double> d(h_alpha, h_alpha + n), e(h_beta, h_beta + std::max(n - 1, 0)), ev, Y;
  if (!hostla::tridiag_eig(d, e, ev, &Y)) {
    set_error("qsi_ftlm_sample: no convergence of the tridiagonal eigensolver");
    return QSI_E_NOCONV;
  }
  if (!std::isfinite(e_shift)) e_shift = ev[0];
  if (e_used) *e_used = e_shift;
  std::vector<double> C(static_cast<size_t>(n) * n_beta, 0.0);  // C[j * n_beta + b]
  for (int b = 0; b < n_beta; ++b)
    for (int k = 0; k < n; ++k) {
      const double f = std::exp(-0.5 * h_betas[b] * (ev[k] - e_shift)) *
                       Y[static_cast<size_t>(k) * n];
      for (int j = 0; j < n; ++j)
        C[static_cast<size_t>(j) * n_beta + b] += Y[static_cast<size_t>(k) * n + j] * f;
    }
  // pass 2: Psi = V C, one column per beta
  const size_t ld = static_cast<size_t>(D);
  QSI_CHECK_HIP(bpsi.alloc(ld * n_beta * sizeof(double)));
  QSI_CHECK_HIP(bc.alloc(C.size() * sizeof(double)));
  QSI_CHECK_HIP(hipMemcpy(bc.d(), C.data(), C.size() * sizeof(double), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemsetAsync(bpsi.d(), 0, ld * n_beta * sizeof(double), s));
  QSI_CHECK_HIP(hipMemcpyAsync(bv.d(), d_r, bytes, hipMemcpyDeviceToDevice, s));
  double n2 = 0.0;
  QSI_TRY(launch_dot(bv.d(), bv.d(), D, bpart.d(), bscal.d(), s));
  QSI_CHECK_HIP(hipMemcpy(&n2, bscal.d(), sizeof(double), hipMemcpyDeviceToHost));
  QSI_TRY(launch_scal(1.0 / std::sqrt(n2), bv.d(), D, s));
  double* v = bv.d();
  double* u = bu.d();
  for (int it = 0; it < n; ++it) {
    QSI_TRY(launch_accum_cols(v, bpsi.d(), static_cast<int64_t>(ld), n_beta,
                              bc.d() + static_cast<size_t>(it) * n_beta, D, s,
                              lanczos_scale(bab.d(), it)));
    if (it + 1 < n) {
      QSI_TRY(lanczos_step(H, v, u, it, bab.d(), bpart.d(), s));
      std::swap(v, u);
    }
  }
  // <psi|psi> and <psi|O|psi>
  QSI_CHECK_HIP(bt.alloc(bytes));
  for (int b = 0; b < n_beta; ++b) {
    const double* psi = bpsi.d() + b * ld;
    QSI_TRY(launch_dot(psi, psi, D, bpart.d(), bscal.d(), s));
    QSI_CHECK_HIP(hipMemcpy(h_norm + b, bscal.d(), sizeof(double), hipMemcpyDeviceToHost));
    for (int l = 0; l < n_ops; ++l) {
      QSI_TRY(launch_apply(ops[l], psi, bt.d(), nullptr, 0.0, nullptr, nullptr, s));
      QSI_TRY(launch_dot(psi, bt.d(), D, bpart.d(), bscal.d(), s));
      QSI_CHECK_HIP(hipMemcpy(h_expect + static_cast<size_t>(l) * n_beta + b, bscal.d(),
                              sizeof(double), hipMemcpyDeviceToHost));
    }
  }
  return QSI_OK;
}

// Cullum-Willoughby selection of the Ritz values of an M-step Lanczos run
// without reorthogonalisation (J. Cullum, R. Willoughby, "Lanczos Algorithms
// for Large Symmetric Eigenvalue Computations", vol. 1, ch. 4).  Lost
// orthogonality makes T_M repeat converged eigenvalues (numerically multiple
// copies: kept, reported once) and adds spurious ones; a spurious eigenvalue
// is simple in T_M and also an eigenvalue of T_M with its first row and
// column removed, because its eigenvector has a vanishing first component.
// The residual of a Ritz pair is |beta_M z_M| (z_M the last component of the
// unit eigenvector of T_M); a cluster reports the smallest of its members'.
int qsi_lanczos_ritz(const double* h_alpha, const double* h_beta, int m, double tol,
                     int max_out, double* vals, double* resid, int* mult, int* n_out) {
  QSI_REQUIRE(h_alpha && h_beta && m >= 1 && max_out >= 0 && n_out,
              "qsi_lanczos_ritz: bad arguments");
  QSI_REQUIRE(max_out == 0 || (vals && resid && mult), "qsi_lanczos_ritz: null output");
  // a beta that underflowed ends the recurrence: T is then exact on an
  // invariant subspace (no copies, no spurious values, zero residuals)
  int n = m;
  for (int j = 0; j < m; ++j) {
    if (!(h_beta[j] > 0.0)) {
      n = j + 1;
      break;
    }
  }
  const double beta_m = n == m ? h_beta[m - 1] : 0.0;
  std::vector<double> d(h_alpha, h_alpha + n), e(h_beta, h_beta + std::max(n - 1, 0));
  double tnorm = 0.0;
  for (int j = 0; j < n; ++j) {
    const double r = std::fabs(d[j]) + (j > 0 ? std::fabs(e[j - 1]) : 0.0) +
                     (j + 1 < n ? std::fabs(e[j]) : 0.0);
    tnorm = std::max(tnorm, r);
  }
  if (!(tol > 0.0)) tol = 10.0 * n * std::numeric_limits<double>::epsilon() * std::max(tnorm, 1e-300);
  std::vector<double> theta, Z;
  if (!hostla::tridiag_eig(d, e, theta, &Z)) {
    set_error("qsi_lanczos_ritz: no convergence of the tridiagonal eigensolver");
    return QSI_E_NOCONV;
  }
  std::vector<double> that;  // eigenvalues of T_M without its first row and column
  if (n > 1 && n == m) {
    std::vector<double> d2(d.begin() + 1, d.end()), e2(e.begin() + 1, e.end()), w2;
    if (!hostla::tridiag_eig(d2, e2, that, nullptr)) {
      set_error("qsi_lanczos_ritz: no convergence of the tridiagonal eigensolver");
      return QSI_E_NOCONV;
    }
  }
  int k = 0;
  for (int a = 0; a < n;) {
    int b = a + 1;
    while (b < n && theta[b] - theta[b - 1] <= tol) ++b;  // cluster [a, b)
    double best_r = std::numeric_limits<double>::infinity(), best_v = theta[a];
    for (int i = a; i < b; ++i) {
      const double r = std::fabs(beta_m * Z[static_cast<size_t>(i) * n + (n - 1)]);
      if (r < best_r) {
        best_r = r;
        best_v = theta[i];
      }
    }
    bool spurious = false;
    if (b - a == 1 && !that.empty()) {
      auto it = std::lower_bound(that.begin(), that.end(), theta[a] - tol);
      spurious = it != that.end() && *it <= theta[a] + tol;
    }
    if (!spurious) {
      if (k < max_out) {
        vals[k] = best_v;
        resid[k] = best_r;
        mult[k] = b - a;
      }
      ++k;
    }
    a = b;
  }
  *n_out = std::min(k, max_out);
  return QSI_OK;
}

int qsi_eigsh(qsi_op_t op, int nev, int ncv, double tol, int max_restarts,
              const double* d_v0, uint64_t seed, double* h_evals,
              double* d_evecs, int n_evecs, int* nconv_out, int* n_applies) {
  QSI_REQUIRE(op && h_evals, "qsi_eigsh: null argument");
  const int64_t D = op->basis->dim;
  QSI_REQUIRE(D > 0, "qsi_eigsh: empty basis");
  QSI_REQUIRE(nev >= 1 && ncv >= 2, "qsi_eigsh: need nev >= 1, ncv >= 2");
  QSI_REQUIRE(ncv <= 63, "qsi_eigsh: ncv must be <= 63");
  QSI_REQUIRE(n_evecs >= 0 && (n_evecs == 0 || d_evecs), "qsi_eigsh: bad eigenvector output");
  // Spectra's sizing (matrix_diag_bits.hpp:64-66): ncv <= n, nev <= ncv - 1
  ncv = static_cast<int>(std::min<int64_t>(ncv, D));
  nev = std::min(nev, std::max(ncv - 1, 1));
  n_evecs = std::min(n_evecs, nev);
  const int64_t ld = (D + 63) / 64 * 64;
  hipStream_t s = nullptr;
  DevBuf bV, bW, bpart, bscal;
  QSI_CHECK_HIP(bV.alloc(static_cast<size_t>(ld) * (ncv + 1) * sizeof(double)));
  QSI_CHECK_HIP(bW.alloc(static_cast<size_t>(ld) * (ncv + 1) * sizeof(double)));
  QSI_CHECK_HIP(bpart.alloc(std::max(max_vec_partials(), multidot_partials()) * sizeof(double)));
  QSI_CHECK_HIP(bscal.alloc(4 * sizeof(double)));
  DevBuf bY;  // Ritz coefficients of a restart (at most ncv x ncv)
  QSI_CHECK_HIP(bY.alloc(static_cast<size_t>(ncv) * ncv * sizeof(double)));
  // projected columns (ncv + 1 per step: h and, from pass 1, |w|^2), per-step
  // beta and 1/beta, the fused pass's sums and the third-pass flag, all
  // device-resident so a restart cycle runs without host synchronisation
  const int colw = ncv + 1;
  DevBuf bT, bbeta, bg, bflag;
  QSI_CHECK_HIP(bT.alloc(static_cast<size_t>(ncv) * colw * sizeof(double)));
  QSI_CHECK_HIP(bbeta.alloc(static_cast<size_t>(2 * ncv) * sizeof(double)));
  QSI_CHECK_HIP(bg.alloc(80 * sizeof(double)));
  QSI_CHECK_HIP(bflag.alloc(sizeof(int)));
  double* V = bV.d();
  double* W = bW.d();
  int applies = 0;

  // start vector
  if (d_v0) {
    QSI_CHECK_HIP(hipMemcpyAsync(V, d_v0, static_cast<size_t>(D) * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
  } else {
    // Spectra's solver takes no seeded start vector (matrix_diag_bits.hpp:75),
    // so none needs reproducing: a counter-based Gaussian made on the device
    QSI_TRY(launch_randn(V, D, seed, s));
  }
  double nrm2 = 0.0;
  QSI_TRY(launch_dot(V, V, D, bpart.d(), bscal.d(), s));
  QSI_CHECK_HIP(hipMemcpy(&nrm2, bscal.d(), sizeof(double), hipMemcpyDeviceToHost));
  QSI_REQUIRE(nrm2 > 0.0, "qsi_eigsh: zero start vector");
  QSI_TRY(launch_scal(1.0 / std::sqrt(nrm2), V, D, s));

  std::vector<double> T(static_cast<size_t>(ncv) * ncv, 0.0);
  auto Tat = [&](int i, int j) -> double& { return T[static_cast<size_t>(j) * ncv + i]; };
  std::vector<double> theta, Y;
  int k = 0;       // kept (locked + thick) vectors
  int nconv = 0;
  int m = ncv;     // current basis size
  double beta_last = 0.0;
  bool exhausted = false;
  std::vector<double> hT(static_cast<size_t>(ncv) * colw), hbeta(2 * ncv);
  for (int restart = 0;; ++restart) {
    m = ncv;
    // the cycle: steps k..ncv-1 write v_{j+1} straight into V[:, j+1]
    for (int j = k; j < ncv; ++j) {
      const double* vj = V + static_cast<int64_t>(j) * ld;
      double* w = V + static_cast<int64_t>(j + 1) * ld;
      double* hcol = bT.d() + static_cast<size_t>(j) * colw;
      QSI_TRY(launch_apply(op, vj, w, nullptr, 0.0, nullptr, nullptr, s));
      ++applies;
      // pass 1: h = V[:, 0..j]^T w (and w . w: column j + 1 is w itself)
      QSI_TRY(launch_multidot(V, ld, j + 2, w, D, bpart.d(), hcol, nullptr, s));
      // pass 2 fused with the norm and the test for pass 3; the conditional
      // third pass and the normalisation share one final pass
      QSI_TRY(launch_gs_reorth(V, ld, j + 1, hcol, w, D, bpart.d(), bg.d(), hcol,
                               static_cast<int*>(bflag.p), bbeta.d() + 2 * j, s));
    }
    QSI_CHECK_HIP(hipMemcpyAsync(hT.data() + static_cast<size_t>(k) * colw,
                                 bT.d() + static_cast<size_t>(k) * colw,
                                 static_cast<size_t>(ncv - k) * colw * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    QSI_CHECK_HIP(hipMemcpyAsync(hbeta.data(), bbeta.d(), 2 * ncv * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    QSI_CHECK_HIP(hipStreamSynchronize(s));
    for (int j = k; j < ncv; ++j) {
      const double* hh = hT.data() + static_cast<size_t>(j) * colw;
      for (int i = 0; i <= j; ++i) {
        if (i < k && j == k) {
          Tat(i, j) = Tat(j, i) = hh[i];  // coupling to the kept Ritz vectors
        } else if (i >= k) {
          Tat(i, j) = Tat(j, i) = hh[i];
        }
      }
      const double beta = hbeta[2 * j];
      double tnorm = 0.0;
      for (int i = 0; i <= j; ++i) tnorm = std::max(tnorm, std::fabs(Tat(i, j)));
      if (beta <= 1e-13 * std::max(tnorm, 1e-300)) {
        // invariant subspace: the projected matrix is exact (the later steps
        // of the cycle worked on a vanishing vector and are dropped)
        m = j + 1;
        beta_last = 0.0;
        exhausted = true;
        break;
      }
      beta_last = beta;
    }
    // Rayleigh-Ritz on the m x m projected matrix
    std::vector<double> Tm(static_cast<size_t>(m) * m);
    for (int j = 0; j < m; ++j)
      for (int i = 0; i < m; ++i) Tm[static_cast<size_t>(j) * m + i] = Tat(i, j);
    hostla::sym_eig(Tm, m, theta, Y);
    const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
    const int want = std::min(nev, m);
    nconv = 0;
    for (int i = 0; i < want; ++i) {
      const double resid = beta_last * std::fabs(Y[static_cast<size_t>(i) * m + (m - 1)]);
      if (resid < tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
      else break;  // count the leading converged run
    }
    if (nconv >= want || exhausted || restart >= max_restarts) break;
    // thick restart: keep the k lowest Ritz vectors plus the residual vector
    int kn = nev + std::min(nconv, (m - nev) / 2);
    kn = std::max(1, std::min(kn, m - 1));
    std::vector<double> Yk(Y.begin(), Y.begin() + static_cast<size_t>(kn) * m);
    QSI_CHECK_HIP(hipMemcpyAsync(bY.d(), Yk.data(), Yk.size() * sizeof(double),
                                 hipMemcpyHostToDevice, s));
    QSI_TRY(launch_combine(V, ld, m, bY.d(), kn, W, ld, D, s));
    QSI_CHECK_HIP(hipMemcpyAsync(W + static_cast<int64_t>(kn) * ld,
                                 V + static_cast<int64_t>(m) * ld,
                                 static_cast<size_t>(D) * sizeof(double),
                                 hipMemcpyDeviceToDevice, s));
    QSI_CHECK_HIP(hipStreamSynchronize(s));
    std::swap(V, W);
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < kn; ++i) Tat(i, i) = theta[i];
    k = kn;
  }
  const int nout = std::min(nev, m);
  for (int i = 0; i < nev; ++i)
    h_evals[i] = i < nout ? theta[i] : std::numeric_limits<double>::quiet_NaN();
  if (n_evecs > 0) {
    const int ne = std::min(n_evecs, m);
    std::vector<double> Yk(Y.begin(), Y.begin() + static_cast<size_t>(ne) * m);
    QSI_CHECK_HIP(hipMemcpyAsync(bY.d(), Yk.data(), Yk.size() * sizeof(double),
                                 hipMemcpyHostToDevice, s));
    QSI_TRY(launch_combine(V, ld, m, bY.d(), ne, d_evecs, D, D, s));
    QSI_CHECK_HIP(hipStreamSynchronize(s));
  }
  // hand the original buffer ownership back so both are freed
  if (V != bV.d()) std::swap(bV.p, bW.p);
  if (nconv_out) *nconv_out = nconv;
  if (n_applies) *n_applies = applies;
  if (nconv < std::min(nev, m)) {
    set_error("qsi_eigsh: not all requested eigenpairs converged");
    return QSI_E_NOCONV;
  }
  return QSI_OK;
}

}  // extern "C"

namespace {

// Y[:, l*n_vec + i] = O_l psi_i for every operator and vector (Y column-major
// with leading dimension ldy), then out[b*P + a] = Y[:, a] . W[:, b] for the
// P = n_ops*n_vec columns a and the n_w columns b of W (ldw), by multi-dots
// of up to 64 columns.  Synchronises.
int observe(const qsi_op_t* ops, int n_ops, const double* d_psi, int64_t ld, int n_vec,
            bool gram, std::vector<double>& out, int* P_out) {
  QSI_REQUIRE(ops && d_psi && n_ops >= 1 && n_vec >= 1, "observables: bad arguments");
  const qsi_basis_s* basis = ops[0] ? ops[0]->basis : nullptr;
  QSI_REQUIRE(basis != nullptr, "observables: null operator");
  for (int l = 0; l < n_ops; ++l)
    QSI_REQUIRE(ops[l] && ops[l]->basis == basis, "observables: operators on different bases");
  const int64_t D = basis->dim;
  QSI_REQUIRE(D > 0 && ld >= D, "observables: ld < dim");
  const int P = n_ops * n_vec;
  const int64_t ldy = (D + 63) / 64 * 64;
  hipStream_t s = nullptr;
  DevBuf bY, bpart, bout;
  QSI_CHECK_HIP(bY.alloc(static_cast<size_t>(ldy) * P * sizeof(double)));
  QSI_CHECK_HIP(bpart.alloc(static_cast<size_t>(multidot_partials()) * sizeof(double)));
  const int n_w = gram ? P : n_vec;
  QSI_CHECK_HIP(bout.alloc(static_cast<size_t>(P) * n_w * sizeof(double)));
  double* Y = bY.d();
  for (int l = 0; l < n_ops; ++l)
    for (int i = 0; i < n_vec; ++i)
      QSI_TRY(launch_apply(ops[l], d_psi + static_cast<int64_t>(i) * ld,
                           Y + static_cast<int64_t>(l * n_vec + i) * ldy, nullptr, 0.0, nullptr,
                           nullptr, s));
  for (int b = 0; b < n_w; ++b) {
    const double* w = gram ? Y + static_cast<int64_t>(b) * ldy : d_psi + static_cast<int64_t>(b) * ld;
    for (int a0 = 0; a0 < P; a0 += 64)
      QSI_TRY(launch_multidot(Y + static_cast<int64_t>(a0) * ldy, ldy, std::min(64, P - a0), w, D,
                              bpart.d(), bout.d() + static_cast<int64_t>(b) * P + a0, nullptr, s));
  }
  out.resize(static_cast<size_t>(P) * n_w);
  QSI_CHECK_HIP(hipMemcpy(out.data(), bout.d(), out.size() * sizeof(double), hipMemcpyDeviceToHost));
  *P_out = P;
  return QSI_OK;
}

}  // namespace

extern "C" {

int qsi_expectations(const qsi_op_t* ops, int n_ops, const double* d_psi, int64_t ld, int n_vec,
                     double* h_out) {
  QSI_REQUIRE(h_out != nullptr, "qsi_expectations: null output");
  std::vector<double> g;
  int P = 0;
  QSI_TRY(observe(ops, n_ops, d_psi, ld, n_vec, false, g, &P));
  for (int l = 0; l < n_ops; ++l)
    for (int i = 0; i < n_vec; ++i)
      for (int j = 0; j < n_vec; ++j)
        h_out[(static_cast<int64_t>(l) * n_vec + i) * n_vec + j] =
            g[static_cast<size_t>(j) * P + l * n_vec + i];
  return QSI_OK;
}

int qsi_cross_correlations(const qsi_op_t* ops, int n_ops, const double* d_psi, int64_t ld,
                           int n_vec, double* h_out) {
  QSI_REQUIRE(h_out != nullptr, "qsi_cross_correlations: null output");
  std::vector<double> g;
  int P = 0;
  QSI_TRY(observe(ops, n_ops, d_psi, ld, n_vec, true, g, &P));
  for (int l = 0; l < n_ops; ++l)
    for (int m = 0; m < n_ops; ++m)
      for (int i = 0; i < n_vec; ++i)
        for (int j = 0; j < n_vec; ++j)
          h_out[((static_cast<int64_t>(l) * n_ops + m) * n_vec + i) * n_vec + j] =
              g[static_cast<size_t>(m * n_vec + j) * P + l * n_vec + i];
  return QSI_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Basis enumeration
// ---------------------------------------------------------------------------

namespace {

using namespace qsi::enumr;

void dfs(const EnumLattice& L, Node root, std::vector<u128>& out) {
  std::vector<Node> stack;
  stack.push_back(root);
  while (!stack.empty()) {
    const Node c = stack.back();
    stack.pop_back();
    if (static_cast<int>(c.spin) == L.n) {
      out.push_back(c.state);
      continue;
    }
    fork(L, c, [&](const Node& x) { stack.push_back(x); });
  }
}

}  // namespace

extern "C" {

int qsi_enumerate_basis(int n_spins, int n_tetra, const int32_t* tetra_members,
                        const int32_t* order, int n_spinon_pairs, int n_threads,
                        uint64_t** out_lohi, int64_t* out_dim) {
  QSI_REQUIRE(out_lohi && out_dim, "qsi_enumerate_basis: null output");
  *out_lohi = nullptr;
  *out_dim = 0;
  QSI_REQUIRE(n_spins >= 1 && n_spins <= 128, "qsi_enumerate_basis: n_spins must be in [1,128]");
  QSI_REQUIRE(n_tetra >= 1 && tetra_members && order, "qsi_enumerate_basis: null lattice");
  QSI_REQUIRE(n_spinon_pairs >= 0, "qsi_enumerate_basis: n_spinon_pairs < 0");
  EnumLattice L;
  {
    const int rc = build_lattice(n_spins, n_tetra, tetra_members, order, n_spinon_pairs, L);
    if (rc != QSI_OK) return rc;
  }
  // breadth-first frontier, then independent depth-first subtrees
  std::vector<Node> front =
      frontier(L, static_cast<size_t>(std::max(1, n_threads)) * 64);
  const int nt = std::max(1, n_threads);
  std::vector<std::vector<u128>> parts(nt);
  std::atomic<size_t> next_node{0};
  auto worker = [&](int tid) {
    for (;;) {
      const size_t i = next_node.fetch_add(1);
      if (i >= front.size()) break;
      dfs(L, front[i], parts[tid]);
    }
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nt; ++t) pool.emplace_back(worker, t);
  worker(0);
  for (auto& th : pool) th.join();
  size_t total = 0;
  for (auto& p : parts) total += p.size();
  std::vector<u128> all;
  all.reserve(total);
  for (auto& p : parts) {
    all.insert(all.end(), p.begin(), p.end());
    std::vector<u128>().swap(p);
  }
  // undo the visiting permutation: bit k -> bit order[k] (bittools permute)
  for (auto& st : all) {
    u128 y = 0;
    for (int kk = 0; kk < n_spins; ++kk)
      if ((st >> kk) & 1) y |= static_cast<u128>(1) << order[kk];
    st = y;
  }
  std::sort(all.begin(), all.end());
  auto* outp = static_cast<uint64_t*>(std::malloc(std::max<size_t>(1, all.size()) * 16));
  if (!outp) {
    set_error("qsi_enumerate_basis: host allocation failed");
    return QSI_E_NOMEM;
  }
  for (size_t i = 0; i < all.size(); ++i) {
    outp[2 * i] = static_cast<uint64_t>(all[i]);
    outp[2 * i + 1] = static_cast<uint64_t>(all[i] >> 64);
  }
  *out_lohi = outp;
  *out_dim = static_cast<int64_t>(all.size());
  return QSI_OK;
}

void qsi_free(void* p) { std::free(p); }

}  // extern "C"
