/*
 * C++ face of the qsi_gpu C ABI (qsi_gpu.h): the reference's operator and
 * solver interfaces over the device library, header only.
 *
 *   qsi::GpuBasis         ZBasisInterp (components/operator/include/operator.hpp:
 *                         153-204): a sorted basis of [lo, hi] words, rank =
 *                         index, held on the device.
 *   qsi::GpuOp            LazyOpSum<double, B> + LazyOpSumProd
 *                         (components/operator/include/operator_matrix.hpp:
 *                         10-139): Scalar, rows(), cols(), perform_op,
 *                         evaluate, evaluate_add on host vectors, so
 *                         Spectra::SymEigsSolver<qsi::GpuOp> compiles unchanged;
 *                         evaluate_device / apply_block on device vectors.
 *                         Borrows the basis (reference-counted), copies the
 *                         terms.
 *   qsi::apply_fn         projED::RealApplyFn (components/linalg_routines/
 *                         include/common_bits.hpp:18-21): out += A in, host
 *                         vectors, for lanczos::eigval0 and friends.
 *   qsi::lanczos_eigval0  lanczos::eigval0 (components/linalg_routines/src/
 *                         lanczos.cpp:184-235) with the recurrence on the device.
 *   qsi::eigsh            compute_spectrum_iterative<LazyOpSumProd>
 *                         (include/matrix_diag_bits.hpp:51-95) with the Krylov
 *                         loop on the device (thick-restart Lanczos, lowest
 *                         eigenvalues).
 *   qsi::FtlmComputer     ftlm_computer (include/ftlm.hpp:22-122):
 *                         (ops, H, T_min, T_max, n_Ts), evolve(rng),
 *                         set_numerical_params, get_beta_grid, write_to_h5,
 *                         over textbook FTLM: each random start vector runs
 *                         krylov_dim Lanczos steps, whose Ritz pairs give
 *                         Z(beta), E(beta), C(beta) and, per operator,
 *                         sums of <psi|O|psi> with psi = exp(-beta H / 2) r.
 *
 * Every failing call throws qsi::Error (a std::runtime_error carrying
 * qsi_last_error()), as the reference's operator code throws
 * std::runtime_error; nothing is thrown across the C ABI itself.
 *
 * Needs the HIP runtime headers for device buffers: compile with hipcc, or
 * with a host compiler and -D__HIP_PLATFORM_AMD__ -I/opt/rocm/include; link
 * libqsi_gpu.so and libamdhip64.so.
 */
#ifndef QSI_GPU_HPP
#define QSI_GPU_HPP

#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <functional>
#include <limits>
#include <memory>
#include <random>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "qsi_gpu.h"

namespace qsi {

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

inline void throw_if(int rc, const char* what) {
  if (rc != QSI_OK) throw Error(rc, std::string(what) + ": " + qsi_last_error());
}

inline void throw_if_hip(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw Error(QSI_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// Device array of doubles (hipMalloc), move-only.
class DeviceVector {
 public:
  DeviceVector() = default;
  explicit DeviceVector(std::size_t n) : n_(n) {
    if (n) throw_if_hip(hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(double)), "hipMalloc");
  }
  DeviceVector(const double* host, std::size_t n) : DeviceVector(n) { upload(host); }
  ~DeviceVector() {
    if (p_) (void)hipFree(p_);
  }
  DeviceVector(const DeviceVector&) = delete;
  DeviceVector& operator=(const DeviceVector&) = delete;
  DeviceVector(DeviceVector&& o) noexcept : p_(o.p_), n_(o.n_) {
    o.p_ = nullptr;
    o.n_ = 0;
  }
  DeviceVector& operator=(DeviceVector&& o) noexcept {
    std::swap(p_, o.p_);
    std::swap(n_, o.n_);
    return *this;
  }
  double* data() { return p_; }
  const double* data() const { return p_; }
  std::size_t size() const { return n_; }
  void upload(const double* host) {
    if (n_) throw_if_hip(hipMemcpy(p_, host, n_ * sizeof(double), hipMemcpyHostToDevice), "upload");
  }
  void download(double* host) const {
    if (n_) throw_if_hip(hipMemcpy(host, p_, n_ * sizeof(double), hipMemcpyDeviceToHost), "download");
  }
  std::vector<double> to_host() const {
    std::vector<double> h(n_);
    download(h.data());
    return h;
  }

 private:
  double* p_ = nullptr;
  std::size_t n_ = 0;
};

// The basis: D rows of [lo, hi] words, strictly increasing as 128-bit
// integers (read_basis_hdf5's layout, components/common/include/
// basis_io_h5.hpp:90-141).  Copied to the device once.  Copies share the
// device basis.
class GpuBasis {
 public:
  GpuBasis(const uint64_t* lohi, int64_t dim, int n_spins, int bucket_bits = 0) {
    qsi_basis_t b = nullptr;
    throw_if(qsi_basis_create(lohi, dim, n_spins, bucket_bits, &b), "qsi_basis_create");
    h_.reset(b, [](qsi_basis_t p) { (void)qsi_basis_destroy(p); });
  }
  GpuBasis(const std::vector<uint64_t>& lohi, int n_spins, int bucket_bits = 0)
      : GpuBasis(lohi.data(), static_cast<int64_t>(lohi.size() / 2), n_spins, bucket_bits) {}
  int64_t dim() const { return qsi_basis_dim(h_.get()); }
  int n_spins() const { return qsi_basis_n_spins(h_.get()); }
  int bucket_bits() const { return qsi_basis_bucket_bits(h_.get()); }
  // The packed low keys (4 bytes per state) that the QSI_APPLY_LOW* kernels
  // bisect on (qsi_basis_build_lowkeys); false drops them.
  void build_lowkeys(bool enable = true) const {
    throw_if(qsi_basis_build_lowkeys(h_.get(), enable ? 1 : 0), "build_lowkeys");
  }
  bool has_lowkeys() const { return qsi_basis_has_lowkeys(h_.get()) != 0; }
  qsi_basis_t handle() const { return h_.get(); }

 private:
  std::shared_ptr<qsi_basis_s> h_;
};

// H over a basis.  perform_op / rows / cols / Scalar are the Spectra MatOp
// members of LazyOpSumProd; evaluate / evaluate_add are LazyOpSum's, on host
// vectors as the reference passes them (each call copies over PCIe and
// synchronises); evaluate_device and apply_block work on device vectors,
// stream ordered.  x and y must not alias.  Copies share the device
// operator, which keeps its basis alive.
class GpuOp {
 public:
  using Scalar = double;
  using Index = std::ptrdiff_t;

  GpuOp(const GpuBasis& basis, const std::vector<qsi_term_t>& terms, int n_slots = 0) {
    qsi_op_t o = nullptr;
    throw_if(qsi_op_create(basis.handle(), terms.data(), static_cast<int>(terms.size()), n_slots,
                           &o),
             "qsi_op_create");
    h_.reset(o, [](qsi_op_t p) { (void)qsi_op_destroy(p); });
  }

  Index rows() const { return static_cast<Index>(qsi_op_dim(h_.get())); }
  Index cols() const { return rows(); }

  // y_out = H x_in (LazyOpSumProd::perform_op).
  void perform_op(const double* x_in, double* y_out) const {
    throw_if(qsi_op_apply_host(h_.get(), x_in, y_out, 0), "perform_op");
  }
  // y = H x, y += H x (LazyOpSum::evaluate / evaluate_add).
  void evaluate(const double* x, double* y) const { perform_op(x, y); }
  void evaluate_add(const double* x, double* y) const {
    throw_if(qsi_op_apply_host(h_.get(), x, y, 1), "evaluate_add");
  }
  // The same on device vectors.
  void evaluate_device(const double* d_x, double* d_y, bool accumulate = false,
                       qsi_stream_t stream = nullptr) const {
    throw_if(qsi_op_apply(h_.get(), d_x, d_y, accumulate ? 1 : 0, stream), "evaluate_device");
  }
  // Y[:, r] (+)= H_r X[:, r], row-major D x R device blocks; d_coeffs
  // (n_slots x R) gives each column its own slot coefficients.
  void apply_block(const double* d_X, int64_t ldx, double* d_Y, int64_t ldy, int R,
                   const double* d_coeffs = nullptr, bool accumulate = false,
                   qsi_stream_t stream = nullptr) const {
    throw_if(qsi_op_apply_block(h_.get(), d_X, ldx, d_Y, ldy, R, d_coeffs, accumulate ? 1 : 0,
                                stream),
             "apply_block");
  }
  // Keep the connected ranks on the device (applies become gathers, same
  // bits); returns the off-diagonal nonzeros.
  int64_t cache_connectivity(bool enable = true) const {
    int64_t nnz = -1;
    throw_if(qsi_op_cache_connectivity(h_.get(), enable ? 1 : 0, &nnz), "cache_connectivity");
    return nnz;
  }
  // Apply kernel (one of the QSI_APPLY_* values; AUTO by default).
  void set_variant(int variant) const {
    throw_if(qsi_op_set_variant(h_.get(), variant), "set_variant");
  }
  int variant() const { return qsi_op_get_variant(h_.get()); }
  // One pass over the rows (qsi_op_verify_lowkeys; the basis needs low
  // keys): 2 when the low-key kernels need no compare, 1 when they compare
  // low keys, 0 when they compare states.
  int verify_lowkeys() const {
    int mode = 0;
    throw_if(qsi_op_verify_lowkeys(h_.get(), &mode), "verify_lowkeys");
    return mode;
  }
  // Times the apply kernels and keeps the fastest (R = 0: single vector).
  int autotune(int R = 0, int reps = 3) const {
    int best = -1;
    double ms = 0.0;
    throw_if(qsi_op_autotune(h_.get(), R, reps, nullptr, nullptr, &best, &ms), "autotune");
    return best;
  }
  qsi_op_t handle() const { return h_.get(); }

 private:
  std::shared_ptr<qsi_op_s> h_;
};

// projED::RealApplyFn: out += A in on host vectors (the functor holds a
// copy of the operator).
inline std::function<void(const double*, double*)> apply_fn(const GpuOp& op) {
  return [op](const double* in, double* out) { op.evaluate_add(in, out); };
}

// lanczos::Settings (components/linalg_routines/include/lanczos.hpp:9-20)
// and the result of lanczos::eigval0.
struct LanczosSettings {
  int64_t krylov_dim = 10;
  double abs_tol = 1e-8;
  double rel_tol = 1e-8;
  int64_t min_iterations = 30;
  int64_t max_iterations = 5000;
  uint32_t x0_seed = 10;
  bool calc_eigenvector = false;
  bool verify_eigenvector = false;
};

struct LanczosResult {
  double eigval = 0.0;
  bool eigval_converged = false;
  bool eigvec_converged = false;
  int64_t n_iterations = 0;
  double eigval_error = 0.0;
  double eigvec_error = 0.0;
  DeviceVector eigvec;  // D doubles when calc_eigenvector
};

inline LanczosResult lanczos_eigval0(const GpuOp& op, const LanczosSettings& st) {
  qsi_lanczos_settings_t c{st.krylov_dim,  st.abs_tol, st.rel_tol,
                           st.min_iterations, st.max_iterations, st.x0_seed,
                           st.calc_eigenvector ? 1 : 0, st.verify_eigenvector ? 1 : 0};
  LanczosResult r;
  if (st.calc_eigenvector) r.eigvec = DeviceVector(static_cast<std::size_t>(op.rows()));
  qsi_lanczos_result_t res{};
  throw_if(qsi_lanczos_eigval0(op.handle(), &c, &r.eigval, r.eigvec.data(), &res),
           "lanczos_eigval0");
  r.eigval_converged = res.eigval_converged != 0;
  r.eigvec_converged = res.eigvec_converged != 0;
  r.n_iterations = res.n_iterations;
  r.eigval_error = res.eigval_error;
  r.eigvec_error = res.eigvec_error;
  return r;
}

struct EigshResult {
  std::vector<double> evals;  // ascending
  int nconv = 0;
  int n_applies = 0;
  DeviceVector evecs;         // D x n_evecs column-major, when asked for
};

// Lowest nev eigenpairs; sizing as compute_spectrum_iterative
// (include/matrix_diag_bits.hpp:64-66: nev <= ncv - 1, ncv <= rows), with at
// least one eigenvalue for a one-state basis.  d_v0:
// device start vector, or null for qsi_vec_randn(seed).
inline EigshResult eigsh(const GpuOp& op, int nev, int ncv, double tol, int max_restarts = 1000,
                         int n_evecs = 0, const double* d_v0 = nullptr, uint64_t seed = 0) {
  const int64_t D = op.rows();
  // qsi_eigsh clamps ncv to the basis size itself; it needs ncv >= 2 on
  // entry, so the clamp here keeps at least 2
  ncv = static_cast<int>(std::max<int64_t>(2, std::min<int64_t>(ncv, D)));
  nev = std::min(nev, std::max(static_cast<int>(std::min<int64_t>(ncv, D)) - 1, 1));
  EigshResult r;
  r.evals.resize(static_cast<std::size_t>(std::max(nev, 0)));
  if (n_evecs > 0) r.evecs = DeviceVector(static_cast<std::size_t>(D) * n_evecs);
  throw_if(qsi_eigsh(op.handle(), nev, ncv, tol, max_restarts, d_v0, seed, r.evals.data(),
                     r.evecs.data(), n_evecs, &r.nconv, &r.n_applies),
           "eigsh");
  return r;
}

// ftlm_computer's interface (include/ftlm.hpp:22-122) over textbook FTLM.
// The beta grid is the reference's: n_Ts points evenly spaced in beta from
// 1/T_max to 1/T_min.  Each start vector r (normalised) runs krylov_dim
// steps of the 3-term recurrence on the device; with the Ritz values e_k and
// weights w_k = <r|phi_k>^2 of T_M, sample r adds
//   Z_r(beta) = sum_k w_k exp(-beta (e_k - e_ref)),  E_r, E2_r likewise
// with w_k e_k and w_k e_k^2, where e_ref (the lowest Ritz value seen so
// far) keeps the exponentials finite.  Over n samples of D-dimensional
// random vectors Z = (D / n) sum_r Z_r exp(-beta e_ref); E = sum E_r /
// sum Z_r; C = beta^2 (sum E2_r / sum Z_r - E^2).  A complete orthonormal
// start set (n = D) gives the exact traces.
//
// With operators (the reference's ops_to_evaluate), each sample also builds
// psi_r(beta) = exp(-beta (H - e_ref) / 2) r on the device (qsi_ftlm_sample:
// the Lanczos vectors of a second, identical pass combined with the
// coefficients of exp(-beta T_M / 2) e_1), and accumulates per operator O and
// beta
//   op_sum[O](beta)   = sum_r <psi_r|O|psi_r>
//   op_sum_2[O](beta) = sum_r <psi_r|O|psi_r>^2
//   norm_sum(beta)    = sum_r <psi_r|psi_r>
// (all with the same shift e_ref), so <O>(beta) = op_sum / norm_sum; with a
// complete orthonormal start set op_sum = Tr(exp(-beta (H - e_ref)) O)
// exactly.  write_to_h5 writes op_sum and op_sum_2 as the datasets
// "<name>_sum" and "<name>_sum_2" (ftlm.hpp:113-124; available when
// <hdf5.h> is included before this header).  The reference's own loop
// (ftlm.hpp:56-91) does not build and divides its running sums by each
// sample's norm; these are the sums its comments describe.  tol is kept for
// the interface and not used.
class FtlmComputer {
 public:
  FtlmComputer(const GpuOp& H, double T_min, double T_max, int n_Ts)
      : FtlmComputer(std::vector<GpuOp>{}, H, T_min, T_max, n_Ts) {}

  FtlmComputer(const std::vector<GpuOp>& ops, const GpuOp& H, double T_min, double T_max,
               int n_Ts)
      : op_(H), H_(H.handle()), ops_(ops) {
    if (T_min >= T_max) throw std::runtime_error("Invalid spec: T_max must be greater than T_min");
    if (n_Ts < 2) throw std::runtime_error("Invalid spec: n_Ts must be at least 2");
    const double b0 = 1.0 / T_max, b1 = 1.0 / T_min, db = (b1 - b0) / (n_Ts - 1);
    for (int j = 0; j < n_Ts; ++j) beta_.push_back(b0 + j * db);
    z_.assign(beta_.size(), 0.0);
    e_.assign(beta_.size(), 0.0);
    e2_.assign(beta_.size(), 0.0);
    norm_.assign(beta_.size(), 0.0);
    for (const GpuOp& o : ops_) {
      if (o.rows() != H.rows())
        throw std::runtime_error("ftlm_computer: operator and Hamiltonian sizes differ");
      op_h_.push_back(o.handle());
    }
    osum_.assign(ops_.size() * beta_.size(), 0.0);
    osum2_.assign(ops_.size() * beta_.size(), 0.0);
  }

  void set_numerical_params(int krylov_dim = 30, double tol = 1e-10) {
    krylov_dim_ = krylov_dim;
    tol_ = tol;
  }

  // One sample: a Gaussian vector drawn from rng as init_random does
  // (include/ftlm.hpp:10-17: a fresh std::normal_distribution<double>,
  // D draws, then normalised).
  void evolve(std::mt19937& rng) { evolve_block(rng, 1); }

  // R samples (R <= 64) from rng in turn; without operators they run side by
  // side in one launch, with operators one after another.
  void evolve_block(std::mt19937& rng, int R) {
    const int64_t D = qsi_op_dim(H_);
    std::vector<double> V(static_cast<std::size_t>(D) * R);
    for (int c = 0; c < R; ++c) {
      std::normal_distribution<double> n;
      double s = 0.0;
      for (int64_t i = 0; i < D; ++i) {
        const double v = n(rng);
        V[static_cast<std::size_t>(i) * R + c] = v;
        s += v * v;
      }
      const double inv = 1.0 / std::sqrt(s);
      for (int64_t i = 0; i < D; ++i) V[static_cast<std::size_t>(i) * R + c] *= inv;
    }
    evolve_vectors(V.data(), R);
  }

  // R given start vectors (host, D x R row-major; normalised on the device).
  void evolve_vectors(const double* h_V0, int R) {
    const int64_t D = qsi_op_dim(H_);
    const int M = krylov_dim_;
    std::vector<double> ev(M), fr(M);
    if (!ops_.empty()) {
      std::vector<double> col(static_cast<std::size_t>(D)), al(M), be(M);
      const std::size_t nb = beta_.size(), no = ops_.size();
      std::vector<double> nrm(nb), ex(no * nb);
      for (int c = 0; c < R; ++c) {
        for (int64_t i = 0; i < D; ++i) col[static_cast<std::size_t>(i)] = h_V0[i * R + c];
        DeviceVector r(col.data(), col.size());
        int m = 0;
        double used = 0.0;
        throw_if(qsi_ftlm_sample(H_, op_h_.data(), static_cast<int>(no), r.data(), M,
                                 beta_.data(), static_cast<int>(nb),
                                 std::numeric_limits<double>::quiet_NaN(), &used, nrm.data(),
                                 ex.data(), al.data(), be.data(), &m),
                 "ftlm evolve");
        throw_if(qsi_tridiag_eig(al.data(), m > 1 ? be.data() : nullptr, m, ev.data(), fr.data()),
                 "ftlm quadrature");
        // the sample's shift is its lowest Ritz value; bring both to the lower
        shift_to(std::min(e_ref_, used));
        for (std::size_t j = 0; j < nb; ++j) {
          const double f = std::exp(-beta_[j] * (used - e_ref_));  // <= 1
          norm_[j] += f * nrm[j];
          for (std::size_t l = 0; l < no; ++l) {
            const double o = f * ex[l * nb + j];
            osum_[l * nb + j] += o;
            osum2_[l * nb + j] += o * o;
          }
        }
        add_sample(ev.data(), fr.data(), m);
      }
    } else {
      DeviceVector V0(h_V0, static_cast<std::size_t>(D) * R);
      std::vector<double> al(static_cast<std::size_t>(R) * M), be(static_cast<std::size_t>(R) * M);
      const double* v0 = V0.data();
      throw_if(qsi_lanczos_block(&H_, 1, R, &v0, nullptr, M, al.data(), be.data(), 1, nullptr),
               "ftlm evolve");
      for (int c = 0; c < R; ++c) {
        const double* a = al.data() + static_cast<std::size_t>(c) * M;
        const double* b = be.data() + static_cast<std::size_t>(c) * M;
        // a beta that underflows ends the sample's recurrence
        int m = 1;
        while (m < M && b[m - 1] != 0.0) ++m;
        throw_if(qsi_tridiag_eig(a, m > 1 ? b : nullptr, m, ev.data(), fr.data()),
                 "ftlm quadrature");
        add_sample(ev.data(), fr.data(), m);
      }
    }
    n_samples_ += R;
    dim_ = D;
  }

  const std::vector<double>& get_beta_grid() const { return beta_; }
  std::vector<double> get_temperature_grid() const {
    std::vector<double> t;
    for (double b : beta_) t.push_back(1.0 / b);
    return t;
  }
  int64_t n_samples() const { return n_samples_; }
  int krylov_dim() const { return krylov_dim_; }
  double tolerance() const { return tol_; }
  double energy_shift() const { return e_ref_; }
  // Z(beta) exp(beta e_ref): (D / n) sum_r Z_r
  std::vector<double> partition_function_shifted() const {
    std::vector<double> z(beta_.size());
    const double f = n_samples_ ? static_cast<double>(dim_) / n_samples_ : 0.0;
    for (std::size_t j = 0; j < z.size(); ++j) z[j] = f * z_[j];
    return z;
  }
  std::vector<double> energy() const {
    std::vector<double> e(beta_.size());
    for (std::size_t j = 0; j < e.size(); ++j) e[j] = e_[j] / z_[j];
    return e;
  }
  std::vector<double> specific_heat() const {
    std::vector<double> c(beta_.size());
    for (std::size_t j = 0; j < c.size(); ++j) {
      const double e = e_[j] / z_[j];
      c[j] = beta_[j] * beta_[j] * (e2_[j] / z_[j] - e * e);
    }
    return c;
  }
  // Operator sums over the samples so far (shift e_ref), per beta.
  std::vector<double> op_sum(int op_idx) const { return slice(osum_, op_idx); }
  std::vector<double> op_sum_2(int op_idx) const { return slice(osum2_, op_idx); }
  const std::vector<double>& norm_sum() const { return norm_; }
  // <O>(beta) = op_sum / norm_sum
  std::vector<double> op_expectation(int op_idx) const {
    std::vector<double> o = op_sum(op_idx);
    for (std::size_t j = 0; j < o.size(); ++j) o[j] /= norm_[j];
    return o;
  }

#ifdef H5_VERS_MAJOR
  // "<name>_sum" and "<name>_sum_2", one value per beta (ftlm.hpp:113-124).
  void write_to_h5(hid_t file_id, int op_idx, const std::string& dset_name) const {
    write_1d(file_id, dset_name + "_sum", op_sum(op_idx));
    write_1d(file_id, dset_name + "_sum_2", op_sum_2(op_idx));
  }
#endif

 private:
  std::vector<double> slice(const std::vector<double>& a, int op_idx) const {
    if (op_idx < 0 || static_cast<std::size_t>(op_idx) >= ops_.size())
      throw std::out_of_range("ftlm_computer: operator index out of range");
    const std::size_t nb = beta_.size();
    return std::vector<double>(a.begin() + op_idx * nb, a.begin() + (op_idx + 1) * nb);
  }

#ifdef H5_VERS_MAJOR
  static void write_1d(hid_t fid, const std::string& name, const std::vector<double>& v) {
    const hsize_t dims[1] = {static_cast<hsize_t>(v.size())};
    const hid_t sp = H5Screate_simple(1, dims, nullptr);
    const hid_t ds = H5Dcreate2(fid, name.c_str(), H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT,
                                H5P_DEFAULT, H5P_DEFAULT);
    const herr_t rc = ds < 0 ? -1 : H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                             v.data());
    if (ds >= 0) H5Dclose(ds);
    H5Sclose(sp);
    if (rc < 0) throw std::runtime_error("ftlm_computer: cannot write dataset " + name);
  }
#endif

  // Moves every sum to the reference energy lo (lo <= e_ref_).
  void shift_to(double lo) {
    if (!(lo < e_ref_)) return;
    if (std::isfinite(e_ref_))
      for (std::size_t j = 0; j < beta_.size(); ++j) {
        const double f = std::exp(-beta_[j] * (e_ref_ - lo));
        z_[j] *= f;
        e_[j] *= f;
        e2_[j] *= f;
        norm_[j] *= f;
        for (std::size_t l = 0; l < ops_.size(); ++l) {
          osum_[l * beta_.size() + j] *= f;
          osum2_[l * beta_.size() + j] *= f * f;
        }
      }
    e_ref_ = lo;
  }

  void add_sample(const double* ev, const double* fr, int m) {
    shift_to(*std::min_element(ev, ev + m));
    for (std::size_t j = 0; j < beta_.size(); ++j) {
      for (int k = 0; k < m; ++k) {
        const double w = fr[k] * fr[k] * std::exp(-beta_[j] * (ev[k] - e_ref_));
        z_[j] += w;
        e_[j] += w * ev[k];
        e2_[j] += w * ev[k] * ev[k];
      }
    }
  }

  GpuOp op_;  // keeps the operator alive
  qsi_op_t H_;
  std::vector<GpuOp> ops_;
  std::vector<qsi_op_t> op_h_;
  int krylov_dim_ = 30;
  double tol_ = 1e-10;
  std::vector<double> beta_, z_, e_, e2_, norm_, osum_, osum2_;
  double e_ref_ = std::numeric_limits<double>::infinity();
  int64_t n_samples_ = 0;
  int64_t dim_ = 0;
};

}  // namespace qsi

#endif  // QSI_GPU_HPP
