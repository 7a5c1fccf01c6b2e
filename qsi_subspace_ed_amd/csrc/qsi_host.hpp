// qsi_host.hpp — header-only C++ face of the C ABI (include/qsi_gpu.h) with
// the reference's operator interfaces, so C++ callers can swap it in:
//
//   qsi::GpuOp      perform_op / evaluate / evaluate_add / rows / cols / Scalar
//                   (LazyOpSum + LazyOpSumProd, components/operator/include/
//                   operator_matrix.hpp:10-139; a Spectra MatOp)
//   qsi::apply_fn   a projED::RealApplyFn-shaped functor computing out += A in
//                   (components/linalg_routines/include/common_bits.hpp:18-21)
//   qsi::GpuBasis   device basis (ZBasisInterp, operator.hpp:197-204)
//
// Errors from the C ABI become std::runtime_error here (the reference throws
// std::runtime_error / std::out_of_range from the same places).
#pragma once

#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qsi_gpu.h"

namespace qsi {

inline void throw_if(int rc, const char* what) {
  if (rc != QSI_OK)
    throw std::runtime_error(std::string(what) + ": " + qsi_last_error());
}

class GpuBasis {
 public:
  // states: D rows of [lo, hi] words, strictly increasing (the HDF5 layout)
  GpuBasis(const std::vector<uint64_t>& lohi, int n_spins, int bucket_bits = 0) {
    qsi_basis_t b = nullptr;
    throw_if(qsi_basis_create(lohi.data(), static_cast<int64_t>(lohi.size() / 2),
                              n_spins, bucket_bits, &b),
             "qsi_basis_create");
    h_.reset(b, [](qsi_basis_t p) { qsi_basis_destroy(p); });
  }
  int64_t dim() const { return qsi_basis_dim(h_.get()); }
  qsi_basis_t handle() const { return h_.get(); }

 private:
  std::shared_ptr<qsi_basis_s> h_;
};

class GpuOp {
 public:
  using Scalar = double;
  GpuOp(const GpuBasis& basis, const std::vector<qsi_term_t>& terms, int n_slots = 0) {
    qsi_op_t o = nullptr;
    throw_if(qsi_op_create(basis.handle(), terms.data(), static_cast<int>(terms.size()),
                           n_slots, &o),
             "qsi_op_create");
    h_.reset(o, [](qsi_op_t p) { qsi_op_destroy(p); });
  }
  int64_t rows() const { return qsi_op_dim(h_.get()); }
  int64_t cols() const { return qsi_op_dim(h_.get()); }

  // Host vectors (Spectra passes host pointers): y = H x.
  void perform_op(const double* x_in, double* y_out) const {
    throw_if(qsi_op_apply_host(h_.get(), x_in, y_out, 0), "qsi_op_apply_host");
  }
  void evaluate(const double* x, double* y) const { perform_op(x, y); }
  void evaluate_add(const double* x, double* y) const {
    throw_if(qsi_op_apply_host(h_.get(), x, y, 1), "qsi_op_apply_host");
  }
  // Device vectors, stream ordered.
  void evaluate_device(const double* d_x, double* d_y, bool accumulate = false,
                       qsi_stream_t stream = nullptr) const {
    throw_if(qsi_op_apply(h_.get(), d_x, d_y, accumulate ? 1 : 0, stream), "qsi_op_apply");
  }
  qsi_op_t handle() const { return h_.get(); }

 private:
  std::shared_ptr<qsi_op_s> h_;
};

// out += A in (projED::RealApplyFn contract) on host vectors.
inline std::function<void(const double*, double*)> apply_fn(const GpuOp& op) {
  return [op](const double* in, double* out) { op.evaluate_add(in, out); };
}

// Lowest eigenpairs on the device (the mfsparse counterpart).  Returns the
// nev eigenvalues ascending.
inline std::vector<double> eigsh(const GpuOp& op, int nev, int ncv, double tol,
                                 int max_restarts = 1000, uint64_t seed = 0) {
  std::vector<double> ev(static_cast<size_t>(nev));
  int nconv = 0, napp = 0;
  throw_if(qsi_eigsh(op.handle(), nev, ncv, tol, max_restarts, nullptr, seed, ev.data(),
                     nullptr, 0, &nconv, &napp),
           "qsi_eigsh");
  return ev;
}

}  // namespace qsi
