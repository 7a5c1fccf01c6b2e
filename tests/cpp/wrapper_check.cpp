// Runs the C++ face of the library (include/qsi_gpu.hpp) on inputs written
// by tests/test_gpu_cpp.py and writes what it computes next to them.
//
// usage: wrapper_check DIR
//   in:  DIR/meta.txt "D n_spins n_terms", DIR/basis.bin (D x 2 uint64),
//        DIR/terms.bin (n_terms qsi_term_t), DIR/x.bin (D doubles)
//   out: y_perform.bin   H x through GpuOp::perform_op (host vectors)
//        y_applyfn.bin   1 + H x through apply_fn (out += A in)
//        y_device.bin    H x through GpuOp::evaluate_device
//        evals.bin       eigsh: 5 lowest, ncv 15, tol 1e-10, seed 0
//        ftlm_exact.bin  [T grid, Z shifted, E, C, e_ref] of FtlmComputer over
//                        the complete unit-vector start set, krylov_dim 80
//        ftlm_rng.bin    [Z shifted, E] of 4 samples from mt19937(7) one at a
//                        time, then of the same 4 as one block
//   stdout: "lowkeys ok MODE" when, after GpuBasis::build_lowkeys and
//           GpuOp::verify_lowkeys (MODE), GpuOp::evaluate_device on each
//           low-key kernel and on AUTO gives y_device.bin's bits;
//           "one-state ok" when eigsh on the basis of the first state alone
//           gives that state's diagonal element;
//           "error ok" when an unsorted basis is refused with qsi::Error
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <random>
#include <string>
#include <vector>

#include "qsi_gpu.hpp"

template <typename T>
static std::vector<T> read_all(const std::string& path, size_t n) {
  std::vector<T> v(n);
  std::ifstream f(path, std::ios::binary);
  f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
  if (!f) throw std::runtime_error("short read: " + path);
  return v;
}

static void write_all(const std::string& path, const std::vector<double>& v) {
  std::ofstream f(path, std::ios::binary);
  f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * 8));
}

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: wrapper_check DIR\n");
    return 2;
  }
  const std::string dir = argv[1];
  int64_t D = 0;
  int n_spins = 0, n_terms = 0;
  {
    std::ifstream m(dir + "/meta.txt");
    m >> D >> n_spins >> n_terms;
    if (!m || D <= 0) throw std::runtime_error("bad meta.txt");
  }
  const auto lohi = read_all<uint64_t>(dir + "/basis.bin", static_cast<size_t>(D) * 2);
  const auto terms = read_all<qsi_term_t>(dir + "/terms.bin", static_cast<size_t>(n_terms));
  const auto x = read_all<double>(dir + "/x.bin", static_cast<size_t>(D));

  qsi::GpuBasis basis(lohi, n_spins);
  qsi::GpuOp op(basis, terms);
  if (op.rows() != D || op.cols() != D) throw std::runtime_error("rows/cols");

  std::vector<double> y(D);
  op.perform_op(x.data(), y.data());
  write_all(dir + "/y_perform.bin", y);

  std::vector<double> y_dev;
  {
    qsi::DeviceVector dx(x.data(), static_cast<size_t>(D)), dy(static_cast<size_t>(D));
    op.evaluate_device(dx.data(), dy.data());
    y_dev = dy.to_host();
    write_all(dir + "/y_device.bin", y_dev);
  }

  // the low-key kernels through the C++ face: the same bits
  {
    basis.build_lowkeys();
    const int mode = op.verify_lowkeys();
    bool same = basis.has_lowkeys();
    qsi::DeviceVector dx(x.data(), static_cast<size_t>(D)), dy(static_cast<size_t>(D));
    for (int v : {QSI_APPLY_LOW2, QSI_APPLY_LOW4, QSI_APPLY_LOW8, QSI_APPLY_AUTO}) {
      op.set_variant(v);
      same = same && op.variant() == v;
      op.evaluate_device(dx.data(), dy.data());
      const std::vector<double> yl = dy.to_host();
      for (int64_t i = 0; i < D && same; ++i) same = yl[i] == y_dev[i];
    }
    basis.build_lowkeys(false);
    std::printf("lowkeys %s %d\n", same && !basis.has_lowkeys() ? "ok" : "bad", mode);
  }

  std::vector<double> y1(D, 1.0);
  auto f = qsi::apply_fn(op);
  f(x.data(), y1.data());
  write_all(dir + "/y_applyfn.bin", y1);

  const auto es = qsi::eigsh(op, 5, 15, 1e-10);
  if (es.nconv != 5) throw std::runtime_error("eigsh did not converge");
  write_all(dir + "/evals.bin", es.evals);

  {
    qsi::FtlmComputer ft(op, 0.5, 5.0, 6);
    ft.set_numerical_params(80);
    const int R = 64;
    for (int64_t c0 = 0; c0 < D; c0 += R) {
      const int r = static_cast<int>(std::min<int64_t>(R, D - c0));
      std::vector<double> V(static_cast<size_t>(D) * r, 0.0);
      for (int c = 0; c < r; ++c) V[static_cast<size_t>(c0 + c) * r + c] = 1.0;
      ft.evolve_vectors(V.data(), r);
    }
    std::vector<double> out = ft.get_temperature_grid();
    for (double v : ft.partition_function_shifted()) out.push_back(v);
    for (double v : ft.energy()) out.push_back(v);
    for (double v : ft.specific_heat()) out.push_back(v);
    out.push_back(ft.energy_shift());
    write_all(dir + "/ftlm_exact.bin", out);
  }
  {
    qsi::FtlmComputer a(op, 0.5, 5.0, 6), b(op, 0.5, 5.0, 6);
    a.set_numerical_params(40);
    b.set_numerical_params(40);
    std::mt19937 ra(7), rb(7);
    for (int k = 0; k < 4; ++k) a.evolve(ra);
    b.evolve_block(rb, 4);
    std::vector<double> out;
    for (auto* c : {&a, &b}) {
      const auto z = c->partition_function_shifted();
      const auto e = c->energy();
      out.insert(out.end(), z.begin(), z.end());
      out.insert(out.end(), e.begin(), e.end());
    }
    write_all(dir + "/ftlm_rng.bin", out);
  }

  // a one-state basis: eigsh gives its one eigenvalue, the diagonal element
  {
    std::vector<uint64_t> one(lohi.begin(), lohi.begin() + 2);
    qsi::GpuBasis b1(one, n_spins);
    qsi::GpuOp op1(b1, terms);
    std::vector<double> e0 = {1.0}, h00 = {0.0};
    op1.perform_op(e0.data(), h00.data());
    const auto r1 = qsi::eigsh(op1, 5, 15, 1e-10);
    const bool ok = r1.evals.size() == 1 &&
                    std::fabs(r1.evals[0] - h00[0]) <= 1e-12 * (1.0 + std::fabs(h00[0]));
    std::printf("one-state %s %.17g %.17g\n", ok ? "ok" : "bad",
                r1.evals.empty() ? 0.0 : r1.evals[0], h00[0]);
  }

  // the error path: an unsorted basis is refused
  try {
    std::vector<uint64_t> bad = {3, 0, 1, 0};
    qsi::GpuBasis b2(bad, n_spins);
    std::printf("error missing\n");
  } catch (const qsi::Error& e) {
    std::printf("error ok (%d): %s\n", e.code(), e.what());
  }
  return 0;
}
