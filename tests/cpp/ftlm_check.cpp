// Runs qsi::FtlmComputer with operators (ftlm_computer(ops, H, T_min, T_max,
// n_Ts), include/ftlm.hpp:22-122) over the complete set of unit start
// vectors and writes its sums through write_to_h5, on inputs written by
// tests/test_gpu_cpp.py.
//
// usage: ftlm_check DIR
//   in:  DIR/meta.txt "D n_spins n_terms n_ops k_0 ... k_{n_ops-1}" (k_l =
//        terms of operator l), DIR/basis.bin (D x 2 uint64), DIR/terms.bin
//        (H, n_terms qsi_term_t), DIR/ops.bin (the operators' terms, one
//        after another)
//   out: DIR/ftlm.h5      "op<l>_sum", "op<l>_sum_2" per operator
//        DIR/ftlm.bin     [T grid, norm_sum, e_ref] (n_Ts, n_Ts, 1 doubles)
#include <hdf5.h>

#include <cstdio>
#include <fstream>
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

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: ftlm_check DIR\n");
    return 2;
  }
  const std::string dir = argv[1];
  int64_t D = 0;
  int n_spins = 0, n_terms = 0, n_ops = 0;
  std::vector<int> k;
  {
    std::ifstream m(dir + "/meta.txt");
    m >> D >> n_spins >> n_terms >> n_ops;
    k.resize(n_ops);
    for (int& v : k) m >> v;
    if (!m || D <= 0) throw std::runtime_error("bad meta.txt");
  }
  int total = 0;
  for (int v : k) total += v;
  const auto lohi = read_all<uint64_t>(dir + "/basis.bin", static_cast<size_t>(D) * 2);
  const auto terms = read_all<qsi_term_t>(dir + "/terms.bin", static_cast<size_t>(n_terms));
  const auto oterms = read_all<qsi_term_t>(dir + "/ops.bin", static_cast<size_t>(total));

  qsi::GpuBasis basis(lohi, n_spins);
  qsi::GpuOp H(basis, terms);
  std::vector<qsi::GpuOp> ops;
  int at = 0;
  for (int l = 0; l < n_ops; ++l) {
    ops.emplace_back(basis, std::vector<qsi_term_t>(oterms.begin() + at, oterms.begin() + at + k[l]));
    at += k[l];
  }
  qsi::FtlmComputer ft(ops, H, 0.5, 5.0, 6);
  ft.set_numerical_params(60);
  std::vector<double> e(static_cast<size_t>(D), 0.0);
  for (int64_t i = 0; i < D; ++i) {
    e[static_cast<size_t>(i)] = 1.0;
    ft.evolve_vectors(e.data(), 1);
    e[static_cast<size_t>(i)] = 0.0;
  }
  const hid_t f = H5Fcreate((dir + "/ftlm.h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (f < 0) throw std::runtime_error("cannot create ftlm.h5");
  for (int l = 0; l < n_ops; ++l) ft.write_to_h5(f, l, "op" + std::to_string(l));
  H5Fclose(f);
  std::vector<double> out = ft.get_temperature_grid();
  for (double v : ft.norm_sum()) out.push_back(v);
  out.push_back(ft.energy_shift());
  std::ofstream o(dir + "/ftlm.bin", std::ios::binary);
  o.write(reinterpret_cast<const char*>(out.data()), static_cast<std::streamsize>(out.size() * 8));
  std::printf("ftlm_check ok: %lld samples\n", static_cast<long long>(ft.n_samples()));
  return 0;
}
