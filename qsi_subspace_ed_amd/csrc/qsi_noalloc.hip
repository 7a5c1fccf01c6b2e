// Allocation-free variant of qsi_lanczos_coeffs for callers that cannot
// allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc).  The caller brings the two Lanczos
// vectors; the small buffers the recurrence needs besides them (the partial
// sums of the fused apply + dot and of the norms, the per-step alpha / beta /
// 1/beta history and two scalars) live in one file-scope device array,
// `scratch`, which every GPU the library runs on holds until the process
// ends.  Each device's copy is found once, on the first call made with that
// device current, by hipGetSymbolAddress; host code only ever passes that
// address, never the array's name.
//
// Every kernel of the run writes the scratch it later reads, so nothing in it
// needs clearing between calls; calls that overlap on one GPU share the array
// and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "qsi_internal.hpp"

namespace qsi {
namespace {

constexpr int64_t kNaPartials = int64_t{1} << 20;  // apply + dot partials: D <= 2^28
constexpr int kNaMaxSteps = 4096;                  // alpha, beta, 1/beta per step
constexpr int64_t kNaHistory = 3 * int64_t{kNaMaxSteps};
constexpr int64_t kNaScalars = 32;
constexpr int64_t kNaDoubles = kNaPartials + kNaHistory + kNaScalars;  // 8.1 MiB

}  // namespace
}  // namespace qsi

static __device__ __align__(256) double scratch[qsi::kNaDoubles];

namespace qsi {
namespace {

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;
double* g_na_ptr[kNaMaxDevices] = {};

// Device address of `scratch` on the current device (resolved once per
// device, with that device current).
int na_scratch(int device, double** out) {
  QSI_REQUIRE(device >= 0 && device < kNaMaxDevices,
              "qsi_lanczos_coeffs_noalloc: device index out of range");
  std::lock_guard<std::mutex> lock(g_na_mu);
  if (!g_na_ptr[device]) {
    void* p = nullptr;
    QSI_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    QSI_REQUIRE(p != nullptr, "qsi_lanczos_coeffs_noalloc: no device address for the scratch");
    g_na_ptr[device] = static_cast<double*>(p);
  }
  *out = g_na_ptr[device];
  return QSI_OK;
}

}  // namespace
}  // namespace qsi

using namespace qsi;

extern "C" int qsi_lanczos_coeffs_noalloc(qsi_op_t op, double* d_v, double* d_u, int M,
                                          double* h_alpha, double* h_beta, int* n_steps) {
  QSI_REQUIRE(op && d_v && d_u && h_alpha && h_beta,
              "qsi_lanczos_coeffs_noalloc: null argument");
  QSI_REQUIRE(d_v != d_u, "qsi_lanczos_coeffs_noalloc: d_v and d_u must be different vectors");
  QSI_REQUIRE(M >= 1 && M <= kNaMaxSteps, "qsi_lanczos_coeffs_noalloc: M must be in [1, 4096]");
  QSI_REQUIRE(op->basis->dim > 0, "qsi_lanczos_coeffs_noalloc: empty basis");
  QSI_REQUIRE(lanczos_coeffs_partials(op) <= kNaPartials,
              "qsi_lanczos_coeffs_noalloc: basis too large for the allocation-free scratch "
              "(at most 2^28 states; use qsi_lanczos_coeffs)");
  int dev = 0;
  QSI_CHECK_HIP(hipGetDevice(&dev));
  QSI_REQUIRE(dev == op->basis->device,
              "qsi_lanczos_coeffs_noalloc: the operator's device is not the current device");
  double* sp = nullptr;
  const int rc = na_scratch(dev, &sp);
  if (rc != QSI_OK) return rc;
  return lanczos_coeffs_run(op, d_v, d_u, sp, sp + kNaPartials, sp + kNaPartials + kNaHistory,
                            M, h_alpha, h_beta, n_steps, nullptr);
}
