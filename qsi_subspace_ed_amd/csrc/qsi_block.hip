// Block Lanczos recurrences for many start vectors on several operators in
// one launch per kernel — the batched FTLM path.
//
// R start vectors per sector (operator + basis) run the reference's 3-term
// recurrence (components/linalg_routines/src/lanczos.cpp:111-146, without the
// convergence checks) side by side, M steps, no reorthogonalisation, which is
// what FTLM needs: the (alpha, beta) coefficients of every sample.  Sectors of
// different basis size share each launch: their rows are cut into tiles of 64
// and one persistent grid of wavefronts walks the concatenated tile list, so
// each basis state's flip tests and rank lookups serve all R columns.
//
// Per step (three rotating D x R blocks A = v_{j-1}, B = v_j, C):
//   k_blk_apply      C = H B, per-wave partials of <B_c, C_c>
//   k_blk_alpha      alpha[s][c]
//   k_blk_update     C -= alpha B + beta A, per-wave partials of |C_c|^2
//   k_blk_beta       beta[s][c], 1/beta (0 when beta underflows)
//   k_blk_scale      C *= 1/beta;  then (A, B, C) <- (B, C, A)
// The five launches of a step are captured once per rotation phase as a HIP
// graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qsi_blockk.hpp"
#include "qsi_hostla.hpp"

namespace qsi {
namespace {

struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n * sizeof(T), 8));
    if (e == hipSuccess) {
      ptrs.push_back(q);
      *p = static_cast<T*>(q);
    }
    return e;
  }
};

}  // namespace
}  // namespace qsi

using namespace qsi;

extern "C" int qsi_lanczos_block(const qsi_op_t* ops, int n_sectors, int R,
                                 const double* const* d_V0,
                                 const double* const* d_coeffs, int M,
                                 double* h_alpha, double* h_beta, int use_graph,
                                 double* elapsed_ms) {
  QSI_REQUIRE(ops && d_V0 && h_alpha && h_beta, "qsi_lanczos_block: null argument");
  QSI_REQUIRE(n_sectors >= 1 && n_sectors <= kMaxSectors,
              "qsi_lanczos_block: n_sectors must be in [1, 8]");
  QSI_REQUIRE(R >= 1 && R <= 64, "qsi_lanczos_block: R must be in [1, 64]");
  QSI_REQUIRE(M >= 1, "qsi_lanczos_block: M must be >= 1");
  const int nw = ops[0]->basis->nw;
  BlkArgs args{};
  args.n_sec = n_sectors;
  args.R = R;
  int G = 64;
  while (G / 2 >= R && G > 1) G >>= 1;
  args.G = G;
  args.ldx = args.ldy = R;
  args.accumulate = 0;
  args.y_colmajor = 0;
  int64_t tiles = 0, rows = 0;
  for (int s = 0; s < n_sectors; ++s) {
    const qsi_op_s* op = ops[s];
    QSI_REQUIRE(op != nullptr && op->basis->dim > 0, "qsi_lanczos_block: empty sector");
    QSI_REQUIRE(op->basis->nw == nw, "qsi_lanczos_block: sectors mix <=64 and >64 spins");
    QSI_REQUIRE(d_V0[s] != nullptr, "qsi_lanczos_block: null start block");
    QSI_REQUIRE(!d_coeffs || !d_coeffs[s] || op->n_slots > 0,
                "qsi_lanczos_block: coefficients given for an operator without slots");
    BlkSector& S = args.sec[s];
    S.b = op->basis->view();
    S.off = op->d_off;
    S.dg = op->d_diag;
    S.n_off = op->n_off;
    S.n_diag = op->n_diag;
    S.coeffs = d_coeffs ? d_coeffs[s] : nullptr;
    S.tile_begin = tiles;
    S.row_off = rows;
    tiles += (op->basis->dim + kTileRows - 1) / kTileRows;
    rows += op->basis->dim;
  }
  args.n_tiles = tiles;
  // persistent grid: enough wavefronts to fill the chip, never more than tiles
  const int64_t want_waves = std::min<int64_t>(tiles, kBlkMaxWaves);
  const int n_blocks = static_cast<int>((want_waves + kBlkWaves - 1) / kBlkWaves);
  const int64_t n_waves = static_cast<int64_t>(n_blocks) * kBlkWaves;
  const size_t blk = static_cast<size_t>(rows) * R;
  DevMem mem;
  double *A = nullptr, *B = nullptr, *C = nullptr, *part = nullptr;
  double *alpha = nullptr, *beta = nullptr, *invb = nullptr, *ha = nullptr, *hb = nullptr;
  int* step = nullptr;
  const int SR = n_sectors * R;
  QSI_CHECK_HIP(mem.alloc(&A, blk));
  QSI_CHECK_HIP(mem.alloc(&B, blk));
  QSI_CHECK_HIP(mem.alloc(&C, blk));
  QSI_CHECK_HIP(mem.alloc(&part, static_cast<size_t>(n_sectors) * n_waves * R));
  QSI_CHECK_HIP(mem.alloc(&alpha, SR));
  QSI_CHECK_HIP(mem.alloc(&beta, SR));
  QSI_CHECK_HIP(mem.alloc(&invb, SR));
  QSI_CHECK_HIP(mem.alloc(&ha, static_cast<size_t>(SR) * M));
  QSI_CHECK_HIP(mem.alloc(&hb, static_cast<size_t>(SR) * M));
  QSI_CHECK_HIP(mem.alloc(&step, 1));
  hipStream_t st = nullptr;
  QSI_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  for (int s = 0; s < n_sectors; ++s) {
    QSI_CHECK_HIP(hipMemcpyAsync(B + args.sec[s].row_off * R, d_V0[s],
                                 static_cast<size_t>(ops[s]->basis->dim) * R * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
  }
  QSI_CHECK_HIP(hipMemsetAsync(A, 0, blk * sizeof(double), st));
  QSI_CHECK_HIP(hipMemsetAsync(beta, 0, SR * sizeof(double), st));
  QSI_CHECK_HIP(hipMemsetAsync(step, 0, sizeof(int), st));
  // normalise the start columns: B = B / |B_c|
  hipLaunchKernelGGL(k_blk_update, dim3(n_blocks), dim3(kBlkThreads), 0, st, args,
                     static_cast<const double*>(nullptr), B, B, static_cast<const double*>(nullptr),
                     static_cast<const double*>(nullptr), part);
  QSI_CHECK_LAUNCH();
  // (C = B - 0*B: the kernel writes back B unchanged and produces |B_c|^2)
  hipLaunchKernelGGL(k_blk_finalize, dim3(SR), dim3(256), 0, st, part, n_waves, R, 1,
                     beta, invb, static_cast<double*>(nullptr), static_cast<const int*>(nullptr), M);
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_blk_scale, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, B, invb,
                     static_cast<int*>(nullptr));
  QSI_CHECK_LAUNCH();
  QSI_CHECK_HIP(hipMemsetAsync(beta, 0, SR * sizeof(double), st));

  auto enqueue_step = [&](double* a, double* b, double* c) -> int {
    if (nw == 1)
      hipLaunchKernelGGL(k_blk_apply<1>, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, b, c, part);
    else
      hipLaunchKernelGGL(k_blk_apply<2>, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, b, c, part);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_finalize, dim3(SR), dim3(256), 0, st, part, n_waves, R, 0,
                       alpha, static_cast<double*>(nullptr), ha, step, M);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_update, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, a, b, c,
                       alpha, beta, part);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_finalize, dim3(SR), dim3(256), 0, st, part, n_waves, R, 1,
                       beta, invb, hb, step, M);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_scale, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, c, invb, step);
    QSI_CHECK_LAUNCH();
    return QSI_OK;
  };

  double* bufs[3] = {A, B, C};
  hipEvent_t e0, e1;
  QSI_CHECK_HIP(hipEventCreate(&e0));
  QSI_CHECK_HIP(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } eg{e0, e1};
  if (use_graph) {
    // one graph per rotation phase p: (A, B, C) = (bufs[p], bufs[p+1], bufs[p+2])
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    struct ExecGuard {
      hipGraphExec_t* e;
      ~ExecGuard() {
        for (int k = 0; k < 3; ++k)
          if (e[k]) (void)hipGraphExecDestroy(e[k]);
      }
    } xg{exec};
    for (int p = 0; p < 3 && p < M; ++p) {
      hipGraph_t g = nullptr;
      QSI_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_step(bufs[p % 3], bufs[(p + 1) % 3], bufs[(p + 2) % 3]);
      hipGraph_t g2 = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &g2);
      g = g2;
      if (rc != QSI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      QSI_CHECK_HIP(ec);
      const hipError_t ei = hipGraphInstantiate(&exec[p], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      QSI_CHECK_HIP(ei);
    }
    QSI_CHECK_HIP(hipEventRecord(e0, st));
    for (int j = 0; j < M; ++j) QSI_CHECK_HIP(hipGraphLaunch(exec[j % 3], st));
    QSI_CHECK_HIP(hipEventRecord(e1, st));
    QSI_CHECK_HIP(hipStreamSynchronize(st));
  } else {
    QSI_CHECK_HIP(hipEventRecord(e0, st));
    for (int j = 0; j < M; ++j) {
      const int p = j % 3;
      const int rc = enqueue_step(bufs[p], bufs[(p + 1) % 3], bufs[(p + 2) % 3]);
      if (rc != QSI_OK) return rc;
    }
    QSI_CHECK_HIP(hipEventRecord(e1, st));
    QSI_CHECK_HIP(hipStreamSynchronize(st));
  }
  if (elapsed_ms) {
    float ms = 0.f;
    QSI_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    *elapsed_ms = ms;
  }
  QSI_CHECK_HIP(hipMemcpy(h_alpha, ha, static_cast<size_t>(SR) * M * sizeof(double),
                          hipMemcpyDeviceToHost));
  QSI_CHECK_HIP(hipMemcpy(h_beta, hb, static_cast<size_t>(SR) * M * sizeof(double),
                          hipMemcpyDeviceToHost));
  return QSI_OK;
}

// Eigenvalues and first eigenvector components of one Lanczos tridiagonal
// (Gauss quadrature nodes and weights for FTLM).
extern "C" int qsi_tridiag_eig(const double* d, const double* e, int n,
                               double* evals, double* first_row) {
  QSI_REQUIRE(d && evals && n >= 1 && (n == 1 || e), "qsi_tridiag_eig: bad arguments");
  std::vector<double> dv(d, d + n), ev(e ? e : d, e ? e + std::max(n - 1, 0) : d), w, fr;
  if (!hostla::tridiag_eig(dv, ev, w, nullptr, first_row ? &fr : nullptr)) {
    set_error("qsi_tridiag_eig: no convergence");
    return QSI_E_NOCONV;
  }
  std::copy(w.begin(), w.end(), evals);
  if (first_row) std::copy(fr.begin(), fr.end(), first_row);
  return QSI_OK;
}
