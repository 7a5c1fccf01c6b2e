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

#include "qsi_device.hpp"
#include "qsi_hostla.hpp"
#include "qsi_internal.hpp"
#include "qsi_reduce.hpp"

namespace qsi {
namespace {

constexpr int kBlkThreads = 256;
constexpr int kBlkWaves = kBlkThreads / 64;
constexpr int kTileRows = 64;
constexpr int kMaxSectors = 8;

struct BlkSector {
  BasisView b;
  const OffTerm* off;
  const DiagTerm* dg;
  const double* coeffs;  // n_slots x R, or null
  int n_off;
  int n_diag;
  int64_t tile_begin;  // first global tile of this sector
  int64_t row_off;     // first row of this sector in the concatenated blocks
};

struct BlkArgs {
  BlkSector sec[kMaxSectors];
  int n_sec;
  int64_t n_tiles;
  int R;
  int G;  // lanes per row group (power of two >= R)
};

__device__ __forceinline__ int find_sector(const BlkArgs& a, int64_t tile) {
  int s = 0;
  for (int k = 1; k < a.n_sec; ++k)
    if (tile >= a.sec[k].tile_begin) s = k;
  return s;
}

// Sum of v over the lanes that share this lane's column (lanes c, c+G, ...).
__device__ __forceinline__ double sum_over_groups(double v, int G) {
  for (int o = G; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <int NW>
__global__ __launch_bounds__(kBlkThreads) void k_blk_apply(
    BlkArgs a, const double* __restrict__ X, double* __restrict__ Y,
    double* __restrict__ partials) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, R = a.R;
  const int grp = lane / G, col = lane % G;
  const bool col_ok = col < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  const int rows_per_pass = 64 / G;
  double dot[kMaxSectors];
#pragma unroll
  for (int k = 0; k < kMaxSectors; ++k) dot[k] = 0.0;
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = find_sector(a, t);
    const BlkSector& S = a.sec[s];
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int64_t r_end = min<int64_t>(r_begin + kTileRows, S.b.dim);
    for (int64_t r0 = r_begin; r0 < r_end; r0 += rows_per_pass) {
      const int64_t i = r0 + grp;
      const bool row_ok = i < r_end;
      const State<NW> st = load_state_cached<NW>(S.b.states, row_ok ? i : r_begin);
      double acc = 0.0;
      for (int t0 = 0; t0 < S.n_off; t0 += G) {
        const int tt = t0 + col;
        const bool ok = row_ok && tt < S.n_off && term_ok<NW>(S.off[tt], st);
        const uint64_t bal = __ballot(ok);
        const uint64_t mine = (G == 64) ? bal : ((bal >> (grp * G)) & ((1ull << G) - 1));
        // the group's connected terms, in term order: flip j is term t0 + (bit j of mine)
        const int nflip = __popcll(mine);
        int maxflip = nflip;
        for (int o = G; o < 64; o <<= 1) maxflip = max(maxflip, __shfl_xor(maxflip, o, 64));
        // lane `col` of the group looks up flip `col` (all lookups in flight together)
        int my_term = -1;
        int64_t my_J = -1;
        if (col < nflip) {
          uint64_t m = mine;
          for (int q = 0; q < col; ++q) m &= m - 1;
          my_term = t0 + __builtin_ctzll(m);
          my_J = basis_lookup<NW>(S.b, xor_state<NW>(st, S.off[my_term].X));
        }
        // every flip: broadcast its rank, then one coalesced read of X[J, 0:R]
        for (int j = 0; j < maxflip; ++j) {
          const int src = grp * G + j;
          const int64_t J = __shfl(my_J, src, 64);
          const int k = __shfl(my_term, src, 64);
          if (j < nflip && J >= 0 && col_ok) {
            const OffTerm& T = S.off[k];
            const double xj = X[(S.row_off + J) * R + col];
            if (fabs(xj) > kApplyTol) {
              double c = (S.coeffs && T.slot >= 0) ? S.coeffs[T.slot * R + col] : T.c;
              c *= T.sgn;
              if (parity_notz<NW>(st, T.Z)) c = -c;
              acc += c * xj;
            }
          }
        }
      }
      if (row_ok && col_ok) {
        const int64_t gi = (S.row_off + i) * R + col;
        const double xi = X[gi];
        for (int k = 0; k < S.n_diag; ++k) {
          const DiagTerm& d = S.dg[k];
          double f;
          if (d.kind == 0) {
            f = diag_ok<NW>(d, st) ? d.sgn * (1 - 2 * parity_notz<NW>(st, d.Z)) : 0.0;
          } else {
            int nz = __popcll(d.Z[0]);
            if constexpr (NW == 2) nz += __popcll(d.Z[1]);
            f = static_cast<double>(2 * popc_and<NW>(st, d.Z) - nz);
          }
          const double c = (S.coeffs && d.slot >= 0) ? S.coeffs[d.slot * R + col] : d.c;
          acc += c * xi * f;
        }
        Y[gi] = acc;
        for (int k = 0; k < kMaxSectors; ++k)
          if (k == s) dot[k] += xi * acc;
      }
    }
  }
  for (int k = 0; k < a.n_sec; ++k) {
    const double v = sum_over_groups(dot[k], G);
    if (grp == 0 && col_ok) partials[(k * n_waves + wave) * R + col] = v;
  }
}

// C -= alpha B + beta A (beta, A skipped when A is null); |C_c|^2 partials.
__global__ __launch_bounds__(kBlkThreads) void k_blk_update(
    BlkArgs a, const double* __restrict__ A, const double* B,
    double* C, const double* __restrict__ alpha,
    const double* __restrict__ beta, double* __restrict__ partials) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, R = a.R;
  const int grp = lane / G, col = lane % G;
  const bool col_ok = col < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  const int rows_per_pass = 64 / G;
  double nrm[kMaxSectors];
#pragma unroll
  for (int k = 0; k < kMaxSectors; ++k) nrm[k] = 0.0;
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = find_sector(a, t);
    const BlkSector& S = a.sec[s];
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int64_t r_end = min<int64_t>(r_begin + kTileRows, S.b.dim);
    const double al = col_ok && alpha ? alpha[s * R + col] : 0.0;
    const double be = col_ok && A ? beta[s * R + col] : 0.0;
    double part = 0.0;
    for (int64_t i = r_begin + grp; i < r_end; i += rows_per_pass) {
      if (!col_ok) continue;
      const int64_t gi = (S.row_off + i) * R + col;
      double w = C[gi] - al * B[gi];
      if (A) w -= be * A[gi];
      C[gi] = w;
      part += w * w;
    }
    for (int k = 0; k < kMaxSectors; ++k)
      if (k == s) nrm[k] += part;
  }
  for (int k = 0; k < a.n_sec; ++k) {
    const double v = sum_over_groups(nrm[k], G);
    if (grp == 0 && col_ok) partials[(k * n_waves + wave) * R + col] = v;
  }
}

// One workgroup per (sector, column): fixed-order sum of the wave partials.
// mode 0 (alpha): out[s*R+c] = sum.  mode 1 (beta): out[s*R+c] = sqrt(sum),
// out2[s*R+c] = 1/beta or 0.  hist (optional) receives the value at
// hist[(s*R + c)*M + step[0]].
__global__ __launch_bounds__(256) void k_blk_finalize(
    const double* __restrict__ partials, int64_t n_waves, int R, int mode,
    double* __restrict__ out, double* __restrict__ out2, double* __restrict__ hist,
    const int* __restrict__ step, int M) {
  const int sc = blockIdx.x;  // s * R + c
  const int s = sc / R, c = sc % R;
  double v = 0.0;
  for (int64_t w = threadIdx.x; w < n_waves; w += 256)
    v += partials[(s * n_waves + w) * R + c];
  v = block_sum<256>(v);
  if (threadIdx.x == 0) {
    double r = v;
    if (mode == 1) {
      r = sqrt(v);
      out2[sc] = r > 1e-300 ? 1.0 / r : 0.0;
    }
    out[sc] = r;
    if (hist) hist[static_cast<int64_t>(sc) * M + step[0]] = r;
  }
}

// C *= 1/beta per (sector, column); optionally advances the step counter.
__global__ __launch_bounds__(kBlkThreads) void k_blk_scale(
    BlkArgs a, double* __restrict__ C, const double* __restrict__ invb,
    int* __restrict__ step) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, R = a.R;
  const int grp = lane / G, col = lane % G;
  const bool col_ok = col < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  const int rows_per_pass = 64 / G;
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = find_sector(a, t);
    const BlkSector& S = a.sec[s];
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int64_t r_end = min<int64_t>(r_begin + kTileRows, S.b.dim);
    if (!col_ok) continue;
    const double f = invb[s * R + col];
    for (int64_t i = r_begin + grp; i < r_end; i += rows_per_pass)
      C[(S.row_off + i) * R + col] *= f;
  }
  if (step && blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1;
}

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
  const int64_t want_waves = std::min<int64_t>(tiles, 256 * 32);
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
