// Continuous-batching eigensolver engine: many independent lowest-eigenvalue
// requests on one device-resident basis share every H·v launch.
//
// Each request is the thick-restart Lanczos solve of qsi_eigsh (the
// counterpart of compute_spectrum_iterative -> Spectra::SymEigsSolver,
// include/matrix_diag_bits.hpp:51-95) with its own coefficients (its own g,
// B) held in one column of the engine.  One engine step advances every active
// column by one Lanczos step:
//   gather    X[:, c] = V_c[:, j_c]                (D x R row-major block)
//   apply     W = H_c X  (one block launch, per-column coefficients)
//   h_c = V_c[:, 0..j_c]^T w_c ; u_c = w_c - V_c h_c ; h2_c = V_c^T u_c
//   if |h2_c| > 1e-12 |u_c|: u_c -= V_c h2_c (second Gram-Schmidt pass)
//   beta_c = |u_c| ; V_c[:, j_c+1] = u_c / beta_c
// Active columns take apply slots 0..n_act in column order, so a step's
// launches cost what its active columns need, not what the engine can hold.
// Basis vectors are stored unnormalised with a scale per vector (vs): the
// Lanczos vector v_i is vs_i V_c[:, i], and every kernel that reads V_c
// applies the scales, so no pass rewrites a vector just to normalise it.
// then the host updates each column's projected matrix and, when a column's
// basis is full, runs its Rayleigh-Ritz step, convergence test and restart.
// Requests join free columns between steps and leave when converged.  Every
// per-column computation uses the same row tiling and summation order
// whatever the other columns hold, so a request's result does not depend on
// the mix it ran in.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <deque>
#include <limits>
#include <new>
#include <vector>

#include "qsi_blockk.hpp"
#include "qsi_hostla.hpp"

namespace qsi {
namespace {

constexpr int kSrvThreads = 256;
// rows per workgroup of the reorthogonalisation kernels: D / 256 rounded up
// to a multiple of 64, between 256 and 2048 (small bases still get many
// workgroups; the grid stays a few hundred workgroups per column)
constexpr int kSrvRowsMin = 256;
constexpr int kSrvRowsMax = 2048;
constexpr int kHMax = 64;
// partials slot holding |u|^2 (projections use slots 0..ncv-1 <= 62)
constexpr int kNormSlot = kHMax - 1;
// the second Gram-Schmidt pass runs when |h2|^2 > kReorthTol2 |u|^2
constexpr double kReorthTol2 = 1e-24;

// Start of a step: for every apply slot sl < n_act (column c = slotcol[sl]),
// the Lanczos vector vs_j V_c[j_c] into the row-major apply input
// X[row * ldx + sl] (ldx >= n_act: rows padded to an aligned size; the
// padding is not written).  Workgroups stride over 64-row tiles; a tile goes
// through LDS so that both the column reads and the row writes coalesce.
__global__ __launch_bounds__(kSrvThreads) void k_srv_prep(
    const double* __restrict__ Vb, int64_t ld, int ncv1, const int* __restrict__ colj,
    const double* __restrict__ vs, const int* __restrict__ slotcol, int n_act, int ldx,
    int64_t D, double* __restrict__ X) {
  // tile[row][sl] with row stride n_act + 1 (dynamic LDS sized to the step's
  // columns, so more workgroups fit on a CU)
  extern __shared__ double tile[];
  const int ts = n_act + 1;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int64_t n_tiles = (D + 63) / 64;
  for (int64_t tl = blockIdx.x; tl < n_tiles; tl += gridDim.x) {
    const int64_t r0 = tl * 64;
    const int64_t r = r0 + lane;
    for (int sl = wid; sl < n_act; sl += kSrvThreads / 64) {
      const int c = slotcol[sl];
      const int j = colj[c];
      double v = 0.0;
      if (r < D) v = Vb[(static_cast<int64_t>(c) * ncv1 + j) * ld + r] * vs[c * kHMax + j];
      tile[lane * ts + sl] = v;
    }
    __syncthreads();
    const int64_t rows = min<int64_t>(64, D - r0);
    for (int t = threadIdx.x; t < rows * n_act; t += kSrvThreads)
      X[(r0 + t / n_act) * ldx + t % n_act] = tile[(t / n_act) * ts + t % n_act];
    __syncthreads();
  }
}

// Per-block partial dot products of column c's w with its basis vectors
// V_c[0..m), m = j_c + 1 <= MMAX, reduced over the block's rows into
// partials[(c*kHMax + i)*nblk + b] (per projection, contiguous over blocks).
template <int MMAX>
__device__ __forceinline__ void srv_block_dots(double (&acc)[MMAX], int m, int c, int nblk,
                                               double* __restrict__ partials) {
  __shared__ double red[kSrvThreads / 64][MMAX];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int i = 0; i < MMAX; ++i) {
    if (i < m) {
      const double v = wave_sum(acc[i]);
      if (lane == 0) red[wid][i] = v;
    }
  }
  __syncthreads();
  if (threadIdx.x < m) {
    double v = 0.0;
    for (int k = 0; k < kSrvThreads / 64; ++k) v += red[k][threadIdx.x];
    partials[(static_cast<int64_t>(c) * kHMax + threadIdx.x) * nblk + blockIdx.x] = v;
  }
}

// pass 1: partials of V_c^T w_c.  Grid (nblk, n_act); w_c is apply slot
// blockIdx.y.
template <int MMAX>
__global__ __launch_bounds__(kSrvThreads) void k_srv_mdot(
    const double* __restrict__ Vb, int64_t ld, int ncv1, const double* __restrict__ Wc,
    const int* __restrict__ colj, const int* __restrict__ slotcol, int64_t D, int rpb,
    double* __restrict__ partials) {
  const int sl = blockIdx.y;
  const int c = slotcol[sl];
  const int j = colj[c];
  const int m = j + 1;
  const double* V = Vb + static_cast<int64_t>(c) * ncv1 * ld;
  const double* w = Wc + static_cast<int64_t>(sl) * ld;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rpb;
  const int64_t r1 = min<int64_t>(r0 + rpb, D);
  double acc[MMAX];
#pragma unroll
  for (int i = 0; i < MMAX; ++i) acc[i] = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kSrvThreads) {
    const double wr = w[r];
    if constexpr (MMAX <= 32) {
      // all loads first, then unpredicated updates (entries i >= m are zero)
      double v[MMAX];
#pragma unroll
      for (int i = 0; i < MMAX; ++i) v[i] = i < m ? V[i * ld + r] : 0.0;
#pragma unroll
      for (int i = 0; i < MMAX; ++i) acc[i] += v[i] * wr;
    } else {
#pragma unroll
      for (int i = 0; i < MMAX; ++i)
        if (i < m) acc[i] += V[i * ld + r] * wr;
    }
  }
  srv_block_dots<MMAX>(acc, m, c, gridDim.x, partials);
}

// pass 1 -> 2: u_c = w_c - V_c h_c (h from pass 1) written to V_c[j_c + 1],
// then partials of V_c^T u_c (slots 0..j_c) and of |u_c|^2 (kNormSlot).
// Grid (nblk, n_act).  The block walks its rows in 64-row tiles; wave k holds
// basis vectors i = k, k + 4, ... in registers, the four waves' parts of
// (V h)[row] meet in LDS, and each wave reuses its V entries for the dots.
template <int MMAX>
__global__ __launch_bounds__(kSrvThreads) void k_srv_gemv_mdot(
    double* __restrict__ Vb, int64_t ld, int ncv1, const double* __restrict__ h,
    const double* __restrict__ vs, const int* __restrict__ colj, const int* __restrict__ slotcol,
    int64_t D, int rpb, const double* __restrict__ Wc, double* __restrict__ partials) {
  constexpr int NWV = kSrvThreads / 64;
  constexpr int Q = (MMAX + NWV - 1) / NWV;
  const int sl = blockIdx.y;
  const int c = slotcol[sl];
  const int j = colj[c];
  const int m = j + 1;
  __shared__ double hs[MMAX];
  __shared__ double ared[NWV][64];
  __shared__ double red[MMAX];
  __shared__ double nsh;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (threadIdx.x < m) hs[threadIdx.x] = h[c * kHMax + threadIdx.x] * vs[c * kHMax + threadIdx.x];
  __syncthreads();
  const double* V = Vb + static_cast<int64_t>(c) * ncv1 * ld;
  double* out = Vb + (static_cast<int64_t>(c) * ncv1 + m) * ld;
  const double* w = Wc + static_cast<int64_t>(sl) * ld;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rpb;
  const int64_t r1 = min<int64_t>(r0 + rpb, D);
  double acc[Q];
#pragma unroll
  for (int q = 0; q < Q; ++q) acc[q] = 0.0;
  double nrm = 0.0;
  for (int64_t rt = r0; rt < r1; rt += 64) {
    const int64_t r = rt + lane;
    const bool ok = r < r1;
    double v[Q];
    double a = 0.0;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int i = wid + NWV * q;
      v[q] = (ok && i < m) ? V[i * ld + r] : 0.0;
      if (i < m) a += v[q] * hs[i];
    }
    ared[wid][lane] = a;
    __syncthreads();
    double ur = 0.0;
    if (ok) {
      double t = 0.0;
#pragma unroll
      for (int k = 0; k < NWV; ++k) t += ared[k][lane];
      ur = w[r] - t;
      if (wid == 0) {
        out[r] = ur;
        nrm += ur * ur;
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) acc[q] += v[q] * ur;
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < Q; ++q) {
    const int i = wid + NWV * q;
    const double t = wave_sum(acc[q]);
    if (lane == 0 && i < m) red[i] = t;
  }
  if (wid == 0) {
    const double t = wave_sum(nrm);
    if (lane == 0) nsh = t;
  }
  __syncthreads();
  const int64_t nb = gridDim.x;
  double* pp = partials + static_cast<int64_t>(c) * kHMax * nb + blockIdx.x;
  if (threadIdx.x < m) pp[threadIdx.x * nb] = red[threadIdx.x];
  if (threadIdx.x == 0) pp[kNormSlot * nb] = nsh;
}

// Pass 1 -> 2 as k_srv_gemv_mdot for MMAX <= 24: one thread per row keeps
// the row's m basis entries in registers for both the update and the dots
// (no workgroup barrier per row tile).
template <int MMAX>
__global__ __launch_bounds__(kSrvThreads) void k_srv_gemv_mdot_rows(
    double* __restrict__ Vb, int64_t ld, int ncv1, const double* __restrict__ h,
    const double* __restrict__ vs, const int* __restrict__ colj, const int* __restrict__ slotcol,
    int64_t D, int rpb, const double* __restrict__ Wc, double* __restrict__ partials) {
  const int sl = blockIdx.y;
  const int c = slotcol[sl];
  const int m = colj[c] + 1;
  // h in registers (uniform); entries i >= m are zero, so the loops below
  // run unpredicated
  double hr[MMAX];
#pragma unroll
  for (int i = 0; i < MMAX; ++i) hr[i] = i < m ? h[c * kHMax + i] * vs[c * kHMax + i] : 0.0;
  const double* V = Vb + static_cast<int64_t>(c) * ncv1 * ld;
  double* out = Vb + (static_cast<int64_t>(c) * ncv1 + m) * ld;
  const double* w = Wc + static_cast<int64_t>(sl) * ld;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rpb;
  const int64_t r1 = min<int64_t>(r0 + rpb, D);
  double acc[MMAX];
#pragma unroll
  for (int i = 0; i < MMAX; ++i) acc[i] = 0.0;
  double nrm = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kSrvThreads) {
    double v[MMAX];
#pragma unroll
    for (int i = 0; i < MMAX; ++i) v[i] = i < m ? V[i * ld + r] : 0.0;
    const double wr = w[r];
    double a = 0.0;
#pragma unroll
    for (int i = 0; i < MMAX; ++i) a += v[i] * hr[i];
    const double ur = wr - a;
    out[r] = ur;
    nrm += ur * ur;
#pragma unroll
    for (int i = 0; i < MMAX; ++i) acc[i] += v[i] * ur;
  }
  srv_block_dots<MMAX>(acc, m, c, gridDim.x, partials);
  nrm = block_sum<kSrvThreads>(nrm);
  if (threadIdx.x == 0)
    partials[(static_cast<int64_t>(c) * kHMax + kNormSlot) * gridDim.x + blockIdx.x] = nrm;
}

// Thick restart of one column, in place: row by row, V[0..kn) = v[0..m) Y
// (Y column-major m x kn; v_i = vsc_i V[i]) and V[kn] = V[m] (the residual
// direction, unscaled).  Every thread reads its row's m + 1 entries before
// writing any of them.
template <int MMAX>
__global__ __launch_bounds__(kSrvThreads) void k_srv_restart(double* __restrict__ V, int64_t ld,
                                                              int m, int kn,
                                                              const double* __restrict__ Y,
                                                              const double* __restrict__ vsc,
                                                              int64_t D) {
  __shared__ double ys[MMAX * MMAX];
  for (int t = threadIdx.x; t < m * kn; t += kSrvThreads) ys[t] = Y[t] * vsc[t % m];
  __syncthreads();
  for (int64_t r = static_cast<int64_t>(blockIdx.x) * kSrvThreads + threadIdx.x; r < D;
       r += static_cast<int64_t>(gridDim.x) * kSrvThreads) {
    double v[MMAX + 1];
#pragma unroll
    for (int i = 0; i <= MMAX; ++i) v[i] = i <= m ? V[i * ld + r] : 0.0;
    for (int k = 0; k < kn; ++k) {
      double a = 0.0;
#pragma unroll
      for (int i = 0; i < MMAX; ++i)
        if (i < m) a += v[i] * ys[k * m + i];
      V[k * ld + r] = a;
    }
    double res = 0.0;
#pragma unroll
    for (int i = 0; i <= MMAX; ++i)
      if (i == m) res = v[i];
    V[kn * ld + r] = res;
  }
}

// After pass 2's projections h2_c are summed: beta_c = |u_c| from the
// kNormSlot partials, the scale 1/beta_c of the new vector u_c, and flag_c = 1
// when |h2_c|^2 > kReorthTol2 |u_c|^2 (u_c needs the second Gram-Schmidt
// pass).  Grid R, one wave each.
__global__ void k_srv_check(const double* __restrict__ partials, int nblk,
                            const int* __restrict__ colj, const double* __restrict__ h,
                            int* __restrict__ flag, double* __restrict__ beta,
                            double* __restrict__ vs) {
  const int c = blockIdx.x;
  const int j = colj[c];
  if (j < 0) return;
  double n2 = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64)
    n2 += partials[(static_cast<int64_t>(c) * kHMax + kNormSlot) * nblk + b];
  n2 = wave_sum(n2);
  double s2 = 0.0;
  for (int i = threadIdx.x; i <= j; i += 64) {
    const double v = h[c * kHMax + i];
    s2 += v * v;
  }
  s2 = wave_sum(s2);
  if (threadIdx.x == 0) {
    flag[c] = s2 > kReorthTol2 * n2 ? 1 : 0;
    const double r = sqrt(n2);
    beta[c] = r;
    vs[c * kHMax + j + 1] = r > 1e-300 ? 1.0 / r : 0.0;
  }
}

// pass 3, flagged columns only: u_c -= V_c h2_c in place in V_c[j_c + 1],
// partials of |u_c|^2 in partials[c*nblk + b].  Grid (nblk, n_act).
template <int MMAX>
__global__ __launch_bounds__(kSrvThreads) void k_srv_gemv_norm(
    double* __restrict__ Vb, int64_t ld, int ncv1, const double* __restrict__ h,
    const double* __restrict__ vs, const int* __restrict__ colj,
    const int* __restrict__ slotcol, const int* __restrict__ flag, int64_t D, int rpb,
    double* __restrict__ partials) {
  const int c = slotcol[blockIdx.y];
  if (!flag[c]) return;
  const int m = colj[c] + 1;
  __shared__ double hs[MMAX];
  if (threadIdx.x < MMAX)
    hs[threadIdx.x] =
        threadIdx.x < m ? h[c * kHMax + threadIdx.x] * vs[c * kHMax + threadIdx.x] : 0.0;
  __syncthreads();
  const double* V = Vb + static_cast<int64_t>(c) * ncv1 * ld;
  double* u = Vb + (static_cast<int64_t>(c) * ncv1 + m) * ld;
  const int64_t r0 = static_cast<int64_t>(blockIdx.x) * rpb;
  const int64_t r1 = min<int64_t>(r0 + rpb, D);
  double nrm = 0.0;
  for (int64_t r = r0 + threadIdx.x; r < r1; r += kSrvThreads) {
    double a = 0.0;
    if constexpr (MMAX <= 32) {
      // all loads first, then the sum in column order (absent columns load
      // as 0 against a zero h), so a lane keeps its column loads in flight
      double v[MMAX];
#pragma unroll
      for (int i = 0; i < MMAX; ++i) v[i] = i < m ? V[i * ld + r] : 0.0;
#pragma unroll
      for (int i = 0; i < MMAX; ++i) a += v[i] * hs[i];
    } else {
#pragma unroll
      for (int i = 0; i < MMAX; ++i)
        if (i < m) a += V[i * ld + r] * hs[i];
    }
    const double ur = u[r] - a;
    u[r] = ur;
    nrm += ur * ur;
  }
  nrm = block_sum<kSrvThreads>(nrm);
  if (threadIdx.x == 0) partials[static_cast<int64_t>(c) * gridDim.x + blockIdx.x] = nrm;
}

// h[c*kHMax + i] = vs_i * sum_b partials (fixed order); acc[c*kHMax + i] = h
// (first pass, add = 0) or += h.  Grid (R, kHMax), one workgroup each.
__global__ __launch_bounds__(kSrvThreads) void k_srv_fin(const double* __restrict__ partials,
                                                         int nblk, const int* __restrict__ colj,
                                                         const double* __restrict__ vs,
                                                         double* __restrict__ h,
                                                         double* __restrict__ acc, int add) {
  const int c = blockIdx.x, i = blockIdx.y;
  const int j = colj[c];
  if (j < 0 || i > j) return;
  const double* p = partials + (static_cast<int64_t>(c) * kHMax + i) * nblk;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblk; b += kSrvThreads) v += p[b];
  v = block_sum<kSrvThreads>(v);
  if (threadIdx.x == 0) {
    v *= vs[c * kHMax + i];
    h[c * kHMax + i] = v;
    acc[c * kHMax + i] = add ? acc[c * kHMax + i] + v : v;
  }
}

// Flagged columns after pass 3: beta[c] = sqrt(sum) and the new vector's
// scale 1/beta (0 if it underflows).  Grid R.
__global__ void k_srv_beta(const double* __restrict__ partials, int nblk,
                           const int* __restrict__ colj, const int* __restrict__ flag,
                           double* __restrict__ beta, double* __restrict__ vs) {
  const int c = blockIdx.x;
  if (colj[c] < 0 || !flag[c]) return;
  double v = 0.0;
  for (int b = threadIdx.x; b < nblk; b += 64) v += partials[static_cast<int64_t>(c) * nblk + b];
  v = wave_sum(v);
  if (threadIdx.x == 0) {
    const double r = sqrt(v);
    beta[c] = r;
    vs[c * kHMax + colj[c] + 1] = r > 1e-300 ? 1.0 / r : 0.0;
  }
}

// The step's results for the host, written by a kernel straight into mapped
// pinned memory (no copy engine): hh = hacc (R x kHMax), bb = beta (R).
__global__ void k_srv_out(const double* __restrict__ hacc, const double* __restrict__ beta,
                          int R, double* hh, double* bb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < R * kHMax) hh[t] = hacc[t];
  if (t < R) bb[t] = beta[t];
}

// vs[c][0] = 1 / sqrt(*n2) (start vector normalisation).
__global__ void k_srv_setinv(const double* __restrict__ n2, int c, double* __restrict__ vs) {
  if (threadIdx.x == 0) vs[c * kHMax] = 1.0 / sqrt(*n2);
}

// After column c's restart: the kept Ritz vectors are normalised and the
// residual direction moved from index m to kn keeps its scale.
__global__ void k_srv_rescale(double* __restrict__ vs, int c, int m, int kn) {
  if (threadIdx.x == 0) {
    const double sm = vs[c * kHMax + m];
    for (int k = 0; k < kn; ++k) vs[c * kHMax + k] = 1.0;
    vs[c * kHMax + kn] = sm;
  }
}

}  // namespace
}  // namespace qsi

using namespace qsi;

struct qsi_engine_s {
  struct Request {
    int64_t id;
    uint64_t seed;
    std::vector<double> coeffs;  // n_slots
  };
  struct Result {
    int64_t id;
    std::vector<double> evals;
    int nconv;
    int steps;
  };
  struct Col {
    bool active = false;
    bool held = false;  // basis full: out of the steps until its restart
    int64_t req = -1;
    int j = 0;        // index of the next vector to expand
    int k = 0;        // kept Ritz vectors after the last restart
    int restarts = 0;
    int steps = 0;
    double beta_last = 0.0;
    std::vector<double> T;  // ncv x ncv column-major
  };
  qsi_op_s* op = nullptr;
  int R = 0, nev = 0, ncv = 0, ncv1 = 0, max_restarts = 0, mmax = 0;
  double tol = 0.0;
  int64_t D = 0, ld = 0;
  int nblk = 0;
  int rpb = 0;  // rows per workgroup of the reorthogonalisation kernels
  hipStream_t st = nullptr;
  // device: Krylov bases (column c: ncv + 1 vectors of stride ld), the
  // row-major apply input X (D x R), the apply output w (column-major),
  // slot coefficients, projections, norms and reduction partials
  double *Vb = nullptr, *X = nullptr, *Wc = nullptr, *coeffs = nullptr;
  double *h = nullptr, *hacc = nullptr, *beta = nullptr, *vs = nullptr, *part = nullptr;
  double* Ydev = nullptr;
  int *colj_d = nullptr, *slot_d = nullptr, *flag_d = nullptr;
  std::vector<void*> allocs;
  // pinned staging of the restarts' Ritz coefficients, kHMax^2 per column
  double* Ystage = nullptr;
  // Up to two steps are in flight; each has its own pinned exchange (column
  // state down, projections and norms up) and completion event.
  struct StepBuf {
    int* colj_p = nullptr;
    int* slot_p = nullptr;  // slotcol: apply slot -> column
    double* coeffs_p = nullptr;
    double* hh_p = nullptr;
    double* bb_p = nullptr;
    double* hh_dev = nullptr;  // device addresses of hh_p / bb_p (mapped)
    double* bb_dev = nullptr;
    hipEvent_t ev = nullptr;   // recorded after the step's last command
    std::vector<int> colj;     // j of each column in this step (-1: not in it)
    std::vector<int64_t> req;  // request of each column when it was enqueued
  };
  StepBuf sb[2];
  void* pin = nullptr;
  std::deque<int> inflight;  // StepBuf indices, oldest first
  int next_buf = 0;
  struct Restart {
    int c, m, kn;
  };
  std::vector<Restart> restarts;
  // host
  std::vector<Col> cols;
  std::vector<int> colj, slotcol;
  std::vector<double> coeffs_h;
  std::deque<Request> queue;
  std::deque<Result> done;
  int64_t next_id = 0;
  BlkArgs args{};
  int n_blocks = 0;

  ~qsi_engine_s() {
    if (st) (void)hipStreamSynchronize(st);
    for (void* p : allocs) (void)hipFree(p);
    if (Ystage) (void)hipHostFree(Ystage);
    if (pin) (void)hipHostFree(pin);
    for (auto& b : sb)
      if (b.ev) (void)hipEventDestroy(b.ev);
    if (st) (void)hipStreamDestroy(st);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n * sizeof(T), 8));
    if (e == hipSuccess) {
      allocs.push_back(q);
      *p = static_cast<T*>(q);
    }
    return e;
  }
};

namespace {

// Free columns take queued requests: coefficients (uploaded per step in slot
// order), a seeded Gaussian start vector and its scale 1/|v0|.
int engine_admit(qsi_engine_s* E) {
  for (int c = 0; c < E->R && !E->queue.empty(); ++c) {
    auto& col = E->cols[c];
    if (col.active) continue;
    qsi_engine_s::Request rq = std::move(E->queue.front());
    E->queue.pop_front();
    col = qsi_engine_s::Col();
    col.active = true;
    col.req = rq.id;
    col.T.assign(static_cast<size_t>(E->ncv) * E->ncv, 0.0);
    for (int s = 0; s < E->op->n_slots; ++s) E->coeffs_h[s * E->R + c] = rq.coeffs[s];
    double* v0 = E->Vb + static_cast<int64_t>(c) * E->ncv1 * E->ld;
    int rc = launch_randn(v0, E->D, rq.seed, E->st);
    if (rc != QSI_OK) return rc;
    rc = launch_dot(v0, v0, E->D, E->part + 8, E->part, E->st);
    if (rc != QSI_OK) return rc;
    hipLaunchKernelGGL(k_srv_setinv, dim3(1), dim3(64), 0, E->st, E->part, c, E->vs);
    QSI_CHECK_LAUNCH();
  }
  return QSI_OK;
}

// Rayleigh-Ritz, convergence test and thick restart of column c whose basis
// holds m vectors (V_c[m], the residual direction, still unnormalised).
int engine_ritz(qsi_engine_s* E, int c, int m, bool exhausted) {
  auto& col = E->cols[c];
  const int ncv = E->ncv;
  std::vector<double> Tm(static_cast<size_t>(m) * m), theta, Y;
  for (int j = 0; j < m; ++j)
    for (int i = 0; i < m; ++i)
      Tm[static_cast<size_t>(j) * m + i] = col.T[static_cast<size_t>(j) * ncv + i];
  hostla::sym_eig(Tm, m, theta, Y);
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
  const int want = std::min(E->nev, m);
  int nconv = 0;
  for (int i = 0; i < want; ++i) {
    const double resid = col.beta_last * std::fabs(Y[static_cast<size_t>(i) * m + (m - 1)]);
    if (resid < E->tol * std::max(eps23, std::fabs(theta[i]))) ++nconv;
    else break;
  }
  if (nconv >= want || exhausted || col.restarts >= E->max_restarts) {
    qsi_engine_s::Result r;
    r.id = col.req;
    r.evals.assign(theta.begin(), theta.begin() + want);
    r.evals.resize(E->nev, std::numeric_limits<double>::quiet_NaN());
    r.nconv = nconv;
    r.steps = col.steps;
    E->done.push_back(std::move(r));
    col.active = false;
    col.held = false;
    return QSI_OK;
  }
  // keep the kn lowest Ritz vectors, then the residual direction; the
  // device side runs after every column's host work (engine_restarts)
  int kn = E->nev + std::min(nconv, (m - E->nev) / 2);
  kn = std::max(1, std::min(kn, m - 1));
  std::copy(Y.begin(), Y.begin() + static_cast<size_t>(kn) * m,
            E->Ystage + static_cast<size_t>(c) * kHMax * kHMax);
  E->restarts.push_back({c, m, kn});
  std::fill(col.T.begin(), col.T.end(), 0.0);
  for (int i = 0; i < kn; ++i) col.T[static_cast<size_t>(i) * ncv + i] = theta[i];
  col.k = kn;
  col.j = kn;
  col.held = false;
  col.restarts += 1;
  return QSI_OK;
}

// V_c[0..kn) = V_c[0..m) Y_c and V_c[kn] = V_c[m] for the staged restarts.
// The staging buffer is rewritten only after the next step synchronises.
int engine_restarts(qsi_engine_s* E) {
  const size_t per = static_cast<size_t>(kHMax) * kHMax;
  const uint32_t grid = static_cast<uint32_t>(
      std::min<int64_t>((E->D + kSrvThreads - 1) / kSrvThreads, 4096));
  for (const auto& r : E->restarts) {
    double* V = E->Vb + static_cast<int64_t>(r.c) * E->ncv1 * E->ld;
    double* Yd = E->Ydev + r.c * per;
    const double* vsc = E->vs + r.c * kHMax;
    QSI_CHECK_HIP(hipMemcpyAsync(Yd, E->Ystage + r.c * per,
                                 static_cast<size_t>(r.kn) * r.m * sizeof(double),
                                 hipMemcpyHostToDevice, E->st));
    if (E->mmax == 16)
      hipLaunchKernelGGL(k_srv_restart<16>, dim3(grid), dim3(kSrvThreads), 0, E->st, V, E->ld,
                         r.m, r.kn, Yd, vsc, E->D);
    else if (E->mmax == 24)
      hipLaunchKernelGGL(k_srv_restart<24>, dim3(grid), dim3(kSrvThreads), 0, E->st, V, E->ld,
                         r.m, r.kn, Yd, vsc, E->D);
    else if (E->mmax == 32)
      hipLaunchKernelGGL(k_srv_restart<32>, dim3(grid), dim3(kSrvThreads), 0, E->st, V, E->ld,
                         r.m, r.kn, Yd, vsc, E->D);
    else
      hipLaunchKernelGGL(k_srv_restart<64>, dim3(grid), dim3(kSrvThreads), 0, E->st, V, E->ld,
                         r.m, r.kn, Yd, vsc, E->D);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_srv_rescale, dim3(1), dim3(64), 0, E->st, E->vs, r.c, r.m, r.kn);
    QSI_CHECK_LAUNCH();
  }
  E->restarts.clear();
  return QSI_OK;
}

template <int MMAX>
void launch_reorth(qsi_engine_s* E, dim3 gcol, hipStream_t s) {
  const int R = E->R;
  hipLaunchKernelGGL(k_srv_mdot<MMAX>, gcol, dim3(kSrvThreads), 0, s, E->Vb, E->ld, E->ncv1,
                     E->Wc, E->colj_d, E->slot_d, E->D, E->rpb, E->part);
  hipLaunchKernelGGL(k_srv_fin, dim3(R, kHMax), dim3(kSrvThreads), 0, s, E->part, E->nblk,
                     E->colj_d, E->vs, E->h, E->hacc, 0);
  if constexpr (MMAX <= 24)
    hipLaunchKernelGGL(k_srv_gemv_mdot_rows<MMAX>, gcol, dim3(kSrvThreads), 0, s, E->Vb, E->ld,
                       E->ncv1, E->h, E->vs, E->colj_d, E->slot_d, E->D, E->rpb, E->Wc,
                       E->part);
  else
    hipLaunchKernelGGL(k_srv_gemv_mdot<MMAX>, gcol, dim3(kSrvThreads), 0, s, E->Vb, E->ld,
                       E->ncv1, E->h, E->vs, E->colj_d, E->slot_d, E->D, E->rpb, E->Wc,
                       E->part);
  hipLaunchKernelGGL(k_srv_fin, dim3(R, kHMax), dim3(kSrvThreads), 0, s, E->part, E->nblk,
                     E->colj_d, E->vs, E->h, E->hacc, 1);
  hipLaunchKernelGGL(k_srv_check, dim3(R), dim3(64), 0, s, E->part, E->nblk, E->colj_d, E->h,
                     E->flag_d, E->beta, E->vs);
  hipLaunchKernelGGL(k_srv_gemv_norm<MMAX>, gcol, dim3(kSrvThreads), 0, s, E->Vb, E->ld,
                     E->ncv1, E->h, E->vs, E->colj_d, E->slot_d, E->flag_d, E->D, E->rpb,
                     E->part);
  hipLaunchKernelGGL(k_srv_beta, dim3(R), dim3(64), 0, s, E->part, E->nblk, E->colj_d,
                     E->flag_d, E->beta, E->vs);
}

}  // namespace

extern "C" {

int qsi_engine_create(qsi_op_t op, int max_cols, int nev, int ncv, double tol,
                      int max_restarts, qsi_engine_t* out) {
  QSI_REQUIRE(out != nullptr && op != nullptr, "qsi_engine_create: null argument");
  *out = nullptr;
  QSI_REQUIRE(max_cols >= 1 && max_cols <= 64, "qsi_engine_create: max_cols must be in [1, 64]");
  QSI_REQUIRE(nev >= 1 && ncv >= 2 && ncv <= 63, "qsi_engine_create: need nev >= 1, 2 <= ncv <= 63");
  QSI_REQUIRE(op->basis->dim > 0, "qsi_engine_create: empty basis");
  auto* E = new (std::nothrow) qsi_engine_s();
  if (!E) {
    set_error("qsi_engine_create: host allocation failed");
    return QSI_E_NOMEM;
  }
  E->op = op;
  E->D = op->basis->dim;
  E->ncv = static_cast<int>(std::min<int64_t>(ncv, E->D));
  E->nev = std::min(nev, std::max(E->ncv - 1, 1));
  E->ncv1 = E->ncv + 1;
  E->mmax = E->ncv <= 16 ? 16 : (E->ncv <= 24 ? 24 : (E->ncv <= 32 ? 32 : 64));
  E->R = max_cols;
  E->tol = tol;
  E->max_restarts = max_restarts;
  E->ld = (E->D + 63) / 64 * 64;
  {
    int64_t rpb = ((E->D + 255) / 256 + 63) / 64 * 64;
    rpb = std::min<int64_t>(std::max<int64_t>(rpb, kSrvRowsMin), kSrvRowsMax);
    E->rpb = static_cast<int>(rpb);
    E->nblk = static_cast<int>((E->D + rpb - 1) / rpb);
  }
  BlkArgs& a = E->args;
  a.n_sec = 1;
  a.R = E->R;
  int G = 64;
  while (G / 2 >= E->R && G > 1) G >>= 1;
  a.G = G;
  a.ldx = E->R;
  a.ldy = E->ld;
  a.accumulate = 0;
  a.y_colmajor = 1;
  a.sec[0].b = op->basis->view();
  a.sec[0].off = op->d_off;
  a.sec[0].dg = op->d_diag;
  a.sec[0].cls = op->d_cls;
  a.sec[0].n_cls = op->n_cls;
  a.sec[0].rowptr = op->d_rowptr;
  a.sec[0].cj = op->d_cj;
  a.sec[0].cm = op->d_cm;
  a.sec[0].n_off = op->n_off;
  a.sec[0].n_diag = op->n_diag;
  a.sec[0].tile_begin = 0;
  a.sec[0].row_off = 0;
  a.n_tiles = (E->D + kTileRows - 1) / kTileRows;
  a.cfg = op->blk_cfg;
  E->n_blocks = blk_apply_blocks(op->basis->nw, a.n_tiles, blk_cached(a), a.cfg);
  const size_t R = static_cast<size_t>(E->R);
  hipError_t e = hipStreamCreateWithFlags(&E->st, hipStreamNonBlocking);
  const int ns = std::max(op->n_slots, 1);
  if (e == hipSuccess) e = E->alloc(&E->Vb, R * E->ncv1 * E->ld);
  if (e == hipSuccess) e = E->alloc(&E->X, static_cast<size_t>(E->D) * R);
  if (e == hipSuccess) e = E->alloc(&E->Wc, R * E->ld);
  // per-step state sent down in one copy: coeffs (ns x R doubles), then colj,
  // slotcol (R ints each), the same layout as the pinned block
  double* down = nullptr;
  if (e == hipSuccess) e = E->alloc(&down, static_cast<size_t>(ns) * R + R);
  if (e == hipSuccess) {
    E->coeffs = down;
    E->colj_d = reinterpret_cast<int*>(down + static_cast<size_t>(ns) * R);
    E->slot_d = E->colj_d + R;
  }
  if (e == hipSuccess) e = E->alloc(&E->h, R * kHMax);
  if (e == hipSuccess) e = E->alloc(&E->hacc, R * kHMax);
  if (e == hipSuccess) e = E->alloc(&E->beta, R);
  if (e == hipSuccess) e = E->alloc(&E->vs, R * kHMax);
  if (e == hipSuccess)
    e = E->alloc(&E->part, std::max<size_t>(R * E->nblk * kHMax, static_cast<size_t>(max_vec_partials()) + 8));
  if (e == hipSuccess) e = E->alloc(&E->Ydev, R * kHMax * kHMax);
  if (e == hipSuccess)
    e = hipHostMalloc(reinterpret_cast<void**>(&E->Ystage), R * kHMax * kHMax * sizeof(double),
                      hipHostMallocDefault);
  // pinned: per step buffer hh (R x kHMax), bb (R), coeffs (ns x R), colj and
  // slotcol (R ints each, in R more doubles)
  const size_t per_buf = R * (kHMax + 2 + ns);
  if (e == hipSuccess)
    e = hipHostMalloc(&E->pin, 2 * per_buf * sizeof(double),
                      hipHostMallocMapped | hipHostMallocCoherent);
  void* dp = nullptr;
  if (e == hipSuccess) e = hipHostGetDevicePointer(&dp, E->pin, 0);
  for (int b = 0; b < 2 && e == hipSuccess; ++b) {
    auto& B = E->sb[b];
    B.hh_p = static_cast<double*>(E->pin) + b * per_buf;
    B.bb_p = B.hh_p + R * kHMax;
    B.coeffs_p = B.bb_p + R;
    B.colj_p = reinterpret_cast<int*>(B.coeffs_p + R * ns);
    B.slot_p = B.colj_p + R;
    B.hh_dev = static_cast<double*>(dp) + b * per_buf;
    B.bb_dev = B.hh_dev + R * kHMax;
    B.colj.assign(R, -1);
    B.req.assign(R, -1);
    e = hipEventCreateWithFlags(&B.ev, hipEventDisableTiming);
  }
  if (e == hipSuccess) e = E->alloc(&E->flag_d, R);
  // on the engine's own stream: it does not wait for the null stream, so a
  // null-stream memset could still be running under the first step
  if (e == hipSuccess) e = hipMemsetAsync(E->Vb, 0, R * E->ncv1 * E->ld * sizeof(double), E->st);
  // first touch of the step buffers here rather than in the first full step
  if (e == hipSuccess) e = hipMemsetAsync(E->X, 0, static_cast<size_t>(E->D) * R * sizeof(double), E->st);
  if (e == hipSuccess) e = hipMemsetAsync(E->Wc, 0, R * E->ld * sizeof(double), E->st);
  if (e == hipSuccess) e = hipMemsetAsync(E->vs, 0, R * kHMax * sizeof(double), E->st);
  if (e == hipSuccess) e = hipStreamSynchronize(E->st);
  if (e != hipSuccess) {
    set_error(std::string("qsi_engine_create: ") + hipGetErrorString(e));
    delete E;
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  }
  E->cols.assign(E->R, qsi_engine_s::Col());
  E->colj.assign(E->R, -1);
  E->slotcol.assign(E->R, 0);
  E->coeffs_h.assign(static_cast<size_t>(ns) * R, 0.0);
  E->args.sec[0].coeffs = op->n_slots > 0 ? E->coeffs : nullptr;
  *out = E;
  return QSI_OK;
}

int qsi_engine_destroy(qsi_engine_t E) {
  if (E) {
    (void)hipStreamSynchronize(E->st);
    delete E;
  }
  return QSI_OK;
}

// Bytes of device memory an engine with these parameters allocates (the
// Krylov bases, X and w; the small arrays aside).
int64_t qsi_engine_footprint(int64_t dim, int max_cols, int ncv) {
  const int64_t ld = (dim + 63) / 64 * 64;
  const int64_t R = max_cols;
  return 8 * (R * (ncv + 1) * ld + dim * R + R * ld);
}

int qsi_engine_submit(qsi_engine_t E, const double* h_coeffs, uint64_t seed, int64_t* id) {
  QSI_REQUIRE(E && id, "qsi_engine_submit: null argument");
  const int ns = E->op->n_slots;
  QSI_REQUIRE(ns == 0 || h_coeffs, "qsi_engine_submit: the operator has slots; coefficients required");
  qsi_engine_s::Request r;
  r.id = E->next_id++;
  r.seed = seed;
  if (ns > 0) r.coeffs.assign(h_coeffs, h_coeffs + ns);
  E->queue.push_back(std::move(r));
  *id = E->queue.back().id;
  return QSI_OK;
}

int qsi_engine_active(qsi_engine_t E) {
  if (!E) return -1;
  int n = 0;
  for (auto& c : E->cols) n += c.active ? 1 : 0;
  return n;
}

int qsi_engine_queued(qsi_engine_t E) { return E ? static_cast<int>(E->queue.size()) : -1; }

// First half of an engine step: admit queued requests into free columns and
// enqueue one Lanczos step for every active column whose basis is not full
// on the engine's stream.  A second step may be enqueued before the first one
// is finished (its columns continue from where the first leaves them; a
// column whose basis the first fills waits for its restart).  *launched
// (optional) = 0 when no column can step (nothing enqueued).  Engines on
// different streams overlap between step_begin and step_end.
int qsi_engine_step_begin(qsi_engine_t E, int* launched) {
  QSI_REQUIRE(E != nullptr, "qsi_engine_step_begin: null engine");
  QSI_REQUIRE(E->inflight.size() < 2, "qsi_engine_step_begin: two steps are already in flight");
  if (launched) *launched = 0;
  int rc = engine_admit(E);
  if (rc != QSI_OK) return rc;
  // stepping columns take apply slots 0..nact in column order
  int nact = 0;
  for (int c = 0; c < E->R; ++c) {
    const auto& col = E->cols[c];
    const bool in = col.active && !col.held;
    E->colj[c] = in ? col.j : -1;
    if (in) E->slotcol[nact++] = c;
  }
  if (nact == 0) return QSI_OK;
  hipStream_t s = E->st;
  const int64_t D = E->D;
  const int R = E->R;
  const int ns = E->op->n_slots;
  const int bi = E->next_buf;
  E->next_buf ^= 1;
  auto& B = E->sb[bi];
  // through pinned memory: an async copy from pageable memory is not ordered
  // with the kernels of a non-blocking stream.  The device copy is read by
  // this step's kernels only: the previous step's kernels come first on the
  // stream.
  std::copy(E->colj.begin(), E->colj.end(), B.colj_p);
  std::copy(E->slotcol.begin(), E->slotcol.begin() + nact, B.slot_p);
  for (int k = 0; k < ns; ++k)
    for (int sl = 0; sl < nact; ++sl)
      B.coeffs_p[k * nact + sl] = E->coeffs_h[static_cast<size_t>(k) * R + E->slotcol[sl]];
  QSI_CHECK_HIP(hipMemcpyAsync(E->coeffs, B.coeffs_p,
                               static_cast<size_t>(std::max(ns, 1)) * R * sizeof(double) +
                                   2 * R * sizeof(int),
                               hipMemcpyHostToDevice, s));
  // X rows padded to a power of two (one aligned 64- or 128-byte segment per
  // gathered row for up to 16 columns) when the buffer allows
  int ldx = 1;
  while (ldx < nact) ldx <<= 1;
  if (ldx > R) ldx = nact;
  hipLaunchKernelGGL(k_srv_prep, dim3(static_cast<uint32_t>(std::min<int64_t>((D + 63) / 64, 8192))),
                     dim3(kSrvThreads), static_cast<size_t>(64) * (nact + 1) * sizeof(double),
                     s, E->Vb, E->ld, E->ncv1, E->colj_d, E->vs, E->slot_d, nact, ldx,
                     D, E->X);
  QSI_CHECK_LAUNCH();
  // the block apply over the nact slots: X is D x ldx row-major, slot sl's
  // output is column sl of Wc
  BlkArgs a = E->args;
  a.R = nact;
  int G = 64;
  while (G / 2 >= nact && G > 1) G >>= 1;
  // the cached kernel takes any row-group width: one lane per live column
  a.G = blk_cached(a) ? nact : G;
  a.ldx = ldx;
  rc = launch_blk_apply(a, E->op->basis->nw, E->n_blocks, E->X, E->Wc, nullptr, s);
  if (rc != QSI_OK) return rc;
  const dim3 gcol(E->nblk, nact);
  // hacc is written by the first projection pass (entries past a column's
  // j are stale and never read)
  if (E->mmax == 16) launch_reorth<16>(E, gcol, s);
  else if (E->mmax == 24) launch_reorth<24>(E, gcol, s);
  else if (E->mmax == 32) launch_reorth<32>(E, gcol, s);
  else launch_reorth<64>(E, gcol, s);
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_srv_out, dim3((R * kHMax + 255) / 256), dim3(256), 0, s, E->hacc, E->beta,
                     R, B.hh_dev, B.bb_dev);
  QSI_CHECK_LAUNCH();
  QSI_CHECK_HIP(hipEventRecord(B.ev, s));
  for (int c = 0; c < R; ++c) {
    B.colj[c] = E->colj[c];
    B.req[c] = E->cols[c].req;
    if (E->colj[c] >= 0) {
      auto& col = E->cols[c];
      col.j += 1;
      if (col.j == E->ncv) col.held = true;
    }
  }
  E->inflight.push_back(bi);
  if (launched) *launched = 1;
  return QSI_OK;
}

// Second half: wait for the oldest enqueued step, update its columns'
// projected matrices, run Rayleigh-Ritz / convergence / restart where a
// column's basis is full or exhausted, and enqueue the restarts.  Results of
// a column that finished in between (its request left) are dropped.
// *n_active (optional) = columns active afterwards.  Without an enqueued step
// it only reports *n_active.
int qsi_engine_step_end(qsi_engine_t E, int* n_active) {
  QSI_REQUIRE(E != nullptr, "qsi_engine_step_end: null engine");
  if (!E->inflight.empty()) {
    const int bi = E->inflight.front();
    E->inflight.pop_front();
    auto& B = E->sb[bi];
    QSI_CHECK_HIP(hipEventSynchronize(B.ev));
    const int ncv = E->ncv;
    for (int c = 0; c < E->R; ++c) {
      auto& col = E->cols[c];
      const int j = B.colj[c];
      if (j < 0 || !col.active || col.req != B.req[c]) continue;
      col.steps += 1;
      // projections onto the kept Ritz vectors only couple to the first new
      // vector (arrowhead); later ones are rounding noise
      for (int i = 0; i <= j; ++i) {
        const double v = B.hh_p[static_cast<size_t>(c) * kHMax + i];
        if ((i < col.k && j == col.k) || i >= col.k) {
          col.T[static_cast<size_t>(j) * ncv + i] = v;
          col.T[static_cast<size_t>(i) * ncv + j] = v;
        }
      }
      double tn = 0.0;
      for (int i = 0; i <= j; ++i)
        tn = std::max(tn, std::fabs(col.T[static_cast<size_t>(j) * ncv + i]));
      const double beta = B.bb_p[c];
      const bool exhausted = beta <= 1e-13 * std::max(tn, 1e-300);
      col.beta_last = exhausted ? 0.0 : beta;
      if (j + 1 == ncv || exhausted) {
        const int rc = engine_ritz(E, c, j + 1, exhausted);
        if (rc != QSI_OK) return rc;
      }
    }
    const int rc = engine_restarts(E);
    if (rc != QSI_OK) return rc;
  }
  int act = 0;
  for (auto& col : E->cols) act += col.active ? 1 : 0;
  if (n_active) *n_active = act;
  return QSI_OK;
}

// 1 when no step is enqueued or the oldest enqueued one has completed
// (step_end will not wait), 0 while it runs, negative on error.
int qsi_engine_step_ready(qsi_engine_t E) {
  if (!E) return QSI_E_INVALID;
  if (E->inflight.empty()) return 1;
  const hipError_t e = hipEventQuery(E->sb[E->inflight.front()].ev);
  if (e == hipSuccess) return 1;
  if (e == hipErrorNotReady) return 0;
  set_error(std::string("qsi_engine_step_ready: ") + hipGetErrorString(e));
  return QSI_E_HIP;
}

// Steps enqueued and not yet finished by step_end (0, 1 or 2).
int qsi_engine_inflight(qsi_engine_t E) {
  return E ? static_cast<int>(E->inflight.size()) : -1;
}

// One engine step: keeps up to two steps enqueued (step_begin while fewer are
// in flight and a column can step), then finishes the oldest (step_end).
// Synchronises with that step.
int qsi_engine_step(qsi_engine_t E, int* n_active) {
  QSI_REQUIRE(E != nullptr, "qsi_engine_step: null engine");
  while (E->inflight.size() < 2) {
    int launched = 0;
    const int rc = qsi_engine_step_begin(E, &launched);
    if (rc != QSI_OK) return rc;
    if (!launched) break;
  }
  return qsi_engine_step_end(E, n_active);
}

// Pops one finished request: returns 1 and fills *id, h_evals (nev values),
// *nconv and *steps (Lanczos steps it took), or 0 when none is finished.
int qsi_engine_poll(qsi_engine_t E, int64_t* id, double* h_evals, int* nconv, int* steps) {
  if (!E || E->done.empty()) return 0;
  auto& r = E->done.front();
  if (id) *id = r.id;
  if (h_evals) std::copy(r.evals.begin(), r.evals.end(), h_evals);
  if (nconv) *nconv = r.nconv;
  if (steps) *steps = r.steps;
  E->done.pop_front();
  return 1;
}

}  // extern "C"
