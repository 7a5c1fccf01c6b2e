// Deterministic FP64 reductions: wavefront (64 lanes) butterfly, then LDS
// across the waves of a workgroup, then one partial per workgroup that a
// single-workgroup kernel sums in a fixed order.
#pragma once

#include "qsi_common.hpp"

namespace qsi {

constexpr int kWave = 64;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Sum over a workgroup of BLOCK threads; result valid in thread 0.
template <int BLOCK>
__device__ __forceinline__ double block_sum(double v) {
  static_assert(BLOCK % kWave == 0, "block must be whole waves");
  __shared__ double red[BLOCK / kWave];
  v = wave_sum(v);
  const int lane = threadIdx.x % kWave;
  const int wid = threadIdx.x / kWave;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int w = 0; w < BLOCK / kWave; ++w) r += red[w];
  }
  return r;
}

// Reduction of per-workgroup partials by one workgroup of BLOCK threads, in
// a fixed order: partials[i * stride] for i < n.  Each thread keeps eight
// interleaved sums (entries t + (8k+u)·BLOCK go to sum u) so eight loads are
// in flight, then the sums are combined in a fixed order and block-summed.
// mode 0: out[0] = sum.  mode 1: out[0] = sqrt(sum), out[1] = 1/out[0]
// (0 if the norm is 0).
template <int BLOCK = 256>
__global__ __launch_bounds__(BLOCK) void k_finalize_sum(
    const double* __restrict__ partials, int64_t n, double* __restrict__ out,
    int mode, int64_t stride = 1) {
  double a[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  int64_t i = threadIdx.x;
  for (; i + 7 * BLOCK < n; i += 8 * BLOCK) {
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] += partials[(i + u * BLOCK) * stride];
  }
  for (; i < n; i += BLOCK) a[0] += partials[i * stride];
  double v = ((a[0] + a[1]) + (a[2] + a[3])) + ((a[4] + a[5]) + (a[6] + a[7]));
  v = block_sum<BLOCK>(v);
  if (threadIdx.x == 0) {
    if (mode == 0) {
      out[0] = v;
    } else {
      const double nrm = sqrt(v);
      out[0] = nrm;
      out[1] = nrm > 0.0 ? 1.0 / nrm : 0.0;
    }
  }
}

// First stage of the two-stage sum used for long partials arrays (a
// 2-billion-row apply leaves 7.8 M): workgroup b sums partials
// [b·chunk, min(n, (b+1)·chunk)) in a fixed order and stores the sum over
// partials[b·chunk].  Only workgroup b reads or writes that range, and the
// store follows the block sum's barrier, so no workgroup reads another's
// output; k_finalize_sum with stride = chunk then sums the chunk results.
template <int BLOCK = 256>
__global__ __launch_bounds__(BLOCK) void k_sum_chunks(double* partials, int64_t n,
                                                      int64_t chunk) {
  const int64_t lo = static_cast<int64_t>(blockIdx.x) * chunk;
  const int64_t hi = lo + chunk < n ? lo + chunk : n;
  double v = 0.0;
  for (int64_t i = lo + threadIdx.x; i < hi; i += BLOCK) v += partials[i];
  v = block_sum<BLOCK>(v);
  if (threadIdx.x == 0) partials[lo] = v;
}

}  // namespace qsi
