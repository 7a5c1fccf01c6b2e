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

// Reduction of per-workgroup partials by one workgroup of 256 threads, in a
// fixed order (thread t sums partials t, t+256, ... then a block sum).
// mode 0: out[0] = sum.  mode 1: out[0] = sqrt(sum), out[1] = 1/out[0]
// (0 if the norm is 0).
template <int BLOCK = 256>
__global__ __launch_bounds__(BLOCK) void k_finalize_sum(
    const double* __restrict__ partials, int64_t n, double* __restrict__ out,
    int mode) {
  double v = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += BLOCK) v += partials[i];
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

}  // namespace qsi
