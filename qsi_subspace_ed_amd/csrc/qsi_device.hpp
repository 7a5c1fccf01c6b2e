// Device-side term evaluation shared by the apply kernels (pull form of
// SymbolicPMROperator::applyState, components/operator/include/operator.hpp:
// 232-254, and LazyOpSum::evaluate_add_diagonal,
// components/operator/src/operator_matrix.cpp:65-82).
#pragma once

#include "qsi_internal.hpp"

namespace qsi {

constexpr double kApplyTol = 1e-15;  // LazyOpSum::APPLY_TOL

template <int NW>
__device__ __forceinline__ bool term_ok(const OffTerm& t, const State<NW>& s) {
  if constexpr (NW == 1) {
    return (s.w[0] & t.m[0]) == t.v[0];
  } else {
    return ((s.w[0] & t.m[0]) == t.v[0]) & ((s.w[1] & t.m[1]) == t.v[1]);
  }
}

template <int NW>
__device__ __forceinline__ int parity_notz(const State<NW>& s,
                                           const uint64_t* Z) {
  int p = __popcll(~s.w[0] & Z[0]);
  if constexpr (NW == 2) p += __popcll(~s.w[1] & Z[1]);
  return p & 1;
}

template <int NW>
__device__ __forceinline__ int popc_and(const State<NW>& s, const uint64_t* Z) {
  int p = __popcll(s.w[0] & Z[0]);
  if constexpr (NW == 2) p += __popcll(s.w[1] & Z[1]);
  return p;
}

template <int NW>
__device__ __forceinline__ State<NW> xor_state(const State<NW>& s,
                                               const uint64_t* X) {
  State<NW> r;
  r.w[0] = s.w[0] ^ X[0];
  if constexpr (NW == 2) r.w[1] = s.w[1] ^ X[1];
  return r;
}

template <int NW>
__device__ __forceinline__ bool diag_ok(const DiagTerm& d, const State<NW>& s) {
  if constexpr (NW == 1) {
    return (s.w[0] & d.m[0]) == d.v[0];
  } else {
    return ((s.w[0] & d.m[0]) == d.v[0]) & ((s.w[1] & d.m[1]) == d.v[1]);
  }
}

// Diagonal value d(s) = sum_t c_t * (+-1 or field count); x-independent.
template <int NW>
__device__ __forceinline__ double diag_accumulate(
    double acc, double xi, const State<NW>& s, const DiagTerm* __restrict__ dg,
    int n_diag) {
  for (int t = 0; t < n_diag; ++t) {
    const DiagTerm& d = dg[t];
    double f;
    if (d.kind == 0) {
      f = diag_ok<NW>(d, s) ? d.sgn * (1 - 2 * parity_notz<NW>(s, d.Z)) : 0.0;
    } else {
      int nz = __popcll(d.Z[0]);
      if constexpr (NW == 2) nz += __popcll(d.Z[1]);
      f = d.sgn * static_cast<double>(2 * popc_and<NW>(s, d.Z) - nz);
    }
    // reference: dy = c * x[i] * applyState(psi)
    acc += d.c * xi * f;
  }
  return acc;
}


}  // namespace qsi
