// Device-side term evaluation shared by the apply kernels (pull form of
// SymbolicPMROperator::applyState, components/operator/include/operator.hpp:
// 232-254, and LazyOpSum::evaluate_add_diagonal,
// components/operator/src/operator_matrix.cpp:65-82).
#pragma once

#include <cstddef>

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

// w = 2 w + (this lane's bit of `lanes`): the lane mask of a compare
// shifted into a per-lane word by one add with carry-in.
__device__ __forceinline__ uint32_t shift_in_lane(uint32_t w, uint64_t lanes) {
  uint32_t r;
  uint64_t carry;
  asm("v_addc_co_u32_e64 %0, %1, %2, %2, %3" : "=v"(r), "=s"(carry) : "v"(w), "s"(lanes));
  return r;
}

// A term's projector mask and value, read as one 32-byte block (the first
// two fields of OffTerm).
struct TermMV {
  uint64_t m[2];
  uint64_t v[2];
};
static_assert(offsetof(OffTerm, m) == 0 && offsetof(OffTerm, v) == sizeof(TermMV::m),
              "OffTerm must start with its m and v words");
template <int NW>
__device__ __forceinline__ TermMV load_mv(const OffTerm* __restrict__ t) {
  return *reinterpret_cast<const TermMV*>(t);
}

// Lanes of the wavefront whose s & m (u) equals v.
template <int NW>
__device__ __forceinline__ uint64_t match_lanes(const uint64_t* u, const uint64_t* v) {
  uint64_t b = __builtin_amdgcn_ballot_w64(u[0] == v[0]);
  if constexpr (NW == 2) b &= __builtin_amdgcn_ballot_w64(u[1] == v[1]);
  return b;
}

// Which of the terms off[0, tn) (tn <= 64) connect row s: bit 63 - t of the
// result for term t, so the terms come out in order by leading zeros.  Every
// lane of the wavefront tests the same term at once against its masks held
// in scalar registers; the compare's lane mask goes into a 32-bit word per
// 32 terms with one add, and a term with the same projector mask m as the
// one before it (a ring flip and its conjugate) reuses s & m.  Called by
// every active lane of the wavefront together.
template <int NW>
__device__ __forceinline__ uint64_t connect_mask(const OffTerm* __restrict__ off, int tn,
                                                 const State<NW>& s) {
  uint32_t word[2];
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int lo = 32 * h, hi = min(tn, lo + 32);
    uint32_t w = 0;
    for (int t = lo; t < hi;) {
      // this term's projector mask and value in one scalar load
      const TermMV A = load_mv<NW>(off + t);
      uint64_t u[NW];
#pragma unroll
      for (int q = 0; q < NW; ++q) u[q] = s.w[q] & A.m[q];
      w = shift_in_lane(w, match_lanes<NW>(u, A.v));
      ++t;
      if (t < hi) {
        const TermMV B = load_mv<NW>(off + t);
        bool same = B.m[0] == A.m[0];
        if constexpr (NW == 2) same = same && B.m[1] == A.m[1];
        if (same) {
          w = shift_in_lane(w, match_lanes<NW>(u, B.v));
          ++t;
        }
      }
    }
    const int n = hi - lo;
    word[h] = n > 0 ? w << (32 - n) : 0u;
  }
  return (static_cast<uint64_t>(word[0]) << 32) | word[1];
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
