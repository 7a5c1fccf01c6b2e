// Device side of the qsi library: device basis + top-bit bucket table,
// matrix-free pull apply (single vector and block), connectivity
// materialiser and streaming FP64 vector kernels.  gfx950 only.
//
// Reference behaviour being replaced (file:line in the reference tree):
//   ZBasisInterp::find_bounds / search   components/operator/src/operator.cpp:38-86
//   SymbolicPMROperator::applyState      components/operator/include/operator.hpp:232-254
//   LazyOpSum::evaluate_add_off_diag_omp components/operator/src/operator_matrix.cpp:28-60
//   LazyOpSum::evaluate_add_diagonal     components/operator/src/operator_matrix.cpp:65-82
//   LazyOpSum::toSparseMatrix            components/operator/include/operator_matrix.hpp:79-105
//
// The reference pushes: for each term and each column j it finds
// J = rank(O s_j) and atomically adds c*sign*x_j to y_J.  Here every output
// row i pulls instead: for each term whose precondition holds on the
// pre-image s_i ^ X, it gathers x at rank(s_i ^ X).  Both compute
// y_i = sum_j H_ij x_j for the same H (no symmetry is assumed), but the pull
// form writes each y_i exactly once, so it needs no atomics and its sum order
// is fixed (term order, then diagonal terms) — results are deterministic.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "qsi_blockk.hpp"
#include "qsi_device.hpp"
#include "qsi_internal.hpp"
#include "qsi_reduce.hpp"

namespace qsi {

constexpr int kApplyBlock = 256;
constexpr double kSparseTol = 1e-14;  // toSparseMatrix default tol

// ============================================================================
// Basis: bucket table
// ============================================================================

// First rank whose bucket key is >= k (a bisection over the basis).
template <int NW>
__device__ __forceinline__ uint32_t table_entry(const BasisView& b, int64_t k) {
  int64_t lo = 0, n = b.dim;
  while (n > 0) {
    const int64_t half = n >> 1;
    const State<NW> s = load_state_cached<NW>(b.states, lo + half);
    if (static_cast<int64_t>(bucket_key<NW>(b, s)) < k) {
      lo = lo + half + 1;
      n = n - half - 1;
    } else {
      n = half;
    }
  }
  return static_cast<uint32_t>(lo);
}

// table[k] = first rank whose bucket key is >= k, for k in [0, 2^K].
// Grid-stride: a dispatch holds fewer than 2^32 work-items, and a 32-bit key
// has 2^32 + 1 entries.
template <int NW>
__global__ void k_build_table(BasisView b, int64_t nkeys, uint32_t* __restrict__ table) {
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k <= nkeys;
       k += stride)
    table[k] = table_entry<NW>(b, k);
}

// mask[p / 64] bit p % 64 = 1 if some neighbouring pair of the sorted basis
// first differs at bit p.  Grid-stride; one global OR per word per block.
template <int NW>
__global__ __launch_bounds__(256) void k_diff_bits(const uint64_t* __restrict__ states,
                                                   int64_t dim,
                                                   unsigned long long* __restrict__ mask) {
  __shared__ unsigned long long sm[2];
  if (threadIdx.x < 2) sm[threadIdx.x] = 0;
  __syncthreads();
  unsigned long long m0 = 0, m1 = 0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i + 1 < dim;
       i += stride) {
    const State<NW> a = load_state_cached<NW>(states, i);
    const State<NW> c = load_state_cached<NW>(states, i + 1);
    if constexpr (NW == 2) {
      const uint64_t hi = a.w[1] ^ c.w[1];
      if (hi) {
        m1 |= 1ull << (63 - __clzll(hi));
        continue;
      }
    }
    const uint64_t lo = a.w[0] ^ c.w[0];
    if (lo) m0 |= 1ull << (63 - __clzll(lo));
  }
  if (m0) atomicOr(&sm[0], m0);
  if (m1) atomicOr(&sm[1], m1);
  __syncthreads();
  if (threadIdx.x < 2 && sm[threadIdx.x]) atomicOr(&mask[threadIdx.x], sm[threadIdx.x]);
}

// Insert every basis state into the hash index (slots pre-filled with ones).
// A thread claims a free slot by a compare-and-swap on its rank word and then
// writes the state; keys are unique, so no insertion compares keys.
template <int NW>
__global__ void k_build_hash(const uint64_t* __restrict__ states, int64_t dim,
                             ulonglong2* __restrict__ slots, uint64_t mask,
                             int hbits) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const State<NW> s = load_state_cached<NW>(states, i);
  uint64_t p = hash_home<NW>(s, hbits);
  for (;;) {
    unsigned long long* rank_word =
        (NW == 1) ? reinterpret_cast<unsigned long long*>(&slots[p].y)
                  : reinterpret_cast<unsigned long long*>(&slots[2 * p + 1].x);
    const unsigned long long old =
        atomicCAS(rank_word, static_cast<unsigned long long>(kHashEmpty),
                  static_cast<unsigned long long>(i));
    if (old == kHashEmpty) {
      if constexpr (NW == 1) {
        slots[p].x = s.w[0];
      } else {
        slots[2 * p].x = s.w[0];
        slots[2 * p].y = s.w[1];
      }
      return;
    }
    p = (p + 1) & mask;
  }
}

// Place every basis state in its bucket's slot region (slots pre-filled with
// kSlotEmpty): linear probing from the state's home, a slot claimed by a
// compare-and-swap.  A bucket above kSlotMaxBucket states sets *too_big and
// places nothing.
// For every row and connecting term of an operator: whether the target's
// low-key lookup ends on a candidate with the target's bucket key and low key
// but another state (a target outside the basis that differs from a basis
// state only at positions fixed in the basis: *bad |= 1), and whether any
// target is outside the basis (*bad |= 2) (qsi_op_verify_lowkeys).
template <int NW>
__global__ void k_verify_lowkeys(BasisView b, const OffTerm* __restrict__ off, int n_off,
                                 unsigned int* __restrict__ bad) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.dim) return;
  const State<NW> s = load_state_cached<NW>(b.states, i);
  bool clash = false, absent = false;
  for (int t = 0; t < n_off; ++t) {
    if (!term_ok<NW>(off[t], s)) continue;
    const State<NW> key = xor_state<NW>(s, off[t].X);
    uint32_t lo, n;
    bucket_bounds(b.table, bucket_key<NW>(b, key), lo, n);
    if (n == 0) {
      absent = true;
      continue;
    }
    const uint32_t r = low_key<NW>(b, key);
    while (n > 1) {
      const uint32_t half = n >> 1;
      lo = b.lowkeys[lo + half] <= r ? lo + half : lo;
      n -= half;
    }
    if (b.lowkeys[lo] != r) absent = true;
    else if (!st_eq<NW>(load_state_cached<NW>(b.states, lo), key)) clash = true;
  }
  if (clash || absent) atomicOr(bad, (clash ? 1u : 0u) | (absent ? 2u : 0u));
}

// lowkeys[i] = low key of state i (qsi_basis_build_lowkeys)
template <int NW>
__global__ void k_build_lowkeys(BasisView b, uint32_t* __restrict__ lowkeys) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.dim) return;
  lowkeys[i] = low_key<NW>(b, load_state_cached<NW>(b.states, i));
}

template <int NW>
__global__ void k_build_slots(BasisView b, uint32_t* __restrict__ slots,
                              unsigned int* __restrict__ too_big) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.dim) return;
  const State<NW> s = load_state_cached<NW>(b.states, i);
  uint32_t lo, n;
  bucket_bounds(b.table, bucket_key<NW>(b, s), lo, n);
  if (n > kSlotMaxBucket) {
    atomicOr(too_big, 1u);
    return;
  }
  const uint64_t h = slot_hash<NW>(s);
  const uint32_t m = 2 * n;
  const uint32_t e = (slot_fp(h) << 16) | static_cast<uint32_t>(i - lo);
  uint32_t* reg = slots + 2 * static_cast<uint64_t>(lo);
  uint32_t p = slot_home(h, m);
  while (atomicCAS(reg + p, kSlotEmpty, e) != kSlotEmpty) p = p + 1 == m ? 0 : p + 1;
}

template <int NW>
__global__ void k_search(BasisView b, const uint64_t* __restrict__ q, int64_t n,
                         int64_t* __restrict__ rank) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  State<NW> s;
  // queries always come as [lo, hi] pairs
  s.w[0] = q[2 * i];
  if constexpr (NW == 2) s.w[1] = q[2 * i + 1];
  bool ok = true;
  if constexpr (NW == 1) ok = (q[2 * i + 1] == 0);
  rank[i] = !ok ? -1 : b.slots ? slot_lookup<NW>(b, s) : basis_lookup<NW>(b, s);
}

// ============================================================================
// Apply: one lane per output row
// ============================================================================

// Flip masks and signed coefficients of the first kLdsTerms off-diagonal
// terms, staged in LDS by every workgroup: lookups re-form their connected
// state and take their coefficient from here instead of gathering each
// lane's term from global memory.
constexpr int kLdsTerms = 256;
template <int NW>
struct TermsLds {
  uint64_t X[NW][kLdsTerms];
  double c[kLdsTerms];  // c * sgn
};

template <int NW>
__device__ __forceinline__ bool stage_terms(TermsLds<NW>& tl, const OffTerm* __restrict__ off,
                                            int n_off) {
  const bool lds = n_off <= kLdsTerms;
  if (lds) {
    for (int t = threadIdx.x; t < n_off; t += blockDim.x) {
#pragma unroll
      for (int w = 0; w < NW; ++w) tl.X[w][t] = off[t].X[w];
      tl.c[t] = off[t].c * off[t].sgn;
    }
  }
  __syncthreads();
  return lds;
}

template <int NW>
__device__ __forceinline__ State<NW> term_target(const TermsLds<NW>& tl, bool lds,
                                                 const OffTerm* __restrict__ off,
                                                 const State<NW>& s, int t) {
  State<NW> r;
#pragma unroll
  for (int w = 0; w < NW; ++w) r.w[w] = s.w[w] ^ (lds ? tl.X[w][t] : off[t].X[w]);
  return r;
}

// Terms as the lane and low-key kernels stage them in LDS, kLdsTerms at a
// time: flip masks, signed coefficients, and each flip mask's bucket key and
// low key.  Both keys select bits of a state, so they commute with a flip:
// key(s ^ X) = key(s) ^ key(X).  A lookup forms its keys from the row's own
// keys with an XOR instead of gathering them bit field by bit field from
// the connected state.
template <int NW>
struct LowTermsLds {
  uint64_t X[NW][kLdsTerms];
  double c[kLdsTerms];   // c * sgn
  uint2 key[kLdsTerms];  // bucket key, low key (0 without low keys) of X
};

// Stages terms [0, n) of `off` (n <= kLdsTerms).  Every thread of the
// workgroup calls it; `reuse` (uniform) first waits until every thread is
// done with the previous chunk.
template <int NW>
__device__ __forceinline__ void stage_low_terms(LowTermsLds<NW>& tl, const BasisView& b,
                                                const OffTerm* __restrict__ off, int n,
                                                bool reuse) {
  if (reuse) __syncthreads();
  for (int t = threadIdx.x; t < n; t += blockDim.x) {
    State<NW> X;
#pragma unroll
    for (int w = 0; w < NW; ++w) {
      X.w[w] = off[t].X[w];
      tl.X[w][t] = X.w[w];
    }
    tl.c[t] = off[t].c * off[t].sgn;
    tl.key[t] = make_uint2(bucket_key<NW>(b, X), b.lowkeys ? low_key<NW>(b, X) : 0u);
  }
  __syncthreads();
}

template <int NW>
__device__ __forceinline__ State<NW> staged_target(const LowTermsLds<NW>& tl,
                                                   const State<NW>& s, uint32_t t) {
  State<NW> r;
#pragma unroll
  for (int w = 0; w < NW; ++w) r.w[w] = s.w[w] ^ tl.X[w][t];
  return r;
}

// y[i] = yscale * y_in[i] + sum_t H_{i,t} x[...] + d(s_i) x[i]
// If DOT, partials[block] = sum over the block's rows of x[i] * y[i].
// Lane variant: every lane walks its own row's connected states and runs a
// full table lookup for each.  G lookups go through the bucket table, the
// bisection and the x gather together, and every stage issues the loads of
// all G (those of unfinished lookups) before it uses any, so a lane keeps G
// loads in flight; contributions are still added in term order.  An empty
// asm statement that reads each loaded value after the last load of the
// stage keeps the compiler from moving a use next to its load, which would
// make every load wait for the one before.
template <int NW, bool HAS_Z, bool DOT, int G>
__global__ __launch_bounds__(kApplyBlock)
__attribute__((amdgpu_waves_per_eu(G >= 8 ? 3 : G >= 4 ? 5 : 8))) void k_apply_lane(
    BasisView b, const OffTerm* __restrict__ off, int n_off,
    const DiagTerm* __restrict__ dg, int n_diag, const double* __restrict__ x,
    double* y, const double* y_in, double yscale,
    const double* __restrict__ yscale_dev, double* __restrict__ partials,
    const double* __restrict__ xs_dev, const double* __restrict__ ys_dev) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if (yscale_dev) yscale *= yscale_dev[0];
  // x and y_in are read as xs * x and ys * y_in.  Only the dot-product
  // instances (the Lanczos step) take scales; elsewhere they are the
  // constant 1 and the multiplies fold away.
  const double xs = (DOT && xs_dev) ? xs_dev[0] : 1.0;
  const double ys = (DOT && ys_dev) ? ys_dev[0] : 1.0;
  __shared__ LowTermsLds<NW> tl;
  const int64_t i = static_cast<int64_t>(blk) * kApplyBlock + threadIdx.x;
  const bool live = i < b.dim;
  State<NW> s = {};
  double xi = 0.0, acc = 0.0, prod = 0.0;
  uint32_t ks = 0;
  if (live) {
    s = load_state_cached<NW>(b.states, i);
    xi = x[i] * xs;
    acc = y_in ? yscale * (y_in[i] * ys) : 0.0;
    ks = bucket_key<NW>(b, s);
  }
  for (int c0 = 0; c0 < n_off; c0 += kLdsTerms) {
    const int cn = min(kLdsTerms, n_off - c0);
    stage_low_terms<NW>(tl, b, off + c0, cn, c0 > 0);
    if (!live) continue;
    for (int t0 = 0; t0 < cn; t0 += 64) {
      const int tn = min(64, cn - t0);
      // which terms connect this row (term t0 + t at bit 63 - t)
      uint64_t mask = connect_mask<NW>(off + c0 + t0, tn, s);
      while (mask) {
        uint32_t tt[G], lo[G], n[G];
        bool act[G];
        uint2 p[G];
        // bucket bounds of all G lookups (each connected state is re-formed
        // from the row and the staged flip mask where it is compared)
#pragma unroll
        for (int g = 0; g < G; ++g) {
          act[g] = mask != 0;
          const uint32_t z = act[g] ? static_cast<uint32_t>(__builtin_clzll(mask)) : 0u;
          tt[g] = t0 + z;
          mask &= ~(uint64_t{1} << 63 >> z);
          p[g] = make_uint2(0u, 0u);
          if (act[g]) __builtin_memcpy(&p[g], b.table + (ks ^ tl.key[tt[g]].x), sizeof(uint2));
        }
#pragma unroll
        for (int g = 0; g < G; ++g) asm volatile("" : "+v"(p[g].x), "+v"(p[g].y));
        bool more = false;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          lo[g] = p[g].x;
          n[g] = act[g] ? p[g].y - p[g].x : 0u;
          more |= n[g] > 1;
        }
        // interleaved bisection down to one candidate per key: the G probes,
        // then the G updates (a finished lookup has half = 0: no change)
        while (more) {
          State<NW> m[G];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            m[g] = State<NW>{};
            if (n[g] > 1) m[g] = load_state_cached<NW>(b.states, lo[g] + (n[g] >> 1));
          }
          // every probe of the group is used only after all were issued
#pragma unroll
          for (int g = 0; g < G; ++g)
#pragma unroll
            for (int w = 0; w < NW; ++w) asm volatile("" : "+v"(m[g].w[w]));
          more = false;
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const uint32_t half = n[g] >> 1;
            lo[g] = st_le<NW>(m[g], staged_target<NW>(tl, s, tt[g])) ? lo[g] + half : lo[g];
            n[g] -= half;
            more |= n[g] > 1;
          }
        }
        // the candidate's state and x are loaded together (lo < dim whenever
        // the bucket is not empty); x counts only if the state matches
        double xj[G], c[G];
        State<NW> f[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          xj[g] = 0.0;
          c[g] = tl.c[tt[g]];
          if (n[g] == 1) {
            f[g] = load_state_cached<NW>(b.states, lo[g]);
            xj[g] = x[lo[g]];
          }
        }
        // the scale is applied where the values are consumed, after every
        // gather of the group has been issued
#pragma unroll
        for (int g = 0; g < G; ++g) xj[g] *= xs;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (n[g] == 1 && st_eq<NW>(f[g], staged_target<NW>(tl, s, tt[g])) &&
              fabs(xj[g]) > kApplyTol) {
            double cg = c[g];
            if constexpr (HAS_Z) {
              if (parity_notz<NW>(s, off[c0 + tt[g]].Z)) cg = -cg;
            }
            acc += cg * xj[g];
          }
        }
      }
    }
  }
  if (live) {
    acc = diag_accumulate<NW>(acc, xi, s, dg, n_diag);
    y[i] = acc;
    if constexpr (DOT) prod = xi * acc;
  }
  if constexpr (DOT) {
    const double bs = block_sum<kApplyBlock>(prod);
    if (threadIdx.x == 0) partials[blk] = bs;
  }
}

// Low-key variant (QSI_APPLY_LOW*): as the lane variant, but the bisection
// of a bucket probes the 4-byte low keys (qsi_basis_build_lowkeys) instead of
// whole states — four times as many candidates per cache line, and the state
// array is touched once per lookup at most, for the candidate.  The
// candidate is the last state of the bucket whose low key is <= the
// target's, the same state the lane variant's bisection ends on, so y is
// bit-identical.  A lookup keeps only its term, bucket start, bucket size and
// (below eight lookups per lane) its low key.  Every stage issues the loads of
// all G lookups (those not yet finished) before it uses any, as in the lane
// variant, so a lane keeps G loads in flight through the table, each
// bisection step and the gathers.
//
// CHECK says how a candidate is confirmed.  kLowCmpState: its state is
// loaded and compared.  kLowCmpKey (qsi_op_verify_lowkeys found that a target
// whose bucket key and low key match a basis state is that state): its low
// key is compared, and the state array is not read.  kLowTrusted (the pass
// also found every connected state in the basis): the candidate is the
// target, with no compare.
constexpr int kLowCmpState = 0, kLowCmpKey = 1, kLowTrusted = 2;
template <int NW, bool HAS_Z, bool DOT, int G, int CHECK>
__global__ __launch_bounds__(kApplyBlock)
__attribute__((amdgpu_waves_per_eu(
    G >= 8 ? ((CHECK == 2 || (CHECK == 1 && NW == 1)) && !HAS_Z ? 8 : 5) : HAS_Z ? 6 : 8)))
void k_apply_low(
    BasisView b, const OffTerm* __restrict__ off, int n_off,
    const DiagTerm* __restrict__ dg, int n_diag, const double* __restrict__ x,
    double* y, const double* y_in, double yscale,
    const double* __restrict__ yscale_dev, double* __restrict__ partials,
    const double* __restrict__ xs_dev, const double* __restrict__ ys_dev) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if (yscale_dev) yscale *= yscale_dev[0];
  // scales as in k_apply_lane (dot-product instances only)
  const double xs = (DOT && xs_dev) ? xs_dev[0] : 1.0;
  const double ys = (DOT && ys_dev) ? ys_dev[0] : 1.0;
  __shared__ LowTermsLds<NW> tl;
  const int64_t i = static_cast<int64_t>(blk) * kApplyBlock + threadIdx.x;
  const bool live = i < b.dim;
  State<NW> s = {};
  double xi = 0.0, acc = 0.0, prod = 0.0;
  uint32_t ks = 0, ls = 0;
  if (live) {
    s = load_state_cached<NW>(b.states, i);
    xi = x[i] * xs;
    acc = y_in ? yscale * (y_in[i] * ys) : 0.0;
    ks = bucket_key<NW>(b, s);
    ls = low_key<NW>(b, s);
  }
  for (int c0 = 0; c0 < n_off; c0 += kLdsTerms) {
    const int cn = min(kLdsTerms, n_off - c0);
    stage_low_terms<NW>(tl, b, off + c0, cn, c0 > 0);
    if (!live) continue;
    for (int t0 = 0; t0 < cn; t0 += 64) {
      const int tn = min(64, cn - t0);
      uint64_t mask = connect_mask<NW>(off + c0 + t0, tn, s);
      while (mask) {
        // per lookup: term, bucket start, bucket size (0: nothing to find),
        // and below eight lookups the target's low key (with eight it is
        // re-formed from the staged keys where it is compared, which keeps
        // the kernel within the registers of eight waves per SIMD); bucket
        // bounds of all G lookups loaded first
        uint32_t tt[G], lo[G], n[G];
        constexpr bool kKeepLow = G < 8;
        uint32_t r[kKeepLow ? G : 1];
        auto target_low = [&](int g) -> uint32_t {
          if constexpr (kKeepLow) return r[g];
          else return ls ^ tl.key[tt[g]].y;
        };
        bool act[G];
        uint2 p[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          act[g] = mask != 0;
          const uint32_t z = act[g] ? static_cast<uint32_t>(__builtin_clzll(mask)) : 0u;
          tt[g] = t0 + z;
          mask &= ~(uint64_t{1} << 63 >> z);
          const uint2 kx = tl.key[tt[g]];
          if constexpr (kKeepLow) r[g] = ls ^ kx.y;
          p[g] = make_uint2(0u, 0u);
          if (act[g]) __builtin_memcpy(&p[g], b.table + (ks ^ kx.x), sizeof(uint2));
        }
#pragma unroll
        for (int g = 0; g < G; ++g) asm volatile("" : "+v"(p[g].x), "+v"(p[g].y));
        bool more = false;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          lo[g] = p[g].x;
          n[g] = act[g] ? p[g].y - p[g].x : 0u;
          more |= n[g] > 1;
        }
        // interleaved bisection on the low keys down to one candidate: the G
        // probes, then the G updates (a finished lookup has half = 0: no
        // change)
        while (more) {
          uint32_t m[G];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            m[g] = 0;
            if (n[g] > 1) m[g] = b.lowkeys[lo[g] + (n[g] >> 1)];
          }
          // every probe of the group is used only after all were issued
#pragma unroll
          for (int g = 0; g < G; ++g) asm volatile("" : "+v"(m[g]));
          more = false;
#pragma unroll
          for (int g = 0; g < G; ++g) {
            const uint32_t half = n[g] >> 1;
            lo[g] = m[g] <= target_low(g) ? lo[g] + half : lo[g];
            n[g] -= half;
            more |= n[g] > 1;
          }
        }
        // candidates in chunks of up to four: every candidate state and x of
        // a chunk first, then its compares and sums (in term order)
        constexpr int CH = G % 4 == 0 ? 4 : G % 3 == 0 ? 3 : G % 2 == 0 ? 2 : 1;
        static_assert(G % CH == 0, "candidate chunks must tile the lookups");
#pragma unroll
        for (int g0 = 0; g0 < G; g0 += CH) {
          State<NW> f[CH];
          uint32_t fk[CH];
          double xj[CH], c[CH];
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            xj[u] = 0.0;
            c[u] = tl.c[tt[g0 + u]];
            if (n[g0 + u] == 1) {
              if constexpr (CHECK == kLowCmpKey) fk[u] = b.lowkeys[lo[g0 + u]];
              if constexpr (CHECK == kLowCmpState) f[u] = load_state_cached<NW>(b.states, lo[g0 + u]);
              xj[u] = x[lo[g0 + u]];
            }
          }
#pragma unroll
          for (int u = 0; u < CH; ++u) xj[u] *= xs;
#pragma unroll
          for (int u = 0; u < CH; ++u) {
            if (n[g0 + u] == 1 && fabs(xj[u]) > kApplyTol) {
              const int t = static_cast<int>(tt[g0 + u]);
              bool hit = true;
              if constexpr (CHECK == kLowCmpKey) hit = fk[u] == target_low(g0 + u);
              if constexpr (CHECK == kLowCmpState) {
                hit = st_eq<NW>(f[u], staged_target<NW>(tl, s, tt[g0 + u]));
              }
              if (hit) {
                double cu = c[u];
                if constexpr (HAS_Z) {
                  if (parity_notz<NW>(s, off[c0 + t].Z)) cu = -cu;
                }
                acc += cu * xj[u];
              }
            }
          }
        }
      }
    }
  }
  if (live) {
    acc = diag_accumulate<NW>(acc, xi, s, dg, n_diag);
    y[i] = acc;
    if constexpr (DOT) prod = xi * acc;
  }
  if constexpr (DOT) {
    const double bs = block_sum<kApplyBlock>(prod);
    if (threadIdx.x == 0) partials[blk] = bs;
  }
}

// Slot variant: as the lane variant, with each connected state found through
// the slot index under its bucket (bucket bounds, then slots from the
// state's home until an empty slot or a fingerprint match, whose state and x
// are loaded together) instead of a bisection of the bucket.  G lookups per
// lane are carried through the stages together; contributions are added in
// term order, so y is bit-identical to the other variants.  A lookup keeps
// only its bucket start, size and fingerprint, probe position, candidate
// rank and term (the connected state is re-formed from the row and the
// term's flip mask where needed), so four lookups fit the register budget of
// seven waves per SIMD.
template <int NW, bool HAS_Z, bool DOT, int G>
__global__ __launch_bounds__(kApplyBlock) void k_apply_slot(
    BasisView b, const OffTerm* __restrict__ off, int n_off,
    const DiagTerm* __restrict__ dg, int n_diag, const double* __restrict__ x,
    double* y, const double* y_in, double yscale,
    const double* __restrict__ yscale_dev, double* __restrict__ partials,
    const double* __restrict__ xs_dev, const double* __restrict__ ys_dev) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if (yscale_dev) yscale *= yscale_dev[0];
  // x and y_in are read as xs * x and ys * y_in.  Only the dot-product
  // instances (the Lanczos step) take scales; elsewhere they are the
  // constant 1 and the multiplies fold away.
  const double xs = (DOT && xs_dev) ? xs_dev[0] : 1.0;
  const double ys = (DOT && ys_dev) ? ys_dev[0] : 1.0;
  __shared__ TermsLds<NW> tl;
  const bool lds = stage_terms<NW>(tl, off, n_off);
  const int64_t i = static_cast<int64_t>(blk) * kApplyBlock + threadIdx.x;
  double prod = 0.0;
  if (i < b.dim) {
    const State<NW> s = load_state_cached<NW>(b.states, i);
    const double xi = x[i] * xs;
    double acc = y_in ? yscale * (y_in[i] * ys) : 0.0;
    for (int t0 = 0; t0 < n_off; t0 += 64) {
      const int tn = min(64, n_off - t0);
      uint64_t mask = 0;
      for (int t = 0; t < tn; ++t) {
        mask |= static_cast<uint64_t>(term_ok<NW>(off[t0 + t], s)) << t;
      }
      while (mask) {
        // per lookup: stage (bits 8-9: 0 absent, 1 probing, 2 candidate,
        // 3 found) | term offset in the chunk (bits 0-7); bucket start;
        // bucket size << 16 | fingerprint; probe position; candidate rank
        uint32_t tg[G], lo[G], nf[G], p[G], J[G];
        double xj[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          const bool act = mask != 0;
          const uint32_t tt = act ? __builtin_ctzll(mask) : 0;
          mask &= mask - 1;
          tg[g] = tt;
          lo[g] = 0;
          nf[g] = 0;
          p[g] = 0;
          J[g] = 0;
          xj[g] = 0.0;
          if (act) {
            const State<NW> key = term_target<NW>(tl, lds, off, s, t0 + tt);
            uint32_t n;
            bucket_bounds(b.table, bucket_key<NW>(b, key), lo[g], n);
            const uint64_t h = slot_hash<NW>(key);
            nf[g] = (n << 16) | slot_fp(h);
            p[g] = slot_home(h, 2 * n);
            tg[g] |= (n ? 1u : 0u) << 8;
          }
        }
        bool more = true;
        while (more) {
          more = false;
          uint32_t e[G];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            e[g] = (tg[g] >> 8) == 1 ? b.slots[2 * static_cast<uint64_t>(lo[g]) + p[g]]
                                     : kSlotEmpty;
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if ((tg[g] >> 8) == 1) {
              if (e[g] == kSlotEmpty) {
                tg[g] &= 0xFFu;
              } else if ((e[g] >> 16) == (nf[g] & 0xFFFFu)) {
                J[g] = lo[g] + (e[g] & 0xFFFFu);
                tg[g] = (tg[g] & 0xFFu) | (2u << 8);
              } else {
                p[g] = p[g] + 1 == 2 * (nf[g] >> 16) ? 0 : p[g] + 1;
                more = true;
              }
            }
          }
          State<NW> f[G];
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if ((tg[g] >> 8) == 2) {
              f[g] = load_state_cached<NW>(b.states, J[g]);
              xj[g] = x[J[g]];
            }
          }
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if ((tg[g] >> 8) == 2) {
              const State<NW> key = term_target<NW>(tl, lds, off, s, t0 + (tg[g] & 0xFFu));
              if (st_eq<NW>(f[g], key)) {
                tg[g] = (tg[g] & 0xFFu) | (3u << 8);
              } else {
                tg[g] = (tg[g] & 0xFFu) | (1u << 8);
                p[g] = p[g] + 1 == 2 * (nf[g] >> 16) ? 0 : p[g] + 1;
                more = true;
              }
            }
          }
        }
#pragma unroll
        for (int g = 0; g < G; ++g) xj[g] *= xs;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if ((tg[g] >> 8) == 3 && fabs(xj[g]) > kApplyTol) {
            const int t = t0 + (tg[g] & 0xFFu);
            double c = lds ? tl.c[t] : off[t].c * off[t].sgn;
            if constexpr (HAS_Z) {
              if (parity_notz<NW>(s, off[t].Z)) c = -c;
            }
            acc += c * xj[g];
          }
        }
      }
    }
    acc = diag_accumulate<NW>(acc, xi, s, dg, n_diag);
    y[i] = acc;
    if constexpr (DOT) prod = xi * acc;
  }
  if constexpr (DOT) {
    const double bs = block_sum<kApplyBlock>(prod);
    if (threadIdx.x == 0) partials[blk] = bs;
  }
}

// Cached variant: the row's connected ranks come from the operator's cached
// connectivity (qsi_op_cache_connectivity), so an apply is a gather with no
// flip tests or searches.  Same terms, order and arithmetic as the lane
// variant, so y is bit-identical.
//
// ALIGNED: the rows are padded to whole groups of four (conn_pad 4), so every
// group the loop loads belongs to the row and no entry needs masking.
template <int NW, bool DOT, int PIPE, bool ALIGNED>
__global__ __launch_bounds__(kApplyBlock) void k_apply_cached(
    BasisView b, const CoefClass* __restrict__ cls, int n_cls,
    const uint64_t* __restrict__ rowptr, const uint32_t* __restrict__ cj,
    const uint8_t* __restrict__ cm, const DiagTerm* __restrict__ dg, int n_diag,
    const double* __restrict__ x, double* y, const double* y_in, double yscale,
    const double* __restrict__ yscale_dev, double* __restrict__ partials,
    const double* __restrict__ xs_dev, const double* __restrict__ ys_dev) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if (yscale_dev) yscale *= yscale_dev[0];
  // x and y_in are read as xs * x and ys * y_in.  Only the dot-product
  // instances (the Lanczos step) take scales; elsewhere they are the
  // constant 1 and the multiplies fold away.
  const double xs = (DOT && xs_dev) ? xs_dev[0] : 1.0;
  const double ys = (DOT && ys_dev) ? ys_dev[0] : 1.0;
  double cc[kMaxCoefClasses];
#pragma unroll
  for (int k = 0; k < kMaxCoefClasses; ++k) cc[k] = k < n_cls ? cls[k].c : 0.0;
  const int64_t i = static_cast<int64_t>(blk) * kApplyBlock + threadIdx.x;
  double prod = 0.0;
  if (i < b.dim) {
    const double xi = x[i] * xs;
    double acc = y_in ? yscale * (y_in[i] * ys) : 0.0;
    // the row's state (for the diagonal) is loaded before the gathers so its
    // latency overlaps theirs
    const State<NW> si = load_state_cached<NW>(b.states, i);
    const uint64_t e0 = rowptr[i], e1 = rowptr[i + 1];
    // columns and class bytes come four at a time from 16- and 4-byte aligned
    // groups (the arrays are allocated in whole groups); without padding,
    // entries of the neighbouring rows in the first and last group are masked
    // off
    for (uint64_t e = ALIGNED ? e0 : (e0 & ~uint64_t{3}); e < e1; e += PIPE) {
      uint32_t J[PIPE];
      uint8_t mt[PIPE];
      double xj[PIPE];
#pragma unroll
      for (int u = 0; u < PIPE; u += 4) {
        const uint64_t g = e + u;
        const bool in = g < e1;
        const uint4 j4 = in ? reinterpret_cast<const uint4*>(cj)[g >> 2]
                           : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        const uint32_t m4 = in ? reinterpret_cast<const uint32_t*>(cm)[g >> 2] : 0u;
        const uint32_t jv[4] = {j4.x, j4.y, j4.z, j4.w};
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          const bool mine = ALIGNED || (g + k >= e0 && g + k < e1);
          J[u + k] = mine ? jv[k] : 0xFFFFFFFFu;
          mt[u + k] = static_cast<uint8_t>(m4 >> (8 * k));
        }
      }
#pragma unroll
      for (int u = 0; u < PIPE; ++u) xj[u] = J[u] != 0xFFFFFFFFu ? x[J[u]] : 0.0;
      // the scale is applied after every gather of the group has been issued
#pragma unroll
      for (int u = 0; u < PIPE; ++u) xj[u] *= xs;
#pragma unroll
      for (int u = 0; u < PIPE; ++u) {
        if (J[u] != 0xFFFFFFFFu && fabs(xj[u]) > kApplyTol) {
          const int k = mt[u] & 0x7F;
          double c = cc[0];
#pragma unroll
          for (int q = 1; q < kMaxCoefClasses; ++q)
            if (k == q) c = cc[q];
          if (mt[u] & 0x80) c = -c;
          acc += c * xj[u];
        }
      }
    }
    if (n_diag > 0) acc = diag_accumulate<NW>(acc, xi, si, dg, n_diag);
    y[i] = acc;
    if constexpr (DOT) prod = xi * acc;
  }
  if constexpr (DOT) {
    const double bs = block_sum<kApplyBlock>(prod);
    if (threadIdx.x == 0) partials[blk] = bs;
  }
}

// Hash variant (default when the basis has a hash index): as the lane
// variant, but each connected state is found with one probe of the hash index
// (one 16-byte slot, ~1.2 probes on average) instead of a bucket bisection.
template <int NW, bool HAS_Z, bool DOT, int G>
__global__ __launch_bounds__(kApplyBlock) void k_apply_hash(
    BasisView b, HashView hv, const OffTerm* __restrict__ off, int n_off,
    const DiagTerm* __restrict__ dg, int n_diag, const double* __restrict__ x,
    double* y, const double* y_in, double yscale,
    const double* __restrict__ yscale_dev, double* __restrict__ partials,
    const double* __restrict__ xs_dev, const double* __restrict__ ys_dev) {
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if (yscale_dev) yscale *= yscale_dev[0];
  // x and y_in are read as xs * x and ys * y_in.  Only the dot-product
  // instances (the Lanczos step) take scales; elsewhere they are the
  // constant 1 and the multiplies fold away.
  const double xs = (DOT && xs_dev) ? xs_dev[0] : 1.0;
  const double ys = (DOT && ys_dev) ? ys_dev[0] : 1.0;
  const int64_t i = static_cast<int64_t>(blk) * kApplyBlock + threadIdx.x;
  double prod = 0.0;
  if (i < b.dim) {
    const State<NW> s = load_state_cached<NW>(b.states, i);
    const double xi = x[i] * xs;
    double acc = y_in ? yscale * (y_in[i] * ys) : 0.0;
    for (int t0 = 0; t0 < n_off; t0 += 64) {
      const int tn = min(64, n_off - t0);
      uint64_t mask = 0;
      for (int t = 0; t < tn; ++t) {
        mask |= static_cast<uint64_t>(term_ok<NW>(off[t0 + t], s)) << t;
      }
      while (mask) {
        int tt[G];
        bool live[G];
        State<NW> key[G];
        uint64_t pos[G];
        int64_t J[G];
#pragma unroll
        for (int g = 0; g < G; ++g) {
          live[g] = mask != 0;
          tt[g] = live[g] ? t0 + __builtin_ctzll(mask) : t0;
          mask &= mask - 1;
          key[g] = xor_state<NW>(s, off[tt[g]].X);
          pos[g] = hash_home<NW>(key[g], hv.hbits);
          J[g] = -1;
        }
        // interleaved probing
        bool more = true;
        while (more) {
          more = false;
#pragma unroll
          for (int g = 0; g < G; ++g) {
            if (live[g]) {
              uint64_t rk, k0, k1 = 0;
              if constexpr (NW == 1) {
                const ulonglong2 e = hv.slots[pos[g]];
                k0 = e.x;
                rk = e.y;
              } else {
                const ulonglong2 e = hv.slots[2 * pos[g]];
                const ulonglong2 r = hv.slots[2 * pos[g] + 1];
                k0 = e.x;
                k1 = e.y;
                rk = r.x;
              }
              bool eq = (k0 == key[g].w[0]);
              if constexpr (NW == 2) eq = eq && (k1 == key[g].w[1]);
              if (rk == kHashEmpty) {
                live[g] = false;
              } else if (eq) {
                J[g] = static_cast<int64_t>(rk);
                live[g] = false;
              } else {
                pos[g] = (pos[g] + 1) & hv.mask;
                more = true;
              }
            }
          }
        }
        double xj[G];
#pragma unroll
        for (int g = 0; g < G; ++g) xj[g] = (J[g] >= 0) ? x[J[g]] : 0.0;
#pragma unroll
        for (int g = 0; g < G; ++g) xj[g] *= xs;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          if (J[g] >= 0 && fabs(xj[g]) > kApplyTol) {
            const OffTerm& T = off[tt[g]];
            double c = T.c * T.sgn;
            if constexpr (HAS_Z) {
              if (parity_notz<NW>(s, T.Z)) c = -c;
            }
            acc += c * xj[g];
          }
        }
      }
    }
    acc = diag_accumulate<NW>(acc, xi, s, dg, n_diag);
    y[i] = acc;
    if constexpr (DOT) prod = xi * acc;
  }
  if constexpr (DOT) {
    const double bs = block_sum<kApplyBlock>(prod);
    if (threadIdx.x == 0) partials[blk] = bs;
  }
}

// ----------------------------------------------------------------------------
// Window variant (default).  A wavefront owns 64 consecutive rows.  For one
// term, the rows it connects share the term's precondition bits, so their
// targets s ^ X are increasing in lane order and fall in a short run of the
// sorted basis (median ~40 ranks on the 64-spin ice basis).  Per chunk of 64
// terms the wavefront
//   1. tests every term on every row; a ballot gives each term's lane mask,
//   2. finds each term's anchor rank (lower bound of its first target) with
//      the 64 lookups spread over the 64 lanes,
//   3. per term, stages basis[r .. r+63] and x[r .. r+63] in LDS with one
//      coalesced load each, and every connected lane bisects the staged window
//      (targets past the window fall back to a full lookup).
// Contributions are added in term order, then the diagonal, exactly as in the
// lane variant, so both variants give bit-identical y.
// ----------------------------------------------------------------------------

constexpr int kWavesPerBlock = kApplyBlock / 64;


__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t readlane64(uint64_t v, int lane) {
  const uint32_t lo = __builtin_amdgcn_readlane(static_cast<uint32_t>(v), lane);
  const uint32_t hi = __builtin_amdgcn_readlane(static_cast<uint32_t>(v >> 32), lane);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

template <int NW, bool HAS_Z, bool DOT, int P>
__global__ __launch_bounds__(kApplyBlock) void k_apply_win(
    BasisView b, const OffTerm* __restrict__ off, int n_off,
    const DiagTerm* __restrict__ dg, int n_diag, const double* __restrict__ x,
    double* y, const double* y_in, double yscale,
    const double* __restrict__ yscale_dev, double* __restrict__ partials,
    const double* __restrict__ xs_dev, const double* __restrict__ ys_dev) {
  __shared__ uint64_t rows_lds[kWavesPerBlock][64 * NW];
  __shared__ uint64_t win_lds[kWavesPerBlock][P][64 * NW];
  __shared__ double xwin_lds[kWavesPerBlock][P][64];
  const int lane = threadIdx.x & 63;
  const int wid = threadIdx.x >> 6;
  const uint32_t blk = xcd_remap(blockIdx.x, gridDim.x);
  if (yscale_dev) yscale *= yscale_dev[0];
  // x and y_in are read as xs * x and ys * y_in.  Only the dot-product
  // instances (the Lanczos step) take scales; elsewhere they are the
  // constant 1 and the multiplies fold away.
  const double xs = (DOT && xs_dev) ? xs_dev[0] : 1.0;
  const double ys = (DOT && ys_dev) ? ys_dev[0] : 1.0;
  const int64_t i = static_cast<int64_t>(blk) * kApplyBlock + threadIdx.x;
  const bool row_ok = i < b.dim;
  const State<NW> s = load_state_cached<NW>(b.states, row_ok ? i : 0);
#pragma unroll
  for (int w = 0; w < NW; ++w) rows_lds[wid][lane * NW + w] = s.w[w];
  double acc = 0.0;
  double xi = 0.0;
  if (row_ok) {
    xi = x[i] * xs;
    acc = y_in ? yscale * (y_in[i] * ys) : 0.0;
  }
  wave_sync();
  for (int t0 = 0; t0 < n_off; t0 += 64) {
    const int tn = min(64, n_off - t0);
    // 1. lane masks: lane k keeps the mask of term t0 + k
    uint64_t my_mask = 0;
    for (int k = 0; k < tn; ++k) {
      const uint64_t m = __ballot(row_ok && term_ok<NW>(off[t0 + k], s));
      my_mask = (lane == k) ? m : my_mask;
    }
    // 2. anchors: lower bound of the first connected row's target, one
    //    lookup per lane, all in flight together
    uint32_t anchor = 0;
    if (lane < tn && my_mask != 0) {
      const int a = __builtin_ctzll(my_mask);
      State<NW> key;
#pragma unroll
      for (int w = 0; w < NW; ++w) key.w[w] = rows_lds[wid][a * NW + w] ^ off[t0 + lane].X[w];
      anchor = static_cast<uint32_t>(basis_lower_bound<NW>(b, key));
    }
    // 3. groups of P terms: stage P windows with back-to-back loads, then run
    //    the P bisections interleaved
    for (int kb = 0; kb < tn; kb += P) {
      uint64_t gm[P];
      int64_t gr[P];
      uint64_t any = 0;
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const int k = kb + j;
        gm[j] = (k < tn) ? readlane64(my_mask, k) : 0;
        gr[j] = gm[j] ? static_cast<int64_t>(__builtin_amdgcn_readlane(anchor, k)) : 0;
        any |= gm[j];
      }
      if (any == 0) continue;
      State<NW> wv[P];
      double xv[P];
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const int64_t idx = gr[j] + lane;
        if (gm[j] != 0 && idx < b.dim) {
          wv[j] = load_state_cached<NW>(b.states, idx);
          xv[j] = x[idx] * xs;
        }
      }
#pragma unroll
      for (int j = 0; j < P; ++j) {
        const int64_t idx = gr[j] + lane;
        if (gm[j] != 0 && idx < b.dim) {
#pragma unroll
          for (int w = 0; w < NW; ++w) win_lds[wid][j][lane * NW + w] = wv[j].w[w];
          xwin_lds[wid][j][lane] = xv[j];
        }
      }
      wave_sync();
      int lo[P], n[P], nv[P];
#pragma unroll
      for (int j = 0; j < P; ++j) {
        nv[j] = gm[j] ? static_cast<int>(min<int64_t>(64, b.dim - gr[j])) : 0;
        lo[j] = 0;
        n[j] = ((gm[j] >> lane) & 1) ? nv[j] : 0;
      }
      // lower-bound bisection of each key in its window (<= 7 rounds)
#pragma unroll
      for (int step = 0; step < 7; ++step) {
#pragma unroll
        for (int j = 0; j < P; ++j) {
          if (n[j] > 0) {
            const State<NW> key = xor_state<NW>(s, off[t0 + kb + j].X);
            const int half = n[j] >> 1;
            State<NW> m;
#pragma unroll
            for (int w = 0; w < NW; ++w) m.w[w] = win_lds[wid][j][(lo[j] + half) * NW + w];
            if (st_lt<NW>(m, key)) {
              lo[j] += half + 1;
              n[j] -= half + 1;
            } else {
              n[j] = half;
            }
          }
        }
      }
      // contributions in term order
#pragma unroll
      for (int j = 0; j < P; ++j) {
        if ((gm[j] >> lane) & 1) {
          const OffTerm& T = off[t0 + kb + j];
          const State<NW> key = xor_state<NW>(s, T.X);
          bool found = false;
          double xj = 0.0;
          if (lo[j] < nv[j]) {
            State<NW> m;
#pragma unroll
            for (int w = 0; w < NW; ++w) m.w[w] = win_lds[wid][j][lo[j] * NW + w];
            if (st_eq<NW>(m, key)) {
              found = true;
              xj = xwin_lds[wid][j][lo[j]];
            }
          } else {
            const int64_t J = basis_lookup<NW>(b, key);
            if (J >= 0) {
              found = true;
              xj = x[J] * xs;
            }
          }
          if (found && fabs(xj) > kApplyTol) {
            double c = T.c * T.sgn;
            if constexpr (HAS_Z) {
              if (parity_notz<NW>(s, T.Z)) c = -c;
            }
            acc += c * xj;
          }
        }
      }
      wave_sync();
    }
  }
  double prod = 0.0;
  if (row_ok) {
    acc = diag_accumulate<NW>(acc, xi, s, dg, n_diag);
    y[i] = acc;
    if constexpr (DOT) prod = xi * acc;
  }
  if constexpr (DOT) {
    const double bs = block_sum<kApplyBlock>(prod);
    if (threadIdx.x == 0) partials[blk] = bs;
  }
}


// ============================================================================
// Connectivity (row-major triplets of the off-diagonal part + diagonal)
// ============================================================================

template <int NW>
__global__ void k_conn(BasisView b, const OffTerm* __restrict__ off, int n_off,
                       const DiagTerm* __restrict__ dg, int n_diag,
                       const int64_t* __restrict__ offsets,
                       int64_t* __restrict__ counts, int64_t* __restrict__ rows,
                       int64_t* __restrict__ cols, double* __restrict__ vals,
                       double* __restrict__ diag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.dim) return;
  const State<NW> s = load_state_cached<NW>(b.states, i);
  int64_t pos = offsets ? offsets[i] : 0;
  int64_t cnt = 0;
  for (int t = 0; t < n_off; ++t) {
    const OffTerm& T = off[t];
    if (!term_ok<NW>(T, s)) continue;
    const int64_t J = basis_lookup<NW>(b, xor_state<NW>(s, T.X));
    if (J < 0) continue;
    double c = T.c * T.sgn;
    if (parity_notz<NW>(s, T.Z)) c = -c;
    if (!(fabs(c) > kSparseTol)) continue;
    if (offsets) {
      rows[pos] = i;
      cols[pos] = J;
      vals[pos] = c;
      ++pos;
    }
    ++cnt;
  }
  if (counts) counts[i] = cnt;
  if (diag) diag[i] = diag_accumulate<NW>(0.0, 1.0, s, dg, n_diag);
}

// ============================================================================
// Streaming FP64 vector kernels
// ============================================================================

constexpr int kVecBlock = 256;

inline int vec_blocks(int64_t n) {
  int64_t nb = (n + kVecBlock * 4 - 1) / (kVecBlock * 4);
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(nb, 8192)));
}

// Workgroups of the projection passes (multi-dots over up to 64 columns):
// about 16 rows per thread, but at least 1024 workgroups while there are 4
// rows per thread to give them, so each workgroup's column sums (one block
// reduction per column) are spread over more rows than in a plain vector
// kernel and the chip still fills.
inline int proj_blocks(int64_t n) {
  const int64_t want = (n + kVecBlock * 16 - 1) / (kVecBlock * 16);
  return static_cast<int>(std::min<int64_t>(vec_blocks(n), std::max<int64_t>(want, 1024)));
}

// The grid-stride dot, axpy-norm and device-scalar scale kernels take
// kVecUnroll of a thread's elements per round, all loads first, and
// accumulate in element order (the sums are the ones a one-element loop
// gives).
constexpr int kVecUnroll = 8;

__global__ __launch_bounds__(kVecBlock) void k_dot_partial(
    const double* __restrict__ a, const double* __restrict__ b, int64_t n,
    double* __restrict__ partials) {
  double v = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBlock;
  int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
  for (; i + (kVecUnroll - 1) * stride < n; i += kVecUnroll * stride) {
    double av[kVecUnroll], bv[kVecUnroll];
#pragma unroll
    for (int k = 0; k < kVecUnroll; ++k) {
      av[k] = a[i + k * stride];
      bv[k] = b[i + k * stride];
    }
#pragma unroll
    for (int k = 0; k < kVecUnroll; ++k) v += av[k] * bv[k];
  }
  for (; i < n; i += stride) v += a[i] * b[i];
  v = block_sum<kVecBlock>(v);
  if (threadIdx.x == 0) partials[blockIdx.x] = v;
}

__global__ __launch_bounds__(kVecBlock) void k_axpy(double alpha,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ y,
                                                    int64_t n,
                                                    const double* __restrict__ xs_dev) {
  const double xs = xs_dev ? xs_dev[0] : 1.0;  // x is read as xs * x
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock)
    y[i] += alpha * (x[i] * xs);
}

__global__ __launch_bounds__(kVecBlock) void k_scal(double alpha,
                                                    double* __restrict__ x,
                                                    int64_t n) {
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock)
    x[i] *= alpha;
}

// u -= (*alpha) v, partials[b] = sum u^2 over the workgroup's share
__global__ __launch_bounds__(kVecBlock) void k_axpy_dev_norm(
    const double* __restrict__ alpha, const double* __restrict__ v,
    double* __restrict__ u, int64_t n, double* __restrict__ partials,
    const double* __restrict__ vs_dev) {
  const double a = alpha[0];
  const double vs = vs_dev ? vs_dev[0] : 1.0;  // v is read as vs * v
  double acc = 0.0;
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBlock;
  int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
  for (; i + (kVecUnroll - 1) * stride < n; i += kVecUnroll * stride) {
    double uv[kVecUnroll], vv[kVecUnroll];
#pragma unroll
    for (int k = 0; k < kVecUnroll; ++k) {
      uv[k] = u[i + k * stride];
      vv[k] = v[i + k * stride] * vs;
    }
#pragma unroll
    for (int k = 0; k < kVecUnroll; ++k) {
      const double w = uv[k] - a * vv[k];
      u[i + k * stride] = w;
      acc += w * w;
    }
  }
  for (; i < n; i += stride) {
    const double w = u[i] - a * (v[i] * vs);
    u[i] = w;
    acc += w * w;
  }
  acc = block_sum<kVecBlock>(acc);
  if (threadIdx.x == 0) partials[blockIdx.x] = acc;
}

// ---- tall-skinny kernels for full reorthogonalisation ----------------------
// V is column-major with leading dimension ldv (column j at V + j*ldv).

constexpr int kMaxCols = 64;

// partials[blk * m + j] = sum over the workgroup's rows of V[:, j] * w
__global__ __launch_bounds__(kVecBlock) void k_multidot(
    const double* __restrict__ V, int64_t ldv, int m, const double* __restrict__ w,
    int64_t n, double* __restrict__ partials) {
  __shared__ double red[kVecBlock / kWave][16];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  for (int j0 = 0; j0 < m; j0 += 16) {
    const int jn = min(16, m - j0);
    double acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = 0.0;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
         i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock) {
      // every load of the row first, then unpredicated updates (absent
      // columns load as 0 and leave their unused sums unchanged), so a lane
      // keeps all its column loads in flight
      const double wi = w[i];
      double v[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) v[j] = j < jn ? V[(j0 + j) * ldv + i] : 0.0;
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] += v[j] * wi;
    }
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const double v = wave_sum(acc[j]);
      if (lane == 0) red[wid][j] = v;
    }
    __syncthreads();
    if (threadIdx.x < jn) {
      double r = 0.0;
      for (int k = 0; k < kVecBlock / kWave; ++k) r += red[k][threadIdx.x];
      partials[static_cast<int64_t>(blockIdx.x) * m + j0 + threadIdx.x] = r;
    }
    __syncthreads();
  }
}

// out[j] = sum_b partials[b * m + j], one workgroup per column j, fixed order
__global__ __launch_bounds__(256) void k_finalize_cols(
    const double* __restrict__ partials, int nb, int m, double* __restrict__ out,
    double* __restrict__ acc_out) {
  const int j = blockIdx.x;
  double v = 0.0;
  for (int b = threadIdx.x; b < nb; b += 256) v += partials[static_cast<int64_t>(b) * m + j];
  v = block_sum<256>(v);
  if (threadIdx.x == 0) {
    out[j] = v;
    if (acc_out) acc_out[j] += v;
  }
}

// w -= V[:, 0:m] h
__global__ __launch_bounds__(kVecBlock) void k_gemv_sub(
    const double* __restrict__ V, int64_t ldv, int m, const double* __restrict__ h,
    double* __restrict__ w, int64_t n) {
  __shared__ double hs[kMaxCols];
  if (threadIdx.x < m) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += V[j * ldv + i] * hs[j];
    w[i] -= acc;
  }
}

// Second Gram-Schmidt pass fused with the check for a third: w -= V[:, 0:m] h
// (the same sum order as k_gemv_sub), and from the same read of V the
// per-workgroup partials of V[:, 0:m]^T w_new (columns 0..m-1) and of
// w_new . w_new (column m), m <= MAXC.
template <int MAXC>
__global__ __launch_bounds__(kVecBlock) void k_gemv_sub_dots(
    const double* __restrict__ V, int64_t ldv, int m, const double* __restrict__ h,
    double* __restrict__ w, int64_t n, double* __restrict__ partials) {
  __shared__ double hs[MAXC];
  __shared__ double red[kVecBlock / kWave][MAXC + 1];
  const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
  if (threadIdx.x < MAXC) hs[threadIdx.x] = threadIdx.x < m ? h[threadIdx.x] : 0.0;
  __syncthreads();
  double acc[MAXC + 1];
#pragma unroll
  for (int j = 0; j <= MAXC; ++j) acc[j] = 0.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock) {
    // every load of the row first (absent columns load as 0, and their h is
    // 0, so the sum and the dots are unchanged), then the sums in column
    // order, so a lane keeps all its column loads in flight
    double vr[MAXC];
#pragma unroll
    for (int j = 0; j < MAXC; ++j) vr[j] = j < m ? V[j * ldv + i] : 0.0;
    const double wo = w[i];
    double sum = 0.0;
#pragma unroll
    for (int j = 0; j < MAXC; ++j) sum += vr[j] * hs[j];
    const double wn = wo - sum;
    w[i] = wn;
#pragma unroll
    for (int j = 0; j < MAXC; ++j) acc[j] += vr[j] * wn;
    acc[MAXC] += wn * wn;
  }
#pragma unroll
  for (int j = 0; j <= MAXC; ++j) {
    const double v = wave_sum(acc[j]);
    if (lane == 0) red[wid][j] = v;
  }
  __syncthreads();
  for (int j = threadIdx.x; j <= m; j += kVecBlock) {
    const int src = j == m ? MAXC : j;
    double r = 0.0;
    for (int k = 0; k < kVecBlock / kWave; ++k) r += red[k][src];
    partials[static_cast<int64_t>(blockIdx.x) * (m + 1) + j] = r;
  }
}

// After the fused pass: g[0..m) = V^T w, g[m] = w . w.  A third pass is
// needed when |V^T w| > 1e-12 |w| (flag[0] = 1); its correction is added to
// the projected column hcol[0..m) and the norm follows it.  beta[0] = |w|,
// beta[1] = 1 / |w| (0 for a vanishing w).
__global__ void k_gs_tail(const double* __restrict__ g, int m, double* __restrict__ hcol,
                          int* __restrict__ flag, double* __restrict__ beta) {
  __shared__ int need;
  if (threadIdx.x == 0) {
    double g2 = 0.0;
    for (int j = 0; j < m; ++j) g2 += g[j] * g[j];
    const double q = g[m];
    need = (q > 0.0 && g2 > 1e-24 * q) ? 1 : 0;
    const double b2 = need ? q - g2 : q;
    const double b = b2 > 0.0 ? sqrt(b2) : 0.0;
    beta[0] = b;
    beta[1] = b > 0.0 ? 1.0 / b : 0.0;
    flag[0] = need;
  }
  __syncthreads();
  if (need)
    for (int j = threadIdx.x; j < m; j += blockDim.x) hcol[j] += g[j];
}

// The conditional third pass fused with the normalisation: w -= V[:, 0:m] h
// when flag[0] is set (a plain column-order sum), then w *= a = beta[1], the
// scale k_gs_tail computed from the pass-2 sums, in one pass over w.
__global__ __launch_bounds__(kVecBlock) void k_gemv_sub_scal(
    const double* __restrict__ V, int64_t ldv, int m, const double* __restrict__ h,
    double* __restrict__ w, int64_t n, const int* __restrict__ flag,
    const double* __restrict__ beta) {
  const bool need = flag[0] != 0;
  const double a = beta[1];
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kVecBlock;
  int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
  if (!need) {
    for (; i + (kVecUnroll - 1) * stride < n; i += kVecUnroll * stride) {
      double xv[kVecUnroll];
#pragma unroll
      for (int k = 0; k < kVecUnroll; ++k) xv[k] = w[i + k * stride];
#pragma unroll
      for (int k = 0; k < kVecUnroll; ++k) w[i + k * stride] = xv[k] * a;
    }
    for (; i < n; i += stride) w[i] *= a;
    return;
  }
  __shared__ double hs[kMaxCols];
  if (threadIdx.x < m) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  for (; i < n; i += stride) {
    double acc = 0.0;
    for (int j = 0; j < m; ++j) acc += V[j * ldv + i] * hs[j];
    w[i] = (w[i] - acc) * a;
  }
}

// out[:, c] = V[:, 0:m] Y[:, c] for c < k (Y column-major m x k, on device);
// output columns in chunks of 8 so V is read ceil(k/8) times.
__global__ __launch_bounds__(kVecBlock) void k_combine(
    const double* __restrict__ V, int64_t ldv, int m, const double* __restrict__ Y,
    int k, double* __restrict__ out, int64_t ldo, int64_t n) {
  __shared__ double ys[kMaxCols * kMaxCols];
  for (int t = threadIdx.x; t < m * k; t += kVecBlock) ys[t] = Y[t];
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock) {
    for (int c0 = 0; c0 < k; c0 += 8) {
      double acc[8];
#pragma unroll
      for (int c = 0; c < 8; ++c) acc[c] = 0.0;
      for (int j = 0; j < m; ++j) {
        const double vj = V[j * ldv + i];
#pragma unroll
        for (int c = 0; c < 8; ++c)
          if (c0 + c < k) acc[c] += vj * ys[(c0 + c) * m + j];
      }
#pragma unroll
      for (int c = 0; c < 8; ++c)
        if (c0 + c < k) out[(c0 + c) * ldo + i] = acc[c];
    }
  }
}

// x[i] = s_dev[0] * y[i]
__global__ __launch_bounds__(kVecBlock) void k_copy_scaled(
    const double* __restrict__ s, const double* __restrict__ y,
    double* __restrict__ x, int64_t n) {
  const double a = s[0];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock)
    x[i] = a * y[i];
}

// y += s_dev[0] * c * x
__global__ __launch_bounds__(kVecBlock) void k_axpy_dev(
    const double* __restrict__ s, double c, const double* __restrict__ x,
    double* __restrict__ y, int64_t n) {
  const double a = c * s[0];
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock)
    y[i] += a * x[i];
}

// Psi[:, b] += c[b] * v for b < nb (Psi column-major, leading dimension ld):
// one read of v feeds every column.  v is read as vs * v.
__global__ __launch_bounds__(kVecBlock) void k_accum_cols(
    const double* __restrict__ v, double* __restrict__ Psi, int64_t ld, int nb,
    const double* __restrict__ c, int64_t n, const double* __restrict__ vs_dev) {
  const double vs = vs_dev ? vs_dev[0] : 1.0;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kVecBlock + threadIdx.x;
       i < n; i += static_cast<int64_t>(gridDim.x) * kVecBlock) {
    const double vi = v[i] * vs;
    for (int b = 0; b < nb; ++b) Psi[b * ld + i] += c[b] * vi;
  }
}

// ============================================================================
// Host helpers
// ============================================================================

namespace {
thread_local std::string g_err;
}

void set_error(const std::string& msg) { g_err = msg; }
const char* get_error() { return g_err.c_str(); }

void basis_retain(qsi_basis_s* b) { b->refs.fetch_add(1); }
void basis_release(qsi_basis_s* b) {
  if (b->refs.fetch_sub(1) == 1) {
    if (b->d_states) (void)hipFree(b->d_states);
    if (b->d_table) (void)hipFree(b->d_table);
    if (b->d_hash) (void)hipFree(b->d_hash);
    if (b->d_slots) (void)hipFree(b->d_slots);
    if (b->d_lowkeys) (void)hipFree(b->d_lowkeys);
    delete b;
  }
}

int64_t apply_num_blocks(const qsi_op_s* op) {
  return (op->basis->dim + kApplyBlock - 1) / kApplyBlock;
}

template <int NW, bool DOT, bool ALIGNED>
static void launch_cached_t(const qsi_op_s* op, dim3 grid, dim3 block, const BasisView& v,
                            const double* x, double* y, const double* y_in, double yscale,
                            const double* yscale_dev, double* partials, const double* xs_dev,
                            const double* ys_dev, hipStream_t s) {
  auto k = op->cached_pipe == 32 ? k_apply_cached<NW, DOT, 32, ALIGNED>
         : op->cached_pipe == 4  ? k_apply_cached<NW, DOT, 4, ALIGNED>
         : op->cached_pipe == 16 ? k_apply_cached<NW, DOT, 16, ALIGNED>
                                 : k_apply_cached<NW, DOT, 8, ALIGNED>;
  hipLaunchKernelGGL(k, grid, block, 0, s, v, op->d_cls, op->n_cls, op->d_rowptr, op->d_cj,
                     op->d_cm, op->d_diag, op->n_diag, x, y, y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
}

template <int NW, bool HAS_Z, bool DOT>
static void launch_apply_t(const qsi_op_s* op, const double* x, double* y,
                           const double* y_in, double yscale,
                           const double* yscale_dev, double* partials,
                           const double* xs_dev, const double* ys_dev, hipStream_t s) {
  const int64_t nb = apply_num_blocks(op);
  const dim3 grid(static_cast<uint32_t>(nb)), block(kApplyBlock);
  const auto v = op->basis->view();
  int variant = op->variant;
  // AUTO: the cached gathers when the connectivity is cached, else LANE4 (the
  // fastest search kernel measured on the 64- and 96-spin bases, with or
  // without the slot index and low keys; the autotuner times them all)
  // AUTO: the cached connectivity when it exists, else the low-key kernel
  // (eight lookups per lane once the candidates need no compare, four
  // otherwise), else the lane kernel
  if (variant == QSI_APPLY_AUTO)
    variant = op->conn_nnz >= 0         ? QSI_APPLY_CACHED
              : !op->basis->d_lowkeys   ? QSI_APPLY_LANE4
              : op->lowkeys_check == kLowTrusted ? QSI_APPLY_LOW8
                                                 : QSI_APPLY_LOW4;
  // a slot or low-key variant chosen while the basis had that index, which
  // has been dropped since, falls back to the lane kernel (LANE4 for SLOT4,
  // LANE2 for SLOT2 and SLOT3; the same depth for LOW*)
  if (!op->basis->d_slots && variant >= QSI_APPLY_SLOT2 && variant <= QSI_APPLY_SLOT3)
    variant = variant == QSI_APPLY_SLOT4 ? QSI_APPLY_LANE4 : QSI_APPLY_LANE2;
  if (!op->basis->d_lowkeys && variant >= QSI_APPLY_LOW2 && variant <= QSI_APPLY_LOW8)
    variant = variant == QSI_APPLY_LOW2 ? QSI_APPLY_LANE2
            : variant == QSI_APPLY_LOW4 ? QSI_APPLY_LANE4 : QSI_APPLY_LANE8;
  switch (variant) {
    case QSI_APPLY_CACHED:
      // gathers in flight per lane: op->cached_pipe (4, 8, 16 or 32; autotuned)
      if (op->conn_pad == 4)
        launch_cached_t<NW, DOT, true>(op, grid, block, v, x, y, y_in, yscale, yscale_dev,
                                       partials, xs_dev, ys_dev, s);
      else
        launch_cached_t<NW, DOT, false>(op, grid, block, v, x, y, y_in, yscale, yscale_dev,
                                        partials, xs_dev, ys_dev, s);
      break;
    case QSI_APPLY_LANE:
      hipLaunchKernelGGL((k_apply_lane<NW, HAS_Z, DOT, 1>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_LANE2:
      hipLaunchKernelGGL((k_apply_lane<NW, HAS_Z, DOT, 2>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_WINDOW:
      hipLaunchKernelGGL((k_apply_win<NW, HAS_Z, DOT, 4>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_LANE8:
      hipLaunchKernelGGL((k_apply_lane<NW, HAS_Z, DOT, 8>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_HASH:
      hipLaunchKernelGGL((k_apply_hash<NW, HAS_Z, DOT, 1>), grid, block, 0, s, v,
                         op->basis->hview(), op->d_off, op->n_off, op->d_diag,
                         op->n_diag, x, y, y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_HASH4:
      hipLaunchKernelGGL((k_apply_hash<NW, HAS_Z, DOT, 4>), grid, block, 0, s, v,
                         op->basis->hview(), op->d_off, op->n_off, op->d_diag,
                         op->n_diag, x, y, y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_SLOT2:
      hipLaunchKernelGGL((k_apply_slot<NW, HAS_Z, DOT, 2>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_SLOT4:
      hipLaunchKernelGGL((k_apply_slot<NW, HAS_Z, DOT, 4>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_SLOT3:
      hipLaunchKernelGGL((k_apply_slot<NW, HAS_Z, DOT, 3>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_LOW2:
      if (op->lowkeys_check == kLowTrusted)
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 2, kLowTrusted>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      else if (op->lowkeys_check == kLowCmpKey)
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 2, kLowCmpKey>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      else
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 2, kLowCmpState>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_LOW4:
      if (op->lowkeys_check == kLowTrusted)
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 4, kLowTrusted>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      else if (op->lowkeys_check == kLowCmpKey)
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 4, kLowCmpKey>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      else
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 4, kLowCmpState>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      break;
    case QSI_APPLY_LOW8:
      if (op->lowkeys_check == kLowTrusted)
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 8, kLowTrusted>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      else if (op->lowkeys_check == kLowCmpKey)
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 8, kLowCmpKey>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      else
        hipLaunchKernelGGL((k_apply_low<NW, HAS_Z, DOT, 8, kLowCmpState>), grid, block, 0, s,
                           v, op->d_off, op->n_off, op->d_diag, op->n_diag, x, y, y_in, yscale,
                           yscale_dev, partials, xs_dev, ys_dev);
      break;
    default:  // QSI_APPLY_LANE4
      hipLaunchKernelGGL((k_apply_lane<NW, HAS_Z, DOT, 4>), grid, block, 0, s, v,
                         op->d_off, op->n_off, op->d_diag, op->n_diag, x, y,
                         y_in, yscale, yscale_dev, partials, xs_dev, ys_dev);
  }

}

int launch_apply(const qsi_op_s* op, const double* x, double* y,
                 const double* y_in, double yscale, double* dot_partials,
                 double* dot_out, hipStream_t s) {
  return launch_apply_dev(op, x, y, y_in, yscale, nullptr, dot_partials,
                          dot_out, s);
}

int launch_apply_dev(const qsi_op_s* op, const double* x, double* y,
                     const double* y_in, double yscale,
                     const double* yscale_dev, double* dot_partials,
                     double* dot_out, hipStream_t s) {
  return launch_apply_scaled(op, x, y, y_in, yscale, yscale_dev, nullptr, nullptr, dot_partials,
                             dot_out, s);
}

// Sums n per-workgroup partials into out (mode as in k_finalize_sum) in a
// fixed order.  Arrays longer than 8 chunks are first reduced chunk by chunk
// in place (k_sum_chunks), so a single workgroup never streams millions of
// partials; the partials are clobbered.
static void launch_finalize(double* partials, int64_t n, double* out, int mode,
                            hipStream_t s) {
  constexpr int64_t kChunk = 4096;
  if (n > 8 * kChunk) {
    const int64_t g = (n + kChunk - 1) / kChunk;
    hipLaunchKernelGGL(k_sum_chunks<256>, dim3(static_cast<uint32_t>(g)), dim3(256), 0, s,
                       partials, n, kChunk);
    hipLaunchKernelGGL(k_finalize_sum<256>, dim3(1), dim3(256), 0, s, partials, g, out, mode,
                       kChunk);
  } else {
    hipLaunchKernelGGL(k_finalize_sum<256>, dim3(1), dim3(256), 0, s, partials, n, out, mode,
                       int64_t{1});
  }
}

int launch_apply_scaled(const qsi_op_s* op, const double* x, double* y,
                        const double* y_in, double yscale, const double* yscale_dev,
                        const double* xs_dev, const double* ys_dev, double* dot_partials,
                        double* dot_out, hipStream_t s) {
  const bool dot = dot_partials != nullptr;
  // the kernels honour the scales only in their dot-product instances
  QSI_REQUIRE(dot || (xs_dev == nullptr && ys_dev == nullptr),
              "launch_apply_scaled: scales need the dot product");
  const int nw = op->basis->nw;
  const bool hz = op->has_z;
#define QSI_LAUNCH(NW_, HZ_, D_) \
  launch_apply_t<NW_, HZ_, D_>(op, x, y, y_in, yscale, yscale_dev, dot_partials, xs_dev, \
                               ys_dev, s)
  if (nw == 1) {
    if (hz) { if (dot) QSI_LAUNCH(1, true, true); else QSI_LAUNCH(1, true, false); }
    else    { if (dot) QSI_LAUNCH(1, false, true); else QSI_LAUNCH(1, false, false); }
  } else {
    if (hz) { if (dot) QSI_LAUNCH(2, true, true); else QSI_LAUNCH(2, true, false); }
    else    { if (dot) QSI_LAUNCH(2, false, true); else QSI_LAUNCH(2, false, false); }
  }
#undef QSI_LAUNCH
  QSI_CHECK_LAUNCH();
  if (dot) {
    launch_finalize(dot_partials, apply_num_blocks(op), dot_out, 0, s);
    QSI_CHECK_LAUNCH();
  }
  return QSI_OK;
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Gaussian entries from a counter-based generator: Box-Muller on two
// splitmix64 (Steele, Lea & Flood 2014) uniforms per entry, so v[i] depends
// on (seed, i) only, whatever the launch shape.
__global__ void k_randn(double* __restrict__ v, int64_t n, uint64_t seed) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t a = splitmix64(seed ^ splitmix64(2 * static_cast<uint64_t>(i)));
  const uint64_t b = splitmix64(seed ^ splitmix64(2 * static_cast<uint64_t>(i) + 1));
  const double u1 = (static_cast<double>(a >> 11) + 0.5) * 0x1.0p-53;
  const double u2 = (static_cast<double>(b >> 11) + 0.5) * 0x1.0p-53;
  v[i] = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

int launch_randn(double* v, int64_t n, uint64_t seed, hipStream_t s) {
  if (n <= 0) return QSI_OK;
  hipLaunchKernelGGL(k_randn, dim3(static_cast<uint32_t>((n + 255) / 256)), dim3(256), 0, s, v,
                     n, seed);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_dot(const double* a, const double* b, int64_t n, double* partials,
               double* out, hipStream_t s) {
  const int nb = vec_blocks(n);
  hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kVecBlock), 0, s, a, b, n,
                     partials);
  QSI_CHECK_LAUNCH();
  launch_finalize(partials, nb, out, 0, s);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_axpy(double alpha, const double* x, double* y, int64_t n,
                hipStream_t s, const double* xs_dev) {
  hipLaunchKernelGGL(k_axpy, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s, alpha,
                     x, y, n, xs_dev);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_scal(double alpha, double* x, int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_scal, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s, alpha,
                     x, n);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

// u -= alpha_dev * v; nrm_out[0] = ||u||, nrm_out[1] = 1/||u||
int launch_axpy_norm(const double* alpha_dev, const double* v, double* u,
                     int64_t n, double* partials, double* nrm_out,
                     hipStream_t s, const double* vs_dev) {
  const int nb = vec_blocks(n);
  hipLaunchKernelGGL(k_axpy_dev_norm, dim3(nb), dim3(kVecBlock), 0, s,
                     alpha_dev, v, u, n, partials, vs_dev);
  QSI_CHECK_LAUNCH();
  launch_finalize(partials, nb, nrm_out, 1, s);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int max_vec_partials() { return 8192; }
int multidot_partials() { return 1 << 17; }

int launch_multidot(const double* V, int64_t ldv, int m, const double* w,
                    int64_t n, double* partials, double* out, double* acc_out,
                    hipStream_t s) {
  if (m <= 0) return QSI_OK;
  if (m > kMaxCols) {
    set_error("launch_multidot: too many columns");
    return QSI_E_INVALID;
  }
  // one partial per (workgroup, column): the caller's buffer holds
  // multidot_partials() doubles
  const int nb = std::max(1, std::min(proj_blocks(n), multidot_partials() / std::max(m, 1)));
  hipLaunchKernelGGL(k_multidot, dim3(nb), dim3(kVecBlock), 0, s, V, ldv, m, w,
                     n, partials);
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finalize_cols, dim3(m), dim3(256), 0, s, partials, nb, m,
                     out, acc_out);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_gemv_sub(const double* V, int64_t ldv, int m, const double* h,
                    double* w, int64_t n, hipStream_t s) {
  if (m <= 0) return QSI_OK;
  hipLaunchKernelGGL(k_gemv_sub, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s, V,
                     ldv, m, h, w, n);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_gs_reorth(const double* V, int64_t ldv, int m, const double* h, double* w,
                     int64_t n, double* partials, double* g, double* hcol, int* flag,
                     double* beta, hipStream_t s) {
  if (m <= 0 || m > kMaxCols) {
    set_error("launch_gs_reorth: column count out of range");
    return QSI_E_INVALID;
  }
  const int nb = std::max(1, std::min(proj_blocks(n), multidot_partials() / (m + 1)));
  if (m <= 16)
    hipLaunchKernelGGL(k_gemv_sub_dots<16>, dim3(nb), dim3(kVecBlock), 0, s, V, ldv, m, h, w, n,
                       partials);
  else if (m <= 32)
    hipLaunchKernelGGL(k_gemv_sub_dots<32>, dim3(nb), dim3(kVecBlock), 0, s, V, ldv, m, h, w, n,
                       partials);
  else
    hipLaunchKernelGGL(k_gemv_sub_dots<64>, dim3(nb), dim3(kVecBlock), 0, s, V, ldv, m, h, w, n,
                       partials);
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_finalize_cols, dim3(m + 1), dim3(256), 0, s, partials, nb, m + 1, g,
                     static_cast<double*>(nullptr));
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_gs_tail, dim3(1), dim3(64), 0, s, g, m, hcol, flag, beta);
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_gemv_sub_scal, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s, V, ldv, m, g,
                     w, n, flag, beta);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_combine(const double* V, int64_t ldv, int m, const double* Y, int k,
                   double* out, int64_t ldo, int64_t n, hipStream_t s) {
  if (m > kMaxCols || k > kMaxCols) {
    set_error("launch_combine: too many columns");
    return QSI_E_INVALID;
  }
  hipLaunchKernelGGL(k_combine, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s, V,
                     ldv, m, Y, k, out, ldo, n);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_copy_scaled(const double* s_dev, const double* y, double* x,
                       int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_copy_scaled, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s,
                     s_dev, y, x, n);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_axpy_dev(const double* s_dev, double c, const double* x, double* y,
                    int64_t n, hipStream_t s) {
  hipLaunchKernelGGL(k_axpy_dev, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s,
                     s_dev, c, x, y, n);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int launch_accum_cols(const double* v, double* Psi, int64_t ld, int nb, const double* c,
                      int64_t n, hipStream_t s, const double* vs_dev) {
  hipLaunchKernelGGL(k_accum_cols, dim3(vec_blocks(n)), dim3(kVecBlock), 0, s, v, Psi, ld, nb,
                     c, n, vs_dev);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

}  // namespace qsi

// ============================================================================
// C ABI: library, basis, operator, apply, connectivity, vector ops
// ============================================================================

using namespace qsi;

extern "C" {

const char* qsi_last_error(void) { return get_error(); }

const char* qsi_version(void) { return "qsi-gpu 0.1 gfx950"; }

int qsi_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

// K = ceil(log2 D) + 7, at most 31 (an 8 GiB table) and at most n_spins.
// Ice-rule bases fill buckets very unevenly, so the table is kept wide: on
// the 64-spin Q=0 basis the 29-bit key leaves a lookup-weighted mean bucket
// of ~3.4 states (the apply 5 % faster than with 27 bits); on the 96-spin
// basis the 31-bit cap still leaves ~115 (the low-key apply 3 % faster than
// with 30 bits, one bisection step fewer).
static int auto_bucket_bits(int64_t dim, int n_spins) {
  int lg = 0;
  while ((int64_t{1} << lg) < dim) ++lg;
  int k = lg + 7;
  k = std::min(k, 31);
  k = std::min(k, n_spins);
  return std::max(k, 1);
}


// Chooses the bucket key fields from the basis's informative bit positions
// (k_diff_bits) and builds the bucket table.  K = bucket_bits, or automatic;
// at most the informative positions (a smaller K then splits the basis into
// single states) and at most kKeySegs fields.  Synchronises.
static hipError_t basis_build_index(qsi_basis_s* b, int bucket_bits) {
  int K = bucket_bits ? bucket_bits : auto_bucket_bits(b->dim, b->n_spins);
  unsigned long long hmask[2] = {0, 0};
  hipError_t e = hipSuccess;
  if (b->dim >= 2) {
    unsigned long long* dmask = nullptr;
    e = hipMalloc(&dmask, 2 * sizeof(unsigned long long));
    if (e != hipSuccess) return e;
    e = hipMemset(dmask, 0, 2 * sizeof(unsigned long long));
    const int64_t want = (b->dim + 255) / 256;
    const uint32_t grid = static_cast<uint32_t>(std::min<int64_t>(want, 8192));
    if (e == hipSuccess) {
      if (b->nw == 1)
        hipLaunchKernelGGL(k_diff_bits<1>, dim3(grid), dim3(256), 0, 0, b->d_states, b->dim, dmask);
      else
        hipLaunchKernelGGL(k_diff_bits<2>, dim3(grid), dim3(256), 0, 0, b->d_states, b->dim, dmask);
      e = hipGetLastError();
    }
    if (e == hipSuccess)
      e = hipMemcpy(hmask, dmask, sizeof(hmask), hipMemcpyDeviceToHost);
    (void)hipFree(dmask);
    if (e != hipSuccess) return e;
  }
  b->infmask[0] = hmask[0];
  b->infmask[1] = hmask[1];
  // informative positions from the top; a run starts a new field, and once
  // kKeySegs fields exist the last one is extended down over the gap (the
  // fixed positions it then covers do not break the ordering)
  b->n_seg = 0;
  int taken = 0;  // key width so far
  for (int p = b->n_spins - 1; p >= 0 && taken < K; --p) {
    if (!((hmask[p >> 6] >> (p & 63)) & 1ull)) continue;
    const int q = b->n_seg - 1;
    if (q >= 0 && b->seg_lo[q] == p + 1) {
      b->seg_lo[q] = static_cast<uint8_t>(p);
      b->seg_w[q] += 1;
      taken += 1;
    } else if (b->n_seg < kKeySegs) {
      b->seg_lo[b->n_seg] = static_cast<uint8_t>(p);
      b->seg_w[b->n_seg] = 1;
      ++b->n_seg;
      taken += 1;
    } else {
      const int grow = b->seg_lo[q] - p;
      if (taken + grow > K) break;
      b->seg_lo[q] = static_cast<uint8_t>(p);
      b->seg_w[q] += grow;
      taken += grow;
    }
  }
  if (taken == 0) {  // no informative position (dim < 2): the top K bits
    K = std::min(K, b->n_spins);
    b->n_seg = 1;
    b->seg_lo[0] = static_cast<uint8_t>(b->n_spins - K);
    b->seg_w[0] = static_cast<uint8_t>(K);
    taken = K;
  }
  b->kbits = taken;
  const size_t nkeys = size_t{1} << b->kbits;
  e = hipMalloc(&b->d_table, (nkeys + 1) * sizeof(uint32_t));
  if (e != hipSuccess) return e;
  const uint32_t grid = static_cast<uint32_t>(std::min<size_t>((nkeys + 1 + 255) / 256, 1u << 20));
  const int64_t nk = static_cast<int64_t>(nkeys);
  if (b->nw == 1)
    hipLaunchKernelGGL(k_build_table<1>, dim3(grid), dim3(256), 0, 0, b->view(), nk, b->d_table);
  else
    hipLaunchKernelGGL(k_build_table<2>, dim3(grid), dim3(256), 0, 0, b->view(), nk, b->d_table);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return e;
}

int qsi_basis_create_ex(const uint64_t* h_lohi, int64_t dim, int n_spins,
                        int bucket_bits, int with_hash, qsi_basis_t* out) {
  QSI_REQUIRE(out != nullptr, "qsi_basis_create: out is null");
  *out = nullptr;
  QSI_REQUIRE(h_lohi != nullptr || dim == 0, "qsi_basis_create: states are null");
  QSI_REQUIRE(dim >= 0 && dim <= kMaxDim,
              "qsi_basis_create: dim must be in [0, 2^32 - 512]");
  QSI_REQUIRE(n_spins >= 1 && n_spins <= 128,
              "qsi_basis_create: n_spins must be in [1, 128]");
  QSI_REQUIRE(bucket_bits >= 0 && bucket_bits <= 32 && bucket_bits <= n_spins,
              "qsi_basis_create: bucket_bits must be in [0, min(32, n_spins)]");
  const int nw = n_spins <= 64 ? 1 : 2;
  // validate: in range and strictly increasing
  for (int64_t i = 0; i < dim; ++i) {
    const uint64_t lo = h_lohi[2 * i], hi = h_lohi[2 * i + 1];
    bool in_range;
    if (n_spins <= 64) in_range = hi == 0 && (n_spins == 64 || (lo >> n_spins) == 0);
    else in_range = (n_spins == 128) || (hi >> (n_spins - 64)) == 0;
    if (!in_range) {
      set_error("qsi_basis_create: state " + std::to_string(i) +
                " has bits at or above n_spins");
      return QSI_E_INVALID;
    }
    if (i > 0) {
      const uint64_t plo = h_lohi[2 * i - 2], phi = h_lohi[2 * i - 1];
      if (!(phi < hi || (phi == hi && plo < lo))) {
        set_error("qsi_basis_create: states are not strictly increasing at row " +
                  std::to_string(i));
        return QSI_E_INVALID;
      }
    }
  }
  auto* b = new (std::nothrow) qsi_basis_s();
  if (!b) {
    set_error("qsi_basis_create: host allocation failed");
    return QSI_E_NOMEM;
  }
  b->dim = dim;
  b->n_spins = n_spins;
  b->nw = nw;
  (void)hipGetDevice(&b->device);
  const size_t words = static_cast<size_t>(std::max<int64_t>(dim, 1)) * nw;
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("qsi_basis_create: ") + what + ": " + hipGetErrorString(e));
    basis_release(b);
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  };
  hipError_t e = hipMalloc(&b->d_states, words * sizeof(uint64_t));
  if (e != hipSuccess) return fail(e, "hipMalloc(states)");
  if (dim > 0) {
    if (nw == 2) {
      e = hipMemcpy(b->d_states, h_lohi, words * sizeof(uint64_t),
                    hipMemcpyHostToDevice);
    } else {
      std::vector<uint64_t> lo(static_cast<size_t>(dim));
      for (int64_t i = 0; i < dim; ++i) lo[i] = h_lohi[2 * i];
      e = hipMemcpy(b->d_states, lo.data(), words * sizeof(uint64_t),
                    hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return fail(e, "hipMemcpy(states)");
  } else {
    e = hipMemset(b->d_states, 0, words * sizeof(uint64_t));
    if (e != hipSuccess) return fail(e, "hipMemset(states)");
  }
  e = basis_build_index(b, bucket_bits);
  if (e != hipSuccess) return fail(e, "bucket table");
  if (with_hash && dim > 0) {
    int hb = 1;
    while ((int64_t{1} << hb) < 2 * dim) ++hb;  // load factor <= 1/2
    const size_t slots = size_t{1} << hb;
    const size_t bytes = slots * sizeof(ulonglong2) * nw;
    e = hipMalloc(&b->d_hash, bytes);
    if (e != hipSuccess) return fail(e, "hipMalloc(hash)");
    e = hipMemset(b->d_hash, 0xFF, bytes);
    if (e != hipSuccess) return fail(e, "hipMemset(hash)");
    b->hbits = hb;
    const uint32_t g2 = static_cast<uint32_t>((dim + 255) / 256);
    const uint64_t mask = slots - 1;
    if (nw == 1)
      hipLaunchKernelGGL(k_build_hash<1>, dim3(g2), dim3(256), 0, 0, b->d_states,
                         dim, b->d_hash, mask, hb);
    else
      hipLaunchKernelGGL(k_build_hash<2>, dim3(g2), dim3(256), 0, 0, b->d_states,
                         dim, b->d_hash, mask, hb);
    e = hipGetLastError();
    if (e != hipSuccess) return fail(e, "k_build_hash launch");
    e = hipDeviceSynchronize();
    if (e != hipSuccess) return fail(e, "k_build_hash");
  }
  *out = b;
  return QSI_OK;
}

int qsi_basis_create(const uint64_t* h_lohi, int64_t dim, int n_spins,
                     int bucket_bits, qsi_basis_t* out) {
  return qsi_basis_create_ex(h_lohi, dim, n_spins, bucket_bits, 0, out);
}

int qsi_basis_has_hash(qsi_basis_t b) { return b && b->d_hash ? 1 : 0; }

int qsi_basis_build_slots(qsi_basis_t b, int enable) {
  QSI_REQUIRE(b != nullptr, "qsi_basis_build_slots: null basis");
  QSI_REQUIRE(enable == 0 || enable == 1, "qsi_basis_build_slots: enable must be 0 or 1");
  if (!enable || b->d_slots) {
    if (!enable && b->d_slots) {
      QSI_CHECK_HIP(hipDeviceSynchronize());  // no apply may still read it
      QSI_CHECK_HIP(hipFree(b->d_slots));
      b->d_slots = nullptr;
    }
    return QSI_OK;
  }
  if (b->dim == 0) return QSI_OK;
  const size_t n = 2 * static_cast<size_t>(b->dim);
  uint32_t* slots = nullptr;
  unsigned int* flag = nullptr;
  hipError_t e = hipMalloc(&slots, n * sizeof(uint32_t));
  if (e != hipSuccess) {
    set_error(std::string("qsi_basis_build_slots: hipMalloc: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  }
  e = hipMalloc(&flag, sizeof(unsigned int));
  if (e == hipSuccess) e = hipMemset(flag, 0, sizeof(unsigned int));
  if (e == hipSuccess) e = hipMemset(slots, 0xFF, n * sizeof(uint32_t));
  if (e == hipSuccess) {
    qsi::BasisView v = b->view();
    const uint32_t grid = static_cast<uint32_t>((b->dim + 255) / 256);
    if (b->nw == 1)
      hipLaunchKernelGGL(qsi::k_build_slots<1>, dim3(grid), dim3(256), 0, 0, v, slots, flag);
    else
      hipLaunchKernelGGL(qsi::k_build_slots<2>, dim3(grid), dim3(256), 0, 0, v, slots, flag);
    e = hipGetLastError();
  }
  unsigned int too_big = 0;
  if (e == hipSuccess) e = hipMemcpy(&too_big, flag, sizeof(too_big), hipMemcpyDeviceToHost);
  if (flag) (void)hipFree(flag);
  if (e != hipSuccess || too_big) {
    (void)hipFree(slots);
    if (e != hipSuccess) {
      set_error(std::string("qsi_basis_build_slots: ") + hipGetErrorString(e));
      return QSI_E_HIP;
    }
    set_error("qsi_basis_build_slots: a bucket holds more than 65535 states "
              "(a wider bucket key is needed)");
    return QSI_E_INVALID;
  }
  b->d_slots = slots;
  return QSI_OK;
}

int qsi_basis_has_slots(qsi_basis_t b) { return b && b->d_slots ? 1 : 0; }

int qsi_basis_build_lowkeys(qsi_basis_t b, int enable) {
  QSI_REQUIRE(b != nullptr, "qsi_basis_build_lowkeys: null basis");
  QSI_REQUIRE(enable == 0 || enable == 1, "qsi_basis_build_lowkeys: enable must be 0 or 1");
  if (!enable || b->d_lowkeys) {
    if (!enable && b->d_lowkeys) {
      QSI_CHECK_HIP(hipDeviceSynchronize());  // no apply may still read it
      QSI_CHECK_HIP(hipFree(b->d_lowkeys));
      b->d_lowkeys = nullptr;
    }
    return QSI_OK;
  }
  if (b->dim == 0) return QSI_OK;
  // the informative positions below the key's lowest bit, as runs from the top
  int key_lo = b->n_spins;
  for (int q = 0; q < b->n_seg; ++q) key_lo = std::min<int>(key_lo, b->seg_lo[q]);
  int n_low = 0, bits = 0;
  uint8_t lo[qsi::kLowSegs] = {}, w[qsi::kLowSegs] = {};
  for (int p = key_lo - 1; p >= 0; --p) {
    if (!((b->infmask[p >> 6] >> (p & 63)) & 1ull)) continue;
    ++bits;
    if (n_low > 0 && lo[n_low - 1] == p + 1) {
      lo[n_low - 1] = static_cast<uint8_t>(p);
      w[n_low - 1] += 1;
    } else {
      if (n_low == qsi::kLowSegs) { n_low = qsi::kLowSegs + 1; break; }
      lo[n_low] = static_cast<uint8_t>(p);
      w[n_low] = 1;
      ++n_low;
    }
  }
  QSI_REQUIRE(bits <= qsi::kLowBits && n_low <= qsi::kLowSegs,
              "qsi_basis_build_lowkeys: more than 32 informative positions lie below the "
              "bucket key (a wider key is needed)");
  uint32_t* keys = nullptr;
  hipError_t e = hipMalloc(&keys, static_cast<size_t>(b->dim) * sizeof(uint32_t));
  if (e != hipSuccess) {
    set_error(std::string("qsi_basis_build_lowkeys: hipMalloc: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  }
  b->n_low = n_low;
  for (int q = 0; q < qsi::kLowSegs; ++q) {
    b->low_lo[q] = lo[q];
    b->low_w[q] = w[q];
  }
  const qsi::BasisView v = b->view();
  const uint32_t grid = static_cast<uint32_t>((b->dim + 255) / 256);
  if (b->nw == 1)
    hipLaunchKernelGGL(qsi::k_build_lowkeys<1>, dim3(grid), dim3(256), 0, 0, v, keys);
  else
    hipLaunchKernelGGL(qsi::k_build_lowkeys<2>, dim3(grid), dim3(256), 0, 0, v, keys);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    (void)hipFree(keys);
    set_error(std::string("qsi_basis_build_lowkeys: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  b->d_lowkeys = keys;
  return QSI_OK;
}

int qsi_basis_has_lowkeys(qsi_basis_t b) { return b && b->d_lowkeys ? 1 : 0; }

}  // extern "C"

namespace qsi {
namespace {

// flag[0] = 1 when some state has bits at or above n_spins, 2 when the
// states are not strictly increasing (any offender's code; plain stores).
template <int NW>
__global__ void k_check_sorted(const uint64_t* __restrict__ st, int64_t dim, int n_spins,
                               int* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= dim) return;
  const State<NW> s = load_state_cached<NW>(st, i);
  bool bad_range;
  if constexpr (NW == 1) bad_range = n_spins < 64 && (s.w[0] >> n_spins) != 0;
  else bad_range = n_spins < 128 && (s.w[1] >> (n_spins - 64)) != 0;
  bool bad_order = false;
  if (i > 0) bad_order = !st_lt<NW>(load_state_cached<NW>(st, i - 1), s);
  if (bad_range || bad_order) flag[0] = bad_range ? 1 : 2;
}

}  // namespace

// A basis over device states the caller hands over (dim x nw words, freed by
// the basis): validates them on the device and builds the bucket table.
int basis_adopt_device(uint64_t* d_states, int64_t dim, int n_spins, int bucket_bits,
                       qsi_basis_s** out) {
  *out = nullptr;
  if (dim > kMaxDim) {
    (void)hipFree(d_states);
    set_error("qsi_basis: more than 2^32 - 512 states");
    return QSI_E_INVALID;
  }
  auto* b = new (std::nothrow) qsi_basis_s();
  if (!b) {
    (void)hipFree(d_states);
    set_error("qsi_basis: host allocation failed");
    return QSI_E_NOMEM;
  }
  b->dim = dim;
  b->n_spins = n_spins;
  b->nw = n_spins <= 64 ? 1 : 2;
  (void)hipGetDevice(&b->device);
  b->d_states = d_states;
  auto fail = [&](hipError_t e, const char* what) {
    set_error(std::string("qsi_basis: ") + what + ": " + hipGetErrorString(e));
    basis_release(b);
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  };
  int* flag = nullptr;
  hipError_t e = hipMalloc(&flag, sizeof(int));
  if (e == hipSuccess) e = hipMemset(flag, 0, sizeof(int));
  if (e != hipSuccess) return fail(e, "validation buffer");
  if (dim > 0) {
    const uint32_t g = static_cast<uint32_t>((dim + 255) / 256);
    if (b->nw == 1)
      hipLaunchKernelGGL(k_check_sorted<1>, dim3(g), dim3(256), 0, 0, d_states, dim, n_spins, flag);
    else
      hipLaunchKernelGGL(k_check_sorted<2>, dim3(g), dim3(256), 0, 0, d_states, dim, n_spins, flag);
  }
  int hflag = 0;
  e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpy(&hflag, flag, sizeof(int), hipMemcpyDeviceToHost);
  (void)hipFree(flag);
  if (e != hipSuccess) return fail(e, "validation");
  if (hflag) {
    set_error(hflag & 1 ? "qsi_basis: a state has bits at or above n_spins"
                        : "qsi_basis: states are not strictly increasing");
    basis_release(b);
    return QSI_E_INVALID;
  }
  e = basis_build_index(b, bucket_bits);
  if (e != hipSuccess) return fail(e, "bucket table");
  *out = b;
  return QSI_OK;
}

}  // namespace qsi

extern "C" {

int qsi_basis_download(qsi_basis_t b, uint64_t* h_lohi) {
  QSI_REQUIRE(b != nullptr && (h_lohi || b->dim == 0), "qsi_basis_download: null argument");
  if (b->dim == 0) return QSI_OK;
  if (b->nw == 2) {
    QSI_CHECK_HIP(hipMemcpy(h_lohi, b->d_states, static_cast<size_t>(b->dim) * 16,
                            hipMemcpyDeviceToHost));
  } else {
    std::vector<uint64_t> lo(static_cast<size_t>(b->dim));
    QSI_CHECK_HIP(hipMemcpy(lo.data(), b->d_states, lo.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < b->dim; ++i) {
      h_lohi[2 * i] = lo[i];
      h_lohi[2 * i + 1] = 0;
    }
  }
  return QSI_OK;
}

int qsi_basis_destroy(qsi_basis_t b) {
  if (b) basis_release(b);
  return QSI_OK;
}

int64_t qsi_basis_dim(qsi_basis_t b) { return b ? b->dim : -1; }
int qsi_basis_n_spins(qsi_basis_t b) { return b ? b->n_spins : -1; }
int qsi_basis_bucket_bits(qsi_basis_t b) { return b ? b->kbits : -1; }

int qsi_basis_key_fields(qsi_basis_t b, int* n_fields, int* lo, int* width) {
  QSI_REQUIRE(b && n_fields && lo && width, "qsi_basis_key_fields: null argument");
  *n_fields = b->n_seg;
  for (int q = 0; q < kKeySegs; ++q) {
    lo[q] = q < b->n_seg ? b->seg_lo[q] : 0;
    width[q] = q < b->n_seg ? b->seg_w[q] : 0;
  }
  return QSI_OK;
}

int qsi_basis_search(qsi_basis_t b, const uint64_t* d_lohi, int64_t n,
                     int64_t* d_rank, qsi_stream_t stream) {
  QSI_REQUIRE(b != nullptr, "qsi_basis_search: null basis");
  QSI_REQUIRE(n >= 0, "qsi_basis_search: n < 0");
  if (n == 0) return QSI_OK;
  QSI_REQUIRE(d_lohi && d_rank, "qsi_basis_search: null pointer");
  const uint32_t grid = static_cast<uint32_t>((n + 255) / 256);
  auto s = static_cast<hipStream_t>(stream);
  if (b->nw == 1)
    hipLaunchKernelGGL(k_search<1>, dim3(grid), dim3(256), 0, s, b->view(),
                       d_lohi, n, d_rank);
  else
    hipLaunchKernelGGL(k_search<2>, dim3(grid), dim3(256), 0, s, b->view(),
                       d_lohi, n, d_rank);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int qsi_op_create(qsi_basis_t b, const qsi_term_t* terms, int n_terms,
                  int n_slots, qsi_op_t* out) {
  QSI_REQUIRE(out != nullptr, "qsi_op_create: out is null");
  *out = nullptr;
  QSI_REQUIRE(b != nullptr, "qsi_op_create: null basis");
  QSI_REQUIRE(n_terms >= 0 && (terms != nullptr || n_terms == 0),
              "qsi_op_create: bad terms");
  QSI_REQUIRE(n_slots >= 0, "qsi_op_create: n_slots < 0");
  const int nw = b->nw;
  auto* op = new (std::nothrow) qsi_op_s();
  if (!op) {
    set_error("qsi_op_create: host allocation failed");
    return QSI_E_NOMEM;
  }
  // merged single-site Z terms: coefficient -> mask (kept in first-seen order)
  std::vector<DiagTerm> fields;
  for (int t = 0; t < n_terms; ++t) {
    const qsi_term_t& T = terms[t];
    const uint64_t bad_hi = (nw == 1) ? (T.X[1] | T.Z[1] | T.up[1] | T.down[1]) : 0;
    if (bad_hi) {
      delete op;
      set_error("qsi_op_create: term " + std::to_string(t) +
                " acts on spins >= 64 but the basis has <= 64 spins");
      return QSI_E_INVALID;
    }
    if (T.slot >= 0 && T.slot >= n_slots) {
      delete op;
      set_error("qsi_op_create: term slot out of range");
      return QSI_E_INVALID;
    }
    // a term whose sign is 0 or whose up/down masks overlap is identically 0
    if (T.sign == 0 || (T.up[0] & T.down[0]) || (T.up[1] & T.down[1])) continue;
    if ((T.X[0] | T.X[1]) == 0) {
      // diagonal: a Z string, possibly times up/down projectors (a ring
      // that visits a spin twice reduces to one)
      const bool proj = (T.up[0] | T.up[1] | T.down[0] | T.down[1]) != 0;
      const int nz = __builtin_popcountll(T.Z[0]) + __builtin_popcountll(T.Z[1]);
      if (nz == 1 && !proj) {
        // single-site Z: c * sign * Z_i.  Unslotted terms merge on equal
        // c * sign; slotted ones on equal (slot, c, sign), so that slot
        // coefficients (which replace c) apply to the whole group.
        const bool slotted = T.slot >= 0;
        const double ce = slotted ? T.coeff : T.coeff * T.sign;
        const double sg = slotted ? static_cast<double>(T.sign) : 1.0;
        bool merged = false;
        for (auto& f : fields) {
          if (f.slot == T.slot && f.c == ce && f.sgn == sg) {
            f.Z[0] |= T.Z[0];
            f.Z[1] |= T.Z[1];
            merged = true;
            break;
          }
        }
        if (!merged) {
          DiagTerm d{};
          d.Z[0] = T.Z[0];
          d.Z[1] = T.Z[1];
          d.c = ce;
          d.sgn = sg;
          d.kind = 1;
          d.slot = T.slot;
          fields.push_back(d);
        }
        continue;
      }
      DiagTerm d{};
      for (int w = 0; w < 2; ++w) {
        d.m[w] = T.up[w] | T.down[w];
        d.v[w] = T.up[w];
      }
      d.Z[0] = T.Z[0];
      d.Z[1] = T.Z[1];
      d.c = T.coeff;
      d.sgn = static_cast<double>(T.sign);
      d.kind = 0;
      d.slot = T.slot;
      op->h_diag.push_back(d);
    } else {
      OffTerm o{};
      for (int w = 0; w < 2; ++w) {
        o.m[w] = T.up[w] | T.down[w];
        o.v[w] = T.up[w] ^ (T.X[w] & o.m[w]);
        o.X[w] = T.X[w];
        o.Z[w] = T.Z[w];
      }
      o.c = T.coeff;
      o.sgn = static_cast<double>(T.sign);
      o.slot = T.slot;
      if (o.Z[0] | o.Z[1]) op->has_z = true;
      op->h_off.push_back(o);
    }
  }
  for (auto& f : fields) op->h_diag.push_back(f);
  // coefficient classes: (slot, c) pairs of the off-diagonal terms
  bool cls_ok = true;
  for (auto& o : op->h_off) {
    int k = 0;
    while (k < static_cast<int>(op->h_cls.size()) &&
           !(op->h_cls[k].slot == o.slot && op->h_cls[k].c == o.c))
      ++k;
    if (k == static_cast<int>(op->h_cls.size())) {
      if (k == kMaxCoefClasses) {
        cls_ok = false;
        break;
      }
      op->h_cls.push_back(CoefClass{o.c, o.slot, 0});
    }
    o.cls = k;
  }
  if (!cls_ok) {
    op->h_cls.clear();
    for (auto& o : op->h_off) o.cls = -1;
  }
  op->n_cls = static_cast<int>(op->h_cls.size());
  op->n_off = static_cast<int>(op->h_off.size());
  op->n_diag = static_cast<int>(op->h_diag.size());
  op->n_slots = n_slots;
  hipError_t e = hipMalloc(&op->d_off, std::max<size_t>(1, op->h_off.size()) * sizeof(OffTerm));
  if (e == hipSuccess)
    e = hipMalloc(&op->d_diag, std::max<size_t>(1, op->h_diag.size()) * sizeof(DiagTerm));
  if (e == hipSuccess && op->n_off)
    e = hipMemcpy(op->d_off, op->h_off.data(), op->h_off.size() * sizeof(OffTerm),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess && op->n_diag)
    e = hipMemcpy(op->d_diag, op->h_diag.data(), op->h_diag.size() * sizeof(DiagTerm),
                  hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMalloc(&op->d_cls, kMaxCoefClasses * sizeof(CoefClass));
  if (e == hipSuccess && op->n_cls)
    e = hipMemcpy(op->d_cls, op->h_cls.data(), op->h_cls.size() * sizeof(CoefClass),
                  hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (op->d_off) (void)hipFree(op->d_off);
    if (op->d_diag) (void)hipFree(op->d_diag);
    if (op->d_cls) (void)hipFree(op->d_cls);
    delete op;
    set_error(std::string("qsi_op_create: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  }
  basis_retain(b);
  op->basis = b;
  *out = op;
  return QSI_OK;
}

int qsi_op_destroy(qsi_op_t op) {
  if (!op) return QSI_OK;
  if (op->d_off) (void)hipFree(op->d_off);
  if (op->d_diag) (void)hipFree(op->d_diag);
  if (op->d_cls) (void)hipFree(op->d_cls);
  if (op->d_rowptr) (void)hipFree(op->d_rowptr);
  if (op->d_cj) (void)hipFree(op->d_cj);
  if (op->d_cm) (void)hipFree(op->d_cm);
  if (op->basis) basis_release(op->basis);
  delete op;
  return QSI_OK;
}

int64_t qsi_op_dim(qsi_op_t op) { return op ? op->basis->dim : -1; }

int qsi_op_set_variant(qsi_op_t op, int variant) {
  QSI_REQUIRE(op != nullptr, "qsi_op_set_variant: null operator");
  QSI_REQUIRE(variant >= QSI_APPLY_AUTO && variant <= QSI_APPLY_LOW8,
              "qsi_op_set_variant: unknown variant");
  QSI_REQUIRE(op->basis->d_slots || variant < QSI_APPLY_SLOT2 || variant > QSI_APPLY_SLOT3,
              "qsi_op_set_variant: the basis has no slot index (qsi_basis_build_slots)");
  QSI_REQUIRE(op->basis->d_lowkeys || variant < QSI_APPLY_LOW2,
              "qsi_op_set_variant: the basis has no low keys (qsi_basis_build_lowkeys)");
  QSI_REQUIRE(variant != QSI_APPLY_CACHED || op->conn_nnz >= 0,
              "qsi_op_set_variant: no cached connectivity (qsi_op_cache_connectivity)");
  QSI_REQUIRE(op->basis->d_hash || (variant != QSI_APPLY_HASH && variant != QSI_APPLY_HASH4),
              "qsi_op_set_variant: the basis has no hash index");
  op->variant = variant;
  return QSI_OK;
}

int qsi_op_verify_lowkeys(qsi_op_t op, int* mode) {
  QSI_REQUIRE(op != nullptr, "qsi_op_verify_lowkeys: null operator");
  QSI_REQUIRE(op->basis->d_lowkeys != nullptr,
              "qsi_op_verify_lowkeys: the basis has no low keys (qsi_basis_build_lowkeys)");
  op->lowkeys_check = qsi::kLowCmpState;
  unsigned int* bad = nullptr;
  unsigned int h_bad = 0;
  QSI_CHECK_HIP(hipMalloc(&bad, sizeof(unsigned int)));
  hipError_t e = hipMemset(bad, 0, sizeof(unsigned int));
  if (e == hipSuccess && op->basis->dim > 0 && op->n_off > 0) {
    const qsi::BasisView v = op->basis->view();
    const uint32_t grid = static_cast<uint32_t>((op->basis->dim + 255) / 256);
    if (op->basis->nw == 1)
      hipLaunchKernelGGL(qsi::k_verify_lowkeys<1>, dim3(grid), dim3(256), 0, 0, v, op->d_off,
                         op->n_off, bad);
    else
      hipLaunchKernelGGL(qsi::k_verify_lowkeys<2>, dim3(grid), dim3(256), 0, 0, v, op->d_off,
                         op->n_off, bad);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpy(&h_bad, bad, sizeof(h_bad), hipMemcpyDeviceToHost);
  (void)hipFree(bad);
  if (e != hipSuccess) {
    set_error(std::string("qsi_op_verify_lowkeys: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  op->lowkeys_check = (h_bad & 1u) ? qsi::kLowCmpState
                    : (h_bad & 2u) ? qsi::kLowCmpKey
                                   : qsi::kLowTrusted;
  if (mode) *mode = op->lowkeys_check;
  return QSI_OK;
}

int qsi_op_get_variant(qsi_op_t op) { return op ? op->variant : -1; }

int qsi_op_set_cached_pipe(qsi_op_t op, int pipe) {
  QSI_REQUIRE(op != nullptr, "qsi_op_set_cached_pipe: null operator");
  QSI_REQUIRE(pipe == 4 || pipe == 8 || pipe == 16 || pipe == 32,
              "qsi_op_set_cached_pipe: pipe must be 4, 8, 16 or 32");
  op->cached_pipe = pipe;
  return QSI_OK;
}

int qsi_op_get_cached_pipe(qsi_op_t op) { return op ? op->cached_pipe : -1; }

int qsi_op_set_block_cfg(qsi_op_t op, int cfg) {
  QSI_REQUIRE(op != nullptr, "qsi_op_set_block_cfg: null operator");
  QSI_REQUIRE(cfg >= 0 && cfg <= 15 && blk_cfg_kernel(cfg) <= 2,
              "qsi_op_set_block_cfg: cfg must be kernel 0-2, plus 4 for a half grid, "
              "plus 8 for column pairs");
  op->blk_cfg = cfg;
  return QSI_OK;
}

int qsi_op_get_block_cfg(qsi_op_t op) { return op ? op->blk_cfg : -1; }
int qsi_op_n_offdiag(qsi_op_t op) { return op ? op->n_off : -1; }
int qsi_op_n_diag(qsi_op_t op) { return op ? op->n_diag : -1; }

int qsi_op_apply(qsi_op_t op, const double* d_x, double* d_y, int accumulate,
                 qsi_stream_t stream) {
  QSI_REQUIRE(op != nullptr, "qsi_op_apply: null operator");
  if (op->basis->dim == 0) return QSI_OK;
  QSI_REQUIRE(d_x && d_y, "qsi_op_apply: null vector");
  QSI_REQUIRE(d_x != d_y, "qsi_op_apply: x and y must not alias");
  return launch_apply(op, d_x, d_y, accumulate ? d_y : nullptr, 1.0, nullptr,
                      nullptr, static_cast<hipStream_t>(stream));
}

int qsi_op_apply_host(qsi_op_t op, const double* h_x, double* h_y,
                      int accumulate) {
  QSI_REQUIRE(op != nullptr, "qsi_op_apply_host: null operator");
  const int64_t D = op->basis->dim;
  if (D == 0) return QSI_OK;
  QSI_REQUIRE(h_x && h_y, "qsi_op_apply_host: null vector");
  double *dx = nullptr, *dy = nullptr;
  const size_t bytes = static_cast<size_t>(D) * sizeof(double);
  QSI_CHECK_HIP(hipMalloc(&dx, bytes));
  hipError_t e = hipMalloc(&dy, bytes);
  if (e == hipSuccess) e = hipMemcpy(dx, h_x, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess && accumulate) e = hipMemcpy(dy, h_y, bytes, hipMemcpyHostToDevice);
  int rc = QSI_OK;
  if (e == hipSuccess) {
    rc = launch_apply(op, dx, dy, accumulate ? dy : nullptr, 1.0, nullptr,
                      nullptr, nullptr);
    if (rc == QSI_OK) e = hipMemcpy(h_y, dy, bytes, hipMemcpyDeviceToHost);
  }
  (void)hipFree(dx);
  if (dy) (void)hipFree(dy);
  if (e != hipSuccess) {
    set_error(std::string("qsi_op_apply_host: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  return rc;
}

int qsi_op_apply_block(qsi_op_t op, const double* d_X, int64_t ldx, double* d_Y,
                       int64_t ldy, int R, const double* d_coeffs,
                       int accumulate, qsi_stream_t stream) {
  QSI_REQUIRE(op != nullptr, "qsi_op_apply_block: null operator");
  QSI_REQUIRE(R >= 1 && R <= 64, "qsi_op_apply_block: R must be in [1, 64]");
  QSI_REQUIRE(ldx >= R && ldy >= R, "qsi_op_apply_block: ld < R");
  QSI_REQUIRE(d_coeffs == nullptr || op->n_slots > 0,
              "qsi_op_apply_block: coefficients given but the operator has no slots");
  if (op->basis->dim == 0) return QSI_OK;
  QSI_REQUIRE(d_X && d_Y, "qsi_op_apply_block: null matrix");
  QSI_REQUIRE(d_X != d_Y, "qsi_op_apply_block: X and Y must not alias");
  BlkArgs a{};
  a.n_sec = 1;
  a.R = R;
  int G = 64;
  while (G / 2 >= R && G > 1) G >>= 1;
  a.G = G;
  a.ldx = ldx;
  a.ldy = ldy;
  a.accumulate = accumulate ? 1 : 0;
  a.y_colmajor = 0;
  a.sec[0].b = op->basis->view();
  a.sec[0].off = op->d_off;
  a.sec[0].dg = op->d_diag;
  a.sec[0].cls = op->d_cls;
  a.sec[0].n_cls = op->n_cls;
  a.sec[0].coeffs = d_coeffs;
  a.sec[0].n_off = op->n_off;
  a.sec[0].n_diag = op->n_diag;
  a.sec[0].rowptr = op->d_rowptr;
  a.sec[0].cj = op->d_cj;
  a.sec[0].cm = op->d_cm;
  a.n_tiles = (op->basis->dim + kTileRows - 1) / kTileRows;
  a.cfg = op->blk_cfg;
  const int nb = blk_apply_blocks(op->basis->nw, a.n_tiles, blk_cached(a), a.cfg);
  return launch_blk_apply(a, op->basis->nw, nb, d_X, d_Y, nullptr,
                          static_cast<hipStream_t>(stream));
}

int qsi_op_connectivity(qsi_op_t op, int64_t* nnz, int64_t* d_rows,
                        int64_t* d_cols, double* d_vals, double* d_diag) {
  QSI_REQUIRE(op != nullptr && nnz != nullptr, "qsi_op_connectivity: null argument");
  const int64_t D = op->basis->dim;
  if (D == 0) {
    *nnz = 0;
    return QSI_OK;
  }
  int64_t *counts = nullptr, *offsets = nullptr;
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  const uint32_t grid = static_cast<uint32_t>((D + 255) / 256);
  auto cleanup = [&]() {
    if (counts) (void)hipFree(counts);
    if (offsets) (void)hipFree(offsets);
    if (tmp) (void)hipFree(tmp);
  };
  hipError_t e = hipMalloc(&counts, static_cast<size_t>(D + 1) * sizeof(int64_t));
  if (e == hipSuccess) e = hipMalloc(&offsets, static_cast<size_t>(D + 1) * sizeof(int64_t));
  if (e == hipSuccess) e = hipMemset(counts, 0, static_cast<size_t>(D + 1) * sizeof(int64_t));
  if (e != hipSuccess) {
    cleanup();
    set_error(std::string("qsi_op_connectivity: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  auto view = op->basis->view();
  if (op->basis->nw == 1)
    hipLaunchKernelGGL(k_conn<1>, dim3(grid), dim3(256), 0, 0, view, op->d_off,
                       op->n_off, op->d_diag, op->n_diag, nullptr, counts,
                       nullptr, nullptr, nullptr, d_rows ? d_diag : nullptr);
  else
    hipLaunchKernelGGL(k_conn<2>, dim3(grid), dim3(256), 0, 0, view, op->d_off,
                       op->n_off, op->d_diag, op->n_diag, nullptr, counts,
                       nullptr, nullptr, nullptr, d_rows ? d_diag : nullptr);
  e = hipGetLastError();
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, offsets, D + 1);
  if (e == hipSuccess) e = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1));
  if (e == hipSuccess)
    e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, offsets, D + 1);
  int64_t total = 0;
  if (e == hipSuccess)
    e = hipMemcpy(&total, offsets + D, sizeof(int64_t), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    cleanup();
    set_error(std::string("qsi_op_connectivity: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  if (d_rows == nullptr) {
    *nnz = total;
    cleanup();
    return QSI_OK;
  }
  if (*nnz < total || !d_cols || !d_vals || !d_diag) {
    cleanup();
    set_error("qsi_op_connectivity: output arrays too small or null");
    return QSI_E_INVALID;
  }
  if (op->basis->nw == 1)
    hipLaunchKernelGGL(k_conn<1>, dim3(grid), dim3(256), 0, 0, view, op->d_off,
                       op->n_off, op->d_diag, op->n_diag, offsets, nullptr,
                       d_rows, d_cols, d_vals, d_diag);
  else
    hipLaunchKernelGGL(k_conn<2>, dim3(grid), dim3(256), 0, 0, view, op->d_off,
                       op->n_off, op->d_diag, op->n_diag, offsets, nullptr,
                       d_rows, d_cols, d_vals, d_diag);
  e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  cleanup();
  if (e != hipSuccess) {
    set_error(std::string("qsi_op_connectivity: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  *nnz = total;
  return QSI_OK;
}

int qsi_vec_dot(const double* d_a, const double* d_b, int64_t n, double* h_out,
                qsi_stream_t stream) {
  QSI_REQUIRE(h_out != nullptr, "qsi_vec_dot: null output");
  if (n <= 0) {
    *h_out = 0.0;
    return QSI_OK;
  }
  QSI_REQUIRE(d_a && d_b, "qsi_vec_dot: null vector");
  auto s = static_cast<hipStream_t>(stream);
  double* buf = nullptr;
  QSI_CHECK_HIP(hipMalloc(&buf, (max_vec_partials() + 2) * sizeof(double)));
  int rc = launch_dot(d_a, d_b, n, buf + 2, buf, s);
  hipError_t e = hipSuccess;
  if (rc == QSI_OK) e = hipMemcpyAsync(h_out, buf, sizeof(double), hipMemcpyDeviceToHost, s);
  if (rc == QSI_OK && e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(buf);
  if (e != hipSuccess) {
    set_error(std::string("qsi_vec_dot: ") + hipGetErrorString(e));
    return QSI_E_HIP;
  }
  return rc;
}

int qsi_vec_axpy(double alpha, const double* d_x, double* d_y, int64_t n,
                 qsi_stream_t stream) {
  if (n <= 0) return QSI_OK;
  QSI_REQUIRE(d_x && d_y, "qsi_vec_axpy: null vector");
  return launch_axpy(alpha, d_x, d_y, n, static_cast<hipStream_t>(stream));
}

int qsi_vec_scal(double alpha, double* d_x, int64_t n, qsi_stream_t stream) {
  if (n <= 0) return QSI_OK;
  QSI_REQUIRE(d_x, "qsi_vec_scal: null vector");
  return launch_scal(alpha, d_x, n, static_cast<hipStream_t>(stream));
}

int qsi_vec_randn(double* d_x, int64_t n, uint64_t seed, qsi_stream_t stream) {
  if (n <= 0) return QSI_OK;
  QSI_REQUIRE(d_x, "qsi_vec_randn: null vector");
  return launch_randn(d_x, n, seed, static_cast<hipStream_t>(stream));
}

}  // extern "C"
