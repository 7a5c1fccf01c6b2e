// Shared device/host helpers for the qsi HIP library (gfx950).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/qsi_gpu.h"

namespace qsi {

// ---- errors ---------------------------------------------------------------
void set_error(const std::string& msg);
const char* get_error();

#define QSI_CHECK_HIP(expr)                                                   \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::qsi::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));   \
      return QSI_E_HIP;                                                       \
    }                                                                         \
  } while (0)

#define QSI_REQUIRE(cond, msg)                                                \
  do {                                                                        \
    if (!(cond)) {                                                            \
      ::qsi::set_error(msg);                                                  \
      return QSI_E_INVALID;                                                   \
    }                                                                         \
  } while (0)

// Checks the launch that was just enqueued.
#define QSI_CHECK_LAUNCH() QSI_CHECK_HIP(hipGetLastError())

// ---- launch geometry --------------------------------------------------------
// MI355X: 256 CUs in 8 XCDs.  Workgroups are handed to XCDs round-robin by
// workgroup id, so workgroup g runs on XCD g % 8.  xcd_remap() turns the
// hardware id into a logical id such that each XCD works on one contiguous
// range of logical ids (rows that are close in the sorted basis share flip
// targets and x entries, so they should share an L2).
constexpr int kNumXcd = 8;

__device__ __forceinline__ uint32_t xcd_remap(uint32_t g, uint32_t n) {
  const uint32_t per = n / kNumXcd;      // full rounds
  const uint32_t rem = n % kNumXcd;
  const uint32_t xcd = g % kNumXcd;
  const uint32_t k = g / kNumXcd;
  // XCD x owns logical ids [start(x), start(x) + cnt(x)), cnt = per + (x < rem)
  const uint32_t start = xcd * per + (xcd < rem ? xcd : rem);
  const uint32_t cnt = per + (xcd < rem ? 1u : 0u);
  return k < cnt ? start + k : g;  // k >= cnt cannot happen when n % 8 handled
}

// ---- basis states -----------------------------------------------------------
// A state of NW 64-bit words, word 0 = spins 0..63.  The device basis stores
// NW words per state, interleaved ([lo] or [lo, hi]).
template <int NW>
struct State {
  uint64_t w[NW];
};

// Plain (cached) load; used for lookups where reuse across lanes matters.
template <int NW>
__device__ __forceinline__ State<NW> load_state_cached(
    const uint64_t* __restrict__ s, int64_t i) {
  State<NW> r;
  if constexpr (NW == 1) {
    r.w[0] = s[i];
  } else {
    const ulonglong2 v = reinterpret_cast<const ulonglong2*>(s)[i];
    r.w[0] = v.x;
    r.w[1] = v.y;
  }
  return r;
}

template <int NW>
__device__ __forceinline__ bool st_le(const State<NW>& a, const State<NW>& b) {
  if constexpr (NW == 1) {
    return a.w[0] <= b.w[0];
  } else {
    return (a.w[1] < b.w[1]) || (a.w[1] == b.w[1] && a.w[0] <= b.w[0]);
  }
}

template <int NW>
__device__ __forceinline__ bool st_eq(const State<NW>& a, const State<NW>& b) {
  if constexpr (NW == 1) {
    return a.w[0] == b.w[0];
  } else {
    return a.w[0] == b.w[0] && a.w[1] == b.w[1];
  }
}

// Bucket keys.  The key of a state is the concatenation, highest field
// first, of up to kKeySegs bit fields [lo, lo + w) of the state.  The fields
// cover the K highest bit positions at which two neighbouring states of the
// sorted basis first differ ("informative" positions): a position where no
// neighbouring pair first differs is fixed by the bits above it in every
// basis state, so leaving it out keeps the key non-decreasing along the basis
// while the same K key bits split the basis into smaller buckets than the top
// K bits do (ice-rule bases fix many positions).  With no informative
// positions the key is the top K bits.
constexpr int kKeySegs = 4;

// Low keys.  The informative positions below the bucket key's lowest bit,
// packed highest first into one 32-bit word per state (at most kLowBits of
// them, in at most kLowSegs runs of adjacent positions).  Two basis states
// always differ at some informative position, so within one bucket (equal
// key bits) the low keys are distinct and ordered like the states: a lookup
// can bisect a bucket on 4-byte low keys instead of whole states, and only
// the candidate's state needs loading (a target outside the basis may share
// its informative bits with a basis state and differ at a fixed position).
constexpr int kLowSegs = 32;
constexpr int kLowBits = 32;

// Device view of a basis: states + bucket table (+ slot index, low keys).
struct BasisView {
  const uint64_t* states;  // D * NW words
  const uint32_t* table;   // 2^K + 1 entries: table[k] = #states with key < k
  const uint32_t* slots;   // slot index (2 D entries) or null
  const uint32_t* lowkeys; // D low keys or null
  int64_t dim;
  int kbits;               // K = sum of the field widths
  int n_spins;
  int n_seg;
  uint8_t seg_lo[kKeySegs];
  uint8_t seg_w[kKeySegs];
  int n_low;               // low-key runs, highest first
  uint8_t low_lo[kLowSegs];
  uint8_t low_w[kLowSegs];
};

template <int NW>
__host__ __device__ __forceinline__ uint32_t state_field(const State<NW>& s, int lo, int w) {
  uint64_t v;
  if constexpr (NW == 1) {
    v = s.w[0] >> lo;
  } else {
    if (lo >= 64) v = s.w[1] >> (lo - 64);
    else if (lo == 0) v = s.w[0];
    else v = (s.w[1] << (64 - lo)) | (s.w[0] >> lo);
  }
  return static_cast<uint32_t>(v & ((uint64_t{1} << w) - 1));
}

template <int NW>
__host__ __device__ __forceinline__ uint32_t bucket_key(const BasisView& b, const State<NW>& s) {
  uint64_t k = 0;  // 64-bit, so a 32-bit field may shift in
#pragma unroll
  for (int q = 0; q < kKeySegs; ++q) {
    if (q < b.n_seg) k = (k << b.seg_w[q]) | state_field<NW>(s, b.seg_lo[q], b.seg_w[q]);
  }
  return static_cast<uint32_t>(k);
}

template <int NW>
__host__ __device__ __forceinline__ uint32_t low_key(const BasisView& b, const State<NW>& s) {
  uint64_t k = 0;  // 64-bit, so a 32-bit run may shift in
  for (int q = 0; q < b.n_low; ++q)
    k = (k << b.low_w[q]) | state_field<NW>(s, b.low_lo[q], b.low_w[q]);
  return static_cast<uint32_t>(k);
}

// true if the state has a bit at or above n_spins (it is then larger than
// every basis state)
template <int NW>
__device__ __forceinline__ bool above_range(const State<NW>& s, int n_spins) {
  if constexpr (NW == 1) {
    return n_spins < 64 && (s.w[0] >> n_spins) != 0;
  } else {
    return n_spins < 128 && (s.w[1] >> (n_spins - 64)) != 0;
  }
}

template <int NW>
__device__ __forceinline__ bool st_lt(const State<NW>& a, const State<NW>& b) {
  if constexpr (NW == 1) {
    return a.w[0] < b.w[0];
  } else {
    return (a.w[1] < b.w[1]) || (a.w[1] == b.w[1] && a.w[0] < b.w[0]);
  }
}

// Bucket k's ranks [lo, lo + n) from table[k], table[k + 1] with one 8-byte
// load (the pair is 4-byte aligned, which global loads allow).
__device__ __forceinline__ void bucket_bounds(const uint32_t* __restrict__ table, uint32_t k,
                                              uint32_t& lo, uint32_t& n) {
  uint2 p;
  __builtin_memcpy(&p, table + k, sizeof(p));
  lo = p.x;
  n = p.y - p.x;
}

// Rank of `key` in the basis, or -1.  The bucket [table[k], table[k+1]) is
// searched with a branch-free lower-bound bisection.  A basis state's key
// selects the bucket that holds it; any other state is reported absent.
template <int NW>
__device__ __forceinline__ int64_t basis_lookup(const BasisView& b,
                                                const State<NW>& key) {
  const uint32_t k = bucket_key<NW>(b, key);
  uint32_t lo = b.table[k];
  uint32_t n = b.table[static_cast<uint64_t>(k) + 1] - lo;
  if (n == 0) return -1;
  while (n > 1) {
    const uint32_t half = n >> 1;
    const State<NW> m = load_state_cached<NW>(b.states, lo + half);
    lo = st_le<NW>(m, key) ? lo + half : lo;
    n -= half;
  }
  const State<NW> f = load_state_cached<NW>(b.states, lo);
  return st_eq<NW>(f, key) ? static_cast<int64_t>(lo) : -1;
}

// ---- slot index ------------------------------------------------------------
// A second level under the bucket table for bases whose buckets stay large
// (the 96-spin ice basis spreads its states over 49 informative positions, so
// a 31-bit key leaves ~115 states in the bucket of an average lookup).  The
// bucket of ranks [lo, lo + n) owns the 2n slots [2 lo, 2 lo + 2n) of one
// uint32 array (so no offsets are stored: the slots of all buckets tile
// [0, 2 D)); each of its states sits in one slot, placed by linear probing
// from a hash of the state:
//   slot = (fingerprint << 16) | (rank - lo),  fingerprint = hash bits 48..63
// and kSlotEmpty marks a free slot (n <= kSlotMaxBucket keeps rank - lo below
// 0xFFFF, so no entry equals it).  A lookup reads the bucket bounds, then
// slots from the state's home until it meets an empty one (absent) or a
// fingerprint match, whose state is compared (and x gathered) at once; load
// factor 1/2 keeps the expected probe count near 1.5, and the slots of one
// bucket are contiguous, so neighbouring rows' lookups share their lines.
constexpr uint32_t kSlotEmpty = 0xFFFFFFFFu;
constexpr uint32_t kSlotMaxBucket = 0xFFFFu;

template <int NW>
__host__ __device__ __forceinline__ uint64_t slot_hash(const State<NW>& s) {
  uint64_t h = s.w[0];
  if constexpr (NW == 2) h ^= s.w[1] * 0x9E3779B97F4A7C15ull;
  // splitmix64 finaliser (Steele, Lea & Flood 2014)
  h ^= h >> 30;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 27;
  h *= 0x94D049BB133111EBull;
  return h ^ (h >> 31);
}

// Home slot (0 <= home < m) of a hash in a bucket region of m slots.
__host__ __device__ __forceinline__ uint32_t slot_home(uint64_t h, uint32_t m) {
  return static_cast<uint32_t>((static_cast<uint64_t>(static_cast<uint32_t>(h)) * m) >> 32);
}

__host__ __device__ __forceinline__ uint32_t slot_fp(uint64_t h) {
  return static_cast<uint32_t>(h >> 48);
}

// Rank of key or -1 through the slot index (b.slots non-null).
template <int NW>
__device__ __forceinline__ int64_t slot_lookup(const BasisView& b, const State<NW>& key) {
  uint32_t lo, n;
  bucket_bounds(b.table, bucket_key<NW>(b, key), lo, n);
  if (n == 0) return -1;
  const uint64_t h = slot_hash<NW>(key);
  const uint32_t m = 2 * n, fp = slot_fp(h);
  const uint32_t* reg = b.slots + 2 * static_cast<uint64_t>(lo);
  uint32_t p = slot_home(h, m);
  for (;;) {
    const uint32_t e = reg[p];
    if (e == kSlotEmpty) return -1;
    if ((e >> 16) == fp) {
      const uint32_t J = lo + (e & 0xFFFFu);
      if (st_eq<NW>(load_state_cached<NW>(b.states, J), key)) return J;
    }
    p = p + 1 == m ? 0 : p + 1;
  }
}

// ---- hash index ------------------------------------------------------------
// Open addressing with linear probing over 2^hbits slots, load <= 1/2.  A slot
// is one 16-byte word pair: NW=1 {state, rank}; NW=2 uses two pairs
// {lo, hi}, {rank, -}.  rank == kHashEmpty marks a free slot.
constexpr uint64_t kHashEmpty = ~0ull;

struct HashView {
  const ulonglong2* slots;
  uint64_t mask;  // 2^hbits - 1
  int hbits;
};

template <int NW>
__host__ __device__ __forceinline__ uint64_t hash_home(const State<NW>& s, int hbits) {
  uint64_t k = s.w[0];
  if constexpr (NW == 2) k ^= s.w[1] * 0xC2B2AE3D27D4EB4Full;
  return (k * 0x9E3779B97F4A7C15ull) >> (64 - hbits);
}

// Rank of key or -1.
template <int NW>
__device__ __forceinline__ int64_t hash_lookup(const HashView& h,
                                               const State<NW>& key) {
  uint64_t p = hash_home<NW>(key, h.hbits);
  for (;;) {
    if constexpr (NW == 1) {
      const ulonglong2 e = h.slots[p];
      if (e.y == kHashEmpty) return -1;
      if (e.x == key.w[0]) return static_cast<int64_t>(e.y);
    } else {
      const ulonglong2 e = h.slots[2 * p];
      const ulonglong2 r = h.slots[2 * p + 1];
      if (r.x == kHashEmpty) return -1;
      if (e.x == key.w[0] && e.y == key.w[1]) return static_cast<int64_t>(r.x);
    }
    p = (p + 1) & h.mask;
  }
}

// First rank whose state is >= key (dim if none): bucket from the table,
// then a lower-bound bisection inside it.  A state outside the basis may
// differ from the basis at a position the key leaves out, so its bucket need
// not bound its place: the result is checked against its neighbours and, if
// it is off, the bisection is redone over the whole basis.
template <int NW>
__device__ __forceinline__ int64_t basis_lower_bound(const BasisView& b,
                                                     const State<NW>& key) {
  if (above_range<NW>(key, b.n_spins)) return b.dim;
  auto bisect = [&](int64_t lo, int64_t n) {
    while (n > 0) {
      const int64_t half = n >> 1;
      const State<NW> m = load_state_cached<NW>(b.states, lo + half);
      if (st_lt<NW>(m, key)) {
        lo += half + 1;
        n -= half + 1;
      } else {
        n = half;
      }
    }
    return lo;
  };
  const uint32_t k = bucket_key<NW>(b, key);
  const int64_t lo0 = b.table[k];
  const int64_t lo = bisect(lo0, static_cast<int64_t>(b.table[static_cast<uint64_t>(k) + 1]) - lo0);
  const bool below_ok = lo == 0 || st_lt<NW>(load_state_cached<NW>(b.states, lo - 1), key);
  const bool above_ok = lo == b.dim || !st_lt<NW>(load_state_cached<NW>(b.states, lo), key);
  if (below_ok && above_ok) return lo;
  return bisect(0, b.dim);
}

}  // namespace qsi
