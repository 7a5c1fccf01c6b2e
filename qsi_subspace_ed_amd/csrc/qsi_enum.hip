// Ice-rule basis enumeration on the device, for bases too large to build on
// the host and copy over (96 spins: ~2e9 states, 35 GB).
//
// The search is the reference's gen_spinon_basis depth-first tree
// (components/build_basis/src/pyro_tree.cpp:20-240; gen_spinon_basis.cpp:
// 16-156): spins in a fixed visiting order, each spin up or down as its two
// tetrahedra allow (2 up spins per complete tetrahedron), up to
// n_spinon_pairs forced violations.  The host expands the tree breadth-first
// to ~kFrontier nodes; then
//   k_enum<false>  one thread per frontier node walks its subtree depth-first
//                  and counts the leaves
//   scan           exclusive sum of the counts -> each subtree's output offset
//   k_enum<true>   the same walk writes every leaf, its bits moved from
//                  visiting position k to spin order[k]
//   sort           two stable 64-bit radix passes (low word, then high word)
//                  give the ascending 128-bit order of ZBasisBase::states
// and the sorted states become a device basis (no host copy).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <vector>

#include "qsi_enum.hpp"
#include "qsi_internal.hpp"

namespace qsi {
namespace {

constexpr size_t kFrontier = size_t{1} << 20;

struct EnumDev {
  const ulonglong2* tmask;  // tetrahedron masks in visiting order
  const int* tmin;
  const int* tmax;
  const int* tsize;
  const int* nb0;  // first two tetrahedra of each visited spin
  const int* nb1;
  const int* order;  // visiting position k -> spin index
  int n;
  int pairs;
};

struct ENode {
  uint64_t lo, hi;
  int spin, pairs;
};

// Bits [0, k) set.
__device__ __forceinline__ void low_mask(int k, uint64_t& lo, uint64_t& hi) {
  if (k >= 128) {
    lo = hi = ~0ull;
  } else if (k >= 64) {
    lo = ~0ull;
    hi = k == 64 ? 0 : (~0ull >> (128 - k));
  } else {
    lo = k == 0 ? 0 : (~0ull >> (64 - k));
    hi = 0;
  }
}

// possible_spin_states (pyro_tree.cpp:20-81) for visiting position `spin`
// with bits [0, spin) of (lo, hi) assigned: bit 0 = down allowed, bit 1 = up.
__device__ int enum_possible(const EnumDev& L, uint64_t lo, uint64_t hi, int spin) {
  uint64_t klo, khi;
  low_mask(spin, klo, khi);
  lo &= klo;
  hi &= khi;
  int res = 3;
  for (int ud = 0; ud < 2; ++ud) {
    uint64_t slo = lo, shi = hi;
    if (ud) {
      if (spin < 64) slo |= 1ull << spin;
      else shi |= 1ull << (spin - 64);
    }
    for (int k = 0; k < 2; ++k) {
      const int t = k == 0 ? L.nb0[spin] : L.nb1[spin];
      const ulonglong2 tm = L.tmask[t];
      const int q = __popcll(slo & tm.x) + __popcll(shi & tm.y);
      const int nk = __popcll(tm.x & klo) + __popcll(tm.y & khi) + 1;
      const int nu = L.tsize[t] - nk;
      if (q + nu < L.tmin[t] || q > L.tmax[t]) {
        res &= ~(1 << ud);
        break;
      }
    }
  }
  return res;
}

// Children of a node (fork_state_impl, pyro_tree.cpp:168-189): bit 0 = down
// child, bit 1 = up child, bit 2 / 3 = that child spends a spinon pair.
__device__ __forceinline__ uint8_t enum_children(const EnumDev& L, uint64_t lo, uint64_t hi,
                                                 int spin, int pairs) {
  const int poss = enum_possible(L, lo, hi, spin);
  const bool may_pair = pairs < L.pairs;
  uint8_t o = 0;
  if (poss & 1) o |= 1;
  else if (may_pair) o |= 1 | 4;
  if (poss & 2) o |= 2;
  else if (may_pair) o |= 2 | 8;
  return o;
}

// Visiting-order bits -> spin-index bits.
__device__ __forceinline__ void unpermute(const EnumDev& L, uint64_t lo, uint64_t hi,
                                          uint64_t& olo, uint64_t& ohi) {
  olo = ohi = 0;
  for (int k = 0; k < L.n; ++k) {
    const uint64_t bit = k < 64 ? (lo >> k) & 1 : (hi >> (k - 64)) & 1;
    const int s = L.order[k];
    if (s < 64) olo |= bit << s;
    else ohi |= bit << (s - 64);
  }
}

// Per-level flags of the walk, one bit per visiting position (< 128) in two
// words, so the walk keeps no per-level arrays (which would live in scratch
// memory) and backtracks with one count-leading-zeros.
struct Bits128 {
  uint64_t w[2];
};

__device__ __forceinline__ void bset(Bits128& b, int k, bool v) {
  const uint64_t m = 1ull << (k & 63);
  uint64_t& x = b.w[k >> 6];
  x = v ? (x | m) : (x & ~m);
}

__device__ __forceinline__ bool bget(const Bits128& b, int k) {
  return (b.w[k >> 6] >> (k & 63)) & 1ull;
}

// Highest set bit below position k, or -1.
__device__ __forceinline__ int bhigh_below(const Bits128& b, int k) {
  uint64_t lo, hi;
  low_mask(k, lo, hi);
  const uint64_t h = b.w[1] & hi, l = b.w[0] & lo;
  if (h) return 127 - __clzll(h);
  if (l) return 63 - __clzll(l);
  return -1;
}

// Set bits below position k.
__device__ __forceinline__ int bcount_below(const Bits128& b, int k) {
  uint64_t lo, hi;
  low_mask(k, lo, hi);
  return __popcll(b.w[0] & lo) + __popcll(b.w[1] & hi);
}

// Depth-first walk of one frontier subtree: counts its leaves (FILL = false)
// or writes them from out[offset[i]] on (FILL = true).  Down child first,
// then up (the order of fork_state_impl); the walk order is the same in both
// passes.  Per level the walk remembers whether the up child is still to
// visit (pend), whether that child spends a spinon pair (pend_pair), and
// whether the child taken spent one (spent); a backtrack jumps straight to
// the deepest level with a pending up child.
template <bool FILL>
__global__ __launch_bounds__(256) void k_enum(EnumDev L, const ENode* __restrict__ nodes,
                                              int64_t n_nodes, uint64_t* __restrict__ counts,
                                              const uint64_t* __restrict__ offsets,
                                              uint64_t* __restrict__ out_lo,
                                              uint64_t* __restrict__ out_hi) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n_nodes) return;
  const ENode root = nodes[i];
  uint64_t lo = root.lo, hi = root.hi;
  uint64_t n_leaves = 0;
  uint64_t pos = FILL ? offsets[i] : 0;
  auto emit = [&]() {
    if constexpr (FILL) {
      uint64_t a, b;
      unpermute(L, lo, hi, a, b);
      out_lo[pos] = a;
      out_hi[pos] = b;
      ++pos;
    } else {
      ++n_leaves;
    }
  };
  auto set_spin = [&](int d, int b) {
    if (d < 64) lo = (lo & ~(1ull << d)) | (static_cast<uint64_t>(b) << d);
    else hi = (hi & ~(1ull << (d - 64))) | (static_cast<uint64_t>(b) << (d - 64));
  };
  const int d0 = root.spin;
  if (d0 == L.n) {
    emit();
  } else {
    Bits128 pend{{0, 0}}, pend_pair{{0, 0}}, spent{{0, 0}};
    int d = d0;
    int pairs = root.pairs;
    while (true) {
      bool back = false;
      if (d == L.n) {
        // bits above n are never set: every branch assigns bits below n only
        emit();
        back = true;
      } else {
        // bits at and above d are left over from other branches and are
        // masked off where they are read
        const uint8_t o = enum_children(L, lo, hi, d, pairs);
        if (o & 1) {
          set_spin(d, 0);
          const bool sp = (o >> 2) & 1;
          bset(spent, d, sp);
          bset(pend, d, (o & 2) != 0);
          bset(pend_pair, d, (o >> 3) & 1);
          pairs += sp;
          ++d;
        } else if (o & 2) {
          set_spin(d, 1);
          const bool sp = (o >> 3) & 1;
          bset(spent, d, sp);
          bset(pend, d, false);
          pairs += sp;
          ++d;
        } else {
          back = true;
        }
      }
      if (back) {
        const int p = bhigh_below(pend, d);
        if (p < d0) break;
        bset(pend, p, false);
        set_spin(p, 1);
        bset(spent, p, bget(pend_pair, p));
        pairs = root.pairs + bcount_below(spent, p + 1);  // no bit below d0 is set
        d = p + 1;
      }
    }
  }
  if constexpr (!FILL) counts[i] = n_leaves;
}

// Interleave sorted (lo, hi) columns into basis rows [lo, hi] (nw = 2).
__global__ void k_interleave(const uint64_t* __restrict__ lo, const uint64_t* __restrict__ hi,
                             int64_t n, uint64_t* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = lo[i];
  out[2 * i + 1] = hi[i];
}

struct DevBufs {
  std::vector<void*> p;
  ~DevBufs() {
    for (void* q : p) (void)hipFree(q);
  }
  template <typename T>
  hipError_t alloc(T** out, size_t n) {
    void* q = nullptr;
    const hipError_t e = hipMalloc(&q, std::max<size_t>(n * sizeof(T), 8));
    if (e == hipSuccess) {
      p.push_back(q);
      *out = static_cast<T*>(q);
    }
    return e;
  }
  void release(void* q) {
    auto it = std::find(p.begin(), p.end(), q);
    if (it != p.end()) {
      (void)hipFree(q);
      p.erase(it);
    }
  }
};

}  // namespace
}  // namespace qsi

using namespace qsi;

extern "C" int qsi_enumerate_basis_device(int n_spins, int n_tetra, const int32_t* tetra_members,
                                          const int32_t* order, int n_spinon_pairs,
                                          int bucket_bits, qsi_basis_t* out, int64_t* out_dim) {
  QSI_REQUIRE(out != nullptr, "qsi_enumerate_basis_device: null output");
  *out = nullptr;
  if (out_dim) *out_dim = 0;
  QSI_REQUIRE(n_spins >= 1 && n_spins <= 128,
              "qsi_enumerate_basis_device: n_spins must be in [1,128]");
  QSI_REQUIRE(n_tetra >= 1 && tetra_members && order,
              "qsi_enumerate_basis_device: null lattice");
  QSI_REQUIRE(n_spinon_pairs >= 0 && n_spinon_pairs < 256,
              "qsi_enumerate_basis_device: n_spinon_pairs must be in [0, 256)");
  QSI_REQUIRE(bucket_bits >= 0 && bucket_bits <= 32 && bucket_bits <= n_spins,
              "qsi_enumerate_basis_device: bucket_bits must be in [0, min(32, n_spins)]");
  enumr::EnumLattice L;
  {
    const int rc = enumr::build_lattice(n_spins, n_tetra, tetra_members, order, n_spinon_pairs, L);
    if (rc != QSI_OK) return rc;
  }
  const std::vector<enumr::Node> front = enumr::frontier(L, kFrontier);
  const int64_t nn = static_cast<int64_t>(front.size());
  std::vector<ENode> hnodes(front.size());
  for (size_t k = 0; k < front.size(); ++k)
    hnodes[k] = ENode{static_cast<uint64_t>(front[k].state),
                      static_cast<uint64_t>(front[k].state >> 64),
                      static_cast<int>(front[k].spin), static_cast<int>(front[k].pairs)};
  std::vector<ulonglong2> htm(n_tetra);
  for (int t = 0; t < n_tetra; ++t)
    htm[t] = make_ulonglong2(static_cast<uint64_t>(L.tmask[t]),
                             static_cast<uint64_t>(L.tmask[t] >> 64));
  DevBufs db;
  ulonglong2* dtm = nullptr;
  int *dtmin = nullptr, *dtmax = nullptr, *dtsize = nullptr, *dnb0 = nullptr, *dnb1 = nullptr,
      *dorder = nullptr;
  ENode* dnodes = nullptr;
  uint64_t *dcount = nullptr, *doff = nullptr;
  QSI_CHECK_HIP(db.alloc(&dtm, n_tetra));
  QSI_CHECK_HIP(db.alloc(&dtmin, n_tetra));
  QSI_CHECK_HIP(db.alloc(&dtmax, n_tetra));
  QSI_CHECK_HIP(db.alloc(&dtsize, n_tetra));
  QSI_CHECK_HIP(db.alloc(&dnb0, n_spins));
  QSI_CHECK_HIP(db.alloc(&dnb1, n_spins));
  QSI_CHECK_HIP(db.alloc(&dorder, n_spins));
  QSI_CHECK_HIP(db.alloc(&dnodes, hnodes.size()));
  QSI_CHECK_HIP(db.alloc(&dcount, hnodes.size() + 1));
  QSI_CHECK_HIP(db.alloc(&doff, hnodes.size() + 1));
  QSI_CHECK_HIP(hipMemcpy(dtm, htm.data(), htm.size() * sizeof(ulonglong2), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dtmin, L.tmin.data(), n_tetra * sizeof(int), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dtmax, L.tmax.data(), n_tetra * sizeof(int), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dtsize, L.tsize.data(), n_tetra * sizeof(int), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dnb0, L.nb0.data(), n_spins * sizeof(int), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dnb1, L.nb1.data(), n_spins * sizeof(int), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dorder, order, n_spins * sizeof(int), hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemcpy(dnodes, hnodes.data(), hnodes.size() * sizeof(ENode),
                          hipMemcpyHostToDevice));
  QSI_CHECK_HIP(hipMemset(dcount + nn, 0, sizeof(uint64_t)));
  EnumDev E{dtm, dtmin, dtmax, dtsize, dnb0, dnb1, dorder, n_spins, n_spinon_pairs};
  const uint32_t grid = static_cast<uint32_t>((nn + 255) / 256);
  hipLaunchKernelGGL(k_enum<false>, dim3(grid), dim3(256), 0, 0, E, dnodes, nn, dcount,
                     static_cast<const uint64_t*>(nullptr), static_cast<uint64_t*>(nullptr),
                     static_cast<uint64_t*>(nullptr));
  QSI_CHECK_LAUNCH();
  // offsets = exclusive sum of the counts; offsets[nn] = D
  size_t tmp_bytes = 0;
  void* tmp = nullptr;
  QSI_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, dcount, doff, nn + 1));
  QSI_CHECK_HIP(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, dcount, doff, nn + 1);
  (void)hipFree(tmp);
  QSI_CHECK_HIP(e);
  uint64_t D = 0;
  QSI_CHECK_HIP(hipMemcpy(&D, doff + nn, sizeof(uint64_t), hipMemcpyDeviceToHost));
  QSI_REQUIRE(D < (uint64_t{1} << 32) - 1,
              "qsi_enumerate_basis_device: more than 2^32 - 2 states (ranks are 32-bit)");
  const int64_t dim = static_cast<int64_t>(D);
  const bool two = n_spins > 64;
  uint64_t *lo_a = nullptr, *lo_b = nullptr, *hi_a = nullptr, *hi_b = nullptr;
  QSI_CHECK_HIP(db.alloc(&lo_a, D));
  QSI_CHECK_HIP(db.alloc(&lo_b, D));
  QSI_CHECK_HIP(db.alloc(&hi_a, D));
  if (two) QSI_CHECK_HIP(db.alloc(&hi_b, D));
  hipLaunchKernelGGL(k_enum<true>, dim3(grid), dim3(256), 0, 0, E, dnodes, nn,
                     static_cast<uint64_t*>(nullptr), doff, lo_a, hi_a);
  QSI_CHECK_LAUNCH();
  uint64_t* sorted_lo = lo_a;
  uint64_t* sorted_hi = hi_a;
  if (dim > 1) {
    // low word first, then a stable pass on the high word (LSD order)
    hipcub::DoubleBuffer<uint64_t> klo(lo_a, lo_b);
    const int lo_bits = std::min(n_spins, 64);
    if (two) {
      hipcub::DoubleBuffer<uint64_t> vhi(hi_a, hi_b);
      tmp_bytes = 0;
      QSI_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, klo, vhi, dim, 0,
                                                       lo_bits));
      QSI_CHECK_HIP(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
      e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, klo, vhi, dim, 0, lo_bits);
      (void)hipFree(tmp);
      QSI_CHECK_HIP(e);
      hipcub::DoubleBuffer<uint64_t> khi(vhi.Current(), vhi.Alternate());
      hipcub::DoubleBuffer<uint64_t> vlo(klo.Current(), klo.Alternate());
      tmp_bytes = 0;
      QSI_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, khi, vlo, dim, 0,
                                                       n_spins - 64));
      QSI_CHECK_HIP(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
      e = hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, khi, vlo, dim, 0, n_spins - 64);
      (void)hipFree(tmp);
      QSI_CHECK_HIP(e);
      sorted_lo = vlo.Current();
      sorted_hi = khi.Current();
    } else {
      tmp_bytes = 0;
      QSI_CHECK_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, tmp_bytes, klo, dim, 0, lo_bits));
      QSI_CHECK_HIP(hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1)));
      e = hipcub::DeviceRadixSort::SortKeys(tmp, tmp_bytes, klo, dim, 0, lo_bits);
      (void)hipFree(tmp);
      QSI_CHECK_HIP(e);
      sorted_lo = klo.Current();
    }
  }
  // the basis owns its states: one word per state (<= 64 spins) or [lo, hi]
  uint64_t* states = nullptr;
  if (two) {
    e = hipMalloc(&states, std::max<size_t>(static_cast<size_t>(dim) * 16, 16));
    if (e != hipSuccess) {
      set_error(std::string("qsi_enumerate_basis_device: hipMalloc(states): ") +
                hipGetErrorString(e));
      return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
    }
    if (dim > 0)
      hipLaunchKernelGGL(k_interleave, dim3(static_cast<uint32_t>((dim + 255) / 256)), dim3(256),
                         0, 0, sorted_lo, sorted_hi, dim, states);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) {
      (void)hipFree(states);
      set_error(std::string("qsi_enumerate_basis_device: ") + hipGetErrorString(e));
      return QSI_E_HIP;
    }
  } else {
    // keep the sorted low-word buffer itself
    QSI_CHECK_HIP(hipDeviceSynchronize());
    auto it = std::find(db.p.begin(), db.p.end(), static_cast<void*>(sorted_lo));
    if (it != db.p.end()) db.p.erase(it);
    states = sorted_lo;
  }
  db.release(lo_a);
  db.release(lo_b);
  db.release(hi_a);
  db.release(hi_b);
  qsi_basis_s* b = nullptr;
  const int rc = basis_adopt_device(states, dim, n_spins, bucket_bits, &b);
  if (rc != QSI_OK) return rc;
  *out = b;
  if (out_dim) *out_dim = dim;
  return QSI_OK;
}
