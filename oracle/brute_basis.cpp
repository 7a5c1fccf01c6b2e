// ORACLE — test infrastructure only.
//
// Brute-force restatement of the basis builder's acceptance rule for lattices
// of at most 32 spins: every one of the 2^W states is replayed along the
// visiting order, and a state is kept when the number of visiting positions
// whose value the rule does not allow is at most n_spinon_pairs.
//
//   possible_spin_states  components/build_basis/src/pyro_tree.cpp:20-81:
//       with the spins before position idx known (the rest 0), value u of
//       spin idx is allowed unless, on one of the spin's first two
//       tetrahedra, the up count Q with spin idx = u satisfies
//       Q + (unknown members) < min or Q > max (min = max = 2 for a
//       4-member tetrahedron, components/common/include/tetra_graph_io.hpp:
//       45-46);
//   fork_state_impl       pyro_tree.cpp:168-189: an allowed value costs
//       nothing, a disallowed one costs one spinon pair while pairs remain.
//
// The depth-first search of the reference visits exactly the states this
// replay accepts (each state is one path of the search tree), so the kept
// states, sorted, are the basis gen_spinon_basis writes
// (components/build_basis/gen_spinon_basis.cpp:16-73, after undoing the
// visiting-order permutation and sorting).
#include <omp.h>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

struct Lat {
  int n = 0;
  // visiting position -> its first two tetrahedra, as member masks over
  // visiting positions
  std::vector<uint32_t> t0, t1;
  std::vector<int> s0, s1;  // member counts
};

int popc(uint32_t x) { return __builtin_popcount(x); }

// bit 0: down allowed, bit 1: up allowed (pyro_tree.cpp:20-81)
int possible(const Lat& L, uint32_t known_bits, int idx) {
  int res = 3;
  const uint32_t known_mask = idx == 0 ? 0u : (idx >= 32 ? ~0u : ((1u << idx) - 1u));
  for (int u = 0; u < 2; ++u) {
    const uint32_t st = (known_bits & known_mask) | (static_cast<uint32_t>(u) << idx);
    for (int k = 0; k < 2; ++k) {
      const uint32_t tm = k == 0 ? L.t0[idx] : L.t1[idx];
      const int size = k == 0 ? L.s0[idx] : L.s1[idx];
      const int q = popc(st & tm);
      const int known = popc(tm & known_mask) + 1;
      const int unknown = size - known;
      const int lo = size / 2, hi = (size + 1) / 2;
      if (q + unknown < lo || q > hi) {
        res &= ~(1 << u);
        break;
      }
    }
  }
  return res;
}

}  // namespace

extern "C" {

// States (spin-index bits, ascending) of the n_pairs sector: written to out
// with their count of disallowed positions in bad (so one scan serves every
// sector up to n_pairs), up to cap entries; returns the count, or -1 on bad
// input.  tetra_members: n_tetra x 4 spin ids; order[k] = spin visited k-th.
int64_t oracle_brute_basis(int n_spins, int n_tetra, const int32_t* tetra_members,
                           const int32_t* order, int n_pairs, uint64_t* out, int8_t* bad_out,
                           int64_t cap) {
  if (n_spins < 1 || n_spins > 32 || n_tetra < 1 || n_pairs < 0) return -1;
  const int W = n_spins;
  std::vector<int> pos(W, -1);
  for (int k = 0; k < W; ++k) {
    if (order[k] < 0 || order[k] >= W || pos[order[k]] >= 0) return -1;
    pos[order[k]] = k;
  }
  Lat L;
  L.n = W;
  L.t0.assign(W, 0);
  L.t1.assign(W, 0);
  L.s0.assign(W, 0);
  L.s1.assign(W, 0);
  std::vector<int> seen(W, 0);
  for (int t = 0; t < n_tetra; ++t) {
    uint32_t m = 0;
    for (int q = 0; q < 4; ++q) {
      const int sid = tetra_members[4 * t + q];
      if (sid < 0 || sid >= W) return -1;
      m |= 1u << pos[sid];
    }
    for (int q = 0; q < 4; ++q) {
      const int p = pos[tetra_members[4 * t + q]];
      if (seen[p] == 0) {
        L.t0[p] = m;
        L.s0[p] = 4;
      } else if (seen[p] == 1) {
        L.t1[p] = m;
        L.s1[p] = 4;
      }
      ++seen[p];
    }
  }
  for (int p = 0; p < W; ++p)
    if (seen[p] < 2) return -1;  // every spin needs two tetrahedra
  const uint64_t total = uint64_t{1} << W;
  const int nthr = omp_get_max_threads();
  std::vector<std::vector<uint64_t>> part(static_cast<size_t>(nthr));
  std::vector<std::vector<int8_t>> partb(static_cast<size_t>(nthr));
#pragma omp parallel num_threads(nthr)
  {
    const int tid = omp_get_thread_num();
    // contiguous ranges keep the concatenation ascending
    const uint64_t a = total * tid / nthr, b = total * (tid + 1) / nthr;
    std::vector<uint64_t>& mine = part[static_cast<size_t>(tid)];
    std::vector<int8_t>& mineb = partb[static_cast<size_t>(tid)];
    for (uint64_t s = a; s < b; ++s) {
      uint32_t vis = 0;  // the state in visiting-position bits
      for (int k = 0; k < W; ++k) vis |= static_cast<uint32_t>((s >> order[k]) & 1u) << k;
      int bad = 0;
      for (int k = 0; k < W && bad <= n_pairs; ++k) {
        const int u = (vis >> k) & 1u;
        if (!((possible(L, vis, k) >> u) & 1)) ++bad;
      }
      if (bad <= n_pairs) {
        mine.push_back(s);
        mineb.push_back(static_cast<int8_t>(bad));
      }
    }
  }
  int64_t n = 0;
  for (size_t t = 0; t < part.size(); ++t)
    for (size_t j = 0; j < part[t].size(); ++j) {
      if (n < cap && out) out[n] = part[t][j];
      if (n < cap && bad_out) bad_out[n] = partb[t][j];
      ++n;
    }
  return n;
}

}  // extern "C"
