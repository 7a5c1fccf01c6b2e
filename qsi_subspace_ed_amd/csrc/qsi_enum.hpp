// Host side of the ice-rule basis enumeration shared by the host enumerator
// (qsi_enumerate_basis) and the device one (qsi_enumerate_basis_device):
// the lattice in visiting order, search-tree nodes, the branching rule and
// the breadth-first frontier that splits the tree into independent subtrees.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "qsi_internal.hpp"

namespace qsi {
namespace enumr {

using u128 = unsigned __int128;

struct EnumLattice {
  int n = 0;
  std::vector<u128> tmask;           // tetra bitmask in visiting order
  std::vector<int> tmin, tmax, tsize;
  std::vector<int> nb0, nb1;         // first two tetras of each (visited) spin
  unsigned pairs = 0;
};

struct Node {
  u128 state;
  unsigned spin;
  unsigned pairs;
};

// possible_spin_states (pyro_tree.cpp:20-81): bit 0 = down allowed, bit 1 =
// up allowed, judged on the spin's two tetrahedra with the spins visited so
// far.
inline int possible(const EnumLattice& L, const Node& c) {
  int res = 3;
  const u128 known = (static_cast<u128>(1) << c.spin) - 1;
  for (int ud = 0; ud < 2; ++ud) {
    const u128 st = ud ? (c.state | (static_cast<u128>(1) << c.spin)) : c.state;
    for (int k = 0; k < 2; ++k) {
      const int t = k == 0 ? L.nb0[c.spin] : L.nb1[c.spin];
      const u128 tm = L.tmask[t];
      const int q = __builtin_popcountll(static_cast<uint64_t>(st & tm)) +
                    __builtin_popcountll(static_cast<uint64_t>((st & tm) >> 64));
      const u128 km = tm & known;
      const int nk = __builtin_popcountll(static_cast<uint64_t>(km)) +
                     __builtin_popcountll(static_cast<uint64_t>(km >> 64)) + 1;
      const int nu = L.tsize[t] - nk;
      if (q + nu < L.tmin[t] || q > L.tmax[t]) {
        res &= ~(1 << ud);
        break;
      }
    }
  }
  return res;
}

// fork_state_impl (pyro_tree.cpp:168-189)
template <typename Push>
inline void fork(const EnumLattice& L, const Node& c, Push&& push) {
  const int poss = possible(L, c);
  const bool may_pair = c.pairs < L.pairs;
  const u128 up = static_cast<u128>(1) << c.spin;
  if (poss & 1) push(Node{c.state, c.spin + 1, c.pairs});
  else if (may_pair) push(Node{c.state, c.spin + 1, c.pairs + 1});
  if (poss & 2) push(Node{c.state | up, c.spin + 1, c.pairs});
  else if (may_pair) push(Node{c.state | up, c.spin + 1, c.pairs + 1});
}


// Lattice in visiting order from the tetrahedron list and the visiting order
// (order[k] = spin visited k-th).  Returns QSI_OK or an error code.
inline int build_lattice(int n_spins, int n_tetra, const int32_t* tetra_members,
                         const int32_t* order, int n_spinon_pairs, EnumLattice& L) {
  std::vector<int> iperm(n_spins, -1);
  for (int k = 0; k < n_spins; ++k) {
    const int sid = order[k];
    QSI_REQUIRE(sid >= 0 && sid < n_spins && iperm[sid] < 0,
                "qsi_enumerate_basis: order is not a permutation");
    iperm[sid] = k;
  }
  L.n = n_spins;
  L.pairs = static_cast<unsigned>(n_spinon_pairs);
  L.tmask.resize(n_tetra);
  L.tmin.resize(n_tetra);
  L.tmax.resize(n_tetra);
  L.tsize.resize(n_tetra);
  std::vector<std::vector<int>> nbs(n_spins);
  for (int t = 0; t < n_tetra; ++t) {
    u128 m = 0;
    for (int q = 0; q < 4; ++q) {
      const int sid = tetra_members[4 * t + q];
      QSI_REQUIRE(sid >= 0 && sid < n_spins, "qsi_enumerate_basis: tetra member out of range");
      m |= static_cast<u128>(1) << iperm[sid];
      nbs[iperm[sid]].push_back(t);
    }
    L.tmask[t] = m;
    L.tsize[t] = 4;
    L.tmin[t] = 4 / 2;        // tetra default rules (tetra_graph_io.hpp:45-46)
    L.tmax[t] = (4 + 1) / 2;
  }
  L.nb0.resize(n_spins);
  L.nb1.resize(n_spins);
  for (int i = 0; i < n_spins; ++i) {
    QSI_REQUIRE(nbs[i].size() >= 2, "qsi_enumerate_basis: every spin needs two tetrahedra");
    L.nb0[i] = nbs[i][0];
    L.nb1[i] = nbs[i][1];
  }
  return QSI_OK;
}

// Breadth-first expansion until at least `target` nodes (or all leaves).
inline std::vector<Node> frontier(const EnumLattice& L, size_t target) {
  std::vector<Node> front{Node{0, 0, 0}};
  while (front.size() < target) {
    std::vector<Node> next;
    bool any = false;
    for (const Node& c : front) {
      if (static_cast<int>(c.spin) == L.n) {
        next.push_back(c);
        continue;
      }
      any = true;
      fork(L, c, [&](const Node& x) { next.push_back(x); });
    }
    front.swap(next);
    if (!any) break;
  }
  return front;
}

}  // namespace enumr
}  // namespace qsi
