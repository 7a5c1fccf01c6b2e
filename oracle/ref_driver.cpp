// ORACLE — test infrastructure only.
//
// Thin C shim over the reference's own operator core, compiled from the
// reference sources where they lie (components/operator/src/operator.cpp and
// its headers) by oracle/Makefile into oracle/_ref/libref.so.  Tests use it to
// pin the restatement in qsi_oracle.cpp: PMR masks and applyState
// (components/operator/include/operator.hpp:215-467) and both basis searches
// (components/operator/src/operator.cpp:4-86).  The reference's apply
// (operator_matrix.cpp) and Lanczos need Eigen/Spectra, which are absent, so
// they are not built here.
#include <cstdint>
#include <cstring>
#include <vector>

#include "operator.hpp"

namespace {

// Derived views that expose the protected members the shim needs.
struct PmrView : SymbolicPMROperator {
  using SymbolicPMROperator::SymbolicPMROperator;
  explicit PmrView(const SymbolicPMROperator& o) : SymbolicPMROperator(o) {}
  void dump(uint64_t* m, int* s) const {
    const ZBasisBase::state_t v[4] = {X_mask, Z_mask, down_mask, up_mask};
    for (int k = 0; k < 4; ++k) {
      m[2 * k] = v[k].uint64[0];
      m[2 * k + 1] = v[k].uint64[1];
    }
    *s = sign;
  }
};

struct RefBasis : ZBasisInterp {
  void set(const uint64_t* lohi, int64_t n) {
    states.resize(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) states[i] = Uint128(lohi[2 * i + 1], lohi[2 * i]);
    find_bounds();
  }
  int search_bst(const state_t& s, idx_t& J) const { return ZBasisBST::search(s, J); }
};

}  // namespace

extern "C" {

// Returns 0, or -1 if the reference constructor throws.
int ref_pmr_from_spec(const char* spec, uint64_t* masks, int* sign) {
  try {
    PmrView p{std::string(spec)};
    p.dump(masks, sign);
    return 0;
  } catch (...) {
    return -1;
  }
}

// Applies the operator given by `spec` to state [lo, hi]; returns the sign.
int ref_apply_spec(const char* spec, uint64_t* state) {
  SymbolicPMROperator op{std::string(spec)};
  ZBasisBase::state_t s(state[1], state[0]);
  const int r = op.applyState(s);
  state[0] = s.uint64[0];
  state[1] = s.uint64[1];
  return r;
}

// a * b (operator*, operator.hpp:332-336) for two specs.
int ref_pmr_mul(const char* a, const char* b, uint64_t* masks, int* sign) {
  try {
    SymbolicPMROperator A{std::string(a)}, B{std::string(b)};
    PmrView c(A * B);
    c.dump(masks, sign);
    return 0;
  } catch (...) {
    return -1;
  }
}

void* ref_basis_new(const uint64_t* lohi, int64_t n) {
  auto* b = new RefBasis();
  b->set(lohi, n);
  return b;
}

void ref_basis_free(void* b) { delete static_cast<RefBasis*>(b); }

// rank[i] = index, -1 missing, -2 when ZBasisInterp::search throws
void ref_basis_search(const void* bp, const uint64_t* q, int64_t n,
                      int64_t* rank, int bst) {
  const RefBasis& b = *static_cast<const RefBasis*>(bp);
  for (int64_t i = 0; i < n; ++i) {
    ZBasisBase::state_t s(q[2 * i + 1], q[2 * i]);
    ZBasisBase::idx_t J = -1;
    try {
      const int f = bst ? b.search_bst(s, J) : b.search(s, J);
      rank[i] = f ? J : -1;
    } catch (...) {
      rank[i] = -2;
    }
  }
}

}  // extern "C"
