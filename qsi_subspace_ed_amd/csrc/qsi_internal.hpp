// Internal objects behind the C ABI handles.
#pragma once

#include <atomic>
#include <vector>

#include "qsi_common.hpp"

namespace qsi {

// Largest basis dimension: ranks are uint32, and a kernel with one thread per
// row (rounded up to whole 256-thread workgroups) must stay below the 2^32
// work-items a dispatch can hold.
constexpr int64_t kMaxDim = (int64_t{1} << 32) - 512;

// Off-diagonal term in pull form.  For a row state s the term contributes
//   c * sgn * (-1)^popcount(~s & Z) * x[rank(s ^ X)]
// when (s & m) == v.  With s' = s ^ X the pre-image, this is the reference's
// precondition (s' & down) == 0 && (s' & up) == up rewritten on s:
//   m = up | down,  v = up ^ (X & m).
struct OffTerm {
  uint64_t m[2];
  uint64_t v[2];
  uint64_t X[2];
  uint64_t Z[2];
  double c;     // coefficient (single-vector apply)
  double sgn;   // +-1 operator sign
  int32_t slot; // coefficient slot for block applies (-1: use c)
  int32_t cls;  // coefficient class (block applies), -1 when classes are off
};

// Off-diagonal terms of one coefficient class share their per-column
// coefficient in a block apply: slot `slot` of the column (or `c` when the
// apply has no per-column coefficients or slot < 0).  Block applies keep the
// classes' coefficients in registers instead of loading each term's.
struct CoefClass {
  double c;
  int32_t slot;
  int32_t pad;
};
constexpr int kMaxCoefClasses = 8;

// Diagonal term.  kind 0: c * sgn * (-1)^popcount(~s & Z) when (s & m) == v
// (a Z string times up/down projectors, m = up | down, v = up), else 0;
// kind 1: c * sgn * (2 popcount(s & Z) - popcount(Z))  (a sum of single-site
// Z terms that share one coefficient and slot, merged on the host).
struct DiagTerm {
  uint64_t m[2];
  uint64_t v[2];
  uint64_t Z[2];
  double c;
  double sgn;
  int32_t kind;
  int32_t slot;
};

}  // namespace qsi

struct qsi_basis_s {
  int64_t dim = 0;
  int n_spins = 0;
  int nw = 1;
  int kbits = 0;
  int n_seg = 0;  // bucket key fields (see BasisView)
  uint8_t seg_lo[qsi::kKeySegs] = {};
  uint8_t seg_w[qsi::kKeySegs] = {};
  int device = 0;
  uint64_t* d_states = nullptr;  // dim * nw words
  uint32_t* d_table = nullptr;   // 2^kbits + 1
  ulonglong2* d_hash = nullptr;  // hash index (null if not built)
  uint32_t* d_slots = nullptr;   // slot index, 2 * dim entries (null if not built)
  uint32_t* d_lowkeys = nullptr; // low keys, dim entries (null if not built)
  int n_low = 0;                 // low-key runs (see BasisView)
  uint8_t low_lo[qsi::kLowSegs] = {};
  uint8_t low_w[qsi::kLowSegs] = {};
  uint64_t infmask[2] = {0, 0};  // informative positions (bit p: spin p)
  int hbits = 0;
  std::atomic<int> refs{1};

  qsi::HashView hview() const {
    qsi::HashView h;
    h.slots = d_hash;
    h.mask = (hbits > 0) ? ((uint64_t{1} << hbits) - 1) : 0;
    h.hbits = hbits;
    return h;
  }

  qsi::BasisView view() const {
    qsi::BasisView v;
    v.states = d_states;
    v.table = d_table;
    v.slots = d_slots;
    v.lowkeys = d_lowkeys;
    v.dim = dim;
    v.kbits = kbits;
    v.n_spins = n_spins;
    v.n_seg = n_seg;
    for (int q = 0; q < qsi::kKeySegs; ++q) {
      v.seg_lo[q] = seg_lo[q];
      v.seg_w[q] = seg_w[q];
    }
    v.n_low = n_low;
    for (int q = 0; q < qsi::kLowSegs; ++q) {
      v.low_lo[q] = low_lo[q];
      v.low_w[q] = low_w[q];
    }
    return v;
  }
};

struct qsi_op_s {
  qsi_basis_s* basis = nullptr;
  int n_off = 0;
  int n_diag = 0;
  int n_slots = 0;
  bool has_z = false;  // any off-diagonal term with a Z string
  int variant = 0;     // apply kernel: one of the QSI_APPLY_* values
  qsi::OffTerm* d_off = nullptr;
  qsi::DiagTerm* d_diag = nullptr;
  std::vector<qsi::OffTerm> h_off;
  std::vector<qsi::DiagTerm> h_diag;
  int n_cls = 0;  // coefficient classes (0: more than kMaxCoefClasses)
  qsi::CoefClass* d_cls = nullptr;
  std::vector<qsi::CoefClass> h_cls;
  // cached connectivity (qsi_op_cache_connectivity): CSR over the rows (64-bit
  // offsets: a 96-spin operator has 3.3e10 nonzeros), the
  // rank of each connected state (0xFFFFFFFF if absent) and its coefficient
  // class | negative-sign bit << 7, off-diagonal terms in term order
  uint64_t* d_rowptr = nullptr;
  uint32_t* d_cj = nullptr;
  uint8_t* d_cm = nullptr;
  int64_t conn_nnz = -1;  // -1: not cached (connected terms; absent ones included)
  int conn_pad = 1;       // rows padded to multiples of conn_pad entries (1 or 4)
  int blk_cfg = 1;        // block-apply configuration (qsi_op_autotune)
  int cached_pipe = 8;    // gathers in flight per lane of the cached apply (4-32)
  // how the low-key kernels confirm a candidate (qsi_op_verify_lowkeys):
  // 0 compare its state; 1 compare its low key (no connected state outside
  // the basis shares a bucket key and low key with a basis state); 2 no
  // compare (every connected state is in the basis).  A property of the terms
  // and the basis, both fixed, so it holds for any build of the keys.
  int lowkeys_check = 0;
};

namespace qsi {

void basis_retain(qsi_basis_s* b);
void basis_release(qsi_basis_s* b);
// takes ownership of d_states (dim x nw words, device)
int basis_adopt_device(uint64_t* d_states, int64_t dim, int n_spins, int bucket_bits,
                       qsi_basis_s** out);

// Launch helpers shared by the solvers (all stream ordered).
// y = yscale * y_in + H x  (y_in may be null or alias y); if dot_out is
// non-null, *dot_out (device) = sum_i x[i] * y[i] after the update.
int launch_apply(const qsi_op_s* op, const double* x, double* y,
                 const double* y_in, double yscale, double* dot_partials,
                 double* dot_out, hipStream_t s);
// as launch_apply, with the y_in scale multiplied by yscale_dev[0] (device)
int launch_apply_dev(const qsi_op_s* op, const double* x, double* y,
                     const double* y_in, double yscale,
                     const double* yscale_dev, double* dot_partials,
                     double* dot_out, hipStream_t s);
// as launch_apply_dev, with x read as xs_dev[0] * x and y_in as
// ys_dev[0] * y_in (either may be null: scale 1).  The products are the
// values a separate scaling pass would have stored, so the result has the
// same bits as applying to the scaled vectors.  Scales need dot_partials
// (only the dot-product kernel instances read them).
int launch_apply_scaled(const qsi_op_s* op, const double* x, double* y,
                        const double* y_in, double yscale, const double* yscale_dev,
                        const double* xs_dev, const double* ys_dev, double* dot_partials,
                        double* dot_out, hipStream_t s);
int64_t apply_num_blocks(const qsi_op_s* op);

// streaming vector kernels; partial buffers need max_vec_partials() doubles
// (multidot_partials() for launch_multidot)
int max_vec_partials();
int multidot_partials();
int launch_dot(const double* a, const double* b, int64_t n, double* partials,
               double* out, hipStream_t s);
int launch_axpy(double alpha, const double* x, double* y, int64_t n,
                hipStream_t s, const double* xs_dev = nullptr);
int launch_scal(double alpha, double* x, int64_t n, hipStream_t s);
// u -= alpha_dev[0] v; nrm_out[0] = ||u||, nrm_out[1] = 1/||u||; v is read as
// vs_dev[0] * v when vs_dev is given
int launch_axpy_norm(const double* alpha_dev, const double* v, double* u,
                     int64_t n, double* partials, double* nrm_out,
                     hipStream_t s, const double* vs_dev = nullptr);
// out[j] = V[:, j] . w for j < m (V column-major, ld ldv); acc_out += out
int launch_multidot(const double* V, int64_t ldv, int m, const double* w,
                    int64_t n, double* partials, double* out, double* acc_out,
                    hipStream_t s);
// w -= V[:, 0:m] h
int launch_gemv_sub(const double* V, int64_t ldv, int m, const double* h,
                    double* w, int64_t n, hipStream_t s);
// Second and (when needed) third classical Gram-Schmidt passes of w against
// V[:, 0:m] after a first pass gave h = V^T w: w -= V h with g[0..m) =
// V^T w_new and g[m] = |w_new|^2 from the same read (g: m + 1 doubles); if
// |g[0..m)| > 1e-12 |w_new| then w -= V g, g is added to hcol[0..m) and
// flag[0] = 1.  beta[0] = |w| at the end, beta[1] = 1 / beta[0] (0 if w
// vanishes), and w is scaled by beta[1] in the same final pass.  All on the
// device (no host synchronisation).
int launch_gs_reorth(const double* V, int64_t ldv, int m, const double* h, double* w,
                     int64_t n, double* partials, double* g, double* hcol, int* flag,
                     double* beta, hipStream_t s);
// out[:, 0:k] = V[:, 0:m] Y  (Y m x k column-major, device)
int launch_combine(const double* V, int64_t ldv, int m, const double* Y, int k,
                   double* out, int64_t ldo, int64_t n, hipStream_t s);
int launch_randn(double* v, int64_t n, uint64_t seed, hipStream_t s);
int launch_copy_scaled(const double* s_dev, const double* y, double* x,
                       int64_t n, hipStream_t s);
// y += c * s_dev[0] * x
int launch_axpy_dev(const double* s_dev, double c, const double* x, double* y,
                    int64_t n, hipStream_t s);

// Psi[:, b] += c[b] * v for b < nb (c on the device; Psi column-major); v is
// read as vs_dev[0] * v when vs_dev is given
int launch_accum_cols(const double* v, double* Psi, int64_t ld, int nb, const double* c,
                      int64_t n, hipStream_t s, const double* vs_dev = nullptr);

// The 3-term Lanczos coefficient run behind qsi_lanczos_coeffs on
// caller-provided device buffers: v (D, the start vector, normalised here;
// clobbered), u (D), partials (lanczos_coeffs_partials(op) doubles), ab
// (3 M doubles), scal (2 doubles).  Allocates nothing; synchronises.
int64_t lanczos_coeffs_partials(const qsi_op_s* op);
int lanczos_coeffs_run(const qsi_op_s* op, double* v, double* u, double* partials, double* ab,
                       double* scal, int M, double* h_alpha, double* h_beta, int* n_steps,
                       hipStream_t s);

}  // namespace qsi
