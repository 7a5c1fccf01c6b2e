/*
 * qsi_gpu.h — C ABI of the MI355X (gfx950) matrix-free ring-exchange H·v and
 * Lanczos library.
 *
 * The reference computes H·v on the host with OpenMP (LazyOpSum, push form with
 * atomics) and drives it from Spectra's SymEigsSolver (`mfsparse`) or its own
 * 3-term Lanczos (`mfeig0`).  This library keeps the same objects — a sorted
 * basis of 128-bit Z-basis states, a list of permutation-matrix ("PMR") terms
 * with coefficients, an operator that applies y = Hx — but every vector and
 * basis lives in GPU memory and every pass over the basis is a HIP kernel.
 *
 * Conventions
 *   - All functions return 0 on success and a negative QSI_E* code on failure;
 *     qsi_last_error() gives a message (thread local).  Nothing throws across
 *     this boundary.
 *   - A "state" is a pair of uint64 words [lo, hi] (bit i = spin i, 1 = up),
 *     exactly the HDF5 (D,2) layout of the reference's basis files
 *     (components/common/include/basis_io_h5.hpp:30-70,90-141).
 *   - Vector arguments named d_* are device pointers; h_* are host pointers.
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).  All
 *     device work is stream ordered; an entry point returns once its work is
 *     enqueued unless its comment says it synchronises.
 *   - A handle (basis, operator, engine) is used from one host thread at a
 *     time; different handles may be used from different threads at once.
 */
#ifndef QSI_GPU_H
#define QSI_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void *qsi_stream_t;
typedef struct qsi_basis_s *qsi_basis_t;
typedef struct qsi_op_s *qsi_op_t;

enum {
  QSI_OK = 0,
  QSI_E_INVALID = -1,   /* bad argument (null pointer, bad size, unsorted basis...) */
  QSI_E_NOMEM = -2,     /* device or host allocation failed                          */
  QSI_E_HIP = -3,       /* a HIP runtime call failed                                 */
  QSI_E_NOTFOUND = -4,  /* a connected state is missing from the basis              */
  QSI_E_NOCONV = -5,    /* an iterative solver hit its iteration limit               */
  QSI_E_UNSUPPORTED = -6
};

/*
 * One operator term in the reference's SymbolicPMROperator representation
 * (components/operator/include/operator.hpp:343-347): it maps |s> to
 *   sign * (-1)^popcount(~s' & Z) * |s'>,  s' = s ^ X,
 * when (s & down) == 0 and (s & up) == up, and to 0 otherwise
 * (applyState, operator.hpp:232-254).  `coeff` is the SymbolicOpSum
 * coefficient (operator.hpp:489-495).  Words are [lo, hi].
 * A term with X == 0 is diagonal.
 */
typedef struct {
  double coeff;
  uint64_t X[2];
  uint64_t Z[2];
  uint64_t up[2];
  uint64_t down[2];
  int32_t sign;
  int32_t slot; /* coefficient slot for block applies with per-column
                   coefficients (qsi_op_apply_block); -1 = use `coeff`   */
} qsi_term_t;

/* ---- library ------------------------------------------------------------ */

/* Message of the last failing call on this thread ("" if none). */
const char *qsi_last_error(void);
/* Version string, e.g. "qsi-gpu 0.1 gfx950". */
const char *qsi_version(void);
/* Number of visible GPUs (0 if none).  Does not fail. */
int qsi_device_count(void);

/* ---- basis ---------------------------------------------------------------
 * Replaces ZBasisInterp / ZBasisBST (components/operator/include/operator.hpp:
 * 153-204; load_from_file + find_bounds, components/operator/src/operator.cpp:
 * 73-86,138-159).  The host array (D rows of [lo,hi], strictly increasing as
 * 128-bit integers) is copied to the device once; a bucket table (2^K+1
 * uint32 ranks over a K-bit key, see qsi_basis_key_fields) is built on the
 * device.  n_spins bounds the state bits (1..128); D <= 2^32 - 512.
 * bucket_bits (K, at most 32 and at most n_spins; a 32-bit key is a 16 GiB
 * table) = 0 picks K automatically (at most 31).  Synchronises. */
int qsi_basis_create(const uint64_t *h_lohi, int64_t dim, int n_spins,
                     int bucket_bits, qsi_basis_t *out);
/* As qsi_basis_create; with_hash = 1 also builds a device hash index
 * (state -> rank, open addressing, load <= 1/2, 16 bytes per slot for <= 64
 * spins, 32 above) used by the QSI_APPLY_HASH* kernels.  qsi_basis_create
 * does not build it (with_hash = 0).  No counterpart in the reference. */
int qsi_basis_create_ex(const uint64_t *h_lohi, int64_t dim, int n_spins,
                        int bucket_bits, int with_hash, qsi_basis_t *out);
int qsi_basis_has_hash(qsi_basis_t b);
/* Builds (enable = 1) or drops (enable = 0) the basis's slot index: a second
 * level under the bucket table for bases whose buckets stay large.  Bucket
 * [lo, lo + n) owns slots [2 lo, 2 lo + 2n) of one uint32 array (8 bytes per
 * state); each state sits in one slot as (16-bit hash fingerprint, rank - lo),
 * placed by linear probing from a hash of the state.  A lookup reads the
 * bucket bounds, then slots until an empty one or a fingerprint match, whose
 * state is compared: three dependent loads where a bisection of an n-state
 * bucket needs log2(n) + 2.  Used by the QSI_APPLY_SLOT* kernels and by
 * qsi_basis_search.  Fails (QSI_E_INVALID) if a bucket holds more than 65535
 * states.  Dropping synchronises the device.  No counterpart in the
 * reference (its buckets are searched by bisection, operator.cpp:38-70). */
int qsi_basis_build_slots(qsi_basis_t b, int enable);
int qsi_basis_has_slots(qsi_basis_t b);
/* Builds (enable = 1) or drops (enable = 0) the basis's low keys: for every
 * state, the informative positions below the bucket key's lowest bit (the
 * positions at which neighbouring basis states first differ) packed into one
 * uint32 (4 bytes per state).  Within a bucket the low keys are distinct and
 * ordered like the states, so the QSI_APPLY_LOW* kernels bisect a bucket on
 * them and load one candidate state per lookup.  Fails (QSI_E_INVALID) if
 * more than 32 informative positions lie below the key.  Dropping
 * synchronises the device.  No counterpart in the reference (its buckets are
 * bisected on whole states, operator.cpp:38-70). */
int qsi_basis_build_lowkeys(qsi_basis_t b, int enable);
int qsi_basis_has_lowkeys(qsi_basis_t b);
int qsi_basis_destroy(qsi_basis_t b);
int64_t qsi_basis_dim(qsi_basis_t b);
int qsi_basis_n_spins(qsi_basis_t b);
int qsi_basis_bucket_bits(qsi_basis_t b);
/* The bucket key: the concatenation, highest first, of *n_fields (<= 4) bit
 * fields [lo[f], lo[f] + width[f]) of a state, K bits in all, covering the
 * highest bit positions at which neighbouring basis states first differ
 * (positions fixed by the bits above them are left out between fields); the
 * top K bits when there are none.  lo and width hold 4 entries.  No
 * counterpart in the reference (its bucket key is the high 64-bit word,
 * operator.cpp:73-86). */
int qsi_basis_key_fields(qsi_basis_t b, int *n_fields, int *lo, int *width);
/* Device-side state -> rank lookup of n states (test hook and building block;
 * replaces ZBasisInterp::search, operator.cpp:38-70).  d_lohi: n rows of
 * [lo,hi] in device memory; d_rank receives the rank or -1 if absent.  Goes
 * through the slot index when the basis has one, else bisects the bucket. */
int qsi_basis_search(qsi_basis_t b, const uint64_t *d_lohi, int64_t n,
                     int64_t *d_rank, qsi_stream_t stream);

/* ---- operator ------------------------------------------------------------
 * Replaces LazyOpSum<double, ZBasisInterp> (components/operator/include/
 * operator_matrix.hpp:10-118).  Borrows the basis (kept alive by a
 * reference count), copies the terms.  n_slots > 0 declares coefficient slots
 * for qsi_op_apply_block. */
int qsi_op_create(qsi_basis_t b, const qsi_term_t *terms, int n_terms,
                  int n_slots, qsi_op_t *out);
int qsi_op_destroy(qsi_op_t op);
int64_t qsi_op_dim(qsi_op_t op);
int qsi_op_n_offdiag(qsi_op_t op);
int qsi_op_n_diag(qsi_op_t op);

/* Apply kernel choice (no counterpart in the reference; a tuning knob).
 * QSI_APPLY_HASH{,4}: each lane owns a row and finds its connected states in
 * the hash index, 1/4 lookups in flight per lane.  QSI_APPLY_LANE{,2,4,8}: the
 * same through the bucket table and a bisection of the bucket.
 * QSI_APPLY_WINDOW: per wavefront, one bucket lookup per term, then the
 * connected rows search a 64-state window staged in LDS.  QSI_APPLY_CACHED:
 * gathers through the cached connectivity (qsi_op_cache_connectivity; setting
 * it requires the cache, dropping the cache resets it to AUTO).
 * QSI_APPLY_SLOT{2,3,4}: as LANE*, through the slot index under each bucket
 * (qsi_basis_build_slots), 2/3/4 lookups in flight per lane; setting one
 * requires the index; if the index is dropped later the apply falls back to
 * LANE4 (SLOT4) or LANE2.  QSI_APPLY_LOW{2,4,8}: as LANE*, bisecting each
 * bucket on the basis's low keys (qsi_basis_build_lowkeys); setting one
 * requires them, and without them later the apply falls back to LANE* of the
 * same depth.  AUTO = CACHED when the connectivity is cached, else, when the
 * basis has low keys, LOW8 after qsi_op_verify_lowkeys found mode 2 and LOW4
 * otherwise, else LANE4.
 * All give bit-identical y. */
enum {
  QSI_APPLY_AUTO = 0,
  QSI_APPLY_WINDOW = 1,
  QSI_APPLY_LANE = 2,
  QSI_APPLY_LANE2 = 3,
  QSI_APPLY_LANE4 = 4,
  QSI_APPLY_LANE8 = 5,
  QSI_APPLY_HASH = 6,
  QSI_APPLY_HASH4 = 7,
  QSI_APPLY_CACHED = 8,
  QSI_APPLY_SLOT2 = 9,
  QSI_APPLY_SLOT4 = 10,
  QSI_APPLY_SLOT3 = 11,
  QSI_APPLY_LOW2 = 12,
  QSI_APPLY_LOW4 = 13,
  QSI_APPLY_LOW8 = 14
};
int qsi_op_set_variant(qsi_op_t op, int variant);
/* Checks, once per operator, by one pass over the rows (about the cost of one
 * apply; synchronises), how the QSI_APPLY_LOW* kernels may confirm the
 * candidate a low-key lookup ends on.  *mode = 1 when every connected state
 * its terms reach from a basis state either is that candidate or differs
 * from it in bucket key or low key: the kernels then compare the candidate's
 * low key instead of loading its state.  *mode = 2 when, in addition, every
 * connected state is in the basis (a closed sector, such as ring flips on a
 * full spinon sector): the kernels take the candidate with no compare.
 * *mode = 0 otherwise: states are compared, as before the check.  The same
 * candidates in every mode, so the same y.  Needs the basis's low keys.  No
 * counterpart in the reference. */
int qsi_op_verify_lowkeys(qsi_op_t op, int *mode);
int qsi_op_get_variant(qsi_op_t op);

/* y = H x (accumulate = 0) or y += H x (accumulate = 1).
 * accumulate = 0 is LazyOpSum::evaluate / LazyOpSumProd::perform_op
 * (operator_matrix.hpp:32-36,129-131); accumulate = 1 is evaluate_add
 * (components/operator/src/operator_matrix.cpp:108-112) and the
 * projED::RealApplyFn contract (components/linalg_routines/include/
 * common_bits.hpp:18-21).  Off-diagonal contributions with |x_j| <= 1e-15
 * are skipped (APPLY_TOL, operator_matrix.hpp:113).  x and y must not alias. */
int qsi_op_apply(qsi_op_t op, const double *d_x, double *d_y, int accumulate,
                 qsi_stream_t stream);

/* Same contract with host vectors (copies in and out; synchronises).  This is
 * the literal drop-in for LazyOpSumProd::perform_op under a host eigensolver;
 * it includes PCIe traffic. */
int qsi_op_apply_host(qsi_op_t op, const double *h_x, double *h_y,
                      int accumulate);

/* Block apply: Y[:, r] (+)= H_r X[:, r] for r < R, X and Y row-major D x R
 * with leading dimensions ldx, ldy >= R.  If d_coeffs is non-null it holds
 * n_slots x R doubles (slot-major): term t with slot k >= 0 uses
 * d_coeffs[k*R + r] for column r instead of its own coeff. */
int qsi_op_apply_block(qsi_op_t op, const double *d_X, int64_t ldx,
                       double *d_Y, int64_t ldy, int R, const double *d_coeffs,
                       int accumulate, qsi_stream_t stream);

/* Keep (enable = 1) or drop (0) the operator's connectivity on the device:
 * for every row, in term order, the rank of each connected state and the
 * term's coefficient class and sign (5 bytes per nonzero plus 8 per row;
 * when the padded arrays leave a quarter of the free device memory unused,
 * rows are padded to multiples of four entries, up to 15 more bytes per
 * row, which the applies read one aligned group at a time).  The
 * single-vector apply (QSI_APPLY_CACHED) and the block applies
 * (qsi_op_apply_block, qsi_lanczos_block, qsi_engine_*) then gather through
 * it instead of testing and searching, with identical results; worth it
 * when the same operator is applied many times.  *nnz (optional) receives
 * the cached nonzero count, padding excluded (-1 when not cached).
 * enable = 2 keeps it with rows unpadded whatever the free memory.  An
 * existing cache keeps its layout (drop it first to rebuild);
 * qsi_op_connectivity_padding reports it: 4 (padded), 1, or -1 (none).
 * Requires at most 8 distinct (slot, coeff) pairs among the off-diagonal
 * terms.  Synchronises.  The counterpart of the reference's `sparse`
 * algorithm, which materialises H (operator_matrix.hpp:79-105) instead. */
int qsi_op_cache_connectivity(qsi_op_t op, int enable, int64_t *nnz);
int qsi_op_connectivity_padding(qsi_op_t op);

/* Gathers of x in flight per lane of the cached apply (4, 8, 16 or 32).
 * Chosen by qsi_op_autotune; set it to pin one depth (results are
 * bit-identical for all).  No counterpart in the reference. */
int qsi_op_set_cached_pipe(qsi_op_t op, int pipe);
int qsi_op_get_cached_pipe(qsi_op_t op);
/* Block-apply configuration (bits 0-1: kernel 0-2, bit 2: half-size grid,
 * bit 3: two columns per lane in the cached kernel, used when X's rows are
 * 16-byte aligned and hold an even number of columns) of
 * qsi_op_apply_block, qsi_lanczos_block (the first sector's) and engines
 * created afterwards.  Chosen by qsi_op_autotune / qsi_block_autotune; set it
 * to pin one.  No counterpart in the reference. */
int qsi_op_set_block_cfg(qsi_op_t op, int cfg);
int qsi_op_get_block_cfg(qsi_op_t op);

/* Autotuner: times the apply kernels on this operator (reps calls each, after
 * a warm-up, with HIP events) and keeps the fastest.  R = 0 tunes the
 * single-vector kernel (the QSI_APPLY_* variants that apply: hash variants
 * when the basis has a hash index, CACHED when the connectivity is cached)
 * and sets the operator's variant; R in [1, 64] tunes the block apply for R
 * columns (kernel depth and grid size) used by qsi_op_apply_block,
 * qsi_lanczos_block and qsi_engine_create.  d_x, d_y: scratch of D (R = 0)
 * or D x R doubles, or NULL to allocate.  *best (optional) receives the
 * chosen variant / configuration, *ms its time per apply.  No counterpart in
 * the reference.  Synchronises. */
int qsi_op_autotune(qsi_op_t op, int R, int reps, double *d_x, double *d_y,
                    int *best, double *ms);
/* Autotuner for the multi-sector block apply of qsi_lanczos_block: times
 * every configuration on a few steps of the block Lanczos recurrence over all
 * n_sectors operators (R columns, Gaussian start blocks and slot coefficients
 * allocated here; reps runs each after a warm-up) and sets the fastest on
 * every operator.  *best receives it, *ms the time per recurrence step.  No
 * counterpart in the reference.  Synchronises. */
int qsi_block_autotune(const qsi_op_t *ops, int n_sectors, int R, int reps, int *best,
                       double *ms);

/* Connectivity materialiser (the matrix-free equivalent of
 * LazyOpSum::toSparseMatrix, operator_matrix.hpp:79-105): writes the nonzero
 * off-diagonal entries H[row, col] = val in row-major order plus the diagonal
 * value of every row.  Call with d_rows == NULL to get the count in *nnz.
 * Otherwise the arrays must hold *nnz entries and d_diag D doubles.
 * Synchronises. */
int qsi_op_connectivity(qsi_op_t op, int64_t *nnz, int64_t *d_rows,
                        int64_t *d_cols, double *d_vals, double *d_diag);

/* ---- Lanczos / eigensolvers ---------------------------------------------- */

/* Settings of the reference's in-house Lanczos (lanczos::Settings,
 * components/linalg_routines/include/lanczos.hpp:9-20). */
typedef struct {
  int64_t krylov_dim;
  double abs_tol;
  double rel_tol;
  int64_t min_iterations;
  int64_t max_iterations;
  uint32_t x0_seed;
  int calc_eigenvector;
  int verify_eigenvector;
} qsi_lanczos_settings_t;

typedef struct {
  int eigval_converged;
  int eigvec_converged;
  int64_t n_iterations;
  double eigval_error;
  double eigvec_error;
} qsi_lanczos_result_t;

/* lanczos::eigval0 (components/linalg_routines/src/lanczos.cpp:184-235) with
 * the vector recurrence on the GPU in FP64.  The start vector is the
 * reference's: std::mt19937(x0_seed) + std::normal_distribution<double>,
 * normalised (lanczos.cpp:55-62,80-81).  If calc_eigenvector, d_eigvec (D
 * doubles, device) receives the Ritz vector.  Synchronises. */
int qsi_lanczos_eigval0(qsi_op_t op, const qsi_lanczos_settings_t *settings,
                        double *eigval, double *d_eigvec,
                        qsi_lanczos_result_t *result);

/* M steps of the reference's 3-term recurrence (lanczos.cpp:111-146, no
 * reorthogonalisation) from d_v0 (D doubles, normalised here) or, when d_v0
 * is NULL, from qsi_vec_randn(seed): h_alpha[j], h_beta[j] for j < M (zero
 * after a beta that underflows, which ends the recurrence; *n_steps, if
 * given, receives the steps taken).  The Ritz values are the eigenvalues of
 * the tridiagonal (qsi_tridiag_eig).  Synchronises. */
int qsi_lanczos_coeffs(qsi_op_t op, const double *d_v0, uint64_t seed, int M,
                       double *h_alpha, double *h_beta, int *n_steps);

/* qsi_lanczos_coeffs without any device allocation, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc); no counterpart in the reference.  The caller brings the
 * two Lanczos vectors: d_v (D doubles) holds the start vector on entry (it is
 * normalised in place and the run clobbers it) and d_u (D doubles) is work
 * space.  The small buffers of the run live in a static device array of
 * 8.1 MiB, one per GPU, held until the process ends: it limits M to 4096 and
 * the basis to 2^28 states (larger requests fail with QSI_E_INVALID).  Same
 * kernels and order as qsi_lanczos_coeffs, so the same h_alpha / h_beta.
 * Calls that overlap on one GPU share that array and race.  Synchronises. */
int qsi_lanczos_coeffs_noalloc(qsi_op_t op, double *d_v, double *d_u, int M,
                               double *h_alpha, double *h_beta, int *n_steps);

/* Lowest-nev eigenpairs with a thick-restart Lanczos (full
 * reorthogonalisation, ncv-dimensional search space), the GPU counterpart of
 * compute_spectrum_iterative<LazyOpSumProd> -> Spectra::SymEigsSolver
 * (include/matrix_diag_bits.hpp:51-95; SortRule::SmallestAlge).  tol is the
 * Spectra-style relative residual tolerance.  d_v0 (D doubles, device) is the
 * start vector or NULL for qsi_vec_randn(seed) (Spectra's init() takes no
 * seeded vector, matrix_diag_bits.hpp:75).  h_evals receives nev
 * ascending eigenvalues, d_evecs (D x n_evecs column-major, may be NULL) the
 * eigenvectors; *nconv the number of converged pairs.  Synchronises. */
int qsi_eigsh(qsi_op_t op, int nev, int ncv, double tol, int max_restarts,
              const double *d_v0, uint64_t seed, double *h_evals,
              double *d_evecs, int n_evecs, int *nconv, int *n_applies);

/* Block Lanczos for FTLM: for each of n_sectors operators (sectors may have
 * different basis sizes; all <= 64 spins or all > 64), R start vectors
 * d_V0[s] (device, D_s x R row-major; columns normalised here) run M steps of
 * the reference's 3-term recurrence (lanczos.cpp:111-146) side by side, with
 * no reorthogonalisation and no convergence test.  All sectors share each
 * launch (one persistent grid over the concatenated row tiles).  d_coeffs
 * (optional, per sector n_slots x R) gives per-column slot coefficients, as in
 * qsi_op_apply_block.  Output: h_alpha[(s*R + c)*M + j] = alpha_j and
 * h_beta[...] = beta_j of column c of sector s; a beta that underflows ends
 * that column (later coefficients are 0).  use_graph = 1 captures each step as
 * a HIP graph.  *elapsed_ms (optional) = device time of the M steps.
 * Synchronises.  No single counterpart in the reference (its FTLM driver,
 * include/ftlm.hpp:56-91, does one sample at a time). */
int qsi_lanczos_block(const qsi_op_t *ops, int n_sectors, int R,
                      const double *const *d_V0, const double *const *d_coeffs,
                      int M, double *h_alpha, double *h_beta, int use_graph,
                      double *elapsed_ms);

/* One FTLM sample with operator expectations, the vector loop of
 * ftlm_computer::evolve (include/ftlm.hpp:56-91) for one start vector, with
 * textbook semantics (the reference's loop does not build: its krylov_expv
 * is undefined and the sign of its step is reversed).  From d_r (D doubles,
 * device; normalised internally, not modified) M steps of the 3-term
 * recurrence give T_n (n = *n_steps <= M, fewer when a beta underflows) =
 * Y diag(e) Y^T; for every beta_b (h_betas, n_beta values)
 *   psi_b = exp(-beta_b (H - e_shift) / 2) r = V Y exp(-beta_b (e - e_shift)/2) Y^T e_1
 * is built on the device by a second, identical pass over the Lanczos vectors
 * V, and
 *   h_norm[b] = <psi_b|psi_b>,  h_expect[l * n_beta + b] = <psi_b|O_l|psi_b>
 * for the n_ops operators ops[l] (on the same basis size as H).  Summed over
 * samples these estimate Tr(exp(-beta (H - e_shift)) O) and the shifted
 * partition function; a complete orthonormal start set gives them exactly.
 * e_shift: a finite value, or NaN for the sample's lowest Ritz value; *e_used
 * (optional) receives the shift used.  h_alpha, h_beta (M doubles) receive the
 * coefficients as qsi_lanczos_coeffs gives them.  Device memory: 3 D + n_beta
 * D doubles.  Synchronises. */
int qsi_ftlm_sample(qsi_op_t H, const qsi_op_t *ops, int n_ops, const double *d_r, int M,
                    const double *h_betas, int n_beta, double e_shift, double *e_used,
                    double *h_norm, double *h_expect, double *h_alpha, double *h_beta,
                    int *n_steps);

/* Eigenvalues (ascending) of the symmetric tridiagonal (d: n, e: n-1) and,
 * if first_row != NULL, the first component of each unit eigenvector (the
 * FTLM quadrature weights are their squares).  Host only. */
int qsi_tridiag_eig(const double *d, const double *e, int n, double *evals,
                    double *first_row);

/* Ritz values of an m-step run of the 3-term recurrence without
 * reorthogonalisation (h_alpha, h_beta as qsi_lanczos_coeffs returns them;
 * h_beta[m-1] is the final residual norm), with the Cullum-Willoughby test:
 * numerically multiple eigenvalues of T_m (ghost copies of a converged one,
 * agreeing within tol) are reported once with their multiplicity, and simple
 * eigenvalues of T_m that are also eigenvalues of T_m without its first row
 * and column (spurious) are dropped.  resid[k] = |beta_m z_m| of the
 * cluster's best member (z_m = last component of its unit eigenvector of
 * T_m), a residual-norm estimate of the Ritz pair.  tol <= 0 picks
 * 10 m eps ||T||_inf.  At most max_out values, ascending; *n_out receives the
 * count written.  Host only.  The reference checks convergence of the lowest
 * Ritz value only (components/linalg_routines/src/lanczos.cpp:14-53); this
 * serves runs asked for several low eigenvalues with no room for a
 * reorthogonalised Krylov basis. */
int qsi_lanczos_ritz(const double *h_alpha, const double *h_beta, int m, double tol,
                     int max_out, double *vals, double *resid, int *mult, int *n_out);

/* ---- observables ----------------------------------------------------------
 * compute_all_expectations (src/expectation_eval.cpp:28-53):
 *   h_out[(l*n_vec + i)*n_vec + j] = psi_j . (O_l psi_i)
 * compute_cross_correlation (src/expectation_eval.cpp:75-107):
 *   h_out[((l*n_ops + m)*n_vec + i)*n_vec + j] = (O_l psi_i) . (O_m psi_j)
 * as eval_observables uses them for the ring, ring-ring and partial-volume
 * observables (src/eval_observables.cpp:136-206).  ops: n_ops operators on
 * one basis (each O_l is its own qsi_op_t, e.g. one ring term); d_psi: n_vec
 * device vectors of the basis dimension, column stride ld.  The reference
 * materialises each operator as a sparse matrix first; here every O_l psi_i
 * is one device apply and the dot products are device multi-dots in a fixed
 * order.  Synchronises. */
int qsi_expectations(const qsi_op_t *ops, int n_ops, const double *d_psi, int64_t ld, int n_vec,
                     double *h_out);
int qsi_cross_correlations(const qsi_op_t *ops, int n_ops, const double *d_psi, int64_t ld,
                           int n_vec, double *h_out);

/* ---- solve engine: continuous batching of eigensolves on one basis -------
 * Each request is one qsi_eigsh-style solve (lowest nev eigenvalues,
 * thick-restart Lanczos with full reorthogonalisation: one Gram-Schmidt pass
 * against the column's basis, a second only when the first leaves
 * projections above 1e-12 of the residual; counterpart of the reference's
 * mfsparse solve, include/matrix_diag_bits.hpp:51-95) with its own slot
 * coefficients, i.e. its own (g, B) point on a shared device-resident basis
 * and term list.  Up to max_cols requests run at once, one per column of a
 * block apply; every qsi_engine_step advances each of them by one Lanczos
 * step in shared launches, admits queued requests into free columns and
 * retires converged ones.  A request's eigenvalues do not depend on the other
 * requests in the engine.  The start vector is a Gaussian generated on the
 * device from the request's seed.  `op` must outlive the engine.  No
 * counterpart in the reference (it solves one Hamiltonian per process). */
typedef struct qsi_engine_s *qsi_engine_t;
int qsi_engine_create(qsi_op_t op, int max_cols, int nev, int ncv, double tol,
                      int max_restarts, qsi_engine_t *out);
int qsi_engine_destroy(qsi_engine_t e);
/* device bytes qsi_engine_create allocates for these sizes */
int64_t qsi_engine_footprint(int64_t dim, int max_cols, int ncv);
/* queue one request; h_coeffs: n_slots values (NULL if the op has no slots) */
int qsi_engine_submit(qsi_engine_t e, const double *h_coeffs, uint64_t seed,
                      int64_t *id);
int qsi_engine_active(qsi_engine_t e);  /* requests holding a column */
int qsi_engine_queued(qsi_engine_t e);  /* requests waiting for one  */
/* keeps up to two steps enqueued (admitting requests), then finishes the
 * oldest: one Lanczos step for every active request; synchronises with it */
int qsi_engine_step(qsi_engine_t e, int *n_active);
/* The halves of qsi_engine_step, so that several engines (each on its own
 * stream) overlap and the device never waits for the host: step_begin admits
 * and enqueues a step (*launched = 0 when no request can step; at most two
 * steps in flight, the second continuing from the first, a request whose
 * Krylov basis the first fills sitting out until its restart), step_end waits
 * for the oldest step and does its host work (projected matrices, Ritz
 * values, restarts, finished requests). */
int qsi_engine_step_begin(qsi_engine_t e, int *launched);
int qsi_engine_step_end(qsi_engine_t e, int *n_active);
/* 1 when step_end would not wait (no step enqueued, or the oldest has
 * completed), 0 while it runs; does not block */
int qsi_engine_step_ready(qsi_engine_t e);
/* steps enqueued and not yet finished by step_end (0, 1 or 2) */
int qsi_engine_inflight(qsi_engine_t e);
/* pop one finished request: returns 1 and fills id, h_evals (nev), nconv and
 * steps (Lanczos steps taken); returns 0 when none has finished */
int qsi_engine_poll(qsi_engine_t e, int64_t *id, double *h_evals, int *nconv,
                    int *steps);

/* ---- vector kernels (streaming FP64, deterministic reductions) ----------- */
int qsi_vec_dot(const double *d_a, const double *d_b, int64_t n, double *h_out,
                qsi_stream_t stream); /* synchronises */
int qsi_vec_axpy(double alpha, const double *d_x, double *d_y, int64_t n,
                 qsi_stream_t stream);
int qsi_vec_scal(double alpha, double *d_x, int64_t n, qsi_stream_t stream);
/* d_x[i] = standard Gaussian draw determined by (seed, i) alone (counter-based
 * generator); the start vectors of qsi_engine requests.  The reference's
 * start vectors come from std::mt19937 on the host (lanczos.cpp:55-62);
 * qsi_eigsh with d_v0 = NULL keeps those. */
int qsi_vec_randn(double *d_x, int64_t n, uint64_t seed, qsi_stream_t stream);

/* ---- basis enumeration (host) --------------------------------------------
 * Ice-rule depth-first enumeration (components/build_basis/src/pyro_tree.cpp:
 * 20-240; gen_spinon_basis.cpp:16-156): spins are visited in the order `order`
 * (n_spins entries; the reference's default is its greedy ordering), each
 * tetrahedron (4 member spins) must hold 2 up spins, and up to
 * n_spinon_pairs forced violations are allowed as in fork_state_impl.  The
 * sorted states are returned in a malloc'd array (*out_lohi, free with
 * qsi_free); *out_dim receives D. */
int qsi_enumerate_basis(int n_spins, int n_tetra, const int32_t *tetra_members,
                        const int32_t *order, int n_spinon_pairs,
                        int n_threads, uint64_t **out_lohi, int64_t *out_dim);
void qsi_free(void *p);

/* The same enumeration on the device, for bases too large for the host
 * (96 spins: ~2e9 states): a breadth-first frontier of the search tree is
 * expanded on the host, every subtree is walked on the device (a counting
 * pass, then a writing pass), and the states are radix-sorted there.  The
 * result is a basis handle as from qsi_basis_create (bucket_bits = 0: auto);
 * *out_dim (optional) receives D.  The states can be read back with
 * qsi_basis_download.  Synchronises. */
int qsi_enumerate_basis_device(int n_spins, int n_tetra,
                               const int32_t *tetra_members,
                               const int32_t *order, int n_spinon_pairs,
                               int bucket_bits, qsi_basis_t *out,
                               int64_t *out_dim);
/* Copies the basis states to h_lohi (D rows of [lo, hi]); synchronises. */
int qsi_basis_download(qsi_basis_t b, uint64_t *h_lohi);

#ifdef __cplusplus
}
#endif
#endif /* QSI_GPU_H */
