// Block Lanczos recurrences for many start vectors on several operators in
// one launch per kernel — the batched FTLM path.
//
// R start vectors per sector (operator + basis) run the reference's 3-term
// recurrence (components/linalg_routines/src/lanczos.cpp:111-146, without the
// convergence checks) side by side, M steps, no reorthogonalisation, which is
// what FTLM needs: the (alpha, beta) coefficients of every sample.  Sectors of
// different basis size share each launch: their rows are cut into tiles of 64
// and one persistent grid of wavefronts walks the concatenated tile list, so
// each basis state's flip tests and rank lookups serve all R columns.
//
// Per step (three rotating D x R blocks holding v_{j-1}, v_j unnormalised,
// A = w_{j-1}, B = w_j with v = scale * column, and C):
//   k_blk_apply      C = H B, per-wave partials of <B_c, C_c>
//   k_blk_finalize   alpha[s][c] = scale_B^2 <B_c, C_c>
//   k_blk_update     C = scale_B (C - alpha B) - beta scale_A A, partials of |C_c|^2
//   k_blk_finalize   beta[s][c] = |C_c|, scale_C = 1/beta (0 when beta underflows)
//   k_blk_step       history index += 1;  then (A, B, C) <- (B, C, A)
// No pass rewrites a block to normalise it.  The five launches of a step are
// captured once per rotation phase as a HIP graph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <vector>

#include <hipcub/hipcub.hpp>

#include "qsi_blockk.hpp"
#include "qsi_hostla.hpp"

namespace qsi {
namespace {

__device__ __forceinline__ int find_sector(const BlkArgs& a, int64_t tile) {
  int s = 0;
  for (int k = 1; k < a.n_sec; ++k)
    if (tile >= a.sec[k].tile_begin) s = k;
  return s;
}

// Sum of v over the lanes that share this lane's column (lanes c, c+G, ...).
__device__ __forceinline__ double sum_over_groups(double v, int G) {
  for (int o = G; o < 64; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

constexpr int kBlkCap = 512;         // LDS flip entries per wavefront and round
constexpr int kBlkSuper = 256;       // off-diagonal terms per round (4 masks)
constexpr int kBlkBatch = 4;         // lookups in flight per lane (phase A)
constexpr int kBlkPipe = 4;          // X loads in flight per lane (phase B)
constexpr uint32_t kNoRank = 0xFFFFFFFFu;

__device__ __forceinline__ void blk_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ranks of up to B keys with all bisections interleaved; kNoRank for
// keys absent from the basis (or inactive slots).  Every stage issues the
// loads of all unfinished lookups before it uses any; an empty asm statement
// that reads each loaded value after the last load keeps the compiler from
// moving a use next to its load (which would make each load wait for the
// one before).
template <int NW, int B>
__device__ __forceinline__ void lookup_batch(const BasisView& b, const State<NW>* key,
                                             const bool* act, uint32_t* rank) {
  uint32_t lo[B], n[B];
  uint2 p[B];
#pragma unroll
  for (int g = 0; g < B; ++g) {
    p[g] = make_uint2(0u, 0u);
    if (act[g]) __builtin_memcpy(&p[g], b.table + bucket_key<NW>(b, key[g]), sizeof(uint2));
  }
#pragma unroll
  for (int g = 0; g < B; ++g) asm volatile("" : "+v"(p[g].x), "+v"(p[g].y));
  bool more = false;
#pragma unroll
  for (int g = 0; g < B; ++g) {
    lo[g] = p[g].x;
    n[g] = act[g] ? p[g].y - p[g].x : 0u;
    more |= n[g] > 1;
  }
  while (more) {
    State<NW> m[B];
#pragma unroll
    for (int g = 0; g < B; ++g) {
      m[g] = State<NW>{};
      if (n[g] > 1) m[g] = load_state_cached<NW>(b.states, lo[g] + (n[g] >> 1));
    }
#pragma unroll
    for (int g = 0; g < B; ++g)
#pragma unroll
      for (int w = 0; w < NW; ++w) asm volatile("" : "+v"(m[g].w[w]));
    more = false;
#pragma unroll
    for (int g = 0; g < B; ++g) {
      const uint32_t half = n[g] >> 1;  // 0 once the lookup has finished
      lo[g] = st_le<NW>(m[g], key[g]) ? lo[g] + half : lo[g];
      n[g] -= half;
      more |= n[g] > 1;
    }
  }
  State<NW> f[B];
#pragma unroll
  for (int g = 0; g < B; ++g) {
    f[g] = State<NW>{};
    if (n[g] == 1) f[g] = load_state_cached<NW>(b.states, lo[g]);
  }
#pragma unroll
  for (int g = 0; g < B; ++g)
#pragma unroll
    for (int w = 0; w < NW; ++w) asm volatile("" : "+v"(f[g].w[w]));
#pragma unroll
  for (int g = 0; g < B; ++g) rank[g] = n[g] == 1 && st_eq<NW>(f[g], key[g]) ? lo[g] : kNoRank;
}

// Y[:, c] = H_c X[:, c] (or Y += H_c X) for R columns of row-major blocks,
// and (partials != null) per-wave, per-sector partials of <X[:, c], Y[:, c]>.
//
// Each wavefront takes a 64-row tile in two phases per round:
//   A  one lane per row finds the row's connected terms, looks up the
//      pre-image ranks BATCH at a time and writes (rank, term, sign)
//      entries to LDS in term order;
//   B  G lanes per row (one per column) walk the row's entries and read
//      X[J, 0:R] as one coalesced segment per connected state.
// A round covers up to kBlkSuper terms and as many rows as fit kBlkCap LDS
// entries; partial sums of rows cut by a round boundary go through Y.
// Contributions are added in term order, then the diagonal terms in order, so
// column c's result depends only on X[:, c] and its coefficients.
template <int NW, int BATCH>
__global__ __launch_bounds__(kBlkThreads) void k_blk_apply(
    BlkArgs a, const double* __restrict__ X, double* __restrict__ Y,
    double* __restrict__ partials) {
  // entry: rank (bits 0-31, kNoRank if absent), term - round base (32-39),
  // coefficient class (40-47), total sign negative (bit 48)
  __shared__ uint64_t lds_ent[kBlkWaves][kBlkCap];
  const int lane = threadIdx.x & 63;
  // wave-uniform values kept in scalar registers
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint64_t* ent = lds_ent[wv];
  const int G = a.G, R = a.R;
  const int lg = __builtin_ctz(G);
  const int grp = lane >> lg, col = lane & (G - 1);
  const bool col_ok = col < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + wv;
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  const int rpp = 64 >> lg;  // rows per phase-B pass
  // tiles are visited in increasing order, so sectors come in order: one
  // running dot product, flushed when the sector changes
  double dot = 0.0;
  int flushed = 0;  // sectors [0, flushed) have their partial written
  auto flush_to = [&](int upto, int cur) {
    for (; flushed < upto; ++flushed) {
      const double v = sum_over_groups(flushed == cur ? dot : 0.0, G);
      if (partials && grp == 0 && col_ok) partials[(flushed * n_waves + wave) * R + col] = v;
    }
  };
  int cur = -1;
  double cc[kMaxCoefClasses];
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = __builtin_amdgcn_readfirstlane(find_sector(a, t));
    const BlkSector& S = a.sec[s];
    if (s != cur) {
      flush_to(s, cur);
      cur = s;
      dot = 0.0;
      // this lane's column coefficient of every class
#pragma unroll
      for (int k = 0; k < kMaxCoefClasses; ++k) {
        cc[k] = 0.0;
        if (k < S.n_cls && col_ok) {
          const CoefClass& K = S.cls[k];
          cc[k] = (S.coeffs && K.slot >= 0) ? S.coeffs[K.slot * R + col] : K.c;
        }
      }
    }
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int n_rows = static_cast<int>(min<int64_t>(kTileRows, S.b.dim - r_begin));
    const bool row_ok = lane < n_rows;
    const State<NW> st = load_state_cached<NW>(S.b.states, r_begin + (row_ok ? lane : 0));
    const int n_sc = max(1, (S.n_off + kBlkSuper - 1) / kBlkSuper);
    for (int sc = 0; sc < n_sc; ++sc) {
      const int tb = sc * kBlkSuper;
      const int tn = min(kBlkSuper, S.n_off - tb);
      // connected terms of this lane's row
      uint64_t conn[4];
      int cnt = 0;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        conn[q] = 0;
        const int qn = row_ok ? min(64, tn - q * 64) : 0;
        for (int u = 0; u < qn; ++u)
          conn[q] |= static_cast<uint64_t>(term_ok<NW>(S.off[tb + q * 64 + u], st)) << u;
        cnt += __popcll(conn[q]);
      }
      int incl = cnt;
      for (int o = 1; o < 64; o <<= 1) {
        const int v = __shfl_up(incl, o, 64);
        if (lane >= o) incl += v;
      }
      const int excl = incl - cnt;
      for (int ra = 0; ra < 64;) {
        const int base = __shfl(excl, ra, 64);
        const uint64_t fits = __ballot(lane >= ra && incl - base <= kBlkCap);
        const int rb = ra + __popcll(fits);  // >= ra + 1: cnt <= kBlkSuper <= kBlkCap
        // phase A: this lane's connected terms, in term order
        if (lane >= ra && lane < rb && cnt > 0) {
          int pos = excl - base;
          for (int q = 0; q < 4; ++q) {
            uint64_t m = conn[q];
            while (m) {
              State<NW> key[BATCH];
              bool act[BATCH];
              int tl[BATCH];
              uint32_t rk[BATCH];
#pragma unroll
              for (int g = 0; g < BATCH; ++g) {
                act[g] = m != 0;
                tl[g] = q * 64 + (act[g] ? __builtin_ctzll(m) : 0);
                m &= m - 1;
                key[g] = xor_state<NW>(st, S.off[tb + tl[g]].X);
              }
              lookup_batch<NW, BATCH>(S.b, key, act, rk);
#pragma unroll
              for (int g = 0; g < BATCH; ++g) {
                if (act[g]) {
                  const OffTerm& T = S.off[tb + tl[g]];
                  const bool ng = (T.sgn < 0.0) != static_cast<bool>(parity_notz<NW>(st, T.Z));
                  ent[pos++] = static_cast<uint64_t>(rk[g]) |
                               (static_cast<uint64_t>(tl[g]) << 32) |
                               (static_cast<uint64_t>(T.cls & 0xFF) << 40) |
                               (static_cast<uint64_t>(ng) << 48);
                }
              }
            }
          }
        }
        blk_wave_sync();
        // phase B: G lanes per row
        const int n_pass = (rb - ra + rpp - 1) / rpp;
        for (int p = 0; p < n_pass; ++p) {
          const int r = ra + p * rpp + grp;
          const int rs = min(r, 63);
          const int e0 = __shfl(excl, rs, 64) - base;
          const int e1 = __shfl(incl, rs, 64) - base;
          if (r < rb && r < n_rows && col_ok) {
            const int64_t row = S.row_off + r_begin + r;
            const int64_t gi = a.y_colmajor ? col * a.ldy + row : row * a.ldy + col;
            double acc = (sc > 0 || a.accumulate) ? Y[gi] : 0.0;
            // kBlkPipe entries at a time: their X and coefficient loads are
            // issued together, the sum is taken in entry order
            for (int e = e0; e < e1; e += kBlkPipe) {
              uint64_t en[kBlkPipe];
              bool ok[kBlkPipe];
              double xj[kBlkPipe], c[kBlkPipe];
#pragma unroll
              for (int u = 0; u < kBlkPipe; ++u) {
                en[u] = e + u < e1 ? ent[e + u] : static_cast<uint64_t>(kNoRank);
                ok[u] = static_cast<uint32_t>(en[u]) != kNoRank;
              }
#pragma unroll
              for (int u = 0; u < kBlkPipe; ++u) {
                const int64_t J = ok[u] ? static_cast<int64_t>(static_cast<uint32_t>(en[u])) : 0;
                xj[u] = ok[u] ? X[(S.row_off + J) * a.ldx + col] : 0.0;
                if (S.n_cls > 0) {
                  const int k = static_cast<int>((en[u] >> 40) & 0xFF);
                  double ck = cc[0];
#pragma unroll
                  for (int q = 1; q < kMaxCoefClasses; ++q)
                    if (k == q) ck = cc[q];
                  c[u] = ck;
                } else {
                  const OffTerm& T = S.off[tb + static_cast<int>((en[u] >> 32) & 0xFF)];
                  c[u] = (S.coeffs && T.slot >= 0) ? S.coeffs[T.slot * R + col] : T.c;
                }
              }
#pragma unroll
              for (int u = 0; u < kBlkPipe; ++u) {
                if (ok[u] && fabs(xj[u]) > kApplyTol) {
                  const double cu = ((en[u] >> 48) & 1) ? -c[u] : c[u];
                  acc += cu * xj[u];
                }
              }
            }
            if (sc == n_sc - 1) {
              const State<NW> sr = load_state_cached<NW>(S.b.states, r_begin + r);
              const double xi = X[row * a.ldx + col];
              for (int k = 0; k < S.n_diag; ++k) {
                const DiagTerm& d = S.dg[k];
                double f;
                if (d.kind == 0) {
                  f = diag_ok<NW>(d, sr) ? d.sgn * (1 - 2 * parity_notz<NW>(sr, d.Z)) : 0.0;
                } else {
                  int nz = __popcll(d.Z[0]);
                  if constexpr (NW == 2) nz += __popcll(d.Z[1]);
                  f = d.sgn * static_cast<double>(2 * popc_and<NW>(sr, d.Z) - nz);
                }
                const double c = (S.coeffs && d.slot >= 0) ? S.coeffs[d.slot * R + col] : d.c;
                acc += c * xi * f;
              }
              dot += xi * acc;
            }
            Y[gi] = acc;
          }
        }
        blk_wave_sync();
        ra = rb;
      }
    }
  }
  flush_to(a.n_sec, cur);
}

// ---- cached connectivity ---------------------------------------------------
// Rows' connected states are fixed by the basis and the term list, so an
// operator applied many times (a solve engine, an FTLM run) can keep them:
// per row, in term order, the rank J (kNoRank if absent) and one byte
// (coefficient class | negative total sign << 7).  A block apply is then a
// pure gather: no flip tests, no searches.

// counts[i] = number of connected off-diagonal terms of row i, rounded up to
// a multiple of `pad` (1 or 4: rows padded to whole groups of four entries).
template <int NW>
__global__ __launch_bounds__(256) void k_cc_count(BasisView b, const OffTerm* __restrict__ off,
                                                  int n_off, int pad,
                                                  uint64_t* __restrict__ counts) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.dim) return;
  const State<NW> st = load_state_cached<NW>(b.states, i);
  uint64_t n = 0;
  for (int t = 0; t < n_off; ++t) n += term_ok<NW>(off[t], st) ? 1u : 0u;
  counts[i] = (n + pad - 1) / pad * pad;
}

// Entries of row i at rowptr[i]...: ranks looked up kBlkBatch at a time.
template <int NW>
__global__ __launch_bounds__(256) void k_cc_fill(BasisView b, const OffTerm* __restrict__ off,
                                                 int n_off, const uint64_t* __restrict__ rowptr,
                                                 uint32_t* __restrict__ cj,
                                                 uint8_t* __restrict__ cm) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= b.dim) return;
  const State<NW> st = load_state_cached<NW>(b.states, i);
  uint64_t pos = rowptr[i];
  for (int t0 = 0; t0 < n_off; t0 += 64) {
    const int tn = min(64, n_off - t0);
    uint64_t m = 0;
    for (int u = 0; u < tn; ++u)
      m |= static_cast<uint64_t>(term_ok<NW>(off[t0 + u], st)) << u;
    while (m) {
      State<NW> key[kBlkBatch];
      bool act[kBlkBatch];
      int tt[kBlkBatch];
      uint32_t rk[kBlkBatch];
#pragma unroll
      for (int g = 0; g < kBlkBatch; ++g) {
        act[g] = m != 0;
        tt[g] = t0 + (act[g] ? __builtin_ctzll(m) : 0);
        m &= m - 1;
        key[g] = xor_state<NW>(st, off[tt[g]].X);
      }
      lookup_batch<NW, kBlkBatch>(b, key, act, rk);
#pragma unroll
      for (int g = 0; g < kBlkBatch; ++g) {
        if (act[g]) {
          const OffTerm& T = off[tt[g]];
          const bool ng = (T.sgn < 0.0) != static_cast<bool>(parity_notz<NW>(st, T.Z));
          cj[pos] = rk[g];
          cm[pos] = static_cast<uint8_t>((T.cls & 0x7F) | (ng ? 0x80 : 0));
          ++pos;
        }
      }
    }
  }
  // padding entries (padded layout): absent, so every apply skips them
  for (const uint64_t end = rowptr[i + 1]; pos < end; ++pos) {
    cj[pos] = kNoRank;
    cm[pos] = 0;
  }
}

// PIPE entries of a row's cached connectivity from entry e (a multiple of 4):
// columns four at a time with one 16-byte load, class bytes with one 4-byte
// load (the arrays are allocated in whole groups); entries outside the row
// [e0, e1) read as absent.
template <int PIPE>
__device__ __forceinline__ void load_entries(const BlkSector& S, uint64_t e, uint64_t e0,
                                             uint64_t e1, uint32_t (&J)[PIPE],
                                             uint8_t (&mt)[PIPE]) {
  static_assert(PIPE % 4 == 0, "groups of four");
#pragma unroll
  for (int u = 0; u < PIPE; u += 4) {
    const uint64_t g = e + u;
    const bool in = g < e1;
    const uint4 j4 = in ? reinterpret_cast<const uint4*>(S.cj)[g >> 2]
                       : make_uint4(kNoRank, kNoRank, kNoRank, kNoRank);
    const uint32_t m4 = in ? reinterpret_cast<const uint32_t*>(S.cm)[g >> 2] : 0u;
    const uint32_t jv[4] = {j4.x, j4.y, j4.z, j4.w};
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const bool mine = g + k >= e0 && g + k < e1;
      J[u + k] = mine ? jv[k] : kNoRank;
      mt[u + k] = static_cast<uint8_t>(m4 >> (8 * k));
    }
  }
}

// The block apply over cached connectivity: same arithmetic, order and
// results as k_blk_apply.  G lanes per row (one per column), PIPE entries'
// gathers in flight.  G may be any width up to 64 when no partials are asked
// for (a wavefront then serves floor(64 / G) rows per pass); with partials it
// must be a power of two (the per-column sums cross row groups by butterfly).
template <int NW, int PIPE>
__global__ __launch_bounds__(kBlkThreads) void k_blk_capply(
    BlkArgs a, const double* __restrict__ X, double* __restrict__ Y,
    double* __restrict__ partials) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int G = a.G, R = a.R;
  const int grp = lane / G, col = lane % G;
  const int rpp = 64 / G;
  const bool col_ok = col < R && grp < rpp;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + wv;
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  double dot = 0.0;
  int flushed = 0;
  auto flush_to = [&](int upto, int cur) {
    for (; flushed < upto; ++flushed) {
      if (!partials) continue;
      const double v = sum_over_groups(flushed == cur ? dot : 0.0, G);
      if (grp == 0 && col_ok) partials[(flushed * n_waves + wave) * R + col] = v;
    }
  };
  int cur = -1;
  double cc[kMaxCoefClasses];
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = __builtin_amdgcn_readfirstlane(find_sector(a, t));
    const BlkSector& S = a.sec[s];
    if (s != cur) {
      flush_to(s, cur);
      cur = s;
      dot = 0.0;
#pragma unroll
      for (int k = 0; k < kMaxCoefClasses; ++k) {
        cc[k] = 0.0;
        if (k < S.n_cls && col_ok) {
          const CoefClass& K = S.cls[k];
          cc[k] = (S.coeffs && K.slot >= 0) ? S.coeffs[K.slot * R + col] : K.c;
        }
      }
    }
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int n_rows = static_cast<int>(min<int64_t>(kTileRows, S.b.dim - r_begin));
    const int n_pass = (n_rows + rpp - 1) / rpp;
    for (int p = 0; p < n_pass; ++p) {
      const int r = p * rpp + grp;
      if (r >= n_rows || !col_ok) continue;
      const int64_t i = r_begin + r;
      const int64_t row = S.row_off + i;
      const int64_t gi = a.y_colmajor ? col * a.ldy + row : row * a.ldy + col;
      double acc = a.accumulate ? Y[gi] : 0.0;
      const uint64_t e0 = S.rowptr[i], e1 = S.rowptr[i + 1];
      for (uint64_t e = e0 & ~uint64_t{3}; e < e1; e += PIPE) {
        uint32_t J[PIPE];
        uint8_t mt[PIPE];
        double xj[PIPE];
        load_entries<PIPE>(S, e, e0, e1, J, mt);
#pragma unroll
        for (int u = 0; u < PIPE; ++u)
          xj[u] = J[u] != kNoRank ? X[(S.row_off + static_cast<int64_t>(J[u])) * a.ldx + col] : 0.0;
#pragma unroll
        for (int u = 0; u < PIPE; ++u) {
          if (J[u] != kNoRank && fabs(xj[u]) > kApplyTol) {
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
      const State<NW> sr = load_state_cached<NW>(S.b.states, i);
      const double xi = X[row * a.ldx + col];
      for (int k = 0; k < S.n_diag; ++k) {
        const DiagTerm& d = S.dg[k];
        double f;
        if (d.kind == 0) {
          f = diag_ok<NW>(d, sr) ? d.sgn * (1 - 2 * parity_notz<NW>(sr, d.Z)) : 0.0;
        } else {
          int nz = __popcll(d.Z[0]);
          if constexpr (NW == 2) nz += __popcll(d.Z[1]);
          f = d.sgn * static_cast<double>(2 * popc_and<NW>(sr, d.Z) - nz);
        }
        const double c = (S.coeffs && d.slot >= 0) ? S.coeffs[d.slot * R + col] : d.c;
        acc += c * xi * f;
      }
      dot += xi * acc;
      Y[gi] = acc;
    }
  }
  flush_to(a.n_sec, cur);
}

// Column pairs: each lane of a row group serves columns c0 = 2q and c0 + 1
// with one 16-byte gather per entry, so a wavefront covers twice the rows per
// pass of k_blk_capply for the same number of memory instructions (G here is
// the number of lanes per row, ceil(R / 2), a power of two when partials are
// asked for).  Needs 16-byte aligned rows of X that hold both columns of
// every lane: even ldx >= R rounded up to even.  Per column the arithmetic
// and its order are those of k_blk_capply, so Y is identical; partials sum
// the row groups in another order.
template <int NW, int PIPE>
__global__ __launch_bounds__(kBlkThreads) void k_blk_capply2(
    BlkArgs a, const double* __restrict__ X, double* __restrict__ Y,
    double* __restrict__ partials) {
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int G = a.G, R = a.R;
  const int grp = lane / G, c0 = 2 * (lane % G);
  const int rpp = 64 / G;
  const bool ok0 = c0 < R && grp < rpp;
  const bool ok1 = ok0 && c0 + 1 < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + wv;
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  double dot0 = 0.0, dot1 = 0.0;
  int flushed = 0;
  auto flush_to = [&](int upto, int cur) {
    for (; flushed < upto; ++flushed) {
      if (!partials) continue;
      const double v0 = sum_over_groups(flushed == cur ? dot0 : 0.0, G);
      const double v1 = sum_over_groups(flushed == cur ? dot1 : 0.0, G);
      if (grp == 0 && ok0) partials[(flushed * n_waves + wave) * R + c0] = v0;
      if (grp == 0 && ok1) partials[(flushed * n_waves + wave) * R + c0 + 1] = v1;
    }
  };
  int cur = -1;
  double cc0[kMaxCoefClasses], cc1[kMaxCoefClasses];
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = __builtin_amdgcn_readfirstlane(find_sector(a, t));
    const BlkSector& S = a.sec[s];
    if (s != cur) {
      flush_to(s, cur);
      cur = s;
      dot0 = dot1 = 0.0;
#pragma unroll
      for (int k = 0; k < kMaxCoefClasses; ++k) {
        cc0[k] = cc1[k] = 0.0;
        if (k < S.n_cls) {
          const CoefClass& K = S.cls[k];
          const bool slot = S.coeffs && K.slot >= 0;
          if (ok0) cc0[k] = slot ? S.coeffs[K.slot * R + c0] : K.c;
          if (ok1) cc1[k] = slot ? S.coeffs[K.slot * R + c0 + 1] : K.c;
        }
      }
    }
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int n_rows = static_cast<int>(min<int64_t>(kTileRows, S.b.dim - r_begin));
    const int n_pass = (n_rows + rpp - 1) / rpp;
    for (int p = 0; p < n_pass; ++p) {
      const int r = p * rpp + grp;
      if (r >= n_rows || !ok0) continue;
      const int64_t i = r_begin + r;
      const int64_t row = S.row_off + i;
      const int64_t g0 = a.y_colmajor ? c0 * a.ldy + row : row * a.ldy + c0;
      const int64_t g1 = a.y_colmajor ? g0 + a.ldy : g0 + 1;
      double acc0 = a.accumulate ? Y[g0] : 0.0;
      double acc1 = a.accumulate && ok1 ? Y[g1] : 0.0;
      const uint64_t e0 = S.rowptr[i], e1 = S.rowptr[i + 1];
      for (uint64_t e = e0 & ~uint64_t{3}; e < e1; e += PIPE) {
        uint32_t J[PIPE];
        uint8_t mt[PIPE];
        double2 xj[PIPE];
        load_entries<PIPE>(S, e, e0, e1, J, mt);
#pragma unroll
        for (int u = 0; u < PIPE; ++u)
          xj[u] = J[u] != kNoRank
                      ? *reinterpret_cast<const double2*>(
                            X + (S.row_off + static_cast<int64_t>(J[u])) * a.ldx + c0)
                      : make_double2(0.0, 0.0);
#pragma unroll
        for (int u = 0; u < PIPE; ++u) {
          if (J[u] != kNoRank) {
            const int k = mt[u] & 0x7F;
            double ca = cc0[0], cb = cc1[0];
#pragma unroll
            for (int q = 1; q < kMaxCoefClasses; ++q)
              if (k == q) {
                ca = cc0[q];
                cb = cc1[q];
              }
            if (mt[u] & 0x80) {
              ca = -ca;
              cb = -cb;
            }
            if (fabs(xj[u].x) > kApplyTol) acc0 += ca * xj[u].x;
            if (fabs(xj[u].y) > kApplyTol) acc1 += cb * xj[u].y;
          }
        }
      }
      const State<NW> sr = load_state_cached<NW>(S.b.states, i);
      const double2 xi = *reinterpret_cast<const double2*>(X + row * a.ldx + c0);
      for (int k = 0; k < S.n_diag; ++k) {
        const DiagTerm& d = S.dg[k];
        double f;
        if (d.kind == 0) {
          f = diag_ok<NW>(d, sr) ? d.sgn * (1 - 2 * parity_notz<NW>(sr, d.Z)) : 0.0;
        } else {
          int nz = __popcll(d.Z[0]);
          if constexpr (NW == 2) nz += __popcll(d.Z[1]);
          f = d.sgn * static_cast<double>(2 * popc_and<NW>(sr, d.Z) - nz);
        }
        const bool slot = S.coeffs && d.slot >= 0;
        const double ca = slot ? S.coeffs[d.slot * R + c0] : d.c;
        const double cb = ok1 ? (slot ? S.coeffs[d.slot * R + c0 + 1] : d.c) : 0.0;
        acc0 += ca * xi.x * f;
        acc1 += cb * xi.y * f;
      }
      dot0 += xi.x * acc0;
      dot1 += xi.y * acc1;
      Y[g0] = acc0;
      if (ok1) Y[g1] = acc1;
    }
  }
  flush_to(a.n_sec, cur);
}

// C = fb C - alpha fb B - beta fa A per column (fb, fa: the scales sB, sA of
// the unnormalised blocks B and A, 1 when null; beta, A skipped when A is
// null); |C_c|^2 partials.
__global__ __launch_bounds__(kBlkThreads) void k_blk_update(
    BlkArgs a, const double* __restrict__ A, const double* B,
    double* C, const double* __restrict__ alpha,
    const double* __restrict__ beta, double* __restrict__ partials,
    const double* __restrict__ sA, const double* __restrict__ sB) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, R = a.R;
  const int grp = lane / G, col = lane % G;
  const bool col_ok = col < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  const int rows_per_pass = 64 / G;
  double nrm[kMaxSectors];
#pragma unroll
  for (int k = 0; k < kMaxSectors; ++k) nrm[k] = 0.0;
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = find_sector(a, t);
    const BlkSector& S = a.sec[s];
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int64_t r_end = min<int64_t>(r_begin + kTileRows, S.b.dim);
    const double al = col_ok && alpha ? alpha[s * R + col] : 0.0;
    const double be = col_ok && A ? beta[s * R + col] : 0.0;
    // the blocks hold unnormalised vectors: v_B = fb B, v_A = fa A, and C
    // holds H B, so H v_B = fb C
    const double fb = col_ok && sB ? sB[s * R + col] : 1.0;
    const double fa = col_ok && sA ? sA[s * R + col] : 1.0;
    const double alb = al * fb, bea = be * fa;
    double part = 0.0;
    // eight rows per round: their loads are issued together; the norm is
    // still accumulated row by row in order
    constexpr int kU = 8;
    for (int64_t i0 = r_begin + grp; i0 < r_end; i0 += kU * rows_per_pass) {
      if (!col_ok) continue;
      double cv[kU], bv[kU], av[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t i = i0 + u * rows_per_pass;
        const int64_t gi = (S.row_off + i) * R + col;
        const bool in = i < r_end;
        cv[u] = in ? C[gi] : 0.0;
        bv[u] = in ? B[gi] : 0.0;
        av[u] = in && A ? A[gi] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t i = i0 + u * rows_per_pass;
        if (i >= r_end) continue;
        const int64_t gi = (S.row_off + i) * R + col;
        double w = fb * cv[u] - alb * bv[u];
        if (A) w -= bea * av[u];
        if (fb == 0.0) w = 0.0;  // a column whose beta underflowed stays zero
        C[gi] = w;
        part += w * w;
      }
    }
    for (int k = 0; k < kMaxSectors; ++k)
      if (k == s) nrm[k] += part;
  }
  for (int k = 0; k < a.n_sec; ++k) {
    const double v = sum_over_groups(nrm[k], G);
    if (grp == 0 && col_ok) partials[(k * n_waves + wave) * R + col] = v;
  }
}

// One workgroup per (sector, column): fixed-order sum of the wave partials.
// mode 0 (alpha): out[s*R+c] = sum, times sq[s*R+c]^2 when sq is given.  mode 1 (beta): out[s*R+c] = sqrt(sum),
// out2[s*R+c] = 1/beta or 0.  hist (optional) receives the value at
// hist[(s*R + c)*M + step[0]].
__global__ __launch_bounds__(256) void k_blk_finalize(
    const double* __restrict__ partials, int64_t n_waves, int R, int mode,
    double* __restrict__ out, double* __restrict__ out2, double* __restrict__ hist,
    const int* __restrict__ step, int M, const double* __restrict__ sq) {
  const int sc = blockIdx.x;  // s * R + c
  const int s = sc / R, c = sc % R;
  double v = 0.0;
  for (int64_t w = threadIdx.x; w < n_waves; w += 256)
    v += partials[(s * n_waves + w) * R + c];
  v = block_sum<256>(v);
  if (threadIdx.x == 0) {
    double r = v;
    if (mode == 0 && sq) r = r * sq[sc] * sq[sc];
    if (mode == 1) {
      r = sqrt(v);
      out2[sc] = r > 1e-300 ? 1.0 / r : 0.0;
    }
    out[sc] = r;
    if (hist) hist[static_cast<int64_t>(sc) * M + step[0]] = r;
  }
}

// Advances the step counter the coefficient history is indexed by.
__global__ void k_blk_step(int* __restrict__ step) {
  if (blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1;
}

// C *= 1/beta per (sector, column); optionally advances the step counter.
__global__ __launch_bounds__(kBlkThreads) void k_blk_scale(
    BlkArgs a, double* __restrict__ C, const double* __restrict__ invb,
    int* __restrict__ step) {
  const int lane = threadIdx.x & 63;
  const int G = a.G, R = a.R;
  const int grp = lane / G, col = lane % G;
  const bool col_ok = col < R;
  const int64_t wave = static_cast<int64_t>(blockIdx.x) * kBlkWaves + (threadIdx.x >> 6);
  const int64_t n_waves = static_cast<int64_t>(gridDim.x) * kBlkWaves;
  const int rows_per_pass = 64 / G;
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = find_sector(a, t);
    const BlkSector& S = a.sec[s];
    const int64_t r_begin = (t - S.tile_begin) * kTileRows;
    const int64_t r_end = min<int64_t>(r_begin + kTileRows, S.b.dim);
    if (!col_ok) continue;
    const double f = invb[s * R + col];
    constexpr int kU = 8;
    for (int64_t i0 = r_begin + grp; i0 < r_end; i0 += kU * rows_per_pass) {
      double cv[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t i = i0 + u * rows_per_pass;
        cv[u] = i < r_end ? C[(S.row_off + i) * R + col] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int64_t i = i0 + u * rows_per_pass;
        if (i < r_end) C[(S.row_off + i) * R + col] = cv[u] * f;
      }
    }
  }
  if (step && blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1;
}



struct DevMem {
  std::vector<void*> ptrs;
  ~DevMem() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  hipError_t alloc(T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, std::max<size_t>(n * sizeof(T), 8));
    if (e == hipSuccess) {
      ptrs.push_back(q);
      *p = static_cast<T*>(q);
    }
    return e;
  }
};

}  // namespace

namespace {

using BlkKernel = void (*)(BlkArgs, const double*, double*, double*);

// Block-apply kernel of a configuration: cached kernels by gathers in flight
// (4, 8, 16), one or two columns per lane; searching kernels by lookups in
// flight (4, 2).
BlkKernel blk_kernel(int nw, bool cached, int cfg, bool pairs) {
  const int k = blk_cfg_kernel(cfg);
  if (cached && pairs) {
    if (nw == 1) return k == 0 ? k_blk_capply2<1, 4> : (k == 2 ? k_blk_capply2<1, 16> : k_blk_capply2<1, 8>);
    return k == 0 ? k_blk_capply2<2, 4> : (k == 2 ? k_blk_capply2<2, 16> : k_blk_capply2<2, 8>);
  }
  if (cached) {
    if (nw == 1) return k == 0 ? k_blk_capply<1, 4> : (k == 2 ? k_blk_capply<1, 16> : k_blk_capply<1, 8>);
    return k == 0 ? k_blk_capply<2, 4> : (k == 2 ? k_blk_capply<2, 16> : k_blk_capply<2, 8>);
  }
  if (nw == 1) return k == 2 ? k_blk_apply<1, 2> : k_blk_apply<1, 4>;
  return k == 2 ? k_blk_apply<2, 2> : k_blk_apply<2, 4>;
}

}  // namespace

bool blk_pairs(const BlkArgs& a, const double* X) {
  return blk_cached(a) && blk_cfg_pairs(a.cfg) && a.R >= 2 && a.ldx % 2 == 0 &&
         a.ldx >= a.R + (a.R & 1) && reinterpret_cast<uintptr_t>(X) % 16 == 0;
}

int blk_apply_blocks(int nw, int64_t n_tiles, bool cached, int cfg) {
  // resident blocks per CU of each kernel: [cached][nw - 1][kernel][pairs];
  // atomics, as host threads may launch concurrently (they store equal
  // values).  The CU count is a function-local static, whose initialisation
  // is thread-safe, so no thread can see an occupancy without it.
  static std::atomic<int> per_cu[2][2][3][2];
  static const int n_cu = [] {
    int dev = 0, cu = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        cu <= 0)
      cu = 256;
    return cu;
  }();
  const bool pairs = cached && blk_cfg_pairs(cfg);
  std::atomic<int>& occ = per_cu[cached ? 1 : 0][nw - 1][blk_cfg_kernel(cfg)][pairs ? 1 : 0];
  int o_cu = occ.load(std::memory_order_relaxed);
  if (o_cu == 0) {
    int o = 0;
    const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &o, blk_kernel(nw, cached, cfg, pairs), kBlkThreads, 0);
    o_cu = (e == hipSuccess && o > 0) ? o : 4;
    occ.store(o_cu, std::memory_order_relaxed);
  }
  int64_t resident = static_cast<int64_t>(o_cu) * n_cu;
  if (blk_cfg_half(cfg)) resident = std::max<int64_t>(1, resident / 2);
  const int64_t need = (n_tiles + kBlkWaves - 1) / kBlkWaves;
  return static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(need, resident)));
}

int launch_blk_apply(const BlkArgs& a, int nw, int n_blocks, const double* X, double* Y,
                     double* partials, hipStream_t s) {
  if (blk_pairs(a, X)) {
    // ceil(R / 2) lanes per row; a power of two when the partials cross
    // row groups by butterfly
    BlkArgs b = a;
    const int w = (a.R + 1) / 2;
    int G = w;
    if (partials) {
      G = 1;
      while (G < w) G <<= 1;
    }
    b.G = G;
    hipLaunchKernelGGL(blk_kernel(nw, true, a.cfg, true), dim3(n_blocks), dim3(kBlkThreads), 0,
                       s, b, X, Y, partials);
    QSI_CHECK_LAUNCH();
    return QSI_OK;
  }
  hipLaunchKernelGGL(blk_kernel(nw, blk_cached(a), a.cfg, false), dim3(n_blocks),
                     dim3(kBlkThreads), 0, s, a, X, Y, partials);
  QSI_CHECK_LAUNCH();
  return QSI_OK;
}

int op_cache_connectivity(qsi_op_s* op, bool enable, bool allow_pad) {
  auto release = [&]() {
    if (op->d_rowptr) (void)hipFree(op->d_rowptr);
    if (op->d_cj) (void)hipFree(op->d_cj);
    if (op->d_cm) (void)hipFree(op->d_cm);
    op->d_rowptr = nullptr;
    op->d_cj = nullptr;
    op->d_cm = nullptr;
    op->conn_nnz = -1;
  };
  if (!enable) {
    if (op->conn_nnz >= 0) QSI_CHECK_HIP(hipDeviceSynchronize());
    release();
    if (op->variant == QSI_APPLY_CACHED) op->variant = QSI_APPLY_AUTO;
    return QSI_OK;
  }
  if (op->conn_nnz >= 0) return QSI_OK;
  QSI_REQUIRE(op->n_cls > 0 || op->n_off == 0,
              "qsi_op_cache_connectivity: more than 8 coefficient classes");
  const int64_t D = op->basis->dim;
  const auto view = op->basis->view();
  const uint32_t grid = static_cast<uint32_t>((D + 255) / 256);
  void* tmp = nullptr;
  size_t tmp_bytes = 0;
  hipError_t e = hipMalloc(&op->d_rowptr, static_cast<size_t>(D + 1) * sizeof(uint64_t));
  // row offsets: counts (rounded up to multiples of pad), then an in-place
  // exclusive scan of the D + 1 values (the last count is 0); returns the total
  auto offsets = [&](int pad, uint64_t* total) -> hipError_t {
    hipError_t r = hipMemset(op->d_rowptr, 0, static_cast<size_t>(D + 1) * sizeof(uint64_t));
    if (r == hipSuccess && D > 0) {
      if (op->basis->nw == 1)
        hipLaunchKernelGGL(k_cc_count<1>, dim3(grid), dim3(256), 0, 0, view, op->d_off, op->n_off,
                           pad, op->d_rowptr);
      else
        hipLaunchKernelGGL(k_cc_count<2>, dim3(grid), dim3(256), 0, 0, view, op->d_off, op->n_off,
                           pad, op->d_rowptr);
      r = hipGetLastError();
    }
    if (r == hipSuccess && !tmp) {
      r = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, op->d_rowptr, op->d_rowptr, D + 1);
      if (r == hipSuccess) r = hipMalloc(&tmp, std::max<size_t>(tmp_bytes, 1));
    }
    if (r == hipSuccess)
      r = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, op->d_rowptr, op->d_rowptr, D + 1);
    if (r == hipSuccess)
      r = hipMemcpy(total, op->d_rowptr + D, sizeof(uint64_t), hipMemcpyDeviceToHost);
    return r;
  };
  uint64_t nnz = 0, slots = 0;
  if (e == hipSuccess) e = offsets(1, &nnz);
  // Rows padded to whole groups of four entries let the cached apply read
  // every group with one aligned load (about 15 % faster at 64 spins); the
  // padding (at most 3 entries, 15 bytes, per row) is spent only when the
  // padded arrays leave a quarter of the free memory (at least 1 GiB) unused.
  if (e == hipSuccess) {
    size_t free_b = 0, total_b = 0;
    const size_t padded = (nnz + 3 * static_cast<size_t>(D)) * 5;
    if (allow_pad && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
        free_b > padded + std::max<size_t>(free_b / 4, size_t{1} << 30)) {
      e = offsets(4, &slots);
      op->conn_pad = 4;
    } else {
      slots = nnz;
      op->conn_pad = 1;
    }
  }
  if (tmp) (void)hipFree(tmp);
  // rounded up to whole 16-byte groups of columns (the cached apply reads the
  // columns four at a time from 16-byte aligned addresses, and the class
  // bytes four at a time)
  const size_t cap = std::max<size_t>((slots + 3) & ~uint64_t{3}, 4);
  if (e == hipSuccess) e = hipMalloc(&op->d_cj, cap * sizeof(uint32_t));
  if (e == hipSuccess) e = hipMalloc(&op->d_cm, cap);
  if (e == hipSuccess && D > 0) {
    if (op->basis->nw == 1)
      hipLaunchKernelGGL(k_cc_fill<1>, dim3(grid), dim3(256), 0, 0, view, op->d_off, op->n_off,
                         op->d_rowptr, op->d_cj, op->d_cm);
    else
      hipLaunchKernelGGL(k_cc_fill<2>, dim3(grid), dim3(256), 0, 0, view, op->d_off, op->n_off,
                         op->d_rowptr, op->d_cj, op->d_cm);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    release();
    set_error(std::string("qsi_op_cache_connectivity: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? QSI_E_NOMEM : QSI_E_HIP;
  }
  op->conn_nnz = nnz;
  return QSI_OK;
}

}  // namespace qsi

using namespace qsi;

namespace {

// Block-apply arguments over the sectors' operators (rows concatenated in
// sector order), R columns with row stride R; *rows = total rows.
BlkArgs sector_args(const qsi_op_t* ops, int n_sectors, int R, const double* const* d_coeffs,
                    int64_t* rows_out) {
  BlkArgs args{};
  args.n_sec = n_sectors;
  args.R = R;
  int G = 64;
  while (G / 2 >= R && G > 1) G >>= 1;
  args.G = G;
  args.ldx = args.ldy = R;
  args.accumulate = 0;
  args.y_colmajor = 0;
  int64_t tiles = 0, rows = 0;
  for (int s = 0; s < n_sectors; ++s) {
    const qsi_op_s* op = ops[s];
    BlkSector& S = args.sec[s];
    S.b = op->basis->view();
    S.off = op->d_off;
    S.dg = op->d_diag;
    S.cls = op->d_cls;
    S.n_cls = op->n_cls;
    S.rowptr = op->d_rowptr;
    S.cj = op->d_cj;
    S.cm = op->d_cm;
    S.n_off = op->n_off;
    S.n_diag = op->n_diag;
    S.coeffs = d_coeffs ? d_coeffs[s] : nullptr;
    S.tile_begin = tiles;
    S.row_off = rows;
    tiles += (op->basis->dim + kTileRows - 1) / kTileRows;
    rows += op->basis->dim;
  }
  args.n_tiles = tiles;
  args.cfg = ops[0]->blk_cfg;
  *rows_out = rows;
  return args;
}

}  // namespace

extern "C" int qsi_lanczos_block(const qsi_op_t* ops, int n_sectors, int R,
                                 const double* const* d_V0,
                                 const double* const* d_coeffs, int M,
                                 double* h_alpha, double* h_beta, int use_graph,
                                 double* elapsed_ms) {
  QSI_REQUIRE(ops && d_V0 && h_alpha && h_beta, "qsi_lanczos_block: null argument");
  QSI_REQUIRE(n_sectors >= 1 && n_sectors <= kMaxSectors,
              "qsi_lanczos_block: n_sectors must be in [1, 8]");
  QSI_REQUIRE(R >= 1 && R <= 64, "qsi_lanczos_block: R must be in [1, 64]");
  QSI_REQUIRE(M >= 1, "qsi_lanczos_block: M must be >= 1");
  const int nw = ops[0]->basis->nw;
  for (int s = 0; s < n_sectors; ++s) {
    QSI_REQUIRE(ops[s] != nullptr && ops[s]->basis->dim > 0, "qsi_lanczos_block: empty sector");
    QSI_REQUIRE(ops[s]->basis->nw == nw, "qsi_lanczos_block: sectors mix <=64 and >64 spins");
    QSI_REQUIRE(d_V0[s] != nullptr, "qsi_lanczos_block: null start block");
    QSI_REQUIRE(!d_coeffs || !d_coeffs[s] || ops[s]->n_slots > 0,
                "qsi_lanczos_block: coefficients given for an operator without slots");
  }
  int64_t rows = 0;
  BlkArgs args = sector_args(ops, n_sectors, R, d_coeffs, &rows);
  const int64_t tiles = args.n_tiles;
  // persistent grid: enough wavefronts to fill the chip, never more than tiles
  args.cfg = ops[0]->blk_cfg;
  const int n_blocks = blk_apply_blocks(nw, tiles, blk_cached(args), args.cfg);
  const int64_t n_waves = static_cast<int64_t>(n_blocks) * kBlkWaves;
  const size_t blk = static_cast<size_t>(rows) * R;
  DevMem mem;
  double *A = nullptr, *B = nullptr, *C = nullptr, *part = nullptr;
  double *alpha = nullptr, *beta = nullptr, *invb = nullptr, *ha = nullptr, *hb = nullptr;
  int* step = nullptr;
  const int SR = n_sectors * R;
  QSI_CHECK_HIP(mem.alloc(&A, blk));
  QSI_CHECK_HIP(mem.alloc(&B, blk));
  QSI_CHECK_HIP(mem.alloc(&C, blk));
  QSI_CHECK_HIP(mem.alloc(&part, static_cast<size_t>(n_sectors) * n_waves * R));
  QSI_CHECK_HIP(mem.alloc(&alpha, SR));
  QSI_CHECK_HIP(mem.alloc(&beta, SR));
  QSI_CHECK_HIP(mem.alloc(&invb, SR));
  QSI_CHECK_HIP(mem.alloc(&ha, static_cast<size_t>(SR) * M));
  QSI_CHECK_HIP(mem.alloc(&hb, static_cast<size_t>(SR) * M));
  QSI_CHECK_HIP(mem.alloc(&step, 1));
  hipStream_t st = nullptr;
  QSI_CHECK_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{st};
  for (int s = 0; s < n_sectors; ++s) {
    QSI_CHECK_HIP(hipMemcpyAsync(B + args.sec[s].row_off * R, d_V0[s],
                                 static_cast<size_t>(ops[s]->basis->dim) * R * sizeof(double),
                                 hipMemcpyDeviceToDevice, st));
  }
  QSI_CHECK_HIP(hipMemsetAsync(A, 0, blk * sizeof(double), st));
  QSI_CHECK_HIP(hipMemsetAsync(beta, 0, SR * sizeof(double), st));
  QSI_CHECK_HIP(hipMemsetAsync(step, 0, sizeof(int), st));
  // normalise the start columns: B = B / |B_c|
  hipLaunchKernelGGL(k_blk_update, dim3(n_blocks), dim3(kBlkThreads), 0, st, args,
                     static_cast<const double*>(nullptr), B, B, static_cast<const double*>(nullptr),
                     static_cast<const double*>(nullptr), part, static_cast<const double*>(nullptr),
                     static_cast<const double*>(nullptr));
  QSI_CHECK_LAUNCH();
  // (C = B - 0*B: the kernel writes back B unchanged and produces |B_c|^2)
  hipLaunchKernelGGL(k_blk_finalize, dim3(SR), dim3(256), 0, st, part, n_waves, R, 1,
                     beta, invb, static_cast<double*>(nullptr), static_cast<const int*>(nullptr), M,
                     static_cast<const double*>(nullptr));
  QSI_CHECK_LAUNCH();
  hipLaunchKernelGGL(k_blk_scale, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, B, invb,
                     static_cast<int*>(nullptr));
  QSI_CHECK_LAUNCH();
  QSI_CHECK_HIP(hipMemsetAsync(beta, 0, SR * sizeof(double), st));

  // The blocks hold the Lanczos vectors unnormalised, each column with its
  // scale (v = scale * column) in scl[k] beside bufs[k]: the step applies H to
  // the unnormalised B, folds B's scale into alpha and the update, and the
  // new column's scale is 1/beta — no pass rewrites a block to normalise it.
  // B starts normalised (scale 1); A is zero.
  double* bufs[3] = {A, B, C};
  double* scl[3] = {nullptr, nullptr, nullptr};
  for (int k = 0; k < 3; ++k) {
    QSI_CHECK_HIP(mem.alloc(&scl[k], SR));
    std::vector<double> ones(static_cast<size_t>(SR), 1.0);
    QSI_CHECK_HIP(hipMemcpyAsync(scl[k], ones.data(), SR * sizeof(double),
                                 hipMemcpyHostToDevice, st));
    QSI_CHECK_HIP(hipStreamSynchronize(st));
  }
  auto enqueue_step = [&](int p) -> int {
    double* a = bufs[p % 3];
    double* b = bufs[(p + 1) % 3];
    double* c = bufs[(p + 2) % 3];
    const double* sa = scl[p % 3];
    const double* sb = scl[(p + 1) % 3];
    double* sc = scl[(p + 2) % 3];
    const int rc = launch_blk_apply(args, nw, n_blocks, b, c, part, st);
    if (rc != QSI_OK) return rc;
    hipLaunchKernelGGL(k_blk_finalize, dim3(SR), dim3(256), 0, st, part, n_waves, R, 0,
                       alpha, static_cast<double*>(nullptr), ha, step, M, sb);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_update, dim3(n_blocks), dim3(kBlkThreads), 0, st, args, a, b, c,
                       alpha, beta, part, sa, sb);
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_finalize, dim3(SR), dim3(256), 0, st, part, n_waves, R, 1,
                       beta, sc, hb, step, M, static_cast<const double*>(nullptr));
    QSI_CHECK_LAUNCH();
    hipLaunchKernelGGL(k_blk_step, dim3(1), dim3(64), 0, st, step);
    QSI_CHECK_LAUNCH();
    return QSI_OK;
  };

  hipEvent_t e0, e1;
  QSI_CHECK_HIP(hipEventCreate(&e0));
  QSI_CHECK_HIP(hipEventCreate(&e1));
  struct EvGuard {
    hipEvent_t a, b;
    ~EvGuard() {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
    }
  } eg{e0, e1};
  if (use_graph) {
    // one graph per rotation phase p: (A, B, C) = (bufs[p], bufs[p+1], bufs[p+2])
    hipGraphExec_t exec[3] = {nullptr, nullptr, nullptr};
    struct ExecGuard {
      hipGraphExec_t* e;
      ~ExecGuard() {
        for (int k = 0; k < 3; ++k)
          if (e[k]) (void)hipGraphExecDestroy(e[k]);
      }
    } xg{exec};
    for (int p = 0; p < 3 && p < M; ++p) {
      hipGraph_t g = nullptr;
      QSI_CHECK_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      const int rc = enqueue_step(p);
      hipGraph_t g2 = nullptr;
      const hipError_t ec = hipStreamEndCapture(st, &g2);
      g = g2;
      if (rc != QSI_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      QSI_CHECK_HIP(ec);
      const hipError_t ei = hipGraphInstantiate(&exec[p], g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      QSI_CHECK_HIP(ei);
    }
    QSI_CHECK_HIP(hipEventRecord(e0, st));
    for (int j = 0; j < M; ++j) QSI_CHECK_HIP(hipGraphLaunch(exec[j % 3], st));
    QSI_CHECK_HIP(hipEventRecord(e1, st));
    QSI_CHECK_HIP(hipStreamSynchronize(st));
  } else {
    QSI_CHECK_HIP(hipEventRecord(e0, st));
    for (int j = 0; j < M; ++j) {
      const int p = j % 3;
      const int rc = enqueue_step(p);
      if (rc != QSI_OK) return rc;
    }
    QSI_CHECK_HIP(hipEventRecord(e1, st));
    QSI_CHECK_HIP(hipStreamSynchronize(st));
  }
  if (elapsed_ms) {
    float ms = 0.f;
    QSI_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
    *elapsed_ms = ms;
  }
  QSI_CHECK_HIP(hipMemcpy(h_alpha, ha, static_cast<size_t>(SR) * M * sizeof(double),
                          hipMemcpyDeviceToHost));
  QSI_CHECK_HIP(hipMemcpy(h_beta, hb, static_cast<size_t>(SR) * M * sizeof(double),
                          hipMemcpyDeviceToHost));
  return QSI_OK;
}

// Eigenvalues and first eigenvector components of one Lanczos tridiagonal
// (Gauss quadrature nodes and weights for FTLM).
namespace {

// Average milliseconds per call of f(), timed with events on stream s: after
// one warm-up call, `reps` calls, repeated with more calls when those took
// under 5 ms.
template <typename F>
int time_ms(F&& f, int reps, hipStream_t s, double* ms) {
  hipEvent_t e0, e1;
  QSI_CHECK_HIP(hipEventCreate(&e0));
  QSI_CHECK_HIP(hipEventCreate(&e1));
  int rc = f();
  for (int pass = 0; pass < 2 && rc == QSI_OK; ++pass) {
    (void)hipEventRecord(e0, s);
    for (int r = 0; r < reps && rc == QSI_OK; ++r) rc = f();
    (void)hipEventRecord(e1, s);
    (void)hipEventSynchronize(e1);
    float t = 0.f;
    (void)hipEventElapsedTime(&t, e0, e1);
    *ms = t / reps;
    if (pass == 0) {
      const int want = *ms > 0.0 ? std::min(200, static_cast<int>(5.0 / *ms) + 1) : reps;
      if (want <= reps) break;
      reps = want;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}

// Times every block-apply configuration (kernel x grid) of `a` on X, Y
// (rows x R) and returns the fastest.
int tune_block(BlkArgs a, int nw, int reps, const double* X, double* Y, int* pick,
               double* pick_ms) {

  const bool cached = blk_cached(a);
  const std::vector<int> kernels = cached ? std::vector<int>{0, 1, 2} : std::vector<int>{1, 2};
  const int n_pairs = cached && a.R >= 2 ? 2 : 1;
  *pick = -1;
  for (int k : kernels) {
    for (int cf = 0; cf < 2 * n_pairs; ++cf) {
      const int half = cf & 1, pairs = cf >> 1;
      a.cfg = k | (half << 2) | (pairs << 3);
      const int nb = blk_apply_blocks(nw, a.n_tiles, cached, a.cfg);
      double ms = 0.0;
      const int rc = time_ms([&]() { return launch_blk_apply(a, nw, nb, X, Y, nullptr, nullptr); },
                             reps, nullptr, &ms);
      if (rc != QSI_OK) return rc;
      if (*pick < 0 || ms < *pick_ms) {
        *pick = a.cfg;
        *pick_ms = ms;
      }
    }
  }
  return QSI_OK;
}

}  // namespace

extern "C" int qsi_block_autotune(const qsi_op_t* ops, int n_sectors, int R, int reps, int* best,
                                  double* best_ms) {
  QSI_REQUIRE(ops != nullptr, "qsi_block_autotune: null argument");
  QSI_REQUIRE(n_sectors >= 1 && n_sectors <= kMaxSectors,
              "qsi_block_autotune: n_sectors must be in [1, 8]");
  QSI_REQUIRE(R >= 1 && R <= 64, "qsi_block_autotune: R must be in [1, 64]");
  const int nw = ops[0]->basis->nw;
  for (int s = 0; s < n_sectors; ++s) {
    QSI_REQUIRE(ops[s] != nullptr && ops[s]->basis->dim > 0, "qsi_block_autotune: empty sector");
    QSI_REQUIRE(ops[s]->basis->nw == nw, "qsi_block_autotune: sectors mix <=64 and >64 spins");
  }
  reps = std::max(1, reps);
  // each configuration runs a few steps of the block Lanczos recurrence
  // itself (the apply between the recurrence's vector kernels behaves
  // differently from back-to-back applies), from Gaussian start blocks
  constexpr int kSteps = 8;
  DevMem mem;
  std::vector<const double*> v0(n_sectors, nullptr), cp(n_sectors, nullptr);
  for (int s = 0; s < n_sectors; ++s) {
    double* v = nullptr;
    const int64_t n = ops[s]->basis->dim * R;
    QSI_CHECK_HIP(mem.alloc(&v, static_cast<size_t>(n)));
    int rc = launch_randn(v, n, 12345 + s, nullptr);
    if (rc != QSI_OK) return rc;
    v0[s] = v;
    if (ops[s]->n_slots > 0) {
      double* c = nullptr;
      QSI_CHECK_HIP(mem.alloc(&c, static_cast<size_t>(ops[s]->n_slots) * R));
      rc = launch_randn(c, static_cast<int64_t>(ops[s]->n_slots) * R, 7 + s, nullptr);
      if (rc != QSI_OK) return rc;
      cp[s] = c;
    }
  }
  QSI_CHECK_HIP(hipDeviceSynchronize());
  std::vector<double> ha(static_cast<size_t>(n_sectors) * R * kSteps),
      hb(static_cast<size_t>(n_sectors) * R * kSteps);
  std::vector<int> saved(n_sectors);
  for (int s = 0; s < n_sectors; ++s) saved[s] = ops[s]->blk_cfg;
  int64_t rows = 0;
  const bool cached = blk_cached(sector_args(ops, n_sectors, R, cp.data(), &rows));
  const std::vector<int> kernels = cached ? std::vector<int>{0, 1, 2} : std::vector<int>{1, 2};
  const int n_pairs = cached && R >= 2 ? 2 : 1;
  int pick = -1;
  double pick_ms = 0.0;
  for (int k : kernels) {
    for (int cf = 0; cf < 2 * n_pairs; ++cf) {
      const int half = cf & 1, pairs = cf >> 1;
      const int cfg = k | (half << 2) | (pairs << 3);
      for (int s = 0; s < n_sectors; ++s) ops[s]->blk_cfg = cfg;
      double ms = 0.0;
      for (int r = 0; r <= reps; ++r) {  // the first run warms up
        double t = 0.0;
        const int rc = qsi_lanczos_block(ops, n_sectors, R, v0.data(), cp.data(), kSteps,
                                         ha.data(), hb.data(), 0, &t);
        if (rc != QSI_OK) {
          for (int q = 0; q < n_sectors; ++q) ops[q]->blk_cfg = saved[q];
          return rc;
        }
        if (r > 0) ms += t / kSteps;
      }
      ms /= reps;
      if (pick < 0 || ms < pick_ms) {
        pick = cfg;
        pick_ms = ms;
      }
    }
  }
  for (int s = 0; s < n_sectors; ++s) ops[s]->blk_cfg = pick;
  if (best) *best = pick;
  if (best_ms) *best_ms = pick_ms;
  return QSI_OK;
}

extern "C" int qsi_op_autotune(qsi_op_t op, int R, int reps, double* d_x, double* d_y,
                               int* best, double* best_ms) {
  QSI_REQUIRE(op != nullptr, "qsi_op_autotune: null operator");
  QSI_REQUIRE(R >= 0 && R <= 64, "qsi_op_autotune: R must be in [0, 64]");
  reps = std::max(1, reps);
  const int64_t D = op->basis->dim;
  if (D == 0) return QSI_OK;
  const size_t n = static_cast<size_t>(D) * std::max(R, 1);
  DevMem mem;
  if (!d_x) {
    QSI_CHECK_HIP(mem.alloc(&d_x, n));
    // Gaussian input: every term's |x_J| is above the apply tolerance
    const int rc = launch_randn(d_x, static_cast<int64_t>(n), 12345, nullptr);
    if (rc != QSI_OK) return rc;
  }
  if (!d_y) QSI_CHECK_HIP(mem.alloc(&d_y, n));
  double* coeffs = nullptr;
  if (R > 0 && op->n_slots > 0) {
    QSI_CHECK_HIP(mem.alloc(&coeffs, static_cast<size_t>(op->n_slots) * R));
    const int rc = launch_randn(coeffs, static_cast<int64_t>(op->n_slots) * R, 7, nullptr);
    if (rc != QSI_OK) return rc;
  }
  QSI_CHECK_HIP(hipDeviceSynchronize());
  int pick = -1;
  double pick_ms = 0.0;
  if (R == 0) {
    std::vector<int> cand = {QSI_APPLY_LANE, QSI_APPLY_LANE2, QSI_APPLY_LANE4, QSI_APPLY_LANE8,
                             QSI_APPLY_WINDOW};
    if (op->basis->d_hash) {
      cand.push_back(QSI_APPLY_HASH);
      cand.push_back(QSI_APPLY_HASH4);
    }
    if (op->basis->d_slots) {
      cand.push_back(QSI_APPLY_SLOT2);
      cand.push_back(QSI_APPLY_SLOT3);
      cand.push_back(QSI_APPLY_SLOT4);
    }
    if (op->basis->d_lowkeys) {
      cand.push_back(QSI_APPLY_LOW2);
      cand.push_back(QSI_APPLY_LOW4);
      cand.push_back(QSI_APPLY_LOW8);
    }
    // the cached kernel once per gather depth (4, 8, 16, 32)
    std::vector<int> pipes(cand.size(), 8);
    if (op->conn_nnz >= 0) {
      for (int p : {4, 8, 16, 32}) {
        cand.push_back(QSI_APPLY_CACHED);
        pipes.push_back(p);
      }
    }
    const int saved = op->variant, saved_pipe = op->cached_pipe;
    int pick_pipe = saved_pipe;
    for (size_t k = 0; k < cand.size(); ++k) {
      op->variant = cand[k];
      op->cached_pipe = pipes[k];
      double ms = 0.0;
      const int rc = time_ms(
          [&]() { return launch_apply(op, d_x, d_y, nullptr, 0.0, nullptr, nullptr, nullptr); },
          reps, nullptr, &ms);
      if (rc != QSI_OK) {
        op->variant = saved;
        op->cached_pipe = saved_pipe;
        return rc;
      }
      if (pick < 0 || ms < pick_ms) {
        pick = cand[k];
        pick_pipe = pipes[k];
        pick_ms = ms;
      }
    }
    op->variant = pick;
    op->cached_pipe = pick_pipe;
  } else {
    const double* cp[1] = {coeffs};
    int64_t rows = 0;
    BlkArgs a = sector_args(&op, 1, R, cp, &rows);
    const int rc = tune_block(a, op->basis->nw, reps, d_x, d_y, &pick, &pick_ms);
    if (rc != QSI_OK) return rc;
    op->blk_cfg = pick;
  }
  if (best) *best = pick;
  if (best_ms) *best_ms = pick_ms;
  return QSI_OK;
}

extern "C" int qsi_op_cache_connectivity(qsi_op_t op, int enable, int64_t* nnz) {
  QSI_REQUIRE(op != nullptr, "qsi_op_cache_connectivity: null operator");
  QSI_REQUIRE(enable >= 0 && enable <= 2, "qsi_op_cache_connectivity: enable must be 0, 1 or 2");
  const int rc = op_cache_connectivity(op, enable != 0, enable == 1);
  if (nnz) *nnz = op->conn_nnz;
  return rc;
}

extern "C" int qsi_op_connectivity_padding(qsi_op_t op) {
  return op && op->conn_nnz >= 0 ? op->conn_pad : -1;
}

extern "C" int qsi_tridiag_eig(const double* d, const double* e, int n,
                               double* evals, double* first_row) {
  QSI_REQUIRE(d && evals && n >= 1 && (n == 1 || e), "qsi_tridiag_eig: bad arguments");
  std::vector<double> dv(d, d + n), ev(e ? e : d, e ? e + std::max(n - 1, 0) : d), w, fr;
  if (!hostla::tridiag_eig(dv, ev, w, nullptr, first_row ? &fr : nullptr)) {
    set_error("qsi_tridiag_eig: no convergence");
    return QSI_E_NOCONV;
  }
  std::copy(w.begin(), w.end(), evals);
  if (first_row) std::copy(fr.begin(), fr.end(), first_row);
  return QSI_OK;
}
