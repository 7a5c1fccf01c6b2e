// Block kernels shared by the FTLM block Lanczos and the serving engine:
// many columns (start vectors or requests) over one or more operators, rows
// cut into 64-row tiles walked by a persistent grid of wavefronts.
#pragma once

#include <hip/hip_runtime.h>

#include "qsi_device.hpp"
#include "qsi_internal.hpp"
#include "qsi_reduce.hpp"

namespace qsi {
namespace {

constexpr int kBlkThreads = 256;
constexpr int kBlkWaves = kBlkThreads / 64;
constexpr int kTileRows = 64;
constexpr int kMaxSectors = 8;

struct BlkSector {
  BasisView b;
  const OffTerm* off;
  const DiagTerm* dg;
  const double* coeffs;  // n_slots x R, or null
  int n_off;
  int n_diag;
  int64_t tile_begin;  // first global tile of this sector
  int64_t row_off;     // first row of this sector in the concatenated blocks
};

struct BlkArgs {
  BlkSector sec[kMaxSectors];
  int n_sec;
  int64_t n_tiles;
  int R;
  int G;            // lanes per row group (power of two >= R)
  int64_t ldx, ldy; // row strides of the X and Y blocks (>= R)
  int accumulate;   // Y += H X instead of Y = H X (apply kernel)
  int y_colmajor;   // Y stored column-major: Y[col * ldy + row] (apply kernel)
};

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

// persistent grid: 256 CUs x 20 resident waves (the apply kernel's occupancy
// with its 32 KB of LDS per block)
constexpr int64_t kBlkMaxWaves = 256 * 20;
constexpr int kBlkCap = 1024;        // LDS flip entries per wavefront and round
constexpr int kBlkSuper = 256;       // off-diagonal terms per round (4 masks)
constexpr int kBlkBatch = 4;         // lookups in flight per lane (phase A)
constexpr int kBlkPipe = 4;          // X loads in flight per lane (phase B)
constexpr uint32_t kNoRank = 0xFFFFFFFFu;

__device__ __forceinline__ void blk_wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
  const uint32_t lo = __shfl(static_cast<uint32_t>(v), src, 64);
  const uint32_t hi = __shfl(static_cast<uint32_t>(v >> 32), src, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

// Ranks of up to kBlkBatch keys with all bisections interleaved; kNoRank for
// keys absent from the basis (or inactive slots).
template <int NW>
__device__ __forceinline__ void lookup_batch(const BasisView& b, const State<NW>* key,
                                             const bool* act, uint32_t* rank) {
  uint32_t lo[kBlkBatch], n[kBlkBatch];
#pragma unroll
  for (int g = 0; g < kBlkBatch; ++g) {
    lo[g] = 0;
    n[g] = 0;
    if (act[g]) {
      const uint32_t k = bucket_key<NW>(key[g], b.shift);
      if (k < (1u << b.kbits)) {
        lo[g] = b.table[k];
        n[g] = b.table[k + 1] - lo[g];
      }
    }
  }
  bool more = true;
  while (more) {
    more = false;
#pragma unroll
    for (int g = 0; g < kBlkBatch; ++g) {
      if (n[g] > 1) {
        const uint32_t half = n[g] >> 1;
        const State<NW> m = load_state_cached<NW>(b.states, lo[g] + half);
        lo[g] = st_le<NW>(m, key[g]) ? lo[g] + half : lo[g];
        n[g] -= half;
        more |= n[g] > 1;
      }
    }
  }
#pragma unroll
  for (int g = 0; g < kBlkBatch; ++g) {
    rank[g] = kNoRank;
    if (n[g] == 1) {
      const State<NW> f = load_state_cached<NW>(b.states, lo[g]);
      if (st_eq<NW>(f, key[g])) rank[g] = lo[g];
    }
  }
}

// Y[:, c] = H_c X[:, c] (or Y += H_c X) for R columns of row-major blocks,
// and (partials != null) per-wave, per-sector partials of <X[:, c], Y[:, c]>.
//
// Each wavefront takes a 64-row tile in two phases per round:
//   A  one lane per row finds the row's connected terms, looks up the
//      pre-image ranks kBlkBatch at a time and writes (rank, term, sign)
//      entries to LDS in term order;
//   B  G lanes per row (one per column) walk the row's entries and read
//      X[J, 0:R] as one coalesced segment per connected state.
// A round covers up to kBlkSuper terms and as many rows as fit kBlkCap LDS
// entries; partial sums of rows cut by a round boundary go through Y.
// Contributions are added in term order, then the diagonal terms in order, so
// column c's result depends only on X[:, c] and its coefficients.
template <int NW>
__global__ __launch_bounds__(kBlkThreads) void k_blk_apply(
    BlkArgs a, const double* __restrict__ X, double* __restrict__ Y,
    double* __restrict__ partials) {
  // entry: rank (bits 0-31, kNoRank if absent), term - round base (32-47),
  // total sign negative (bit 48)
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
  for (int64_t t = wave; t < a.n_tiles; t += n_waves) {
    const int s = __builtin_amdgcn_readfirstlane(find_sector(a, t));
    if (s != cur) {
      flush_to(s, cur);
      cur = s;
      dot = 0.0;
    }
    const BlkSector& S = a.sec[s];
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
              State<NW> key[kBlkBatch];
              bool act[kBlkBatch];
              int tl[kBlkBatch];
              uint32_t rk[kBlkBatch];
#pragma unroll
              for (int g = 0; g < kBlkBatch; ++g) {
                act[g] = m != 0;
                tl[g] = q * 64 + (act[g] ? __builtin_ctzll(m) : 0);
                m &= m - 1;
                key[g] = xor_state<NW>(st, S.off[tb + tl[g]].X);
              }
              lookup_batch<NW>(S.b, key, act, rk);
#pragma unroll
              for (int g = 0; g < kBlkBatch; ++g) {
                if (act[g]) {
                  const OffTerm& T = S.off[tb + tl[g]];
                  const bool ng = (T.sgn < 0.0) != static_cast<bool>(parity_notz<NW>(st, T.Z));
                  ent[pos++] = static_cast<uint64_t>(rk[g]) |
                               (static_cast<uint64_t>(tl[g]) << 32) |
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
                const OffTerm& T = S.off[tb + static_cast<int>((en[u] >> 32) & 0xFFFF)];
                c[u] = (S.coeffs && T.slot >= 0) ? S.coeffs[T.slot * R + col] : T.c;
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

// C -= alpha B + beta A (beta, A skipped when A is null); |C_c|^2 partials.
__global__ __launch_bounds__(kBlkThreads) __attribute__((unused)) void k_blk_update(
    BlkArgs a, const double* __restrict__ A, const double* B,
    double* C, const double* __restrict__ alpha,
    const double* __restrict__ beta, double* __restrict__ partials) {
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
    double part = 0.0;
    for (int64_t i = r_begin + grp; i < r_end; i += rows_per_pass) {
      if (!col_ok) continue;
      const int64_t gi = (S.row_off + i) * R + col;
      double w = C[gi] - al * B[gi];
      if (A) w -= be * A[gi];
      C[gi] = w;
      part += w * w;
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
// mode 0 (alpha): out[s*R+c] = sum.  mode 1 (beta): out[s*R+c] = sqrt(sum),
// out2[s*R+c] = 1/beta or 0.  hist (optional) receives the value at
// hist[(s*R + c)*M + step[0]].
__global__ __launch_bounds__(256) __attribute__((unused)) void k_blk_finalize(
    const double* __restrict__ partials, int64_t n_waves, int R, int mode,
    double* __restrict__ out, double* __restrict__ out2, double* __restrict__ hist,
    const int* __restrict__ step, int M) {
  const int sc = blockIdx.x;  // s * R + c
  const int s = sc / R, c = sc % R;
  double v = 0.0;
  for (int64_t w = threadIdx.x; w < n_waves; w += 256)
    v += partials[(s * n_waves + w) * R + c];
  v = block_sum<256>(v);
  if (threadIdx.x == 0) {
    double r = v;
    if (mode == 1) {
      r = sqrt(v);
      out2[sc] = r > 1e-300 ? 1.0 / r : 0.0;
    }
    out[sc] = r;
    if (hist) hist[static_cast<int64_t>(sc) * M + step[0]] = r;
  }
}

// C *= 1/beta per (sector, column); optionally advances the step counter.
__global__ __launch_bounds__(kBlkThreads) __attribute__((unused)) void k_blk_scale(
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
    for (int64_t i = r_begin + grp; i < r_end; i += rows_per_pass)
      C[(S.row_off + i) * R + col] *= f;
  }
  if (step && blockIdx.x == 0 && threadIdx.x == 0) step[0] += 1;
}

}  // namespace
}  // namespace qsi
