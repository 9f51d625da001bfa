// IC row sampling by geometric gaps: the IC part of the per-sample stream
// contract (common.h), shared by host and device so both compute the same
// live edges bit for bit.
//
// The reference draws one uniform per in-edge and keeps the edge when
// u <= w (add_rrrset.h:128-134).  Here the live edges of a row are drawn
// as gaps between candidates, so a row costs draws in proportion to its
// live edges rather than to its length; every edge is still live
// independently with probability t / 2^24, t = floor(w 2^24) its integer
// threshold (ic_threshold in common.h).
//
// Row v = in-edges [b, b + d), thresholds t_e, T = max_e t_e:
//   T = 0: no edge is live; the row draws nothing.
//   block size B = 2^beta, beta = the least value in [4, 20] with T 2^beta >= 2^25
//   (about two candidates per block); block k holds edges b + kB + [0, len),
//   len = min(B, d - kB), and draws elements 1 + 2(b + kB) + j of the sample's
//   stream, j = 0, 1, ... (at most 2 len of them).
//   Candidate i of the block: gap element g_i = element 2i, acceptance element
//   a_i = element 2i + 1.  With U = hi32(g_i) + 1 in [1, 2^32],
//       gap = floor(NL(U) / q),  NL(U) = -log2(U / 2^32) in units of 2^-58,
//       q = NL((2^24 - T) 2^8) (-log2(1 - T / 2^24)); q = 0 when T = 2^24
//   (then every edge is a candidate, gap 0).  The candidate is the edge
//   `gap` positions after the previous one (or after the block start); a gap
//   reaching past the block ends it.
//   A candidate edge e is live iff hi32(a_i) T < t_e 2^32, i.e. it is kept
//   with probability t_e / T (always when t_e = T: rows whose edges share one
//   threshold, such as weighted cascade, never read a_i or t_e).
// So each edge is a candidate independently with probability T / 2^24
// (geometric gaps) and live with probability t_e / 2^24.
//
// NL is a fixed-point log: a 256-entry table of log2(1 + i/256) and a
// four-term series for the rest, integer arithmetic only, error below
// 2^-41; it is strictly decreasing in U (consecutive U differ by more than
// 2^-32 in -log2), so "gap >= B" is exactly "U <= ufull" for a per-row
// threshold ufull, which lets a block without candidates cost one compare.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rip {

#include "ic_gap_tables.inc"

constexpr uint32_t kIcOne = 1u << 24;          // threshold of w >= 1
constexpr uint32_t kIcBetaMin = 4, kIcBetaMax = 20;

__host__ __device__ inline uint64_t mulhi_u64(uint64_t a, uint64_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __umul64hi(a, b);
#else
  return static_cast<uint64_t>((static_cast<unsigned __int128>(a) * b) >> 64);
#endif
}

__host__ __device__ inline int clz_u64(uint64_t x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __clzll(static_cast<long long>(x));
#else
  return __builtin_clzll(x);
#endif
}

// -log2(u / 2^32) for u in [1, 2^32], in units of 2^-58 (0 for u >= 2^32),
// with the tables kLog2C and kRecipC read from log2c and recipc (kernels pass
// copies in LDS).
__host__ __device__ inline uint64_t ic_neg_log2_tab(uint64_t u, const uint64_t *log2c,
                                                    const uint64_t *recipc) {
  if (u >= (1ull << 32)) return 0;
  const int e = 63 - clz_u64(u);                   // floor(log2 u), 0..31
  const uint64_t m = u << (63 - e);                // mantissa in [2^63, 2^64)
  const uint32_t i = static_cast<uint32_t>(m >> 55) & 255u;
  const uint64_t r = m & ((1ull << 55) - 1);       // m = 2^63 (1 + i/256) + r
  // x = r 2^-63 / (1 + i/256) < 2^-8, held as X = x 2^64
  const uint64_t c = recipc[i];
  const uint64_t X = (mulhi_u64(r, c) << 2) | ((r * c) >> 62);
  const uint64_t X2 = mulhi_u64(X, X), X3 = mulhi_u64(X2, X), X4 = mulhi_u64(X3, X);
  const uint64_t ln = X - X2 / 2 + X3 / 3 - X4 / 4;    // ln(1 + x), Q0.64
  const uint64_t frac = log2c[i] + (mulhi_u64(ln, RIP_INV_LN2_Q63) << 1);
  return (static_cast<uint64_t>(32 - e) << 58) - (frac >> 6);
}

__host__ __device__ inline uint64_t ic_neg_log2(uint64_t u) {
  return ic_neg_log2_tab(u, kLog2C, kRecipC);
}

// The tables of ic_neg_log2, for a copy in LDS.
struct IcLogTables {
  uint64_t log2c[256], recipc[256];
};

// Fills t (LDS) from the constant tables; every thread of the workgroup calls
// it, the caller synchronises.
__device__ inline void ic_load_tables(IcLogTables *t) {
  for (uint32_t i = threadIdx.x; i < 256; i += blockDim.x) {
    t->log2c[i] = kLog2C[i];
    t->recipc[i] = kRecipC[i];
  }
}

// q of a row whose largest threshold is T (1 <= T <= 2^24).
__host__ __device__ inline uint64_t ic_gap_q(uint32_t T) {
  if (T >= kIcOne) return 0;
  return ic_neg_log2(static_cast<uint64_t>(kIcOne - T) << 8);
}

// log2 of the block size of a row whose largest threshold is T >= 1.
__host__ __device__ inline uint32_t ic_block_log2(uint32_t T) {
  uint32_t beta = kIcBetaMin;
  while (beta < kIcBetaMax && (static_cast<uint64_t>(T) << beta) < (1ull << 25)) ++beta;
  return beta;
}

// Gap drawn from stream element x.
__host__ __device__ inline uint64_t ic_gap(uint64_t x, uint64_t q) {
  if (q == 0) return 0;
  return ic_neg_log2((x >> 32) + 1) / q;
}

// Candidate edge with threshold t kept, acceptance element a, row maximum T.
__host__ __device__ inline bool ic_accept(uint64_t a, uint32_t T, uint32_t t) {
  return (a >> 32) * T < (static_cast<uint64_t>(t) << 32);
}

// The largest U = hi32(x) + 1 whose gap is at least len (a block of len
// edges without candidates), 0 if there is none.  NL is strictly decreasing,
// so gap >= len holds exactly for U <= this value.
__host__ __device__ inline uint32_t ic_skip_len(uint64_t q, uint32_t len) {
  if (q == 0 || len == 0) return 0;
  const uint64_t top = ic_neg_log2(1);               // NL(1) = 32 * 2^58
  if (q > top / len) return 0;                       // q len > NL(1)
  const uint64_t need = q * len;
  uint64_t lo = 1, hi = 1ull << 32;                  // NL(lo) >= need > NL(hi)
  if (ic_neg_log2(lo) < need) return 0;
  while (hi - lo > 1) {
    const uint64_t mid = lo + (hi - lo) / 2;
    if (ic_neg_log2(mid) >= need)
      lo = mid;
    else
      hi = mid;
  }
  return static_cast<uint32_t>(lo);
}

// The same for a full block of 2^beta edges.
__host__ __device__ inline uint32_t ic_full_skip(uint64_t q, uint32_t beta) {
  return ic_skip_len(q, 1u << beta);
}

// floor(2^64 / q) for q >= 2 (0 for q = 0): the reciprocal ic_gap_fast
// divides by.  Every row has q = 0 or q >= NL(2^32 - 2^8) > 2^34.
__host__ __device__ inline uint64_t ic_gap_recip(uint64_t q) {
  if (q < 2) return 0;
  const uint64_t r0 = ~0ull / q;                     // floor((2^64 - 1) / q)
  const uint64_t rem = ~0ull - r0 * q;               // (2^64 - 1) mod q
  return rem == q - 1 ? r0 + 1 : r0;                 // q divides 2^64
}

// ic_gap(x, q) with the division done as a multiplication by rq =
// ic_gap_recip(q): with N = NL(U) < 2^64, N rq / 2^64 lies in (N/q - 1, N/q],
// so its floor is floor(N/q) or one less, and one compare settles which;
// (g + 1) q <= N + q stays below 2^64 since N, q <= 2^63.
__host__ __device__ inline uint64_t ic_gap_fast(uint64_t x, uint64_t q, uint64_t rq,
                                                const IcLogTables *t) {
  if (q == 0) return 0;
  const uint64_t N = ic_neg_log2_tab((x >> 32) + 1, t->log2c, t->recipc);
  uint64_t g = mulhi_u64(N, rq);
  if ((g + 1) * q <= N) ++g;
  return g;
}

}  // namespace rip
