// Breakpoint form of the LT threshold scan.
//
// The sequential rule (include/ripples/add_rrrset.h:136-150) is
//     t <- fl(t - w_i)  for i = begin, begin + 1, ...
// stopping at the first i with !(t > 0), with t and w_i float.  The sampler's
// thresholds are u = k * 2^-24 with k in [1, 2^24], and the chain is
// monotone in its start (fl(t - w) is non-decreasing in t), so the edge at
// which it crosses zero is a non-decreasing function of k.  Row breakpoints
//     B[j] = min { k : the chain from k * 2^-24 stays > 0 through edge j }
// (kLtNever when even k = 2^24 does not) are therefore non-decreasing, and the
// crossing edge for k is the first j with B[j] > k (none when every
// B[j] <= k): a binary search replaces the subtraction loop and gives the same
// edge for every k.  Rows with negative, infinite or NaN weights keep the
// subtraction loop.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace rip {

constexpr uint32_t kLtNever = (1u << 24) + 1;

// Runs the float chain of the row's weights `w` from u = k * 2^-24 through
// edge j.  Returns whether it survives; `slack` is the value after edge j
// when it does, and w_i - t_i at the crossing edge i when it does not.
// Weights are read kChainGroup at a time so the loads of a group are in
// flight together (the chain itself is a dependent sequence of subtractions,
// so the group size sets how many memory round trips a long row costs).
constexpr uint32_t kChainGroup = 32;

__host__ __device__ inline bool lt_chain_eval(const uint32_t *w, uint32_t j, uint32_t k,
                                              double &slack) {
  float t = static_cast<float>(k) * 0x1p-24f;
  for (uint32_t i0 = 0; i0 <= j; i0 += kChainGroup) {
    const uint32_t cnt = j - i0 + 1 < kChainGroup ? j - i0 + 1 : kChainGroup;
    uint32_t wb[kChainGroup];
#pragma unroll
    for (uint32_t u = 0; u < kChainGroup; ++u) wb[u] = u < cnt ? w[i0 + u] : 0u;
#pragma unroll
    for (uint32_t u = 0; u < kChainGroup; ++u) {
      if (u >= cnt) break;
      float wi;
      __builtin_memcpy(&wi, &wb[u], 4);
      const float nt = t - wi;
      if (!(nt > 0)) {
        slack = static_cast<double>(wi) - static_cast<double>(t);
        return false;
      }
      t = nt;
    }
  }
  slack = static_cast<double>(t);
  return true;
}

// Next candidate of the breakpoint search from the last evaluation, kept in
// [lo + 1, hi - 1]: where the chain value at edge j is estimated to reach zero
// (one unit of k moves it by 2^-24 along an unchanged rounding path), or the
// middle of the bracket once the estimates have had a few rounds.
__host__ __device__ inline uint32_t lt_bp_next(uint32_t lo, uint32_t hi, uint32_t at, bool ok,
                                               double slack, int round) {
  double g;
  if (round >= 4 || slack != slack) {
    g = static_cast<double>(lo) + static_cast<double>(hi - lo) * 0.5;
  } else if (ok) {
    g = floor(static_cast<double>(at) - slack * 16777216.0) + 1.0;
  } else {
    g = floor(static_cast<double>(at) + slack * 16777216.0) + 1.0;
  }
  const double glo = static_cast<double>(lo) + 1.0, ghi = static_cast<double>(hi) - 1.0;
  g = g < glo ? glo : (g > ghi ? ghi : g);
  return static_cast<uint32_t>(g);
}

// B[j] for the row whose weights start at w.  The bracket (lo, hi] holds
// B[j]: the chain fails from lo (k = 0 stands for "no start") and survives
// from hi (kLtNever stands for "never").  The first two rounds evaluate one
// candidate each (the top k = 2^24, then the estimate from its chain value,
// which is typically tens of units off on long rows); later candidates are
// evaluated in pairs (k, k - 1), so an exact estimate closes the bracket in
// one round.  Every round shrinks the bracket, and after four rounds of
// estimates it is bisected.
__host__ __device__ inline uint32_t lt_breakpoint(const uint32_t *w, uint32_t j) {
  constexpr int kPairFrom = 2;
  uint32_t lo = 0, hi = kLtNever;
  uint32_t k = 1u << 24;
  for (int round = 0;; ++round) {
    double s1 = 0.0, s2 = 0.0;
    const bool a = lt_chain_eval(w, j, k, s1);
    const bool b = a && round >= kPairFrom && k - 1 > lo && lt_chain_eval(w, j, k - 1, s2);
    uint32_t at;
    bool ok;
    double slack;
    if (b) {
      hi = k - 1;
      at = k - 1, ok = true, slack = s2;
    } else if (a) {
      if (round >= kPairFrom) lo = k - 1;  // k - 1 was evaluated and failed
      hi = k;
      at = k, ok = true, slack = s1;
    } else {
      lo = k;
      at = k, ok = false, slack = s1;
    }
    if (hi - lo <= 1) return hi;
    k = lt_bp_next(lo, hi, at, ok, slack, round);
  }
}

// Crossing edge of threshold k from the row's breakpoints (d = row length):
// the first j with B[j] > k, or d when there is none.
__host__ __device__ inline uint32_t lt_bp_cross(const uint32_t *B, uint32_t d, uint32_t k) {
  uint32_t lo = 0, hi = d;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (B[mid] > k)
      hi = mid;
    else
      lo = mid + 1;
  }
  return lo;
}

// ------------------------------------------------------------ row index
// Rows longer than 16 edges carry an index over their breakpoints: level 0
// holds the last breakpoint of each block of 16, level l + 1 the last entry
// of each block of 16 of level l, up to a top level of at most 16 entries.
// Levels are stored top first, each padded to a multiple of 16 entries with
// 0xFFFFFFFF (above every threshold), so a search reads one aligned 64-byte
// node per level and a last block of at most 16 breakpoints.
constexpr uint32_t kLtFan = 16;
// Blocks of 16 breakpoints and their 16 in-neighbours after the last row
// of the graph (a search of an empty row reads one).
constexpr uint64_t kBpPad = 32;

__host__ __device__ inline uint32_t lt_pad16(uint32_t x) { return (x + 15u) & ~15u; }

// Number of levels of a row of d edges (0 when d <= kLtFan) and their sizes,
// sz[0] = level 0.
__host__ __device__ inline int lt_index_levels(uint32_t d, uint32_t *sz) {
  if (d <= kLtFan) return 0;
  int h = 0;
  sz[0] = (d + 15u) >> 4;
  while (sz[h] > kLtFan) {
    sz[h + 1] = (sz[h] + 15u) >> 4;
    ++h;
  }
  return h + 1;
}

__host__ __device__ inline uint32_t lt_bp_at(const uint32_t *B, uint32_t j) { return B[j]; }
__host__ __device__ inline uint32_t lt_bp_at(uint32_t *B, uint32_t j) { return B[j]; }
template <typename F>
__host__ __device__ inline uint32_t lt_bp_at(const F &B, uint32_t j) { return B(j); }

// Index entries (with padding) of a row of d edges.
__host__ __device__ inline uint32_t lt_index_size(uint32_t d) {
  uint32_t sz[8];
  const int nl = lt_index_levels(d, sz);
  uint32_t total = 0;
  for (int l = 0; l < nl; ++l) total += lt_pad16(sz[l]);
  return total;
}

// Writes the index of row breakpoints B(0) .. B(d - 1) at idx
// (lt_index_size(d) entries, top level first); B is an array or a callable
// that returns breakpoint j.
template <typename Bp>
__host__ __device__ inline void lt_index_build(const Bp &B, uint32_t d, uint32_t *idx) {
  uint32_t sz[8];
  const int nl = lt_index_levels(d, sz);
  if (nl == 0) return;
  uint32_t start[8];
  uint32_t at = 0;
  for (int l = nl - 1; l >= 0; --l) {
    start[l] = at;
    at += lt_pad16(sz[l]);
  }
  for (int l = 0; l < nl; ++l) {
    const uint32_t below = l == 0 ? d : sz[l - 1];
    for (uint32_t b = 0; b < lt_pad16(sz[l]); ++b) {
      uint32_t v = 0xFFFFFFFFu;
      if (b < sz[l]) {
        const uint32_t last = (16u * b + 15u < below) ? 16u * b + 15u : below - 1u;
        v = l == 0 ? lt_bp_at(B, last) : idx[start[l - 1] + last];
      }
      idx[start[l] + b] = v;
    }
  }
}

// Crossing edge of threshold k (first j with B[j] > k, or d) through the
// index: one node of 16 entries per level, then one block of B.  `probes`
// counts the nodes and blocks read.
__host__ __device__ inline uint32_t lt_index_cross(const uint32_t *B, uint32_t d,
                                                   const uint32_t *idx, uint32_t k,
                                                   unsigned long long *probes) {
  uint32_t sz[8];
  const int nl = lt_index_levels(d, sz);
  uint32_t node = 0;
  if (nl > 0) {
    uint32_t start = 0;  // the top level comes first
    for (int l = nl - 1; l >= 0; --l) {
      const uint32_t *e = idx + start + 16u * node;
      uint32_t p = 0;
#pragma unroll
      for (int i = 0; i < 16; ++i) p += e[i] <= k ? 1u : 0u;
      ++*probes;
      const uint32_t child = 16u * node + p;
      if (child >= sz[l]) return d;  // every breakpoint is <= k
      node = child;
      start += lt_pad16(sz[l]);
    }
  }
  const uint32_t lo = 16u * node;
  uint32_t c = 0;
#pragma unroll
  for (uint32_t i = 0; i < 16; ++i) c += (lo + i < d && B[lo + i] <= k) ? 1u : 0u;
  ++*probes;
  return lo + c;
}

// Weights without a breakpoint form: sign bit set (other than -0) or not
// finite.  (The chain stays monotone for them, but the search's estimates
// assume a decreasing chain of finite values.)
__host__ __device__ inline bool lt_weight_irregular(uint32_t wbits) {
  return ((wbits >> 23) & 0xFFu) == 0xFFu || (wbits > 0x80000000u);
}

}  // namespace rip
