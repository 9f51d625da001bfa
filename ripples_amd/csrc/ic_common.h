// Device helpers shared by the IC samplers (wave, lane and large-set
// kernels): vertex hashing, wave-wide bitonic sort, the table of stream
// jumps between gap blocks and the work counters of a launch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.h"
#include "device_util.h"
#include "internal.h"

namespace rip {

constexpr uint32_t kEmptyVertex = 0xFFFFFFFFu;
constexpr uint32_t kLenFail = 0xFFFFFFFFu;

__device__ inline uint32_t hash_vertex(uint32_t v) {
  v ^= v >> 16;
  v *= 0x7feb352du;
  v ^= v >> 15;
  v *= 0x846ca68bu;
  v ^= v >> 16;
  return v;
}

// Bitonic sort of a[0, len) (len <= cap2, a power of two) padded with
// kEmptyVertex; one wave, lanes over compare-exchange pairs.
template <typename Ptr>
__device__ inline void wave_bitonic(Ptr a, uint32_t len, uint32_t cap2) {
  const uint32_t lane = lane_id();
  for (uint32_t i = len + lane; i < cap2; i += 64) a[i] = kEmptyVertex;
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  for (uint32_t k = 2; k <= cap2; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t t = lane; t < cap2 / 2; t += 64) {
        // pair (i, i ^ j) with i having bit j clear
        const uint32_t i = ((t & ~(j - 1)) << 1) | (t & (j - 1));
        const uint32_t p = i | j;
        const uint32_t x = a[i], y = a[p];
        const bool up = (i & k) == 0;
        if ((x > y) == up) {
          a[i] = y;
          a[p] = x;
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
  }
}

// Algorithmic bytes of the gap samplers: the 64-byte record of every row
// expanded, the threshold of every candidate edge of a row whose edges do
// not share one (4), the in-neighbour of every live edge (4) and every
// output id (4).
constexpr uint32_t kIcRecBytes = 64;

// Adds a wave's work counters to the launch's (algorithmic bytes, candidate
// edges, draws, row expansions, live edges; each lane passes its own counts).
__device__ inline void report_stats(unsigned long long *ctl, unsigned long long bytes,
                                    unsigned long long cand, unsigned long long draws,
                                    unsigned long long exp, unsigned long long live) {
  for (int d = 32; d > 0; d >>= 1) {
    bytes += __shfl_down(bytes, d, 64);
    cand += __shfl_down(cand, d, 64);
    draws += __shfl_down(draws, d, 64);
    exp += __shfl_down(exp, d, 64);
    live += __shfl_down(live, d, 64);
  }
  if (lane_id() == 0) {
    if (bytes) atomicAdd(&ctl[kCtlStatBytes], bytes);
    if (cand) atomicAdd(&ctl[kCtlStatEdges], cand);
    if (draws) atomicAdd(&ctl[kCtlStatCoins], draws);
    if (exp) atomicAdd(&ctl[kCtlStatExpansions], exp);
    if (live) atomicAdd(&ctl[kCtlStatLevels], live);
  }
}

// P^(2^j) for j < 64 into p2 (one thread; the caller synchronises), so any
// jump below 2^64 elements is a product of entries.
__device__ inline void fill_step_powers(Affine step, Affine *p2) {
  p2[0] = step;
  for (int j = 1; j < 64; ++j) p2[j] = aff_then(p2[j - 1], p2[j - 1]);
}

}  // namespace rip
