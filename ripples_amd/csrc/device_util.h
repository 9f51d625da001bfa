// Wave64 / workgroup helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rip {

__device__ inline uint32_t lane_id() { return __lane_id(); }

// Values every lane of the wave holds alike, moved to scalar registers so
// that branches on them are scalar branches.
__device__ inline uint32_t uniform_u32(uint32_t x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ inline uint64_t uniform_u64(uint64_t x) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x));
  const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(x >> 32));
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

__device__ inline uint64_t lanes_below() {
  return (1ull << lane_id()) - 1ull;
}

// Inclusive scan across the 64 lanes of a wave.
__device__ inline uint32_t wave_inclusive_scan(uint32_t x) {
  const uint32_t lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t y = __shfl_up(x, d, 64);
    if (lane >= static_cast<uint32_t>(d)) x += y;
  }
  return x;
}

__device__ inline uint32_t wave_sum(uint32_t x) {
  return __shfl(wave_inclusive_scan(x), 63, 64);
}

// Exclusive scan over a workgroup of BT threads.  Every thread must call it.
// s_tmp needs BT/64 words.  Returns the prefix; *total gets the sum.
template <int BT>
__device__ inline uint32_t block_exclusive_scan(uint32_t v, uint32_t *s_tmp,
                                                uint32_t *total) {
  constexpr int NW = BT / 64;
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x >> 6;
  uint32_t x = wave_inclusive_scan(v);
  if (lane == 63) s_tmp[wid] = x;
  __syncthreads();
  if (wid == 0) {
    uint32_t w = lane < NW ? s_tmp[lane] : 0;
    w = wave_inclusive_scan(w);
    if (lane < NW) s_tmp[lane] = w;
  }
  __syncthreads();
  uint32_t off = wid ? s_tmp[wid - 1] : 0;
  *total = s_tmp[NW - 1];
  __syncthreads();
  return off + x - v;
}

// The same in 64 bits (scans whose tile sums may pass 2^32).
template <int BT>
__device__ inline unsigned long long block_exclusive_scan_u64(unsigned long long v,
                                                              unsigned long long *s_tmp,
                                                              unsigned long long *total) {
  constexpr int NW = BT / 64;
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x >> 6;
  unsigned long long x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const unsigned long long y = __shfl_up(x, d, 64);
    if (lane >= static_cast<uint32_t>(d)) x += y;
  }
  if (lane == 63) s_tmp[wid] = x;
  __syncthreads();
  if (wid == 0) {
    unsigned long long w = lane < NW ? s_tmp[lane] : 0ull;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const unsigned long long y = __shfl_up(w, d, 64);
      if (lane >= static_cast<uint32_t>(d)) w += y;
    }
    if (lane < NW) s_tmp[lane] = w;
  }
  __syncthreads();
  const unsigned long long off = wid ? s_tmp[wid - 1] : 0ull;
  *total = s_tmp[NW - 1];
  __syncthreads();
  return off + x - v;
}

// Append `val` to out[] for every lane with pred set; one LDS atomic per wave.
// All active lanes of the wave must call it together.
__device__ inline void wave_append(bool pred, uint32_t val, uint32_t *s_count,
                                   uint32_t *out) {
  uint64_t m = __ballot(pred);
  if (m == 0) return;
  const uint32_t leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  uint32_t base = 0;
  if (lane_id() == leader) base = atomicAdd(s_count, static_cast<uint32_t>(__popcll(m)));
  base = __shfl(base, static_cast<int>(leader), 64);
  if (pred) out[base + __popcll(m & lanes_below())] = val;
}

}  // namespace rip
