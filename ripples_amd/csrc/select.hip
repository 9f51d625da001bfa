// Greedy max-coverage seed selection on the device.
//
// Same result as FindMostInfluentialSet (include/ripples/find_most_influential.h:85-174,
// streaming_find_most_influential.h:607-643) with lowest-id ties as
// GPUMaxElement (src/gpu/find_most_influential.cc:65-72):
//   1. counts[v] = number of sets containing v          (counting.h:65-97)
//   2. inverted index vertex -> sets by counting sort
//   3. k times: argmax of counts (key = count<<32 | ~v, so the largest key is
//      the largest count with the lowest id), then a kill pass that visits
//      only the sets containing the new seed, marks the uncovered ones
//      covered and decrements the counts of their members.
// Every set is killed at most once, so the kill passes cost O(T) over all k
// seeds, where the reference recounts every uncovered set per seed
// (gpu/find_most_influential.h:98-120).  The 2k kernels of the seed loop run
// as one HIP graph with no host synchronisation; the host reads the k keys
// once.
#include <algorithm>
#include <vector>

#include "device_util.h"
#include "internal.h"

namespace rip {
namespace {

constexpr uint32_t kBT = 256;
constexpr uint32_t kArgmaxBlocks = 1024;  // most blocks of an argmax pass

// Each lane of a wave holds one run (off, len) of `data`; the wave visits
// every element of the 64 runs once, lanes taking elements of any run:
// f(value, run index).
template <typename F>
__device__ inline void for_each_run_element(unsigned long long off, uint32_t len,
                                            const uint32_t *data, uint32_t *s_pre,
                                            unsigned long long *s_off, F &&f) {
  const uint32_t lane = lane_id();
  s_off[lane] = off;
  const uint32_t incl = wave_inclusive_scan(len);
  s_pre[lane + 1] = incl;
  if (lane == 0) s_pre[0] = 0;
  const uint32_t total = __shfl(incl, 63, 64);
  __builtin_amdgcn_wave_barrier();
  for (uint32_t i = lane; i < total; i += 64) {
    uint32_t lo = 0, hi = 64;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_pre[mid + 1] <= i) lo = mid + 1;
      else hi = mid;
    }
    f(data[s_off[lo] + (i - s_pre[lo])], lo);
  }
  __builtin_amdgcn_wave_barrier();
}

// Each wave walks 64 consecutive sets; lanes take elements of any of them.
template <typename F>
__device__ inline void for_each_element(const unsigned long long *set_off,
                                        const uint32_t *set_len,
                                        const uint32_t *data, uint64_t num_sets,
                                        uint32_t (*s_pre)[65],
                                        unsigned long long (*s_off)[64], F &&f) {
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kBT / 64);
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * (kBT / 64) + wid;
       w * 64 < num_sets; w += nwaves) {
    const uint64_t s = w * 64 + lane;
    const uint32_t len = s < num_sets ? set_len[s] : 0u;
    const unsigned long long off = s < num_sets ? set_off[s] : 0ull;
    for_each_run_element(off, len, data, s_pre[wid], s_off[wid],
                         [&](uint32_t v, uint32_t j) { f(v, w * 64 + j); });
  }
}

__global__ void __launch_bounds__(kBT)
count_kernel(const unsigned long long *set_off, const uint32_t *set_len,
             const uint32_t *data, uint64_t num_sets, uint32_t *counts) {
  __shared__ uint32_t s_pre[kBT / 64][65];
  __shared__ unsigned long long s_off[kBT / 64][64];
  for_each_element(set_off, set_len, data, num_sets, s_pre, s_off,
                   [&](uint32_t v, uint64_t) { atomicAdd(&counts[v], 1u); });
}

__global__ void __launch_bounds__(kBT)
fill_inverted_kernel(const unsigned long long *set_off, const uint32_t *set_len,
                     const uint32_t *data, uint64_t num_sets,
                     unsigned long long *cursor, uint32_t *inv) {
  __shared__ uint32_t s_pre[kBT / 64][65];
  __shared__ unsigned long long s_off[kBT / 64][64];
  for_each_element(set_off, set_len, data, num_sets, s_pre, s_off,
                   [&](uint32_t v, uint64_t s) {
                     const unsigned long long p = atomicAdd(&cursor[v], 1ull);
                     inv[p] = static_cast<uint32_t>(s);
                   });
}

// ------------------------------------------------------------ sorted index
// Sparse pools of kSortMinPairs to kSortMaxPairs elements build the inverted
// index by sorting their (vertex, set) pairs by vertex: no atomic per
// element, where the count and fill passes above make one each (and the hub
// vertices of R-MAT graphs share their counters).  Below the minimum the
// sort's fixed costs (its buffers, a scan) outweigh the atomics: IMM on
// R-MAT scale 20 LT (pools of up to 3.5M elements) took 19 ms with the sort
// against 16 ms without; at 60M elements the index takes 2.1 ms against 13.9.
constexpr uint64_t kSortMinPairs = 1ull << 23;
constexpr uint64_t kSortMaxPairs = 1ull << 31;

// The pairs of sets [0, num_sets) in set order: pair soff[s] + j is
// (member j of s, s).  A wave takes 64 consecutive sets, whose pairs are
// consecutive from soff of the first.
__global__ void __launch_bounds__(kBT)
pairs_kernel(const unsigned long long *set_off, const uint32_t *set_len, const uint32_t *data,
             uint64_t num_sets, const unsigned long long *soff, uint32_t *keys, uint32_t *vals) {
  __shared__ uint32_t s_pre[kBT / 64][65];
  __shared__ unsigned long long s_off[kBT / 64][64];
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kBT / 64);
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * (kBT / 64) + wid; w * 64 < num_sets;
       w += nwaves) {
    const uint64_t s = w * 64 + lane;
    const uint32_t len = s < num_sets ? set_len[s] : 0u;
    s_off[wid][lane] = s < num_sets ? set_off[s] : 0ull;
    const uint32_t incl = wave_inclusive_scan(len);
    s_pre[wid][lane + 1] = incl;
    if (lane == 0) s_pre[wid][0] = 0;
    const uint32_t total = __shfl(incl, 63, 64);
    const unsigned long long base = soff[w * 64];
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < total; i += 64) {
      uint32_t lo = 0, hi = 64;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pre[wid][mid + 1] <= i) lo = mid + 1;
        else hi = mid;
      }
      keys[base + i] = data[s_off[wid][lo] + (i - s_pre[wid][lo])];
      vals[base + i] = static_cast<uint32_t>(w * 64 + lo);
    }
    __builtin_amdgcn_wave_barrier();
  }
}

// inv_start[v] = the first of the `count` sorted pairs whose vertex is >= v,
// for v in [0, n]: pair i starts the vertices after its predecessor's up to
// its own (i = count: up to n).
__global__ void starts_kernel(const uint32_t *keys, uint64_t count, uint32_t n,
                              unsigned long long *inv_start) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i <= count;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t hi = i < count ? keys[i] : n;
    const uint64_t lo = i == 0 ? 0ull : static_cast<uint64_t>(keys[i - 1]) + 1;
    for (uint64_t v = lo; v <= hi; ++v) inv_start[v] = i;
  }
}

__global__ void counts_from_starts_kernel(const unsigned long long *inv_start, uint32_t n,
                                          uint32_t *counts) {
  for (uint32_t v = blockIdx.x * blockDim.x + threadIdx.x; v < n; v += gridDim.x * blockDim.x)
    counts[v] = static_cast<uint32_t>(inv_start[v + 1] - inv_start[v]);
}

// ------------------------------------------------------------ dense pools
// Pools whose sets are long compared with the number of vertex ranges below
// build no inverted index.  Counts come from LDS histograms: block (set chunk
// c, vertex range r) visits the sets of chunk c and, in each, only the run of
// vertices inside range r, located by binary search (sets are sorted and
// duplicate-free), then adds its histogram to the global counts with one
// atomic per vertex.  A seed round finds the uncovered sets holding the seed
// by binary search in each of them, and subtracts the killed sets' members the
// same way (or with per-element atomics when the killed sets are small).
constexpr uint32_t kDenseBT = 512;
constexpr uint32_t kDenseRange = 16384;        // u32 counters: 64 KiB of LDS
constexpr uint32_t kDenseLong = 256;           // runs this long are read by the whole wave
constexpr uint32_t kDenseBlocks = 2048;        // blocks per histogram pass
constexpr unsigned long long kDenseDecMin = 1ull << 21;  // killed elements for LDS decrements
constexpr uint32_t kScanGrid = 1024;

// First i in [0, len) with a[i] >= x (len if none).
__device__ inline uint32_t lower_bound_u32(const uint32_t *a, uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// The block's (chunk, range).  Blocks are dealt to the 8 XCDs round-robin;
// the remap gives each XCD a run of consecutive logical blocks, so the
// ranges of one chunk share an L2.
__device__ inline void dense_block(uint32_t ranges, uint32_t *c, uint32_t *r) {
  uint32_t b = blockIdx.x;
  const uint32_t g = gridDim.x;
  if ((g & 7u) == 0) b = (b & 7u) * (g >> 3) + (b >> 3);
  *c = b / ranges;
  *r = b % ranges;
}

// Lane j of the wave holds set j of a group (its data at o, length L; live
// false for no set).  f(v) for every member v of the 64 sets with
// v_lo <= v < v_lo + width.  Runs of kDenseLong or more are read by the whole
// wave with eight loads in flight per lane, the rest lane-mapped.
template <typename F>
__device__ inline void dense_group(unsigned long long o, uint32_t L, bool live,
                                   const uint32_t *data, uint32_t v_lo, uint32_t width,
                                   uint32_t *s_pre, unsigned long long *s_off, F &&f) {
  unsigned long long off = 0;
  uint32_t len = 0;
  if (live) {
    const uint32_t b = lower_bound_u32(data + o, L, v_lo);
    // at most `width` distinct vertices fall in the range
    const uint32_t span = L - b < width ? L - b : width;
    off = o + b;
    len = lower_bound_u32(data + off, span, v_lo + width);
  }
  const uint32_t lane = lane_id();
  uint64_t longs = __ballot(len >= kDenseLong);
  while (longs) {
    const int j = __builtin_ctzll(longs);
    longs &= longs - 1;
    const uint32_t lo32 = __builtin_amdgcn_readlane(static_cast<uint32_t>(off), j);
    const uint32_t hi32 = __builtin_amdgcn_readlane(static_cast<uint32_t>(off >> 32), j);
    const uint32_t rl = __builtin_amdgcn_readlane(len, j);
    const uint32_t *run = data + ((static_cast<unsigned long long>(hi32) << 32) | lo32);
    for (uint32_t q = lane; q < rl; q += 64 * 8) {
      uint32_t v[8];
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u) v[u] = q + 64 * u < rl ? run[q + 64 * u] : 0u;
#pragma unroll
      for (uint32_t u = 0; u < 8; ++u)
        if (q + 64 * u < rl) f(v[u]);
    }
  }
  for_each_run_element(off, len >= kDenseLong ? 0u : len, data, s_pre, s_off,
                       [&](uint32_t v, uint32_t) { f(v); });
}

// counts[v] += (kSub: -=) the number of sets holding v among sets [0, total)
// (list == nullptr) or list[0, total).  Grid: chunks x ranges blocks.
template <bool kSub>
__device__ inline void dense_hist(const unsigned long long *set_off, const uint32_t *set_len,
                                  const uint32_t *data, const uint32_t *list, uint64_t total,
                                  uint32_t n, uint32_t ranges, uint32_t *counts) {
  __shared__ uint32_t hist[kDenseRange];
  __shared__ uint32_t s_pre[kDenseBT / 64][65];
  __shared__ unsigned long long s_off[kDenseBT / 64][64];
  const uint32_t chunks = gridDim.x / ranges;
  uint32_t c, r;
  dense_block(ranges, &c, &r);
  const uint64_t per = (total + chunks - 1) / chunks;
  const uint64_t sb = c * per;
  const uint64_t se = sb + per < total ? sb + per : total;
  if (sb >= se) return;  // block-uniform
  const uint32_t v_lo = r * kDenseRange;
  const uint32_t width = n - v_lo < kDenseRange ? n - v_lo : kDenseRange;
  for (uint32_t i = threadIdx.x; i < width; i += kDenseBT) hist[i] = 0;
  __syncthreads();
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  for (uint64_t base = sb + static_cast<uint64_t>(wid) * 64; base < se; base += kDenseBT) {
    const uint64_t t = base + lane;
    const bool live = t < se;
    const uint64_t s = live ? (list ? list[t] : t) : 0;
    const unsigned long long o = live ? set_off[s] : 0ull;
    const uint32_t L = live ? set_len[s] : 0u;
    // the range test also keeps an unsorted set from indexing past LDS
    dense_group(o, L, live, data, v_lo, width, s_pre[wid], s_off[wid], [&](uint32_t v) {
      if (v - v_lo < width) atomicAdd(&hist[v - v_lo], 1u);
    });
  }
  __syncthreads();
  for (uint32_t i = threadIdx.x; i < width; i += kDenseBT) {
    const uint32_t h = hist[i];
    if (h == 0) continue;
    if (kSub) atomicSub(&counts[v_lo + i], h);
    else atomicAdd(&counts[v_lo + i], h);
  }
}

__global__ void __launch_bounds__(kDenseBT)
dense_count_kernel(const unsigned long long *set_off, const uint32_t *set_len,
                   const uint32_t *data, uint64_t num_sets, uint32_t n, uint32_t ranges,
                   uint32_t *counts) {
  dense_hist<false>(set_off, set_len, data, nullptr, num_sets, n, ranges, counts);
}

// Seed round, dense pools, last step for large kills: subtract the members of
// the killed sets through LDS histograms.
__global__ void __launch_bounds__(kDenseBT) dense_dec_kernel(const Args *a, uint32_t ranges) {
  if (a->ctl[kCtlKillElems] <= kDenseDecMin) return;
  dense_hist<true>(a->set_off, a->set_len, a->data, a->kill, a->ctl[kCtlKillSets], a->n,
                   ranges, a->counts);
}

__global__ void fill_u32_kernel(uint32_t *p, uint64_t n, uint32_t v) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}

__global__ void fill_u8_kernel(uint8_t *p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

__global__ void copy_u64_kernel(unsigned long long *dst,
                                const unsigned long long *src, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    dst[i] = src[i];
}

__global__ void zero_word_kernel(unsigned long long *w) { *w = 0; }

__global__ void clear_keys_kernel(const Args *a, uint32_t k) {
  for (uint32_t i = threadIdx.x; i < k; i += blockDim.x) a->keys[i] = 0;
  if (threadIdx.x == 0) a->ctl[kCtlCovered] = 0;
}

__device__ inline unsigned long long wave_max_u64(unsigned long long x) {
  for (int d = 32; d > 0; d >>= 1) {
    const unsigned long long y = __shfl_xor(x, d, 64);
    x = y > x ? y : x;
  }
  return x;
}

__device__ inline unsigned long long block_max_u64(unsigned long long x,
                                                   unsigned long long *s_w) {
  x = wave_max_u64(x);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = x;
  __syncthreads();
  x = s_w[0];
  for (uint32_t w = 1; w < kBT / 64; ++w) x = s_w[w] > x ? s_w[w] : x;
  return x;
}

// Seed round, first half: partial[b] = max over block b's vertices v of
// counts[v] << 32 | ~v.  The maxima are reduced by the second half, where
// one atomicMax per block on a single key measured as a serial tail.
__global__ void __launch_bounds__(kBT) argmax_kernel(const Args *a) {
  __shared__ unsigned long long s_w[kBT / 64];
  const uint32_t *counts = a->counts;
  const uint32_t n = a->n;
  unsigned long long best = 0;
  for (uint32_t v = blockIdx.x * kBT + threadIdx.x; v < n; v += gridDim.x * kBT) {
    const unsigned long long key =
        (static_cast<unsigned long long>(counts[v]) << 32) | (0xFFFFFFFFu - v);
    best = key > best ? key : best;
  }
  best = block_max_u64(best, s_w);
  if (threadIdx.x == 0) {
    a->partial[blockIdx.x] = best;
    if (blockIdx.x == 0) {  // the dense pools' kill list of this round
      a->ctl[kCtlKillSets] = 0;
      a->ctl[kCtlKillElems] = 0;
      a->ctl[kCtlKillGroups] = 0;
    }
  }
}

// Seed round i, second half, first step in every block: the round's key from
// the argmax blocks' maxima; block 0 records it in keys[i] (0 when no vertex
// covers an uncovered set).
__device__ inline unsigned long long round_key(const Args *a, uint32_t i, uint32_t parts) {
  __shared__ unsigned long long s_w[kBT / 64];
  unsigned long long best = 0;
  for (uint32_t b = threadIdx.x; b < parts; b += kBT)
    best = a->partial[b] > best ? a->partial[b] : best;
  best = block_max_u64(best, s_w);
  if (best >> 32 == 0) best = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) a->keys[i] = best;
  return best;
}

// Seed round i, second half: every uncovered set containing the seed is
// marked covered once and its members' counts drop by one.
__global__ void __launch_bounds__(kBT) kill_kernel(const Args *a, uint32_t i, uint32_t parts) {
  const unsigned long long key = round_key(a, i, parts);
  const uint32_t gain = static_cast<uint32_t>(key >> 32);
  if (gain == 0) return;
  const uint32_t seed = 0xFFFFFFFFu - static_cast<uint32_t>(key);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a->ctl[kCtlCovered], gain);
  const uint32_t lane = lane_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kBT / 64);
  const unsigned long long *inv_start = a->inv_start;
  const uint32_t *inv = a->inv;
  const unsigned long long *set_off = a->set_off;
  const uint32_t *set_len = a->set_len;
  const uint32_t *data = a->data;
  uint8_t *covered = a->covered;
  uint32_t *counts = a->counts;
  const unsigned long long b = inv_start[seed], e = inv_start[seed + 1];
  for (unsigned long long p = b + blockIdx.x * (kBT / 64) + (threadIdx.x >> 6); p < e;
       p += nwaves) {
    const uint32_t s = inv[p];
    if (covered[s]) continue;  // wave-uniform
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) covered[s] = 1;
    const unsigned long long off = set_off[s];
    const uint32_t len = set_len[s];
    for (uint32_t q = lane; q < len; q += 64) atomicSub(&counts[data[off + q]], 1u);
  }
}

// Seed round i, dense pools, first step: every uncovered set holding the
// seed (binary search; sets are sorted) is marked covered and appended to the
// kill list; the killed elements are summed.
__global__ void __launch_bounds__(kBT) scan_kill_kernel(const Args *a, uint32_t i,
                                                        uint32_t parts) {
  const unsigned long long key = round_key(a, i, parts);
  const uint32_t gain = static_cast<uint32_t>(key >> 32);
  if (gain == 0) return;
  const uint32_t seed = 0xFFFFFFFFu - static_cast<uint32_t>(key);
  if (blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(&a->ctl[kCtlCovered], gain);
  const uint64_t num_sets = a->num_sets;
  const uint8_t *covered_in = a->covered;
  uint8_t *covered = a->covered;
  const unsigned long long *set_off = a->set_off;
  const uint32_t *set_len = a->set_len;
  const uint32_t *data = a->data;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBT;
  const uint64_t lane = lane_id();
  // column store: the seed's place among the hot vertices
  const uint32_t seed_hot = a->groups ? a->hot_index[seed] : kNotHot;
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * kBT + threadIdx.x - lane; w < num_sets;
       w += stride) {
    const uint64_t s = w + lane;
    bool hit = false;
    uint32_t len = 0;
    if (s < num_sets && !covered_in[s]) {
      const unsigned long long off = set_off[s];
      len = set_len[s];
      const uint32_t pos = lower_bound_u32(data + off, len, seed);
      hit = pos < len && data[off + pos] == seed;
      const uint32_t e = a->groups ? a->gref[s] : kNoGroup;
      if (e != kNoGroup) {
        // a set of the column store: its hot members are bit c of its
        // group's column; its hot members leave the counts in hot_dec_kernel
        if (!hit && seed_hot != kNotHot) hit = (a->colptr[e >> 6][seed_hot] >> (e & 63u)) & 1ull;
        if (hit) {
          const unsigned long long prev = atomicOr(&a->gkill[e >> 6], 1ull << (e & 63u));
          if (prev == 0) {
            const unsigned long long at = atomicAdd(&a->ctl[kCtlKillGroups], 1ull);
            a->kgroups[at] = e >> 6;
          }
        }
      }
      if (hit) covered[s] = 1;
    }
    const uint64_t m = __ballot(hit);
    if (m == 0) continue;  // wave-uniform
    unsigned long long elems = hit ? len : 0u;
    for (int d = 32; d > 0; d >>= 1) elems += __shfl_xor(elems, d, 64);
    const uint32_t leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    unsigned long long base = 0;
    if (lane == leader) {
      base = atomicAdd(&a->ctl[kCtlKillSets], static_cast<unsigned long long>(__popcll(m)));
      atomicAdd(&a->ctl[kCtlKillElems], elems);
    }
    base = __shfl(base, static_cast<int>(leader), 64);
    if (hit) a->kill[base + __popcll(m & lanes_below())] = static_cast<uint32_t>(s);
  }
}

// Seed round, dense pools, last step for small kills: one atomic per element.
__global__ void __launch_bounds__(kBT) dec_small_kernel(const Args *a) {
  __shared__ uint32_t s_pre[kBT / 64][65];
  __shared__ unsigned long long s_off[kBT / 64][64];
  if (a->ctl[kCtlKillElems] > kDenseDecMin) return;
  const uint64_t cnt = a->ctl[kCtlKillSets];
  const uint32_t *kill = a->kill;
  const unsigned long long *set_off = a->set_off;
  const uint32_t *set_len = a->set_len;
  uint32_t *counts = a->counts;
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kBT / 64);
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * (kBT / 64) + wid; w * 64 < cnt;
       w += nwaves) {
    const uint64_t t = w * 64 + lane;
    const uint32_t s = t < cnt ? kill[t] : 0u;
    for_each_run_element(t < cnt ? set_off[s] : 0ull, t < cnt ? set_len[s] : 0u, a->data,
                         s_pre[wid], s_off[wid],
                         [&](uint32_t v, uint32_t) { atomicSub(&counts[v], 1u); });
  }
}

// ------------------------------------------------------------ column store
// Sets held in the column store keep their hot members as bit c of their
// group's column: word h of the column holds, for the group's sets, whether
// hot vertex h is a member.

// gactive[g] = the colours of group g whose sample id is below num_sets.
__global__ void group_active_kernel(const uint32_t *gsample, uint32_t groups, uint64_t num_sets,
                                    unsigned long long *gactive, unsigned long long *gkill) {
  for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < groups;
       g += gridDim.x * blockDim.x) {
    unsigned long long m = 0;
    for (uint32_t c = 0; c < 64; ++c) {
      const uint32_t sid = gsample[64ull * g + c];
      if (sid != kNoGroup && sid < num_sets) m |= 1ull << c;
    }
    gactive[g] = m;
    gkill[g] = 0;
  }
}

// counts[hot vertex h] += its memberships among the active column sets.
__global__ void __launch_bounds__(kBT) hot_count_kernel(unsigned long long *const *colptr,
                                                        const unsigned long long *gactive,
                                                        uint32_t groups, const uint32_t *hot_vertex,
                                                        uint32_t hot_n, uint32_t *counts) {
  for (uint32_t h = blockIdx.x * kBT + threadIdx.x; h < hot_n; h += gridDim.x * kBT) {
    uint32_t acc = 0;
    for (uint32_t g = 0; g < groups; ++g) acc += __popcll(colptr[g][h] & gactive[g]);
    if (acc) counts[hot_vertex[h]] += acc;
  }
}

// Seed round, column store: the hot members of the column sets killed this
// round leave the counts (one pass over the columns of the groups with
// kills; none in most rounds).
__global__ void __launch_bounds__(kBT) hot_dec_kernel(const Args *a) {
  const uint32_t nk = static_cast<uint32_t>(a->ctl[kCtlKillGroups]);
  if (nk == 0) return;
  const uint32_t *kg = a->kgroups;
  for (uint32_t h = blockIdx.x * kBT + threadIdx.x; h < a->hot_n; h += gridDim.x * kBT) {
    uint32_t d = 0;
    for (uint32_t i = 0; i < nk; ++i) {
      const uint32_t g = kg[i];
      d += __popcll(a->colptr[g][h] & a->gkill[g]);
    }
    if (d) a->counts[a->hot_vertex[h]] -= d;
  }
}

__global__ void clear_gkill_kernel(const Args *a) {
  const uint32_t nk = static_cast<uint32_t>(a->ctl[kCtlKillGroups]);
  for (uint32_t i = threadIdx.x; i < nk; i += blockDim.x) a->gkill[a->kgroups[i]] = 0;
}

// Coverage of the column sets not already hit through their lists: a hit
// when a flagged seed is a hot member.
__global__ void __launch_bounds__(kBT) coverage_hot_kernel(const uint32_t *gref,
                                                           unsigned long long *const *colptr,
                                                           const uint32_t *hot_index,
                                                           uint64_t num_sets, const uint32_t *seeds,
                                                           uint32_t num_seeds, uint8_t *hit) {
  for (uint64_t s = static_cast<uint64_t>(blockIdx.x) * kBT + threadIdx.x; s < num_sets;
       s += static_cast<uint64_t>(gridDim.x) * kBT) {
    const uint32_t e = gref[s];
    if (e == kNoGroup || hit[s]) continue;
    const unsigned long long *col = colptr[e >> 6];
    for (uint32_t j = 0; j < num_seeds; ++j) {
      const uint32_t h = hot_index[seeds[j]];
      if (h != kNotHot && ((col[h] >> (e & 63u)) & 1ull)) {
        hit[s] = 1;
        break;
      }
    }
  }
}

__global__ void __launch_bounds__(kBT) count_u8_kernel(const uint8_t *v, uint64_t n,
                                                       unsigned long long *out) {
  __shared__ unsigned long long s_w[kBT / 64];
  unsigned long long acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBT + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kBT)
    acc += v[i];
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    acc = 0;
    for (uint32_t w = 0; w < kBT / 64; ++w) acc += s_w[w];
    if (acc) atomicAdd(out, acc);
  }
}

// Sets [0, num_sets) that contain a flagged vertex (FindCoverage,
// include/ripples/opimc.h:100-118): each wave walks 64 sets, marks hits in
// LDS and adds their popcount to *out once per wave.  Per-set hits also go
// to hit_out (one byte per set) or hit_bits (word w = sets 64w .. 64w + 63:
// a wave's 64 sets are one word).
template <bool kBits>
__global__ void __launch_bounds__(kBT)
coverage_kernel(const unsigned long long *set_off, const uint32_t *set_len,
                const uint32_t *data, uint64_t num_sets, const uint32_t *flags,
                unsigned long long *out, uint8_t *hit_out = nullptr,
                unsigned long long *hit_bits = nullptr) {
  __shared__ uint32_t s_pre[kBT / 64][65];
  __shared__ unsigned long long s_off[kBT / 64][64];
  __shared__ uint32_t s_hit[kBT / 64][64];
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kBT / 64);
  unsigned long long acc = 0;
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * (kBT / 64) + wid; w * 64 < num_sets;
       w += nwaves) {
    const uint64_t s = w * 64 + lane;
    const uint32_t len = s < num_sets ? set_len[s] : 0u;
    s_off[wid][lane] = s < num_sets ? set_off[s] : 0ull;
    s_hit[wid][lane] = 0;
    const uint32_t incl = wave_inclusive_scan(len);
    s_pre[wid][lane + 1] = incl;
    if (lane == 0) s_pre[wid][0] = 0;
    const uint32_t total = __shfl(incl, 63, 64);
    __builtin_amdgcn_wave_barrier();
    for (uint32_t i = lane; i < total; i += 64) {
      uint32_t lo = 0, hi = 64;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_pre[wid][mid + 1] <= i) lo = mid + 1;
        else hi = mid;
      }
      const uint32_t v = data[s_off[wid][lo] + (i - s_pre[wid][lo])];
      if (kBits ? (flags[v >> 5] >> (v & 31)) & 1u : flags[v]) s_hit[wid][lo] = 1;
    }
    __builtin_amdgcn_wave_barrier();
    const uint64_t hm = __ballot(s_hit[wid][lane] != 0);
    acc += __popcll(hm);
    if (hit_out && s < num_sets) hit_out[s] = s_hit[wid][lane] != 0;
    if (hit_bits && lane == 0) hit_bits[w] = hm;
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0 && acc) atomicAdd(out, acc);
}

// Column sets not already hit through their lists, with per-set hit bits
// (word w = sets 64w .. 64w + 63, one wave per word): a hit when a seed is a
// hot member.
__global__ void __launch_bounds__(kBT) coverage_hot_bits_kernel(
    const uint32_t *gref, unsigned long long *const *colptr, const uint32_t *hot_index,
    uint64_t num_sets, const uint32_t *seeds, uint32_t num_seeds,
    unsigned long long *hit_bits) {
  const uint32_t lane = lane_id();
  const uint64_t nwaves = static_cast<uint64_t>(gridDim.x) * (kBT / 64);
  for (uint64_t w = static_cast<uint64_t>(blockIdx.x) * (kBT / 64) + (threadIdx.x >> 6);
       w * 64 < num_sets; w += nwaves) {
    const uint64_t s = w * 64 + lane;
    const unsigned long long word = hit_bits[w];
    bool hit = false;
    const uint32_t e = s < num_sets ? gref[s] : kNoGroup;
    if (e != kNoGroup && !((word >> lane) & 1ull)) {
      const unsigned long long *col = colptr[e >> 6];
      for (uint32_t j = 0; j < num_seeds && !hit; ++j) {
        const uint32_t h = hot_index[seeds[j]];
        hit = h != kNotHot && ((col[h] >> (e & 63u)) & 1ull);
      }
    }
    const uint64_t m = __ballot(hit);
    if (lane == 0 && m) hit_bits[w] = word | m;
  }
}

// *out += the set bits of words[0, n).
__global__ void __launch_bounds__(kBT) popcount_u64_kernel(const unsigned long long *words,
                                                           uint64_t n, unsigned long long *out) {
  __shared__ unsigned long long s_w[kBT / 64];
  unsigned long long acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBT + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * kBT)
    acc += __popcll(words[i]);
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    acc = 0;
    for (uint32_t w = 0; w < kBT / 64; ++w) acc += s_w[w];
    if (acc) atomicAdd(out, acc);
  }
}

// Sets the bit of every seed in a vertex bitmap.
__global__ void flag_seed_bits_kernel(const uint32_t *seeds, uint32_t k, uint32_t *bits) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
    atomicOr(&bits[seeds[i] >> 5], 1u << (seeds[i] & 31u));
}

__global__ void fill_u64_kernel(unsigned long long *p, uint64_t n) {
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < n; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = 0;
}

__global__ void flag_seeds_kernel(const uint32_t *seeds, uint32_t k, uint32_t *flags) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
    flags[seeds[i]] = 1;
}



// inv_elems: elements of the inverted index (0: none); kill_sets: length of
// the dense pools' kill list (0: none).
void ensure_select_scratch(rip_pool *p, uint64_t num_sets, uint64_t inv_elems,
                           uint64_t kill_sets, uint32_t k) {
  SelectScratch &s = p->sel;
  const uint32_t n = p->g->n;
  if (s.n != n) {
    dev_free(s.d_counts);
    dev_free(s.d_inv_start);
    dev_free(s.d_cursor);
    s.d_counts = dev_alloc<uint32_t>(n);
    s.d_inv_start = dev_alloc<unsigned long long>(static_cast<uint64_t>(n) + 1);
    s.d_cursor = dev_alloc<unsigned long long>(n);
    s.n = n;
  }
  if (s.cap_sets < num_sets) {
    dev_free(s.d_covered);
    dev_free(s.d_soff);
    s.cap_sets = std::max<uint64_t>(num_sets, s.cap_sets + s.cap_sets / 2);
    s.d_covered = dev_alloc<uint8_t>(s.cap_sets);
    s.d_soff = dev_alloc<unsigned long long>(s.cap_sets + 1);
  }
  if (s.cap_elems < inv_elems) {
    dev_free(s.d_inv);
    s.cap_elems = std::max<uint64_t>(inv_elems, s.cap_elems + s.cap_elems / 2);
    s.d_inv = dev_alloc<uint32_t>(s.cap_elems);
  }
  if (s.cap_kill < kill_sets) {
    dev_free(s.d_kill);
    s.cap_kill = std::max<uint64_t>(kill_sets, s.cap_kill + s.cap_kill / 2);
    s.d_kill = dev_alloc<uint32_t>(s.cap_kill);
  }
  if (s.cap_k < k) {
    dev_free(s.d_keys);
    s.cap_k = k;
    s.d_keys = dev_alloc<unsigned long long>(k);
  }
  if (s.cap_groups < p->groups) {
    dev_free(s.d_gactive);
    dev_free(s.d_gkill);
    dev_free(s.d_kgroups);
    s.cap_groups = 0;
    const uint32_t cap = std::max<uint32_t>(p->groups, p->cap_groups);
    s.d_gactive = dev_alloc<unsigned long long>(cap);
    s.d_gkill = dev_alloc<unsigned long long>(cap);
    s.d_kgroups = dev_alloc<uint32_t>(cap);
    s.cap_groups = cap;
  }
  if (!s.d_args) s.d_args = dev_alloc<Args>(1);
  if (!s.d_partial) s.d_partial = dev_alloc<unsigned long long>(kArgmaxBlocks);
}

uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap) {
  uint64_t g = (work + per_block - 1) / per_block;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

constexpr uint64_t kDenseMinElements = 1ull << 22;

uint32_t ranges_of(uint32_t n) { return (n + kDenseRange - 1) / kDenseRange; }

// Dense pools: at least 4M elements and a mean set length of at least 64
// per vertex range, since the range searches cost O(num_sets x ranges)
// loads where the per-element kernels make one global atomic per element.
bool is_dense(uint32_t n, uint64_t num_sets, uint64_t total) {
  if (num_sets == 0 || n == 0 || total < kDenseMinElements) return false;
  return total / num_sets >= 64ull * ranges_of(n);
}

// Blocks per vertex range of a histogram pass over `sets` sets.
uint32_t dense_chunks(uint32_t ranges, uint64_t sets) {
  const uint64_t want = (kDenseBlocks + ranges - 1) / ranges;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(want, sets / kDenseBT)));
}

enum class Need { Counts, Select };

// The inverted index of sparse sets [0, num_sets) (total elements) by sorting
// their (vertex, set) pairs: counts, inv_start and the index itself, in
// d_inv or d_pval; returns where the index is.
const uint32_t *inverted_by_sort(rip_pool *p, uint64_t num_sets, uint64_t total) {
  SelectScratch &s = p->sel;
  hipStream_t st = p->stream;
  const uint32_t n = p->g->n;
  if (s.cap_pairs < total) {
    const uint64_t cap = std::max<uint64_t>(total, s.cap_pairs + s.cap_pairs / 2);
    dev_free(s.d_pkey[0]);
    dev_free(s.d_pkey[1]);
    dev_free(s.d_pval);
    s.cap_pairs = 0;
    s.d_pkey[0] = dev_alloc<uint32_t>(cap);
    s.d_pkey[1] = dev_alloc<uint32_t>(cap);
    s.d_pval = dev_alloc<uint32_t>(cap);
    s.cap_pairs = cap;
  }
  exclusive_scan_u32(p->d_len, s.d_soff, num_sets, st);
  hipLaunchKernelGGL(pairs_kernel, dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)), dim3(kBT), 0,
                     st, p->d_off, p->d_len, p->d_data, num_sets, s.d_soff, s.d_pkey[0], s.d_inv);
  RIP_HIP(hipGetLastError());
  uint32_t bits = 1;
  while (bits < 32 && (static_cast<uint64_t>(n) - 1) >> bits) ++bits;
  const int at = sort_pairs_u32(s.d_pkey[0], s.d_pkey[1], s.d_inv, s.d_pval, total, bits,
                                &s.d_stemp, &s.stemp_bytes, st);
  hipLaunchKernelGGL(starts_kernel, dim3(grid_for(total + 1, kBT, 4096)), dim3(kBT), 0, st,
                     s.d_pkey[at], total, n, s.d_inv_start);
  hipLaunchKernelGGL(counts_from_starts_kernel, dim3(grid_for(n, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_inv_start, n, s.d_counts);
  RIP_HIP(hipGetLastError());
  return at == 0 ? s.d_inv : s.d_pval;
}

// counts over sets [0, num_sets) and the scratch a selection needs; returns
// whether the pool takes the dense path.  *sorted_inv (selections only): the
// inverted index when it was built here by sorting, else nullptr.
bool build_counts(rip_pool *p, uint64_t num_sets, uint32_t k, Need need,
                  const uint32_t **sorted_inv = nullptr) {
  if (sorted_inv) *sorted_inv = nullptr;
  RIP_REQUIRE(num_sets <= p->num_sets, -1, "select: num_sets beyond the pool");
  const uint32_t n = p->g->n;
  hipStream_t st = p->stream;
  RIP_HIP(hipSetDevice(p->g->device));
  // total element count
  hipLaunchKernelGGL(zero_word_kernel, dim3(1), dim3(1), 0, st, p->d_ctl + kCtlWords - 1);
  launch_sum_u32(p->d_len, num_sets, p->d_ctl + kCtlWords - 1, st);
  unsigned long long total = 0;
  RIP_HIP(hipMemcpyAsync(&total, p->d_ctl + kCtlWords - 1, sizeof(total),
                         hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  // pools with column sets take the dense kernels, whose seed rounds scan
  // every uncovered set and so also test the column sets' hot members
  const bool cols = p->groups > 0;
  const bool dense = cols ? num_sets > 0 : is_dense(n, num_sets, total);
  const bool sel = need == Need::Select;
  ensure_select_scratch(p, sel ? num_sets : 0, sel && !dense ? total : 0,
                        sel && dense ? num_sets : 0, k);
  SelectScratch &s = p->sel;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_counts, static_cast<uint64_t>(n), 0u);
  if (dense) {
    const uint32_t ranges = ranges_of(n);
    hipLaunchKernelGGL(dense_count_kernel, dim3(dense_chunks(ranges, num_sets) * ranges),
                       dim3(kDenseBT), 0, st, p->d_off, p->d_len, p->d_data, num_sets, n,
                       ranges, s.d_counts);
    if (cols) {
      hipLaunchKernelGGL(group_active_kernel, dim3(grid_for(p->groups, kBT, 1024)), dim3(kBT), 0,
                         st, p->d_gsample, p->groups, num_sets, s.d_gactive, s.d_gkill);
      hipLaunchKernelGGL(hot_count_kernel, dim3(grid_for(p->g->hot_n, kBT, 8192)), dim3(kBT), 0,
                         st, p->d_colptr, s.d_gactive, p->groups, p->g->d_hot_vertex,
                         p->g->hot_n, s.d_counts);
    }
  } else if (num_sets) {
    if (sel && sorted_inv && total >= kSortMinPairs && total <= kSortMaxPairs) {
      *sorted_inv = inverted_by_sort(p, num_sets, total);
    } else {
      hipLaunchKernelGGL(count_kernel, dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)),
                         dim3(kBT), 0, st, p->d_off, p->d_len, p->d_data, num_sets,
                         s.d_counts);
    }
  }
  RIP_HIP(hipGetLastError());
  return dense;
}

// Captures clear + k seed rounds on the pool's stream as an executable graph:
// (argmax, kill) per round, or for dense pools (argmax, scan, two decrement
// kernels of which one does the work).  Only kernels are recorded (no memset
// nodes); the capture mode is thread-local so other host threads' streams
// are unaffected.
void capture_rounds(rip_pool *p, uint32_t k, bool dense, bool cols) {
  SelectScratch &s = p->sel;
  if (s.graph) {
    (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
    s.graph_k = 0;
  }
  hipStream_t st = p->stream;
  const uint32_t n = p->g->n;
  const uint32_t argmax_grid = grid_for(n, kBT * 8, kArgmaxBlocks);
  const uint32_t ranges = ranges_of(n);
  const uint32_t dec_grid = ranges * dense_chunks(ranges, ~0ull);
  RIP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(clear_keys_kernel, dim3(1), dim3(256), 0, st, s.d_args, k);
  for (uint32_t i = 0; i < k; ++i) {
    hipLaunchKernelGGL(argmax_kernel, dim3(argmax_grid), dim3(kBT), 0, st, s.d_args);
    if (dense) {
      hipLaunchKernelGGL(scan_kill_kernel, dim3(kScanGrid), dim3(kBT), 0, st, s.d_args, i,
                         argmax_grid);
      hipLaunchKernelGGL(dec_small_kernel, dim3(256), dim3(kBT), 0, st, s.d_args);
      hipLaunchKernelGGL(dense_dec_kernel, dim3(dec_grid), dim3(kDenseBT), 0, st, s.d_args,
                         ranges);
      if (cols) {
        hipLaunchKernelGGL(hot_dec_kernel, dim3(grid_for(p->g->hot_n, kBT, 8192)), dim3(kBT), 0,
                           st, s.d_args);
        hipLaunchKernelGGL(clear_gkill_kernel, dim3(1), dim3(256), 0, st, s.d_args);
      }
    } else {
      hipLaunchKernelGGL(kill_kernel, dim3(256), dim3(kBT), 0, st, s.d_args, i, argmax_grid);
    }
  }
  const hipError_t launch = hipGetLastError();
  hipGraph_t graph = nullptr;
  const hipError_t end = hipStreamEndCapture(st, &graph);
  RIP_HIP(launch);
  RIP_HIP(end);
  hipGraphExec_t exec = nullptr;
  const hipError_t inst = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  RIP_HIP(inst);
  s.graph = exec;
  s.graph_k = k;
  s.graph_dense = dense;
  s.graph_cols = cols;
}

}  // namespace

uint32_t select_run(rip_pool *p, uint64_t num_sets, uint32_t k, uint32_t *seeds,
                    uint64_t *covered) {
  const uint32_t n = p->g->n;
  hipStream_t st = p->stream;
  const uint32_t *sorted_inv = nullptr;
  const bool dense = build_counts(p, num_sets, std::max<uint32_t>(k, 1), Need::Select,
                                  &sorted_inv);
  SelectScratch &s = p->sel;
  if (!dense && !sorted_inv) {
    exclusive_scan_u32(s.d_counts, s.d_inv_start, n, st);
    hipLaunchKernelGGL(copy_u64_kernel, dim3(grid_for(n, kBT, 2048)), dim3(kBT), 0, st,
                       s.d_cursor, s.d_inv_start, static_cast<uint64_t>(n));
    if (num_sets)
      hipLaunchKernelGGL(fill_inverted_kernel,
                         dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)), dim3(kBT), 0, st,
                         p->d_off, p->d_len, p->d_data, num_sets, s.d_cursor, s.d_inv);
  }
  hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(num_sets, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_covered, num_sets);
  // the seed rounds run as one graph; this call's buffers travel in d_args
  // (`args` stays alive until the stream is synchronised below)
  const bool cols = p->groups > 0;
  const Args args{s.d_counts, s.d_inv_start, sorted_inv ? sorted_inv : s.d_inv, p->d_off,
                  p->d_len, p->d_data,
                  s.d_covered, s.d_keys, p->d_ctl, s.d_kill, s.d_partial, num_sets, n,
                  p->d_gref, p->d_colptr, p->g->d_hot_index, p->g->d_hot_vertex, s.d_gkill,
                  s.d_kgroups, cols ? p->groups : 0u, p->g->hot_n};
  RIP_HIP(hipMemcpyAsync(s.d_args, &args, sizeof(Args), hipMemcpyHostToDevice, st));
  if (k) {
    if (!s.graph || s.graph_k != k || s.graph_dense != dense || s.graph_cols != cols)
      capture_rounds(p, k, dense, cols);
    RIP_HIP(hipGraphLaunch(s.graph, st));
  } else {
    hipLaunchKernelGGL(clear_keys_kernel, dim3(1), dim3(256), 0, st, s.d_args, 0u);
  }
  RIP_HIP(hipGetLastError());
  std::vector<unsigned long long> keys(std::max<uint32_t>(k, 1));
  unsigned long long cov = 0;
  if (k)
    RIP_HIP(hipMemcpyAsync(keys.data(), s.d_keys, sizeof(unsigned long long) * k,
                           hipMemcpyDeviceToHost, st));
  RIP_HIP(hipMemcpyAsync(&cov, p->d_ctl + kCtlCovered, sizeof(cov),
                         hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  uint32_t got = 0;
  for (uint32_t i = 0; i < k; ++i) {
    if ((keys[i] >> 32) == 0) break;
    seeds[got++] = 0xFFFFFFFFu - static_cast<uint32_t>(keys[i]);
  }
  if (covered) *covered = cov;
  return got;
}

uint64_t coverage_run(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                      uint32_t num_seeds) {
  RIP_REQUIRE(num_sets <= p->num_sets, -1, "coverage: num_sets beyond the pool");
  const uint32_t n = p->g->n;
  for (uint32_t i = 0; i < num_seeds; ++i)
    RIP_REQUIRE(seeds[i] < n, -1, "coverage: seed vertex out of range");
  hipStream_t st = p->stream;
  RIP_HIP(hipSetDevice(p->g->device));
  // column pools also mark per-set hits (in the covered flags)
  ensure_select_scratch(p, p->groups ? num_sets : 0, 0, 0, std::max<uint32_t>(num_seeds, 1));
  SelectScratch &s = p->sel;
  // the seeds travel through the key buffer (k >= num_seeds words of 8 bytes)
  if (num_seeds)
    RIP_HIP(hipMemcpyAsync(s.d_keys, seeds, sizeof(uint32_t) * num_seeds,
                           hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_counts, static_cast<uint64_t>(n), 0u);
  if (num_seeds)
    hipLaunchKernelGGL(flag_seeds_kernel, dim3(grid_for(num_seeds, kBT, 64)), dim3(kBT), 0, st,
                       reinterpret_cast<const uint32_t *>(s.d_keys), num_seeds, s.d_counts);
  unsigned long long *out = p->d_ctl + kCtlWords - 1;
  hipLaunchKernelGGL(zero_word_kernel, dim3(1), dim3(1), 0, st, out);
  if (num_sets && num_seeds && p->groups) {
    // column sets: hits through their lists or their hot members, counted
    // once per set
    hipLaunchKernelGGL(coverage_kernel<false>, dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)),
                       dim3(kBT), 0, st, p->d_off, p->d_len, p->d_data, num_sets, s.d_counts,
                       p->d_ctl + kCtlScratch2, s.d_covered);
    hipLaunchKernelGGL(coverage_hot_kernel, dim3(grid_for(num_sets, kBT, 4096)), dim3(kBT), 0, st,
                       p->d_gref, p->d_colptr, p->g->d_hot_index, num_sets,
                       reinterpret_cast<const uint32_t *>(s.d_keys), num_seeds, s.d_covered);
    hipLaunchKernelGGL(count_u8_kernel, dim3(grid_for(num_sets, kBT, 1024)), dim3(kBT), 0, st,
                       s.d_covered, num_sets, out);
  } else if (num_sets && num_seeds) {
    hipLaunchKernelGGL(coverage_kernel<false>, dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)),
                       dim3(kBT), 0, st, p->d_off, p->d_len, p->d_data, num_sets, s.d_counts,
                       out);
  }
  RIP_HIP(hipGetLastError());
  unsigned long long cov = 0;
  RIP_HIP(hipMemcpyAsync(&cov, out, sizeof(cov), hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  return cov;
}

// ------------------------------------------------- coverage, no allocation
// rip_pool_coverage_noalloc works in this fixed device array instead of the
// pool's selector buffers: a vertex bitmap (kNaMaxNodes bits), the seeds
// (kNaMaxSeeds), per-set hit bits of column pools (kNaMaxColumnSets bits) and
// the result counter, 8.3 MB in all.  Each device holds its own copy, found
// with hipGetSymbolAddress on the device's first call; every call on one
// device shares it, so calls that overlap on a device race.
constexpr uint64_t kNaMaxNodes = 1ull << 25;
constexpr uint64_t kNaMaxSeeds = 1ull << 16;
constexpr uint64_t kNaMaxColumnSets = 1ull << 25;
constexpr uint64_t kNaBitmapWords = kNaMaxNodes / 64;         // u64 words
constexpr uint64_t kNaSeedWords = kNaMaxSeeds / 2;
constexpr uint64_t kNaHitWords = kNaMaxColumnSets / 64;
constexpr uint64_t kNaSeedAt = kNaBitmapWords;
constexpr uint64_t kNaHitAt = kNaSeedAt + kNaSeedWords;
constexpr uint64_t kNaCounterAt = kNaHitAt + kNaHitWords;
constexpr uint64_t kNaWords = kNaCounterAt + 32;

static __device__ __align__(256) unsigned long long scratch[kNaWords];

namespace {
constexpr int kNaMaxDevices = 64;
std::mutex na_mu;
unsigned long long *na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on the current device.
unsigned long long *noalloc_scratch(int device) {
  RIP_REQUIRE(device >= 0 && device < kNaMaxDevices, RIP_EINVAL, "coverage_noalloc: device id");
  std::lock_guard<std::mutex> hold(na_mu);
  if (!na_ptr[device]) {
    void *ptr = nullptr;
    RIP_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    RIP_REQUIRE(ptr, RIP_EDEVICE, "coverage_noalloc: no device address for the scratch array");
    na_ptr[device] = static_cast<unsigned long long *>(ptr);
  }
  return na_ptr[device];
}
}  // namespace

uint64_t coverage_run_noalloc(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                              uint32_t num_seeds) {
  const uint32_t n = p->g->n;
  RIP_REQUIRE(num_sets <= p->num_sets, RIP_EINVAL, "coverage_noalloc: num_sets beyond the pool");
  RIP_REQUIRE(n <= kNaMaxNodes, RIP_EINVAL,
              "coverage_noalloc: graphs of at most 2^25 vertices (the fixed scratch)");
  RIP_REQUIRE(num_seeds <= kNaMaxSeeds, RIP_EINVAL,
              "coverage_noalloc: at most 65536 seeds (the fixed scratch)");
  RIP_REQUIRE(p->groups == 0 || num_sets <= kNaMaxColumnSets, RIP_EINVAL,
              "coverage_noalloc: pools with column sets: at most 2^25 sets (the fixed scratch)");
  for (uint32_t i = 0; i < num_seeds; ++i)
    RIP_REQUIRE(seeds[i] < n, RIP_EINVAL, "coverage_noalloc: seed vertex out of range");
  RIP_HIP(hipSetDevice(p->g->device));
  unsigned long long *base = noalloc_scratch(p->g->device);
  uint32_t *bits = reinterpret_cast<uint32_t *>(base);
  uint32_t *d_seeds = reinterpret_cast<uint32_t *>(base + kNaSeedAt);
  unsigned long long *hit = base + kNaHitAt;
  unsigned long long *out = base + kNaCounterAt;
  hipStream_t st = p->stream;
  // the previous call's contents stay in the array: clear what this call
  // reads (kernels, not memsets, so the sequence can sit in a captured graph)
  const uint64_t nwords = (static_cast<uint64_t>(n) + 63) / 64;
  const uint64_t hwords = p->groups ? (num_sets + 63) / 64 : 0;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(nwords, kBT, 1024)), dim3(kBT), 0, st, base,
                     nwords);
  hipLaunchKernelGGL(fill_u64_kernel, dim3(1), dim3(64), 0, st, out, 2ull);  // + a dummy
  if (hwords)
    hipLaunchKernelGGL(fill_u64_kernel, dim3(grid_for(hwords, kBT, 1024)), dim3(kBT), 0, st, hit,
                       hwords);
  if (num_seeds) {
    RIP_HIP(hipMemcpyAsync(d_seeds, seeds, sizeof(uint32_t) * num_seeds, hipMemcpyHostToDevice,
                           st));
    hipLaunchKernelGGL(flag_seed_bits_kernel, dim3(grid_for(num_seeds, kBT, 64)), dim3(kBT), 0,
                       st, d_seeds, num_seeds, bits);
  }
  if (num_sets && num_seeds) {
    const dim3 grid(grid_for(num_sets, 64 * (kBT / 64), 8192));
    if (p->groups) {
      // column sets: hits through their lists, then through their hot
      // members; each set counted once from the hit bits
      hipLaunchKernelGGL(coverage_kernel<true>, grid, dim3(kBT), 0, st, p->d_off, p->d_len,
                         p->d_data, num_sets, bits, base + kNaCounterAt + 1, nullptr, hit);
      hipLaunchKernelGGL(coverage_hot_bits_kernel, grid, dim3(kBT), 0, st, p->d_gref,
                         p->d_colptr, p->g->d_hot_index, num_sets, d_seeds, num_seeds, hit);
      hipLaunchKernelGGL(popcount_u64_kernel, dim3(grid_for(hwords, kBT, 1024)), dim3(kBT), 0,
                         st, hit, hwords, out);
    } else {
      hipLaunchKernelGGL(coverage_kernel<true>, grid, dim3(kBT), 0, st, p->d_off, p->d_len,
                         p->d_data, num_sets, bits, out, nullptr, nullptr);
    }
  }
  RIP_HIP(hipGetLastError());
  unsigned long long cov = 0;
  RIP_HIP(hipMemcpyAsync(&cov, out, sizeof(cov), hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  return cov;
}

void count_run(rip_pool *p, uint64_t num_sets, uint32_t *counts_host) {
  build_counts(p, num_sets, 1, Need::Counts);
  RIP_HIP(hipMemcpyAsync(counts_host, p->sel.d_counts, sizeof(uint32_t) * p->g->n,
                         hipMemcpyDeviceToHost, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
}

}  // namespace rip
