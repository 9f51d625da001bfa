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
    s_off[wid][lane] = s < num_sets ? set_off[s] : 0ull;
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
      f(data[s_off[wid][lo] + (i - s_pre[wid][lo])], w * 64 + lo);
    }
    __builtin_amdgcn_wave_barrier();
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

__global__ void sum_len_kernel(const uint32_t *set_len, uint64_t num_sets,
                               unsigned long long *out) {
  unsigned long long acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < num_sets; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    acc += set_len[i];
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if (lane_id() == 0 && acc) atomicAdd(out, acc);
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

// Seed round i, first half: keys[i] = max over v of counts[v] << 32 | ~v.
__global__ void __launch_bounds__(kBT) argmax_kernel(const Args *a, uint32_t i) {
  __shared__ unsigned long long s_w[kBT / 64];
  const uint32_t *counts = a->counts;
  const uint32_t n = a->n;
  unsigned long long best = 0;
  for (uint32_t v = blockIdx.x * kBT + threadIdx.x; v < n; v += gridDim.x * kBT) {
    const unsigned long long key =
        (static_cast<unsigned long long>(counts[v]) << 32) | (0xFFFFFFFFu - v);
    best = key > best ? key : best;
  }
  best = wave_max_u64(best);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (uint32_t w = 1; w < kBT / 64; ++w) best = s_w[w] > best ? s_w[w] : best;
    if (best >> 32) atomicMax(a->keys + i, best);
  }
}

// Seed round i, second half: every uncovered set containing the seed is
// marked covered once and its members' counts drop by one.
__global__ void __launch_bounds__(kBT) kill_kernel(const Args *a, uint32_t i) {
  const unsigned long long key = a->keys[i];
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

// Sets [0, num_sets) that contain a flagged vertex (FindCoverage,
// include/ripples/opimc.h:100-118): each wave walks 64 sets, marks hits in
// LDS and adds their popcount to *out once per wave.
__global__ void __launch_bounds__(kBT)
coverage_kernel(const unsigned long long *set_off, const uint32_t *set_len,
                const uint32_t *data, uint64_t num_sets, const uint32_t *flags,
                unsigned long long *out) {
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
      if (flags[v]) s_hit[wid][lo] = 1;
    }
    __builtin_amdgcn_wave_barrier();
    acc += __popcll(__ballot(s_hit[wid][lane] != 0));
    __builtin_amdgcn_wave_barrier();
  }
  if (lane == 0 && acc) atomicAdd(out, acc);
}

__global__ void flag_seeds_kernel(const uint32_t *seeds, uint32_t k, uint32_t *flags) {
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < k; i += gridDim.x * blockDim.x)
    flags[seeds[i]] = 1;
}

void ensure_select_scratch(rip_pool *p, uint64_t num_sets, uint64_t total,
                           uint32_t k) {
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
    s.cap_sets = std::max<uint64_t>(num_sets, s.cap_sets + s.cap_sets / 2);
    s.d_covered = dev_alloc<uint8_t>(s.cap_sets);
  }
  if (s.cap_elems < total) {
    dev_free(s.d_inv);
    s.cap_elems = std::max<uint64_t>(total, s.cap_elems + s.cap_elems / 2);
    s.d_inv = dev_alloc<uint32_t>(s.cap_elems);
  }
  if (s.cap_k < k) {
    dev_free(s.d_keys);
    s.cap_k = k;
    s.d_keys = dev_alloc<unsigned long long>(k);
  }
  if (!s.d_args) s.d_args = dev_alloc<Args>(1);
}

uint32_t grid_for(uint64_t work, uint32_t per_block, uint32_t cap) {
  uint64_t g = (work + per_block - 1) / per_block;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(g, cap)));
}

// counts over sets [0, num_sets); returns the element total
uint64_t build_counts(rip_pool *p, uint64_t num_sets, uint32_t k) {
  RIP_REQUIRE(num_sets <= p->num_sets, -1, "select: num_sets beyond the pool");
  const uint32_t n = p->g->n;
  hipStream_t st = p->stream;
  RIP_HIP(hipSetDevice(p->g->device));
  // total element count
  hipLaunchKernelGGL(zero_word_kernel, dim3(1), dim3(1), 0, st, p->d_ctl + kCtlWords - 1);
  hipLaunchKernelGGL(sum_len_kernel, dim3(grid_for(num_sets, kBT, 1024)), dim3(kBT), 0, st,
                     p->d_len, num_sets, p->d_ctl + kCtlWords - 1);
  unsigned long long total = 0;
  RIP_HIP(hipMemcpyAsync(&total, p->d_ctl + kCtlWords - 1, sizeof(total),
                         hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  ensure_select_scratch(p, num_sets, total, k);
  SelectScratch &s = p->sel;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(grid_for(n, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_counts, static_cast<uint64_t>(n), 0u);
  if (num_sets)
    hipLaunchKernelGGL(count_kernel, dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)),
                       dim3(kBT), 0, st, p->d_off, p->d_len, p->d_data, num_sets,
                       s.d_counts);
  RIP_HIP(hipGetLastError());
  return total;
}

// Captures clear + k x (argmax, kill) on the pool's stream as an executable
// graph.  Only kernels are recorded (no memset nodes); the capture mode is
// thread-local so other host threads' streams are unaffected.
void capture_rounds(rip_pool *p, uint32_t k) {
  SelectScratch &s = p->sel;
  if (s.graph) {
    (void)hipGraphExecDestroy(s.graph);
    s.graph = nullptr;
    s.graph_k = 0;
  }
  hipStream_t st = p->stream;
  const uint32_t argmax_grid = grid_for(p->g->n, kBT * 8, 1024);
  RIP_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  hipLaunchKernelGGL(clear_keys_kernel, dim3(1), dim3(256), 0, st, s.d_args, k);
  for (uint32_t i = 0; i < k; ++i) {
    hipLaunchKernelGGL(argmax_kernel, dim3(argmax_grid), dim3(kBT), 0, st, s.d_args, i);
    hipLaunchKernelGGL(kill_kernel, dim3(256), dim3(kBT), 0, st, s.d_args, i);
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
}

}  // namespace

uint32_t select_run(rip_pool *p, uint64_t num_sets, uint32_t k, uint32_t *seeds,
                    uint64_t *covered) {
  const uint32_t n = p->g->n;
  hipStream_t st = p->stream;
  build_counts(p, num_sets, std::max<uint32_t>(k, 1));
  SelectScratch &s = p->sel;
  exclusive_scan_u32(s.d_counts, s.d_inv_start, n, st);
  hipLaunchKernelGGL(copy_u64_kernel, dim3(grid_for(n, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_cursor, s.d_inv_start, static_cast<uint64_t>(n));
  if (num_sets)
    hipLaunchKernelGGL(fill_inverted_kernel,
                       dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)), dim3(kBT), 0, st,
                       p->d_off, p->d_len, p->d_data, num_sets, s.d_cursor, s.d_inv);
  hipLaunchKernelGGL(fill_u8_kernel, dim3(grid_for(num_sets, kBT, 2048)), dim3(kBT), 0, st,
                     s.d_covered, num_sets);
  // the seed rounds run as one graph; this call's buffers travel in d_args
  // (`args` stays alive until the stream is synchronised below)
  const Args args{s.d_counts, s.d_inv_start, s.d_inv, p->d_off, p->d_len, p->d_data,
                  s.d_covered, s.d_keys, p->d_ctl, n};
  RIP_HIP(hipMemcpyAsync(s.d_args, &args, sizeof(Args), hipMemcpyHostToDevice, st));
  if (k) {
    if (!s.graph || s.graph_k != k) capture_rounds(p, k);
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
  ensure_select_scratch(p, 0, 0, std::max<uint32_t>(num_seeds, 1));
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
  if (num_sets && num_seeds)
    hipLaunchKernelGGL(coverage_kernel, dim3(grid_for(num_sets, 64 * (kBT / 64), 8192)),
                       dim3(kBT), 0, st, p->d_off, p->d_len, p->d_data, num_sets, s.d_counts,
                       out);
  RIP_HIP(hipGetLastError());
  unsigned long long cov = 0;
  RIP_HIP(hipMemcpyAsync(&cov, out, sizeof(cov), hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  return cov;
}

void count_run(rip_pool *p, uint64_t num_sets, uint32_t *counts_host) {
  build_counts(p, num_sets, 1);
  RIP_HIP(hipMemcpyAsync(counts_host, p->sel.d_counts, sizeof(uint32_t) * p->g->n,
                         hipMemcpyDeviceToHost, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
}

}  // namespace rip
