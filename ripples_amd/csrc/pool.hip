// HBM-resident RRR pool.
//
// Sets of one (graph, seed) stream are indexed by sample id: d_off[s] and
// d_len[s] locate sample s's sorted vertex list inside d_data.  Samplers
// reserve space with one atomic bump per colour group (IC) or per wave (LT),
// so the lists of a group are contiguous but groups land in claim order;
// export compacts a range into CSR order.  Replaces the RRRsets vector of
// include/ripples/rrr_sets.h:81-105 and the flattened device copy built for
// selection (streaming_find_most_influential.h:139-244).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <cstring>
#include <vector>

#include "device_util.h"
#include "internal.h"

namespace rip {
namespace {

constexpr uint32_t kScanBT = 256;
constexpr uint32_t kScanItems = 8;
constexpr uint32_t kScanTile = kScanBT * kScanItems;

// The scans add in 64 bits: a tile of 2048 counts of up to 2^32 - 1 each
// (vertex occurrence counts, set sizes) can pass 2^32.
__global__ void scan_tile_sums(const uint32_t *in, uint64_t n,
                               unsigned long long *tile_sum) {
  __shared__ unsigned long long s_tmp[kScanBT / 64];
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile;
  unsigned long long acc = 0;
  for (uint32_t i = 0; i < kScanItems; ++i) {
    const uint64_t j = base + static_cast<uint64_t>(i) * kScanBT + threadIdx.x;
    if (j < n) acc += in[j];
  }
  unsigned long long total;
  (void)block_exclusive_scan_u64<kScanBT>(acc, s_tmp, &total);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = total;
}

// single workgroup: exclusive scan of the tile sums (in place)
__global__ void scan_tile_prefix(unsigned long long *tile_sum, uint64_t tiles) {
  __shared__ unsigned long long s_w[kScanBT / 64];
  __shared__ unsigned long long s_carry;
  if (threadIdx.x == 0) s_carry = 0;
  __syncthreads();
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  for (uint64_t b = 0; b < tiles; b += kScanBT) {
    const uint64_t j = b + threadIdx.x;
    const unsigned long long v = j < tiles ? tile_sum[j] : 0ull;
    unsigned long long x = v;
    for (int d = 1; d < 64; d <<= 1) {
      unsigned long long y = __shfl_up(x, d, 64);
      if (lane >= static_cast<uint32_t>(d)) x += y;
    }
    if (lane == 63) s_w[wid] = x;
    __syncthreads();
    unsigned long long off = s_carry;
    for (uint32_t w = 0; w < wid; ++w) off += s_w[w];
    if (j < tiles) tile_sum[j] = off + x - v;
    __syncthreads();
    if (threadIdx.x == kScanBT - 1) s_carry = off + x;
    __syncthreads();
  }
}

__global__ void scan_apply(const uint32_t *in, uint64_t n,
                           const unsigned long long *tile_prefix,
                           unsigned long long *out) {
  __shared__ unsigned long long s_tmp[kScanBT / 64];
  // each thread owns kScanItems consecutive elements of the tile
  const uint64_t base = static_cast<uint64_t>(blockIdx.x) * kScanTile +
                        static_cast<uint64_t>(threadIdx.x) * kScanItems;
  uint32_t v[kScanItems];
  unsigned long long acc = 0;
  for (uint32_t i = 0; i < kScanItems; ++i) {
    v[i] = (base + i < n) ? in[base + i] : 0u;
    acc += v[i];
  }
  unsigned long long total;
  const unsigned long long pre = block_exclusive_scan_u64<kScanBT>(acc, s_tmp, &total);
  unsigned long long run = tile_prefix[blockIdx.x] + pre;
  for (uint32_t i = 0; i < kScanItems; ++i) {
    if (base + i < n) out[base + i] = run;
    run += v[i];
  }
  if (base + kScanItems >= n && base < n) out[n] = run;  // thread holding the last element
}

__global__ void set_ctl_kernel(unsigned long long *ctl, unsigned long long bump) {
  ctl[kCtlDataBump] = bump;
  ctl[kCtlOverflow] = 0;
  ctl[kCtlGroupCounter] = 0;
  ctl[kCtlLongCount] = 0;
  ctl[kCtlElems] = 0;
  ctl[kCtlGroupCounter2] = 0;
  for (int i = kCtlStatBytes; i <= kCtlStatLevels; ++i) ctl[i] = 0;
}

__global__ void zero_elems_kernel(unsigned long long *ctl) { ctl[kCtlElems] = 0; }

// *out += sum of v[0, n): one atomic per block (one per wave measured as a
// serial tail of tens of microseconds on the single target word)
__global__ void __launch_bounds__(256) sum_u32_kernel(const uint32_t *v, uint64_t n,
                                                      unsigned long long *out) {
  __shared__ unsigned long long s_w[4];
  unsigned long long acc = 0;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * 256 + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * 256)
    acc += v[i];
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    acc = s_w[0] + s_w[1] + s_w[2] + s_w[3];
    if (acc) atomicAdd(out, acc);
  }
}

// Copy sets [first, first+count) into CSR order at out (prefix from scan).
// A wave handles 64 consecutive sets: a wave scan of their sizes lets every
// lane copy elements of any of the 64 sets, so small and large sets balance.
__global__ void gather_sets(const unsigned long long *set_off,
                            const uint32_t *set_len, const uint32_t *data,
                            uint64_t first, uint64_t count,
                            const unsigned long long *prefix, uint32_t *out) {
  __shared__ uint32_t s_pre[4][65];
  __shared__ unsigned long long s_off[4][64];
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint64_t w0 = (static_cast<uint64_t>(blockIdx.x) * 4 + wid) * 64;
  if (w0 >= count) return;
  const uint64_t j = w0 + lane;
  const uint32_t len = j < count ? set_len[first + j] : 0u;
  s_off[wid][lane] = j < count ? set_off[first + j] : 0ull;
  const uint32_t incl = wave_inclusive_scan(len);
  s_pre[wid][lane + 1] = incl;
  if (lane == 0) s_pre[wid][0] = 0;
  const uint32_t total = __shfl(incl, 63, 64);
  __builtin_amdgcn_wave_barrier();
  const unsigned long long obase = prefix[w0];
  for (uint32_t f = lane; f < total; f += 64) {
    uint32_t lo = 0, hi = 64;
    while (lo < hi) {
      uint32_t mid = (lo + hi) >> 1;
      if (s_pre[wid][mid + 1] <= f) lo = mid + 1;
      else hi = mid;
    }
    out[obase + f] = data[s_off[wid][lo] + (f - s_pre[wid][lo])];
  }
}

// Column store export: hot members of giant set j (entry g << 6 | c).
constexpr uint32_t kHotBT = 1024;

__global__ void __launch_bounds__(kHotBT) hot_size_kernel(unsigned long long *const *colptr,
                                                          const uint32_t *ent, uint32_t hot_n,
                                                          unsigned long long *sizes) {
  __shared__ unsigned long long s_tmp[kHotBT / 64];
  const uint32_t e = ent[blockIdx.x];
  const unsigned long long *col = colptr[e >> 6];
  const uint32_t c = e & 63u;
  unsigned long long cnt = 0;
  for (uint32_t h = threadIdx.x; h < hot_n; h += kHotBT) cnt += (col[h] >> c) & 1ull;
  unsigned long long tot;
  (void)block_exclusive_scan_u64<kHotBT>(cnt, s_tmp, &tot);
  if (threadIdx.x == 0) sizes[blockIdx.x] = tot;
}

// The hot members of set j in ascending vertex order (hot vertices are
// numbered in vertex order) at out[hoff[j] ...].
__global__ void __launch_bounds__(kHotBT) hot_extract_kernel(unsigned long long *const *colptr,
                                                             const uint32_t *ent, uint32_t hot_n,
                                                             const uint32_t *hot_vertex,
                                                             const unsigned long long *hoff,
                                                             uint32_t *out) {
  __shared__ unsigned long long s_tmp[kHotBT / 64];
  const uint32_t e = ent[blockIdx.x];
  const unsigned long long *col = colptr[e >> 6];
  const uint32_t c = e & 63u;
  unsigned long long run = hoff[blockIdx.x];
  for (uint32_t h0 = 0; h0 < hot_n; h0 += kHotBT) {
    const uint32_t h = h0 + threadIdx.x;
    const unsigned long long bit = h < hot_n ? (col[h] >> c) & 1ull : 0ull;
    unsigned long long tot;
    const unsigned long long pre = block_exclusive_scan_u64<kHotBT>(bit, s_tmp, &tot);
    if (bit) out[run + pre] = hot_vertex[h];
    run += tot;
  }
}

}  // namespace

void exclusive_scan_u32(const uint32_t *d_in, unsigned long long *d_out,
                        uint64_t n, hipStream_t stream) {
  if (n == 0) {
    const unsigned long long zero = 0;
    RIP_HIP(hipMemcpyAsync(d_out, &zero, sizeof(zero), hipMemcpyHostToDevice, stream));
    RIP_HIP(hipStreamSynchronize(stream));
    return;
  }
  const uint64_t tiles = (n + kScanTile - 1) / kScanTile;
  unsigned long long *tmp = dev_alloc<unsigned long long>(tiles);
  hipLaunchKernelGGL(scan_tile_sums, dim3(static_cast<uint32_t>(tiles)), dim3(kScanBT),
                     0, stream, d_in, n, tmp);
  hipLaunchKernelGGL(scan_tile_prefix, dim3(1), dim3(kScanBT), 0, stream, tmp, tiles);
  hipLaunchKernelGGL(scan_apply, dim3(static_cast<uint32_t>(tiles)), dim3(kScanBT), 0,
                     stream, d_in, n, tmp, d_out);
  hipError_t e = hipGetLastError();
  RIP_HIP(hipStreamSynchronize(stream));
  dev_free(tmp);
  RIP_HIP(e);
}

void pool_create(rip_pool *p, rip_graph *g, uint64_t seed) {
  RIP_HIP(hipSetDevice(g->device));
  p->g = g;
  p->seed = seed;
  try {
    pool_create_resources(p);
  } catch (...) {
    // release whatever was allocated before the failure; the caller may
    // retry on the same object
    pool_destroy(p);
    p->g = g;
    throw;
  }
}

void pool_create_resources(rip_pool *p) {
  RIP_HIP(hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking));
  RIP_HIP(hipEventCreate(&p->ev0));
  RIP_HIP(hipEventCreate(&p->ev1));
  RIP_HIP(hipEventCreate(&p->ev2));
  RIP_HIP(hipEventCreate(&p->ev3));
  RIP_HIP(hipEventCreate(&p->ev4));
  p->d_ctl = dev_alloc<unsigned long long>(kCtlWords);
  // pinned, so reading the control words back after a launch is one DMA
  // rather than a copy staged through a driver buffer
  RIP_HIP(hipHostMalloc(reinterpret_cast<void **>(&p->h_ctl),
                        sizeof(unsigned long long) * kCtlWords, hipHostMallocDefault));
  RIP_HIP(hipMemsetAsync(p->d_ctl, 0, sizeof(unsigned long long) * kCtlWords, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
}

void pool_destroy(rip_pool *p) {
  if (!p || !p->g) return;
  (void)hipSetDevice(p->g->device);
  if (p->stream) (void)hipStreamSynchronize(p->stream);
  dev_free(p->d_off);
  dev_free(p->d_len);
  dev_free(p->d_data);
  dev_free(p->d_ctl);
  dev_free(p->d_gref);
  dev_free(p->d_colptr);
  dev_free(p->d_gsample);
  for (auto *b : p->col_blocks) dev_free(b);
  p->col_blocks.clear();
  p->groups = p->cap_groups = 0;
  dev_free(p->sel.d_counts);
  dev_free(p->sel.d_inv_start);
  dev_free(p->sel.d_cursor);
  dev_free(p->sel.d_inv);
  dev_free(p->sel.d_pkey[0]);
  dev_free(p->sel.d_pkey[1]);
  dev_free(p->sel.d_pval);
  dev_free(p->sel.d_soff);
  dev_free(p->sel.d_stemp);
  dev_free(p->sel.d_kill);
  dev_free(p->sel.d_covered);
  dev_free(p->sel.d_keys);
  dev_free(p->sel.d_partial);
  dev_free(p->sel.d_args);
  dev_free(p->sel.d_gactive);
  dev_free(p->sel.d_gkill);
  dev_free(p->sel.d_kgroups);
  if (p->sel.graph) (void)hipGraphExecDestroy(p->sel.graph);
  if (p->h_ctl) (void)hipHostFree(p->h_ctl);
  p->h_ctl = nullptr;
  if (p->ev0) (void)hipEventDestroy(p->ev0);
  if (p->ev1) (void)hipEventDestroy(p->ev1);
  if (p->ev2) (void)hipEventDestroy(p->ev2);
  if (p->ev3) (void)hipEventDestroy(p->ev3);
  if (p->ev4) (void)hipEventDestroy(p->ev4);
  p->ev0 = p->ev1 = p->ev2 = p->ev3 = p->ev4 = nullptr;
  if (p->stream) (void)hipStreamDestroy(p->stream);
  p->stream = nullptr;
  p->sel = SelectScratch{};
  p->cap_sets = p->cap_data = p->used_data = p->num_sets = p->num_elements = 0;
}

void pool_reserve_sets(rip_pool *p, uint64_t cap) {
  if (cap <= p->cap_sets) return;
  cap = std::max(cap, 2 * p->cap_sets);
  auto *off = dev_alloc<unsigned long long>(cap);
  uint32_t *len = nullptr, *gref = nullptr;
  try {
    len = dev_alloc<uint32_t>(cap);
    if (p->d_gref) gref = dev_alloc<uint32_t>(cap);
  } catch (...) {
    dev_free(off);
    dev_free(len);
    throw;
  }
  if (p->num_sets) {
    RIP_HIP(hipMemcpyAsync(off, p->d_off, sizeof(unsigned long long) * p->num_sets,
                           hipMemcpyDeviceToDevice, p->stream));
    RIP_HIP(hipMemcpyAsync(len, p->d_len, sizeof(uint32_t) * p->num_sets,
                           hipMemcpyDeviceToDevice, p->stream));
    if (gref)
      RIP_HIP(hipMemcpyAsync(gref, p->d_gref, sizeof(uint32_t) * p->num_sets,
                             hipMemcpyDeviceToDevice, p->stream));
  }
  RIP_HIP(hipStreamSynchronize(p->stream));
  dev_free(p->d_off);
  dev_free(p->d_len);
  dev_free(p->d_gref);
  p->d_off = off;
  p->d_len = len;
  p->d_gref = gref;
  p->cap_sets = cap;
}

namespace {
__global__ void fill_u32_range_kernel(uint32_t *p, uint64_t first, uint64_t end, uint32_t v) {
  for (uint64_t i = first + static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
       i < end; i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    p[i] = v;
}
}  // namespace

void pool_clear_gref(rip_pool *p, uint64_t first, uint64_t end) {
  if (!p->d_gref || end <= first) return;
  const uint64_t blocks = std::min<uint64_t>((end - first + 255) / 256, 4096);
  hipLaunchKernelGGL(fill_u32_range_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0,
                     p->stream, p->d_gref, first, end, kNoGroup);
  RIP_HIP(hipGetLastError());
}

// Room for `need` column groups: columns are allocated kColBlock at a time,
// so the store grows without moving them; the pointer and sample tables are
// reallocated.  The first call also sets up d_gref (every set so far a plain
// list).
uint64_t pool_device_bytes(const rip_pool *p) {
  uint64_t b = 4ull * p->cap_data + 12ull * p->cap_sets;      // data, d_off, d_len
  if (p->d_gref) b += 4ull * p->cap_sets;
  b += 8ull * kColBlock * p->g->hot_n * p->col_blocks.size();
  b += (8ull + 4ull * 64) * p->cap_groups;                     // d_colptr, d_gsample
  return b;
}

void pool_reserve_groups(rip_pool *p, uint32_t need) {
  graph_ensure_hot(p->g, p->stream);
  if (!p->d_gref) {
    p->d_gref = dev_alloc<uint32_t>(std::max<uint64_t>(p->cap_sets, 1));
    pool_clear_gref(p, 0, std::max<uint64_t>(p->cap_sets, 1));
  }
  if (need <= p->cap_groups) return;
  const uint32_t cap = (need + kColBlock - 1) / kColBlock * kColBlock;
  const uint64_t hot = p->g->hot_n;
  std::vector<unsigned long long *> ptrs(cap);
  if (p->cap_groups)
    RIP_HIP(hipMemcpyAsync(ptrs.data(), p->d_colptr, sizeof(void *) * p->cap_groups,
                           hipMemcpyDeviceToHost, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
  for (uint32_t b = p->cap_groups; b < cap; b += kColBlock) {
    unsigned long long *blk = dev_alloc<unsigned long long>(kColBlock * hot);
    p->col_blocks.push_back(blk);
    for (uint32_t i = 0; i < kColBlock; ++i) ptrs[b + i] = blk + i * hot;
  }
  unsigned long long **colptr = dev_alloc<unsigned long long *>(cap);
  uint32_t *gsample = nullptr;
  try {
    gsample = dev_alloc<uint32_t>(64ull * cap);
    RIP_HIP(hipMemcpyAsync(colptr, ptrs.data(), sizeof(void *) * cap, hipMemcpyHostToDevice,
                           p->stream));
    if (p->cap_groups)
      RIP_HIP(hipMemcpyAsync(gsample, p->d_gsample, sizeof(uint32_t) * 64 * p->cap_groups,
                             hipMemcpyDeviceToDevice, p->stream));
    RIP_HIP(hipStreamSynchronize(p->stream));
  } catch (...) {
    dev_free(colptr);
    dev_free(gsample);
    throw;
  }
  dev_free(p->d_colptr);
  dev_free(p->d_gsample);
  p->d_colptr = colptr;
  p->d_gsample = gsample;
  p->cap_groups = cap;
}

int blocks_per_cu(const void *kernel, int block) {
  static std::mutex mu;
  static std::map<std::tuple<int, const void *, int>, int> cache;
  int dev = 0;
  RIP_HIP(hipGetDevice(&dev));
  const auto key = std::make_tuple(dev, kernel, block);
  {
    std::lock_guard<std::mutex> hold(mu);
    const auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  std::lock_guard<std::mutex> hold(mu);
  cache[key] = per_cu;
  return per_cu;
}

int cu_count() {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0;
  RIP_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> hold(mu);
  const auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  int cus = 1;
  RIP_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  cache[dev] = std::max(cus, 1);
  return cache[dev];
}

void launch_sum_u32(const uint32_t *v, uint64_t n, unsigned long long *out, hipStream_t st) {
  if (n == 0) return;
  const uint64_t blocks = std::min<uint64_t>((n + 4095) / 4096, 512);
  hipLaunchKernelGGL(sum_u32_kernel, dim3(static_cast<uint32_t>(blocks)), dim3(256), 0, st, v, n,
                     out);
  RIP_HIP(hipGetLastError());
}

// A copy on the pool's stream, finished before returning (no legacy-stream
// operation: other pools' streams may be capturing a graph meanwhile).
static void copy_sync(rip_pool *p, void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
  RIP_HIP(hipMemcpyAsync(dst, src, bytes, kind, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
}

void pool_reserve_data(rip_pool *p, uint64_t cap) {
  if (cap <= p->cap_data) return;
  auto *data = dev_alloc<uint32_t>(cap);
  if (p->used_data)
    RIP_HIP(hipMemcpyAsync(data, p->d_data, sizeof(uint32_t) * p->used_data,
                           hipMemcpyDeviceToDevice, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
  dev_free(p->d_data);
  p->d_data = data;
  p->cap_data = cap;
}

void pool_read_ctl(rip_pool *p, bool main) {
  RIP_HIP(hipMemcpyAsync(p->h_ctl, p->d_ctl, sizeof(unsigned long long) * kCtlWords,
                         hipMemcpyDeviceToHost, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
  if (main)
    for (int i = 0; i < RIP_STAT_COUNT; ++i) p->main_stats[i] = p->h_ctl[kCtlStatBytes + i];
}

// Samples of the first call of a pool drawn apart to size its data.
constexpr uint64_t kSizingPrefix = 16384;
// Calls of at least this many samples set the graph's set-size hint.
constexpr uint64_t kSizeHintMin = 1024;

void pool_sample(rip_pool *p, uint64_t count, const rip_tune *tune) {
  if (count == 0) return;
  if (!p->sized) {
    // a pool's first call: the set size another pool of this graph measured
    // (an earlier query's, the autotuner's probes) sizes its data, where the
    // default guess would have the call overflow and be drawn again
    std::lock_guard<std::mutex> hold(p->g->sampler_mu);
    if (p->g->set_size_hint > 0.0) {
      p->avg_set_size = p->g->set_size_hint;
      p->sized = true;
    }
  }
  if (!p->sized && count > 4 * kSizingPrefix) {
    // with no set size measured yet the data capacity is a guess, and a call
    // that overflows it is drawn again: a prefix of the call measures the
    // size first (the same sample ids in two calls give the same sets)
    pool_sample(p, kSizingPrefix, tune);
    pool_sample(p, count - kSizingPrefix, tune);
    return;
  }
  rip_tune tuned;
  if (!tune) {   // the autotuner's choice, copied under the lock it is written under
    std::lock_guard<std::mutex> hold(p->g->sampler_mu);
    if (p->g->tuned_set) {
      tuned = p->g->tuned;
      tune = &tuned;
    }
  }
  RIP_HIP(hipSetDevice(p->g->device));
  RIP_REQUIRE(p->num_sets + count < (1ull << 32), -1,
              "pool: at most 2^32-1 samples per stream");
  pool_reserve_sets(p, p->num_sets + count);
  double est = static_cast<double>(count) * p->avg_set_size * 1.25 + 65536.0;
  est += static_cast<double>(p->g->model == RIP_LT ? lt_output_slack(p, count, tune)
                                                  : ic_output_slack(p, count, tune));
  if (p->used_data + static_cast<uint64_t>(est) > p->cap_data)
    pool_reserve_data(p, std::max<uint64_t>(p->used_data + static_cast<uint64_t>(est),
                                            p->cap_data + p->cap_data / 2));
  IcLaunch ic{};
  if (p->g->model == RIP_IC) ic = ic_tune(p, tune);
  const uint64_t first = p->num_sets;
  std::lock_guard<std::mutex> hold(p->g->sampler_mu);  // graph-owned sampler scratch
  for (int attempt = 0;; ++attempt) {
    RIP_REQUIRE(attempt < 40, -2, "pool: data capacity did not converge");
    p->grid_ms = 0.0;     // timings of the attempt that succeeds
    p->grid_levels = 0;
    p->wave_ms = 0.0;
    p->wave_ran = false;
    hipLaunchKernelGGL(set_ctl_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl,
                       static_cast<unsigned long long>(p->used_data));
    pool_clear_gref(p, first, first + count);
    // the launchers bracket their main kernel with ev0 / ev1
    uint32_t launches;
    if (p->g->model == RIP_IC)
      launches = ic_sample(p, first, count, ic, tune);
    else
      launches = lt_launch(p, first, count, tune);
    RIP_HIP(hipEventRecord(p->ev2, p->stream));
    launch_sum_u32(p->d_len + first, count, p->d_ctl + kCtlElems, p->stream);
    unsigned long long *ctl = p->h_ctl;
    const size_t ctl_bytes = sizeof(unsigned long long) * kCtlWords;
    RIP_HIP(hipMemcpyAsync(ctl, p->d_ctl, ctl_bytes, hipMemcpyDeviceToHost, p->stream));
    RIP_HIP(hipStreamSynchronize(p->stream));
    // LT: the walk kernel is the main one and the stats so far are its own;
    // IC colour groups alone: one kernel
    if (p->g->model == RIP_LT || ic.sampler == RIP_SAMPLER_COLOUR)
      for (int i = 0; i < RIP_STAT_COUNT; ++i) p->main_stats[i] = ctl[kCtlStatBytes + i];
    if (p->g->model == RIP_LT && ctl[kCtlOverflow] == 0 && ctl[kCtlLongCount] != 0) {
      // walks longer than the walk kernel's lists, then the sizes again
      launches += lt_finish_long(p, first, count, ctl[kCtlLongCount]);
      RIP_HIP(hipEventRecord(p->ev2, p->stream));
      hipLaunchKernelGGL(zero_elems_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
      launch_sum_u32(p->d_len + first, count, p->d_ctl + kCtlElems, p->stream);
      RIP_HIP(hipMemcpyAsync(ctl, p->d_ctl, ctl_bytes, hipMemcpyDeviceToHost, p->stream));
      RIP_HIP(hipStreamSynchronize(p->stream));
    }
    if (ctl[kCtlOverflow] == 0) {
      float ms = 0.0f, call = 0.0f;
      RIP_HIP(hipEventElapsedTime(&ms, p->ev0, p->ev1));
      RIP_HIP(hipEventElapsedTime(&call, p->ev0, p->ev2));
      p->last_kernel_ms = ms;
      p->last_call_ms = call;
      p->last_launches = launches;
      for (int i = 0; i < RIP_STAT_COUNT; ++i) p->last_stats[i] = ctl[kCtlStatBytes + i];
      // a wave's last output chunk may reach past cap_data with no write
      // there (only writes past it flag an overflow): every set lies below
      // cap_data, so the cursor is clamped to it
      const uint64_t bump = std::min<uint64_t>(ctl[kCtlDataBump], p->cap_data);
      const uint64_t added = bump - p->used_data;
      p->used_data = bump;
      p->num_elements += ctl[kCtlElems];
      p->groups += static_cast<uint32_t>(ctl[kCtlGroupCounter2]);
      p->avg_set_size = std::max(1.0, static_cast<double>(added) / static_cast<double>(count));
      p->sized = true;
      if (count >= kSizeHintMin) p->g->set_size_hint = p->avg_set_size;
      p->num_sets += count;
      return;
    }
    // grow to what the failed attempt asked for, with headroom, and redo
    const uint64_t want = ctl[kCtlDataBump];
    pool_reserve_data(p, std::max<uint64_t>(want + want / 4, 2 * p->cap_data));
  }
}

void pool_append_host(rip_pool *p, uint64_t count, const uint64_t *offsets,
                      const uint32_t *vertices) {
  if (count == 0) return;
  RIP_REQUIRE(offsets && offsets[0] == 0, -1, "append: offsets[0] must be 0");
  RIP_REQUIRE(p->num_sets + count < (1ull << 32), -1,
              "pool: at most 2^32-1 samples per stream");
  const uint64_t total = offsets[count];
  RIP_REQUIRE(total == 0 || vertices, -1, "append: vertices is NULL");
  std::vector<unsigned long long> off(count);
  std::vector<uint32_t> len(count);
  for (uint64_t i = 0; i < count; ++i) {
    RIP_REQUIRE(offsets[i + 1] >= offsets[i], -1, "append: offsets not monotone");
    RIP_REQUIRE(offsets[i + 1] - offsets[i] < 0xFFFFFFF0ull, -1, "append: set too large");
    for (uint64_t j = offsets[i]; j < offsets[i + 1]; ++j) {
      RIP_REQUIRE(vertices[j] < p->g->n, -1, "append: vertex id out of range");
      RIP_REQUIRE(j == offsets[i] || vertices[j - 1] < vertices[j], -1,
                  "append: sets must be sorted and duplicate-free");
    }
    off[i] = p->used_data + offsets[i];
    len[i] = static_cast<uint32_t>(offsets[i + 1] - offsets[i]);
  }
  RIP_HIP(hipSetDevice(p->g->device));
  pool_reserve_sets(p, p->num_sets + count);
  if (p->used_data + total > p->cap_data)
    pool_reserve_data(p, std::max<uint64_t>(p->used_data + total,
                                            p->cap_data + p->cap_data / 2));
  if (total)
    copy_sync(p, p->d_data + p->used_data, vertices, sizeof(uint32_t) * total,
                      hipMemcpyHostToDevice);
  copy_sync(p, p->d_off + p->num_sets, off.data(), sizeof(unsigned long long) * count,
                    hipMemcpyHostToDevice);
  copy_sync(p, p->d_len + p->num_sets, len.data(), sizeof(uint32_t) * count,
                    hipMemcpyHostToDevice);
  pool_clear_gref(p, p->num_sets, p->num_sets + count);
  p->num_sets += count;
  p->used_data += total;
  p->num_elements += total;
}

// Sets [first, first + count) in CSR form on the host.  A set held in the
// column store is its list (the members outside the hot vertices) merged
// with its hot members, both ascending.
void pool_export(rip_pool *p, uint64_t first, uint64_t count, uint64_t *offsets,
                 uint32_t *vertices, uint64_t cap) {
  RIP_REQUIRE(first + count <= p->num_sets && first + count >= first, -1,
              "export: range beyond the pool");
  RIP_REQUIRE(offsets != nullptr, -1, "export: offsets is NULL");
  RIP_HIP(hipSetDevice(p->g->device));
  // sets of the range held partly in columns: (index in the range, entry)
  std::vector<uint64_t> gidx;
  std::vector<uint32_t> gent;
  if (p->groups && p->d_gref && count) {
    std::vector<uint32_t> gref(count);
    copy_sync(p, gref.data(), p->d_gref + first, sizeof(uint32_t) * count,
                      hipMemcpyDeviceToHost);
    for (uint64_t i = 0; i < count; ++i)
      if (gref[i] != kNoGroup) {
        gidx.push_back(i);
        gent.push_back(gref[i]);
      }
  }
  auto *pre = dev_alloc<unsigned long long>(count + 1);
  uint32_t *out = nullptr, *d_ent = nullptr, *hot = nullptr;
  unsigned long long *hsz = nullptr;
  try {
    exclusive_scan_u32(p->d_len + first, pre, count, p->stream);
    std::vector<uint64_t> coff(count + 1);
    copy_sync(p, coff.data(), pre, sizeof(uint64_t) * (count + 1), hipMemcpyDeviceToHost);
    const uint64_t cold_total = coff[count];
    // hot member counts of the column-held sets
    const uint64_t ng = gidx.size();
    std::vector<unsigned long long> hoff(ng + 1, 0);
    if (ng) {
      d_ent = dev_alloc<uint32_t>(ng);
      hsz = dev_alloc<unsigned long long>(ng + 1);
      copy_sync(p, d_ent, gent.data(), sizeof(uint32_t) * ng, hipMemcpyHostToDevice);
      hipLaunchKernelGGL(hot_size_kernel, dim3(static_cast<uint32_t>(ng)), dim3(kHotBT), 0,
                         p->stream, p->d_colptr, d_ent, p->g->hot_n, hsz);
      RIP_HIP(hipGetLastError());
      std::vector<unsigned long long> sz(ng);
      RIP_HIP(hipMemcpyAsync(sz.data(), hsz, sizeof(unsigned long long) * ng,
                             hipMemcpyDeviceToHost, p->stream));
      RIP_HIP(hipStreamSynchronize(p->stream));
      for (uint64_t j = 0; j < ng; ++j) hoff[j + 1] = hoff[j] + sz[j];
    }
    // final offsets: list part plus hot part of each set
    offsets[0] = 0;
    for (uint64_t i = 0, j = 0; i < count; ++i) {
      uint64_t add = coff[i + 1] - coff[i];
      if (j < ng && gidx[j] == i) {
        add += hoff[j + 1] - hoff[j];
        ++j;
      }
      offsets[i + 1] = offsets[i] + add;
    }
    const uint64_t total = offsets[count];
    if (vertices) {
      RIP_REQUIRE(total <= cap, -1, "export: vertices buffer too small");
      out = dev_alloc<uint32_t>(cold_total);
      const uint64_t waves = (count + 63) / 64;
      if (waves) {
        hipLaunchKernelGGL(gather_sets, dim3(static_cast<uint32_t>((waves + 3) / 4)),
                           dim3(256), 0, p->stream, p->d_off, p->d_len, p->d_data,
                           first, count, pre, out);
        RIP_HIP(hipGetLastError());
      }
      std::vector<uint32_t> hot_h;
      if (ng) {
        hot = dev_alloc<uint32_t>(hoff[ng]);
        copy_sync(p, hsz, hoff.data(), sizeof(unsigned long long) * (ng + 1),
                          hipMemcpyHostToDevice);
        hipLaunchKernelGGL(hot_extract_kernel, dim3(static_cast<uint32_t>(ng)), dim3(kHotBT), 0,
                           p->stream, p->d_colptr, d_ent, p->g->hot_n, p->g->d_hot_vertex, hsz,
                           hot);
        RIP_HIP(hipGetLastError());
        hot_h.resize(hoff[ng]);
      }
      RIP_HIP(hipStreamSynchronize(p->stream));
      if (!ng) {
        if (total)
          copy_sync(p, vertices, out, sizeof(uint32_t) * total, hipMemcpyDeviceToHost);
      } else {
        std::vector<uint32_t> cold_h(cold_total);
        if (cold_total)
          copy_sync(p, cold_h.data(), out, sizeof(uint32_t) * cold_total,
                            hipMemcpyDeviceToHost);
        if (!hot_h.empty())
          copy_sync(p, hot_h.data(), hot, sizeof(uint32_t) * hot_h.size(),
                            hipMemcpyDeviceToHost);
        for (uint64_t i = 0, j = 0; i < count; ++i) {
          const uint32_t *c0 = cold_h.data() + coff[i], *c1 = cold_h.data() + coff[i + 1];
          if (j < ng && gidx[j] == i) {
            std::merge(c0, c1, hot_h.data() + hoff[j], hot_h.data() + hoff[j + 1],
                       vertices + offsets[i]);
            ++j;
          } else {
            std::copy(c0, c1, vertices + offsets[i]);
          }
        }
      }
    }
  } catch (...) {
    dev_free(pre);
    dev_free(out);
    dev_free(d_ent);
    dev_free(hsz);
    dev_free(hot);
    throw;
  }
  dev_free(pre);
  dev_free(out);
  dev_free(d_ent);
  dev_free(hsz);
  dev_free(hot);
}

}  // namespace rip
