// Lane sampler: one IC reverse BFS per lane, 64 samples per wave, for the
// many small sets of a launch.
//
// Every lane runs its own sample as a state machine, one step per loop
// round: pop the next vertex of its list and load its record, move to the
// row's next gap block, or draw one gap (ic_gap.h) and, for a live edge,
// read the in-neighbour and insert it into the lane's visited table.  Lanes
// whose sample finished take the next sample id in the same round, so a
// wave keeps 64 BFS of very uneven size in flight without waiting for its
// slowest.  Semantics are those of AddRRRSet (add_rrrset.h:105-157) under
// the stream contract (common.h, ic_gap.h): the set is the reverse closure
// of the root over live edges, identical to the other samplers' sets.
//
// Per lane (HBM): a list of `cap` vertices (BFS queue, then the set) and an
// open-addressing table of 2 cap 64-bit entries (epoch << 32 | vertex,
// never cleared: an entry counts only under the lane's current epoch).  A
// sample that would pass `cap` vertices is listed for the wave sampler.
// Finished sets of up to 16 vertices are sorted by the lane in registers
// and written through the wave's output chunk; longer ones are written
// unsorted, listed, and sorted by ic_lane_sort_kernel afterwards.
#include <algorithm>

#include "device_util.h"
#include "ic_common.h"
#include "ic_gap.h"
#include "internal.h"

namespace rip {
namespace {

constexpr uint32_t kLaneSmall = 16;     // sets sorted in registers
constexpr uint32_t kLaneClaim = 64;     // sample ids a wave claims at a time

struct LaneParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  Affine step;
  uint64_t seed;
  uint64_t first;
  uint64_t count;
  unsigned long long *table;   // lanes x (2 cap)
  uint32_t *list;              // lanes x cap
  uint32_t *epoch;             // per lane
  uint32_t cap, tmask;
  uint32_t *over_ids;          // samples left for the wave sampler (kCtlLongCount)
  uint32_t *sort_ids;          // sets written unsorted (kCtlSortCount)
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
  uint32_t chunk;              // pool elements a wave reserves at a time
};

// Insert v into a lane's own table (no other lane touches it): true if new.
__device__ inline bool lane_insert(unsigned long long *table, uint32_t mask,
                                   unsigned long long tag, uint32_t v) {
  uint32_t h = hash_vertex(v) & mask;
  for (;;) {
    const unsigned long long cur = table[h];
    if ((cur & 0xFFFFFFFF00000000ull) != tag) {
      table[h] = tag | v;
      return true;
    }
    if (static_cast<uint32_t>(cur) == v) return false;
    h = (h + 1) & mask;
  }
}

__device__ inline void cswap(uint32_t &a, uint32_t &b) {
  const uint32_t lo = min(a, b), hi = max(a, b);
  a = lo;
  b = hi;
}

// Sorts list[0, len) for len <= 16 in registers (bitonic network over 16
// slots padded with kEmptyVertex; every index is a compile-time constant).
__device__ inline void sort16(uint32_t *list, uint32_t len) {
  uint32_t r[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) r[i] = i < static_cast<int>(len) ? list[i] : kEmptyVertex;
#pragma unroll
  for (int k = 2; k <= 16; k <<= 1) {
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1) {
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int p = i ^ j;
        if (p > i) {
          if ((i & k) == 0)
            cswap(r[i], r[p]);
          else
            cswap(r[p], r[i]);
        }
      }
    }
  }
#pragma unroll
  for (int i = 0; i < 16; ++i)
    if (i < static_cast<int>(len)) list[i] = r[i];
}

enum : uint32_t { kIdle = 0, kPop = 1, kRow = 2, kDone = 3 };

template <int BT>
__global__ void __launch_bounds__(BT) ic_lane_kernel(LaneParams P) {
  __shared__ Affine s_p2[64];
  __shared__ IcLogTables s_tab;
  if (threadIdx.x == 0) fill_step_powers(P.step, s_p2);
  ic_load_tables(&s_tab);
  __syncthreads();
  const uint32_t lane = lane_id();
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * (BT / 64) + (threadIdx.x >> 6);
  const uint64_t gl = wave * 64 + lane;
  unsigned long long *table = P.table + gl * (static_cast<uint64_t>(P.tmask) + 1);
  uint32_t *list = P.list + gl * P.cap;
  uint32_t epoch = P.epoch[gl];
  const Affine Pst = s_p2[0];
  // wave-uniform: the claimed range of sample indices and the output chunk
  uint64_t next = 0, hi = 0;
  bool exhausted = false;
  unsigned long long out_cur = 0, out_end = 0;
  // lane state
  uint32_t state = kIdle;
  uint64_t idx = 0, z = 0;
  unsigned long long tag = 0;
  uint32_t head = 0, tail = 0;
  uint32_t rbegin = 0, deg = 0, T = 0, beta = 0, nblk = 0, k = 0, pos = 0, len = 0;
  uint64_t q = 0, rq = 0, ufull = 0, ulast = 0, xb = 0, x = 0;
  bool uniform = true;
  unsigned long long st_cand = 0, st_draws = 0, st_exp = 0, st_bytes = 0, st_live = 0;

  for (;;) {
    // idle lanes take the next ids of the wave's claimed range
    const uint64_t im = __ballot(state == kIdle);
    if (im) {
      if (next >= hi && !exhausted) {
        unsigned long long c = 0;
        if (lane == 0) c = atomicAdd(&P.ctl[kCtlGroupCounter], static_cast<unsigned long long>(kLaneClaim));
        c = uniform_u64(__shfl(c, 0, 64));
        if (c >= P.count) {
          exhausted = true;
        } else {
          next = c;
          hi = std::min<uint64_t>(c + kLaneClaim, P.count);
        }
      }
      if (state == kIdle) {
        const uint64_t my = next + __popcll(im & lanes_below());
        if (my < hi) {
          idx = my;
          z = sample_z(P.seed, P.first + idx);
          const uint32_t root = uniform_index(z, P.n);
          ++epoch;
          tag = static_cast<unsigned long long>(epoch) << 32;
          lane_insert(table, P.tmask, tag, root);
          list[0] = root;
          head = 0;
          tail = 1;
          state = kPop;
        } else if (exhausted) {
          state = kDone;
        }
      }
      next = std::min<uint64_t>(next + __popcll(im), hi);
    }
    if (__ballot(state != kDone) == 0) break;

    bool finish = false;
    if (state == kPop) {
      if (head < tail) {
        const uint32_t v = list[head++];
        const VertexRec rv = P.vrec[v];
        ++st_exp;
        T = rv.thr;
        if (T != 0) {
          rbegin = rv.begin;
          deg = rv.end - rv.begin;
          uniform = (rv.flags & kRowUniform) != 0;
          beta = (rv.flags >> kBetaShift) & 31u;
          q = rv.q;
          ufull = rv.ufull;
          ulast = rv.ulast;
          rq = rv.rq;
          nblk = ((deg - 1) >> beta) + 1;
          k = 0;
          xb = rv.ja * z + rv.jb;                  // the row's first element
          x = xb;
          pos = 0;
          len = min(deg, 1u << beta);
          ++st_draws;
          // a block without candidates costs one compare
          if ((xb >> 32) + 1 <= (len == (1u << beta) ? ufull : ulast)) len = 0;
          state = kRow;
        }
      } else {
        finish = true;
      }
    } else if (state == kRow) {
      if (pos >= len) {
        if (++k < nblk) {
          const Affine jb = s_p2[beta + 1];          // 2B elements: the next block
          xb = jb.a * xb + jb.b;
          x = xb;
          pos = 0;
          len = min(deg - (k << beta), 1u << beta);
          ++st_draws;
          if ((xb >> 32) + 1 <= (len == (1u << beta) ? ufull : ulast)) len = 0;
        } else {
          state = kPop;
        }
      } else {
        const uint64_t gap = ic_gap_fast(x, q, rq, &s_tab);
        if (gap >= len - pos) {
          pos = len;
        } else {
          pos += static_cast<uint32_t>(gap);
          const uint32_t e = rbegin + (k << beta) + pos;
          const uint64_t a = Pst.a * x + Pst.b;      // acceptance element
          ++st_cand;
          if (!uniform) {
            ++st_draws;
            st_bytes += 4;
          }
          const bool live = uniform || ic_accept(a, T, P.val[e]);
          ++pos;
          x = Pst.a * a + Pst.b;                     // next gap element
          if (pos < len) ++st_draws;
          if (live) {
            ++st_live;
            st_bytes += 4;
            const uint32_t u = P.ids[e];
            if (lane_insert(table, P.tmask, tag, u)) {
              if (tail == P.cap) {
                // too large for a lane: the wave sampler redoes it
                const unsigned long long b = atomicAdd(&P.ctl[kCtlLongCount], 1ull);
                P.over_ids[b] = static_cast<uint32_t>(idx);
                state = kIdle;
              } else {
                list[tail++] = u;
              }
            }
          }
        }
      }
    }

    // finished sets go out through the wave's output chunk
    if (__ballot(finish)) {
      const uint32_t flen = finish ? tail : 0u;
      if (finish && tail <= kLaneSmall) sort16(list, tail);
      const uint32_t incl = wave_inclusive_scan(flen);
      const uint32_t total = uniform_u32(__shfl(incl, 63, 64));
      if (out_end - out_cur < total) {
        const unsigned long long want = total > P.chunk ? total : P.chunk;
        unsigned long long nb = 0;
        if (lane == 0) nb = atomicAdd(&P.ctl[kCtlDataBump], want);
        nb = uniform_u64(__shfl(nb, 0, 64));
        out_cur = nb;
        out_end = nb + want;
      }
      const unsigned long long base = out_cur + incl - flen;
      out_cur += total;
      const uint64_t s = P.first + idx;
      bool unsorted = false;
      if (finish) {
        if (base + flen > P.cap_data) {
          atomicOr(&P.ctl[kCtlOverflow], 1ull);
          P.set_len[s] = kLenFail;
        } else {
          for (uint32_t i = 0; i < flen; ++i) P.data[base + i] = list[i];
          P.set_off[s] = base;
          P.set_len[s] = flen;
          st_bytes += 4ull * flen;
          unsorted = flen > kLaneSmall;
        }
        state = kIdle;
      }
      // longer sets are sorted afterwards: one counter bump per wave
      const uint64_t um = __ballot(unsorted);
      if (um) {
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(um)) - 1;
        unsigned long long b = 0;
        if (lane == leader) b = atomicAdd(&P.ctl[kCtlSortCount], static_cast<unsigned long long>(__popcll(um)));
        b = __shfl(b, static_cast<int>(leader), 64);
        if (unsorted) P.sort_ids[b + __popcll(um & lanes_below())] = static_cast<uint32_t>(idx);
      }
    }
  }
  P.epoch[gl] = epoch;
  st_bytes += kIcRecBytes * st_exp;
  report_stats(P.ctl, st_bytes, st_cand, st_draws, st_exp, st_live);
}

// Sorts the sets the lane kernel wrote unsorted (17 .. cap vertices), one
// wave per set, bitonic in LDS.
template <int BT, int CAP>
__global__ void __launch_bounds__(BT) ic_lane_sort_kernel(const uint32_t *sort_ids,
                                                          const unsigned long long *ctl,
                                                          uint64_t first,
                                                          const unsigned long long *set_off,
                                                          const uint32_t *set_len,
                                                          uint32_t *data) {
  __shared__ uint32_t s_a[BT / 64][CAP];
  const uint32_t lane = lane_id(), wid = threadIdx.x >> 6;
  const uint64_t nsort = ctl[kCtlSortCount];
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (BT / 64);
  uint32_t *a = s_a[wid];
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * (BT / 64) + wid; i < nsort; i += waves) {
    const uint64_t s = first + sort_ids[i];
    const unsigned long long off = set_off[s];
    const uint32_t len = set_len[s];
    uint32_t cap2 = 32;
    while (cap2 < len) cap2 <<= 1;
    for (uint32_t j = lane; j < len; j += 64) a[j] = data[off + j];
    wave_bitonic(a, len, cap2);
    for (uint32_t j = lane; j < len; j += 64) data[off + j] = a[j];
    __builtin_amdgcn_wave_barrier();
  }
}

__global__ void reset_lane_kernel(unsigned long long *ctl) {
  ctl[kCtlGroupCounter] = 0;
  ctl[kCtlLongCount] = 0;
  ctl[kCtlSortCount] = 0;
}

constexpr int kLaneBT = 256;

}  // namespace

uint32_t ic_lane_cap(const rip_tune *tune) {
  const uint32_t c = tune && tune->queue_cap ? tune->queue_cap : kLaneCapDefault;
  RIP_REQUIRE(c == 64 || c == 256 || c == 1024 || c == 2048 || c == 4096, -1,
              "tune: queue_cap must be 64, 256, 1024, 2048 or 4096");
  return c;
}

// Threads per workgroup of the lane kernel (tune->block; 256 by default).
static uint32_t ic_lane_block(const rip_tune *tune) {
  const uint32_t b = tune && tune->block ? tune->block : kLaneBT;
  RIP_REQUIRE(b == 256 || b == 512 || b == 1024, -1, "tune: block must be 256, 512 or 1024");
  return b;
}

static const void *lane_kernel_ptr(uint32_t bt) {
  switch (bt) {
    case 512: return reinterpret_cast<const void *>(&ic_lane_kernel<512>);
    case 1024: return reinterpret_cast<const void *>(&ic_lane_kernel<1024>);
    default: return reinterpret_cast<const void *>(&ic_lane_kernel<256>);
  }
}

// Lanes of a lane-sampler launch of `count` samples: every resident lane,
// at most one per sample, within 1/16 of the device memory; a multiple of
// the block size bt.
static uint64_t lane_lanes(uint64_t count, uint32_t cap, uint32_t bt) {
  const int per_cu = blocks_per_cu(lane_kernel_ptr(bt), static_cast<int>(bt));
  size_t free_b = 0, total_b = 0;
  RIP_HIP(hipMemGetInfo(&free_b, &total_b));
  const uint64_t per_lane = 20ull * cap + 4;       // table 16 cap + list 4 cap + epoch
  uint64_t lanes = static_cast<uint64_t>(per_cu) * cu_count() * bt;
  lanes = std::min<uint64_t>(lanes, std::max<uint64_t>(bt, (total_b / 16) / per_lane));
  lanes = std::min<uint64_t>(lanes, (count + bt - 1) / bt * bt);
  return std::max<uint64_t>(bt, lanes / bt * bt);
}

uint64_t ic_lane_slack(const rip_pool *p, uint64_t count, const rip_tune *tune) {
  const uint64_t waves = lane_lanes(count, ic_lane_cap(tune), ic_lane_block(tune)) / 64;
  const uint32_t chunk = tune && tune->out_chunk ? tune->out_chunk : 0;
  return waves * (chunk ? chunk : 4096u);
}

// Samples [first, first + count) with the lane kernel; returns how many
// outgrew a lane (their local indices are in g->wave.d_over).
uint64_t ic_lane_launch(rip_pool *p, uint64_t first, uint64_t count, const rip_tune *tune) {
  rip_graph *g = p->g;
  WaveScratch &w = g->wave;
  const uint32_t cap = ic_lane_cap(tune);
  const uint32_t bt = ic_lane_block(tune);
  const uint64_t lanes = lane_lanes(count, cap, bt);
  if (w.lane_lanes < lanes || w.lane_cap != cap) {
    dev_free(w.d_ltable);
    dev_free(w.d_llist);
    dev_free(w.d_lepoch);
    w.lane_lanes = 0;
    w.d_ltable = dev_alloc<unsigned long long>(lanes * 2 * cap);
    w.d_llist = dev_alloc<uint32_t>(lanes * cap);
    w.d_lepoch = dev_alloc<uint32_t>(lanes);
    RIP_HIP(hipMemsetAsync(w.d_ltable, 0, sizeof(unsigned long long) * lanes * 2 * cap, p->stream));
    RIP_HIP(hipMemsetAsync(w.d_lepoch, 0, sizeof(uint32_t) * lanes, p->stream));
    w.lane_lanes = lanes;
    w.lane_cap = cap;
  }
  if (w.cap_over < count) {
    dev_free(w.d_over);
    dev_free(w.d_lsort);
    w.cap_over = 0;
    w.d_over = dev_alloc<uint32_t>(count);
    w.d_lsort = dev_alloc<uint32_t>(count);
    w.cap_over = count;
  }
  LaneParams P;
  P.vrec = g->d_vrec;
  P.ids = g->d_ids;
  P.val = g->d_val;
  P.n = g->n;
  P.step = g->step;
  P.seed = p->seed;
  P.first = first;
  P.count = count;
  P.table = w.d_ltable;
  P.list = w.d_llist;
  P.epoch = w.d_lepoch;
  P.cap = cap;
  P.tmask = 2 * cap - 1;
  P.over_ids = w.d_over;
  P.sort_ids = w.d_lsort;
  P.set_off = p->d_off;
  P.set_len = p->d_len;
  P.data = p->d_data;
  P.cap_data = p->cap_data;
  P.ctl = p->d_ctl;
  const uint64_t waves = lanes / 64;
  P.chunk = tune && tune->out_chunk ? tune->out_chunk
                                     : out_chunk_rule(static_cast<double>(count) *
                                                          std::min(p->avg_set_size, 1024.0) /
                                                          static_cast<double>(waves),
                                                      4096);
  hipLaunchKernelGGL(reset_lane_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  RIP_HIP(hipEventRecord(p->ev0, p->stream));
  const dim3 grid(static_cast<uint32_t>(lanes / bt));
  switch (bt) {
    case 512: hipLaunchKernelGGL((ic_lane_kernel<512>), grid, dim3(512), 0, p->stream, P); break;
    case 1024: hipLaunchKernelGGL((ic_lane_kernel<1024>), grid, dim3(1024), 0, p->stream, P); break;
    default: hipLaunchKernelGGL((ic_lane_kernel<256>), grid, dim3(256), 0, p->stream, P); break;
  }
  RIP_HIP(hipGetLastError());
  RIP_HIP(hipEventRecord(p->ev1, p->stream));
  switch (cap) {
    case 64:
      hipLaunchKernelGGL((ic_lane_sort_kernel<256, 64>), dim3(1024), dim3(256), 0, p->stream,
                         w.d_lsort, p->d_ctl, first, p->d_off, p->d_len, p->d_data);
      break;
    case 256:
      hipLaunchKernelGGL((ic_lane_sort_kernel<256, 256>), dim3(1024), dim3(256), 0, p->stream,
                         w.d_lsort, p->d_ctl, first, p->d_off, p->d_len, p->d_data);
      break;
    case 1024:
      hipLaunchKernelGGL((ic_lane_sort_kernel<256, 1024>), dim3(1024), dim3(256), 0, p->stream,
                         w.d_lsort, p->d_ctl, first, p->d_off, p->d_len, p->d_data);
      break;
    case 2048:
      hipLaunchKernelGGL((ic_lane_sort_kernel<256, 2048>), dim3(1024), dim3(256), 0, p->stream,
                         w.d_lsort, p->d_ctl, first, p->d_off, p->d_len, p->d_data);
      break;
    default:   // 4096: 64 KB of LDS per workgroup
      hipLaunchKernelGGL((ic_lane_sort_kernel<256, 4096>), dim3(1024), dim3(256), 0, p->stream,
                         w.d_lsort, p->d_ctl, first, p->d_off, p->d_len, p->d_data);
      break;
  }
  RIP_HIP(hipGetLastError());
  pool_read_ctl(p, true);
  return p->h_ctl[kCtlLongCount];
}

}  // namespace rip
