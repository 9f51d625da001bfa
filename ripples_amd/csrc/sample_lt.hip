// LT sampler: one reverse random walk per lane.
//
// Semantics of AddRRRSet under linear_threshold_tag
// (include/ripples/add_rrrset.h:136-150): at vertex v draw one threshold
// U in (0,1], subtract the in-edge weights in CSR order with float
// arithmetic, and take the first in-neighbour at which the threshold drops
// to <= 0; stop when no edge is taken or the taken vertex was already
// visited.  The threshold of (sample s, vertex v) is element 1 + begin(v)
// of sample s's stream, reached with the vertex record's jump.
//
// Walks up to CAP vertices keep their visited list in LDS (column-major, so
// lanes hit distinct banks) and are sorted in place; the reference's GPU
// kernel (include/ripples/gpu/lt_kernel.h:59-130) caps walks at 8 vertices
// and finishes longer ones on the host.  Here longer walks are finished on
// the device by lt_long_kernel, which keeps a visited list and bitmap per
// lane in HBM and heap-sorts the list before writing it out.
#include <algorithm>

#include "device_util.h"
#include "internal.h"

namespace rip {
namespace {

constexpr uint32_t kLenFail = 0xFFFFFFFFu;
constexpr uint32_t kLenLong = 0xFFFFFFFEu;

struct LtParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  uint64_t seed;
  uint64_t first;
  uint64_t count;
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
  uint32_t *long_ids;
  uint32_t *spill;
};

// Persistent walkers: every lane runs its own walk and, when it ends, claims
// the next sample id from a device counter (one atomic per wave and round),
// so short walks never wait for the long ones of the same wave.  Each round
// scans up to UNROLL in-edges of the lane's current row with 16-byte loads
// (two edges each) issued together, then reduces the float threshold in CSR
// order.  A walk's visited list lives in LDS for its first CAP vertices and
// in a per-lane HBM spill area for the next SPILL; longer walks are handed to
// lt_long_kernel.
template <int BT, int CAP, int SPILL, int UNROLL>
__global__ void __launch_bounds__(BT) lt_walk_kernel(LtParams P) {
  static_assert(UNROLL % 4 == 0, "groups of four weights per load");
  __shared__ uint32_t s_list[CAP * BT];
  uint32_t *list = s_list + threadIdx.x;  // element i < CAP at list[i * BT]
  uint32_t *spill = P.spill + (static_cast<uint64_t>(blockIdx.x) * BT + threadIdx.x) * SPILL;
  auto at = [&](uint32_t i) -> uint32_t & {
    return i < static_cast<uint32_t>(CAP) ? list[i * BT] : spill[i - CAP];
  };
  const uint4 *val4 = reinterpret_cast<const uint4 *>(P.val);
  const uint32_t lane = lane_id();
  bool busy = false, exhausted = false;
  uint64_t gid = 0, z = 0;
  uint32_t len = 0, e = 0, end = 0;
  float t = 0.0f;
  unsigned long long bytes = 0;  // 8 per visited vertex, 8 per scanned edge, 4 per output
  unsigned long long edges = 0;

  for (;;) {
    // ---- refill idle lanes
    const bool want = !busy && !exhausted;
    const uint64_t wm = __ballot(want);
    if (wm) {
      const uint32_t leader = __ffsll(static_cast<unsigned long long>(wm)) - 1;
      unsigned long long base = 0;
      if (lane == leader)
        base = atomicAdd(&P.ctl[kCtlGroupCounter], static_cast<unsigned long long>(__popcll(wm)));
      base = __shfl(base, static_cast<int>(leader), 64);
      if (want) {
        gid = base + __popcll(wm & lanes_below());
        if (gid < P.count) {
          z = sample_z(P.seed, P.first + gid);
          const uint32_t root = uniform_index(z, P.n);
          list[0] = root;
          len = 1;
          const VertexRec r = P.vrec[root];
          t = u01(r.ja * z + r.jb);
          e = r.begin;
          end = r.end;
          bytes += 8;
          busy = true;
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(busy) == 0) break;

    // ---- one scan round
    bool done = false, is_long = false;
    if (busy) {
      if (e >= end) {
        done = true;  // no in-edge crosses the threshold: the walk stops
      } else {
        const uint32_t base = e & ~3u;      // 16-byte aligned group of 4 weights
        const uint32_t last_quad = (end - 1) >> 2;
        uint4 q[UNROLL / 4];
#pragma unroll
        for (int kk = 0; kk < UNROLL / 4; ++kk)
          q[kk] = val4[min((base >> 2) + kk, last_quad)];
        const uint32_t lo = e - base;       // 0..3
        const uint32_t hi = min(static_cast<uint32_t>(UNROLL), end - base);
        bool found = false;
        uint32_t stop = hi;
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
          const uint4 &g4 = q[k >> 2];
          const uint32_t wbits = (k & 3) == 0 ? g4.x : (k & 3) == 1 ? g4.y
                                 : (k & 3) == 2 ? g4.z : g4.w;
          if (!found && static_cast<uint32_t>(k) >= lo && static_cast<uint32_t>(k) < hi) {
            t -= __uint_as_float(wbits);
            if (!(t > 0)) {
              found = true;
              stop = k + 1;
            }
          }
        }
        const uint32_t scanned = stop - lo;
        bytes += 8ull * scanned;
        edges += scanned;
        e = base + stop;
        if (found) {
          const uint32_t u = P.ids[e - 1];  // the chosen in-neighbour
          bool seen = false;
          for (uint32_t i = 0; i < len; ++i)
            if (at(i) == u) { seen = true; break; }
          if (seen) {
            done = true;
          } else if (len == CAP + SPILL) {
            done = true;
            is_long = true;
          } else {
            at(len) = u;
            ++len;
            const VertexRec r = P.vrec[u];
            t = u01(r.ja * z + r.jb);
            e = r.begin;
            end = r.end;
            bytes += 8;
          }
        } else if (e >= end) {
          done = true;
        }
      }
    }

    // ---- retire finished walks: sort, reserve pool space per wave, write
    if (__ballot(done)) {
      const bool emit = done && !is_long;
      if (emit) {
        for (uint32_t i = 1; i < len; ++i) {
          const uint32_t key = at(i);
          uint32_t j = i;
          while (j > 0 && at(j - 1) > key) {
            at(j) = at(j - 1);
            --j;
          }
          at(j) = key;
        }
      }
      const uint32_t size = emit ? len : 0;
      const uint32_t incl = wave_inclusive_scan(size);
      const uint32_t total = __shfl(incl, 63, 64);
      unsigned long long base = 0;
      if (lane == 0 && total)
        base = atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(total));
      base = __shfl(base, 0, 64);
      const bool fail = total && base + total > P.cap_data;
      if (fail && lane == 0) atomicOr(&P.ctl[kCtlOverflow], 1ull);
      if (done) {
        const uint64_t s = P.first + gid;
        if (is_long) {
          const unsigned long long idx = atomicAdd(&P.ctl[kCtlLongCount], 1ull);
          P.long_ids[idx] = static_cast<uint32_t>(gid);
          P.set_len[s] = kLenLong;
        } else if (fail) {
          P.set_len[s] = kLenFail;
        } else {
          const unsigned long long off = base + incl - size;
          for (uint32_t i = 0; i < len; ++i) P.data[off + i] = at(i);
          P.set_off[s] = off;
          P.set_len[s] = len;
          bytes += 4ull * len;
        }
        busy = false;
        len = 0;
      }
    }
  }
  for (int d = 32; d > 0; d >>= 1) {
    bytes += __shfl_down(bytes, d, 64);
    edges += __shfl_down(edges, d, 64);
  }
  if (lane == 0 && bytes) atomicAdd(&P.ctl[kCtlStatBytes], bytes);
  if (lane == 0 && edges) atomicAdd(&P.ctl[kCtlStatEdges], edges);
}

template <int BT, int CAP, int SPILL, int UNROLL>
uint32_t lt_grid(uint64_t count) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &per_cu, reinterpret_cast<const void *>(&lt_walk_kernel<BT, CAP, SPILL, UNROLL>), BT, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  int dev = 0, cus = 1;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  const uint64_t resident = static_cast<uint64_t>(per_cu) * static_cast<uint64_t>(cus);
  const uint64_t needed = (count + BT - 1) / BT;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min(resident, needed)));
}

// Unbounded walks: one lane per walk, visited list + bitmap in HBM.
__device__ inline void sift_down(uint32_t *a, uint32_t root, uint32_t end) {
  for (;;) {
    uint32_t child = 2 * root + 1;
    if (child >= end) return;
    if (child + 1 < end && a[child] < a[child + 1]) ++child;
    if (a[root] >= a[child]) return;
    const uint32_t t = a[root];
    a[root] = a[child];
    a[child] = t;
    root = child;
  }
}

__global__ void lt_long_kernel(LtParams P, uint32_t num_long, uint32_t lo,
                               uint32_t lanes, uint32_t *lists,
                               uint32_t *bitmaps) {
  const uint32_t lane_g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t idx = lo + lane_g;
  if (lane_g >= lanes || idx >= num_long) return;
  const uint64_t nw = (P.n + 31) / 32;
  uint32_t *list = lists + static_cast<uint64_t>(lane_g) * P.n;
  uint32_t *bm = bitmaps + static_cast<uint64_t>(lane_g) * nw;
  const uint64_t s = P.first + P.long_ids[idx];
  const uint64_t z = sample_z(P.seed, s);
  uint32_t v = uniform_index(z, P.n);
  uint32_t len = 0;
  list[len++] = v;
  bm[v >> 5] |= 1u << (v & 31);
  for (;;) {
    const VertexRec r = P.vrec[v];
    float t = u01(r.ja * z + r.jb);
    uint32_t next = kLenFail;
    for (uint32_t e = r.begin; e < r.end; ++e) {
      t -= __uint_as_float(P.val[e]);
      if (t > 0) continue;
      const uint32_t u = P.ids[e];
      if (!((bm[u >> 5] >> (u & 31)) & 1u)) next = u;
      break;
    }
    if (next == kLenFail) break;
    list[len++] = next;
    bm[next >> 5] |= 1u << (next & 31);
    v = next;
  }
  for (uint32_t i = 0; i < len; ++i) bm[list[i] >> 5] = 0;
  for (uint32_t i = len / 2; i-- > 0;) sift_down(list, i, len);
  for (uint32_t end = len; end-- > 1;) {
    const uint32_t t = list[0];
    list[0] = list[end];
    list[end] = t;
    sift_down(list, 0, end);
  }
  const unsigned long long base =
      atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(len));
  if (base + len > P.cap_data) {
    atomicOr(&P.ctl[kCtlOverflow], 1ull);
    P.set_len[s] = kLenFail;
    return;
  }
  for (uint32_t i = 0; i < len; ++i) P.data[base + i] = list[i];
  P.set_off[s] = base;
  P.set_len[s] = len;
}

__global__ void reset_long_kernel(unsigned long long *ctl) {
  ctl[kCtlLongCount] = 0;
}

}  // namespace

uint32_t lt_launch(rip_pool *p, uint64_t first, uint64_t count,
                   const rip_tune *tune) {
  (void)tune;
  constexpr int BT = 256, CAP = 32, SPILL = 96, UNROLL = 32;
  const rip_graph *g = p->g;
  LtParams P;
  P.vrec = g->d_vrec;
  P.ids = g->d_ids;
  P.val = g->d_val;
  P.n = g->n;
  P.seed = p->seed;
  P.first = first;
  P.count = count;
  P.set_off = p->d_off;
  P.set_len = p->d_len;
  P.data = p->d_data;
  P.cap_data = p->cap_data;
  P.ctl = p->d_ctl;
  // long-walk ids: at most `count`
  if (p->g->lt.cap_ids < count) {
    dev_free(p->g->lt.d_long_ids);
    p->g->lt.d_long_ids = dev_alloc<uint32_t>(count);
    p->g->lt.cap_ids = count;
  }
  P.long_ids = p->g->lt.d_long_ids;
  hipLaunchKernelGGL(reset_long_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  const uint32_t blocks = lt_grid<BT, CAP, SPILL, UNROLL>(count);
  const uint64_t spill_words = static_cast<uint64_t>(blocks) * BT * SPILL;
  if (p->g->lt.cap_spill < spill_words) {
    dev_free(p->g->lt.d_spill);
    p->g->lt.d_spill = dev_alloc<uint32_t>(spill_words);
    p->g->lt.cap_spill = spill_words;
  }
  P.spill = p->g->lt.d_spill;
  RIP_HIP(hipEventRecord(p->ev0, p->stream));
  hipLaunchKernelGGL((lt_walk_kernel<BT, CAP, SPILL, UNROLL>), dim3(blocks), dim3(BT), 0,
                     p->stream, P);
  RIP_HIP(hipGetLastError());
  RIP_HIP(hipEventRecord(p->ev1, p->stream));
  uint32_t launches = 1;
  unsigned long long nlong = 0;
  RIP_HIP(hipMemcpyAsync(&nlong, p->d_ctl + kCtlLongCount, sizeof(nlong),
                         hipMemcpyDeviceToHost, p->stream));
  RIP_HIP(hipStreamSynchronize(p->stream));
  if (nlong) {
    // lanes of unbounded walks: a visited list (n words) and bitmap each
    const uint64_t per_lane = 4ull * g->n + 4ull * ((g->n + 31) / 32);
    const uint64_t lanes = std::max<uint64_t>(
        1, std::min<uint64_t>({nlong, 256, (static_cast<uint64_t>(1) << 29) / per_lane}));
    uint32_t *lists = dev_alloc<uint32_t>(lanes * g->n);
    uint32_t *bms = dev_alloc<uint32_t>(lanes * ((g->n + 31) / 32));
    RIP_HIP(hipMemsetAsync(bms, 0, sizeof(uint32_t) * lanes * ((g->n + 31) / 32),
                           p->stream));
    for (uint64_t lo = 0; lo < nlong; lo += lanes) {
      const uint32_t blocks_l = static_cast<uint32_t>((lanes + 63) / 64);
      hipLaunchKernelGGL(lt_long_kernel, dim3(blocks_l), dim3(64), 0, p->stream,
                         P, static_cast<uint32_t>(nlong), static_cast<uint32_t>(lo),
                         static_cast<uint32_t>(lanes), lists, bms);
      RIP_HIP(hipGetLastError());
      ++launches;
    }
    RIP_HIP(hipStreamSynchronize(p->stream));
    dev_free(lists);
    dev_free(bms);
  }
  return launches;
}

}  // namespace rip
