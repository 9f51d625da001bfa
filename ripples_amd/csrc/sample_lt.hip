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
  const uint2 *adj;
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
};

template <int BT, int CAP>
__global__ void __launch_bounds__(BT) lt_walk_kernel(LtParams P) {
  __shared__ uint32_t s_list[CAP * BT];
  const uint32_t tid = threadIdx.x;
  const uint64_t gid = static_cast<uint64_t>(blockIdx.x) * BT + tid;
  const bool active = gid < P.count;
  const uint64_t s = P.first + gid;
  uint32_t len = 0;
  bool is_long = false;
  unsigned long long bytes = 0;  // algorithmic bytes: 8 per vertex + 8 per scanned edge
  if (active) {
    const uint64_t z = sample_z(P.seed, s);
    uint32_t v = uniform_index(z, P.n);
    s_list[tid] = v;
    len = 1;
    for (;;) {
      const VertexRec r = P.vrec[v];
      float t = u01(r.ja * z + r.jb);
      uint32_t next = kLenFail;
      bytes += 8;
      for (uint32_t e = r.begin; e < r.end; ++e) {
        const uint2 a = P.adj[e];
        bytes += 8;
        t -= __uint_as_float(a.y);
        if (t > 0) continue;
        bool seen = false;
        for (uint32_t i = 0; i < len; ++i)
          if (s_list[i * BT + tid] == a.x) { seen = true; break; }
        if (!seen) next = a.x;
        break;
      }
      if (next == kLenFail) break;
      if (len == CAP) { is_long = true; break; }
      s_list[len * BT + tid] = next;
      ++len;
      v = next;
    }
    if (!is_long) {
      for (uint32_t i = 1; i < len; ++i) {
        const uint32_t key = s_list[i * BT + tid];
        uint32_t j = i;
        while (j > 0 && s_list[(j - 1) * BT + tid] > key) {
          s_list[j * BT + tid] = s_list[(j - 1) * BT + tid];
          --j;
        }
        s_list[j * BT + tid] = key;
      }
    }
  }
  // one pool reservation per wave
  const uint32_t size = (active && !is_long) ? len : 0;
  const uint32_t incl = wave_inclusive_scan(size);
  const uint32_t total = __shfl(incl, 63, 64);
  unsigned long long base = 0;
  if (lane_id() == 0 && total)
    base = atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(total));
  base = __shfl(base, 0, 64);
  const bool fail = total && base + total > P.cap_data;
  if (fail && lane_id() == 0) atomicOr(&P.ctl[kCtlOverflow], 1ull);
  bytes += 4ull * size;
  for (int d = 32; d > 0; d >>= 1) bytes += __shfl_down(bytes, d, 64);
  if (lane_id() == 0 && bytes) atomicAdd(&P.ctl[kCtlStatBytes], bytes);
  if (!active) return;
  if (is_long) {
    const unsigned long long idx = atomicAdd(&P.ctl[kCtlLongCount], 1ull);
    P.long_ids[idx] = static_cast<uint32_t>(gid);
    P.set_len[s] = kLenLong;
  } else if (fail) {
    P.set_len[s] = kLenFail;
  } else {
    const unsigned long long off = base + incl - size;
    for (uint32_t i = 0; i < len; ++i) P.data[off + i] = s_list[i * BT + tid];
    P.set_off[s] = off;
    P.set_len[s] = len;
  }
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
      const uint2 a = P.adj[e];
      t -= __uint_as_float(a.y);
      if (t > 0) continue;
      if (!((bm[a.x >> 5] >> (a.x & 31)) & 1u)) next = a.x;
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
  constexpr int BT = 128, CAP = 64;
  const rip_graph *g = p->g;
  LtParams P;
  P.vrec = g->d_vrec;
  P.adj = g->d_adj;
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
  if (p->lt.cap_ids < count) {
    dev_free(p->lt.d_long_ids);
    p->lt.d_long_ids = dev_alloc<uint32_t>(count);
    p->lt.cap_ids = count;
  }
  P.long_ids = p->lt.d_long_ids;
  hipLaunchKernelGGL(reset_long_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  const uint64_t blocks = (count + BT - 1) / BT;
  RIP_HIP(hipEventRecord(p->ev0, p->stream));
  hipLaunchKernelGGL((lt_walk_kernel<BT, CAP>), dim3(static_cast<uint32_t>(blocks)),
                     dim3(BT), 0, p->stream, P);
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
        1, std::min<uint64_t>(nlong, (static_cast<uint64_t>(1) << 30) / per_lane));
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
