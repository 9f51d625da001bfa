// Device copy of a transposed CSR graph (GraphBwd), laid out for the reverse
// traversals of the samplers:
//   d_vrec[v] : 64-byte record {jump to the vertex's first stream element,
//               begin, end, IC gap parameters of the row (ic_gap.h), flags}
//               -- one aligned line gives everything a row expansion needs
//               before its candidate edges
//   d_val[e]  : IC integer threshold | LT float weight bits
//   d_ids[e]  : in-neighbour (read only for live / chosen edges)
//   d_lrec[v] : LT walk record (common.h), one 128-byte line per vertex
//   d_bp      : LT rows in blocks of 16 edges, one aligned 128-byte line
//               each: the 16 row breakpoints of the float chain (lt_scan.h)
//               then the 16 in-neighbours, so the block a walk step searches
//               also holds the chosen in-neighbour.  Edge j of the row of v is
//               slot p % 16 of block p / 16 for p = boff(v) + j; rows are
//               padded to whole blocks (breakpoint 0xFFFFFFFF).
//   d_bidx    : LT fanout-16 index over the breakpoints of rows longer than 16
// The id and value arrays carry kEdgePad zero elements for 16-byte loads.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ic_gap.h"
#include "internal.h"
#include "lt_scan.h"

namespace rip {
namespace {

// IC: the row's gap parameters from its largest threshold (ic_gap.h);
// LT: the jump to the threshold element and the breakpoint index offset.
__global__ void build_vrec_kernel(const unsigned long long *index, const uint32_t *thr,
                                  const uint32_t *flags, uint32_t n, Affine step, int model,
                                  VertexRec *vrec) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint64_t b = index[v], e = index[v + 1];
  Affine j = aff_pow(step, model == RIP_IC ? 1 + 2 * b : 1 + b);
  VertexRec r;
  r.ja = j.a;
  r.jb = j.b;
  r.q = 0;
  r.begin = static_cast<uint32_t>(b);
  r.end = static_cast<uint32_t>(e);
  r.thr = thr[v];
  r.ufull = 0;
  r.ulast = 0;
  r.rq = 0;
  r.flags = flags[v];
  if (model == RIP_IC && r.thr != 0) {
    const uint32_t beta = ic_block_log2(r.thr);
    r.q = ic_gap_q(r.thr);
    r.rq = ic_gap_recip(r.q);
    r.ufull = ic_full_skip(r.q, beta);
    // the last block holds d - (nblk - 1) B edges
    const uint32_t d = static_cast<uint32_t>(e - b);
    const uint32_t last = d == 0 ? 0u : d - (((d - 1) >> beta) << beta);
    r.ulast = ic_skip_len(r.q, last);
    r.flags |= beta << kBetaShift;
  }
  for (int i = 0; i < 2; ++i) r.pad[i] = 0;
  vrec[v] = r;
}

__global__ void build_lrec_kernel(const VertexRec *vrec, const uint32_t *boff, uint32_t n,
                                  LtRec *lrec) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  const VertexRec r = vrec[v];
  LtRec o;
  o.ja = r.ja;
  o.jb = r.jb;
  o.boff = boff[v];
  o.d = r.end - r.begin;
  o.ioff = r.thr;
  o.flags = r.flags;
  for (int i = 0; i < 16; ++i) o.top[i] = 0xFFFFFFFFu;
  o.begin = r.begin;
  for (int i = 0; i < 7; ++i) o.pad[i] = 0;
  lrec[v] = o;
}

// One thread per edge e of an LT row: the row breakpoint B[e - begin] of
// lt_scan.h (kLtNever in rows with irregular weights) and the in-neighbour,
// in slot p % 16 of block p / 16 for p = boff + e - begin.  Threads of a
// wave take consecutive edges, so they read the same weights in step.
__global__ void build_bp_kernel(const unsigned long long *index, const uint32_t *val,
                                const uint32_t *ids, const VertexRec *vrec, const LtRec *lrec,
                                uint32_t n, uint64_t m, uint32_t *bp) {
  const uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= m) return;
  uint32_t lo = 0, hi = n;  // row of e: the last v with index[v] <= e
  while (hi - lo > 1) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (index[mid] <= e)
      lo = mid;
    else
      hi = mid;
  }
  const VertexRec r = vrec[lo];
  const uint64_t p = lrec[lo].boff + (e - r.begin);
  uint32_t *blk = bp + 32 * (p >> 4);
  blk[16 + (p & 15)] = ids[e];
  if (r.flags & kRowScalar) {
    blk[p & 15] = kLtNever;
    return;
  }
  blk[p & 15] = lt_breakpoint(val + r.begin, static_cast<uint32_t>(e) - r.begin);
}

// One thread per indexed LT row: the fanout-16 index over its breakpoints,
// top level first, and a copy of the top level in the walk record.
__global__ void build_bidx_kernel(const uint32_t *rows, uint32_t nrows, LtRec *lrec,
                                  const uint32_t *bp, uint32_t *bidx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  LtRec &r = lrec[rows[i]];
  const uint32_t *rows_bp = bp + 2ull * r.boff;  // the row's blocks (boff is a multiple of 16)
  lt_index_build([&](uint32_t j) { return rows_bp[32 * (j >> 4) + (j & 15)]; }, r.d,
                 bidx + r.ioff);
  for (int t = 0; t < 16; ++t) r.top[t] = bidx[r.ioff + t];
}

// key[u] += the IC threshold of every transposed entry with source u: the
// expected number of live out-edges of u, in units of 2^-24 (integer, so
// the sums do not depend on the order of the atomics).
__global__ void out_weight_kernel(const uint32_t *ids, const uint32_t *val, uint64_t m,
                                  unsigned long long *key) {
  for (uint64_t e = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; e < m;
       e += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    if (val[e]) atomicAdd(&key[ids[e]], static_cast<unsigned long long>(val[e]));
}

}  // namespace

// The column store's hot vertices: a quarter of the vertices (a multiple of
// 64, at least min(n, 64)), those most likely to be reached from others,
// ranked by the sum of their out-edge thresholds.  On R-MAT scale 24 with
// U[0,0.02] weights they hold 96% of the members of the giant sets.
void graph_ensure_hot(rip_graph *g, hipStream_t st) {
  if (g->d_hot_index) return;
  RIP_REQUIRE(g->model == RIP_IC, -1, "graph: the column store is for IC graphs");
  const uint32_t n = g->n;
  const uint32_t hot = std::max<uint32_t>(std::min<uint32_t>(n, 64u), (n / 4) & ~63u);
  std::vector<unsigned long long> key(n);
  unsigned long long *d_key = dev_alloc<unsigned long long>(n);
  try {
    RIP_HIP(hipMemsetAsync(d_key, 0, sizeof(unsigned long long) * n, st));
    if (g->m)
      hipLaunchKernelGGL(out_weight_kernel, dim3(2048), dim3(256), 0, st, g->d_ids, g->d_val,
                         g->m, d_key);
    RIP_HIP(hipGetLastError());
    RIP_HIP(hipMemcpyAsync(key.data(), d_key, sizeof(unsigned long long) * n,
                           hipMemcpyDeviceToHost, st));
    RIP_HIP(hipStreamSynchronize(st));
  } catch (...) {
    dev_free(d_key);
    throw;
  }
  dev_free(d_key);
  std::vector<uint32_t> order(n);
  for (uint32_t v = 0; v < n; ++v) order[v] = v;
  auto better = [&](uint32_t a, uint32_t b) { return key[a] != key[b] ? key[a] > key[b] : a < b; };
  if (hot < n) std::nth_element(order.begin(), order.begin() + hot, order.end(), better);
  std::vector<uint32_t> index(n, kNotHot), vertex(order.begin(), order.begin() + hot);
  std::sort(vertex.begin(), vertex.end());
  for (uint32_t h = 0; h < hot; ++h) index[vertex[h]] = h;
  uint32_t *d_index = dev_alloc<uint32_t>(n);
  uint32_t *d_vertex = nullptr;
  try {
    d_vertex = dev_alloc<uint32_t>(hot);
    RIP_HIP(hipMemcpyAsync(d_index, index.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice,
                           st));
    RIP_HIP(hipMemcpyAsync(d_vertex, vertex.data(), sizeof(uint32_t) * hot,
                           hipMemcpyHostToDevice, st));
    RIP_HIP(hipStreamSynchronize(st));
  } catch (...) {
    dev_free(d_index);
    dev_free(d_vertex);
    throw;
  }
  g->d_hot_index = d_index;
  g->d_hot_vertex = d_vertex;
  g->hot_n = hot;
}

void graph_create(rip_graph *g, int device, int model, uint32_t n, uint64_t m,
                  const uint64_t *index, const uint32_t *edges,
                  const float *weights) {
  RIP_REQUIRE(model == RIP_IC || model == RIP_LT, -1, "graph: unknown model");
  RIP_REQUIRE(n > 0, -1, "graph: num_nodes must be > 0");
  RIP_REQUIRE(index != nullptr, -1, "graph: index is NULL");
  RIP_REQUIRE(m == 0 || (edges != nullptr && weights != nullptr), -1,
              "graph: edges/weights are NULL");
  RIP_REQUIRE(m < (1ull << 32) - kEdgePad, -1, "graph: more than 2^32-5 edges");
  RIP_REQUIRE(index[0] == 0 && index[n] == m, -1,
              "graph: index[0] must be 0 and index[n] the edge count");
  uint32_t maxdeg = 0;
  uint64_t deg_ge[33] = {};
  for (uint32_t v = 0; v < n; ++v) {
    RIP_REQUIRE(index[v + 1] >= index[v], -1, "graph: index is not monotone");
    uint64_t d = index[v + 1] - index[v];
    if (d > maxdeg) maxdeg = static_cast<uint32_t>(d);
    for (int b = 0; b < 33 && (1ull << b) <= d; ++b) ++deg_ge[b];
  }
  std::vector<uint32_t> ids(m + kEdgePad, 0), val(m + kEdgePad, 0);
  for (uint64_t e = 0; e < m; ++e) {
    RIP_REQUIRE(edges[e] < n, -1, "graph: edge endpoint out of range");
    ids[e] = edges[e];
    if (model == RIP_IC) {
      val[e] = ic_threshold(weights[e]);
    } else {
      std::memcpy(&val[e], &weights[e], 4);
    }
  }
  // IC: each row's largest threshold, and whether every edge has it (then
  // a candidate needs no acceptance draw); LT rows holding a weight without
  // a breakpoint form are marked
  std::vector<uint32_t> uthr(n, 0u), flags(n, 0u);
  for (uint32_t v = 0; v < n; ++v) {
    const uint64_t b = index[v], e = index[v + 1];
    if (b == e) continue;
    if (model == RIP_IC) {
      uint32_t T = 0;
      bool same = true;
      for (uint64_t i = b; i < e; ++i) {
        T = std::max(T, val[i]);
        same = same && val[i] == val[b];
      }
      uthr[v] = T;
      if (same) flags[v] |= kRowUniform;
    } else {
      for (uint64_t i = b; i < e; ++i)
        if (lt_weight_irregular(val[i])) {
          flags[v] |= kRowScalar;
          break;
        }
    }
  }

  // LT rows longer than 16 edges (without irregular weights) get a breakpoint
  // index; its offset takes the place of the IC threshold in the record
  std::vector<uint32_t> index_rows;
  uint64_t index_size = 0;
  if (model == RIP_LT) {
    for (uint32_t v = 0; v < n; ++v) {
      const uint64_t d = index[v + 1] - index[v];
      if (d <= kLtFan || (flags[v] & kRowScalar)) continue;
      RIP_REQUIRE(index_size + lt_index_size(static_cast<uint32_t>(d)) < (1ull << 32), -1,
                  "graph: LT breakpoint index exceeds 2^32 entries");
      index_rows.push_back(v);
      uthr[v] = static_cast<uint32_t>(index_size);
      index_size += lt_index_size(static_cast<uint32_t>(d));
    }
  }
  int ndev = 0;
  RIP_HIP(hipGetDeviceCount(&ndev));
  RIP_REQUIRE(device >= 0 && device < ndev, -1, "graph: no such device");
  RIP_HIP(hipSetDevice(device));

  g->device = device;
  g->model = model;
  g->n = n;
  g->m = m;
  g->max_degree = maxdeg;
  g->step = sample_step();
  for (int b = 0; b < 33; ++b) g->deg_ge_pow2[b] = deg_ge[b];

  unsigned long long *d_index = nullptr;
  uint32_t *d_uthr = nullptr, *d_flags = nullptr, *d_rows = nullptr, *d_boff = nullptr;
  try {
    g->d_vrec = dev_alloc<VertexRec>(n);
    g->d_ids = dev_alloc<uint32_t>(m + kEdgePad);
    g->d_val = dev_alloc<uint32_t>(m + kEdgePad);
    d_index = dev_alloc<unsigned long long>(static_cast<size_t>(n) + 1);
    d_uthr = dev_alloc<uint32_t>(n);
    d_flags = dev_alloc<uint32_t>(n);
    RIP_HIP(hipMemcpy(d_index, index, sizeof(uint64_t) * (static_cast<size_t>(n) + 1),
                      hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(d_uthr, uthr.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(d_flags, flags.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(g->d_ids, ids.data(), sizeof(uint32_t) * (m + kEdgePad),
                      hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(g->d_val, val.data(), sizeof(uint32_t) * (m + kEdgePad),
                      hipMemcpyHostToDevice));
    dim3 grid((n + 255) / 256);
    hipLaunchKernelGGL(build_vrec_kernel, grid, dim3(256), 0, 0, d_index, d_uthr, d_flags,
                       n, g->step, model, g->d_vrec);
    RIP_HIP(hipGetLastError());
    if (model == RIP_LT && m > 0) {
      // padded so a search may read 16-byte groups past the last row
      // pair offsets: every row padded to whole blocks of 16 pairs
      std::vector<uint32_t> boff(n);
      uint64_t pairs = 0;
      for (uint32_t v = 0; v < n; ++v) {
        boff[v] = static_cast<uint32_t>(pairs);
        pairs += lt_pad16(static_cast<uint32_t>(index[v + 1] - index[v]));
        RIP_REQUIRE(pairs + kBpPad < (1ull << 32), -1, "graph: more than 2^32 LT pairs");
      }
      d_boff = dev_alloc<uint32_t>(n);
      RIP_HIP(hipMemcpy(d_boff, boff.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
      g->d_lrec = dev_alloc<LtRec>(n);
      hipLaunchKernelGGL(build_lrec_kernel, grid, dim3(256), 0, 0, g->d_vrec, d_boff, n,
                         g->d_lrec);
      g->d_bp = dev_alloc<uint32_t>(2 * (pairs + kBpPad));
      RIP_HIP(hipMemset(g->d_bp, 0xFF, sizeof(uint32_t) * 2 * (pairs + kBpPad)));
      hipLaunchKernelGGL(build_bp_kernel, dim3(static_cast<uint32_t>((m + 255) / 256)),
                         dim3(256), 0, 0, d_index, g->d_val, g->d_ids, g->d_vrec, g->d_lrec, n,
                         m, g->d_bp);
      RIP_HIP(hipGetLastError());
      if (!index_rows.empty()) {
        const uint32_t nrows = static_cast<uint32_t>(index_rows.size());
        g->d_bidx = dev_alloc<uint32_t>(index_size);
        d_rows = dev_alloc<uint32_t>(nrows);
        RIP_HIP(hipMemcpy(d_rows, index_rows.data(), sizeof(uint32_t) * nrows,
                          hipMemcpyHostToDevice));
        hipLaunchKernelGGL(build_bidx_kernel, dim3((nrows + 255) / 256), dim3(256), 0, 0, d_rows,
                           nrows, g->d_lrec, g->d_bp, g->d_bidx);
        RIP_HIP(hipGetLastError());
      }
    }
    RIP_HIP(hipDeviceSynchronize());
  } catch (...) {
    dev_free(d_index);
    dev_free(d_uthr);
    dev_free(d_flags);
    dev_free(d_rows);
    dev_free(d_boff);
    graph_destroy(g);
    throw;
  }
  dev_free(d_index);
  dev_free(d_uthr);
  dev_free(d_flags);
  dev_free(d_rows);
  dev_free(d_boff);
}

void graph_destroy(rip_graph *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  dev_free(g->d_vrec);
  dev_free(g->d_hot_index);
  dev_free(g->d_hot_vertex);
  dev_free(g->d_ids);
  dev_free(g->d_val);
  dev_free(g->ic.d_vp);
  dev_free(g->ic.d_bm0);
  dev_free(g->ic.d_bm1);
  dev_free(g->ic.d_queue);
  dev_free(g->ic.d_hub);
  dev_free(g->wave.d_table);
  dev_free(g->wave.d_list);
  dev_free(g->wave.d_epoch);
  dev_free(g->wave.d_big);
  dev_free(g->wave.d_bbits);
  dev_free(g->wave.d_bsum);
  dev_free(g->wave.d_blist);
  dev_free(g->wave.d_ltable);
  dev_free(g->wave.d_llist);
  dev_free(g->wave.d_lepoch);
  dev_free(g->wave.d_over);
  dev_free(g->wave.d_lsort);
  grid_scratch_free(g->grid);
  dev_free(g->lt.d_long_ids);
  dev_free(g->lt.d_spill);
  dev_free(g->d_bp);
  dev_free(g->d_lrec);
  dev_free(g->d_bidx);
  dev_free(g->fwd.d_rec);
  dev_free(g->fwd.d_dst);
  dev_free(g->fwd.d_val);
  dev_free(g->fwd.d_scratch);
}

}  // namespace rip
