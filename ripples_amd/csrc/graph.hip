// Device copy of a transposed CSR graph (GraphBwd), laid out for the reverse
// traversals of the samplers:
//   d_vrec[v] : 32-byte record {P^(1+begin) jump, begin, end} -- one aligned
//               load gives a vertex's row and the jump to its first coin
//   d_adj[e]  : (in-neighbour, IC integer threshold | LT float weight bits),
//               padded with two zero edges for pairwise 16-byte loads
#include <cstring>
#include <vector>

#include "internal.h"

namespace rip {
namespace {

__global__ void build_vrec_kernel(const unsigned long long *index, uint32_t n,
                                  Affine step, VertexRec *vrec) {
  uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= n) return;
  uint64_t b = index[v], e = index[v + 1];
  Affine j = aff_pow(step, 1 + b);
  VertexRec r;
  r.ja = j.a;
  r.jb = j.b;
  r.begin = static_cast<uint32_t>(b);
  r.end = static_cast<uint32_t>(e);
  r.pad0 = 0;
  r.pad1 = 0;
  vrec[v] = r;
}

}  // namespace

void graph_create(rip_graph *g, int device, int model, uint32_t n, uint64_t m,
                  const uint64_t *index, const uint32_t *edges,
                  const float *weights) {
  RIP_REQUIRE(model == RIP_IC || model == RIP_LT, -1, "graph: unknown model");
  RIP_REQUIRE(n > 0, -1, "graph: num_nodes must be > 0");
  RIP_REQUIRE(index != nullptr, -1, "graph: index is NULL");
  RIP_REQUIRE(m == 0 || (edges != nullptr && weights != nullptr), -1,
              "graph: edges/weights are NULL");
  RIP_REQUIRE(m < (1ull << 32), -1, "graph: more than 2^32-1 edges");
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
  // two zero edges of padding: 16-byte loads read edges in aligned pairs
  std::vector<uint2> adj(m + 2, make_uint2(0, 0));
  for (uint64_t e = 0; e < m; ++e) {
    RIP_REQUIRE(edges[e] < n, -1, "graph: edge endpoint out of range");
    uint32_t y;
    if (model == RIP_IC) {
      y = ic_threshold(weights[e]);
    } else {
      std::memcpy(&y, &weights[e], 4);
    }
    adj[e] = make_uint2(edges[e], y);
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
  try {
    g->d_vrec = dev_alloc<VertexRec>(n);
    g->d_adj = dev_alloc<uint2>(m + 2);
    d_index = dev_alloc<unsigned long long>(static_cast<size_t>(n) + 1);
    RIP_HIP(hipMemcpy(d_index, index, sizeof(uint64_t) * (static_cast<size_t>(n) + 1),
                      hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(g->d_adj, adj.data(), sizeof(uint2) * (m + 2),
                      hipMemcpyHostToDevice));
    dim3 grid((n + 255) / 256);
    hipLaunchKernelGGL(build_vrec_kernel, grid, dim3(256), 0, 0, d_index, n,
                       g->step, g->d_vrec);
    RIP_HIP(hipGetLastError());
    RIP_HIP(hipDeviceSynchronize());
  } catch (...) {
    dev_free(d_index);
    graph_destroy(g);
    throw;
  }
  dev_free(d_index);
}

void graph_destroy(rip_graph *g) {
  if (!g) return;
  (void)hipSetDevice(g->device);
  dev_free(g->d_vrec);
  dev_free(g->d_adj);
  dev_free(g->ic.d_vp);
  dev_free(g->ic.d_bm0);
  dev_free(g->ic.d_bm1);
  dev_free(g->ic.d_queue);
  dev_free(g->ic.d_hub);
  dev_free(g->wave.d_table);
  dev_free(g->wave.d_list);
  dev_free(g->wave.d_epoch);
  dev_free(g->wave.d_big);
  dev_free(g->lt.d_long_ids);
  dev_free(g->lt.d_spill);
}

}  // namespace rip
