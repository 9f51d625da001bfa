// extern "C" boundary: argument checks, exception -> status code mapping and
// the thread-local error message.  The reference reports failures by
// asserts, `throw` and ignored HIP status codes
// (include/ripples/gpu/gpu_runtime_trait.h:168-243); here every failure is a
// negative return value with a message.
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "internal.h"

namespace rip {
namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string &msg) { g_last_error = msg; }

template <typename F>
int guarded(F &&f) {
  try {
    f();
    g_last_error.clear();
    return RIP_OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::bad_alloc &) {
    g_last_error = "host allocation failed";
    return RIP_ENOMEM;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return RIP_EINVAL;
  } catch (...) {
    g_last_error = "unknown error";
    return RIP_EINVAL;
  }
}

}  // namespace rip

using rip::Error;
using rip::guarded;

extern "C" {

const char *rip_last_error(void) { return rip::g_last_error.c_str(); }

int rip_version(void) { return 1; }

int rip_device_count(int *count) {
  return guarded([&] {
    RIP_REQUIRE(count, RIP_EINVAL, "count is NULL");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
  });
}

int rip_graph_create(int device, int model, uint32_t num_nodes,
                     uint64_t num_edges, const uint64_t *index,
                     const uint32_t *edges, const float *weights,
                     rip_graph **out) {
  return guarded([&] {
    RIP_REQUIRE(out, RIP_EINVAL, "out is NULL");
    *out = nullptr;
    auto *g = new rip_graph();
    try {
      rip::graph_create(g, device, model, num_nodes, num_edges, index, edges, weights);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int rip_graph_destroy(rip_graph *g) {
  return guarded([&] {
    if (!g) return;
    rip::graph_destroy(g);
    delete g;
  });
}

int rip_graph_info(const rip_graph *g, uint32_t *num_nodes, uint64_t *num_edges,
                   int *model) {
  return guarded([&] {
    RIP_REQUIRE(g, RIP_EINVAL, "graph is NULL");
    if (num_nodes) *num_nodes = g->n;
    if (num_edges) *num_edges = g->m;
    if (model) *model = g->model;
  });
}

int rip_pool_create(rip_graph *g, uint64_t seed, rip_pool **out) {
  return guarded([&] {
    RIP_REQUIRE(g && out, RIP_EINVAL, "graph or out is NULL");
    *out = nullptr;
    auto *p = new rip_pool();
    try {
      rip::pool_create(p, g, seed);
    } catch (...) {
      rip::pool_destroy(p);
      delete p;
      throw;
    }
    *out = p;
  });
}

int rip_pool_destroy(rip_pool *p) {
  return guarded([&] {
    if (!p) return;
    rip::pool_destroy(p);
    delete p;
  });
}

int rip_pool_sample(rip_pool *p, uint64_t count, const rip_tune *tune) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    rip::pool_sample(p, count, tune);
  });
}

int rip_pool_append_host(rip_pool *p, uint64_t count, const uint64_t *offsets,
                         const uint32_t *vertices) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    rip::pool_append_host(p, count, offsets, vertices);
  });
}

int rip_autotune(rip_graph *g, uint64_t probe_sets, int force, rip_tune *best,
                 double *best_ms, uint32_t *tried) {
  return guarded([&] {
    RIP_REQUIRE(g && best, RIP_EINVAL, "graph or best is NULL");
    rip::autotune(g, probe_sets, force != 0, best, best_ms, tried);
  });
}

int rip_pool_reset(rip_pool *p) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    p->num_sets = 0;
    p->used_data = 0;
    p->num_elements = 0;
    p->groups = 0;
  });
}

int rip_pool_columns(const rip_pool *p, uint32_t *groups, uint32_t *hot_nodes,
                     uint64_t *column_bytes) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    if (groups) *groups = p->groups;
    if (hot_nodes) *hot_nodes = p->g->hot_n;
    if (column_bytes) *column_bytes = 8ull * p->g->hot_n * p->cap_groups;
  });
}

int rip_pool_device_bytes(const rip_pool *p, uint64_t *bytes) {
  return guarded([&] {
    RIP_REQUIRE(p && bytes, RIP_EINVAL, "pool or bytes is NULL");
    *bytes = rip::pool_device_bytes(p);
  });
}

int rip_pool_size(const rip_pool *p, uint64_t *num_sets, uint64_t *num_elements) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    if (num_sets) *num_sets = p->num_sets;
    if (num_elements) *num_elements = p->num_elements;
  });
}

int rip_pool_export(rip_pool *p, uint64_t first, uint64_t count, uint64_t *offsets,
                    uint32_t *vertices, uint64_t cap) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    rip::pool_export(p, first, count, offsets, vertices, cap);
  });
}

int rip_pool_select(rip_pool *p, uint64_t num_sets, uint32_t k, uint32_t *seeds,
                    uint32_t *num_seeds, uint64_t *covered) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    RIP_REQUIRE(k == 0 || seeds, RIP_EINVAL, "seeds is NULL");
    RIP_REQUIRE(num_sets <= p->num_sets, RIP_EINVAL, "num_sets beyond the pool");
    uint32_t got = rip::select_run(p, num_sets, k, seeds, covered);
    if (num_seeds) *num_seeds = got;
  });
}

int rip_pool_count(rip_pool *p, uint64_t num_sets, uint32_t *counts) {
  return guarded([&] {
    RIP_REQUIRE(p && counts, RIP_EINVAL, "pool or counts is NULL");
    RIP_REQUIRE(num_sets <= p->num_sets, RIP_EINVAL, "num_sets beyond the pool");
    rip::count_run(p, num_sets, counts);
  });
}

int rip_pool_coverage(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                      uint32_t num_seeds, uint64_t *covered) {
  return guarded([&] {
    RIP_REQUIRE(p && covered, RIP_EINVAL, "pool or covered is NULL");
    RIP_REQUIRE(num_seeds == 0 || seeds, RIP_EINVAL, "seeds is NULL");
    RIP_REQUIRE(num_sets <= p->num_sets, RIP_EINVAL, "num_sets beyond the pool");
    *covered = rip::coverage_run(p, num_sets, seeds, num_seeds);
  });
}

int rip_pool_coverage_noalloc(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                              uint32_t num_seeds, uint64_t *covered) {
  return guarded([&] {
    RIP_REQUIRE(p && covered, RIP_EINVAL, "pool or covered is NULL");
    RIP_REQUIRE(num_seeds == 0 || seeds, RIP_EINVAL, "seeds is NULL");
    *covered = rip::coverage_run_noalloc(p, num_sets, seeds, num_seeds);
  });
}

int rip_pool_last_kernel_ms(const rip_pool *p, double *ms, uint32_t *launches) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    if (ms) *ms = p->last_kernel_ms;
    if (launches) *launches = p->last_launches;
  });
}

int rip_pool_last_grid(const rip_pool *p, double *ms, uint32_t *levels, uint64_t *stats,
                       uint32_t count) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    RIP_REQUIRE(count == 0 || stats, RIP_EINVAL, "stats is NULL");
    if (ms) *ms = p->grid_ms;
    if (levels) *levels = p->grid_levels;
    const uint64_t *before = p->wave_ran ? p->wave_stats : p->main_stats;
    for (uint32_t i = 0; i < count; ++i)
      stats[i] = i < RIP_STAT_COUNT && p->grid_levels ? p->last_stats[i] - before[i] : 0;
  });
}

int rip_pool_last_wave(const rip_pool *p, double *ms, uint64_t *stats, uint32_t count) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    RIP_REQUIRE(count == 0 || stats, RIP_EINVAL, "stats is NULL");
    if (ms) *ms = p->wave_ran ? p->wave_ms : 0.0;
    for (uint32_t i = 0; i < count; ++i)
      stats[i] = i < RIP_STAT_COUNT && p->wave_ran ? p->wave_stats[i] - p->main_stats[i] : 0;
  });
}

int rip_pool_last_call_ms(const rip_pool *p, double *ms) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    RIP_REQUIRE(ms, RIP_EINVAL, "ms is NULL");
    *ms = p->last_call_ms;
  });
}

int rip_pool_last_kernel_stats(const rip_pool *p, uint64_t *stats, uint32_t count) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    RIP_REQUIRE(count == 0 || stats, RIP_EINVAL, "stats is NULL");
    for (uint32_t i = 0; i < count; ++i)
      stats[i] = i < RIP_STAT_COUNT ? p->main_stats[i] : 0;
  });
}

int rip_pool_last_stats(const rip_pool *p, uint64_t *stats, uint32_t count) {
  return guarded([&] {
    RIP_REQUIRE(p, RIP_EINVAL, "pool is NULL");
    RIP_REQUIRE(count == 0 || stats, RIP_EINVAL, "stats is NULL");
    for (uint32_t i = 0; i < count; ++i)
      stats[i] = i < RIP_STAT_COUNT ? p->last_stats[i] : 0;
  });
}

int rip_imm(rip_graph *g, uint32_t k, double epsilon, uint64_t seed,
            const rip_tune *tune, uint32_t *seeds, rip_imm_record *rec) {
  return guarded([&] {
    RIP_REQUIRE(g, RIP_EINVAL, "graph is NULL");
    RIP_REQUIRE(k == 0 || seeds, RIP_EINVAL, "seeds is NULL");
    rip::imm_run(g, k, epsilon, seed, tune, seeds, rec);
  });
}

int rip_opimc(rip_graph *g, uint32_t k, double epsilon, double delta, uint64_t seed,
              const rip_tune *tune, uint32_t *seeds, rip_opimc_record *rec) {
  return guarded([&] {
    RIP_REQUIRE(g && seeds, RIP_EINVAL, "graph or seeds is NULL");
    rip::opimc_run(g, k, epsilon, delta, seed, tune, seeds, rec);
  });
}

uint64_t rip_opimc_r2_seed(uint64_t seed) { return rip::opimc_r2_seed(seed); }

int rip_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b,
                   double c, uint64_t seed, uint64_t *num_edges, uint32_t *src,
                   uint32_t *dst, uint64_t cap) {
  return guarded([&] {
    RIP_REQUIRE(num_edges, RIP_EINVAL, "num_edges is NULL");
    std::vector<uint32_t> s, d;
    rip::rmat_edges(scale, edge_factor, a, b, c, seed, s, d);
    *num_edges = s.size();
    if (!src) return;
    RIP_REQUIRE(dst, RIP_EINVAL, "dst is NULL");
    RIP_REQUIRE(cap >= s.size(), RIP_EINVAL, "edge buffers too small");
    std::memcpy(src, s.data(), sizeof(uint32_t) * s.size());
    std::memcpy(dst, d.data(), sizeof(uint32_t) * d.size());
  });
}

int rip_csr_from_edges(uint32_t num_nodes, uint64_t num_edges, const uint32_t *src,
                       const uint32_t *dst, int scheme, float scale, uint64_t seed,
                       uint64_t *index, uint32_t *edges, float *weights) {
  return guarded([&] {
    RIP_REQUIRE(index && (num_edges == 0 || (src && dst && edges && weights)),
                RIP_EINVAL, "NULL buffer");
    rip::csr_from_edges(num_nodes, num_edges, src, dst, scheme, scale, seed, index,
                        edges, weights);
  });
}

int rip_spread(rip_graph *g, const uint32_t *seeds, uint32_t num_seeds, uint64_t seed,
               uint64_t first, uint64_t count, uint64_t *activated, uint32_t *levels) {
  return guarded([&] {
    RIP_REQUIRE(g, RIP_EINVAL, "graph is NULL");
    rip::spread_run(g, seeds, num_seeds, seed, first, count, activated, levels);
  });
}

int rip_server_create(uint32_t slots, uint32_t workers, uint64_t budget_bytes,
                      rip_server **out) {
  return guarded([&] {
    RIP_REQUIRE(out, RIP_EINVAL, "out is NULL");
    *out = rip::server_create(slots, workers, budget_bytes);
  });
}

int rip_server_add_graph(rip_server *s, rip_graph *g, uint32_t *graph_id) {
  return guarded([&] {
    RIP_REQUIRE(s, RIP_EINVAL, "server is NULL");
    const uint32_t id = rip::server_add_graph(s, g);
    if (graph_id) *graph_id = id;
  });
}

int rip_server_submit(rip_server *s, uint32_t graph_id, uint32_t k, double epsilon,
                      uint64_t seed, uint64_t *query_id) {
  return guarded([&] {
    RIP_REQUIRE(s && query_id, RIP_EINVAL, "server or query_id is NULL");
    *query_id = rip::server_submit(s, graph_id, k, epsilon, seed);
  });
}

int rip_server_poll(rip_server *s, uint64_t query_id, double wait_ms, int *done, uint32_t *seeds,
                    uint32_t cap, uint32_t *num_seeds, rip_imm_record *rec, double *latency_ms) {
  return guarded([&] {
    RIP_REQUIRE(s && done, RIP_EINVAL, "server or done is NULL");
    *done = rip::server_poll(s, query_id, wait_ms, seeds, cap, num_seeds, rec, latency_ms) ? 1 : 0;
  });
}

int rip_server_stats(rip_server *s, uint64_t *out, uint32_t count) {
  return guarded([&] {
    RIP_REQUIRE(s && (out || count == 0), RIP_EINVAL, "server or out is NULL");
    rip::server_stats(s, out, count);
  });
}

int rip_server_destroy(rip_server *s) {
  return guarded([&] {
    if (s) rip::server_destroy(s);
  });
}

int rip_load_edge_list(const char *path, const rip_load_options *opts, rip_host_graph **out) {
  return guarded([&] {
    RIP_REQUIRE(out && opts, RIP_EINVAL, "opts or out is NULL");
    *out = nullptr;
    auto *g = new rip_host_graph();
    try {
      rip::load_edge_list(path, *opts, g);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int rip_load_binary(const char *path, float scale_factor, rip_host_graph **out) {
  return guarded([&] {
    RIP_REQUIRE(out, RIP_EINVAL, "out is NULL");
    *out = nullptr;
    auto *g = new rip_host_graph();
    try {
      rip::load_binary(path, scale_factor, g);
    } catch (...) {
      delete g;
      throw;
    }
    *out = g;
  });
}

int rip_dump_binary(const rip_host_graph *g, const char *path) {
  return guarded([&] {
    RIP_REQUIRE(g, RIP_EINVAL, "graph is NULL");
    rip::dump_binary(g, path);
  });
}

int rip_host_graph_info(const rip_host_graph *g, uint32_t *num_nodes, uint64_t *num_edges) {
  return guarded([&] {
    RIP_REQUIRE(g, RIP_EINVAL, "graph is NULL");
    if (num_nodes) *num_nodes = static_cast<uint32_t>(g->reverse_map.size());
    if (num_edges) *num_edges = g->edges.size();
  });
}

int rip_host_graph_arrays(const rip_host_graph *g, const uint64_t **index, const uint32_t **edges,
                          const float **weights, const uint32_t **reverse_map) {
  return guarded([&] {
    RIP_REQUIRE(g, RIP_EINVAL, "graph is NULL");
    if (index) *index = g->index.data();
    if (edges) *edges = g->edges.data();
    if (weights) *weights = g->weights.data();
    if (reverse_map) *reverse_map = g->reverse_map.data();
  });
}

int rip_host_graph_destroy(rip_host_graph *g) {
  return guarded([&] { delete g; });
}

}  // extern "C"
