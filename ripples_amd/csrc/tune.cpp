// Autotuner of the samplers: the settings that draw a probe batch fastest on
// a graph, remembered for the graph and, keyed by its degree skew, for graphs
// alike.  The reference picks its GPU batch size by a similar probe
// (streaming_rrr_generator.h:1150-1241, called at tools/imm.cc:238-244);
// every setting tried here gives the same sets, so only the time differs.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "internal.h"

namespace rip {
namespace {

constexpr uint64_t kProbeBase = 16384;   // sets of the first probe (probe_sets = 0)
constexpr double kProbeTargetMs = 20.0;    // time per candidate the probe is scaled to

std::mutex g_tune_mu;
std::map<std::tuple<int, int, int>, rip_tune> g_tune_cache;

// (model, log2 of the largest in-degree over the mean, log2 n): graphs alike
// in these share their settings.
std::tuple<int, int, int> skew_key(const rip_graph *g) {
  const double mean = g->n ? static_cast<double>(g->m) / g->n : 0.0;
  const double skew = mean > 0 ? g->max_degree / mean : 1.0;
  return std::make_tuple(g->model, static_cast<int>(std::lround(std::log2(std::max(1.0, skew)))),
                         static_cast<int>(std::lround(std::log2(std::max(1.0, double(g->n))))));
}

// IC: the lane list capacity 64..4096 (how large a sample grows before the grid
// sampler takes it) and the lane kernel's workgroup size, and three settings
// with the wave stage between the lanes and the grid; LT: the output chunk of
// the walk kernel's waves.
std::vector<rip_tune> candidates(int model) {
  std::vector<rip_tune> c;
  if (model == RIP_IC) {
    for (uint32_t cap : {64u, 256u, 1024u, 2048u, 4096u})
      for (uint32_t block : {256u, 512u, 1024u}) {
        rip_tune t{};
        t.queue_cap = cap;
        t.block = block;
        c.push_back(t);
      }
    // the wave stage between the lanes and the grid sampler
    for (uint32_t cap : {64u, 256u, 1024u}) {
      rip_tune t{};
      t.queue_cap = cap;
      t.block = 256;
      t.wave_stage = 1;
      c.push_back(t);
    }
  } else {
    for (uint32_t chunk : {0u, 256u, 1024u, 4096u}) {
      rip_tune t{};
      t.out_chunk = chunk;
      c.push_back(t);
    }
  }
  return c;
}

void pool_clear(rip_pool *p) {
  p->num_sets = 0;
  p->used_data = 0;
  p->num_elements = 0;
  p->groups = 0;
}

}  // namespace

void autotune(rip_graph *g, uint64_t probe, bool force, rip_tune *best, double *best_ms,
              uint32_t *tried) {
  const auto key = skew_key(g);
  if (!force) {
    std::lock_guard<std::mutex> hold(g_tune_mu);
    const auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      *best = it->second;
      if (best_ms) *best_ms = 0.0;
      if (tried) *tried = 0;
      std::lock_guard<std::mutex> graph_hold(g->sampler_mu);
      g->tuned = it->second;
      g->tuned_set = true;
      return;
    }
  }
  const std::vector<rip_tune> cand = candidates(g->model);
  rip_pool pool;
  pool_create(&pool, g, 0x5eedull);
  if (probe == 0) {
    // about 20 ms of sampling per candidate: a probe much shorter than the
    // device's fill (hundreds of thousands of lanes) would favour settings
    // that only win on tiny batches
    try {
      double t = 0.0;
      for (int rep = 0; rep < 2; ++rep) {
        pool_clear(&pool);
        const auto t0 = std::chrono::steady_clock::now();
        pool_sample(&pool, kProbeBase, nullptr);
        t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
      }
      const double f = std::floor(kProbeTargetMs / std::max(t, 1e-3));
      probe = kProbeBase * static_cast<uint64_t>(std::min(16.0, std::max(1.0, f)));
    } catch (...) {
      pool_destroy(&pool);
      throw;
    }
  }
  double best_t = 0.0;
  rip_tune pick{};
  uint32_t extra = 0;   // settings timed after the lane candidates
  try {
    for (size_t i = 0; i < cand.size(); ++i) {
      double t = 0.0;
      for (int rep = 0; rep < 2; ++rep) {   // the first run grows the scratch
        pool_clear(&pool);
        const auto t0 = std::chrono::steady_clock::now();
        pool_sample(&pool, probe, &cand[i]);
        t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                .count();
      }
      if (i == 0 || t < best_t) {
        best_t = t;
        pick = cand[i];
      }
    }
    if (g->model == RIP_IC) {
      // then the samples per colour group of the grid sampler, with the
      // lane setting chosen
      for (uint32_t colors : {32u}) {
        rip_tune c = pick;
        c.colors = colors;
        double t = 0.0;
        for (int rep = 0; rep < 2; ++rep) {
          pool_clear(&pool);
          const auto t0 = std::chrono::steady_clock::now();
          pool_sample(&pool, probe, &c);
          t = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                  .count();
        }
        ++extra;
        if (t < best_t) {
          best_t = t;
          pick = c;
        }
      }
    }
  } catch (...) {
    pool_destroy(&pool);
    throw;
  }
  pool_destroy(&pool);
  {
    std::lock_guard<std::mutex> hold(g_tune_mu);
    g_tune_cache[key] = pick;
  }
  {
    // pools of the graph read the choice under the same lock
    std::lock_guard<std::mutex> hold(g->sampler_mu);
    g->tuned = pick;
    g->tuned_set = true;
  }
  *best = pick;
  if (best_ms) *best_ms = best_t;
  if (tried) *tried = static_cast<uint32_t>(cand.size()) + extra;
}

}  // namespace rip
