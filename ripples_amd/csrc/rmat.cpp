// Synthetic inputs: R-MAT edge lists and transposed CSR construction with the
// weight schemes of the IMM workloads.  Host code; deterministic for a given
// seed regardless of the number of threads.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <thread>
#include <vector>

#include "common.h"
#include "internal.h"
#include "ripples_amd.h"

namespace rip {
namespace {

unsigned num_threads() {
  unsigned t = std::thread::hardware_concurrency();
  if (t == 0) t = 1;
  return std::min(t, 32u);
}

template <typename F>
void parallel_for(uint64_t n, F &&body) {
  unsigned nt = num_threads();
  if (n < 4096) nt = 1;
  std::vector<std::thread> pool;
  for (unsigned t = 0; t < nt; ++t) {
    uint64_t lo = n * t / nt, hi = n * (t + 1) / nt;
    pool.emplace_back([&, lo, hi, t] { body(lo, hi, t); });
  }
  for (auto &th : pool) th.join();
}

inline uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Edge i of an R-MAT graph: one quadrant draw per level, MSB first.
inline void rmat_edge(uint32_t scale, double a, double ab, double abc,
                      uint64_t seed, uint64_t i, uint32_t *s, uint32_t *d) {
  uint32_t u = 0, v = 0;
  uint64_t base = splitmix64(seed ^ 0x5DEECE66Dull) + i * 0x100000001B3ull;
  for (uint32_t l = 0; l < scale; ++l) {
    uint64_t h = splitmix64(base + l * 0xD1B54A32D192ED03ull);
    double r = static_cast<double>(h >> 11) * 0x1p-53;
    uint32_t bu, bv;
    if (r < a) { bu = 0; bv = 0; }
    else if (r < ab) { bu = 0; bv = 1; }
    else if (r < abc) { bu = 1; bv = 0; }
    else { bu = 1; bv = 1; }
    u = (u << 1) | bu;
    v = (v << 1) | bv;
  }
  *s = u;
  *d = v;
}

// Bucket (key, payload) pairs by key with atomic counters, then sort each
// bucket so the result does not depend on thread interleaving.
struct Buckets {
  std::vector<uint64_t> start;  // n + 1
  std::vector<uint32_t> item;   // payload per slot
};

}  // namespace

// Forward edge list: sorted by (src, dst), deduplicated, no self loops.
void rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b,
                double c, uint64_t seed, std::vector<uint32_t> &src,
                std::vector<uint32_t> &dst) {
  RIP_REQUIRE(scale >= 1 && scale <= 31, -1, "rmat: scale must be in [1, 31]");
  RIP_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, -1,
              "rmat: probabilities must be >= 0 and a+b+c <= 1");
  const uint64_t n = 1ull << scale;
  const uint64_t m0 = static_cast<uint64_t>(edge_factor) << scale;
  const double ab = a + b, abc = a + b + c;
  std::vector<uint32_t> es(m0), ed(m0);
  std::vector<std::atomic<uint32_t>> cnt(n);
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) cnt[i].store(0, std::memory_order_relaxed);
  });
  parallel_for(m0, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      rmat_edge(scale, a, ab, abc, seed, i, &es[i], &ed[i]);
      if (es[i] != ed[i]) cnt[es[i]].fetch_add(1, std::memory_order_relaxed);
    }
  });
  std::vector<uint64_t> start(n + 1, 0);
  for (uint64_t v = 0; v < n; ++v) start[v + 1] = start[v] + cnt[v].load();
  std::vector<uint32_t> bucket(start[n]);
  std::vector<std::atomic<uint64_t>> cur(n);
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t v = lo; v < hi; ++v) cur[v].store(start[v], std::memory_order_relaxed);
  });
  parallel_for(m0, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) {
      if (es[i] == ed[i]) continue;
      uint64_t p = cur[es[i]].fetch_add(1, std::memory_order_relaxed);
      bucket[p] = ed[i];
    }
  });
  es.clear();
  es.shrink_to_fit();
  ed.clear();
  ed.shrink_to_fit();
  // sort + dedup each source row, count survivors
  std::vector<uint64_t> keep(n + 1, 0);
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t v = lo; v < hi; ++v) {
      auto b0 = bucket.begin() + start[v], b1 = bucket.begin() + start[v + 1];
      std::sort(b0, b1);
      keep[v + 1] = std::unique(b0, b1) - b0;
    }
  });
  for (uint64_t v = 0; v < n; ++v) keep[v + 1] += keep[v];
  src.assign(keep[n], 0);
  dst.assign(keep[n], 0);
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t v = lo; v < hi; ++v) {
      uint64_t k = keep[v + 1] - keep[v];
      for (uint64_t j = 0; j < k; ++j) {
        src[keep[v] + j] = static_cast<uint32_t>(v);
        dst[keep[v] + j] = bucket[start[v] + j];
      }
    }
  });
}

// Transposed CSR (row v = in-neighbours of v, ascending) with weights.
void csr_from_edges(uint32_t n, uint64_t m, const uint32_t *src,
                    const uint32_t *dst, int scheme, float scale,
                    uint64_t seed, uint64_t *index, uint32_t *edges,
                    float *weights) {
  RIP_REQUIRE(scheme >= RIP_W_UNIFORM && scheme <= RIP_W_CONST, -1,
              "csr: unknown weight scheme");
  for (uint64_t i = 0; i < m; ++i)
    RIP_REQUIRE(src[i] < n && dst[i] < n, -1, "csr: vertex id out of range");
  // per-input-edge draw from the weights stream master.split(2,0):
  // edge i takes master output 2i = M^(2i+1)(seed)
  std::vector<float> draw;
  if (scheme == RIP_W_UNIFORM || scheme == RIP_W_LT_NORM) {
    draw.resize(m);
    const Affine step2 = aff_pow(master_step(), 2);
    parallel_for(m, [&](uint64_t lo, uint64_t hi, unsigned) {
      uint64_t x = aff_apply(aff_pow(master_step(), 2 * lo + 1), seed);
      for (uint64_t i = lo; i < hi; ++i) {
        draw[i] = u01(x);
        x = aff_apply(step2, x);
      }
    });
  }
  std::vector<std::atomic<uint64_t>> cnt(n);
  for (uint32_t v = 0; v < n; ++v) cnt[v].store(0, std::memory_order_relaxed);
  parallel_for(m, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t i = lo; i < hi; ++i) cnt[dst[i]].fetch_add(1, std::memory_order_relaxed);
  });
  index[0] = 0;
  for (uint32_t v = 0; v < n; ++v) index[v + 1] = index[v] + cnt[v].load();
  std::vector<uint64_t> order(m);
  {
    std::vector<std::atomic<uint64_t>> cur(n);
    for (uint32_t v = 0; v < n; ++v) cur[v].store(index[v], std::memory_order_relaxed);
    parallel_for(m, [&](uint64_t lo, uint64_t hi, unsigned) {
      for (uint64_t i = lo; i < hi; ++i)
        order[cur[dst[i]].fetch_add(1, std::memory_order_relaxed)] = i;
    });
  }
  parallel_for(n, [&](uint64_t lo, uint64_t hi, unsigned) {
    for (uint64_t v = lo; v < hi; ++v) {
      auto o0 = order.begin() + index[v], o1 = order.begin() + index[v + 1];
      std::sort(o0, o1, [&](uint64_t x, uint64_t y) {
        return src[x] != src[y] ? src[x] < src[y] : x < y;
      });
      uint64_t deg = index[v + 1] - index[v];
      float total = 0.0f;
      if (scheme == RIP_W_LT_NORM)
        for (uint64_t j = index[v]; j < index[v + 1]; ++j) total += draw[order[j]];
      for (uint64_t j = index[v]; j < index[v + 1]; ++j) {
        uint64_t i = order[j];
        edges[j] = src[i];
        float w;
        switch (scheme) {
          case RIP_W_UNIFORM: w = scale * draw[i]; break;
          case RIP_W_CASCADE: w = 1.0f / static_cast<float>(deg); break;
          case RIP_W_LT_NORM: w = (draw[i] / total) * scale; break;
          default: w = scale; break;
        }
        weights[j] = w;
      }
    }
  });
}

}  // namespace rip
