// Serving loop: IMM queries against resident graphs, with continuous
// batching.  There is no counterpart in the reference (which runs one IMM per
// process, tools/imm.cc); each query follows the IMM control flow of imm_run
// (include/ripples/imm.h:145-260, 768-860) and gives the same seeds as
// rip_imm(g, k, epsilon, seed) alone.
//
// RRR pools stay resident in HBM, keyed by (graph, seed): the pool of a key
// is the sample stream rip_imm draws for that seed.  A query only ever uses
// a prefix of its pool -- N = theta'(x) in round x of the estimation and
// theta at the end, exactly the sets a solo run draws -- so every prefix
// already drawn for an earlier or concurrent query with the same key serves it
// unchanged (theta-sufficient reuse), and queries with different k and
// epsilon on one key share each sampling launch.
//
// A query is admitted from the arrival queue while fewer than `slots` queries
// are active and, if it needs a new pool, while the resident pools fit the HBM
// budget.  A new key takes over the least recently used idle pool of its graph
// (buffers kept, contents dropped) once the graph holds `slots` pools or the
// budget is reached; idle pools of other graphs are freed, least recently used
// first, to make room.
// `workers` threads repeat one scheduling step each on an idle pool that has
// active queries: extend the pool once to the largest prefix any of them waits
// for, then run the selection of every query whose prefix is drawn and move it
// to its next round.  Each pool has its own stream, so steps of different
// pools overlap on the device; sampling launches of pools of one graph take
// turns on the graph's sampler scratch (pool_sample holds its lock).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <utility>
#include <vector>

#include "internal.h"

namespace rip {
namespace {

using clk = std::chrono::steady_clock;

double ms_between(clk::time_point a, clk::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Entry;

struct Query {
  uint64_t id = 0;
  uint32_t graph = 0, k = 0;
  double eps = 0;
  uint64_t seed = 0;
  Entry *entry = nullptr;
  // IMM state
  int64_t x = 1;
  double l = 0, eps_prime = 0, lb = 0;
  bool final_round = false, done = false;
  uint64_t need = 0;   // prefix of the pool the next selection uses
  uint64_t drawn = 0;  // largest prefix selected on so far (a solo run's pool size)
  std::vector<uint32_t> seeds;
  rip_imm_record rec{};
  int error = RIP_OK;
  std::string message;
  clk::time_point submitted, admitted, finished;
};

// One resident pool and the active queries that select on it.
struct Entry {
  uint32_t graph = 0;
  uint64_t seed = 0;
  rip_pool pool;
  bool created = false;  // pool_create ran (by the first worker that steps it)
  bool reseed = false;   // re-keyed to `seed`: the worker empties the pool first
  bool busy = false;     // a worker is stepping it
  std::vector<Query *> active;
  uint64_t bytes = 0;    // device bytes, refreshed after every step
  clk::time_point last_used;
};

}  // namespace
}  // namespace rip

struct rip_server {
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  uint32_t slots = 8;
  uint64_t budget = 0;
  uint64_t next_id = 1;
  uint32_t active = 0;  // admitted, unfinished queries
  std::vector<rip_graph *> graphs;
  std::map<std::pair<uint32_t, uint64_t>, std::unique_ptr<rip::Entry>> pools;
  std::deque<rip::Query *> inbox;
  std::map<uint64_t, std::unique_ptr<rip::Query>> queries;
  std::vector<std::thread> workers;
  uint64_t launches = 0, sets_drawn = 0, selections = 0, reused = 0, evicted = 0, recycled = 0;
};

namespace rip {
namespace {

uint64_t pool_bytes(const rip_pool &p) { return pool_device_bytes(&p); }

// The callers below hold the server lock.
uint64_t resident_bytes(const rip_server *s) {
  uint64_t b = 0;
  for (const auto &e : s->pools) b += e.second->bytes;
  return b;
}

// Idle pools (no active query, not being stepped), least recently used first,
// move to `out` until the resident bytes fall under the budget.
void evict_idle(rip_server *s, std::vector<std::unique_ptr<Entry>> *out) {
  while (resident_bytes(s) >= s->budget) {
    auto victim = s->pools.end();
    for (auto it = s->pools.begin(); it != s->pools.end(); ++it) {
      const Entry &e = *it->second;
      if (e.busy || !e.active.empty()) continue;
      if (victim == s->pools.end() || e.last_used < victim->second->last_used) victim = it;
    }
    if (victim == s->pools.end()) return;
    out->push_back(std::move(victim->second));
    s->pools.erase(victim);
    s->evicted++;
  }
}

// A pool for a new key: an idle pool of the same graph is re-keyed (its
// buffers are kept, so no device allocation or free) once the graph holds
// `slots` pools or the budget is reached; otherwise a new one is made, after
// evicting idle pools while over budget.  End of the map when the query has
// to wait for memory.
std::map<std::pair<uint32_t, uint64_t>, std::unique_ptr<Entry>>::iterator new_pool(
    rip_server *s, Query *q, std::vector<std::unique_ptr<Entry>> *evicted) {
  const auto key = std::make_pair(q->graph, q->seed);
  auto victim = s->pools.end();
  uint32_t of_graph = 0;
  for (auto it = s->pools.begin(); it != s->pools.end(); ++it) {
    const Entry &e = *it->second;
    if (e.graph != q->graph) continue;
    ++of_graph;
    if (e.busy || !e.active.empty()) continue;
    if (victim == s->pools.end() || e.last_used < victim->second->last_used) victim = it;
  }
  const bool over = s->budget != 0 && resident_bytes(s) >= s->budget;
  if (victim != s->pools.end() && (of_graph >= s->slots || over)) {
    auto node = s->pools.extract(victim);
    node.key() = key;
    node.mapped()->seed = q->seed;
    node.mapped()->reseed = true;
    s->recycled++;
    return s->pools.insert(std::move(node)).position;
  }
  if (over) {
    evict_idle(s, evicted);
    if (s->active != 0 && resident_bytes(s) >= s->budget) return s->pools.end();
  }
  auto e = std::make_unique<Entry>();
  e->graph = q->graph;
  e->seed = q->seed;
  return s->pools.emplace(key, std::move(e)).first;
}

void admit(rip_server *s, std::vector<std::unique_ptr<Entry>> *evicted) {
  while (!s->inbox.empty() && s->active < s->slots) {
    Query *q = s->inbox.front();
    auto it = s->pools.find(std::make_pair(q->graph, q->seed));
    if (it == s->pools.end()) {
      it = new_pool(s, q, evicted);
      if (it == s->pools.end()) return;
    } else {
      s->reused++;
    }
    s->inbox.pop_front();
    q->entry = it->second.get();
    q->admitted = clk::now();
    q->entry->active.push_back(q);
    s->active++;
  }
}

// The idle pool whose oldest active query arrived first.
Entry *pick(rip_server *s) {
  Entry *best = nullptr;
  clk::time_point when;
  for (auto &kv : s->pools) {
    Entry *e = kv.second.get();
    if (e->busy || e->active.empty()) continue;
    clk::time_point t = e->active.front()->submitted;
    for (Query *q : e->active) t = std::min(t, q->submitted);
    if (!best || t < when) {
      best = e;
      when = t;
    }
  }
  return best;
}

// The first round's prefix, or, past the last round, the final theta
// (imm_run's loop bounds and closed forms).
void start_query(Query *q, uint64_t n) {
  q->l = 1.0 * (1 + 1 / std::log2(static_cast<double>(n)));
  q->eps_prime = 1.4142135623730951 * q->eps;
  q->x = 1;
  if (static_cast<double>(q->x) < std::log2(static_cast<double>(n))) {
    q->need = static_cast<uint64_t>(theta_prime(q->x, q->eps_prime, q->l, q->k, n));
  } else {
    q->final_round = true;
    q->need = theta_final(q->eps, q->l, q->k, 0.0, n);
    RIP_REQUIRE(q->need < (1ull << 32), -1, "imm: theta exceeds 2^32-1 samples");
  }
}

// One selection of a query on its prefix, and the move to its next prefix.
void step_query(rip_pool *pool, uint64_t n, Query *q) {
  std::vector<uint32_t> tmp(std::max<uint32_t>(q->k, 1));
  uint64_t cov = 0;
  const auto t0 = clk::now();
  q->drawn = std::max(q->drawn, q->need);
  if (!q->final_round) {
    select_run(pool, q->need, q->k, tmp.data(), &cov);
    const double f = q->need ? static_cast<double>(cov) / static_cast<double>(q->need) : 0.0;
    rip_imm_record &rec = q->rec;
    if (rec.rounds < 64) {
      rec.theta_prime[rec.rounds] = static_cast<int64_t>(q->need);
      rec.f[rec.rounds] = f;
      rec.round_select_ms[rec.rounds] = ms_between(t0, clk::now());
    }
    rec.rounds++;
    if (f >= std::pow(2, static_cast<double>(-q->x))) {
      q->lb = (static_cast<double>(n) * f) / (1 + q->eps_prime);
      q->final_round = true;
      q->need = theta_final(q->eps, q->l, q->k, q->lb, n);
    } else if (static_cast<double>(++q->x) < std::log2(static_cast<double>(n))) {
      q->need = static_cast<uint64_t>(theta_prime(q->x, q->eps_prime, q->l, q->k, n));
    } else {
      q->final_round = true;
      q->need = theta_final(q->eps, q->l, q->k, q->lb, n);
    }
    if (q->final_round) {
      RIP_REQUIRE(q->need < (1ull << 32), -1, "imm: theta exceeds 2^32-1 samples");
      q->rec.estimation_ms = ms_between(q->admitted, clk::now());
    }
    return;
  }
  q->seeds.assign(std::max<uint32_t>(q->k, 1), 0);
  const uint32_t got = select_run(pool, q->need, q->k, q->seeds.data(), &cov);
  q->seeds.resize(got);
  rip_imm_record &rec = q->rec;
  rec.select_ms = ms_between(t0, clk::now());
  rec.theta = q->need;
  rec.num_sets = q->drawn;
  rec.rrr_elements = pool->num_elements;  // elements of the shared pool
  rec.pool_bytes = pool_device_bytes(pool);
  rec.num_seeds = got;
  rec.lb = q->lb;
  rec.final_f = q->need ? static_cast<double>(cov) / static_cast<double>(q->need) : 0.0;
  q->done = true;
}

void fail(const std::vector<Query *> &qs, int code, const char *msg) {
  for (Query *q : qs) {
    q->error = code;
    q->message = msg;
    q->done = true;
  }
}

void worker_loop(rip_server *s) {
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    std::vector<std::unique_ptr<Entry>> evicted;
    admit(s, &evicted);
    if (!evicted.empty()) {
      lk.unlock();
      for (auto &v : evicted)
        if (v->created) {
          (void)hipSetDevice(s->graphs[v->graph]->device);
          pool_destroy(&v->pool);
        }
      evicted.clear();
      lk.lock();
      continue;  // the lock was dropped: look again
    }
    Entry *e = pick(s);
    if (!e) {
      if (s->stop && s->active == 0 && s->inbox.empty()) return;
      s->cv_work.wait(lk);
      continue;
    }
    e->busy = true;
    const std::vector<Query *> batch = e->active;
    rip_graph *g = s->graphs[e->graph];
    lk.unlock();

    uint64_t launches = 0, drawn = 0, selections = 0;
    try {
      RIP_HIP(hipSetDevice(g->device));
      if (!e->created) {
        pool_create(&e->pool, g, e->seed);
        e->created = true;
      } else if (e->reseed) {
        e->pool.seed = e->seed;
        e->pool.num_sets = 0;
        e->pool.used_data = 0;
        e->pool.groups = 0;
        e->pool.num_elements = 0;
      }
      e->reseed = false;
      // one shared launch extends the pool to the largest prefix waited for
      uint64_t want = 0;
      for (Query *q : batch) want = std::max(want, q->need);
      if (want > e->pool.num_sets) {
        const uint64_t before = e->pool.num_sets;
        const auto t0 = clk::now();
        pool_sample(&e->pool, want - before, nullptr);
        const double ms = ms_between(t0, clk::now());
        launches++;
        drawn += want - before;
        for (Query *q : batch) q->rec.generate_ms += ms;
      }
      for (Query *q : batch)
        if (q->need <= e->pool.num_sets) {
          step_query(&e->pool, g->n, q);
          selections++;
        }
    } catch (const Error &err) {
      fail(batch, err.code, err.what());
    } catch (const std::exception &err) {
      fail(batch, -1, err.what());
    }

    lk.lock();
    e->busy = false;
    e->bytes = e->created ? pool_bytes(e->pool) : 0;
    e->last_used = clk::now();
    s->launches += launches;
    s->sets_drawn += drawn;
    s->selections += selections;
    const auto now = clk::now();
    auto it = std::remove_if(e->active.begin(), e->active.end(), [&](Query *q) {
      if (!q->done) return false;
      q->finished = now;
      q->rec.total_ms = ms_between(q->admitted, now);
      return true;
    });
    const size_t retired = static_cast<size_t>(e->active.end() - it);
    e->active.erase(it, e->active.end());
    s->active -= static_cast<uint32_t>(retired);
    if (retired) s->cv_done.notify_all();
    s->cv_work.notify_all();  // the pool is idle again; slots may have freed
  }
}

}  // namespace

rip_server *server_create(uint32_t slots, uint32_t workers, uint64_t budget) {
  RIP_REQUIRE(slots >= 1, -1, "server: at least one slot");
  RIP_REQUIRE(workers >= 1 && workers <= 64, -1, "server: 1 to 64 workers");
  auto *s = new rip_server();
  s->slots = slots;
  s->budget = budget;
  for (uint32_t i = 0; i < workers; ++i) s->workers.emplace_back(worker_loop, s);
  return s;
}

uint32_t server_add_graph(rip_server *s, rip_graph *g) {
  RIP_REQUIRE(g != nullptr, -1, "server: graph is NULL");
  RIP_REQUIRE(g->n >= 2, -1, "server: the graph needs at least two vertices");
  std::lock_guard<std::mutex> lk(s->mu);
  s->graphs.push_back(g);
  return static_cast<uint32_t>(s->graphs.size() - 1);
}

uint64_t server_submit(rip_server *s, uint32_t graph, uint32_t k, double eps, uint64_t seed) {
  RIP_REQUIRE(eps > 0, -1, "imm: epsilon must be > 0");
  std::lock_guard<std::mutex> lk(s->mu);
  RIP_REQUIRE(graph < s->graphs.size(), -1, "server: no such graph");
  RIP_REQUIRE(!s->stop, -4, "server: shutting down");
  auto q = std::make_unique<Query>();
  q->id = s->next_id++;
  q->graph = graph;
  q->k = k;
  q->eps = eps;
  q->seed = seed;
  q->submitted = clk::now();
  start_query(q.get(), s->graphs[graph]->n);
  Query *qp = q.get();
  s->queries[q->id] = std::move(q);
  s->inbox.push_back(qp);
  s->cv_work.notify_all();
  return qp->id;
}

bool server_poll(rip_server *s, uint64_t id, double wait_ms, uint32_t *seeds, uint32_t cap,
                 uint32_t *num_seeds, rip_imm_record *rec, double *latency_ms) {
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->queries.find(id);
  RIP_REQUIRE(it != s->queries.end(), -1, "server: no such query (or already collected)");
  Query *q = it->second.get();
  // `finished` is written under the lock when a worker retires the query
  auto finished = [&] { return q->finished != clk::time_point(); };
  if (!finished() && wait_ms > 0)
    s->cv_done.wait_for(lk, std::chrono::duration<double, std::milli>(wait_ms), finished);
  if (!finished()) return false;
  if (q->error != RIP_OK) {
    const int code = q->error;
    const std::string msg = q->message;
    s->queries.erase(it);
    throw Error(code, msg);
  }
  RIP_REQUIRE(!seeds || cap >= q->seeds.size(), -1, "server: seed buffer too small");
  if (seeds) std::copy(q->seeds.begin(), q->seeds.end(), seeds);
  if (num_seeds) *num_seeds = static_cast<uint32_t>(q->seeds.size());
  if (rec) *rec = q->rec;
  if (latency_ms) *latency_ms = ms_between(q->submitted, q->finished);
  s->queries.erase(it);
  return true;
}

void server_stats(rip_server *s, uint64_t *out, uint32_t count) {
  std::lock_guard<std::mutex> lk(s->mu);
  const uint64_t v[RIP_SERVER_STAT_COUNT] = {
      s->launches, s->sets_drawn, s->selections, resident_bytes(s),
      static_cast<uint64_t>(s->pools.size()), s->reused, s->evicted, s->recycled};
  for (uint32_t i = 0; i < count && i < RIP_SERVER_STAT_COUNT; ++i) out[i] = v[i];
}

void server_destroy(rip_server *s) {
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  for (auto &t : s->workers) t.join();
  for (auto &kv : s->pools)
    if (kv.second->created) {
      (void)hipSetDevice(s->graphs[kv.second->graph]->device);
      pool_destroy(&kv.second->pool);
    }
  delete s;
}

}  // namespace rip
