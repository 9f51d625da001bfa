// IMM driver: theta estimation by doubling, final theta, final sampling and
// selection -- the control flow of Sampling + IMM in include/ripples/imm.h
// (:145-260 and :768-860) with the closed forms of :77-128 and
// include/ripples/utility.h:84-87, evaluated in the same order so the
// double -> integer truncations agree.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "internal.h"

namespace rip {

double log_binomial(uint64_t n, uint64_t k) {
  return std::lgamma(static_cast<double>(n + 1)) -
         std::lgamma(static_cast<double>(n - k + 1)) -
         std::lgamma(static_cast<double>(k + 1));
}

int64_t theta_prime(int64_t x, double eps_prime, double l, uint64_t k,
                    uint64_t n) {
  k = std::min(k, n / 2);
  return static_cast<int64_t>(
      (2 + 2. / 3. * eps_prime) *
      (l * std::log(static_cast<double>(n)) + log_binomial(n, k) +
       std::log(std::log2(static_cast<double>(n)))) *
      std::pow(2.0, static_cast<double>(x)) / (eps_prime * eps_prime));
}

uint64_t theta_final(double eps, double l, uint64_t k, double lb, uint64_t n) {
  if (lb == 0) return 0;
  k = std::min(k, n / 2);
  const double term1 = 0.6321205588285577;  // 1 - 1/e
  const double alpha = std::sqrt(l * std::log(static_cast<double>(n)) + std::log(2));
  const double beta = std::sqrt(
      term1 * (log_binomial(n, k) + l * std::log(static_cast<double>(n)) + std::log(2)));
  const double lamda = static_cast<double>(2 * n) * (term1 * alpha + beta) *
                       (term1 * alpha + beta) * std::pow(eps, -2);
  return static_cast<uint64_t>(lamda / lb);
}

namespace {

using clock_t_ = std::chrono::steady_clock;

double ms_since(clock_t_::time_point t0) {
  return std::chrono::duration<double, std::milli>(clock_t_::now() - t0).count();
}

struct PoolGuard {
  rip_pool pool;
  ~PoolGuard() { pool_destroy(&pool); }
};

// RIPPLES_AMD_LOG=1: one line per estimation round and for the final round
// on stderr (the reference logs its rounds through spdlog).
bool imm_log() {
  static const bool on = [] {
    const char *v = std::getenv("RIPPLES_AMD_LOG");
    return v && *v && *v != '0';
  }();
  return on;
}

}  // namespace

void imm_run(rip_graph *g, uint32_t k, double eps, uint64_t seed,
             const rip_tune *tune, uint32_t *seeds, rip_imm_record *rec) {
  RIP_REQUIRE(eps > 0, -1, "imm: epsilon must be > 0");
  RIP_REQUIRE(g->n >= 2, -1, "imm: the graph needs at least two vertices");
  rip_imm_record r{};
  const auto t_total = clock_t_::now();
  PoolGuard guard;
  rip_pool &pool = guard.pool;
  pool_create(&pool, g, seed);
  const uint64_t n = g->n;
  std::vector<uint32_t> tmp(std::max<uint32_t>(k, 1));

  const double l = 1.0 * (1 + 1 / std::log2(static_cast<double>(n)));
  const double eps_prime = 1.4142135623730951 * eps;
  double lb = 0;
  for (int64_t x = 1; x < std::log2(static_cast<double>(n)); ++x) {
    int64_t tp = theta_prime(x, eps_prime, l, k, n);
    const auto tg = clock_t_::now();
    if (tp <= static_cast<int64_t>(pool.num_sets))
      tp = static_cast<int64_t>(pool.num_sets);
    else
      pool_sample(&pool, static_cast<uint64_t>(tp) - pool.num_sets, tune);
    const double gen_ms = ms_since(tg);
    const auto ts = clock_t_::now();
    uint64_t cov = 0;
    select_run(&pool, pool.num_sets, k, tmp.data(), &cov);
    const double sel_ms = ms_since(ts);
    const double f = static_cast<double>(cov) / static_cast<double>(pool.num_sets);
    if (r.rounds < 64) {
      r.theta_prime[r.rounds] = tp;
      r.f[r.rounds] = f;
      r.round_generate_ms[r.rounds] = gen_ms;
      r.round_select_ms[r.rounds] = sel_ms;
    }
    r.rounds++;
    if (imm_log())
      std::fprintf(stderr, "[imm] round %lld: theta' %lld, sampled in %.1f ms, selected in %.1f ms, "
                   "f %.6f\n", static_cast<long long>(x), static_cast<long long>(tp), gen_ms, sel_ms, f);
    if (f >= std::pow(2, static_cast<double>(-x))) {
      lb = (static_cast<double>(n) * f) / (1 + eps_prime);
      break;
    }
  }
  const uint64_t theta = theta_final(eps, l, k, lb, n);
  r.estimation_ms = ms_since(t_total);
  RIP_REQUIRE(theta < (1ull << 32), -1, "imm: theta exceeds 2^32-1 samples");
  const auto tg = clock_t_::now();
  if (theta > pool.num_sets) pool_sample(&pool, theta - pool.num_sets, tune);
  r.generate_ms = ms_since(tg);
  const auto ts = clock_t_::now();
  uint64_t cov = 0;
  const uint32_t got = select_run(&pool, theta, k, seeds, &cov);
  r.select_ms = ms_since(ts);
  r.theta = theta;
  r.num_sets = pool.num_sets;
  r.rrr_elements = pool.num_elements;
  r.num_seeds = got;
  r.pool_bytes = pool_device_bytes(&pool);
  r.lb = lb;
  r.final_f = theta ? static_cast<double>(cov) / static_cast<double>(theta) : 0.0;
  r.total_ms = ms_since(t_total);
  if (imm_log())
    std::fprintf(stderr, "[imm] final: theta %llu, sampled in %.1f ms, selected in %.1f ms, "
                 "total %.1f ms\n", static_cast<unsigned long long>(theta), r.generate_ms,
                 r.select_ms, r.total_ms);
  if (rec) *rec = r;
}

}  // namespace rip
