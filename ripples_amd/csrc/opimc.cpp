// OPIM-C on the device pool and selector: a second consumer of the sampling
// and selection path, following OPIMC (include/ripples/opimc.h:150-241).
//
// Two RRR collections are kept: R1 drives the greedy selection, R2 checks the
// chosen seeds.  With the per-sample stream contract, R1 is sample ids
// [0, theta) of the seed's sampling stream and R2 is ids [2^31, 2^31 + theta):
// a pool created from the master state 2^32 steps after `seed` draws exactly
// those (sample s element j is master output 1 + 2(s + jS)), so no sampler
// needs an id offset.  Each round selects k seeds on R1, counts the R2 sets
// the seeds cover (FindCoverage, opimc.h:100-118), and stops when the bound
// ratio reaches 1 - 1/e - epsilon; otherwise both collections double.
//
// The bound arithmetic is the reference's, in single precision where its
// variables are float (ThetaHelper, ThetaMax, ThetaZero, LowerBound,
// UpperBound: opimc.h:26-70), with the transcendental calls taken in double
// and rounded on assignment.  The covered count handed to UpperBound is the
// selector's coverage fraction times |R1|, truncated, as in opimc.h:206.
#include <chrono>
#include <cmath>
#include <vector>

#include "internal.h"

namespace rip {

namespace {

constexpr float kOneMinusInvE = 0.6321203113733684f;  // 1 - 1/e

using clk = std::chrono::steady_clock;

double ms_since(clk::time_point t0) {
  return std::chrono::duration<double, std::milli>(clk::now() - t0).count();
}

struct Pools {
  rip_pool r1, r2;
  ~Pools() {
    pool_destroy(&r1);
    pool_destroy(&r2);
  }
};

}  // namespace

float opimc_theta_helper(uint64_t n, float epsilon, uint64_t k, float delta) {
  (void)epsilon;
  const float B = static_cast<float>(kOneMinusInvE * std::sqrt(std::log(6.0 / delta)));
  const float C = static_cast<float>(
      std::sqrt(kOneMinusInvE * (log_binomial(n, k) + std::log(6.0 / delta))));
  const float D = (B + C) * (B + C);
  return 2 * D;
}

uint64_t opimc_theta_max(uint64_t n, float epsilon, uint64_t k, float delta) {
  const float num = static_cast<float>(n) * opimc_theta_helper(n, epsilon, k, delta);
  const float den = epsilon * epsilon * static_cast<float>(k);
  return static_cast<uint64_t>(std::ceil(static_cast<double>(num / den)));
}

uint64_t opimc_theta_zero(uint64_t n, float epsilon, uint64_t k, float delta) {
  return static_cast<uint64_t>(
      std::ceil(static_cast<double>(opimc_theta_helper(n, epsilon, k, delta))));
}

float opimc_lower_bound(uint64_t coverage, float delta2, uint64_t n, uint64_t theta) {
  const float A = static_cast<float>(std::log(static_cast<double>(1 / delta2)));
  const float B = static_cast<float>(std::sqrt(static_cast<double>(
      static_cast<float>(coverage) + 2 * A / 9)));
  const float C = static_cast<float>(std::sqrt(static_cast<double>(A / 2)));
  const float D = (B - C) * (B - C);
  const float E = A / 18;
  return (static_cast<float>(n) / static_cast<float>(theta)) * (D - E);
}

float opimc_upper_bound(uint64_t coverage, float delta1, uint64_t n, uint64_t theta) {
  const float B = static_cast<float>(coverage) / kOneMinusInvE;
  const float C = static_cast<float>(std::log(static_cast<double>(1 / delta1)) / 2);
  const float D = static_cast<float>(std::sqrt(static_cast<double>(B + C)));
  const double sc = std::sqrt(static_cast<double>(C));
  const float E = static_cast<float>((D + sc) * (D + sc));
  return (static_cast<float>(n) / static_cast<float>(theta)) * E;
}

uint64_t opimc_r2_seed(uint64_t seed) {
  return aff_apply(aff_pow(master_step(), 1ull << 32), seed);
}

void opimc_run(rip_graph *g, uint32_t k, double epsilon, double delta_cfg, uint64_t seed,
               const rip_tune *tune, uint32_t *seeds, rip_opimc_record *rec) {
  RIP_REQUIRE(epsilon > 0 && epsilon < kOneMinusInvE, -1,
              "opimc: epsilon must be in (0, 1 - 1/e)");
  RIP_REQUIRE(delta_cfg > 0 && delta_cfg < 1, -1, "opimc: delta must be in (0, 1)");
  RIP_REQUIRE(k >= 1 && k <= g->n, -1, "opimc: k must be in [1, n]");
  const auto t_total = clk::now();
  rip_opimc_record r{};
  const uint64_t n = g->n;
  const float eps_f = static_cast<float>(epsilon), delta_f = static_cast<float>(delta_cfg);
  uint64_t theta = opimc_theta_zero(n, eps_f, k, delta_f);
  const uint64_t theta_max = opimc_theta_max(n, eps_f, k, delta_f);
  r.theta_zero = theta;
  r.theta_max = theta_max;
  RIP_REQUIRE(theta_max / theta >= 2, -1, "opimc: theta_max / theta_zero must be at least 2");
  const uint64_t i_max =
      static_cast<uint64_t>(std::ceil(std::log2(static_cast<double>(theta_max / theta))));
  const float delta = delta_f / static_cast<float>(3 * i_max);

  Pools P;
  pool_create(&P.r1, g, seed);
  pool_create(&P.r2, g, opimc_r2_seed(seed));
  std::vector<uint32_t> chosen(k);
  for (uint32_t round = 0;; ++round) {
    RIP_REQUIRE(theta < (1ull << 31), -1, "opimc: theta reached 2^31 sets per collection");
    const auto tg = clk::now();
    const uint64_t grow = theta - P.r1.num_sets;
    pool_sample(&P.r1, grow, tune);
    pool_sample(&P.r2, grow, tune);
    const double gen_ms = ms_since(tg);
    const auto ts = clk::now();
    uint64_t cov1 = 0;
    const uint32_t got = select_run(&P.r1, theta, k, chosen.data(), &cov1);
    const double sel_ms = ms_since(ts);
    const uint64_t cov2 = coverage_run(&P.r2, theta, chosen.data(), got);
    // the selector reports a fraction; OPIMC multiplies it back by |R1|
    const double f1 = static_cast<double>(cov1) / static_cast<double>(theta);
    const uint64_t c1 = static_cast<uint64_t>(f1 * static_cast<double>(theta));
    const float upper = opimc_upper_bound(c1, delta, n, theta);
    const float lower = opimc_lower_bound(cov2, delta, n, theta);
    const float alpha = lower / upper;
    if (round < 64) {
      r.rrr_sets_generated[round] = 2 * grow;
      r.generate_ms[round] = gen_ms;
      r.select_ms[round] = sel_ms;
      r.coverage1[round] = c1;
      r.coverage2[round] = cov2;
      r.lower[round] = lower;
      r.upper[round] = upper;
      r.alpha[round] = alpha;
    }
    r.rounds = round + 1;
    r.num_seeds = got;
    r.theta = theta;
    if (static_cast<double>(alpha) >=
        static_cast<double>(kOneMinusInvE) - epsilon) {
      std::copy(chosen.begin(), chosen.begin() + got, seeds);
      break;
    }
    theta *= 2;
  }
  r.rrr_elements = P.r1.num_elements + P.r2.num_elements;
  r.total_ms = ms_since(t_total);
  if (rec) *rec = r;
}

}  // namespace rip
