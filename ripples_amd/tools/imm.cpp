// imm: command-line IMM on one GPU through the C ABI, with the flags of the
// reference tool (tools/imm.cc; flags from include/ripples/configuration.h,
// imm_configuration.h, tim_configuration.h) and its JSON experiment record
// (tools/imm.cc:101-136): an array holding one object, times in
// milliseconds, seeds as original vertex ids.
//
//   imm -i graph.tsv -k 50 -e 0.5 -d IC [-w] [-u] [--reload-binary]
//       [--distribution uniform|normal|const] [--mean M] [--variance V]
//       [--scale-factor S] [--disable-renumbering] [--rrsets N]
//       [-o output.json] [--seed S] [--device D]
//
// Sampling and selection always run on the device; the reference's engine
// knobs (-p, --streaming-*, --gpu-batch-size, --cpu-batch-size,
// --pause-threshold, --seed-select-max-*) are accepted and echoed in the
// record.  --rrsets N samples N RRR sets and reports only their generation
// time, with k copies of vertex 1 as the seeds, as tools/imm.cc:247-281 does.
// --seed (absent from the reference, which seeds both streams with 0) seeds
// the weights and sampling streams.
#include "cli_common.h"

namespace cli {
const char *kTool = "imm";
const char *kUsage =
    "usage: imm -i GRAPH -k K -e EPS -d IC|LT [-w] [-u] [--reload-binary]\n"
    "           [--distribution uniform|normal|const] [--mean M] [--variance V]\n"
    "           [--scale-factor S] [--disable-renumbering] [--rrsets N]\n"
    "           [-o OUT.json] [--seed S] [--device D]\n";
}  // namespace cli

using namespace cli;

int main(int argc, char **argv) {
  const Options o = parse(argc, argv, false);
  const int model = o.model == "IC" ? RIP_IC : RIP_LT;
  Loaded L;
  load(o, model, &L);
  rip_graph *g = L.g;

  std::vector<uint32_t> seeds;
  std::vector<uint64_t> deltas;
  std::vector<double> est_gen, est_sel;
  double total = 0, estimation = 0, generate = 0, select = 0;
  uint64_t theta = 0, elements = 0;
  if (o.rrsets) {
    rip_pool *p = nullptr;
    check(rip_pool_create(g, o.seed, &p), "creating the pool");
    const auto t = std::chrono::steady_clock::now();
    check(rip_pool_sample(p, o.rrsets, nullptr), "sampling");
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
    uint64_t sets = 0;
    check(rip_pool_size(p, &sets, &elements), "pool size");
    rip_pool_destroy(p);
    theta = o.rrsets;
    deltas.push_back(o.rrsets);
    est_gen.push_back(ms);
    est_sel.push_back(0.0);
    total = ms;
    seeds.assign(o.k, L.n > 1 ? 1u : 0u);
  } else {
    rip_imm_record rec{};
    seeds.resize(o.k ? o.k : 1);
    check(rip_imm(g, static_cast<uint32_t>(o.k), o.epsilon, o.seed, nullptr, seeds.data(), &rec),
          "IMM");
    seeds.resize(rec.num_seeds);
    uint64_t have = 0;
    for (uint32_t r = 0; r < rec.rounds && r < 64; ++r) {
      const uint64_t tp = static_cast<uint64_t>(rec.theta_prime[r]);
      deltas.push_back(tp > have ? tp - have : 0);
      have = tp > have ? tp : have;
      est_gen.push_back(rec.round_generate_ms[r]);
      est_sel.push_back(rec.round_select_ms[r]);
    }
    if (rec.theta > have) deltas.push_back(rec.theta - have);
    theta = rec.theta;
    elements = rec.rrr_elements;
    total = rec.total_ms;
    estimation = rec.estimation_ms;
    generate = rec.generate_ms;
    select = rec.select_ms;
  }
  for (uint32_t &s : seeds) s = L.reverse[s];  // original ids, as G.convertID does

  Record R;
  R.field("Algorithm", json_string("IMM"));
  R.field("Input", json_string(o.input));
  R.field("Output", json_string(o.output));
  R.field("DiffusionModel", json_string(o.model));
  R.field("Epsilon", num(o.epsilon));
  R.field("K", std::to_string(o.k));
  R.field("L", "1");
  R.field("NumThreads", "1");
  R.field("NumWalkWorkers", std::to_string(o.gpu_workers));
  R.field("NumCPUTeams", std::to_string(o.cpu_teams));
  R.field("NumGPUWalkWorkers", std::to_string(o.gpu_workers));
  R.field("PauseThreshold", num(o.pause));
  R.field("Total", num(total));
  R.field("ThetaPrimeDeltas", json_list(deltas));
  R.field("ThetaEstimation", num(estimation));
  R.field("ThetaEstimationGenerateRRR", num_list(est_gen));
  R.field("ThetaEstimationMostInfluential", num_list(est_sel));
  R.field("Theta", std::to_string(theta));
  R.field("Counting", "[]");
  R.field("Pivoting", "[]");
  R.field("Microbenchmarking", "0");
  R.field("CPUBatchSize", std::to_string(o.cpu_batch));
  R.field("GPUBatchSize", std::to_string(o.gpu_batch));
  R.field("RRRSetSizeBytes", std::to_string(4 * elements));
  R.field("GenerateRRRSets", num(generate));
  R.field("FindMostInfluentialSet", num(select));
  R.field("Seeds", json_list(seeds));
  if (!R.write(o.output)) {
    std::fprintf(stderr, "imm: cannot write %s\n", o.output.c_str());
    return 1;
  }
  std::fprintf(stderr, "imm: theta %llu, %zu seeds, %.1f ms; record in %s\n",
               static_cast<unsigned long long>(theta), seeds.size(), total, o.output.c_str());
  return 0;
}
