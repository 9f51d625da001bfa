// opimc: command-line OPIM-C on one GPU through the C ABI, with the flags of
// the reference tool (tools/opimc.cc: the imm flags plus --delta from
// include/ripples/opimc_configuration.h:47-56) and its JSON experiment record
// (tools/opimc.cc:63-97): an array holding one object, times in
// milliseconds, seeds as original vertex ids.
//
//   opimc -i graph.tsv -k 50 -e 0.5 -d IC [--delta 0.05] [-w] [-u]
//         [--reload-binary] [--distribution ...] [-o output.json] [--seed S]
//
// --seed (absent from the reference, which seeds both streams with 0) seeds
// the weights and sampling streams.
#include "cli_common.h"

namespace cli {
const char *kTool = "opimc";
const char *kUsage =
    "usage: opimc -i GRAPH -k K -e EPS -d IC|LT [--delta D] [-w] [-u] [--reload-binary]\n"
    "             [--distribution uniform|normal|const] [--mean M] [--variance V]\n"
    "             [--scale-factor S] [--disable-renumbering]\n"
    "             [-o OUT.json] [--seed S] [--device D]\n";
}  // namespace cli

using namespace cli;

int main(int argc, char **argv) {
  const Options o = parse(argc, argv, true);
  const int model = o.model == "IC" ? RIP_IC : RIP_LT;
  Loaded L;
  load(o, model, &L);

  rip_opimc_record rec{};
  std::vector<uint32_t> seeds(o.k ? o.k : 1);
  check(rip_opimc(L.g, static_cast<uint32_t>(o.k), o.epsilon, o.delta, o.seed, nullptr,
                  seeds.data(), &rec),
        "OPIM-C");
  seeds.resize(rec.num_seeds);
  for (uint32_t &s : seeds) s = L.reverse[s];  // original ids, as G.convertID does
  const uint32_t rounds = rec.rounds < 64 ? rec.rounds : 64;
  std::vector<double> gen(rec.generate_ms, rec.generate_ms + rounds);
  std::vector<double> sel(rec.select_ms, rec.select_ms + rounds);
  std::vector<uint64_t> generated(rec.rrr_sets_generated, rec.rrr_sets_generated + rounds);

  Record R;
  R.field("Algorithm", json_string("OPIMC"));
  R.field("Input", json_string(o.input));
  R.field("Output", json_string(o.output));
  R.field("DiffusionModel", json_string(o.model));
  R.field("Epsilon", num(o.epsilon));
  R.field("K", std::to_string(o.k));
  R.field("Delta", num(o.delta));
  R.field("NumThreads", "1");
  R.field("NumWalkWorkers", std::to_string(o.gpu_workers));
  R.field("NumCPUTeams", std::to_string(o.cpu_teams));
  R.field("NumGPUWalkWorkers", std::to_string(o.gpu_workers));
  R.field("PauseThreshold", num(o.pause));
  R.field("Total", num(rec.total_ms));
  R.field("CPUBatchSize", std::to_string(o.cpu_batch));
  R.field("GPUBatchSize", std::to_string(o.gpu_batch));
  R.field("RRRSetSizeBytes", std::to_string(4 * rec.rrr_elements));
  R.field("GenerateRRRSets", num_list(gen));
  R.field("FindMostInfluentialSet", num_list(sel));
  R.field("Microbenchmarking", "0");
  R.field("ThetaZero", std::to_string(rec.theta_zero));
  R.field("ThetaMax", std::to_string(rec.theta_max));
  R.field("RRRSetsGenerated", json_list(generated));
  R.field("Seeds", json_list(seeds));
  if (!R.write(o.output)) {
    std::fprintf(stderr, "opimc: cannot write %s\n", o.output.c_str());
    return 1;
  }
  std::fprintf(stderr, "opimc: %u rounds, theta %llu, %zu seeds, %.1f ms; record in %s\n",
               rec.rounds, static_cast<unsigned long long>(rec.theta), seeds.size(),
               rec.total_ms, o.output.c_str());
  return 0;
}
