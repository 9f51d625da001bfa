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
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ripples_amd.h"

namespace {

struct Options {
  std::string input, output = "output.json", model = "IC", distribution = "uniform";
  bool weighted = false, undirected = false, reload = false, renumber = true, parallel = false;
  float mean = 0.5f, variance = 1.0f, scale = 1.0f, pause = 1.0f;
  uint64_t k = 10, rrsets = 0, seed = 0;
  double epsilon = 0.5;
  int device = 0;
  long cpu_teams = 0, gpu_workers = 1, gpu_batch = 64, cpu_batch = 64;
  bool have_k = false, have_eps = false, have_model = false;
};

[[noreturn]] void usage(const char *msg) {
  std::fprintf(stderr, "imm: %s\n", msg);
  std::fprintf(stderr,
               "usage: imm -i GRAPH -k K -e EPS -d IC|LT [-w] [-u] [--reload-binary]\n"
               "           [--distribution uniform|normal|const] [--mean M] [--variance V]\n"
               "           [--scale-factor S] [--disable-renumbering] [--rrsets N]\n"
               "           [-o OUT.json] [--seed S] [--device D]\n");
  std::exit(2);
}

Options parse(int argc, char **argv) {
  Options o;
  auto value = [&](int &i) -> const char * {
    if (i + 1 >= argc) usage((std::string("missing value for ") + argv[i]).c_str());
    return argv[++i];
  };
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "-i" || a == "--input-graph") o.input = value(i);
    else if (a == "-o" || a == "--output") o.output = value(i);
    else if (a == "-k" || a == "--seed-set-size") o.k = std::strtoull(value(i), nullptr, 10), o.have_k = true;
    else if (a == "-e" || a == "--epsilon") o.epsilon = std::strtod(value(i), nullptr), o.have_eps = true;
    else if (a == "-d" || a == "--diffusion-model") o.model = value(i), o.have_model = true;
    else if (a == "-w" || a == "--weighted") o.weighted = true;
    else if (a == "-u" || a == "--undirected") o.undirected = true;
    else if (a == "-p" || a == "--parallel") o.parallel = true;
    else if (a == "--reload-binary") o.reload = true;
    else if (a == "--disable-renumbering") o.renumber = false;
    else if (a == "--distribution") o.distribution = value(i);
    else if (a == "--mean") o.mean = std::strtof(value(i), nullptr);
    else if (a == "--variance") o.variance = std::strtof(value(i), nullptr);
    else if (a == "--scale-factor") o.scale = std::strtof(value(i), nullptr);
    else if (a == "--rrsets") o.rrsets = std::strtoull(value(i), nullptr, 10);
    else if (a == "--seed") o.seed = std::strtoull(value(i), nullptr, 10);
    else if (a == "--device") o.device = std::atoi(value(i));
    else if (a == "--pause-threshold") o.pause = std::strtof(value(i), nullptr);
    else if (a == "--streaming-cpu-teams") o.cpu_teams = std::atol(value(i));
    else if (a == "--streaming-gpu-workers") o.gpu_workers = std::atol(value(i));
    else if (a == "--gpu-batch-size") o.gpu_batch = std::atol(value(i));
    else if (a == "--cpu-batch-size") o.cpu_batch = std::atol(value(i));
    else if (a == "--streaming-gpu-mapping" || a == "--seed-select-max-workers" ||
             a == "--seed-select-max-gpu-workers" || a == "--metall-store-dir" ||
             a == "--rr-store-dir")
      (void)value(i);
    else if (a == "-h" || a == "--help") usage("help");
    else usage(("unknown option " + a).c_str());
  }
  if (o.input.empty()) usage("--input-graph is required");
  if (!o.have_k) usage("--seed-set-size is required");
  if (!o.have_eps) usage("--epsilon is required");
  if (!o.have_model) usage("--diffusion-model is required");
  if (o.model != "IC" && o.model != "LT") usage("--diffusion-model must be IC or LT");
  if (o.distribution != "uniform" && o.distribution != "normal" && o.distribution != "const")
    usage("--distribution must be uniform, normal or const");
  return o;
}

void check(int rc, const char *what) {
  if (rc != RIP_OK) {
    std::fprintf(stderr, "imm: %s failed: %s (code %d)\n", what, rip_last_error(), rc);
    std::exit(1);
  }
}

std::string json_string(const std::string &s) {
  std::string r = "\"";
  for (char c : s) {
    if (c == '"' || c == '\\') r += '\\', r += c;
    else if (static_cast<unsigned char>(c) < 0x20) {
      char b[8];
      std::snprintf(b, sizeof(b), "\\u%04x", c);
      r += b;
    } else r += c;
  }
  return r + "\"";
}

template <typename T>
std::string json_list(const std::vector<T> &v) {
  std::string r = "[";
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) r += ", ";
    r += std::to_string(v[i]);
  }
  return r + "]";
}

std::string num(double x) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", x);
  return b;
}

}  // namespace

int main(int argc, char **argv) {
  const Options o = parse(argc, argv);
  const int model = o.model == "IC" ? RIP_IC : RIP_LT;

  const auto t_load = std::chrono::steady_clock::now();
  rip_host_graph *hg = nullptr;
  if (o.reload) {
    check(rip_load_binary(o.input.c_str(), o.scale, &hg), "loading the binary graph");
  } else {
    rip_load_options lo{};
    lo.model = model;
    lo.weighted = o.weighted;
    lo.undirected = o.undirected;
    lo.renumber = o.renumber;
    lo.distribution = o.distribution == "uniform" ? RIP_DIST_UNIFORM
                      : o.distribution == "normal" ? RIP_DIST_NORMAL : RIP_DIST_CONST;
    lo.mean = o.mean;
    lo.variance = o.variance;
    lo.scale_factor = o.scale;
    lo.seed = o.seed;
    check(rip_load_edge_list(o.input.c_str(), &lo, &hg), "loading the edge list");
  }
  uint32_t n = 0;
  uint64_t m = 0;
  check(rip_host_graph_info(hg, &n, &m), "graph info");
  const uint64_t *index;
  const uint32_t *edges, *rev;
  const float *weights;
  check(rip_host_graph_arrays(hg, &index, &edges, &weights, &rev), "graph arrays");
  rip_graph *g = nullptr;
  check(rip_graph_create(o.device, model, n, m, index, edges, weights, &g), "creating the device graph");
  const double load_ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_load).count();
  std::fprintf(stderr, "imm: %u nodes, %llu edges, loaded in %.0f ms\n", n,
               static_cast<unsigned long long>(m), load_ms);

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
    seeds.assign(o.k, n > 1 ? 1u : 0u);
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
  for (uint32_t &s : seeds) s = rev[s];  // original ids, as G.convertID does

  std::string js = "[\n  {\n";
  auto field = [&](const char *k, const std::string &v, bool last = false) {
    js += "    \"" + std::string(k) + "\": " + v + (last ? "\n" : ",\n");
  };
  field("Algorithm", json_string("IMM"));
  field("Input", json_string(o.input));
  field("Output", json_string(o.output));
  field("DiffusionModel", json_string(o.model));
  field("Epsilon", num(o.epsilon));
  field("K", std::to_string(o.k));
  field("L", "1");
  field("NumThreads", "1");
  field("NumWalkWorkers", std::to_string(o.gpu_workers));
  field("NumCPUTeams", std::to_string(o.cpu_teams));
  field("NumGPUWalkWorkers", std::to_string(o.gpu_workers));
  field("PauseThreshold", num(o.pause));
  field("Total", num(total));
  field("ThetaPrimeDeltas", json_list(deltas));
  field("ThetaEstimation", num(estimation));
  std::string g1 = "[", g2 = "[";
  for (size_t i = 0; i < est_gen.size(); ++i) {
    g1 += (i ? ", " : "") + num(est_gen[i]);
    g2 += (i ? ", " : "") + num(est_sel[i]);
  }
  field("ThetaEstimationGenerateRRR", g1 + "]");
  field("ThetaEstimationMostInfluential", g2 + "]");
  field("Theta", std::to_string(theta));
  field("Counting", "[]");
  field("Pivoting", "[]");
  field("Microbenchmarking", "0");
  field("CPUBatchSize", std::to_string(o.cpu_batch));
  field("GPUBatchSize", std::to_string(o.gpu_batch));
  field("RRRSetSizeBytes", std::to_string(4 * elements));
  field("GenerateRRRSets", num(generate));
  field("FindMostInfluentialSet", num(select));
  field("Seeds", json_list(seeds), true);
  js += "  }\n]\n";

  FILE *f = std::fopen(o.output.c_str(), "w");
  if (!f || std::fputs(js.c_str(), f) < 0 || std::fclose(f) != 0) {
    std::fprintf(stderr, "imm: cannot write %s\n", o.output.c_str());
    return 1;
  }
  std::fprintf(stderr, "imm: theta %llu, %zu seeds, %.1f ms; record in %s\n",
               static_cast<unsigned long long>(theta), seeds.size(), total, o.output.c_str());
  rip_graph_destroy(g);
  rip_host_graph_destroy(hg);
  return 0;
}
