// Shared parts of the imm and opimc command-line tools: the reference's
// flags (include/ripples/configuration.h, imm_configuration.h,
// tim_configuration.h, opimc_configuration.h), graph loading through the C
// ABI, and the small JSON writer of the experiment records.
#pragma once

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "ripples_amd.h"

namespace cli {

extern const char *kTool;   // program name in messages
extern const char *kUsage;  // usage lines

struct Options {
  std::string input, output = "output.json", model = "IC", distribution = "uniform";
  bool weighted = false, undirected = false, reload = false, renumber = true, parallel = false;
  float mean = 0.5f, variance = 1.0f, scale = 1.0f, pause = 1.0f, delta = 0.05f;
  uint64_t k = 10, rrsets = 0, seed = 0;
  double epsilon = 0.5;
  int device = 0;
  long cpu_teams = 0, gpu_workers = 1, gpu_batch = 64, cpu_batch = 64;
  bool have_k = false, have_eps = false, have_model = false;
};

[[noreturn]] inline void usage(const char *msg) {
  std::fprintf(stderr, "%s: %s\n", kTool, msg);
  std::fprintf(stderr, "%s", kUsage);
  std::exit(2);
}

// with_delta: accept --delta (OPIMCConfiguration, opimc_configuration.h:47-56)
inline Options parse(int argc, char **argv, bool with_delta) {
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
    else if (a == "--rrsets" && !with_delta) o.rrsets = std::strtoull(value(i), nullptr, 10);
    else if (a == "--delta" && with_delta) o.delta = std::strtof(value(i), nullptr);
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

inline void check(int rc, const char *what) {
  if (rc != RIP_OK) {
    std::fprintf(stderr, "%s: %s failed: %s (code %d)\n", kTool, what, rip_last_error(), rc);
    std::exit(1);
  }
}

inline std::string json_string(const std::string &s) {
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

inline std::string num(double x) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", x);
  return b;
}

inline std::string num_list(const std::vector<double> &v) {
  std::string r = "[";
  for (size_t i = 0; i < v.size(); ++i) r += (i ? ", " : "") + num(v[i]);
  return r + "]";
}

// The input graph as the reference's loadGraph reads it, on the device.
struct Loaded {
  rip_host_graph *hg = nullptr;
  rip_graph *g = nullptr;
  uint32_t n = 0;
  uint64_t m = 0;
  const uint32_t *reverse = nullptr;  // internal id -> original id
  ~Loaded() {
    if (g) rip_graph_destroy(g);
    if (hg) rip_host_graph_destroy(hg);
  }
};

inline void load(const Options &o, int model, Loaded *L) {
  const auto t_load = std::chrono::steady_clock::now();
  if (o.reload) {
    check(rip_load_binary(o.input.c_str(), o.scale, &L->hg), "loading the binary graph");
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
    check(rip_load_edge_list(o.input.c_str(), &lo, &L->hg), "loading the edge list");
  }
  check(rip_host_graph_info(L->hg, &L->n, &L->m), "graph info");
  const uint64_t *index;
  const uint32_t *edges;
  const float *weights;
  check(rip_host_graph_arrays(L->hg, &index, &edges, &weights, &L->reverse), "graph arrays");
  check(rip_graph_create(o.device, model, L->n, L->m, index, edges, weights, &L->g),
        "creating the device graph");
  const double ms =
      std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_load).count();
  std::fprintf(stderr, "%s: %u nodes, %llu edges, loaded in %.0f ms\n", kTool, L->n,
               static_cast<unsigned long long>(L->m), ms);
}

// A JSON object of one experiment, fields in insertion order.
struct Record {
  std::string js = "[\n  {\n";
  bool first = true;
  void field(const char *k, const std::string &v) {
    if (!first) js += ",\n";
    first = false;
    js += "    \"" + std::string(k) + "\": " + v;
  }
  bool write(const std::string &path) const {
    const std::string out = js + "\n  }\n]\n";
    FILE *f = std::fopen(path.c_str(), "w");
    return f && std::fputs(out.c_str(), f) >= 0 && std::fclose(f) == 0;
  }
};

}  // namespace cli
