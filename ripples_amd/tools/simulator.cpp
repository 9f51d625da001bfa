// simulator: Monte Carlo spread of the seed sets recorded by imm, through the
// C ABI (rip_spread), with the flags and output of the reference tool
// (tools/simulator.cc:60-171): the graph options of imm, -e/--experiment-file
// (the imm JSON record), --replicas, -o.  The model comes from the record's
// DiffusionModel; seeds are original vertex ids.  The output is an array with
// one object per input record: Input, Algorithm, DiffusionModel, Epsilon, K,
// Seeds, and Simulations = [[activated, steps], ...] per replica.
// --seed (absent from the reference) seeds the weights and replica streams.
#include <algorithm>
#include <cctype>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "ripples_amd.h"

namespace {

// ---- a small JSON reader: enough for the imm experiment record
struct Json {
  enum Kind { Null, Bool, Number, String, Array, Object } kind = Null;
  double num = 0;
  bool b = false;
  std::string str;
  std::vector<Json> arr;
  std::vector<std::pair<std::string, Json>> obj;
  const Json *get(const std::string &k) const {
    for (const auto &kv : obj)
      if (kv.first == k) return &kv.second;
    return nullptr;
  }
};

class Reader {
 public:
  explicit Reader(const std::string &s) : s_(s) {}
  Json parse() {
    Json j = value();
    ws();
    if (i_ != s_.size()) fail("trailing characters");
    return j;
  }

 private:
  [[noreturn]] void fail(const char *what) {
    std::fprintf(stderr, "simulator: bad experiment file: %s at byte %zu\n", what, i_);
    std::exit(1);
  }
  void ws() {
    while (i_ < s_.size() && std::isspace(static_cast<unsigned char>(s_[i_]))) ++i_;
  }
  bool eat(char c) {
    ws();
    if (i_ < s_.size() && s_[i_] == c) {
      ++i_;
      return true;
    }
    return false;
  }
  std::string string_lit() {
    if (!eat('"')) fail("expected a string");
    std::string r;
    while (i_ < s_.size() && s_[i_] != '"') {
      char c = s_[i_++];
      if (c == '\\' && i_ < s_.size()) {
        const char e = s_[i_++];
        switch (e) {
          case 'n': c = '\n'; break;
          case 't': c = '\t'; break;
          case 'r': c = '\r'; break;
          case 'b': c = '\b'; break;
          case 'f': c = '\f'; break;
          case 'u': {
            if (i_ + 4 > s_.size()) fail("bad escape");
            c = static_cast<char>(std::strtol(s_.substr(i_, 4).c_str(), nullptr, 16));
            i_ += 4;
            break;
          }
          default: c = e;
        }
      }
      r += c;
    }
    if (i_ >= s_.size()) fail("unterminated string");
    ++i_;
    return r;
  }
  Json value() {
    ws();
    if (i_ >= s_.size()) fail("unexpected end");
    Json j;
    const char c = s_[i_];
    if (c == '{') {
      ++i_;
      j.kind = Json::Object;
      if (eat('}')) return j;
      do {
        std::string k = string_lit();
        if (!eat(':')) fail("expected ':'");
        j.obj.emplace_back(k, value());
      } while (eat(','));
      if (!eat('}')) fail("expected '}'");
    } else if (c == '[') {
      ++i_;
      j.kind = Json::Array;
      if (eat(']')) return j;
      do j.arr.push_back(value());
      while (eat(','));
      if (!eat(']')) fail("expected ']'");
    } else if (c == '"') {
      j.kind = Json::String;
      j.str = string_lit();
    } else if (s_.compare(i_, 4, "true") == 0) {
      j.kind = Json::Bool, j.b = true, i_ += 4;
    } else if (s_.compare(i_, 5, "false") == 0) {
      j.kind = Json::Bool, i_ += 5;
    } else if (s_.compare(i_, 4, "null") == 0) {
      i_ += 4;
    } else {
      char *end = nullptr;
      j.num = std::strtod(s_.c_str() + i_, &end);
      if (end == s_.c_str() + i_) fail("bad value");
      j.kind = Json::Number;
      i_ = static_cast<size_t>(end - s_.c_str());
    }
    return j;
  }
  const std::string &s_;
  size_t i_ = 0;
};

std::string slurp(const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "rb");
  if (!f) {
    std::fprintf(stderr, "simulator: cannot open %s\n", path.c_str());
    std::exit(1);
  }
  std::string d;
  char b[65536];
  size_t n;
  while ((n = std::fread(b, 1, sizeof(b), f)) > 0) d.append(b, n);
  std::fclose(f);
  return d;
}

std::string jstr(const std::string &s) {
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

std::string jnum(double x) {
  char b[64];
  std::snprintf(b, sizeof(b), "%.17g", x);
  return b;
}

[[noreturn]] void usage(const char *msg) {
  std::fprintf(stderr, "simulator: %s\n", msg);
  std::fprintf(stderr,
               "usage: simulator -i GRAPH -e IMM.json --replicas R [-o OUT.json] [-w] [-u]\n"
               "                 [--reload-binary] [--distribution uniform|normal|const]\n"
               "                 [--mean M] [--variance V] [--scale-factor S]\n"
               "                 [--disable-renumbering] [--seed S] [--device D]\n");
  std::exit(2);
}

void check(int rc, const char *what) {
  if (rc != RIP_OK) {
    std::fprintf(stderr, "simulator: %s failed: %s (code %d)\n", what, rip_last_error(), rc);
    std::exit(1);
  }
}

}  // namespace

int main(int argc, char **argv) {
  std::string input, exp_file, output = "output.json", distribution = "uniform";
  bool weighted = false, undirected = false, reload = false, renumber = true;
  float mean = 0.5f, variance = 1.0f, scale = 1.0f;
  uint64_t replicas = 0, seed = 0;
  int device = 0;
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto value = [&]() -> const char * {
      if (i + 1 >= argc) usage(("missing value for " + a).c_str());
      return argv[++i];
    };
    if (a == "-i" || a == "--input-graph") input = value();
    else if (a == "-e" || a == "--experiment-file") exp_file = value();
    else if (a == "-o" || a == "--output") output = value();
    else if (a == "--replicas") replicas = std::strtoull(value(), nullptr, 10);
    else if (a == "-w" || a == "--weighted") weighted = true;
    else if (a == "-u" || a == "--undirected") undirected = true;
    else if (a == "--reload-binary") reload = true;
    else if (a == "--disable-renumbering") renumber = false;
    else if (a == "--distribution") distribution = value();
    else if (a == "--mean") mean = std::strtof(value(), nullptr);
    else if (a == "--variance") variance = std::strtof(value(), nullptr);
    else if (a == "--scale-factor") scale = std::strtof(value(), nullptr);
    else if (a == "--seed") seed = std::strtoull(value(), nullptr, 10);
    else if (a == "--device") device = std::atoi(value());
    else if (a == "-h" || a == "--help") usage("help");
    else usage(("unknown option " + a).c_str());
  }
  if (input.empty()) usage("--input-graph is required");
  if (exp_file.empty()) usage("--experiment-file is required");
  if (replicas == 0) usage("--replicas is required");

  const std::string text = slurp(exp_file);
  const Json records = Reader(text).parse();
  if (records.kind != Json::Array || records.arr.empty()) {
    std::fprintf(stderr, "simulator: the experiment file holds no records\n");
    return 1;
  }
  const Json *dm = records.arr[0].get("DiffusionModel");
  if (!dm || dm->kind != Json::String || (dm->str != "IC" && dm->str != "LT")) {
    std::fprintf(stderr, "simulator: the record's DiffusionModel must be IC or LT\n");
    return 1;
  }
  const int model = dm->str == "IC" ? RIP_IC : RIP_LT;

  rip_host_graph *hg = nullptr;
  if (reload) {
    check(rip_load_binary(input.c_str(), scale, &hg), "loading the binary graph");
  } else {
    rip_load_options lo{};
    lo.model = model;
    lo.weighted = weighted;
    lo.undirected = undirected;
    lo.renumber = renumber;
    lo.distribution = distribution == "normal" ? RIP_DIST_NORMAL
                      : distribution == "const" ? RIP_DIST_CONST : RIP_DIST_UNIFORM;
    lo.mean = mean;
    lo.variance = variance;
    lo.scale_factor = scale;
    lo.seed = seed;
    check(rip_load_edge_list(input.c_str(), &lo, &hg), "loading the edge list");
  }
  uint32_t n = 0;
  uint64_t m = 0;
  check(rip_host_graph_info(hg, &n, &m), "graph info");
  const uint64_t *index;
  const uint32_t *edges, *rev;
  const float *weights;
  check(rip_host_graph_arrays(hg, &index, &edges, &weights, &rev), "graph arrays");
  std::map<uint32_t, uint32_t> to_internal;  // original id -> internal id
  for (uint32_t v = 0; v < n; ++v) to_internal[rev[v]] = v;
  rip_graph *g = nullptr;
  check(rip_graph_create(device, model, n, m, index, edges, weights, &g), "creating the device graph");

  std::string out = "[\n";
  for (size_t r = 0; r < records.arr.size(); ++r) {
    const Json &rec = records.arr[r];
    const Json *sj = rec.get("Seeds");
    if (!sj || sj->kind != Json::Array) {
      std::fprintf(stderr, "simulator: record %zu has no Seeds\n", r);
      return 1;
    }
    std::vector<uint32_t> seeds;
    std::string seeds_txt = "[";
    for (size_t i = 0; i < sj->arr.size(); ++i) {
      const uint32_t orig = static_cast<uint32_t>(sj->arr[i].num);
      auto it = to_internal.find(orig);
      if (it == to_internal.end()) {
        std::fprintf(stderr, "simulator: seed %u is not a vertex of the graph\n", orig);
        return 1;
      }
      seeds.push_back(it->second);
      seeds_txt += (i ? ", " : "") + std::to_string(orig);
    }
    seeds_txt += "]";
    std::vector<uint64_t> act(replicas);
    std::vector<uint32_t> lev(replicas);
    check(rip_spread(g, seeds.data(), static_cast<uint32_t>(seeds.size()), seed, 0, replicas,
                     act.data(), lev.data()),
          "spread");
    auto field = [&](const char *k) -> std::string {
      const Json *v = rec.get(k);
      if (!v) return "null";
      if (v->kind == Json::String) return jstr(v->str);
      if (v->kind == Json::Number) return jnum(v->num);
      return "null";
    };
    std::string sims = "[";
    double sum = 0;
    for (uint64_t i = 0; i < replicas; ++i) {
      sims += (i ? ", [" : "[") + std::to_string(act[i]) + ", " + std::to_string(lev[i]) + "]";
      sum += static_cast<double>(act[i]);
    }
    sims += "]";
    out += "  {\n    \"Input\": " + field("Input") + ",\n    \"Algorithm\": " + field("Algorithm") +
           ",\n    \"DiffusionModel\": " + jstr(dm->str) + ",\n    \"Epsilon\": " + field("Epsilon") +
           ",\n    \"K\": " + field("K") + ",\n    \"Seeds\": " + seeds_txt +
           ",\n    \"Simulations\": " + sims + "\n  }" + (r + 1 < records.arr.size() ? ",\n" : "\n");
    std::fprintf(stderr, "simulator: record %zu: mean spread %.2f over %llu replicas\n", r,
                 sum / static_cast<double>(replicas), static_cast<unsigned long long>(replicas));
  }
  out += "]\n";
  FILE *f = std::fopen(output.c_str(), "w");
  if (!f || std::fputs(out.c_str(), f) < 0 || std::fclose(f) != 0) {
    std::fprintf(stderr, "simulator: cannot write %s\n", output.c_str());
    return 1;
  }
  rip_graph_destroy(g);
  rip_host_graph_destroy(hg);
  return 0;
}
