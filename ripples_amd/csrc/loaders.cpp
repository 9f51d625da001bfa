// Graph inputs: edge-list files (plain and weighted TSV) and the reference
// binary CSR format, turned into the host transposed CSR (GraphBwd) that
// rip_graph_create takes.  Host code.
//
// Edge lists (include/ripples/loaders.h:138-296): one edge "src dst" or
// "src dst weight" per line; empty lines and lines containing '%' or '#' are
// skipped.  Undirected inputs add the reverse edge after each line.
// Unweighted inputs draw each weight from the weights stream
// (master.split(2,0), tools/imm.cc:172-174: draw i is master output 2i) in
// file order; under LT the edges are then grouped by source in ascending
// order and each group is divided by (one more draw + the group's sum), as
// loaders.h:169-191 does.  Weighted inputs are scaled by scale_factor (<= 1).
//
// Graph construction (include/ripples/graph.h:506-599): with renumbering,
// the vertex ids that occur are mapped to their rank (reverse map = the
// sorted ids); without it ids are kept and n = max id + 1.  Row v of the
// transposed CSR lists the sources of the edges into v in input order
// (duplicates kept), which is the order the reference's construction gives
// with one thread.
//
// Binary CSR (graph.h:700-771, 854-995): u64 0xc0ffee, u64 direction
// (forward 0xf0cacc1a, backward 0xa1ccac0f), u64 weight type (32 float,
// 16 uint16, 8 uint8), u64 n, u64 m, u32 reverse map[n], u64 index[n+1],
// u32 edges[m], weights[m]; little endian.  IMM needs the backward
// direction.  Integer weights are fractions of their type's maximum.
#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <numeric>
#include <string>
#include <vector>

#include "common.h"
#include "internal.h"
#include "ripples_amd.h"

namespace rip {
namespace {

constexpr uint64_t kBinMagic = 0xc0ffeeull;
constexpr uint64_t kBinForward = 0xf0cacc1aull;
constexpr uint64_t kBinBackward = 0xa1ccac0full;

struct RawEdge {
  uint32_t src, dst;
  float w;
};

std::string read_file(const char *path) {
  FILE *f = std::fopen(path, "rb");
  RIP_REQUIRE(f != nullptr, -1, std::string("load: cannot open ") + path + ": " + std::strerror(errno));
  std::string data;
  char buf[1 << 16];
  size_t got;
  while ((got = std::fread(buf, 1, sizeof(buf), f)) > 0) data.append(buf, got);
  const bool err = std::ferror(f) != 0;
  std::fclose(f);
  RIP_REQUIRE(!err, -1, std::string("load: read error on ") + path);
  return data;
}

// Inverse of the standard normal CDF: Acklam's rational approximation
// (relative error below 1.2e-9), refined by one Halley step on erfc.
double inv_norm_cdf(double p) {
  static const double a[] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                             1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
  static const double b[] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                             6.680131188771972e+01, -1.328068155288572e+01};
  static const double c[] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                             -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
  static const double d[] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                             3.754408661907416e+00};
  const double lo = 0.02425, hi = 1 - lo;
  double x;
  if (p < lo) {
    const double q = std::sqrt(-2 * std::log(p));
    x = (((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  } else if (p <= hi) {
    const double q = p - 0.5, r = q * q;
    x = (((((a[0] * r + a[1]) * r + a[2]) * r + a[3]) * r + a[4]) * r + a[5]) * q /
        (((((b[0] * r + b[1]) * r + b[2]) * r + b[3]) * r + b[4]) * r + 1);
  } else {
    const double q = std::sqrt(-2 * std::log(1 - p));
    x = -(((((c[0] * q + c[1]) * q + c[2]) * q + c[3]) * q + c[4]) * q + c[5]) /
        ((((d[0] * q + d[1]) * q + d[2]) * q + d[3]) * q + 1);
  }
  const double e = 0.5 * std::erfc(-x / std::sqrt(2.0)) - p;
  const double u = e * std::sqrt(2 * M_PI) * std::exp(x * x / 2);
  return x - u / (1 + x * u / 2);
}

// Weight draws for unweighted inputs: draw i is master output 2i of the
// seed's stream, mapped to (0, 1] and shaped by the distribution.
class WeightStream {
 public:
  explicit WeightStream(const rip_load_options &o)
      : o_(o), x_(aff_apply(master_step(), o.seed)), step2_(aff_pow(master_step(), 2)) {
    if (o.distribution == RIP_DIST_NORMAL) {
      const double s = o.variance;
      pa_ = 0.5 * std::erfc((o.mean - 0.0) / s / std::sqrt(2.0));  // Phi((0 - mu) / s)
      pb_ = 0.5 * std::erfc((o.mean - 1.0) / s / std::sqrt(2.0));  // Phi((1 - mu) / s)
    }
  }
  float next() {
    const float u = u01(x_);
    x_ = aff_apply(step2_, x_);
    switch (o_.distribution) {
      case RIP_DIST_CONST: return o_.mean;
      case RIP_DIST_NORMAL: {
        // normal(mean, variance as its scale) truncated to [0, 1], by
        // inverting the CDF at a uniform point of [Phi(a), Phi(b)]
        const double p = pa_ + static_cast<double>(u) * (pb_ - pa_);
        double v = o_.mean + o_.variance * inv_norm_cdf(std::min(std::max(p, 1e-300), 1 - 1e-16));
        v = std::min(std::max(v, 0.0), 1.0);
        return static_cast<float>(v) * o_.scale_factor;
      }
      default: return u * o_.scale_factor;
    }
  }

 private:
  rip_load_options o_;
  uint64_t x_;
  Affine step2_;
  double pa_ = 0, pb_ = 1;
};

bool parse_u32(const char *&p, const char *end, uint32_t *out) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  if (p == end || *p < '0' || *p > '9') return false;
  uint64_t v = 0;
  while (p < end && *p >= '0' && *p <= '9') {
    v = v * 10 + static_cast<uint64_t>(*p - '0');
    if (v > 0xFFFFFFFFull) return false;
    ++p;
  }
  *out = static_cast<uint32_t>(v);
  return true;
}

bool parse_float(const char *&p, const char *end, float *out) {
  while (p < end && (*p == ' ' || *p == '\t' || *p == '\r')) ++p;
  if (p == end) return false;
  char tmp[64];
  size_t n = 0;
  while (p + n < end && n < sizeof(tmp) - 1 && p[n] != ' ' && p[n] != '\t' && p[n] != '\r' &&
         p[n] != '\n')
    ++n;
  std::memcpy(tmp, p, n);
  tmp[n] = 0;
  char *stop = nullptr;
  const float v = std::strtof(tmp, &stop);
  if (stop == tmp) return false;
  p += stop - tmp;
  *out = v;
  return true;
}

void check_options(const rip_load_options &o) {
  RIP_REQUIRE(o.model == RIP_IC || o.model == RIP_LT, -1, "load: model must be RIP_IC or RIP_LT");
  RIP_REQUIRE(o.scale_factor > 0.0f && o.scale_factor <= 1.0f, -1,
              "load: scale factor must be in (0, 1]");
  RIP_REQUIRE(o.distribution == RIP_DIST_UNIFORM || o.distribution == RIP_DIST_NORMAL ||
                  o.distribution == RIP_DIST_CONST,
              -1, "load: unknown weight distribution");
  if (o.distribution == RIP_DIST_CONST)
    RIP_REQUIRE(o.mean >= 0.0f && o.mean <= 1.0f, -1, "load: const weight must be in [0, 1]");
  if (o.distribution == RIP_DIST_NORMAL)
    RIP_REQUIRE(o.variance > 0.0f, -1, "load: normal distribution needs variance > 0");
}

void build_csr(const std::vector<RawEdge> &list, bool renumber, rip_host_graph *g) {
  uint64_t maxid = 0;
  std::vector<uint32_t> ids;
  ids.reserve(list.size() * 2);
  for (const RawEdge &e : list) {
    ids.push_back(e.src);
    ids.push_back(e.dst);
    maxid = std::max<uint64_t>(maxid, std::max(e.src, e.dst));
  }
  std::sort(ids.begin(), ids.end());
  ids.erase(std::unique(ids.begin(), ids.end()), ids.end());
  auto map = [&](uint32_t v) -> uint32_t {
    return renumber ? static_cast<uint32_t>(std::lower_bound(ids.begin(), ids.end(), v) - ids.begin())
                    : v;
  };
  uint64_t n;
  if (renumber) {
    g->reverse_map = ids;
    n = ids.size();
  } else {
    n = list.empty() ? 0 : maxid + 1;
    g->reverse_map.resize(n);
    std::iota(g->reverse_map.begin(), g->reverse_map.end(), 0u);
  }
  RIP_REQUIRE(n < 0xFFFFFFFFull, -1, "load: more than 2^32-1 vertices");
  const uint64_t m = list.size();
  g->index.assign(n + 1, 0);
  std::vector<uint32_t> src(m), dst(m);
  for (uint64_t i = 0; i < m; ++i) {
    src[i] = map(list[i].src);
    dst[i] = map(list[i].dst);
    ++g->index[dst[i] + 1];
  }
  for (uint64_t v = 0; v < n; ++v) g->index[v + 1] += g->index[v];
  std::vector<uint64_t> cur(g->index.begin(), g->index.end() - (n ? 1 : 0));
  g->edges.resize(m);
  g->weights.resize(m);
  for (uint64_t i = 0; i < m; ++i) {
    const uint64_t at = cur[dst[i]]++;
    g->edges[at] = src[i];
    g->weights[at] = list[i].w;
  }
}

}  // namespace

void load_edge_list(const char *path, const rip_load_options &o, rip_host_graph *g) {
  RIP_REQUIRE(path != nullptr, -1, "load: path is NULL");
  check_options(o);
  const std::string data = read_file(path);
  std::vector<RawEdge> list;
  const char *p = data.data(), *end = p + data.size();
  uint64_t line_no = 0;
  while (p < end) {
    const char *eol = static_cast<const char *>(std::memchr(p, '\n', end - p));
    if (!eol) eol = end;
    ++line_no;
    const char *q = p;
    bool skip = true;
    for (const char *r = p; r < eol; ++r) {
      if (*r == '%' || *r == '#') {
        skip = true;
        break;
      }
      if (*r != ' ' && *r != '\t' && *r != '\r') skip = false;
    }
    if (!skip) {
      RawEdge e{0, 0, 0.0f};
      bool ok = parse_u32(q, eol, &e.src) && parse_u32(q, eol, &e.dst);
      if (ok && o.weighted) ok = parse_float(q, eol, &e.w);
      RIP_REQUIRE(ok, -1, "load: malformed edge on line " + std::to_string(line_no) + " of " + path);
      list.push_back(e);
      if (o.undirected) list.push_back(RawEdge{e.dst, e.src, e.w});
    }
    p = eol + 1;
  }
  if (o.weighted) {
    if (o.scale_factor < 1.0f)
      for (RawEdge &e : list) e.w *= o.scale_factor;
  } else {
    WeightStream gen(o);
    for (RawEdge &e : list) e.w = gen.next();
    if (o.model == RIP_LT) {
      std::stable_sort(list.begin(), list.end(),
                       [](const RawEdge &a, const RawEdge &b) { return a.src < b.src; });
      for (size_t b = 0; b < list.size();) {
        size_t e = b;
        while (e < list.size() && list[e].src == list[b].src) ++e;
        float total = gen.next();  // the share of not being influenced
        for (size_t i = b; i < e; ++i) total += list[i].w;
        for (size_t i = b; i < e; ++i) list[i].w /= total;
        b = e;
      }
    }
  }
  build_csr(list, o.renumber != 0, g);
}

void load_binary(const char *path, float scale_factor, rip_host_graph *g) {
  RIP_REQUIRE(path != nullptr, -1, "load: path is NULL");
  RIP_REQUIRE(scale_factor > 0.0f && scale_factor <= 1.0f, -1, "load: scale factor must be in (0, 1]");
  const std::string data = read_file(path);
  size_t at = 0;
  auto take = [&](void *dst, size_t bytes) {
    RIP_REQUIRE(at + bytes <= data.size(), -1, std::string("load: truncated binary graph ") + path);
    std::memcpy(dst, data.data() + at, bytes);
    at += bytes;
  };
  uint64_t hdr[5];
  take(hdr, sizeof(hdr));
  RIP_REQUIRE(hdr[0] == kBinMagic, -1, "load: not a binary graph (endianness check failed)");
  RIP_REQUIRE(hdr[1] == kBinBackward || hdr[1] == kBinForward, -1, "load: unknown direction policy");
  RIP_REQUIRE(hdr[1] == kBinBackward, -1,
              "load: IMM needs the transposed (backward) graph; regenerate with --transpose");
  RIP_REQUIRE(hdr[2] == 32 || hdr[2] == 16 || hdr[2] == 8, -1, "load: unknown weight type");
  const uint64_t n = hdr[3], m = hdr[4];
  RIP_REQUIRE(n < 0xFFFFFFFFull && m < (1ull << 40), -1, "load: implausible graph size");
  g->reverse_map.resize(n);
  take(g->reverse_map.data(), 4 * n);
  g->index.resize(n + 1);
  take(g->index.data(), 8 * (n + 1));
  g->edges.resize(m);
  take(g->edges.data(), 4 * m);
  g->weights.resize(m);
  if (hdr[2] == 32) {
    take(g->weights.data(), 4 * m);
  } else if (hdr[2] == 16) {
    std::vector<uint16_t> w(m);
    take(w.data(), 2 * m);
    for (uint64_t i = 0; i < m; ++i) g->weights[i] = static_cast<float>(w[i]) / 65535.0f;
  } else {
    std::vector<uint8_t> w(m);
    take(w.data(), m);
    for (uint64_t i = 0; i < m; ++i) g->weights[i] = static_cast<float>(w[i]) / 255.0f;
  }
  RIP_REQUIRE(g->index[0] == 0 && g->index[n] == m, -1, "load: inconsistent CSR index");
  for (uint64_t v = 0; v < n; ++v)
    RIP_REQUIRE(g->index[v + 1] >= g->index[v], -1, "load: CSR index is not monotone");
  for (uint64_t i = 0; i < m; ++i) RIP_REQUIRE(g->edges[i] < n, -1, "load: edge endpoint out of range");
  if (scale_factor < 1.0f)
    for (float &w : g->weights) w *= scale_factor;
}

void dump_binary(const rip_host_graph *g, const char *path) {
  RIP_REQUIRE(path != nullptr, -1, "dump: path is NULL");
  FILE *f = std::fopen(path, "wb");
  RIP_REQUIRE(f != nullptr, -1, std::string("dump: cannot open ") + path + ": " + std::strerror(errno));
  const uint64_t n = g->reverse_map.size(), m = g->edges.size();
  const uint64_t hdr[5] = {kBinMagic, kBinBackward, 32, n, m};
  bool ok = std::fwrite(hdr, sizeof(hdr), 1, f) == 1;
  ok = ok && (n == 0 || std::fwrite(g->reverse_map.data(), 4, n, f) == n);
  ok = ok && std::fwrite(g->index.data(), 8, n + 1, f) == n + 1;
  ok = ok && (m == 0 || std::fwrite(g->edges.data(), 4, m, f) == m);
  ok = ok && (m == 0 || std::fwrite(g->weights.data(), 4, m, f) == m);
  ok = (std::fclose(f) == 0) && ok;
  RIP_REQUIRE(ok, -1, std::string("dump: write error on ") + path);
}

}  // namespace rip
