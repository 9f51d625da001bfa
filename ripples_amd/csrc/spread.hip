// Forward Monte Carlo spread of a seed set under IC or LT: the check of a
// seed set's influence (include/ripples/diffusion_simulation.h:73-178, driven
// by tools/simulator.cc:100-171), one replica per persistent workgroup at a
// time.
//
// Replica r draws from the sample stream r of the given seed (the same
// per-sample construction the samplers use): element 1 + f is the IC coin of
// forward edge f, live iff (x >> 40) < floor(w * 2^24); element 1 + v is the
// LT threshold of v, U = ((x >> 40) + 1) * 2^-24.  The forward CSR lists, for
// each u, the targets of the transposed entries with source u in transposed
// position order; it is built once per graph on first use.
//
// IC: level-synchronous breadth-first search; each wave takes frontier
// vertices and its lanes take 64 consecutive out-edges per round.  Levels
// count the layers processed, the seed layer included.
// LT: per level, the inactive out-neighbours of the vertices activated in
// the previous level become candidates (deduplicated by a bitmap); a wave per
// candidate reads its in-row 64 edges at a time, and the weights of active
// in-neighbours are added in CSR order; a candidate activates when the sum
// reaches its threshold.  Activations take effect at the end of the level.
// Levels count the iterations, the last one (activating nothing) included.
#include <algorithm>
#include <vector>

#include "device_util.h"
#include "internal.h"

namespace rip {
namespace {

struct alignas(16) FwdRec {
  uint64_t ja, jb;  // P^(1 + fbegin(u)): coin of u's first out-edge
  uint64_t ta, tb;  // P^(1 + u): LT threshold of u
  uint32_t begin, end;
  uint32_t pad[2];
};

__global__ void build_fwd_kernel(const unsigned long long *fidx, uint32_t n, Affine step,
                                 FwdRec *rec) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= n) return;
  FwdRec r;
  const Affine j = aff_pow(step, 1 + fidx[u]);
  const Affine t = aff_pow(step, 1 + static_cast<uint64_t>(u));
  r.ja = j.a;
  r.jb = j.b;
  r.ta = t.a;
  r.tb = t.b;
  r.begin = static_cast<uint32_t>(fidx[u]);
  r.end = static_cast<uint32_t>(fidx[u + 1]);
  r.pad[0] = r.pad[1] = 0;
  rec[u] = r;
}

struct SpreadParams {
  const FwdRec *fwd;
  const uint32_t *fdst;
  const uint32_t *fval;  // IC threshold of each forward edge
  const VertexRec *vrec; // transposed rows (LT in-edges)
  const uint32_t *ids;
  const uint32_t *val;   // LT weight bits of each transposed edge
  uint32_t n;
  int model;
  const uint32_t *seeds;
  uint32_t nseeds;
  uint64_t seed;
  uint64_t first;
  uint64_t count;
  unsigned long long *act_out;
  uint32_t *lev_out;
  unsigned long long *ctl;  // [0]: replica counter
  uint32_t *scratch;        // per workgroup: act bitmap, cand bitmap, queue, candidates
  uint64_t words;           // 32-bit words per bitmap
  Affine step;
};

__device__ inline bool test_bit(const uint32_t *bm, uint32_t v) { return (bm[v >> 5] >> (v & 31)) & 1u; }

// Sets bit v; true when this call set it.
__device__ inline bool set_bit(uint32_t *bm, uint32_t v) {
  const uint32_t b = 1u << (v & 31);
  return !(atomicOr(&bm[v >> 5], b) & b);
}

template <int BT>
__global__ void __launch_bounds__(BT) spread_kernel(SpreadParams P) {
  constexpr uint32_t NW = BT / 64;
  __shared__ uint32_t s_len, s_nc;
  __shared__ unsigned long long s_r;
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x >> 6;
  const uint64_t nn = P.n;
  uint32_t *act = P.scratch + static_cast<uint64_t>(blockIdx.x) * (2 * P.words + 2 * nn);
  uint32_t *cand = act + P.words;
  uint32_t *queue = cand + P.words;
  uint32_t *clist = queue + nn;
  const Affine jl = aff_pow(P.step, lane);
  const Affine j64 = aff_pow(P.step, 64);

  for (;;) {
    if (threadIdx.x == 0) {
      s_r = atomicAdd(&P.ctl[0], 1ull);
      s_len = 0;
      s_nc = 0;
    }
    __syncthreads();
    const unsigned long long r = s_r;
    if (r >= P.count) break;
    const uint64_t z = sample_z(P.seed, P.first + r);
    for (uint32_t i = threadIdx.x; i < P.nseeds; i += BT) {
      const uint32_t v = P.seeds[i];
      if (set_bit(act, v)) queue[atomicAdd(&s_len, 1u)] = v;
    }
    __syncthreads();
    uint32_t lo = 0, hi = s_len, lev = 0;
    __syncthreads();
    if (P.model == RIP_IC) {
      while (lo < hi) {
        ++lev;
        for (uint32_t q = lo + wid; q < hi; q += NW) {
          const uint32_t u = queue[q];
          const FwdRec fr = P.fwd[u];
          uint64_t x = fr.ja * z + fr.jb;
          x = jl.a * x + jl.b;  // coin of edge begin + lane
          for (uint32_t f0 = fr.begin; f0 < fr.end; f0 += 64) {
            const uint32_t f = f0 + lane;
            if (f < fr.end && static_cast<uint32_t>(x >> 40) < P.fval[f]) {
              const uint32_t v = P.fdst[f];
              if (!test_bit(act, v) && set_bit(act, v)) queue[atomicAdd(&s_len, 1u)] = v;
            }
            x = j64.a * x + j64.b;
          }
        }
        __syncthreads();
        lo = hi;
        hi = s_len;
        __syncthreads();
      }
    } else {
      do {
        ++lev;
        // candidates: inactive out-neighbours of the last level's activations
        for (uint32_t q = lo + wid; q < hi; q += NW) {
          const FwdRec fr = P.fwd[queue[q]];
          for (uint32_t f = fr.begin + lane; f < fr.end; f += 64) {
            const uint32_t v = P.fdst[f];
            if (!test_bit(act, v) && !test_bit(cand, v) && set_bit(cand, v))
              clist[atomicAdd(&s_nc, 1u)] = v;
          }
        }
        __syncthreads();
        const uint32_t nc = s_nc;
        // a wave per candidate: weights of active in-neighbours in CSR order
        for (uint32_t c = wid; c < nc; c += NW) {
          const uint32_t v = clist[c];
          const VertexRec vr = P.vrec[v];
          float total = 0.0f;
          for (uint32_t e0 = vr.begin; e0 < vr.end; e0 += 64) {
            const uint32_t e = e0 + lane;
            const bool on = e < vr.end && test_bit(act, P.ids[e]);
            const float wv = on ? __uint_as_float(P.val[e]) : 0.0f;
            uint64_t bal = __ballot(on);
            while (bal) {
              const int b = __ffsll(static_cast<unsigned long long>(bal)) - 1;
              total += __shfl(wv, b, 64);
              bal &= bal - 1;
            }
          }
          const FwdRec fr = P.fwd[v];
          const float th = u01(fr.ta * z + fr.tb);
          if (lane == 0 && total >= th) queue[atomicAdd(&s_len, 1u)] = v;
        }
        __syncthreads();
        const uint32_t len = s_len;
        for (uint32_t i = hi + threadIdx.x; i < len; i += BT) set_bit(act, queue[i]);
        for (uint32_t i = threadIdx.x; i < nc; i += BT) cand[clist[i] >> 5] = 0u;
        __syncthreads();
        if (threadIdx.x == 0) s_nc = 0;
        lo = hi;
        hi = len;
        __syncthreads();
      } while (lo < hi);
    }
    const uint32_t len = s_len;
    if (threadIdx.x == 0) {
      P.act_out[r] = len;
      P.lev_out[r] = lev;
    }
    for (uint32_t i = threadIdx.x; i < len; i += BT) act[queue[i] >> 5] = 0u;
    __syncthreads();
  }
}

}  // namespace

// Forward CSR of the graph (once): rows by source in transposed-position
// order, with the per-vertex stream jumps.
static void ensure_forward(rip_graph *g) {
  if (g->fwd.d_rec) return;
  const uint32_t n = g->n;
  const uint64_t m = g->m;
  std::vector<VertexRec> vr(n);
  std::vector<uint32_t> ids(m + kEdgePad), val(m + kEdgePad);
  RIP_HIP(hipMemcpy(vr.data(), g->d_vrec, sizeof(VertexRec) * n, hipMemcpyDeviceToHost));
  RIP_HIP(hipMemcpy(ids.data(), g->d_ids, sizeof(uint32_t) * (m + kEdgePad), hipMemcpyDeviceToHost));
  RIP_HIP(hipMemcpy(val.data(), g->d_val, sizeof(uint32_t) * (m + kEdgePad), hipMemcpyDeviceToHost));
  std::vector<unsigned long long> fidx(static_cast<size_t>(n) + 1, 0);
  for (uint64_t i = 0; i < m; ++i) ++fidx[ids[i] + 1];
  for (uint32_t u = 0; u < n; ++u) fidx[u + 1] += fidx[u];
  std::vector<unsigned long long> cur(fidx.begin(), fidx.end() - 1);
  std::vector<uint32_t> fdst(m + 1), fval(m + 1);
  for (uint32_t v = 0; v < n; ++v)
    for (uint32_t i = vr[v].begin; i < vr[v].end; ++i) {
      const unsigned long long f = cur[ids[i]]++;
      fdst[f] = v;
      fval[f] = val[i];  // IC: the edge's integer threshold
    }
  unsigned long long *d_fidx = nullptr;
  try {
    g->fwd.d_rec = dev_alloc<uint8_t>(sizeof(FwdRec) * n);
    g->fwd.d_dst = dev_alloc<uint32_t>(m + 1);
    g->fwd.d_val = dev_alloc<uint32_t>(m + 1);
    d_fidx = dev_alloc<unsigned long long>(static_cast<size_t>(n) + 1);
    RIP_HIP(hipMemcpy(d_fidx, fidx.data(), sizeof(unsigned long long) * (n + 1), hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(g->fwd.d_dst, fdst.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice));
    RIP_HIP(hipMemcpy(g->fwd.d_val, fval.data(), sizeof(uint32_t) * (m + 1), hipMemcpyHostToDevice));
    hipLaunchKernelGGL(build_fwd_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d_fidx, n, g->step,
                       reinterpret_cast<FwdRec *>(g->fwd.d_rec));
    RIP_HIP(hipGetLastError());
    RIP_HIP(hipDeviceSynchronize());
  } catch (...) {
    dev_free(d_fidx);
    dev_free(g->fwd.d_rec);
    dev_free(g->fwd.d_dst);
    dev_free(g->fwd.d_val);
    throw;
  }
  dev_free(d_fidx);
}

void spread_run(rip_graph *g, const uint32_t *seeds, uint32_t nseeds, uint64_t seed,
                uint64_t first, uint64_t count, uint64_t *activated, uint32_t *levels) {
  RIP_REQUIRE(nseeds == 0 || seeds, -1, "spread: seeds is NULL");
  RIP_REQUIRE(count == 0 || (activated && levels), -1, "spread: output arrays are NULL");
  for (uint32_t i = 0; i < nseeds; ++i)
    RIP_REQUIRE(seeds[i] < g->n, -1, "spread: seed vertex out of range");
  if (count == 0) return;
  std::lock_guard<std::mutex> lock(g->sampler_mu);
  RIP_HIP(hipSetDevice(g->device));
  ensure_forward(g);
  constexpr int BT = 256;
  const uint64_t words = (static_cast<uint64_t>(g->n) + 31) / 32;
  const uint64_t per_wg = 2 * words + 2 * static_cast<uint64_t>(g->n);  // 32-bit words
  size_t free_b = 0, total_b = 0;
  RIP_HIP(hipMemGetInfo(&free_b, &total_b));
  const uint64_t held = g->fwd.wgs * g->fwd.per_wg * 4;
  const uint64_t budget = std::min<uint64_t>((free_b + held) / 8, 4ull << 30);
  uint64_t wgs = std::max<uint64_t>(1, std::min<uint64_t>({count, 1024, budget / (per_wg * 4)}));
  if (g->fwd.per_wg != per_wg || g->fwd.wgs < wgs) {
    dev_free(g->fwd.d_scratch);
    g->fwd.d_scratch = dev_alloc<uint32_t>(wgs * per_wg);
    RIP_HIP(hipMemset(g->fwd.d_scratch, 0, sizeof(uint32_t) * wgs * per_wg));
    g->fwd.wgs = wgs;
    g->fwd.per_wg = per_wg;
  }
  wgs = std::min<uint64_t>(wgs, g->fwd.wgs);
  unsigned long long *d_ctl = nullptr, *d_act = nullptr;
  uint32_t *d_lev = nullptr, *d_seeds = nullptr;
  try {
    d_ctl = dev_alloc<unsigned long long>(1);
    d_act = dev_alloc<unsigned long long>(count);
    d_lev = dev_alloc<uint32_t>(count);
    d_seeds = dev_alloc<uint32_t>(nseeds);
    RIP_HIP(hipMemset(d_ctl, 0, sizeof(unsigned long long)));
    if (nseeds)
      RIP_HIP(hipMemcpy(d_seeds, seeds, sizeof(uint32_t) * nseeds, hipMemcpyHostToDevice));
    SpreadParams P;
    P.fwd = reinterpret_cast<const FwdRec *>(g->fwd.d_rec);
    P.fdst = g->fwd.d_dst;
    P.fval = g->fwd.d_val;
    P.vrec = g->d_vrec;
    P.ids = g->d_ids;
    P.val = g->d_val;
    P.n = g->n;
    P.model = g->model;
    P.seeds = d_seeds;
    P.nseeds = nseeds;
    P.seed = seed;
    P.first = first;
    P.count = count;
    P.act_out = d_act;
    P.lev_out = d_lev;
    P.ctl = d_ctl;
    P.scratch = g->fwd.d_scratch;
    P.words = words;
    P.step = g->step;
    hipLaunchKernelGGL((spread_kernel<BT>), dim3(static_cast<uint32_t>(wgs)), dim3(BT), 0, 0, P);
    RIP_HIP(hipGetLastError());
    RIP_HIP(hipDeviceSynchronize());
    std::vector<unsigned long long> a(count);
    RIP_HIP(hipMemcpy(a.data(), d_act, sizeof(unsigned long long) * count, hipMemcpyDeviceToHost));
    RIP_HIP(hipMemcpy(levels, d_lev, sizeof(uint32_t) * count, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < count; ++i) activated[i] = a[i];
  } catch (...) {
    dev_free(d_ctl);
    dev_free(d_act);
    dev_free(d_lev);
    dev_free(d_seeds);
    throw;
  }
  dev_free(d_ctl);
  dev_free(d_act);
  dev_free(d_lev);
  dev_free(d_seeds);
}

}  // namespace rip
