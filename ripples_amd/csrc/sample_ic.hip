// Fused multi-colour IC sampler: reverse BFS of up to 64 RRR samples at once.
//
// Follows the IC semantics of AddRRRSet (include/ripples/add_rrrset.h:105-157:
// an in-edge (u -> v) is live when uniform01 <= w) and the colour-per-sample
// idea of BatchedBFSNeighborColor (batched_add_rrrset.h:184-257) and
// GPUBatchedBFSMultiColorFused (include/ripples/gpu/bfs.h:1338-1680), but with
// the per-sample stream contract of common.h: the coin of (sample s, edge e)
// depends only on (seed, s, e), so the set is the reverse reachability of the
// root over live edges, independent of traversal order.
//
// Execution: a persistent grid, one workgroup per resident colour group
// ("slot").  Each workgroup claims groups of `colors` consecutive samples
// from a device counter and runs a level-synchronous BFS for all colours:
//   visited[v] / pending[v] : 64-bit colour masks per vertex (per slot)
//   frontier                : sparse queue of vertices whose pending mask
//                             went 0 -> nonzero (each vertex at most once per
//                             level); masks are read and cleared when a chunk
//                             of the queue is staged into LDS
//   expansion               : the chunk's rows are split into their gap
//                             blocks (ic_gap.h), balanced over all lanes by a
//                             workgroup scan + binary search in LDS; a lane
//                             draws the candidates of its block for each
//                             colour pending at the row, one live edge per
//                             step of the wave
//   discovery               : atomicOr on visited, then on pending; queue and
//                             union bookkeeping are wave-aggregated
//                             (ballot/popcount) into LDS counters
//   output                  : the union of the group is kept in a two-level
//                             bitmap, read back in ascending vertex order,
//                             and each colour's sorted list is scattered to
//                             the pool with ballot ranks.  Everything touched
//                             is cleared again, so no O(n) reset per group.
#include <algorithm>

#include "device_util.h"
#include "ic_common.h"
#include "ic_gap.h"
#include "internal.h"

namespace rip {
namespace {

struct IcParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  Affine step;
  uint64_t seed;
  uint64_t first;
  uint64_t count;
  uint32_t colors;
  uint64_t ngroups;
  unsigned long long *vp;
  uint32_t *bm0;
  uint32_t *bm1;
  uint32_t *queue;
  uint32_t nw0, nw1;
  uint64_t npad;  // queue words per buffer (even)
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
  uint32_t heavy_deg;   // rows at least this long wait for the hub phase
  uint32_t *hub;        // slots x 2 x hub_cap deferred-hub lists
  uint64_t hub_cap;
  const uint32_t *sample_ids;  // optional: local index -> sample index (minus first)
  uint32_t precheck;    // read visited masks before their atomics (giant components)
  // column store (nullptr: every set a list): a group whose sets hold more
  // than 2 hot_n hot members in all takes the next column (columns
  // group_base, group_base + 1, ... as counted in ctl[kCtlGroupCounter2])
  unsigned long long *const *colptr;
  uint32_t *gsample;
  uint32_t *gref;
  const uint32_t *hot_index;
  uint32_t hot_n;
  uint32_t group_base;
};

__device__ inline uint64_t sample_id(const IcParams &P, uint64_t local) {
  return P.first + (P.sample_ids ? P.sample_ids[local] : local);
}

// Mark (u, bits) discovered.  Every lane of the wave must call it.  With
// precheck the visited mask is read first and the atomic skipped when no
// colour is new: within a group masks only gain bits, so a read that lags
// can only cost an extra atomic.  In a giant component most edges reach
// vertices every colour has visited, and their atomics, many on the same hub
// words, were most of the expansion time; elsewhere the read is one more
// round trip.  The read is a device-scope atomic load, served by L2 like the
// atomics: a plain load could meet a vector-L1 line from an earlier group,
// whose masks have since been cleared.
template <bool kPrecheck>
__device__ inline void discover(unsigned long long *vp, uint32_t *bm0, uint32_t *bm1, uint32_t u,
                                unsigned long long bits, uint32_t *s_union, uint32_t *s_qn,
                                uint32_t *qnext) {
  bool first = false, push = false;
  if (kPrecheck && bits)
    bits &= ~__hip_atomic_load(&vp[2ull * u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (bits) {
    unsigned long long old = atomicOr(&vp[2ull * u], bits);
    unsigned long long fresh = bits & ~old;
    if (fresh) {
      first = (old == 0);
      unsigned long long pold = atomicOr(&vp[2ull * u + 1], fresh);
      push = (pold == 0);
    }
  }
  if (first) {
    uint32_t o0 = atomicOr(&bm0[u >> 5], 1u << (u & 31));
    if (o0 == 0) atomicOr(&bm1[u >> 10], 1u << ((u >> 5) & 31));
  }
  uint64_t fm = __ballot(first);
  if (fm && lane_id() == static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(fm)) - 1))
    atomicAdd(s_union, static_cast<uint32_t>(__popcll(fm)));
  wave_append(push, u, s_qn, qnext);
}

template <int BT, bool kPrecheck>
__global__ void __launch_bounds__(BT)
ic_fused_kernel(IcParams P) {
  constexpr int NW = BT / 64;
  __shared__ unsigned long long s_z[64];
  __shared__ uint32_t s_cnt[64];
  __shared__ uint32_t s_cofs[64];
  __shared__ uint32_t s_tmp[NW];
  __shared__ unsigned long long s_group;
  __shared__ unsigned long long s_base;
  __shared__ uint32_t s_fail;
  __shared__ uint32_t s_qn[2];   // light frontier queues
  __shared__ uint32_t s_hn[2];   // deferred-hub lists
  __shared__ uint32_t s_union;
  __shared__ Affine s_p2[64];      // P^(2^j): jumps between gap blocks
  __shared__ IcLogTables s_tab;
  __shared__ uint32_t s_ebeg[BT];
  __shared__ uint32_t s_edeg[BT];
  __shared__ uint32_t s_eT[BT];    // the row's largest threshold
  __shared__ uint32_t s_eflags[BT];
  __shared__ uint32_t s_eufull[BT];
  __shared__ uint32_t s_eulast[BT];
  __shared__ unsigned long long s_erq[BT];
  __shared__ uint32_t s_epre[BT];
  __shared__ unsigned long long s_eq[BT];
  __shared__ unsigned long long s_ea[BT];
  __shared__ unsigned long long s_eb[BT];
  __shared__ unsigned long long s_em[BT];
  __shared__ unsigned long long s_hot[BT / 64];
  __shared__ uint32_t s_col;       // the group's column, or kNoGroup

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = lane_id();
  const uint32_t wid = tid >> 6;
  const uint64_t slot = blockIdx.x;
  unsigned long long *vp = P.vp + slot * 2ull * P.n;
  uint32_t *bm0 = P.bm0 + slot * P.nw0;
  uint32_t *bm1 = P.bm1 + slot * P.nw1;
  // (selected with ?: rather than indexed arrays so they stay in registers)
  uint32_t *const q0 = P.queue + slot * 3ull * P.npad;
  uint32_t *const q1 = q0 + P.npad;
  // output staging after the BFS: sorted union + masks
  uint32_t *list = q0;
  unsigned long long *masks = reinterpret_cast<unsigned long long *>(q1);
  uint32_t *const h0 = P.hub + slot * 2ull * P.hub_cap;
  uint32_t *const h1 = h0 + P.hub_cap;
  // work counters (candidate edges, draws, expansions, live edges, bytes)
  unsigned long long st_edges = 0, st_coins = 0, st_exp = 0, st_live = 0, st_bytes = 0;
  if (tid == 0) {
    s_p2[0] = P.step;
    for (int j = 1; j < 64; ++j) s_p2[j] = aff_then(s_p2[j - 1], s_p2[j - 1]);
  }
  ic_load_tables(&s_tab);

  for (;;) {
    if (tid == 0) s_group = atomicAdd(&P.ctl[kCtlGroupCounter], 1ull);
    __syncthreads();
    const uint64_t g = s_group;
    if (g >= P.ngroups) break;
    const uint64_t s0 = g * P.colors;
    const uint32_t nb = static_cast<uint32_t>(
        (P.count - s0) < P.colors ? (P.count - s0) : P.colors);

    if (tid < 64) s_cnt[tid] = 0;
    if (tid == 0) {
      s_qn[0] = 0;
      s_qn[1] = 0;
      s_hn[0] = 0;
      s_hn[1] = 0;
      s_union = 0;
      s_fail = 0;
    }
    uint32_t root = 0;
    unsigned long long rbits = 0;
    if (tid < nb) {
      unsigned long long z = sample_z(P.seed, sample_id(P, s0 + tid));
      s_z[tid] = z;
      root = uniform_index(z, P.n);
      rbits = 1ull << tid;
    }
    __syncthreads();
    if (tid < 64)
      discover<kPrecheck>(vp, bm0, bm1, root, rbits, &s_union, &s_qn[0], q0);

    // ---------------------------------------------------------- BFS levels
    // Light levels expand the frontier queue; rows of at least heavy_deg
    // in-edges met there are parked in a hub list with their pending colours
    // still accumulating.  When the light frontier runs dry the hub list is
    // expanded as one level with every colour that reached each hub.  The
    // RRR set is the reverse closure over live edges, so the schedule does
    // not change the result, only how often a hub row is read.
    int lcur = 0, hfill = 0;
    for (;;) {
      __syncthreads();
      uint32_t qn = s_qn[lcur];
      const bool hub_phase = (qn == 0);
      if (hub_phase) qn = s_hn[hfill];
      __syncthreads();
      if (qn == 0) break;
      const int lnext = lcur ^ 1;
      if (tid == 0) {
        s_qn[lnext] = 0;
        if (hub_phase) s_hn[hfill ^ 1] = 0;
      }
      uint32_t *const hcur = hfill ? h1 : h0;
      const uint32_t *qsrc = hub_phase ? hcur : (lcur ? q1 : q0);
      uint32_t *qnext = lnext ? q1 : q0;
      uint32_t *hdefer = hcur;
      for (uint32_t cb = 0; cb < qn; cb += BT) {
        const uint32_t cl = min(static_cast<uint32_t>(BT), qn - cb);
        uint32_t runs = 0, v = 0;
        bool defer = false;
        if (tid < cl) {
          v = qsrc[cb + tid];
          const VertexRec r = P.vrec[v];
          const uint32_t deg = r.end - r.begin;
          unsigned long long M = 0;
          if (!hub_phase && deg >= P.heavy_deg) {
            defer = true;  // pending colours stay set: no re-push until expanded
          } else {
            M = atomicExch(&vp[2ull * v + 1], 0ull);
          }
          s_ebeg[tid] = r.begin;
          s_edeg[tid] = deg;
          s_eT[tid] = r.thr;
          s_eflags[tid] = r.flags;
          s_eufull[tid] = r.ufull;
          s_eulast[tid] = r.ulast;
          s_erq[tid] = r.rq;
          s_eq[tid] = r.q;
          s_ea[tid] = r.ja;
          s_eb[tid] = r.jb;
          s_em[tid] = M;
          // one work item per gap block of the row, for all its colours; a
          // row no edge of which can be live (T = 0) makes none
          runs = (M && r.thr) ? ((deg - 1) >> ((r.flags >> kBetaShift) & 31u)) + 1 : 0;
          if (M) {
            ++st_exp;
            st_bytes += 64;
          }
        }
        wave_append(defer, v, &s_hn[hfill], hdefer);
        uint32_t total;
        const uint32_t pre = block_exclusive_scan<BT>(runs, s_tmp, &total);
        if (tid < cl) s_epre[tid] = pre;
        __syncthreads();
        for (uint32_t rb = 0; rb < total; rb += BT) {
          const uint32_t rr = rb + tid;
          // this lane's item: block k of entry ent, for the colours in mm
          unsigned long long mm = 0;
          Affine E{1, 0};
          uint32_t e0 = 0, len = 0, B = 0, T = 0, uskip = 0;
          uint64_t q = 0, rq = 0;
          bool uniform = true;
          if (rr < total) {
            uint32_t lo = 0, hi = cl;
            while (lo < hi) {
              uint32_t mid = (lo + hi) >> 1;
              if (s_epre[mid] <= rr) lo = mid + 1;
              else hi = mid;
            }
            const uint32_t ent = lo - 1;
            const uint32_t k = rr - s_epre[ent];
            const uint32_t fl = s_eflags[ent];
            const uint32_t beta = (fl >> kBetaShift) & 31u;
            B = 1u << beta;
            e0 = s_ebeg[ent] + (k << beta);
            len = min(s_edeg[ent] - (k << beta), B);
            T = s_eT[ent];
            q = s_eq[ent];
            // a block without candidates costs one compare: the full or
            // the partial last block's skip threshold
            uskip = len == B ? s_eufull[ent] : s_eulast[ent];
            rq = s_erq[ent];
            uniform = (fl & kRowUniform) != 0;
            mm = s_em[ent];
            // block k's first element: P^(2 B k) after the row's first
            E = Affine{s_ea[ent], s_eb[ent]};
            for (uint32_t kk = k, bit = beta + 1; kk; kk >>= 1, ++bit)
              if (kk & 1) E = aff_then(E, s_p2[bit]);
          }
          // one live (edge, colour) per lane per step; every lane takes part
          // in each discovery
          uint32_t pos = len, c = 0;
          uint64_t x = 0;
          for (;;) {
            uint32_t src = 0;
            unsigned long long bits = 0;
            for (;;) {
              if (pos >= len) {
                if (mm == 0) break;
                c = static_cast<uint32_t>(__ffsll(mm) - 1);
                mm &= mm - 1;
                x = E.a * s_z[c] + E.b;
                pos = 0;
                ++st_coins;
                if ((x >> 32) + 1 <= uskip) pos = len;
                continue;
              }
              const uint64_t gap = ic_gap_fast(x, q, rq, &s_tab);
              if (gap >= len - pos) {
                pos = len;
                continue;
              }
              pos += static_cast<uint32_t>(gap);
              const uint32_t e = e0 + pos;
              const uint64_t a = P.step.a * x + P.step.b;   // acceptance element
              ++st_edges;
              if (!uniform) {
                ++st_coins;
                st_bytes += 4;
              }
              const bool live = uniform || ic_accept(a, T, P.val[e]);
              ++pos;
              x = P.step.a * a + P.step.b;                  // next gap element
              if (pos < len) ++st_coins;
              if (live) {
                src = P.ids[e];
                bits = 1ull << c;
                ++st_live;
                st_bytes += 4;
                break;
              }
            }
            if (__ballot(bits != 0) == 0) break;
            discover<kPrecheck>(vp, bm0, bm1, src, bits, &s_union, &s_qn[lnext], qnext);
          }
        }
        __syncthreads();
      }
      lcur = lnext;
      if (hub_phase) hfill ^= 1;
    }

    // ---------------------------------------------- sorted union + colours
    const uint32_t U = s_union;
    uint32_t running = 0;
    // The bitmaps and visited masks were written by atomics, which complete
    // in L2; read and clear them with atomics too so no stale L1 line of an
    // earlier group is ever observed.
    for (uint32_t t0 = 0; t0 < P.nw1; t0 += BT) {
      const uint32_t w1i = t0 + tid;
      const uint32_t w1 =
          w1i < P.nw1 ? __hip_atomic_load(&bm1[w1i], __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT)
                      : 0u;
      uint32_t cnt = 0;
      for (uint32_t b = w1; b; b &= b - 1)
        cnt += __popc(__hip_atomic_load(&bm0[w1i * 32 + (__ffs(b) - 1)],
                                        __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
      uint32_t tot;
      uint32_t p = running + block_exclusive_scan<BT>(cnt, s_tmp, &tot);
      for (uint32_t b = w1; b; b &= b - 1) {
        const uint32_t w0i = w1i * 32 + (__ffs(b) - 1);
        const uint32_t w0 = atomicExch(&bm0[w0i], 0u);
        for (uint32_t bb = w0; bb; bb &= bb - 1) {
          const uint32_t v = w0i * 32 + (__ffs(bb) - 1);
          list[p] = v;
          masks[p] = atomicExch(&vp[2ull * v], 0ull);
          ++p;
        }
      }
      if (w1) atomicExch(&bm1[w1i], 0u);
      running += tot;
    }
    __syncthreads();
    if (P.colptr) {
      // per-group choice of representation: a column costs 8 bytes per hot
      // vertex, lists 4 per member; the group's hot members decide
      unsigned long long hot = 0;
      for (uint32_t i = tid; i < U; i += BT)
        if (P.hot_index[list[i]] != kNotHot) hot += __popcll(masks[i]);
      unsigned long long tot_hot;
      (void)block_exclusive_scan_u64<BT>(hot, s_hot, &tot_hot);
      if (tid == 0) {
        s_col = kNoGroup;
        if (tot_hot > 2ull * P.hot_n)
          s_col = P.group_base +
                  static_cast<uint32_t>(atomicAdd(&P.ctl[kCtlGroupCounter2], 1ull));
      }
      __syncthreads();
      const uint32_t colg = s_col;
      if (colg != kNoGroup) {
        unsigned long long *col = P.colptr[colg];
        for (uint32_t h = tid; h < P.hot_n; h += BT) col[h] = 0ull;
        __syncthreads();
        // hot members go to the column, which the lists then leave out
        for (uint32_t i = tid; i < U; i += BT) {
          const uint32_t h = P.hot_index[list[i]];
          if (h != kNotHot) {
            col[h] = masks[i];
            masks[i] = 0ull;
          }
        }
        st_bytes += 8ull * ((P.hot_n + BT - 1 - tid) / BT);
        // the pool's element count includes the members held in the column
        if (tid == 0) atomicAdd(&P.ctl[kCtlElems], tot_hot);
        if (tid < 64) {
          uint32_t sid = kNoGroup;
          if (tid < nb) {
            sid = static_cast<uint32_t>(sample_id(P, s0 + tid));
            P.gref[sid] = (colg << 6) | tid;
          }
          P.gsample[64ull * colg + tid] = sid;
        }
      }
      __syncthreads();
    }
    for (uint32_t i = tid; i < U; i += BT)
      for (unsigned long long mk = masks[i]; mk; mk &= mk - 1)
        atomicAdd(&s_cnt[__ffsll(mk) - 1], 1u);
    __syncthreads();
    if (tid == 0) {
      uint32_t acc = 0;
      for (uint32_t c = 0; c < nb; ++c) {
        s_cofs[c] = acc;
        acc += s_cnt[c];
      }
      unsigned long long base = atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(acc));
      s_base = base;
      if (base + acc > P.cap_data) {
        s_fail = 1;
        atomicOr(&P.ctl[kCtlOverflow], 1ull);
      }
    }
    __syncthreads();
    if (!s_fail) {
      const unsigned long long base = s_base;
      for (uint32_t c = wid; c < nb; c += NW) {
        uint32_t run = 0;
        const unsigned long long dst0 = base + s_cofs[c];
        for (uint32_t i0 = 0; i0 < U; i0 += 64) {
          const uint32_t i = i0 + lane;
          const bool bit = i < U && ((masks[i] >> c) & 1ull);
          const uint64_t bm = __ballot(bit);
          if (bit) P.data[dst0 + run + __popcll(bm & lanes_below())] = list[i];
          run += __popcll(bm);
        }
      }
      if (tid < nb) {
        const uint64_t s = sample_id(P, s0 + tid);
        P.set_off[s] = base + s_cofs[tid];
        P.set_len[s] = s_cnt[tid];
        st_bytes += 4ull * s_cnt[tid];
      }
    } else if (tid < nb) {
      P.set_len[sample_id(P, s0 + tid)] = 0xFFFFFFFFu;
    }
    __syncthreads();
  }
  report_stats(P.ctl, st_bytes, st_edges, st_coins, st_exp, st_live);
}

template <int BT>
void launch_one(const IcParams &P, uint32_t slots, hipStream_t stream) {
  if (P.precheck)
    hipLaunchKernelGGL((ic_fused_kernel<BT, true>), dim3(slots), dim3(BT), 0, stream, P);
  else
    hipLaunchKernelGGL((ic_fused_kernel<BT, false>), dim3(slots), dim3(BT), 0, stream, P);
}

__global__ void reset_counter_kernel(unsigned long long *ctl) {
  ctl[kCtlGroupCounter] = 0;
}

// Workgroups of the persistent kernel that fit on one CU at once.
template <int BT>
int occupancy_of() {
  return blocks_per_cu(reinterpret_cast<const void *>(&ic_fused_kernel<BT, false>), BT);
}

int resident_blocks(uint32_t block) {
  switch (block) {
    case 256: return occupancy_of<256>();
    case 512: return occupancy_of<512>();
    default: return occupancy_of<1024>();
  }
}

// Upper bound of the number of vertices with at least `deg` in-edges.
uint64_t hub_capacity(const rip_graph *g, uint32_t deg) {
  if (deg <= 1) return g->n;
  uint32_t b = 31 - __builtin_clz(deg);  // 2^b <= deg
  return std::max<uint64_t>(1, g->deg_ge_pow2[b]);
}

size_t slot_bytes_for(const rip_graph *g, uint32_t heavy_deg) {
  const uint32_t n = g->n;
  const uint64_t npad = (static_cast<uint64_t>(n) + 63) & ~63ull;
  const uint64_t nw0 = (n + 31) / 32, nw1 = (n + 1023) / 1024;
  return 16ull * n + 4ull * nw0 + 4ull * nw1 + 12ull * npad +
         8ull * hub_capacity(g, heavy_deg);
}

}  // namespace

IcLaunch ic_tune(const rip_pool *p, const rip_tune *tune) {
  // defaults from a sweep on R-MAT scale 22 weighted cascade: 8-edge runs and
  // 256-thread groups (more independent groups per CU); hub deferral off,
  // because colours of one group rarely meet at the same hub there
  IcLaunch c{64, 256, 0, 0, 0xFFFFFFFFu, RIP_SAMPLER_AUTO, true, 0, true, true};
  if (tune) {
    if (tune->colors) {
      c.colors = tune->colors;
      c.colors_auto = false;
    }
    if (tune->block) {
      c.block = tune->block;
      c.block_auto = false;
    }
    if (tune->slots) {
      c.slots = tune->slots;
      c.slots_auto = false;
    }
    c.queue_cap = tune->queue_cap;
    if (tune->heavy_degree) c.heavy_degree = tune->heavy_degree;
    c.sampler = tune->sampler;
    c.out_chunk = tune->out_chunk;
  }
  RIP_REQUIRE(c.colors >= 1 && c.colors <= 64, -1, "tune: colors must be in [1, 64]");
  RIP_REQUIRE(c.sampler <= RIP_SAMPLER_LANE, -1, "tune: unknown sampler kind");
  (void)ic_lane_cap(tune);  // validates queue_cap
  RIP_REQUIRE(c.block == 256 || c.block == 512 || c.block == 1024, -1,
              "tune: block must be 256, 512 or 1024");
  if (c.slots == 0) {
    size_t free_b = 0, total_b = 0;
    RIP_HIP(hipMemGetInfo(&free_b, &total_b));
    const size_t per = slot_bytes_for(p->g, c.heavy_degree);
    // keep the sampler scratch within an eighth of the memory it may use
    // (free plus what it already holds), 24 GiB at most
    size_t held = 0;
    if (p->g->ic.slots && p->g->ic.n == p->g->n) held = p->g->ic.slot_bytes * p->g->ic.slots;
    const size_t budget = std::min((free_b + held) / 8, static_cast<size_t>(24) << 30);
    uint64_t by_mem = std::max<uint64_t>(1, budget / per);
    uint64_t by_cu = static_cast<uint64_t>(cu_count()) * resident_blocks(c.block);
    c.slots = static_cast<uint32_t>(std::min(by_mem, by_cu));
  }
  return c;
}

void ic_ensure_slots(rip_pool *p, uint32_t slots, uint32_t heavy_deg) {
  IcSlots &s = p->g->ic;
  const uint64_t hub_cap = hub_capacity(p->g, heavy_deg);
  if (s.slots >= slots && s.n == p->g->n && s.hub_cap >= hub_cap) return;
  dev_free(s.d_vp);
  dev_free(s.d_bm0);
  dev_free(s.d_bm1);
  dev_free(s.d_queue);
  dev_free(s.d_hub);
  s.slots = 0;
  const uint32_t n = p->g->n;
  const uint64_t npad = (static_cast<uint64_t>(n) + 63) & ~63ull;
  const uint64_t nw0 = (n + 31) / 32, nw1 = (n + 1023) / 1024;
  s.d_vp = dev_alloc<unsigned long long>(2ull * n * slots);
  s.d_bm0 = dev_alloc<uint32_t>(nw0 * slots);
  s.d_bm1 = dev_alloc<uint32_t>(nw1 * slots);
  s.d_queue = dev_alloc<uint32_t>(3ull * npad * slots);
  s.d_hub = dev_alloc<uint32_t>(2ull * hub_cap * slots);
  RIP_HIP(hipMemsetAsync(s.d_vp, 0, sizeof(unsigned long long) * 2ull * n * slots, p->stream));
  RIP_HIP(hipMemsetAsync(s.d_bm0, 0, sizeof(uint32_t) * nw0 * slots, p->stream));
  RIP_HIP(hipMemsetAsync(s.d_bm1, 0, sizeof(uint32_t) * nw1 * slots, p->stream));
  s.slots = slots;
  s.n = n;
  s.hub_cap = hub_cap;
  s.slot_bytes = slot_bytes_for(p->g, heavy_deg);
}

uint32_t ic_launch(rip_pool *p, uint64_t first, uint64_t count,
                   const IcLaunch &cfg, const uint32_t *ids, bool record_events,
                   bool columns) {
  // no more resident groups than there are groups; scratch only ever grows
  const uint64_t groups = (count + cfg.colors - 1) / cfg.colors;
  ic_ensure_slots(p, static_cast<uint32_t>(std::min<uint64_t>(cfg.slots, groups)),
                  cfg.heavy_degree);
  const rip_graph *g = p->g;
  IcParams P;
  P.vrec = g->d_vrec;
  P.ids = g->d_ids;
  P.val = g->d_val;
  P.n = g->n;
  P.step = g->step;
  P.seed = p->seed;
  P.first = first;
  P.count = count;
  P.colors = cfg.colors;
  P.ngroups = (count + cfg.colors - 1) / cfg.colors;
  P.vp = p->g->ic.d_vp;
  P.bm0 = p->g->ic.d_bm0;
  P.bm1 = p->g->ic.d_bm1;
  P.queue = p->g->ic.d_queue;
  P.nw0 = (g->n + 31) / 32;
  P.nw1 = (g->n + 1023) / 1024;
  P.npad = (static_cast<uint64_t>(g->n) + 63) & ~63ull;
  P.set_off = p->d_off;
  P.set_len = p->d_len;
  P.data = p->d_data;
  P.cap_data = p->cap_data;
  P.ctl = p->d_ctl;
  P.heavy_deg = cfg.heavy_degree;
  P.hub = p->g->ic.d_hub;
  P.hub_cap = p->g->ic.hub_cap;
  P.sample_ids = ids;
  P.colptr = nullptr;
  P.gsample = nullptr;
  P.gref = nullptr;
  P.hot_index = nullptr;
  P.hot_n = 0;
  P.group_base = 0;
  if (columns) {
    // room for a column per group of the launch; groups whose sets are too
    // small for one keep lists and leave theirs unused
    pool_reserve_groups(p, p->groups + static_cast<uint32_t>(P.ngroups));
    P.colptr = p->d_colptr;
    P.gsample = p->d_gsample;
    P.gref = p->d_gref;
    P.hot_index = g->d_hot_index;
    P.hot_n = g->hot_n;
    P.group_base = p->groups;
  }
  // the large-set fallback walks giant components, where most edges reach
  // vertices every colour has visited; a whole launch of mostly small sets
  // (weighted cascade) gains nothing from the extra read
  P.precheck = ids != nullptr;
  const uint32_t slots =
      static_cast<uint32_t>(std::min<uint64_t>(cfg.slots, P.ngroups));
  hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  if (record_events) RIP_HIP(hipEventRecord(p->ev0, p->stream));
  switch (cfg.block) {
    case 256: launch_one<256>(P, slots, p->stream); break;
    case 1024: launch_one<1024>(P, slots, p->stream); break;
    default: launch_one<512>(P, slots, p->stream); break;
  }
  RIP_HIP(hipGetLastError());
  if (record_events) RIP_HIP(hipEventRecord(p->ev1, p->stream));
  return 2;
}

uint32_t ic_sample(rip_pool *p, uint64_t first, uint64_t count, const IcLaunch &cfg,
                   const rip_tune *tune) {
  if (cfg.sampler == RIP_SAMPLER_COLOUR) return ic_launch(p, first, count, cfg);
  uint32_t launches = 0;
  uint64_t wave_count = count;
  const uint32_t *wave_ids = nullptr;
  const bool lanes = cfg.sampler == RIP_SAMPLER_AUTO || cfg.sampler == RIP_SAMPLER_LANE;
  if (lanes) {
    // one sample per lane first; the few that outgrow a lane's queue go to
    // the grid sampler (AUTO) or to one BFS per wave (LANE)
    wave_count = ic_lane_launch(p, first, count, tune);
    launches += 3;
    if (wave_count == 0) return launches;
    wave_ids = p->g->wave.d_over;
    if (cfg.sampler == RIP_SAMPLER_AUTO) {
      const uint32_t frontier_cap = tune ? tune->frontier_cap : 0u;
      const uint32_t colors = tune && tune->colors ? tune->colors : 64u;
      if (tune && tune->wave_stage) {
        // one BFS per wave, 64 frontier vertices at a time, for the samples
        // up to the wave list; the grid sampler takes the larger ones
        const uint64_t big = ic_wave_launch(p, first, wave_count, wave_ids, cfg.out_chunk, false,
                                            true, true);
        launches += 2;
        if (big == 0) return launches;
        return launches + ic_grid_launch(p, first, big, p->g->wave.d_big, true, frontier_cap,
                                         colors);
      }
      return launches + ic_grid_launch(p, first, wave_count, wave_ids, true, frontier_cap, colors);
    }
  }
  const uint64_t big = ic_wave_launch(p, first, wave_count, wave_ids, cfg.out_chunk, !lanes);
  launches += 2;
  if (big == 0) return launches;
  // sets that outgrew a wave's list: by default colour groups over the listed
  // sample indices (samples that reach a giant component share most of their
  // rows, which the groups read once for 64 of them); WAVE and LANE finish
  // them with the bitmap-visited BFS per wave instead
  if (cfg.sampler == RIP_SAMPLER_WAVE || cfg.sampler == RIP_SAMPLER_LANE) {
    ic_big_launch(p, first, big);
  } else {
    // the groups are latency-bound, one level at a time: 512-thread groups
    // (on R-MAT scale 18 with U[0,0.1] weights, 100,000 samples: 0.119 s at
    // 64 colours against 0.150 s with 256 threads and 0.123 s with 1024) and
    // about one group per two resident slots, i.e. colours near 2 big / slots
    // (rounded to a power of two), measured best on R-MAT scales 16 and 18
    // with U[0,0.1] weights from 1,000 to 17,000 large samples
    IcLaunch fb = cfg;
    if (cfg.block_auto) {
      fb.block = 512;
      if (cfg.slots_auto) {
        fb.slots = std::min<uint32_t>(
            cfg.slots, static_cast<uint32_t>(cu_count()) * resident_blocks(fb.block));
      }
    }
    if (cfg.colors_auto) {
      const double want = 2.0 * static_cast<double>(big) / std::max<uint32_t>(fb.slots, 1);
      uint32_t c = 1;
      while (c < 64 && static_cast<double>(c) * 1.41421356 < want) c <<= 1;
      fb.colors = c;
    }
    ic_launch(p, first, big, fb, p->g->wave.d_big, false, true);
  }
  return launches + 2;
}

}  // namespace rip
