// Wave sampler: one IC reverse BFS per wave, many waves per CU, no
// workgroup barriers.
//
// Semantics are those of AddRRRSet (include/ripples/add_rrrset.h:105-157)
// under the per-sample stream contract of common.h and ic_gap.h: the live
// in-edges of a row depend only on (seed, sample, row), so the set is the
// reverse closure of the root over live edges and equals the other
// samplers' and the oracle's bit for bit.
//
// Layout per resident wave (HBM):
//   table[TCAP]  visited set: open addressing, 64-bit entries
//                (epoch << 32 | vertex); an entry counts only if its epoch
//                is the wave's current one, so the table is never cleared
//   list[CAP]    BFS queue and output list (FIFO in discovery order)
// Expansion of a vertex: lane l draws the candidates of the row's gap
// blocks l, l + 64, ... (one compare for a full block without candidates);
// only live edges read their in-neighbour and touch the table (in the
// vertex-parallel form, an LDS table first).  The finished list is sorted in
// LDS (bitonic, 1024 entries per wave, 2048 in the vertex-parallel form) or,
// if longer, in its HBM buffer.  Samples whose sets outgrow CAP are listed for
// the large-set kernels.
#include <algorithm>

#include "device_util.h"
#include "ic_common.h"
#include "ic_gap.h"
#include "internal.h"

namespace rip {
namespace {

constexpr uint32_t kSortCap = 1024;
// Vertex-parallel form: the first kWaveLdsLimit vertices of a sample are
// marked in a per-wave LDS table of kWaveLdsSlots 32-bit slots (vertex + 1,
// 0 empty), cleared per sample; only later ones go to the wave's HBM table.
// Most sets of this stage end below the limit, so most of its visited checks
// never leave the CU.  The slots double as the wave's sort buffer once the
// BFS is done.
constexpr uint32_t kWaveLdsSlots = 2048;
constexpr uint32_t kWaveLdsLimit = kWaveLdsSlots / 2;

struct WaveParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  Affine step;
  uint64_t seed;
  uint64_t first;
  uint64_t count;
  unsigned long long *table;   // waves x tcap
  uint32_t *list;              // waves x cap
  uint32_t *epoch;             // per resident wave
  uint32_t cap, tcap;          // tcap a power of two >= 2 cap
  uint32_t *big_ids;           // samples left for the colour sampler
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
  uint32_t chunk;              // pool elements a wave reserves at a time
  const uint32_t *sample_ids;  // optional: local index -> sample index (minus first)
};

// Insert v into the wave's table; true if it was not there yet.
__device__ inline bool table_insert(unsigned long long *table, uint32_t mask,
                                    unsigned long long tag, uint32_t v) {
  const unsigned long long want = tag | v;
  uint32_t h = hash_vertex(v) & mask;
  for (;;) {
    unsigned long long cur = __hip_atomic_load(&table[h], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    for (;;) {
      if ((cur & 0xFFFFFFFF00000000ull) != tag) {  // stale or empty slot
        const unsigned long long seen = atomicCAS(&table[h], cur, want);
        if (seen == cur) return true;
        cur = seen;
        continue;
      }
      if (static_cast<uint32_t>(cur) == v) return false;
      break;  // occupied by another vertex of this sample
    }
    h = (h + 1) & mask;
  }
}

// Insert v into a wave's LDS table (power-of-two slots, fewer entries than
// slots); true if it was not there yet.
__device__ inline bool lds_insert(uint32_t *t, uint32_t mask, uint32_t v) {
  const uint32_t want = v + 1u;
  uint32_t h = hash_vertex(v) & mask;
  for (;;) {
    uint32_t cur = __hip_atomic_load(&t[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (cur == 0) {
      cur = atomicCAS(&t[h], 0u, want);
      if (cur == 0) return true;
    }
    if (cur == want) return false;
    h = (h + 1) & mask;
  }
}

// Whether v is in a wave's LDS table.
__device__ inline bool lds_find(const uint32_t *t, uint32_t mask, uint32_t v) {
  const uint32_t want = v + 1u;
  uint32_t h = hash_vertex(v) & mask;
  for (;;) {
    const uint32_t cur = __hip_atomic_load(&t[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
    if (cur == want) return true;
    if (cur == 0) return false;
    h = (h + 1) & mask;
  }
}

// Expands row v of the sample whose element 0 is z (ic_gap.h): lane l draws
// the candidates of gap blocks l, l + 64, ... (block k's first element is
// P^(2 B k) applied to the row's first, P^(1 + 2 begin) z); the in-neighbours
// of live edges for which visit(src) reports a first visit are appended to
// list[tail ...].  Every lane of the wave must call it.  Returns false,
// leaving the row unfinished, when the live edges of a step (an upper bound
// on the new vertices) would take the list past cap.
template <typename Visit>
__device__ inline bool expand_row(const VertexRec *vrec, const uint32_t *ids,
                                  const uint32_t *val, uint32_t v, uint64_t z,
                                  const Affine *p2, const IcLogTables *tab, uint32_t *list,
                                  uint32_t &tail, uint32_t cap,
                                  unsigned long long &st_cand, unsigned long long &st_draws,
                                  unsigned long long &st_bytes, unsigned long long &st_live,
                                  Visit &&visit) {
  const uint32_t lane = lane_id();
  // the row is the same for every lane: keep it in scalar registers
  const VertexRec rv = vrec[v];
  const uint32_t T = uniform_u32(rv.thr);
  if (T == 0) return true;                         // no edge of the row can be live
  const uint32_t rbegin = uniform_u32(rv.begin);
  const uint32_t deg = uniform_u32(rv.end) - rbegin;
  const uint32_t flags = uniform_u32(rv.flags);
  const bool uniform = (flags & kRowUniform) != 0;
  const uint32_t beta = (flags >> kBetaShift) & 31u;
  const uint32_t B = 1u << beta;
  const uint64_t q = uniform_u64(rv.q);
  const uint64_t ufull = uniform_u32(rv.ufull);
  const uint64_t ulast = uniform_u32(rv.ulast);
  const uint64_t rq = uniform_u64(rv.rq);
  const uint32_t nblk = ((deg - 1) >> beta) + 1;
  const Affine P = p2[0];
  const uint64_t x0 = uniform_u64(rv.ja) * z + uniform_u64(rv.jb);  // row's first element
  // this lane's first block: P^(2B lane) from the binary digits of lane
  Affine jl{1, 0};
  for (uint32_t k = lane, bit = beta + 1; k; k >>= 1, ++bit)
    if (k & 1) jl = aff_then(jl, p2[bit]);
  const Affine j64 = p2[beta + 7];                 // 64 blocks further
  uint64_t xb = aff_apply(jl, x0);
  for (uint32_t k0 = 0; k0 < nblk; k0 += 64, xb = aff_apply(j64, xb)) {
    const uint32_t k = k0 + lane;
    uint32_t len = 0;
    if (k < nblk) {
      len = min(deg - (k << beta), B);
      ++st_draws;
      // a block without candidates costs one compare
      if ((xb >> 32) + 1 <= (len == B ? ufull : ulast)) len = 0;
    }
    uint32_t pos = 0;
    uint64_t x = xb;
    bool first = true;
    for (;;) {
      bool live = false;
      uint32_t e = 0;
      if (pos < len) {
        if (!first) ++st_draws;
        first = false;
        const uint64_t gap = ic_gap_fast(x, q, rq, tab);
        if (gap >= len - pos) {
          pos = len;
        } else {
          pos += static_cast<uint32_t>(gap);
          e = rbegin + (k << beta) + pos;
          const uint64_t a = aff_apply(P, x);     // acceptance element
          ++st_cand;
          if (!uniform) {
            ++st_draws;
            st_bytes += 4;
          }
          live = uniform || ic_accept(a, T, val[e]);
          if (live) {
            ++st_live;
            st_bytes += 4;
          }
          ++pos;
          x = aff_apply(P, a);                     // next gap element
        }
      }
      const uint64_t lm = __ballot(live);
      if (lm) {
        if (tail + __popcll(lm) > cap) return false;
        // the in-neighbour id is read only for live edges
        const uint32_t src = live ? ids[e] : 0u;
        const bool fresh = live && visit(src);
        const uint64_t fm = __ballot(fresh);
        if (fresh) list[tail + __popcll(fm & lanes_below())] = src;
        tail += __popcll(fm);
      }
      if (__ballot(pos < len) == 0) break;
    }
  }
  return true;
}

// Expands up to 64 queued vertices of the sample whose element 0 is z at
// once, lane l taking the vertex whose record is rv (when `have`): the lane
// walks its row's gap blocks one draw per step, as a lane of the lane
// sampler does; the live edges of a step go through visit() and the first
// visits are appended to list[tail ...].  Every lane of the wave must call
// it.  Returns false when the live edges of a step could take the list past
// cap.  Rows of a few edges (weighted cascade: one candidate per row) keep
// 64 lanes busy here, where expand_row would give each row a whole wave.
template <typename Visit>
__device__ inline bool expand_rows_lanes(const uint32_t *ids, const uint32_t *val, bool have,
                                         const VertexRec &rv, uint64_t z,
                                         const Affine *p2, const IcLogTables *tab,
                                         uint32_t *list, uint32_t &tail, uint32_t cap,
                                         unsigned long long &st_cand,
                                         unsigned long long &st_draws,
                                         unsigned long long &st_bytes,
                                         unsigned long long &st_live, Visit &&visit) {
  uint32_t T = 0, rbegin = 0, deg = 0, beta = 0, nblk = 0, k = 0, len = 0, pos = 0;
  bool uniform = true;
  uint64_t q = 0, rq = 0, ufull = 0, ulast = 0, xb = 0, x = 0;
  if (have) {
    T = rv.thr;
    have = T != 0;                               // no edge of the row can be live
    if (have) {
      rbegin = rv.begin;
      deg = rv.end - rv.begin;
      uniform = (rv.flags & kRowUniform) != 0;
      beta = (rv.flags >> kBetaShift) & 31u;
      q = rv.q;
      rq = rv.rq;
      ufull = rv.ufull;
      ulast = rv.ulast;
      nblk = ((deg - 1) >> beta) + 1;
      xb = rv.ja * z + rv.jb;                    // the row's first element
    }
  }
  // block k of the row: its length, and no candidate when one compare says so
  auto open_block = [&]() {
    len = min(deg - (k << beta), 1u << beta);
    x = xb;
    pos = 0;
    ++st_draws;
    if ((xb >> 32) + 1 <= (len == (1u << beta) ? ufull : ulast)) len = 0;
  };
  if (have) open_block();
  const Affine P = p2[0];
  for (;;) {
    bool live = false;
    uint32_t e = 0;
    if (have) {
      if (pos >= len) {
        if (++k < nblk) {
          const Affine jb = p2[beta + 1];          // 2B elements: the next block
          xb = aff_apply(jb, xb);
          open_block();
        } else {
          have = false;
        }
      } else {
        const uint64_t gap = ic_gap_fast(x, q, rq, tab);
        if (gap >= len - pos) {
          pos = len;
        } else {
          pos += static_cast<uint32_t>(gap);
          e = rbegin + (k << beta) + pos;
          const uint64_t a = aff_apply(P, x);      // acceptance element
          ++st_cand;
          if (!uniform) {
            ++st_draws;
            st_bytes += 4;
          }
          live = uniform || ic_accept(a, T, val[e]);
          if (live) {
            ++st_live;
            st_bytes += 4;
          }
          ++pos;
          x = aff_apply(P, a);                     // next gap element
          if (pos < len) ++st_draws;
        }
      }
    }
    const uint64_t lm = __ballot(live);
    if (lm) {
      if (tail + __popcll(lm) > cap) return false;
      const uint32_t src = live ? ids[e] : 0u;
      const bool fresh = live && visit(src);
      const uint64_t fm = __ballot(fresh);
      if (fresh) list[tail + __popcll(fm & lanes_below())] = src;
      tail += __popcll(fm);
    }
    if (__ballot(have) == 0) break;
  }
  return true;
}

// VP: expand 64 queued vertices at a time (expand_rows_lanes) instead of
// one row per step over its gap blocks (expand_row, the plain form), with
// the first vertices of each sample marked in LDS.
template <int BT, bool VP>
__global__ void __launch_bounds__(BT) ic_wave_kernel(WaveParams P) {
  constexpr int NW = BT / 64;
  constexpr uint32_t kBuf = VP && kWaveLdsSlots > kSortCap ? kWaveLdsSlots : kSortCap;
  __shared__ uint32_t s_sort[NW][kBuf];
  const uint32_t lane = lane_id();
  const uint32_t wid = threadIdx.x >> 6;
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * NW + wid;
  unsigned long long *table = P.table + wave * P.tcap;
  uint32_t *list = P.list + wave * P.cap;
  const uint32_t tmask = P.tcap - 1;
  uint32_t epoch = P.epoch[wave];
  // the wave's reserved output [out_cur, out_end) (wave-uniform); the rest
  // of a chunk that cannot take the next set stays a gap in the pool's data
  unsigned long long out_cur = 0, out_end = 0;
  __shared__ Affine s_p2[64];                 // P^(2^j): jumps between gap blocks
  __shared__ IcLogTables s_tab;
  if (threadIdx.x == 0) fill_step_powers(P.step, s_p2);
  ic_load_tables(&s_tab);
  __syncthreads();
  unsigned long long st_cand = 0, st_draws = 0, st_exp = 0, st_bytes = 0, st_live = 0;

  for (;;) {
    unsigned long long idx = 0;
    if (lane == 0) idx = atomicAdd(&P.ctl[kCtlGroupCounter], 1ull);
    idx = uniform_u64(__shfl(idx, 0, 64));
    if (idx >= P.count) break;
    const uint32_t local = P.sample_ids ? P.sample_ids[idx] : static_cast<uint32_t>(idx);
    const uint64_t s = P.first + local;
    ++epoch;
    const unsigned long long tag = static_cast<unsigned long long>(epoch) << 32;
    const uint64_t z = sample_z(P.seed, s);
    const uint32_t root = uniform_index(z, P.n);
    uint32_t *ltab = s_sort[wid];
    if (VP) {
      for (uint32_t i = 4 * lane; i < kWaveLdsSlots; i += 256)
        *reinterpret_cast<uint4 *>(&ltab[i]) = make_uint4(0, 0, 0, 0);
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    if (lane == 0) {
      if (VP)
        lds_insert(ltab, kWaveLdsSlots - 1, root);
      else
        table_insert(table, tmask, tag, root);
      list[0] = root;
    }
    uint32_t head = 0, tail = 1;
    bool big = false;
    // vertex-parallel rounds: the records of the next round's first npre
    // vertices, loaded with the current round's where the queue held them
    uint32_t npre = 0;
    VertexRec prec{};
    // a visit call adds at most 64 vertices: while the sample stays within
    // the LDS limit they go to the LDS table, then to the HBM table (after a
    // look in the LDS one)
    auto visit_vp = [&](uint32_t src) {
      if (tail + 64u <= kWaveLdsLimit) return lds_insert(ltab, kWaveLdsSlots - 1, src);
      if (lds_find(ltab, kWaveLdsSlots - 1, src)) return false;
      return table_insert(table, tmask, tag, src);
    };
    while (head < tail && !big) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      if (VP) {
        const uint32_t take = min(tail - head, 64u);
        const bool have = lane < take;
        const uint32_t nhead = head + take;
        // the next round's vertices already queued: their records load with
        // this round's, off the next round's chain of dependent loads
        const uint32_t nnext = min(tail - nhead, 64u);
        const bool now = have && lane >= npre;
        const bool next = lane < nnext;
        const uint32_t v = now ? list[head + lane] : 0u;
        const uint32_t vn = next ? list[nhead + lane] : 0u;
        VertexRec rec = prec;
        if (now) rec = P.vrec[v];
        VertexRec recn{};
        if (next) recn = P.vrec[vn];
        head = nhead;
        st_exp += take;
        if (!expand_rows_lanes(P.ids, P.val, have, rec, z, s_p2, &s_tab, list, tail, P.cap,
                               st_cand, st_draws, st_bytes, st_live, visit_vp))
          big = true;
        prec = recn;
        npre = nnext;
      } else {
        const uint32_t v = uniform_u32(list[head++]);
        ++st_exp;
        if (!expand_row(P.vrec, P.ids, P.val, v, z, s_p2, &s_tab, list, tail, P.cap, st_cand,
                        st_draws, st_bytes, st_live,
                        [&](uint32_t src) { return table_insert(table, tmask, tag, src); }))
          big = true;
      }
    }
    if (big) {
      if (lane == 0) {
        const unsigned long long b = atomicAdd(&P.ctl[kCtlLongCount], 1ull);
        P.big_ids[b] = local;
      }
      continue;
    }
    // sort the set and copy it to the pool
    const uint32_t len = tail;
    uint32_t cap2 = 64;
    while (cap2 < len) cap2 <<= 1;
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const uint32_t *sorted;
    if (cap2 <= kBuf) {
      uint32_t *a = s_sort[wid];
      for (uint32_t i = lane; i < len; i += 64) a[i] = list[i];
      wave_bitonic(a, len, cap2);
      sorted = a;
    } else {
      // the HBM list has room for cap entries; cap is a power of two
      wave_bitonic(list, len, cap2);
      sorted = list;
    }
    if (out_end - out_cur < len) {
      const unsigned long long want = len > P.chunk ? len : P.chunk;
      unsigned long long nb = 0;
      if (lane == 0) nb = atomicAdd(&P.ctl[kCtlDataBump], want);
      nb = uniform_u64(__shfl(nb, 0, 64));
      out_cur = nb;
      out_end = nb + want;
    }
    const unsigned long long base = out_cur;
    out_cur += len;
    if (base + len > P.cap_data) {
      if (lane == 0) {
        atomicOr(&P.ctl[kCtlOverflow], 1ull);
        P.set_len[s] = kLenFail;
      }
      continue;
    }
    for (uint32_t i = lane; i < len; i += 64) P.data[base + i] = sorted[i];
    if (lane == 0) {
      P.set_off[s] = base;
      P.set_len[s] = len;
      st_bytes += 4ull * len;
    }
  }
  if (lane == 0) P.epoch[wave] = epoch;
  if (lane == 0) st_bytes += kIcRecBytes * st_exp;
  report_stats(P.ctl, st_bytes, st_cand, st_draws, lane == 0 ? st_exp : 0ull, st_live);
}

// ------------------------------------------------------------ large sets
// Samples whose sets outgrow the wave list are finished here, again one
// reverse BFS per wave, with a visited bitmap of n bits instead of the hash
// table and a list of up to n entries.  A summary bitmap (one bit per bitmap
// word) is set with the first bit of each word; the finished set is written
// in vertex order by walking the summary, then the marked words, and both
// bitmaps are cleared on the way, so they are all zero between samples.
struct BigParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  Affine step;
  uint64_t seed;
  uint64_t first;
  const uint32_t *big_ids;  // local sample indices
  uint64_t count;
  uint32_t *bits;           // waves x words
  uint32_t *sum;            // waves x swords
  uint32_t *list;           // waves x n
  uint32_t words, swords;
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
};

template <int BT>
__global__ void __launch_bounds__(BT) ic_big_kernel(BigParams P) {
  constexpr int NW = BT / 64;
  const uint32_t lane = lane_id();
  const uint64_t wave = static_cast<uint64_t>(blockIdx.x) * NW + (threadIdx.x >> 6);
  uint32_t *bits = P.bits + wave * P.words;
  uint32_t *sum = P.sum + wave * P.swords;
  uint32_t *list = P.list + wave * P.n;
  __shared__ Affine s_p2[64];
  __shared__ IcLogTables s_tab;
  if (threadIdx.x == 0) fill_step_powers(P.step, s_p2);
  ic_load_tables(&s_tab);
  __syncthreads();
  unsigned long long st_cand = 0, st_draws = 0, st_exp = 0, st_bytes = 0, st_live = 0;
  auto visit = [&](uint32_t u) {
    const uint32_t w = u >> 5, bit = 1u << (u & 31);
    const uint32_t old = atomicOr(&bits[w], bit);
    if (old == 0) atomicOr(&sum[w >> 5], 1u << (w & 31));
    return (old & bit) == 0;
  };
  for (;;) {
    unsigned long long idx = 0;
    if (lane == 0) idx = atomicAdd(&P.ctl[kCtlBigCounter], 1ull);
    idx = uniform_u64(__shfl(idx, 0, 64));
    if (idx >= P.count) break;
    const uint64_t s = P.first + P.big_ids[idx];
    const uint64_t z = sample_z(P.seed, s);
    const uint32_t root = uniform_index(z, P.n);
    if (lane == 0) {
      visit(root);
      list[0] = root;
    }
    uint32_t head = 0, tail = 1;
    while (head < tail) {
      __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
      const uint32_t v = uniform_u32(list[head++]);
      ++st_exp;
      // every vertex enters the list once, so it never passes n entries; the
      // cap test of expand_row counts live edges, not new vertices, so it
      // gets no bound here (with cap n it could stop a row near the end)
      expand_row(P.vrec, P.ids, P.val, v, z, s_p2, &s_tab, list, tail, 0xFFFFFFFFu, st_cand, st_draws,
                 st_bytes, st_live, visit);
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    const uint32_t len = tail;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(len));
    base = uniform_u64(__shfl(base, 0, 64));
    const bool fits = base + len <= P.cap_data;
    // 1. indices of the marked words, in order, into the (finished) list
    uint32_t nw = 0;
    for (uint32_t s0 = 0; s0 < P.swords; s0 += 64) {
      const uint32_t sw = s0 + lane;
      // the bitmaps are written by device-scope atomics (in L2): read them
      // past the vector L1, which may hold lines from an earlier sample
      uint32_t m = sw < P.swords ? __hip_atomic_load(&sum[sw], __ATOMIC_RELAXED,
                                                      __HIP_MEMORY_SCOPE_AGENT)
                                 : 0u;
      const uint32_t c = __popc(m);
      const uint32_t incl = wave_inclusive_scan(c);
      const uint32_t total = uniform_u32(__shfl(incl, 63, 64));
      if (total == 0) continue;
      uint32_t at = nw + incl - c;
      if (m) __hip_atomic_store(&sum[sw], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      while (m) {
        const uint32_t b = __ffs(m) - 1;
        list[at++] = sw * 32 + b;
        m &= m - 1;
      }
      nw += total;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    // 2. the marked words in order: their set bits are the sorted set
    uint32_t out = 0;
    for (uint32_t i0 = 0; i0 < nw; i0 += 64) {
      const uint32_t i = i0 + lane;
      const uint32_t w = i < nw ? list[i] : 0u;
      uint32_t m = i < nw ? __hip_atomic_load(&bits[w], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                          : 0u;
      const uint32_t c = __popc(m);
      const uint32_t incl = wave_inclusive_scan(c);
      const uint32_t total = uniform_u32(__shfl(incl, 63, 64));
      uint32_t at = out + incl - c;
      if (i < nw) __hip_atomic_store(&bits[w], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      if (fits) {
        while (m) {
          const uint32_t b = __ffs(m) - 1;
          P.data[base + at++] = w * 32 + b;
          m &= m - 1;
        }
      }
      out += total;
    }
    if (lane == 0) {
      if (fits) {
        P.set_off[s] = base;
        P.set_len[s] = len;
        st_bytes += 4ull * len;
      } else {
        atomicOr(&P.ctl[kCtlOverflow], 1ull);
        P.set_len[s] = kLenFail;
      }
    }
  }
  if (lane == 0) st_bytes += kIcRecBytes * st_exp;
  report_stats(P.ctl, st_bytes, st_cand, st_draws, lane == 0 ? st_exp : 0ull, st_live);
}

__global__ void reset_big_kernel(unsigned long long *ctl) { ctl[kCtlBigCounter] = 0; }

template <int BT>
uint32_t wave_grid() {
  return static_cast<uint32_t>(
      blocks_per_cu(reinterpret_cast<const void *>(&ic_wave_kernel<BT, false>), BT) * cu_count());
}

}  // namespace

// Returns the number of samples left for the colour sampler (their local
// indices are in g->wave.d_big).
constexpr int kWaveBT = 256;

uint32_t ic_wave_blocks(uint64_t count) {
  return std::max<uint32_t>(
      1, std::min<uint64_t>(wave_grid<kWaveBT>(),
                            (count + kWaveBT / 64 - 1) / (kWaveBT / 64)));
}

uint32_t ic_chunk(const rip_pool *p, uint64_t count, uint32_t out_chunk, uint64_t waves) {
  if (out_chunk) return out_chunk;
  // rows past the wave's list go to the fallback, so sets written here are
  // at most 4096 long
  return out_chunk_rule(static_cast<double>(count) * std::min(p->avg_set_size, 4096.0) /
                            static_cast<double>(waves),
                        4096);
}

uint64_t ic_output_slack(const rip_pool *p, uint64_t count, const rip_tune *tune) {
  const uint64_t waves = static_cast<uint64_t>(ic_wave_blocks(count)) * (kWaveBT / 64);
  uint64_t slack = waves * ic_chunk(p, count, tune ? tune->out_chunk : 0, waves);
  // the lane kernel runs first unless a sampler kind without it is asked for
  const uint32_t kind = tune ? tune->sampler : RIP_SAMPLER_AUTO;
  if (kind == RIP_SAMPLER_AUTO || kind == RIP_SAMPLER_LANE) slack += ic_lane_slack(p, count, tune);
  return slack;
}

__global__ void reset_wave_kernel(unsigned long long *ctl) {
  ctl[kCtlGroupCounter] = 0;
  ctl[kCtlLongCount] = 0;
}

uint64_t ic_wave_launch(rip_pool *p, uint64_t first, uint64_t count, const uint32_t *ids,
                        uint32_t out_chunk, bool record_events, bool vertex_parallel,
                        bool stage) {
  constexpr int BT = kWaveBT;
  constexpr uint32_t CAP = 4096, TCAP = 8192;
  rip_graph *g = p->g;
  WaveScratch &w = g->wave;
  const uint32_t blocks = ic_wave_blocks(count);
  const uint64_t waves = static_cast<uint64_t>(blocks) * (BT / 64);
  if (w.waves < waves) {
    dev_free(w.d_table);
    dev_free(w.d_list);
    dev_free(w.d_epoch);
    w.waves = 0;
    w.d_table = dev_alloc<unsigned long long>(waves * TCAP);
    w.d_list = dev_alloc<uint32_t>(waves * CAP);
    w.d_epoch = dev_alloc<uint32_t>(waves);
    RIP_HIP(hipMemsetAsync(w.d_table, 0, sizeof(unsigned long long) * waves * TCAP, p->stream));
    RIP_HIP(hipMemsetAsync(w.d_epoch, 0, sizeof(uint32_t) * waves, p->stream));
    w.waves = waves;
  }
  if (w.cap_big < count) {
    dev_free(w.d_big);
    w.d_big = dev_alloc<uint32_t>(count);
    w.cap_big = count;
  }
  WaveParams P;
  P.vrec = g->d_vrec;
  P.ids = g->d_ids;
  P.val = g->d_val;
  P.n = g->n;
  P.step = g->step;
  P.seed = p->seed;
  P.first = first;
  P.count = count;
  P.table = w.d_table;
  P.list = w.d_list;
  P.epoch = w.d_epoch;
  P.cap = CAP;
  P.tcap = TCAP;
  P.big_ids = w.d_big;
  P.set_off = p->d_off;
  P.set_len = p->d_len;
  P.data = p->d_data;
  P.cap_data = p->cap_data;
  P.ctl = p->d_ctl;
  P.chunk = ic_chunk(p, count, out_chunk, waves);
  P.sample_ids = ids;
  hipLaunchKernelGGL(reset_wave_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  if (record_events) RIP_HIP(hipEventRecord(p->ev0, p->stream));
  if (stage) RIP_HIP(hipEventRecord(p->ev3, p->stream));
  if (vertex_parallel)
    hipLaunchKernelGGL((ic_wave_kernel<BT, true>), dim3(blocks), dim3(BT), 0, p->stream, P);
  else
    hipLaunchKernelGGL((ic_wave_kernel<BT, false>), dim3(blocks), dim3(BT), 0, p->stream, P);
  RIP_HIP(hipGetLastError());
  if (record_events) RIP_HIP(hipEventRecord(p->ev1, p->stream));
  if (stage) RIP_HIP(hipEventRecord(p->ev4, p->stream));
  pool_read_ctl(p, record_events);
  if (stage) {
    float ms = 0.0f;
    RIP_HIP(hipEventElapsedTime(&ms, p->ev3, p->ev4));
    p->wave_ms = ms;
    p->wave_ran = true;
    for (int i = 0; i < RIP_STAT_COUNT; ++i) p->wave_stats[i] = p->h_ctl[kCtlStatBytes + i];
  }
  return p->h_ctl[kCtlLongCount];
}

}  // namespace rip

namespace rip {

// Finishes the `count` samples listed in g->wave.d_big (local indices from
// `first`) with ic_big_kernel.  Per resident wave it needs n + n/32 + n/1024
// words; the number of waves is bounded by occupancy and by 1/16 of the
// device memory.
void ic_big_launch(rip_pool *p, uint64_t first, uint64_t count) {
  constexpr int BT = 256;
  rip_graph *g = p->g;
  WaveScratch &w = g->wave;
  const uint32_t words = (g->n + 31) / 32, swords = (words + 31) / 32;
  const uint64_t per_wave = 4ull * (static_cast<uint64_t>(g->n) + words + swords);
  size_t free_b = 0, total_b = 0;
  RIP_HIP(hipMemGetInfo(&free_b, &total_b));
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(&ic_big_kernel<BT>), BT);
  const int cus = cu_count();
  uint64_t waves = std::min<uint64_t>(static_cast<uint64_t>(per_cu) * cus * (BT / 64),
                                      std::max<uint64_t>(1, (total_b / 16) / per_wave));
  waves = std::min<uint64_t>(waves, (count + BT / 64 - 1) / (BT / 64) * (BT / 64));
  waves = std::max<uint64_t>(BT / 64, waves / (BT / 64) * (BT / 64));
  if (w.big_waves < waves || w.big_n != g->n) {
    dev_free(w.d_bbits);
    dev_free(w.d_bsum);
    dev_free(w.d_blist);
    w.big_waves = 0;
    w.d_bbits = dev_alloc<uint32_t>(waves * words);
    w.d_bsum = dev_alloc<uint32_t>(waves * swords);
    w.d_blist = dev_alloc<uint32_t>(waves * g->n);
    RIP_HIP(hipMemsetAsync(w.d_bbits, 0, sizeof(uint32_t) * waves * words, p->stream));
    RIP_HIP(hipMemsetAsync(w.d_bsum, 0, sizeof(uint32_t) * waves * swords, p->stream));
    w.big_waves = waves;
    w.big_n = g->n;
  }
  BigParams P;
  P.vrec = g->d_vrec;
  P.ids = g->d_ids;
  P.val = g->d_val;
  P.n = g->n;
  P.step = g->step;
  P.seed = p->seed;
  P.first = first;
  P.big_ids = w.d_big;
  P.count = count;
  P.bits = w.d_bbits;
  P.sum = w.d_bsum;
  P.list = w.d_blist;
  P.words = words;
  P.swords = swords;
  P.set_off = p->d_off;
  P.set_len = p->d_len;
  P.data = p->d_data;
  P.cap_data = p->cap_data;
  P.ctl = p->d_ctl;
  hipLaunchKernelGGL(reset_big_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  hipLaunchKernelGGL((ic_big_kernel<BT>), dim3(static_cast<uint32_t>(waves / (BT / 64))),
                     dim3(BT), 0, p->stream, P);
  RIP_HIP(hipGetLastError());
}

}  // namespace rip
