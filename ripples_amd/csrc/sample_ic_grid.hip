// Grid sampler: the IC samples that outgrow a lane's list, 64 per colour
// group and many groups per batch, expanded by the whole GPU one BFS level at
// a time.
//
// Semantics are those of AddRRRSet (add_rrrset.h:105-157) under the stream
// contract (common.h, ic_gap.h): the set is the reverse closure of the root
// over live edges, so any expansion order gives the sets of the other
// samplers and the oracle bit for bit.  The colour-per-sample idea is that of
// BatchedBFSNeighborColor (batched_add_rrrset.h:184-257); the reference's GPU
// form runs one BFS level per kernel as well (gpu/bfs.h:1338-1680).
//
// Samples that reach a giant component (R-MAT with U[0,0.1] or U[0,0.02]
// weights) visit hundreds of thousands of vertices each.  Per group g (64
// samples, colour c = bit c) the scratch holds
//   vis[g][v], pend[g][v] : colour masks of the visited and of the not yet
//                           expanded colours of vertex v (dense, n words each)
//   bm0[g], bm1[g]        : bitmap of the group's union and its summary
// so the groups of a batch cost 16 bytes per vertex each, which bounds the
// batch by memory rather than by the number of workgroups: every level is
// one persistent kernel over the frontier items (g, v) of all groups of the
// batch, pulling chunks of items from a global counter.  A chunk's rows are
// split into units of one colour and up to eight gap blocks, balanced over
// the workgroup by a scan; a lane draws its unit's candidates one live edge
// per step, and every live edge is one discovery: a read of vis (most edges
// in a giant component reach vertices the colour has visited), then atomicOr
// on vis and pend, and (g, u) joins the next frontier when its pend word was
// zero.  A frontier longer than its buffer sets an overflow flag; the
// vertices it could not list keep their pend bits, and a scan of pend finds
// them once the listed frontier is done.
//
// After the last level, one workgroup per group reads its union in vertex
// order from the bitmaps (clearing them and vis), keeps the group's hot
// members in a column of the pool's column store when they are many, and
// writes each colour's sorted list through one reservation of pool space.
#include <algorithm>
#include <vector>

#include "device_util.h"
#include "ic_common.h"
#include "ic_gap.h"
#include "internal.h"

namespace rip {
namespace {

constexpr int kExpBT = 256;          // expand kernel
constexpr int kOutBT = 1024;         // output kernel, one workgroup per group
// rows pending for this many colours or more are expanded a block per wave,
// lane c drawing colour c, so a live edge is one discovery for all colours
constexpr uint32_t kDenseColours = 8;
// level counters kept on the device (a ring), and levels enqueued per host
// synchronisation: a BFS of hundreds of narrow levels costs a launch per
// level, not a round trip
constexpr uint32_t kRing = 64;
constexpr uint32_t kLevelsPerSync = 16;

// control words of the grid scratch
enum : int { kGOver = 0, kGLevels = 1, kGCount = 2, kGWords = 4 };

struct GridParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  Affine step;
  unsigned long long *vis;         // groups x n
  unsigned long long *pend;        // groups x n
  uint32_t *bm0, *bm1;             // groups x nw0, groups x nw1
  uint32_t nw0, nw1;
  const unsigned long long *z;     // groups x 64: element 0 of each colour's sample
  const unsigned long long *cur;   // frontier items g << 32 | v
  uint64_t ncur;
  unsigned long long *next;        // next frontier
  unsigned long long *next_cnt;    // its length (counts every append, also past cap)
  uint64_t cap;                    // capacity of next
  unsigned long long *gctl;        // control words: overflow flag, level count
  unsigned long long *ring;        // per level L: ring[L % kRing] frontier length,
                                   // ring[kRing + L % kRing] chunk counter
  unsigned long long *buf[2];      // level L reads buf[L % 2], writes buf[(L + 1) % 2]
  unsigned long long *ctl;         // the pool's control words (work counters)
};

// The j-th set bit of m (j < popcount(m)).
__device__ inline uint32_t nth_bit(uint64_t m, uint32_t j) {
  uint32_t pos = 0;
  uint32_t c = __popc(static_cast<uint32_t>(m));
  uint32_t w = static_cast<uint32_t>(m);
  if (j >= c) {
    j -= c;
    w = static_cast<uint32_t>(m >> 32);
    pos = 32;
  }
  c = __popc(w & 0xFFFFu);
  if (j >= c) { j -= c; w >>= 16; pos += 16; }
  c = __popc(w & 0xFFu);
  if (j >= c) { j -= c; w >>= 8; pos += 8; }
  c = __popc(w & 0xFu);
  if (j >= c) { j -= c; w >>= 4; pos += 4; }
  c = __popc(w & 0x3u);
  if (j >= c) { j -= c; w >>= 2; pos += 2; }
  if (j >= (w & 1u)) pos += 1;
  return pos;
}

// Marks the colours `bits` of group g discovered at u (for lanes with live
// set); (g, u) joins the next frontier when its pend word was zero.  Every
// lane of the wave must call it.
__device__ inline void grid_discover(const GridParams &P, bool live, uint32_t g, uint32_t u,
                                     unsigned long long bits) {
  bool push = false;
  if (live) {
    const uint64_t idx = static_cast<uint64_t>(g) * P.n + u;
    // vis is written by atomics, which complete in L2; the read goes there too
    unsigned long long old = __hip_atomic_load(&P.vis[idx], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    if (bits & ~old) {
      old = atomicOr(&P.vis[idx], bits);
      const unsigned long long fresh = bits & ~old;
      if (fresh) {
        if (old == 0) {
          uint32_t *b0 = P.bm0 + static_cast<uint64_t>(g) * P.nw0;
          const uint32_t o0 = atomicOr(&b0[u >> 5], 1u << (u & 31u));
          if (o0 == 0)
            atomicOr(&P.bm1[static_cast<uint64_t>(g) * P.nw1 + (u >> 10)], 1u << ((u >> 5) & 31u));
        }
        push = atomicOr(&P.pend[idx], fresh) == 0;
      }
    }
  }
  const uint64_t m = __ballot(push);
  if (m == 0) return;
  const uint32_t leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
  unsigned long long base = 0;
  if (lane_id() == leader) base = atomicAdd(P.next_cnt, static_cast<unsigned long long>(__popcll(m)));
  base = __shfl(base, static_cast<int>(leader), 64);
  if (push) {
    const unsigned long long at = base + __popcll(m & lanes_below());
    if (at < P.cap) P.next[at] = (static_cast<unsigned long long>(g) << 32) | u;
  }
}

__device__ inline uint32_t wave_min_u32(uint32_t v) {
  for (int d = 32; d > 0; d >>= 1) v = min(v, static_cast<uint32_t>(__shfl_xor(v, d, 64)));
  return v;
}

__device__ inline unsigned long long wave_or_u64(unsigned long long v) {
  for (int d = 32; d > 0; d >>= 1) v |= __shfl_xor(v, d, 64);
  return v;
}

// Batch start: the root of every colour of every group (one 64-lane block
// per group; lanes past the group's colours have no sample).
__global__ void __launch_bounds__(64) grid_seed_kernel(GridParams P, const uint32_t *nb) {
  const uint32_t g = blockIdx.x, c = threadIdx.x;
  const bool has = c < nb[g];
  const unsigned long long z = P.z[64ull * g + c];
  const uint32_t root = has ? uniform_index(z, P.n) : 0u;
  grid_discover(P, has, g, root, 1ull << c);
}

// Shared memory of the expansion: the step powers and, per item of a chunk,
// its row and pending colours with the prefix of its work.
template <int BT>
struct ExpandLds {
  Affine p2[64];
  IcLogTables tab;
  uint32_t tmp[BT / 64];
  uint32_t g[BT], beg[BT], deg[BT], T[BT], fl[BT], uf[BT], ul[BT], pre[BT], nch[BT], dpre[BT];
  unsigned long long q[BT], rq[BT], ja[BT], jb[BT], m[BT];
  uint32_t ucur;   // next unit of the chunk not claimed by a wave
};

struct ExpandStats {
  unsigned long long cand = 0, draws = 0, exp = 0, bytes = 0, live = 0;
};

// Expands the frontier items cur[c0, c0 + cl), cl <= BT: every thread of
// the workgroup calls it.  A chunk's rows are split into units of one colour
// and up to eight gap blocks (rows pending for few colours) or into wave
// tasks of one block for every colour (rows pending for many).
template <int BT>
__device__ void expand_chunk(const GridParams &P, const unsigned long long *cur, uint64_t c0,
                             uint32_t cl, ExpandLds<BT> &S, ExpandStats &st) {
  constexpr int NW = BT / 64;
  const uint32_t tid = threadIdx.x;
  __syncthreads();   // the previous chunk's waves are done with the tables
  Affine *s_p2 = S.p2;
  uint32_t *s_tmp = S.tmp, *s_g = S.g, *s_beg = S.beg, *s_deg = S.deg, *s_T = S.T, *s_fl = S.fl;
  uint32_t *s_uf = S.uf, *s_ul = S.ul, *s_pre = S.pre, *s_nch = S.nch, *s_dpre = S.dpre;
  unsigned long long *s_q = S.q, *s_rq = S.rq, *s_ja = S.ja, *s_jb = S.jb, *s_m = S.m;
  unsigned long long &st_cand = st.cand, &st_draws = st.draws, &st_exp = st.exp;
  unsigned long long &st_bytes = st.bytes, &st_live = st.live;
  uint32_t units = 0, ntask = 0;
  if (tid < cl) {
    const unsigned long long item = cur[c0 + tid];
    const uint32_t g = static_cast<uint32_t>(item >> 32), v = static_cast<uint32_t>(item);
    // the colours still to expand at v; a repeated item finds none
    const unsigned long long M = atomicExch(&P.pend[static_cast<uint64_t>(g) * P.n + v], 0ull);
    uint32_t nch = 0;
    if (M) {
      const VertexRec r = P.vrec[v];
      ++st_exp;
      st_bytes += kIcRecBytes;
      s_g[tid] = g;
      s_beg[tid] = r.begin;
      s_deg[tid] = r.end - r.begin;
      s_T[tid] = r.thr;
      s_fl[tid] = r.flags;
      s_uf[tid] = r.ufull;
      s_ul[tid] = r.ulast;
      s_q[tid] = r.q;
      s_rq[tid] = r.rq;
      s_ja[tid] = r.ja;
      s_jb[tid] = r.jb;
      s_m[tid] = M;
      if (r.thr) {
        const uint32_t beta = (r.flags >> kBetaShift) & 31u;
        const uint32_t nblk = ((r.end - r.begin - 1) >> beta) + 1;
        if (__popcll(M) >= kDenseColours) {
          ntask = nblk;   // one wave per block, a lane per colour
        } else {
          nch = nblk;   // a unit per (colour, block)
          units = static_cast<uint32_t>(__popcll(M)) * nch;
        }
      }
    }
    s_nch[tid] = nch;
  }
  uint32_t total, dtotal;
  const uint32_t pre = block_exclusive_scan<BT>(units, s_tmp, &total);
  const uint32_t dpre = block_exclusive_scan<BT>(ntask, s_tmp, &dtotal);
  if (tid < cl) {
    s_pre[tid] = pre;
    s_dpre[tid] = dpre;
  }
  if (tid == 0) S.ucur = 0;
  __syncthreads();
  // Rows pending for few colours: unit u is block k of colour j of entry
  // ent.  Each wave claims 64 units at a time from the chunk's counter; a
  // lane with no unit takes the next one, and every step each lane with a
  // unit makes one draw (the block's first element and its first gap, or the
  // next gap) and tests its candidate, so the lanes stay busy whatever the
  // lengths of their units; a live edge is discovered in the same step.
  const uint32_t lane = lane_id(), wid = tid >> 6;
  {
    const Affine Pst = s_p2[0];
    uint32_t nxt = 0, hi = 0;   // the wave's claimed units [nxt, hi) (wave-uniform)
    bool has = false, started = false;
    uint32_t g = 0, e0 = 0, len = 0, pos = 0, T = 0, uskip = 0;
    unsigned long long bit = 0;
    uint64_t q = 0, rq = 0, x = 0;
    bool uniform = true;
    for (;;) {
      const uint64_t idle = __ballot(!has);
      if (idle) {
        if (nxt >= hi) {
          uint32_t c0 = 0;
          if (lane == 0) c0 = atomicAdd(&S.ucur, 64u);
          c0 = uniform_u32(__shfl(c0, 0, 64));
          nxt = min(c0, total);
          hi = min(c0 + 64u, total);
        }
        if (!has) {
          const uint32_t u = nxt + static_cast<uint32_t>(__popcll(idle & lanes_below()));
          if (u < hi) {
            uint32_t lo = 0, up = cl;
            while (lo < up) {
              const uint32_t mid = (lo + up) >> 1;
              if (s_pre[mid] <= u) lo = mid + 1;
              else up = mid;
            }
            const uint32_t ent = lo - 1;   // the last entry starting at or before u
            const uint32_t r = u - s_pre[ent];
            const uint32_t nblk = s_nch[ent];
            const uint32_t j = r / nblk, k = r - j * nblk;
            g = s_g[ent];
            const uint32_t c = nth_bit(s_m[ent], j);
            bit = 1ull << c;
            const uint32_t fl = s_fl[ent];
            const uint32_t beta = (fl >> kBetaShift) & 31u;
            const uint32_t B = 1u << beta;
            e0 = s_beg[ent] + (k << beta);
            len = min(s_deg[ent] - (k << beta), B);
            T = s_T[ent];
            uskip = len == B ? s_uf[ent] : s_ul[ent];
            q = s_q[ent];
            rq = s_rq[ent];
            uniform = (fl & kRowUniform) != 0;
            // block k's first element: P^(2 B k) after the row's first
            Affine E{s_ja[ent], s_jb[ent]};
            for (uint32_t kk = k, b2 = beta + 1; kk; kk >>= 1, ++b2)
              if (kk & 1) E = aff_then(E, s_p2[b2]);
            x = E.a * P.z[64ull * g + c] + E.b;
            pos = 0;
            has = true;
            started = false;
          }
        }
        nxt = min(nxt + static_cast<uint32_t>(__popcll(idle)), hi);
      }
      if (__ballot(has) == 0) {
        if (nxt >= total) break;
        continue;
      }
      bool live = false;
      uint32_t src = 0;
      if (has) {
        if (!started) {
          started = true;
          ++st_draws;
          // a block without candidates costs one compare
          if ((x >> 32) + 1 <= uskip) has = false;
        }
        if (has) {
          const uint64_t gap = ic_gap_fast(x, q, rq, &S.tab);
          if (gap >= len - pos) {
            has = false;
          } else {
            pos += static_cast<uint32_t>(gap);
            const uint32_t e = e0 + pos;
            const uint64_t a = Pst.a * x + Pst.b;        // acceptance element
            ++st_cand;
            if (!uniform) {
              ++st_draws;
              st_bytes += 4;
            }
            live = uniform || ic_accept(a, T, P.val[e]);
            ++pos;
            x = Pst.a * a + Pst.b;                       // next gap element
            if (pos < len) ++st_draws;
            else has = false;
            if (live) {
              src = P.ids[e];
              ++st_live;
              st_bytes += 4;
            }
          }
        }
      }
      grid_discover(P, live, g, src, bit);
    }
  }
  // Rows pending for many colours: wave task t is block k of entry ent.
  // Lane c walks colour c's candidates of the block; the wave takes them a
  // window of 64 positions at a time (from the smallest candidate left),
  // one ballot per live position gives the colours that position is live
  // for, and each position is one discovery with all of them.
  for (uint32_t t = wid; t < dtotal; t += NW) {
    uint32_t lo = 0, hi = cl;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (s_dpre[mid] <= t) lo = mid + 1;
      else hi = mid;
    }
    const uint32_t ent = uniform_u32(lo - 1);   // the last entry starting at or before t
    const uint32_t k = t - s_dpre[ent];
    const uint32_t g = s_g[ent];
    const unsigned long long M = s_m[ent];
    const uint32_t fl = s_fl[ent];
    const uint32_t beta = (fl >> kBetaShift) & 31u;
    const uint32_t B = 1u << beta;
    const uint32_t deg = s_deg[ent];
    const uint32_t T = s_T[ent];
    const uint64_t q = s_q[ent], rq = s_rq[ent];
    const bool uniform = (fl & kRowUniform) != 0;
    const uint32_t e0 = s_beg[ent] + (k << beta);
    const uint32_t len = min(deg - (k << beta), B);
    Affine E{s_ja[ent], s_jb[ent]};
    for (uint32_t kk = k, bit = beta + 1; kk; kk >>= 1, ++bit)
      if (kk & 1) E = aff_then(E, s_p2[bit]);
    const Affine Pst = s_p2[0];
    const bool act = (M >> lane) & 1ull;
    // lane state: next search position, stream element, current candidate
    // cp (len: none left) and whether it is live
    uint32_t pos = 0, cp = len;
    bool clive = false;
    uint64_t x = 0;
    auto advance = [&]() {
      const uint64_t gap = ic_gap_fast(x, q, rq, &S.tab);
      if (gap >= len - pos) {
        cp = len;
        return;
      }
      pos += static_cast<uint32_t>(gap);
      cp = pos;
      const uint64_t a = Pst.a * x + Pst.b;      // acceptance element
      ++st_cand;
      if (!uniform) {
        ++st_draws;
        st_bytes += 4;
      }
      clive = uniform || ic_accept(a, T, P.val[e0 + cp]);
      if (clive) ++st_live;
      ++pos;
      x = Pst.a * a + Pst.b;                     // next gap element
      if (pos < len) ++st_draws;
    };
    if (act) {
      x = E.a * P.z[64ull * g + lane] + E.b;
      ++st_draws;
      // a block without candidates costs one compare
      if ((x >> 32) + 1 > (len == B ? s_uf[ent] : s_ul[ent])) advance();
    }
    for (;;) {
      const uint32_t w0 = uniform_u32(wave_min_u32(cp));
      if (w0 >= len) break;
      unsigned long long L = 0;
      while (cp < w0 + 64u && cp < len) {
        if (clive) L |= 1ull << (cp - w0);
        advance();
      }
      const unsigned long long U = uniform_u64(wave_or_u64(L));
      const uint32_t npos = static_cast<uint32_t>(__popcll(U));
      uint32_t mypos = 0;
      unsigned long long mybits = 0;
      for (uint32_t j = 0; j < npos; ++j) {
        const uint32_t p = nth_bit(U, j);
        const unsigned long long m = __ballot((L >> p) & 1ull);
        if (lane == j) {
          mypos = p;
          mybits = m;
        }
      }
      const bool lv = lane < npos;
      const uint32_t u = lv ? P.ids[e0 + w0 + mypos] : 0u;
      if (lv) st_bytes += 4;
      grid_discover(P, lv, g, u, mybits);
    }
  }
}

// BFS level L: the frontier is buf[L % 2][0, n) with n = ring[L % kRing]
// (its first cap entries when it overflowed, which raises the overflow
// flag); every workgroup with a chunk to take pulls chunks of BT items from
// the level's counter.  Levels are enqueued ahead of the host: a level whose
// frontier is empty costs its launch only.
template <int BT>
__global__ void __launch_bounds__(BT) grid_level_kernel(GridParams P, uint32_t L) {
  __shared__ ExpandLds<BT> S;
  __shared__ unsigned long long s_chunk;
  const uint32_t tid = threadIdx.x;
  const uint64_t raw = __hip_atomic_load(&P.ring[L % kRing], __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
  const uint64_t ncur = raw < P.cap ? raw : P.cap;
  if (blockIdx.x == 0 && tid == 0 && ncur) {
    P.gctl[kGLevels] += 1;
    if (raw > P.cap) P.gctl[kGOver] = 1;
  }
  if (static_cast<uint64_t>(blockIdx.x) * BT >= ncur) return;   // no chunk for this workgroup
  GridParams Q = P;
  Q.cur = P.buf[L & 1u];
  Q.ncur = ncur;
  Q.next = P.buf[(L + 1) & 1u];
  Q.next_cnt = &P.ring[(L + 1) % kRing];
  unsigned long long *head = &P.ring[kRing + L % kRing];
  if (tid == 0) fill_step_powers(P.step, S.p2);
  ic_load_tables(&S.tab);
  ExpandStats st;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_chunk = atomicAdd(head, static_cast<unsigned long long>(BT));
    __syncthreads();
    const uint64_t c0 = s_chunk;
    if (c0 >= ncur) break;
    expand_chunk<BT>(Q, Q.cur, c0, static_cast<uint32_t>(min(static_cast<uint64_t>(BT), ncur - c0)),
                     S, st);
  }
  report_stats(P.ctl, st.bytes, st.cand, st.draws, st.exp, st.live);
}

// Zeroes the frontier lengths of levels a .. a + k - 1 and the chunk
// counters of levels a - 1 .. a + k - 2 (ring indices mod kRing), or, with
// `all`, the whole ring and the control words.
__global__ void grid_ring_kernel(unsigned long long *ring, unsigned long long *gctl, uint32_t a,
                                 uint32_t k, uint32_t all) {
  const uint32_t i = threadIdx.x;
  if (all) {
    if (i < 2 * kRing) ring[i] = 0;
    if (i < kGWords) gctl[i] = 0;
    return;
  }
  if (i < k) {
    ring[(a + i) % kRing] = 0;
    ring[kRing + (a + i - 1) % kRing] = 0;
  }
}

// Frontier from the pend words (after an overflow): every (g, v) of groups
// [0, groups) with colours still to expand.
__global__ void grid_rebuild_kernel(GridParams P, uint32_t groups) {
  const uint64_t total = static_cast<uint64_t>(groups) * P.n;
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i0 = static_cast<uint64_t>(blockIdx.x) * blockDim.x; i0 < total; i0 += stride) {
    const uint64_t i = i0 + threadIdx.x;
    const bool push = i < total && __hip_atomic_load(&P.pend[i], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT) != 0;
    const uint64_t m = __ballot(push);
    if (m == 0) continue;
    const uint32_t leader = __ffsll(static_cast<unsigned long long>(m)) - 1;
    unsigned long long base = 0;
    if (lane_id() == leader)
      base = atomicAdd(P.next_cnt, static_cast<unsigned long long>(__popcll(m)));
    base = __shfl(base, static_cast<int>(leader), 64);
    if (push) {
      const unsigned long long at = base + __popcll(m & lanes_below());
      const uint64_t g = i / P.n, v = i - g * P.n;
      if (at < P.cap) P.next[at] = (g << 32) | v;
    }
  }
}


// Union size of every group: the set bits of its bitmap.
__global__ void __launch_bounds__(256) grid_union_kernel(const uint32_t *bm0, uint32_t nw0,
                                                         unsigned long long *out) {
  __shared__ unsigned long long s_w[4];
  const uint32_t *b = bm0 + static_cast<uint64_t>(blockIdx.x) * nw0;
  unsigned long long acc = 0;
  for (uint32_t i = threadIdx.x; i < nw0; i += 256)
    acc += __popc(__hip_atomic_load(&b[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
  for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) out[blockIdx.x] = s_w[0] + s_w[1] + s_w[2] + s_w[3];
}

struct OutParams {
  uint32_t n, nw0, nw1;
  unsigned long long *vis;
  uint32_t *bm0, *bm1;
  uint32_t g0;                          // first group of this launch
  const unsigned long long *soff;       // per group of the launch: staging offset
  uint32_t *slist;                      // staging: union lists
  unsigned long long *smask;            // staging: union masks
  const uint32_t *sid;                  // groups x 64 sample ids (kNoGroup: none)
  const uint32_t *nb;                   // colours per group
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
  // column store (colptr == nullptr: every set a list)
  unsigned long long *const *colptr;
  uint32_t *gsample;
  uint32_t *gref;
  const uint32_t *hot_index;
  uint32_t hot_n;
  uint32_t group_base;
};

template <int BT>
__global__ void __launch_bounds__(BT) grid_output_kernel(OutParams P) {
  constexpr int NW = BT / 64;
  __shared__ uint32_t s_tmp[NW];
  __shared__ unsigned long long s_hot[NW];
  __shared__ uint32_t s_cnt[64], s_cofs[64];
  __shared__ unsigned long long s_base;
  __shared__ uint32_t s_fail, s_col;
  const uint32_t tid = threadIdx.x, lane = lane_id(), wid = tid >> 6;
  const uint32_t g = P.g0 + blockIdx.x;
  const uint32_t nb = P.nb[g];
  uint32_t *bm0 = P.bm0 + static_cast<uint64_t>(g) * P.nw0;
  uint32_t *bm1 = P.bm1 + static_cast<uint64_t>(g) * P.nw1;
  unsigned long long *vis = P.vis + static_cast<uint64_t>(g) * P.n;
  uint32_t *list = P.slist + P.soff[blockIdx.x];
  unsigned long long *masks = P.smask + P.soff[blockIdx.x];
  unsigned long long st_bytes = 0;
  if (tid < 64) s_cnt[tid] = 0;
  // the union in vertex order; bitmaps and vis read and cleared with atomics
  uint32_t running = 0;
  for (uint32_t t0 = 0; t0 < P.nw1; t0 += BT) {
    const uint32_t w1i = t0 + tid;
    const uint32_t w1 = w1i < P.nw1 ? __hip_atomic_load(&bm1[w1i], __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT)
                                    : 0u;
    uint32_t cnt = 0;
    for (uint32_t b = w1; b; b &= b - 1)
      cnt += __popc(__hip_atomic_load(&bm0[w1i * 32 + (__ffs(b) - 1)], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT));
    uint32_t tot;
    uint32_t p = running + block_exclusive_scan<BT>(cnt, s_tmp, &tot);
    for (uint32_t b = w1; b; b &= b - 1) {
      const uint32_t w0i = w1i * 32 + (__ffs(b) - 1);
      const uint32_t w0 = atomicExch(&bm0[w0i], 0u);
      for (uint32_t bb = w0; bb; bb &= bb - 1) {
        const uint32_t v = w0i * 32 + (__ffs(bb) - 1);
        list[p] = v;
        masks[p] = atomicExch(&vis[v], 0ull);
        ++p;
      }
    }
    if (w1) atomicExch(&bm1[w1i], 0u);
    running += tot;
  }
  __syncthreads();
  const uint32_t U = running;
  if (P.colptr) {
    // a column costs 8 bytes per hot vertex, lists 4 per member: the group's
    // hot members decide
    unsigned long long hot = 0;
    for (uint32_t i = tid; i < U; i += BT)
      if (P.hot_index[list[i]] != kNotHot) hot += __popcll(masks[i]);
    unsigned long long tot_hot;
    (void)block_exclusive_scan_u64<BT>(hot, s_hot, &tot_hot);
    if (tid == 0) {
      s_col = kNoGroup;
      if (tot_hot > 2ull * P.hot_n)
        s_col = P.group_base + static_cast<uint32_t>(atomicAdd(&P.ctl[kCtlGroupCounter2], 1ull));
    }
    __syncthreads();
    const uint32_t colg = s_col;
    if (colg != kNoGroup) {
      unsigned long long *col = P.colptr[colg];
      for (uint32_t h = tid; h < P.hot_n; h += BT) col[h] = 0ull;
      __syncthreads();
      for (uint32_t i = tid; i < U; i += BT) {
        const uint32_t h = P.hot_index[list[i]];
        if (h != kNotHot) {
          col[h] = masks[i];
          masks[i] = 0ull;
        }
      }
      st_bytes += 8ull * ((P.hot_n + BT - 1 - tid) / BT);
      if (tid == 0) atomicAdd(&P.ctl[kCtlElems], tot_hot);
      if (tid < 64) {
        const uint32_t sid = tid < nb ? P.sid[64ull * g + tid] : kNoGroup;
        if (tid < nb) P.gref[sid] = (colg << 6) | tid;
        P.gsample[64ull * colg + tid] = sid;
      }
    }
    __syncthreads();
  }
  for (uint32_t i = tid; i < U; i += BT)
    for (unsigned long long mk = masks[i]; mk; mk &= mk - 1) atomicAdd(&s_cnt[__ffsll(mk) - 1], 1u);
  __syncthreads();
  if (tid == 0) {
    uint32_t acc = 0;
    for (uint32_t c = 0; c < nb; ++c) {
      s_cofs[c] = acc;
      acc += s_cnt[c];
    }
    const unsigned long long base =
        atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(acc));
    s_base = base;
    s_fail = 0;
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
      const uint32_t s = P.sid[64ull * g + tid];
      P.set_off[s] = base + s_cofs[tid];
      P.set_len[s] = s_cnt[tid];
      st_bytes += 4ull * s_cnt[tid];
    }
  } else if (tid < nb) {
    P.set_len[P.sid[64ull * g + tid]] = kLenFail;
  }
  report_stats(P.ctl, st_bytes, 0, 0, 0, 0);
}

// Element 0 and sample id of every colour of the batch's groups.
__global__ void grid_samples_kernel(uint64_t seed, uint64_t first, const uint32_t *ids,
                                    uint64_t b0, uint64_t count, uint32_t groups,
                                    unsigned long long *z, uint32_t *sid, uint32_t *nb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 64u * groups) return;
  const uint64_t t = b0 + i;
  const bool has = t < count;
  const uint64_t s = has ? first + (ids ? ids[t] : t) : 0;
  z[i] = has ? sample_z(seed, s) : 0ull;
  sid[i] = has ? static_cast<uint32_t>(s) : kNoGroup;
  if ((i & 63u) == 0) {
    const uint64_t left = count > t ? count - t : 0;
    nb[i >> 6] = static_cast<uint32_t>(left < 64 ? left : 64);
  }
}

// Bytes of one group's scratch: vis and pend, the union bitmap and summary.
uint64_t group_bytes(uint32_t n) {
  return 16ull * n + 4ull * ((n + 31) / 32) + 4ull * ((n + 1023) / 1024);
}

}  // namespace

// Groups per batch and frontier capacity for `count` samples.
static void grid_plan(const rip_graph *g, uint64_t count, uint32_t *groups, uint64_t *fcap) {
  size_t free_b = 0, total_b = 0;
  RIP_HIP(hipMemGetInfo(&free_b, &total_b));
  const GridScratch &s = g->grid;
  const uint64_t held = s.n == g->n ? s.groups * group_bytes(g->n) + 16ull * s.fcap : 0;
  // the scratch stays within an eighth of the memory it may use, 32 GiB at most
  const uint64_t budget = std::min<uint64_t>((free_b + held) / 8, 32ull << 30);
  const uint64_t per = group_bytes(g->n) + 16ull * std::min<uint64_t>(g->n, 1u << 21);
  uint64_t G = std::max<uint64_t>(1, budget / per);
  G = std::min<uint64_t>({G, 256, (count + 63) / 64});
  *groups = static_cast<uint32_t>(G);
  // room for every group's frontier up to 2^21 vertices, and for one whole
  // union (the output staging reuses the frontier buffers)
  *fcap = std::max<uint64_t>(g->n, G * std::min<uint64_t>(g->n, 1u << 21));
}

static void grid_ensure(rip_graph *g, uint32_t groups, uint64_t fcap, hipStream_t st) {
  GridScratch &s = g->grid;
  const uint32_t n = g->n;
  if (s.n != n || s.groups < groups) {
    dev_free(s.d_vis);
    dev_free(s.d_pend);
    dev_free(s.d_bm0);
    dev_free(s.d_bm1);
    dev_free(s.d_z);
    dev_free(s.d_sid);
    dev_free(s.d_nb);
    s.groups = 0;
    const uint64_t nw0 = (n + 31) / 32, nw1 = (n + 1023) / 1024;
    s.d_vis = dev_alloc<unsigned long long>(static_cast<uint64_t>(groups) * n);
    s.d_pend = dev_alloc<unsigned long long>(static_cast<uint64_t>(groups) * n);
    s.d_bm0 = dev_alloc<uint32_t>(groups * nw0);
    s.d_bm1 = dev_alloc<uint32_t>(groups * nw1);
    s.d_z = dev_alloc<unsigned long long>(64ull * groups);
    s.d_sid = dev_alloc<uint32_t>(64ull * groups);
    s.d_nb = dev_alloc<uint32_t>(groups);
    // zero once; every batch leaves them zero again
    RIP_HIP(hipMemsetAsync(s.d_vis, 0, 8ull * groups * n, st));
    RIP_HIP(hipMemsetAsync(s.d_pend, 0, 8ull * groups * n, st));
    RIP_HIP(hipMemsetAsync(s.d_bm0, 0, 4ull * groups * nw0, st));
    RIP_HIP(hipMemsetAsync(s.d_bm1, 0, 4ull * groups * nw1, st));
    s.groups = groups;
    s.n = n;
  }
  if (s.fcap < fcap) {
    dev_free(s.d_fa);
    dev_free(s.d_fb);
    s.fcap = 0;
    s.d_fa = dev_alloc<unsigned long long>(fcap);
    s.d_fb = dev_alloc<unsigned long long>(fcap);
    s.fcap = fcap;
  }
  if (!s.d_gctl) {
    s.d_ring = dev_alloc<unsigned long long>(2 * kRing);
    s.d_gctl = dev_alloc<unsigned long long>(kGWords);
    s.d_union = dev_alloc<unsigned long long>(256);
    RIP_HIP(hipHostMalloc(reinterpret_cast<void **>(&s.h_ctl), sizeof(unsigned long long) * 256,
                          hipHostMallocDefault));
  }
}

// Reads `count` words of device memory into the scratch's pinned buffer.
static const unsigned long long *grid_read(GridScratch &s, const unsigned long long *src,
                                           uint32_t count, hipStream_t st) {
  RIP_HIP(hipMemcpyAsync(s.h_ctl, src, sizeof(unsigned long long) * count,
                         hipMemcpyDeviceToHost, st));
  RIP_HIP(hipStreamSynchronize(st));
  return s.h_ctl;
}

void grid_scratch_free(GridScratch &s) {
  dev_free(s.d_vis);
  dev_free(s.d_pend);
  dev_free(s.d_bm0);
  dev_free(s.d_bm1);
  dev_free(s.d_z);
  dev_free(s.d_sid);
  dev_free(s.d_nb);
  dev_free(s.d_fa);
  dev_free(s.d_fb);
  dev_free(s.d_gctl);
  dev_free(s.d_union);
  dev_free(s.d_ring);
  if (s.h_ctl) (void)hipHostFree(s.h_ctl);
  s.h_ctl = nullptr;
  s.groups = 0;
  s.fcap = 0;
  s.n = 0;
}

// Samples first + ids[i] (first + i without ids), i in [0, count), with the
// grid sampler; with columns, groups with many hot members go to the pool's
// column store.  Returns the kernel launches made.
uint32_t ic_grid_launch(rip_pool *p, uint64_t first, uint64_t count, const uint32_t *ids,
                        bool columns, uint32_t frontier_cap) {
  if (count == 0) return 0;
  rip_graph *gr = p->g;
  hipStream_t st = p->stream;
  uint32_t G = 0;
  uint64_t fcap = 0;
  grid_plan(gr, count, &G, &fcap);
  grid_ensure(gr, G, fcap, st);
  GridScratch &s = gr->grid;
  const uint64_t total_groups = (count + 63) / 64;
  if (columns) pool_reserve_groups(p, p->groups + static_cast<uint32_t>(total_groups));
  const uint32_t n = gr->n;
  GridParams P;
  P.vrec = gr->d_vrec;
  P.ids = gr->d_ids;
  P.val = gr->d_val;
  P.n = n;
  P.step = gr->step;
  P.vis = s.d_vis;
  P.pend = s.d_pend;
  P.bm0 = s.d_bm0;
  P.bm1 = s.d_bm1;
  P.nw0 = (n + 31) / 32;
  P.nw1 = (n + 1023) / 1024;
  P.z = s.d_z;
  // entries a level may list: the buffer, or fewer when asked (the buffers
  // keep room for one whole union, the output staging)
  const uint64_t lcap = frontier_cap ? std::min<uint64_t>(frontier_cap, s.fcap) : s.fcap;
  P.cap = lcap;
  P.gctl = s.d_gctl;
  P.ctl = p->d_ctl;
  P.ring = s.d_ring;
  P.buf[0] = s.d_fa;
  P.buf[1] = s.d_fb;
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(&grid_level_kernel<kExpBT>),
                                   kExpBT);
  const uint32_t exp_grid = static_cast<uint32_t>(per_cu * cu_count());
  uint32_t launches = 0;
  for (uint64_t b0 = 0; b0 < count; b0 += 64ull * G) {
    const uint32_t Gb = static_cast<uint32_t>(std::min<uint64_t>(G, (count - b0 + 63) / 64));
    hipLaunchKernelGGL(grid_samples_kernel, dim3((64 * Gb + 255) / 256), dim3(256), 0, st,
                       p->seed, first, ids, b0, count, Gb, s.d_z, s.d_sid, s.d_nb);
    hipLaunchKernelGGL(grid_ring_kernel, dim3(1), dim3(2 * kRing), 0, st, s.d_ring, s.d_gctl, 0u,
                       0u, 1u);
    // level 0: the roots, in buf[0]
    P.next = s.d_fa;
    P.next_cnt = &s.d_ring[0];
    hipLaunchKernelGGL(grid_seed_kernel, dim3(Gb), dim3(64), 0, st, P,
                       static_cast<const uint32_t *>(s.d_nb));
    launches += 3;
    uint32_t L = 0;            // the level whose frontier is ring[L % kRing]
    bool dirty = false;        // some vertices with pend bits are in no list
    const uint32_t levels0 = p->grid_levels;
    for (;;) {
      // kLevelsPerSync levels, then one look at the frontier length
      hipLaunchKernelGGL(grid_ring_kernel, dim3(1), dim3(kLevelsPerSync), 0, st, s.d_ring,
                         s.d_gctl, L + 1, kLevelsPerSync, 0u);
      RIP_HIP(hipEventRecord(p->ev3, st));
      for (uint32_t i = 0; i < kLevelsPerSync; ++i)
        hipLaunchKernelGGL((grid_level_kernel<kExpBT>), dim3(exp_grid), dim3(kExpBT), 0, st, P,
                           L + i);
      RIP_HIP(hipGetLastError());
      RIP_HIP(hipEventRecord(p->ev4, st));
      launches += 1 + kLevelsPerSync;
      L += kLevelsPerSync;
      const unsigned long long *h = grid_read(s, s.d_gctl, kGWords, st);
      const bool over = h[kGOver] != 0;
      const uint32_t lv = static_cast<uint32_t>(h[kGLevels]);
      uint64_t pending = 0;
      RIP_HIP(hipMemcpyAsync(s.h_ctl + kGWords, s.d_ring + L % kRing, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, st));
      RIP_HIP(hipStreamSynchronize(st));
      pending = s.h_ctl[kGWords];
      float ms = 0.0f;
      RIP_HIP(hipEventElapsedTime(&ms, p->ev3, p->ev4));
      p->grid_ms += ms;
      p->grid_levels = levels0 + lv;
      if (over) dirty = true;
      if (pending > lcap) dirty = true;
      if (pending != 0) continue;
      if (!dirty) break;
      // the vertices the overflowing lists could not hold, as level L
      P.next = P.buf[L & 1u];
      P.next_cnt = &s.d_ring[L % kRing];
      hipLaunchKernelGGL(grid_rebuild_kernel, dim3(2048), dim3(256), 0, st, P, Gb);
      launches += 1;
      RIP_HIP(hipMemcpyAsync(s.h_ctl + kGWords, s.d_ring + L % kRing, sizeof(unsigned long long),
                             hipMemcpyDeviceToHost, st));
      RIP_HIP(hipStreamSynchronize(st));
      pending = s.h_ctl[kGWords];
      dirty = pending > lcap;
      if (pending == 0) break;
    }
    // output: groups whose unions fit the staging together share a launch
    hipLaunchKernelGGL(grid_union_kernel, dim3(Gb), dim3(256), 0, st, s.d_bm0, P.nw0, s.d_union);
    std::vector<unsigned long long> usz(Gb);
    const unsigned long long *h = grid_read(s, s.d_union, Gb, st);
    std::copy(h, h + Gb, usz.begin());
    OutParams O;
    O.n = n;
    O.nw0 = P.nw0;
    O.nw1 = P.nw1;
    O.vis = s.d_vis;
    O.bm0 = s.d_bm0;
    O.bm1 = s.d_bm1;
    O.slist = reinterpret_cast<uint32_t *>(s.d_fa);
    O.smask = s.d_fb;
    O.sid = s.d_sid;
    O.nb = s.d_nb;
    O.set_off = p->d_off;
    O.set_len = p->d_len;
    O.data = p->d_data;
    O.cap_data = p->cap_data;
    O.ctl = p->d_ctl;
    O.colptr = columns ? p->d_colptr : nullptr;
    O.gsample = columns ? p->d_gsample : nullptr;
    O.gref = columns ? p->d_gref : nullptr;
    O.hot_index = columns ? gr->d_hot_index : nullptr;
    O.hot_n = columns ? gr->hot_n : 0;
    O.group_base = p->groups;
    for (uint32_t a = 0; a < Gb;) {
      // staging offsets of groups a, a + 1, ... while they fit (one always does)
      std::vector<unsigned long long> off;
      unsigned long long acc = 0;
      uint32_t e = a;
      while (e < Gb && (e == a || acc + usz[e] <= s.fcap)) {
        off.push_back(acc);
        acc += usz[e];
        ++e;
      }
      // the offsets travel in the union buffer (its counts are on the host)
      RIP_HIP(hipMemcpyAsync(s.d_union, off.data(), sizeof(unsigned long long) * off.size(),
                             hipMemcpyHostToDevice, st));
      O.g0 = a;
      O.soff = s.d_union;
      hipLaunchKernelGGL((grid_output_kernel<kOutBT>), dim3(e - a), dim3(kOutBT), 0, st, O);
      RIP_HIP(hipGetLastError());
      RIP_HIP(hipStreamSynchronize(st));   // `off` and the staging are reused
      launches += 1;
      a = e;
    }
    launches += 1;
  }
  return launches;
}

}  // namespace rip
