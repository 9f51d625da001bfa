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
//   vp[g][v] = (vis, pend) : colour masks of the visited and of the not yet
//                            expanded colours of vertex v, side by side in
//                            one 16-byte pair (dense, n pairs per group), so
//                            a discovery's two atomics touch one cache line
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
// frontier items per chunk of the expand kernel: half a workgroup, which
// keeps the chunk's tables small enough in LDS for six workgroups per CU
constexpr uint32_t kExpItems = 128;
// rows pending for this many colours or more are expanded a block per wave,
// lane c drawing colour c, so a live edge is one discovery for all colours
constexpr uint32_t kDenseColours = 8;
constexpr uint32_t kDenseSpan = 128;   // largest block of the dense path (edges)
constexpr uint32_t kDenseSlots = 8;    // dense tasks per round of a workgroup
constexpr uint32_t kDenseBatch = 4;    // candidates a dense lane reads thresholds for at once
constexpr uint32_t kSparseSteps = 8;   // steps of a sparse lane between discoveries
// level counters kept on the device (a ring), and levels enqueued per host
// synchronisation: a BFS of hundreds of narrow levels costs a launch per
// level, not a round trip
constexpr uint32_t kRing = 64;
constexpr uint32_t kLevelsPerSync = 16;
// a frontier of at most this many items (one chunk) is expanded by one
// workgroup that runs level after level with a barrier between them (no
// launch per level): the long thin tails of deep samples.  Wider frontiers go
// to the level kernels: one workgroup would take their chunks one by one.
constexpr uint64_t kTailItems = kExpItems;

// control words of the grid scratch
enum : int { kGOver = 0, kGLevels = 1, kGCount = 2, kGTail = 3, kGWords = 4 };

struct GridParams {
  const VertexRec *vrec;
  const uint32_t *ids;
  const uint32_t *val;
  uint32_t n;
  Affine step;
  unsigned long long *vp;          // groups x n pairs (visited, pending colours)
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
  uint32_t chunk;                  // frontier items a level kernel's workgroup takes at a time
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
    unsigned long long old = __hip_atomic_load(&P.vp[2 * idx], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT);
    if (bits & ~old) {
      old = atomicOr(&P.vp[2 * idx], bits);
      const unsigned long long fresh = bits & ~old;
      if (fresh) {
        if (old == 0) {
          uint32_t *b0 = P.bm0 + static_cast<uint64_t>(g) * P.nw0;
          const uint32_t o0 = atomicOr(&b0[u >> 5], 1u << (u & 31u));
          if (o0 == 0)
            atomicOr(&P.bm1[static_cast<uint64_t>(g) * P.nw1 + (u >> 10)], 1u << ((u >> 5) & 31u));
        }
        push = atomicOr(&P.vp[2 * idx + 1], fresh) == 0;
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

// grid_discover for K entries per lane (bits[k] == 0: no entry), stage by
// stage over the entries: the K reads of vis, then the K atomics on vis, the
// bitmaps and the pend words, so each lane keeps K independent requests in
// flight instead of one chain; one frontier reservation per wave for all.
// Entries of one lane may name the same (g, u): the atomics on vis settle
// which of them finds each colour fresh, as between lanes.
template <int K>
__device__ inline void grid_discover_k(const GridParams &P, const uint32_t (&g)[K],
                                       const uint32_t (&u)[K], const unsigned long long (&bits)[K]) {
  uint64_t idx[K];
  unsigned long long old[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    idx[k] = static_cast<uint64_t>(g[k]) * P.n + u[k];
    old[k] = ~0ull;
    if (bits[k])
      old[k] = __hip_atomic_load(&P.vp[2 * idx[k]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (bits[k] & ~old[k]) old[k] = atomicOr(&P.vp[2 * idx[k]], bits[k]);
    else old[k] = ~0ull;
  }
  unsigned long long fresh[K];
  uint32_t o0[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    fresh[k] = bits[k] & ~old[k];
    o0[k] = 1u;
    if (fresh[k] && old[k] == 0)
      o0[k] = atomicOr(&P.bm0[static_cast<uint64_t>(g[k]) * P.nw0 + (u[k] >> 5)], 1u << (u[k] & 31u));
  }
  // the pend atomics go out with the bitmap ones (neither waits on the
  // other); the summary bits, which need the bitmap's old word, follow
  unsigned long long pp[K];
#pragma unroll
  for (int k = 0; k < K; ++k) pp[k] = fresh[k] ? atomicOr(&P.vp[2 * idx[k] + 1], fresh[k]) : 1ull;
  bool push[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (o0[k] == 0)
      atomicOr(&P.bm1[static_cast<uint64_t>(g[k]) * P.nw1 + (u[k] >> 10)], 1u << ((u[k] >> 5) & 31u));
    push[k] = pp[k] == 0;
  }
  uint64_t m[K];
  uint32_t total = 0;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    m[k] = __ballot(push[k]);
    total += static_cast<uint32_t>(__popcll(m[k]));
  }
  if (total == 0) return;
  unsigned long long base = 0;
  if (lane_id() == 0) base = atomicAdd(P.next_cnt, static_cast<unsigned long long>(total));
  base = __shfl(base, 0, 64);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (push[k]) {
      const unsigned long long at = base + __popcll(m[k] & lanes_below());
      if (at < P.cap) P.next[at] = (static_cast<unsigned long long>(g[k]) << 32) | u[k];
    }
    base += __popcll(m[k]);
  }
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
  uint32_t g[kExpItems], beg[kExpItems], deg[kExpItems], T[kExpItems], fl[kExpItems];
  uint32_t uf[kExpItems], ul[kExpItems], pre[kExpItems], nch[kExpItems], dpre[kExpItems];
  unsigned long long q[kExpItems], rq[kExpItems], ja[kExpItems], jb[kExpItems], m[kExpItems];
  uint32_t ucur;   // next unit of the chunk not claimed by a wave
  // the dense round's tasks: entry, block, first element, prefix of colours,
  // and per task and position the colours live there
  uint32_t tent[kDenseSlots], tk[kDenseSlots], tpre[kDenseSlots];
  Affine tE[kDenseSlots];
  unsigned long long pos[kDenseSlots * kDenseSpan];
};

struct ExpandStats {
  unsigned long long cand = 0, draws = 0, exp = 0, bytes = 0, live = 0;
};

// Expands the frontier items cur[c0, c0 + cl), cl <= kExpItems: every thread of
// the workgroup calls it.  A chunk's rows are split into units of one colour
// and up to eight gap blocks (rows pending for few colours) or into wave
// tasks of one block for every colour (rows pending for many).
template <int BT>
__device__ void expand_chunk(const GridParams &P, const unsigned long long *cur, uint64_t c0,
                             uint32_t cl, ExpandLds<BT> &S, ExpandStats &st) {
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
    const unsigned long long M =
        atomicExch(&P.vp[2 * (static_cast<uint64_t>(g) * P.n + v) + 1], 0ull);
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
        if (__popcll(M) >= kDenseColours && (1u << beta) <= kDenseSpan) {
          ntask = nblk;   // a round task per block, a lane per (block, colour)
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
  // lengths of their units; the live edges of kSparseSteps steps are then
  // discovered together.
  const uint32_t lane = lane_id();
  {
    const Affine Pst = s_p2[0];
    uint32_t nxt = 0, hi = 0;   // the wave's claimed units [nxt, hi) (wave-uniform)
    bool has = false, started = false;
    uint32_t g = 0, e0 = 0, len = 0, pos = 0, T = 0, uskip = 0;
    unsigned long long bit = 0;
    uint64_t q = 0, rq = 0, x = 0;
    bool uniform = true;
    // kSparseSteps steps per lane, then their live edges discovered together
    for (bool finished = false; !finished;) {
      uint32_t kg[kSparseSteps], ku[kSparseSteps];
      unsigned long long kb[kSparseSteps];
#pragma unroll
      for (uint32_t j = 0; j < kSparseSteps; ++j) {
        kg[j] = 0;
        ku[j] = 0;
        kb[j] = 0;
        if (finished) continue;   // wave-uniform
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
              const uint32_t cj = r / nblk, k = r - cj * nblk;
              g = s_g[ent];
              const uint32_t c = nth_bit(s_m[ent], cj);
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
          if (nxt >= total) finished = true;
          continue;
        }
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
              const bool live = uniform || ic_accept(a, T, P.val[e]);
              ++pos;
              x = Pst.a * a + Pst.b;                       // next gap element
              if (pos < len) ++st_draws;
              else has = false;
              if (live) {
                kg[j] = g;
                ku[j] = P.ids[e];
                kb[j] = bit;
                ++st_live;
                st_bytes += 4;
              }
            }
          }
        }
      }
      grid_discover_k<kSparseSteps>(P, kg, ku, kb);
    }
  }
  // Rows pending for many colours (blocks of at most kDenseSpan edges): task
  // t is block k of entry ent, kDenseSlots tasks at a time.  Phase 1: a lane
  // per (task, colour) walks the colour's candidates of the block and ORs the
  // colour's bit into the LDS mask of each live position; phase 2: a lane per
  // position makes one discovery with all the colours live there.  The
  // atomics of a phase-2 step are independent of each other (many in flight
  // per wave), and colours reaching one edge together cost one.
  for (uint32_t r0 = 0; r0 < dtotal; r0 += kDenseSlots) {
    const uint32_t nt = min(kDenseSlots, dtotal - r0);
    // phase 0: the round's tasks, their colours and first elements
    uint32_t tu = 0;
    if (tid < nt) {
      const uint32_t t = r0 + tid;
      uint32_t lo = 0, hi = cl;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (s_dpre[mid] <= t) lo = mid + 1;
        else hi = mid;
      }
      const uint32_t ent = lo - 1;   // the last entry starting at or before t
      const uint32_t k = t - s_dpre[ent];
      const uint32_t beta = (s_fl[ent] >> kBetaShift) & 31u;
      Affine E{s_ja[ent], s_jb[ent]};
      for (uint32_t kk = k, b2 = beta + 1; kk; kk >>= 1, ++b2)
        if (kk & 1) E = aff_then(E, s_p2[b2]);
      S.tent[tid] = ent;
      S.tk[tid] = k;
      S.tE[tid] = E;
      tu = static_cast<uint32_t>(__popcll(s_m[ent]));
    }
    uint32_t units;
    const uint32_t tpre = block_exclusive_scan<BT>(tu, s_tmp, &units);
    if (tid < nt) S.tpre[tid] = tpre;
    __syncthreads();
    // phase 1
    for (uint32_t u = tid; u < units; u += BT) {
      uint32_t lo = 0, hi = nt;
      while (lo < hi) {
        const uint32_t mid = (lo + hi) >> 1;
        if (S.tpre[mid] <= u) lo = mid + 1;
        else hi = mid;
      }
      const uint32_t t = lo - 1;
      const uint32_t ent = S.tent[t], k = S.tk[t];
      const uint32_t c = nth_bit(s_m[ent], u - S.tpre[t]);
      const uint32_t g = s_g[ent];
      const uint32_t fl = s_fl[ent];
      const uint32_t beta = (fl >> kBetaShift) & 31u;
      const uint32_t B = 1u << beta;
      const uint32_t len = min(s_deg[ent] - (k << beta), B);
      const uint32_t e0 = s_beg[ent] + (k << beta);
      const uint32_t T = s_T[ent];
      const uint64_t q = s_q[ent], rq = s_rq[ent];
      const bool uniform = (fl & kRowUniform) != 0;
      const Affine Pst = s_p2[0];
      const Affine E = S.tE[t];
      uint64_t x = E.a * P.z[64ull * g + c] + E.b;
      ++st_draws;
      // a block without candidates costs one compare
      if ((x >> 32) + 1 <= (len == B ? s_uf[ent] : s_ul[ent])) continue;
      unsigned long long *pm = S.pos + t * kDenseSpan;
      // candidates kDenseBatch at a time: their positions first (arithmetic
      // only), then their thresholds read together, then the tests
      uint32_t pos = 0;
      for (bool more = true; more;) {
        uint32_t cp[kDenseBatch];
        uint64_t ca[kDenseBatch];
        uint32_t nc = 0;
#pragma unroll
        for (uint32_t j = 0; j < kDenseBatch; ++j) {
          cp[j] = 0;
          ca[j] = 0;
          if (more) {
            const uint64_t gap = ic_gap_fast(x, q, rq, &S.tab);
            if (gap >= len - pos) {
              more = false;
            } else {
              pos += static_cast<uint32_t>(gap);
              const uint64_t a = Pst.a * x + Pst.b;    // acceptance element
              ++st_cand;
              if (!uniform) {
                ++st_draws;
                st_bytes += 4;
              }
              cp[j] = pos;
              ca[j] = a;
              nc = j + 1;
              ++pos;
              if (pos >= len) {
                more = false;
              } else {
                x = Pst.a * a + Pst.b;                 // next gap element
                ++st_draws;
              }
            }
          }
        }
        uint32_t tv[kDenseBatch];
#pragma unroll
        for (uint32_t j = 0; j < kDenseBatch; ++j)
          tv[j] = (!uniform && j < nc) ? P.val[e0 + cp[j]] : 0u;
#pragma unroll
        for (uint32_t j = 0; j < kDenseBatch; ++j) {
          if (j < nc && (uniform || ic_accept(ca[j], T, tv[j]))) {
            atomicOr(&pm[cp[j]], 1ull << c);
            ++st_live;
          }
        }
      }
    }
    __syncthreads();
    // phase 2: every thread takes KP positions and discovers them together
    constexpr int KP = static_cast<int>(kDenseSlots * kDenseSpan / BT);
    static_assert(KP * BT == kDenseSlots * kDenseSpan, "positions split evenly");
    uint32_t dg[KP], du[KP];
    unsigned long long dm[KP];
#pragma unroll
    for (int j = 0; j < KP; ++j) {
      const uint32_t pi = static_cast<uint32_t>(j) * BT + tid;
      const uint32_t t = pi / kDenseSpan, pos = pi % kDenseSpan;
      dm[j] = 0;
      dg[j] = 0;
      du[j] = 0;
      if (t < nt) {
        dm[j] = S.pos[pi];
        if (dm[j]) {
          S.pos[pi] = 0;
          const uint32_t ent = S.tent[t];
          dg[j] = s_g[ent];
          du[j] = P.ids[s_beg[ent] + (S.tk[t] << ((s_fl[ent] >> kBetaShift) & 31u)) + pos];
          st_bytes += 4;
        }
      }
    }
    grid_discover_k<KP>(P, dg, du, dm);
    __syncthreads();
  }
}

// BFS level L: the frontier is buf[L % 2][0, n) with n = ring[L % kRing]
// (its first cap entries when it overflowed, which raises the overflow
// flag); every workgroup with a chunk to take pulls chunks of P.chunk items
// (at most kExpItems) from the level's counter.  Levels are enqueued ahead of
// the host: a level whose frontier is empty costs its launch only.
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
  const uint32_t chunk = P.chunk;
  if (static_cast<uint64_t>(blockIdx.x) * chunk >= ncur) return;   // no chunk for it
  GridParams Q = P;
  Q.cur = P.buf[L & 1u];
  Q.ncur = ncur;
  Q.next = P.buf[(L + 1) & 1u];
  Q.next_cnt = &P.ring[(L + 1) % kRing];
  unsigned long long *head = &P.ring[kRing + L % kRing];
  if (tid == 0) fill_step_powers(P.step, S.p2);
  ic_load_tables(&S.tab);
  for (uint32_t i = tid; i < kDenseSlots * kDenseSpan; i += BT) S.pos[i] = 0;  // phase 2 keeps it zero
  ExpandStats st;
  for (;;) {
    __syncthreads();
    if (tid == 0) s_chunk = atomicAdd(head, static_cast<unsigned long long>(chunk));
    __syncthreads();
    const uint64_t c0 = s_chunk;
    if (c0 >= ncur) break;
    expand_chunk<BT>(Q, Q.cur, c0,
                     static_cast<uint32_t>(min(static_cast<uint64_t>(chunk), ncur - c0)), S,
                     st);
  }
  report_stats(P.ctl, st.bytes, st.cand, st.draws, st.exp, st.live);
}

// Levels L0, L0 + 1, ... by one workgroup while the frontier holds between 1
// and tail_cap items: each level zeroes the next level's length, expands its
// items chunk by chunk and meets the next level at a barrier.  Stops at the
// first level whose frontier is empty or longer than tail_cap and leaves its
// number in gctl[kGTail]; that level's length stays in the ring for the
// level kernels.
template <int BT>
__global__ void __launch_bounds__(BT) grid_tail_kernel(GridParams P, uint32_t L0,
                                                       uint64_t tail_cap) {
  __shared__ ExpandLds<BT> S;
  __shared__ unsigned long long s_n;
  const uint32_t tid = threadIdx.x;
  if (tid == 0) fill_step_powers(P.step, S.p2);
  ic_load_tables(&S.tab);
  for (uint32_t i = tid; i < kDenseSlots * kDenseSpan; i += BT) S.pos[i] = 0;
  ExpandStats st;
  uint32_t L = L0;
  for (;; ++L) {
    __syncthreads();   // the previous level's appends are done
    if (tid == 0) {
      const unsigned long long raw = __hip_atomic_load(&P.ring[L % kRing], __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
      s_n = raw;
      if (raw != 0 && raw <= tail_cap) {
        __hip_atomic_store(&P.ring[(L + 1) % kRing], 0ull, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        P.gctl[kGLevels] += 1;
      }
    }
    __syncthreads();
    const uint64_t ncur = s_n;
    if (ncur == 0 || ncur > tail_cap) break;
    GridParams Q = P;
    Q.cur = P.buf[L & 1u];
    Q.ncur = ncur;
    Q.next = P.buf[(L + 1) & 1u];
    Q.next_cnt = &P.ring[(L + 1) % kRing];
    for (uint64_t c0 = 0; c0 < ncur; c0 += kExpItems)
      expand_chunk<BT>(Q, Q.cur, c0,
                       static_cast<uint32_t>(min(static_cast<uint64_t>(kExpItems), ncur - c0)),
                       S, st);
  }
  if (tid == 0) P.gctl[kGTail] = L;
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
    const bool push = i < total && __hip_atomic_load(&P.vp[2 * i + 1], __ATOMIC_RELAXED,
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


// ------------------------------------------------------------------ output
// After a batch's last level the groups' sets are written in four steps,
// each spread over many workgroups per group:
//   1. union lists: the group's bitmap is cut into kRanges ranges of summary
//      words; one kernel counts each range's members, the next writes them in
//      vertex order into the group's staging (list + colour masks), clearing
//      bitmaps and vis on the way;
//   2. per chunk of kChunk staging entries (one wave each): members of each
//      colour, all and hot ones;
//   3. per group: the column decision, each colour's offsets over the chunks,
//      one reservation of pool space, the sets' offsets and lengths;
//   4. per chunk: hot members into the group's column (when it has one), the
//      rest scattered to their colours' sorted lists.
constexpr uint32_t kRanges = 64;
constexpr uint32_t kChunk = 2048;   // staging entries per wave of steps 2 and 4

struct OutParams {
  uint32_t n, nw0, nw1;
  unsigned long long *vp;
  uint32_t *bm0, *bm1;
  uint32_t g0;                          // first group of this sub-batch
  uint32_t groups;                      // groups of the sub-batch
  const unsigned long long *soff;       // per group of the sub-batch: staging offset
  uint32_t *slist;                      // staging: union lists
  unsigned long long *smask;            // staging: union masks
  unsigned long long *rc;               // groups x kRanges: range member counts
  const unsigned long long *cbase;      // groups + 1: first chunk of each group
  const unsigned long long *usz;        // groups: union sizes
  uint32_t *cc;                         // chunks x 64: members per colour (then offsets)
  uint32_t *cch;                        // chunks x 64: hot members per colour
  unsigned long long *dbase;            // groups x 64: pool offset of each colour's list
  uint32_t *gcol;                       // groups: column of the group, or kNoGroup
  uint32_t *gfail;                      // groups: pool space ran out (nothing is written)
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

// Summary words [lo, hi) of range r of a group with nw1 of them.
__device__ inline void range_of(uint32_t nw1, uint32_t r, uint32_t *lo, uint32_t *hi) {
  *lo = static_cast<uint32_t>(static_cast<uint64_t>(nw1) * r / kRanges);
  *hi = static_cast<uint32_t>(static_cast<uint64_t>(nw1) * (r + 1) / kRanges);
}

// Step 1a: members of each (group, range).
template <int BT>
__global__ void __launch_bounds__(BT) out_range_count_kernel(OutParams P) {
  __shared__ unsigned long long s_w[BT / 64];
  const uint32_t i = blockIdx.x / kRanges, r = blockIdx.x % kRanges;
  const uint32_t g = P.g0 + i;
  const uint32_t *bm0 = P.bm0 + static_cast<uint64_t>(g) * P.nw0;
  const uint32_t *bm1 = P.bm1 + static_cast<uint64_t>(g) * P.nw1;
  uint32_t lo, hi;
  range_of(P.nw1, r, &lo, &hi);
  unsigned long long cnt = 0;
  for (uint32_t w1i = lo + threadIdx.x; w1i < hi; w1i += BT) {
    const uint32_t w1 = __hip_atomic_load(&bm1[w1i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (uint32_t b = w1; b; b &= b - 1)
      cnt += __popc(__hip_atomic_load(&bm0[w1i * 32 + (__ffs(b) - 1)], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT));
  }
  for (int d = 32; d > 0; d >>= 1) cnt += __shfl_down(cnt, d, 64);
  if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = cnt;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int w = 0; w < BT / 64; ++w) t += s_w[w];
    P.rc[blockIdx.x] = t;
  }
}

// Step 1b: the members of each (group, range) into the staging in vertex
// order; bitmaps and vis read and cleared with atomics (they were written by
// atomics, which complete in L2).
template <int BT>
__global__ void __launch_bounds__(BT) out_range_write_kernel(OutParams P) {
  __shared__ uint32_t s_tmp[BT / 64];
  __shared__ unsigned long long s_pre;
  const uint32_t i = blockIdx.x / kRanges, r = blockIdx.x % kRanges;
  const uint32_t g = P.g0 + i;
  const uint32_t tid = threadIdx.x;
  uint32_t *bm0 = P.bm0 + static_cast<uint64_t>(g) * P.nw0;
  uint32_t *bm1 = P.bm1 + static_cast<uint64_t>(g) * P.nw1;
  unsigned long long *vp = P.vp + 2 * static_cast<uint64_t>(g) * P.n;
  if (tid == 0) {
    unsigned long long t = 0;
    for (uint32_t q = 0; q < r; ++q) t += P.rc[i * kRanges + q];
    s_pre = t;
  }
  __syncthreads();
  uint32_t *list = P.slist + P.soff[i];
  unsigned long long *masks = P.smask + P.soff[i];
  uint32_t lo, hi;
  range_of(P.nw1, r, &lo, &hi);
  uint64_t running = s_pre;
  for (uint32_t t0 = lo; t0 < hi; t0 += BT) {
    const uint32_t w1i = t0 + tid;
    const uint32_t w1 = w1i < hi ? __hip_atomic_load(&bm1[w1i], __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)
                                 : 0u;
    uint32_t cnt = 0;
    for (uint32_t b = w1; b; b &= b - 1)
      cnt += __popc(__hip_atomic_load(&bm0[w1i * 32 + (__ffs(b) - 1)], __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT));
    uint32_t tot;
    uint64_t p = running + block_exclusive_scan<BT>(cnt, s_tmp, &tot);
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
}

// The group of chunk c (cbase is ascending, groups + 1 entries).
__device__ inline uint32_t group_of_chunk(const unsigned long long *cbase, uint32_t groups,
                                          uint64_t c) {
  uint32_t lo = 0, hi = groups;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (cbase[mid + 1] <= c) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Step 2: per chunk (one wave), members of each colour, all and hot.
__global__ void __launch_bounds__(64) out_chunk_count_kernel(OutParams P) {
  const uint64_t c = blockIdx.x;
  const uint32_t i = group_of_chunk(P.cbase, P.groups, c);
  const uint32_t lane = lane_id();
  const uint64_t e0 = (c - P.cbase[i]) * kChunk;
  const uint64_t e1 = min(e0 + kChunk, P.usz[i]);
  const uint32_t *list = P.slist + P.soff[i];
  const unsigned long long *masks = P.smask + P.soff[i];
  uint32_t all = 0, hot = 0;   // lane j: colour j's counts
  for (uint64_t t = e0; t < e1; t += 64) {
    const uint64_t e = t + lane;
    const unsigned long long m = e < e1 ? masks[e] : 0ull;
    const bool h = P.colptr && e < e1 && P.hot_index[list[e]] != kNotHot;
    for (unsigned long long u = uniform_u64(wave_or_u64(m)); u; u &= u - 1) {
      const uint32_t col = static_cast<uint32_t>(__ffsll(u) - 1);
      const uint64_t b = __ballot((m >> col) & 1ull);
      const uint64_t bh = __ballot(h && ((m >> col) & 1ull));
      if (lane == col) {
        all += static_cast<uint32_t>(__popcll(b));
        hot += static_cast<uint32_t>(__popcll(bh));
      }
    }
  }
  P.cc[c * 64 + lane] = all;
  P.cch[c * 64 + lane] = hot;
}

// Step 3: per group (64 threads, thread j = colour j): the column decision,
// colour j's list offsets over the group's chunks, one pool reservation.
__global__ void __launch_bounds__(64) out_group_kernel(OutParams P) {
  __shared__ unsigned long long s_hot, s_base;
  __shared__ uint32_t s_col, s_fail;
  __shared__ unsigned long long s_tot[64];
  const uint32_t i = blockIdx.x, g = P.g0 + i, j = threadIdx.x;
  const uint32_t nb = P.nb[g];
  const uint64_t c0 = P.cbase[i], c1 = P.cbase[i + 1];
  unsigned long long hot = 0;
  for (uint64_t c = c0; c < c1; ++c) hot += P.cch[c * 64 + j];
  for (int d = 32; d > 0; d >>= 1) hot += __shfl_xor(hot, d, 64);
  if (j == 0) {
    s_hot = hot;
    s_col = kNoGroup;
    // a column costs 8 bytes per hot vertex, lists 4 per member
    if (P.colptr && hot > 2ull * P.hot_n)
      s_col = P.group_base + static_cast<uint32_t>(atomicAdd(&P.ctl[kCtlGroupCounter2], 1ull));
  }
  __syncthreads();
  const bool col = s_col != kNoGroup;
  // colour j's offsets in its list, chunk by chunk
  unsigned long long run = 0;
  for (uint64_t c = c0; c < c1; ++c) {
    const uint32_t k = P.cc[c * 64 + j] - (col ? P.cch[c * 64 + j] : 0u);
    P.cc[c * 64 + j] = static_cast<uint32_t>(run);
    run += k;
  }
  s_tot[j] = run;
  __syncthreads();
  if (j == 0) {
    unsigned long long acc = 0;
    for (uint32_t q = 0; q < nb; ++q) acc += s_tot[q];
    const unsigned long long base = atomicAdd(&P.ctl[kCtlDataBump], acc);
    s_base = base;
    s_fail = 0;
    if (base + acc > P.cap_data) {
      s_fail = 1;
      atomicOr(&P.ctl[kCtlOverflow], 1ull);
    }
    if (col) atomicAdd(&P.ctl[kCtlElems], s_hot);
    P.gcol[i] = s_col;
    P.gfail[i] = s_fail;
  }
  __syncthreads();
  unsigned long long mine = s_base;
  for (uint32_t q = 0; q < j; ++q) mine += s_tot[q];
  P.dbase[i * 64ull + j] = mine;
  if (j < nb) {
    const uint32_t sid = P.sid[64ull * g + j];
    if (s_fail) {
      P.set_len[sid] = kLenFail;
    } else {
      P.set_off[sid] = mine;
      P.set_len[sid] = static_cast<uint32_t>(s_tot[j]);
    }
    if (col) P.gref[sid] = (s_col << 6) | j;
  }
  if (col) P.gsample[64ull * s_col + j] = j < nb ? P.sid[64ull * g + j] : kNoGroup;
  unsigned long long bytes = s_fail ? 0ull : 4ull * s_tot[j];
  report_stats(P.ctl, bytes, 0, 0, 0, 0);
}

// Step 3b: the columns of the sub-batch's column groups start at zero.
__global__ void out_column_clear_kernel(OutParams P) {
  const uint64_t total = static_cast<uint64_t>(P.groups) * P.hot_n;
  for (uint64_t t = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t i = static_cast<uint32_t>(t / P.hot_n);
    const uint32_t cg = P.gcol[i];
    if (cg != kNoGroup && !P.gfail[i]) P.colptr[cg][t - static_cast<uint64_t>(i) * P.hot_n] = 0ull;
  }
}

// Step 4: per chunk (one wave), hot members into the column (column
// groups), the rest to their colours' lists in vertex order.
__global__ void __launch_bounds__(64) out_scatter_kernel(OutParams P) {
  const uint64_t c = blockIdx.x;
  const uint32_t i = group_of_chunk(P.cbase, P.groups, c);
  if (P.gfail[i]) return;   // the call is redone with more pool space
  const uint32_t lane = lane_id();
  const uint32_t cg = P.gcol[i];
  const unsigned long long base = P.dbase[i * 64ull + lane];   // lane j: colour j
  const uint64_t e0 = (c - P.cbase[i]) * kChunk;
  const uint64_t e1 = min(e0 + kChunk, P.usz[i]);
  const uint32_t *list = P.slist + P.soff[i];
  const unsigned long long *masks = P.smask + P.soff[i];
  unsigned long long *colv = cg != kNoGroup ? P.colptr[cg] : nullptr;
  unsigned long long pos = base + P.cc[c * 64 + lane];   // lane j: next slot of colour j
  for (uint64_t t = e0; t < e1; t += 64) {
    const uint64_t e = t + lane;
    unsigned long long m = 0;
    uint32_t v = 0;
    if (e < e1) {
      m = masks[e];
      v = list[e];
      if (colv) {
        const uint32_t h = P.hot_index[v];
        if (h != kNotHot) {
          colv[h] = m;
          m = 0;
        }
      }
    }
    for (unsigned long long u = uniform_u64(wave_or_u64(m)); u; u &= u - 1) {
      const uint32_t col = static_cast<uint32_t>(__ffsll(u) - 1);
      const uint64_t b = __ballot((m >> col) & 1ull);
      const unsigned long long at = __shfl(pos, static_cast<int>(col), 64);
      if ((m >> col) & 1ull) P.data[at + __popcll(b & lanes_below())] = v;
      if (lane == col) pos += __popcll(b);
    }
  }
}

// Element 0 and sample id of every colour of the batch's groups: group g
// holds samples b0 + g C + c, c < C (C = colors); colours c >= C have none.
__global__ void grid_samples_kernel(uint64_t seed, uint64_t first, const uint32_t *ids,
                                    uint64_t b0, uint64_t count, uint32_t groups, uint32_t C,
                                    unsigned long long *z, uint32_t *sid, uint32_t *nb) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 64u * groups) return;
  const uint32_t g = i >> 6, c = i & 63u;
  const uint64_t t0 = b0 + static_cast<uint64_t>(g) * C;   // the group's first sample
  const uint64_t t = t0 + c;
  const bool has = c < C && t < count;
  const uint64_t s = has ? first + (ids ? ids[t] : t) : 0;
  z[i] = has ? sample_z(seed, s) : 0ull;
  sid[i] = has ? static_cast<uint32_t>(s) : kNoGroup;
  if (c == 0) {
    const uint64_t left = count > t0 ? count - t0 : 0;
    nb[g] = static_cast<uint32_t>(left < C ? left : C);
  }
}

// Bytes of one group's scratch: vis and pend, the union bitmap and summary.
uint64_t group_bytes(uint32_t n) {
  return 16ull * n + 4ull * ((n + 31) / 32) + 4ull * ((n + 1023) / 1024);
}

}  // namespace

// Groups per batch and frontier capacity for `count` samples.
static void grid_plan(const rip_graph *g, uint64_t count, uint32_t colors, uint32_t *groups,
                      uint64_t *fcap) {
  size_t free_b = 0, total_b = 0;
  RIP_HIP(hipMemGetInfo(&free_b, &total_b));
  const GridScratch &s = g->grid;
  const uint64_t held = s.n == g->n ? s.groups * group_bytes(g->n) + 16ull * s.fcap : 0;
  // the scratch stays within an eighth of the memory it may use, 32 GiB at most
  const uint64_t budget = std::min<uint64_t>((free_b + held) / 8, 32ull << 30);
  const uint64_t per = group_bytes(g->n) + 16ull * std::min<uint64_t>(g->n, 1u << 22);
  uint64_t G = std::max<uint64_t>(1, budget / per);
  G = std::min<uint64_t>({G, 256, (count + colors - 1) / colors});
  *groups = static_cast<uint32_t>(G);
  // room for every group's frontier up to 2^22 vertices, and for one whole
  // union (the output staging reuses the frontier buffers)
  *fcap = std::max<uint64_t>(g->n, G * std::min<uint64_t>(g->n, 1u << 22));
}

static void grid_ensure(rip_graph *g, uint32_t groups, uint64_t fcap, hipStream_t st) {
  GridScratch &s = g->grid;
  const uint32_t n = g->n;
  if (s.n != n || s.groups < groups) {
    dev_free(s.d_vis);
    dev_free(s.d_bm0);
    dev_free(s.d_bm1);
    dev_free(s.d_z);
    dev_free(s.d_sid);
    dev_free(s.d_nb);
    s.groups = 0;
    const uint64_t nw0 = (n + 31) / 32, nw1 = (n + 1023) / 1024;
    s.d_vis = dev_alloc<unsigned long long>(2ull * groups * n);
    s.d_bm0 = dev_alloc<uint32_t>(groups * nw0);
    s.d_bm1 = dev_alloc<uint32_t>(groups * nw1);
    s.d_z = dev_alloc<unsigned long long>(64ull * groups);
    s.d_sid = dev_alloc<uint32_t>(64ull * groups);
    s.d_nb = dev_alloc<uint32_t>(groups);
    // zero once; every batch leaves them zero again
    RIP_HIP(hipMemsetAsync(s.d_vis, 0, 16ull * groups * n, st));
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

// Room for `words` 64-bit words of output tables.
static void grid_out_arena(GridScratch &s, uint64_t words) {
  if (s.out_words >= words) return;
  dev_free(s.d_out);
  s.out_words = 0;
  s.d_out = dev_alloc<unsigned long long>(words);
  s.out_words = words;
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
  dev_free(s.d_out);
  s.out_words = 0;
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
                        bool columns, uint32_t frontier_cap, uint32_t colors) {
  if (count == 0) return 0;
  RIP_REQUIRE(colors >= 1 && colors <= 64, RIP_EINVAL, "tune: colors must be in [1, 64]");
  const uint32_t C = colors;
  rip_graph *gr = p->g;
  hipStream_t st = p->stream;
  uint32_t G = 0;
  uint64_t fcap = 0;
  grid_plan(gr, count, C, &G, &fcap);
  grid_ensure(gr, G, fcap, st);
  GridScratch &s = gr->grid;
  const uint64_t total_groups = (count + C - 1) / C;
  if (columns) pool_reserve_groups(p, p->groups + static_cast<uint32_t>(total_groups));
  const uint32_t n = gr->n;
  GridParams P;
  P.vrec = gr->d_vrec;
  P.ids = gr->d_ids;
  P.val = gr->d_val;
  P.n = n;
  P.step = gr->step;
  P.vp = s.d_vis;
  P.bm0 = s.d_bm0;
  P.bm1 = s.d_bm1;
  P.nw0 = (n + 31) / 32;
  P.nw1 = (n + 1023) / 1024;
  P.z = s.d_z;
  // entries a level may list: the buffer, or fewer when asked (the buffers
  // keep room for one whole union, the output staging)
  const uint64_t lcap = frontier_cap ? std::min<uint64_t>(frontier_cap, s.fcap) : s.fcap;
  P.cap = lcap;
  const uint64_t tail_cap = std::min<uint64_t>(kTailItems, lcap);
  P.gctl = s.d_gctl;
  P.ctl = p->d_ctl;
  P.ring = s.d_ring;
  P.buf[0] = s.d_fa;
  P.buf[1] = s.d_fb;
  const int per_cu = blocks_per_cu(reinterpret_cast<const void *>(&grid_level_kernel<kExpBT>),
                                   kExpBT);
  const uint32_t exp_grid = static_cast<uint32_t>(per_cu * cu_count());
  uint32_t launches = 0;
  for (uint64_t b0 = 0; b0 < count; b0 += static_cast<uint64_t>(C) * G) {
    const uint32_t Gb = static_cast<uint32_t>(std::min<uint64_t>(G, (count - b0 + C - 1) / C));
    // small batches on small graphs take their levels in chunks of 32 items:
    // a level's hub rows, discovered together, sit in few chunks, and with
    // chunks of 128 the workgroups holding them set the level's time.  On
    // R-MAT scale 16 IC (n Gb of 4M-25M), IMM's sampling calls of 4K-8K sets
    // took 10-11.6 ms with 128 and 6.3-8.6 ms with 32; with 32 on every batch,
    // scale 18 (100,000 sets) and scale 24 (20,000 sets) ran 4-5% slower.
    P.chunk = static_cast<uint64_t>(n) * Gb <= (1ull << 25) ? kExpItems / 4 : kExpItems;
    hipLaunchKernelGGL(grid_samples_kernel, dim3((64 * Gb + 255) / 256), dim3(256), 0, st,
                       p->seed, first, ids, b0, count, Gb, C, s.d_z, s.d_sid, s.d_nb);
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
    auto frontier_len = [&](uint32_t level) {
      RIP_HIP(hipMemcpyAsync(s.h_ctl + kGWords, s.d_ring + level % kRing,
                             sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
      RIP_HIP(hipStreamSynchronize(st));
      return static_cast<uint64_t>(s.h_ctl[kGWords]);
    };
    uint64_t pending = frontier_len(L);
    for (;;) {
      if (pending == 0) {
        if (!dirty) break;
        // the vertices the overflowing lists could not hold, as level L
        P.next = P.buf[L & 1u];
        P.next_cnt = &s.d_ring[L % kRing];
        hipLaunchKernelGGL(grid_rebuild_kernel, dim3(2048), dim3(256), 0, st, P, Gb);
        launches += 1;
        pending = frontier_len(L);
        dirty = pending > lcap;
        if (pending == 0) break;
        continue;
      }
      RIP_HIP(hipEventRecord(p->ev3, st));
      if (pending <= tail_cap) {
        // a thin frontier: one workgroup runs the levels until it widens
        hipLaunchKernelGGL((grid_tail_kernel<kExpBT>), dim3(1), dim3(kExpBT), 0, st, P, L,
                           tail_cap);
        launches += 1;
      } else {
        // kLevelsPerSync levels, then one look at the frontier length
        hipLaunchKernelGGL(grid_ring_kernel, dim3(1), dim3(kLevelsPerSync), 0, st, s.d_ring,
                           s.d_gctl, L + 1, kLevelsPerSync, 0u);
        for (uint32_t i = 0; i < kLevelsPerSync; ++i)
          hipLaunchKernelGGL((grid_level_kernel<kExpBT>), dim3(exp_grid), dim3(kExpBT), 0, st, P,
                             L + i);
        launches += 1 + kLevelsPerSync;
      }
      RIP_HIP(hipGetLastError());
      RIP_HIP(hipEventRecord(p->ev4, st));
      const unsigned long long *h = grid_read(s, s.d_gctl, kGWords, st);
      if (pending <= tail_cap) {
        L = static_cast<uint32_t>(h[kGTail]);
      } else {
        L += kLevelsPerSync;
        if (h[kGOver] != 0) dirty = true;
      }
      p->grid_levels = levels0 + static_cast<uint32_t>(h[kGLevels]);
      float ms = 0.0f;
      RIP_HIP(hipEventElapsedTime(&ms, p->ev3, p->ev4));
      p->grid_ms += ms;
      pending = frontier_len(L);
      if (pending > lcap) dirty = true;
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
    O.vp = s.d_vis;
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
      // groups a, a + 1, ... while their unions fit the staging together (one
      // always does); their staging offsets and chunks
      std::vector<unsigned long long> off, cb(1, 0);
      unsigned long long acc = 0;
      uint32_t e = a;
      while (e < Gb && (e == a || acc + usz[e] <= s.fcap)) {
        off.push_back(acc);
        acc += usz[e];
        cb.push_back(cb.back() + (usz[e] + kChunk - 1) / kChunk);
        ++e;
      }
      const uint32_t m = e - a;
      const uint64_t chunks = cb.back();
      // device tables of the sub-batch, carved from the output arena
      const uint64_t words = 4ull * m + kRanges * static_cast<uint64_t>(m) + 64ull * m +
                             64ull * chunks + 2;   // u64 words
      grid_out_arena(s, words);
      unsigned long long *w = s.d_out;
      unsigned long long *d_soff = w;          w += m;
      unsigned long long *d_cbase = w;         w += m + 1;
      unsigned long long *d_usz = w;           w += m;
      unsigned long long *d_rc = w;            w += static_cast<uint64_t>(kRanges) * m;
      unsigned long long *d_dbase = w;         w += 64ull * m;
      uint32_t *d_cc = reinterpret_cast<uint32_t *>(w);
      uint32_t *d_cch = d_cc + 64ull * chunks; w += 64ull * chunks;
      uint32_t *d_gcol = reinterpret_cast<uint32_t *>(w);
      uint32_t *d_gfail = d_gcol + m;          w += m;
      std::vector<unsigned long long> up(3ull * m + 1);
      std::copy(off.begin(), off.end(), up.begin());
      std::copy(cb.begin(), cb.end(), up.begin() + m);
      std::copy(usz.begin() + a, usz.begin() + e, up.begin() + 2 * m + 1);
      RIP_HIP(hipMemcpyAsync(s.d_out, up.data(), sizeof(unsigned long long) * up.size(),
                             hipMemcpyHostToDevice, st));
      O.g0 = a;
      O.groups = m;
      O.soff = d_soff;
      O.cbase = d_cbase;
      O.usz = d_usz;
      O.rc = d_rc;
      O.dbase = d_dbase;
      O.cc = d_cc;
      O.cch = d_cch;
      O.gcol = d_gcol;
      O.gfail = d_gfail;
      hipLaunchKernelGGL((out_range_count_kernel<256>), dim3(m * kRanges), dim3(256), 0, st, O);
      hipLaunchKernelGGL((out_range_write_kernel<256>), dim3(m * kRanges), dim3(256), 0, st, O);
      hipLaunchKernelGGL(out_chunk_count_kernel, dim3(static_cast<uint32_t>(chunks)), dim3(64), 0,
                         st, O);
      hipLaunchKernelGGL(out_group_kernel, dim3(m), dim3(64), 0, st, O);
      if (columns)
        hipLaunchKernelGGL(out_column_clear_kernel, dim3(4096), dim3(256), 0, st, O);
      hipLaunchKernelGGL(out_scatter_kernel, dim3(static_cast<uint32_t>(chunks)), dim3(64), 0, st,
                         O);
      RIP_HIP(hipGetLastError());
      RIP_HIP(hipStreamSynchronize(st));   // `up` and the staging are reused
      launches += 6;
      a = e;
    }
    launches += 1;
  }
  return launches;
}

}  // namespace rip
