// LT sampler: one reverse random walk per lane.
//
// Semantics of AddRRRSet under linear_threshold_tag
// (include/ripples/add_rrrset.h:136-150): at vertex v draw one threshold
// U in (0,1], subtract the in-edge weights in CSR order with float
// arithmetic, and take the first in-neighbour at which the threshold drops
// to <= 0; stop when no edge is taken or the taken vertex was already
// visited.  The threshold of (sample s, vertex v) is element 1 + begin(v)
// of sample s's stream, reached with the vertex record's jump; it is
// U = k * 2^-24 with k = (x >> 40) + 1.
//
// The float chain's crossing edge is a non-decreasing function of k, so the
// graph holds per-row breakpoints and a fanout-16 index over them (lt_scan.h,
// built once with the graph), and a step is one search for the first
// breakpoint above k -- a 64-byte node per index level and one block of
// breakpoints -- giving the same edge as the subtraction loop.  Rows holding negative, infinite or NaN
// weights have no breakpoints and are scanned with the float chain itself,
// UNROLL edges per round.
//
// Walks up to CAP vertices keep their visited list in LDS (column-major, so
// lanes hit distinct banks) and are sorted in place; the reference's GPU
// kernel (include/ripples/gpu/lt_kernel.h:59-130) caps walks at 8 vertices
// and finishes longer ones on the host.  Here longer walks are finished on
// the device by lt_long_kernel, which keeps a visited list and bitmap per
// lane in HBM and heap-sorts the list before writing it out.
#include <algorithm>

#include "device_util.h"
#include "internal.h"
#include "lt_scan.h"

namespace rip {
namespace {

constexpr uint32_t kLenFail = 0xFFFFFFFFu;
constexpr uint32_t kLenLong = 0xFFFFFFFEu;

struct LtParams {
  const LtRec *lrec;
  const uint32_t *ids;
  const uint32_t *val;
  const uint32_t *bp;
  const uint32_t *bidx;
  uint32_t n;
  uint64_t seed;
  uint64_t first;
  uint64_t count;
  unsigned long long *set_off;
  uint32_t *set_len;
  uint32_t *data;
  uint64_t cap_data;
  unsigned long long *ctl;
  uint32_t *long_ids;
  uint32_t *spill;
  uint32_t chunk;  // pool elements a wave reserves at a time
  uint32_t claim;  // sample ids a wave claims at a time, 64..kClaim
};

// Resident blocks of the walk kernel per CU (at most).
constexpr int kLtBlocksPerCu = 2;

// Most sample ids a wave claims from the work counter at a time; the batch
// (LtParams::claim) is at least a wave's worth, since one refill hands out
// up to 64 ids.
constexpr unsigned long long kClaim = 512;

// LT threshold of the walk at a vertex (ja, jb from its record): U = k * 2^-24.
__device__ inline uint32_t lt_threshold(uint64_t ja, uint64_t jb, uint64_t z) {
  return static_cast<uint32_t>((ja * z + jb) >> 40) + 1u;
}

// The first 32 bytes of an LT record: jump, pair offset, degree, index
// offset, flags.
struct LtHead {
  uint64_t ja, jb;
  uint32_t boff, d, ioff, flags;
};

__device__ inline LtHead lt_head(const LtRec *lrec, uint32_t v) {
  const uint4 *q = reinterpret_cast<const uint4 *>(lrec + v);
  const uint4 a = q[0], b = q[1];
  LtHead h;
  h.ja = (static_cast<uint64_t>(a.y) << 32) | a.x;
  h.jb = (static_cast<uint64_t>(a.w) << 32) | a.z;
  h.boff = b.x;
  h.d = b.y;
  h.ioff = b.z;
  h.flags = b.w;
  return h;
}

__device__ inline uint32_t count_le(uint4 q, uint32_t k) {
  return (q.x <= k ? 1u : 0u) + (q.y <= k ? 1u : 0u) + (q.z <= k ? 1u : 0u) + (q.w <= k ? 1u : 0u);
}

// Crossing edge of threshold k in the row of vertex v (d edges, first pair
// boff): the first j with B[j] > k, or d.  The top level of the row's
// fanout-16 index comes from v's record (the same line as its head); each
// level below it is one aligned 64-byte node of the index array; the last
// reads are the 16 breakpoints of one aligned 128-byte block and, in the same
// line, *id, the in-neighbour of the crossing edge when there is
// one (read from the block's line after the search).  *probes counts the
// nodes and blocks read (lt_index_cross of lt_scan.h is the same search over
// plain arrays).  Loading the top level with the head, and the ids with the
// breakpoints, measured no faster: the walk is bound by the lines in flight,
// not by the number of dependent loads.
__device__ inline uint32_t lt_cross(const LtRec *lrec, uint32_t v, const uint32_t *bp,
                                    const uint32_t *bidx, uint32_t boff, uint32_t d,
                                    uint32_t ioff, uint32_t k, uint32_t *probes, uint32_t *id) {
  // level l holds ceil(d / 16^(l+1)) entries; the top is the first level of
  // at most 16
  int nl = 0;
  while (d > kLtFan && ((static_cast<uint64_t>(d) + (1ull << (4 * nl + 4)) - 1) >> (4 * nl + 4)) >
                           kLtFan)
    ++nl;
  if (d > kLtFan) ++nl;
  uint32_t node = 0;
  if (nl > 0) {
    const uint4 *q = reinterpret_cast<const uint4 *>(lrec[v].top);
    const uint4 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3];
    const uint32_t p = count_le(q0, k) + count_le(q1, k) + count_le(q2, k) + count_le(q3, k);
    const uint32_t szt = static_cast<uint32_t>(
        (static_cast<uint64_t>(d) + (1ull << (4 * (nl - 1) + 4)) - 1) >> (4 * (nl - 1) + 4));
    ++*probes;
    if (p >= szt) return d;  // every breakpoint is <= k
    node = p;
    uint32_t start = ioff + lt_pad16(szt);  // levels below the top
    for (int l = nl - 2; l >= 0; --l) {
      const uint32_t sz = static_cast<uint32_t>(
          (static_cast<uint64_t>(d) + (1ull << (4 * l + 4)) - 1) >> (4 * l + 4));
      const uint4 *r = reinterpret_cast<const uint4 *>(bidx + start + 16u * node);
      const uint4 r0 = r[0], r1 = r[1], r2 = r[2], r3 = r[3];
      const uint32_t pl = count_le(r0, k) + count_le(r1, k) + count_le(r2, k) + count_le(r3, k);
      ++*probes;
      const uint32_t child = 16u * node + pl;
      if (child >= sz) return d;
      node = child;
      start += lt_pad16(sz);
    }
  }
  // block `node` of the row: 16 breakpoints, then their 16 in-neighbours,
  // in one aligned line; breakpoints rise along the row, so the crossing is
  // the number of them <= k
  const uint32_t lo = 16u * node;
  const uint32_t cnt = d - lo < 16u ? d - lo : 16u;
  const uint32_t *blk = bp + 2ull * (boff + lo);
  const uint4 *q = reinterpret_cast<const uint4 *>(blk);
  const uint4 b0 = q[0], b1 = q[1], b2 = q[2], b3 = q[3];
  const uint32_t c = count_le(b0, k) + count_le(b1, k) + count_le(b2, k) + count_le(b3, k);
  ++*probes;
  // padding breakpoints (0xFFFFFFFF) are above every threshold, so c <= cnt
  if (c < cnt) *id = blk[16 + c];
  return lo + c;
}

// A walk's visited filter: 512 bits in eight 64-bit registers (a 64-vertex
// walk meets a false positive about one step in nine, where 128 bits gave
// one in three and sent long walks to their list search most steps).  Words
// are picked with masks on named members: selects between them were turned
// into an indexed stack array by the compiler.
struct LtFilter {
  uint64_t w0, w1, w2, w3, w4, w5, w6, w7;
};

__device__ inline uint32_t lt_filter_bit(uint32_t v) { return (v * 0x9E3779B1u) >> 23; }

// all ones when word w is the one, else zero
__device__ inline uint64_t lt_word_mask(uint32_t q, uint32_t w) {
  return 0ull - static_cast<uint64_t>(q == w);
}

__device__ inline void lt_filter_add(LtFilter &f, uint32_t v) {
  const uint32_t b = lt_filter_bit(v), q = b >> 6;
  const uint64_t bit = 1ull << (b & 63u);
  f.w0 |= bit & lt_word_mask(q, 0);
  f.w1 |= bit & lt_word_mask(q, 1);
  f.w2 |= bit & lt_word_mask(q, 2);
  f.w3 |= bit & lt_word_mask(q, 3);
  f.w4 |= bit & lt_word_mask(q, 4);
  f.w5 |= bit & lt_word_mask(q, 5);
  f.w6 |= bit & lt_word_mask(q, 6);
  f.w7 |= bit & lt_word_mask(q, 7);
}

__device__ inline bool lt_filter_has(const LtFilter &f, uint32_t v) {
  const uint32_t b = lt_filter_bit(v), q = b >> 6;
  const uint64_t word = (f.w0 & lt_word_mask(q, 0)) | (f.w1 & lt_word_mask(q, 1)) |
                        (f.w2 & lt_word_mask(q, 2)) | (f.w3 & lt_word_mask(q, 3)) |
                        (f.w4 & lt_word_mask(q, 4)) | (f.w5 & lt_word_mask(q, 5)) |
                        (f.w6 & lt_word_mask(q, 6)) | (f.w7 & lt_word_mask(q, 7));
  return ((word >> (b & 63u)) & 1ull) != 0;
}

// x of lane (l ^ j) for j = 1, 2, 4 or 8, within each 16-lane row, from DPP
// moves (a few cycles each, where a general shuffle goes through the LDS
// crossbar): quad permutes give xor 1 and 2; xor 4 is the half-row mirror
// (l -> 7 - l within 8) followed by reversing each quad, and xor 8 the row
// mirror (l -> 15 - l) followed by the half-row mirror.
__device__ inline uint32_t xor16(uint32_t x, uint32_t j) {
  constexpr int kQuadSwap1 = 0xB1;    // quad_perm [1, 0, 3, 2]
  constexpr int kQuadSwap2 = 0x4E;    // quad_perm [2, 3, 0, 1]
  constexpr int kQuadReverse = 0x1B;  // quad_perm [3, 2, 1, 0]
  constexpr int kRowMirror = 0x140;
  constexpr int kRowHalfMirror = 0x141;
  const int v = static_cast<int>(x);
  int r;
  if (j == 1) {
    r = __builtin_amdgcn_mov_dpp(v, kQuadSwap1, 0xF, 0xF, false);
  } else if (j == 2) {
    r = __builtin_amdgcn_mov_dpp(v, kQuadSwap2, 0xF, 0xF, false);
  } else if (j == 4) {
    r = __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(v, kRowHalfMirror, 0xF, 0xF, false),
                                 kQuadReverse, 0xF, 0xF, false);
  } else {
    r = __builtin_amdgcn_mov_dpp(__builtin_amdgcn_mov_dpp(v, kRowMirror, 0xF, 0xF, false),
                                 kRowHalfMirror, 0xF, 0xF, false);
  }
  return static_cast<uint32_t>(r);
}

// Ascending bitonic sort of the first w keys of a wave (w = 32, 64 or 128,
// wave-uniform), held two per lane as a[lane], a[lane + 64]; for w < 128
// only x0 takes part, and for w == 32 lanes 32..63 sort a sequence of their
// own that the caller ignores.  Smaller w saves stages: 15, 21 or 28.
__device__ inline void wave_sort(uint32_t &x0, uint32_t &x1, uint32_t w) {
  const uint32_t lane = lane_id();
#pragma unroll 1
  for (uint32_t k = 2; k <= w; k <<= 1) {
#pragma unroll 1
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      if (j == 64) {
        // partners a[lane] and a[lane + 64] sit in one lane; k == 128 is ascending
        const uint32_t lo = min(x0, x1), hi = max(x0, x1);
        x0 = lo;
        x1 = hi;
      } else {
        const bool lower = (lane & j) == 0;
        const uint32_t p0 = __shfl_xor(x0, static_cast<int>(j), 64);
        const bool asc0 = (lane & k) == 0;
        x0 = (lower == asc0) ? min(x0, p0) : max(x0, p0);
        if (w == 128) {
          const uint32_t p1 = __shfl_xor(x1, static_cast<int>(j), 64);
          const bool asc1 = ((lane + 64) & k) == 0;
          x1 = (lower == asc1) ? min(x1, p1) : max(x1, p1);
        }
      }
    }
  }
}

// Persistent walkers: every lane runs its own walk and, when it ends, takes
// the next sample id from its wave's batch (claimed P.claim at a time from a
// device counter), so short walks never wait for the long ones of the wave.  Each round
// takes one step of every busy lane (a breakpoint search), or scans up to
// UNROLL in-edges of a row without breakpoints with 16-byte loads issued
// together.  A walk's visited list lives in LDS for its first CAP vertices
// and in a per-lane HBM spill area for the next SPILL, with a 512-bit filter
// in registers so the list is searched only when the filter has the vertex;
// longer walks are handed to lt_long_kernel.  Finished lists up to 16 are
// sorted four at a time in 16-lane register groups; longer ones are copied
// by the whole wave and sorted by lt_sort_kernel.
// The compiler's own register budget (about 100 registers, four waves per
// SIMD, no scratch) is kept: forcing six to eight waves per SIMD measured no
// faster, and eight waves spilled, on R-MAT scale 20.
template <int BT, int CAP, int SPILL, int UNROLL>
__global__ void __launch_bounds__(BT) lt_walk_kernel(LtParams P) {
  static_assert(UNROLL % 4 == 0, "groups of four weights per load");
  static_assert(CAP >= 16 && CAP + SPILL <= 128, "16-lane group sort and 128-key wave sort");
  constexpr uint32_t kSort = 16;  // lists this short are sorted in registers
  __shared__ uint32_t s_list[CAP * BT];
  // s_pow[r] = M^(2r): from the stream start of a batch's first sample id to
  // that of the id r places later (element 0 of sample s is M^(2s+2)(seed))
  __shared__ Affine s_pow[kClaim];
  for (uint32_t r = threadIdx.x; r < P.claim; r += BT) s_pow[r] = aff_pow(master_step(), 2ull * r);
  __syncthreads();
  uint32_t *list = s_list + threadIdx.x;  // element i < CAP at list[i * BT]
  uint32_t *spill = P.spill + (static_cast<uint64_t>(blockIdx.x) * BT + threadIdx.x) * SPILL;
  // Whether u is among the walk's first len vertices.  The LDS part and the
  // spill are searched with their own loads, four at a time (one access path
  // choosing between them per element compiles to dependent flat loads).
  auto listed = [&](uint32_t u, uint32_t len) -> bool {
    const uint32_t nl = len < static_cast<uint32_t>(CAP) ? len : static_cast<uint32_t>(CAP);
    bool hit = false;
    for (uint32_t i = 0; i < nl && !hit; i += 4) {
      uint32_t x[4];
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) x[q] = i + q < nl ? list[(i + q) * BT] : ~u;
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) hit = hit || x[q] == u;
    }
    for (uint32_t i = CAP; i < len && !hit; i += 4) {
      uint32_t x[4];
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) x[q] = i + q < len ? spill[i + q - CAP] : ~u;
#pragma unroll
      for (uint32_t q = 0; q < 4; ++q) hit = hit || x[q] == u;
    }
    return hit;
  };
  const uint4 *val4 = reinterpret_cast<const uint4 *>(P.val);
  const uint32_t lane = lane_id();
  bool busy = false, exhausted = false;
  uint64_t gid = 0, z = 0;
  // sample ids claimed by this wave and not yet handed out (wave-uniform);
  // ids are handed out in increasing order, so a wave only stops once every
  // id it still holds is past `count`
  uint64_t claim_lo = 0, claim_hi = 0;
  // the batch [claim_base, claim_hi) and the stream start of its first id
  uint64_t claim_base = 0, z_base = 0;
  uint32_t len = 0, e = 0, end = 0, rflags = 0, kthr = 0, ridx = 0;
  LtFilter filt{};  // visited filter
  // the wave's reserved output [out_cur, out_end) (wave-uniform): finished
  // walks are written there, and a new chunk is reserved only when they do
  // not fit, so waves rarely meet on the pool's one allocation counter
  unsigned long long out_cur = 0, out_end = 0;

  float t = 0.0f;                 // float chain of rows without breakpoints
  // work counters: in-edges the subtraction loop scans, index nodes and
  // breakpoint blocks read, vertices written out (every visited vertex is)
  uint32_t edges = 0, probes = 0, outs = 0;
  // busy-lane rounds (walk steps) and loop rounds of the wave
  uint32_t steps = 0, rounds = 0;

  // e, end: the row's pairs for a breakpoint search, its transposed-CSR
  // range for rows scanned by weight
  uint32_t vcur = 0;  // the walk's current vertex
  auto enter = [&](uint32_t v) {
    const LtHead h = lt_head(P.lrec, v);
    kthr = lt_threshold(h.ja, h.jb, z);
    t = static_cast<float>(kthr) * 0x1p-24f;
    e = (h.flags & kRowScalar) ? P.lrec[v].begin : h.boff;
    end = e + h.d;
    rflags = h.flags;
    ridx = h.ioff;
    vcur = v;
  };

  for (;;) {
    // ---- refill idle lanes from the wave's batch of claimed sample ids
    const bool want = !busy && !exhausted;
    const uint64_t wm = __ballot(want);
    if (wm) {
      const uint32_t need = __popcll(wm);
      const uint32_t rank = __popcll(wm & lanes_below());
      const uint64_t have = claim_hi - claim_lo;
      uint64_t id, base = claim_base, zb = z_base;
      if (have >= need) {
        id = claim_lo + rank;
        claim_lo += need;
      } else {
        // a new batch; the rest of the current one is handed out first
        const uint32_t leader = __ffsll(static_cast<unsigned long long>(wm)) - 1;
        unsigned long long nb = 0;
        if (lane == leader)
          nb = atomicAdd(&P.ctl[kCtlGroupCounter], static_cast<unsigned long long>(P.claim));
        nb = __shfl(nb, static_cast<int>(leader), 64);
        const uint64_t znb = sample_z(P.seed, P.first + nb);  // once per batch
        id = rank < have ? claim_lo + rank : nb + (rank - have);
        if (rank >= have) {
          base = nb;
          zb = znb;
        }
        claim_lo = nb + (need - have);
        claim_hi = nb + P.claim;
        claim_base = nb;
        z_base = znb;
      }
      if (want) {
        gid = id;
        if (gid < P.count) {
          z = aff_apply(s_pow[gid - base], zb);  // = sample_z(P.seed, P.first + gid)
          const uint32_t root = uniform_index(z, P.n);
          list[0] = root;
          len = 1;
          filt = LtFilter{};
          lt_filter_add(filt, root);
          enter(root);
          busy = true;
        } else {
          exhausted = true;
        }
      }
    }
    if (__ballot(busy) == 0) break;
    ++rounds;

    // ---- one step (or one scan round of a row without breakpoints)
    bool done = false, is_long = false;
    if (busy) {
      ++steps;
      bool found = false;
      uint32_t u = 0;  // the chosen in-neighbour
      if (e >= end) {
        done = true;  // no in-edge crosses the threshold: the walk stops
      } else if (!(rflags & kRowScalar)) {
        const uint32_t d = end - e;  // e is the row's first pair
        const uint32_t c = lt_cross(P.lrec, vcur, P.bp, P.bidx, e, d, ridx, kthr, &probes, &u);
        edges += c < d ? c + 1 : d;
        if (c < d) {
          found = true;
          e += c + 1;
        } else {
          done = true;
        }
      } else {
        const uint32_t base = e & ~3u;  // 16-byte aligned group of 4 weights
        const uint32_t last_quad = (end - 1) >> 2;
        uint4 q[UNROLL / 4];
#pragma unroll
        for (int kk = 0; kk < UNROLL / 4; ++kk)
          q[kk] = val4[min((base >> 2) + kk, last_quad)];
        const uint32_t lo = e - base;  // 0..3
        const uint32_t hi = min(static_cast<uint32_t>(UNROLL), end - base);
        uint32_t stop = hi;
#pragma unroll
        for (int k = 0; k < UNROLL; ++k) {
          const uint4 &g4 = q[k >> 2];
          const uint32_t wbits = (k & 3) == 0 ? g4.x : (k & 3) == 1 ? g4.y
                                 : (k & 3) == 2 ? g4.z : g4.w;
          if (!found && static_cast<uint32_t>(k) >= lo && static_cast<uint32_t>(k) < hi) {
            t -= __uint_as_float(wbits);
            if (!(t > 0)) {
              found = true;
              stop = k + 1;
            }
          }
        }
        edges += stop - lo;
        e = base + stop;
        if (!found && e >= end) done = true;
        if (found) u = P.ids[e - 1];
      }
      if (found) {
        const bool seen = lt_filter_has(filt, u) && listed(u, len);
        if (seen) {
          done = true;
        } else if (len == CAP + SPILL) {
          done = true;
          is_long = true;
        } else {
          if (len < static_cast<uint32_t>(CAP))
            list[len * BT] = u;
          else
            spill[len - CAP] = u;
          ++len;
          lt_filter_add(filt, u);
          enter(u);
        }
      }
    }

    // ---- retire finished walks: reserve pool space per wave, sort, write
    if (__ballot(done)) {
      const bool emit = done && !is_long;
      const uint32_t size = emit ? len : 0;
      const uint32_t incl = wave_inclusive_scan(size);
      const uint32_t total = __shfl(incl, 63, 64);
      unsigned long long base = 0;
      if (total) {
        if (out_end - out_cur < total) {
          // the rest of the old chunk stays a gap in the pool's data
          const unsigned long long want = total > P.chunk ? total : P.chunk;
          unsigned long long nb = 0;
          if (lane == 0) nb = atomicAdd(&P.ctl[kCtlDataBump], want);
          nb = __shfl(nb, 0, 64);
          out_cur = nb;
          out_end = nb + want;
        }
        base = out_cur;
        out_cur += total;
      }
      const bool fail = total && base + total > P.cap_data;
      if (fail && lane == 0) atomicOr(&P.ctl[kCtlOverflow], 1ull);
      const unsigned long long off = base + incl - size;
      if (done) {
        const uint64_t s = P.first + gid;
        if (is_long) {
          const unsigned long long idx = atomicAdd(&P.ctl[kCtlLongCount], 1ull);
          P.long_ids[idx] = static_cast<uint32_t>(gid);
          P.set_len[s] = kLenLong;
        } else if (fail) {
          P.set_len[s] = kLenFail;
        } else {
          if (len > kSort) {
            // written unsorted below; lt_sort_kernel sorts it in place
            const unsigned long long idx = atomicAdd(&P.ctl[kCtlSortCount], 1ull);
            P.long_ids[P.count - 1 - idx] = static_cast<uint32_t>(gid);
          }
          P.set_off[s] = off;
          P.set_len[s] = len;
          outs += len;
        }
      }
      const uint32_t tid0 = threadIdx.x & ~63u;
      // lists of at most kSort ids: four at a time, one per 16-lane group,
      // sorted in registers and written in order
      uint64_t sm = __ballot(emit && !fail && len <= kSort);
      const uint32_t g = lane >> 4, gi = lane & 15u;
      while (sm) {
        uint32_t pick = 64;  // this group's lane, 64 = none
#pragma unroll
        for (uint32_t q = 0; q < 4; ++q) {
          const uint32_t l = sm ? static_cast<uint32_t>(__ffsll(static_cast<unsigned long long>(sm)) - 1) : 64u;
          if (sm) sm &= sm - 1;
          if (q == g) pick = l;
        }
        const uint32_t src = pick < 64 ? pick : 0u;
        const uint32_t n_g = static_cast<uint32_t>(__shfl(static_cast<int>(len), static_cast<int>(src), 64));
        const unsigned long long off_g = __shfl(off, static_cast<int>(src), 64);
        const bool have = pick < 64 && gi < n_g;
        uint32_t x = have ? s_list[gi * BT + tid0 + pick] : 0xFFFFFFFFu;
#pragma unroll
        for (uint32_t k = 2; k <= 16; k <<= 1) {
#pragma unroll
          for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            const uint32_t y = xor16(x, j);
            const bool lower = (gi & j) == 0, asc = (gi & k) == 0;
            x = (lower == asc) ? min(x, y) : max(x, y);
          }
        }
        if (have) P.data[off_g + gi] = x;
      }
      // longer lists: copied by the whole wave, sorted by lt_sort_kernel
      uint64_t lm = __ballot(emit && !fail && len > kSort);
      while (lm) {
        const uint32_t L = __ffsll(static_cast<unsigned long long>(lm)) - 1;
        lm &= lm - 1;
        const uint32_t n_l = static_cast<uint32_t>(__shfl(static_cast<int>(len), static_cast<int>(L), 64));
        const unsigned long long off_l = __shfl(off, static_cast<int>(L), 64);
        const uint32_t *spill_l = P.spill + (static_cast<uint64_t>(blockIdx.x) * BT + tid0 + L) * SPILL;
        for (uint32_t i = lane; i < n_l; i += 64)
          P.data[off_l + i] = i < static_cast<uint32_t>(CAP) ? s_list[i * BT + tid0 + L] : spill_l[i - CAP];
      }
      if (done) {
        busy = false;
        len = 0;
      }
    }
  }
  unsigned long long e64 = edges, p64 = probes, o64 = outs, s64 = steps;
  for (int d = 32; d > 0; d >>= 1) {
    e64 += __shfl_down(e64, d, 64);
    p64 += __shfl_down(p64, d, 64);
    o64 += __shfl_down(o64, d, 64);
    s64 += __shfl_down(s64, d, 64);
  }
  if (lane == 0 && s64) atomicAdd(&P.ctl[kCtlStatExpansions], s64);
  if (lane == 0 && rounds) atomicAdd(&P.ctl[kCtlStatLevels], static_cast<unsigned long long>(rounds));
  // algorithmic bytes of the subtraction loop: 8 per visited vertex (its CSR
  // offsets), 8 per scanned in-edge (id + weight), 4 per output id
  if (lane == 0 && (e64 || o64)) atomicAdd(&P.ctl[kCtlStatBytes], 12ull * o64 + 8ull * e64);
  if (lane == 0 && e64) atomicAdd(&P.ctl[kCtlStatEdges], e64);
  if (lane == 0 && p64) atomicAdd(&P.ctl[kCtlStatCoins], p64);
}

template <int BT, int CAP, int SPILL, int UNROLL>
uint32_t lt_grid(uint64_t count) {
  int per_cu =
      blocks_per_cu(reinterpret_cast<const void *>(&lt_walk_kernel<BT, CAP, SPILL, UNROLL>), BT);
  // Walks are bound by the lines in flight, not by walker count: on R-MAT
  // scale 20 two blocks per CU (eight waves) measured 0.93 ms per launch, one
  // block 1.03 ms and the four that fit 1.05 ms.
  per_cu = std::min(per_cu, kLtBlocksPerCu);
  const uint64_t resident = static_cast<uint64_t>(per_cu) * static_cast<uint64_t>(cu_count());
  const uint64_t needed = (count + BT - 1) / BT;
  return static_cast<uint32_t>(std::max<uint64_t>(1, std::min(resident, needed)));
}

// Unbounded walks: one lane per walk, visited list + bitmap in HBM.
__device__ inline void sift_down(uint32_t *a, uint32_t root, uint32_t end) {
  for (;;) {
    uint32_t child = 2 * root + 1;
    if (child >= end) return;
    if (child + 1 < end && a[child] < a[child + 1]) ++child;
    if (a[root] >= a[child]) return;
    const uint32_t t = a[root];
    a[root] = a[child];
    a[child] = t;
    root = child;
  }
}

__global__ void lt_long_kernel(LtParams P, uint32_t num_long, uint32_t lo,
                               uint32_t lanes, uint32_t *lists,
                               uint32_t *bitmaps) {
  const uint32_t lane_g = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t idx = lo + lane_g;
  if (lane_g >= lanes || idx >= num_long) return;
  const uint64_t nw = (P.n + 31) / 32;
  uint32_t *list = lists + static_cast<uint64_t>(lane_g) * P.n;
  uint32_t *bm = bitmaps + static_cast<uint64_t>(lane_g) * nw;
  const uint64_t s = P.first + P.long_ids[idx];
  const uint64_t z = sample_z(P.seed, s);
  uint32_t v = uniform_index(z, P.n);
  uint32_t len = 0;
  uint32_t probes = 0;
  list[len++] = v;
  bm[v >> 5] |= 1u << (v & 31);
  for (;;) {
    const LtHead h = lt_head(P.lrec, v);
    const uint32_t k = lt_threshold(h.ja, h.jb, z);
    uint32_t c = h.d, u = 0;  // crossing edge of the row, d when none
    if (!(h.flags & kRowScalar)) {
      c = lt_cross(P.lrec, v, P.bp, P.bidx, h.boff, h.d, h.ioff, k, &probes, &u);
    } else {
      const uint32_t begin = P.lrec[v].begin;
      float t = static_cast<float>(k) * 0x1p-24f;
      for (uint32_t i = 0; i < h.d; ++i) {
        t -= __uint_as_float(P.val[begin + i]);
        if (!(t > 0)) {
          c = i;
          u = P.ids[begin + i];
          break;
        }
      }
    }
    if (c >= h.d) break;
    if ((bm[u >> 5] >> (u & 31)) & 1u) break;
    list[len++] = u;
    bm[u >> 5] |= 1u << (u & 31);
    v = u;
  }
  for (uint32_t i = 0; i < len; ++i) bm[list[i] >> 5] = 0;
  for (uint32_t i = len / 2; i-- > 0;) sift_down(list, i, len);
  for (uint32_t end = len; end-- > 1;) {
    const uint32_t t = list[0];
    list[0] = list[end];
    list[end] = t;
    sift_down(list, 0, end);
  }
  const unsigned long long base =
      atomicAdd(&P.ctl[kCtlDataBump], static_cast<unsigned long long>(len));
  if (base + len > P.cap_data) {
    atomicOr(&P.ctl[kCtlOverflow], 1ull);
    P.set_len[s] = kLenFail;
    return;
  }
  for (uint32_t i = 0; i < len; ++i) P.data[base + i] = list[i];
  P.set_off[s] = base;
  P.set_len[s] = len;
}

// Sorts the long lists the walk kernel recorded, one wave per list, in
// registers (lists hold at most 128 ids).  Waves stride over the recorded
// lists, whose count the walk kernel left in the control block.
__global__ void lt_sort_kernel(LtParams P) {
  const uint32_t lane = lane_id();
  const uint64_t nsort = P.ctl[kCtlSortCount];
  const uint64_t waves = static_cast<uint64_t>(gridDim.x) * (blockDim.x >> 6);
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * (blockDim.x >> 6) + (threadIdx.x >> 6);
       i < nsort; i += waves) {
    const uint64_t s = P.first + P.long_ids[P.count - 1 - i];
    const unsigned long long off = P.set_off[s];
    const uint32_t len = P.set_len[s];
    uint32_t x0 = 0xFFFFFFFFu, x1 = 0xFFFFFFFFu;
    if (lane < len) x0 = P.data[off + lane];
    if (lane + 64 < len) x1 = P.data[off + lane + 64];
    wave_sort(x0, x1, len <= 32 ? 32u : len <= 64 ? 64u : 128u);
    if (lane < len) P.data[off + lane] = x0;
    if (lane + 64 < len) P.data[off + lane + 64] = x1;
  }
}

__global__ void reset_long_kernel(unsigned long long *ctl) {
  ctl[kCtlLongCount] = 0;
  ctl[kCtlSortCount] = 0;
}

}  // namespace

// 64 list entries per lane in LDS (64 KiB per block, two blocks per CU):
// long walks fill their filter, so their revisit checks search the
// list most steps, and entries past the LDS part are HBM reads (16 entries
// in LDS measured 0.93 ms per launch on R-MAT scale 20, 64 entries 0.90 ms).
constexpr int kLtBT = 256, kLtCap = 64, kLtSpill = 64, kLtUnroll = 16;

// Ids per claim: an even share of the launch per wave, so every wave gets
// one batch, rounded up to whole waves of lanes.
uint32_t lt_claim(uint64_t count, uint32_t blocks) {
  const uint64_t waves = static_cast<uint64_t>(blocks) * (kLtBT / 64);
  const uint64_t share = (count + waves - 1) / waves;
  return static_cast<uint32_t>(std::min<uint64_t>(std::max<uint64_t>((share + 63) / 64 * 64, 64),
                                                  kClaim));
}

// Output elements a wave reserves at a time.  On R-MAT scale 20 (657,104
// walks of 5.3 vertices) chunks of 1, 256 and 512 elements measured 1.69,
// 1.18 and 1.16 ms per launch.
uint32_t lt_chunk(const rip_pool *p, uint64_t count, const rip_tune *tune, uint32_t blocks) {
  if (tune && tune->out_chunk) return tune->out_chunk;
  // only waves that get a batch of ids write anything
  const uint32_t claim = lt_claim(count, blocks);
  const uint64_t waves = std::max<uint64_t>(
      1, std::min<uint64_t>(static_cast<uint64_t>(blocks) * (kLtBT / 64),
                            (count + claim - 1) / claim));
  return out_chunk_rule(static_cast<double>(count) * p->avg_set_size / static_cast<double>(waves),
                        2048);
}

uint64_t lt_output_slack(const rip_pool *p, uint64_t count, const rip_tune *tune) {
  const uint32_t blocks = lt_grid<kLtBT, kLtCap, kLtSpill, kLtUnroll>(count);
  return static_cast<uint64_t>(blocks) * (kLtBT / 64) * lt_chunk(p, count, tune, blocks);
}

// Kernel parameters shared by the walk, sort and long-walk kernels.
LtParams lt_params(rip_pool *p, uint64_t first, uint64_t count) {
  const rip_graph *g = p->g;
  LtParams P{};
  P.lrec = g->d_lrec;
  P.ids = g->d_ids;
  P.val = g->d_val;
  P.bp = g->d_bp;
  P.bidx = g->d_bidx;
  P.n = g->n;
  P.seed = p->seed;
  P.first = first;
  P.count = count;
  P.set_off = p->d_off;
  P.set_len = p->d_len;
  P.data = p->d_data;
  P.cap_data = p->cap_data;
  P.ctl = p->d_ctl;
  P.long_ids = g->lt.d_long_ids;
  P.spill = g->lt.d_spill;
  return P;
}

// Enqueues the walk and sort kernels of samples [first, first+count); no
// host synchronisation.  Walks longer than the lists are counted in the
// pool's control words and finished by lt_finish_long.
uint32_t lt_launch(rip_pool *p, uint64_t first, uint64_t count,
                   const rip_tune *tune) {
  constexpr int BT = kLtBT, CAP = kLtCap, SPILL = kLtSpill, UNROLL = kLtUnroll;
  // long-walk ids: at most `count`
  if (p->g->lt.cap_ids < count) {
    dev_free(p->g->lt.d_long_ids);
    p->g->lt.d_long_ids = dev_alloc<uint32_t>(count);
    p->g->lt.cap_ids = count;
  }
  hipLaunchKernelGGL(reset_long_kernel, dim3(1), dim3(1), 0, p->stream, p->d_ctl);
  const uint32_t blocks = lt_grid<BT, CAP, SPILL, UNROLL>(count);
  const uint64_t spill_words = static_cast<uint64_t>(blocks) * BT * SPILL;
  if (p->g->lt.cap_spill < spill_words) {
    dev_free(p->g->lt.d_spill);
    p->g->lt.d_spill = dev_alloc<uint32_t>(spill_words);
    p->g->lt.cap_spill = spill_words;
  }
  LtParams P = lt_params(p, first, count);
  P.chunk = lt_chunk(p, count, tune, blocks);
  P.claim = lt_claim(count, blocks);
  RIP_HIP(hipEventRecord(p->ev0, p->stream));
  hipLaunchKernelGGL((lt_walk_kernel<BT, CAP, SPILL, UNROLL>), dim3(blocks), dim3(BT), 0,
                     p->stream, P);
  RIP_HIP(hipGetLastError());
  RIP_HIP(hipEventRecord(p->ev1, p->stream));  // the events time the walk kernel alone
  hipLaunchKernelGGL(lt_sort_kernel, dim3(1024), dim3(256), 0, p->stream, P);
  RIP_HIP(hipGetLastError());
  return 2;
}

// Walks that outgrew the walk kernel's 128-entry lists (nlong of them, from
// the pool's control words after lt_launch): one lane per walk with an HBM
// list and bitmap each.  Returns the launches made.
uint32_t lt_finish_long(rip_pool *p, uint64_t first, uint64_t count, uint64_t nlong) {
  const rip_graph *g = p->g;
  const LtParams P = lt_params(p, first, count);
  // lanes of unbounded walks: a visited list (n words) and bitmap each
  const uint64_t per_lane = 4ull * g->n + 4ull * ((g->n + 31) / 32);
  const uint64_t lanes = std::max<uint64_t>(
      1, std::min<uint64_t>({nlong, 256, (static_cast<uint64_t>(1) << 29) / per_lane}));
  uint32_t *lists = dev_alloc<uint32_t>(lanes * g->n);
  uint32_t *bms = dev_alloc<uint32_t>(lanes * ((g->n + 31) / 32));
  RIP_HIP(hipMemsetAsync(bms, 0, sizeof(uint32_t) * lanes * ((g->n + 31) / 32), p->stream));
  uint32_t launches = 0;
  for (uint64_t lo = 0; lo < nlong; lo += lanes) {
    const uint32_t blocks_l = static_cast<uint32_t>((lanes + 63) / 64);
    hipLaunchKernelGGL(lt_long_kernel, dim3(blocks_l), dim3(64), 0, p->stream, P,
                       static_cast<uint32_t>(nlong), static_cast<uint32_t>(lo),
                       static_cast<uint32_t>(lanes), lists, bms);
    RIP_HIP(hipGetLastError());
    ++launches;
  }
  RIP_HIP(hipStreamSynchronize(p->stream));
  dev_free(lists);
  dev_free(bms);
  return launches;
}

}  // namespace rip
