// Shared definitions of libripples_amd: the lcg64 affine algebra used by the
// per-sample stream contract, error plumbing of the C ABI, and HIP helpers.
//
// Stream contract (identical on host and device):
//   master lcg64 r <- a*r + b mod 2^64, a = 18145460002477866997, b = 1,
//   seeded with r = seed (tools/imm.cc:172-174, 212-214 seed it with 0).
//   The sampling stream is master.split(2, 1); sample s owns the leapfrog
//   sub-stream sampling.split(S, s) with S = 2^32 - 1.  Element j of sample s
//   is master output 1 + 2(s + jS):
//     j = 0               -> root = (hi32(x) * n) >> 32
//     j = 1 + 2 off[v] + i -> IC draws of v's in-edge row: gaps between
//                           candidate edges and their acceptances (ic_gap.h)
//     j = 1 + off[v]      -> LT threshold at vertex v
//   with u = ((x >> 40) + 1) * 2^-24 in (0, 1] for LT thresholds.  An IC edge
//   of weight w is live with probability floor(w 2^24) / 2^24, independently
//   of every other edge (the reference's u <= w, add_rrrset.h:128-134).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>

namespace rip {

constexpr uint64_t kLcgA = 18145460002477866997ull;
constexpr uint64_t kLcgB = 1ull;
constexpr uint64_t kSampleStride = 4294967295ull;  // S = 2^32 - 1

// Affine map x -> a*x + b (mod 2^64).
struct Affine {
  uint64_t a, b;
};

__host__ __device__ inline Affine aff_then(Affine first, Affine second) {
  // second(first(x))
  return Affine{second.a * first.a, second.a * first.b + second.b};
}

__host__ __device__ inline Affine aff_pow(Affine f, uint64_t n) {
  Affine r{1, 0};
  while (n) {
    if (n & 1) r = aff_then(r, f);
    f = aff_then(f, f);
    n >>= 1;
  }
  return r;
}

__host__ __device__ inline uint64_t aff_apply(Affine f, uint64_t x) {
  return f.a * x + f.b;
}

__host__ __device__ inline Affine master_step() { return Affine{kLcgA, kLcgB}; }

// One step of a sample stream = 2S steps of the master stream.
__host__ __device__ inline Affine sample_step() {
  return aff_pow(master_step(), 2 * kSampleStride);
}

// Element 0 of sample s: master output 1 + 2s = M^(2s+2)(seed).
__host__ __device__ inline uint64_t sample_z(uint64_t seed, uint64_t s) {
  return aff_apply(aff_pow(master_step(), 2 * s + 2), seed);
}

__host__ __device__ inline uint32_t uniform_index(uint64_t x, uint32_t n) {
  return static_cast<uint32_t>(((x >> 32) * static_cast<uint64_t>(n)) >> 32);
}

__host__ __device__ inline float u01(uint64_t x) {
  return static_cast<float>((x >> 40) + 1) * 0x1p-24f;
}

// Integer threshold of an IC edge: live iff (x >> 40) < thr.
inline uint32_t ic_threshold(float w) {
  if (!(w > 0.0f)) return 0;  // also NaN
  if (w >= 1.0f) return 1u << 24;
  float scaled = w * 16777216.0f;  // exact: power-of-two scaling
  return static_cast<uint32_t>(scaled);  // floor for positive values
}

// Device vertex record, one aligned 64-byte line: the jump to the vertex's
// first stream element (IC: P^(1 + 2 begin), the row's first draw; LT:
// P^(1 + begin), the threshold) and its CSR row; for IC the row's gap
// parameters (ic_gap.h): largest threshold T, q and its reciprocal, log2 of
// the block size and the skip thresholds of a full and of the last block.
struct alignas(64) VertexRec {
  uint64_t ja, jb;       // jump to the first element, as an affine map
  uint64_t q;            // IC: gap divisor of the row (0 when T = 2^24)
  uint32_t begin, end;   // transposed-CSR row [begin, end)
  uint32_t thr;          // IC: the row's largest threshold T (0: nothing live);
                         // LT: offset of the row's breakpoint index
  uint32_t ufull;        // IC: a full block has no candidate iff hi32(x) + 1 <= ufull
  uint32_t flags;        // kRowUniform, kRowScalar; IC: beta << kBetaShift
  uint32_t ulast;        // IC: the same for the row's last block when it is partial
  uint64_t rq;           // IC: floor(2^64 / q), the reciprocal of the gap divisor
  uint32_t pad[2];
};
static_assert(sizeof(VertexRec) == 64, "one line per vertex record");

// LT walk record of a vertex: one aligned 128-byte line holds what a walk
// step reads before its breakpoint block.  The row's (breakpoint,
// in-neighbour) pairs start at pair boff, a multiple of 16, so each block of
// 16 pairs is one aligned 128-byte line; `top` is the top level of the row's
// breakpoint index (rows longer than 16 edges), the levels below it sit in
// the graph's index array at ioff, after their own copy of the top level.
struct alignas(128) LtRec {
  uint64_t ja, jb;        // P^(1 + begin): jump to the vertex's threshold element
  uint32_t boff;          // first pair of the row
  uint32_t d;             // in-degree
  uint32_t ioff;          // the row's index in the index array
  uint32_t flags;         // kRowScalar
  uint32_t top[16];       // top index level, padded with 0xFFFFFFFF
  uint32_t begin;         // transposed-CSR row start (rows scanned by weight)
  uint32_t pad[7];
};
static_assert(sizeof(LtRec) == 128, "one line per LT record");

constexpr uint32_t kRowUniform = 1u;  // IC: every edge of the row has threshold T
constexpr uint32_t kBetaShift = 8;     // IC: flags bits 8..12 hold log2 of the block size
constexpr uint32_t kRowScalar = 2u;   // LT: no breakpoints, scan the float chain
// Extra zero elements after the per-edge arrays, so 16-byte loads of an
// aligned group of four never leave the allocation.
constexpr uint64_t kEdgePad = 4;

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &msg);

#define RIP_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess)                                                    \
      throw ::rip::Error(-3, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define RIP_REQUIRE(cond, code, msg)                                         \
  do {                                                                       \
    if (!(cond)) throw ::rip::Error((code), (msg));                          \
  } while (0)

}  // namespace rip
