// Host check of the IC gap arithmetic (ripples_amd/csrc/ic_gap.h):
//  1. ic_gap_recip(q) = floor(2^64 / q), by 128-bit division;
//  2. ic_gap_fast(x, q, rq) = ic_gap(x, q) (multiplication by the
//     reciprocal and one correction against the exact 64-bit division);
//  3. ic_skip_len(q, len) is the exact boundary: gap >= len at U = the
//     returned value and below, gap < len just above it;
// for the q of every threshold class a row can have (random T in [1, 2^24),
// T near 1, 2^23 and 2^24, powers of two) and uniforms at the ends of the
// range (U = 1, 2^32) and at random.
// Usage: ic_gap_check <seed> <rows>; prints "ok" and exits 0 when all agree.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../ripples_amd/csrc/ic_gap.h"

using namespace rip;

static int fails = 0;
static void fail(const char *what, uint32_t T, uint64_t a, uint64_t b) {
  if (fails++ < 10)
    std::printf("FAIL %s T=%u %llu %llu\n", what, T, (unsigned long long)a, (unsigned long long)b);
}

static uint64_t x_of(uint64_t U, uint64_t low) { return ((U - 1) << 32) | (low & 0xFFFFFFFFull); }

int main(int argc, char **argv) {
  const unsigned seed = argc > 1 ? std::atoi(argv[1]) : 1;
  const int rows = argc > 2 ? std::atoi(argv[2]) : 1000;
  std::mt19937_64 rng(seed);
  IcLogTables tab;
  for (int i = 0; i < 256; ++i) {
    tab.log2c[i] = kLog2C[i];
    tab.recipc[i] = kRecipC[i];
  }
  for (int r = 0; r < rows; ++r) {
    uint32_t T;
    switch (r % 6) {
      case 0: T = 1 + static_cast<uint32_t>(rng() % 64); break;
      case 1: T = (1u << 24) - 1 - static_cast<uint32_t>(rng() % 64); break;
      case 2: T = (1u << 23) + static_cast<uint32_t>(rng() % 129) - 64; break;
      case 3: T = 1u << (rng() % 24); break;
      default: T = 1 + static_cast<uint32_t>(rng() % ((1u << 24) - 1)); break;
    }
    const uint64_t q = ic_gap_q(T);
    const uint64_t rq = ic_gap_recip(q);
    const unsigned __int128 exact = (static_cast<unsigned __int128>(1) << 64) / q;
    if (rq != static_cast<uint64_t>(exact)) fail("recip", T, rq, static_cast<uint64_t>(exact));
    for (int i = 0; i < 200; ++i) {
      uint64_t U;
      switch (i % 5) {
        case 0: U = 1 + (rng() % 4); break;
        case 1: U = (1ull << 32) - (rng() % 4); break;
        default: U = 1 + (rng() % (1ull << 32)); break;
      }
      const uint64_t x = x_of(U, rng());
      const uint64_t g0 = ic_gap(x, q), g1 = ic_gap_fast(x, q, rq, &tab);
      if (g0 != g1) fail("gap", T, g0, g1);
    }
    const uint32_t beta = ic_block_log2(T);
    const uint32_t lens[4] = {1u, 1u + static_cast<uint32_t>(rng() % (1u << beta)), 1u << beta,
                              static_cast<uint32_t>(1 + rng() % 17)};
    for (uint32_t len : lens) {
      const uint64_t u = ic_skip_len(q, len);
      if (u > 0 && ic_gap(x_of(u, rng()), q) < len) fail("skip below", T, u, len);
      if (u < (1ull << 32) && ic_gap(x_of(u + 1, rng()), q) >= len) fail("skip above", T, u, len);
      if (u == 0 && ic_gap(x_of(1, 0), q) >= len) fail("skip none", T, u, len);
    }
    if (ic_full_skip(q, beta) != ic_skip_len(q, 1u << beta)) fail("full", T, beta, 0);
  }
  if (fails) {
    std::printf("%d failures\n", fails);
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
