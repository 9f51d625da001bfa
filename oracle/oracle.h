/*
 * CPU oracle for the IMM hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This library restates, in plain C, what the reference computes on the IMM
 * path (RRR-set sampling under IC and LT, greedy max-coverage seed selection,
 * and the IMM theta-estimation driver) under the same per-sample stream
 * contract the HIP implementation uses.  It is the checker: only the tests,
 * the smoke check and the CPU-baseline leg of the benchmark may load it.  The
 * product (libripples_amd.so) never links, loads or calls it.
 *
 * Stream contract (shared with the device code, written out here so the two
 * can be compared):
 *   - master generator: TRNG lcg64 recurrence r <- a*r + b (mod 2^64) with
 *     a = 18145460002477866997, b = 1, seeded with r = seed; the t-th output
 *     (t = 0, 1, ...) is x_t = M^(t+1)(seed) where M is the affine step.
 *   - split(s, i) is TRNG's leapfrog split: the i-th of s sub-streams yields
 *     the parent's outputs i, i+s, i+2s, ...
 *   - weights stream  = master.split(2, 0)   (tools/imm.cc:172-174)
 *     sampling stream = master.split(2, 1)   (tools/imm.cc:212-214)
 *   - sample s draws from sampling.split(S, s) with S = 2^32 - 1:
 *       element 0      -> the root,  root = (hi32(x) * n) >> 32
 *       elements 1 + 2 index[v] + j -> IC draws of vertex v's row (gaps
 *                         between candidate edges and their acceptances,
 *                         see the IC section of oracle.c)
 *       element 1 + index[v] -> LT threshold of vertex v
 *     so element j of sample s is master output 1 + 2*(s + j*S).
 *   - uniform01: u = ((x >> 40) + 1) * 2^-24, a float in (0, 1].
 *   - An IC edge of weight w is live with probability t / 2^24, t =
 *     floor(w 2^24) (u <= w on the 2^-24 grid, add_rrrset.h:130), each edge
 *     independently: a row's candidates are drawn by geometric gaps at the
 *     row's largest threshold and thinned to each edge's own.
 *
 * Parity status of each function is listed in the header of oracle.c.
 */
#ifndef RIPPLES_ORACLE_H
#define RIPPLES_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_LCG_A 18145460002477866997ull
#define ORC_LCG_B 1ull
#define ORC_SAMPLE_STRIDE 4294967295ull /* S = 2^32 - 1 */

enum { ORC_IC = 0, ORC_LT = 1 };

/* ---- lcg64 restatement (TRNG 4.25 lcg64, published recurrence) ---- */
typedef struct { uint64_t a, b, r; } orc_lcg64;

void orc_lcg64_seed(orc_lcg64 *g, uint64_t s);
uint64_t orc_lcg64_next(orc_lcg64 *g);
void orc_lcg64_jump(orc_lcg64 *g, uint64_t steps);
int orc_lcg64_split(orc_lcg64 *g, uint64_t s, uint64_t i);
/* x_t of the master stream seeded with `seed`, by square-and-multiply. */
uint64_t orc_master_output(uint64_t seed, uint64_t t);
float orc_u01(uint64_t x);
uint32_t orc_uniform_index(uint64_t x, uint32_t n);

/* ---- IC gap sampling (the contract's integer functions, exposed for tests) ----
 * orc_ic_neg_log2(u): -log2(u / 2^32) for u in [1, 2^32], units of 2^-58.
 * orc_ic_threshold(w): floor(w 2^24) clamped to [0, 2^24] (0 for NaN).
 * orc_ic_gap_q(T): NL((2^24 - T) 2^8), 0 for T = 2^24.
 * orc_ic_block_log2(T): least beta in [4, 20] with T 2^beta >= 2^25.
 * orc_ic_gap(x, q): floor(NL(hi32(x) + 1) / q), 0 for q = 0. */
uint64_t orc_ic_neg_log2(uint64_t u);
uint32_t orc_ic_threshold(float w);
uint64_t orc_ic_gap_q(uint32_t T);
uint32_t orc_ic_block_log2(uint32_t T);
uint64_t orc_ic_gap(uint64_t x, uint64_t q);

/* element j of sample s's stream, computed from the definition */
uint64_t orc_sample_element(uint64_t seed, uint64_t s, uint64_t j);
uint32_t orc_sample_root(uint64_t seed, uint64_t s, uint32_t n);

/* ---- RRR sampling over a transposed CSR (in-neighbours of each vertex) ----
 * index: uint64[n+1], src: uint32[m] (in-neighbour ids), w: float[m].
 * The result handle holds count sorted vertex lists in sample order. */
typedef struct orc_sets orc_sets;

/* algo: 0 = one FIFO BFS / walk per sample (add_rrrset.h:105-157),
 *       1 = 64-colour batched sweep (batched_add_rrrset.h:184-257), IC only,
 *       2 = per-coin evaluation from the stream definition (slow; tiny tests) */
orc_sets *orc_sample(int model, int algo, uint32_t n, const uint64_t *index,
                     const uint32_t *src, const float *w, uint64_t seed,
                     uint64_t first, uint64_t count, int nthreads);
uint64_t orc_sets_count(const orc_sets *h);
uint64_t orc_sets_total(const orc_sets *h);
/* offsets: uint64[count+1]; vertices: uint32[total] */
void orc_sets_export(const orc_sets *h, uint64_t *offsets, uint32_t *vertices);
void orc_sets_free(orc_sets *h);

/* ---- greedy max-coverage selection, lowest vertex id on ties ----
 * Returns the number of seeds written (<= k); *covered = sets covered. */
uint32_t orc_select(uint32_t n, uint64_t nsets, const uint64_t *offsets,
                    const uint32_t *vertices, uint32_t k, uint32_t *seeds,
                    uint64_t *covered);
/* Initial occurrence counts (counting.h:65-97). counts: uint32[n] */
void orc_count(uint32_t n, uint64_t nsets, const uint64_t *offsets,
               const uint32_t *vertices, uint32_t *counts);

/* ---- IMM closed forms (imm.h:77-128, utility.h:84-87) ---- */
double orc_log_binomial(uint64_t n, uint64_t k);
int64_t orc_theta_prime(int64_t x, double eps_prime, double l, uint64_t k,
                        uint64_t n);
uint64_t orc_theta(double eps, double l, uint64_t k, double lb, uint64_t n);

/* ---- IMM driver (imm.h:145-260, 768-860) ---- */
typedef struct {
  uint64_t theta;          /* final theta */
  uint64_t num_sets;       /* |RR| at the end */
  uint32_t rounds;         /* estimation rounds run */
  int64_t theta_prime[64]; /* theta' per round */
  double f[64];            /* coverage fraction per round */
  double lb;               /* lower bound LB */
  double final_f;          /* coverage of the final seeds on RR[0, theta) */
} orc_imm_record;

uint32_t orc_imm(int model, uint32_t n, const uint64_t *index,
                 const uint32_t *src, const float *w, uint32_t k, double eps,
                 uint64_t seed, int nthreads, uint32_t *seeds,
                 orc_imm_record *rec);

/* ---- forward Monte Carlo spread (diffusion_simulation.h:73-178) ----
 * Replica r of a seed set draws from the sample stream r of `seed` (element j
 * = master output 1 + 2(r + jS)): element 1 + f is the IC coin of forward
 * edge f, element 1 + v the LT threshold of vertex v.  The forward CSR lists,
 * for each vertex u, the targets v of the transposed entries with source u,
 * in transposed-position order.
 *   IC: breadth-first from the seeds; edge f is live when u01 <= w; levels =
 *       breadth-first layers processed (the seed layer included).
 *   LT: per level, every inactive out-neighbour of the vertices activated in
 *       the previous level sums, in transposed CSR order, the weights of its
 *       active in-neighbours and activates when the sum >= its threshold;
 *       levels = iterations, the last one (activating nothing) included.
 * activated[i] / levels[i] belong to replica first + i. */
void orc_spread(int model, uint32_t n, const uint64_t *index, const uint32_t *src,
                const float *w, const uint32_t *seeds, uint32_t nseeds, uint64_t seed,
                uint64_t first, uint64_t count, uint64_t *activated, uint32_t *levels,
                int nthreads);

/* ---- OPIM-C (opimc.h:26-241) ---- */
/* Sets among [0, nsets) containing any of the seeds (FindCoverage). */
uint64_t orc_coverage(uint32_t n, uint64_t nsets, const uint64_t *offsets,
                      const uint32_t *vertices, const uint32_t *seeds, uint32_t nseeds);
float orc_opimc_theta_helper(uint64_t n, float epsilon, uint64_t k, float delta);
uint64_t orc_opimc_theta_max(uint64_t n, float epsilon, uint64_t k, float delta);
uint64_t orc_opimc_theta_zero(uint64_t n, float epsilon, uint64_t k, float delta);
float orc_opimc_lower_bound(uint64_t coverage, float delta2, uint64_t n, uint64_t theta);
float orc_opimc_upper_bound(uint64_t coverage, float delta1, uint64_t n, uint64_t theta);

typedef struct {
  uint64_t theta_zero, theta_max, theta;
  uint32_t rounds, num_seeds;
  uint64_t rrr_sets_generated[64];
  uint64_t coverage1[64], coverage2[64];
  float lower[64], upper[64], alpha[64];
} orc_opimc_record;

/* R1 = sample ids [0, theta) of `seed`, R2 = ids [2^31, 2^31 + theta). */
uint32_t orc_opimc(int model, uint32_t n, const uint64_t *index, const uint32_t *src,
                   const float *w, uint32_t k, double eps, double delta, uint64_t seed,
                   int nthreads, uint32_t *seeds, orc_opimc_record *rec);

#ifdef __cplusplus
}
#endif
#endif
