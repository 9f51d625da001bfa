/*
 * CPU oracle for the IMM hot path -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Only the tests, the smoke check and the benchmark's CPU-baseline leg may
 * load this library.  Nothing in the product links it.
 *
 * What pins each function (the tests under tests/ carry the checks):
 *   orc_lcg64_* / orc_master_output / orc_sample_element
 *       algebra: jump(a)+jump(b) == jump(a+b); split leapfrog == parent's
 *       outputs; square-and-multiply == stepping; split(s,i) rejects i >= s.
 *       Equality with TRNG 4.25's own output: parity unpinned (TRNG is not
 *       vendored in the reference and no reference test checks an RNG value).
 *   orc_sample (IC)   w = 0 -> {root}; w = 1 -> reverse reachability (scipy
 *       BFS); three independently written traversals (FIFO BFS, 64-colour
 *       sweep, per-coin definition) agree; exhaustive live-edge enumeration
 *       on tiny graphs matches the inclusion frequencies.
 *   orc_sample (LT)   single in-edge of weight 1 -> deterministic path;
 *       per-coin definition agrees; walk-step distribution matches weights.
 *   orc_select        known-answer fixture of test/cuda_find_most_influential.cc
 *       and brute-force greedy on tiny pools.
 *   orc_theta*, orc_log_binomial   log(C(n,k)) against math.comb; closed-form
 *       values recomputed with Python's math module in the tests.
 *   orc_imm           round structure checked against the closed forms and
 *       against orc_select/orc_sample run step by step.
 *   orc_spread (IC)   w = 1 -> forward reachability; w = 0 -> the seeds;
 *       exact expected spread by live-edge enumeration on tiny graphs;
 *       equals the reverse formulation (v is reached iff v's RRR set, drawn
 *       with the same coins, meets the seeds) on graphs where both are run.
 *   orc_spread (LT)   path graphs with unit weights -> deterministic chains;
 *       exact activation probabilities of single in-edges; threshold rule
 *       checked against a direct Python restatement on small graphs.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* affine maps x -> a*x + b (mod 2^64): the LCG step and all its jumps       */
/* ------------------------------------------------------------------------ */
typedef struct { uint64_t a, b; } aff;

static aff aff_then(aff first, aff second) {
  /* second(first(x)) */
  aff r;
  r.a = second.a * first.a;
  r.b = second.a * first.b + second.b;
  return r;
}

static aff aff_pow(aff f, uint64_t n) {
  aff r = {1, 0};
  aff p = f;
  while (n) {
    if (n & 1) r = aff_then(r, p);
    p = aff_then(p, p);
    n >>= 1;
  }
  return r;
}

static uint64_t aff_apply(aff f, uint64_t x) { return f.a * x + f.b; }

static const aff MASTER = {ORC_LCG_A, ORC_LCG_B};

/* ---- lcg64 object, following TRNG's documented semantics ---- */
void orc_lcg64_seed(orc_lcg64 *g, uint64_t s) {
  g->a = ORC_LCG_A;
  g->b = ORC_LCG_B;
  g->r = s;
}

uint64_t orc_lcg64_next(orc_lcg64 *g) {
  g->r = g->a * g->r + g->b;
  return g->r;
}

void orc_lcg64_jump(orc_lcg64 *g, uint64_t steps) {
  aff f = {g->a, g->b};
  g->r = aff_apply(aff_pow(f, steps), g->r);
}

static uint64_t inv_odd(uint64_t a) {
  uint64_t x = a; /* correct to 3 bits for odd a; Newton doubles the bits */
  for (int i = 0; i < 6; ++i) x *= 2 - a * x;
  return x;
}

/* Leapfrog split: afterwards the generator yields the parent's outputs
 * i, i+s, i+2s, ... (TRNG lcg64::split(s, i)). */
int orc_lcg64_split(orc_lcg64 *g, uint64_t s, uint64_t i) {
  if (s < 1 || i >= s) return -1;
  if (s == 1) return 0;
  aff f = {g->a, g->b};
  orc_lcg64_jump(g, i + 1); /* r = parent output i */
  aff fs = aff_pow(f, s);
  g->a = fs.a;
  g->b = fs.b;
  /* step back once with the new parameters so the next output is r */
  g->r = (g->r - g->b) * inv_odd(g->a);
  return 0;
}

uint64_t orc_master_output(uint64_t seed, uint64_t t) {
  return aff_apply(aff_pow(MASTER, t + 1), seed);
}

float orc_u01(uint64_t x) {
  return (float)((x >> 40) + 1) * 0x1p-24f;
}

uint32_t orc_uniform_index(uint64_t x, uint32_t n) {
  return (uint32_t)(((x >> 32) * (uint64_t)n) >> 32);
}

uint64_t orc_sample_element(uint64_t seed, uint64_t s, uint64_t j) {
  /* master output index 1 + 2(s + j S); exponents wrap mod 2^64 because the
   * full-period LCG satisfies M^(2^64) = identity. */
  uint64_t t = 1 + 2 * (s + j * ORC_SAMPLE_STRIDE);
  return orc_master_output(seed, t);
}

uint32_t orc_sample_root(uint64_t seed, uint64_t s, uint32_t n) {
  return orc_uniform_index(orc_sample_element(seed, s, 0), n);
}

/* ------------------------------------------------------------------------ */
/* IC row sampling by geometric gaps (the IC part of the stream contract).   */
/* Row v = in-edges [b, b+d) with integer thresholds t_e = floor(w 2^24)     */
/* and T = max t_e.  T = 0: nothing is live.  Block size 2^beta, beta the    */
/* least value in [4, 20] with T 2^beta >= 2^25; block k = edges b + kB +    */
/* [0, len), len = min(B, d - kB), draws elements 1 + 2(b + kB) + j of the   */
/* sample's stream: candidate i uses element 2i for its gap and 2i + 1 for  */
/* its acceptance.  gap = floor(NL(hi32(x) + 1) / q), NL(U) = -log2(U 2^-32)*/
/* in units of 2^-58, q = NL((2^24 - T) 2^8) (q = 0, gap 0, when T = 2^24). */
/* A candidate e is live iff hi32(a) T < t_e 2^32.  Each edge is then live   */
/* independently with probability t_e / 2^24, the reference's u <= w test    */
/* (add_rrrset.h:128-134) with u on the 2^-24 grid.                         */
/* ------------------------------------------------------------------------ */
#include "ic_gap_tables.inc"

static uint64_t mulhi64(uint64_t a, uint64_t b) {
  return (uint64_t)(((unsigned __int128)a * b) >> 64);
}

uint64_t orc_ic_neg_log2(uint64_t u) {
  if (u >= (1ull << 32)) return 0;
  int e = 63 - __builtin_clzll(u);
  uint64_t m = u << (63 - e);
  uint32_t i = (uint32_t)(m >> 55) & 255u;
  uint64_t r = m & ((1ull << 55) - 1);
  /* x = r 2^-63 / (1 + i/256) as X = x 2^64 = (r * kRecipC[i]) >> 62 */
  uint64_t X = (uint64_t)(((unsigned __int128)r * kRecipC[i]) >> 62);
  uint64_t X2 = mulhi64(X, X), X3 = mulhi64(X2, X), X4 = mulhi64(X3, X);
  uint64_t ln = X - X2 / 2 + X3 / 3 - X4 / 4;
  uint64_t frac = kLog2C[i] + (mulhi64(ln, RIP_INV_LN2_Q63) << 1);
  return ((uint64_t)(32 - e) << 58) - (frac >> 6);
}

uint32_t orc_ic_threshold(float w) {
  if (!(w > 0.0f)) return 0;
  if (w >= 1.0f) return 1u << 24;
  return (uint32_t)(w * 16777216.0f);
}

uint64_t orc_ic_gap_q(uint32_t T) {
  if (T >= (1u << 24)) return 0;
  return orc_ic_neg_log2((uint64_t)((1u << 24) - T) << 8);
}

uint32_t orc_ic_block_log2(uint32_t T) {
  uint32_t beta = 4;
  while (beta < 20 && ((uint64_t)T << beta) < (1ull << 25)) ++beta;
  return beta;
}

uint64_t orc_ic_gap(uint64_t x, uint64_t q) {
  return q == 0 ? 0 : orc_ic_neg_log2((x >> 32) + 1) / q;
}

/* ------------------------------------------------------------------------ */
/* fast evaluation of the same stream: per-vertex jump + stepping            */
/* ------------------------------------------------------------------------ */
typedef struct {
  aff P;         /* one step of a sample stream: M^(2S) */
  aff P2[32];    /* P^(2^j) */
  aff *vjump;    /* IC: P^(1 + 2 index[v]) (the row's first element);
                    LT: P^(1 + index[v]) (the threshold element) */
  uint32_t *thr; /* IC: integer threshold of every edge */
  uint32_t *rowT;/* IC: the row's largest threshold */
} stream_ctx;

static void stream_ctx_init(stream_ctx *c, int model, uint32_t n, const uint64_t *index,
                            const float *w) {
  c->P = aff_pow(MASTER, 2 * ORC_SAMPLE_STRIDE);
  c->P2[0] = c->P;
  for (int j = 1; j < 32; ++j) c->P2[j] = aff_then(c->P2[j - 1], c->P2[j - 1]);
  c->vjump = (aff *)malloc(sizeof(aff) * (size_t)(n ? n : 1));
  c->thr = NULL;
  c->rowT = NULL;
#pragma omp parallel for schedule(static)
  for (int64_t v = 0; v < (int64_t)n; ++v)
    c->vjump[v] = aff_pow(c->P, model == ORC_IC ? 1 + 2 * index[v] : 1 + index[v]);
  if (model == ORC_IC) {
    uint64_t m = index[n];
    c->thr = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(m ? m : 1));
    c->rowT = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(n ? n : 1));
#pragma omp parallel for schedule(static)
    for (int64_t v = 0; v < (int64_t)n; ++v) {
      uint32_t T = 0;
      for (uint64_t e = index[v]; e < index[v + 1]; ++e) {
        c->thr[e] = orc_ic_threshold(w[e]);
        if (c->thr[e] > T) T = c->thr[e];
      }
      c->rowT[v] = T;
    }
  }
}

static void stream_ctx_free(stream_ctx *c) {
  free(c->vjump);
  free(c->thr);
  free(c->rowT);
}

/* Live in-edges of row v for one sample: x0 is the sample's element
 * 1 + 2 index[v] (NULL-free fast path) or, with definitional set, every
 * element is recomputed from (seed, s).  Writes the live edge positions in
 * increasing order to live[] and returns their count (at most the degree). */
static uint64_t ic_row_live(const stream_ctx *c, const uint64_t *index, uint32_t v, uint64_t x0,
                            int definitional, uint64_t seed, uint64_t s, uint64_t *live) {
  uint32_t T = c->rowT[v];
  uint64_t b = index[v], d = index[v + 1] - b, nl = 0;
  if (T == 0 || d == 0) return 0;
  uint64_t q = orc_ic_gap_q(T);
  uint32_t beta = orc_ic_block_log2(T);
  uint64_t B = 1ull << beta;
  aff blk_step = c->P2[beta + 1]; /* P^(2B): one block further */
  uint64_t xb = x0;
  for (uint64_t k0 = 0; k0 < d; k0 += B, xb = aff_apply(blk_step, xb)) {
    uint64_t len = d - k0 < B ? d - k0 : B, pos = 0, i = 0;
    uint64_t x = xb;
    while (pos < len) {
      uint64_t gx = definitional ? orc_sample_element(seed, s, 1 + 2 * (b + k0) + 2 * i) : x;
      uint64_t g = orc_ic_gap(gx, q);
      if (g >= len - pos) break;
      pos += g;
      uint64_t e = b + k0 + pos;
      uint64_t y = definitional ? orc_sample_element(seed, s, 1 + 2 * (b + k0) + 2 * i + 1)
                                : aff_apply(c->P, x);
      if ((y >> 32) * T < ((uint64_t)c->thr[e] << 32)) live[nl++] = e;
      pos += 1;
      ++i;
      x = aff_apply(c->P, aff_apply(c->P, x));
    }
  }
  return nl;
}

static uint64_t sample_z(uint64_t seed, uint64_t s) {
  return aff_apply(aff_pow(MASTER, 2 * s + 2), seed);
}

/* ------------------------------------------------------------------------ */
/* per-thread output arenas                                                  */
/* ------------------------------------------------------------------------ */
struct orc_sets {
  uint64_t count;
  uint64_t total;
  uint64_t *offsets; /* count + 1 */
  uint32_t *data;
};

typedef struct {
  uint32_t *buf;
  uint64_t len, cap;
} arena;

static void arena_push(arena *a, uint32_t v) {
  if (a->len == a->cap) {
    a->cap = a->cap ? 2 * a->cap : 4096;
    a->buf = (uint32_t *)realloc(a->buf, sizeof(uint32_t) * a->cap);
  }
  a->buf[a->len++] = v;
}

static int cmp_u32(const void *x, const void *y) {
  uint32_t a = *(const uint32_t *)x, b = *(const uint32_t *)y;
  return (a > b) - (a < b);
}

/* ---- one reverse BFS (IC) or walk (LT) per sample: add_rrrset.h:105-157 ---- */
static void rrr_one(int model, uint32_t n, const uint64_t *index,
                    const uint32_t *src, const float *w, const stream_ctx *c,
                    uint64_t seed, uint64_t s, uint8_t *visited, arena *out,
                    int definitional, uint64_t *live) {
  uint64_t start = out->len;
  uint64_t z = sample_z(seed, s);
  uint32_t root = orc_uniform_index(z, n);
  arena_push(out, root);
  visited[root] = 1;
  if (model == ORC_IC) {
    /* FIFO reverse BFS; the live in-edges of a row depend only on (seed, s,
     * row), so the set is the reverse closure of the root over live edges */
    for (uint64_t head = start; head < out->len; ++head) {
      uint32_t v = out->buf[head];
      uint64_t nl = ic_row_live(c, index, v, aff_apply(c->vjump[v], z), definitional, seed, s,
                                live);
      for (uint64_t j = 0; j < nl; ++j) {
        uint32_t u = src[live[j]];
        if (!visited[u]) {
          visited[u] = 1;
          arena_push(out, u);
        }
      }
    }
  } else {
    uint32_t v = root;
    for (;;) {
      uint64_t x = definitional ? orc_sample_element(seed, s, 1 + index[v])
                                : aff_apply(c->vjump[v], z);
      float threshold = orc_u01(x);
      int64_t next = -1;
      for (uint64_t e = index[v]; e < index[v + 1]; ++e) {
        threshold -= w[e];
        if (threshold > 0) continue;
        uint32_t u = src[e];
        if (!visited[u]) {
          visited[u] = 1;
          arena_push(out, u);
          next = u;
        }
        break;
      }
      if (next < 0) break;
      v = (uint32_t)next;
    }
  }
  for (uint64_t i = start; i < out->len; ++i) visited[out->buf[i]] = 0;
  qsort(out->buf + start, out->len - start, sizeof(uint32_t), cmp_u32);
}

/* ---- 64-colour sweep, as BatchedBFSNeighborColor (batched_add_rrrset.h:184-257):
 * full-vertex sweeps until no mask changes; colour i of the batch is bit 63-i.
 * Each newly active colour at v draws its own live in-edges of v's row. */
static void rrr_batched64(uint32_t n, const uint64_t *index, const uint32_t *src,
                          const float *w, const stream_ctx *c, uint64_t seed,
                          uint64_t s0, uint32_t nb, uint64_t *oldm,
                          uint64_t *newm, arena *out, uint64_t *sizes, uint64_t *live) {
  uint64_t z[64];
  memset(oldm, 0, sizeof(uint64_t) * n);
  memset(newm, 0, sizeof(uint64_t) * n);
  for (uint32_t i = 0; i < nb; ++i) {
    z[i] = sample_z(seed, s0 + i);
    newm[orc_uniform_index(z[i], n)] |= 1ull << (63 - i);
  }
  int found = 1;
  while (found) {
    found = 0;
    for (uint32_t v = 0; v < n; ++v) {
      if (oldm[v] == newm[v]) continue;
      found = 1;
      uint64_t colors = newm[v] ^ oldm[v];
      oldm[v] |= newm[v];
      while (colors) {
        int i = __builtin_clzll(colors); /* bit 63-i <-> colour i */
        uint64_t bit = 1ull << (63 - i);
        colors &= ~bit;
        uint64_t nl = ic_row_live(c, index, v, aff_apply(c->vjump[v], z[i]), 0, seed, s0 + i,
                                  live);
        for (uint64_t j = 0; j < nl; ++j) newm[src[live[j]]] |= bit;
      }
    }
  }
  /* decode: per colour, vertices in ascending order */
  for (uint32_t i = 0; i < nb; ++i) sizes[i] = 0;
  for (uint32_t v = 0; v < n; ++v) {
    uint64_t m = newm[v];
    while (m) {
      int i = __builtin_clzll(m);
      m &= ~(1ull << (63 - i));
      ++sizes[i];
    }
  }
  uint64_t base = out->len, acc = 0;
  uint64_t pos[64];
  for (uint32_t i = 0; i < nb; ++i) {
    pos[i] = base + acc;
    acc += sizes[i];
  }
  for (uint64_t i = 0; i < acc; ++i) arena_push(out, 0);
  for (uint32_t v = 0; v < n; ++v) {
    uint64_t m = newm[v];
    while (m) {
      int i = __builtin_clzll(m);
      m &= ~(1ull << (63 - i));
      out->buf[pos[i]++] = v;
    }
  }
}

orc_sets *orc_sample(int model, int algo, uint32_t n, const uint64_t *index,
                     const uint32_t *src, const float *w, uint64_t seed,
                     uint64_t first, uint64_t count, int nthreads) {
  if (n == 0 || (algo == 1 && model != ORC_IC)) return NULL;
  if (nthreads < 1) nthreads = 1;
  stream_ctx c;
  stream_ctx_init(&c, model, n, index, w);
  uint32_t maxdeg = 0;
  for (uint32_t v = 0; v < n; ++v)
    if (index[v + 1] - index[v] > maxdeg) maxdeg = (uint32_t)(index[v + 1] - index[v]);
  /* record which arena holds each sample */
  uint32_t *owner = (uint32_t *)malloc(sizeof(uint32_t) * (count ? count : 1));
  uint64_t *where = (uint64_t *)malloc(sizeof(uint64_t) * (count ? count : 1));
  uint64_t *size = (uint64_t *)malloc(sizeof(uint64_t) * (count ? count : 1));
  arena *arenas = (arena *)calloc((size_t)nthreads, sizeof(arena));
  uint64_t nbatches = (count + 63) / 64;

#pragma omp parallel num_threads(nthreads)
  {
    int tid = 0;
#ifdef _OPENMP
    tid = omp_get_thread_num();
#endif
    arena *a = &arenas[tid];
    uint64_t *live = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)maxdeg + 1));
    if (algo == 1) {
      uint64_t *oldm = (uint64_t *)malloc(sizeof(uint64_t) * n);
      uint64_t *newm = (uint64_t *)malloc(sizeof(uint64_t) * n);
#pragma omp for schedule(dynamic, 1)
      for (int64_t b = 0; b < (int64_t)nbatches; ++b) {
        uint64_t s0 = (uint64_t)b * 64;
        uint32_t nb = (uint32_t)((count - s0) < 64 ? (count - s0) : 64);
        uint64_t sizes[64];
        uint64_t base = a->len;
        rrr_batched64(n, index, src, w, &c, seed, first + s0, nb, oldm, newm, a,
                      sizes, live);
        for (uint32_t i = 0; i < nb; ++i) {
          owner[s0 + i] = (uint32_t)tid;
          where[s0 + i] = base;
          size[s0 + i] = sizes[i];
          base += sizes[i];
        }
      }
      free(oldm);
      free(newm);
    } else {
      uint8_t *visited = (uint8_t *)calloc(n, 1);
#pragma omp for schedule(dynamic, 64)
      for (int64_t i = 0; i < (int64_t)count; ++i) {
        uint64_t base = a->len;
        rrr_one(model, n, index, src, w, &c, seed, first + (uint64_t)i,
                visited, a, algo == 2, live);
        owner[i] = (uint32_t)tid;
        where[i] = base;
        size[i] = a->len - base;
      }
      free(visited);
    }
    free(live);
  }

  orc_sets *h = (orc_sets *)calloc(1, sizeof(orc_sets));
  h->count = count;
  h->offsets = (uint64_t *)malloc(sizeof(uint64_t) * (count + 1));
  h->offsets[0] = 0;
  for (uint64_t i = 0; i < count; ++i) h->offsets[i + 1] = h->offsets[i] + size[i];
  h->total = h->offsets[count];
  h->data = (uint32_t *)malloc(sizeof(uint32_t) * (h->total ? h->total : 1));
#pragma omp parallel for schedule(static) num_threads(nthreads)
  for (int64_t i = 0; i < (int64_t)count; ++i)
    memcpy(h->data + h->offsets[i], arenas[owner[i]].buf + where[i],
           sizeof(uint32_t) * size[i]);
  for (int t = 0; t < nthreads; ++t) free(arenas[t].buf);
  free(arenas);
  free(owner);
  free(where);
  free(size);
  stream_ctx_free(&c);
  return h;
}

uint64_t orc_sets_count(const orc_sets *h) { return h ? h->count : 0; }
uint64_t orc_sets_total(const orc_sets *h) { return h ? h->total : 0; }

void orc_sets_export(const orc_sets *h, uint64_t *offsets, uint32_t *vertices) {
  memcpy(offsets, h->offsets, sizeof(uint64_t) * (h->count + 1));
  memcpy(vertices, h->data, sizeof(uint32_t) * h->total);
}

void orc_sets_free(orc_sets *h) {
  if (!h) return;
  free(h->offsets);
  free(h->data);
  free(h);
}

/* ------------------------------------------------------------------------ */
/* greedy max coverage (find_most_influential.h:85-174, lowest id on ties as  */
/* src/gpu/find_most_influential.cc:65-72)                                    */
/* ------------------------------------------------------------------------ */
void orc_count(uint32_t n, uint64_t nsets, const uint64_t *offsets,
               const uint32_t *vertices, uint32_t *counts) {
  memset(counts, 0, sizeof(uint32_t) * n);
  for (uint64_t i = 0; i < offsets[nsets]; ++i) counts[vertices[i]]++;
}

uint32_t orc_select(uint32_t n, uint64_t nsets, const uint64_t *offsets,
                    const uint32_t *vertices, uint32_t k, uint32_t *seeds,
                    uint64_t *covered) {
  uint32_t *counts = (uint32_t *)malloc(sizeof(uint32_t) * (n ? n : 1));
  orc_count(n, nsets, offsets, vertices, counts);
  /* inverted index vertex -> sets by counting sort */
  uint64_t *start = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  start[0] = 0;
  for (uint32_t v = 0; v < n; ++v) start[v + 1] = start[v] + counts[v];
  uint64_t total = start[n];
  uint64_t *inv = (uint64_t *)malloc(sizeof(uint64_t) * (total ? total : 1));
  uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * ((size_t)n + 1));
  memcpy(cur, start, sizeof(uint64_t) * ((size_t)n + 1));
  for (uint64_t s = 0; s < nsets; ++s)
    for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i)
      inv[cur[vertices[i]]++] = s;
  uint8_t *dead = (uint8_t *)calloc(nsets ? nsets : 1, 1);

  uint64_t uncovered = nsets;
  uint32_t got = 0;
  while (got < k && uncovered != 0) {
    uint32_t best = 0;
    for (uint32_t v = 1; v < n; ++v)
      if (counts[v] > counts[best]) best = v;
    if (counts[best] == 0) break; /* only empty sets remain */
    uncovered -= counts[best];
    seeds[got++] = best;
    for (uint64_t p = start[best]; p < start[best + 1]; ++p) {
      uint64_t s = inv[p];
      if (dead[s]) continue;
      dead[s] = 1;
      for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i) counts[vertices[i]]--;
    }
  }
  if (covered) *covered = nsets - uncovered;
  free(counts);
  free(start);
  free(inv);
  free(cur);
  free(dead);
  return got;
}

/* ------------------------------------------------------------------------ */
/* IMM closed forms                                                          */
/* ------------------------------------------------------------------------ */
double orc_log_binomial(uint64_t n, uint64_t k) {
  return lgamma((double)(n + 1)) - lgamma((double)(n - k + 1)) -
         lgamma((double)(k + 1));
}

int64_t orc_theta_prime(int64_t x, double eps_prime, double l, uint64_t k,
                        uint64_t n) {
  if (k > n / 2) k = n / 2;
  return (int64_t)((2 + 2. / 3. * eps_prime) *
                   (l * log((double)n) + orc_log_binomial(n, k) +
                    log(log2((double)n))) *
                   pow(2.0, (double)x) / (eps_prime * eps_prime));
}

uint64_t orc_theta(double eps, double l, uint64_t k, double lb, uint64_t n) {
  if (lb == 0) return 0;
  if (k > n / 2) k = n / 2;
  double term1 = 0.6321205588285577;
  double alpha = sqrt(l * log((double)n) + log(2));
  double beta = sqrt(term1 * (orc_log_binomial(n, k) + l * log((double)n) + log(2)));
  double lamda = (double)(2 * n) * (term1 * alpha + beta) * (term1 * alpha + beta) *
                 pow(eps, -2);
  return (uint64_t)(lamda / lb);
}

/* ------------------------------------------------------------------------ */
/* IMM driver                                                                */
/* ------------------------------------------------------------------------ */
typedef struct {
  uint64_t count, total, cap_sets, cap_data;
  uint64_t *offsets;
  uint32_t *data;
} pool;

static void pool_append(pool *p, orc_sets *h) {
  if (p->count + h->count + 1 > p->cap_sets) {
    p->cap_sets = 2 * (p->count + h->count + 1);
    p->offsets = (uint64_t *)realloc(p->offsets, sizeof(uint64_t) * p->cap_sets);
  }
  if (p->total + h->total > p->cap_data) {
    p->cap_data = 2 * (p->total + h->total) + 1;
    p->data = (uint32_t *)realloc(p->data, sizeof(uint32_t) * p->cap_data);
  }
  if (p->count == 0) p->offsets[0] = 0;
  for (uint64_t i = 0; i < h->count; ++i)
    p->offsets[p->count + 1 + i] = p->total + h->offsets[i + 1];
  memcpy(p->data + p->total, h->data, sizeof(uint32_t) * h->total);
  p->count += h->count;
  p->total += h->total;
}

static void pool_grow_to(pool *p, int model, uint32_t n, const uint64_t *index,
                         const uint32_t *src, const float *w, uint64_t seed,
                         uint64_t target, int nthreads) {
  if (target <= p->count) return;
  orc_sets *h = orc_sample(model, 0, n, index, src, w, seed, p->count,
                           target - p->count, nthreads);
  pool_append(p, h);
  orc_sets_free(h);
}

/* sets [base + p->count, base + target) appended: R2 of OPIM-C lives at base 2^31 */
static void pool_grow_at(pool *p, int model, uint32_t n, const uint64_t *index,
                         const uint32_t *src, const float *w, uint64_t seed, uint64_t base,
                         uint64_t target, int nthreads) {
  if (target <= p->count) return;
  orc_sets *h = orc_sample(model, 0, n, index, src, w, seed, base + p->count,
                           target - p->count, nthreads);
  pool_append(p, h);
  orc_sets_free(h);
}

uint32_t orc_imm(int model, uint32_t n, const uint64_t *index,
                 const uint32_t *src, const float *w, uint32_t k, double eps,
                 uint64_t seed, int nthreads, uint32_t *seeds,
                 orc_imm_record *rec) {
  pool p;
  memset(&p, 0, sizeof(p));
  p.cap_sets = 1;
  p.offsets = (uint64_t *)malloc(sizeof(uint64_t));
  p.offsets[0] = 0;
  orc_imm_record r;
  memset(&r, 0, sizeof(r));
  uint32_t *tmp = (uint32_t *)malloc(sizeof(uint32_t) * (k ? k : 1));

  double l = 1.0 * (1 + 1 / log2((double)n));
  double eps_prime = 1.4142135623730951 * eps;
  double lb = 0;
  for (int64_t x = 1; x < log2((double)n); ++x) {
    int64_t tp = orc_theta_prime(x, eps_prime, l, k, n);
    if (tp <= (int64_t)p.count) tp = (int64_t)p.count;
    else pool_grow_to(&p, model, n, index, src, w, seed, (uint64_t)tp, nthreads);
    uint64_t cov = 0;
    orc_select(n, p.count, p.offsets, p.data, k, tmp, &cov);
    double f = (double)cov / (double)p.count;
    if (r.rounds < 64) {
      r.theta_prime[r.rounds] = tp;
      r.f[r.rounds] = f;
    }
    r.rounds++;
    if (f >= pow(2, (double)-x)) {
      lb = ((double)n * f) / (1 + eps_prime);
      break;
    }
  }
  uint64_t theta = orc_theta(eps, l, k, lb, n);
  pool_grow_to(&p, model, n, index, src, w, seed, theta, nthreads);
  uint64_t cov = 0;
  uint32_t got = orc_select(n, theta, p.offsets, p.data, k, seeds, &cov);
  r.theta = theta;
  r.num_sets = p.count;
  r.lb = lb;
  r.final_f = theta ? (double)cov / (double)theta : 0.0;
  if (rec) *rec = r;
  free(tmp);
  free(p.offsets);
  free(p.data);
  return got;
}

/* ------------------------------------------------------------------------ */
/* forward Monte Carlo spread: diffusion_simulation.h:73-178                 */
/* ------------------------------------------------------------------------ */
void orc_spread(int model, uint32_t n, const uint64_t *index, const uint32_t *src,
                const float *w, const uint32_t *seeds, uint32_t nseeds, uint64_t seed,
                uint64_t first, uint64_t count, uint64_t *activated, uint32_t *levels,
                int nthreads) {
  const uint64_t m = index[n];
  /* forward CSR: row u = targets of the transposed entries with source u */
  uint64_t *fidx = (uint64_t *)calloc((size_t)n + 1, sizeof(uint64_t));
  uint32_t *fdst = (uint32_t *)malloc(sizeof(uint32_t) * (size_t)(m ? m : 1));
  float *fw = (float *)malloc(sizeof(float) * (size_t)(m ? m : 1));
  for (uint64_t i = 0; i < m; ++i) fidx[src[i] + 1]++;
  for (uint32_t u = 0; u < n; ++u) fidx[u + 1] += fidx[u];
  {
    uint64_t *cur = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)(n ? n : 1));
    memcpy(cur, fidx, sizeof(uint64_t) * n);
    for (uint32_t v = 0; v < n; ++v)
      for (uint64_t i = index[v]; i < index[v + 1]; ++i) {
        const uint64_t f = cur[src[i]]++;
        fdst[f] = v;
        fw[f] = w[i];
      }
    free(cur);
  }
  const aff P = aff_pow(MASTER, 2 * ORC_SAMPLE_STRIDE);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
#pragma omp parallel
  {
    uint8_t *act = (uint8_t *)calloc((size_t)n ? n : 1, 1);
    uint8_t *cand = (uint8_t *)calloc((size_t)n ? n : 1, 1);
    uint32_t *queue = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n ? n : 1));
    uint32_t *cl = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)n ? n : 1));
#pragma omp for schedule(dynamic, 1)
    for (int64_t ri = 0; ri < (int64_t)count; ++ri) {
      const uint64_t z = sample_z(seed, first + (uint64_t)ri);
      uint64_t len = 0;
      for (uint32_t i = 0; i < nseeds; ++i)
        if (!act[seeds[i]]) {
          act[seeds[i]] = 1;
          queue[len++] = seeds[i];
        }
      uint64_t lo = 0, hi = len;
      uint32_t lev = 0;
      if (model == ORC_IC) {
        while (lo < hi) {
          ++lev;
          for (uint64_t q = lo; q < hi; ++q) {
            const uint32_t u = queue[q];
            uint64_t x = aff_apply(aff_pow(P, 1 + fidx[u]), z);
            for (uint64_t f = fidx[u]; f < fidx[u + 1]; ++f, x = aff_apply(P, x)) {
              const uint32_t v = fdst[f];
              if (!act[v] && orc_u01(x) <= fw[f]) {
                act[v] = 1;
                queue[len++] = v;
              }
            }
          }
          lo = hi;
          hi = len;
        }
      } else {
        do {
          ++lev;
          uint64_t nc = 0;
          for (uint64_t q = lo; q < hi; ++q) {
            const uint32_t u = queue[q];
            for (uint64_t f = fidx[u]; f < fidx[u + 1]; ++f) {
              const uint32_t v = fdst[f];
              if (!act[v] && !cand[v]) {
                cand[v] = 1;
                cl[nc++] = v;
              }
            }
          }
          const uint64_t before = len;
          for (uint64_t c = 0; c < nc; ++c) {
            const uint32_t v = cl[c];
            float total = 0.0f;
            for (uint64_t i = index[v]; i < index[v + 1]; ++i)
              if (act[src[i]]) total += w[i];
            const float th = orc_u01(aff_apply(aff_pow(P, 1 + (uint64_t)v), z));
            if (total >= th) queue[len++] = v;
          }
          for (uint64_t c = 0; c < nc; ++c) cand[cl[c]] = 0;
          for (uint64_t q = before; q < len; ++q) act[queue[q]] = 1;
          lo = before;
          hi = len;
        } while (lo < hi);
      }
      activated[ri] = len;
      levels[ri] = lev;
      for (uint64_t q = 0; q < len; ++q) act[queue[q]] = 0;
    }
    free(act);
    free(cand);
    free(queue);
    free(cl);
  }
  free(fidx);
  free(fdst);
  free(fw);
}

/* ------------------------------------------------------------------------ */
/* OPIM-C (opimc.h)                                                          */
/* ------------------------------------------------------------------------ */
uint64_t orc_coverage(uint32_t n, uint64_t nsets, const uint64_t *offsets,
                      const uint32_t *vertices, const uint32_t *seeds, uint32_t nseeds) {
  /* opimc.h:100-118: a set counts once if any of its vertices is a seed */
  uint8_t *is_seed = (uint8_t *)calloc(n ? n : 1, 1);
  for (uint32_t i = 0; i < nseeds; ++i) is_seed[seeds[i]] = 1;
  uint64_t hits = 0;
  for (uint64_t s = 0; s < nsets; ++s) {
    for (uint64_t i = offsets[s]; i < offsets[s + 1]; ++i)
      if (is_seed[vertices[i]]) {
        ++hits;
        break;
      }
  }
  free(is_seed);
  return hits;
}

/* 1 - 1/e as the reference's float constant */
#define ORC_OPIMC_A 0.6321203113733684f

/* opimc.h:26-33.  C's log/sqrt are the double functions; the float
 * variables round on assignment as in the reference. */
float orc_opimc_theta_helper(uint64_t n, float epsilon, uint64_t k, float delta) {
  (void)epsilon;
  float B = ORC_OPIMC_A * sqrt(log(6.0 / delta));
  float C = sqrt(ORC_OPIMC_A * (orc_log_binomial(n, k) + log(6.0 / delta)));
  float D = (B + C) * (B + C);
  return 2 * D;
}

/* Equation (16), opimc.h:36-38 */
uint64_t orc_opimc_theta_max(uint64_t n, float epsilon, uint64_t k, float delta) {
  float top = (float)n * orc_opimc_theta_helper(n, epsilon, k, delta);
  float bottom = epsilon * epsilon * (float)k;
  return (uint64_t)ceil(top / bottom);
}

/* Equation (17), opimc.h:41-43 */
uint64_t orc_opimc_theta_zero(uint64_t n, float epsilon, uint64_t k, float delta) {
  return (uint64_t)ceil(orc_opimc_theta_helper(n, epsilon, k, delta));
}

/* Equation (5), opimc.h:46-54 */
float orc_opimc_lower_bound(uint64_t coverage, float delta2, uint64_t n, uint64_t theta) {
  float a = log(1 / delta2);
  float b = sqrt((float)coverage + 2 * a / 9);
  float c = sqrt(a / 2);
  float d = (b - c) * (b - c);
  float e = a / 18;
  return ((float)n / (float)theta) * (d - e);
}

/* Equation (8), opimc.h:57-65 */
float orc_opimc_upper_bound(uint64_t coverage, float delta1, uint64_t n, uint64_t theta) {
  float b = (float)coverage / ORC_OPIMC_A;
  float c = log(1 / delta1) / 2;
  float d = sqrt(b + c);
  float e = (d + sqrt(c)) * (d + sqrt(c));
  return ((float)n / (float)theta) * e;
}

/* OPIMC (opimc.h:150-241): R1 = sample ids [0, theta) of `seed`, R2 = ids
 * [2^31, 2^31 + theta); both double each round until alpha >= 1-1/e-eps. */
uint32_t orc_opimc(int model, uint32_t n, const uint64_t *index, const uint32_t *src,
                   const float *w, uint32_t k, double eps, double delta_cfg, uint64_t seed,
                   int nthreads, uint32_t *seeds, orc_opimc_record *rec) {
  orc_opimc_record r;
  memset(&r, 0, sizeof(r));
  float eps_f = (float)eps, delta_f = (float)delta_cfg;
  uint64_t theta = orc_opimc_theta_zero(n, eps_f, k, delta_f);
  r.theta_zero = theta;
  r.theta_max = orc_opimc_theta_max(n, eps_f, k, delta_f);
  uint64_t imax = (uint64_t)ceil(log2((double)(r.theta_max / theta)));
  float delta = delta_f / (float)(3 * imax);
  pool r1, r2;
  memset(&r1, 0, sizeof(r1));
  memset(&r2, 0, sizeof(r2));
  r1.cap_sets = r2.cap_sets = 1;
  r1.offsets = (uint64_t *)calloc(1, sizeof(uint64_t));
  r2.offsets = (uint64_t *)calloc(1, sizeof(uint64_t));
  uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * (k ? k : 1));
  uint32_t got = 0;
  for (uint32_t round = 0;; ++round) {
    uint64_t before = r1.count;
    pool_grow_at(&r1, model, n, index, src, w, seed, 0, theta, nthreads);
    pool_grow_at(&r2, model, n, index, src, w, seed, 1ull << 31, theta, nthreads);
    uint64_t cov = 0;
    got = orc_select(n, r1.count, r1.offsets, r1.data, k, cur, &cov);
    double coverage1 = (double)cov / (double)r1.count;
    uint64_t c1 = (uint64_t)(coverage1 * (double)r1.count);
    uint64_t c2 = orc_coverage(n, r2.count, r2.offsets, r2.data, cur, got);
    float upper = orc_opimc_upper_bound(c1, delta, n, r1.count);
    float lower = orc_opimc_lower_bound(c2, delta, n, r2.count);
    float alpha = lower / upper;
    if (round < 64) {
      r.rrr_sets_generated[round] = 2 * (r1.count - before);
      r.coverage1[round] = c1;
      r.coverage2[round] = c2;
      r.lower[round] = lower;
      r.upper[round] = upper;
      r.alpha[round] = alpha;
    }
    r.rounds = round + 1;
    r.theta = theta;
    if (alpha >= (ORC_OPIMC_A - eps)) break;
    theta *= 2;
  }
  memcpy(seeds, cur, sizeof(uint32_t) * got);
  r.num_seeds = got;
  if (rec) *rec = r;
  free(cur);
  free(r1.offsets);
  free(r1.data);
  free(r2.offsets);
  free(r2.data);
  return got;
}
