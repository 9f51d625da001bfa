/*
 * libripples_amd -- C ABI of the MI355X-native IMM hot path.
 *
 * Plain C types only: every entry point returns 0 on success or a negative
 * RIP_E* code; rip_last_error() returns a thread-local message for the last
 * failure on the calling thread.  No exception crosses the boundary.
 *
 * Object model (reference citations are relative to the pnnl/ripples tree):
 *   rip_graph  device copy of a transposed CSR with model-specific edge data.
 *              Replaces make_gpu_graph / gpu_graph (src/gpu/gpu_graph.cc:48-70,
 *              include/ripples/gpu/gpu_graph.h:47-90) of a GraphBwd.
 *   rip_pool   HBM-resident store of RRR sets of one (graph, seed) stream.
 *              Replaces the RRRsets vector (include/ripples/rrr_sets.h:81-105)
 *              filled by the streaming generator and the device copy made for
 *              selection (streaming_find_most_influential.h:139-244).
 * Sample s of a pool is always the same set for the same (graph, seed): the
 * RNG stream of a sample depends only on (seed, s), never on scheduling.
 *
 * Threading: one object is used by one host thread at a time; distinct
 * objects may be used concurrently.  Each pool owns one HIP stream.
 */
#ifndef RIPPLES_AMD_H
#define RIPPLES_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RIP_OK 0
#define RIP_EINVAL (-1)   /* bad argument */
#define RIP_ENOMEM (-2)   /* device or host allocation failed */
#define RIP_EDEVICE (-3)  /* HIP runtime error */
#define RIP_ESTATE (-4)   /* object in the wrong state for the call */

/* Diffusion models: independent_cascade_tag / linear_threshold_tag
 * (include/ripples/diffusion_simulation.h:56-59). */
#define RIP_IC 0
#define RIP_LT 1

typedef struct rip_graph rip_graph;
typedef struct rip_pool rip_pool;

/* Sampler tuning (gpu_batch_size / streaming engine knobs of
 * include/ripples/imm_configuration.h:64-109).  Zero fields = automatic. */
typedef struct {
  uint32_t colors;       /* IC: samples fused per colour group, 1..64 (grid
                            sampler and colour sampler; 0 = 64) */
  uint32_t block;        /* IC: threads per workgroup: 256, 512 or 1024 */
  uint32_t slots;        /* IC: resident colour groups (persistent workgroups) */
  uint32_t queue_cap;    /* IC lane sampler: vertices one lane's sample may
                            reach before the wave sampler takes it over:
                            64, 256, 1024, 2048 or 4096 (0 = 1024) */
  uint32_t heavy_degree; /* IC: rows this long wait for the hub phase */
  uint32_t sampler;      /* IC: RIP_SAMPLER_AUTO, _WAVE or _COLOUR */
  uint32_t out_chunk;    /* LT and IC wave sampler: pool elements a wave
                            reserves at a time (the pool's data then holds
                            gaps between chunks) */
  uint32_t frontier_cap; /* IC grid sampler (AUTO): frontier entries a BFS
                            level may list (0 = automatic); vertices past it
                            are found again by a scan of the pending masks */
  uint32_t wave_stage;   /* IC AUTO: nonzero = samples passing queue_cap are
                            first run one per wave, 64 frontier vertices at a
                            time, up to 4096 vertices, and only the larger
                            ones go to the grid sampler */
} rip_tune;

/* IC sampler kinds (all give the same sets):
 *   AUTO    one reverse BFS per lane (64 samples per wave); samples passing
 *           queue_cap vertices go to the grid sampler: 64-sample colour groups,
 *           many groups per batch, one BFS level of all of them per kernel
 *           (with rip_tune.wave_stage, first to one BFS per wave expanding 64
 *           queued vertices at a time, and to the grid past 4096 vertices)
 *   WAVE    one reverse BFS per wave for every sample (the plain form), sets
 *           passing 4096 vertices finished by a bitmap-visited BFS per wave
 *   COLOUR  64-sample colour groups for every sample
 *   LANE    one reverse BFS per lane, then one per wave for samples passing
 *           queue_cap vertices and the bitmap-visited BFS per wave for those
 *           passing 4096 */
#define RIP_SAMPLER_AUTO 0
#define RIP_SAMPLER_WAVE 1
#define RIP_SAMPLER_COLOUR 2
#define RIP_SAMPLER_LANE 3

/* Timing and size record of one IMM run; times in milliseconds.  Mirrors
 * IMMExecutionRecord (include/ripples/imm_execution_record.h:55-92). */
typedef struct {
  uint64_t theta;            /* Theta */
  uint64_t num_sets;         /* |RR| at the end */
  uint64_t rrr_elements;     /* total vertices stored in the pool */
  uint32_t rounds;           /* theta-estimation rounds run */
  uint32_t num_seeds;
  double lb;                 /* lower bound LB of OPT */
  double final_f;            /* coverage fraction of the seeds on RR[0,theta) */
  double total_ms;           /* Total */
  double estimation_ms;      /* ThetaEstimation */
  double generate_ms;        /* GenerateRRRSets (final round) */
  double select_ms;          /* FindMostInfluentialSet (final) */
  int64_t theta_prime[64];   /* ThetaPrimeDeltas are the differences */
  double f[64];              /* coverage fraction per estimation round */
  double round_generate_ms[64];  /* ThetaEstimationGenerateRRR */
  double round_select_ms[64];    /* ThetaEstimationMostInfluential */
  uint64_t pool_bytes;       /* device bytes of the pool at the end (rip_pool_device_bytes) */
} rip_imm_record;

/* OPIMCExecutionRecord (include/ripples/opimc_execution_record.h) plus the
 * per-round bound values.  Round i: rrr_sets_generated[i] sets drawn into
 * R1 and R2 before it, covered sets of R1 (as OPIMC passes it to UpperBound)
 * and of R2, the two bounds and their ratio alpha. */
typedef struct rip_opimc_record {
  uint64_t theta_zero;           /* ThetaZero */
  uint64_t theta_max;            /* ThetaMax */
  uint64_t theta;                /* |R1| = |R2| in the last round */
  uint64_t rrr_elements;         /* elements of R1 and R2 */
  uint32_t rounds;
  uint32_t num_seeds;
  double total_ms;               /* Total */
  uint64_t rrr_sets_generated[64];  /* RRRSetsGenerated */
  double generate_ms[64];        /* GenerateRRRSets */
  double select_ms[64];          /* FindMostInfluentialSet */
  uint64_t coverage1[64];
  uint64_t coverage2[64];
  float lower[64];
  float upper[64];
  float alpha[64];
} rip_opimc_record;

/* ---------------------------------------------------------------- basics */
const char *rip_last_error(void);
int rip_version(void);
/* Number of visible HIP devices (0 without a GPU). */
int rip_device_count(int *count);

/* ---------------------------------------------------------------- graph */
/* Device graph from a transposed CSR: row v lists the in-neighbours u of v
 * (an edge u -> v of the input) with its weight, as GraphBwd stores it
 * (include/ripples/graph.h:107-130, 364-1086; neighbors() at :616-618).
 * index: uint64[num_nodes+1] (graph.h index_type), edges: uint32[num_edges],
 * weights: float[num_edges].  model selects the edge encoding (RIP_IC or
 * RIP_LT); for RIP_LT the call also derives each row's threshold
 * breakpoints on the device (4 bytes per edge), which the LT walk searches
 * instead of subtracting weights edge by edge.  The caller keeps ownership of
 * the host arrays. */
int rip_graph_create(int device, int model, uint32_t num_nodes,
                     uint64_t num_edges, const uint64_t *index,
                     const uint32_t *edges, const float *weights,
                     rip_graph **out);
int rip_graph_destroy(rip_graph *g);
int rip_graph_info(const rip_graph *g, uint32_t *num_nodes,
                   uint64_t *num_edges, int *model);

/* ---------------------------------------------------------------- pool */
/* Empty RRR pool bound to graph g and the sampling stream of `seed`
 * (lcg64.seed(seed).split(2,1), tools/imm.cc:212-214). */
int rip_pool_create(rip_graph *g, uint64_t seed, rip_pool **out);
int rip_pool_destroy(rip_pool *p);
/* Append the RRR sets of samples [size, size + count) -- the generator
 * call GenerateRRRSets(G, gen, begin, end, ...) of
 * include/ripples/generate_rrr_sets.h:217-224 /
 * StreamingRRRGenerator::generate (streaming_rrr_generator.h:1026-1146). */
int rip_pool_sample(rip_pool *p, uint64_t count, const rip_tune *tune);
/* Append `count` host-side sets (CSR: offsets uint64[count+1] with
 * offsets[0] = 0, vertices uint32[offsets[count]]); each set must be sorted,
 * duplicate-free and hold ids < num_nodes.  The copy of an existing pool to
 * the device -- LoadDataToDevice / MoveRRRSets
 * (streaming_find_most_influential.h:139-244). */
int rip_pool_append_host(rip_pool *p, uint64_t count, const uint64_t *offsets,
                         const uint32_t *vertices);
/* Autotuner: draws probe_sets samples (0: 16384 times the factor, 1 to 16,
 * that brings one candidate's draw to about 20 ms) with each candidate setting (IC: the
 * lane list capacity 64/256/1024/2048/4096 x the lane kernel's workgroup size,
 * capacities 64/256/1024 with the wave stage, then 32 against 64 samples per
 * colour group of the grid sampler
 * 256/512/1024; LT: the walk's output chunk) and keeps the fastest, which
 * the graph then uses for every sampling call given no tune (rip_pool_sample,
 * rip_imm, the server).  The choice is remembered per (model, degree skew =
 * round(log2(max in-degree / mean)), round(log2 n)); a later call on a graph
 * of the same key takes it without probing unless force is set.  *best_ms:
 * the winner's time per probe batch (0 when taken from the cache); *tried:
 * settings timed.  No counterpart call in the reference, whose GPU batch
 * size comes from a similar probe (streaming_rrr_generator.h:1150-1241). */
int rip_autotune(rip_graph *g, uint64_t probe_sets, int force, rip_tune *best,
                 double *best_ms, uint32_t *tried);
/* Forget all sets (keeps the allocations). */
int rip_pool_reset(rip_pool *p);
int rip_pool_size(const rip_pool *p, uint64_t *num_sets,
                  uint64_t *num_elements);
/* Device bytes the pool holds (sorted lists, per-set offsets and lengths,
 * the column store and its tables): what the serving budget counts.  The
 * reference's RRRsets are host vectors (rrr_sets.h:81-105); no counterpart
 * call. */
int rip_pool_device_bytes(const rip_pool *p, uint64_t *bytes);
/* The pool's store of large sets: IC samples that outgrow one BFS per wave
 * are drawn 64 at a time by colour groups, and a group whose sets hold more
 * than 2 * hot_nodes members among the graph's hot vertices (a quarter of
 * the vertices, those with the most out-edge weight) keeps those members as
 * one bit per (hot vertex, set) in a column of 64-bit words, the rest as
 * sorted lists (the reference keeps every set as a sorted list,
 * rrr_sets.h:81-105).  *groups = columns in use, *hot_nodes = column height,
 * *column_bytes = device bytes of the columns allocated. */
int rip_pool_columns(const rip_pool *p, uint32_t *groups, uint32_t *hot_nodes,
                     uint64_t *column_bytes);
/* Copy sets [first, first+count) to host memory as sorted uint32 lists in
 * CSR form: offsets uint64[count+1] (offsets[0] = 0), vertices uint32[...].
 * With vertices == NULL only offsets are written (a size query).  cap is the
 * capacity of vertices in elements. */
int rip_pool_export(rip_pool *p, uint64_t first, uint64_t count,
                    uint64_t *offsets, uint32_t *vertices, uint64_t cap);
/* Greedy max coverage over sets [0, num_sets) -- FindMostInfluentialSet
 * (include/ripples/find_most_influential.h:85-196,
 * streaming_find_most_influential.h:607-643).  Ties go to the lowest vertex
 * id, the rule of GPUMaxElement (src/gpu/find_most_influential.cc:65-72).
 * Stops after k seeds or when no vertex covers an uncovered set: every set
 * covered, or only empty sets left.  In the latter case the reference's
 * sequential loop (find_most_influential.h:129-160) goes on adding zero-gain
 * vertices in its heap order; sampled pools hold no empty sets (every set
 * holds its root), so only host-appended pools can tell the two apart.
 * seeds: uint32[k]. */
int rip_pool_select(rip_pool *p, uint64_t num_sets, uint32_t k,
                    uint32_t *seeds, uint32_t *num_seeds, uint64_t *covered);
/* Occurrence counts of every vertex over sets [0, num_sets) --
 * CountOccurrencies (include/ripples/counting.h:65-97). counts: uint32[n]. */
int rip_pool_count(rip_pool *p, uint64_t num_sets, uint32_t *counts);
/* Sets among [0, num_sets) that contain at least one of seeds[0 ..
 * num_seeds) -- FindCoverage (include/ripples/opimc.h:100-118). */
int rip_pool_coverage(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                      uint32_t num_seeds, uint64_t *covered);
/* rip_pool_coverage without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * call allocates and frees nothing: it works in a fixed 8.3 MB device array
 * of the library, one per GPU, held until the process ends.  Limits (larger
 * inputs return RIP_EINVAL): graphs of at most 2^25 vertices, at most 65536
 * seeds, and for pools with column sets (rip_pool_columns) at most 2^25 sets.
 * Every call on one GPU shares that array, so calls that overlap on a GPU
 * race; calls on one pool are ordered by its stream. */
int rip_pool_coverage_noalloc(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                              uint32_t num_seeds, uint64_t *covered);
/* Device time of the last rip_pool_sample call's main sampling kernel (IC:
 * the lane kernel, or the wave kernel when the sampler kind has no lane
 * stage; LT: the walk kernel), measured with HIP events on the pool's
 * stream, and the number of kernel launches of the call. */
int rip_pool_last_kernel_ms(const rip_pool *p, double *ms, uint32_t *launches);
/* The grid sampler's share of the last rip_pool_sample call (IC, AUTO: the
 * samples that outgrew a lane): the summed device time of its BFS level
 * kernels (HIP events on the pool's stream), how many levels ran (0 when no
 * sample reached it), and the work counters of its kernels (RIP_STAT_*
 * order, as rip_pool_last_stats) into stats[0 .. count). */
int rip_pool_last_grid(const rip_pool *p, double *ms, uint32_t *levels, uint64_t *stats,
                       uint32_t count);
/* The wave stage's share of the last rip_pool_sample call (IC, AUTO with
 * rip_tune.wave_stage): the device time of its kernel (HIP events on the
 * pool's stream, 0 when it did not run) and its work counters (RIP_STAT_*
 * order) into stats[0 .. count).  No counterpart call in the reference. */
int rip_pool_last_wave(const rip_pool *p, double *ms, uint64_t *stats, uint32_t count);
/* Device time from the start of the last rip_pool_sample call's main kernel
 * to the end of its last kernel (fallback kernels, sorts and the host round
 * trips between them included), measured with HIP events. */
int rip_pool_last_call_ms(const rip_pool *p, double *ms);
/* Work counters of the last rip_pool_sample call's kernels, written to
 * stats[0 .. count) in this order (no counterpart in the reference; used for
 * roofline accounting):
 *   RIP_STAT_ALGO_BYTES  IC: bytes the gap samplers' algorithm reads and
 *                        writes: 64 per row expanded (its vertex record),
 *                        4 per candidate edge of a row whose edges do not
 *                        share one threshold, 4 per live edge (the
 *                        in-neighbour), 4 per output id;
 *                        LT: the reference's subtraction scan, 8 per visited
 *                        vertex (its CSR offsets), 8 per in-edge the loop
 *                        scans (id + weight) and 4 per output id
 *   RIP_STAT_EDGES       IC: candidate edges; LT: in-edges the subtraction
 *                        loop scans (up to and including the chosen one)
 *   RIP_STAT_COINS       IC: stream draws (gaps and acceptances);
 *                        LT: row breakpoints read
 *   RIP_STAT_EXPANSIONS  IC: rows expanded; LT: walk steps
 *   RIP_STAT_LEVELS      IC: live edges; LT: loop rounds summed over waves
 *                        (64 x this over steps is the walk kernel's lane
 *                        utilisation) */
#define RIP_STAT_ALGO_BYTES 0
#define RIP_STAT_EDGES 1
#define RIP_STAT_COINS 2
#define RIP_STAT_EXPANSIONS 3
#define RIP_STAT_LEVELS 4
#define RIP_STAT_COUNT 5
int rip_pool_last_stats(const rip_pool *p, uint64_t *stats, uint32_t count);
/* The same counters for the main kernel alone (the one rip_pool_last_kernel_ms
 * times). */
int rip_pool_last_kernel_stats(const rip_pool *p, uint64_t *stats, uint32_t count);

/* ---------------------------------------------------------------- IMM */
/* The IMM algorithm -- IMM(G, CFG, l=1, gen, record, model, omp) of
 * include/ripples/imm.h:768-860 with Sampling at :145-260: theta estimation
 * by doubling, the final theta, the final sampling and selection.
 * seeds: uint32[k]. */
int rip_imm(rip_graph *g, uint32_t k, double epsilon, uint64_t seed,
            const rip_tune *tune, uint32_t *seeds, rip_imm_record *rec);

/* OPIM-C -- OPIMC(G, CFG, 1, gen, record, model, omp) (include/ripples/
 * opimc.h:150-241) with CFG.delta = delta (the reference's default is 0.05).
 * R1 is sample ids [0, theta) of the seed's sampling stream, R2 ids
 * [2^31, 2^31 + theta): the pool of seed rip_opimc_r2_seed(seed) draws them.
 * seeds: uint32[k]; *num_seeds in the record. */
int rip_opimc(rip_graph *g, uint32_t k, double epsilon, double delta, uint64_t seed,
              const rip_tune *tune, uint32_t *seeds, rip_opimc_record *rec);
/* The seed whose pool draws sample ids 2^31, 2^31 + 1, ... of `seed`'s stream. */
uint64_t rip_opimc_r2_seed(uint64_t seed);

/* ------------------------------------------------ synthetic input (host) */
/* Directed R-MAT edge list with 2^scale vertices and edge_factor * 2^scale
 * draws, self loops and duplicates removed, sorted by (src, dst) -- the
 * generator the reference benchmarks use
 * (test/benchmarks/generate_rr_sets.cc:208-210).  With src == NULL only
 * *num_edges is computed.  Deterministic in `seed`. */
int rip_rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b,
                   double c, uint64_t seed, uint64_t *num_edges,
                   uint32_t *src, uint32_t *dst, uint64_t cap);

/* Weight schemes for rip_csr_from_edges */
#define RIP_W_UNIFORM 0   /* w = scale * u01, drawn per edge in input order */
#define RIP_W_CASCADE 1   /* w = 1 / indeg(v) (weighted cascade) */
#define RIP_W_LT_NORM 2   /* u01 per edge, normalised per target to `scale` */
#define RIP_W_CONST 3     /* w = scale */
/* Transposed CSR of an edge list: index uint64[n+1], edges uint32[m] (the
 * sources), weights float[m]; rows sorted by source.  Uniform draws come from
 * lcg64.seed(seed).split(2,0) in input-edge order (tools/imm.cc:172-174,
 * include/ripples/loaders.h:138-167). */
int rip_csr_from_edges(uint32_t num_nodes, uint64_t num_edges,
                       const uint32_t *src, const uint32_t *dst, int scheme,
                       float scale, uint64_t seed, uint64_t *index,
                       uint32_t *edges, float *weights);

/* --------------------------------------------------------------- spread */
/* Monte Carlo spread of a seed set: replicas [first, first + count) of the
 * forward IC or LT process of the graph's model from seeds[0 .. num_seeds)
 * (internal ids; duplicates allowed).  Replaces simulate() of
 * include/ripples/diffusion_simulation.h:73-178 as driven by
 * tools/simulator.cc:100-171.  Replica r draws from sample stream r of
 * `seed` (the samplers' per-sample construction): element 1 + f is the IC
 * coin of forward edge f, element 1 + v the LT threshold of vertex v; the
 * forward edges of u are the transposed entries with source u, in transposed
 * order.  activated[i] = vertices active at the end of replica first + i,
 * levels[i] = its steps (IC: breadth-first layers including the seeds'; LT:
 * iterations including the last, which activates nothing).  The first call
 * on a graph builds its forward CSR on the device. */
int rip_spread(rip_graph *g, const uint32_t *seeds, uint32_t num_seeds, uint64_t seed,
               uint64_t first, uint64_t count, uint64_t *activated, uint32_t *levels);

/* -------------------------------------------------------------- serving */
/* IMM queries served against resident graphs with continuous batching (no
 * counterpart in the reference, which runs one IMM per process).  A query
 * (graph, k, epsilon, seed) returns the seeds rip_imm(g, k, epsilon, seed)
 * returns alone.  RRR pools stay resident, one per (graph, seed): a query
 * selects only on the pool prefixes a solo run draws (theta'(x) per
 * estimation round, then theta), so prefixes drawn for earlier or concurrent
 * queries with the same key serve it unchanged and such queries share each
 * sampling launch.  Queries are admitted from the arrival order while fewer
 * than `slots` are active and, when a new pool is needed, while the resident
 * pools stay under `budget_bytes` (0 = no limit).  A new key re-uses the
 * buffers of the least recently used idle pool of its graph once the graph
 * holds `slots` pools or the budget is reached; idle pools of other graphs
 * are freed, least recently used first, while over budget.  `workers` threads step different pools concurrently,
 * each pool on its own stream.  In a served query's record, rrr_elements
 * counts the shared pool and the timings are the query's own; num_sets is the
 * pool size a solo run would have drawn.  Graphs must outlive the server. */
typedef struct rip_server rip_server;
int rip_server_create(uint32_t slots, uint32_t workers, uint64_t budget_bytes, rip_server **out);
int rip_server_add_graph(rip_server *s, rip_graph *g, uint32_t *graph_id);
int rip_server_submit(rip_server *s, uint32_t graph_id, uint32_t k, double epsilon,
                      uint64_t seed, uint64_t *query_id);
/* Waits up to wait_ms for the query; *done = 1 when its result was written
 * (seeds[0 .. *num_seeds), the record, the latency from submission), after
 * which the query is forgotten; *done = 0 on time-out. */
int rip_server_poll(rip_server *s, uint64_t query_id, double wait_ms, int *done, uint32_t *seeds,
                    uint32_t cap, uint32_t *num_seeds, rip_imm_record *rec, double *latency_ms);
/* Counters in this order (rip_server_stats copies the first `count`). */
enum {
  RIP_SERVER_LAUNCHES = 0,   /* sampling launches */
  RIP_SERVER_SETS_DRAWN,     /* RRR sets drawn */
  RIP_SERVER_SELECTIONS,     /* selector runs */
  RIP_SERVER_POOL_BYTES,     /* device bytes of the resident pools */
  RIP_SERVER_POOLS,          /* resident pools */
  RIP_SERVER_REUSED,         /* queries admitted onto an existing pool */
  RIP_SERVER_EVICTED,        /* pools freed for the budget */
  RIP_SERVER_RECYCLED,       /* idle pools re-keyed to a new seed */
  RIP_SERVER_STAT_COUNT
};
int rip_server_stats(rip_server *s, uint64_t *out, uint32_t count);
/* Finishes every submitted query, then frees the pools. */
int rip_server_destroy(rip_server *s);

/* ---------------------------------------------------------------- files */
/* Host transposed CSR read from a file: the arrays rip_graph_create takes,
 * plus the map from internal to original vertex ids.  Replaces loadGraph /
 * loadEdgeList (include/ripples/loaders.h:322-418) and Graph::load_binary /
 * dump_binary (include/ripples/graph.h:700-771, 854-995). */
typedef struct rip_host_graph rip_host_graph;

/* Weight distributions for unweighted edge lists (--distribution). */
#define RIP_DIST_UNIFORM 0 /* u01 * scale_factor */
#define RIP_DIST_NORMAL 1  /* normal(mean, scale = variance) truncated to [0,1],
                              times scale_factor */
#define RIP_DIST_CONST 2   /* mean (in [0, 1]) */

typedef struct {
  int model;          /* RIP_IC or RIP_LT: under LT the drawn weights of each
                         source are divided by (their sum + one more draw),
                         loaders.h:169-191 */
  int weighted;       /* the third column of each line is the weight (-w) */
  int undirected;     /* each line also gives the reverse edge (-u) */
  int renumber;       /* ids mapped to their rank (default); 0 keeps ids and
                         n = max id + 1 (--disable-renumbering) */
  int distribution;   /* RIP_DIST_* for unweighted inputs */
  float mean;         /* --mean (0.5) */
  float variance;     /* --variance (1.0) */
  float scale_factor; /* --scale-factor, in (0, 1] */
  uint64_t seed;      /* weights stream: lcg64(seed).split(2, 0), draw i =
                         master output 2i (tools/imm.cc:172-174 use 0) */
} rip_load_options;

/* Edge list in TSV form ("src dst" or, weighted, "src dst w" per line; empty
 * lines and lines containing '%' or '#' are skipped; a malformed line is an
 * error).  Rows list the sources of the edges into each vertex in input
 * order, duplicates kept (graph.h:506-599 with one thread). */
int rip_load_edge_list(const char *path, const rip_load_options *opts,
                       rip_host_graph **out);
/* Reference binary CSR (--reload-binary); must be the backward (transposed)
 * direction.  Weights are multiplied by scale_factor (in (0, 1]). */
int rip_load_binary(const char *path, float scale_factor, rip_host_graph **out);
/* Writes the reference binary CSR (backward direction, float weights). */
int rip_dump_binary(const rip_host_graph *g, const char *path);
/* Sizes and arrays of a host graph; the pointers stay valid until
 * rip_host_graph_destroy.  reverse_map[v] is v's original id. */
int rip_host_graph_info(const rip_host_graph *g, uint32_t *num_nodes, uint64_t *num_edges);
int rip_host_graph_arrays(const rip_host_graph *g, const uint64_t **index,
                          const uint32_t **edges, const float **weights,
                          const uint32_t **reverse_map);
int rip_host_graph_destroy(rip_host_graph *g);

#ifdef __cplusplus
}
#endif
#endif
