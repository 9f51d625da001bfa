// Internal objects behind the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "common.h"
#include "lt_scan.h"
#include "ripples_amd.h"

// Host transposed CSR read from a file (loaders.cpp).
struct rip_host_graph {
  std::vector<uint64_t> index;
  std::vector<uint32_t> edges;
  std::vector<float> weights;
  std::vector<uint32_t> reverse_map;
};

namespace rip {

// Scratch of the fused IC sampler: one slot per resident colour group.
struct IcSlots {
  uint32_t slots = 0;
  uint32_t n = 0;
  size_t slot_bytes = 0;
  unsigned long long *d_vp = nullptr;  // slots x n x {visited, pending}
  uint32_t *d_bm0 = nullptr;           // slots x ceil(n/32)   union bitmap
  uint32_t *d_bm1 = nullptr;           // slots x ceil(n/1024) bitmap summary
  uint32_t *d_queue = nullptr;         // slots x 3n words: frontier queues / output staging
  uint32_t *d_hub = nullptr;           // slots x 2 x hub_cap deferred-hub lists
  uint64_t hub_cap = 0;
};

// Scratch of the wave IC sampler: per resident wave a visited table and a
// BFS list; samples too large for them are handed to the colour sampler.
struct WaveScratch {
  uint64_t waves = 0;
  unsigned long long *d_table = nullptr;  // waves x table capacity
  uint32_t *d_list = nullptr;             // waves x list capacity
  uint32_t *d_epoch = nullptr;            // waves
  uint64_t cap_big = 0;
  uint32_t *d_big = nullptr;              // sample indices left over
  // large-set kernel: per resident wave a bitmap of n bits, its summary and
  // a list of up to n vertices
  uint64_t big_waves = 0;
  uint32_t big_n = 0;
  uint32_t *d_bbits = nullptr;
  uint32_t *d_bsum = nullptr;
  uint32_t *d_blist = nullptr;
  // lane kernel: per lane a visited table of 2 cap entries, a list of cap
  // vertices and an epoch; samples that outgrew a lane, sets to sort
  uint64_t lane_lanes = 0;
  uint32_t lane_cap = 0;
  unsigned long long *d_ltable = nullptr;
  uint32_t *d_llist = nullptr;
  uint32_t *d_lepoch = nullptr;
  uint64_t cap_over = 0;
  uint32_t *d_over = nullptr;
  uint32_t *d_lsort = nullptr;
};

// Scratch of the grid IC sampler (sample_ic_grid.hip): per colour group of a
// batch the visited and pending colour masks of every vertex and the union
// bitmaps; two frontier buffers of fcap items (g << 32 | v), which the output
// reuses as staging; control words and a pinned copy for the host.
struct GridScratch {
  uint32_t groups = 0, n = 0;
  uint64_t fcap = 0;
  unsigned long long *d_vis = nullptr;  // groups x n pairs (visited, pending colours)
  uint32_t *d_bm0 = nullptr, *d_bm1 = nullptr;
  unsigned long long *d_z = nullptr;   // groups x 64
  uint32_t *d_sid = nullptr;           // groups x 64
  uint32_t *d_nb = nullptr;            // groups
  unsigned long long *d_fa = nullptr, *d_fb = nullptr;
  unsigned long long *d_gctl = nullptr, *d_union = nullptr;
  unsigned long long *d_ring = nullptr;  // per-level frontier lengths and chunk counters
  unsigned long long *d_out = nullptr;   // output tables of a sub-batch
  uint64_t out_words = 0;
  unsigned long long *h_ctl = nullptr;
};

// Forward CSR and per-workgroup scratch of the spread simulator, built on
// first use (spread.hip).
struct FwdGraph {
  uint8_t *d_rec = nullptr;       // n forward records
  uint32_t *d_dst = nullptr;      // m + 1: target of each forward edge
  uint32_t *d_val = nullptr;      // m + 1: IC threshold of each forward edge
  uint32_t *d_scratch = nullptr;  // wgs x per_wg words
  uint64_t wgs = 0, per_wg = 0;
};

struct LtScratch {
  uint64_t cap_ids = 0;
  uint32_t *d_long_ids = nullptr;      // walks that outgrew the LDS + spill list (from
                                       // the front), unsorted long lists (from the back)
  uint64_t cap_spill = 0;
  uint32_t *d_spill = nullptr;         // per-lane visited-list spill area
};


// Device-resident arguments of the selector's seed rounds (select.hip).
struct Args {
  uint32_t *counts;
  const unsigned long long *inv_start;
  const uint32_t *inv;
  const unsigned long long *set_off;
  const uint32_t *set_len;
  const uint32_t *data;
  uint8_t *covered;
  unsigned long long *keys;
  unsigned long long *ctl;
  uint32_t *kill;      // dense pools: sets killed this round
  unsigned long long *partial;  // per-block maxima of the argmax pass
  uint64_t num_sets;
  uint32_t n;
  // column store (groups == 0: none)
  const uint32_t *gref;
  unsigned long long *const *colptr;
  const uint32_t *hot_index;
  const uint32_t *hot_vertex;
  unsigned long long *gkill;    // per group: colours killed this round
  uint32_t *kgroups;            // the groups with gkill != 0 (ctl[kCtlKillGroups])
  uint32_t groups;
  uint32_t hot_n;
};

struct SelectScratch {
  uint32_t n = 0;
  uint64_t cap_sets = 0, cap_elems = 0;
  uint32_t *d_counts = nullptr;        // n
  unsigned long long *d_inv_start = nullptr;  // n + 1
  unsigned long long *d_cursor = nullptr;     // n
  uint32_t *d_inv = nullptr;           // total elements: set ids by vertex
  // sparse pools: the (vertex, set) pairs of the sets, sorted by vertex into
  // the inverted index (keys in two buffers, set ids in d_inv and d_pval)
  uint64_t cap_pairs = 0;
  uint32_t *d_pkey[2] = {nullptr, nullptr};
  uint32_t *d_pval = nullptr;
  unsigned long long *d_soff = nullptr;  // cap_sets + 1: each set's first pair
  void *d_stemp = nullptr;               // the sort's temporary storage
  size_t stemp_bytes = 0;
  uint32_t *d_kill = nullptr;          // cap_kill: dense pools' kill list
  uint64_t cap_kill = 0;
  uint8_t *d_covered = nullptr;        // cap_sets
  unsigned long long *d_keys = nullptr;  // per-seed argmax keys (k)
  unsigned long long *d_partial = nullptr;  // kArgmaxBlocks per-block maxima
  uint32_t cap_k = 0;
  // The k rounds of argmax + kill as one HIP graph, captured for graph_k
  // seeds; its kernels read the buffer pointers from d_args, so the graph
  // stays valid when the buffers above are reallocated.
  struct Args *d_args = nullptr;
  hipGraphExec_t graph = nullptr;
  uint32_t graph_k = 0;
  bool graph_dense = false;
  bool graph_cols = false;
  // column store: per group the colours of the selected prefix, the colours
  // killed in the current round, and the list of groups with such kills
  uint32_t cap_groups = 0;
  unsigned long long *d_gactive = nullptr;
  unsigned long long *d_gkill = nullptr;
  uint32_t *d_kgroups = nullptr;
};

}  // namespace rip

struct rip_graph {
  int device = 0;
  int model = RIP_IC;
  uint32_t n = 0;
  uint64_t m = 0;
  uint32_t max_degree = 0;
  rip::Affine step{1, 0};           // one step of a sample stream
  rip::VertexRec *d_vrec = nullptr;  // n records
  uint32_t *d_ids = nullptr;         // m + kEdgePad: in-neighbour of each edge
  uint32_t *d_val = nullptr;         // m + kEdgePad: IC threshold | LT weight bits
  rip::LtRec *d_lrec = nullptr;      // LT: n walk records
  uint32_t *d_bp = nullptr;          // LT: (row breakpoint, in-neighbour) pairs, rows
                                     // padded to multiples of 16 pairs
  uint32_t *d_bidx = nullptr;        // LT: fanout-16 index of rows longer than 16
  uint64_t deg_ge_pow2[33] = {};     // vertices with in-degree >= 2^b
  // IC: the "hot" vertices of the column store (pool.hip): the hot_n
  // vertices with the largest sum of out-edge thresholds (ties to the lower
  // id), numbered in ascending vertex order; built on first use
  uint32_t hot_n = 0;
  uint32_t *d_hot_index = nullptr;   // n: index among the hot vertices, or kNotHot
  uint32_t *d_hot_vertex = nullptr;  // hot_n: the hot vertices, ascending
  // Sampler settings chosen by the autotuner (tune.cpp), used by sampling
  // calls that pass none.
  rip_tune tuned{};
  bool tuned_set = false;
  // Mean set size (pool elements per set) of the last sampling call of at
  // least kSizeHintMin sets by any pool of this graph, 0 before one: a new
  // pool sizes its first call's data with it.  Guarded by sampler_mu.
  double set_size_hint = 0.0;
  // Sampler scratch shared by every pool of this graph; pool_sample holds
  // sampler_mu while a launch uses it.
  std::mutex sampler_mu;
  rip::IcSlots ic;
  rip::WaveScratch wave;
  rip::GridScratch grid;
  rip::LtScratch lt;
  rip::FwdGraph fwd;
};

namespace rip {
constexpr uint32_t kNotHot = 0xFFFFFFFFu;
constexpr uint32_t kNoGroup = 0xFFFFFFFFu;
constexpr uint32_t kColBlock = 16;  // column groups per device allocation
}  // namespace rip

struct rip_pool {
  rip_graph *g = nullptr;
  uint64_t seed = 0;
  hipStream_t stream = nullptr;
  uint64_t num_sets = 0;
  uint64_t cap_sets = 0;
  unsigned long long *d_off = nullptr;  // per-sample start in d_data
  uint32_t *d_len = nullptr;            // per-sample size
  uint32_t *d_data = nullptr;
  uint64_t cap_data = 0;
  uint64_t used_data = 0;     // next free element of d_data
  uint64_t num_elements = 0;  // sum of set sizes: used_data less the gaps the
                              // LT sampler leaves after its waves' output chunks
  // column store (pool.hip): groups of up to 64 large sets; set s of a
  // group keeps its members among the graph's hot vertices as bit c of the
  // group's column (one 64-bit word per hot vertex) and the rest as its list
  uint32_t groups = 0, cap_groups = 0;
  std::vector<unsigned long long *> col_blocks;  // kColBlock columns each
  unsigned long long **d_colptr = nullptr;       // cap_groups column pointers
  uint32_t *d_gsample = nullptr;                 // cap_groups x 64 sample ids (kNoGroup: none)
  uint32_t *d_gref = nullptr;                    // cap_sets: group << 6 | colour, or kNoGroup
  unsigned long long *d_ctl = nullptr;  // control words, see pool.hip
  unsigned long long *h_ctl = nullptr;  // pinned host copy of the control words
  rip::SelectScratch sel;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  hipEvent_t ev3 = nullptr, ev4 = nullptr;  // around each grid-sampler level
  double grid_ms = 0.0;       // last sampling call: the grid sampler's level kernels
  uint32_t grid_levels = 0;   // and how many ran
  double wave_ms = 0.0;       // last sampling call: the wave stage's kernel (AUTO)
  bool wave_ran = false;
  uint64_t wave_stats[RIP_STAT_COUNT] = {};  // last_stats as they stood after it
  double last_kernel_ms = 0.0;
  double last_call_ms = 0.0;
  uint32_t last_launches = 0;
  uint64_t last_stats[RIP_STAT_COUNT] = {};
  uint64_t main_stats[RIP_STAT_COUNT] = {};  // the main kernel's share of last_stats
  double avg_set_size = 8.0;            // running estimate for capacity planning
  bool sized = false;                   // avg_set_size measured by a call
};

namespace rip {

// control word indices in rip_pool::d_ctl
enum : int {
  kCtlGroupCounter = 0,  // work counter of the persistent samplers
  kCtlDataBump = 1,      // used elements of d_data
  kCtlOverflow = 2,      // nonzero when d_data ran out
  kCtlLongCount = 3,     // LT: walks that outgrew the per-lane list
  kCtlCovered = 4,       // selector: sets covered so far
  kCtlStatBytes = 5,     // stats of the last sampler launch (RIP_STAT_* order)
  kCtlStatEdges = 6,
  kCtlStatCoins = 7,
  kCtlStatExpansions = 8,
  kCtlStatLevels = 9,
  kCtlSortCount = 10,    // LT: finished lists longer than the LDS part, to sort
  kCtlBigCounter = 11,   // IC: work counter of the large-set kernel
  kCtlKillSets = 12,     // selector, dense pools: sets killed this round
  kCtlKillElems = 13,    // selector, dense pools: their elements
  kCtlElems = 14,        // elements of the sets of the last sampling launch
  kCtlKillGroups = 15,   // selector, column store: groups with sets killed this round
  kCtlGroupCounter2 = 16,  // IC colour groups: columns taken by the launch
  kCtlScratch2 = 17,     // a word kernels may count into and nobody reads
  kCtlWords = 19,        // the last word is scratch too
};

// graph.hip
void graph_create(rip_graph *g, int device, int model, uint32_t n, uint64_t m,
                  const uint64_t *index, const uint32_t *edges,
                  const float *weights);
void graph_destroy(rip_graph *g);

// pool.hip
void pool_create(rip_pool *p, rip_graph *g, uint64_t seed);
void pool_create_resources(rip_pool *p);  // stream, events, control words
void pool_destroy(rip_pool *p);
// Device bytes the pool holds: lists, per-set offsets, lengths and group
// references, columns and their tables.
uint64_t pool_device_bytes(const rip_pool *p);
void pool_reserve_sets(rip_pool *p, uint64_t cap_sets);
void pool_reserve_data(rip_pool *p, uint64_t cap_data);
// Column store: the graph's hot vertices (built on first use), room for
// `groups` column groups, and d_gref set to kNoGroup for sets [first, end).
void graph_ensure_hot(rip_graph *g, hipStream_t st);
void pool_reserve_groups(rip_pool *p, uint32_t groups);
void pool_clear_gref(rip_pool *p, uint64_t first, uint64_t end);
// Resident blocks per CU of a kernel at a block size (at least 1), and the
// CU count of the current device; both cached, so launchers can ask on
// every call.
int blocks_per_cu(const void *kernel, int block);
int cu_count();
// *out += sum of v[0, n) on stream st (out is device memory)
void launch_sum_u32(const uint32_t *v, uint64_t n, unsigned long long *out, hipStream_t st);
void pool_sample(rip_pool *p, uint64_t count, const rip_tune *tune);
// Reads the control words into p->h_ctl (synchronising the pool's stream);
// with main set, also keeps the sampling counters as the main kernel's.
void pool_read_ctl(rip_pool *p, bool main);
void pool_append_host(rip_pool *p, uint64_t count, const uint64_t *offsets,
                      const uint32_t *vertices);
void pool_export(rip_pool *p, uint64_t first, uint64_t count,
                 uint64_t *offsets, uint32_t *vertices, uint64_t cap);

// sample_ic.hip
struct IcLaunch {
  uint32_t colors, block, slots, queue_cap, heavy_degree, sampler;
  bool colors_auto;    // colours not set by the caller
  uint32_t out_chunk;  // wave sampler's output chunk, 0 = automatic
  bool block_auto;     // block size not set by the caller
  bool slots_auto;     // slots not set by the caller
};

// Pool elements a sampler wave reserves at a time: a power of two in
// [64, max_chunk], at most half of what one busy wave writes, so a wave
// reserves a few times and leaves at most one partial chunk unused.  Waves
// that took one global atomic per finished set or walk spent a large part of
// their time waiting on that one address.
inline uint32_t out_chunk_rule(double elems_per_wave, uint32_t max_chunk) {
  uint32_t c = 64;
  while (c < max_chunk && 4.0 * c <= elems_per_wave) c *= 2;
  return c;
}
IcLaunch ic_tune(const rip_pool *p, const rip_tune *tune);
void ic_ensure_slots(rip_pool *p, uint32_t slots, uint32_t heavy_deg);
// Colour sampler over samples first + i, i in [0, count), or first + ids[i].
// With columns, groups whose sets hold many hot members store them in the
// pool's column store (pool.hip).
uint32_t ic_launch(rip_pool *p, uint64_t first, uint64_t count,
                   const IcLaunch &cfg, const uint32_t *ids = nullptr,
                   bool record_events = true, bool columns = false);

// IC sampling of samples [first, first+count) with the configured sampler
uint32_t ic_sample(rip_pool *p, uint64_t first, uint64_t count, const IcLaunch &cfg,
                   const rip_tune *tune);

// sample_ic_wave.hip: one BFS per wave over samples first + i, i in
// [0, count), or first + ids[i]; returns how many samples outgrew the wave's
// list, their indices (relative to first) in g->wave.d_big.  record_events:
// bracket the kernel with the pool's ev0 / ev1.  vertex_parallel: expand 64
// queued vertices at a time, a lane each, instead of one row per step.
// stage: the wave stage of AUTO, timed into p->wave_ms / p->wave_stats.
uint64_t ic_wave_launch(rip_pool *p, uint64_t first, uint64_t count, const uint32_t *ids,
                        uint32_t out_chunk, bool record_events, bool vertex_parallel = false,
                        bool stage = false);
// Most pool elements an IC sampling call of `count` samples can leave
// unused (the rest of each wave's last output chunk).
uint64_t ic_output_slack(const rip_pool *p, uint64_t count, const rip_tune *tune);
void ic_big_launch(rip_pool *p, uint64_t first, uint64_t count);

// sample_ic_grid.hip: samples first + ids[i] (first + i without ids), i in
// [0, count), 64 per colour group and many groups per batch, one BFS level
// of all groups per kernel; with columns, groups with many hot members go to
// the pool's column store.  frontier_cap limits the entries a level lists
// (0 = automatic); colors (1..64) samples share a group.  Returns the kernel
// launches made.
uint32_t ic_grid_launch(rip_pool *p, uint64_t first, uint64_t count, const uint32_t *ids,
                        bool columns, uint32_t frontier_cap, uint32_t colors);
void grid_scratch_free(GridScratch &s);

// sample_ic_lane.hip: one BFS per lane over samples [first, first + count);
// returns how many outgrew a lane's queue, their indices (relative to first)
// in g->wave.d_over.  Brackets its main kernel with the pool's ev0 / ev1.
constexpr uint32_t kLaneCapDefault = 1024;
uint32_t ic_lane_cap(const rip_tune *tune);
uint64_t ic_lane_launch(rip_pool *p, uint64_t first, uint64_t count, const rip_tune *tune);
uint64_t ic_lane_slack(const rip_pool *p, uint64_t count, const rip_tune *tune);

// sample_lt.hip
uint32_t lt_launch(rip_pool *p, uint64_t first, uint64_t count,
                   const rip_tune *tune);
// Finishes the nlong walks lt_launch left longer than its lists; returns the
// launches made (synchronises the pool's stream).
uint32_t lt_finish_long(rip_pool *p, uint64_t first, uint64_t count, uint64_t nlong);
// Most pool elements an LT launch of `count` samples can leave unused (the
// rest of each wave's last output chunk).
uint64_t lt_output_slack(const rip_pool *p, uint64_t count, const rip_tune *tune);

// select.hip
uint32_t select_run(rip_pool *p, uint64_t num_sets, uint32_t k,
                    uint32_t *seeds, uint64_t *covered);
void count_run(rip_pool *p, uint64_t num_sets, uint32_t *counts_host);
uint64_t coverage_run(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                      uint32_t num_seeds);
// coverage_run in a fixed device array instead of the pool's selector
// buffers: no device memory is allocated or freed
uint64_t coverage_run_noalloc(rip_pool *p, uint64_t num_sets, const uint32_t *seeds,
                              uint32_t num_seeds);

// scan.hip: exclusive scan of uint32 values into uint64 prefixes (n + 1 out)
// select_sort.hip: sorts `count` (key, value) pairs by the low `bits` bits
// of the key in the double buffers (k0, k1), (v0, v1); returns the buffer
// holding the result (0 or 1).  *temp grows to the sort's temporary storage.
int sort_pairs_u32(uint32_t *k0, uint32_t *k1, uint32_t *v0, uint32_t *v1, uint64_t count,
                   uint32_t bits, void **temp, size_t *temp_bytes, hipStream_t st);
void exclusive_scan_u32(const uint32_t *d_in, unsigned long long *d_out,
                        uint64_t n, hipStream_t stream);

// tune.cpp: the candidate settings that draw `probe` samples fastest on g
// (or, unless force, those found before for a graph of the same model and
// degree skew); g then uses them when a sampling call passes none.
void autotune(rip_graph *g, uint64_t probe, bool force, rip_tune *best, double *best_ms,
              uint32_t *tried);

// imm.cpp
void imm_run(rip_graph *g, uint32_t k, double eps, uint64_t seed,
             const rip_tune *tune, uint32_t *seeds, rip_imm_record *rec);
double log_binomial(uint64_t n, uint64_t k);
int64_t theta_prime(int64_t x, double eps_prime, double l, uint64_t k,
                    uint64_t n);
uint64_t theta_final(double eps, double l, uint64_t k, double lb, uint64_t n);

// rmat.cpp
void rmat_edges(uint32_t scale, uint32_t edge_factor, double a, double b,
                double c, uint64_t seed, std::vector<uint32_t> &src,
                std::vector<uint32_t> &dst);
void csr_from_edges(uint32_t n, uint64_t m, const uint32_t *src,
                    const uint32_t *dst, int scheme, float scale,
                    uint64_t seed, uint64_t *index, uint32_t *edges,
                    float *weights);

// device helpers
template <typename T>
T *dev_alloc(size_t count) {
  if (count == 0) count = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, count * sizeof(T));
  if (e != hipSuccess)
    throw Error(-2, std::string("hipMalloc of ") +
                        std::to_string(count * sizeof(T)) +
                        " bytes failed: " + hipGetErrorString(e));
  return static_cast<T *>(p);
}

template <typename T>
void dev_free(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// spread.hip
void spread_run(rip_graph *g, const uint32_t *seeds, uint32_t nseeds, uint64_t seed,
                uint64_t first, uint64_t count, uint64_t *activated, uint32_t *levels);

// opimc.cpp
float opimc_theta_helper(uint64_t n, float epsilon, uint64_t k, float delta);
uint64_t opimc_theta_max(uint64_t n, float epsilon, uint64_t k, float delta);
uint64_t opimc_theta_zero(uint64_t n, float epsilon, uint64_t k, float delta);
float opimc_lower_bound(uint64_t coverage, float delta2, uint64_t n, uint64_t theta);
float opimc_upper_bound(uint64_t coverage, float delta1, uint64_t n, uint64_t theta);
uint64_t opimc_r2_seed(uint64_t seed);
void opimc_run(rip_graph *g, uint32_t k, double epsilon, double delta, uint64_t seed,
               const rip_tune *tune, uint32_t *seeds, rip_opimc_record *rec);

// server.cpp
rip_server *server_create(uint32_t slots, uint32_t workers, uint64_t budget);
uint32_t server_add_graph(rip_server *s, rip_graph *g);
uint64_t server_submit(rip_server *s, uint32_t graph, uint32_t k, double eps, uint64_t seed);
bool server_poll(rip_server *s, uint64_t id, double wait_ms, uint32_t *seeds, uint32_t cap,
                 uint32_t *num_seeds, rip_imm_record *rec, double *latency_ms);
void server_stats(rip_server *s, uint64_t *out, uint32_t count);
void server_destroy(rip_server *s);

// loaders.cpp
void load_edge_list(const char *path, const rip_load_options &o, rip_host_graph *g);
void load_binary(const char *path, float scale_factor, rip_host_graph *g);
void dump_binary(const rip_host_graph *g, const char *path);

}  // namespace rip
