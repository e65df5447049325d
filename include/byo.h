/*
 * byo.h -- C ABI of the MI355X-native BYO hot path (libbyo.so).
 *
 * One handle (byo_graph_t) owns one device-resident graph container -- a
 * static CSR or a CPMA-style packed-memory-array -- plus the HIP stream and
 * scratch memory its algorithms use.  Every entry point returns an int
 * status (BYO_OK = 0, negative on error) and never throws; the message of
 * the last error on the calling thread is returned by byo_last_error().
 * Pointers are plain host pointers unless the name says "dev".  Calls on
 * one handle are serialised by the library; different handles may be used
 * from different threads.
 *
 * Each declaration names the reference interface it replaces (paths are
 * relative to the reference repository root).  Algorithms run device-side
 * through whole-algorithm entry points because the reference's functor
 * arguments (F in map_neighbors(i, F)) cannot cross a C ABI; the host C++
 * classes in byo_graph.hpp put the functor API back on top.
 */
#ifndef BYO_H
#define BYO_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BYO_OK 0
#define BYO_E_INVAL -1 /* bad argument (null pointer, id out of range, ...) */
#define BYO_E_OOM -2   /* device allocation failed                           */
#define BYO_E_HIP -3   /* HIP runtime error                                  */
#define BYO_E_STATE -4 /* operation not valid for this container             */
#define BYO_E_NODEV -5 /* no usable gfx950 device                            */

#define BYO_NONE 0xFFFFFFFFu /* UINT_E_MAX (gbbs/macros.h): unreached / none */

typedef struct byo_graph *byo_graph_t;

/* Message of the last failing call on this thread ("" if none). */
const char *byo_last_error(void);
/* Library version string. */
const char *byo_version(void);
/* Number of visible devices (0 when none); never fails. */
int byo_device_count(void);

/* ---------------------------------------------------------------------- */
/* Containers                                                              */
/* ---------------------------------------------------------------------- */

/* Static CSR from host arrays (copied to HBM).  Replaces
 * CSR::CSR(v_data, n, m, deletion_fn, e0) (benchmarks/run_structures/
 * run_csr.cc:87-106): offsets[n+1] (uint64, -DLONG) and ids[m] (uint32),
 * each adjacency list sorted ascending without duplicates. */
int byo_csr_create(const uint64_t *offsets, const uint32_t *ids, uint64_t n,
                   uint64_t m, byo_graph_t *out);

/* Static CSR generated on the device: 2^scale vertices, edge_factor*2^scale
 * RMAT edges (a, b, c; d = 1-a-b-c) from this library's counter-based
 * generator with the given seed, symmetrised, deduplicated, self loops
 * dropped.  The graph a reference run would read from an AdjacencyGraph file
 * produced by gbbs_io (gbbs/graph_io.h:829-847 normalisation). */
int byo_csr_create_rmat(int scale, uint64_t edge_factor, double a, double b,
                        double c, uint64_t seed, byo_graph_t *out);

/* Held-out graph families for the direction selector, generated on the
 * device, symmetrised, deduplicated, self loops dropped: Erdos-Renyi with
 * 2^scale vertices and num_edges uniform random edges (splitmix64 stream of
 * seed), and the side^3 6-neighbour grid (id = x + side*(y + side*z)). */
int byo_csr_create_er(int scale, uint64_t num_edges, uint64_t seed, byo_graph_t *out);
int byo_csr_create_grid3d(uint32_t side, byo_graph_t *out);

/* CPMA-style packed-memory-array graph from host CSR arrays.  Replaces
 * PMA_graph::PMA_graph(v_data, n, m, ...) (benchmarks/run_structures/
 * run_single_pma.cc:75-89): keys (src<<32)|dst. */
int byo_cpma_create(const uint64_t *offsets, const uint32_t *ids, uint64_t n,
                    uint64_t m, byo_graph_t *out);
/* CPMA built from an existing device container's edge set (device to device). */
int byo_cpma_create_from(byo_graph_t src, byo_graph_t *out);

/* Frees the container and everything it owns (the container destructor,
 * run_csr.cc:159). */
int byo_graph_destroy(byo_graph_t g);

/* 0 = CSR, 1 = CPMA. */
int byo_graph_kind(byo_graph_t g, int *kind);
/* Graph::N() / num_vertices() (gbbs/graph.h:2179-2183, run_csr.cc:31). */
int byo_num_vertices(byo_graph_t g, uint64_t *n);
/* Graph::M() / num_edges() (gbbs/graph.h:2266-2271, run_csr.cc:32). */
int byo_num_edges(byo_graph_t g, uint64_t *m);
/* degree(i) for all vertices into deg[n] (run_csr.cc:34-36; the stored
 * degrees of Graph::finilize, gbbs/graph.h:2559-2588, always current). */
int byo_degrees(byo_graph_t g, uint32_t *deg);
/* degree(i) of one vertex. */
int byo_degree(byo_graph_t g, uint32_t v, uint32_t *deg);
/* Neighbours of v in ascending order: the sequence map_neighbors(v, F)
 * visits (run_csr.cc:38-43, run_single_pma.cc:46-61).  Writes at most cap
 * ids and the full degree to *count. */
int byo_neighbors(byo_graph_t g, uint32_t v, uint32_t *out, uint64_t cap,
                  uint64_t *count);
/* Whole edge set as host CSR: offsets[n+1], ids[m]. */
int byo_export_csr(byo_graph_t g, uint64_t *offsets, uint32_t *ids);
/* Device bytes held by the container (get_memory_size, run_csr.cc:156-158,
 * run_single_pma.cc:126-131). */
int byo_memory_size(byo_graph_t g, uint64_t *bytes);

/* AdjacencyGraph files, host only (no device needed).  Text: the tokens
 * "AdjacencyGraph" n m, n offsets, m ids, as parse_unweighted_graph reads
 * them (gbbs/graph_io.cc:35-100) and write_graph_to_file writes them
 * (gbbs/graph_io.h:634-687), one number per line.  Binary (binary != 0):
 * the -b layout of graph_io.cc:86-97 with -DLONG -- int64 {n, m, 4m},
 * uint64 offsets[n+1], uint32 ids[m].  byo_adjacency_info reads n and m
 * from the header; byo_adjacency_read fills caller arrays of n+1 and m
 * entries (n and m must match the header) and checks the CSR (monotone
 * offsets, ids < n); malformed files give BYO_E_INVAL.  byo_adjacency_write
 * writes a host CSR (e.g. from byo_export_csr). */
int byo_adjacency_info(const char *path, int binary, uint64_t *n, uint64_t *m);
int byo_adjacency_read(const char *path, int binary, uint64_t n, uint64_t m,
                       uint64_t *offsets, uint32_t *ids);
int byo_adjacency_write(const char *path, int binary, const uint64_t *offsets,
                        const uint32_t *ids, uint64_t n, uint64_t m);

/* Batch updates.  Replace Graph::insert_batch / remove_batch
 * (gbbs/graph.h:3071-3124) over PMA_graph::insert_sorted_batch /
 * remove_sorted_batch (run_single_pma.cc:63-70).  keys[b] use the
 * std::tuple<uint32,uint32> byte layout of the reference, i.e. the uint64
 * (src<<32)|dst, so a reference harness can pass its tuple array directly.
 * The batch may be unsorted with duplicates (the library sorts and
 * dedupes, as graph.h:3090-3092 does).  Edges are directed, as in the
 * reference.  Every vertex id must be < n (BYO_E_INVAL otherwise, container
 * unchanged).  CPMA only (BYO_E_STATE for CSR). */
int byo_insert_batch(byo_graph_t g, const uint64_t *keys, uint64_t b);
int byo_remove_batch(byo_graph_t g, const uint64_t *keys, uint64_t b);
/* The same with the batch already in device memory (keys_dev, b entries). */
int byo_insert_batch_dev(byo_graph_t g, const uint64_t *keys_dev, uint64_t b);
int byo_remove_batch_dev(byo_graph_t g, const uint64_t *keys_dev, uint64_t b);
/* Update statistics of a CPMA: *touched_bytes = PMA slot bytes rewritten by
 * all updates so far (8 B per key of every merged leaf and rebuilt window
 * after an insert, of every compacted leaf before a remove, of the whole
 * array on a global rebalance; the S_touched of the update byte model
 * 24 b + 2 S_touched), *rebalances = windowed or global rebalances so far,
 * *leaves = current leaf count (512 slots each).  Any pointer may be NULL.
 * No counterpart in the reference (CPMA internals are not exposed there). */
int byo_cpma_stats(byo_graph_t g, uint64_t *touched_bytes, uint64_t *rebalances, uint64_t *leaves);
/* Allocation-free variant of byo_insert_batch_dev, for callers that cannot
 * allocate device or pinned memory (tight serving loops, code under HIP
 * graph capture, harnesses that forbid hipMalloc).  No counterpart in the
 * reference.  Same result as byo_insert_batch_dev for b <= 4096 keys; the
 * sorted batch, segments and counters live in a fixed 86,528-byte device
 * array of the library (one per GPU, held until the process ends), so the
 * call issues no hipMalloc/hipFree/hipHostMalloc.  Errors: b > 4096 ->
 * BYO_E_INVAL; an id >= n -> BYO_E_INVAL; a batch that would overflow a
 * leaf (it needs a rebalance, whose buffers scale with the container) ->
 * BYO_E_STATE.  On any error the container is unchanged.  All calls on one
 * GPU share that array: calls that overlap on one GPU (from different
 * threads or handles) race and must be serialised by the caller. */
int byo_insert_batch_dev_noalloc(byo_graph_t g, const uint64_t *keys_dev, uint64_t b);
/* nbatches insert batches held back to back in device memory (batch i is
 * keys_dev[offsets[i] .. offsets[i+1]), offsets on the host), applied in
 * order as ONE captured HIP graph with no host round trip between batches:
 * each batch is sorted, deduplicated, segmented and merged into its leaves
 * on the device.  A batch that would overflow a leaf, or holds an id >= n,
 * stops the graph's updates from that batch on; the host then applies it as
 * byo_insert_batch_dev does (global rebalance, or BYO_E_INVAL with the
 * earlier batches applied) and continues with a new graph.  The result
 * equals byo_insert_batch_dev over the batches in order.  *graphs (may be
 * NULL) receives the number of graph launches and *device_ms (may be NULL)
 * the device time of the updates, graph capture excluded.  No single
 * reference call: it runs Graph::insert_batch (gbbs/graph.h:3071-3095) over a
 * stream of batches. */
int byo_insert_batches_dev(byo_graph_t g, const uint64_t *keys_dev,
                           const uint64_t *offsets, uint32_t nbatches,
                           uint32_t *graphs, double *device_ms);

/* Device buffers for callers that keep update batches resident in HBM
 * (the *_dev entry points above).  No counterpart in the reference, whose
 * batches live in host memory. */
int byo_device_alloc(uint64_t bytes, void **dev_ptr);
int byo_device_free(void *dev_ptr);
int byo_copy_to_device(void *dev_ptr, const void *host, uint64_t bytes);

/* Update stream of Batch_insert_runner (benchmarks/run_unweighted.h:781-806,
 * rMat at :730-778), generated on the device: keys[k] = (src<<32)|dst of
 * updates first .. first+count-1 for a graph of graph_n vertices.  The
 * reference's arithmetic is kept exactly (uint32 hash32 / INT32_MAX, so the
 * effective quadrant weights are skewed and endpoints stay below
 * 2^(ceil(log2 graph_n)-1)). */
int byo_update_rmat_keys(uint64_t graph_n, uint64_t first, uint64_t count,
                         uint64_t *keys);

/* ---------------------------------------------------------------------- */
/* Algorithms                                                              */
/* ---------------------------------------------------------------------- */

/* Per-level record of one BFS: frontier size, sum of frontier out-degrees
 * and the direction taken (0 sparse push, 1 dense pull, 2 empty because
 * the frontier has no out-edges -- edge_map_data.h:289). */
typedef struct {
  uint32_t levels;      /* edgeMap calls made                               */
  uint32_t reached;     /* vertices with a parent                           */
  uint64_t edges_reached; /* sum of degrees of reached vertices (m_r)       */
  double device_ms;     /* device time of the traversal (HIP events)        */
  uint64_t fsize[64];
  uint64_t fdeg[64];
  uint8_t dir[64];
  float level_ms[64];   /* per-level device time when profiling is on      */
} byo_bfs_stats;

/* Direction policies for edgeMap (edge_map_data.h:254-307). */
#define BYO_DIR_THRESHOLD 0 /* GBBS rule: dense iff |F|+deg(F) > M/20 ...    */
#define BYO_DIR_MLP 1       /* learned selector, rule as fallback            */
#define BYO_DIR_SPARSE 2    /* always sparse push (edgeMap no_dense)         */
#define BYO_DIR_DENSE 3     /* always dense pull (edgeMap dense_only)        */
#define BYO_DIR_SCHEDULE 4  /* per-level list set by byo_set_direction_schedule */
/* threshold < 0 means M/20 (the -1 default of edgeMap). */
int byo_set_direction_policy(byo_graph_t g, int policy, int64_t threshold);
/* Explicit per-level directions (0 sparse push, 1 dense pull, 3 dense
 * forward, 4 bitmap pull, 255 = the threshold rule) for levels 0..len-1
 * (len <= 64), the rule beyond; selects BYO_DIR_SCHEDULE.  Bitmap pull is
 * the dense pull that tests each in-neighbour against a 1-bit union
 * frontier before reading its per-source frontier mask (the shared
 * multi-source traversal; the one-source traversal's pull already reads a
 * bitmap, so 4 means 1 there).  The learned selector's third class is 3 for
 * the one-source engine and 4 for the shared traversal.  Used to time every
 * direction at every level when training the learned selector. */
int byo_set_direction_schedule(byo_graph_t g, const uint8_t *dirs, uint32_t len);

/* BFS from src.  Replaces BFS(G, src) (benchmarks/BFS/NonDeterministicBFS/
 * BFS.h:47-64): parents[n] with parents[src] = src and BYO_NONE for
 * unreached vertices (parents chosen by push rounds are any valid BFS
 * parent, as CAS order makes them in the reference).  parents and stats
 * may be NULL (result stays on the device). */
int byo_bfs(byo_graph_t g, uint32_t src, uint32_t *parents,
            byo_bfs_stats *stats);
/* k independent BFS runs in one call (one device launch sequence);
 * parents[k*n] (may be NULL), stats[k] (may be NULL). */
int byo_bfs_multi(byo_graph_t g, const uint32_t *srcs, uint32_t k,
                  uint32_t *parents, byo_bfs_stats *stats);
/* BFS levels (hop distance, BYO_NONE unreached) of the last byo_bfs call,
 * derived on the device from its parent tree (run_unweighted.h:142-163). */
int byo_bfs_levels(byo_graph_t g, uint32_t *levels);

#define BYO_PR_FP64 0 /* FP64 Jacobi                                        */
#define BYO_PR_FP32 1 /* FP32 values, Kahan per-vertex sums, FP64 residual  */
/* PageRank.  Replaces PageRank(G, eps, max_iters) (benchmarks/PageRank/
 * PageRank.h:125-193) with the Jacobi semantics of PageRank_edgeMap
 * (:80-123): p0 = 1/n, p' = damping * sum_{s in N(i)} p[s]/deg(s) +
 * (1-damping)/n, stop after max_iters or when the L1 change < eps
 * (eps <= 0: never early).  p[n] (FP64 on output for both precisions, may
 * be NULL), l1_log[max_iters] (may be NULL), *iters_run and *device_ms
 * (may be NULL). */
int byo_pagerank(byo_graph_t g, double damping, int max_iters, double eps,
                 int precision, double *p, double *l1_log, int *iters_run,
                 double *device_ms);
/* Personalized PageRank for k seeds in one launch sequence (FP32 values,
 * Kahan sums): p0 = e_seed, p' = damping * sum p[s]/deg(s) +
 * (1-damping)[i == seed].  p[k*n] as FP32 (may be NULL). */
int byo_ppr_multi(byo_graph_t g, const uint32_t *seeds, uint32_t k,
                  double damping, int iters, float *p, double *device_ms);
/* Batched job in ONE persistent cooperative launch: k_bfs <= 64 BFS sources
 * sharing one traversal (as byo_bfs_multi) and then k_ppr <= 64
 * personalized-PageRank seeds for ppr_iters iterations (as byo_ppr_multi),
 * over the same graph.  No single reference call covers this: it batches the
 * per-source BFS(G, src) of BFS.h:47-64 and a seeded PageRank_edgeMap
 * (PageRank.h:80-123) the way run_all (run_unweighted.h:930-1034) would run
 * them one after another.  parents[k_bfs*n], stats[k_bfs], ppr[k_ppr*n]
 * (FP32) and device_ms may be NULL. */
int byo_batch_job(byo_graph_t g, const uint32_t *bfs_srcs, uint32_t k_bfs,
                  const uint32_t *ppr_seeds, uint32_t k_ppr, double damping,
                  int ppr_iters, uint32_t *parents, byo_bfs_stats *stats,
                  float *ppr, double *device_ms);

/* Connected components.  Replaces SimpleUnionAsync(G) (benchmarks/
 * Connectivity/SimpleUnionAsync/Connectivity.h:122-136): labels[n] with
 * label = smallest vertex id of the component. */
int byo_cc(byo_graph_t g, uint32_t *labels, double *device_ms);
/* Min-label propagation (LabelPropagation/Connectivity.h:132-146) over
 * edgeMap dense-forward rounds; same labels; *rounds may be NULL. */
int byo_cc_labelprop(byo_graph_t g, uint32_t *labels, uint32_t *rounds,
                     double *device_ms);

/* Triangle count.  Replaces Triangle_degree_ordering(G, f)
 * (benchmarks/TriangleCounting/ShunTangwongsan15/Triangle.h:167-204):
 * rank by (degree, id), orient low to high rank, count sorted-list
 * intersections.  CSR only (the reference never runs it on the CPMA). */
int byo_tc(byo_graph_t g, uint64_t *count, double *device_ms);

/* Single-source betweenness centrality, bc_bfs::SSBetweennessCentrality_BFS
 * (benchmarks/SSBetweenessCentrality/Brandes/SSBetweennessCentrality.h:
 * 332-454), the variant run_all runs: out[v] = the dependency of src on v
 * in FP64 (sigma(v) * Dep(v) - 1; for src itself the vertices reached - 1),
 * NaN for vertices src does not reach, as in the reference.  device_ms
 * covers the BFS, both passes and the final formula. */
int byo_bc(byo_graph_t g, uint32_t src, double *out, double *device_ms);

/* delta::PageRankDelta (benchmarks/PageRank/PageRank.h:335-396), run_all's
 * PageRank with -delta: FP64, damping 0.85; stops when the sum of |delta|
 * falls below eps or after max_iters rounds (*rounds_run).  p[n] may be
 * null.  The reference's semantics are kept, including the first round's
 * signed test (only positive deltas stay active), so the result is not the
 * PageRank vector of byo_pagerank. */
int byo_pagerank_delta(byo_graph_t g, double eps, double local_eps, int max_iters,
                       double *p, int *rounds_run, double *device_ms);

/* Low-diameter decomposition, LDD (benchmarks/LowDiameterDecomposition/
 * MPX13/LowDiameterDecomposition.h:98-186) without permutation, as run_all
 * runs it (beta 0.2): cluster[v] = the id of v's cluster centre.  Vertices
 * become centres in id order at rounds set by generate_shifts (:12-25);
 * a vertex reached from several clusters in one round joins the smallest
 * cluster id (one of the outcomes of the reference's CAS race), so the
 * result is deterministic.  *rounds_run may be null. */
int byo_ldd(byo_graph_t g, double beta, uint32_t *cluster, uint32_t *rounds_run,
            double *device_ms);

/* Maximal independent set, MaximalIndependentSet_rootset
 * (benchmarks/MaximalIndependentSet/RandomGreedy/MaximalIndependentSet.h:
 * 119-170, run_all's "MIS"): in_mis[n] = 1 for members.  The reference
 * orders vertices by a random permutation; here the order is that of the
 * keys (high 32 bits of splitmix64(v), v), and the result is the
 * lexicographically first MIS of that order (deterministic).  *rounds the
 * rootset rounds run (may be null).  Symmetric graphs, as in the reference. */
int byo_mis(byo_graph_t g, uint8_t *in_mis, uint32_t *rounds, double *device_ms);

/* Spanner (benchmarks/Spanner/MPXV15/Spanner.h:242-261, run_all's
 * "Spanner" with k = 4): beta = ln(n) / (2k); the LDD of byo_ldd, each
 * non-centre's edge to its parent (the smallest neighbour of its cluster
 * claimed one round before it) and, for every pair of adjacent clusters
 * c1 < c2, the edge (u, v) with cluster(u) = c1, cluster(v) = c2 smallest
 * as u << 32 | v (the reference leaves both choices to CAS races).
 * *count = the number of edges; when edges is not null it receives them
 * as sorted keys (u << 32 | v) and must hold *count (else BYO_E_INVAL with
 * *count set).  Symmetric graphs. */
int byo_spanner(byo_graph_t g, uint32_t k, uint64_t *edges, uint64_t capacity, uint64_t *count,
                double *device_ms);

/* Approximate densest subgraph by peeling, WorkEfficientDensestSubgraph
 * (benchmarks/ApproximateDensestSubgraph/ApproxPeelingBKV12/
 * DensestSubgraph.h:8-145, run_all's default eps 0.001): *density = the
 * largest sum_R D / |R| over the peeling rounds (both directions of an edge
 * counted, as the reference prints it), *rounds the rounds run (may be
 * null).  Symmetric graphs. */
int byo_densest_subgraph(byo_graph_t g, double eps, double *density, uint32_t *rounds, double *device_ms);

/* Greedy colouring, Coloring (benchmarks/GraphColoring/Hasenplaugh14/
 * GraphColoring.h:71-146, run_all's "Coloring", its default LLF order):
 * color[n].  Vertices are coloured in the order larger log2_up(degree)
 * first, ties by the keys of byo_mis (the reference breaks them by a random
 * permutation), each taking coloring::color's value: the smallest colour
 * in [0, deg) no neighbour holds, deg + 1 when all are taken, 0 when
 * isolated.  *rounds the rootset rounds run (may be null).  Symmetric
 * graphs, as in the reference. */
int byo_coloring(byo_graph_t g, uint32_t *color, uint32_t *rounds, double *device_ms);

/* Coreness of every vertex, KCore (benchmarks/KCore/JulienneDBS17/
 * KCore.h:8-56, run_all's "KCore"): core[n]; *k_max the largest coreness
 * (the reference prints it), *rounds the peeling rounds run (either may be
 * null).  Symmetric graphs, as in the reference. */
int byo_kcore(byo_graph_t g, uint32_t *core, uint32_t *k_max, uint64_t *rounds,
              double *device_ms);

/* Test_Map (benchmarks/Map/Base/Map.h:8-18): sum of dst over all edges. */
int byo_map_sum(byo_graph_t g, uint64_t *sum, double *device_ms);
/* Test_Map_Remote (benchmarks/Map/RemoteRead/Map.h:22-34, run as
 * Map_withRemote_<read_size> by run_unweighted.h:50-125): the sum over all
 * edges of remote[dst].  remote is a host array of n elements of read_size
 * bytes (0, 1, 2, 4, 8, 16, 32 or 64; 0 = one bit per vertex, bit v at byte
 * v/8, bit v%8; elements wider than 8 bytes count the little-endian uint64
 * at their start, as wide_int<size_t, read_size> does).  It is copied to
 * the device first; device_ms covers the edge scan only. */
int byo_map_remote(byo_graph_t g, const void *remote, int read_size, uint64_t *sum,
                   double *device_ms);

/* ---------------------------------------------------------------------- */
/* Profiling: HIP events around each launch of the named device phases.    */
/* ---------------------------------------------------------------------- */
int byo_profile_enable(byo_graph_t g, int on);
/* Accumulated launches and device milliseconds of one phase ("bfs",
 * "pr_fp32", "pr_fp64", "cc", "tc_count", "map", "cpma_insert", ...). */
int byo_profile_get(byo_graph_t g, const char *phase, uint64_t *launches,
                    double *total_ms);
int byo_profile_reset(byo_graph_t g);

/* ---------------------------------------------------------------------- */
/* Learned direction selector (MLP 6->32->32->3, FP32)                     */
/* ---------------------------------------------------------------------- */
#define BYO_MLP_NPARAM 1379
/* Tuning knobs of a handle -- block and tile sizes picked per graph by the
 * project's autotuner (byo_amd/autotune.py), which times candidates on the
 * device and stores the winners in byo_amd/tuning.json.  No counterpart in
 * the reference (its block sizes are compile-time constants).  Keys:
 *   coop_blocks_per_cu  1..4  blocks per CU of the persistent traversals (4)
 *   ppr_blocks_per_cu   1..8  blocks per CU of persistent PPR (4)
 *   pieces              1..16 word pieces per wave in vertex passes (8)
 *   ppr_pieces          1..16 word pieces per wave in the personalized
 *                       PageRank sweeps (4)
 *   pr_lanes            0,4,8,16,32 PageRank lanes per vertex (0: per
 *                       tile of 64 vertices from its largest degree)
 *   pr_grid             256..65536 PageRank pull grid (4096)
 *   pr_seg_log2         10..40 PageRank source-segment width (40: one)
 *   pr_relabel          0/1 PageRank and PageRankDelta on a copy of the
 *                       graph relabelled by descending degree, built on
 *                       the first call after each change of the graph;
 *                       ranks return in the original ids (0)
 *   pr_xcd              0/1 PageRank pull split into 8 source slices of
 *                       equal edge count, slice x summed by the workgroups
 *                       that run on XCD x (blockIdx % 8), partial sums
 *                       combined in slice order; slices built on the first
 *                       call after each change of the graph (0)
 *   kcore_blocks_per_cu 1..8  blocks per CU of the persistent k-core peel (3)
 *   sel_range_x100      0..1000 bound on the selector's normalised features
 *                       for it to override the rule, deviations x100 (300)
 *   sel_margin_x100     0..100000 logit margin x100 by which the learned
 *                       selector must beat the rule (one-source engine, 100)
 *   sel_margin_batch_x100  the same for the shared traversal (100)
 *   bfs_engine          0..2 several-source BFS: 0 shared traversal per 64
 *                       sources, 1 one persistent launch taking them in
 *                       turn, 2 host-driven level loop (0)
 * Unknown keys and out-of-range values give BYO_E_INVAL. */
int byo_set_tuning(byo_graph_t g, const char *key, int64_t value);
int byo_get_tuning(byo_graph_t g, const char *key, int64_t *value);

/* Loads selector weights (layout W1[32x6] b1 W2[32x32] b2 W3[3x32] b3,
 * row-major) and the 6-feature normalisation (mean[6], inv_std[6]). */
int byo_selector_load(byo_graph_t g, const float *params, const float *mean,
                      const float *inv_std);
/* The same for one traversal engine: BYO_SEL_SINGLE is the one-source
 * traversal (byo_bfs; byo_selector_load loads this one), BYO_SEL_BATCH the
 * shared multi-source traversal of byo_bfs_multi, whose per-level costs
 * differ and which is trained on its own timings.  An engine without
 * weights falls back to the threshold rule under BYO_DIR_MLP. */
#define BYO_SEL_SINGLE 0
#define BYO_SEL_BATCH 1
int byo_selector_load_engine(byo_graph_t g, int engine, const float *params,
                             const float *mean, const float *inv_std);
/* One training step on the device: forward, softmax cross-entropy,
 * backward and an Adam update over a minibatch x[B*6], labels[B].  The
 * parameter and moment state live in a device-side trainer object. */
typedef struct byo_trainer *byo_trainer_t;
int byo_trainer_create(const float *init_params, float lr, byo_trainer_t *out);
int byo_trainer_destroy(byo_trainer_t t);
int byo_trainer_step(byo_trainer_t t, const float *x, const int32_t *labels,
                     int batch, float *loss);
/* Forward + backward only (no update): grads[NPARAM], logits[B*3]. */
int byo_trainer_grad(byo_trainer_t t, const float *x, const int32_t *labels,
                     int batch, float *loss, float *grads, float *logits);
int byo_trainer_params(byo_trainer_t t, float *params);
/* Predicted class per row of x[B*6] (argmax of logits). */
int byo_trainer_predict(byo_trainer_t t, const float *x, int batch,
                        int32_t *cls);

#ifdef __cplusplus
}
#endif
#endif
