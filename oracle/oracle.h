/*
 * oracle.h -- CPU restatement of the BYO hot path (GBBS edgeMap + BFS,
 * PageRank, Connectivity, TriangleCounting, CPMA set semantics, update RMAT).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is linked by the
 * tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and by
 * nothing else.  The product path (byo_amd/, libbyo.so) never calls it.
 *
 * Every function cites the reference file:line whose semantics it restates.
 * Graph layout everywhere: CSR with uint64 offsets[n+1] and uint32 ids[m],
 * symmetric, each adjacency list sorted ascending, no duplicates, no self
 * loops (run_csr.cc:28-191 with -DLONG).
 */
#ifndef BYO_ORACLE_H
#define BYO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NONE 0xFFFFFFFFu /* UINT_E_MAX: unreached (BFS.h:50) */

/* ---- hashing / generators ------------------------------------------- */
uint64_t orc_splitmix64(uint64_t x);
/* batch_insert_helpers::hash64 / hash32 (run_unweighted.h:691-728) */
uint64_t orc_ref_hash64(uint64_t u);
uint32_t orc_ref_hash32(uint32_t a);

/* This framework's seeded RMAT edge generator for graph construction.
 * Edge i, level l (0 = most significant bit): r = u01(splitmix64(base + i*scale + l)),
 * base = splitmix64(seed).  r < a: (0,0); r < ab: (0,1); r < abc: (1,0);
 * else (1,1).  ab = a+b and abc = a+b+c are passed in precomputed. */
void orc_rmat_edges(int scale, uint64_t num_edges, double a, double ab,
                    double abc, uint64_t seed, uint64_t first_edge,
                    uint32_t *src, uint32_t *dst);

/* Erdos-Renyi and 3-D grid edge lists (this framework's held-out graph
 * families for the direction selector). */
void orc_er_edges(uint64_t n, uint64_t num_edges, uint64_t seed, uint32_t *src, uint32_t *dst);
uint64_t orc_grid3d_edges(uint64_t side, uint32_t *src, uint32_t *dst);

/* Reference-exact update generator of Batch_insert_runner
 * (run_unweighted.h:781-806, rMat at :730-778): nn = 2^(ceil(log2 n)-1),
 * seed = (uint32)hash64(0), r = hash32(x) / INT32_MAX. */
void orc_update_rmat(uint64_t graph_n, uint64_t first, uint64_t count,
                     uint32_t *src, uint32_t *dst);

/* ---- containers ------------------------------------------------------ */
/* Sort + unique uint64 keys in place; returns the new length. */
uint64_t orc_sort_unique_u64(uint64_t *keys, uint64_t n);
/* Symmetrize an edge list: keep u != v, emit (u,v),(v,u), sort, dedupe,
 * build CSR over exactly n vertices (graph_io.h:829-847 semantics with n
 * fixed).  ids must hold 2*ne entries.  Returns m. */
uint64_t orc_build_sym_csr(uint64_t n, uint64_t ne, const uint32_t *src,
                           const uint32_t *dst, uint64_t *offsets,
                           uint32_t *ids);
/* Keys (src<<32)|dst of a CSR in CSR order (run_single_pma.cc:83). */
void orc_csr_to_keys(uint64_t n, const uint64_t *off, const uint32_t *ids, uint64_t *keys);
/* Graph::insert_batch / remove_batch (gbbs/graph.h:3071-3124) over a sorted
 * array: batch sorted + deduplicated in place, merged with the sorted unique
 * E[m] into out (E u B, or E minus B); returns |out| (<= m + b).  OpenMP. */
uint64_t orc_batch_insert(const uint64_t *E, uint64_t m, uint64_t *batch, uint64_t b, uint64_t *out);
uint64_t orc_batch_remove(const uint64_t *E, uint64_t m, uint64_t *batch, uint64_t b, uint64_t *out);
/* CSR over sorted unique keys (src<<32|dst) -> offsets[n+1], ids[len]. */
void orc_keys_to_csr(uint64_t n, const uint64_t *keys, uint64_t len,
                     uint64_t *offsets, uint32_t *ids);

/* ---- BFS ------------------------------------------------------------- */
/* Textbook FIFO BFS levels (ORC_NONE for unreached). */
void orc_bfs_levels(uint64_t n, const uint64_t *off, const uint32_t *ids,
                    uint32_t src, uint32_t *level);
/* Direction-optimizing BFS restating BFS.h:47-64 over edgeMapData
 * (edge_map_data.h:254-307), OpenMP-parallel.  threshold < 0 means M/20.
 * dirlog[k] = 1 if level k ran dense pull, 0 if sparse push.  Also logs
 * |F| and sum of out-degrees of F per level (fsize, fdeg; may be NULL).
 * Returns the number of edgeMap calls (levels). */
int orc_bfs_diropt(uint64_t n, uint64_t m, const uint64_t *off,
                   const uint32_t *ids, uint32_t src, int64_t threshold,
                   uint32_t *parents, uint8_t *dirlog, uint64_t *fsize,
                   uint64_t *fdeg, int max_log);
/* Validity of a BFS parent tree against exact levels: parent[src]=src,
 * parent[v] in N(v) with level[parent]=level[v]-1, unreached = ORC_NONE.
 * Returns 0 when valid, else 1 + index of the first bad vertex. */
uint64_t orc_check_bfs_tree(uint64_t n, const uint64_t *off,
                            const uint32_t *ids, uint32_t src,
                            const uint32_t *parents, const uint32_t *level);

/* ---- PageRank -------------------------------------------------------- */
/* FP64 Jacobi PageRank restating PageRank_edgeMap (PageRank.h:80-123),
 * evaluated as pull: p0 = 1/n, p'[i] = d * sum_{s in N(i)} p[s]/deg(s)
 * + (1-d)/n.  Stops after max_iters or when L1 < eps (eps <= 0 never stops
 * early).  Writes the last iterate to p, per-iteration L1 to l1_log (may be
 * NULL).  Returns iterations run. */
int orc_pagerank_f64(uint64_t n, const uint64_t *off, const uint32_t *ids,
                     double damping, int max_iters, double eps, double *p,
                     double *l1_log);
/* Personalized PageRank from one seed (teleport to the seed vertex):
 * p0 = e_seed, p'[i] = d * sum p[s]/deg(s) + (1-d)[i == seed]. */
int orc_ppr_f64(uint64_t n, const uint64_t *off, const uint32_t *ids,
                uint32_t seed, double damping, int iters, double *p);
/* PageRank() (PageRank.h:125-193) under a sequential schedule: the
 * Gauss-Seidel-like in-sweep update of p_div. */
int orc_pagerank_ref_order(uint64_t n, const uint64_t *off,
                           const uint32_t *ids, double damping, int max_iters,
                           double eps, double *p);

/* ---- Connectivity ---------------------------------------------------- */
/* SimpleUnionAsync (SimpleUnionAsync/Connectivity.h:32-136), sequential:
 * link larger root under smaller, finish() compresses -> min id label. */
void orc_cc_union_find(uint64_t n, const uint64_t *off, const uint32_t *ids,
                       uint32_t *labels);
/* Min-label propagation (LabelPropagation/Connectivity.h:78-146). */
int orc_cc_labelprop(uint64_t n, const uint64_t *off, const uint32_t *ids,
                     uint32_t *labels);

/* ---- Triangle counting ----------------------------------------------- */
/* Triangle_degree_ordering (Triangle.h:167-204): rank by (degree, id),
 * orient low->high rank, count sum |N+(u) & N+(v)| by sorted merge
 * (intersect.h:68-150).  OpenMP-parallel. */
uint64_t orc_tc(uint64_t n, const uint64_t *off, const uint32_t *ids);
/* Brute force: sum over u<v<w with all three edges present. */
uint64_t orc_tc_brute(uint64_t n, const uint64_t *off, const uint32_t *ids);
/* Oriented graph DG as built by filterGraph (graph_mutation.h:15-81) with
 * rank (deg,id): dg_off[n+1], dg_ids[m/2]. */
void orc_orient(uint64_t n, const uint64_t *off, const uint32_t *ids,
                uint64_t *dg_off, uint32_t *dg_ids);

/* ---- Map ------------------------------------------------------------- */
/* Test_Map (Map/Base/Map.h:8-18): sum of dst over all CSR entries. */
void orc_bc(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t src, double *out);
void orc_kcore(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t *core);
/* MaximalIndependentSet_rootset (MaximalIndependentSet.h:119-170) over the
 * priority order of orc_mis_key: the lexicographically first MIS. */
uint64_t orc_mis_key(uint32_t v);
void orc_mis(uint64_t n, const uint64_t *off, const uint32_t *ids, uint8_t *in_mis);
/* Coloring(G, lf = false) (GraphColoring/Hasenplaugh14/GraphColoring.h:
 * 71-146): greedy colouring in the LLF order with coloring::color's rule. */
void orc_coloring(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t *color);
int orc_llf_before(const uint64_t *off, uint32_t u, uint32_t v);
/* WorkEfficientDensestSubgraph (ApproxPeelingBKV12/DensestSubgraph.h:8-145):
 * the largest density (2|E(R)|/|R|) over the peeling rounds. */
double orc_densest(uint64_t n, const uint64_t *off, const uint32_t *ids, double eps, uint32_t *rounds);
/* Spanner (Spanner/MPXV15/Spanner.h:242-261), beta = ln(n)/(2k): sorted
 * edge keys (u << 32 | v) into out (capacity n + m); returns the count. */
uint64_t orc_spanner(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t k, uint64_t *out);
int orc_pagerank_delta(uint64_t n, const uint64_t *off, const uint32_t *ids, double eps, double local_eps,
                       int max_iters, double *p);
void orc_ldd(uint64_t n, const uint64_t *off, const uint32_t *ids, double beta, uint32_t *cluster, uint32_t *round);
uint64_t orc_map_sum(uint64_t m, const uint32_t *ids);

/* ---- Learned selector: MLP 6->32->32->3 reference math --------------- */
/* Forward + backward of one minibatch: softmax cross-entropy over 3
 * classes, ReLU hidden layers.  params layout: W1[32*6] b1[32] W2[32*32]
 * b2[32] W3[3*32] b3[3] (row-major, out x in).  grads has the same layout.
 * Returns mean loss.  logits (may be NULL) receives [B*3]. */
#define ORC_MLP_NPARAM (32 * 6 + 32 + 32 * 32 + 32 + 3 * 32 + 3)
double orc_mlp_forward_backward(const float *params, const float *x,
                                const int32_t *labels, int batch,
                                float *grads, float *logits);

#ifdef __cplusplus
}
#endif
#endif
