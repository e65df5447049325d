/*
 * oracle.c -- CPU restatement of the BYO hot path.  TEST INFRASTRUCTURE
 * ONLY: linked by tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg as the checker and the CPU baseline, never by the
 * product library.  Each function cites the reference file:line it follows
 * (paths relative to the reference repository root).
 *
 * Pinning (how each function is checked independently of its own formula):
 *   orc_ref_hash32/64, orc_update_rmat : hand-evaluated known answers for
 *       i = 0, 1 and a committed golden prefix (tests/golden/).
 *   orc_build_sym_csr : the reference's own graph_io_test.cc normalisation
 *       cases and the inputs/ fixtures (n, m, degrees).
 *   orc_bfs_levels    : textbook FIFO BFS; inputs/ facts (reached count,
 *       level count); test_bfs.cc parent cases.
 *   orc_bfs_diropt    : levels equal orc_bfs_levels on every graph;
 *       direction log equals the closed-form rule of edge_map_data.h:257-306.
 *   orc_pagerank_f64  : closed form on a star (2x2 linear system), dense
 *       power iteration on <=128-vertex fixtures, per-iteration mass law.
 *   orc_cc_*          : brute-force DSU; test_cc.cc partitions.
 *   orc_tc            : orc_tc_brute (independent triple enumeration) and
 *       inputs/ facts (107, 2, 0 triangles).
 *   orc_mlp_*         : central finite differences of the loss.
 *   orc_kcore         : test_kcore.cc vectors; brute force of the definition
 *       (largest k whose k-core, by repeated removal, holds the vertex).
 *   orc_bc            : brute-force dependencies from all-pairs shortest-path
 *       counts (textbook definition) on small graphs; star and path closed
 *       forms; the source's own value = reached - 1.
 *   orc_spanner       : checked without its own code on small graphs: a
 *       subgraph of G with G's components, one tree edge per non-centre to
 *       a neighbour of its LDD cluster one round nearer the centre, and for
 *       every pair of adjacent clusters exactly their smallest edge.
 *   orc_densest       : closed forms (clique k - 1, star 2n/(n+1)); brute
 *       force over all subsets of small graphs: the result is the density
 *       of a subset (<= the optimum) and within 2(1+eps) of the optimum.
 *   orc_coloring      : a proper colouring, and every colour equals the
 *       rule applied to the neighbours earlier in the LLF order (checked
 *       directly, which fixes the colouring); hand-worked triangle and star.
 *   orc_mis           : test_mis.cc cases; the characterisation of the
 *       lexicographically first MIS checked directly (independent; every
 *       vertex outside has a neighbour inside of smaller key; every vertex
 *       inside has none of smaller key inside) -- it fixes the set.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ===================================================================== */
/* hashing / generators                                                  */
/* ===================================================================== */

uint64_t orc_splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* run_unweighted.h:691-702 */
uint64_t orc_ref_hash64(uint64_t u) {
  uint64_t v = u * 3935559000370003845ull + 2691343689449507681ull;
  v ^= v >> 21;
  v ^= v << 37;
  v ^= v >> 4;
  v *= 4768777513237032717ull;
  v ^= v << 20;
  v ^= v >> 41;
  v ^= v << 5;
  return v;
}

/* run_unweighted.h:720-728 */
uint32_t orc_ref_hash32(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

void orc_rmat_edges(int scale, uint64_t num_edges, double a, double ab,
                    double abc, uint64_t seed, uint64_t first_edge,
                    uint32_t *src, uint32_t *dst) {
  const uint64_t base = orc_splitmix64(seed);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)num_edges; k++) {
    uint64_t i = first_edge + (uint64_t)k;
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; l++) {
      uint64_t h = orc_splitmix64(base + i * (uint64_t)scale + (uint64_t)l);
      double r = (double)(h >> 11) * 0x1.0p-53;
      uint32_t bit = 1u << (scale - 1 - l);
      if (r < a) {
      } else if (r < ab) {
        v |= bit;
      } else if (r < abc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
    src[k] = u;
    dst[k] = v;
  }
}

/* Erdos-Renyi edge i: (splitmix64(base+2i) mod n, splitmix64(base+2i+1) mod n),
 * base = splitmix64(seed).  This framework's generator (no reference
 * counterpart); pinned by its degree distribution in the tests. */
void orc_er_edges(uint64_t n, uint64_t num_edges, uint64_t seed, uint32_t *src, uint32_t *dst) {
  const uint64_t base = orc_splitmix64(seed);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)num_edges; i++) {
    src[i] = (uint32_t)(orc_splitmix64(base + 2 * (uint64_t)i) % n);
    dst[i] = (uint32_t)(orc_splitmix64(base + 2 * (uint64_t)i + 1) % n);
  }
}

/* side^3 grid, id = x + side*(y + side*z): the +x, +y, +z edges. */
uint64_t orc_grid3d_edges(uint64_t side, uint32_t *src, uint32_t *dst) {
  uint64_t k = 0;
  for (uint64_t z = 0; z < side; z++)
    for (uint64_t y = 0; y < side; y++)
      for (uint64_t x = 0; x < side; x++) {
        uint64_t id = x + side * (y + side * z);
        if (x + 1 < side) { src[k] = (uint32_t)id; dst[k++] = (uint32_t)(id + 1); }
        if (y + 1 < side) { src[k] = (uint32_t)id; dst[k++] = (uint32_t)(id + side); }
        if (z + 1 < side) { src[k] = (uint32_t)id; dst[k++] = (uint32_t)(id + side * side); }
      }
  return k;
}

static int log2_up(uint64_t n) { /* parlay::log2_up: ceil(log2 n) */
  int r = 0;
  while ((1ull << r) < n) r++;
  return r;
}

/* Batch_insert_runner's update stream (run_unweighted.h:797-806) through
 * rMat::operator() / rMatRec (run_unweighted.h:755-777).  All uint32
 * arithmetic wraps exactly as intT = uintE does there. */
void orc_update_rmat(uint64_t graph_n, uint64_t first, uint64_t count,
                     uint32_t *src, uint32_t *dst) {
  const double a = 0.5, b = 0.1, c = 0.1;
  const double ab = a + b, abc = a + b + c;
  const uint32_t nn = (uint32_t)(1ull << (log2_up(graph_n) - 1));
  const uint32_t seed = (uint32_t)orc_ref_hash64(0); /* r.ith_rand(0) */
  const uint32_t h = orc_ref_hash32(seed);
#pragma omp parallel for schedule(static)
  for (int64_t k = 0; k < (int64_t)count; k++) {
    uint32_t i = (uint32_t)(first + (uint64_t)k);
    uint32_t start = orc_ref_hash32((uint32_t)(2u * i) * h);
    uint32_t stride = orc_ref_hash32((uint32_t)(2u * i + 1u) * h);
    uint32_t row = 0, col = 0;
    for (uint32_t size = nn; size > 1; size >>= 1) {
      double r = (double)orc_ref_hash32(start) / 2147483647.0;
      if (r < a) {
      } else if (r < ab) {
        col += size >> 1;
      } else if (r < abc) {
        row += size >> 1;
      } else {
        row += size >> 1;
        col += size >> 1;
      }
      start += stride;
    }
    src[k] = row;
    dst[k] = col;
  }
}

/* ===================================================================== */
/* containers                                                            */
/* ===================================================================== */

static void radix_sort_u64(uint64_t *keys, uint64_t n) {
  if (n < 2) return;
  uint64_t *tmp = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint64_t *cnt = (uint64_t *)malloc(65536 * sizeof(uint64_t));
  uint64_t *from = keys, *to = tmp;
  for (int pass = 0; pass < 4; pass++) {
    int sh = pass * 16;
    memset(cnt, 0, 65536 * sizeof(uint64_t));
    for (uint64_t i = 0; i < n; i++) cnt[(from[i] >> sh) & 0xFFFF]++;
    int trivial = 0;
    for (int d = 0; d < 65536; d++)
      if (cnt[d] == n) trivial = 1;
    if (trivial) continue;
    uint64_t s = 0;
    for (int d = 0; d < 65536; d++) {
      uint64_t c = cnt[d];
      cnt[d] = s;
      s += c;
    }
    for (uint64_t i = 0; i < n; i++) to[cnt[(from[i] >> sh) & 0xFFFF]++] = from[i];
    uint64_t *t = from;
    from = to;
    to = t;
  }
  if (from != keys) memcpy(keys, from, n * sizeof(uint64_t));
  free(tmp);
  free(cnt);
}

uint64_t orc_sort_unique_u64(uint64_t *keys, uint64_t n) {
  radix_sort_u64(keys, n);
  if (n == 0) return 0;
  uint64_t w = 1;
  for (uint64_t i = 1; i < n; i++)
    if (keys[i] != keys[w - 1]) keys[w++] = keys[i];
  return w;
}

void orc_keys_to_csr(uint64_t n, const uint64_t *keys, uint64_t len,
                     uint64_t *offsets, uint32_t *ids) {
  memset(offsets, 0, (n + 1) * sizeof(uint64_t));
  for (uint64_t i = 0; i < len; i++) offsets[(keys[i] >> 32) + 1]++;
  for (uint64_t v = 0; v < n; v++) offsets[v + 1] += offsets[v];
  for (uint64_t i = 0; i < len; i++) ids[i] = (uint32_t)keys[i];
}

/* (src<<32)|dst keys of a CSR, in CSR order (sorted when the lists are):
 * the key layout of run_single_pma.cc:83. */
void orc_csr_to_keys(uint64_t n, const uint64_t *off, const uint32_t *ids, uint64_t *keys) {
#pragma omp parallel for schedule(dynamic, 4096)
  for (uint64_t v = 0; v < n; v++)
    for (uint64_t e = off[v]; e < off[v + 1]; e++) keys[e] = (v << 32) | ids[e];
}

/* First index of a sorted array a[0..len) holding a value >= x. */
static uint64_t lower_u64(const uint64_t *a, uint64_t len, uint64_t x) {
  uint64_t lo = 0, hi = len;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

/* Graph::insert_batch / remove_batch over a sorted-array container
 * (gbbs/graph.h:3071-3124: parlay::unique(parlay::sort(batch)), then
 * insert_sorted_batch / remove_sorted_batch): batch[b] is sorted and
 * deduplicated in place, then merged with the sorted unique edge set E[m]
 * into out (insert: E u B; remove: E minus B).  The merge is split over the
 * threads at equal positions of E (each split point found in B by binary
 * search), so every thread writes a disjoint run of out.  Returns the size
 * of out (at most m + b entries).  The CPU baseline of the CPMA updates. */
static uint64_t batch_merge(const uint64_t *E, uint64_t m, uint64_t *batch, uint64_t b, uint64_t *out,
                            int insert) {
  const uint64_t bu = orc_sort_unique_u64(batch, b);
  if (m == 0) {
    if (insert) memcpy(out, batch, bu * sizeof(uint64_t));
    return insert ? bu : 0;
  }
  const int T = omp_get_max_threads();
  uint64_t *cnt = (uint64_t *)calloc((size_t)T + 1, sizeof(uint64_t));
  uint64_t total = 0;
#pragma omp parallel num_threads(T)
  {
    const int t = omp_get_thread_num(), nt = omp_get_num_threads();
    const uint64_t e0 = m * (uint64_t)t / nt, e1 = m * (uint64_t)(t + 1) / nt;
    /* batch keys below E[e0] belong to the previous thread (or, for t = 0,
     * to this one); the last thread also takes the keys past E's end */
    const uint64_t b0 = t == 0 ? 0 : lower_u64(batch, bu, E[e0]);
    const uint64_t b1 = t == nt - 1 ? bu : lower_u64(batch, bu, E[e1]);
    for (int pass = 0; pass < 2; pass++) {
      uint64_t i = e0, j = b0, w = 0;
      uint64_t *o = NULL;
      if (pass == 1) {
#pragma omp barrier
#pragma omp single
        {
          for (int x = 0; x < nt; x++) {
            uint64_t c = cnt[x];
            cnt[x] = total;
            total += c;
          }
        }
        o = out + cnt[t];
      }
      while (i < e1 || j < b1) {
        uint64_t k;
        int keep;
        if (j >= b1 || (i < e1 && E[i] < batch[j])) {
          k = E[i++];
          keep = 1;
        } else if (i >= e1 || batch[j] < E[i]) {
          k = batch[j++];
          keep = insert;
        } else {
          k = E[i++];
          j++;
          keep = insert;
        }
        if (keep) {
          if (o) o[w] = k;
          w++;
        }
      }
      if (pass == 0) cnt[t] = w;
    }
  }
  free(cnt);
  return total;
}

uint64_t orc_batch_insert(const uint64_t *E, uint64_t m, uint64_t *batch, uint64_t b, uint64_t *out) {
  return batch_merge(E, m, batch, b, out, 1);
}

uint64_t orc_batch_remove(const uint64_t *E, uint64_t m, uint64_t *batch, uint64_t b, uint64_t *out) {
  return batch_merge(E, m, batch, b, out, 0);
}

/* graph_io.h:829-847 (sym_graph_from_edges: symmetrize, sort, dedupe, drop
 * self loops), with n fixed by the caller instead of inferred. */
uint64_t orc_build_sym_csr(uint64_t n, uint64_t ne, const uint32_t *src,
                           const uint32_t *dst, uint64_t *offsets,
                           uint32_t *ids) {
  uint64_t *keys = (uint64_t *)malloc((2 * ne + 1) * sizeof(uint64_t));
  uint64_t k = 0;
  for (uint64_t i = 0; i < ne; i++) {
    uint64_t u = src[i], v = dst[i];
    if (u == v) continue;
    keys[k++] = (u << 32) | v;
    keys[k++] = (v << 32) | u;
  }
  uint64_t m = orc_sort_unique_u64(keys, k);
  orc_keys_to_csr(n, keys, m, offsets, ids);
  free(keys);
  return m;
}

/* ===================================================================== */
/* BFS                                                                   */
/* ===================================================================== */

void orc_bfs_levels(uint64_t n, const uint64_t *off, const uint32_t *ids,
                    uint32_t src, uint32_t *level) {
  for (uint64_t i = 0; i < n; i++) level[i] = ORC_NONE;
  if (n == 0) return;
  uint32_t *q = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t head = 0, tail = 0;
  level[src] = 0;
  q[tail++] = src;
  while (head < tail) {
    uint32_t u = q[head++];
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t t = ids[e];
      if (level[t] == ORC_NONE) {
        level[t] = level[u] + 1;
        q[tail++] = t;
      }
    }
  }
  free(q);
}

static inline int bm_get(const uint32_t *bm, uint32_t i) {
  return (bm[i >> 5] >> (i & 31)) & 1u;
}

/* BFS.h:47-64 driving edgeMapData (edge_map_data.h:254-307):
 *   threshold = M/20; dense if (isDense && |F| > n/10), else compute
 *   sum of out-degrees of F; empty result when it is 0; dense when
 *   |F| + sum > threshold; otherwise sparse push (edgeMapSparse,
 *   edge_map_blocked.h:18-92).  Dense pull (edgeMapDense :47-175): every
 *   unvisited v scans N(v) in order and takes the first frontier member as
 *   parent (early exit).  Sparse push: CAS(parents[t], NONE, u) per edge. */
int orc_bfs_diropt(uint64_t n, uint64_t m, const uint64_t *off,
                   const uint32_t *ids, uint32_t src, int64_t threshold,
                   uint32_t *parents, uint8_t *dirlog, uint64_t *fsize,
                   uint64_t *fdeg, int max_log) {
  const uint64_t thr = threshold < 0 ? m / 20 : (uint64_t)threshold;
  const uint64_t nw = (n + 31) / 32;
  for (uint64_t i = 0; i < n; i++) parents[i] = ORC_NONE;
  if (n == 0) return 0;
  parents[src] = src;
  uint32_t *cur = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *nxt = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *bm = (uint32_t *)calloc(nw, sizeof(uint32_t));
  uint32_t *bm2 = (uint32_t *)calloc(nw, sizeof(uint32_t));
  uint64_t fs = 1;
  cur[0] = src;
  int is_dense = 0, have_sparse = 1, have_dense = 0;
  int lv = 0;
  int nthreads = 1;
#ifdef _OPENMP
  nthreads = omp_get_max_threads();
#endif
  uint32_t **tbuf = (uint32_t **)calloc((size_t)nthreads, sizeof(uint32_t *));
  uint64_t *tcnt = (uint64_t *)calloc((size_t)nthreads + 1, sizeof(uint64_t));
  uint64_t *tcap = (uint64_t *)calloc((size_t)nthreads, sizeof(uint64_t));

  while (fs > 0) {
    /* out-degree sum of F from whichever representation exists */
    uint64_t od = 0;
    if (have_sparse) {
#pragma omp parallel for reduction(+ : od) schedule(static)
      for (int64_t i = 0; i < (int64_t)fs; i++)
        od += off[cur[i] + 1] - off[cur[i]];
    } else {
#pragma omp parallel for reduction(+ : od) schedule(static)
      for (int64_t v = 0; v < (int64_t)n; v++)
        if (bm_get(bm, (uint32_t)v)) od += off[v + 1] - off[v];
    }
    int dense;
    if (is_dense && fs > n / 10) {
      dense = 1;
    } else if (od == 0) {
      dense = 2; /* edgeMapData returns an empty subset (:289) */
    } else {
      dense = (fs + od > thr) ? 1 : 0;
    }
    if (lv < max_log) {
      if (dirlog) dirlog[lv] = (uint8_t)dense;
      if (fsize) fsize[lv] = fs;
      if (fdeg) fdeg[lv] = od;
    }
    lv++;
    if (dense == 2) break;
    uint64_t count = 0;
    if (dense) {
      if (!have_dense) { /* toDense: set_atomic per element (:357-364) */
        memset(bm, 0, nw * sizeof(uint32_t));
        for (uint64_t i = 0; i < fs; i++) bm[cur[i] >> 5] |= 1u << (cur[i] & 31);
      }
      memset(bm2, 0, nw * sizeof(uint32_t));
#pragma omp parallel for reduction(+ : count) schedule(dynamic, 1024)
      for (int64_t v = 0; v < (int64_t)n; v++) {
        if (parents[v] != ORC_NONE) continue;
        for (uint64_t e = off[v]; e < off[v + 1]; e++) {
          uint32_t u = ids[e];
          if (bm_get(bm, u)) {
            parents[v] = u;
            __atomic_fetch_or(&bm2[v >> 5], 1u << (v & 31), __ATOMIC_RELAXED);
            count++;
            break;
          }
        }
      }
      uint32_t *t = bm;
      bm = bm2;
      bm2 = t;
      is_dense = 1;
      have_dense = 1;
      have_sparse = 0;
    } else {
      if (!have_sparse) { /* toSparse: pack_index, ascending (:339-354) */
        uint64_t k = 0;
        for (uint64_t v = 0; v < n; v++)
          if (bm_get(bm, (uint32_t)v)) cur[k++] = (uint32_t)v;
      }
#pragma omp parallel
      {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        tcnt[tid] = 0;
#pragma omp for schedule(dynamic, 64)
        for (int64_t i = 0; i < (int64_t)fs; i++) {
          uint32_t u = cur[i];
          for (uint64_t e = off[u]; e < off[u + 1]; e++) {
            uint32_t t = ids[e];
            uint32_t expect = ORC_NONE;
            if (__atomic_load_n(&parents[t], __ATOMIC_RELAXED) == ORC_NONE &&
                __atomic_compare_exchange_n(&parents[t], &expect, u, 0,
                                            __ATOMIC_RELAXED,
                                            __ATOMIC_RELAXED)) {
              if (tcnt[tid] == tcap[tid]) {
                tcap[tid] = tcap[tid] ? 2 * tcap[tid] : 1024;
                tbuf[tid] = (uint32_t *)realloc(tbuf[tid],
                                                tcap[tid] * sizeof(uint32_t));
              }
              tbuf[tid][tcnt[tid]++] = t;
            }
          }
        }
      }
      /* Reducer_Vector::get_sequence: concatenate in worker order
       * (reducer.hpp:228-240) */
      for (int t = 0; t < nthreads; t++) {
        memcpy(nxt + count, tbuf[t], tcnt[t] * sizeof(uint32_t));
        count += tcnt[t];
      }
      uint32_t *t = cur;
      cur = nxt;
      nxt = t;
      is_dense = 0;
      have_sparse = 1;
      have_dense = 0;
    }
    fs = count;
  }
  for (int t = 0; t < nthreads; t++) free(tbuf[t]);
  free(tbuf);
  free(tcnt);
  free(tcap);
  free(cur);
  free(nxt);
  free(bm);
  free(bm2);
  return lv;
}

uint64_t orc_check_bfs_tree(uint64_t n, const uint64_t *off,
                            const uint32_t *ids, uint32_t src,
                            const uint32_t *parents, const uint32_t *level) {
  for (uint64_t v = 0; v < n; v++) {
    uint32_t p = parents[v];
    if (level[v] == ORC_NONE) {
      if (p != ORC_NONE) return v + 1;
      continue;
    }
    if (v == src) {
      if (p != src) return v + 1;
      continue;
    }
    if (p == ORC_NONE || p >= n) return v + 1;
    if (level[p] == ORC_NONE || level[p] + 1 != level[v]) return v + 1;
    /* p must be a neighbour of v (adjacency sorted: binary search) */
    uint64_t lo = off[v], hi = off[v + 1];
    while (lo < hi) {
      uint64_t mid = (lo + hi) / 2;
      if (ids[mid] < p) lo = mid + 1; else hi = mid;
    }
    if (lo == off[v + 1] || ids[lo] != p) return v + 1;
  }
  return 0;
}

/* ===================================================================== */
/* PageRank                                                              */
/* ===================================================================== */

/* PageRank.h:80-123 (PR_F / PR_Vertex_F, Jacobi), evaluated as a pull over
 * the symmetric adjacency: identical arithmetic result up to summation
 * order. */
int orc_pagerank_f64(uint64_t n, const uint64_t *off, const uint32_t *ids,
                     double damping, int max_iters, double eps, double *p,
                     double *l1_log) {
  if (n == 0) return 0;
  double *contrib = (double *)malloc(n * sizeof(double));
  double *pn = (double *)malloc(n * sizeof(double));
  const double added = (1.0 - damping) * (1.0 / (double)n);
  for (uint64_t i = 0; i < n; i++) p[i] = 1.0 / (double)n;
  int it = 0;
  while (it < max_iters) {
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < (int64_t)n; s++) {
      uint64_t d = off[s + 1] - off[s];
      contrib[s] = d ? p[s] / (double)d : 0.0;
    }
    double l1 = 0.0;
#pragma omp parallel for reduction(+ : l1) schedule(dynamic, 1024)
    for (int64_t i = 0; i < (int64_t)n; i++) {
      double acc = 0.0;
      for (uint64_t e = off[i]; e < off[i + 1]; e++) acc += contrib[ids[e]];
      double v = damping * acc + added;
      pn[i] = v;
      l1 += fabs(v - p[i]);
    }
    memcpy(p, pn, n * sizeof(double));
    if (l1_log) l1_log[it] = l1;
    it++;
    if (eps > 0 && l1 < eps) break;
  }
  free(contrib);
  free(pn);
  return it;
}

int orc_ppr_f64(uint64_t n, const uint64_t *off, const uint32_t *ids,
                uint32_t seed, double damping, int iters, double *p) {
  if (n == 0) return 0;
  double *contrib = (double *)malloc(n * sizeof(double));
  double *pn = (double *)malloc(n * sizeof(double));
  for (uint64_t i = 0; i < n; i++) p[i] = 0.0;
  p[seed] = 1.0;
  for (int it = 0; it < iters; it++) {
    for (uint64_t s = 0; s < n; s++) {
      uint64_t d = off[s + 1] - off[s];
      contrib[s] = d ? p[s] / (double)d : 0.0;
    }
    for (uint64_t i = 0; i < n; i++) {
      double acc = 0.0;
      for (uint64_t e = off[i]; e < off[i + 1]; e++) acc += contrib[ids[e]];
      pn[i] = damping * acc + (i == seed ? 1.0 - damping : 0.0);
    }
    memcpy(p, pn, n * sizeof(double));
  }
  free(contrib);
  free(pn);
  return iters;
}

/* PageRank.h:125-193 with parallel_for run in index order: the apply
 * function overwrites p_div[u] inside the sweep that reads p_div[s]. */
int orc_pagerank_ref_order(uint64_t n, const uint64_t *off,
                           const uint32_t *ids, double damping, int max_iters,
                           double eps, double *p) {
  if (n == 0) return 0;
  const double added = (1.0 - damping) * (1.0 / (double)n);
  double *pc = (double *)malloc(n * sizeof(double));
  double *pn = (double *)malloc(n * sizeof(double));
  double *pdiv = (double *)malloc(n * sizeof(double));
  for (uint64_t i = 0; i < n; i++) {
    pc[i] = 1.0 / (double)n;
    pdiv[i] = (1.0 / (double)n) / (double)(off[i + 1] - off[i]);
  }
  int it = 0;
  while (it < max_iters) {
    for (uint64_t u = 0; u < n; u++) {
      double acc = 0.0;
      for (uint64_t e = off[u]; e < off[u + 1]; e++) acc += pdiv[ids[e]];
      pn[u] = damping * acc + added;
      pdiv[u] = pn[u] / (double)(off[u + 1] - off[u]);
    }
    double l1 = 0.0;
    for (uint64_t i = 0; i < n; i++) l1 += fabs(pc[i] - pn[i]);
    memcpy(pc, pn, n * sizeof(double));
    it++;
    if (l1 < eps) break;
  }
  memcpy(p, pc, n * sizeof(double));
  free(pc);
  free(pn);
  free(pdiv);
  return it;
}

/* ===================================================================== */
/* Connectivity                                                          */
/* ===================================================================== */

/* find_compress (SimpleUnionAsync/Connectivity.h:32-47) */
static uint32_t uf_find(uint32_t *par, uint32_t i) {
  uint32_t j = i;
  while (par[j] != j) j = par[j];
  uint32_t tmp;
  while ((tmp = par[i]) > j) {
    par[i] = j;
    i = tmp;
  }
  return j;
}

/* unite_impl (:49-65) run for every CSR entry in vertex order, then
 * finish() (:113-116). */
void orc_cc_union_find(uint64_t n, const uint64_t *off, const uint32_t *ids,
                       uint32_t *labels) {
  for (uint64_t i = 0; i < n; i++) labels[i] = (uint32_t)i;
  for (uint64_t u = 0; u < n; u++) {
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t a = uf_find(labels, (uint32_t)u);
      uint32_t b = uf_find(labels, ids[e]);
      if (a == b) continue;
      if (a > b) labels[a] = b; else labels[b] = a;
    }
  }
  for (uint64_t i = 0; i < n; i++) uf_find(labels, (uint32_t)i);
}

/* labelprop_cc::CC (LabelPropagation/Connectivity.h:78-146): repeat
 * label[t] = min(label[t], label[s]) over edges until no change. */
int orc_cc_labelprop(uint64_t n, const uint64_t *off, const uint32_t *ids,
                     uint32_t *labels) {
  for (uint64_t i = 0; i < n; i++) labels[i] = (uint32_t)i;
  int rounds = 0;
  int changed = 1;
  while (changed) {
    changed = 0;
    rounds++;
    for (uint64_t u = 0; u < n; u++) {
      for (uint64_t e = off[u]; e < off[u + 1]; e++) {
        uint32_t t = ids[e];
        if (labels[u] < labels[t]) {
          labels[t] = labels[u];
          changed = 1;
        }
      }
    }
  }
  return rounds;
}

/* ===================================================================== */
/* Triangle counting                                                     */
/* ===================================================================== */

static inline int rank_less(const uint64_t *off, uint32_t u, uint32_t v) {
  uint64_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
  return du < dv || (du == dv && u < v);
}

/* filterGraph(G, rank[u] < rank[v]) (Triangle.h:185-190 ->
 * graph_mutation.h:15-81): count pass, scan, order-preserving fill. */
void orc_orient(uint64_t n, const uint64_t *off, const uint32_t *ids,
                uint64_t *dg_off, uint32_t *dg_ids) {
  dg_off[0] = 0;
  for (uint64_t u = 0; u < n; u++) {
    uint64_t c = 0;
    for (uint64_t e = off[u]; e < off[u + 1]; e++)
      c += rank_less(off, (uint32_t)u, ids[e]);
    dg_off[u + 1] = dg_off[u] + c;
  }
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    uint64_t w = dg_off[u];
    for (uint64_t e = off[u]; e < off[u + 1]; e++)
      if (rank_less(off, (uint32_t)u, ids[e])) dg_ids[w++] = ids[e];
  }
}

/* seq_merge_full (intersect.h:68-90) */
static uint64_t merge_count(const uint32_t *a, uint64_t na, const uint32_t *b,
                            uint64_t nb) {
  uint64_t i = 0, j = 0, c = 0;
  while (i < na && j < nb) {
    if (a[i] == b[j]) {
      c++;
      i++;
      j++;
    } else if (a[i] < b[j]) {
      i++;
    } else {
      j++;
    }
  }
  return c;
}

/* CountDirectedBalanced (Triangle.h:99-150) without the block split. */
uint64_t orc_tc(uint64_t n, const uint64_t *off, const uint32_t *ids) {
  uint64_t m = n ? off[n] : 0;
  uint64_t *dg_off = (uint64_t *)malloc((n + 1) * sizeof(uint64_t));
  uint32_t *dg_ids = (uint32_t *)malloc((m / 2 + 1) * sizeof(uint32_t));
  orc_orient(n, off, ids, dg_off, dg_ids);
  uint64_t total = 0;
#pragma omp parallel for reduction(+ : total) schedule(dynamic, 64)
  for (int64_t u = 0; u < (int64_t)n; u++) {
    const uint32_t *nu = dg_ids + dg_off[u];
    uint64_t du = dg_off[u + 1] - dg_off[u];
    for (uint64_t k = 0; k < du; k++) {
      uint32_t v = nu[k];
      total += merge_count(nu, du, dg_ids + dg_off[v], dg_off[v + 1] - dg_off[v]);
    }
  }
  free(dg_off);
  free(dg_ids);
  return total;
}

static int has_edge(const uint64_t *off, const uint32_t *ids, uint32_t u,
                    uint32_t w) {
  uint64_t lo = off[u], hi = off[u + 1];
  while (lo < hi) {
    uint64_t mid = (lo + hi) / 2;
    if (ids[mid] < w) lo = mid + 1; else hi = mid;
  }
  return lo < off[u + 1] && ids[lo] == w;
}

uint64_t orc_tc_brute(uint64_t n, const uint64_t *off, const uint32_t *ids) {
  uint64_t c = 0;
  for (uint64_t u = 0; u < n; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      uint32_t v = ids[e];
      if (v <= u) continue;
      for (uint64_t f = off[v]; f < off[v + 1]; f++) {
        uint32_t w = ids[f];
        if (w <= v) continue;
        c += has_edge(off, ids, (uint32_t)u, w);
      }
    }
  return c;
}

/* ===================================================================== */
/* Map                                                                   */
/* ===================================================================== */

/* LDD with permute = false (LowDiameterDecomposition.h:98-186 and
 * ldd_utils::generate_shifts :12-25; run_all's beta 0.2, ldd_permute
 * false): round r first makes every still unclaimed vertex in
 * [shifts[r], min(shifts[r+1], n)) a centre of its own cluster, then one
 * BFS step from the frontier (last round's claims and the new centres)
 * claims the unclaimed neighbours.  A vertex reached from several clusters
 * in one step joins the smallest cluster id -- one of the outcomes the
 * reference's CAS race allows (its pull would pick the first frontier
 * neighbour in adjacency order).  cluster[v] = the centre's id; round[v] =
 * the round whose frontier holds v. */
void orc_ldd(uint64_t n, const uint64_t *off, const uint32_t *ids, double beta, uint32_t *cluster, uint32_t *round) {
  if (n == 0) return;
  const uint64_t last = (uint64_t)fmin((double)(n + 1), 2.0 + ceil(log((double)n) / beta));
  uint64_t *shifts = (uint64_t *)calloc(last + 1, sizeof(uint64_t));
  for (uint64_t i = 0; i < last; i++) shifts[i] = (uint64_t)floor(exp((double)i * beta));
  uint64_t acc = 0;
  for (uint64_t i = 0; i <= last; i++) {
    const uint64_t x = shifts[i];
    shifts[i] = acc;
    acc += x;
  }
  uint32_t *fr = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *nx = (uint32_t *)malloc(n * sizeof(uint32_t));
  for (uint64_t v = 0; v < n; v++) {
    cluster[v] = ORC_NONE;
    round[v] = ORC_NONE;
  }
  uint64_t nf = 0, visited = 0;
  for (uint64_t r = 0; visited < n; r++) {
    const uint64_t start = shifts[r], end = shifts[r + 1] < n ? shifts[r + 1] : n;
    for (uint64_t v = start; v < end; v++)
      if (cluster[v] == ORC_NONE) {
        cluster[v] = (uint32_t)v;
        round[v] = (uint32_t)r;
        fr[nf++] = (uint32_t)v;
      }
    visited += nf;
    if (visited >= n) break;
    uint64_t nn = 0;
    for (uint64_t i = 0; i < nf; i++) {
      const uint32_t u = fr[i];
      for (uint64_t e = off[u]; e < off[u + 1]; e++) {
        const uint32_t v = ids[e];
        if (round[v] == ORC_NONE) {
          round[v] = (uint32_t)(r + 1);
          cluster[v] = cluster[u];
          nx[nn++] = v;
        } else if (round[v] == r + 1 && cluster[u] < cluster[v]) {
          cluster[v] = cluster[u];
        }
      }
    }
    uint32_t *t = fr;
    fr = nx;
    nx = t;
    nf = nn;
  }
  free(shifts);
  free(fr);
  free(nx);
}

/* delta::PageRankDelta (benchmarks/PageRank/PageRank.h:335-396), FP64.
 * Every round both branches of sparse_or_dense (:230-269) give
 * nghSum[v] = sum of Delta[s].delta_over_degree over v's neighbours s in
 * the frontier (sums here in adjacency order); the first round's filter
 * (:271-297) adds damping * nghSum + (1-d)/n to p and keeps
 * delta = that - 1/n, active when delta > local_eps * p; later rounds
 * (:305-327) take delta = damping * nghSum for every vertex and add it to p
 * only where |delta| > local_eps * p, which is also the next frontier.
 * Stops when sum |delta| < eps or after max_iters rounds; returns the
 * rounds run.  Isolated vertices divide by degree 0, as the reference does
 * (their quotient is never read). */
int orc_pagerank_delta(uint64_t n, const uint64_t *off, const uint32_t *ids, double eps, double local_eps,
                       int max_iters, double *p) {
  const double damping = 0.85, one_over_n = 1.0 / (double)n;
  double *dod = (double *)malloc((n ? n : 1) * sizeof(double));
  double *ngh = (double *)calloc(n ? n : 1, sizeof(double));
  uint8_t *front = (uint8_t *)malloc(n ? n : 1);
  uint8_t *act = (uint8_t *)malloc(n ? n : 1);
  for (uint64_t i = 0; i < n; i++) {
    p[i] = 0.0;
    dod[i] = one_over_n / (double)(off[i + 1] - off[i]);
    front[i] = 1;
  }
  int round = 0, ran = 0;
  while (round++ < max_iters) {
    ran = round;
    for (uint64_t v = 0; v < n; v++) {
      double sum = 0.0;
      for (uint64_t e = off[v]; e < off[v + 1]; e++)
        if (front[ids[e]]) sum += dod[ids[e]];
      ngh[v] = sum;
    }
    double l1 = 0.0;
    for (uint64_t i = 0; i < n; i++) {
      const double deg = (double)(off[i + 1] - off[i]);
      double nd;
      if (round == 1) {
        const double pre = damping * ngh[i] + (1.0 - damping) * one_over_n;
        p[i] += pre;
        nd = pre - one_over_n;
        act[i] = nd > local_eps * p[i];
      } else {
        nd = ngh[i] * damping;
        act[i] = fabs(nd) > local_eps * p[i];
        if (act[i]) p[i] += nd;
      }
      dod[i] = nd / deg;
      l1 += fabs(nd);
    }
    if (l1 < eps) break;
    memcpy(front, act, n);
  }
  free(dod);
  free(ngh);
  free(front);
  free(act);
  return ran;
}

/* bc_bfs::SSBetweennessCentrality_BFS (benchmarks/SSBetweenessCentrality/
 * Brandes/SSBetweennessCentrality.h:332-454), level-synchronous, FP64:
 * forward, NumPaths[u] = sum of NumPaths over neighbours in earlier levels
 * (Visited == 2); backward from the deepest level, Dep[u] = sum of Dep over
 * neighbours in deeper levels, then Dep[u] += 1/NumPaths[u]; result
 * (Dep - 1/NumPaths) / (1/NumPaths) -- NaN for unreached vertices, exactly
 * as the reference's arithmetic gives.  Sums run in adjacency order. */
void orc_bc(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t src, double *out) {
  uint32_t *lv = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  double *np = (double *)calloc(n ? n : 1, sizeof(double));
  double *dep = (double *)calloc(n ? n : 1, sizeof(double));
  uint32_t *order = (uint32_t *)malloc((n ? n : 1) * sizeof(uint32_t));
  orc_bfs_levels(n, off, ids, src, lv);
  /* vertices in BFS order (levels ascending) */
  uint64_t cnt = 0;
  if (n) {
    uint64_t head = 0;
    order[cnt++] = src;
    uint8_t *seen = (uint8_t *)calloc(n, 1);
    seen[src] = 1;
    while (head < cnt) {
      const uint32_t u = order[head++];
      for (uint64_t e = off[u]; e < off[u + 1]; e++)
        if (!seen[ids[e]]) {
          seen[ids[e]] = 1;
          order[cnt++] = ids[e];
        }
    }
    free(seen);
    np[src] = 1.0;
  }
  for (uint64_t i = 1; i < cnt; i++) {
    const uint32_t u = order[i];
    double s = 0.0;
    for (uint64_t e = off[u]; e < off[u + 1]; e++)
      if (lv[ids[e]] < lv[u]) s += np[ids[e]];
    np[u] = s;
  }
  for (uint64_t i = 0; i < n; i++) np[i] = 1.0 / np[i];
  /* deepest level first; within a level any order (sums read deeper only) */
  for (uint64_t i = cnt; i-- > 0;) {
    const uint32_t u = order[i];
    double s = 0.0;
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      const uint32_t d = lv[ids[e]];
      if (d != ORC_NONE && d > lv[u]) s += dep[ids[e]];
    }
    dep[u] = s + np[u];
  }
  for (uint64_t i = 0; i < n; i++) out[i] = (dep[i] - np[i]) / np[i];
  free(lv);
  free(np);
  free(dep);
  free(order);
}

/* KCore (benchmarks/KCore/JulienneDBS17/KCore.h:8-56) returns every
 * vertex's coreness: the bucket k at which peeling removes it, degrees
 * floored at the current k.  Restated as the sequential bucket peel of
 * Batagelj and Zaversnik (vertices in order of current degree, each
 * removal decrementing its higher-degree neighbours). */
void orc_kcore(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t *core) {
  if (n == 0) return;
  uint32_t md = 0;
  uint32_t *deg = (uint32_t *)malloc(n * sizeof(uint32_t));
  for (uint64_t v = 0; v < n; v++) {
    deg[v] = (uint32_t)(off[v + 1] - off[v]);
    if (deg[v] > md) md = deg[v];
  }
  uint64_t *bin = (uint64_t *)calloc((size_t)md + 2, sizeof(uint64_t));
  uint32_t *vert = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t *pos = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t v = 0; v < n; v++) bin[deg[v]]++;
  uint64_t start = 0;
  for (uint32_t d = 0; d <= md; d++) {
    const uint64_t c = bin[d];
    bin[d] = start;
    start += c;
  }
  for (uint64_t v = 0; v < n; v++) {
    pos[v] = bin[deg[v]];
    vert[pos[v]] = (uint32_t)v;
    bin[deg[v]]++;
  }
  for (uint32_t d = md; d >= 1; d--) bin[d] = bin[d - 1];
  bin[0] = 0;
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t v = vert[i];
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      const uint32_t u = ids[e];
      if (deg[u] > deg[v]) {
        const uint32_t du = deg[u];
        const uint64_t pu = pos[u], pw = bin[du];
        const uint32_t w = vert[pw];
        if (u != w) {
          pos[u] = pw;
          vert[pu] = w;
          pos[w] = pu;
          vert[pw] = u;
        }
        bin[du]++;
        deg[u]--;
      }
    }
  }
  for (uint64_t v = 0; v < n; v++) core[v] = deg[v];
  free(deg);
  free(bin);
  free(vert);
  free(pos);
}

uint64_t orc_map_sum(uint64_t m, const uint32_t *ids) {
  uint64_t s = 0;
#pragma omp parallel for reduction(+ : s) schedule(static)
  for (int64_t i = 0; i < (int64_t)m; i++) s += ids[i];
  return s;
}

/* ===================================================================== */
/* MLP 6->32->32->3 forward/backward (double accumulation)               */
/* ===================================================================== */

double orc_mlp_forward_backward(const float *P, const float *x,
                                const int32_t *labels, int batch, float *G,
                                float *logits) {
  const float *W1 = P, *b1 = W1 + 32 * 6, *W2 = b1 + 32, *b2 = W2 + 32 * 32,
              *W3 = b2 + 32, *b3 = W3 + 3 * 32;
  double g[ORC_MLP_NPARAM];
  memset(g, 0, sizeof(g));
  double *gW1 = g, *gb1 = gW1 + 32 * 6, *gW2 = gb1 + 32, *gb2 = gW2 + 32 * 32,
         *gW3 = gb2 + 32, *gb3 = gW3 + 3 * 32;
  double loss = 0.0;
  for (int s = 0; s < batch; s++) {
    const float *xs = x + 6 * s;
    double h1[32], h2[32], z[3];
    for (int o = 0; o < 32; o++) {
      double a = b1[o];
      for (int i = 0; i < 6; i++) a += (double)W1[o * 6 + i] * xs[i];
      h1[o] = a > 0 ? a : 0;
    }
    for (int o = 0; o < 32; o++) {
      double a = b2[o];
      for (int i = 0; i < 32; i++) a += (double)W2[o * 32 + i] * h1[i];
      h2[o] = a > 0 ? a : 0;
    }
    for (int o = 0; o < 3; o++) {
      double a = b3[o];
      for (int i = 0; i < 32; i++) a += (double)W3[o * 32 + i] * h2[i];
      z[o] = a;
      if (logits) logits[3 * s + o] = (float)a;
    }
    double mx = z[0] > z[1] ? z[0] : z[1];
    if (z[2] > mx) mx = z[2];
    double se = exp(z[0] - mx) + exp(z[1] - mx) + exp(z[2] - mx);
    int y = labels[s];
    loss += -(z[y] - mx - log(se));
    double dz[3];
    for (int o = 0; o < 3; o++)
      dz[o] = (exp(z[o] - mx) / se - (o == y ? 1.0 : 0.0)) / batch;
    double dh2[32], dh1[32];
    for (int i = 0; i < 32; i++) dh2[i] = 0;
    for (int o = 0; o < 3; o++) {
      gb3[o] += dz[o];
      for (int i = 0; i < 32; i++) {
        gW3[o * 32 + i] += dz[o] * h2[i];
        dh2[i] += W3[o * 32 + i] * dz[o];
      }
    }
    for (int i = 0; i < 32; i++) {
      dh2[i] = h2[i] > 0 ? dh2[i] : 0;
      dh1[i] = 0;
    }
    for (int o = 0; o < 32; o++) {
      gb2[o] += dh2[o];
      for (int i = 0; i < 32; i++) {
        gW2[o * 32 + i] += dh2[o] * h1[i];
        dh1[i] += W2[o * 32 + i] * dh2[o];
      }
    }
    for (int o = 0; o < 32; o++) {
      double d = h1[o] > 0 ? dh1[o] : 0;
      gb1[o] += d;
      for (int i = 0; i < 6; i++) gW1[o * 6 + i] += d * xs[i];
    }
  }
  if (G)
    for (int k = 0; k < ORC_MLP_NPARAM; k++) G[k] = (float)g[k];
  return loss / batch;
}

/* ===================================================================== */
/* MaximalIndependentSet (rootset)                                       */
/* ===================================================================== */

/* Priority key of vertex v: the reference draws a random permutation
 * (parlay::random_permutation, MaximalIndependentSet.h:127); here the order
 * is that of (high 32 bits of splitmix64(v), v), a fixed pseudo-random
 * permutation that the device uses too. */
uint64_t orc_mis_key(uint32_t v) {
  return ((orc_splitmix64(v) >> 32) << 32) | (uint64_t)v;
}

/* MaximalIndependentSet_rootset::MaximalIndependentSet
 * (MaximalIndependentSet.h:119-170) computes the lexicographically first
 * MIS of the priority order: the rootset rounds add exactly the vertices
 * that sequential greedy in that order adds.  Restated as the sequential
 * greedy: vertices in increasing key order join unless a neighbour already
 * has.  in_mis[n] = 0/1. */
void orc_mis(uint64_t n, const uint64_t *off, const uint32_t *ids, uint8_t *in_mis) {
  if (n == 0) return;
  uint64_t *key = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t v = 0; v < n; v++) key[v] = orc_mis_key((uint32_t)v);
  /* the low 32 bits of a key are the vertex id */
  radix_sort_u64(key, n);
  memset(in_mis, 0, n);
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t v = (uint32_t)(key[i] & 0xFFFFFFFFull);
    int blocked = 0;
    for (uint64_t e = off[v]; e < off[v + 1] && !blocked; e++) blocked = in_mis[ids[e]];
    in_mis[v] = (uint8_t)!blocked;
  }
  free(key);
}

/* ===================================================================== */
/* Coloring (Hasenplaugh14, LLF)                                         */
/* ===================================================================== */

static uint32_t log2_up_u64(uint64_t x) { /* parlay::log2_up: ceil(log2 x), 0 for x <= 1 */
  uint32_t r = 0;
  while (x > ((uint64_t)1 << r)) r++;
  return r;
}

/* u comes before v in the LLF order of Coloring (GraphColoring.h:97-112):
 * larger log2_up(degree) first, ties by the permutation, which here is the
 * MIS key order (orc_mis_key). */
static int llf_before(const uint64_t *off, uint32_t u, uint32_t v) {
  const uint64_t du = off[u + 1] - off[u], dv = off[v + 1] - off[v];
  const uint32_t lu = du ? log2_up_u64(du) : 0, lv = dv ? log2_up_u64(dv) : 0;
  return lu > lv || (lu == lv && orc_mis_key(u) < orc_mis_key(v));
}

/* Coloring(G, lf = false) (GraphColoring.h:71-146) colours the rootset of
 * the LLF order round by round; a root's colour is coloring::color
 * (:30-55): the smallest colour in [0, deg) no neighbour holds, or deg + 1
 * when all of [0, deg) are taken (the reference's value, kept), 0 for an
 * isolated vertex.  A root's coloured neighbours are exactly the ones
 * before it, so this is greedy colouring in LLF order, restated
 * sequentially. */
void orc_coloring(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t *color) {
  if (n == 0) return;
  /* LLF order: sort by permutation key, then a stable counting sort by
   * 63 - log2_up(degree) */
  uint64_t *key = (uint64_t *)malloc(n * sizeof(uint64_t));
  for (uint64_t v = 0; v < n; v++) key[v] = orc_mis_key((uint32_t)v);
  radix_sort_u64(key, n); /* by permutation key */
  uint32_t *order = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint64_t cntl[66] = {0};
  for (uint64_t v = 0; v < n; v++) {
    const uint64_t d = off[v + 1] - off[v];
    cntl[63 - (d ? log2_up_u64(d) : 0)]++;
  }
  uint64_t start[66], acc = 0;
  for (int l = 0; l < 66; l++) {
    start[l] = acc;
    acc += cntl[l];
  }
  for (uint64_t i = 0; i < n; i++) { /* stable counting sort by level */
    const uint32_t v = (uint32_t)(key[i] & 0xFFFFFFFFull);
    const uint64_t d = off[v + 1] - off[v];
    order[start[63 - (d ? log2_up_u64(d) : 0)]++] = v;
  }
  free(key);
  for (uint64_t v = 0; v < n; v++) color[v] = 0xFFFFFFFFu;
  uint8_t *used = (uint8_t *)calloc(n + 2, 1);
  for (uint64_t i = 0; i < n; i++) {
    const uint32_t v = order[i];
    const uint64_t deg = off[v + 1] - off[v];
    if (deg == 0) {
      color[v] = 0;
      continue;
    }
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      const uint32_t c = color[ids[e]];
      if (c < deg) used[c] = 1;
    }
    uint64_t c = 0;
    while (c < deg && used[c]) c++;
    color[v] = (uint32_t)(c < deg ? c : deg + 1);
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      const uint32_t cc = color[ids[e]];
      if (cc < deg) used[cc] = 0;
    }
  }
  free(used);
  free(order);
}

/* 1 where u comes before v in the LLF order (for the tests). */
int orc_llf_before(const uint64_t *off, uint32_t u, uint32_t v) { return llf_before(off, u, v); }

/* ===================================================================== */
/* WorkEfficientDensestSubgraph (ApproxPeelingBKV12)                     */
/* ===================================================================== */

/* WorkEfficientDensestSubgraph (ApproximateDensestSubgraph/ApproxPeelingBKV12/
 * DensestSubgraph.h:8-145): rounds over the alive set R with D[v] = alive
 * neighbours of v; density = sum_R D / |R| (both directions of an edge
 * counted), the maximum over rounds returned; every alive vertex with
 * D[v] <= (1 + eps) * density is removed and its alive neighbours lose one
 * per removed neighbour.  Returns the maximum density; *rounds the rounds
 * run. */
double orc_densest(uint64_t n, const uint64_t *off, const uint32_t *ids, double eps, uint32_t *rounds) {
  uint32_t r = 0;
  double best = 0.0;
  if (n == 0) {
    if (rounds) *rounds = 0;
    return 0.0;
  }
  uint64_t *D = (uint64_t *)malloc(n * sizeof(uint64_t));
  uint8_t *alive = (uint8_t *)malloc(n);
  uint32_t *rem = (uint32_t *)malloc(n * sizeof(uint32_t));
  for (uint64_t v = 0; v < n; v++) {
    D[v] = off[v + 1] - off[v];
    alive[v] = 1;
  }
  uint64_t left = n;
  while (left > 0) {
    uint64_t edges = 0;
    for (uint64_t v = 0; v < n; v++)
      if (alive[v]) edges += D[v];
    const double density = (double)edges / (double)left;
    const double target = ((1.0 + eps) * (double)edges) / (double)left;
    if (density > best) best = density;
    uint64_t k = 0;
    for (uint64_t v = 0; v < n; v++)
      if (alive[v] && (double)D[v] <= target) rem[k++] = (uint32_t)v;
    for (uint64_t i = 0; i < k; i++) alive[rem[i]] = 0;
    left -= k;
    for (uint64_t i = 0; i < k; i++) {
      const uint32_t u = rem[i];
      for (uint64_t e = off[u]; e < off[u + 1]; e++)
        if (alive[ids[e]]) D[ids[e]]--;
    }
    r++;
  }
  free(D);
  free(alive);
  free(rem);
  if (rounds) *rounds = r;
  return best;
}

/* ===================================================================== */
/* Spanner (MPXV15)                                                      */
/* ===================================================================== */

typedef struct {
  uint64_t pair, edge;
} sp_cand;

static int sp_cand_cmp(const void *a, const void *b) {
  const sp_cand *x = (const sp_cand *)a, *y = (const sp_cand *)b;
  if (x->pair != y->pair) return x->pair < y->pair ? -1 : 1;
  if (x->edge != y->edge) return x->edge < y->edge ? -1 : 1;
  return 0;
}

static int u64_cmp(const void *a, const void *b) {
  const uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
  return x < y ? -1 : x > y;
}

/* Spanner (Spanner/MPXV15/Spanner.h:242-261) with beta = ln(n) / (2k) as
 * Spanner_runner sets it: LDD_parents, then tree_and_intercluster_edges.
 * The reference's LDD permutes the centres and lets CAS races pick parents
 * and one edge per cluster pair; restated deterministically: the LDD of
 * orc_ldd (identity order, smallest cluster), v's parent the smallest-id
 * neighbour in v's cluster claimed one round before v, and for every pair
 * of adjacent clusters c1 < c2 the edge (u, v) with cl[u] = c1, cl[v] = c2
 * smallest as (u << 32 | v).  out (capacity n + m) receives the edges as
 * sorted keys (u << 32 | v); returns their number. */
uint64_t orc_spanner(uint64_t n, const uint64_t *off, const uint32_t *ids, uint32_t k, uint64_t *out) {
  if (n == 0) return 0;
  const double beta = log((double)n) / (2.0 * (double)k);
  uint32_t *cl = (uint32_t *)malloc(n * sizeof(uint32_t));
  uint32_t *rd = (uint32_t *)malloc(n * sizeof(uint32_t));
  orc_ldd(n, off, ids, beta, cl, rd);
  uint64_t cnt = 0;
  for (uint64_t v = 0; v < n; v++) {
    if (cl[v] == v) continue;
    uint32_t best = 0xFFFFFFFFu;
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      const uint32_t u = ids[e];
      if (cl[u] == cl[v] && rd[u] + 1 == rd[v] && u < best) best = u;
    }
    out[cnt++] = (v << 32) | best;
  }
  const uint64_t m = off[n];
  sp_cand *c = (sp_cand *)malloc((m + 1) * sizeof(sp_cand));
  uint64_t nc = 0;
  for (uint64_t u = 0; u < n; u++)
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      const uint32_t v = ids[e];
      if (cl[u] < cl[v]) {
        c[nc].pair = ((uint64_t)cl[u] << 32) | cl[v];
        c[nc].edge = (u << 32) | v;
        nc++;
      }
    }
  qsort(c, nc, sizeof(sp_cand), sp_cand_cmp);
  for (uint64_t i = 0; i < nc; i++)
    if (i == 0 || c[i].pair != c[i - 1].pair) out[cnt++] = c[i].edge;
  free(c);
  free(cl);
  free(rd);
  qsort(out, cnt, sizeof(uint64_t), u64_cmp);
  return cnt;
}
