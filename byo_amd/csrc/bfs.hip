// bfs.hip -- GBBS edgeMap traversal on gfx950 and the BFS it drives.
//
// Frontier (vertexSubsetData<empty>, gbbs/vertex_subset.h:203-371) is
// either a sparse queue of uint32 ids or a dense bitmap of uint32 words
// (BitArray, gbbs/bit_array.h:9-122, bit i = word[i/32] >> (i%32)).
// Each edgeMap call (gbbs/edge_map_data.h:254-307) is one of
//   sparse push  (edgeMapSparse, gbbs/edge_map_blocked.h:18-92): the edges
//                of the queue are split evenly over waves by a prefix sum of
//                frontier degrees; CAS on parents claims a vertex and the
//                winners are appended with one atomic per wave (ballot
//                compaction replaces Reducer_Vector, reducer.hpp:163-392);
//   dense pull   (edgeMapDense, edge_map_data.h:47-175): one lane per
//                unvisited vertex scans its neighbours until the first
//                frontier member (early exit); a wave owns 64 consecutive
//                vertices, so the next bitmap is written as whole 64-bit
//                words from a ballot, without atomics;
//   dense forward (edgeMapDenseForward, :180-247): push over the bitmap
//                frontier's vertices, one wave per frontier vertex.
// A visited bitmap mirrors "parents[v] != NONE" so the push's cond() test
// hits a 0.5 MiB (scale 22) array that stays in L2.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "primitives.hpp"

namespace byo {

// logits of the selector for features x; true when every normalised
// feature lies within range deviations of the training set (selector.hip)
bool mlp_eval_host(const std::vector<float> &p, const float *x, float *logits, float range);
void bfs_batch_run(byo_graph *g, const uint32_t *srcs, uint32_t k, byo_bfs_stats *stats, uint64_t maxdeg,
                   double *ms_out);  // bfs_batch.hip
void bfs_coop_run(byo_graph *g, const uint32_t *srcs, uint32_t k, bool keep_all, byo_bfs_stats *stats,
                  uint64_t maxdeg, double *ms_out);  // bfs_coop.hip

namespace {

struct BfsCtl {
  unsigned long long nsize;  // next frontier size
  unsigned long long ndeg;   // sum of out-degrees of the next frontier
  unsigned long long pad[6];
};

__device__ __forceinline__ bool bit_get(const uint32_t *bm, uint32_t i) {
  return (bm[i >> 5] >> (i & 31)) & 1u;
}

__global__ void k_bfs_init(uint32_t *parents, uint64_t n, uint64_t *visited64, uint64_t nw64) {
  for (uint64_t i = global_tid(); i < n; i += grid_threads()) parents[i] = kNone;
  for (uint64_t i = global_tid(); i < nw64; i += grid_threads()) visited64[i] = 0;
}

__global__ void k_bfs_seed(uint32_t *parents, uint32_t *visited, uint32_t *q, uint32_t src,
                           const uint64_t *off, BfsCtl *ctl) {
  parents[src] = src;
  visited[src >> 5] |= 1u << (src & 31);
  q[0] = src;
  ctl->nsize = 1;
  ctl->ndeg = off[src + 1] - off[src];
}

// deg[i] = out-degree of q[i]
__global__ void k_frontier_degrees(const uint32_t *q, uint64_t fs, const uint64_t *off, uint64_t *deg) {
  for (uint64_t i = global_tid(); i < fs; i += grid_threads()) {
    uint32_t v = q[i];
    deg[i] = off[v + 1] - off[v];
  }
}

constexpr int kPushPerLane = 4;
constexpr uint64_t kPushChunk = 64 * kPushPerLane;

// Sparse push over the edges of queue q (scan = exclusive degree prefix,
// fs+1 entries, scan[fs] = total edges E).  Each wave takes chunks of 256
// consecutive edges.
__global__ void __launch_bounds__(256) k_push(const uint64_t *off, const uint32_t *ids, const uint32_t *q,
                                              const uint64_t *scan, uint64_t fs, uint64_t E,
                                              uint32_t *parents, uint32_t *visited, uint32_t *qn,
                                              BfsCtl *ctl) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const uint64_t nchunks = (E + kPushChunk - 1) / kPushChunk;
  for (uint64_t c = wave; c < nchunks; c += nwaves) {
    const uint64_t e0 = c * kPushChunk;
    // first frontier index of the chunk: last i with scan[i] <= e0
    uint64_t lo = 0, hi = fs;  // invariant scan[lo] <= e0 < scan[hi]
    while (hi - lo > 1) {
      uint64_t mid = (lo + hi) >> 1;
      if (scan[mid] <= e0) lo = mid; else hi = mid;
    }
    // frontier vertices of out-degree 0 (possible on a directed CPMA) repeat
    // scan values, so the search range of the lanes stays [ibase, fs)
    const uint64_t ibase = lo;
    const uint64_t ilim = fs;
#pragma unroll
    for (int t = 0; t < kPushPerLane; t++) {
      const uint64_t e = e0 + (uint64_t)t * 64 + lane;
      bool win = false;
      uint32_t tgt = 0;
      if (e < E) {
        uint64_t a = ibase, b = ilim;  // scan[a] <= e < scan[b]
        while (b - a > 1) {
          uint64_t mid = (a + b) >> 1;
          if (scan[mid] <= e) a = mid; else b = mid;
        }
        const uint32_t u = q[a];
        tgt = ids[off[u] + (e - scan[a])];
        if (!bit_get(visited, tgt)) {
          if (atomicCAS(&parents[tgt], kNone, u) == kNone) {
            win = true;
            atomicOr(&visited[tgt >> 5], 1u << (tgt & 31));
          }
        }
      }
      const uint64_t mask = __ballot(win);
      if (mask) {
        uint64_t d = win ? off[tgt + 1] - off[tgt] : 0;
        d = wave_sum(d);
        unsigned long long base = 0;
        if (lane == 0) {
          base = atomicAdd(&ctl->nsize, (unsigned long long)__popcll(mask));
          atomicAdd(&ctl->ndeg, (unsigned long long)d);
        }
        base = __shfl(base, 0, 64);
        if (win) qn[base + mbcnt(mask)] = tgt;
      }
    }
  }
}

// Dense pull: wave w owns vertices [64w, 64w+64) == visited64[w].
__global__ void __launch_bounds__(256) k_pull(const uint64_t *off, const uint32_t *ids, uint64_t n,
                                              const uint32_t *fbm, uint64_t *nbm64, uint64_t *visited64,
                                              uint32_t *parents, BfsCtl *ctl) {
  const int lane = lane_id();
  const uint64_t nwords = (n + 63) >> 6;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t w = wave; w < nwords; w += nwaves) {
    const uint64_t vis = visited64[w];
    const uint64_t v = w * 64 + lane;
    bool found = false;
    uint64_t dv = 0;
    if (v < n && !((vis >> lane) & 1)) {
      const uint64_t b = off[v], e = off[v + 1];
      dv = e - b;
      for (uint64_t k = b; k < e; k++) {
        const uint32_t u = ids[k];
        if (bit_get(fbm, u)) {
          parents[v] = u;
          found = true;
          break;
        }
      }
    }
    const uint64_t mask = __ballot(found);
    uint64_t d = wave_sum(found ? dv : 0);
    if (lane == 0) {
      nbm64[w] = mask;
      if (mask) {
        visited64[w] = vis | mask;
        atomicAdd(&ctl->nsize, (unsigned long long)__popcll(mask));
        atomicAdd(&ctl->ndeg, (unsigned long long)d);
      }
    }
  }
}

// Dense forward: one wave per frontier vertex (bitmap), push to neighbours.
__global__ void __launch_bounds__(256) k_forward(const uint64_t *off, const uint32_t *ids, uint64_t n,
                                                 const uint64_t *fbm64, uint64_t *nbm64_atomic,
                                                 uint32_t *visited, uint32_t *parents, BfsCtl *ctl) {
  const int lane = lane_id();
  const uint64_t nwords = (n + 63) >> 6;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t w = wave; w < nwords; w += nwaves) {
    uint64_t word = fbm64[w];
    while (word) {
      const int bitpos = __builtin_ctzll(word);
      word &= word - 1;
      const uint32_t u = (uint32_t)(w * 64 + bitpos);
      const uint64_t b = off[u], e = off[u + 1];
      for (uint64_t k0 = b; k0 < e; k0 += 64) {
        const uint64_t k = k0 + lane;
        bool win = false;
        uint32_t t = 0;
        if (k < e) {
          t = ids[k];
          if (!bit_get(visited, t) && atomicCAS(&parents[t], kNone, u) == kNone) {
            win = true;
            atomicOr(&visited[t >> 5], 1u << (t & 31));
            atomicOr(reinterpret_cast<unsigned long long *>(&nbm64_atomic[t >> 6]), 1ull << (t & 63));
          }
        }
        const uint64_t mask = __ballot(win);
        if (mask) {
          uint64_t d = wave_sum(win ? off[t + 1] - off[t] : 0);
          if (lane == 0) {
            atomicAdd(&ctl->nsize, (unsigned long long)__popcll(mask));
            atomicAdd(&ctl->ndeg, (unsigned long long)d);
          }
        }
      }
    }
  }
}

__global__ void k_clear64(uint64_t *p, uint64_t n) {
  for (uint64_t i = global_tid(); i < n; i += grid_threads()) p[i] = 0;
}

// toDense (vertex_subset.h:357-364): set_atomic per queue element.
__global__ void k_q2bm(const uint32_t *q, uint64_t fs, uint32_t *bm) {
  for (uint64_t i = global_tid(); i < fs; i += grid_threads()) {
    uint32_t v = q[i];
    atomicOr(&bm[v >> 5], 1u << (v & 31));
  }
}

// toSparse (vertex_subset.h:339-354): pack bitmap to a queue (any order).
__global__ void k_bm2q(const uint64_t *bm64, uint64_t nwords, uint32_t *q, unsigned long long *cursor) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t w0 = wave * 64; w0 < nwords; w0 += nwaves * 64) {
    const uint64_t w = w0 + lane;
    uint64_t word = w < nwords ? bm64[w] : 0;
    uint32_t c = __popcll(word);
    uint32_t incl = wave_incl_scan(c);
    uint32_t tot = __shfl(incl, 63, 64);
    unsigned long long base = 0;
    if (lane == 0 && tot) base = atomicAdd(cursor, (unsigned long long)tot);
    base = __shfl(base, 0, 64);
    uint64_t pos = base + incl - c;
    while (word) {
      int b = __builtin_ctzll(word);
      word &= word - 1;
      q[pos++] = (uint32_t)(w * 64 + b);
    }
  }
}

// Levels from a parent tree by pointer jumping:
// depth[v] = hops to src; anc[v] = ancestor 2^k hops up (or src).
__global__ void k_lv_init(const uint32_t *parents, uint64_t n, uint32_t src, uint32_t *anc, uint32_t *dep) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    uint32_t p = parents[v];
    anc[v] = p;
    dep[v] = (p == kNone) ? kNone : (v == src ? 0u : 1u);
  }
}
__global__ void k_lv_jump(const uint32_t *anc, const uint32_t *dep, uint32_t *anc2, uint32_t *dep2,
                          uint64_t n, uint32_t src) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    uint32_t a = anc[v];
    if (a == kNone || a == src) {
      anc2[v] = a;
      dep2[v] = dep[v];
    } else {
      anc2[v] = anc[a];
      dep2[v] = dep[v] + dep[a];
    }
  }
}

}  // namespace

// Per-vertex levels of the tree of the last BFS (g->last_tree, rooted at
// g->last_src) by pointer jumping: every vertex doubles the distance to its
// known ancestor each pass, so ceil(log2 n) + 1 passes reach the root.
const uint32_t *bfs_levels_dev(byo_graph *g) {
  const uint32_t src = (uint32_t)g->last_src;
  const uint64_t n = g->n;
  const uint32_t *parents = g->last_tree;
  BYO_REQUIRE(parents, BYO_E_STATE, "no BFS tree on this handle");
  uint32_t *anc = g->buf("lv_anc", n * 4 + 64).as<uint32_t>();
  uint32_t *dep = g->buf("lv_dep", n * 4 + 64).as<uint32_t>();
  uint32_t *anc2 = g->buf("lv_anc2", n * 4 + 64).as<uint32_t>();
  uint32_t *dep2 = g->buf("lv_dep2", n * 4 + 64).as<uint32_t>();
  hipStream_t s = g->stream;
  unsigned grid = byo_grid(n, 256, 16384);
  k_lv_init<<<grid, 256, 0, s>>>(parents, n, src, anc, dep);
  int jumps = 1;
  while ((1ull << jumps) < n + 1) jumps++;
  for (int j = 0; j < jumps + 1; j++) {
    k_lv_jump<<<grid, 256, 0, s>>>(anc, dep, anc2, dep2, n, src);
    std::swap(anc, anc2);
    std::swap(dep, dep2);
  }
  BYO_LAUNCH_CHECK();
  return dep;
}

// Host-side state of one BFS on a handle.
struct BfsRun {
  byo_graph *g;
  uint64_t n, m, nw64;
  uint32_t *parents, *visited, *q, *qn;
  uint64_t *fbm64, *nbm64, *deg, *scan;
  BfsCtl *ctl;
  void *scan_tmp;
  BfsCtl *hctl;
};

static BfsRun bfs_setup(byo_graph *g) {
  BfsRun r{};
  r.g = g;
  r.n = g->n;
  r.m = g->m;
  r.nw64 = (r.n + 63) / 64 + 1;
  r.parents = g->buf("bfs_parents", r.n * 4 + 64).as<uint32_t>();
  r.visited = g->buf("bfs_visited", r.nw64 * 8).as<uint32_t>();
  r.q = g->buf("bfs_q", r.n * 4 + 64).as<uint32_t>();
  r.qn = g->buf("bfs_qn", r.n * 4 + 64).as<uint32_t>();
  r.fbm64 = g->buf("bfs_fbm", r.nw64 * 8).as<uint64_t>();
  r.nbm64 = g->buf("bfs_nbm", r.nw64 * 8).as<uint64_t>();
  r.deg = g->buf("bfs_deg", r.n * 8 + 64).as<uint64_t>();
  r.scan = g->buf("bfs_scan", r.n * 8 + 64).as<uint64_t>();
  r.ctl = g->buf("bfs_ctl", sizeof(BfsCtl)).as<BfsCtl>();
  r.scan_tmp = g->buf("bfs_scan_tmp", prim::scan_tmp_bytes(r.n + 1)).p;
  r.hctl = static_cast<BfsCtl *>(pinned(g, sizeof(BfsCtl)));
  return r;
}

// Direction choice for one edgeMap call.  Returns 0 sparse, 1 dense pull,
// 2 empty (no out-edges), 3 dense forward.
static int choose_direction(byo_graph *g, bool is_dense, uint64_t fs, uint64_t fdeg, uint64_t visited_deg,
                            int level, uint64_t max_deg) {
  const uint64_t n = g->n, m = g->m;
  const uint64_t thr = g->dir_threshold < 0 ? m / 20 : (uint64_t)g->dir_threshold;
  // edge_map_data.h:262 -- an already dense, large frontier stays dense
  int rule;
  if (is_dense && fs > n / 10) {
    rule = 1;
  } else if (fdeg == 0) {
    return 2;  // :289
  } else {
    rule = (fs + fdeg > thr) ? 1 : 0;  // :291
  }
  switch (g->dir_policy) {
    case BYO_DIR_SPARSE: return 0;
    case BYO_DIR_DENSE: return 1;
    case BYO_DIR_SCHEDULE:
      // a scheduled bitmap pull (4) is this engine's pull
      if (level < 64 && g->dir_sched[level] != 255) return g->dir_sched[level] == 4 ? 1 : g->dir_sched[level];
      return rule;
    case BYO_DIR_MLP: {
      if (!g->sel_loaded) return rule;
      if (rule == 0 && !g->tuning("sel_densify")) return rule;  // sparse levels stay sparse
      float x[6];
      const double avg = n ? (double)m / (double)n : 0.0;
      x[0] = (float)std::log2(1.0 + (double)fs);
      x[1] = (float)std::log2(1.0 + (double)fdeg);
      x[2] = (float)std::log2(1.0 + (double)(m > visited_deg ? m - visited_deg : 0));
      x[3] = (float)level;
      x[4] = (float)std::log2(1.0 + avg);
      x[5] = (float)std::log2(1.0 + (double)max_deg);
      float logits[3];
      const bool in_range =
          mlp_eval_host(g->sel_host, x, logits, std::min(kSelClamp, (float)g->tuning("sel_range_x100") / 100.0f));
      int best = 0;
      for (int k = 1; k < 3; k++)
        if (logits[k] > logits[best]) best = k;
      if (!std::isfinite(logits[0] + logits[1] + logits[2])) return rule;  // fallback
      const float margin = (float)g->tuning("sel_margin_x100") / 100.0f;
      if (!in_range || logits[best] - logits[dir_class(rule)] < margin) return rule;
      return best == 0 ? 0 : (best == 1 ? 1 : 3);
    }
    default: return rule;
  }
}

// One full BFS from src.  Result stays in r.parents.
static void bfs_run(BfsRun &r, uint32_t src, byo_bfs_stats *st, uint64_t max_deg) {
  byo_graph *g = r.g;
  hipStream_t s = g->stream;
  const unsigned G = 2048;
  k_bfs_init<<<byo_grid(r.n, 256, 16384), 256, 0, s>>>(r.parents, r.n, reinterpret_cast<uint64_t *>(r.visited),
                                                       r.nw64);
  k_bfs_seed<<<1, 1, 0, s>>>(r.parents, r.visited, r.q, src, g->off.as<uint64_t>(), r.ctl);
  BYO_LAUNCH_CHECK();
  BYO_HIP(hipMemcpyAsync(r.hctl, r.ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, s));
  BYO_HIP(hipStreamSynchronize(s));
  uint64_t fs = r.hctl->nsize, fdeg = r.hctl->ndeg;
  bool is_dense = false;      // representation of the current frontier
  int rep = 0;                // 0 queue valid, 1 bitmap valid
  uint64_t visited_deg = 0;
  uint32_t level = 0;
  uint64_t reached = 0, edges_reached = 0;
  const uint64_t *off = g->off.as<uint64_t>();
  const uint32_t *ids = g->ids.as<uint32_t>();
  while (fs > 0) {
    reached += fs;
    edges_reached += fdeg;
    int dir = choose_direction(g, is_dense, fs, fdeg, visited_deg, (int)level, max_deg);
    visited_deg += fdeg;
    PhaseTimer lt(s, st && g->prof.on);
    if (st && level < 64) {
      st->fsize[level] = fs;
      st->fdeg[level] = fdeg;
      st->dir[level] = (uint8_t)dir;
    }
    level++;
    if (dir == 2) break;
    BYO_HIP(hipMemsetAsync(r.ctl, 0, 16, s));
    if (dir == 1 || dir == 3) {
      if (rep == 0) {
        k_clear64<<<byo_grid(r.nw64, 256, 4096), 256, 0, s>>>(r.fbm64, r.nw64);
        k_q2bm<<<byo_grid(fs, 256, 4096), 256, 0, s>>>(r.q, fs, reinterpret_cast<uint32_t *>(r.fbm64));
      }
      if (dir == 1) {
        k_pull<<<G, 256, 0, s>>>(off, ids, r.n, reinterpret_cast<const uint32_t *>(r.fbm64), r.nbm64,
                                 reinterpret_cast<uint64_t *>(r.visited), r.parents, r.ctl);
      } else {
        k_clear64<<<byo_grid(r.nw64, 256, 4096), 256, 0, s>>>(r.nbm64, r.nw64);
        k_forward<<<G, 256, 0, s>>>(off, ids, r.n, r.fbm64, r.nbm64, r.visited, r.parents, r.ctl);
      }
      BYO_LAUNCH_CHECK();
      std::swap(r.fbm64, r.nbm64);
      rep = 1;
      is_dense = true;
    } else {
      if (rep == 1) {
        BYO_HIP(hipMemsetAsync(&r.ctl->pad[0], 0, 8, s));
        k_bm2q<<<byo_grid((r.nw64 + 63) / 64 * 64, 256, 2048), 256, 0, s>>>(
            r.fbm64, (r.n + 63) / 64, r.q, &r.ctl->pad[0]);
      }
      k_frontier_degrees<<<byo_grid(fs, 256, 8192), 256, 0, s>>>(r.q, fs, off, r.deg);
      prim::exclusive_scan_u64(r.deg, r.scan, fs, true, r.scan_tmp, s);
      const uint64_t E = fdeg;
      const uint64_t chunks = (E + kPushChunk - 1) / kPushChunk;
      const unsigned grid = (unsigned)std::min<uint64_t>(8192, std::max<uint64_t>(1, (chunks + 3) / 4));
      k_push<<<grid, 256, 0, s>>>(off, ids, r.q, r.scan, fs, E, r.parents, r.visited, r.qn, r.ctl);
      BYO_LAUNCH_CHECK();
      std::swap(r.q, r.qn);
      rep = 0;
      is_dense = false;
    }
    BYO_HIP(hipMemcpyAsync(r.hctl, r.ctl, sizeof(BfsCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    if (st && level - 1 < 64) st->level_ms[level - 1] = (float)lt.stop();
    fs = r.hctl->nsize;
    fdeg = r.hctl->ndeg;
  }
  if (st) {
    st->levels = level;
    st->reached = (uint32_t)reached;
    st->edges_reached = edges_reached;
  }
}

namespace {
__global__ void k_max_deg(const uint64_t *off, uint64_t n, unsigned long long *out) {
  uint64_t mx = 0;
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    uint64_t d = off[v + 1] - off[v];
    mx = d > mx ? d : mx;
  }
  mx = wave_max(mx);
  if (lane_id() == 0 && mx) atomicMax(out, (unsigned long long)mx);
}
__global__ void k_max_deg_u32(const uint32_t *deg, uint64_t n, unsigned long long *out) {
  uint64_t mx = 0;
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) mx = deg[v] > mx ? deg[v] : mx;
  mx = wave_max(mx);
  if (lane_id() == 0 && mx) atomicMax(out, (unsigned long long)mx);
}
}  // namespace
// Maximum degree of the graph (from the container's own degrees).
uint64_t graph_max_degree(byo_graph *g) {
  if (g->maxdeg_valid) return g->maxdeg;
  auto &b = g->buf("maxdeg", 64);
  BYO_HIP(hipMemsetAsync(b.p, 0, 8, g->stream));
  const AdjView v = adj_view(g);
  if (g->n && v.deg)
    k_max_deg_u32<<<byo_grid(g->n, 256, 4096), 256, 0, g->stream>>>(v.deg, g->n, b.as<unsigned long long>());
  else if (g->n)
    k_max_deg<<<byo_grid(g->n, 256, 4096), 256, 0, g->stream>>>(v.off, g->n, b.as<unsigned long long>());
  BYO_LAUNCH_CHECK();
  uint64_t *h = static_cast<uint64_t *>(pinned(g, 64));
  BYO_HIP(hipMemcpyAsync(h, b.p, 8, hipMemcpyDeviceToHost, g->stream));
  BYO_HIP(hipStreamSynchronize(g->stream));
  g->maxdeg = h[0];
  g->maxdeg_valid = true;
  return g->maxdeg;
}
}  // namespace byo

using namespace byo;

extern "C" {

int byo_set_direction_policy(byo_graph_t g, int policy, int64_t threshold) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    BYO_REQUIRE(policy >= 0 && policy <= 4, BYO_E_INVAL, "unknown direction policy");
    g->dir_policy = policy;
    g->dir_threshold = threshold;
  });
}

int byo_set_direction_schedule(byo_graph_t g, const uint8_t *dirs, uint32_t len) {
  return guard([&] {
    BYO_REQUIRE(g && (dirs || len == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(len <= 64, BYO_E_INVAL, "schedule longer than 64 levels");
    for (uint32_t i = 0; i < len; i++)
      BYO_REQUIRE(dirs[i] == 0 || dirs[i] == 1 || dirs[i] == 3 || dirs[i] == 4 || dirs[i] == 255, BYO_E_INVAL,
                  "schedule entries must be 0, 1, 3, 4 or 255");
    std::memset(g->dir_sched, 255, sizeof(g->dir_sched));
    if (len) std::memcpy(g->dir_sched, dirs, len);
    g->dir_policy = BYO_DIR_SCHEDULE;
  });
}

int byo_bfs_multi(byo_graph_t g, const uint32_t *srcs, uint32_t k, uint32_t *parents, byo_bfs_stats *stats) {
  return guard([&] {
    BYO_REQUIRE(g && (srcs || k == 0), BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    for (uint32_t i = 0; i < k; i++) BYO_REQUIRE(srcs[i] < g->n, BYO_E_INVAL, "source out of range");
    if (k == 0) return;
    const uint64_t mx = (g->dir_policy == BYO_DIR_MLP) ? graph_max_degree(g) : 0;
    const int64_t engine = g->tuning("bfs_engine");
    const bool levels_mode = engine == 2;
    // the persistent engines read the container itself (adj_view); the
    // host-driven level loop reads a CSR (exported from a CPMA)
    if (levels_mode) ensure_csr(g);
    const bool seq_mode = engine == 1;
    if (k > 1 && !levels_mode && !seq_mode) {
      // default for several sources: up to 64 at a time share one traversal
      const uint64_t n = g->n;
      for (uint32_t c0 = 0; c0 < k; c0 += 64) {
        const uint32_t kc = std::min<uint32_t>(64, k - c0);
        double ms = 0;
        bfs_batch_run(g, srcs + c0, kc, stats ? stats + c0 : nullptr, mx, &ms);
        if (g->prof.on) {
          auto &acc = g->prof.acc["bfs"];
          acc.launches += kc;
          acc.ms += ms;
          g->account("bfs_launch", ms);
        }
        uint32_t *all = g->scratch["bfs_parents_all"].as<uint32_t>();
        if (parents)
          BYO_HIP(hipMemcpyAsync(parents + (uint64_t)c0 * n, all, (uint64_t)kc * n * 4, hipMemcpyDeviceToHost,
                                 g->stream));
        if (c0 + kc == k) g->last_tree = all + (uint64_t)(kc - 1) * n;
        BYO_HIP(hipStreamSynchronize(g->stream));
      }
      g->last_src = srcs[k - 1];
      return;
    }
    if (!levels_mode) {
      // default: every source in one persistent cooperative launch
      const bool keep_all = parents != nullptr && k > 1;
      std::vector<byo_bfs_stats> tmp(stats ? 0 : k);
      byo_bfs_stats *st = stats ? stats : tmp.data();
      double ms = 0;
      bfs_coop_run(g, srcs, k, keep_all, st, mx, &ms);
      if (g->prof.on) {
        auto &acc = g->prof.acc["bfs"];
        acc.launches += k;
        acc.ms += ms;
        g->account("bfs_launch", ms);
      }
      const uint64_t n = g->n;
      if (keep_all) {
        uint32_t *all = g->scratch["bfs_parents_all"].as<uint32_t>();
        BYO_HIP(hipMemcpyAsync(parents, all, (uint64_t)k * n * 4, hipMemcpyDeviceToHost, g->stream));
      } else if (parents) {
        BYO_HIP(hipMemcpyAsync(parents, g->last_tree, n * 4, hipMemcpyDeviceToHost, g->stream));
      }
      g->last_src = srcs[k - 1];
      return;
    }
    BfsRun r = bfs_setup(g);
    for (uint32_t i = 0; i < k; i++) {
      byo_bfs_stats *st = stats ? &stats[i] : nullptr;
      if (st) std::memset(st, 0, sizeof(*st));
      PhaseTimer t(g->stream, true);
      bfs_run(r, srcs[i], st, mx);
      double ms = t.stop();
      if (st) st->device_ms = ms;
      if (g->prof.on) g->account("bfs", ms);
      if (parents) {
        BYO_HIP(hipMemcpyAsync(parents + (uint64_t)i * g->n, r.parents, g->n * sizeof(uint32_t),
                               hipMemcpyDeviceToHost, g->stream));
        BYO_HIP(hipStreamSynchronize(g->stream));
      }
    }
    BYO_HIP(hipStreamSynchronize(g->stream));
    g->last_src = srcs[k - 1];
    g->last_tree = r.parents;
  });
}

int byo_bfs(byo_graph_t g, uint32_t src, uint32_t *parents, byo_bfs_stats *stats) {
  return byo_bfs_multi(g, &src, 1, parents, stats);
}

int byo_bfs_levels(byo_graph_t g, uint32_t *levels) {
  return guard([&] {
    BYO_REQUIRE(g && levels, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    BYO_REQUIRE(g->last_src >= 0, BYO_E_STATE, "no BFS has run on this handle");
    const uint32_t *dep = bfs_levels_dev(g);
    BYO_HIP(hipMemcpyAsync(levels, dep, g->n * 4, hipMemcpyDeviceToHost, g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

}  // extern "C"
