// bfs_coop.hip -- the whole BFS (all edgeMap levels, for k sources) in one
// persistent cooperative launch.
//
// Same traversal as the level-by-level path in bfs.hip, restating BFS.h:47-64
// over edgeMapData (gbbs/edge_map_data.h:254-307), but every level's
// direction is decided on the device from counters the previous level left
// in HBM, so no host round trip happens between levels.
//
// Frontier representations (vertexSubsetData, gbbs/vertex_subset.h:203-371):
// the bitmap (BitArray) of the next frontier is produced by every direction
// (ballot bytes in pull, atomicOr by CAS winners in push/forward); the queue
// (sparse ids) by push and forward.  A sparse level after a dense one packs
// the bitmap into a queue first (toSparse, vertex_subset.h:339-354).
//
// The kernels are shaped for latency hiding rather than for one thread per
// vertex:
//  * sparse push (edgeMapSparse, gbbs/edge_map_blocked.h:18-92): the edges
//    of the frontier are cut into chunks of 256; the scan pass also records
//    the first frontier index of every chunk, so a wave loads the (scan,
//    adjacency base) window of its chunk into LDS with coalesced loads (64
//    entries, up to 256 for frontiers of low degree) and resolves each
//    edge's owner by an LDS binary search;
//  * dense pull (edgeMapDense, edge_map_data.h:47-175): a wave owns one
//    64-vertex bitmap word per step; 8 lanes per unvisited vertex read 8
//    consecutive neighbours at once and test them against the bitmap,
//    exiting at the first frontier member (the lowest such neighbour in
//    adjacency order becomes the parent, as the reference's sequential
//    early-exit scan picks it); the first 8 neighbours of all 64 vertices are
//    probed with independent loads, and the next-frontier and visited words
//    are written whole, without atomics;
//  * dense forward (edgeMapDenseForward, :180-247): one wave per frontier
//    vertex, 64 neighbours per step.
//
// Grid barriers per level: push 3 (1 for frontiers of at most 4 vertices
// per block; +1 when packing), pull 1, forward 1.
// Frontier counters rotate over three slots so that resetting the slot for
// level L+2 never races with readers of level L or writers of level L+1.
#include <algorithm>
#include <cmath>

#include "coop.hpp"
#include "primitives.hpp"

namespace byo {

namespace {

struct Ctr {
  unsigned long long size;  // frontier size (queue cursor in push/forward)
  unsigned long long deg;   // sum of out-degrees of the frontier
  unsigned long long cursor;  // pack cursor of the level using this slot
  unsigned long long pad;
};

struct LevelLog {
  unsigned long long fsize, fdeg;
  uint32_t dir, pad;
  unsigned long long t0;  // wall clock (100 MHz) at the start of the level
};

struct CoopArgs {
  const uint64_t *off;  // adjacency view (AdjView): CSR, or the PMA's slot starts
  const uint32_t *ids;
  const uint32_t *deg;  // true degrees when the lists hold empty slots
  uint64_t n, m, nw64;
  uint32_t *parents;        // parents of source j at parents + j * pstride
  uint64_t pstride;
  uint32_t *q0, *q1;
  uint64_t *bm;              // three rotating bitmaps of nw64 words
  uint64_t *visited64;
  uint64_t *scan;
  uint32_t *cstart;         // first frontier index of each 256-edge chunk
  unsigned long long *bsum;
  Ctr *ctr;                 // 3 slots
  const uint32_t *srcs;
  uint32_t k;
  LevelLog *logs;           // k x 64
  uint32_t *nlevels;        // k
  unsigned long long *tot;  // k x {vertices reached, sum of their degrees}
  DirCtl dc;
  Pieces pieces;            // weight-balanced word pieces of the grid
  GridBar *bar;
};

constexpr int kBlock = kCoopBlock;
constexpr int kWaves = kCoopWaves;
constexpr int kPerLane = 4;
constexpr uint64_t kChunk = 64 * kPerLane;
constexpr int kG = 8;  // lanes per vertex in the pull

__device__ __forceinline__ bool bit_get(const uint64_t *bm, uint32_t i) { return (bm[i >> 6] >> (i & 63)) & 1ull; }

// A per-block filter in LDS of vertices known to be claimed: direct-mapped
// by a multiplicative hash, an entry holding a vertex id.  An id is stored
// only once its bit has been seen set in the next-frontier bitmap (or won),
// so a hit means the vertex is discovered for good and the global tests
// are skipped.  RMAT hubs take thousands of edges of one frontier; their
// bitmap words would otherwise be read in L2 by every such edge, and all of
// those reads queue on the one L2 channel that holds the word.
constexpr int kFiltLog = 11;
constexpr int kFilt = 1 << kFiltLog;
__device__ __forceinline__ uint32_t filt_slot(uint32_t t) { return (t * 2654435761u) >> (32 - kFiltLog); }

// Claims vertex t for the next frontier: t is in neither the visited bitmap
// nor the current frontier's, and this thread's atomicOr set its bit in the
// next frontier's bitmap (one winner per vertex, which then writes the
// parent with a plain store).  The visited bitmap lags by one level: a push
// or forward level folds the current frontier into it while it runs, a pull
// level as it writes each word.
// A relaxed device-scope read of the next bitmap first skips the atomic for a
// vertex already claimed this level.
__device__ __forceinline__ bool claim(const uint64_t *visited, const uint64_t *fbm, uint64_t *nbm, uint32_t t,
                                      uint32_t *filt) {
  const uint32_t slot = filt_slot(t);
  if (filt[slot] == t) return false;
  const uint64_t w = t >> 6, bit = 1ull << (t & 63);
  if ((visited[w] | fbm[w]) & bit) return false;
  unsigned long long *p = reinterpret_cast<unsigned long long *>(&nbm[w]);
  const bool won = !(__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) & bit) &&
                   !(atomicOr(p, (unsigned long long)bit) & bit);
  filt[slot] = t;
  return won;
}

// claim() over K candidate targets of one lane, stage by stage: all K filter
// reads, then all K discovered tests, then all K relaxed reads of the next
// bitmap, then all K atomicOr, so a lane waits for each stage once.  ok[t]
// is in: t is a candidate; out: t was claimed by this thread.  Empty slots
// (kNone) are no candidates when G.
template <int K, bool G>
__device__ __forceinline__ void claim_staged(const uint64_t *visited, const uint64_t *fbm, uint64_t *nbm,
                                             const uint32_t (&tg)[K], bool (&ok)[K], uint32_t *filt) {
  uint64_t x[K], y[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    if (G && tg[t] == kNone) ok[t] = false;
    if (ok[t] && filt[filt_slot(tg[t])] == tg[t]) ok[t] = false;
  }
#pragma unroll
  for (int t = 0; t < K; t++) {
    x[t] = 0;
    y[t] = 0;
    if (ok[t]) {
      x[t] = visited[tg[t] >> 6];
      y[t] = fbm[tg[t] >> 6];
    }
  }
  bool tried[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    const uint64_t bit = 1ull << (tg[t] & 63);
    ok[t] = ok[t] && !((x[t] | y[t]) & bit);
    tried[t] = ok[t];
    if (ok[t])
      x[t] = __hip_atomic_load(reinterpret_cast<unsigned long long *>(&nbm[tg[t] >> 6]), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
  }
#pragma unroll
  for (int t = 0; t < K; t++) {
    const uint64_t bit = 1ull << (tg[t] & 63);
    ok[t] = ok[t] && !(x[t] & bit);
    if (ok[t]) x[t] = atomicOr(reinterpret_cast<unsigned long long *>(&nbm[tg[t] >> 6]), (unsigned long long)bit);
  }
#pragma unroll
  for (int t = 0; t < K; t++) {
    ok[t] = ok[t] && !(x[t] & (1ull << (tg[t] & 63)));
    if (tried[t]) filt[filt_slot(tg[t])] = tg[t];
  }
}

// Push levels with at least this many frontier edges mark instead of
// claiming.  The bitmap claim costs an L2 atomic per candidate edge, and
// those atomics queue on the words they share (64 vertices a word, hubs
// among them): 3.8 M candidates of an RMAT-22 level took 0.3 ms.  The mark
// is the parent itself: a plain store of whichever frontier neighbour
// stores last (each is a valid BFS parent: the target is undiscovered and
// the neighbour in the frontier), into a parent array that holds kNone for
// every undiscovered vertex; one pass over the parents of the vertices not
// yet discovered then builds the next bitmap and queue.
constexpr uint64_t kMarkMinEdges = 1ull << 20;

// The mark counterpart of claim_staged: targets that are in neither the
// visited bitmap nor the current frontier get parents[t] = their owner.  The
// LDS filter skips targets this block has marked already (and any claimed
// before).
template <int K, bool G>
__device__ __forceinline__ void mark_staged(const uint64_t *visited, const uint64_t *fbm, uint32_t *parents,
                                            const uint32_t (&tg)[K], const uint32_t (&own)[K], bool (&ok)[K],
                                            uint32_t *filt) {
  uint64_t x[K];
#pragma unroll
  for (int t = 0; t < K; t++) {
    if (G && tg[t] == kNone) ok[t] = false;
    if (ok[t] && filt[filt_slot(tg[t])] == tg[t]) ok[t] = false;
  }
#pragma unroll
  for (int t = 0; t < K; t++) {
    x[t] = ~0ull;
    if (ok[t]) x[t] = visited[tg[t] >> 6] | fbm[tg[t] >> 6];
  }
#pragma unroll
  for (int t = 0; t < K; t++) {
    if (ok[t] && !((x[t] >> (tg[t] & 63)) & 1)) {
      parents[tg[t]] = own[t];
      filt[filt_slot(tg[t])] = tg[t];
    }
  }
}

// After a marking push: every wave takes runs of 32 bitmap words; a vertex
// (lane) of a run is new when it has a parent but is neither visited nor in
// the current frontier; the pass writes the next frontier's bitmap words
// whole and appends the new vertices to the next queue (in id order within a
// run, one cursor atomic per run).  Returns the lane's share of the next
// frontier's degree sum.
template <bool G>
__device__ __forceinline__ unsigned long long pack_marks(const CoopArgs &a, const uint32_t *parents,
                                                         const uint64_t *fbm, uint64_t *nbm, uint32_t *qn,
                                                         unsigned long long *cursor, uint64_t nwords, uint64_t wave,
                                                         uint64_t nwaves, int lane) {
  constexpr int kRun = 32;
  unsigned long long dsum = 0;
  for (uint64_t r0 = wave * kRun; r0 < nwords; r0 += nwaves * kRun) {
    uint64_t mine = 0;  // lane s keeps the word of step s
#pragma unroll 4
    for (int st = 0; st < kRun; st++) {
      const uint64_t w = r0 + st;
      if (w >= nwords) break;
      const uint64_t v = w * 64 + lane;
      const uint64_t old = a.visited64[w] | fbm[w];
      const bool on = v < a.n && parents[v] != kNone && !((old >> lane) & 1);
      const uint64_t word = __ballot(on);
      if (on) dsum += vdeg<G>(a.off, a.deg, v);
      if (lane == st) mine = word;
      if (lane == 0) nbm[w] = word;
    }
    const uint32_t c = (uint32_t)__popcll(mine);
    const uint32_t incl = wave_incl_scan(c);
    const uint32_t tot = __shfl(incl, 63, 64);
    if (!tot) continue;
    unsigned long long base = 0;
    if (lane == 0) base = atomicAdd(cursor, (unsigned long long)tot);
    base = __shfl(base, 0, 64) + (incl - c);  // lane s: start of step s's vertices
    for (int st = 0; st < kRun; st++) {
      const uint64_t word = __shfl(mine, st, 64);
      const unsigned long long b = __shfl(base, st, 64);
      if ((word >> lane) & 1) qn[b + mbcnt(word)] = (uint32_t)((r0 + st) * 64 + lane);
    }
  }
  return dsum;
}

// (scan, adjacency base) of up to kChunk frontier entries per wave: a chunk
// of kChunk edges spans at most kChunk entries of positive degree
struct PushLds {
  uint64_t sc[kWaves][kChunk];
  int64_t base[kWaves][kChunk];
};

// Adds the wave-uniform totals x to *px and y to *py for the whole block,
// one atomic per counter and block (thousands of same-address atomics at
// the end of a level, one per wave, queue in one L2 channel).  Every thread
// of the block calls it; px may be null.
__device__ __forceinline__ void block_add2(unsigned long long *px, unsigned long long x, unsigned long long *py,
                                           unsigned long long y, int lane) {
  __shared__ unsigned long long red[2][kWaves];
  if (lane == 0) {
    red[0][threadIdx.x >> 6] = x;
    red[1][threadIdx.x >> 6] = y;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long sx = 0, sy = 0;
    for (int w = 0; w < kWaves; w++) {
      sx += red[0][w];
      sy += red[1][w];
    }
    if (px && sx) atomicAdd(px, sx);
    if (sy) atomicAdd(py, sy);
  }
  __syncthreads();  // red is rewritten by the next call
}

// G: the adjacency lists may hold empty slots (kNone), skipped everywhere.
template <bool G>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) k_bfs_coop(CoopArgs a) {
  __shared__ unsigned long long lds[8];
  __shared__ MlpLds mlp;
  __shared__ PushLds win;
  __shared__ uint32_t filt[kFilt];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint64_t nb = gridDim.x;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const uint64_t nwords = (a.n + 63) >> 6;
  const uint64_t nbytes = (a.n + 7) >> 3;

  pieces_init(a.pieces, a.off, a.n, nwords, nwaves);
  mlp_stage(a.dc, mlp);
  for (uint32_t j = 0; j < a.k; j++) {
    uint32_t *parents = a.parents + (uint64_t)j * a.pstride;
    const uint32_t src = a.srcs[j];
    // one pass seeds the BFS: parents[src] = src, the source's bit in the
    // visited and frontier bitmaps, q[0] = src, counters of level 0.  Both
    // bitmaps start clear elsewhere: the pull writes only the bytes of real
    // vertices, so the padding of the last word must already be zero.
    for (uint64_t i = global_tid(); i < a.n; i += grid_threads()) parents[i] = (i == src) ? src : kNone;
    for (int i = threadIdx.x; i < kFilt; i += kBlock) filt[i] = kNone;  // synchronised by the barrier below
    const uint64_t sw = src >> 6, sbit = 1ull << (src & 63);
    for (uint64_t i = global_tid(); i < a.nw64; i += grid_threads()) {
      a.visited64[i] = (i == sw) ? sbit : 0;
      a.bm[i] = (i == sw) ? sbit : 0;
      a.bm[a.nw64 + i] = 0;
      a.bm[2 * a.nw64 + i] = 0;
    }
    if (leader) {
      a.q0[0] = src;
      a.ctr[0] = Ctr{1, vdeg<G>(a.off, a.deg, src), 0, 0};
      a.ctr[1] = Ctr{0, 0, 0, 0};
      a.ctr[2] = Ctr{0, 0, 0, 0};
    }
    grid_barrier(a.bar, (unsigned)nb);

    uint32_t *q = a.q0, *qn = a.q1;
    bool is_dense = false, have_q = true;
    uint64_t visited_deg = 0, reached = 0;
    int L = 0;
    for (;; L++) {
      Ctr *cur = &a.ctr[L % 3];
      const uint64_t fs = ld_relaxed(&cur->size);
      const uint64_t fdeg = ld_relaxed(&cur->deg);
      if (fs == 0) break;
      reached += fs;
      // this engine's dense pull already tests a bitmap frontier: a
      // scheduled bitmap pull (4) is its pull
      const int dsel = decide(a.dc, is_dense, fs, fdeg, visited_deg, L, mlp);
      const int dir = dsel == 4 ? 1 : dsel;
      if (leader && L < 64) a.logs[(uint64_t)j * 64 + L] = LevelLog{fs, fdeg, (uint32_t)dir, 0, (unsigned long long)wall_clock64()};
      if (dir == 2) {
        visited_deg += fdeg;
        L++;
        break;
      }
      Ctr *nx = &a.ctr[(L + 1) % 3];
      if (leader) a.ctr[(L + 2) % 3] = Ctr{0, 0, 0, 0};
      // bitmaps rotate like the counters: B[L%3] is this frontier, the next
      // one is built in B[(L+1)%3], and B[(L+2)%3] -- last level's
      // frontier, read by nobody now -- is cleared for level L+1
      // offsets from one base keep these global (a runtime index into an
      // array of pointers would make every access a flat one)
      uint64_t *const fbm = a.bm + (uint64_t)(L % 3) * a.nw64, *const nbm = a.bm + (uint64_t)((L + 1) % 3) * a.nw64;
      {
        uint64_t *clr = a.bm + (uint64_t)((L + 2) % 3) * a.nw64;
        for (uint64_t i = global_tid(); i < a.nw64; i += grid_threads()) clr[i] = 0;
      }
      if (dir != 1) {
        // push and forward claim against visited | frontier; the frontier is
        // folded into the visited bitmap while they run (no other writer)
        for (uint64_t i = global_tid(); i < a.nw64; i += grid_threads()) {
          const uint64_t f = fbm[i];
          if (f) a.visited64[i] |= f;
        }
      }

      if (dir == 0) {
        if (!have_q) {
          // toSparse: pack the bitmap into the queue (any order)
          for (uint64_t w0 = wave * 64; w0 < nwords; w0 += nwaves * 64) {
            const uint64_t w = w0 + lane;
            uint64_t word = w < nwords ? fbm[w] : 0;
            const uint32_t c = __popcll(word);
            const uint32_t incl = wave_incl_scan(c);
            const uint32_t tot = __shfl(incl, 63, 64);
            unsigned long long base = 0;
            if (lane == 0 && tot) base = atomicAdd(&cur->cursor, (unsigned long long)tot);
            base = __shfl(base, 0, 64);
            uint64_t pos = base + incl - c;
            while (word) {
              const int bp = __builtin_ctzll(word);
              word &= word - 1;
              q[pos++] = (uint32_t)(w * 64 + bp);
            }
          }
          grid_barrier(a.bar, (unsigned)nb);
        }
        if (fs <= 4 * nb || (fs <= 16 * nwaves && fdeg <= 32 * fs)) {
          // small frontier, or a moderate one of low average degree: one
          // block (fs <= #blocks) or one wave per frontier vertex, 1 barrier
          // instead of 3
          const bool per_block = fs <= nb;
          const uint64_t unit = per_block ? blockIdx.x : wave, nunits = per_block ? nb : nwaves;
          const int width = per_block ? kBlock : 64;
          const int lid = per_block ? (int)threadIdx.x : lane;
          unsigned long long dsum = 0;
          for (uint64_t i = unit; i < fs; i += nunits) {
            const uint32_t u = q[i];
            const uint64_t b = a.off[u], e = a.off[u + 1];
            for (uint64_t k0 = b; k0 < e; k0 += width) {
              const uint64_t kk = k0 + lid;
              bool win_ = false;
              uint32_t tgt = 0;
              if (kk < e) {
                tgt = a.ids[kk];
                if ((!G || tgt != kNone) && claim(a.visited64, fbm, nbm, tgt, filt)) {
                  win_ = true;
                  parents[tgt] = u;
                  dsum += vdeg<G>(a.off, a.deg, tgt);
                }
              }
              const uint64_t mask = __ballot(win_);
              if (mask) {
                unsigned long long base = 0;
                if (lane == 0) base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
                base = __shfl(base, 0, 64);
                if (win_) qn[base + mbcnt(mask)] = tgt;
              }
            }
          }
          dsum = wave_sum(dsum);
          block_add2(nullptr, 0, &nx->deg, dsum, lane);
          grid_barrier(a.bar, (unsigned)nb);
          is_dense = false;
          have_q = true;
          visited_deg += fdeg;
          uint32_t *tq = q;
          q = qn;
          qn = tq;
          continue;
        }
        // (1) slice sums of frontier degrees
        const uint64_t S = (fs + nb - 1) / nb;
        const uint64_t s0 = std::min<uint64_t>(fs, blockIdx.x * S), s1 = std::min<uint64_t>(fs, s0 + S);
        unsigned long long part = 0;
        for (uint64_t i = s0 + threadIdx.x; i < s1; i += kBlock) {
          const uint32_t v = q[i];
          part += a.off[v + 1] - a.off[v];
        }
        part = block_sum(part, lds);
        if (threadIdx.x == 0) a.bsum[blockIdx.x] = part;
        grid_barrier(a.bar, (unsigned)nb);
        // (2) exclusive scan of the queue degrees and the chunk starts
        unsigned long long pre = 0;
        for (uint64_t b = threadIdx.x; b < blockIdx.x; b += kBlock) pre += a.bsum[b];
        pre = block_sum(pre, lds);
        for (uint64_t c0 = s0; c0 < s1; c0 += kBlock) {
          const uint64_t i = c0 + threadIdx.x;
          unsigned long long d = 0;
          if (i < s1) {
            const uint32_t v = q[i];
            d = a.off[v + 1] - a.off[v];
          }
          const unsigned long long incl = wave_incl_scan(d);
          __syncthreads();
          if (lane == 63) lds[wv] = incl;
          __syncthreads();
          unsigned long long wpre = 0, tot = 0;
          for (int w = 0; w < kWaves; w++) {
            if (w < wv) wpre += lds[w];
            tot += lds[w];
          }
          if (i < s1) {
            const uint64_t sc = pre + wpre + incl - d;
            a.scan[i] = sc;
            // chunks whose first edge falls in [sc, sc + d) start at i
            for (uint64_t c = (sc + kChunk - 1) / kChunk; c * kChunk < sc + d; c++) a.cstart[c] = (uint32_t)i;
          }
          pre += tot;
        }
        // the scan's total: the frontier's edges, or its slots when the lists
        // hold empty slots (written by the block holding the last slice)
        if constexpr (G) {
          if (threadIdx.x == 0 && s0 < s1 && s1 == fs) a.scan[fs] = pre;
        } else {
          if (leader) a.scan[fs] = fdeg;
        }
        grid_barrier(a.bar, (unsigned)nb);
        // (3) edge-balanced push, 256 edges (slots) per wave step
        const uint64_t E = G ? ld_relaxed(reinterpret_cast<const unsigned long long *>(&a.scan[fs])) : fdeg;
        const uint64_t nch = (E + kChunk - 1) / kChunk;
        const bool marks = E >= kMarkMinEdges;
        unsigned long long dsum = 0;
        for (uint64_t c = wave; c < nch; c += nwaves) {
          const uint64_t e0 = c * kChunk;
          const uint64_t i0 = a.cstart[c];
          // the window: 64 entries, and 64 more while the last one loaded
          // still starts before the chunk's end (low-degree frontiers)
          int nwin = 0;
          for (int r = 0; r < (int)(kChunk / 64); r++) {
            const uint64_t idx = i0 + (uint64_t)r * 64 + lane;
            uint64_t sc = ~0ull;
            int64_t bs = 0;
            if (idx < fs) {
              sc = a.scan[idx];
              bs = (int64_t)a.off[q[idx]] - (int64_t)sc;
            }
            win.sc[wv][r * 64 + lane] = sc;
            win.base[wv][r * 64 + lane] = bs;
            nwin += 64;
            if (__shfl(sc, 63, 64) >= e0 + kChunk) break;
          }
          wave_lds_sync();
          // the lane's kPerLane edges advance through each stage of the
          // claim together (ids, discovered test, claim, parent and degree),
          // so every stage has all its loads or atomics in flight at once
          uint32_t tg[kPerLane], own[kPerLane];
          bool wn[kPerLane];
#pragma unroll
          for (int t = 0; t < kPerLane; t++) {
            const uint64_t e = e0 + (uint64_t)t * 64 + lane;
            wn[t] = e < E;
            own[t] = 0;
            tg[t] = 0;
            if (wn[t]) {
              int lo = 0, hi = nwin;  // sc[lo] <= e (sc[0] <= e0 by construction)
              while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (win.sc[wv][mid] <= e) lo = mid; else hi = mid;
              }
              int64_t bs;
              if (lo == nwin - 1 && i0 + nwin < fs && a.scan[i0 + nwin] <= e) {
                // the chunk spans more than kChunk frontier entries (entries
                // of degree 0: a directed graph)
                uint64_t x = i0 + nwin, y = fs;
                while (y - x > 1) {
                  const uint64_t mid = (x + y) >> 1;
                  if (a.scan[mid] <= e) x = mid; else y = mid;
                }
                own[t] = q[x];
                bs = (int64_t)a.off[own[t]] - (int64_t)a.scan[x];
              } else {
                own[t] = q[i0 + lo];
                bs = win.base[wv][lo];
              }
              tg[t] = a.ids[(uint64_t)(bs + (int64_t)e)];
            }
          }
          if (marks) {
            // large frontier: mark, do not claim (the pack pass below builds
            // the next frontier from the marks)
            mark_staged<kPerLane, G>(a.visited64, fbm, parents, tg, own, wn, filt);
            wave_lds_sync();  // the window is rewritten by the next chunk
            continue;
          }
          claim_staged<kPerLane, G>(a.visited64, fbm, nbm, tg, wn, filt);
          uint32_t total = 0;
#pragma unroll
          for (int t = 0; t < kPerLane; t++) {
            if (wn[t]) {
              parents[tg[t]] = own[t];
              dsum += vdeg<G>(a.off, a.deg, tg[t]);
            }
            total += __popcll(__ballot(wn[t]));
          }
          if (total) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&nx->size, (unsigned long long)total);
            base = __shfl(base, 0, 64);
#pragma unroll
            for (int t = 0; t < kPerLane; t++) {
              const uint64_t mask = __ballot(wn[t]);
              if (wn[t]) qn[base + mbcnt(mask)] = tg[t];
              base += __popcll(mask);
            }
          }
          wave_lds_sync();  // the window is rewritten by the next chunk
        }
        if (marks) {
          grid_barrier(a.bar, (unsigned)nb);
          dsum += pack_marks<G>(a, parents, fbm, nbm, qn, &nx->size, nwords, wave, nwaves, lane);
        }
        dsum = wave_sum(dsum);
        block_add2(nullptr, 0, &nx->deg, dsum, lane);
        grid_barrier(a.bar, (unsigned)nb);
        is_dense = false;
        have_q = true;
      } else if (dir == 1) {
        // dense pull: a wave owns one 64-vertex bitmap word per step.
        // Offsets of the 64 vertices are loaded once (lane = vertex); then
        // every unvisited vertex probes its first 8 neighbours with 8 lanes
        // (8 vertices x 8 steps, all loads independent); vertices still
        // unresolved continue 8 neighbours per step, 8 vertices at a time.
        const int g = lane >> 3, sub = lane & 7;
        unsigned long long cnt = 0, dsum = 0;
        for_my_words(a.pieces, wave, nwaves, [&](uint64_t w) {
          // discovered = visited or in the current frontier (the bitmap
          // lags by one level after a push)
          const uint64_t v0 = a.visited64[w];
          const uint64_t vis = v0 | fbm[w];
          const uint64_t vl = w * 64 + lane;
          uint64_t bl = 0, el = 0;
          if (vl < a.n) {
            bl = a.off[vl];
            el = a.off[vl + 1];
          }
          const uint64_t todo = __ballot(vl < a.n && !((vis >> lane) & 1));
          uint64_t fmask = 0;
          if (todo) {
            uint32_t u[8];
            bool h[8];
#pragma unroll
            for (int st = 0; st < 8; st++) {
              const int x = st * 8 + g;
              const uint64_t b = __shfl(bl, x, 64), e = __shfl(el, x, 64);
              const uint64_t kk = b + sub;
              h[st] = ((todo >> x) & 1) && kk < e;
              u[st] = a.ids[h[st] ? kk : 0];  // unconditional: all 8 loads in flight
            }
            if constexpr (G) {
              // an empty slot (or an empty slot 0 read by an idle lane) is no
              // neighbour; probe word 0 instead of an out-of-range one
#pragma unroll
              for (int st = 0; st < 8; st++)
                if (u[st] == kNone) {
                  u[st] = 0;
                  h[st] = false;
                }
            }
            uint64_t fw[8];
#pragma unroll
            for (int st = 0; st < 8; st++) fw[st] = fbm[u[st] >> 6];
#pragma unroll
            for (int st = 0; st < 8; st++) h[st] = h[st] && ((fw[st] >> (u[st] & 63)) & 1);
#pragma unroll
            for (int st = 0; st < 8; st++) {
              const uint32_t gm = (uint32_t)(__ballot(h[st]) >> (g * 8)) & 0xFFu;
              const uint32_t pu = __shfl(u[st], g * 8 + (gm ? __builtin_ctz(gm) : 0), 64);
              if (gm && sub == 0) parents[w * 64 + st * 8 + g] = pu;
              const uint64_t fb = __ballot(gm != 0 && sub == 0);
              uint64_t byte = 0;
#pragma unroll
              for (int t = 0; t < 8; t++) byte |= ((fb >> (t * 8)) & 1ull) << t;
              fmask |= byte << (st * 8);
            }
            // vertices with more than 8 neighbours and no hit yet
            uint64_t rem = todo & ~fmask & __ballot(el - bl > 8);
            while (rem) {
              int x = -1;
              uint64_t mm = rem;
              for (int t = 0; t < 8 && mm; t++) {
                const int bp = __builtin_ctzll(mm);
                mm &= mm - 1;
                if (t == g) x = bp;
              }
              rem = mm;
              const bool has = x >= 0;
              const uint64_t b = __shfl(bl, has ? x : 0, 64) + 8, e = __shfl(el, has ? x : 0, 64);
              bool found = false;
              uint32_t par = 0;
              for (uint64_t k0 = b;; k0 += 8) {
                const bool active = has && !found && k0 < e;
                if (!__ballot(active)) break;
                const uint64_t kk = k0 + sub;
                uint32_t uu = 0;
                bool hit = false;
                if (active && kk < e) {
                  uu = a.ids[kk];
                  hit = (!G || uu != kNone) && bit_get(fbm, uu);
                }
                const uint32_t gm = (uint32_t)(__ballot(hit) >> (g * 8)) & 0xFFu;
                const uint32_t pu = __shfl(uu, g * 8 + (gm ? __builtin_ctz(gm) : 0), 64);
                if (active && gm) {
                  found = true;
                  par = pu;
                }
              }
              uint64_t mine = 0;
              if (found && sub == 0) {
                parents[w * 64 + x] = par;
                mine = 1ull << x;
              }
#pragma unroll
              for (int o = 32; o > 0; o >>= 1) mine |= __shfl_xor(mine, o, 64);
              fmask |= mine;
            }
          }
          if (lane == 0) {
            nbm[w] = fmask;
            if ((vis | fmask) != v0) a.visited64[w] = vis | fmask;
          }
          cnt += __popcll(fmask);
          if ((fmask >> lane) & 1) dsum += vdeg<G>(a.off, a.deg, vl);
        });
        dsum = wave_sum(dsum);
        block_add2(&nx->size, cnt, &nx->deg, dsum, lane);
        grid_barrier(a.bar, (unsigned)nb);
        is_dense = true;
        have_q = false;
      } else {
        // dense forward: one wave per frontier vertex of each bitmap word
        for_my_words(a.pieces, wave, nwaves, [&](uint64_t w) {
          uint64_t word = fbm[w];
          while (word) {
            const int bp = __builtin_ctzll(word);
            word &= word - 1;
            const uint32_t u = (uint32_t)(w * 64 + bp);
            const uint64_t b = a.off[u], e = a.off[u + 1];
            for (uint64_t k0 = b; k0 < e; k0 += 64) {
              const uint64_t kk = k0 + lane;
              bool win_ = false;
              uint32_t t = 0;
              if (kk < e) {
                t = a.ids[kk];
                if ((!G || t != kNone) && claim(a.visited64, fbm, nbm, t, filt)) {
                  win_ = true;
                  parents[t] = u;
                }
              }
              const uint64_t mask = __ballot(win_);
              if (mask) {
                const unsigned long long d = wave_sum(win_ ? (unsigned long long)vdeg<G>(a.off, a.deg, t) : 0ull);
                unsigned long long base = 0;
                if (lane == 0) {
                  base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
                  atomicAdd(&nx->deg, d);
                }
                base = __shfl(base, 0, 64);
                if (win_) qn[base + mbcnt(mask)] = t;
              }
            }
          }
        });
        grid_barrier(a.bar, (unsigned)nb);
        is_dense = true;
        have_q = true;
      }
      visited_deg += fdeg;
      uint32_t *tq = q;
      q = qn;
      qn = tq;
    }
    if (leader) {
      if (L < 64) a.logs[(uint64_t)j * 64 + L].t0 = wall_clock64();
      a.nlevels[j] = (uint32_t)L;
      a.tot[2 * j] = reached;
      a.tot[2 * j + 1] = visited_deg;
    }
    grid_barrier(a.bar, (unsigned)nb);
  }
}

}  // namespace

GridBar *coop_bar(byo_graph *g) {
  GridBar *b = g->buf("bfs_coop_bar", sizeof(GridBar)).as<GridBar>();
  if (!g->coop_bar_ready) {
    BYO_HIP(hipMemsetAsync(b, 0, sizeof(GridBar), g->stream));
    g->coop_bar_ready = true;
  }
  return b;
}

DirCtl dir_ctl(byo_graph *g, uint64_t maxdeg, hipStream_t s, bool batched) {
  DirCtl d{};
  d.n = g->n;
  d.m = g->m;
  d.policy = g->dir_policy;
  d.thr = g->dir_threshold < 0 ? g->m / 20 : (uint64_t)g->dir_threshold;
  if (g->dir_policy == BYO_DIR_MLP) {
    if (batched && g->sel_loaded_b) d.sel = g->sel_params_b.as<float>();
    if (!batched && g->sel_loaded) d.sel = g->sel_params.as<float>();
  }
  d.maxdeg = maxdeg;
  d.margin = (float)g->tuning(batched ? "sel_margin_batch_x100" : "sel_margin_x100") / 100.0f;
  d.third = batched ? 4 : 3;
  d.densify = g->tuning("sel_densify") != 0;
  d.range = std::min(kSelClamp, (float)g->tuning("sel_range_x100") / 100.0f);
  uint8_t *dsched = g->buf("bfs_coop_sched", 64).as<uint8_t>();
  if (g->dir_policy == BYO_DIR_SCHEDULE) BYO_HIP(hipMemcpyAsync(dsched, g->dir_sched, 64, hipMemcpyHostToDevice, s));
  d.sched = dsched;
  return d;
}

int coop_grid(const void *kernel, int block, int cap) {
  static std::mutex mu;
  static std::map<std::pair<int, const void *>, std::pair<int, int>> cache;  // (CUs, resident blocks per CU)
  int dev = 0;
  BYO_HIP(hipGetDevice(&dev));
  std::pair<int, int> cp;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find({dev, kernel});
    if (it == cache.end()) {
      int cus = 0, per = 0;
      BYO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
      BYO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, kernel, block, 0));
      BYO_REQUIRE(per >= 1, BYO_E_HIP, "cooperative kernel cannot be resident");
      it = cache.emplace(std::make_pair(dev, kernel), std::make_pair(cus, per)).first;
    }
    cp = it->second;
  }
  const int nb = cp.first * std::max(1, std::min(cp.second, cap));
  BYO_REQUIRE(nb <= 2048, BYO_E_HIP, "grid barrier sized for at most 2048 blocks");
  return nb;
}

int traversal_grid(byo_graph *g, const void *kernel, int block, uint64_t slots) {
  const int nb = coop_grid(kernel, block, (int)g->tuning("coop_blocks_per_cu"));
  const uint64_t div = (uint64_t)g->tuning("coop_grid_sqrt_div");
  if (!div) return nb;
  const uint64_t want = std::max<uint64_t>(8, ((uint64_t)std::sqrt((double)slots) / div + 7) / 8 * 8);
  return (int)std::min<uint64_t>((uint64_t)nb, want);
}

Pieces make_pieces(byo_graph *g, int nblocks, PieceSet set) {
  Pieces pc;
  pc.per_wave = (uint32_t)g->tuning(set == kPprPieces ? "ppr_pieces" : "pieces");
  pc.bound = g->buf(set == kPprPieces ? "ppr_pieces" : (set == kPmaPieces ? "pma_pieces" : "bfs_pieces"),
                    ((uint64_t)kMaxPieces * 2048 * 4 + 1) * 8 + 64)
                 .as<uint64_t>();
  const byo_graph::PieceCache &c = g->pcache[set];
  pc.ready = c.valid && c.nblocks == nblocks && c.per_wave == pc.per_wave;
  return pc;
}

void pieces_launched(byo_graph *g, const Pieces &pc, int nblocks, PieceSet set) {
  byo_graph::PieceCache &c = g->pcache[set];
  c.valid = true;
  c.nblocks = nblocks;
  c.per_wave = pc.per_wave;
}

// Runs k BFS in one cooperative launch.  Results: k*n parents in scratch
// "bfs_parents_all" (keep_all) or the last source's tree in "bfs_parents".
// Fills stats[k] when given.
void bfs_coop_run(byo_graph *g, const uint32_t *srcs, uint32_t k, bool keep_all, byo_bfs_stats *stats,
                  uint64_t maxdeg, double *ms_out) {
  const uint64_t n = g->n;
  hipStream_t s = g->stream;
  CoopArgs a{};
  const AdjView v = adj_view(g);
  const bool gaps = v.deg != nullptr;
  a.off = v.off;
  a.ids = v.ids;
  a.deg = v.deg;
  a.n = n;
  a.m = g->m;
  a.nw64 = (n + 63) / 64 + 1;
  a.pstride = keep_all ? n : 0;
  a.parents = g->buf(keep_all ? "bfs_parents_all" : "bfs_parents", (keep_all ? k : 1) * n * 4 + 64).as<uint32_t>();
  g->last_tree = a.parents + (keep_all ? (uint64_t)(k - 1) * n : 0);
  a.q0 = g->buf("bfs_q", n * 4 + 64).as<uint32_t>();
  a.q1 = g->buf("bfs_qn", n * 4 + 64).as<uint32_t>();
  a.bm = g->buf("bfs_bms", 3 * a.nw64 * 8).as<uint64_t>();
  a.visited64 = g->buf("bfs_visited", a.nw64 * 8).as<uint64_t>();
  a.scan = g->buf("bfs_scan", (n + 1) * 8 + 64).as<uint64_t>();
  a.cstart = g->buf("bfs_cstart", (v.nslots / kChunk + 2) * 4 + 64).as<uint32_t>();
  a.ctr = g->buf("bfs_coop_ctr", 3 * sizeof(Ctr) + 64).as<Ctr>();
  a.k = k;
  uint32_t *dsrc = g->buf("bfs_coop_srcs", (size_t)k * 4 + 64).as<uint32_t>();
  a.srcs = dsrc;
  a.logs = g->buf("bfs_coop_logs", (size_t)k * 65 * sizeof(LevelLog) + 64).as<LevelLog>();
  a.nlevels = g->buf("bfs_coop_nlev", (size_t)k * 4 + 64).as<uint32_t>();
  a.tot = g->buf("bfs_coop_tot", (size_t)k * 16 + 64).as<unsigned long long>();
  a.dc = dir_ctl(g, maxdeg, s, false);
  const void *kern = gaps ? reinterpret_cast<const void *>(k_bfs_coop<true>)
                         : reinterpret_cast<const void *>(k_bfs_coop<false>);
  const PieceSet pset = gaps ? kPmaPieces : kTraversalPieces;
  const int nblocks = traversal_grid(g, kern, kBlock, v.nslots);
  a.pieces = make_pieces(g, nblocks, pset);
  a.bsum = g->buf("bfs_coop_bsum", (size_t)nblocks * 8 + 64).as<unsigned long long>();
  a.bar = coop_bar(g);
  BYO_HIP(hipMemcpyAsync(dsrc, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
  void *args[] = {&a};
  PhaseTimer t(s, true);
  BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kBlock), args, 0, s));
  pieces_launched(g, a.pieces, nblocks, pset);
  const double ms = t.stop();
  if (ms_out) *ms_out = ms;
  if (stats) {
    int wall_khz = 0;
    BYO_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, g->device));
    const double wall_hz = wall_khz > 0 ? wall_khz * 1e3 : 1e8;
    std::vector<LevelLog> logs((size_t)k * 64);
    std::vector<uint32_t> nlev(k);
    std::vector<unsigned long long> tot(2 * (size_t)k);
    BYO_HIP(hipMemcpyAsync(tot.data(), a.tot, tot.size() * 8, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipMemcpyAsync(logs.data(), a.logs, logs.size() * sizeof(LevelLog), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipMemcpyAsync(nlev.data(), a.nlevels, k * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    for (uint32_t j = 0; j < k; j++) {
      byo_bfs_stats &st = stats[j];
      std::memset(&st, 0, sizeof(st));
      st.levels = nlev[j];
      for (uint32_t L = 0; L < std::min<uint32_t>(nlev[j], 64); L++) {
        const LevelLog &lg = logs[(size_t)j * 64 + L];
        st.fsize[L] = lg.fsize;
        st.fdeg[L] = lg.fdeg;
        st.dir[L] = (uint8_t)lg.dir;
        // the next record's start (or the end stamp) closes this level
        if (L + 1 < 64 && L + 1 <= nlev[j])
          st.level_ms[L] = (float)((double)(logs[(size_t)j * 64 + L + 1].t0 - lg.t0) / wall_hz * 1e3);
      }
      st.reached = (uint32_t)tot[2 * j];
      st.edges_reached = tot[2 * j + 1];
      st.device_ms = ms / k;
    }
  }
}

}  // namespace byo
