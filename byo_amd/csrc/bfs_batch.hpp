// bfs_batch.hpp -- up to 64 BFS sources traversed together in one persistent
// cooperative launch (the traversal body; bfs_batch.hip launches it alone,
// job.hip together with personalized PageRank).
//
// Each source's BFS is BFS.h:47-64 over edgeMap (gbbs/edge_map_data.h:
// 254-307), unchanged in what it produces: parents[j][v] is a BFS parent of
// v in the tree of source j, and every vertex gets the level it has in that
// source's own BFS.  What is shared is the frontier.  A vertex carries a
// mask with bit j set when it belongs to source j's frontier (and a second
// mask of the sources that have reached it), so one edgeMap pass over the
// union frontier advances every source by one level:
//
//   push  (edgeMapSparse, gbbs/edge_map_blocked.h:18-92): for a frontier
//         vertex u with mask f and an edge u->v, the bits of f that v has
//         not seen are claimed by one atomicOr on next[v]; each bit j that
//         this edge wins sets parents[j][v] = u, and the thread that makes
//         next[v] nonzero appends v to the next queue;
//   pull  (edgeMapDense, :47-175): a vertex with bits U still unreached
//         (restricted to sources whose frontier is not empty) scans its
//         neighbours in adjacency order and takes, for each bit of U, the
//         first neighbour whose frontier mask has it, stopping once all of
//         U is covered -- per bit the same parent the reference's
//         early-exit scan picks for that source;
//   forward (edgeMapDenseForward, :180-247): push driven by the frontier
//         masks instead of a queue.
//
// A bit of v's mask is set at the first level at which some frontier vertex
// of source j has an edge to v, so levels are exactly the per-source BFS
// levels.  Direction is decided per level from the union frontier (size,
// sum of out-degrees) by the same rule or learned selector as the
// single-source kernel.
//
// Data per vertex: seen and three rotating frontier masks (1, 4 or 8 bytes
// for <= 8, <= 32, <= 64 sources); queues of the union frontier rotate over
// three buffers so the frontier of level L-1 can be cleared through its
// queue while level L runs.  A push claims through the next mask alone, so
// seen lags one level behind: at level L the vertices reached are those of
// seen | frontier_L, and level L folds its frontier into seen (each frontier
// vertex by the one thread that takes it; a pull while it rewrites seen).
#pragma once
#include <algorithm>
#include <type_traits>

#include "coop.hpp"

namespace byo {
namespace batch {

struct BCtr {
  unsigned long long size;    // vertices in the union frontier (queue cursor)
  unsigned long long deg;     // sum of their out-degrees
  unsigned long long cursor;  // pack cursor of the level using this slot
  unsigned long long active;  // sources with a nonempty frontier
};

struct BLog {
  unsigned long long fsize, fdeg, active;
  uint32_t dir, pad;
  unsigned long long t0;  // wall clock at the start of the level
};

struct BatchArgs {
  const uint64_t *off;  // adjacency view (AdjView): CSR, or the PMA's slot starts
  const uint32_t *ids;
  const uint32_t *deg;  // true degrees when the lists hold empty slots
  const uint32_t *fn;   // first neighbour of every vertex (kNone without one), n entries
  uint64_t n, m;
  uint64_t nm;              // mask entries (n rounded up to 64)
  uint32_t *parents;        // parents of source j at parents + j * pstride
  uint64_t pstride;
  void *seen;
  void *fm;                 // three rotating frontier-mask arrays of nm entries
  uint64_t *ubm;            // three rotating union-frontier bitmaps of nm / 64 words
                            // (bit v: v is in some source's frontier)
  uint32_t *q;              // three rotating queues of qstride entries
  uint64_t qstride;
  uint64_t *scan;
  uint32_t *cstart;
  unsigned long long *bsum;  // per-block frontier degree sums (push)
  BCtr *ctr;                // 3 slots
  uint32_t srcs[64];         // the sources, passed in the kernel arguments
  uint32_t k;
  BLog *logs;               // 64
  uint32_t *lastlev;        // k: last level with a nonempty frontier
  uint32_t *nlev;           // levels of the whole traversal
  unsigned long long *tot;  // k x {vertices reached, sum of their degrees}
  DirCtl dc;
  Pieces pieces;            // weight-balanced word pieces of the grid
  unsigned int *dyn_ctr;    // 3 rotating sets of 8 piece counters (set at + 8192 s, counter q at + 1024 q)
  uint32_t dyn;             // pieces per wave dealt by the counters in dense levels (0: all static)
  GridBar *bar;
};

constexpr int kBlock = kCoopBlock;
constexpr int kWaves = kCoopWaves;
constexpr int kPerLane = 4;
constexpr uint64_t kChunk = 64 * kPerLane;
// loads per lane of one pull window (the pull's remaining edges are
// latency-bound per wave, so a window keeps this many in flight)
constexpr int kPullPerLane = 8;

// Per-vertex source masks.  One-byte masks are updated through the 32-bit
// word that holds them.
template <class M>
__device__ __forceinline__ M mask_or(M *a, uint64_t v, M bits);
template <>
__device__ __forceinline__ uint8_t mask_or<uint8_t>(uint8_t *a, uint64_t v, uint8_t bits) {
  uint32_t *w = reinterpret_cast<uint32_t *>(a + (v & ~3ull));
  const int sh = (int)(v & 3) * 8;
  return (uint8_t)(atomicOr(w, (uint32_t)bits << sh) >> sh);
}
template <>
__device__ __forceinline__ uint32_t mask_or<uint32_t>(uint32_t *a, uint64_t v, uint32_t bits) {
  return atomicOr(a + v, bits);
}
template <>
__device__ __forceinline__ uint64_t mask_or<uint64_t>(uint64_t *a, uint64_t v, uint64_t bits) {
  return atomicOr(reinterpret_cast<unsigned long long *>(a + v), (unsigned long long)bits);
}

// Masks travel through shuffles as 32- or 64-bit values.
template <class M>
using Wide = typename std::conditional<sizeof(M) == 8, unsigned long long, unsigned int>::type;

template <class M>
__device__ __forceinline__ M shfl_m(M v, int src) {
  return (M)__shfl((Wide<M>)v, src, 64);
}
template <class M>
__device__ __forceinline__ M shfl_up_m(M v, int d) {
  return (M)__shfl_up((Wide<M>)v, (unsigned)d, 64);
}
template <class M>
__device__ __forceinline__ M shfl_up8(M v, int d) {
  return (M)__shfl_up((Wide<M>)v, (unsigned)d, 8);
}
template <class M>
__device__ __forceinline__ M wave_or(M v) {
  Wide<M> x = (Wide<M>)v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x |= __shfl_xor(x, o, 64);
  return (M)x;
}

// Claims the bits of f that v has not reached (seen | this frontier, read
// plainly: neither changes during a push) with one atomicOr on v's next
// mask, which settles races between claims of the same level; writes the
// parent of every bit won and returns true for the one claim that makes v's
// next mask nonzero.
template <class M>
__device__ __forceinline__ bool relax(const BatchArgs &a, const M *seen, const M *fcur, M *nmask, uint32_t u, M f,
                                      uint32_t v, M &won) {
  const M nb = f & (M)~(seen[v] | fcur[v]);
  if (!nb) return false;
  const M old = mask_or<M>(nmask, v, nb);
  M g = nb & (M)~old;
  if (!g) return false;
  won |= g;
  while (g) {
    const int b = __builtin_ctzll((unsigned long long)g);
    g &= (M)(g - 1);
    a.parents[(uint64_t)b * a.pstride + v] = u;
  }
  return old == 0;
}

struct WinLds {
  uint64_t sc[kWaves][64];
  int64_t base[kWaves][64];
  uint64_t f[kWaves][64];
};

// The whole shared traversal, run by every thread of a cooperative grid.
// G: the adjacency lists may hold empty slots (kNone), skipped everywhere.
template <class M, bool G = false>
__device__ __forceinline__ void bfs_batch_body(const BatchArgs &a) {
  __shared__ unsigned long long lds[8];
  __shared__ MlpLds mlp;
  __shared__ WinLds win;
  __shared__ uint64_t res[kWaves][64];
  __shared__ unsigned long long tl[2][64];
  __shared__ unsigned long long lvl[3][kWaves];
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const uint64_t nb = gridDim.x;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const uint64_t nwords = a.nm >> 6;  // 64-vertex words
  // the rotating arrays are addressed by offset from one base (a runtime
  // index into an array of pointers would make every access a flat one)
  M *const seen = static_cast<M *>(a.seen);
  M *const fmb = static_cast<M *>(a.fm);
  uint32_t *const qb = a.q;

  // logs[64] holds wall-clock marks: start, seeded, traversal end, tallied
  if (leader) a.logs[64].fsize = wall_clock64();
  pieces_init(a.pieces, a.off, a.n, nwords, nwaves);
  mlp_stage(a.dc, mlp);
  // ---- seed: masks cleared, then every source's bit set.  (Parents are
  // not cleared: the tally pass writes kNone for every source that did not
  // reach a vertex, the traversal writes all the others.)
  {
    const uint64_t nvec = a.nm * sizeof(M) / 16;
    uint4 *s4 = reinterpret_cast<uint4 *>(seen);
    uint4 *f4 = reinterpret_cast<uint4 *>(fmb);
    const uint4 z = make_uint4(0, 0, 0, 0);
    for (uint64_t i = global_tid(); i < nvec; i += grid_threads()) {
      s4[i] = z;
      f4[i] = z;
      f4[nvec + i] = z;
      f4[2 * nvec + i] = z;
    }
    for (uint64_t i = global_tid(); i < 3 * nwords; i += grid_threads()) a.ubm[i] = 0;
    if (global_tid() < 24) a.dyn_ctr[(global_tid() >> 3) * 8192 + (global_tid() & 7) * 1024] = 0;
  }
  grid_barrier(a.bar, (unsigned)nb);
  if (blockIdx.x == 0) {
    if (threadIdx.x < a.k) {
      const uint32_t j = threadIdx.x, s = a.srcs[j];
      a.parents[(uint64_t)j * a.pstride + s] = s;
      mask_or<M>(seen, s, (M)((M)1 << j));
      mask_or<M>(fmb, s, (M)((M)1 << j));
      atomicOr(reinterpret_cast<unsigned long long *>(&a.ubm[s >> 6]), 1ull << (s & 63));
      a.lastlev[j] = 0;
      a.tot[2 * j] = 0;
      a.tot[2 * j + 1] = 0;
    }
    if (threadIdx.x < 64) {
      // the first frontier: lane j holds source j unless an earlier source
      // is the same vertex; the queue keeps first occurrences in order
      bool first = false;
      unsigned long long d = 0;
      uint32_t sj = 0;
      if ((uint32_t)lane < a.k) {
        sj = a.srcs[lane];
        first = true;
        for (int i = 0; i < lane; i++) first = first && a.srcs[i] != sj;
        if (first) d = vdeg<G>(a.off, a.deg, sj);
      }
      const uint64_t fm = __ballot(first);
      if (first) qb[mbcnt(fm)] = sj;
      const unsigned long long deg = wave_sum(d);
      if (lane == 0) {
        const unsigned long long all = a.k >= 64 ? ~0ull : ((1ull << a.k) - 1);
        a.ctr[0] = BCtr{(unsigned long long)__popcll(fm), deg, 0, all};
        a.ctr[1] = BCtr{0, 0, 0, 0};
        a.ctr[2] = BCtr{0, 0, 0, 0};
      }
    }
  }
  grid_barrier(a.bar, (unsigned)nb);
  if (leader) a.logs[64].fdeg = wall_clock64();

  bool is_dense = false, have_q = true, prev_have_q = true;
  bool ubm_ok = true;     // this frontier's union bitmap is written (seeding, or a pull before)
  const M *flast = nullptr;  // a frontier left unfolded into seen when the traversal stops
  uint64_t prev_fs = 0, visited_deg = 0;
  int L = 0;
  for (;; L++) {
    BCtr *cur = &a.ctr[L % 3];
    const uint64_t fs = ld_relaxed(&cur->size);
    const uint64_t fdeg = ld_relaxed(&cur->deg);
    const M act = (M)ld_relaxed(&cur->active);
    if (fs == 0) break;
    if (leader) {
      for (M b = act; b; b &= (M)(b - 1)) a.lastlev[__builtin_ctzll((unsigned long long)b)] = (uint32_t)L;
    }
    const int dir = decide(a.dc, is_dense, fs, fdeg, visited_deg, L, mlp);
    if (leader && L < 64)
      a.logs[L] = BLog{fs, fdeg, (unsigned long long)act, (uint32_t)dir, 0, (unsigned long long)wall_clock64()};
    if (dir == 2) {
      flast = fmb + (uint64_t)(L % 3) * a.nm;
      L++;
      break;
    }
    BCtr *nx = &a.ctr[(L + 1) % 3];
    if (leader) a.ctr[(L + 2) % 3] = BCtr{0, 0, 0, 0};
    // piece counters of this level's dense pass, and the set of level L+2
    // cleared (its last use, level L-1, has passed the barrier)
    unsigned int *const dctr = a.dyn_ctr + (uint64_t)(L % 3) * 8192;
    // the dynamic tail pays in the first dense level after a sparse one (its
    // lists are scanned whole, and their lengths are what the static weights
    // balance least well); later dense levels keep the static share-out,
    // where claims cost more than they balance
    const uint32_t ldyn = is_dense ? 0 : a.dyn;
    if (blockIdx.x == 0 && threadIdx.x < 8) a.dyn_ctr[(uint64_t)((L + 2) % 3) * 8192 + threadIdx.x * 1024] = 0;
    M *const fmask = fmb + (uint64_t)(L % 3) * a.nm, *const nmask = fmb + (uint64_t)((L + 1) % 3) * a.nm;
    uint32_t *const q = qb + (uint64_t)(L % 3) * a.qstride, *const qn = qb + (uint64_t)((L + 1) % 3) * a.qstride;
    // union bitmaps rotate with the masks: this frontier's, the next one's
    // (whole words written by a pull; after a push or forward level a bitmap
    // pull writes its frontier's words first), and last level's, cleared now
    // for level L+1
    uint64_t *const cubm = a.ubm + (uint64_t)(L % 3) * nwords;
    uint64_t *const nubm = a.ubm + (uint64_t)((L + 1) % 3) * nwords;
    {
      uint64_t *clr = a.ubm + (uint64_t)((L + 2) % 3) * nwords;
      for (uint64_t i = global_tid(); i < nwords; i += grid_threads()) clr[i] = 0;
    }
    {
      // last level's frontier masks, read by nobody now, are cleared for
      // level L+1 -- through that frontier's queue when it has one
      M *clr = fmb + (uint64_t)((L + 2) % 3) * a.nm;
      if (prev_have_q) {
        const uint32_t *pq = qb + (uint64_t)((L + 2) % 3) * a.qstride;
        for (uint64_t i = global_tid(); i < prev_fs; i += grid_threads()) clr[pq[i]] = 0;
      } else {
        const uint64_t nvec = a.nm * sizeof(M) / 16;
        uint4 *c4 = reinterpret_cast<uint4 *>(clr);
        for (uint64_t i = global_tid(); i < nvec; i += grid_threads()) c4[i] = make_uint4(0, 0, 0, 0);
      }
    }
    if (dir == 4 && !ubm_ok) {
      for (uint64_t w = wave; w < nwords; w += nwaves) {
        const uint64_t vl = w * 64 + lane;
        const uint64_t bits = __ballot(vl < a.n && fmask[vl] != 0);
        if (lane == 0) cubm[w] = bits;
      }
      grid_barrier(a.bar, (unsigned)nb);
    }
    unsigned long long dsum = 0, addsize = 0;
    M won = 0;

    if (dir == 0) {
      if (!have_q) {
        // toSparse: a block packs runs of kBlock 16-byte vectors of mask
        // entries (one load per thread), counts them in LDS and takes its
        // place in the queue with one cursor atomic per run; the queue holds
        // each run's vertices in id order
        constexpr int E = 16 / sizeof(M);
        const uint64_t nvec = a.nm * sizeof(M) / 16;
        const uint4 *const f4 = reinterpret_cast<const uint4 *>(fmask);
        for (uint64_t v0 = blockIdx.x * (uint64_t)kBlock; v0 < nvec; v0 += nb * kBlock) {
          const uint64_t i = v0 + threadIdx.x;
          uint4 x = make_uint4(0, 0, 0, 0);
          if (i < nvec) x = f4[i];
          M e[E];
          __builtin_memcpy(e, &x, 16);
          uint32_t c = 0;
#pragma unroll
          for (int t = 0; t < E; t++) c += e[t] != 0;
          const uint32_t incl = wave_incl_scan(c);
          if (lane == 63) lds[wv] = incl;
          __syncthreads();
          unsigned long long wpre = 0, tot = 0;
          for (int w = 0; w < kWaves; w++) {
            if (w < wv) wpre += lds[w];
            tot += lds[w];
          }
          if (threadIdx.x == 0 && tot) lds[kWaves] = atomicAdd(&cur->cursor, tot);
          __syncthreads();
          if (tot) {
            uint64_t pos = lds[kWaves] + wpre + incl - c;
#pragma unroll
            for (int t = 0; t < E; t++)
              if (e[t]) q[pos++] = (uint32_t)(i * E + t);
          }
          __syncthreads();  // lds is rewritten by the next run
        }
        grid_barrier(a.bar, (unsigned)nb);
      }
      if (fs <= 4 * nb || (fs <= 16 * nwaves && fdeg <= 32 * fs)) {
        // one block per frontier vertex (fs <= #blocks), or gl lanes per
        // frontier vertex, gl the power of two >= the frontier's average
        // degree: a frontier of short lists is then served in one round of
        // loads instead of a wave per vertex, one after another
        int gl = kBlock;
        if (fs > nb) {
          gl = 1;
          while (gl < 64 && (uint64_t)gl * fs < fdeg) gl <<= 1;
        }
        const int per = gl >= 64 ? 1 : 64 / gl;
        const int gi = gl >= 64 ? 0 : lane / gl;
        const int lid = gl == kBlock ? (int)threadIdx.x : (gl == 64 ? lane : lane % gl);
        const uint64_t first = gl == kBlock ? blockIdx.x : wave * per, step = gl == kBlock ? nb : nwaves * per;
        const int width = gl == kBlock ? kBlock : gl;
        for (uint64_t i0 = first; i0 < fs; i0 += step) {
          const uint64_t i = i0 + gi;
          uint32_t u = 0;
          M f = 0;
          uint64_t b = 0, e = 0;
          if (i < fs) {
            u = q[i];
            f = fmask[u];
            b = a.off[u];
            e = a.off[u + 1];
            if (lid == 0) seen[u] = seen[u] | f;  // fold u's frontier bits (u is this group's alone)
          }
          // a wave-uniform trip count (the ballots below need every lane)
          const uint64_t trips = wave_max<uint64_t>((e - b + width - 1) / width);
          for (uint64_t t = 0; t < trips; t++) {
            const uint64_t kk = b + t * width + lid;
            bool add = false;
            uint32_t tgt = 0;
            if (kk < e) {
              tgt = a.ids[kk];
              add = (!G || tgt != kNone) && relax<M>(a, seen, fmask, nmask, u, f, tgt, won);
              if (add) dsum += vdeg<G>(a.off, a.deg, tgt);
            }
            const uint64_t mask = __ballot(add);
            if (mask) {
              unsigned long long base = 0;
              if (lane == 0) base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
              base = __shfl(base, 0, 64);
              if (add) qn[base + mbcnt(mask)] = tgt;
            }
          }
        }
      } else {
        // (1) slice sums of frontier degrees
        const uint64_t S = (fs + nb - 1) / nb;
        const uint64_t s0 = std::min<uint64_t>(fs, blockIdx.x * S), s1 = std::min<uint64_t>(fs, s0 + S);
        unsigned long long part = 0;
        for (uint64_t i = s0 + threadIdx.x; i < s1; i += kBlock) {
          const uint32_t v = q[i];
          part += a.off[v + 1] - a.off[v];
          seen[v] = seen[v] | fmask[v];  // fold v's frontier bits (each queue entry once)
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
        // (3) edge-balanced push, 256 edges (slots) per wave step; the LDS
        // window holds (scan, adjacency base, frontier mask) of 64 queue entries
        const uint64_t E = G ? ld_relaxed(reinterpret_cast<const unsigned long long *>(&a.scan[fs])) : fdeg;
        const uint64_t nch = (E + kChunk - 1) / kChunk;
        for (uint64_t c = wave; c < nch; c += nwaves) {
          const uint64_t e0 = c * kChunk;
          const uint64_t i0 = a.cstart[c];
          {
            const uint64_t idx = i0 + lane;
            uint64_t sc = ~0ull, fm = 0;
            int64_t bs = 0;
            if (idx < fs) {
              const uint32_t u = q[idx];
              sc = a.scan[idx];
              bs = (int64_t)a.off[u] - (int64_t)sc;
              fm = (uint64_t)fmask[u];
            }
            win.sc[wv][lane] = sc;
            win.base[wv][lane] = bs;
            win.f[wv][lane] = fm;
          }
          wave_lds_sync();
          uint32_t tg[kPerLane];
          bool ad[kPerLane];
          uint32_t total = 0;
#pragma unroll
          for (int t = 0; t < kPerLane; t++) {
            const uint64_t e = e0 + (uint64_t)t * 64 + lane;
            bool add = false;
            uint32_t tgt = 0;
            if (e < E) {
              int lo = 0, hi = 64;  // sc[lo] <= e (sc[0] <= e0 by construction)
              while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (win.sc[wv][mid] <= e) lo = mid; else hi = mid;
              }
              uint32_t owner;
              int64_t bs;
              // read the window first: a select between an LDS and a global
              // address would turn the load into a flat one
              M f = (M)win.f[wv][lo];
              if (lo == 63 && i0 + 64 < fs && a.scan[i0 + 64] <= e) {
                // the chunk spans more than 64 frontier vertices
                uint64_t x = i0 + 64, y = fs;
                while (y - x > 1) {
                  const uint64_t mid = (x + y) >> 1;
                  if (a.scan[mid] <= e) x = mid; else y = mid;
                }
                owner = q[x];
                bs = (int64_t)a.off[owner] - (int64_t)a.scan[x];
                f = fmask[owner];
              } else {
                owner = q[i0 + lo];
                bs = win.base[wv][lo];
              }
              tgt = a.ids[(uint64_t)(bs + (int64_t)e)];
              add = (!G || tgt != kNone) && relax<M>(a, seen, fmask, nmask, owner, f, tgt, won);
              if (add) dsum += vdeg<G>(a.off, a.deg, tgt);
            }
            tg[t] = tgt;
            ad[t] = add;
            total += __popcll(__ballot(add));
          }
          if (total) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(&nx->size, (unsigned long long)total);
            base = __shfl(base, 0, 64);
#pragma unroll
            for (int t = 0; t < kPerLane; t++) {
              const uint64_t mask = __ballot(ad[t]);
              if (ad[t]) qn[base + mbcnt(mask)] = tg[t];
              base += __popcll(mask);
            }
          }
          wave_lds_sync();  // the window is rewritten by the next chunk
        }
      }
      is_dense = false;
      prev_have_q = true;  // this frontier has (or was packed into) a queue
      have_q = true;
    } else if (dir == 1 || dir == 4) {
      // dense pull: a wave owns 64 consecutive vertices per step (lane =
      // vertex for offsets and masks).  Groups of 8 lanes scan 8 neighbours
      // of one vertex at a time; within a group the exclusive prefix-OR of
      // the hit masks gives each bit's first neighbour in adjacency order.
      // Direction 4 (bitmap pull) tests each neighbour against the union
      // bitmap of the frontier first (n/8 bytes, small enough to stay in
      // L2) and gathers the neighbour's source mask only on a hit;
      // direction 1 gathers every neighbour's mask.
      const bool bmp = dir == 4;
      // after a push or forward level seen lags by this frontier (claims go
      // through the next masks alone); after a pull it is whole
      const bool lag = !is_dense || !ubm_ok;
      // the one-neighbour round runs in dense levels after the first with up
      // to 8 sources: there most vertices finish at their first neighbour
      // (with more sources the many open bits keep them scanning, and the
      // extra round trip costs more than the gathers it saves)
      const bool first1 = is_dense && sizeof(M) == 1;
      const int g = lane >> 3, sub = lane & 7;
      unsigned long long cnt = 0;
      for_my_words_dyn(a.pieces, wave, nwaves, dctr, ldyn, [&](uint64_t w) {
        const uint64_t vl = w * 64 + lane;
        uint64_t bl = 0, el = 0;
        M s = (M)~(M)0;
        uint32_t f1 = kNone;
        if (vl < a.n) {
          bl = a.off[vl];
          el = a.off[vl + 1];
          s = seen[vl];
          if (lag) s |= fmask[vl];  // this frontier's bits, not folded yet, are reached too
          if (first1) f1 = a.fn[vl];  // loaded with the word (coalesced), not after its offsets
        }
        const M U = (M)(~s & act);
        const uint64_t todo = __ballot(U != 0);
        M acc = 0;
        if (todo) {
          // first neighbour of every unfinished vertex (lane = vertex): in
          // the dense levels after the first, most vertices finish here, and
          // the 8-wide probe below then gathers only for the others
          M h1 = 0;
          if (first1) {
            const bool p1 = ((todo >> lane) & 1) && f1 != kNone;
            const uint32_t u1 = p1 ? f1 : 0;
            h1 = p1 ? U : (M)0;
            if (bmp && !((cubm[u1 >> 6] >> (u1 & 63)) & 1)) h1 = 0;
            h1 &= (!bmp || h1) ? fmask[u1] : (M)0;
            M nw = h1;
            while (nw) {
              const int b = __builtin_ctzll((unsigned long long)nw);
              nw &= (M)(nw - 1);
              a.parents[(uint64_t)b * a.pstride + vl] = u1;
            }
          }
          const M U2 = (M)(U & (M)~h1);
          const uint64_t todo2 = __ballot(((todo >> lane) & 1) && U2 != 0 && el - bl > (first1 ? 1 : 0));
          res[wv][lane] = (uint64_t)h1;
          wave_lds_sync();
          if (todo2) {
            uint32_t u[8];
            M h[8];
#pragma unroll
            for (int st = 0; st < 8; st++) {
              const int x = st * 8 + g;
              const uint64_t b = __shfl(bl, x, 64) + (first1 ? 1 : 0), e = __shfl(el, x, 64);
              const uint64_t kk = b + sub;
              const bool on = ((todo2 >> x) & 1) && kk < e;
              u[st] = a.ids[on ? kk : 0];  // unconditional: all 8 loads in flight
              const M Ux = shfl_m<M>(U2, x);  // every lane joins the shuffle
              h[st] = on ? Ux : (M)0;
            }
            if constexpr (G) {
              // an empty slot (or an empty slot 0 read by an idle lane) is no
              // neighbour; read entry 0's mask instead of an out-of-range one
#pragma unroll
              for (int st = 0; st < 8; st++)
                if (u[st] == kNone) {
                  u[st] = 0;
                  h[st] = 0;
                }
            }
            M fv[8];
            if (bmp) {
              uint64_t bw[8];
#pragma unroll
              for (int st = 0; st < 8; st++) bw[st] = cubm[u[st] >> 6];
#pragma unroll
              for (int st = 0; st < 8; st++)
                if (!((bw[st] >> (u[st] & 63)) & 1)) h[st] = 0;
#pragma unroll
              for (int st = 0; st < 8; st++) fv[st] = h[st] ? fmask[u[st]] : (M)0;
            } else {
#pragma unroll
              for (int st = 0; st < 8; st++) fv[st] = fmask[u[st]];
            }
#pragma unroll
            for (int st = 0; st < 8; st++) h[st] &= fv[st];
#pragma unroll
            for (int st = 0; st < 8; st++) {
              M incl = h[st];
#pragma unroll
              for (int d = 1; d < 8; d <<= 1) {
                const M t = shfl_up8<M>(incl, d);
                if (sub >= d) incl |= t;
              }
              M excl = shfl_up8<M>(incl, 1);
              if (sub == 0) excl = 0;
              M nw = h[st] & (M)~excl;
              const uint64_t v = w * 64 + st * 8 + g;
              while (nw) {
                const int b = __builtin_ctzll((unsigned long long)nw);
                nw &= (M)(nw - 1);
                a.parents[(uint64_t)b * a.pstride + v] = u[st];
              }
              if (sub == 7) res[wv][st * 8 + g] |= (uint64_t)incl;
            }
          }
          wave_lds_sync();
          acc = (M)res[wv][lane];
          // the remaining edges of vertices with bits still uncovered are
          // scanned wave-wide: the word's unfinished vertices lay their
          // edges end to end (exclusive prefix of remaining degrees in LDS,
          // owner found by a 6-step search) and every window of 512 edges
          // issues its 8 loads per lane before using any of them.  Each
          // window starts at the first vertex that still has bits to find;
          // bits found are merged into the vertex's LDS mask by atomicOr, and
          // any lane holding a frontier neighbour may write that bit's parent
          const uint64_t pre = first1 ? 9 : 8;  // neighbours checked by the two probes
          const bool more = ((todo >> lane) & 1) && acc != U && el - bl > pre;
          const uint64_t r = more ? el - bl - pre : 0;
          const uint64_t rin = wave_incl_scan(r);
          const uint64_t R = __shfl(rin, 63, 64);
          if (R) {
            const uint64_t scl = rin - r;
            win.sc[wv][lane] = scl;
            win.base[wv][lane] = (int64_t)(bl + pre) - (int64_t)scl;
            win.f[wv][lane] = (uint64_t)U;
            wave_lds_sync();
            uint64_t t0 = 0;
            for (;;) {
              const uint64_t inc = __ballot(r > 0 && (M)res[wv][lane] != U && scl + r > t0);
              if (!inc) break;
              t0 = std::max<uint64_t>(t0, __shfl(scl, __builtin_ctzll(inc), 64));
              // loads are unconditional (idle lanes read edge 0 / vertex
              // ids[0]) so that all of a window's loads are in flight at once
              // the 8 owner searches advance in lockstep so each step's LDS
              // reads are issued together
              uint32_t uu[kPullPerLane];
              int xx[kPullPerLane];
              M hm[kPullPerLane];
#pragma unroll
              for (int t = 0; t < kPullPerLane; t++) xx[t] = 0;
#pragma unroll
              for (int st = 32; st > 0; st >>= 1) {
                uint64_t sv[kPullPerLane];
#pragma unroll
                for (int t = 0; t < kPullPerLane; t++) sv[t] = win.sc[wv][xx[t] + st];
#pragma unroll
                for (int t = 0; t < kPullPerLane; t++)
                  xx[t] += (sv[t] <= t0 + (uint64_t)t * 64 + lane) ? st : 0;
              }
#pragma unroll
              for (int t = 0; t < kPullPerLane; t++) {
                const uint64_t idx = t0 + (uint64_t)t * 64 + lane;
                const int x = xx[t];
                const M Ux = (M)win.f[wv][x], ax = (M)res[wv][x];
                const int64_t bs = win.base[wv][x];
                const bool ok = idx < R && ax != Ux;
                hm[t] = ok ? (M)(Ux & (M)~ax) : (M)0;
                uu[t] = a.ids[ok ? (uint64_t)(bs + (int64_t)idx) : 0];
                if (G && uu[t] == kNone) {
                  uu[t] = 0;
                  hm[t] = 0;
                }
              }
              M fv[kPullPerLane];
              if (bmp) {
                uint64_t bw[kPullPerLane];
#pragma unroll
                for (int t = 0; t < kPullPerLane; t++) bw[t] = cubm[uu[t] >> 6];
#pragma unroll
                for (int t = 0; t < kPullPerLane; t++)
                  if (!((bw[t] >> (uu[t] & 63)) & 1)) hm[t] = 0;
#pragma unroll
                for (int t = 0; t < kPullPerLane; t++) fv[t] = hm[t] ? fmask[uu[t]] : (M)0;
              } else {
#pragma unroll
                for (int t = 0; t < kPullPerLane; t++) fv[t] = fmask[uu[t]];
              }
#pragma unroll
              for (int t = 0; t < kPullPerLane; t++) hm[t] &= fv[t];
#pragma unroll
              for (int t = 0; t < kPullPerLane; t++) {
                if (hm[t]) {
                  atomicOr(reinterpret_cast<unsigned long long *>(&res[wv][xx[t]]), (unsigned long long)hm[t]);
                  M nw = hm[t];
                  while (nw) {
                    const int bb = __builtin_ctzll((unsigned long long)nw);
                    nw &= (M)(nw - 1);
                    a.parents[(uint64_t)bb * a.pstride + w * 64 + xx[t]] = uu[t];
                  }
                }
              }
              wave_lds_sync();
              t0 += 64 * kPullPerLane;
            }
          }
          wave_lds_sync();
          acc = (M)res[wv][lane];
          wave_lds_sync();  // res is rewritten by the next word
        }
        if (vl < a.n) {
          nmask[vl] = acc;
          if (acc || lag) seen[vl] = s | acc;
        }
        const uint64_t nword = __ballot(acc != 0);
        if (lane == 0) nubm[w] = nword;
        cnt += __popcll(nword);
        if (acc) dsum += vdeg<G>(a.off, a.deg, vl);
        won |= acc;
      });
      addsize = cnt;  // wave-uniform (a sum of ballot counts)
      is_dense = true;
      prev_have_q = have_q;
      have_q = false;
    } else {
      // dense forward: one wave per frontier vertex of each 64-vertex word
      for_my_words_dyn(a.pieces, wave, nwaves, dctr, ldyn, [&](uint64_t w) {
        const uint64_t vl = w * 64 + lane;
        const M fv = vl < a.n ? fmask[vl] : (M)0;
        if (fv) seen[vl] = seen[vl] | fv;  // fold (lane = vertex)
        uint64_t word = __ballot(fv != 0);
        while (word) {
          const int bp = __builtin_ctzll(word);
          word &= word - 1;
          const uint32_t u = (uint32_t)(w * 64 + bp);
          const M f = shfl_m<M>(fv, bp);
          const uint64_t b = a.off[u], e = a.off[u + 1];
          for (uint64_t k0 = b; k0 < e; k0 += 64) {
            const uint64_t kk = k0 + lane;
            bool add = false;
            uint32_t t = 0;
            if (kk < e) {
              t = a.ids[kk];
              add = (!G || t != kNone) && relax<M>(a, seen, fmask, nmask, u, f, t, won);
              if (add) dsum += vdeg<G>(a.off, a.deg, t);
            }
            const uint64_t mask = __ballot(add);
            if (mask) {
              unsigned long long base = 0;
              if (lane == 0) base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
              base = __shfl(base, 0, 64);
              if (add) qn[base + mbcnt(mask)] = t;
            }
          }
        }
      });
      is_dense = true;
      prev_have_q = have_q;
      have_q = true;
    }
    dsum = wave_sum(dsum);
    won = wave_or<M>(won);
    // the level's counters, summed per workgroup first: one atomic per block
    // and counter instead of one per wave (thousands of same-address atomics
    // at the end of a level queue in one L2 channel)
    if (lane == 0) {
      lvl[0][wv] = addsize;
      lvl[1][wv] = dsum;
      lvl[2][wv] = (unsigned long long)won;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      unsigned long long sz = 0, dg = 0, ac = 0;
      for (int w = 0; w < kWaves; w++) {
        sz += lvl[0][w];
        dg += lvl[1][w];
        ac |= lvl[2][w];
      }
      if (sz) atomicAdd(&nx->size, sz);
      if (dg) atomicAdd(&nx->deg, dg);
      if (ac) atomicOr(&nx->active, ac);
    }
    grid_barrier(a.bar, (unsigned)nb);
    prev_fs = fs;
    visited_deg += fdeg;
    ubm_ok = dir == 1 || dir == 4;
  }
  if (leader) {
    const unsigned long long t = wall_clock64();
    if (L < 64) a.logs[L].t0 = t;
    a.logs[64].active = t;
  }

  // ---- per-source tallies: vertices reached and the sum of their degrees,
  // and kNone as the parent of every (source, vertex) pair not reached.
  // Up to 8 sources: every lane counts its own vertices for all sources in
  // registers, and the wave sums them once at the end.  More: lane j counts
  // source j over each word's 64 masks staged in LDS.  Blocks combine lanes
  // in LDS and add their totals with one atomic per source.
  if constexpr (sizeof(M) == 1) {
    // a thread takes 16 consecutive vertices at a time (one 16-byte load of
    // their masks and one run of their offsets), so the pass is a couple of
    // load rounds per thread rather than one per 64-vertex word
    unsigned long long rc[8] = {}, rd[8] = {};
    const uint64_t ngrp = (a.n + 15) / 16;
    for (uint64_t t = global_tid(); t < ngrp; t += grid_threads()) {
      const uint64_t v0 = t * 16;
      uint4 sv = reinterpret_cast<const uint4 *>(seen)[t];  // nm is a multiple of 64
      if (flast) {
        const uint4 fl = reinterpret_cast<const uint4 *>(flast)[t];
        sv = make_uint4(sv.x | fl.x, sv.y | fl.y, sv.z | fl.z, sv.w | fl.w);
      }
      uint8_t sm[16];
      __builtin_memcpy(sm, &sv, 16);
      uint64_t d[16];
      if constexpr (G) {
#pragma unroll
        for (int i = 0; i < 16; i++) d[i] = v0 + i < a.n ? a.deg[v0 + i] : 0;
      } else {
        uint64_t o[17];
#pragma unroll
        for (int i = 0; i <= 16; i++) o[i] = a.off[std::min<uint64_t>(v0 + i, a.n)];
#pragma unroll
        for (int i = 0; i < 16; i++) d[i] = o[i + 1] - o[i];
      }
#pragma unroll
      for (int i = 0; i < 16; i++) {
        if (v0 + i >= a.n) break;
        const uint32_t m8 = sm[i];
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const unsigned long long bit = (m8 >> j) & 1;
          rc[j] += bit;
          rd[j] += bit * d[i];
        }
        const uint32_t miss = ~m8 & ((a.k >= 8) ? 0xffu : ((1u << a.k) - 1));
        for (uint32_t mm = miss; mm; mm &= mm - 1)
          a.parents[(uint64_t)__builtin_ctz(mm) * a.pstride + v0 + i] = kNone;
      }
    }
    if (threadIdx.x < 64) {
      tl[0][threadIdx.x] = 0;
      tl[1][threadIdx.x] = 0;
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const unsigned long long c = wave_sum(rc[j]), d = wave_sum(rd[j]);
      if (lane == 0 && c) {
        atomicAdd(&tl[0][j], c);
        atomicAdd(&tl[1][j], d);
      }
    }
    __syncthreads();
    if (threadIdx.x < a.k && tl[0][threadIdx.x]) {
      atomicAdd(&a.tot[2 * threadIdx.x], tl[0][threadIdx.x]);
      atomicAdd(&a.tot[2 * threadIdx.x + 1], tl[1][threadIdx.x]);
    }
  } else {
    unsigned long long rc = 0, rd = 0;
    for_my_words(a.pieces, wave, nwaves, [&](uint64_t w) {
      const uint64_t vl = w * 64 + lane;
      uint64_t sm = 0, d = 0;
      if (vl < a.n) {
        sm = (uint64_t)(flast ? (M)(seen[vl] | flast[vl]) : seen[vl]);
        d = vdeg<G>(a.off, a.deg, vl);
        const uint64_t kmask = a.k >= 64 ? ~0ull : ((1ull << a.k) - 1);
        for (uint64_t mm = ~sm & kmask; mm; mm &= mm - 1)
          a.parents[(uint64_t)__builtin_ctzll(mm) * a.pstride + vl] = kNone;
      }
      if (!__ballot(sm != 0)) return;
      res[wv][lane] = sm;
      win.sc[wv][lane] = d;
      wave_lds_sync();
      if ((uint32_t)lane < a.k) {
#pragma unroll 16
        for (int x = 0; x < 64; x++) {
          const unsigned long long bit = (res[wv][x] >> lane) & 1;
          rc += bit;
          rd += bit * win.sc[wv][x];
        }
      }
      wave_lds_sync();
    });
    if (threadIdx.x < 64) {
      tl[0][threadIdx.x] = 0;
      tl[1][threadIdx.x] = 0;
    }
    __syncthreads();
    if ((uint32_t)lane < a.k && rc) {
      atomicAdd(&tl[0][lane], rc);
      atomicAdd(&tl[1][lane], rd);
    }
    __syncthreads();
    if (threadIdx.x < a.k && tl[0][threadIdx.x]) {
      atomicAdd(&a.tot[2 * threadIdx.x], tl[0][threadIdx.x]);
      atomicAdd(&a.tot[2 * threadIdx.x + 1], tl[1][threadIdx.x]);
    }
  }
  if (leader) {
    *a.nlev = (uint32_t)L;
    a.logs[64].t0 = wall_clock64();
  }
}

// Host side (bfs_batch.hip): buffers and arguments of one launch for k
// sources (sources staged, barrier cleared, on g's stream), and the read-back
// of its per-source statistics.
// The traversal reads adjacency view v (G = v.deg != null).
void prepare(byo_graph *g, const AdjView &v, const uint32_t *srcs, uint32_t k, uint64_t maxdeg, BatchArgs &a,
             int mask_bytes, int nblocks);
void finish(byo_graph *g, const BatchArgs &a, byo_bfs_stats *stats, double ms);

}  // namespace batch
}  // namespace byo
