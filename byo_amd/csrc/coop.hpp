// coop.hpp -- pieces shared by the persistent cooperative traversal kernels
// (bfs_coop.hip: one source at a time; bfs_batch.hip: up to 64 sources per
// traversal): the grid barrier, block reductions and the on-device direction
// decision (rule of gbbs/edge_map_data.h:257-306 or the learned selector).
#pragma once
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace byo {

constexpr int kCoopBlock = 256;
constexpr int kCoopWaves = kCoopBlock / 64;

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide barrier for a cooperative launch (all blocks co-resident).
// Arrivals are counted per group of 32 blocks, then once per group on the
// top counter, so no address sees more than 32 + nblocks/32 atomics; the
// last arrival bumps the generation word the others poll.  Agent-scope
// release/acquire fences make every block's writes visible across XCDs.
struct GridBar {
  unsigned int top;
  unsigned int pad0[31];
  unsigned int gen;  // polled word on a line of its own
  unsigned int pad1[31];
  unsigned int group[1024];  // counter of 32-block group g at group[16 g]
};

__device__ __forceinline__ void grid_barrier(GridBar *b, unsigned int nblocks) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const unsigned int gen = __hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    const unsigned int grp = blockIdx.x >> 5;
    const unsigned int gsize = min(32u, nblocks - (grp << 5));
    const unsigned int ngroups = (nblocks + 31) >> 5;
    bool last = false;
    const unsigned int a = __hip_atomic_fetch_add(&b->group[grp * 16], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (a == gsize - 1) {
      // the group's last arrival carries the group's writes forward
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __hip_atomic_store(&b->group[grp * 16], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      const unsigned int t = __hip_atomic_fetch_add(&b->top, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
      if (t == ngroups - 1) {
        __hip_atomic_store(&b->top, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_fetch_add(&b->gen, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        last = true;
      }
    }
    // poll relaxed: an acquire per poll would invalidate this XCD's L2 and
    // evict the data the still-working blocks are using
    if (!last)
      while (__hip_atomic_load(&b->gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == gen) __builtin_amdgcn_s_sleep(2);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
}

// Degree of v over an adjacency view (G: the lists may hold empty slots, so
// the true degree comes from deg[]).
template <bool G>
__device__ __forceinline__ uint64_t vdeg(const uint64_t *off, const uint32_t *deg, uint64_t v) {
  if constexpr (G) return deg[v];
  else return off[v + 1] - off[v];
}

// Block-wide sum (all threads get the result); lds holds kCoopWaves slots.
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long *lds) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long s = 0;
  for (int w = 0; w < kCoopWaves; w++) s += lds[w];
  return s;
}

// Share-out of the vertex-driven passes (pull, forward, tallies) over
// 64-vertex words.  The words are cut into kPieces pieces per wave of equal
// weight (edges + 8 per vertex, read off the offsets), and wave w takes
// pieces w, w + nwaves, ..., one from each 1/kPieces of the id space.  Equal
// weights balance the passes that scan the whole graph; drawing each wave's
// pieces from every region of the ids averages out the regional skew of the
// later levels, whose work sits in whatever vertices are still unvisited.
// (Dealing single words out round-robin lines RMAT's heavy low-id words up
// on the same few waves; claiming words through atomics costs a round trip
// per word.)
// Pieces per wave is a tuning knob ("pieces", 8 by default, at most
// kMaxPieces).
constexpr uint32_t kMaxPieces = 64;
struct Pieces {
  uint64_t *bound;  // per_wave * nwaves + 1 word bounds
  uint32_t per_wave;
  uint32_t ready;   // bounds already hold this graph's pieces for this grid
};

// Piece bounds, computed once per launch by the whole grid before its first
// barrier: bound[i] = first word w with W(w) >= i * W(nwords) / P, where
// W(w) = off[v] + 8 v, v = min(64 w, n), and P = per_wave * nwaves.
__device__ inline void pieces_init(const Pieces &pc, const uint64_t *off, uint64_t n, uint64_t nwords,
                                   uint64_t nwaves) {
  if (pc.ready) return;
  uint64_t *bound = pc.bound;
  const uint64_t P = (uint64_t)pc.per_wave * nwaves;
  const uint64_t total = off[n] + 8 * n;
  for (uint64_t i = global_tid(); i <= P; i += grid_threads()) {
    const uint64_t target = (uint64_t)(((unsigned __int128)total * i) / P);
    uint64_t lo = 0, hi = nwords;  // first w in [lo, hi] with W(w) >= target
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      const uint64_t v = std::min<uint64_t>(mid * 64, n);
      if (off[v] + 8 * v >= target) hi = mid; else lo = mid + 1;
    }
    bound[i] = (i == P) ? nwords : lo;
  }
}

// Runs body(w) for every word of this wave's pieces; all 64 lanes call it
// together and body runs wave-uniformly.
template <class F>
__device__ __forceinline__ void for_my_words(const Pieces &pc, uint64_t wave, uint64_t nwaves, F &&body) {
  for (uint64_t i = 0; i < pc.per_wave; i++) {
    const uint64_t p = wave + i * nwaves;
    const uint64_t w1 = pc.bound[p + 1];
    for (uint64_t w = pc.bound[p]; w < w1; w++) body(w);
  }
}

// The same share-out with a dynamic tail: of a wave's pieces, the first
// per_wave - dyn are its own as above, and the last dyn * nwaves pieces of the
// grid are handed out by 8 counters (counter q deals pieces q, q + 8, ... of
// the tail; the counters sit 4 KB apart, ctr[q * 1024], so their atomics
// queue at different memory channels).  A wave claims from the counter of its
// block's XCD (workgroups are dealt to the 8 XCDs round-robin) and moves to
// the first other one that still has pieces when that runs dry (one read of
// all 8 counters finds it, so a dry counter costs no atomic).  The claim of the next piece is in flight while the
// wave runs the current one.  The static share balances the edge weights; the
// tail absorbs what they miss -- the per-word cost of a sweep that finds
// little to do, or lists that end early.  The counters must be zero on entry.
template <class F>
__device__ __forceinline__ void for_my_words_dyn(const Pieces &pc, uint64_t wave, uint64_t nwaves,
                                                 unsigned int *ctr, uint32_t dyn, F &&body) {
  const uint32_t own = dyn < pc.per_wave ? pc.per_wave - dyn : 0;
  const uint64_t base = (uint64_t)own * nwaves, ntail = (uint64_t)pc.per_wave * nwaves - base;
  const int lane = threadIdx.x & 63;
  const unsigned int home = blockIdx.x & 7;
  // one call site of body (a second inlined copy of a large body costs
  // registers): the static pieces first, then the claimed ones
  uint64_t i = 0, cnt = 0;
  unsigned int q = 8, tn = 0;  // q: counter being drained (8: none); tn: lane 0's claim in flight
  for (;;) {
    uint64_t p;
    if (i < own) {
      p = wave + i * nwaves;
      i++;
    } else {
      if (!ntail) break;
      unsigned int t = q < 8 ? __shfl(tn, 0, 64) : 0;
      bool done = false;
      while (q == 8 || t >= cnt) {
        // lanes 0-7 read the 8 counters (from the home XCD's on) in one
        // round trip; the wave drains the first that still has pieces
        const unsigned int ql = (home + (unsigned int)lane) & 7;
        const uint64_t cl = ntail > ql ? (ntail - ql + 7) / 8 : 0;
        const bool avail =
            lane < 8 && __hip_atomic_load(ctr + ql * 1024, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < cl;
        const uint64_t am = __ballot(avail);
        if (!am) {
          done = true;
          break;
        }
        const int src = __builtin_ctzll(am);
        q = (home + (unsigned int)src) & 7;
        cnt = __shfl(cl, src, 64);
        unsigned int c0 = 0;
        if (lane == 0) c0 = atomicAdd(ctr + q * 1024, 1u);
        t = __shfl(c0, 0, 64);
      }
      if (done) break;
      if (lane == 0) tn = atomicAdd(ctr + q * 1024, 1u);  // the next claim, in flight meanwhile
      p = base + q + 8ull * t;
    }
    const uint64_t w1 = pc.bound[p + 1];
    for (uint64_t w = pc.bound[p]; w < w1; w++) body(w);
  }
}

// Inputs of the per-level direction decision.
struct DirCtl {
  uint64_t n, m;
  int policy;
  uint64_t thr;
  const float *sel;      // 1379 params + mean[6] + inv_std[6], or null
  uint64_t maxdeg;
  const uint8_t *sched;  // 64 per-level directions (BYO_DIR_SCHEDULE)
  float margin;          // logit margin the network must beat the rule by
  float range;           // in-range bound on the normalised features
  int third;             // direction of selector class 2 (3 forward, 4 bitmap pull)
  bool densify;          // the network may turn a level the rule keeps sparse dense
};

struct MlpLds {
  float w[BYO_MLP_NPARAM + 12];  // the selector's parameters and normalisation, staged once per launch
  float z[2][3];  // the logits of the last two decisions (by level parity)
  bool in_range[2];
};

// Stages the selector's parameters in the block's LDS (every thread calls
// it once, before the first decide); a no-op unless the learned policy runs.
// The 32x32 hidden-layer matrix is stored transposed in LDS (entry (j, i) at
// 224 + i*32 + j), so the 32 lanes computing units j read consecutive words
// for each input i (no bank conflicts); the other parameters keep their
// layout.
__device__ inline void mlp_stage(const DirCtl &a, MlpLds &s) {
  if (a.policy != BYO_DIR_MLP || !a.sel) return;
  for (int i = threadIdx.x; i < BYO_MLP_NPARAM + 12; i += blockDim.x) {
    int dst = i;
    if (i >= 224 && i < 224 + 1024) {
      const int r = i - 224;  // (j, i') = (r / 32, r % 32)
      dst = 224 + (r & 31) * 32 + (r >> 5);
    }
    s.w[dst] = a.sel[i];
  }
  __syncthreads();
}

// Direction of one edgeMap call: 0 sparse, 1 dense pull, 2 empty, 3 dense
// forward, 4 bitmap pull (the same decision as choose_direction in bfs.hip;
// the rule picks 0 or 1, the network may pick the engine's third direction).  Called by
// every thread of the block (the selector MLP runs one neuron per thread,
// parameter layout as in selector.hip); returns the same value in all
// threads.
__device__ inline int decide(const DirCtl &a, bool is_dense, uint64_t fs, uint64_t fdeg, uint64_t visited_deg,
                             int level, MlpLds &s) {
  constexpr int kNP = BYO_MLP_NPARAM;
  int rule;
  if (is_dense && fs > a.n / 10) {
    rule = 1;
  } else if (fdeg == 0) {
    return 2;
  } else {
    rule = (fs + fdeg > a.thr) ? 1 : 0;
  }
  if (a.policy == BYO_DIR_SPARSE) return 0;
  if (a.policy == BYO_DIR_DENSE) return 1;
  if (a.policy == BYO_DIR_SCHEDULE) return (level < 64 && a.sched[level] != 255) ? a.sched[level] : rule;
  if (a.policy != BYO_DIR_MLP || !a.sel) return rule;
  // a level the rule keeps sparse stays sparse unless densify is set (a
  // wrong dense choice scans every unvisited edge, a wrong sparse one only
  // the frontier's): then there is nothing for the network to decide
  if (rule == 0 && !a.densify) return rule;
  // the network decides where the rule turns dense (the level after sparse
  // ones: pull now, later, or from the bitmap); once the traversal is dense
  // the rule's pull stands, and the network is not evaluated again
  if (is_dense && rule == 1 && !a.densify) return rule;
  const float *p = s.w;  // staged by mlp_stage
  // one wave evaluates the network (lanes 0-31 hold the hidden units, the
  // second half repeats them), reading the hidden values with v_readlane
  // (scalar, no LDS round trip); the logits go to the LDS buffer of this
  // level's parity, so no barrier is needed before writing them (the
  // previous decide's readers are separated from this one by the grid
  // barrier ending that level), and one barrier publishes them
  const int par = level & 1;
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x, j = lane & 31;
    const double avg = a.n ? (double)a.m / (double)a.n : 0.0;
    const float xr[6] = {log2f(1.0f + (float)fs), log2f(1.0f + (float)fdeg),
                         log2f(1.0f + (float)(a.m > visited_deg ? a.m - visited_deg : 0)), (float)level,
                         log2f(1.0f + (float)avg), log2f(1.0f + (float)a.maxdeg)};
    float x[6];
    bool in_range = true;
#pragma unroll
    for (int i = 0; i < 6; i++) {
      const float v = (xr[i] - p[kNP + i]) * p[kNP + 6 + i];
      in_range = in_range && fabsf(v) <= a.range;
      x[i] = fminf(fmaxf(v, -kSelClamp), kSelClamp);
    }
    float h1 = p[192 + j];
#pragma unroll
    for (int i = 0; i < 6; i++) h1 += p[j * 6 + i] * x[i];
    h1 = h1 > 0.f ? h1 : 0.f;
    float h2 = p[1248 + j];
#pragma unroll
    for (int i = 0; i < 32; i++)
      h2 += p[224 + i * 32 + j] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h1), i));
    h2 = h2 > 0.f ? h2 : 0.f;
    // every lane sums the 32 hidden units of each output in unit order
    float z0 = p[1376], z1 = p[1377], z2 = p[1378];
#pragma unroll
    for (int i = 0; i < 32; i++) {
      const float hi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(h2), i));
      z0 += p[1280 + i] * hi;
      z1 += p[1312 + i] * hi;
      z2 += p[1344 + i] * hi;
    }
    if (lane == 0) {
      s.z[par][0] = z0;
      s.z[par][1] = z1;
      s.z[par][2] = z2;
      s.in_range[par] = in_range;
    }
  }
  __syncthreads();
  const float z0 = s.z[par][0], z1 = s.z[par][1], z2 = s.z[par][2];
  if (!isfinite(z0 + z1 + z2)) return rule;
  const int best = (z1 > z0) ? ((z2 > z1) ? 2 : 1) : ((z2 > z0) ? 2 : 0);
  // the rule stands unless the frontier is in the training range and the
  // network prefers another direction by at least the margin
  if (!s.in_range[par] || s.z[par][best] - s.z[par][dir_class(rule)] < a.margin) return rule;
  return best == 0 ? 0 : (best == 1 ? 1 : a.third);
}

// The grid barrier of the handle's persistent launches ("bfs_coop_bar"),
// cleared once after it is allocated: a barrier's last arrivals reset its
// counters and only the generation word moves on, so every launch that
// returns leaves it ready for the next (no clear queued per launch).
GridBar *coop_bar(byo_graph *g);

// Direction-decision inputs of a handle (policy, threshold, and the
// selector weights of the single-source or the batched traversal).
DirCtl dir_ctl(byo_graph *g, uint64_t maxdeg, hipStream_t s, bool batched);

// Grid size of a cooperative kernel: CUs x min(resident blocks per CU, cap)
// (cap from the handle's tuning, at most 8; at most 2048 blocks).
int coop_grid(const void *kernel, int block, int cap);
// The traversal grid for an adjacency of `slots` entries: coop_grid, cut to
// sqrt(slots) / coop_grid_sqrt_div blocks rounded up to a multiple of 8 (one
// per XCD) when that knob is set.  A level costs about a * blocks (its grid
// barrier and per-block setup) + b * slots / blocks (its work), which is
// least at blocks ~ sqrt(slots); measured optima: ~120 blocks at RMAT-16,
// 248 at 18, 504 at 20, the whole resident grid (768) from 22 on.
int traversal_grid(byo_graph *g, const void *kernel, int block, uint64_t slots);

// Word pieces of a cooperative launch of nblocks blocks, for the traversal
// passes (buffer "bfs_pieces", knob "pieces") or the personalized PageRank
// sweeps (buffer "ppr_pieces", knob "ppr_pieces": their gathers balance
// best at fewer pieces).  The bounds depend only on the graph and the grid,
// so they are computed by the first launch after either changes and reused
// by the next ones (ready set); a launcher reports a queued launch with
// pieces_launched.
enum PieceSet { kTraversalPieces = 0, kPprPieces = 1, kPmaPieces = 2 };
Pieces make_pieces(byo_graph *g, int nblocks, PieceSet set = kTraversalPieces);
// Records that a launch computing pc's bounds has been queued.
void pieces_launched(byo_graph *g, const Pieces &pc, int nblocks, PieceSet set = kTraversalPieces);

}  // namespace byo
