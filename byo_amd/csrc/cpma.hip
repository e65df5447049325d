// cpma.hip -- CPMA-style packed-memory-array graph container on the device.
//
// Replaces PMA_graph over CPMA<PMA_SETTINGS<uint64_t>> (benchmarks/
// run_structures/run_single_pma.cc:29-140): the edge set is the set of
// uint64 keys (src<<32)|dst (run_single_pma.cc:83), kept sorted in fixed
// size leaves of kLeaf slots.  Leaf l holds cnt[l] keys packed at its start
// (the rest are the empty key ~0), all in [sep[l], sep[l+1]).  The slots are
// stored as two word arrays, src (high word) and dst (low word), so that a
// neighbour scan over the PMA reads only 4-byte dst words: with the
// per-vertex slot starts (cpma_index), the dst array is a CSR whose lists
// may hold empty slots (dst ~0) at leaf ends.
//
// insert_batch / remove_batch (Graph::insert_batch, gbbs/graph.h:3071-3124:
// sort + unique, then insert_sorted_batch):
//   1. the batch is radix sorted and deduplicated on the device;
//   2. every key finds its leaf by binary search over sep[];
//   3. the sorted batch splits into one contiguous segment per touched leaf;
//   4. insert: a leaf whose keys plus new keys fit is merged in place by one
//      workgroup in LDS (only touched leaves move); if any leaf would
//      overflow, the whole array is rebalanced: leaves are compacted, merged
//      with the batch by merge path, deduplicated and spread evenly over a
//      (possibly grown) leaf array at the target density;
//      remove: each touched leaf drops the keys found in its segment (binary
//      search) and compacts in place; the array shrinks and rebalances when
//      its density falls below kMinDensity.
// Degrees are maintained per update (the stale-degree problem of
// Graph::finilize caching, gbbs/graph.h:2559-2598, does not arise), and a
// CSR view (offsets + ids) for the traversal kernels is materialised from
// the leaves when an algorithm asks for it after an update.
#include <algorithm>

#include "primitives.hpp"

namespace byo {

constexpr uint32_t kLeaf = 512;
constexpr double kTargetDensity = 0.6;
constexpr double kMinDensity = 0.2;
constexpr uint64_t kEmpty = ~0ull;

struct Cpma {
  uint64_t nleaves = 0;
  uint64_t total = 0;
  DevBuf hi;     // uint32[nleaves * kLeaf]: src word of every slot (~0 empty)
  DevBuf lo;     // uint32[nleaves * kLeaf]: dst word of every slot (~0 empty)
  DevBuf cnt;    // uint32[nleaves]
  DevBuf sep;    // uint64[nleaves + 1]
  DevBuf deg;    // uint32[n]
  uint64_t rebalances = 0;
  uint64_t touched = 0;  // PMA slot bytes rewritten by all updates so far (S_touched)
  // per-vertex slot starts (n + 1; getExtraData's per-vertex index,
  // run_single_pma.cc:88): v's keys lie in slots [vstart[v], vstart[v+1]),
  // with empty slots at leaf ends in between; rebuilt after an update
  DevBuf vstart;
  bool index_valid = false;
};

static Cpma *state(byo_graph *g) { return static_cast<Cpma *>(g->cpma); }

// Slot access over the two word arrays.
struct Leaves {
  uint32_t *hi, *lo;
  __device__ __forceinline__ uint64_t get(uint64_t s) const { return ((uint64_t)hi[s] << 32) | lo[s]; }
  __device__ __forceinline__ void put(uint64_t s, uint64_t k) const {
    hi[s] = (uint32_t)(k >> 32);
    lo[s] = (uint32_t)k;
  }
  // x among the len sorted keys from slot base
  __device__ __forceinline__ bool has(uint64_t base, uint64_t len, uint64_t x) const {
    uint64_t a = 0, b = len;
    while (a < b) {
      const uint64_t mid = (a + b) >> 1;
      if (get(base + mid) < x) a = mid + 1; else b = mid;
    }
    return a < len && get(base + a) == x;
  }
};

static Leaves leaves(Cpma *c) { return Leaves{c->hi.as<uint32_t>(), c->lo.as<uint32_t>()}; }

// Every view derived from the edge set is stale after a change.
static void cpma_changed(byo_graph *g) {
  g->csr_valid = false;
  g->maxdeg_valid = false;
  g->pieces_stale();
  g->rl_valid = false;
  g->xs_valid = false;
  g->pb_valid = false;
  state(g)->index_valid = false;
}

namespace {

struct UpdCtl {
  unsigned long long changed;   // keys inserted / removed
  unsigned long long overflow;  // a leaf would overflow
  unsigned long long bad;       // an id out of range
  // pad[0] unique batch keys, pad[1] merged count of a global rebalance,
  // pad[2] global rebalance needed, pad[3] root windows, pad[4] PMA slot
  // bytes the update rewrote (S_touched: keys of merged leaves and rebuilt
  // windows after an insert, keys of compacted leaves before a remove, 8 B each)
  unsigned long long pad[5];
};

__global__ void k_check_ids(const uint64_t *keys, uint64_t b, uint64_t n, UpdCtl *ctl) {
  for (uint64_t i = global_tid(); i < b; i += grid_threads()) {
    const uint64_t k = keys[i];
    if ((k >> 32) >= n || (k & 0xFFFFFFFFull) >= n) atomicOr(&ctl->bad, 1ull);
  }
}

// Leaf l receives keys [l*m/L, (l+1)*m/L) of the sorted unique array.
__global__ void k_spread(const uint64_t *keys, uint64_t m, uint64_t nleaves, Leaves lv, uint32_t *cnt,
                         uint64_t *sep) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t l = wave; l < nleaves; l += nwaves) {
    const uint64_t s = (uint64_t)(((unsigned __int128)m * l) / nleaves);
    const uint64_t e = (uint64_t)(((unsigned __int128)m * (l + 1)) / nleaves);
    for (uint32_t i = lane; i < kLeaf; i += 64) lv.put(l * kLeaf + i, (s + i < e) ? keys[s + i] : kEmpty);
    if (lane == 0) {
      cnt[l] = (uint32_t)(e - s);
      sep[l] = (l == 0) ? 0ull : (s < m ? keys[s] : kEmpty);
      if (l == nleaves - 1) sep[nleaves] = kEmpty;
    }
  }
}

__global__ void k_degrees_from_off(const uint64_t *off, uint64_t n, uint32_t *deg) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) deg[v] = (uint32_t)(off[v + 1] - off[v]);
}

// leaf index of each batch key: last l with sep[l] <= key.  The unique key
// count *bu is read on the device (keys at and past it are skipped), so the
// same launch serves a batch whose size is known only after deduplication.
// The keys are sorted, so the leaves of a wave's 64 keys lie between the
// leaves of its first and last key: lanes 0 and 1 search those two over all
// leaves, and every lane then searches only that range.
__device__ __forceinline__ uint64_t leaf_search(const uint64_t *sep, uint64_t lo, uint64_t hi, uint64_t k) {
  // sep[lo] <= k < sep[hi] (sep[nleaves] = ~0 > any key)
  while (hi - lo > 1) {
    const uint64_t mid = (lo + hi) >> 1;
    if (sep[mid] <= k) lo = mid; else hi = mid;
  }
  return lo;
}

__global__ void k_leaf_of(const uint64_t *keys, const unsigned long long *bu, const uint64_t *sep, uint64_t nleaves,
                          uint32_t *leaf) {
  const uint64_t b = *bu;
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t base = wave * 64; base < b; base += nwaves * 64) {
    const uint64_t i = base + lane;
    const uint64_t last = std::min<uint64_t>(b - 1, base + 63);
    const uint64_t k = keys[i < b ? i : last];
    // lane 0: the leaf of the wave's first key, lane 1: of its last
    const uint64_t kend = lane == 1 ? keys[last] : k;
    const uint64_t mine = lane < 2 ? leaf_search(sep, 0, nleaves, kend) : 0;
    const uint64_t l0 = __shfl(mine, 0, 64), l1 = __shfl(mine, 1, 64);
    if (i < b) leaf[i] = (uint32_t)leaf_search(sep, l0, l1 + 1, k);
  }
}

// Segment heads over the first *bu keys; the flags up to cap are written
// (zero past *bu) so a scan over cap entries counts the segments.
__global__ void k_seg_flags(const uint32_t *leaf, const unsigned long long *bu, uint64_t cap, uint32_t *flag) {
  const uint64_t b = *bu;
  for (uint64_t i = global_tid(); i < cap; i += grid_threads())
    flag[i] = (i < b && (i == 0 || leaf[i] != leaf[i - 1])) ? 1u : 0u;
}

__global__ void k_seg_starts(const uint32_t *flag, const uint32_t *pos, uint64_t cap, uint32_t *starts) {
  for (uint64_t i = global_tid(); i < cap; i += grid_threads())
    if (flag[i]) starts[pos[i]] = (uint32_t)i;
}

// Device-side counts of one update: unique keys *bu (ctl->pad[0]) and
// segments *nseg (the scan total of the flags).  halt, when given, is the
// captured-sequence stop flag: a set flag turns the CPMA-writing kernels of
// this and every later batch into no-ops.
struct UpdDev {
  const unsigned long long *bu;
  const uint32_t *nseg;
  const uint32_t *halt;
};

__device__ __forceinline__ bool in_sorted(const uint64_t *a, uint64_t len, uint64_t x) {
  uint64_t lo = 0, hi = len;
  while (lo < hi) {
    uint64_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < len && a[lo] == x;
}

__device__ __forceinline__ uint32_t lower_rank(const uint64_t *a, uint32_t len, uint64_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// The three per-leaf update passes give each touched leaf (one segment of
// the sorted batch) to one wave: the leaf's keys are staged in the wave's
// LDS with one coalesced pass and every lookup is an LDS binary search, and
// the wave synchronises only itself.  (One workgroup per leaf spent most of
// its time in barriers and dependent global round trips: a batch touches
// hundreds of thousands of leaves, each holding a few hundred keys.)
constexpr int kUpdWaves = 4;  // waves per 256-thread workgroup
// Workgroups of a per-leaf pass: enough to fill every CU at the passes'
// LDS footprint (16-32 KB per workgroup), few enough that each wave takes
// many leaves and the block tallies stay few.
constexpr uint64_t kUpdGrid = 4096;

// x among the len sorted keys of an LDS array.
__device__ __forceinline__ bool lds_has(const uint64_t *a, uint32_t len, uint64_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < len && a[lo] == x;
}

// A wave's segments are j, j + stride, ...; the bounds of the next one are
// loaded while the current one is worked on, and its leaf right after the
// current leaf's slots have been requested, so the dependent chain
// starts[] -> leaf_of[] -> slots costs one round trip per segment, not three.
__device__ __forceinline__ void seg_bounds(const uint32_t *starts, uint64_t nseg, uint64_t b, uint64_t j,
                                           uint64_t &s, uint64_t &e) {
  if (j < nseg) {
    s = starts[j];
    e = (j + 1 < nseg) ? starts[j + 1] : b;
  }
}

// All kLeaf slots of the leaf at slot `leaf` into registers (slot t*64 +
// lane in h[t], l[t]): the loads do not wait for the leaf's count, which
// comes in parallel.
__device__ __forceinline__ void load_leaf(const Leaves &lv, uint64_t leaf, uint32_t (&h)[kLeaf / 64],
                                          uint32_t (&l)[kLeaf / 64], int lane) {
#pragma unroll
  for (int t = 0; t < (int)(kLeaf / 64); t++) {
    h[t] = lv.hi[leaf + t * 64 + lane];
    l[t] = lv.lo[leaf + t * 64 + lane];
  }
}

// The first c keys of a leaf loaded by load_leaf into the wave's LDS array.
__device__ __forceinline__ void put_leaf(const uint32_t (&h)[kLeaf / 64], const uint32_t (&l)[kLeaf / 64], uint32_t c,
                                         uint64_t *dst, int lane) {
#pragma unroll
  for (int t = 0; t < (int)(kLeaf / 64); t++) {
    const uint32_t i = t * 64 + lane;
    if (i < c) dst[i] = ((uint64_t)h[t] << 32) | l[t];
  }
}

// Loads up to kLeaf consecutive batch keys [s0, min(s0 + kLeaf, e)) into
// registers, lane-strided (key s0 + t*64 + lane in k[t]), all loads in flight.
__device__ __forceinline__ void load_keys(const uint64_t *keys, uint64_t s0, uint64_t e, uint64_t (&k)[kLeaf / 64],
                                          int lane) {
#pragma unroll
  for (int t = 0; t < (int)(kLeaf / 64); t++) {
    const uint64_t i = s0 + (uint64_t)t * 64 + lane;
    k[t] = i < e ? keys[i] : 0;
  }
}

// deg[src] += sign * (number of keys of src) over the n sorted keys of an
// LDS array: the first key of each source's run adds the run's length, so
// a hub with thousands of keys in a batch costs one atomic per leaf it
// touches, not one per key (per-key atomics on one address queue in L2).
__device__ __forceinline__ void add_degree_runs(const uint64_t *a, uint32_t n, uint32_t *deg, int sign, int lane) {
  for (uint32_t i = lane; i < n; i += 64) {
    const uint32_t src = (uint32_t)(a[i] >> 32);
    if (i > 0 && (uint32_t)(a[i - 1] >> 32) == src) continue;
    const uint32_t end = src == 0xFFFFFFFFu ? n : lower_rank(a, n, ((uint64_t)src + 1) << 32);
    const uint32_t len = end - i;
    if (sign > 0) atomicAdd(&deg[src], len); else atomicSub(&deg[src], len);
  }
}

// Adds the block's keys changed and slot bytes touched to the update's
// control block: one pair of atomics per workgroup.  (One pair per wave put
// 65,536 atomics on the same two words for a 1e7-key batch; they queue in
// one L2 channel and took about a third of the pass.)  Every thread of the
// block calls it.
__device__ __forceinline__ void block_tally(UpdCtl *ctl, unsigned long long changed, unsigned long long touched,
                                            int lane) {
  __shared__ unsigned long long red[2][kUpdWaves];
  if (lane == 0) {
    red[0][threadIdx.x >> 6] = changed;
    red[1][threadIdx.x >> 6] = touched;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long c = 0, t = 0;
    for (int w = 0; w < kUpdWaves; w++) {
      c += red[0][w];
      t += red[1][w];
    }
    if (c) {
      atomicAdd(&ctl->changed, c);
      atomicAdd(&ctl->pad[4], t);
    }
  }
}

// Pass 1 of an insert: count the segment keys absent from their leaf.
// done (optional): segments merged already by the one-pass merge (their new
// keys are in the leaf: 0 counted, nothing loaded).
__global__ void __launch_bounds__(256) k_ins_count(const uint64_t *keys, UpdDev d, const uint32_t *starts,
                                                   const uint32_t *leaf_of, Leaves lv,
                                                   const uint32_t *cnt, uint32_t *newc, UpdCtl *ctl,
                                                   const uint8_t *done = nullptr) {
  __shared__ uint64_t stage[kUpdWaves][kLeaf];
  if (d.halt && *d.halt) return;
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  uint64_t *old = stage[wv];
  const uint64_t b = *d.bu, nseg = *d.nseg, stride = grid_threads() >> 6;
  uint64_t j = global_tid() >> 6, s = 0, e = 0;
  uint32_t l = 0;
  seg_bounds(starts, nseg, b, j, s, e);
  if (j < nseg) l = leaf_of[s];
  for (; j < nseg; j += stride) {
    uint64_t sn = 0, en = 0;
    seg_bounds(starts, nseg, b, j + stride, sn, en);
    if (done && done[j]) {
      if (lane == 0) newc[j] = 0;
      s = sn;
      e = en;
      l = (j + stride < nseg) ? leaf_of[sn] : 0;
      continue;
    }
    const uint32_t c = cnt[l];
    uint32_t h[kLeaf / 64], lo[kLeaf / 64];
    load_leaf(lv, (uint64_t)l * kLeaf, h, lo, lane);
    uint64_t k[kLeaf / 64];
    load_keys(keys, s, e, k, lane);
    const uint32_t ln = (j + stride < nseg) ? leaf_of[sn] : 0;
    put_leaf(h, lo, c, old, lane);
    wave_lds_sync();
    uint32_t mine = 0;
    for (uint64_t s0 = s; s0 < e; s0 += kLeaf) {
      if (s0 != s) load_keys(keys, s0, e, k, lane);
#pragma unroll
      for (int t = 0; t < (int)(kLeaf / 64); t++)
        if (s0 + (uint64_t)t * 64 + lane < e) mine += !lds_has(old, c, k[t]);
    }
    mine = wave_sum(mine);
    if (lane == 0) {
      newc[j] = mine;
      if ((uint64_t)c + mine > kLeaf) atomicOr(&ctl->overflow, 1ull);
    }
    wave_lds_sync();  // the stage is rewritten by the next segment
    s = sn;
    e = en;
    l = ln;
  }
}

// Pass 2 of an insert (no leaf overflows): merge in LDS and write back.
// With done: first == true is the one-pass merge, run before any count: it
// merges every segment whose leaf holds room for all of the segment's keys
// (c + segment length <= kLeaf, so it cannot overflow, whatever the count)
// and marks it done; first == false skips the done segments.
__global__ void __launch_bounds__(256) k_ins_merge(const uint64_t *keys, UpdDev d, const uint32_t *starts,
                                                   const uint32_t *leaf_of, Leaves lv, uint32_t *cnt,
                                                   uint32_t *deg, UpdCtl *ctl, const uint32_t *cand,
                                                   uint64_t nleaves, uint8_t *done = nullptr, bool first = false) {
  // one leaf-sized LDS array per wave: the leaf's c keys at its start, the
  // absent batch keys after them (c + their number <= kLeaf: no leaf
  // overflows here), so 8 workgroups fit a CU instead of 5
  __shared__ uint64_t stage[kUpdWaves][kLeaf];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  uint64_t *old_keys = stage[wv];
  // stop on a halted sequence or a global rebalance; the leaves of rebuilt
  // windows (candidate windows, when a leaf overflowed) are skipped
  if ((d.halt && *d.halt) || ctl->pad[2] || ctl->bad) return;
  const bool windows = ctl->overflow != 0;
  const uint64_t b = *d.bu, nseg = *d.nseg, stride = grid_threads() >> 6;
  unsigned long long changed = 0, touched = 0;
  uint64_t j = global_tid() >> 6, s = 0, e = 0;
  uint32_t l = 0;
  seg_bounds(starts, nseg, b, j, s, e);
  if (j < nseg) l = leaf_of[s];
  for (; j < nseg; j += stride) {
    uint64_t sn = 0, en = 0;
    seg_bounds(starts, nseg, b, j + stride, sn, en);
    bool skip = false;
    if (windows) {
      // levels 1.. of the aligned windows holding l, one per lane
      const int k = lane + 1;
      bool inw = false;
      if (k < 40 && (1ull << (k - 1)) < nleaves) inw = cand[l & ~((1ull << k) - 1)] >= (uint32_t)k;
      skip = __ballot(inw) != 0;
    }
    const uint64_t leaf = (uint64_t)l * kLeaf;
    uint32_t c = 0, na = 0;
    uint32_t h[kLeaf / 64], lo[kLeaf / 64];
    uint64_t k[kLeaf / 64];
    if (done) {
      if (first) {
        const bool fits = (uint64_t)cnt[l] + (e - s) <= kLeaf;
        if (lane == 0) done[j] = fits ? 1 : 0;
        skip = !fits;
      } else {
        skip = skip || done[j];
      }
    }
    if (!skip) {
      c = cnt[l];
      load_leaf(lv, leaf, h, lo, lane);
      load_keys(keys, s, e, k, lane);
    }
    const uint32_t ln = (j + stride < nseg) ? leaf_of[sn] : 0;
    if (!skip) {
      put_leaf(h, lo, c, old_keys, lane);
      uint64_t *add = old_keys + c;
      wave_lds_sync();
      // stable compaction of the absent keys, 64 at a time (a segment holds
      // at most kLeaf keys: the unique absent ones fit the leaf, the rest
      // are in it; a longer one is taken kLeaf keys at a time)
      for (uint64_t s0 = s; s0 < e; s0 += kLeaf) {
        if (s0 != s) load_keys(keys, s0, e, k, lane);
#pragma unroll
        for (int t = 0; t < (int)(kLeaf / 64); t++) {
          const uint64_t i = s0 + (uint64_t)t * 64 + lane;
          const bool keep = i < e && !lds_has(old_keys, c, k[t]);
          const uint64_t mask = __ballot(keep);
          if (keep) add[na + mbcnt(mask)] = k[t];
          na += (uint32_t)__popcll(mask);
        }
      }
      wave_lds_sync();
      // merged position of each key = own index + rank in the other list
      for (uint32_t i = lane; i < c; i += 64) {
        const uint64_t x = old_keys[i];
        lv.put(leaf + i + lower_rank(add, na, x), x);
      }
      for (uint32_t i = lane; i < na; i += 64) {
        const uint64_t x = add[i];
        lv.put(leaf + i + lower_rank(old_keys, c, x), x);
      }
      add_degree_runs(add, na, deg, 1, lane);
      if (lane == 0) cnt[l] = c + na;
      if (na) {
        changed += na;
        touched += 8ull * (c + na);
      }
      wave_lds_sync();  // the stages are rewritten by the next segment
    }
    s = sn;
    e = en;
    l = ln;
  }
  block_tally(ctl, changed, touched, lane);
}

// Removal: each touched leaf keeps the keys absent from its segment.
__global__ void __launch_bounds__(256) k_remove(const uint64_t *keys, UpdDev d, const uint32_t *starts,
                                                const uint32_t *leaf_of, Leaves lv, uint32_t *cnt,
                                                uint32_t *deg, UpdCtl *ctl) {
  __shared__ uint64_t stage[kUpdWaves][kLeaf];
  __shared__ uint64_t stage_rm[kUpdWaves][kLeaf];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  uint64_t *kept = stage[wv], *rm = stage_rm[wv];
  if ((d.halt && *d.halt) || ctl->bad) return;
  const uint64_t b = *d.bu, nseg = *d.nseg, stride = grid_threads() >> 6;
  unsigned long long changed = 0, touched = 0;
  uint64_t j = global_tid() >> 6, s = 0, e = 0;
  uint32_t l = 0;
  seg_bounds(starts, nseg, b, j, s, e);
  if (j < nseg) l = leaf_of[s];
  for (; j < nseg; j += stride) {
    uint64_t sn = 0, en = 0;
    seg_bounds(starts, nseg, b, j + stride, sn, en);
    const uint64_t leaf = (uint64_t)l * kLeaf;
    const uint32_t c = cnt[l];
    // the leaf's slots in registers and the segment's keys in LDS (when they
    // fit; a longer segment is searched in global memory), all loads issued
    // before any is used
    constexpr int kSteps = kLeaf / 64;
    uint32_t h[kSteps], lo[kSteps];
    load_leaf(lv, leaf, h, lo, lane);
    const uint64_t ns = e - s;
    const bool in_lds = ns <= kLeaf;
    uint64_t k[kLeaf / 64];
    if (in_lds) load_keys(keys, s, e, k, lane);
    const uint32_t ln = (j + stride < nseg) ? leaf_of[sn] : 0;
    if (in_lds) {
#pragma unroll
      for (int t = 0; t < kSteps; t++)
        if ((uint64_t)t * 64 + lane < ns) rm[t * 64 + lane] = k[t];
    }
    wave_lds_sync();
    uint32_t run = 0;
    bool gone[kSteps];
#pragma unroll
    for (int t = 0; t < kSteps; t++) {
      const uint32_t i = t * 64 + lane;
      const uint64_t x = ((uint64_t)h[t] << 32) | lo[t];
      bool keep = false;
      gone[t] = false;
      if (i < c) {
        gone[t] = in_lds ? lds_has(rm, (uint32_t)ns, x) : in_sorted(keys + s, ns, x);
        keep = !gone[t];
      }
      const uint64_t mask = __ballot(keep);
      if (keep) kept[run + mbcnt(mask)] = x;
      run += (uint32_t)__popcll(mask);
    }
    wave_lds_sync();
    // the removed keys, compacted in order over the segment's LDS copy (read
    // by nobody now), lower the degrees one atomic per source run
    uint32_t ng = 0;
#pragma unroll
    for (int t = 0; t < kSteps; t++) {
      const uint64_t mask = __ballot(gone[t]);
      if (gone[t]) rm[ng + mbcnt(mask)] = ((uint64_t)h[t] << 32) | lo[t];
      ng += (uint32_t)__popcll(mask);
    }
    wave_lds_sync();
    add_degree_runs(rm, ng, deg, -1, lane);
    if (run < c) {
      for (uint32_t i = lane; i < c; i += 64) lv.put(leaf + i, i < run ? kept[i] : kEmpty);
      if (lane == 0) cnt[l] = run;
      changed += c - run;
      touched += 8ull * c;
    }
    wave_lds_sync();  // the stage is rewritten by the next segment
    s = sn;
    e = en;
    l = ln;
  }
  block_tally(ctl, changed, touched, lane);
}

// ------------------------------------------- windowed rebalance (insert)
// A leaf that would overflow is fixed locally, as a packed-memory array
// does: the smallest aligned window of 2^k leaves (k >= 1) whose keys plus
// new keys fit at kWinDensity fill is rebuilt -- its old keys and the
// batch keys that fall in it merged and spread evenly over its leaves --
// and only when the whole array is too dense (ctl->pad[2]) does the update
// fall back to a global rebalance.  Windows nest (aligned), so the roots
// are the candidate windows no larger candidate contains; the other
// segments merge in their leaves as before.
constexpr double kWinDensity = 0.75;

struct WinArgs {
  const uint64_t *keys;  // sorted unique batch
  UpdDev d;
  const uint32_t *starts, *leaf_of, *newc;
  Leaves lv;
  uint32_t *cnt;
  uint64_t *sep;
  uint32_t *deg;
  uint64_t nleaves;
  uint64_t *dem;        // nleaves: keys + new keys per leaf
  uint64_t *P;          // nleaves + 1: exclusive prefix of dem
  uint32_t *cand;       // nleaves: largest candidate window level based at the leaf
  uint64_t *roots;      // (base << 8) | level of each root window
  uint64_t *stage;      // old keys of a window, at its base * kLeaf
  uint64_t *stage_new;  // new keys of all windows, in batch order
  UpdCtl *ctl;
  // grid-wide rebuild
  uint64_t *cnt64;      // nleaves: counts widened for the scan
  uint64_t *pc;         // nleaves + 1: exclusive prefix of the counts
  uint32_t *nflag;      // cap: batch key is new and in a window
  uint32_t *npos;       // cap + 1: exclusive prefix of nflag
  uint64_t *rinfo;      // nleaves x 4 per root: base, end, first new key, new-key count
  uint64_t *rsize;      // nleaves: merged size of root r (0 past the roots)
  uint64_t *rpre;       // nleaves + 1: exclusive prefix of rsize
  uint64_t cap;         // batch capacity
};

// Root window (base, level) containing leaf l, or level 0: the largest
// candidate window that holds it.
__device__ __forceinline__ uint32_t root_of(const WinArgs &w, uint64_t l, uint64_t *base) {
  int kmax = 1;
  while (kmax < 40 && (1ull << kmax) < w.nleaves) kmax++;
  for (int k = kmax; k >= 1; k--) {
    const uint64_t b = l & ~((1ull << k) - 1);
    const uint32_t c = w.cand[b];
    if (c >= (uint32_t)k) {
      *base = b;
      return c;
    }
  }
  return 0;
}

__device__ __forceinline__ bool win_skip(const WinArgs &w) {
  return (w.d.halt && *w.d.halt) || !w.ctl->overflow || w.ctl->bad;
}

// True when leaf l lies in a candidate window (hence in a root window).
__device__ __forceinline__ bool in_window(const WinArgs &w, uint64_t l) {
  for (int k = 1; k < 40 && (1ull << (k - 1)) < w.nleaves; k++) {
    const uint64_t b = l & ~((1ull << k) - 1);
    if (w.cand[b] >= (uint32_t)k) return true;
  }
  return false;
}

__global__ void k_win_dem(WinArgs w) {
  if (win_skip(w)) return;
  for (uint64_t l = global_tid(); l < w.nleaves; l += grid_threads()) {
    w.dem[l] = w.cnt[l];
    w.cand[l] = 0;
  }
}

__global__ void k_win_add(WinArgs w) {
  if (win_skip(w)) return;
  const uint64_t nseg = *w.d.nseg;
  for (uint64_t j = global_tid(); j < nseg; j += grid_threads()) {
    const uint32_t l = w.leaf_of[w.starts[j]];
    w.dem[l] = (uint64_t)w.cnt[l] + w.newc[j];
  }
}

__global__ void k_win_pick(WinArgs w) {
  if (win_skip(w)) return;
  const uint64_t nseg = *w.d.nseg;
  for (uint64_t j = global_tid(); j < nseg; j += grid_threads()) {
    const uint64_t l = w.leaf_of[w.starts[j]];
    if ((uint64_t)w.cnt[l] + w.newc[j] <= kLeaf) continue;
    int k = 1;
    for (;; k++) {
      const uint64_t size = 1ull << k;
      const uint64_t b = l & ~(size - 1);
      const uint64_t e = std::min<uint64_t>(b + size, w.nleaves);
      if ((double)(w.P[e] - w.P[b]) <= kWinDensity * (double)((e - b) * kLeaf)) {
        atomicMax(&w.cand[b], (uint32_t)k);
        break;
      }
      if (b == 0 && e == w.nleaves) {  // the whole array is too dense
        atomicOr(&w.ctl->pad[2], 1ull);
        break;
      }
    }
  }
}

__global__ void k_win_roots(WinArgs w) {
  if (win_skip(w) || w.ctl->pad[2]) return;
  for (uint64_t b = global_tid(); b < w.nleaves; b += grid_threads()) {
    const uint32_t k = w.cand[b];
    if (!k) continue;
    bool contained = false;
    for (int k2 = (int)k + 1; k2 < 40 && (1ull << (k2 - 1)) < w.nleaves; k2++) {
      const uint64_t b2 = b & ~((1ull << k2) - 1);
      if (w.cand[b2] >= (uint32_t)k2) {
        contained = true;
        break;
      }
    }
    if (!contained) w.roots[atomicAdd(&w.ctl->pad[3], 1ull)] = (b << 8) | k;
  }
}

// Rebuild of the root windows by the whole grid, in kernels:
//  w1  counts widened (then scanned into pc);
//  w2  every leaf of a window copies its keys to the window's staging run;
//  w3  a batch key is new when it is in a window and its own leaf lacks it
//      (then scanned into npos; w4 compacts them, in batch order, and counts
//      them into the degrees);
//  w5  per root: its old and new key ranges and merged size (scanned into
//      rpre);
//  w6  merged positions of all windows, 64 per thread, one merge-path search
//      per run: leaf i of a window of W leaves and T keys takes positions
//      [i T / W, (i+1) T / W);
//  w7  counts and empty tails of the window leaves.
__device__ __forceinline__ bool win_live(const WinArgs &w) { return !win_skip(w) && !w.ctl->pad[2]; }

__global__ void k_win_w1(WinArgs w) {
  if (!win_live(w)) return;
  for (uint64_t l = global_tid(); l < w.nleaves; l += grid_threads()) w.cnt64[l] = w.cnt[l];
}

__global__ void k_win_w2(WinArgs w) {
  if (!win_live(w)) return;
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t l = wave; l < w.nleaves; l += nwaves) {
    uint64_t b = 0;
    if (!root_of(w, l, &b)) continue;
    const uint32_t c = w.cnt[l];
    uint64_t *dst = w.stage + b * kLeaf + (w.pc[l] - w.pc[b]);
    for (uint32_t i = lane; i < c; i += 64) dst[i] = w.lv.get(l * kLeaf + i);
  }
}

__global__ void k_win_w3(WinArgs w) {
  if (!win_live(w)) return;
  const uint64_t bu = *w.d.bu;
  for (uint64_t i = global_tid(); i < w.cap; i += grid_threads()) {
    uint32_t f = 0;
    if (i < bu) {
      const uint32_t l = w.leaf_of[i];
      uint64_t b = 0;
      if (root_of(w, l, &b)) f = !w.lv.has((uint64_t)l * kLeaf, w.cnt[l], w.keys[i]);
    }
    w.nflag[i] = f;
  }
}

__global__ void k_win_w4(WinArgs w) {
  if (!win_live(w)) return;
  const uint64_t bu = *w.d.bu;
  for (uint64_t i = global_tid(); i < bu; i += grid_threads())
    if (w.nflag[i]) {
      const uint64_t key = w.keys[i];
      w.stage_new[w.npos[i]] = key;
      atomicAdd(&w.deg[key >> 32], 1u);
    }
}

__global__ void k_win_w5(WinArgs w) {
  if (!win_live(w)) return;
  const uint64_t nroots = w.ctl->pad[3];
  const uint64_t bu = *w.d.bu, nseg = *w.d.nseg;
  for (uint64_t r = global_tid(); r < w.nleaves; r += grid_threads()) {
    if (r >= nroots) {
      w.rsize[r] = 0;
      continue;
    }
    const uint64_t b = w.roots[r] >> 8;
    const uint64_t e = std::min<uint64_t>(b + (1ull << (w.roots[r] & 0xFF)), w.nleaves);
    uint64_t jlo = 0, jhi = nseg;  // first segment with leaf >= b
    while (jlo < jhi) {
      const uint64_t mid = (jlo + jhi) >> 1;
      if (w.leaf_of[w.starts[mid]] < b) jlo = mid + 1; else jhi = mid;
    }
    uint64_t j2 = jlo, j3 = nseg;  // first segment with leaf >= e
    while (j2 < j3) {
      const uint64_t mid = (j2 + j3) >> 1;
      if (w.leaf_of[w.starts[mid]] < e) j2 = mid + 1; else j3 = mid;
    }
    const uint64_t ks = jlo < nseg ? w.starts[jlo] : bu, ke = j2 < nseg ? w.starts[j2] : bu;
    const uint64_t n0 = w.npos[ks], nn = w.npos[ke] - n0;
    w.rinfo[4 * r] = b;
    w.rinfo[4 * r + 1] = e;
    w.rinfo[4 * r + 2] = n0;
    w.rinfo[4 * r + 3] = nn;
    w.rsize[r] = (w.pc[e] - w.pc[b]) + nn;
    atomicAdd(&w.ctl->pad[4], 8ull * ((w.pc[e] - w.pc[b]) + nn));
    w.cnt64[b] = (w.pc[e] - w.pc[b]) + nn;  // merged size by root base (cnt64 is free after w1's scan)
    if (nn) atomicAdd(&w.ctl->changed, (unsigned long long)nn);
  }
}

__global__ void k_win_w6(WinArgs w) {
  if (!win_live(w)) return;
  const uint64_t nroots = w.ctl->pad[3];
  if (!nroots) return;
  const uint64_t total = w.rpre[nroots];
  constexpr uint64_t kRun = 64;
  for (uint64_t p0 = global_tid() * kRun; p0 < total; p0 += grid_threads() * kRun) {
    uint64_t p = p0;
    const uint64_t pend = std::min<uint64_t>(p0 + kRun, total);
    while (p < pend) {
      uint64_t lo = 0, hi = nroots;  // root r with rpre[r] <= p < rpre[r + 1]
      while (hi - lo > 1) {
        const uint64_t mid = (lo + hi) >> 1;
        if (w.rpre[mid] <= p) lo = mid; else hi = mid;
      }
      const uint64_t r = lo;
      const uint64_t b = w.rinfo[4 * r], e = w.rinfo[4 * r + 1];
      const uint64_t W = e - b;
      const uint64_t T_old = w.pc[e] - w.pc[b], T_new = w.rinfo[4 * r + 3];
      const uint64_t T = T_old + T_new;
      const uint64_t *old = w.stage + b * kLeaf;
      const uint64_t *nw = w.stage_new + w.rinfo[4 * r + 2];
      const uint64_t q0 = p - w.rpre[r];
      const uint64_t q1 = std::min<uint64_t>(T, q0 + (pend - p));
      uint64_t x = q0 > T_new ? q0 - T_new : 0, y = q0 < T_old ? q0 : T_old;
      while (x < y) {
        const uint64_t mid = (x + y) >> 1;
        if (old[mid] < nw[q0 - mid - 1]) x = mid + 1; else y = mid;
      }
      uint64_t a = x, c = q0 - x;
      uint64_t i = (uint64_t)(((unsigned __int128)q0 * W) / T);
      while (i + 1 < W && (uint64_t)(((unsigned __int128)(i + 1) * T) / W) <= q0) i++;
      while (i > 0 && (uint64_t)(((unsigned __int128)i * T) / W) > q0) i--;
      uint64_t si = (uint64_t)(((unsigned __int128)i * T) / W);
      uint64_t snext = (uint64_t)(((unsigned __int128)(i + 1) * T) / W);
      for (uint64_t q = q0; q < q1; q++) {
        uint64_t key;
        if (c >= T_new || (a < T_old && old[a] < nw[c])) key = old[a++];
        else key = nw[c++];
        while (q >= snext) {
          i++;
          si = snext;
          snext = (uint64_t)(((unsigned __int128)(i + 1) * T) / W);
        }
        w.lv.put((b + i) * kLeaf + (q - si), key);
        if (q == si && i > 0) w.sep[b + i] = key;
      }
      p += q1 - q0;
    }
  }
}

__global__ void k_win_w7(WinArgs w) {
  if (!win_live(w)) return;
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t l = wave; l < w.nleaves; l += nwaves) {
    uint64_t b = 0;
    const uint32_t k = root_of(w, l, &b);
    if (!k) continue;
    // the root's merged size: old keys from pc, new keys from npos over the
    // window's batch range (recomputed from the segment search in w5's
    // record, found by the root's base)
    const uint64_t e = std::min<uint64_t>(b + (1ull << k), w.nleaves);
    const uint64_t W = e - b;
    const uint64_t T = w.cnt64[b];  // the root's merged size (w5)
    const uint64_t i = l - b;
    const uint64_t si = (uint64_t)(((unsigned __int128)i * T) / W);
    const uint64_t ei = (uint64_t)(((unsigned __int128)(i + 1) * T) / W);
    for (uint64_t x = (ei - si) + lane; x < kLeaf; x += 64) w.lv.put(l * kLeaf + x, kEmpty);
    if (lane == 0) w.cnt[l] = (uint32_t)(ei - si);
  }
}

// Small batches (b <= kSmallBatch): one workgroup checks the ids, sorts the
// keys in LDS (bitonic), deduplicates them and cuts the segments -- the work
// of ~40 launches of the general path (radix passes, scans) in one.
constexpr uint32_t kSmallBatch = 4096;
constexpr int kSmallThreads = 1024;

__device__ __forceinline__ uint32_t block_excl_scan4k(uint32_t *v, uint32_t *part, uint32_t *total) {
  // v[4096] in LDS, 4 consecutive entries per thread; returns nothing, v
  // becomes its exclusive prefix and *total the sum
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  uint32_t a0 = v[4 * t], a1 = v[4 * t + 1], a2 = v[4 * t + 2], a3 = v[4 * t + 3];
  const uint32_t sum = a0 + a1 + a2 + a3;
  const uint32_t incl = wave_incl_scan(sum);
  if (lane == 63) part[wv] = incl;
  __syncthreads();
  uint32_t before = 0, tot = 0;
  for (int x = 0; x < kSmallThreads / 64; x++) {
    if (x < wv) before += part[x];
    tot += part[x];
  }
  uint32_t p = before + incl - sum;
  v[4 * t] = p;
  p += a0;
  v[4 * t + 1] = p;
  p += a1;
  v[4 * t + 2] = p;
  p += a2;
  v[4 * t + 3] = p;
  if (t == 0) *total = tot;
  __syncthreads();
  return tot;
}

__global__ void __launch_bounds__(kSmallThreads) k_small_batch(const uint64_t *keys_in, uint64_t b, uint64_t n,
                                                                const uint64_t *sep, uint64_t nleaves,
                                                                uint64_t *uniq, uint32_t *leaf_of,
                                                                uint32_t *starts, uint32_t *pos_total,
                                                                UpdCtl *ctl) {
  __shared__ uint64_t k[kSmallBatch];
  __shared__ uint32_t f[kSmallBatch];
  __shared__ uint32_t lf[kSmallBatch];
  __shared__ uint32_t part[kSmallThreads / 64];
  __shared__ uint32_t tot;
  const int t = threadIdx.x;
  for (uint32_t i = t; i < kSmallBatch; i += kSmallThreads) {
    uint64_t v = kEmpty;
    if (i < b) {
      v = keys_in[i];
      if ((v >> 32) >= n || (v & 0xFFFFFFFFull) >= n) atomicOr(&ctl->bad, 1ull);
    }
    k[i] = v;
  }
  __syncthreads();
  uint32_t N = 64;
  while (N < b) N <<= 1;
  for (uint32_t size = 2; size <= N; size <<= 1)
    for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
      for (uint32_t i = t; i < N; i += kSmallThreads) {
        const uint32_t j = i ^ stride;
        if (j > i) {
          const bool up = (i & size) == 0;
          const uint64_t x = k[i], y = k[j];
          if ((x > y) == up) {
            k[i] = y;
            k[j] = x;
          }
        }
      }
      __syncthreads();
    }
  // unique keys (the empty key, padding or a bad id, is dropped)
  for (uint32_t i = t; i < kSmallBatch; i += kSmallThreads)
    f[i] = (i < N && k[i] != kEmpty && (i == 0 || k[i] != k[i - 1])) ? 1u : 0u;
  __syncthreads();
  for (uint32_t i = t; i < kSmallBatch; i += kSmallThreads) lf[i] = f[i];
  __syncthreads();
  const uint32_t bu = block_excl_scan4k(f, part, &tot);
  for (uint32_t i = t; i < N; i += kSmallThreads)
    if (lf[i]) uniq[f[i]] = k[i];
  __syncthreads();
  // leaves of the unique keys, then one segment per run of equal leaves
  for (uint32_t j = t; j < bu; j += kSmallThreads) {
    const uint64_t key = uniq[j];
    uint64_t lo = 0, hi = nleaves;
    while (hi - lo > 1) {
      const uint64_t mid = (lo + hi) >> 1;
      if (sep[mid] <= key) lo = mid; else hi = mid;
    }
    lf[j] = (uint32_t)lo;
    leaf_of[j] = (uint32_t)lo;
  }
  __syncthreads();
  for (uint32_t j = t; j < kSmallBatch; j += kSmallThreads) f[j] = (j < bu && (j == 0 || lf[j] != lf[j - 1])) ? 1u : 0u;
  __syncthreads();
  for (uint32_t j = t; j < bu; j += kSmallThreads) k[j] = f[j];  // keep the flags
  __syncthreads();
  const uint32_t nseg = block_excl_scan4k(f, part, &tot);
  for (uint32_t j = t; j < bu; j += kSmallThreads)
    if (k[j]) starts[f[j]] = j;
  if (t == 0) {
    ctl->pad[0] = bu;
    *pos_total = nseg;
  }
}

__global__ void k_widen(const uint32_t *a, uint64_t *b, uint64_t n) {
  for (uint64_t i = global_tid(); i < n; i += grid_threads()) b[i] = a[i];
}

// Compaction of all leaves into a flat sorted array (pos = prefix of cnt).
__global__ void k_compact(Leaves lv, const uint32_t *cnt, const uint64_t *pos, uint64_t nleaves,
                          uint64_t *flat) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t l = wave; l < nleaves; l += nwaves) {
    const uint32_t c = cnt[l];
    const uint64_t p = pos[l];
    for (uint32_t i = lane; i < c; i += 64) flat[p + i] = lv.get(l * kLeaf + i);
  }
}

// Merge path of sorted A[na] and B[nb] into out[na+nb] (A first on ties).
__global__ void k_merge(const uint64_t *A, uint64_t na, const uint64_t *B, uint64_t nb, uint64_t *out) {
  constexpr int K = 8;
  const uint64_t total = na + nb;
  for (uint64_t d = global_tid() * K; d < total; d += grid_threads() * K) {
    uint64_t lo = d > nb ? d - nb : 0, hi = d < na ? d : na;
    while (lo < hi) {
      uint64_t mid = (lo + hi) >> 1;
      if (A[mid] <= B[d - mid - 1]) lo = mid + 1; else hi = mid;
    }
    uint64_t i = lo, j = d - lo;
    for (int k = 0; k < K && d + k < total; k++) {
      if (j >= nb || (i < na && A[i] <= B[j])) out[d + k] = A[i++];
      else out[d + k] = B[j++];
    }
  }
}

__global__ void k_slot_keys_to_csr_off(const uint64_t *flat, uint64_t m, uint64_t n, uint64_t *off, uint32_t *ids) {
  for (uint64_t i = global_tid(); i < m; i += grid_threads()) {
    const uint64_t key = flat[i];
    const uint64_t s = key >> 32;
    ids[i] = (uint32_t)key;
    const uint64_t lo = (i == 0) ? 0 : (flat[i - 1] >> 32) + 1;
    for (uint64_t v = lo; v <= s; v++) off[v] = i;
    if (i == m - 1)
      for (uint64_t v = s + 1; v <= n; v++) off[v] = m;
  }
}

}  // namespace

// ------------------------------------------------------------ host side

static uint64_t leaves_for(uint64_t m) {
  uint64_t nl = (uint64_t)((double)m / (kLeaf * kTargetDensity)) + 1;
  return nl;
}

// (Re)builds the leaves from a sorted unique flat key array; also refreshes
// the CSR view and the degrees.  flat must hold m keys.
static void cpma_build(byo_graph *g, const uint64_t *flat, uint64_t m) {
  Cpma *c = state(g);
  hipStream_t s = g->stream;
  const uint64_t nl = leaves_for(m);
  if (c->nleaves != nl || !c->lo.p) {
    c->hi.alloc(nl * kLeaf * sizeof(uint32_t));
    c->lo.alloc(nl * kLeaf * sizeof(uint32_t));
    c->cnt.alloc(nl * sizeof(uint32_t));
    c->sep.alloc((nl + 1) * sizeof(uint64_t));
    c->nleaves = nl;
  }
  k_spread<<<byo_grid(nl * 64, 256, 16384), 256, 0, s>>>(flat, m, nl, leaves(c), c->cnt.as<uint32_t>(),
                                                         c->sep.as<uint64_t>());
  BYO_LAUNCH_CHECK();
  c->total = m;
  g->m = m;
  // CSR view and degrees straight from the flat array
  g->off.ensure((g->n + 1) * sizeof(uint64_t));
  g->ids.ensure(m * sizeof(uint32_t) + 16);
  prim::keys_to_csr(flat, m, g->n, g->off.as<uint64_t>(), g->ids.as<uint32_t>(), s);
  c->deg.ensure(g->n * sizeof(uint32_t) + 64);
  if (g->n)
    k_degrees_from_off<<<byo_grid(g->n, 256, 16384), 256, 0, s>>>(g->off.as<uint64_t>(), g->n, c->deg.as<uint32_t>());
  BYO_LAUNCH_CHECK();
  cpma_changed(g);
  g->csr_valid = true;
}

// Compacts the leaves into flat[total] (sorted, unique).
static void cpma_flatten(byo_graph *g, uint64_t *flat) {
  Cpma *c = state(g);
  hipStream_t s = g->stream;
  uint64_t *cnt64 = g->buf("cpma_cnt64", (c->nleaves + 1) * 8).as<uint64_t>();
  uint64_t *pos = g->buf("cpma_pos", (c->nleaves + 2) * 8).as<uint64_t>();
  void *tmp = g->buf("cpma_scan_tmp", prim::scan_tmp_bytes(c->nleaves + 1)).p;
  // widen counts to uint64 for the scan
  k_widen<<<byo_grid(c->nleaves, 256, 16384), 256, 0, s>>>(c->cnt.as<uint32_t>(), cnt64, c->nleaves);
  prim::exclusive_scan_u64(cnt64, pos, c->nleaves, true, tmp, s);
  k_compact<<<byo_grid(c->nleaves * 64, 256, 16384), 256, 0, s>>>(leaves(c), c->cnt.as<uint32_t>(), pos,
                                                                  c->nleaves, flat);
  BYO_LAUNCH_CHECK();
}

// The CSR export of the CPMA (byo_export_csr, byo_neighbors and the
// consumers that read a CSR); timed as profile phase "cpma_csr_export".
void cpma_ensure_csr(byo_graph *g) {
  Cpma *c = state(g);
  hipStream_t s = g->stream;
  PhaseTimer t(s, g->prof.on);
  uint64_t *flat = g->buf("cpma_flat", c->total * 8 + 64).as<uint64_t>();
  cpma_flatten(g, flat);
  g->off.ensure((g->n + 1) * sizeof(uint64_t));
  g->ids.ensure(c->total * sizeof(uint32_t) + 16);
  prim::keys_to_csr(flat, c->total, g->n, g->off.as<uint64_t>(), g->ids.as<uint32_t>(), s);
  BYO_HIP(hipStreamSynchronize(s));
  if (g->prof.on) g->account("cpma_csr_export", t.stop());
  g->csr_valid = true;
  g->maxdeg_valid = false;
  g->pieces_stale();
  g->rl_valid = false;
  g->xs_valid = false;
  g->pb_valid = false;
}

namespace {
// vstart[v] = slot of the key of rank rank[v] (its first key; the next
// vertex's first key when v has none); ranks at or past the total map to
// the end of the array.  pc[l] = keys before leaf l (pc[nleaves] = total).
__global__ void k_vstart(const uint64_t *rank, uint64_t n, const uint64_t *pc, uint64_t nleaves, uint64_t total,
                         uint64_t *vstart) {
  for (uint64_t v = global_tid(); v <= n; v += grid_threads()) {
    const uint64_t r = rank[v];
    if (r >= total) {
      vstart[v] = nleaves * kLeaf;
      continue;
    }
    uint64_t lo = 0, hi = nleaves - 1;  // first leaf l with pc[l + 1] > r
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (pc[mid + 1] > r) hi = mid; else lo = mid + 1;
    }
    vstart[v] = lo * kLeaf + (r - pc[lo]);
  }
}
}  // namespace

AdjView cpma_view(byo_graph *g) {
  Cpma *c = state(g);
  if (!c->index_valid) {
    hipStream_t s = g->stream;
    PhaseTimer t(s, g->prof.on);
    const uint64_t n = g->n, nl = c->nleaves;
    c->vstart.ensure((n + 1) * 8 + 64);
    uint64_t *w = g->buf("pidx_w", (std::max(n, nl) + 1) * 8 + 64).as<uint64_t>();
    uint64_t *rank = g->buf("pidx_rank", (n + 2) * 8 + 64).as<uint64_t>();
    uint64_t *pc = g->buf("pidx_pc", (nl + 2) * 8 + 64).as<uint64_t>();
    void *tmp = g->buf("pidx_tmp", prim::scan_tmp_bytes(std::max(n, nl) + 1)).p;
    // ranks: exclusive prefix of the degrees; leaf starts: of the counts
    if (n) {
      k_widen<<<byo_grid(n, 256, 16384), 256, 0, s>>>(c->deg.as<uint32_t>(), w, n);
      prim::exclusive_scan_u64(w, rank, n, true, tmp, s);
    } else {
      prim::fill_u64(rank, 0, 1, s);
    }
    k_widen<<<byo_grid(nl, 256, 16384), 256, 0, s>>>(c->cnt.as<uint32_t>(), w, nl);
    prim::exclusive_scan_u64(w, pc, nl, true, tmp, s);
    k_vstart<<<byo_grid(n + 1, 256, 16384), 256, 0, s>>>(rank, n, pc, nl, c->total, c->vstart.as<uint64_t>());
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipStreamSynchronize(s));
    if (g->prof.on) g->account("cpma_index", t.stop());
    c->index_valid = true;
  }
  return AdjView{c->vstart.as<uint64_t>(), c->lo.as<uint32_t>(), c->deg.as<uint32_t>(), c->nleaves * kLeaf};
}

const uint32_t *cpma_degrees(byo_graph *g) { return state(g)->deg.as<uint32_t>(); }

void cpma_destroy(byo_graph *g) {
  delete state(g);
  g->cpma = nullptr;
}

uint64_t cpma_bytes(byo_graph *g) {
  Cpma *c = state(g);
  // leaves + counts + separators + degrees + slot starts (get_memory_size
  // counts the PMA and two per-vertex arrays, run_single_pma.cc:126-131)
  return c->nleaves * kLeaf * 8 + c->nleaves * 4 + (c->nleaves + 1) * 8 + g->n * 4 + (g->n + 1) * 8;
}

namespace {
__global__ void k_zero_ctl(UpdCtl *ctl, uint64_t count) {
  for (uint64_t i = global_tid(); i < count; i++) ctl[i] = UpdCtl{0, 0, 0, {0, 0, 0, 0, 0}};
}
__global__ void k_zero_u32(uint32_t *p) { *p = 0; }
// Stops a captured sequence at batch `index` (1-based) when its keys hold an
// id >= n or the array is too dense for a windowed rebalance.
__global__ void k_gate(const UpdCtl *ctl, uint32_t *halt, uint32_t index) {
  if (*halt == 0 && (ctl->pad[2] || ctl->bad)) *halt = index;
}
}  // namespace

// Buffers of one update for batches of up to cap keys.
struct UpdBufs {
  uint64_t *k0, *k1;
  void *sort_tmp, *sel_tmp, *scan_tmp;
  uint32_t *leaf, *flag, *pos, *starts, *newc;
  uint8_t *done;  // per segment: merged by the one-pass merge
};

static UpdBufs upd_bufs(byo_graph *g, uint64_t cap) {
  UpdBufs u;
  u.k0 = g->buf("upd_k0", cap * 8 + 64).as<uint64_t>();
  u.k1 = g->buf("upd_k1", cap * 8 + 64).as<uint64_t>();
  u.sort_tmp = g->buf("upd_sort_tmp", prim::radix_tmp_bytes(cap)).p;
  u.sel_tmp = g->buf("upd_sel_tmp", prim::select_tmp_bytes(cap) + 256).p;
  u.scan_tmp = g->buf("upd_scan_tmp", prim::scan_tmp_bytes(cap + 1)).p;
  u.leaf = g->buf("upd_leaf", cap * 4 + 64).as<uint32_t>();
  u.flag = g->buf("upd_flag", cap * 4 + 64).as<uint32_t>();
  u.pos = g->buf("upd_pos", (cap + 1) * 4 + 64).as<uint32_t>();
  u.starts = g->buf("upd_starts", cap * 4 + 64).as<uint32_t>();
  u.newc = g->buf("upd_newc", cap * 4 + 64).as<uint32_t>();
  u.done = g->buf("upd_done", cap + 64).as<uint8_t>();
  return u;
}

// Device-only part of an update (no host round trip, so it can be captured):
// ids checked (ctl->bad), keys_in[b] radix sorted and deduplicated (count in
// ctl->pad[0]), each unique key's leaf found and the batch cut into one
// segment per touched leaf (count at u.pos[b]).  ctl must be zero.
static const uint64_t *sort_and_segment(byo_graph *g, const uint64_t *keys_in, uint64_t b, UpdCtl *ctl,
                                        const UpdBufs &u) {
  Cpma *c = state(g);
  hipStream_t s = g->stream;
  if (b <= kSmallBatch) {
    // u.pos + b holds the segment count, as after the general path's scan
    k_small_batch<<<1, kSmallThreads, 0, s>>>(keys_in, b, g->n, c->sep.as<uint64_t>(), c->nleaves, u.k1, u.leaf,
                                              u.starts, u.pos + b, ctl);
    BYO_LAUNCH_CHECK();
    return u.k1;
  }
  k_check_ids<<<byo_grid(b, 256, 8192), 256, 0, s>>>(keys_in, b, g->n, ctl);
  BYO_HIP(hipMemcpyAsync(u.k0, keys_in, b * 8, hipMemcpyDeviceToDevice, s));
  int idbits = 1;
  while ((1ull << idbits) < g->n) idbits++;
  int shifts[16];
  const int np = prim::edge_key_shifts(idbits, shifts);
  uint64_t *sorted = prim::radix_sort_u64(u.k0, u.k1, b, shifts, np, u.sort_tmp, s);
  uint64_t *uniq = (sorted == u.k0) ? u.k1 : u.k0;
  const unsigned long long *bu = &ctl->pad[0];
  prim::unique_u64(sorted, uniq, b, kEmpty, reinterpret_cast<uint64_t *>(&ctl->pad[0]), u.sel_tmp, s);
  const unsigned grid = byo_grid(b, 256, 8192);
  k_leaf_of<<<grid, 256, 0, s>>>(uniq, bu, c->sep.as<uint64_t>(), c->nleaves, u.leaf);
  k_seg_flags<<<grid, 256, 0, s>>>(u.leaf, bu, b, u.flag);
  prim::exclusive_scan_u32(u.flag, u.pos, b, true, u.scan_tmp, s);
  k_seg_starts<<<grid, 256, 0, s>>>(u.flag, u.pos, b, u.starts);
  BYO_LAUNCH_CHECK();
  return uniq;
}

// Buffers and arguments of the windowed rebalance for batches of up to cap
// keys (the staging area is as large as the leaves).
static WinArgs win_args(byo_graph *g, const uint64_t *batch, const UpdDev &d, const UpdBufs &u, UpdCtl *ctl,
                        uint64_t cap) {
  Cpma *c = state(g);
  const uint64_t nl = c->nleaves;
  WinArgs w{};
  w.keys = batch;
  w.d = d;
  w.starts = u.starts;
  w.leaf_of = u.leaf;
  w.newc = u.newc;
  w.lv = leaves(c);
  w.cnt = c->cnt.as<uint32_t>();
  w.sep = c->sep.as<uint64_t>();
  w.deg = c->deg.as<uint32_t>();
  w.nleaves = nl;
  w.dem = g->buf("win_dem", (nl + 1) * 8 + 64).as<uint64_t>();
  w.P = g->buf("win_prefix", (nl + 2) * 8 + 64).as<uint64_t>();
  w.cand = g->buf("win_cand", nl * 4 + 64).as<uint32_t>();
  w.roots = g->buf("win_roots", nl * 8 + 64).as<uint64_t>();
  w.stage = g->buf("win_stage", nl * kLeaf * 8 + 64).as<uint64_t>();
  w.stage_new = g->buf("win_stage_new", cap * 8 + 64).as<uint64_t>();
  w.ctl = ctl;
  w.cnt64 = g->buf("win_cnt64", (nl + 1) * 8 + 64).as<uint64_t>();
  w.pc = g->buf("win_pc", (nl + 2) * 8 + 64).as<uint64_t>();
  w.nflag = g->buf("win_nflag", (cap + 1) * 4 + 64).as<uint32_t>();
  w.npos = g->buf("win_npos", (cap + 2) * 4 + 64).as<uint32_t>();
  w.rinfo = g->buf("win_rinfo", nl * 32 + 64).as<uint64_t>();
  w.rsize = g->buf("win_rsize", (nl + 1) * 8 + 64).as<uint64_t>();
  w.rpre = g->buf("win_rpre", (nl + 2) * 8 + 64).as<uint64_t>();
  w.cap = cap;
  g->buf("win_scan_tmp", prim::scan_tmp_bytes(nl + 1));
  g->buf("win_scan_tmp2", prim::scan_tmp_bytes(cap + 1));
  return w;
}

// Candidate windows and their roots (then, in window_rebuild, the rebuilt
// windows); every kernel returns at once unless a leaf overflowed.
static void window_pick(byo_graph *g, const WinArgs &w, uint64_t cap) {
  hipStream_t s = g->stream;
  const unsigned gl = byo_grid(w.nleaves, 256, 16384), gs = byo_grid(cap, 256, 16384);
  k_win_dem<<<gl, 256, 0, s>>>(w);
  k_win_add<<<gs, 256, 0, s>>>(w);
  prim::exclusive_scan_u64(w.dem, w.P, w.nleaves, true, g->scratch["win_scan_tmp"].p, s);
  k_win_pick<<<gs, 256, 0, s>>>(w);
  k_win_roots<<<gl, 256, 0, s>>>(w);
  BYO_LAUNCH_CHECK();
}

static void window_rebuild(byo_graph *g, const WinArgs &w) {
  hipStream_t s = g->stream;
  void *t1 = g->scratch["win_scan_tmp"].p, *t2 = g->scratch["win_scan_tmp2"].p;
  const unsigned gl = byo_grid(w.nleaves, 256, 16384), gw = byo_grid(w.nleaves * 64, 256, 16384),
                 gc = byo_grid(w.cap, 256, 16384);
  k_win_w1<<<gl, 256, 0, s>>>(w);
  prim::exclusive_scan_u64(w.cnt64, w.pc, w.nleaves, true, t1, s);
  k_win_w2<<<gw, 256, 0, s>>>(w);
  k_win_w3<<<gc, 256, 0, s>>>(w);
  prim::exclusive_scan_u32(w.nflag, w.npos, w.cap, true, t2, s);
  k_win_w4<<<gc, 256, 0, s>>>(w);
  k_win_w5<<<gl, 256, 0, s>>>(w);
  prim::exclusive_scan_u64(w.rsize, w.rpre, w.nleaves, true, t1, s);
  k_win_w6<<<4096, 256, 0, s>>>(w);
  k_win_w7<<<gw, 256, 0, s>>>(w);
  BYO_LAUNCH_CHECK();
}

static void cpma_update(byo_graph *g, const uint64_t *keys_dev, uint64_t b, bool insert) {
  Cpma *c = state(g);
  hipStream_t s = g->stream;
  if (b == 0) return;
  UpdCtl *ctl = g->buf("upd_ctl", sizeof(UpdCtl)).as<UpdCtl>();
  UpdCtl *hctl = static_cast<UpdCtl *>(pinned(g, sizeof(UpdCtl)));
  PhaseTimer t(s, g->prof.on);
  const UpdBufs u = upd_bufs(g, b);
  k_zero_ctl<<<1, 1, 0, s>>>(ctl, 1);
  const uint64_t *batch = sort_and_segment(g, keys_dev, b, ctl, u);
  // counts stay on the device; the first read-back also reports a bad id
  // (no kernel writes the container when ctl->bad is set)
  const UpdDev d{&ctl->pad[0], u.pos + b, nullptr};
  const unsigned grid = (unsigned)std::min<uint64_t>(b, kUpdGrid);
  if (insert) {
    // one pass for the segments whose leaf has room for all their keys; the
    // count (and, on an overflow, the windows) and the merge then cover the
    // rest (the first pass writes nothing when an id is bad)
    k_ins_merge<<<grid, 256, 0, s>>>(batch, d, u.starts, u.leaf, leaves(c), c->cnt.as<uint32_t>(),
                                     c->deg.as<uint32_t>(), ctl, nullptr, c->nleaves, u.done, true);
    k_ins_count<<<grid, 256, 0, s>>>(batch, d, u.starts, u.leaf, leaves(c), c->cnt.as<uint32_t>(),
                                     u.newc, ctl, u.done);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(hctl, ctl, sizeof(UpdCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(hctl->bad == 0, BYO_E_INVAL, "batch holds a vertex id >= n");
    const uint64_t bu = hctl->pad[0];
    if (bu == 0) return;
    const uint32_t *cand = nullptr;
    if (hctl->overflow) {
      // leaves that would overflow: rebuild their windows on the device
      const WinArgs w = win_args(g, batch, d, u, ctl, b);
      window_pick(g, w, b);
      window_rebuild(g, w);
      cand = w.cand;
    }
    k_ins_merge<<<grid, 256, 0, s>>>(batch, d, u.starts, u.leaf, leaves(c), c->cnt.as<uint32_t>(),
                                     c->deg.as<uint32_t>(), ctl, cand, c->nleaves, u.done, false);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(hctl, ctl, sizeof(UpdCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    if (!hctl->pad[2]) {
      if (hctl->overflow) c->rebalances++;
      c->touched += hctl->pad[4];
      c->total += hctl->changed;
      g->m = c->total;
      if (hctl->changed) cpma_changed(g);
    } else {
      // global rebalance: flatten, merge with the batch, dedupe, respread
      // (the leaves hold the one-pass merge's keys already: counted first)
      c->rebalances++;
      c->total += hctl->changed;
      c->touched += hctl->pad[4];
      const uint64_t tot = c->total + bu;
      uint64_t *flat = g->buf("cpma_flat", c->total * 8 + 64).as<uint64_t>();
      uint64_t *merged = g->buf("cpma_merged", tot * 8 + 64).as<uint64_t>();
      uint64_t *uniq = g->buf("cpma_uniq", tot * 8 + 64).as<uint64_t>();
      cpma_flatten(g, flat);
      k_merge<<<byo_grid((tot + 7) / 8, 256, 16384), 256, 0, s>>>(flat, c->total, batch, bu, merged);
      BYO_LAUNCH_CHECK();
      void *stmp = g->buf("cpma_sel_tmp", prim::select_tmp_bytes(tot) + 256).p;
      prim::unique_u64(merged, uniq, tot, kEmpty, reinterpret_cast<uint64_t *>(&ctl->pad[1]), stmp, s);
      BYO_HIP(hipMemcpyAsync(hctl, ctl, sizeof(UpdCtl), hipMemcpyDeviceToHost, s));
      BYO_HIP(hipStreamSynchronize(s));
      const uint64_t m2 = hctl->pad[1];
      c->touched += 8 * m2;  // the whole array is rewritten
      cpma_build(g, uniq, m2);
      BYO_HIP(hipStreamSynchronize(s));
    }
  } else {
    k_remove<<<grid, 256, 0, s>>>(batch, d, u.starts, u.leaf, leaves(c), c->cnt.as<uint32_t>(),
                                  c->deg.as<uint32_t>(), ctl);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(hctl, ctl, sizeof(UpdCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(hctl->bad == 0, BYO_E_INVAL, "batch holds a vertex id >= n");
    c->touched += hctl->pad[4];
    c->total -= hctl->changed;
    g->m = c->total;
    if (hctl->changed) {
      cpma_changed(g);
    }
    if (c->nleaves > 1 && (double)c->total < kMinDensity * (double)(c->nleaves * kLeaf)) {
      c->rebalances++;
      c->touched += 8 * c->total;
      uint64_t *flat = g->buf("cpma_flat", c->total * 8 + 64).as<uint64_t>();
      cpma_flatten(g, flat);
      cpma_build(g, flat, c->total);
      BYO_HIP(hipStreamSynchronize(s));
    }
  }
  if (g->prof.on) g->account(insert ? "cpma_insert" : "cpma_remove", t.stop());
}

// Insert batches [first, nb) as one captured HIP graph; returns the 1-based
// index of the batch that stopped the sequence (0: all applied) and adds the
// graph's launch time to *ms.
static uint32_t insert_graph(byo_graph *g, const uint64_t *keys, const uint64_t *offs, uint32_t first, uint32_t nb,
                             const UpdBufs &u, UpdCtl *ctls, uint32_t *halt, double *ms) {
  Cpma *c = state(g);
  hipStream_t s = g->stream;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  uint64_t cap = 0;
  for (uint32_t i = first; i < nb; i++) cap = std::max<uint64_t>(cap, offs[i + 1] - offs[i]);
  // the windowed-rebalance buffers for the current leaf count exist before
  // capture (nothing is allocated while capturing)
  (void)win_args(g, nullptr, UpdDev{}, u, ctls, cap);
  BYO_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeRelaxed));
  try {
    k_zero_u32<<<1, 1, 0, s>>>(halt);
    k_zero_ctl<<<1, 1, 0, s>>>(ctls + first, nb - first);
    for (uint32_t i = first; i < nb; i++) {
      const uint64_t b = offs[i + 1] - offs[i];
      if (b == 0) continue;
      UpdCtl *ctl = ctls + i;
      const uint64_t *batch = sort_and_segment(g, keys + offs[i], b, ctl, u);
      const UpdDev d{&ctl->pad[0], u.pos + b, halt};
      const unsigned grid = (unsigned)std::min<uint64_t>(b, kUpdGrid);
      k_ins_count<<<grid, 256, 0, s>>>(batch, d, u.starts, u.leaf, leaves(c), c->cnt.as<uint32_t>(),
                                       u.newc, ctl);
      const WinArgs w = win_args(g, batch, d, u, ctl, cap);
      window_pick(g, w, b);
      k_gate<<<1, 1, 0, s>>>(ctl, halt, i + 1);
      window_rebuild(g, w);
      k_ins_merge<<<grid, 256, 0, s>>>(batch, d, u.starts, u.leaf, leaves(c), c->cnt.as<uint32_t>(),
                                       c->deg.as<uint32_t>(), ctl, w.cand, c->nleaves);
    }
    BYO_LAUNCH_CHECK();
  } catch (...) {
    (void)hipStreamEndCapture(s, &graph);
    if (graph) (void)hipGraphDestroy(graph);
    throw;
  }
  BYO_HIP(hipStreamEndCapture(s, &graph));
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(graph);
    BYO_HIP(ie);
  }
  double t_ms = 0;
  {
    PhaseTimer t(s, true);
    const hipError_t le = hipGraphLaunch(exec, s);
    if (le != hipSuccess) {
      (void)hipGraphExecDestroy(exec);
      (void)hipGraphDestroy(graph);
      BYO_HIP(le);
    }
    t_ms = t.stop();
  }
  (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  *ms += t_ms;
  // read back the stop index and the per-batch counts
  const size_t nbytes = (size_t)(nb - first) * sizeof(UpdCtl);
  char *h = static_cast<char *>(pinned(g, 64 + nbytes));
  BYO_HIP(hipMemcpyAsync(h, halt, 4, hipMemcpyDeviceToHost, s));
  BYO_HIP(hipMemcpyAsync(h + 64, ctls + first, nbytes, hipMemcpyDeviceToHost, s));
  BYO_HIP(hipStreamSynchronize(s));
  const uint32_t stop = *reinterpret_cast<uint32_t *>(h);
  const UpdCtl *hc = reinterpret_cast<const UpdCtl *>(h + 64);
  const uint32_t applied_end = stop ? stop - 1 : nb;
  uint64_t changed = 0;
  for (uint32_t i = first; i < applied_end; i++) {
    changed += hc[i - first].changed;
    c->touched += hc[i - first].pad[4];
  }
  c->total += changed;
  g->m = c->total;
  if (changed) {
    cpma_changed(g);
  }
  return stop;
}

}  // namespace byo

using namespace byo;

namespace {
byo_graph *new_cpma_handle(uint64_t n) {
  byo_graph *g = new_handle(1);
  g->n = n;
  g->cpma = new Cpma();
  return g;
}
}  // namespace

extern "C" {

int byo_cpma_create(const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m, byo_graph_t *out) {
  return guard([&] {
    BYO_REQUIRE(out != nullptr, BYO_E_INVAL, "out is null");
    check_csr_input(offsets, ids, n, m);
    for (uint64_t v = 0; v < n; v++)
      for (uint64_t e = offsets[v] + 1; e < offsets[v + 1]; e++)
        BYO_REQUIRE(ids[e - 1] < ids[e], BYO_E_INVAL, "adjacency lists must be sorted without duplicates");
    byo_graph *g = new_cpma_handle(n);
    try {
      csr_upload(g, offsets, ids, n, m);
      uint64_t *flat = g->buf("cpma_flat", m * 8 + 64).as<uint64_t>();
      prim::csr_to_keys(g->off.as<uint64_t>(), g->ids.as<uint32_t>(), n, m, flat, g->stream);
      cpma_build(g, flat, m);
      BYO_HIP(hipStreamSynchronize(g->stream));
    } catch (...) {
      byo_graph_destroy(g);
      throw;
    }
    *out = g;
  });
}

int byo_cpma_create_from(byo_graph_t src, byo_graph_t *out) {
  return guard([&] {
    BYO_REQUIRE(src && out, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(src->mu);
    bind(src);
    ensure_csr(src);
    byo_graph *g = new_cpma_handle(src->n);
    try {
      const uint64_t m = src->m;
      uint64_t *flat = g->buf("cpma_flat", m * 8 + 64).as<uint64_t>();
      BYO_HIP(hipStreamSynchronize(src->stream));
      prim::csr_to_keys(src->off.as<uint64_t>(), src->ids.as<uint32_t>(), src->n, m, flat, g->stream);
      cpma_build(g, flat, m);
      BYO_HIP(hipStreamSynchronize(g->stream));
    } catch (...) {
      byo_graph_destroy(g);
      throw;
    }
    *out = g;
  });
}

int byo_insert_batches_dev(byo_graph_t g, const uint64_t *keys_dev, const uint64_t *offsets, uint32_t nbatches,
                           uint32_t *graphs, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && offsets && (keys_dev || nbatches == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(g->kind == 1, BYO_E_STATE, "insert_batch needs a CPMA container");
    uint64_t cap = 0;
    for (uint32_t i = 0; i < nbatches; i++) {
      BYO_REQUIRE(offsets[i + 1] >= offsets[i], BYO_E_INVAL, "batch offsets must not decrease");
      cap = std::max<uint64_t>(cap, offsets[i + 1] - offsets[i]);
    }
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    if (graphs) *graphs = 0;
    if (device_ms) *device_ms = 0;
    if (cap == 0) return;
    // every buffer the captured kernels touch exists before capture starts
    const UpdBufs u = upd_bufs(g, cap);
    UpdCtl *ctls = g->buf("upd_ctls", (size_t)nbatches * sizeof(UpdCtl) + 64).as<UpdCtl>();
    uint32_t *halt = g->buf("upd_halt", 64).as<uint32_t>();
    pinned(g, 64 + (size_t)nbatches * sizeof(UpdCtl));
    double ms = 0;
    uint32_t first = 0, launches = 0;
    while (first < nbatches) {
      const uint32_t stop = insert_graph(g, keys_dev, offsets, first, nbatches, u, ctls, halt, &ms);
      launches++;
      if (stop == 0) break;
      // the stopping batch goes through the host-driven path (global
      // rebalance, or the error for an id >= n); the rest follows in a new graph
      const uint32_t h = stop - 1;
      PhaseTimer t(g->stream, true);
      cpma_update(g, keys_dev + offsets[h], offsets[h + 1] - offsets[h], true);
      ms += t.stop();
      first = h + 1;
    }
    if (graphs) *graphs = launches;
    if (device_ms) *device_ms = ms;
    if (g->prof.on) g->account("cpma_insert_graph", ms);
  });
}

int byo_insert_batch_dev(byo_graph_t g, const uint64_t *keys_dev, uint64_t b) {
  return guard([&] {
    BYO_REQUIRE(g && (keys_dev || b == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(g->kind == 1, BYO_E_STATE, "insert_batch needs a CPMA container");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    cpma_update(g, keys_dev, b, true);
  });
}

int byo_remove_batch_dev(byo_graph_t g, const uint64_t *keys_dev, uint64_t b) {
  return guard([&] {
    BYO_REQUIRE(g && (keys_dev || b == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(g->kind == 1, BYO_E_STATE, "remove_batch needs a CPMA container");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    cpma_update(g, keys_dev, b, false);
  });
}

static int host_batch(byo_graph_t g, const uint64_t *keys, uint64_t b, bool insert) {
  return guard([&] {
    BYO_REQUIRE(g && (keys || b == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(g->kind == 1, BYO_E_STATE, "batch updates need a CPMA container");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    if (b == 0) return;
    uint64_t *d = g->buf("upd_host_in", b * 8 + 64).as<uint64_t>();
    BYO_HIP(hipMemcpyAsync(d, keys, b * 8, hipMemcpyHostToDevice, g->stream));
    cpma_update(g, d, b, insert);
  });
}

int byo_cpma_stats(byo_graph_t g, uint64_t *touched_bytes, uint64_t *rebalances, uint64_t *leaves) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    BYO_REQUIRE(g->kind == 1, BYO_E_STATE, "CPMA statistics need a CPMA container");
    std::lock_guard<std::mutex> lk(g->mu);
    const Cpma *c = state(g);
    if (touched_bytes) *touched_bytes = c->touched;
    if (rebalances) *rebalances = c->rebalances;
    if (leaves) *leaves = c->nleaves;
  });
}

int byo_insert_batch(byo_graph_t g, const uint64_t *keys, uint64_t b) { return host_batch(g, keys, b, true); }
int byo_remove_batch(byo_graph_t g, const uint64_t *keys, uint64_t b) { return host_batch(g, keys, b, false); }

}  // extern "C"

// ---------------------------------------------- allocation-free small insert
// byo_insert_batch_dev_noalloc: the small-batch insert path (one workgroup
// sorts, dedupes and segments; one pass counts; one pass merges in place)
// with every buffer beyond the container and the caller's keys carved from
// one file-scope device array.  Each GPU holds its own copy of the array
// until the process ends; all calls on a GPU share it (the last call's
// contents stay), so calls that overlap on one GPU race.
namespace byo {
namespace {
constexpr uint64_t kNaMaxBatch = kSmallBatch;
// layout: ctl | uniq u64[max] | leaf_of u32[max] | starts u32[max] |
// newc u32[max] | segment count u32
constexpr size_t kNaCtl = 0;
constexpr size_t kNaUniq = 256;
constexpr size_t kNaLeaf = kNaUniq + kNaMaxBatch * 8;
constexpr size_t kNaStarts = kNaLeaf + kNaMaxBatch * 4;
constexpr size_t kNaNewc = kNaStarts + kNaMaxBatch * 4;
constexpr size_t kNaNseg = kNaNewc + kNaMaxBatch * 4;
constexpr size_t kNaBytes = kNaNseg + 256;  // 86,528 bytes
static_assert(sizeof(UpdCtl) <= kNaUniq - kNaCtl, "control block fits its slot");
}  // namespace
}  // namespace byo

static __device__ __align__(256) unsigned char scratch[byo::kNaBytes];

namespace byo {
namespace {
constexpr int kNaMaxDevices = 64;
std::mutex na_mu;
unsigned char *na_ptr[kNaMaxDevices] = {};

// Device address of this GPU's copy of `scratch`, looked up once per device
// with that device current (the symbol's host-side name is not a device
// address and is never passed to a kernel or copy).
unsigned char *noalloc_scratch(int device) {
  BYO_REQUIRE(device >= 0 && device < kNaMaxDevices, BYO_E_INVAL, "device index out of range");
  std::lock_guard<std::mutex> lk(na_mu);
  if (!na_ptr[device]) {
    void *p = nullptr;
    BYO_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    BYO_REQUIRE(p != nullptr, BYO_E_HIP, "hipGetSymbolAddress returned no address for the scratch array");
    na_ptr[device] = static_cast<unsigned char *>(p);
  }
  return na_ptr[device];
}
}  // namespace
}  // namespace byo

extern "C" int byo_insert_batch_dev_noalloc(byo_graph_t g, const uint64_t *keys_dev, uint64_t b) {
  return guard([&] {
    BYO_REQUIRE(g && (keys_dev || b == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(g->kind == 1, BYO_E_STATE, "insert_batch needs a CPMA container");
    BYO_REQUIRE(b <= kNaMaxBatch, BYO_E_INVAL,
                "byo_insert_batch_dev_noalloc takes at most 4096 keys per batch");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    if (b == 0) return;
    unsigned char *base = noalloc_scratch(g->device);
    Cpma *c = state(g);
    hipStream_t s = g->stream;
    UpdCtl *ctl = reinterpret_cast<UpdCtl *>(base + kNaCtl);
    uint64_t *uniq = reinterpret_cast<uint64_t *>(base + kNaUniq);
    uint32_t *leaf = reinterpret_cast<uint32_t *>(base + kNaLeaf);
    uint32_t *starts = reinterpret_cast<uint32_t *>(base + kNaStarts);
    uint32_t *newc = reinterpret_cast<uint32_t *>(base + kNaNewc);
    uint32_t *nseg = reinterpret_cast<uint32_t *>(base + kNaNseg);
    // the control block keeps the previous call's counts: clear it first
    k_zero_ctl<<<1, 1, 0, s>>>(ctl, 1);
    k_small_batch<<<1, kSmallThreads, 0, s>>>(keys_dev, b, g->n, c->sep.as<uint64_t>(), c->nleaves, uniq, leaf,
                                              starts, nseg, ctl);
    const UpdDev d{&ctl->pad[0], nseg, nullptr};
    const unsigned grid = (unsigned)std::min<uint64_t>(b, kUpdGrid);
    k_ins_count<<<grid, 256, 0, s>>>(uniq, d, starts, leaf, leaves(c), c->cnt.as<uint32_t>(), newc,
                                     ctl);
    BYO_LAUNCH_CHECK();
    UpdCtl h{};
    BYO_HIP(hipMemcpyAsync(&h, ctl, sizeof(UpdCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(h.bad == 0, BYO_E_INVAL, "batch holds a vertex id >= n");
    if (h.pad[0] == 0) return;
    // a leaf that would overflow needs a rebalance, whose buffers scale with
    // the container: refuse, with the container unchanged
    BYO_REQUIRE(h.overflow == 0, BYO_E_STATE,
                "a leaf would overflow: this batch needs a rebalance, which allocates (use byo_insert_batch_dev)");
    k_ins_merge<<<grid, 256, 0, s>>>(uniq, d, starts, leaf, leaves(c), c->cnt.as<uint32_t>(),
                                     c->deg.as<uint32_t>(), ctl, nullptr, c->nleaves);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(&h, ctl, sizeof(UpdCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    c->total += h.changed;
    c->touched += h.pad[4];
    g->m = c->total;
    if (h.changed) {
      cpma_changed(g);
    }
  });
}
