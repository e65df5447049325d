// pagerank.hip -- PageRank pull (EdgeMap::edgeMapReduce_dense,
// gbbs/edge_map_reduce.h:458-519, driven by benchmarks/PageRank/
// PageRank.h:125-193) with the Jacobi semantics of PageRank_edgeMap
// (PageRank.h:80-123), and batched personalized PageRank.
//
// Two ways to an iteration.  The pull: FP32 ranks and contributions p/deg,
// each vertex summing its in-neighbour contributions with Kahan
// compensation per lane and a TwoSum-compensated combine across the lanes
// of its group (FP64: double values, no compensation); one kernel reads
// offsets (8 B/vertex), ids (4 B/edge), one gathered contribution per edge
// (w B), the old rank (w B/vertex), and writes the new rank and contribution
// (2w B/vertex).  Propagation blocking (below, knob pr_pb): contributions in
// FP32 or FP64 moved through destination bins in two streaming passes and
// summed exactly in 64-bit fixed point.  Either way the global L1 residual is
// reduced in FP64 in a fixed order (deterministic).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "kahan.hpp"
#include "primitives.hpp"

namespace byo {
namespace {

// One pass of the pull over the neighbours of every vertex whose ids lie in
// one source segment [lo_pos[v], hi_pos[v]).  Adjacency lists are sorted,
// so a segment is a contiguous sub-range of each list and the gathered
// contributions of one pass stay inside one window of the rank vector.  A
// wave takes a tile of 64 consecutive vertices: per-vertex state (segment
// bounds, the (sum, compensation) carried between passes, ranks) moves with
// coalesced lane-per-vertex loads and stores, while the gathers run G lanes
// per vertex over G rounds of 64/G vertices.  The last pass finalises
// p' = d * sum + (1-d)/n, the next contributions p'/deg and the block's share
// of the FP64 L1 residual.
template <class T, int G, bool kComp, bool kFirst, bool kLast>
__global__ void __launch_bounds__(256) k_pr_seg(const uint64_t *lo_pos, const uint64_t *hi_pos, const uint32_t *ids,
                                                const uint64_t *off, uint64_t n, const T *contrib_in, T *acc_s,
                                                T *acc_c, const T *p_old, T *p_new, T *contrib_out, double damping,
                                                double addc, double *partial) {
  constexpr int VPR = 64 / G;  // vertices per round
  __shared__ double lds[4];
  const int lane = lane_id();
  const int sub = lane & (G - 1), grp = lane / G;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const uint64_t ntiles = (n + 63) >> 6;
  double resid = 0.0;
  const T d = (T)damping, t = (T)addc;
  for (uint64_t tile = wave; tile < ntiles; tile += nwaves) {
    const uint64_t v = tile * 64 + lane;
    const bool valid = v < n;
    const uint64_t lo = valid ? lo_pos[v] : 0, hi = valid ? hi_pos[v] : 0;
    Acc<T> mine{0, 0};
#pragma unroll 2
    for (int r = 0; r < G; r++) {
      const int owner = r * VPR + grp;
      const uint64_t b = __shfl(lo, owner, 64), e = __shfl(hi, owner, 64);
      Acc<T> a{0, 0};
      uint64_t k = b + sub;
      for (; k + G < e; k += 2 * G) {
        const T x0 = contrib_in[ids[k]];
        const T x1 = contrib_in[ids[k + G]];
        if (kComp) {
          kahan_add(a, x0);
          kahan_add(a, x1);
        } else {
          a.s += x0;
          a.s += x1;
        }
      }
      if (k < e) {
        const T x0 = contrib_in[ids[k]];
        if (kComp) kahan_add(a, x0); else a.s += x0;
      }
#pragma unroll
      for (int o = G / 2; o > 0; o >>= 1) {
        Acc<T> other{__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64)};
        if (kComp) a = combine(a, other); else a.s += other.s;
      }
      // the group leader of vertex r*VPR + i is lane i*G
      const int rel = lane - r * VPR;
      const int src = (rel >= 0 && rel < VPR) ? rel * G : 0;
      const T gs = __shfl(a.s, src, 64), gc = __shfl(a.c, src, 64);
      if (rel >= 0 && rel < VPR) mine = Acc<T>{gs, gc};
    }
    if (valid) {
      if (!kFirst) {
        if (kComp) mine = combine(Acc<T>{acc_s[v], acc_c[v]}, mine); else mine.s += acc_s[v];
      }
      if (!kLast) {
        acc_s[v] = mine.s;
        if (kComp) acc_c[v] = mine.c;
      } else {
        const T sum = kComp ? (mine.s - mine.c) : mine.s;
        const T pv = d * sum + t;
        const uint64_t deg = off[v + 1] - off[v];
        p_new[v] = pv;
        contrib_out[v] = deg ? pv / (T)deg : (T)0;
        resid += fabs((double)pv - (double)p_old[v]);
      }
    }
  }
  if (kLast) {
    resid = wave_sum(resid);
    if (lane == 0) lds[threadIdx.x >> 6] = resid;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
  }
}

// Pull over whole lists (one source segment) with the lanes per vertex
// picked per tile: a wave takes 64 consecutive vertices and gives each
// G = the smallest power of two with G * kPrUnroll >= the tile's largest
// degree (at most 64) lanes, so a tile of leaves runs one vertex per lane
// and a tile of hubs one vertex per wave.  Every lane keeps kPrUnroll
// gathers in flight: the ids of a step are loaded first (clamped to the
// vertex's first edge past its end, so the loads need no branch), then the
// contributions, then the Kahan adds of the ones in range.  Tiles of
// similar degrees -- vertices numbered by degree, as the relabelled copy
// is -- waste the fewest lanes.
constexpr int kPrUnroll = 16;

// H: the lists may hold empty slots (kNone, the PMA's leaf ends): skipped.
template <class T, int G, bool kComp, bool H, int U = kPrUnroll>
__device__ __forceinline__ Acc<T> pr_tile(uint64_t lo, uint64_t hi, const uint32_t *ids, const T *contrib_in,
                                          int lane) {
  constexpr int VPR = 64 / G;  // vertices per round
  const int sub = lane & (G - 1), grp = lane / G;
  Acc<T> mine{0, 0};
  for (int r = 0; r < G; r++) {
    const int owner = r * VPR + grp;
    const uint64_t b = G == 1 ? lo : __shfl(lo, owner, 64);
    const uint64_t e = G == 1 ? hi : __shfl(hi, owner, 64);
    Acc<T> a{0, 0};
    for (uint64_t k = b + sub; k < e; k += (uint64_t)U * G) {
      uint32_t u[U];
      T x[U];
      bool hole[U];
#pragma unroll
      for (int j = 0; j < U; j++) {
        const uint64_t kk = k + (uint64_t)j * G;
        u[j] = ids[kk < e ? kk : b];
        hole[j] = H && u[j] == kNone;
        if (H && hole[j]) u[j] = 0;
      }
#pragma unroll
      for (int j = 0; j < U; j++) x[j] = contrib_in[u[j]];
#pragma unroll
      for (int j = 0; j < U; j++) {
        if (k + (uint64_t)j * G < e && !hole[j]) {
          if (kComp) kahan_add(a, x[j]); else a.s += x[j];
        }
      }
    }
    if (G == 1) return a;
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      Acc<T> other{__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64)};
      if (kComp) a = combine(a, other); else a.s += other.s;
    }
    // the group leader of vertex r*VPR + i is lane i*G
    const int rel = lane - r * VPR;
    const int src = (rel >= 0 && rel < VPR) ? rel * G : 0;
    const T gs = __shfl(a.s, src, 64), gc = __shfl(a.c, src, 64);
    if (rel >= 0 && rel < VPR) mine = Acc<T>{gs, gc};
  }
  return mine;
}

// Lanes per vertex for a tile whose longest list is dmax: the smallest G
// with G * U >= dmax (a tile of short lists runs a vertex per lane).
template <class T, bool kComp, bool H, int U>
__device__ __forceinline__ Acc<T> pr_tile_any(uint32_t dmax, uint64_t lo, uint64_t hi, const uint32_t *ids,
                                              const T *contrib_in, int lane) {
  if (dmax <= U) return pr_tile<T, 1, kComp, H, U>(lo, hi, ids, contrib_in, lane);
  if (dmax <= 2 * U) return pr_tile<T, 2, kComp, H, U>(lo, hi, ids, contrib_in, lane);
  if (dmax <= 4 * U) return pr_tile<T, 4, kComp, H, U>(lo, hi, ids, contrib_in, lane);
  if (dmax <= 8 * U) return pr_tile<T, 8, kComp, H, U>(lo, hi, ids, contrib_in, lane);
  if (dmax <= 16 * U) return pr_tile<T, 16, kComp, H, U>(lo, hi, ids, contrib_in, lane);
  if (dmax <= 32 * U) return pr_tile<T, 32, kComp, H, U>(lo, hi, ids, contrib_in, lane);
  return pr_tile<T, 64, kComp, H, U>(lo, hi, ids, contrib_in, lane);
}

// H: the adjacency view holds empty slots (the CPMA); degrees come from
// deg[] (null otherwise).
template <class T, bool kComp, bool H = false>
__global__ void __launch_bounds__(256) k_pr_tiles(const uint64_t *off, const uint32_t *ids, uint64_t n,
                                                  const T *contrib_in, const T *p_old, T *p_new, T *contrib_out,
                                                  double damping, double addc, double *partial,
                                                  const uint32_t *degs = nullptr, const uint32_t *stop = nullptr) {
  __shared__ double lds[4];
  if (stop && *stop) return;  // a batched iteration after convergence
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const uint64_t ntiles = (n + 63) >> 6;
  double resid = 0.0;
  const T d = (T)damping, t = (T)addc;
  for (uint64_t tile = wave; tile < ntiles; tile += nwaves) {
    const uint64_t v = tile * 64 + lane;
    const bool valid = v < n;
    const uint64_t lo = off[valid ? v : n], hi = off[valid ? v + 1 : n];
    const uint32_t dmax = wave_max((uint32_t)(hi - lo));
    Acc<T> mine;
    if (dmax <= kPrUnroll) mine = pr_tile<T, 1, kComp, H>(lo, hi, ids, contrib_in, lane);
    else if (dmax <= 2 * kPrUnroll) mine = pr_tile<T, 2, kComp, H>(lo, hi, ids, contrib_in, lane);
    else if (dmax <= 4 * kPrUnroll) mine = pr_tile<T, 4, kComp, H>(lo, hi, ids, contrib_in, lane);
    else if (dmax <= 8 * kPrUnroll) mine = pr_tile<T, 8, kComp, H>(lo, hi, ids, contrib_in, lane);
    else if (dmax <= 16 * kPrUnroll) mine = pr_tile<T, 16, kComp, H>(lo, hi, ids, contrib_in, lane);
    else if (dmax <= 32 * kPrUnroll) mine = pr_tile<T, 32, kComp, H>(lo, hi, ids, contrib_in, lane);
    else mine = pr_tile<T, 64, kComp, H>(lo, hi, ids, contrib_in, lane);
    if (valid) {
      const T sum = kComp ? (mine.s - mine.c) : mine.s;
      const T pv = d * sum + t;
      const uint64_t deg = H ? (uint64_t)degs[v] : hi - lo;
      p_new[v] = pv;
      contrib_out[v] = deg ? pv / (T)deg : (T)0;
      resid += fabs((double)pv - (double)p_old[v]);
    }
  }
  resid = wave_sum(resid);
  if (lane == 0) lds[threadIdx.x >> 6] = resid;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// ------------------------------------------------ pull split over the XCDs
// (tuning knob pr_xcd).  Workgroups are dealt to the chip's 8 XCDs
// round-robin, so workgroup b runs on XCD b % 8; XCD x sums, for every
// vertex, only the neighbours in source slice x -- a range of ids holding
// 1/8 of the edges, or of the edges plus pr_xcd_vw per vertex (a slice of
// many low-degree sources has a footprint larger than one L2, and runs
// slower per edge) -- so the contributions it gathers form one slice of the
// vector, sized to stay in that XCD's own L2 instead of the whole vector
// competing for every L2.  The slices' neighbour lists are stored apart
// (slice x of vertex v at so[x * (n+1) + v] .. so[x * (n+1) + v + 1] of
// ids_s), each XCD writes its (sum, compensation) per vertex, and a second
// kernel combines the 8 partials in slice order (deterministic) and
// finalises the ranks.
constexpr int kXcds = 8;
constexpr int kXcdUnroll = 2;  // gathers in flight per lane in a slice's short lists (2 measured best of 1-4, 8)

// Sums slice x's neighbours of the 64 vertices of one tile into the
// slice's partials.
template <class T, bool kComp>
__device__ __forceinline__ void pr_xcd_tile(const uint32_t *so, const uint32_t *ids_s, uint64_t n,
                                            const T *contrib_in, T *part_s, T *part_c, uint32_t x, uint64_t tile,
                                            int lane) {
  const uint32_t *sox = so + (uint64_t)x * (n + 1);
  const uint64_t v = tile * 64 + lane;
  const bool valid = v < n;
  const uint64_t lo = sox[valid ? v : n], hi = sox[valid ? v + 1 : n];
  const uint32_t dmax = wave_max((uint32_t)(hi - lo));
  // slices hold a few edges per vertex: short unrolled steps
  const Acc<T> mine = pr_tile_any<T, kComp, false, kXcdUnroll>(dmax, lo, hi, ids_s, contrib_in, lane);
  if (valid) {
    part_s[(uint64_t)x * n + v] = mine.s;
    if (kComp) part_c[(uint64_t)x * n + v] = mine.c;
  }
}

// XCD x takes every tile of slice x in a fixed order.
template <class T, bool kComp>
__global__ void __launch_bounds__(256) k_pr_xcd(const uint32_t *so, const uint32_t *ids_s, uint64_t n,
                                                const T *contrib_in, T *part_s, T *part_c, const uint32_t *stop) {
  if (stop && *stop) return;  // a batched iteration after convergence
  const int lane = lane_id();
  const uint32_t x = blockIdx.x % kXcds;
  const uint64_t ntiles = (n + 63) >> 6;
  const uint64_t wave = (uint64_t)(blockIdx.x / kXcds) * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)(gridDim.x / kXcds) * (blockDim.x >> 6);
  for (uint64_t tile = wave; tile < ntiles; tile += nwaves)
    pr_xcd_tile<T, kComp>(so, ids_s, n, contrib_in, part_s, part_c, x, tile, lane);
}

// degs: true degrees when the view's lists hold empty slots (else null:
// the offsets give them).
template <class T, bool kComp>
__global__ void __launch_bounds__(256) k_pr_xcd_combine(const T *part_s, const T *part_c, const uint64_t *off,
                                                        uint64_t n, const T *p_old, T *p_new, T *contrib_out,
                                                        double damping, double addc, double *partial,
                                                        const uint32_t *degs, const uint32_t *stop) {
  __shared__ double lds[4];
  if (stop && *stop) return;  // a batched iteration after convergence
  const T d = (T)damping, t = (T)addc;
  double resid = 0.0;
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    Acc<T> a{part_s[v], kComp ? part_c[v] : (T)0};
#pragma unroll
    for (int x = 1; x < kXcds; x++) {
      if (kComp) a = combine(a, Acc<T>{part_s[(uint64_t)x * n + v], part_c[(uint64_t)x * n + v]});
      else a.s += part_s[(uint64_t)x * n + v];
    }
    const T sum = kComp ? (a.s - a.c) : a.s;
    const T pv = d * sum + t;
    const uint64_t deg = degs ? (uint64_t)degs[v] : off[v + 1] - off[v];
    p_new[v] = pv;
    contrib_out[v] = deg ? pv / (T)deg : (T)0;
    resid += fabs((double)pv - (double)p_old[v]);
  }
  resid = wave_sum(resid);
  if (lane_id() == 0) lds[threadIdx.x >> 6] = resid;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// Slice bounds: bnd[x] = first vertex v with w(v) >= x * w(n) / 8, where
// w(v) = off[v] + vw * v weighs each source by its edges plus vw (a slice's
// gathers cost more as its vertex count, the footprint of its
// contributions, outgrows one L2).
__global__ void k_xs_bounds(const uint64_t *off, uint64_t n, uint64_t vw, uint32_t *bnd) {
  const int x = threadIdx.x;
  if (x > kXcds) return;
  const uint64_t total = off[n] + vw * n;
  const uint64_t target = total / kXcds * x + (total % kXcds) * x / kXcds;
  uint64_t lo = 0, hi = n;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (off[mid] + vw * mid >= target) hi = mid; else lo = mid + 1;
  }
  bnd[x] = x == kXcds ? (uint32_t)n : (uint32_t)lo;
}

struct XcdBounds {
  uint32_t b[kXcds + 1];
};

// First index of v's sorted list [lo, hi) holding an id >= bound.
__device__ __forceinline__ uint64_t list_lower_bound(const uint32_t *ids, uint64_t lo, uint64_t hi, uint32_t bound) {
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (ids[mid] < bound) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// cnt[x * (n+1) + v] = neighbours of v in slice x (0 at v = n).  H: the
// lists may hold empty slots (kNone, a PMA slot view), which break their
// order for a binary search: the slots are walked and the holes skipped.
// Lists longer than kXsShort are left to one wave each (k_xs_long_*): a
// thread walking an RMAT hub's list alone held the whole pass.
constexpr uint64_t kXsShort = 256;

__device__ __forceinline__ int xs_slice(uint32_t u, const XcdBounds &bd) {
  int x = 0;
#pragma unroll
  for (int y = 1; y < kXcds; y++) x += u >= bd.b[y];
  return x;
}

template <bool H>
__global__ void k_xs_count(const uint64_t *off, const uint32_t *ids, uint64_t n, XcdBounds bd, uint32_t *cnt) {
  for (uint64_t v = global_tid(); v <= n; v += grid_threads()) {
    if (v == n) {
      for (int x = 0; x < kXcds; x++) cnt[(uint64_t)x * (n + 1) + n] = 0;
      continue;
    }
    uint64_t a = off[v];
    const uint64_t e = off[v + 1];
    if (e - a > kXsShort) continue;  // k_xs_long_count
    if constexpr (H) {
      uint32_t c[kXcds] = {};
      int x = 0;
      for (uint64_t k = a; k < e; k++) {
        const uint32_t u = ids[k];
        if (u == kNone) continue;
        while (x < kXcds - 1 && u >= bd.b[x + 1]) x++;
        c[x]++;
      }
      for (int y = 0; y < kXcds; y++) cnt[(uint64_t)y * (n + 1) + v] = c[y];
    } else {
      for (int x = 0; x < kXcds; x++) {
        const uint64_t b = x == kXcds - 1 ? e : list_lower_bound(ids, a, e, bd.b[x + 1]);
        cnt[(uint64_t)x * (n + 1) + v] = (uint32_t)(b - a);
        a = b;
      }
    }
  }
}

template <bool H>
__global__ void k_xs_scatter(const uint64_t *off, const uint32_t *ids, uint64_t n, XcdBounds bd, const uint32_t *so,
                             uint32_t *ids_s) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    if (off[v + 1] - off[v] > kXsShort) continue;  // k_xs_long_scatter
    int x = 0;
    uint64_t pos = so[v];
    for (uint64_t e = off[v]; e < off[v + 1]; e++) {
      const uint32_t u = ids[e];
      if (H && u == kNone) continue;
      while (x < kXcds - 1 && u >= bd.b[x + 1]) {
        x++;
        pos = so[(uint64_t)x * (n + 1) + v];
      }
      ids_s[pos++] = u;
    }
  }
}

// One wave per long list: every 64 slots, a ballot per slice counts (and
// ranks) the ids of that slice; slots are visited in order, so each
// slice's run keeps the list's order.
template <bool H>
__global__ void k_xs_long_count(const uint64_t *off, const uint32_t *ids, uint64_t n, XcdBounds bd, uint32_t *cnt) {
  const int lane = lane_id();
  const uint64_t wstride = (grid_threads() >> 6) * 64;
  for (uint64_t v0 = (global_tid() >> 6) * 64; v0 < n; v0 += wstride) {
    // 64 vertices' bounds at once; the long lists among them one by one
    const uint64_t vl = v0 + lane;
    const uint64_t al = vl < n ? off[vl] : 0, el = vl < n ? off[vl + 1] : 0;
    uint64_t longm = __ballot(vl < n && el - al > kXsShort);
    while (longm) {
      const int l = __builtin_ctzll(longm);
      longm &= longm - 1;
      const uint64_t v = v0 + l, a = __shfl(al, l, 64), e = __shfl(el, l, 64);
      uint32_t c = 0;  // lane x < kXcds: the count of slice x
      for (uint64_t k0 = a; k0 < e; k0 += 64) {
        const uint64_t k = k0 + lane;
        const uint32_t u = k < e ? ids[k] : kNone;
        const bool real = k < e && (!H || u != kNone);
        const int x = real ? xs_slice(u, bd) : -1;
#pragma unroll
        for (int y = 0; y < kXcds; y++) {
          const uint32_t t = (uint32_t)__popcll(__ballot(x == y));
          if (lane == y) c += t;
        }
      }
      if (lane < kXcds) cnt[(uint64_t)lane * (n + 1) + v] = c;
    }
  }
}

template <bool H>
__global__ void k_xs_long_scatter(const uint64_t *off, const uint32_t *ids, uint64_t n, XcdBounds bd,
                                  const uint32_t *so, uint32_t *ids_s) {
  const int lane = lane_id();
  const uint64_t wstride = (grid_threads() >> 6) * 64;
  for (uint64_t v0 = (global_tid() >> 6) * 64; v0 < n; v0 += wstride) {
    const uint64_t vl = v0 + lane;
    const uint64_t al = vl < n ? off[vl] : 0, el = vl < n ? off[vl + 1] : 0;
    uint64_t longm = __ballot(vl < n && el - al > kXsShort);
    while (longm) {
      const int l = __builtin_ctzll(longm);
      longm &= longm - 1;
      const uint64_t v = v0 + l, a = __shfl(al, l, 64), e = __shfl(el, l, 64);
      // lane x < kXcds holds slice x's next output position
      uint64_t pos = lane < kXcds ? so[(uint64_t)lane * (n + 1) + v] : 0;
      for (uint64_t k0 = a; k0 < e; k0 += 64) {
        const uint64_t k = k0 + lane;
        const uint32_t u = k < e ? ids[k] : kNone;
        const bool real = k < e && (!H || u != kNone);
        const int x = real ? xs_slice(u, bd) : -1;
#pragma unroll
        for (int y = 0; y < kXcds; y++) {
          const uint64_t bal = __ballot(x == y);
          const uint64_t py = __shfl(pos, y, 64);
          if (x == y) ids_s[py + mbcnt(bal)] = u;
          if (lane == y) pos += (uint64_t)__popcll(bal);
        }
      }
    }
  }
}

// pos[(s-1)*n + v] = first index of v's list holding an id >= s*W
// (s = 1 .. S-1).
__global__ void k_pr_bounds(const uint64_t *off, const uint32_t *ids, uint64_t n, int S, uint64_t W, uint64_t *pos) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    uint64_t lo = off[v];
    const uint64_t end = off[v + 1];
    for (int s = 1; s < S; s++) {
      const uint64_t bound = (uint64_t)s * W;
      uint64_t a = lo, b = end;
      while (a < b) {
        const uint64_t mid = (a + b) >> 1;
        if (ids[mid] < bound) a = mid + 1; else b = mid;
      }
      pos[(uint64_t)(s - 1) * n + v] = a;
      lo = a;
    }
  }
}

// With stop: an iteration after the one that converged returns at once, and
// the first L1 below eps records its iteration count (it1) in *stop.
__global__ void k_sum_doubles(const double *partial, int nb, double *out, uint32_t *stop = nullptr, double eps = 0.0,
                              uint32_t it1 = 0) {
  // one block, fixed association order
  __shared__ double lds[1024];
  if (stop && *stop) return;
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    *out = lds[0];
    if (stop && lds[0] < eps) *stop = it1;
  }
}

template <class T>
__global__ void k_pr_init(const uint64_t *off, uint64_t n, T *p, T *contrib, const uint32_t *degs) {
  const double p0 = 1.0 / (double)n;
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    const uint64_t deg = degs ? (uint64_t)degs[v] : off[v + 1] - off[v];
    p[v] = (T)p0;
    contrib[v] = deg ? (T)(p0 / (double)deg) : (T)0;
  }
}

template <class T>
__global__ void k_to_double(const T *in, double *out, uint64_t n) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) out[v] = (double)in[v];
}

// Rank of vertex v from the relabelled vector (perm: old id -> new id).
template <class T>
__global__ void k_unpermute(const T *in, const uint32_t *perm, double *out, uint64_t n) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) out[v] = (double)in[perm[v]];
}

template <class T, int G, bool kComp>
void pr_pass(bool first, bool last, const uint64_t *lo, const uint64_t *hi, const uint32_t *ids, const uint64_t *off,
             uint64_t n, const T *cin, T *as, T *ac, const T *pold, T *pnew, T *cout, double damping, double addc,
             double *partial, unsigned grid, hipStream_t s) {
#define BYO_PR_PASS(F, L) \
  k_pr_seg<T, G, kComp, F, L><<<grid, 256, 0, s>>>(lo, hi, ids, off, n, cin, as, ac, pold, pnew, cout, damping, addc, partial)
  if (first && last) BYO_PR_PASS(true, true);
  else if (first) BYO_PR_PASS(true, false);
  else if (last) BYO_PR_PASS(false, true);
  else BYO_PR_PASS(false, false);
#undef BYO_PR_PASS
}

template <class T, bool kComp>
void pr_pass_g(int G, bool first, bool last, const uint64_t *lo, const uint64_t *hi, const uint32_t *ids,
               const uint64_t *off, uint64_t n, const T *cin, T *as, T *ac, const T *pold, T *pnew, T *cout,
               double damping, double addc, double *partial, unsigned grid, hipStream_t s) {
  switch (G) {
    case 4: pr_pass<T, 4, kComp>(first, last, lo, hi, ids, off, n, cin, as, ac, pold, pnew, cout, damping, addc, partial, grid, s); break;
    case 8: pr_pass<T, 8, kComp>(first, last, lo, hi, ids, off, n, cin, as, ac, pold, pnew, cout, damping, addc, partial, grid, s); break;
    case 16: pr_pass<T, 16, kComp>(first, last, lo, hi, ids, off, n, cin, as, ac, pold, pnew, cout, damping, addc, partial, grid, s); break;
    default: pr_pass<T, 32, kComp>(first, last, lo, hi, ids, off, n, cin, as, ac, pold, pnew, cout, damping, addc, partial, grid, s); break;
  }
}

// Source-segment width in vertices (tuning knob pr_seg_log2; one segment
// by default).
template <class T>
uint64_t segment_width(byo_graph *g, uint64_t n) {
  const int lg = (int)g->tuning("pr_seg_log2");
  uint64_t W = 1ull << lg;
  return W < n ? W : n;
}

// The graph the pull runs on: as it is, or the degree-relabelled copy
// (tuning knob pr_relabel), whose contributions the gathers hit most sit
// together at low ids, where they stay in L2.
RelabelView pr_view(byo_graph *g) {
  if (!g->tuning("pr_relabel") || g->n == 0)
    return RelabelView{g->off.as<uint64_t>(), g->ids.as<uint32_t>(), nullptr};
  return relabel_view(g);
}

// Per-slice neighbour lists of the pull's graph view for k_pr_xcd, built
// once per graph and layout (relabelled or not).
struct XcdSlices {
  const uint32_t *so, *ids_s;
};

XcdSlices xcd_slices(byo_graph *g, const RelabelView &view, bool holes) {
  const uint64_t n = g->n, m = g->m;
  uint32_t *so = g->buf("pr_xs_so", (uint64_t)kXcds * (n + 1) * 4 + 64).as<uint32_t>();
  uint32_t *ids_s = g->buf("pr_xs_ids", m * 4 + 64).as<uint32_t>();
  const int layout = holes ? 2 : (view.perm ? 1 : 0);
  const int64_t vw = g->tuning("pr_xcd_vw");
  if (!g->xs_valid || g->xs_layout != layout || g->xs_vw != vw) {
    hipStream_t s = g->stream;
    PhaseTimer t(s, g->prof.on);
    uint32_t *dbnd = g->buf("pr_xs_bnd", 64 * 4).as<uint32_t>();
    k_xs_bounds<<<1, 64, 0, s>>>(view.off, n, (uint64_t)vw, dbnd);
    XcdBounds bd{};
    BYO_HIP(hipMemcpyAsync(bd.b, dbnd, sizeof(bd.b), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    const uint64_t cnt_n = (uint64_t)kXcds * (n + 1);
    DevBuf cnt(cnt_n * 4 + 64), tmp(prim::scan_tmp_bytes(cnt_n));
    if (holes) {
      k_xs_count<true><<<byo_grid(n + 1, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd, cnt.as<uint32_t>());
      k_xs_long_count<true><<<byo_grid(n, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd,
                                                                          cnt.as<uint32_t>());
    } else {
      k_xs_count<false><<<byo_grid(n + 1, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd, cnt.as<uint32_t>());
      k_xs_long_count<false><<<byo_grid(n, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd,
                                                                           cnt.as<uint32_t>());
    }
    prim::exclusive_scan_u32(cnt.as<uint32_t>(), so, cnt_n, false, tmp.p, s);
    if (holes) {
      k_xs_scatter<true><<<byo_grid(n, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd, so, ids_s);
      k_xs_long_scatter<true><<<byo_grid(n, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd, so, ids_s);
    } else {
      k_xs_scatter<false><<<byo_grid(n, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd, so, ids_s);
      k_xs_long_scatter<false><<<byo_grid(n, 256, 16384), 256, 0, s>>>(view.off, view.ids, n, bd, so, ids_s);
    }
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipStreamSynchronize(s));
    if (g->prof.on) g->account("pr_xcd_slices", t.stop());
    g->xs_valid = true;
    g->xs_layout = layout;
    g->xs_vw = vw;
  }
  return XcdSlices{so, ids_s};
}

// ------------------------------------------------ propagation blocking
// (tuning knob pr_pb).  The pull above pays one scattered L2 request per
// edge for a 4- or 8-byte contribution.  Here an iteration moves the
// contributions in two streaming passes over a layout built once per graph:
//  * destinations are cut into bins of at most kPbV consecutive vertices
//    and about kPbE in-edges (a bin around a hub holds more and is summed by
//    several gather items); every edge has a slot in a bin-major array,
//    ordered by the bin of its destination, then by source, and dl[slot]
//    (16 bits) is the destination's offset in its bin;
//  * sources are cut into blocks of at most kPbSV consecutive vertices and
//    about kPbSE out-edges; a block's slots form one run per bin, and fq /
//    fs list them in slot order (fq the slot, fs its source's offset in the
//    block);
//  * scatter: one workgroup per source block finalises its sources' ranks
//    from the previous gather, stages their contributions in LDS and writes
//    its runs slot after slot (consecutive lanes, consecutive slots); the
//    blocks of one XCD are consecutive, so the runs that meet in a bin meet
//    in that XCD's L2;
//  * gather: one workgroup per item streams the item's slots and adds each
//    value into its destination's accumulator in LDS.
// The sums are kept in 64-bit fixed point (FP32 values: one word in 2^-62
// units; FP64: two words, 2^-62 and 2^-94 units).  Every contribution is at
// most 1 and every in-sum at most the total rank, 1, so the words do not
// overflow; the LDS atomics add exactly, and the result does not depend on
// their order (bit-identical between runs).
#ifndef BYO_PB_E_LOG2
#define BYO_PB_E_LOG2 17
#endif
#ifndef BYO_PB_SE_LOG2
#define BYO_PB_SE_LOG2 16
#endif
#ifndef BYO_PB_GU
#define BYO_PB_GU 4
#endif
#ifndef BYO_PB_SU
#define BYO_PB_SU 4
#endif
#ifndef BYO_PB_SV
#define BYO_PB_SV 8192
#endif
#ifndef BYO_PB_SB32
#define BYO_PB_SB32 1024
#endif
#ifndef BYO_PB_SB64
#define BYO_PB_SB64 512
#endif
#ifndef BYO_PB_GB32
#define BYO_PB_GB32 1024
#endif
#ifndef BYO_PB_GB64
#define BYO_PB_GB64 1024
#endif
constexpr uint32_t kPbV = 4096;                    // destinations per bin
constexpr uint64_t kPbE = 1ull << BYO_PB_E_LOG2;   // slots per gather item
constexpr uint32_t kPbSV = BYO_PB_SV;              // sources per scatter block
constexpr uint64_t kPbSE = 1ull << BYO_PB_SE_LOG2; // slots per scatter block (a hub's block holds more)
// threads per scatter block and per gather item, by value width
template <class T>
struct PbThreads {
  static constexpr int scatter = sizeof(T) == 4 ? BYO_PB_SB32 : BYO_PB_SB64;
  static constexpr int gather = sizeof(T) == 4 ? BYO_PB_GB32 : BYO_PB_GB64;
};
constexpr int kPbGU = BYO_PB_GU;                   // slot loads in flight per thread (gather)
constexpr int kPbSU = BYO_PB_SU;                   // slot stores in flight per thread (scatter)

struct PbItem {
  uint32_t lo, nv;      // first destination of the bin, destinations in it
  uint64_t e0, e1;      // slot range
  uint32_t split, pad;  // 1: the bin spans several items (sums added atomically)
};

template <class T>
struct PbWords {
  static constexpr int k = sizeof(T) == 8 ? 2 : 1;
};

// x in [0, 1] in fixed point: w[0] = floor(x 2^62); FP64 also w[1] = the
// rest in 2^-94 units (the products and differences below are exact).
template <class T>
__device__ __forceinline__ void pb_fix(T x, unsigned long long *w) {
  const double q = (double)x * 0x1p62;
  const unsigned long long hi = (unsigned long long)q;
  w[0] = hi;
  if (PbWords<T>::k == 2) w[1] = (unsigned long long)((q - (double)hi) * 0x1p32);
}

// One workgroup per source block b = [sb[b], sb[b+1]) (workgroup i takes
// block (i % 8) * per + i / 8, per = ceil(blocks / 8): XCD i % 8 a
// contiguous range).  kInit: the contributions of the ranks in p_old
// (iteration 0).  Otherwise p_new = d * (sum gathered in acc) + (1-d)/n,
// the block's share of the L1 residual, acc cleared for the next gather,
// and (write_vals) the contributions p_new/deg written to the block's
// slots.
template <class T, bool kInit>
__global__ void __launch_bounds__(PbThreads<T>::scatter) k_pb_scatter(const uint64_t *off, const uint32_t *sb, uint32_t nsb,
                                                         const uint32_t *fq, const uint16_t *fs, uint64_t n,
                                                         unsigned long long *acc, const T *p_old, T *p_new, T *val,
                                                         double damping, double addc, double *partial,
                                                         int write_vals, const uint32_t *stop) {
  __shared__ T cs[kPbSV];
  __shared__ double lds[PbThreads<T>::scatter / 64];
  if (stop && *stop) return;  // a batched iteration after convergence
  const uint32_t per = (nsb + kXcds - 1) / kXcds;
  const uint32_t b = (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
  double resid = 0.0;
  if (b < nsb) {
    const uint32_t v0 = sb[b], v1 = sb[b + 1];
    for (uint32_t i = threadIdx.x; i < v1 - v0; i += PbThreads<T>::scatter) {
      const uint64_t v = v0 + i;
      T pv;
      if (kInit) {
        pv = p_old[v];
      } else {
        double sum = (double)acc[v] * 0x1p-62;
        if (PbWords<T>::k == 2) sum += (double)acc[n + v] * 0x1p-94;
        pv = (T)(damping * sum + addc);
        p_new[v] = pv;
        resid += fabs((double)pv - (double)p_old[v]);
        acc[v] = 0;
        if (PbWords<T>::k == 2) acc[n + v] = 0;
      }
      const uint64_t deg = off[v + 1] - off[v];
      cs[i] = deg ? pv / (T)deg : (T)0;
    }
    if (write_vals) {
      __syncthreads();
      // the block's slots: fq/fs entries [off[v0], off[v1]) (a block's
      // slot count is its sources' edge count)
      const uint64_t j0 = off[v0], j1 = off[v1];
      for (uint64_t j = j0 + threadIdx.x; j < j1; j += (uint64_t)PbThreads<T>::scatter * kPbSU) {
        uint32_t q[kPbSU];
        uint16_t l[kPbSU];
#pragma unroll
        for (int u = 0; u < kPbSU; u++) {
          const uint64_t jj = j + (uint64_t)u * PbThreads<T>::scatter;
          const uint64_t at = jj < j1 ? jj : j;
          q[u] = fq[at];
          l[u] = fs[at];
        }
#pragma unroll
        for (int u = 0; u < kPbSU; u++)
          if (j + (uint64_t)u * PbThreads<T>::scatter < j1) val[q[u]] = cs[l[u]];
      }
    }
  }
  if (kInit) return;
  resid = wave_sum(resid);
  if (lane_id() == 0) lds[threadIdx.x >> 6] = resid;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < PbThreads<T>::scatter / 64; w++) t += lds[w];
    partial[blockIdx.x] = t;
  }
}

// One workgroup per item: the item's slots added into LDS accumulators of
// its bin's destinations, then written to acc (word w of destination v at
// acc[w * n + v]); a split bin's items add theirs atomically (acc is clear).
template <class T>
__global__ void __launch_bounds__(PbThreads<T>::gather) k_pb_gather(const PbItem *items, const T *val, const uint16_t *dl,
                                                        unsigned long long *acc, uint64_t n, const uint32_t *stop) {
  constexpr int W = PbWords<T>::k;
  __shared__ unsigned long long s[W * kPbV];
  if (stop && *stop) return;  // a batched iteration after convergence
  const PbItem it = items[blockIdx.x];
#pragma unroll
  for (int w = 0; w < W; w++)
    for (uint32_t i = threadIdx.x; i < it.nv; i += PbThreads<T>::gather) s[w * kPbV + i] = 0;
  __syncthreads();
  for (uint64_t e = it.e0 + threadIdx.x; e < it.e1; e += (uint64_t)PbThreads<T>::gather * kPbGU) {
    T xv[kPbGU];
    uint32_t d[kPbGU];
#pragma unroll
    for (int j = 0; j < kPbGU; j++) {
      const uint64_t kk = e + (uint64_t)j * PbThreads<T>::gather;
      const uint64_t at = kk < it.e1 ? kk : e;
      xv[j] = val[at];
      d[j] = dl[at];
    }
#pragma unroll
    for (int j = 0; j < kPbGU; j++) {
      if (e + (uint64_t)j * PbThreads<T>::gather < it.e1) {
        unsigned long long w[2];
        pb_fix<T>(xv[j], w);
        atomicAdd(&s[d[j]], w[0]);
        if (W == 2) atomicAdd(&s[kPbV + d[j]], w[1]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int w = 0; w < W; w++) {
    for (uint32_t i = threadIdx.x; i < it.nv; i += PbThreads<T>::gather) {
      const unsigned long long a = s[w * kPbV + i];
      unsigned long long *dst = acc + (uint64_t)w * n + it.lo + i;
      if (!it.split) *dst = a;
      else if (a) atomicAdd(dst, a);
    }
  }
}

// Layout build.  cut[j] = first vertex v with off[v] >= j * step.
__global__ void k_pb_cuts(const uint64_t *off, uint64_t n, uint64_t step, uint64_t ncut, uint32_t *cut) {
  for (uint64_t j = global_tid(); j < ncut; j += grid_threads()) {
    const uint64_t target = j * step;
    uint64_t lo = 0, hi = n;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (off[mid] >= target) hi = mid; else lo = mid + 1;
    }
    cut[j] = (uint32_t)lo;
  }
}
// eb[k] = off[lo[k]]; which[v] = the range k with lo[k] <= v < lo[k+1].
__global__ void k_pb_range_edges(const uint64_t *off, const uint32_t *lo, uint64_t nr1, uint64_t *eb) {
  for (uint64_t k = global_tid(); k < nr1; k += grid_threads()) eb[k] = off[lo[k]];
}
__global__ void k_pb_which(const uint32_t *lo, uint64_t nr, uint64_t n, uint32_t *which) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    uint64_t a = 0, b = nr;  // last k with lo[k] <= v
    while (b - a > 1) {
      const uint64_t mid = (a + b) >> 1;
      if (lo[mid] <= v) a = mid; else b = mid;
    }
    which[v] = (uint32_t)a;
  }
}
// keys[e] = (bin of ids[e]) << 32 | e: in edge order, so a stable sort on
// the bin digits leaves each bin's slots in source order.
__global__ void k_pb_keys(const uint32_t *ids, uint64_t m, const uint32_t *vbin, uint64_t *keys) {
  for (uint64_t e = global_tid(); e < m; e += grid_threads()) keys[e] = ((uint64_t)vbin[ids[e]] << 32) | e;
}
// Slot q holds edge e = low word of sorted[q] (bin k = high word): its
// destination offset dl[q], and the key (source block << 32 | q) of the
// second sort, which lists every source block's slots in slot order.
__global__ void k_pb_slots(const uint64_t *sorted, uint64_t m, const uint32_t *ids, const uint64_t *ekeys,
                           const uint32_t *bin_lo, const uint32_t *vsb, uint16_t *dl, uint32_t *ssrc,
                           uint64_t *keys2) {
  for (uint64_t q = global_tid(); q < m; q += grid_threads()) {
    const uint64_t key = sorted[q];
    const uint32_t e = (uint32_t)key, k = (uint32_t)(key >> 32);
    const uint32_t src = (uint32_t)(ekeys[e] >> 32);
    dl[q] = (uint16_t)(ids[e] - bin_lo[k]);
    ssrc[q] = src;
    keys2[q] = ((uint64_t)vsb[src] << 32) | q;
  }
}
__global__ void k_pb_flist(const uint64_t *sorted2, uint64_t m, const uint32_t *ssrc, const uint32_t *sb_lo,
                           uint32_t *fq, uint16_t *fs) {
  for (uint64_t j = global_tid(); j < m; j += grid_threads()) {
    const uint64_t key = sorted2[j];
    const uint32_t q = (uint32_t)key, b = (uint32_t)(key >> 32);
    fq[j] = q;
    fs[j] = (uint16_t)(ssrc[q] - sb_lo[b]);
  }
}

struct PbLayout {
  const PbItem *items;
  const uint32_t *sb, *fq;
  const uint16_t *dl, *fs;
  uint32_t nitems, nsb;
};

// Cuts [0, n) into ranges of about step edges and at most width vertices
// (the ranges' first vertices, ending with n).
std::vector<uint32_t> pb_ranges(byo_graph *g, const uint64_t *off, uint64_t step, uint64_t width) {
  const uint64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  const uint64_t ncut = (m + step - 1) / step + 1;
  DevBuf dcut(ncut * 4 + 64);
  k_pb_cuts<<<byo_grid(ncut, 256, 4096), 256, 0, s>>>(off, n, step, ncut, dcut.as<uint32_t>());
  BYO_LAUNCH_CHECK();
  std::vector<uint32_t> cut(ncut);
  BYO_HIP(hipMemcpyAsync(cut.data(), dcut.p, ncut * 4, hipMemcpyDeviceToHost, s));
  BYO_HIP(hipStreamSynchronize(s));
  cut.push_back((uint32_t)n);
  std::vector<uint32_t> lo{0};
  uint64_t prev = 0;
  for (uint32_t c : cut) {
    if (c <= prev) continue;
    const uint64_t w = c - prev, parts = (w + width - 1) / width;
    for (uint64_t i = 1; i < parts; i++) lo.push_back((uint32_t)(prev + w * i / parts));
    lo.push_back(c);
    prev = c;
  }
  return lo;
}

// Stable sort of keys[m] on the digits of the high word below 2^bits;
// returns the sorted buffer (keys or alt).
const uint64_t *pb_sort_high(uint64_t *keys, uint64_t *alt, uint64_t m, uint64_t maxhigh, void *tmp,
                             hipStream_t s) {
  int shifts[4], np = 0;
  for (int b = 0; b < 32 && (maxhigh >> b); b += 8) shifts[np++] = 32 + b;
  return np ? prim::radix_sort_u64(keys, alt, m, shifts, np, tmp, s) : keys;
}

// The propagation-blocking layout of a CSR view, built once per graph and
// layout (0: the CSR, 1: its relabelled copy, 2: the CSR export of a CPMA).
PbLayout pb_layout(byo_graph *g, const RelabelView &view, int layout) {
  const uint64_t n = g->n, m = g->m;
  uint16_t *dl = g->buf("pr_pb_dl", m * 2 + 64).as<uint16_t>();
  uint32_t *fq = g->buf("pr_pb_fq", m * 4 + 64).as<uint32_t>();
  uint16_t *fs = g->buf("pr_pb_fs", m * 2 + 64).as<uint16_t>();
  if (!g->pb_valid || g->pb_layout != layout) {
    hipStream_t s = g->stream;
    PhaseTimer t(s, g->prof.on);
    const std::vector<uint32_t> lo = pb_ranges(g, view.off, kPbE, kPbV);     // destination bins
    const std::vector<uint32_t> sbl = pb_ranges(g, view.off, kPbSE, kPbSV);  // source blocks
    const uint64_t nbins = lo.size() - 1, nsb = sbl.size() - 1;
    DevBuf dlo(lo.size() * 4 + 64), deb(lo.size() * 8 + 64);
    uint32_t *dsb = g->buf("pr_pb_sb", sbl.size() * 4 + 64).as<uint32_t>();
    BYO_HIP(hipMemcpyAsync(dlo.p, lo.data(), lo.size() * 4, hipMemcpyHostToDevice, s));
    BYO_HIP(hipMemcpyAsync(dsb, sbl.data(), sbl.size() * 4, hipMemcpyHostToDevice, s));
    k_pb_range_edges<<<byo_grid(lo.size(), 256, 4096), 256, 0, s>>>(view.off, dlo.as<uint32_t>(), lo.size(),
                                                                  deb.as<uint64_t>());
    BYO_LAUNCH_CHECK();
    std::vector<uint64_t> eb(lo.size());
    BYO_HIP(hipMemcpyAsync(eb.data(), deb.p, lo.size() * 8, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    std::vector<PbItem> items;
    for (uint64_t k = 0; k < nbins; k++) {
      const uint64_t e0 = eb[k], e1 = eb[k + 1];
      if (e1 == e0) continue;  // no in-edges: its sums stay 0
      const uint64_t parts = (e1 - e0 + kPbE - 1) / kPbE;
      for (uint64_t i = 0; i < parts; i++) {
        PbItem it{};
        it.lo = lo[k];
        it.nv = lo[k + 1] - lo[k];
        it.e0 = e0 + (e1 - e0) * i / parts;
        it.e1 = e0 + (e1 - e0) * (i + 1) / parts;
        it.split = parts > 1;
        items.push_back(it);
      }
    }
    PbItem *ditems = g->buf("pr_pb_items", items.size() * sizeof(PbItem) + 64).as<PbItem>();
    BYO_HIP(hipMemcpyAsync(ditems, items.data(), items.size() * sizeof(PbItem), hipMemcpyHostToDevice, s));
    {
      DevBuf vbin(n * 4 + 64), vsb(n * 4 + 64), ekeys(m * 8 + 64), keys(m * 8 + 64), alt(m * 8 + 64),
          ssrc(m * 4 + 64), tmp(prim::radix_tmp_bytes(m));
      k_pb_which<<<byo_grid(n, 256, 16384), 256, 0, s>>>(dlo.as<uint32_t>(), nbins, n, vbin.as<uint32_t>());
      k_pb_which<<<byo_grid(n, 256, 16384), 256, 0, s>>>(dsb, nsb, n, vsb.as<uint32_t>());
      k_pb_keys<<<byo_grid(m, 256, 16384), 256, 0, s>>>(view.ids, m, vbin.as<uint32_t>(), keys.as<uint64_t>());
      prim::csr_to_keys(view.off, view.ids, n, m, ekeys.as<uint64_t>(), s);  // edge sources
      BYO_LAUNCH_CHECK();
      // slots: the edges sorted (stably) on their destination's bin
      const uint64_t *sorted = pb_sort_high(keys.as<uint64_t>(), alt.as<uint64_t>(), m, nbins - 1, tmp.p, s);
      uint64_t *keys2 = sorted == keys.as<uint64_t>() ? alt.as<uint64_t>() : keys.as<uint64_t>();
      k_pb_slots<<<byo_grid(m, 256, 16384), 256, 0, s>>>(sorted, m, view.ids, ekeys.as<uint64_t>(),
                                                        dlo.as<uint32_t>(), vsb.as<uint32_t>(), dl,
                                                        ssrc.as<uint32_t>(), keys2);
      BYO_LAUNCH_CHECK();
      // every source block's slots in slot order: a stable sort on the block
      uint64_t *other = keys2 == keys.as<uint64_t>() ? alt.as<uint64_t>() : keys.as<uint64_t>();
      const uint64_t *sorted2 = pb_sort_high(keys2, other, m, nsb - 1, tmp.p, s);
      k_pb_flist<<<byo_grid(m, 256, 16384), 256, 0, s>>>(sorted2, m, ssrc.as<uint32_t>(), dsb, fq, fs);
      BYO_LAUNCH_CHECK();
      BYO_HIP(hipStreamSynchronize(s));
    }
    g->pb_items = (uint32_t)items.size();
    g->pb_nsb = (uint32_t)nsb;
    if (g->prof.on) g->account("pr_pb_layout", t.stop());
    g->pb_valid = true;
    g->pb_layout = layout;
  }
  PbLayout L{};
  L.items = g->buf("pr_pb_items", 64).as<PbItem>();
  L.sb = g->buf("pr_pb_sb", 64).as<uint32_t>();
  L.fq = fq;
  L.fs = fs;
  L.dl = dl;
  L.nitems = g->pb_items;
  L.nsb = g->pb_nsb;
  return L;
}

template <class T, bool kComp>
void pagerank_impl(byo_graph *g, double damping, int max_iters, double eps, double *p_out, double *l1_log,
                   int *iters_run, double *device_ms, const char *phase) {
  const uint64_t n = g->n;
  hipStream_t s = g->stream;
  // a CPMA is pulled in place (its slot index and true degrees); the
  // relabelled, per-XCD and segmented layouts are copies of a CSR
  const AdjView av = adj_view(g);
  const bool pma = av.deg != nullptr;
  const RelabelView view = pma ? RelabelView{av.off, av.ids, nullptr} : pr_view(g);
  T *p[2], *c[2];
  p[0] = g->buf("pr_p0", n * sizeof(T) + 64).as<T>();
  p[1] = g->buf("pr_p1", n * sizeof(T) + 64).as<T>();
  c[0] = g->buf("pr_c0", n * sizeof(T) + 64).as<T>();
  c[1] = g->buf("pr_c1", n * sizeof(T) + 64).as<T>();
  const uint64_t W = pma ? n : segment_width<T>(g, n);
  const int S = (int)((n + W - 1) / W);
  T *acc_s = S > 1 ? g->buf("pr_acc_s", n * sizeof(T) + 64).as<T>() : nullptr;
  T *acc_c = (S > 1 && kComp) ? g->buf("pr_acc_c", n * sizeof(T) + 64).as<T>() : nullptr;
  uint64_t *pos = S > 1 ? g->buf("pr_pos", (uint64_t)(S - 1) * n * 8 + 64).as<uint64_t>() : nullptr;
  // lanes per vertex: picked per tile with one segment and pr_lanes 0;
  // otherwise pr_lanes, or about half the per-segment average degree, in
  // [4, 32]
  const bool per_tile = pma || (S == 1 && !g->tuning("pr_lanes"));
  // the XCD split needs the tile pull and 32-bit edge positions; on a PMA
  // its slices are built from the slot view (holes dropped)
  // propagation blocking (CSR views with 32-bit slot numbers) replaces the
  // pull, and the XCD split with it
  const bool use_pb = per_tile && g->tuning("pr_pb") && g->m > 0 && g->m < 0xFFFFFFFFull;
  const bool per_xcd = !use_pb && per_tile && g->tuning("pr_xcd") && g->m < (1ull << 32) && n > 0;
  const double avg = n ? (double)g->m / (double)n / S : 0.0;
  int G = 4;
  while (G < 32 && G < avg / 2) G *= 2;
  if (g->tuning("pr_lanes")) G = (int)g->tuning("pr_lanes");
  const unsigned grid = (unsigned)g->tuning("pr_grid");
  const unsigned cgrid = (unsigned)byo_grid(n, 256, 4096);
  double *partial = g->buf("pr_partial", (std::max(grid, cgrid) + 8) * sizeof(double)).as<double>();
  XcdSlices xs{};
  T *part_s = nullptr, *part_c = nullptr;
  if (per_xcd) {
    xs = xcd_slices(g, view, pma);
    part_s = g->buf("pr_xs_part_s", (uint64_t)kXcds * n * sizeof(T) + 64).as<T>();
    if (kComp) part_c = g->buf("pr_xs_part_c", (uint64_t)kXcds * n * sizeof(T) + 64).as<T>();
  }
  PbLayout pb{};
  const uint64_t *pb_off = nullptr;
  T *pb_val = nullptr;
  unsigned long long *pb_acc = nullptr;
  if (use_pb) {
    // a CPMA's layout is cut from its CSR export (built once per update, like
    // the layout); the iterations read the layout and its compact offsets
    RelabelView pbv = view;
    if (pma) {
      ensure_csr(g);
      pbv = RelabelView{g->off.as<uint64_t>(), g->ids.as<uint32_t>(), nullptr};
    }
    pb = pb_layout(g, pbv, pma ? 2 : (view.perm ? 1 : 0));
    pb_off = pbv.off;
    // the scatter's residual partials: one per source-block workgroup
    partial = g->buf("pr_partial", ((size_t)std::max({grid, cgrid, (pb.nsb + kXcds - 1) / kXcds * kXcds}) + 8) *
                                       sizeof(double)).as<double>();
    pb_val = g->buf("pr_pb_val", g->m * sizeof(T) + 64).as<T>();
    pb_acc = g->buf("pr_pb_acc", (uint64_t)PbWords<T>::k * n * 8 + 64).as<unsigned long long>();
    prim::fill_u64(reinterpret_cast<uint64_t *>(pb_acc), 0, (uint64_t)PbWords<T>::k * n, s);
  }
  double *l1dev = g->buf("pr_l1", (size_t)(max_iters + 1) * sizeof(double)).as<double>();
  const uint64_t *off = view.off;
  const uint32_t *ids = view.ids;
  const double addc = (1.0 - damping) * (1.0 / (double)n);
  PhaseTimer timer(s, true);
  k_pr_init<T><<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, n, p[0], c[0], av.deg);
  if (S > 1) k_pr_bounds<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, ids, n, S, W, pos);
  const unsigned sgrid = use_pb ? (pb.nsb + kXcds - 1) / kXcds * kXcds : 0;
  if (use_pb)
    k_pb_scatter<T, true><<<sgrid, PbThreads<T>::scatter, 0, s>>>(pb_off, pb.sb, pb.nsb, pb.fq, pb.fs, n, pb_acc, p[0], nullptr,
                                                     pb_val, damping, addc, partial, 1, nullptr);
  BYO_LAUNCH_CHECK();
  double *hl1 = static_cast<double *>(pinned(g, 64));
  // With eps > 0 the iterations are queued in batches of kSpec with a device
  // stop word: the residual kernel of the first iteration below eps records
  // the iteration count there, the iterations queued after it return at once,
  // and the host reads the word once per batch instead of waiting for every
  // residual (the segmented pull keeps the per-iteration check).
  constexpr int kSpec = 8;
  const bool spec = eps > 0 && S == 1;
  uint32_t *stop = spec ? g->buf("pr_stop", 64).as<uint32_t>() : nullptr;
  if (spec) prim::fill_u32(stop, 0, 1, s);
  int it = 0, cur = 0;
  while (it < max_iters) {
    PhaseTimer it_t(s, g->prof.on);
    unsigned nparts = grid;
    if (use_pb) {
      k_pb_gather<T><<<pb.nitems, PbThreads<T>::gather, 0, s>>>(pb.items, pb_val, pb.dl, pb_acc, n, stop);
      k_pb_scatter<T, false><<<sgrid, PbThreads<T>::scatter, 0, s>>>(pb_off, pb.sb, pb.nsb, pb.fq, pb.fs, n, pb_acc, p[cur],
                                                        p[cur ^ 1], pb_val, damping, addc, partial,
                                                        it + 1 < max_iters, stop);
      nparts = sgrid;
    } else if (per_xcd) {
      k_pr_xcd<T, kComp><<<grid / kXcds * kXcds, 256, 0, s>>>(xs.so, xs.ids_s, n, c[cur], part_s, part_c, stop);
      k_pr_xcd_combine<T, kComp><<<cgrid, 256, 0, s>>>(part_s, part_c, off, n, p[cur], p[cur ^ 1], c[cur ^ 1],
                                                       damping, addc, partial, av.deg, stop);
      nparts = cgrid;
    } else if (pma) {
      k_pr_tiles<T, kComp, true><<<grid, 256, 0, s>>>(off, ids, n, c[cur], p[cur], p[cur ^ 1], c[cur ^ 1], damping,
                                                      addc, partial, av.deg, stop);
    } else if (per_tile) {
      k_pr_tiles<T, kComp><<<grid, 256, 0, s>>>(off, ids, n, c[cur], p[cur], p[cur ^ 1], c[cur ^ 1], damping, addc,
                                                partial, nullptr, stop);
    } else {
      for (int sg = 0; sg < S; sg++) {
        const uint64_t *lo = sg == 0 ? off : pos + (uint64_t)(sg - 1) * n;
        const uint64_t *hi = sg == S - 1 ? off + 1 : pos + (uint64_t)sg * n;
        pr_pass_g<T, kComp>(G, sg == 0, sg == S - 1, lo, hi, ids, off, n, c[cur], acc_s, acc_c, p[cur],
                            p[cur ^ 1], c[cur ^ 1], damping, addc, partial, grid, s);
      }
    }
    k_sum_doubles<<<1, 1024, 0, s>>>(partial, nparts, l1dev + it, stop, eps, (uint32_t)(it + 1));
    BYO_LAUNCH_CHECK();
    if (g->prof.on) g->account(phase, it_t.stop());
    cur ^= 1;
    it++;
    if (spec) {
      if (it % kSpec == 0 || it == max_iters) {
        uint32_t *hs = reinterpret_cast<uint32_t *>(hl1);
        BYO_HIP(hipMemcpyAsync(hs, stop, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        BYO_HIP(hipStreamSynchronize(s));
        if (*hs) {
          it = (int)*hs;  // the iterate of iteration it is in p[it % 2]
          cur = it & 1;
          break;
        }
      }
    } else if (eps > 0) {
      BYO_HIP(hipMemcpyAsync(hl1, l1dev + it - 1, sizeof(double), hipMemcpyDeviceToHost, s));
      BYO_HIP(hipStreamSynchronize(s));
      if (*hl1 < eps) break;
    }
  }
  const double ms = timer.stop();
  if (device_ms) *device_ms = ms;
  if (iters_run) *iters_run = it;
  if (l1_log && it) {
    BYO_HIP(hipMemcpyAsync(l1_log, l1dev, it * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (p_out) {
    double *pd = g->buf("pr_out_f64", n * sizeof(double) + 64).as<double>();
    if (view.perm)
      k_unpermute<T><<<byo_grid(n, 256, 16384), 256, 0, s>>>(p[cur], view.perm, pd, n);
    else
      k_to_double<T><<<byo_grid(n, 256, 16384), 256, 0, s>>>(p[cur], pd, n);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(p_out, pd, n * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  BYO_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_pagerank(byo_graph_t g, double damping, int max_iters, double eps, int precision, double *p, double *l1_log,
                 int *iters_run, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    BYO_REQUIRE(max_iters >= 0, BYO_E_INVAL, "max_iters < 0");
    BYO_REQUIRE(damping >= 0 && damping <= 1, BYO_E_INVAL, "damping not in [0,1]");
    BYO_REQUIRE(precision == BYO_PR_FP32 || precision == BYO_PR_FP64, BYO_E_INVAL, "unknown precision");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    if (g->n == 0) {
      if (iters_run) *iters_run = 0;
      return;
    }
    if (precision == BYO_PR_FP32)
      pagerank_impl<float, true>(g, damping, max_iters, eps, p, l1_log, iters_run, device_ms, "pr_fp32");
    else
      pagerank_impl<double, false>(g, damping, max_iters, eps, p, l1_log, iters_run, device_ms, "pr_fp64");
  });
}

}  // extern "C"
