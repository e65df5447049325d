// tc.hip -- degree-ordered triangle counting (benchmarks/TriangleCounting/
// ShunTangwongsan15/Triangle.h:57-204).
//
// 1. rank by (min(degree, 65535), id) -- rankNodes sorts by degree; any
//    strict total order gives the same count, and this one makes the
//    orientation deterministic;
// 2. filterGraph(G, rank[u] < rank[v]) (gbbs/interface.h:81-99 ->
//    graph_mutation.h:15-81), order-preserving, giving the oriented graph
//    DG (m/2 edges of a symmetric graph, each list sorted); its lists are
//    packed in place at the front of each CSR list (one pass, no count or
//    scan), so u's list is dg_ids[off[u] .. dg_end[u]);
// 3. count sum_{u->v in DG} |N+(u) & N+(v)| (CountDirectedBalanced,
//    Triangle.h:99-150, with intersect.h's sorted-list intersection): one
//    wave per u puts N+(u) in an LDS hash table; the pairs (v, w in N+(v))
//    of all v in N+(u) are flattened over the lanes, four per lane with
//    their loads in flight together, and each w is looked up in the table;
//    the total is a uint64.  (Membership of w in N+(u) is what the sorted
//    intersection counts, so the count is the same.)
#include <algorithm>

#include "primitives.hpp"

namespace byo {
namespace {

// Degree of every vertex as 16 bits, degrees of 65535 or more held as
// 65535: (that degree, id) is still a strict total order, which is all the
// orientation needs for a correct count, and it orders by degree every
// vertex below the cap (vertices above it, among themselves, by id).  The
// rank tests then gather from a table a quarter the size of the offsets.
__global__ void k_tc_deg(const uint64_t *off, const uint32_t *degs, uint64_t n, uint16_t *deg) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    const uint64_t d = degs ? (uint64_t)degs[v] : off[v + 1] - off[v];
    deg[v] = d < 0xFFFFull ? (uint16_t)d : (uint16_t)0xFFFFu;
  }
}

constexpr int kG = 16;      // lanes per vertex in the orientation pass
constexpr int kOrientU = 8;  // steps of kG edges whose loads are in flight together

// filterGraph in one pass, in the container's own index space (CSR
// positions, or PMA slots whose empty entries, kNone, are dropped): the edges
// of u that DG keeps are written, in adjacency order, to dg_ids[off[u] ...],
// and dg_end[u] = off[u] + their number.  Every lane of the wave runs the same
// number of steps (ballots below); the ids and degrees of kOrientU steps are
// loaded at clamped positions before any is used.
__global__ void k_tc_orient(const uint64_t *off, const uint32_t *ids, const uint16_t *deg, uint64_t n,
                            uint32_t *dg_ids, uint64_t *dg_end) {
  const int lane = lane_id();
  const int sub = lane & (kG - 1);
  const int gshift = lane & ~(kG - 1);
  const uint64_t group = global_tid() / kG, ngroups = grid_threads() / kG;
  const uint64_t nrounds = (n + ngroups - 1) / ngroups;
  for (uint64_t r = 0; r < nrounds; r++) {
    const uint64_t u = r * ngroups + group;
    const bool active = u < n;
    const uint64_t b = active ? off[u] : 0, e = active ? off[u + 1] : 0;
    const uint32_t du = active ? deg[u] : 0;
    const uint64_t last = e > b ? e - 1 : 0;  // ids[0] exists: the graph has edges
    uint64_t w = b;
    const uint64_t steps = wave_max((e - b + kG - 1) / kG);
    for (uint64_t s0 = 0; s0 < steps; s0 += kOrientU) {
      uint32_t v[kOrientU], dv[kOrientU];
      bool hole[kOrientU];
#pragma unroll
      for (int t = 0; t < kOrientU; t++) {
        v[t] = ids[std::min<uint64_t>(b + (s0 + t) * kG + sub, last)];
        hole[t] = v[t] == kNone;  // an empty PMA slot
        if (hole[t]) v[t] = 0;
      }
#pragma unroll
      for (int t = 0; t < kOrientU; t++) dv[t] = deg[v[t]];
#pragma unroll
      for (int t = 0; t < kOrientU; t++) {
        const uint64_t k = b + (s0 + t) * kG + sub;
        const bool kp = k < e && !hole[t] && (du < dv[t] || (du == dv[t] && (uint32_t)u < v[t]));
        const uint64_t mask = __ballot(kp);
        const uint64_t gm = (mask >> gshift) & ((1ull << kG) - 1);
        if (kp) dg_ids[w + __popcll(gm & ((1ull << sub) - 1))] = v[t];
        w += __popcll(gm);
      }
    }
    if (active && sub == 0) dg_end[u] = w;
  }
}

// N+(u) is held in an LDS hash table per wave (open addressing, linear
// probing) of S slots, S the smallest power of two >= 4|N+(u)| and >= 64,
// so a lookup costs about one LDS read where a binary search over the list
// costs log2 |N+(u)| dependent reads; the load of at most 1/4 keeps the
// longest probe of a wave's 64 lanes short, which is what the wave waits
// for.  Lists longer than kHashCap are searched in global memory instead.
constexpr int kSlots = 1024;  // hash slots per wave (4 KiB)
constexpr uint32_t kHashCap = kSlots / 4;
constexpr uint32_t kEmpty = 0xFFFFFFFFu;
constexpr int kTcWaves = 4;
constexpr int kPairsPerLane = 4;  // pair lookups per lane with loads in flight together
constexpr int kWin = 1024;        // pairs per owner table (16 per lane)

__device__ __forceinline__ bool find_sorted(const uint32_t *a, uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < len && a[lo] == x;
}

// First slot of x's probe sequence: a multiple of 4, so the sequence
// visits aligned groups of 4 slots and a lookup reads one group per step.
__device__ __forceinline__ uint32_t slot_of(uint32_t x, uint32_t mask) {
  return ((x * 0x9E3779B1u) >> 11) << 2 & mask;  // Fibonacci hashing, bits 11..31
}

__global__ void __launch_bounds__(256) k_tc_count(const uint64_t *off, const uint64_t *dg_end, const uint32_t *dg_ids,
                                                  uint64_t n,
                                                  unsigned long long *next_u, unsigned long long *total) {
  __shared__ __align__(16) uint32_t tab[kTcWaves][kSlots];
  __shared__ int64_t base[kTcWaves][64];
  __shared__ __align__(16) uint8_t own[kTcWaves][kWin];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  uint64_t mycount = 0;
  constexpr unsigned kGrab = 32;  // vertices claimed per atomic, one per lane
  for (;;) {
    unsigned long long ubase = 0;
    if (lane == 0) ubase = atomicAdd(next_u, (unsigned long long)kGrab);
    ubase = __shfl(ubase, 0, 64);
    if (ubase >= n) break;
    // lane l < kGrab holds vertex ubase + l: its list start and DG degree
    uint64_t myb = 0, myd = 0;
    if (lane < (int)kGrab && ubase + lane < n) {
      myb = off[ubase + lane];
      myd = dg_end[ubase + lane] - myb;
    }
    uint64_t todo = __ballot(myd >= 2);  // a triangle needs two out-neighbours of u
    while (todo) {
      const int l = __builtin_ctzll(todo);
      todo &= todo - 1;
      const uint64_t bu = __shfl(myb, l, 64), du = __shfl(myd, l, 64);
      const bool hashed = du <= kHashCap;
      uint32_t S = 64;
      while (S < 4 * du && S < (uint32_t)kSlots) S <<= 1;
      const uint32_t mask = S - 1;
      if (hashed) {
        for (uint32_t i = lane; i < S; i += 64) tab[wv][i] = kEmpty;
        wave_lds_sync();
      }
      for (uint64_t c0 = 0; c0 < du; c0 += 64) {
        const uint64_t j = c0 + lane;
        uint64_t dv = 0, bv = 0;
        if (j < du) {
          const uint32_t v = dg_ids[bu + j];
          bv = off[v];
          dv = dg_end[v] - bv;
          // ids of one list are distinct, so every insert ends in an empty slot
          if (hashed && c0 == 0) {
            uint32_t h = slot_of(v, mask);
            while (atomicCAS(&tab[wv][h], kEmpty, v) != kEmpty) h = (h + 1) & mask;
          }
        }
        if (c0 == 0) {
          if (hashed)
            for (uint64_t i = 64 + lane; i < du; i += 64) {
              const uint32_t x = dg_ids[bu + i];
              uint32_t h = slot_of(x, mask);
              while (atomicCAS(&tab[wv][h], kEmpty, x) != kEmpty) h = (h + 1) & mask;
            }
          // LDS writes by this wave must land before its lanes read them
          wave_lds_sync();
        }
        // the chunk's pairs laid end to end: pair x of slot s (x in
        // [excl_s, incl_s)) reads dg_ids[base[s] + x]
        const uint64_t incl = wave_incl_scan(dv);
        const uint64_t excl = incl - dv;
        const uint64_t T = __shfl(incl, 63, 64);
        base[wv][lane] = (int64_t)bv - (int64_t)excl;
        for (uint64_t X = 0; X < T; X += kWin) {
          // owner table of pairs [X, X + kWin): each slot marks where its run
          // starts (slot + 1), a prefix maximum fills the rest; the slot that
          // owns pair X when its run started earlier is the carry-in
          const uint64_t Xe = std::min<uint64_t>(X + kWin, T);
          const uint64_t held = __ballot(dv > 0 && excl <= X && X < incl);
          const uint32_t carry = held ? (uint32_t)__builtin_ctzll(held) + 1 : 0;
          reinterpret_cast<uint4 *>(own[wv])[lane] = make_uint4(0, 0, 0, 0);
          wave_lds_sync();
          if (dv > 0 && excl > X && excl < Xe) own[wv][excl - X] = (uint8_t)(lane + 1);
          wave_lds_sync();
          {
            uint4 q = reinterpret_cast<uint4 *>(own[wv])[lane];
            uint32_t wd[4] = {q.x, q.y, q.z, q.w};
            uint32_t run = 0;  // running maximum over this lane's 16 bytes
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
              for (int k = 0; k < 4; k++) run = std::max(run, (wd[i] >> (8 * k)) & 0xFFu);
            // maximum over the lanes before this one, then the carry-in
            uint32_t lm = run;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
              const uint32_t t = __shfl_up(lm, o, 64);
              if (lane >= o) lm = std::max(lm, t);
            }
            uint32_t cur = __shfl_up(lm, 1, 64);
            cur = std::max(lane ? cur : 0u, carry);
#pragma unroll
            for (int i = 0; i < 4; i++) {
              uint32_t outw = 0;
#pragma unroll
              for (int k = 0; k < 4; k++) {
                cur = std::max(cur, (wd[i] >> (8 * k)) & 0xFFu);
                outw |= cur << (8 * k);
              }
              wd[i] = outw;
            }
            reinterpret_cast<uint4 *>(own[wv])[lane] = make_uint4(wd[0], wd[1], wd[2], wd[3]);
          }
          wave_lds_sync();
          for (uint64_t x0 = X; x0 < Xe; x0 += 64 * kPairsPerLane) {
            uint32_t w[kPairsPerLane];
#pragma unroll
            for (int t = 0; t < kPairsPerLane; t++) {
              const uint64_t x = x0 + (uint64_t)t * 64 + lane;
              const uint64_t xc = x < Xe ? x : X;  // past the end: reread pair X
              const int o = (int)own[wv][xc - X] - 1;
              w[t] = dg_ids[(uint64_t)(base[wv][o] + (int64_t)xc)];
            }
            if (hashed) {
              // the lane's probes advance together, one LDS read each per
              // step (the table is indexed directly, so no probe becomes a
              // flat load)
              uint32_t h[kPairsPerLane];
              bool live[kPairsPerLane];
#pragma unroll
              for (int t = 0; t < kPairsPerLane; t++) {
                live[t] = x0 + (uint64_t)t * 64 + lane < Xe;
                h[t] = slot_of(w[t], mask);
              }
              for (;;) {
                bool any = false;
#pragma unroll
                for (int t = 0; t < kPairsPerLane; t++) {
                  if (!live[t]) continue;
                  const uint4 y = *reinterpret_cast<const uint4 *>(&tab[wv][h[t]]);
                  if (y.x == w[t] || y.y == w[t] || y.z == w[t] || y.w == w[t]) {
                    mycount++;
                    live[t] = false;
                  } else if (y.w == kEmpty) {  // slots fill in order, so an empty slot ends the group's run
                    live[t] = false;
                  } else {
                    h[t] = (h[t] + 4) & mask;
                  }
                  any |= live[t];
                }
                if (!any) break;
              }
            } else {
#pragma unroll
              for (int t = 0; t < kPairsPerLane; t++)
                if (x0 + (uint64_t)t * 64 + lane < Xe) mycount += find_sorted(dg_ids + bu, (uint32_t)du, w[t]);
            }
          }
          wave_lds_sync();  // own[] is rewritten by the next window
        }
        wave_lds_sync();  // base[] is rewritten by the next chunk
      }
      wave_lds_sync();  // tab[] is rewritten by the next u
    }
  }
  mycount = wave_sum(mycount);
  if (lane == 0 && mycount) atomicAdd(total, (unsigned long long)mycount);
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_tc(byo_graph_t g, uint64_t *count, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && count, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    const uint64_t n = g->n;
    *count = 0;
    if (n == 0 || g->m == 0) return;
    hipStream_t s = g->stream;
    const AdjView av = adj_view(g);
    const uint64_t *off = av.off;
    const uint32_t *ids = av.ids;
    // DG lives in the container's index space: u's kept edges at off[u] ... dg_end[u]
    uint32_t *dg_ids = g->buf("tc_dgids", (av.nslots + 64) * 4).as<uint32_t>();
    uint64_t *dg_end = g->buf("tc_dgend", n * 8 + 64).as<uint64_t>();
    uint16_t *deg = g->buf("tc_deg", n * 2 + 64).as<uint16_t>();
    unsigned long long *ctr = g->buf("tc_ctr", 64).as<unsigned long long>();
    PhaseTimer t(s, true);
    k_tc_deg<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, av.deg, n, deg);
    k_tc_orient<<<byo_grid(n * kG, 256, 16384), 256, 0, s>>>(off, ids, deg, n, dg_ids, dg_end);
    BYO_HIP(hipMemsetAsync(ctr, 0, 16, s));
    BYO_LAUNCH_CHECK();
    PhaseTimer tc(s, g->prof.on);
    k_tc_count<<<2048, 256, 0, s>>>(off, dg_end, dg_ids, n, ctr, ctr + 1);
    BYO_LAUNCH_CHECK();
    if (g->prof.on) g->account("tc_count", tc.stop());
    double ms = t.stop();
    if (g->prof.on) g->account("tc", ms);
    if (device_ms) *device_ms = ms;
    uint64_t *h = static_cast<uint64_t *>(pinned(g, 64));
    BYO_HIP(hipMemcpyAsync(h, ctr + 1, 8, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    *count = h[0];
  });
}

}  // extern "C"
