// tc.hip -- degree-ordered triangle counting (benchmarks/TriangleCounting/
// ShunTangwongsan15/Triangle.h:57-204).
//
// 1. rank by (degree, id) -- rankNodes sorts by degree; any tie order gives
//    the same count, and (degree, id) makes the orientation deterministic;
// 2. filterGraph(G, rank[u] < rank[v]) (gbbs/interface.h:81-99 ->
//    graph_mutation.h:15-81): count pass, exclusive scan, order-preserving
//    fill, giving the oriented graph DG (m/2 ids, each list sorted);
// 3. count sum_{u->v in DG} |N+(u) & N+(v)| (CountDirectedBalanced,
//    Triangle.h:99-150, with intersect.h's sorted-list intersection): one
//    wave per u stages N+(u) in LDS; the pairs (v, w in N+(v)) of all
//    v in N+(u) are flattened over the lanes and each w is looked up in the
//    staged list by binary search; the total is a uint64.
#include <algorithm>

#include "primitives.hpp"

namespace byo {
namespace {

__device__ __forceinline__ bool rank_less(const uint64_t *off, uint32_t u, uint64_t du, uint32_t v) {
  const uint64_t dv = off[v + 1] - off[v];
  return du < dv || (du == dv && u < v);
}

constexpr int kG = 16;  // lanes per vertex in the orientation passes

__global__ void k_tc_outdeg(const uint64_t *off, const uint32_t *ids, uint64_t n, uint64_t *cnt) {
  const int sub = lane_id() & (kG - 1);
  const uint64_t group = global_tid() / kG, ngroups = grid_threads() / kG;
  const uint64_t nrounds = (n + ngroups - 1) / ngroups;
  for (uint64_t r = 0; r < nrounds; r++) {
    const uint64_t u = r * ngroups + group;
    uint64_t c = 0;
    if (u < n) {
      const uint64_t b = off[u], e = off[u + 1], du = e - b;
      for (uint64_t k = b + sub; k < e; k += kG) c += rank_less(off, (uint32_t)u, du, ids[k]);
    }
#pragma unroll
    for (int o = kG / 2; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
    if (u < n && sub == 0) cnt[u] = c;
  }
}

__global__ void k_tc_fill(const uint64_t *off, const uint32_t *ids, uint64_t n, const uint64_t *dg_off,
                          uint32_t *dg_ids) {
  const int lane = lane_id();
  const int sub = lane & (kG - 1);
  const int gshift = lane & ~(kG - 1);
  const uint64_t group = global_tid() / kG, ngroups = grid_threads() / kG;
  const uint64_t nrounds = (n + ngroups - 1) / ngroups;
  for (uint64_t r = 0; r < nrounds; r++) {
    const uint64_t u = r * ngroups + group;
    const bool active = u < n;
    const uint64_t b = active ? off[u] : 0, e = active ? off[u + 1] : 0, du = e - b;
    uint64_t w = active ? dg_off[u] : 0;
    // every lane of the wave runs the same number of steps (ballot below)
    uint64_t steps = (du + kG - 1) / kG;
    steps = wave_max(steps);
    for (uint64_t s = 0; s < steps; s++) {
      const uint64_t k = b + s * kG + sub;
      bool keep = false;
      uint32_t v = 0;
      if (k < e) {
        v = ids[k];
        keep = rank_less(off, (uint32_t)u, du, v);
      }
      const uint64_t mask = __ballot(keep);
      const uint64_t gm = (mask >> gshift) & ((1ull << kG) - 1);
      if (keep) dg_ids[w + __popcll(gm & ((1ull << sub) - 1))] = v;
      w += __popcll(gm);
    }
  }
}

constexpr int kLdsPerWave = 2048;  // staged N+(u) entries per wave (8 KiB)
constexpr int kTcWaves = 4;

__device__ __forceinline__ bool find_sorted(const uint32_t *a, uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < len && a[lo] == x;
}

// The same search over the wave's LDS stage, indexed directly: a pointer
// that may point to LDS or to global memory would make every probe a flat
// load.
template <int N>
__device__ __forceinline__ bool find_sorted_lds(const uint32_t (&a)[N], uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1; else hi = mid;
  }
  return lo < len && a[lo] == x;
}

__global__ void __launch_bounds__(256) k_tc_count(const uint64_t *dg_off, const uint32_t *dg_ids, uint64_t n,
                                                  unsigned long long *next_u, unsigned long long *total) {
  __shared__ uint32_t stage[kTcWaves][kLdsPerWave];
  __shared__ uint64_t pre[kTcWaves][64];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  uint64_t mycount = 0;
  constexpr unsigned long long kGrab = 8;  // vertices claimed per atomic
  unsigned long long ubase = 0, uoff = kGrab;
  while (true) {
    if (uoff == kGrab) {
      if (lane == 0) ubase = atomicAdd(next_u, kGrab);
      ubase = __shfl(ubase, 0, 64);
      uoff = 0;
    }
    const unsigned long long u = ubase + uoff++;
    if (u >= n) break;
    const uint64_t bu = dg_off[u], du = dg_off[u + 1] - bu;
    if (du < 2) continue;  // a triangle needs two out-neighbours of u
    const bool in_lds = du <= kLdsPerWave;
    if (in_lds)
      for (uint64_t i = lane; i < du; i += 64) stage[wv][i] = dg_ids[bu + i];
    // LDS writes by this wave must land before its lanes read them
    wave_lds_sync();
    for (uint64_t c0 = 0; c0 < du; c0 += 64) {
      const uint64_t j = c0 + lane;
      uint32_t v = 0;
      uint64_t dv = 0, bv = 0;
      if (j < du) {
        v = dg_ids[bu + j];
        bv = dg_off[v];
        dv = dg_off[v + 1] - bv;
      }
      const uint64_t incl = wave_incl_scan(dv);
      pre[wv][lane] = incl;
      const uint64_t T = __shfl(incl, 63, 64);
      wave_lds_sync();
      for (uint64_t x0 = 0; x0 < T; x0 += 64) {
        const uint64_t x = x0 + lane;
        // owner slot of pair x: first s with pre[s] > x (all lanes then
        // fetch the owner's list start with one shuffle)
        int owner = 0;
        uint64_t startx = 0;
        if (x < T) {
          int lo = 0, hi = 63;
          while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (pre[wv][mid] > x) hi = mid; else lo = mid + 1;
          }
          owner = lo;
          startx = lo ? pre[wv][lo - 1] : 0;
        }
        const uint64_t obv = __shfl(bv, owner, 64);
        if (x < T) {
          const uint32_t w = dg_ids[obv + (x - startx)];
          if (in_lds)
            mycount += find_sorted_lds(stage[wv], (uint32_t)du, w);
          else
            mycount += find_sorted(dg_ids + bu, (uint32_t)du, w);
        }
      }
      wave_lds_sync();  // pre[] is rewritten by the next chunk
    }
    wave_lds_sync();  // stage[] is rewritten by the next u
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
    ensure_csr(g);
    const uint64_t n = g->n;
    *count = 0;
    if (n == 0 || g->m == 0) return;
    hipStream_t s = g->stream;
    const uint64_t *off = g->off.as<uint64_t>();
    const uint32_t *ids = g->ids.as<uint32_t>();
    uint64_t *cnt = g->buf("tc_cnt", (n + 1) * 8 + 64).as<uint64_t>();
    uint64_t *dg_off = g->buf("tc_dgoff", (n + 2) * 8 + 64).as<uint64_t>();
    uint32_t *dg_ids = g->buf("tc_dgids", (g->m / 2 + 64) * 4).as<uint32_t>();
    void *stmp = g->buf("tc_scan_tmp", prim::scan_tmp_bytes(n + 1)).p;
    unsigned long long *ctr = g->buf("tc_ctr", 64).as<unsigned long long>();
    PhaseTimer t(s, true);
    const unsigned grid = byo_grid(n * kG, 256, 16384);
    k_tc_outdeg<<<grid, 256, 0, s>>>(off, ids, n, cnt);
    prim::exclusive_scan_u64(cnt, dg_off, n, true, stmp, s);
    k_tc_fill<<<grid, 256, 0, s>>>(off, ids, n, dg_off, dg_ids);
    BYO_HIP(hipMemsetAsync(ctr, 0, 16, s));
    BYO_LAUNCH_CHECK();
    PhaseTimer tc(s, g->prof.on);
    k_tc_count<<<2048, 256, 0, s>>>(dg_off, dg_ids, n, ctr, ctr + 1);
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
