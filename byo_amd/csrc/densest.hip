// densest.hip -- approximate densest subgraph by peeling, the edgeMap
// consumer WorkEfficientDensestSubgraph (benchmarks/ApproximateDensestSubgraph/
// ApproxPeelingBKV12/DensestSubgraph.h:8-145, run_all's
// "WorkEfficientDensestSubgraph", eps 0.001).
//
// Rounds over the alive set R, D[v] the alive neighbours of v: the density
// is sum_R D / |R| (both directions of an edge counted, as the reference
// does) and the largest over the rounds is the result; every alive vertex
// with D[v] <= (1 + eps) * density leaves R, and each of its alive
// neighbours loses one (nghCount, :64-73).  The sums are kept exact in
// integers: a round's removals subtract their D, its decrements subtract
// one each, so no pass re-reduces D; the density and threshold are the
// reference's double expressions of those integers, and the result is
// bit-identical to a sequential run.
//
// One cooperative launch, two grid barriers per round (removal, then
// decrements).  Removals are appended with one counter atomic per wave and
// their D summed per wave; a decrement phase serves four removed vertices
// per wave step, 16 lanes each.
#include "coop.hpp"

namespace byo {
namespace {

constexpr int kDsLanes = 16;
constexpr int kDsVpw = 64 / kDsLanes;

struct DsCtl {
  unsigned int cnt[3];  // vertices removed in a phase (rotating)
  unsigned int pad0;
  unsigned long long sum[3];  // D removed / decrements applied in a phase (rotating)
  double best;
  unsigned int rounds, pad1;
};

struct DsArgs {
  const uint64_t *off;
  const uint32_t *ids;
  uint64_t n, m;
  double eps;
  uint32_t *D;
  uint32_t *alive;
  uint32_t *q;
  DsCtl *ctl;
  GridBar *bar;
};

__device__ __forceinline__ unsigned int ld_u32(const unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long ld_u64(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(kCoopBlock) k_densest_persist(DsArgs a) {
  const int lane = lane_id();
  const int dsub = lane & (kDsLanes - 1), dgrp = lane / kDsLanes;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const unsigned nb = gridDim.x;
  DsCtl *c = a.ctl;
  unsigned p = 0;
  auto next_phase = [&] {
    if (leader) {
      c->cnt[(p + 1) % 3] = 0;
      c->sum[(p + 1) % 3] = 0;
    }
    grid_barrier(a.bar, nb);
    p++;
  };
  for (uint64_t v = global_tid(); v < a.n; v += grid_threads()) {
    a.D[v] = (uint32_t)(a.off[v + 1] - a.off[v]);
    a.alive[v] = 1;
  }
  next_phase();
  uint64_t edges = a.m, left = a.n;
  double best = 0.0;
  uint32_t rounds = 0;
  while (left > 0) {
    const double density = (double)edges / (double)left;
    const double target = ((1.0 + a.eps) * (double)edges) / (double)left;
    if (density > best) best = density;
    // removal: alive vertices at or below the threshold
    unsigned long long dsum = 0;
    for (uint64_t base = wave * 64; base < a.n; base += nwaves * 64) {
      const uint64_t v = base + lane;
      bool take = false;
      uint32_t dv = 0;
      if (v < a.n && a.alive[v]) {
        dv = a.D[v];
        take = (double)dv <= target;
        if (take) a.alive[v] = 0;
      }
      if (take) dsum += dv;
      const uint64_t bal = __ballot(take);
      if (bal) {
        unsigned int at = 0;
        if (lane == 0) at = atomicAdd(&c->cnt[p % 3], (unsigned int)__popcll(bal));
        at = __shfl(at, 0, 64);
        if (take) a.q[at + mbcnt(bal)] = (uint32_t)v;
      }
    }
    dsum = wave_sum(dsum);
    if (lane == 0 && dsum) atomicAdd(&c->sum[p % 3], dsum);
    next_phase();
    const uint32_t removed = ld_u32(&c->cnt[(p - 1) % 3]);
    edges -= ld_u64(&c->sum[(p - 1) % 3]);
    left -= removed;
    // decrements: each alive neighbour of a removed vertex loses one
    unsigned long long decs = 0;
    if (left > 0) {
      for (uint64_t i0 = wave * kDsVpw; i0 < removed; i0 += nwaves * kDsVpw) {
        const uint64_t i = i0 + dgrp;
        uint64_t e0 = 0, e1 = 0;
        if (i < removed) {
          const uint32_t u = a.q[i];
          e0 = a.off[u];
          e1 = a.off[u + 1];
        }
        for (uint64_t e = e0 + dsub; e < e1; e += kDsLanes) {
          const uint32_t v = a.ids[e];
          if (a.alive[v]) {
            atomicSub(&a.D[v], 1u);
            decs++;
          }
        }
      }
    }
    decs = wave_sum(decs);
    if (lane == 0 && decs) atomicAdd(&c->sum[p % 3], decs);
    next_phase();
    edges -= ld_u64(&c->sum[(p - 1) % 3]);
    rounds++;
  }
  if (leader) {
    c->best = best;
    c->rounds = rounds;
  }
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_densest_subgraph(byo_graph_t g, double eps, double *density, uint32_t *rounds, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && density, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(eps >= 0.0, BYO_E_INVAL, "eps < 0");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    *density = 0.0;
    if (rounds) *rounds = 0;
    if (device_ms) *device_ms = 0.0;
    if (n == 0) return;
    hipStream_t s = g->stream;
    DsArgs a{};
    a.off = g->off.as<uint64_t>();
    a.ids = g->ids.as<uint32_t>();
    a.n = n;
    a.m = g->m;
    a.eps = eps;
    a.D = g->buf("ds_deg", n * 4 + 64).as<uint32_t>();
    a.alive = g->buf("ds_alive", n * 4 + 64).as<uint32_t>();
    a.q = g->buf("ds_q", n * 4 + 64).as<uint32_t>();
    a.ctl = g->buf("ds_ctl", sizeof(DsCtl) + 64).as<DsCtl>();
    a.bar = g->buf("ds_bar", sizeof(GridBar)).as<GridBar>();
    DsCtl *h = static_cast<DsCtl *>(pinned(g, sizeof(DsCtl)));
    BYO_HIP(hipMemsetAsync(a.bar, 0, sizeof(GridBar), s));
    BYO_HIP(hipMemsetAsync(a.ctl, 0, sizeof(DsCtl), s));
    const void *kern = reinterpret_cast<const void *>(k_densest_persist);
    const int nblocks = coop_grid(kern, kCoopBlock, (int)g->tuning("coop_blocks_per_cu"));
    void *args[] = {&a};
    PhaseTimer t(s, true);
    BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kCoopBlock), args, 0, s));
    const double ms = t.stop();
    BYO_HIP(hipMemcpyAsync(h, a.ctl, sizeof(DsCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    if (g->prof.on) g->account("densest", ms);
    *density = h->best;
    if (rounds) *rounds = h->rounds;
    if (device_ms) *device_ms = ms;
  });
}

}  // extern "C"
