// kcore.hip -- coreness of every vertex, the edgeMap consumer KCore
// (benchmarks/KCore/JulienneDBS17/KCore.h:8-56, run_all's "KCore"): the
// reference takes the lowest non-empty degree bucket k, removes its
// vertices, and lowers each live neighbour's degree by the edges removed,
// floored at k, so the degree a vertex holds when it is removed is its
// coreness.
//
// Device peel in one cooperative launch, by rounds.  The frontier is the
// live vertices of degree <= k; a round has every wave take four frontier
// vertices at a time and decrement their live neighbours with atomics.  The thread
// whose decrement takes a neighbour from k+1 to k pushes it to the next
// frontier and marks it dead with coreness k right away -- it would be
// removed next round, and later decrements can only lower its degree, which
// no longer matters -- so each round's frontier is all live vertices of
// degree <= k without a scan, and a round costs one grid barrier.  An empty
// frontier moves k to the smallest live degree (a reduction) and refills
// the frontier with one scan.  Counters rotate over three slots, the one
// two phases ahead cleared by the leader, so no phase needs a reset barrier.
#include <vector>

#include "coop.hpp"

namespace byo {
namespace {

struct KcCtl {
  unsigned int cnt[3];   // vertices pushed in a phase (rotating)
  unsigned int pad0;
  unsigned int mn[3];    // smallest live degree (rotating)
  unsigned int pad1;
  unsigned int kmax, rounds, bad, pad2;
};

struct KcArgs {
  const uint64_t *off;
  const uint32_t *ids;
  uint64_t n;
  uint32_t *D, *core, *qa, *qb;
  uint8_t *dead;
  KcCtl *ctl;
  GridBar *bar;
};

// Lanes per frontier vertex in a round, and frontier vertices per wave.
constexpr int kKcLanes = 16;
constexpr int kKcVpw = 64 / kKcLanes;

__device__ __forceinline__ unsigned int ld_u32(const unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void __launch_bounds__(kCoopBlock) k_kc_persist(KcArgs a) {
  const int lane = lane_id();
  const int ksub = lane & (kKcLanes - 1), kgrp = lane / kKcLanes;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const unsigned nb = gridDim.x;
  KcCtl *c = a.ctl;
  for (uint64_t v = global_tid(); v < a.n; v += grid_threads()) {
    a.D[v] = (uint32_t)(a.off[v + 1] - a.off[v]);
    a.dead[v] = 0;
  }
  grid_barrier(a.bar, nb);
  uint32_t k = 0, cnt = 0, kmax = 0, rounds = 0;
  uint64_t removed = 0;
  uint32_t *q = a.qa, *qn = a.qb;
  unsigned p = 0;  // phase index: slot p % 3 is this phase's counter
  auto next_phase = [&] {
    if (leader) {
      c->cnt[(p + 1) % 3] = 0;
      c->mn[(p + 1) % 3] = 0xFFFFFFFFu;
    }
    grid_barrier(a.bar, nb);
    p++;
  };
  while (removed < a.n) {
    if (cnt == 0) {
      // next bucket: the smallest live degree ... (scans keep 8 vertices
      // per lane in flight)
      constexpr int U = 8;
      uint32_t m = 0xFFFFFFFFu;
      for (uint64_t base = wave * 64 * U; base < a.n; base += nwaves * 64 * U) {
        uint8_t dv[U];
        uint32_t Dv[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
          const uint64_t v = min(base + (uint64_t)j * 64 + lane, a.n - 1);
          dv[j] = a.dead[v];
          Dv[j] = a.D[v];
        }
#pragma unroll
        for (int j = 0; j < U; j++)
          if (!dv[j]) m = min(m, Dv[j]);
      }
      m = wave_min(m);
      if (lane == 0 && m != 0xFFFFFFFFu) atomicMin(&c->mn[p % 3], m);
      next_phase();
      m = ld_u32(&c->mn[(p - 1) % 3]);
      if (m == 0xFFFFFFFFu) {
        if (leader) c->bad = 1;  // vertices left but none live: cannot happen on a valid CSR
        break;
      }
      k = max(k, m);
      // ... and its vertices, dead with coreness k
      for (uint64_t base = wave * 64 * U; base < a.n; base += nwaves * 64 * U) {
        uint8_t dv[U];
        uint32_t Dv[U];
#pragma unroll
        for (int j = 0; j < U; j++) {
          const uint64_t v = min(base + (uint64_t)j * 64 + lane, a.n - 1);
          dv[j] = a.dead[v];
          Dv[j] = a.D[v];
        }
#pragma unroll
        for (int j = 0; j < U; j++) {
          const uint64_t v = base + (uint64_t)j * 64 + lane;
          const bool take = v < a.n && !dv[j] && Dv[j] <= k;
          const uint64_t bal = __ballot(take);
          if (!bal) continue;
          unsigned int at = 0;
          if (lane == 0) at = atomicAdd(&c->cnt[p % 3], (unsigned int)__popcll(bal));
          at = __shfl(at, 0, 64);
          if (take) {
            q[at + mbcnt(bal)] = (uint32_t)v;
            a.dead[v] = 1;
            a.core[v] = k;
          }
        }
      }
      next_phase();
      cnt = ld_u32(&c->cnt[(p - 1) % 3]);
    }
    kmax = max(kmax, k);
    // one round: each frontier vertex gets 16 lanes of a wave (four per
    // wave step) that decrement its live neighbours, and the vertices that
    // reach k are appended with one counter atomic per wave and step (one
    // per push would serialise a round's millions of pushes on that one
    // address).  Four lists per instruction fill the lanes that short lists
    // left idle under a wave per vertex (Erdos-Renyi 1.4x faster); serving
    // 64 frontier vertices per instruction instead makes the decrements
    // collide on the hubs they share (2.4x slower on RMAT, even with the
    // appends aggregated).  Dead neighbours are skipped to save atomics; a stale
    // flag is harmless, as a dead vertex's degree no longer matters and,
    // degrees only falling, one that died at level k0 <= k cannot pass from
    // k+1 to k again.
    for (uint64_t i0 = wave * kKcVpw; i0 < cnt; i0 += nwaves * kKcVpw) {
      // kKcVpw frontier vertices per wave, kKcLanes lanes each; a step
      // serves kKcLanes edges of every one of them
      const uint64_t i = i0 + kgrp;
      uint64_t e0 = 0, e1 = 0;
      if (i < cnt) {
        const uint32_t u = q[i];
        e0 = a.off[u];
        e1 = a.off[u + 1];
      }
      const uint64_t steps = wave_max((e1 - e0 + kKcLanes - 1) / kKcLanes);
      for (uint64_t st = 0; st < steps; st++) {
        const uint64_t e = e0 + st * kKcLanes + ksub;
        bool push = false;
        uint32_t v = 0;
        if (e < e1) {
          v = a.ids[e];
          push = !a.dead[v] && atomicSub(&a.D[v], 1u) == k + 1;
        }
        const uint64_t bal = __ballot(push);
        if (!bal) continue;
        unsigned int at = 0;
        if (lane == 0) at = atomicAdd(&c->cnt[p % 3], (unsigned int)__popcll(bal));
        at = __shfl(at, 0, 64);
        if (push) {
          qn[at + mbcnt(bal)] = v;
          a.dead[v] = 1;
          a.core[v] = k;
        }
      }
    }
    next_phase();
    removed += cnt;
    rounds++;
    cnt = ld_u32(&c->cnt[(p - 1) % 3]);

    uint32_t *t = q;
    q = qn;
    qn = t;
  }
  if (leader) {
    c->kmax = kmax;
    c->rounds = rounds;
  }
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_kcore(byo_graph_t g, uint32_t *core, uint32_t *k_max, uint64_t *rounds, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && core, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    hipStream_t s = g->stream;
    KcArgs a{};
    a.off = g->off.as<uint64_t>();
    a.ids = g->ids.as<uint32_t>();
    a.n = n;
    a.D = g->buf("kc_deg", n * 4 + 64).as<uint32_t>();
    a.core = g->buf("kc_core", n * 4 + 64).as<uint32_t>();
    a.qa = g->buf("kc_q0", n * 4 + 64).as<uint32_t>();
    a.qb = g->buf("kc_q1", n * 4 + 64).as<uint32_t>();
    a.dead = g->buf("kc_dead", n + 64).as<uint8_t>();
    a.ctl = g->buf("kc_ctl", sizeof(KcCtl) + 64).as<KcCtl>();
    a.bar = g->buf("kc_bar", sizeof(GridBar)).as<GridBar>();
    KcCtl *h = static_cast<KcCtl *>(pinned(g, sizeof(KcCtl)));
    BYO_HIP(hipMemsetAsync(a.bar, 0, sizeof(GridBar), s));
    BYO_HIP(hipMemsetAsync(a.ctl, 0, sizeof(KcCtl), s));
    BYO_HIP(hipMemsetAsync(a.ctl->mn, 0xFF, sizeof(a.ctl->mn), s));
    const void *kern = reinterpret_cast<const void *>(k_kc_persist);
    // latency-bound per frontier vertex: as many resident waves as fit
    const int nblocks = coop_grid(kern, kCoopBlock, (int)g->tuning("kcore_blocks_per_cu"));
    void *args[] = {&a};
    PhaseTimer t(s, true);
    BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kCoopBlock), args, 0, s));
    const double ms = t.stop();
    BYO_HIP(hipMemcpyAsync(h, a.ctl, sizeof(KcCtl), hipMemcpyDeviceToHost, s));
    if (n) BYO_HIP(hipMemcpyAsync(core, a.core, n * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(!h->bad, BYO_E_STATE, "k-core peel found no live vertex while vertices remained");
    if (g->prof.on) g->account("kcore", ms);
    if (device_ms) *device_ms = ms;
    if (k_max) *k_max = h->kmax;
    if (rounds) *rounds = h->rounds;
  });
}

}  // extern "C"
