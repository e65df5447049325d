// mis.hip -- maximal independent set, the edgeMap consumer MIS
// (benchmarks/MaximalIndependentSet/RandomGreedy/MaximalIndependentSet.h:
// 119-170, MaximalIndependentSet_rootset, run_all's "MIS").
//
// The rootset algorithm over a priority order: every vertex counts its
// neighbours that come before it; the vertices with none are the roots.  A
// round puts the roots in the set, removes their undecided neighbours (CAS
// on the state, get_nghs :62-67), and every removed vertex decrements the
// counters of its undecided neighbours that come after it (mis_f :79-97);
// those whose counter reaches zero are the next roots.  The result is the
// lexicographically first MIS of the order, the set sequential greedy in
// that order builds.  The reference's order is a random permutation
// (parlay::random_permutation, :127); here it is the order of the keys
// (high 32 bits of splitmix64(v), v), so the set is deterministic.
//
// One cooperative launch: the count pass, then per round a remove phase and
// a decrement phase, each a grid barrier apart.  A wave takes four queued
// vertices at a time, 16 lanes each walking one list; the vertices it adds
// to a queue are appended with one counter atomic per wave and step.  Counters
// rotate over three slots, the one two phases ahead cleared by the leader.
#include <vector>

#include "coop.hpp"

namespace byo {
namespace {

constexpr uint32_t kUndecided = 0, kIn = 1, kOut = 2;
// Lanes per queued vertex in a round's phases, and queued vertices per wave
// step (most lists are short: a wave per vertex would leave most lanes idle).
constexpr int kMisLanes = 16;
constexpr int kMisVpw = 64 / kMisLanes;

struct MisCtl {
  unsigned int cnt[3];  // vertices appended in a phase (rotating)
  unsigned int pad0;
  unsigned int rounds, decided, bad, pad1;
};

struct MisArgs {
  const uint64_t *off;
  const uint32_t *ids;
  uint64_t n;
  uint32_t *cnt;    // neighbours before v still undecided
  uint32_t *state;  // kUndecided / kIn / kOut
  uint32_t *qa, *qb, *qr;  // roots, next roots, removed
  MisCtl *ctl;
  GridBar *bar;
};

__device__ __forceinline__ uint64_t mis_key(uint32_t v) {
  uint64_t z = (uint64_t)v + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((z >> 32) << 32) | (uint64_t)v;
}

__device__ __forceinline__ unsigned int ld_u32(const unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Appends v (where take) to q at the slot's counter: one atomic per wave.
__device__ __forceinline__ void wave_append(bool take, uint32_t v, uint32_t *q, unsigned int *slot) {
  const uint64_t bal = __ballot(take);
  if (!bal) return;
  unsigned int at = 0;
  if (lane_id() == 0) at = atomicAdd(slot, (unsigned int)__popcll(bal));
  at = __shfl(at, 0, 64);
  if (take) q[at + mbcnt(bal)] = v;
}

__global__ void __launch_bounds__(kCoopBlock) k_mis_persist(MisArgs a) {
  const int lane = lane_id();
  const int msub = lane & (kMisLanes - 1), mgrp = lane / kMisLanes;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const unsigned nb = gridDim.x;
  MisCtl *c = a.ctl;
  unsigned p = 0;  // phase index: slot p % 3 is this phase's counter
  auto next_phase = [&] {
    if (leader) c->cnt[(p + 1) % 3] = 0;
    grid_barrier(a.bar, nb);
    p++;
  };
  // count pass: a lane per vertex counts the neighbours with a smaller key;
  // vertices with none are the first roots (joined right away)
  for (uint64_t base = wave * 64; base < a.n; base += nwaves * 64) {
    const uint64_t v = base + lane;
    uint32_t before = 0;
    if (v < a.n) {
      const uint64_t kv = mis_key((uint32_t)v);
      for (uint64_t e = a.off[v]; e < a.off[v + 1]; e++) before += mis_key(a.ids[e]) < kv;
      a.cnt[v] = before;
      a.state[v] = before ? kUndecided : kIn;
    }
    wave_append(v < a.n && before == 0, (uint32_t)v, a.qa, &c->cnt[p % 3]);
  }
  next_phase();
  uint32_t *q = a.qa, *qn = a.qb;
  uint32_t roots = ld_u32(&c->cnt[(p - 1) % 3]);
  uint64_t decided = 0;
  uint32_t rounds = 0;
  while (roots) {
    decided += roots;
    // remove the roots' undecided neighbours
    for (uint64_t i0 = wave * kMisVpw; i0 < roots; i0 += nwaves * kMisVpw) {
      const uint64_t i = i0 + mgrp;
      uint64_t e0 = 0, e1 = 0;
      if (i < roots) {
        const uint32_t r = q[i];
        e0 = a.off[r];
        e1 = a.off[r + 1];
      }
      const uint64_t steps = wave_max((e1 - e0 + kMisLanes - 1) / kMisLanes);
      for (uint64_t st = 0; st < steps; st++) {
        const uint64_t e = e0 + st * kMisLanes + msub;
        bool took = false;
        uint32_t d = 0;
        if (e < e1) {
          d = a.ids[e];
          took = a.state[d] == kUndecided && atomicCAS(&a.state[d], kUndecided, kOut) == kUndecided;
        }
        wave_append(took, d, a.qr, &c->cnt[p % 3]);
      }
    }
    next_phase();
    const uint32_t removed = ld_u32(&c->cnt[(p - 1) % 3]);
    decided += removed;
    // every removed vertex releases its undecided neighbours after it; the
    // decrement that reaches zero makes the neighbour a root
    for (uint64_t i0 = wave * kMisVpw; i0 < removed; i0 += nwaves * kMisVpw) {
      const uint64_t i = i0 + mgrp;
      uint64_t e0 = 0, e1 = 0, ks = 0;
      if (i < removed) {
        const uint32_t sv = a.qr[i];
        ks = mis_key(sv);
        e0 = a.off[sv];
        e1 = a.off[sv + 1];
      }
      const uint64_t steps = wave_max((e1 - e0 + kMisLanes - 1) / kMisLanes);
      for (uint64_t st = 0; st < steps; st++) {
        const uint64_t e = e0 + st * kMisLanes + msub;
        bool root = false;
        uint32_t d = 0;
        if (e < e1) {
          d = a.ids[e];
          if (a.state[d] == kUndecided && ks < mis_key(d)) {
            root = atomicSub(&a.cnt[d], 1u) == 1u;
            if (root) a.state[d] = kIn;
          }
        }
        wave_append(root, d, qn, &c->cnt[p % 3]);
      }
    }
    next_phase();
    roots = ld_u32(&c->cnt[(p - 1) % 3]);
    rounds++;
    uint32_t *t = q;
    q = qn;
    qn = t;
  }
  if (leader) {
    c->rounds = rounds;
    c->decided = (unsigned int)decided;
    c->bad = decided != a.n;  // every vertex is decided when no root is left
  }
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_mis(byo_graph_t g, uint8_t *in_mis, uint32_t *rounds, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && in_mis, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    if (rounds) *rounds = 0;
    if (device_ms) *device_ms = 0.0;
    if (n == 0) return;
    hipStream_t s = g->stream;
    MisArgs a{};
    a.off = g->off.as<uint64_t>();
    a.ids = g->ids.as<uint32_t>();
    a.n = n;
    a.cnt = g->buf("mis_cnt", n * 4 + 64).as<uint32_t>();
    a.state = g->buf("mis_state", n * 4 + 64).as<uint32_t>();
    a.qa = g->buf("mis_q0", n * 4 + 64).as<uint32_t>();
    a.qb = g->buf("mis_q1", n * 4 + 64).as<uint32_t>();
    a.qr = g->buf("mis_qr", n * 4 + 64).as<uint32_t>();
    a.ctl = g->buf("mis_ctl", sizeof(MisCtl) + 64).as<MisCtl>();
    a.bar = g->buf("mis_bar", sizeof(GridBar)).as<GridBar>();
    MisCtl *h = static_cast<MisCtl *>(pinned(g, sizeof(MisCtl)));
    BYO_HIP(hipMemsetAsync(a.bar, 0, sizeof(GridBar), s));
    BYO_HIP(hipMemsetAsync(a.ctl, 0, sizeof(MisCtl), s));
    const void *kern = reinterpret_cast<const void *>(k_mis_persist);
    const int nblocks = coop_grid(kern, kCoopBlock, (int)g->tuning("coop_blocks_per_cu"));
    void *args[] = {&a};
    PhaseTimer t(s, true);
    BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kCoopBlock), args, 0, s));
    const double ms = t.stop();
    BYO_HIP(hipMemcpyAsync(h, a.ctl, sizeof(MisCtl), hipMemcpyDeviceToHost, s));
    std::vector<uint32_t> st(n);
    BYO_HIP(hipMemcpyAsync(st.data(), a.state, n * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(!h->bad, BYO_E_STATE, "MIS rounds ended with undecided vertices");
    for (uint64_t v = 0; v < n; v++) in_mis[v] = st[v] == kIn;
    if (g->prof.on) g->account("mis", ms);
    if (device_ms) *device_ms = ms;
    if (rounds) *rounds = h->rounds;
  });
}

}  // extern "C"
