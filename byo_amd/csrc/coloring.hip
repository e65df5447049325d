// coloring.hip -- greedy graph colouring, the edgeMap consumer Coloring
// (benchmarks/GraphColoring/Hasenplaugh14/GraphColoring.h:71-146, run_all's
// "Coloring" with its default LLF order).
//
// Every vertex counts its neighbours that come before it in the LLF order
// (larger log2_up(degree) first, ties by a permutation, :97-112); a round
// colours the roots (count 0) and decrements the counters of their
// uncoloured neighbours (coloring_f :56-69), the decrement to zero making
// the next roots.  A root's colour follows coloring::color (:30-55): the
// smallest colour in [0, deg) no neighbour holds, deg + 1 when all of
// [0, deg) are taken, 0 when isolated.  The roots of a round are never
// adjacent and a root's coloured neighbours are exactly the ones before
// it, so the result is greedy colouring in LLF order.  The reference's
// tie-break is a random permutation (parlay::random_permutation); here it
// is the order of the keys (high 32 bits of splitmix64(v), v), so the
// colouring is deterministic.
//
// One cooperative launch: the count pass, then one grid barrier per round
// (a root's colour depends only on earlier rounds' colours, so colouring
// and decrementing share a phase).  A wave takes one root at a time: it
// marks its neighbours' colours in a window of 4096 colours (an LDS bitmap
// of 64 words, one per lane), takes the first free slot by a wave minimum,
// and moves to the next window only when this one is full.
#include "coop.hpp"

namespace byo {
namespace {

constexpr uint32_t kNoColor = 0xFFFFFFFFu;
constexpr uint32_t kColWin = 64 * 64;  // colours per window (one bit each)

struct ColCtl {
  unsigned int cnt[3];  // roots appended in a phase (rotating)
  unsigned int pad0;
  unsigned int rounds, colored, bad, pad1;
};

struct ColArgs {
  const uint64_t *off;
  const uint32_t *ids;
  uint64_t n;
  uint32_t *cnt;    // neighbours before v not yet coloured
  uint32_t *color;  // kNoColor until coloured
  uint32_t *qa, *qb;
  ColCtl *ctl;
  GridBar *bar;
};

__device__ __forceinline__ uint64_t perm_key(uint32_t v) {
  uint64_t z = (uint64_t)v + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return ((z >> 32) << 32) | (uint64_t)v;
}

// parlay::log2_up: ceil(log2 x), 0 for x <= 1.
__device__ __forceinline__ uint32_t log2_up(uint64_t x) { return x <= 1 ? 0u : 64u - (uint32_t)__clzll(x - 1); }

__device__ __forceinline__ unsigned int ld_u32(const unsigned int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void wave_append(bool take, uint32_t v, uint32_t *q, unsigned int *slot) {
  const uint64_t bal = __ballot(take);
  if (!bal) return;
  unsigned int at = 0;
  if (lane_id() == 0) at = atomicAdd(slot, (unsigned int)__popcll(bal));
  at = __shfl(at, 0, 64);
  if (take) q[at + mbcnt(bal)] = v;
}

__global__ void __launch_bounds__(kCoopBlock) k_color_persist(ColArgs a) {
  __shared__ unsigned long long bm[kCoopWaves][64];
  const int lane = lane_id(), wv = threadIdx.x >> 6;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const unsigned nb = gridDim.x;
  ColCtl *c = a.ctl;
  unsigned p = 0;
  auto next_phase = [&] {
    if (leader) c->cnt[(p + 1) % 3] = 0;
    grid_barrier(a.bar, nb);
    p++;
  };
  // count pass: a lane per vertex counts the neighbours before it
  for (uint64_t base = wave * 64; base < a.n; base += nwaves * 64) {
    const uint64_t v = base + lane;
    uint32_t before = 0;
    if (v < a.n) {
      const uint64_t b = a.off[v], e = a.off[v + 1];
      const uint32_t lv = log2_up(e - b);
      const uint64_t kv = perm_key((uint32_t)v);
      for (uint64_t k = b; k < e; k++) {
        const uint32_t u = a.ids[k];
        const uint32_t lu = log2_up(a.off[u + 1] - a.off[u]);
        before += lu > lv || (lu == lv && perm_key(u) < kv);
      }
      a.cnt[v] = before;
      a.color[v] = kNoColor;
    }
    wave_append(v < a.n && before == 0, (uint32_t)v, a.qa, &c->cnt[p % 3]);
  }
  next_phase();
  uint32_t *q = a.qa, *qn = a.qb;
  uint32_t roots = ld_u32(&c->cnt[(p - 1) % 3]);
  uint64_t colored = 0;
  uint32_t rounds = 0;
  while (roots) {
    colored += roots;
    for (uint64_t i = wave; i < roots; i += nwaves) {
      const uint32_t r = q[i];
      const uint64_t e0 = a.off[r], e1 = a.off[r + 1], deg = e1 - e0;
      // colour: first free slot of [0, deg), window by window
      uint64_t col = deg ? deg + 1 : 0;
      for (uint64_t w0 = 0; w0 < deg; w0 += kColWin) {
        bm[wv][lane] = 0;
        wave_lds_sync();
        const uint64_t w1 = w0 + kColWin < deg ? w0 + kColWin : deg;
        for (uint64_t c0 = e0; c0 < e1; c0 += 64) {
          const uint64_t e = c0 + lane;
          if (e < e1) {
            const uint32_t cu = a.color[a.ids[e]];
            if (cu >= w0 && cu < w1)
              atomicOr(&bm[wv][(cu - w0) >> 6], 1ull << ((cu - w0) & 63));
          }
        }
        wave_lds_sync();
        // lane l holds colours w0 + 64 l ... w0 + 64 l + 63; bits past w1 count as taken
        const uint64_t lo = w0 + 64 * (uint64_t)lane;
        uint64_t word = bm[wv][lane];
        if (lo >= w1) word = ~0ull;
        else if (w1 - lo < 64) word |= ~0ull << (w1 - lo);
        const uint64_t freec = word == ~0ull ? ~0ull : lo + (uint64_t)__builtin_ctzll(~word);
        const uint64_t best = wave_min(freec);
        wave_lds_sync();  // bm is rewritten by the next window or root
        if (best != ~0ull) {
          col = best;
          break;
        }
      }
      if (lane == 0) a.color[r] = (uint32_t)col;
      // release the uncoloured neighbours (all after r)
      for (uint64_t c0 = e0; c0 < e1; c0 += 64) {
        const uint64_t e = c0 + lane;
        bool root = false;
        uint32_t d = 0;
        if (e < e1) {
          d = a.ids[e];
          if (a.color[d] == kNoColor) root = atomicSub(&a.cnt[d], 1u) == 1u;
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
    c->colored = (unsigned int)colored;
    c->bad = colored != a.n;  // every vertex is coloured when no root is left
  }
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_coloring(byo_graph_t g, uint32_t *color, uint32_t *rounds, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && color, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    if (rounds) *rounds = 0;
    if (device_ms) *device_ms = 0.0;
    if (n == 0) return;
    hipStream_t s = g->stream;
    ColArgs a{};
    a.off = g->off.as<uint64_t>();
    a.ids = g->ids.as<uint32_t>();
    a.n = n;
    a.cnt = g->buf("col_cnt", n * 4 + 64).as<uint32_t>();
    a.color = g->buf("col_color", n * 4 + 64).as<uint32_t>();
    a.qa = g->buf("col_q0", n * 4 + 64).as<uint32_t>();
    a.qb = g->buf("col_q1", n * 4 + 64).as<uint32_t>();
    a.ctl = g->buf("col_ctl", sizeof(ColCtl) + 64).as<ColCtl>();
    a.bar = g->buf("col_bar", sizeof(GridBar)).as<GridBar>();
    ColCtl *h = static_cast<ColCtl *>(pinned(g, sizeof(ColCtl)));
    BYO_HIP(hipMemsetAsync(a.bar, 0, sizeof(GridBar), s));
    BYO_HIP(hipMemsetAsync(a.ctl, 0, sizeof(ColCtl), s));
    const void *kern = reinterpret_cast<const void *>(k_color_persist);
    const int nblocks = coop_grid(kern, kCoopBlock, (int)g->tuning("coop_blocks_per_cu"));
    void *args[] = {&a};
    PhaseTimer t(s, true);
    BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kCoopBlock), args, 0, s));
    const double ms = t.stop();
    BYO_HIP(hipMemcpyAsync(h, a.ctl, sizeof(ColCtl), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipMemcpyAsync(color, a.color, n * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(!h->bad, BYO_E_STATE, "colouring rounds ended with uncoloured vertices");
    if (g->prof.on) g->account("coloring", ms);
    if (device_ms) *device_ms = ms;
    if (rounds) *rounds = h->rounds;
  });
}

}  // extern "C"
