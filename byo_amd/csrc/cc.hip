// cc.hip -- connected components with min-id labels.
//
// byo_cc restates SimpleUnionAsync (benchmarks/Connectivity/
// SimpleUnionAsync/Connectivity.h:32-136): lock-free union-find where a
// link always hangs the larger root under the smaller one (CAS on the root's
// parent), so every tree's root is its smallest vertex and the compressed
// parent is the component's minimum id -- the reference's output exactly.
// Scheduling is MI355X-shaped rather than one task per vertex: first the
// first two neighbours of every vertex are linked (a cheap sample), the
// most frequent root is found from 1024 sampled vertices, and the remaining
// edges are processed only from vertices outside that component (for a
// symmetric graph every edge leaving the giant component is seen from its
// other end).  The labels do not depend on the schedule.
//
// Both read the container's own adjacency (AdjView): a CSR, or the CPMA's
// slot index and dst words, whose empty slots (kNone) are skipped.
//
// byo_cc_labelprop restates labelprop_cc::CC (LabelPropagation/
// Connectivity.h:78-146): rounds of dense-forward min propagation over the
// changed set until it is empty.
#include <algorithm>
#include <unordered_map>

#include "primitives.hpp"

namespace byo {
namespace {

__device__ __forceinline__ uint32_t uf_find(uint32_t *par, uint32_t x) {
  while (true) {
    uint32_t p = par[x];
    if (p == x) return x;
    uint32_t gp = par[p];
    if (gp != p) par[x] = gp;  // path halving; gp is an ancestor of x
    x = gp;
  }
}

__device__ __forceinline__ void uf_unite(uint32_t *par, uint32_t u, uint32_t v) {
  while (true) {
    u = uf_find(par, u);
    v = uf_find(par, v);
    if (u == v) return;
    if (u < v) {
      uint32_t t = u;
      u = v;
      v = t;
    }
    // u > v: hang root u under v
    uint32_t old = atomicCAS(&par[u], u, v);
    if (old == u) return;
    u = old;
  }
}

__global__ void k_cc_init(uint32_t *par, uint64_t n) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) par[v] = (uint32_t)v;
}

__global__ void k_cc_link_sample(const uint64_t *off, const uint32_t *ids, uint64_t n, uint32_t *par, int r) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    const uint64_t b = off[v];
    if (off[v + 1] - b > (uint64_t)r) {
      const uint32_t u = ids[b + r];
      if (u != kNone) uf_unite(par, (uint32_t)v, u);
    }
  }
}

__global__ void k_cc_compress(uint32_t *par, uint64_t n) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) par[v] = uf_find(par, (uint32_t)v);
}

__global__ void k_cc_gather(const uint32_t *par, uint64_t n, uint32_t *out, int k) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) out[i] = par[((uint64_t)i * 2654435761ull + 12345ull) % n];
}

// Remaining edges of vertices outside the giant component.  A wave takes 64
// vertices, one per lane, finds which are outside the giant component (the
// finds run side by side), then gives each of those its neighbours strided
// over the lanes.  (A vertex the wave's earlier unites have just joined to
// the giant component is still processed: unites are idempotent.)
__global__ void k_cc_link_rest(const uint64_t *off, const uint32_t *ids, uint64_t n, uint32_t *par, uint32_t giant,
                               int r0) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t base = wave * 64; base < n; base += nwaves * 64) {
    const uint64_t mine = base + lane;
    unsigned long long out = __ballot(mine < n && uf_find(par, (uint32_t)mine) != giant);
    while (out) {
      const uint64_t v = base + __builtin_ctzll(out);
      out &= out - 1;
      const uint64_t b = off[v] + r0, e = off[v + 1];
      for (uint64_t k = b + lane; k < e; k += 64) {
        const uint32_t u = ids[k];
        if (u != kNone) uf_unite(par, (uint32_t)v, u);
      }
    }
  }
}

// ------------------------------------------------------- label propagation
__global__ void k_lp_init(uint32_t *lab, uint64_t n, uint64_t *fbm64, uint64_t nw) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) lab[v] = (uint32_t)v;
  for (uint64_t w = global_tid(); w < nw; w += grid_threads()) {
    uint64_t lo = w * 64;
    fbm64[w] = (lo + 64 <= n) ? ~0ull : (lo < n ? ((1ull << (n - lo)) - 1) : 0ull);
  }
}

// edgeMapDenseForward with lp_f (LabelPropagation/Connectivity.h:45-76):
// a frontier vertex u lowers each neighbour's label to label[u]; a lowered
// neighbour joins the next frontier.
__global__ void k_lp_round(const uint64_t *off, const uint32_t *ids, uint64_t n, uint32_t *lab, const uint64_t *fbm64,
                           uint64_t *nbm64, unsigned long long *changed) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const uint64_t nw = (n + 63) >> 6;
  for (uint64_t w = wave; w < nw; w += nwaves) {
    uint64_t word = fbm64[w];
    while (word) {
      const int bp = __builtin_ctzll(word);
      word &= word - 1;
      const uint64_t u = w * 64 + bp;
      const uint32_t lu = __hip_atomic_load(&lab[u], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      bool any = false;
      for (uint64_t k = off[u] + lane; k < off[u + 1]; k += 64) {
        const uint32_t t = ids[k];
        if (t != kNone && lu < lab[t]) {
          const uint32_t old = atomicMin(&lab[t], lu);
          if (lu < old) {
            atomicOr(reinterpret_cast<unsigned long long *>(&nbm64[t >> 6]), 1ull << (t & 63));
            any = true;
          }
        }
      }
      if (__ballot(any) && lane == 0) atomicAdd(changed, 1ull);
    }
  }
}

__global__ void k_clear64(uint64_t *p, uint64_t n) {
  for (uint64_t i = global_tid(); i < n; i += grid_threads()) p[i] = 0;
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_cc(byo_graph_t g, uint32_t *labels, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    const uint64_t n = g->n;
    if (n == 0) return;
    const AdjView av = adj_view(g);
    hipStream_t s = g->stream;
    uint32_t *par = g->buf("cc_par", n * 4 + 64).as<uint32_t>();
    uint32_t *samp = g->buf("cc_samp", 1024 * 4).as<uint32_t>();
    const uint64_t *off = av.off;
    const uint32_t *ids = av.ids;
    const unsigned grid = byo_grid(n, 256, 16384);
    PhaseTimer t(s, true);
    k_cc_init<<<grid, 256, 0, s>>>(par, n);
    const int kSample = 2;
    for (int r = 0; r < kSample; r++) {
      k_cc_link_sample<<<grid, 256, 0, s>>>(off, ids, n, par, r);
      k_cc_compress<<<grid, 256, 0, s>>>(par, n);
    }
    const int ks = (int)std::min<uint64_t>(1024, n);
    k_cc_gather<<<(ks + 255) / 256, 256, 0, s>>>(par, n, samp, ks);
    BYO_LAUNCH_CHECK();
    uint32_t *hs = static_cast<uint32_t *>(pinned(g, 1024 * 4));
    BYO_HIP(hipMemcpyAsync(hs, samp, ks * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    std::unordered_map<uint32_t, int> freq;
    uint32_t giant = hs[0];
    int best = 0;
    for (int i = 0; i < ks; i++) {
      int c = ++freq[hs[i]];
      if (c > best || (c == best && hs[i] < giant)) {
        best = c;
        giant = hs[i];
      }
    }
    k_cc_link_rest<<<byo_grid(n * 64, 256, 16384), 256, 0, s>>>(off, ids, n, par, giant, kSample);
    k_cc_compress<<<grid, 256, 0, s>>>(par, n);
    BYO_LAUNCH_CHECK();
    double ms = t.stop();
    if (g->prof.on) g->account("cc", ms);
    if (device_ms) *device_ms = ms;
    if (labels) BYO_HIP(hipMemcpyAsync(labels, par, n * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
  });
}

int byo_cc_labelprop(byo_graph_t g, uint32_t *labels, uint32_t *rounds, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    const uint64_t n = g->n;
    if (n == 0) {
      if (rounds) *rounds = 0;
      return;
    }
    const AdjView av = adj_view(g);
    hipStream_t s = g->stream;
    const uint64_t nw = (n + 63) / 64 + 1;
    uint32_t *lab = g->buf("lp_lab", n * 4 + 64).as<uint32_t>();
    uint64_t *f0 = g->buf("lp_f0", nw * 8).as<uint64_t>();
    uint64_t *f1 = g->buf("lp_f1", nw * 8).as<uint64_t>();
    unsigned long long *chg = g->buf("lp_chg", 64).as<unsigned long long>();
    unsigned long long *h = static_cast<unsigned long long *>(pinned(g, 64));
    const uint64_t *off = av.off;
    const uint32_t *ids = av.ids;
    PhaseTimer t(s, true);
    k_lp_init<<<byo_grid(n, 256, 16384), 256, 0, s>>>(lab, n, f0, nw);
    uint32_t r = 0;
    while (true) {
      k_clear64<<<byo_grid(nw, 256, 4096), 256, 0, s>>>(f1, nw);
      BYO_HIP(hipMemsetAsync(chg, 0, 8, s));
      k_lp_round<<<8192, 256, 0, s>>>(off, ids, n, lab, f0, f1, chg);
      BYO_LAUNCH_CHECK();
      BYO_HIP(hipMemcpyAsync(h, chg, 8, hipMemcpyDeviceToHost, s));
      BYO_HIP(hipStreamSynchronize(s));
      r++;
      std::swap(f0, f1);
      if (h[0] == 0) break;
    }
    double ms = t.stop();
    if (g->prof.on) g->account("cc_labelprop", ms);
    if (device_ms) *device_ms = ms;
    if (rounds) *rounds = r;
    if (labels) BYO_HIP(hipMemcpyAsync(labels, lab, n * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
