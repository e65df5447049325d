// spanner.hip -- the MPX spanner, the edgeMap consumer Spanner
// (benchmarks/Spanner/MPXV15/Spanner.h:242-261, run_all's "Spanner", k = 4,
// beta = ln(n) / (2k) as Spanner_runner sets it).
//
// A low-diameter decomposition gives clusters; the spanner is every
// vertex's edge to its parent in its cluster's BFS tree plus one edge per
// pair of adjacent clusters (tree_and_intercluster_edges, :147-176).  The
// reference permutes the centres and lets CAS races pick parents and the
// edge of a cluster pair; here the choices are fixed: the LDD of ldd.hip
// (centres in id order, smallest cluster on ties), v's parent the smallest
// neighbour in v's cluster claimed one round before v, and for clusters
// c1 < c2 the edge (u, v), cl[u] = c1, cl[v] = c2, smallest as u << 32 | v.
// The smallest edge of each pair is kept by an open-addressing table keyed
// by the pair (atomicMin on the edge); the result is sorted by key.
#include <cmath>

#include "primitives.hpp"

namespace byo {
namespace {

constexpr uint64_t kEmptyPair = ~0ull;

__device__ __forceinline__ void wave_append64(bool take, uint64_t x, uint64_t *out, unsigned long long *cnt) {
  const uint64_t bal = __ballot(take);
  if (!bal) return;
  unsigned long long at = 0;
  if (lane_id() == 0) at = atomicAdd(cnt, (unsigned long long)__popcll(bal));
  at = __shfl(at, 0, 64);
  if (take) out[at + mbcnt(bal)] = x;
}

// Tree edges (v, parent) of the non-centres, appended to out.
__global__ void k_sp_tree(const uint64_t *off, const uint32_t *ids, uint64_t n, const uint32_t *cl,
                          const uint32_t *rd, uint64_t *out, unsigned long long *cnt) {
  for (uint64_t base = (global_tid() >> 6) * 64; base < n; base += (grid_threads() >> 6) * 64) {
    const uint64_t v = base + lane_id();
    bool take = false;
    uint64_t key = 0;
    if (v < n && cl[v] != v) {
      uint32_t best = kNone;
      for (uint64_t e = off[v]; e < off[v + 1]; e++) {
        const uint32_t u = ids[e];
        if (cl[u] == cl[v] && rd[u] + 1 == rd[v] && u < best) best = u;
      }
      take = best != kNone;  // an LDD claim always has such a neighbour
      if (!take) atomicOr(reinterpret_cast<unsigned int *>(cnt + 2), 1u);
      key = (v << 32) | best;
    }
    wave_append64(take, key, out, cnt);
  }
}

// Number of edge slots u -> v with cl[u] < cl[v].
__global__ void k_sp_count(const uint64_t *off, const uint32_t *ids, uint64_t n, const uint32_t *cl,
                           unsigned long long *cnt) {
  unsigned long long c = 0;
  for (uint64_t u = global_tid(); u < n; u += grid_threads())
    for (uint64_t e = off[u]; e < off[u + 1]; e++) c += cl[u] < cl[ids[e]];
  c = wave_sum(c);
  if (lane_id() == 0 && c) atomicAdd(cnt, c);
}

__global__ void k_sp_table_init(uint64_t *tkey, unsigned long long *tval, uint64_t T) {
  for (uint64_t i = global_tid(); i < T; i += grid_threads()) {
    tkey[i] = kEmptyPair;
    tval[i] = ~0ull;
  }
}

__device__ __forceinline__ uint64_t pair_slot(uint64_t p, uint64_t mask) {
  p ^= p >> 33;
  p *= 0xFF51AFD7ED558CCDull;
  p ^= p >> 33;
  return p & mask;
}

// Every edge u -> v with cl[u] < cl[v] keeps the smallest edge of its pair.
__global__ void k_sp_insert(const uint64_t *off, const uint32_t *ids, uint64_t n, const uint32_t *cl,
                            uint64_t *tkey, unsigned long long *tval, uint64_t mask) {
  for (uint64_t u = global_tid(); u < n; u += grid_threads()) {
    const uint32_t cu = cl[u];
    for (uint64_t e = off[u]; e < off[u + 1]; e++) {
      const uint32_t v = ids[e];
      const uint32_t cv = cl[v];
      if (cu >= cv) continue;
      const uint64_t p = ((uint64_t)cu << 32) | cv, edge = (u << 32) | v;
      // the table holds at least twice as many slots as edges, so every
      // probe ends at the pair's slot or an empty one
      for (uint64_t h = pair_slot(p, mask);; h = (h + 1) & mask) {
        uint64_t k = tkey[h];
        if (k == kEmptyPair) k = atomicCAS(reinterpret_cast<unsigned long long *>(&tkey[h]), kEmptyPair, p);
        if (k == kEmptyPair || k == p) {
          atomicMin(&tval[h], (unsigned long long)edge);
          break;
        }
      }
    }
  }
}

__global__ void k_sp_collect(const uint64_t *tkey, const unsigned long long *tval, uint64_t T, uint64_t *out,
                             unsigned long long *cnt) {
  for (uint64_t base = (global_tid() >> 6) * 64; base < T; base += (grid_threads() >> 6) * 64) {
    const uint64_t i = base + lane_id();
    const bool take = i < T && tkey[i] != kEmptyPair;
    wave_append64(take, take ? (uint64_t)tval[i] : 0, out, cnt);
  }
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_spanner(byo_graph_t g, uint32_t k, uint64_t *edges, uint64_t capacity, uint64_t *count,
                double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && count, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(k >= 1, BYO_E_INVAL, "k must be at least 1");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    *count = 0;
    if (device_ms) *device_ms = 0.0;
    if (n < 2) return;  // no edges to keep
    const double beta = std::log((double)n) / (2.0 * (double)k);
    BYO_REQUIRE(beta < 64, BYO_E_INVAL, "k too small for this graph");
    hipStream_t s = g->stream;
    const uint64_t *off = g->off.as<uint64_t>();
    const uint32_t *ids = g->ids.as<uint32_t>();
    PhaseTimer t(s, true);
    uint32_t rounds = 0;
    ldd_run(g, beta, &rounds);
    const uint32_t *cl = g->buf("ldd_cl", n * 4 + 64).as<uint32_t>();
    const uint32_t *rd = g->buf("ldd_rd", n * 4 + 64).as<uint32_t>();
    unsigned long long *ctr = g->buf("sp_ctr", 64).as<unsigned long long>();
    unsigned long long *hc = static_cast<unsigned long long *>(pinned(g, 64));
    BYO_HIP(hipMemsetAsync(ctr, 0, 24, s));  // edge count, candidate count, no-parent flag
    k_sp_count<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, ids, n, cl, ctr + 1);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(hc, ctr + 1, 8, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    const uint64_t cand = hc[0];
    uint64_t T = 1024;
    while (T < 2 * cand) T <<= 1;
    uint64_t *out = g->buf("sp_out", (n + cand) * 8 + 64).as<uint64_t>();
    uint64_t *alt = g->buf("sp_alt", (n + cand) * 8 + 64).as<uint64_t>();
    uint64_t *tkey = g->buf("sp_tkey", T * 8 + 64).as<uint64_t>();
    unsigned long long *tval = g->buf("sp_tval", T * 8 + 64).as<unsigned long long>();
    k_sp_tree<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, ids, n, cl, rd, out, ctr);
    k_sp_table_init<<<byo_grid(T, 256, 16384), 256, 0, s>>>(tkey, tval, T);
    k_sp_insert<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, ids, n, cl, tkey, tval, T - 1);
    k_sp_collect<<<byo_grid(T, 256, 16384), 256, 0, s>>>(tkey, tval, T, out, ctr);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(hc, ctr, 24, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    BYO_REQUIRE(hc[2] == 0, BYO_E_STATE, "a clustered vertex has no neighbour one round nearer its centre");
    const uint64_t total = hc[0];
    BYO_REQUIRE(total <= n + cand, BYO_E_STATE, "spanner edge count out of range");
    int idbits = 1;
    while ((1ull << idbits) < n) idbits++;
    int shifts[16];
    const int np = prim::edge_key_shifts(idbits, shifts);
    void *tmp = g->buf("sp_sort_tmp", prim::radix_tmp_bytes(total + 1)).p;
    const uint64_t *sorted = total ? prim::radix_sort_u64(out, alt, total, shifts, np, tmp, s) : out;
    const double ms = t.stop();
    *count = total;
    if (g->prof.on) g->account("spanner", ms);
    if (device_ms) *device_ms = ms;
    if (edges) {
      BYO_REQUIRE(capacity >= total, BYO_E_INVAL, "edges buffer too small (count holds the size needed)");
      if (total) BYO_HIP(hipMemcpyAsync(edges, sorted, total * 8, hipMemcpyDeviceToHost, s));
    }
    BYO_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
