// relabel.hip -- degree-ordered relabelled copy of a graph, the layout the
// PageRank pull can run on (tuning knob pr_relabel).
//
// Vertices are renumbered by descending degree, ties by id: perm maps an
// original id to its new id.  The relabelled CSR keeps every
// list sorted by new id, so hubs come first in each list and hub data sits
// together at low ids.  Built once per graph on the device -- a radix sort
// of (maxdeg - deg, v) gives the order, every edge (u, v) becomes
// (perm u, perm v), a radix sort of those keys and keys_to_csr give the
// CSR -- and rebuilt after the graph changes.
#include <algorithm>

#include "primitives.hpp"

namespace byo {
namespace {

__global__ void k_rl_vkeys(const uint64_t *off, uint64_t n, uint64_t maxdeg, uint64_t *keys) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) keys[v] = ((maxdeg - (off[v + 1] - off[v])) << 32) | v;
}

__global__ void k_rl_perm(const uint64_t *sorted, uint64_t n, uint32_t *perm) {
  for (uint64_t i = global_tid(); i < n; i += grid_threads()) perm[sorted[i] & 0xFFFFFFFFull] = (uint32_t)i;
}

__global__ void k_rl_ekeys(const uint64_t *off, const uint32_t *ids, uint64_t n, const uint32_t *perm,
                           uint64_t *keys) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t u = wave; u < n; u += nwaves) {
    const uint64_t pu = (uint64_t)perm[u] << 32;
    for (uint64_t k = off[u] + lane; k < off[u + 1]; k += 64) keys[k] = pu | perm[ids[k]];
  }
}

}  // namespace

RelabelView relabel_view(byo_graph *g) {
  const uint64_t n = g->n, m = g->m;
  uint64_t *off2 = g->buf("rl_off", (n + 1) * 8 + 64).as<uint64_t>();
  uint32_t *ids2 = g->buf("rl_ids", m * 4 + 64).as<uint32_t>();
  uint32_t *perm = g->buf("rl_perm", n * 4 + 64).as<uint32_t>();
  if (!g->rl_valid && n) {
    hipStream_t s = g->stream;
    PhaseTimer t(s, g->prof.on);
    const uint64_t maxdeg = graph_max_degree(g);
    const uint64_t cap = std::max<uint64_t>(n, m);
    DevBuf k0(cap * 8 + 64), k1(cap * 8 + 64), tmp(prim::radix_tmp_bytes(cap));
    int idbits = 1, vbits = 1;
    while ((1ull << idbits) < n) idbits++;
    while ((1ull << vbits) < std::max<uint64_t>(n, maxdeg + 1)) vbits++;
    int shifts[16], vshifts[16];
    const int np = prim::edge_key_shifts(idbits, shifts);
    const int nvp = prim::edge_key_shifts(vbits, vshifts);
    k_rl_vkeys<<<byo_grid(n, 256, 16384), 256, 0, s>>>(g->off.as<uint64_t>(), n, maxdeg, k0.as<uint64_t>());
    uint64_t *sv = prim::radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), n, vshifts, nvp, tmp.p, s);
    k_rl_perm<<<byo_grid(n, 256, 16384), 256, 0, s>>>(sv, n, perm);
    k_rl_ekeys<<<byo_grid(n * 64, 256, 16384), 256, 0, s>>>(g->off.as<uint64_t>(), g->ids.as<uint32_t>(), n, perm,
                                                             k0.as<uint64_t>());
    uint64_t *se = prim::radix_sort_u64(k0.as<uint64_t>(), k1.as<uint64_t>(), m, shifts, np, tmp.p, s);
    prim::keys_to_csr(se, m, n, off2, ids2, s);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipStreamSynchronize(s));
    if (g->prof.on) g->account("relabel", t.stop());
    g->rl_valid = true;
  }
  return RelabelView{off2, ids2, perm};
}

}  // namespace byo
