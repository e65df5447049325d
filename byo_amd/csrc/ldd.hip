// ldd.hip -- low-diameter decomposition, the edgeMap consumer LDD of
// run_all (benchmarks/LowDiameterDecomposition/MPX13/
// LowDiameterDecomposition.h:98-186 with beta 0.2 and permute false):
// clusters grown by BFS from centres that start at staggered rounds.
//
// Round r first makes every still unclaimed vertex in [shifts[r],
// min(shifts[r+1], n)) a centre (shifts: exclusive prefix sums of
// floor(exp(i * beta)), generate_shifts :12-25), then one BFS step from the
// frontier -- last step's claims plus the new centres -- claims their
// unclaimed neighbours.  A vertex reached from several clusters in one step
// joins the smallest cluster id, one of the outcomes of the reference's CAS
// race, so the result is deterministic.  The step is a push from the
// frontier list (compare-and-swap of the claim round, then atomicMin of the
// cluster) when the frontier's edges are few, else a pull over the
// unclaimed vertices (minimum cluster among neighbours claimed last round),
// by the M/20 rule of edge_map_data.h:291.
#include <cmath>
#include <vector>

#include "common.hpp"

namespace byo {
namespace {

struct LddCtl {
  unsigned int cnt;          // vertices appended to the list being built
  unsigned int pad0;
  unsigned long long deg;    // sum of their degrees
  unsigned long long pad[6];
};

__global__ void k_ldd_init(uint64_t n, uint32_t *cl, uint32_t *rd) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    cl[v] = kNone;
    rd[v] = kNone;
  }
}

// New centres of round r, appended after the frontier already in q (one
// counter atomic per wave, degrees summed per wave).
__global__ void k_ldd_centres(const uint64_t *off, uint64_t start, uint64_t end, uint32_t r, uint32_t *cl,
                              uint32_t *rd, uint32_t *q, LddCtl *ctl) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  unsigned long long dsum = 0;
  for (uint64_t base = start + wave * 64; base < end; base += nwaves * 64) {
    const uint64_t v = base + lane;
    const bool take = v < end && rd[v] == kNone;
    const uint64_t bal = __ballot(take);
    if (!bal) continue;
    unsigned int at = 0;
    if (lane == 0) at = atomicAdd(&ctl->cnt, (unsigned int)__popcll(bal));
    at = __shfl(at, 0, 64);
    if (take) {
      rd[v] = r;
      cl[v] = (uint32_t)v;
      q[at + mbcnt(bal)] = (uint32_t)v;
      dsum += off[v + 1] - off[v];
    }
  }
  dsum = wave_sum(dsum);
  if (lane == 0 && dsum) atomicAdd(&ctl->deg, dsum);
}

// A wave per frontier vertex, 64 neighbours at a time; claims are appended
// with one counter atomic per wave and step, degrees summed per wave.
__global__ void k_ldd_push(const uint64_t *off, const uint32_t *ids, const uint32_t *q, uint32_t cnt, uint32_t r,
                           uint32_t *cl, uint32_t *rd, uint32_t *qn, LddCtl *ctl) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  unsigned long long dsum = 0;
  for (uint64_t i = wave; i < cnt; i += nwaves) {
    const uint32_t u = q[i];
    const uint32_t c = cl[u];
    const uint64_t e1 = off[u + 1];
    for (uint64_t c0 = off[u]; c0 < e1; c0 += 64) {
      const uint64_t e = c0 + lane;
      bool won = false;
      uint32_t v = 0;
      if (e < e1) {
        v = ids[e];
        const uint32_t seen = rd[v];
        if (seen == kNone || seen == r + 1) {
          const uint32_t old = atomicCAS(&rd[v], kNone, r + 1);
          won = old == kNone;
          if (won || old == r + 1) atomicMin(&cl[v], c);
        }
      }
      const uint64_t bal = __ballot(won);
      if (!bal) continue;
      unsigned int at = 0;
      if (lane == 0) at = atomicAdd(&ctl->cnt, (unsigned int)__popcll(bal));
      at = __shfl(at, 0, 64);
      if (won) {
        qn[at + mbcnt(bal)] = v;
        dsum += off[v + 1] - off[v];
      }
    }
  }
  dsum = wave_sum(dsum);
  if (lane == 0 && dsum) atomicAdd(&ctl->deg, dsum);
}

// Pull: every unclaimed vertex takes the smallest cluster among its
// neighbours claimed in round r (a claim made now stamps r + 1, never r).
__global__ void k_ldd_pull(const uint64_t *off, const uint32_t *ids, uint64_t n, uint32_t r, uint32_t *cl,
                           uint32_t *rd, uint32_t *qn, LddCtl *ctl) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  unsigned long long dsum = 0;
  for (uint64_t base = wave * 64; base < n; base += nwaves * 64) {
    const uint64_t v = base + lane;
    uint32_t best = kNone;
    if (v < n && rd[v] == kNone) {
      for (uint64_t e = off[v]; e < off[v + 1]; e++) {
        const uint32_t u = ids[e];
        if (rd[u] == r) best = min(best, cl[u]);
      }
    }
    const bool take = best != kNone;
    const uint64_t bal = __ballot(take);
    if (!bal) continue;
    unsigned int at = 0;
    if (lane == 0) at = atomicAdd(&ctl->cnt, (unsigned int)__popcll(bal));
    at = __shfl(at, 0, 64);
    if (take) {
      rd[v] = r + 1;
      cl[v] = best;
      qn[at + mbcnt(bal)] = (uint32_t)v;
      dsum += off[v + 1] - off[v];
    }
  }
  dsum = wave_sum(dsum);
  if (lane == 0 && dsum) atomicAdd(&ctl->deg, dsum);
}

__global__ void k_ldd_ctl(LddCtl *ctl, uint32_t cnt) {
  ctl->cnt = cnt;
  ctl->deg = 0;
}

}  // namespace

// The LDD rounds; clusters and claim rounds stay in device buffers "ldd_cl"
// and "ldd_rd" (u32[n]).  Returns the device time; *rounds_run the rounds.
// The caller holds the handle's lock with the CSR view current.
double ldd_run(byo_graph *g, double beta, uint32_t *rounds_run) {
  const uint64_t n = g->n, m = g->m;
  hipStream_t s = g->stream;
  // generate_shifts: floor(exp(i * beta)) for i < last, exclusive scan
  std::vector<uint64_t> shifts;
  if (n) {
    const uint64_t last = (uint64_t)std::fmin((double)(n + 1), 2.0 + std::ceil(std::log((double)n) / beta));
    shifts.assign(last + 1, 0);
    uint64_t acc = 0;
    for (uint64_t i = 0; i <= last; i++) {
      const uint64_t x = i < last ? (uint64_t)std::floor(std::exp((double)i * beta)) : 0;
      shifts[i] = acc;
      acc += x;
    }
  }
  uint32_t *cl = g->buf("ldd_cl", n * 4 + 64).as<uint32_t>();
  uint32_t *rd = g->buf("ldd_rd", n * 4 + 64).as<uint32_t>();
  uint32_t *q = g->buf("ldd_q0", n * 4 + 64).as<uint32_t>();
  uint32_t *qn = g->buf("ldd_q1", n * 4 + 64).as<uint32_t>();
  LddCtl *ctl = g->buf("ldd_ctl", sizeof(LddCtl) + 64).as<LddCtl>();
  LddCtl *h = static_cast<LddCtl *>(pinned(g, sizeof(LddCtl)));
  const uint64_t *off = g->off.as<uint64_t>();
  const uint32_t *ids = g->ids.as<uint32_t>();
  PhaseTimer t(s, true);
  uint64_t visited = 0, nf = 0, fdeg = 0;
  uint32_t r = 0;
  if (n) {
    k_ldd_init<<<byo_grid(n, 256, 16384), 256, 0, s>>>(n, cl, rd);
    for (r = 0; visited < n; r++) {
      BYO_REQUIRE(r + 1 < shifts.size(), BYO_E_STATE, "LDD ran past its last round");
      const uint64_t start = shifts[r], end = std::min<uint64_t>(shifts[r + 1], n);
      if (end > start) {
        k_ldd_ctl<<<1, 1, 0, s>>>(ctl, (uint32_t)nf);
        k_ldd_centres<<<byo_grid(end - start, 256, 16384), 256, 0, s>>>(off, start, end, r, cl, rd, q, ctl);
        BYO_LAUNCH_CHECK();
        BYO_HIP(hipMemcpyAsync(h, ctl, sizeof(LddCtl), hipMemcpyDeviceToHost, s));
        BYO_HIP(hipStreamSynchronize(s));
        nf = h->cnt;
        fdeg += h->deg;
      }
      visited += nf;
      if (visited >= n) break;
      k_ldd_ctl<<<1, 1, 0, s>>>(ctl, 0);
      if (nf + fdeg > m / 20)
        k_ldd_pull<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, ids, n, r, cl, rd, qn, ctl);
      else if (nf)
        k_ldd_push<<<byo_grid(nf * 64, 256, 16384), 256, 0, s>>>(off, ids, q, (uint32_t)nf, r, cl, rd, qn, ctl);
      BYO_LAUNCH_CHECK();
      BYO_HIP(hipMemcpyAsync(h, ctl, sizeof(LddCtl), hipMemcpyDeviceToHost, s));
      BYO_HIP(hipStreamSynchronize(s));
      nf = h->cnt;
      fdeg = h->deg;
      std::swap(q, qn);
    }
  }
  if (rounds_run) *rounds_run = r;
  return t.stop();
}

}  // namespace byo

using namespace byo;

extern "C" {

int byo_ldd(byo_graph_t g, double beta, uint32_t *cluster, uint32_t *rounds_run, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && cluster, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(beta > 0 && beta < 64, BYO_E_INVAL, "beta must be in (0, 64)");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    const double ms = ldd_run(g, beta, rounds_run);
    if (g->prof.on) g->account("ldd", ms);
    if (device_ms) *device_ms = ms;
    if (n) BYO_HIP(hipMemcpyAsync(cluster, g->buf("ldd_cl", n * 4 + 64).as<uint32_t>(), n * 4, hipMemcpyDeviceToHost, g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

}  // extern "C"
