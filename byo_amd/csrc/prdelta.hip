// prdelta.hip -- delta::PageRankDelta (benchmarks/PageRank/PageRank.h:
// 335-396), run_all's PageRank with -delta: FP64 ranks that grow by
// propagated deltas, only the vertices whose delta is large against their
// rank staying in the frontier.
//
// Per round (the reference's sparse_or_dense, :230-269, then its vertex
// filters, :271-327):
//   nghSum   v's sum of delta/degree over its neighbours in the frontier --
//            a pull over every vertex when the frontier holds more than n/5
//            vertices (the reference's dense edgeMapReduce), otherwise a
//            push of the frontier's lists with FP64 atomics (its sparse
//            edgeMap);
//   filter   round 1: p += d * nghSum + (1-d)/n, delta = that - 1/n, active
//            when delta > local_eps * p (signed, as the reference tests it);
//            later rounds: delta = d * nghSum for every vertex, added to p
//            and active where |delta| > local_eps * p;
//   stop     sum |delta| < eps (FP64, reduced in a fixed order), or
//            max_iters rounds.
// The pull gathers one double per edge from dodF = delta/degree of frontier
// vertices and 0 elsewhere, so frontier membership costs no second gather;
// lanes per vertex are picked per tile as in the PageRank pull.  With the
// pr_relabel knob the rounds run on the degree-relabelled graph (the same
// values per vertex up to summation order) and ranks are mapped back.
#include <algorithm>

#include "common.hpp"

namespace byo {
namespace {

struct PdCtl {
  double l1;
  unsigned int cnt;
  unsigned int pad[13];
};

__global__ void k_pd_init(const uint64_t *off, uint64_t n, double *p, double *dodf, double *ngh) {
  const double inv_n = 1.0 / (double)n;
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    p[v] = 0.0;
    dodf[v] = inv_n / (double)(off[v + 1] - off[v]);  // every vertex starts in the frontier
    ngh[v] = 0.0;
  }
}

constexpr int kPdUnroll = 8;

template <int G>
__device__ __forceinline__ double pd_tile(uint64_t lo, uint64_t hi, const uint32_t *ids, const double *dodf,
                                          int lane) {
  constexpr int VPR = 64 / G;
  const int sub = lane & (G - 1), grp = lane / G;
  double mine = 0.0;
  for (int r = 0; r < G; r++) {
    const int owner = r * VPR + grp;
    const uint64_t b = G == 1 ? lo : __shfl(lo, owner, 64);
    const uint64_t e = G == 1 ? hi : __shfl(hi, owner, 64);
    double a = 0.0;
    for (uint64_t k = b + sub; k < e; k += (uint64_t)kPdUnroll * G) {
      uint32_t u[kPdUnroll];
      double x[kPdUnroll];
#pragma unroll
      for (int j = 0; j < kPdUnroll; j++) {
        const uint64_t kk = k + (uint64_t)j * G;
        u[j] = ids[kk < e ? kk : b];
      }
#pragma unroll
      for (int j = 0; j < kPdUnroll; j++) x[j] = dodf[u[j]];
#pragma unroll
      for (int j = 0; j < kPdUnroll; j++) a += (k + (uint64_t)j * G < e) ? x[j] : 0.0;
    }
    if (G == 1) return a;
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    const int rel = lane - r * VPR;
    const double got = __shfl(a, (rel >= 0 && rel < VPR) ? rel * G : 0, 64);
    if (rel >= 0 && rel < VPR) mine = got;
  }
  return mine;
}

__global__ void __launch_bounds__(256) k_pd_pull(const uint64_t *off, const uint32_t *ids, uint64_t n,
                                                 const double *dodf, double *ngh) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t t0 = wave * 64; t0 < n; t0 += nwaves * 64) {
    const uint64_t v = t0 + lane;
    const bool valid = v < n;
    const uint64_t lo = off[valid ? v : n], hi = off[valid ? v + 1 : n];
    const uint32_t dmax = wave_max((uint32_t)(hi - lo));
    double s;
    if (dmax <= kPdUnroll) s = pd_tile<1>(lo, hi, ids, dodf, lane);
    else if (dmax <= 2 * kPdUnroll) s = pd_tile<2>(lo, hi, ids, dodf, lane);
    else if (dmax <= 4 * kPdUnroll) s = pd_tile<4>(lo, hi, ids, dodf, lane);
    else if (dmax <= 8 * kPdUnroll) s = pd_tile<8>(lo, hi, ids, dodf, lane);
    else if (dmax <= 16 * kPdUnroll) s = pd_tile<16>(lo, hi, ids, dodf, lane);
    else if (dmax <= 32 * kPdUnroll) s = pd_tile<32>(lo, hi, ids, dodf, lane);
    else s = pd_tile<64>(lo, hi, ids, dodf, lane);
    if (valid) ngh[v] = s;
  }
}

// Sparse round: a wave per frontier vertex adds its delta/degree to its
// neighbours' sums (ngh was cleared by the last filter).
__global__ void k_pd_push(const uint64_t *off, const uint32_t *ids, const uint32_t *q, uint32_t cnt,
                          const double *dodf, double *ngh) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t i = wave; i < cnt; i += nwaves) {
    const uint32_t s = q[i];
    const double x = dodf[s];
    for (uint64_t e = off[s] + lane; e < off[s + 1]; e += 64) atomicAdd(&ngh[ids[e]], x);
  }
}

__global__ void __launch_bounds__(256) k_pd_filter(const uint64_t *off, uint64_t n, bool first, double local_eps,
                                                   double *p, double *dodf, double *ngh, uint32_t *q, PdCtl *ctl,
                                                   double *partial) {
  __shared__ double lds[4];
  const double damping = 0.85, inv_n = 1.0 / (double)n;
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  double l1 = 0.0;
  for (uint64_t base = wave * 64; base < n; base += nwaves * 64) {
    const uint64_t v = base + lane;
    bool act = false;
    if (v < n) {
      const double deg = (double)(off[v + 1] - off[v]);
      const double s = ngh[v];
      double pv = p[v], nd;
      if (first) {
        const double pre = damping * s + (1.0 - damping) * inv_n;
        pv += pre;
        nd = pre - inv_n;
        act = nd > local_eps * pv;
      } else {
        nd = s * damping;
        act = fabs(nd) > local_eps * pv;
        if (act) pv += nd;
      }
      p[v] = pv;
      dodf[v] = act ? nd / deg : 0.0;
      ngh[v] = 0.0;
      l1 += fabs(nd);
    }
    const uint64_t bal = __ballot(act);
    if (bal) {
      unsigned int at = 0;
      if (lane == 0) at = atomicAdd(&ctl->cnt, (unsigned int)__popcll(bal));
      at = __shfl(at, 0, 64);
      if (act) q[at + mbcnt(bal)] = (uint32_t)v;
    }
  }
  l1 = wave_sum(l1);
  if (lane == 0) lds[threadIdx.x >> 6] = l1;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

// Fixed-order sum of the filter's block partials.
__global__ void k_pd_l1(const double *partial, int nb, PdCtl *ctl) {
  __shared__ double lds[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) ctl->l1 = lds[0];
}

__global__ void k_pd_clear(PdCtl *ctl) { ctl->cnt = 0; }

// Rank of vertex v from the relabelled vector (perm: old id -> new id).
__global__ void k_pd_unpermute(const double *in, const uint32_t *perm, uint64_t n, double *out) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) out[v] = in[perm[v]];
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_pagerank_delta(byo_graph_t g, double eps, double local_eps, int max_iters, double *p, int *rounds_run,
                       double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    BYO_REQUIRE(max_iters >= 0, BYO_E_INVAL, "max_iters < 0");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    hipStream_t s = g->stream;
    double *pd = g->buf("pd_p", n * 8 + 64).as<double>();
    double *dodf = g->buf("pd_dodf", n * 8 + 64).as<double>();
    double *ngh = g->buf("pd_ngh", n * 8 + 64).as<double>();
    uint32_t *q = g->buf("pd_q", n * 4 + 64).as<uint32_t>();
    PdCtl *ctl = g->buf("pd_ctl", sizeof(PdCtl) + 64).as<PdCtl>();
    const unsigned fgrid = (unsigned)byo_grid(n, 256, 4096);
    double *partial = g->buf("pd_partial", (fgrid + 8) * 8).as<double>();
    PdCtl *h = static_cast<PdCtl *>(pinned(g, sizeof(PdCtl)));
    // the graph the rounds run on: as it is, or its degree-relabelled copy
    // (tuning knob pr_relabel, shared with PageRank), whose hot deltas sit
    // together at low ids where they stay in L2
    const RelabelView view = g->tuning("pr_relabel") && n
                                 ? relabel_view(g)
                                 : RelabelView{g->off.as<uint64_t>(), g->ids.as<uint32_t>(), nullptr};
    const uint64_t *off = view.off;
    const uint32_t *ids = view.ids;
    PhaseTimer t(s, true);
    int round = 0, ran = 0;
    uint64_t fs = n;  // every vertex starts in the frontier
    if (n) {
      k_pd_init<<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, n, pd, dodf, ngh);
      const unsigned pgrid = (unsigned)g->tuning("pr_grid");
      while (round++ < max_iters) {
        ran = round;
        if (fs > n / 5)
          k_pd_pull<<<pgrid, 256, 0, s>>>(off, ids, n, dodf, ngh);
        else if (fs)
          k_pd_push<<<byo_grid(fs * 64, 256, 16384), 256, 0, s>>>(off, ids, q, (uint32_t)fs, dodf, ngh);
        k_pd_clear<<<1, 1, 0, s>>>(ctl);
        k_pd_filter<<<fgrid, 256, 0, s>>>(off, n, round == 1, local_eps, pd, dodf, ngh, q, ctl, partial);
        k_pd_l1<<<1, 1024, 0, s>>>(partial, (int)fgrid, ctl);
        BYO_LAUNCH_CHECK();
        BYO_HIP(hipMemcpyAsync(h, ctl, sizeof(PdCtl), hipMemcpyDeviceToHost, s));
        BYO_HIP(hipStreamSynchronize(s));
        if (h->l1 < eps) break;
        fs = h->cnt;
      }
    }
    const double ms = t.stop();
    if (g->prof.on) g->account("pagerank_delta", ms);
    if (device_ms) *device_ms = ms;
    if (rounds_run) *rounds_run = ran;
    if (p && n) {
      const double *res = pd;
      if (view.perm) {
        double *un = g->buf("pd_unperm", n * 8 + 64).as<double>();
        k_pd_unpermute<<<byo_grid(n, 256, 16384), 256, 0, s>>>(pd, view.perm, n, un);
        BYO_LAUNCH_CHECK();
        res = un;
      }
      BYO_HIP(hipMemcpyAsync(p, res, n * 8, hipMemcpyDeviceToHost, s));
    }
    BYO_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
