// job.hip -- personalized PageRank as one persistent cooperative launch, and
// the batched job: up to 64 BFS sources (one shared traversal,
// bfs_batch.hpp) followed by up to 64 personalized-PageRank seeds, all
// iterations, in ONE cooperative launch.
//
// Personalized PageRank restates the Jacobi PageRank of PageRank_edgeMap
// (benchmarks/PageRank/PageRank.h:80-123) with the teleport mass on the seed:
// p0 = e_seed, p'[v] = damping * sum_{u in N(v)} p[u]/deg(u) +
// (1 - damping) [v == seed], FP32 values and per-vertex Kahan sums.  The K
// seeds are kept vertex-major (p[v*K + j]), so every edge gathers the K
// contributions of its source as one contiguous run; W = K rounded up to a
// power of two lanes serve one vertex, 64/W vertices per wave step, and
// each lane sums its seed's contributions in adjacency order (eight
// gathers in flight per step with the next eight ids behind them, the adds
// in order), so the result is the same as one thread summing the list.  The
// vertices of a sweep are shared out by weight-balanced word pieces of their
// own (knob ppr_pieces; the sweeps balance best at fewer pieces per wave
// than the traversal); iterations are separated by the grid barrier.
#include "bfs_batch.hpp"
#include "kahan.hpp"

namespace byo {
namespace {

// Neighbours whose ids and contributions are loaded before any is added
// (the adds stay in adjacency order).
constexpr int kPprUnroll = 8;

struct PprArgs {
  const uint64_t *off;
  const uint32_t *ids;
  uint64_t n;
  const uint32_t *seeds;
  uint32_t K;  // seeds
  uint32_t W;  // lanes per vertex: K rounded up to a power of two
  int iters;
  float damping;
  float *p;       // n x K values of the last iteration
  float *c0, *c1;  // n x K contributions p/deg, alternating
  Pieces pieces;
};

// Every thread of the grid.  pieces must already hold the word pieces of
// this grid (pieces_init before a barrier).
__device__ __forceinline__ void ppr_body(const PprArgs &q, GridBar *bar) {
  const uint64_t nb = gridDim.x;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const uint64_t n = q.n, K = q.K;
  for (uint64_t i = global_tid(); i < n * K; i += grid_threads()) {
    const uint64_t v = i / K;
    const uint32_t j = (uint32_t)(i - v * K);
    const uint64_t deg = q.off[v + 1] - q.off[v];
    const float x = (q.seeds[j] == v) ? 1.0f : 0.0f;
    q.p[i] = x;
    q.c0[i] = deg ? x / (float)deg : 0.0f;
  }
  grid_barrier(bar, (unsigned)nb);
  const int lane = lane_id();
  const uint32_t W = q.W, G = 64 / W;
  const uint32_t sub = lane & (W - 1), grp = lane / W;
  const uint32_t my_seed = sub < K ? q.seeds[sub] : kNone;
  const float teleport = 1.0f - q.damping;
  for (int it = 0; it < q.iters; it++) {
    const float *cin = (it & 1) ? q.c1 : q.c0;
    float *cout = (it & 1) ? q.c0 : q.c1;
    for_my_words(q.pieces, wave, nwaves, [&](uint64_t w) {
      for (uint32_t x0 = 0; x0 < 64; x0 += G) {
        const uint64_t v = w * 64 + x0 + grp;
        if (v < n && sub < K) {
          const uint64_t b = q.off[v], e = q.off[v + 1];
          Acc<float> a{0.f, 0.f};
          if (e > b) {
            // Steps of kPprUnroll edges, the last one partial: ids are read
            // at clamped positions (so every load is issued), the next
            // step's ids are requested while this step's gathers are in
            // flight, and only the edges of the list are added.
            const uint64_t last = e - 1;
            uint32_t u[kPprUnroll];
#pragma unroll
            for (int t = 0; t < kPprUnroll; t++) u[t] = q.ids[std::min<uint64_t>(b + t, last)];
            for (uint64_t k = b; k < e; k += kPprUnroll) {
              float f[kPprUnroll];
#pragma unroll
              for (int t = 0; t < kPprUnroll; t++) f[t] = cin[(uint64_t)u[t] * K + sub];
#pragma unroll
              for (int t = 0; t < kPprUnroll; t++) u[t] = q.ids[std::min<uint64_t>(k + kPprUnroll + t, last)];
#pragma unroll
              for (int t = 0; t < kPprUnroll; t++)
                if (k + t < e) kahan_add(a, f[t]);
            }
          }
          const float pv = q.damping * (a.s - a.c) + (my_seed == v ? teleport : 0.0f);
          q.p[v * K + sub] = pv;
          cout[v * K + sub] = (e > b) ? pv / (float)(e - b) : 0.0f;
        }
      }
    });
    grid_barrier(bar, (unsigned)nb);
  }
}

__global__ void __launch_bounds__(kCoopBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) k_ppr_persist(PprArgs q,
                                                                                                       GridBar *bar) {
  const uint64_t nwaves = grid_threads() >> 6;
  pieces_init(q.pieces, q.off, q.n, (q.n + 63) >> 6, nwaves);
  grid_barrier(bar, gridDim.x);
  ppr_body(q, bar);
}

// The batched job: the shared BFS traversal, the PageRank pieces, a
// barrier, then the PageRank iterations.
template <class M>
__global__ void __launch_bounds__(kCoopBlock) __attribute__((amdgpu_waves_per_eu(4, 8)))
    k_job(batch::BatchArgs a, PprArgs q) {
  if (a.k) batch::bfs_batch_body<M>(a);
  if (q.K) pieces_init(q.pieces, q.off, q.n, (q.n + 63) >> 6, grid_threads() >> 6);
  grid_barrier(a.bar, gridDim.x);
  if (q.K) ppr_body(q, a.bar);
}

PprArgs ppr_args(byo_graph *g, const uint32_t *seeds, uint32_t k, double damping, int iters) {
  PprArgs q{};
  const uint64_t n = g->n;
  q.off = g->off.as<uint64_t>();
  q.ids = g->ids.as<uint32_t>();
  q.n = n;
  q.K = k;
  q.W = 1;
  while (q.W < k) q.W <<= 1;
  q.iters = iters;
  q.damping = (float)damping;
  const uint64_t sz = n * std::max<uint32_t>(k, 1) * sizeof(float) + 64;
  q.p = g->buf("ppr_p", sz).as<float>();
  q.c0 = g->buf("ppr_c0", sz).as<float>();
  q.c1 = g->buf("ppr_c1", sz).as<float>();
  uint32_t *dseeds = g->buf("ppr_seeds", 64 * 4 + 64).as<uint32_t>();
  if (k) BYO_HIP(hipMemcpyAsync(dseeds, seeds, k * 4, hipMemcpyHostToDevice, g->stream));
  q.seeds = dseeds;
  return q;
}

// p[j*n + v] from the vertex-major device state.
void ppr_readback(byo_graph *g, const PprArgs &q, float *p) {
  const uint64_t n = q.n, k = q.K;
  std::vector<float> h(n * k);
  BYO_HIP(hipMemcpyAsync(h.data(), q.p, n * k * sizeof(float), hipMemcpyDeviceToHost, g->stream));
  BYO_HIP(hipStreamSynchronize(g->stream));
  for (uint64_t v = 0; v < n; v++)
    for (uint64_t j = 0; j < k; j++) p[j * n + v] = h[v * k + j];
}

template <class M>
double job_launch(byo_graph *g, const uint32_t *srcs, uint32_t kb, uint64_t maxdeg, batch::BatchArgs &a,
                  PprArgs &q) {
  const void *kern = reinterpret_cast<const void *>(k_job<M>);
  // the PageRank sweeps dominate a job: it runs at the larger of the two
  // tuned residencies (the traversal body caps it at its own occupancy)
  const int cap = (int)std::max(g->tuning("coop_blocks_per_cu"), g->tuning("ppr_blocks_per_cu"));
  const int nblocks = coop_grid(kern, kCoopBlock, cap);
  // the job reads the CSR view (exported from a CPMA by the caller)
  const AdjView csr{g->off.as<uint64_t>(), g->ids.as<uint32_t>(), nullptr, g->m};
  batch::prepare(g, csr, srcs, std::max<uint32_t>(kb, 1), maxdeg, a, (int)sizeof(M), nblocks);
  a.k = kb;
  q.pieces = make_pieces(g, nblocks, kPprPieces);
  void *args[] = {&a, &q};
  PhaseTimer t(g->stream, true);
  BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kCoopBlock), args, 0, g->stream));
  if (kb) pieces_launched(g, a.pieces, nblocks);
  if (q.K) pieces_launched(g, q.pieces, nblocks, kPprPieces);
  return t.stop();
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_ppr_multi(byo_graph_t g, const uint32_t *seeds, uint32_t k, double damping, int iters, float *p,
                  double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && (seeds || k == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(k <= 64, BYO_E_INVAL, "at most 64 seeds per call");
    BYO_REQUIRE(iters >= 0, BYO_E_INVAL, "iters < 0");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    for (uint32_t j = 0; j < k; j++) BYO_REQUIRE(seeds[j] < g->n, BYO_E_INVAL, "seed out of range");
    if (k == 0 || g->n == 0) return;
    PprArgs q = ppr_args(g, seeds, k, damping, iters);
    const void *kern = reinterpret_cast<const void *>(k_ppr_persist);
    const int nblocks = coop_grid(kern, kCoopBlock, (int)g->tuning("ppr_blocks_per_cu"));
    q.pieces = make_pieces(g, nblocks, kPprPieces);
    GridBar *bar = coop_bar(g);
    void *args[] = {&q, &bar};
    PhaseTimer t(g->stream, true);
    BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kCoopBlock), args, 0, g->stream));
    pieces_launched(g, q.pieces, nblocks, kPprPieces);
    const double ms = t.stop();
    if (g->prof.on) g->account("ppr", ms);
    if (device_ms) *device_ms = ms;
    if (p) ppr_readback(g, q, p);
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

int byo_batch_job(byo_graph_t g, const uint32_t *bfs_srcs, uint32_t k_bfs, const uint32_t *ppr_seeds,
                  uint32_t k_ppr, double damping, int ppr_iters, uint32_t *parents, byo_bfs_stats *stats,
                  float *ppr, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && (bfs_srcs || k_bfs == 0) && (ppr_seeds || k_ppr == 0), BYO_E_INVAL, "null argument");
    BYO_REQUIRE(k_bfs <= 64 && k_ppr <= 64, BYO_E_INVAL, "at most 64 BFS sources and 64 seeds per job");
    BYO_REQUIRE(ppr_iters >= 0, BYO_E_INVAL, "ppr_iters < 0");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    for (uint32_t j = 0; j < k_bfs; j++) BYO_REQUIRE(bfs_srcs[j] < n, BYO_E_INVAL, "source out of range");
    for (uint32_t j = 0; j < k_ppr; j++) BYO_REQUIRE(ppr_seeds[j] < n, BYO_E_INVAL, "seed out of range");
    if (n == 0 || (k_bfs == 0 && k_ppr == 0)) return;
    uint64_t maxdeg = 0;
    if (g->dir_policy == BYO_DIR_MLP) maxdeg = graph_max_degree(g);
    PprArgs q = ppr_args(g, ppr_seeds, k_ppr, damping, ppr_iters);
    batch::BatchArgs a{};
    const uint32_t dummy = 0;
    const uint32_t *srcs = k_bfs ? bfs_srcs : &dummy;
    double ms;
    if (k_bfs <= 8)
      ms = job_launch<uint8_t>(g, srcs, k_bfs, maxdeg, a, q);
    else if (k_bfs <= 32)
      ms = job_launch<uint32_t>(g, srcs, k_bfs, maxdeg, a, q);
    else
      ms = job_launch<uint64_t>(g, srcs, k_bfs, maxdeg, a, q);
    if (g->prof.on) g->account("job", ms);
    if (device_ms) *device_ms = ms;
    if (k_bfs) {
      batch::finish(g, a, stats, ms);
      if (parents)
        BYO_HIP(hipMemcpyAsync(parents, a.parents, (uint64_t)k_bfs * n * 4, hipMemcpyDeviceToHost, g->stream));
      g->last_tree = a.parents + (uint64_t)(k_bfs - 1) * n;
      g->last_src = bfs_srcs[k_bfs - 1];
    }
    if (ppr && k_ppr) ppr_readback(g, q, ppr);
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

}  // extern "C"
