// bfs_batch.hip -- launches the shared multi-source traversal of
// bfs_batch.hpp on its own (byo_bfs_multi) and reads its results back.
#include "bfs_batch.hpp"

#ifndef BYO_BATCH_WPE
#define BYO_BATCH_WPE 3
#endif

namespace byo {
namespace batch {

template <class M, bool G>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(BYO_BATCH_WPE, 8))) k_bfs_batch(BatchArgs a) {
  bfs_batch_body<M, G>(a);
}


// Per-launch results (level logs, last levels, level count, tallies) sit in
// one device block so they come back with one copy.
constexpr size_t kLogBytes = 65 * sizeof(BLog), kLastOff = kLogBytes, kNlevOff = kLastOff + 64 * 4,
                 kTotOff = kNlevOff + 8, kOutBytes = kTotOff + 64 * 16;

namespace {
// fn[v] = the first entry of v's list (a PMA list starts with a key), kNone
// for an empty list.
__global__ void k_first_ids(const uint64_t *off, const uint32_t *ids, uint64_t n, uint32_t *fn) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) fn[v] = off[v] < off[v + 1] ? ids[off[v]] : kNone;
}
}  // namespace

// The first-neighbour array of view v, built once per adjacency.
const uint32_t *first_ids(byo_graph *g, const AdjView &v) {
  uint32_t *fn = g->buf("bfs_first_ids", g->n * 4 + 64).as<uint32_t>();
  const int view = v.deg ? 1 : 0;
  if (!g->fn_valid || g->fn_view != view) {
    k_first_ids<<<byo_grid(g->n, 256, 16384), 256, 0, g->stream>>>(v.off, v.ids, g->n, fn);
    BYO_LAUNCH_CHECK();
    g->fn_valid = true;
    g->fn_view = view;
  }
  return fn;
}

void prepare(byo_graph *g, const AdjView &v, const uint32_t *srcs, uint32_t k, uint64_t maxdeg, BatchArgs &a,
             int mask_bytes, int nblocks) {
  BYO_REQUIRE(k >= 1 && k <= 64, BYO_E_INVAL, "batched BFS takes 1..64 sources per launch");
  const uint64_t n = g->n;
  hipStream_t s = g->stream;
  a = BatchArgs{};
  a.off = v.off;
  a.ids = v.ids;
  a.deg = v.deg;
  a.fn = first_ids(g, v);
  a.n = n;
  a.m = g->m;
  a.nm = (n + 63) / 64 * 64;
  a.pstride = n;
  a.parents = g->buf("bfs_parents_all", (size_t)k * n * 4 + 64).as<uint32_t>();
  a.qstride = (n + 63) / 64 * 64;
  a.q = g->buf("bfs_batch_q", 3 * a.qstride * 4 + 64).as<uint32_t>();
  a.scan = g->buf("bfs_scan", (n + 1) * 8 + 64).as<uint64_t>();
  a.cstart = g->buf("bfs_cstart", (v.nslots / kChunk + 2) * 4 + 64).as<uint32_t>();
  a.ctr = g->buf("bfs_batch_ctr", 3 * sizeof(BCtr) + 64).as<BCtr>();
  a.k = k;
  char *out = g->buf("bfs_batch_out", kOutBytes + 64).as<char>();
  a.logs = reinterpret_cast<BLog *>(out);
  a.lastlev = reinterpret_cast<uint32_t *>(out + kLastOff);
  a.nlev = reinterpret_cast<uint32_t *>(out + kNlevOff);
  a.tot = reinterpret_cast<unsigned long long *>(out + kTotOff);
  a.dc = dir_ctl(g, maxdeg, s, true);
  a.pieces = make_pieces(g, nblocks, v.deg ? kPmaPieces : kTraversalPieces);
  a.bar = coop_bar(g);
  a.dyn_ctr = g->buf("bfs_batch_dyn", 3 * 8192 * 4 + 64).as<unsigned int>();
  a.dyn = (uint32_t)std::min<int64_t>(g->tuning("pull_dyn"), a.pieces.per_wave);
  a.bsum = g->buf("bfs_batch_bsum", (size_t)nblocks * 8 + 64).as<unsigned long long>();
  a.seen = g->buf("bfs_batch_seen", a.nm * mask_bytes + 64).p;
  a.fm = g->buf("bfs_batch_fm", 3 * a.nm * mask_bytes + 64).p;
  a.ubm = g->buf("bfs_batch_ubm", 3 * (a.nm / 64) * 8 + 64).as<uint64_t>();
  // the sources travel in the kernel arguments (no copy queued before the
  // launch); the pinned staging of the results (at 256) is sized now, before
  // any copy is queued
  std::memcpy(a.srcs, srcs, (size_t)k * 4);
  (void)pinned(g, 256 + kOutBytes);
  (void)s;
}

void finish(byo_graph *g, const BatchArgs &a, byo_bfs_stats *stats, double ms) {
  if (!stats) return;
  hipStream_t s = g->stream;
  const uint32_t k = a.k;
  int wall_khz = 0;
  BYO_HIP(hipDeviceGetAttribute(&wall_khz, hipDeviceAttributeWallClockRate, g->device));
  const double wall_hz = wall_khz > 0 ? wall_khz * 1e3 : 1e8;
  char *ho = static_cast<char *>(pinned(g, 256 + kOutBytes)) + 256;
  BYO_HIP(hipMemcpyAsync(ho, a.logs, kOutBytes, hipMemcpyDeviceToHost, s));
  BYO_HIP(hipStreamSynchronize(s));
  const BLog *logs = reinterpret_cast<const BLog *>(ho);
  const uint32_t *last = reinterpret_cast<const uint32_t *>(ho + kLastOff);
  const uint32_t nlev = *reinterpret_cast<const uint32_t *>(ho + kNlevOff);
  const unsigned long long *tot = reinterpret_cast<const unsigned long long *>(ho + kTotOff);
  if (g->prof.on) {
    const BLog &mk = logs[64];
    g->account("bfs_batch_seed", (double)(mk.fdeg - mk.fsize) / wall_hz * 1e3);
    g->account("bfs_batch_levels", (double)(mk.active - mk.fdeg) / wall_hz * 1e3);
    g->account("bfs_batch_tally", (double)(mk.t0 - mk.active) / wall_hz * 1e3);
  }
  for (uint32_t j = 0; j < k; j++) {
    byo_bfs_stats &st = stats[j];
    std::memset(&st, 0, sizeof(st));
    st.levels = last[j] + 1;
    for (uint32_t L = 0; L < std::min<uint32_t>(nlev, 64); L++) {
      const BLog &lg = logs[L];
      st.fsize[L] = lg.fsize;
      st.fdeg[L] = lg.fdeg;
      st.dir[L] = (uint8_t)lg.dir;
      if (L + 1 < 64 && L + 1 <= nlev) st.level_ms[L] = (float)((double)(logs[L + 1].t0 - lg.t0) / wall_hz * 1e3);
    }
    st.reached = (uint32_t)tot[2 * j];
    st.edges_reached = tot[2 * j + 1];
    st.device_ms = ms / k;
  }
}

namespace {
template <class M>
double launch(byo_graph *g, const uint32_t *srcs, uint32_t k, uint64_t maxdeg, BatchArgs &a) {
  const AdjView v = adj_view(g);
  const bool gaps = v.deg != nullptr;
  const void *kern = gaps ? reinterpret_cast<const void *>(k_bfs_batch<M, true>)
                          : reinterpret_cast<const void *>(k_bfs_batch<M, false>);
  const int nblocks = traversal_grid(g, kern, kBlock, v.nslots);
  prepare(g, v, srcs, k, maxdeg, a, (int)sizeof(M), nblocks);
  void *args[] = {&a};
  PhaseTimer t(g->stream, true);
  BYO_HIP(hipLaunchCooperativeKernel(kern, dim3(nblocks), dim3(kBlock), args, 0, g->stream));
  pieces_launched(g, a.pieces, nblocks, gaps ? kPmaPieces : kTraversalPieces);
  return t.stop();
}
}  // namespace

}  // namespace batch

// Runs k <= 64 BFS as one shared traversal in one cooperative launch.
// Parents of source j end in scratch "bfs_parents_all" at j * n.  Fills
// stats[k] when given: levels, reached and edges_reached per source; the
// per-level records (frontier size, degree sum, direction, time) are those
// of the union frontier and are the same for every source.
void bfs_batch_run(byo_graph *g, const uint32_t *srcs, uint32_t k, byo_bfs_stats *stats, uint64_t maxdeg,
                   double *ms_out) {
  batch::BatchArgs a{};
  double ms;
  if (k <= 8)
    ms = batch::launch<uint8_t>(g, srcs, k, maxdeg, a);
  else if (k <= 32)
    ms = batch::launch<uint32_t>(g, srcs, k, maxdeg, a);
  else
    ms = batch::launch<uint64_t>(g, srcs, k, maxdeg, a);
  if (ms_out) *ms_out = ms;
  batch::finish(g, a, stats, ms);
}

}  // namespace byo
