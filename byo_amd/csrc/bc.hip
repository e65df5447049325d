// bc.hip -- single-source betweenness centrality, the edgeMap consumer
// bc_bfs::SSBetweennessCentrality_BFS (benchmarks/SSBetweenessCentrality/
// Brandes/SSBetweennessCentrality.h:332-454), the variant run_all runs
// (run_unweighted.h:599-610 with ssbetween_fa = ssbetween_ligra = false).
//
// The reference's level-synchronous Brandes, in FP64:
//   levels   the BFS from src (the direction-optimizing persistent
//            traversal, then levels from its parent tree);
//   forward  for every level l >= 1, NumPaths[u] = sum of NumPaths[d] over
//            u's neighbours d in earlier levels (Visited[d] == 2);
//   backward NumPaths inverted; from the deepest level up, Dep[u] = sum of
//            Dep[d] over neighbours in deeper levels, plus 1/NumPaths[u];
//   result   (Dep - 1/NumPaths) / (1/NumPaths): sigma * Dep - 1 for reached
//            vertices, NaN for unreached ones (0 - inf) / inf, as the
//            reference's arithmetic gives.
//
// Vertices are bucketed by level (counting sort through per-block LDS
// histograms), and each level is one
// pass over its bucket's adjacency lists.  A vertex's level and value share
// a 16-byte record, so every edge costs one gather; a wave takes 64 bucket
// entries and picks the lanes per vertex from their largest degree, with 8
// gathers in flight per lane.  A neighbour in the level being written is
// masked out by its level, so the write race on its value is harmless.
#include <vector>

#include "primitives.hpp"

namespace byo {

void bfs_coop_run(byo_graph *g, const uint32_t *srcs, uint32_t k, bool keep_all, byo_bfs_stats *stats,
                  uint64_t maxdeg, double *ms_out);  // bfs_coop.hip

namespace {

struct alignas(16) BcState {
  double val;    // NumPaths in the forward pass, Dep in the backward pass
  uint32_t lev;  // BFS level, kNone when unreached
  uint32_t pad;
};

constexpr int kBcUnroll = 16;

// Levels are few (tens on small-world graphs) and every reached vertex
// counts into one, so counting and bucketing go through a per-block LDS
// histogram: one global atomic per level per block instead of one per
// vertex on a handful of addresses.  Graphs deeper than kBcLdsLevels levels
// (long paths, grids) use global atomics, spread over many levels anyway.
constexpr uint32_t kBcLdsLevels = 4096;
constexpr int kBcBlock = 256;

__global__ void __launch_bounds__(kBcBlock) k_bc_init(const uint32_t *lev, uint64_t n, uint32_t src, uint32_t R,
                                                      BcState *st, uint32_t *cnt) {
  __shared__ uint32_t h[kBcLdsLevels];
  const bool local = R <= kBcLdsLevels;
  if (local)
    for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) h[i] = 0;
  __syncthreads();
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    const uint32_t l = lev[v];
    st[v] = BcState{v == src ? 1.0 : 0.0, l, 0};
    if (l != kNone) {
      if (local) atomicAdd(&h[l], 1u);
      else atomicAdd(&cnt[l], 1u);
    }
  }
  __syncthreads();
  if (local)
    for (uint32_t i = threadIdx.x; i < R; i += blockDim.x)
      if (h[i]) atomicAdd(&cnt[i], h[i]);
}

// Bucket fill: pos[l] starts at the bucket's offset (exclusive scan).  Each
// block takes kBcBlock vertices at a time, ranks them per level in LDS,
// reserves one range per level, then writes.
__global__ void __launch_bounds__(kBcBlock) k_bc_bucket(const uint32_t *lev, uint64_t n, uint32_t R, uint32_t *pos,
                                                        uint32_t *order) {
  __shared__ uint32_t h[kBcLdsLevels];
  if (R > kBcLdsLevels) {
    for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
      const uint32_t l = lev[v];
      if (l != kNone) order[atomicAdd(&pos[l], 1u)] = (uint32_t)v;
    }
    return;
  }
  for (uint64_t c0 = (uint64_t)blockIdx.x * kBcBlock; c0 < n; c0 += (uint64_t)gridDim.x * kBcBlock) {
    for (uint32_t i = threadIdx.x; i < R; i += blockDim.x) h[i] = 0;
    __syncthreads();
    const uint64_t v = c0 + threadIdx.x;
    const uint32_t l = v < n ? lev[v] : kNone;
    const uint32_t rank = l != kNone ? atomicAdd(&h[l], 1u) : 0;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < R; i += blockDim.x)
      if (h[i]) h[i] = atomicAdd(&pos[i], h[i]);  // count -> this block's base
    __syncthreads();
    if (l != kNone) order[h[l] + rank] = (uint32_t)v;
    __syncthreads();
  }
}

// Sum of the masked neighbour values of the vertices of one tile, G lanes
// per vertex (G a power of two); returns the sum of the lane's own vertex.
template <int G, bool kFwd>
__device__ __forceinline__ double bc_tile(uint64_t lo, uint64_t hi, const uint32_t *ids, const BcState *st,
                                          uint32_t level, int lane) {
  constexpr int VPR = 64 / G;
  const int sub = lane & (G - 1), grp = lane / G;
  double mine = 0.0;
  for (int r = 0; r < G; r++) {
    const int owner = r * VPR + grp;
    const uint64_t b = G == 1 ? lo : __shfl(lo, owner, 64);
    const uint64_t e = G == 1 ? hi : __shfl(hi, owner, 64);
    double a = 0.0;
    for (uint64_t k = b + sub; k < e; k += (uint64_t)kBcUnroll * G) {
      uint32_t u[kBcUnroll];
      BcState x[kBcUnroll];
#pragma unroll
      for (int j = 0; j < kBcUnroll; j++) {
        const uint64_t kk = k + (uint64_t)j * G;
        u[j] = ids[kk < e ? kk : b];
      }
#pragma unroll
      for (int j = 0; j < kBcUnroll; j++) x[j] = st[u[j]];
#pragma unroll
      for (int j = 0; j < kBcUnroll; j++) {
        const bool in = k + (uint64_t)j * G < e;
        const bool take = kFwd ? x[j].lev < level : (x[j].lev != kNone && x[j].lev > level);
        a += (in && take) ? x[j].val : 0.0;
      }
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

// One level: forward (kFwd) NumPaths, or backward Dep plus 1/NumPaths.
template <bool kFwd>
__global__ void __launch_bounds__(256) k_bc_level(const uint64_t *off, const uint32_t *ids, const uint32_t *order,
                                                  uint64_t count, uint32_t level, BcState *st, const double *inv) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  for (uint64_t t0 = wave * 64; t0 < count; t0 += nwaves * 64) {
    const uint64_t i = t0 + lane;
    const bool valid = i < count;
    const uint32_t u = valid ? order[i] : 0;
    const uint64_t lo = valid ? off[u] : 0, hi = valid ? off[u + 1] : 0;
    const uint32_t dmax = wave_max((uint32_t)(hi - lo));
    double s;
    if (dmax <= kBcUnroll) s = bc_tile<1, kFwd>(lo, hi, ids, st, level, lane);
    else if (dmax <= 2 * kBcUnroll) s = bc_tile<2, kFwd>(lo, hi, ids, st, level, lane);
    else if (dmax <= 4 * kBcUnroll) s = bc_tile<4, kFwd>(lo, hi, ids, st, level, lane);
    else if (dmax <= 8 * kBcUnroll) s = bc_tile<8, kFwd>(lo, hi, ids, st, level, lane);
    else if (dmax <= 16 * kBcUnroll) s = bc_tile<16, kFwd>(lo, hi, ids, st, level, lane);
    else if (dmax <= 32 * kBcUnroll) s = bc_tile<32, kFwd>(lo, hi, ids, st, level, lane);
    else s = bc_tile<64, kFwd>(lo, hi, ids, st, level, lane);
    if (valid) st[u].val = kFwd ? s : s + inv[u];
  }
}

// NumPaths inverted (1/0 = inf when unreached), Dep cleared.
__global__ void k_bc_invert(BcState *st, uint64_t n, double *inv) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    inv[v] = 1.0 / st[v].val;
    st[v].val = 0.0;
  }
}

__global__ void k_bc_final(const BcState *st, const double *inv, uint64_t n, double *out) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) out[v] = (st[v].val - inv[v]) / inv[v];
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_bc(byo_graph_t g, uint32_t src, double *out, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && out, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t n = g->n;
    BYO_REQUIRE(src < n, BYO_E_INVAL, "source out of range");
    hipStream_t s = g->stream;
    const uint64_t mx = (g->dir_policy == BYO_DIR_MLP) ? graph_max_degree(g) : 0;
    PhaseTimer t(s, true);
    byo_bfs_stats bst;
    double bms = 0;
    bfs_coop_run(g, &src, 1, false, &bst, mx, &bms);
    g->last_src = src;
    const uint32_t *lev = bfs_levels_dev(g);
    const uint32_t R = bst.levels;  // levels 0 .. R-1
    BcState *st = g->buf("bc_state", n * sizeof(BcState) + 64).as<BcState>();
    double *inv = g->buf("bc_inv", n * 8 + 64).as<double>();
    double *dout = g->buf("bc_out", n * 8 + 64).as<double>();
    uint32_t *order = g->buf("bc_order", n * 4 + 64).as<uint32_t>();
    uint32_t *cnt = g->buf("bc_cnt", ((uint64_t)R + 1) * 4 + 64).as<uint32_t>();
    uint32_t *pos = g->buf("bc_pos", ((uint64_t)R + 1) * 4 + 64).as<uint32_t>();
    DevBuf tmp(prim::scan_tmp_bytes(R + 1) + 64);
    const unsigned grid = byo_grid(n, 256, 16384);
    prim::fill_u32(cnt, 0, R + 1, s);
    const unsigned hgrid = (unsigned)byo_grid(n, kBcBlock, 2048);
    k_bc_init<<<hgrid, kBcBlock, 0, s>>>(lev, n, src, R, st, cnt);
    prim::exclusive_scan_u32(cnt, pos, R + 1, false, tmp.as<uint32_t>(), s);
    std::vector<uint32_t> start(R + 1);
    BYO_HIP(hipMemcpyAsync(start.data(), pos, (R + 1) * 4, hipMemcpyDeviceToHost, s));
    k_bc_bucket<<<hgrid, kBcBlock, 0, s>>>(lev, n, R, pos, order);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipStreamSynchronize(s));
    // bucket l holds order[start[l] .. start[l+1]); start[R] = vertices reached
    auto bucket = [&](uint32_t l, uint64_t *cnt_l) {
      *cnt_l = (uint64_t)start[l + 1] - start[l];
      return order + start[l];
    };
    for (uint32_t l = 1; l < R; l++) {
      uint64_t c;
      const uint32_t *o = bucket(l, &c);
      k_bc_level<true><<<byo_grid(c * 64, 256, 16384), 256, 0, s>>>(g->off.as<uint64_t>(), g->ids.as<uint32_t>(), o,
                                                                    c, l, st, inv);
    }
    k_bc_invert<<<grid, 256, 0, s>>>(st, n, inv);
    for (uint32_t l = R; l-- > 0;) {
      uint64_t c;
      const uint32_t *o = bucket(l, &c);
      k_bc_level<false><<<byo_grid(c * 64, 256, 16384), 256, 0, s>>>(g->off.as<uint64_t>(), g->ids.as<uint32_t>(),
                                                                     o, c, l, st, inv);
    }
    k_bc_final<<<grid, 256, 0, s>>>(st, inv, n, dout);
    BYO_LAUNCH_CHECK();
    const double ms = t.stop();
    if (g->prof.on) g->account("bc", ms);
    if (device_ms) *device_ms = ms;
    BYO_HIP(hipMemcpyAsync(out, dout, n * 8, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
