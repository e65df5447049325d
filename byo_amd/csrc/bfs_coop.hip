// bfs_coop.hip -- the whole BFS (all edgeMap levels, for k sources) in one
// persistent cooperative launch.
//
// Same traversal as the level-by-level path in bfs.hip, restating BFS.h:47-64
// over edgeMapData (gbbs/edge_map_data.h:254-307), but every level's
// direction is decided on the device from counters the previous level left
// in HBM, so no host round trip happens between levels.  Both frontier
// representations are kept current at every level -- the queue (sparse
// vertexSubset) is appended by one atomic per wave in push and in pull, and
// the bitmap (BitArray) is written by ballot words in pull and set by the
// CAS winners in push -- so no toSparse / toDense conversion pass exists.
//
// Per level:  sparse push  = 3 grid barriers (slice sums of frontier
//             degrees; exclusive scan; edge-balanced push),
//             dense pull   = 1 barrier,
//             dense forward = 2 barriers.
// Counters of the frontier sizes rotate over three slots so that resetting
// the slot for level L+2 never races with readers of level L or writers of
// level L+1.
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cmath>

#include "primitives.hpp"

namespace cg = cooperative_groups;

namespace byo {

namespace {

struct Ctr {
  unsigned long long size;  // frontier size == queue cursor
  unsigned long long deg;   // sum of out-degrees of the frontier
};

struct LevelLog {
  unsigned long long fsize, fdeg;
  uint32_t dir, pad;
};

struct CoopArgs {
  const uint64_t *off;
  const uint32_t *ids;
  uint64_t n, m, nw64;
  uint32_t *parents;        // parents of source j at parents + j * pstride
  uint64_t pstride;
  uint32_t *q0, *q1;
  uint64_t *bm0, *bm1, *visited64;
  uint64_t *scan;
  unsigned long long *bsum;
  Ctr *ctr;                 // 3 slots
  const uint32_t *srcs;
  uint32_t k;
  LevelLog *logs;           // k x 64
  uint32_t *nlevels;        // k
  unsigned long long *tot;  // k x {vertices reached, sum of their degrees}
  int policy;
  uint64_t thr;
  const float *sel;         // 1379 params + mean[6] + inv_std[6], or null
  uint64_t maxdeg;
};

constexpr int kBlock = 256;
constexpr int kPerLane = 4;
constexpr uint64_t kChunk = 64 * kPerLane;
constexpr int kNP = BYO_MLP_NPARAM;

__device__ __forceinline__ unsigned long long ld_relaxed(const unsigned long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool bit_get(const uint64_t *bm, uint32_t i) { return (bm[i >> 6] >> (i & 63)) & 1ull; }

struct MlpLds {
  float x[6], h1[32], h2[32], z[3];
};

// Direction of one edgeMap call: 0 sparse, 1 dense pull, 2 empty, 3 dense
// forward (the same decision as choose_direction in bfs.hip).  Called by
// every thread of the block (the selector MLP runs one neuron per thread,
// layout as in selector.hip); returns the same value in all threads.
__device__ int decide(const CoopArgs &a, bool is_dense, uint64_t fs, uint64_t fdeg, uint64_t visited_deg, int level,
                      MlpLds &s) {
  int rule;
  if (is_dense && fs > a.n / 10) {
    rule = 1;
  } else if (fdeg == 0) {
    return 2;
  } else {
    rule = (fs + fdeg > a.thr) ? 1 : 0;
  }
  if (a.policy == BYO_DIR_SPARSE) return 0;
  if (a.policy == BYO_DIR_DENSE) return 1;
  if (a.policy != BYO_DIR_MLP || !a.sel) return rule;
  const float *p = a.sel;
  const int t = threadIdx.x;
  __syncthreads();
  if (t == 0) {
    const double avg = a.n ? (double)a.m / (double)a.n : 0.0;
    float x[6];
    x[0] = (float)log2(1.0 + (double)fs);
    x[1] = (float)log2(1.0 + (double)fdeg);
    x[2] = (float)log2(1.0 + (double)(a.m > visited_deg ? a.m - visited_deg : 0));
    x[3] = (float)level;
    x[4] = (float)log2(1.0 + avg);
    x[5] = (float)log2(1.0 + (double)a.maxdeg);
    for (int i = 0; i < 6; i++) s.x[i] = (x[i] - p[kNP + i]) * p[kNP + 6 + i];
  }
  __syncthreads();
  if (t < 32) {
    float v = p[192 + t];
    for (int i = 0; i < 6; i++) v += p[t * 6 + i] * s.x[i];
    s.h1[t] = v > 0.f ? v : 0.f;
  }
  __syncthreads();
  if (t < 32) {
    float v = p[1248 + t];
    for (int i = 0; i < 32; i++) v += p[224 + t * 32 + i] * s.h1[i];
    s.h2[t] = v > 0.f ? v : 0.f;
  }
  __syncthreads();
  if (t < 3) {
    float v = p[1376 + t];
    for (int i = 0; i < 32; i++) v += p[1280 + t * 32 + i] * s.h2[i];
    s.z[t] = v;
  }
  __syncthreads();
  const float z0 = s.z[0], z1 = s.z[1], z2 = s.z[2];
  if (!isfinite(z0 + z1 + z2)) return rule;
  const int best = (z1 > z0) ? ((z2 > z1) ? 2 : 1) : ((z2 > z0) ? 2 : 0);
  return best == 0 ? 0 : (best == 1 ? 1 : 3);
}

// Block-wide sum (all threads get the result).
__device__ __forceinline__ unsigned long long block_sum(unsigned long long v, unsigned long long *lds) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = v;
  __syncthreads();
  unsigned long long s = 0;
  for (int w = 0; w < kBlock / 64; w++) s += lds[w];
  return s;
}

__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4, 8))) k_bfs_coop(CoopArgs a) {
  cg::grid_group grid = cg::this_grid();
  __shared__ unsigned long long lds[8];
  __shared__ MlpLds mlp;
  const int lane = threadIdx.x & 63;
  const uint64_t nb = gridDim.x;
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const bool leader = blockIdx.x == 0 && threadIdx.x == 0;
  const uint64_t nwords = (a.n + 63) >> 6;

  for (uint32_t j = 0; j < a.k; j++) {
    uint32_t *parents = a.parents + (uint64_t)j * a.pstride;
    const uint32_t src = a.srcs[j];
    for (uint64_t i = global_tid(); i < a.n; i += grid_threads()) parents[i] = kNone;
    for (uint64_t i = global_tid(); i < a.nw64; i += grid_threads()) {
      a.visited64[i] = 0;
      a.bm0[i] = 0;
    }
    if (leader)
      for (int s = 0; s < 3; s++) a.ctr[s] = Ctr{0, 0};
    grid.sync();
    if (leader) {
      parents[src] = src;
      a.visited64[src >> 6] |= 1ull << (src & 63);
      a.bm0[src >> 6] |= 1ull << (src & 63);
      a.q0[0] = src;
      a.ctr[0] = Ctr{1, a.off[src + 1] - a.off[src]};
    }
    grid.sync();

    uint32_t *q = a.q0, *qn = a.q1;
    uint64_t *fbm = a.bm0, *nbm = a.bm1;
    bool is_dense = false;
    uint64_t visited_deg = 0, reached = 0;
    int L = 0;
    for (;; L++) {
      const uint64_t fs = ld_relaxed(&a.ctr[L % 3].size);
      const uint64_t fdeg = ld_relaxed(&a.ctr[L % 3].deg);
      if (fs == 0) break;
      reached += fs;
      const int dir = decide(a, is_dense, fs, fdeg, visited_deg, L, mlp);
      if (leader && L < 64) a.logs[(uint64_t)j * 64 + L] = LevelLog{fs, fdeg, (uint32_t)dir, 0};
      if (dir == 2) {
        visited_deg += fdeg;
        L++;
        break;
      }
      Ctr *nx = &a.ctr[(L + 1) % 3];
      if (leader) a.ctr[(L + 2) % 3] = Ctr{0, 0};

      if (dir == 0) {
        // (1) slice sums of frontier degrees; clear the next bitmap
        const uint64_t S = (fs + nb - 1) / nb;
        const uint64_t s0 = std::min<uint64_t>(fs, blockIdx.x * S), s1 = std::min<uint64_t>(fs, s0 + S);
        unsigned long long part = 0;
        for (uint64_t i = s0 + threadIdx.x; i < s1; i += kBlock) {
          const uint32_t v = q[i];
          part += a.off[v + 1] - a.off[v];
        }
        part = block_sum(part, lds);
        if (threadIdx.x == 0) a.bsum[blockIdx.x] = part;
        for (uint64_t i = global_tid(); i < a.nw64; i += grid_threads()) nbm[i] = 0;
        grid.sync();
        // (2) exclusive scan of the degrees of the queue
        unsigned long long pre = 0;
        for (uint64_t b = threadIdx.x; b < blockIdx.x; b += kBlock) pre += a.bsum[b];
        pre = block_sum(pre, lds);
        for (uint64_t c0 = s0; c0 < s1; c0 += kBlock) {
          const uint64_t i = c0 + threadIdx.x;
          unsigned long long d = 0;
          if (i < s1) {
            const uint32_t v = q[i];
            d = a.off[v + 1] - a.off[v];
          }
          const unsigned long long incl = wave_incl_scan(d);
          __syncthreads();
          if (lane == 63) lds[threadIdx.x >> 6] = incl;
          __syncthreads();
          unsigned long long wpre = 0, tot = 0;
          for (int w = 0; w < kBlock / 64; w++) {
            if (w < (int)(threadIdx.x >> 6)) wpre += lds[w];
            tot += lds[w];
          }
          if (i < s1) a.scan[i] = pre + wpre + incl - d;
          pre += tot;
        }
        if (leader) a.scan[fs] = fdeg;
        grid.sync();
        // (3) edge-balanced push: chunks of 256 edges per wave
        const uint64_t E = fdeg;
        const uint64_t nch = (E + kChunk - 1) / kChunk;
        for (uint64_t c = wave; c < nch; c += nwaves) {
          const uint64_t e0 = c * kChunk;
          uint64_t lo = 0, hi = fs;  // scan[lo] <= e0 < scan[hi]
          while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (a.scan[mid] <= e0) lo = mid; else hi = mid;
          }
          const uint64_t ib = lo;
#pragma unroll
          for (int t = 0; t < kPerLane; t++) {
            const uint64_t e = e0 + (uint64_t)t * 64 + lane;
            bool win = false;
            uint32_t tgt = 0;
            if (e < E) {
              uint64_t x = ib, y = fs;
              while (y - x > 1) {
                const uint64_t mid = (x + y) >> 1;
                if (a.scan[mid] <= e) x = mid; else y = mid;
              }
              const uint32_t u = q[x];
              tgt = a.ids[a.off[u] + (e - a.scan[x])];
              if (!bit_get(a.visited64, tgt) && atomicCAS(&parents[tgt], kNone, u) == kNone) {
                win = true;
                atomicOr(reinterpret_cast<unsigned long long *>(&a.visited64[tgt >> 6]), 1ull << (tgt & 63));
                atomicOr(reinterpret_cast<unsigned long long *>(&nbm[tgt >> 6]), 1ull << (tgt & 63));
              }
            }
            const uint64_t mask = __ballot(win);
            if (mask) {
              unsigned long long d = win ? a.off[tgt + 1] - a.off[tgt] : 0;
              d = wave_sum(d);
              unsigned long long base = 0;
              if (lane == 0) {
                base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
                atomicAdd(&nx->deg, d);
              }
              base = __shfl(base, 0, 64);
              if (win) qn[base + mbcnt(mask)] = tgt;
            }
          }
        }
        grid.sync();
        is_dense = false;
      } else {
        if (dir == 3) {
          for (uint64_t i = global_tid(); i < a.nw64; i += grid_threads()) nbm[i] = 0;
          grid.sync();
          // dense forward: one wave per frontier word
          for (uint64_t w = wave; w < nwords; w += nwaves) {
            uint64_t word = fbm[w];
            while (word) {
              const int bp = __builtin_ctzll(word);
              word &= word - 1;
              const uint32_t u = (uint32_t)(w * 64 + bp);
              const uint64_t b = a.off[u], e = a.off[u + 1];
              for (uint64_t k0 = b; k0 < e; k0 += 64) {
                const uint64_t kk = k0 + lane;
                bool win = false;
                uint32_t t = 0;
                if (kk < e) {
                  t = a.ids[kk];
                  if (!bit_get(a.visited64, t) && atomicCAS(&parents[t], kNone, u) == kNone) {
                    win = true;
                    atomicOr(reinterpret_cast<unsigned long long *>(&a.visited64[t >> 6]), 1ull << (t & 63));
                    atomicOr(reinterpret_cast<unsigned long long *>(&nbm[t >> 6]), 1ull << (t & 63));
                  }
                }
                const uint64_t mask = __ballot(win);
                if (mask) {
                  unsigned long long d = wave_sum(win ? (unsigned long long)(a.off[t + 1] - a.off[t]) : 0ull);
                  unsigned long long base = 0;
                  if (lane == 0) {
                    base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
                    atomicAdd(&nx->deg, d);
                  }
                  base = __shfl(base, 0, 64);
                  if (win) qn[base + mbcnt(mask)] = t;
                }
              }
            }
          }
        } else {
          // dense pull: a block owns 64 consecutive bitmap words (4096
          // vertices); its wave i takes words i, i+4, ...
          const uint64_t nchunks = (nwords + 63) / 64;
          const int wv = threadIdx.x >> 6;
          for (uint64_t c = blockIdx.x; c < nchunks; c += nb) {
            for (uint64_t w = c * 64 + wv; w < std::min<uint64_t>(nwords, c * 64 + 64); w += kBlock / 64) {
              const uint64_t vis = a.visited64[w];
              const uint64_t v = w * 64 + lane;
              bool found = false;
              uint64_t dv = 0;
              if (v < a.n && !((vis >> lane) & 1)) {
                const uint64_t b = a.off[v], e = a.off[v + 1];
                dv = e - b;
                for (uint64_t kk = b; kk < e; kk++) {
                  const uint32_t u = a.ids[kk];
                  if (bit_get(fbm, u)) {
                    parents[v] = u;
                    found = true;
                    break;
                  }
                }
              }
              const uint64_t mask = __ballot(found);
              const unsigned long long d = wave_sum(found ? dv : 0ull);
              unsigned long long base = 0;
              if (lane == 0) {
                nbm[w] = mask;
                if (mask) {
                  a.visited64[w] = vis | mask;
                  base = atomicAdd(&nx->size, (unsigned long long)__popcll(mask));
                  atomicAdd(&nx->deg, d);
                }
              }
              base = __shfl(base, 0, 64);
              if (found) qn[base + mbcnt(mask)] = (uint32_t)v;
            }
          }
        }
        grid.sync();
        is_dense = true;
      }
      visited_deg += fdeg;
      uint32_t *tq = q;
      q = qn;
      qn = tq;
      uint64_t *tb = fbm;
      fbm = nbm;
      nbm = tb;
    }
    if (leader) {
      a.nlevels[j] = (uint32_t)L;
      a.tot[2 * j] = reached;
      a.tot[2 * j + 1] = visited_deg;
    }
    grid.sync();
  }
}

}  // namespace

// Runs k BFS in one cooperative launch.  parents_dev: k*n results (or one
// reused n-buffer when keep_all is false; the last source's tree then
// remains).  Fills stats[k] when given.
void bfs_coop_run(byo_graph *g, const uint32_t *srcs, uint32_t k, bool keep_all, byo_bfs_stats *stats,
                  uint64_t maxdeg, double *ms_out) {
  const uint64_t n = g->n;
  hipStream_t s = g->stream;
  CoopArgs a{};
  a.off = g->off.as<uint64_t>();
  a.ids = g->ids.as<uint32_t>();
  a.n = n;
  a.m = g->m;
  a.nw64 = (n + 63) / 64 + 1;
  a.pstride = keep_all ? n : 0;
  a.parents = g->buf(keep_all ? "bfs_parents_all" : "bfs_parents", (keep_all ? k : 1) * n * 4 + 64).as<uint32_t>();
  a.q0 = g->buf("bfs_q", n * 4 + 64).as<uint32_t>();
  a.q1 = g->buf("bfs_qn", n * 4 + 64).as<uint32_t>();
  a.bm0 = g->buf("bfs_fbm", a.nw64 * 8).as<uint64_t>();
  a.bm1 = g->buf("bfs_nbm", a.nw64 * 8).as<uint64_t>();
  a.visited64 = g->buf("bfs_visited", a.nw64 * 8).as<uint64_t>();
  a.scan = g->buf("bfs_scan", (n + 1) * 8 + 64).as<uint64_t>();
  a.ctr = g->buf("bfs_coop_ctr", 3 * sizeof(Ctr) + 64).as<Ctr>();
  a.k = k;
  uint32_t *dsrc = g->buf("bfs_coop_srcs", (size_t)k * 4 + 64).as<uint32_t>();
  a.srcs = dsrc;
  a.logs = g->buf("bfs_coop_logs", (size_t)k * 64 * sizeof(LevelLog) + 64).as<LevelLog>();
  a.nlevels = g->buf("bfs_coop_nlev", (size_t)k * 4 + 64).as<uint32_t>();
  a.tot = g->buf("bfs_coop_tot", (size_t)k * 16 + 64).as<unsigned long long>();
  a.policy = g->dir_policy;
  a.thr = g->dir_threshold < 0 ? g->m / 20 : (uint64_t)g->dir_threshold;
  a.sel = (g->dir_policy == BYO_DIR_MLP && g->sel_loaded) ? g->sel_params.as<float>() : nullptr;
  a.maxdeg = maxdeg;

  static int grid_cache = 0;
  if (grid_cache == 0) {
    int dev = 0, cus = 0, per = 0;
    BYO_HIP(hipGetDevice(&dev));
    BYO_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    BYO_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, reinterpret_cast<const void *>(k_bfs_coop), kBlock, 0));
    BYO_REQUIRE(per >= 1, BYO_E_HIP, "cooperative BFS kernel cannot be resident");
    grid_cache = cus * std::min(per, 4);
  }
  const int nblocks = grid_cache;
  a.bsum = g->buf("bfs_coop_bsum", (size_t)nblocks * 8 + 64).as<unsigned long long>();
  BYO_HIP(hipMemcpyAsync(dsrc, srcs, (size_t)k * 4, hipMemcpyHostToDevice, s));
  void *args[] = {&a};
  PhaseTimer t(s, true);
  BYO_HIP(hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_bfs_coop), dim3(nblocks), dim3(kBlock), args, 0,
                                     s));
  const double ms = t.stop();
  if (ms_out) *ms_out = ms;
  if (stats) {
    std::vector<LevelLog> logs((size_t)k * 64);
    std::vector<uint32_t> nlev(k);
    std::vector<unsigned long long> tot(2 * (size_t)k);
    BYO_HIP(hipMemcpyAsync(tot.data(), a.tot, tot.size() * 8, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipMemcpyAsync(logs.data(), a.logs, logs.size() * sizeof(LevelLog), hipMemcpyDeviceToHost, s));
    BYO_HIP(hipMemcpyAsync(nlev.data(), a.nlevels, k * 4, hipMemcpyDeviceToHost, s));
    BYO_HIP(hipStreamSynchronize(s));
    for (uint32_t j = 0; j < k; j++) {
      byo_bfs_stats &st = stats[j];
      std::memset(&st, 0, sizeof(st));
      st.levels = nlev[j];
      for (uint32_t L = 0; L < std::min<uint32_t>(nlev[j], 64); L++) {
        const LevelLog &lg = logs[(size_t)j * 64 + L];
        st.fsize[L] = lg.fsize;
        st.fdeg[L] = lg.fdeg;
        st.dir[L] = (uint8_t)lg.dir;
      }
      st.reached = (uint32_t)tot[2 * j];
      st.edges_reached = tot[2 * j + 1];
      st.device_ms = ms / k;
    }
  }
}

}  // namespace byo
