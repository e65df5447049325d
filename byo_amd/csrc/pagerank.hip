// pagerank.hip -- PageRank pull (EdgeMap::edgeMapReduce_dense,
// gbbs/edge_map_reduce.h:458-519, driven by benchmarks/PageRank/
// PageRank.h:125-193) with the Jacobi semantics of PageRank_edgeMap
// (PageRank.h:80-123), and batched personalized PageRank.
//
// FP32 path: ranks and contributions p/deg are FP32; each vertex sums its
// in-neighbour contributions with Kahan compensation per lane and a
// TwoSum-compensated combine across the lanes of its group; the global L1
// residual is reduced in FP64 in a fixed order (deterministic).  FP64 path:
// the same kernel with double values and no compensation.
//
// Per iteration one kernel reads offsets (8 B/vertex), ids (4 B/edge), one
// contribution per edge (w B), the old rank (w B/vertex), and writes the new
// rank and contribution (2w B/vertex).
#include <algorithm>
#include <cmath>

#include "primitives.hpp"

namespace byo {
namespace {

template <class T>
struct Acc {
  T s, c;
};

// Kahan step (per lane).
template <class T>
__device__ __forceinline__ void kahan_add(Acc<T> &a, T x) {
  T y = x - a.c;
  T t = a.s + y;
  a.c = (t - a.s) - y;
  a.s = t;
}

// Compensated combine of two partial sums (TwoSum on the heads).
template <class T>
__device__ __forceinline__ Acc<T> combine(Acc<T> a, Acc<T> b) {
  T s = a.s + b.s;
  T bp = s - a.s;
  T err = (a.s - (s - bp)) + (b.s - bp);
  // Kahan carries the negated error in c (value = s - c)
  return Acc<T>{s, a.c + b.c - err};
}

template <class T, int G, bool kCompensated>
__global__ void __launch_bounds__(256) k_pr_iter(const uint64_t *off, const uint32_t *ids, uint64_t n,
                                                 const T *contrib_in, const T *p_old, T *p_new, T *contrib_out,
                                                 double damping, double addc, double *partial) {
  __shared__ double lds[4];
  const int lane = lane_id();
  const int sub = lane & (G - 1);
  const uint64_t group = global_tid() / G, ngroups = grid_threads() / G;
  double resid = 0.0;
  const T d = (T)damping, t = (T)addc;
  const uint64_t nrounds = (n + ngroups - 1) / ngroups;
  for (uint64_t r = 0; r < nrounds; r++) {
    const uint64_t v = r * ngroups + group;
    Acc<T> a{0, 0};
    uint64_t b = 0, e = 0;
    if (v < n) {
      b = off[v];
      e = off[v + 1];
      for (uint64_t k = b + sub; k < e; k += G) {
        T x = contrib_in[ids[k]];
        if (kCompensated) kahan_add(a, x); else a.s += x;
      }
    }
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
      Acc<T> other{__shfl_xor(a.s, o, 64), __shfl_xor(a.c, o, 64)};
      if (kCompensated) a = combine(a, other); else a.s += other.s;
    }
    if (v < n && sub == 0) {
      const T sum = kCompensated ? (a.s - a.c) : a.s;
      const T pv = d * sum + t;
      const uint64_t deg = e - b;
      p_new[v] = pv;
      contrib_out[v] = deg ? pv / (T)deg : (T)0;
      resid += fabs((double)pv - (double)p_old[v]);
    }
  }
  resid = wave_sum(resid);
  if (lane == 0) lds[threadIdx.x >> 6] = resid;
  __syncthreads();
  if (threadIdx.x == 0) partial[blockIdx.x] = (lds[0] + lds[1]) + (lds[2] + lds[3]);
}

__global__ void k_sum_doubles(const double *partial, int nb, double *out) {
  // one block, fixed association order
  __shared__ double lds[1024];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) s += partial[i];
  lds[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) lds[threadIdx.x] += lds[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = lds[0];
}

template <class T>
__global__ void k_pr_init(const uint64_t *off, uint64_t n, T *p, T *contrib) {
  const double p0 = 1.0 / (double)n;
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) {
    const uint64_t deg = off[v + 1] - off[v];
    p[v] = (T)p0;
    contrib[v] = deg ? (T)(p0 / (double)deg) : (T)0;
  }
}

template <class T>
__global__ void k_to_double(const T *in, double *out, uint64_t n) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) out[v] = (double)in[v];
}

// ---------------------------------------------- personalized PageRank, k seeds
// Vertex-major state p[v*K + j]: one wave per vertex, lane j = seed j, so
// every edge gathers K consecutive floats (one 128 B line for K = 32).
__global__ void k_ppr_init(const uint64_t *off, uint64_t n, const uint32_t *seeds, int K, float *p, float *contrib) {
  for (uint64_t i = global_tid(); i < n * (uint64_t)K; i += grid_threads()) {
    const uint64_t v = i / K;
    const int j = (int)(i % K);
    const uint64_t deg = off[v + 1] - off[v];
    const float x = (seeds[j] == v) ? 1.0f : 0.0f;
    p[i] = x;
    contrib[i] = deg ? x / (float)deg : 0.0f;
  }
}

__global__ void __launch_bounds__(256) k_ppr_iter(const uint64_t *off, const uint32_t *ids, uint64_t n,
                                                  const uint32_t *seeds, int K, const float *contrib_in,
                                                  float *p_new, float *contrib_out, float damping) {
  const int lane = lane_id();
  const uint64_t wave = global_tid() >> 6, nwaves = grid_threads() >> 6;
  const uint32_t my_seed = lane < K ? seeds[lane] : kNone;
  for (uint64_t v = wave; v < n; v += nwaves) {
    const uint64_t b = off[v], e = off[v + 1];
    Acc<float> a{0.f, 0.f};
    if (lane < K) {
      for (uint64_t k = b; k < e; k++) kahan_add(a, contrib_in[(uint64_t)ids[k] * K + lane]);
      const float pv = damping * (a.s - a.c) + (my_seed == v ? 1.0f - damping : 0.0f);
      p_new[v * K + lane] = pv;
      contrib_out[v * K + lane] = (e > b) ? pv / (float)(e - b) : 0.0f;
    }
  }
}

template <class T>
struct PrState {
  T *p[2], *c[2];
};

template <class T, bool kComp>
void pr_launch(int G, const uint64_t *off, const uint32_t *ids, uint64_t n, const T *cin, const T *pold, T *pnew,
               T *cout, double damping, double addc, double *partial, unsigned grid, hipStream_t s) {
  switch (G) {
    case 4: k_pr_iter<T, 4, kComp><<<grid, 256, 0, s>>>(off, ids, n, cin, pold, pnew, cout, damping, addc, partial); break;
    case 8: k_pr_iter<T, 8, kComp><<<grid, 256, 0, s>>>(off, ids, n, cin, pold, pnew, cout, damping, addc, partial); break;
    case 16: k_pr_iter<T, 16, kComp><<<grid, 256, 0, s>>>(off, ids, n, cin, pold, pnew, cout, damping, addc, partial); break;
    default: k_pr_iter<T, 32, kComp><<<grid, 256, 0, s>>>(off, ids, n, cin, pold, pnew, cout, damping, addc, partial); break;
  }
}

template <class T, bool kComp>
void pagerank_impl(byo_graph *g, double damping, int max_iters, double eps, double *p_out, double *l1_log,
                   int *iters_run, double *device_ms, const char *phase) {
  const uint64_t n = g->n;
  hipStream_t s = g->stream;
  PrState<T> st;
  st.p[0] = g->buf("pr_p0", n * sizeof(T) + 64).as<T>();
  st.p[1] = g->buf("pr_p1", n * sizeof(T) + 64).as<T>();
  st.c[0] = g->buf("pr_c0", n * sizeof(T) + 64).as<T>();
  st.c[1] = g->buf("pr_c1", n * sizeof(T) + 64).as<T>();
  const unsigned grid = 4096;
  double *partial = g->buf("pr_partial", (grid + 8) * sizeof(double)).as<double>();
  double *l1dev = g->buf("pr_l1", (size_t)(max_iters + 1) * sizeof(double)).as<double>();
  const uint64_t *off = g->off.as<uint64_t>();
  const uint32_t *ids = g->ids.as<uint32_t>();
  // lanes per vertex: about half the average degree, in [4, 32]
  const double avg = n ? (double)g->m / (double)n : 0.0;
  int G = 4;
  while (G < 32 && G < avg / 2) G *= 2;
  const double addc = (1.0 - damping) * (1.0 / (double)n);
  k_pr_init<T><<<byo_grid(n, 256, 16384), 256, 0, s>>>(off, n, st.p[0], st.c[0]);
  BYO_LAUNCH_CHECK();
  double *hl1 = static_cast<double *>(pinned(g, 64));
  int it = 0, cur = 0;
  PhaseTimer timer(s, true);
  double ms_sum = 0.0;
  while (it < max_iters) {
    PhaseTimer it_t(s, g->prof.on);
    pr_launch<T, kComp>(G, off, ids, n, st.c[cur], st.p[cur], st.p[cur ^ 1], st.c[cur ^ 1], damping, addc, partial,
                        grid, s);
    k_sum_doubles<<<1, 1024, 0, s>>>(partial, grid, l1dev + it);
    BYO_LAUNCH_CHECK();
    if (g->prof.on) g->account(phase, it_t.stop());
    cur ^= 1;
    it++;
    if (eps > 0) {
      BYO_HIP(hipMemcpyAsync(hl1, l1dev + it - 1, sizeof(double), hipMemcpyDeviceToHost, s));
      BYO_HIP(hipStreamSynchronize(s));
      if (*hl1 < eps) break;
    }
  }
  ms_sum = timer.stop();
  if (device_ms) *device_ms = ms_sum;
  if (iters_run) *iters_run = it;
  if (l1_log && it) {
    BYO_HIP(hipMemcpyAsync(l1_log, l1dev, it * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  if (p_out) {
    double *pd = g->buf("pr_out_f64", n * sizeof(double) + 64).as<double>();
    k_to_double<T><<<byo_grid(n, 256, 16384), 256, 0, s>>>(st.p[cur], pd, n);
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(p_out, pd, n * sizeof(double), hipMemcpyDeviceToHost, s));
  }
  BYO_HIP(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace byo

using namespace byo;

extern "C" {

int byo_pagerank(byo_graph_t g, double damping, int max_iters, double eps, int precision, double *p, double *l1_log,
                 int *iters_run, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    BYO_REQUIRE(max_iters >= 0, BYO_E_INVAL, "max_iters < 0");
    BYO_REQUIRE(damping >= 0 && damping <= 1, BYO_E_INVAL, "damping not in [0,1]");
    BYO_REQUIRE(precision == BYO_PR_FP32 || precision == BYO_PR_FP64, BYO_E_INVAL, "unknown precision");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    if (g->n == 0) {
      if (iters_run) *iters_run = 0;
      return;
    }
    if (precision == BYO_PR_FP32)
      pagerank_impl<float, true>(g, damping, max_iters, eps, p, l1_log, iters_run, device_ms, "pr_fp32");
    else
      pagerank_impl<double, false>(g, damping, max_iters, eps, p, l1_log, iters_run, device_ms, "pr_fp64");
  });
}

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
    const uint64_t n = g->n;
    hipStream_t s = g->stream;
    const uint64_t sz = n * k * sizeof(float) + 64;
    float *pa = g->buf("ppr_p", sz).as<float>();
    float *c0 = g->buf("ppr_c0", sz).as<float>();
    float *c1 = g->buf("ppr_c1", sz).as<float>();
    uint32_t *dseeds = g->buf("ppr_seeds", 64 * 4).as<uint32_t>();
    BYO_HIP(hipMemcpyAsync(dseeds, seeds, k * 4, hipMemcpyHostToDevice, s));
    const uint64_t *off = g->off.as<uint64_t>();
    const uint32_t *ids = g->ids.as<uint32_t>();
    k_ppr_init<<<byo_grid(n * k, 256, 16384), 256, 0, s>>>(off, n, dseeds, (int)k, pa, c0);
    BYO_LAUNCH_CHECK();
    PhaseTimer t(s, true);
    float *cin = c0, *cout = c1;
    for (int it = 0; it < iters; it++) {
      k_ppr_iter<<<8192, 256, 0, s>>>(off, ids, n, dseeds, (int)k, cin, pa, cout, (float)damping);
      std::swap(cin, cout);
    }
    BYO_LAUNCH_CHECK();
    double ms = t.stop();
    if (g->prof.on) g->account("ppr", ms);
    if (device_ms) *device_ms = ms;
    if (p) {
      std::vector<float> h(n * k);
      BYO_HIP(hipMemcpyAsync(h.data(), pa, n * k * sizeof(float), hipMemcpyDeviceToHost, s));
      BYO_HIP(hipStreamSynchronize(s));
      for (uint64_t v = 0; v < n; v++)
        for (uint32_t j = 0; j < k; j++) p[(uint64_t)j * n + v] = h[v * k + j];
    }
    BYO_HIP(hipStreamSynchronize(s));
  });
}

}  // extern "C"
