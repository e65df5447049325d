// The host concept classes under the reference's Graph_API_use ablations
// (gbbs/graph.h:2134-2152, benchmarks/run_api/): a Graph<>-style harness
// detects the container's members with requires-expressions, as graph.h
// does (:2163-2420), and an API policy withholds degree, M, parallel maps
// or early exit, which the harness then derives from map_neighbors.  Each
// variant runs a direction-optimizing BFS written against the harness (top
// down through map / parallel map, bottom up through the early-exit map,
// M/20 rule) plus degree and M queries; every variant must give the same
// answers.  Prints one JSON object per variant.
#include <cstdio>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

#include "byo_graph.hpp"

template <bool degree_, bool M_, bool parallel_map_, bool early_exit_>
struct Api {
  static constexpr bool degree = degree_, M = M_, parallel_map = parallel_map_, early_exit = early_exit_;
};
// the reference's variants (graph.h:2147-2152) restricted to these four knobs
using full_api = Api<true, true, true, true>;
using minimal_api = Api<false, false, false, false>;
using no_early_exit = Api<true, true, true, false>;
using no_parallel_map = Api<true, true, false, true>;
using no_degree = Api<false, true, true, true>;
using no_M = Api<true, false, true, true>;

template <class G, class A>
struct Harness {
  static constexpr bool native_degree = requires(const G &g, uint64_t i) { g.degree(i); };
  static constexpr bool native_M = requires(const G &g) { g.M(); };
  static constexpr bool native_early = requires(const G &g, size_t i) {
    g.map_neighbors_early_exit(i, [](uint32_t, uint32_t, byo::empty) { return true; });
  };
  static constexpr bool native_parallel = requires(const G &g, size_t i) {
    g.parallel_map_neighbors(i, [](uint32_t, uint32_t, byo::empty) {});
  };
  static_assert(requires(const G &g) { g.N(); }, "the container must give N()");
  static_assert(requires(const G &g, size_t i) { g.map_neighbors(i, [](uint32_t, uint32_t, byo::empty) {}); },
                "the container must give map_neighbors");

  const G &g;
  std::vector<uint32_t> deg;
  size_t m = 0;

  explicit Harness(const G &gr) : g(gr) {
    if constexpr (!(A::degree && native_degree)) {
      deg.assign(g.N(), 0);
      for (size_t i = 0; i < g.N(); i++) g.map_neighbors(i, [&](uint32_t, uint32_t, byo::empty) { deg[i]++; });
    }
    if constexpr (A::M && native_M) {
      m = g.M();
    } else {
      for (size_t i = 0; i < g.N(); i++) m += degree(i);
    }
  }
  size_t N() const { return g.N(); }
  size_t M() const { return m; }
  size_t degree(size_t i) const {
    if constexpr (A::degree && native_degree) return g.degree(i);
    else return deg[i];
  }
  template <class F>
  void map(size_t i, F f) const {
    if constexpr (A::parallel_map && native_parallel) g.parallel_map_neighbors(i, f);
    else g.map_neighbors(i, f);
  }
  template <class F>
  void map_early_exit(size_t i, F f) const {
    if constexpr (A::early_exit && native_early) {
      g.map_neighbors_early_exit(i, f);
    } else {
      bool stop = false;
      g.map_neighbors(i, [&](uint32_t s, uint32_t d, byo::empty w) {
        if (!stop) stop = f(s, d, w);
      });
    }
  }
};

// Direction-optimizing BFS over the harness (edge_map_data.h:257-306 rule).
template <class H>
std::vector<uint32_t> bfs_levels(const H &h, uint32_t src, size_t *pulls) {
  const size_t n = h.N();
  std::vector<uint32_t> lv(n, BYO_NONE);
  std::vector<uint32_t> front{src};
  lv[src] = 0;
  *pulls = 0;
  for (uint32_t L = 0; !front.empty(); L++) {
    size_t fdeg = 0;
    for (uint32_t u : front) fdeg += h.degree(u);
    std::vector<uint32_t> next;
    if (front.size() + fdeg > h.M() / 20) {
      (*pulls)++;
      for (size_t v = 0; v < n; v++) {
        if (lv[v] != BYO_NONE) continue;
        h.map_early_exit(v, [&](uint32_t, uint32_t u, byo::empty) {
          if (lv[u] == L) {
            lv[v] = L + 1;
            next.push_back((uint32_t)v);
            return true;
          }
          return false;
        });
      }
    } else {
      for (uint32_t u : front)
        h.map(u, [&](uint32_t, uint32_t v, byo::empty) {
          if (lv[v] == BYO_NONE) {
            lv[v] = L + 1;
            next.push_back(v);
          }
        });
    }
    front.swap(next);
  }
  return lv;
}

template <class A>
void run(const char *name, const byo::CsrGraph &g, uint32_t src) {
  Harness<byo::CsrGraph, A> h(g);
  size_t pulls = 0;
  auto lv = bfs_levels(h, src, &pulls);
  unsigned long long lvsum = 0, reached = 0, degsum = 0;
  for (auto x : lv)
    if (x != BYO_NONE) {
      lvsum += x;
      reached++;
    }
  for (size_t i = 0; i < h.N(); i++) degsum += h.degree(i);
  std::printf("{\"api\": \"%s\", \"n\": %zu, \"m\": %zu, \"degsum\": %llu, \"reached\": %llu, \"level_sum\": %llu, "
              "\"pull_levels\": %zu}\n",
              name, h.N(), h.M(), degsum, reached, lvsum, pulls);
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  std::ifstream in(argv[1]);
  const uint32_t src = (uint32_t)std::stoul(argv[2]);
  std::string hdr;
  size_t n, m;
  in >> hdr >> n >> m;
  std::vector<byo::vertex_data> v(n);
  std::vector<uint64_t> off(n + 1);
  for (size_t i = 0; i < n; i++) in >> off[i];
  off[n] = m;
  std::vector<std::tuple<uint32_t, byo::empty>> e0(m);
  for (size_t e = 0; e < m; e++) {
    uint32_t x;
    in >> x;
    e0[e] = std::make_tuple(x, byo::empty{});
  }
  for (size_t i = 0; i < n; i++) v[i] = byo::vertex_data{off[i], (uint32_t)(off[i + 1] - off[i])};
  byo::CsrGraph g(v.data(), n, m, nullptr, e0.data());
  run<full_api>("full_api", g, src);
  run<minimal_api>("minimal_api", g, src);
  run<no_early_exit>("no_early_exit", g, src);
  run<no_parallel_map>("no_parallel_map", g, src);
  run<no_degree>("no_degree", g, src);
  run<no_M>("no_M", g, src);
  return 0;
}
