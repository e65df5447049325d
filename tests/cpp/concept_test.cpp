// Drives libbyo through the host concept classes (include/byo_graph.hpp) the
// way the reference harness drives a container: build from
// (vertex_data, e0), query N/M/degree, map_neighbors, run the algorithms,
// insert/remove a tuple<u32,u32> batch.  Prints one JSON object; the pytest
// wrapper checks it against the oracle.
#include <cstdio>
#include <fstream>
#include <string>
#include <tuple>
#include <vector>

#include "byo_graph.hpp"

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  std::ifstream in(argv[1]);
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
  bool deleted = false;
  byo::CsrGraph g(v.data(), n, m, [&] { deleted = true; }, e0.data());
  uint64_t degsum = 0, dstsum = 0;
  for (size_t i = 0; i < g.N(); i++) {
    degsum += g.degree(i);
    g.map_neighbors(i, [&](uint32_t, uint32_t d, byo::empty) { dstsum += d; });
  }
  size_t first_stop = 0;
  g.map_neighbors_early_exit(0, [&](uint32_t, uint32_t, byo::empty) { return ++first_stop == 1; });
  auto par = g.bfs(0);
  auto lab = g.cc();
  uint64_t tri = g.tc();
  uint64_t ms = g.map_sum();
  uint32_t reached = 0;
  for (auto p : par) reached += p != BYO_NONE;

  byo::CpmaGraph c(v.data(), n, m, nullptr, e0.data());
  std::vector<std::tuple<uint32_t, uint32_t>> batch = {{0, 5}, {5, 0}, {1, 2}, {2, 1}, {0, 5}};
  c.insert_batch(batch.data(), batch.size());
  size_t m_ins = c.M();
  c.remove_batch(batch.data(), batch.size());
  size_t m_rem = c.M();
  uint32_t ncomp = 0;
  for (size_t i = 0; i < lab.size(); i++) ncomp += lab[i] == i;
  std::printf("{\"n\": %zu, \"m\": %zu, \"deleted\": %d, \"degsum\": %llu, \"dstsum\": %llu, \"first_stop\": %zu, "
              "\"reached\": %u, \"components\": %u, \"triangles\": %llu, \"map_sum\": %llu, \"m_after_insert\": %zu, "
              "\"m_after_remove\": %zu, \"mem\": %zu}\n",
              g.N(), g.M(), deleted ? 1 : 0, (unsigned long long)degsum, (unsigned long long)dstsum, first_stop, reached,
              ncomp, (unsigned long long)tri, (unsigned long long)ms, m_ins, m_rem, g.get_memory_size());
  return 0;
}
