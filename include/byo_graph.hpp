// byo_graph.hpp -- header-only C++ host classes that put the reference's
// container concept back on top of the libbyo C ABI (byo.h).
//
// byo::CsrGraph and byo::CpmaGraph satisfy the requirements Graph<> detects
// with requires-expressions (gbbs/graph.h:2163-2420): N(), M(), degree(i),
// map_neighbors(i, F), map_neighbors_early_exit(i, F),
// parallel_map_neighbors(i, F), insert_batch / remove_batch
// (std::tuple<uint32_t, uint32_t>*, size_t) for the CPMA, get_memory_size(),
// weight_type = empty.  Their constructor takes the same (v_data, n, m,
// deletion_fn, e0) arguments as CSR / PMA_graph (benchmarks/run_structures/
// run_csr.cc:87-106, run_single_pma.cc:75-89).  The graph lives in HBM; the
// map_* functors run on the host over a CSR snapshot fetched once and
// refreshed after updates (the verification path), while the algorithms
// (bfs, pagerank, cc, tc, map_sum) run on the device.
//
// Errors from the C ABI are thrown as byo::Error (message from
// byo_last_error()).
#pragma once
#include <cstdint>
#include <cstring>
#include <functional>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "byo.h"

namespace byo {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc) {
  if (rc != BYO_OK) throw Error(rc, byo_last_error());
}

struct empty {};

// {offset, degree} per vertex, the vertex_data of gbbs/graph_io.h:117-124.
struct vertex_data {
  uint64_t offset;
  uint32_t degree;
};

class GraphBase {
 public:
  using weight_type = empty;
  using vertex_weight_type = double;

  GraphBase(const GraphBase &) = delete;
  GraphBase &operator=(const GraphBase &) = delete;
  GraphBase(GraphBase &&o) noexcept : h_(o.h_), snap_valid_(o.snap_valid_) {
    o.h_ = nullptr;
    off_.swap(o.off_);
    ids_.swap(o.ids_);
  }
  virtual ~GraphBase() {
    if (h_) byo_graph_destroy(h_);
  }

  size_t N() const {
    uint64_t n = 0;
    check(byo_num_vertices(h_, &n));
    return n;
  }
  size_t M() const {
    uint64_t m = 0;
    check(byo_num_edges(h_, &m));
    return m;
  }
  size_t num_vertices() const { return N(); }
  size_t num_edges() const { return M(); }
  size_t degree(size_t i) const {
    uint32_t d = 0;
    check(byo_degree(h_, (uint32_t)i, &d));
    return d;
  }
  size_t get_memory_size() const {
    uint64_t b = 0;
    check(byo_memory_size(h_, &b));
    return b;
  }

  // F(src, dst, weight) -> void
  template <class F>
  void map_neighbors(size_t i, F f) const {
    snapshot();
    weight_type w{};
    for (uint64_t e = off_[i]; e < off_[i + 1]; e++) f((uint32_t)i, ids_[e], w);
  }
  // F(src, dst, weight) -> bool (true = stop)
  template <class F>
  void map_neighbors_early_exit(size_t i, F f) const {
    snapshot();
    weight_type w{};
    for (uint64_t e = off_[i]; e < off_[i + 1]; e++)
      if (f((uint32_t)i, ids_[e], w)) break;
  }
  template <class F>
  void parallel_map_neighbors(size_t i, F f) const {
    map_neighbors(i, f);
  }

  // ---- whole-algorithm entry points (device) ----
  std::vector<uint32_t> bfs(uint32_t src, byo_bfs_stats *stats = nullptr) const {
    std::vector<uint32_t> p(N());
    check(byo_bfs(h_, src, p.data(), stats));
    return p;
  }
  std::vector<double> pagerank(double damping = 0.85, int iters = 20, double eps = 0.0,
                               int precision = BYO_PR_FP32) const {
    std::vector<double> p(N());
    int it = 0;
    check(byo_pagerank(h_, damping, iters, eps, precision, p.data(), nullptr, &it, nullptr));
    return p;
  }
  std::vector<uint32_t> cc() const {
    std::vector<uint32_t> l(N());
    check(byo_cc(h_, l.data(), nullptr));
    return l;
  }
  uint64_t tc() const {
    uint64_t c = 0;
    check(byo_tc(h_, &c, nullptr));
    return c;
  }
  uint64_t map_sum() const {
    uint64_t s = 0;
    check(byo_map_sum(h_, &s, nullptr));
    return s;
  }
  byo_graph_t handle() const { return h_; }

 protected:
  GraphBase() = default;
  static void to_csr(const vertex_data *v, size_t n, size_t m, const std::tuple<uint32_t, empty> *e0,
                     std::vector<uint64_t> &off, std::vector<uint32_t> &ids) {
    off.assign(n + 1, 0);
    ids.resize(m);
    for (size_t i = 0; i < n; i++) {
      off[i] = v[i].offset;
      for (uint32_t j = 0; j < v[i].degree; j++) ids[v[i].offset + j] = std::get<0>(e0[v[i].offset + j]);
    }
    off[n] = m;
  }
  void snapshot() const {
    if (snap_valid_) return;
    off_.resize(N() + 1);
    ids_.resize(M());
    check(byo_export_csr(h_, off_.data(), ids_.data()));
    snap_valid_ = true;
  }
  byo_graph_t h_ = nullptr;
  mutable std::vector<uint64_t> off_;
  mutable std::vector<uint32_t> ids_;
  mutable bool snap_valid_ = false;
};

// Static CSR (run_csr.cc:28-191).
class CsrGraph : public GraphBase {
 public:
  CsrGraph(const vertex_data *v_data, size_t n, size_t m, std::function<void()> deletion_fn,
           const std::tuple<uint32_t, empty> *e0, vertex_weight_type * = nullptr) {
    std::vector<uint64_t> off;
    std::vector<uint32_t> ids;
    to_csr(v_data, n, m, e0, off, ids);
    if (deletion_fn) deletion_fn();
    check(byo_csr_create(off.data(), ids.data(), n, m, &h_));
  }
  CsrGraph(const uint64_t *offsets, const uint32_t *ids, size_t n, size_t m) {
    check(byo_csr_create(offsets, ids, n, m, &h_));
  }
};

// CPMA-style packed-memory-array graph (run_single_pma.cc:29-140).
class CpmaGraph : public GraphBase {
 public:
  CpmaGraph(const vertex_data *v_data, size_t n, size_t m, std::function<void()> deletion_fn,
            const std::tuple<uint32_t, empty> *e0, vertex_weight_type * = nullptr) {
    std::vector<uint64_t> off;
    std::vector<uint32_t> ids;
    to_csr(v_data, n, m, e0, off, ids);
    if (deletion_fn) deletion_fn();
    check(byo_cpma_create(off.data(), ids.data(), n, m, &h_));
  }
  CpmaGraph(const uint64_t *offsets, const uint32_t *ids, size_t n, size_t m) {
    check(byo_cpma_create(offsets, ids, n, m, &h_));
  }
  // Graph::insert_batch / remove_batch (gbbs/graph.h:3071-3124): the tuple
  // array is passed as is -- its bytes are the uint64 keys (src<<32)|dst.
  void insert_batch(std::tuple<uint32_t, uint32_t> *es, size_t n) {
    check(byo_insert_batch(h_, keys_of(es, n), n));
    snap_valid_ = false;
  }
  void remove_batch(std::tuple<uint32_t, uint32_t> *es, size_t n) {
    check(byo_remove_batch(h_, keys_of(es, n), n));
    snap_valid_ = false;
  }

 private:
  // libstdc++ stores the last tuple element first, so on little-endian
  // hosts the 8 bytes of tuple<src, dst> read as (src << 32) | dst.  This
  // is verified once at run time; other layouts are repacked.
  const uint64_t *keys_of(std::tuple<uint32_t, uint32_t> *es, size_t n) {
    static_assert(sizeof(std::tuple<uint32_t, uint32_t>) == 8, "tuple<u32,u32> must be 8 bytes");
    std::tuple<uint32_t, uint32_t> probe{1u, 2u};
    uint64_t k = 0;
    std::memcpy(&k, &probe, 8);
    if (k == ((1ull << 32) | 2ull)) return reinterpret_cast<const uint64_t *>(es);
    buf_.resize(n);
    for (size_t i = 0; i < n; i++) buf_[i] = ((uint64_t)std::get<0>(es[i]) << 32) | std::get<1>(es[i]);
    return buf_.data();
  }
  std::vector<uint64_t> buf_;
};

}  // namespace byo
