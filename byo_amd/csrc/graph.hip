// graph.hip -- handle management, the static CSR container, the device
// RMAT generator, container queries and the Map edge-scan kernel.
#include <cstdio>

#include "primitives.hpp"

namespace byo {

static thread_local std::string g_last_error;
void set_last_error(const std::string &m) { g_last_error = m; }

void require_device() {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    throw Error(BYO_E_NODEV, "no HIP device visible");
  }
}

void bind(byo_graph *g) { BYO_HIP(hipSetDevice(g->device)); }

byo_graph *new_handle(int kind) {
  require_device();
  auto *g = new byo_graph();
  g->kind = kind;
  BYO_HIP(hipGetDevice(&g->device));
  BYO_HIP(hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking));
  return g;
}

void *pinned(byo_graph *g, size_t bytes) {
  if (g->pinned_bytes < bytes) {
    if (g->pinned) (void)hipHostFree(g->pinned);
    g->pinned = nullptr;
    g->pinned_bytes = 0;
    BYO_HIP(hipHostMalloc(&g->pinned, bytes, hipHostMallocDefault));
    g->pinned_bytes = bytes;
  }
  return g->pinned;
}

namespace {

// ---------------------------------------------------------- RMAT generator
__device__ __forceinline__ uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Edge i of the seeded RMAT stream, emitted as its two directed keys
// (self loops become two copies of the sentinel ~0).
__global__ void k_rmat_sym_keys(int scale, uint64_t ne, double a, double ab, double abc,
                                uint64_t base, uint64_t *keys) {
  for (uint64_t i = global_tid(); i < ne; i += grid_threads()) {
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; l++) {
      uint64_t h = splitmix64(base + i * (uint64_t)scale + (uint64_t)l);
      double r = (double)(h >> 11) * 0x1.0p-53;
      uint32_t bit = 1u << (scale - 1 - l);
      if (r < a) {
      } else if (r < ab) {
        v |= bit;
      } else if (r < abc) {
        u |= bit;
      } else {
        u |= bit;
        v |= bit;
      }
    }
    bool loop = (u == v);
    keys[2 * i] = loop ? ~0ull : ((uint64_t)u << 32) | v;
    keys[2 * i + 1] = loop ? ~0ull : ((uint64_t)v << 32) | u;
  }
}

// Erdos-Renyi G(n, M): edge i joins splitmix64(base + 2i) mod n and
// splitmix64(base + 2i + 1) mod n; emitted as its two directed keys.
__global__ void k_er_sym_keys(uint64_t n, uint64_t ne, uint64_t base, uint64_t *keys) {
  for (uint64_t i = global_tid(); i < ne; i += grid_threads()) {
    const uint64_t u = splitmix64(base + 2 * i) % n, v = splitmix64(base + 2 * i + 1) % n;
    const bool loop = (u == v);
    keys[2 * i] = loop ? ~0ull : (u << 32) | v;
    keys[2 * i + 1] = loop ? ~0ull : (v << 32) | u;
  }
}

// side^3 grid, vertex id x + side*(y + side*z), 6-neighbour: each vertex
// emits its +x, +y, +z edges in both directions (6 keys, ~0 off the grid).
__global__ void k_grid3d_sym_keys(uint64_t side, uint64_t *keys) {
  const uint64_t n = side * side * side;
  for (uint64_t id = global_tid(); id < n; id += grid_threads()) {
    const uint64_t x = id % side, y = (id / side) % side, z = id / (side * side);
    const uint64_t nb[3] = {x + 1 < side ? id + 1 : ~0ull, y + 1 < side ? id + side : ~0ull,
                            z + 1 < side ? id + side * side : ~0ull};
    for (int k = 0; k < 3; k++) {
      const uint64_t t = nb[k];
      keys[6 * id + 2 * k] = t == ~0ull ? ~0ull : (id << 32) | t;
      keys[6 * id + 2 * k + 1] = t == ~0ull ? ~0ull : (t << 32) | id;
    }
  }
}

// batch_insert_helpers::hash32 (benchmarks/run_unweighted.h:720-728)
__device__ __forceinline__ uint32_t ref_hash32(uint32_t a) {
  a = (a + 0x7ed55d16u) + (a << 12);
  a = (a ^ 0xc761c23cu) ^ (a >> 19);
  a = (a + 0x165667b1u) + (a << 5);
  a = (a + 0xd3a2646cu) ^ (a << 9);
  a = (a + 0xfd7046c5u) + (a << 3);
  a = (a ^ 0xb55a4f09u) ^ (a >> 16);
  return a;
}

// rMat::operator()(i) / rMatRec (run_unweighted.h:755-777), iterated from
// the top level down; a = 0.5, b = c = 0.1 as Batch_insert_runner sets them.
__global__ void k_update_rmat(uint32_t nn, uint32_t h, uint64_t first, uint64_t count, uint64_t *keys) {
  const double a = 0.5, ab = 0.5 + 0.1, abc = 0.5 + 0.1 + 0.1;
  for (uint64_t k = global_tid(); k < count; k += grid_threads()) {
    const uint32_t i = (uint32_t)(first + k);
    uint32_t start = ref_hash32((uint32_t)(2u * i) * h);
    const uint32_t stride = ref_hash32((uint32_t)(2u * i + 1u) * h);
    uint32_t row = 0, col = 0;
    for (uint32_t size = nn; size > 1; size >>= 1) {
      const double r = (double)ref_hash32(start) / 2147483647.0;
      if (r < a) {
      } else if (r < ab) {
        col += size >> 1;
      } else if (r < abc) {
        row += size >> 1;
      } else {
        row += size >> 1;
        col += size >> 1;
      }
      start += stride;
    }
    keys[k] = ((uint64_t)row << 32) | col;
  }
}

// H: ids is the CPMA's dst word array, whose empty slots (kNone) add nothing.
template <bool H>
__device__ __forceinline__ uint64_t id_val(uint32_t x) {
  return (H && x == kNone) ? 0 : (uint64_t)x;
}
template <bool H>
__global__ void k_map_sum(const uint32_t *ids, uint64_t m, uint64_t *partial) {
  __shared__ uint64_t lds[16];
  uint64_t acc = 0;
  const uint64_t nvec = m / 4;
  const uint4 *v4 = reinterpret_cast<const uint4 *>(ids);
  for (uint64_t i = global_tid(); i < nvec; i += grid_threads()) {
    uint4 q = v4[i];
    acc += id_val<H>(q.x) + id_val<H>(q.y) + id_val<H>(q.z) + id_val<H>(q.w);
  }
  for (uint64_t i = nvec * 4 + global_tid(); i < m; i += grid_threads()) acc += id_val<H>(ids[i]);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

// Test_Map_Remote (benchmarks/Map/RemoteRead/Map.h:22-34): the sum over all
// edges of remote[dst], remote holding n elements of W bytes (W = 0: one bit
// per vertex, bit v at byte v/8, bit v%8).  Elements wider than 8 bytes are
// read as the uint64 at their start, the value of the reference's
// wide_int<size_t, W>; their stride still spreads the gathers over W*n
// bytes.  Every thread gathers 8 elements at a time from 8 coalesced ids.
template <int W>
__device__ __forceinline__ uint64_t remote_at(const uint8_t *r, uint32_t d) {
  if constexpr (W == 0) return (r[d >> 3] >> (d & 7)) & 1;
  else if constexpr (W == 1) return r[d];
  else if constexpr (W == 2) return reinterpret_cast<const uint16_t *>(r)[d];
  else if constexpr (W == 4) return reinterpret_cast<const uint32_t *>(r)[d];
  else return *reinterpret_cast<const uint64_t *>(r + (uint64_t)d * W);
}

template <int W>
__global__ void k_map_remote(const uint32_t *ids, uint64_t m, const uint8_t *remote, uint64_t *partial) {
  __shared__ uint64_t lds[16];
  uint64_t acc = 0;
  const uint64_t n8 = m / 8;
  const uint4 *v4 = reinterpret_cast<const uint4 *>(ids);
  for (uint64_t i = global_tid(); i < n8; i += grid_threads()) {
    const uint4 a = v4[2 * i], b = v4[2 * i + 1];
    const uint32_t d[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint64_t x[8];
#pragma unroll
    for (int j = 0; j < 8; j++) x[j] = remote_at<W>(remote, d[j]);
#pragma unroll
    for (int j = 0; j < 8; j++) acc += x[j];
  }
  for (uint64_t i = n8 * 8 + global_tid(); i < m; i += grid_threads()) acc += remote_at<W>(remote, ids[i]);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    partial[blockIdx.x] = s;
  }
}

__global__ void k_sum_partials(const uint64_t *partial, int nb, uint64_t *out) {
  __shared__ uint64_t lds[16];
  uint64_t acc = 0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += partial[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint64_t s = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += lds[w];
    *out = s;
  }
}

__global__ void k_degrees(const uint64_t *off, uint64_t n, uint32_t *deg) {
  for (uint64_t v = global_tid(); v < n; v += grid_threads()) deg[v] = (uint32_t)(off[v + 1] - off[v]);
}

void check_host_csr(const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m) {
  BYO_REQUIRE(offsets != nullptr, BYO_E_INVAL, "offsets is null");
  BYO_REQUIRE(m == 0 || ids != nullptr, BYO_E_INVAL, "ids is null");
  BYO_REQUIRE(n < 0xFFFFFFFFull, BYO_E_INVAL, "n must be below 2^32-1 (uint32 vertex ids)");
  BYO_REQUIRE(offsets[0] == 0 && offsets[n] == m, BYO_E_INVAL, "offsets[0] != 0 or offsets[n] != m");
  for (uint64_t v = 0; v < n; v++) {
    BYO_REQUIRE(offsets[v] <= offsets[v + 1], BYO_E_INVAL, "offsets not monotone");
  }
  for (uint64_t e = 0; e < m; e++) BYO_REQUIRE(ids[e] < n, BYO_E_INVAL, "neighbour id out of range");
}

}  // namespace

// Builds a CSR in g->off / g->ids from unsorted keys (may contain the ~0
// sentinel and duplicates).  keys and alt hold nk entries each.
void csr_from_keys(byo_graph *g, uint64_t *keys, uint64_t *alt, uint64_t nk, uint64_t n) {
  hipStream_t s = g->stream;
  int idbits = 1;
  while ((1ull << idbits) < n) idbits++;
  int shifts[16];
  int np = prim::edge_key_shifts(idbits, shifts);
  auto &tmp = g->buf("sort_tmp", prim::radix_tmp_bytes(nk));
  uint64_t *sorted = prim::radix_sort_u64(keys, alt, nk, shifts, np, tmp.p, s);
  uint64_t *uniq = (sorted == keys) ? alt : keys;
  auto &stmp = g->buf("select_tmp", prim::select_tmp_bytes(nk) + 64);
  uint64_t *count_dev = stmp.as<uint64_t>();
  prim::unique_u64(sorted, uniq, nk, ~0ull, count_dev, stmp.as<char>() + 256, s);
  uint64_t *h = static_cast<uint64_t *>(pinned(g, 64));
  BYO_HIP(hipMemcpyAsync(h, count_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
  BYO_HIP(hipStreamSynchronize(s));
  uint64_t m = h[0];
  g->n = n;
  g->m = m;
  g->off.alloc((n + 1) * sizeof(uint64_t));
  g->ids.alloc(m * sizeof(uint32_t) + 16);
  prim::keys_to_csr(uniq, m, n, g->off.as<uint64_t>(), g->ids.as<uint32_t>(), s);
  BYO_HIP(hipStreamSynchronize(s));
  g->csr_valid = true;
  g->maxdeg_valid = false;
  g->pieces_stale();
  g->rl_valid = false;
  g->xs_valid = false;
  g->pb_valid = false;
}

void csr_upload(byo_graph *g, const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m) {
  g->n = n;
  g->m = m;
  g->off.alloc((n + 1) * sizeof(uint64_t));
  g->ids.alloc(m * sizeof(uint32_t) + 16);
  BYO_HIP(hipMemcpyAsync(g->off.p, offsets, (n + 1) * sizeof(uint64_t), hipMemcpyHostToDevice, g->stream));
  if (m)
    BYO_HIP(hipMemcpyAsync(g->ids.p, ids, m * sizeof(uint32_t), hipMemcpyHostToDevice, g->stream));
  BYO_HIP(hipStreamSynchronize(g->stream));
  g->csr_valid = true;
  g->maxdeg_valid = false;
  g->pieces_stale();
  g->rl_valid = false;
  g->xs_valid = false;
  g->pb_valid = false;
}

void check_csr_input(const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m) {
  check_host_csr(offsets, ids, n, m);
}

// Current CSR view of any container (the CPMA materialises one lazily).
void ensure_csr(byo_graph *g) {
  if (g->kind == 1 && !g->csr_valid) cpma_ensure_csr(g);
}

AdjView adj_view(byo_graph *g) {
  if (g->kind == 1) return cpma_view(g);
  return AdjView{g->off.as<uint64_t>(), g->ids.as<uint32_t>(), nullptr, g->m};
}

}  // namespace byo

using namespace byo;

extern "C" {

const char *byo_last_error(void) { return g_last_error.c_str(); }
const char *byo_version(void) { return "byo-mi355x 0.1 (gfx950)"; }

int byo_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

int byo_csr_create(const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m,
                   byo_graph_t *out) {
  return guard([&] {
    BYO_REQUIRE(out != nullptr, BYO_E_INVAL, "out is null");
    check_host_csr(offsets, ids, n, m);
    byo_graph *g = new_handle(0);
    try {
      csr_upload(g, offsets, ids, n, m);
    } catch (...) {
      byo_graph_destroy(g);
      throw;
    }
    *out = g;
  });
}

int byo_csr_create_rmat(int scale, uint64_t edge_factor, double a, double b, double c, uint64_t seed,
                        byo_graph_t *out) {
  return guard([&] {
    BYO_REQUIRE(out != nullptr, BYO_E_INVAL, "out is null");
    BYO_REQUIRE(scale >= 1 && scale <= 31, BYO_E_INVAL, "scale must be in [1, 31]");
    BYO_REQUIRE(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0, BYO_E_INVAL, "bad RMAT probabilities");
    uint64_t n = 1ull << scale;
    uint64_t ne = edge_factor * n;
    BYO_REQUIRE(2 * ne < 0xFFFFFFFFull, BYO_E_INVAL, "too many edges for one sort (2^32)");
    byo_graph *g = new_handle(0);
    try {
      DevBuf keys(2 * ne * sizeof(uint64_t) + 64), alt(2 * ne * sizeof(uint64_t) + 64);
      uint64_t base = 0;
      {
        // base = splitmix64(seed), computed identically on the host
        uint64_t z = seed + 0x9E3779B97F4A7C15ull;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        base = z ^ (z >> 31);
      }
      k_rmat_sym_keys<<<byo_grid(ne, 256, 32768), 256, 0, g->stream>>>(scale, ne, a, a + b, a + b + c,
                                                                         base, keys.as<uint64_t>());
      BYO_LAUNCH_CHECK();
      csr_from_keys(g, keys.as<uint64_t>(), alt.as<uint64_t>(), 2 * ne, n);
      g->scratch.erase("sort_tmp");
      g->scratch.erase("select_tmp");
    } catch (...) {
      byo_graph_destroy(g);
      throw;
    }
    *out = g;
  });
}

int byo_device_alloc(uint64_t bytes, void **dev_ptr) {
  return guard([&] {
    BYO_REQUIRE(dev_ptr, BYO_E_INVAL, "dev_ptr is null");
    require_device();
    *dev_ptr = nullptr;
    BYO_HIP(hipMalloc(dev_ptr, bytes ? bytes : 256));
  });
}

int byo_device_free(void *dev_ptr) {
  return guard([&] {
    if (dev_ptr) BYO_HIP(hipFree(dev_ptr));
  });
}

int byo_copy_to_device(void *dev_ptr, const void *host, uint64_t bytes) {
  return guard([&] {
    BYO_REQUIRE(dev_ptr && (host || bytes == 0), BYO_E_INVAL, "null argument");
    if (bytes) BYO_HIP(hipMemcpy(dev_ptr, host, bytes, hipMemcpyHostToDevice));
  });
}

int byo_update_rmat_keys(uint64_t graph_n, uint64_t first, uint64_t count, uint64_t *keys) {
  return guard([&] {
    BYO_REQUIRE(keys || count == 0, BYO_E_INVAL, "keys is null");
    BYO_REQUIRE(graph_n >= 2 && graph_n <= (1ull << 32), BYO_E_INVAL, "graph_n must be in [2, 2^32]");
    if (count == 0) return;
    require_device();
    int lg = 0;  // parlay::log2_up
    while ((1ull << lg) < graph_n) lg++;
    const uint32_t nn = (uint32_t)(1ull << (lg - 1));
    // batch_insert_helpers::hash64(0) truncated to the uint32 seed of rMat
    uint64_t v = 0ull * 3935559000370003845ull + 2691343689449507681ull;
    v ^= v >> 21;
    v ^= v << 37;
    v ^= v >> 4;
    v *= 4768777513237032717ull;
    v ^= v << 20;
    v ^= v >> 41;
    v ^= v << 5;
    uint32_t a = (uint32_t)v;  // hash32 on the host
    a = (a + 0x7ed55d16u) + (a << 12);
    a = (a ^ 0xc761c23cu) ^ (a >> 19);
    a = (a + 0x165667b1u) + (a << 5);
    a = (a + 0xd3a2646cu) ^ (a << 9);
    a = (a + 0xfd7046c5u) + (a << 3);
    a = (a ^ 0xb55a4f09u) ^ (a >> 16);
    DevBuf d(count * sizeof(uint64_t));
    k_update_rmat<<<byo_grid(count, 256, 16384), 256>>>(nn, a, first, count, d.as<uint64_t>());
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpy(keys, d.p, count * sizeof(uint64_t), hipMemcpyDeviceToHost));
  });
}

int byo_csr_create_er(int scale, uint64_t num_edges, uint64_t seed, byo_graph_t *out) {
  return guard([&] {
    BYO_REQUIRE(out != nullptr, BYO_E_INVAL, "out is null");
    BYO_REQUIRE(scale >= 1 && scale <= 31, BYO_E_INVAL, "scale must be in [1, 31]");
    BYO_REQUIRE(2 * num_edges < 0xFFFFFFFFull, BYO_E_INVAL, "too many edges for one sort (2^32)");
    const uint64_t n = 1ull << scale;
    byo_graph *g = new_handle(0);
    try {
      DevBuf keys(2 * num_edges * 8 + 64), alt(2 * num_edges * 8 + 64);
      uint64_t z = seed + 0x9E3779B97F4A7C15ull;  // splitmix64(seed) on the host
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      k_er_sym_keys<<<byo_grid(num_edges, 256, 32768), 256, 0, g->stream>>>(n, num_edges, z ^ (z >> 31),
                                                                             keys.as<uint64_t>());
      BYO_LAUNCH_CHECK();
      csr_from_keys(g, keys.as<uint64_t>(), alt.as<uint64_t>(), 2 * num_edges, n);
      g->scratch.erase("sort_tmp");
      g->scratch.erase("select_tmp");
    } catch (...) {
      byo_graph_destroy(g);
      throw;
    }
    *out = g;
  });
}

int byo_csr_create_grid3d(uint32_t side, byo_graph_t *out) {
  return guard([&] {
    BYO_REQUIRE(out != nullptr, BYO_E_INVAL, "out is null");
    BYO_REQUIRE(side >= 2 && side <= 1024, BYO_E_INVAL, "side must be in [2, 1024]");
    const uint64_t n = (uint64_t)side * side * side;
    BYO_REQUIRE(6 * n < 0xFFFFFFFFull, BYO_E_INVAL, "grid too large for one sort (2^32 keys)");
    byo_graph *g = new_handle(0);
    try {
      DevBuf keys(6 * n * 8 + 64), alt(6 * n * 8 + 64);
      k_grid3d_sym_keys<<<byo_grid(n, 256, 32768), 256, 0, g->stream>>>(side, keys.as<uint64_t>());
      BYO_LAUNCH_CHECK();
      csr_from_keys(g, keys.as<uint64_t>(), alt.as<uint64_t>(), 6 * n, n);
      g->scratch.erase("sort_tmp");
      g->scratch.erase("select_tmp");
    } catch (...) {
      byo_graph_destroy(g);
      throw;
    }
    *out = g;
  });
}

int byo_graph_destroy(byo_graph_t g) {
  if (!g) return BYO_OK;
  return guard([&] {
    (void)hipSetDevice(g->device);
    if (g->stream) (void)hipStreamSynchronize(g->stream);
    if (g->kind == 1) cpma_destroy(g);
    g->scratch.clear();
    g->off.release();
    g->ids.release();
    g->sel_params.release();
    g->sel_params_b.release();
    if (g->pinned) (void)hipHostFree(g->pinned);
    if (g->stream) (void)hipStreamDestroy(g->stream);
    delete g;
  });
}

int byo_graph_kind(byo_graph_t g, int *kind) {
  return guard([&] {
    BYO_REQUIRE(g && kind, BYO_E_INVAL, "null argument");
    *kind = g->kind;
  });
}

int byo_num_vertices(byo_graph_t g, uint64_t *n) {
  return guard([&] {
    BYO_REQUIRE(g && n, BYO_E_INVAL, "null argument");
    *n = g->n;
  });
}

int byo_num_edges(byo_graph_t g, uint64_t *m) {
  return guard([&] {
    BYO_REQUIRE(g && m, BYO_E_INVAL, "null argument");
    *m = g->m;
  });
}

int byo_degrees(byo_graph_t g, uint32_t *deg) {
  return guard([&] {
    BYO_REQUIRE(g && deg, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    if (g->n == 0) return;
    if (g->kind == 1) {
      // the CPMA keeps its degrees up to date (no CSR view needed)
      BYO_HIP(hipMemcpyAsync(deg, cpma_degrees(g), g->n * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
      BYO_HIP(hipStreamSynchronize(g->stream));
      return;
    }
    auto &d = g->buf("degrees_out", g->n * sizeof(uint32_t));
    k_degrees<<<byo_grid(g->n, 256, 16384), 256, 0, g->stream>>>(g->off.as<uint64_t>(), g->n,
                                                                  d.as<uint32_t>());
    BYO_LAUNCH_CHECK();
    BYO_HIP(hipMemcpyAsync(deg, d.p, g->n * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

int byo_degree(byo_graph_t g, uint32_t v, uint32_t *deg) {
  return guard([&] {
    BYO_REQUIRE(g && deg, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(v < g->n, BYO_E_INVAL, "vertex out of range");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    uint64_t *h = static_cast<uint64_t *>(pinned(g, 64));
    if (g->kind == 1) {
      BYO_HIP(hipMemcpyAsync(h, cpma_degrees(g) + v, sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
      BYO_HIP(hipStreamSynchronize(g->stream));
      *deg = *reinterpret_cast<uint32_t *>(h);
      return;
    }
    BYO_HIP(hipMemcpyAsync(h, g->off.as<uint64_t>() + v, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
    *deg = (uint32_t)(h[1] - h[0]);
  });
}

int byo_neighbors(byo_graph_t g, uint32_t v, uint32_t *out, uint64_t cap, uint64_t *count) {
  return guard([&] {
    BYO_REQUIRE(g && count, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(v < g->n, BYO_E_INVAL, "vertex out of range");
    BYO_REQUIRE(cap == 0 || out, BYO_E_INVAL, "out is null");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    uint64_t *h = static_cast<uint64_t *>(pinned(g, 64));
    BYO_HIP(hipMemcpyAsync(h, g->off.as<uint64_t>() + v, 2 * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
    uint64_t d = h[1] - h[0];
    *count = d;
    uint64_t k = d < cap ? d : cap;
    if (k)
      BYO_HIP(hipMemcpyAsync(out, g->ids.as<uint32_t>() + h[0], k * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

int byo_export_csr(byo_graph_t g, uint64_t *offsets, uint32_t *ids) {
  return guard([&] {
    BYO_REQUIRE(g && offsets, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    BYO_HIP(hipMemcpyAsync(offsets, g->off.p, (g->n + 1) * sizeof(uint64_t), hipMemcpyDeviceToHost,
                           g->stream));
    if (g->m) {
      BYO_REQUIRE(ids != nullptr, BYO_E_INVAL, "ids is null");
      BYO_HIP(hipMemcpyAsync(ids, g->ids.p, g->m * sizeof(uint32_t), hipMemcpyDeviceToHost, g->stream));
    }
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

int byo_memory_size(byo_graph_t g, uint64_t *bytes) {
  return guard([&] {
    BYO_REQUIRE(g && bytes, BYO_E_INVAL, "null argument");
    if (g->kind == 0)
      *bytes = (g->n + 1) * sizeof(uint64_t) + g->m * sizeof(uint32_t);
    else
      *bytes = cpma_bytes(g);
  });
}

int byo_map_sum(byo_graph_t g, uint64_t *sum, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && sum, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    // the container's own adjacency: the CSR ids, or every slot of the PMA
    const AdjView av = adj_view(g);
    const int nb = 2048;
    auto &part = g->buf("map_partial", (nb + 1) * sizeof(uint64_t));
    PhaseTimer t(g->stream, true);
    if (av.deg)
      k_map_sum<true><<<nb, 256, 0, g->stream>>>(av.ids, av.nslots, part.as<uint64_t>());
    else
      k_map_sum<false><<<nb, 256, 0, g->stream>>>(av.ids, g->m, part.as<uint64_t>());
    k_sum_partials<<<1, 1024, 0, g->stream>>>(part.as<uint64_t>(), nb, part.as<uint64_t>() + nb);
    BYO_LAUNCH_CHECK();
    double ms = t.stop();
    if (g->prof.on) g->account("map", ms);
    if (device_ms) *device_ms = ms;
    BYO_HIP(hipMemcpyAsync(sum, part.as<uint64_t>() + nb, sizeof(uint64_t), hipMemcpyDeviceToHost,
                           g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

int byo_map_remote(byo_graph_t g, const void *remote, int read_size, uint64_t *sum, double *device_ms) {
  return guard([&] {
    BYO_REQUIRE(g && remote && sum, BYO_E_INVAL, "null argument");
    BYO_REQUIRE(read_size == 0 || read_size == 1 || read_size == 2 || read_size == 4 || read_size == 8 ||
                    read_size == 16 || read_size == 32 || read_size == 64,
                BYO_E_INVAL, "read_size must be 0, 1, 2, 4, 8, 16, 32 or 64");
    std::lock_guard<std::mutex> lk(g->mu);
    bind(g);
    ensure_csr(g);
    const uint64_t bytes = read_size == 0 ? (g->n + 7) / 8 : g->n * (uint64_t)read_size;
    uint8_t *dr = g->buf("map_remote", bytes + 64).as<uint8_t>();
    BYO_HIP(hipMemcpyAsync(dr, remote, bytes, hipMemcpyHostToDevice, g->stream));
    const int nb = 4096;
    auto &part = g->buf("map_partial", (nb + 1) * sizeof(uint64_t));
    const uint32_t *ids = g->ids.as<uint32_t>();
    uint64_t *pp = part.as<uint64_t>();
    PhaseTimer t(g->stream, true);
    switch (read_size) {
      case 0: k_map_remote<0><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      case 1: k_map_remote<1><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      case 2: k_map_remote<2><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      case 4: k_map_remote<4><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      case 8: k_map_remote<8><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      case 16: k_map_remote<16><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      case 32: k_map_remote<32><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
      default: k_map_remote<64><<<nb, 256, 0, g->stream>>>(ids, g->m, dr, pp); break;
    }
    k_sum_partials<<<1, 1024, 0, g->stream>>>(pp, nb, pp + nb);
    BYO_LAUNCH_CHECK();
    const double ms = t.stop();
    if (g->prof.on) g->account("map_remote", ms);
    if (device_ms) *device_ms = ms;
    BYO_HIP(hipMemcpyAsync(sum, pp + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, g->stream));
    BYO_HIP(hipStreamSynchronize(g->stream));
  });
}

int byo_set_tuning(byo_graph_t g, const char *key, int64_t value) {
  return guard([&] {
    BYO_REQUIRE(g && key, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    for (const auto &t : tune_specs()) {
      if (std::strcmp(t.key, key) != 0) continue;
      BYO_REQUIRE(value >= t.lo && value <= t.hi, BYO_E_INVAL, std::string("tuning value out of range for ") + key);
      if (std::strcmp(key, "pr_lanes") == 0)
        BYO_REQUIRE(value == 0 || value == 4 || value == 8 || value == 16 || value == 32, BYO_E_INVAL,
                    "pr_lanes must be 0, 4, 8, 16 or 32");
      g->tune[key] = value;
      return;
    }
    throw Error(BYO_E_INVAL, std::string("unknown tuning knob ") + key);
  });
}

int byo_get_tuning(byo_graph_t g, const char *key, int64_t *value) {
  return guard([&] {
    BYO_REQUIRE(g && key && value, BYO_E_INVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    for (const auto &t : tune_specs())
      if (std::strcmp(t.key, key) == 0) {
        *value = g->tuning(key);
        return;
      }
    throw Error(BYO_E_INVAL, std::string("unknown tuning knob ") + key);
  });
}

int byo_profile_enable(byo_graph_t g, int on) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    g->prof.on = on != 0;
  });
}

int byo_profile_get(byo_graph_t g, const char *phase, uint64_t *launches, double *total_ms) {
  return guard([&] {
    BYO_REQUIRE(g && phase, BYO_E_INVAL, "null argument");
    auto it = g->prof.acc.find(phase);
    uint64_t l = 0;
    double ms = 0;
    if (it != g->prof.acc.end()) {
      l = it->second.launches;
      ms = it->second.ms;
    }
    if (launches) *launches = l;
    if (total_ms) *total_ms = ms;
  });
}

int byo_profile_reset(byo_graph_t g) {
  return guard([&] {
    BYO_REQUIRE(g, BYO_E_INVAL, "null handle");
    g->prof.acc.clear();
  });
}

}  // extern "C"
