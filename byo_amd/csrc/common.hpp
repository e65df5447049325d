// common.hpp -- shared internals of libbyo: error plumbing, device buffers,
// the graph handle, and wave64 helpers for gfx950.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/byo.h"

namespace byo {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// Bound on the selector's normalised features, in standard deviations of
// the training set (the trainer applies the same clamp).
constexpr float kSelClamp = 3.0f;
// The learned selector overrides the threshold rule only for frontiers whose
// features all lie within a bound (tuning knob sel_range_x100, at most
// kSelClamp) of the training set's deviations and when
// its logit for another direction beats the rule's by at least a margin
// (tuning knobs sel_margin_x100 and sel_margin_batch_x100, per traversal
// engine; the trainer stores the value it validated with the weights).

// Selector class of a direction: 0 push (direction 0), 1 pull (1), 2 the
// engine's third choice -- dense forward (3) for the one-source traversal,
// bitmap pull (4) for the shared multi-source traversal.
__host__ __device__ inline int dir_class(int dir) { return (dir == 3 || dir == 4) ? 2 : (dir == 1 ? 1 : 0); }
constexpr int kWave = 64;

// ---------------------------------------------------------------- errors
struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string &m);

#define BYO_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess)                                                    \
      throw ::byo::Error(_e == hipErrorOutOfMemory ? BYO_E_OOM : BYO_E_HIP,  \
                         std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define BYO_REQUIRE(cond, code, msg)                    \
  do {                                                  \
    if (!(cond)) throw ::byo::Error((code), (msg));     \
  } while (0)

#define BYO_LAUNCH_CHECK() BYO_HIP(hipGetLastError())

// Runs f, mapping exceptions to a status code for the C ABI.
template <class F>
int guard(F &&f) {
  try {
    f();
    set_last_error("");
    return BYO_OK;
  } catch (const Error &e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc &) {
    set_last_error("host allocation failed");
    return BYO_E_OOM;
  } catch (const std::exception &e) {
    set_last_error(e.what());
    return BYO_E_HIP;
  }
}

// ---------------------------------------------------------- device memory
// Owning device allocation.
struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  explicit DevBuf(size_t b) { alloc(b); }
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) { o.p = nullptr; o.bytes = 0; }
  DevBuf &operator=(DevBuf &&o) noexcept {
    if (this != &o) { release(); p = o.p; bytes = o.bytes; o.p = nullptr; o.bytes = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t b) {
    release();
    if (b == 0) b = 256;
    BYO_HIP(hipMalloc(&p, b));
    bytes = b;
  }
  // Grows (discarding contents) when smaller than b.
  void ensure(size_t b) {
    if (bytes < b) alloc(b + b / 8);
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
  template <class T> T *as() const { return static_cast<T *>(p); }
};

// Event-timed phases (device time on the handle's stream).
struct Profiler {
  bool on = false;
  struct Acc { uint64_t launches = 0; double ms = 0; };
  std::map<std::string, Acc> acc;
};

// RAII timing of one phase with HIP events on stream s.
struct PhaseTimer {
  hipEvent_t a{}, b{};
  hipStream_t s;
  bool active;
  PhaseTimer(hipStream_t s_, bool active_) : s(s_), active(active_) {
    if (active) {
      BYO_HIP(hipEventCreate(&a));
      BYO_HIP(hipEventCreate(&b));
      BYO_HIP(hipEventRecord(a, s));
    }
  }
  // Records the end event and returns elapsed ms (synchronises on it).
  double stop() {
    if (!active) return 0.0;
    BYO_HIP(hipEventRecord(b, s));
    BYO_HIP(hipEventSynchronize(b));
    float ms = 0;
    BYO_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
  }
  ~PhaseTimer() {
    if (active) { (void)hipEventDestroy(a); (void)hipEventDestroy(b); }
  }
};

}  // namespace byo

// ------------------------------------------------------------ the handle
namespace byo {
// Tuning knobs of a handle: block and tile sizes the autotuner
// (byo_amd/autotune.py) picks per graph and sets through byo_set_tuning.
struct TuneSpec {
  const char *key;
  int64_t dflt, lo, hi;
  const char *what;
};
inline const std::vector<TuneSpec> &tune_specs() {
  static const std::vector<TuneSpec> specs = {
      {"coop_blocks_per_cu", 4, 1, 4, "blocks per CU of the persistent traversal launches"},
      {"ppr_blocks_per_cu", 4, 1, 8, "blocks per CU of the persistent personalized PageRank launch"},
      {"pieces", 8, 1, 64, "weight-balanced word pieces per wave in vertex-driven passes"},
      {"coop_grid_sqrt_div", 14, 0, 1 << 20,
       "traversal grid at most sqrt(adjacency slots) / this blocks, a multiple of 8 (0: the whole resident grid)"},
      {"pull_dyn", 8, 0, 64,
       "word pieces per wave dealt dynamically (the tail of the share-out) in the shared traversal's dense levels"},
      {"ppr_pieces", 4, 1, 16, "weight-balanced word pieces per wave in the personalized PageRank sweeps"},
      {"pr_lanes", 0, 0, 32, "PageRank lanes per vertex (4, 8, 16, 32; 0: per tile from its largest degree)"},
      {"pr_grid", 4096, 256, 65536, "PageRank pull grid (blocks of 256)"},
      {"pr_seg_log2", 40, 10, 40, "log2 of the PageRank source-segment width (40: one segment)"},
      {"kcore_blocks_per_cu", 3, 1, 8, "blocks per CU of the persistent k-core peel"},
      {"sel_range_x100", 300, 0, 1000, "learned selector in-range bound, training deviations x100"},
      {"sel_margin_x100", 100, 0, 100000, "learned selector margin, logits x100 (one-source engine)"},
      {"sel_margin_batch_x100", 100, 0, 100000, "learned selector margin, logits x100 (shared traversal)"},
      {"sel_densify", 0, 0, 1, "learned selector may turn a level the rule keeps sparse dense (0: never)"},
      {"pr_relabel", 0, 0, 1, "PageRank on a copy relabelled by descending degree (built once per graph)"},
      {"pr_xcd", 0, 0, 1, "PageRank pull split into one source slice per XCD (built once per graph)"},
      {"pr_pb", 0, 0, 1,
       "PageRank by propagation blocking: contributions scattered by source into destination bins, each bin "
       "summed in LDS (built once per graph)"},
      {"pr_xcd_vw", 0, 0, 1024,
       "weight of a vertex, in edges, when the per-XCD source slices are cut to equal weight (0: equal edges)"},
      {"bfs_engine", 0, 0, 2,
       "BFS of several sources: 0 one shared traversal per 64, 1 one persistent launch taking them in turn, "
       "2 host-driven level loop"},
  };
  return specs;
}
}  // namespace byo

struct byo_graph {
  int kind = 0;  // 0 CSR, 1 CPMA
  int device = 0;
  uint64_t n = 0, m = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;

  // CSR arrays (kind 0), or the CSR view exported from the CPMA (kind 1).
  byo::DevBuf off;  // uint64[n+1]
  byo::DevBuf ids;  // uint32[m]
  bool csr_valid = false;
  uint64_t maxdeg = 0;        // maximum degree of the current CSR view
  bool maxdeg_valid = false;  // cleared whenever the CSR view is rebuilt
  bool rl_valid = false;      // degree-relabelled copy (relabel_view) is current
  // word pieces cached per set (0: traversals, buffer "bfs_pieces"; 1:
  // personalized PageRank sweeps, "ppr_pieces"): the buffer holds the pieces
  // of this grid and pieces per wave
  struct PieceCache {
    bool valid = false;
    int nblocks = 0;
    uint32_t per_wave = 0;
  } pcache[3];
  // after any change of the adjacency: the word pieces and the
  // first-neighbour array are rebuilt by their next user
  void pieces_stale() {
    pcache[0].valid = pcache[1].valid = pcache[2].valid = false;
    fn_valid = false;
  }
  bool fn_valid = false;      // "bfs_first_ids" holds the first neighbours of view fn_view (0 CSR, 1 PMA)
  int fn_view = 0;
  bool xs_valid = false;      // PageRank per-XCD slices are current (for layout xs_layout)
  int xs_layout = 0;
  int64_t xs_vw = 0;          // pr_xcd_vw the slices were cut with
  // PageRank propagation-blocking layout (pr_pb): current for layout
  // pb_layout; pb_items gather items, pb_nsb scatter source blocks
  bool pb_valid = false;
  int pb_layout = 0;
  uint32_t pb_items = 0, pb_nsb = 0;
  int64_t last_src = -1;      // last source of byo_bfs / byo_bfs_multi
  const uint32_t *last_tree = nullptr;  // device parents of last_src's BFS (in a scratch buffer)
  bool coop_bar_ready = false;  // "bfs_coop_bar" cleared once: every launch leaves it clear

  // CPMA state (kind 1); owned by cpma.hip.
  void *cpma = nullptr;

  // Direction policy.
  int dir_policy = BYO_DIR_THRESHOLD;
  int64_t dir_threshold = -1;
  uint8_t dir_sched[64] = {};  // BYO_DIR_SCHEDULE: per-level directions (255 = rule)
  // Learned selector weights (device, 1379 floats + 12 normalisation).
  byo::DevBuf sel_params;
  std::vector<float> sel_host;  // host copy for the host-side decision
  bool sel_loaded = false;
  byo::DevBuf sel_params_b;  // weights for the shared multi-source traversal
  bool sel_loaded_b = false;

  // Scratch owned by algorithms (grown on demand, kept across calls).
  std::map<std::string, byo::DevBuf> scratch;
  byo::DevBuf &buf(const std::string &name, size_t bytes) {
    auto &b = scratch[name];
    b.ensure(bytes);
    return b;
  }
  // Pinned host staging for small readbacks.
  void *pinned = nullptr;
  size_t pinned_bytes = 0;

  std::map<std::string, int64_t> tune;  // knobs set through byo_set_tuning
  int64_t tuning(const char *key) const {
    auto it = tune.find(key);
    if (it != tune.end()) return it->second;
    for (const auto &t : byo::tune_specs())
      if (std::strcmp(t.key, key) == 0) return t.dflt;
    return 0;
  }

  byo::Profiler prof;
  void account(const char *phase, double ms) {
    auto &a = prof.acc[phase];
    a.launches++;
    a.ms += ms;
  }
};

namespace byo {
// Ensures a device exists and selects the handle's device on this thread.
void require_device();
void bind(byo_graph *g);
byo_graph *new_handle(int kind);
void *pinned(byo_graph *g, size_t bytes);
// Maximum degree of the current CSR view (cached until it is rebuilt).
uint64_t graph_max_degree(byo_graph *g);
// Device levels (n entries, kNone when unreached) of the last BFS tree left
// in "bfs_parents" (bfs.hip); valid until the next call.
const uint32_t *bfs_levels_dev(byo_graph *g);
// Copy of the current CSR view with vertices renumbered by descending
// degree (relabel.hip, used by the PageRank pull): perm[old] = new.  Built
// on the first call after each change of the graph; the CSR view must be
// current.
struct RelabelView {
  const uint64_t *off;
  const uint32_t *ids;
  const uint32_t *perm;  // null when the graph is used as it is
};
RelabelView relabel_view(byo_graph *g);
// LDD rounds (ldd.hip); leaves clusters and claim rounds in device buffers
// "ldd_cl" / "ldd_rd".  Caller holds the lock with the CSR view current.
double ldd_run(byo_graph *g, double beta, uint32_t *rounds_run);
// CSR construction (graph.hip).
void csr_from_keys(byo_graph *g, uint64_t *keys, uint64_t *alt, uint64_t nk, uint64_t n);
void csr_upload(byo_graph *g, const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m);
void check_csr_input(const uint64_t *offsets, const uint32_t *ids, uint64_t n, uint64_t m);
// Current CSR view of any container (the CPMA materialises one lazily).
void ensure_csr(byo_graph *g);
// Adjacency as the traversal kernels read it: v's neighbour list is
// ids[off[v] .. off[v+1]).  For a CSR the lists are dense and deg is null
// (degree = off[v+1] - off[v]).  For the CPMA it is the packed-memory array
// itself: off holds the per-vertex slot starts, ids the PMA's dst words,
// lists may hold empty slots (kNone) at leaf ends, and deg holds the true
// degrees.  nslots = off[n].
struct AdjView {
  const uint64_t *off;
  const uint32_t *ids;
  const uint32_t *deg;
  uint64_t nslots;
};
// The container's own adjacency (CPMA: slot index built on demand after an
// update); the caller holds the handle's lock.
AdjView adj_view(byo_graph *g);
AdjView cpma_view(byo_graph *g);
// The CPMA's maintained degrees (uint32[n], device).
const uint32_t *cpma_degrees(byo_graph *g);
// Internal hooks implemented in other translation units.
void cpma_destroy(byo_graph *g);
void cpma_ensure_csr(byo_graph *g);  // refresh g->off/g->ids from the PMA
uint64_t cpma_bytes(byo_graph *g);
}  // namespace byo

// ------------------------------------------------------ device helpers
namespace byo {
__device__ __forceinline__ int lane_id() { return __lane_id(); }
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
}
template <class T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
template <class T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}
template <class T>
__device__ __forceinline__ T wave_min(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    T w = __shfl_xor(v, o, 64);
    v = w < v ? w : v;
  }
  return v;
}
// Inclusive prefix sum across the wave.
template <class T>
__device__ __forceinline__ T wave_incl_scan(T v) {
  const int l = __lane_id();
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T w = __shfl_up(v, o, 64);
    if (l >= o) v += w;
  }
  return v;
}
// Orders this wave's LDS writes before its later LDS reads (and the
// reverse) for the compiler and the hardware; no cross-wave effect.
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
__device__ __forceinline__ uint64_t grid_threads() {
  return static_cast<uint64_t>(gridDim.x) * blockDim.x;
}
__device__ __forceinline__ uint64_t global_tid() {
  return static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
}
}  // namespace byo

// Launch helper: grid size capped for grid-stride kernels.
inline unsigned byo_grid(uint64_t work, unsigned block, unsigned cap = 8192) {
  uint64_t g = (work + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}
