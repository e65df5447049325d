// primitives.hip -- scan, radix sort, stable selection and CSR/key
// conversion kernels for gfx950 (wave64).
#include "primitives.hpp"

namespace byo {
namespace prim {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr uint64_t kScanTile = kScanBlock * kScanItems;  // 2048

// Block-wide exclusive scan of one value per thread; returns the exclusive
// prefix and writes the block total to *total.
template <class T>
__device__ T block_excl_scan(T v, T *lds_waves, T *total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int nwaves = blockDim.x >> 6;
  T incl = wave_incl_scan(v);
  if (lane == 63) lds_waves[wave] = incl;
  __syncthreads();
  if (threadIdx.x == 0) {
    T run = 0;
    for (int w = 0; w < nwaves; w++) {
      T t = lds_waves[w];
      lds_waves[w] = run;
      run += t;
    }
    lds_waves[nwaves] = run;
  }
  __syncthreads();
  T r = lds_waves[wave] + incl - v;
  *total = lds_waves[nwaves];
  __syncthreads();
  return r;
}

template <class T>
__global__ void __launch_bounds__(kScanBlock) k_scan_reduce(const T *in, uint64_t n, T *bsum) {
  __shared__ T lds[kScanBlock / 64 + 1];
  const uint64_t base = blockIdx.x * kScanTile;
  T acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    uint64_t i = base + k * kScanBlock + threadIdx.x;
    if (i < n) acc += in[i];
  }
  T tot;
  (void)block_excl_scan<T>(acc, lds, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Single block: exclusive scan of bsum[nb] in place.
template <class T>
__global__ void __launch_bounds__(1024) k_scan_blocksums(T *bsum, uint64_t nb) {
  __shared__ T lds[1024 / 64 + 1];
  T carry = 0;
  for (uint64_t b0 = 0; b0 < nb; b0 += 1024) {
    uint64_t i = b0 + threadIdx.x;
    T v = i < nb ? bsum[i] : 0;
    T tot;
    T ex = block_excl_scan<T>(v, lds, &tot);
    if (i < nb) bsum[i] = carry + ex;
    carry += tot;
  }
}

template <class T>
__global__ void __launch_bounds__(kScanBlock) k_scan_final(const T *in, T *out, uint64_t n,
                                                           const T *bsum, int write_total) {
  __shared__ T lds[kScanBlock / 64 + 1];
  const uint64_t base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  T v[kScanItems];
  T acc = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    uint64_t i = base + k;
    v[k] = i < n ? in[i] : 0;
    acc += v[k];
  }
  T tot;
  T ex = block_excl_scan<T>(acc, lds, &tot) + bsum[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; k++) {
    uint64_t i = base + k;
    if (i < n) out[i] = ex;
    ex += v[k];
  }
  if (write_total && blockIdx.x == gridDim.x - 1 && threadIdx.x == blockDim.x - 1) out[n] = ex;
}

// One workgroup scans the whole array in chunks of 1024 x kScanItems
// (thread-contiguous items, read before any is written, so in == out is
// safe).  For arrays up to kSmallScan one launch instead of three: a launch
// costs ~10 us of dispatch here, more than the scan of such an array.
constexpr uint64_t kSmallScan = 1ull << 16;

template <class T>
__global__ void __launch_bounds__(1024) k_scan_small(const T *in, T *out, uint64_t n, int write_total) {
  __shared__ T lds[1024 / 64 + 1];
  T carry = 0;
  for (uint64_t b0 = 0; b0 < n; b0 += 1024 * kScanItems) {
    const uint64_t base = b0 + threadIdx.x * kScanItems;
    T v[kScanItems];
    T acc = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      const uint64_t i = base + k;
      v[k] = i < n ? in[i] : 0;
      acc += v[k];
    }
    T tot;
    T ex = block_excl_scan<T>(acc, lds, &tot) + carry;
#pragma unroll
    for (int k = 0; k < kScanItems; k++) {
      const uint64_t i = base + k;
      if (i < n) out[i] = ex;
      ex += v[k];
    }
    carry += tot;
  }
  if (write_total && threadIdx.x == 0) out[n] = carry;
}

template <class T>
void scan_impl(const T *in, T *out, uint64_t n, bool write_total, void *tmp, hipStream_t s) {
  if (n == 0) {
    if (write_total) BYO_HIP(hipMemsetAsync(out, 0, sizeof(T), s));
    return;
  }
  if (n <= kSmallScan) {
    k_scan_small<T><<<1, 1024, 0, s>>>(in, out, n, write_total ? 1 : 0);
    BYO_LAUNCH_CHECK();
    return;
  }
  uint64_t nb = (n + kScanTile - 1) / kScanTile;
  T *bsum = static_cast<T *>(tmp);
  // The reduce pass reads `in` with a tile-major layout; the final pass
  // scans thread-contiguous items -- both cover the same tile [b*2048, +2048).
  k_scan_reduce<T><<<nb, kScanBlock, 0, s>>>(in, n, bsum);
  k_scan_blocksums<T><<<1, 1024, 0, s>>>(bsum, nb);
  k_scan_final<T><<<nb, kScanBlock, 0, s>>>(in, out, n, bsum, write_total ? 1 : 0);
  BYO_LAUNCH_CHECK();
}

// ------------------------------------------------------------ radix sort
constexpr int kSortBlock = 256;
constexpr int kSortPerThread = 16;
constexpr uint64_t kSortTile = kSortBlock * kSortPerThread;  // 4096
constexpr int kWavesPerSortBlock = kSortBlock / 64;

__global__ void __launch_bounds__(kSortBlock) k_radix_hist(const uint64_t *keys, uint64_t n, int shift,
                                                            uint32_t *counts, uint32_t ntiles) {
  __shared__ uint32_t h[256];
  h[threadIdx.x] = 0;
  __syncthreads();
  const uint64_t base = blockIdx.x * kSortTile;
#pragma unroll 4
  for (int k = 0; k < kSortPerThread; k++) {
    uint64_t i = base + k * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[(keys[i] >> shift) & 0xFF], 1u);
  }
  __syncthreads();
  counts[(uint64_t)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// Stable scatter: wave w of the block owns keys [tile + w*1024, +1024),
// visited in 16 rounds of 64 consecutive keys, so ranks follow input order.
// The tile is first ordered by digit in LDS (its 4096 keys, stable), then
// written out by consecutive threads, so the keys of one digit leave as runs
// of consecutive addresses (about 16 keys, one 128-byte line, per digit and
// tile) instead of one scattered 8-byte store per key.
__global__ void __launch_bounds__(kSortBlock) k_radix_scatter(const uint64_t *keys, uint64_t *out,
                                                               uint64_t n, int shift,
                                                               const uint32_t *base_tbl,
                                                               uint32_t ntiles) {
  __shared__ uint32_t cnt[kWavesPerSortBlock][256];
  __shared__ uint32_t dstart[256];  // first position of each digit in the tile's sorted order
  __shared__ uint32_t scan_w[kWavesPerSortBlock + 1];
  __shared__ uint64_t sorted[kSortTile];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int d = lane; d < 256; d += 64) cnt[wave][d] = 0;
  __syncthreads();
  const uint64_t tile0 = blockIdx.x * kSortTile;
  const uint64_t wbase = tile0 + (uint64_t)wave * 64 * kSortPerThread;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t k_reg[kSortPerThread];
  uint32_t r_reg[kSortPerThread];
#pragma unroll
  for (int r = 0; r < kSortPerThread; r++) {
    uint64_t i = wbase + (uint64_t)r * 64 + lane;
    bool valid = i < n;
    uint64_t key = valid ? keys[i] : 0;
    uint32_t d = (uint32_t)(key >> shift) & 0xFF;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; b++) {
      bool bit = (d >> b) & 1;
      uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    uint32_t pre = valid ? cnt[wave][d] : 0;
    uint32_t rank = pre + __popcll(peers & lt_mask);
    // the lowest peer publishes the new count for this digit
    if (valid && (peers & lt_mask) == 0) cnt[wave][d] = pre + __popcll(peers);
    k_reg[r] = key;
    r_reg[r] = valid ? rank : 0xFFFFFFFFu;
  }
  __syncthreads();
  // per digit (thread = digit): wave offsets within the digit, then the
  // digit's start in the tile's sorted order
  uint32_t run = 0;
  for (int w = 0; w < kWavesPerSortBlock; w++) {
    uint32_t t = cnt[w][threadIdx.x];
    cnt[w][threadIdx.x] = run;
    run += t;
  }
  uint32_t tot;
  dstart[threadIdx.x] = block_excl_scan<uint32_t>(run, scan_w, &tot);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kSortPerThread; r++) {
    if (r_reg[r] != 0xFFFFFFFFu) {
      uint32_t d = (uint32_t)(k_reg[r] >> shift) & 0xFF;
      sorted[dstart[d] + cnt[wave][d] + r_reg[r]] = k_reg[r];
    }
  }
  __syncthreads();
  // position i of the sorted tile (digit d) goes to base_tbl[d] + (i - dstart[d])
  for (uint32_t i = threadIdx.x; i < tot; i += kSortBlock) {
    const uint64_t key = sorted[i];
    const uint32_t d = (uint32_t)(key >> shift) & 0xFF;
    out[(uint64_t)base_tbl[(uint64_t)d * ntiles + blockIdx.x] + (i - dstart[d])] = key;
  }
}

// ------------------------------------------------------------- selection
constexpr int kSelBlock = 256;
constexpr int kSelItems = 8;
constexpr uint64_t kSelTile = kSelBlock * kSelItems;

__device__ __forceinline__ bool keep_unique(const uint64_t *k, uint64_t i, uint64_t drop) {
  uint64_t v = k[i];
  return v != drop && (i == 0 || k[i - 1] != v);
}

__global__ void __launch_bounds__(kSelBlock) k_select_count(const uint64_t *k, uint64_t n, uint64_t drop,
                                                             uint64_t *bcount) {
  __shared__ uint64_t lds[kSelBlock / 64 + 1];
  const uint64_t base = blockIdx.x * kSelTile;
  uint64_t c = 0;
  for (int r = 0; r < kSelItems; r++) {
    uint64_t i = base + r * kSelBlock + threadIdx.x;
    if (i < n && keep_unique(k, i, drop)) c++;
  }
  uint64_t tot;
  (void)block_excl_scan<uint64_t>(c, lds, &tot);
  if (threadIdx.x == 0) bcount[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(kSelBlock) k_select_write(const uint64_t *k, uint64_t *out, uint64_t n,
                                                             uint64_t drop, const uint64_t *bbase,
                                                             uint64_t *count_dev) {
  __shared__ uint32_t wcnt[kSelBlock / 64];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t base = blockIdx.x * kSelTile;
  uint64_t run = bbase[blockIdx.x];
  for (int r = 0; r < kSelItems; r++) {
    uint64_t i = base + r * kSelBlock + threadIdx.x;
    bool keep = i < n && keep_unique(k, i, drop);
    uint64_t mask = __ballot(keep);
    if (lane == 0) wcnt[wave] = __popcll(mask);
    __syncthreads();
    uint32_t before = 0, total = 0;
    for (int w = 0; w < kSelBlock / 64; w++) {
      if (w < wave) before += wcnt[w];
      total += wcnt[w];
    }
    if (keep) out[run + before + mbcnt(mask)] = k[i];
    run += total;
    __syncthreads();
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) *count_dev = run;
}

__global__ void k_keys_to_csr(const uint64_t *keys, uint64_t m, uint64_t n, uint64_t *off,
                              uint32_t *ids) {
  for (uint64_t i = global_tid(); i < m; i += grid_threads()) {
    uint64_t key = keys[i];
    uint64_t s = key >> 32;
    ids[i] = (uint32_t)key;
    uint64_t lo = (i == 0) ? 0 : (keys[i - 1] >> 32) + 1;
    for (uint64_t v = lo; v <= s; v++) off[v] = i;
    if (i == m - 1)
      for (uint64_t v = s + 1; v <= n; v++) off[v] = m;
  }
}

__global__ void k_csr_to_keys(const uint64_t *off, const uint32_t *ids, uint64_t n, uint64_t *keys) {
  // one wave per vertex, grid-stride
  const uint64_t wid = global_tid() >> 6, nw = grid_threads() >> 6;
  const int lane = lane_id();
  for (uint64_t v = wid; v < n; v += nw) {
    for (uint64_t e = off[v] + lane; e < off[v + 1]; e += 64) keys[e] = (v << 32) | ids[e];
  }
}

template <class T>
__global__ void k_fill(T *p, T v, uint64_t n) {
  for (uint64_t i = global_tid(); i < n; i += grid_threads()) p[i] = v;
}

}  // namespace

size_t scan_tmp_bytes(uint64_t n) { return ((n + kScanTile - 1) / kScanTile + 1) * sizeof(uint64_t); }

void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n, bool write_total, void *tmp,
                        hipStream_t s) {
  scan_impl<uint64_t>(in, out, n, write_total, tmp, s);
}
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n, bool write_total, void *tmp,
                        hipStream_t s) {
  scan_impl<uint32_t>(in, out, n, write_total, tmp, s);
}

size_t radix_tmp_bytes(uint64_t n) {
  uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
  uint64_t cnt = 256 * ntiles;
  // counts + scanned bases + scan scratch
  return 2 * cnt * sizeof(uint32_t) + scan_tmp_bytes(cnt) + 256;
}

uint64_t *radix_sort_u64(uint64_t *keys, uint64_t *alt, uint64_t n, const int *shifts, int npasses,
                         void *tmp, hipStream_t s) {
  if (n <= 1 || npasses == 0) return keys;
  uint64_t ntiles = (n + kSortTile - 1) / kSortTile;
  BYO_REQUIRE(ntiles * 256 < 0xFFFFFFFFull && n < 0xFFFFFFFFull, BYO_E_INVAL,
              "radix sort: more than 2^32 keys");
  uint64_t cnt = 256 * ntiles;
  uint32_t *counts = static_cast<uint32_t *>(tmp);
  uint32_t *bases = counts + cnt;
  void *stmp = reinterpret_cast<void *>(
      (reinterpret_cast<uintptr_t>(bases + cnt) + 255) & ~uintptr_t(255));
  uint64_t *from = keys, *to = alt;
  for (int p = 0; p < npasses; p++) {
    k_radix_hist<<<ntiles, kSortBlock, 0, s>>>(from, n, shifts[p], counts, (uint32_t)ntiles);
    exclusive_scan_u32(counts, bases, cnt, false, stmp, s);
    k_radix_scatter<<<ntiles, kSortBlock, 0, s>>>(from, to, n, shifts[p], bases, (uint32_t)ntiles);
    BYO_LAUNCH_CHECK();
    uint64_t *t = from;
    from = to;
    to = t;
  }
  return from;
}

int edge_key_shifts(int idbits, int *shifts) {
  int np = 0;
  for (int b = 0; b < idbits; b += 8) shifts[np++] = b;
  for (int b = 0; b < idbits; b += 8) shifts[np++] = 32 + b;
  return np;
}

size_t select_tmp_bytes(uint64_t n) {
  uint64_t nb = (n + kSelTile - 1) / kSelTile + 1;
  return 2 * nb * sizeof(uint64_t) + scan_tmp_bytes(nb) + 256;
}

void unique_u64(const uint64_t *keys, uint64_t *out, uint64_t n, uint64_t drop, uint64_t *count_dev,
                void *tmp, hipStream_t s) {
  if (n == 0) {
    BYO_HIP(hipMemsetAsync(count_dev, 0, sizeof(uint64_t), s));
    return;
  }
  uint64_t nb = (n + kSelTile - 1) / kSelTile;
  uint64_t *bcount = static_cast<uint64_t *>(tmp);
  uint64_t *bbase = bcount + nb + 1;
  void *stmp = reinterpret_cast<void *>(
      (reinterpret_cast<uintptr_t>(bbase + nb + 1) + 255) & ~uintptr_t(255));
  k_select_count<<<nb, kSelBlock, 0, s>>>(keys, n, drop, bcount);
  exclusive_scan_u64(bcount, bbase, nb, false, stmp, s);
  k_select_write<<<nb, kSelBlock, 0, s>>>(keys, out, n, drop, bbase, count_dev);
  BYO_LAUNCH_CHECK();
}

void keys_to_csr(const uint64_t *keys, uint64_t m, uint64_t n, uint64_t *off, uint32_t *ids,
                 hipStream_t s) {
  if (m == 0) {
    fill_u64(off, 0, n + 1, s);
    return;
  }
  k_keys_to_csr<<<byo_grid(m, 256, 16384), 256, 0, s>>>(keys, m, n, off, ids);
  BYO_LAUNCH_CHECK();
}

void csr_to_keys(const uint64_t *off, const uint32_t *ids, uint64_t n, uint64_t m, uint64_t *keys,
                 hipStream_t s) {
  if (m == 0 || n == 0) return;
  k_csr_to_keys<<<byo_grid(n * 64, 256, 16384), 256, 0, s>>>(off, ids, n, keys);
  BYO_LAUNCH_CHECK();
}

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_fill<uint32_t><<<byo_grid(n, 256, 16384), 256, 0, s>>>(p, v, n);
  BYO_LAUNCH_CHECK();
}
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s) {
  if (!n) return;
  k_fill<uint64_t><<<byo_grid(n, 256, 16384), 256, 0, s>>>(p, v, n);
  BYO_LAUNCH_CHECK();
}

}  // namespace prim
}  // namespace byo
