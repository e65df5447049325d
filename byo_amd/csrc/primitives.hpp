// primitives.hpp -- device-wide building blocks used by the containers and
// algorithms: exclusive scan, stable LSD radix sort of uint64 keys, stable
// selection (unique / filter), CSR offsets from sorted keys.  All run on the
// caller's stream and take caller-provided scratch from the handle.
#pragma once
#include "common.hpp"

namespace byo {
namespace prim {

// out[i] = sum_{j<i} in[i]; out[n] = total when out has n+1 slots
// (write_total).  tmp: scratch of scan_tmp_bytes(n).
size_t scan_tmp_bytes(uint64_t n);
void exclusive_scan_u64(const uint64_t *in, uint64_t *out, uint64_t n,
                        bool write_total, void *tmp, hipStream_t s);
void exclusive_scan_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                        bool write_total, void *tmp, hipStream_t s);

// Stable LSD radix sort of keys[n] on 8-bit digits at the given shifts
// (ascending significance).  Returns the buffer holding the result (keys
// or alt).  tmp: scratch of radix_tmp_bytes(n).
size_t radix_tmp_bytes(uint64_t n);
uint64_t *radix_sort_u64(uint64_t *keys, uint64_t *alt, uint64_t n,
                         const int *shifts, int npasses, void *tmp,
                         hipStream_t s);
// Digit shifts covering bits [0, idbits) and [32, 32 + idbits) of
// (src<<32)|dst keys with ids below 2^idbits.
int edge_key_shifts(int idbits, int *shifts);

// Stable: out <- keys[i] where keys[i] != keys[i-1] and keys[i] != drop.
// The count is written to *count_dev.  tmp: select_tmp_bytes(n).
size_t select_tmp_bytes(uint64_t n);
void unique_u64(const uint64_t *keys, uint64_t *out, uint64_t n,
                uint64_t drop, uint64_t *count_dev, void *tmp, hipStream_t s);

// off[n+1] and ids[m] from sorted unique keys (src<<32)|dst.
void keys_to_csr(const uint64_t *keys, uint64_t m, uint64_t n, uint64_t *off,
                 uint32_t *ids, hipStream_t s);
// keys[e] = (src<<32)|ids[e] for a CSR.
void csr_to_keys(const uint64_t *off, const uint32_t *ids, uint64_t n,
                 uint64_t m, uint64_t *keys, hipStream_t s);

void fill_u32(uint32_t *p, uint32_t v, uint64_t n, hipStream_t s);
void fill_u64(uint64_t *p, uint64_t v, uint64_t n, hipStream_t s);

}  // namespace prim
}  // namespace byo
