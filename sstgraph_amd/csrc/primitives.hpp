// Device-wide primitives used by the update pipeline and the traversals:
// exclusive scans, stream compaction and an LSD radix sort of 64-bit keys.
#pragma once

#include "common.hpp"

namespace sstg {

// out[i] = sum_{j<i} in[i]; in == out allowed.  If d_total is not null the
// grand total is written there (device memory).
int scan_exclusive_u64(const uint64_t *in, uint64_t *out, uint64_t n,
                       uint64_t *d_total, Workspace &ws, int ws_base,
                       hipStream_t s);
int scan_exclusive_u32(const uint32_t *in, uint32_t *out, uint64_t n,
                       uint32_t *d_total, Workspace &ws, int ws_base,
                       hipStream_t s);

// Stream compaction driven by a byte flag per element.  Writes the kept
// elements (values[i], or i itself when values is null) in order to out and
// their number to *d_count (device).
int select_flagged_u64(const uint8_t *flags, const uint64_t *values,
                       uint64_t n, uint64_t *out, uint64_t *d_count,
                       Workspace &ws, int ws_base, hipStream_t s);
// select_flagged_u64 that also compacts values2 into out2 in the same order.
int select_flagged_u64_u32(const uint8_t *flags, const uint64_t *values,
                           const uint32_t *values2, uint64_t n, uint64_t *out,
                           uint32_t *out2, uint64_t *d_count, Workspace &ws,
                           int ws_base, hipStream_t s);
// The same for 32-bit values (values null: the indices, which must fit).
int select_flagged_u32(const uint8_t *flags, const uint32_t *values,
                       uint64_t n, uint32_t *out, uint64_t *d_count,
                       Workspace &ws, int ws_base, hipStream_t s);

// Stable LSD radix sort of n keys on 8-bit digits starting at the given bit
// shifts (applied in order).  keys and tmp alternate as source and
// destination; *result points at the buffer holding the sorted keys.
// Requires n < 2^32.
int radix_sort_u64(uint64_t *keys, uint64_t *tmp, uint64_t n,
                   const int *shifts, int npass, Workspace &ws, int ws_base,
                   hipStream_t s, uint64_t **result);

// Full sort of n 64-bit keys whose set bits all lie in the given 8-bit
// digits: up to 2^20 keys, whole-key sorts of 4K-key chunks (one block each)
// and merge passes (one launch each); beyond that
// radix_sort_u64 over the digits.  Same contract for keys, tmp and *result.
int sort_keys_u64(uint64_t *keys, uint64_t *tmp, uint64_t n, const int *shifts,
                  int npass, Workspace &ws, int ws_base, hipStream_t s,
                  uint64_t **result);

// keys[i] = src[i] << 32 | dst[i], sorted as sort_keys_u64 sorts them.  A
// pair outside [0, n_rows) x [0, n_cols) sets *err to 1 and packs as (0, 0).
// On the radix path the packing kernel also counts the first digit of its
// tile, so the first pass reads no keys for its histogram.
int pack_sort_pairs(const uint32_t *src, const uint32_t *dst, uint64_t n,
                    uint32_t n_rows, uint32_t n_cols, uint64_t *err,
                    uint64_t *keys, uint64_t *tmp, const int *shifts,
                    int npass, Workspace &ws, int ws_base, hipStream_t s,
                    uint64_t **result);

// Workspace slots used by the primitives, starting at ws_base.
constexpr int kPrimSlots = 8;

}  // namespace sstg
