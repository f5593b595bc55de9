// Device-wide primitives: stable LSD radix sort of (u64 key, u32 value) pairs and exclusive
// scans of u32 arrays, hand-written for gfx950 (wave64 multisplit ranking, LDS tiles).
#pragma once

#include "fh_common.hpp"

namespace fh {

// Stable sort of (keys, vals) by key bits [0, end_bit); bits at and above end_bit are ignored
// (they may carry payload that moves with the key). Sorted data is returned in keys/vals (the
// alternate buffers are scratch of the same sizes).
void radix_sort_pairs(uint64_t* keys, uint32_t* vals, uint64_t* keys_alt, uint32_t* vals_alt, size_t n,
                      int end_bit, hipStream_t st);

// The same on owned buffers: when the passes leave the result in the alternate buffers, the
// buffers are swapped instead of copied back (keys/vals then own the sorted data, the alternates
// the scratch). All four hold at least n elements.
void radix_sort_pairs(DBuf<uint64_t>& keys, DBuf<uint32_t>& vals, DBuf<uint64_t>& keys_alt, DBuf<uint32_t>& vals_alt,
                      size_t n, int end_bit, hipStream_t st);

// out[i] = sum_{j<i} in[i]; *total_dev (may be null) = sum of all. in and out may alias.
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, size_t n, uint32_t* total_dev, hipStream_t st);

// 1 when keys[i] < keys[i-1] for some i (not sorted), written to *flag_dev (zeroed first).
void check_sorted_u64(const uint64_t* keys, size_t n, uint32_t* flag_dev, hipStream_t st);

}  // namespace fh
