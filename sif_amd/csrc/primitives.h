// Device-wide integer primitives (primitives.hip). All run on the library
// stream and are deterministic.
#pragma once
#include <cstdint>

namespace sif_amd {
// out[i] = sum of in[0..i); *total_dev (device pointer, optional) = sum of all.
int exclusive_scan_u32(const uint32_t* in, uint64_t n, uint64_t* out, uint64_t* total_dev);
int exclusive_scan_u64(const uint64_t* in, uint64_t n, uint64_t* out, uint64_t* total_dev);
// Stable ascending sort of (key, value) pairs on the low `bits` bits of key.
// differ: the bits on which the keys may differ; a 4-bit digit outside it is
// the same in every key and its pass is skipped.
int radix_sort_u32(uint32_t* keys, uint32_t* vals, uint32_t n, int bits,
  uint32_t differ = 0xffffffffu);
}  // namespace sif_amd
