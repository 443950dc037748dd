// Host-side launchers for the sort / scan primitives (prims.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nbx {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kSortBins = 256;

inline int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Words of scratch the scan of n elements needs.
inline size_t scan_scratch_words(int64_t n) { return (size_t)div_up(n, kScanTile) + 2; }

// Exclusive prefix sum of n uint32 values (in may equal out).  total (device
// pointer, may be null) receives the sum.
void scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint32_t* scratch,
              hipStream_t s);

// Words of scratch a radix sort of n elements needs.
inline size_t sort_scratch_words(int64_t n) {
  int64_t tiles = div_up(n, kSortTile);
  return (size_t)(kSortBins * tiles) + kSortBins;
}

// Stable LSD radix sort of (key, value) pairs over bits [0, bits) in 8-bit
// digits.  Ping-pongs between (k0, v0) and (k1, v1); returns true when the
// result ended in (k1, v1).
bool radix_sort_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, int64_t n, int bits,
                    uint32_t* scratch, hipStream_t s);
bool radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int bits,
                    uint32_t* scratch, hipStream_t s);

}  // namespace nbx
