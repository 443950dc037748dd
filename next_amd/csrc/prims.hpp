// Host-side launchers for the sort / scan primitives (prims.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace nbx {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanBlock * kScanItems;

constexpr int kSortBins = 256;

// Radix sort: tiles of kSortTile keys (kSortBlock threads x kSortItems),
// one launch per 8-bit digit after one histogram launch for all digits.
constexpr int kSortBlock = 512;
constexpr int kSortItems = 8;
constexpr int kSortTile = kSortBlock * kSortItems;
constexpr int kSortMaxPasses = 8;
constexpr int kSortHistSlots = 16;   // the digit counts are spread over slots summed by the passes
constexpr int kSortGroup = 16;       // tiles per look-back group
// scratch words before the look-back words: digit counts per slot and pass,
// tile tickets per pass, spare
constexpr int kSortHeader = kSortHistSlots * kSortMaxPasses * kSortBins + 64;

constexpr int64_t div_up(int64_t a, int64_t b) { return (a + b - 1) / b; }

// Words of scratch the scan of n elements needs.
constexpr size_t scan_scratch_words(int64_t n) { return (size_t)div_up(n, kScanTile) + 2; }

// Look-back words of one pass: one per (tile, digit) and per (group, digit).
constexpr size_t sort_pass_words(int64_t n) {
  return (size_t)kSortBins * (size_t)(div_up(n, kSortTile) + div_up(div_up(n, kSortTile), kSortGroup));
}

// Words of scratch a radix sort of n elements needs.
constexpr size_t sort_scratch_words(int64_t n) { return (size_t)kSortHeader + (size_t)kSortMaxPasses * sort_pass_words(n); }

// Exclusive prefix sum of n uint32 values (in may equal out).  total (device
// pointer, may be null) receives the sum.
void scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint32_t* scratch,
              hipStream_t s);

// Stable LSD radix sort of (key, value) pairs by bits [lo, lo + bits) of
// the keys in 8-bit digits (the order of bits below lo is kept), n < 2^30.
// Ping-pongs between (k0, v0) and (k1, v1); returns true when the result
// ended in (k1, v1).  Every tile waits for the tiles before it in a bounded
// loop; if a wait runs out, bit kSortErrTimeout is set in *err (a device
// word, may be null) and the order is not valid.
constexpr int32_t kSortErrTimeout = 2;
constexpr int32_t kSortErrPatch = 4;
bool radix_sort_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, int64_t n, int lo, int bits,
                    uint32_t* scratch, int32_t* err, hipStream_t s);
bool radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int lo, int bits,
                    uint32_t* scratch, int32_t* err, hipStream_t s);

// Sort of keys that are nearly sorted already (the order of the last sort,
// a few keys moved): k (n keys) into kout, with perm[j] the position in k of
// kout[j].  The descents k[i] > k[i+1] are collected (*ndesc, a device word,
// must be zero on entry); both ends of every descent are taken out, and the
// rest is in order when, across each run of taken positions, the key before
// it is at most the key after it.  The taken keys (at most 2 kPatchMaxDesc)
// are sorted in one workgroup and merged with the rest by binary search:
// keys in ascending order, equal keys in no particular order.  With more
// than kPatchMaxDesc descents, or the rest out of order, bit kSortErrPatch
// is set in *err, kout is a copy of k and perm the identity: the caller sorts
// again with radix_sort_u64.  Scratch: patch_scratch_words() words.
constexpr int kPatchMaxDesc = 1024;
constexpr size_t patch_scratch_words() { return 16 + 9 * (size_t)kPatchMaxDesc; }
void patch_sort_u64(const uint64_t* k, uint64_t* kout, uint32_t* perm, int64_t n, uint32_t* ndesc,
                    uint32_t* scratch, int32_t* err, hipStream_t s);

}  // namespace nbx
