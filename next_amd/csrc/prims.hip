// Device-wide exclusive scan and stable LSD radix sort for gfx950.
//
// Scan: reduce-then-scan in two launches (tile sums; tile-local scan plus the
// sum of the preceding tile sums, which every tile forms itself).  No
// inter-workgroup waiting.
//
// Radix sort: 8-bit digits, one launch per digit after one launch that
// counts every digit of every key (the counts of each digit position are all
// a pass needs to place its digit groups).  A pass takes tiles in the order of
// a ticket counter, ranks its keys stably inside the tile with wave64 ballots
// (8 ballots give each lane the set of lanes holding the same digit), and
// learns how many keys of each digit the tiles before it hold by a two-level
// decoupled look-back: each tile publishes its own digit counts at once; the
// last tile of each group of kSortGroup tiles adds up its group and publishes
// the group's counts, then the running totals up to its group's end when it
// has them, one word per (tile or group, digit) with the kind in the top two
// bits.  A tile adds the words of the tiles before it in its group and walks
// back over the groups before its own until it meets a running total, so it
// reads a few dozen words per digit however many tiles run at once.  A tile
// only waits for tiles holding earlier tickets, which are resident and wait on
// nobody later, so the waits end; they are still bounded, and a wait that runs
// out sets an error bit instead of hanging.  The tile is staged in LDS in
// digit order and written out in runs, so consecutive lanes store consecutive
// addresses.
//
// Nearly sorted keys (patch_sort_u64): keys that were sorted one step ago
// and of which a few moved.  The descents are collected, both ends of each
// are taken out, the rest is checked to be in order, the taken keys are
// sorted in one workgroup, and every key is placed by a merge: three
// launches, and a refusal (too many descents, or the rest out of order)
// that the caller answers with the radix sort.
#include <algorithm>

#include "common.hpp"
#include "prims.hpp"

namespace nbx {
namespace {

__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const uint32_t* __restrict__ in, int64_t n,
                                                            uint32_t* __restrict__ bsum) {
  __shared__ uint32_t scr[kScanBlock / kWave + 1];
  const int64_t base = (int64_t)blockIdx.x * kScanTile;
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    int64_t i = base + j * kScanBlock + threadIdx.x;
    if (i < n) s += in[i];
  }
  uint32_t tot;
  block_excl_scan_u32<kScanBlock>(s, scr, &tot);
  if (threadIdx.x == 0) bsum[blockIdx.x] = tot;
}

// Sum of bsum[0, k) by one workgroup (every thread gets it).
__device__ uint32_t block_prefix(const uint32_t* __restrict__ bsum, int64_t k, uint32_t* scr) {
  uint32_t s = 0;
  for (int64_t i = threadIdx.x; i < k; i += kScanBlock) s += bsum[i];
  uint32_t tot;
  (void)block_excl_scan_u32<kScanBlock>(s, scr, &tot);
  return tot;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_down(const uint32_t* in, uint32_t* out, int64_t n,
                                                          const uint32_t* __restrict__ bsum, int64_t nb,
                                                          uint32_t* __restrict__ total) {
  __shared__ uint32_t scr[kScanBlock / kWave + 1];
  const uint32_t carry = block_prefix(bsum, blockIdx.x, scr);
  if (total && blockIdx.x == nb - 1 && threadIdx.x == 0) *total = carry + bsum[blockIdx.x];
  const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    v[j] = (base + j < n) ? in[base + j] : 0u;
    s += v[j];
  }
  uint32_t tot;
  uint32_t run = carry + block_excl_scan_u32<kScanBlock>(s, scr, &tot);
#pragma unroll
  for (int j = 0; j < kScanItems; ++j) {
    if (base + j < n) out[base + j] = run;
    run += v[j];
  }
}

constexpr uint32_t kLbAggregate = 1u << 30;   // the tile's (group's) own count
constexpr uint32_t kLbPrefix = 2u << 30;      // count of this group and every earlier one
constexpr uint32_t kLbKind = 3u << 30;
constexpr uint32_t kLbCount = (1u << 30) - 1;
constexpr uint32_t kLbSpinLimit = 1u << 20;   // sleeps per lane per look-back, far above any real wait
constexpr int kLbWindow = 16;                 // group words loaded per round of the group look-back
constexpr int kHistRounds = 8;                // keys per thread per round of the counting launch
constexpr int kHistStride = kSortMaxPasses * kSortBins;   // words between two count slots

__global__ __launch_bounds__(kScanBlock) void k_sort_clear(uint32_t* __restrict__ w, int64_t nw) {
  for (int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x; i < nw; i += (int64_t)gridDim.x * kScanBlock)
    w[i] = 0u;
}

// Counts of every digit position (hist[p * 256 + d]), kHistRounds keys per
// thread per round of loads.  Keys arrive nearly sorted (the bodies' last order), so a
// wave's high digits are mostly all alike: such a digit is added once for
// the wave instead of once per lane.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_sort_hist(const K* __restrict__ keys, int64_t n, int lo, int passes,
                                                          uint32_t* __restrict__ hist) {
  __shared__ uint32_t h[kSortMaxPasses * kSortBins];
  for (int i = threadIdx.x; i < kSortMaxPasses * kSortBins; i += kSortBlock) h[i] = 0;
  __syncthreads();
  const int l = threadIdx.x % kWave;
  const int64_t span = (int64_t)kSortBlock * kHistRounds;
  for (int64_t base = (int64_t)blockIdx.x * span; base < n; base += (int64_t)gridDim.x * span) {
    K kr[kHistRounds];                        // all loads in flight before the counting
#pragma unroll
    for (int r = 0; r < kHistRounds; ++r) {
      const int64_t i = base + (int64_t)r * kSortBlock + threadIdx.x;
      kr[r] = i < n ? keys[i] : K(0);
    }
#pragma unroll
    for (int r = 0; r < kHistRounds; ++r) {
      const bool valid = base + (int64_t)r * kSortBlock + threadIdx.x < n;   // invalid lanes: a wave's last ones
      const uint64_t act = __ballot(valid);
      for (int p = 0; p < passes; ++p) {
        const uint32_t d = (uint32_t)(kr[r] >> (lo + 8 * p)) & 255u;
        const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
        if (__ballot(valid && d == d0) == act) {
          if (l == 0 && act) atomicAdd(&h[p * kSortBins + d0], (uint32_t)__popcll(act));
        } else if (valid) {
          atomicAdd(&h[p * kSortBins + d], 1u);
        }
      }
    }
  }
  __syncthreads();
  uint32_t* slot = hist + (blockIdx.x % kSortHistSlots) * kHistStride;
  for (int i = threadIdx.x; i < passes * kSortBins; i += kSortBlock)
    if (h[i]) atomicAdd(&slot[i], h[i]);
}

// One digit pass: the tile with ticket `tile` moves keys [tile * kSortTile,
// +kSortTile) to their places.  hist: this digit's count slots over all keys
// (slot k at hist + k * kHistStride); ticket: this pass's tile counter;
// tw, gw: this pass's words per (tile, digit) and (group, digit); ticket, tw
// and gw zero before the launch.  Grid: exactly ntiles = div_up(n, kSortTile)
// workgroups.
template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_sort_pass(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                          K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                          int64_t n, int64_t ntiles, int shift,
                                                          const uint32_t* __restrict__ hist,
                                                          uint32_t* __restrict__ ticket, uint32_t* __restrict__ tw,
                                                          uint32_t* __restrict__ gw, int32_t* __restrict__ err) {
  constexpr int W = kSortBlock / kWave;
  __shared__ K sk[kSortTile];
  __shared__ uint32_t sv[kSortTile];
  __shared__ uint32_t wc[W][kSortBins];
  __shared__ uint32_t dstart[kSortBins];
  __shared__ uint32_t gbase[kSortBins];
  __shared__ uint32_t scr[W + 1];
  __shared__ uint32_t hc[kSortBins];
  __shared__ uint32_t tile_s;
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  uint32_t* wcf = &wc[0][0];
  for (int i = t; i < W * kSortBins; i += kSortBlock) wcf[i] = 0;
  if (t < kSortBins) hc[t] = 0;
  if (t == 0) tile_s = atomicAdd(ticket, 1u);
  __syncthreads();
  const int64_t tile = tile_s;
  const int64_t g = tile / kSortGroup, g0 = g * kSortGroup;
  const bool last = tile - g0 == kSortGroup - 1 || tile == ntiles - 1;   // publishes its group's count

  // Wave w ranks elements [tile + w*64*ITEMS, +64*ITEMS) in order: item j of
  // lane l is element j*64 + l, so per-wave counters see them in input order.
  const int64_t tbase = tile * kSortTile;
  const int64_t wbase = tbase + (int64_t)w * kWave * kSortItems;
  const uint64_t ltmask = lanemask_lt();
  K k[kSortItems];
  uint32_t v[kSortItems], r[kSortItems];
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t idx = wbase + j * kWave + l;
    const bool valid = idx < n;
    k[j] = valid ? kin[idx] : K(0);
    v[j] = valid ? vin[idx] : 0u;
  }
  // The tile's digit counts first (a wave adds a digit all its lanes share
  // once), published before the ranking so that later tiles find them.
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const bool valid = wbase + j * kWave + l < n;   // the invalid lanes of a wave are its last ones
    const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
    const uint32_t d0 = __builtin_amdgcn_readfirstlane(d);
    const uint64_t act = __ballot(valid);
    if (__ballot(valid && d == d0) == act) {
      if (l == 0 && act) atomicAdd(&hc[d0], (uint32_t)__popcll(act));
    } else if (valid) {
      atomicAdd(&hc[d], 1u);
    }
  }
  __syncthreads();
  uint32_t spins = 0;
  bool out = false;   // a wait ran out
  auto wait = [&]() {
    if (++spins > kLbSpinLimit) {
      if (err) atomicOr(err, kSortErrTimeout);
      out = true;
    } else {
      __builtin_amdgcn_s_sleep(1);
    }
  };
  // keys of digit t in the earlier tiles of the group (all loaded at once)
  uint32_t ip = 0;
  auto in_group = [&]() {
    for (int64_t q = g0; q < tile && !out;) {
      uint32_t wv[kSortGroup - 1];
#pragma unroll
      for (int j = 0; j < kSortGroup - 1; ++j)
        wv[j] = q + j < tile ? __hip_atomic_load(&tw[(q + j) * kSortBins + t], __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)
                             : 0u;
      int j = 0;
      for (; j < kSortGroup - 1 && q + j < tile; ++j) {
        if ((wv[j] & kLbKind) == 0u) break;   // not published yet: reload from here
        ip += wv[j] & kLbCount;
      }
      q += j;
      if (q < tile) wait();
    }
  };
  if (t < kSortBins) {
    __hip_atomic_store(&tw[tile * kSortBins + t], kLbAggregate | hc[t], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (last) {   // the group's count, as early as the group's tiles allow
      in_group();
      __hip_atomic_store(&gw[g * kSortBins + t], (g == 0 ? kLbPrefix : kLbAggregate) | (ip + hc[t]),
                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }

#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t idx = wbase + j * kWave + l;
    const bool valid = idx < n;
    const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (d >> b) & 1u;
      const uint64_t bb = __ballot(bit);
      peers &= bit ? bb : ~bb;
    }
    const uint32_t before = (uint32_t)__popcll(peers & ltmask);
    const uint32_t cnt = (uint32_t)__popcll(peers);
    uint32_t prev = valid ? wc[w][d] : 0u;
    if (valid && before == 0) wc[w][d] = prev + cnt;
    r[j] = prev + before;
  }
  __syncthreads();

  // Thread t < 256 handles digit t: the waves' offsets inside the tile and
  // the digit's start in the tile.
  uint32_t run = 0, total = 0;
  if (t < kSortBins) {
#pragma unroll
    for (int q = 0; q < W; ++q) { const uint32_t c = wc[q][t]; wc[q][t] = run; run += c; }
#pragma unroll
    for (int q = 0; q < kSortHistSlots; ++q) total += hist[q * kHistStride + t];
  }
  uint32_t tot;
  const uint32_t ds = block_excl_scan_u32<kSortBlock>(run, scr, &tot);
  const uint32_t dbase = block_excl_scan_u32<kSortBlock>(total, scr, &tot);
  if (t < kSortBins) {
    dstart[t] = ds;
    if (!last) in_group();
    // keys of digit t in the earlier groups: their words backwards until a
    // running total, kLbWindow per round of loads
    uint32_t gp = 0;
    for (int64_t q = g - 1; q >= 0 && !out;) {
      uint32_t wv[kLbWindow];
#pragma unroll
      for (int j = 0; j < kLbWindow; ++j)
        wv[j] = q - j >= 0 ? __hip_atomic_load(&gw[(q - j) * kSortBins + t], __ATOMIC_RELAXED,
                                               __HIP_MEMORY_SCOPE_AGENT)
                           : kLbPrefix;   // before group 0: nothing
      int j = 0;
      bool done = false;
      for (; j < kLbWindow; ++j) {
        const uint32_t kind = wv[j] & kLbKind;
        if (kind == 0u) break;             // not published yet: reload from here
        gp += wv[j] & kLbCount;
        if (kind == kLbPrefix) { done = true; break; }
      }
      if (done) break;
      q -= j;
      if (j < kLbWindow) wait();
    }
    if (last && g > 0)
      __hip_atomic_store(&gw[g * kSortBins + t], kLbPrefix | (gp + ip + run), __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
    gbase[t] = dbase + gp + ip;
  }
  __syncthreads();
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int64_t idx = wbase + j * kWave + l;
    if (idx < n) {
      const uint32_t d = (uint32_t)(k[j] >> shift) & 255u;
      const uint32_t pos = dstart[d] + wc[w][d] + r[j];
      sk[pos] = k[j];
      sv[pos] = v[j];
    }
  }
  __syncthreads();
  const int64_t rem = n - tbase;
  const int cntv = rem < kSortTile ? (int)rem : kSortTile;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    const int p = j * kSortBlock + t;
    if (p < cntv) {
      const K kk = sk[p];
      const uint32_t d = (uint32_t)(kk >> shift) & 255u;
      const int64_t o = (int64_t)gbase[d] + (p - (int64_t)dstart[d]);
      kout[o] = kk;
      vout[o] = sv[p];
    }
  }
}

template <typename K>
bool radix_sort_impl(K* k0, uint32_t* v0, K* k1, uint32_t* v1, int64_t n, int lo, int bits, uint32_t* scratch,
                     int32_t* err, hipStream_t s) {
  const int passes = (int)div_up(bits, 8);
  if (n <= 0 || passes <= 0) return false;
  const int64_t ntiles = div_up(n, kSortTile);
  const size_t pw = sort_pass_words(n);
  uint32_t* hist = scratch;                                                 // [slots][passes][256]
  uint32_t* tickets = scratch + kSortHistSlots * kHistStride;               // [passes]
  uint32_t* words = scratch + kSortHeader;                                  // [passes][pw]
  const int64_t nw = kSortHeader + (int64_t)passes * (int64_t)pw;
  hipLaunchKernelGGL(k_sort_clear, dim3((unsigned)std::min<int64_t>(div_up(nw, kScanBlock), 1024)), dim3(kScanBlock),
                     0, s, scratch, nw);
  hipLaunchKernelGGL(k_sort_hist<K>, dim3((unsigned)std::min<int64_t>(div_up(n, (int64_t)kSortBlock * kHistRounds), 256)),
                     dim3(kSortBlock), 0, s, k0, n, lo, passes, hist);
  K* ki = k0; K* ko = k1;
  uint32_t* vi = v0; uint32_t* vo = v1;
  for (int p = 0; p < passes; ++p) {
    uint32_t* tw = words + (size_t)p * pw;
    hipLaunchKernelGGL(k_sort_pass<K>, dim3((unsigned)ntiles), dim3(kSortBlock), 0, s, ki, vi, ko, vo, n, ntiles,
                       lo + 8 * p, hist + p * kSortBins, tickets + p, tw, tw + ntiles * kSortBins, err);
    K* tk = ki; ki = ko; ko = tk;
    uint32_t* tv = vi; vi = vo; vo = tv;
  }
  return (passes & 1) != 0;
}

// ---- nearly sorted keys (patch_sort_u64) ----
// Scratch words: [0] taken positions m, [1] refused, then the descent list,
// the taken positions in order, their keys in order (8-byte words) and the
// positions those keys came from.
constexpr int kPatchBlock = 1024;
constexpr int kPatchTake = 2 * kPatchMaxDesc;         // taken positions at most
constexpr size_t kPatchDesc = 16;
constexpr size_t kPatchM = kPatchDesc + kPatchMaxDesc;
constexpr size_t kPatchDk = kPatchM + kPatchTake;     // even: 8-byte aligned
constexpr size_t kPatchDp = kPatchDk + 2 * kPatchTake;
static_assert(kPatchDp + kPatchTake == patch_scratch_words(), "patch scratch layout");
static_assert(kPatchDk % 2 == 0, "patch key alignment");

__global__ __launch_bounds__(kScanBlock) void k_patch_desc(const uint64_t* __restrict__ k, int64_t n,
                                                           uint32_t* __restrict__ ndesc, uint32_t* __restrict__ desc) {
  const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  if (i + 1 < n && k[i] > k[i + 1]) {
    const uint32_t j = atomicAdd(ndesc, 1u);
    if (j < (uint32_t)kPatchMaxDesc) desc[j] = (uint32_t)i;
  }
}

// Bitonic sort of len (a power of two, at most kPatchTake) entries in LDS by
// (a, b), all threads of the block taking part.
template <typename A>
__device__ void lds_bitonic(A* a, uint32_t* b, int len) {
  for (int kk = 2; kk <= len; kk <<= 1) {
    for (int j = kk >> 1; j > 0; j >>= 1) {
      for (int i = threadIdx.x; i < len; i += kPatchBlock) {
        const int l = i ^ j;
        if (l > i) {
          const bool up = (i & kk) == 0;
          const bool gt = a[i] > a[l] || (a[i] == a[l] && b[i] > b[l]);
          if (gt == up) {
            const A ta = a[i]; a[i] = a[l]; a[l] = ta;
            const uint32_t tb = b[i]; b[i] = b[l]; b[l] = tb;
          }
        }
      }
      __syncthreads();
    }
  }
}

// One workgroup: the taken positions (both ends of every descent, in order,
// each once), the check that the untaken keys are in order, and the taken
// keys sorted with their positions.
__global__ __launch_bounds__(kPatchBlock) void k_patch_plan(const uint64_t* __restrict__ k, int64_t n,
                                                            const uint32_t* __restrict__ ndesc,
                                                            uint32_t* __restrict__ plan, int32_t* __restrict__ err) {
  __shared__ uint32_t pos[kPatchTake];
  __shared__ uint32_t aux[kPatchTake];
  __shared__ uint64_t key[kPatchTake];
  __shared__ uint32_t scr[kPatchBlock / kWave + 1];
  __shared__ int bad;
  const int t = threadIdx.x;
  const uint32_t nd = *ndesc;
  if (t == 0) bad = nd > (uint32_t)kPatchMaxDesc;
  __syncthreads();
  if (bad) {
    if (t == 0) { plan[0] = 0; plan[1] = 1; atomicOr(err, kSortErrPatch); }
    return;
  }
  int len = 1;
  while (len < 2 * (int)nd) len <<= 1;
  const uint32_t* desc = plan + kPatchDesc;
  for (int i = t; i < len; i += kPatchBlock) {
    pos[i] = i < 2 * (int)nd ? desc[i >> 1] + (uint32_t)(i & 1) : 0xffffffffu;
    aux[i] = 0;
  }
  __syncthreads();
  lds_bitonic(pos, aux, len);
  // each position once (the ends of neighbouring descents coincide)
  const int per = (len + kPatchBlock - 1) / kPatchBlock;      // at most 2
  uint32_t f[2] = {0, 0};
  for (int q = 0; q < per; ++q) {
    const int i = t * per + q;
    if (i < len) f[q] = pos[i] != 0xffffffffu && (i == 0 || pos[i] != pos[i - 1]);
  }
  uint32_t tot;
  const uint32_t base = block_excl_scan_u32<kPatchBlock>(f[0] + f[1], scr, &tot);
  uint32_t m_pos[2];
  for (int q = 0; q < per; ++q) m_pos[q] = t * per + q < len ? pos[t * per + q] : 0u;
  __syncthreads();
  {
    uint32_t o = base;
    for (int q = 0; q < per; ++q)
      if (f[q]) aux[o++] = m_pos[q];
  }
  __syncthreads();
  const int m = (int)tot;
  // across each run of taken positions [a, b]: k[a - 1] <= k[b + 1]
  for (int i = t; i < m; i += kPatchBlock) {
    const uint32_t a = aux[i];
    if (i > 0 && aux[i - 1] + 1 == a) continue;              // not a run start
    int j = i;
    while (j + 1 < m && aux[j + 1] == aux[j] + 1) ++j;
    const uint32_t b = aux[j];
    if (a > 0 && (int64_t)b + 1 < n && k[a - 1] > k[b + 1]) bad = 1;
  }
  __syncthreads();
  if (bad) {
    if (t == 0) { plan[0] = 0; plan[1] = 1; atomicOr(err, kSortErrPatch); }
    return;
  }
  uint32_t* M = plan + kPatchM;
  int lenm = 1;
  while (lenm < m) lenm <<= 1;
  for (int i = t; i < lenm; i += kPatchBlock) {
    if (i < m) {
      M[i] = aux[i];
      key[i] = k[aux[i]];
      pos[i] = aux[i];
    } else {
      key[i] = ~0ull;
      pos[i] = 0xffffffffu;
    }
  }
  __syncthreads();
  lds_bitonic(key, pos, lenm);
  uint64_t* Dk = reinterpret_cast<uint64_t*>(plan + kPatchDk);
  uint32_t* Dp = plan + kPatchDp;
  for (int i = t; i < m; i += kPatchBlock) { Dk[i] = key[i]; Dp[i] = pos[i]; }
  if (t == 0) { plan[0] = (uint32_t)m; plan[1] = 0; }
}

// Positions taken below p: the taken positions are M[0, m), in order.
__device__ __forceinline__ int count_below(const uint32_t* M, int m, uint32_t p) {
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (M[mid] < p) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Taken keys below key: the taken keys are Dk[0, m), in order.
__device__ __forceinline__ int keys_below(const uint64_t* Dk, int m, uint64_t key) {
  int lo = 0, hi = m;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (Dk[mid] < key) lo = mid + 1; else hi = mid;
  }
  return lo;
}

// Taken entries a workgroup copies to LDS for its searches (more: the
// searches read the plan in global memory).
constexpr int kPatchLds = 512;

// Threads [0, n): the untaken keys, each to (untaken keys before it) +
// (taken keys below it).  Threads [n, n + m): taken key r to r + (untaken
// keys at most it), found by a binary search over positions in which a taken
// position stands for the nearest untaken one below it.
__global__ __launch_bounds__(kScanBlock) void k_patch_apply(const uint64_t* __restrict__ k, int64_t n,
                                                            const uint32_t* __restrict__ plan,
                                                            uint64_t* __restrict__ kout, uint32_t* __restrict__ perm) {
  __shared__ uint32_t sM[kPatchLds];
  __shared__ uint64_t sD[kPatchLds];
  const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  const bool refused = plan[1] != 0;
  const int m = refused ? 0 : (int)plan[0];
  const uint32_t* M = plan + kPatchM;
  const uint64_t* Dk = reinterpret_cast<const uint64_t*>(plan + kPatchDk);
  const uint32_t* Dp = plan + kPatchDp;
  const bool lds = m <= kPatchLds;                          // uniform
  if (lds && (int64_t)blockIdx.x * kScanBlock < n) {
    for (int j = threadIdx.x; j < m; j += kScanBlock) { sM[j] = M[j]; sD[j] = Dk[j]; }
    __syncthreads();
  }
  if (i < n) {
    if (refused) { kout[i] = k[i]; perm[i] = (uint32_t)i; return; }
    const int below = lds ? count_below(sM, m, (uint32_t)i) : count_below(M, m, (uint32_t)i);
    const uint32_t* Mb = lds ? sM : M;
    if (below < m && Mb[below] == (uint32_t)i) return;       // taken: placed by its own thread
    const uint64_t key = k[i];
    const int lo = lds ? keys_below(sD, m, key) : keys_below(Dk, m, key);
    const int64_t o = i - below + lo;
    kout[o] = key;
    perm[o] = (uint32_t)i;
    return;
  }
  const int64_t r = i - n;
  if (r >= m) return;
  const uint64_t key = Dk[r];
  // nearest untaken position at or below p (-1: none)
  auto untaken_at_or_below = [&](int64_t p) {
    int j = count_below(M, m, (uint32_t)(p + 1)) - 1;        // last taken position <= p
    while (j >= 0 && (int64_t)M[j] == p) { --p; --j; }
    return p;
  };
  int64_t lo = -1, hi = n;                                   // f(lo) true, f(hi) false
  while (hi - lo > 1) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    const int64_t u = untaken_at_or_below(mid);
    if (u < 0 || k[u] <= key) lo = mid; else hi = mid;
  }
  const int64_t u = lo < 0 ? -1 : untaken_at_or_below(lo);
  const int64_t cnt = u < 0 ? 0 : u + 1 - count_below(M, m, (uint32_t)u);
  const int64_t o = r + cnt;
  kout[o] = key;
  perm[o] = Dp[r];
}

}  // namespace

void scan_u32(const uint32_t* in, uint32_t* out, int64_t n, uint32_t* total, uint32_t* scratch, hipStream_t s) {
  if (n <= 0) {
    if (total) (void)hipMemsetAsync(total, 0, sizeof(uint32_t), s);
    return;
  }
  const int64_t nb = div_up(n, kScanTile);
  hipLaunchKernelGGL(k_scan_reduce, dim3((unsigned)nb), dim3(kScanBlock), 0, s, in, n, scratch);
  hipLaunchKernelGGL(k_scan_down, dim3((unsigned)nb), dim3(kScanBlock), 0, s, in, out, n, scratch, nb, total);
}

bool radix_sort_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, int64_t n, int lo, int bits,
                    uint32_t* scratch, int32_t* err, hipStream_t s) {
  return radix_sort_impl<uint64_t>(k0, v0, k1, v1, n, lo, bits, scratch, err, s);
}

bool radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int lo, int bits,
                    uint32_t* scratch, int32_t* err, hipStream_t s) {
  return radix_sort_impl<uint32_t>(k0, v0, k1, v1, n, lo, bits, scratch, err, s);
}

void patch_sort_u64(const uint64_t* k, uint64_t* kout, uint32_t* perm, int64_t n, uint32_t* ndesc,
                    uint32_t* scratch, int32_t* err, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_patch_desc, dim3((unsigned)div_up(n, kScanBlock)), dim3(kScanBlock), 0, s, k, n, ndesc,
                     scratch + kPatchDesc);
  hipLaunchKernelGGL(k_patch_plan, dim3(1), dim3(kPatchBlock), 0, s, k, n, ndesc, scratch, err);
  hipLaunchKernelGGL(k_patch_apply, dim3((unsigned)div_up(n + kPatchTake, kScanBlock)), dim3(kScanBlock), 0, s, k, n,
                     scratch, kout, perm);
}

}  // namespace nbx
