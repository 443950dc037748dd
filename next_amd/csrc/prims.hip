// Device-wide exclusive scan and stable LSD radix sort for gfx950.
//
// Scan: reduce-then-scan in two launches (tile sums; tile-local scan plus the
// sum of the preceding tile sums, which every tile forms itself).  No
// inter-workgroup waiting.
//
// Radix sort: 8-bit digits, three launches per digit (tile histograms, one
// workgroup per digit scanning that digit's column over the tiles, scatter;
// the scatter adds the exclusive prefix of the digit totals, which every tile
// forms itself from 256 values).  The scatter ranks keys
// stably inside a tile with wave64 ballots (8 ballots give each lane the set
// of lanes holding the same digit), stages the tile in LDS in digit order and
// writes it out in runs, so consecutive lanes store consecutive addresses.
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

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const K* __restrict__ keys, int64_t n, int shift,
                                                           uint32_t* __restrict__ hist, int64_t ntiles) {
  __shared__ uint32_t h[kSortBlock / kWave][kSortBins];
  uint32_t* hf = &h[0][0];
  for (int i = threadIdx.x; i < (kSortBlock / kWave) * kSortBins; i += kSortBlock) hf[i] = 0;
  __syncthreads();
  const int w = threadIdx.x / kWave;
  const int64_t base = (int64_t)blockIdx.x * kSortTile;
#pragma unroll
  for (int j = 0; j < kSortItems; ++j) {
    int64_t i = base + j * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&h[w][(uint32_t)(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  uint32_t c = 0;
#pragma unroll
  for (int k = 0; k < kSortBlock / kWave; ++k) c += h[k][threadIdx.x];
  hist[(int64_t)threadIdx.x * ntiles + blockIdx.x] = c;
}

// Digit d's column of the tile histograms (hist[d * ntiles + t]) becomes its
// exclusive prefix over the tiles; coltot[d] receives the column total.
__global__ __launch_bounds__(kSortBlock) void k_radix_colscan(uint32_t* __restrict__ hist, int64_t ntiles,
                                                              uint32_t* __restrict__ coltot) {
  __shared__ uint32_t scr[kSortBlock / kWave + 1];
  uint32_t* col = hist + (int64_t)blockIdx.x * ntiles;
  uint32_t carry = 0;
  for (int64_t c = 0; c < ntiles; c += kSortBlock) {
    const int64_t i = c + threadIdx.x;
    const uint32_t v = i < ntiles ? col[i] : 0u;
    uint32_t tot;
    const uint32_t ex = block_excl_scan_u32<kSortBlock>(v, scr, &tot);
    if (i < ntiles) col[i] = carry + ex;
    carry += tot;
  }
  if (threadIdx.x == 0) coltot[blockIdx.x] = carry;
}

template <typename K>
__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(const K* __restrict__ kin,
                                                              const uint32_t* __restrict__ vin,
                                                              K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                              int64_t n, int shift,
                                                              const uint32_t* __restrict__ off, int64_t ntiles,
                                                              const uint32_t* __restrict__ coltot) {
  constexpr int W = kSortBlock / kWave;
  __shared__ K sk[kSortTile];
  __shared__ uint32_t sv[kSortTile];
  __shared__ uint32_t wc[W][kSortBins];
  __shared__ uint32_t dstart[kSortBins];
  __shared__ uint32_t gbase[kSortBins];
  __shared__ uint32_t scr[W + 1];
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  uint32_t* wcf = &wc[0][0];
  for (int i = t; i < W * kSortBins; i += kSortBlock) wcf[i] = 0;
  __syncthreads();

  // Wave w ranks elements [tile + w*64*ITEMS, +64*ITEMS) in order: item j of
  // lane l is element j*64 + l, so per-wave counters see them in input order.
  const int64_t tbase = (int64_t)blockIdx.x * kSortTile;
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

  {
    uint32_t c[W];
    uint32_t run = 0;
#pragma unroll
    for (int q = 0; q < W; ++q) { c[q] = wc[q][t]; wc[q][t] = run; run += c[q]; }
    uint32_t tot;
    uint32_t ds = block_excl_scan_u32<kSortBlock>(run, scr, &tot);
    dstart[t] = ds;
    const uint32_t dbase = block_excl_scan_u32<kSortBlock>(coltot[t], scr, &tot);   // keys of smaller digits
    gbase[t] = dbase + off[(int64_t)t * ntiles + blockIdx.x];
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
bool radix_sort_impl(K* k0, uint32_t* v0, K* k1, uint32_t* v1, int64_t n, int bits, uint32_t* scratch,
                     hipStream_t s) {
  if (n <= 0) return false;
  const int64_t ntiles = div_up(n, kSortTile);
  uint32_t* hist = scratch;
  uint32_t* coltot = scratch + kSortBins * ntiles;
  const int passes = (int)div_up(bits, 8);
  K* ki = k0; K* ko = k1;
  uint32_t* vi = v0; uint32_t* vo = v1;
  for (int p = 0; p < passes; ++p) {
    const int shift = 8 * p;
    hipLaunchKernelGGL(k_radix_hist<K>, dim3((unsigned)ntiles), dim3(kSortBlock), 0, s, ki, n, shift, hist,
                       ntiles);
    hipLaunchKernelGGL(k_radix_colscan, dim3(kSortBins), dim3(kSortBlock), 0, s, hist, ntiles, coltot);
    hipLaunchKernelGGL(k_radix_scatter<K>, dim3((unsigned)ntiles), dim3(kSortBlock), 0, s, ki, vi, ko, vo, n,
                       shift, hist, ntiles, coltot);
    K* tk = ki; ki = ko; ko = tk;
    uint32_t* tv = vi; vi = vo; vo = tv;
  }
  return (passes & 1) != 0;
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

bool radix_sort_u64(uint64_t* k0, uint32_t* v0, uint64_t* k1, uint32_t* v1, int64_t n, int bits,
                    uint32_t* scratch, hipStream_t s) {
  return radix_sort_impl<uint64_t>(k0, v0, k1, v1, n, bits, scratch, s);
}

bool radix_sort_u32(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, int64_t n, int bits,
                    uint32_t* scratch, hipStream_t s) {
  return radix_sort_impl<uint32_t>(k0, v0, k1, v1, n, bits, scratch, s);
}

}  // namespace nbx
