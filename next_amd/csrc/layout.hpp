// Placement of internal tree nodes in the per-depth node list and the list
// of their skip targets, shared by the reference and physics layouts.
//
// With lcp[i] = digits shared by sorted bodies i and i+1 (-1 between
// systems and past the end), body i starts the internal nodes at depths
// (lcp[i-1], lcp[i]] and ends those at depths (lcp[i], lcp[i-1]].  The k-th
// internal node of depth d in start order is the k-th in end order (nodes of
// one depth are disjoint and nest nothing of their own depth), so placing
// both kinds at
//     (nodes of smaller depth) + (starts / ends at depth d by earlier bodies)
// pairs every node with its skip target, and the node list is sorted by
// depth and, within a depth, by preorder: the upward pass's order.
//
// The count of earlier bodies comes from the scanned per-block counts of
// k_lcp_count (columns 0..63 starts, 64..127 ends, per depth; column c of
// block b at part[c * blocks + b], with the same blocks as k_lcp_count) plus
// a rank inside the block from wave ballots, in body order.
#pragma once

#include "common.hpp"
#include "ref_kernels.hpp"

namespace nbx {

// Called by every thread of a kBlock-thread block (valid: the thread holds
// a body).  lp, li: lcp of the body with its predecessor and successor;
// pre: preorder index of its first started node; skipv: the preorder index
// after the nodes it ends.  part: the scanned counts (this block's at
// part[c * gridDim.x + blockIdx.x]).
// Entries at or beyond cap are not written.
__device__ __forceinline__ void place_internal_nodes(bool valid, int lp, int li, uint32_t pre, uint32_t skipv,
                                                     const uint32_t* __restrict__ part,
                                                     const DevScalars* __restrict__ sc,
                                                     int32_t* __restrict__ level_list,
                                                     uint32_t* __restrict__ end_list, uint32_t cap) {
  constexpr int W = kBlock / kWave;
  constexpr int D = kMaxLevels + 1;                 // internal node depths 0..41 (one spare)
  __shared__ uint32_t off[kMaxDepthSlots];         // nodes of smaller depth
  __shared__ uint32_t wc[2][W][D];                 // per wave: starts / ends at depth d, then their prefix
  __shared__ int drange[2];
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  if (t < kWave) {
    const uint32_t v = (uint32_t)sc->level_count[t];
    const uint32_t inc = wave_incl_scan_u32(v);
    off[t] = inc - v;
  }
  if (t == 0) { drange[0] = D; drange[1] = -1; }
  __syncthreads();
  int lo = D, hi = -1;
  if (valid) {
    lo = (lp < li ? lp : li) + 1;
    hi = lp > li ? lp : li;
  }
  lo = wave_min_i32(lo);
  hi = -wave_min_i32(-hi);
  if (l == 0) {
    atomicMin(&drange[0], lo);
    atomicMax(&drange[1], hi);
  }
  __syncthreads();
  const int dlo = drange[0], dhi = drange[1];
  const uint64_t lt = lanemask_lt();
  for (int d = dlo; d <= dhi; ++d) {
    const bool fs = valid && lp < d && d <= li;
    const bool fe = valid && li < d && d <= lp;
    const uint64_t ms = __builtin_amdgcn_ballot_w64(fs), me = __builtin_amdgcn_ballot_w64(fe);
    if (l == 0) {
      wc[0][w][d] = (uint32_t)__popcll(ms);
      wc[1][w][d] = (uint32_t)__popcll(me);
    }
  }
  __syncthreads();
  // exclusive prefix over the waves, one thread per (kind, depth)
  for (int j = t; j < 2 * D; j += kBlock) {
    const int kind = j / D, d = j % D;
    if (d >= dlo && d <= dhi) {
      uint32_t run = 0;
      for (int q = 0; q < W; ++q) {
        const uint32_t c = wc[kind][q][d];
        wc[kind][q][d] = run;
        run += c;
      }
    }
  }
  __syncthreads();
  for (int d = dlo; d <= dhi; ++d) {
    const bool fs = valid && lp < d && d <= li;
    const bool fe = valid && li < d && d <= lp;
    const uint64_t ms = __builtin_amdgcn_ballot_w64(fs), me = __builtin_amdgcn_ballot_w64(fe);
    if (fs) {
      const uint32_t pos = off[d] + part[(int64_t)d * gridDim.x + blockIdx.x] + wc[0][w][d] +
                           (uint32_t)__popcll(ms & lt);
      if (pos < cap) level_list[pos] = (int32_t)(pre + (uint32_t)(d - lp - 1));
    }
    if (fe) {
      const uint32_t pos = off[d] + part[(int64_t)(kMaxDepthSlots + d) * gridDim.x + blockIdx.x] + wc[1][w][d] +
                           (uint32_t)__popcll(me & lt);
      if (pos < cap) end_list[pos] = skipv;
    }
  }
}

// Internal nodes of the last build: the sum of the per-depth counts.
inline int64_t internal_count(const DevScalars& s) {
  int64_t m = 0;
  for (int d = 0; d < kMaxDepthSlots; ++d) m += (uint32_t)s.level_count[d];
  return m;
}

}  // namespace nbx
