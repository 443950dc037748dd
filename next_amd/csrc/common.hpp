// Shared device-side definitions for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace nbx {

constexpr int kWave = 64;          // CDNA wavefront width
constexpr int kKeyDigits = 21;     // octree digits per 64-bit key word
constexpr int kMaxLevels = 42;     // two key words
constexpr int kMaxDepthSlots = 64; // histogram slots for node depths

// One tree node in depth-first preorder, children in digit order: what the
// upward pass and the walk's evaluation read (the traversal reads the
// node's WalkHead below).  128 bytes = one L2 line.
constexpr int32_t kLinkZeroMass = 1 << 30;   // link flag: zero mass, skipped whole
struct alignas(128) RefNode {
  double cx, cy, cz, m;    // centre of mass, mass
  int32_t skip;            // preorder index of the first node after this subtree
  int32_t body;            // leaf: sorted body index; internal: -1
  int32_t depth;
  int32_t first;           // first body (sorted index) of the subtree
  double size;             // half-width: root half-width * 2^-depth
  double esum;             // size*0.015 + cbrt(m)*0.002 (softening numerator)
  double qxx, qyy, qzz, qxy, qxz, qyz;
  double pad_[2];
};
static_assert(sizeof(RefNode) == 128, "RefNode layout");

// The per-visit fields of a node for the walk's traversal, in an array of
// their own beside the node records (two per 64-byte scalar cache line): a
// visited node costs one 32-byte scalar load.  acc2 is the upper edge of the
// opening test's band in FP32, (size/theta)^2 (1 + 2e-6) rounded up, or NaN
// where that is out of FP32's normal range (every decision then exact).
struct alignas(32) WalkHead {
  double cx, cy, cz;       // centre of mass
  float acc2;
  // leaf: ~body (negative); internal: skip; with kLinkZeroMass set (and the
  // skip target) when the node's mass is zero
  int32_t link;
};
static_assert(sizeof(WalkHead) == 32, "WalkHead layout");

// Small per-context device block: flags and the per-level node counts the
// host needs to launch the upward pass.
struct DevScalars {
  int32_t err;              // bit 0: non-finite position; bit 1: radix sort wait ran out; bit 2: patch sort refused
  int32_t collisions;       // adjacent equal primary keys
  int32_t coincident;       // adjacent equal primary and secondary keys
  int32_t total_nodes;              // (internal nodes: the sum of level_count)
  int32_t max_leaf_depth;
  int32_t level_count[kMaxDepthSlots];  // internal nodes per depth
  unsigned long long n_pp, n_pc;        // interaction counters
  unsigned long long n_open, n_iter;    // opened nodes, wave traversal iterations
  unsigned long long n_eval;            // wave evaluation steps (deferred interactions)
  uint32_t tree_groups;                 // walk groups built from the tree (physics mode)
  // Builds that never read counts back to the host (captured steps): each
  // depth's first entry in the node list, and a sticky flag set when a
  // build does not fit what its launches were sized for (kAbort*), after
  // which the tree kernels of the run do nothing.
  int32_t abort_flags;
  int32_t level_off[kMaxDepthSlots];
  int32_t step_count;                   // next entry of the dt log of a captured run
  uint32_t ndesc;                       // descents of nearly sorted keys (patch_sort_u64), cleared per build
};

constexpr int32_t kAbortCapacity = 1;   // nodes or node-list entries beyond the buffers
constexpr int32_t kAbortDepth = 2;      // deeper than the levels launched
constexpr int32_t kAbortCollision = 4;  // bodies sharing all 42 digits (the eager build reports it)
constexpr int32_t kAbortNonFinite = 8;  // a position is NaN or infinite
constexpr int32_t kAbortSort = 16;      // the radix sort's wait for earlier tiles ran out
constexpr int32_t kAbortPatch = 32;     // the nearly-sorted path refused the keys (the bodies moved too much)

// Per-system arrays: a context holds one or more independent particle
// systems, stored one after another in key order.
struct SysArrays {
  double* root;                 // 4 per system: centre x, y, z, half-width
  double* dt;                   // time step of the current step
  unsigned long long* v2max;    // bit pattern of max |v|^2 (non-negative doubles order as integers)
  int32_t* soff;                // first body of each system; soff[nsys] = n
  int32_t* sroot;               // preorder index of each system's root
  int32_t* send;                // preorder index after each system's tree
  int32_t nsys;
};

// One wave of walk targets: up to 64 consecutive bodies of one system.
struct Group {
  int32_t first, count, sys, pad;
};

// A block of the segmented bounding-box reduction: bodies [begin, end) of one system.
struct BoxBlock {
  int32_t sys, begin, end, pad;
};

// ---- wave primitives ------------------------------------------------------

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int wave_min_i32(int v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    int o = __shfl_xor(v, off, kWave);
    v = o < v ? o : v;
  }
  return v;
}

__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int l = lane_id();
  return l == 0 ? 0ull : (~0ull >> (64 - l));
}

// Inclusive scan within a wave (Hillis-Steele over shuffles).
__device__ __forceinline__ uint32_t wave_incl_scan_u32(uint32_t v) {
  const int l = lane_id();
#pragma unroll
  for (int off = 1; off < kWave; off <<= 1) {
    uint32_t o = __shfl_up(v, off, kWave);
    if (l >= off) v += o;
  }
  return v;
}

// Exclusive scan of one value per thread across a block of BLOCK threads
// (BLOCK a multiple of 64, at most 1024).  Returns the exclusive prefix;
// *total receives the block sum.  `scratch` holds BLOCK/64 words.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_excl_scan_u32(uint32_t v, uint32_t* scratch, uint32_t* total) {
  constexpr int W = BLOCK / kWave;
  const int t = threadIdx.x, w = t / kWave, l = t % kWave;
  uint32_t inc = wave_incl_scan_u32(v);
  if (l == kWave - 1) scratch[w] = inc;
  __syncthreads();
  if (w == 0) {
    uint32_t s = l < W ? scratch[l] : 0u;
    uint32_t si = wave_incl_scan_u32(s);
    if (l < W) scratch[l] = si - s;
    if (l == W - 1) scratch[W] = si;
  }
  __syncthreads();
  uint32_t r = scratch[w] + inc - v;
  *total = scratch[W];
  __syncthreads();
  return r;
}

}  // namespace nbx
