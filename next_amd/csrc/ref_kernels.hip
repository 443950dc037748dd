// Reference-mode kernels: NEXT's FP64 gravity + time-step path on gfx950.
//
// Compiled with -ffp-contract=off: every product and sum is rounded as
// written, in the order NEXT writes it, so the GPU reproduces the reference
// arithmetic operation by operation.
//
// Pipeline of one tree build (positions in `st`):
//   bbox -> root cube (step.h:68-102)
//   keys: each body replays the FP64 node-centre recurrence of
//         octree.h:49-64 for 21 levels -> 63-bit digit key
//   stable radix sort of keys, gather of the state into key order
//   lcp of neighbouring keys -> the one-body-per-leaf octree of
//         octree.h:69-87 in depth-first preorder (children 0..7), built from
//         the rule "an internal node exists for every digit prefix shared by
//         two or more bodies", with skip pointers for stackless traversal
//   upward pass, one launch per depth, deepest first (octree.h:92-127)
//   walk: one wave per 64 consecutive targets; the wave visits the union of
//         its lanes' depth-first paths in preorder, each lane applying its own
//         acceptance test (octree.h:133-179), so every lane accumulates the
//         same terms in the same order as the recursive reference.
#include "common.hpp"
#include "layout.hpp"
#include "prims.hpp"
#include "ref_kernels.hpp"

#include <algorithm>
#include <climits>
#include <type_traits>

namespace nbx {
namespace {

__device__ __forceinline__ void lo_upd(double& lo, double p) { if (p < lo) lo = p; }   // std::min(lo, p)
__device__ __forceinline__ void hi_upd(double& hi, double p) { if (hi < p) hi = p; }   // std::max(hi, p)

// Block-wide min/max of 6 values (3 lows, 3 highs) written as one partial.
__device__ void block_bbox_write(double lo[3], double hi[3], double* part) {
  __shared__ double sm[kBlock / kWave][6];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      lo_upd(lo[k], __shfl_xor(lo[k], off, kWave));
      hi_upd(hi[k], __shfl_xor(hi[k], off, kWave));
    }
  }
  const int w = threadIdx.x / kWave;
  if (lane_id() == 0) {
    for (int k = 0; k < 3; ++k) { sm[w][k] = lo[k]; sm[w][3 + k] = hi[k]; }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < kBlock / kWave; ++q)
      for (int k = 0; k < 3; ++k) { lo_upd(lo[k], sm[q][k]); hi_upd(hi[k], sm[q][3 + k]); }
    for (int k = 0; k < 3; ++k) { part[6 * blockIdx.x + k] = lo[k]; part[6 * blockIdx.x + 3 + k] = hi[k]; }
  }
}

__global__ __launch_bounds__(kBlock) void k_fill_scalars(DevScalars* sc, int what) {
  const int t = threadIdx.x;
  if (what & 1) {
    if (t < kMaxDepthSlots) sc->level_count[t] = 0;
    if (t == 0) { sc->collisions = 0; sc->coincident = 0; sc->max_leaf_depth = 0; sc->ndesc = 0; }
  }
  if ((what & 4) && t == 0) sc->err = 0;
  if ((what & 8) && t == 0) { sc->n_pp = 0ull; sc->n_pc = 0ull; sc->n_open = 0ull; sc->n_iter = 0ull; sc->n_eval = 0ull; }
}

// Start of a captured run: no abort yet, dt log entries from `step` on.
__global__ void k_set_run(DevScalars* sc, int32_t step) {
  sc->abort_flags = 0;
  sc->step_count = step;
}

__global__ __launch_bounds__(kBlock) void k_fill_u32(int64_t n, uint32_t* __restrict__ a, uint32_t v) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) a[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_fill_v2max(SysArrays sy) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s < sy.nsys) sy.v2max[s] = 0ull;
}

__global__ __launch_bounds__(kBlock) void k_set_dt(SysArrays sy, double dt) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s < sy.nsys) sy.dt[s] = dt;
}

// Min/max of the bodies [begin, end) of one system (one block).
__global__ __launch_bounds__(kBlock) void k_bbox_partial(StateArrays st, const BoxBlock* __restrict__ blocks,
                                                         double* part) {
  const BoxBlock b = blocks[blockIdx.x];
  double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
  for (int64_t i = b.begin + threadIdx.x; i < b.end; i += kBlock) {
    const double p[3] = {st.x[i], st.y[i], st.z[i]};
    for (int k = 0; k < 3; ++k) { lo_upd(lo[k], p[k]); hi_upd(hi[k], p[k]); }
  }
  block_bbox_write(lo, hi, part);
}

// Root centre and half-width of each system from its blocks' partials
// (step.h:93-100), one wave per system; min and max are exact, so the
// order does not matter.  clear (may be null): scalars to reset for the
// tree build that follows.
__global__ __launch_bounds__(kWave) void k_bbox_final(const double* __restrict__ part,
                                                      const int32_t* __restrict__ block_off, SysArrays sy,
                                                      DevScalars* clear) {
  const int s = blockIdx.x;
  if (clear && s == 0) {     // a tree build's counts and flags (k_fill_scalars, what = 1 | 4)
    const int t = threadIdx.x;
    clear->level_count[t] = 0;
    if (t == 0) {
      clear->collisions = 0; clear->coincident = 0; clear->max_leaf_depth = 0; clear->err = 0; clear->ndesc = 0;
    }
  }
  if (s >= sy.nsys) return;
  double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
  for (int b = block_off[s] + (int)threadIdx.x; b < block_off[s + 1]; b += kWave)
    for (int k = 0; k < 3; ++k) { lo_upd(lo[k], part[6 * b + k]); hi_upd(hi[k], part[6 * b + 3 + k]); }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
    for (int k = 0; k < 3; ++k) {
      lo_upd(lo[k], __shfl_xor(lo[k], off, kWave));
      hi_upd(hi[k], __shfl_xor(hi[k], off, kWave));
    }
  if (threadIdx.x != 0) return;
  const double cx = (lo[0] + hi[0]) * 0.5;
  const double cy = (lo[1] + hi[1]) * 0.5;
  const double cz = (lo[2] + hi[2]) * 0.5;
  double e = hi[0] - lo[0];
  const double ey = hi[1] - lo[1], ez = hi[2] - lo[2];
  if (e < ey) e = ey;
  if (e < ez) e = ez;
  double h = e * 0.5;
  if (h <= 0.0) h = 1.0;
  double* r = sy.root + 4 * s;
  r[0] = cx; r[1] = cy; r[2] = cz; r[3] = h;
}

// Digits of one body along the node-centre recurrence (octree.h:49-64),
// levels [l0, l1), starting from the root.
__device__ __forceinline__ void key_digits(double px, double py, double pz, const double* root, int l0, int l1,
                                           uint64_t& hi, uint64_t& lo) {
  double cx = root[0], cy = root[1], cz = root[2], s = root[3];
  hi = 0; lo = 0;
  for (int l = 0; l < l1; ++l) {
    const uint32_t o = (px > cx ? 1u : 0u) | (py > cy ? 2u : 0u) | (pz > cz ? 4u : 0u);
    if (l >= l0) {
      if (l < kKeyDigits) hi = (hi << 3) | o; else lo = (lo << 3) | o;
    }
    const double hs = s * 0.5;
    cx = cx + ((o & 1u) ? hs : -hs);
    cy = cy + ((o & 2u) ? hs : -hs);
    cz = cz + ((o & 4u) ? hs : -hs);
    s = hs;
  }
}

__device__ __forceinline__ void sys_root(const SysArrays& sy, int32_t s, double root[4]) {
  const double* r = sy.root + 4 * s;
  root[0] = r[0]; root[1] = r[1]; root[2] = r[2]; root[3] = r[3];
}

__global__ __launch_bounds__(kBlock) void k_keys(int64_t n, StateArrays st, SysArrays sy,
                                                 uint64_t* __restrict__ key, uint32_t* __restrict__ idx,
                                                 DevScalars* scw) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double px = st.x[i], py = st.y[i], pz = st.z[i];
  if (!(__builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz))) atomicOr(&scw->err, 1);
  double root[4];
  sys_root(sy, st.sys[i], root);
  uint64_t hi, lo;
  key_digits(px, py, pz, root, 0, kKeyDigits, hi, lo);
  key[i] = hi;
  idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_keys2(int64_t n, StateArrays st, SysArrays sy,
                                                  uint64_t* __restrict__ key2) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double root[4];
  sys_root(sy, st.sys[i], root);
  uint64_t hi, lo;
  key_digits(st.x[i], st.y[i], st.z[i], root, kKeyDigits, kMaxLevels, hi, lo);
  key2[i] = lo;
}

// The sort orders bodies by bits [lo, 63) of the first key word (its top
// digits: enough to separate the bodies of a tree of the last build's depth).
// Bodies of one system whose words agree on those bits sit next to each
// other, in their previous order.  Each such run is sorted in place by the
// whole first word, then the second word (digits 21..41), stably: the order a
// sort by (system, first word, second word) with ties in the previous order
// gives.  The second word is written for the runs' bodies only, the only ones
// whose second word is ever read (where two neighbours' first words are
// equal).  Runs are disjoint and a swap inside one keeps its bits [lo, 63), so
// the run bounds every thread finds do not depend on the others' swaps.
__global__ __launch_bounds__(kBlock) void k_fix_runs(int64_t n, StateArrays st, SysArrays sy, uint64_t* key,
                                                     uint64_t* __restrict__ key2, uint32_t* __restrict__ amag,
                                                     int lo) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t s = st.sys[i];
  const uint64_t top = key[i] >> lo;
  auto same = [&](int64_t j) { return st.sys[j] == s && (key[j] >> lo) == top; };
  const bool eq_next = i + 1 < n && same(i + 1);
  const bool eq_prev = i > 0 && same(i - 1);
  if (!eq_next || eq_prev) return;                           // not the first body of a run
  int64_t j = i + 1;                                         // the run is [i, j]
  while (j + 1 < n && same(j + 1)) ++j;
  double root[4];
  sys_root(sy, s, root);
  for (int64_t t = i; t <= j; ++t) {
    uint64_t hi, lo2;
    key_digits(st.x[t], st.y[t], st.z[t], root, kKeyDigits, kMaxLevels, hi, lo2);
    key2[t] = lo2;
  }
  for (int64_t t = i + 1; t <= j; ++t)
    for (int64_t u = t; u > i && (key[u - 1] > key[u] || (key[u - 1] == key[u] && key2[u - 1] > key2[u])); --u) {
      const int64_t v = u - 1;
      auto sw_d = [&](double* a) { const double w = a[u]; a[u] = a[v]; a[v] = w; };
      auto sw_i = [&](int32_t* a) { const int32_t w = a[u]; a[u] = a[v]; a[v] = w; };
      auto sw_k = [&](uint64_t* a) { const uint64_t w = a[u]; a[u] = a[v]; a[v] = w; };
      sw_d(st.x); sw_d(st.y); sw_d(st.z); sw_d(st.vx); sw_d(st.vy); sw_d(st.vz); sw_d(st.m);
      sw_i(st.type); sw_i(st.id);
      sw_k(key); sw_k(key2);
      if (amag) { const uint32_t a = amag[u]; amag[u] = amag[v]; amag[v] = a; }
    }
}

__global__ __launch_bounds__(kBlock) void k_keys_export(int64_t n, StateArrays st, SysArrays sy,
                                                        uint64_t* __restrict__ hi_out,
                                                        uint64_t* __restrict__ lo_out) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  double root[4];
  sys_root(sy, st.sys[i], root);
  uint64_t hi, lo;
  key_digits(st.x[i], st.y[i], st.z[i], root, 0, kMaxLevels, hi, lo);
  const int32_t o = st.id[i];
  hi_out[o] = hi;
  lo_out[o] = lo;
}

__global__ __launch_bounds__(kBlock) void k_iota(int64_t n, uint32_t* idx) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) idx[i] = (uint32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_gather_u64(int64_t n, const uint32_t* __restrict__ perm,
                                                       const uint64_t* __restrict__ src, uint64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

__global__ __launch_bounds__(kBlock) void k_gather_state(int64_t n, const uint32_t* __restrict__ perm, StateArrays a,
                                                         StateArrays b) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  b.x[i] = a.x[p]; b.y[i] = a.y[p]; b.z[i] = a.z[p];
  b.vx[i] = a.vx[p]; b.vy[i] = a.vy[p]; b.vz[i] = a.vz[p];
  b.m[i] = a.m[p]; b.type[i] = a.type[p]; b.id[i] = a.id[p]; b.sys[i] = a.sys[p];
}

__global__ __launch_bounds__(kBlock) void k_gather_sys(int64_t n, const uint32_t* __restrict__ perm,
                                                       const int32_t* __restrict__ sys, uint32_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = (uint32_t)sys[perm[i]];
}

// Map for removing the bodies [cut, cut + len) of one system: slot i of the
// shortened arrays takes old slot perm[i].
__global__ __launch_bounds__(kBlock) void k_remove_map(int64_t n_new, int64_t cut, int64_t len,
                                                       uint32_t* __restrict__ perm) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n_new) perm[i] = (uint32_t)(i < cut ? i : i + len);
}

__global__ __launch_bounds__(kBlock) void k_relabel(int64_t n, int32_t* __restrict__ sys, int32_t removed) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n && sys[i] > removed) sys[i] -= 1;
}

__global__ __launch_bounds__(kBlock) void k_gather_f64(int64_t n, const uint32_t* __restrict__ perm,
                                                       const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) dst[i] = src[perm[i]];
}

// Number of leading octree digits two keys share (0..21 per word).
__device__ __forceinline__ int lcp_word(uint64_t a, uint64_t b) {
  const uint64_t d = a ^ b;
  return d ? (__clzll((long long)d) - 1) / 3 : kKeyDigits;
}

// Bodies of different systems share nothing: -1 separates their trees.
__device__ __forceinline__ int lcp_pair(const uint64_t* key, const uint64_t* key2, const int32_t* sys, int64_t i,
                                        int& equal) {
  equal = 0;
  if (sys[i] != sys[i + 1]) return -1;
  int l = lcp_word(key[i], key[i + 1]);
  if (l == kKeyDigits) {
    if (key2) {
      l += lcp_word(key2[i], key2[i + 1]);
      if (l == kMaxLevels) equal = 1;
    } else {
      equal = 1;
    }
  }
  return l;
}

// lcp[i] = digits shared by sorted bodies i and i+1 (-1 past the end).
// Body i starts the internal nodes at depths (lcp[i-1], lcp[i]] and its leaf;
// it ends the internal nodes at depths (lcp[i], lcp[i-1]].  Each block also
// writes one row of partial counts (internal nodes started and ended per
// depth, deepest leaf, equal-key pairs) for k_lcp_scan: plain stores, no
// global atomics.
__global__ __launch_bounds__(kBlock) void k_lcp_count(int64_t n, const uint64_t* __restrict__ key,
                                                      const uint64_t* __restrict__ key2,
                                                      const int32_t* __restrict__ sys, int32_t* __restrict__ lcp,
                                                      uint32_t* __restrict__ cnt_s, uint32_t* __restrict__ part) {
  __shared__ uint32_t hist[kPartW];
  if (threadIdx.x < kPartW) hist[threadIdx.x] = 0;
  __syncthreads();
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    int eq_i = 0, eq_p = 0;
    const int li = (i + 1 < n) ? lcp_pair(key, key2, sys, i, eq_i) : -1;
    const int lp = (i > 0) ? lcp_pair(key, key2, sys, i - 1, eq_p) : -1;
    lcp[i] = li;
    cnt_s[i] = 1u + (uint32_t)(li > lp ? li - lp : 0);
    for (int d = lp + 1; d <= li; ++d) atomicAdd(&hist[d], 1u);
    for (int d = li + 1; d <= lp; ++d) atomicAdd(&hist[kPartEnds + d], 1u);
    atomicMax(&hist[kPartMaxDepth], (uint32_t)(1 + (li > lp ? li : lp)));
    if (eq_i) atomicAdd(&hist[kPartEqual], 1u);
  }
  __syncthreads();
  if (threadIdx.x < kPartW) part[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = hist[threadIdx.x];
}

// Column `col` of the partial counts, one block per column: a per-depth
// column becomes its exclusive prefix over the blocks (in place) and its
// total is the node count of that depth; the last two columns reduce to the
// deepest leaf and the equal-key pairs.
__global__ __launch_bounds__(kBlock) void k_lcp_scan(uint32_t* __restrict__ part, int64_t rows, int has_key2,
                                                     DevScalars* sc) {
  __shared__ uint32_t scr[kBlock / kWave + 1];
  const int col = blockIdx.x;
  uint32_t* cp = part + (int64_t)col * rows;
  if (col < kPartMaxDepth) {
    // thread t takes rows [t*per, (t+1)*per) of the column: their sum, one
    // block scan of the sums, then the rows' running prefix
    const int64_t per = (rows + kBlock - 1) / kBlock;
    const int64_t r0 = (int64_t)threadIdx.x * per;
    const int64_t r1 = r0 + per < rows ? r0 + per : rows;
    uint32_t sum = 0;
    for (int64_t r = r0; r < r1; ++r) sum += cp[r];
    uint32_t tot;
    uint32_t run = block_excl_scan_u32<kBlock>(sum, scr, &tot);
    for (int64_t r = r0; r < r1; ++r) {
      const uint32_t v = cp[r];
      cp[r] = run;
      run += v;
    }
    if (threadIdx.x == 0 && col < kPartEnds) sc->level_count[col] = (int32_t)tot;
    return;
  }
  __shared__ uint32_t red[kBlock / kWave];
  const bool is_max = col == kPartMaxDepth;
  uint32_t v = 0;
  for (int64_t r = threadIdx.x; r < rows; r += kBlock) {
    const uint32_t x = cp[r];
    v = is_max ? (x > v ? x : v) : v + x;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const uint32_t o = __shfl_xor(v, off, kWave);
    v = is_max ? (o > v ? o : v) : v + o;
  }
  if (lane_id() == 0) red[threadIdx.x / kWave] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; ++w) v = is_max ? (red[w] > v ? red[w] : v) : v + red[w];
    if (is_max) sc->max_leaf_depth = (int32_t)v;
    else if (has_key2) sc->coincident = (int32_t)v;
    else sc->collisions = (int32_t)v;
  }
}

// Preorder index of each system's root (its first body starts at depth 0)
// and of the node after its tree.
__global__ __launch_bounds__(kBlock) void k_sys_roots(SysArrays sy, const uint32_t* __restrict__ pre_off,
                                                      const DevScalars* __restrict__ sc) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= sy.nsys) return;
  sy.sroot[s] = (int32_t)pre_off[sy.soff[s]];
  sy.send[s] = s + 1 < sy.nsys ? (int32_t)pre_off[sy.soff[s + 1]] : sc->total_nodes;
}

__global__ __launch_bounds__(kBlock) void k_emit(int64_t n, const int32_t* __restrict__ lcp,
                                                 const uint32_t* __restrict__ pre_off,
                                                 const uint32_t* __restrict__ part, StateArrays st, SysArrays sy,
                                                 const DevScalars* __restrict__ sc, RefNode* __restrict__ nodes,
                                                 WalkHead* __restrict__ heads, int32_t* __restrict__ level_list,
                                                 uint32_t* __restrict__ end_list, uint32_t cap_nodes,
                                                 uint32_t cap_int) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  int li = -1, lp = -1;
  uint32_t pre = 0, skipv = 0;
  if (valid) {
    li = lcp[i];
    lp = i > 0 ? lcp[i - 1] : -1;
    pre = pre_off[i];
    skipv = (i + 1 < n) ? pre_off[i + 1] : (uint32_t)sc->total_nodes;
  }
  place_internal_nodes(valid, lp, li, pre, skipv, part, sc, level_list, end_list,
                       cap_int);
  if (!valid) return;
  // One all-zero record after the last node: the walk's evaluation reads it
  // for a lane with no entry left and for a leaf's quadrupole.
  if (i == n - 1) {
    const uint32_t z = (uint32_t)sc->total_nodes;
    if (z < cap_nodes) nodes[z] = RefNode{};
  }
  // Internal nodes are written by the upward pass (their records, skip
  // pointers included, come from the node lists placed above).
  const int ns = li > lp ? li - lp : 0;
  if (pre + ns < cap_nodes) {
    const uint32_t p = pre + ns;
    const int d = 1 + (li > lp ? li : lp);
    const double m = st.m[i];
    const double size = ldexp(sy.root[4 * st.sys[i] + 3], -d);
    // Leaves carry no quadrupole; the walk never reads a leaf's Q fields
    // (the first 80 bytes of the record are written, in one piece).
    const double px = st.x[i], py = st.y[i], pz = st.z[i];
    RefNode& nd = nodes[p];
    nd.cx = px; nd.cy = py; nd.cz = pz;
    nd.m = m;
    nd.skip = (int32_t)(p + 1);
    nd.body = (int32_t)i;
    nd.depth = d;
    nd.first = (int32_t)i;
    nd.size = size;
    // softening.h:26-27: size*0.015 + cbrt(m)*0.002, summed before the taper
    nd.esum = size * 0.015 + cbrt(m) * 0.002;
    // a leaf is accepted whatever its distance: no threshold
    heads[p] = WalkHead{px, py, pz, 0.0f, m == 0.0 ? (int32_t)(p + 1) | kLinkZeroMass : ~(int32_t)i};
  }
}

// A build whose counts stay on the device: each depth's offset in the node
// list, and the abort flag if the tree does not fit the buffers and launches
// sized beforehand (node records, node-list entries, depth levels), needs
// the two-word keys, or has a non-finite position.  One wave.
__global__ __launch_bounds__(kWave) void k_check(DevScalars* sc, int64_t cap_records, int64_t cap_list, int dmax) {
  const int t = threadIdx.x;
  const uint32_t v = (uint32_t)sc->level_count[t];
  const uint32_t inc = wave_incl_scan_u32(v);
  sc->level_off[t] = (int32_t)(inc - v);
  if (t == kWave - 1) {
    int32_t f = 0;
    if ((int64_t)(uint32_t)sc->total_nodes + 1 > cap_records || (int64_t)inc > cap_list) f |= kAbortCapacity;
    if (sc->max_leaf_depth - 1 > dmax) f |= kAbortDepth;
    if (sc->collisions > 0 || sc->coincident > 0) f |= kAbortCollision;
    if (sc->err & 1) f |= kAbortNonFinite;
    if (sc->err & kSortErrTimeout) f |= kAbortSort;
    if (sc->err & kSortErrPatch) f |= kAbortPatch;
    if (f) sc->abort_flags |= f;
  }
}

// Mass, centre of mass and one-level softened quadrupole of an internal node
// of depth d from its children (octree.h:104-126), and the squared-distance
// thresholds of the walk's opening test for this theta: the node's record
// (all of it but the padding) and its walk head.  end: the preorder index after its
// subtree (its skip pointer, from the list of skip targets); the first body
// comes from the first child, the half-width from the system's root.
__device__ __forceinline__ void upward_node(int32_t id, int32_t end, int d, RefNode* __restrict__ nodes,
                                            WalkHead* __restrict__ heads, double inv_theta,
                                            const int32_t* __restrict__ sys, const double* __restrict__ root) {
  // One walk of the child chain (at most 8; the bound also keeps a malformed
  // chain finite) gathers each child's mass and centre into registers; a
  // slot past the last child keeps mass 0, which both sums skip.
  double cmv[8], cxv[8], cyv[8], czv[8];
  int32_t c = id + 1, first = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    cmv[j] = 0.0; cxv[j] = 0.0; cyv[j] = 0.0; czv[j] = 0.0;
    if (c < end) {
      const RefNode& ch = nodes[c];
      cmv[j] = ch.m; cxv[j] = ch.cx; cyv[j] = ch.cy; czv[j] = ch.cz;
      if (j == 0) first = ch.first;
      c = ch.skip;
    }
  }
  double M = 0.0, sx = 0.0, sy = 0.0, sz = 0.0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const double cm = cmv[j];
    if (cm == 0.0) continue;
    M = M + cm;
    sx = sx + cxv[j] * cm;
    sy = sy + cyv[j] * cm;
    sz = sz + czv[j] * cm;
  }
  double cx = sx, cy = sy, cz = sz;
  if (M > 0.0) { cx = sx / M; cy = sy / M; cz = sz / M; }
  const double size = ldexp(root[4 * sys[first] + 3], -d);
  double qxx = 0, qyy = 0, qzz = 0, qxy = 0, qxz = 0, qyz = 0;
  const double soft = size * size * 0.01;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const double mc = cmv[j];
    if (mc == 0.0) continue;
    const double rx = cxv[j] - cx, ry = cyv[j] - cy, rz = czv[j] - cz;
    const double r2 = rx * rx + ry * ry + rz * rz + soft;
    qxx = qxx + mc * (3.0 * rx * rx - r2);
    qyy = qyy + mc * (3.0 * ry * ry - r2);
    qzz = qzz + mc * (3.0 * rz * rz - r2);
    qxy = qxy + mc * (3.0 * rx * ry);
    qxz = qxz + mc * (3.0 * rx * rz);
    qyz = qyz + mc * (3.0 * ry * rz);
  }
  RefNode p;
  p.cx = cx; p.cy = cy; p.cz = cz; p.m = M;
  p.skip = end;
  p.body = -1;
  p.depth = d;
  p.first = first;
  p.size = size;
  p.esum = size * 0.015 + cbrt(M) * 0.002;
  p.qxx = qxx; p.qyy = qyy; p.qzz = qzz; p.qxy = qxy; p.qxz = qxz; p.qyz = qyz;
  RefNode& out = nodes[id];                  // the record but its padding
  out.cx = p.cx; out.cy = p.cy; out.cz = p.cz; out.m = p.m;
  out.skip = p.skip; out.body = p.body; out.depth = p.depth; out.first = p.first;
  out.size = p.size; out.esum = p.esum;
  out.qxx = p.qxx; out.qyy = p.qyy; out.qzz = p.qzz; out.qxy = p.qxy; out.qxz = p.qxz; out.qyz = p.qyz;
  // Upper edge of the opening test's band for r^2 (NEXT tests r^2 + 1e-20,
  // so the 1e-20 is taken off the edge); NaN, and every decision exact, where
  // (size/theta)^2 <= 1e-16, where the 1e-20 would eat the band's margin.
  const double q = size * inv_theta;
  const double q2 = q * q;
  const float acc2 =
      (q2 > 1e-16 && q2 < 1e36) ? __double2float_ru(q2 * (1.0 + 2e-6) - 1e-20) : __builtin_nanf("");
  heads[id] = WalkHead{cx, cy, cz, acc2, M == 0.0 ? (end | kLinkZeroMass) : end};
}

// The internal nodes [beg, beg + cnt) of the node list, all of depth d.
__global__ __launch_bounds__(kBlock) void k_upward(const int32_t* __restrict__ level_list,
                                                   const uint32_t* __restrict__ end_list, int64_t beg, int64_t cnt,
                                                   int d, RefNode* __restrict__ nodes, WalkHead* __restrict__ heads,
                                                   double inv_theta, const int32_t* __restrict__ sys,
                                                   const double* __restrict__ root) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= cnt) return;
  upward_node(level_list[beg + k], (int32_t)end_list[beg + k], d, nodes, heads, inv_theta, sys, root);
}

// Depths d_lo .. 41 of a build whose counts stay on the device, deepest
// first, in one workgroup (a barrier between depths: the waves share the
// CU's cache, so a depth's results are visible to the next): the deep end
// of a tree holds few nodes, chains mostly, and its depth varies from build
// to build.
__global__ __launch_bounds__(kBlock) void k_upward_tail(int d_lo, const int32_t* __restrict__ level_list,
                                                        const uint32_t* __restrict__ end_list,
                                                        RefNode* __restrict__ nodes, WalkHead* __restrict__ heads,
                                                        double inv_theta, const int32_t* __restrict__ sys,
                                                        const double* __restrict__ root,
                                                        const DevScalars* __restrict__ sc) {
  if (sc->abort_flags) return;
  for (int d = kMaxLevels - 1; d >= d_lo; --d) {
    const int64_t beg = sc->level_off[d], cnt = sc->level_count[d];
    if (cnt == 0) continue;
    for (int64_t k = threadIdx.x; k < cnt; k += kBlock)
      upward_node(level_list[beg + k], (int32_t)end_list[beg + k], d, nodes, heads, inv_theta, sys, root);
    __syncthreads();
  }
}

// The same for depth d of a build whose counts stay on the device: the
// depth's extent in the node list comes from the scalars (k_check), the grid
// is fixed and strides over it.
__global__ __launch_bounds__(kBlock) void k_upward_level(int d, const int32_t* __restrict__ level_list,
                                                         const uint32_t* __restrict__ end_list,
                                                         RefNode* __restrict__ nodes, WalkHead* __restrict__ heads,
                                                         double inv_theta, const int32_t* __restrict__ sys,
                                                         const double* __restrict__ root,
                                                         const DevScalars* __restrict__ sc) {
  if (sc->abort_flags) return;
  const int64_t beg = sc->level_off[d], cnt = sc->level_count[d];
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < cnt; k += (int64_t)gridDim.x * kBlock)
    upward_node(level_list[beg + k], (int32_t)end_list[beg + k], d, nodes, heads, inv_theta, sys, root);
}

// Blocks are dealt to the 8 XCDs round-robin; remap so that each XCD walks
// one contiguous (spatially coherent) range of targets and its L2 holds
// the part of the tree those targets touch.
__device__ __forceinline__ int64_t xcd_block(int nb) {
  const int per = (nb + 7) / 8;
  return (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
}

// Reciprocal square root and reciprocal with a relative error of a few
// units in the last place: the v_rsq_f64 / v_rcp_f64 estimates refined by
// one Newton step each (the residual formed with an FMA).  The walk's
// arguments are positive, finite and mid-range: r^2 + 1e-20, r^2 + eps^2
// with eps >= 1e-4, and 1 + 10 dist >= 1.
__device__ __forceinline__ double rsq_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-(x * y), y, 1.0);        // 1 - x y^2
  return __builtin_fma(0.5 * y, e, y);
}

__device__ __forceinline__ double rcp_nr(double x) {
  const double y = __builtin_amdgcn_rcp(x);
  const double e = __builtin_fma(-x, y, 1.0);              // 1 - x y
  return __builtin_fma(y, e, y);
}

// Softening floor of a dark-matter target (octree.h:146-148), out of line:
// expanded in place, the FP64 pow raises the walk's register count by about
// half for every target.
__device__ __noinline__ double dm_softening(double eps, double size, double m, double mi) {
  const double alt = 2.0 * size / pow(m / mi, 0.333333333);
  return eps < alt ? alt : eps;
}

// k < 2^25, checked on the host: the byte offset fits 32 bits and becomes the
// scalar load's offset register.
// The per-visit fields of a node as the traversal holds them: the head,
// loaded as one 8-dword value (one scalar load), with the FP32 threshold
// kept as its bits until it is used.  (Read as a float straight from the
// loaded vector, the threshold was folded to element 0 instead of 6 by the
// ROCm 7.2 compiler, which wrecked every opening decision; the integer
// extraction compiles correctly.)
typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
struct HeadRegs {
  double cx, cy, cz;
  uint32_t acc2_bits;
  int32_t link;
};
__device__ __forceinline__ HeadRegs load_head(const WalkHead* __restrict__ heads, uint32_t k) {
  const u32x8 r = *reinterpret_cast<const u32x8*>(reinterpret_cast<const char*>(heads) + (k << 5));
  HeadRegs h;
  h.cx = __builtin_bit_cast(double, (uint64_t)r[0] | ((uint64_t)r[1] << 32));
  h.cy = __builtin_bit_cast(double, (uint64_t)r[2] | ((uint64_t)r[3] << 32));
  h.cz = __builtin_bit_cast(double, (uint64_t)r[4] | ((uint64_t)r[5] << 32));
  h.acc2_bits = r[6];
  h.link = (int32_t)r[7];
  return h;
}

// Lower edge of the opening test's band from its upper edge acc2 (FP32,
// (size/theta)^2 (1 + 2e-6) - 1e-20 rounded up, for (size/theta)^2 > 1e-16):
// r^2 below it has r^2 + 1e-20 below (size/theta)^2 (1 - 2e-6).  Squared
// distances between the edges are decided by NEXT's expression.
constexpr double kBandLow = (1.0 - 2e-6) / (1.0 + 2.2e-6);

// Accepted nodes a lane holds (a ring in LDS), and how many of each lane's
// oldest entries one evaluation round takes.  Evaluating only the oldest
// entries when some lane's ring is full keeps more lanes busy per
// evaluation step than draining every ring: lanes fill at different rates.
// The walk is bound by the latency of its dependent node loads, so resident
// waves count: 24 entries (6 KB of LDS per wave) with the registers held to
// 80 (a few spills outside the loops) give 6 waves per SIMD, 6 % faster
// than 32 entries at 5; 20 entries at 7 waves (17 spills) were slower, and so
// were 16 (fewer lanes busy per evaluation step) and 64 (2.5 waves).
constexpr int kWalkList = 24;
// One wave per workgroup: a wave's registers and LDS are released as soon as
// its group is done instead of when the slowest wave of a larger workgroup
// is (measured 2.39 -> 2.34 ms against four waves per workgroup).
constexpr int kWalkBlock = 64;
constexpr int kWalkFlush = 4;
static_assert(kWalkFlush <= kWalkList, "walk ring");

// Node records come in through the scalar cache: the traversal loads one
// 32-byte WalkHead per visited node, the evaluation a lane's accepted
// records by vector loads.  Loading the next entry's record while the
// current one is evaluated measured slower (it costs the registers of a
// sixth wave).
//
// Two phases per wave.  Traversal: every lane follows its own depth-first
// path (open: go to the first child, cur+1; accept or skip: go past the
// subtree, skip[cur]).  The wave visits node `cur` and the lanes whose own
// next node is `cur` decide it.  No lane that is waiting can point inside
// cur's subtree (its pointer came from a skip of an ancestor or an earlier
// sibling, both of which end at or after skip[cur]), so the wave's next node
// is cur+1 when any lane opened cur and skip[cur] otherwise: one ballot, no
// reduction.  An accepted node is only appended to the lane's ring in LDS.
// Evaluation: when some lane's ring is full, the wave evaluates up to
// kWalkFlush of every lane's oldest entries in step, entry k of every lane
// at once (and everything at the end), so the costly interaction (two square
// roots, two divides, the quadrupole) runs with most lanes busy instead of
// only the lanes that accepted the node being visited.  Each lane still adds
// its interactions in its own depth-first order, the order of
// octree.h:158-177, so the sums are NEXT's bit for bit.
//
// Opening test of octree.h:153, (size/dist) < theta with
// dist = sqrt(r2 + 1e-20): decided from r2 against the node's FP32 band
// edges around (size/theta)^2 - 1e-20 (relative width about 4e-6; rounding
// errors of either form are below 1e-15), and by NEXT's exact expression
// inside the band, so the decision is NEXT's bit for bit and opened nodes
// cost no sqrt or divide.
template <bool KICK, bool COUNT>
__global__ __launch_bounds__(kWalkBlock) __attribute__((amdgpu_waves_per_eu(6)))
void k_walk(const Group* __restrict__ groups, int64_t ngroups,
                                                 const RefNode* __restrict__ nodes,
                                                 const WalkHead* __restrict__ heads, StateArrays st, SysArrays sy,
                                                 double theta, double* __restrict__ ax, double* __restrict__ ay,
                                                 double* __restrict__ az, DevScalars* sc) {
  __shared__ int32_t lists[kWalkBlock / kWave][kWalkList][kWave];
  // kWalkBlock / 64 waves per block, one group per wave.  The wave index goes through
  // readfirstlane so the compiler sees it as wave-uniform: the group, its
  // system's tree bounds and every node visited are then uniform too, and
  // the node records come in through scalar loads into scalar registers.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
  const int64_t gi = xcd_block(gridDim.x) * (kWalkBlock / kWave) + wave;
  if (gi >= ngroups || sc->abort_flags) return;            // whole wave past the end, or an aborted build
  int32_t* const mine = &lists[wave][0][lane_id()];      // ring slot k at mine[k * kWave]
  const Group g = groups[gi];
  const int64_t i = g.first + lane_id();
  const bool valid = lane_id() < g.count;
  const int32_t total = sy.send[g.sys];                    // this system's tree: [sroot, send)
  double px = 0, py = 0, pz = 0, pm = 1;
  int pt = 0;
  if (valid) { px = st.x[i]; py = st.y[i]; pz = st.z[i]; pm = st.m[i]; pt = st.type[i]; }
  const int32_t self_link = valid ? ~(int32_t)i : INT_MAX;   // link of this lane's own leaf
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  uint32_t npp = 0, npc = 0, nop = 0, nit = 0, nev = 0;
  int ws = 0, len = 0;                                     // this lane's ring: next free slot, entries before it
  const uint32_t zoff = (uint32_t)sc->total_nodes << 7;   // the all-zero record (k_emit)

  // Evaluate up to `limit` of the oldest entries (octree.h:139-176 for an
  // accepted node); an entry is the node index, with bit 31 set for a leaf
  // (Q = 0).  DM: some target of the wave is dark matter (its softening
  // floor, octree.h:146-148, is a call); without one the loop is straight
  // code.
  auto evaluate = [&](auto dm, int limit) {
    const int take = len < limit ? len : limit;
    int rs = ws - len;                                     // the oldest entry's slot
    if (rs < 0) rs += kWalkList;
    for (int k = 0; __builtin_amdgcn_ballot_w64(k < take); ++k) {
      if (COUNT) ++nev;
      // Every lane runs every round: a lane with no entry left reads the
      // all-zero record after the last node, and so does a leaf entry for
      // its quadrupole (leaf records carry none); both add exact zeros
      // (m = 0, Q = 0, every factor finite).  Byte offsets: e << 7 drops
      // the leaf bit.
      const int32_t e = mine[rs * kWave];
      rs = rs == kWalkList - 1 ? 0 : rs + 1;
      const bool on = k < take;
      const uint32_t off = on ? (uint32_t)e << 7 : zoff;
      const uint32_t qoff = (on && e >= 0) ? off : zoff;
      const RefNode* nd = reinterpret_cast<const RefNode*>(reinterpret_cast<const char*>(nodes) + off);
      const RefNode* qn = reinterpret_cast<const RefNode*>(reinterpret_cast<const char*>(nodes) + qoff);
      const double dx = nd->cx - px, dy = nd->cy - py, dz = nd->cz - pz;
      const double m = nd->m;
      // octree.h:139-172 with products fused into FMAs and the square
      // roots and divides faithfully rounded (a few ulp per interaction):
      // r2 = |d|^2, dist = sqrt(r2 + 1e-20), eps from softening.h:26-36,
      // dinv = 1/sqrt(r2 + eps^2).
      const double r2 = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
      const double r2e = r2 + 1e-20;
      const double dist = r2e * rsq_nr(r2e);
      double eps = nd->esum * rcp_nr(__builtin_fma(dist, 10.0, 1.0));
      if (eps < 1e-4) eps = 1e-4;
      if (decltype(dm)::value && pt == 1) eps = dm_softening(eps, nd->size, m, pm);   // octree.h:146-148
      const double dinv = rsq_nr(__builtin_fma(eps, eps, r2));
      const double inv2 = dinv * dinv;
      const double inv3 = inv2 * dinv;
      const double inv5 = inv3 * inv2;
      const double inv7 = inv5 * inv2;
      // quadrupole term 0.5 (2 (Q d) inv5 - 5 q inv7 d), q = d.Q.d
      // (octree.h:159-172), as (Q d) inv5 + d (-(2.5 q) inv7)
      const double qxx = qn->qxx, qyy = qn->qyy, qzz = qn->qzz;
      const double qxy = qn->qxy, qxz = qn->qxz, qyz = qn->qyz;
      const double sx = __builtin_fma(qxz, dz, __builtin_fma(qxy, dy, qxx * dx));
      const double sy = __builtin_fma(qyz, dz, __builtin_fma(qyy, dy, qxy * dx));
      const double sz = __builtin_fma(qzz, dz, __builtin_fma(qyz, dy, qxz * dx));
      const double q = __builtin_fma(dz, sz, __builtin_fma(dy, sy, dx * sx));
      const double c = __builtin_fma(-2.5 * q, inv7, m * inv3);   // G = 1
      a0 = __builtin_fma(dx, c, __builtin_fma(sx, inv5, a0));
      a1 = __builtin_fma(dy, c, __builtin_fma(sy, inv5, a1));
      a2 = __builtin_fma(dz, c, __builtin_fma(sz, inv5, a2));
    }
    len -= take;
  };
  // dark-matter targets in this wave (wave-uniform)
  const bool any_dm = __builtin_amdgcn_ballot_w64(pt == 1) != 0;

  const int32_t root = sy.sroot[g.sys];
  int32_t next = valid ? root : INT_MAX;
  int32_t cur = root;
  HeadRegs h = load_head(heads, root);
  // Traversal runs until some lane's ring is full (or the tree is done),
  // then one evaluation round; the running sums are not touched inside the
  // traversal loop, which keeps them out of its loop-carried registers.
  for (;;) {
    while (cur < total) {
      // successors, loaded while this node is processed
      // (the node array has one spare record after the last node, so both
      // successors, at most total, are readable)
      const bool leaf = h.link < 0;
      const int32_t lk = leaf ? cur + 1 : (h.link & ~kLinkZeroMass);
      const int32_t sk = lk > cur ? lk : cur + 1;           // skip > cur by construction
      const HeadRegs h1 = load_head(heads, (uint32_t)cur + 1u);
      const HeadRegs h2 = load_head(heads, (uint32_t)sk);
      // This lane decides cur if its own path is at cur.  The decision is
      // formed without branching on it: every lane computes the test, and
      // every lane writes cur into its ring's free slot ws (each ring holds
      // fewer than kWalkList entries inside this loop), which becomes an
      // entry only if the lane accepts.
      const bool here = next == cur;
      const bool pass = (h.link >= 0 && (h.link & kLinkZeroMass)) || h.link == self_link;   // octree.h:134-135
      bool accept = leaf;
      if (!leaf) {                                            // wave-uniform
        const double dx = h.cx - px, dy = h.cy - py, dz = h.cz - pz;
        // Against the band edges (for r^2, about 4e-6 wide) any rounding of
        // r^2 decides alike: FMAs here, NEXT's expression, as NEXT rounds
        // it, inside the band.
        const double r2f = __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
        const double hi = (double)__uint_as_float(h.acc2_bits);   // NaN: no fast decision
        accept = r2f > hi;
        const bool band = here && !accept && !(r2f < hi * kBandLow);
        if (band) {                                           // rare: NEXT's exact expression
          const double r2e = (dx * dx + dy * dy + dz * dz) + 1e-20;
          accept = (nodes[cur].size / sqrt(r2e)) < theta;
        }
      }
      const bool take = here && !pass && accept;
      const bool open = here && !pass && !accept;
      mine[ws * kWave] = leaf ? (int32_t)(cur | 0x80000000u) : cur;
      const int tk = take ? 1 : 0;
      len += tk;
      ws += tk;
      if (ws == kWalkList) ws = 0;
      // an opening lane is here (next == cur): its next node is next + 1
      const int op = open ? 1 : 0;
      next = (here && !open) ? sk : next + op;
      if (COUNT) {
        npp += (take && leaf) ? 1u : 0u;
        npc += (take && !leaf) ? 1u : 0u;
        nop += open ? 1u : 0u;
      }
      if (COUNT) ++nit;
      if (__builtin_amdgcn_ballot_w64(op != 0)) { cur = cur + 1; h = h1; }
      else { cur = sk; h = h2; }
      if (__builtin_amdgcn_ballot_w64(len == kWalkList)) break;
    }
    if (cur >= total) break;
    if (any_dm) evaluate(std::true_type{}, kWalkFlush);
    else evaluate(std::false_type{}, kWalkFlush);
  }
  if (any_dm) evaluate(std::true_type{}, kWalkList);
  else evaluate(std::false_type{}, kWalkList);
  (void)h;
  double v2 = 0.0;
  if (valid) {
    ax[i] = a0; ay[i] = a1; az[i] = a2;
    if (KICK) {                                             // step.h:171-173
      const double half = sy.dt[g.sys] * 0.5;
      const double vx = st.vx[i] + a0 * half;
      const double vy = st.vy[i] + a1 * half;
      const double vz = st.vz[i] + a2 * half;
      st.vx[i] = vx; st.vy[i] = vy; st.vz[i] = vz;
      v2 = vx * vx + vy * vy + vz * vz;
      if (!(v2 > 0.0)) v2 = 0.0;
    }
  }
  if (KICK) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
      const double o = __shfl_xor(v2, off, kWave);
      if (o > v2) v2 = o;
    }
    if (lane_id() == 0 && v2 > 0.0) atomicMax(&sy.v2max[g.sys], (unsigned long long)__double_as_longlong(v2));
  }
  if (COUNT) {
    const uint32_t sp = wave_sum_u32(npp), sc2 = wave_sum_u32(npc), so = wave_sum_u32(nop);
    if (lane_id() == 0) {
      atomicAdd(&sc->n_pp, (unsigned long long)sp);
      atomicAdd(&sc->n_pc, (unsigned long long)sc2);
      atomicAdd(&sc->n_open, (unsigned long long)so);
      atomicAdd(&sc->n_iter, (unsigned long long)nit);
      atomicAdd(&sc->n_eval, (unsigned long long)nev);
    }
  }
}

// Opening half kick and drift (step.h:120-122, 141-146), each system with its dt.
// amag (may be null) receives |a| of each body, in FP32, for an opening
// criterion relative to the previous acceleration.
__global__ __launch_bounds__(kBlock) void k_kick_drift(int64_t n, StateArrays st, const double* __restrict__ ax,
                                                       const double* __restrict__ ay, const double* __restrict__ az,
                                                       SysArrays sy, float* __restrict__ amag) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const double dt = sy.dt[st.sys[i]];
  const double half = dt * 0.5;
  const double gx = ax[i], gy = ay[i], gz = az[i];
  const double vx = st.vx[i] + gx * half;
  const double vy = st.vy[i] + gy * half;
  const double vz = st.vz[i] + gz * half;
  st.vx[i] = vx; st.vy[i] = vy; st.vz[i] = vz;
  st.x[i] = st.x[i] + vx * dt;
  st.y[i] = st.y[i] + vy * dt;
  st.z[i] = st.z[i] + vz * dt;
  if (amag) amag[i] = (float)sqrt(gx * gx + gy * gy + gz * gz);
}

// Max |v|^2 per system (adaptive.h:31-37).  Each thread takes kMaxv2Items
// bodies (block-strided), a block whose bodies all belong to one system
// reduces in registers and LDS and makes one atomic; otherwise every body
// makes its own.
constexpr int kMaxv2Items = 4;
__global__ __launch_bounds__(kBlock) void k_maxv2(int64_t n, StateArrays st, SysArrays sy) {
  __shared__ double red[kBlock / kWave];
  const int64_t base = (int64_t)blockIdx.x * kBlock * kMaxv2Items;
  const int64_t last = base + (int64_t)kBlock * kMaxv2Items - 1 < n ? base + (int64_t)kBlock * kMaxv2Items - 1 : n - 1;
  const int32_t s0 = st.sys[base];
  const bool one = st.sys[last] == s0;                    // sorted by system within the block's range
  double v2 = 0.0;
  for (int k = 0; k < kMaxv2Items; ++k) {
    const int64_t i = base + (int64_t)k * kBlock + threadIdx.x;
    if (i >= n) break;
    double w = st.vx[i] * st.vx[i] + st.vy[i] * st.vy[i] + st.vz[i] * st.vz[i];
    if (!(w > 0.0)) w = 0.0;
    if (one) {
      if (w > v2) v2 = w;
    } else if (w > 0.0) {
      atomicMax(&sy.v2max[st.sys[i]], (unsigned long long)__double_as_longlong(w));
    }
  }
  if (!one) return;                                        // block-uniform
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    const double o = __shfl_xor(v2, off, kWave);
    if (o > v2) v2 = o;
  }
  if (lane_id() == 0) red[threadIdx.x / kWave] = v2;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kBlock / kWave; ++w) v2 = red[w] > v2 ? red[w] : v2;
    if (v2 > 0.0) atomicMax(&sy.v2max[s0], (unsigned long long)__double_as_longlong(v2));
  }
}

// computeAdaptiveDt (adaptive.h:38-51) for each system from its max |v|^2;
// dts (may be null) receives system 0's value for `step`.
__global__ __launch_bounds__(kBlock) void k_dt(SysArrays sy, double base, double* dts, int64_t step,
                                               DevScalars* sc) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= sy.nsys) return;
  const double v2 = __longlong_as_double((long long)sy.v2max[s]);
  sy.v2max[s] = 0ull;   // read once: the next kicks gather the next step's maximum
  double v = sqrt(v2);
  if (1e4 < v) v = 1e4;
  double dt;
  if (v < 1e-8) {
    dt = base;
  } else {
    dt = base / (1.0 + v);
    if (dt < base * 0.01) dt = base * 0.01;
    if (base * 1.0 < dt) dt = base * 1.0;
  }
  sy.dt[s] = dt;
  if (dts && s == 0) {
    // step < 0: the entry comes from the device's step counter (captured runs)
    const int64_t at = step >= 0 ? step : (int64_t)(sc->step_count++);
    dts[at] = dt;
  }
}

__device__ __forceinline__ uint32_t digit_at(const uint64_t* key, const uint64_t* key2, int32_t f, int l) {
  if (l < kKeyDigits) return (uint32_t)(key[f] >> (3 * (kKeyDigits - 1 - l))) & 7u;
  return key2 ? (uint32_t)(key2[f] >> (3 * (kMaxLevels - 1 - l))) & 7u : 0u;
}

__global__ __launch_bounds__(kBlock) void k_export(int64_t total, const RefNode* __restrict__ nodes,
                                                   const uint64_t* __restrict__ key, const uint64_t* __restrict__ key2,
                                                   const int32_t* __restrict__ id, SysArrays sy,
                                                   int32_t* depth, int32_t* leaf, int32_t* body, int32_t* cmask,
                                                   double* geo4, double* mass, double* com3, double* q6) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= total) return;
  const RefNode nd = nodes[k];
  depth[k] = nd.depth;
  leaf[k] = nd.body >= 0 ? 1 : 0;
  body[k] = nd.body >= 0 ? id[nd.body] : -1;
  double cx = sy.root[0], cy = sy.root[1], cz = sy.root[2], s = sy.root[3];
  for (int l = 0; l < nd.depth; ++l) {
    const uint32_t o = digit_at(key, key2, nd.first, l);
    const double hs = s * 0.5;
    cx = cx + ((o & 1u) ? hs : -hs);
    cy = cy + ((o & 2u) ? hs : -hs);
    cz = cz + ((o & 4u) ? hs : -hs);
    s = hs;
  }
  geo4[4 * k] = cx; geo4[4 * k + 1] = cy; geo4[4 * k + 2] = cz; geo4[4 * k + 3] = s;
  int32_t mask = 0;
  if (nd.body < 0)
    for (int32_t c = (int32_t)k + 1, it = 0; it < 8 && c < nd.skip; ++it, c = nodes[c].skip)
      mask |= 1 << digit_at(key, key2, nodes[c].first, nd.depth);
  cmask[k] = mask;
  mass[k] = nd.m;
  com3[3 * k] = nd.cx; com3[3 * k + 1] = nd.cy; com3[3 * k + 2] = nd.cz;
  const bool lf = nd.body >= 0;
  q6[6 * k] = lf ? 0.0 : nd.qxx; q6[6 * k + 1] = lf ? 0.0 : nd.qyy; q6[6 * k + 2] = lf ? 0.0 : nd.qzz;
  q6[6 * k + 3] = lf ? 0.0 : nd.qxy; q6[6 * k + 4] = lf ? 0.0 : nd.qxz; q6[6 * k + 5] = lf ? 0.0 : nd.qyz;
}

inline unsigned grid_of(int64_t n) { return (unsigned)div_up(n, kBlock); }

}  // namespace

inline unsigned sys_grid(const SysArrays& sy) { return (unsigned)div_up(std::max(sy.nsys, 1), kBlock); }

void launch_fill_scalars(DevScalars* sc, int what, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_scalars, dim3(1), dim3(kBlock), 0, s, sc, what);
}
void launch_fill_v2max(const SysArrays& sy, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_v2max, dim3(sys_grid(sy)), dim3(kBlock), 0, s, sy);
}
void launch_set_dt(const SysArrays& sy, double dt, hipStream_t s) {
  hipLaunchKernelGGL(k_set_dt, dim3(sys_grid(sy)), dim3(kBlock), 0, s, sy, dt);
}
void launch_bbox(const StateArrays& st, const BoxBlock* blocks, int nblocks, const int32_t* block_off, double* part,
                 const SysArrays& sy, DevScalars* clear, hipStream_t s) {
  if (nblocks > 0) hipLaunchKernelGGL(k_bbox_partial, dim3(nblocks), dim3(kBlock), 0, s, st, blocks, part);
  hipLaunchKernelGGL(k_bbox_final, dim3((unsigned)std::max(sy.nsys, 1)), dim3(kWave), 0, s, part, block_off, sy,
                     clear);
}
void launch_keys(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* key, uint32_t* idx,
                 DevScalars* scw, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_keys, dim3(grid_of(n)), dim3(kBlock), 0, s, n, st, sy, key, idx, scw);
}
void launch_keys2(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* key2, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_keys2, dim3(grid_of(n)), dim3(kBlock), 0, s, n, st, sy, key2);
}
void launch_keys_export(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* hi, uint64_t* lo,
                        hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_keys_export, dim3(grid_of(n)), dim3(kBlock), 0, s, n, st, sy, hi, lo);
}
void launch_fix_runs(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* key, uint64_t* key2,
                     uint32_t* amag, int lo, hipStream_t s) {
  if (n > 1) hipLaunchKernelGGL(k_fix_runs, dim3(grid_of(n)), dim3(kBlock), 0, s, n, st, sy, key, key2, amag, lo);
}
void launch_iota(int64_t n, uint32_t* idx, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_iota, dim3(grid_of(n)), dim3(kBlock), 0, s, n, idx);
}
void launch_gather_u64(int64_t n, const uint32_t* perm, const uint64_t* src, uint64_t* dst, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_u64, dim3(grid_of(n)), dim3(kBlock), 0, s, n, perm, src, dst);
}
void launch_gather_sys(int64_t n, const uint32_t* perm, const int32_t* sys, uint32_t* dst, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_sys, dim3(grid_of(n)), dim3(kBlock), 0, s, n, perm, sys, dst);
}
void launch_gather_u32(int64_t n, const uint32_t* perm, const uint32_t* src, uint32_t* dst, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_gather_sys, dim3(grid_of(n)), dim3(kBlock), 0, s, n, perm,
                       reinterpret_cast<const int32_t*>(src), dst);
}
void launch_gather_state(int64_t n, const uint32_t* perm, const StateArrays& src, const StateArrays& dst,
                         hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_state, dim3(grid_of(n)), dim3(kBlock), 0, s, n, perm, src, dst);
}
void launch_lcp_count(int64_t n, const uint64_t* key, const uint64_t* key2, const int32_t* sys, int32_t* lcp,
                      uint32_t* cnt_s, uint32_t* part, DevScalars* sc, hipStream_t s) {
  if (n <= 0) return;
  const int64_t rows = grid_of(n);
  hipLaunchKernelGGL(k_lcp_count, dim3((unsigned)rows), dim3(kBlock), 0, s, n, key, key2, sys, lcp, cnt_s, part);
  hipLaunchKernelGGL(k_lcp_scan, dim3(kPartW), dim3(kBlock), 0, s, part, rows, key2 ? 1 : 0, sc);
}
void launch_remove_map(int64_t n_new, int64_t cut, int64_t len, uint32_t* perm, hipStream_t s) {
  if (n_new > 0) hipLaunchKernelGGL(k_remove_map, dim3(grid_of(n_new)), dim3(kBlock), 0, s, n_new, cut, len, perm);
}
void launch_relabel(int64_t n, int32_t* sys, int32_t removed, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_relabel, dim3(grid_of(n)), dim3(kBlock), 0, s, n, sys, removed);
}
void launch_gather_f64(int64_t n, const uint32_t* perm, const double* src, double* dst, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_gather_f64, dim3(grid_of(n)), dim3(kBlock), 0, s, n, perm, src, dst);
}
void launch_sys_roots(const SysArrays& sy, const uint32_t* pre_off, const DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_sys_roots, dim3(sys_grid(sy)), dim3(kBlock), 0, s, sy, pre_off, sc);
}
void launch_emit(int64_t n, const int32_t* lcp, const uint32_t* pre_off, const uint32_t* part,
                 const StateArrays& st, const SysArrays& sy, const DevScalars* sc, RefNode* nodes, WalkHead* heads,
                 int32_t* level_list, uint32_t* end_list, uint32_t cap_nodes, uint32_t cap_int, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_emit, dim3(grid_of(n)), dim3(kBlock), 0, s, n, lcp, pre_off, part, st, sy, sc, nodes, heads,
                       level_list, end_list, cap_nodes, cap_int);
}
void launch_set_run(DevScalars* sc, int32_t step, hipStream_t s) {
  hipLaunchKernelGGL(k_set_run, dim3(1), dim3(1), 0, s, sc, step);
}
void launch_check(DevScalars* sc, int64_t cap_records, int64_t cap_list, int dmax, hipStream_t s) {
  hipLaunchKernelGGL(k_check, dim3(1), dim3(kWave), 0, s, sc, cap_records, cap_list, dmax);
}
void launch_upward_tail(int d_lo, const int32_t* level_list, const uint32_t* end_list, RefNode* nodes,
                        WalkHead* heads, double theta, const int32_t* sys, const double* root, const DevScalars* sc,
                        hipStream_t s) {
  if (d_lo < kMaxLevels)
    hipLaunchKernelGGL(k_upward_tail, dim3(1), dim3(kBlock), 0, s, d_lo, level_list, end_list, nodes, heads,
                       1.0 / theta, sys, root, sc);
}
void launch_upward_level(unsigned grid, int d, const int32_t* level_list, const uint32_t* end_list, RefNode* nodes,
                         WalkHead* heads, double theta, const int32_t* sys, const double* root, const DevScalars* sc,
                         hipStream_t s) {
  hipLaunchKernelGGL(k_upward_level, dim3(grid), dim3(kBlock), 0, s, d, level_list, end_list, nodes, heads,
                     1.0 / theta, sys, root, sc);
}
void launch_upward(const int32_t* level_list, const uint32_t* end_list, int64_t beg, int64_t cnt, int d,
                   RefNode* nodes, WalkHead* heads, double theta, const int32_t* sys, const double* root,
                   hipStream_t s) {
  if (cnt > 0)
    hipLaunchKernelGGL(k_upward, dim3(grid_of(cnt)), dim3(kBlock), 0, s, level_list, end_list, beg, cnt, d, nodes,
                       heads, 1.0 / theta, sys, root);
}
void launch_upward_all(const int32_t* counts, const int32_t* level_list, const uint32_t* end_list, RefNode* nodes,
                       WalkHead* heads, double theta, const int32_t* sys, const double* root, hipStream_t s) {
  int64_t off = 0;
  for (int d = 0; d < kMaxDepthSlots; ++d) off += counts[d];
  for (int d = kMaxDepthSlots - 1; d >= 0; --d) {
    off -= counts[d];
    launch_upward(level_list, end_list, off, counts[d], d, nodes, heads, theta, sys, root, s);
  }
}
void launch_walk(const Group* groups, int64_t ngroups, const RefNode* nodes, const WalkHead* heads,
                 const StateArrays& st,
                 const SysArrays& sy, double theta, double* ax, double* ay, double* az, DevScalars* sc, bool kick,
                 bool count, hipStream_t s) {
  if (ngroups <= 0) return;
  const int wpb = kWalkBlock / kWave;
  const unsigned nb = (unsigned)(div_up(div_up(ngroups, wpb), 8) * 8);
#define NBX_RW(K, C) \
  hipLaunchKernelGGL((k_walk<K, C>), dim3(nb), dim3(kWalkBlock), 0, s, groups, ngroups, nodes, heads, st, sy, theta, \
                     ax, ay, az, sc)
  if (kick && count) NBX_RW(true, true);
  else if (kick) NBX_RW(true, false);
  else if (count) NBX_RW(false, true);
  else NBX_RW(false, false);
#undef NBX_RW
}
void launch_fill_u32(int64_t n, uint32_t* a, uint32_t v, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_fill_u32, dim3(grid_of(n)), dim3(kBlock), 0, s, n, a, v);
}
void launch_kick_drift(int64_t n, const StateArrays& st, const double* ax, const double* ay, const double* az,
                       const SysArrays& sy, float* amag, hipStream_t s) {
  if (n > 0) hipLaunchKernelGGL(k_kick_drift, dim3(grid_of(n)), dim3(kBlock), 0, s, n, st, ax, ay, az, sy, amag);
}
void launch_maxv2(int64_t n, const StateArrays& st, const SysArrays& sy, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_maxv2, dim3((unsigned)div_up(n, (int64_t)kBlock * kMaxv2Items)), dim3(kBlock), 0, s, n, st,
                       sy);
}
void launch_dt(const SysArrays& sy, double base, double* dts, int64_t step, DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_dt, dim3(sys_grid(sy)), dim3(kBlock), 0, s, sy, base, dts, step, sc);
}
void launch_export(int64_t total, const RefNode* nodes, const uint64_t* key, const uint64_t* key2,
                   const int32_t* id, const SysArrays& sy, int32_t* depth, int32_t* leaf, int32_t* body,
                   int32_t* cmask, double* geo4, double* mass, double* com3, double* q6, hipStream_t s) {
  if (total > 0)
    hipLaunchKernelGGL(k_export, dim3(grid_of(total)), dim3(kBlock), 0, s, total, nodes, key, key2, id, sy, depth,
                       leaf, body, cmask, geo4, mass, com3, q6);
}

}  // namespace nbx
