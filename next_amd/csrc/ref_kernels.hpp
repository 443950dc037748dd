// Launchers for the reference-mode (FP64, NEXT-exact) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace nbx {

// Particle state, structure of arrays (particle.h:22-52) plus `id`, the
// upload-order index of the body held in each slot, and `sys`, its system.
struct StateArrays {
  double *x, *y, *z, *vx, *vy, *vz, *m;
  int32_t *type, *id, *sys;
};

constexpr int kBlock = 256;
constexpr int kBoxSpan = 2048;      // bodies per block of the bounding-box reduction
// Row of per-block partial counts written by the lcp kernel: internal nodes
// started per depth (columns 0..63), ended per depth (64..127), deepest leaf,
// equal-key pairs.  k_lcp_scan turns the per-depth columns into exclusive
// prefixes over the blocks.
constexpr int kPartEnds = kMaxDepthSlots;
constexpr int kPartMaxDepth = 2 * kMaxDepthSlots;  // column: deepest leaf
constexpr int kPartEqual = 2 * kMaxDepthSlots + 1; // column: equal-key pairs
constexpr int kPartW = 2 * kMaxDepthSlots + 2;
inline int64_t lcp_part_words(int64_t n) { return ((n + kBlock - 1) / kBlock) * kPartW; }

void launch_fill_scalars(DevScalars* sc, int what, hipStream_t s);  // what: 1 build counters, 4 err, 8 walk counters
// v2max (the largest |v|^2 per system, which the closing kicks raise) to 0;
// k_dt also resets it once it has read it.
void launch_fill_v2max(const SysArrays& sy, hipStream_t s);
void launch_set_dt(const SysArrays& sy, double dt, hipStream_t s);
// Roots of the systems from their bodies' bounding boxes; clear (may be
// null): the scalars a tree build counts into, reset as launch_fill_scalars
// with what = 1 | 4 would.
void launch_bbox(const StateArrays& st, const BoxBlock* blocks, int nblocks, const int32_t* block_off, double* part,
                 const SysArrays& sy, DevScalars* clear, hipStream_t s);
void launch_keys(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* key, uint32_t* idx,
                 DevScalars* scw, hipStream_t s);
void launch_keys2(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* key2, hipStream_t s);
// Runs of bodies of one system with equal first key words (in the order of
// the sort by the first word) sorted in place by the second word, stably;
// key2 written for those bodies (amag, may be null, moves with them).
void launch_fix_runs(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* key, uint64_t* key2,
                     uint32_t* amag, int lo, hipStream_t s);
void launch_iota(int64_t n, uint32_t* idx, hipStream_t s);
void launch_gather_u64(int64_t n, const uint32_t* perm, const uint64_t* src, uint64_t* dst, hipStream_t s);
void launch_gather_sys(int64_t n, const uint32_t* perm, const int32_t* sys, uint32_t* dst, hipStream_t s);
// dst[i] = src[perm[i]] for 32-bit words.
void launch_gather_u32(int64_t n, const uint32_t* perm, const uint32_t* src, uint32_t* dst, hipStream_t s);
void launch_gather_state(int64_t n, const uint32_t* perm, const StateArrays& src, const StateArrays& dst,
                         hipStream_t s);
// lcp of the sorted keys, each body's node count (its started internal
// nodes and its leaf) in cnt_s, the per-block rows of per-depth counts
// (scanned over the blocks) in part, and in sc the internal nodes per depth,
// the deepest leaf and the equal-key pairs.
void launch_lcp_count(int64_t n, const uint64_t* key, const uint64_t* key2, const int32_t* sys, int32_t* lcp,
                      uint32_t* cnt_s, uint32_t* part, DevScalars* sc, hipStream_t s);
void launch_sys_roots(const SysArrays& sy, const uint32_t* pre_off, const DevScalars* sc, hipStream_t s);
void launch_remove_map(int64_t n_new, int64_t cut, int64_t len, uint32_t* perm, hipStream_t s);
void launch_relabel(int64_t n, int32_t* sys, int32_t removed, hipStream_t s);
void launch_gather_f64(int64_t n, const uint32_t* perm, const double* src, double* dst, hipStream_t s);
// Leaf records and walk heads from the sorted keys' lcp and the scan of
// cnt_s (pre_off), and every internal node's place in the per-depth node list
// and the list of skip targets (layout.hpp); writes records below cap_nodes
// and list entries below cap_int only (the buffers' capacities), so it can
// run before the host knows the node counts.  The internal nodes' records
// come from the upward pass.
void launch_emit(int64_t n, const int32_t* lcp, const uint32_t* pre_off, const uint32_t* part,
                 const StateArrays& st, const SysArrays& sy, const DevScalars* sc, RefNode* nodes, WalkHead* heads,
                 int32_t* level_list, uint32_t* end_list, uint32_t cap_nodes, uint32_t cap_int, hipStream_t s);
// Builds whose counts stay on the device (no host read-back; a step of them
// can be captured): k_check derives each depth's offset and raises the abort
// flag when the tree does not fit cap_records node records, cap_list
// node-list entries or depth levels 0..dmax; the per-depth upward kernels
// take their counts from the scalars with fixed grids, and every tree kernel
// does nothing once the flag is set.
void launch_check(DevScalars* sc, int64_t cap_records, int64_t cap_list, int dmax, hipStream_t s);
// Start of a captured run: abort flag cleared, dt log entries from `step` on.
void launch_set_run(DevScalars* sc, int32_t step, hipStream_t s);
// Internal node records and heads of depth d (the list entries [beg, beg +
// cnt)), from their children's records; sys and root: the bodies' systems
// and the systems' roots (4 doubles each).
void launch_upward_level(unsigned grid, int d, const int32_t* level_list, const uint32_t* end_list, RefNode* nodes,
                         WalkHead* heads, double theta, const int32_t* sys, const double* root, const DevScalars* sc,
                         hipStream_t s);
// Depths d_lo .. 41 (whatever the build holds there) in one workgroup.
void launch_upward_tail(int d_lo, const int32_t* level_list, const uint32_t* end_list, RefNode* nodes,
                        WalkHead* heads, double theta, const int32_t* sys, const double* root, const DevScalars* sc,
                        hipStream_t s);
void launch_upward(const int32_t* level_list, const uint32_t* end_list, int64_t beg, int64_t cnt, int d,
                   RefNode* nodes, WalkHead* heads, double theta, const int32_t* sys, const double* root,
                   hipStream_t s);
// Every depth of a build whose per-depth counts the host holds (counts[d],
// the node list in depth order), one launch per depth, deepest first.  (A
// narrow level costs the latency of its nodes' child chains, not the launch:
// running the narrow levels at either end of the tree in one workgroup each
// measured slower.)
void launch_upward_all(const int32_t* counts, const int32_t* level_list, const uint32_t* end_list, RefNode* nodes,
                       WalkHead* heads, double theta, const int32_t* sys, const double* root, hipStream_t s);
void launch_fill_u32(int64_t n, uint32_t* a, uint32_t v, hipStream_t s);
void launch_walk(const Group* groups, int64_t ngroups, const RefNode* nodes, const WalkHead* heads,
                 const StateArrays& st,
                 const SysArrays& sy, double theta, double* ax, double* ay, double* az, DevScalars* sc, bool kick,
                 bool count, hipStream_t s);
// amag (may be null): |a| per body in FP32.
void launch_kick_drift(int64_t n, const StateArrays& st, const double* ax, const double* ay, const double* az,
                       const SysArrays& sy, float* amag, hipStream_t s);
void launch_maxv2(int64_t n, const StateArrays& st, const SysArrays& sy, hipStream_t s);
// dts (may be null) receives system 0's dt at entry `step`, or, for step < 0,
// at the entry the device's step counter gives (which it advances).
void launch_dt(const SysArrays& sy, double base, double* dts, int64_t step, DevScalars* sc, hipStream_t s);
void launch_export(int64_t total, const RefNode* nodes, const uint64_t* key, const uint64_t* key2,
                   const int32_t* id, const SysArrays& sy, int32_t* depth, int32_t* leaf, int32_t* body,
                   int32_t* cmask, double* geo4, double* mass, double* com3, double* q6, hipStream_t s);
void launch_keys_export(int64_t n, const StateArrays& st, const SysArrays& sy, uint64_t* hi, uint64_t* lo,
                        hipStream_t s);

}  // namespace nbx
