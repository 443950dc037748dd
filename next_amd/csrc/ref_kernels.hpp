// Launchers for the reference-mode (FP64, NEXT-exact) kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"

namespace nbx {

// Particle state, structure of arrays (particle.h:22-52) plus `id`, the
// upload-order index of the body held in each slot.
struct StateArrays {
  double *x, *y, *z, *vx, *vy, *vz, *m;
  int32_t *type, *id;
};

constexpr int kBlock = 256;
constexpr int kBboxBlocks = 1024;   // partial min/max blocks (grid-stride)
// Row of per-block partial counts written by the lcp kernel.
constexpr int kPartMaxDepth = kMaxDepthSlots;      // column: deepest leaf
constexpr int kPartEqual = kMaxDepthSlots + 1;     // column: equal-key pairs
constexpr int kPartW = kMaxDepthSlots + 2;
inline int64_t lcp_part_words(int64_t n) { return ((n + kBlock - 1) / kBlock) * kPartW; }

void launch_fill_scalars(DevScalars* sc, int what, hipStream_t s);  // what: 1 build counters, 2 v2max, 4 err
void launch_set_dt(DevScalars* sc, double dt, hipStream_t s);
void launch_bbox(int64_t n, const StateArrays& st, double* part, DevScalars* sc, hipStream_t s);
void launch_keys(int64_t n, const StateArrays& st, const DevScalars* sc, uint64_t* key, uint32_t* idx,
                 DevScalars* scw, hipStream_t s);
void launch_keys2(int64_t n, const StateArrays& st, const DevScalars* sc, uint64_t* key2, hipStream_t s);
void launch_iota(int64_t n, uint32_t* idx, hipStream_t s);
void launch_gather_u64(int64_t n, const uint32_t* perm, const uint64_t* src, uint64_t* dst, hipStream_t s);
void launch_gather_state(int64_t n, const uint32_t* perm, const StateArrays& src, const StateArrays& dst,
                         hipStream_t s);
void launch_lcp_count(int64_t n, const uint64_t* key, const uint64_t* key2, int32_t* lcp, uint32_t* cnt_s,
                      uint32_t* cnt_e, uint32_t* part, DevScalars* sc, hipStream_t s);
void launch_emit(int64_t n, const int32_t* lcp, const uint32_t* pre_off, const uint32_t* end_off,
                 const StateArrays& st, const DevScalars* sc, RefNode* nodes, uint32_t* st_key,
                 uint32_t* st_val, uint32_t* en_key, uint32_t* en_val, hipStream_t s);
void launch_pair(int64_t m, const uint32_t* st_val, const uint32_t* en_val, RefNode* nodes, int32_t* level_list,
                 hipStream_t s);
void launch_upward(const int32_t* level_list, int64_t beg, int64_t cnt, RefNode* nodes, hipStream_t s);
void launch_walk(int64_t n, int64_t total, const RefNode* nodes, const StateArrays& st, double theta, double* ax,
                 double* ay, double* az, DevScalars* sc, bool kick, bool count, hipStream_t s);
void launch_kick_drift(int64_t n, const StateArrays& st, const double* ax, const double* ay, const double* az,
                       const DevScalars* sc, double* bbox_part, hipStream_t s);
void launch_bbox_final(const double* part, int nparts, DevScalars* sc, hipStream_t s);
void launch_maxv2(int64_t n, const StateArrays& st, DevScalars* sc, hipStream_t s);
void launch_dt(DevScalars* sc, double base, double* dts, int64_t step, hipStream_t s);
void launch_export(int64_t total, const RefNode* nodes, const uint64_t* key, const uint64_t* key2,
                   const int32_t* id, const DevScalars* sc, int32_t* depth, int32_t* leaf, int32_t* body,
                   int32_t* cmask, double* geo4, double* mass, double* com3, double* q6, hipStream_t s);
void launch_keys_export(int64_t n, const StateArrays& st, const DevScalars* sc, uint64_t* hi, uint64_t* lo,
                        hipStream_t s);

}  // namespace nbx
