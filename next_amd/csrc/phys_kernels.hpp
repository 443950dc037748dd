// Launchers for the physics-mode kernels (phys_kernels.hip): bucket-leaf
// group walk with traceless quadrupoles and Plummer softening, and the FP64
// direct sum used for force errors and energies.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "common.hpp"
#include "ref_kernels.hpp"

namespace nbx {

// Walk record of a node (64 bytes): FP64 centre of mass for the
// group-relative conversion, FP32 moments, and the traversal links.
struct alignas(64) PhysNode {
  double cx, cy, cz;
  float m;
  float rc2;                  // (bmax / theta)^2; +inf: never accepted as a cell
  float qxx, qyy, qxy, qxz, qyz;   // traceless: qzz = -(qxx + qyy)
  int32_t skip;
  int32_t first;              // first body (sorted index)
  int32_t count;              // bodies in the subtree
};
static_assert(sizeof(PhysNode) == 64, "PhysNode layout");

// A leaf is never accepted as a cell (rc2 = +inf), so its first two
// quadrupole slots hold its FP64 mass, which the upward pass reads.
__host__ __device__ inline void set_leaf_mass(PhysNode& nd, double m) { __builtin_memcpy(&nd.qxx, &m, 8); }
__host__ __device__ inline double leaf_mass(const PhysNode& nd) {
  double m;
  __builtin_memcpy(&m, &nd.qxx, 8);
  return m;
}

constexpr int kPhysBlock = 64;      // one wave = one group of targets
// Per-wave LDS (7.4 KB with these sizes) bounds the walk's occupancy at
// about 5 waves per SIMD; a 256-entry stack and body list (11 KB) allowed 3.
// Entries of the per-wave LDS lists (accepted cells, bodies of accepted
// buckets) a walk collects before it evaluates them: a walk setting
// (nbx_set_walk_list), one of kPhysLists.
constexpr int kPhysListDefault = 128;
constexpr int kPhysLists[] = {96, 128, 192, 256};
constexpr int kPhysMaxLeaf = 64;    // bucket size limit (one body per lane)
constexpr int kPhysStack = 32;      // per-wave LDS stack of donated node ranges
constexpr int kPhysAuxBytes = 128;  // FP64 moments per node (mass, bmax, Q, box)

// Walk groups are marked per first body in gcnt (body count) and gmark
// (1), both zeroed beforehand, then compacted in body order.
// Node layout from the sorted keys (the arguments of launch_emit, capacities
// included): leaf records complete, internal nodes' first body, their places
// in the per-depth node list and the list of skip targets; a leaf that is a
// system root is a one-body group.
void launch_phys_emit(int64_t n, const int32_t* lcp, const uint32_t* pre_off, const uint32_t* part,
                      const StateArrays& st, const DevScalars* sc, PhysNode* pn, int32_t* level_list,
                      uint32_t* end_list, int32_t* gcnt, uint32_t* gmark, uint32_t cap_nodes, uint32_t cap_int,
                      hipStream_t s);
// Skip pointers of the m internal nodes from the two lists.
void launch_phys_pair(int64_t m, const int32_t* level_list, const uint32_t* end_list, PhysNode* pn, hipStream_t s);
// The same two passes for builds whose counts stay on the device (see
// launch_check): fixed grids, counts from the scalars.
void launch_phys_pair_dev(unsigned grid, const int32_t* level_list, const uint32_t* end_list, PhysNode* pn,
                          const DevScalars* sc, hipStream_t s);
void launch_phys_upward_level(unsigned grid, int d, const int32_t* level_list, PhysNode* pn, void* aux, double theta,
                              int group_size, int32_t* gcnt, uint32_t* gmark, const DevScalars* sc, hipStream_t s);
void launch_phys_upward_tail(int d_lo, const int32_t* level_list, PhysNode* pn, void* aux, double theta,
                             int group_size, int32_t* gcnt, uint32_t* gmark, const DevScalars* sc, hipStream_t s);
// Internal nodes of one depth from their children (FP64 accumulation), and
// the walk groups of at most group_size bodies each node packs (see
// k_phys_upward).  `aux`: kPhysAuxBytes per node of FP64 moments (mass,
// bmax, quadrupole, bounding box of the bodies), internal nodes only.
void launch_phys_upward(const int32_t* level_list, int64_t beg, int64_t cnt, int depth, PhysNode* pn, void* aux,
                        double theta, int group_size, int32_t* gcnt, uint32_t* gmark, hipStream_t s);
// The group table from the marks (gmark is overwritten by its scan);
// *ngroups (device) receives the count; scratch as for scan_u32 over n.
void launch_group_compact(int64_t n, const int32_t* gcnt, uint32_t* gmark, const int32_t* body_sys, Group* groups,
                          uint32_t* ngroups, uint32_t* scratch, hipStream_t s);
// Group walk over the group table: accelerations (and potentials if
// pot != null), optionally followed by the closing half kick.  fp64 selects
// the FP64 evaluation.  With amag (|a| of each body from the previous
// evaluation, FP32) and alpha > 0, a cell that passes the geometric test
// (bmax / theta) is accepted only if m bmax^3 <= alpha |a|min d^5, |a|min
// the smallest |a| of the group and d the distance to the group's box.  The
// kernel takes the group count from sc->tree_groups; grid_groups (> 0) only
// sizes the grid, and a grid below the count walks the rest in more rounds.
void launch_phys_walk(const Group* groups, int64_t grid_groups, const PhysNode* pn, const StateArrays& st,
                      const SysArrays& sy, double eps, int nleaf, double* ax, double* ay, double* az, double* pot,
                      DevScalars* sc, bool kick, bool count, bool fp64, const float* amag, double alpha,
                      double theta, int list, hipStream_t s);
// FP64 direct sum with Plummer softening over all pairs of each system (G = 1).
void launch_direct64(int64_t n, const StateArrays& st, const SysArrays& sy, double eps, double* ax, double* ay,
                     double* az, double* pot, hipStream_t s);
// Kinetic energy, potential energy (0.5 sum m phi) and momentum of each
// system, reduced in a fixed order over the bounding-box blocks:
// out (device, 5 per system) = {ekin, epot, px, py, pz}; part holds 5 doubles
// per block.
void launch_energy(const StateArrays& st, const double* pot, const BoxBlock* blocks, int nblocks,
                   const int32_t* block_off, int nsys, double* part, double* out, hipStream_t s);

}  // namespace nbx
