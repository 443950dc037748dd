// Physics-mode kernels on gfx950: a bucket-leaf Barnes-Hut walk with
// traceless quadrupoles, Plummer softening and a group acceptance test, plus
// an FP64 direct sum.
//
// Tree: the same depth-first preorder octree as reference mode (one body per
// leaf, skip pointers), read through a compact 64-byte PhysNode.  A node
// whose body count is at most `nleaf` that must be opened is treated as a
// bucket: its bodies go straight to the direct-sum list.
//
// Multipoles (FP64): m, centre of mass, Q_ij = sum m (3 r_i r_j - r^2 d_ij)
// about the centre of mass, propagated with the parallel-axis rule
//   Q = sum_c [Q_c + m_c (3 r_c r_c^T - |r_c|^2 I)],  r_c = com_c - com,
// and bmax, a bound on the distance from the centre of mass to any body: the
// smaller of the distance to the farthest corner of the bodies' bounding box
// and max_c (|r_c| + bmax_c).  A refit (bodies moved, same order and
// topology) recomputes all of these from the current positions.
//
// Walk: one wave = one group of at most 64 targets packed from the tree
// (consecutive sibling subtrees inside one node).  The wave walks the
// preorder tree with a decision shared by the whole group: a node is
// accepted when the distance from its centre of mass to the group's bounding
// box exceeds bmax/theta (and, with the relative criterion, when its
// octupole-order term is small against the group's previous |a|).  Lanes
// walk disjoint ranges of the preorder array
// and hand work to idle lanes (k_phys_walk).  Accepted cells and bucket
// bodies are appended, relative to the group centre, to per-wave LDS lists;
// a full list is evaluated by all the group's lanes at once (entries read as
// LDS broadcasts).  Each flush's sum is formed in the evaluation precision
// and added to FP64 totals.
//
// Acceleration on a target at distance vector d = com - x from a cell:
//   a = m d / r^3 - Q d / r^5 + (5/2) (d^T Q d) d / r^7,  r^2 = |d|^2 + eps^2
// (the gradient of phi = -m/r - (1/2) x^T Q x / r^5 with x = -d).
#include "common.hpp"
#include "layout.hpp"
#include "phys_kernels.hpp"
#include "prims.hpp"

#include <algorithm>
#include <cmath>

namespace nbx {
namespace {

struct PhysAux {
  double m, bmax;
  double q[6];          // xx yy zz xy xz yz
  double lo[3], hi[3];  // bounding box of the node's bodies
  double pad[2];
};
static_assert(sizeof(PhysAux) == kPhysAuxBytes, "PhysAux layout");

inline unsigned grid_of(int64_t n, int b) { return (unsigned)div_up(n, b); }

// A walk group of c bodies starts at sorted body f: marked in two zeroed
// per-body arrays, compacted in body (Morton) order afterwards.
__device__ __forceinline__ void mark_group(int32_t f, int32_t c, int32_t* gcnt, uint32_t* gmark) {
  gcnt[f] = c;
  gmark[f] = 1u;
}

// Node layout from the sorted keys, as k_emit does for the reference tree
// but writing walk records: for body i, the internal nodes that start at it
// (first body only; the skip pointer comes from k_phys_pair, the moments from
// k_phys_upward) and its leaf, complete.  A leaf is never accepted as a cell
// (rc2 = +inf), so its quadrupole slots carry its FP64 mass for the upward
// pass.  A leaf that is a system's root is that system's walk group.
__global__ __launch_bounds__(kBlock) void k_phys_emit(int64_t n, const int32_t* __restrict__ lcp,
                                                      const uint32_t* __restrict__ pre_off,
                                                      const uint32_t* __restrict__ part, StateArrays st,
                                                      const DevScalars* __restrict__ sc, PhysNode* __restrict__ pn,
                                                      int32_t* __restrict__ level_list,
                                                      uint32_t* __restrict__ end_list,
                                                      int32_t* __restrict__ gcnt, uint32_t* __restrict__ gmark,
                                                      uint32_t cap_nodes, uint32_t cap_int) {
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
  const int ns = li > lp ? li - lp : 0;
  for (int t = 0; t < ns; ++t) {
    const uint32_t p = pre + t;
    if (p >= cap_nodes) break;
    pn[p].first = (int32_t)i;
  }
  if (pre + ns < cap_nodes) {
    const uint32_t p = pre + ns;
    PhysNode nd;
    nd.cx = st.x[i]; nd.cy = st.y[i]; nd.cz = st.z[i];
    const double m = st.m[i];
    nd.m = (float)m;
    nd.rc2 = INFINITY;
    set_leaf_mass(nd, m);
    nd.qxy = nd.qxz = nd.qyz = 0.f;
    nd.skip = (int32_t)(p + 1);
    nd.first = (int32_t)i;
    nd.count = 1;
    pn[p] = nd;
  }
  if (li < 0 && lp < 0) mark_group((int32_t)i, 1, gcnt, gmark);
}

// Skip pointer of the k-th internal node of the per-depth list.
__global__ __launch_bounds__(kBlock) void k_phys_pair(int64_t m, const int32_t* __restrict__ level_list,
                                                      const uint32_t* __restrict__ end_list,
                                                      PhysNode* __restrict__ pn) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= m) return;
  pn[level_list[k]].skip = (int32_t)end_list[k];
}

// The same with the count on the device (fixed grid striding over it).
__global__ __launch_bounds__(kBlock) void k_phys_pair_dev(const int32_t* __restrict__ level_list,
                                                          const uint32_t* __restrict__ end_list,
                                                          PhysNode* __restrict__ pn,
                                                          const DevScalars* __restrict__ sc) {
  if (sc->abort_flags) return;
  const int64_t m = (int64_t)sc->level_off[kMaxDepthSlots - 1] + sc->level_count[kMaxDepthSlots - 1];
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < m; k += (int64_t)gridDim.x * kBlock)
    pn[level_list[k]].skip = (int32_t)end_list[k];
}

// Internal nodes of one depth from their children (FP64), and the walk
// groups this node packs: a node of more than G bodies packs its children of
// at most G bodies into groups, consecutive siblings together while the sum
// stays within G; a system root of at most G bodies is a group of its own.
// Every body lands in exactly one group and a group's bodies lie inside one
// node, so groups are compact in space where runs of G consecutive Morton
// keys are not (such a run can straddle two distant cells).
__device__ __forceinline__ void phys_upward_node(int32_t id, bool root_level, PhysNode* __restrict__ pn,
                                                 PhysAux* __restrict__ aux, double inv_theta, int G,
                                                 int32_t* __restrict__ gcnt, uint32_t* __restrict__ gmark) {
  const int32_t end = pn[id].skip;
  const int32_t first = pn[id].first;
  // One walk of the child chain (at most 8 children) records the child
  // indices (-1: no child) and sizes and sums mass and mass moment; the
  // second pass reads the same (now cached) records again, which keeps the
  // kernel's registers, and so its occupancy, where a latency-bound pass
  // needs them.
  int32_t chv[8], cnv[8];
  double M = 0, sx = 0, sy = 0, sz = 0;
  int32_t count = 0;
  int32_t c = id + 1;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    chv[j] = -1; cnv[j] = 0;
    if (c < end) {
      const PhysNode& ch = pn[c];
      chv[j] = c;
      cnv[j] = ch.count;
      const double cm = ch.count == 1 ? leaf_mass(ch) : aux[c].m;
      M += cm;
      sx += ch.cx * cm; sy += ch.cy * cm; sz += ch.cz * cm;
      count += ch.count;
      c = ch.skip;
    }
  }
  double cx, cy, cz;
  if (M > 0) { cx = sx / M; cy = sy / M; cz = sz / M; }
  else { cx = pn[id + 1].cx; cy = pn[id + 1].cy; cz = pn[id + 1].cz; }
  double q[6] = {0, 0, 0, 0, 0, 0};
  double bsum = 0;
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    if (chv[j] < 0) continue;
    const PhysNode& ch = pn[chv[j]];
    const double px = ch.cx, py = ch.cy, pz = ch.cz;
    const double rx = px - cx, ry = py - cy, rz = pz - cz;
    const double r2 = rx * rx + ry * ry + rz * rz;
    double b = sqrt(r2);
    double mc;
    if (cnv[j] == 1) {                                     // a body: a point
      mc = leaf_mass(ch);
      lo[0] = fmin(lo[0], px); hi[0] = fmax(hi[0], px);
      lo[1] = fmin(lo[1], py); hi[1] = fmax(hi[1], py);
      lo[2] = fmin(lo[2], pz); hi[2] = fmax(hi[2], pz);
    } else {                                               // a cell: its moments and box
      const PhysAux& a = aux[chv[j]];
      mc = a.m;
      for (int t = 0; t < 6; ++t) q[t] += a.q[t];
      b += a.bmax;
      for (int d = 0; d < 3; ++d) {
        lo[d] = fmin(lo[d], a.lo[d]);
        hi[d] = fmax(hi[d], a.hi[d]);
      }
    }
    q[0] += mc * (3 * rx * rx - r2);
    q[1] += mc * (3 * ry * ry - r2);
    q[2] += mc * (3 * rz * rz - r2);
    q[3] += mc * 3 * rx * ry;
    q[4] += mc * 3 * rx * rz;
    q[5] += mc * 3 * ry * rz;
    if (b > bsum) bsum = b;
  }
  // bmax: the farthest corner of the bodies' box from the centre of mass,
  // or the children's bound if that is smaller.
  const double ex = fmax(cx - lo[0], hi[0] - cx), ey = fmax(cy - lo[1], hi[1] - cy), ez = fmax(cz - lo[2], hi[2] - cz);
  double bmax = sqrt(ex * ex + ey * ey + ez * ez);
  if (bsum < bmax) bmax = bsum;
  PhysNode nd;
  nd.cx = cx; nd.cy = cy; nd.cz = cz;
  nd.m = (float)M;
  nd.count = count;
  // Cells of very few bodies cost more as a multipole than as bodies.
  const double rc = bmax * inv_theta;
  nd.rc2 = count < 4 ? INFINITY : (float)(rc * rc * (1.0 + 1e-6));
  nd.qxx = (float)q[0]; nd.qyy = (float)q[1];
  nd.qxy = (float)q[3]; nd.qxz = (float)q[4]; nd.qyz = (float)q[5];
  nd.skip = end;
  nd.first = first;
  pn[id] = nd;
  // walk groups
  if (count > G) {
    int32_t run_first = 0, run = 0, f = first;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int32_t cc = cnv[j];
      const int32_t cf = f;                                // children's bodies follow one another
      f += cc;
      if (cc == 0) continue;
      if (cc > G) {                                        // packs its own children
        if (run > 0) mark_group(run_first, run, gcnt, gmark);
        run = 0;
        continue;
      }
      if (run > 0 && run + cc > G) { mark_group(run_first, run, gcnt, gmark); run = 0; }
      if (run == 0) run_first = cf;
      run += cc;
    }
    if (run > 0) mark_group(run_first, run, gcnt, gmark);
  } else if (root_level) {
    mark_group(first, count, gcnt, gmark);
  }
  PhysAux a;
  a.m = M;
  a.bmax = bmax;
  for (int j = 0; j < 6; ++j) a.q[j] = q[j];
  for (int d = 0; d < 3; ++d) { a.lo[d] = lo[d]; a.hi[d] = hi[d]; }
  a.pad[0] = a.pad[1] = 0.0;
  aux[id] = a;
}

__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void k_phys_upward(const int32_t* __restrict__ level_list, int64_t beg, int64_t cnt, bool root_level,
                   PhysNode* __restrict__ pn, PhysAux* __restrict__ aux, double inv_theta, int G,
                   int32_t* __restrict__ gcnt, uint32_t* __restrict__ gmark) {
  const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (k >= cnt) return;
  phys_upward_node(level_list[beg + k], root_level, pn, aux, inv_theta, G, gcnt, gmark);
}

// Depths d_lo .. 41 of a build whose counts stay on the device in one
// workgroup, deepest first, a barrier between depths (see k_upward_tail).
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void k_phys_upward_tail(int d_lo, const int32_t* __restrict__ level_list, PhysNode* __restrict__ pn,
                        PhysAux* __restrict__ aux, double inv_theta, int G, int32_t* __restrict__ gcnt,
                        uint32_t* __restrict__ gmark, const DevScalars* __restrict__ sc) {
  if (sc->abort_flags) return;
  for (int d = kMaxLevels - 1; d >= d_lo; --d) {
    const int64_t beg = sc->level_off[d], cnt = sc->level_count[d];
    if (cnt == 0) continue;
    for (int64_t k = threadIdx.x; k < cnt; k += kBlock)
      phys_upward_node(level_list[beg + k], d == 0, pn, aux, inv_theta, G, gcnt, gmark);
    __syncthreads();
  }
}

// Depth d of a build whose counts stay on the device (k_check): fixed grid
// striding over the depth's extent in the node list.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(5)))
void k_phys_upward_level(int d, const int32_t* __restrict__ level_list, PhysNode* __restrict__ pn,
                         PhysAux* __restrict__ aux, double inv_theta, int G, int32_t* __restrict__ gcnt,
                         uint32_t* __restrict__ gmark, const DevScalars* __restrict__ sc) {
  if (sc->abort_flags) return;
  const int64_t beg = sc->level_off[d], cnt = sc->level_count[d];
  for (int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x; k < cnt; k += (int64_t)gridDim.x * kBlock)
    phys_upward_node(level_list[beg + k], d == 0, pn, aux, inv_theta, G, gcnt, gmark);
}

template <typename T>
struct alignas(4 * sizeof(T)) V4 {
  T x, y, z, w;
};
template <typename T>
struct alignas(2 * sizeof(T)) V2 {
  T x, y;
};

// Sum over the lanes that share a target slot (lanes slot + p * width).
template <typename T>
__device__ __forceinline__ void combine_parts(T& fx, T& fy, T& fz, T& fp, int width) {
  for (int off = width; off < kWave; off <<= 1) {
    fx += __shfl_xor(fx, off, kWave);
    fy += __shfl_xor(fy, off, kWave);
    fz += __shfl_xor(fz, off, kWave);
    fp += __shfl_xor(fp, off, kWave);
  }
}

template <typename T>
__device__ __forceinline__ T rsq(T x);
template <>
__device__ __forceinline__ float rsq<float>(float x) { return __frsqrt_rn(x); }
template <>
__device__ __forceinline__ double rsq<double>(double x) { return rsqrt(x); }

// Group walk with work donation.  Every lane walks a contiguous range of
// the preorder node array with skip pointers (accept or bucket: jump past the
// subtree; open: step to the first child).  Lane 0 starts with the whole
// tree; when a lane opens a node, it hands the rest of its range after that
// node's subtree, [skip, end), to a per-wave LDS stack from which idle lanes
// take work (last in, first out).  Handing work over is optional, so a full
// stack only means a lane keeps walking its range itself: the stack cannot
// overflow.  All decisions are made in lockstep by one wave and compacted in
// lane order, so the lists, and every sum, come out in the same order on
// every run.
template <typename T, int LIST, bool KICK, bool POT, bool COUNT>
__global__ __launch_bounds__(kPhysBlock) void k_phys_walk(const Group* __restrict__ groups,
                                                          const PhysNode* __restrict__ pn, StateArrays st,
                                                          SysArrays sy, double eps, int nleaf,
                                                          double* __restrict__ ax, double* __restrict__ ay,
                                                          double* __restrict__ az, double* __restrict__ pot,
                                                          DevScalars* sc, const float* __restrict__ amag,
                                                          float rel2, float th2) {
  __shared__ int32_t stk_b[kPhysStack];  // donated ranges [stk_b, stk_e)
  __shared__ int32_t stk_e[kPhysStack];
  __shared__ int32_t bfirst[kWave];      // buckets of the current step: first body,
  __shared__ int32_t bstart[kWave];      //   offset in the step's flattened body list
  __shared__ V4<T> ca[LIST];             // cell: relative centre of mass, mass
  __shared__ V4<T> cb[LIST];             // cell: qxx qyy qxy qxz
  __shared__ V2<T> cq[LIST];             // cell: qyz qzz
  __shared__ V4<T> pl[LIST];             // body: relative position, mass

  const int lane = threadIdx.x;
  // The group count comes from the device (the compaction that wrote it ran
  // earlier on this stream), so the host need not read it back first; the
  // grid (a multiple of 8) is sized from a bound.  Blocks are dealt to the 8
  // XCDs round-robin: XCD x walks the contiguous (spatially coherent) range
  // [x per, (x + 1) per) of the groups with its gridDim / 8 blocks, so its L2
  // holds the part of the tree those groups touch, whatever the grid size.
  const int64_t ngroups = sc->abort_flags ? 0 : (int64_t)sc->tree_groups;   // nothing after an aborted build
  const int64_t per = (ngroups + 7) / 8;
  const int64_t xbase = (int64_t)(blockIdx.x % 8) * per;
  for (int64_t j = blockIdx.x / 8; j < per; j += gridDim.x / 8) {
    const int64_t gi = xbase + j;
    if (gi >= ngroups) break;                              // the last range is short
    const Group g = groups[gi];
    // A group of at most 32 (16, ...) targets puts 2 (4, ...) lanes on each
    // target: lane = slot + part * width, and the parts evaluate interleaved
    // entries of every list, their sums combined after each flush.
    int width = kWave;
    while (width > 1 && (width >> 1) >= g.count) width >>= 1;
    const int parts = kWave / width;
    const int slot = lane & (width - 1), part = lane / width;
    const int64_t i = g.first + slot;
    const bool valid = slot < g.count && part == 0;        // the lane that owns the result
    const int64_t ii = slot < g.count ? i : g.first;
    const double px = st.x[ii], py = st.y[ii], pz = st.z[ii];
    const uint64_t ltmask = lanemask_lt();
    const int32_t root = sy.sroot[g.sys], total = sy.send[g.sys];   // this system's tree

    // group bounding box
    double lo[3] = {px, py, pz}, hi[3] = {px, py, pz};
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const double a = __shfl_xor(lo[k], off, kWave), b = __shfl_xor(hi[k], off, kWave);
        lo[k] = a < lo[k] ? a : lo[k];
        hi[k] = b > hi[k] ? b : hi[k];
      }
    }
    const double gcx = 0.5 * (lo[0] + hi[0]), gcy = 0.5 * (lo[1] + hi[1]), gcz = 0.5 * (lo[2] + hi[2]);
    const float hx = (float)(0.5 * (hi[0] - lo[0])) * (1.0f + 1e-6f);
    const float hy = (float)(0.5 * (hi[1] - lo[1])) * (1.0f + 1e-6f);
    const float hz = (float)(0.5 * (hi[2] - lo[2])) * (1.0f + 1e-6f);
    const T tx = (T)(px - gcx), ty = (T)(py - gcy), tz = (T)(pz - gcz);
    const T e2 = (T)(eps * eps);
    // Relative criterion (amag != null): a cell that passes the geometric test
    // is accepted only if its first neglected (octupole) term, ~ m bmax^3 / d^5,
    // is at most alpha times the smallest previous |a| of the group:
    // m^2 bmax^6 <= alpha^2 amin^2 d^10, with bmax^2 = rc2 * theta^2.
    float arel = 0.f;
    if (amag) {
      float am = amag[ii];
#pragma unroll
      for (int off = 32; off >= 1; off >>= 1) am = fminf(am, __shfl_xor(am, off, kWave));
      arel = rel2 * am * am;
    }

    double Ax = 0, Ay = 0, Az = 0, P = 0;
    uint32_t ncell_tot = 0, npart_tot = 0, niter = 0, nopen = 0;
    int nc = 0, np = 0;

    // The terms of list entry k added to the running sums, every product
    // fused into the sum it feeds: a = m d / r^3 + 2.5 (d.Q.d) d / r^7 - Q d / r^5
    // and phi = -m / r - 0.5 (d.Q.d) / r^5, d from the target to the cell.
    auto cell_term = [&](int k, T& fx, T& fy, T& fz, T& fp) {
      const V4<T> a = ca[k];
      const V4<T> b = cb[k];
      const V2<T> q = cq[k];
      const T dx = a.x - tx, dy = a.y - ty, dz = a.z - tz;
      const T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, e2)));
      const T ri = rsq<T>(r2);
      const T ri2 = ri * ri;
      const T ri3 = ri2 * ri;
      const T ri5 = ri3 * ri2;
      const T qdx = fma(b.w, dz, fma(b.z, dy, b.x * dx));
      const T qdy = fma(q.x, dz, fma(b.y, dy, b.z * dx));
      const T qdz = fma(q.y, dz, fma(q.x, dy, b.w * dx));
      const T dqd = fma(dz, qdz, fma(dy, qdy, dx * qdx));
      const T u = dqd * ri5;
      const T s = fma((T)2.5 * ri2, u, a.w * ri3);
      fx = fma(-ri5, qdx, fma(s, dx, fx));
      fy = fma(-ri5, qdy, fma(s, dy, fy));
      fz = fma(-ri5, qdz, fma(s, dz, fz));
      if (POT) fp = fma((T)-0.5, u, fma(-a.w, ri, fp));
    };
    auto body_term = [&](int k, T& fx, T& fy, T& fz, T& fp) {
      const V4<T> b = pl[k];
      const T dx = b.x - tx, dy = b.y - ty, dz = b.z - tz;
      const T r2 = fma(dz, dz, fma(dy, dy, fma(dx, dx, e2)));
      const T ri = rsq<T>(r2);
      const T mri = b.w * ri;
      const T mr3 = mri * (ri * ri);
      fx = fma(mr3, dx, fx);
      fy = fma(mr3, dy, fy);
      fz = fma(mr3, dz, fz);
      if (POT) fp -= mri;
    };
    // A group of more than 32 targets has one lane per target, and every
    // lane takes every entry: the loop is wave-uniform (scalar control, one
    // LDS address for all lanes).
    auto flush_cells = [&]() {
      __syncthreads();
      T fx = 0, fy = 0, fz = 0, fp = 0;
      if (parts == 1) {
#pragma unroll 2
        for (int k = 0; k < nc; ++k) cell_term(k, fx, fy, fz, fp);
      } else {
        for (int k = part; k < nc; k += parts) cell_term(k, fx, fy, fz, fp);
        combine_parts(fx, fy, fz, fp, width);
      }
      Ax += (double)fx; Ay += (double)fy; Az += (double)fz;
      if (POT) P += (double)fp;
      if (COUNT) ncell_tot += nc;
      __syncthreads();
      nc = 0;
    };
    auto flush_parts = [&]() {
      __syncthreads();
      T fx = 0, fy = 0, fz = 0, fp = 0;
      if (parts == 1) {
#pragma unroll 2
        for (int k = 0; k < np; ++k) body_term(k, fx, fy, fz, fp);
      } else {
        for (int k = part; k < np; k += parts) body_term(k, fx, fy, fz, fp);
        combine_parts(fx, fy, fz, fp, width);
      }
      Ax += (double)fx; Ay += (double)fy; Az += (double)fz;
      if (POT) P += (double)fp;
      if (COUNT) npart_tot += np;
      __syncthreads();
      np = 0;
    };

    int32_t cur = lane == 0 ? root : total;                  // this lane's range [cur, end)
    int32_t end = total;
    int sp = 0;                                              // donated ranges on the stack
    for (;;) {
      // idle lanes take ranges from the top of the stack
      const bool idle = cur >= end;
      const uint64_t im = __ballot(idle);
      const int take = __popcll(im) < sp ? (int)__popcll(im) : sp;
      if (idle) {
        const int r = (int)__popcll(im & ltmask);
        if (r < take) { cur = stk_b[sp - 1 - r]; end = stk_e[sp - 1 - r]; }
      }
      sp -= take;
      const bool busy = cur < end;
      if (!__ballot(busy)) break;
      if (COUNT) ++niter;
      if (nc > LIST - kWave) flush_cells();

      bool acc = false, bucket = false, donate = false;
      int32_t cnt = 0, first = 0, nskip = 0;
      double ncx = 0, ncy = 0, ncz = 0;
      if (busy) {
        const PhysNode nd = pn[cur];
        ncx = nd.cx; ncy = nd.cy; ncz = nd.cz;
        const float rx = (float)(ncx - gcx), ry = (float)(ncy - gcy), rz = (float)(ncz - gcz);
        const float dx = fmaxf(fabsf(rx) - hx, 0.f), dy = fmaxf(fabsf(ry) - hy, 0.f), dz = fmaxf(fabsf(rz) - hz, 0.f);
        const float d2 = dx * dx + dy * dy + dz * dz;
        cnt = nd.count;
        first = nd.first;
        nskip = nd.skip > cur ? nd.skip : cur + 1;           // skip > cur by construction
        if (d2 > nd.rc2) {
          acc = true;
          if (amag) {
            const float b2 = nd.rc2 * th2;
            const float d4 = d2 * d2;
            acc = nd.m * nd.m * (b2 * b2 * b2) <= arel * (d4 * d4 * d2);
          }
        }
        if (!acc) {
          if (cnt <= nleaf) bucket = true;
          else donate = nskip < end;                         // open; the rest of the range can go
        }
      }
      // accepted cells -> cell list
      const uint64_t am = __ballot(acc);
      if (acc) {
        const int pos = nc + (int)__popcll(am & ltmask);
        const PhysNode& nd = pn[cur];
        ca[pos] = V4<T>{(T)(ncx - gcx), (T)(ncy - gcy), (T)(ncz - gcz), (T)nd.m};
        cb[pos] = V4<T>{(T)nd.qxx, (T)nd.qyy, (T)nd.qxy, (T)nd.qxz};
        cq[pos] = V2<T>{(T)nd.qyz, (T)(-((double)nd.qxx + (double)nd.qyy))};
      }
      nc += (int)__popcll(am);
      // opened nodes hand [skip, end) to the stack while it has room
      const uint64_t dm = __ballot(donate);
      const int room = kPhysStack - sp;
      if (donate) {
        const int r = (int)__popcll(dm & ltmask);
        if (r < room) { stk_b[sp + r] = nskip; stk_e[sp + r] = end; end = nskip; }
      }
      sp += (int)__popcll(dm) < room ? (int)__popcll(dm) : room;
      if (busy) cur = (acc || bucket) ? nskip : cur + 1;
      if (COUNT) nopen += (uint32_t)__popcll(__ballot(busy && !acc && !bucket));

      // buckets -> body list, all lanes loading bodies in parallel
      const uint32_t cb_n = bucket ? (uint32_t)cnt : 0u;
      const uint32_t bincl = wave_incl_scan_u32(cb_n);
      const int nbody = __builtin_amdgcn_readlane((int)bincl, kWave - 1);   // wave-uniform
      const uint64_t bm = __ballot(bucket);
      const int nb = (int)__popcll(bm);
      if (bucket) {
        const int j = (int)__popcll(bm & ltmask);
        bfirst[j] = first;
        bstart[j] = (int)(bincl - cb_n);
      }
      __syncthreads();
      for (int base = 0; base < nbody; base += kWave) {
        if (np > LIST - kWave) flush_parts();
        const int t = base + lane;
        if (t < nbody) {
          int lo_j = 0, hi_j = nb - 1;                       // last bucket with bstart <= t
          while (lo_j < hi_j) {
            const int mid = (lo_j + hi_j + 1) >> 1;
            if (bstart[mid] <= t) lo_j = mid; else hi_j = mid - 1;
          }
          const int32_t b = bfirst[lo_j] + (t - bstart[lo_j]);
          pl[np + lane] = V4<T>{(T)(st.x[b] - gcx), (T)(st.y[b] - gcy), (T)(st.z[b] - gcz), (T)st.m[b]};
        }
        np += nbody - base < kWave ? nbody - base : kWave;
      }
      __syncthreads();
    }
    if (nc) flush_cells();
    if (np) flush_parts();

    double v2 = 0.0;
    if (valid) {
      ax[i] = Ax; ay[i] = Ay; az[i] = Az;
      if (POT) pot[i] = P + st.m[i] / eps;                   // drop the self pair, -m_i/eps
      if (KICK) {
        const double half = sy.dt[g.sys] * 0.5;
        const double vx = st.vx[i] + Ax * half;
        const double vy = st.vy[i] + Ay * half;
        const double vz = st.vz[i] + Az * half;
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
      if (lane == 0 && v2 > 0.0) atomicMax(&sy.v2max[g.sys], (unsigned long long)__double_as_longlong(v2));
    }
    if (COUNT && lane == 0) {
      const uint32_t nv = (uint32_t)g.count;
      atomicAdd(&sc->n_pc, (unsigned long long)ncell_tot * nv);
      atomicAdd(&sc->n_pp, (unsigned long long)npart_tot * nv);
      atomicAdd(&sc->n_open, (unsigned long long)nopen);
      atomicAdd(&sc->n_iter, (unsigned long long)niter);
    }
    __syncthreads();                                       // the next group reuses the LDS lists
  }
}

// FP64 direct sum: each thread owns one target; source bodies are staged
// through LDS in tiles of one block.  Only bodies of the target's own system
// act on it; the block loops over the union of its targets' systems.
__global__ __launch_bounds__(kBlock) void k_direct64(int64_t n, StateArrays st, SysArrays sy, double e2,
                                                     double* __restrict__ ax, double* __restrict__ ay,
                                                     double* __restrict__ az, double* __restrict__ pot) {
  __shared__ double sx[kBlock], sy_[kBlock], sz[kBlock], sm[kBlock];
  __shared__ int32_t ss[kBlock];
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  const bool valid = i < n;
  const int64_t last = ((int64_t)blockIdx.x + 1) * kBlock - 1 < n ? ((int64_t)blockIdx.x + 1) * kBlock - 1 : n - 1;
  const int32_t me = valid ? st.sys[i] : -1;
  const int64_t lo = sy.soff[st.sys[(int64_t)blockIdx.x * kBlock]];
  const int64_t hi = sy.soff[st.sys[last] + 1];
  const double px = valid ? st.x[i] : 0, py = valid ? st.y[i] : 0, pz = valid ? st.z[i] : 0;
  double a0 = 0, a1 = 0, a2 = 0, p = 0;
  for (int64_t t = lo; t < hi; t += kBlock) {
    const int64_t j = t + threadIdx.x;
    const bool in = j < hi;
    sx[threadIdx.x] = in ? st.x[j] : 0;
    sy_[threadIdx.x] = in ? st.y[j] : 0;
    sz[threadIdx.x] = in ? st.z[j] : 0;
    sm[threadIdx.x] = in ? st.m[j] : 0;
    ss[threadIdx.x] = in ? st.sys[j] : -2;
    __syncthreads();
    const int lim = (int)(hi - t < kBlock ? hi - t : kBlock);
    for (int k = 0; k < lim; ++k) {
      if (t + k == i || ss[k] != me) continue;
      const double dx = sx[k] - px, dy = sy_[k] - py, dz = sz[k] - pz;
      const double r2 = dx * dx + dy * dy + dz * dz + e2;
      const double ri = 1.0 / sqrt(r2);
      const double mr3 = sm[k] * ri * ri * ri;
      a0 += dx * mr3; a1 += dy * mr3; a2 += dz * mr3;
      p -= sm[k] * ri;
    }
    __syncthreads();
  }
  if (valid) {
    ax[i] = a0; ay[i] = a1; az[i] = a2;
    if (pot) pot[i] = p;
  }
}

// Energy and momentum per system, over the segmented blocks of the
// bounding-box reduction; each system's blocks are summed in block order.
__global__ __launch_bounds__(kBlock) void k_energy_partial(StateArrays st, const double* __restrict__ pot,
                                                           const BoxBlock* __restrict__ blocks,
                                                           double* __restrict__ part) {
  __shared__ double red[kBlock / kWave][5];
  const BoxBlock b = blocks[blockIdx.x];
  double v[5] = {0, 0, 0, 0, 0};
  for (int64_t i = b.begin + threadIdx.x; i < b.end; i += kBlock) {
    const double m = st.m[i], vx = st.vx[i], vy = st.vy[i], vz = st.vz[i];
    v[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    if (pot) v[1] += 0.5 * m * pot[i];
    v[2] += m * vx; v[3] += m * vy; v[4] += m * vz;
  }
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1)
    for (int k = 0; k < 5; ++k) v[k] += __shfl_xor(v[k], off, kWave);
  if (lane_id() == 0)
    for (int k = 0; k < 5; ++k) red[threadIdx.x / kWave][k] = v[k];
  __syncthreads();
  if (threadIdx.x < 5) {
    double t = 0;
    for (int w = 0; w < kBlock / kWave; ++w) t += red[w][threadIdx.x];
    part[5 * blockIdx.x + threadIdx.x] = t;
  }
}

__global__ __launch_bounds__(kBlock) void k_energy_final(const double* __restrict__ part,
                                                         const int32_t* __restrict__ block_off, int nsys,
                                                         double* out) {
  const int s = blockIdx.x * kBlock + threadIdx.x;
  if (s >= nsys) return;
  double t[5] = {0, 0, 0, 0, 0};
  for (int b = block_off[s]; b < block_off[s + 1]; ++b)
    for (int k = 0; k < 5; ++k) t[k] += part[5 * b + k];
  for (int k = 0; k < 5; ++k) out[5 * s + k] = t[k];
}

template <typename T, int LIST>
void walk_dispatch(unsigned nb, const Group* groups, const PhysNode* pn, const StateArrays& st, const SysArrays& sy,
                   double eps, int nleaf, double* ax, double* ay, double* az, double* pot, DevScalars* sc, bool kick,
                   bool count, const float* amag, float rel2, float th2, hipStream_t s) {
#define NBX_PW(K, P, C)                                                                                          \
  hipLaunchKernelGGL((k_phys_walk<T, LIST, K, P, C>), dim3(nb), dim3(kPhysBlock), 0, s, groups, pn, st, sy, eps, \
                     nleaf, ax, ay, az, pot, sc, amag, rel2, th2)
  const bool p = pot != nullptr;
  if (kick) {
    if (count) NBX_PW(true, false, true); else NBX_PW(true, false, false);
  } else if (p) {
    if (count) NBX_PW(false, true, true); else NBX_PW(false, true, false);
  } else {
    if (count) NBX_PW(false, false, true); else NBX_PW(false, false, false);
  }
#undef NBX_PW
}

template <typename T>
void walk_dispatch_list(int list, unsigned nb, const Group* groups, const PhysNode* pn, const StateArrays& st,
                        const SysArrays& sy, double eps, int nleaf, double* ax, double* ay, double* az, double* pot,
                        DevScalars* sc, bool kick, bool count, const float* amag, float rel2, float th2,
                        hipStream_t s) {
  switch (list) {
    case 96: walk_dispatch<T, 96>(nb, groups, pn, st, sy, eps, nleaf, ax, ay, az, pot, sc, kick, count, amag, rel2,
                                  th2, s); break;
    case 192: walk_dispatch<T, 192>(nb, groups, pn, st, sy, eps, nleaf, ax, ay, az, pot, sc, kick, count, amag,
                                    rel2, th2, s); break;
    case 256: walk_dispatch<T, 256>(nb, groups, pn, st, sy, eps, nleaf, ax, ay, az, pot, sc, kick, count, amag,
                                    rel2, th2, s); break;
    default: walk_dispatch<T, 128>(nb, groups, pn, st, sy, eps, nleaf, ax, ay, az, pot, sc, kick, count, amag,
                                   rel2, th2, s); break;
  }
}

}  // namespace

// Group table in body order from the marks (gmark holds the exclusive scan
// of the marks by now; gcnt != 0 marks a group start).
__global__ __launch_bounds__(kBlock) void k_group_compact(int64_t n, const int32_t* __restrict__ gcnt,
                                                          const uint32_t* __restrict__ gidx,
                                                          const int32_t* __restrict__ body_sys,
                                                          Group* __restrict__ groups) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t c = gcnt[i];
  if (c != 0) groups[gidx[i]] = Group{(int32_t)i, c, body_sys[i], 0};
}

void launch_phys_emit(int64_t n, const int32_t* lcp, const uint32_t* pre_off, const uint32_t* part,
                      const StateArrays& st, const DevScalars* sc, PhysNode* pn, int32_t* level_list,
                      uint32_t* end_list, int32_t* gcnt, uint32_t* gmark, uint32_t cap_nodes, uint32_t cap_int,
                      hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_phys_emit, dim3(grid_of(n, kBlock)), dim3(kBlock), 0, s, n, lcp, pre_off, part, st, sc, pn,
                       level_list, end_list, gcnt, gmark, cap_nodes, cap_int);
}

void launch_phys_pair(int64_t m, const int32_t* level_list, const uint32_t* end_list, PhysNode* pn, hipStream_t s) {
  if (m > 0)
    hipLaunchKernelGGL(k_phys_pair, dim3(grid_of(m, kBlock)), dim3(kBlock), 0, s, m, level_list, end_list, pn);
}

void launch_phys_pair_dev(unsigned grid, const int32_t* level_list, const uint32_t* end_list, PhysNode* pn,
                          const DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_phys_pair_dev, dim3(grid), dim3(kBlock), 0, s, level_list, end_list, pn, sc);
}

void launch_phys_upward_tail(int d_lo, const int32_t* level_list, PhysNode* pn, void* aux, double theta,
                             int group_size, int32_t* gcnt, uint32_t* gmark, const DevScalars* sc, hipStream_t s) {
  if (d_lo < kMaxLevels)
    hipLaunchKernelGGL(k_phys_upward_tail, dim3(1), dim3(kBlock), 0, s, d_lo, level_list, pn,
                       reinterpret_cast<PhysAux*>(aux), 1.0 / theta, group_size, gcnt, gmark, sc);
}

void launch_phys_upward_level(unsigned grid, int d, const int32_t* level_list, PhysNode* pn, void* aux, double theta,
                              int group_size, int32_t* gcnt, uint32_t* gmark, const DevScalars* sc, hipStream_t s) {
  hipLaunchKernelGGL(k_phys_upward_level, dim3(grid), dim3(kBlock), 0, s, d, level_list, pn,
                     reinterpret_cast<PhysAux*>(aux), 1.0 / theta, group_size, gcnt, gmark, sc);
}

void launch_phys_upward(const int32_t* level_list, int64_t beg, int64_t cnt, int depth, PhysNode* pn, void* aux,
                        double theta, int group_size, int32_t* gcnt, uint32_t* gmark, hipStream_t s) {
  if (cnt > 0)
    hipLaunchKernelGGL(k_phys_upward, dim3(grid_of(cnt, kBlock)), dim3(kBlock), 0, s, level_list, beg, cnt,
                       depth == 0, pn, reinterpret_cast<PhysAux*>(aux), 1.0 / theta, group_size, gcnt, gmark);
}

void launch_group_compact(int64_t n, const int32_t* gcnt, uint32_t* gmark, const int32_t* body_sys, Group* groups,
                          uint32_t* ngroups, uint32_t* scratch, hipStream_t s) {
  if (n <= 0) return;
  scan_u32(gmark, gmark, n, ngroups, scratch, s);
  hipLaunchKernelGGL(k_group_compact, dim3(grid_of(n, kBlock)), dim3(kBlock), 0, s, n, gcnt, gmark, body_sys, groups);
}

void launch_phys_walk(const Group* groups, int64_t grid_groups, const PhysNode* pn, const StateArrays& st,
                      const SysArrays& sy, double eps, int nleaf, double* ax, double* ay, double* az, double* pot,
                      DevScalars* sc, bool kick, bool count, bool fp64, const float* amag, double alpha,
                      double theta, int list, hipStream_t s) {
  if (grid_groups <= 0) return;
  const unsigned nb = (unsigned)(div_up(grid_groups, 8) * 8);
  const float rel2 = (float)(alpha * alpha), th2 = (float)(theta * theta);
  if (alpha <= 0.0) amag = nullptr;
  if (fp64)
    walk_dispatch_list<double>(list, nb, groups, pn, st, sy, eps, nleaf, ax, ay, az, pot, sc, kick, count, amag, rel2,
                               th2, s);
  else
    walk_dispatch_list<float>(list, nb, groups, pn, st, sy, eps, nleaf, ax, ay, az, pot, sc, kick, count, amag, rel2,
                              th2, s);
}

void launch_direct64(int64_t n, const StateArrays& st, const SysArrays& sy, double eps, double* ax, double* ay,
                     double* az, double* pot, hipStream_t s) {
  if (n > 0)
    hipLaunchKernelGGL(k_direct64, dim3(grid_of(n, kBlock)), dim3(kBlock), 0, s, n, st, sy, eps * eps, ax, ay, az,
                       pot);
}

void launch_energy(const StateArrays& st, const double* pot, const BoxBlock* blocks, int nblocks,
                   const int32_t* block_off, int nsys, double* part, double* out, hipStream_t s) {
  if (nblocks > 0) hipLaunchKernelGGL(k_energy_partial, dim3(nblocks), dim3(kBlock), 0, s, st, pot, blocks, part);
  hipLaunchKernelGGL(k_energy_final, dim3(grid_of(nsys, kBlock)), dim3(kBlock), 0, s, part, block_off, nsys, out);
}

}  // namespace nbx
