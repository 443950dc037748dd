// nbx_accel_noalloc: reference-mode accelerations (bhAccel over all targets,
// octree.h:133-179, the kernels of nbx_accel) of bodies the caller already
// holds in device memory, for callers that cannot allocate device memory.
//
// Every buffer beyond the caller's inputs and outputs lives in `scratch`, a
// static device array of this file (kNoAllocBytes, held by each GPU the
// library runs on until the process ends).  Its device address is looked up
// once per device with hipGetSymbolAddress, with that device current, and
// only that address is handed to kernels: in host code `scratch` names a
// host-side stand-in, not device memory.  Nothing here calls hipMalloc,
// hipFree or an allocating library routine.
//
// All calls on one device share its array; a lock per device serialises
// them, and a call returns only after its work on the stream has finished,
// so no two calls in this process use the array at once.  The array keeps
// the last call's contents: everything a call reads is written by that call
// first (the scalar block is cleared by a kernel at the start).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>

#include "common.hpp"
#include "layout.hpp"
#include "nbx.h"
#include "prims.hpp"
#include "ref_kernels.hpp"

namespace nbx {
namespace {

constexpr int64_t kNaBodies = NBX_NOALLOC_MAX_BODIES;
constexpr int64_t kNaNodes = NBX_NOALLOC_MAX_NODES;   // tree nodes, leaves included
constexpr int64_t kNaGroups = (kNaBodies + kWave - 1) / kWave;
constexpr int64_t kNaBoxes = (kNaBodies + kBoxSpan - 1) / kBoxSpan;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of every buffer inside the array, for kNaBodies bodies and
// kNaNodes nodes.
struct Layout {
  size_t sc, st[2][10], acc[3], key[2], idx[2], key2[2], lcp, cnt_s, part, sortscr, nodes, heads, level_list, end_list;
  size_t root, dt, v2, soff, sroot, send, groups, boxes, box_off, box_part, total;
};

constexpr Layout make_layout() {
  Layout L{};
  size_t o = 0;
  auto take = [&o](size_t bytes) { const size_t at = o; o += align256(bytes); return at; };
  L.sc = take(sizeof(DevScalars));
  for (int w = 0; w < 2; ++w)
    for (int k = 0; k < 10; ++k) L.st[w][k] = take((size_t)kNaBodies * (k < 7 ? 8 : 4));
  for (int k = 0; k < 3; ++k) L.acc[k] = take((size_t)kNaBodies * 8);
  for (int w = 0; w < 2; ++w) {
    L.key[w] = take((size_t)kNaBodies * 8);
    L.idx[w] = take((size_t)kNaBodies * 4);
    L.key2[w] = take((size_t)kNaBodies * 8);
  }
  L.lcp = take((size_t)kNaBodies * 4);
  L.cnt_s = take((size_t)kNaBodies * 4);
  L.part = take((size_t)(((kNaBodies + kBlock - 1) / kBlock) * kPartW) * 4);
  const int64_t big = kNaBodies > kNaNodes ? kNaBodies : kNaNodes;
  const size_t sw = sort_scratch_words(big);
  const size_t cw = (size_t)((big + kScanTile - 1) / kScanTile + 2);
  L.sortscr = take((sw > cw ? sw : cw) * 4);
  L.nodes = take((size_t)(kNaNodes + 1) * sizeof(RefNode));   // one spare record for the walk
  L.heads = take((size_t)(kNaNodes + 1) * sizeof(WalkHead));
  L.level_list = take((size_t)kNaNodes * 4);
  L.end_list = take((size_t)kNaNodes * 4);
  L.root = take(4 * 8);
  L.dt = take(8);
  L.v2 = take(8);
  L.soff = take(2 * 4);
  L.sroot = take(4);
  L.send = take(4);
  L.groups = take((size_t)kNaGroups * sizeof(Group));
  L.boxes = take((size_t)kNaBoxes * sizeof(BoxBlock));
  L.box_off = take(2 * 4);
  L.box_part = take((size_t)kNaBoxes * 6 * 8);
  L.total = o;
  return L;
}

constexpr Layout kLayout = make_layout();
constexpr size_t kNoAllocBytes = 16u << 20;
static_assert(kLayout.total <= kNoAllocBytes, "scratch layout exceeds the array");

}  // namespace
}  // namespace nbx

static __device__ __align__(256) unsigned char scratch[nbx::kNoAllocBytes];

namespace nbx {
namespace {

constexpr int kMaxDevices = 64;
std::mutex g_lock[kMaxDevices];
unsigned char* g_ptr[kMaxDevices] = {};   // device address of `scratch` per device (under g_lock)

// Caller's bodies into the first state buffers (velocities zero, upload
// index, system 0), and the one-system tables: walk groups of 64 bodies,
// bounding-box blocks, body offsets.
__global__ __launch_bounds__(kBlock) void k_na_setup(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ y, const double* __restrict__ z,
                                                     const double* __restrict__ m, const int32_t* __restrict__ type,
                                                     StateArrays st, Group* groups, BoxBlock* boxes,
                                                     int32_t* box_off, int32_t* soff) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) {
    st.x[i] = x[i]; st.y[i] = y[i]; st.z[i] = z[i];
    st.vx[i] = 0.0; st.vy[i] = 0.0; st.vz[i] = 0.0;
    st.m[i] = m[i]; st.type[i] = type[i];
    st.id[i] = (int32_t)i; st.sys[i] = 0;
  }
  const int64_t ng = (n + kWave - 1) / kWave, nb = (n + kBoxSpan - 1) / kBoxSpan;
  if (i < ng) {
    const int32_t f = (int32_t)(i * kWave);
    groups[i] = Group{f, (int32_t)std::min<int64_t>(kWave, n - f), 0, 0};
  }
  if (i < nb) {
    const int32_t b = (int32_t)(i * kBoxSpan);
    boxes[i] = BoxBlock{0, b, (int32_t)std::min<int64_t>(b + kBoxSpan, n), 0};
  }
  if (i == 0) { box_off[0] = 0; box_off[1] = (int32_t)nb; soff[0] = 0; soff[1] = (int32_t)n; }
}

// Accelerations from key order back to the caller's order.
__global__ __launch_bounds__(kBlock) void k_na_scatter(int64_t n, const int32_t* __restrict__ id,
                                                       const double* __restrict__ a0, const double* __restrict__ a1,
                                                       const double* __restrict__ a2, double* __restrict__ ax,
                                                       double* __restrict__ ay, double* __restrict__ az) {
  const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= n) return;
  const int32_t o = id[i];
  ax[o] = a0[i]; ay[o] = a1[i]; az[o] = a2[i];
}

struct Arena {
  unsigned char* base;
  template <class T>
  T* at(size_t off) const { return reinterpret_cast<T*>(base + off); }
  StateArrays state(int w) const {
    const size_t* o = kLayout.st[w];
    return StateArrays{at<double>(o[0]), at<double>(o[1]), at<double>(o[2]), at<double>(o[3]), at<double>(o[4]),
                       at<double>(o[5]), at<double>(o[6]), at<int32_t>(o[7]), at<int32_t>(o[8]), at<int32_t>(o[9])};
  }
  SysArrays sys() const {
    return SysArrays{at<double>(kLayout.root), at<double>(kLayout.dt), at<unsigned long long>(kLayout.v2),
                     at<int32_t>(kLayout.soff), at<int32_t>(kLayout.sroot), at<int32_t>(kLayout.send), 1};
  }
};

nbx_status hipst(hipError_t e) {
  if (e == hipSuccess) return NBX_OK;
  (void)hipGetLastError();
  return NBX_ERR_HIP;
}

#define NA_HIP(expr)                              \
  do {                                            \
    const nbx_status s_ = hipst(expr);            \
    if (s_ != NBX_OK) return s_;                  \
  } while (0)

// Scalars of the build into host memory the caller's thread owns (the
// stack); the copy is complete when the stream has drained.
nbx_status read_scalars(const DevScalars* sc, DevScalars* h, hipStream_t s) {
  NA_HIP(hipMemcpyAsync(h, sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
  NA_HIP(hipStreamSynchronize(s));
  return NBX_OK;
}

// One tree build and walk of n bodies (n >= 1) in the arena: the sequence
// of build_pre / build_post / walk of the context path, with every size
// checked against the arena instead of grown.
nbx_status accel_in_arena(const Arena& A, int64_t n, const double* x, const double* y, const double* z,
                          const double* m, const int32_t* type, double theta, double* ax, double* ay, double* az,
                          hipStream_t s) {
  DevScalars* sc = A.at<DevScalars>(kLayout.sc);
  SysArrays sy = A.sys();
  uint32_t* scr = A.at<uint32_t>(kLayout.sortscr);
  int cur = 0;
  StateArrays St = A.state(cur);
  const int64_t nb = (n + kBoxSpan - 1) / kBoxSpan;
  hipLaunchKernelGGL(k_na_setup, dim3((unsigned)div_up(n, kBlock)), dim3(kBlock), 0, s,
                     n, x, y, z, m, type, St, A.at<Group>(kLayout.groups), A.at<BoxBlock>(kLayout.boxes),
                     A.at<int32_t>(kLayout.box_off), A.at<int32_t>(kLayout.soff));
  launch_bbox(St, A.at<BoxBlock>(kLayout.boxes), (int)nb, A.at<int32_t>(kLayout.box_off),
              A.at<double>(kLayout.box_part), sy, nullptr, s);
  launch_fill_scalars(sc, 1 | 4 | 8, s);
  launch_set_run(sc, 0, s);                                // no abort flag: the walk runs
  uint64_t* key[2] = {A.at<uint64_t>(kLayout.key[0]), A.at<uint64_t>(kLayout.key[1])};
  uint32_t* idx[2] = {A.at<uint32_t>(kLayout.idx[0]), A.at<uint32_t>(kLayout.idx[1])};
  uint64_t* key2[2] = {A.at<uint64_t>(kLayout.key2[0]), A.at<uint64_t>(kLayout.key2[1])};
  launch_keys(n, St, sy, key[0], idx[0], sc, s);
  const bool in1 = radix_sort_u64(key[0], idx[0], key[1], idx[1], n, 0, 63, scr, &sc->err, s);
  launch_gather_state(n, idx[in1 ? 1 : 0], St, A.state(1 - cur), s);
  cur = 1 - cur;
  St = A.state(cur);
  uint64_t* skey = key[in1 ? 1 : 0];
  int32_t* lcp = A.at<int32_t>(kLayout.lcp);
  uint32_t* cnt_s = A.at<uint32_t>(kLayout.cnt_s);
  uint32_t* part = A.at<uint32_t>(kLayout.part);
  uint32_t* tot_nodes = reinterpret_cast<uint32_t*>(&sc->total_nodes);
  // bodies sharing all digits of the first word: ordered by the second
  launch_fix_runs(n, St, sy, skey, key2[0], nullptr, 0, s);
  launch_lcp_count(n, skey, key2[0], St.sys, lcp, cnt_s, part, sc, s);
  scan_u32(cnt_s, cnt_s, n, tot_nodes, scr, s);
  DevScalars h;
  {
    const nbx_status r = read_scalars(sc, &h, s);
    if (r != NBX_OK) return r;
  }
  if (h.err & 1) return NBX_ERR_NONFINITE;
  if (h.err & kSortErrTimeout) return NBX_ERR_HIP;
  if (h.coincident > 0) return NBX_ERR_DEPTH;
  const int64_t total = (uint32_t)h.total_nodes;
  const int64_t mi = internal_count(h);
  if (total > kNaNodes || mi > kNaNodes) return NBX_ERR_NOMEM;   // the tree does not fit the array
  launch_sys_roots(sy, cnt_s, sc, s);
  RefNode* nodes = A.at<RefNode>(kLayout.nodes);
  WalkHead* heads = A.at<WalkHead>(kLayout.heads);
  int32_t* level_list = A.at<int32_t>(kLayout.level_list);
  uint32_t* end_list = A.at<uint32_t>(kLayout.end_list);
  launch_emit(n, lcp, cnt_s, part, St, sy, sc, nodes, heads, level_list, end_list, (uint32_t)(kNaNodes + 1),
              (uint32_t)kNaNodes, s);
  launch_upward_all(h.level_count, level_list, end_list, nodes, heads, theta, St.sys, sy.root, s);
  double* acc[3] = {A.at<double>(kLayout.acc[0]), A.at<double>(kLayout.acc[1]), A.at<double>(kLayout.acc[2])};
  launch_walk(A.at<Group>(kLayout.groups), div_up(n, kWave), nodes, heads, St, sy, theta, acc[0], acc[1], acc[2], sc,
              false, false, s);
  hipLaunchKernelGGL(k_na_scatter, dim3((unsigned)div_up(n, kBlock)), dim3(kBlock), 0, s, n, St.id, acc[0], acc[1],
                     acc[2], ax, ay, az);
  NA_HIP(hipGetLastError());
  NA_HIP(hipStreamSynchronize(s));
  return NBX_OK;
}

}  // namespace
}  // namespace nbx

extern "C" nbx_status nbx_accel_noalloc(void* stream, int64_t n, const double* x, const double* y, const double* z,
                                        const double* m, const int32_t* type, double theta, double* ax, double* ay,
                                        double* az) {
  using namespace nbx;
  if (n < 0 || n > kNaBodies || !(theta >= 0.0) || theta == INFINITY) return NBX_ERR_ARG;
  if (n == 0) return NBX_OK;
  if (!(x && y && z && m && type && ax && ay && az)) return NBX_ERR_ARG;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
    (void)hipGetLastError();
    return NBX_ERR_NODEVICE;
  }
  if (dev >= kMaxDevices) return NBX_ERR_NODEVICE;
  std::lock_guard<std::mutex> lk(g_lock[dev]);
  if (!g_ptr[dev]) {
    // Device address of this device's copy of the array; `scratch` itself
    // is only a host-side handle for the lookup.
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return NBX_ERR_HIP;
    }
    g_ptr[dev] = static_cast<unsigned char*>(p);
  }
  return accel_in_arena(Arena{g_ptr[dev]}, n, x, y, z, m, type, theta, ax, ay, az,
                        static_cast<hipStream_t>(stream));
}
