// C ABI of the engine (include/nbx.h): context, device buffers and the
// host-side orchestration of one KDK step on one HIP stream.
//
// A context holds one or more independent particle systems, stored one after
// another.  One step of all of them (step.h:25-204, with the second tree of
// step n reused as the first tree of step n+1: the walk is deterministic, so
// a(x_{n+1}) from the closing kick of step n is bit-identical to the
// opening-kick force of step n+1):
//
//   kick(a_n, dt/2) + drift(dt)                    one kernel
//   per-system root cube, keys, radix sort by key and then by system,
//   state gather                                   tree of x_{n+1}
//   lcp/counts, scans  -> host reads node counts (one small copy)
//   node emit, depth sorts, skip pairing, upward pass per depth
//   walk over the group table + closing kick(a_{n+1}, dt/2) + max |v|^2
//
// Systems never share a tree node or a walk group, so each system's result
// is the same alone and in a batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "layout.hpp"
#include "nbx.h"
#include "phys_kernels.hpp"
#include "prims.hpp"
#include "ref_kernels.hpp"

using namespace nbx;

namespace {

struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

enum { EV_START, EV_KICK, EV_KEYS, EV_SORT, EV_BUILD, EV_UPWARD, EV_WALK0, EV_WALK, EV_COUNT };
constexpr int kStateArrays = 10;   // x y z vx vy vz m type id sys

}  // namespace

struct nbx_ctx {
  nbx_config cfg{};
  hipStream_t stream = nullptr;
  std::string err;
  int64_t n = 0;
  int64_t cap = 0;                 // particle capacity of the state buffers
  // state in key order, double-buffered for the gather after each sort
  DBuf st[2][kStateArrays];
  int cur = 0;
  DBuf acc[3];
  DBuf key[2], idx[2], key2[2];
  DBuf lcp, cnt_s;
  DBuf nodes, heads;          // reference-mode node records and their walk heads
  DBuf level_list, end_list;  // internal nodes by depth, and their skip targets
  DBuf scratch, dts, lvl_part;
  DBuf pnodes, paux, pot, epart, eout;     // physics-mode tree, moments, potential, energy
  DBuf dacc;                  // forces of the diagnostic passes (potential, direct energy)
  // systems
  std::vector<int64_t> sys_n;              // bodies per system
  DBuf s_root, s_dt, s_v2, s_soff, s_sroot, s_send;
  DBuf groups, boxes, box_off, box_part;
  int64_t ngroups = 0;
  DBuf tgroups;               // physics mode: walk groups packed from the tree
  DBuf gmarks;                // physics mode: per-body group counts and marks
  // Physics mode: the count of walk groups lives on the device
  // (sc->tree_groups); the host knows it after a read-back (else -1) and
  // sizes the walk's grid from the bound meanwhile.
  int64_t ntgroups = 0;
  int64_t groups_bound = 0;
  int nboxes = 0;
  DevScalars* sc = nullptr;   // device
  DevScalars* hsc = nullptr;  // pinned host mirror
  void* hstage = nullptr;     // pinned host staging for device-to-host copies
  size_t hstage_bytes = 0;
  bool physics = false;       // NBX_MODE_PHYSICS
  double alpha = 0.0;         // physics mode: relative opening criterion (0: geometric only)
  int walk_list = kPhysListDefault;   // physics mode: entries of the walk's LDS lists
  DBuf amag[2];               // |a| per body from the opening kick, in key order (FP32)
  int acur = 0;
  bool amag_live = false;     // amag[acur] follows the bodies through this step's sort
  bool acc_valid = false;     // acc holds a(x) of the current positions
  bool tree_valid = false;    // nodes hold the tree of the current positions
  // Physics mode: the body order, node layout and walk groups of the last
  // build still describe the bodies (positions may have moved since), so
  // the tree can be refitted instead of rebuilt.
  bool layout_valid = false;
  int32_t rebuild_every = 1;  // physics mode: full build every k-th step, refits between
  int32_t since_build = 0;    // steps since the last full build
  const uint64_t* skey = nullptr;
  const uint64_t* skey2 = nullptr;
  int64_t total_nodes = 0;
  int32_t max_depth = 0;
  // The bodies are in the key order of the last build (nearly the order of
  // the next one), so the next eager build of a single system may try the
  // patch sort; sort_patched: this build did.
  bool patch_ok = false;
  bool sort_patched = false;
  // After a refusal the next patch_wait builds sort in full; the wait
  // doubles with each refusal in a row (up to 64 builds), so bodies that
  // keep moving too much (a Plummer sphere: about 1 % of them out of order
  // per step) rarely pay for an attempt.
  int32_t patch_wait = 0, patch_backoff = 1;
  int32_t patch_streak = 0;   // eager builds in a row that the path sorted
  int prof = 0;
  hipEvent_t ev[EV_COUNT] = {};
  hipEvent_t ev_scalars = nullptr;  // the build's scalar read-back has landed
  // The layout kernel ran before the host knew the node counts, bounded by
  // these capacities; valid if the tree fits them.
  bool spec_emit = false;
  uint32_t spec_nodes = 0, spec_int = 0;
  nbx_timing tim{};
  // Captured runs (nbx_set_graph): a period of steps whose tree builds never
  // read counts back, replayed from a HIP graph.
  int graph = 0;              // 0 off, 1 on, 2 on without headroom in the bounded sizes
  uint64_t gen = 0;           // bumped when a buffer moves or the systems or settings change
  hipGraphExec_t gexec = nullptr;
  struct GraphKey {
    uint64_t gen;
    double base_dt;
    int adaptive, period;
    int64_t cap_records, cap_list, groups_bound;
    int dmax;
    uint64_t grids;             // hash of the per-depth upward grids and the pair grid
    int patch;                  // builds sorted by the nearly-sorted path
    bool operator==(const GraphKey& o) const {
      return gen == o.gen && base_dt == o.base_dt && adaptive == o.adaptive && period == o.period &&
             cap_records == o.cap_records && cap_list == o.cap_list && groups_bound == o.groups_bound &&
             dmax == o.dmax && grids == o.grids && patch == o.patch;
    }
  } gkey{};
  // Captured builds sort by the nearly-sorted path (bounded_patch) when the
  // last two eager builds did and graph_patch holds: a run in which it
  // refused the keys aborts (kAbortPatch) and is redone step by step, and
  // later captures use the radix sort.
  bool graph_patch = true, bounded_patch = false;
  int64_t gnodes = 0;         // kernels in the captured period
  // sizes of the bounded build (launch_check's limits and the fixed grids)
  int64_t b_cap_records = 0, b_cap_list = 0;
  int b_dmax = -1;
  unsigned b_grid_up[kMaxDepthSlots] = {};   // per depth: blocks of the upward pass
  unsigned b_grid_pair = 1;
  // the state a captured run starts from, restored if one of its builds aborts
  DBuf snap[kStateArrays], snap_acc[3], snap_amag;
};

namespace {

template <class T>
T* P(const DBuf& b) { return static_cast<T*>(b.p); }

nbx_status fail(nbx_ctx* c, nbx_status s, const std::string& msg) {
  if (c) c->err = msg;
  return s;
}

#define HIPCHK(ctx, expr)                                                                        \
  do {                                                                                           \
    hipError_t e_ = (expr);                                                                      \
    if (e_ != hipSuccess)                                                                        \
      return fail(ctx, NBX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));          \
  } while (0)

#define NBXCHK(expr)                          \
  do {                                        \
    nbx_status s_ = (expr);                   \
    if (s_ != NBX_OK) return s_;              \
  } while (0)

// Grow a device buffer to at least `bytes`; its first `keep` bytes are
// copied over (0: contents not kept).  A buffer that has to grow gets 1/8
// extra: tree sizes change a little from step to step, and every
// reallocation (hipMalloc/hipFree) stalls the device for hundreds of
// microseconds.
nbx_status ensure(nbx_ctx* c, DBuf& b, size_t bytes, size_t keep = 0) {
  if (bytes <= b.bytes && b.p) return NBX_OK;
  const size_t want = std::max<size_t>(b.p ? bytes + bytes / 8 : bytes, 256);
  void* p = nullptr;
  hipError_t e = hipMalloc(&p, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(c, NBX_ERR_NOMEM, "device allocation of " + std::to_string(want) + " bytes failed");
  }
  if (keep && b.p) {
    HIPCHK(c, hipMemcpyAsync(p, b.p, std::min(keep, b.bytes), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
  }
  if (b.p) (void)hipFree(b.p);
  b.p = p;
  b.bytes = want;
  ++c->gen;
  return NBX_OK;
}

void release(DBuf& b) {
  if (b.p) (void)hipFree(b.p);   // (callers that keep using the context bump its generation)
  b.p = nullptr;
  b.bytes = 0;
}

// Every device buffer of a context.
template <class F>
void each_buf(nbx_ctx* c, F f) {
  for (auto& w : c->st) for (auto& b : w) f(b);
  for (DBuf* a : {c->acc, c->key, c->idx, c->key2, c->amag})
    for (int k = 0; k < (a == c->acc ? 3 : 2); ++k) f(a[k]);
  for (DBuf* b : {&c->lcp, &c->cnt_s, &c->nodes, &c->heads, &c->level_list, &c->end_list, &c->scratch, &c->dts,
                  &c->lvl_part, &c->pnodes, &c->paux, &c->pot, &c->dacc, &c->epart, &c->eout, &c->s_root, &c->s_dt,
                  &c->s_v2, &c->s_soff, &c->s_sroot, &c->s_send, &c->groups, &c->boxes, &c->box_off, &c->box_part,
                  &c->tgroups, &c->gmarks, &c->snap_amag})
    f(*b);
  for (DBuf& b : c->snap) f(b);
  for (DBuf& b : c->snap_acc) f(b);
}

size_t elem_bytes(int k) { return k < 7 ? 8 : 4; }

uint32_t cap_words(const DBuf& b, size_t esz) { return (uint32_t)std::min<size_t>(b.bytes / esz, UINT32_MAX); }

// Entries the four depth-sort key arrays can hold.
uint32_t cap_int(const nbx_ctx* c) {
  return std::min(cap_words(c->level_list, 4), cap_words(c->end_list, 4));
}

StateArrays view(nbx_ctx* c, int which) {
  DBuf* s = c->st[which];
  return StateArrays{P<double>(s[0]), P<double>(s[1]), P<double>(s[2]), P<double>(s[3]), P<double>(s[4]),
                     P<double>(s[5]), P<double>(s[6]), P<int32_t>(s[7]), P<int32_t>(s[8]), P<int32_t>(s[9])};
}

SysArrays sysa(nbx_ctx* c) {
  return SysArrays{P<double>(c->s_root), P<double>(c->s_dt), P<unsigned long long>(c->s_v2), P<int32_t>(c->s_soff),
                   P<int32_t>(c->s_sroot), P<int32_t>(c->s_send), (int32_t)c->sys_n.size()};
}

void record(nbx_ctx* c, int e) {
  if (c->prof & 1) (void)hipEventRecord(c->ev[e], c->stream);
}

nbx_status sync_scalars(nbx_ctx* c) {
  HIPCHK(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  ++c->tim.host_waits;
  return NBX_OK;
}

// Buffers sized by the particle count; the current state (and the cached
// accelerations) are kept when they grow.
nbx_status ensure_particles(nbx_ctx* c, int64_t n, int64_t keep_n) {
  const int64_t want = std::max<int64_t>(n, 1);
  if (want > c->cap) {
    const int64_t cap = std::max<int64_t>(want, c->cap + c->cap / 2);
    for (int w = 0; w < 2; ++w)
      for (int k = 0; k < kStateArrays; ++k)
        NBXCHK(ensure(c, c->st[w][k], (size_t)cap * elem_bytes(k), w == c->cur ? (size_t)keep_n * elem_bytes(k) : 0));
    for (int k = 0; k < 3; ++k) NBXCHK(ensure(c, c->acc[k], (size_t)cap * 8, (size_t)keep_n * 8));
    for (int w = 0; w < 2; ++w) {
      NBXCHK(ensure(c, c->key[w], (size_t)cap * 8));
      NBXCHK(ensure(c, c->idx[w], (size_t)cap * 4));
      NBXCHK(ensure(c, c->key2[w], (size_t)cap * 8));
    }
    NBXCHK(ensure(c, c->lcp, (size_t)cap * 4));
    NBXCHK(ensure(c, c->cnt_s, (size_t)cap * 4));
    NBXCHK(ensure(c, c->lvl_part, (size_t)lcp_part_words(cap) * 4));
    NBXCHK(ensure(c, c->scratch, std::max(sort_scratch_words(cap), scan_scratch_words(cap)) * 4));
    c->cap = cap;
  }
  return NBX_OK;
}

// Host-built tables for the current systems: body offsets, walk groups (64
// bodies of one system each), bounding-box blocks.
nbx_status set_systems(nbx_ctx* c) {
  const int S = (int)c->sys_n.size();
  std::vector<int32_t> soff(S + 1, 0);
  for (int s = 0; s < S; ++s) soff[s + 1] = soff[s] + (int32_t)c->sys_n[s];
  std::vector<Group> groups;
  std::vector<BoxBlock> boxes;
  std::vector<int32_t> box_off(S + 1, 0);
  for (int s = 0; s < S; ++s) {
    for (int32_t b = soff[s]; b < soff[s + 1]; b += kWave)
      groups.push_back(Group{b, std::min<int32_t>(kWave, soff[s + 1] - b), s, 0});
    for (int32_t b = soff[s]; b < soff[s + 1]; b += kBoxSpan)
      boxes.push_back(BoxBlock{s, b, std::min<int32_t>(b + kBoxSpan, soff[s + 1]), 0});
    box_off[s + 1] = (int32_t)boxes.size();
  }
  const size_t SS = (size_t)std::max(S, 1);
  NBXCHK(ensure(c, c->s_root, SS * 4 * 8));
  NBXCHK(ensure(c, c->s_dt, SS * 8));
  NBXCHK(ensure(c, c->s_v2, SS * 8));
  NBXCHK(ensure(c, c->s_soff, (SS + 1) * 4));
  NBXCHK(ensure(c, c->s_sroot, SS * 4));
  NBXCHK(ensure(c, c->s_send, SS * 4));
  NBXCHK(ensure(c, c->groups, std::max<size_t>(groups.size(), 1) * sizeof(Group)));
  NBXCHK(ensure(c, c->boxes, std::max<size_t>(boxes.size(), 1) * sizeof(BoxBlock)));
  NBXCHK(ensure(c, c->box_off, (SS + 1) * 4));
  NBXCHK(ensure(c, c->box_part, std::max<size_t>(boxes.size(), 1) * 6 * 8));
  NBXCHK(ensure(c, c->epart, std::max<size_t>(boxes.size(), 1) * 5 * 8));
  NBXCHK(ensure(c, c->eout, SS * 5 * 8));
  HIPCHK(c, hipMemcpyAsync(c->s_soff.p, soff.data(), soff.size() * 4, hipMemcpyHostToDevice, c->stream));
  if (!groups.empty())
    HIPCHK(c, hipMemcpyAsync(c->groups.p, groups.data(), groups.size() * sizeof(Group), hipMemcpyHostToDevice,
                             c->stream));
  if (!boxes.empty())
    HIPCHK(c, hipMemcpyAsync(c->boxes.p, boxes.data(), boxes.size() * sizeof(BoxBlock), hipMemcpyHostToDevice,
                             c->stream));
  HIPCHK(c, hipMemcpyAsync(c->box_off.p, box_off.data(), box_off.size() * 4, hipMemcpyHostToDevice, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));   // the host vectors go out of scope
  c->ngroups = (int64_t)groups.size();
  c->nboxes = (int)boxes.size();
  ++c->gen;
  c->acc_valid = false;
  c->tree_valid = false;
  c->layout_valid = false;
  c->max_depth = 0;           // the next build sorts every digit
  c->patch_ok = false;        // bodies in the order they came in
  c->patch_wait = 0;
  c->patch_backoff = 1;
  c->patch_streak = 0;
  c->graph_patch = true;
  return NBX_OK;
}

int sys_bits(int nsys) {
  int b = 0;
  while ((1 << b) < nsys) ++b;
  return b;
}

// Bounding boxes, keys, sort by key and then system, gather of the state
// into key order (flips c->cur): the keys of the new order.  The sort covers
// the first `digits` octree digits of the keys (the depth of a recent build:
// bodies in different leaves of such a tree differ there; 0: all 21 digits of
// the first word), rounded up to whole 8-bit passes; k_fix_runs then orders
// the runs of bodies that agree on those bits by their whole keys, so the
// order is the full keys' whatever the depth of this build.  patch (a single
// system whose bodies are in the last build's key order): patch_sort_u64
// instead of the radix passes, which sorts all 63 bits; when it refuses
// (kSortErrPatch, seen with the node counts) the build is redone with the
// radix sort from the same state (build_tree).
const uint64_t* sort_bodies(nbx_ctx* c, int digits, bool patch) {
  const int64_t n = c->n;
  const int S = (int)c->sys_n.size();
  hipStream_t s = c->stream;
  SysArrays sy = sysa(c);
  StateArrays St = view(c, c->cur);
  launch_bbox(St, P<BoxBlock>(c->boxes), c->nboxes, P<int32_t>(c->box_off), P<double>(c->box_part), sy, c->sc, s);
  launch_keys(n, St, sy, P<uint64_t>(c->key[0]), P<uint32_t>(c->idx[0]), c->sc, s);
  record(c, EV_KEYS);
  uint32_t* scr = P<uint32_t>(c->scratch);
  c->sort_patched = patch && S == 1;
  const int passes = digits > 0 && digits < kKeyDigits ? (3 * digits + 7) / 8 : 8;
  const int lo = c->sort_patched ? 0 : std::max(0, 63 - 8 * passes);
  bool in1 = true;
  if (c->sort_patched)
    patch_sort_u64(P<uint64_t>(c->key[0]), P<uint64_t>(c->key[1]), P<uint32_t>(c->idx[1]), n, &c->sc->ndesc, scr,
                   &c->sc->err, s);
  else
    in1 = radix_sort_u64(P<uint64_t>(c->key[0]), P<uint32_t>(c->idx[0]), P<uint64_t>(c->key[1]),
                         P<uint32_t>(c->idx[1]), n, lo, 63 - lo, scr, &c->sc->err, s);
  uint64_t* skey = P<uint64_t>(c->key[in1 ? 1 : 0]);
  const uint32_t* perm = P<uint32_t>(c->idx[in1 ? 1 : 0]);
  if (S > 1) {
    // stable pass by system: system-major, key order within each system
    uint32_t* pa = P<uint32_t>(c->idx[in1 ? 1 : 0]);
    uint32_t* pb = P<uint32_t>(c->idx[in1 ? 0 : 1]);
    uint32_t* ka = P<uint32_t>(c->key[in1 ? 0 : 1]);       // the free key buffer, as 2n words
    uint32_t* kb = ka + n;
    launch_gather_sys(n, pa, St.sys, ka, s);
    const bool r = radix_sort_u32(ka, pa, kb, pb, n, 0, sys_bits(S), scr, &c->sc->err, s);
    perm = r ? pb : pa;
  }
  launch_gather_state(n, perm, St, view(c, 1 - c->cur), s);
  c->cur = 1 - c->cur;
  St = view(c, c->cur);
  if (c->amag_live) {
    launch_gather_u32(n, perm, P<uint32_t>(c->amag[c->acur]), P<uint32_t>(c->amag[1 - c->acur]), s);
    c->acur = 1 - c->acur;
  }
  if (S > 1) {   // keys of the system-major order
    launch_keys(n, St, sy, P<uint64_t>(c->key[0]), P<uint32_t>(c->idx[0]), c->sc, s);
    skey = P<uint64_t>(c->key[0]);
  }
  // bodies agreeing on the sorted bits: ordered by their whole keys
  launch_fix_runs(n, St, sy, skey, P<uint64_t>(c->key2[0]), c->amag_live ? P<uint32_t>(c->amag[c->acur]) : nullptr,
                  lo, s);
  record(c, EV_SORT);
  return skey;
}

nbx_status build_pre(nbx_ctx* c, const uint64_t** skey_out, bool patch) {
  const int64_t n = c->n;
  hipStream_t s = c->stream;
  const uint64_t* skey = sort_bodies(c, c->max_depth, patch);
  SysArrays sy = sysa(c);
  StateArrays St = view(c, c->cur);
  uint32_t* scr = P<uint32_t>(c->scratch);
  uint32_t* tot_nodes = reinterpret_cast<uint32_t*>(&c->sc->total_nodes);
  launch_lcp_count(n, skey, P<uint64_t>(c->key2[0]), St.sys, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s),
                   P<uint32_t>(c->lvl_part), c->sc, s);
  scan_u32(P<uint32_t>(c->cnt_s), P<uint32_t>(c->cnt_s), n, tot_nodes, scr, s);
  HIPCHK(c, hipMemcpyAsync(c->hsc, c->sc, sizeof(DevScalars), hipMemcpyDeviceToHost, s));
  HIPCHK(c, hipEventRecord(c->ev_scalars, s));
  // Lay out the nodes now, into the buffers of the last build, while the
  // host waits for the node counts; build_post keeps the result if the tree
  // fits those buffers (the kernel writes nothing beyond them).
  c->spec_emit = false;
  const uint32_t ci = cap_int(c);
  if (c->physics && c->pnodes.p && ci > 0) {
    NBXCHK(ensure(c, c->gmarks, (size_t)n * 8));
    launch_fill_u32(2 * n, P<uint32_t>(c->gmarks), 0u, s);
    c->spec_nodes = cap_words(c->pnodes, sizeof(PhysNode));
    c->spec_int = ci;
    launch_phys_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), St, c->sc,
                     P<PhysNode>(c->pnodes), P<int32_t>(c->level_list), P<uint32_t>(c->end_list),
                     P<int32_t>(c->gmarks), P<uint32_t>(c->gmarks) + n, c->spec_nodes,
                     c->spec_int, s);
    c->spec_emit = true;
  } else if (!c->physics && c->nodes.p && c->heads.p && ci > 0) {
    c->spec_nodes = std::min(cap_words(c->nodes, sizeof(RefNode)), cap_words(c->heads, sizeof(WalkHead)));
    c->spec_int = ci;
    launch_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), St, sy, c->sc,
                P<RefNode>(c->nodes), P<WalkHead>(c->heads), P<int32_t>(c->level_list), P<uint32_t>(c->end_list),
                c->spec_nodes, c->spec_int, s);
    c->spec_emit = true;
  }
  *skey_out = skey;
  return NBX_OK;
}

// Reference node records of the current tree (layout and skip pointers;
// the moments come from the upward pass): k_emit (unless `emitted`: already
// run into buffers that hold the tree).  Needs the scalars of the last
// build on the host.
nbx_status ref_layout(nbx_ctx* c, bool emitted) {
  const int64_t n = c->n, total = c->total_nodes;
  hipStream_t s = c->stream;
  // The walk reads node records at 32-bit byte offsets, and one record past
  // the last node (a successor it loads but never uses).
  NBXCHK(ensure(c, c->nodes, (size_t)(total + 1) * sizeof(RefNode)));
  NBXCHK(ensure(c, c->heads, (size_t)(total + 1) * sizeof(WalkHead)));
  RefNode* nodes = P<RefNode>(c->nodes);
  if (!emitted)
    launch_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), view(c, c->cur), sysa(c),
                c->sc, nodes, P<WalkHead>(c->heads), P<int32_t>(c->level_list), P<uint32_t>(c->end_list),
                std::min(cap_words(c->nodes, sizeof(RefNode)), cap_words(c->heads, sizeof(WalkHead))), cap_int(c), s);
  return NBX_OK;
}

// Second half: wait for the scalars, then the node layout and multipoles.
nbx_status build_post(nbx_ctx* c, const uint64_t* skey) {
  const int64_t n = c->n;
  hipStream_t s = c->stream;
  SysArrays sy = sysa(c);
  StateArrays St = view(c, c->cur);
  uint32_t* scr = P<uint32_t>(c->scratch);
  HIPCHK(c, hipEventSynchronize(c->ev_scalars));
  ++c->tim.host_waits;
  if (c->hsc->err & 1) return fail(c, NBX_ERR_NONFINITE, "a particle position is NaN or infinite");
  if (c->hsc->err & kSortErrTimeout)
    return fail(c, NBX_ERR_HIP, "radix sort: a tile's wait for the tiles before it ran out");

  if (c->hsc->coincident > 0)
    return fail(c, NBX_ERR_DEPTH,
                std::to_string(c->hsc->coincident) +
                    " pair(s) of bodies share all 42 octree digits (the tree is limited to 42 levels)");
  const uint64_t* skey2 = P<uint64_t>(c->key2[0]);
  c->skey = skey;
  c->skey2 = skey2;
  launch_sys_roots(sy, P<uint32_t>(c->cnt_s), c->sc, s);

  const int64_t total = (uint32_t)c->hsc->total_nodes;
  const int64_t m = internal_count(*c->hsc);
  c->total_nodes = total;
  c->max_depth = c->hsc->max_leaf_depth;
  // The reference walk reads 128-byte node records at 32-bit byte offsets.
  if (total >= (int64_t)1 << 25)
    return fail(c, NBX_ERR_ARG, "the tree has " + std::to_string(total) + " nodes; at most 33554431 are supported");
  const size_t mm = (size_t)std::max<int64_t>(m, 1) * 4;
  for (DBuf* b : {&c->level_list, &c->end_list}) NBXCHK(ensure(c, *b, mm));
  NBXCHK(ensure(c, c->scratch,
                std::max(sort_scratch_words(std::max(c->cap, m)), scan_scratch_words(std::max(c->cap, total))) * 4));
  int64_t off[kMaxDepthSlots + 1];
  off[0] = 0;
  for (int d = 0; d < kMaxDepthSlots; ++d) off[d + 1] = off[d] + c->hsc->level_count[d];
  if (c->physics) {
    NBXCHK(ensure(c, c->pnodes, (size_t)total * sizeof(PhysNode)));
    NBXCHK(ensure(c, c->paux, (size_t)total * kPhysAuxBytes));
    NBXCHK(ensure(c, c->tgroups, (size_t)n * sizeof(Group)));
    NBXCHK(ensure(c, c->gmarks, (size_t)n * 8));
    int32_t* gcnt = P<int32_t>(c->gmarks);
    uint32_t* gmark = P<uint32_t>(c->gmarks) + n;
    PhysNode* pn = P<PhysNode>(c->pnodes);
    const bool emitted = c->spec_emit && total <= (int64_t)c->spec_nodes && m <= (int64_t)c->spec_int;
    if (!emitted) {
      launch_fill_u32(2 * n, P<uint32_t>(c->gmarks), 0u, s);
      launch_phys_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), St, c->sc, pn,
                       P<int32_t>(c->level_list), P<uint32_t>(c->end_list), gcnt, gmark,
                       cap_words(c->pnodes, sizeof(PhysNode)), cap_int(c), s);
    }
    launch_phys_pair(m, P<int32_t>(c->level_list), P<uint32_t>(c->end_list), pn, s);
    record(c, EV_BUILD);
    for (int d = kMaxDepthSlots - 1; d >= 0; --d)
      if (c->hsc->level_count[d] > 0)
        launch_phys_upward(P<int32_t>(c->level_list), off[d], c->hsc->level_count[d], d, pn, c->paux.p,
                           c->cfg.theta, c->cfg.group_size, gcnt, gmark, s);
    // The walk reads the group count on the device.  After a build of the
    // same systems and settings its grid comes from that build's count (it
    // reached the host with this build's scalars), with room to grow; the
    // first build of a layout reads the count back.
    const bool bounded = c->layout_valid;
    const int64_t last = c->hsc->tree_groups;
    launch_group_compact(n, gcnt, gmark, St.sys, P<Group>(c->tgroups), &c->sc->tree_groups, scr, s);
    if (bounded) {
      c->ntgroups = -1;
      c->groups_bound = last + last / 4 + 64;
    } else {
      NBXCHK(sync_scalars(c));
      c->ntgroups = c->hsc->tree_groups;
    }
    c->layout_valid = true;
  } else {
    NBXCHK(ref_layout(c, c->spec_emit && total + 1 <= (int64_t)c->spec_nodes && m <= (int64_t)c->spec_int));
    record(c, EV_BUILD);
    RefNode* nodes = P<RefNode>(c->nodes);
    launch_upward_all(c->hsc->level_count, P<int32_t>(c->level_list), P<uint32_t>(c->end_list), nodes,
                      P<WalkHead>(c->heads), c->cfg.theta, St.sys, sy.root, s);
  }
  record(c, EV_UPWARD);
  c->tree_valid = true;
  return NBX_OK;
}

// Physics mode: moments of the existing tree for the current positions.
// The bodies keep their order, so the node layout, skip pointers and walk
// groups of the last build still hold; leaf records, centres of mass,
// quadrupoles and bmax are recomputed from the bodies (bmax from their
// actual box), so every opening decision stays conservative however far
// the bodies have moved.  The group marks the passes write are not used.
nbx_status refit_tree(nbx_ctx* c) {
  const int64_t n = c->n;
  hipStream_t s = c->stream;
  record(c, EV_KEYS);
  record(c, EV_SORT);
  int32_t* gcnt = P<int32_t>(c->gmarks);
  uint32_t* gmark = P<uint32_t>(c->gmarks) + n;
  PhysNode* pn = P<PhysNode>(c->pnodes);
  // leaf records from the current positions; the node lists of the last
  // build stand (list capacity 0: nothing written to them)
  launch_phys_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), view(c, c->cur), c->sc, pn,
                   P<int32_t>(c->level_list), P<uint32_t>(c->end_list), gcnt, gmark,
                   cap_words(c->pnodes, sizeof(PhysNode)), 0u, s);
  record(c, EV_BUILD);
  int64_t off[kMaxDepthSlots + 1];
  off[0] = 0;
  for (int d = 0; d < kMaxDepthSlots; ++d) off[d + 1] = off[d] + c->hsc->level_count[d];
  for (int d = kMaxDepthSlots - 1; d >= 0; --d)
    if (c->hsc->level_count[d] > 0)
      launch_phys_upward(P<int32_t>(c->level_list), off[d], c->hsc->level_count[d], d, pn, c->paux.p, c->cfg.theta,
                         c->cfg.group_size, gcnt, gmark, s);
  record(c, EV_UPWARD);
  c->tree_valid = true;
  return NBX_OK;
}

// Tree of the current positions: per-system roots, keys, sort, gather,
// layout, multipoles.
nbx_status build_tree(nbx_ctx* c) {
  const uint64_t* skey = nullptr;
  const bool patch = c->patch_ok && c->patch_wait == 0;
  if (c->patch_wait > 0) --c->patch_wait;
  NBXCHK(build_pre(c, &skey, patch));
  if (c->sort_patched) {
    // (build_post waits for the same read-back first)
    HIPCHK(c, hipEventSynchronize(c->ev_scalars));
    ++c->tim.sort_patched;
    if (c->hsc->err & kSortErrPatch) {                     // the bodies moved too much: radix sort
      c->cur = 1 - c->cur;                                 // the state before the gather
      if (c->amag_live) c->acur = 1 - c->acur;
      c->patch_ok = false;
      c->patch_backoff = std::min(2 * c->patch_backoff, 64);
      c->patch_wait = c->patch_backoff;
      ++c->tim.sort_redos;
      c->patch_streak = 0;
      NBXCHK(build_pre(c, &skey, false));
    } else {
      c->patch_backoff = 1;
      ++c->patch_streak;
    }
  } else {
    c->patch_streak = 0;
  }
  const nbx_status r = build_post(c, skey);
  c->patch_ok = r == NBX_OK && c->sys_n.size() == 1;
  return r;
}

// Grid of the physics walk: the group count if the host has it, else its bound.
int64_t group_grid(const nbx_ctx* c) { return c->ntgroups >= 0 ? c->ntgroups : c->groups_bound; }

nbx_status walk(nbx_ctx* c, bool kick) {
  const bool count = (c->prof & 2) != 0;
  if (count) launch_fill_scalars(c->sc, 8, c->stream);
  record(c, EV_WALK0);
  if (c->physics)
    launch_phys_walk(P<Group>(c->tgroups), group_grid(c), P<PhysNode>(c->pnodes), view(c, c->cur), sysa(c), c->cfg.eps,
                     c->cfg.leaf_size, P<double>(c->acc[0]), P<double>(c->acc[1]), P<double>(c->acc[2]), nullptr,
                     c->sc, kick, count, c->cfg.precision == NBX_PREC_FP64,
                     c->amag_live ? P<float>(c->amag[c->acur]) : nullptr, c->alpha, c->cfg.theta, c->walk_list,
                     c->stream);
  else
    launch_walk(P<Group>(c->groups), c->ngroups, P<RefNode>(c->nodes), P<WalkHead>(c->heads), view(c, c->cur),
                sysa(c), c->cfg.theta,
                P<double>(c->acc[0]), P<double>(c->acc[1]), P<double>(c->acc[2]), c->sc, kick, count, c->stream);
  record(c, EV_WALK);
  HIPCHK(c, hipGetLastError());
  c->acc_valid = true;
  return NBX_OK;
}

float elapsed(nbx_ctx* c, int a, int b) {
  float ms = 0.f;
  if (hipEventElapsedTime(&ms, c->ev[a], c->ev[b]) != hipSuccess) return 0.f;
  return ms;
}

nbx_status finish_timing(nbx_ctx* c, bool with_kick) {
  if (c->prof & 1) {
    HIPCHK(c, hipEventSynchronize(c->ev[EV_WALK]));
    const int k0 = with_kick ? EV_KICK : EV_START;
    if (with_kick) c->tim.kick_ms += elapsed(c, EV_START, EV_KICK);
    c->tim.keys_ms += elapsed(c, k0, EV_KEYS);
    c->tim.sort_ms += elapsed(c, EV_KEYS, EV_SORT);
    c->tim.build_ms += elapsed(c, EV_SORT, EV_BUILD);
    c->tim.upward_ms += elapsed(c, EV_BUILD, EV_UPWARD);
    c->tim.walk_ms += elapsed(c, EV_WALK0, EV_WALK);
    c->tim.total_ms += elapsed(c, EV_START, EV_WALK);
  }
  if (c->prof & 2) {
    NBXCHK(sync_scalars(c));
    c->tim.n_pp += (int64_t)c->hsc->n_pp;
    c->tim.n_pc += (int64_t)c->hsc->n_pc;
    c->tim.n_open += (int64_t)c->hsc->n_open;
    c->tim.n_iter += (int64_t)c->hsc->n_iter;
    c->tim.n_eval += (int64_t)c->hsc->n_eval;
  }
  c->tim.steps += 1;
  c->tim.nodes = c->total_nodes;
  c->tim.depth = c->max_depth;
  return NBX_OK;
}

// a(x) of the current positions, without a kick.
nbx_status ensure_accel(nbx_ctx* c) {
  if (c->acc_valid) return NBX_OK;
  record(c, EV_START);
  NBXCHK(build_tree(c));
  c->since_build = 0;                   // the refit schedule counts from this build
  NBXCHK(walk(c, false));
  return NBX_OK;
}

// One KDK step of every system with the dt held on the device.
nbx_status step_device(nbx_ctx* c) {
  if (c->n == 0) return NBX_OK;
  c->amag_live = false;                 // the first evaluation uses the geometric criterion
  NBXCHK(ensure_accel(c));
  const bool rel = c->physics && c->alpha > 0.0;
  if (rel)
    for (DBuf& b : c->amag) NBXCHK(ensure(c, b, (size_t)c->n * sizeof(float)));
  record(c, EV_START);
  launch_kick_drift(c->n, view(c, c->cur), P<double>(c->acc[0]), P<double>(c->acc[1]), P<double>(c->acc[2]), sysa(c),
                    rel ? P<float>(c->amag[c->acur]) : nullptr, c->stream);
  c->acc_valid = false;
  c->tree_valid = false;
  c->amag_live = rel;
  record(c, EV_KICK);
  const bool refit = c->physics && c->layout_valid && c->since_build + 1 < c->rebuild_every;
  const nbx_status st = refit ? refit_tree(c) : build_tree(c);
  c->since_build = refit ? c->since_build + 1 : 0;
  if (st != NBX_OK) { c->amag_live = false; return st; }
  const nbx_status sw = walk(c, true);
  c->amag_live = false;
  NBXCHK(sw);
  return finish_timing(c, true);
}

// ---- captured runs: steps whose tree builds keep their counts on the device

// Steps in a captured period: the state and the |a| arrays are double
// buffered and flip at every build, so the period holds an even number of
// builds and starts with one.
int graph_period(const nbx_ctx* c) { return c->physics ? 2 * c->rebuild_every : 2; }

// Limits and grids of the bounded builds, from the last build: kept while
// the last build fits them with a margin, else sized with a quarter of
// headroom (no headroom with graph == 2).  The buffers are grown to match.
nbx_status prepare_bounded(nbx_ctx* c, int32_t nsteps) {
  const bool tight = c->graph == 2;
  const int64_t n = c->n, total = c->total_nodes, m = internal_count(*c->hsc);
  const int dneed = c->max_depth - 1;                      // deepest internal node of the last build
  bool keep = !tight && c->b_cap_records >= total + total / 16 + 64 && c->b_cap_list >= m + m / 16 + 64 &&
              c->b_dmax >= dneed;
  for (int d = 0; keep && d <= dneed; ++d)
    keep = (int64_t)c->b_grid_up[d] * kBlock >= (uint32_t)c->hsc->level_count[d];
  if (!keep) {
    const int64_t rec = tight ? total + 1 : total + total / 4 + 1024;
    const int64_t lst = tight ? std::max<int64_t>(m, 1) : m + m / 4 + 1024;
    if (c->physics) {
      NBXCHK(ensure(c, c->pnodes, (size_t)rec * sizeof(PhysNode)));
      NBXCHK(ensure(c, c->paux, (size_t)rec * kPhysAuxBytes));
    } else {
      NBXCHK(ensure(c, c->nodes, (size_t)rec * sizeof(RefNode)));
      NBXCHK(ensure(c, c->heads, (size_t)rec * sizeof(WalkHead)));
    }
    NBXCHK(ensure(c, c->level_list, (size_t)lst * 4));
    NBXCHK(ensure(c, c->end_list, (size_t)lst * 4));
    // depths 0 .. b_dmax launch one grid each; deeper ones run in the tail kernel
    c->b_dmax = tight ? dneed : std::min(dneed + 1, kMaxLevels - 1);
    // each depth's grid covers its count in the last build with an eighth to
    // spare (the kernels stride over whatever count the device holds)
    for (int d = 0; d < kMaxDepthSlots; ++d) {
      const int64_t cnt = (uint32_t)c->hsc->level_count[d];
      const int64_t want = tight ? cnt : cnt + cnt / 8 + 32;
      c->b_grid_up[d] = (unsigned)std::min<int64_t>(std::max<int64_t>(div_up(want, kBlock), 1), 4096);
    }
    c->b_grid_pair = (unsigned)std::min<int64_t>(std::max<int64_t>(div_up(tight ? m : m + m / 8 + 64, kBlock), 1),
                                                 8192);
  }
  c->b_cap_records = c->physics ? std::min<int64_t>(cap_words(c->pnodes, sizeof(PhysNode)), cap_words(c->paux, kPhysAuxBytes))
                                : std::min<int64_t>(cap_words(c->nodes, sizeof(RefNode)), cap_words(c->heads, sizeof(WalkHead)));
  c->b_cap_list = cap_int(c);
  NBXCHK(ensure(c, c->scratch, std::max(sort_scratch_words(std::max(c->cap, c->b_cap_list)),
                                        scan_scratch_words(std::max(c->cap, c->b_cap_records))) * 4));
  NBXCHK(ensure(c, c->dts, (size_t)std::max(nsteps, 1) * 8));
  if (c->physics) {
    NBXCHK(ensure(c, c->tgroups, (size_t)n * sizeof(Group)));
    NBXCHK(ensure(c, c->gmarks, (size_t)n * 8));
    const int64_t last = c->hsc->tree_groups;
    if (tight || c->groups_bound < last + last / 16 + 64)
      c->groups_bound = std::min<int64_t>(tight ? std::max<int64_t>(last, 1) : last + last / 4 + 64, n);
    if (c->alpha > 0.0)
      for (DBuf& b : c->amag) NBXCHK(ensure(c, b, (size_t)n * sizeof(float)));
  }
  return NBX_OK;
}

// Tree of the current positions without a host read-back: launch_check
// compares the counts with the bounded sizes on the device and the tree
// kernels after it do nothing if they do not fit.
void build_bounded(nbx_ctx* c) {
  const int64_t n = c->n;
  hipStream_t s = c->stream;
  const uint64_t* skey = sort_bodies(c, c->max_depth, c->bounded_patch);   // the depth when the run was captured
  SysArrays sy = sysa(c);
  StateArrays St = view(c, c->cur);
  uint32_t* scr = P<uint32_t>(c->scratch);
  launch_lcp_count(n, skey, P<uint64_t>(c->key2[0]), St.sys, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s),
                   P<uint32_t>(c->lvl_part), c->sc, s);
  scan_u32(P<uint32_t>(c->cnt_s), P<uint32_t>(c->cnt_s), n, reinterpret_cast<uint32_t*>(&c->sc->total_nodes), scr,
           s);
  launch_check(c->sc, c->b_cap_records, c->b_cap_list, kMaxLevels, s);   // any depth: the tail kernel
  launch_sys_roots(sy, P<uint32_t>(c->cnt_s), c->sc, s);
  const uint32_t cap_rec = (uint32_t)c->b_cap_records, cap_lst = (uint32_t)c->b_cap_list;
  if (c->physics) {
    int32_t* gcnt = P<int32_t>(c->gmarks);
    uint32_t* gmark = P<uint32_t>(c->gmarks) + n;
    PhysNode* pn = P<PhysNode>(c->pnodes);
    launch_fill_u32(2 * n, P<uint32_t>(c->gmarks), 0u, s);
    launch_phys_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), St, c->sc, pn,
                     P<int32_t>(c->level_list), P<uint32_t>(c->end_list), gcnt, gmark, cap_rec, cap_lst, s);
    launch_phys_pair_dev(c->b_grid_pair, P<int32_t>(c->level_list), P<uint32_t>(c->end_list), pn, c->sc, s);
    launch_phys_upward_tail(c->b_dmax + 1, P<int32_t>(c->level_list), pn, c->paux.p, c->cfg.theta, c->cfg.group_size,
                            gcnt, gmark, c->sc, s);
    for (int d = c->b_dmax; d >= 0; --d)
      launch_phys_upward_level(c->b_grid_up[d], d, P<int32_t>(c->level_list), pn, c->paux.p, c->cfg.theta,
                               c->cfg.group_size, gcnt, gmark, c->sc, s);
    launch_group_compact(n, gcnt, gmark, St.sys, P<Group>(c->tgroups), &c->sc->tree_groups, scr, s);
    c->ntgroups = -1;                                     // the walk's grid is groups_bound
    c->layout_valid = true;
  } else {
    RefNode* nodes = P<RefNode>(c->nodes);
    launch_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), St, sy, c->sc, nodes,
                P<WalkHead>(c->heads), P<int32_t>(c->level_list), P<uint32_t>(c->end_list), cap_rec, cap_lst, s);
    launch_upward_tail(c->b_dmax + 1, P<int32_t>(c->level_list), P<uint32_t>(c->end_list), nodes,
                       P<WalkHead>(c->heads), c->cfg.theta, St.sys, sy.root, c->sc, s);
    for (int d = c->b_dmax; d >= 0; --d)
      launch_upward_level(c->b_grid_up[d], d, P<int32_t>(c->level_list), P<uint32_t>(c->end_list), nodes,
                          P<WalkHead>(c->heads), c->cfg.theta, St.sys, sy.root, c->sc, s);
  }
  c->skey = skey;
  c->skey2 = P<uint64_t>(c->key2[0]);
  c->spec_emit = false;
  c->tree_valid = true;
}

// Physics refit of the last bounded build's tree (its per-depth offsets are
// on the device).
void refit_bounded(nbx_ctx* c) {
  const int64_t n = c->n;
  hipStream_t s = c->stream;
  int32_t* gcnt = P<int32_t>(c->gmarks);
  uint32_t* gmark = P<uint32_t>(c->gmarks) + n;
  PhysNode* pn = P<PhysNode>(c->pnodes);
  launch_phys_emit(n, P<int32_t>(c->lcp), P<uint32_t>(c->cnt_s), P<uint32_t>(c->lvl_part), view(c, c->cur), c->sc, pn,
                   P<int32_t>(c->level_list), P<uint32_t>(c->end_list), gcnt, gmark, (uint32_t)c->b_cap_records, 0u,
                   s);
  launch_phys_upward_tail(c->b_dmax + 1, P<int32_t>(c->level_list), pn, c->paux.p, c->cfg.theta, c->cfg.group_size,
                          gcnt, gmark, c->sc, s);
  for (int d = c->b_dmax; d >= 0; --d)
    launch_phys_upward_level(c->b_grid_up[d], d, P<int32_t>(c->level_list), pn, c->paux.p, c->cfg.theta,
                             c->cfg.group_size, gcnt, gmark, c->sc, s);
  c->tree_valid = true;
}

// One KDK step with no host wait: step_device's sequence with the bounded
// build.  Needs the forces of the current positions and prepare_bounded.
nbx_status step_bounded(nbx_ctx* c) {
  const bool rel = c->physics && c->alpha > 0.0;
  launch_kick_drift(c->n, view(c, c->cur), P<double>(c->acc[0]), P<double>(c->acc[1]), P<double>(c->acc[2]), sysa(c),
                    rel ? P<float>(c->amag[c->acur]) : nullptr, c->stream);
  c->acc_valid = false;
  c->tree_valid = false;
  c->amag_live = rel;
  const bool refit = c->physics && c->since_build + 1 < c->rebuild_every;
  if (refit) refit_bounded(c);
  else build_bounded(c);
  c->since_build = refit ? c->since_build + 1 : 0;
  const nbx_status sw = walk(c, true);
  c->amag_live = false;
  return sw;
}

// The state a captured run starts from (positions, velocities and the rest
// of the body arrays in their current slots, the cached forces and |a|).
struct Snapshot {
  int cur, acur, since_build;
};

nbx_status snapshot(nbx_ctx* c, Snapshot* sn) {
  const size_t n = (size_t)c->n;
  for (int k = 0; k < kStateArrays; ++k) {
    NBXCHK(ensure(c, c->snap[k], n * elem_bytes(k)));
    HIPCHK(c, hipMemcpyAsync(c->snap[k].p, c->st[c->cur][k].p, n * elem_bytes(k), hipMemcpyDeviceToDevice, c->stream));
  }
  for (int k = 0; k < 3; ++k) {
    NBXCHK(ensure(c, c->snap_acc[k], n * 8));
    HIPCHK(c, hipMemcpyAsync(c->snap_acc[k].p, c->acc[k].p, n * 8, hipMemcpyDeviceToDevice, c->stream));
  }
  if (c->physics && c->alpha > 0.0) {
    NBXCHK(ensure(c, c->snap_amag, n * sizeof(float)));
    HIPCHK(c, hipMemcpyAsync(c->snap_amag.p, c->amag[c->acur].p, n * sizeof(float), hipMemcpyDeviceToDevice,
                             c->stream));
  }
  *sn = Snapshot{c->cur, c->acur, c->since_build};
  return NBX_OK;
}

nbx_status restore(nbx_ctx* c, const Snapshot& sn) {
  const size_t n = (size_t)c->n;
  c->cur = sn.cur;
  c->acur = sn.acur;
  for (int k = 0; k < kStateArrays; ++k)
    HIPCHK(c, hipMemcpyAsync(c->st[c->cur][k].p, c->snap[k].p, n * elem_bytes(k), hipMemcpyDeviceToDevice, c->stream));
  for (int k = 0; k < 3; ++k)
    HIPCHK(c, hipMemcpyAsync(c->acc[k].p, c->snap_acc[k].p, n * 8, hipMemcpyDeviceToDevice, c->stream));
  if (c->physics && c->alpha > 0.0)
    HIPCHK(c, hipMemcpyAsync(c->amag[c->acur].p, c->snap_amag.p, n * sizeof(float), hipMemcpyDeviceToDevice,
                             c->stream));
  c->since_build = sn.since_build;   // the next step is a build, as it was
  c->acc_valid = true;
  c->tree_valid = false;
  c->layout_valid = false;
  c->spec_emit = false;
  return NBX_OK;
}

void drop_graph(nbx_ctx* c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  c->gexec = nullptr;
}

// A graph of one period of steps (each: dt from the speeds if adaptive, then
// the step) captured from the context's stream; the host-side state (buffer
// parity, refit phase) ends where it began.
nbx_status capture(nbx_ctx* c, double base_dt, int period) {
  drop_graph(c);
  const int prof = c->prof;
  c->prof = 0;
  hipStreamCaptureMode mode = hipStreamCaptureModeThreadLocal;
  HIPCHK(c, hipThreadExchangeStreamCaptureMode(&mode));
  hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  nbx_status st = e == hipSuccess ? NBX_OK : fail(c, NBX_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
  SysArrays sy = sysa(c);
  for (int p = 0; p < period && st == NBX_OK && e == hipSuccess; ++p) {
    if (c->cfg.adaptive_dt) launch_dt(sy, base_dt, P<double>(c->dts), -1, c->sc, c->stream);
    st = step_bounded(c);
  }
  hipGraph_t g = nullptr;
  if (e == hipSuccess) e = hipStreamEndCapture(c->stream, &g);
  (void)hipThreadExchangeStreamCaptureMode(&mode);
  c->prof = prof;
  if (st == NBX_OK && e != hipSuccess) st = fail(c, NBX_ERR_HIP, std::string("capture: ") + hipGetErrorString(e));
  if (st == NBX_OK) {
    size_t nodes = 0;
    e = hipGraphGetNodes(g, nullptr, &nodes);
    if (e == hipSuccess) e = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) st = fail(c, NBX_ERR_HIP, std::string("graph: ") + hipGetErrorString(e));
    c->gnodes = (int64_t)nodes;
  }
  if (g) (void)hipGraphDestroy(g);
  if (st != NBX_OK) {
    (void)hipGetLastError();
    drop_graph(c);
  }
  return st;
}

// A tree built from the current positions, for the inspection hooks.  A
// refitted physics tree still has the cells of its last build, so it is
// rebuilt here.
nbx_status built_tree(nbx_ctx* c) {
  if (c->tree_valid && !(c->physics && c->since_build > 0)) return NBX_OK;
  NBXCHK(build_tree(c));
  c->acc_valid = false;   // the build reordered the bodies
  c->since_build = 0;
  return NBX_OK;
}

nbx_status check_ctx(nbx_ctx* c) {
  if (!c) return NBX_ERR_ARG;
  HIPCHK(c, hipSetDevice(c->cfg.device));
  return NBX_OK;
}

int64_t sys_offset(const nbx_ctx* c, int s) {
  int64_t o = 0;
  for (int k = 0; k < s; ++k) o += c->sys_n[(size_t)k];
  return o;
}

// Device-to-host copy through the context's pinned staging buffer: the
// stream is drained before the host reads, whatever the runtime does with
// copies into pageable memory.  Returns the staged bytes.
nbx_status fetch(nbx_ctx* c, const void* dev, size_t bytes, const void** staged) {
  if (bytes > c->hstage_bytes) {
    if (c->hstage) (void)hipHostFree(c->hstage);
    c->hstage = nullptr;
    c->hstage_bytes = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 8, 4096);
    HIPCHK(c, hipHostMalloc(&c->hstage, want, hipHostMallocDefault));
    c->hstage_bytes = want;
  }
  if (bytes) HIPCHK(c, hipMemcpyAsync(c->hstage, dev, bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  *staged = c->hstage;
  return NBX_OK;
}

// Copy bodies [off, off + cnt) of a key-order device array out, each placed
// at its upload index minus `base` (`id` from read_ids, checked there to be a
// permutation of [base, base + cnt)).
nbx_status download_range(nbx_ctx* c, const void* dev, size_t esz, int64_t off, int64_t cnt, int64_t base,
                          const std::vector<int32_t>& id, void* out) {
  const void* staged = nullptr;
  NBXCHK(fetch(c, static_cast<const char*>(dev) + off * esz, (size_t)cnt * esz, &staged));
  const char* src = static_cast<const char*>(staged);
  char* o = static_cast<char*>(out);
  for (int64_t k = 0; k < cnt; ++k)
    std::memcpy(o + ((int64_t)id[(size_t)k] - base) * esz, src + (size_t)k * esz, esz);
  return NBX_OK;
}

// Upload indices (`id`) of the bodies in slots [off, off + cnt), which must
// be a permutation of [base, base + cnt): every later scatter by upload index
// relies on it, so anything else is refused (NBX_ERR_STATE) before a byte is
// written to the caller's arrays.
nbx_status read_ids(nbx_ctx* c, int64_t off, int64_t cnt, int64_t base, std::vector<int32_t>& id) {
  const void* staged = nullptr;
  NBXCHK(fetch(c, P<int32_t>(c->st[c->cur][8]) + off, (size_t)cnt * 4, &staged));
  id.assign(static_cast<const int32_t*>(staged), static_cast<const int32_t*>(staged) + cnt);
  std::vector<uint8_t> seen((size_t)cnt, 0);
  for (int64_t k = 0; k < cnt; ++k) {
    const int64_t at = (int64_t)id[(size_t)k] - base;
    if (at < 0 || at >= cnt || seen[(size_t)at])
      return fail(c, NBX_ERR_STATE,
                  "body indices on the device are not a permutation (slot " + std::to_string(off + k) + " holds " +
                      std::to_string(id[(size_t)k]) + ")");
    seen[(size_t)at] = 1;
  }
  return NBX_OK;
}

// Write n bodies of system `sys` into slots [at, at + n); upload indices
// id0, id0 + 1, ...
nbx_status put_bodies(nbx_ctx* c, int64_t at, int64_t n, int32_t sys, int64_t id0, const double* x, const double* y,
                      const double* z, const double* vx, const double* vy, const double* vz, const double* m,
                      const int32_t* type) {
  if (n == 0) return NBX_OK;
  const void* src[8] = {x, y, z, vx, vy, vz, m, type};
  for (int k = 0; k < 8; ++k)
    HIPCHK(c, hipMemcpyAsync(static_cast<char*>(c->st[c->cur][k].p) + at * elem_bytes(k), src[k],
                             (size_t)n * elem_bytes(k), hipMemcpyHostToDevice, c->stream));
  std::vector<int32_t> id((size_t)n), sv((size_t)n, sys);
  for (int64_t i = 0; i < n; ++i) id[(size_t)i] = (int32_t)(id0 + i);
  HIPCHK(c, hipMemcpyAsync(P<int32_t>(c->st[c->cur][8]) + at, id.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipMemcpyAsync(P<int32_t>(c->st[c->cur][9]) + at, sv.data(), (size_t)n * 4, hipMemcpyHostToDevice,
                           c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBX_OK;
}

}  // namespace

extern "C" {

void nbx_config_default(nbx_config* cfg) {
  if (!cfg) return;
  cfg->device = 0;
  cfg->mode = NBX_MODE_REFERENCE;
  cfg->precision = NBX_PREC_FP64;
  cfg->theta = 0.5;
  cfg->eps = 0.01;
  cfg->leaf_size = 16;
  cfg->group_size = 64;
  cfg->adaptive_dt = 1;
}

nbx_status nbx_create(nbx_ctx** out, const nbx_config* cfg) {
  if (!out) return NBX_ERR_ARG;
  *out = nullptr;
  nbx_config c0;
  nbx_config_default(&c0);
  if (cfg) c0 = *cfg;
  if (!(c0.theta >= 0.0)) return NBX_ERR_ARG;
  if (c0.mode != NBX_MODE_REFERENCE && c0.mode != NBX_MODE_PHYSICS) return NBX_ERR_ARG;
  if (c0.mode == NBX_MODE_PHYSICS) {
    // The group test needs theta < 1 (a cell holding a target is then never
    // accepted for that target's group) and a positive softening length.
    if (!(c0.theta > 0.0 && c0.theta < 1.0) || !(c0.eps > 0.0)) return NBX_ERR_ARG;
    if (c0.leaf_size < 1 || c0.leaf_size > kPhysMaxLeaf || c0.group_size < 1 || c0.group_size > kPhysBlock)
      return NBX_ERR_ARG;
    if (c0.precision != NBX_PREC_FP64 && c0.precision != NBX_PREC_MIXED) return NBX_ERR_ARG;
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return NBX_ERR_NODEVICE;
  }
  if (c0.device < 0 || c0.device >= ndev) return NBX_ERR_NODEVICE;
  if (hipSetDevice(c0.device) != hipSuccess) return NBX_ERR_HIP;
  nbx_ctx* c = new nbx_ctx;
  c->cfg = c0;
  c->physics = c0.mode == NBX_MODE_PHYSICS;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
      hipMalloc(&c->sc, sizeof(DevScalars)) != hipSuccess ||
      hipHostMalloc(reinterpret_cast<void**>(&c->hsc), sizeof(DevScalars), hipHostMallocDefault) != hipSuccess) {
    nbx_destroy(c);
    return NBX_ERR_HIP;
  }
  for (auto& e : c->ev)
    if (hipEventCreate(&e) != hipSuccess) { nbx_destroy(c); return NBX_ERR_HIP; }
  if (hipEventCreateWithFlags(&c->ev_scalars, hipEventDisableTiming) != hipSuccess) {
    nbx_destroy(c);
    return NBX_ERR_HIP;
  }
  std::memset(c->hsc, 0, sizeof(DevScalars));
  if (hipMemcpyAsync(c->sc, c->hsc, sizeof(DevScalars), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess || set_systems(c) != NBX_OK) {
    nbx_destroy(c);
    return NBX_ERR_HIP;
  }
  *out = c;
  return NBX_OK;
}

void nbx_destroy(nbx_ctx* c) {
  if (!c) return;
  (void)hipSetDevice(c->cfg.device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  drop_graph(c);
  each_buf(c, release);
  for (auto& e : c->ev) if (e) (void)hipEventDestroy(e);
  if (c->ev_scalars) (void)hipEventDestroy(c->ev_scalars);
  if (c->sc) (void)hipFree(c->sc);
  if (c->hsc) (void)hipHostFree(c->hsc);
  if (c->hstage) (void)hipHostFree(c->hstage);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* nbx_last_error(const nbx_ctx* c) { return c ? c->err.c_str() : "null context"; }

nbx_status nbx_trim(nbx_ctx* c) {
  NBXCHK(check_ctx(c));
  if (c->n != 0) return fail(c, NBX_ERR_STATE, "only a context without bodies can release its buffers");
  HIPCHK(c, hipStreamSynchronize(c->stream));
  each_buf(c, release);
  c->cap = 0;
  c->skey = c->skey2 = nullptr;
  c->spec_emit = false;
  c->total_nodes = 0;
  return set_systems(c);   // the (empty) per-system tables, and every valid flag cleared
}

uint64_t nbx_device_bytes(const nbx_ctx* c) {
  if (!c) return 0;
  uint64_t total = sizeof(DevScalars);
  each_buf(const_cast<nbx_ctx*>(c), [&](DBuf& b) { total += b.bytes; });
  return total;
}

bool valid_walk_list(int32_t list) {
  for (int v : kPhysLists)
    if (v == list) return true;
  return false;
}

nbx_status nbx_autotune(nbx_ctx* c, const int32_t* leaf, const int32_t* group, int32_t ncand, double dt,
                        int32_t steps, double* ms_out) {
  return nbx_autotune_lists(c, leaf, group, nullptr, ncand, dt, steps, ms_out);
}

nbx_status nbx_autotune_lists(nbx_ctx* c, const int32_t* leaf, const int32_t* group, const int32_t* list,
                              int32_t ncand, double dt, int32_t steps, double* ms_out) {
  NBXCHK(check_ctx(c));
  if (!c->physics) return fail(c, NBX_ERR_ARG, "the autotuner chooses physics-mode leaf and group sizes");
  if (ncand <= 0 || !leaf || !group || steps <= 0 || !(dt > 0.0)) return fail(c, NBX_ERR_ARG, "bad candidates");
  for (int32_t k = 0; k < ncand; ++k)
    if (leaf[k] < 1 || leaf[k] > kPhysMaxLeaf || group[k] < 1 || group[k] > kPhysBlock ||
        (list && !valid_walk_list(list[k])))
      return fail(c, NBX_ERR_ARG, "candidate out of range");
  if (c->n == 0) return NBX_OK;
  // The state every trial starts from (and that the call leaves behind).
  const int cur0 = c->cur;
  DBuf saved[kStateArrays];
  auto release_all = [&]() { for (auto& b : saved) release(b); };
  for (int k = 0; k < kStateArrays; ++k) {
    const size_t bytes = (size_t)c->n * elem_bytes(k);
    nbx_status st = ensure(c, saved[k], bytes);
    if (st == NBX_OK && hipMemcpyAsync(saved[k].p, c->st[cur0][k].p, bytes, hipMemcpyDeviceToDevice, c->stream) !=
                            hipSuccess)
      st = fail(c, NBX_ERR_HIP, "state copy failed");
    if (st != NBX_OK) { release_all(); return st; }
  }
  auto restore = [&]() -> nbx_status {
    for (int k = 0; k < kStateArrays; ++k)
      HIPCHK(c, hipMemcpyAsync(c->st[cur0][k].p, saved[k].p, (size_t)c->n * elem_bytes(k), hipMemcpyDeviceToDevice,
                               c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->cur = cur0;
    c->acc_valid = false;
    c->tree_valid = false;
    c->layout_valid = false;
    return NBX_OK;
  };
  const nbx_config keep = c->cfg;
  const int keep_list = c->walk_list;
  const nbx_timing tim0 = c->tim;                          // trials do not count as steps
  int32_t best = -1;
  double best_ms = 0.0;
  for (int32_t k = 0; k < ncand; ++k) {
    c->cfg.leaf_size = leaf[k];
    c->cfg.group_size = group[k];
    if (list) c->walk_list = list[k];
    ++c->gen;
    c->layout_valid = false;                               // groups of these sizes need a full build
    nbx_status st = nbx_run(c, dt, 1, nullptr);            // warm-up: buffers, first tree
    const auto t0 = std::chrono::steady_clock::now();
    if (st == NBX_OK) st = nbx_run(c, dt, steps, nullptr);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count() / steps;
    const nbx_status rs = restore();
    if (st != NBX_OK || rs != NBX_OK) {
      c->cfg = keep;
      c->walk_list = keep_list;
      c->tim = tim0;
      release_all();
      return st != NBX_OK ? st : rs;
    }
    if (ms_out) ms_out[k] = ms;
    if (best < 0 || ms < best_ms) { best = k; best_ms = ms; }
  }
  release_all();
  c->tim = tim0;
  c->cfg = keep;
  c->cfg.leaf_size = leaf[best];
  c->cfg.group_size = group[best];
  c->walk_list = list ? list[best] : keep_list;
  ++c->gen;
  return NBX_OK;
}

nbx_status nbx_set_alpha(nbx_ctx* c, double alpha) {
  if (!c) return NBX_ERR_ARG;
  if (!(alpha >= 0.0 && alpha < INFINITY)) return fail(c, NBX_ERR_ARG, "alpha must be finite and >= 0");
  if (alpha > 0.0 && !c->physics) return fail(c, NBX_ERR_ARG, "the relative opening criterion is a physics-mode option");
  c->alpha = alpha;
  ++c->gen;
  return NBX_OK;
}

nbx_status nbx_set_walk_list(nbx_ctx* c, int32_t entries) {
  if (!c) return NBX_ERR_ARG;
  if (!valid_walk_list(entries)) return fail(c, NBX_ERR_ARG, "the walk list holds 96, 128, 192 or 256 entries");
  if (!c->physics) return fail(c, NBX_ERR_ARG, "the walk list size is a physics-mode setting");
  if (entries != c->walk_list) c->acc_valid = false;   // forces of this setting from the next walk on
  c->walk_list = entries;
  ++c->gen;
  return NBX_OK;
}

nbx_status nbx_set_rebuild_interval(nbx_ctx* c, int32_t every) {
  if (!c) return NBX_ERR_ARG;
  if (every < 1 || every > 1000) return fail(c, NBX_ERR_ARG, "the rebuild interval must be 1..1000 steps");
  if (every > 1 && !c->physics) return fail(c, NBX_ERR_ARG, "tree refits are a physics-mode option");
  c->rebuild_every = every;
  c->since_build = 0;
  ++c->gen;
  return NBX_OK;
}

nbx_status nbx_set_theta(nbx_ctx* c, double theta) {
  if (!c || !(theta >= 0.0)) return NBX_ERR_ARG;
  if (c->physics && !(theta > 0.0 && theta < 1.0)) return fail(c, NBX_ERR_ARG, "physics mode needs 0 < theta < 1");
  c->tree_valid = false;
  c->layout_valid = false;
  c->cfg.theta = theta;
  c->acc_valid = false;
  ++c->gen;
  return NBX_OK;
}

int64_t nbx_size(const nbx_ctx* c) { return c ? c->n : -1; }
int32_t nbx_num_systems(const nbx_ctx* c) { return c ? (int32_t)c->sys_n.size() : -1; }
int64_t nbx_system_size(const nbx_ctx* c, int32_t s) {
  return (c && s >= 0 && s < (int32_t)c->sys_n.size()) ? c->sys_n[(size_t)s] : -1;
}

nbx_status nbx_upload_batch(nbx_ctx* c, int32_t nsys, const int64_t* counts, const double* x, const double* y,
                            const double* z, const double* vx, const double* vy, const double* vz, const double* m,
                            const int32_t* type) {
  NBXCHK(check_ctx(c));
  if (nsys < 0 || (nsys > 0 && !counts)) return fail(c, NBX_ERR_ARG, "bad system count");
  int64_t n = 0;
  for (int s = 0; s < nsys; ++s) {
    if (counts[s] <= 0) return fail(c, NBX_ERR_ARG, "every system needs at least one body");
    n += counts[s];
  }
  if (n > (int64_t)INT32_MAX / 64) return fail(c, NBX_ERR_ARG, "particle count out of range");
  if (n > 0 && !(x && y && z && vx && vy && vz && m && type)) return fail(c, NBX_ERR_ARG, "null particle array");
  c->cur = 0;
  NBXCHK(ensure_particles(c, n, 0));
  c->n = n;
  c->sys_n.assign(counts, counts + nsys);
  int64_t at = 0;
  for (int s = 0; s < nsys; ++s) {
    NBXCHK(put_bodies(c, at, counts[s], s, at, x + at, y + at, z + at, vx + at, vy + at, vz + at, m + at, type + at));
    at += counts[s];
  }
  return set_systems(c);
}

nbx_status nbx_upload(nbx_ctx* c, int64_t n, const double* x, const double* y, const double* z, const double* vx,
                      const double* vy, const double* vz, const double* m, const int32_t* type) {
  if (!c) return NBX_ERR_ARG;
  if (n < 0) return fail(c, NBX_ERR_ARG, "particle count out of range");
  if (n == 0) return nbx_upload_batch(c, 0, nullptr, x, y, z, vx, vy, vz, m, type);
  return nbx_upload_batch(c, 1, &n, x, y, z, vx, vy, vz, m, type);
}

nbx_status nbx_add_system(nbx_ctx* c, int64_t n, const double* x, const double* y, const double* z, const double* vx,
                          const double* vy, const double* vz, const double* m, const int32_t* type, int32_t* sys_out) {
  NBXCHK(check_ctx(c));
  if (n <= 0 || c->n + n > (int64_t)INT32_MAX / 64) return fail(c, NBX_ERR_ARG, "bad particle count");
  if (!(x && y && z && vx && vy && vz && m && type)) return fail(c, NBX_ERR_ARG, "null particle array");
  NBXCHK(ensure_particles(c, c->n + n, c->n));
  const int32_t s = (int32_t)c->sys_n.size();
  NBXCHK(put_bodies(c, c->n, n, s, c->n, x, y, z, vx, vy, vz, m, type));
  c->n += n;
  c->sys_n.push_back(n);
  if (sys_out) *sys_out = s;
  return set_systems(c);
}

nbx_status nbx_download_system(nbx_ctx* c, int32_t s, double* x, double* y, double* z, double* vx, double* vy,
                               double* vz) {
  NBXCHK(check_ctx(c));
  if (s < 0 || s >= (int32_t)c->sys_n.size()) return fail(c, NBX_ERR_ARG, "no such system");
  const int64_t off = sys_offset(c, s), cnt = c->sys_n[(size_t)s];
  std::vector<int32_t> id;
  NBXCHK(read_ids(c, off, cnt, off, id));
  double* outs[6] = {x, y, z, vx, vy, vz};
  for (int k = 0; k < 6; ++k)
    if (outs[k]) NBXCHK(download_range(c, c->st[c->cur][k].p, 8, off, cnt, off, id, outs[k]));
  return NBX_OK;
}

nbx_status nbx_remove_system(nbx_ctx* c, int32_t s, double* x, double* y, double* z, double* vx, double* vy,
                             double* vz) {
  NBXCHK(nbx_download_system(c, s, x, y, z, vx, vy, vz));
  const int64_t off = sys_offset(c, s), cnt = c->sys_n[(size_t)s];
  const int64_t n_new = c->n - cnt;
  hipStream_t st = c->stream;
  uint32_t* perm = P<uint32_t>(c->idx[0]);
  launch_remove_map(n_new, off, cnt, perm, st);
  launch_gather_state(n_new, perm, view(c, c->cur), view(c, 1 - c->cur), st);
  c->cur = 1 - c->cur;
  // the other systems keep their cached forces
  for (int k = 0; k < 3; ++k) {
    launch_gather_f64(n_new, perm, P<double>(c->acc[k]), P<double>(c->key[1]), st);
    if (n_new > 0)
      HIPCHK(c, hipMemcpyAsync(c->acc[k].p, c->key[1].p, (size_t)n_new * 8, hipMemcpyDeviceToDevice, st));
  }
  launch_relabel(n_new, P<int32_t>(c->st[c->cur][9]), s, st);
  // upload indices of later systems move down by cnt
  if (n_new > 0) {
    const void* staged = nullptr;
    NBXCHK(fetch(c, c->st[c->cur][8].p, (size_t)n_new * 4, &staged));
    int32_t* ids = static_cast<int32_t*>(const_cast<void*>(staged));
    for (int64_t k = 0; k < n_new; ++k)
      if (ids[k] >= off + cnt) ids[k] -= (int32_t)cnt;
    HIPCHK(c, hipMemcpyAsync(c->st[c->cur][8].p, ids, (size_t)n_new * 4, hipMemcpyHostToDevice, st));
  }
  HIPCHK(c, hipStreamSynchronize(st));
  const bool kept = c->acc_valid;
  c->n = n_new;
  c->sys_n.erase(c->sys_n.begin() + s);
  NBXCHK(set_systems(c));
  c->acc_valid = kept && n_new > 0;   // the tree is rebuilt on the next step; the others' forces are unchanged
  return NBX_OK;
}

nbx_status nbx_download(nbx_ctx* c, double* x, double* y, double* z, double* vx, double* vy, double* vz) {
  NBXCHK(check_ctx(c));
  const int64_t n = c->n;
  if (n == 0) return NBX_OK;
  std::vector<int32_t> id;
  NBXCHK(read_ids(c, 0, n, 0, id));
  double* outs[6] = {x, y, z, vx, vy, vz};
  for (int k = 0; k < 6; ++k)
    if (outs[k]) NBXCHK(download_range(c, c->st[c->cur][k].p, 8, 0, n, 0, id, outs[k]));
  return NBX_OK;
}

nbx_status nbx_adaptive_dt(nbx_ctx* c, double base_dt, double* dt_out) {
  NBXCHK(check_ctx(c));
  if (!dt_out) return fail(c, NBX_ERR_ARG, "null dt_out");
  if (c->n == 0) { *dt_out = base_dt; return NBX_OK; }   // computeAdaptiveDt of no particles
  SysArrays sy = sysa(c);
  launch_fill_v2max(sy, c->stream);
  launch_maxv2(c->n, view(c, c->cur), sy, c->stream);
  launch_dt(sy, base_dt, nullptr, 0, c->sc, c->stream);
  HIPCHK(c, hipMemcpyAsync(dt_out, c->s_dt.p, 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBX_OK;
}

nbx_status nbx_step(nbx_ctx* c, double dt) {
  NBXCHK(check_ctx(c));
  if (c->n == 0) return NBX_OK;
  launch_set_dt(sysa(c), dt, c->stream);
  NBXCHK(step_device(c));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBX_OK;
}

nbx_status nbx_run(nbx_ctx* c, double base_dt, int32_t nsteps, double* dts_out) {
  NBXCHK(check_ctx(c));
  if (nsteps < 0) return fail(c, NBX_ERR_ARG, "negative step count");
  if (nsteps == 0) return NBX_OK;
  if (c->n == 0 || !c->cfg.adaptive_dt) {   // no bodies: computeAdaptiveDt gives base_dt; fixed dt: base_dt
    if (dts_out) for (int s = 0; s < nsteps; ++s) dts_out[s] = base_dt;
    if (c->n == 0) return NBX_OK;
  }
  NBXCHK(ensure(c, c->dts, (size_t)nsteps * 8));
  SysArrays sy = sysa(c);
  // dt of the first step from the current speeds (adaptive), or the fixed dt
  auto first_dt = [&]() {
    if (c->cfg.adaptive_dt) {
      launch_fill_v2max(sy, c->stream);
      launch_maxv2(c->n, view(c, c->cur), sy, c->stream);
    } else {
      launch_set_dt(sy, base_dt, c->stream);
    }
  };
  first_dt();
  int32_t s = 0;
  const int period = graph_period(c);
  if (c->graph && nsteps >= period) {
    NBXCHK(ensure_accel(c));
    // step by step up to a step that rebuilds the tree (physics refits)
    for (; s < nsteps && c->physics && c->layout_valid && c->since_build + 1 < c->rebuild_every; ++s) {
      if (c->cfg.adaptive_dt) launch_dt(sy, base_dt, P<double>(c->dts), s, c->sc, c->stream);
      NBXCHK(step_device(c));
    }
    const int32_t reps = (nsteps - s) / period;
    if (reps > 0) {
      NBXCHK(prepare_bounded(c, nsteps));
      Snapshot sn{};
      NBXCHK(snapshot(c, &sn));
      uint64_t grids = 1469598103934665603ull ^ c->b_grid_pair;   // FNV-1a over the grids
      for (unsigned g : c->b_grid_up) grids = (grids ^ g) * 1099511628211ull;
      c->bounded_patch = c->graph_patch && c->patch_ok && c->patch_streak >= 2 && c->sys_n.size() == 1;
      const nbx_ctx::GraphKey key{c->gen,        base_dt,         c->cfg.adaptive_dt, period, c->b_cap_records,
                                  c->b_cap_list, c->groups_bound, c->b_dmax,          grids, c->bounded_patch};
      if (!c->gexec || !(key == c->gkey)) {
        NBXCHK(capture(c, base_dt, period));
        c->gkey = key;
      }
      launch_set_run(c->sc, s, c->stream);
      for (int32_t r = 0; r < reps; ++r) HIPCHK(c, hipGraphLaunch(c->gexec, c->stream));
      NBXCHK(sync_scalars(c));                            // the run's one wait: did every build fit?
      if (c->hsc->abort_flags) {
        ++c->tim.graph_aborts;
        if (c->hsc->abort_flags & kAbortPatch) c->graph_patch = false;
        NBXCHK(restore(c, sn));
        launch_set_run(c->sc, 0, c->stream);              // the flag would stop the walks of the steps below
        first_dt();                                       // the speeds of the restored state
      } else {
        const int32_t done = reps * period;
        s += done;
        c->tim.steps += done;
        c->tim.graph_steps += done;
        c->tim.graph_launches += reps;
        c->tim.graph_nodes = c->gnodes;
        c->tim.graph_period = period;
        c->total_nodes = (uint32_t)c->hsc->total_nodes;
        c->max_depth = c->hsc->max_leaf_depth;
        c->tim.nodes = c->total_nodes;
        c->tim.depth = c->max_depth;
        c->acc_valid = true;                              // the last walk's forces, at the current positions
      }
    }
  }
  for (; s < nsteps; ++s) {
    if (c->cfg.adaptive_dt) launch_dt(sy, base_dt, P<double>(c->dts), s, c->sc, c->stream);
    NBXCHK(step_device(c));
  }
  if (dts_out && c->cfg.adaptive_dt)
    HIPCHK(c, hipMemcpyAsync(dts_out, c->dts.p, (size_t)nsteps * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  ++c->tim.host_waits;
  return NBX_OK;
}

nbx_status nbx_set_graph(nbx_ctx* c, int on) {
  if (!c) return NBX_ERR_ARG;
  if (on < 0 || on > 2) return fail(c, NBX_ERR_ARG, "graph mode must be 0, 1 or 2");
  if (on != c->graph) {
    c->graph = on;
    drop_graph(c);
    c->b_cap_records = c->b_cap_list = 0;
    c->b_dmax = -1;
  }
  return NBX_OK;
}

nbx_status nbx_accel(nbx_ctx* c, double* ax, double* ay, double* az) {
  NBXCHK(check_ctx(c));
  if (c->n == 0) return NBX_OK;
  if (!(ax && ay && az)) return fail(c, NBX_ERR_ARG, "null output");
  const bool fresh = !c->acc_valid;
  NBXCHK(ensure_accel(c));
  if (fresh) NBXCHK(finish_timing(c, false));
  std::vector<int32_t> id;
  NBXCHK(read_ids(c, 0, c->n, 0, id));
  double* outs[3] = {ax, ay, az};
  for (int k = 0; k < 3; ++k) NBXCHK(download_range(c, c->acc[k].p, 8, 0, c->n, 0, id, outs[k]));
  return NBX_OK;
}

static nbx_status download_unpermuted(nbx_ctx* c, const void* dev, double* out) {
  std::vector<int32_t> id;
  NBXCHK(read_ids(c, 0, c->n, 0, id));
  return download_range(c, dev, 8, 0, c->n, 0, id, out);
}

// Tree potential of the current positions (physics mode).  Its walk writes
// forces into the diagnostic buffer: the cached accelerations the next
// step's opening kick uses stay untouched (with the relative criterion they
// came from a walk that depended on the previous |a|), so a diagnostic call
// between steps leaves the trajectory bit for bit as it was.
static nbx_status tree_potential(nbx_ctx* c) {
  const size_t n = (size_t)std::max<int64_t>(c->n, 1);
  NBXCHK(ensure(c, c->pot, n * 8));
  NBXCHK(ensure(c, c->dacc, n * 3 * 8));
  if (!c->tree_valid) {
    NBXCHK(build_tree(c));
    c->acc_valid = false;   // the build reordered the bodies
  }
  double* da = P<double>(c->dacc);
  launch_phys_walk(P<Group>(c->tgroups), group_grid(c), P<PhysNode>(c->pnodes), view(c, c->cur), sysa(c), c->cfg.eps,
                   c->cfg.leaf_size, da, da + n, da + 2 * n, P<double>(c->pot), c->sc, false, false,
                   c->cfg.precision == NBX_PREC_FP64, nullptr, 0.0, c->cfg.theta, c->walk_list, c->stream);
  HIPCHK(c, hipGetLastError());
  return NBX_OK;
}

nbx_status nbx_potential(nbx_ctx* c, double* pot) {
  NBXCHK(check_ctx(c));
  if (!c->physics) return fail(c, NBX_ERR_ARG, "the tree potential exists in physics mode only");
  if (c->n == 0) return NBX_OK;
  if (!pot) return fail(c, NBX_ERR_ARG, "null output");
  NBXCHK(tree_potential(c));
  return download_unpermuted(c, c->pot.p, pot);
}

nbx_status nbx_direct(nbx_ctx* c, double* ax, double* ay, double* az, double* pot) {
  NBXCHK(check_ctx(c));
  if (c->n == 0) return NBX_OK;
  if (!(c->cfg.eps >= 0.0)) return fail(c, NBX_ERR_ARG, "negative softening");
  DBuf tmp;
  NBXCHK(ensure(c, tmp, (size_t)c->n * 8 * 4));
  double* d = P<double>(tmp);
  launch_direct64(c->n, view(c, c->cur), sysa(c), c->cfg.eps, d, d + c->n, d + 2 * c->n, d + 3 * c->n, c->stream);
  nbx_status st = NBX_OK;
  double* outs[4] = {ax, ay, az, pot};
  for (int k = 0; k < 4 && st == NBX_OK; ++k)
    if (outs[k]) st = download_unpermuted(c, d + k * c->n, outs[k]);
  (void)hipStreamSynchronize(c->stream);
  release(tmp);
  return st;
}

nbx_status nbx_energy_systems(nbx_ctx* c, int source, double* out) {
  NBXCHK(check_ctx(c));
  if (!out) return fail(c, NBX_ERR_ARG, "null output");
  if (source != 0 && source != 1) return fail(c, NBX_ERR_ARG, "source must be 0 (tree) or 1 (direct)");
  if (source == 0 && !c->physics) return fail(c, NBX_ERR_ARG, "the tree potential exists in physics mode only");
  const int S = (int)c->sys_n.size();
  if (c->n == 0) return NBX_OK;
  NBXCHK(ensure(c, c->pot, (size_t)c->n * 8));
  if (source == 0) {
    NBXCHK(tree_potential(c));
  } else {
    NBXCHK(ensure(c, c->dacc, (size_t)c->n * 3 * 8));
    double* da = P<double>(c->dacc);
    launch_direct64(c->n, view(c, c->cur), sysa(c), c->cfg.eps, da, da + c->n, da + 2 * c->n, P<double>(c->pot),
                    c->stream);
  }
  launch_energy(view(c, c->cur), P<double>(c->pot), P<BoxBlock>(c->boxes), c->nboxes, P<int32_t>(c->box_off), S,
                P<double>(c->epart), P<double>(c->eout), c->stream);
  HIPCHK(c, hipMemcpyAsync(out, c->eout.p, (size_t)S * 5 * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  return NBX_OK;
}

nbx_status nbx_energy(nbx_ctx* c, int source, double* out5) {
  NBXCHK(check_ctx(c));
  if (!out5) return fail(c, NBX_ERR_ARG, "null output");
  if (source != 0 && source != 1) return fail(c, NBX_ERR_ARG, "source must be 0 (tree) or 1 (direct)");
  for (int k = 0; k < 5; ++k) out5[k] = 0.0;
  const int S = (int)c->sys_n.size();
  if (c->n == 0) return NBX_OK;
  std::vector<double> per((size_t)S * 5);
  NBXCHK(nbx_energy_systems(c, source, per.data()));
  for (int s = 0; s < S; ++s)
    for (int k = 0; k < 5; ++k) out5[k] += per[(size_t)s * 5 + k];
  return NBX_OK;
}

nbx_status nbx_get_timing(nbx_ctx* c, nbx_timing* out) {
  if (!c || !out) return NBX_ERR_ARG;
  *out = c->tim;
  return NBX_OK;
}

nbx_status nbx_reset_timing(nbx_ctx* c) {
  if (!c) return NBX_ERR_ARG;
  c->tim = nbx_timing{};
  return NBX_OK;
}

nbx_status nbx_set_profiling(nbx_ctx* c, int flags) {
  if (!c) return NBX_ERR_ARG;
  c->prof = flags;
  return NBX_OK;
}

nbx_status nbx_debug_keys(nbx_ctx* c, double* root4, uint64_t* hi, uint64_t* lo) {
  NBXCHK(check_ctx(c));
  const int64_t n = c->n;
  if (!root4) return fail(c, NBX_ERR_ARG, "null root4");
  if (n == 0) return NBX_OK;
  if (c->sys_n.size() != 1) return fail(c, NBX_ERR_ARG, "parity hooks need a single system");
  if (!(hi && lo)) return fail(c, NBX_ERR_ARG, "null key output");
  SysArrays sy = sysa(c);
  launch_bbox(view(c, c->cur), P<BoxBlock>(c->boxes), c->nboxes, P<int32_t>(c->box_off), P<double>(c->box_part), sy,
              nullptr, c->stream);
  for (int w = 0; w < 2; ++w) NBXCHK(ensure(c, c->key2[w], (size_t)c->cap * 8));
  launch_keys_export(n, view(c, c->cur), sy, P<uint64_t>(c->key2[0]), P<uint64_t>(c->key2[1]), c->stream);
  HIPCHK(c, hipMemcpyAsync(root4, c->s_root.p, 32, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(hi, c->key2[0].p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipMemcpyAsync(lo, c->key2[1].p, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  c->skey2 = nullptr;  // the second-word buffers were reused
  c->tree_valid = false;
  c->layout_valid = false;
  return NBX_OK;
}

nbx_status nbx_debug_order(nbx_ctx* c, int32_t* id, int32_t* sys) {
  NBXCHK(check_ctx(c));
  if (c->n == 0) return NBX_OK;
  if (!(id && sys)) return fail(c, NBX_ERR_ARG, "null output");
  const void* staged = nullptr;
  NBXCHK(fetch(c, c->st[c->cur][8].p, (size_t)c->n * 4, &staged));
  std::memcpy(id, staged, (size_t)c->n * 4);
  NBXCHK(fetch(c, c->st[c->cur][9].p, (size_t)c->n * 4, &staged));
  std::memcpy(sys, staged, (size_t)c->n * 4);
  return NBX_OK;
}

nbx_status nbx_debug_tree_size(nbx_ctx* c, int64_t* nodes) {
  NBXCHK(check_ctx(c));
  if (!nodes) return fail(c, NBX_ERR_ARG, "null output");
  if (c->n == 0) { *nodes = 0; return NBX_OK; }
  if (c->sys_n.size() != 1) return fail(c, NBX_ERR_ARG, "parity hooks need a single system");
  NBXCHK(built_tree(c));
  *nodes = c->total_nodes;
  return NBX_OK;
}

nbx_status nbx_debug_groups(nbx_ctx* c, int64_t cap, int32_t* first, int32_t* count, int32_t* sys,
                            int64_t* ngroups) {
  NBXCHK(check_ctx(c));
  if (!ngroups || cap < 0 || (cap > 0 && !(first && count && sys))) return fail(c, NBX_ERR_ARG, "bad output");
  *ngroups = 0;
  if (c->n == 0) return NBX_OK;
  NBXCHK(built_tree(c));
  if (c->physics && c->ntgroups < 0) {
    NBXCHK(sync_scalars(c));
    c->ntgroups = c->hsc->tree_groups;
  }
  const int64_t ng = c->physics ? c->ntgroups : c->ngroups;
  const Group* dg = P<Group>(c->physics ? c->tgroups : c->groups);
  *ngroups = ng;
  const int64_t k = std::min(ng, cap);
  if (k == 0) return NBX_OK;
  std::vector<Group> h((size_t)k);
  HIPCHK(c, hipMemcpyAsync(h.data(), dg, (size_t)k * sizeof(Group), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(c, hipStreamSynchronize(c->stream));
  for (int64_t j = 0; j < k; ++j) {
    first[j] = h[(size_t)j].first;
    count[j] = h[(size_t)j].count;
    sys[j] = h[(size_t)j].sys;
  }
  return NBX_OK;
}

nbx_status nbx_debug_tree(nbx_ctx* c, int32_t* depth, int32_t* leaf, int32_t* body, int32_t* cmask, double* geo4,
                          double* mass, double* com3, double* q6, int32_t* perm) {
  int64_t total = 0;
  NBXCHK(nbx_debug_tree_size(c, &total));
  if (total == 0) return NBX_OK;
  if (!(depth && leaf && body && cmask && geo4 && mass && com3 && q6 && perm))
    return fail(c, NBX_ERR_ARG, "null output");
  if (c->physics) {
    // The physics build writes walk records only: lay out the reference
    // records of the same tree, with NEXT's moments.
    NBXCHK(ref_layout(c, false));
    launch_upward_all(c->hsc->level_count, P<int32_t>(c->level_list), P<uint32_t>(c->end_list), P<RefNode>(c->nodes),
                      P<WalkHead>(c->heads), c->cfg.theta, view(c, c->cur).sys, sysa(c).root, c->stream);
  }
  // one staging allocation for all outputs
  const size_t ni = (size_t)total * 4, nd = (size_t)total * 8;
  DBuf tmp;
  NBXCHK(ensure(c, tmp, 4 * ni + 14 * nd));
  char* b = P<char>(tmp);
  int32_t* d_depth = reinterpret_cast<int32_t*>(b);
  int32_t* d_leaf = d_depth + total;
  int32_t* d_body = d_leaf + total;
  int32_t* d_cmask = d_body + total;
  double* d_geo = reinterpret_cast<double*>(b + 4 * ni);
  double* d_mass = d_geo + 4 * total;
  double* d_com = d_mass + total;
  double* d_q = d_com + 3 * total;
  launch_export(total, P<RefNode>(c->nodes), c->skey, c->skey2, P<int32_t>(c->st[c->cur][8]), sysa(c), d_depth,
                d_leaf, d_body, d_cmask, d_geo, d_mass, d_com, d_q, c->stream);
  hipError_t e = hipMemcpyAsync(depth, d_depth, ni, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(leaf, d_leaf, ni, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(body, d_body, ni, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(cmask, d_cmask, ni, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(geo4, d_geo, 4 * nd, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(mass, d_mass, nd, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(com3, d_com, 3 * nd, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(q6, d_q, 6 * nd, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(perm, c->st[c->cur][8].p, (size_t)c->n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  release(tmp);
  if (e != hipSuccess) return fail(c, NBX_ERR_HIP, std::string("tree export: ") + hipGetErrorString(e));
  return NBX_OK;
}

}  // extern "C"
