/*
 * nbx.h - C ABI of the MI355X Barnes-Hut gravity + KDK time-step engine.
 *
 * This is the drop-in boundary for NEXT's gravity and time-step path.  NEXT
 * has no FFI; its path is two header-only C++ calls made by the driver loop
 * (begrun.cpp:96-99):
 *
 *     real computeAdaptiveDt(const Particle&, real base_dt);   adaptive.h:23
 *     void Step(ParticleSystem&, real dt);                     step.h:25
 *
 * on a caller-owned structure of arrays (particle.h:22-52).  Here the state
 * lives on the GPU between calls: nbx_upload copies the caller's arrays in,
 * nbx_adaptive_dt / nbx_step / nbx_run replace the two calls, and
 * nbx_download copies positions and velocities back (for snapshots).
 *
 * Conventions
 *  - Every function returns nbx_status (0 = OK, negative = error) unless
 *    stated.  No C++ exception crosses this boundary.  The text of the last
 *    error is kept per context (nbx_last_error).
 *  - Arrays are plain host pointers owned by the caller; they are copied.
 *    Particle arrays are in the caller's order (the order of the IC file);
 *    the engine reorders bodies internally along its tree and always
 *    returns results in the caller's order.
 *  - A context owns its device buffers and one HIP stream.  A context is
 *    used from one host thread at a time; several contexts may coexist.
 */
#ifndef NBX_H
#define NBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum nbx_status {
  NBX_OK = 0,
  NBX_ERR_ARG = -1,        /* bad argument: null pointer, n < 0, theta < 0 ... */
  NBX_ERR_NONFINITE = -2,  /* a position is NaN or infinite */
  NBX_ERR_DEPTH = -3,      /* two bodies of one system share all 42 octree digits: the tree
                              would need leaves deeper than 42 levels (cells below 2^-42 of the
                              root's half-width).  NEXT's insertion (octree.h:69-87) keeps
                              splitting for distinct bodies and recurses without end for
                              coincident ones; this engine stops at 42 levels either way. */
  NBX_ERR_COINCIDENT = NBX_ERR_DEPTH, /* former name of NBX_ERR_DEPTH */
  NBX_ERR_NOMEM = -4,      /* device allocation failed or budget exceeded */
  NBX_ERR_HIP = -5,        /* HIP runtime error */
  NBX_ERR_NODEVICE = -6,   /* no usable GPU */
  NBX_ERR_STATE = -7       /* call out of order (e.g. no particles uploaded) */
} nbx_status;

typedef enum nbx_mode {
  /* NEXT's algorithm in FP64: one body per leaf, one-level softened
   * quadrupole, per-interaction adaptive softening, per-target MAC
   * s/d < theta (octree.h:25-179, softening.h:21-35). */
  NBX_MODE_REFERENCE = 0,
  /* Bucket leaves, traceless quadrupole with parallel-axis propagation,
   * Plummer softening eps, group MAC; FP32 or FP64 walk. */
  NBX_MODE_PHYSICS = 1
} nbx_mode;

typedef enum nbx_precision {
  NBX_PREC_FP64 = 0,   /* FP64 walk */
  NBX_PREC_MIXED = 1   /* FP32 walk (each list flush summed in FP32, added to FP64
                          totals); FP64 multipoles and state */
} nbx_precision;

typedef struct nbx_config {
  int device;        /* HIP device ordinal */
  int mode;          /* nbx_mode */
  int precision;     /* nbx_precision (NBX_MODE_PHYSICS only) */
  double theta;      /* opening angle; NEXT fixes 0.5 (step.h:32) */
  double eps;        /* Plummer softening length (NBX_MODE_PHYSICS only) */
  int leaf_size;     /* max bodies per leaf bucket, 1..64 (NBX_MODE_PHYSICS only) */
  int group_size;    /* at most this many targets per walk group, 1..64 (NBX_MODE_PHYSICS
                        only; reference mode walks runs of 64) */
  int adaptive_dt;   /* nbx_run: 1 = NEXT's speed-adaptive dt (adaptive.h), 0 = fixed dt */
} nbx_config;

typedef struct nbx_ctx nbx_ctx;

/* Per-phase device times (ms), summed over the steps / tree builds since
 * the last nbx_reset_timing, measured with HIP events on the context's
 * stream (profiling bit 1), and walk interaction counters (bit 2). */
typedef struct nbx_timing {
  int64_t steps;       /* steps (or accel evaluations) accumulated */
  double keys_ms;      /* bounding box + Morton keys */
  double sort_ms;      /* sort (radix, or nearly-sorted path) + reorder of the state */
  double build_ms;     /* node layout from sorted keys */
  double upward_ms;    /* multipole pass */
  double walk_ms;      /* tree walk (+ fused closing half kick) */
  double kick_ms;      /* opening half kick + drift */
  double total_ms;     /* whole step */
  int64_t nodes;       /* tree nodes of the last build */
  int32_t depth;       /* deepest leaf of the last build */
  int64_t n_pp;        /* particle-particle interactions (if counting) */
  int64_t n_pc;        /* particle-cell interactions (if counting) */
  int64_t n_open;      /* opened cells, per target (reference) or per group (physics) (if counting) */
  int64_t n_iter;      /* walk traversal steps summed over waves (if counting) */
  int64_t n_eval;      /* interaction-evaluation steps summed over waves (if counting) */
  int64_t graph_steps;    /* steps replayed from a captured graph (nbx_set_graph) */
  int64_t graph_launches; /* graph launches */
  int64_t graph_nodes;    /* kernels in the captured period of steps */
  int64_t graph_period;   /* steps in the captured period */
  int64_t graph_aborts;   /* captured runs redone step by step (a build did not fit) */
  int64_t host_waits;     /* host waits on the device inside steps and runs */
  int64_t sort_patched;   /* tree builds sorted by the nearly-sorted path (bodies in the last build's order) */
  int64_t sort_redos;     /* of those, builds redone with the radix sort (the bodies had moved too much) */
} nbx_timing;

/* Fill *cfg with NEXT's settings: device 0, reference mode, FP64, theta 0.5,
 * adaptive dt; physics defaults eps 0.01, leaf_size 16, group_size 64. */
void nbx_config_default(nbx_config* cfg);

nbx_status nbx_create(nbx_ctx** out, const nbx_config* cfg);
void nbx_destroy(nbx_ctx* ctx);
const char* nbx_last_error(const nbx_ctx* ctx);
nbx_status nbx_set_theta(nbx_ctx* ctx, double theta);

/* Device memory (no counterpart in NEXT).  Buffers grow with the bodies a
 * context holds and are kept for reuse; nbx_trim frees them once the context
 * holds no bodies (NBX_ERR_STATE otherwise), and nbx_device_bytes reports
 * what the context holds now.  The request server trims a batch that
 * empties, so idle batches hold no memory. */
nbx_status nbx_trim(nbx_ctx* ctx);
uint64_t nbx_device_bytes(const nbx_ctx* ctx);

/* Physics mode: opening criterion relative to each group's previous
 * acceleration (no counterpart in NEXT, whose criterion is geometric).  A
 * cell that passes the geometric test (distance to the group's box beyond
 * bmax / theta) is accepted only if m bmax^3 <= alpha |a|min d^5, where
 * |a|min is the smallest |a| of the group at the start of the step and d the
 * distance.  Applies to the walks inside nbx_step / nbx_run after the first
 * evaluation (which has no previous acceleration); alpha = 0 (the default)
 * turns it off.  NBX_ERR_ARG for alpha < 0, non-finite, or reference mode. */
nbx_status nbx_set_alpha(nbx_ctx* ctx, double alpha);

/* Physics mode: rebuild the tree (keys, sort, layout, walk groups) only on
 * every k-th step and refit it in between: the bodies keep their order and
 * the tree its topology, while leaf records, centres of mass, quadrupoles
 * and bmax are recomputed from the current positions, so the opening test
 * stays conservative (no counterpart in NEXT, which rebuilds every step).
 * k = 1 (the default) rebuilds every step.  NBX_ERR_ARG for k outside
 * 1..1000, or k > 1 in reference mode. */
nbx_status nbx_set_rebuild_interval(nbx_ctx* ctx, int32_t k);

/* Copy n particles in (Particle SoA of particle.h:22-52).  n may be 0. */
nbx_status nbx_upload(nbx_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                      const double* vx, const double* vy, const double* vz, const double* m,
                      const int32_t* type);

/* Copy positions and velocities out, in upload order.  Any pointer may be
 * null to skip that array. */
nbx_status nbx_download(nbx_ctx* ctx, double* x, double* y, double* z, double* vx, double* vy,
                        double* vz);

int64_t nbx_size(const nbx_ctx* ctx);

/* ---- several independent systems in one context (ensembles, serving) ----
 * No counterpart in NEXT, which runs one system per process.  Systems never
 * interact; every system's trajectory is the same as in a context of its
 * own.  nbx_step / nbx_run / nbx_accel / nbx_download act on all of them. */

/* Copy nsys systems in: counts[s] bodies each (> 0), arrays concatenated in
 * system order.  Replaces any previous contents. */
nbx_status nbx_upload_batch(nbx_ctx* ctx, int32_t nsys, const int64_t* counts, const double* x, const double* y,
                            const double* z, const double* vx, const double* vy, const double* vz,
                            const double* m, const int32_t* type);
/* Append one system of n bodies; *sys_out (may be null) receives its index. */
nbx_status nbx_add_system(nbx_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                          const double* vx, const double* vy, const double* vz, const double* m,
                          const int32_t* type, int32_t* sys_out);
/* Copy system s out (its own upload order); nbx_remove_system also removes
 * it, and the systems after it move down by one index. */
nbx_status nbx_download_system(nbx_ctx* ctx, int32_t s, double* x, double* y, double* z, double* vx, double* vy,
                               double* vz);
nbx_status nbx_remove_system(nbx_ctx* ctx, int32_t s, double* x, double* y, double* z, double* vx, double* vy,
                             double* vz);
int32_t nbx_num_systems(const nbx_ctx* ctx);
int64_t nbx_system_size(const nbx_ctx* ctx, int32_t s);

/* = computeAdaptiveDt (adaptive.h:23-52) on the device state. */
nbx_status nbx_adaptive_dt(nbx_ctx* ctx, double base_dt, double* dt_out);

/* = Step (step.h:25-204): one kick-drift-kick step of size dt (every system). */
nbx_status nbx_step(nbx_ctx* ctx, double dt);

/* nsteps iterations of the driver loop body (begrun.cpp:97-99): dt from
 * the current speeds (each system its own), then one step, without host
 * round trips; with cfg.adaptive_dt = 0 every step is base_dt.  dts_out (may
 * be null) receives the nsteps time steps of system 0. */
nbx_status nbx_run(nbx_ctx* ctx, double base_dt, int32_t nsteps, double* dts_out);

/* Accelerations at the current positions (bhAccel over all targets,
 * octree.h:133-179), upload order. */
nbx_status nbx_accel(nbx_ctx* ctx, double* ax, double* ay, double* az);

/* No counterpart in NEXT.  For callers that cannot allocate device memory
 * (serving loops and harnesses that forbid hipMalloc): nbx_accel's
 * reference-mode accelerations (same kernels, same results) of n bodies the
 * caller holds in device memory, written to device arrays ax/ay/az, without
 * a context.  Every other buffer is a static device array of the library
 * (16 MiB, held by each GPU it runs on until the process ends), sized for at
 * most NBX_NOALLOC_MAX_BODIES bodies and NBX_NOALLOC_MAX_NODES tree nodes.
 * Runs on the calling thread's current device and the given stream (null:
 * the null stream); synchronises that stream once for the node counts and
 * returns after the results are written.  Calls on one device share the
 * array and are serialised by a lock.  The call itself allocates no memory
 * and is not capturable (it waits on the stream).  Errors: NBX_ERR_ARG (null
 * pointer, n < 0 or above the maximum, theta < 0), NBX_ERR_NOMEM (the tree
 * needs more nodes than the maximum), NBX_ERR_NONFINITE, NBX_ERR_DEPTH,
 * NBX_ERR_NODEVICE, NBX_ERR_HIP. */
#define NBX_NOALLOC_MAX_BODIES 16384
#define NBX_NOALLOC_MAX_NODES 65536
nbx_status nbx_accel_noalloc(void* stream, int64_t n, const double* x, const double* y, const double* z,
                             const double* m, const int32_t* type, double theta, double* ax, double* ay,
                             double* az);

/* NBX_MODE_PHYSICS: gravitational potential of every body at the current
 * positions from the tree walk (self pair excluded), upload order. */
nbx_status nbx_potential(nbx_ctx* ctx, double* pot);

/* FP64 direct sum over all pairs with Plummer softening cfg.eps (G = 1):
 * the force reference for either mode.  Any output may be null. */
nbx_status nbx_direct(nbx_ctx* ctx, double* ax, double* ay, double* az, double* pot);

/* Energy and momentum of the current state, FP64 with a fixed reduction
 * order: out5 = {kinetic, potential, px, py, pz}.  source 0: potential from
 * the tree walk (NBX_MODE_PHYSICS), 1: from the FP64 direct sum.  Like
 * nbx_potential, it keeps the forces cached for the next step: a run with
 * these calls between steps is bit for bit the run without them. */
nbx_status nbx_energy(nbx_ctx* ctx, int source, double* out5);
/* The same per system: out receives 5 values per system. */
nbx_status nbx_energy_systems(nbx_ctx* ctx, int source, double* out);

/* Autotuner (physics mode; no counterpart in NEXT).  Times `steps` fixed-dt
 * steps of the current systems for each of the ncand candidate
 * (leaf[k], group[k]) pairs, restoring the state after every trial, and
 * keeps the fastest pair in the context's configuration.  ms_out (may be
 * null, ncand entries) receives each candidate's ms per step.  The state
 * the call leaves behind is the one it started from. */
nbx_status nbx_autotune(nbx_ctx* ctx, const int32_t* leaf, const int32_t* group, int32_t ncand, double dt,
                        int32_t steps, double* ms_out);
/* The same with the walk's list size as a third coordinate: candidate k is
 * (leaf[k], group[k], list[k]) with list[k] as nbx_set_walk_list takes it
 * (list may be null: the context's list size for every candidate). */
nbx_status nbx_autotune_lists(nbx_ctx* ctx, const int32_t* leaf, const int32_t* group, const int32_t* list,
                              int32_t ncand, double dt, int32_t steps, double* ms_out);

/* Physics-mode walk setting (no counterpart in NEXT): entries of the
 * per-wave LDS lists of accepted cells and bucket bodies that the walk
 * collects before it evaluates them, 96, 128 (the default), 192 or 256.
 * Longer lists evaluate more entries per flush with more lanes busy but hold
 * more LDS per wave (fewer resident waves).  The FP32 flush sums change with
 * it (the FP64 totals add the flushes).  NBX_ERR_ARG for another size or a
 * reference-mode context. */
nbx_status nbx_set_walk_list(nbx_ctx* ctx, int32_t entries);

/* Captured runs (no counterpart in NEXT).  With on = 1, nbx_run replays its
 * steps from a HIP graph of a period of steps (2, or 2k in physics mode with
 * the tree rebuilt every k-th step) whose tree builds never read counts back
 * to the host: buffers and launches are sized from the last build with
 * headroom, and a build that does not fit them (or needs the two-word keys,
 * or meets a non-finite position) raises a device flag after which the
 * run's tree kernels do nothing.  One host wait per run checks the flag; if
 * it is set, the state the run started from is restored and the run is made
 * step by step instead, so results are those of nbx_run without capture,
 * bit for bit, either way.  Timings per phase are not recorded inside a
 * captured run.  on = 2 sizes the bounded builds without headroom (a test of
 * the fallback); 0 turns capture off.  NBX_ERR_ARG for other values. */
nbx_status nbx_set_graph(nbx_ctx* ctx, int on);

/* Accumulated timings and counters. */
nbx_status nbx_get_timing(nbx_ctx* ctx, nbx_timing* out);
nbx_status nbx_reset_timing(nbx_ctx* ctx);
/* flags: 1 = record per-phase HIP events, 2 = count interactions. */
nbx_status nbx_set_profiling(nbx_ctx* ctx, int flags);

/* ---- parity hooks: the tree of the current positions ------------------ */

/* Root centre and half-width (step.h:93-102) and each body's octree
 * digits (octree.h:49-64): digits 0..20 in hi, 21..41 in lo, 3 bits each,
 * first digit most significant; upload order. */
nbx_status nbx_debug_keys(nbx_ctx* ctx, double* root4, uint64_t* hi, uint64_t* lo);

/* The body held in each device slot, in the engine's internal (key) order:
 * its upload index within the context (id) and its system (sys), n entries
 * each.  The ids of each system's slots form a permutation of that system's
 * upload indices; the download calls check this and refuse anything else
 * (NBX_ERR_STATE).  Test hook: no counterpart in NEXT. */
nbx_status nbx_debug_order(nbx_ctx* ctx, int32_t* id, int32_t* sys);

/* Number of tree nodes of the current positions. */
nbx_status nbx_debug_tree_size(nbx_ctx* ctx, int64_t* nodes);

/* The node table in depth-first preorder, children in digit order 0..7
 * (the order of NEXT's recursion), with: depth, leaf flag, body index
 * (upload order, -1 for internal nodes), child-digit mask, geometric
 * centre and half-width (geo4), mass, centre of mass (com3), quadrupole
 * (q6: xx yy zz xy xz yz), and perm = bodies in leaf order.  The moments
 * are NEXT's (octree.h:92-127) in both modes; physics mode lays these
 * records out on demand from its own build of the same tree (a tree
 * refitted since its last build is rebuilt first). */
nbx_status nbx_debug_tree(nbx_ctx* ctx, int32_t* depth, int32_t* leaf, int32_t* body,
                          int32_t* cmask, double* geo4, double* mass, double* com3, double* q6,
                          int32_t* perm);

/* The walk groups of the current positions, at most cap of them: first
 * body (an index into the leaf order of nbx_debug_tree's perm), body count
 * and system.  *ngroups receives the full count.  Reference mode: runs of
 * 64 consecutive bodies of a system; physics mode: groups packed from the
 * tree (at most group_size bodies inside one node). */
nbx_status nbx_debug_groups(nbx_ctx* ctx, int64_t cap, int32_t* first, int32_t* count, int32_t* sys,
                            int64_t* ngroups);

/* ---- serving: a continuous-batching request loop -----------------------
 * No counterpart in NEXT.  Each request is an independent simulation with
 * its own settings.  Requests are admitted first come, first served while
 * their estimated device memory fits the budget beside the admitted
 * requests' estimates, or beside what the batch contexts hold if that is
 * more (buffers grow with slack and are freed when a batch empties).
 * Admitted requests with the same settings share a batch context and step
 * together, joining and leaving between steps; batches with different
 * settings overlap on the device.  A request's result does not depend on the
 * others.  A batch context (a stream and a host thread) that has had no
 * request for 2 seconds is retired, so a stream of ever new settings keeps
 * only the batches it uses.  An engine failure inside a batch fails every
 * request of that batch (NBX_ERR_* of the failing call) and resets it; no
 * request receives another's bodies. */

typedef struct nbx_server nbx_server;

typedef struct nbx_request {
  int64_t n;
  const double *x, *y, *z, *vx, *vy, *vz, *m; /* copied at submission */
  const int32_t* type;
  int32_t nsteps;     /* KDK steps to make */
  double dt;          /* step (base step if adaptive_dt) */
  int mode;           /* nbx_mode */
  int precision;      /* nbx_precision */
  double theta;
  double eps;
  int leaf_size;
  int adaptive_dt;
} nbx_request;

typedef struct nbx_request_timing {
  double queue_ms;    /* submission to first step */
  double latency_ms;  /* submission to completion */
} nbx_request_timing;

typedef struct nbx_server_info {
  int64_t queued, running, batches;
  uint64_t bytes_in_use;  /* admission estimate of the admitted requests */
  int64_t body_steps;     /* bodies x steps completed */
  int64_t batch_steps;    /* batched steps launched */
  uint64_t device_bytes;  /* device memory the batch contexts hold, as of their last step */
} nbx_server_info;

nbx_status nbx_server_create(nbx_server** out, int device, uint64_t budget_bytes);
/* Stops the server: requests not yet done fail with NBX_ERR_STATE, threads
 * already inside nbx_server_wait on them return, and only then is the server
 * freed.  No call on the server may start once destroy has been called. */
void nbx_server_destroy(nbx_server* srv);
nbx_status nbx_server_submit(nbx_server* srv, const nbx_request* req, int64_t* id);
/* 0 queued, 1 running, 2 done, 3 failed; negative: unknown id. */
int nbx_server_poll(nbx_server* srv, int64_t id);
/* Block until the request is done, copy its final positions and velocities
 * (any pointer may be null) and release it; returns the request's status.
 * Of several threads waiting on one id, one receives the result and the
 * others NBX_ERR_ARG (the id is released). */
nbx_status nbx_server_wait(nbx_server* srv, int64_t id, double* x, double* y, double* z, double* vx, double* vy,
                           double* vz, nbx_request_timing* timing);
nbx_status nbx_server_stats(nbx_server* srv, nbx_server_info* out);

#ifdef __cplusplus
}
#endif

#endif /* NBX_H */
