/*
 * nbx_serve_ops.h - the engine calls the request server makes, as a table.
 *
 * No counterpart in NEXT (which runs one simulation per process).  The
 * request server (nbx_server_* in nbx.h) drives batch contexts through
 * exactly these nine calls.  nbx_server_create uses this library's engine;
 * nbx_server_create_with_ops puts the same scheduler (admission by budget,
 * batching, joining and leaving between steps, failure handling, idle-batch
 * retirement) over another table: a wrapper around the engine (logging,
 * accounting) or a host-only stand-in for testing the scheduler on a machine
 * without a GPU.  Every entry must be set; the calls have the signatures and
 * meaning of the nbx.h functions of the same names, and each context is used
 * from one thread at a time.
 */
#ifndef NBX_SERVE_OPS_H
#define NBX_SERVE_OPS_H

#include "nbx.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct nbx_engine_ops {
  nbx_status (*create)(nbx_ctx** out, const nbx_config* cfg);
  void (*destroy)(nbx_ctx* ctx);
  const char* (*last_error)(const nbx_ctx* ctx);
  nbx_status (*add_system)(nbx_ctx* ctx, int64_t n, const double* x, const double* y, const double* z,
                           const double* vx, const double* vy, const double* vz, const double* m,
                           const int32_t* type, int32_t* sys_out);
  nbx_status (*remove_system)(nbx_ctx* ctx, int32_t s, double* x, double* y, double* z, double* vx, double* vy,
                              double* vz);
  nbx_status (*run)(nbx_ctx* ctx, double base_dt, int32_t nsteps, double* dts_out);
  nbx_status (*upload_batch)(nbx_ctx* ctx, int32_t nsys, const int64_t* counts, const double* x, const double* y,
                             const double* z, const double* vx, const double* vy, const double* vz,
                             const double* m, const int32_t* type);
  nbx_status (*trim)(nbx_ctx* ctx);
  uint64_t (*device_bytes)(const nbx_ctx* ctx);
} nbx_engine_ops;

/* As nbx_server_create, over the given engine table (copied). */
nbx_status nbx_server_create_with_ops(nbx_server** out, int device, uint64_t budget_bytes,
                                      const nbx_engine_ops* ops);

#ifdef __cplusplus
}
#endif

#endif /* NBX_SERVE_OPS_H */
