// nbx_next.hpp - NEXT's gravity + time-step interface over the GPU engine.
//
// Header-only C++ mirror of the two calls NEXT's driver makes each iteration
// (begrun.cpp:97-98):
//
//   real computeAdaptiveDt(const Particle& p, real base_dt);   adaptive.h:23
//   void Step(ParticleSystem& ps, real dt);                    step.h:25
//
// with the same argument meaning, on any structure of arrays with NEXT's
// fields (x, y, z, vx, vy, vz, m as contiguous doubles, type as ints;
// particle.h:22-52).  The particle state stays on the GPU between calls:
// the first call uploads it, Step downloads positions and velocities back
// into the caller's arrays after the step so the host copy stays current,
// exactly as NEXT's in-place update.  The host arrays are re-uploaded only
// when the particle count changes or `invalidate()` is called.
//
// Errors: NEXT has no error returns; this mirror throws nbx_next::Error
// (std::runtime_error) carrying the nbx_status and message.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "nbx.h"

namespace nbx_next {

struct Error : std::runtime_error {
  nbx_status status;
  Error(nbx_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

// One device-resident particle system.
class Session {
 public:
  explicit Session(double theta = 0.5, int device = 0) {
    nbx_config cfg;
    nbx_config_default(&cfg);
    cfg.theta = theta;
    cfg.device = device;
    const nbx_status s = nbx_create(&ctx_, &cfg);
    if (s != NBX_OK) throw Error(s, "nbx_create failed");
  }
  ~Session() { nbx_destroy(ctx_); }
  Session(const Session&) = delete;
  Session& operator=(const Session&) = delete;

  void invalidate() { n_ = -1; }

  template <class PS>
  void sync_in(const PS& p) {
    const int64_t n = (int64_t)p.x.size();
    if (n == n_) return;
    type_.assign(p.type.begin(), p.type.end());
    check(nbx_upload(ctx_, n, p.x.data(), p.y.data(), p.z.data(), p.vx.data(), p.vy.data(), p.vz.data(),
                     p.m.data(), type_.data()));
    n_ = n;
  }

  // computeAdaptiveDt (adaptive.h:23-52).
  template <class PS>
  double adaptive_dt(const PS& p, double base_dt) {
    sync_in(p);
    double dt = base_dt;
    check(nbx_adaptive_dt(ctx_, base_dt, &dt));
    return dt;
  }

  // Step (step.h:25-204): one KDK step; positions and velocities are
  // written back into the caller's arrays.
  template <class PS>
  void step(PS& ps, double dt) {
    sync_in(ps);
    if (n_ == 0) return;
    check(nbx_step(ctx_, dt));
    check(nbx_download(ctx_, ps.x.data(), ps.y.data(), ps.z.data(), ps.vx.data(), ps.vy.data(), ps.vz.data()));
  }

  nbx_ctx* handle() { return ctx_; }

 private:
  void check(nbx_status s) {
    if (s != NBX_OK) throw Error(s, nbx_last_error(ctx_));
  }
  nbx_ctx* ctx_ = nullptr;
  int64_t n_ = -1;
  std::vector<int32_t> type_;
};

inline Session& default_session() {
  static Session s;
  return s;
}

// The two free functions with NEXT's signatures, on the process-wide session.
template <class PS>
double computeAdaptiveDt(const PS& p, double base_dt) {
  return default_session().adaptive_dt(p, base_dt);
}

template <class PS>
void Step(PS& ps, double dt) {
  default_session().step(ps, dt);
}

}  // namespace nbx_next
