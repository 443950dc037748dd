// CPU oracle: an FP64 restatement of NEXT's gravity + time-step path.
//
// TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
// cpu_baseline leg of bench.py load this library, and only as the checker or
// the reported CPU baseline.  The product (libnbx.so and the `next` binary)
// never links, loads or calls it.
//
// Every function restates the reference arithmetic in the reference's order
// of evaluation; build with -ffp-contract=off so no product is fused into an
// FMA.  The tree is built the way the reference builds it (one body per leaf,
// recursive splitting, octree.h:69-87), from an index pool instead of
// per-node heap allocations; the GPU path builds the same tree from sorted
// keys, so the two constructions check each other.
//
// Pinning (what checks each function, beyond the golden vectors written by
// the compiled reference in tests/golden/):
//   orc_root          closed form min/max; golden root.f64
//   orc_keys          golden keys_hi/keys_lo; tie-on-centre-plane cases
//   orc_tree_*        golden preorder node table (topology, m, COM, Q);
//                     2- and 3-body closed forms; canonical-prefix rule
//   orc_tree_accel    golden acc.f64 at theta 0.3/0.5/0.7; theta=0 equals
//                     brute force with leaf softening; N=2 analytic
//   orc_adaptive_dt   closed-form clamps; golden dt sequence
//   orc_step/orc_run  golden multi-step state and dt sequence
//   orc_direct        Plummer-softened direct sum: two-body closed form and
//                     Newton's third law (total force ~ 0)
//   orc_direct_subset equality with orc_direct on the same targets
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

struct Node {
  double gx, gy, gz, s;          // geometric centre and half-width
  double m, cx, cy, cz;          // mass and centre of mass
  double qxx, qyy, qzz, qxy, qxz, qyz;
  int32_t child[8];
  int32_t body;                  // -1: none
  bool leaf;
};

constexpr int kMaxDepth = 256;   // the reference recurses without bound

struct Tree {
  std::vector<Node> nodes;
  int max_depth = 0;
  int err = 0;                   // 1: coincident bodies / depth cap

  int make(double gx, double gy, double gz, double s) {
    Node n;
    n.gx = gx; n.gy = gy; n.gz = gz; n.s = s;
    n.m = 0; n.cx = n.cy = n.cz = 0;
    n.qxx = n.qyy = n.qzz = n.qxy = n.qxz = n.qyz = 0;
    for (int c = 0; c < 8; ++c) n.child[c] = -1;
    n.body = -1;
    n.leaf = true;
    nodes.push_back(n);
    return (int)nodes.size() - 1;
  }
};

// Octant digit against the node's geometric centre, ties low (octree.h:49-51).
inline int octant(const Node& n, double px, double py, double pz) {
  return (px > n.gx ? 1 : 0) + (py > n.gy ? 2 : 0) + (pz > n.gz ? 4 : 0);
}

// Child of octant o: centre = parent centre +/- half of the half-width
// (octree.h:56-64).
inline int child_of(Tree& t, int p, int o) {
  if (t.nodes[p].child[o] < 0) {
    const Node& n = t.nodes[p];
    double hs = n.s * 0.5;
    double gx = n.gx + ((o & 1) ? hs : -hs);
    double gy = n.gy + ((o & 2) ? hs : -hs);
    double gz = n.gz + ((o & 4) ? hs : -hs);
    int c = t.make(gx, gy, gz, hs);
    t.nodes[p].child[o] = c;
  }
  return t.nodes[p].child[o];
}

// Insertion with splitting (octree.h:69-87): an empty leaf takes the body; an
// occupied leaf becomes internal, pushes its body down, then the new body
// descends.  Written as a loop with a depth cap.
void insert(Tree& t, int i, const double* x, const double* y, const double* z) {
  int node = 0, depth = 0;
  for (;;) {
    if (depth > kMaxDepth) { t.err = 1; return; }
    if (t.nodes[node].leaf && t.nodes[node].body < 0) {
      t.nodes[node].body = i;
      t.max_depth = std::max(t.max_depth, depth);
      return;
    }
    if (t.nodes[node].leaf) {
      t.nodes[node].leaf = false;
      int old = t.nodes[node].body;
      t.nodes[node].body = -1;
      int o = octant(t.nodes[node], x[old], y[old], z[old]);
      int c = child_of(t, node, o);
      t.nodes[c].body = old;        // fresh child: an empty leaf
      t.max_depth = std::max(t.max_depth, depth + 1);
    }
    int o = octant(t.nodes[node], x[i], y[i], z[i]);
    node = child_of(t, node, o);
    ++depth;
  }
}

// Post-order mass, centre of mass and the one-level softened quadrupole
// (octree.h:92-127).
void upward(Tree& t, int id, const double* x, const double* y, const double* z, const double* m) {
  Node& n = t.nodes[id];
  if (n.leaf) {
    if (n.body >= 0) { n.m = m[n.body]; n.cx = x[n.body]; n.cy = y[n.body]; n.cz = z[n.body]; }
    else { n.m = 0; n.cx = n.cy = n.cz = 0; }
    n.qxx = n.qyy = n.qzz = n.qxy = n.qxz = n.qyz = 0;
    return;
  }
  double M = 0, sx = 0, sy = 0, sz = 0;
  for (int c = 0; c < 8; ++c) {
    int k = t.nodes[id].child[c];
    if (k < 0) continue;
    upward(t, k, x, y, z, m);
    const Node& ch = t.nodes[k];
    if (ch.m == 0) continue;
    M += ch.m;
    sx += ch.cx * ch.m; sy += ch.cy * ch.m; sz += ch.cz * ch.m;
  }
  Node& p = t.nodes[id];
  p.m = M; p.cx = sx; p.cy = sy; p.cz = sz;
  if (M > 0) { p.cx = sx / M; p.cy = sy / M; p.cz = sz / M; }
  double qxx = 0, qyy = 0, qzz = 0, qxy = 0, qxz = 0, qyz = 0;
  const double soft = p.s * p.s * 0.01;
  for (int c = 0; c < 8; ++c) {
    int k = p.child[c];
    if (k < 0) continue;
    const Node& ch = t.nodes[k];
    if (ch.m == 0) continue;
    double rx = ch.cx - p.cx, ry = ch.cy - p.cy, rz = ch.cz - p.cz;
    double r2 = rx * rx + ry * ry + rz * rz + soft;
    double mc = ch.m;
    qxx += mc * (3 * rx * rx - r2);
    qyy += mc * (3 * ry * ry - r2);
    qzz += mc * (3 * rz * rz - r2);
    qxy += mc * (3 * rx * ry);
    qxz += mc * (3 * rx * rz);
    qyz += mc * (3 * ry * rz);
  }
  p.qxx = qxx; p.qyy = qyy; p.qzz = qzz; p.qxy = qxy; p.qxz = qxz; p.qyz = qyz;
}

// Root geometry of the tree-build lambda (step.h:68-102).
void root_geometry(int64_t n, const double* x, const double* y, const double* z, double out[4]) {
  double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
  for (int64_t i = 0; i < n; ++i) {
    const double p[3] = {x[i], y[i], z[i]};
    for (int k = 0; k < 3; ++k) {
      if (p[k] < lo[k]) lo[k] = p[k];     // std::min(lo, p) keeps lo unless p < lo
      if (hi[k] < p[k]) hi[k] = p[k];     // std::max(hi, p) keeps hi unless hi < p
    }
  }
  out[0] = (lo[0] + hi[0]) * 0.5;
  out[1] = (lo[1] + hi[1]) * 0.5;
  out[2] = (lo[2] + hi[2]) * 0.5;
  double e = hi[0] - lo[0];
  if (e < hi[1] - lo[1]) e = hi[1] - lo[1];
  if (e < hi[2] - lo[2]) e = hi[2] - lo[2];
  double s = e * 0.5;
  if (s <= 0) s = 1.0;
  out[3] = s;
}

Tree build(int64_t n, const double* x, const double* y, const double* z, const double* m) {
  Tree t;
  double r[4];
  root_geometry(n, x, y, z, r);
  t.nodes.reserve((size_t)(2 * n + 16));
  t.make(r[0], r[1], r[2], r[3]);
  for (int64_t i = 0; i < n && !t.err; ++i) insert(t, (int)i, x, y, z);
  if (!t.err && n > 0) upward(t, 0, x, y, z, m);
  return t;
}

// Per-interaction softening (softening.h:21-35).
inline double softening(double size, double mass, double dist) {
  double e_size = size * 0.015;
  double e_mass = std::cbrt(mass) * 0.002;
  double taper = 1.0 / (1.0 + dist * 10.0);
  double e = (e_size + e_mass) * taper;
  return e < 1e-4 ? 1e-4 : e;             // std::max(e, 1e-4)
}

// Depth-first walk for target i (octree.h:133-179).
void walk(const Tree& t, int id, int i, const double* x, const double* y, const double* z,
          const double* m, const int32_t* type, double theta, double& ax, double& ay, double& az) {
  const Node& n = t.nodes[id];
  if (n.m == 0) return;
  if (n.leaf && n.body == i) return;
  double dx = n.cx - x[i], dy = n.cy - y[i], dz = n.cz - z[i];
  double r2 = dx * dx + dy * dy + dz * dz;
  double dist = std::sqrt(r2 + 1e-20);
  double eps = softening(n.s, n.m, dist);
  if (type[i] == 1) {
    double alt = 2.0 * n.s / std::pow(n.m / m[i], 0.333333333);
    if (eps < alt) eps = alt;
  }
  double r2s = r2 + eps * eps;
  double dinv = 1.0 / std::sqrt(r2s);
  if (n.leaf || (n.s / dist) < theta) {
    double inv3 = dinv * dinv * dinv;
    double fac = 1.0 * n.m * inv3;
    ax += dx * fac; ay += dy * fac; az += dz * fac;
    double inv5 = inv3 * (dinv * dinv);
    double inv7 = inv5 * (dinv * dinv);
    double q = n.qxx * dx * dx + n.qyy * dy * dy + n.qzz * dz * dz +
               2 * (n.qxy * dx * dy + n.qxz * dx * dz + n.qyz * dy * dz);
    double qrx = 2 * (n.qxx * dx + n.qxy * dy + n.qxz * dz);
    double qry = 2 * (n.qxy * dx + n.qyy * dy + n.qyz * dz);
    double qrz = 2 * (n.qxz * dx + n.qyz * dy + n.qzz * dz);
    ax += (1.0 * 0.5) * (qrx * inv5 - 5 * q * inv7 * dx);
    ay += (1.0 * 0.5) * (qry * inv5 - 5 * q * inv7 * dy);
    az += (1.0 * 0.5) * (qrz * inv5 - 5 * q * inv7 * dz);
    return;
  }
  for (int c = 0; c < 8; ++c)
    if (n.child[c] >= 0) walk(t, n.child[c], i, x, y, z, m, type, theta, ax, ay, az);
}

void preorder(const Tree& t, int id, int depth, int64_t& k, int32_t* dep, int32_t* leaf,
              int32_t* body, int32_t* cmask, double* geo, double* mass, double* com, double* q,
              int32_t* perm, int64_t& pk) {
  const Node& n = t.nodes[id];
  int mask = 0;
  for (int c = 0; c < 8; ++c) if (n.child[c] >= 0) mask |= 1 << c;
  dep[k] = depth; leaf[k] = n.leaf; body[k] = n.body; cmask[k] = mask;
  geo[4 * k] = n.gx; geo[4 * k + 1] = n.gy; geo[4 * k + 2] = n.gz; geo[4 * k + 3] = n.s;
  mass[k] = n.m;
  com[3 * k] = n.cx; com[3 * k + 1] = n.cy; com[3 * k + 2] = n.cz;
  double* qq = q + 6 * k;
  qq[0] = n.qxx; qq[1] = n.qyy; qq[2] = n.qzz; qq[3] = n.qxy; qq[4] = n.qxz; qq[5] = n.qyz;
  ++k;
  if (n.leaf && n.body >= 0) perm[pk++] = n.body;
  for (int c = 0; c < 8; ++c)
    if (n.child[c] >= 0)
      preorder(t, n.child[c], depth + 1, k, dep, leaf, body, cmask, geo, mass, com, q, perm, pk);
}

void accel_all(const Tree& t, int64_t n, const double* x, const double* y, const double* z,
               const double* m, const int32_t* type, double theta, double* ax, double* ay,
               double* az, int threads) {
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (int64_t i = 0; i < n; ++i) {
    double a0 = 0, a1 = 0, a2 = 0;
    walk(t, 0, (int)i, x, y, z, m, type, theta, a0, a1, a2);
    ax[i] = a0; ay[i] = a1; az[i] = a2;
  }
}

}  // namespace

extern "C" {

struct orc_tree { Tree t; };

void orc_root(int64_t n, const double* x, const double* y, const double* z, double* out4) {
  root_geometry(n, x, y, z, out4);
}

// Digit keys by replaying the node-centre recurrence from the root
// (octree.h:49-64): digit l of particle i is its octant in the depth-l node on
// its path; digits 0..20 go to hi (3 bits each, first digit most significant),
// digits 21..41 to lo.
void orc_keys(int64_t n, const double* x, const double* y, const double* z, const double* root4,
              int levels, uint64_t* hi, uint64_t* lo) {
  for (int64_t i = 0; i < n; ++i) {
    double cx = root4[0], cy = root4[1], cz = root4[2], s = root4[3];
    uint64_t h = 0, l = 0;
    for (int d = 0; d < levels && d < 42; ++d) {
      int o = (x[i] > cx ? 1 : 0) + (y[i] > cy ? 2 : 0) + (z[i] > cz ? 4 : 0);
      if (d < 21) h = (h << 3) | (uint64_t)o; else l = (l << 3) | (uint64_t)o;
      double hs = s * 0.5;
      cx = cx + ((o & 1) ? hs : -hs);
      cy = cy + ((o & 2) ? hs : -hs);
      cz = cz + ((o & 4) ? hs : -hs);
      s = hs;
    }
    hi[i] = h;
    lo[i] = l;
  }
}

orc_tree* orc_tree_build(int64_t n, const double* x, const double* y, const double* z, const double* m) {
  orc_tree* t = new orc_tree;
  t->t = build(n, x, y, z, m);
  return t;
}
int orc_tree_error(const orc_tree* t) { return t->t.err; }
int64_t orc_tree_nodes(const orc_tree* t) { return (int64_t)t->t.nodes.size(); }
int orc_tree_max_depth(const orc_tree* t) { return t->t.max_depth; }

void orc_tree_export(const orc_tree* t, int32_t* depth, int32_t* leaf, int32_t* body, int32_t* cmask,
                     double* geo4, double* mass, double* com3, double* q6, int32_t* perm) {
  int64_t k = 0, pk = 0;
  preorder(t->t, 0, 0, k, depth, leaf, body, cmask, geo4, mass, com3, q6, perm, pk);
}

void orc_tree_accel(const orc_tree* t, int64_t n, const double* x, const double* y, const double* z,
                    const double* m, const int32_t* type, double theta, double* ax, double* ay,
                    double* az, int threads) {
  accel_all(t->t, n, x, y, z, m, type, theta, ax, ay, az, threads);
}

// The walk for a subset of targets (indices into the particle arrays);
// results go to ax[k], ay[k], az[k] for targets[k].
void orc_tree_accel_subset(const orc_tree* t, const double* x, const double* y, const double* z,
                           const double* m, const int32_t* type, double theta, int64_t nt,
                           const int64_t* targets, double* ax, double* ay, double* az, int threads) {
#pragma omp parallel for schedule(dynamic, 16) num_threads(threads)
  for (int64_t k = 0; k < nt; ++k) {
    double a0 = 0, a1 = 0, a2 = 0;
    walk(t->t, 0, (int)targets[k], x, y, z, m, type, theta, a0, a1, a2);
    ax[k] = a0; ay[k] = a1; az[k] = a2;
  }
}

void orc_tree_free(orc_tree* t) { delete t; }

// Global time step from the largest speed (adaptive.h:23-52).
double orc_adaptive_dt(int64_t n, const double* vx, const double* vy, const double* vz, double base) {
  double v2max = 0;
  for (int64_t i = 0; i < n; ++i) {
    double s = vx[i] * vx[i] + vy[i] * vy[i] + vz[i] * vz[i];
    if (s > v2max) v2max = s;
  }
  double v = std::sqrt(v2max);
  if (1e4 < v) v = 1e4;                    // std::min(v, 1e4)
  if (v < 1e-8) return base;
  double dt = base / (1 + v);
  if (dt < base * 0.01) dt = base * 0.01;  // std::max(dt, 0.01 base)
  if (base * 1.0 < dt) dt = base * 1.0;    // std::min(dt, base)
  return dt;
}

// One KDK step with the reference's structure (step.h:25-204, non-MPI):
// build + walk + half kick, drift, build + walk + half kick.
// Returns 0, or 1 if a build hit the depth cap (coincident bodies).
int orc_step(int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
             const double* m, const int32_t* type, double dt, double theta, int threads) {
  if (n == 0) return 0;
  const double half = dt * 0.5;
  std::vector<double> ax(n), ay(n), az(n);
  for (int pass = 0; pass < 2; ++pass) {
    Tree t = build(n, x, y, z, m);
    if (t.err) return 1;
    accel_all(t, n, x, y, z, m, type, theta, ax.data(), ay.data(), az.data(), threads);
    for (int64_t i = 0; i < n; ++i) {
      vx[i] += ax[i] * half; vy[i] += ay[i] * half; vz[i] += az[i] * half;
    }
    if (pass == 0) {
#pragma omp parallel for schedule(static) num_threads(threads)
      for (int64_t i = 0; i < n; ++i) {
        x[i] += vx[i] * dt; y[i] += vy[i] * dt; z[i] += vz[i] * dt;
      }
    }
  }
  return 0;
}

// The driver loop body (begrun.cpp:96-99): dt from the current speeds, then
// one step; dts_out (may be null) receives each dt.
int orc_run(int64_t n, double* x, double* y, double* z, double* vx, double* vy, double* vz,
            const double* m, const int32_t* type, double base_dt, int nsteps, double theta,
            int threads, double* dts_out) {
  for (int s = 0; s < nsteps; ++s) {
    double dt = orc_adaptive_dt(n, vx, vy, vz, base_dt);
    if (dts_out) dts_out[s] = dt;
    if (orc_step(n, x, y, z, vx, vy, vz, m, type, dt, theta, threads)) return 1;
  }
  return 0;
}

// FP64 direct sum with Plummer softening eps (G = 1): the force reference for
// the mixed-precision tree.  pot may be null.
static void direct_one(int64_t n, const double* x, const double* y, const double* z, const double* m, double e2,
                       int64_t i, double* out4) {
  double a0 = 0, a1 = 0, a2 = 0, p = 0;
  for (int64_t j = 0; j < n; ++j) {
    if (j == i) continue;
    double dx = x[j] - x[i], dy = y[j] - y[i], dz = z[j] - z[i];
    double r2 = dx * dx + dy * dy + dz * dz + e2;
    double rinv = 1.0 / std::sqrt(r2);
    double mr3 = m[j] * rinv * rinv * rinv;
    a0 += dx * mr3; a1 += dy * mr3; a2 += dz * mr3;
    p -= m[j] * rinv;
  }
  out4[0] = a0; out4[1] = a1; out4[2] = a2; out4[3] = p;
}

void orc_direct(int64_t n, const double* x, const double* y, const double* z, const double* m,
                double eps, double* ax, double* ay, double* az, double* pot, int threads) {
  const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads)
  for (int64_t i = 0; i < n; ++i) {
    double o[4];
    direct_one(n, x, y, z, m, e2, i, o);
    ax[i] = o[0]; ay[i] = o[1]; az[i] = o[2];
    if (pot) pot[i] = o[3];
  }
}

// The same sum for the targets tg[0..nt) only (bodies of all n act on each),
// so the force error of a large system can be sampled without the O(n^2)
// sum over every target.  Pinned by equality with orc_direct on those
// targets (the same per-target loop).
void orc_direct_subset(int64_t n, const double* x, const double* y, const double* z, const double* m,
                       double eps, int64_t nt, const int64_t* tg, double* ax, double* ay, double* az, double* pot,
                       int threads) {
  const double e2 = eps * eps;
#pragma omp parallel for schedule(dynamic, 4) num_threads(threads)
  for (int64_t k = 0; k < nt; ++k) {
    double o[4];
    direct_one(n, x, y, z, m, e2, tg[k], o);
    ax[k] = o[0]; ay[k] = o[1]; az[k] = o[2];
    if (pot) pot[k] = o[3];
  }
}

}  // extern "C"
