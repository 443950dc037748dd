// Driver around NEXT's own FP64 gravity + time-step headers.
//
// TEST INFRASTRUCTURE ONLY.  This file is compiled by oracle/Makefile against
// the reference headers where they lie (src/gravity/step.h, src/gravity/octree.h,
// src/dt/adaptive.h, src/io/vtu_save.h, src/io/vtk_save.h) into oracle/_ref/.
// It is used to write golden vectors under tests/golden/ and as the timed
// "reference" CPU baseline.  Nothing in the product links or runs it.
//
// Sub-commands (all FP64, NEXT_FP64):
//   tree  <ic.txt> <theta> <levels> <outdir>
//         root geometry, per-particle digit keys (root->leaf path of the
//         reference tree, extended with Octree::getOctant/createChild to
//         <levels> digits), DFS leaf order, preorder node table and the
//         accelerations of bhAccel at <theta>.
//   steps <ic.txt> <base_dt> <nsteps> <threads> <outdir>
//         the begrun.cpp:96-99 loop body (computeAdaptiveDt + Step) nsteps
//         times; writes the final state and the dt sequence.
//   time  <ic.txt> <threads> <base_dt> <warmup> <nsteps>
//         the same loop, timed per step; prints one JSON line.
//   vtu|vtk <ic.txt> <out>
//         SaveVTU / SaveVTK of the loaded particles.
#include "gravity/step.h"
#include "dt/adaptive.h"
#include "io/vtu_save.h"
#include "io/vtk_save.h"

#include <omp.h>
#include <algorithm>
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

namespace {

// Text IC reader with the stop-at-first-failure semantics of
// load_particle.hpp:100-116 (the header itself pulls in HDF5).  A path
// ending in ".bin" is read as raw arrays instead: int64 n, then x y z vx vy
// vz m (n doubles each) and type (n int32).
Particle load_text(const char* path) {
  Particle p;
  const std::string ps(path);
  if (ps.size() > 4 && ps.compare(ps.size() - 4, 4, ".bin") == 0) {
    FILE* f = std::fopen(path, "rb");
    int64_t n = 0;
    if (!f || std::fread(&n, 8, 1, f) != 1) { if (f) std::fclose(f); return p; }
    std::vector<double> cols(7 * (size_t)n);
    std::vector<int32_t> t((size_t)n);
    if (std::fread(cols.data(), 8, cols.size(), f) != cols.size() ||
        std::fread(t.data(), 4, t.size(), f) != t.size()) { std::fclose(f); return p; }
    std::fclose(f);
    for (int64_t i = 0; i < n; ++i)
      p.addParticle(cols[i], cols[n + i], cols[2 * n + i], cols[3 * n + i], cols[4 * n + i],
                    cols[5 * n + i], cols[6 * n + i], t[i]);
    return p;
  }
  std::ifstream in(path);
  double a[7];
  int t;
  while (in >> a[0] >> a[1] >> a[2] >> a[3] >> a[4] >> a[5] >> a[6] >> t)
    p.addParticle(a[0], a[1], a[2], a[3], a[4], a[5], a[6], t);
  return p;
}

template <class T>
void dump(const std::string& dir, const char* name, const std::vector<T>& v) {
  std::string path = dir + "/" + name;
  FILE* f = std::fopen(path.c_str(), "wb");
  if (!f) { std::perror(path.c_str()); std::exit(2); }
  if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
  std::fclose(f);
}

// Root set-up of the buildTree lambda (step.h:67-102), restated because the
// lambda is local to Step.
std::unique_ptr<Octree> make_root(const Particle& ps, double out[4]) {
  const int n = (int)ps.size();
  double lo[3] = {1e30, 1e30, 1e30}, hi[3] = {-1e30, -1e30, -1e30};
  for (int i = 0; i < n; ++i) {
    lo[0] = std::min(lo[0], ps.x[i]); lo[1] = std::min(lo[1], ps.y[i]);
    lo[2] = std::min(lo[2], ps.z[i]);
    hi[0] = std::max(hi[0], ps.x[i]); hi[1] = std::max(hi[1], ps.y[i]);
    hi[2] = std::max(hi[2], ps.z[i]);
  }
  double cx = (lo[0] + hi[0]) * 0.5, cy = (lo[1] + hi[1]) * 0.5, cz = (lo[2] + hi[2]) * 0.5;
  double s = std::max({hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]}) * 0.5;
  if (s <= 0) s = 1.0;
  out[0] = cx; out[1] = cy; out[2] = cz; out[3] = s;
  return std::make_unique<Octree>(cx, cy, cz, s);
}

struct NodeDump {
  std::vector<int32_t> depth, leaf, body, cmask;
  std::vector<double> geo, mass, com, q;
};

void preorder(const Octree* n, int d, NodeDump& nd, std::vector<int32_t>& perm) {
  int mask = 0;
  for (int c = 0; c < 8; ++c) if (n->child[c]) mask |= 1 << c;
  nd.depth.push_back(d);
  nd.leaf.push_back(n->leaf ? 1 : 0);
  nd.body.push_back(n->bodyIdx);
  nd.cmask.push_back(mask);
  nd.geo.insert(nd.geo.end(), {n->x, n->y, n->z, n->size});
  nd.mass.push_back(n->m);
  nd.com.insert(nd.com.end(), {n->cx, n->cy, n->cz});
  nd.q.insert(nd.q.end(), {n->Qxx, n->Qyy, n->Qzz, n->Qxy, n->Qxz, n->Qyz});
  if (n->leaf && n->bodyIdx >= 0) perm.push_back(n->bodyIdx);
  for (int c = 0; c < 8; ++c)
    if (n->child[c]) preorder(n->child[c].get(), d + 1, nd, perm);
}

int cmd_tree(int argc, char** argv) {
  if (argc != 6) return 1;
  Particle ps = load_text(argv[2]);
  const double theta = std::atof(argv[3]);
  const int levels = std::atoi(argv[4]);
  const std::string out = argv[5];
  const int n = (int)ps.size();
  std::vector<double> root(4, 0.0);
  auto tree = make_root(ps, root.data());
  for (int i = 0; i < n; ++i) tree->insert(i, ps);
  tree->computeMass(ps);
  dump(out, "root.f64", root);

  // Keys: digits of the root->leaf path, then extended with the reference's
  // own octant / child-centre arithmetic down to `levels` digits.
  std::vector<uint64_t> khi(n, 0), klo(n, 0);
  std::vector<int32_t> leafdepth(n, 0);
  for (int i = 0; i < n; ++i) {
    std::vector<int> dig;
    const Octree* node = tree.get();
    while (!node->leaf) {
      int o = node->getOctant(ps.x[i], ps.y[i], ps.z[i]);
      dig.push_back(o);
      node = node->child[o].get();
    }
    leafdepth[i] = (int)dig.size();
    auto cur = std::make_unique<Octree>(node->x, node->y, node->z, node->size);
    while ((int)dig.size() < levels) {
      int o = cur->getOctant(ps.x[i], ps.y[i], ps.z[i]);
      dig.push_back(o);
      cur = cur->createChild(o);
    }
    for (int l = 0; l < levels && l < 42; ++l) {
      if (l < 21) khi[i] = (khi[i] << 3) | (uint64_t)dig[l];
      else klo[i] = (klo[i] << 3) | (uint64_t)dig[l];
    }
  }
  dump(out, "keys_hi.u64", khi);
  dump(out, "keys_lo.u64", klo);
  dump(out, "leafdepth.i32", leafdepth);

  NodeDump nd;
  std::vector<int32_t> perm;
  if (n > 0) preorder(tree.get(), 0, nd, perm);
  dump(out, "perm.i32", perm);
  dump(out, "node_depth.i32", nd.depth);
  dump(out, "node_leaf.i32", nd.leaf);
  dump(out, "node_body.i32", nd.body);
  dump(out, "node_cmask.i32", nd.cmask);
  dump(out, "node_geo.f64", nd.geo);
  dump(out, "node_mass.f64", nd.mass);
  dump(out, "node_com.f64", nd.com);
  dump(out, "node_q.f64", nd.q);

  std::vector<double> acc(3 * (size_t)n, 0.0);
  for (int i = 0; i < n; ++i) {
    double ax = 0, ay = 0, az = 0;
    bhAccel(tree.get(), i, ps, theta, ax, ay, az);
    acc[3 * i] = ax; acc[3 * i + 1] = ay; acc[3 * i + 2] = az;
  }
  dump(out, "acc.f64", acc);
  return 0;
}

int cmd_steps(int argc, char** argv) {
  if (argc != 7) return 1;
  Particle ps = load_text(argv[2]);
  const double base = std::atof(argv[3]);
  const int nsteps = std::atoi(argv[4]);
  omp_set_num_threads(std::atoi(argv[5]));
  const std::string out = argv[6];
  std::vector<double> dts;
  for (int s = 0; s < nsteps; ++s) {
    double dt = computeAdaptiveDt(ps, base);
    Step(ps, dt);
    dts.push_back(dt);
  }
  dump(out, "dts.f64", dts);
  std::vector<double> st;
  for (auto* v : {&ps.x, &ps.y, &ps.z, &ps.vx, &ps.vy, &ps.vz}) st.insert(st.end(), v->begin(), v->end());
  dump(out, "state.f64", st);
  return 0;
}

int cmd_time(int argc, char** argv) {
  if (argc != 7) return 1;
  Particle ps = load_text(argv[2]);
  const int threads = std::atoi(argv[3]);
  const double base = std::atof(argv[4]);
  const int warm = std::atoi(argv[5]), nsteps = std::atoi(argv[6]);
  omp_set_num_threads(threads);
  for (int s = 0; s < warm; ++s) Step(ps, computeAdaptiveDt(ps, base));
  std::vector<double> ms;
  for (int s = 0; s < nsteps; ++s) {
    auto t0 = std::chrono::steady_clock::now();
    Step(ps, computeAdaptiveDt(ps, base));
    auto t1 = std::chrono::steady_clock::now();
    ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
  }
  std::vector<double> s = ms;
  std::sort(s.begin(), s.end());
  double med = s.empty() ? 0 : s[s.size() / 2];
  double tot = 0;
  for (double v : ms) tot += v;
  std::printf("{\"n\": %zu, \"threads\": %d, \"steps\": %d, \"ms_median\": %.6f, \"ms_total\": %.6f}\n",
              ps.size(), threads, nsteps, med, tot);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 2) { std::fprintf(stderr, "usage: ref_harness tree|steps|time|vtu|vtk ...\n"); return 1; }
  std::string c = argv[1];
  int rc = 1;
  if (c == "tree") rc = cmd_tree(argc, argv);
  else if (c == "steps") rc = cmd_steps(argc, argv);
  else if (c == "time") rc = cmd_time(argc, argv);
  else if ((c == "vtu" || c == "vtk") && argc == 4) {
    Particle ps = load_text(argv[2]);
    if (c == "vtu") SaveVTU(ps, argv[3]); else SaveVTK(ps, argv[3]);
    rc = 0;
  }
  if (rc) std::fprintf(stderr, "ref_harness: bad arguments for '%s'\n", c.c_str());
  return rc;
}
