// Host-side particle set, text IC loader and VTU / legacy-VTK writers of the
// NEXT-compatible driver.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace nextio {

// Structure of arrays with the fields of NEXT's Particle (particle.h:22-52).
struct ParticleSet {
  std::vector<double> x, y, z, vx, vy, vz, m;
  std::vector<int32_t> type;
  size_t size() const { return x.size(); }
};

// 8 whitespace-separated columns per particle, "x y z vx vy vz m type",
// read with stream extraction until the first group that fails to parse
// (load_particle.hpp:100-116).  A missing file gives zero particles.
ParticleSet LoadText(const std::string& path);

// ASCII VTU UnstructuredGrid of vertices with type / velocity / mass point
// data, default stream formatting (vtu_save.h:22-78).  Returns false if the
// file cannot be opened.
bool SaveVTU(const ParticleSet& p, const std::string& path);

// Legacy VTK POLYDATA with the same fields (vtk_save.h:21-75), FP64 types.
bool SaveVTK(const ParticleSet& p, const std::string& path);

}  // namespace nextio
