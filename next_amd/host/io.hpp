// Host-side particle set, IC loaders and snapshot writers (VTU, legacy VTK,
// HDF5) of the NEXT-compatible driver.
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

// HDF5 snapshots (built with libhdf5, NBX_HDF5).  LoadHDF5 returns false if
// the file is not HDF5; otherwise it appends PartType1 as type 1 and
// PartType4 as type 0, reading coordinates, velocities and masses as FP32
// (load_particle.hpp:23-98).  SaveHDF5 writes every body to PartType1
// (FP32 Coordinates and Velocities, FP64 Masses, ParticleIDs 1..N) plus an
// XDMF sidecar beside it (hdf5_save.cpp:23-99).
#ifdef NBX_HDF5
bool LoadHDF5(const std::string& path, ParticleSet& p);
bool SaveHDF5(const ParticleSet& p, const std::string& path);
#endif

// NEXT's loader: HDF5 when the file is HDF5 (and the build has it), the
// 8-column text format otherwise.
ParticleSet LoadParticles(const std::string& path);

}  // namespace nextio
