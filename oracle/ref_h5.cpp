// Driver around NEXT's HDF5 loader and writer.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (src/io/load_particle.hpp,
// src/io/hdf5_save.cpp, src/io/vtk_save.h) and the image's libhdf5 into
// oracle/_ref/.  It writes the HDF5 golden files under tests/golden/.
// Nothing in the product links or runs it.
//
// Sub-commands:
//   h5save <ic> <out.hdf5>   LoadParticlesFromFile(ic), then SaveHDF5
//                            (also writes the .xdmf sidecar)
//   h5load <in> <out.vtk>    LoadParticlesFromFile(in), then SaveVTK
#include "io/load_particle.hpp"
#include "io/hdf5_save.h"
#include "io/vtk_save.h"

#include <cstdio>
#include <string>

int main(int argc, char** argv) {
  if (argc != 4) { std::fprintf(stderr, "usage: ref_h5 h5save|h5load <in> <out>\n"); return 1; }
  const std::string c = argv[1];
  Particle p = LoadParticlesFromFile(argv[2]);
  if (c == "h5save") SaveHDF5(p, argv[3]);
  else if (c == "h5load") SaveVTK(p, argv[3]);
  else { std::fprintf(stderr, "unknown command\n"); return 1; }
  std::printf("%zu\n", p.x.size());
  return 0;
}
