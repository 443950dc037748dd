// Text IC loader and VTU / VTK snapshot writers.
//
// The snapshot text must be byte-identical to NEXT's for the same values, so
// the element order, separators and number formatting (an std::ostream at
// its default precision of 6 significant digits) follow vtu_save.h:22-78 and
// vtk_save.h:21-75.  Points are written as the reference writes them: one
// long line per array in VTU, one line per particle in VTK.
#include "io.hpp"

#include <fstream>
#include <sstream>

namespace nextio {

ParticleSet LoadText(const std::string& path) {
  ParticleSet p;
  std::ifstream in(path);
  double v[7];
  int t;
  while (in >> v[0] >> v[1] >> v[2] >> v[3] >> v[4] >> v[5] >> v[6] >> t) {
    p.x.push_back(v[0]); p.y.push_back(v[1]); p.z.push_back(v[2]);
    p.vx.push_back(v[3]); p.vy.push_back(v[4]); p.vz.push_back(v[5]);
    p.m.push_back(v[6]);
    p.type.push_back(t);
  }
  return p;
}

namespace {

// Writes through a string buffer and one file write: the text is large
// (about 100 bytes per particle) and this keeps snapshotting I/O-bound.
bool flush_to(const std::string& path, const std::ostringstream& s) {
  std::ofstream out(path, std::ios::binary);
  if (!out) return false;
  const std::string& b = s.str();
  out.write(b.data(), (std::streamsize)b.size());
  return (bool)out;
}

}  // namespace

bool SaveVTU(const ParticleSet& p, const std::string& path) {
  const size_t n = p.size();
  std::ostringstream o;
  o << "<?xml version=\"1.0\"?>\n"
    << "<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">\n"
    << "  <UnstructuredGrid>\n"
    << "    <Piece NumberOfPoints=\"" << n << "\" NumberOfCells=\"" << n << "\">\n"
    << "      <Points>\n"
    << "        <DataArray type=\"Float32\" NumberOfComponents=\"3\" format=\"ascii\">\n          ";
  for (size_t i = 0; i < n; ++i) o << p.x[i] << ' ' << p.y[i] << ' ' << p.z[i] << ' ';
  o << "\n        </DataArray>\n      </Points>\n"
    << "      <Cells>\n"
    << "        <DataArray type=\"Int32\" Name=\"connectivity\" format=\"ascii\">\n          ";
  for (size_t i = 0; i < n; ++i) o << i << ' ';
  o << "\n        </DataArray>\n"
    << "        <DataArray type=\"Int32\" Name=\"offsets\" format=\"ascii\">\n          ";
  for (size_t i = 1; i <= n; ++i) o << i << ' ';
  o << "\n        </DataArray>\n"
    << "        <DataArray type=\"UInt8\" Name=\"types\" format=\"ascii\">\n          ";
  for (size_t i = 0; i < n; ++i) o << "1 ";
  o << "\n        </DataArray>\n      </Cells>\n"
    << "      <PointData>\n"
    << "        <DataArray type=\"Int32\" Name=\"type\" format=\"ascii\">\n          ";
  for (size_t i = 0; i < n; ++i) o << p.type[i] << ' ';
  o << "\n        </DataArray>\n"
    << "        <DataArray type=\"Float32\" Name=\"velocity\" NumberOfComponents=\"3\" format=\"ascii\">\n          ";
  for (size_t i = 0; i < n; ++i) o << p.vx[i] << ' ' << p.vy[i] << ' ' << p.vz[i] << ' ';
  o << "\n        </DataArray>\n"
    << "        <DataArray type=\"Float32\" Name=\"mass\" format=\"ascii\">\n          ";
  for (size_t i = 0; i < n; ++i) o << p.m[i] << ' ';
  o << "\n        </DataArray>\n"
    << "      </PointData>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
  return flush_to(path, o);
}

bool SaveVTK(const ParticleSet& p, const std::string& path) {
  const size_t n = p.size();
  std::ostringstream o;
  o << "# vtk DataFile Version 3.0\nNEXT snapshot\nASCII\nDATASET POLYDATA\n";
  o << "POINTS " << n << " double\n";
  for (size_t i = 0; i < n; ++i) o << p.x[i] << ' ' << p.y[i] << ' ' << p.z[i] << '\n';
  o << "VERTICES " << n << ' ' << n * 2 << '\n';
  for (size_t i = 0; i < n; ++i) o << "1 " << i << '\n';
  o << "POINT_DATA " << n << '\n';
  o << "SCALARS type int 1\nLOOKUP_TABLE default\n";
  for (size_t i = 0; i < n; ++i) o << p.type[i] << '\n';
  o << "VECTORS velocity double\n";
  for (size_t i = 0; i < n; ++i) o << p.vx[i] << ' ' << p.vy[i] << ' ' << p.vz[i] << '\n';
  o << "SCALARS mass double 1\nLOOKUP_TABLE default\n";
  for (size_t i = 0; i < n; ++i) o << p.m[i] << '\n';
  return flush_to(path, o);
}

}  // namespace nextio
