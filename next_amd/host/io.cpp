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

#ifdef NBX_HDF5
#include <hdf5.h>
#endif

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

#ifdef NBX_HDF5

namespace {

// One PartTypeN group into the set (load_particle.hpp:23-85): Coordinates
// and Velocities (N x 3) and Masses (N) read as FP32, every body of the
// group gets `type`.  A group without Coordinates adds nothing.
void load_part_type(hid_t file, const std::string& group, int type, ParticleSet& p) {
  if (H5Lexists(file, (group + "/Coordinates").c_str(), H5P_DEFAULT) <= 0) return;
  const hid_t dc = H5Dopen(file, (group + "/Coordinates").c_str(), H5P_DEFAULT);
  const hid_t sc = H5Dget_space(dc);
  hsize_t dims[2] = {0, 0};
  H5Sget_simple_extent_dims(sc, dims, nullptr);
  const size_t n = (size_t)dims[0];
  std::vector<float> pos(n * 3), vel(n * 3), mass(n);
  H5Dread(dc, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, pos.data());
  const hid_t dv = H5Dopen(file, (group + "/Velocities").c_str(), H5P_DEFAULT);
  H5Dread(dv, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, vel.data());
  const hid_t dm = H5Dopen(file, (group + "/Masses").c_str(), H5P_DEFAULT);
  H5Dread(dm, H5T_NATIVE_FLOAT, H5S_ALL, H5S_ALL, H5P_DEFAULT, mass.data());
  for (size_t i = 0; i < n; ++i) {
    p.x.push_back(pos[3 * i]); p.y.push_back(pos[3 * i + 1]); p.z.push_back(pos[3 * i + 2]);
    p.vx.push_back(vel[3 * i]); p.vy.push_back(vel[3 * i + 1]); p.vz.push_back(vel[3 * i + 2]);
    p.m.push_back(mass[i]);
    p.type.push_back(type);
  }
  H5Dclose(dc);
  H5Dclose(dv);
  H5Dclose(dm);
  H5Sclose(sc);
}

}  // namespace

bool LoadHDF5(const std::string& path, ParticleSet& p) {
  H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);   // a text file is not an error here
  const hid_t file = H5Fopen(path.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0) return false;
  load_part_type(file, "PartType1", 1, p);   // dark matter
  load_part_type(file, "PartType4", 0, p);   // stars
  H5Fclose(file);
  return true;
}

bool SaveHDF5(const ParticleSet& p, const std::string& path) {
  const size_t n = p.size();
  if (n == 0) return true;                   // hdf5_save.cpp:27: nothing written
  const hid_t file = H5Fcreate(path.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0) return false;
  const hid_t group = H5Gcreate(file, "PartType1", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  std::vector<float> pos(n * 3), vel(n * 3);
  std::vector<int> ids(n);
  for (size_t i = 0; i < n; ++i) {
    pos[3 * i] = (float)p.x[i]; pos[3 * i + 1] = (float)p.y[i]; pos[3 * i + 2] = (float)p.z[i];
    vel[3 * i] = (float)p.vx[i]; vel[3 * i + 1] = (float)p.vy[i]; vel[3 * i + 2] = (float)p.vz[i];
    ids[i] = (int)i + 1;
  }
  const hsize_t d3[2] = {n, 3}, d1[1] = {n};
  const hid_t s3 = H5Screate_simple(2, d3, nullptr), s1 = H5Screate_simple(1, d1, nullptr);
  auto put = [&](const char* name, hid_t type, hid_t space, const void* data) {
    const hid_t d = H5Dcreate(group, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Dwrite(d, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data);
    H5Dclose(d);
  };
  put("Coordinates", H5T_NATIVE_FLOAT, s3, pos.data());
  put("Velocities", H5T_NATIVE_FLOAT, s3, vel.data());
  put("Masses", H5T_NATIVE_DOUBLE, s1, p.m.data());
  put("ParticleIDs", H5T_NATIVE_INT, s1, ids.data());
  H5Sclose(s3);
  H5Sclose(s1);
  H5Gclose(group);
  H5Fclose(file);
  // XDMF sidecar next to it (hdf5_save.cpp:80-99)
  std::ostringstream x;
  x << "<?xml version=\"1.0\" ?>\n<Xdmf Version=\"3.0\">\n  <Domain>\n"
    << "    <Grid Name=\"Particles\" GridType=\"Uniform\">\n"
    << "      <Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" << n << "\"/>\n"
    << "      <Geometry GeometryType=\"XYZ\">\n"
    << "        <DataItem Dimensions=\"" << n << " 3\" NumberType=\"Float\" Precision=\"4\" Format=\"HDF\">\n"
    << "          " << path << ":/PartType1/Coordinates\n"
    << "        </DataItem>\n      </Geometry>\n"
    << "      <Attribute Name=\"Mass\" AttributeType=\"Scalar\" Center=\"Node\">\n"
    << "        <DataItem Dimensions=\"" << n << "\" NumberType=\"Float\" Precision=\"8\" Format=\"HDF\">\n"
    << "          " << path << ":/PartType1/Masses\n"
    << "        </DataItem>\n      </Attribute>\n"
    << "    </Grid>\n  </Domain>\n</Xdmf>\n";
  return flush_to(path.substr(0, path.find_last_of('.')) + ".xdmf", x);
}

#endif

ParticleSet LoadParticles(const std::string& path) {
  ParticleSet p;
#ifdef NBX_HDF5
  if (LoadHDF5(path, p)) return p;
#endif
  return LoadText(path);
}

}  // namespace nextio
