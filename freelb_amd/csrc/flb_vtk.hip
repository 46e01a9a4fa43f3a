// Macroscopic read-out and the on-disk format (host code, C ABI).
//
// flb_store_moments is the read-out task moment::rhoU<CELL, true> applied to
// the cells whose flag matches a mask (lbm/moment.h:359-377, as
// examples/cavity3d/cavity3d.cpp:192-194 runs it before writing): rho and u
// of those cells go to the RHO / VELOCITY fields.
//
// flb_write_vtm writes blocks the way FreeLB's vtmo::vtmWriter::WriteBinary
// does (io/vtm_writer.h:288-430 with vtino::vtiManager, io/vti_writer.h:424-987):
// one ImageData .vti per block, `strip` layers cut off each side (the
// writer's overlap), Origin the first written cell centre, point data as
// base64 "binary" arrays (a uint32 byte count and then the values, each
// encoded on its own, io/base64.h) in the order RHO, VELOCITY (components
// interleaved), flags; and a .vtm that lists the .vti files. Names and the
// text of the files match FreeLB's, so ParaView sessions that open FreeLB's
// output open these unchanged.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/flb_abi.h"
#include "flb_host.hpp"

using namespace flbh;

namespace {

// RFC 4648 base64 with '=' padding, each call one complete encoding
// (Base64Encoder, io/base64.h: a block of 3 bytes to 4 characters).
void base64(std::ostream& os, const unsigned char* p, size_t n) {
  static const char tab[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
  size_t i = 0;
  for (; i + 3 <= n; i += 3) {
    const unsigned v = (unsigned(p[i]) << 16) | (unsigned(p[i + 1]) << 8) | unsigned(p[i + 2]);
    os << tab[(v >> 18) & 63] << tab[(v >> 12) & 63] << tab[(v >> 6) & 63] << tab[v & 63];
  }
  if (i < n) {
    const unsigned b0 = p[i], b1 = i + 1 < n ? p[i + 1] : 0u;
    const unsigned v = (b0 << 16) | (b1 << 8);
    os << tab[(v >> 18) & 63] << tab[(v >> 12) & 63] << (i + 1 < n ? tab[(v >> 6) & 63] : '=') << '=';
  }
}

template <class V>
void data_array(std::ostream& os, const char* type, const char* name, int ncomp, const std::vector<V>& v) {
  os << "<DataArray type=\"" << type << "\" Name=\"" << name << "\" "
     << "format=\"binary\" encoding=\"base64\" " << "NumberOfComponents=\"" << ncomp << "\">\n";
  const unsigned int bytes = (unsigned int)(v.size() * sizeof(V));
  base64(os, reinterpret_cast<const unsigned char*>(&bytes), sizeof bytes);
  base64(os, reinterpret_cast<const unsigned char*>(v.data()), v.size() * sizeof(V));
  os << "\n</DataArray>\n";
}

// T: the lattice's float type (how FreeLB prints origin and spacing)
template <class T>
int write_vti(flb_block b, const std::string& path, const double* min_centre, double voxel, int strip,
              const char* rho_name, const char* u_name, const char* flag_name) {
  const int d = b->d;
  const int nx = b->desc.nx, ny = b->desc.ny, nz = b->desc.nz;
  const int ext[3] = {nx - 1 - 2 * strip, ny - 1 - 2 * strip, d == 3 ? nz - 1 - 2 * strip : 0};
  if (ext[0] < 0 || ext[1] < 0 || ext[2] < 0) return fail(FLB_ERR_ARG, "strip %d leaves no cells", strip);
  const size_t N = size_t(b->N);
  std::vector<double> fields((size_t(d) + 1) * N);
  int rc = flb_download_fields(b, fields.data());
  if (rc) return rc;
  std::vector<T> rho, u;
  std::vector<uint8_t> fl;
  const int z0 = d == 3 ? strip : 0, z1 = d == 3 ? nz - strip : 1;
  for (int z = z0; z < z1; ++z)
    for (int y = strip; y < ny - strip; ++y)
      for (int x = strip; x < nx - strip; ++x) {
        const size_t id = size_t(x) + size_t(nx) * (size_t(y) + size_t(ny) * z);
        rho.push_back(T(fields[id]));
        for (int c = 0; c < d; ++c) u.push_back(T(fields[(size_t(c) + 1) * N + id]));
        fl.push_back(b->flags_set ? b->h_flags[id] : uint8_t(0));
      }
  const T org[3] = {T(min_centre[0] + strip * voxel), T(min_centre[1] + strip * voxel),
                    T(min_centre[2] + strip * voxel)};
  const T vox = T(voxel);
  std::ostringstream os;
  os << "<?xml version=\"1.0\"?>\n";
  os << "<VTKFile type=\"ImageData\" version=\"0.1\" " << "byte_order=\"LittleEndian\">\n";
  if (d == 2) {
    os << "<ImageData WholeExtent=\"" << 0 << " " << ext[0] << " " << 0 << " " << ext[1] << " " << 0 << " " << 0
       << "\" Origin=\"" << org[0] << " " << org[1] << " " << 0 << "\" Spacing=\"" << vox << " " << vox << " "
       << vox << "\">\n";
    os << "<Piece Extent=\"" << 0 << " " << ext[0] << " " << 0 << " " << ext[1] << " " << 0 << " " << 0
       << "\">\n";
  } else {
    os << "<ImageData WholeExtent=\"" << 0 << " " << ext[0] << " " << 0 << " " << ext[1] << " " << 0 << " "
       << ext[2] << "\" Origin=\"" << org[0] << " " << org[1] << " " << org[2] << "\" Spacing=\"" << vox << " "
       << vox << " " << vox << "\">\n";
    os << "<Piece Extent=\"" << 0 << " " << ext[0] << " " << 0 << " " << ext[1] << " " << 0 << " " << ext[2]
       << "\">\n";
  }
  os << "<PointData>\n";
  const char* ftype = sizeof(T) == 8 ? "Float64" : "Float32";
  if (rho_name) data_array(os, ftype, rho_name, 1, rho);
  if (u_name) data_array(os, ftype, u_name, d, u);
  if (flag_name) data_array(os, "UInt8", flag_name, 1, fl);
  os << "</PointData>\n" << "</Piece>\n" << "</ImageData>\n" << "</VTKFile>\n";
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) return fail(FLB_ERR_ARG, "cannot open %s for writing", path.c_str());
  const std::string text = os.str();
  f.write(text.data(), std::streamsize(text.size()));
  return f ? FLB_OK : fail(FLB_ERR_ARG, "write to %s failed", path.c_str());
}

}  // namespace

extern "C" {

int flb_store_moments(flb_block b, uint8_t mask) {
  CHECK_BLOCK(b);
  if (!b->flags_set) return fail(FLB_ERR_STATE, "flags not uploaded");
  const size_t N = size_t(b->N), d = size_t(b->d);
  std::vector<double> fields((d + 1) * N);
  int rc = flb_download_fields(b, fields.data());
  if (rc) return rc;
  std::vector<char> mom((d + 1) * N * b->real);
  if ((rc = flb_moments(b, mask, mom.data()))) return rc;
  for (size_t id = 0; id < N; ++id) {
    if (!(b->h_flags[id] & mask)) continue;
    for (size_t c = 0; c <= d; ++c) {
      const size_t k = c * N + id;
      fields[k] = b->real == 8 ? reinterpret_cast<const double*>(mom.data())[k]
                               : double(reinterpret_cast<const float*>(mom.data())[k]);
    }
  }
  if ((rc = flb_set_field_rho(b, fields.data()))) return rc;
  return flb_set_field_u(b, fields.data() + N);
}

int flb_write_vtm(const flb_block* blocks, int n, const double* min_centres, const double* voxels,
                  const int* strips, const char* dir, const char* name, int step, const char* rho_name,
                  const char* u_name, const char* flag_name) {
  if (!blocks || n < 1 || !min_centres || !voxels || !dir || !name)
    return fail(FLB_ERR_ARG, "null argument or no blocks");
  const std::string base = std::string(dir) + (std::strlen(dir) && dir[std::strlen(dir) - 1] != '/' ? "/" : "");
  std::vector<std::string> vtis;
  for (int k = 0; k < n; ++k) {
    if (!blocks[k]) return fail(FLB_ERR_ARG, "block %d is null", k);
    if (blocks[k]->pending) return fail(FLB_ERR_STATE, "block %d is between ApplyCellDynamics and Stream", k);
    const std::string vti = std::string(name) + (step >= 0 ? "_T" + std::to_string(step) : std::string()) + "_B" +
                            std::to_string(k) + ".vti";
    const int strip = strips ? strips[k] : 0;
    const int rc = blocks[k]->real == 8
                       ? write_vti<double>(blocks[k], base + vti, min_centres + 3 * k, voxels[k], strip, rho_name,
                                           u_name, flag_name)
                       : write_vti<float>(blocks[k], base + vti, min_centres + 3 * k, voxels[k], strip, rho_name,
                                          u_name, flag_name);
    if (rc) return rc;
    vtis.push_back(vti);
  }
  const std::string vtm = base + name + (step >= 0 ? std::to_string(step) : std::string()) + ".vtm";
  std::ofstream f(vtm, std::ios::trunc);
  if (!f) return fail(FLB_ERR_ARG, "cannot open %s for writing", vtm.c_str());
  f << "<?xml version=\"1.0\"?>\n";
  f << "<VTKFile type=\"vtkMultiBlockDataSet\" version=\"1.0\" " << "byte_order=\"LittleEndian\">\n"
    << "<vtkMultiBlockDataSet>\n";
  for (int k = 0; k < n; ++k) {
    f << "<Block index=\"" << k << "\" >\n";
    f << "<DataSet index= \"0\" " << "file=\"" << vtis[size_t(k)] << "\">\n" << "</DataSet>\n";
    f << "</Block>\n";
  }
  f << "</vtkMultiBlockDataSet>\n";
  f << "</VTKFile>\n";
  return f ? FLB_OK : fail(FLB_ERR_ARG, "write to %s failed", vtm.c_str());
}

}  // extern "C"
