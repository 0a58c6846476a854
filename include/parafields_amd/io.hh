// io.hh -- the on-disk format of a random field, host side: the
// `<name>.stoch.h5` HDF5 file of the stochastic part (dataset /stochastic,
// IEEE 64-bit little-endian, dims slowest axis first), its `<name>.xdmf`
// description for visualisation, the `<name>.trend` coefficients and the
// `<name>.field` configuration (io.hh:55-402 and randomfield.hh:706-713 of
// the reference, which write the same files through parallel HDF5; one GPU
// holds the whole field, so the file is written serially).
//
// HDF5 is found through <hdf5.h>; without it writeStochasticHDF5 and
// readStochasticHDF5 throw, as the reference does without HDF5 support.
#pragma once

#include <array>
#include <fstream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#if defined(__has_include)
#if __has_include(<hdf5.h>)
#include <hdf5.h>
#define PARAFIELDS_AMD_HAVE_HDF5 1
#endif
#endif

namespace parafields_amd {

inline bool fileExists(const std::string& filename)
{
  std::ifstream f(filename);
  return f.good();
}

#ifdef PARAFIELDS_AMD_HAVE_HDF5
template <typename RF>
inline hid_t hdf5_memtype();
template <>
inline hid_t hdf5_memtype<float>()
{
  return H5T_NATIVE_FLOAT;
}
template <>
inline hid_t hdf5_memtype<double>()
{
  return H5T_NATIVE_DOUBLE;
}
#endif

// writeParallelToHDF5 (io.hh:169-282) on one process: dataset `name` of the
// global size `cells` (x fastest in memory; the file lists dims slowest
// first), stored as IEEE F64LE whatever RF is.
template <typename RF, size_t dim>
void writeStochasticHDF5(const std::array<unsigned int, dim>& cells, const std::vector<RF>& data,
                         const std::string& name, const std::string& filename)
{
#ifdef PARAFIELDS_AMD_HAVE_HDF5
  size_t n = 1;
  hsize_t dims[dim];
  for (size_t i = 0; i < dim; i++) {
    dims[dim - 1 - i] = cells[i];
    n *= cells[i];
  }
  if (data.size() != n)
    throw std::runtime_error("field size does not match the grid");
  const hid_t file = H5Fcreate(filename.c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0)
    throw std::runtime_error("cannot create " + filename);
  const hid_t space = H5Screate_simple((int)dim, dims, nullptr);
  const hid_t dset =
    H5Dcreate2(file, name.c_str(), H5T_IEEE_F64LE, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  herr_t st = dset < 0 ? -1
                       : H5Dwrite(dset, hdf5_memtype<RF>(), H5S_ALL, H5S_ALL, H5P_DEFAULT,
                                  data.data());
  if (dset >= 0)
    H5Dclose(dset);
  H5Sclose(space);
  H5Fclose(file);
  if (st < 0)
    throw std::runtime_error("writing " + name + " to " + filename + " failed");
#else
  (void)cells;
  (void)data;
  (void)name;
  (void)filename;
  throw std::runtime_error("Writing and reading field files requires HDF5 support");
#endif
}

// readParallelFromHDF5 (io.hh:55-151) on one process: the whole dataset,
// whose dims must equal `cells` (slowest first in the file).
template <typename RF, size_t dim>
std::vector<RF> readStochasticHDF5(const std::array<unsigned int, dim>& cells,
                                   const std::string& name, const std::string& filename)
{
#ifdef PARAFIELDS_AMD_HAVE_HDF5
  if (!fileExists(filename))
    throw std::runtime_error("File is missing: " + filename);
  const hid_t file = H5Fopen(filename.c_str(), H5F_ACC_RDONLY, H5P_DEFAULT);
  if (file < 0)
    throw std::runtime_error("cannot open " + filename);
  const hid_t dset = H5Dopen2(file, name.c_str(), H5P_DEFAULT);
  if (dset < 0) {
    H5Fclose(file);
    throw std::runtime_error("no dataset " + name + " in " + filename);
  }
  const hid_t space = H5Dget_space(dset);
  const int nd = H5Sget_simple_extent_ndims(space);
  hsize_t dims[8] = {};
  bool ok = nd == (int)dim && H5Sget_simple_extent_dims(space, dims, nullptr) >= 0;
  size_t n = 1;
  for (size_t i = 0; ok && i < dim; i++) {
    ok = dims[dim - 1 - i] == cells[i];
    n *= cells[i];
  }
  std::vector<RF> out;
  if (ok) {
    out.resize(n);
    ok = H5Dread(dset, hdf5_memtype<RF>(), H5S_ALL, H5S_ALL, H5P_DEFAULT, out.data()) >= 0;
  }
  H5Sclose(space);
  H5Dclose(dset);
  H5Fclose(file);
  if (!ok)
    throw std::runtime_error("dataset " + name + " in " + filename +
                             " does not match the grid of the field");
  return out;
#else
  (void)cells;
  (void)name;
  (void)filename;
  throw std::runtime_error("Writing and reading field files requires HDF5 support");
#endif
}

// writeToXDMF (io.hh:304-402): a structured-grid description of the HDF5
// dataset with the cell sizes, for dims 1 to 3 (lower dims padded with the
// first cell size) and the temporal collection of 3D slices for dim 4.
template <typename RF, size_t dim>
bool writeToXDMF(const std::array<unsigned int, dim>& cells, const std::array<RF, dim>& ext,
                 const std::string& fileName)
{
  if (dim > 4)
    return false;
  std::ofstream f(fileName + ".xdmf", std::ofstream::trunc);
  f << "<?xml version=\"1.0\" ?>\n<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n"
    << "<Xdmf Version=\"2.0\">\n <Domain>\n";
  auto reversed = [&](size_t m) {
    std::ostringstream s;
    for (size_t i = 0; i < m; i++)
      s << cells[m - 1 - i] << " ";
    return s.str();
  };
  auto spacing = [&](size_t m) {
    std::ostringstream s;
    for (size_t i = 0; i < m; i++)
      s << ext[i] / cells[i] << " ";
    return s.str();
  };
  if (dim < 4) {
    std::ostringstream pad;
    for (size_t i = 0; i < 3 - dim; i++)
      pad << ext[0] / cells[0] << " ";
    f << "  <Grid Name=\"StructuredGrid\">\n"
      << "   <Topology TopologyType=\"3DRectMesh\" NumberOfElements=\"" << reversed(dim)
      << "\"/>\n"
      << "   <Geometry GeometryType=\"origin_dxdydz\">\n"
      << "    <DataItem Dimensions=\"3\">\n     0. 0. 0.\n    </DataItem>\n"
      << "    <DataItem Dimensions=\"3\">\n     " << pad.str() << spacing(dim) << "\n"
      << "    </DataItem>\n   </Geometry>\n"
      << "   <Attribute Name=\"" << fileName << "\" Center=\"Cell\">\n"
      << "    <DataItem Dimensions=\"" << reversed(dim) << "\" Format=\"HDF\">\n"
      << "     " << fileName << ".stoch.h5:/stochastic\n"
      << "    </DataItem>\n   </Attribute>\n  </Grid>\n";
  } else {
    f << "  <Grid Name=\"GridTime\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
    for (unsigned int j = 0; j < cells[dim - 1]; j++) {
      std::ostringstream slab;
      for (size_t i = 0; i + 1 < dim; i++)
        slab << cells[dim - 2 - i] << " ";
      f << "   <Grid Name=\"StructuredGrid\">\n"
        << "    <Time Value=\"" << j << "\"/>\n"
        << "    <Topology TopologyType=\"3DRectMesh\" NumberOfElements=\"" << slab.str()
        << "\"/>\n"
        << "    <Geometry GeometryType=\"origin_dxdydz\">\n"
        << "     <DataItem Dimensions=\"3\">\n      0. 0. 0.\n     </DataItem>\n"
        << "     <DataItem Dimensions=\"3\">\n      " << spacing(dim - 1) << "\n"
        << "     </DataItem>\n    </Geometry>\n"
        << "    <Attribute Name=\"" << fileName << "\" Center=\"Cell\">\n"
        << "     <DataItem ItemType=\"HyperSlab\" Dimensions=\"" << slab.str() << "1\">\n"
        << "      <DataItem Dimensions=\"3 4\">\n       " << j << " 0 0 0\n"
        << "       1 1 1 1\n       1 " << slab.str() << "\n      </DataItem>\n"
        << "      <DataItem Dimensions=\"" << reversed(dim) << "\" Format=\"HDF\">\n"
        << "       " << fileName << ".stoch.h5:/stochastic\n"
        << "      </DataItem>\n     </DataItem>\n    </Attribute>\n   </Grid>\n";
    }
    f << "  </Grid>\n";
  }
  f << " </Domain>\n</Xdmf>" << std::endl;
  return true;
}

} // namespace parafields_amd
