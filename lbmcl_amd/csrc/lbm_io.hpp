// lbm_io.hpp -- the three dump formats of LBMCL, byte-compatible with the
// reference writers (lbmcl.hpp:159-334), written from host arrays.
//
//   write_map   map.dump      (lbmcl.hpp:159-203)
//   write_f     f_<it>.dump   (lbmcl.hpp:206-258)
//   write_vti   lbmcl.<it>.vti (lbmcl.hpp:261-334)
//
// Formatting goes through the same iostream manipulators as the reference so
// the bytes match: VTI values use std::scientific with 16 digits, f values
// std::fixed with 6 digits in a field of 8, and the iteration number in file
// names is zero-padded to the digit count of the iteration total.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <type_traits>

namespace lbm_io {

constexpr int Q = 19;
constexpr int DUMP_PRECISION = 6;   // lbmcl.hpp:15
constexpr int VTK_PRECISION = 16;   // lbmcl.hpp:16

// Number of decimal digits of val, 1 for 0 (lbmcl.hpp:13).
inline size_t digits(size_t val) { return val > 0 ? size_t(std::log10(double(val))) + 1 : 1; }

// map flags (common.h:7-20) and the dump codes of lbmcl.hpp:188-193
inline int map_code(int32_t t) {
    const int32_t boundary = 0x10 | 0x20 | 0x40 | 0x80 | 0x100 | 0x200;
    int val = 0;
    if (t == 0x1) val = 1;         // fluid
    if (t & 0x2) val = 2;          // moving
    if (t & boundary) val = 3;     // boundary (the lid, FRONT|MOVING, also prints 3)
    if (t == 0x8) val = 4;         // wall
    if (t == 0x4) val = 5;         // corner
    return val;
}

inline std::string numbered_path(const std::string& dir, const char* stem, size_t iteration, size_t iterations,
                                 const char* ext) {
    std::stringstream b;
    b << dir << "/" << stem << std::setw(int(digits(iterations))) << std::setfill('0') << iteration << ext;
    return b.str();
}

inline void write_map(const std::string& path, const int32_t* map, size_t dim) {
    std::ofstream dump(path);
    dump << "# FLUID       1" << std::endl
         << "# MOVING      2" << std::endl
         << "# BOUNDARY    3" << std::endl
         << "# WALL        4" << std::endl
         << "# CORNER      5" << std::endl
         << std::endl;
    for (size_t z = 0; z < dim; ++z) {
        for (size_t y = 0; y < dim; ++y) {
            for (size_t x = 0; x < dim; ++x) dump << map_code(map[x + y * dim + z * dim * dim]) << " ";
            dump << std::endl;
        }
        dump << std::endl;
    }
    dump << std::endl;
}

// f in the reference's CSoA layout with the given stride.
template <class T>
void write_f(const std::string& path, const T* f, size_t dim, size_t stride) {
    const size_t dim_digits = digits(dim);
    const size_t coord_spaces = dim_digits * 3 + 5;
    std::ofstream dump(path);
    for (size_t z = 0; z < dim; ++z) {
        for (size_t y = 0; y < dim; ++y) {
            for (size_t s = 0; s < coord_spaces; ++s) dump << " ";
            for (size_t q = 0; q < size_t(Q); ++q) dump << std::setw(DUMP_PRECISION + 2) << q << " ";
            dump << std::endl;
            for (size_t x = 0; x < dim; ++x) {
                const size_t id = x + y * dim + z * dim * dim;
                dump << std::setw(int(dim_digits)) << "(" << x << "," << y << "," << z << ") ";
                for (size_t q = 0; q < size_t(Q); ++q) {
                    const size_t slot = ((id / stride) * Q + q) * stride + (id & (stride - 1));
                    dump << std::fixed << std::setw(DUMP_PRECISION + 2) << std::setprecision(DUMP_PRECISION)
                         << f[slot] << " ";
                }
                dump << std::endl;
            }
            dump << std::endl;
        }
        dump << std::endl;
    }
    dump << std::endl;
}

// rho (dim^3) and u (3 dim^3, planar) on the window x, y, z in [1, dim-2].
template <class T>
void write_vti(const std::string& path, const T* rho, const T* u, size_t dim) {
    const size_t from = 1;
    const size_t to = dim - 1;
    const size_t extent = to - from - 1;
    const size_t n3 = dim * dim * dim;
    const std::string type = std::is_same<T, float>::value ? "Float32" : "Float64";
    std::ofstream vtk(path);
    vtk << "<?xml version=\"1.0\"?>\n"
        << "<VTKFile type=\"ImageData\" version=\"0.1\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
        << "  <ImageData WholeExtent=\"0 " << extent << " 0 " << extent << " 0 " << extent
        << "\" Origin=\"0 0 0\" Spacing=\"1 1 1\">\n"
        << "    <Piece Extent=\"0 " << extent << " 0 " << extent << " 0 " << extent << "\">\n"
        << "      <PointData Scalars=\"rho\">\n"
        << "        <DataArray type=\"" << type << "\" Name=\"rho\" NumberOfComponents=\"1\" format=\"ascii\">\n";
    for (size_t z = from; z < to; ++z) {
        for (size_t y = from; y < to; ++y) {
            for (size_t x = from; x < to; ++x)
                vtk << std::scientific << std::setprecision(VTK_PRECISION) << rho[x + y * dim + z * dim * dim] << " ";
            vtk << "\n";
        }
    }
    vtk << "        </DataArray>\n"
        << "        <DataArray type=\"" << type << "\" Name=\"v\" NumberOfComponents=\"3\" format=\"ascii\">\n";
    for (size_t z = from; z < to; ++z) {
        for (size_t y = from; y < to; ++y) {
            for (size_t x = from; x < to; ++x) {
                const size_t id = x + y * dim + z * dim * dim;
                vtk << std::scientific << std::setprecision(VTK_PRECISION) << u[id] << " "
                    << std::scientific << std::setprecision(VTK_PRECISION) << u[n3 + id] << " "
                    << std::scientific << std::setprecision(VTK_PRECISION) << u[2 * n3 + id] << " ";
            }
            vtk << "\n";
        }
    }
    vtk << "        </DataArray>\n"
        << "      </PointData>\n"
        << "    </Piece>\n"
        << "  </ImageData>\n"
        << "</VTKFile>\n";
}

}  // namespace lbm_io
