// lbm_io_reference.hpp -- the reference's three dump writers restated with the
// same iostream manipulators (test infrastructure: the byte-format reference
// that tests/test_formats.py compares the product's threaded writers with).
//
// Portions restate LBMCL's writers (blackwut/LBMCL, lbmcl.hpp:159-334,
// GPL-3.0-or-later; see LICENSE at the repository root):
//
//   write_map   map.dump      (lbmcl.hpp:159-203)
//   write_f     f_<it>.dump   (lbmcl.hpp:206-258)
//   write_vti   lbmcl.<it>.vti (lbmcl.hpp:261-334)
//
// Formatting goes through the same iostream manipulators as the reference so
// the bytes match: VTI values use std::scientific with 16 digits, f values
// std::fixed with 6 digits in a field of 8.
#pragma once

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <type_traits>

#include "lbm_io.hpp"

namespace lbm_io {

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
