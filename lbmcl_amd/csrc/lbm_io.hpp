// lbm_io.hpp -- names and codes shared by the dump writers of the lbmcl host
// (lbm_output.hpp): the dump file names (the iteration number zero-padded to
// the digit count of the iteration total, lbmcl.hpp:13, 159-334) and the
// map.dump cell codes (lbmcl.hpp:188-193).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <string>

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

// dir/stem<iteration, zero-padded to the digits of iterations>ext
inline std::string numbered_path(const std::string& dir, const char* stem, size_t iteration, size_t iterations,
                                 const char* ext) {
    std::string num = std::to_string(iteration);
    const size_t width = digits(iterations);
    if (num.size() < width) num.insert(0, width - num.size(), '0');
    return dir + "/" + stem + num + ext;
}

}  // namespace lbm_io
