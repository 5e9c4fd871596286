// lbm_lattice.hpp -- D3Q19 lattice tables and cell classification, shared by
// host and device code of liblbm_hip.so.
//
// Semantics follow the reference: cell flags common.h:7-20, predicates
// common.h:23-66, classification kernels.cl:236-266, velocities and opposites
// kernels.cl:131-226, weights kernels.cl:110-128, CSoA slot kernels.cl:64.
#pragma once

#include <cstdint>

#ifndef LBM_HD
#define LBM_HD __host__ __device__ __forceinline__
#endif

namespace lbm {

constexpr int Q = 19;

// Cell flags, bit-compatible with the reference's map buffer (common.h:7-18).
constexpr int32_t NONE = 0x000;
constexpr int32_t FLUID = 0x001;
constexpr int32_t MOVING = 0x002;
constexpr int32_t CORNER = 0x004;
constexpr int32_t WALL = 0x008;
constexpr int32_t LEFT = 0x010;
constexpr int32_t RIGHT = 0x020;
constexpr int32_t TOP = 0x040;
constexpr int32_t BOTTOM = 0x080;
constexpr int32_t FRONT = 0x100;
constexpr int32_t BACK = 0x200;
constexpr int32_t BOUNDARY_MASK = LEFT | RIGHT | BOTTOM | TOP | BACK | FRONT;

LBM_HD bool is_fluid(int32_t t) { return t == FLUID; }
LBM_HD bool is_wall(int32_t t) { return t == WALL; }
LBM_HD bool is_boundary(int32_t t) { return (t & BOUNDARY_MASK) != 0; }
LBM_HD bool is_moving(int32_t t) { return (t & MOVING) != 0; }
LBM_HD bool is_moving_init(int32_t t) { return (t & FRONT) != 0; }
LBM_HD bool is_store_macro(int32_t t) { return t == FLUID || (t & MOVING) != 0; }
LBM_HD bool is_bounceback(int32_t t) { return is_boundary(t) && !(t & MOVING); }

// Flags of the x coordinate alone and of the (y, z) row alone; a cell's type is
// finish_type(xflags | rowflags).  Wall coordinates yield WALL.
LBM_HD int32_t x_flags(int32_t x, int32_t n) {
    if (x == 0 || x == n - 1) return WALL;
    return (x == 1 ? LEFT : 0) | (x == n - 2 ? RIGHT : 0);
}
LBM_HD int32_t row_flags(int32_t y, int32_t z, int32_t n) {
    if (y == 0 || y == n - 1 || z == 0 || z == n - 1) return WALL;
    return (y == 1 ? BOTTOM : 0) | (y == n - 2 ? TOP : 0) | (z == 1 ? BACK : 0) | (z == n - 2 ? FRONT : 0);
}
LBM_HD int32_t finish_type(int32_t t) {
    if (t & WALL) return WALL;
    if (t == (LEFT | BACK | BOTTOM) || t == (RIGHT | BACK | BOTTOM) || t == (LEFT | BACK | TOP) ||
        t == (RIGHT | BACK | TOP))
        return CORNER;
    if (t == FRONT) return FRONT | MOVING;
    if (t == NONE) return FLUID;
    return t;
}
LBM_HD int32_t cell_type(int32_t x, int32_t y, int32_t z, int32_t n) {
    return finish_type(x_flags(x, n) | row_flags(y, z, n));
}

// CSoA slot (kernels.cl:64) in 64-bit arithmetic, written as
// 19 c - 18 (c mod S) + q S, which equals ((c / S) * 19 + q) * S + c mod S.
LBM_HD int64_t slot(int64_t c, int q, int64_t stride) {
    return 19 * c - 18 * (c & (stride - 1)) + int64_t(q) * stride;
}

}  // namespace lbm
