// lbm_io_selftest.cpp -- writes the three dump formats from deterministic
// synthetic arrays, plus the arrays themselves (rho.bin, u.bin, f.bin,
// map.bin), so the formats can be checked without a GPU.  Each file is
// written twice: by the iostream writers (lbm_io_reference.hpp) and, under the prefix
// "fast_", by the threaded writers of the output path (lbm_output.hpp) from
// the snapshot layouts the device produces (packed VTI window, cell-major f).
// The arrays include NaN of both signs, infinities, signed zero, subnormals,
// large magnitudes and values whose 17th significant digit is a rounding tie.
//
// usage: lbm_io_selftest <outdir> <dim> <stride> <f32|f64> <iteration> <iterations> [threads]
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "lbm_io_reference.hpp"
#include "lbm_output.hpp"

// special values planted at fixed positions
template <class T> T special(size_t k) {
    const T specials[] = {-std::numeric_limits<T>::quiet_NaN(),
                          std::numeric_limits<T>::infinity(),
                          -std::numeric_limits<T>::infinity(),
                          T(-0.0),
                          std::ldexp(T(1), -25),   // 2.98023223876953125e-08: a tie at the 18th digit
                          std::ldexp(T(3), -26),
                          std::numeric_limits<T>::denorm_min(),
                          std::numeric_limits<T>::max(),
                          T(1e30),
                          T(-123456.75),
                          T(0.0000005),              // rounds at the 6th decimal
                          T(-0.0000004)};
    return specials[k % (sizeof specials / sizeof specials[0])];
}

template <class T>
int run(const std::string& dir, size_t dim, size_t stride, size_t it, size_t its, unsigned threads) {
    const size_t n3 = dim * dim * dim;
    std::vector<T> rho(n3), u(3 * n3), f(19 * n3);
    std::vector<int32_t> map(n3);
    const int32_t types[] = {0x1, 0x8, 0x4, 0x102, 0x10, 0x120, 0x280, 0x200};
    for (size_t c = 0; c < n3; ++c) {
        rho[c] = (c % 7 == 3) ? std::numeric_limits<T>::quiet_NaN() : T(1) + T(c) * T(1e-7) - T(0.001) * T(c % 5);
        for (int k = 0; k < 3; ++k) u[k * n3 + c] = T(0.05) * std::sin(T(c + 11 * k)) * ((c % 3) ? T(1) : T(-1e-9));
        map[c] = types[c % 8];
    }
    for (size_t s = 0; s < 19 * n3; ++s)
        f[s] = (s % 13 == 5) ? std::numeric_limits<T>::quiet_NaN() : T(1) / T(36) + T(1e-4) * std::cos(T(s));
    {
        // in the VTI window: the first 12 window points of rho and of u_y
        const size_t w = dim - 2;
        for (size_t k = 0; k < 12 && k < w * w * w; ++k) {
            const size_t c = (1 + k % w) + dim * ((1 + (k / w) % w) + dim * (1 + k / (w * w)));
            rho[c] = special<T>(k);
            u[n3 + c] = special<T>(k + 3);
        }
        for (size_t k = 0; k < 12 && 37 * k + 50 < 19 * n3; ++k) f[37 * k + 50] = special<T>(k);
    }
    auto raw = [&](const char* name, const void* p, size_t bytes) {
        std::FILE* fh = std::fopen((dir + "/" + name).c_str(), "wb");
        if (fh) {
            std::fwrite(p, 1, bytes, fh);
            std::fclose(fh);
        }
    };
    raw("rho.bin", rho.data(), rho.size() * sizeof(T));
    raw("u.bin", u.data(), u.size() * sizeof(T));
    raw("f.bin", f.data(), f.size() * sizeof(T));
    raw("map.bin", map.data(), map.size() * sizeof(int32_t));
    lbm_io::write_map(dir + "/map.dump", map.data(), dim);
    lbm_io::write_f(lbm_io::numbered_path(dir, "f_", it, its, ".dump"), f.data(), dim, stride);
    lbm_io::write_vti(lbm_io::numbered_path(dir, "lbmcl.", it, its, ".vti"), rho.data(), u.data(), dim);

    // the same files from the snapshot layouts
    const size_t w = dim - 2, w3 = w * w * w;
    std::vector<T> macro(4 * w3), cells(19 * n3);
    for (size_t z = 0; z < w; ++z)
        for (size_t y = 0; y < w; ++y)
            for (size_t x = 0; x < w; ++x) {
                const size_t k = x + w * (y + w * z), c = (x + 1) + dim * ((y + 1) + dim * (z + 1));
                macro[k] = rho[c];
                for (size_t d = 0; d < 3; ++d) macro[w3 + 3 * k + d] = u[d * n3 + c];
            }
    for (size_t c = 0; c < n3; ++c)
        for (size_t q = 0; q < 19; ++q) cells[c * 19 + q] = f[((c / stride) * 19 + q) * stride + (c & (stride - 1))];
    lbm_out::write_map(dir + "/fast_map.dump", map.data(), dim, threads);
    lbm_out::write_f(lbm_io::numbered_path(dir, "fast_f_", it, its, ".dump"), cells.data(), dim, threads);
    lbm_out::write_vti(lbm_io::numbered_path(dir, "fast_lbmcl.", it, its, ".vti"), macro.data(), dim, threads);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 7 && argc != 8) {
        std::cerr << "usage: lbm_io_selftest <outdir> <dim> <stride> <f32|f64> <iteration> <iterations> [threads]"
                  << std::endl;
        return 2;
    }
    const unsigned threads = argc == 8 ? unsigned(std::strtoul(argv[7], nullptr, 10)) : 4u;
    const std::string dir = argv[1];
    const size_t dim = std::strtoul(argv[2], nullptr, 10), stride = std::strtoul(argv[3], nullptr, 10);
    const size_t it = std::strtoul(argv[5], nullptr, 10), its = std::strtoul(argv[6], nullptr, 10);
    if (!std::strcmp(argv[4], "f64")) return run<double>(dir, dim, stride, it, its, threads);
    return run<float>(dir, dim, stride, it, its, threads);
}
