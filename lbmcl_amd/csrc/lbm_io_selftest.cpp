// lbm_io_selftest.cpp -- writes the three dump formats from deterministic
// synthetic arrays, plus the arrays themselves (rho.bin, u.bin, f.bin,
// map.bin), so the formats can be checked without a GPU.
//
// usage: lbm_io_selftest <outdir> <dim> <stride> <f32|f64> <iteration> <iterations>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <limits>
#include <string>
#include <vector>

#include "lbm_io.hpp"

template <class T>
int run(const std::string& dir, size_t dim, size_t stride, size_t it, size_t its) {
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
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 7) {
        std::cerr << "usage: lbm_io_selftest <outdir> <dim> <stride> <f32|f64> <iteration> <iterations>" << std::endl;
        return 2;
    }
    const std::string dir = argv[1];
    const size_t dim = std::strtoul(argv[2], nullptr, 10), stride = std::strtoul(argv[3], nullptr, 10);
    const size_t it = std::strtoul(argv[5], nullptr, 10), its = std::strtoul(argv[6], nullptr, 10);
    if (!std::strcmp(argv[4], "f64")) return run<double>(dir, dim, stride, it, its);
    return run<float>(dir, dim, stride, it, its);
}
