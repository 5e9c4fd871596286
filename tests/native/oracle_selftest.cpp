// oracle_selftest.cpp -- checks of the CPU oracle that run natively, built with
// and without sanitizers (tests/test_sanitizers.py builds this file with
// -fsanitize=address,undefined and runs it).
//
// It compiles the oracle's translation unit (oracle/lbm_oracle.cpp) into this
// program and checks, on small cavities over several CSoA strides:
//   * per-slot write counts of one push step (kernels.cl:420-425): slot (d, i)
//     is written exactly once if its source d - e_i is a non-wall cell, and
//     never otherwise (those are the reference's stale slots);
//   * per-slot write counts of one pull step: a non-wall cell writes its own
//     19 slots once, a wall cell none;
//   * push and pull agree bit for bit on the lattice view and on rho, u over
//     several steps (FP64), and the 16-bit codecs round-trip their own values.
// Every array access of these steps runs under the sanitizers' bounds checks.
// Exit status 0 when every check passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../oracle/lbm_oracle.cpp"

namespace {

int failures = 0;

void check(bool ok, const char* what, long long n, long long stride) {
    if (!ok) {
        std::fprintf(stderr, "FAIL %s (N=%lld stride=%lld)\n", what, n, stride);
        ++failures;
    }
}

bool inside(int64_t n, int64_t x, int64_t y, int64_t z) { return x >= 0 && y >= 0 && z >= 0 && x < n && y < n && z < n; }

bool bits_equal(double a, double b) {
    uint64_t ua, ub;
    std::memcpy(&ua, &a, 8);
    std::memcpy(&ub, &b, 8);
    return ua == ub || (std::isnan(a) && std::isnan(b));
}

void write_counts(int64_t n, int64_t stride) {
    lbmo_params p{};
    p.dim = int32_t(n);
    p.stride = stride;
    p.viscosity = 0.0089;
    p.velocity = 0.05;
    p.threads = 2;
    const int64_t n3 = n * n * n, ns = Q * n3;
    std::vector<double> fs(ns), fc(ns), rho(n3), u(3 * n3);
    std::vector<int32_t> map(n3), writes(ns, 0);
    lbmo_initialize_f64(&p, fs.data(), fc.data(), rho.data(), u.data(), map.data());
    lbmo_compute_counted_f64(&p, fs.data(), fc.data(), rho.data(), u.data(), map.data(), writes.data());
    bool ok = true;
    for (int64_t z = 0; z < n; ++z)
        for (int64_t y = 0; y < n; ++y)
            for (int64_t x = 0; x < n; ++x)
                for (int i = 0; i < Q; ++i) {
                    const int64_t sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
                    const bool fed = inside(n, sx, sy, sz) && !is_wall(cell_type(n, sx, sy, sz));
                    const int32_t w = writes[size_t(slot(x + y * n + z * n * n, i, stride))];
                    ok = ok && (w == (fed ? 1 : 0));
                }
    check(ok, "push step: every slot written once if its source is a non-wall cell, else never", n, stride);

    // pull form: each non-wall cell writes its own 19 slots
    std::vector<double> pa(ns), pb(ns);
    std::fill(writes.begin(), writes.end(), 0);
    lbmo_pull_initialize_f64(&p, pa.data(), pb.data(), rho.data(), u.data(), map.data());
    lbmo_pull_compute_counted_f64(&p, pa.data(), pb.data(), rho.data(), u.data(), writes.data());
    ok = true;
    for (int64_t c = 0; c < n3; ++c) {
        const int64_t x = c % n, y = (c / n) % n, z = c / (n * n);
        const int32_t want = is_wall(cell_type(n, x, y, z)) ? 0 : 1;
        for (int i = 0; i < Q; ++i) ok = ok && writes[size_t(slot(c, i, stride))] == want;
    }
    check(ok, "pull step: a non-wall cell writes its own 19 slots once, a wall cell none", n, stride);
}

void push_equals_pull(int64_t n, int64_t stride, int steps) {
    lbmo_params p{};
    p.dim = int32_t(n);
    p.stride = stride;
    p.viscosity = 0.0089;
    p.velocity = 0.05;
    p.threads = 2;
    const int64_t n3 = n * n * n, ns = Q * n3;
    std::vector<double> fs(ns), fc(ns), r1(n3), u1(3 * n3), pa(ns), pb(ns), r2(n3), u2(3 * n3), view(ns);
    std::vector<int32_t> map(n3);
    lbmo_initialize_f64(&p, fs.data(), fc.data(), r1.data(), u1.data(), map.data());
    lbmo_pull_initialize_f64(&p, pa.data(), pb.data(), r2.data(), u2.data(), map.data());
    double* push_in = fc.data();
    double* push_out = fs.data();
    void* pull_in = pa.data();
    void* pull_out = pb.data();
    bool ok = true;
    for (int s = 0; s < steps; ++s) {
        lbmo_compute_f64(&p, push_out, push_in, r1.data(), u1.data(), map.data(), 1, nullptr);
        lbmo_pull_compute_f64(&p, pull_in, pull_out, r2.data(), u2.data(), 1, nullptr, nullptr);
        std::swap(push_in, push_out);
        std::swap(pull_in, pull_out);
        lbmo_pull_view_f64(&p, pull_in, view.data());
        // the push lattice and the pull view agree on every slot the next step reads
        for (int64_t c = 0; c < n3; ++c) {
            const int64_t x = c % n, y = (c / n) % n, z = c / (n * n);
            if (is_wall(cell_type(n, x, y, z))) continue;
            for (int i = 0; i < Q; ++i) {
                const size_t k = size_t(slot(c, i, stride));
                ok = ok && bits_equal(push_in[k], view[k]);
            }
        }
        for (int64_t c = 0; c < n3; ++c) ok = ok && bits_equal(r1[size_t(c)], r2[size_t(c)]);
        for (int64_t c = 0; c < 3 * n3; ++c) ok = ok && bits_equal(u1[size_t(c)], u2[size_t(c)]);
    }
    check(ok, "push and pull agree bit for bit (lattice view, rho, u)", n, stride);
}

void codecs() {
    bool ok = true;
    for (uint32_t h = 0; h < 0x10000u; ++h) {
        const uint16_t v = uint16_t(h);
        const float f = f16_to_f32(v);
        if (!std::isnan(f)) ok = ok && f32_to_f16_rne(f) == v;
        const float b = bf16_to_f32(v);
        if (!std::isnan(b)) ok = ok && f32_to_bf16_rne(b) == v;
    }
    check(ok, "FP16 and BF16 codecs round-trip every non-NaN value", 0, 0);
}

}  // namespace

int main() {
    for (int64_t n : {6, 8, 16})
        for (int64_t stride : {int64_t(1), int64_t(2), int64_t(4), n, n * n * n})
            if ((stride & (stride - 1)) == 0) write_counts(n, stride);
    push_equals_pull(8, 8, 12);
    push_equals_pull(16, 4, 6);
    push_equals_pull(16, 16 * 16 * 16, 6);
    codecs();
    if (failures) {
        std::fprintf(stderr, "%d check(s) failed\n", failures);
        return 1;
    }
    std::printf("oracle self-test: all checks passed\n");
    return 0;
}
