// membw.hip -- calibration kernels: a plain device-to-device copy of known
// size, and the data movement of one D3Q19 step without its arithmetic.
//
// Used by bench.py for two things: the sustained copy bandwidth of the GPU it
// runs on (read + write bytes / kernel time, HIP events on the kernel's own
// stream), which is the practical ceiling for any streaming kernel, and as the
// calibration point of the rocprofv3 FETCH_SIZE / WRITE_SIZE counters (a copy
// of B bytes must fetch and write B bytes).  One 16-byte element per thread,
// streaming (nontemporal) loads and stores, XCD-contiguous block order: the
// fastest copy shape measured on MI355X (6.6 TB/s; plain block order 6.5,
// cached loads 6.33, four elements per thread 6.0; read-only 7.1, write-only
// 6.8).
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

typedef uint32_t U4 __attribute__((ext_vector_type(4)));

// Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous range of blocks (the same order the LBM kernels use).
__device__ __forceinline__ uint32_t xcd_contiguous(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7u, xcd = b & 7u, k = b >> 3;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

__global__ __launch_bounds__(256) void copy_kernel(const U4* __restrict__ src, U4* __restrict__ dst, int64_t n) {
    const int64_t i = int64_t(xcd_contiguous(blockIdx.x, gridDim.x)) * blockDim.x + threadIdx.x;
    if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// The D3Q19 pull step's data movement without its arithmetic: one block per
// interior row of an N^3 lattice in the CSoA layout with stride S, each lane
// two consecutive cells (8-byte lanes); 19 aligned loads from the source rows
// (y - e_y, z - e_z) of each population, then 19 aligned stores to the row's
// own slots -- the access pattern of the collide-stream kernels at full
// occupancy or, with `lds_bytes` of dynamic LDS per block, at fewer waves.
// The values are moved, not combined (no e_x shift: the shifted element moves
// within the same cache lines).
__global__ __launch_bounds__(256) void pattern_kernel(const float* __restrict__ src, float* __restrict__ dst, int32_t n,
                                                      int64_t stride) {
    extern __shared__ float pad[];
    typedef float F2 __attribute__((ext_vector_type(2)));
    const uint32_t nrow = uint32_t(n - 2) * uint32_t(n - 2);
    const uint32_t b = xcd_contiguous(blockIdx.x, gridDim.x);
    if (b >= nrow) return;
    const int64_t y = 1 + b % uint32_t(n - 2), z = 1 + b / uint32_t(n - 2);
    const int64_t r = y + int64_t(n) * z;
    const int64_t smask = stride - 1;
    const uint32_t x0 = threadIdx.x * 2u;
    const uint32_t off = 19u * x0 - 18u * (x0 & uint32_t(smask));
    const int ey[19] = {0, 0, 1, 0, -1, 0, 0, 1, 1, -1, -1, 0, 1, 0, -1, 0, 1, 0, -1};
    const int ez[19] = {0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, -1, -1, -1, 1, 1, 1, 1};
    F2 v[19];
#pragma unroll
    for (int i = 0; i < 19; ++i) {
        const int64_t c0 = int64_t(n) * (r - ey[i] - int64_t(n) * ez[i]);
        v[i] = __builtin_nontemporal_load(reinterpret_cast<const F2*>(src + 19 * c0 - 18 * (c0 & smask) + i * stride + off));
    }
    if (threadIdx.x == 0xffffffffu) pad[0] = v[0][0];  // keeps the dynamic LDS (occupancy control) allocated
    const int64_t oc0 = int64_t(n) * r;
#pragma unroll
    for (int i = 0; i < 19; ++i)
        __builtin_nontemporal_store(v[i], reinterpret_cast<F2*>(dst + 19 * oc0 - 18 * (oc0 & smask) + i * stride + off));
}

}  // namespace

extern "C" {

// The access pattern of one FP32 collide-stream step on an n^3 lattice (n = 512:
// one 256-thread block per row) with CSoA stride `stride`, `reps` times between
// two fresh lattices; ms[r] holds the duration of repetition r.  lds_bytes of
// dynamic LDS per block limit the waves per SIMD (0: none).
int membw_pattern(int device, int n, int64_t stride, int lds_bytes, int reps, double* ms) {
    if (n != 512 || stride < 2 || (stride & (stride - 1)) || reps <= 0 || !ms || lds_bytes < 0 || lds_bytes > 65536)
        return -1;
    if (hipSetDevice(device) != hipSuccess) return -2;
    const size_t bytes = size_t(19) * size_t(n) * size_t(n) * size_t(n) * sizeof(float);
    const unsigned rows = unsigned(n - 2) * unsigned(n - 2);
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = -3;
    }
    if (!rc && (hipMemsetAsync(a, 0, bytes, s) != hipSuccess || hipMemsetAsync(b, 0, bytes, s) != hipSuccess)) rc = -4;
    for (int r = 0; !rc && r < reps; ++r) {
        if (hipEventRecord(e0, s) != hipSuccess) rc = -4;
        hipLaunchKernelGGL(pattern_kernel, dim3(rows), dim3(256), size_t(lds_bytes), s,
                           static_cast<const float*>((r & 1) ? b : a), static_cast<float*>((r & 1) ? a : b), n, stride);
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess)
            rc = -4;
        float t = 0.f;
        if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = -4;
        ms[r] = double(t);
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

// Copy `bytes` (a multiple of 16, at most 2^31 * 256 * 16) `reps` times between
// two fresh device buffers; on success ms[r] holds the duration of repetition r.
int membw_copy(int device, int64_t bytes, int reps, double* ms) {
    if (bytes <= 0 || (bytes % 16) != 0 || reps <= 0 || !ms) return -1;
    const int64_t n = bytes / 16;
    const int64_t blocks = (n + 255) / 256;
    if (blocks > 0x7fffffff) return -1;
    if (hipSetDevice(device) != hipSuccess) return -2;
    void *a = nullptr, *b = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = 0;
    if (hipMalloc(&a, size_t(bytes)) != hipSuccess || hipMalloc(&b, size_t(bytes)) != hipSuccess ||
        hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = -3;
    }
    if (!rc && hipMemsetAsync(a, 1, size_t(bytes), s) != hipSuccess) rc = -4;
    for (int r = 0; !rc && r < reps; ++r) {
        if (hipEventRecord(e0, s) != hipSuccess) rc = -4;
        hipLaunchKernelGGL(copy_kernel, dim3(unsigned(blocks)), dim3(256), 0, s, static_cast<const U4*>(a),
                           static_cast<U4*>(b), n);
        if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
            hipEventSynchronize(e1) != hipSuccess)
            rc = -4;
        float t = 0.f;
        if (!rc && hipEventElapsedTime(&t, e0, e1) != hipSuccess) rc = -4;
        ms[r] = double(t);
    }
    if (s) (void)hipStreamSynchronize(s);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (s) (void)hipStreamDestroy(s);
    if (a) (void)hipFree(a);
    if (b) (void)hipFree(b);
    return rc;
}

}  // extern "C"

#ifdef MEMBW_MAIN
// Stand-alone form for counter calibration (rocprofv3 --pmc on a process that
// runs nothing else):  membw copy           -- five 4 GiB copies
//                      membw pattern LDS    -- six runs of the 512^3 FP32 step's
//                                              access pattern, LDS bytes per block
#include <cstdio>
#include <cstdlib>
#include <cstring>

int main(int argc, char** argv) {
    double ms[6] = {};
    if (argc >= 2 && !std::strcmp(argv[1], "copy")) {
        const int64_t bytes = int64_t(4) << 30;
        const int rc = membw_copy(0, bytes, 5, ms);
        if (rc) return std::fprintf(stderr, "membw_copy failed (%d)\n", rc), 1;
        for (int r = 0; r < 5; ++r) std::printf("copy 4 GiB: %.4f ms, %.1f GB/s\n", ms[r], 2.0 * bytes / (ms[r] * 1e6));
        return 0;
    }
    if (argc >= 3 && !std::strcmp(argv[1], "pattern")) {
        const int lds = std::atoi(argv[2]);
        const int rc = membw_pattern(0, 512, 4096, lds, 6, ms);
        if (rc) return std::fprintf(stderr, "membw_pattern failed (%d)\n", rc), 1;
        for (int r = 0; r < 6; ++r)
            std::printf("pattern 512^3 stride 4096 lds %d: %.4f ms, %.1f GB/s\n", lds, ms[r],
                        510.0 * 510.0 * 510.0 * 152.0 / (ms[r] * 1e6));
        return 0;
    }
    std::fprintf(stderr, "usage: membw copy | membw pattern LDS_BYTES\n");
    return 2;
}
#endif
