// membw.hip -- a plain device-to-device copy kernel of known size.
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

}  // namespace

extern "C" {

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
