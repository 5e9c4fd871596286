/* ref_runner.c -- test infrastructure: runs the REFERENCE's own OpenCL kernels
 * (kernels.cl `initialize` and `compute`, compiled offline for the local GPU by
 * oracle/ref_build.py from the reference's source with the options that
 * lbmcl.hpp:131-156 passes to clBuildProgram) through the ROCm OpenCL runtime,
 * with the host flow of lbmcl.hpp:392-521:
 *
 *   - five buffers f_stream, f_collide, density, u, map (lbmcl.hpp:403-417);
 *   - `initialize(f_stream, f_collide, density, u, map)` over the NDRange
 *     (N, N, N) with the given work-group size (lbmcl.hpp:423-430, 493-498);
 *   - iteration it = 1..iterations runs `compute(dst, src, density, u, map,
 *     update_macro)` with dst/src = f_collide/f_stream when it is even and
 *     f_stream/f_collide when it is odd, and update_macro = (every != 0 &&
 *     it % every == 0) (lbmcl.hpp:435-467, 505-511);
 *   - every kernel has a profiling event; the kernel time is the sum of the
 *     compute events, as kernelsTimeMS (lbmcl.hpp:562-574).
 *
 * Writes <out>.rho (N^3), <out>.u (3 N^3, planar) and <out>.f (the lattice that
 * iteration iterations + 1 would read, 19 N^3 in the CSoA layout) as raw values
 * of the arithmetic type, then prints one line
 *   "ref_runner <device> dim D iterations I timed T kernels_ms K mlups M"
 * where MLUPS counts (N-2)^3 updates per timed iteration (lbmcl.hpp:606, 619)
 * and the timed iterations are those after the first `skip`.
 *
 * usage: ref_runner <code object> f32|f64 <dim> <stride> <lx,ly,lz> <iterations> <every> <skip> <out|->
 *        ref_runner --probe        (prints the OpenCL GPU device, exit 3 when there is none)
 *
 * The reference computes slot indices in `int` (kernels.cl:64-65), so the
 * runner refuses 19 N^3 >= 2^31 (N = 512) instead of launching a kernel that
 * would index out of bounds. */
#define CL_TARGET_OPENCL_VERSION 120
#define CL_USE_DEPRECATED_OPENCL_1_2_APIS
#include <CL/cl.h>

#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define Q 19

static void die(const char* what, cl_int err) {
    fprintf(stderr, "ref_runner: %s failed (%d)\n", what, (int)err);
    exit(1);
}
#define CL_CHECK(call)                       \
    do {                                     \
        cl_int e_ = (call);                  \
        if (e_ != CL_SUCCESS) die(#call, e_); \
    } while (0)

static int is_pow2(long v) { return v > 0 && (v & (v - 1)) == 0; }

/* First GPU device of any platform; 0 when there is none. */
static int find_gpu(cl_platform_id* plat, cl_device_id* dev) {
    cl_uint np = 0;
    if (clGetPlatformIDs(0, NULL, &np) != CL_SUCCESS || np == 0) return 0;
    cl_platform_id ps[16];
    if (np > 16) np = 16;
    CL_CHECK(clGetPlatformIDs(np, ps, NULL));
    for (cl_uint i = 0; i < np; ++i) {
        cl_uint nd = 0;
        if (clGetDeviceIDs(ps[i], CL_DEVICE_TYPE_GPU, 1, dev, &nd) == CL_SUCCESS && nd > 0) {
            *plat = ps[i];
            return 1;
        }
    }
    return 0;
}

static void write_file(const char* prefix, const char* ext, const void* data, size_t bytes) {
    char path[4096];
    snprintf(path, sizeof path, "%s.%s", prefix, ext);
    FILE* fh = fopen(path, "wb");
    if (!fh || fwrite(data, 1, bytes, fh) != bytes) {
        fprintf(stderr, "ref_runner: cannot write %s\n", path);
        exit(1);
    }
    fclose(fh);
}

int main(int argc, char** argv) {
    cl_platform_id plat;
    cl_device_id dev;
    char name[256] = {0};
    if (argc == 2 && strcmp(argv[1], "--probe") == 0) {
        if (!find_gpu(&plat, &dev)) {
            printf("no OpenCL GPU device\n");
            return 3;
        }
        CL_CHECK(clGetDeviceInfo(dev, CL_DEVICE_NAME, sizeof name - 1, name, NULL));
        printf("%s\n", name);
        return 0;
    }
    if (argc != 10) {
        fprintf(stderr,
                "usage: ref_runner <code object> f32|f64 <dim> <stride> <lx,ly,lz> <iterations> <every> <skip> <out|->\n"
                "       ref_runner --probe\n");
        return 2;
    }
    const char* co_path = argv[1];
    const int fp64 = strcmp(argv[2], "f64") == 0;
    if (!fp64 && strcmp(argv[2], "f32") != 0) {
        fprintf(stderr, "ref_runner: precision must be f32 or f64\n");
        return 2;
    }
    const long n = atol(argv[3]), stride = atol(argv[4]);
    size_t lws[3];
    if (sscanf(argv[5], "%zu,%zu,%zu", &lws[0], &lws[1], &lws[2]) != 3) {
        fprintf(stderr, "ref_runner: work-group size must be x,y,z\n");
        return 2;
    }
    const long iterations = atol(argv[6]), every = atol(argv[7]), skip = atol(argv[8]);
    const char* out = argv[9];
    const long n3 = n * n * n;
    /* the shapes the reference accepts (lbmcl.hpp:366-387), and its int index range */
    if (!is_pow2(n) || n < 4 || !is_pow2(stride) || stride > n3 || (long long)Q * n3 >= (1LL << 31)) {
        fprintf(stderr, "ref_runner: dim must be a power of two with 19 dim^3 < 2^31 and stride a power of two <= dim^3\n");
        return 2;
    }
    for (int k = 0; k < 3; ++k)
        if (lws[k] == 0 || (size_t)n % lws[k] != 0) {
            fprintf(stderr, "ref_runner: work-group size must divide dim\n");
            return 2;
        }
    if (lws[0] * lws[1] * lws[2] > 1024 || iterations < 1 || every < 0 || skip < 0 || skip >= iterations) {
        fprintf(stderr, "ref_runner: bad work-group size, iterations, every or skip\n");
        return 2;
    }

    if (!find_gpu(&plat, &dev)) {
        fprintf(stderr, "ref_runner: no OpenCL GPU device\n");
        return 3;
    }
    CL_CHECK(clGetDeviceInfo(dev, CL_DEVICE_NAME, sizeof name - 1, name, NULL));

    FILE* fh = fopen(co_path, "rb");
    if (!fh) {
        fprintf(stderr, "ref_runner: cannot open %s\n", co_path);
        return 1;
    }
    fseek(fh, 0, SEEK_END);
    const size_t co_bytes = (size_t)ftell(fh);
    fseek(fh, 0, SEEK_SET);
    unsigned char* co = (unsigned char*)malloc(co_bytes);
    if (!co || fread(co, 1, co_bytes, fh) != co_bytes) {
        fprintf(stderr, "ref_runner: cannot read %s\n", co_path);
        return 1;
    }
    fclose(fh);

    cl_int err;
    cl_context ctx = clCreateContext(NULL, 1, &dev, NULL, NULL, &err);
    if (err != CL_SUCCESS) die("clCreateContext", err);
    cl_command_queue queue = clCreateCommandQueue(ctx, dev, CL_QUEUE_PROFILING_ENABLE, &err);
    if (err != CL_SUCCESS) die("clCreateCommandQueue", err);
    cl_int bin_status;
    const unsigned char* bins[1] = {co};
    cl_program prog = clCreateProgramWithBinary(ctx, 1, &dev, &co_bytes, bins, &bin_status, &err);
    if (err != CL_SUCCESS || bin_status != CL_SUCCESS) die("clCreateProgramWithBinary", err ? err : bin_status);
    err = clBuildProgram(prog, 1, &dev, "", NULL, NULL);
    if (err != CL_SUCCESS) {
        char log[4096] = {0};
        clGetProgramBuildInfo(prog, dev, CL_PROGRAM_BUILD_LOG, sizeof log - 1, log, NULL);
        fprintf(stderr, "build log: %s\n", log);
        die("clBuildProgram", err);
    }

    const size_t tb = fp64 ? 8 : 4;
    const size_t f_bytes = (size_t)Q * n3 * tb, rho_bytes = n3 * tb, u_bytes = 3 * n3 * tb, map_bytes = n3 * 4;
    cl_mem f_stream = clCreateBuffer(ctx, CL_MEM_READ_WRITE, f_bytes, NULL, &err);
    if (err != CL_SUCCESS) die("clCreateBuffer(f_stream)", err);
    cl_mem f_collide = clCreateBuffer(ctx, CL_MEM_READ_WRITE, f_bytes, NULL, &err);
    if (err != CL_SUCCESS) die("clCreateBuffer(f_collide)", err);
    cl_mem rho = clCreateBuffer(ctx, CL_MEM_READ_WRITE, rho_bytes, NULL, &err);
    if (err != CL_SUCCESS) die("clCreateBuffer(rho)", err);
    cl_mem u = clCreateBuffer(ctx, CL_MEM_READ_WRITE, u_bytes, NULL, &err);
    if (err != CL_SUCCESS) die("clCreateBuffer(u)", err);
    cl_mem map = clCreateBuffer(ctx, CL_MEM_READ_WRITE, map_bytes, NULL, &err);
    if (err != CL_SUCCESS) die("clCreateBuffer(map)", err);

    cl_kernel init_k = clCreateKernel(prog, "initialize", &err);
    if (err != CL_SUCCESS) die("clCreateKernel(initialize)", err);
    cl_kernel comp_k = clCreateKernel(prog, "compute", &err);
    if (err != CL_SUCCESS) die("clCreateKernel(compute)", err);

    const size_t gws[3] = {(size_t)n, (size_t)n, (size_t)n};
    CL_CHECK(clSetKernelArg(init_k, 0, sizeof(cl_mem), &f_stream));
    CL_CHECK(clSetKernelArg(init_k, 1, sizeof(cl_mem), &f_collide));
    CL_CHECK(clSetKernelArg(init_k, 2, sizeof(cl_mem), &rho));
    CL_CHECK(clSetKernelArg(init_k, 3, sizeof(cl_mem), &u));
    CL_CHECK(clSetKernelArg(init_k, 4, sizeof(cl_mem), &map));
    CL_CHECK(clEnqueueNDRangeKernel(queue, init_k, 3, NULL, gws, lws, 0, NULL, NULL));

    cl_event* evs = (cl_event*)calloc((size_t)iterations, sizeof(cl_event));
    CL_CHECK(clSetKernelArg(comp_k, 2, sizeof(cl_mem), &rho));
    CL_CHECK(clSetKernelArg(comp_k, 3, sizeof(cl_mem), &u));
    CL_CHECK(clSetKernelArg(comp_k, 4, sizeof(cl_mem), &map));
    for (long it = 1; it <= iterations; ++it) {
        const int swap = (it % 2 == 0);
        const cl_int store = (every != 0 && it % every == 0) ? 1 : 0;
        /* the kernel arguments are captured at enqueue time */
        CL_CHECK(clSetKernelArg(comp_k, 0, sizeof(cl_mem), swap ? &f_collide : &f_stream));
        CL_CHECK(clSetKernelArg(comp_k, 1, sizeof(cl_mem), swap ? &f_stream : &f_collide));
        CL_CHECK(clSetKernelArg(comp_k, 5, sizeof(cl_int), &store));
        CL_CHECK(clEnqueueNDRangeKernel(queue, comp_k, 3, NULL, gws, lws, 0, NULL, &evs[it - 1]));
    }
    CL_CHECK(clFinish(queue));

    double kernels_ms = 0.0;
    for (long it = 1; it <= iterations; ++it) {
        cl_ulong t0 = 0, t1 = 0;
        CL_CHECK(clGetEventProfilingInfo(evs[it - 1], CL_PROFILING_COMMAND_START, sizeof t0, &t0, NULL));
        CL_CHECK(clGetEventProfilingInfo(evs[it - 1], CL_PROFILING_COMMAND_END, sizeof t1, &t1, NULL));
        if (it > skip) kernels_ms += (double)(t1 - t0) * 1e-6;
        clReleaseEvent(evs[it - 1]);
    }
    free(evs);

    if (strcmp(out, "-") != 0) {
        void* host = malloc(f_bytes);
        if (!host) die("malloc", 0);
        CL_CHECK(clEnqueueReadBuffer(queue, rho, CL_TRUE, 0, rho_bytes, host, 0, NULL, NULL));
        write_file(out, "rho", host, rho_bytes);
        CL_CHECK(clEnqueueReadBuffer(queue, u, CL_TRUE, 0, u_bytes, host, 0, NULL, NULL));
        write_file(out, "u", host, u_bytes);
        /* iteration it + 1 reads f_stream when it + 1 is even (lbmcl.hpp:518) */
        const cl_mem next_src = ((iterations + 1) % 2 == 0) ? f_stream : f_collide;
        CL_CHECK(clEnqueueReadBuffer(queue, next_src, CL_TRUE, 0, f_bytes, host, 0, NULL, NULL));
        write_file(out, "f", host, f_bytes);
        free(host);
    }

    const double wet = (double)(n - 2) * (double)(n - 2) * (double)(n - 2);
    const long timed = iterations - skip;
    printf("ref_runner %s dim %ld iterations %ld timed %ld kernels_ms %.6f mlups %.1f\n", name, n, iterations, timed,
           kernels_ms, wet * (double)timed / (kernels_ms * 1e3));

    clReleaseKernel(init_k);
    clReleaseKernel(comp_k);
    clReleaseMemObject(f_stream);
    clReleaseMemObject(f_collide);
    clReleaseMemObject(rho);
    clReleaseMemObject(u);
    clReleaseMemObject(map);
    clReleaseProgram(prog);
    clReleaseCommandQueue(queue);
    clReleaseContext(ctx);
    free(co);
    return 0;
}
