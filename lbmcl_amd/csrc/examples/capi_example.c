/* capi_example.c -- the lbmcl workflow through the C ABI alone (plain C99):
 * create a context, run a cavity, write a dump asynchronously, read the
 * reference's timing figures, destroy.  Built and run by the tests; a caller
 * of liblbm_hip.so from C, Go (cgo) or Rust (FFI) follows the same calls.
 *
 * usage: capi_example [dim] [iterations]          (prints one summary line) */
#include <stdio.h>
#include <stdlib.h>

#include "lbm_hip.h"

#define CHECK(call)                                                                  \
    do {                                                                             \
        int rc_ = (call);                                                            \
        if (rc_ != LBM_OK) {                                                         \
            fprintf(stderr, "%s failed (%d): %s\n", #call, rc_, lbm_last_error());   \
            return 1;                                                                \
        }                                                                            \
    } while (0)

int main(int argc, char** argv) {
    const int dim = argc > 1 ? atoi(argv[1]) : 64;
    const int iterations = argc > 2 ? atoi(argv[2]) : 100;
    lbm_config cfg = LBM_CONFIG_INIT; /* struct_size set; optimize 0 is the reference's arithmetic */
    cfg.dim = dim;
    cfg.precision = LBM_F32;
    cfg.storage = LBM_STORE_NATIVE;
    cfg.device = 0;
    cfg.stride = 64;
    cfg.viscosity = 0.0089;
    cfg.velocity = 0.05;
    cfg.kernel = LBM_KERNEL_AUTO;
    cfg.profile = 1;
    cfg.xcd_remap = 1;

    lbm_ctx* ctx = NULL;
    CHECK(lbm_create(&cfg, &ctx));
    CHECK(lbm_initialize(ctx));
    for (int64_t it = 1; it <= iterations; ++it) CHECK(lbm_compute(ctx, it, it == iterations, it == iterations));

    /* the macroscopic fields of the last iteration, without blocking the stream */
    int64_t ticket = 0, bytes = 0;
    const void* data = NULL;
    CHECK(lbm_snapshot(ctx, LBM_SNAPSHOT_MACRO, 0, &ticket));
    CHECK(lbm_snapshot_wait(ctx, ticket, &data, &bytes));
    const float* rho = (const float*)data;  /* rho of the VTI window, then (ux, uy, uz) triples */
    const int64_t w = dim - 2, centre = (w / 2) + w * ((w / 2) + w * (w / 2));
    const float rho_centre = rho[centre];
    CHECK(lbm_snapshot_release(ctx, ticket));

    double mass = 0.0, energy = 0.0, total_ms = 0.0, kernels_ms = 0.0;
    CHECK(lbm_read_diagnostics(ctx, &mass, &energy));
    CHECK(lbm_total_time_ms(ctx, &total_ms));
    CHECK(lbm_kernels_time_ms(ctx, &kernels_ms));
    char name[128];
    CHECK(lbm_get_device_name(0, name, sizeof name));
    const double wet = (double)(dim - 2) * (dim - 2) * (dim - 2);
    printf("%s dim %d iterations %d: rho(centre) %.7f mass %.6f energy %.6e kernels MLUPS %.0f (%lld snapshot bytes)\n",
           name, dim, iterations, rho_centre, mass, energy, wet * iterations / (kernels_ms * 1e3), (long long)bytes);
    CHECK(lbm_destroy(ctx));
    return 0;
}
