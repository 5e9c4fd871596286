// lbm_noalloc.hip -- lbm_compute_noalloc: one collide-stream iteration with
// diagnostics on caller-owned device buffers, allocating nothing.
//
// lbm_compute(ctx, it, update_macro, 1) keeps its per-block (mass, energy)
// partial sums in a buffer the context allocated.  This entry point runs the
// same kernels (the same selection, launch shapes and reduction, hence the
// same bits) on buffers the caller passes, and keeps the partial sums in a
// fixed device array of this library, `scratch`, of 2^19 blocks (8 MiB).
// Each device holds its own copy of the array, at its own address, for the
// life of the process; that address is looked up once per device with
// hipGetSymbolAddress and used for every launch.  Every block of a launch
// writes its two partials before the reduction reads them, so nothing needs
// clearing between calls; calls that overlap on one device share the array
// and race.
#include "lbm_ctx.hpp"

using namespace lbm_host;

namespace {

constexpr int64_t kScratchBlocks = int64_t(1) << 19;  // blocks of one launch: 2 doubles each, 8 MiB
static __device__ __align__(256) double scratch[2 * kScratchBlocks];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

// The address of `scratch` on `dev`, which must be the current device.
int scratch_on(int dev, double** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            (void)hipGetLastError();
            return fail(LBM_ERR_RUNTIME, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        }
        g_scratch[dev] = static_cast<double*>(p);
    }
    *out = g_scratch[dev];
    return LBM_OK;
}

int run(StepPlan* plan, const void* src, void* dst, int update_macro, double* diag) {
    int rc = LBM_DISPATCH(plan, launch_compute, plan, src, dst, update_macro, 1);
    if (rc) return rc;
    hipLaunchKernelGGL(lbm::lbm_reduce_partials, dim3(1), dim3(1024), 0, plan->stream, plan->partials,
                       compute_blocks(plan), diag);
    LBM_HIP_CHECK(hipGetLastError(), "diagnostics reduction launch");
    return LBM_OK;
}

}  // namespace

extern "C" int lbm_compute_noalloc(const lbm_config* cfg, const void* src, void* dst, void* rho, void* u,
                                   int32_t update_macro, double* diag, void* stream) {
    if (!cfg || !src || !dst || !diag) return fail(LBM_ERR_INVALID, "null argument");
    if (update_macro && (!rho || !u)) return fail(LBM_ERR_INVALID, "update_macro needs rho and u");
    if (src == dst) return fail(LBM_ERR_INVALID, "src and dst must be different lattices");
    StepPlan plan;
    int rc = plan_from_config(*cfg, &plan);
    if (rc) return rc;
    const int64_t nblocks = compute_blocks(&plan);
    if (nblocks > kScratchBlocks)
        return fail(LBM_ERR_INVALID, "lbm_compute_noalloc: a launch of " + std::to_string(nblocks) +
                                         " blocks exceeds the fixed scratch of " + std::to_string(kScratchBlocks) +
                                         " blocks (dim 512 fits; use lbm_create/lbm_compute for larger lattices)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) {
        (void)hipGetLastError();
        ndev = 0;
    }
    if (cfg->device < 0 || cfg->device >= ndev || cfg->device >= kMaxDevices)
        return fail(LBM_ERR_DEVICE, "no HIP device " + std::to_string(cfg->device));
    int prev = 0;
    LBM_HIP_CHECK(hipGetDevice(&prev), "hipGetDevice");
    LBM_HIP_CHECK(hipSetDevice(cfg->device), "hipSetDevice");
    rc = scratch_on(cfg->device, &plan.partials);
    if (!rc) {
        plan.stream = static_cast<hipStream_t>(stream);
        plan.rho = rho;
        plan.u = u;
        plan.force = nullptr;
        rc = run(&plan, src, dst, update_macro ? 1 : 0, diag);
    }
    const hipError_t e = hipSetDevice(prev);
    if (!rc && e != hipSuccess) return fail(LBM_ERR_RUNTIME, std::string("hipSetDevice: ") + hipGetErrorString(e));
    return rc;
}
