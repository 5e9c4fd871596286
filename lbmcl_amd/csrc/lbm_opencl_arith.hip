// lbm_opencl_arith.hip -- compute launches in the LBM_ARITH_OPENCL form.
//
// The collide-stream kernels of lbm_kernels.hpp instantiated with
// collide_opencl (the contractions clang forms for OpenCL C), in the one
// translation unit compiled with -fno-hip-fp32-correctly-rounded-divide-sqrt:
// FP32 divisions then carry OpenCL's 2.5 ulp bound and lower to the same
// instruction sequence as kernels.cl's `jx / rho` (kernels.cl:377-379) under
// ROCm's OpenCL compiler.  FP64 division is correctly rounded in both.
#include "lbm_ctx.hpp"

namespace lbm_host {

int launch_compute_opencl_f32(StepPlan* c, const void* src, void* dst, int update_macro, int diag, hipEvent_t e0,
                              hipEvent_t e1) {
    return launch_compute_impl<float, lbm::ST_NATIVE, lbm::AR_OPENCL>(c, src, dst, update_macro, diag, e0, e1);
}

int launch_compute_opencl_f64(StepPlan* c, const void* src, void* dst, int update_macro, int diag, hipEvent_t e0,
                              hipEvent_t e1) {
    return launch_compute_impl<double, lbm::ST_NATIVE, lbm::AR_OPENCL>(c, src, dst, update_macro, diag, e0, e1);
}

}  // namespace lbm_host
