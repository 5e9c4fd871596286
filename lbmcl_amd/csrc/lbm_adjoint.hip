// lbm_adjoint.hip -- gradients of a final-state objective with respect to the
// lid velocity U, the viscosity nu and a per-cell body force g(x), by the
// adjoint (reverse) collide-stream step with checkpointing.
//
// Schedule for T steps with checkpoints every K steps (M = ceil(T / K)
// segments):
//   forward   P_0 -> P_T; P_{mK} is kept for every segment m; the states of
//             the last segment are kept whole (the tape), so it needs no
//             recomputation.
//   seed      lambda_T = dJ/dP_T from the objective on the gathered P_T.
//   reverse   for m = M-1 .. 0: recompute P_{mK+1} .. P_{mK+len-1} from the
//             checkpoint into the tape (not for the last segment), then step
//             lambda backwards through the segment.
// Device memory: M checkpoints + (K - 1) tape states + 2 lattices + 2 adjoint
// lattices, i.e. about 2 sqrt(T) states for K = sqrt(T).
// No counterpart in the reference.
#include <type_traits>

#include "lbm_adjoint.hpp"
#include "lbm_ctx.hpp"

using namespace lbm_host;

namespace {

struct AdjointBuffers {
    std::vector<void*> owned;
    ~AdjointBuffers() {
        for (void* p : owned)
            if (p) (void)hipFree(p);
    }
    int alloc(void** p, size_t bytes, const char* what) {
        const int rc = dev_alloc(p, bytes, what);
        if (!rc) owned.push_back(*p);
        return rc;
    }
};

// Cells per lane of the vector adjoint kernel, 0 if it cannot run: it needs
// whole wavefronts per row (N / V >= 64), at most 512 threads per row, and V
// consecutive cells contiguous (stride >= V).  8-byte lanes where possible
// (FP32: V = 2), else one cell per lane.  LBM_KERNEL_GENERIC and
// LBM_KERNEL_ROW select the generic adjoint kernel.
template <class T> int64_t adjoint_vec_width(const lbm_ctx* c) {
    if (c->cfg.kernel == LBM_KERNEL_GENERIC || c->cfg.kernel == LBM_KERNEL_ROW) return 0;
    for (int64_t V : {int64_t(8 / sizeof(T)), int64_t(1)})
        if (c->n / V >= 64 && c->n / V <= 512 && c->stride >= V) return V;
    return 0;
}

template <class T>
int run_adjoint(lbm_ctx* c, const lbm_adjoint_config& cfg, lbm_adjoint_result* res, double* grad_g_host) {
    const int64_t T_steps = cfg.steps;
    int64_t K = cfg.checkpoint_every;
    if (K <= 0) K = int64_t(std::llround(std::sqrt(double(T_steps))));
    if (K < 1) K = 1;
    if (K > T_steps) K = T_steps;
    const int64_t M = (T_steps + K - 1) / K;
    const size_t S = size_t(c->n3) * lbm::Q * sizeof(T);
    AdjointBuffers buf;
    double device_bytes = 2.0 * S + double(c->n3) * (4 * sizeof(T) + 4);

    // states: ckpt[0] is the context's lattice 0, the final state goes to lattice 1
    std::vector<void*> ckpt(size_t(M), nullptr);
    ckpt[0] = c->lat[0];
    for (int64_t m = 1; m < M; ++m) {
        const int rc = buf.alloc(&ckpt[size_t(m)], S, "adjoint checkpoint");
        if (rc) return rc;
        device_bytes += double(S);
    }
    const int64_t ntape = K - 1;
    std::vector<void*> tape(size_t(ntape), nullptr);
    for (int64_t j = 0; j < ntape; ++j) {
        const int rc = buf.alloc(&tape[size_t(j)], S, "adjoint tape");
        if (rc) return rc;
        device_bytes += double(S);
    }
    void* lam[2] = {nullptr, nullptr};
    for (int j = 0; j < 2; ++j) {
        const int rc = buf.alloc(&lam[j], S, "adjoint state");
        if (rc) return rc;
        device_bytes += double(S);
    }
    // dJ/dg: FP64 sum; FP32 runs accumulate in an FP32 buffer flushed into it
    // every ADJ_FLUSH steps (half the read-modify-write traffic per step)
    double* grad_g = nullptr;
    T* grad_acc = nullptr;
    constexpr bool staged = std::is_same<T, float>::value;
    if (grad_g_host) {
        const size_t n = size_t(3 * c->n3);
        int rc = buf.alloc(reinterpret_cast<void**>(&grad_g), n * sizeof(double), "dJ/dg");
        if (rc) return rc;
        device_bytes += double(n) * sizeof(double);
        LBM_HIP_CHECK(hipMemsetAsync(grad_g, 0, n * sizeof(double), c->stream), "hipMemsetAsync");
        if (staged) {
            rc = buf.alloc(reinterpret_cast<void**>(&grad_acc), n * sizeof(T), "dJ/dg accumulator");
            if (rc) return rc;
            device_bytes += double(n) * sizeof(T);
            LBM_HIP_CHECK(hipMemsetAsync(grad_acc, 0, n * sizeof(T), c->stream), "hipMemsetAsync");
        } else {
            grad_acc = reinterpret_cast<T*>(grad_g);
        }
    }
    // vector adjoint kernel: one block of N / V threads per interior row
    const int64_t V = adjoint_vec_width<T>(c);
    const bool vec = V > 0;
    const int block = vec ? int(c->n / V) : 256;
    const int64_t nblocks = vec ? int64_t(c->n - 2) * (c->n - 2) : (c->n3 + 255) / 256;
    double* partials = nullptr;
    double* acc = nullptr;
    {
        int rc = buf.alloc(reinterpret_cast<void**>(&partials), size_t(3 * nblocks) * sizeof(double), "partials");
        if (rc) return rc;
        rc = buf.alloc(reinterpret_cast<void**>(&acc), 3 * sizeof(double), "accumulators");
        if (rc) return rc;
        LBM_HIP_CHECK(hipMemsetAsync(partials, 0, size_t(3 * nblocks) * sizeof(double), c->stream), "hipMemsetAsync");
        LBM_HIP_CHECK(hipMemsetAsync(acc, 0, 3 * sizeof(double), c->stream), "hipMemsetAsync");
        // slots that no step writes (fed by wall cells) must read as zero
        for (int j = 0; j < 2; ++j) LBM_HIP_CHECK(hipMemsetAsync(lam[j], 0, S, c->stream), "hipMemsetAsync");
    }

    hipEvent_t e0, e1, e2;
    LBM_HIP_CHECK(hipEventCreate(&e0), "hipEventCreate");
    LBM_HIP_CHECK(hipEventCreate(&e1), "hipEventCreate");
    LBM_HIP_CHECK(hipEventCreate(&e2), "hipEventCreate");
    struct EvGuard {
        hipEvent_t* e;
        ~EvGuard() {
            for (int k = 0; k < 3; ++k) (void)hipEventDestroy(e[k]);
        }
    };
    hipEvent_t evs[3] = {e0, e1, e2};
    EvGuard guard{evs};

    // P_0 in both context lattices, then its wall slots into every other state
    int rc = launch_initialize<T, lbm::ST_NATIVE>(c);
    if (rc) return rc;
    for (void* p : ckpt)
        if (p != c->lat[0]) LBM_HIP_CHECK(hipMemcpyAsync(p, c->lat[0], S, hipMemcpyDeviceToDevice, c->stream), "D2D");
    for (void* p : tape) LBM_HIP_CHECK(hipMemcpyAsync(p, c->lat[0], S, hipMemcpyDeviceToDevice, c->stream), "D2D");

    // forward with checkpoints
    LBM_HIP_CHECK(hipEventRecord(e0, c->stream), "hipEventRecord");
    void* fin = c->lat[1];
    for (int64_t m = 0; m < M; ++m) {
        const int64_t s0 = m * K;
        const int64_t len = std::min(K, T_steps - s0);
        const bool last = (m == M - 1);
        const void* src = ckpt[size_t(m)];
        for (int64_t j = 1; j <= len; ++j) {  // produce P_{s0 + j}
            void* dst;
            if (j == len)
                dst = last ? fin : ckpt[size_t(m + 1)];
            else
                dst = last ? tape[size_t(j - 1)] : tape[size_t((j - 1) % 2)];
            rc = launch_compute<T, lbm::ST_NATIVE>(c, src, dst, 0, 0);
            if (rc) return rc;
            src = dst;
        }
    }
    LBM_HIP_CHECK(hipEventRecord(e1, c->stream), "hipEventRecord");

    lbm::AdjArgs<T> a{};
    a.grad_acc = grad_acc;
    a.partials = partials;
    a.n3 = c->n3;
    a.stride = c->stride;
    a.n = c->n;
    a.log2n = c->log2n;
    a.inv_tau = T(c->inv_tau);
    a.U = T(c->U);
    a.objective = cfg.objective;
    a.plane_z = cfg.plane_z;
    a.xcd_remap = c->cfg.xcd_remap >= 0 ? 1 : 0;
    {
        // number of FLUID/MOVING cells in the plane z = plane_z (all interior cells of that plane but the
        // boundary-layer ones); counted on the host
        int64_t np = 0;
        for (int32_t y = 0; y < c->n; ++y)
            for (int32_t x = 0; x < c->n; ++x)
                if (lbm::is_store_macro(lbm::cell_type(x, y, cfg.plane_z, c->n))) ++np;
        a.plane_scale = np > 0 ? 1.0 / double(np) : 0.0;
    }
    auto adj_step = [&](const void* P, const void* Lnext, void* Lprev, int seed, int first) -> int {
        a.P = static_cast<const T*>(P);
        a.Lnext = static_cast<const T*>(Lnext);
        a.Lprev = static_cast<T*>(Lprev);
        a.seed = seed;
        a.first = first;
        if (vec && V == 2) {
            if constexpr (sizeof(T) == 4)
                hipLaunchKernelGGL((lbm::lbm_adjoint_vec<T, 2>), dim3(unsigned(nblocks)), dim3(block), 0, c->stream,
                                   a);
        } else if (vec)
            hipLaunchKernelGGL((lbm::lbm_adjoint_vec<T, 1>), dim3(unsigned(nblocks)), dim3(block), 0, c->stream, a);
        else
            hipLaunchKernelGGL((lbm::lbm_adjoint_step<T>), dim3(unsigned(nblocks)), dim3(block), 0, c->stream, a);
        LBM_HIP_CHECK(hipGetLastError(), "adjoint step launch");
        return LBM_OK;
    };

    // seed: lambda_T from the objective evaluated on P_T
    int cur = 0;
    rc = adj_step(fin, nullptr, lam[cur], 1, T_steps == 0 ? 1 : 0);
    if (rc) return rc;
    // reverse sweep
    for (int64_t m = M - 1; m >= 0; --m) {
        const int64_t s0 = m * K;
        const int64_t len = std::min(K, T_steps - s0);
        if (m != M - 1) {  // recompute P_{s0+1} .. P_{s0+len-1}
            const void* src = ckpt[size_t(m)];
            for (int64_t j = 1; j < len; ++j) {
                rc = launch_compute<T, lbm::ST_NATIVE>(c, src, tape[size_t(j - 1)], 0, 0);
                if (rc) return rc;
                src = tape[size_t(j - 1)];
            }
        }
        for (int64_t j = len - 1; j >= 0; --j) {  // step k = s0 + j reads P_k
            const void* Pk = (j == 0) ? ckpt[size_t(m)] : tape[size_t(j - 1)];
            rc = adj_step(Pk, lam[cur], lam[1 - cur], 0, (s0 + j == 0) ? 1 : 0);
            if (rc) return rc;
            cur = 1 - cur;
            if (staged && grad_acc && (s0 + j) % lbm::ADJ_FLUSH == 0) {
                const int64_t n = 3 * c->n3;
                hipLaunchKernelGGL(lbm::lbm_adjoint_flush, dim3(unsigned((n + 255) / 256)), dim3(256), 0, c->stream,
                                   reinterpret_cast<float*>(grad_acc), grad_g, n);
                LBM_HIP_CHECK(hipGetLastError(), "adjoint flush launch");
            }
        }
    }
    hipLaunchKernelGGL(lbm::lbm_adjoint_reduce, dim3(1), dim3(1024), 0, c->stream, partials, nblocks, acc);
    LBM_HIP_CHECK(hipGetLastError(), "adjoint reduce launch");
    LBM_HIP_CHECK(hipEventRecord(e2, c->stream), "hipEventRecord");

    double h[3];
    LBM_HIP_CHECK(hipMemcpyAsync(h, acc, sizeof h, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync D2H");
    if (grad_g_host)
        LBM_HIP_CHECK(hipMemcpyAsync(grad_g_host, grad_g, size_t(3 * c->n3) * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream),
                      "hipMemcpyAsync D2H");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    float fms = 0.f, ams = 0.f;
    LBM_HIP_CHECK(hipEventElapsedTime(&fms, e0, e1), "hipEventElapsedTime");
    LBM_HIP_CHECK(hipEventElapsedTime(&ams, e1, e2), "hipEventElapsedTime");
    const double om = c->inv_tau;
    res->objective = h[2];
    res->d_velocity = h[0];
    res->d_viscosity = h[1] * (-3.0 * om * om);  // omega = 1 / (3 nu + 1/2)
    res->forward_ms = double(fms);
    res->adjoint_ms = double(ams);
    res->device_bytes = device_bytes;
    res->checkpoints = M;
    res->segment = K;
    c->initialized = false;  // the lattices no longer hold a resumable state
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_set_force(lbm_ctx* c, const void* force) {
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    int rc = set_device(c);
    if (rc) return rc;
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    if (!force) {
        if (c->force) (void)hipFree(c->force);
        c->force = nullptr;
        return LBM_OK;
    }
    const size_t bytes = size_t(3 * c->n3) * c->tsize;
    if (!c->force) {
        rc = dev_alloc(&c->force, bytes, "body force");
        if (rc) return rc;
    }
    LBM_HIP_CHECK(hipMemcpyAsync(c->force, force, bytes, hipMemcpyHostToDevice, c->stream), "hipMemcpyAsync H2D");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_adjoint(lbm_ctx* c, const lbm_adjoint_config* cfg, lbm_adjoint_result* res, double* grad_force) {
    if (!cfg || !res) return fail(LBM_ERR_INVALID, "null argument");
    // the layout check comes first: it needs no device
    if (cfg->struct_size != sizeof(lbm_adjoint_config))
        return fail(LBM_ERR_INVALID, "lbm_adjoint_config.struct_size is " + std::to_string(cfg->struct_size) +
                                         ", this library's lbm_adjoint_config has " +
                                         std::to_string(sizeof(lbm_adjoint_config)) +
                                         " bytes (set it to sizeof(lbm_adjoint_config), e.g. with "
                                         "LBM_ADJOINT_CONFIG_INIT)");
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    int rc = set_device(c);
    if (rc) return rc;
    if (c->cfg.storage != LBM_STORE_NATIVE) return fail(LBM_ERR_INVALID, "the adjoint needs native storage");
    if (cfg->steps < 1) return fail(LBM_ERR_INVALID, "adjoint needs steps >= 1");
    if (cfg->objective != LBM_OBJECTIVE_KINETIC_ENERGY && cfg->objective != LBM_OBJECTIVE_MEAN_UX_PLANE)
        return fail(LBM_ERR_INVALID, "unknown objective");
    if (cfg->objective == LBM_OBJECTIVE_MEAN_UX_PLANE && (cfg->plane_z < 1 || cfg->plane_z > c->n - 2))
        return fail(LBM_ERR_INVALID, "plane_z must be an interior plane");
    if (c->n < 4) return fail(LBM_ERR_INVALID, "the adjoint needs dim >= 4");
    std::memset(res, 0, sizeof *res);
    if (c->cfg.precision == LBM_F64) return run_adjoint<double>(c, *cfg, res, grad_force);
    return run_adjoint<float>(c, *cfg, res, grad_force);
}

}  // extern "C"
