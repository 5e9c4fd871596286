// lbm_ctx.hpp -- host-side internals of liblbm_hip.so shared by its
// translation units (lbm_capi.hip, lbm_adjoint.hip): the context object, error
// reporting, allocation, kernel selection and the kernel launchers.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <condition_variable>
#include <mutex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <new>
#include <utility>
#include <string>
#include <vector>

#include "../../include/lbm_hip.h"
#include "lbm_kernels.hpp"
#include "lbm_slots.hpp"

namespace lbm_host {

inline thread_local std::string g_last_error;

inline int fail(int code, const std::string& what) {
    g_last_error = what;
    return code;
}

#define LBM_HIP_CHECK(call, what)                                                                 \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess)                                                                     \
            return fail(LBM_ERR_RUNTIME, std::string(what) + ": " + hipGetErrorString(e_));       \
    } while (0)

inline bool is_pow2(int64_t v) { return v > 0 && (v & (v - 1)) == 0; }

inline int ilog2(int64_t v) {
    int k = 0;
    while ((int64_t(1) << k) < v) ++k;
    return k;
}

// The reference passes nu and U to the OpenCL compiler as -D literals printed
// by an ostream with 6 significant digits (lbmcl.hpp:140-141); in FP32 builds
// the literal is single precision (lbmcl.hpp:146).  Reproduce that rounding.
inline double kernel_literal(double value, int precision) {
    char buf[64];
    if (precision == LBM_F32) {
        std::snprintf(buf, sizeof buf, "%g", double(float(value)));
        return double(std::strtof(buf, nullptr));
    }
    std::snprintf(buf, sizeof buf, "%g", value);
    return std::strtod(buf, nullptr);
}

// Power-of-two scale of the shifted storage: 2^floor(log2(1024 / (0.05 + 2|U|))),
// clamped to [2^0, 2^24].  |f_i - w_i| stays far below 1 in the cavity, so
// scaled values stay in binary16's normal range with ample headroom.
inline double shift_scale(double U) {
    const double a = 0.05 + 2.0 * std::fabs(U);
    int k = int(std::floor(std::log2(1024.0 / a)));
    if (k < 0) k = 0;
    if (k > 24) k = 24;
    return std::ldexp(1.0, k);
}

struct EventRec {
    int32_t kind;
    hipEvent_t start;
    hipEvent_t stop;
};

// Event timing with recycled events.  Pairs are kept in enqueue order; once a
// pair is complete its duration is read and its events go back to the pool
// (creating two events per launch would otherwise dominate the host's enqueue
// cost of small lattices).  The newest pair and the first start event are
// always kept: the total span needs them.
struct EventLog {
    std::vector<hipEvent_t> pool;
    std::deque<EventRec> pending;                  // not yet read, in enqueue order
    std::vector<std::pair<int32_t, float>> done;   // (kind, ms) of the read pairs, in order
    hipEvent_t first = nullptr;                    // start of the first pair, once it was read
};

// One staging slot of the asynchronous snapshots: device buffer, pinned host
// buffer, and the event that marks the end of its device-to-host copy.  Its
// state and ticket live in the context's SlotTable.
struct SnapSlot {
    void* dev = nullptr;
    void* host = nullptr;
    size_t cap = 0;        // bytes allocated in both buffers
    size_t bytes = 0;      // bytes of the current snapshot
    hipEvent_t copied = nullptr;
    hipEvent_t packed = nullptr;
};

// What one compute launch needs: the configuration, the derived constants, the
// stream and the device buffers other than the two lattices.  A context holds
// one; lbm_compute_noalloc builds one on the stack around caller buffers.
struct StepPlan {
    lbm_config cfg{};
    int32_t n = 0;
    int32_t log2n = 0;
    int64_t n3 = 0;
    int64_t stride = 0;
    size_t tsize = 0;   // arithmetic type size
    size_t ssize = 0;   // storage element size
    hipStream_t stream = nullptr;
    void* rho = nullptr;
    void* u = nullptr;
    double* partials = nullptr;  // 2 doubles per block of a compute launch (diagnostics)
    void* force = nullptr;       // body force, 3 N^3 of the arithmetic type (null: none)
    double nu = 0, inv_tau = 0, U = 0, scale = 1;
};

}  // namespace lbm_host

struct lbm_ctx : lbm_host::StepPlan {
    void* lat[2] = {nullptr, nullptr};  // lat[0] read by odd iterations, lat[1] by even
    int32_t* map = nullptr;      // cell flags, allocated on first read (ensure_map)
    bool map_valid = false;
    int64_t partials_cap = 0;  // in blocks
    double* diag = nullptr;    // 2 doubles
    bool diag_valid = false;
    bool initialized = false;
    lbm_host::EventLog events;
    // asynchronous snapshots: staging slots and their state machine
    hipStream_t copy_stream = nullptr;
    lbm_host::SnapSlot snaps[LBM_SNAPSHOT_SLOTS];
    lbm_host::SlotTable<LBM_SNAPSHOT_SLOTS> snap_table;
};

namespace lbm_host {

inline int set_device(lbm_ctx* c) {
    LBM_HIP_CHECK(hipSetDevice(c->cfg.device), "hipSetDevice");
    return LBM_OK;
}

inline int ev_acquire(lbm_ctx* c, hipEvent_t* e) {
    if (!c->events.pool.empty()) {
        *e = c->events.pool.back();
        c->events.pool.pop_back();
        return LBM_OK;
    }
    LBM_HIP_CHECK(hipEventCreate(e), "hipEventCreate");
    return LBM_OK;
}

// An event pair for a launch that stamps it itself (launch_on).
inline int ev_create(lbm_ctx* c, int kind, EventRec& r) {
    r.kind = kind;
    r.start = r.stop = nullptr;
    if (!c->cfg.profile) return LBM_OK;
    int rc = ev_acquire(c, &r.start);
    if (!rc) rc = ev_acquire(c, &r.stop);
    return rc;
}

// An event pair around enqueued work: start recorded now, stop by ev_end.
inline int ev_begin(lbm_ctx* c, int kind, EventRec& r) {
    int rc = ev_create(c, kind, r);
    if (rc || !c->cfg.profile) return rc;
    LBM_HIP_CHECK(hipEventRecord(r.start, c->stream), "hipEventRecord");
    return LBM_OK;
}

// Read the durations of the oldest complete pairs and recycle their events.
inline int ev_harvest(lbm_ctx* c) {
    EventLog& L = c->events;
    while (L.pending.size() > 1) {
        const EventRec& e = L.pending.front();
        const hipError_t q = hipEventQuery(e.stop);
        if (q == hipErrorNotReady) {
            (void)hipGetLastError();
            break;
        }
        if (q != hipSuccess) return fail(LBM_ERR_RUNTIME, std::string("hipEventQuery: ") + hipGetErrorString(q));
        float t = 0.f;
        LBM_HIP_CHECK(hipEventElapsedTime(&t, e.start, e.stop), "hipEventElapsedTime");
        L.done.emplace_back(e.kind, t);
        if (!L.first && L.done.size() == 1)
            L.first = e.start;
        else
            L.pool.push_back(e.start);
        L.pool.push_back(e.stop);
        L.pending.pop_front();
    }
    return LBM_OK;
}

inline int ev_keep(lbm_ctx* c, const EventRec& r) {
    if (!c->cfg.profile) return LBM_OK;
    c->events.pending.push_back(r);
    return c->events.pending.size() > 64 ? ev_harvest(c) : LBM_OK;
}

inline int ev_end(lbm_ctx* c, EventRec& r) {
    if (!c->cfg.profile) return LBM_OK;
    LBM_HIP_CHECK(hipEventRecord(r.stop, c->stream), "hipEventRecord");
    return ev_keep(c, r);
}

inline void free_events(lbm_ctx* c) {
    EventLog& L = c->events;
    for (auto& e : L.pending) {
        if (e.start) (void)hipEventDestroy(e.start);
        if (e.stop) (void)hipEventDestroy(e.stop);
    }
    for (hipEvent_t e : L.pool) (void)hipEventDestroy(e);
    if (L.first) (void)hipEventDestroy(L.first);
    L.pending.clear();
    L.pool.clear();
    L.done.clear();
    L.first = nullptr;
}

inline void release(lbm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->cfg.device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
    for (SnapSlot& s : c->snaps) {
        if (s.dev) (void)hipFree(s.dev);
        if (s.host) (void)hipHostFree(s.host);
        if (s.copied) (void)hipEventDestroy(s.copied);
        if (s.packed) (void)hipEventDestroy(s.packed);
    }
    if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
    free_events(c);
    for (void* p : {c->lat[0], c->lat[1], c->rho, c->u, (void*)c->map, (void*)c->partials, (void*)c->diag, c->force})
        if (p) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

inline int dev_alloc(void** p, size_t bytes, const char* what) {
    *p = nullptr;
    const hipError_t e = hipMalloc(p, bytes);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(LBM_ERR_MEMORY, std::string("hipMalloc(") + what + ", " + std::to_string(bytes) +
                                        " bytes): " + hipGetErrorString(e));
    }
    return LBM_OK;
}

enum KernelKind { K_GENERIC, K_ROW, K_VEC, K_VEC8, K_TILE, K_LDS };

// The row-pipeline kernel (lbm_collide_stream_lds): whole rows of 128 to 1024
// cells staged in two LDS buffers, which must fit one CU's 160 KB; its direct
// 16-byte global-to-LDS loads need stride >= 16 / bytes per stored value.
inline int64_t lds_buffer_bytes(int64_t n, int64_t ssize) { return 2 * 19 * (n + 2 * (16 / ssize)) * ssize; }
inline bool lds_possible_for(int64_t n, int64_t stride, int64_t ssize) {
    return (n == 128 || n == 256 || n == 512 || n == 1024) && n * ssize >= 512 &&  // a wavefront per row at least
           stride >= 16 / ssize && lds_buffer_bytes(n, ssize) <= 160 * 1024;
}
// its persistent grid: the blocks (at most 8) that fit one CU's LDS, times the
// MI355X's 256 CUs
inline int64_t lds_grid(int64_t n, int64_t ssize) {
    const int64_t per_cu = std::min<int64_t>(8, (160 * 1024) / (lds_buffer_bytes(n, ssize) + 256));
    return 256 * std::max<int64_t>(1, per_cu);
}

// The tile kernel's shape for the stored population size.
template <class S> inline void tile_dims(int64_t* d) {
    d[0] = lbm::TileShape<S>::B;
    d[1] = lbm::TileShape<S>::TY;
    d[2] = lbm::TileShape<S>::TZ;
}
inline int64_t tile_dim(const StepPlan* c, int k) {
    int64_t d[3];
    if (c->ssize == 8) tile_dims<double>(d);
    else if (c->ssize == 4) tile_dims<float>(d);
    else tile_dims<uint16_t>(d);
    return d[k];
}
inline int64_t tile_b(const StepPlan* c) { return tile_dim(c, 0); }
inline int64_t tile_ty(const StepPlan* c) { return tile_dim(c, 1); }
inline int64_t tile_tz(const StepPlan* c) { return tile_dim(c, 2); }
// The tile kernel needs the x extent to be whole stride blocks (stride <= B)
// and the lattice to hold whole tiles.
inline bool tile_possible(const StepPlan* c) {
    return c->stride <= tile_b(c) && c->n >= tile_b(c) && c->n >= tile_ty(c) && c->n >= tile_tz(c);
}
// AUTO picks it where a stride block of one population is at most 32 bytes:
// measured on MI355X, there it beats the vector kernels, above it loses.
inline bool tile_preferred(const StepPlan* c) { return c->stride * int64_t(c->ssize) <= 32 && tile_possible(c); }

// bytes per lane and population of the vector kernels
inline int64_t lane_bytes(KernelKind k) { return k == K_VEC8 ? 8 : 16; }

inline bool vec_possible(const StepPlan* c, KernelKind k) {
    const int64_t v = lane_bytes(k) / int64_t(c->ssize);  // cells per lane
    return v >= 1 && c->stride >= v && c->n / v >= 64;
}

// Preferred vector width (measured on MI355X at 512^3, stride 4096): 8-byte
// lanes for every storage type.  16-byte lanes need too many registers for
// four waves per SIMD in FP32 and 16-bit storage, and in FP64 (two cells per
// lane) they were 1-2.5% slower than one cell per lane once the wavefront ends
// came through LDS.
inline KernelKind preferred_vec(const StepPlan*) { return K_VEC8; }

inline KernelKind kernel_kind(const StepPlan* c) {
    switch (c->cfg.kernel) {
        case LBM_KERNEL_GENERIC: return K_GENERIC;
        case LBM_KERNEL_ROW: return K_ROW;
        case LBM_KERNEL_VEC: return K_VEC;
        case LBM_KERNEL_VEC8: return K_VEC8;
        case LBM_KERNEL_TILE: return K_TILE;
        case LBM_KERNEL_LDS: return K_LDS;
        default: {
            if (tile_preferred(c)) return K_TILE;
            const KernelKind p = preferred_vec(c);
            if (vec_possible(c, p)) return p;
            if (vec_possible(c, K_VEC8)) return K_VEC8;
            // rows too short for whole wavefronts of vectors (e.g. 16-bit
            // storage below 256^3): the tile kernel when the stride allows it
            if (tile_possible(c)) return K_TILE;
            return c->n >= 64 ? K_ROW : K_GENERIC;
        }
    }
}

// cells per lane of the selected kernel
inline int64_t vec_width(const StepPlan* c) {
    const KernelKind k = kernel_kind(c);
    return (k == K_VEC || k == K_VEC8) ? lane_bytes(k) / int64_t(c->ssize) : 1;
}

// threads per block and blocks per row of the row kernels
inline int64_t row_block(const StepPlan* c) {
    const int64_t per_row = c->n / vec_width(c);
    return per_row < 256 ? per_row : 256;
}
inline int64_t row_gx(const StepPlan* c) { return (c->n / vec_width(c)) / row_block(c); }
// Store policy of the 8-byte-lane vector kernel: lattices up to 384 MiB are stored through
// the caches, so that much of what a step writes is still in the MI355X's
// 256 MB memory-attached last-level cache when the next step reads it
// (measured at 128^3: +13% FP32, +15% FP64); larger ones with streaming
// stores (cached stores cost 6% at 512^3).
inline bool cached_stores(const StepPlan* c) {
    return double(c->n3) * lbm::Q * double(c->ssize) <= 384.0 * 1024 * 1024;
}

// Rows per block of the vector kernels: rows shorter than 256 threads share a
// block (at most LBM_VEC_ROWS_PER_BLOCK of them), so short-row lattices launch
// fewer, larger workgroups.
#ifndef LBM_VEC_ROWS_PER_BLOCK
#define LBM_VEC_ROWS_PER_BLOCK 4
#endif
inline int64_t vec_rows_per_block(const StepPlan* c) {
    int64_t r = 256 / row_block(c);
    if (r > LBM_VEC_ROWS_PER_BLOCK) r = LBM_VEC_ROWS_PER_BLOCK;
    return r < 1 ? 1 : r;
}

inline int generic_block(const StepPlan* c) {
    int64_t b = 1;
    for (int d = 0; d < 3; ++d) b *= (c->cfg.lws[d] > 0 ? c->cfg.lws[d] : 1);
    if (c->cfg.lws[0] <= 0 && c->cfg.lws[1] <= 0 && c->cfg.lws[2] <= 0) b = 256;
    return int(b);
}

// Number of blocks of one compute launch (for the diagnostics partials).
inline int64_t compute_blocks(const StepPlan* c) {
    const KernelKind k = kernel_kind(c);
    if (k == K_TILE) return (c->n / tile_b(c)) * (c->n / tile_ty(c)) * (c->n / tile_tz(c));
    if (k == K_LDS) return lds_grid(c->n, int64_t(c->ssize));
    if (k == K_VEC || k == K_VEC8) return row_gx(c) * int64_t(c->n) * int64_t(c->n) / vec_rows_per_block(c);  // all rows
    if (k == K_ROW) return row_gx(c) * int64_t(c->n - 2) * int64_t(c->n - 2);         // interior rows
    const int64_t b = generic_block(c);
    return (c->n3 + b - 1) / b;
}

// Arithmetic form of the collision: the deviation form for 16-bit storage
// always; for native storage the fused form with -o (LBM_ARITH_FUSED), the
// default (LBM_ARITH_OPENCL) evaluating kernels.cl's expressions the way ROCm's
// OpenCL compiler does, or LBM_ARITH_EXACT, the CPU oracle's form.
inline int arith_of(const StepPlan* c) {
    if (c->cfg.storage != LBM_STORE_NATIVE) return lbm::AR_DELTA;
    if (c->cfg.optimize == LBM_ARITH_FUSED) return lbm::AR_FUSED;
    if (c->cfg.optimize == LBM_ARITH_EXACT) return lbm::AR_EXACT;
    return lbm::AR_OPENCL;
}

// Check a configuration and derive the step constants into `p` (no device
// access): the validation of lbm_create, shared with lbm_compute_noalloc.
inline int plan_from_config(const lbm_config& k, StepPlan* p) {
    if (k.struct_size != sizeof(lbm_config))
        return fail(LBM_ERR_INVALID, "lbm_config.struct_size is " + std::to_string(k.struct_size) +
                                         ", this library's lbm_config has " + std::to_string(sizeof(lbm_config)) +
                                         " bytes (set it to sizeof(lbm_config), e.g. with LBM_CONFIG_INIT)");
    if (!is_pow2(k.dim)) return fail(LBM_ERR_INVALID, "dim must be a positive power of two");
    if (k.dim > 1024) return fail(LBM_ERR_INVALID, "dim above 1024 is not supported");
    const int64_t n3 = int64_t(k.dim) * k.dim * k.dim;
    if (!is_pow2(k.stride) || k.stride > n3)
        return fail(LBM_ERR_INVALID, "stride must be a power of two in [1, dim^3]");
    if (k.precision != LBM_F32 && k.precision != LBM_F64) return fail(LBM_ERR_INVALID, "unknown precision");
    if (k.storage != LBM_STORE_NATIVE && k.storage != LBM_STORE_FP16_SHIFT && k.storage != LBM_STORE_BF16_SHIFT)
        return fail(LBM_ERR_INVALID, "unknown storage");
    if (k.storage != LBM_STORE_NATIVE && k.precision != LBM_F32)
        return fail(LBM_ERR_INVALID, "shifted 16-bit storage computes in FP32");
    if (k.optimize != LBM_ARITH_REFERENCE && k.optimize != LBM_ARITH_FUSED && k.optimize != LBM_ARITH_EXACT)
        return fail(LBM_ERR_INVALID,
                    "optimize must be LBM_ARITH_REFERENCE (LBM_ARITH_OPENCL), LBM_ARITH_FUSED or LBM_ARITH_EXACT");
    if (k.optimize == LBM_ARITH_EXACT && k.storage != LBM_STORE_NATIVE)
        return fail(LBM_ERR_INVALID, "LBM_ARITH_EXACT needs native storage (16-bit storage uses the deviation form)");
    if (k.kernel != LBM_KERNEL_AUTO && k.kernel != LBM_KERNEL_GENERIC && k.kernel != LBM_KERNEL_ROW &&
        k.kernel != LBM_KERNEL_VEC && k.kernel != LBM_KERNEL_VEC8 && k.kernel != LBM_KERNEL_TILE &&
        k.kernel != LBM_KERNEL_LDS)
        return fail(LBM_ERR_INVALID, "unknown kernel selection");
    if (k.kernel == LBM_KERNEL_LDS) {
        const int64_t ss = (k.storage != LBM_STORE_NATIVE) ? 2 : (k.precision == LBM_F64 ? 8 : 4);
        if (!lds_possible_for(k.dim, k.stride, ss))
            return fail(LBM_ERR_INVALID, "the row-pipeline kernel needs dim 128, 256, 512 or 1024 (two rows of every "
                                         "population within 160 KB of LDS) and stride >= 16 / bytes per stored value");
    }
    if (k.kernel == LBM_KERNEL_TILE) {
        int64_t d[3];
        if (k.storage != LBM_STORE_NATIVE) tile_dims<uint16_t>(d);
        else if (k.precision == LBM_F64) tile_dims<double>(d);
        else tile_dims<float>(d);
        if (k.stride > d[0] || k.dim < d[0] || k.dim < d[1] || k.dim < d[2])
            return fail(LBM_ERR_INVALID, "the tile kernel needs stride <= its x extent (" + std::to_string(d[0]) +
                                             " cells) and dim >= that extent");
    }
    if (k.kernel == LBM_KERNEL_ROW && k.dim < 64) return fail(LBM_ERR_INVALID, "the row kernel needs dim >= 64");
    if (k.kernel == LBM_KERNEL_VEC || k.kernel == LBM_KERNEL_VEC8) {
        const int64_t lb = (k.kernel == LBM_KERNEL_VEC) ? 16 : 8;
        const int64_t v = lb / ((k.storage != LBM_STORE_NATIVE) ? 2 : (k.precision == LBM_F64 ? 8 : 4));
        if (k.stride < v || k.dim / v < 64)
            return fail(LBM_ERR_INVALID, "the vector kernel needs stride >= V and dim / V >= 64 (V = lane bytes / "
                                         "bytes per stored population)");
    }
    if (!(k.viscosity >= 0.0) || !std::isfinite(k.viscosity)) return fail(LBM_ERR_INVALID, "invalid viscosity");
    if (!std::isfinite(k.velocity)) return fail(LBM_ERR_INVALID, "invalid velocity");
    int64_t lws = 1;
    for (int d = 0; d < 3; ++d) {
        if (k.lws[d] < 0) return fail(LBM_ERR_INVALID, "negative work-group size");
        lws *= (k.lws[d] > 0 ? k.lws[d] : 1);
    }
    if (lws > 1024) return fail(LBM_ERR_INVALID, "work-group size above 1024");

    p->cfg = k;
    p->n = k.dim;
    p->log2n = ilog2(k.dim);
    p->n3 = n3;
    p->stride = k.stride;
    p->tsize = (k.precision == LBM_F64) ? 8 : 4;
    p->ssize = (k.storage == LBM_STORE_NATIVE) ? p->tsize : 2;
    p->nu = k.raw_parameters ? (k.precision == LBM_F32 ? double(float(k.viscosity)) : k.viscosity)
                             : kernel_literal(k.viscosity, k.precision);
    p->U = k.raw_parameters ? (k.precision == LBM_F32 ? double(float(k.velocity)) : k.velocity)
                            : kernel_literal(k.velocity, k.precision);
    if (k.precision == LBM_F32) {
        const float nu = float(p->nu);
        const float tau = (3.0f * nu) + 0.5f;  // kernels.cl:61, folded in float
        p->inv_tau = double(1.0f / tau);       // kernels.cl:62
    } else {
        const double tau = (3.0 * p->nu) + 0.5;
        p->inv_tau = 1.0 / tau;
    }
    p->scale = (k.storage == LBM_STORE_NATIVE) ? 1.0 : shift_scale(p->U);
    return LBM_OK;
}

// Compute launches in the OpenCL arithmetic form, defined in
// lbm_opencl_arith.hip (built with OpenCL's FP32 division precision).
int launch_compute_opencl_f32(StepPlan* c, const void* src, void* dst, int update_macro, int diag, hipEvent_t e0,
                              hipEvent_t e1);
int launch_compute_opencl_f64(StepPlan* c, const void* src, void* dst, int update_macro, int diag, hipEvent_t e0,
                              hipEvent_t e1);

// Launch on the context's stream; with events, through hipExtLaunchKernel,
// which stamps the kernel's own start and end into them (no separate marker
// packets between launches, like the reference's profiling-queue events).
template <class K, class A>
inline void launch_on(StepPlan* c, K kernel, dim3 grid, dim3 block, hipEvent_t start, hipEvent_t stop, const A& a,
                      size_t shmem = 0) {
    if (start)
        hipExtLaunchKernelGGL(kernel, grid, block, shmem, c->stream, start, stop, 0, a);
    else
        hipLaunchKernelGGL(kernel, grid, block, shmem, c->stream, a);
}

// Occupancy cap of the vector kernel on lattices stored through the caches
// (128^3): a few launch rounds of blocks per CU, where fewer, fuller rounds
// beat the register-limited maximum -- measured at 128^3, FP32 4 waves per
// SIMD instead of 5 (+1.2%), FP64 5 instead of 6 (+1.3%).  The cap is unused
// dynamic LDS: 160 KiB / B - 2 KiB per block admits B blocks per CU, not B + 1,
// for kernels with at most 2 KiB of static LDS (the vector kernels use
// 0.6-1.5 KiB).
inline size_t small_lattice_lds_pad(const StepPlan* c, unsigned block_threads) {
    const unsigned waves_per_simd = c->tsize == 4 ? 4u : 5u;
    const unsigned blocks_per_cu = waves_per_simd * 4u / (block_threads / 64u);
    if (blocks_per_cu == 0) return 0;
    return size_t(160 * 1024) / blocks_per_cu - 2048;
}

template <class T, int KIND, int ARITH>
inline int launch_compute_impl(StepPlan* c, const void* src, void* dst, int update_macro, int diag, hipEvent_t e0,
                               hipEvent_t e1) {
    lbm::StepArgs<T> a{};
    a.src = src;
    a.dst = dst;
    a.rho = static_cast<T*>(c->rho);
    a.u = static_cast<T*>(c->u);
    a.partials = c->partials;
    a.n3 = c->n3;
    a.stride = c->stride;
    a.n = c->n;
    a.log2n = c->log2n;
    a.update_macro = update_macro;
    a.diagnostics = diag;
    a.xcd_remap = c->cfg.xcd_remap >= 0 ? 1 : 0;
    a.k.inv_tau = T(c->inv_tau);
    a.k.U = T(c->U);
    a.k.om1 = T(1) - T(c->inv_tau);
    a.force = static_cast<const T*>(c->force);
    a.scale = float(c->scale);
    a.inv_scale = float(1.0 / c->scale);
    a.k.dscale = a.scale;
    a.k.dinv = a.inv_scale;
    {
        // lid-layer initial populations, same expression as the initialize kernel
        const T U = T(c->U);
        if constexpr (ARITH == lbm::AR_DELTA) {
            float d[lbm::Q];
            lbm::delta_lid_initial(float(c->U), d);
            for (int i = 0; i < lbm::Q; ++i) a.lid_eq[i] = T(d[i] * float(c->scale));  // scaled, as the kernels hold them
        } else if constexpr (ARITH == lbm::AR_OPENCL) {
            const T u2 = lbm::u2_contracted<T>(U, T(0), T(0));
            for (int i = 0; i < lbm::Q; ++i) a.lid_eq[i] = lbm::feq_contracted<T>(i, T(1), U, T(0), T(0), u2);
        } else {
            const T u2 = ((U * U) + (T(0) * T(0))) + (T(0) * T(0));
            for (int i = 0; i < lbm::Q; ++i) a.lid_eq[i] = lbm::feq<T>(i, T(1), U, T(0), T(0), u2);
        }
    }
    const KernelKind kk = kernel_kind(c);
    if (kk == K_LDS) {
        const dim3 grid(unsigned(lds_grid(c->n, int64_t(c->ssize))));
        using S = typename lbm::Codec<T, KIND>::S;
        if (c->n == 128) {
            if constexpr (lbm::LdsRow<S, 128>::WAVES >= 1)
                launch_on(c, lbm::lbm_collide_stream_lds<T, KIND, ARITH, 128>, grid,
                          dim3(unsigned(lbm::LdsRow<S, 128>::THREADS)), e0, e1, a);
        } else if (c->n == 256)
            launch_on(c, lbm::lbm_collide_stream_lds<T, KIND, ARITH, 256>, grid,
                      dim3(unsigned(lbm::LdsRow<S, 256>::THREADS)), e0, e1, a);
        else if (c->n == 512)
            launch_on(c, lbm::lbm_collide_stream_lds<T, KIND, ARITH, 512>, grid,
                      dim3(unsigned(lbm::LdsRow<S, 512>::THREADS)), e0, e1, a);
        else if constexpr (2 * 19 * (1024 + 2 * (16 / sizeof(S))) * sizeof(S) <= 160 * 1024)
            launch_on(c, lbm::lbm_collide_stream_lds<T, KIND, ARITH, 1024>, grid,
                      dim3(unsigned(lbm::LdsRow<S, 1024>::THREADS)), e0, e1, a);
    } else if (kk == K_TILE) {
        const dim3 grid(unsigned(c->n / tile_b(c)), unsigned(c->n / tile_ty(c)), unsigned(c->n / tile_tz(c)));
        const dim3 block(unsigned(tile_b(c) * tile_ty(c) * tile_tz(c)));
        launch_on(c, lbm::lbm_collide_stream_tile<T, KIND, ARITH>, grid, block, e0, e1, a);
    } else if (kk != K_GENERIC) {
        a.gx = int32_t(row_gx(c));
        a.log2gx = ilog2(a.gx);
        const int64_t nb = compute_blocks(c);
        const unsigned bx = unsigned(row_block(c));
        if (kk == K_VEC || kk == K_VEC8) {
            const int64_t rpb = vec_rows_per_block(c);
            a.log2tpr = ilog2(int64_t(bx));
            a.log2rpb = ilog2(rpb);
            const dim3 block(unsigned(int64_t(bx) * rpb));
            if (kk == K_VEC)
                launch_on(c, lbm::lbm_collide_stream_vec<T, KIND, 16, ARITH>, dim3(unsigned(nb)), block, e0, e1, a);
            else if (cached_stores(c))
                launch_on(c, lbm::lbm_collide_stream_vec<T, KIND, 8, ARITH, true>, dim3(unsigned(nb)), block, e0, e1, a,
                          small_lattice_lds_pad(c, block.x));
            else
                launch_on(c, lbm::lbm_collide_stream_vec<T, KIND, 8, ARITH, false>, dim3(unsigned(nb)), block, e0, e1, a);
        } else {
            launch_on(c, lbm::lbm_collide_stream_row<T, KIND, ARITH>, dim3(unsigned(nb)), dim3(bx), e0, e1, a);
        }
    } else {
        const int b = generic_block(c);
        const int64_t nb = (c->n3 + b - 1) / b;
        a.gx = 0;
        launch_on(c, lbm::lbm_collide_stream_generic<T, KIND, ARITH>, dim3(unsigned(nb)), dim3(b), e0, e1, a);
    }
    LBM_HIP_CHECK(hipGetLastError(), "compute launch");
    return LBM_OK;
}

// e0/e1: events stamped with the kernel's start and end (null: none)
template <class T, int KIND>
inline int launch_compute(StepPlan* c, const void* src, void* dst, int update_macro, int diag, hipEvent_t e0 = nullptr,
                          hipEvent_t e1 = nullptr) {
    if constexpr (KIND != lbm::ST_NATIVE) {
        return launch_compute_impl<T, KIND, lbm::AR_DELTA>(c, src, dst, update_macro, diag, e0, e1);
    } else {
        const int ar = arith_of(c);
        if (ar == lbm::AR_FUSED)
            return launch_compute_impl<T, KIND, lbm::AR_FUSED>(c, src, dst, update_macro, diag, e0, e1);
        if (ar == lbm::AR_OPENCL) {
            if constexpr (sizeof(T) == 4) return launch_compute_opencl_f32(c, src, dst, update_macro, diag, e0, e1);
            else return launch_compute_opencl_f64(c, src, dst, update_macro, diag, e0, e1);
        }
        return launch_compute_impl<T, KIND, lbm::AR_EXACT>(c, src, dst, update_macro, diag, e0, e1);
    }
}

template <class T, int KIND>
inline int launch_initialize(lbm_ctx* c) {
    lbm::InitArgs<T> a{};
    a.a = c->lat[0];
    a.b = c->lat[1];
    a.rho = static_cast<T*>(c->rho);
    a.u = static_cast<T*>(c->u);
    a.n3 = c->n3;
    a.stride = c->stride;
    a.n = c->n;
    a.log2n = c->log2n;
    a.U = T(c->U);
    a.scale = float(c->scale);
    a.inv_scale = float(1.0 / c->scale);
    a.contracted = arith_of(c) == lbm::AR_OPENCL;
    if constexpr (KIND != lbm::ST_NATIVE) {
        float d[lbm::Q];
        lbm::delta_lid_initial(float(c->U), d);
        for (int i = 0; i < lbm::Q; ++i) a.lid_delta[i] = T(d[i]);
    }
    const int64_t nb = (c->n3 + 255) / 256;
    hipLaunchKernelGGL((lbm::lbm_initialize_kernel<T, KIND>), dim3(unsigned(nb)), dim3(256), 0, c->stream, a);
    LBM_HIP_CHECK(hipGetLastError(), "initialize launch");
    return LBM_OK;
}

template <class T, int KIND>
inline int launch_view(lbm_ctx* c, const void* p, T* out, int64_t out_stride) {
    lbm::ViewArgs<T> a{};
    a.p = p;
    a.out = out;
    a.n3 = c->n3;
    a.stride = c->stride;
    a.out_stride = out_stride;
    a.n = c->n;
    a.log2n = c->log2n;
    a.scale = float(c->scale);
    a.inv_scale = float(1.0 / c->scale);
    const int64_t nb = (c->n3 + 255) / 256;
    hipLaunchKernelGGL((lbm::lbm_view_kernel<T, KIND>), dim3(unsigned(nb)), dim3(256), 0, c->stream, a);
    LBM_HIP_CHECK(hipGetLastError(), "view launch");
    return LBM_OK;
}

// Dispatch on (precision, storage).
#define LBM_DISPATCH(c, FN, ...)                                                        \
    ((c)->cfg.precision == LBM_F64 ? FN<double, lbm::ST_NATIVE>(__VA_ARGS__)            \
     : (c)->cfg.storage == LBM_STORE_FP16_SHIFT ? FN<float, lbm::ST_FP16>(__VA_ARGS__)  \
     : (c)->cfg.storage == LBM_STORE_BF16_SHIFT ? FN<float, lbm::ST_BF16>(__VA_ARGS__)  \
                                                : FN<float, lbm::ST_NATIVE>(__VA_ARGS__))

inline int read_device(lbm_ctx* c, int kind, void* dst, const void* src, size_t bytes) {
    EventRec r;
    int rc = ev_begin(c, kind, r);
    if (rc) return rc;
    LBM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync D2H");
    rc = ev_end(c, r);
    if (rc) return rc;
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return LBM_OK;
}

// The map buffer, allocated and filled on first use (lbm_read_map, map
// snapshots); the initialize and compute kernels do not need it.
inline int ensure_map(lbm_ctx* c) {
    if (c->map_valid) return LBM_OK;
    if (!c->map) {
        const int rc = dev_alloc(reinterpret_cast<void**>(&c->map), size_t(c->n3) * 4, "map");
        if (rc) return rc;
    }
    const int64_t nb = (c->n3 + 255) / 256;
    hipLaunchKernelGGL(lbm::lbm_map_kernel, dim3(unsigned(nb)), dim3(256), 0, c->stream, c->map, c->n3, c->n, c->log2n);
    LBM_HIP_CHECK(hipGetLastError(), "map launch");
    c->map_valid = true;
    return LBM_OK;
}

template <class T>
inline int launch_pack_macro(lbm_ctx* c, T* out) {
    lbm::PackArgs<T> a{};
    a.rho = static_cast<const T*>(c->rho);
    a.u = static_cast<const T*>(c->u);
    a.out = out;
    a.n = c->n;
    a.n3 = c->n3;
    a.w = c->n - 2;
    a.w3 = int64_t(a.w) * a.w * a.w;
    const int64_t nb = (a.w3 + 255) / 256;
    hipLaunchKernelGGL((lbm::lbm_pack_macro_kernel<T>), dim3(unsigned(nb)), dim3(256), 0, c->stream, a);
    LBM_HIP_CHECK(hipGetLastError(), "pack launch");
    return LBM_OK;
}

#define LBM_CHECK_CTX(c)                                                   \
    do {                                                                   \
        if (!(c)) return fail(LBM_ERR_INVALID, "null context");            \
        int rc_ = set_device(c);                                           \
        if (rc_) return rc_;                                               \
    } while (0)

#define LBM_CHECK_INIT(c)                                                                \
    do {                                                                                 \
        if (!(c)->initialized) return fail(LBM_ERR_STATE, "lbm_initialize not called");  \
    } while (0)

}  // namespace lbm_host
