// lbm_capi.hip -- implementation of the C ABI declared in include/lbm_hip.h.
//
// Host side of liblbm_hip.so: device selection, buffers, the ping-pong
// schedule of the compute kernel, blocking reads and event timing.  It takes
// the place of the OpenCL device layer of LBMCL (lbmcl.hpp:392-593,
// libs/CLUtil.hpp).
#include "lbm_ctx.hpp"

using namespace lbm_host;

namespace {

// The push view of lattice `buffer` into device memory `out` with CSoA stride
// out_stride, on the context's stream.
int enqueue_view(lbm_ctx* c, int32_t buffer, void* out, int64_t out_stride) {
    if (c->cfg.precision == LBM_F64)
        return launch_view<double, lbm::ST_NATIVE>(c, c->lat[buffer], static_cast<double*>(out), out_stride);
    if (c->cfg.storage == LBM_STORE_FP16_SHIFT)
        return launch_view<float, lbm::ST_FP16>(c, c->lat[buffer], static_cast<float*>(out), out_stride);
    if (c->cfg.storage == LBM_STORE_BF16_SHIFT)
        return launch_view<float, lbm::ST_BF16>(c, c->lat[buffer], static_cast<float*>(out), out_stride);
    return launch_view<float, lbm::ST_NATIVE>(c, c->lat[buffer], static_cast<float*>(out), out_stride);
}

}  // namespace

extern "C" {

const char* lbm_last_error(void) { return g_last_error.c_str(); }

int lbm_get_device_count(int32_t* count) {
    if (!count) return fail(LBM_ERR_INVALID, "null count");
    int n = 0;
    const hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *count = 0;
        if (e == hipErrorNoDevice) return LBM_OK;
        return fail(LBM_ERR_RUNTIME, std::string("hipGetDeviceCount: ") + hipGetErrorString(e));
    }
    *count = n;
    return LBM_OK;
}

int lbm_get_device_name(int32_t device, char* name, size_t len) {
    if (!name || len == 0) return fail(LBM_ERR_INVALID, "null name buffer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n) {
        (void)hipGetLastError();
        return fail(LBM_ERR_DEVICE, "no HIP device " + std::to_string(device));
    }
    hipDeviceProp_t prop;
    LBM_HIP_CHECK(hipGetDeviceProperties(&prop, device), "hipGetDeviceProperties");
    // the architecture name, which is what ROCm's OpenCL reports as
    // CL_DEVICE_NAME for the same GPU (the marketing name can be empty when
    // the system's GPU id table is missing)
    std::snprintf(name, len, "%s", prop.gcnArchName);
    return LBM_OK;
}

int lbm_create(const lbm_config* cfg, lbm_ctx** out) {
    if (!cfg || !out) return fail(LBM_ERR_INVALID, "null argument");
    *out = nullptr;
    const lbm_config& k = *cfg;
    StepPlan plan;
    const int prc = plan_from_config(k, &plan);
    if (prc) return prc;
    const int64_t n3 = plan.n3;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) {
        (void)hipGetLastError();
        ndev = 0;
    }
    if (k.device < 0 || k.device >= ndev)
        return fail(LBM_ERR_DEVICE, "no HIP device " + std::to_string(k.device) + " (" + std::to_string(ndev) +
                                        " present)");

    lbm_ctx* c = new (std::nothrow) lbm_ctx();
    if (!c) return fail(LBM_ERR_MEMORY, "host allocation failed");
    static_cast<StepPlan&>(*c) = plan;

    int rc = set_device(c);
    if (rc) {
        delete c;
        return rc;
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        c->stream = nullptr;
        release(c);
        return fail(LBM_ERR_RUNTIME, "hipStreamCreate failed");
    }
    const size_t lat_bytes = size_t(n3) * lbm::Q * c->ssize;
    c->partials_cap = compute_blocks(c);
    if ((rc = dev_alloc(&c->lat[0], lat_bytes, "lattice 0")) || (rc = dev_alloc(&c->lat[1], lat_bytes, "lattice 1")) ||
        (rc = dev_alloc(&c->rho, size_t(n3) * c->tsize, "rho")) ||
        (rc = dev_alloc(&c->u, size_t(3 * n3) * c->tsize, "u")) ||
        (rc = dev_alloc(reinterpret_cast<void**>(&c->partials), size_t(c->partials_cap) * 2 * sizeof(double),
                        "partials")) ||
        (rc = dev_alloc(reinterpret_cast<void**>(&c->diag), 2 * sizeof(double), "diagnostics"))) {
        std::string msg = g_last_error;
        release(c);
        return fail(rc, msg);
    }
    *out = c;
    return LBM_OK;
}

int lbm_destroy(lbm_ctx* ctx) {
    if (!ctx) return fail(LBM_ERR_INVALID, "null context");
    release(ctx);
    return LBM_OK;
}

int lbm_initialize(lbm_ctx* c) {
    LBM_CHECK_CTX(c);
    EventRec r;
    int rc = ev_begin(c, LBM_EVENT_INITIALIZE, r);
    if (rc) return rc;
    rc = LBM_DISPATCH(c, launch_initialize, c);
    if (rc) return rc;
    rc = ev_end(c, r);
    if (rc) return rc;
    c->initialized = true;
    c->diag_valid = false;
    return LBM_OK;
}

int lbm_compute(lbm_ctx* c, int64_t iteration, int32_t update_macro, int32_t diagnostics) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (iteration < 1) return fail(LBM_ERR_INVALID, "iterations are numbered from 1");
    const bool odd = (iteration % 2) != 0;  // lbmcl.hpp:437, 456-457
    const void* src = odd ? c->lat[0] : c->lat[1];
    void* dst = odd ? c->lat[1] : c->lat[0];
    EventRec r;
    int rc = ev_create(c, LBM_EVENT_COMPUTE, r);
    if (rc) return rc;
    rc = LBM_DISPATCH(c, launch_compute, c, src, dst, update_macro ? 1 : 0, diagnostics ? 1 : 0, r.start, r.stop);
    if (rc) {
        if (r.start) (void)hipEventDestroy(r.start);
        if (r.stop) (void)hipEventDestroy(r.stop);
        return rc;
    }
    rc = ev_keep(c, r);
    if (rc) return rc;
    if (diagnostics) {
        EventRec d;
        rc = ev_begin(c, LBM_EVENT_DIAGNOSTICS, d);
        if (rc) return rc;
        hipLaunchKernelGGL(lbm::lbm_reduce_partials, dim3(1), dim3(1024), 0, c->stream, c->partials,
                           compute_blocks(c), c->diag);
        LBM_HIP_CHECK(hipGetLastError(), "diagnostics reduction launch");
        rc = ev_end(c, d);
        if (rc) return rc;
        c->diag_valid = true;
    }
    return LBM_OK;
}

int lbm_compute_range(lbm_ctx* c, int64_t first, int64_t count, int64_t every) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (first < 1 || count < 0 || every < 0) return fail(LBM_ERR_INVALID, "bad iteration range");
    for (int64_t it = first; it < first + count; ++it) {
        const int um = (every > 0 && it % every == 0) ? 1 : 0;
        const int rc = lbm_compute(c, it, um, 0);
        if (rc) return rc;
    }
    return LBM_OK;
}

int lbm_finish(lbm_ctx* c) {
    LBM_CHECK_CTX(c);
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_read_map(lbm_ctx* c, int32_t* out) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    EventRec r;
    int rc = ev_begin(c, LBM_EVENT_READ_MAP, r);
    if (!rc) rc = ensure_map(c);
    if (rc) return rc;
    LBM_HIP_CHECK(hipMemcpyAsync(out, c->map, size_t(c->n3) * 4, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync D2H");
    rc = ev_end(c, r);
    if (rc) return rc;
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_read_rho(lbm_ctx* c, void* out) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    return read_device(c, LBM_EVENT_READ_RHO, out, c->rho, size_t(c->n3) * c->tsize);
}

int lbm_read_u(lbm_ctx* c, void* out) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    return read_device(c, LBM_EVENT_READ_U, out, c->u, size_t(3 * c->n3) * c->tsize);
}

int lbm_read_f(lbm_ctx* c, int32_t buffer, void* out) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    if (buffer != 0 && buffer != 1) return fail(LBM_ERR_INVALID, "buffer must be 0 or 1");
    const size_t bytes = size_t(c->n3) * lbm::Q * c->tsize;
    void* tmp = nullptr;
    int rc = dev_alloc(&tmp, bytes, "f view");
    if (rc) return rc;
    EventRec r;
    rc = ev_begin(c, LBM_EVENT_READ_F, r);
    if (!rc) rc = enqueue_view(c, buffer, tmp, c->stride);
    if (!rc) {
        const hipError_t e = hipMemcpyAsync(out, tmp, bytes, hipMemcpyDeviceToHost, c->stream);
        if (e != hipSuccess) rc = fail(LBM_ERR_RUNTIME, std::string("hipMemcpyAsync D2H: ") + hipGetErrorString(e));
    }
    if (!rc) rc = ev_end(c, r);
    const hipError_t se = hipStreamSynchronize(c->stream);
    (void)hipFree(tmp);
    if (rc) return rc;
    if (se != hipSuccess) return fail(LBM_ERR_RUNTIME, std::string("hipStreamSynchronize: ") + hipGetErrorString(se));
    return LBM_OK;
}

int lbm_snapshot(lbm_ctx* c, int32_t what, int32_t buffer, int64_t* ticket) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!ticket) return fail(LBM_ERR_INVALID, "null ticket");
    size_t bytes = 0;
    int kind = 0;
    switch (what) {
        case LBM_SNAPSHOT_MACRO:
            bytes = size_t(4) * size_t(c->n - 2) * size_t(c->n - 2) * size_t(c->n - 2) * c->tsize;
            kind = LBM_EVENT_READ_RHO;
            break;
        case LBM_SNAPSHOT_F:
            if (buffer != 0 && buffer != 1) return fail(LBM_ERR_INVALID, "buffer must be 0 or 1");
            bytes = size_t(c->n3) * lbm::Q * c->tsize;
            kind = LBM_EVENT_READ_F;
            break;
        case LBM_SNAPSHOT_MAP:
            bytes = size_t(c->n3) * 4;
            kind = LBM_EVENT_READ_MAP;
            break;
        default:
            return fail(LBM_ERR_INVALID, "unknown snapshot kind");
    }
    if (!c->copy_stream) LBM_HIP_CHECK(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking), "hipStreamCreate");
    int idx = -1;
    int64_t t = 0;
    // blocks while every slot is held; fails at once after a failed wait, whose
    // slot may never be released
    if (c->snap_table.take(&idx, &t) != SlotTable<LBM_SNAPSHOT_SLOTS>::OK)
        return fail(LBM_ERR_STATE, "a snapshot copy failed earlier on this context; no new snapshots");
    SnapSlot* s = &c->snaps[idx];
    auto give_back = [&](int rc) {
        c->snap_table.give_back(idx);
        return rc;
    };
    if (s->cap < bytes) {
        if (s->dev) (void)hipFree(s->dev);
        if (s->host) (void)hipHostFree(s->host);
        s->dev = s->host = nullptr;
        s->cap = 0;
        int rc = dev_alloc(&s->dev, bytes, "snapshot staging");
        if (rc) return give_back(rc);
        if (hipHostMalloc(&s->host, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            s->host = nullptr;
            return give_back(fail(LBM_ERR_MEMORY, "hipHostMalloc(snapshot, " + std::to_string(bytes) + " bytes) failed"));
        }
        s->cap = bytes;
    }
    if (!s->copied && hipEventCreateWithFlags(&s->copied, hipEventDisableTiming) != hipSuccess)
        return give_back(fail(LBM_ERR_RUNTIME, "hipEventCreate failed"));
    if (!s->packed && hipEventCreateWithFlags(&s->packed, hipEventDisableTiming) != hipSuccess)
        return give_back(fail(LBM_ERR_RUNTIME, "hipEventCreate failed"));
    s->bytes = bytes;
    EventRec r;
    int rc = ev_begin(c, kind, r);
    if (rc) return give_back(rc);
    if (what == LBM_SNAPSHOT_MACRO) {
        rc = (c->cfg.precision == LBM_F64) ? launch_pack_macro<double>(c, static_cast<double*>(s->dev))
                                           : launch_pack_macro<float>(c, static_cast<float*>(s->dev));
    } else if (what == LBM_SNAPSHOT_F) {
        rc = enqueue_view(c, buffer, s->dev, 1);
    } else {
        rc = ensure_map(c);
        const hipError_t e = rc ? hipSuccess : hipMemcpyAsync(s->dev, c->map, bytes, hipMemcpyDeviceToDevice, c->stream);
        if (e != hipSuccess) rc = fail(LBM_ERR_RUNTIME, std::string("hipMemcpyAsync D2D: ") + hipGetErrorString(e));
    }
    if (!rc) rc = ev_end(c, r);
    if (rc) return give_back(rc);
    hipError_t e = hipEventRecord(s->packed, c->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(c->copy_stream, s->packed, 0);
    if (e == hipSuccess) e = hipMemcpyAsync(s->host, s->dev, bytes, hipMemcpyDeviceToHost, c->copy_stream);
    if (e == hipSuccess) e = hipEventRecord(s->copied, c->copy_stream);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(c->copy_stream);
        return give_back(fail(LBM_ERR_RUNTIME, std::string("snapshot copy: ") + hipGetErrorString(e)));
    }
    *ticket = t;
    return LBM_OK;
}

int lbm_snapshot_wait(lbm_ctx* c, int64_t ticket, const void** data, int64_t* bytes) {
    LBM_CHECK_CTX(c);
    if (!data || !bytes) return fail(LBM_ERR_INVALID, "null output");
    const int idx = c->snap_table.find(ticket);
    if (idx < 0) return fail(LBM_ERR_INVALID, "unknown or released snapshot ticket");
    SnapSlot* s = &c->snaps[idx];
    const hipError_t e = hipEventSynchronize(s->copied);
    // a failed copy leaves the slot FAILED: lbm_snapshot_release still frees it,
    // and later lbm_snapshot calls report the failure instead of waiting
    c->snap_table.waited(idx, e == hipSuccess);
    if (e != hipSuccess) return fail(LBM_ERR_RUNTIME, std::string("hipEventSynchronize: ") + hipGetErrorString(e));
    *data = s->host;
    *bytes = int64_t(s->bytes);
    return LBM_OK;
}

int lbm_snapshot_release(lbm_ctx* c, int64_t ticket) {
    if (!c) return fail(LBM_ERR_INVALID, "null context");
    switch (c->snap_table.release(ticket)) {
        case SlotTable<LBM_SNAPSHOT_SLOTS>::OK: return LBM_OK;
        case SlotTable<LBM_SNAPSHOT_SLOTS>::NOT_WAITED:
            return fail(LBM_ERR_STATE, "snapshot released before lbm_snapshot_wait");
        default: return fail(LBM_ERR_INVALID, "unknown or released snapshot ticket");
    }
}

int lbm_read_state(lbm_ctx* c, int32_t buffer, void* out) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!out) return fail(LBM_ERR_INVALID, "null output");
    if (buffer != 0 && buffer != 1) return fail(LBM_ERR_INVALID, "buffer must be 0 or 1");
    return read_device(c, LBM_EVENT_READ_F, out, c->lat[buffer], size_t(c->n3) * lbm::Q * c->ssize);
}

int lbm_write_state(lbm_ctx* c, int32_t buffer, const void* in) {
    LBM_CHECK_CTX(c);
    LBM_CHECK_INIT(c);
    if (!in) return fail(LBM_ERR_INVALID, "null input");
    if (buffer != 0 && buffer != 1) return fail(LBM_ERR_INVALID, "buffer must be 0 or 1");
    LBM_HIP_CHECK(hipMemcpyAsync(c->lat[buffer], in, size_t(c->n3) * lbm::Q * c->ssize, hipMemcpyHostToDevice,
                                 c->stream),
                  "hipMemcpyAsync H2D");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    return LBM_OK;
}

int lbm_read_diagnostics(lbm_ctx* c, double* mass, double* kinetic_energy) {
    LBM_CHECK_CTX(c);
    if (!mass || !kinetic_energy) return fail(LBM_ERR_INVALID, "null output");
    if (!c->diag_valid) return fail(LBM_ERR_STATE, "no compute with diagnostics enqueued");
    double h[2];
    LBM_HIP_CHECK(hipMemcpyAsync(h, c->diag, sizeof h, hipMemcpyDeviceToHost, c->stream), "hipMemcpyAsync D2H");
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    *mass = h[0];
    *kinetic_energy = h[1];
    return LBM_OK;
}

int lbm_get_constants(lbm_ctx* c, double* out) {
    if (!c || !out) return fail(LBM_ERR_INVALID, "null argument");
    out[0] = c->inv_tau;
    out[1] = c->U;
    out[2] = c->nu;
    out[3] = c->scale;
    return LBM_OK;
}

int lbm_event_count(lbm_ctx* c, int64_t* count) {
    if (!c || !count) return fail(LBM_ERR_INVALID, "null argument");
    *count = int64_t(c->events.done.size() + c->events.pending.size());
    return LBM_OK;
}

int lbm_event_get(lbm_ctx* c, int64_t index, int32_t* kind, double* ms) {
    LBM_CHECK_CTX(c);
    const EventLog& L = c->events;
    const int64_t nd = int64_t(L.done.size());
    if (index < 0 || index >= nd + int64_t(L.pending.size())) return fail(LBM_ERR_INVALID, "event index out of range");
    if (index < nd) {
        if (kind) *kind = L.done[size_t(index)].first;
        if (ms) *ms = double(L.done[size_t(index)].second);
        return LBM_OK;
    }
    const EventRec& e = L.pending[size_t(index - nd)];
    LBM_HIP_CHECK(hipEventSynchronize(e.stop), "hipEventSynchronize");
    float t = 0.f;
    LBM_HIP_CHECK(hipEventElapsedTime(&t, e.start, e.stop), "hipEventElapsedTime");
    if (kind) *kind = e.kind;
    if (ms) *ms = double(t);
    return LBM_OK;
}

int lbm_total_time_ms(lbm_ctx* c, double* ms) {
    LBM_CHECK_CTX(c);
    if (!ms) return fail(LBM_ERR_INVALID, "null output");
    *ms = 0.0;
    const EventLog& L = c->events;
    if (L.pending.empty()) return LBM_OK;  // the newest pair is never recycled
    hipEvent_t start = L.first ? L.first : L.pending.front().start;
    LBM_HIP_CHECK(hipEventSynchronize(L.pending.back().stop), "hipEventSynchronize");
    float t = 0.f;
    LBM_HIP_CHECK(hipEventElapsedTime(&t, start, L.pending.back().stop), "hipEventElapsedTime");
    *ms = double(t);
    return LBM_OK;
}

int lbm_kernels_time_ms(lbm_ctx* c, double* ms) {
    LBM_CHECK_CTX(c);
    if (!ms) return fail(LBM_ERR_INVALID, "null output");
    double total = 0.0;
    for (const auto& d : c->events.done)
        if (d.first == LBM_EVENT_COMPUTE) total += double(d.second);
    for (const EventRec& e : c->events.pending) {
        if (e.kind != LBM_EVENT_COMPUTE) continue;
        LBM_HIP_CHECK(hipEventSynchronize(e.stop), "hipEventSynchronize");
        float t = 0.f;
        LBM_HIP_CHECK(hipEventElapsedTime(&t, e.start, e.stop), "hipEventElapsedTime");
        total += double(t);
    }
    *ms = total;
    return LBM_OK;
}

int lbm_clear_events(lbm_ctx* c) {
    LBM_CHECK_CTX(c);
    LBM_HIP_CHECK(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    free_events(c);
    return LBM_OK;
}

}  // extern "C"
