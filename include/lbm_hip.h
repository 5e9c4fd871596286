/*
 * lbm_hip.h -- C ABI of liblbm_hip.so, the MI355X (gfx950) device layer of the
 * D3Q19 lid-driven-cavity solver.
 *
 * It replaces the OpenCL device layer of LBMCL: platform/device selection,
 * context and profiling queue (libs/CLUtil.hpp:119-199), the JIT build of
 * kernels.cl with -D options (lbmcl.hpp:131-156, CLUtil.hpp:201-229), the five
 * cl::Buffer objects (lbmcl.hpp:403-417), the `initialize` and `compute`
 * kernels (kernels.cl:277-426) with their ping-pong argument binding
 * (lbmcl.hpp:420-467, 490-521), the blocking buffer reads behind the dumps
 * (lbmcl.hpp:159-334) and the event timing (lbmcl.hpp:545-593,
 * CLUtil.hpp:231-243).
 *
 * Conventions: every entry point returns an int status (LBM_OK == 0); no C++
 * exception crosses the ABI; the text of the last error of the calling thread
 * is available from lbm_last_error().  The library owns device memory, the
 * caller owns host buffers.  One context per host thread; each context owns one
 * HIP stream and every enqueue goes to it in order.
 *
 * Device layout of the populations: the CSoA layout of the reference
 * (kernels.cl:64), slot(c, q) = ((c / stride) * 19 + q) * stride + c % stride
 * for cell c = x + y*dim + z*dim^2, with 64-bit slot arithmetic (the reference's
 * int arithmetic overflows at dim = 512).  The device buffers hold the
 * populations in pull ("gather") form -- the post-collision value at its source
 * cell -- and lbm_read_f() returns them in the reference's push view, so the
 * values a caller sees are the reference's.
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define LBM_OK 0
#define LBM_ERR_INVALID (-1)   /* bad argument or configuration */
#define LBM_ERR_DEVICE (-2)    /* no such device */
#define LBM_ERR_MEMORY (-3)    /* device allocation failed */
#define LBM_ERR_RUNTIME (-4)   /* HIP runtime error */
#define LBM_ERR_STATE (-5)     /* call out of order (e.g. compute before initialize) */

/* arithmetic precision (lbm_options.hpp:77 use_double, lbmcl.hpp:144-149) */
#define LBM_F32 0
#define LBM_F64 1

/* population storage; the shifted forms store (f_i - w_i) * scale in 16 bits
 * with FP32 arithmetic and have no counterpart in the reference */
#define LBM_STORE_NATIVE 0
#define LBM_STORE_FP16_SHIFT 1
#define LBM_STORE_BF16_SHIFT 2

/* kernel selection.  The vector row kernels give each lane 16 (VEC) or 8
 * (VEC8) bytes of every population: V = lane bytes / bytes per stored
 * population consecutive cells; they need stride >= V and dim / V >= 64.
 * LBM_KERNEL_AUTO picks the tile kernel for small strides (below), else VEC8
 * when possible, else the tile kernel when possible, else the row kernel when
 * dim >= 64, else the generic kernel.  All kernels compute identical results. */
#define LBM_KERNEL_AUTO 0
#define LBM_KERNEL_GENERIC 1
#define LBM_KERNEL_ROW 2
#define LBM_KERNEL_VEC 3
#define LBM_KERNEL_VEC8 4
/* the tile kernel for small strides (e.g. AoS): tiles of rows staged through
 * LDS with 16-byte accesses; needs stride <= its x extent (16 cells in FP32,
 * 8 in FP64, 32 with 16-bit storage) and dim >= max(that extent, 8).
 * LBM_KERNEL_AUTO selects it when stride * stored bytes <= 32. */
#define LBM_KERNEL_TILE 5
/* the row-pipeline kernel: a persistent grid walking whole rows, each row's
 * 19 source spans staged in LDS by direct global-to-LDS loads issued one row
 * ahead; needs dim 128, 256, 512 or 1024 with two rows of every population
 * within 160 KB of LDS and a wavefront per row (FP32 and FP64 from 128, 16-bit
 * from 256, all up to 512, FP32 and 16-bit at 1024) and
 * stride >= 16 / bytes per stored value. */
#define LBM_KERNEL_LDS 6

/* lbm_config.optimize: arithmetic form of the collision (see lbm_compute).
 * The zero value is the reference's own arithmetic: with native storage,
 * kernels.cl's expressions as ROCm's OpenCL compiler evaluates them (the
 * reference's results on this GPU, bit for bit); 16-bit storage has no
 * counterpart in the reference and always computes in its own deviation form
 * (see lbm_compute), whatever optimize says (LBM_ARITH_EXACT is rejected). */
#define LBM_ARITH_REFERENCE 0 /* default: LBM_ARITH_OPENCL for native storage; 16-bit: the deviation form */
#define LBM_ARITH_OPENCL 0    /* kernels.cl's expressions as ROCm's OpenCL compiler evaluates them */
#define LBM_ARITH_FUSED 1     /* -o: the fused BGK form (native storage) */
#define LBM_ARITH_EXACT 2     /* the CPU oracle's form: kernels.cl's expressions operation by operation,
                                 no contraction, correctly rounded division; a checking aid, not the
                                 reference's rounding (native storage only) */

/* event kinds, named like the reference's events (lbmcl.hpp:26-31) */
#define LBM_EVENT_INITIALIZE 0
#define LBM_EVENT_COMPUTE 1
#define LBM_EVENT_READ_MAP 2
#define LBM_EVENT_READ_F 3
#define LBM_EVENT_READ_RHO 4
#define LBM_EVENT_READ_U 5
#define LBM_EVENT_DIAGNOSTICS 6

/* Both configuration structs start with their own size: set struct_size to
 * sizeof(lbm_config) (LBM_CONFIG_INIT does) and the library rejects a struct
 * of any other size with LBM_ERR_INVALID, so a caller built against another
 * layout fails loudly instead of passing shifted fields.  Every other field's
 * zero value is its default (dim and stride have none). */
typedef struct lbm_config {
    uint32_t struct_size; /* sizeof(lbm_config) */
    int32_t dim;         /* lattice cube dimension, a power of two (lbmcl.hpp:366-369) */
    int32_t precision;   /* LBM_F32 or LBM_F64 */
    int32_t storage;     /* LBM_STORE_*; shifted storage requires LBM_F32 */
    int32_t device;      /* HIP device ordinal (replaces -P/-D, CLUtil.hpp:119-178) */
    int64_t stride;      /* CSoA stride, a power of two in [1, dim^3] (lbmcl.hpp:384-387) */
    double viscosity;    /* kinematic viscosity nu; tau = 3 nu + 1/2 (kernels.cl:61-62) */
    double velocity;     /* lid x-velocity U (kernels.cl:57) */
    int32_t lws[3];      /* work-group size of the generic kernel (product <= 1024); 0 = default */
    int32_t optimize;    /* LBM_ARITH_*: 0 the reference's arithmetic, 1 is -o (fused collision) */
    int32_t kernel;      /* LBM_KERNEL_* */
    int32_t profile;     /* nonzero: record a start/stop event pair around every enqueue */
    int32_t xcd_remap;   /* 0 (default) or positive: consecutive row blocks go to one XCD (row and
                            vector kernels); negative: the plain block order */
    int32_t raw_parameters; /* nonzero: use viscosity and velocity as given, without the 6-digit
                               rounding of the reference's -D literals (for gradient checks) */
} lbm_config;

/* lbm_config cfg = LBM_CONFIG_INIT;  -- struct_size set, every other field zero
 * (a designated initializer in C, so -Wmissing-field-initializers stays quiet) */
#ifdef __cplusplus
#define LBM_CONFIG_INIT {(uint32_t)sizeof(lbm_config)}
#else
#define LBM_CONFIG_INIT {.struct_size = (uint32_t)sizeof(lbm_config)}
#endif

/* objectives of lbm_adjoint (no counterpart in the reference) */
#define LBM_OBJECTIVE_KINETIC_ENERGY 0  /* J = sum over FLUID/MOVING cells of rho |u|^2 / 2 */
#define LBM_OBJECTIVE_MEAN_UX_PLANE 1   /* J = mean u_x over the FLUID/MOVING cells of plane z */

typedef struct lbm_adjoint_config {
    uint32_t struct_size;      /* sizeof(lbm_adjoint_config) */
    int32_t objective;         /* LBM_OBJECTIVE_* */
    int32_t plane_z;           /* plane of LBM_OBJECTIVE_MEAN_UX_PLANE, 1 .. dim-2 */
    int64_t steps;             /* T: the objective is evaluated on the state after T iterations */
    int64_t checkpoint_every;  /* K: forward states kept every K steps; 0 = round(sqrt(T)) */
} lbm_adjoint_config;

#ifdef __cplusplus
#define LBM_ADJOINT_CONFIG_INIT {(uint32_t)sizeof(lbm_adjoint_config)}
#else
#define LBM_ADJOINT_CONFIG_INIT {.struct_size = (uint32_t)sizeof(lbm_adjoint_config)}
#endif

typedef struct lbm_adjoint_result {
    double objective;     /* J */
    double d_velocity;    /* dJ/dU (lid velocity) */
    double d_viscosity;   /* dJ/dnu */
    double forward_ms;    /* forward pass with checkpointing */
    double adjoint_ms;    /* objective seed, recomputation and reverse sweep */
    double device_bytes;  /* device memory held during the run */
    int64_t checkpoints;  /* number of checkpoints (segments) */
    int64_t segment;      /* K */
} lbm_adjoint_result;

typedef struct lbm_ctx lbm_ctx;

/* Number of HIP devices (CLUtil.hpp:145-178 lists devices). */
int lbm_get_device_count(int32_t* count);

/* Device name (replaces CL_DEVICE_NAME, lbmcl.hpp:626, 651): the HIP
 * architecture name, e.g. "gfx950:sramecc+:xnack-", as ROCm's OpenCL reports it. */
int lbm_get_device_name(int32_t device, char* name, size_t len);

/* Create a context: select the device, create the stream, allocate the two
 * population lattices, rho, u and map (lbmcl.hpp:392-417).  viscosity and
 * velocity are rounded the way the reference's -D build options round them:
 * the value is converted to the arithmetic type, printed with 6 significant
 * digits and parsed back as a literal of that type (lbmcl.hpp:140-146). */
int lbm_create(const lbm_config* config, lbm_ctx** ctx);

/* Release the context and all its device memory (lbmcl.hpp:672-678). */
int lbm_destroy(lbm_ctx* ctx);

/* Enqueue the `initialize` kernel (kernels.cl:277-318, lbmcl.hpp:493-498):
 * map, rho = 1, u = (U,0,0) on FRONT cells and 0 elsewhere (NaN outside
 * FLUID/MOVING cells), equilibrium populations in both lattices. */
int lbm_initialize(lbm_ctx* ctx);

/* Enqueue iteration `iteration` (>= 1) of the `compute` kernel
 * (kernels.cl:321-426, lbmcl.hpp:505-511): odd iterations read lattice 0 and
 * write lattice 1, even iterations the reverse (lbmcl.hpp:437, 456-457).
 * update_macro stores rho and u of FLUID/MOVING cells (kernels.cl:383-388).
 * diagnostics != 0 also reduces the mass sum(rho) and kinetic energy
 * sum(rho |u|^2 / 2) over FLUID/MOVING cells in FP64; read them with
 * lbm_read_diagnostics (no counterpart in the reference).
 * Arithmetic (lbm_config.optimize): LBM_ARITH_OPENCL, the default for native
 * storage, evaluates kernels.cl:343-418 as ROCm's OpenCL compiler builds it
 * without -o: products added or subtracted within one expression are
 * contracted into fused multiply-adds (OpenCL C's default FP_CONTRACT ON, as
 * clang forms them) and FP32 divisions round to OpenCL's 2.5 ulp;
 * lbm_initialize then uses the same contracted equilibrium.  This is the
 * reference's own arithmetic on this GPU.  LBM_ARITH_FUSED (the -o option, the
 * reference's -cl-fast-relaxed-math) uses a
 * fused BGK form, f' = (1 - 1/tau) f + (rho w_i / tau) T_i with
 * T_i = (1 - 1.5 u^2) + e (3 + 4.5 e), e = e_i . u, evaluated with fused
 * multiply-adds; rho and u are unchanged, results differ by rounding.
 * LBM_ARITH_EXACT evaluates each expression operation by operation without
 * contraction and with correctly rounded division: the CPU oracle's form, bit
 * for bit, whose FP32 results differ from the reference's by rounding.
 * 16-bit storage computes in FP32 on the deviations d_i = f_i - w_i that its
 * lattice stores: rho = 1 + sum d_i, j = sum e_i d_i, and the BGK update
 * d_i' = (w_i / tau) (rho tau_i + (rho - 1)) + (1 - 1/tau) d_i with
 * tau_i = T_i - 1, so no population is ever rounded next to w_i in FP32
 * (which biases the mass by ~3e-7 of itself per step with 16-bit inputs). */
int lbm_compute(lbm_ctx* ctx, int64_t iteration, int32_t update_macro, int32_t diagnostics);

/* Enqueue iterations first .. first+count-1; update_macro is set on those
 * divisible by `every` when every > 0 (lbmcl.hpp:436). */
int lbm_compute_range(lbm_ctx* ctx, int64_t first, int64_t count, int64_t every);

/* Block until the stream is idle (queue.finish(), lbmcl.hpp:525-532). */
int lbm_finish(lbm_ctx* ctx);

/* Blocking reads into caller memory (lbmcl.hpp:159-166, 261-277).  map: dim^3
 * int32 cell flags (common.h:7-20); rho: dim^3 values, u: 3*dim^3 values
 * [ux | uy | uz], of the arithmetic type (float for shifted storage). */
int lbm_read_map(lbm_ctx* ctx, int32_t* out);
int lbm_read_rho(lbm_ctx* ctx, void* out);
int lbm_read_u(lbm_ctx* ctx, void* out);

/* Blocking read of lattice `buffer` (0: read by odd iterations, the
 * reference's f_collide; 1: read by even iterations, its f_stream) in the
 * reference's CSoA push view, 19*dim^3 values of the arithmetic type
 * (lbmcl.hpp:206-214, 503, 517-519). */
int lbm_read_f(lbm_ctx* ctx, int32_t buffer, void* out);

/* Blocking read of lattice `buffer` as stored on the device (pull form, storage
 * type: 4 or 8 bytes per slot native, 2 bytes shifted).  For checkpoints and
 * tests; no counterpart in the reference. */
int lbm_read_state(lbm_ctx* ctx, int32_t buffer, void* out);

/* Blocking write of lattice `buffer` in stored (pull) form; restores a state
 * read by lbm_read_state.  No counterpart in the reference. */
int lbm_write_state(lbm_ctx* ctx, int32_t buffer, const void* in);

/* Asynchronous snapshots for the output path.  They replace the blocking
 * enqueueReadBuffer calls of the reference's dump writers (lbmcl.hpp:164, 211,
 * 268, 274) where the caller formats files on other threads while the device
 * keeps computing.
 *
 * lbm_snapshot enqueues, after all work enqueued so far on the context's
 * stream, a pack of the requested data into a device staging slot and the
 * copy of that slot into pinned host memory on a second stream, and returns a
 * ticket without waiting.  The context has LBM_SNAPSHOT_SLOTS slots; when all
 * are taken (enqueued, or waited and not yet released) lbm_snapshot blocks
 * until another thread releases one.  lbm_snapshot_wait blocks until the copy
 * of `ticket` is complete and returns its host data, valid until
 * lbm_snapshot_release(ticket).  lbm_snapshot_wait and lbm_snapshot_release
 * may be called from any thread; lbm_destroy waits for outstanding copies.
 *
 *   LBM_SNAPSHOT_MACRO  rho and u on the VTI window x, y, z in [1, dim-2]
 *                       (x fastest, W = dim-2): W^3 rho values, then W^3
 *                       (ux, uy, uz) triples; arithmetic type; `buffer` unused.
 *   LBM_SNAPSHOT_F      lattice `buffer` in the push view of lbm_read_f, cell
 *                       major: value (c, i) at c*19 + i, dim^3 cells.
 *   LBM_SNAPSHOT_MAP    the dim^3 int32 cell flags of lbm_read_map.
 * Packing is recorded as a READ_RHO, READ_F or READ_MAP event. */
#define LBM_SNAPSHOT_MACRO 0
#define LBM_SNAPSHOT_F 1
#define LBM_SNAPSHOT_MAP 2
#define LBM_SNAPSHOT_SLOTS 3
int lbm_snapshot(lbm_ctx* ctx, int32_t what, int32_t buffer, int64_t* ticket);
int lbm_snapshot_wait(lbm_ctx* ctx, int64_t ticket, const void** data, int64_t* bytes);
int lbm_snapshot_release(lbm_ctx* ctx, int64_t ticket);

/* Mass and kinetic energy of the last lbm_compute with diagnostics != 0
 * (blocking). */
int lbm_read_diagnostics(lbm_ctx* ctx, double* mass, double* kinetic_energy);

/* Constants the kernels use, widened to double: out[0] = 1/tau, out[1] = U,
 * out[2] = nu, out[3] = storage scale (1 for native storage). */
int lbm_get_constants(lbm_ctx* ctx, double* out);

/* Event timing (profile != 0), the reference's named cl::Event list
 * (lbmcl.hpp:74, 545-593). */
int lbm_event_count(lbm_ctx* ctx, int64_t* count);
int lbm_event_get(lbm_ctx* ctx, int64_t index, int32_t* kind, double* ms);
/* First event start to last event end (lbmcl.hpp:548-556). */
int lbm_total_time_ms(lbm_ctx* ctx, double* ms);
/* Sum of the compute events (lbmcl.hpp:562-574). */
int lbm_kernels_time_ms(lbm_ctx* ctx, double* ms);
/* Forget recorded events (keeps the context). */
int lbm_clear_events(lbm_ctx* ctx);

/* Body force g(x) on FLUID cells: 3*dim^3 values [gx | gy | gz] of the
 * arithmetic type, copied to the device; each collision then adds
 * 3 w_i (e_i . g) to the populations (a zero field changes nothing).  NULL
 * removes the force.  No counterpart in the reference. */
int lbm_set_force(lbm_ctx* ctx, const void* force);

/* Run the cavity from its initial state for config->steps iterations and
 * return the objective J and its gradients with respect to the lid velocity,
 * the viscosity and (if grad_force is not NULL: 3*dim^3 doubles) the body
 * force, by the adjoint collide-stream step with checkpoints every
 * config->checkpoint_every steps and recomputation of the segments.  Native
 * storage only; the arithmetic type of the context is used for the adjoint
 * states, FP64 for the gradient sums.  The context's lattices are left
 * unspecified (call lbm_initialize before computing again).  No counterpart
 * in the reference. */
int lbm_adjoint(lbm_ctx* ctx, const lbm_adjoint_config* config, lbm_adjoint_result* result, double* grad_force);

/* One collide-stream iteration that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * It runs the kernels of lbm_compute(ctx, it, update_macro, 1) -- same kernel
 * selection (cfg->kernel), arithmetic (cfg->optimize) and results, bit for
 * bit -- on caller-owned device memory and enqueues on `stream` (a
 * hipStream_t; NULL: the null stream) without waiting:
 *   src, dst  the lattice read and the lattice written, pull form as stored
 *             (19 * dim^3 values of the storage type, CSoA with cfg->stride;
 *             lbm_read_state/lbm_write_state move the same bytes), distinct;
 *   rho, u    dim^3 and 3 * dim^3 values of the arithmetic type, written on
 *             FLUID/MOVING cells when update_macro != 0 (may be NULL otherwise);
 *   diag      2 doubles of device memory: mass and kinetic energy (FP64).
 * cfg is validated like lbm_create's (struct_size, dim, stride, ...) and
 * cfg->device selects the device (the caller's current device is restored).
 * The per-block partial sums live in a fixed array of the library on each
 * device, 2^19 blocks (8 MiB, enough for every form at dim 512); a
 * configuration whose launch has more blocks is rejected with
 * LBM_ERR_INVALID.  Calls that overlap on one device share that array and
 * race: serialise them (one stream, or events).  No body force. */
int lbm_compute_noalloc(const lbm_config* cfg, const void* src, void* dst, void* rho, void* u, int32_t update_macro,
                        double* diag, void* stream);

/* Text of the calling thread's last error ("" if none). */
const char* lbm_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* LBM_HIP_H */
