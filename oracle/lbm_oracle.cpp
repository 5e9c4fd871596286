// lbm_oracle.cpp -- CPU oracle for the D3Q19 lid-driven-cavity collide-stream path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
// check in __graft_entry__.py and the cpu_baseline leg of bench.py load it.  The
// product (lbmcl_amd/liblbm_hip.so and the lbmcl executable) never links or
// calls it.
//
// It is a scalar restatement of the reference algorithm (LBMCL, kernels.cl and
// common.h), written from the reference's behaviour:
//   * cell classification          kernels.cl:236-266, common.h:7-66
//   * initialize kernel            kernels.cl:277-318
//   * compute kernel (push form)   kernels.cl:321-426
//   * lattice tables               kernels.cl:110-226
//   * CSoA slot index              kernels.cl:64-70, lbmcl.hpp:19-23
//   * JIT constant folding         lbmcl.hpp:131-156 (-cl-single-precision-constant
//                                  makes every literal a float in FP32 builds)
// Arithmetic follows the reference expressions operation by operation, left to
// right, without contraction (the file is built with -ffp-contract=off).
//
// It also restates the pull ("gather") form the HIP kernels use, so that the
// equivalence of the two forms is checked on the CPU bit for bit, plus the
// shifted FP16/BF16 storage form, a body-force term and the mass / kinetic
// energy diagnostics, which have no counterpart in the reference.
//
// Everything is exported with C linkage and plain pointers for ctypes.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <cstdlib>
#include <vector>
#include <limits>
#include <omp.h>

#ifndef LBMO_MUTANT
#define LBMO_MUTANT 0  // planted-mistake variants for the oracle's own tests; 0 = the oracle
#endif

namespace {

constexpr int Q = 19;

// Lattice velocities and opposites (kernels.cl:131-226).
constexpr int EX[Q] = {0, +1, 0, -1, 0, 0, 0, +1, -1, -1, +1, +1, 0, -1, 0, +1, 0, -1, 0};
constexpr int EY[Q] = {0, 0, +1, 0, -1, 0, 0, +1, +1, -1, -1, 0, +1, 0, -1, 0, +1, 0, -1};
constexpr int EZ[Q] = {0, 0, 0, 0, 0, -1, +1, 0, 0, 0, 0, -1, -1, -1, -1, +1, +1, +1, +1};
constexpr int OPP[Q] = {0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12};

// Cell flags (common.h:7-18).
constexpr int32_t NONE = 0x000, FLUID = 0x001, MOVING = 0x002, CORNER = 0x004, WALL = 0x008;
constexpr int32_t LEFT = 0x010, RIGHT = 0x020, TOP = 0x040, BOTTOM = 0x080, FRONT = 0x100, BACK = 0x200;
constexpr int32_t BOUNDARY_MASK = LEFT | RIGHT | BOTTOM | TOP | BACK | FRONT;

inline bool is_fluid(int32_t t) { return t == FLUID; }            // common.h:28
inline bool is_wall(int32_t t) { return t == WALL; }              // common.h:33
inline bool is_boundary(int32_t t) { return (t & BOUNDARY_MASK) != 0; }  // common.h:43
inline bool is_moving(int32_t t) { return (t & MOVING) != 0; }    // common.h:48
inline bool is_moving_init(int32_t t) { return (t & FRONT) != 0; }       // common.h:23
inline bool is_store_macro(int32_t t) { return is_fluid(t) || is_moving(t); }  // common.h:53
inline bool is_collision(int32_t t) { return is_fluid(t) || is_moving(t); }    // common.h:58
inline bool is_bounceback(int32_t t) { return is_boundary(t) && !(t & MOVING); } // common.h:63

// kernels.cl:236-266
int32_t cell_type(int64_t N, int64_t x, int64_t y, int64_t z) {
    int32_t t = NONE;
    if (x == 1) t |= LEFT;
    if (x == N - 2) t |= RIGHT;
    if (y == 1) t |= BOTTOM;
    if (y == N - 2) t |= TOP;
    if (z == 1) t |= BACK;
    if (z == N - 2) t |= FRONT;
    if (x == 0 || x == N - 1 || y == 0 || y == N - 1 || z == 0 || z == N - 1) t = WALL;
    if (t == (LEFT | BACK | BOTTOM) || t == (RIGHT | BACK | BOTTOM) ||
        t == (LEFT | BACK | TOP) || t == (RIGHT | BACK | TOP))
        t = CORNER;
    if (t == FRONT) t |= MOVING;
    if (t == NONE) t = FLUID;
    return t;
}

struct Params {
    int32_t dim;
    int32_t threads;
    int64_t stride;
    double viscosity;   // value of the VISCOSITY literal
    double velocity;    // value of the VELOCITY literal
    int32_t fast;       // 0: kernels.cl's expressions as written; 1: the fused collision (no
                        // counterpart in the reference); 2: kernels.cl's expressions contracted as
                        // clang compiles OpenCL C (FP_CONTRACT ON)
    int32_t reserved;
};

// Slot of population q of cell id in the CSoA layout (kernels.cl:64).
inline int64_t slot(int64_t id, int q, int64_t stride) {
    return ((id / stride) * Q + q) * stride + (id & (stride - 1));
}

// Constants exactly as the JIT build folds them (kernels.cl:56-62, 110-128).
// In FP32 builds every literal is single precision (lbmcl.hpp:146).
template <class T> struct Consts {
    T w[Q];
    T inv_tau;
    T U;
    T om1;       // 1 - inv_tau
    bool fast;
    bool contracted;
    explicit Consts(const Params& p) {
        const T one = T(1), three = T(3), half = T(0.5);
        const T nu = T(p.viscosity);
        U = T(p.velocity);
        const T tau = (three * nu) + half;
        inv_tau = one / tau;
        om1 = one - inv_tau;
        fast = p.fast == 1;
        contracted = p.fast == 2;
        const T w0 = one / T(3), w1 = one / T(18), w7 = one / T(36);
        for (int i = 0; i < Q; ++i) w[i] = (i == 0) ? w0 : (i <= 6 ? w1 : w7);
    }
};

// Equilibrium of population i, expression of kernels.cl:307-308, 396-397, 415-416:
//   (rho * w_i) * (1 + 3 eu + 4.5 eu eu - 1.5 u2), evaluated left to right.
template <class T>
inline T feq(const Consts<T>& k, int i, T rho, T ux, T uy, T uz, T u2) {
    const T eu = ((ux * T(EX[i])) + (uy * T(EY[i]))) + (uz * T(EZ[i]));
    return (rho * k.w[i]) * (((T(1) + (T(3) * eu)) + ((T(4.5) * eu) * eu)) - (T(1.5) * u2));
}

// The same expressions with clang's OpenCL contractions (FP_CONTRACT ON):
// inside one expression x + a*b and x - a*b become fma(a, b, x) and
// fma(-a, b, x), the left operand's product taken when both are products.
// e_i . u multiplies by 0 and +-1 only, so it is exact in either form.
template <class T>
inline T feq_contracted(const Consts<T>& k, int i, T rho, T ux, T uy, T uz, T u2) {
    const T eu = ((ux * T(EX[i])) + (uy * T(EY[i]))) + (uz * T(EZ[i]));
    const T a = std::fma(T(3), eu, T(1));
    const T b = std::fma(T(4.5) * eu, eu, a);
    return (rho * k.w[i]) * std::fma(-T(1.5), u2, b);
}
// ((ux ux) + (uy uy)) + (uz uz) (kernels.cl:390) contracted
template <class T>
inline T u2_contracted(T ux, T uy, T uz) {
    return std::fma(uz, uz, std::fma(ux, ux, uy * uy));
}

// rho and u moments, kernels.cl:355, 376-378.
template <class T>
inline T moment_rho(const T* f) {
    T r = f[0];
    for (int i = 1; i < Q; ++i) r = r + f[i];
    return r;
}
template <class T>
inline void moment_u(const T* f, T rho, T& ux, T& uy, T& uz) {
#if LBMO_MUTANT == 5  // planted mistake: population 10 counted in the wrong direction
    ux = ((((f[1] + f[7]) + f[8]) + f[11]) + f[15]) - ((((f[3] + f[10]) + f[9]) + f[13]) + f[17]);
#else
    ux = ((((f[1] + f[7]) + f[10]) + f[11]) + f[15]) - ((((f[3] + f[8]) + f[9]) + f[13]) + f[17]);
#endif
    uy = ((((f[2] + f[7]) + f[8]) + f[12]) + f[16]) - ((((f[4] + f[9]) + f[10]) + f[14]) + f[18]);
    uz = ((((f[6] + f[15]) + f[16]) + f[17]) + f[18]) - ((((f[5] + f[11]) + f[12]) + f[13]) + f[14]);
    ux = ux / rho;
    uy = uy / rho;
    uz = uz / rho;
}

// Fused collision (no counterpart in the reference; the HIP path uses it for
// 16-bit storage and, on request, for native storage): the same moments and
// boundary handling, and the BGK update written as
//   f_i' = (1 - 1/tau) f_i + (rho w_i / tau) T_i,  T_i = K + e (3 + 4.5 e),
//   K = 1 - 1.5 u2,  e = e_i . u,
// with fused multiply-adds, T shared between opposite populations as
// A +- 3e, A = 4.5 e e + K.  Each std::fma is a single IEEE rounding.
template <class T>
inline void collide_fused(const Consts<T>& k, int32_t t, T* f, T rho, T ux, T uy, T uz, const T* force) {
    const bool moving = is_moving(t);
    if (!moving && !is_fluid(t)) {
        if (is_bounceback(t)) {
            const int half[9] = {1, 2, 5, 7, 8, 11, 12, 13, 14};
            for (int h : half) {
                const T tmp = f[h];
                f[h] = f[OPP[h]];
                f[OPP[h]] = tmp;
            }
        }
        return;
    }
    const T u2 = std::fma(uz, uz, std::fma(uy, uy, ux * ux));
    const T K = std::fma(T(-1.5), u2, T(1));
    const T s = moving ? rho : k.inv_tau * rho;
    const T c0 = s * k.w[0], c1 = s * k.w[1], c7 = s * k.w[7];
    const T keep = moving ? T(0) : k.om1;
    const int PI[9] = {1, 2, 6, 7, 10, 15, 16, 11, 12};
    const int PO[9] = {3, 4, 5, 9, 8, 13, 14, 17, 18};
    f[0] = std::fma(c0, K, keep * f[0]);
    for (int p = 0; p < 9; ++p) {
        const int i = PI[p], o = PO[p];
        const T e = ((ux * T(EX[i])) + (uy * T(EY[i]))) + (uz * T(EZ[i]));
        const T A = std::fma(T(4.5) * e, e, K);
        const T B = T(3) * e;
        const T c = (i <= 6) ? c1 : c7;
        f[i] = std::fma(c, A + B, keep * f[i]);
        f[o] = std::fma(c, A - B, keep * f[o]);
    }
    if (force && is_fluid(t)) {
        for (int i = 0; i < Q; ++i) {
            const T eg = ((force[0] * T(EX[i])) + (force[1] * T(EY[i]))) + (force[2] * T(EZ[i]));
            f[i] = f[i] + (T(3) * k.w[i]) * eg;
        }
    }
}

// Deviation ("delta") form of the step, used with 16-bit storage (no
// counterpart in the reference).  The populations are held as d_i = f_i - w_i
// from load to store -- the stored 16-bit values are scaled d_i -- so no value
// near w_i is ever rounded in FP32: the sums, the equilibrium and the
// relaxation work on small deviations.  The moments are
//   drho = d_0 + ... + d_18 (rho = 1 + drho),  j = sum e_i d_i  (sum e_i w_i = 0),
// the lid fix-up and bounce-back move d_i as the reference moves f_i, and the
// BGK update f' = (1 - w) f + w rho w_i T_i is written for the deviations:
//   d_i' = (w w_i) (rho tau_i + drho) + (1 - w) d_i,   tau_i = T_i - 1,
//   tau_i = Km1 + e (3 + 4.5 e),  Km1 = -1.5 u2,  e = e_i . u,
// (rho tau_i + drho = rho T_i - 1), with tau shared between opposite
// populations as A +- 3e, A = 4.5 e e + Km1; the lid sets d_i = w_i (rho tau_i +
// drho).  Each std::fma is one IEEE rounding; the GPU kernels evaluate the
// same operations in the same order.
inline void delta_eq_all(const Consts<float>& k, float rho, float drho, float ux, float uy, float uz, float cw0,
                         float cw1, float cw7, float keep, float* d) {
    // d_i <- fma(cw_i, fma(rho, tau_i, drho), keep * d_i)
    const float u2 = std::fma(uz, uz, std::fma(uy, uy, ux * ux));
    const float Km1 = -1.5f * u2;
    const int PI[9] = {1, 2, 6, 7, 10, 15, 16, 11, 12};
    const int PO[9] = {3, 4, 5, 9, 8, 13, 14, 17, 18};
    d[0] = std::fma(cw0, std::fma(rho, Km1, drho), keep * d[0]);
    for (int p = 0; p < 9; ++p) {
        const int i = PI[p], o = PO[p];
        const float e = ((ux * float(EX[i])) + (uy * float(EY[i]))) + (uz * float(EZ[i]));
        const float A = std::fma(4.5f * e, e, Km1);
        const float B = 3.0f * e;
        const float c = (i <= 6) ? cw1 : cw7;
        d[i] = std::fma(c, std::fma(rho, A + B, drho), keep * d[i]);
        d[o] = std::fma(c, std::fma(rho, A - B, drho), keep * d[o]);
    }
    (void)k;
}

// The local step in deviation form; on entry d holds the gathered deviations,
// on return the ones to store; rho (= 1 + drho) and u as the reference defines them.
inline void collide_delta(const Consts<float>& k, int32_t t, float* d, float* rho_out, float* ux_out,
                          float* uy_out, float* uz_out, const float* force) {
    const bool moving = is_moving(t);
    if (moving) {  // kernels.cl:343-349
        d[5] = d[6];
        d[11] = d[17];
        d[12] = d[18];
        d[13] = d[15];
        d[14] = d[16];
    }
    const float drho = moment_rho(d);
    const float rho = 1.0f + drho;
    float ux, uy, uz;
    if (moving) {  // kernels.cl:362-365
        ux = k.U;
        uy = 0.0f;
        uz = 0.0f;
    } else {
        moment_u(d, rho, ux, uy, uz);
    }
    *rho_out = rho;
    *ux_out = ux;
    *uy_out = uy;
    *uz_out = uz;
    if (!moving && !is_fluid(t)) {
        if (is_bounceback(t)) {
            const int half[9] = {1, 2, 5, 7, 8, 11, 12, 13, 14};
            for (int h : half) {
                const float tmp = d[h];
                d[h] = d[OPP[h]];
                d[OPP[h]] = tmp;
            }
        }
        return;
    }
    if (moving) {
        delta_eq_all(k, rho, drho, ux, uy, uz, k.w[0], k.w[1], k.w[7], 0.0f, d);
    } else {
        const float om = k.inv_tau;
        delta_eq_all(k, rho, drho, ux, uy, uz, om * k.w[0], om * k.w[1], om * k.w[7], k.om1, d);
        if (force) {
            for (int i = 0; i < Q; ++i) {
                const float eg = ((force[0] * float(EX[i])) + (force[1] * float(EY[i]))) + (force[2] * float(EZ[i]));
                d[i] = d[i] + (3.0f * k.w[i]) * eg;
            }
        }
    }
}

// Initial deviations of a cell: w_i tau_i at rho = 1 (drho = 0) and
// u = (U, 0, 0) on FRONT cells, 0 elsewhere; NaN for walls.
inline void init_cell_delta(const Consts<float>& k, int32_t t, float* d) {
    if (is_wall(t)) {
        for (int i = 0; i < Q; ++i) d[i] = std::numeric_limits<float>::quiet_NaN();
        return;
    }
    for (int i = 0; i < Q; ++i) d[i] = 0.0f;
    if (is_moving_init(t)) delta_eq_all(k, 1.0f, 0.0f, k.U, 0.0f, 0.0f, k.w[0], k.w[1], k.w[7], 0.0f, d);
}

// The local part of the compute kernel (kernels.cl:343-418): lid fix-up,
// moments, optional macro store, boundary conditions and BGK collision.
// `f` holds the 19 populations that arrived at the cell; on return it holds the
// values the cell sends.  `force` (3 components, may be null) adds
// 3 w_i (e_i . g) to the post-collision populations of FLUID cells; with a zero
// or absent force the result is bit-identical to the reference.
template <class T>
inline void collide(const Consts<T>& k, int32_t t, T* f, T* rho_out, T* ux_out, T* uy_out, T* uz_out,
                    const T* force) {
    if (is_moving(t)) {  // kernels.cl:343-349
        f[5] = f[6];
        f[11] = f[17];
        f[12] = f[18];
#if LBMO_MUTANT == 1  // planted mistake: wrong partner in the lid fix-up
        f[13] = f[17];
#else
        f[13] = f[15];
#endif
        f[14] = f[16];
    }
    const T rho = moment_rho(f);
    T ux, uy, uz;
    if (is_moving(t)) {  // kernels.cl:362-365
        ux = k.U;
        uy = T(0);
        uz = T(0);
    } else {
        moment_u(f, rho, ux, uy, uz);
    }
    if (rho_out) {
        *rho_out = rho;
        *ux_out = ux;
        *uy_out = uy;
        *uz_out = uz;
    }
    if (k.fast) {
        collide_fused(k, t, f, rho, ux, uy, uz, force);
        return;
    }
    const T u2 = k.contracted ? u2_contracted(ux, uy, uz) : ((ux * ux) + (uy * uy)) + (uz * uz);  // kernels.cl:390
    auto eq = [&](int i) { return k.contracted ? feq_contracted(k, i, rho, ux, uy, uz, u2)
                                               : feq(k, i, rho, ux, uy, uz, u2); };
    if (is_moving(t)) {  // kernels.cl:393-398
        for (int i = 0; i < Q; ++i) f[i] = eq(i);
    } else if (is_bounceback(t)
#if LBMO_MUTANT == 2  // planted mistake: corners bounce back too
               || t == CORNER
#endif
    ) {  // kernels.cl:400-408
#if LBMO_MUTANT == 3  // planted mistake: one wrong opposite
        const int opp[Q] = {0, 3, 4, 1, 2, 6, 5, 10, 9, 8, 7, 17, 18, 15, 16, 13, 14, 11, 12};
        const int half[9] = {1, 2, 5, 7, 9, 11, 12, 13, 14};
#else
        const int* opp = OPP;
        const int half[9] = {1, 2, 5, 7, 8, 11, 12, 13, 14};
#endif
        for (int h : half) {
            const T tmp = f[h];
            f[h] = f[opp[h]];
            f[opp[h]] = tmp;
        }
    }
    if (is_collision(t)) {  // kernels.cl:412-418, 270-273
#if LBMO_MUTANT == 4  // planted mistake: relax with tau instead of 1/tau
        const T omega = T(1) / k.inv_tau;
#else
        const T omega = k.inv_tau;
#endif
        for (int i = 0; i < Q; ++i) {
            const T e = eq(i);
            f[i] = k.contracted ? std::fma(omega, e - f[i], f[i]) : f[i] + omega * (e - f[i]);
        }
        if (force && is_fluid(t)) {
            for (int i = 0; i < Q; ++i) {
                const T eg = ((force[0] * T(EX[i])) + (force[1] * T(EY[i]))) + (force[2] * T(EZ[i]));
                f[i] = f[i] + (T(3) * k.w[i]) * eg;
            }
        }
    }
}

template <class T>
inline T nan_v() { return std::numeric_limits<T>::quiet_NaN(); }

// Initial populations of a cell (kernels.cl:292-317); NaN for walls.
template <class T>
inline void init_cell(const Consts<T>& k, int32_t t, T* f, T& rho, T& ux, T& uy, T& uz) {
    rho = T(1);
#if LBMO_MUTANT == 6  // planted mistake: only lid cells start moving (not the lid edges)
    ux = is_moving(t) ? k.U : T(0);
#else
    ux = is_moving_init(t) ? k.U : T(0);
#endif
    uy = T(0);
    uz = T(0);
    if (k.contracted) {
        const T u2 = u2_contracted(ux, uy, uz);
        for (int i = 0; i < Q; ++i) f[i] = is_wall(t) ? nan_v<T>() : feq_contracted(k, i, rho, ux, uy, uz, u2);
        return;
    }
    const T u2 = ((ux * ux) + (uy * uy)) + (uz * uz);
    for (int i = 0; i < Q; ++i) f[i] = is_wall(t) ? nan_v<T>() : feq(k, i, rho, ux, uy, uz, u2);
}

// ---------------------------------------------------------------------------
// Reference (push) form.
// ---------------------------------------------------------------------------
template <class T>
void push_initialize(const Params& p, T* f_stream, T* f_collide, T* rho, T* u, int32_t* map) {
    const int64_t N = p.dim, N3 = N * N * N, S = p.stride;
    const Consts<T> k(p);
#pragma omp parallel for num_threads(p.threads) schedule(static)
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x) {
                const int64_t id = x + y * N + z * N * N;
                const int32_t t = cell_type(N, x, y, z);
                map[id] = t;
                T f[Q], r, ux, uy, uz;
                init_cell(k, t, f, r, ux, uy, uz);
                const bool st = is_store_macro(t);
                rho[id] = st ? r : nan_v<T>();
                u[id] = st ? ux : nan_v<T>();
                u[N3 + id] = st ? uy : nan_v<T>();
                u[2 * N3 + id] = st ? uz : nan_v<T>();
                for (int i = 0; i < Q; ++i) {
                    f_collide[slot(id, i, S)] = f[i];
                    f_stream[slot(id, i, S)] = f[i];
                }
            }
}

// writes (may be null): one counter per slot of f_stream, incremented by every
// store of this step; the push scheme writes each slot at most once
// (kernels.cl:423: slot (d, i) is written only by its source d - e_i).
template <class T>
void push_compute(const Params& p, T* f_stream, const T* f_collide, T* rho, T* u, const int32_t* map,
                  int update_macro, const T* force, int32_t* writes = nullptr) {
    const int64_t N = p.dim, N3 = N * N * N, S = p.stride;
    const Consts<T> k(p);
#pragma omp parallel for num_threads(p.threads) schedule(static)
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x) {
                const int64_t id = x + y * N + z * N * N;
                const int32_t t = map[id];
                T f[Q];
                for (int i = 0; i < Q; ++i) f[i] = f_collide[slot(id, i, S)];
                T r, ux, uy, uz;
                T g[3];
                const T* gp = nullptr;
                if (force) {
                    g[0] = force[id];
                    g[1] = force[N3 + id];
                    g[2] = force[2 * N3 + id];
                    gp = g;
                }
                collide(k, t, f, &r, &ux, &uy, &uz, gp);
                if (update_macro && is_store_macro(t)) {  // kernels.cl:383-388
                    rho[id] = r;
                    u[id] = ux;
                    u[N3 + id] = uy;
                    u[2 * N3 + id] = uz;
                }
                if (is_wall(t)) continue;  // kernels.cl:421
                for (int i = 0; i < Q; ++i) {  // kernels.cl:423
                    const int64_t d = (x + EX[i]) + (y + EY[i]) * N + (z + EZ[i]) * N * N;
                    f_stream[slot(d, i, S)] = f[i];
                    if (writes) {
#pragma omp atomic
                        ++writes[slot(d, i, S)];
                    }
                }
            }
}

// ---------------------------------------------------------------------------
// Storage codecs for the pull form.  Native stores T; the shifted forms store
// (f - w_i) * scale as IEEE binary16 or bfloat16, rounded to nearest even, and
// decode as w_i + h * (1/scale).  scale is a power of two.
// ---------------------------------------------------------------------------
inline uint32_t fbits(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float bitsf(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

uint16_t f32_to_f16_rne(float f) {
    const uint32_t x = fbits(f);
    const uint32_t sign = (x >> 16) & 0x8000u;
    const uint32_t a = x & 0x7fffffffu;
    if (a >= 0x7f800000u)  // inf or NaN
        return uint16_t(sign | 0x7c00u | (a > 0x7f800000u ? (0x200u | ((a >> 13) & 0x3ffu)) : 0u));
    if (a >= 0x477ff000u) return uint16_t(sign | 0x7c00u);  // rounds to infinity
    if (a >= 0x38800000u) {  // normal binary16
        const uint32_t odd = (a >> 13) & 1u;
        return uint16_t(sign | ((a - (112u << 23) + 0xfffu + odd) >> 13));
    }
    const uint32_t e = a >> 23;  // subnormal binary16 (or zero)
    if (e == 0) return uint16_t(sign);
    const uint32_t m = (a & 0x7fffffu) | 0x800000u;
    const uint32_t shift = 126u - e;  // value * 2^24 = m * 2^(e-126)
    if (shift > 24) return uint16_t(sign);
    uint32_t q = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u), halfway = 1u << (shift - 1u);
    if (rem > halfway || (rem == halfway && (q & 1u))) ++q;
    return uint16_t(sign | q);
}

float f16_to_f32(uint16_t h) {
    const uint32_t sign = uint32_t(h & 0x8000u) << 16;
    const uint32_t e = (h >> 10) & 0x1fu, m = h & 0x3ffu;
    if (e == 0x1f) return bitsf(sign | 0x7f800000u | (m << 13));
    if (e != 0) return bitsf(sign | ((e + 112u) << 23) | (m << 13));
    if (m == 0) return bitsf(sign);
    // subnormal: m * 2^-24, exact in binary32
    const float v = float(m) * bitsf(0x33800000u);  // 2^-24
    return sign ? -v : v;
}

uint16_t f32_to_bf16_rne(float f) {
    const uint32_t x = fbits(f);
    if ((x & 0x7fffffffu) > 0x7f800000u) return uint16_t((x >> 16) | 0x40u);
    return uint16_t((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}

float bf16_to_f32(uint16_t h) { return bitsf(uint32_t(h) << 16); }

enum Storage { STORE_NATIVE = 0, STORE_FP16 = 1, STORE_BF16 = 2 };

// Power-of-two scale of the shifted storage for lid velocity U:
// 2^floor(log2(1024 / (0.05 + 2|U|))), clamped to [2^0, 2^24].
double shift_scale(double U) {
    const double a = 0.05 + 2.0 * std::fabs(U);
    int k = int(std::floor(std::log2(1024.0 / a)));
    if (k < 0) k = 0;
    if (k > 24) k = 24;
    return std::ldexp(1.0, k);
}

template <class T>
struct Codec {
    int kind;
    T w[Q];
    float scale, inv_scale;
    Codec(int kind_, const Consts<T>& k, double U) : kind(kind_) {
        for (int i = 0; i < Q; ++i) w[i] = k.w[i];
        const double s = shift_scale(U);
        scale = float(s);
        inv_scale = float(1.0 / s);
    }
    inline T load(const void* base, int64_t s, int i) const {
        if (kind == STORE_NATIVE) return static_cast<const T*>(base)[s];
        const uint16_t h = static_cast<const uint16_t*>(base)[s];
        const float v = (kind == STORE_FP16) ? f16_to_f32(h) : bf16_to_f32(h);
        return T(std::fma(v, inv_scale, float(w[i])));  // = w + v / scale: the product is exact
    }
    // 16-bit storage holds the deviation d = f - w_i scaled: h = RN16(d * scale),
    // d = h * (1/scale); both products are exact (scale is a power of two)
    inline float load_delta(const void* base, int64_t s) const {
        const uint16_t h = static_cast<const uint16_t*>(base)[s];
        return ((kind == STORE_FP16) ? f16_to_f32(h) : bf16_to_f32(h)) * inv_scale;
    }
    inline void store_delta(void* base, int64_t s, float d) const {
        const float v = d * scale;
        static_cast<uint16_t*>(base)[s] = (kind == STORE_FP16) ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
    }
    inline void store(void* base, int64_t s, int i, T f) const {
        if (kind == STORE_NATIVE) {
            static_cast<T*>(base)[s] = f;
            return;
        }
        const float v = std::fma(float(f), scale, -(float(w[i]) * scale));  // = (f - w) * scale, one rounding
        static_cast<uint16_t*>(base)[s] = (kind == STORE_FP16) ? f32_to_f16_rne(v) : f32_to_bf16_rne(v);
    }
};

// ---------------------------------------------------------------------------
// Pull form.  P holds post-collision populations at their source cell.  Both
// buffers start as P0[s][i] = F0[s + e_i][i] (F0 = the reference's initial
// lattice, NaN for walls and out of range), wall cells are never written, and a
// non-wall cell d gathers g_i = P[d - e_i][i].  The gathered g equals the
// reference's f_collide[d][i] at every step, including the stale slots whose
// source is a wall (they keep F0[d][i]).
// ---------------------------------------------------------------------------
template <class T>
void pull_initialize(const Params& p, int storage, void* Pa, void* Pb, T* rho, T* u, int32_t* map) {
    const int64_t N = p.dim, N3 = N * N * N, S = p.stride;
    const Consts<T> k(p);
    const Codec<T> cd(storage, k, p.velocity);
#pragma omp parallel for num_threads(p.threads) schedule(static)
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x) {
                const int64_t id = x + y * N + z * N * N;
                const int32_t t = cell_type(N, x, y, z);
                map[id] = t;
                T f[Q], r, ux, uy, uz;
                init_cell(k, t, f, r, ux, uy, uz);
                const bool st = is_store_macro(t);
                rho[id] = st ? r : nan_v<T>();
                u[id] = st ? ux : nan_v<T>();
                u[N3 + id] = st ? uy : nan_v<T>();
                u[2 * N3 + id] = st ? uz : nan_v<T>();
                float d[Q] = {};
                if constexpr (std::is_same<T, float>::value) {
                    if (storage != STORE_NATIVE) init_cell_delta(k, t, d);
                }
                for (int i = 0; i < Q; ++i) {
                    const int64_t sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
                    if (sx < 0 || sy < 0 || sz < 0 || sx >= N || sy >= N || sz >= N) continue;
                    const int64_t s = slot(sx + sy * N + sz * N * N, i, S);
                    if (storage != STORE_NATIVE) {
                        cd.store_delta(Pa, s, d[i]);
                        cd.store_delta(Pb, s, d[i]);
                        continue;
                    }
                    cd.store(Pa, s, i, f[i]);
                    cd.store(Pb, s, i, f[i]);
                }
            }
    // slots whose destination is outside the lattice: NaN
#pragma omp parallel for num_threads(p.threads) schedule(static)
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x)
                for (int i = 0; i < Q; ++i) {
                    const int64_t dx = x + EX[i], dy = y + EY[i], dz = z + EZ[i];
                    if (dx >= 0 && dy >= 0 && dz >= 0 && dx < N && dy < N && dz < N) continue;
                    const int64_t s = slot(x + y * N + z * N * N, i, S);
                    if (storage != STORE_NATIVE) {
                        cd.store_delta(Pa, s, std::numeric_limits<float>::quiet_NaN());
                        cd.store_delta(Pb, s, std::numeric_limits<float>::quiet_NaN());
                        continue;
                    }
                    cd.store(Pa, s, i, nan_v<T>());
                    cd.store(Pb, s, i, nan_v<T>());
                }
}

// writes (may be null): per-slot store counters of Pout, as in push_compute.
template <class T>
void pull_compute(const Params& p, int storage, const void* Pin, void* Pout, T* rho, T* u, int update_macro,
                  const T* force, double* diag, int32_t* writes = nullptr) {
    const int64_t N = p.dim, N3 = N * N * N, S = p.stride;
    const Consts<T> k(p);
    const Codec<T> cd(storage, k, p.velocity);
    double mass = 0.0, ke = 0.0;
#pragma omp parallel for num_threads(p.threads) schedule(static) reduction(+ : mass, ke)
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x) {
                const int64_t id = x + y * N + z * N * N;
                const int32_t t = cell_type(N, x, y, z);
                if (is_wall(t)) continue;
                T f[Q];
                T r, ux, uy, uz, g[3];
                const T* gp = nullptr;
                if (force) {
                    g[0] = force[id];
                    g[1] = force[N3 + id];
                    g[2] = force[2 * N3 + id];
                    gp = g;
                }
                if constexpr (std::is_same<T, float>::value) {
                    if (storage != STORE_NATIVE) {  // deviation form
                        for (int i = 0; i < Q; ++i) {
                            const int64_t sid = (x - EX[i]) + (y - EY[i]) * N + (z - EZ[i]) * N * N;
                            f[i] = cd.load_delta(Pin, slot(sid, i, S));
                        }
                        collide_delta(k, t, f, &r, &ux, &uy, &uz, gp);
                    }
                }
                if (storage == STORE_NATIVE) {
                    for (int i = 0; i < Q; ++i) {
                        const int64_t sid = (x - EX[i]) + (y - EY[i]) * N + (z - EZ[i]) * N * N;
                        f[i] = cd.load(Pin, slot(sid, i, S), i);
                    }
                    collide(k, t, f, &r, &ux, &uy, &uz, gp);
                }
                if (is_store_macro(t)) {
                    if (update_macro) {
                        rho[id] = r;
                        u[id] = ux;
                        u[N3 + id] = uy;
                        u[2 * N3 + id] = uz;
                    }
                    const double dr = double(r), dx = double(ux), dy = double(uy), dz = double(uz);
                    mass += dr;
                    ke += 0.5 * dr * (((dx * dx) + (dy * dy)) + (dz * dz));
                }
                for (int i = 0; i < Q; ++i) {
                    if (storage != STORE_NATIVE) cd.store_delta(Pout, slot(id, i, S), float(f[i]));
                    else cd.store(Pout, slot(id, i, S), i, f[i]);
                    if (writes) {
#pragma omp atomic
                        ++writes[slot(id, i, S)];
                    }
                }
            }
    if (diag) {
        diag[0] = mass;
        diag[1] = ke;
    }
}

// The reference's CSoA lattice as it would be after the step that wrote P
// (push view): view[d][i] = P[d - e_i][i], NaN where d - e_i is outside.
template <class T>
void pull_view(const Params& p, int storage, const void* P, T* out) {
    const int64_t N = p.dim, S = p.stride;
    const Consts<T> k(p);
    const Codec<T> cd(storage, k, p.velocity);
#pragma omp parallel for num_threads(p.threads) schedule(static)
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x)
                for (int i = 0; i < Q; ++i) {
                    const int64_t id = x + y * N + z * N * N;
                    const int64_t sx = x - EX[i], sy = y - EY[i], sz = z - EZ[i];
                    T v = nan_v<T>();
                    if (sx >= 0 && sy >= 0 && sz >= 0 && sx < N && sy < N && sz < N)
                        v = cd.load(P, slot(sx + sy * N + sz * N * N, i, S), i);
                    out[slot(id, i, S)] = v;
                }
}

// Mass and kinetic energy of the stored macroscopic fields, in FP64, summed in
// cell order.
template <class T>
void diagnostics(const Params& p, const T* rho, const T* u, double* out) {
    const int64_t N = p.dim, N3 = N * N * N;
    double mass = 0.0, ke = 0.0;
    for (int64_t z = 0; z < N; ++z)
        for (int64_t y = 0; y < N; ++y)
            for (int64_t x = 0; x < N; ++x) {
                const int64_t id = x + y * N + z * N * N;
                if (!is_store_macro(cell_type(N, x, y, z))) continue;
                const double r = rho[id], a = u[id], b = u[N3 + id], c = u[2 * N3 + id];
                mass += r;
                ke += 0.5 * r * (((a * a) + (b * b)) + (c * c));
            }
    out[0] = mass;
    out[1] = ke;
}

template <class T>
void constants(const Params& p, T* out) {
    const Consts<T> k(p);
    out[0] = k.inv_tau;
    out[1] = k.U;
    for (int i = 0; i < Q; ++i) out[2 + i] = k.w[i];
}

}  // namespace

extern "C" {

typedef Params lbmo_params;

int lbmo_cell_type(int32_t N, int32_t x, int32_t y, int32_t z) { return cell_type(N, x, y, z); }
int64_t lbmo_slot(int64_t id, int32_t q, int64_t stride) { return slot(id, q, stride); }
double lbmo_shift_scale(double U) { return shift_scale(U); }

void lbmo_constants_f32(const lbmo_params* p, float* out) { constants<float>(*p, out); }
void lbmo_constants_f64(const lbmo_params* p, double* out) { constants<double>(*p, out); }

void lbmo_initialize_f32(const lbmo_params* p, float* fs, float* fc, float* rho, float* u, int32_t* map) {
    push_initialize<float>(*p, fs, fc, rho, u, map);
}
void lbmo_initialize_f64(const lbmo_params* p, double* fs, double* fc, double* rho, double* u, int32_t* map) {
    push_initialize<double>(*p, fs, fc, rho, u, map);
}
void lbmo_compute_f32(const lbmo_params* p, float* fs, const float* fc, float* rho, float* u, const int32_t* map,
                      int32_t update_macro, const float* force) {
    push_compute<float>(*p, fs, fc, rho, u, map, update_macro, force);
}
void lbmo_compute_f64(const lbmo_params* p, double* fs, const double* fc, double* rho, double* u,
                      const int32_t* map, int32_t update_macro, const double* force) {
    push_compute<double>(*p, fs, fc, rho, u, map, update_macro, force);
}

// One push / pull step in FP64 that also counts the stores into each slot of
// the written lattice (writes: 19 N^3 counters, added to).
void lbmo_compute_counted_f64(const lbmo_params* p, double* fs, const double* fc, double* rho, double* u,
                              const int32_t* map, int32_t* writes) {
    push_compute<double>(*p, fs, fc, rho, u, map, 0, nullptr, writes);
}
void lbmo_pull_compute_counted_f64(const lbmo_params* p, const void* pin, void* pout, double* rho, double* u,
                                   int32_t* writes) {
    pull_compute<double>(*p, STORE_NATIVE, pin, pout, rho, u, 0, nullptr, nullptr, writes);
}

void lbmo_pull_initialize_f32(const lbmo_params* p, int32_t storage, void* pa, void* pb, float* rho, float* u,
                              int32_t* map) {
    pull_initialize<float>(*p, storage, pa, pb, rho, u, map);
}
void lbmo_pull_initialize_f64(const lbmo_params* p, void* pa, void* pb, double* rho, double* u, int32_t* map) {
    pull_initialize<double>(*p, STORE_NATIVE, pa, pb, rho, u, map);
}
void lbmo_pull_compute_f32(const lbmo_params* p, int32_t storage, const void* pin, void* pout, float* rho,
                           float* u, int32_t update_macro, const float* force, double* diag) {
    pull_compute<float>(*p, storage, pin, pout, rho, u, update_macro, force, diag);
}
void lbmo_pull_compute_f64(const lbmo_params* p, const void* pin, void* pout, double* rho, double* u,
                           int32_t update_macro, const double* force, double* diag) {
    pull_compute<double>(*p, STORE_NATIVE, pin, pout, rho, u, update_macro, force, diag);
}
void lbmo_pull_view_f32(const lbmo_params* p, int32_t storage, const void* P, float* out) {
    pull_view<float>(*p, storage, P, out);
}
void lbmo_pull_view_f64(const lbmo_params* p, const void* P, double* out) {
    pull_view<double>(*p, STORE_NATIVE, P, out);
}

void lbmo_diagnostics_f32(const lbmo_params* p, const float* rho, const float* u, double* out) {
    diagnostics<float>(*p, rho, u, out);
}
void lbmo_diagnostics_f64(const lbmo_params* p, const double* rho, const double* u, double* out) {
    diagnostics<double>(*p, rho, u, out);
}

// One cell's local step in two forms, for pinning the deviation form against
// the full-value form: f (19 populations) through collide in FP64 (exact
// expressions) into out_full, and d = f - w_i through collide_delta (FP32) into
// out_delta; force may be null.  out_macro: rho, ux, uy, uz of each form.
void lbmo_collide_pair(const lbmo_params* p, int32_t t, const double* f, const float* force, double* out_full,
                       float* out_delta, double* out_macro) {
    Params q = *p;
    q.fast = 0;
    const Consts<double> kd(q);
    const Consts<float> kf(q);
    double g[3] = {0, 0, 0};
    if (force) for (int j = 0; j < 3; ++j) g[j] = force[j];
    for (int i = 0; i < Q; ++i) out_full[i] = f[i];
    double r, ux, uy, uz;
    collide(kd, t, out_full, &r, &ux, &uy, &uz, force ? g : nullptr);
    for (int i = 0; i < Q; ++i) out_delta[i] = float(f[i] - double(kf.w[i]));
    float rf, xf, yf, zf;
    collide_delta(kf, t, out_delta, &rf, &xf, &yf, &zf, force);
    const double m[8] = {r, ux, uy, uz, rf, xf, yf, zf};
    for (int j = 0; j < 8; ++j) out_macro[j] = m[j];
}

// Element-wise codec helpers (for pinning the conversions against numpy/torch).
void lbmo_f32_to_f16(const float* in, uint16_t* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = f32_to_f16_rne(in[i]);
}
void lbmo_f16_to_f32(const uint16_t* in, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = f16_to_f32(in[i]);
}
void lbmo_f32_to_bf16(const float* in, uint16_t* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = f32_to_bf16_rne(in[i]);
}
void lbmo_bf16_to_f32(const uint16_t* in, float* out, int64_t n) {
    for (int64_t i = 0; i < n; ++i) out[i] = bf16_to_f32(in[i]);
}

}  // extern "C"
