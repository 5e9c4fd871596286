// lbm_kernels.hpp -- gfx950 kernels of the D3Q19 lid-driven cavity.
//
// Pull ("gather") form of the reference's push kernel (kernels.cl:321-426):
// lattice P holds post-collision populations at their source cell, so a cell d
// reads g_i = P[d - e_i][i] (misaligned by at most one element along x) and
// writes its own 19 slots aligned.  Both lattices start as
// P0[s][i] = F0[s + e_i][i] (F0 = the reference's initial lattice) and wall
// cells are never written, so a slot whose source is a wall keeps supplying
// the reference's stale initial value (kernels.cl:293-317, 421).  The gathered
// populations, and hence rho, u and every later state, equal the reference's.
//
// Three forms of the step with identical results: the vector row kernel (each
// lane owns 8 or 16 bytes of every population, i.e. 1-8 consecutive cells),
// the row kernel and the generic kernel (one thread per cell).  HBM-bound (152
// B/cell/step in FP32, 304 in FP64, 76 with 16-bit shifted storage); no MFMA
// -- nothing here is a contraction.  Everything is compiled without
// floating-point contraction, so each value is the reference's expression
// evaluated operation by operation; the fused collision (-o, and 16-bit
// storage) is a different expression that uses fma explicitly.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

#include "lbm_lattice.hpp"

#pragma clang fp contract(off)

// Loads and stores of the lattices in the vector and row-pipeline kernels:
// streaming (nontemporal) by default -- every slot is read and written once
// per step.  A build can define LBM_LATTICE_CACHED to compare cached accesses.
#if defined(LBM_LATTICE_CACHED) || defined(LBM_LATTICE_CACHED_LOADS)
#define LBM_LAT_LOAD(p) (*(p))
#else
#define LBM_LAT_LOAD(p) __builtin_nontemporal_load(p)
#endif
#if defined(LBM_LATTICE_CACHED) || defined(LBM_LATTICE_CACHED_STORES)
#define LBM_LAT_STORE(v, p) (*(p) = (v))
#else
#define LBM_LAT_STORE(v, p) __builtin_nontemporal_store(v, p)
#endif

namespace lbm {

// ---------------------------------------------------------------------------
// Storage codecs.  Native stores T.  The shifted forms store
// h = (f - w_i) * scale rounded to nearest even in binary16 or bfloat16 and
// decode f = w_i + h * (1/scale); scale is a power of two, so both products are
// exact and only the subtraction/addition and the 16-bit rounding round.
// ---------------------------------------------------------------------------
enum StorageKind { ST_NATIVE = 0, ST_FP16 = 1, ST_BF16 = 2 };

template <class T, int KIND> struct Codec;

template <class T> struct Codec<T, ST_NATIVE> {
    using S = T;
    float scale, inv_scale;
    __device__ __forceinline__ T decode(S s, T) const { return s; }
    __device__ __forceinline__ S encode(T v, T) const { return v; }
    __device__ __forceinline__ T load(const S* p, T w) const { return decode(*p, w); }
    __device__ __forceinline__ void store(S* p, T w, T v) const { __builtin_nontemporal_store(encode(v, w), p); }
};

template <> struct Codec<float, ST_FP16> {
    using S = _Float16;
    float scale, inv_scale;
    // h * inv_scale is exact, so the fma equals w + h * inv_scale; v * scale - w * scale
    // is (v - w) * scale exactly before its single rounding (scale is a power of two)
    __device__ __forceinline__ float decode(S s, float w) const { return __builtin_fmaf(float(s), inv_scale, w); }
    __device__ __forceinline__ S encode(float v, float w) const {
        return _Float16(__builtin_fmaf(v, scale, -(w * scale)));
    }
    __device__ __forceinline__ float load(const S* p, float w) const { return decode(*p, w); }
    __device__ __forceinline__ void store(S* p, float w, float v) const { __builtin_nontemporal_store(encode(v, w), p); }
    // deviation form: d = f - w_i itself, h = RN16(d * scale), d = h * (1/scale) (both products exact)
    __device__ __forceinline__ float decode_delta(S s) const { return float(s) * inv_scale; }
    __device__ __forceinline__ S encode_delta(float d) const { return _Float16(d * scale); }
    // the collision's own domain: the scaled deviation d * scale, exactly the stored value
    __device__ __forceinline__ float raw(S s) const { return float(s); }
    __device__ __forceinline__ S from_raw(float v) const { return _Float16(v); }
};

template <> struct Codec<float, ST_BF16> {
    using S = uint16_t;
    float scale, inv_scale;
    __device__ __forceinline__ float decode(S s, float w) const {
        return __builtin_fmaf(__uint_as_float(uint32_t(s) << 16), inv_scale, w);
    }
    __device__ __forceinline__ S encode(float v, float w) const {
        // bfloat16 conversion rounds to nearest even (a NaN stays a NaN)
        const __bf16 h = __bf16(__builtin_fmaf(v, scale, -(w * scale)));
        return __builtin_bit_cast(uint16_t, h);
    }
    __device__ __forceinline__ float load(const S* p, float w) const { return decode(*p, w); }
    __device__ __forceinline__ void store(S* p, float w, float v) const { __builtin_nontemporal_store(encode(v, w), p); }
    __device__ __forceinline__ float decode_delta(S s) const { return __uint_as_float(uint32_t(s) << 16) * inv_scale; }
    __device__ __forceinline__ S encode_delta(float d) const { return __builtin_bit_cast(uint16_t, __bf16(d * scale)); }
    __device__ __forceinline__ float raw(S s) const { return __uint_as_float(uint32_t(s) << 16); }
    __device__ __forceinline__ S from_raw(float v) const { return __builtin_bit_cast(uint16_t, __bf16(v)); }
};

// ---------------------------------------------------------------------------
// Lattice constants in the arithmetic type (kernels.cl:110-128; with
// -cl-single-precision-constant the reference folds them in float for FP32).
// ---------------------------------------------------------------------------
template <class T> struct W {
    static constexpr T w0 = T(1) / T(3);
    static constexpr T w1 = T(1) / T(18);
    static constexpr T w7 = T(1) / T(36);
};

__host__ __device__ constexpr int ex_of(int i) {
    return (i == 1 || i == 7 || i == 10 || i == 11 || i == 15) ? 1
           : (i == 3 || i == 8 || i == 9 || i == 13 || i == 17) ? -1
                                                                 : 0;
}
__host__ __device__ constexpr int ey_of(int i) {
    return (i == 2 || i == 7 || i == 8 || i == 12 || i == 16) ? 1
           : (i == 4 || i == 9 || i == 10 || i == 14 || i == 18) ? -1
                                                                  : 0;
}
__host__ __device__ constexpr int ez_of(int i) {
    return (i == 6 || i == 15 || i == 16 || i == 17 || i == 18) ? 1
           : (i == 5 || i == 11 || i == 12 || i == 13 || i == 14) ? -1
                                                                   : 0;
}
template <class T> __host__ __device__ __forceinline__ T w_of(int i) {
    return i == 0 ? W<T>::w0 : (i <= 6 ? W<T>::w1 : W<T>::w7);
}

template <class T> struct StepConsts {
    T inv_tau;  // 1 / (3 nu + 1/2) folded in T (kernels.cl:61-62)
    T U;        // lid velocity literal (kernels.cl:57)
    T om1;      // 1 - inv_tau, for the fused collision
    float dscale, dinv;  // deviation form: storage scale (a power of two) and its inverse
};

// e_i . u with the reference's ((ux ex + uy ey) + uz ez) evaluated exactly:
// products with 0 contribute +-0 and products with -1 are negations, so the
// expression reduces to the signed sum of the non-zero components in order.
template <class T> __host__ __device__ __forceinline__ T edot(int i, T ux, T uy, T uz) {
    switch (i) {
        case 0: return T(0);
        case 1: return ux;
        case 2: return uy;
        case 3: return -ux;
        case 4: return -uy;
        case 5: return -uz;
        case 6: return uz;
        case 7: return ux + uy;
        case 8: return uy - ux;
        case 9: return -(ux + uy);
        case 10: return ux - uy;
        case 11: return ux - uz;
        case 12: return uy - uz;
        case 13: return -(ux + uz);
        case 14: return -(uy + uz);
        case 15: return ux + uz;
        case 16: return uy + uz;
        case 17: return uz - ux;
        default: return uz - uy;
    }
}

// (rho w_i) (1 + 3 eu + 4.5 eu eu - 1.5 u2), left to right (kernels.cl:397, 416).
template <class T> __host__ __device__ __forceinline__ T feq(int i, T rho, T ux, T uy, T uz, T u2) {
    const T eu = edot<T>(i, ux, uy, uz);
    return (rho * w_of<T>(i)) * (((T(1) + (T(3) * eu)) + ((T(4.5) * eu) * eu)) - (T(1.5) * u2));
}

// The same expressions as clang compiles OpenCL C, whose default is
// FP_CONTRACT ON: within one expression x + a*b and x - a*b become fma(a, b, x)
// and fma(-a, b, x), the left operand's product taken when both operands are
// products.  In kernels.cl that touches u2, the equilibrium and the BGK
// update; e_i . u multiplies by 0 and +-1 only, so it is exact in both forms.
template <class T> __host__ __device__ __forceinline__ T fma_any(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}
// ((ux ux) + (uy uy)) + (uz uz) (kernels.cl:390)
template <class T> __host__ __device__ __forceinline__ T u2_contracted(T ux, T uy, T uz) {
    return fma_any<T>(uz, uz, fma_any<T>(ux, ux, uy * uy));
}
// (rho w_i) (((1 + 3 eu) + 4.5 eu eu) - 1.5 u2) (kernels.cl:397, 416)
template <class T> __host__ __device__ __forceinline__ T feq_contracted(int i, T rho, T ux, T uy, T uz, T u2) {
    const T eu = edot<T>(i, ux, uy, uz);
    const T a = fma_any<T>(T(3), eu, T(1));
    const T b = fma_any<T>(T(4.5) * eu, eu, a);
    return (rho * w_of<T>(i)) * fma_any<T>(-T(1.5), u2, b);
}

// Arithmetic forms of the local step: the reference's expressions evaluated
// operation by operation (IEEE, no contraction; the CPU oracle bit for bit),
// the fused collision (-o, 16-bit storage), and the reference's expressions as
// ROCm's OpenCL compiler evaluates them (contracted as above, FP32 division to
// OpenCL's 2.5 ulp; those kernels live in lbm_opencl_arith.hip, the one
// translation unit built with that division precision).
enum Arith { AR_EXACT = 0, AR_FUSED = 1, AR_OPENCL = 2, AR_DELTA = 3 };

// Lid fix-up and macroscopic values of a non-wall cell (kernels.cl:343-379):
// on return f holds the fixed-up populations, rho and u the moments.
template <class T>
__device__ __forceinline__ void moments(int32_t t, T (&f)[Q], T& rho, T& ux, T& uy, T& uz,
                                        const StepConsts<T>& k) {
    const bool moving = is_moving(t);
    if (moving) {  // kernels.cl:343-349
        f[5] = f[6];
        f[11] = f[17];
        f[12] = f[18];
        f[13] = f[15];
        f[14] = f[16];
    }
    rho = f[0];  // kernels.cl:355
#pragma unroll
    for (int i = 1; i < Q; ++i) rho = rho + f[i];
    if (moving) {  // kernels.cl:362-365
        ux = k.U;
        uy = T(0);
        uz = T(0);
    } else {  // kernels.cl:376-378
        const T jx = ((((f[1] + f[7]) + f[10]) + f[11]) + f[15]) - ((((f[3] + f[8]) + f[9]) + f[13]) + f[17]);
        const T jy = ((((f[2] + f[7]) + f[8]) + f[12]) + f[16]) - ((((f[4] + f[9]) + f[10]) + f[14]) + f[18]);
        const T jz = ((((f[6] + f[15]) + f[16]) + f[17]) + f[18]) - ((((f[5] + f[11]) + f[12]) + f[13]) + f[14]);
        ux = jx / rho;
        uy = jy / rho;
        uz = jz / rho;
    }
}

// Half-way bounce-back: swap opposite populations (kernels.cl:400-408).
template <class T> __device__ __forceinline__ void bounce_back(T (&f)[Q]) {
    T tmp;
    tmp = f[1]; f[1] = f[3]; f[3] = tmp;
    tmp = f[2]; f[2] = f[4]; f[4] = tmp;
    tmp = f[5]; f[5] = f[6]; f[6] = tmp;
    tmp = f[7]; f[7] = f[9]; f[9] = tmp;
    tmp = f[8]; f[8] = f[10]; f[10] = tmp;
    tmp = f[11]; f[11] = f[17]; f[17] = tmp;
    tmp = f[12]; f[12] = f[18]; f[18] = tmp;
    tmp = f[13]; f[13] = f[15]; f[15] = tmp;
    tmp = f[14]; f[14] = f[16]; f[16] = tmp;
}

// Local part of `compute` (kernels.cl:343-418) for a non-wall cell of type t.
// On entry f holds the gathered populations, on return the ones to store.
template <class T>
__device__ __forceinline__ void collide(int32_t t, T (&f)[Q], T& rho, T& ux, T& uy, T& uz,
                                        const StepConsts<T>& k) {
    moments<T>(t, f, rho, ux, uy, uz, k);
    const T u2 = ((ux * ux) + (uy * uy)) + (uz * uz);  // kernels.cl:390
    if (is_moving(t)) {
        // kernels.cl:393-398 sets f_i = feq_i; the BGK step that follows
        // (kernels.cl:412-418) then adds inv_tau * (feq_i - f_i) = 0 exactly.
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = feq<T>(i, rho, ux, uy, uz, u2);
    } else if (t == FLUID) {  // kernels.cl:412-418, 270-273
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T e = feq<T>(i, rho, ux, uy, uz, u2);
            f[i] = f[i] + k.inv_tau * (e - f[i]);
        }
    } else if (is_bounceback(t)) {
        bounce_back<T>(f);
    }
    // CORNER: pure propagation (kernels.cl:254-260 with no BC and no collision).
}

// collide() with the contractions of ROCm's OpenCL compiler (see
// feq_contracted); the divisions round as the translation unit compiles them.
template <class T>
__device__ __forceinline__ void collide_opencl(int32_t t, T (&f)[Q], T& rho, T& ux, T& uy, T& uz,
                                               const StepConsts<T>& k) {
    moments<T>(t, f, rho, ux, uy, uz, k);
    const T u2 = u2_contracted<T>(ux, uy, uz);
    if (is_moving(t)) {
        // f_i = feq_i, then f_i + inv_tau (feq_i - f_i) = fma(inv_tau, 0, feq_i) = feq_i
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = feq_contracted<T>(i, rho, ux, uy, uz, u2);
    } else if (t == FLUID) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const T e = feq_contracted<T>(i, rho, ux, uy, uz, u2);
            f[i] = fma_any<T>(k.inv_tau, e - f[i], f[i]);  // kernels.cl:270-273
        }
    } else if (is_bounceback(t)) {
        bounce_back<T>(f);
    }
}

// Fused form of the same local step, used by the 16-bit storage forms and by
// the native forms when fast math is requested (-o).  Moments, macroscopic
// values and boundary handling are those of collide(); the BGK update is
// rewritten as f' = (1 - w) f + (w rho w_i) T_i with
//   T_i = K + e (3 + 4.5 e),  K = 1 - 1.5 u2,  e = e_i . u,
// built with fused multiply-adds and shared between opposite populations
// (T for -e is A - 3e with A = 4.5 e e + K).  Every operation is a single IEEE
// rounding, so the CPU oracle reproduces it with std::fma.
template <class T> __device__ __forceinline__ T fmaT(T a, T b, T c) {
    if constexpr (sizeof(T) == 4) return __builtin_fmaf(a, b, c);
    else return __builtin_fma(a, b, c);
}

template <class T>
__device__ __forceinline__ void collide_fast(int32_t t, T (&f)[Q], T& rho, T& ux, T& uy, T& uz,
                                             const StepConsts<T>& k) {
    const bool moving = is_moving(t);
    moments<T>(t, f, rho, ux, uy, uz, k);
    if (!moving && t != FLUID) {
        if (is_bounceback(t)) bounce_back<T>(f);
        return;
    }
    const T u2 = fmaT<T>(uz, uz, fmaT<T>(uy, uy, ux * ux));
    const T K = fmaT<T>(T(-1.5), u2, T(1));
    // moving: f_i = (rho w_i) T_i;  fluid: f_i = (inv_tau rho w_i) T_i + om1 f_i
    const T s = moving ? rho : k.inv_tau * rho;
    const T c0 = s * W<T>::w0, c1 = s * W<T>::w1, c7 = s * W<T>::w7;
    const T keep = moving ? T(0) : k.om1;
    constexpr int PI[9] = {1, 2, 6, 7, 10, 15, 16, 11, 12};    // e_i
    constexpr int PO[9] = {3, 4, 5, 9, 8, 13, 14, 17, 18};     // -e_i
    f[0] = fmaT<T>(c0, K, keep * f[0]);
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int i = PI[p], o = PO[p];
        const T e = edot<T>(i, ux, uy, uz);
        const T A = fmaT<T>(T(4.5) * e, e, K);
        const T B = T(3) * e;
        const T c = (i <= 6) ? c1 : c7;
        f[i] = fmaT<T>(c, A + B, keep * f[i]);
        f[o] = fmaT<T>(c, A - B, keep * f[o]);
    }
}

// Body force on FLUID cells: f_i += 3 w_i (e_i . g) after the collision (a
// first-order forcing term; no counterpart in the reference, and a zero g
// leaves every value unchanged).
template <class T>
__device__ __forceinline__ void add_force(int32_t t, T (&f)[Q], T gx, T gy, T gz) {
    if (t != FLUID) return;
#pragma unroll
    for (int i = 0; i < Q; ++i) f[i] = f[i] + (T(3) * w_of<T>(i)) * edot<T>(i, gx, gy, gz);
}

// ---------------------------------------------------------------------------
// Deviation ("delta") form of the local step, the arithmetic of the 16-bit
// storage forms (no counterpart in the reference).  The populations stay
// deviations d_i = f_i - w_i from load to store -- what the 16-bit lattice
// holds, scaled -- so no value near w_i is ever rounded in FP32.  (Rounding
// f_i itself, whose low bits all cells of a smooth flow share with w_i,
// gives correlated errors: with 16-bit-decoded inputs the total mass then
// drifted by ~3e-7 of itself per step.)  Moments:
//   drho = d_0 + ... + d_18 (rho = 1 + drho),  j = sum e_i d_i (sum e_i w_i = 0);
// the lid fix-up and bounce-back move d_i as the reference moves f_i; BGK
// f' = (1 - w) f + w rho w_i T_i becomes, with tau_i = T_i - 1,
//   d_i' = (w w_i) fma(rho, tau_i, drho) + (1 - w) d_i,
//   tau_i = Km1 + e (3 + 4.5 e),  Km1 = -1.5 u2,  e = e_i . u,
// tau shared between opposite populations as A +- 3e, A = 4.5 e e + Km1; the
// lid sets d_i = w_i fma(rho, tau_i, drho).  The CPU oracle evaluates the same
// operations in the same order.
//
// The kernels keep the deviations scaled, h_i = d_i * scale -- the stored
// values as they are -- and fold the power-of-two scale into the constants:
// drho = (sum h_i) / scale, j = (sum e_i h_i) / scale, h_i' = fma(cw_i scale,
// fma(rho, tau_i, drho), (1 - w) h_i).  Multiplying by a power of two commutes
// with every rounding here (all values are far from FP32's subnormal and
// overflow ranges), so these are the oracle's unscaled results times scale,
// bit for bit, at two multiplications per population less.
// ---------------------------------------------------------------------------
// d_i <- fma(cw_i, fma(rho, tau_i, drho), keep * d_i); cw_i = c0, c1, c7 by |e_i|^2
__host__ __device__ __forceinline__ void delta_relax(float rho, float drho, float ux, float uy, float uz, float c0,
                                                     float c1, float c7, float keep, float (&d)[Q]) {
    const float u2 = __builtin_fmaf(uz, uz, __builtin_fmaf(uy, uy, ux * ux));
    const float Km1 = -1.5f * u2;
    constexpr int PI[9] = {1, 2, 6, 7, 10, 15, 16, 11, 12};  // e_i
    constexpr int PO[9] = {3, 4, 5, 9, 8, 13, 14, 17, 18};   // -e_i
    d[0] = __builtin_fmaf(c0, __builtin_fmaf(rho, Km1, drho), keep * d[0]);
#pragma unroll
    for (int p = 0; p < 9; ++p) {
        const int i = PI[p], o = PO[p];
        const float e = edot<float>(i, ux, uy, uz);
        const float A = __builtin_fmaf(4.5f * e, e, Km1);
        const float B = 3.0f * e;
        const float c = (i <= 6) ? c1 : c7;
        d[i] = __builtin_fmaf(c, __builtin_fmaf(rho, A + B, drho), keep * d[i]);
        d[o] = __builtin_fmaf(c, __builtin_fmaf(rho, A - B, drho), keep * d[o]);
    }
}

// Initial deviations of a non-wall cell with u = (U, 0, 0) (FRONT cells):
// w_i tau_i at rho = 1; other cells start at d = 0.
__host__ __device__ __forceinline__ void delta_lid_initial(float U, float (&d)[Q]) {
#pragma unroll
    for (int i = 0; i < Q; ++i) d[i] = 0.0f;
    delta_relax(1.0f, 0.0f, U, 0.0f, 0.0f, W<float>::w0, W<float>::w1, W<float>::w7, 0.0f, d);
}

template <class T>
__device__ __forceinline__ void collide_delta(int32_t t, T (&d)[Q], T& rho, T& ux, T& uy, T& uz,
                                              const StepConsts<T>& k) {
    static_assert(sizeof(T) == 4, "the deviation form computes in FP32");
    const bool moving = is_moving(t);
    if (moving) {  // kernels.cl:343-349
        d[5] = d[6];
        d[11] = d[17];
        d[12] = d[18];
        d[13] = d[15];
        d[14] = d[16];
    }
    T hrho = d[0];  // d[] holds scaled deviations h_i = d_i * scale
#pragma unroll
    for (int i = 1; i < Q; ++i) hrho = hrho + d[i];
    const T drho = hrho * k.dinv;
    rho = T(1) + drho;
    if (moving) {  // kernels.cl:362-365
        ux = k.U;
        uy = T(0);
        uz = T(0);
    } else {  // kernels.cl:376-378 on the deviations
        const T jx = ((((d[1] + d[7]) + d[10]) + d[11]) + d[15]) - ((((d[3] + d[8]) + d[9]) + d[13]) + d[17]);
        const T jy = ((((d[2] + d[7]) + d[8]) + d[12]) + d[16]) - ((((d[4] + d[9]) + d[10]) + d[14]) + d[18]);
        const T jz = ((((d[6] + d[15]) + d[16]) + d[17]) + d[18]) - ((((d[5] + d[11]) + d[12]) + d[13]) + d[14]);
        ux = (jx * k.dinv) / rho;
        uy = (jy * k.dinv) / rho;
        uz = (jz * k.dinv) / rho;
    }
    if (!moving && t != FLUID) {
        if (is_bounceback(t)) bounce_back<T>(d);
        return;  // CORNER: pure propagation
    }
    if (moving)
        delta_relax(rho, drho, ux, uy, uz, W<T>::w0 * k.dscale, W<T>::w1 * k.dscale, W<T>::w7 * k.dscale, T(0), d);
    else
        delta_relax(rho, drho, ux, uy, uz, (k.inv_tau * W<T>::w0) * k.dscale, (k.inv_tau * W<T>::w1) * k.dscale,
                    (k.inv_tau * W<T>::w7) * k.dscale, k.om1, d);
}

template <class T, int ARITH>
__device__ __forceinline__ void collide_any(int32_t t, T (&f)[Q], T& rho, T& ux, T& uy, T& uz,
                                            const StepConsts<T>& k) {
    if constexpr (ARITH == AR_FUSED) collide_fast<T>(t, f, rho, ux, uy, uz, k);
    else if constexpr (ARITH == AR_OPENCL) collide_opencl<T>(t, f, rho, ux, uy, uz, k);
    else if constexpr (ARITH == AR_DELTA) collide_delta<T>(t, f, rho, ux, uy, uz, k);
    else collide<T>(t, f, rho, ux, uy, uz, k);
}

// Storage <-> collision values: the population, or in the deviation form its
// scaled deviation from w_i (the stored value itself).
template <int ARITH, class C, class S, class T>
__device__ __forceinline__ T value_in(const C& codec, S s, T w) {
    if constexpr (ARITH == AR_DELTA) return codec.raw(s);
    else return codec.decode(s, w);
}
template <int ARITH, class C, class T>
__device__ __forceinline__ typename C::S value_out(const C& codec, T v, T w) {
    if constexpr (ARITH == AR_DELTA) return codec.from_raw(v);
    else return codec.encode(v, w);
}

// Body force on collision values: 3 w_i (e_i . g), scaled in the deviation form.
template <class T, int ARITH>
__device__ __forceinline__ void add_force_any(int32_t t, T (&f)[Q], T gx, T gy, T gz, const StepConsts<T>& k) {
    if constexpr (ARITH == AR_DELTA) add_force<T>(t, f, gx * k.dscale, gy * k.dscale, gz * k.dscale);
    else add_force<T>(t, f, gx, gy, gz);
}

// ---------------------------------------------------------------------------
// Block reduction of (mass, kinetic energy) in FP64: wavefront shuffles, then
// LDS across the block's wavefronts.  Result valid in thread 0.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void block_reduce2(double& a, double& b, double* lds) {
    if (blockDim.x & 63u) {
        // block sizes that are not whole wavefronts: serial sum through LDS
        __syncthreads();
        lds[2 * threadIdx.x] = a;
        lds[2 * threadIdx.x + 1] = b;
        __syncthreads();
        if (threadIdx.x == 0) {
            double sa = lds[0], sb = lds[1];
            for (uint32_t j = 1; j < blockDim.x; ++j) {
                sa += lds[2 * j];
                sb += lds[2 * j + 1];
            }
            a = sa;
            b = sb;
        }
        return;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = (blockDim.x + 63) >> 6;
    if (lane == 0) {
        lds[2 * wave] = a;
        lds[2 * wave + 1] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = lds[0], sb = lds[1];
        for (int w = 1; w < nwaves; ++w) {
            sa += lds[2 * w];
            sb += lds[2 * w + 1];
        }
        a = sa;
        b = sb;
    }
}

template <class T>
__device__ __forceinline__ void diag_terms(T rho, T ux, T uy, T uz, double& m, double& e) {
    const double dr = double(rho), dx = double(ux), dy = double(uy), dz = double(uz);
    m = dr;
    e = 0.5 * dr * (((dx * dx) + (dy * dy)) + (dz * dz));
}

// ---------------------------------------------------------------------------
// Kernel arguments.
// ---------------------------------------------------------------------------
template <class T> struct StepArgs {
    const void* src;      // lattice read by this iteration (pull form)
    void* dst;            // lattice written
    T* rho;               // N^3
    T* u;                 // 3 N^3, [ux | uy | uz]
    double* partials;     // 2 doubles per block (diagnostics)
    int64_t n3;           // N^3
    int64_t stride;       // CSoA stride S
    int32_t n;            // N
    int32_t log2n;
    int32_t update_macro;
    int32_t diagnostics;
    int32_t xcd_remap;
    int32_t gx;           // row kernels: blocks per row
    int32_t log2gx;       // vector kernel: log2(gx)
    int32_t log2tpr;      // vector kernel: log2(threads per row segment of a block)
    int32_t log2rpb;      // vector kernel: log2(rows per block), > 0 for rows shorter than a block
    StepConsts<T> k;
    float scale, inv_scale;
    T lid_eq[Q];          // feq_i(1, (U, 0, 0)): initial populations of lid-layer cells (deviation form:
                          // their deviations from w_i times the storage scale)
    const T* force;       // body force g, 3 N^3 [gx | gy | gz], or null
};

// ---------------------------------------------------------------------------
// Row kernel (N >= 64).  A block covers blockDim.x consecutive cells of one
// (y, z) row with 1 <= y, z <= N-2; wall rows are not launched.  Because
// slot(c, q) = v(x) + slot(N r, q) with v(x) = 19 x - 18 (x mod S) for
// c = x + N r, every gather address is a per-row (wavefront-uniform, scalar)
// 64-bit base plus one of three small per-thread offsets v(x-1), v(x), v(x+1):
// the loads use a scalar base and a 32-bit vector offset.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t xcd_logical_block(uint32_t b, uint32_t nb, int remap) {
    if (!remap) return b;
    // Workgroups are dispatched round-robin over the 8 XCDs; give each XCD a
    // contiguous range of logical blocks so neighbouring rows share an L2.
    const uint32_t q = nb >> 3, r = nb & 7u;
    const uint32_t xcd = b & 7u, k = b >> 3;
    return xcd < r ? xcd * (q + 1) + k : r * (q + 1) + (xcd - r) * q + k;
}

template <class T, int KIND, int ARITH>
__global__ __launch_bounds__(256) void lbm_collide_stream_row(StepArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    const C codec{a.scale, a.inv_scale};
    const int32_t n = a.n;
    const uint32_t gx = uint32_t(a.gx);
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const uint32_t xb = lb % gx;
    const uint32_t row = lb / gx;                   // 0 .. (N-2)^2 - 1
    const int32_t y = int32_t(row % uint32_t(n - 2)) + 1;
    const int32_t z = int32_t(row / uint32_t(n - 2)) + 1;
    const int32_t x = int32_t(xb * blockDim.x + threadIdx.x);
    const int64_t r = int64_t(y) + int64_t(n) * z;  // row index of cell c = x + N r
    const int32_t t = finish_type(x_flags(x, n) | row_flags(y, z, n));
    const bool active = (t != WALL);
    const int64_t smask = a.stride - 1;
    const uint32_t smask32 = uint32_t(smask);

    // per-thread offsets v(x - ex) for ex = +1, 0, -1 (valid for active lanes)
    const uint32_t xm = uint32_t(x - 1), x0 = uint32_t(x), xp = uint32_t(x + 1);
    const uint32_t v_m = 19u * xm - 18u * (xm & smask32);
    const uint32_t v_0 = 19u * x0 - 18u * (x0 & smask32);
    const uint32_t v_p = 19u * xp - 18u * (xp & smask32);

    const S* __restrict__ src = static_cast<const S*>(a.src);
    S* __restrict__ dst = static_cast<S*>(a.dst);

    T rho = T(0), ux = T(0), uy = T(0), uz = T(0);
    if (active) {
        T f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int ex = ex_of(i), ey = ey_of(i), ez = ez_of(i);
            const int64_t c0 = int64_t(n) * (r - ey - int64_t(n) * ez);  // first cell of the source row
            const int64_t base = 19 * c0 - 18 * (c0 & smask) + int64_t(i) * a.stride;
            const uint32_t off = ex == 1 ? v_m : (ex == -1 ? v_p : v_0);
            f[i] = value_in<ARITH>(codec, src[base + off], w_of<T>(i));
        }
        collide_any<T, ARITH>(t, f, rho, ux, uy, uz, a.k);
        if (a.force) {
            const int64_t c = int64_t(x) + int64_t(n) * r;
            add_force_any<T, ARITH>(t, f, a.force[c], a.force[a.n3 + c], a.force[2 * a.n3 + c], a.k);
        }
        const int64_t c0 = int64_t(n) * r;
        const int64_t obase = 19 * c0 - 18 * (c0 & smask);
#pragma unroll
        for (int i = 0; i < Q; ++i)
            __builtin_nontemporal_store(value_out<ARITH>(codec, f[i], w_of<T>(i)), dst + obase + int64_t(i) * a.stride + v_0);
        if (a.update_macro && is_store_macro(t)) {
            const int64_t c = int64_t(x) + int64_t(n) * r;
            a.rho[c] = rho;
            a.u[c] = ux;
            a.u[a.n3 + c] = uy;
            a.u[2 * a.n3 + c] = uz;
        }
    }
    if (a.diagnostics) {
        __shared__ double lds[2 * 16];
        double m = 0.0, e = 0.0;
        if (active && is_store_macro(t)) diag_terms<T>(rho, ux, uy, uz, m, e);
        block_reduce2(m, e, lds);
        if (threadIdx.x == 0) {
            a.partials[2 * lb] = m;
            a.partials[2 * lb + 1] = e;
        }
    }
}

// ---------------------------------------------------------------------------
// Vector row kernel.  Each lane owns V = LB / sizeof(S) consecutive cells of a
// row (LB = 16 or 8 bytes: 4 or 2 cells in FP32, 2 or 1 in FP64, 8 or 4 with
// 16-bit storage), so every population moves as one LB-byte access per lane:
// an aligned vector load of the lane's own slots, and for populations with
// e_x = +-1 the one element that crosses into the neighbouring lane comes from
// that lane by a cross-lane shuffle; at a wavefront's ends it comes from the
// neighbouring wavefront of the block through LDS, and at the block's ends by
// a wavefront-uniform load.  Stores are whole aligned vectors: the two wall
// cells of a row (x = 0, N-1) get their initial pull-form values back, which
// is what they hold already, so wall slots keep P0.
// Requirements: stride >= V (V consecutive cells are contiguous and LB-byte
// aligned for every population) and N / V >= 64 (a wavefront lies in one row).
// ---------------------------------------------------------------------------
// Cross-lane moves by one lane within the wavefront (DPP wave_shr:1 /
// wave_shl:1).  The lane at the wavefront's end, which has no source lane,
// receives `edge`.
__device__ __forceinline__ uint32_t dpp_from_below(uint32_t v, uint32_t edge) {  // lane l gets lane l-1
    return uint32_t(__builtin_amdgcn_update_dpp(int(edge), int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ uint32_t dpp_from_above(uint32_t v, uint32_t edge) {  // lane l gets lane l+1
    return uint32_t(__builtin_amdgcn_update_dpp(int(edge), int(v), 0x130, 0xf, 0xf, false));
}
template <class S> __device__ __forceinline__ S lane_shift(S v, S edge, bool from_below) {
    if constexpr (sizeof(S) == 8) {
        const uint64_t b = __builtin_bit_cast(uint64_t, v), e = __builtin_bit_cast(uint64_t, edge);
        const uint32_t lo = from_below ? dpp_from_below(uint32_t(b), uint32_t(e)) : dpp_from_above(uint32_t(b), uint32_t(e));
        const uint32_t hi = from_below ? dpp_from_below(uint32_t(b >> 32), uint32_t(e >> 32))
                                       : dpp_from_above(uint32_t(b >> 32), uint32_t(e >> 32));
        return __builtin_bit_cast(S, (uint64_t(hi) << 32) | lo);
    } else if constexpr (sizeof(S) == 4) {
        const uint32_t b = __builtin_bit_cast(uint32_t, v), e = __builtin_bit_cast(uint32_t, edge);
        return __builtin_bit_cast(S, from_below ? dpp_from_below(b, e) : dpp_from_above(b, e));
    } else {
        const uint32_t b = __builtin_bit_cast(uint16_t, v), e = __builtin_bit_cast(uint16_t, edge);
        return __builtin_bit_cast(S, uint16_t(from_below ? dpp_from_below(b, e) : dpp_from_above(b, e)));
    }
}

// Initial pull-form value of population i at the wall cell x = 0 (side -1) or
// x = N-1 (side +1) of row (y, z): F0 of the destination (x + e_x, y + e_y,
// z + e_z), i.e. NaN unless the destination is a fluid-side cell.
// In the deviation form (ARITH == AR_DELTA) the value is (F0 - w_i) * scale:
// lid_eq then holds the scaled lid-layer deviations and the rest value is 0.
template <class T, int ARITH = AR_EXACT>
__device__ __forceinline__ T wall_p0(int i, int side, int32_t y, int32_t z, int32_t n, const T* lid_eq) {
    const int32_t dy = y + ey_of(i), dz = z + ez_of(i);
    const bool into = (side < 0) ? (ex_of(i) == 1) : (ex_of(i) == -1);
    if (!into || dy < 1 || dy > n - 2 || dz < 1 || dz > n - 2) return T(__builtin_nan(""));
    return dz == n - 2 ? lid_eq[i] : (ARITH == AR_DELTA ? T(0) : w_of<T>(i));
}

// Occupancy target of the vector kernel: the waves per SIMD its register
// budget is fitted to (512 VGPRs per SIMD lane / waves).
#ifndef LBM_VEC8_NATIVE_WAVES
#define LBM_VEC8_NATIVE_WAVES 4
#endif
// base + uoff + off elements as (uniform address) + (32-bit byte offset): the
// form the global load/store instructions take with the address in scalar
// registers (saddr), so no 64-bit per-lane address arithmetic.  uoff is
// wavefront-uniform; it goes through readfirstlane, which keeps the compiler
// from folding the per-lane offset into it (bases that differ by multiples of
// the stride were otherwise turned into one per-lane base plus 64-bit
// multiply-adds per access).  off * sizeof(E) must be below 2^32 (row offsets
// here are below 19 * 1024 elements).
template <class P, class E>
__device__ __forceinline__ P* at_offset(E* base, int64_t uoff, uint32_t off) {
    using B = typename std::conditional<std::is_const<E>::value, const char, char>::type;
    const uint64_t v = uint64_t(uoff);
    const int64_t u = int64_t(uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(v)))) |
                              (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(uint32_t(v >> 32)))) << 32));
    return reinterpret_cast<P*>(reinterpret_cast<B*>(base + u) + off * uint32_t(sizeof(E)));
}

#ifndef LBM_VEC8_16BIT_WAVES
#define LBM_VEC8_16BIT_WAVES 3
#endif
template <class T, int KIND, int LB> struct VecWaves {
    static constexpr int value =
        (LB == 16) ? (sizeof(T) == 8 ? 3 : 2) : (KIND == ST_NATIVE ? LBM_VEC8_NATIVE_WAVES : LBM_VEC8_16BIT_WAVES);
};

// CACHED: store the lattice through the caches (lattices small enough to stay
// in the 256 MB memory-attached last-level cache until the next step reads
// them), else with streaming stores.  A template parameter: a run-time branch
// around the 19 stores cost 9% at 256^3.
template <class T, int KIND, int LB, int ARITH, bool CACHED = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(VecWaves<T, KIND, LB>::value)))
void lbm_collide_stream_vec(StepArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    constexpr int V = int(LB / sizeof(S));  // cells per lane
    typedef S SV __attribute__((ext_vector_type(V)));
    const C codec{a.scale, a.inv_scale};
    const int32_t n = a.n;
    // grid: gx blocks per row, N^2 rows (wall rows exit at once); all powers of two
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x, a.xcd_remap);
    // a block covers 2^log2rpb rows of 2^log2tpr threads each (rows shorter than
    // a block), or one segment of 2^log2tpr threads of a row (gx segments)
    // row of the block: wavefront-uniform (a row segment holds whole wavefronts,
    // 2^log2tpr >= 64), made explicit so the row's addresses stay scalar
    const uint32_t seg = __builtin_amdgcn_readfirstlane(threadIdx.x >> a.log2tpr);
    const uint32_t tid = threadIdx.x & ((1u << a.log2tpr) - 1u);        // thread in its row segment
    const uint32_t tpr = 1u << a.log2tpr;
    const uint32_t xb = lb & uint32_t(a.gx - 1);
    const uint32_t row = ((lb >> a.log2gx) << a.log2rpb) + seg;
    const int32_t y = int32_t(row & uint32_t(n - 1));
    const int32_t z = int32_t(row >> a.log2n);
    const bool wall_row = (y == 0 || y == n - 1 || z == 0 || z == n - 1);  // uniform per wavefront
    if (a.log2rpb == 0 && wall_row) {  // the whole block: nothing to do
        if (a.diagnostics && threadIdx.x == 0) {
            a.partials[2 * lb] = 0.0;
            a.partials[2 * lb + 1] = 0.0;
        }
        return;
    }
    const int32_t x0 = int32_t((xb * tpr + tid) * V);
    // first cell of this wavefront, made explicitly wavefront-uniform
    const int32_t xw = __builtin_amdgcn_readfirstlane(int32_t((xb * tpr + (tid & ~63u)) * V));
    const int64_t r = int64_t(y) + int64_t(n) * z;
    const int32_t rf = row_flags(y, z, n);
    const int64_t smask = a.stride - 1;
    const uint32_t smask32 = uint32_t(smask);
    auto voff = [&](int32_t xx) -> uint32_t {  // v(x) = slot(N r + x) - slot(N r)
        const uint32_t u = uint32_t(xx);
        return 19u * u - 18u * (u & smask32);
    };
    auto src_off = [&](int i) -> int64_t {  // slot 0 of the source row of population i
        const int64_t c0 = int64_t(n) * (r - ey_of(i) - int64_t(n) * ez_of(i));
        return 19 * c0 - 18 * (c0 & smask) + int64_t(i) * a.stride;
    };
    const S* __restrict__ src = static_cast<const S*>(a.src);
    auto src_row = [&](int i) -> const S* { return src + src_off(i); };
    const uint32_t off0 = voff(x0);
    const bool has_left = xw > 0;               // cell xw-1 exists (it is not needed when xw = 0)
    const bool has_right = xw + 64 * V < n;     // cell xw+64V exists
    S* __restrict__ dst = static_cast<S*>(a.dst);

    // all 19 vector loads first, then the wavefront-end elements, then the
    // shifts.  Every slot is read by exactly one cell per step, so the loads
    // are streaming (nontemporal: 1.5-2.5% faster than cached loads at 512^3)
    SV A[Q];
    if (!wall_row) {
#pragma unroll
        for (int i = 0; i < Q; ++i) A[i] = LBM_LAT_LOAD(at_offset<const SV>(src, src_off(i), off0));
    } else {
#pragma unroll
        for (int i = 0; i < Q; ++i) A[i] = SV{};
    }
    // wavefront-end elements: from the neighbouring wavefront of the row
    // segment through LDS, from memory only at the segment's ends
    __shared__ S xe_up[16][5];  // per wavefront: lane 63's last element, e_x = +1 populations
    __shared__ S xe_dn[16][5];  // per wavefront: lane 0's first element, e_x = -1 populations
    const uint32_t wave = tid >> 6, lane = tid & 63u, nwaves = tpr >> 6;  // within the row segment
    const uint32_t gw = threadIdx.x >> 6;                                  // within the block
    // Rows of one wavefront (nwaves == 1, block-uniform) exchange nothing: their
    // ends come from memory, and their block -- several such rows -- needs no barrier.
    if (nwaves > 1) {
        int ku = 0, kd = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (ex_of(i) == 1) {
                if (lane == 63) xe_up[gw][ku] = A[i][V - 1];
                ++ku;
            } else if (ex_of(i) == -1) {
                if (lane == 0) xe_dn[gw][kd] = A[i][0];
                ++kd;
            }
        }
        __syncthreads();
    }
    S edge[Q];
    {
        int ku = 0, kd = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            edge[i] = S(0);
            if (ex_of(i) == 1) {
                if (wave > 0) edge[i] = xe_up[gw - 1][ku];
                else if (has_left && !wall_row) edge[i] = src_row(i)[voff(xw - 1)];
                ++ku;
            } else if (ex_of(i) == -1) {
                if (wave + 1 < nwaves) edge[i] = xe_dn[gw + 1][kd];
                else if (has_right && !wall_row) edge[i] = src_row(i)[voff(xw + 64 * V)];
                ++kd;
            }
        }
    }
    // gather: g[i][j] is population i arriving at cell x0 + j
    S g[Q][V];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int ex = ex_of(i);
        if (ex == 0) {
#pragma unroll
            for (int j = 0; j < V; ++j) g[i][j] = A[i][j];
        } else if (ex == 1) {  // source x - 1
            g[i][0] = lane_shift<S>(A[i][V - 1], edge[i], true);
#pragma unroll
            for (int j = 1; j < V; ++j) g[i][j] = A[i][j - 1];
        } else {  // source x + 1
#pragma unroll
            for (int j = 0; j < V - 1; ++j) g[i][j] = A[i][j + 1];
            g[i][V - 1] = lane_shift<S>(A[i][0], edge[i], false);
        }
    }

    T rho_v[V], ux_v[V], uy_v[V], uz_v[V];
    double m = 0.0, e = 0.0;
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const int32_t t = finish_type(x_flags(x0 + j, n) | rf);
        T f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = value_in<ARITH>(codec, g[i][j], w_of<T>(i));
        T rho = T(0), ux = T(0), uy = T(0), uz = T(0);
        collide_any<T, ARITH>(t, f, rho, ux, uy, uz, a.k);  // WALL: passes f through unchanged
        if (a.force) {
            const int64_t c = int64_t(x0 + j) + int64_t(n) * r;
            add_force_any<T, ARITH>(t, f, a.force[c], a.force[a.n3 + c], a.force[2 * a.n3 + c], a.k);
        }
        const bool store = is_store_macro(t);
        const T qnan = T(__builtin_nan(""));
        rho_v[j] = store ? rho : qnan;
        ux_v[j] = store ? ux : qnan;
        uy_v[j] = store ? uy : qnan;
        uz_v[j] = store ? uz : qnan;
        if (a.diagnostics && store) {
            double dm, de;
            diag_terms<T>(rho, ux, uy, uz, dm, de);
            m += dm;
            e += de;
        }
        if (j == 0 || j == V - 1) {
            // the row's wall cells keep their initial values
            const bool wl = (j == 0) && (x0 == 0);
            const bool wr = (j == V - 1) && (x0 + V - 1 == n - 1);
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if (wl) f[i] = wall_p0<T, ARITH>(i, -1, y, z, n, a.lid_eq);
                if (wr) f[i] = wall_p0<T, ARITH>(i, +1, y, z, n, a.lid_eq);
            }
        }
#pragma unroll
        for (int i = 0; i < Q; ++i) g[i][j] = value_out<ARITH>(codec, f[i], w_of<T>(i));
    }

    const int64_t oc0 = int64_t(n) * r;
    const int64_t orow = 19 * oc0 - 18 * (oc0 & smask);
    if (!wall_row) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            SV o;
#pragma unroll
            for (int j = 0; j < V; ++j) o[j] = g[i][j];
            SV* p = at_offset<SV>(dst, orow + int64_t(i) * a.stride, off0);
            if constexpr (CACHED) *p = o;
            else LBM_LAT_STORE(o, p);
        }
    }

    if (a.update_macro && !wall_row) {
        // rho and u of the lane's V cells, in chunks of at most 16 bytes
        const int64_t c = int64_t(x0) + oc0;
        constexpr int CH = int(16 / sizeof(T)) < V ? int(16 / sizeof(T)) : V;
#pragma unroll
        for (int j0 = 0; j0 < V; j0 += CH) {
            if constexpr (CH == 1) {
                a.rho[c + j0] = rho_v[j0];
                a.u[c + j0] = ux_v[j0];
                a.u[a.n3 + c + j0] = uy_v[j0];
                a.u[2 * a.n3 + c + j0] = uz_v[j0];
            } else {
                typedef T TC __attribute__((ext_vector_type(CH)));
                TC vr, vx, vy, vz;
#pragma unroll
                for (int k = 0; k < CH; ++k) {
                    vr[k] = rho_v[j0 + k];
                    vx[k] = ux_v[j0 + k];
                    vy[k] = uy_v[j0 + k];
                    vz[k] = uz_v[j0 + k];
                }
                *reinterpret_cast<TC*>(a.rho + c + j0) = vr;
                *reinterpret_cast<TC*>(a.u + c + j0) = vx;
                *reinterpret_cast<TC*>(a.u + a.n3 + c + j0) = vy;
                *reinterpret_cast<TC*>(a.u + 2 * a.n3 + c + j0) = vz;
            }
        }
    }
    if (a.diagnostics) {
        __shared__ double lds[2 * 16];
        block_reduce2(m, e, lds);
        if (threadIdx.x == 0) {
            a.partials[2 * lb] = m;
            a.partials[2 * lb + 1] = e;
        }
    }
}

// ---------------------------------------------------------------------------
// Row-pipeline kernel: a persistent grid whose blocks walk interior rows of
// NX = N cells, staging each row's 19 source spans in LDS with gfx950's
// direct global-to-LDS loads (global_load_lds_dwordx4).  Those loads hold no
// VGPRs while in flight, so a block issues the next row's loads before it
// computes the current one: the memory system always has a row per block in
// flight, at two or four waves per SIMD, where the vector kernel keeps its
// loads in registers and needs four waves per SIMD to hide its compute.
//
// LDS: two buffers of 19 spans; span i holds the source row of population i,
// cells 0 .. NX-1, between 16-byte pads, so a cell reads its population i at
// x - e_x without bounds tests (x = 0 and x = N-1 are wall cells whose
// gathered values are not used: they store their initial values, wall_p0).
// One load instruction moves 1 KB: 64 lanes x 16 bytes of one span.
//
// Per row k of the block: wait for the loads of row k (issued during row
// k-1; the stores of row k-1 may still be in flight), barrier, copy the row's
// gathers from LDS into registers, issue the loads of row k+1 into the other
// buffer (everyone read it before this barrier), compute, store.  Wall rows
// are not visited; block b takes interior rows b, b + grid, b + 2 grid, ...
// (any grid size is correct: blocks do not depend on each other).  Each
// block's diagnostics cover all its rows.
// Requirements: N = NX, stride >= 16 / sizeof(stored) (whole 16-byte pieces
// are contiguous).
// ---------------------------------------------------------------------------
// One direct global-to-LDS load of 16 bytes per lane (gfx950
// global_load_lds_dwordx4): lane l's bytes land at LDS byte address
// lds_base + 16 l (lds_base wavefront-uniform, through M0).  Written as inline
// assembly so that the compiler, which does not see the LDS write, inserts no
// wait of its own for it: the kernel waits with s_waitcnt vmcnt before the
// barrier that precedes the LDS reads.  The instruction writes no VGPR.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"  // m0 is reserved: the clobber documents the write
__device__ __forceinline__ void dma16(const void* src, uint32_t lds_base) {
    asm volatile("s_mov_b32 m0, %1\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(src), "s"(lds_base)
                 : "memory", "m0");
}
#pragma clang diagnostic pop
template <class S> __device__ __forceinline__ uint32_t lds_address(S* p) {
    return uint32_t(reinterpret_cast<uintptr_t>((__attribute__((address_space(3))) S*)(p)));
}

template <class S, int NX> struct LdsRow {
    static constexpr int PAD = 16 / int(sizeof(S));             // elements of a 16-byte pad
    static constexpr int PITCH = NX + 2 * PAD;                    // elements per span
    static constexpr int V = 8 / int(sizeof(S));                  // cells per lane (8-byte lanes)
    static constexpr int THREADS = NX / V;
    static constexpr int SPAN_LANES = NX * int(sizeof(S)) / 16;   // 16-byte pieces per span
    static constexpr int LANES = SPAN_LANES < 64 ? SPAN_LANES : 64;  // active lanes per load instruction
    static constexpr int CPI = LANES * 16 / int(sizeof(S));       // cells per load instruction
    static constexpr int CHUNKS = NX / CPI;                       // instructions per span
    static constexpr int LOADS = Q * CHUNKS;                      // instructions per row
    static constexpr int WAVES = THREADS / 64;
    static constexpr int BUF = Q * PITCH;                         // elements per buffer
};

template <class T, int KIND, int ARITH, int NX>
__global__ __launch_bounds__((LdsRow<typename Codec<T, KIND>::S, NX>::THREADS)) void lbm_collide_stream_lds(StepArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    using L = LdsRow<S, NX>;
    constexpr int V = L::V;
    typedef S SV __attribute__((ext_vector_type(V)));
    __shared__ __align__(16) S lds[2 * L::BUF];
    const C codec{a.scale, a.inv_scale};
    constexpr int32_t n = NX;
    const int64_t smask = a.stride - 1;
    const uint32_t smask32 = uint32_t(smask);
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const int32_t x0 = int32_t(tid) * V;
    auto voff = [&](int32_t xx) -> uint32_t {  // slot(N r + x) - slot(N r)
        const uint32_t u = uint32_t(xx);
        return 19u * u - 18u * (u & smask32);
    };
    const uint32_t off0 = voff(x0);
    const int64_t rows = int64_t(n - 2) * (n - 2);
    const int64_t first = blockIdx.x, step = gridDim.x;

    // loads of row j (interior index) into buffer b: wave w issues instructions m = w, w + WAVES, ...
    const int wave_u = __builtin_amdgcn_readfirstlane(int(wave));
    auto issue = [&](int64_t j, int b) {
        const int64_t r = (1 + j % (n - 2)) + int64_t(n) * (1 + j / (n - 2));  // y + N z
#pragma unroll
        for (int m = 0; m < L::LOADS; ++m) {
            if (m % L::WAVES == wave_u) {
                const int i = m / L::CHUNKS, h = m % L::CHUNKS;
                // source row of population i (wavefront-uniform)
                const int64_t c0 = int64_t(n) * (r - ey_of(i) - int64_t(n) * ez_of(i));
                const S* row = static_cast<const S*>(a.src) + (19 * c0 - 18 * (c0 & smask) + int64_t(i) * a.stride);
                const int32_t xl = h * L::CPI + int32_t(lane) * (16 / int(sizeof(S)));
                S* ldst = lds + b * L::BUF + i * L::PITCH + L::PAD + h * L::CPI;
                if (L::LANES == 64 || int(lane) < L::LANES) dma16(row + voff(xl), lds_address(ldst));
            }
        }
    };

    double m_acc = 0.0, e_acc = 0.0;
    int k = 0;
    if (first < rows) issue(first, 0);
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");  // the first row's loads
    for (int64_t j = first; j < rows; j += step, ++k) {
        const int b = k & 1;
        // Vector memory operations complete in issue order, and each wave
        // issued this row's loads before the previous row's 19 stores: once at
        // most 19 operations are outstanding, the loads have landed.  The
        // barrier then makes every wave's part visible.  (Inline assembly: one
        // s_barrier on every path, and no compiler-inserted wait for the stores.)
        asm volatile("s_waitcnt vmcnt(19)\n\ts_barrier" ::: "memory");
        const int32_t y = int32_t(1 + j % (n - 2)), z = int32_t(1 + j / (n - 2));
        const int64_t r = int64_t(y) + int64_t(n) * z;
        // gathers: g[i][c] is population i arriving at cell x0 + c, read at x0 + c - e_x
        S g[Q][V];
        const S* lb = lds + b * L::BUF + L::PAD;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const S* sp = lb + i * L::PITCH + x0 - ex_of(i);
            if (ex_of(i) == 0) {
                const SV v = *reinterpret_cast<const SV*>(sp);
#pragma unroll
                for (int c = 0; c < V; ++c) g[i][c] = v[c];
            } else {
#pragma unroll
                for (int c = 0; c < V; ++c) g[i][c] = sp[c];
            }
        }
        if (j + step < rows) issue(j + step, b ^ 1);

        const int32_t rf = row_flags(y, z, n);
#pragma unroll
        for (int c = 0; c < V; ++c) {
            const int32_t t = finish_type(x_flags(x0 + c, n) | rf);
            T f[Q];
#pragma unroll
            for (int i = 0; i < Q; ++i) f[i] = value_in<ARITH>(codec, g[i][c], w_of<T>(i));
            T rho = T(0), ux = T(0), uy = T(0), uz = T(0);
            collide_any<T, ARITH>(t, f, rho, ux, uy, uz, a.k);  // WALL: passes f through unchanged
            if (a.force) {
                const int64_t cc = int64_t(x0 + c) + int64_t(n) * r;
                add_force_any<T, ARITH>(t, f, a.force[cc], a.force[a.n3 + cc], a.force[2 * a.n3 + cc], a.k);
            }
            const bool store = is_store_macro(t);
            if (a.diagnostics && store) {
                double dm, de;
                diag_terms<T>(rho, ux, uy, uz, dm, de);
                m_acc += dm;
                e_acc += de;
            }
            if (a.update_macro) {
                const int64_t cc = int64_t(x0 + c) + int64_t(n) * r;
                const T qnan = T(__builtin_nan(""));
                a.rho[cc] = store ? rho : qnan;
                a.u[cc] = store ? ux : qnan;
                a.u[a.n3 + cc] = store ? uy : qnan;
                a.u[2 * a.n3 + cc] = store ? uz : qnan;
            }
            const bool wl = (x0 + c == 0), wr = (x0 + c == n - 1);  // the row's wall cells keep their initial values
#pragma unroll
            for (int i = 0; i < Q; ++i) {
                if (wl) f[i] = wall_p0<T, ARITH>(i, -1, y, z, n, a.lid_eq);
                if (wr) f[i] = wall_p0<T, ARITH>(i, +1, y, z, n, a.lid_eq);
            }
#pragma unroll
            for (int i = 0; i < Q; ++i) g[i][c] = value_out<ARITH>(codec, f[i], w_of<T>(i));
        }
        const int64_t oc0 = int64_t(n) * r;
        S* orow = static_cast<S*>(a.dst) + (19 * oc0 - 18 * (oc0 & smask));
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            SV o;
#pragma unroll
            for (int c = 0; c < V; ++c) o[c] = g[i][c];
            LBM_LAT_STORE(o, reinterpret_cast<SV*>(orow + int64_t(i) * a.stride + off0));
        }
    }
    if (a.diagnostics) {
        __shared__ double red[2 * 16];
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        block_reduce2(m_acc, e_acc, red);
        if (threadIdx.x == 0) {
            a.partials[2 * blockIdx.x] = m_acc;
            a.partials[2 * blockIdx.x + 1] = e_acc;
        }
    }
}

// ---------------------------------------------------------------------------
// Tile kernel for small CSoA strides: a stride block of one population of at
// most 32 bytes (S = 1 is AoS).  There the populations of neighbouring cells
// share cache lines, so per-population accesses fetch each line many times
// from L2.  A block takes TY x TZ rows and B cells along x (B a multiple of
// the stride): it loads the B-cell span of each of the (TY + 2)(TZ + 2) source
// rows once -- 19 B contiguous stored values, since the span starts a stride
// block -- with 16-byte loads into LDS, plus the x halo cells x0 - 1 and
// x0 + B for the five populations that cross each side, gathers every cell's
// populations from LDS, and writes each destination span back through LDS
// with 16-byte stores.  Wall rows inside a tile are neither gathered nor
// written; the row's wall cells get their initial pull-form values back.
// ---------------------------------------------------------------------------
// Tile shapes (B cells along x, TY x TZ rows) per stored size, measured on
// MI355X at 512^3 (FP32, 16-bit) and 256^3 (FP64) over strides 1-16: thin x
// extents with tall y-z tiles (60 source rows for 32 destination rows) and two
// blocks per CU beat the 128-byte-wide tiles by 1.5x.
#ifndef LBM_TILE_4B
#define LBM_TILE_4B 16, 4, 8
#endif
#ifndef LBM_TILE_8B
#define LBM_TILE_8B 8, 4, 8
#endif
#ifndef LBM_TILE_2B
#define LBM_TILE_2B 32, 4, 4
#endif
template <int B_, int TY_, int TZ_> struct TileDims {
    static constexpr int B = B_, TY = TY_, TZ = TZ_;
};
template <class S> struct TileShape;
template <> struct TileShape<float> : TileDims<LBM_TILE_4B> {};
template <> struct TileShape<double> : TileDims<LBM_TILE_8B> {};
template <> struct TileShape<_Float16> : TileDims<LBM_TILE_2B> {};
template <> struct TileShape<uint16_t> : TileDims<LBM_TILE_2B> {};

// populations with e_x = +1 and e_x = -1, in index order
__device__ __forceinline__ constexpr int xp_slot(int i) {
    return i == 1 ? 0 : i == 7 ? 1 : i == 10 ? 2 : i == 11 ? 3 : i == 15 ? 4 : -1;
}
__device__ __forceinline__ constexpr int xm_slot(int i) {
    return i == 3 ? 0 : i == 8 ? 1 : i == 9 ? 2 : i == 13 ? 3 : i == 17 ? 4 : -1;
}

template <class T, int KIND> struct TileThreads {
    using TS = TileShape<typename Codec<T, KIND>::S>;
    static constexpr int value = TS::B * TS::TY * TS::TZ;
};

template <class T, int KIND, int ARITH>
__global__ __launch_bounds__((TileThreads<T, KIND>::value)) void lbm_collide_stream_tile(StepArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    using TS = TileShape<S>;
    constexpr int B = TS::B, TY = TS::TY, TZ = TS::TZ;
    constexpr int NT = B * TY * TZ;
    constexpr int SY = TY + 2, NR = (TY + 2) * (TZ + 2);  // source rows of the tile
    constexpr int SPAN = Q * B;                            // stored values of one row span
    constexpr int VE = 16 / int(sizeof(S));                // values per 16-byte access
    constexpr int CH = SPAN / VE;                          // 16-byte chunks per span
    static_assert(SPAN % VE == 0, "span must be whole 16-byte chunks");
    constexpr int PER = (NR * CH + NT - 1) / NT;           // chunks per thread
    __shared__ __align__(16) S span[NR][SPAN];
    __shared__ S halo_l[NR][5];  // cell x0 - 1: populations with e_x = +1
    __shared__ S halo_r[NR][5];  // cell x0 + B: populations with e_x = -1
    __shared__ double lds_red[2 * (NT / 64)];

    const C codec{a.scale, a.inv_scale};
    const int32_t n = a.n;
    const int32_t x0 = int32_t(blockIdx.x) * B, y0 = int32_t(blockIdx.y) * TY, z0 = int32_t(blockIdx.z) * TZ;
    const int tid = int(threadIdx.x);
    const S* __restrict__ src = static_cast<const S*>(a.src);
    S* __restrict__ dst = static_cast<S*>(a.dst);
    const int ls = __builtin_ctzll(uint64_t(a.stride));  // log2 stride (<= log2 B)
    const int smask = (1 << ls) - 1;
    auto row_first = [&](int32_t yy, int32_t zz) -> int64_t {  // slot of (x0, yy, zz), population 0
        return 19 * (int64_t(x0) + int64_t(n) * (int64_t(yy) + int64_t(n) * zz));
    };
    auto csoa = [&](int xs, int i) -> int {  // position of (cell x0 + xs, population i) in a span
        return (((xs >> ls) * Q) << ls) + (i << ls) + (xs & smask);
    };

    // 1. source spans: all loads first, then the LDS writes
    typedef uint32_t U4 __attribute__((ext_vector_type(4)));
    U4 buf[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q = tid + k * NT;
        const int r = q / CH, ch = q - (q / CH) * CH;
        const int32_t yy = y0 - 1 + r % SY, zz = z0 - 1 + r / SY;
        buf[k] = U4{0u, 0u, 0u, 0u};
        if (q < NR * CH && yy >= 0 && yy < n && zz >= 0 && zz < n)
            buf[k] = *reinterpret_cast<const U4*>(src + row_first(yy, zz) + ch * VE);
    }
    // x halos: 5 populations per side and source row
    constexpr int NH = NR * 10, HPER = (NH + NT - 1) / NT;
    S hv[HPER];
#pragma unroll
    for (int k = 0; k < HPER; ++k) {
        const int hq = tid + k * NT;
        hv[k] = S(0);
        const int r = hq / 10, side = (hq / 5) % 2, j = hq % 5;
        const int32_t yy = y0 - 1 + r % SY, zz = z0 - 1 + r / SY;
        const int32_t xx = side == 0 ? x0 - 1 : x0 + B;
        const int i = side == 0 ? (j == 0 ? 1 : j == 1 ? 7 : j == 2 ? 10 : j == 3 ? 11 : 15)
                                : (j == 0 ? 3 : j == 1 ? 8 : j == 2 ? 9 : j == 3 ? 13 : 17);
        if (hq < NH && yy >= 0 && yy < n && zz >= 0 && zz < n && xx >= 0 && xx < n)
            hv[k] = src[slot(int64_t(xx) + int64_t(n) * (int64_t(yy) + int64_t(n) * zz), i, a.stride)];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const int q = tid + k * NT;
        if (q < NR * CH) {
            const int r = q / CH, ch = q - (q / CH) * CH;
            *reinterpret_cast<U4*>(&span[r][ch * VE]) = buf[k];
        }
    }
#pragma unroll
    for (int k = 0; k < HPER; ++k) {
        const int hq = tid + k * NT;
        if (hq < NH) {
            const int r = hq / 10, side = (hq / 5) % 2, j = hq % 5;
            if (side == 0) halo_l[r][j] = hv[k];
            else halo_r[r][j] = hv[k];
        }
    }
    __syncthreads();

    // 2. one cell per thread
    const int xl = tid % B, ty = (tid / B) % TY, tz = tid / (B * TY);
    const int32_t x = x0 + xl, y = y0 + ty, z = z0 + tz;
    const bool row_ok = y >= 1 && y <= n - 2 && z >= 1 && z <= n - 2;
    const int32_t t = finish_type(x_flags(x, n) | row_flags(y, z, n));
    T f[Q];
    T rho = T(0), ux = T(0), uy = T(0), uz = T(0);
    if (row_ok && t != WALL) {
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int r = (ty + 1 - ey_of(i)) + SY * (tz + 1 - ez_of(i));
            const int xs = xl - ex_of(i);
            S v;
            if (ex_of(i) == 1 && xs < 0) v = halo_l[r][xp_slot(i)];
            else if (ex_of(i) == -1 && xs >= B) v = halo_r[r][xm_slot(i)];
            else v = span[r][csoa(xs, i)];
            f[i] = value_in<ARITH>(codec, v, w_of<T>(i));
        }
        collide_any<T, ARITH>(t, f, rho, ux, uy, uz, a.k);
        const int64_t c = int64_t(x) + int64_t(n) * (int64_t(y) + int64_t(n) * z);
        if (a.force) add_force_any<T, ARITH>(t, f, a.force[c], a.force[a.n3 + c], a.force[2 * a.n3 + c], a.k);
        if (a.update_macro && is_store_macro(t)) {
            a.rho[c] = rho;
            a.u[c] = ux;
            a.u[a.n3 + c] = uy;
            a.u[2 * a.n3 + c] = uz;
        }
    } else if (row_ok) {  // the row's wall cells keep their initial values
#pragma unroll
        for (int i = 0; i < Q; ++i) f[i] = wall_p0<T, ARITH>(i, x == 0 ? -1 : +1, y, z, n, a.lid_eq);
    }
    __syncthreads();  // every gather is done: the spans are reused for the output

    // 3. outputs through LDS (destination row (ty, tz) uses span row ty + TY tz)
    const int ro = ty + TY * tz;
    if (row_ok) {
#pragma unroll
        for (int i = 0; i < Q; ++i) span[ro][csoa(xl, i)] = value_out<ARITH>(codec, f[i], w_of<T>(i));
    }
    __syncthreads();
    constexpr int PERO = (TY * TZ * CH + NT - 1) / NT;
#pragma unroll
    for (int k = 0; k < PERO; ++k) {
        const int q = tid + k * NT;
        if (q < TY * TZ * CH) {
            const int r = q / CH, ch = q - (q / CH) * CH;
            const int32_t yy = y0 + r % TY, zz = z0 + r / TY;
            if (yy >= 1 && yy <= n - 2 && zz >= 1 && zz <= n - 2)
                __builtin_nontemporal_store(*reinterpret_cast<const U4*>(&span[r][ch * VE]),
                                            reinterpret_cast<U4*>(dst + row_first(yy, zz) + ch * VE));
        }
    }

    if (a.diagnostics) {
        double m = 0.0, e = 0.0;
        if (row_ok && t != WALL && is_store_macro(t)) diag_terms<T>(rho, ux, uy, uz, m, e);
        block_reduce2(m, e, lds_red);
        if (threadIdx.x == 0) {
            const uint32_t lb = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
            a.partials[2 * lb] = m;
            a.partials[2 * lb + 1] = e;
        }
    }
}

// ---------------------------------------------------------------------------
// Generic kernel: any N, any stride, any block size.  One thread per cell of a
// 1-D grid over all N^3 cells, full 64-bit slot arithmetic per population.
// ---------------------------------------------------------------------------
template <class T, int KIND, int ARITH>
__global__ __launch_bounds__(1024) void lbm_collide_stream_generic(StepArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    const C codec{a.scale, a.inv_scale};
    const int32_t n = a.n;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool inside = c < a.n3;
    const int32_t x = int32_t(c & (n - 1));
    const int32_t y = int32_t((c >> a.log2n) & (n - 1));
    const int32_t z = int32_t(c >> (2 * a.log2n));
    const int32_t t = inside ? cell_type(x, y, z, n) : WALL;
    const bool active = (t != WALL);
    const S* __restrict__ src = static_cast<const S*>(a.src);
    S* __restrict__ dst = static_cast<S*>(a.dst);
    T rho = T(0), ux = T(0), uy = T(0), uz = T(0);
    if (active) {
        T f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int64_t s = c - (ex_of(i) + int64_t(n) * (ey_of(i) + int64_t(n) * ez_of(i)));
            f[i] = value_in<ARITH>(codec, src[slot(s, i, a.stride)], w_of<T>(i));
        }
        collide_any<T, ARITH>(t, f, rho, ux, uy, uz, a.k);
        if (a.force) add_force_any<T, ARITH>(t, f, a.force[c], a.force[a.n3 + c], a.force[2 * a.n3 + c], a.k);
#pragma unroll
        for (int i = 0; i < Q; ++i)
            __builtin_nontemporal_store(value_out<ARITH>(codec, f[i], w_of<T>(i)), dst + slot(c, i, a.stride));
        if (a.update_macro && is_store_macro(t)) {
            a.rho[c] = rho;
            a.u[c] = ux;
            a.u[a.n3 + c] = uy;
            a.u[2 * a.n3 + c] = uz;
        }
    }
    if (a.diagnostics) {
        __shared__ double lds[2 * 1024];
        double m = 0.0, e = 0.0;
        if (active && is_store_macro(t)) diag_terms<T>(rho, ux, uy, uz, m, e);
        block_reduce2(m, e, lds);
        if (threadIdx.x == 0) {
            a.partials[2 * blockIdx.x] = m;
            a.partials[2 * blockIdx.x + 1] = e;
        }
    }
}

// ---------------------------------------------------------------------------
// initialize (kernels.cl:277-318) in pull form: cell s writes map, rho, u and
// P0[s][i] = F0[s + e_i][i] into both lattices (NaN where s + e_i is outside
// the lattice or a wall).
// ---------------------------------------------------------------------------
template <class T> struct InitArgs {
    void* a;
    void* b;
    T* rho;
    T* u;
    int64_t n3;
    int64_t stride;
    int32_t n;
    int32_t log2n;
    T U;
    float scale, inv_scale;
    int32_t contracted;  // nonzero: the equilibrium as ROCm's OpenCL compiler evaluates it
    T lid_delta[Q];      // 16-bit storage (deviation form): F0 - w_i of FRONT cells
};

template <class T, int KIND>
__global__ __launch_bounds__(256) void lbm_initialize_kernel(InitArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    const C codec{a.scale, a.inv_scale};
    const int32_t n = a.n;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= a.n3) return;
    const int32_t x = int32_t(c & (n - 1));
    const int32_t y = int32_t((c >> a.log2n) & (n - 1));
    const int32_t z = int32_t(c >> (2 * a.log2n));
    const int32_t t = cell_type(x, y, z, n);
    const T qnan = __builtin_nan("");
    const bool st = is_store_macro(t);
    a.rho[c] = st ? T(1) : qnan;
    a.u[c] = st ? (is_moving_init(t) ? a.U : T(0)) : qnan;
    a.u[a.n3 + c] = st ? T(0) : qnan;
    a.u[2 * a.n3 + c] = st ? T(0) : qnan;
    S* pa = static_cast<S*>(a.a);
    S* pb = static_cast<S*>(a.b);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int32_t dx = x + ex_of(i), dy = y + ey_of(i), dz = z + ez_of(i);
        T v = qnan;
        if (dx >= 0 && dy >= 0 && dz >= 0 && dx < n && dy < n && dz < n) {
            const int32_t td = cell_type(dx, dy, dz, n);
            if (!is_wall(td) && KIND != ST_NATIVE) {
                v = is_moving_init(td) ? a.lid_delta[i] : T(0);  // deviation form
            } else if (!is_wall(td)) {
                // kernels.cl:292-308 with rho = 1, u = (U or 0, 0, 0)
                const T ux = is_moving_init(td) ? a.U : T(0);
                if (a.contracted) {
                    v = feq_contracted<T>(i, T(1), ux, T(0), T(0), u2_contracted<T>(ux, T(0), T(0)));
                } else {
                    const T u2 = ((ux * ux) + (T(0) * T(0))) + (T(0) * T(0));
                    v = feq<T>(i, T(1), ux, T(0), T(0), u2);
                }
            }
        }
        const int64_t s = slot(c, i, a.stride);
        if constexpr (KIND != ST_NATIVE) {
            const S h = codec.encode_delta(v);
            __builtin_nontemporal_store(h, pa + s);
            __builtin_nontemporal_store(h, pb + s);
        } else {
            codec.store(pa + s, w_of<T>(i), v);
            codec.store(pb + s, w_of<T>(i), v);
        }
    }
}

// The cell flags (common.h:7-20), materialised only when a caller reads the
// map: the compute kernels derive them from coordinates in registers.
static __global__ __launch_bounds__(256) void lbm_map_kernel(int32_t* map, int64_t n3, int32_t n, int32_t log2n) {
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= n3) return;
    const int32_t x = int32_t(c & (n - 1));
    const int32_t y = int32_t((c >> log2n) & (n - 1));
    const int32_t z = int32_t(c >> (2 * log2n));
    map[c] = cell_type(x, y, z, n);
}

// ---------------------------------------------------------------------------
// Push view for lbm_read_f: out[slot(d, i)] = P[slot(d - e_i, i)] decoded,
// NaN where d - e_i is outside the lattice.
// ---------------------------------------------------------------------------
template <class T> struct ViewArgs {
    const void* p;
    T* out;
    int64_t n3;
    int64_t stride;      // CSoA stride of the lattice read
    int64_t out_stride;  // CSoA stride of the view written (1: cell major)
    int32_t n;
    int32_t log2n;
    float scale, inv_scale;
};

template <class T, int KIND>
__global__ __launch_bounds__(256) void lbm_view_kernel(ViewArgs<T> a) {
    using C = Codec<T, KIND>;
    using S = typename C::S;
    const C codec{a.scale, a.inv_scale};
    const int32_t n = a.n;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= a.n3) return;
    const int32_t x = int32_t(c & (n - 1));
    const int32_t y = int32_t((c >> a.log2n) & (n - 1));
    const int32_t z = int32_t(c >> (2 * a.log2n));
    const S* p = static_cast<const S*>(a.p);
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int32_t sx = x - ex_of(i), sy = y - ey_of(i), sz = z - ez_of(i);
        T v = __builtin_nan("");
        if (sx >= 0 && sy >= 0 && sz >= 0 && sx < n && sy < n && sz < n) {
            const int64_t s = int64_t(sx) + int64_t(n) * (int64_t(sy) + int64_t(n) * sz);
            v = codec.load(p + slot(s, i, a.stride), w_of<T>(i));
        }
        a.out[slot(c, i, a.out_stride)] = v;
    }
}

// rho and u of the VTI window x, y, z in [1, N-2] (W = N - 2, x fastest):
// out[k] = rho, out[W^3 + 3k + d] = u_d, for window point k.
template <class T> struct PackArgs {
    const T* rho;
    const T* u;
    T* out;
    int32_t n;
    int32_t w;
    int64_t n3;
    int64_t w3;
};

template <class T>
__global__ __launch_bounds__(256) void lbm_pack_macro_kernel(PackArgs<T> a) {
    const int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= a.w3) return;
    const int64_t w = a.w;
    const int64_t x = k % w + 1, y = (k / w) % w + 1, z = k / (w * w) + 1;
    const int64_t c = x + int64_t(a.n) * (y + int64_t(a.n) * z);
    a.out[k] = a.rho[c];
    T* v = a.out + a.w3 + 3 * k;
    v[0] = a.u[c];
    v[1] = a.u[a.n3 + c];
    v[2] = a.u[2 * a.n3 + c];
}

// Final reduction of the per-block (mass, energy) partials, in a fixed order:
// thread j sums blocks j, j + 1024, ... then a tree over the 1024 threads.
static __global__ __launch_bounds__(1024) void lbm_reduce_partials(const double* partials, int64_t nblocks, double* out) {
    __shared__ double sm[1024];
    __shared__ double se[1024];
    double m = 0.0, e = 0.0;
    for (int64_t b = threadIdx.x; b < nblocks; b += 1024) {
        m += partials[2 * b];
        e += partials[2 * b + 1];
    }
    sm[threadIdx.x] = m;
    se[threadIdx.x] = e;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (int(threadIdx.x) < s) {
            sm[threadIdx.x] += sm[threadIdx.x + s];
            se[threadIdx.x] += se[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = sm[0];
        out[1] = se[0];
    }
}

}  // namespace lbm
