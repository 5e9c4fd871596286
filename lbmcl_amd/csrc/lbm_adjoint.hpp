// lbm_adjoint.hpp -- reverse (adjoint) kernels of the collide-stream step.
//
// Forward model, in pull form: P_{k+1}[d] = Phi_d(G_k[d]; U, omega, g_d) for
// every non-wall cell d, with G_k[d][i] = P_k[d - e_i][i] (wall cells keep
// their initial values, so a gather from a wall reads the stale constant
// F0[d][i], which depends on U through the lid row's initial equilibrium).
// The adjoint state lambda_k has the layout of P_k: lambda_k[s][i] =
// dJ/dP_k[s][i].  One backward step, for each non-wall cell d:
//   L'_j    = lambda_{k+1}[d][j]                     (own slots, aligned)
//   lam_i   = sum_j L'_j dPhi_j / dG_i                (local vector-Jacobian product)
//   lambda_k[d - e_i][i] = lam_i                     (scatter = the forward gather)
// plus the parameter terms, accumulated in FP64:
//   dJ/domega += sum_j L'_j (feq_j - f_j)            (FLUID)
//   dJ/dU     += sum_j L'_j rho w_j dc_j/dU           (lid cells, feq(rho, (U,0,0)))
//             += lam_i dF0_i(d)/dU                    (gathers from walls; at k = 0 all gathers)
//   dJ/dg_d   += 3 sum_j L'_j w_j e_j                 (FLUID, body force)
// Local Jacobians, with f = G, rho = sum f, u = sum e f / rho:
//   FLUID   feq_j = rho w_j (1 + 3 e_j.u + 4.5 (e_j.u)^2 - 1.5 u^2), out_j = f_j + omega (feq_j - f_j)
//           lam_i = (1 - omega) L'_i + omega (A + (B.e_i - B.u) / rho),
//           A = sum_j L'_j feq_j / rho,  B_a = sum_j L'_j rho w_j (3 e_ja + 9 (e_j.u) e_ja - 3 u_a)
//   MOVING  out_j = rho~ w_j c_j(U) with rho~ summed after the lid fix-up (kernels.cl:343-349):
//           lam_i = m_i sum_j L'_j w_j c_j, m_i = 2 for i in {6,15,16,17,18}, 0 for {5,11,12,13,14}, else 1
//   bounce-back lam_i = L'_opp(i);  corner lam_i = L'_i
// Own slots whose destination is a wall are read by nobody: set to zero.
#pragma once

#include "lbm_kernels.hpp"

namespace lbm {

enum Objective { OBJ_KINETIC_ENERGY = 0, OBJ_MEAN_UX_PLANE = 1 };

template <class T> struct AdjArgs {
    const T* P;        // forward state P_k (native storage, pull form)
    const T* Lnext;    // lambda_{k+1}
    T* Lprev;          // lambda_k
    double* grad_g;    // 3 N^3, accumulated (may be null)
    double* partials;  // 3 doubles per block: dJ/dU, dJ/domega, J
    int64_t n3;
    int64_t stride;
    int32_t n;
    int32_t log2n;
    T inv_tau;         // omega
    T U;
    int32_t seed;      // 1: lambda_T from the objective (no lambda_{k+1})
    int32_t first;     // 1: the step that reads the initial state P_0 (adds dP_0/dU)
    int32_t objective;
    int32_t plane_z;
    double plane_scale;  // 1 / number of cells averaged by OBJ_MEAN_UX_PLANE
};

__device__ __forceinline__ bool wall_coord(int32_t x, int32_t y, int32_t z, int32_t n) {
    return x == 0 || y == 0 || z == 0 || x == n - 1 || y == n - 1 || z == n - 1;
}

// d feq_i(1, (U, 0, 0)) / dU, the initial value of a lid-row population
template <class T> __device__ __forceinline__ T dF0_dU(int i, T U) {
    const T ex = T(ex_of(i));
    return w_of<T>(i) * ((T(3) * ex + T(9) * ex * ex * U) - T(3) * U);
}

__device__ __forceinline__ int lid_multiplicity(int i) {
    return (i == 6 || i == 15 || i == 16 || i == 17 || i == 18) ? 2
           : (i == 5 || i == 11 || i == 12 || i == 13 || i == 14) ? 0
                                                                   : 1;
}

__device__ __forceinline__ void block_reduce3(double& a, double& b, double& c, double* lds) {
    // serial LDS sum (deterministic for any block size)
    __syncthreads();
    lds[3 * threadIdx.x] = a;
    lds[3 * threadIdx.x + 1] = b;
    lds[3 * threadIdx.x + 2] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = 0.0, sb = 0.0, sc = 0.0;
        for (uint32_t j = 0; j < blockDim.x; ++j) {
            sa += lds[3 * j];
            sb += lds[3 * j + 1];
            sc += lds[3 * j + 2];
        }
        a = sa;
        b = sb;
        c = sc;
    }
}

template <class T>
__global__ __launch_bounds__(256) void lbm_adjoint_step(AdjArgs<T> a) {
    __shared__ double lds[3 * 256];
    const int32_t n = a.n;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool inside = c < a.n3;
    const int32_t x = int32_t(c & (n - 1));
    const int32_t y = int32_t((c >> a.log2n) & (n - 1));
    const int32_t z = int32_t(c >> (2 * a.log2n));
    const int32_t t = inside ? cell_type(x, y, z, n) : WALL;
    double dU = 0.0, dW = 0.0, J = 0.0;
    if (t != WALL) {
        T f[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int64_t s = c - (ex_of(i) + int64_t(n) * (ey_of(i) + int64_t(n) * ez_of(i)));
            f[i] = a.P[slot(s, i, a.stride)];
        }
        T lam[Q];
        const bool moving = is_moving(t);
        const T U = a.U;
        if (moving) {  // lid fix-up (kernels.cl:343-349)
            f[5] = f[6];
            f[11] = f[17];
            f[12] = f[18];
            f[13] = f[15];
            f[14] = f[16];
        }
        T rho = T(0), jx = T(0), jy = T(0), jz = T(0);
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            rho += f[i];
            jx += T(ex_of(i)) * f[i];
            jy += T(ey_of(i)) * f[i];
            jz += T(ez_of(i)) * f[i];
        }
        const T ux = jx / rho, uy = jy / rho, uz = jz / rho;
        if (a.seed) {
            // lambda_T = dJ / dG_T for the objective
            bool any = false;
            if (a.objective == OBJ_KINETIC_ENERGY && is_store_macro(t)) {
                if (moving) {
                    J = 0.5 * double(rho) * double(U) * double(U);
                    dU = double(rho) * double(U);
#pragma unroll
                    for (int i = 0; i < Q; ++i) lam[i] = T(lid_multiplicity(i)) * (T(0.5) * U * U);
                } else {
                    const T u2 = ux * ux + uy * uy + uz * uz;
                    J = 0.5 * double(rho) * double(u2);
#pragma unroll
                    for (int i = 0; i < Q; ++i) lam[i] = edot<T>(i, ux, uy, uz) - T(0.5) * u2;
                }
                any = true;
            } else if (a.objective == OBJ_MEAN_UX_PLANE && is_store_macro(t) && z == a.plane_z) {
                const T sc = T(a.plane_scale);
                if (moving) {
                    J = a.plane_scale * double(U);
                    dU = a.plane_scale;
#pragma unroll
                    for (int i = 0; i < Q; ++i) lam[i] = T(0);
                } else {
                    J = a.plane_scale * double(ux);
#pragma unroll
                    for (int i = 0; i < Q; ++i) lam[i] = sc * (T(ex_of(i)) - ux) / rho;
                }
                any = true;
            }
            if (!any) {
#pragma unroll
                for (int i = 0; i < Q; ++i) lam[i] = T(0);
            }
        } else {
            T L[Q];
#pragma unroll
            for (int j = 0; j < Q; ++j) L[j] = a.Lnext[slot(c, j, a.stride)];
            const T om = a.inv_tau;
            if (t == FLUID) {
                const T u2 = ux * ux + uy * uy + uz * uz;
                T A = T(0), Bx = T(0), By = T(0), Bz = T(0), dw = T(0);
                double ggx = 0.0, ggy = 0.0, ggz = 0.0;
#pragma unroll
                for (int j = 0; j < Q; ++j) {
                    const T eu = edot<T>(j, ux, uy, uz);
                    const T w = w_of<T>(j);
                    const T feq = rho * w * (T(1) + T(3) * eu + T(4.5) * eu * eu - T(1.5) * u2);
                    A += L[j] * feq;
                    const T q = L[j] * rho * w;
                    const T k9 = T(9) * eu;
                    Bx += q * ((T(3) + k9) * T(ex_of(j)) - T(3) * ux);
                    By += q * ((T(3) + k9) * T(ey_of(j)) - T(3) * uy);
                    Bz += q * ((T(3) + k9) * T(ez_of(j)) - T(3) * uz);
                    dw += L[j] * (feq - f[j]);
                    const double lw = 3.0 * double(L[j]) * double(w);
                    ggx += lw * double(ex_of(j));
                    ggy += lw * double(ey_of(j));
                    ggz += lw * double(ez_of(j));
                }
                A = A / rho;
                const T Bu = Bx * ux + By * uy + Bz * uz;
#pragma unroll
                for (int i = 0; i < Q; ++i)
                    lam[i] = (T(1) - om) * L[i] + om * (A + (edot<T>(i, Bx, By, Bz) - Bu) / rho);
                dW = double(dw);
                if (a.grad_g) {
                    a.grad_g[c] += ggx;
                    a.grad_g[a.n3 + c] += ggy;
                    a.grad_g[2 * a.n3 + c] += ggz;
                }
            } else if (moving) {
                T S = T(0), dUl = T(0);
#pragma unroll
                for (int j = 0; j < Q; ++j) {
                    const T ex = T(ex_of(j));
                    const T w = w_of<T>(j);
                    const T cj = T(1) + T(3) * ex * U + T(4.5) * ex * ex * U * U - T(1.5) * U * U;
                    S += L[j] * w * cj;
                    dUl += L[j] * rho * w * ((T(3) * ex + T(9) * ex * ex * U) - T(3) * U);
                }
#pragma unroll
                for (int i = 0; i < Q; ++i) lam[i] = T(lid_multiplicity(i)) * S;
                dU = double(dUl);
            } else if (is_bounceback(t)) {
                constexpr int OPPI[Q] = {0, 3, 4, 1, 2, 6, 5, 9, 10, 7, 8, 17, 18, 15, 16, 13, 14, 11, 12};
#pragma unroll
                for (int i = 0; i < Q; ++i) lam[i] = L[OPPI[i]];
            } else {  // corner
#pragma unroll
                for (int i = 0; i < Q; ++i) lam[i] = L[i];
            }
        }
        // scatter to the gather positions; gathers from walls and (first step)
        // from the initial lattice depend on U through F0[d][i]
        const bool front = (t & FRONT) != 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int32_t sx = x - ex_of(i), sy = y - ey_of(i), sz = z - ez_of(i);
            const bool wall_src = wall_coord(sx, sy, sz, n);
            if (front && (wall_src || a.first)) dU += double(lam[i]) * double(dF0_dU<T>(i, U));
            if (!wall_src) {
                const int64_t s = c - (ex_of(i) + int64_t(n) * (ey_of(i) + int64_t(n) * ez_of(i)));
                a.Lprev[slot(s, i, a.stride)] = lam[i];
            }
            const int32_t dx = x + ex_of(i), dy = y + ey_of(i), dz = z + ez_of(i);
            if (wall_coord(dx, dy, dz, n)) a.Lprev[slot(c, i, a.stride)] = T(0);
        }
    }
    block_reduce3(dU, dW, J, lds);
    if (threadIdx.x == 0) {
        a.partials[3 * blockIdx.x] = dU;
        a.partials[3 * blockIdx.x + 1] = dW;
        a.partials[3 * blockIdx.x + 2] = J;
    }
}

// acc[0..2] += sum of the per-block partials, in a fixed order
static __global__ __launch_bounds__(1024) void lbm_adjoint_reduce(const double* partials, int64_t nblocks, double* acc) {
    __shared__ double s[3][1024];
    double v[3] = {0.0, 0.0, 0.0};
    for (int64_t b = threadIdx.x; b < nblocks; b += 1024)
        for (int k = 0; k < 3; ++k) v[k] += partials[3 * b + k];
    for (int k = 0; k < 3; ++k) s[k][threadIdx.x] = v[k];
    __syncthreads();
    for (int st = 512; st > 0; st >>= 1) {
        if (int(threadIdx.x) < st)
            for (int k = 0; k < 3; ++k) s[k][threadIdx.x] += s[k][threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 3; ++k) acc[k] += s[k][0];
}

}  // namespace lbm
