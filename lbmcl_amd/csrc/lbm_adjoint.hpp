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
    T* grad_acc;       // 3 N^3 dJ/dg accumulator of the arithmetic type (may be null); FP32 runs
                       // flush it into the FP64 sum every ADJ_FLUSH steps
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
    int32_t xcd_remap;
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

// Local vector-Jacobian product of one non-wall cell of type t in plane z.
// f: the populations the cell gathered (the lid fix-up is applied here); L:
// lambda_{k+1} at the cell's own slots (unused when seeding).  Writes lam
// (dJ/dG_k of the gathered populations) and the cell's parameter terms:
// dU and dW (d/domega) and J are set, gg = d/dg of the cell.
template <class T>
__device__ __forceinline__ void adjoint_cell(const AdjArgs<T>& a, int32_t t, int32_t z, T (&f)[Q], const T (&L)[Q],
                                             T (&lam)[Q], double& dU, double& dW, double& J, double (&gg)[3]) {
    dU = dW = J = 0.0;
    gg[0] = gg[1] = gg[2] = 0.0;
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
        return;
    }
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
        for (int i = 0; i < Q; ++i) lam[i] = (T(1) - om) * L[i] + om * (A + (edot<T>(i, Bx, By, Bz) - Bu) / rho);
        dW = double(dw);
        gg[0] = ggx;
        gg[1] = ggy;
        gg[2] = ggz;
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

// dJ/dU through the gathers of cell (x, y, z) that read initial values: from
// wall cells, and at the first step from every cell.  Only lid-layer cells
// (FRONT) gather initial values that depend on U.
template <class T>
__device__ __forceinline__ void adjoint_initial_sources(const AdjArgs<T>& a, int32_t t, int32_t x, int32_t y,
                                                        int32_t z, const T (&lam)[Q], double& dU) {
    if (!(t & FRONT)) return;
    const int32_t n = a.n;
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const bool wall_src = wall_coord(x - ex_of(i), y - ey_of(i), z - ez_of(i), n);
        if (wall_src || a.first) dU += double(lam[i]) * double(dF0_dU<T>(i, a.U));
    }
}

// Block sums of (dU, dW, J) in FP64 added to the block's own partials slot
// (one slot per block, accumulated over all steps of the sweep; the order of
// every sum is fixed).  Blocks are whole wavefronts.
__device__ __forceinline__ void block_accumulate3(double a, double b, double c, double* lds, double* slot) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        a += __shfl_xor(a, off, 64);
        b += __shfl_xor(b, off, 64);
        c += __shfl_xor(c, off, 64);
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nwaves = blockDim.x >> 6;
    if (lane == 0) {
        lds[3 * wave] = a;
        lds[3 * wave + 1] = b;
        lds[3 * wave + 2] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double sa = lds[0], sb = lds[1], sc = lds[2];
        for (int w = 1; w < nwaves; ++w) {
            sa += lds[3 * w];
            sb += lds[3 * w + 1];
            sc += lds[3 * w + 2];
        }
        slot[0] += sa;
        slot[1] += sb;
        slot[2] += sc;
    }
}

// Generic adjoint step: one thread per cell of a 1-D grid over all N^3 cells
// (block size 256), 64-bit slot arithmetic per population.
template <class T>
__global__ __launch_bounds__(256) void lbm_adjoint_step(AdjArgs<T> a) {
    __shared__ double lds[3 * 4];
    const int32_t n = a.n;
    const int64_t c = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool inside = c < a.n3;
    const int32_t x = int32_t(c & (n - 1));
    const int32_t y = int32_t((c >> a.log2n) & (n - 1));
    const int32_t z = int32_t(c >> (2 * a.log2n));
    const int32_t t = inside ? cell_type(x, y, z, n) : WALL;
    double dU = 0.0, dW = 0.0, J = 0.0;
    if (t != WALL) {
        T f[Q], L[Q], lam[Q];
        double gg[3];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int64_t s = c - (ex_of(i) + int64_t(n) * (ey_of(i) + int64_t(n) * ez_of(i)));
            f[i] = a.P[slot(s, i, a.stride)];
        }
#pragma unroll
        for (int j = 0; j < Q; ++j) L[j] = a.seed ? T(0) : a.Lnext[slot(c, j, a.stride)];
        adjoint_cell<T>(a, t, z, f, L, lam, dU, dW, J, gg);
        if (a.grad_acc && t == FLUID && !a.seed) {
            a.grad_acc[c] += T(gg[0]);
            a.grad_acc[a.n3 + c] += T(gg[1]);
            a.grad_acc[2 * a.n3 + c] += T(gg[2]);
        }
        adjoint_initial_sources<T>(a, t, x, y, z, lam, dU);
        // scatter to the gather positions; own slots whose destination is a
        // wall are read by nobody and get zero
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (!wall_coord(x - ex_of(i), y - ey_of(i), z - ez_of(i), n)) {
                const int64_t s = c - (ex_of(i) + int64_t(n) * (ey_of(i) + int64_t(n) * ez_of(i)));
                a.Lprev[slot(s, i, a.stride)] = lam[i];
            }
            if (wall_coord(x + ex_of(i), y + ey_of(i), z + ez_of(i), n)) a.Lprev[slot(c, i, a.stride)] = T(0);
        }
    }
    block_accumulate3(dU, dW, J, lds, a.partials + 3 * blockIdx.x);
}

// Vector adjoint step (V = 2 cells per lane in FP32 when the row has at least
// 128 cells and the stride is at least 2, else 1).  One block per interior row
// (y, z), N / V threads, so every cross-lane move stays inside the block.  The
// gather of P_k is the forward vector kernel's (aligned V-element loads, DPP
// shifts by one lane, wavefront ends through LDS); lambda_{k+1} is read at the
// lane's own aligned slots.  The scatter
// lambda_k[d - e_i][i] = lam_i(d) is the gather reversed: the vector written at
// the lane's aligned position of population i's source row holds lam_i of the
// cells x + e_x, moved by one lane the other way, and the element that crosses
// a wavefront end comes from the neighbouring wavefront through LDS.  Wall
// cells of the row scatter zeros, which is what the slots they feed must hold;
// slots fed by wall rows are never written and keep the zeros both adjoint
// lattices start with.
template <class T, int V>
__global__ __launch_bounds__(512) void lbm_adjoint_vec(AdjArgs<T> a) {
    typedef T SV __attribute__((ext_vector_type(V)));
    __shared__ T xup[8][5];  // per wavefront: lane 0's lam_i of its first cell, e_x(i) = +1
    __shared__ T xdn[8][5];  // per wavefront: lane 63's lam_i of its last cell, e_x(i) = -1
    __shared__ T pe_up[8][5];  // per wavefront: lane 63's last gathered-row element, e_x(i) = +1
    __shared__ T pe_dn[8][5];  // per wavefront: lane 0's first gathered-row element, e_x(i) = -1
    __shared__ double lds[3 * 8];
    const int32_t n = a.n;
    const uint32_t rows = uint32_t(n - 2);
    const uint32_t lb = xcd_logical_block(blockIdx.x, gridDim.x, a.xcd_remap);
    const int32_t y = int32_t(lb % rows) + 1;
    const int32_t z = int32_t(lb / rows) + 1;
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    const uint32_t nwaves = blockDim.x >> 6;
    const int32_t x0 = int32_t(tid * V);
    const int32_t xw = __builtin_amdgcn_readfirstlane(int32_t((tid & ~63u) * V));
    const int64_t r = int64_t(y) + int64_t(n) * z;
    const int32_t rf = row_flags(y, z, n);
    const int64_t smask = a.stride - 1;
    const uint32_t smask32 = uint32_t(smask);
    auto voff = [&](int32_t xx) -> uint32_t {  // v(x) = slot(N r + x) - slot(N r)
        const uint32_t u = uint32_t(xx);
        return 19u * u - 18u * (u & smask32);
    };
    auto row_slot = [&](int i) -> int64_t {  // slot 0 of population i's source row (= its scatter row)
        const int64_t c0 = int64_t(n) * (r - ey_of(i) - int64_t(n) * ez_of(i));
        return 19 * c0 - 18 * (c0 & smask) + int64_t(i) * a.stride;
    };
    const int64_t oc0 = int64_t(n) * r;
    const int64_t own = 19 * oc0 - 18 * (oc0 & smask);
    const uint32_t off0 = voff(x0);
    const bool has_left = xw > 0;            // wave > 0 (one block per row)
    const bool has_right = xw + 64 * V < n;  // wave < nwaves - 1

    SV A[Q];
#pragma unroll
    for (int i = 0; i < Q; ++i)  // each slot is read once per step: streaming loads
        A[i] = __builtin_nontemporal_load(reinterpret_cast<const SV*>(a.P + row_slot(i) + off0));
    SV Lv[Q];
#pragma unroll
    for (int j = 0; j < Q; ++j) {
        if (a.seed) {
#pragma unroll
            for (int k = 0; k < V; ++k) Lv[j][k] = T(0);
        } else {
            Lv[j] = __builtin_nontemporal_load(reinterpret_cast<const SV*>(a.Lnext + own + int64_t(j) * a.stride + off0));
        }
    }
    // wavefront-end elements of the gather from the neighbouring wavefront
    // (the block is the whole row: the row's ends are walls, which need none)
    {
        int ku = 0, kd = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (ex_of(i) == 1) {
                if (lane == 63) pe_up[wave][ku] = A[i][V - 1];
                ++ku;
            } else if (ex_of(i) == -1) {
                if (lane == 0) pe_dn[wave][kd] = A[i][0];
                ++kd;
            }
        }
    }
    __syncthreads();
    T edge[Q];
    {
        int ku = 0, kd = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            edge[i] = T(0);
            if (ex_of(i) == 1) {
                if (has_left) edge[i] = pe_up[wave - 1][ku];
                ++ku;
            } else if (ex_of(i) == -1) {
                if (has_right) edge[i] = pe_dn[wave + 1][kd];
                ++kd;
            }
        }
    }
    T g[Q][V];
#pragma unroll
    for (int i = 0; i < Q; ++i) {
        const int ex = ex_of(i);
        if (ex == 0) {
#pragma unroll
            for (int j = 0; j < V; ++j) g[i][j] = A[i][j];
        } else if (ex == 1) {
            g[i][0] = lane_shift<T>(A[i][V - 1], edge[i], true);
#pragma unroll
            for (int j = 1; j < V; ++j) g[i][j] = A[i][j - 1];
        } else {
#pragma unroll
            for (int j = 0; j < V - 1; ++j) g[i][j] = A[i][j + 1];
            g[i][V - 1] = lane_shift<T>(A[i][0], edge[i], false);
        }
    }

    T lam[Q][V];
    double dU = 0.0, dW = 0.0, J = 0.0;
    double gv[3][V];
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const int32_t t = finish_type(x_flags(x0 + j, n) | rf);
        T f[Q], L[Q], lc[Q];
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            f[i] = g[i][j];
            L[i] = Lv[i][j];
            lc[i] = T(0);
        }
        double cu = 0.0, cw = 0.0, cj = 0.0, gg[3] = {0.0, 0.0, 0.0};
        if (t != WALL) {
            adjoint_cell<T>(a, t, z, f, L, lc, cu, cw, cj, gg);
            adjoint_initial_sources<T>(a, t, x0 + j, y, z, lc, cu);
        }
        dU += cu;
        dW += cw;
        J += cj;
#pragma unroll
        for (int i = 0; i < Q; ++i) lam[i][j] = lc[i];
        const bool gfl = (t == FLUID) && !a.seed;
#pragma unroll
        for (int k = 0; k < 3; ++k) gv[k][j] = gfl ? gg[k] : 0.0;
    }
    if (a.grad_acc && !a.seed) {
        typedef T DV __attribute__((ext_vector_type(V)));
        const int64_t c = int64_t(x0) + oc0;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            DV* p = reinterpret_cast<DV*>(a.grad_acc + int64_t(k) * a.n3 + c);
            DV v = *p;
#pragma unroll
            for (int j = 0; j < V; ++j) v[j] += T(gv[k][j]);
            *p = v;
        }
    }

    // wavefront-end elements of the scatter through LDS
    {
        int ku = 0, kd = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            if (ex_of(i) == 1) {
                if (lane == 0) xup[wave][ku] = lam[i][0];
                ++ku;
            } else if (ex_of(i) == -1) {
                if (lane == 63) xdn[wave][kd] = lam[i][V - 1];
                ++kd;
            }
        }
    }
    __syncthreads();
    {
        int ku = 0, kd = 0;
#pragma unroll
        for (int i = 0; i < Q; ++i) {
            const int ex = ex_of(i);
            SV o;
            if (ex == 0) {
#pragma unroll
                for (int j = 0; j < V; ++j) o[j] = lam[i][j];
            } else if (ex == 1) {  // position x holds lam_i(x + 1)
                const T e = (wave + 1 < nwaves) ? xup[wave + 1][ku] : T(0);
#pragma unroll
                for (int j = 0; j < V - 1; ++j) o[j] = lam[i][j + 1];
                o[V - 1] = lane_shift<T>(lam[i][0], e, false);
                ++ku;
            } else {  // position x holds lam_i(x - 1)
                const T e = (wave > 0) ? xdn[wave - 1][kd] : T(0);
                o[0] = lane_shift<T>(lam[i][V - 1], e, true);
#pragma unroll
                for (int j = 1; j < V; ++j) o[j] = lam[i][j - 1];
                ++kd;
            }
            __builtin_nontemporal_store(o, reinterpret_cast<SV*>(a.Lprev + row_slot(i) + off0));
        }
    }
    block_accumulate3(dU, dW, J, lds, a.partials + 3 * lb);
}

// Steps between flushes of an FP32 dJ/dg accumulator into the FP64 sum.  The
// flushes follow the global step index (k % ADJ_FLUSH == 0), so results do
// not depend on the checkpoint interval.
constexpr int ADJ_FLUSH = 32;

// g[i] += s[i], s[i] = 0 for i < n (FP32 accumulator into the FP64 sum)
static __global__ __launch_bounds__(256) void lbm_adjoint_flush(float* s, double* g, int64_t n) {
    const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        g[i] += double(s[i]);
        s[i] = 0.0f;
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
