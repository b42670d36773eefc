// vt_launch.hip -- dispatch of the stage launchers over the quadrature size,
// and the batch Roe-flux kernel used to check the device flux on its own.
#include <hip/hip_runtime.h>

#include "vt_stage.hpp"
#include "vt_small.hpp"

namespace vt {

// Batch Roe flux, used to check the device flux function on its own.
__global__ void flux_kernel(const double* __restrict__ in, double* __restrict__ out, int* __restrict__ codes, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double* r = in + 10 * static_cast<size_t>(i);
    double UL[4] = {r[0], r[1], r[2], r[3]}, UR[4] = {r[4], r[5], r[6], r[7]}, F[4];
    codes[i] = roe_flux(UL, UR, r[8], r[9], F);
    for (int k = 0; k < 4; ++k) out[4 * static_cast<size_t>(i) + k] = F[k];
}

// Sum over the 64 lanes of a wavefront by a butterfly of lane swaps: at each
// level both lanes of a pair add the same two values, so every lane ends with
// the same bits, in an order fixed by the lane numbers.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// The block's sums of NV values per thread, into out[0..NV): each wavefront's
// sums by wave_sum, then the BS / 64 wavefront sums added in wavefront order
// (a fixed shape, so the result is the same run to run).  red holds NV * BS / 64
// doubles; every thread of the block calls this.
template <int NV, int BS>
__device__ __forceinline__ void block_sums(double (&v)[NV], double* red, double* out) {
    static_assert(BS % 64 == 0 && NV <= BS, "whole wavefronts, one output per thread");
    constexpr int NW = BS / 64;
    const int lane = static_cast<int>(threadIdx.x) & 63, wave = static_cast<int>(threadIdx.x) >> 6;
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < NV; ++k) red[k * NW + wave] = v[k];
    __syncthreads();
    if (static_cast<int>(threadIdx.x) < NV) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < NW; ++w) t += red[threadIdx.x * NW + w];
        out[threadIdx.x] = t;
    }
}

// The objective (see ObjParams).  Per element the sum runs over nodes in order,
// as (detJ drho_i) (M_ref drho)_i; the block sum is block_sums', so J is
// reproducible run to run.
template <int P>
__global__ void __launch_bounds__(kObjBlock) objective_kernel(ObjParams prm) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    __shared__ double red[kObjBlock / 64];
    const int e = blockIdx.x * kObjBlock + static_cast<int>(threadIdx.x);
    double je = 0.0;
    if (e < prm.ne) {
        const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
        const unsigned oe = static_cast<unsigned>(e) * 8u;
        const ctab_t M = const_table(prm.mref);
        const double x0 = plane_at(prm.vtx, 0, ld8, oe), y0 = plane_at(prm.vtx, 1, ld8, oe);
        const double x1 = plane_at(prm.vtx, 2, ld8, oe), y1 = plane_at(prm.vtx, 3, ld8, oe);
        const double x2 = plane_at(prm.vtx, 4, ld8, oe), y2 = plane_at(prm.vtx, 5, ld8, oe);
        const double det = (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0);
        double d[NP];
#pragma unroll
        for (int i = 0; i < NP; ++i) d[i] = plane_at(prm.U, i, ld8, oe) - plane_at(prm.tgt, i, ld8, oe);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            double a = 0.0;
#pragma unroll
            for (int j = 0; j < NP; ++j) a += M[i * NP + j] * d[j];
            je += det * d[i] * a;
            plane_at(prm.lam, i, ld8, oe) = 2.0 * det * a;
        }
#pragma unroll
        for (int v = NP; v < 4 * NP; ++v) plane_at(prm.lam, v, ld8, oe) = 0.0;
    }
    double v[1] = {je};
    block_sums<1, kObjBlock>(v, red, prm.part + blockIdx.x);
}

// The L2 state-error metric (see L2Params).  Per element and component:
// sum over the rule's points of w_i * sum_j phi_j(x_i) (U - Uref)_j^2, in that
// order, times detJ; the block sums are block_sums'.
template <int P>
__global__ void __launch_bounds__(kL2Block) l2_kernel(L2Params prm) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    __shared__ double red[4 * kL2Block / 64];
    const int e = blockIdx.x * kL2Block + static_cast<int>(threadIdx.x);
    double ik[4] = {0.0, 0.0, 0.0, 0.0};
    if (e < prm.ne) {
        const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
        const unsigned oe = static_cast<unsigned>(e) * 8u;
        const ctab_t B = const_table(prm.tab);
        const double x0 = plane_at(prm.vtx, 0, ld8, oe), y0 = plane_at(prm.vtx, 1, ld8, oe);
        const double x1 = plane_at(prm.vtx, 2, ld8, oe), y1 = plane_at(prm.vtx, 3, ld8, oe);
        const double x2 = plane_at(prm.vtx, 4, ld8, oe), y2 = plane_at(prm.vtx, 5, ld8, oe);
        const double det = (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            double du[NP];
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const double d = plane_at(prm.U, k * NP + j, ld8, oe) - plane_at(prm.Uref, k * NP + j, ld8, oe);
                du[j] = d * d;
            }
            double sum = 0.0;
#pragma unroll 5
            for (int i = 0; i < kL2Points; ++i) {
                double dq = 0.0;
#pragma unroll
                for (int j = 0; j < NP; ++j) dq += du[j] * B[i * NP + j];
                sum += B[kL2Points * NP + i] * dq;
            }
            ik[k] = det * sum;
        }
    }
    block_sums<4, kL2Block>(ik, red, prm.part + 4 * blockIdx.x);
}

// dJ/d(eps, x0, y0) (see VGradParams).  The tangent of the vortex initial
// state of Element.cpp:134-160 at each node, as the host's vortex_tangent
// forms it; per element the sum runs over nodes and variables in order, the
// block sums are block_sums'.
template <int P>
__global__ void __launch_bounds__(kVGradBlock) vortex_grad_kernel(VGradParams prm) {
    constexpr int NP = (P + 1) * (P + 2) / 2;
    __shared__ double red[3 * kVGradBlock / 64];
    const int e = blockIdx.x * kVGradBlock + static_cast<int>(threadIdx.x);
    double ga[3] = {0.0, 0.0, 0.0};
    if (e < prm.ne) {
        constexpr double gamma = 1.4, M_infty = 0.5, p_infty = 20.0 / 7.0, rc = 1.0;
        const double U_infty = 1.0 / sqrt(2.0), V_infty = 1.0 / sqrt(2.0);
        const double A = (gamma - 1) * M_infty * M_infty / (8.0 * M_PI * M_PI);
        const double B = (U_infty * U_infty + V_infty * V_infty) / (2.0 * M_PI * rc);
        const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
        const unsigned oe = static_cast<unsigned>(e) * 8u;
        const double vx0 = plane_at(prm.vtx, 0, ld8, oe), vy0 = plane_at(prm.vtx, 1, ld8, oe);
        const double vx1 = plane_at(prm.vtx, 2, ld8, oe), vy1 = plane_at(prm.vtx, 3, ld8, oe);
        const double vx2 = plane_at(prm.vtx, 4, ld8, oe), vy2 = plane_at(prm.vtx, 5, ld8, oe);
        const double eps = prm.eps;
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double xi = prm.nodes[2 * i], eta = prm.nodes[2 * i + 1], f = 1.0 - xi - eta;
            const double dx = (vx0 * f + vx1 * xi + vx2 * eta) - prm.x0 - U_infty * prm.t;
            const double dy = (vy0 * f + vy1 * xi + vy2 * eta) - prm.y0 - V_infty * prm.t;
            const double f0 = 1.0 - (dx * dx + dy * dy) / (rc * rc);
            const double ex = exp(f0), eh = exp(0.5 * f0);
            const double f1 = 1.0 - eps * eps * A * ex;
            const double f2 = eps * B * eh;
            const double rho = pow(f1, 1.0 / (gamma - 1.0)), pr = p_infty * pow(f1, gamma / (gamma - 1.0));
            const double u = U_infty - f2 * dy, v = V_infty + f2 * dx;
            const double df0[3] = {0.0, 2.0 * dx / (rc * rc), 2.0 * dy / (rc * rc)};
            double l[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) l[k] = plane_at(prm.lam, k * NP + i, ld8, oe);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double df1 = -(a == 0 ? 2.0 * eps : 0.0) * A * ex - eps * eps * A * ex * df0[a];
                const double df2 = (a == 0 ? B * eh : 0.0) + f2 * 0.5 * df0[a];
                const double drho = rho / ((gamma - 1.0) * f1) * df1;
                const double dp = pr * gamma / ((gamma - 1.0) * f1) * df1;
                const double ddx = a == 1 ? -1.0 : 0.0, ddy = a == 2 ? -1.0 : 0.0;
                const double du = -df2 * dy - f2 * ddy, dv = df2 * dx + f2 * ddx;
                ga[a] += drho * l[0];
                ga[a] += (drho * u + rho * du) * l[1];
                ga[a] += (drho * v + rho * dv) * l[2];
                ga[a] += (dp / (gamma - 1.0) + 0.5 * drho * (u * u + v * v) + rho * (u * du + v * dv)) * l[3];
            }
        }
    }
    block_sums<3, kVGradBlock>(ga, red, prm.part + 3 * blockIdx.x);
}

// The vortex of Element.cpp:134-160 at (x, y, t): rho, rho u, rho v, rho E,
// operation by operation as vt::vortex_point computes it on the host (the same
// operand order, contraction off, so each step rounds the same); exp and pow
// come from the device library (an ulp or so from the host's) and pow(a, 2.0)
// is a * a.
__device__ __forceinline__ void vortex_at(const VortexParams& prm, double x, double y, double (&U)[4]) {
#pragma clang fp contract(off)
    const double Uinf = prm.k[0], Vinf = prm.k[1];
    const double dx = x - prm.x0 - Uinf * prm.t, dy = y - prm.y0 - Vinf * prm.t;
    const double f0 = 1.0 - (dx * dx + dy * dy) / 1.0;
    const double f1 = 1.0 - prm.k[2] * exp(f0) / prm.k[3];
    const double f2 = prm.k[4] * exp(f0 / 2.0) / prm.k[5];
    const double rho = 1.0 * pow(f1, prm.k[6]);
    const double u = Uinf - f2 * dy;
    const double v = Vinf + f2 * dx;
    const double p = 20.0 / 7.0 * pow(f1, prm.k[7]);
    const double E = p / (rho * (1.4 - 1.0)) + (u * u + v * v) / 2.0;
    U[0] = rho;
    U[1] = rho * u;
    U[2] = rho * v;
    U[3] = rho * E;
}

template <int P>
__global__ void __launch_bounds__(kVortexBlock) vortex_state_kernel(VortexParams prm) {
#pragma clang fp contract(off)
    constexpr int NP = (P + 1) * (P + 2) / 2;
    const int e = blockIdx.x * kVortexBlock + static_cast<int>(threadIdx.x);
    if (e >= prm.ne) return;
    const size_t ld8 = static_cast<size_t>(prm.ld) * 8;
    const unsigned oe = static_cast<unsigned>(e) * 8u;
    const double x0 = plane_at(prm.vtx, 0, ld8, oe), y0 = plane_at(prm.vtx, 1, ld8, oe);
    const double x1 = plane_at(prm.vtx, 2, ld8, oe), y1 = plane_at(prm.vtx, 3, ld8, oe);
    const double x2 = plane_at(prm.vtx, 4, ld8, oe), y2 = plane_at(prm.vtx, 5, ld8, oe);
    if (prm.type == 0) {
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double xi = prm.nodes[2 * i], eta = prm.nodes[2 * i + 1];
            const double f = 1.0 - xi - eta;
            double U[4];
            vortex_at(prm, x0 * f + x1 * xi + x2 * eta, y0 * f + y1 * xi + y2 * eta, U);
            if (prm.density_only) {
                plane_at(prm.out, i, ld8, oe) = U[0];
            } else {
#pragma unroll
                for (int k = 0; k < 4; ++k) plane_at(prm.out, k * NP + i, ld8, oe) = U[k];
            }
        }
        return;
    }
    // Least squares (Element.cpp:180-304): b_i = sum_q detJ phi_i(q) U(q) w_q,
    // U_i = sum_j (Minv_ij / detJ) b_j, each sum in ascending order from 0.
    const ctab_t R = const_table(prm.rule);
    const ctab_t phi = R + 3 * prm.nqa;
    const ctab_t Minv = phi + NP * prm.nqa;
    const double detJ = (x1 - x0) * (y2 - y0) - (y1 - y0) * (x2 - x0);
    double b[NP][4];
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) b[i][c] = 0.0;
    for (int q = 0; q < prm.nqa; ++q) {
        const double xi = R[3 * q], eta = R[3 * q + 1], w = R[3 * q + 2];
        const double f = 1.0 - xi - eta;
        double U[4];
        vortex_at(prm, x0 * f + x1 * xi + x2 * eta, y0 * f + y1 * xi + y2 * eta, U);
#pragma unroll
        for (int i = 0; i < NP; ++i) {
            const double dp = detJ * phi[i * prm.nqa + q];
#pragma unroll
            for (int c = 0; c < 4; ++c) b[i][c] += dp * U[c] * w;
        }
    }
#pragma unroll
    for (int i = 0; i < NP; ++i)
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            double s = 0.0;
#pragma unroll
            for (int j = 0; j < NP; ++j) s += Minv[i * NP + j] / detJ * b[j][c];
            plane_at(prm.out, c * NP + i, ld8, oe) = s;
        }
}

#define VT_NQ_CASES(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10)

// The launchers (and their kernels) are instantiated in vt_kernels_nq.hip,
// one translation unit per quadrature size; none are instantiated here.
#define X(n) extern template struct NqLaunch<n>;
VT_NQ_CASES(X)
#undef X

hipError_t launch_stage(const StageParams& prm, int p, int nq, const LaunchConfig& cfg, hipStream_t st) {
    switch (nq) {
#define X(n) case n: return NqLaunch<n>::stage(prm, p, cfg, st);
        VT_NQ_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_batched(const BatchParams& bp, int nq, const LaunchConfig& cfg, hipStream_t st) {
    switch (nq) {
#define X(n) case n: return NqLaunch<n>::batched(bp, cfg, st);
        VT_NQ_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_small(const SmallParams& prm, int p, int nq, hipStream_t st) {
    switch (nq) {
#define X(n) case n: return NqLaunch<n>::small(prm, p, st);
        VT_NQ_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

size_t small_lds_bytes(int p, int nq, int ne) {
    return sizeof(double) * static_cast<size_t>(small_layout((p + 1) * (p + 2) / 2, p + 1, nq, small_pad(ne)).total);
}

hipError_t launch_adjoint(const AdjParams& prm, int p, int nq, hipStream_t st) {
    switch (nq) {
#define X(n) case n: return NqLaunch<n>::adjoint(prm, p, st);
        VT_NQ_CASES(X)
#undef X
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_flux(const double* in, double* out, int* codes, int n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(flux_kernel, dim3((n + 255) / 256), dim3(256), 0, st, in, out, codes, n);
    return hipGetLastError();
}

hipError_t launch_l2(const L2Params& prm, int p, hipStream_t st) {
    const dim3 grid((prm.ne + kL2Block - 1) / kL2Block), block(kL2Block);
    switch (p) {
    case 0: hipLaunchKernelGGL(l2_kernel<0>, grid, block, 0, st, prm); break;
    case 1: hipLaunchKernelGGL(l2_kernel<1>, grid, block, 0, st, prm); break;
    case 2: hipLaunchKernelGGL(l2_kernel<2>, grid, block, 0, st, prm); break;
    case 3: hipLaunchKernelGGL(l2_kernel<3>, grid, block, 0, st, prm); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_vortex_grad(const VGradParams& prm, int p, hipStream_t st) {
    const dim3 grid((prm.ne + kVGradBlock - 1) / kVGradBlock), block(kVGradBlock);
    switch (p) {
    case 0: hipLaunchKernelGGL(vortex_grad_kernel<0>, grid, block, 0, st, prm); break;
    case 1: hipLaunchKernelGGL(vortex_grad_kernel<1>, grid, block, 0, st, prm); break;
    case 2: hipLaunchKernelGGL(vortex_grad_kernel<2>, grid, block, 0, st, prm); break;
    case 3: hipLaunchKernelGGL(vortex_grad_kernel<3>, grid, block, 0, st, prm); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_vortex_state(const VortexParams& prm, int p, hipStream_t st) {
    if (prm.ne <= 0) return hipSuccess;
    const dim3 grid((prm.ne + kVortexBlock - 1) / kVortexBlock), block(kVortexBlock);
    switch (p) {
    case 0: hipLaunchKernelGGL(vortex_state_kernel<0>, grid, block, 0, st, prm); break;
    case 1: hipLaunchKernelGGL(vortex_state_kernel<1>, grid, block, 0, st, prm); break;
    case 2: hipLaunchKernelGGL(vortex_state_kernel<2>, grid, block, 0, st, prm); break;
    case 3: hipLaunchKernelGGL(vortex_state_kernel<3>, grid, block, 0, st, prm); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_objective(const ObjParams& prm, int p, hipStream_t st) {
    const dim3 grid((prm.ne + kObjBlock - 1) / kObjBlock), block(kObjBlock);
    switch (p) {
    case 0: hipLaunchKernelGGL(objective_kernel<0>, grid, block, 0, st, prm); break;
    case 1: hipLaunchKernelGGL(objective_kernel<1>, grid, block, 0, st, prm); break;
    case 2: hipLaunchKernelGGL(objective_kernel<2>, grid, block, 0, st, prm); break;
    case 3: hipLaunchKernelGGL(objective_kernel<3>, grid, block, 0, st, prm); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace vt
