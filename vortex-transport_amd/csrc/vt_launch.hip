// vt_launch.hip -- dispatch of the stage launchers over the quadrature size,
// and the batch Roe-flux kernel used to check the device flux on its own.
#include <hip/hip_runtime.h>

#include "vt_stage.hpp"

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

}  // namespace vt
