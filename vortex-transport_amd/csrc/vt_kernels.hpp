// vt_kernels.hpp -- launch interface of the stage kernels (vt_kernels.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace vt {

#ifndef VT_BLOCK
#define VT_BLOCK 256
#endif
constexpr int kBlock = VT_BLOCK;

enum StageMode : int {
    kModeRhs = 0,    // Uout = dU/dt                         (evolve_elem)
    kModeRk1 = 1,    // acc = k1, Uout = Un + k1/2           (Timestepping.cpp:136-171)
    kModeRk2 = 2,    // acc += 2 k2, Uout = Un + k2/2        (:178-213)
    kModeRk3 = 3,    // acc = Un + (acc + 2 k3)/6, Uout = Un + k3  (:220-255)
    kModeRk4 = 4,    // Uout = acc + k4/6                    (:262-315)
    kModeEuler = 5,  // Uout = Uin + dt dU/dt                (Timestepping.cpp:50-93)
};

struct StageParams {
    const double* Uin;    // stage input, planes [4*NP][ld]
    const double* Un;     // step start state (modes 2, 3)
    double* acc;          // RK accumulator (modes 1-4)
    double* Uout;         // stage output
    const double* vtx;    // vertices, planes [6][ld]: x0 y0 x1 y1 x2 y2
    const int* nbr;       // neighbour across side 1, 2, 3: planes [3][ld]
    const double* tab;    // packed operator table (vt_device.hpp)
    unsigned long long* err;
    double dt;
    int ne, ld, mode, stage_id;
    int blocks_per_xcd;
};

hipError_t launch_stage(const StageParams& prm, int p, int nq, hipStream_t st);
hipError_t launch_flux(const double* in, double* out, int* codes, int n, hipStream_t st);

inline int blocks_per_xcd(int ne) { return ((ne + kBlock - 1) / kBlock + 7) / 8; }

}  // namespace vt
