// vt_kernels.hpp -- launch interface of the stage kernels (vt_stage.hpp,
// instantiated by vt_kernels_nq.hip, dispatched by vt_launch.hip).
#pragma once
#include <hip/hip_runtime.h>

namespace vt {

// Default elements (threads) per workgroup: one wavefront.  128 and 256 are
// the other supported tiles (more in-block neighbours, fewer workgroups).
constexpr int kBlock = 64;
constexpr int kMaxGroups = 4;

enum StageMode : int {
    kModeRhs = 0,    // Uout = dU/dt                         (evolve_elem)
    kModeRk1 = 1,    // Uout = Un + k1/2 (Uin = Un)                      (Timestepping.cpp:136-171)
    kModeRk2 = 2,    // Uout = Un + k2/2                                 (:178-213)
    kModeRk3 = 3,    // acc = Un + (k1 + 2 k2 + 2 k3)/6 with k1 = 2 (U2 - Un),
                     // k2 = 2 (Uin - Un); Uout = Un + k3                (:220-255)
    kModeRk4 = 4,    // Uout = acc + k4/6                                (:262-315)
    kModeEuler = 5,  // Uout = Uin + dt dU/dt                (Timestepping.cpp:50-93)
};

struct StageParams {
    const double* Uin;    // stage input, planes [4*NP][ld]
    const double* Un;     // step start state (modes 2, 3)
    const double* U2;     // the step's second stage state, Un + k1/2 (mode 3; may be Uout)
    double* acc;          // RK accumulator (modes 3, 4)
    double* Uout;         // stage output
    const double* vtx;    // vertices, planes [6][ld]: x0 y0 x1 y1 x2 y2
    const int* nbr;       // neighbour across side 1, 2, 3: planes [3][ld]
    const double* tab;    // packed operator table (vt_device.hpp)
    unsigned long long* err;
    double dt;
    int ne, ld, mode, stage_id;
    int blocks_per_xcd;   // tiles per XCD band: ceil(ceil(ne / block) / 8)
    int err_base;         // added to the element index of an error report
    int pairs;            // 1: the neighbour across side s = 1 of every element e is e ^ 1
    int plain_order;      // 1: workgroup b works on tile b; 0: XCD band order (stage_kernel)
};

// Launch shape.  block: elements per workgroup (128 or 256).  persistent:
// 8 * slots_per_xcd resident workgroups loop over the tiles; otherwise one
// workgroup per tile.
struct LaunchConfig {
    int block = kBlock;
    int persistent = 0;
    int slots_per_xcd = 0;
};

// Several element groups in one launch (instances packed by polynomial
// order, all with the same quadrature size).  Each group's tiles are split
// into 8 XCD bands of tpx[g] = ceil(tiles[g] / 8) tiles; XCD x walks the
// concatenation over groups of its bands, item k of that walk being tile
// x * tpx[g] + (k - kbegin[g]) of the group g with kbegin[g] <= k <
// kbegin[g + 1].  Every XCD so gets the same share of each polynomial order.
// g[i].blocks_per_xcd is unused here.
struct BatchParams {
    StageParams g[kMaxGroups];
    int p[kMaxGroups];
    int tiles[kMaxGroups];
    int tpx[kMaxGroups];
    int kbegin[kMaxGroups + 1];
    int ngroups;
};

// Discrete adjoint (vt_adjoint.hpp).  Modes: kAdjRaw writes g = f'(U)^T w to
// Yout; kAdj4..kAdj1 are the four stages of one reverse RK4 step; kAdjSeed
// writes Yout = scale * Minv_ref^T lam / detJ (no flux work).
enum AdjMode : int { kAdjRaw = 0, kAdj1 = 1, kAdj2 = 2, kAdj3 = 3, kAdj4 = 4, kAdjSeed = 5 };
constexpr int kAdjBlock = 64;

struct AdjParams {
    const double* U;      // stage state the right-hand side is linearised at
    const double* Y;      // y = Minv_ref^T w / detJ of the stage's cotangent w
    double* Yout;         // y of the next stage's cotangent (kAdjRaw: g)
    double* lam;          // dJ/dU_{n+1}; kAdj1 overwrites it with dJ/dU_n; kAdjSeed input
    double* acc;          // sum of the stage contributions
    const double* vtx;
    const int* nbr;
    const double* tab;
    double dt, scale;
    int ne, ld, mode, blocks_per_xcd;
    int pairs;            // 1: the neighbour across side s = 1 of every element e is e ^ 1
};

// Objective of the adjoint, J = sum_e detJ_e drho_e^T M_ref drho_e with
// drho = rho(U) - rho_target at the nodes: one thread per element writes
// lam = dJ/dU (density rows 2 detJ M_ref drho, momentum and energy rows 0)
// and each block of kObjBlock elements its partial sum of J to part[block].
constexpr int kObjBlock = 256;

struct ObjParams {
    const double* U;     // state planes
    const double* tgt;   // target density planes [Np][ld]
    const double* mref;  // M_ref [Np][Np], row-major
    const double* vtx;   // vertex planes
    double* lam;         // dJ/dU planes
    double* part;        // partial sums of J, one per block
    int ne, ld;
};

// L2 state-error metric (plottools/l2stateerror.py:31-99): per element the
// squared nodal difference of two states is interpolated to the 25 points of
// the order-10 area rule, weighted and scaled by detJ; each block of kL2Block
// elements writes its four component sums to part[4 * block + k].
constexpr int kL2Points = 25;
constexpr int kL2Block = 256;

struct L2Params {
    const double* U;     // state planes
    const double* Uref;  // reference state planes
    const double* vtx;   // vertex planes
    const double* tab;   // basis values [kL2Points][Np], then the rule's weights [kL2Points]
    double* part;        // [blocks][4]
    int ne, ld;
};

// Gradient of a scalar with respect to the vortex parameters (epsilon, x0,
// y0) of the initial state, given its gradient lam = dJ/dU0 on the device:
// per element the closed-form tangent dU0/d(eps, x0, y0) at the nodes dotted
// with lam; each block of kVGradBlock elements writes its three partial sums
// to part[3 * block + a].
constexpr int kVGradBlock = 256;

struct VGradParams {
    const double* lam;   // dJ/dU0 planes
    const double* vtx;   // vertex planes
    double* part;        // [blocks][3]
    double nodes[2 * 10];  // reference nodes (xi, eta)
    double t, eps, x0, y0;
    int ne, ld;
};

// Isentropic-vortex state computed on the device (Element.cpp:119-326): per
// element the type 0 nodal values, or the type 1 least-squares projection
// over the area rule, written into state planes; with density_only, only the
// nodal density into Np planes (an objective's target).  k holds the
// point-independent factors of the host formula (vt::vortex_constants).
constexpr int kVortexBlock = 128;

struct VortexParams {
    double* out;          // planes [4*Np][ld] (or [Np][ld] with density_only), element 0 of the range
    const double* vtx;    // vertex planes, element 0 of the range
    const double* rule;   // type 1: rule [nqa][3] (xi, eta, w), phi [Np][nqa], Minv_ref [Np][Np]
    double nodes[2 * 10];  // reference nodes (xi, eta)
    double k[8];
    double t, x0, y0;
    int ne, ld, type, nqa, density_only;
};

// Small-mesh path (vt_small.hpp): one workgroup of kSmallThreads threads
// runs nsteps steps of a whole mesh of at most kSmallMaxElements elements
// with the stage state in LDS.  faces lists each shared side once, as
// e | nb << 10 | s << 20 with e < nb the two elements and s the side (the
// same side index on both), ordered by s: faces[fbeg[s] .. fbeg[s + 1]).
// Un holds the state before and after the launch; the first failure goes to
// err with stage ids counted from 0 in the launch.
constexpr int kSmallThreads = 1024;
constexpr int kSmallItems = 512;                     // role B threads, one per (variable, element)
constexpr int kSmallMaxElements = kSmallItems / 4;

struct SmallParams {
    double* Un;           // state planes [4*NP][ld]
    const double* vtx;    // vertex planes
    const double* tab;    // packed operator table
    const int* faces;     // [nfaces]
    unsigned long long* err;
    double dt;
    int ne, ld, nfaces, nsteps, method;  // method: 1 RK4, 0 forward Euler
    int fbeg[4];
};

hipError_t launch_stage(const StageParams& prm, int p, int nq, const LaunchConfig& cfg, hipStream_t st);
hipError_t launch_adjoint(const AdjParams& prm, int p, int nq, hipStream_t st);
hipError_t launch_small(const SmallParams& prm, int p, int nq, hipStream_t st);
size_t small_lds_bytes(int p, int nq, int ne);
hipError_t launch_batched(const BatchParams& bp, int nq, const LaunchConfig& cfg, hipStream_t st);
hipError_t launch_flux(const double* in, double* out, int* codes, int n, hipStream_t st);
hipError_t launch_objective(const ObjParams& prm, int p, hipStream_t st);
hipError_t launch_l2(const L2Params& prm, int p, hipStream_t st);
hipError_t launch_vortex_grad(const VGradParams& prm, int p, hipStream_t st);
hipError_t launch_vortex_state(const VortexParams& prm, int p, hipStream_t st);

inline int tiles_per_xcd(int ne, int block) { return ((ne + block - 1) / block + 7) / 8; }
inline int blocks_per_xcd(int ne) { return tiles_per_xcd(ne, kBlock); }

}  // namespace vt
