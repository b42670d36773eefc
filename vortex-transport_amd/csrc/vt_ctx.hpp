// vt_ctx.hpp -- internal: the per-mesh context behind the C ABI and the
// helpers shared by vt_api.hip and vt_batch.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <map>
#include <tuple>
#include <vector>

#include "vt_api.h"
#include "vt_kernels.hpp"
#include "vt_setup.hpp"

struct vt_ctx {
    int ne = 0, ld = 0, p = 0, np = 0, nq = 0, order = 0, device = 0;
    int pairs = 0;                           // side s = 1 of every element e faces e ^ 1
    int plain_order = 0;                     // bit i: plain tile order for tiles of 64 << i elements
    hipStream_t stream = nullptr;
    double *Un = nullptr, *Ua = nullptr, *Ub = nullptr, *acc = nullptr, *vtx = nullptr, *tab = nullptr;
    int* nbr = nullptr;
    unsigned long long* err = nullptr;       // device error key
    unsigned long long* err_init = nullptr;  // pinned host, holds kNoError
    unsigned long long* err_host = nullptr;  // pinned host, one slot per replay
    size_t err_host_cap = 0;
    double* l2ref = nullptr;                 // reference state planes of the L2 metric
    double* l2tab = nullptr;                 // basis at the metric's 25 points, weights
    double* l2part = nullptr;                // metric partial sums, [blocks][4]
    double* host_stage = nullptr;            // pinned host staging for plane transfers
    size_t host_stage_cap = 0;               // in doubles
    std::vector<double> vertices;
    double time = 0.0;
    vt::LaunchConfig cfg;
    std::map<std::tuple<int, unsigned long long, int>, hipGraphExec_t> graphs;
    int last_err[6] = {-1, -1, -1, 0, 0, 0};
    // Discrete adjoint (vt_adjoint.hip): checkpoints of the recorded forward
    // run, the stage states of one segment, and lam / acc / y ping-pong planes.
    std::vector<double*> ckpt;
    double* seg = nullptr;
    double* adj = nullptr;
    double* mref = nullptr;                  // M_ref [Np][Np] of the objective
    double* tgt = nullptr;                   // target density planes [Np][ld]
    double* part = nullptr;                  // objective partial sums, one per block
    int part_cap = 0;
    int seg_slots = 0;
    int rec_steps = 0, rec_every = 0;
    double rec_dt = 0.0;
    bool recorded = false;
    bool rec_last_stored = false;            // seg holds the last segment's stage states from the recording
    unsigned long long rec_id = 0;           // number of vt_adjoint_forward calls so far
    // A second segment slot set and a stream of its own: when both sets fit in
    // device memory, the reverse sweep recomputes the next segment's stage
    // states on rstream while the adjoint kernels sweep the current segment.
    bool adj_overlap = true;
    double* seg2 = nullptr;
    int seg2_slots = 0;
    hipStream_t rstream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_rec[2] = {nullptr, nullptr}, ev_swept[2] = {nullptr, nullptr};
    // Device set-up of the vortex state: area rule, basis values and Minv_ref
    // of the least-squares type (uploaded on first use), and the objective's
    // target vortex {t, eps, x0, y0} set by vt_set_target_vortex.
    double* vrule = nullptr;
    // Small-mesh path (vt_small.hpp): each shared side once, and whether
    // vt_step runs through it (small_ok: the mesh qualifies).
    int* faces = nullptr;
    int nfaces = 0, fbeg[4] = {0, 0, 0, 0};
    bool small_ok = false, small_on = false;
    bool tgt_vortex = false;
    double tgt_par[4] = {0.0, 0.0, 0.0, 0.0};
};

namespace vti {

// Records a failed HIP call for vt_last_hip_error(); true when e is hipSuccess.
bool hip_ok(hipError_t e, const char* what);

// Plane stride: element count rounded to 256 plus a 64-element skew, so the
// 4*Np planes of one element do not all start at the same HBM channel offset.
inline int plane_stride(int ne) { return ((ne + 255) / 256) * 256 + 64; }

// Whether the stage kernels of context c run in plain tile order for tiles
// of `block` elements (vt_ctx::plain_order, chosen at creation).
inline int plain_order_for(const vt_ctx* c, int block) {
    const int bit = block == 64 ? 0 : block == 128 ? 1 : block == 256 ? 2 : -1;
    return bit >= 0 ? (c->plain_order >> bit) & 1 : 0;
}

// Largest mesh of one context: kernels address an element within a plane by
// a 32-bit byte offset (e * 8) and the neighbour table by the int s * ld + e;
// plane bases are 64-bit, so a plane set may exceed 4 GiB.  2^28 elements keep
// both offsets below 2^31 (the device memory runs out first: 86 GB per state
// at p = 3).
constexpr int kMaxElements = 1 << 28;

// Graph steps per replay: one graph covers up to kGraphSteps steps.
constexpr int kGraphSteps = 64;

// Allocates and uploads one mesh (vertices [n][3][2], neighbours [n][3] R,L,V,
// types [n]) with the given operator table; state zeroed.
int create_ctx(int n, const double* vertices, const int* neighbors, const int* type, int p, const vt::OperatorTable& tab,
               vt_ctx** out);

// [n][Np][4] state <-> planes, for elements [e0, e0 + n) of the context.
void to_planes(const vt_ctx* c, const double* U, int e0, int n, double* planes);
void from_planes(const vt_ctx* c, const double* planes, int e0, int n, double* U);

// Pinned host staging of at least n doubles, kept by the context.
double* host_stage(vt_ctx* c, size_t n);

// Whole-state transfers through the staging buffer: [ne][Np][4] host state
// to / from a device plane set of the context (padding columns zeroed).
int upload_state(vt_ctx* c, const double* U, double* dev);
int download_state(vt_ctx* c, const double* dev, double* U);

// Steps per graph replay for nsteps steps (Euler replays hold an even count,
// except a single trailing step).
std::vector<int> replay_plan(int method, int nsteps);

// nsteps RK4 (method 1) or Euler (0) steps through graph replays, with
// optional events around them; VT_ERR_NONPHYSICAL locates the first failure.
int run_steps(vt_ctx* c, int method, double dt, int nsteps, hipEvent_t ev0, hipEvent_t ev1);

// Frees the adjoint recording of a context.
void free_adjoint(vt_ctx* c);

// The adjoint's lam / acc / y plane sets (allocated and zeroed once).
int ensure_adj_planes(vt_ctx* c);

// Objective of the adjoint over nr element ranges [e0[r], e0[r] + n[r]) of the
// current state, with target densities rho[r] [n[r]][Np] from the host, or,
// with rho null, the nodal density of the vortex vpar[r] = {t, eps, x0, y0}
// computed on the device: J[r] per range (block partial sums added in block
// order) and dJ/dU of every range written into the lam planes.
int objective(vt_ctx* c, int nr, const int* e0, const int* n, const double* const* rho, const double* vpar, double* J);

// The vortex state on the device for element ranges [e0[r], e0[r] + n[r])
// with parameters par[r] = {t, eps, x0, y0}: the type 0 or type 1 state into
// the state planes `out`, or, with density_only, the type 0 nodal density
// into the Np planes `out`.
int vortex_device(vt_ctx* c, int nr, const int* e0, const int* n, const double* par, int type, double* out,
                  bool density_only);

// The reverse sweep over the recorded forward run, from lam = dJ/dU_T in the
// lam planes to lam = dJ/dU_0 in the same planes.
int backward_sweep(vt_ctx* c);

// Fills last_err from an error key: element, point, side, code, the step
// (step_base + the key's stage / per_step + 1) and the stage within it.
void decode_error(vt_ctx* c, unsigned long long key, int step_base, int per_step);

// Pinned host error slots, at least n of them.
int ensure_err_slots(unsigned long long** slots, size_t* cap, size_t n);

// Ends a stream capture whose launches returned launch_err, instantiates the
// captured graph into *ex and uploads it to the device on the stream (the
// first replay of a graph otherwise pays for that upload), waiting for the
// upload.  The captured graph is destroyed in every case.
int finish_capture(hipStream_t st, hipError_t launch_err, hipGraphExec_t* ex);

}  // namespace vti

#define HIP_TRY(expr)                                        \
    do {                                                     \
        if (!vti::hip_ok((expr), #expr)) return VT_ERR_HIP;  \
    } while (0)
