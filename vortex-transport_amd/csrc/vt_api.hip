// vt_api.hip -- C ABI over the gfx950 kernels: device buffers, set-up upload,
// RK4 / forward-Euler stepping through captured HIP graphs, error reporting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "vt_ctx.hpp"
#include "vt_device.hpp"

namespace vti {

thread_local char g_hip_error[256] = "";

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    std::snprintf(g_hip_error, sizeof(g_hip_error), "%s: %s", what, hipGetErrorString(e));
    return false;
}

void to_planes(const vt_ctx* c, const double* U, int e0, int n, double* planes) {
    vt::aos_to_planes(n, c->np, c->ld, e0, U, planes);
}

void from_planes(const vt_ctx* c, const double* planes, int e0, int n, double* U) {
    vt::planes_to_aos(n, c->np, c->ld, e0, planes, U);
}

double* host_stage(vt_ctx* c, size_t n) {
    if (c->host_stage_cap >= n) return c->host_stage;
    if (c->host_stage) (void)hipHostFree(c->host_stage);
    c->host_stage = nullptr;
    c->host_stage_cap = 0;
    if (!hip_ok(hipHostMalloc(reinterpret_cast<void**>(&c->host_stage), sizeof(double) * n, hipHostMallocDefault),
                "hipHostMalloc staging"))
        return nullptr;
    c->host_stage_cap = n;
    return c->host_stage;
}

int upload_state(vt_ctx* c, const double* U, double* dev) {
    const size_t n = static_cast<size_t>(4) * c->np * c->ld;
    double* h = host_stage(c, n);
    if (!h) return VT_ERR_HIP;
    to_planes(c, U, 0, c->ne, h);
    for (int v = 0; v < 4 * c->np; ++v)
        std::memset(h + static_cast<size_t>(v) * c->ld + c->ne, 0, sizeof(double) * (c->ld - c->ne));
    HIP_TRY(hipMemcpyAsync(dev, h, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VT_OK;
}

int download_state(vt_ctx* c, const double* dev, double* U) {
    const size_t n = static_cast<size_t>(4) * c->np * c->ld;
    double* h = host_stage(c, n);
    if (!h) return VT_ERR_HIP;
    HIP_TRY(hipMemcpyAsync(h, dev, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    from_planes(c, h, 0, c->ne, U);
    return VT_OK;
}

std::vector<int> replay_plan(int method, int nsteps) {
    std::vector<int> reps;
    int left = nsteps;
    while (left > 0) {
        int k = left < kGraphSteps ? left : kGraphSteps;
        if (method == 0 && k > 1 && (k % 2)) k -= 1;
        reps.push_back(k);
        left -= k;
    }
    return reps;
}

int finish_capture(hipStream_t st, hipError_t launch_err, hipGraphExec_t* ex) {
    hipGraph_t g = nullptr;
    const hipError_t e2 = hipStreamEndCapture(st, &g);
    const bool ok = hip_ok(launch_err, "graph capture launch") && hip_ok(e2, "hipStreamEndCapture");
    if (!ok) {
        if (e2 == hipSuccess && g) (void)hipGraphDestroy(g);
        return VT_ERR_HIP;
    }
    const hipError_t e3 = hipGraphInstantiate(ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (!hip_ok(e3, "hipGraphInstantiate")) return VT_ERR_HIP;
    if (!hip_ok(hipGraphUpload(*ex, st), "hipGraphUpload") || !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize")) {
        (void)hipGraphExecDestroy(*ex);
        return VT_ERR_HIP;
    }
    return VT_OK;
}

int ensure_err_slots(unsigned long long** slots, size_t* cap, size_t n) {
    if (n <= *cap) return VT_OK;
    if (*slots) hipHostFree(*slots);
    *slots = nullptr;
    *cap = 0;
    HIP_TRY(hipHostMalloc(reinterpret_cast<void**>(slots), n * sizeof(unsigned long long), hipHostMallocDefault));
    *cap = n;
    return VT_OK;
}

}  // namespace vti

using vti::hip_ok;
using vti::plane_stride;
using vti::kGraphSteps;

namespace {

vt::StageParams base_params(vt_ctx* c) {
    vt::StageParams prm{};
    prm.vtx = c->vtx;
    prm.nbr = c->nbr;
    prm.tab = c->tab;
    prm.err = c->err;
    prm.ne = c->ne;
    prm.ld = c->ld;
    prm.blocks_per_xcd = vt::tiles_per_xcd(c->ne, c->cfg.block);
    prm.pairs = c->pairs;
    prm.plain_order = vti::plain_order_for(c, c->cfg.block);
    return prm;
}

// Launches the kernels of one RK4 step (4 stages) or one Euler step on the
// context stream.  Un holds the state before and after the step, except for
// an Euler step, which writes Un -> dst.
hipError_t enqueue_rk4_step(vt_ctx* c, double dt, int stage_base) {
    vt::StageParams prm = base_params(c);
    prm.dt = dt;
    struct S { const double* in; double* out; int mode; };
    const S st[4] = {{c->Un, c->Ua, vt::kModeRk1}, {c->Ua, c->Ub, vt::kModeRk2},
                     {c->Ub, c->Ua, vt::kModeRk3}, {c->Ua, c->Un, vt::kModeRk4}};
    for (int s = 0; s < 4; ++s) {
        prm.Uin = st[s].in;
        prm.Uout = st[s].out;
        prm.Un = (s == 1 || s == 2) ? c->Un : nullptr;
        prm.U2 = s == 2 ? c->Ua : nullptr;
        prm.acc = c->acc;
        prm.mode = st[s].mode;
        prm.stage_id = stage_base + s;
        hipError_t e = vt::launch_stage(prm, c->p, c->nq, c->cfg, c->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t enqueue_euler_step(vt_ctx* c, double dt, const double* src, double* dst, int stage_id) {
    vt::StageParams prm = base_params(c);
    prm.dt = dt;
    prm.Uin = src;
    prm.Uout = dst;
    prm.mode = vt::kModeEuler;
    prm.stage_id = stage_id;
    return vt::launch_stage(prm, c->p, c->nq, c->cfg, c->stream);
}

// Graph of k steps (Euler graphs hold an even number of steps, alternating
// Un -> Ua -> Un).
int get_graph(vt_ctx* c, int method, double dt, int k, hipGraphExec_t* out) {
    unsigned long long dbits;
    std::memcpy(&dbits, &dt, sizeof(dbits));
    auto key = std::make_tuple(method, dbits, k);
    auto it = c->graphs.find(key);
    if (it != c->graphs.end()) { *out = it->second; return VT_OK; }
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < k && e == hipSuccess; ++i) {
        if (method == 1) e = enqueue_rk4_step(c, dt, 4 * i);
        else e = enqueue_euler_step(c, dt, (i % 2) ? c->Ua : c->Un, (i % 2) ? c->Un : c->Ua, i);
    }
    hipGraphExec_t ex;
    const int rc = vti::finish_capture(c->stream, e, &ex);
    if (rc) return rc;
    c->graphs[key] = ex;
    *out = ex;
    return VT_OK;
}

int read_planes(vt_ctx* c, const double* dev, double* U) { return vti::download_state(c, dev, U); }

}  // namespace

void vti::decode_error(vt_ctx* c, unsigned long long key, int step_base, int per_step) {
    const unsigned long long stage = key >> 40;
    unsigned long long t = key & ((1ull << 40) - 1);
    c->last_err[3] = static_cast<int>(t % 4); t /= 4;
    c->last_err[2] = static_cast<int>(t % 4); t /= 4;
    c->last_err[1] = static_cast<int>(t % 16); t /= 16;
    c->last_err[0] = static_cast<int>(t);
    c->last_err[4] = step_base + static_cast<int>(stage / per_step) + 1;
    c->last_err[5] = static_cast<int>(stage % per_step) + 1;
}

int vti::run_steps(vt_ctx* c, int method, double dt, int nsteps, hipEvent_t ev0, hipEvent_t ev1) {
    if (method != 0 && method != 1) return VT_ERR_METHOD;
    if (nsteps < 0) return VT_ERR_ARG;
    const std::vector<int> reps = vti::replay_plan(method, nsteps);
    int rc = vti::ensure_err_slots(&c->err_host, &c->err_host_cap, reps.size() + 1);
    if (rc) return rc;
    std::vector<hipGraphExec_t> ex(reps.size(), nullptr);
    for (size_t r = 0; r < reps.size() && !c->small_on; ++r)
        if (!(method == 0 && reps[r] == 1)) {
            rc = get_graph(c, method, dt, reps[r], &ex[r]);
            if (rc) return rc;
        }
    if (ev0) HIP_TRY(hipEventRecord(ev0, c->stream));
    for (size_t r = 0; r < reps.size(); ++r) {
        HIP_TRY(hipMemcpyAsync(c->err, c->err_init, sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
        if (c->small_on) {  // the whole chunk in one launch, state in LDS
            vt::SmallParams sp{};
            sp.Un = c->Un;
            sp.vtx = c->vtx;
            sp.tab = c->tab;
            sp.faces = c->faces;
            sp.err = c->err;
            sp.dt = dt;
            sp.ne = c->ne;
            sp.ld = c->ld;
            sp.nfaces = c->nfaces;
            sp.nsteps = reps[r];
            sp.method = method;
            for (int g = 0; g < 4; ++g) sp.fbeg[g] = c->fbeg[g];
            HIP_TRY(vt::launch_small(sp, c->p, c->nq, c->stream));
        } else if (ex[r]) {
            HIP_TRY(hipGraphLaunch(ex[r], c->stream));
        } else {  // single Euler step: Un -> Ua, then copy back
            HIP_TRY(enqueue_euler_step(c, dt, c->Un, c->Ua, 0));
            HIP_TRY(hipMemcpyAsync(c->Un, c->Ua, sizeof(double) * 4 * c->np * c->ld, hipMemcpyDeviceToDevice, c->stream));
        }
        HIP_TRY(hipMemcpyAsync(c->err_host + r, c->err, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
    }
    if (ev1) HIP_TRY(hipEventRecord(ev1, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    int base = 0;
    for (size_t r = 0; r < reps.size(); ++r) {
        // Element::time advances by dt once per step (Timestepping.cpp:91,315).
        if (c->err_host[r] != vt::kNoError) {
            vti::decode_error(c, c->err_host[r], base, method == 1 ? 4 : 1);
            return VT_ERR_NONPHYSICAL;
        }
        for (int i = 0; i < reps[r]; ++i) c->time += dt;
        base += reps[r];
    }
    return VT_OK;
}

extern "C" {

const char* vt_status_string(int s) {
    switch (s) {
    case VT_OK: return "ok";
    case VT_ERR_ARG: return "invalid argument";
    case VT_ERR_UNSUPPORTED_P: return "unsupported p order";
    case VT_ERR_QUAD: return "unsupported quadrature order";
    case VT_ERR_NONPHYSICAL: return "Non-physical state";
    case VT_ERR_HIP: return "HIP error";
    case VT_ERR_INIT_TYPE: return "unsupported initialization type";
    case VT_ERR_METHOD: return "unsupported time stepping method";
    case VT_ERR_MESH: return "unsupported mesh size";
    default: return "unknown status";
    }
}

const char* vt_last_hip_error(void) { return vti::g_hip_error; }

int vt_mesh_generate(int nx, int ny, float dx, float dy, float perturb, double* v, int* nb, int* t) {
    if (!v || !nb || !t) return VT_ERR_ARG;
    return vt::generate_mesh(nx, ny, dx, dy, perturb, v, nb, t) ? VT_ERR_MESH : VT_OK;
}

int vt_mesh_generate_seeded(int nx, int ny, float dx, float dy, double amp, uint64_t seed, double* v, int* nb, int* t) {
    if (!v || !nb || !t) return VT_ERR_ARG;
    return vt::generate_mesh_seeded(nx, ny, dx, dy, amp, seed, v, nb, t) ? VT_ERR_MESH : VT_OK;
}

int vt_nodes_reference(int p, double* nodes) {
    if (!nodes) return VT_ERR_ARG;
    return vt::reference_nodes(p, nodes) ? VT_OK : VT_ERR_UNSUPPORTED_P;
}

int vt_reference_operators(int p, int order, double* M, double* Minv, double* S) {
    if (p < 0 || p > vt::kMaxP) return VT_ERR_UNSUPPORTED_P;
    if (vt::area_rule(order, nullptr) < 0) return VT_ERR_QUAD;
    return vt::reference_operators(p, order, M, Minv, S) ? VT_OK : VT_ERR_ARG;
}

int vt_quadrature(int order, int* nq_line, double* line_xw, int* n_area, double* area_xyw) {
    std::vector<double> a, b;
    const int nl = vt::line_rule(order, &a), na = vt::area_rule(order, &b);
    if (nl < 0 || na < 0) return VT_ERR_QUAD;
    if (nq_line) *nq_line = nl;
    if (n_area) *n_area = na;
    if (line_xw) std::memcpy(line_xw, a.data(), sizeof(double) * a.size());
    if (area_xyw) std::memcpy(area_xyw, b.data(), sizeof(double) * b.size());
    return VT_OK;
}

}  // extern "C"

namespace vti {

int create_ctx(int n, const double* vertices, const int* neighbors, const int* type, int p, const vt::OperatorTable& tab,
               vt_ctx** out) {
    // Size limit first (nothing of the arrays is read for a refused size).
    if (n > kMaxElements) return VT_ERR_ARG;
    for (int e = 0; e < n; ++e) {
        if (type[e] != 0 && type[e] != 1) return VT_ERR_ARG;
        for (int s = 0; s < 3; ++s)
            if (neighbors[3 * e + s] < 0 || neighbors[3 * e + s] >= n) return VT_ERR_ARG;
    }
    vt_ctx* c = new vt_ctx();
    c->ne = n; c->ld = plane_stride(n); c->p = p; c->np = vt::num_nodes(p); c->nq = tab.nq; c->order = tab.order;
    c->vertices.assign(vertices, vertices + 6 * static_cast<size_t>(n));
    auto fail = [&](int rc) { vt_destroy(c); return rc; };
    if (!hip_ok(hipGetDevice(&c->device), "hipGetDevice")) return fail(VT_ERR_HIP);
    if (!hip_ok(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate")) return fail(VT_ERR_HIP);
    const size_t sbytes = sizeof(double) * 4 * c->np * c->ld;
    for (double** b : {&c->Un, &c->Ua, &c->Ub, &c->acc})
        if (!hip_ok(hipMalloc(reinterpret_cast<void**>(b), sbytes), "hipMalloc state")) return fail(VT_ERR_HIP);
    if (!hip_ok(hipMalloc(reinterpret_cast<void**>(&c->vtx), sizeof(double) * 6 * c->ld), "hipMalloc vtx") ||
        !hip_ok(hipMalloc(reinterpret_cast<void**>(&c->nbr), sizeof(int) * 3 * c->ld), "hipMalloc nbr") ||
        !hip_ok(hipMalloc(reinterpret_cast<void**>(&c->tab), sizeof(double) * tab.data.size()), "hipMalloc tab") ||
        !hip_ok(hipMalloc(reinterpret_cast<void**>(&c->err), sizeof(unsigned long long)), "hipMalloc err") ||
        !hip_ok(hipHostMalloc(reinterpret_cast<void**>(&c->err_init), sizeof(unsigned long long), hipHostMallocDefault), "hipHostMalloc"))
        return fail(VT_ERR_HIP);
    *c->err_init = vt::kNoError;
    std::vector<double> hv(static_cast<size_t>(6) * c->ld, 0.0);
    std::vector<int> ns(static_cast<size_t>(3) * n), hn(static_cast<size_t>(3) * c->ld, 0);
    vt::side_neighbours(n, neighbors, type, ns.data());
    // The reference's meshes number the two triangles of a cell 2c, 2c + 1 and
    // put their common diagonal on side s = 1 of both; the stage kernel then
    // evaluates each diagonal once (half of its points on each side).
    c->pairs = n % 2 == 0 ? 1 : 0;
    for (int e = 0; e < n && c->pairs; ++e)
        if (ns[3 * static_cast<size_t>(e) + 1] != (e ^ 1)) c->pairs = 0;
    // Tile order per tile size: plain when at least 90% of the neighbour pairs
    // whose tiles are more than one tile apart (the rows above and below)
    // are a multiple of 8 tiles apart, i.e. on the same XCD in plain order.
    {
        long long far[3] = {0, 0, 0}, same[3] = {0, 0, 0};  // tiles of 64, 128, 256 elements
        for (int e = 0; e < n; ++e)
            for (int s = 0; s < 3; ++s) {
                const int nb = ns[3 * static_cast<size_t>(e) + s];
                for (int bit = 0; bit < 3; ++bit) {
                    const int d = (nb >> (6 + bit)) - (e >> (6 + bit));
                    if (d > 1 || d < -1) {
                        ++far[bit];
                        if ((d & 7) == 0) ++same[bit];
                    }
                }
            }
        for (int bit = 0; bit < 3; ++bit)
            if (far[bit] > 0 && same[bit] * 10 >= far[bit] * 9) c->plain_order |= 1 << bit;
    }
    // Small-mesh path: the mesh fits one workgroup's LDS and every side's
    // neighbour faces back across the same side, so each shared side can be
    // listed once (as e | nb << 10 | s << 20 with e < nb).
    std::vector<int> faces;
    if (n <= vt::kSmallMaxElements && c->nq >= 1) {
        bool ok = true;
        for (int s = 0; s < 3 && ok; ++s) {  // listed by side index
            c->fbeg[s] = static_cast<int>(faces.size());
            for (int e = 0; e < n && ok; ++e) {
                const int nb = ns[3 * static_cast<size_t>(e) + s];
                if (nb == e || ns[3 * static_cast<size_t>(nb) + s] != e) ok = false;
                else if (e < nb) faces.push_back(e | (nb << 10) | (s << 20));
            }
        }
        c->fbeg[3] = static_cast<int>(faces.size());
        int lim = 0, optin = 0;
        if (!hip_ok(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device), "attribute") ||
            !hip_ok(hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, c->device), "attribute"))
            return fail(VT_ERR_HIP);
        c->small_ok = ok && static_cast<int>(faces.size()) * 2 == 3 * n &&
                      vt::small_lds_bytes(p, c->nq, n) <= static_cast<size_t>(std::max(lim, optin));
        if (!c->small_ok) faces.clear();
    }
    if (c->small_ok) {
        if (!hip_ok(hipMalloc(reinterpret_cast<void**>(&c->faces), sizeof(int) * faces.size()), "hipMalloc faces") ||
            !hip_ok(hipMemcpy(c->faces, faces.data(), sizeof(int) * faces.size(), hipMemcpyHostToDevice), "upload faces"))
            return fail(VT_ERR_HIP);
        c->nfaces = static_cast<int>(faces.size());
        c->small_on = true;
    }
    for (int e = 0; e < n; ++e) {
        for (int k = 0; k < 6; ++k) hv[static_cast<size_t>(k) * c->ld + e] = vertices[6 * static_cast<size_t>(e) + k];
        for (int s = 0; s < 3; ++s) hn[static_cast<size_t>(s) * c->ld + e] = ns[3 * static_cast<size_t>(e) + s];
    }
    if (!hip_ok(hipMemcpy(c->vtx, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice), "upload vtx") ||
        !hip_ok(hipMemcpy(c->nbr, hn.data(), sizeof(int) * hn.size(), hipMemcpyHostToDevice), "upload nbr") ||
        !hip_ok(hipMemcpy(c->tab, tab.data.data(), sizeof(double) * tab.data.size(), hipMemcpyHostToDevice), "upload tab") ||
        // Stream-ordered clears: the context stream does not synchronise with the
        // null stream, so a plain hipMemset could land after later uploads.
        !hip_ok(hipMemsetAsync(c->Un, 0, sbytes, c->stream), "memset") ||
        !hip_ok(hipMemsetAsync(c->Ua, 0, sbytes, c->stream), "memset") ||
        !hip_ok(hipMemsetAsync(c->Ub, 0, sbytes, c->stream), "memset") ||
        !hip_ok(hipMemsetAsync(c->acc, 0, sbytes, c->stream), "memset") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return fail(VT_ERR_HIP);
    *out = c;
    return VT_OK;
}

int vortex_device(vt_ctx* c, int nr, const int* e0, const int* n, const double* par, int type, double* out,
                  bool density_only) {
    if (type != 0 && type != 1) return VT_ERR_INIT_TYPE;
    vt::VortexParams v{};
    if (!vt::reference_nodes(c->p, v.nodes)) return VT_ERR_UNSUPPORTED_P;
    v.ld = c->ld;
    v.type = type;
    v.density_only = density_only ? 1 : 0;
    if (type == 1) {
        if (c->order < 0) return VT_ERR_QUAD;  // no area rule known
        std::vector<double> q;
        const int nqa = vt::area_rule(c->order, &q);
        if (nqa < 1) return VT_ERR_QUAD;
        v.nqa = nqa;
        if (!c->vrule) {
            // rule [nqa][3], phi [Np][nqa], Minv_ref [Np][Np]
            const int np = c->np;
            std::vector<double> t(static_cast<size_t>(3 + np) * nqa + static_cast<size_t>(np) * np);
            std::copy(q.begin(), q.end(), t.begin());
            for (int k = 0; k < nqa; ++k) {
                double b[vt::kMaxNp];
                vt::basis(c->p, q[3 * k], q[3 * k + 1], b);
                for (int j = 0; j < np; ++j) t[3 * static_cast<size_t>(nqa) + static_cast<size_t>(j) * nqa + k] = b[j];
            }
            if (!vt::reference_operators(c->p, c->order, nullptr, t.data() + static_cast<size_t>(3 + np) * nqa, nullptr))
                return VT_ERR_QUAD;
            HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->vrule), sizeof(double) * t.size()));
            HIP_TRY(hipMemcpyAsync(c->vrule, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, c->stream));
            HIP_TRY(hipStreamSynchronize(c->stream));
        }
        v.rule = c->vrule;
    }
    for (int r = 0; r < nr; ++r) {
        v.out = out + e0[r];
        v.vtx = c->vtx + e0[r];
        v.ne = n[r];
        v.t = par[4 * r];
        vt::vortex_constants(par[4 * r + 1], v.k);
        v.x0 = par[4 * r + 2];
        v.y0 = par[4 * r + 3];
        HIP_TRY(vt::launch_vortex_state(v, c->p, c->stream));
    }
    return VT_OK;
}

}  // namespace vti

extern "C" {

int vt_create(int n, const double* vertices, const int* neighbors, const int* type, int p, int order, vt_ctx** out) {
    if (!out) return VT_ERR_ARG;
    *out = nullptr;
    if (n < 1 || !vertices || !neighbors || !type) return VT_ERR_ARG;
    if (p < 0 || p > vt::kMaxP) return VT_ERR_UNSUPPORTED_P;
    vt::OperatorTable tab;
    if (vt::line_rule(order, nullptr) < 0 || !vt::build_operator_table(p, order, &tab)) return VT_ERR_QUAD;
    return vti::create_ctx(n, vertices, neighbors, type, p, tab, out);
}

int vt_create_with_operators(int n, const double* vertices, const int* neighbors, const int* type, int p, int nq,
                             const double* line_xw, const double* Minv_ref, const double* S_ref, vt_ctx** out) {
    if (!out) return VT_ERR_ARG;
    *out = nullptr;
    if (n < 1 || !vertices || !neighbors || !type || !line_xw || !Minv_ref || !S_ref) return VT_ERR_ARG;
    if (p < 0 || p > vt::kMaxP) return VT_ERR_UNSUPPORTED_P;
    if (nq < 1 || nq > vt::kMaxNq) return VT_ERR_QUAD;
    vt::OperatorTable tab;
    if (!vt::build_operator_table(p, nq, line_xw, Minv_ref, S_ref, &tab)) return VT_ERR_ARG;
    tab.order = -1;
    return vti::create_ctx(n, vertices, neighbors, type, p, tab, out);
}

void vt_destroy(vt_ctx* c) {
    if (!c) return;
    vti::free_adjoint(c);
    for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
    for (void* b : {static_cast<void*>(c->Un), static_cast<void*>(c->Ua), static_cast<void*>(c->Ub),
                    static_cast<void*>(c->acc), static_cast<void*>(c->vtx), static_cast<void*>(c->nbr),
                    static_cast<void*>(c->tab), static_cast<void*>(c->err), static_cast<void*>(c->l2ref),
                    static_cast<void*>(c->l2tab), static_cast<void*>(c->l2part), static_cast<void*>(c->vrule),
                    static_cast<void*>(c->faces)})
        if (b) hipFree(b);
    if (c->err_init) hipHostFree(c->err_init);
    if (c->err_host) hipHostFree(c->err_host);
    if (c->host_stage) hipHostFree(c->host_stage);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int vt_num_elements(const vt_ctx* c) { return c ? c->ne : 0; }
int vt_num_nodes(const vt_ctx* c) { return c ? c->np : 0; }

int vt_init_vortex(vt_ctx* c, int init_type, double epsilon, double x0, double y0) {
    if (!c) return VT_ERR_ARG;
    if (init_type == 1 && c->order < 0) return VT_ERR_QUAD;  // no area rule known
    std::vector<double> U(static_cast<size_t>(c->ne) * c->np * 4);
    if (!vt::initial_state(c->ne, c->vertices.data(), c->p, c->order, init_type, epsilon, x0, y0, c->time, U.data()))
        return VT_ERR_INIT_TYPE;
    return vt_set_state(c, U.data());
}

int vt_init_vortex_device(vt_ctx* c, int init_type, double epsilon, double x0, double y0) {
    if (!c) return VT_ERR_ARG;
    const int e0 = 0;
    const double par[4] = {c->time, epsilon, x0, y0};
    const int rc = vti::vortex_device(c, 1, &e0, &c->ne, par, init_type, c->Un, false);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VT_OK;
}

int vt_set_small_path(vt_ctx* c, int enable) {
    if (!c) return VT_ERR_ARG;
    if (enable && !c->small_ok) return VT_ERR_MESH;
    c->small_on = enable != 0;
    return VT_OK;
}

int vt_get_small_path(const vt_ctx* c) { return c && c->small_on ? 1 : 0; }

int vt_set_target_vortex(vt_ctx* c, double t, double epsilon, double x0, double y0) {
    if (!c) return VT_ERR_ARG;
    c->tgt_par[0] = t;
    c->tgt_par[1] = epsilon;
    c->tgt_par[2] = x0;
    c->tgt_par[3] = y0;
    c->tgt_vortex = true;
    return VT_OK;
}

int vt_exact_vortex(const vt_ctx* c, double t, double epsilon, double x0, double y0, double* U) {
    if (!c || !U) return VT_ERR_ARG;
    vt::initial_state(c->ne, c->vertices.data(), c->p, c->order, 0, epsilon, x0, y0, t, U);
    return VT_OK;
}

int vt_set_state(vt_ctx* c, const double* U) {
    if (!c || !U) return VT_ERR_ARG;
    return vti::upload_state(c, U, c->Un);
}

int vt_get_state(vt_ctx* c, double* U) {
    if (!c || !U) return VT_ERR_ARG;
    return read_planes(c, c->Un, U);
}

int vt_l2_state_error(vt_ctx* c, const double* U_ref, double* err) {
    if (!c || !U_ref || !err) return VT_ERR_ARG;
    const int nblk = (c->ne + vt::kL2Block - 1) / vt::kL2Block;
    if (!c->l2ref) {
        std::vector<double> q;
        if (vt::area_rule(10, &q) != vt::kL2Points) return VT_ERR_QUAD;
        std::vector<double> t(static_cast<size_t>(vt::kL2Points) * (c->np + 1));
        for (int i = 0; i < vt::kL2Points; ++i) {
            vt::basis(c->p, q[3 * i], q[3 * i + 1], &t[static_cast<size_t>(i) * c->np]);
            t[static_cast<size_t>(vt::kL2Points) * c->np + i] = q[3 * i + 2];
        }
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->l2ref), sizeof(double) * 4 * c->np * c->ld));
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->l2tab), sizeof(double) * t.size()));
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->l2part), sizeof(double) * 4 * nblk));
        HIP_TRY(hipMemcpyAsync(c->l2tab, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    int rc = vti::upload_state(c, U_ref, c->l2ref);
    if (rc) return rc;
    vt::L2Params l{};
    l.U = c->Un;
    l.Uref = c->l2ref;
    l.vtx = c->vtx;
    l.tab = c->l2tab;
    l.part = c->l2part;
    l.ne = c->ne;
    l.ld = c->ld;
    HIP_TRY(vt::launch_l2(l, c->p, c->stream));
    double* h = vti::host_stage(c, 4 * static_cast<size_t>(nblk));
    if (!h) return VT_ERR_HIP;
    HIP_TRY(hipMemcpyAsync(h, c->l2part, sizeof(double) * 4 * nblk, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int k = 0; k < 4; ++k) {
        double s = 0.0;
        for (int b = 0; b < nblk; ++b) s += h[4 * b + k];
        err[k] = std::sqrt((1.0 / 25.0) * s);
    }
    return VT_OK;
}

int vt_get_flux(vt_ctx* c, double* F) {
    if (!c || !F) return VT_ERR_ARG;
    std::vector<double> U(static_cast<size_t>(c->ne) * c->np * 4);
    int rc = read_planes(c, c->Un, U.data());
    if (rc) return rc;
    vt::nodal_flux_n(static_cast<size_t>(c->ne) * c->np, U.data(), F);
    return VT_OK;
}

double vt_get_time(const vt_ctx* c) { return c ? c->time : 0.0; }
int vt_set_time(vt_ctx* c, double t) {
    if (!c) return VT_ERR_ARG;
    c->time = t;
    return VT_OK;
}

int vt_rhs(vt_ctx* c, double* dUdt) {
    if (!c || !dUdt) return VT_ERR_ARG;
    vt::StageParams prm = base_params(c);
    prm.Uin = c->Un;
    prm.Uout = c->Ub;
    prm.mode = vt::kModeRhs;
    prm.stage_id = 0;
    HIP_TRY(hipMemcpyAsync(c->err, c->err_init, sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(vt::launch_stage(prm, c->p, c->nq, c->cfg, c->stream));
    int rc = read_planes(c, c->Ub, dUdt);
    if (rc) return rc;
    unsigned long long key = 0;
    HIP_TRY(hipMemcpy(&key, c->err, sizeof(key), hipMemcpyDeviceToHost));
    if (key != vt::kNoError) {
        vti::decode_error(c, key, -1, 1);
        return VT_ERR_NONPHYSICAL;
    }
    return VT_OK;
}

int vt_step(vt_ctx* c, int method, double dt, int nsteps) {
    if (!c) return VT_ERR_ARG;
    return vti::run_steps(c, method, dt, nsteps, nullptr, nullptr);
}

int vt_prepare_steps(vt_ctx* c, int method, double dt, int nsteps) {
    if (!c) return VT_ERR_ARG;
    if (method != 0 && method != 1) return VT_ERR_METHOD;
    if (nsteps < 0) return VT_ERR_ARG;
    for (int k : vti::replay_plan(method, nsteps)) {
        if (method == 0 && k == 1) continue;  // a single Euler step is launched directly
        hipGraphExec_t g;
        const int rc = get_graph(c, method, dt, k, &g);
        if (rc) return rc;
    }
    return VT_OK;
}

int vt_step_timed(vt_ctx* c, int method, double dt, int nsteps, double* ms) {
    if (!c || !ms) return VT_ERR_ARG;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = vti::run_steps(c, method, dt, nsteps, e0, e1);
    float f = 0.f;
    if (rc == VT_OK && !hip_ok(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime")) rc = VT_ERR_HIP;
    *ms = f;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int vt_stage_kernel_times(vt_ctx* c, double dt, int nsteps, double* ms) {
    if (!c || !ms || nsteps < 1) return VT_ERR_ARG;
    std::vector<hipEvent_t> ev(8 * static_cast<size_t>(nsteps));
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    vt::StageParams prm = base_params(c);
    prm.dt = dt;
    const double* in[4] = {c->Un, c->Ua, c->Ub, c->Ua};
    double* outp[4] = {c->Ua, c->Ub, c->Ua, c->Un};
    HIP_TRY(hipMemcpyAsync(c->err, c->err_init, sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
    for (int i = 0; i < nsteps; ++i)
        for (int s = 0; s < 4; ++s) {
            prm.Uin = in[s];
            prm.Uout = outp[s];
            prm.Un = (s == 1 || s == 2) ? c->Un : nullptr;
            prm.U2 = s == 2 ? c->Ua : nullptr;
            prm.acc = c->acc;
            prm.mode = vt::kModeRk1 + s;
            prm.stage_id = 4 * i + s;
            HIP_TRY(hipEventRecord(ev[8 * i + 2 * s], c->stream));
            HIP_TRY(vt::launch_stage(prm, c->p, c->nq, c->cfg, c->stream));
            HIP_TRY(hipEventRecord(ev[8 * i + 2 * s + 1], c->stream));
        }
    HIP_TRY(hipStreamSynchronize(c->stream));
    double acc[4] = {0, 0, 0, 0};
    for (int i = 0; i < nsteps; ++i)
        for (int s = 0; s < 4; ++s) {
            float f = 0.f;
            HIP_TRY(hipEventElapsedTime(&f, ev[8 * i + 2 * s], ev[8 * i + 2 * s + 1]));
            acc[s] += f;
        }
    for (auto& e : ev) hipEventDestroy(e);
    for (int s = 0; s < 4; ++s) ms[s] = acc[s] / nsteps;
    for (int i = 0; i < nsteps; ++i) c->time += dt;
    unsigned long long key = 0;
    HIP_TRY(hipMemcpy(&key, c->err, sizeof(key), hipMemcpyDeviceToHost));
    if (key != vt::kNoError) {
        vti::decode_error(c, key, 0, 4);
        return VT_ERR_NONPHYSICAL;
    }
    return VT_OK;
}

int vt_set_launch(vt_ctx* c, int block, int persistent, int slots_per_xcd) {
    if (!c || (block != 64 && block != 128 && block != 256) || slots_per_xcd < 0 || (persistent && slots_per_xcd < 1))
        return VT_ERR_ARG;
    c->cfg.block = block;
    c->cfg.persistent = persistent ? 1 : 0;
    c->cfg.slots_per_xcd = slots_per_xcd;
    for (auto& kv : c->graphs) hipGraphExecDestroy(kv.second);
    c->graphs.clear();
    return VT_OK;
}

int vt_get_launch(const vt_ctx* c, int* cfg) {
    if (!c || !cfg) return VT_ERR_ARG;
    cfg[0] = c->cfg.block;
    cfg[1] = c->cfg.persistent;
    cfg[2] = c->cfg.slots_per_xcd;
    return VT_OK;
}

int vt_last_error(const vt_ctx* c, int* info) {
    if (!c || !info) return VT_ERR_ARG;
    std::memcpy(info, c->last_err, sizeof(c->last_err));
    return VT_OK;
}

int vt_numerical_flux(const double* rec, double* flux, int* codes, int n) {
    if (n < 0 || (n > 0 && (!rec || !flux || !codes))) return VT_ERR_ARG;
    if (n == 0) return VT_OK;
    double *din = nullptr, *dout = nullptr;
    int* dc = nullptr;
    int rc = VT_OK;
    if (!hip_ok(hipMalloc(reinterpret_cast<void**>(&din), sizeof(double) * 10 * n), "hipMalloc") ||
        !hip_ok(hipMalloc(reinterpret_cast<void**>(&dout), sizeof(double) * 4 * n), "hipMalloc") ||
        !hip_ok(hipMalloc(reinterpret_cast<void**>(&dc), sizeof(int) * n), "hipMalloc") ||
        !hip_ok(hipMemcpy(din, rec, sizeof(double) * 10 * n, hipMemcpyHostToDevice), "upload") ||
        !hip_ok(vt::launch_flux(din, dout, dc, n, nullptr), "flux kernel") ||
        !hip_ok(hipMemcpy(flux, dout, sizeof(double) * 4 * n, hipMemcpyDeviceToHost), "download") ||
        !hip_ok(hipMemcpy(codes, dc, sizeof(int) * n, hipMemcpyDeviceToHost), "download"))
        rc = VT_ERR_HIP;
    if (din) hipFree(din);
    if (dout) hipFree(dout);
    if (dc) hipFree(dc);
    return rc;
}

}  // extern "C"
