// vt_batch.hip -- the ensemble: many independent vortex problems of uneven
// mesh size and polynomial order advanced together.
//
// Instances with the same (p, integration order) are packed into one group:
// a disjoint union of their meshes in one set of SoA planes, with each
// instance's neighbour indices shifted by its element offset, so the
// connectivity is block-diagonal and the single-mesh stage code applies
// unchanged.  Every RK stage is one launch of the batched stage kernel per
// quadrature size (normally one), covering the tiles of all groups; the
// steps are captured in HIP graphs as for a single mesh.  Non-physical
// states are reported through one error key whose element index is global
// over the groups (StageParams::err_base).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <tuple>
#include <vector>

#include "vt_ctx.hpp"
#include "vt_device.hpp"

struct vt_batch {
    int n = 0;
    std::vector<int> ne, p, order, grp, eoff;  // per instance
    std::vector<vt_ctx*> g;                    // groups
    std::vector<long long> gbase;              // first global element of each group
    std::vector<std::vector<int>> members;     // instances of each group, by element offset
    std::vector<std::vector<int>> launches;    // groups of each launch (same quadrature size)
    hipStream_t stream = nullptr;
    unsigned long long* err = nullptr;
    unsigned long long* err_init = nullptr;
    unsigned long long* err_host = nullptr;
    size_t err_host_cap = 0;
    vt::LaunchConfig cfg;
    std::map<std::tuple<int, unsigned long long, int>, hipGraphExec_t> graphs;
    double time = 0.0;
    int last_err[7] = {-1, -1, -1, -1, 0, 0, 0};
    std::vector<double> tgt_par;               // [n][4] target vortex {t, eps, x0, y0}, or empty
    // Launch mode 2 (cfg.persistent == 2): each group's stage is its own
    // single-mesh launch on the group's stream, the groups forked from and
    // joined back to the batch stream, so they run concurrently, each kernel
    // at its own occupancy.
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
};

namespace {

using vti::hip_ok;

void drop_graphs(vt_batch* b) {
    for (auto& kv : b->graphs) hipGraphExecDestroy(kv.second);
    b->graphs.clear();
}

// One batched launch per quadrature size for one stage (or, in launch mode 2,
// one launch per group on concurrent streams).  bufs(c) gives the stage's
// {Uin, Uout, Un, U2} of group context c.
template <class Bufs>
hipError_t enqueue_stage(vt_batch* b, double dt, int mode, int stage_id, Bufs bufs) {
    if (b->cfg.persistent == 2) {
        hipError_t e = hipEventRecord(b->fork, b->stream);
        vt::LaunchConfig one;
        one.block = b->cfg.block;
        for (size_t gi = 0; gi < b->g.size() && e == hipSuccess; ++gi) {
            vt_ctx* c = b->g[gi];
            const double* io[4];
            bufs(c, io);
            vt::StageParams prm{};
            prm.Uin = io[0];
            prm.Uout = const_cast<double*>(io[1]);
            prm.Un = io[2];
            prm.U2 = io[3];
            prm.acc = c->acc;
            prm.vtx = c->vtx;
            prm.nbr = c->nbr;
            prm.tab = c->tab;
            prm.err = b->err;
            prm.dt = dt;
            prm.ne = c->ne;
            prm.ld = c->ld;
            prm.mode = mode;
            prm.stage_id = stage_id;
            prm.err_base = static_cast<int>(b->gbase[gi]);
            prm.pairs = c->pairs;
            prm.plain_order = vti::plain_order_for(c, one.block);
            prm.blocks_per_xcd = vt::tiles_per_xcd(c->ne, one.block);
            e = hipStreamWaitEvent(c->stream, b->fork, 0);
            if (e == hipSuccess) e = vt::launch_stage(prm, c->p, c->nq, one, c->stream);
            if (e == hipSuccess) e = hipEventRecord(b->join[gi], c->stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join[gi], 0);
        }
        return e;
    }
    for (const auto& L : b->launches) {
        vt::BatchParams bp{};
        bp.ngroups = static_cast<int>(L.size());
        bp.kbegin[0] = 0;
        for (int i = 0; i < bp.ngroups; ++i) {
            vt_ctx* c = b->g[L[i]];
            vt::StageParams& prm = bp.g[i];
            const double* io[4];
            bufs(c, io);
            prm.Uin = io[0];
            prm.Uout = const_cast<double*>(io[1]);
            prm.Un = io[2];
            prm.U2 = io[3];
            prm.acc = c->acc;
            prm.vtx = c->vtx;
            prm.nbr = c->nbr;
            prm.tab = c->tab;
            prm.err = b->err;
            prm.dt = dt;
            prm.ne = c->ne;
            prm.ld = c->ld;
            prm.mode = mode;
            prm.stage_id = stage_id;
            prm.err_base = static_cast<int>(b->gbase[L[i]]);
            prm.pairs = c->pairs;
            bp.p[i] = c->p;
            bp.tiles[i] = (c->ne + b->cfg.block - 1) / b->cfg.block;
            bp.tpx[i] = (bp.tiles[i] + 7) / 8;
            bp.kbegin[i + 1] = bp.kbegin[i] + bp.tpx[i];
        }
        hipError_t e = vt::launch_batched(bp, b->g[L[0]]->nq, b->cfg, b->stream);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t enqueue_rk4_step(vt_batch* b, double dt, int stage_base) {
    // Un -> Ua -> Ub -> Ua -> Un, as for a single mesh.
    const int modes[4] = {vt::kModeRk1, vt::kModeRk2, vt::kModeRk3, vt::kModeRk4};
    for (int s = 0; s < 4; ++s) {
        hipError_t e = enqueue_stage(b, dt, modes[s], stage_base + s, [s](vt_ctx* c, const double** io) {
            const double* in[4] = {c->Un, c->Ua, c->Ub, c->Ua};
            double* out[4] = {c->Ua, c->Ub, c->Ua, c->Un};
            io[0] = in[s];
            io[1] = out[s];
            io[2] = (s == 1 || s == 2) ? c->Un : nullptr;
            io[3] = s == 2 ? c->Ua : nullptr;
        });
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t enqueue_euler_step(vt_batch* b, double dt, bool from_a, int stage_id) {
    return enqueue_stage(b, dt, vt::kModeEuler, stage_id, [from_a](vt_ctx* c, const double** io) {
        io[0] = from_a ? c->Ua : c->Un;
        io[1] = from_a ? c->Un : c->Ua;
        io[2] = nullptr;
        io[3] = nullptr;
    });
}

int get_graph(vt_batch* b, int method, double dt, int k, hipGraphExec_t* out) {
    unsigned long long dbits;
    std::memcpy(&dbits, &dt, sizeof(dbits));
    auto key = std::make_tuple(method, dbits, k);
    auto it = b->graphs.find(key);
    if (it != b->graphs.end()) { *out = it->second; return VT_OK; }
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int i = 0; i < k && e == hipSuccess; ++i) {
        if (method == 1) e = enqueue_rk4_step(b, dt, 4 * i);
        else e = enqueue_euler_step(b, dt, (i % 2) != 0, i);
    }
    hipGraphExec_t ex;
    const int rc = vti::finish_capture(b->stream, e, &ex);
    if (rc) return rc;
    b->graphs[key] = ex;
    *out = ex;
    return VT_OK;
}

void decode_error(vt_batch* b, unsigned long long key, int step_base, int per_step) {
    const unsigned long long stage = key >> 40;
    unsigned long long t = key & ((1ull << 40) - 1);
    b->last_err[4] = static_cast<int>(t % 4); t /= 4;
    b->last_err[3] = static_cast<int>(t % 4); t /= 4;
    b->last_err[2] = static_cast<int>(t % 16); t /= 16;
    const long long ge = static_cast<long long>(t);
    int gi = 0;
    for (size_t i = 1; i < b->g.size(); ++i)
        if (ge >= b->gbase[i]) gi = static_cast<int>(i);
    const long long le = ge - b->gbase[gi];
    int inst = b->members[gi][0];
    for (int m : b->members[gi])
        if (le >= b->eoff[m]) inst = m;
    b->last_err[0] = inst;
    b->last_err[1] = static_cast<int>(le - b->eoff[inst]);
    b->last_err[5] = step_base + static_cast<int>(stage / per_step) + 1;
    b->last_err[6] = static_cast<int>(stage % per_step) + 1;
}

int run_steps(vt_batch* b, int method, double dt, int nsteps, hipEvent_t ev0, hipEvent_t ev1) {
    if (method != 0 && method != 1) return VT_ERR_METHOD;
    if (nsteps < 0) return VT_ERR_ARG;
    const std::vector<int> reps = vti::replay_plan(method, nsteps);
    int rc = vti::ensure_err_slots(&b->err_host, &b->err_host_cap, reps.size() + 1);
    if (rc) return rc;
    std::vector<hipGraphExec_t> ex(reps.size(), nullptr);
    for (size_t r = 0; r < reps.size(); ++r)
        if (!(method == 0 && reps[r] == 1)) {
            rc = get_graph(b, method, dt, reps[r], &ex[r]);
            if (rc) return rc;
        }
    if (ev0) HIP_TRY(hipEventRecord(ev0, b->stream));
    for (size_t r = 0; r < reps.size(); ++r) {
        HIP_TRY(hipMemcpyAsync(b->err, b->err_init, sizeof(unsigned long long), hipMemcpyHostToDevice, b->stream));
        if (ex[r]) {
            HIP_TRY(hipGraphLaunch(ex[r], b->stream));
        } else {  // single Euler step: Un -> Ua, then copy back
            HIP_TRY(enqueue_euler_step(b, dt, false, 0));
            for (vt_ctx* c : b->g)
                HIP_TRY(hipMemcpyAsync(c->Un, c->Ua, sizeof(double) * 4 * c->np * c->ld, hipMemcpyDeviceToDevice,
                                       b->stream));
        }
        HIP_TRY(hipMemcpyAsync(b->err_host + r, b->err, sizeof(unsigned long long), hipMemcpyDeviceToHost, b->stream));
    }
    if (ev1) HIP_TRY(hipEventRecord(ev1, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    int base = 0;
    for (size_t r = 0; r < reps.size(); ++r) {
        if (b->err_host[r] != vt::kNoError) {
            decode_error(b, b->err_host[r], base, method == 1 ? 4 : 1);
            return VT_ERR_NONPHYSICAL;
        }
        for (int i = 0; i < reps[r]; ++i) b->time += dt;
        base += reps[r];
    }
    return VT_OK;
}

int check_inst(const vt_batch* b, int i) { return (b && i >= 0 && i < b->n) ? VT_OK : VT_ERR_ARG; }

// First node of each instance in the concatenated [sum n_i][Np_i] arrays.
std::vector<size_t> node_offsets(const vt_batch* b) {
    std::vector<size_t> off(b->n + 1, 0);
    for (int i = 0; i < b->n; ++i) off[i + 1] = off[i] + static_cast<size_t>(b->ne[i]) * vt::num_nodes(b->p[i]);
    return off;
}

// Objective of every instance of group gi at its current state: J of each
// member, dJ/dU of the group in its lam planes.
// rho_target null: the members' target vortices set by vt_batch_set_target_vortex,
// computed on the device.
int group_objective(vt_batch* b, int gi, const double* rho_target, const std::vector<size_t>& off, double* J) {
    vt_ctx* c = b->g[gi];
    const std::vector<int>& mem = b->members[gi];
    std::vector<int> e0(mem.size()), n(mem.size());
    std::vector<const double*> rho(mem.size());
    std::vector<double> par(4 * mem.size());
    if (!rho_target && b->tgt_par.empty()) return VT_ERR_ARG;
    for (size_t k = 0; k < mem.size(); ++k) {
        const int i = mem[k];
        e0[k] = b->eoff[i];
        n[k] = b->ne[i];
        if (rho_target) rho[k] = rho_target + off[i];
        else std::copy(b->tgt_par.begin() + 4 * i, b->tgt_par.begin() + 4 * i + 4, par.begin() + 4 * k);
    }
    std::vector<double> Jg(mem.size());
    const int rc = vti::objective(c, static_cast<int>(mem.size()), e0.data(), n.data(), rho_target ? rho.data() : nullptr,
                                  par.data(), Jg.data());
    if (rc) return rc;
    for (size_t k = 0; k < mem.size(); ++k) J[mem[k]] = Jg[k];
    return VT_OK;
}

// dJ/dU of group gi (its lam planes) into the members' slices of dJdU,
// converted from the planes straight into each slice.
int scatter_group_gradient(vt_batch* b, int gi, const std::vector<size_t>& off, double* dJdU) {
    vt_ctx* c = b->g[gi];
    const size_t n = static_cast<size_t>(4) * c->np * c->ld;
    double* h = vti::host_stage(c, n);
    if (!h) return VT_ERR_HIP;
    HIP_TRY(hipMemcpyAsync(h, c->adj, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int i : b->members[gi]) vt::planes_to_aos(b->ne[i], c->np, c->ld, b->eoff[i], h, dJdU + 4 * off[i]);
    return VT_OK;
}

// dJ_i/d(eps_i, x0_i, y0_i) of every member of group gi from its lam planes,
// on the device: one launch per member's element range, the block partial
// sums added on the host in block order.  t: the time of the initial state.
int group_vortex_gradient(vt_batch* b, int gi, double t, const double* eps, const double* x0, const double* y0,
                          double* grad) {
    vt_ctx* c = b->g[gi];
    const std::vector<int>& mem = b->members[gi];
    std::vector<int> boff(mem.size() + 1, 0);
    for (size_t k = 0; k < mem.size(); ++k) boff[k + 1] = boff[k] + (b->ne[mem[k]] + vt::kVGradBlock - 1) / vt::kVGradBlock;
    const int need = 3 * boff[mem.size()];
    if (c->part_cap < need) {
        if (c->part) (void)hipFree(c->part);
        c->part = nullptr;
        c->part_cap = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->part), sizeof(double) * need));
        c->part_cap = need;
    }
    vt::VGradParams prm{};
    if (!vt::reference_nodes(c->p, prm.nodes)) return VT_ERR_UNSUPPORTED_P;
    prm.ld = c->ld;
    prm.t = t;
    for (size_t k = 0; k < mem.size(); ++k) {
        const int i = mem[k];
        prm.lam = c->adj + b->eoff[i];
        prm.vtx = c->vtx + b->eoff[i];
        prm.part = c->part + 3 * boff[k];
        prm.ne = b->ne[i];
        prm.eps = eps[i];
        prm.x0 = x0[i];
        prm.y0 = y0[i];
        HIP_TRY(vt::launch_vortex_grad(prm, c->p, c->stream));
    }
    double* h = vti::host_stage(c, static_cast<size_t>(need));
    if (!h) return VT_ERR_HIP;
    HIP_TRY(hipMemcpyAsync(h, c->part, sizeof(double) * need, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (size_t k = 0; k < mem.size(); ++k) {
        double g[3] = {0.0, 0.0, 0.0};
        for (int bl = boff[k]; bl < boff[k + 1]; ++bl)
            for (int a = 0; a < 3; ++a) g[a] += h[3 * bl + a];
        for (int a = 0; a < 3; ++a) grad[3 * mem[k] + a] = g[a];
    }
    return VT_OK;
}

// The ensemble adjoint, group by group: each group is one mesh (the disjoint
// union of its instances), recorded, evaluated and swept by the single-mesh
// adjoint; gather(gi) then reads the group's dJ/dU0 from its lam planes.  A
// group's recording (checkpoints, segment states, adjoint planes: the bulk of
// device memory here) is released before another group records, so one is
// held at a time; the last one is kept and re-used by the next call (freeing
// and re-allocating some hundred GB costs seconds), until
// vt_batch_adjoint_release or vt_batch_destroy.
template <class Gather>
int batch_adjoint(vt_batch* b, double dt, int nsteps, int every, const double* rho_target, double* J, Gather gather) {
    const std::vector<size_t> off = node_offsets(b);
    const int ng = static_cast<int>(b->g.size());
    for (int gi = 0; gi < ng; ++gi) {
        vt_ctx* c = b->g[gi];
        for (int gj = 0; gj < ng; ++gj)
            if (gj != gi) vti::free_adjoint(b->g[gj]);
        int rc = vt_adjoint_forward(c, dt, nsteps, every);
        if (rc == VT_ERR_NONPHYSICAL) {
            const int le = c->last_err[0];
            int inst = b->members[gi][0];
            for (int m : b->members[gi])
                if (le >= b->eoff[m]) inst = m;
            b->last_err[0] = inst;
            b->last_err[1] = le - b->eoff[inst];
            for (int k = 1; k < 6; ++k) b->last_err[k + 1] = c->last_err[k];
        }
        if (!rc) rc = group_objective(b, gi, rho_target, off, J);
        if (!rc) rc = vti::backward_sweep(c);
        if (!rc) rc = gather(gi, off);
        if (rc) {
            vti::free_adjoint(c);
            return rc;
        }
    }
    for (int s = 0; s < nsteps; ++s) b->time += dt;
    return VT_OK;
}

}  // namespace

extern "C" {

int vt_batch_create(int n_inst, const int* n_elem, const int* p, const int* order, const double* vertices,
                    const int* neighbors, const int* types, vt_batch** out) {
    if (!out) return VT_ERR_ARG;
    *out = nullptr;
    if (n_inst < 1 || !n_elem || !p || !order || !vertices || !neighbors || !types) return VT_ERR_ARG;
    std::vector<long long> first(n_inst + 1, 0);
    for (int i = 0; i < n_inst; ++i) {
        if (n_elem[i] < 1) return VT_ERR_ARG;
        if (p[i] < 0 || p[i] > vt::kMaxP) return VT_ERR_UNSUPPORTED_P;
        if (vt::line_rule(order[i], nullptr) < 0) return VT_ERR_QUAD;
        first[i + 1] = first[i] + n_elem[i];
    }
    // Whole-ensemble size limit (int element indices, 40-bit error key) before
    // any array is read; each group's own limit (kMaxElements) is checked when
    // it is created.
    if (first[n_inst] >= (1ll << 31) || first[n_inst] * 256 >= (1ll << 40)) return VT_ERR_ARG;
    for (int i = 0; i < n_inst; ++i) {
        for (long long e = first[i]; e < first[i + 1]; ++e)
            for (int s = 0; s < 3; ++s)
                if (neighbors[3 * e + s] < 0 || neighbors[3 * e + s] >= n_elem[i]) return VT_ERR_ARG;
    }
    vt_batch* b = new vt_batch();
    b->n = n_inst;
    b->ne.assign(n_elem, n_elem + n_inst);
    b->p.assign(p, p + n_inst);
    b->order.assign(order, order + n_inst);
    b->grp.assign(n_inst, -1);
    b->eoff.assign(n_inst, 0);
    auto fail = [&](int rc) { vt_batch_destroy(b); return rc; };
    // Groups by (quadrature size, p, order), so that the groups of one launch are adjacent.
    std::map<std::tuple<int, int, int>, std::vector<int>> by_key;
    for (int i = 0; i < n_inst; ++i) by_key[std::make_tuple(vt::line_rule(order[i], nullptr), p[i], order[i])].push_back(i);
    long long gb = 0;
    for (auto& kv : by_key) {
        const std::vector<int>& ins = kv.second;
        long long tot = 0;
        for (int i : ins) tot += n_elem[i];
        if (tot + gb >= (1ll << 31)) return fail(VT_ERR_ARG);
        std::vector<double> v(6 * static_cast<size_t>(tot));
        std::vector<int> nb(3 * static_cast<size_t>(tot)), ty(static_cast<size_t>(tot));
        long long off = 0;
        for (int i : ins) {
            b->grp[i] = static_cast<int>(b->g.size());
            b->eoff[i] = static_cast<int>(off);
            std::memcpy(&v[6 * off], vertices + 6 * first[i], sizeof(double) * 6 * n_elem[i]);
            std::memcpy(&ty[off], types + first[i], sizeof(int) * n_elem[i]);
            for (long long k = 0; k < 3ll * n_elem[i]; ++k) nb[3 * off + k] = neighbors[3 * first[i] + k] + static_cast<int>(off);
            off += n_elem[i];
        }
        vt::OperatorTable tab;
        if (!vt::build_operator_table(std::get<1>(kv.first), std::get<2>(kv.first), &tab)) return fail(VT_ERR_QUAD);
        vt_ctx* c = nullptr;
        int rc = vti::create_ctx(static_cast<int>(tot), v.data(), nb.data(), ty.data(), std::get<1>(kv.first), tab, &c);
        if (rc) return fail(rc);
        b->g.push_back(c);
        b->gbase.push_back(gb);
        b->members.push_back(ins);
        gb += tot;
    }
    if (gb * 256 >= (1ll << 40)) return fail(VT_ERR_ARG);  // 40-bit error key
    for (size_t i = 0; i < b->g.size(); ++i) {
        if (!b->launches.empty()) {
            std::vector<int>& L = b->launches.back();
            if (b->g[L[0]]->nq == b->g[i]->nq && static_cast<int>(L.size()) < vt::kMaxGroups) {
                L.push_back(static_cast<int>(i));
                continue;
            }
        }
        b->launches.push_back({static_cast<int>(i)});
    }
    b->join.assign(b->g.size(), nullptr);
    if (!hip_ok(hipEventCreateWithFlags(&b->fork, hipEventDisableTiming), "hipEventCreate"))
        return fail(VT_ERR_HIP);
    for (hipEvent_t& ev : b->join)
        if (!hip_ok(hipEventCreateWithFlags(&ev, hipEventDisableTiming), "hipEventCreate")) return fail(VT_ERR_HIP);
    if (!hip_ok(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate") ||
        !hip_ok(hipMalloc(reinterpret_cast<void**>(&b->err), sizeof(unsigned long long)), "hipMalloc err") ||
        !hip_ok(hipHostMalloc(reinterpret_cast<void**>(&b->err_init), sizeof(unsigned long long), hipHostMallocDefault),
                "hipHostMalloc"))
        return fail(VT_ERR_HIP);
    *b->err_init = vt::kNoError;
    *out = b;
    return VT_OK;
}

void vt_batch_destroy(vt_batch* b) {
    if (!b) return;
    drop_graphs(b);
    for (vt_ctx* c : b->g) vt_destroy(c);
    if (b->err) hipFree(b->err);
    if (b->err_init) hipHostFree(b->err_init);
    if (b->err_host) hipHostFree(b->err_host);
    if (b->fork) hipEventDestroy(b->fork);
    for (hipEvent_t ev : b->join)
        if (ev) hipEventDestroy(ev);
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

int vt_batch_num_instances(const vt_batch* b) { return b ? b->n : 0; }
int vt_batch_num_launches(const vt_batch* b) {
    if (!b) return 0;
    return static_cast<int>(b->cfg.persistent == 2 ? b->g.size() : b->launches.size());
}

int vt_batch_init_vortex(vt_batch* b, int init_type, const double* eps, const double* x0, const double* y0) {
    if (!b || !eps || !x0 || !y0) return VT_ERR_ARG;
    for (int i = 0; i < b->n; ++i) {
        vt_ctx* c = b->g[b->grp[i]];
        std::vector<double> U(static_cast<size_t>(b->ne[i]) * c->np * 4);
        if (!vt::initial_state(b->ne[i], c->vertices.data() + 6 * static_cast<size_t>(b->eoff[i]), c->p, c->order, init_type,
                               eps[i], x0[i], y0[i], b->time, U.data()))
            return VT_ERR_INIT_TYPE;
        int rc = vt_batch_set_state(b, i, U.data());
        if (rc) return rc;
    }
    return VT_OK;
}

int vt_batch_init_vortex_device(vt_batch* b, int init_type, const double* eps, const double* x0, const double* y0) {
    if (!b || !eps || !x0 || !y0) return VT_ERR_ARG;
    for (int gi = 0; gi < static_cast<int>(b->g.size()); ++gi) {
        vt_ctx* c = b->g[gi];
        const std::vector<int>& mem = b->members[gi];
        std::vector<int> e0(mem.size()), n(mem.size());
        std::vector<double> par(4 * mem.size());
        for (size_t k = 0; k < mem.size(); ++k) {
            const int i = mem[k];
            e0[k] = b->eoff[i];
            n[k] = b->ne[i];
            par[4 * k] = b->time;
            par[4 * k + 1] = eps[i];
            par[4 * k + 2] = x0[i];
            par[4 * k + 3] = y0[i];
        }
        int rc = vti::vortex_device(c, static_cast<int>(mem.size()), e0.data(), n.data(), par.data(), init_type, c->Un,
                                    false);
        if (rc) return rc;
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return VT_OK;
}

int vt_batch_set_target_vortex(vt_batch* b, double t, const double* eps, const double* x0, const double* y0) {
    if (!b || !eps || !x0 || !y0) return VT_ERR_ARG;
    b->tgt_par.resize(4 * static_cast<size_t>(b->n));
    for (int i = 0; i < b->n; ++i) {
        b->tgt_par[4 * i] = t;
        b->tgt_par[4 * i + 1] = eps[i];
        b->tgt_par[4 * i + 2] = x0[i];
        b->tgt_par[4 * i + 3] = y0[i];
    }
    return VT_OK;
}

int vt_batch_set_state(vt_batch* b, int i, const double* U) {
    if (check_inst(b, i) || !U) return VT_ERR_ARG;
    vt_ctx* c = b->g[b->grp[i]];
    // Each plane's slice of this instance is contiguous: one strided 2-D copy.
    std::vector<double> h(static_cast<size_t>(4) * c->np * b->ne[i]);
    for (int e = 0; e < b->ne[i]; ++e)
        for (int j = 0; j < c->np; ++j)
            for (int k = 0; k < 4; ++k)
                h[static_cast<size_t>(k * c->np + j) * b->ne[i] + e] = U[(static_cast<size_t>(e) * c->np + j) * 4 + k];
    HIP_TRY(hipMemcpy2DAsync(c->Un + b->eoff[i], sizeof(double) * c->ld, h.data(), sizeof(double) * b->ne[i],
                             sizeof(double) * b->ne[i], 4 * c->np, hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return VT_OK;
}

int vt_batch_get_state(vt_batch* b, int i, double* U) {
    if (check_inst(b, i) || !U) return VT_ERR_ARG;
    vt_ctx* c = b->g[b->grp[i]];
    std::vector<double> h(static_cast<size_t>(4) * c->np * b->ne[i]);
    HIP_TRY(hipMemcpy2DAsync(h.data(), sizeof(double) * b->ne[i], c->Un + b->eoff[i], sizeof(double) * c->ld,
                             sizeof(double) * b->ne[i], 4 * c->np, hipMemcpyDeviceToHost, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (int e = 0; e < b->ne[i]; ++e)
        for (int j = 0; j < c->np; ++j)
            for (int k = 0; k < 4; ++k)
                U[(static_cast<size_t>(e) * c->np + j) * 4 + k] = h[static_cast<size_t>(k * c->np + j) * b->ne[i] + e];
    return VT_OK;
}

double vt_batch_get_time(const vt_batch* b) { return b ? b->time : 0.0; }
int vt_batch_set_time(vt_batch* b, double t) {
    if (!b) return VT_ERR_ARG;
    b->time = t;
    return VT_OK;
}

int vt_batch_step(vt_batch* b, int method, double dt, int nsteps) {
    if (!b) return VT_ERR_ARG;
    return run_steps(b, method, dt, nsteps, nullptr, nullptr);
}

int vt_batch_prepare_steps(vt_batch* b, int method, double dt, int nsteps) {
    if (!b) return VT_ERR_ARG;
    if (method != 0 && method != 1) return VT_ERR_METHOD;
    if (nsteps < 0) return VT_ERR_ARG;
    for (int k : vti::replay_plan(method, nsteps)) {
        if (method == 0 && k == 1) continue;  // a single Euler step is launched directly
        hipGraphExec_t g;
        const int rc = get_graph(b, method, dt, k, &g);
        if (rc) return rc;
    }
    return VT_OK;
}

int vt_batch_step_timed(vt_batch* b, int method, double dt, int nsteps, double* ms) {
    if (!b || !ms) return VT_ERR_ARG;
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0));
    HIP_TRY(hipEventCreate(&e1));
    int rc = run_steps(b, method, dt, nsteps, e0, e1);
    float f = 0.f;
    if (rc == VT_OK && !hip_ok(hipEventElapsedTime(&f, e0, e1), "hipEventElapsedTime")) rc = VT_ERR_HIP;
    *ms = f;
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return rc;
}

int vt_batch_last_error(const vt_batch* b, int* info) {
    if (!b || !info) return VT_ERR_ARG;
    std::memcpy(info, b->last_err, sizeof(b->last_err));
    return VT_OK;
}

int vt_batch_set_launch(vt_batch* b, int block, int persistent, int slots_per_xcd) {
    if (!b || (block != 64 && block != 128 && block != 256) || slots_per_xcd < 0 || persistent < 0 || persistent > 2 ||
        (persistent == 1 && slots_per_xcd < 1))
        return VT_ERR_ARG;
    b->cfg.block = block;
    b->cfg.persistent = persistent;
    b->cfg.slots_per_xcd = slots_per_xcd;
    drop_graphs(b);
    return VT_OK;
}

int vt_batch_get_launch(const vt_batch* b, int* cfg) {
    if (!b || !cfg) return VT_ERR_ARG;
    cfg[0] = b->cfg.block;
    cfg[1] = b->cfg.persistent;
    cfg[2] = b->cfg.slots_per_xcd;
    return VT_OK;
}

// Autotuner: times trial_steps RK4 steps (after one untimed replay of the
// same graph) for each candidate launch shape, from the current state each
// time, and keeps the fastest.  State and time are restored afterwards.
int vt_batch_autotune(vt_batch* b, double dt, int trial_steps, int* best_cfg, double* best_ms) {
    if (!b || trial_steps < 1 || trial_steps > vti::kGraphSteps) return VT_ERR_ARG;
    struct Cand { int block, persistent, slots; };
    // Workgroups resident per CU at p = 3 (LDS 320 B and ~250 VGPRs per
    // element): 2 of 256 threads, 4 of 128, 8 of 64; 32 CUs per XCD.
    const Cand cands[] = {{64, 0, 0},    {128, 0, 0},  {256, 0, 0}, {64, 1, 256}, {128, 1, 128},
                          {256, 1, 64},  {128, 1, 64}, {256, 1, 32}, {64, 2, 0},   {128, 2, 0}};
    std::vector<double*> saved(b->g.size(), nullptr);
    auto cleanup = [&]() { for (double* s : saved) if (s) hipFree(s); };
    for (size_t i = 0; i < b->g.size(); ++i) {
        vt_ctx* c = b->g[i];
        if (!hip_ok(hipMalloc(reinterpret_cast<void**>(&saved[i]), sizeof(double) * 4 * c->np * c->ld), "hipMalloc")) {
            cleanup();
            return VT_ERR_HIP;
        }
        if (!hip_ok(hipMemcpyAsync(saved[i], c->Un, sizeof(double) * 4 * c->np * c->ld, hipMemcpyDeviceToDevice, b->stream),
                    "save state")) {
            cleanup();
            return VT_ERR_HIP;
        }
    }
    auto restore = [&]() {
        for (size_t i = 0; i < b->g.size(); ++i) {
            vt_ctx* c = b->g[i];
            if (!hip_ok(hipMemcpyAsync(c->Un, saved[i], sizeof(double) * 4 * c->np * c->ld, hipMemcpyDeviceToDevice, b->stream),
                        "restore state"))
                return false;
        }
        return hip_ok(hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    };
    const double t0 = b->time;
    double best = 1e300;
    Cand win = cands[0];
    int rc = VT_OK;
    for (const Cand& cd : cands) {
        vt_batch_set_launch(b, cd.block, cd.persistent, cd.slots);
        double ms = 0.0;
        rc = run_steps(b, 1, dt, trial_steps, nullptr, nullptr);
        if (rc == VT_OK) rc = vt_batch_step_timed(b, 1, dt, trial_steps, &ms);
        if (!restore()) rc = VT_ERR_HIP;
        b->time = t0;
        if (rc) break;
        if (ms < best) { best = ms; win = cd; }
    }
    cleanup();
    if (rc) return rc;
    vt_batch_set_launch(b, win.block, win.persistent, win.slots);
    if (best_cfg) { best_cfg[0] = win.block; best_cfg[1] = win.persistent; best_cfg[2] = win.slots; }
    if (best_ms) *best_ms = best / trial_steps;
    return VT_OK;
}

int vt_batch_exact_vortex(const vt_batch* b, int i, double t, double epsilon, double x0, double y0, double* U) {
    if (check_inst(b, i) || !U) return VT_ERR_ARG;
    const vt_ctx* c = b->g[b->grp[i]];
    vt::initial_state(b->ne[i], c->vertices.data() + 6 * static_cast<size_t>(b->eoff[i]), c->p, c->order, 0, epsilon,
                      x0, y0, t, U);
    return VT_OK;
}

int vt_batch_mass_l2_objective(vt_batch* b, const double* rho_target, double* J, double* dJdU) {
    if (!b || !J) return VT_ERR_ARG;
    const std::vector<size_t> off = node_offsets(b);
    for (int gi = 0; gi < static_cast<int>(b->g.size()); ++gi) {
        int rc = group_objective(b, gi, rho_target, off, J);
        if (!rc && dJdU) rc = scatter_group_gradient(b, gi, off, dJdU);
        if (rc) return rc;
    }
    return VT_OK;
}

int vt_batch_adjoint_release(vt_batch* b) {
    if (!b) return VT_ERR_ARG;
    for (vt_ctx* c : b->g) vti::free_adjoint(c);
    return VT_OK;
}

int vt_batch_adjoint(vt_batch* b, double dt, int nsteps, int every, const double* rho_target, double* J,
                     double* dJdU0) {
    if (!b || !J || !dJdU0 || nsteps < 1 || every < 1 || (!rho_target && b->tgt_par.empty())) return VT_ERR_ARG;
    return batch_adjoint(b, dt, nsteps, every, rho_target, J, [&](int gi, const std::vector<size_t>& off) {
        return scatter_group_gradient(b, gi, off, dJdU0);
    });
}

int vt_batch_adjoint_vortex(vt_batch* b, double dt, int nsteps, int every, const double* rho_target,
                            const double* eps, const double* x0, const double* y0, double* J, double* grad) {
    if (!b || !eps || !x0 || !y0 || !J || !grad || nsteps < 1 || every < 1 || (!rho_target && b->tgt_par.empty()))
        return VT_ERR_ARG;
    const double t0 = b->time;
    return batch_adjoint(b, dt, nsteps, every, rho_target, J, [&](int gi, const std::vector<size_t>&) {
        return group_vortex_gradient(b, gi, t0, eps, x0, y0, grad);
    });
}

int vt_batch_vortex_gradient(const vt_batch* b, double t, const double* eps, const double* x0, const double* y0,
                             const double* dJdU0, double* grad) {
    if (!b || !eps || !x0 || !y0 || !dJdU0 || !grad) return VT_ERR_ARG;
    const std::vector<size_t> off = node_offsets(b);
    for (int i = 0; i < b->n; ++i) {
        const vt_ctx* c = b->g[b->grp[i]];
        vt::vortex_gradient(b->ne[i], c->vertices.data() + 6 * static_cast<size_t>(b->eoff[i]), c->p, t, eps[i], x0[i],
                            y0[i], dJdU0 + 4 * off[i], grad + 3 * i);
    }
    return VT_OK;
}

}  // extern "C"
