// vt_adjoint.hip -- C ABI of the discrete adjoint of the RK4 loop: recorded
// forward runs with checkpoints, the checkpointed reverse sweep, the adjoint
// of a single right-hand side, and host helpers (vortex initial-state tangent,
// mass-weighted L2 objective).  No reference counterpart: the reference has no
// derivatives.
//
// Memory (one "slot" = one state, 4 * Np * ld doubles):
//   ckpt      ceil(nsteps / every) slots, the state at the start of each segment;
//   seg       4 * every slots, the four stage states of every step of one
//             segment (recomputed from its checkpoint during the reverse sweep;
//             the checkpoint itself serves as the first step's start);
//   adj       4 slots: lam, acc, y (ping-pong pair).
// The reverse sweep launches, per segment, the forward stage kernels of the
// segment with their outputs directed into seg (no extra copies), then per step
// the four adjoint stage kernels.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "vt_ctx.hpp"
#include "vt_device.hpp"

namespace {

using vti::hip_ok;

size_t slot_elems(const vt_ctx* c) { return static_cast<size_t>(4) * c->np * c->ld; }

}  // namespace

int vti::ensure_adj_planes(vt_ctx* c) {
    if (c->adj) return VT_OK;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->adj), sizeof(double) * 4 * slot_elems(c)));
    // On the context stream: it does not synchronise with the null stream.
    HIP_TRY(hipMemsetAsync(c->adj, 0, sizeof(double) * 4 * slot_elems(c), c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return VT_OK;
}

int vti::objective(vt_ctx* c, int nr, const int* e0, const int* n, const double* const* rho, const double* vpar,
                   double* J) {
    if (c->order < 0) return VT_ERR_QUAD;  // M_ref comes from the integration order's area rule
    int rc = ensure_adj_planes(c);
    if (rc) return rc;
    std::vector<int> boff(nr + 1, 0);
    for (int r = 0; r < nr; ++r) boff[r + 1] = boff[r] + (n[r] + vt::kObjBlock - 1) / vt::kObjBlock;
    if (!c->mref) {
        std::vector<double> M(static_cast<size_t>(c->np) * c->np);
        if (!vt::reference_operators(c->p, c->order, M.data(), nullptr, nullptr)) return VT_ERR_QUAD;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->mref), sizeof(double) * M.size()));
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->tgt), sizeof(double) * c->np * c->ld));
        HIP_TRY(hipMemcpyAsync(c->mref, M.data(), sizeof(double) * M.size(), hipMemcpyHostToDevice, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    if (c->part_cap < boff[nr]) {
        if (c->part) (void)hipFree(c->part);
        c->part = nullptr;
        c->part_cap = 0;
        HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->part), sizeof(double) * boff[nr]));
        c->part_cap = boff[nr];
    }
    double* h = host_stage(c, rho ? std::max(static_cast<size_t>(c->np) * c->ld, static_cast<size_t>(boff[nr]))
                                  : static_cast<size_t>(boff[nr]));
    if (!h) return VT_ERR_HIP;
    if (rho) {
        for (int r = 0; r < nr; ++r) vt::density_to_planes(n[r], c->np, c->ld, rho[r], h + e0[r]);
        HIP_TRY(hipMemcpyAsync(c->tgt, h, sizeof(double) * c->np * c->ld, hipMemcpyHostToDevice, c->stream));
    } else {
        if (!vpar) return VT_ERR_ARG;
        rc = vortex_device(c, nr, e0, n, vpar, 0, c->tgt, true);
        if (rc) return rc;
    }
    // One launch per element range: the plane base pointers move to the
    // range's first element, the plane stride stays ld.
    for (int r = 0; r < nr; ++r) {
        vt::ObjParams o{};
        o.U = c->Un + e0[r];
        o.tgt = c->tgt + e0[r];
        o.mref = c->mref;
        o.vtx = c->vtx + e0[r];
        o.lam = c->adj + e0[r];
        o.part = c->part + boff[r];
        o.ne = n[r];
        o.ld = c->ld;
        HIP_TRY(vt::launch_objective(o, c->p, c->stream));
    }
    HIP_TRY(hipMemcpyAsync(h, c->part, sizeof(double) * boff[nr], hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int r = 0; r < nr; ++r) {
        double sum = 0.0;
        for (int b = boff[r]; b < boff[r + 1]; ++b) sum += h[b];
        J[r] = sum;
    }
    return VT_OK;
}

namespace {

vt::AdjParams adj_params(vt_ctx* c, int mode) {
    vt::AdjParams a{};
    a.vtx = c->vtx;
    a.nbr = c->nbr;
    a.tab = c->tab;
    a.ne = c->ne;
    a.ld = c->ld;
    a.mode = mode;
    a.blocks_per_xcd = ((c->ne + vt::kAdjBlock - 1) / vt::kAdjBlock + 7) / 8;
    a.pairs = c->pairs;
    return a;
}

int upload_planes(vt_ctx* c, const double* U, double* dev) { return vti::upload_state(c, U, dev); }

int download_planes(vt_ctx* c, const double* dev, double* U) { return vti::download_state(c, dev, U); }

// Forward stage kernels of one step from slot S0 into slots S1, S2, S3 and
// next; with next null the fourth stage (whose output the reverse sweep of the
// step does not read) is skipped.  Failures carry the stage ids
// stage_base + 0..3.
hipError_t recompute_step(vt_ctx* c, double dt, double* S0, double* S1, double* S2, double* S3, double* next,
                          int stage_base, hipStream_t st) {
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
    prm.dt = dt;
    prm.acc = c->acc;
    const double* in[4] = {S0, S1, S2, S3};
    double* out[4] = {S1, S2, S3, next};
    for (int s = 0; s < (next ? 4 : 3); ++s) {
        prm.Uin = in[s];
        prm.Uout = out[s];
        prm.Un = (s == 1 || s == 2) ? S0 : nullptr;
        prm.U2 = s == 2 ? S1 : nullptr;
        prm.mode = vt::kModeRk1 + s;
        prm.stage_id = stage_base + s;
        hipError_t e = vt::launch_stage(prm, c->p, c->nq, c->cfg, st);
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

// The stage states of one segment: 4 * min(nsteps, every) slots.
int ensure_seg(vt_ctx* c, int nsteps, int every) {
    const int need = 4 * (nsteps < every ? nsteps : every);
    if (c->seg_slots >= need) return VT_OK;
    c->rec_last_stored = false;
    if (c->seg) (void)hipFree(c->seg);
    c->seg = nullptr;
    c->seg_slots = 0;
    HIP_TRY(hipMalloc(reinterpret_cast<void**>(&c->seg), sizeof(double) * slot_elems(c) * need));
    c->seg_slots = need;
    return VT_OK;
}

// Headroom left on the device when the second slot set is allocated.
constexpr size_t kOverlapHeadroom = size_t(2) << 30;

// The second slot set (as large as seg), the recompute stream and its events.
// Returns false, and the sweep runs on one stream, when the set does not fit
// beside everything else (the 32-instance ensemble adjoint does not).
bool ensure_overlap(vt_ctx* c) {
    if (!c->rstream) {
        hipEvent_t* evs[] = {&c->ev_fork, &c->ev_rec[0], &c->ev_rec[1], &c->ev_swept[0], &c->ev_swept[1]};
        bool ok = hip_ok(hipStreamCreateWithFlags(&c->rstream, hipStreamNonBlocking), "hipStreamCreate");
        for (hipEvent_t* ev : evs)
            ok = ok && hip_ok(hipEventCreateWithFlags(ev, hipEventDisableTiming), "hipEventCreate");
        if (!ok) {  // all or nothing, so that a later call starts over
            for (hipEvent_t* ev : evs) {
                if (*ev) (void)hipEventDestroy(*ev);
                *ev = nullptr;
            }
            if (c->rstream) (void)hipStreamDestroy(c->rstream);
            c->rstream = nullptr;
            return false;
        }
    }
    if (c->seg2 && c->seg2_slots == c->seg_slots) return true;
    if (c->seg2) (void)hipFree(c->seg2);
    c->seg2 = nullptr;
    c->seg2_slots = 0;
    const size_t bytes = sizeof(double) * slot_elems(c) * c->seg_slots;
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || free_b < bytes + kOverlapHeadroom) return false;
    if (hipMalloc(reinterpret_cast<void**>(&c->seg2), bytes) != hipSuccess) {
        c->seg2 = nullptr;
        (void)hipGetLastError();
        return false;
    }
    c->seg2_slots = c->seg_slots;
    return true;
}

}  // namespace

int vti::backward_sweep(vt_ctx* c) {
    const int every = c->rec_every, nsteps = c->rec_steps, nseg = (nsteps + every - 1) / every;
    const double dt = c->rec_dt;
    const size_t n = slot_elems(c);
    int rc = ensure_seg(c, nsteps, every);
    if (rc) return rc;
    const bool overlap = c->adj_overlap && nseg > 1 && ensure_overlap(c);
    double *lam = c->adj, *acc = c->adj + n, *y[2] = {c->adj + 2 * n, c->adj + 3 * n};
    vt::AdjParams a = adj_params(c, vt::kAdjSeed);
    a.lam = lam;
    a.Yout = y[0];
    a.scale = 1.0 / 6.0;
    HIP_TRY(vt::launch_adjoint(a, c->p, c->nq, c->stream));
    int cur = 0;
    // Segment i's stage states: its first step starts from its checkpoint in
    // place; the rest sit in seg, or with overlap in seg and seg2 alternately
    // (the last segment, which the recording may have stored, in seg).
    auto set_of = [&](int i) { return (overlap && ((nseg - 1 - i) & 1)) ? c->seg2 : c->seg; };
    auto slot = [&](int i, int step, int stage) {
        return (step == 0 && stage == 0) ? c->ckpt[i] : set_of(i) + (static_cast<size_t>(4) * step + stage) * n;
    };
    auto seg_len = [&](int i) { return (i + 1 < nseg) ? every : nsteps - i * every; };
    auto recompute = [&](int i, hipStream_t st) -> hipError_t {
        const int len = seg_len(i);
        for (int s = 0; s < len; ++s) {
            const hipError_t e = recompute_step(c, dt, slot(i, s, 0), slot(i, s, 1), slot(i, s, 2), slot(i, s, 3),
                                                s + 1 < len ? slot(i, s + 1, 0) : nullptr, 0, st);
            if (e != hipSuccess) return e;
        }
        return hipSuccess;
    };
    auto sweep = [&](int i) -> hipError_t {
        for (int s = seg_len(i) - 1; s >= 0; --s)
            for (int st = 3; st >= 0; --st) {
                vt::AdjParams b = adj_params(c, vt::kAdj1 + st);
                b.U = slot(i, s, st);
                b.Y = y[cur];
                b.Yout = y[cur ^ 1];
                b.lam = lam;
                b.acc = acc;
                b.dt = dt;
                const hipError_t e = vt::launch_adjoint(b, c->p, c->nq, c->stream);
                if (e != hipSuccess) return e;
                cur ^= 1;
            }
        return hipSuccess;
    };
    // The last segment's stage states were kept by the recording itself
    // (until a recompute of an earlier segment overwrites them).
    const bool stored = c->rec_last_stored;
    if (!overlap) {
        for (int i = nseg - 1; i >= 0; --i) {
            if (!(i == nseg - 1 && stored)) {
                c->rec_last_stored = false;
                HIP_TRY(recompute(i, c->stream));
            }
            HIP_TRY(sweep(i));
        }
        return VT_OK;
    }
    // Overlapped: while segment i is swept on the context stream, segment
    // i - 1 is recomputed on rstream into the other set, once the sweep of
    // segment i + 1 (the set's previous user) is done; the sweep of segment
    // i - 1 waits for its recompute.  The recompute stream starts after all
    // work queued so far on the context stream.
    c->rec_last_stored = false;
    HIP_TRY(hipEventRecord(c->ev_fork, c->stream));
    HIP_TRY(hipStreamWaitEvent(c->rstream, c->ev_fork, 0));
    if (!stored) {
        HIP_TRY(recompute(nseg - 1, c->rstream));
        HIP_TRY(hipEventRecord(c->ev_rec[0], c->rstream));
        HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_rec[0], 0));
    }
    for (int i = nseg - 1; i >= 0; --i) {
        const int b = (nseg - 1 - i) & 1;
        if (i > 0) {
            if (i + 1 < nseg) HIP_TRY(hipStreamWaitEvent(c->rstream, c->ev_swept[b ^ 1], 0));
            HIP_TRY(recompute(i - 1, c->rstream));
            HIP_TRY(hipEventRecord(c->ev_rec[b ^ 1], c->rstream));
        }
        HIP_TRY(sweep(i));
        HIP_TRY(hipEventRecord(c->ev_swept[b], c->stream));
        if (i > 0) HIP_TRY(hipStreamWaitEvent(c->stream, c->ev_rec[b ^ 1], 0));
    }
    return VT_OK;
}

void vti::free_adjoint(vt_ctx* c) {
    for (double* p : c->ckpt)
        if (p) (void)hipFree(p);
    c->ckpt.clear();
    if (c->seg) (void)hipFree(c->seg);
    if (c->seg2) (void)hipFree(c->seg2);
    if (c->adj) (void)hipFree(c->adj);
    for (double* p : {c->mref, c->tgt, c->part})
        if (p) (void)hipFree(p);
    for (hipEvent_t* ev : {&c->ev_fork, &c->ev_rec[0], &c->ev_rec[1], &c->ev_swept[0], &c->ev_swept[1]}) {
        if (*ev) (void)hipEventDestroy(*ev);
        *ev = nullptr;
    }
    if (c->rstream) (void)hipStreamDestroy(c->rstream);
    c->rstream = nullptr;
    c->seg2 = nullptr;
    c->seg2_slots = 0;
    c->seg = nullptr;
    c->adj = nullptr;
    c->mref = c->tgt = c->part = nullptr;
    c->part_cap = 0;
    c->seg_slots = 0;
    c->recorded = false;
    c->rec_last_stored = false;
}

extern "C" {

int vt_rhs_vjp(vt_ctx* c, const double* w, double* g) {
    if (!c || !w || !g) return VT_ERR_ARG;
    int rc = vti::ensure_adj_planes(c);
    if (rc) return rc;
    const size_t n = slot_elems(c);
    double *lam = c->adj, *ya = c->adj + 2 * n, *yb = c->adj + 3 * n;
    rc = upload_planes(c, w, lam);
    if (rc) return rc;
    vt::AdjParams a = adj_params(c, vt::kAdjSeed);
    a.lam = lam;
    a.Yout = ya;
    a.scale = 1.0;
    HIP_TRY(vt::launch_adjoint(a, c->p, c->nq, c->stream));
    a = adj_params(c, vt::kAdjRaw);
    a.U = c->Un;
    a.Y = ya;
    a.Yout = yb;
    HIP_TRY(vt::launch_adjoint(a, c->p, c->nq, c->stream));
    return download_planes(c, yb, g);
}

int vt_adjoint_forward(vt_ctx* c, double dt, int nsteps, int every) {
    if (!c || nsteps < 1 || every < 1) return VT_ERR_ARG;
    c->recorded = false;
    ++c->rec_id;
    const int nseg = (nsteps + every - 1) / every;
    const size_t n = slot_elems(c);
    if (static_cast<int>(c->ckpt.size()) < nseg) {
        for (int i = static_cast<int>(c->ckpt.size()); i < nseg; ++i) {
            double* p = nullptr;
            HIP_TRY(hipMalloc(reinterpret_cast<void**>(&p), sizeof(double) * n));
            c->ckpt.push_back(p);
        }
    }
    // The recording runs the per-stage kernels, whose stage states the reverse
    // sweep recomputes from the checkpoints (not the small-mesh path, which
    // rounds the RK update differently).
    struct StageKernels {  // restores the context's choice on every return
        vt_ctx* c;
        bool was;
        explicit StageKernels(vt_ctx* x) : c(x), was(x->small_on) { c->small_on = false; }
        ~StageKernels() { c->small_on = was; }
    } guard(c);
    c->rec_last_stored = false;
    int rc = ensure_seg(c, nsteps, every);
    if (rc) return rc;
    for (int i = 0; i < nseg; ++i) {
        const int len = (i + 1 < nseg) ? every : nsteps - i * every;
        HIP_TRY(hipMemcpyAsync(c->ckpt[i], c->Un, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        if (i + 1 < nseg) {
            rc = vti::run_steps(c, 1, dt, len, nullptr, nullptr);
            if (rc) {
                if (rc == VT_ERR_NONPHYSICAL) c->last_err[4] += i * every;
                return rc;
            }
            continue;
        }
        // The last segment, which the reverse sweep takes first, runs with its
        // stage states written straight into the segment slots, so the sweep
        // need not recompute it (with every >= nsteps: the whole run).
        auto slot = [&](int step, int stage) {
            return (step == 0 && stage == 0) ? c->ckpt[i] : c->seg + (static_cast<size_t>(4) * step + stage) * n;
        };
        rc = vti::ensure_err_slots(&c->err_host, &c->err_host_cap, 1);
        if (rc) return rc;
        HIP_TRY(hipMemcpyAsync(c->err, c->err_init, sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream));
        for (int s = 0; s < len; ++s)
            HIP_TRY(recompute_step(c, dt, slot(s, 0), slot(s, 1), slot(s, 2), slot(s, 3),
                                   s + 1 < len ? slot(s + 1, 0) : c->Un, 4 * s, c->stream));
        HIP_TRY(hipMemcpyAsync(c->err_host, c->err, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        if (c->err_host[0] != vt::kNoError) {
            vti::decode_error(c, c->err_host[0], i * every, 4);
            return VT_ERR_NONPHYSICAL;
        }
        for (int s = 0; s < len; ++s) c->time += dt;
        c->rec_last_stored = true;
    }
    c->rec_steps = nsteps;
    c->rec_every = every;
    c->rec_dt = dt;
    c->recorded = true;
    return VT_OK;
}

int vt_adjoint_release(vt_ctx* c) {
    if (!c) return VT_ERR_ARG;
    vti::free_adjoint(c);
    return VT_OK;
}

int vt_set_adjoint_overlap(vt_ctx* c, int on) {
    if (!c) return VT_ERR_ARG;
    c->adj_overlap = on != 0;
    return VT_OK;
}

int vt_adjoint_recording(const vt_ctx* c, unsigned long long* id) {
    if (!c || !id) return VT_ERR_ARG;
    *id = c->recorded ? c->rec_id : 0;
    return VT_OK;
}

int vt_adjoint_backward(vt_ctx* c, const double* lamT, double* lam0) {
    if (!c || !lamT || !lam0) return VT_ERR_ARG;
    if (!c->recorded) return VT_ERR_ARG;
    int rc = vti::ensure_adj_planes(c);
    if (rc) return rc;
    rc = upload_planes(c, lamT, c->adj);
    if (rc) return rc;
    rc = vti::backward_sweep(c);
    if (rc) return rc;
    return download_planes(c, c->adj, lam0);
}

int vt_vortex_tangent(const vt_ctx* c, double t, double epsilon, double x0, double y0, double* dU) {
    if (!c || !dU) return VT_ERR_ARG;
    vt::vortex_tangent(c->ne, c->vertices.data(), c->p, t, epsilon, x0, y0, dU);
    return VT_OK;
}

int vt_mass_l2_objective(vt_ctx* c, const double* rho_target, double* J, double* dJdU) {
    if (!c || !J) return VT_ERR_ARG;
    const int e0 = 0;
    if (!rho_target && !c->tgt_vortex) return VT_ERR_ARG;
    int rc = vti::objective(c, 1, &e0, &c->ne, rho_target ? &rho_target : nullptr, c->tgt_par, J);
    if (rc || !dJdU) return rc;
    return download_planes(c, c->adj, dJdU);
}

int vt_adjoint(vt_ctx* c, double dt, int nsteps, int every, const double* rho_target, double* J, double* dJdU0) {
    if (!c || !J || !dJdU0 || (!rho_target && !c->tgt_vortex)) return VT_ERR_ARG;
    int rc = vt_adjoint_forward(c, dt, nsteps, every);
    if (rc) return rc;
    // The seed dJ/dU_T goes from the objective kernel straight into lam.
    const int e0 = 0;
    rc = vti::objective(c, 1, &e0, &c->ne, rho_target ? &rho_target : nullptr, c->tgt_par, J);
    if (rc) return rc;
    rc = vti::backward_sweep(c);
    if (rc) return rc;
    return download_planes(c, c->adj, dJdU0);
}

}  // extern "C"
