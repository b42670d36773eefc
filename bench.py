"""Benchmark: RK4 DOF-stage updates/s of the HIP path at p = 3, FP64.

Workload (one GPU): the isentropic vortex on 512 x 512 cells (524,288
triangles) of a 20 x 20 periodic domain, interior vertices moved by
0.02 h U(-1,1) (seed 2), p = 3, integration_order = 8, classic RK4.  A step
is one RK4 step over the whole mesh (four fused stage kernels); the state is
resident in HBM before timing starts.

dt = 1e-4: at h = 20/512 the configured dt = 1e-3 makes the reference
numerics leave the physical range in step 4 (its Roe eigenvalues use
ci = c2^2 ~ 16), after which the reference exits.  The work per step does not
depend on dt.

Beside the plain figure the line carries "batched" (the 64-instance
ensemble, one launch per RK stage, against per-instance plain solvers),
"adjoint" (forward + checkpointed discrete adjoint on 256 x 256 cells, against
the forward solve alone), "batched_adjoint" (32 uneven instances up to
512 x 512 cells, recorded forward + reverse sweep + per-instance parameter
gradients, against the batched forward), "large_mesh" (a forward-only
1024 x 1024-cell run), "full_run" (the main workload's 2,000 configured
steps to the end), "c1_full_run" and "shipped_input_run" (the
reference's own 8 x 8 case to the end at integration orders 8 and 6, the
latter its shipped input, beside the rk4 times the reference recorded for it).

Prints one JSON line (rank 0).  --gpus N > 1 exits: the implementation runs on one GPU.
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 DOF-stage updates/s at p=3 FP64; batched & adjoint vs plain at 1e-12"
HBM_PEAK_GBS = 8000.0
# Algorithmic HBM bytes per element per RK4 step in the 3-array form
# (U_n, stage state, accumulator): 960 + 1600 + 1600 + 960 = 5120 B at p = 3,
# i.e. 1280 B per element per stage launch on average.
BYTES_PER_ELEM_STEP = 5120.0
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "pmc_stage_kernel_p3_512x512.txt")
FP64_FILE = os.path.join(ROOT, "profiles", "pmc_fp64_stage_kernel_p3_512x512.txt")
TRACE_FILE = os.path.join(ROOT, "profiles", "stage_kernel_p3_512x512_kernel_trace.md")
FP64_PEAK_TFLOPS = 78.6
# Algorithmic FP64 operations per element per stage at p = 3 (traces, 15 Roe
# fluxes, lift, volume, nodal flux, residual, M^-1, update; FMA = 2).
FLOPS_PER_ELEM_STAGE = 7565.0


def parse_args():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--cells", type=int, default=512)
    ap.add_argument("--dt", type=float, default=1e-4)
    ap.add_argument("--kernel-steps", type=int, default=64, help="RK4 steps replayed for the per-launch duration")
    ap.add_argument("--settle-s", type=float, default=0.5,
                    help="seconds of untimed stepping between the warmup and the timed steps (clock transient)")
    ap.add_argument("--block", type=int, default=0, help="elements per workgroup of the stage kernel (0: library default)")
    ap.add_argument("--cpu-steps", type=int, default=3, help="steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--orders", action="store_true",
                    help="also time the 512 x 512 mesh at p = 1 and 2 (ns per element-stage per order)")
    ap.add_argument("--size-sweep", action="store_true",
                    help="per-order ms per step of single meshes from 32^2 to 1024^2 cells and the ensemble's speed-up ceiling")
    ap.add_argument("--batched-by-order", action="store_true",
                    help="also time the ensemble's instances of each order p batched alone, against their plain runs")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--batched-steps", type=int, default=64)
    ap.add_argument("--adjoint-steps", type=int, default=1000)
    ap.add_argument("--no-batched-adjoint", action="store_true")
    ap.add_argument("--batched-adjoint-steps", type=int, default=500)
    ap.add_argument("--batched-adjoint-every", type=int, default=16, help="checkpoint interval of the ensemble adjoint")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--no-full-run", action="store_true", help="skip the 2,000-step run of the main workload")
    ap.add_argument("--no-shard", action="store_true")
    ap.add_argument("--no-pcie", action="store_true")
    ap.add_argument("--no-c1", action="store_true")
    ap.add_argument("--large-steps", type=int, default=20)
    ap.add_argument("--profile-pass", action="store_true",
                    help="for rocprofv3 --pmc: 2 steps of per-launch stage kernels, then a plain device copy "
                         "of 2^27 doubles (1 GiB read + 1 GiB written) to calibrate FETCH_SIZE / WRITE_SIZE")
    return ap.parse_args()


def settle_clocks(s, dt, seconds):
    """Untimed RK4 steps (16 per call) until `seconds` of stepping have passed,
    run right before the timed steps.  After an idle gap the stage kernels run
    up to 15% slower for the first ~100 launches (a clock/power transient seen
    in kernel traces: fast for a few launches, slower, then back to a steady
    rate), so a timed window of a few steps that starts cold measures that
    transient rather than the steady rate of a long run.  Returns the steps run."""
    n = 0
    s.prepare(1, dt, 16)
    t_end = time.perf_counter() + seconds
    while time.perf_counter() < t_end:
        s.rk4(dt, 16)
        n += 16
    return n


def stream_rates(n=1 << 28, reps=10):
    """Practical HBM rates of plain streaming kernels on the same device, as a
    yardstick beside the published 8 TB/s: torch's copy (read + write), triad
    c = a + 2 b (two reads, one write), sum (read) and fill (write) over
    2 GiB float64 arrays, HIP events around `reps` back-to-back launches."""
    import torch
    a = torch.ones(n, dtype=torch.float64, device="cuda")
    b = torch.ones_like(a)
    c = torch.empty_like(a)
    out = {}
    for name, f, nbytes in (("copy", lambda: c.copy_(a), 16 * n), ("triad", lambda: torch.add(a, b, alpha=2.0, out=c), 24 * n),
                            ("read", lambda: a.sum(), 8 * n), ("write", lambda: c.fill_(1.0), 8 * n)):
        for _ in range(3):
            f()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            f()
        e1.record()
        torch.cuda.synchronize()
        out[name + "_GBps"] = nbytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9
    del a, b, c
    torch.cuda.empty_cache()
    return out


def profile_pass(s, dt):
    import torch
    s.stage_kernel_times(dt, 2)
    a = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"profile_pass": True, "stage_launches": 8, "copy_bytes_read": a.numel() * 8,
                      "copy_bytes_written": b.numel() * 8, "elements": s.ne}))


def profile_value(path, key, n_elem):
    """A per-launch figure from a committed rocprofv3 --pmc summary, or None
    when the summary is absent or for another mesh size."""
    try:
        txt = open(path).read()
    except OSError:
        return None
    m = re.search(r"^%s:\s*([0-9.eE+]+)" % key, txt, re.M)
    e = re.search(r"^elements:\s*(\d+)", txt, re.M)
    if not m or not e or int(e.group(1)) != n_elem:
        return None
    return float(m.group(1))


def traffic_from_profile(n_elem):
    """HBM bytes per stage launch measured with rocprofv3 --pmc (FETCH_SIZE,
    WRITE_SIZE), scaled by the copy-kernel calibration."""
    return profile_value(TRAFFIC_FILE, "bytes_per_launch_avg", n_elem)


def trace_avg_launch_ms(n_elem, block):
    """Average duration of vt::stage_kernel<3, 5, block> in the committed
    rocprofv3 --kernel-trace --stats summary, or None when the summary is
    absent or for another mesh size."""
    try:
        txt = open(TRACE_FILE).read()
    except OSError:
        return None
    e = re.search(r"\(([0-9,]+) triangles\)", txt)
    m = re.search(r"stage_kernel<3, 5, %d>\(vt::StageParams\)` \| [0-9,]+ \| ([0-9,]+) \|" % block, txt)
    if not e or not m or int(e.group(1).replace(",", "")) != n_elem:
        return None
    return float(m.group(1).replace(",", "")) * 1e-6


def batched_figure(vt, steps, dt=1e-4):
    """Ensemble of 64 instances (seed 3: cells per side 32..256, p 1..3,
    epsilon 0.1..0.4, centre in [-2.5, 2.5]^2 of a 10 x 10 domain), every
    instance's RK stage in one graph with the autotuned launch shape (one
    launch for all groups, or one per (p, order) group run concurrently),
    against the same 64
    instances stepped one plain solver after another; medians of three runs
    of each."""
    import torch
    draws = vt.ensemble_draws(3, 64)
    meshes = [vt.generate_mesh(nx=d["cells"], ny=d["cells"]) for d in draws]
    ps = [d["p"] for d in draws]
    args = ([d["epsilon"] for d in draws], [d["x0"] for d in draws], [d["y0"] for d in draws])
    b = vt.Batch(meshes, ps, 8)
    b.initialize_hydrodinamics(0, *args)
    cfg, _ = b.autotune(dt, 8)
    b.rk4(dt, 4)
    solvers = []
    for m, d in zip(meshes, draws):
        s = vt.Solver(m, d["p"], 8)
        s.initialize_hydrodinamics(0, d["epsilon"], d["x0"], d["y0"])
        s.rk4(dt, 4)
        solvers.append(s)
    # three alternating runs of each (every instance advances by the same steps
    # in both forms), medians reported
    wbs, wps = [], []
    for _ in range(3):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.rk4(dt, steps)
        wbs.append(time.perf_counter() - t0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for s in solvers:
            s.rk4(dt, steps)
        wps.append(time.perf_counter() - t0)
    wb, wp = float(np.median(wbs)), float(np.median(wps))
    worst = max(float(np.linalg.norm(b.state(i) - s.state()) / np.linalg.norm(s.state())) for i, s in enumerate(solvers))
    dof = sum(2 * d["cells"] ** 2 * (d["p"] + 1) * (d["p"] + 2) // 2 * 16 for d in draws)
    # algorithmic HBM bytes per stage launch: 1,280 B per p = 3 element (the
    # three-array RK4 form), i.e. 128 B per node, over all instances
    stage_bytes = sum(2 * d["cells"] ** 2 * (d["p"] + 1) * (d["p"] + 2) // 2 * 128 for d in draws)
    for s in solvers:
        s.close()
    launches = b.num_launches
    b.close()
    return {"workload": "64 instances (seed 3): cells 32..256 per side, p 1..3, eps 0.1..0.4, 10x10 domain, "
                        "integration_order 8, RK4, dt %g" % dt,
            "elements": sum(2 * d["cells"] ** 2 for d in draws), "steps": steps,
            "ms_per_step": wb / steps * 1e3, "value": dof * steps / wb, "unit": "DOF-stage updates/s",
            "per_instance_ms_per_step": wp / steps * 1e3, "speedup_vs_per_instance": wp / wb,
            "runs_ms_per_step": {"batched": [w / steps * 1e3 for w in wbs], "per_instance": [w / steps * 1e3 for w in wps]},
            "hbm_algorithmic_GBps": stage_bytes / (wb / steps / 4) / 1e9,
            "hbm_frac": stage_bytes / (wb / steps / 4) / 8e12,
            "launch": cfg, "launches_per_stage": launches, "max_rel_l2_vs_per_instance": worst}


def batched_by_order_figure(vt, steps, dt=1e-4, block=64):
    """The 64-instance ensemble's instances of one order p at a time (and all
    together), batched with 64-element tiles, against the same instances
    stepped one plain solver after another; medians of three runs of each."""
    import torch
    draws = vt.ensemble_draws(3, 64)
    out = {}
    for label, sel in (("p1", (1,)), ("p2", (2,)), ("p3", (3,)), ("all", (1, 2, 3))):
        ds = [d for d in draws if d["p"] in sel]
        meshes = [vt.generate_mesh(nx=d["cells"], ny=d["cells"]) for d in ds]
        b = vt.Batch(meshes, [d["p"] for d in ds], 8)
        b.initialize_hydrodinamics(0, [d["epsilon"] for d in ds], [d["x0"] for d in ds], [d["y0"] for d in ds])
        b.set_launch(block)
        b.rk4(dt, 4)
        solvers = []
        for m, d in zip(meshes, ds):
            s = vt.Solver(m, d["p"], 8)
            s.initialize_hydrodinamics(0, d["epsilon"], d["x0"], d["y0"])
            s.rk4(dt, 4)
            solvers.append(s)
        # three alternating runs of each, medians (the first builds the graphs)
        wbs, wps = [], []
        for _ in range(3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            b.rk4(dt, steps)
            torch.cuda.synchronize()
            wbs.append(time.perf_counter() - t0)
            t0 = time.perf_counter()
            for s in solvers:
                s.rk4(dt, steps)
            torch.cuda.synchronize()
            wps.append(time.perf_counter() - t0)
        wb, wp = float(np.median(wbs)), float(np.median(wps))
        b.close()
        for s in solvers:
            s.close()
        ne = sum(2 * d["cells"] ** 2 for d in ds)
        out[label] = {"instances": len(ds), "elements": ne, "batched_ms_per_step": wb / steps * 1e3,
                      "per_instance_ms_per_step": wp / steps * 1e3, "speedup": wp / wb,
                      "ns_per_element_stage": wb / steps / 4 / ne * 1e9}
    return out


def adjoint_figure(vt, steps, dt=1e-4, n=256, every=32):
    """Forward + discrete adjoint (checkpoints every 32 steps) of the
    mass-weighted L2 misfit to the advected target vortex (eps 0.32, centre
    (0.1, -0.05)) on n x n cells of a 10 x 10 domain, p = 3, against the plain
    forward solve of the same steps; medians of three runs of each.  The
    reverse sweep recomputes the next segment on a second stream while it
    sweeps the current one (the library default); the one-stream sweep of the
    same run is timed beside it."""
    import torch
    mesh = vt.generate_mesh(nx=n, ny=n)
    s = vt.Solver(mesh, 3, 8)
    tgt = s.exact_vortex(dt * steps, 0.32, 0.1, -0.05)[:, :, 0]
    s.initialize_hydrodinamics(0)
    U0 = s.state()
    s.rk4(dt, 64)
    # three alternating runs of each, medians reported
    tfs, tas, tss = [], [], []
    for _ in range(3):
        s.set_state(U0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.rk4(dt, steps)
        tfs.append(time.perf_counter() - t0)
        s.set_state(U0)
        t0 = time.perf_counter()
        J, lam0 = s.adjoint(dt, steps, tgt, every)
        tas.append(time.perf_counter() - t0)
        s.set_adjoint_overlap(False)
        s.set_state(U0)
        t0 = time.perf_counter()
        J1, lam1 = s.adjoint(dt, steps, tgt, every)
        tss.append(time.perf_counter() - t0)
        s.set_adjoint_overlap(True)
    serial_equal = bool(J1 == J and np.array_equal(lam1, lam0))
    # The same adjoint with every stage state of the run kept in HBM (one
    # segment of 4 x steps states, 168 GB at 256^2 x 1000 steps): no recompute.
    s.set_state(U0)
    t0 = time.perf_counter()
    Js, lam_s = s.adjoint(dt, steps, tgt, steps)
    t_stored = time.perf_counter() - t0
    stored_bytes = 4 * steps * s.ne * 10 * 4 * 8
    s.close()
    tf, ta = float(np.median(tfs)), float(np.median(tas))
    return {"workload": "%dx%d cells, 10x10 domain, p=3, integration_order 8, %d RK4 steps of dt %g, "
                        "checkpoints every %d steps, J = mass-weighted L2 misfit to the advected vortex "
                        "(eps 0.32, centre (0.1, -0.05))" % (n, n, steps, dt, every),
            "elements": 2 * n * n, "forward_s": tf, "forward_plus_adjoint_s": ta, "ratio": ta / tf,
            "runs": {"forward_s": tfs, "forward_plus_adjoint_s": tas}, "J": J,
            "one_stream_sweep": {"forward_plus_adjoint_s": float(np.median(tss)),
                                 "ratio": float(np.median(tss)) / tf, "runs": tss, "same_bits": serial_equal},
            "grad_norm": float(np.linalg.norm(lam0)),
            "all_stages_stored": {"every": steps, "stage_states_gb": stored_bytes / 1e9,
                                  "forward_plus_adjoint_s": t_stored, "ratio": t_stored / tf,
                                  "rel_l2_vs_checkpointed": float(np.linalg.norm(lam_s - lam0) / np.linalg.norm(lam0)),
                                  "J_equal": bool(Js == J)}}


def batched_adjoint_figure(vt, steps, dt=5e-5, every=16):
    """Ensemble adjoint with memory stress: 32 instances (seed 5: cells per
    side 64..512, p = 3, epsilon 0.1..0.4, centre in [-2.5, 2.5]^2 of a
    10 x 10 domain), the recorded forward run, per-instance objectives (misfit
    to the advected target vortex eps 0.32, centre (0.1, -0.05)), the reverse
    sweep and dJ_i/d(eps_i, x0_i, y0_i) (contracted on the device, dJ/dU0
    never leaves it), all instances in one launch per stage; against the
    batched forward of the same steps.  dt = 5e-5 keeps
    dt / h of the 256-cell instances of the 64-instance ensemble at h = 10/512."""
    import torch
    draws = vt.ensemble_draws(5, 32, cells=(64, 512), orders_p=(3,))
    meshes = [vt.generate_mesh(nx=d["cells"], ny=d["cells"]) for d in draws]
    eps, x0, y0 = ([d[k] for d in draws] for k in ("epsilon", "x0", "y0"))
    b = vt.Batch(meshes, 3, 8)
    c0 = time.perf_counter()
    b.initialize_hydrodinamics(0, eps, x0, y0)
    c1 = time.perf_counter()
    tg = [b.exact_vortex(i, dt * steps, 0.32, 0.1, -0.05)[:, :, 0] for i in range(b.n)]
    c2 = time.perf_counter()
    del tg
    # the run itself starts from the device set-up: initial states and the
    # objectives' target vortices computed on the device
    b.initialize_hydrodinamics(0, eps, x0, y0, device=True)
    b.set_target_vortex(dt * steps, [0.32] * b.n, [0.1] * b.n, [-0.05] * b.n)
    torch.cuda.synchronize()
    c3 = time.perf_counter()
    U0 = [b.state(i) for i in range(b.n)]
    b.rk4(dt, 4)
    b.prepare(1, dt, steps)
    for i in range(b.n):
        b.set_state(i, U0[i])
    b.time = 0.0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.rk4(dt, steps)
    tf = time.perf_counter() - t0
    # Two calls: the first allocates the recording (checkpoints and segment
    # stage states, ~200 GB), which the batch keeps for the next call; the
    # second is the steady state of repeated adjoint solves.
    tas, res = [], []
    for _ in range(2):
        for i in range(b.n):
            b.set_state(i, U0[i])
        b.time = 0.0
        t0 = time.perf_counter()
        res.append(b.adjoint_vortex(dt, steps, None, eps, x0, y0, every=every))
        tas.append(time.perf_counter() - t0)
    (J, grad), (J2, grad2) = res
    ta = tas[1]
    ne = sum(b.ne)
    slot_bytes = ne * 10 * 4 * 8
    slots = -(-steps // every) + 4 * every + 8
    b.close()
    return {"workload": "32 instances (seed 5): cells 64..512 per side, p=3, eps 0.1..0.4, 10x10 domain, "
                        "integration_order 8, %d RK4 steps of dt %g, checkpoints every %d steps" % (steps, dt, every),
            "elements": ne, "steps": steps, "forward_s": tf, "forward_ms_per_step": tf / steps * 1e3,
            "forward_plus_adjoint_s": ta, "ratio": ta / tf, "launches_per_stage": 1,
            "first_call": {"forward_plus_adjoint_s": tas[0], "ratio": tas[0] / tf,
                           "why": "allocates the recording that the second (timed) call re-uses"},
            "repeat_same_bits": bool(np.array_equal(J, J2) and np.array_equal(grad, grad2)),
            "parameter_gradient": "dJ_i/d(eps_i, x0_i, y0_i) contracted on the device",
            "device_state_gb": slots * slot_bytes / 1e9, "J_sum": float(np.sum(J)),
            "setup_s": {"initial_states_host": c1 - c0, "targets_host": c2 - c1,
                        "initial_states_and_targets_device": c3 - c2},
            "grad_finite": bool(np.isfinite(grad).all()), "grad_abs_max": float(np.abs(grad).max())}


def orders_figure(vt, mesh, dt, steps=256):
    """The main figure's mesh at p = 1, 2 and 3: ms per RK4 step and ns per
    element-stage, the per-order costs an ensemble adds up."""
    import torch
    out = {}
    for p in (1, 2, 3):
        s = vt.Solver(mesh, p, 8)
        s.initialize_hydrodinamics(0)
        s.rk4(dt, 64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        s.rk4(dt, steps)
        torch.cuda.synchronize()
        w = time.perf_counter() - t0
        out["p%d" % p] = {"ms_per_step": w / steps * 1e3, "ns_per_element_stage": w / steps / 4 / s.ne * 1e9}
        s.close()
    return out


SWEEP_CELLS = (32, 64, 96, 128, 181, 256, 362, 512, 724, 1024)


def size_sweep_figure(vt, dt=2e-5):
    """Cost of the per-stage kernels against mesh size, per order: ms per RK4
    step and ns per element-stage of single n x n-cell meshes (10 x 10 domain,
    nodal vortex) for p = 1, 2, 3, medians of three timed runs of 16·k steps
    (graphs built and 0.2 s of untimed steps first).

    From these rates, the ceiling of the ensemble's speed-up: the 64 instances
    of the batched figure (seed 3) stepped one plain solver after another cost
    sum_i 4 ne_i r_p(ne_i) per step (rate interpolated in log ne between the
    measured sizes), and the same elements each at the rate of its order on the
    largest mesh cost sum_i 4 ne_i r_p(2 * 1024^2); the quotient is the speed-up
    a batched step running every element at large-mesh efficiency shows."""
    import torch
    rates = {}
    out = {}
    for p in (1, 2, 3):
        pts = []
        for n in SWEEP_CELLS:
            mesh = vt.generate_mesh(nx=n, ny=n)
            s = vt.Solver(mesh, p, 8)
            s.initialize_hydrodinamics(0, device=True)
            steps = 16 * max(2, min(64, (1 << 20) // s.ne))
            s.prepare(1, dt, steps)
            settle_clocks(s, dt, 0.2)
            ws = []
            for _ in range(3):
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                s.rk4(dt, steps)
                torch.cuda.synchronize()
                ws.append(time.perf_counter() - t0)
            w = float(np.median(ws))
            pts.append((s.ne, w / steps / 4 / s.ne * 1e9))
            out["p%d_%d" % (p, n)] = {"elements": s.ne, "steps": steps, "ms_per_step": w / steps * 1e3,
                                      "ns_per_element_stage": w / steps / 4 / s.ne * 1e9}
            s.close()
        rates[p] = pts
    draws = vt.ensemble_draws(3, 64)

    def rate(p, ne):
        x = np.log([q[0] for q in rates[p]])
        return float(np.interp(np.log(ne), x, [q[1] for q in rates[p]]))

    per_inst = sum(4 * 2 * d["cells"] ** 2 * rate(d["p"], 2 * d["cells"] ** 2) for d in draws) * 1e-6
    at_large = sum(4 * 2 * d["cells"] ** 2 * rates[d["p"]][-1][1] for d in draws) * 1e-6
    out["ensemble_seed3_64"] = {"per_instance_ms_per_step_from_rates": per_inst,
                                "all_at_largest_mesh_rate_ms_per_step": at_large,
                                "speedup_ceiling": per_inst / at_large}
    return out


def large_mesh_figure(vt, steps, dt=5e-5, n=1024):
    """Forward-only p = 3 run on n x n cells (2 n^2 triangles) of the 20 x 20
    perturbed mesh (seed 2): device-resident state sizing and throughput."""
    import torch
    c0 = time.perf_counter()
    mesh = vt.generate_mesh_seeded(n, n, 20.0, 20.0, 0.02, 2)
    c1 = time.perf_counter()
    s = vt.Solver(mesh, 3, 8)
    c2 = time.perf_counter()
    s.initialize_hydrodinamics(0)
    c3 = time.perf_counter()
    s.initialize_hydrodinamics(0, device=True)
    torch.cuda.synchronize()
    c4 = time.perf_counter()
    s.initialize_hydrodinamics(0)
    s.rk4(dt, 4)
    s.prepare(1, dt, steps)
    settle_clocks(s, dt, 0.2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.rk4(dt, steps)
    wall = time.perf_counter() - t0
    free, total = torch.cuda.mem_get_info()
    ne = s.ne
    ok = bool(np.isfinite(s.state()).all())
    s.close()
    return {"workload": "%dx%d cells (%d triangles), 20x20 perturbed mesh (seed 2), p=3, integration_order 8, "
                        "RK4, dt %g" % (n, n, ne, dt),
            "elements": ne, "steps": steps, "ms_per_step": wall / steps * 1e3,
            "value": ne * 10 * 4 * 4 * steps / wall, "unit": "DOF-stage updates/s",
            "state_bytes": ne * 10 * 4 * 8, "device_used_gb": (total - free) / 1e9, "finite": ok,
            "setup_s": {"mesh": c1 - c0, "create": c2 - c1, "initial_state": c3 - c2,
                        "initial_state_device": c4 - c3}}


def full_run_figure(vt, steps=2000, dt=1e-4, n=512):
    """The main figure's workload run to its configured 2,000 steps (dt 1e-4,
    T = 0.2) on the resident state: time to solution, the run's DOF-stage
    rate, the L2 state error against the exactly advected vortex at T and the
    relative drift of the total mass sum_e detJ_e w^T rho_e (w the column
    sums of M_ref) over the run."""
    import torch
    mesh = vt.generate_mesh_seeded(n, n, 20.0, 20.0, 0.02, 2)
    s = vt.Solver(mesh, 3, 8)
    s.initialize_hydrodinamics(0)
    s.rk4(dt, 4)
    s.prepare(1, dt, steps)
    settle_clocks(s, dt, 0.2)
    s.time = 0.0
    s.initialize_hydrodinamics(0)
    v = mesh.element_coordinates
    detJ = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0])
    w = vt.reference_operators(3, 8)[0].sum(axis=0)
    mass = lambda U: float(np.sum(detJ[:, None] * (U[:, :, 0] * w[None, :])))
    m0 = mass(s.state())
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.rk4(dt, steps)
    wall = time.perf_counter() - t0
    U = s.state()
    err = s.l2_state_error(s.exact_vortex(s.time))
    ne = s.ne
    s.close()
    return {"workload": "%dx%d cells (%d triangles), 20x20 perturbed mesh (seed 2), p=3, integration_order 8, "
                        "%d RK4 steps of dt %g" % (n, n, ne, steps, dt),
            "steps": steps, "seconds": wall, "ms_per_step": wall / steps * 1e3,
            "value": ne * 10 * 4 * 4 * steps / wall, "unit": "DOF-stage updates/s",
            "l2_state_error_vs_exact": [float(x) for x in err],
            "mass_drift_rel": abs(mass(U) - m0) / abs(m0), "finite": bool(np.isfinite(U).all())}


def pcie_figure(vt, steps, dt=1e-4, n=512):
    """The workload stepped as the drop-in's per-step rk4() does it: the host
    state goes to the device, one RK4 step runs, and U and the nodal flux F come
    back, every step (PCIe-inclusive; not the resident figure)."""
    mesh = vt.generate_mesh_seeded(n, n, 20.0, 20.0, 0.02, 2)
    s = vt.Solver(mesh, 3, 8)
    s.initialize_hydrodinamics(0)
    U = s.state()
    s.rk4(dt, 1)
    t0 = time.perf_counter()
    for _ in range(steps):
        s.set_state(U)
        s.rk4(dt, 1)
        U = s.state()
        F = s.flux()
    wall = time.perf_counter() - t0
    ne = s.ne
    s.close()
    assert np.isfinite(F).all()
    return {"workload": "%dx%d cells, as the main figure, one RK4 step per call with the state uploaded and U, F "
                        "downloaded around it" % (n, n),
            "steps": steps, "ms_per_step": wall / steps * 1e3,
            "value": ne * 10 * 4 * 4 * steps / wall, "unit": "DOF-stage updates/s",
            "bytes_moved_per_step": ne * 10 * 4 * 8 * 2 + ne * 10 * 8 * 8}


def shard_figure(vt, steps, dt=1e-4, nx=512, ny=64):
    """One GPU's share of the 512 x 512 workload split into 8 horizontal
    strips (512 x 64 cells, 65,536 triangles) run as its own periodic mesh:
    the kernel at the shape each GPU would run; the exchange of the strips'
    boundary traces is not part of it."""
    import torch
    mesh = vt.generate_mesh_seeded(nx, ny, 20.0, 20.0 * ny / nx, 0.02, 2)
    s = vt.Solver(mesh, 3, 8)
    s.initialize_hydrodinamics(0)
    s.rk4(dt, 8)
    s.prepare(1, dt, steps)
    settle_clocks(s, dt, 0.2)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.rk4(dt, steps)
    wall = time.perf_counter() - t0
    ne = s.ne
    s.close()
    return {"workload": "%dx%d cells (%d triangles): one of 8 strips of the 512x512 mesh, p=3, integration_order 8, "
                        "RK4, dt %g" % (nx, ny, ne, dt),
            "elements": ne, "steps": steps, "ms_per_step": wall / steps * 1e3,
            "value": ne * 10 * 4 * 4 * steps / wall, "unit": "DOF-stage updates/s"}


# The reference's own recorded serial profiles of its shipped input
# (input_files/input: 8 x 8 cells, p = 3, integration_order 6, 10,000 RK4
# steps), rk4 inclusive seconds (performance_history/profile_88dbe79_*.txt:90,
# profile_4d2e0da_*.txt:90, profile_16a08ee_*.txt:10098; CPU not recorded).
REFERENCE_RECORDED_RK4_S = {"88dbe79": 32.349, "4d2e0da": 29.354, "16a08ee": 35.736}


def c1_figure(vt, with_cpu, order=8):
    """The reference's own small case run to the end: 8 x 8 cells of the
    10 x 10 domain (128 triangles), p = 3, 10,000 RK4 steps of
    dt = 1.414213562373e-3 (T = 10 sqrt 2) at the given integration order
    (8: the 8 x 8 case at integration order 8; 6: the reference's shipped input), time to
    solution on the GPU and, beside it, the reference's translation units on
    the host."""
    import torch
    prm = {"p": 3, "nx": 8, "ny": 8, "integration_order": order, "simulation_time": "14.14213562373",
           "number_time_steps": 10000}
    dt = float(prm["simulation_time"]) / prm["number_time_steps"]
    mesh = vt.generate_mesh(nx=8, ny=8)
    s = vt.Solver(mesh, 3, order)
    s.initialize_hydrodinamics(0)
    s.rk4(dt, 64)
    s.time = 0.0  # the initial state is set at the solver's time
    s.initialize_hydrodinamics(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.rk4(dt, prm["number_time_steps"])
    gpu_s = time.perf_counter() - t0
    U0 = s.exact_vortex(0.0)
    err = s.l2_state_error(U0)
    s.close()
    out = {"workload": "8x8 cells (128 triangles), 10x10 domain, p=3, integration_order %d, 10000 RK4 steps of dt %.12g"
                       % (order, dt), "gpu_s": gpu_s, "l2_state_error": [float(x) for x in err]}
    if order == 6:
        out["reference_recorded_rk4_s"] = REFERENCE_RECORDED_RK4_S
        out["speedup_vs_recorded"] = min(REFERENCE_RECORDED_RK4_S.values()) / gpu_s
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        if os.path.exists(O.REF_RUN):
            threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
            with tempfile.TemporaryDirectory() as d:
                r = O.run_reference(d, prm, prm["number_time_steps"], threads=threads, no_state=True,
                                    env_extra={"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_SCHEDULE": "static"})
            cpu_s = float(re.search(r"stepping_sec=([0-9.]+)", r["stdout"]).group(1))
            out.update({"cpu_reference_s": cpu_s, "cpu_threads": threads, "speedup": cpu_s / gpu_s})
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(mesh, p, order, steps):
    """The reference's own translation units (oracle/_ref/vt_ref_run, OpenMP,
    OMP_PROC_BIND=close, OMP_PLACES=cores, OMP_SCHEDULE=static) timed on this
    host over `steps` RK4 steps of the same mesh on all threads, and over one
    step on one thread; beside it the same sample of their -O3 x86-64-v3
    build (oracle/_ref/vt_ref_run_o3); falls back to the oracle port when the
    reference build is absent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O

    ne = mesh.num_elements
    npn = (p + 1) * (p + 2) // 2
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    omp = {"OMP_PROC_BIND": "close", "OMP_PLACES": "cores", "OMP_SCHEDULE": "static"}
    host = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    dof = ne * npn * 4 * 4

    def ref_seconds(nsteps, nthreads, binary=None):
        with tempfile.TemporaryDirectory() as d:
            r = O.run_reference(d, dict(p=p, nx=1, ny=1, integration_order=order, simulation_time="0.2",
                                        number_time_steps=2000), nsteps,
                                mesh=(mesh.element_coordinates, mesh.elements_at_boundary, mesh.element_type),
                                threads=nthreads, no_state=True, env_extra=omp, binary=binary)
        return float(re.search(r"stepping_sec=([0-9.]+)", r["stdout"]).group(1))

    if os.path.exists(O.REF_RUN):
        sec = ref_seconds(steps, threads)
        sec1 = ref_seconds(1, 1)
        out = dict({"value": dof * steps / sec, "unit": "DOF-stage updates/s", "cores": threads, "kind": "reference",
                    "sample": "%d RK4 steps of the same %d-element mesh" % (steps, ne), "seconds": round(sec, 3),
                    "build": "-O2 (the reference's profiling build), no FMA contraction",
                    "single_thread": {"value": dof / sec1, "sample": "1 RK4 step, 1 thread", "seconds": round(sec1, 3)},
                    "omp": omp}, **host)
        if os.path.exists(O.REF_RUN_O3):
            # the same translation units built for speed: the -O2 figure does
            # not understate what the reference's code does on this host
            sec3 = ref_seconds(steps, threads, O.REF_RUN_O3)
            out["o3_x86_64_v3"] = {"value": dof * steps / sec3, "seconds": round(sec3, 3),
                                   "build": "-O3 -march=x86-64-v3 (AVX2, FMA contraction)"}
        return out
    s = O.Solver(mesh.element_coordinates, mesh.elements_at_boundary, mesh.element_type, p, order)
    s.init_vortex(0)
    t0 = time.perf_counter()
    s.run(1, 1e-4, steps)
    sec = time.perf_counter() - t0
    return dict({"value": dof * steps / sec, "unit": "DOF-stage updates/s", "cores": threads, "kind": "port",
                 "sample": "%d RK4 steps of the same %d-element mesh" % (steps, ne), "seconds": round(sec, 3)}, **host)


def main():
    a = parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import __graft_entry__ as ge

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.set_device(0)
    vt = ge.load_package()
    p, order, n = 3, 8, a.cells
    mesh = vt.generate_mesh_seeded(n, n, 20.0, 20.0, 0.02, 2)
    s = vt.Solver(mesh, p, order)
    if a.block:
        s.set_launch(a.block)
    s.initialize_hydrodinamics(0)
    ne, npn = s.ne, s.np
    block = s.launch()["block"]
    if a.profile_pass:
        s.rk4(a.dt, 1)
        profile_pass(s, a.dt)
        return

    s.rk4(a.dt, a.warmup)
    s.prepare(1, a.dt, a.steps)   # the timed steps' graphs are built outside the timed region
    settle = settle_clocks(s, a.dt, a.settle_s)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    s.rk4(a.dt, a.steps)          # graph replays on the library stream; returns after it drains
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    U = s.state()
    assert np.isfinite(U).all()

    # Per-launch duration of the stage kernel: HIP events on the library's stream
    # around graph replays of kernel_steps RK4 steps (4 back-to-back stage
    # launches per step, nothing else in the graphs), divided by the launches.
    # The per-stage split comes from launches timed one by one, which also
    # count each launch's own start-up.
    s.rk4(a.dt, a.kernel_steps)
    replay_ms = s.step_timed(1, a.dt, a.kernel_steps)
    avg_launch_s = replay_ms * 1e-3 / (4 * a.kernel_steps)
    stage_ms = s.stage_kernel_times(a.dt, 5)
    bytes_per_launch = ne * BYTES_PER_ELEM_STEP / 4
    achieved = bytes_per_launch / avg_launch_s / 1e9
    traffic = traffic_from_profile(ne)
    trace_ms = trace_avg_launch_ms(ne, block)
    fp64_ops = profile_value(FP64_FILE, "fp64_ops_per_launch_avg", ne)

    dof_stage = ne * npn * 4 * 4
    line = {
        "metric": METRIC,
        "value": dof_stage * a.steps / wall,
        "unit": "DOF-stage updates/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "settle": {"steps": settle, "seconds": a.settle_s,
                   "why": "untimed steps right before the timed ones: the first ~100 stage launches after an idle "
                          "gap run up to 15% slower (clock/power transient in kernel traces)"},
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "isentropic vortex, %dx%d cells (%d triangles), 20x20 periodic domain, seeded perturbed "
                               "mesh (0.02 h, seed 2), p=3, integration_order=8, RK4" % (n, n, ne),
                   "elements": ne, "p": p, "integration_order": order, "dt": a.dt,
                   "dof_stage_per_step": dof_stage},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "vt::stage_kernel<3,5,%d> (4 launches per step)" % block,
                     "algorithmic_bytes_per_launch": bytes_per_launch,
                     "avg_launch_ms": avg_launch_s * 1e3,
                     "rocprof_avg_launch_ms": trace_ms,
                     "frac_rocprof": None if trace_ms is None else bytes_per_launch / (trace_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "stage_ms_single_launches": [float(x) for x in stage_ms]},
        "fp64": {"algorithmic": FLOPS_PER_ELEM_STAGE * ne / avg_launch_s / 1e12,
                 "executed": None if fp64_ops is None else fp64_ops / avg_launch_s / 1e12,
                 "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                 "frac": FLOPS_PER_ELEM_STAGE * ne / avg_launch_s / 1e12 / FP64_PEAK_TFLOPS,
                 "executed_ops_per_launch": fp64_ops},
    }
    s.close()
    st = stream_rates()
    st["stage_kernel_algorithmic_vs_triad"] = achieved / st["triad_GBps"]
    if traffic is not None:
        st["stage_kernel_traffic_vs_triad"] = traffic / avg_launch_s / 1e9 / st["triad_GBps"]
    line["hbm_stream"] = st
    if not a.no_batched:
        line["batched"] = batched_figure(vt, a.batched_steps)
    if a.orders:
        line["orders"] = orders_figure(vt, mesh, a.dt)
    if a.batched_by_order:
        line["batched_by_order"] = batched_by_order_figure(vt, a.batched_steps)
    if a.size_sweep:
        line["size_sweep"] = size_sweep_figure(vt)
    if not a.no_adjoint:
        line["adjoint"] = adjoint_figure(vt, a.adjoint_steps)
    if not a.no_batched_adjoint:
        line["batched_adjoint"] = batched_adjoint_figure(vt, a.batched_adjoint_steps, every=a.batched_adjoint_every)
    if not a.no_large:
        line["large_mesh"] = large_mesh_figure(vt, a.large_steps)
    if not a.no_full_run:
        line["full_run"] = full_run_figure(vt)
    if not a.no_shard:
        line["shard"] = shard_figure(vt, 200)
    if not a.no_pcie:
        line["pcie_inclusive"] = pcie_figure(vt, 5)
    if not a.no_c1:
        line["c1_full_run"] = c1_figure(vt, not a.no_cpu)
        line["shipped_input_run"] = c1_figure(vt, not a.no_cpu, order=6)
    if not a.no_cpu:
        line["cpu_baseline"] = cpu_baseline(mesh, p, order, a.cpu_steps)
    print(json.dumps(line))


if __name__ == "__main__":
    main()
