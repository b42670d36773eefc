"""Ensemble adjoint (vt_batch_adjoint, vt_batch_mass_l2_objective,
vt_batch_vortex_gradient, vt_batch_adjoint_vortex, vt_batch_exact_vortex).

Every instance's objective J_i and gradient dJ_i/dU0_i must equal the
single-mesh adjoint of the same instance (vt_adjoint), which
test_gpu_adjoint.py pins to the oracle's adjoint and to central differences.
Both paths run the same per-element kernels; "equal" is relative L2 <= 1e-14
(the FMA contraction of a kernel may differ in the last bit).  The ensemble
path is also checked directly against central differences (h = 1e-6,
relative 1e-6) through its own forward runs, and the parameter gradient
against the single-mesh vortex tangent (itself pinned to the oracle's
dual-number tangent) contracted in numpy: relative 1e-12 (summation order).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


SAME = 1e-14
DT, NSTEPS, EVERY = 1e-3, 20, 8
TARGET = (0.32, 0.1, -0.05)  # epsilon, x0, y0 of the advected target vortex

# Mixed sizes, orders p and two integration orders: four (p, order) groups.
CFG = [  # cells, p, order, epsilon, x0, y0
    (6, 3, 8, 0.3, 0.0, 0.0),
    (9, 2, 8, 0.2, 1.0, -0.5),
    (5, 3, 8, 0.35, -1.5, 2.0),
    (7, 1, 6, 0.1, 0.3, 0.3),
    (12, 3, 6, 0.25, 2.0, 1.0),
    (4, 3, 8, 0.4, -2.0, -2.0),
]


def _build(vt, cfg):
    meshes = [vt.generate_mesh(nx=c[0], ny=c[0]) for c in cfg]
    b = vt.Batch(meshes, [c[1] for c in cfg], [c[2] for c in cfg])
    b.initialize_hydrodinamics(0, [c[3] for c in cfg], [c[4] for c in cfg], [c[5] for c in cfg])
    return meshes, b


def _targets(b, T):
    return [b.exact_vortex(i, T, *TARGET)[:, :, 0] for i in range(b.n)]


def test_batch_exact_vortex_equals_single(vt, gpu):
    meshes, b = _build(vt, CFG)
    for i, (m, c) in enumerate(zip(meshes, CFG)):
        s = vt.Solver(m, c[1], c[2])
        assert np.array_equal(b.exact_vortex(i, 0.7, 0.3, 0.2, -0.1), s.exact_vortex(0.7, 0.3, 0.2, -0.1))
        assert np.array_equal(b.exact_vortex(i, 0.0, c[3], c[4], c[5]), b.state(i))
        s.close()


def test_batch_objective_equals_single(vt, gpu):
    meshes, b = _build(vt, CFG)
    tg = _targets(b, 0.05)
    J, g = b.mass_l2_objective(tg)
    for i, (m, c) in enumerate(zip(meshes, CFG)):
        s = vt.Solver(m, c[1], c[2])
        s.set_state(b.state(i))
        Js, gs = s.mass_l2_objective(tg[i])
        assert abs(J[i] - Js) <= SAME * abs(Js)
        assert rel(g[i], gs) <= SAME
        s.close()


def test_batch_adjoint_equals_single_instances(vt, gpu):
    meshes, b = _build(vt, CFG)
    U0 = [b.state(i) for i in range(b.n)]
    tg = _targets(b, DT * NSTEPS)
    J, lam = b.adjoint(DT, NSTEPS, tg, every=EVERY)
    assert b.time == pytest.approx(DT * NSTEPS, rel=1e-12)
    for i, (m, c) in enumerate(zip(meshes, CFG)):
        s = vt.Solver(m, c[1], c[2])
        s.set_state(U0[i])
        Js, ls = s.adjoint(DT, NSTEPS, tg[i], every=EVERY)
        assert rel(b.state(i), s.state()) <= SAME
        assert abs(J[i] - Js) <= SAME * abs(Js), (i, J[i], Js)
        assert rel(lam[i], ls) <= SAME, (i, rel(lam[i], ls))
        s.close()
    # a second sweep over a fresh recording gives the same gradients
    for i in range(b.n):
        b.set_state(i, U0[i])
    J2, lam2 = b.adjoint(DT, NSTEPS, tg, every=EVERY)
    assert np.array_equal(J, J2) and all(np.array_equal(x, y) for x, y in zip(lam, lam2))


def test_batch_adjoint_memory_kept_resized_and_released(vt, gpu):
    # The last group's recording stays allocated for the next call: calls
    # that need more of it (a longer segment), less, or none kept at all (after
    # release_adjoint) give the gradients of a fresh batch.
    meshes, b = _build(vt, CFG)
    U0 = [b.state(i) for i in range(b.n)]
    tg = _targets(b, DT * NSTEPS)

    def run(bb, every, nsteps=NSTEPS):
        for i in range(bb.n):
            bb.set_state(i, U0[i])
        bb.time = 0.0
        return bb.adjoint(DT, nsteps, tg, every=every)

    def same(x, y):
        return np.array_equal(x[0], y[0]) and all(np.array_equal(a, c) for a, c in zip(x[1], y[1]))

    base = run(b, EVERY)
    assert same(run(b, 16), base)      # longer segments than the kept recording holds
    assert same(run(b, 4), base)       # more checkpoints, shorter segments
    b.release_adjoint()
    assert same(run(b, EVERY), base)
    b.release_adjoint()
    b.release_adjoint()                # releasing nothing is allowed
    _, fresh = _build(vt, CFG)
    assert same(run(fresh, 16), base)
    fresh.close()
    b.close()


def test_single_mesh_release_adjoint(vt, gpu):
    m = vt.generate_mesh(nx=6, ny=6)
    s = vt.Solver(m, 3, 8)
    s.initialize_hydrodinamics(0)
    U0 = s.state()
    lamT = np.random.default_rng(3).standard_normal(U0.shape)
    s.adjoint_forward(DT, NSTEPS, EVERY)
    assert s.adjoint_recording() != 0
    lam0 = s.adjoint_backward(lamT)
    s.release_adjoint()
    assert s.adjoint_recording() == 0
    with pytest.raises(vt.VortexError):
        s.adjoint_backward(lamT)
    s.set_state(U0)
    s.adjoint_forward(DT, NSTEPS, EVERY)
    assert np.array_equal(s.adjoint_backward(lamT), lam0)
    s.close()


def test_batch_parameter_gradient(vt, gpu):
    meshes, b = _build(vt, CFG)
    eps, x0, y0 = ([c[k] for c in CFG] for k in (3, 4, 5))
    tg = _targets(b, DT * NSTEPS)
    J, lam = b.adjoint(DT, NSTEPS, tg, every=EVERY)
    grad = b.vortex_gradient(0.0, eps, x0, y0, lam)
    assert grad.shape == (len(CFG), 3)
    for i, (m, c) in enumerate(zip(meshes, CFG)):
        s = vt.Solver(m, c[1], c[2])
        T = s.vortex_tangent(0.0, c[3], c[4], c[5])
        ref = np.tensordot(T, lam[i], axes=([1, 2, 3], [0, 1, 2]))
        assert np.allclose(grad[i], ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max()), (i, grad[i], ref)
        s.close()
    # central differences in every parameter of one instance, through the
    # ensemble's own forward runs (the other instances are unchanged)
    k, h = 4, 1e-6
    for a in range(3):
        Jpm = []
        for sgn in (1.0, -1.0):
            par = [list(eps), list(x0), list(y0)]
            par[a][k] += sgn * h
            b.time = 0.0  # the vortex is set at the ensemble's time
            b.initialize_hydrodinamics(0, *par)
            b.rk4(DT, NSTEPS)
            Jpm.append(b.mass_l2_objective(tg)[0][k])
        fd = (Jpm[0] - Jpm[1]) / (2 * h)
        assert abs(fd - grad[k, a]) < 1e-6 * abs(grad[k, a]) + 1e-12, (a, fd, grad[k, a])


def test_batch_parameter_gradient_on_the_device(vt, gpu):
    # vt_batch_adjoint_vortex: the same J and the same contraction as
    # adjoint() + vortex_gradient() on the host, with dJ/dU0 left on the
    # device (device exp / pow and the block-tree summation order: relative
    # 1e-11), and the state advanced to U_T as by adjoint().
    meshes, b = _build(vt, CFG)
    eps, x0, y0 = ([c[k] for c in CFG] for k in (3, 4, 5))
    tg = _targets(b, DT * NSTEPS)
    U0 = [b.state(i) for i in range(b.n)]
    J, lam = b.adjoint(DT, NSTEPS, tg, every=EVERY)
    UT = [b.state(i) for i in range(b.n)]
    ref = b.vortex_gradient(0.0, eps, x0, y0, lam)
    for i in range(b.n):
        b.set_state(i, U0[i])
    b.time = 0.0
    J2, grad = b.adjoint_vortex(DT, NSTEPS, tg, eps, x0, y0, every=EVERY)
    assert np.array_equal(J2, J)
    for i in range(b.n):
        assert np.array_equal(b.state(i), UT[i])
        assert np.allclose(grad[i], ref[i], rtol=1e-11, atol=1e-13 * np.abs(ref[i]).max()), (i, grad[i], ref[i])
    b.close()


def test_batch_adjoint_names_the_failing_instance(vt, oracle, gpu):
    # p = 3 with integration_order 4: the reference's mass matrix is singular
    # and its run exits in step 1, stage 2 (see test_gpu_batch.py).
    cfg = CFG[:3] + [(6, 3, 4, 0.4, -2.0, -2.0)] + CFG[3:]
    meshes, b = _build(vt, cfg)
    tg = _targets(b, 0.01)
    with pytest.raises(vt.NonPhysicalStateError) as e:
        b.adjoint(1e-3, 10, tg, every=4)
    assert e.value.instance == 3
    m, c = meshes[3], cfg[3]
    o = oracle.Solver(m.element_coordinates, m.elements_at_boundary, m.element_type, c[1], c[2])
    o.init_vortex(0, c[3], c[4], c[5])
    assert o.run(1, 1e-3, 10) != 0
    assert tuple(e.value.info) == o.last_error()


def test_batch_adjoint_rejects_bad_arguments(vt, gpu):
    _, b = _build(vt, CFG[:2])
    tg = _targets(b, 0.0)
    with pytest.raises(ValueError):
        b.adjoint(DT, 2, tg[:1])
    with pytest.raises(vt.VortexError):
        b.adjoint(DT, 0, tg)
    with pytest.raises(vt.VortexError):
        b.adjoint(DT, 2, tg, every=0)
