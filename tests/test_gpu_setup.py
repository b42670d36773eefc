"""The vortex set-up on the device (Element.cpp:119-326: the type 0 nodal
state, the type 1 least-squares projection, and the nodal density of an
objective's target vortex) against the host set-up, whose bits equal the
reference's (test_gpu_parity.test_initial_state_bit_exact, the golden
fixtures).

The device kernel performs the host formula's operations in the same order
with contraction off, so every step rounds as on the host, except exp and pow,
which are the device library's (pow(a, 2.0) is a * a on the device; glibc's
pow(a, 2.0) differs from a * a in about 0.1% of arguments).  Tolerances:
  * type 0: every value within 8 ulp of the host's (observed a few);
  * type 1: the nodal values come from a projection whose sums carry the
    quadrature-point differences through M_ref^-1, so relative 1e-13 of each
    component's largest magnitude (observed ~1e-15);
  * objectives with the device target: J within relative 1e-9 (J sums
    squared differences of ~1e-4 between the state and the target, whose
    ulps move J at ~1e-12) and dJ/dU0 within relative L2 1e-9.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ulps(a, b):
    return float(np.max(np.abs(a - b) / np.spacing(np.maximum(np.abs(a), np.abs(b)))))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _compare(dev, host, init_type):
    assert np.isfinite(dev).all()
    if init_type == 0:
        u = ulps(dev, host)
        same = float(np.mean(dev == host))
        print("type 0: max %.0f ulp, %.4f of the values bit-identical" % (u, same), flush=True)
        assert u <= 8, u
        assert same > 0.5, same
    else:
        scale = np.abs(host).reshape(-1, 4).max(axis=0)
        err = float(np.max(np.abs(dev - host).reshape(-1, 4) / scale))
        print("type 1: max difference %.3e of the component's scale" % err, flush=True)
        assert err <= 1e-13, err


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("init_type,order", [(0, 8), (1, 6), (1, 8), (1, 19)])
def test_device_initial_state_matches_host(vt, gpu, p, init_type, order):
    # type 1 at orders whose area rule has at least Np points (12, 16, 73):
    # with fewer, M_ref is rank-deficient and its computed inverse is round-off
    # noise, which the projection amplifies on host and device alike
    mesh = vt.generate_mesh_seeded(24, 20, 10.0, 10.0, 0.1, 7)
    s = vt.Solver(mesh, p, order)
    for eps, x0, y0, t in ((0.3, 0.0, 0.0, 0.0), (0.37, -1.2, 0.8, 0.45)):
        s.time = t
        s.initialize_hydrodinamics(init_type, eps, x0, y0)
        host = s.state()
        s.set_state(np.zeros_like(host))
        s.initialize_hydrodinamics(init_type, eps, x0, y0, device=True)
        _compare(s.state(), host, init_type)
    s.close()


def test_device_initial_state_reference_mesh(vt, gpu):
    """The reference's own mesh generator (8 x 8 cells, perturb_grid 1) and the
    default vortex, type 0 and type 1 at p = 3."""
    mesh = vt.generate_mesh(nx=8, ny=8, perturb_grid=1.0)
    s = vt.Solver(mesh, 3, 8)
    for init_type in (0, 1):
        s.initialize_hydrodinamics(init_type)
        host = s.state()
        s.initialize_hydrodinamics(init_type, device=True)
        _compare(s.state(), host, init_type)
    s.close()


def test_device_initial_state_ensemble(vt, gpu):
    draws = vt.ensemble_draws(11, 9, cells=(6, 20))
    meshes = [vt.generate_mesh(nx=d["cells"], ny=d["cells"]) for d in draws]
    eps, x0, y0 = ([d[k] for d in draws] for k in ("epsilon", "x0", "y0"))
    b = vt.Batch(meshes, [d["p"] for d in draws], 8)
    b.time = 0.3
    for init_type in (0, 1):
        b.initialize_hydrodinamics(init_type, eps, x0, y0)
        host = [b.state(i) for i in range(b.n)]
        b.initialize_hydrodinamics(init_type, eps, x0, y0, device=True)
        for i in range(b.n):
            _compare(b.state(i), host[i], init_type)
    b.close()


def test_device_target_objective_and_adjoint(vt, gpu):
    dt, n = 1e-3, 12
    mesh = vt.generate_mesh(nx=10, ny=10, perturb_grid=1.0)
    s = vt.Solver(mesh, 3, 8)
    T, tp = dt * n, (0.32, 0.1, -0.05)
    host_target = s.exact_vortex(T, *tp)[:, :, 0]
    s.initialize_hydrodinamics(0)
    U0 = s.state()
    J1, g1 = s.mass_l2_objective(host_target)
    with pytest.raises(vt.VortexError):
        s.mass_l2_objective(None)  # no target set yet
    s.set_target_vortex(T, *tp)
    J2, g2 = s.mass_l2_objective(None)
    print("objective: host target %.15e, device target %.15e" % (J1, J2), flush=True)
    assert abs(J2 - J1) <= 1e-9 * J1 and rel(g2, g1) <= 1e-9
    Ja, ga = s.adjoint(dt, n, host_target, every=4)
    s.set_state(U0)
    s.time = 0.0
    Jb, gb = s.adjoint(dt, n, None, every=4)
    print("adjoint: J %.3e relative, dJ/dU0 %.3e relative L2" % (abs(Jb - Ja) / Ja, rel(gb, ga)), flush=True)
    assert abs(Jb - Ja) <= 1e-9 * Ja and rel(gb, ga) <= 1e-9
    s.close()


def test_device_targets_ensemble_adjoint(vt, gpu):
    dt, n = 5e-4, 10
    draws = vt.ensemble_draws(5, 6, cells=(8, 16), orders_p=(3,))
    meshes = [vt.generate_mesh(nx=d["cells"], ny=d["cells"]) for d in draws]
    eps, x0, y0 = ([d[k] for d in draws] for k in ("epsilon", "x0", "y0"))
    b = vt.Batch(meshes, 3, 8)
    b.initialize_hydrodinamics(0, eps, x0, y0, device=True)
    U0 = [b.state(i) for i in range(b.n)]
    T, tp = dt * n, (0.32, 0.1, -0.05)
    tg = [b.exact_vortex(i, T, *tp)[:, :, 0] for i in range(b.n)]
    J1, g1 = b.adjoint_vortex(dt, n, tg, eps, x0, y0, every=4)
    for i in range(b.n):
        b.set_state(i, U0[i])
    b.time = 0.0
    b.set_target_vortex(T, [tp[0]] * b.n, [tp[1]] * b.n, [tp[2]] * b.n)
    J2, g2 = b.adjoint_vortex(dt, n, None, eps, x0, y0, every=4)
    print("ensemble: J %.3e, gradient %.3e relative" % (np.max(np.abs(J2 - J1) / J1), rel(g2, g1)), flush=True)
    assert np.all(np.abs(J2 - J1) <= 1e-9 * J1)
    assert rel(g2, g1) <= 1e-9
    b.close()
