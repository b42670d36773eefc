"""The configured workloads at their own sizes, on the HIP path:

  * 512 x 512 cells of the 20 x 20 seeded perturbed mesh (524,288 elements,
    p = 3): 8 RK4 steps against the oracle;
  * 1024 x 1024 cells of the same domain (2,097,152 elements): one RK4 step
    against the oracle;
  * the 32-instance ensemble adjoint (seed 5: 64..512 cells per side, p = 3,
    6,193,560 elements in one group), cut to 16 RK4 steps of dt = 5e-5 with a
    checkpoint every 4 steps: J_i and dJ_i/dU0_i of four instances (smallest,
    largest and two more) against the single-mesh adjoint of the instance, the
    device-contracted vortex-parameter gradient against the host contraction,
    and dJ_i/d(eps_i, x0_i, y0_i) of one instance against central differences
    through the ensemble's own forward runs.

Tolerances:
  * states vs the oracle: relative L2 1e-13 (observed 2-7e-16) and every value
    within 32 ulp of the oracle's (observed a few); the departure from the free
    stream U - U_inf, relative L2 2e-12 (observed 2-6e-13);
  * the increment U - U0 vs the oracle's: relative L2 5e-9 (observed 8e-10
    after 8 steps at 512^2, 1.9e-9 after one step at 1024^2).  The GPU's
    regrouped RK4 update rounds each value differently from the reference's
    U_n + (k1 + 2 k2 + 2 k3 + k4)/6, by about one ulp of U (~1e-15 for
    rho E ~ 7.6), while the increment is ~dt |dU/dt| ~ 1e-5 inside the vortex
    and zero outside it, so the ratio is ~1e-10 sqrt(N / N_vortex) per step;
  * ensemble vs single-mesh adjoint: relative 1e-14 (same kernels, one launch);
  * device vs host parameter contraction: relative 1e-11 (summation order);
  * central differences h = 1e-6: relative 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _seeded_pair(vt, oracle, n):
    mesh = vt.generate_mesh_seeded(n, n, 20.0, 20.0, 0.02, 2)
    s = vt.Solver(mesh, 3, 8)
    o = oracle.Solver(mesh.element_coordinates, mesh.elements_at_boundary, mesh.element_type, 3, 8)
    return mesh, s, o


@pytest.mark.parametrize("n,steps", [(512, 8), (1024, 1)])
def test_seeded_mesh_steps_match_oracle(vt, oracle, gpu, n, steps):
    print("seeded %d^2 mesh, %d RK4 steps of dt 1e-4 vs the oracle" % (n, steps), flush=True)
    mesh, s, o = _seeded_pair(vt, oracle, n)
    s.initialize_hydrodinamics(0)
    U0 = s.state()
    o.init_vortex(0)
    assert np.array_equal(o.state(), U0)
    s.rk4(1e-4, steps)
    assert o.run(1, 1e-4, steps) == 0
    U, Uo = s.state(), o.state()
    r_state, r_inc = rel(U, Uo), rel(U - U0, Uo - U0)
    # the stricter diagnostic: the departure from the free stream
    Uinf = np.array([1.0, 1.0 / np.sqrt(2.0), 1.0 / np.sqrt(2.0), 20 / 7 / 0.4 + 0.5])
    r_dev = rel(U - Uinf, Uo - Uinf)
    ulps = float(np.max(np.abs(U - Uo) / np.spacing(np.abs(Uo))))
    print("state %.3e, increment %.3e, departure from free stream %.3e, max %.0f ulp"
          % (r_state, r_inc, r_dev, ulps), flush=True)
    assert r_state < 1e-13, r_state
    assert ulps <= 32, ulps
    assert r_dev < 2e-12, r_dev
    assert r_inc < 5e-9, r_inc
    s.close()


# --- the 32-instance ensemble adjoint ----------------------------------------

DT5, STEPS5, EVERY5 = 5e-5, 16, 4
TARGET = (0.32, 0.1, -0.05)


@pytest.fixture(scope="module")
def c5(vt, gpu):
    draws = vt.ensemble_draws(5, 32, cells=(64, 512), orders_p=(3,))
    meshes = [vt.generate_mesh(nx=d["cells"], ny=d["cells"]) for d in draws]
    eps, x0, y0 = ([d[k] for d in draws] for k in ("epsilon", "x0", "y0"))
    b = vt.Batch(meshes, 3, 8)
    assert b.num_launches == 1
    b.initialize_hydrodinamics(0, eps, x0, y0)
    U0 = [b.state(i) for i in range(b.n)]
    tg = [b.exact_vortex(i, DT5 * STEPS5, *TARGET)[:, :, 0] for i in range(b.n)]
    cells = [d["cells"] for d in draws]
    sample = sorted({int(np.argmin(cells)), int(np.argmax(cells)), 7, 21})
    yield draws, meshes, b, U0, tg, (eps, x0, y0), sample
    b.close()


def _reset(b, U0):
    for i in range(b.n):
        b.set_state(i, U0[i])
    b.time = 0.0


def test_c5_ensemble_adjoint_equals_single_instances(vt, c5):
    draws, meshes, b, U0, tg, _, sample = c5
    print("ensemble adjoint: %d instances, %d elements, %d steps" % (b.n, sum(b.ne), STEPS5), flush=True)
    _reset(b, U0)
    J, lam = b.adjoint(DT5, STEPS5, tg, every=EVERY5)
    assert np.all(np.isfinite(J)) and np.all(J > 0)
    for i in sample:
        print("instance %d: %d cells per side" % (i, draws[i]["cells"]), flush=True)
        s = vt.Solver(meshes[i], 3, 8)
        s.set_state(U0[i])
        Js, ls = s.adjoint(DT5, STEPS5, tg[i], every=EVERY5)
        assert rel(b.state(i), s.state()) <= 1e-14
        assert abs(J[i] - Js) <= 1e-14 * abs(Js), (i, J[i], Js)
        assert rel(lam[i], ls) <= 1e-14, (i, rel(lam[i], ls))
        s.close()


def test_c5_ensemble_parameter_gradient(vt, c5):
    draws, meshes, b, U0, tg, (eps, x0, y0), sample = c5
    _reset(b, U0)
    J, lam = b.adjoint(DT5, STEPS5, tg, every=EVERY5)
    host = b.vortex_gradient(0.0, eps, x0, y0, lam)
    del lam
    _reset(b, U0)
    J2, grad = b.adjoint_vortex(DT5, STEPS5, tg, eps, x0, y0, every=EVERY5)
    assert np.array_equal(J, J2)
    assert np.all(np.abs(grad - host) <= 1e-11 * np.abs(host).max(axis=1, keepdims=True)), np.abs(grad - host).max()
    # central differences in each parameter of one instance, through the
    # ensemble's own forward runs (the other instances unchanged)
    k, h = sample[-1], 1e-6
    for a in range(3):
        Jpm = []
        for sgn in (1.0, -1.0):
            par = [list(eps), list(x0), list(y0)]
            par[a][k] += sgn * h
            b.time = 0.0  # the vortex is set at the ensemble's time
            b.initialize_hydrodinamics(0, *par)
            b.rk4(DT5, STEPS5)
            Jpm.append(b.mass_l2_objective(tg)[0][k])
        fd = (Jpm[0] - Jpm[1]) / (2 * h)
        print("instance %d, parameter %d: adjoint %.9e, central difference %.9e" % (k, a, grad[k, a], fd), flush=True)
        assert abs(fd - grad[k, a]) < 1e-6 * abs(grad[k, a]) + 1e-12, (a, fd, grad[k, a])
