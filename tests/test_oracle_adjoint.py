"""Pins of the oracle's derivatives (the CPU reference gradient).

Tangents (forward mode) and adjoints (reverse mode) are separate codes:
  * each tangent is checked against central finite differences of the
    oracle's own forward function (the Roe flux, the right-hand side, RK4
    steps, the vortex initial state);
  * each adjoint is checked against its tangent by the dot-product identity
    <w, J v> = <J^T w, v>, which fails for a dropped term, a sign or a
    transposed index in either code;
  * the objective against its closed form for a constant density offset.
"""
import numpy as np
import pytest


def _mesh(oracle, n, p=3, order=8, perturb=0.0):
    v, b, t = oracle.generate_mesh(n, n, perturb_grid=perturb)
    return oracle.Solver(v, b, t, p, order)


def _states(rng, m):
    rho = 1.0 + 0.2 * rng.random(m)
    u, v = rng.standard_normal(m), rng.standard_normal(m)
    p = 2.0 + rng.random(m)
    return np.stack([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)], 1)


def test_roe_tangent_matches_finite_differences(oracle):
    rng = np.random.default_rng(11)
    UL, UR = _states(rng, 50), _states(rng, 50)
    th = rng.random(50) * 2 * np.pi
    N = np.stack([np.cos(th), np.sin(th)], 1)
    worst = 0.0
    for i in range(50):
        dL, dR = rng.standard_normal(4), rng.standard_normal(4)
        t = oracle.roe_flux_jvp(UL[i], UR[i], N[i], dL, dR)
        h = 1e-6
        fp, _ = oracle.numerical_flux(UL[i] + h * dL, UR[i] + h * dR, N[i])
        fm, _ = oracle.numerical_flux(UL[i] - h * dL, UR[i] - h * dR, N[i])
        fd = (fp - fm) / (2 * h)
        worst = max(worst, np.abs(t - fd).max() / (np.abs(fd).max() + 1.0))
    assert worst < 1e-6, worst


def test_roe_adjoint_dot_product(oracle):
    rng = np.random.default_rng(12)
    UL, UR = _states(rng, 200), _states(rng, 200)
    # both eigenvalue branches: the vortex regime (|u.n| << c2^2) and strong flows
    UL[::3, 1:3] *= 30.0
    th = rng.random(200) * 2 * np.pi
    N = np.stack([np.cos(th), np.sin(th)], 1)
    for i in range(200):
        dL, dR, bF = rng.standard_normal(4), rng.standard_normal(4), rng.standard_normal(4)
        t = oracle.roe_flux_jvp(UL[i], UR[i], N[i], dL, dR)
        bl, br = oracle.roe_flux_vjp(UL[i], UR[i], N[i], bF)
        a, b = bF @ t, bl @ dL + br @ dR
        assert abs(a - b) <= 1e-12 * (abs(a) + np.abs(bF).max() * np.abs(t).max() + 1.0), (i, a, b)


def test_rhs_tangent_matches_finite_differences(oracle):
    s = _mesh(oracle, 6, perturb=1.0)
    s.init_vortex(0, 0.35, 0.5, -0.5)
    U = s.state()
    rng = np.random.default_rng(13)
    dU = rng.standard_normal(U.shape) * np.array([1e-3, 1e-2, 1e-2, 1e-1])
    t = s.rhs_jvp(dU)
    h = 1e-4
    s.set_state(U + h * dU); fp, _ = s.rhs()
    s.set_state(U - h * dU); fm, _ = s.rhs()
    fd = (fp - fm) / (2 * h)
    assert np.linalg.norm(t - fd) / np.linalg.norm(fd) < 1e-6


@pytest.mark.parametrize("p,order", [(0, 2), (1, 4), (2, 6), (3, 8)])
def test_rhs_adjoint_dot_product(oracle, p, order):
    s = _mesh(oracle, 5, p, order)
    s.init_vortex(0, 0.3, 0.2, 0.1)
    rng = np.random.default_rng(14 + p)
    U = s.state()
    v, w = rng.standard_normal(U.shape), rng.standard_normal(U.shape)
    a = np.vdot(w, s.rhs_jvp(v))
    b = np.vdot(s.rhs_vjp(w), v)
    assert abs(a - b) <= 1e-12 * abs(a), (a, b)


def test_rk4_tangent_matches_finite_differences(oracle):
    s = _mesh(oracle, 4)
    s.init_vortex(0)
    U0 = s.state()
    rng = np.random.default_rng(15)
    dU = rng.standard_normal(U0.shape) * 1e-2
    dt, n = 0.005, 5
    t = s.rk4_tangent(dt, n, dU)
    h = 1e-5
    s.set_state(U0 + h * dU); assert s.run(1, dt, n) == 0; up = s.state()
    s.set_state(U0 - h * dU); assert s.run(1, dt, n) == 0; um = s.state()
    fd = (up - um) / (2 * h)
    assert np.linalg.norm(t - fd) / np.linalg.norm(fd) < 1e-6


def test_rk4_adjoint_dot_product_and_forward_state(oracle):
    s = _mesh(oracle, 4, perturb=1.0)
    s.init_vortex(0)
    U0 = s.state()
    rng = np.random.default_rng(16)
    v, lamT = rng.standard_normal(U0.shape), rng.standard_normal(U0.shape)
    dt, n = 0.005, 7
    t = s.rk4_tangent(dt, n, v)
    UT = s.state()
    s.set_state(U0)
    lam0 = s.rk4_adjoint(dt, n, lamT)
    assert np.array_equal(s.state(), UT)       # the adjoint's forward sweep is the RK4 of oro_run
    r = _mesh(oracle, 4, perturb=1.0); r.set_state(U0)
    assert r.run(1, dt, n) == 0
    assert np.array_equal(r.state(), UT)
    a, b = np.vdot(lamT, t), np.vdot(lam0, v)
    assert abs(a - b) <= 1e-12 * abs(a), (a, b)


def test_vortex_tangent_matches_finite_differences(oracle):
    s = _mesh(oracle, 6)
    base = dict(epsilon=0.3, x0=0.4, y0=-0.3)
    T = s.vortex_tangent(0.7, **base)
    for i, key in enumerate(("epsilon", "x0", "y0")):
        h = 1e-6
        a, b = dict(base), dict(base)
        a[key] += h; b[key] -= h
        fd = (s.exact_vortex(0.7, **a) - s.exact_vortex(0.7, **b)) / (2 * h)
        assert np.abs(T[i] - fd).max() < 1e-7 * np.abs(fd).max(), key


def test_objective_closed_form_and_gradient(oracle):
    # Constant density offset c: J = c^2 * sum_e detJ_e 1^T M_ref 1 = c^2 * 2 * area / 2
    # = c^2 * domain area (the mass matrix integrates 1 over the reference triangle, 1/2).
    s = _mesh(oracle, 5, perturb=1.0)
    s.init_vortex(0)
    U = s.state()
    c = 0.01
    J, g = s.mass_l2_objective(U[:, :, 0] - c)
    assert abs(J - c * c * 100.0) < 1e-12
    rng = np.random.default_rng(17)
    tgt = U[:, :, 0] + 0.01 * rng.standard_normal(U.shape[:2])
    J, g = s.mass_l2_objective(tgt)
    d = rng.standard_normal(U.shape)
    h = 1e-6
    s.set_state(U + h * d); Jp, _ = s.mass_l2_objective(tgt)
    s.set_state(U - h * d); Jm, _ = s.mass_l2_objective(tgt)
    assert abs((Jp - Jm) / (2 * h) - np.vdot(g, d)) < 1e-7 * abs(np.vdot(g, d))
    assert np.all(g[:, :, 1:] == 0)
