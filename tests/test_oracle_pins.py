"""Pins the CPU oracle (oracle/vt_oracle.c) before it is trusted as the checker.

Each check is chosen so that a plausible slip (a wrong coefficient, sign,
index or transposed operand) fails it:
  * golden fixtures from the reference's own translation units, bit-exact;
  * closed forms: monomial exactness of the quadrature tables, exact rational
    mass and stiffness matrices, Kronecker-delta basis, flux consistency and
    antisymmetry, mesh neighbour symmetry;
  * invariants: free-stream preservation, discrete conservation, RK4 order.
"""
import glob
import math
import os
from fractions import Fraction

import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# --- quadrature (Quadraturerule.cpp) -----------------------------------------

LINE_DEGREE = {1: 1, 3: 3, 5: 5, 7: 7, 9: 9, 11: 11, 13: 13, 15: 15, 17: 17, 19: 19}
LINE_RULE_OF_ORDER = [1, 1, 3, 3, 5, 5, 7, 7, 9, 9, 11, 11, 13, 13, 15, 15, 17, 17, 19, 19]
AREA_RULE_OF_ORDER = [1, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 12, 12, 13, 14, 17, 17, 17, 19, 19]


@pytest.mark.parametrize("order", range(20))
def test_line_rule_integrates_monomials(oracle, order):
    xw = oracle.line_rule(order)
    deg = LINE_RULE_OF_ORDER[order]
    assert len(xw) == (deg + 1) // 2
    for k in range(deg + 1):
        exact = 1.0 / (k + 1)
        assert abs(np.sum(xw[:, 1] * xw[:, 0] ** k) - exact) < 2e-14, (order, k)
    # symmetric about 1/2 up to the tables' 15-decimal rounding
    assert np.allclose(np.sort(xw[:, 0]), np.sort(1 - xw[:, 0]), atol=2e-15)


@pytest.mark.parametrize("order", range(20))
def test_area_rule_integrates_monomials(oracle, order):
    q = oracle.area_rule(order)
    deg = AREA_RULE_OF_ORDER[order]
    for a in range(deg + 1):
        for b in range(deg + 1 - a):
            exact = math.factorial(a) * math.factorial(b) / math.factorial(a + b + 2)
            got = np.sum(q[:, 2] * q[:, 0] ** a * q[:, 1] ** b)
            assert abs(got - exact) < 3e-14, (order, a, b, got, exact)


def test_unsupported_orders(oracle):
    for order in (-1, 20):
        with pytest.raises(ValueError):
            oracle.line_rule(order)
        with pytest.raises(ValueError):
            oracle.area_rule(order)


# --- basis and reference operators (Lagrangebasis.cpp, Preevolve.cpp) ----------

def exact_basis_coefficients(p):
    """Monomial coefficients of the nodal Lagrange basis, derived from the node
    set alone by solving the Vandermonde system in exact rationals."""
    nodes = [(Fraction(i, p) if p else Fraction(0), Fraction(j, p) if p else Fraction(0))
             for j in range(p + 1) for i in range(p + 1 - j)]
    mons = [(a, b) for a in range(p + 1) for b in range(p + 1 - a)]
    n = len(nodes)
    A = [[x ** a * y ** b for (a, b) in mons] for (x, y) in nodes]
    # invert A exactly (Gauss-Jordan on fractions): phi_i = sum_m C[m][i] mon_m
    M = [row[:] + [Fraction(int(i == j)) for j in range(n)] for i, row in enumerate(A)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        d = M[c][c]
        M[c] = [v / d for v in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    inv = [row[n:] for row in M]
    return nodes, mons, inv  # phi_i(x) = sum_m inv[m][i] * mon_m(x)


def exact_tri(a, b):
    return Fraction(math.factorial(a) * math.factorial(b), math.factorial(a + b + 2))


def exact_operators(p):
    nodes, mons, C = exact_basis_coefficients(p)
    n = len(nodes)
    coef = [[C[m][i] for m in range(len(mons))] for i in range(n)]
    M = [[sum(coef[i][m1] * coef[j][m2] * exact_tri(mons[m1][0] + mons[m2][0], mons[m1][1] + mons[m2][1])
              for m1 in range(len(mons)) for m2 in range(len(mons))) for j in range(n)] for i in range(n)]
    S = [[[Fraction(0)] * n for _ in range(n)] for _ in range(2)]
    for i in range(n):
        for j in range(n):
            for m1, (a, b) in enumerate(mons):
                for m2, (c, d) in enumerate(mons):
                    w = coef[i][m1] * coef[j][m2]
                    if w == 0:
                        continue
                    if a > 0:
                        S[0][i][j] += w * a * exact_tri(a - 1 + c, b + d)
                    if b > 0:
                        S[1][i][j] += w * b * exact_tri(a + c, b - 1 + d)
    return M, S


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_basis_kronecker_and_partition(oracle, p):
    nodes = oracle.nodes_reference(p)
    nodes_exact, mons, C = exact_basis_coefficients(p)
    assert np.allclose(nodes, np.array(nodes_exact, dtype=float), atol=0, rtol=0) or np.allclose(nodes, np.array(nodes_exact, float), atol=1e-16)
    for i, (x, y) in enumerate(nodes):
        phi = oracle.basis(p, x, y)
        e = np.zeros(len(nodes)); e[i] = 1
        assert np.allclose(phi, e, atol=1e-14), (p, i, phi)
    rng = np.random.default_rng(1)
    for _ in range(50):
        x, y = rng.uniform(0, 1, 2)
        phi = oracle.basis(p, x, y)
        ex = [sum(float(C[m][i]) * x ** a * y ** b for m, (a, b) in enumerate(mons)) for i in range(len(nodes))]
        assert np.allclose(phi, ex, atol=1e-12)
        assert abs(phi.sum() - 1) < 1e-13
        g = oracle.basis_grad(p, x, y)
        assert np.allclose(g.sum(axis=0), 0, atol=1e-12)
        h = 1e-6
        fd = np.stack([(oracle.basis(p, x + h, y) - oracle.basis(p, x - h, y)) / (2 * h),
                       (oracle.basis(p, x, y + h) - oracle.basis(p, x, y - h)) / (2 * h)], axis=1)
        assert np.allclose(g, fd, atol=1e-7)


@pytest.mark.parametrize("p,order", [(0, 2), (1, 4), (2, 6), (3, 6), (3, 8), (3, 12)])
def test_reference_operators_exact(oracle, p, order):
    Mx, Sx = exact_operators(p)
    M = oracle.mass_matrix_ref(p, order)
    S = oracle.stiffness_matrix_ref(p, order)
    Mi = oracle.inverse_mass_matrix_ref(p, order)
    Mf = np.array(Mx, dtype=float)
    assert np.abs(M - Mf).max() < 1e-15 * 50
    assert np.abs(S - np.array(Sx, dtype=float)).max() < 1e-14
    n = len(Mx)
    assert np.abs(Mi @ M - np.eye(n)).max() < 1e-11
    # against the exact rational inverse
    import numpy.linalg as la
    Mi_exact = la.inv(Mf)
    assert np.abs(Mi - Mi_exact).max() / np.abs(Mi_exact).max() < 1e-12


# --- Roe flux (Numericalflux.cpp) -----------------------------------------------

def physical_flux(U, n):
    rho, mx, my, rE = U
    u, v = mx / rho, my / rho
    p = 0.4 * (rE - 0.5 * rho * (u * u + v * v))
    un = u * n[0] + v * n[1]
    return np.array([rho * un, mx * un + p * n[0], my * un + p * n[1], (rE + p) * un])


def test_roe_flux_matches_reference_known_answers(oracle):
    d = np.load(os.path.join(GOLD, "roe_flux_known_answers.npz"))
    r, f = d["records"], d["flux"]
    fixes = 0
    for i in range(len(r)):
        got, err = oracle.numerical_flux(r[i, :4], r[i, 4:8], r[i, 8:])
        assert err == 0
        assert np.array_equal(got, f[i]), i
        # count records where the entropy fix of an acoustic eigenvalue applies
        UL, UR, n = r[i, :4], r[i, 4:8], r[i, 8:]
        di = math.sqrt(UR[0] / UL[0])
        ui = (di * UR[1] / UR[0] + UL[1] / UL[0]) / (1 + di)
        vi = (di * UR[2] / UR[0] + UL[2] / UL[0]) / (1 + di)
        pL = 0.4 * (UL[3] - 0.5 * (UL[1] ** 2 + UL[2] ** 2) / UL[0])
        pR = 0.4 * (UR[3] - 0.5 * (UR[1] ** 2 + UR[2] ** 2) / UR[0])
        Hi = (di * (UR[3] + pR) / UR[0] + (UL[3] + pL) / UL[0]) / (1 + di)
        ci = (0.4 * (Hi - 0.5 * (ui * ui + vi * vi))) ** 2
        ucp = ui * n[0] + vi * n[1]
        fixes += abs(abs(ucp) - ci) < 0.1 * ci
    assert fixes > 20


def test_roe_flux_consistency_and_antisymmetry(oracle):
    rng = np.random.default_rng(5)
    for _ in range(300):
        rho, p = rng.uniform(0.5, 2), rng.uniform(0.5, 5)
        u, v = rng.uniform(-1, 1, 2)
        U = np.array([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
        th = rng.uniform(0, 2 * np.pi)
        n = np.array([math.cos(th), math.sin(th)])
        F, e = oracle.numerical_flux(U, U, n)
        assert e == 0
        assert np.allclose(F, physical_flux(U, n), rtol=1e-13, atol=1e-13)
        rho2, p2 = rng.uniform(0.5, 2), rng.uniform(0.5, 5)
        u2, v2 = rng.uniform(-1, 1, 2)
        U2 = np.array([rho2, rho2 * u2, rho2 * v2, p2 / 0.4 + 0.5 * rho2 * (u2 * u2 + v2 * v2)])
        Fa, _ = oracle.numerical_flux(U, U2, n)
        Fb, _ = oracle.numerical_flux(U2, U, -n)
        assert np.allclose(Fa, -Fb, rtol=1e-12, atol=1e-12)


def test_roe_flux_error_codes(oracle):
    good = np.array([1.0, 0.7, 0.7, 7.6])
    n = np.array([1.0, 0.0])
    assert oracle.numerical_flux(good, good, n)[1] == 0
    bad_rho = good * np.array([-1, 1, 1, 1])
    bad_p = np.array([1.0, 3.0, 3.0, 1.0])
    assert oracle.numerical_flux(bad_rho, good, n)[1] == 1
    assert oracle.numerical_flux(bad_p, good, n)[1] == 1
    assert oracle.numerical_flux(good, bad_p, n)[1] == 2
    assert oracle.numerical_flux(bad_p, bad_p, n)[1] == 1  # left checked first
    # NaN states pass every check (comparisons are false), as in the reference
    nan = np.full(4, np.nan)
    assert oracle.numerical_flux(nan, good, n)[1] == 0


# --- mesh (Meshgeneration.cpp) --------------------------------------------------

@pytest.mark.parametrize("nx,ny", [(1, 2), (2, 2), (3, 2), (8, 8), (5, 7), (37, 11)])
def test_mesh_connectivity_invariants(oracle, nx, ny):
    v, b, t = oracle.generate_mesh(nx, ny, 10.0, 10.0, 0.0)
    ne = 2 * nx * ny
    assert np.array_equal(t, np.arange(ne) % 2)
    # side k of e is shared with the neighbour's side k (Evolve.cpp:167-202)
    side_nb = np.stack([b[:, 2], np.where(t == 0, b[:, 0], b[:, 1]), np.where(t == 0, b[:, 1], b[:, 0])], 1)
    L = 10.0
    for e in range(ne):
        for s in range(3):
            nb = side_nb[e, s]
            assert side_nb[nb, s] == e, (e, s)
            a0, a1 = v[e, s], v[e, (s + 1) % 3]
            c0, c1 = v[nb, s], v[nb, (s + 1) % 3]
            # shared side runs the other way, up to a periodic shift
            d0, d1 = (a0 - c1), (a1 - c0)
            assert np.allclose(d0, d1, atol=1e-12)
            assert np.allclose(np.round(d0 / L) * L, d0, atol=1e-12)
    # counter-clockwise, right angle at vertex 0
    J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]], axis=2)
    assert (np.linalg.det(J) > 0).all()


def test_mesh_float_cell_size(oracle):
    v, b, t = oracle.generate_mesh(37, 5, 10.0, 10.0, 0.0)
    h = float(np.float32(10.0) / np.float32(37))  # float division, as the reference
    x0 = float(np.float32(-10.0) / np.float32(2))
    assert v[0, 0, 0] == x0
    assert v[0, 1, 0] == x0 + h * 1
    assert h != 10.0 / 37


def test_mesh_perturbation_and_nan_trap(oracle):
    v0, _, _ = oracle.generate_mesh(8, 8, 10.0, 10.0, 0.0)
    v1, _, _ = oracle.generate_mesh(8, 8, 10.0, 10.0, 1.0)
    d = np.abs(v1 - v0)
    assert d.max() > 0 and d.max() <= 0.2 * 1.25 + 1e-12
    # half-width above 6: pow(negative, non-integer) is NaN and 0 * NaN = NaN
    vn, _, _ = oracle.generate_mesh(8, 8, 20.0, 20.0, 0.0)
    assert np.isnan(vn).any()


# --- whole-path invariants ---------------------------------------------------------

def _mesh_solver(oracle, nx, p, order, perturb=0.0):
    v, b, t = oracle.generate_mesh(nx, nx, 10.0, 10.0, perturb)
    return oracle.Solver(v, b, t, p, order), v


@pytest.mark.parametrize("perturb", [0.0, 1.0])
@pytest.mark.parametrize("p", [1, 3])
def test_free_stream_preservation(oracle, p, perturb):
    s, v = _mesh_solver(oracle, 6, p, 8, perturb)
    U = np.zeros((s.ne, s.np, 4))
    U[..., 0] = 1.0; U[..., 1] = 0.7; U[..., 2] = -0.3; U[..., 3] = 20 / 7 / 0.4 + 0.5 * (0.49 + 0.09)
    s.set_state(U)
    d, rc = s.rhs()
    assert rc == 0
    assert np.abs(d).max() < 1e-11


@pytest.mark.parametrize("p", [0, 1, 2, 3])
def test_discrete_conservation(oracle, p):
    s, v = _mesh_solver(oracle, 6, p, 8, 1.0)
    s.init_vortex(0)
    d, rc = s.rhs()
    M = oracle.mass_matrix_ref(p, 8)
    J = np.stack([v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]], axis=2)
    detJ = np.linalg.det(J)
    total = np.einsum("e,ij,ejk->k", detJ, M, d)
    scale = np.einsum("e,ij,ejk->k", detJ, np.abs(M), np.abs(d))
    assert (np.abs(total) <= 1e-12 * scale + 1e-14).all(), (total, scale)


def test_rk4_fourth_order(oracle):
    # error of a short run at dt against dt/8, halving dt: ratio ~ 2^4
    def run(dt, n):
        s, _ = _mesh_solver(oracle, 4, 2, 6)
        s.init_vortex(0)
        assert s.run(1, dt, n) == 0
        return s.state()
    T = 0.02
    ref = run(T / 64, 64)
    e1 = np.abs(run(T / 4, 4) - ref).max()
    e2 = np.abs(run(T / 8, 8) - ref).max()
    assert 12 < e1 / e2 < 24, e1 / e2  # 4th order: ~16 (3rd: 8, 5th: 32)


def test_l2_metric_closed_form(oracle):
    v, b, t = oracle.generate_mesh(4, 4, 10.0, 10.0, 0.0)
    U = np.random.default_rng(0).uniform(size=(v.shape[0], 10, 4))
    assert np.array_equal(oracle.l2_state_error(v, U, U), np.zeros(4))
    delta = np.array([1e-3, 2e-3, 3e-3, 4e-3])
    # constant difference: sqrt(delta^2 * area / 25) with area 100
    got = oracle.l2_state_error(v, U, U + delta)
    assert np.allclose(got, 2 * delta, rtol=1e-12)


def test_l2_metric_cubic_closed_form(oracle):
    """A cubic nodal difference separates the script's square-then-interpolate
    formula from interpolate-then-square; the oracle must give the former."""
    import l2_closed_form as L
    from fractions import Fraction
    c = L.basis_integrals()
    assert [c[j] for j in (0, 3, 9)] == [Fraction(1, 60)] * 3 and c[5] == Fraction(9, 40)
    assert all(c[j] == Fraction(3, 80) for j in (1, 2, 4, 6, 7, 8)) and sum(c) == Fraction(1, 2)
    v, b, t = oracle.generate_mesh(6, 5, 10.0, 10.0, 1.0)
    d = L.cubic_difference(v)
    want, other = L.metric(v, d)
    # the two formulas differ by ~1e-3 relative here, 1e7 times the tolerance below
    assert np.all(np.abs(want - other) > 1e-4 * want), (want, other)
    base = np.zeros_like(d)
    got = oracle.l2_state_error(v, base, d)
    assert np.allclose(got, want, rtol=1e-13, atol=0), (got, want)


# --- against the reference's own translation units (golden fixtures) --------------

def _fixtures():
    return sorted(f for f in glob.glob(os.path.join(GOLD, "*.npz")) if "roe_flux" not in f)


@pytest.mark.parametrize("path", _fixtures(), ids=lambda p: os.path.basename(p)[:-4])
def test_oracle_bit_exact_against_reference(oracle, path):
    d = np.load(path)
    p, order, steps = int(d["p"]), int(d["integration_order"]), int(d["steps"])
    nx, ny = int(d["nx"]), int(d["ny"])
    if "full" in path:
        steps_run = steps
    else:
        steps_run = steps
    v, b, t = oracle.generate_mesh(nx, ny, 10.0, 10.0, float(d["perturb_grid"]))
    assert np.array_equal(v, d["vertices"])
    assert np.array_equal(b, d["neighbors"])
    assert np.array_equal(t, d["type"])
    s = oracle.Solver(v, b, t, p, order)
    s.init_vortex(int(d["init_type"]))
    assert np.array_equal(s.state(), d["U0"])
    if "F0" in d:
        assert np.array_equal(s.flux(), d["F0"])
        g = np.array([s.geometry(e) for e in range(s.ne)])
        assert np.array_equal(g, d["geometry"])
    assert s.run(int(d["stepping_method"]), float(d["dt"]), steps_run) == 0
    assert np.array_equal(s.state(), d["U"])
    if "F" in d:
        assert np.array_equal(s.flux(), d["F"])
    assert s.time(0) == float(d["time"])
    if "dUdt" in d:
        dd, rc = s.rhs()
        assert np.array_equal(dd, d["dUdt"])


def test_singular_mass_matrix_is_refused(oracle):
    # A one-point area rule (integration orders 0 and 1) makes M_ref rank 1 for
    # p >= 1: exactly singular, so there is no inverse to step with.
    v, b, t = oracle.generate_mesh(4, 4, 10.0, 10.0, 0.0)
    for p in (1, 2, 3):
        for order in (0, 1):
            with pytest.raises(ValueError):
                oracle.Solver(v, b, t, p, order)
    oracle.Solver(v, b, t, 0, 0)  # p = 0: M_ref is 1 x 1 and regular
