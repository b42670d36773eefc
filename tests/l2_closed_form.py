"""Closed form of the L2 state-error metric of plottools/l2stateerror.py:31-99
for a nodal difference that is a cubic polynomial, derived here from exact
rational arithmetic (not from the oracle or the library).

The script squares the nodal differences d_j first and interpolates the
squares with the p = 3 basis (l2stateerror.py:65-89); with a rule exact for
degree 10 the element integral is then

    I_e = detJ_e * sum_j (int_T phi_j) d_j^2,      err = sqrt(sum_e I_e / 25),

where int_T phi_j over the reference triangle is 1/60 at the vertex nodes,
3/80 at the edge nodes and 9/40 at the centroid.  Squaring the interpolated
difference instead gives detJ_e * d^T M d (M the exact mass matrix), which a
cubic d tells apart (a constant d cannot: both give d^2 * area).
"""
from fractions import Fraction
from math import factorial

import numpy as np

P = 3
NODES = [(Fraction(i, P), Fraction(j, P)) for j in range(P + 1) for i in range(P + 1 - j)]
MONOMIALS = [(a, b) for a in range(P + 1) for b in range(P + 1 - a)]


def _solve(A, B):
    """Exact Gauss-Jordan solve A X = B (lists of Fractions)."""
    n = len(A)
    M = [list(A[r]) + list(B[r]) for r in range(n)]
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        d = M[c][c]
        M[c] = [v / d for v in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [a - f * b for a, b in zip(M[r], M[c])]
    return [row[n:] for row in M]


def _monomial_integral(a, b):
    """int over the reference triangle of xi^a eta^b = a! b! / (a + b + 2)!."""
    return Fraction(factorial(a) * factorial(b), factorial(a + b + 2))


def basis_coefficients():
    """C[m][j]: coefficient of monomial m in the Lagrange basis function of node j."""
    V = [[x ** a * y ** b for (a, b) in MONOMIALS] for (x, y) in NODES]   # V[node][monomial]
    ident = [[Fraction(int(r == c)) for c in range(len(NODES))] for r in range(len(NODES))]
    return _solve(V, ident)  # V C = I


def basis_integrals():
    C = basis_coefficients()
    return [sum(C[m][j] * _monomial_integral(*MONOMIALS[m]) for m in range(len(MONOMIALS))) for j in range(len(NODES))]


def mass_matrix():
    C = basis_coefficients()
    n = len(NODES)
    M = [[Fraction(0)] * n for _ in range(n)]
    for i in range(n):
        for j in range(n):
            M[i][j] = sum(C[m][i] * C[q][j] * _monomial_integral(MONOMIALS[m][0] + MONOMIALS[q][0],
                                                                  MONOMIALS[m][1] + MONOMIALS[q][1])
                          for m in range(n) for q in range(n))
    return M


def physical_nodes(vertices):
    """[n][10][2] physical p = 3 nodes of elements with vertices [n][3][2]."""
    xi = np.array([float(x) for x, _ in NODES])
    eta = np.array([float(y) for _, y in NODES])
    v = np.asarray(vertices, dtype=np.float64)
    f = 1.0 - xi - eta
    return (v[:, None, 0, :] * f[None, :, None] + v[:, None, 1, :] * xi[None, :, None]
            + v[:, None, 2, :] * eta[None, :, None])


def cubic_difference(vertices):
    """A different cubic polynomial per conserved variable at the nodes: [n][10][4]."""
    x = physical_nodes(vertices)
    X, Y = x[..., 0] / 2.5, x[..., 1] / 2.5
    d = np.stack([X ** 3 - 0.5 * X * Y * Y,
                  0.7 * Y ** 3 + X * X * Y - 0.3 * X,
                  X * Y * Y - Y,
                  0.2 * X ** 3 + 0.9 * X * Y * Y - 0.6 * Y], axis=-1)
    return 1e-3 * d


def metric(vertices, d):
    """(err [4] of the script's formula, err [4] with the square taken after
    the interpolation) for the nodal difference d [n][10][4]."""
    v = np.asarray(vertices, dtype=np.float64)
    det = (v[:, 1, 0] - v[:, 0, 0]) * (v[:, 2, 1] - v[:, 0, 1]) - (v[:, 1, 1] - v[:, 0, 1]) * (v[:, 2, 0] - v[:, 0, 0])
    c = np.array([float(w) for w in basis_integrals()])
    M = np.array([[float(w) for w in row] for row in mass_matrix()])
    script = np.einsum("e,j,ejk->k", det, c, d * d)
    squared_after = np.einsum("e,ejk,ji,eik->k", det, d, M, d)
    return np.sqrt(script / 25.0), np.sqrt(squared_after / 25.0)
