"""A mesh whose plane sets exceed 4 GiB: 4096 x 4096 cells (33,554,432
triangles, 10.7 GB per state at p = 3).  Element offsets reach past 2^24 and
plane offsets past 2^32 bytes, so the kernels' 64-bit plane bases and 32-bit
element offsets are exercised at a size the 512 x 512 tests do not reach.

The check is exact translation symmetry.  On the unperturbed 10 x 10 mesh the
grid spacing 10/4096 = 5 * 2^-11 and all vertex coordinates are exact in
float, so shifting by half the domain in y maps every vertex exactly onto a
vertex 2048 rows up, i.e. element e onto e + 2048 * 8192.  A state periodic
with half the domain's period therefore evolves identically in the two halves,
and the second half sits entirely at offsets the first half never uses.  The
two halves must agree to the rounding of the initial state's node coordinates
(relative 1e-12)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_half_domain_translation_symmetry_beyond_4gib(vt, gpu):
    n = 4096
    mesh = vt.generate_mesh(nx=n, ny=n)
    ne = mesh.num_elements
    half = (n // 2) * (2 * n)
    v = mesh.element_coordinates.reshape(ne, 6)
    assert np.array_equal(v[half:, 1::2], v[:half, 1::2] + 5.0)  # the mesh really is shifted exactly
    s = vt.Solver(mesh, 3, 8)
    # node coordinates of the nodal p = 3 basis: v0 + xi (v1 - v0) + eta (v2 - v0)
    nodes = np.array([[i / 3, j / 3] for j in range(4) for i in range(4 - j)])
    x = v[:, None, 0] + nodes[None, :, 0] * (v[:, None, 2] - v[:, None, 0]) + nodes[None, :, 1] * (v[:, None, 4] - v[:, None, 0])
    y = v[:, None, 1] + nodes[None, :, 0] * (v[:, None, 3] - v[:, None, 1]) + nodes[None, :, 1] * (v[:, None, 5] - v[:, None, 1])
    del v
    k = 2 * np.pi / 5.0  # period 5 in x and y
    rho = 1.0 + 0.05 * np.sin(k * x) * np.cos(k * y)
    u = 0.5 + 0.05 * np.cos(k * y)
    w = -0.25 + 0.05 * np.sin(k * x)
    p = 20.0 / 7.0 * (1.0 + 0.05 * np.cos(k * (x + y)))
    del x, y
    U = np.empty((ne, 10, 4))
    U[..., 0] = rho
    U[..., 1] = rho * u
    U[..., 2] = rho * w
    U[..., 3] = p / 0.4 + 0.5 * rho * (u * u + w * w)
    del rho, u, w, p
    s.set_state(U)
    s.rk4(1e-5, 2)
    V = s.state()
    s.close()
    assert np.isfinite(V).all()
    assert np.abs(V - U).max() > 1e-6  # it moved
    d = np.abs(V[half:] - V[:half]).max()
    assert d <= 1e-12 * np.abs(V).max(), d
