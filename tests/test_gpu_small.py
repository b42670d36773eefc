"""The small-mesh path of vt_step (one workgroup, stage state in LDS, each
shared side's flux once per Gauss point, the reference's RK update order)
against the per-stage kernels and the oracle.

Tolerances: states relative L2 <= 1e-13 against the per-stage kernels and
<= 1e-12 against the oracle (both differ from it at the ulp level: the
per-stage kernels in the regrouped RK update, the small path in evaluating
each flux point in one orientation); first failures equal to the oracle's.
The reference's own 8 x 8 case (10,000 steps) runs through this path in
test_gpu_parity.test_run_matches_reference[small].
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_selection(vt, gpu):
    small = vt.Solver(vt.generate_mesh(nx=8, ny=8), 3, 8)
    assert small.small_path
    big = vt.Solver(vt.generate_mesh(nx=9, ny=8), 3, 8)  # 144 elements
    assert not big.small_path
    with pytest.raises(vt.VortexError):
        big.small_path = True
    small.small_path = False
    assert not small.small_path
    small.small_path = True
    assert small.small_path
    small.close()
    big.close()


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("order", [2, 5, 8, 13, 19])
@pytest.mark.parametrize("nx,ny", [(8, 8), (5, 3)])
def test_small_equals_stage_kernels_and_oracle(vt, oracle, gpu, p, order, nx, ny):
    mesh = vt.generate_mesh(nx=nx, ny=ny, perturb_grid=1.0)
    a, b = vt.Solver(mesh, p, order), vt.Solver(mesh, p, order)
    if not a.small_path:
        # p = 3 with 10 Gauss points per side needs 163,848 bytes of LDS at 128 elements
        assert (p, order, nx * ny) == (3, 19, 64)
        pytest.skip("stage state and fluxes exceed one workgroup's LDS")
    b.small_path = False
    o = oracle.Solver(mesh.element_coordinates, mesh.elements_at_boundary, mesh.element_type, p, order)
    for s in (a, b):
        s.initialize_hydrodinamics(0, 0.3, 0.4, -0.3)
    o.set_state(a.state())
    dt = 1e-3
    if o.run(1, dt, 70) != 0:
        # under-integrated pairs (a rank-deficient M_ref) fail as the reference does
        for s in (a, b):
            with pytest.raises(vt.NonPhysicalStateError) as e:
                s.rk4(dt, 70)
            assert tuple(e.value.info) == o.last_error()
        return
    a.rk4(dt, 70)              # 64 + 6 steps: two launches
    b.rk4(dt, 70)
    ra, rb = rel(a.state(), b.state()), rel(a.state(), o.state())
    print("p %d order %d %dx%d: small vs stage %.2e, vs oracle %.2e" % (p, order, nx, ny, ra, rb), flush=True)
    assert ra <= 1e-13 and rb <= 1e-12
    assert a.time == b.time
    # forward Euler, an odd step count
    a.forward_euler(dt, 33)
    b.forward_euler(dt, 33)
    assert o.run(0, dt, 33) == 0
    assert rel(a.state(), b.state()) <= 1e-13 and rel(a.state(), o.state()) <= 1e-12
    a.close()
    b.close()


@pytest.mark.parametrize("kernels", ["small", "stage"])
def test_first_failure_like_the_oracle(vt, oracle, gpu, kernels):
    """A negative density planted in element 37 before an RK4 run: the run
    stops at the oracle's first failure (element, point, side, code, step,
    stage) on either path."""
    mesh = vt.generate_mesh(nx=8, ny=8)
    s = vt.Solver(mesh, 3, 8)
    s.small_path = kernels == "small"
    s.initialize_hydrodinamics(0)
    U = s.state()
    U[37, 4, 0] = -0.5
    s.set_state(U)
    o = oracle.Solver(mesh.element_coordinates, mesh.elements_at_boundary, mesh.element_type, 3, 8)
    o.set_state(U)
    assert o.run(1, 1e-3, 5) != 0
    with pytest.raises(vt.NonPhysicalStateError) as e:
        s.rk4(1e-3, 5)
    print(kernels, tuple(e.value.info), o.last_error(), flush=True)
    assert tuple(e.value.info) == o.last_error()
    s.close()


def test_chunking_does_not_change_the_run(vt, gpu):
    mesh = vt.generate_mesh(nx=6, ny=5, perturb_grid=1.0)
    a, b = vt.Solver(mesh, 3, 8), vt.Solver(mesh, 3, 8)
    for s in (a, b):
        s.initialize_hydrodinamics(0)
    a.rk4(1e-3, 150)
    for _ in range(15):
        b.rk4(1e-3, 10)
    assert np.array_equal(a.state(), b.state())
    a.close()
    b.close()
