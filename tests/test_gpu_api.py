"""The measurement entry points of the C ABI (vt_step_timed,
vt_stage_kernel_times, vt_batch_step_timed) advance the state exactly as the
plain stepping calls do: the same kernels on the same data, so the states are
compared bit for bit, and the times they report are positive."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _solver(vt, n=12, p=3):
    mesh = vt.generate_mesh(nx=n, ny=n, perturb_grid=1.0)
    s = vt.Solver(mesh, p, 8)
    s.initialize_hydrodinamics(0)
    return mesh, s


@pytest.mark.parametrize("method,steps", [(1, 7), (0, 5)])
def test_step_timed_equals_step(vt, gpu, method, steps):
    _, a = _solver(vt)
    _, b = _solver(vt)
    a.step(method, 1e-3, steps)
    ms = b.step_timed(method, 1e-3, steps)
    assert ms > 0
    assert np.array_equal(a.state(), b.state())
    assert a.time == b.time


def test_stage_kernel_times_equals_rk4(vt, gpu):
    _, a = _solver(vt)
    _, b = _solver(vt)
    a.rk4(1e-3, 3)
    ms = b.stage_kernel_times(1e-3, 3)
    assert ms.shape == (4,) and np.all(ms > 0)
    assert np.array_equal(a.state(), b.state())
    assert a.time == b.time


def test_batch_step_timed_equals_step(vt, gpu):
    meshes = [vt.generate_mesh(nx=n, ny=n) for n in (5, 9, 7)]
    ps = [1, 3, 2]
    bs = []
    for _ in range(2):
        b = vt.Batch(meshes, ps, 8)
        b.initialize_hydrodinamics(0, 0.3, 0.1, -0.2)
        bs.append(b)
    assert bs[0].num_instances == 3
    bs[0].rk4(1e-3, 6)
    assert bs[1].step_timed(1, 1e-3, 6) > 0
    for i in range(3):
        assert np.array_equal(bs[0].state(i), bs[1].state(i))
    assert bs[0].time == bs[1].time


def test_prepare_steps_builds_graphs_without_stepping(vt, gpu):
    mesh = vt.generate_mesh(nx=6, ny=6)
    a = vt.Solver(mesh, 3, 8)
    a.initialize_hydrodinamics(0)
    b = vt.Solver(mesh, 3, 8)
    b.initialize_hydrodinamics(0)
    U0, t0 = b.state(), b.time
    b.prepare(1, 1e-3, 75)
    b.prepare(0, 1e-3, 7)
    assert np.array_equal(b.state(), U0) and b.time == t0
    a.rk4(1e-3, 75)
    b.rk4(1e-3, 75)
    assert np.array_equal(a.state(), b.state())
    with pytest.raises(vt.VortexError):
        b.prepare(2, 1e-3, 4)
    a.close()
    b.close()
