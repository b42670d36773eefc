"""CPU checks of the C-ABI library: it loads, exports every symbol that
include/vt_api.h declares, and its host-side set-up (mesh, reference
operators, quadrature, input parsing, argument errors) matches the oracle and
the reference.  No GPU computation happens here."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "vt_api.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(vt_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(vt):
    L = vt.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing


def test_status_strings(vt):
    L = vt.lib()
    assert L.vt_status_string(4).decode() == "Non-physical state"
    assert L.vt_status_string(0).decode() == "ok"


@pytest.mark.parametrize("nx,ny,L,perturb", [(8, 8, 10.0, 0.0), (8, 8, 10.0, 1.0), (37, 5, 10.0, 0.0),
                                            (256, 3, 10.0, 0.0), (3, 2, 7.5, 2.0)])
def test_mesh_bit_exact_with_oracle(vt, oracle, nx, ny, L, perturb):
    m = vt.generate_mesh(nx=nx, ny=ny, domain_x=L, domain_y=L, perturb_grid=perturb)
    v, b, t = oracle.generate_mesh(nx, ny, L, L, perturb)
    assert np.array_equal(m.element_coordinates, v)
    assert np.array_equal(m.elements_at_boundary, b)
    assert np.array_equal(m.element_type, t)


def test_mesh_rejects_single_row(vt):
    with pytest.raises(vt.VortexError):
        vt.generate_mesh(nx=4, ny=1)


def test_seeded_mesh(vt, oracle):
    n, L, amp = 16, 20.0, 0.02
    m = vt.generate_mesh_seeded(n, n, L, L, amp, 2)
    v, b, t = oracle.generate_mesh(n, n, 10.0, 10.0, 0.0)  # connectivity only
    assert np.array_equal(m.elements_at_boundary, b) and np.array_equal(m.element_type, t)
    base = vt.generate_mesh_seeded(n, n, L, L, 0.0, 2).element_coordinates
    d = m.element_coordinates - base
    h = L / n
    assert np.abs(d).max() <= amp * h + 1e-15 and np.abs(d).max() > 0.5 * amp * h
    # boundary vertices stay on the grid
    on_edge = np.isclose(np.abs(base), L / 2)
    assert np.array_equal(d[on_edge], np.zeros(on_edge.sum()))
    m2 = vt.generate_mesh_seeded(n, n, L, L, amp, 2)
    assert np.array_equal(m.element_coordinates, m2.element_coordinates)
    assert not np.array_equal(m.element_coordinates, vt.generate_mesh_seeded(n, n, L, L, amp, 3).element_coordinates)


@pytest.mark.parametrize("p", [0, 1, 2, 3])
@pytest.mark.parametrize("order", [2, 6, 8, 13, 19])
def test_reference_operators_bit_exact(vt, oracle, p, order):
    M, Mi, S = vt.reference_operators(p, order)
    assert np.array_equal(M, oracle.mass_matrix_ref(p, order))
    assert np.array_equal(Mi, oracle.inverse_mass_matrix_ref(p, order))
    assert np.array_equal(S, oracle.stiffness_matrix_ref(p, order))
    assert np.array_equal(vt.generate_nodes_reference_space(p), oracle.nodes_reference(p))


@pytest.mark.parametrize("order", range(20))
def test_quadrature_tables(vt, oracle, order):
    xw, xyw = vt.quadrature(order)
    assert np.array_equal(xw, oracle.line_rule(order))
    assert np.array_equal(xyw, oracle.area_rule(order))


def test_unsupported_p_and_order(vt):
    with pytest.raises(vt.VortexError) as e:
        vt.reference_operators(4, 8)
    assert e.value.status == vt.VT_ERR_UNSUPPORTED_P
    with pytest.raises(vt.VortexError) as e:
        vt.reference_operators(3, 20)
    assert e.value.status == vt.VT_ERR_QUAD


def test_create_rejects_bad_arguments_before_touching_the_device(vt):
    m = vt.generate_mesh(nx=4, ny=4)
    for p, order, status in [(4, 8, vt.VT_ERR_UNSUPPORTED_P), (-1, 8, vt.VT_ERR_UNSUPPORTED_P), (3, 25, vt.VT_ERR_QUAD)]:
        with pytest.raises(vt.VortexError) as e:
            vt.Solver(m, p, order)
        assert e.value.status == status
    bad = vt.Mesh(m.element_coordinates, m.elements_at_boundary.copy(), m.element_type)
    bad.elements_at_boundary[3, 1] = 10_000
    with pytest.raises(vt.VortexError) as e:
        vt.Solver(bad, 3, 8)
    assert e.value.status == vt.VT_ERR_ARG


def test_read_input_files_matches_reference_parsing(vt, tmp_path):
    # the shipped input file, with its trailing comments (input_files/input)
    src = ("// lagrange polinomial order\np=3\n\n// physical domain\ndomain_x=10.0 //cm\ndomain_y=10.0 //cm\n\n"
           "num_element_in_x=8\nnum_element_in_y=8\n\nperturb_grid=0.1\n\nintegration_order=6\n\n"
           "simulation_time=14.14213562373\n\nnumber_time_steps=10000\n\nwrite_every_steps=250\n\n"
           "// 0 : direct interpolation\nU_initialization_type=0\n\n// 1 : rk4\nstepping_method=1")
    f = tmp_path / "input"
    f.write_text(src)
    prm = vt.read_input_files(str(f))
    assert prm.p == 3 and prm.num_element_in_x == 8 and prm.integration_order == 6
    assert prm.domain_x == 10.0
    assert prm.perturb_grid == 0.0            # std::stoi("0.1") == 0 (Parameters.cpp:55)
    assert prm.time_step == 14.14213562373 / 10000
    assert prm.simulation_time == float(np.float32(14.14213562373))
    f.write_text(src.replace("p=3", "p = 3"))
    with pytest.raises(ValueError):          # key "p " does not match "p": stoi("") throws
        vt.read_input_files(str(f))


def test_splitmix64_known_answers(vt):
    # Published splitmix64 outputs for seed 0 (Vigna's reference generator).
    g = vt.splitmix64(0)
    assert [next(g) for _ in range(3)] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_ensemble_draws_ranges_and_determinism(vt):
    a = vt.ensemble_draws(3, 64)
    assert a == vt.ensemble_draws(3, 64)
    assert a != vt.ensemble_draws(4, 64)
    assert all(32 <= d["cells"] <= 256 and d["p"] in (1, 2, 3) for d in a)
    assert all(0.1 <= d["epsilon"] < 0.4 and -2.5 <= d["x0"] < 2.5 and -2.5 <= d["y0"] < 2.5 for d in a)
    # all three orders and a spread of sizes occur in the 64 draws
    assert {d["p"] for d in a} == {1, 2, 3}
    assert max(d["cells"] for d in a) - min(d["cells"] for d in a) > 150


def test_batch_create_rejects_bad_arguments_before_touching_the_device(vt):
    m = vt.generate_mesh(nx=4, ny=4)
    for p, order, status in [([3, 4], 8, vt.VT_ERR_UNSUPPORTED_P), ([3, 3], 25, vt.VT_ERR_QUAD)]:
        with pytest.raises(vt.VortexError) as e:
            vt.Batch([m, m], p, order)
        assert e.value.status == status
    bad = vt.Mesh(m.element_coordinates, m.elements_at_boundary.copy(), m.element_type)
    bad.elements_at_boundary[3, 1] = m.num_elements  # local indices: one past the instance
    with pytest.raises(vt.VortexError) as e:
        vt.Batch([m, bad], 3, 8)
    assert e.value.status == vt.VT_ERR_ARG
    with pytest.raises(vt.VortexError) as e:
        vt.Batch([], 3, 8)
    assert e.value.status == vt.VT_ERR_ARG


def test_batch_adjoint_entry_points_reject_null_handles(vt):
    # argument checks only: nothing reaches the GPU
    L = vt.lib()
    d = np.zeros(8)
    assert L.vt_batch_adjoint(None, 1e-3, 4, 2, vt._d(d), vt._d(d), vt._d(d)) == vt.VT_ERR_ARG
    assert L.vt_batch_mass_l2_objective(None, vt._d(d), vt._d(d), vt._d(d)) == vt.VT_ERR_ARG
    assert L.vt_batch_vortex_gradient(None, 0.0, vt._d(d), vt._d(d), vt._d(d), vt._d(d), vt._d(d)) == vt.VT_ERR_ARG
    assert L.vt_batch_adjoint_vortex(None, 1e-3, 4, 2, vt._d(d), vt._d(d), vt._d(d), vt._d(d), vt._d(d),
                                     vt._d(d)) == vt.VT_ERR_ARG
    assert L.vt_batch_exact_vortex(None, 0, 0.0, 0.3, 0.0, 0.0, vt._d(d)) == vt.VT_ERR_ARG


def test_sizes_beyond_the_offset_limit_are_refused_before_any_array_is_read(vt):
    # At most 2^28 elements per context (element byte offsets and neighbour
    # indices stay below 2^31).  The limit is checked before the arrays are
    # touched, so one-element stand-ins suffice (nothing reaches the GPU).
    L = vt.lib()
    v = np.zeros(6)
    b = np.zeros(3, dtype=np.int32)
    t = np.zeros(1, dtype=np.int32)
    h = vt.ctypes.c_void_p()
    assert L.vt_create(300_000_000, vt._d(v), vt._i(b), vt._i(t), 3, 8, vt.ctypes.byref(h)) == vt.VT_ERR_ARG
    assert not h.value
    n = np.array([1_500_000_000, 1_000_000_000], dtype=np.int32)
    pp = np.array([3, 3], dtype=np.int32)
    oo = np.array([8, 8], dtype=np.int32)
    assert L.vt_batch_create(2, vt._i(n), vt._i(pp), vt._i(oo), vt._d(v), vt._i(b), vt._i(t),
                             vt.ctypes.byref(h)) == vt.VT_ERR_ARG


def test_views_of_one_buffer_are_passed_back_without_a_copy(vt):
    flat = np.arange(4 * (3 * 10 + 2 * 6), dtype=np.float64)
    parts = [flat[0:120].reshape(3, 10, 4), flat[120:168].reshape(2, 6, 4)]
    assert vt._joined(parts, flat.size) is flat
    assert vt._joined([p.copy() for p in parts], flat.size) is None   # not views
    assert vt._joined(parts[::-1], flat.size) is None                 # out of order
    assert vt._joined(parts[:1], flat.size) is None                   # incomplete


def test_prepare_steps_rejects_bad_arguments(vt):
    L = vt.lib()
    assert L.vt_prepare_steps(None, 1, 1e-3, 10) == vt.VT_ERR_ARG
    assert L.vt_batch_prepare_steps(None, 1, 1e-3, 10) == vt.VT_ERR_ARG


def test_bench_refuses_more_than_one_gpu():
    # one GPU only: --gpus 2 (or a multi-rank launch) exits with a message
    # before touching a device
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
