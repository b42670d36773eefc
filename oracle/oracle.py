"""ctypes binding of the CPU oracle (oracle/vt_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the parity checker or the timed CPU baseline.
The product path never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_RUN = os.path.join(HERE, "_ref", "vt_ref_run")
# The same translation units at -O3 with AVX2/FMA: a CPU timing baseline only
# (FMA contraction rounds differently from the reference's profiling build).
REF_RUN_O3 = os.path.join(HERE, "_ref", "vt_ref_run_o3")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oro_create.restype = ctypes.c_void_p
        L.oro_create.argtypes = [ctypes.c_int, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int, _dp]
        L.oro_destroy.argtypes = [ctypes.c_void_p]
        for name in ("oro_init_vortex",):
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]
        L.oro_exact_vortex.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
        L.oro_set_state.argtypes = [ctypes.c_void_p, _dp]
        L.oro_set_state_flux.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oro_get_state.argtypes = [ctypes.c_void_p, _dp]
        L.oro_get_flux.argtypes = [ctypes.c_void_p, _dp]
        L.oro_get_time.restype = ctypes.c_double
        L.oro_get_time.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.oro_get_geometry.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.oro_evolve_elem.argtypes = [ctypes.c_void_p, _dp]
        L.oro_forward_euler.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oro_rk4.argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.oro_run.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int]
        L.oro_last_error.argtypes = [ctypes.c_void_p, _ip]
        L.oro_numerical_flux.argtypes = [_dp, _dp, _dp, _dp]
        L.oro_line_rule.argtypes = [ctypes.c_int, _dp]
        L.oro_area_rule.argtypes = [ctypes.c_int, _dp]
        L.oro_nodes_reference.argtypes = [ctypes.c_int, _dp]
        L.oro_basis.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.oro_basis_grad.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.oro_mass_matrix_ref.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.oro_inverse_mass_matrix_ref.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.oro_stiffness_matrix_ref.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
        L.oro_invert.argtypes = [ctypes.c_int, _dp, _dp]
        L.oro_generate_mesh.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _dp, _ip, _ip]
        L.oro_l2_state_error.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp]
        L.oro_set_threads.argtypes = [ctypes.c_int]
        L.oro_roe_flux_jvp.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.oro_roe_flux_vjp.argtypes = [_dp, _dp, _dp, _dp, _dp, _dp]
        L.oro_rhs_jvp.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oro_rhs_vjp.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.oro_rk4_tangent.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.oro_rk4_adjoint.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int, _dp, _dp]
        L.oro_vortex_tangent.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_double, _dp]
        L.oro_mass_l2_objective.restype = ctypes.c_double
        L.oro_mass_l2_objective.argtypes = [ctypes.c_void_p, _dp, _dp]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def np_of(p):
    return (p + 1) * (p + 2) // 2


def line_rule(order):
    n = lib().oro_line_rule(order, None)
    if n < 0:
        raise ValueError("unsupported quadrature order %d" % order)
    xw = np.zeros((n, 2))
    lib().oro_line_rule(order, _d(xw))
    return xw


def area_rule(order):
    n = lib().oro_area_rule(order, None)
    if n < 0:
        raise ValueError("unsupported quadrature order %d" % order)
    q = np.zeros((n, 3))
    lib().oro_area_rule(order, _d(q))
    return q


def nodes_reference(p):
    x = np.zeros((np_of(p), 2))
    if lib().oro_nodes_reference(p, _d(x)):
        raise ValueError("unsupported p")
    return x


def basis(p, xi, eta):
    out = np.zeros(np_of(p))
    if lib().oro_basis(p, xi, eta, _d(out)):
        raise ValueError("unsupported p")
    return out


def basis_grad(p, xi, eta):
    out = np.zeros((np_of(p), 2))
    if lib().oro_basis_grad(p, xi, eta, _d(out)):
        raise ValueError("unsupported p")
    return out


def mass_matrix_ref(p, order):
    m = np.zeros((np_of(p), np_of(p)))
    if lib().oro_mass_matrix_ref(p, order, _d(m)):
        raise ValueError("bad p/order")
    return m


def inverse_mass_matrix_ref(p, order):
    m = np.zeros((np_of(p), np_of(p)))
    if lib().oro_inverse_mass_matrix_ref(p, order, _d(m)):
        raise ValueError("bad p/order")
    return m


def stiffness_matrix_ref(p, order):
    s = np.zeros((2, np_of(p), np_of(p)))
    if lib().oro_stiffness_matrix_ref(p, order, _d(s)):
        raise ValueError("bad p/order")
    return s


def generate_mesh(nx, ny, domain_x=10.0, domain_y=10.0, perturb_grid=0.0):
    ne = 2 * nx * ny
    v = np.zeros((ne, 3, 2))
    b = np.zeros((ne, 3), dtype=np.int32)
    t = np.zeros(ne, dtype=np.int32)
    rc = lib().oro_generate_mesh(nx, ny, domain_x, domain_y, perturb_grid, _d(v), _i(b), _i(t))
    if rc:
        raise ValueError("invalid mesh (nx, ny)")
    return v, b, t


def numerical_flux(UL, UR, n):
    UL = np.ascontiguousarray(UL, dtype=np.float64)
    UR = np.ascontiguousarray(UR, dtype=np.float64)
    n = np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(4)
    err = lib().oro_numerical_flux(_d(UL), _d(UR), _d(n), _d(out))
    return out, err


def roe_flux_jvp(UL, UR, n, dUL, dUR):
    """Tangent of the Roe flux (forward mode)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (UL, UR, n, dUL, dUR)]
    out = np.zeros(4)
    lib().oro_roe_flux_jvp(*[_d(x) for x in a], _d(out))
    return out


def roe_flux_vjp(UL, UR, n, bF):
    """Adjoint of the Roe flux (reverse mode): (bUL, bUR)."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (UL, UR, n, bF)]
    bl, br = np.zeros(4), np.zeros(4)
    lib().oro_roe_flux_vjp(*[_d(x) for x in a], _d(bl), _d(br))
    return bl, br


def l2_state_error(vertices, U_first, U_last):
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    a = np.ascontiguousarray(U_first, dtype=np.float64)
    b = np.ascontiguousarray(U_last, dtype=np.float64)
    err = np.zeros(4)
    lib().oro_l2_state_error(v.shape[0], _d(v), _d(a), _d(b), _d(err))
    return err


class Solver:
    """The oracle's element array: geometry, operators, U, F and time."""

    def __init__(self, vertices, bnd, etype, p, order, minv_ref=None):
        self.v = np.ascontiguousarray(vertices, dtype=np.float64)
        self.b = np.ascontiguousarray(bnd, dtype=np.int32)
        self.t = np.ascontiguousarray(etype, dtype=np.int32)
        self.ne = self.v.shape[0]
        self.p = p
        self.np = np_of(p)
        mi = None if minv_ref is None else np.ascontiguousarray(minv_ref, dtype=np.float64)
        self.h = lib().oro_create(self.ne, _d(self.v), _i(self.b), _i(self.t), p, order, None if mi is None else _d(mi))
        if not self.h:
            raise ValueError("oracle could not be created (p, order or connectivity invalid)")

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().oro_destroy(h)
            self.h = None

    def init_vortex(self, init_type=0, epsilon=0.3, x0=0.0, y0=0.0):
        rc = lib().oro_init_vortex(self.h, init_type, epsilon, x0, y0)
        if rc:
            raise ValueError("unsupported initialization type")

    def exact_vortex(self, t, epsilon=0.3, x0=0.0, y0=0.0):
        U = np.zeros((self.ne, self.np, 4))
        lib().oro_exact_vortex(self.h, t, epsilon, x0, y0, _d(U))
        return U

    def set_state(self, U, F=None):
        U = np.ascontiguousarray(U, dtype=np.float64)
        if F is None:
            lib().oro_set_state(self.h, _d(U))
        else:
            F = np.ascontiguousarray(F, dtype=np.float64)
            lib().oro_set_state_flux(self.h, _d(U), _d(F))

    def state(self):
        U = np.zeros((self.ne, self.np, 4))
        lib().oro_get_state(self.h, _d(U))
        return U

    def flux(self):
        F = np.zeros((self.ne, self.np, 2, 4))
        lib().oro_get_flux(self.h, _d(F))
        return F

    def time(self, e=0):
        return lib().oro_get_time(self.h, e)

    def geometry(self, e):
        out = np.zeros(19 + 2 * self.np)
        lib().oro_get_geometry(self.h, e, _d(out))
        return out

    def rhs(self):
        d = np.zeros((self.ne, self.np, 4))
        rc = lib().oro_evolve_elem(self.h, _d(d))
        return d, rc

    def rhs_jvp(self, dU):
        dU = np.ascontiguousarray(dU, dtype=np.float64)
        out = np.zeros_like(dU)
        lib().oro_rhs_jvp(self.h, _d(dU), _d(out))
        return out

    def rhs_vjp(self, w):
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros_like(w)
        lib().oro_rhs_vjp(self.h, _d(w), _d(out))
        return out

    def rk4_tangent(self, dt, nsteps, dU0):
        dU0 = np.ascontiguousarray(dU0, dtype=np.float64)
        out = np.zeros_like(dU0)
        lib().oro_rk4_tangent(self.h, dt, nsteps, _d(dU0), _d(out))
        return out

    def rk4_adjoint(self, dt, nsteps, lamT):
        lamT = np.ascontiguousarray(lamT, dtype=np.float64)
        out = np.zeros_like(lamT)
        assert lib().oro_rk4_adjoint(self.h, dt, nsteps, _d(lamT), _d(out)) == 0
        return out

    def vortex_tangent(self, t=0.0, epsilon=0.3, x0=0.0, y0=0.0):
        out = np.zeros((3, self.ne, self.np, 4))
        lib().oro_vortex_tangent(self.h, t, epsilon, x0, y0, _d(out))
        return out

    def mass_l2_objective(self, rho_target):
        rt = np.ascontiguousarray(rho_target, dtype=np.float64).reshape(self.ne, self.np)
        g = np.zeros((self.ne, self.np, 4))
        J = lib().oro_mass_l2_objective(self.h, _d(rt), _d(g))
        return J, g

    def run(self, method, dt, nsteps):
        return lib().oro_run(self.h, method, dt, nsteps)

    def last_error(self):
        info = np.zeros(6, dtype=np.int32)
        lib().oro_last_error(self.h, _i(info))
        return tuple(int(x) for x in info)


def write_input(path, p=3, nx=8, ny=8, domain_x="10.0", domain_y="10.0", perturb_grid=0, integration_order=8,
                simulation_time="14.14213562373", number_time_steps=10000, write_every_steps=250,
                U_initialization_type=0, stepping_method=1):
    """Writes a key=value input file in the reference's format (input_files/input)."""
    with open(path, "w") as f:
        f.write("// lagrange polinomial order\np=%d\n\n" % p)
        f.write("domain_x=%s\ndomain_y=%s\n\n" % (domain_x, domain_y))
        f.write("num_element_in_x=%d\nnum_element_in_y=%d\n\n" % (nx, ny))
        f.write("perturb_grid=%s\n\nintegration_order=%d\n\n" % (perturb_grid, integration_order))
        f.write("simulation_time=%s\n\nnumber_time_steps=%d\n\n" % (simulation_time, number_time_steps))
        f.write("write_every_steps=%d\n\nU_initialization_type=%d\n\nstepping_method=%d" %
                (write_every_steps, U_initialization_type, stepping_method))


def run_reference(workdir, input_kwargs, steps, rhs_at_end=False, mesh=None, threads=None, no_state=False,
                  env_extra=None, write_dumps=False, binary=None):
    """Runs oracle/_ref/vt_ref_run (the reference's own TUs; binary: another
    build of them, such as REF_RUN_O3) in workdir; returns a dict of arrays."""
    exe = binary or REF_RUN
    if not os.path.exists(exe):
        raise FileNotFoundError(exe)
    p = input_kwargs.get("p", 3)
    order = input_kwargs.get("integration_order", 8)
    npn = np_of(p)
    workdir = os.path.abspath(workdir)
    inp = os.path.join(workdir, "input")
    write_input(inp, **input_kwargs)
    minv = os.path.join(workdir, "minv.bin")
    stiff = os.path.join(workdir, "stiff.bin")
    inverse_mass_matrix_ref(p, order).tofile(minv)
    stiffness_matrix_ref(p, order).tofile(stiff)
    out = os.path.join(workdir, "out.bin")
    cmd = [exe, "run", inp, out, "--minv", minv, "--stiff", stiff, "--steps", str(steps)]
    if rhs_at_end:
        cmd.append("--rhs-at-end")
    if no_state:
        cmd.append("--no-state")
    if write_dumps:
        cmd.append("--write-dumps")
    if mesh is not None:
        v, b, t = mesh
        ne = v.shape[0]
        rec = np.zeros((ne, 10))
        rec[:, :6] = v.reshape(ne, 6)
        rec[:, 6:9] = b
        rec[:, 9] = t
        mp = os.path.join(workdir, "mesh.bin")
        np.concatenate([[float(ne)], rec.ravel()]).tofile(mp)
        cmd += ["--mesh", mp]
    env = dict(os.environ)
    if threads:
        env["OMP_NUM_THREADS"] = str(threads)
    if env_extra:
        env.update(env_extra)
    r = subprocess.run(cmd, cwd=workdir, capture_output=True, text=True, env=env)
    if r.returncode != 0:
        raise RuntimeError("reference run failed: %s %s" % (r.stdout[-2000:], r.stderr[-2000:]))
    if no_state:
        return {"stdout": r.stdout}
    d = np.fromfile(out)
    nx, ny = input_kwargs.get("nx", 8), input_kwargs.get("ny", 8)
    ne = mesh[0].shape[0] if mesh is not None else 2 * nx * ny
    rl = 4 + 6 + 10 + 6 + 3 + 2 * npn
    o = 0
    res = {}
    res["setup"] = d[o:o + ne * rl].reshape(ne, rl); o += ne * rl
    res["U0"] = d[o:o + ne * npn * 4].reshape(ne, npn, 4); o += ne * npn * 4
    res["F0"] = d[o:o + ne * npn * 8].reshape(ne, npn, 2, 4); o += ne * npn * 8
    res["U"] = d[o:o + ne * npn * 4].reshape(ne, npn, 4); o += ne * npn * 4
    res["F"] = d[o:o + ne * npn * 8].reshape(ne, npn, 2, 4); o += ne * npn * 8
    res["time"] = d[o]; o += 1
    if rhs_at_end:
        res["dUdt"] = d[o:o + ne * npn * 4].reshape(ne, npn, 4); o += ne * npn * 4
    res["stdout"] = r.stdout
    return res
