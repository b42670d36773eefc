"""MI355X (gfx950) implementation of Vortex-Transport's DG right-hand side and RK4 loop.

Python host mirror of the reference's call surface over the C ABI in
include/vt_api.h (libvt_hip.so, built in-tree by build.py):

  read_input_files(path)        Parameters.cpp:12-64   -> Parameters
  generate_mesh(params)         Meshgeneration.cpp:16-182 -> Mesh
  generate_nodes_reference_space(params)  Meshgeneration.cpp:188-219
  Solver(mesh, p, order)        the Element / Evolve_element arrays of main.cpp:115-147
    .initialize_hydrodinamics() Element.cpp:119-333
    .evolve_elem()              Timestepping.cpp:18-36 (returns dU/dt)
    .rk4(dt) / .forward_euler(dt)  Timestepping.cpp:85-305 / 39-82
    .step(method, dt, nsteps)   the time loop of main.cpp:154-170
  numerical_flux(UL, UR, n)     Numericalflux.cpp:13-156 (device)
  Batch(meshes, p, order)       many independent instances, one launch per RK
                                stage (no reference counterpart)

Non-physical states raise NonPhysicalStateError where the reference prints
"N. Error: Non-physical state!" and exits (Numericalflux.cpp:51-124).  The
package never falls back to a CPU path: if the HIP library cannot be loaded,
import of the solver fails.
"""
import ctypes
import math
import os
import re
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvt_hip.so")

VT_OK, VT_ERR_ARG, VT_ERR_UNSUPPORTED_P, VT_ERR_QUAD, VT_ERR_NONPHYSICAL = 0, 1, 2, 3, 4
VT_ERR_HIP, VT_ERR_INIT_TYPE, VT_ERR_METHOD, VT_ERR_MESH = 5, 6, 7, 8

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


class VortexError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(msg)
        self.status = status


class NonPhysicalStateError(VortexError):
    """Raised where the reference exits on rho <= 0, p <= 0 or c2 <= 0.

    info = (element, quadrature point, side, code, step, stage); for an
    ensemble, `instance` names the instance the element belongs to."""

    def __init__(self, info, instance=None):
        elem, q, side, code, step, stage = info
        msg = "%d. Error: Non-physical state! (element %d, side %d, quadrature point %d, step %d, stage %d)" % (
            code, elem, side + 1, q, step, stage)
        if instance is not None:
            msg += " in instance %d" % instance
        super().__init__(VT_ERR_NONPHYSICAL, msg)
        self.info = info
        self.instance = instance


def lib():
    """Loads libvt_hip.so (fails loudly when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError("libvt_hip.so not built: run __graft_entry__.build() (or make -C vortex-transport_amd/csrc)")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        sig = {
            "vt_status_string": (ctypes.c_char_p, [ctypes.c_int]),
            "vt_last_hip_error": (ctypes.c_char_p, []),
            "vt_mesh_generate": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float, _dp, _ip, _ip]),
            "vt_mesh_generate_seeded": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, ctypes.c_double,
                                                       ctypes.c_uint64, _dp, _ip, _ip]),
            "vt_nodes_reference": (ctypes.c_int, [ctypes.c_int, _dp]),
            "vt_reference_operators": (ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
            "vt_quadrature": (ctypes.c_int, [ctypes.c_int, _ip, _dp, _ip, _dp]),
            "vt_create": (ctypes.c_int, [ctypes.c_int, _dp, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.POINTER(vp)]),
            "vt_destroy": (None, [vp]),
            "vt_num_elements": (ctypes.c_int, [vp]),
            "vt_num_nodes": (ctypes.c_int, [vp]),
            "vt_init_vortex": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
            "vt_init_vortex_device": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double]),
            "vt_set_small_path": (ctypes.c_int, [vp, ctypes.c_int]),
            "vt_get_small_path": (ctypes.c_int, [vp]),
            "vt_set_target_vortex": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                    ctypes.c_double]),
            "vt_batch_init_vortex_device": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp]),
            "vt_batch_set_target_vortex": (ctypes.c_int, [vp, ctypes.c_double, _dp, _dp, _dp]),
            "vt_exact_vortex": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
            "vt_set_state": (ctypes.c_int, [vp, _dp]),
            "vt_get_state": (ctypes.c_int, [vp, _dp]),
            "vt_get_flux": (ctypes.c_int, [vp, _dp]),
            "vt_get_time": (ctypes.c_double, [vp]),
            "vt_set_time": (ctypes.c_int, [vp, ctypes.c_double]),
            "vt_rhs": (ctypes.c_int, [vp, _dp]),
            "vt_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "vt_prepare_steps": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "vt_batch_prepare_steps": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "vt_last_error": (ctypes.c_int, [vp, _ip]),
            "vt_step_timed": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
            "vt_stage_kernel_times": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, _dp]),
            "vt_numerical_flux": (ctypes.c_int, [_dp, _dp, _ip, ctypes.c_int]),
            "vt_rhs_vjp": (ctypes.c_int, [vp, _dp, _dp]),
            "vt_adjoint_forward": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int]),
            "vt_adjoint_backward": (ctypes.c_int, [vp, _dp, _dp]),
            "vt_adjoint_recording": (ctypes.c_int, [vp, ctypes.POINTER(ctypes.c_ulonglong)]),
            "vt_set_adjoint_overlap": (ctypes.c_int, [vp, ctypes.c_int]),
            "vt_adjoint_release": (ctypes.c_int, [vp]),
            "vt_batch_adjoint_release": (ctypes.c_int, [vp]),
            "vt_vortex_tangent": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]),
            "vt_mass_l2_objective": (ctypes.c_int, [vp, _dp, _dp, _dp]),
            "vt_adjoint": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
            "vt_set_launch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "vt_l2_state_error": (ctypes.c_int, [vp, _dp, _dp]),
            "vt_get_launch": (ctypes.c_int, [vp, _ip]),
            "vt_batch_create": (ctypes.c_int, [ctypes.c_int, _ip, _ip, _ip, _dp, _ip, _ip, ctypes.POINTER(vp)]),
            "vt_batch_destroy": (None, [vp]),
            "vt_batch_num_instances": (ctypes.c_int, [vp]),
            "vt_batch_num_launches": (ctypes.c_int, [vp]),
            "vt_batch_init_vortex": (ctypes.c_int, [vp, ctypes.c_int, _dp, _dp, _dp]),
            "vt_batch_set_state": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
            "vt_batch_get_state": (ctypes.c_int, [vp, ctypes.c_int, _dp]),
            "vt_batch_get_time": (ctypes.c_double, [vp]),
            "vt_batch_set_time": (ctypes.c_int, [vp, ctypes.c_double]),
            "vt_batch_step": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int]),
            "vt_batch_step_timed": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int, _dp]),
            "vt_batch_last_error": (ctypes.c_int, [vp, _ip]),
            "vt_batch_set_launch": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "vt_batch_get_launch": (ctypes.c_int, [vp, _ip]),
            "vt_batch_autotune": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, _ip, _dp]),
            "vt_batch_exact_vortex": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                     ctypes.c_double, ctypes.c_double, _dp]),
            "vt_batch_mass_l2_objective": (ctypes.c_int, [vp, _dp, _dp, _dp]),
            "vt_batch_adjoint": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]),
            "vt_batch_vortex_gradient": (ctypes.c_int, [vp, ctypes.c_double, _dp, _dp, _dp, _dp, _dp]),
            "vt_batch_adjoint_vortex": (ctypes.c_int, [vp, ctypes.c_double, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                                       _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def _check(rc, solver=None):
    if rc == VT_OK:
        return
    if rc == VT_ERR_NONPHYSICAL and isinstance(solver, Batch):
        info = np.zeros(7, dtype=np.int32)
        lib().vt_batch_last_error(solver._h, _i(info))
        raise NonPhysicalStateError(tuple(int(x) for x in info[1:]), instance=int(info[0]))
    if rc == VT_ERR_NONPHYSICAL and solver is not None:
        info = np.zeros(6, dtype=np.int32)
        lib().vt_last_error(solver._h, _i(info))
        raise NonPhysicalStateError(tuple(int(x) for x in info))
    msg = lib().vt_status_string(rc).decode()
    if rc == VT_ERR_HIP:
        msg += ": " + lib().vt_last_hip_error().decode()
    raise VortexError(rc, msg)


def num_nodes(p):
    return (p + 1) * (p + 2) // 2


# ---------------------------------------------------------------------------
# Parameters (Parameters.H:5-19, Parameters.cpp:12-64)
# ---------------------------------------------------------------------------

_INT_RE = re.compile(r"\s*([+-]?\d+)")
_FLT_RE = re.compile(r"\s*([+-]?(?:\d+\.?\d*|\.\d+)(?:[eE][+-]?\d+)?|[+-]?(?:inf(?:inity)?|nan))", re.I)


def _stoi(s):
    m = _INT_RE.match(s)
    if not m:
        raise ValueError("stoi: no conversion for %r" % s)
    v = int(m.group(1))
    if not -2**31 <= v < 2**31:
        raise ValueError("stoi: out of range %r" % s)
    return v


def _stod(s):
    m = _FLT_RE.match(s)
    if not m:
        raise ValueError("stod: no conversion for %r" % s)
    return float(m.group(1))


def _stof(s):
    return float(np.float32(_stod(s)))


@dataclass
class Parameters:
    """struct parameters (Parameters.H:5-19); float fields hold float32 values."""
    p: int
    domain_x: float
    domain_y: float
    num_element_in_x: int
    num_element_in_y: int
    perturb_grid: float
    integration_order: int
    simulation_time: float
    number_time_steps: int
    time_step: float
    write_every_steps: int
    U_initialization_type: int
    stepping_method: int


def read_input_files(path):
    """Parses a key=value input file as Parameters.cpp:12-64 does: a line
    without '=' is ignored, values keep trailing text that stoi/stof/stod
    tolerate, perturb_grid goes through stoi, and
    time_step = stod(simulation_time) / stod(number_time_steps)."""
    kv = {}
    with open(path) as f:
        for line in f.read().split("\n"):
            pos = line.find("=")
            if pos >= 0:
                kv[line[:pos]] = line[pos + 1:]
    g = lambda k: kv.get(k, "")
    return Parameters(
        p=_stoi(g("p")), domain_x=_stof(g("domain_x")), domain_y=_stof(g("domain_y")),
        num_element_in_x=_stoi(g("num_element_in_x")), num_element_in_y=_stoi(g("num_element_in_y")),
        perturb_grid=float(_stoi(g("perturb_grid"))), integration_order=_stoi(g("integration_order")),
        simulation_time=_stof(g("simulation_time")), number_time_steps=_stoi(g("number_time_steps")),
        time_step=_stod(g("simulation_time")) / _stod(g("number_time_steps")),
        write_every_steps=_stoi(g("write_every_steps")), U_initialization_type=_stoi(g("U_initialization_type")),
        stepping_method=_stoi(g("stepping_method")))


# ---------------------------------------------------------------------------
# Mesh (Meshgeneration.H:7-17)
# ---------------------------------------------------------------------------

@dataclass
class Mesh:
    """struct mesh: vertices [n][3][2], elements_at_boundary [n][3] (right,
    left, vertical), element_type [n]."""
    element_coordinates: np.ndarray
    elements_at_boundary: np.ndarray
    element_type: np.ndarray

    @property
    def num_elements(self):
        return self.element_coordinates.shape[0]


def generate_mesh(params=None, nx=None, ny=None, domain_x=10.0, domain_y=10.0, perturb_grid=0.0):
    """Structured periodic mesh with the reference connectivity (Meshgeneration.cpp:16-182)."""
    if params is not None:
        nx, ny = params.num_element_in_x, params.num_element_in_y
        domain_x, domain_y, perturb_grid = params.domain_x, params.domain_y, params.perturb_grid
    ne = 2 * nx * ny
    v = np.zeros((ne, 3, 2))
    b = np.zeros((ne, 3), dtype=np.int32)
    t = np.zeros(ne, dtype=np.int32)
    _check(lib().vt_mesh_generate(nx, ny, domain_x, domain_y, perturb_grid, _d(v), _i(b), _i(t)))
    return Mesh(v, b, t)


def generate_mesh_seeded(nx, ny, domain_x, domain_y, amplitude, seed):
    """Reference connectivity with seeded uniform interior-vertex offsets of amplitude*h."""
    ne = 2 * nx * ny
    v = np.zeros((ne, 3, 2))
    b = np.zeros((ne, 3), dtype=np.int32)
    t = np.zeros(ne, dtype=np.int32)
    _check(lib().vt_mesh_generate_seeded(nx, ny, domain_x, domain_y, amplitude, seed, _d(v), _i(b), _i(t)))
    return Mesh(v, b, t)


def generate_nodes_reference_space(p):
    if not isinstance(p, int):
        p = p.p
    x = np.zeros((num_nodes(p), 2))
    _check(lib().vt_nodes_reference(p, _d(x)))
    return x


def reference_operators(p, order):
    """(M_ref, M_ref^-1, S_ref[2]) as Preevolve.cpp:15-144 computes them."""
    n = num_nodes(p)
    M, Mi, S = np.zeros((n, n)), np.zeros((n, n)), np.zeros((2, n, n))
    _check(lib().vt_reference_operators(p, order, _d(M), _d(Mi), _d(S)))
    return M, Mi, S


def quadrature(order):
    nl, na = ctypes.c_int(), ctypes.c_int()
    _check(lib().vt_quadrature(order, ctypes.byref(nl), None, ctypes.byref(na), None))
    xw, xyw = np.zeros((nl.value, 2)), np.zeros((na.value, 3))
    _check(lib().vt_quadrature(order, None, _d(xw), None, _d(xyw)))
    return xw, xyw


def numerical_flux(UL, UR, n):
    """Device Roe flux on arrays UL[...,4], UR[...,4], n[...,2]; returns (F, codes)."""
    UL, UR, n = np.asarray(UL, float), np.asarray(UR, float), np.asarray(n, float)
    shape = UL.shape[:-1]
    cnt = int(np.prod(shape)) if shape else 1
    rec = np.ascontiguousarray(np.concatenate([UL.reshape(cnt, 4), UR.reshape(cnt, 4), n.reshape(cnt, 2)], axis=1))
    F = np.zeros((cnt, 4))
    codes = np.zeros(cnt, dtype=np.int32)
    _check(lib().vt_numerical_flux(_d(rec), _d(F), _i(codes), cnt))
    return F.reshape(shape + (4,)), codes.reshape(shape)


# ---------------------------------------------------------------------------
# Solver (the Element / Evolve_element arrays of main.cpp:115-147)
# ---------------------------------------------------------------------------

class Solver:
    """Device-resident DG solver for one mesh.

    The state stays in HBM between calls; state()/flux() copy it out in the
    reference's [n][Np][4] / [n][Np][2][4] order."""

    def __init__(self, mesh, p, integration_order):
        self.mesh = mesh
        self.p = p
        self.order = integration_order
        self.np = num_nodes(p)
        self.ne = mesh.num_elements
        self._v = np.ascontiguousarray(mesh.element_coordinates, dtype=np.float64)
        self._b = np.ascontiguousarray(mesh.elements_at_boundary, dtype=np.int32)
        self._t = np.ascontiguousarray(mesh.element_type, dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().vt_create(self.ne, _d(self._v), _i(self._b), _i(self._t), p, integration_order, ctypes.byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.vt_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    def initialize_hydrodinamics(self, init_type=0, epsilon=0.3, x0=0.0, y0=0.0, device=False):
        """Vortex initial state (Element::initialize_hydrodinamics).  device:
        computed by a device kernel instead of on the host (the host form has
        the reference's bits; the device form differs in the last ulps where
        the device exp / pow round differently)."""
        f = lib().vt_init_vortex_device if device else lib().vt_init_vortex
        _check(f(self._h, init_type, epsilon, x0, y0), self)

    def set_target_vortex(self, t, epsilon=0.3, x0=0.0, y0=0.0):
        """Objective target = nodal density of the vortex at time t, computed
        on the device; mass_l2_objective() and adjoint() then take
        rho_target=None."""
        _check(lib().vt_set_target_vortex(self._h, t, epsilon, x0, y0))

    def exact_vortex(self, t, epsilon=0.3, x0=0.0, y0=0.0):
        U = np.zeros((self.ne, self.np, 4))
        _check(lib().vt_exact_vortex(self._h, t, epsilon, x0, y0, _d(U)))
        return U

    def set_state(self, U):
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.ne, self.np, 4)
        _check(lib().vt_set_state(self._h, _d(U)))

    def state(self):
        U = np.zeros((self.ne, self.np, 4))
        _check(lib().vt_get_state(self._h, _d(U)))
        return U

    def flux(self):
        F = np.zeros((self.ne, self.np, 2, 4))
        _check(lib().vt_get_flux(self._h, _d(F)))
        return F

    @property
    def time(self):
        return lib().vt_get_time(self._h)

    @time.setter
    def time(self, t):
        _check(lib().vt_set_time(self._h, t))

    def evolve_elem(self):
        d = np.zeros((self.ne, self.np, 4))
        _check(lib().vt_rhs(self._h, _d(d)), self)
        return d

    def step(self, method, dt, nsteps=1):
        _check(lib().vt_step(self._h, method, dt, nsteps), self)

    def rk4(self, dt, nsteps=1):
        self.step(1, dt, nsteps)

    def prepare(self, method, dt, nsteps):
        """Builds the HIP graphs step(method, dt, nsteps) replays, without
        running them."""
        _check(lib().vt_prepare_steps(self._h, method, dt, nsteps), self)

    def forward_euler(self, dt, nsteps=1):
        self.step(0, dt, nsteps)

    def step_timed(self, method, dt, nsteps):
        ms = ctypes.c_double()
        _check(lib().vt_step_timed(self._h, method, dt, nsteps, ctypes.byref(ms)), self)
        return ms.value

    def stage_kernel_times(self, dt, nsteps):
        ms = np.zeros(4)
        _check(lib().vt_stage_kernel_times(self._h, dt, nsteps, _d(ms)), self)
        return ms

    def set_launch(self, block=64, persistent=False, slots_per_xcd=0):
        _check(lib().vt_set_launch(self._h, block, int(bool(persistent)), slots_per_xcd))

    # --- discrete adjoint (no reference counterpart) ---------------------

    def rhs_vjp(self, w):
        """g = f'(U)^T w at the current state, f = dU/dt of evolve_elem."""
        w = np.ascontiguousarray(w, dtype=np.float64).reshape(self.ne, self.np, 4)
        g = np.zeros_like(w)
        _check(lib().vt_rhs_vjp(self._h, _d(w), _d(g)))
        return g

    def adjoint_forward(self, dt, nsteps, every=32):
        """RK4 steps as rk4(), recording a checkpoint every `every` steps."""
        _check(lib().vt_adjoint_forward(self._h, dt, nsteps, every), self)

    def adjoint_recording(self):
        """Id of the recording adjoint_backward sweeps (0: none); it changes
        with every adjoint_forward / adjoint call."""
        r = ctypes.c_ulonglong()
        _check(lib().vt_adjoint_recording(self._h, ctypes.byref(r)))
        return r.value

    def set_adjoint_overlap(self, on):
        """Whether the reverse sweep recomputes the next segment on a second
        stream while sweeping the current one (default on; the same results)."""
        _check(lib().vt_set_adjoint_overlap(self._h, int(bool(on))))

    def release_adjoint(self):
        """Frees the adjoint memory (checkpoints, segment stage states) the
        solver otherwise keeps for re-use until close(); the recording is
        gone, so adjoint_backward() needs a new adjoint_forward()."""
        _check(lib().vt_adjoint_release(self._h))

    def adjoint_backward(self, lamT):
        """(dU_T/dU_0)^T lamT over the last recorded forward run."""
        lamT = np.ascontiguousarray(lamT, dtype=np.float64).reshape(self.ne, self.np, 4)
        lam0 = np.zeros_like(lamT)
        _check(lib().vt_adjoint_backward(self._h, _d(lamT), _d(lam0)))
        return lam0

    def vortex_tangent(self, t=0.0, epsilon=0.3, x0=0.0, y0=0.0):
        """dU0/d(epsilon, x0, y0) of the type 0 vortex state: [3][n][Np][4]."""
        out = np.zeros((3, self.ne, self.np, 4))
        _check(lib().vt_vortex_tangent(self._h, t, epsilon, x0, y0, _d(out)))
        return out

    def l2_state_error(self, U_ref):
        """The L2 state-error metric of plottools/l2stateerror.py between the
        current state and U_ref [n][Np][4] (four components), on the device."""
        U = np.ascontiguousarray(U_ref, dtype=np.float64).reshape(self.ne, self.np, 4)
        err = np.zeros(4)
        _check(lib().vt_l2_state_error(self._h, _d(U), _d(err)))
        return err

    def _target(self, rho_target):
        if rho_target is None:
            return None  # the device target of set_target_vortex
        return _d(np.ascontiguousarray(rho_target, dtype=np.float64).reshape(self.ne, self.np))

    def mass_l2_objective(self, rho_target=None):
        """(J, dJ/dU) of J = sum_e detJ_e drho^T M_ref drho at the current
        state; rho_target None: the target of set_target_vortex()."""
        J = ctypes.c_double()
        g = np.zeros((self.ne, self.np, 4))
        _check(lib().vt_mass_l2_objective(self._h, self._target(rho_target), ctypes.byref(J), _d(g)))
        return J.value, g

    def adjoint(self, dt, nsteps, rho_target=None, every=32):
        """Forward run, objective and reverse sweep: (J, dJ/dU0); rho_target
        None: the target of set_target_vortex()."""
        J = ctypes.c_double()
        g = np.zeros((self.ne, self.np, 4))
        _check(lib().vt_adjoint(self._h, dt, nsteps, every, self._target(rho_target), ctypes.byref(J), _d(g)), self)
        return J.value, g

    def launch(self):
        c = np.zeros(3, dtype=np.int32)
        _check(lib().vt_get_launch(self._h, _i(c)))
        return {"block": int(c[0]), "persistent": bool(c[1]), "slots_per_xcd": int(c[2])}

    @property
    def small_path(self):
        """True when rk4() / forward_euler() run the small-mesh path (one
        workgroup, state in LDS, up to 64 steps per launch)."""
        return bool(lib().vt_get_small_path(self._h))

    @small_path.setter
    def small_path(self, on):
        _check(lib().vt_set_small_path(self._h, 1 if on else 0))


# ---------------------------------------------------------------------------
# Ensemble: many independent meshes in one launch per RK stage (no reference
# counterpart; each instance is what Solver builds from its own mesh)
# ---------------------------------------------------------------------------

def _launch_dict(c):
    return {"block": int(c[0]), "persistent": int(c[1]) == 1, "concurrent": int(c[1]) == 2,
            "slots_per_xcd": int(c[2])}


class Batch:
    """Device-resident ensemble of independent vortex problems.

    meshes: list of Mesh; p, integration_order: one int each or one per
    instance.  Instances share dt, step count and time."""

    def __init__(self, meshes, p, integration_order):
        n = len(meshes)
        if n < 1:
            raise VortexError(VT_ERR_ARG, "an ensemble needs at least one instance")
        ps = [p] * n if np.isscalar(p) else list(p)
        orders = [integration_order] * n if np.isscalar(integration_order) else list(integration_order)
        if len(ps) != n or len(orders) != n:
            raise ValueError("p and integration_order need one entry per instance")
        self.n = n
        self.p = [int(x) for x in ps]
        self.order = [int(x) for x in orders]
        self.ne = [m.num_elements for m in meshes]
        self.np = [num_nodes(x) for x in self.p]
        self._ne = np.array(self.ne, dtype=np.int32)
        self._p = np.array(self.p, dtype=np.int32)
        self._o = np.array(self.order, dtype=np.int32)
        v = np.ascontiguousarray(np.concatenate([m.element_coordinates.reshape(-1) for m in meshes]), dtype=np.float64)
        b = np.ascontiguousarray(np.concatenate([m.elements_at_boundary.reshape(-1) for m in meshes]), dtype=np.int32)
        t = np.ascontiguousarray(np.concatenate([m.element_type.reshape(-1) for m in meshes]), dtype=np.int32)
        h = ctypes.c_void_p()
        _check(lib().vt_batch_create(n, _i(self._ne), _i(self._p), _i(self._o), _d(v), _i(b), _i(t), ctypes.byref(h)))
        self._h = h

    def close(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.vt_batch_destroy(h)
        self._h = None

    def __del__(self):
        self.close()

    @property
    def num_launches(self):
        """Kernel launches per RK stage."""
        return lib().vt_batch_num_launches(self._h)

    @property
    def num_instances(self):
        return lib().vt_batch_num_instances(self._h)

    def _per_instance(self, *vals):
        return [np.ascontiguousarray(np.broadcast_to(np.asarray(v, dtype=np.float64), (self.n,))) for v in vals]

    def initialize_hydrodinamics(self, init_type, epsilon, x0, y0, device=False):
        """Vortex initial state of every instance; device: computed by a
        device kernel (Solver.initialize_hydrodinamics)."""
        e, a, c = self._per_instance(epsilon, x0, y0)
        f = lib().vt_batch_init_vortex_device if device else lib().vt_batch_init_vortex
        _check(f(self._h, init_type, _d(e), _d(a), _d(c)), self)

    def set_target_vortex(self, t, epsilon, x0, y0):
        """Every instance's objective target = nodal density of its vortex
        (epsilon[i], x0[i], y0[i]) at time t, computed on the device; the
        objective and adjoint calls then take rho_targets=None."""
        e, a, c = self._per_instance(epsilon, x0, y0)
        _check(lib().vt_batch_set_target_vortex(self._h, t, _d(e), _d(a), _d(c)))

    def _inst(self, i):
        if not 0 <= i < self.n:
            raise IndexError("instance %d out of range (ensemble of %d)" % (i, self.n))
        return i

    def set_state(self, i, U):
        i = self._inst(i)
        U = np.ascontiguousarray(U, dtype=np.float64).reshape(self.ne[i], self.np[i], 4)
        _check(lib().vt_batch_set_state(self._h, i, _d(U)))

    def state(self, i):
        i = self._inst(i)
        U = np.zeros((self.ne[i], self.np[i], 4))
        _check(lib().vt_batch_get_state(self._h, i, _d(U)))
        return U

    @property
    def time(self):
        return lib().vt_batch_get_time(self._h)

    @time.setter
    def time(self, t):
        _check(lib().vt_batch_set_time(self._h, t))

    def step(self, method, dt, nsteps=1):
        _check(lib().vt_batch_step(self._h, method, dt, nsteps), self)

    def rk4(self, dt, nsteps=1):
        self.step(1, dt, nsteps)

    def prepare(self, method, dt, nsteps):
        """Builds the HIP graphs step(method, dt, nsteps) replays, without
        running them."""
        _check(lib().vt_batch_prepare_steps(self._h, method, dt, nsteps), self)

    def forward_euler(self, dt, nsteps=1):
        self.step(0, dt, nsteps)

    def step_timed(self, method, dt, nsteps):
        ms = ctypes.c_double()
        _check(lib().vt_batch_step_timed(self._h, method, dt, nsteps, ctypes.byref(ms)), self)
        return ms.value

    def set_launch(self, block=64, persistent=False, slots_per_xcd=0, concurrent=False):
        """concurrent: one launch per (p, order) group and stage, the groups
        on streams of their own, instead of one launch per quadrature size."""
        mode = 2 if concurrent else int(bool(persistent))
        _check(lib().vt_batch_set_launch(self._h, block, mode, slots_per_xcd))

    def launch(self):
        c = np.zeros(3, dtype=np.int32)
        _check(lib().vt_batch_get_launch(self._h, _i(c)))
        return _launch_dict(c)

    def autotune(self, dt, trial_steps=4):
        """Picks the fastest launch shape over trial RK4 steps; state and
        time are left unchanged.  Returns (launch dict, ms per step)."""
        c = np.zeros(3, dtype=np.int32)
        ms = ctypes.c_double()
        _check(lib().vt_batch_autotune(self._h, dt, trial_steps, _i(c), ctypes.byref(ms)), self)
        return _launch_dict(c), ms.value

    # --- ensemble adjoint ---------------------------------------------------

    def exact_vortex(self, i, t, epsilon=0.3, x0=0.0, y0=0.0):
        """Type 0 vortex state [n_i][Np_i][4] of instance i at time t."""
        i = self._inst(i)
        U = np.zeros((self.ne[i], self.np[i], 4))
        _check(lib().vt_batch_exact_vortex(self._h, i, t, epsilon, x0, y0, _d(U)))
        return U

    def _nodes(self):
        return sum(self.ne[i] * self.np[i] for i in range(self.n))

    def _targets(self, rho_targets):
        if rho_targets is None:
            return None  # the device targets of set_target_vortex
        if len(rho_targets) != self.n:
            raise ValueError("one target density per instance")
        return np.ascontiguousarray(np.concatenate(
            [np.asarray(r, dtype=np.float64).reshape(self.ne[i] * self.np[i]) for i, r in enumerate(rho_targets)]))

    def _split(self, flat):
        out, o = [], 0
        for i in range(self.n):
            k = self.ne[i] * self.np[i] * 4
            out.append(flat[o:o + k].reshape(self.ne[i], self.np[i], 4))
            o += k
        return out

    def mass_l2_objective(self, rho_targets=None):
        """(J [n], [dJ_i/dU_i]) at the current states; rho_targets: one
        [n_i][Np_i] density per instance, or None for the targets of
        set_target_vortex()."""
        rt = self._targets(rho_targets)
        J = np.zeros(self.n)
        g = np.empty(4 * self._nodes())  # every entry is written
        _check(lib().vt_batch_mass_l2_objective(self._h, None if rt is None else _d(rt), _d(J), _d(g)), self)
        return J, self._split(g)

    def adjoint(self, dt, nsteps, rho_targets=None, every=32):
        """Recorded forward run of every instance, objectives at the end and
        the reverse sweep: (J [n], [dJ_i/dU0_i])."""
        rt = self._targets(rho_targets)
        J = np.zeros(self.n)
        g = np.empty(4 * self._nodes())  # every entry is written
        _check(lib().vt_batch_adjoint(self._h, dt, nsteps, every, None if rt is None else _d(rt), _d(J), _d(g)), self)
        return J, self._split(g)

    def adjoint_vortex(self, dt, nsteps, rho_targets, epsilon, x0, y0, every=32):
        """adjoint() and vortex_gradient() at the current time in one call,
        with dJ_i/dU0_i kept on the device: (J [n], dJ_i/d(epsilon_i, x0_i,
        y0_i) [n][3]) for type 0 vortex initial states; rho_targets None:
        the targets of set_target_vortex()."""
        rt = self._targets(rho_targets)
        e, a, c = self._per_instance(epsilon, x0, y0)
        J = np.zeros(self.n)
        grad = np.zeros((self.n, 3))
        _check(lib().vt_batch_adjoint_vortex(self._h, dt, nsteps, every, None if rt is None else _d(rt), _d(e), _d(a),
                                             _d(c), _d(J), _d(grad)), self)
        return J, grad

    def release_adjoint(self):
        """Frees the adjoint memory adjoint() and adjoint_vortex() keep for
        re-use by the next call (the last group's recording) until close()."""
        _check(lib().vt_batch_adjoint_release(self._h))

    def vortex_gradient(self, t, epsilon, x0, y0, dJdU0):
        """dJ_i/d(epsilon_i, x0_i, y0_i) [n][3] from the per-instance dJ/dU0
        of adjoint(), for type 0 vortex initial states set at time t."""
        e = np.ascontiguousarray(np.broadcast_to(np.asarray(epsilon, dtype=np.float64), (self.n,)))
        a = np.ascontiguousarray(np.broadcast_to(np.asarray(x0, dtype=np.float64), (self.n,)))
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(y0, dtype=np.float64), (self.n,)))
        total = 4 * sum(self.ne[i] * self.np[i] for i in range(self.n))
        lam = _joined(dJdU0, total)
        if lam is None:
            lam = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in dJdU0]))
        if lam.size != total:
            raise ValueError("dJdU0 needs one [n_i][Np_i][4] array per instance")
        grad = np.zeros((self.n, 3))
        _check(lib().vt_batch_vortex_gradient(self._h, t, _d(e), _d(a), _d(c), _d(lam), _d(grad)))
        return grad


def _joined(arrays, total):
    """The flat array that a list of consecutive views was split from (as
    Batch.adjoint returns them), or None: lets a caller pass the list back
    without a copy."""
    base = getattr(arrays[0], "base", None) if len(arrays) else None
    if not isinstance(base, np.ndarray) or base.dtype != np.float64 or base.ndim != 1 or base.size != total:
        return None
    if not base.flags.c_contiguous:
        return None
    addr = base.__array_interface__["data"][0]
    for x in arrays:
        if not isinstance(x, np.ndarray) or x.base is not base or x.__array_interface__["data"][0] != addr:
            return None
        addr += x.nbytes
    return base if addr == base.__array_interface__["data"][0] + base.nbytes else None


def splitmix64(seed):
    """splitmix64 stream: yields uint64 values (Python ints)."""
    mask = (1 << 64) - 1
    x = seed & mask
    while True:
        x = (x + 0x9E3779B97F4A7C15) & mask
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & mask
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & mask
        yield z ^ (z >> 31)


def ensemble_draws(seed, n_inst, cells=(32, 256), orders_p=(1, 2, 3), eps=(0.1, 0.4), centre=(-2.5, 2.5)):
    """Per-instance draws of an ensemble: cells per side uniform in
    cells[0]..cells[1], p uniform over orders_p, epsilon uniform in eps and
    the centre uniform in centre^2.  Uniform doubles are (x >> 11) * 2^-53 of
    a splitmix64 stream; integers are lo + floor(u * count).  Draw order per
    instance: cells, p, epsilon, x0, y0."""
    g = splitmix64(seed)

    def u():
        return (next(g) >> 11) * (1.0 / (1 << 53))

    out = []
    for _ in range(n_inst):
        n = cells[0] + int(u() * (cells[1] - cells[0] + 1))
        p = orders_p[int(u() * len(orders_p))]
        e = eps[0] + (eps[1] - eps[0]) * u()
        x0 = centre[0] + (centre[1] - centre[0]) * u()
        y0 = centre[0] + (centre[1] - centre[0]) * u()
        out.append({"cells": n, "p": p, "epsilon": e, "x0": x0, "y0": y0})
    return out
