"""ctypes binding of the FP64 CPU oracle (oracle/smoke_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline leg, and only as the checker. The product path never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "diff-solver_amd"))

from ds_smoke import abi  # noqa: E402
from ds_smoke.scenes import cell_shape, face_shape  # noqa: E402

LIB_PATH = os.path.join(HERE, "libsmoke_oracle.so")
_lib = None

DP = C.POINTER(C.c_double)


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.oracle_create.restype = C.c_void_p
        L.oracle_create.argtypes = [C.POINTER(abi.Config)]
        L.oracle_last_error.restype = C.c_char_p
        for name in ["oracle_destroy", "oracle_domain_info", "oracle_masks", "oracle_collider_fields",
                     "oracle_set_solver", "oracle_source_forward", "oracle_source_vjp",
                     "oracle_buoyancy_forward", "oracle_buoyancy_vjp", "oracle_integrate_forward",
                     "oracle_integrate_vjp", "oracle_advect_velocity_forward", "oracle_advect_velocity_vjp",
                     "oracle_constrain_forward", "oracle_constrain_vjp", "oracle_advect_scalar_forward",
                     "oracle_advect_scalar_vjp", "oracle_pressure_rhs", "oracle_apply_matrix", "oracle_pcg",
                     "oracle_rbgs", "oracle_project", "oracle_projection_forward", "oracle_projection_vjp",
                     "oracle_forward_step", "oracle_vjp_step", "oracle_loss_gradient"]:
            getattr(L, name).restype = C.c_int if name not in ("oracle_destroy", "oracle_domain_info", "oracle_masks",
                                                                  "oracle_collider_fields", "oracle_set_solver") else None
        L.oracle_set_solver.argtypes = [C.c_void_p, C.c_uint32, C.c_double, C.c_uint32, C.c_uint32]
        L.oracle_pcg.argtypes = [C.c_void_p, DP, DP, DP, C.c_double, C.c_uint32, C.POINTER(abi.SolveInfo)]
        L.oracle_rbgs.argtypes = [C.c_void_p, DP, DP, C.c_uint32]
        L.oracle_vorticity_forward.restype = None
        L.oracle_vorticity_forward.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        L.oracle_vorticity_vjp.restype = None
        L.oracle_vorticity_vjp.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p, C.c_void_p, DP]
        for name in ("oracle_buoyancy_jvp", "oracle_advect_velocity_jvp", "oracle_constrain_jvp", "oracle_advect_scalar_jvp"):
            getattr(L, name).restype = None
        L.oracle_vorticity_jvp.restype = None
        L.oracle_vorticity_jvp.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_double, C.c_void_p]
        for name in ("oracle_projection_jvp", "oracle_jvp_step", "oracle_trajectory_jvp"):
            getattr(L, name).restype = C.c_int
        L.oracle_example_write_control.restype = None
        L.oracle_example_write_control.argtypes = [C.c_void_p, C.c_uint64, DP, DP, DP, C.c_void_p]
        L.oracle_example_reduce_metrics.restype = None
        L.oracle_example_reduce_metrics.argtypes = [C.c_void_p, DP, DP, C.c_void_p, C.c_void_p, DP]
        L.oracle_set_threads.restype = C.c_int
        L.oracle_set_threads.argtypes = [C.c_int]
        _lib = L
    return _lib


def set_threads(n: int) -> int:
    return lib().oracle_set_threads(int(n))


def _p(a):
    if a is None:
        return None
    assert a.dtype == np.float64 and a.flags.c_contiguous, (a.dtype, a.flags)
    return a.ctypes.data_as(DP)


def _arr3(arrs):
    # the pointer array only copies addresses: keep the (possibly temporary) arrays alive with it
    arrs = list(arrs)
    out = (DP * 3)(*[_p(a) for a in arrs])
    out._keep = arrs
    return out


def _arrn(arrs):
    arrs = list(arrs)
    out = (DP * len(arrs))(*[_p(a) if a is not None else None for a in arrs])
    out._keep = arrs
    return out


def f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


class Oracle:
    """One oracle context (Model::allocate_context restated, model.cpp:32-182)."""

    def __init__(self, scene, tolerance: float = 1e-10, solver: int = abi.PROJECTION_PCG,
                 rbgs_iterations: int = 80, max_iterations: int = 0, config=None):
        self.scene = scene
        self.cfg = config if config is not None else scene.config(tolerance=tolerance, solver=solver,
                                                                  rbgs_iterations=rbgs_iterations,
                                                                  max_iterations=max_iterations)
        self.h = lib().oracle_create(C.byref(self.cfg))
        if not self.h:
            raise ValueError("oracle_create: " + lib().oracle_last_error().decode())
        self.n = tuple(int(v) for v in self.cfg.resolution)

    def __del__(self):
        if getattr(self, "h", None):
            try:
                lib().oracle_destroy(C.c_void_p(self.h))
            except TypeError:   # interpreter shutdown: module globals are already gone
                pass
            self.h = None

    # -- shapes
    @property
    def cshape(self):
        return cell_shape(self.n)

    def fshape(self, a):
        return face_shape(self.n, a)

    def zeros_cell(self):
        return np.zeros(self.cshape)

    def zeros_faces(self):
        return tuple(np.zeros(self.fshape(a)) for a in range(3))

    def info(self) -> abi.DomainInfo:
        d = abi.DomainInfo()
        lib().oracle_domain_info(C.c_void_p(self.h), C.byref(d))
        return d

    def masks(self):
        m = np.zeros(self.cshape, dtype=np.uint8)
        c = np.zeros(self.cshape, dtype=np.uint8)
        f = np.zeros(self.cshape)
        lib().oracle_masks(C.c_void_p(self.h), m.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p), _p(f))
        return m, c, f

    def collider_fields(self):
        v = self.zeros_faces()
        d = self.zeros_cell()
        t = self.zeros_cell()
        lib().oracle_collider_fields(C.c_void_p(self.h), _arr3(v), _p(d), _p(t))
        return v, d, t

    def set_solver(self, solver=abi.PROJECTION_PCG, tolerance=0.0, max_iterations=0, rbgs_iterations=0):
        lib().oracle_set_solver(C.c_void_p(self.h), solver, tolerance, max_iterations, rbgs_iterations)

    # -- operators
    def source_forward(self, s, src):
        out = self.zeros_cell()
        lib().oracle_source_forward(C.c_void_p(self.h), _p(f64(s)), _p(f64(src)), _p(out))
        return out

    def source_vjp(self, oa, sa, srca):
        lib().oracle_source_vjp(C.c_void_p(self.h), _p(f64(oa)), _p(sa), _p(srca))

    def buoyancy_forward(self, rho, T, e, params):
        f = tuple(self.zeros_cell() for _ in range(3))
        pr = f64(params)
        lib().oracle_buoyancy_forward(C.c_void_p(self.h), _p(f64(rho)), _p(f64(T)), _arr3([f64(x) for x in e]), _p(pr), _arr3(f))
        return f

    def buoyancy_vjp(self, rho, T, params, fa, rho_adj, T_adj, e_adj, param_adj):
        lib().oracle_buoyancy_vjp(C.c_void_p(self.h), _p(f64(rho)), _p(f64(T)), _p(f64(params)),
                                  _arr3([f64(x) for x in fa]), _p(rho_adj), _p(T_adj), _arr3(e_adj), _p(param_adj))

    def vorticity_forward(self, u, confinement, f=None):
        """f + confinement * h * (N x w) on fluid cells (f defaults to zero force)."""
        f = tuple(self.zeros_cell() for _ in range(3)) if f is None else tuple(f64(x).copy() for x in f)
        lib().oracle_vorticity_forward(C.c_void_p(self.h), _arr3([f64(x) for x in u]), float(confinement), _arr3(f))
        return f

    def vorticity_vjp(self, u, confinement, fa, ua, conf_adj):
        """Accumulates the velocity adjoint into ua (faces) and the confinement adjoint into conf_adj[0]."""
        lib().oracle_vorticity_vjp(C.c_void_p(self.h), _arr3([f64(x) for x in u]), float(confinement),
                                   _arr3([f64(x) for x in fa]), _arr3(ua), _p(conf_adj))

    # -- forward mode (tangents)
    def buoyancy_jvp(self, rho, T, rho_t, T_t, e_t, params, params_t):
        f = tuple(self.zeros_cell() for _ in range(3))
        lib().oracle_buoyancy_jvp(C.c_void_p(self.h), _p(f64(rho)), _p(f64(T)), _p(f64(rho_t)), _p(f64(T_t)),
                                  _arr3([f64(x) for x in e_t]), _p(f64(params)), _p(f64(params_t)), _arr3(f))
        return f

    def vorticity_jvp(self, u, u_t, confinement, confinement_t, f_t=None):
        f_t = tuple(self.zeros_cell() for _ in range(3)) if f_t is None else tuple(f64(x).copy() for x in f_t)
        lib().oracle_vorticity_jvp(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _arr3([f64(x) for x in u_t]),
                                   float(confinement), float(confinement_t), _arr3(f_t))
        return f_t

    def advect_velocity_jvp(self, u, u_t):
        out = self.zeros_faces()
        lib().oracle_advect_velocity_jvp(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _arr3([f64(x) for x in u_t]), _arr3(out))
        return out

    def constrain_jvp(self, u_t):
        out = self.zeros_faces()
        lib().oracle_constrain_jvp(C.c_void_p(self.h), _arr3([f64(x) for x in u_t]), _arr3(out))
        return out

    def advect_scalar_jvp(self, field, src, src_t, u, u_t):
        out = self.zeros_cell()
        lib().oracle_advect_scalar_jvp(C.c_void_p(self.h), int(field), _p(f64(src)), _p(f64(src_t)), _arr3([f64(x) for x in u]),
                                       _arr3([f64(x) for x in u_t]), _p(out))
        return out

    def projection_jvp(self, u_t):
        out = self.zeros_faces()
        info = abi.SolveInfo()
        st = lib().oracle_projection_jvp(C.c_void_p(self.h), _arr3([f64(x) for x in u_t]), _arr3(out), C.byref(info))
        assert st == 0, lib().oracle_last_error()
        return out, info

    def integrate_forward(self, u, f):
        out = self.zeros_faces()
        lib().oracle_integrate_forward(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _arr3([f64(x) for x in f]), _arr3(out))
        return out

    def integrate_vjp(self, oa, ua, fa):
        lib().oracle_integrate_vjp(C.c_void_p(self.h), _arr3([f64(x) for x in oa]), _arr3(ua), _arr3(fa))

    def advect_velocity_forward(self, u):
        out = self.zeros_faces()
        lib().oracle_advect_velocity_forward(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _arr3(out))
        return out

    def advect_velocity_vjp(self, u, oa, ua):
        lib().oracle_advect_velocity_vjp(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _arr3([f64(x) for x in oa]), _arr3(ua))

    def constrain_forward(self, u):
        out = self.zeros_faces()
        lib().oracle_constrain_forward(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _arr3(out))
        return out

    def constrain_vjp(self, oa, ua):
        lib().oracle_constrain_vjp(C.c_void_p(self.h), _arr3([f64(x) for x in oa]), _arr3(ua))

    def advect_scalar_forward(self, field, src, u):
        out = self.zeros_cell()
        lib().oracle_advect_scalar_forward(C.c_void_p(self.h), int(field), _p(f64(src)), _arr3([f64(x) for x in u]), _p(out))
        return out

    def advect_scalar_vjp(self, field, src, u, oa, src_adj, ua):
        lib().oracle_advect_scalar_vjp(C.c_void_p(self.h), int(field), _p(f64(src)), _arr3([f64(x) for x in u]),
                                       _p(f64(oa)), _p(src_adj), _arr3(ua))

    def pressure_rhs(self, u):
        out = self.zeros_cell()
        lib().oracle_pressure_rhs(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _p(out))
        return out

    def apply_matrix(self, x):
        y = self.zeros_cell()
        lib().oracle_apply_matrix(C.c_void_p(self.h), _p(f64(x)), _p(y))
        return y

    def pcg(self, b, x0=None, tol=1e-10, max_it=100000):
        x = self.zeros_cell()
        info = abi.SolveInfo()
        st = lib().oracle_pcg(C.c_void_p(self.h), _p(f64(b)), _p(f64(x0)) if x0 is not None else None, _p(x),
                              tol, max_it, C.byref(info))
        assert st == 0
        return x, info

    def rbgs(self, rhs, iterations):
        p = self.zeros_cell()
        lib().oracle_rbgs(C.c_void_p(self.h), _p(f64(rhs)), _p(p), iterations)
        return p

    def project(self, u, p):
        out = self.zeros_faces()
        lib().oracle_project(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _p(f64(p)), _arr3(out))
        return out

    def projection_forward(self, u, guess=None):
        p = self.zeros_cell()
        rhs = self.zeros_cell()
        out = self.zeros_faces()
        info = abi.SolveInfo()
        st = lib().oracle_projection_forward(C.c_void_p(self.h), _arr3([f64(x) for x in u]), _p(p), _p(rhs),
                                             _p(f64(guess)) if guess is not None else None, _arr3(out), C.byref(info))
        assert st == 0, lib().oracle_last_error()
        return out, p, rhs, info

    def projection_vjp(self, oa):
        ua = self.zeros_faces()
        info = abi.SolveInfo()
        st = lib().oracle_projection_vjp(C.c_void_p(self.h), _arr3([f64(x) for x in oa]), _arr3(ua), C.byref(info))
        assert st == 0
        return ua, info

    # -- model steps
    def forward_step(self, state, control, params, guess=None):
        s = [f64(state["density"]), f64(state["temperature"])] + [f64(v) for v in state["velocity"]]
        c = [f64(control["density_source"]), f64(control["temperature_source"])] + \
            [f64(v) for v in control["external_acceleration"]]
        nxt = [self.zeros_cell(), self.zeros_cell()] + list(self.zeros_faces())
        cache = [self.zeros_cell(), self.zeros_cell()] + [self.zeros_cell() for _ in range(3)] + \
            list(self.zeros_faces()) + list(self.zeros_faces()) + [self.zeros_cell()]
        info = abi.SolveInfo()
        st = lib().oracle_forward_step(C.c_void_p(self.h), _arrn(s), _arrn(c), _p(f64(params)), _arrn(nxt),
                                       _arrn(cache), _p(f64(guess)) if guess is not None else None, C.byref(info))
        assert st == 0, lib().oracle_last_error()
        ns = {"density": nxt[0], "temperature": nxt[1], "velocity": tuple(nxt[2:5])}
        ch = {"sourced_density": cache[0], "sourced_temperature": cache[1], "force": tuple(cache[2:5]),
              "forced_velocity": tuple(cache[5:8]), "advected_velocity": tuple(cache[8:11]), "pressure": cache[11]}
        return ns, ch, info

    def vjp_step(self, next_state, cache, params, next_adj, prev_adj=None, control_adj=None, param_adj=None, state=None):
        """state: the step's input (its velocity is read with vorticity confinement on)."""
        st_arr = None
        if state is not None:
            st_arr = [f64(state["density"]), f64(state["temperature"])] + [f64(v) for v in state["velocity"]]
        nxt = [f64(next_state["density"]), f64(next_state["temperature"])] + [f64(v) for v in next_state["velocity"]]
        ch = [f64(cache["sourced_density"]), f64(cache["sourced_temperature"]), None, None, None] + \
            [f64(v) for v in cache["forced_velocity"]] + [None, None, None, None]
        na = [f64(next_adj["density"]), f64(next_adj["temperature"])] + [f64(v) for v in next_adj["velocity"]]
        if prev_adj is None:
            prev_adj = {"density": self.zeros_cell(), "temperature": self.zeros_cell(), "velocity": self.zeros_faces()}
        pa = [prev_adj["density"], prev_adj["temperature"]] + list(prev_adj["velocity"])
        ca = None
        if control_adj is not None:
            ca = [control_adj["density_source"], control_adj["temperature_source"]] + list(control_adj["external_acceleration"])
        if param_adj is None:
            param_adj = np.zeros(4)
        info = abi.SolveInfo()
        st = lib().oracle_vjp_step(C.c_void_p(self.h), _arrn(st_arr) if st_arr is not None else None, _arrn(nxt), _arrn(ch), _p(f64(params)), _arrn(na), _arrn(pa),
                                   _arrn(ca) if ca is not None else None, _p(param_adj), C.byref(info))
        assert st == 0, lib().oracle_last_error()
        return prev_adj, info

    def jvp_step(self, state, next_state, cache, params, state_t, control_t, params_t):
        """Model::jvp_step: the tangent of the next state (state / next_state / cache from forward_step)."""
        fl = lambda d: [f64(d["density"]), f64(d["temperature"])] + [f64(v) for v in d["velocity"]]
        fc = lambda d: [f64(d["density_source"]), f64(d["temperature_source"])] + [f64(v) for v in d["external_acceleration"]]
        ch = [f64(cache["sourced_density"]), f64(cache["sourced_temperature"]), None, None, None] + \
            [f64(v) for v in cache["forced_velocity"]] + [None, None, None, None]
        nt = [self.zeros_cell(), self.zeros_cell()] + list(self.zeros_faces())
        info = abi.SolveInfo()
        st = lib().oracle_jvp_step(C.c_void_p(self.h), _arrn(fl(state)), _arrn(fl(next_state)), _arrn(ch), _p(f64(params)),
                                   _arrn(fl(state_t)), _arrn(fc(control_t)), _p(f64(params_t)), _arrn(nt), C.byref(info))
        assert st == 0, lib().oracle_last_error()
        return {"density": nt[0], "temperature": nt[1], "velocity": tuple(nt[2:5])}, info

    def trajectory_jvp(self, state, control, params, steps, state_t, control_t, params_t):
        """solver::jvp in one sweep: (final state, final state tangent, TrajectoryInfo)."""
        fl = lambda d: [f64(d["density"]), f64(d["temperature"])] + [f64(v) for v in d["velocity"]]
        fc = lambda d: [f64(d["density_source"]), f64(d["temperature_source"])] + [f64(v) for v in d["external_acceleration"]]
        fin = [self.zeros_cell(), self.zeros_cell()] + list(self.zeros_faces())
        fin_t = [self.zeros_cell(), self.zeros_cell()] + list(self.zeros_faces())
        info = abi.TrajectoryInfo()
        st = lib().oracle_trajectory_jvp(C.c_void_p(self.h), _arrn(fl(state)), _arrn(fc(control)), _p(f64(params)), int(steps),
                                         _arrn(fl(state_t)), _arrn(fc(control_t)), _p(f64(params_t)), _arrn(fin), _arrn(fin_t),
                                         C.byref(info))
        assert st == 0, lib().oracle_last_error()
        as_state = lambda v: {"density": v[0], "temperature": v[1], "velocity": tuple(v[2:5])}
        return as_state(fin), as_state(fin_t), info

    def loss_gradient(self, state, control, params, steps, k, target, want_control=False):
        s = [f64(state["density"]), f64(state["temperature"])] + [f64(v) for v in state["velocity"]]
        c = [f64(control["density_source"]), f64(control["temperature_source"])] + \
            [f64(v) for v in control["external_acceleration"]]
        fin = [self.zeros_cell(), self.zeros_cell()] + list(self.zeros_faces())
        ga = [self.zeros_cell(), self.zeros_cell()] + list(self.zeros_faces())
        ca = [self.zeros_cell() for _ in range(5)] if want_control else None
        pa = np.zeros(4)
        loss = C.c_double()
        info = abi.TrajectoryInfo()
        st = lib().oracle_loss_gradient(C.c_void_p(self.h), _arrn(s), _arrn(c), _p(f64(params)), int(steps), int(k),
                                        _p(f64(target)), _arrn(fin), C.byref(loss), _arrn(ga),
                                        _arrn(ca) if ca is not None else None, _p(pa), C.byref(info))
        assert st == 0, lib().oracle_last_error()
        out = {"loss": loss.value, "final": {"density": fin[0], "temperature": fin[1], "velocity": tuple(fin[2:5])},
               "grad": {"density": ga[0], "temperature": ga[1], "velocity": tuple(ga[2:5])},
               "params": pa, "info": info}
        if ca is not None:
            out["control"] = {"density_source": ca[0], "temperature_source": ca[1], "external_acceleration": tuple(ca[2:5])}
        return out

    # -- forward example harness (examples/smoke/forward/simulation.cu)
    def example_write_control(self, step, jet):
        """(density source, temperature source, acceleration) of the double jet at `step`; jet: the
        16 values of ds_jet_control in its field order."""
        s_rho, s_t = self.zeros_cell(), self.zeros_cell()
        e = tuple(self.zeros_cell() for _ in range(3))
        j = f64(jet)
        assert j.size == 16
        lib().oracle_example_write_control(C.c_void_p(self.h), int(step), _p(j), _p(s_rho), _p(s_t), _arr3(e))
        return s_rho, s_t, e

    def example_reduce_metrics(self, rho, T, pre, post):
        """The six raw metrics of reduce_metrics_kernel over the fluid cells."""
        out = np.zeros(6)
        lib().oracle_example_reduce_metrics(C.c_void_p(self.h), _p(f64(rho)), _p(f64(T)), _arr3([f64(v) for v in pre]),
                                            _arr3([f64(v) for v in post]), _p(out))
        return out

    def simulate(self, state, control, params, steps):
        st = state
        infos = []
        for _ in range(steps):
            st, _, info = self.forward_step(st, control, params)
            infos.append(info.as_dict())
        return st, infos
