"""Python host side of the MI355X smoke solver: a ctypes binding of libds_smoke.so.

Mirrors the reference's Model interface (solvers/smoke/model.ixx:51-96: allocate_context,
allocate_state, forward_step, vjp_step) and trajectory driver (solvers/solver.ixx:53-150) over the
C ABI declared in include/ds/ds_smoke.h. PyTorch provides device memory and the stream; every
computation runs in the HIP library, and loading fails loudly when the library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi
from .scenes import cell_shape, face_shape

_LIB = None


class SolverError(RuntimeError):
    def __init__(self, status: int, where: str, message: str):
        super().__init__(f"{where}: {abi.STATUS_NAMES.get(status, status)} ({status}): {message}")
        self.status = status


def load_library(path: str | None = None):
    """Load the in-tree HIP library (built by __graft_entry__.build()); DS_SMOKE_LIBRARY overrides the path."""
    global _LIB
    if _LIB is not None:
        return _LIB
    # torch first: the library then binds to the HIP runtime torch has loaded (one runtime per
    # process; loading the library first can leave torch and the library on different runtimes)
    import torch  # noqa: F401

    path = path or os.environ.get("DS_SMOKE_LIBRARY") or abi.library_path()
    if not os.path.exists(path):
        raise RuntimeError(f"HIP smoke library not found at {path}; run __graft_entry__.build() first")
    L = C.CDLL(path)
    P = C.c_void_p
    U32 = C.c_uint32
    sigs = {
        "ds_last_error": (C.c_char_p, []),
        "ds_config_cell_count": (C.c_uint64, [C.POINTER(abi.Config)]),
        "ds_config_face_count": (C.c_uint64, [C.POINTER(abi.Config), C.c_uint32]),
        "ds_context_create": (C.c_int32, [C.POINTER(abi.Config), C.c_uint32, P, C.POINTER(P)]),
        "ds_context_destroy": (C.c_int32, [P]),
        "ds_context_stream": (P, [P]),
        "ds_context_domain_info": (C.c_int32, [P, C.POINTER(abi.DomainInfo)]),
        "ds_context_domain_masks": (C.c_int32, [P, P, P]),
        "ds_allocate": (C.c_int32, [P, C.c_size_t, C.POINTER(P)]),
        "ds_release": (C.c_int32, [P, P]),
        "ds_copy_from_host": (C.c_int32, [P, P, P, C.c_size_t]),
        "ds_copy_to_host": (C.c_int32, [P, P, P, C.c_size_t]),
        "ds_copy_device": (C.c_int32, [P, P, P, C.c_size_t]),
        "ds_zero": (C.c_int32, [P, P, C.c_size_t]),
        "ds_synchronize": (C.c_int32, [P]),
        "ds_source_forward": (C.c_int32, [P, P, P, P]),
        "ds_source_vjp": (C.c_int32, [P, P, P, P]),
        "ds_buoyancy_forward": (C.c_int32, [P, P, P, C.POINTER(P), C.POINTER(abi.Parameters), C.POINTER(P)]),
        "ds_buoyancy_vjp": (C.c_int32, [P, P, P, C.POINTER(abi.Parameters), C.POINTER(P), P, P, C.POINTER(P),
                                        C.POINTER(abi.ParameterAdjoint)]),
        "ds_vorticity_forward": (C.c_int32, [P, C.POINTER(P), C.POINTER(abi.Parameters), C.POINTER(P)]),
        "ds_vorticity_vjp": (C.c_int32, [P, C.POINTER(P), C.POINTER(abi.Parameters), C.POINTER(P), C.POINTER(P),
                                         C.POINTER(abi.ParameterAdjoint)]),
        "ds_integrate_velocity_forward": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "ds_integrate_velocity_vjp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "ds_advect_velocity_forward": (C.c_int32, [P, C.POINTER(P), C.POINTER(P)]),
        "ds_advect_velocity_vjp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "ds_constrain_velocity_forward": (C.c_int32, [P, C.POINTER(P), C.POINTER(P)]),
        "ds_constrain_velocity_vjp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P)]),
        "ds_advect_scalar_forward": (C.c_int32, [P, C.c_uint32, P, C.POINTER(P), P]),
        "ds_advect_scalar_vjp": (C.c_int32, [P, C.c_uint32, P, C.POINTER(P), P, P, C.POINTER(P)]),
        "ds_projection_forward": (C.c_int32, [P, C.POINTER(P), P, P, P, C.POINTER(P), C.POINTER(abi.SolveInfo)]),
        "ds_projection_vjp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(abi.SolveInfo)]),
        "ds_pcg_solve": (C.c_int32, [P, P, P, P, C.c_double, C.c_uint32, C.POINTER(abi.SolveInfo)]),
        "ds_apply_pressure_matrix": (C.c_int32, [P, P, P]),
        "ds_stream_copy": (C.c_int32, [P, P, P, C.c_uint64]),
        "ds_profile_enable": (C.c_int32, [P, C.c_uint32]),
        "ds_profile_read": (C.c_int32, [P, C.POINTER(abi.KernelTiming), C.c_uint32, C.POINTER(C.c_uint32)]),
        "ds_copy_state": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.State)]),
        "ds_copy_state_adjoint": (C.c_int32, [P, C.POINTER(abi.StateAdjoint), C.POINTER(abi.StateAdjoint)]),
        "ds_accumulate_state_adjoint": (C.c_int32, [P, C.POINTER(abi.StateAdjoint), C.POINTER(abi.StateAdjoint)]),
        "ds_forward_step": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.POINTER(abi.Parameters),
                                        C.POINTER(abi.State), C.POINTER(abi.StepCache), C.POINTER(abi.StepInfo)]),
        "ds_vjp_step": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.POINTER(abi.Parameters),
                                    C.POINTER(abi.State), C.POINTER(abi.StepCache), C.POINTER(abi.StateAdjoint),
                                    C.POINTER(abi.StateAdjoint), C.POINTER(abi.ControlAdjoint),
                                    C.POINTER(abi.ParameterAdjoint), C.POINTER(abi.StepInfo)]),
        "ds_simulate": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.c_uint32,
                                    C.POINTER(abi.Parameters), C.c_uint32, C.POINTER(abi.State),
                                    C.POINTER(abi.TrajectoryInfo)]),
        "ds_loss_gradient": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.c_uint32,
                                         C.POINTER(abi.Parameters), C.c_uint32, C.c_uint32, P, C.POINTER(abi.State),
                                         C.POINTER(abi.StateAdjoint), C.POINTER(abi.ControlAdjoint),
                                         C.POINTER(abi.ParameterAdjoint), C.POINTER(abi.TrajectoryInfo)]),
        "ds_context_last_solution": (C.c_int32, [P, C.c_uint32, P]),
        "ds_warm_start_parameter_count": (C.c_uint64, []),
        "ds_warm_start_load": (C.c_int32, [P, P, C.c_uint64, C.c_double]),
        "ds_warm_start_infer": (C.c_int32, [P, P, P]),
        "ds_jvp_step": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.POINTER(abi.Parameters),
                                    C.POINTER(abi.State), C.POINTER(abi.StepCache), C.POINTER(abi.State),
                                    C.POINTER(abi.Control), C.POINTER(abi.Parameters), C.POINTER(abi.State),
                                    C.POINTER(abi.StepInfo)]),
        "ds_trajectory_jvp": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.c_uint32,
                                          C.POINTER(abi.Parameters), C.c_uint32, C.POINTER(abi.State), C.POINTER(abi.Control),
                                          C.POINTER(abi.Parameters), C.POINTER(abi.State), C.POINTER(abi.State),
                                          C.POINTER(abi.TrajectoryInfo)]),
        "ds_buoyancy_jvp": (C.c_int32, [P, P, P, P, P, C.POINTER(P), C.POINTER(abi.Parameters), C.POINTER(abi.Parameters),
                                        C.POINTER(P)]),
        "ds_vorticity_jvp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(abi.Parameters), C.POINTER(abi.Parameters),
                                         C.POINTER(P)]),
        "ds_advect_velocity_jvp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(P)]),
        "ds_constrain_velocity_jvp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P)]),
        "ds_advect_scalar_jvp": (C.c_int32, [P, C.c_uint32, P, P, C.POINTER(P), C.POINTER(P), P]),
        "ds_projection_jvp": (C.c_int32, [P, C.POINTER(P), C.POINTER(P), C.POINTER(abi.SolveInfo)]),
        "ds_trajectory_vjp": (C.c_int32, [P, C.POINTER(abi.State), C.POINTER(abi.Control), C.c_uint32,
                                          C.POINTER(abi.Parameters), C.c_uint32, C.c_uint32, C.POINTER(abi.StateAdjoint),
                                          C.POINTER(abi.State), C.POINTER(abi.StateAdjoint), C.POINTER(abi.ControlAdjoint),
                                          C.POINTER(abi.ParameterAdjoint), C.POINTER(abi.TrajectoryInfo)]),
        "ds_example_write_control": (C.c_int32, [P, C.c_uint64, C.POINTER(abi.JetControl), P, P, C.POINTER(P)]),
        "ds_example_reduce_metrics": (C.c_int32, [P, P, P, C.POINTER(P), C.POINTER(P), P]),
        "ds_example_reduce_metrics_static": (C.c_int32, [C.POINTER(abi.Config), P, P, P, C.POINTER(P), C.POINTER(P), P, P]),
        "ds_warm_start_parameter_count": (C.c_uint64, []),
        "ds_warm_start_load": (C.c_int32, [P, P, C.c_uint64, C.c_double]),
        "ds_warm_start_infer": (C.c_int32, [P, P, P]),
        "ds_warm_start_guess": (C.c_int32, [P, P, P, P]),
        "ds_warm_start_configure": (C.c_int32, [P, C.c_uint32]),
        "ds_warm_start_pair": (C.c_int32, [P, P, P, P]),
        "ds_unet_trainer_create": (C.c_int32, [P, C.c_uint32, P, C.c_uint64, C.POINTER(P)]),
        "ds_unet_trainer_destroy": (C.c_int32, [P]),
        "ds_unet_trainer_step": (C.c_int32, [P, P, C.c_uint32, C.c_float, P]),
        "ds_unet_trainer_read": (C.c_int32, [P, C.c_uint32, P]),
        "ds_unet_conv3_forward": (C.c_int32, [U32] * 8 + [P] * 5 + [P]),
        "ds_unet_conv3_dgrad": (C.c_int32, [U32] * 7 + [P] * 6),
        "ds_unet_conv3_wgrad": (C.c_int32, [U32] * 7 + [P] * 6),
        "ds_unet_tconv2_forward": (C.c_int32, [U32] * 6 + [P] * 5),
        "ds_unet_tconv2_dgrad": (C.c_int32, [U32] * 6 + [P] * 5),
        "ds_unet_tconv2_wgrad": (C.c_int32, [U32] * 6 + [P] * 5),
        "ds_unet_maxpool_forward": (C.c_int32, [U32] * 5 + [P] * 3),
        "ds_unet_maxpool_backward": (C.c_int32, [U32] * 5 + [P] * 5),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = L
    return L


def exported_symbols() -> list[str]:
    """Entry points declared by include/ds/ds_smoke.h that this binding knows."""
    load_library()
    return sorted(n for n in dir(_LIB) if n.startswith("ds_"))


def _ptr(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _p3(ts):
    return (C.c_void_p * 3)(*[t.data_ptr() if t is not None else None for t in ts])


class Context:
    """An execution context (Model::allocate_context, model.cpp:32-182) on one GPU.

    Fields are torch tensors in the reference layout: cell fields (nz, ny, nx), face component a
    with extent n_a + 1 on its own axis. All work runs on this context's torch stream.
    """

    def __init__(self, config: abi.Config, mode: int = abi.MODE_DIFFERENTIABLE, device: int = 0):
        import torch

        self.torch = torch
        self.lib = load_library()
        self.config = config
        self.device = torch.device("cuda", device)
        torch.cuda.set_device(self.device)
        self.stream = torch.cuda.Stream(device=self.device)
        h = C.c_void_p()
        st = self.lib.ds_context_create(C.byref(config), mode, C.c_void_p(self.stream.cuda_stream), C.byref(h))
        self._check(st, "ds_context_create")
        self.h = h
        self.mode = mode
        self.n = tuple(int(v) for v in config.resolution)
        self.domain = abi.DomainInfo()
        self._check(self.lib.ds_context_domain_info(self.h, C.byref(self.domain)), "ds_context_domain_info")

    def _check(self, status, where):
        if status != abi.DS_OK:
            raise SolverError(status, where, self.lib.ds_last_error().decode())

    def close(self):
        if getattr(self, "h", None):
            self.lib.ds_context_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- buffers
    def _zeros(self, shape, dtype=None):
        with self.torch.cuda.stream(self.stream):
            return self.torch.zeros(shape, dtype=dtype or self.torch.float32, device=self.device)

    def zeros_cell(self):
        return self._zeros(cell_shape(self.n))

    def zeros_faces(self):
        return tuple(self._zeros(face_shape(self.n, a)) for a in range(3))

    def to_device(self, a):
        with self.torch.cuda.stream(self.stream):
            return self.torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32)).to(self.device, non_blocking=False)

    def synchronize(self):
        self.stream.synchronize()

    def allocate_state(self):
        return {"density": self.zeros_cell(), "temperature": self.zeros_cell(), "velocity": self.zeros_faces()}

    allocate_state_adjoint = allocate_state

    def allocate_control(self):
        return {"density_source": self.zeros_cell(), "temperature_source": self.zeros_cell(),
                "external_acceleration": tuple(self.zeros_cell() for _ in range(3))}

    allocate_control_adjoint = allocate_control

    def allocate_parameters(self, values=(0.0, 0.0, 0.0, 0.0)):
        return self.to_device(np.asarray(values, dtype=np.float32))

    def allocate_parameter_adjoint(self):
        return self._zeros((4,), dtype=self.torch.float64)

    def allocate_step_cache(self, pressure=True):
        return {"sourced_density": self.zeros_cell(), "sourced_temperature": self.zeros_cell(),
                "force": tuple(self.zeros_cell() for _ in range(3)), "forced_velocity": self.zeros_faces(),
                "advected_velocity": self.zeros_faces(),
                "pressure": self._zeros((2,) + cell_shape(self.n)) if pressure else None, "pressure_guess": None}

    def state_from_host(self, s):
        return {"density": self.to_device(s["density"]), "temperature": self.to_device(s["temperature"]),
                "velocity": tuple(self.to_device(v) for v in s["velocity"])}

    def control_from_host(self, k):
        return {"density_source": self.to_device(k["density_source"]),
                "temperature_source": self.to_device(k["temperature_source"]),
                "external_acceleration": tuple(self.to_device(v) for v in k["external_acceleration"])}

    def to_host(self, d):
        self.synchronize()
        if isinstance(d, dict):
            return {k: self.to_host(v) for k, v in d.items()}
        if isinstance(d, (tuple, list)):
            return tuple(self.to_host(v) for v in d)
        if d is None:
            return None
        return d.detach().cpu().numpy()

    # ---- C views
    @staticmethod
    def state_view(s):
        return abi.State(s["density"].data_ptr(), s["temperature"].data_ptr(), (C.c_void_p * 3)(*[v.data_ptr() for v in s["velocity"]]))

    @staticmethod
    def control_view(k):
        return abi.Control(k["density_source"].data_ptr(), k["temperature_source"].data_ptr(),
                           (C.c_void_p * 3)(*[v.data_ptr() for v in k["external_acceleration"]]))

    @staticmethod
    def params_view(p):
        base = p.data_ptr()
        return abi.Parameters(base, base + 4, base + 8, base + 12)

    @staticmethod
    def param_adjoint_view(p):
        base = p.data_ptr()
        return abi.ParameterAdjoint(base, base + 8, base + 16, base + 24)

    def cache_view(self, ch):
        v = abi.StepCache()
        v.sourced_density = ch["sourced_density"].data_ptr()
        v.sourced_temperature = ch["sourced_temperature"].data_ptr()
        for a in range(3):
            v.force[a] = ch["force"][a].data_ptr() if ch.get("force") is not None else None
            v.forced_velocity[a] = ch["forced_velocity"][a].data_ptr()
            v.advected_velocity[a] = ch["advected_velocity"][a].data_ptr() if ch.get("advected_velocity") is not None else None
        v.pressure = ch["pressure"].data_ptr() if ch.get("pressure") is not None else None
        v.pressure_guess = ch["pressure_guess"].data_ptr() if ch.get("pressure_guess") is not None else None
        return v

    # ---- forward example harness (examples/smoke/forward/simulation.cu)
    def example_write_control(self, step, jet, control=None):
        """write_control_kernel: the double-jet control of `step` written into control."""
        control = control or self.allocate_control()
        self._check(self.lib.ds_example_write_control(self.h, int(step), C.byref(jet), _ptr(control["density_source"]),
                                                      _ptr(control["temperature_source"]), _p3(control["external_acceleration"])),
                    "ds_example_write_control")
        return control

    def example_reduce_metrics(self, density, temperature, pre_velocity, post_velocity, out=None):
        """reduce_metrics_kernel: the six raw metrics as a float64 device tensor."""
        out = out if out is not None else self._zeros((6,), dtype=self.torch.float64)
        self._check(self.lib.ds_example_reduce_metrics(self.h, _ptr(density), _ptr(temperature), _p3(pre_velocity),
                                                       _p3(post_velocity), _ptr(out)), "ds_example_reduce_metrics")
        return out

    # ---- model steps (model.cpp:341-389)
    def forward_step(self, state, control, params, next_state=None, cache=None, check=True):
        next_state = next_state or self.allocate_state()
        cache = cache or self.allocate_step_cache()
        info = abi.StepInfo()
        st = self.lib.ds_forward_step(self.h, C.byref(self.state_view(state)), C.byref(self.control_view(control)),
                                      C.byref(self.params_view(params)), C.byref(self.state_view(next_state)),
                                      C.byref(self.cache_view(cache)), C.byref(info))
        if check:
            self._check(st, "ds_forward_step")
        return next_state, cache, info.projection

    def vjp_step(self, state, control, params, next_state, cache, next_adj, prev_adj=None, control_adj=None,
                 param_adj=None, check=True):
        prev_adj = prev_adj or self.allocate_state_adjoint()
        info = abi.StepInfo()
        ca = C.byref(self.control_view(control_adj)) if control_adj is not None else None
        pa = C.byref(self.param_adjoint_view(param_adj)) if param_adj is not None else None
        st = self.lib.ds_vjp_step(self.h, C.byref(self.state_view(state)), C.byref(self.control_view(control)),
                                  C.byref(self.params_view(params)), C.byref(self.state_view(next_state)),
                                  C.byref(self.cache_view(cache)), C.byref(self.state_view(next_adj)),
                                  C.byref(self.state_view(prev_adj)), ca, pa, C.byref(info))
        if check:
            self._check(st, "ds_vjp_step")
        return prev_adj, info.projection

    def jvp_step(self, state, control, params, next_state, cache, state_t, control_t, params_t, next_t=None, check=True):
        """Model::jvp_step: the tangent of the next state (state / next_state / cache from forward_step)."""
        next_t = next_t or self.allocate_state()
        info = abi.StepInfo()
        st = self.lib.ds_jvp_step(self.h, C.byref(self.state_view(state)), C.byref(self.control_view(control)),
                                  C.byref(self.params_view(params)), C.byref(self.state_view(next_state)),
                                  C.byref(self.cache_view(cache)), C.byref(self.state_view(state_t)),
                                  C.byref(self.control_view(control_t)), C.byref(self.params_view(params_t)),
                                  C.byref(self.state_view(next_t)), C.byref(info))
        if check:
            self._check(st, "ds_jvp_step")
        return next_t, info.projection

    def accumulate_state_adjoint(self, src, dst):
        self._check(self.lib.ds_accumulate_state_adjoint(self.h, C.byref(self.state_view(src)),
                                                         C.byref(self.state_view(dst))), "ds_accumulate_state_adjoint")

    def copy_state(self, src, dst):
        self._check(self.lib.ds_copy_state(self.h, C.byref(self.state_view(src)), C.byref(self.state_view(dst))),
                    "ds_copy_state")

    # ---- trajectory (solver.ixx:53-150)
    def simulate(self, initial, control, params, steps, final_state=None, check=True):
        final_state = final_state or self.allocate_state()
        info = abi.TrajectoryInfo()
        st = self.lib.ds_simulate(self.h, C.byref(self.state_view(initial)), C.byref(self.control_view(control)), 1,
                                  C.byref(self.params_view(params)), steps, C.byref(self.state_view(final_state)),
                                  C.byref(info))
        if check:
            self._check(st, "ds_simulate")
        return final_state, info

    def loss_gradient(self, initial, control, params, steps, checkpoint_interval, target, initial_adj=None,
                      control_adj=None, param_adj=None, final_state=None, check=True):
        """0.5 ||rho_T - target||^2 and its gradient (checkpoint every k steps, recompute between)."""
        initial_adj = initial_adj or self.allocate_state_adjoint()
        info = abi.TrajectoryInfo()
        ca = C.byref(self.control_view(control_adj)) if control_adj is not None else None
        pa = C.byref(self.param_adjoint_view(param_adj)) if param_adj is not None else None
        fs = C.byref(self.state_view(final_state)) if final_state is not None else None
        st = self.lib.ds_loss_gradient(self.h, C.byref(self.state_view(initial)), C.byref(self.control_view(control)), 1,
                                       C.byref(self.params_view(params)), steps, checkpoint_interval, _ptr(target), fs,
                                       C.byref(self.state_view(initial_adj)), ca, pa, C.byref(info))
        if check:
            self._check(st, "ds_loss_gradient")
        return info.loss, initial_adj, info

    def trajectory_vjp(self, initial, control, params, steps, checkpoint_interval, final_adj, initial_adj=None,
                       control_adj=None, param_adj=None, final_state=None, check=True):
        """Reverse sweep for a given adjoint of the final state (accumulates into the adjoints)."""
        initial_adj = initial_adj or self.allocate_state_adjoint()
        info = abi.TrajectoryInfo()
        ca = C.byref(self.control_view(control_adj)) if control_adj is not None else None
        pa = C.byref(self.param_adjoint_view(param_adj)) if param_adj is not None else None
        fs = C.byref(self.state_view(final_state)) if final_state is not None else None
        st = self.lib.ds_trajectory_vjp(self.h, C.byref(self.state_view(initial)), C.byref(self.control_view(control)), 1,
                                        C.byref(self.params_view(params)), steps, checkpoint_interval,
                                        C.byref(self.state_view(final_adj)), fs, C.byref(self.state_view(initial_adj)), ca, pa,
                                        C.byref(info))
        if check:
            self._check(st, "ds_trajectory_vjp")
        return initial_adj, info

    def trajectory_jvp(self, initial, control, params, steps, initial_t, control_t, params_t, final_state=None, final_t=None,
                       check=True):
        """Forward mode over `steps` steps in one sweep: (state_T, its tangent, TrajectoryInfo)."""
        final_state = final_state or self.allocate_state()
        final_t = final_t or self.allocate_state()
        info = abi.TrajectoryInfo()
        st = self.lib.ds_trajectory_jvp(self.h, C.byref(self.state_view(initial)), C.byref(self.control_view(control)), 1,
                                        C.byref(self.params_view(params)), steps, C.byref(self.state_view(initial_t)),
                                        C.byref(self.control_view(control_t)), C.byref(self.params_view(params_t)),
                                        C.byref(self.state_view(final_state)), C.byref(self.state_view(final_t)), C.byref(info))
        if check:
            self._check(st, "ds_trajectory_jvp")
        return final_state, final_t, info

    # ---- operators (tests)
    def masks(self):
        m = np.zeros(cell_shape(self.n), dtype=np.uint8)
        c = np.zeros(cell_shape(self.n), dtype=np.uint8)
        self._check(self.lib.ds_context_domain_masks(self.h, m.ctypes.data_as(C.c_void_p), c.ctypes.data_as(C.c_void_p)),
                    "ds_context_domain_masks")
        return m, c

    def source_forward(self, s, src):
        out = self.zeros_cell()
        self._check(self.lib.ds_source_forward(self.h, _ptr(s), _ptr(src), _ptr(out)), "ds_source_forward")
        return out

    def source_vjp(self, oa, sa, srca):
        self._check(self.lib.ds_source_vjp(self.h, _ptr(oa), _ptr(sa), _ptr(srca)), "ds_source_vjp")

    def buoyancy_forward(self, rho, T, e, params):
        f = tuple(self.zeros_cell() for _ in range(3))
        self._check(self.lib.ds_buoyancy_forward(self.h, _ptr(rho), _ptr(T), _p3(e), C.byref(self.params_view(params)),
                                                 _p3(f)), "ds_buoyancy_forward")
        return f

    def buoyancy_vjp(self, rho, T, params, fa, rho_adj, T_adj, e_adj, param_adj):
        self._check(self.lib.ds_buoyancy_vjp(self.h, _ptr(rho), _ptr(T), C.byref(self.params_view(params)), _p3(fa),
                                             _ptr(rho_adj), _ptr(T_adj), _p3(e_adj),
                                             C.byref(self.param_adjoint_view(param_adj))), "ds_buoyancy_vjp")

    def vorticity_forward(self, u, params, f=None):
        """f += the vorticity-confinement force of the face velocity u (f defaults to zeros)."""
        f = tuple(self.zeros_cell() for _ in range(3)) if f is None else f
        self._check(self.lib.ds_vorticity_forward(self.h, _p3(u), C.byref(self.params_view(params)), _p3(f)), "ds_vorticity_forward")
        return f

    def vorticity_vjp(self, u, params, fa, ua, param_adj=None):
        """ua += the velocity adjoint for the cell force adjoint fa; param_adj[3] += the confinement adjoint."""
        pa = C.byref(self.param_adjoint_view(param_adj)) if param_adj is not None else None
        self._check(self.lib.ds_vorticity_vjp(self.h, _p3(u), C.byref(self.params_view(params)), _p3(fa), _p3(ua), pa),
                    "ds_vorticity_vjp")

    def buoyancy_jvp(self, rho, T, rho_t, T_t, e_t, params, params_t):
        f = tuple(self.zeros_cell() for _ in range(3))
        self._check(self.lib.ds_buoyancy_jvp(self.h, _ptr(rho), _ptr(T), _ptr(rho_t), _ptr(T_t), _p3(e_t),
                                             C.byref(self.params_view(params)), C.byref(self.params_view(params_t)), _p3(f)),
                    "ds_buoyancy_jvp")
        return f

    def vorticity_jvp(self, u, u_t, params, params_t, f_t=None):
        f_t = tuple(self.zeros_cell() for _ in range(3)) if f_t is None else f_t
        self._check(self.lib.ds_vorticity_jvp(self.h, _p3(u), _p3(u_t), C.byref(self.params_view(params)),
                                              C.byref(self.params_view(params_t)), _p3(f_t)), "ds_vorticity_jvp")
        return f_t

    def advect_velocity_jvp(self, u, u_t):
        out = self.zeros_faces()
        self._check(self.lib.ds_advect_velocity_jvp(self.h, _p3(u), _p3(u_t), _p3(out)), "ds_advect_velocity_jvp")
        return out

    def constrain_jvp(self, u_t):
        out = self.zeros_faces()
        self._check(self.lib.ds_constrain_velocity_jvp(self.h, _p3(u_t), _p3(out)), "ds_constrain_velocity_jvp")
        return out

    def advect_scalar_jvp(self, field, src, src_t, u, u_t):
        out = self.zeros_cell()
        self._check(self.lib.ds_advect_scalar_jvp(self.h, field, _ptr(src), _ptr(src_t), _p3(u), _p3(u_t), _ptr(out)),
                    "ds_advect_scalar_jvp")
        return out

    def projection_jvp(self, u_t, check=True):
        out = self.zeros_faces()
        info = abi.SolveInfo()
        st = self.lib.ds_projection_jvp(self.h, _p3(u_t), _p3(out), C.byref(info))
        if check:
            self._check(st, "ds_projection_jvp")
        return out, info

    def integrate_forward(self, u, f):
        out = self.zeros_faces()
        self._check(self.lib.ds_integrate_velocity_forward(self.h, _p3(u), _p3(f), _p3(out)), "ds_integrate_velocity_forward")
        return out

    def integrate_vjp(self, oa, ua, fa):
        self._check(self.lib.ds_integrate_velocity_vjp(self.h, _p3(oa), _p3(ua), _p3(fa)), "ds_integrate_velocity_vjp")

    def advect_velocity_forward(self, u):
        out = self.zeros_faces()
        self._check(self.lib.ds_advect_velocity_forward(self.h, _p3(u), _p3(out)), "ds_advect_velocity_forward")
        return out

    def advect_velocity_vjp(self, u, oa, ua):
        self._check(self.lib.ds_advect_velocity_vjp(self.h, _p3(u), _p3(oa), _p3(ua)), "ds_advect_velocity_vjp")

    def constrain_forward(self, u):
        out = self.zeros_faces()
        self._check(self.lib.ds_constrain_velocity_forward(self.h, _p3(u), _p3(out)), "ds_constrain_velocity_forward")
        return out

    def constrain_vjp(self, oa, ua):
        self._check(self.lib.ds_constrain_velocity_vjp(self.h, _p3(oa), _p3(ua)), "ds_constrain_velocity_vjp")

    def advect_scalar_forward(self, field, src, u):
        out = self.zeros_cell()
        self._check(self.lib.ds_advect_scalar_forward(self.h, field, _ptr(src), _p3(u), _ptr(out)), "ds_advect_scalar_forward")
        return out

    def advect_scalar_vjp(self, field, src, u, oa, src_adj, ua):
        self._check(self.lib.ds_advect_scalar_vjp(self.h, field, _ptr(src), _p3(u), _ptr(oa), _ptr(src_adj), _p3(ua)),
                    "ds_advect_scalar_vjp")

    def projection_forward(self, u, guess=None, check=True):
        out = self.zeros_faces()
        p = self.zeros_cell()
        rhs = self.zeros_cell()
        info = abi.SolveInfo()
        st = self.lib.ds_projection_forward(self.h, _p3(u), _ptr(p), _ptr(rhs), _ptr(guess), _p3(out), C.byref(info))
        if check:
            self._check(st, "ds_projection_forward")
        return out, p, rhs, info

    def projection_vjp(self, oa, ua=None, check=True):
        ua = ua or self.zeros_faces()
        info = abi.SolveInfo()
        st = self.lib.ds_projection_vjp(self.h, _p3(oa), _p3(ua), C.byref(info))
        if check:
            self._check(st, "ds_projection_vjp")
        return ua, info

    def pcg_solve(self, b, x0=None, tolerance=0.0, max_iterations=0, check=True):
        x = self.zeros_cell()
        info = abi.SolveInfo()
        st = self.lib.ds_pcg_solve(self.h, _ptr(b), _ptr(x), _ptr(x0), tolerance, max_iterations, C.byref(info))
        if check:
            self._check(st, "ds_pcg_solve")
        return x, info, st

    def warm_start_load(self, params: np.ndarray, output_scale: float = 1.0, levels: int = 0):
        """Load the warm-start network's FP32 parameters (state_dict order); levels: run it on the
        grid coarsened 2^levels times (ds_warm_start_configure)."""
        self._check(self.lib.ds_warm_start_configure(self.h, int(levels)), "ds_warm_start_configure")
        flat = np.ascontiguousarray(params, dtype=np.float32).ravel()
        self._check(self.lib.ds_warm_start_load(self.h, flat.ctypes.data_as(C.c_void_p), flat.size, float(output_scale)),
                    "ds_warm_start_load")

    def warm_start_infer(self, b):
        x = self.zeros_cell()
        self._check(self.lib.ds_warm_start_infer(self.h, _ptr(b), _ptr(x)), "ds_warm_start_infer")
        return x

    def last_solution(self, which: int, out=None):
        """The last forward (0) or adjoint (1) solution as an FP32 cell field."""
        x = self.zeros_cell() if out is None else out
        self._check(self.lib.ds_context_last_solution(self.h, int(which), _ptr(x)), "ds_context_last_solution")
        return x

    def warm_start_guess(self, b, base=None):
        """The network's guess base + gamma N(b - A base) (base None: gamma N(b))."""
        x = self.zeros_cell()
        self._check(self.lib.ds_warm_start_guess(self.h, _ptr(b), _ptr(base), _ptr(x)), "ds_warm_start_guess")
        return x

    def warm_start_pair(self, error, residual=None, energy=None):
        """A training pair of the network: r = A e and <e, A e> (device float64 scalar tensor)."""
        r = self.zeros_cell() if residual is None else residual
        en = self._zeros((1,), dtype=self.torch.float64) if energy is None else energy
        self._check(self.lib.ds_warm_start_pair(self.h, _ptr(error), _ptr(r), _ptr(en)), "ds_warm_start_pair")
        return r, en

    def profile_enable(self, sample_every: int):
        """Bracket every `sample_every`-th launch of the PCG iteration kernels with HIP events."""
        self._check(self.lib.ds_profile_enable(self.h, sample_every), "ds_profile_enable")

    def profile_read(self):
        """{kernel name: (launches, sampled launches, average ms per sampled launch)}."""
        out = (abi.KernelTiming * 8)()
        n = C.c_uint32()
        self._check(self.lib.ds_profile_read(self.h, out, 8, C.byref(n)), "ds_profile_read")
        return {out[i].name.decode(): (out[i].launches, out[i].samples,
                                       out[i].total_ms / out[i].samples if out[i].samples else float("nan"))
                for i in range(n.value)}

    def stream_copy(self, src, dst):
        """dst = src (float32 device tensors of equal size) by 4-byte accesses per lane."""
        assert src.numel() == dst.numel() and src.dtype == dst.dtype
        self._check(self.lib.ds_stream_copy(self.h, _ptr(src), _ptr(dst), src.numel()), "ds_stream_copy")

    def apply_matrix(self, x):
        y = self.zeros_cell()
        self._check(self.lib.ds_apply_pressure_matrix(self.h, _ptr(x), _ptr(y)), "ds_apply_pressure_matrix")
        return y
