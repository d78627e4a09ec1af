"""ctypes mirror of include/ds/ds_smoke.h (the C ABI of the HIP smoke solver).

The structs follow the header field by field; the header cites the reference interfaces
(solvers/smoke/data.ixx, cuda_types.h, operators.h, model.ixx, solver.ixx) each one replaces.
"""
from __future__ import annotations

import ctypes as C
import os

DS_OK = 0
(DS_ERROR_INVALID_ARGUMENT, DS_ERROR_OUT_OF_MEMORY, DS_ERROR_LAUNCH, DS_ERROR_NOT_CONVERGED, DS_ERROR_NON_FINITE,
 DS_ERROR_RUNTIME, DS_ERROR_UNSUPPORTED) = range(1, 8)
STATUS_NAMES = {
    0: "ok", 1: "invalid argument", 2: "out of memory", 3: "launch failure",
    4: "not converged", 5: "non-finite", 6: "runtime error", 7: "unsupported",
}

SCALAR_FIXED_VALUE, SCALAR_ZERO_GRADIENT, SCALAR_PERIODIC = 0, 1, 2
VELOCITY_FIXED_VALUE, VELOCITY_ZERO_GRADIENT, VELOCITY_NORMAL_FIXED, VELOCITY_PERIODIC = 0, 1, 2, 3
SHAPE_ELLIPSOID, SHAPE_BOX = 0, 1
PROJECTION_PCG, PROJECTION_RBGS = 0, 1
WARM_START_NONE, WARM_START_PREVIOUS, WARM_START_CNN, WARM_START_EXTRAPOLATE, WARM_START_PROJECTION, WARM_START_CNN_ALONE = 0, 1, 2, 3, 4, 5
MODE_FORWARD, MODE_DIFFERENTIABLE = 0, 1
CG_SINGLE_REDUCTION, CG_TWO_REDUCTION, CG_SINGLE_REDUCTION_RESIDENT, CG_S_STEP, CG_MULTIGRID = 0, 1, 2, 3, 4

FloatP = C.POINTER(C.c_float)
DoubleP = C.POINTER(C.c_double)


class ScalarBoundary(C.Structure):
    _fields_ = [("modes", C.c_uint32 * 6), ("values", C.c_float * 6)]


class VelocityBoundary(C.Structure):
    _fields_ = [("modes", C.c_uint32 * 6), ("values", C.c_float * 18)]


class Collider(C.Structure):
    _fields_ = [("shape", C.c_uint32), ("center", C.c_float * 3), ("size", C.c_float * 3),
                ("velocity", C.c_float * 3), ("density", C.c_float), ("temperature", C.c_float)]


class Config(C.Structure):
    _fields_ = [
        ("resolution", C.c_uint32 * 3),
        ("cell_size", C.c_float),
        ("time_step", C.c_float),
        ("pressure_iterations", C.c_uint32),
        ("velocity_boundary", VelocityBoundary),
        ("pressure_boundary", ScalarBoundary),
        ("density_boundary", ScalarBoundary),
        ("temperature_boundary", ScalarBoundary),
        ("colliders", C.POINTER(Collider)),
        ("collider_count", C.c_uint32),
        ("vorticity_confinement_enabled", C.c_uint32),
        ("projection_solver", C.c_uint32),
        ("cg_max_iterations", C.c_uint32),
        ("cg_tolerance", C.c_double),
        ("warm_start", C.c_uint32),
        ("residual_replacement_interval", C.c_uint32),
        ("cg_variant", C.c_uint32),
    ]


class State(C.Structure):
    _fields_ = [("density", C.c_void_p), ("temperature", C.c_void_p), ("velocity", C.c_void_p * 3)]


class Control(C.Structure):
    _fields_ = [("density_source", C.c_void_p), ("temperature_source", C.c_void_p),
                ("external_acceleration", C.c_void_p * 3)]


class Parameters(C.Structure):
    _fields_ = [("ambient_temperature", C.c_void_p), ("density_buoyancy", C.c_void_p),
                ("temperature_buoyancy", C.c_void_p), ("vorticity_confinement", C.c_void_p)]


class StepCache(C.Structure):
    _fields_ = [("sourced_density", C.c_void_p), ("sourced_temperature", C.c_void_p),
                ("force", C.c_void_p * 3), ("forced_velocity", C.c_void_p * 3),
                ("advected_velocity", C.c_void_p * 3), ("pressure", C.c_void_p), ("pressure_guess", C.c_void_p)]


StateAdjoint = State
ControlAdjoint = Control
ParameterAdjoint = Parameters


class SolveInfo(C.Structure):
    _fields_ = [("iterations", C.c_uint32), ("converged", C.c_uint32), ("warm_started", C.c_uint32),
                ("guess_rejected", C.c_uint32), ("rhs_norm", C.c_double),
                ("initial_residual", C.c_double), ("final_residual", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class StepInfo(C.Structure):
    _fields_ = [("projection", SolveInfo)]


class TrajectoryInfo(C.Structure):
    _fields_ = [("steps", C.c_uint32), ("checkpoint_interval", C.c_uint32),
                ("forward_solves", C.c_uint32), ("recompute_solves", C.c_uint32),
                ("adjoint_solves", C.c_uint32), ("not_converged", C.c_uint32),
                ("forward_iterations", C.c_uint64), ("recompute_iterations", C.c_uint64),
                ("adjoint_iterations", C.c_uint64), ("max_iterations", C.c_uint32),
                ("guesses_rejected", C.c_uint32), ("max_relative_residual", C.c_double),
                ("loss", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class JetControl(C.Structure):
    """ds_jet_control: the double-jet control of the forward example (simulation.h:10)."""
    _fields_ = [("pulse_period", C.c_float), ("left_source_center", C.c_float * 3), ("right_source_center", C.c_float * 3),
                ("source_radius", C.c_float), ("density_source_rate", C.c_float), ("temperature_source_rate", C.c_float),
                ("left_acceleration", C.c_float * 3), ("right_acceleration", C.c_float * 3)]

    def values(self):
        """The 16 values in field order."""
        return ([self.pulse_period] + list(self.left_source_center) + list(self.right_source_center)
                + [self.source_radius, self.density_source_rate, self.temperature_source_rate]
                + list(self.left_acceleration) + list(self.right_acceleration))


class KernelTiming(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("samples", C.c_uint64), ("total_ms", C.c_double)]


class DomainInfo(C.Structure):
    _fields_ = [("cell_count", C.c_uint64), ("face_count", C.c_uint64 * 3),
                ("unknown_count", C.c_uint64), ("solid_count", C.c_uint64),
                ("pressure_anchor", C.c_uint64), ("extra_anchor_count", C.c_uint32),
                ("has_fixed_pressure", C.c_uint32)]


def package_dir() -> str:
    return os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def library_path() -> str:
    """Path of the in-tree HIP library built by __graft_entry__.build()."""
    return os.path.join(package_dir(), "lib", "libds_smoke.so")
