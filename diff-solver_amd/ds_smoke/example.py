"""The reference's forward example (examples/smoke/forward/): a pulsed double-jet thermal plume
with vorticity confinement in a box open at the top, stepped by ForwardSimulation with per-step
metrics (density mass and maxima, CFL, divergence RMS before and after the projection).

  python -m ds_smoke.example --steps 960 --report 60 [--solver rbgs|pcg] [--resolution 128 192 128]

ForwardSimulationOptions / ForwardSimulationMetrics / ForwardSimulation mirror
simulation.ixx:9-68 and simulation.cpp:61-120; the control writer and the metric reduction are the
HIP kernels behind ds_example_write_control and ds_example_reduce_metrics. The C++ twin is
cpp/xayah/forward_simulation.hpp with the program cpp/examples/smoke_forward.cpp.
"""
from __future__ import annotations

import argparse
import dataclasses
import sys
import time

import numpy as np

from . import abi, scenes

F32 = np.float32


@dataclasses.dataclass
class ForwardSimulationOptions:
    """simulation.ixx:9-27, same fields and defaults (FP32 values as the reference stores them),
    plus the pressure solver: the reference's red-black Gauss-Seidel with pressure_iterations
    sweeps by default, or PCG to cg_tolerance warm-started from the previous step."""
    resolution: tuple = (128, 192, 128)
    cell_size: float = float(F32(1.0 / 128.0))
    time_step: float = float(F32(1.0 / 120.0))
    pressure_iterations: int = 160
    density_source_rate: float = 4.0
    temperature_source_rate: float = 8.0
    ambient_temperature: float = 0.0
    density_buoyancy: float = float(F32(-0.1))
    temperature_buoyancy: float = 1.0
    vorticity_confinement: float = 2.0
    left_source_center: tuple = (float(F32(0.25)), float(F32(0.10)), float(F32(0.36)))
    right_source_center: tuple = (float(F32(0.75)), float(F32(0.10)), float(F32(0.64)))
    source_radius: float = float(F32(0.055))
    left_acceleration: tuple = (3.5, 5.0, float(F32(1.8)))
    right_acceleration: tuple = (-3.5, 5.0, float(F32(-1.8)))
    pulse_period: float = float(F32(0.9))
    host_metrics: bool = True
    projection_solver: int = abi.PROJECTION_RBGS
    cg_tolerance: float = 1e-6
    warm_start: int = abi.WARM_START_PREVIOUS

    def jet(self) -> abi.JetControl:
        j = abi.JetControl()
        j.pulse_period = self.pulse_period
        j.left_source_center[:] = list(self.left_source_center)
        j.right_source_center[:] = list(self.right_source_center)
        j.source_radius = self.source_radius
        j.density_source_rate = self.density_source_rate
        j.temperature_source_rate = self.temperature_source_rate
        j.left_acceleration[:] = list(self.left_acceleration)
        j.right_acceleration[:] = list(self.right_acceleration)
        return j

    def scene(self) -> scenes.Scene:
        """make_configuration (simulation.cpp:17-35): normal-fixed / tangent-free walls on five
        sides, an open top (zero-gradient velocity; pressure, density and temperature fixed at 0
        on y_max), vorticity confinement on."""
        sc = scenes.Scene(resolution=tuple(int(v) for v in self.resolution), cell_size=self.cell_size, time_step=self.time_step,
                          name="double_jet")
        sc.velocity_modes = (2, 2, 2, 1, 2, 2)
        sc.pressure_modes = (1, 1, 1, 0, 1, 1)
        sc.density_modes = (1, 1, 1, 0, 1, 1)
        sc.temperature_modes = (1, 1, 1, 0, 1, 1)
        sc.vorticity = True
        sc.params = (self.ambient_temperature, self.density_buoyancy, self.temperature_buoyancy, self.vorticity_confinement)
        return sc

    def config(self) -> abi.Config:
        warm = self.warm_start if self.projection_solver == abi.PROJECTION_PCG else abi.WARM_START_NONE
        return self.scene().config(tolerance=self.cg_tolerance, solver=self.projection_solver,
                                   rbgs_iterations=self.pressure_iterations, warm_start=warm)


@dataclasses.dataclass
class ForwardSimulationMetrics:
    """simulation.ixx:29-42, plus the last pressure solve's iteration count."""
    step: int = 0
    physical_time: float = 0.0
    density_mass: float = 0.0
    density_maximum: float = 0.0
    temperature_maximum: float = 0.0
    maximum_velocity: float = 0.0
    cfl: float = 0.0
    pre_projection_divergence_rms: float = 0.0
    post_projection_divergence_rms: float = 0.0
    divergence_ratio: float = 0.0
    step_milliseconds: float = 0.0
    average_step_milliseconds: float = 0.0
    iterations: int = 0


def finish_metrics(options: ForwardSimulationOptions, values, step_ms: float, m: ForwardSimulationMetrics) -> None:
    """The derived metrics of one step from the six raw values (simulation.cpp:104-119); FP64
    with the FP32 options promoted, and IEEE division (a zero pre-projection RMS gives nan/inf)."""
    v = [np.float64(x) for x in values]
    h = np.float64(F32(options.cell_size))
    previous = np.float64(m.step - 1)
    m.density_mass = float(v[0] * h * h * h)
    m.density_maximum = float(v[1])
    m.temperature_maximum = float(v[2])
    m.maximum_velocity = float(v[3])
    m.cfl = float(v[3] * np.float64(F32(options.time_step)) / h)
    cells = np.float64(options.resolution[0]) * options.resolution[1] * options.resolution[2]
    m.pre_projection_divergence_rms = float(np.sqrt(v[4] / cells))
    m.post_projection_divergence_rms = float(np.sqrt(v[5] / cells))
    with np.errstate(divide="ignore", invalid="ignore"):
        m.divergence_ratio = float(np.float64(m.post_projection_divergence_rms) / np.float64(m.pre_projection_divergence_rms))
    m.step_milliseconds = step_ms
    m.average_step_milliseconds = float((previous * m.average_step_milliseconds + step_ms) / np.float64(m.step))


class ForwardSimulation:
    """simulation.ixx:44-68 / simulation.cpp:61-120 over one forward-mode context."""

    def __init__(self, options: ForwardSimulationOptions | None = None, device: int = 0):
        from .solver import Context

        self.options = options or ForwardSimulationOptions()
        self.context = Context(self.options.config(), abi.MODE_FORWARD, device)
        ctx = self.context
        self.current_state = ctx.allocate_state()
        self.metrics = ForwardSimulationMetrics()
        self.device_metrics = ctx._zeros((6,), dtype=ctx.torch.float64)
        self._next = ctx.allocate_state()
        self.control = ctx.allocate_control()
        o = self.options
        self._params = ctx.allocate_parameters((o.ambient_temperature, o.density_buoyancy, o.temperature_buoyancy, o.vorticity_confinement))
        self._cache = ctx.allocate_step_cache(pressure=False)
        self._jet = o.jet()
        self.reset()

    def reset(self) -> None:
        """simulation.cpp:74-88: zero both states and the metrics."""
        ctx = self.context
        with ctx.torch.cuda.stream(ctx.stream):
            for s in (self.current_state, self._next):
                s["density"].zero_()
                s["temperature"].zero_()
                for v in s["velocity"]:
                    v.zero_()
            self.device_metrics.zero_()
        ctx.synchronize()
        self.metrics = ForwardSimulationMetrics()

    def step(self) -> None:
        """simulation.cpp:90-120: the jets' control for this step's time, one forward step, the
        metrics over the fluid cells, then swap the states."""
        ctx = self.context
        begin = time.perf_counter() if self.options.host_metrics else 0.0
        ctx.example_write_control(self.metrics.step, self._jet, self.control)
        _, _, info = ctx.forward_step(self.current_state, self.control, self._params, next_state=self._next, cache=self._cache)
        ctx.example_reduce_metrics(self._next["density"], self._next["temperature"], self._cache["advected_velocity"],
                                   self._next["velocity"], out=self.device_metrics)
        values = None
        if self.options.host_metrics:
            values = ctx.to_host(self.device_metrics)
        self.current_state, self._next = self._next, self.current_state
        m = self.metrics
        m.step += 1
        m.physical_time = float(np.float64(m.step) * np.float64(F32(self.options.time_step)))
        m.iterations = int(info.iterations)
        if not self.options.host_metrics:
            return
        finish_metrics(self.options, values, 1e3 * (time.perf_counter() - begin), m)

    def close(self) -> None:
        self.context.close()


def format_metrics(m: ForwardSimulationMetrics) -> str:
    """The line main.cpp:19-22 prints."""
    return (f"step {m.step:3d}  time {m.physical_time:.3f} s  density mass {m.density_mass:.6f} max {m.density_maximum:.5f}  "
            f"temperature max {m.temperature_maximum:.5f}  velocity max {m.maximum_velocity:.5f}  CFL {m.cfl:.4f}  "
            f"divergence {m.pre_projection_divergence_rms:.4e} -> {m.post_projection_divergence_rms:.4e} ({m.divergence_ratio:.4f})  "
            f"step {m.step_milliseconds:.3f} ms  average {m.average_step_milliseconds:.3f} ms")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--steps", type=int, default=960)
    ap.add_argument("--report", type=int, default=60)
    ap.add_argument("--solver", choices=["rbgs", "pcg"], default="rbgs")
    ap.add_argument("--resolution", type=int, nargs=3, default=(128, 192, 128))
    args = ap.parse_args(argv)
    import torch

    opts = ForwardSimulationOptions(resolution=tuple(args.resolution),
                                    projection_solver=abi.PROJECTION_RBGS if args.solver == "rbgs" else abi.PROJECTION_PCG)
    free0 = torch.cuda.mem_get_info()[0]
    sim = ForwardSimulation(opts)
    free1 = torch.cuda.mem_get_info()[0]
    print(f"{opts.resolution[0]} x {opts.resolution[1]} x {opts.resolution[2]} double-jet thermal smoke ({args.solver})", flush=True)
    print(f"GPU allocation: {(free0 - free1) / 2 ** 20:.1f} MiB", flush=True)
    for _ in range(args.steps):
        sim.step()
        if sim.metrics.step % max(1, args.report) == 0:
            print(format_metrics(sim.metrics), flush=True)
    sim.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
