"""Synthetic plume scenes: configurations, seeded initial conditions, controls and targets.

A scene is the reference's smoke set-up (solvers/smoke/data.ixx:72-83, the example's
parameter convention at examples/smoke/forward/simulation.ixx:9-27) with the inputs the
benchmark workloads name:

* a closed no-slip box (velocity fixed_value 0 on all six faces; pressure, density and
  temperature zero_gradient, so the pressure anchor is cell 0, model.cpp:72-77);
* h = 1/n, dt = h/2; buoyancy alpha = 0.1 -> density_buoyancy = -0.1, beta = 1.0 ->
  temperature_buoyancy = +1.0 (buoyancy acts on +y, operators.cu:311);
* a source sphere centred at (0.5, 0.125, 0.5) of radius 0.09375 in domain units (6 cells at
  64^3) injecting density = temperature = 1 per step: the reference source is additive
  (rho + dt*s, operators.cu:289), so s = indicator / dt; membership tests cell centres;
* initial velocity: every face DOF uniform in (-0.01, 0.01) from a splitmix64 stream;
* a target density made of seeded isotropic Gaussian blobs in the upper half of the box;
* optional seeded solid spheres (colliders of ellipsoid shape with equal radii).

All randomness is the project's own splitmix64 -> U[0,1) so the GPU path and the oracle get
bit-identical inputs.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64_uniform(seed: int, count: int, offset: int = 0) -> np.ndarray:
    """U[0,1) doubles: element i is splitmix64 output number (offset + i + 1) of `seed`."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed) + _GOLDEN * (np.arange(offset + 1, offset + count + 1, dtype=np.uint64))
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


class Rng:
    """Sequential splitmix64 stream."""

    def __init__(self, seed: int):
        self.seed = int(seed)
        self.offset = 0

    def uniform(self, count: int, low: float = 0.0, high: float = 1.0) -> np.ndarray:
        u = splitmix64_uniform(self.seed, count, self.offset)
        self.offset += count
        return low + (high - low) * u


def face_shape(n, axis):
    """(z, y, x) numpy shape of an axis face array for resolution n = (nx, ny, nz)."""
    ex = [n[0], n[1], n[2]]
    ex[axis] += 1
    return (ex[2], ex[1], ex[0])


def cell_shape(n):
    return (n[2], n[1], n[0])


@dataclass
class Scene:
    """A configuration plus its host-side inputs (numpy, float32, reference layout)."""

    resolution: tuple
    cell_size: float
    time_step: float
    colliders: list = field(default_factory=list)
    params: tuple = (0.0, -0.1, 1.0, 0.0)   # ambient, density_buoyancy, temperature_buoyancy, confinement
    density_source: np.ndarray | None = None
    temperature_source: np.ndarray | None = None
    initial_velocity: tuple | None = None
    target: np.ndarray | None = None
    velocity_modes: tuple = (0, 0, 0, 0, 0, 0)
    velocity_values: tuple = (0.0,) * 18
    pressure_modes: tuple = (1, 1, 1, 1, 1, 1)
    pressure_values: tuple = (0.0,) * 6
    density_modes: tuple = (1, 1, 1, 1, 1, 1)
    density_values: tuple = (0.0,) * 6
    temperature_modes: tuple = (1, 1, 1, 1, 1, 1)
    temperature_values: tuple = (0.0,) * 6
    name: str = "plume"
    vorticity: bool = False                 # vorticity_confinement_enabled; strength = params[3]

    @property
    def cell_count(self) -> int:
        n = self.resolution
        return n[0] * n[1] * n[2]

    def face_count(self, axis: int) -> int:
        return int(np.prod(face_shape(self.resolution, axis)))

    def config(self, tolerance: float = 1e-6, max_iterations: int = 0, solver: int = abi.PROJECTION_PCG,
               warm_start: int = abi.WARM_START_NONE, rbgs_iterations: int = 80,
               replacement_interval: int = 0, cg_variant: int = abi.CG_MULTIGRID) -> abi.Config:
        cfg = abi.Config()
        cfg.resolution[:] = list(self.resolution)
        cfg.cell_size = self.cell_size
        cfg.time_step = self.time_step
        cfg.pressure_iterations = rbgs_iterations
        cfg.velocity_boundary.modes[:] = list(self.velocity_modes)
        cfg.velocity_boundary.values[:] = list(self.velocity_values)
        cfg.pressure_boundary.modes[:] = list(self.pressure_modes)
        cfg.pressure_boundary.values[:] = list(self.pressure_values)
        cfg.density_boundary.modes[:] = list(self.density_modes)
        cfg.density_boundary.values[:] = list(self.density_values)
        cfg.temperature_boundary.modes[:] = list(self.temperature_modes)
        cfg.temperature_boundary.values[:] = list(self.temperature_values)
        arr = (abi.Collider * max(1, len(self.colliders)))()
        for i, c in enumerate(self.colliders):
            arr[i] = c
        cfg._collider_keepalive = arr  # keep the array alive with the struct
        cfg.colliders = C.cast(arr, C.POINTER(abi.Collider))
        cfg.collider_count = len(self.colliders)
        cfg.vorticity_confinement_enabled = 1 if self.vorticity else 0
        cfg.projection_solver = solver
        cfg.cg_max_iterations = max_iterations
        cfg.cg_tolerance = tolerance
        cfg.warm_start = warm_start
        cfg.residual_replacement_interval = replacement_interval
        cfg.cg_variant = cg_variant
        return cfg


def cell_centres(n, h):
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    return (x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h


def sphere_indicator(n, h, centre, radius) -> np.ndarray:
    px, py, pz = cell_centres(n, h)
    d2 = (px - centre[0]) ** 2 + (py - centre[1]) ** 2 + (pz - centre[2]) ** 2
    return (d2 <= radius * radius).astype(np.float32)


def gaussian_target(n, seed: int, blobs: int = 4, sigma_range=(6.0, 12.0)) -> np.ndarray:
    """Sum of seeded isotropic Gaussians (sigma in cells) in the upper half of the box."""
    rng = Rng(seed)
    nx, ny, nz = n
    z, y, x = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5, indexing="ij")
    t = np.zeros(cell_shape(n), dtype=np.float64)
    for _ in range(blobs):
        sigma = rng.uniform(1, *sigma_range)[0]
        cx = rng.uniform(1, sigma, nx - sigma)[0]
        cy = rng.uniform(1, ny / 2.0, ny - sigma)[0]
        cz = rng.uniform(1, sigma, nz - sigma)[0]
        t += np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2.0 * sigma * sigma))
    return t.astype(np.float32)


def random_spheres(n, seed: int, count: int, radius_range, avoid_centre, avoid_radius, h):
    """Seeded solid spheres (ellipsoid colliders with equal radii), rejected when they overlap
    the source region; radii in cells."""
    rng = Rng(seed)
    out = []
    tries = 0
    while len(out) < count and tries < 10000:
        tries += 1
        r = rng.uniform(1, *radius_range)[0]
        c = [rng.uniform(1, r, n[k] - r)[0] for k in range(3)]
        ac = [avoid_centre[k] / h for k in range(3)]
        if np.sqrt(sum((c[k] - ac[k]) ** 2 for k in range(3))) < r + avoid_radius / h + 4.0:
            continue
        col = abi.Collider()
        col.shape = abi.SHAPE_ELLIPSOID
        col.center[:] = [c[0] * h, c[1] * h, c[2] * h]
        col.size[:] = [r * h, r * h, r * h]
        col.velocity[:] = [0.0, 0.0, 0.0]
        col.density = 0.0
        col.temperature = 0.0
        out.append(col)
    return out


SOURCE_CENTRE = (0.5, 0.125, 0.5)
SOURCE_RADIUS = 6.0 / 64.0


def plume(n: int, noise_seed: int = 1, target_seed: int = 2, solids: int = 0, solid_seed: int = 4,
          noise_amplitude: float = 0.01, source_centre=SOURCE_CENTRE, source_radius=SOURCE_RADIUS,
          with_target: bool = True) -> Scene:
    """The benchmark plume at n^3 (64^3: config 1; 128^3: configs 2/3/5; 256^3 + 8 solids: 4)."""
    res = (n, n, n)
    h = np.float32(1.0 / n)
    dt = np.float32(h * np.float32(0.5))
    sc = Scene(resolution=res, cell_size=float(h), time_step=float(dt), name=f"plume{n}")
    ind = sphere_indicator(res, float(h), source_centre, source_radius)
    src = (ind / np.float32(dt)).astype(np.float32)
    sc.density_source = src
    sc.temperature_source = src.copy()
    rng = Rng(noise_seed)
    vel = []
    for a in range(3):
        shp = face_shape(res, a)
        vel.append(rng.uniform(int(np.prod(shp)), -noise_amplitude, noise_amplitude).astype(np.float32).reshape(shp))
    sc.initial_velocity = tuple(vel)
    if with_target:
        sc.target = gaussian_target(res, target_seed)
    if solids:
        # radius U[8, 20] cells at 256^3, scaled with the grid below that
        scale = min(1.0, n / 256.0)
        # placed against the canonical source, so plumes with a jittered source share the spheres
        sc.colliders = random_spheres(res, solid_seed, solids, (8.0 * scale, 20.0 * scale), SOURCE_CENTRE,
                                      source_radius, float(h))
    sc.name = f"plume{n}" + (f"_solids{solids}" if solids else "")
    return sc


def initial_state(sc: Scene):
    """Host arrays of the initial state: rho = T = 0 and the seeded noise velocity."""
    z = np.zeros(cell_shape(sc.resolution), dtype=np.float32)
    return {"density": z, "temperature": z.copy(), "velocity": tuple(v.copy() for v in sc.initial_velocity)}


def control(sc: Scene):
    z = np.zeros(cell_shape(sc.resolution), dtype=np.float32)
    return {"density_source": sc.density_source, "temperature_source": sc.temperature_source,
            "external_acceleration": (z, z.copy(), z.copy())}
