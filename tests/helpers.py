"""Shared scene builders and comparison helpers for the oracle and GPU parity tests."""
import numpy as np

from ds_smoke import abi, scenes
from ds_smoke.scenes import cell_shape, face_shape


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (d if d > 0 else 1.0)


def collider(shape, center, size, velocity=(0.0, 0.0, 0.0), density=0.0, temperature=0.0):
    c = abi.Collider()
    c.shape = shape
    c.center[:] = list(center)
    c.size[:] = list(size)
    c.velocity[:] = list(velocity)
    c.density = density
    c.temperature = temperature
    return c


def scene_closed(n=16):
    """Closed no-slip box plume (the benchmark set-up at a small size)."""
    return scenes.plume(n)


def scene_solids(res=(16, 12, 10)):
    """Non-cubic grid with a moving box collider and a sphere, mixed velocity boundaries and a
    fixed-value pressure face (no anchor)."""
    h = 1.0 / 16.0
    sc = scenes.Scene(resolution=res, cell_size=h, time_step=0.5 * h, name="solids")
    sc.colliders = [collider(abi.SHAPE_BOX, (0.3, 0.25, 0.3), (0.08, 0.1, 0.12), velocity=(0.2, -0.1, 0.05),
                             density=0.5, temperature=0.25),
                    collider(abi.SHAPE_ELLIPSOID, (0.7, 0.45, 0.35), (0.12, 0.1, 0.14))]
    sc.velocity_modes = (2, 2, 2, 1, 0, 0)       # normal-fixed x/y_min, zero-gradient y_max, fixed z
    vals = [0.0] * 18
    vals[3 * 4 + 0] = 0.05                        # z_min face carries a tangential velocity
    sc.velocity_values = tuple(vals)
    sc.pressure_modes = (1, 1, 1, 0, 1, 1)        # fixed-value pressure on y_max
    sc.pressure_values = (0.0, 0.0, 0.0, 0.1, 0.0, 0.0)
    sc.density_modes = (1, 1, 1, 0, 1, 1)
    sc.density_values = (0.0, 0.0, 0.0, 0.3, 0.0, 0.0)
    sc.temperature_modes = (0, 1, 1, 1, 1, 1)
    sc.temperature_values = (0.2, 0.0, 0.0, 0.0, 0.0, 0.0)
    sc.params = (0.1, -0.1, 1.0, 0.0)
    return sc


def scene_periodic(res=(12, 10, 8)):
    """Periodic in x (velocity and scalars), zero-gradient pressure."""
    h = 1.0 / 12.0
    sc = scenes.Scene(resolution=res, cell_size=h, time_step=0.5 * h, name="periodic")
    sc.velocity_modes = (3, 3, 0, 0, 0, 0)
    sc.pressure_modes = (2, 2, 1, 1, 1, 1)
    sc.density_modes = (2, 2, 1, 1, 1, 1)
    sc.temperature_modes = (2, 2, 1, 1, 1, 1)
    sc.colliders = [collider(abi.SHAPE_ELLIPSOID, (0.5, 0.4, 0.35), (0.15, 0.12, 0.1))]
    return sc


def scene_periodic_xy(res=(16, 12, 8)):
    """Periodic in x and y with a power-of-two cell size (the velocity reverse gathers its fixed
    start stencils there, and the periodic seams take the scatter path), fixed z faces."""
    h = 1.0 / 16.0
    sc = scenes.Scene(resolution=res, cell_size=h, time_step=0.5 * h, name="periodic_xy")
    sc.velocity_modes = (3, 3, 3, 3, 0, 2)
    sc.pressure_modes = (2, 2, 2, 2, 1, 1)
    sc.density_modes = (2, 2, 2, 2, 1, 1)
    sc.temperature_modes = (2, 2, 2, 2, 1, 1)
    sc.colliders = [collider(abi.SHAPE_BOX, (0.45, 0.35, 0.25), (0.12, 0.1, 0.08))]
    return sc


SCENES = {"closed": scene_closed, "solids": scene_solids, "periodic": scene_periodic, "periodic_xy": scene_periodic_xy}


def random_fields(sc, seed=0, vel_scale=0.6):
    """Seeded random inputs of every kind the operators read (FP32 values)."""
    rng = np.random.default_rng(seed)
    n = sc.resolution
    cs = cell_shape(n)

    def cell(scale=1.0, offset=0.0):
        return (offset + scale * rng.standard_normal(cs)).astype(np.float32)

    vel = tuple((vel_scale * rng.standard_normal(face_shape(n, a))).astype(np.float32) for a in range(3))
    return {
        "rho": np.abs(cell()), "T": cell(), "s_rho": cell(), "s_T": cell(),
        "e": (cell(0.3), cell(0.3), cell(0.3)), "vel": vel,
        "cell_adj": cell(), "cell_adj2": cell(),
        "face_adj": tuple(rng.standard_normal(face_shape(n, a)).astype(np.float32) for a in range(3)),
        "force": (cell(), cell(), cell()),
    }


def plume_state(sc):
    return scenes.initial_state(sc), scenes.control(sc)
