"""Failure paths of ds_context_create that need no GPU: an invalid configuration is rejected with
DS_ERROR_INVALID_ARGUMENT, a message in ds_last_error() and no context, before the library touches
the device. The reference's Model constructor (solvers/smoke/model.cpp:30) takes its configuration
unchecked; the checks are this library's (the error contract of include/ds/ds_smoke.h). Also runs
the FP64 oracle under AddressSanitizer and UndefinedBehaviorSanitizer (oracle/sanitize_check.cpp)."""
import ctypes as C
import math
import os
import shutil
import subprocess

import pytest

from ds_smoke import abi, scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(abi.library_path()):
        pytest.skip("HIP library not built")
    from ds_smoke.solver import load_library

    return load_library()


def _create(lib, cfg, mode=abi.MODE_DIFFERENTIABLE):
    h = C.c_void_p(12345)   # must come back null
    st = lib.ds_context_create(C.byref(cfg) if cfg is not None else None, mode, None, C.byref(h))
    return st, h.value, lib.ds_last_error().decode()


def _plume():
    return scenes.plume(8).config()


def _with_colliders(cfg, cols):
    arr = (abi.Collider * len(cols))()
    for i, c in enumerate(cols):
        arr[i] = c
    cfg._collider_keepalive = arr
    cfg.colliders = C.cast(arr, C.POINTER(abi.Collider))
    cfg.collider_count = len(cols)
    return cfg


def _collider(shape, size):
    c = abi.Collider()
    c.shape = shape
    c.center[:] = [0.5, 0.5, 0.5]
    c.size[:] = list(size)
    return c


def _mutate(field, value):
    def f():
        cfg = _plume()
        if field.startswith("resolution"):
            cfg.resolution[int(field[-1])] = value
        elif field.endswith("_mode"):
            kind, face = field.split(".")[0], int(field.split(".")[1].split("_")[0])
            getattr(cfg, kind).modes[face] = value
        else:
            setattr(cfg, field, value)
        return cfg
    return f


CASES = {
    "zero resolution x": (_mutate("resolution0", 0), "resolution"),
    "zero resolution y": (_mutate("resolution1", 0), "resolution"),
    "zero resolution z": (_mutate("resolution2", 0), "resolution"),
    "resolution above 1024": (_mutate("resolution0", 1025), "resolution"),
    "zero cell size": (_mutate("cell_size", 0.0), "cell_size"),
    "negative cell size": (_mutate("cell_size", -0.1), "cell_size"),
    "NaN cell size": (_mutate("cell_size", math.nan), "cell_size"),
    "infinite cell size": (_mutate("cell_size", math.inf), "cell_size"),
    "zero time step": (_mutate("time_step", 0.0), "time_step"),
    "negative time step": (_mutate("time_step", -0.01), "time_step"),
    "NaN time step": (_mutate("time_step", math.nan), "time_step"),
    "velocity mode out of range": (_mutate("velocity_boundary.2_mode", 4), "velocity boundary mode"),
    "pressure mode out of range": (_mutate("pressure_boundary.5_mode", 3), "scalar boundary mode"),
    "density mode out of range": (_mutate("density_boundary.0_mode", 7), "scalar boundary mode"),
    "colliders null with a count": (lambda: (lambda c: (setattr(c, "colliders", None), setattr(c, "collider_count", 2), c)[-1])(_plume()),
                                    "colliders is null"),
    "collider shape out of range": (lambda: _with_colliders(_plume(), [_collider(2, (0.1, 0.1, 0.1))]), "shape"),
    "ellipsoid with a zero radius": (lambda: _with_colliders(_plume(), [_collider(abi.SHAPE_ELLIPSOID, (0.1, 0.0, 0.1))]), "radii"),
    "box with a negative half extent": (lambda: _with_colliders(_plume(), [_collider(abi.SHAPE_BOX, (0.1, -0.1, 0.1))]),
                                        "half extents"),
    "collider size not finite": (lambda: _with_colliders(_plume(), [_collider(abi.SHAPE_BOX, (0.1, math.nan, 0.1))]), "finite"),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_invalid_config_is_rejected_before_the_device(lib, case):
    make, words = CASES[case]
    st, h, msg = _create(lib, make())
    assert st == abi.DS_ERROR_INVALID_ARGUMENT, (case, st, msg)
    assert h is None, case
    assert words in msg, (case, msg)


def test_invalid_mode_and_null_arguments(lib):
    st, h, msg = _create(lib, _plume(), mode=2)
    assert st == abi.DS_ERROR_INVALID_ARGUMENT and h is None and "mode" in msg
    st, h, msg = _create(lib, None)
    assert st == abi.DS_ERROR_INVALID_ARGUMENT and h is None and "null" in msg
    assert lib.ds_context_create(C.byref(_plume()), abi.MODE_FORWARD, None, None) == abi.DS_ERROR_INVALID_ARGUMENT


def test_oracle_is_clean_under_address_and_undefined_behaviour_sanitizers():
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    build = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], capture_output=True, text=True, timeout=600)
    if build.returncode != 0 and ("asan" in build.stderr.lower() or "ubsan" in build.stderr.lower()):
        pytest.skip("sanitizer runtimes not available: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="halt_on_error=1:detect_leaks=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    run = subprocess.run([os.path.join(ROOT, "oracle", "sanitize_check")], capture_output=True, text=True, timeout=600, env=env)
    print(run.stdout)
    assert run.returncode == 0, run.stdout[-1000:] + run.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in run.stderr and "runtime error" not in run.stderr, run.stderr[-3000:]
    assert run.stdout.count("forward + VJP ok") == 3
