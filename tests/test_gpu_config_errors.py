"""Failure paths that reach the device: a domain whose every cell is solid is rejected at context
creation (DS_ERROR_INVALID_ARGUMENT, no context), and an allocation the device cannot hold comes
back as DS_ERROR_OUT_OF_MEMORY with the context still usable (the next operator runs and checks its
launch normally). The reference allocates through cudaMallocAsync (cuda.cpp:39-43) and reports a
failed call through its error check (operators.cu:28-30)."""
import ctypes as C

import numpy as np
import pytest

from ds_smoke import abi, scenes

pytestmark = pytest.mark.gpu


def test_all_solid_domain_is_rejected():
    from ds_smoke.solver import Context, SolverError

    print("all-solid 8^3 box", flush=True)
    sc = scenes.Scene(resolution=(8, 8, 8), cell_size=1.0 / 8.0, time_step=1.0 / 16.0, name="all_solid")
    box = abi.Collider()
    box.shape = abi.SHAPE_BOX
    box.center[:] = [0.5, 0.5, 0.5]
    box.size[:] = [1.0, 1.0, 1.0]
    sc.colliders = [box]
    with pytest.raises(SolverError) as e:
        Context(sc.config(), abi.MODE_DIFFERENTIABLE)
    assert e.value.status == abi.DS_ERROR_INVALID_ARGUMENT
    assert "every cell is solid" in str(e.value)


def test_out_of_memory_is_a_status_and_the_context_stays_usable():
    from ds_smoke.solver import Context

    print("plume 16^3, a 1 PiB allocation", flush=True)
    sc = scenes.plume(16)
    ctx = Context(sc.config(), abi.MODE_DIFFERENTIABLE)
    p = C.c_void_p()
    st = ctx.lib.ds_allocate(ctx.h, C.c_size_t(1 << 50), C.byref(p))
    msg = ctx.lib.ds_last_error().decode()
    assert st == abi.DS_ERROR_OUT_OF_MEMORY, (st, msg)
    assert "hipMallocAsync" in msg, msg
    assert not p.value
    # the failed call leaves nothing behind: an allocation and an operator with its launch check
    q = C.c_void_p()
    assert ctx.lib.ds_allocate(ctx.h, C.c_size_t(4096), C.byref(q)) == abi.DS_OK
    assert ctx.lib.ds_release(ctx.h, q) == abi.DS_OK
    _, code = ctx.masks()
    x = np.where(code > 0, 1.0, 0.0).astype(np.float32)
    y = ctx.to_host(ctx.apply_matrix(ctx.to_device(x)))
    assert np.all(np.isfinite(y))
    ctx.close()
