"""ds_example_reduce_metrics_static, the entry point that allocates nothing, against
ds_example_reduce_metrics (the context's reduction) through the C ABI.

Two calls in a row on one GPU, the second on a smaller grid: each result equals the context
entry point's bit for bit (the same kernels, the same fixed-order reduction; the static partials
array keeps the first call's contents, which the second call must overwrite). A call with a mask
and one without (NULL = no solids) are both covered, and an over-large grid is rejected.
"""
import ctypes as C

import numpy as np
import pytest

from helpers import collider

pytestmark = pytest.mark.gpu


def _lib():
    from ds_smoke import abi
    from ds_smoke.solver import load_library

    L = load_library()
    fn = L.ds_example_reduce_metrics_static
    fn.restype = C.c_int32
    fn.argtypes = [C.POINTER(abi.Config), C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                   C.c_void_p, C.c_void_p]
    return L, fn


def _scene(res, solids):
    from ds_smoke import abi, scenes

    h = 1.0 / res[0]
    sc = scenes.Scene(resolution=res, cell_size=h, time_step=0.5 * h, name=f"metrics{res[0]}")
    if solids:
        sc.colliders = [collider(abi.SHAPE_ELLIPSOID, (0.4, 0.5, 0.45), (0.15, 0.2, 0.12)),
                        collider(abi.SHAPE_BOX, (0.75, 0.3, 0.6), (0.1, 0.08, 0.12))]
    return sc


def _fields(ctx, sc, seed):
    from ds_smoke.scenes import cell_shape, face_shape

    rng = np.random.default_rng(seed)
    n = sc.resolution
    rho = ctx.to_device(np.abs(rng.standard_normal(cell_shape(n))).astype(np.float32))
    T = ctx.to_device(rng.standard_normal(cell_shape(n)).astype(np.float32))
    pre = tuple(ctx.to_device(rng.standard_normal(face_shape(n, a)).astype(np.float32)) for a in range(3))
    post = tuple(ctx.to_device((0.1 * rng.standard_normal(face_shape(n, a))).astype(np.float32)) for a in range(3))
    return rho, T, pre, post


def _static_call(fn, ctx, sc, mask, rho, T, pre, post):
    import torch

    cfg = sc.config(tolerance=1e-6)
    out = torch.full((6,), float("nan"), dtype=torch.float64, device=ctx.device)
    ptrs = lambda ts: (C.c_void_p * 3)(*[t.data_ptr() for t in ts])
    st = fn(C.byref(cfg), C.c_void_p(mask.data_ptr()) if mask is not None else None, C.c_void_p(rho.data_ptr()),
            C.c_void_p(T.data_ptr()), ptrs(pre), ptrs(post), C.c_void_p(out.data_ptr()), C.c_void_p(ctx.stream.cuda_stream))
    assert st == 0, ctx.lib.ds_last_error().decode()
    ctx.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("solids", [True, False], ids=["solids", "no_solids"])
def test_static_metrics_equal_the_context_entry_point(solids):
    import torch
    from ds_smoke import abi
    from ds_smoke.solver import Context

    _, fn = _lib()
    results = []
    for res, seed in (((48, 40, 36), 1), ((20, 16, 12), 2)):   # the second call is smaller
        print(f"static metrics: grid {res}, solids {solids}", flush=True)
        sc = _scene(res, solids)
        ctx = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD)
        rho, T, pre, post = _fields(ctx, sc, seed)
        ref = ctx.example_reduce_metrics(rho, T, pre, post).cpu().numpy()
        mask = None
        if solids:
            m, _ = ctx.masks()
            assert m.any()
            mask = torch.from_numpy(np.ascontiguousarray(m)).to(ctx.device)
        got = _static_call(fn, ctx, sc, mask, rho, T, pre, post)
        print(f"  context {ref}\n  static  {got}")
        assert np.all(np.isfinite(got))
        np.testing.assert_array_equal(got, ref)
        results.append(got)
        ctx.close()
    assert not np.array_equal(results[0], results[1])


def test_static_metrics_reject_a_grid_beyond_32_bit_indices():
    import torch
    from ds_smoke import abi
    from ds_smoke.solver import Context

    L, fn = _lib()
    sc = _scene((8, 8, 8), False)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD)
    rho, T, pre, post = _fields(ctx, sc, 3)
    big = sc.config(tolerance=1e-6)
    big.resolution[0] = big.resolution[1] = big.resolution[2] = 2048
    out = torch.zeros(6, dtype=torch.float64, device=ctx.device)
    ptrs = lambda ts: (C.c_void_p * 3)(*[t.data_ptr() for t in ts])
    st = fn(C.byref(big), None, C.c_void_p(rho.data_ptr()), C.c_void_p(T.data_ptr()), ptrs(pre), ptrs(post),
            C.c_void_p(out.data_ptr()), None)
    assert st == abi.DS_ERROR_INVALID_ARGUMENT
    assert b"too large" in L.ds_last_error()
    ctx.close()
