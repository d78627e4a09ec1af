"""The warm-start network (csrc/unet.hip: BF16 MFMA implicit-GEMM convolutions) against the same
network in PyTorch float32 (ds_smoke/warm_start.py) with the same seeded parameters (both guesses
scaled to their A-norm-optimal multiple), and the
DS_WARM_START_CNN solve path: the guess feeds the PCG, which still converges to 1e-6 and agrees
with the plain solve. BF16 activations through 13 layers: relative L2 tolerance 3e-2."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def _setup(n=64, seed=3):
    import torch
    from ds_smoke import abi, scenes, warm_start
    from ds_smoke.solver import Context

    torch.manual_seed(seed)
    model = warm_start.build_unet()
    flat = warm_start.flatten_parameters(model)
    sc = scenes.plume(n, with_target=False)
    ctx = Context(sc.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN), abi.MODE_FORWARD)
    ctx.warm_start_load(flat, 2.5)
    # a right-hand side of the solver: the projection of the plume's noisy initial velocity
    state = ctx.state_from_host(scenes.initial_state(sc))
    _, code = ctx.masks()
    rng = np.random.default_rng(seed)
    b = np.where(code > 0, rng.standard_normal(code.shape), 0).astype(np.float32)
    return sc, ctx, model, flat, code, b, state


def test_network_matches_pytorch_float32():
    import torch
    from ds_smoke import warm_start

    print("warm-start network 64^3 vs PyTorch", flush=True)
    sc, ctx, model, flat, code, b, _ = _setup()
    got = ctx.to_host(ctx.warm_start_infer(ctx.to_device(b)))
    with torch.no_grad():
        dev = ctx.device
        unknown = torch.from_numpy(code > 0).to(dev)
        bt = torch.from_numpy(b).to(dev)
        x, s = warm_start.network_input(bt, unknown)
        raw = (model.to(dev)(x)[0, 0] * float(s) * 2.5 * unknown.float()).contiguous()
    torch.cuda.synchronize()   # raw was produced on torch's stream; the library runs on its own
    # the library scales its guess g to the A-norm-optimal multiple <g, b> / <g, A g>
    ag = ctx.to_host(ctx.apply_matrix(raw)).astype(np.float64)
    raw = raw.cpu().numpy().astype(np.float64)
    ref = raw * (np.dot(raw.ravel(), b.astype(np.float64).ravel()) / np.dot(raw.ravel(), ag.ravel()))
    err = rel(got, ref)
    print(f"relative L2 {err:.3e}")
    assert np.all(got[code == 0] == 0)
    assert err < 3e-2, err


def test_cnn_warm_started_solve_converges_and_agrees():
    from ds_smoke import abi
    from ds_smoke.solver import Context

    sc, ctx, model, flat, code, b, _ = _setup()
    x, info, st = ctx.pcg_solve(ctx.to_device(b))
    assert st == 0 and info.converged and info.warm_started
    plain = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD)
    xp, ip, _ = plain.pcg_solve(plain.to_device(b))
    print(f"CNN-started {info.iterations} iterations (rejected {info.guess_rejected}), plain {ip.iterations}")
    assert rel(ctx.to_host(x), ctx.to_host(xp)) < 5e-5


def test_cnn_warm_start_needs_a_loaded_network():
    from ds_smoke import abi
    from ds_smoke.solver import Context, SolverError

    sc, ctx, model, flat, code, b, _ = _setup()
    fresh = Context(sc.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN), abi.MODE_FORWARD)
    with pytest.raises(SolverError):
        fresh.pcg_solve(fresh.to_device(b))
    with pytest.raises(SolverError):
        fresh.warm_start_load(flat[:-1], 1.0)
