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


def _prolong(dc, f, shape):
    """Cell-centred trilinear interpolation of a coarse field to the fine grid (numpy), clamped."""
    out = dc.astype(np.float64)
    for axis, n in zip((0, 1, 2), shape):   # (z, y, x)
        nc = out.shape[axis]
        X = (np.arange(n) + 0.5) / f - 0.5
        fl = np.floor(X)
        t = X - fl
        i0 = np.clip(fl.astype(int), 0, nc - 1)
        i1 = np.clip(fl.astype(int) + 1, 0, nc - 1)
        a, b = np.take(out, i0, axis=axis), np.take(out, i1, axis=axis)
        tt = t.reshape([-1 if k == axis else 1 for k in range(3)])
        out = a + tt * (b - a)
    return out


def test_network_on_the_coarsened_grid():
    """levels = 1 at 128^3: restrict (sum of 8 children), the network at 64^3, trilinear prolongation,
    the A-norm-best step, against the same chain assembled from the 64^3 network's own guess and
    numpy; and the guess never increases the A-norm error of the base guess."""
    import oracle as O
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context
    from ds_smoke.unet import init_parameters

    params = init_parameters(11)
    sc = scenes.plume(128, with_target=False)
    ctx = Context(sc.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN_ALONE), abi.MODE_FORWARD)
    ctx.warm_start_load(params, 1.0, levels=1)
    sc64 = scenes.plume(64, with_target=False)
    c64 = Context(sc64.config(tolerance=1e-6, warm_start=abi.WARM_START_CNN_ALONE), abi.MODE_FORWARD)
    c64.warm_start_load(params, 1.0)
    _, code = ctx.masks()
    unk = code > 0
    rng = np.random.default_rng(128)
    # a smooth right-hand side (the solver's are): a few Gaussian bumps
    z, y, x = np.meshgrid(*(np.arange(128) + 0.5,) * 3, indexing="ij")
    b = sum(rng.uniform(-1, 1) * np.exp(-((x - cx) ** 2 + (y - cy) ** 2 + (z - cz) ** 2) / (2 * s * s))
            for cx, cy, cz, s in rng.uniform(20, 108, (5, 4)) * [1, 1, 1, 0.15])
    b = np.where(unk, b, 0.0).astype(np.float32)
    got = ctx.to_host(ctx.warm_start_guess(ctx.to_device(b))).astype(np.float64)
    rc = b.astype(np.float64).reshape(64, 2, 64, 2, 64, 2).sum((1, 3, 5)).astype(np.float32)
    d64 = c64.to_host(c64.warm_start_infer(c64.to_device(rc))).astype(np.float64)
    d = np.where(unk, _prolong(d64, 2, b.shape), 0.0)
    orc = O.Oracle(sc, tolerance=1e-6)
    ad = orc.apply_matrix(d)
    ref = d * (np.dot(d.ravel(), b.astype(np.float64).ravel()) / np.dot(d.ravel(), ad.ravel()))
    err = rel(got, ref)
    print(f"levels 1 at 128^3 against restrict / 64^3 network / prolong: relative L2 {err:.2e}")
    assert np.all(got[~unk] == 0)
    assert err < 2e-2, err
    # a base guess: the network's correction never raises the A-norm error (<e, A e> = x'Ax - 2 x'b + const)
    base = (0.5 * ref).astype(np.float32)
    corr = ctx.to_host(ctx.warm_start_guess(ctx.to_device(b), base=ctx.to_device(base))).astype(np.float64)
    energy = lambda v: float(v.ravel() @ orc.apply_matrix(v).ravel() - 2.0 * v.ravel() @ b.astype(np.float64).ravel())
    print(f"  energy: base {energy(base.astype(np.float64)):.6e}, corrected {energy(corr):.6e}")
    assert energy(corr) <= energy(base.astype(np.float64)) + 1e-9 * abs(energy(base.astype(np.float64)))
