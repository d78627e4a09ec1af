"""The torch.autograd.Function over the C ABI (ds_smoke.autograd): forward = ds_simulate,
backward = ds_trajectory_vjp (solver::vjp, solvers/solver.ixx:115-150, with checkpoints).

Gradients of a loss computed in torch on the rollout's outputs are checked, for a terminal
adjoint that touches every output field, against the same reverse sweep chained by hand from
ds_forward_step / ds_vjp_step (same FP32 arithmetic: tight), and against the FP64 oracle's chain
(loose: with a random terminal adjoint on every face, FP32 and FP64 trajectories that land on
different sides of a cell face give locally different trilinear derivatives); and against
ds_loss_gradient for the density loss.
"""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def oracle_terminal_vjp(orc, state, control, params, steps, final_adj):
    """Reverse sweep of the oracle for dL/d(final state) = final_adj; returns (final, prev_adj, control_adj, param_adj)."""
    states, caches = [state], []
    for _ in range(steps):
        nxt, cache, _ = orc.forward_step(states[-1], control, params)
        states.append(nxt)
        caches.append(cache)
    ca = {"density_source": orc.zeros_cell(), "temperature_source": orc.zeros_cell(),
          "external_acceleration": tuple(orc.zeros_cell() for _ in range(3))}
    pa = np.zeros(4)
    adj = final_adj
    for s in reversed(range(steps)):
        adj, _ = orc.vjp_step(states[s + 1], caches[s], params, adj, control_adj=ca, param_adj=pa)
    return states[-1], adj, ca, pa


def test_rollout_gradients_match_oracle():
    import oracle as O
    import torch
    from ds_smoke import abi, scenes
    from ds_smoke.autograd import rollout
    from ds_smoke.solver import Context

    sc = scenes.plume(12)
    steps, k = 5, 2
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    rng = np.random.default_rng(5)
    s0["density"] = np.abs(rng.standard_normal(s0["density"].shape)).astype(np.float32) * 0.1
    state = {k_: (v.requires_grad_() if isinstance(v, torch.Tensor) else v) for k_, v in ctx.state_from_host(s0).items()}
    state["velocity"] = tuple(v.requires_grad_() for v in state["velocity"])
    control = ctx.control_from_host(ctl)
    control["density_source"].requires_grad_()
    params = ctx.allocate_parameters(sc.params).requires_grad_()
    w_u = [rng.standard_normal(v.shape).astype(np.float32) for v in s0["velocity"]]
    w_t = rng.standard_normal(s0["temperature"].shape).astype(np.float32)
    target = torch.from_numpy(sc.target).to(ctx.device)
    print("autograd rollout 12^3, 5 steps, checkpoint 2", flush=True)
    out = rollout(ctx, params, state, control, steps, k)
    loss = 0.5 * ((out["density"] - target) ** 2).sum() + (out["temperature"] * torch.from_numpy(w_t).to(ctx.device)).sum()
    for a in range(3):
        loss = loss + (out["velocity"][a] * torch.from_numpy(w_u[a]).to(ctx.device)).sum()
    loss.backward()
    torch.cuda.synchronize()

    orc = O.Oracle(sc, tolerance=1e-10)
    states = [s0]
    for _ in range(steps):
        states.append(orc.forward_step(states[-1], ctl, sc.params)[0])
    final_adj = {"density": states[-1]["density"] - sc.target, "temperature": w_t.astype(np.float64),
                 "velocity": tuple(w.astype(np.float64) for w in w_u)}
    fin, adj, ca, pa = oracle_terminal_vjp(orc, s0, ctl, sc.params, steps, final_adj)
    assert rel(out["density"].detach().cpu().numpy(), fin["density"]) < 1e-5
    # the same sweep chained by hand on the GPU from the per-step entry points
    gpu_states, gpu_caches = [ctx.state_from_host(s0)], []
    for _ in range(steps):
        nxt, cache, _ = ctx.forward_step(gpu_states[-1], control, params.detach())
        gpu_states.append(nxt)
        gpu_caches.append(cache)
    g_adj = {"density": out["density"].detach() - target, "temperature": torch.from_numpy(w_t).to(ctx.device),
             "velocity": tuple(torch.from_numpy(w).to(ctx.device) for w in w_u)}
    g_ca, g_pa = ctx.allocate_control_adjoint(), ctx.allocate_parameter_adjoint()
    for s in reversed(range(steps)):
        g_adj = ctx.vjp_step(gpu_states[s], control, params.detach(), gpu_states[s + 1], gpu_caches[s], g_adj,
                             control_adj=g_ca, param_adj=g_pa)[0]
    chained = ctx.to_host(g_adj)
    assert rel(state["density"].grad.cpu().numpy(), chained["density"]) < 1e-4
    for a in range(3):
        assert rel(state["velocity"][a].grad.cpu().numpy(), chained["velocity"][a]) < 1e-4
    assert rel(control["density_source"].grad.cpu().numpy(), ctx.to_host(g_ca["density_source"])) < 1e-4
    np.testing.assert_allclose(params.grad.cpu().numpy()[:3], ctx.to_host(g_pa)[:3], rtol=1e-4, atol=1e-6)
    # FP64 oracle (loose, see above)
    tol = 2e-2
    assert rel(state["density"].grad.cpu().numpy(), adj["density"]) < tol
    for a in range(3):
        g = state["velocity"][a].grad.cpu().numpy()
        assert rel(g, adj["velocity"][a]) < tol, (a, rel(g, adj["velocity"][a]))
    assert rel(control["density_source"].grad.cpu().numpy(), ca["density_source"]) < tol


def test_density_loss_through_autograd_equals_loss_gradient():
    import torch
    from ds_smoke import abi, scenes
    from ds_smoke.autograd import rollout
    from ds_smoke.solver import Context

    sc = scenes.plume(16)
    steps, k = 8, 3
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    s0, ctl = ctx.state_from_host(scenes.initial_state(sc)), ctx.control_from_host(scenes.control(sc))
    params = ctx.allocate_parameters(sc.params)
    target = ctx.to_device(sc.target)
    loss_ref, g_ref, _ = ctx.loss_gradient(s0, ctl, params, steps, k, target)
    state = {"density": s0["density"].clone(), "temperature": s0["temperature"].clone(),
             "velocity": tuple(v.clone().requires_grad_() for v in s0["velocity"])}
    out = rollout(ctx, params, state, ctl, steps, k)
    loss = 0.5 * ((out["density"] - target) ** 2).sum(dtype=torch.float64)
    loss.backward()
    assert abs(loss.item() - loss_ref) <= 1e-6 * abs(loss_ref)
    ref = ctx.to_host(g_ref)["velocity"]
    for a in range(3):
        assert rel(state["velocity"][a].grad.cpu().numpy(), ref[a]) < 1e-5


def test_forward_mode_through_autograd():
    """torch.autograd.forward_ad on the rollout: the tangent equals ds_trajectory_jvp's, and its
    directional derivative of the density loss equals the reverse sweep's."""
    import torch
    import torch.autograd.forward_ad as fwAD
    from ds_smoke import abi, autograd, scenes
    from ds_smoke.solver import Context

    sc = scenes.plume(16)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    s0, ctl = ctx.state_from_host(scenes.initial_state(sc)), ctx.control_from_host(scenes.control(sc))
    params = ctx.allocate_parameters(sc.params)
    rng = np.random.default_rng(9)
    v = tuple(ctx.to_device(rng.standard_normal(tuple(x.shape)).astype(np.float32)) for x in s0["velocity"])
    steps = 6
    with fwAD.dual_level():
        dual = dict(s0, velocity=tuple(fwAD.make_dual(u, t) for u, t in zip(s0["velocity"], v)))
        out = autograd.rollout(ctx, params, dual, ctl, steps)
        primal, tangent = fwAD.unpack_dual(out["density"])
    torch.cuda.synchronize()
    zeros = lambda d: {k: (tuple(torch.zeros_like(x) for x in val) if isinstance(val, tuple) else torch.zeros_like(val)) for k, val in d.items()}
    st_t = dict(zeros(s0), velocity=v)
    fin, fin_t, _ = ctx.trajectory_jvp(s0, ctl, params, steps, st_t, zeros(ctl), torch.zeros_like(params))
    np.testing.assert_array_equal(tangent.cpu().numpy(), ctx.to_host(fin_t)["density"])
    np.testing.assert_array_equal(primal.cpu().numpy(), ctx.to_host(fin)["density"])
    target = ctx.to_device(sc.target)
    _, grad, _ = ctx.loss_gradient(s0, ctl, params, steps, 3, target)
    fwd = float(((primal.double() - target.double()) * tangent.double()).sum())
    rev = sum(float((g.double() * t.double()).sum()) for g, t in zip(grad["velocity"], v))
    print(f"forward mode {fwd:.8e}, reverse mode {rev:.8e}")
    assert abs(fwd - rev) < 2e-3 * abs(rev)
