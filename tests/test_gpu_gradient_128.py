"""The benchmark gradient at full size: the 128^3 plume, 64 steps, checkpoint every 8, loss
0.5 ||rho_64 - target||^2 with respect to the initial velocity (all 6,340,608 face DOFs).

32 DOFs drawn uniformly over the face DOFs: each gradient entry from the reverse sweep
(ds_loss_gradient) is compared with the directional derivative along that DOF from forward mode
(ds_trajectory_jvp: <rho_64 - target, d rho_64 / d u_i>), a separately written linearisation of the
same step. Central differences of the FP32 forward cannot resolve single DOFs here: the loss is
~6e6 and a pressure solve to 1e-6 moves it by ~1e-6 relative between runs, so a difference with
eps = 1e-3 carries noise of order 10 against derivatives of order 0.1-1. The finite-difference
check of the same gradient runs in FP64 in the oracle (tests/test_oracle.py, trajectory), and the
GPU gradient matches the oracle's at 16^3 (tests/test_gpu_step.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_gradient_entries_match_forward_mode_at_128():
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    print("128^3 gradient: 64 steps, k = 8, 32 DOFs", flush=True)
    sc = scenes.plume(128)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    s0h, kh = scenes.initial_state(sc), scenes.control(sc)
    s0, k = ctx.state_from_host(s0h), ctx.control_from_host(kh)
    params = ctx.allocate_parameters(sc.params)
    steps = 64
    loss, grad, info = ctx.loss_gradient(s0, k, params, steps, 8, ctx.to_device(sc.target))
    assert info.not_converged == 0
    g = ctx.to_host(grad)["velocity"]
    sizes = [int(np.prod(v.shape)) for v in s0h["velocity"]]
    assert sum(sizes) == 6340608
    rng = np.random.default_rng(20261202)
    dofs = np.sort(rng.choice(sum(sizes), 32, replace=False))
    zero_control = ctx.control_from_host({key: (tuple(np.zeros_like(x) for x in val) if isinstance(val, tuple) else np.zeros_like(val))
                                          for key, val in kh.items()})
    zero_params = ctx.allocate_parameters((0.0, 0.0, 0.0, 0.0))
    reverse, forward = [], []
    fin = None
    for d in dofs:
        a = int(np.searchsorted(np.cumsum(sizes), d, side="right"))
        local = int(d - sum(sizes[:a]))
        one_hot = [np.zeros_like(v) for v in s0h["velocity"]]
        one_hot[a].reshape(-1)[local] = 1.0
        tangent = ctx.state_from_host({"density": np.zeros_like(s0h["density"]), "temperature": np.zeros_like(s0h["temperature"]),
                                       "velocity": tuple(one_hot)})
        fin, fin_t, _ = ctx.trajectory_jvp(s0, k, params, steps, tangent, zero_control, zero_params, final_state=fin)
        rho = ctx.to_host(fin)["density"].astype(np.float64)
        d_rho = ctx.to_host(fin_t)["density"].astype(np.float64)
        forward.append(float(np.sum((rho - sc.target) * d_rho)))
        reverse.append(float(g[a].reshape(-1)[local]))
    forward, reverse = np.array(forward), np.array(reverse)
    for d, f, r in zip(dofs, forward, reverse):
        print(f"  dof {d:8d}: reverse {r: .6e}  forward {f: .6e}")
    err = np.linalg.norm(forward - reverse) / np.linalg.norm(reverse)
    print(f"  relative L2 over the 32 DOFs {err:.2e}")
    assert np.linalg.norm(reverse) > 0.0
    assert err < 1e-2, err
