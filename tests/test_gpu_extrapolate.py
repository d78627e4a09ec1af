"""DS_WARM_START_EXTRAPOLATE (each forward pressure solve starts from 2 p_{n-1} - p_{n-2}, each
adjoint solve from the previous adjoint solution) and DS_WARM_START_PROJECTION (each solve starts
from the A-norm-best combination of the last 8 solutions of its sequence). The guess changes only
where the iterations start, so the trajectory gradient must match the x0 = 0 one to the solve
tolerance, every solve must converge, and a guess out of the step sequence (the recompute pass's
taped pressure) must not be extrapolated across."""
import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu


def test_extrapolated_guess_keeps_the_gradient():
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    sc = scenes.plume(32)
    print("extrapolated warm start: plume 32^3, 16 steps + VJP, k = 4", flush=True)
    s0, kk = scenes.initial_state(sc), scenes.control(sc)
    out = {}
    for mode in (abi.WARM_START_NONE, abi.WARM_START_PREVIOUS, abi.WARM_START_EXTRAPOLATE, abi.WARM_START_PROJECTION):
        ctx = Context(sc.config(tolerance=1e-6, warm_start=mode), abi.MODE_DIFFERENTIABLE)
        for _ in range(2):   # the second sweep starts after a finished one: no stale sequence
            loss, grad, info = ctx.loss_gradient(ctx.state_from_host(s0), ctx.control_from_host(kk), ctx.allocate_parameters(sc.params),
                                                 16, 4, ctx.to_device(sc.target))
        assert info.not_converged == 0
        out[mode] = (loss, ctx.to_host(grad)["velocity"], info)
        print(f"  mode {mode}: CG per forward solve {info.forward_iterations / info.forward_solves:.1f}, "
              f"per adjoint solve {info.adjoint_iterations / info.adjoint_solves:.1f}, rejected {info.guesses_rejected}")
        ctx.close()
    l0, g0, _ = out[abi.WARM_START_NONE]
    for mode in (abi.WARM_START_PREVIOUS, abi.WARM_START_EXTRAPOLATE, abi.WARM_START_PROJECTION):
        l1, g1, _ = out[mode]
        assert abs(l1 - l0) <= 1e-6 * abs(l0)
        for a in range(3):
            err = rel(g1[a], g0[a])
            assert err < 1e-4, (mode, a, err)
