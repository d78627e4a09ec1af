"""The 64^3 plume after 100 forward steps on the GPU against the committed oracle fixture.

Fixture: tests/golden/plume64_100.npz, written by tests/golden/make_golden.py from the FP64 oracle
with the pressure solved to 1e-10. The GPU solves to the workload's 1e-6 with FP32 fields; the
density after 100 steps must stay within 1e-5 relative L2 of the fixture.
"""
import os

import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "plume64_100.npz")


def test_plume64_100_steps_density_matches_fixture():
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    ref = np.load(GOLDEN)
    sc = scenes.plume(64, with_target=False)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_FORWARD)
    s0 = ctx.state_from_host(scenes.initial_state(sc))
    ctl = ctx.control_from_host(scenes.control(sc))
    params = ctx.allocate_parameters(sc.params)
    final, info = ctx.simulate(s0, ctl, params, 100)
    got = ctx.to_host(final)
    assert info.not_converged == 0
    assert info.forward_solves == 100
    err = rel(got["density"], ref["density"])
    print(f"64^3 plume, 100 steps: density rel L2 {err:.3e}, CG iterations/solve "
          f"{info.forward_iterations / 100:.1f} (oracle at 1e-10: {ref['cg_iterations'].mean():.1f})")
    assert err <= 1e-5, err
    assert rel(got["temperature"], ref["density"]) <= 1e-5
