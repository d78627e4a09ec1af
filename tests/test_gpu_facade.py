"""The C++ facade (xayah/smoke.hpp, xayah/solver.hpp) driven as reference code would drive
xayah::smoke::Model and xayah::solver (model.ixx:51-96, solver.ixx:53-150), checked against the
C-ABI trajectory driver on the same plume: the final density is bit-identical (the forward path is
deterministic) and both gradients (solver::vjp over a stored trajectory, and the checkpointed
pair) match ds_loss_gradient and the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BINARY = os.path.join(ROOT, "diff-solver_amd", "lib", "facade_check")


def test_facade_trajectory_and_gradients(tmp_path):
    import oracle as O
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    if not os.path.exists(BINARY):
        pytest.fail(f"{BINARY} is missing: run __graft_entry__.build() first")
    n, steps, k = 16, 8, 3
    sc = scenes.plume(n)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    for a, name in enumerate("xyz"):
        s0["velocity"][a].astype(np.float32).tofile(tmp_path / f"u0_{name}.f32")
    ctl["density_source"].astype(np.float32).tofile(tmp_path / "source.f32")
    sc.target.astype(np.float32).tofile(tmp_path / "target.f32")
    print(f"facade_check n={n} steps={steps} k={k}", flush=True)
    out = subprocess.run([BINARY, str(tmp_path), str(n), str(steps), str(k)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    final = np.fromfile(tmp_path / "final_density.f32", dtype=np.float32)
    grad_store = [np.fromfile(tmp_path / f"grad_store_{c}.f32", dtype=np.float32) for c in "xyz"]
    grad_ckpt = [np.fromfile(tmp_path / f"grad_ckpt_{c}.f32", dtype=np.float32) for c in "xyz"]

    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    fin = ctx.allocate_state()
    loss, g, info = ctx.loss_gradient(ctx.state_from_host(s0), ctx.control_from_host(ctl), ctx.allocate_parameters(sc.params),
                                      steps, k, ctx.to_device(sc.target), final_state=fin)
    got = ctx.to_host(fin)
    np.testing.assert_array_equal(final, got["density"].ravel())
    gv = ctx.to_host(g)["velocity"]
    ref = O.Oracle(sc, tolerance=1e-10).loss_gradient(s0, ctl, sc.params, steps, k, sc.target)
    for a in range(3):
        assert rel(grad_store[a], gv[a].ravel()) < 1e-5, (a, rel(grad_store[a], gv[a].ravel()))
        assert rel(grad_ckpt[a], gv[a].ravel()) < 1e-5
        assert rel(grad_store[a], ref["grad"]["velocity"][a].ravel()) < 2e-3
    assert abs(float((tmp_path / "loss.txt").read_text()) - loss) <= 1e-6 * abs(loss)
