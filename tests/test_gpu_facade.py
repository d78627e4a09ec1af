"""The C++ facade (xayah/smoke.hpp, xayah/solver.hpp) driven as reference code would drive
xayah::smoke::Model and xayah::solver (model.ixx:51-96, solver.ixx:53-150), checked against the
C-ABI trajectory driver on the same plume: the final density is bit-identical (the forward path is
deterministic), both gradients (solver::vjp over a stored trajectory, and the checkpointed
pair) match ds_loss_gradient and the FP64 oracle, and both forward-mode drivers (solver::jvp over
the stored trajectory, solver::jvp_streaming) match ds_trajectory_jvp bit for bit and give the
directional derivative that the gradient gives."""
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
    rng = np.random.default_rng(8)
    v = [rng.standard_normal(x.shape).astype(np.float32) for x in s0["velocity"]]
    for a, name in enumerate("xyz"):
        v[a].tofile(tmp_path / f"v_{name}.f32")
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
    # forward mode through the facade
    t_store = np.fromfile(tmp_path / "tangent_store.f32", dtype=np.float32)
    t_stream = np.fromfile(tmp_path / "tangent_stream.f32", dtype=np.float32)
    vt = {"density": np.zeros_like(s0["density"]), "temperature": np.zeros_like(s0["temperature"]), "velocity": tuple(v)}
    zc = {"density_source": np.zeros_like(ctl["density_source"]), "temperature_source": np.zeros_like(ctl["temperature_source"]),
          "external_acceleration": tuple(np.zeros_like(x) for x in ctl["external_acceleration"])}
    _, ft, _ = ctx.trajectory_jvp(ctx.state_from_host(s0), ctx.control_from_host(ctl), ctx.allocate_parameters(sc.params), steps,
                                  ctx.state_from_host(vt), ctx.control_from_host(zc), ctx.allocate_parameters((0.0, 0.0, 0.0, 0.0)))
    ref_t = ctx.to_host(ft)["density"].ravel()
    np.testing.assert_array_equal(t_store, ref_t)
    np.testing.assert_array_equal(t_stream, ref_t)
    fwd = float(np.dot((final.astype(np.float64) - sc.target.ravel()), t_store.astype(np.float64)))
    rev = sum(float(np.dot(gv[a].ravel().astype(np.float64), v[a].ravel().astype(np.float64))) for a in range(3))
    assert abs(fwd - rev) < 2e-3 * abs(rev), (fwd, rev)
