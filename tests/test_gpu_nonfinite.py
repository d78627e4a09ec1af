"""Non-finite and huge velocities through the trace kernels (a diverged solve upstream produces them).

A trace position of +inf or beyond 2^31 cells used to convert to INT_MAX, whose x0 + 1 wrapped
negative and passed the interior test of the trilinear gather: an index near 2^31 faces, far out
of the arrays. Trace positions are now clamped to +-2^24 cells before the conversion (device.hpp,
corners and solid_at). Here the velocity carries +-inf, NaN and +-1e30 on a few faces of each axis:
the advection operators (forward and reverse, velocity and scalars) and a whole forward step must
complete, and the faces whose traces do not meet a poisoned value stay finite.
"""
import numpy as np
import pytest

from helpers import SCENES, random_fields

pytestmark = pytest.mark.gpu

POISON = (np.inf, -np.inf, np.nan, 1e30, -1e30)


@pytest.fixture(scope="module", params=sorted(SCENES))
def poisoned(request):
    from ds_smoke import abi
    from ds_smoke.solver import Context

    sc = SCENES[request.param]()
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    f = random_fields(sc, seed=5)
    vel = [np.array(v, dtype=np.float32) for v in f["vel"]]
    rng = np.random.default_rng(7)
    for a in range(3):
        flat = vel[a].reshape(-1)
        for k, value in enumerate(POISON):
            flat[rng.integers(flat.size)] = value
    return request.param, sc, ctx, f, vel


def test_advection_with_poisoned_velocity_completes(poisoned):
    name, sc, ctx, f, vel = poisoned
    print(f"non-finite velocity: {name}", flush=True)
    u = tuple(ctx.to_device(v) for v in vel)
    out = ctx.to_host(ctx.advect_velocity_forward(u))
    ua = ctx.zeros_faces()
    ctx.advect_velocity_vjp(u, tuple(ctx.to_device(x) for x in f["face_adj"]), ua)
    src = ctx.to_device(f["rho"])
    rho = ctx.to_host(ctx.advect_scalar_forward(0, src, u))
    src_adj, ua2 = ctx.zeros_cell(), ctx.zeros_faces()
    ctx.advect_scalar_vjp(0, src, u, ctx.to_device(f["cell_adj"]), src_adj, ua2)
    ctx.synchronize()
    # most faces' traces stay away from the handful of poisoned faces
    for a in range(3):
        assert np.mean(np.isfinite(out[a])) > 0.5, (name, a)
        assert np.mean(np.isfinite(ctx.to_host(ua[a]))) > 0.5, (name, a)
    assert np.mean(np.isfinite(rho)) > 0.5, name


def test_step_with_poisoned_velocity_completes(poisoned):
    from ds_smoke import scenes

    name, sc, ctx, f, vel = poisoned
    z = np.zeros(scenes.cell_shape(sc.resolution), dtype=np.float32)
    state = ctx.state_from_host({"density": f["rho"], "temperature": f["T"], "velocity": tuple(vel)})
    params = ctx.allocate_parameters(sc.params)
    ctl = ctx.control_from_host({"density_source": f["s_rho"], "temperature_source": f["s_T"],
                                 "external_acceleration": (z, z.copy(), z.copy())})
    # the pressure solve of a non-finite right-hand side does not converge: the step reports it
    nxt, _, info = ctx.forward_step(state, ctl, params, check=False)
    ctx.synchronize()
    print(f"  {name}: step solve converged {info.converged}, final residual {info.final_residual}", flush=True)
    assert not info.converged or np.isfinite(info.final_residual)
