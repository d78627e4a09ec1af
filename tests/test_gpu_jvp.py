"""Forward mode on the HIP path (csrc/jvp.hip; Model::jvp_step, model.cpp:351-359; solver::jvp,
solver.ixx:83-113) against the FP64 oracle's forward mode, which is pinned by differences and by
transposition against its reverse (tests/test_oracle_jvp.py).

Per operator (three scenes): the buoyancy, velocity-advection, constraint, scalar-advection,
projection and confinement tangents. Per step: jvp_step against the oracle, and the device-side
dot-product test <J v, w> (GPU jvp_step) = <v, J^T w> (GPU vjp_step). Trajectory: the one-sweep
ds_trajectory_jvp against the oracle, and its directional derivative of the loss against the
checkpointed reverse sweep (ds_loss_gradient). FP32 tangents against FP64: relative L2 1e-4 per
operator and per step, 2e-3 over trajectories (the gradient tests' bound).
"""
import numpy as np
import pytest

from helpers import SCENES, random_fields, rel

pytestmark = pytest.mark.gpu


def _dot(a, b):
    if isinstance(a, dict):
        return sum(_dot(a[k], b[k]) for k in a)
    if isinstance(a, (tuple, list)):
        return sum(_dot(x, y) for x, y in zip(a, b))
    return float(np.sum(np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)))


def _scene(name, vorticity=False):
    sc = SCENES[name]()
    if vorticity:
        sc.vorticity = True
        sc.params = tuple(sc.params[:3]) + (2.0,)
    return sc


@pytest.fixture(scope="module", params=sorted(SCENES))
def jvp_setup(request):
    import oracle as O
    from ds_smoke import abi
    from ds_smoke.solver import Context

    sc = _scene(request.param, vorticity=True)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    orc = O.Oracle(sc, tolerance=1e-10)
    return request.param, sc, ctx, orc


def _dev3(ctx, t):
    return tuple(ctx.to_device(x) for x in t)


def test_operator_tangents_match_oracle(jvp_setup):
    name, sc, ctx, orc = jvp_setup
    print(f"operator tangents {name} {sc.resolution}", flush=True)
    f = random_fields(sc, seed=41, vel_scale=0.4)
    ft = random_fields(sc, seed=42, vel_scale=1.0)
    u, ut = _dev3(ctx, f["vel"]), _dev3(ctx, ft["vel"])
    # velocity self-advection (raw) and the constraint
    got = ctx.to_host(ctx.advect_velocity_jvp(u, ut))
    ref = orc.advect_velocity_jvp(f["vel"], ft["vel"])
    for a in range(3):
        err = rel(got[a], ref[a])
        print(f"  advect velocity [{a}] {err:.2e}")
        assert err < 1e-4, (name, a, err)
    got = ctx.to_host(ctx.constrain_jvp(ut))
    ref = orc.constrain_jvp(ft["vel"])
    for a in range(3):
        np.testing.assert_array_equal(got[a], ref[a].astype(np.float32))
    # scalar advection in both arguments, density and temperature boundaries
    for field in (0, 1):
        got = ctx.to_host(ctx.advect_scalar_jvp(field, ctx.to_device(f["rho"]), ctx.to_device(ft["rho"]), u, ut))
        ref = orc.advect_scalar_jvp(field, f["rho"], ft["rho"], f["vel"], ft["vel"])
        err = rel(got, ref)
        print(f"  advect scalar field {field} {err:.2e}")
        assert err < 1e-4, (name, field, err)
    # buoyancy with parameter tangents
    params = ctx.allocate_parameters((0.3, -0.2, 1.5, 2.0))
    params_t = ctx.allocate_parameters((0.7, -0.4, 0.9, 0.6))
    got = ctx.to_host(ctx.buoyancy_jvp(ctx.to_device(f["rho"]), ctx.to_device(f["T"]), ctx.to_device(ft["rho"]), ctx.to_device(ft["T"]),
                                       _dev3(ctx, ft["e"]), params, params_t))
    ref = orc.buoyancy_jvp(f["rho"], f["T"], ft["rho"], ft["T"], ft["e"], np.array([0.3, -0.2, 1.5, 2.0]),
                           np.array([0.7, -0.4, 0.9, 0.6]))
    for a in range(3):
        assert rel(got[a], ref[a]) < 1e-6, (name, a, rel(got[a], ref[a]))
    # confinement force tangent
    got = ctx.to_host(ctx.vorticity_jvp(u, ut, params, params_t))
    ref = orc.vorticity_jvp(f["vel"], ft["vel"], 2.0, 0.6)
    for a in range(3):
        err = rel(got[a], ref[a])
        print(f"  confinement [{a}] {err:.2e}")
        assert err < 1e-4, (name, a, err)
    # projection of a tangent (no fixed-face term)
    pv, info = ctx.projection_jvp(ut)
    assert info.converged
    ref, _ = orc.projection_jvp(ft["vel"])
    got = ctx.to_host(pv)
    for a in range(3):
        err = rel(got[a], ref[a])
        print(f"  projection [{a}] {err:.2e}")
        assert err < 1e-4, (name, a, err)


def _random_state(sc, seed, vel=0.4):
    f = random_fields(sc, seed=seed, vel_scale=vel)
    return ({"density": f["rho"], "temperature": f["T"], "velocity": f["vel"]},
            {"density_source": f["s_rho"], "temperature_source": f["s_T"], "external_acceleration": f["e"]})


@pytest.mark.parametrize("vorticity", [False, True])
@pytest.mark.parametrize("name", sorted(SCENES))
def test_step_jvp_matches_oracle_and_is_the_transpose_of_the_vjp(name, vorticity):
    import oracle as O
    from ds_smoke import abi
    from ds_smoke.solver import Context

    print(f"step tangent {name} vorticity={vorticity}", flush=True)
    sc = _scene(name, vorticity)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    orc = O.Oracle(sc, tolerance=1e-10)
    s, k = _random_state(sc, 43)
    st_h, kt_h = _random_state(sc, 44, vel=1.0)
    pv = tuple(sc.params[:3]) + ((2.0,) if vorticity else (0.0,))
    pt_h = (0.3, -0.2, 0.5, 0.7 if vorticity else 0.0)
    params, params_t = ctx.allocate_parameters(pv), ctx.allocate_parameters(pt_h)
    st, kt = ctx.state_from_host(s), ctx.control_from_host(k)
    nxt, cache, _ = ctx.forward_step(st, kt, params)
    nt, info = ctx.jvp_step(st, kt, params, nxt, cache, ctx.state_from_host(st_h), ctx.control_from_host(kt_h), params_t)
    assert info.converged
    ref_n, ref_cache, _ = orc.forward_step(s, k, np.array(pv))
    ref_t, _ = orc.jvp_step(s, ref_n, ref_cache, np.array(pv), st_h, kt_h, np.array(pt_h))
    g = ctx.to_host(nt)
    for key in ("density", "temperature"):
        err = rel(g[key], ref_t[key])
        print(f"  {key} tangent {err:.2e}")
        assert err < 1e-4, (name, key, err)
    for a in range(3):
        err = rel(g["velocity"][a], ref_t["velocity"][a])
        print(f"  velocity tangent [{a}] {err:.2e}")
        assert err < 1e-4, (name, a, err)
    # device-side dot-product test against the GPU reverse step
    rng = np.random.default_rng(45)
    w = {"density": rng.standard_normal(s["density"].shape).astype(np.float32),
         "temperature": rng.standard_normal(s["density"].shape).astype(np.float32),
         "velocity": tuple(rng.standard_normal(v.shape).astype(np.float32) for v in s["velocity"])}
    ca, pa = ctx.allocate_control_adjoint(), ctx.allocate_parameter_adjoint()
    prev, vinfo = ctx.vjp_step(st, kt, params, nxt, cache, ctx.state_from_host(w), control_adj=ca, param_adj=pa)
    assert vinfo.converged
    lhs = _dot(g, w)
    rhs = _dot(st_h, ctx.to_host(prev)) + _dot(kt_h, ctx.to_host(ca)) + float(np.dot(np.array(pt_h), ctx.to_host(pa)))
    print(f"  <J v, w> = {lhs:.8e}, <v, J^T w> = {rhs:.8e}")
    assert abs(lhs - rhs) < 1e-4 * (abs(lhs) + abs(rhs)), (lhs, rhs)


@pytest.mark.parametrize("vorticity", [False, True])
def test_trajectory_jvp_matches_oracle_and_the_reverse_sweep(vorticity):
    import oracle as O
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    print(f"trajectory tangent vorticity={vorticity}", flush=True)
    sc = scenes.plume(16)
    if vorticity:
        sc.vorticity = True
        sc.params = (0.0, -0.1, 1.0, 2.0)
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    orc = O.Oracle(sc, tolerance=1e-10)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    rng = np.random.default_rng(46)
    n = sc.resolution
    v = {"density": np.zeros_like(s0["density"]), "temperature": np.zeros_like(s0["temperature"]),
         "velocity": tuple(rng.standard_normal(x.shape).astype(np.float32) for x in s0["velocity"])}
    zero_c = {"density_source": np.zeros_like(ctl["density_source"]), "temperature_source": np.zeros_like(ctl["temperature_source"]),
              "external_acceleration": tuple(np.zeros_like(x) for x in ctl["external_acceleration"])}
    pt_h = (0.0, 0.0, 0.0, 0.5 if vorticity else 0.0)
    steps = 8
    params, params_t = ctx.allocate_parameters(sc.params), ctx.allocate_parameters(pt_h)
    fin, fin_t, info = ctx.trajectory_jvp(ctx.state_from_host(s0), ctx.control_from_host(ctl), params, steps, ctx.state_from_host(v),
                                          ctx.control_from_host(zero_c), params_t)
    assert info.forward_solves == steps and info.adjoint_solves == steps and info.not_converged == 0
    ref_fin, ref_t, _ = orc.trajectory_jvp(s0, ctl, sc.params, steps, v, zero_c, np.array(pt_h))
    g, gt = ctx.to_host(fin), ctx.to_host(fin_t)
    assert rel(g["density"], ref_fin["density"]) < 1e-5
    for key in ("density", "temperature"):
        err = rel(gt[key], ref_t[key])
        print(f"  final {key} tangent {err:.2e}")
        assert err < 2e-3, (key, err)
    for a in range(3):
        err = rel(gt["velocity"][a], ref_t["velocity"][a])
        print(f"  final velocity tangent [{a}] {err:.2e}")
        assert err < 2e-3, (a, err)
    # forward mode and reverse mode give the same directional derivative of the loss
    pa = ctx.allocate_parameter_adjoint()
    _, grad, _ = ctx.loss_gradient(ctx.state_from_host(s0), ctx.control_from_host(ctl), params, steps, 4, ctx.to_device(sc.target),
                                   param_adj=pa)
    fwd = _dot(g["density"].astype(np.float64) - sc.target, gt["density"])
    rev = _dot(ctx.to_host(grad)["velocity"], v["velocity"]) + float(np.dot(np.array(pt_h), ctx.to_host(pa)))
    print(f"  forward mode {fwd:.8e}, reverse mode {rev:.8e}")
    assert abs(fwd - rev) < 2e-3 * abs(rev), (fwd, rev)
