"""The reference's own projection on the HIP path (DS_PROJECTION_RBGS, csrc/rbgs.hip): red-black
Gauss-Seidel from p = 0 for `pressure_iterations` sweeps of each colour (operators.cu:885-902,
1004-1013), its unrolled reverse (:935-955, 1020-1031) and its tangent (:1015-1018), against the
oracle's RBGS mode (pinned by transposition, tests/test_oracle_jvp.py, and by convergence to the
PCG solution, tests/test_oracle.py).

Per operator and per step on three scenes (closed box with an anchor; colliders with mixed
boundaries and a fixed-value pressure face, no anchor; periodic), the device-side dot-product test
<J v, w> = <v, J^T w> of the GPU step tangent and reverse, the trajectory gradient, and
determinism (a fixed-cost sweep in fixed order: two runs are bit-identical). FP32 against FP64:
relative L2 1e-5 forward, 1e-4 reverse and tangent.
"""
import numpy as np
import pytest

from helpers import SCENES, random_fields, rel

pytestmark = pytest.mark.gpu

ITERATIONS = 40


@pytest.fixture(scope="module", params=sorted(SCENES))
def rbgs_setup(request):
    import oracle as O
    from ds_smoke import abi
    from ds_smoke.solver import Context

    sc = SCENES[request.param]()
    ctx = Context(sc.config(solver=abi.PROJECTION_RBGS, rbgs_iterations=ITERATIONS), abi.MODE_DIFFERENTIABLE)
    orc = O.Oracle(sc, solver=abi.PROJECTION_RBGS, rbgs_iterations=ITERATIONS)
    return request.param, sc, ctx, orc


def _dot(a, b):
    if isinstance(a, dict):
        return sum(_dot(a[k], b[k]) for k in a)
    if isinstance(a, (tuple, list)):
        return sum(_dot(x, y) for x, y in zip(a, b))
    return float(np.sum(np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)))


def test_rbgs_projection_matches_oracle(rbgs_setup):
    name, sc, ctx, orc = rbgs_setup
    print(f"RBGS projection {name} {sc.resolution}", flush=True)
    f = random_fields(sc, seed=51, vel_scale=0.5)
    u = tuple(ctx.to_device(v) for v in f["vel"])
    out, p, rhs, info = ctx.projection_forward(u)
    assert info.iterations == ITERATIONS and info.converged
    ref, ref_p, ref_rhs, _ = orc.projection_forward(f["vel"])
    assert rel(ctx.to_host(rhs), ref_rhs) < 1e-6
    err_p = rel(ctx.to_host(p), ref_p)
    print(f"  pressure {err_p:.2e}")
    assert err_p < 1e-5, (name, err_p)
    g = ctx.to_host(out)
    for a in range(3):
        assert rel(g[a], ref[a]) < 1e-5, (name, a, rel(g[a], ref[a]))
    # determinism: a second run is bit-identical
    out2, p2, _, _ = ctx.projection_forward(u)
    np.testing.assert_array_equal(ctx.to_host(p2), ctx.to_host(p))
    # reverse: the unrolled adjoint of the sweeps
    w = tuple(ctx.to_device(v) for v in f["face_adj"])
    ua, _ = ctx.projection_vjp(w)
    ref_ua, _ = orc.projection_vjp(f["face_adj"])
    ga = ctx.to_host(ua)
    for a in range(3):
        err = rel(ga[a], ref_ua[a])
        print(f"  reverse [{a}] {err:.2e}")
        assert err < 1e-4, (name, a, err)
    # tangent: the sweeps on the tangent with zero boundary values
    tv, _ = ctx.projection_jvp(w)
    ref_t, _ = orc.projection_jvp(f["face_adj"])
    gt = ctx.to_host(tv)
    for a in range(3):
        err = rel(gt[a], ref_t[a])
        print(f"  tangent [{a}] {err:.2e}")
        assert err < 1e-4, (name, a, err)
    # on the GPU, the reverse is the transpose of the tangent map
    v = tuple(ctx.to_device(x) for x in random_fields(sc, seed=52)["face_adj"])
    jv, _ = ctx.projection_jvp(v)
    lhs, rhs_ = _dot(ctx.to_host(jv), f["face_adj"]), _dot(ctx.to_host(v), ga)
    assert abs(lhs - rhs_) < 1e-5 * (abs(lhs) + abs(rhs_)), (lhs, rhs_)


def test_rbgs_step_forward_reverse_tangent(rbgs_setup):
    name, sc, ctx, orc = rbgs_setup
    print(f"RBGS step {name}", flush=True)
    f = random_fields(sc, seed=53, vel_scale=0.4)
    s = {"density": f["rho"], "temperature": f["T"], "velocity": f["vel"]}
    k = {"density_source": f["s_rho"], "temperature_source": f["s_T"], "external_acceleration": f["e"]}
    params = ctx.allocate_parameters(sc.params)
    st, kt = ctx.state_from_host(s), ctx.control_from_host(k)
    nxt, cache, info = ctx.forward_step(st, kt, params)
    assert info.iterations == ITERATIONS
    ref, ref_cache, _ = orc.forward_step(s, k, sc.params)
    got = ctx.to_host(nxt)
    for a in range(3):
        assert rel(got["velocity"][a], ref["velocity"][a]) < 1e-5, (name, a, rel(got["velocity"][a], ref["velocity"][a]))
    assert rel(got["density"], ref["density"]) < 1e-5
    w = {"density": f["cell_adj"], "temperature": f["cell_adj2"], "velocity": f["face_adj"]}
    pa, ca = ctx.allocate_parameter_adjoint(), ctx.allocate_control_adjoint()
    prev, _ = ctx.vjp_step(st, kt, params, nxt, cache, ctx.state_from_host(w), control_adj=ca, param_adj=pa)
    opa = np.zeros(4)
    ref_prev, _ = orc.vjp_step(ref, ref_cache, sc.params, w, param_adj=opa, state=s)
    g = ctx.to_host(prev)
    # the collider scene can meet a kink of the trilinear stencil (a trace point on a stencil
    # plane, where FP32 and FP64 may take different one-sided derivatives): looser bound there
    tol = 1e-3 if sc.colliders else 1e-4
    for a in range(3):
        err = rel(g["velocity"][a], ref_prev["velocity"][a])
        print(f"  reverse velocity [{a}] {err:.2e}")
        assert err < tol, (name, a, err)
    ft = random_fields(sc, seed=54, vel_scale=1.0)
    s_t = {"density": ft["rho"], "temperature": ft["T"], "velocity": ft["vel"]}
    k_t = {"density_source": ft["s_rho"], "temperature_source": ft["s_T"], "external_acceleration": ft["e"]}
    pt_h = (0.3, -0.2, 0.5, 0.0)
    nt, _ = ctx.jvp_step(st, kt, params, nxt, cache, ctx.state_from_host(s_t), ctx.control_from_host(k_t), ctx.allocate_parameters(pt_h))
    ref_t, _ = orc.jvp_step(s, ref, ref_cache, np.array(sc.params), s_t, k_t, np.array(pt_h))
    gt = ctx.to_host(nt)
    for a in range(3):
        err = rel(gt["velocity"][a], ref_t["velocity"][a])
        print(f"  tangent velocity [{a}] {err:.2e}")
        assert err < tol, (name, a, err)
    lhs = _dot(gt, w)
    rhs = _dot(s_t, g) + _dot(k_t, ctx.to_host(ca)) + float(np.dot(np.array(pt_h), ctx.to_host(pa)))
    print(f"  <J v, w> = {lhs:.8e}, <v, J^T w> = {rhs:.8e}")
    assert abs(lhs - rhs) < 1e-4 * (abs(lhs) + abs(rhs)), (lhs, rhs)


def test_rbgs_loss_gradient_matches_oracle():
    import oracle as O
    from ds_smoke import abi, scenes
    from ds_smoke.solver import Context

    print("RBGS loss gradient plume 16^3", flush=True)
    sc = scenes.plume(16)
    ctx = Context(sc.config(solver=abi.PROJECTION_RBGS, rbgs_iterations=80), abi.MODE_DIFFERENTIABLE)
    orc = O.Oracle(sc, solver=abi.PROJECTION_RBGS, rbgs_iterations=80)
    s0, kk = scenes.initial_state(sc), scenes.control(sc)
    params = ctx.allocate_parameters(sc.params)
    pa = ctx.allocate_parameter_adjoint()
    loss, grad, info = ctx.loss_gradient(ctx.state_from_host(s0), ctx.control_from_host(kk), params, 8, 3, ctx.to_device(sc.target),
                                         param_adj=pa)
    ref = orc.loss_gradient(s0, kk, sc.params, 8, 3, sc.target)
    assert abs(loss - ref["loss"]) / ref["loss"] < 1e-5
    g = ctx.to_host(grad)
    for a in range(3):
        err = rel(g["velocity"][a], ref["grad"]["velocity"][a])
        print(f"  grad u[{a}] {err:.2e}")
        assert err < 2e-3, (a, err)
    np.testing.assert_allclose(ctx.to_host(pa)[:3], ref["params"][:3], rtol=2e-3)
