"""Pins the oracle's forward mode (jvp_step, model.cpp:351-359; the *_jvp kernels,
operators.cu:315-324,390-413,509-535, and the forwards they reuse, :575-638,1015-1018).

The tangents are written from the reference's JVP kernels, independently of the reverse code, so:
  * each operator's JVP agrees with FP64 central differences of its forward (nonlinear maps) or
    with forward(u + v) - forward(u) (affine maps);
  * per step, <J v, w> from jvp_step equals <v, J^T w> from vjp_step on three scenes (and with
    vorticity confinement): the forward and reverse linearisations are each other's transpose;
  * the step JVP agrees with central differences of forward_step in every input (state, control,
    parameters), and the trajectory JVP gives the directional derivative of the loss that the
    checkpointed reverse sweep gives (<grad L, v> = <rho_T - target, d rho_T>).
"""
import numpy as np
import pytest

import oracle as O
from ds_smoke import scenes
from ds_smoke.scenes import cell_shape, face_shape
from helpers import rel, scene_periodic, scene_solids


def _dot(a, b):
    if isinstance(a, dict):
        return sum(_dot(a[k], b[k]) for k in a)
    if isinstance(a, (tuple, list)):
        return sum(_dot(x, y) for x, y in zip(a, b))
    return float(np.sum(np.asarray(a, dtype=np.float64) * np.asarray(b, dtype=np.float64)))


def _axpy(a, b, s):
    if isinstance(a, dict):
        return {k: _axpy(a[k], b[k], s) for k in a}
    if isinstance(a, (tuple, list)):
        return tuple(_axpy(x, y, s) for x, y in zip(a, b))
    return np.asarray(a, dtype=np.float64) + s * np.asarray(b, dtype=np.float64)


def _scene(name, vorticity=False):
    sc = {"closed": lambda: scenes.plume(8), "solids": lambda: scene_solids((7, 6, 5)),
          "periodic": lambda: scene_periodic((6, 5, 4))}[name]()
    if vorticity:
        sc.vorticity = True
        sc.params = tuple(sc.params[:3]) + (2.0,)
    return sc


def _random_state(sc, rng, vel=0.3):
    n = sc.resolution
    return {"density": np.abs(rng.standard_normal(cell_shape(n))), "temperature": rng.standard_normal(cell_shape(n)),
            "velocity": tuple(vel * rng.standard_normal(face_shape(n, a)) for a in range(3))}


def _random_control(sc, rng, acc=0.3):
    n = sc.resolution
    return {"density_source": rng.standard_normal(cell_shape(n)), "temperature_source": rng.standard_normal(cell_shape(n)),
            "external_acceleration": tuple(acc * rng.standard_normal(cell_shape(n)) for _ in range(3))}


@pytest.mark.parametrize("name", ["closed", "solids", "periodic"])
def test_operator_jvps_match_differences(name):
    sc = _scene(name)
    o = O.Oracle(sc, tolerance=1e-13)
    rng = np.random.default_rng(31)
    n = sc.resolution
    faces = lambda s=1.0: tuple(s * rng.standard_normal(face_shape(n, a)) for a in range(3))
    cell = lambda: rng.standard_normal(cell_shape(n))
    eps = 1e-6
    add = lambda a, b, s: tuple(x + s * y for x, y in zip(a, b))
    u, du = faces(0.3), faces()
    # velocity self-advection (nonlinear)
    jv = o.advect_velocity_jvp(u, du)
    fd = tuple((p - m) / (2 * eps) for p, m in zip(o.advect_velocity_forward(add(u, du, eps)), o.advect_velocity_forward(add(u, du, -eps))))
    for a in range(3):
        assert rel(jv[a], fd[a]) < 1e-6, (name, a, rel(jv[a], fd[a]))
    # scalar advection in both arguments
    src, dsrc = np.abs(cell()), cell()
    jv = o.advect_scalar_jvp(0, src, dsrc, u, du)
    fd = (o.advect_scalar_forward(0, src + eps * dsrc, add(u, du, eps)) - o.advect_scalar_forward(0, src - eps * dsrc, add(u, du, -eps))) / (2 * eps)
    assert rel(jv, fd) < 1e-6, (name, rel(jv, fd))
    # constraint and projection (affine): J v = F(u + v) - F(u)
    cj = o.constrain_jvp(du)
    cd = tuple(p - m for p, m in zip(o.constrain_forward(add(u, du, 1.0)), o.constrain_forward(u)))
    for a in range(3):
        np.testing.assert_allclose(cj[a], cd[a], atol=1e-12)
    pj, _ = o.projection_jvp(du)
    pd = tuple(p - m for p, m in zip(o.projection_forward(add(u, du, 1.0))[0], o.projection_forward(u)[0]))
    for a in range(3):
        assert rel(pj[a], pd[a]) < 1e-9, (name, a, rel(pj[a], pd[a]))
    # buoyancy with parameter tangents
    rho, T, drho, dT = np.abs(cell()), cell(), cell(), cell()
    e, de = (cell(), cell(), cell()), (cell(), cell(), cell())
    params, dp = np.array([0.3, -0.2, 1.5, 0.0]), np.array([0.7, -0.4, 0.9, 0.0])
    bj = o.buoyancy_jvp(rho, T, drho, dT, de, params, dp)
    fp = o.buoyancy_forward(rho + eps * drho, T + eps * dT, add(e, de, eps), params + eps * dp)
    fm = o.buoyancy_forward(rho - eps * drho, T - eps * dT, add(e, de, -eps), params - eps * dp)
    for a in range(3):
        np.testing.assert_allclose(bj[a], (fp[a] - fm[a]) / (2 * eps), rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("name", ["closed", "solids", "periodic"])
def test_vorticity_jvp_matches_differences(name):
    sc = _scene(name, vorticity=True)
    o = O.Oracle(sc)
    rng = np.random.default_rng(32)
    n = sc.resolution
    u = tuple(0.3 * rng.standard_normal(face_shape(n, a)) for a in range(3))
    du = tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))
    conf, dconf, eps = 2.0, 0.6, 1e-6
    jv = o.vorticity_jvp(u, du, conf, dconf)
    fp = o.vorticity_forward(tuple(x + eps * y for x, y in zip(u, du)), conf + eps * dconf)
    fm = o.vorticity_forward(tuple(x - eps * y for x, y in zip(u, du)), conf - eps * dconf)
    for a in range(3):
        fd = (fp[a] - fm[a]) / (2 * eps)
        assert rel(jv[a], fd) < 1e-6, (name, a, rel(jv[a], fd))


@pytest.mark.parametrize("name,vorticity", [("closed", False), ("solids", False), ("periodic", False), ("solids", True)])
def test_step_jvp_is_transpose_of_vjp_and_matches_differences(name, vorticity):
    sc = _scene(name, vorticity)
    o = O.Oracle(sc, tolerance=1e-13)
    rng = np.random.default_rng(33)
    state, control = _random_state(sc, rng), _random_control(sc, rng)
    params = np.array(sc.params, dtype=np.float64) + np.array([0.1, 0.0, 0.0, 0.0])
    nxt, cache, _ = o.forward_step(state, control, params)
    st_t, ct_t = _random_state(sc, rng, 1.0), _random_control(sc, rng, 1.0)
    pt = np.array([0.3, -0.2, 0.5, 0.7 if vorticity else 0.0])
    jv, _ = o.jvp_step(state, nxt, cache, params, st_t, ct_t, pt)
    # <J v, w> = <v, J^T w>
    w = _random_state(sc, rng, 1.0)
    ca = {"density_source": o.zeros_cell(), "temperature_source": o.zeros_cell(),
          "external_acceleration": tuple(o.zeros_cell() for _ in range(3))}
    pa = np.zeros(4)
    prev, _ = o.vjp_step(nxt, cache, params, w, control_adj=ca, param_adj=pa, state=state)
    lhs = _dot(jv, w)
    rhs = _dot(st_t, prev) + _dot(ct_t, ca) + float(np.dot(pt, pa))
    assert abs(lhs - rhs) < 1e-9 * (abs(lhs) + 1.0), (lhs, rhs)
    # central differences in all inputs at once
    eps = 1e-6
    sp, sm = _axpy(state, st_t, eps), _axpy(state, st_t, -eps)
    cp, cm = _axpy(control, ct_t, eps), _axpy(control, ct_t, -eps)
    fp, _, _ = o.forward_step(sp, cp, params + eps * pt)
    fm, _, _ = o.forward_step(sm, cm, params - eps * pt)
    for key in ("density", "temperature"):
        fd = (fp[key] - fm[key]) / (2 * eps)
        assert rel(jv[key], fd) < 1e-5, (name, key, rel(jv[key], fd))
    for a in range(3):
        fd = (fp["velocity"][a] - fm["velocity"][a]) / (2 * eps)
        assert rel(jv["velocity"][a], fd) < 1e-5, (name, a, rel(jv["velocity"][a], fd))


@pytest.mark.parametrize("vorticity", [False, True])
def test_trajectory_jvp_gives_the_reverse_sweeps_directional_derivative(vorticity):
    sc = scenes.plume(12)
    if vorticity:
        sc.vorticity = True
        sc.params = (0.0, -0.1, 1.0, 2.0)
    o = O.Oracle(sc, tolerance=1e-13)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    steps = 4
    rng = np.random.default_rng(34)
    v = {"density": np.zeros(cell_shape(sc.resolution)), "temperature": np.zeros(cell_shape(sc.resolution)),
         "velocity": tuple(rng.standard_normal(face_shape(sc.resolution, a)) for a in range(3))}
    zero_c = {k: (tuple(np.zeros_like(x) for x in val) if isinstance(val, tuple) else np.zeros_like(val)) for k, val in ctl.items()}
    pt = np.array([0.0, 0.0, 0.0, 0.5 if vorticity else 0.0])
    fin, fin_t, info = o.trajectory_jvp(s0, ctl, sc.params, steps, v, zero_c, pt)
    assert info.forward_solves == steps and info.adjoint_solves == steps
    ref = o.loss_gradient(s0, ctl, sc.params, steps, 2, sc.target)
    np.testing.assert_allclose(fin["density"], ref["final"]["density"], rtol=1e-12, atol=1e-14)
    forward_mode = _dot(fin["density"] - sc.target, fin_t["density"])
    reverse_mode = _dot(ref["grad"]["velocity"], v["velocity"]) + float(np.dot(ref["params"], pt))
    assert abs(forward_mode - reverse_mode) < 1e-8 * abs(reverse_mode), (forward_mode, reverse_mode)


@pytest.mark.parametrize("name", ["closed", "solids", "periodic"])
def test_rbgs_mode_projection_reverse_and_tangent(name):
    """The reference's fixed-cost projection (red-black Gauss-Seidel from p = 0, 7 sweeps of each
    colour): its tangent equals projection(u + v) - projection(u) (an affine map), its reverse
    is the transpose of that map, and a full step's JVP and VJP are each other's transpose."""
    from ds_smoke import abi

    sc = _scene(name)
    o = O.Oracle(sc, solver=abi.PROJECTION_RBGS, rbgs_iterations=7)
    rng = np.random.default_rng(35)
    n = sc.resolution
    u = tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))
    du = tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))
    w = tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))
    pj, _ = o.projection_jvp(du)
    base = o.projection_forward(u)[0]
    pd = tuple(p - b for p, b in zip(o.projection_forward(tuple(x + y for x, y in zip(u, du)))[0], base))
    for a in range(3):
        assert rel(pj[a], pd[a]) < 1e-10, (name, a, rel(pj[a], pd[a]))
    pa, _ = o.projection_vjp(w)
    assert abs(_dot(pj, w) - _dot(du, pa)) < 1e-10 * abs(_dot(pj, w))
    state, control = _random_state(sc, rng), _random_control(sc, rng)
    params = np.array(sc.params, dtype=np.float64)
    nxt, cache, _ = o.forward_step(state, control, params)
    st_t, ct_t = _random_state(sc, rng, 1.0), _random_control(sc, rng, 1.0)
    pt = np.array([0.3, -0.2, 0.5, 0.0])
    jv, _ = o.jvp_step(state, nxt, cache, params, st_t, ct_t, pt)
    ws = _random_state(sc, rng, 1.0)
    ca = {"density_source": o.zeros_cell(), "temperature_source": o.zeros_cell(),
          "external_acceleration": tuple(o.zeros_cell() for _ in range(3))}
    qa = np.zeros(4)
    prev, _ = o.vjp_step(nxt, cache, params, ws, control_adj=ca, param_adj=qa, state=state)
    lhs, rhs = _dot(jv, ws), _dot(st_t, prev) + _dot(ct_t, ca) + float(np.dot(pt, qa))
    assert abs(lhs - rhs) < 1e-9 * (abs(lhs) + 1.0), (lhs, rhs)
