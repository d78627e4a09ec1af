"""Pins the oracle's vorticity confinement (operators.cu:662-856; wired at operators.cpp:125-138).

The reference has no tests or fixtures for it, so:
  * the forward force is checked against closed forms: a shear u = c x^2 e_y (and its two cyclic
    rotations) has the discrete vorticity 2 c x e_z away from the clamped ends, a magnitude
    gradient along +x, so N = e_x and f = -confinement * h * 2 c x e_y; this fixes the sign and
    the component order of the curl and of N x w, the h and 0.5/h scalings, and the clamped end
    cell;
  * the VJP agrees with FP64 central differences of that forward (velocity and confinement
    adjoints), on a closed box, with colliders and with periodic boundaries;
  * the full step and the checkpointed trajectory with confinement on agree with central
    differences and across checkpoint intervals.
"""
import numpy as np
import pytest

import oracle as O
from ds_smoke import scenes
from ds_smoke.scenes import cell_shape, face_shape
from helpers import scene_periodic, scene_solids


def _dot(a, b):
    if isinstance(a, (tuple, list)):
        return sum(float(np.sum(x * y)) for x, y in zip(a, b))
    return float(np.sum(a * b))


def _box(n=(8, 9, 10), vorticity=True):
    sc = scenes.Scene(resolution=n, cell_size=1.0 / 8.0, time_step=1.0 / 16.0, name="box")
    sc.vorticity = vorticity
    return sc


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_shear_flow_force_matches_closed_form(axis):
    """u_{a+1} = c * (x_a)^2 (a cyclic): w = 2 c x_a e_{a+2}, N = e_a, f = -conf h w_{a+2} e_{a+1}."""
    sc = _box()
    n, h = sc.resolution, sc.cell_size
    o = O.Oracle(sc)
    c, conf = 1.7, 2.0
    b, w_axis = (axis + 1) % 3, (axis + 2) % 3
    u = list(o.zeros_faces())
    shp = face_shape(n, b)
    idx = np.indices(shp)                     # (z, y, x) indices of the faces of axis b
    coord = (idx[2 - axis] + 0.5) * h         # the cell-centre coordinate along `axis`
    u[b] = c * coord ** 2
    f = o.vorticity_forward(tuple(u), conf)
    m = n[axis]
    xc = (np.arange(m) + 0.5) * h
    # discrete curl: central difference of c x^2 is exact (2 c x); the first cell clamps its lower
    # neighbour to itself: (c x_1^2 - c x_0^2) / (2h) = c h
    w = 2.0 * c * xc
    w[0] = c * h
    expect = -conf * h * w                    # N = +e_axis on cells 0 .. m-3
    got = f[b]
    sl = [slice(None)] * 3
    for i in range(m - 2):
        sl[2 - axis] = i
        np.testing.assert_allclose(got[tuple(sl)], expect[i], rtol=1e-9, atol=1e-14)
    for a in (axis, w_axis):                  # the other two components vanish
        np.testing.assert_allclose(f[a], 0.0, atol=1e-14)


def test_uniform_and_rigid_translation_give_no_force():
    sc = _box()
    o = O.Oracle(sc)
    f = o.vorticity_forward(tuple(np.full(face_shape(sc.resolution, a), 0.3 * (a + 1)) for a in range(3)), 2.0)
    for a in range(3):
        np.testing.assert_allclose(f[a], 0.0, atol=1e-14)


def test_solid_cells_get_no_force():
    sc = scene_solids((7, 6, 5))
    sc.vorticity = True
    o = O.Oracle(sc)
    mask, _, _ = o.masks()
    rng = np.random.default_rng(3)
    u = tuple(rng.standard_normal(face_shape(sc.resolution, a)) for a in range(3))
    f = o.vorticity_forward(u, 2.0)
    assert mask.sum() > 0
    for a in range(3):
        assert np.all(f[a][mask != 0] == 0.0)
        assert np.abs(f[a][mask == 0]).max() > 0.0


@pytest.mark.parametrize("name", ["box", "solids", "periodic"])
def test_vorticity_vjp_matches_central_differences(name):
    sc = {"box": lambda: _box((7, 6, 5)), "solids": lambda: scene_solids((7, 6, 5)),
          "periodic": lambda: scene_periodic((6, 5, 4))}[name]()
    sc.vorticity = True
    o = O.Oracle(sc)
    rng = np.random.default_rng(11)
    n = sc.resolution
    faces = lambda s=1.0: tuple(s * rng.standard_normal(face_shape(n, a)) for a in range(3))
    u, du = faces(0.3), faces()
    fa = tuple(rng.standard_normal(cell_shape(n)) for _ in range(3))
    conf, dconf = 2.0, 0.7
    ua, ca = o.zeros_faces(), np.zeros(1)
    o.vorticity_vjp(u, conf, fa, ua, ca)
    eps = 1e-6
    add = lambda a, b, s: tuple(x + s * y for x, y in zip(a, b))
    J = lambda uu, cc: _dot(o.vorticity_forward(uu, cc), fa)
    fd_u = (J(add(u, du, eps), conf) - J(add(u, du, -eps), conf)) / (2 * eps)
    fd_c = (J(u, conf + eps * dconf) - J(u, conf - eps * dconf)) / (2 * eps)
    an_u = _dot(du, ua)
    assert abs(fd_u - an_u) < 1e-6 * abs(fd_u), (fd_u, an_u)
    assert abs(fd_c - dconf * ca[0]) < 1e-8 * abs(fd_c), (fd_c, dconf * ca[0])
    # the confinement enters linearly: its adjoint is <f(conf = 1), fa>
    assert abs(ca[0] - J(u, 1.0)) < 1e-10 * abs(ca[0])


def _random_step_inputs(sc, rng):
    n = sc.resolution
    state = {"density": np.abs(rng.standard_normal(cell_shape(n))), "temperature": rng.standard_normal(cell_shape(n)),
             "velocity": tuple(0.3 * rng.standard_normal(face_shape(n, a)) for a in range(3))}
    control = {"density_source": rng.standard_normal(cell_shape(n)), "temperature_source": rng.standard_normal(cell_shape(n)),
               "external_acceleration": tuple(0.3 * rng.standard_normal(cell_shape(n)) for _ in range(3))}
    return state, control


def test_full_step_vjp_with_confinement_matches_central_differences():
    sc = scene_solids((7, 6, 5))
    sc.vorticity = True
    o = O.Oracle(sc, tolerance=1e-13)
    rng = np.random.default_rng(12)
    n = sc.resolution
    state, control = _random_step_inputs(sc, rng)
    params = np.array([0.1, -0.1, 1.0, 2.0])
    wt = {"density": rng.standard_normal(cell_shape(n)), "temperature": rng.standard_normal(cell_shape(n)),
          "velocity": tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))}

    def J(st, prm):
        nx_, _, _ = o.forward_step(st, control, prm)
        return _dot(nx_["density"], wt["density"]) + _dot(nx_["temperature"], wt["temperature"]) + _dot(nx_["velocity"], wt["velocity"])

    nxt, cache, _ = o.forward_step(state, control, params)
    pa = np.zeros(4)
    with pytest.raises(AssertionError):      # the reverse of the confinement needs the step's input
        o.vjp_step(nxt, cache, params, wt, param_adj=np.zeros(4))
    prev, _ = o.vjp_step(nxt, cache, params, wt, param_adj=pa, state=state)
    dv = tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))
    eps = 1e-6
    pert = lambda s: dict(state, velocity=tuple(v + s * d for v, d in zip(state["velocity"], dv)))
    fd = (J(pert(eps), params) - J(pert(-eps), params)) / (2 * eps)
    an = _dot(prev["velocity"], dv)
    assert abs(fd - an) < 1e-5 * abs(fd), (fd, an)
    dp = np.array([0.0, 0.0, 0.0, 1.0])
    fd_c = (J(state, params + eps * dp) - J(state, params - eps * dp)) / (2 * eps)
    assert abs(pa[3]) > 0.0
    assert abs(fd_c - pa[3]) < 1e-5 * abs(fd_c), (fd_c, pa[3])
    # confinement changes the step: the velocity differs from the step without it
    nxt0, _, _ = o.forward_step(state, control, np.array([0.1, -0.1, 1.0, 0.0]))
    assert max(np.abs(a - b).max() for a, b in zip(nxt["velocity"], nxt0["velocity"])) > 1e-4


@pytest.mark.slow
def test_trajectory_with_confinement_matches_central_differences_and_checkpoints_agree():
    sc = scenes.plume(12)
    sc.vorticity = True
    sc.params = (0.0, -0.1, 1.0, 2.0)
    o = O.Oracle(sc, tolerance=1e-13)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    steps = 4
    ref = o.loss_gradient(s0, ctl, sc.params, steps, steps, sc.target)
    g = o.loss_gradient(s0, ctl, sc.params, steps, 1, sc.target)
    for a in range(3):
        np.testing.assert_allclose(g["grad"]["velocity"][a], ref["grad"]["velocity"][a], rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(g["params"], ref["params"], rtol=1e-10)
    # the loss is O(1e4) and the confinement's share of it small: a step of 1e-3 keeps the
    # cancellation error (1e-6 relative at 1e-3, 2e-5 at 1e-5) below the truncation error
    eps = 1e-3
    prm = np.array(sc.params, dtype=np.float64)
    lp = o.loss_gradient(s0, ctl, prm + [0, 0, 0, eps], steps, steps, sc.target)["loss"]
    lm = o.loss_gradient(s0, ctl, prm - [0, 0, 0, eps], steps, steps, sc.target)["loss"]
    fd = (lp - lm) / (2 * eps)
    assert abs(ref["params"][3]) > 0.0
    assert abs(fd - ref["params"][3]) < 1e-5 * abs(fd), (fd, ref["params"][3])
