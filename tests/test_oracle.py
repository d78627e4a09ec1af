"""Pins the FP64 oracle (oracle/smoke_oracle.cpp) against properties the mathematics fixes.

The reference has no tests, fixtures or golden vectors for this path, so each oracle function is
checked against something independent of its own formula:
  * trilinear sampling reproduces trilinear fields exactly and the RK2 trace of a uniform velocity
    is an exact shift (operators.cu:136-194,256-276);
  * the pressure matrix equals a graph Laplacian assembled here from the solid mask alone, and is
    symmetric positive definite (operators.cu:852-902, model.cpp:72-77);
  * PCG matches a dense solve, RBGS with many sweeps converges to the same solution;
  * the projection leaves a discretely divergence-free field and is idempotent;
  * every VJP is the transpose of its forward map (dot-product test; affine maps) or agrees with
    FP64 central differences (nonlinear maps), for operators, the full step and the trajectory.
A plausible mistake (a dropped term, a wrong sign or index, a transposed operand) fails one of them.
"""
import numpy as np
import pytest

import oracle as O
from ds_smoke import abi, scenes
from ds_smoke.scenes import cell_shape, face_shape
from helpers import collider, rel, scene_closed, scene_periodic, scene_solids


def small_plume(n=8):
    return scenes.plume(n)


def cell_coords(n, h):
    z, y, x = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
    return (x + 0.5) * h, (y + 0.5) * h, (z + 0.5) * h


def uniform_velocity(n, U):
    return tuple(np.full(face_shape(n, a), U[a]) for a in range(3))


# ---- sampling and trace --------------------------------------------------------------------------
def test_zero_velocity_advection_is_identity():
    sc = scene_closed(12)
    o = O.Oracle(sc)
    rng = np.random.default_rng(0)
    src = rng.standard_normal(cell_shape(sc.resolution))
    out = o.advect_scalar_forward(0, src, o.zeros_faces())
    np.testing.assert_allclose(out, src, rtol=0, atol=1e-14)


def test_uniform_velocity_trace_is_exact_shift_of_trilinear_field():
    """Interior cells: rho'(x) = f(x - dt U) exactly for a trilinear f and uniform U."""
    sc = scenes.Scene(resolution=(12, 10, 9), cell_size=0.1, time_step=0.05)
    sc.velocity_modes = (1,) * 6
    o = O.Oracle(sc)
    h, dt = sc.cell_size, sc.time_step
    U = (0.7, -1.3, 0.4)
    f = lambda x, y, z: 0.3 + 1.1 * x - 0.7 * y + 0.4 * z + 0.9 * x * y - 0.5 * y * z + 0.2 * x * z + 1.3 * x * y * z
    X, Y, Z = cell_coords(sc.resolution, h)
    out = o.advect_scalar_forward(0, f(X, Y, Z), uniform_velocity(sc.resolution, U))
    ref = f(X - dt * U[0], Y - dt * U[1], Z - dt * U[2])
    inner = (slice(2, -2), slice(2, -2), slice(2, -2))
    np.testing.assert_allclose(out[inner], ref[inner], rtol=1e-12, atol=1e-12)


def test_uniform_velocity_self_advection_is_uniform():
    sc = scenes.Scene(resolution=(8, 7, 6), cell_size=0.1, time_step=0.05)
    sc.velocity_modes = (1,) * 6
    o = O.Oracle(sc)
    U = (0.3, -0.2, 0.5)
    out = o.advect_velocity_forward(uniform_velocity(sc.resolution, U))
    for a in range(3):
        np.testing.assert_allclose(out[a], U[a], rtol=0, atol=1e-14)


def test_fixed_value_boundary_is_sampled_outside():
    """A trace that leaves through a fixed-value face samples the face value (load_scalar)."""
    sc = scenes.Scene(resolution=(6, 6, 6), cell_size=0.1, time_step=0.1)
    sc.velocity_modes = (1,) * 6
    sc.density_modes = (0, 1, 1, 1, 1, 1)
    sc.density_values = (2.5, 0, 0, 0, 0, 0)
    o = O.Oracle(sc)
    # velocity +x of 2 cells per step: cell x=0 traces to x = 0.5h - 2h < 0 -> clamped to 0,
    # whose trilinear stencil mixes the fixed x_min value 2.5 with cell x=0 (weight 1/2 each)
    U = (2.0 * sc.cell_size / sc.time_step, 0.0, 0.0)
    src = np.zeros(cell_shape(sc.resolution))
    out = o.advect_scalar_forward(0, src, uniform_velocity(sc.resolution, U))
    np.testing.assert_allclose(out[:, :, 0], 0.5 * 2.5, rtol=1e-13)
    np.testing.assert_allclose(out[:, :, 3:], 0.0, atol=1e-15)


# ---- elementwise operators (hand-computed on a 3 x 2 x 1 grid) --------------------------------------
def tiny_scene():
    sc = scenes.Scene(resolution=(3, 2, 1), cell_size=0.5, time_step=0.25)
    sc.velocity_modes = (1,) * 6
    # one solid cell: (1, 0, 0), centre (0.75, 0.25, 0.25)
    sc.colliders = [collider(abi.SHAPE_BOX, (0.75, 0.25, 0.25), (0.1, 0.1, 0.1), velocity=(2.0, 0.0, 0.0),
                             density=7.0, temperature=3.0)]
    return sc


def test_tiny_source_buoyancy_integrate():
    sc = tiny_scene()
    o = O.Oracle(sc)
    m, code, _ = o.masks()
    assert m.tolist() == [[[0, 1, 0], [0, 0, 0]]]
    rho = np.array([[[1.0, 2.0, 3.0], [4.0, 5.0, 6.0]]])
    src = np.array([[[8.0, 8.0, 8.0], [0.0, -4.0, 4.0]]])
    out = o.source_forward(rho, src)
    np.testing.assert_allclose(out, [[[3.0, 0.0, 5.0], [4.0, 4.0, 7.0]]])
    params = (0.5, -0.1, 2.0, 0.0)   # ambient, alpha_rho, alpha_T
    T = np.array([[[1.0, 1.0, 2.0], [0.0, 0.5, 1.5]]])
    e = (np.ones((1, 2, 3)), np.full((1, 2, 3), 3.0), np.zeros((1, 2, 3)))
    f = o.buoyancy_forward(rho, T, e, params)
    fy = 3.0 - 0.1 * rho + 2.0 * (T - 0.5)
    fy[0, 0, 1] = 0.0
    np.testing.assert_allclose(f[1], fy)
    assert f[0][0, 0, 1] == 0.0 and f[0][0, 1, 1] == 1.0
    # x faces (4 x 2 x 1): face (1,0) sits between fluid (0,0) and solid (1,0) -> uses only (0,0)
    u = tuple(np.zeros(face_shape(sc.resolution, a)) for a in range(3))
    fx = np.array([[[1.0, 10.0, 2.0], [3.0, 4.0, 5.0]]])
    us = o.integrate_forward(u, (fx, np.zeros((1, 2, 3)), np.zeros((1, 2, 3))))
    dt = sc.time_step
    np.testing.assert_allclose(us[0][0, 0], dt * np.array([1.0, 1.0, 2.0, 2.0]))
    np.testing.assert_allclose(us[0][0, 1], dt * np.array([3.0, 3.5, 4.5, 5.0]))


def test_tiny_constrain_and_collider_face_velocity():
    sc = tiny_scene()
    sc.velocity_modes = (0, 2, 1, 1, 1, 1)
    sc.velocity_values = tuple([0.0, 0.0, 0.0] + [0.0] * 15)
    o = O.Oracle(sc)
    cv, cd, ct = o.collider_fields()
    # x faces adjacent to the solid cell (1,0,0): faces 1 and 2 of row y=0 -> collider u_x = 2
    assert cv[0][0, 0].tolist() == [0.0, 2.0, 2.0, 0.0]
    assert cd[0, 0, 1] == 7.0 and ct[0, 0, 1] == 3.0
    u = tuple(np.full(face_shape(sc.resolution, a), 9.0) for a in range(3))
    out = o.constrain_forward(u)
    # x_min fixed (0), x_max normal-fixed (0), solid-adjacent faces take 2, others keep 9
    assert out[0][0, 0].tolist() == [0.0, 2.0, 2.0, 0.0]
    assert out[0][0, 1].tolist() == [0.0, 9.0, 9.0, 0.0]


def test_divergence_of_linear_field():
    """rhs = -h^2 div(u)/dt with div of (a x, b y, c z) = a + b + c on every fluid cell but the anchor."""
    sc = scenes.Scene(resolution=(5, 4, 3), cell_size=0.2, time_step=0.1)
    o = O.Oracle(sc)
    a, b, c = 1.5, -0.5, 2.0
    h, dt = sc.cell_size, sc.time_step
    n = sc.resolution
    u = []
    for ax, coef in enumerate((a, b, c)):
        shp = face_shape(n, ax)
        zz, yy, xx = np.meshgrid(*[np.arange(s) for s in shp], indexing="ij")
        coord = (xx, yy, zz)[ax] * h
        u.append(coef * coord)
    rhs = o.pressure_rhs(u)
    expected = np.full(cell_shape(n), -h * h * (a + b + c) / dt)
    expected.flat[0] = 0.0   # anchor
    np.testing.assert_allclose(rhs, expected, rtol=1e-12)


# ---- pressure matrix -------------------------------------------------------------------------------
def dense_matrix(o):
    code = o.masks()[1]
    idx = np.flatnonzero(code.ravel())
    A = np.zeros((idx.size, idx.size))
    for k, i in enumerate(idx):
        e = np.zeros(code.size)
        e[i] = 1.0
        A[:, k] = o.apply_matrix(e.reshape(code.shape)).ravel()[idx]
    return A, idx


def reference_laplacian(sc, mask, anchors):
    """Graph Laplacian of the fluid cells built here from the mask alone: -1 per fluid neighbour
    (periodic wrap), +1 on the diagonal per fluid neighbour or fixed-value pressure face; the
    anchor rows and columns removed."""
    nx, ny, nz = sc.resolution
    pm = sc.pressure_modes
    per = [pm[2 * d] == 2 and pm[2 * d + 1] == 2 for d in range(3)]
    fluid = [(x, y, z) for z in range(nz) for y in range(ny) for x in range(nx) if mask[z, y, x] == 0]
    lin = lambda x, y, z: x + nx * (y + ny * z)
    unknown = [lin(*c) for c in fluid if lin(*c) not in anchors]
    pos = {c: k for k, c in enumerate(unknown)}
    A = np.zeros((len(unknown), len(unknown)))
    for (x, y, z) in fluid:
        i = lin(x, y, z)
        if i not in pos:
            continue
        for d in range(3):
            for s in (-1, 1):
                c = [x, y, z]
                c[d] += s
                size = (nx, ny, nz)[d]
                if c[d] < 0 or c[d] >= size:
                    if per[d]:
                        c[d] %= size
                    else:
                        if pm[2 * d + (s > 0)] == 0:
                            A[pos[i], pos[i]] += 1.0
                        continue
                if mask[c[2], c[1], c[0]] != 0:
                    continue
                j = lin(*c)
                A[pos[i], pos[i]] += 1.0
                if j in pos:
                    A[pos[i], pos[j]] -= 1.0
    return A, np.array(unknown)


MATRIX_SCENES = {
    "closed": lambda: scenes.Scene(resolution=(4, 3, 3), cell_size=0.25, time_step=0.1),
    "solids": lambda: scene_solids((7, 6, 5)),
    "periodic": lambda: scene_periodic((6, 5, 4)),
}


@pytest.mark.parametrize("name", sorted(MATRIX_SCENES))
def test_matrix_is_independent_laplacian_and_spd(name):
    sc = MATRIX_SCENES[name]()
    o = O.Oracle(sc)
    mask, code, _ = o.masks()
    A, idx = dense_matrix(o)
    info = o.info()
    anchors = set(np.flatnonzero((mask.ravel() == 0) & (code.ravel() == 0)).tolist())
    assert (info.pressure_anchor in anchors) or info.has_fixed_pressure
    R, ridx = reference_laplacian(sc, mask, anchors)
    np.testing.assert_array_equal(idx, ridx)
    np.testing.assert_array_equal(A, R)
    np.testing.assert_array_equal(A, A.T)
    np.linalg.cholesky(A)   # positive definite
    np.testing.assert_array_equal(np.diag(A), code.ravel()[idx])


def test_pocket_gets_its_own_anchor():
    """A fluid pocket sealed by solids gets an anchor; the matrix stays SPD."""
    sc = scenes.Scene(resolution=(6, 6, 6), cell_size=1.0 / 6.0, time_step=0.05)
    h = sc.cell_size
    walls = []
    # a hollow box of solid cells around cell (3,3,3): six slabs
    for d in range(3):
        for s in (-1, 1):
            c = [3.5 * h, 3.5 * h, 3.5 * h]
            c[d] += s * h
            half = [1.4 * h, 1.4 * h, 1.4 * h]
            half[d] = 0.4 * h
            walls.append(collider(abi.SHAPE_BOX, c, half))
    sc.colliders = walls
    o = O.Oracle(sc)
    info = o.info()
    assert info.extra_anchor_count == 1
    A, _ = dense_matrix(o)
    np.linalg.cholesky(A)


@pytest.mark.parametrize("name", sorted(MATRIX_SCENES))
def test_pcg_matches_dense_solve(name):
    sc = MATRIX_SCENES[name]()
    o = O.Oracle(sc)
    A, idx = dense_matrix(o)
    rng = np.random.default_rng(1)
    b = np.zeros(int(np.prod(cell_shape(sc.resolution))))
    b[idx] = rng.standard_normal(idx.size)
    x, info = o.pcg(b.reshape(cell_shape(sc.resolution)), tol=1e-13)
    assert info.converged
    ref = np.linalg.solve(A, b[idx])
    np.testing.assert_allclose(x.ravel()[idx], ref, rtol=1e-9, atol=1e-9 * np.abs(ref).max())
    # warm start from the solution stops at once
    x2, info2 = o.pcg(b.reshape(cell_shape(sc.resolution)), x0=x, tol=1e-10)
    assert info2.iterations == 0


def test_rbgs_converges_to_pcg_solution():
    sc = scenes.Scene(resolution=(6, 5, 4), cell_size=0.2, time_step=0.1)
    o = O.Oracle(sc)
    rng = np.random.default_rng(2)
    u = tuple(rng.standard_normal(face_shape(sc.resolution, a)) for a in range(3))
    rhs = o.pressure_rhs(u)
    _, code, fixed = o.masks()
    p_pcg, _ = o.pcg(np.where(code > 0, rhs + fixed, 0.0), tol=1e-13)
    p_rbgs = o.rbgs(rhs, 4000)
    np.testing.assert_allclose(p_rbgs, p_pcg, rtol=1e-8, atol=1e-8 * np.abs(p_pcg).max())


@pytest.mark.parametrize("name", sorted(MATRIX_SCENES))
def test_projection_divergence_free_and_idempotent(name):
    sc = MATRIX_SCENES[name]()
    o = O.Oracle(sc, tolerance=1e-13)
    rng = np.random.default_rng(3)
    u = tuple(rng.standard_normal(face_shape(sc.resolution, a)) for a in range(3))
    u = o.constrain_forward(u)
    out, p, rhs, info = o.projection_forward(u)
    _, code, fixed = o.masks()
    rhs2 = o.pressure_rhs(out)
    # A p = b + fixed  ->  the projected field has zero divergence on unknowns apart from the
    # fixed-face forcing, which projection cannot remove. The wrap faces of a periodic pressure
    # axis are not projectable (operators.cu:904-910: both cells must lie inside the box), so the
    # two cells next to the seam keep part of their divergence: they are left out.
    keep = code > 0
    pm = sc.pressure_modes
    for d in range(3):
        if pm[2 * d] == 2 and pm[2 * d + 1] == 2:
            sl = [slice(None)] * 3
            for edge in (0, -1):
                sl[2 - d] = edge
                keep[tuple(sl)] = False
    resid = np.where(keep, rhs2, 0.0)
    scale = np.abs(np.where(code > 0, rhs, 0.0)).max()
    if not np.any(fixed):
        assert np.abs(resid).max() <= 1e-9 * scale
    periodic = any(pm[2 * d] == 2 and pm[2 * d + 1] == 2 for d in range(3))
    out2, _, _, _ = o.projection_forward(out)
    if not np.any(fixed) and not periodic:
        for a in range(3):
            np.testing.assert_allclose(out2[a], out[a], rtol=0, atol=1e-9 * np.abs(out[a]).max())


# ---- adjoints: dot-product tests and finite differences ------------------------------------------
def _dot(a, b):
    if isinstance(a, (tuple, list)):
        return sum(float(np.sum(x * y)) for x, y in zip(a, b))
    return float(np.sum(a * b))


@pytest.mark.parametrize("name", ["closed", "solids", "periodic"])
def test_linear_operator_vjps_are_transposes(name):
    sc = {"closed": lambda: small_plume(8), "solids": lambda: scene_solids((7, 6, 5)),
          "periodic": lambda: scene_periodic((6, 5, 4))}[name]()
    o = O.Oracle(sc, tolerance=1e-13)
    rng = np.random.default_rng(4)
    n = sc.resolution
    cell = lambda: rng.standard_normal(cell_shape(n))
    faces = lambda: tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))
    params = np.array([0.3, -0.2, 1.5, 0.0])
    # source: affine in (state, source)
    s, src, w = cell(), cell(), cell()
    sa, srca = o.zeros_cell(), o.zeros_cell()
    o.source_vjp(w, sa, srca)
    z = o.zeros_cell()
    J = lambda a, b: o.source_forward(a, b) - o.source_forward(z, z)
    assert abs(_dot(J(s, src), w) - (_dot(s, sa) + _dot(src, srca))) < 1e-10 * abs(_dot(J(s, src), w)) + 1e-12
    # buoyancy: linear in (rho, T, e) at fixed parameters; parameter adjoint by differences
    rho, T, e = cell(), cell(), (cell(), cell(), cell())
    fw = (cell(), cell(), cell())
    ra, ta, ea, pa = o.zeros_cell(), o.zeros_cell(), (o.zeros_cell(), o.zeros_cell(), o.zeros_cell()), np.zeros(4)
    o.buoyancy_vjp(rho, T, params, fw, ra, ta, ea, pa)
    zc3 = (z, z, z)
    f0 = o.buoyancy_forward(z, z, zc3, params)
    Jf = tuple(fa - fb for fa, fb in zip(o.buoyancy_forward(rho, T, e, params), f0))
    lhs = _dot(Jf, fw)
    assert abs(lhs - (_dot(rho, ra) + _dot(T, ta) + _dot(e, ea))) < 1e-10 * abs(lhs)
    for k in range(3):
        dp = np.zeros(4)
        dp[k] = 1e-6
        fd = (_dot(o.buoyancy_forward(rho, T, e, params + dp), fw) - _dot(o.buoyancy_forward(rho, T, e, params - dp), fw)) / 2e-6
        assert abs(fd - pa[k]) < 1e-6 * (abs(pa[k]) + 1.0)
    # integrate: affine in (u, f)
    u, f, wf = faces(), (cell(), cell(), cell()), faces()
    ua, fa = o.zeros_faces(), (o.zeros_cell(), o.zeros_cell(), o.zeros_cell())
    o.integrate_vjp(wf, ua, fa)
    Ji = o.integrate_forward(u, f)
    assert abs(_dot(Ji, wf) - (_dot(u, ua) + _dot(f, fa))) < 1e-10 * abs(_dot(Ji, wf))
    # constrain: affine (boundary and collider values are constants)
    zf = o.zeros_faces()
    c0 = o.constrain_forward(zf)
    Jc = tuple(a - b for a, b in zip(o.constrain_forward(u), c0))
    ca = o.zeros_faces()
    o.constrain_vjp(wf, ca)
    assert abs(_dot(Jc, wf) - _dot(u, ca)) < 1e-10 * abs(_dot(Jc, wf))
    # projection: affine in u (fixed-face values are constants)
    p0 = o.projection_forward(zf)[0]
    Jp = tuple(a - b for a, b in zip(o.projection_forward(u)[0], p0))
    pa_, _ = o.projection_vjp(wf)
    assert abs(_dot(Jp, wf) - _dot(u, pa_)) < 1e-8 * abs(_dot(Jp, wf))


@pytest.mark.parametrize("name", ["closed", "solids", "periodic"])
def test_advection_vjps_match_central_differences(name):
    sc = {"closed": lambda: small_plume(8), "solids": lambda: scene_solids((7, 6, 5)),
          "periodic": lambda: scene_periodic((6, 5, 4))}[name]()
    o = O.Oracle(sc)
    rng = np.random.default_rng(5)
    n = sc.resolution
    faces = lambda s=1.0: tuple(s * rng.standard_normal(face_shape(n, a)) for a in range(3))
    u, du, w = faces(0.3), faces(), faces()
    src, dsrc, wc = rng.standard_normal(cell_shape(n)), rng.standard_normal(cell_shape(n)), rng.standard_normal(cell_shape(n))
    eps = 1e-6
    add = lambda a, b, s: tuple(x + s * y for x, y in zip(a, b))
    # velocity self-advection
    ua = o.zeros_faces()
    o.advect_velocity_vjp(u, w, ua)
    fd = (_dot(o.advect_velocity_forward(add(u, du, eps)), w) - _dot(o.advect_velocity_forward(add(u, du, -eps)), w)) / (2 * eps)
    assert abs(fd - _dot(du, ua)) < 1e-5 * (abs(fd) + 1e-3), (fd, _dot(du, ua))
    # scalar advection, both arguments
    sa, va = o.zeros_cell(), o.zeros_faces()
    o.advect_scalar_vjp(0, src, u, wc, sa, va)
    f = lambda s_, u_: _dot(o.advect_scalar_forward(0, s_, u_), wc)
    fd_u = (f(src, add(u, du, eps)) - f(src, add(u, du, -eps))) / (2 * eps)
    fd_s = (f(src + eps * dsrc, u) - f(src - eps * dsrc, u)) / (2 * eps)
    assert abs(fd_u - _dot(du, va)) < 1e-5 * (abs(fd_u) + 1e-3), (fd_u, _dot(du, va))
    assert abs(fd_s - _dot(dsrc, sa)) < 1e-8 * (abs(fd_s) + 1e-3)


def test_full_step_vjp_matches_central_differences():
    sc = scene_solids((7, 6, 5))
    o = O.Oracle(sc, tolerance=1e-13)
    rng = np.random.default_rng(6)
    n = sc.resolution
    state = {"density": np.abs(rng.standard_normal(cell_shape(n))), "temperature": rng.standard_normal(cell_shape(n)),
             "velocity": tuple(0.3 * rng.standard_normal(face_shape(n, a)) for a in range(3))}
    control = {"density_source": rng.standard_normal(cell_shape(n)), "temperature_source": rng.standard_normal(cell_shape(n)),
               "external_acceleration": tuple(0.3 * rng.standard_normal(cell_shape(n)) for _ in range(3))}
    params = np.array(sc.params)
    wt = {"density": rng.standard_normal(cell_shape(n)), "temperature": rng.standard_normal(cell_shape(n)),
          "velocity": tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))}

    def J(st, ct):
        nx_, _, _ = o.forward_step(st, ct, params)
        return _dot(nx_["density"], wt["density"]) + _dot(nx_["temperature"], wt["temperature"]) + _dot(nx_["velocity"], wt["velocity"])

    nxt, cache, _ = o.forward_step(state, control, params)
    ca = {"density_source": o.zeros_cell(), "temperature_source": o.zeros_cell(),
          "external_acceleration": tuple(o.zeros_cell() for _ in range(3))}
    pa = np.zeros(4)
    prev, _ = o.vjp_step(nxt, cache, params, wt, control_adj=ca, param_adj=pa)
    dstate = {"density": rng.standard_normal(cell_shape(n)), "temperature": rng.standard_normal(cell_shape(n)),
              "velocity": tuple(rng.standard_normal(face_shape(n, a)) for a in range(3))}
    eps = 1e-6
    pert = lambda s: {"density": state["density"] + s * dstate["density"], "temperature": state["temperature"] + s * dstate["temperature"],
                      "velocity": tuple(v + s * dv for v, dv in zip(state["velocity"], dstate["velocity"]))}
    fd = (J(pert(eps), control) - J(pert(-eps), control)) / (2 * eps)
    an = _dot(prev["density"], dstate["density"]) + _dot(prev["temperature"], dstate["temperature"]) + _dot(prev["velocity"], dstate["velocity"])
    assert abs(fd - an) < 1e-5 * abs(fd), (fd, an)
    dctl = rng.standard_normal(cell_shape(n))
    c_p = dict(control, density_source=control["density_source"] + eps * dctl)
    c_m = dict(control, density_source=control["density_source"] - eps * dctl)
    fd_c = (J(state, c_p) - J(state, c_m)) / (2 * eps)
    assert abs(fd_c - _dot(ca["density_source"], dctl)) < 1e-5 * abs(fd_c)


@pytest.mark.slow
def test_trajectory_gradient_matches_central_differences_and_checkpoints_agree():
    sc = small_plume(12)   # at 8^3 the source sphere holds no cell centre
    o = O.Oracle(sc, tolerance=1e-13)
    s0, ctl = scenes.initial_state(sc), scenes.control(sc)
    steps = 6
    ref = o.loss_gradient(s0, ctl, sc.params, steps, steps, sc.target)
    for k in (1, 2, 4):
        g = o.loss_gradient(s0, ctl, sc.params, steps, k, sc.target)
        for a in range(3):
            np.testing.assert_allclose(g["grad"]["velocity"][a], ref["grad"]["velocity"][a], rtol=1e-10, atol=1e-14)
    rng = np.random.default_rng(7)
    d = tuple(rng.standard_normal(face_shape(sc.resolution, a)) for a in range(3))
    eps = 1e-6   # the error falls as eps^2: 3e-3 at 1e-4, 1e-5 at 1e-5, 3e-7 at 1e-6

    def loss(sign):
        s = dict(s0)
        s["velocity"] = tuple(v.astype(np.float64) + sign * eps * dv for v, dv in zip(s0["velocity"], d))
        return o.loss_gradient(s, ctl, sc.params, steps, steps, sc.target)["loss"]

    fd = (loss(1) - loss(-1)) / (2 * eps)
    an = _dot(ref["grad"]["velocity"], d)
    assert abs(an) > 0.0
    assert abs(fd - an) < 2e-6 * abs(fd), (fd, an)
