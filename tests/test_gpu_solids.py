"""Solid obstacles on the GPU: the hand-computed bisection and rasterisation cases, and the
8-sphere plume (the 256^3 obstacle workload) against the oracle at a reduced size and by its own
properties at full size.

* The slab cases of tests/test_oracle_solids.py run through the HIP kernels: bisected end points
  (trace_rk2's 8 halvings, operators.cu:256-276), the derivative lo times the clamp mask, the
  collider value in solid cells (operators.cu:499-507) and the inclusive ellipsoid and box
  membership of the rasterisation (model.cpp:17-26).
* 48^3 with the 8 seeded spheres scaled to the grid: 8 steps forward plus the checkpointed loss
  gradient (k = 4) against the FP64 oracle.
* 256^3 with 8 spheres of radius 8-20 cells, 32 steps forward plus the VJP (k = 8): every forward
  and adjoint solve converges to 1e-6, and forward mode along a seeded velocity direction gives the
  directional derivative <g, v> of the reverse sweep's gradient.
* A cell size that is not a power of two (h = 1/96) with solids: one forward step and its VJP
  against the oracle (positions divide by h, as operators.cu:137-139 does).
"""
import numpy as np
import pytest

from helpers import collider, rel
from test_oracle_solids import CENTRE, COLLIDER_DENSITY, DT, LO, RES, cells_of, linear_field, rast_scene, slab_scene

pytestmark = pytest.mark.gpu


def near_solid_faces(mask, reach):
    """Per axis, the faces (z, y, x arrays) whose adjacent cells lie within `reach` cells (max norm)
    of a solid cell."""
    grown = mask.astype(bool)
    for _ in range(reach):
        for ax in range(3):   # one axis after the other: the cube of side 2 reach + 1
            g = grown.copy()
            g[(slice(None),) * ax + (slice(1, None),)] |= grown[(slice(None),) * ax + (slice(None, -1),)]
            g[(slice(None),) * ax + (slice(None, -1),)] |= grown[(slice(None),) * ax + (slice(1, None),)]
            grown = g
    out = []
    for a in range(3):   # face axis a: numpy axis 2 - a; face i sits between cells i - 1 and i
        ax = 2 - a
        pad_lo = np.concatenate([np.take(grown, [0], axis=ax), grown], axis=ax)
        pad_hi = np.concatenate([grown, np.take(grown, [-1], axis=ax)], axis=ax)
        out.append(pad_lo | pad_hi)
    return out


def _ctx(sc, **kw):
    from ds_smoke import abi
    from ds_smoke.solver import Context

    return Context(sc.config(tolerance=1e-6, **kw), abi.MODE_DIFFERENTIABLE)


def test_slab_bisection_on_gpu():
    print("slab bisection", flush=True)
    sc, vel = slab_scene()
    ctx = _ctx(sc)
    u = tuple(ctx.to_device(v.astype(np.float32)) for v in vel)
    src = ctx.to_device(linear_field().astype(np.float32))
    out = ctx.to_host(ctx.advect_scalar_forward(0, src, u)).astype(np.float64)
    for x, lo in LO.items():
        np.testing.assert_allclose(out[:, :, x], x + 0.5 + lo * 3.375, rtol=0, atol=2e-6, err_msg=f"cell x={x}")
    for x in range(0, 7):
        np.testing.assert_allclose(out[:, :, x], x + 0.5 + 3.375, rtol=0, atol=2e-6)
    assert np.all(out[:, :, 10:] == COLLIDER_DENSITY)
    for x, lo in LO.items():
        oa = np.zeros(out.shape, dtype=np.float32)
        oa[1, 1, x] = 1.0
        src_adj = ctx.zeros_cell()
        ua = ctx.zeros_faces()
        ctx.advect_scalar_vjp(0, src, u, ctx.to_device(oa), src_adj, ua)
        g = [ctx.to_host(a).astype(np.float64) for a in ua]
        np.testing.assert_allclose(g[0].sum(), -lo * DT, rtol=1e-6)
        assert np.abs(g[1]).max() == 0.0 and np.abs(g[2]).max() == 0.0
        np.testing.assert_allclose(ctx.to_host(src_adj).astype(np.float64).sum(), 1.0, rtol=1e-6)
    # solid cells carry no adjoint
    oa = np.zeros(out.shape, dtype=np.float32)
    oa[:, :, 10:] = 1.0
    src_adj = ctx.zeros_cell()
    ua = ctx.zeros_faces()
    ctx.advect_scalar_vjp(0, src, u, ctx.to_device(oa), src_adj, ua)
    assert not ctx.to_host(src_adj).any() and not any(ctx.to_host(a).any() for a in ua)


def test_slab_bisection_with_clamped_axis_on_gpu():
    sc, vel = slab_scene(vy=6.0)
    ctx = _ctx(sc)
    u = tuple(ctx.to_device(v.astype(np.float32)) for v in vel)
    src = ctx.to_device(linear_field(1.0, 1.0).astype(np.float32))
    out = ctx.to_host(ctx.advect_scalar_forward(0, src, u)).astype(np.float64)
    lo = LO[7]
    np.testing.assert_allclose(out[1, 2, 7], (7.5 + lo * 3.375) + 2.5 * (1.0 - lo), rtol=0, atol=2e-6)
    oa = np.zeros(out.shape, dtype=np.float32)
    oa[1, 2, 7] = 1.0
    src_adj = ctx.zeros_cell()
    ua = ctx.zeros_faces()
    ctx.advect_scalar_vjp(0, src, u, ctx.to_device(oa), src_adj, ua)
    np.testing.assert_allclose(ctx.to_host(ua[0]).astype(np.float64).sum(), -lo * DT, rtol=1e-6)
    assert np.abs(ctx.to_host(ua[1])).max() == 0.0


def test_rasterisation_on_gpu():
    from ds_smoke import abi

    m, _ = _ctx(rast_scene([collider(abi.SHAPE_ELLIPSOID, CENTRE, (0.5, 0.25, 0.25))])).masks()
    assert cells_of(m) == sorted([(1, 2, 2), (2, 2, 2), (3, 2, 2), (4, 2, 2), (5, 2, 2), (3, 1, 2), (3, 3, 2), (3, 2, 1), (3, 2, 3)])
    m, _ = _ctx(rast_scene([collider(abi.SHAPE_BOX, CENTRE, (0.25, 0.0, 0.25))])).masks()
    assert cells_of(m) == sorted((x, 2, z) for x in (2, 3, 4) for z in (1, 2, 3))


def test_spheres_48_gradient_matches_oracle():
    """The 8-sphere plume at 48^3: 8 steps, loss gradient with a checkpoint every 4 steps."""
    import oracle as O
    from ds_smoke import scenes

    sc = scenes.plume(48, solids=8)
    assert len(sc.colliders) == 8
    print("spheres 48^3: 8 steps + VJP, k = 4", flush=True)
    ctx = _ctx(sc)
    orc = O.Oracle(sc, tolerance=1e-10)
    m, _ = ctx.masks()
    om, _, _ = orc.masks()
    np.testing.assert_array_equal(m, om)
    assert m.sum() > 0
    s0, k = scenes.initial_state(sc), scenes.control(sc)
    params = ctx.allocate_parameters(sc.params)
    pa = ctx.allocate_parameter_adjoint()
    loss, grad, info = ctx.loss_gradient(ctx.state_from_host(s0), ctx.control_from_host(k), params, 8, 4,
                                         ctx.to_device(sc.target), param_adj=pa)
    assert info.not_converged == 0
    ref = orc.loss_gradient(s0, k, sc.params, 8, 4, sc.target)
    fin, _ = ctx.simulate(ctx.state_from_host(s0), ctx.control_from_host(k), params, 8)
    ref_fin, _ = orc.simulate(s0, k, sc.params, 8)
    e_rho = rel(ctx.to_host(fin)["density"], ref_fin["density"])
    e_loss = abs(loss - ref["loss"]) / ref["loss"]
    g = ctx.to_host(grad)
    e_grad = [rel(g["velocity"][a], ref["grad"]["velocity"][a]) for a in range(3)]
    e_par = rel(ctx.to_host(pa)[:3], ref["params"][:3])
    # faces within 2 cells of a solid: where a trace can end inside a sphere and be bisected
    near = near_solid_faces(m, 2)
    e_far = [rel(np.where(near[a], 0.0, g["velocity"][a]), np.where(near[a], 0.0, ref["grad"]["velocity"][a])) for a in range(3)]
    diff = [np.abs(g["velocity"][a].astype(np.float64) - ref["grad"]["velocity"][a]) for a in range(3)]
    worst = [np.unravel_index(int(np.argmax(d)), d.shape) for d in diff]
    print(f"  density {e_rho:.2e}  loss {e_loss:.2e}  velocity gradient {max(e_grad):.2e} "
          f"(faces more than 2 cells from a solid {max(e_far):.2e})  parameters {e_par:.2e}")
    for a in range(3):
        print(f"  axis {a}: largest difference {diff[a].max():.3e} at (z, y, x) = {tuple(int(v) for v in worst[a])}, "
              f"near a solid: {bool(near[a][worst[a]])}; gradient scale {np.abs(ref['grad']['velocity'][a]).max():.3e}")
    assert e_rho < 1e-5
    assert e_loss < 1e-5
    assert max(e_grad) < 1e-3, e_grad
    assert e_par < 1e-4


def test_spheres_256_trajectory_converges_and_forward_mode_agrees():
    """The 256^3 obstacle workload: 32 steps + VJP (k = 8) with the extrapolated warm start."""
    from ds_smoke import abi, scenes

    sc = scenes.plume(256, solids=8)
    print("spheres 256^3: 32 steps + VJP, k = 8", flush=True)
    ctx = _ctx(sc, warm_start=abi.WARM_START_EXTRAPOLATE)
    s0h, kh = scenes.initial_state(sc), scenes.control(sc)
    s0, k = ctx.state_from_host(s0h), ctx.control_from_host(kh)
    params = ctx.allocate_parameters(sc.params)
    loss, grad, info = ctx.loss_gradient(s0, k, params, 32, 8, ctx.to_device(sc.target))
    assert info.forward_solves == 32 and info.adjoint_solves == 32
    assert info.not_converged == 0
    print(f"  loss {loss:.6e}; CG per forward solve {info.forward_iterations / 32:.0f}, "
          f"per adjoint solve {info.adjoint_iterations / 32:.0f}", flush=True)
    g = ctx.to_host(grad)["velocity"]
    rng = np.random.default_rng(256)
    v = tuple(rng.standard_normal(x.shape).astype(np.float32) for x in s0h["velocity"])
    zero_cell = np.zeros_like(s0h["density"])
    tangent = ctx.state_from_host({"density": zero_cell, "temperature": zero_cell, "velocity": v})
    zero_control = ctx.control_from_host({"density_source": zero_cell, "temperature_source": zero_cell,
                                          "external_acceleration": (zero_cell, zero_cell, zero_cell)})
    fin, fin_t, tinfo = ctx.trajectory_jvp(s0, k, params, 32, tangent, zero_control, ctx.allocate_parameters((0.0,) * 4))
    rho = ctx.to_host(fin)["density"].astype(np.float64)
    d_rho = ctx.to_host(fin_t)["density"].astype(np.float64)
    forward = float(np.sum((rho - sc.target) * d_rho))
    reverse = float(sum(np.sum(g[a].astype(np.float64) * v[a]) for a in range(3)))
    err = abs(forward - reverse) / abs(reverse)
    print(f"  <g, v> reverse {reverse:.8e} forward {forward:.8e} relative {err:.2e}")
    assert err < 1e-4


def test_non_power_of_two_cell_size_with_solids():
    """h = 1/96: one forward step and its VJP against the oracle."""
    import oracle as O
    from ds_smoke import abi, scenes

    h = 1.0 / 96.0
    sc = scenes.Scene(resolution=(24, 20, 16), cell_size=h, time_step=0.5 * h, name="h96")
    sc.colliders = [collider(abi.SHAPE_ELLIPSOID, (12 * h, 9 * h, 8 * h), (4.3 * h, 3.7 * h, 3.1 * h)),
                    collider(abi.SHAPE_BOX, (5 * h, 15 * h, 6 * h), (2.2 * h, 1.9 * h, 2.6 * h), velocity=(0.1, 0.0, -0.05))]
    rng = np.random.default_rng(96)
    from ds_smoke.scenes import cell_shape, face_shape

    n = sc.resolution
    s = {"density": np.abs(rng.standard_normal(cell_shape(n))).astype(np.float32),
         "temperature": rng.standard_normal(cell_shape(n)).astype(np.float32),
         "velocity": tuple((0.5 * rng.standard_normal(face_shape(n, a))).astype(np.float32) for a in range(3))}
    z = np.zeros(cell_shape(n), dtype=np.float32)
    k = {"density_source": z, "temperature_source": z, "external_acceleration": (z, z, z)}
    ctx = _ctx(sc)
    orc = O.Oracle(sc, tolerance=1e-10)
    np.testing.assert_array_equal(ctx.masks()[0], orc.masks()[0])
    params = ctx.allocate_parameters(sc.params)
    st, kt = ctx.state_from_host(s), ctx.control_from_host(k)
    nxt, cache, info = ctx.forward_step(st, kt, params)
    ref, ref_cache, _ = orc.forward_step(s, k, sc.params)
    got = ctx.to_host(nxt)
    errs = [rel(got["density"], ref["density"])] + [rel(got["velocity"][a], ref["velocity"][a]) for a in range(3)]
    adj = {"density": rng.standard_normal(cell_shape(n)).astype(np.float32), "temperature": z,
           "velocity": tuple(rng.standard_normal(face_shape(n, a)).astype(np.float32) for a in range(3))}
    prev, _ = ctx.vjp_step(st, kt, params, nxt, cache, ctx.state_from_host(adj))
    ref_prev, _ = orc.vjp_step(ref, ref_cache, sc.params, adj)
    gp = ctx.to_host(prev)
    aerrs = [rel(gp["velocity"][a], ref_prev["velocity"][a]) for a in range(3)]
    print(f"h = 1/96: forward {max(errs):.2e}, reverse {max(aerrs):.2e}")
    assert max(errs) < 2e-5, errs
    assert max(aerrs) < 2e-4, aerrs
