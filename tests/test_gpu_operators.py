"""Per-operator parity of the HIP kernels (through the C ABI) against the FP64 oracle.

Each operator of the smoke step and its VJP (solvers/smoke/operators.cu:286-561,875-969) runs on
seeded random inputs on three scenes: the closed plume box, a non-cubic grid with solid colliders,
mixed velocity boundaries and a fixed-value pressure face, and a grid periodic in x. Tolerances
are relative L2 between the FP32 GPU result and the FP64 oracle.
"""
import numpy as np
import pytest

from helpers import SCENES, random_fields, rel

pytestmark = pytest.mark.gpu

FWD_TOL = 2e-6
ADV_TOL = 5e-5     # trilinear gathers after an FP32 trace
ADJ_TOL = 1e-4     # FP32 atomics against FP64 scatters


@pytest.fixture(scope="module", params=sorted(SCENES))
def setup(request):
    import oracle as O
    from ds_smoke import abi
    from ds_smoke.solver import Context

    sc = SCENES[request.param]()
    ctx = Context(sc.config(tolerance=1e-6), abi.MODE_DIFFERENTIABLE)
    orc = O.Oracle(sc, tolerance=1e-10)
    return request.param, sc, ctx, orc, random_fields(sc, seed=7)


def test_domain_matches_oracle(setup):
    name, sc, ctx, orc, f = setup
    m, c = ctx.masks()
    om, oc, _ = orc.masks()
    np.testing.assert_array_equal(m, om)
    np.testing.assert_array_equal(c, oc)
    gi, oi = ctx.domain, orc.info()
    assert (gi.unknown_count, gi.solid_count, gi.pressure_anchor, gi.extra_anchor_count) == \
        (oi.unknown_count, oi.solid_count, oi.pressure_anchor, oi.extra_anchor_count)


def test_source(setup):
    name, sc, ctx, orc, f = setup
    got = ctx.to_host(ctx.source_forward(ctx.to_device(f["rho"]), ctx.to_device(f["s_rho"])))
    assert rel(got, orc.source_forward(f["rho"], f["s_rho"])) < FWD_TOL
    sa, srca = ctx.zeros_cell(), ctx.zeros_cell()
    ctx.source_vjp(ctx.to_device(f["cell_adj"]), sa, srca)
    osa, osrca = orc.zeros_cell(), orc.zeros_cell()
    orc.source_vjp(f["cell_adj"], osa, osrca)
    assert rel(ctx.to_host(sa), osa) < FWD_TOL and rel(ctx.to_host(srca), osrca) < FWD_TOL


def test_buoyancy(setup):
    name, sc, ctx, orc, f = setup
    params = ctx.allocate_parameters(sc.params)
    got = ctx.to_host(ctx.buoyancy_forward(ctx.to_device(f["rho"]), ctx.to_device(f["T"]),
                                           tuple(ctx.to_device(e) for e in f["e"]), params))
    ref = orc.buoyancy_forward(f["rho"], f["T"], f["e"], sc.params)
    for a in range(3):
        assert rel(got[a], ref[a]) < FWD_TOL
    ra, ta = ctx.zeros_cell(), ctx.zeros_cell()
    ea = tuple(ctx.zeros_cell() for _ in range(3))
    pa = ctx.allocate_parameter_adjoint()
    fa = tuple(ctx.to_device(x) for x in f["force"])
    ctx.buoyancy_vjp(ctx.to_device(f["rho"]), ctx.to_device(f["T"]), params, fa, ra, ta, ea, pa)
    ora, ota = orc.zeros_cell(), orc.zeros_cell()
    oea = tuple(orc.zeros_cell() for _ in range(3))
    opa = np.zeros(4)
    orc.buoyancy_vjp(f["rho"], f["T"], sc.params, f["force"], ora, ota, oea, opa)
    assert rel(ctx.to_host(ra), ora) < FWD_TOL and rel(ctx.to_host(ta), ota) < FWD_TOL
    for a in range(3):
        assert rel(ctx.to_host(ea[a]), oea[a]) < FWD_TOL
    np.testing.assert_allclose(ctx.to_host(pa)[:3], opa[:3], rtol=1e-5, atol=1e-9)


def test_integrate(setup):
    name, sc, ctx, orc, f = setup
    u = tuple(ctx.to_device(v) for v in f["vel"])
    fc = tuple(ctx.to_device(x) for x in f["force"])
    got = ctx.to_host(ctx.integrate_forward(u, fc))
    ref = orc.integrate_forward(f["vel"], f["force"])
    for a in range(3):
        assert rel(got[a], ref[a]) < FWD_TOL
    ua, fa = ctx.zeros_faces(), tuple(ctx.zeros_cell() for _ in range(3))
    ctx.integrate_vjp(tuple(ctx.to_device(x) for x in f["face_adj"]), ua, fa)
    oua, ofa = orc.zeros_faces(), tuple(orc.zeros_cell() for _ in range(3))
    orc.integrate_vjp(f["face_adj"], oua, ofa)
    for a in range(3):
        assert rel(ctx.to_host(ua[a]), oua[a]) < FWD_TOL
        assert rel(ctx.to_host(fa[a]), ofa[a]) < FWD_TOL


def test_advect_velocity(setup):
    name, sc, ctx, orc, f = setup
    u = tuple(ctx.to_device(v) for v in f["vel"])
    got = ctx.to_host(ctx.advect_velocity_forward(u))
    ref = orc.advect_velocity_forward(f["vel"])
    for a in range(3):
        assert rel(got[a], ref[a]) < ADV_TOL, (name, a, rel(got[a], ref[a]))
    ua = ctx.zeros_faces()
    ctx.advect_velocity_vjp(u, tuple(ctx.to_device(x) for x in f["face_adj"]), ua)
    oua = orc.zeros_faces()
    orc.advect_velocity_vjp(f["vel"], f["face_adj"], oua)
    for a in range(3):
        assert rel(ctx.to_host(ua[a]), oua[a]) < ADJ_TOL, (name, a, rel(ctx.to_host(ua[a]), oua[a]))


def test_constrain(setup):
    name, sc, ctx, orc, f = setup
    u = tuple(ctx.to_device(v) for v in f["vel"])
    got = ctx.to_host(ctx.constrain_forward(u))
    ref = orc.constrain_forward(f["vel"])
    for a in range(3):
        assert rel(got[a], ref[a]) < FWD_TOL
    ua = ctx.zeros_faces()
    ctx.constrain_vjp(tuple(ctx.to_device(x) for x in f["face_adj"]), ua)
    oua = orc.zeros_faces()
    orc.constrain_vjp(f["face_adj"], oua)
    for a in range(3):
        assert rel(ctx.to_host(ua[a]), oua[a]) < FWD_TOL


@pytest.mark.parametrize("field", [0, 1])
def test_advect_scalar(setup, field):
    name, sc, ctx, orc, f = setup
    u = tuple(ctx.to_device(v) for v in f["vel"])
    src = ctx.to_device(f["rho"])
    got = ctx.to_host(ctx.advect_scalar_forward(field, src, u))
    ref = orc.advect_scalar_forward(field, f["rho"], f["vel"])
    assert rel(got, ref) < ADV_TOL, (name, rel(got, ref))
    sa, ua = ctx.zeros_cell(), ctx.zeros_faces()
    ctx.advect_scalar_vjp(field, src, u, ctx.to_device(f["cell_adj"]), sa, ua)
    osa, oua = orc.zeros_cell(), orc.zeros_faces()
    orc.advect_scalar_vjp(field, f["rho"], f["vel"], f["cell_adj"], osa, oua)
    assert rel(ctx.to_host(sa), osa) < ADJ_TOL
    for a in range(3):
        assert rel(ctx.to_host(ua[a]), oua[a]) < ADJ_TOL, (name, a, rel(ctx.to_host(ua[a]), oua[a]))


def test_apply_matrix(setup):
    name, sc, ctx, orc, f = setup
    _, code = ctx.masks()
    x = np.where(code > 0, f["T"], 0).astype(np.float32)
    got = ctx.to_host(ctx.apply_matrix(ctx.to_device(x)))
    assert rel(got, orc.apply_matrix(x)) < FWD_TOL


def test_projection(setup):
    name, sc, ctx, orc, f = setup
    u = tuple(ctx.to_device(v) for v in f["vel"])
    out, p, rhs, info = ctx.projection_forward(u)
    assert info.converged and info.final_residual <= 1e-6 * info.rhs_norm
    ref_out, ref_p, ref_rhs, _ = orc.projection_forward(f["vel"])
    _, code, fixed = orc.masks()
    b_ref = np.where(code > 0, ref_rhs + fixed, 0)
    assert rel(ctx.to_host(rhs), b_ref) < FWD_TOL
    assert rel(ctx.to_host(p), ref_p) < 1e-4
    for a in range(3):
        assert rel(ctx.to_host(out[a]), ref_out[a]) < 1e-5
    ua, ainfo = ctx.projection_vjp(tuple(ctx.to_device(x) for x in f["face_adj"]))
    assert ainfo.converged
    oua, _ = orc.projection_vjp(f["face_adj"])
    for a in range(3):
        assert rel(ctx.to_host(ua[a]), oua[a]) < 1e-4, (name, a, rel(ctx.to_host(ua[a]), oua[a]))
