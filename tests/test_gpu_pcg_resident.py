"""The resident single-reduction PCG (DS_CG_SINGLE_REDUCTION_RESIDENT, csrc/pcg_resident.hpp: one
cooperative launch per residual-replacement interval, vectors in registers and LDS, grid barrier
between iterations) against the streamed form of the same recurrence (DS_CG_SINGLE_REDUCTION,
one launch per iteration).

Both forms run the same recurrence with the same per-block partials reduced in the same order;
they differ in the last bit of a few FP32 operations at chunk boundaries (measured: after 3
iterations on the 64^3 plume, 2113 cells differ by at most one unit in the last place, and both
forms stay within 1.1e-7 of the FP64 oracle's textbook PCG iterates). So the solves must agree in
iteration count to within 1% and, as two solves to ||r|| <= 1e-6 ||b|| whose rounding differs, in
solution to 5e-4 (the smoothest error mode, nearly a constant here, converges last: two such
solves of the 64^3 plume differ by 4e-5), and the profiler must show that the resident kernel
ran.
Grids: 64^3 closed box, 64^3 with seeded solids, a 64 x 32 x 40 box periodic in x (all eligible:
nx a multiple of 64, ny of 8, at most 8 planes per block), and the 128^3 benchmark plume.
"""
import numpy as np
import pytest

from helpers import collider, rel

pytestmark = pytest.mark.gpu


def _scenes():
    from ds_smoke import abi, scenes

    out = {"plume64": scenes.plume(64), "plume64_solids": scenes.plume(64, solids=6)}
    h = 1.0 / 64.0
    sc = scenes.Scene(resolution=(64, 32, 40), cell_size=h, time_step=0.5 * h, name="periodic_x")
    sc.velocity_modes = (3, 3, 0, 0, 0, 0)
    sc.pressure_modes = (2, 2, 1, 1, 1, 1)
    sc.density_modes = (2, 2, 1, 1, 1, 1)
    sc.temperature_modes = (2, 2, 1, 1, 1, 1)
    sc.colliders = [collider(abi.SHAPE_ELLIPSOID, (0.5, 0.2, 0.3), (0.1, 0.08, 0.12))]
    out["periodic_x"] = sc
    return out


SCENE_NAMES = ["plume64", "plume64_solids", "periodic_x"]


def _pair(sc):
    from ds_smoke import abi
    from ds_smoke.solver import Context

    res = Context(sc.config(tolerance=1e-6, cg_variant=abi.CG_SINGLE_REDUCTION_RESIDENT), abi.MODE_DIFFERENTIABLE)
    stm = Context(sc.config(tolerance=1e-6, cg_variant=abi.CG_SINGLE_REDUCTION), abi.MODE_DIFFERENTIABLE)
    return res, stm


@pytest.mark.parametrize("name", SCENE_NAMES)
def test_resident_solve_matches_the_streamed_solve(name):
    sc = _scenes()[name]
    print(f"resident vs streamed PCG: {name} {sc.resolution}", flush=True)
    res, stm = _pair(sc)
    _, code = res.masks()
    rng = np.random.default_rng(5)
    for trial in range(2):
        b = np.where(code > 0, rng.standard_normal(code.shape), 0.0).astype(np.float32)
        res.profile_enable(1)
        x1, i1, s1 = res.pcg_solve(res.to_device(b))
        prof = res.profile_read()
        res.profile_enable(0)
        x2, i2, s2 = stm.pcg_solve(stm.to_device(b))
        print(f"  trial {trial}: resident {i1.iterations} iterations, streamed {i2.iterations}; kernels {sorted(prof)}")
        assert s1 == 0 and s2 == 0 and i1.converged and i2.converged
        assert "pcg_resident_kernel" in prof and prof["pcg_resident_kernel"][0] > 0
        assert abs(int(i1.iterations) - int(i2.iterations)) <= max(2, 0.01 * i2.iterations)
        assert i1.final_residual <= 1e-6 * i1.rhs_norm * (1 + 1e-9)
        err = rel(res.to_host(x1), stm.to_host(x2))
        print(f"  solution difference {err:.2e}")
        assert err < 5e-4
    res.close()
    stm.close()


def test_resident_trajectory_gradient_matches_the_streamed_one():
    from ds_smoke import scenes

    sc = scenes.plume(64)
    print("resident vs streamed: plume 64^3, 6 steps + VJP, k = 3", flush=True)
    res, stm = _pair(sc)
    s0, kk = scenes.initial_state(sc), scenes.control(sc)
    out = []
    for ctx in (res, stm):
        loss, grad, info = ctx.loss_gradient(ctx.state_from_host(s0), ctx.control_from_host(kk), ctx.allocate_parameters(sc.params), 6,
                                             3, ctx.to_device(sc.target))
        assert info.not_converged == 0
        out.append((loss, ctx.to_host(grad), info))
    (l1, g1, n1), (l2, g2, n2) = out
    print(f"  loss {l1!r} / {l2!r}; iterations fwd {n1.forward_iterations} / {n2.forward_iterations}, "
          f"adj {n1.adjoint_iterations} / {n2.adjoint_iterations}")
    assert abs(l1 - l2) <= 1e-6 * abs(l2)
    assert abs(int(n1.forward_iterations) - int(n2.forward_iterations)) <= 0.02 * n2.forward_iterations
    assert abs(int(n1.adjoint_iterations) - int(n2.adjoint_iterations)) <= 0.02 * n2.adjoint_iterations
    for a in range(3):
        assert rel(g1["velocity"][a], g2["velocity"][a]) < 1e-5
    res.close()
    stm.close()


def test_resident_solve_at_128():
    from ds_smoke import scenes

    sc = scenes.plume(128)
    print("resident vs streamed PCG: plume 128^3 forward step", flush=True)
    res, stm = _pair(sc)
    s0, kk = scenes.initial_state(sc), scenes.control(sc)
    results = []
    for ctx in (res, stm):
        st, ctl = ctx.state_from_host(s0), ctx.control_from_host(kk)
        nxt, _, info = ctx.forward_step(st, ctl, ctx.allocate_parameters(sc.params))
        results.append((ctx.to_host(nxt), info))
    (a, ia), (b, ib) = results
    print(f"  iterations resident {ia.iterations}, streamed {ib.iterations}")
    assert abs(int(ia.iterations) - int(ib.iterations)) <= max(2, 0.01 * ib.iterations) and ia.converged
    assert rel(a["density"], b["density"]) < 1e-6
    for k in range(3):
        assert rel(a["velocity"][k], b["velocity"][k]) < 1e-4
    res.close()
    stm.close()
