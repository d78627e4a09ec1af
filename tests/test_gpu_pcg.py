"""PCG on the projection matrix (solvers/smoke/operators.cu:852-902): Jacobi-preconditioned in three
recurrences (the single-reduction form, one fused kernel per iteration; the two-reduction form; the
s-step form, two iterations per fused kernel and per reduction) and preconditioned by the
aggregation multigrid V-cycle (DS_CG_MULTIGRID).

Checked against the FP64 oracle's solve of the same system on the three scenes, plus the
solver's own contract: the true residual meets the tolerance, a guess equal to the solution
stops at once, a non-finite guess is rejected, b = 0 gives x = 0, an iteration cap that is too
small reports DS_ERROR_NOT_CONVERGED (an odd cap included), the Jacobi recurrences need about the
same iteration count and the multigrid form at most half of it.
"""
import numpy as np
import pytest

from helpers import SCENES, rel

pytestmark = pytest.mark.gpu

VARIANTS = {"single": 0, "two": 1, "s2": 3, "mg": 4}
# the recurrence each variant's iteration count is compared with
OTHER = {"single": 1, "two": 0, "s2": 0, "mg": 0}


@pytest.fixture(scope="module", params=[(s, v) for s in sorted(SCENES) for v in sorted(VARIANTS)],
                ids=lambda p: f"{p[0]}-{p[1]}")
def solver(request):
    import oracle as O
    from ds_smoke import abi
    from ds_smoke.solver import Context

    scene, variant = request.param
    sc = SCENES[scene]()
    ctx = Context(sc.config(tolerance=1e-6, cg_variant=VARIANTS[variant]), abi.MODE_FORWARD)
    orc = O.Oracle(sc, tolerance=1e-12)
    _, code = ctx.masks()
    rng = np.random.default_rng(11)
    b = np.where(code > 0, rng.standard_normal(code.shape), 0.0).astype(np.float32)
    return scene, variant, sc, ctx, orc, code, b


def test_solution_matches_oracle(solver):
    scene, variant, sc, ctx, orc, code, b = solver
    print(f"pcg {scene} {variant}", flush=True)
    x, info, st = ctx.pcg_solve(ctx.to_device(b))
    assert st == 0 and info.converged
    assert info.final_residual <= 1e-6 * info.rhs_norm * (1 + 1e-9)
    ref, rinfo = orc.pcg(b.astype(np.float64), tol=1e-12)
    # (the output x is the solver's compensated pair rounded to FP32: that rounding alone leaves a
    # residual above 1e-6 ||b||, so the residual bound is the solver's own FP64 check above)
    assert rel(ctx.to_host(x), ref) < 5e-5, rel(ctx.to_host(x), ref)


def test_variants_take_similar_iterations(solver):
    scene, variant, sc, ctx, orc, code, b = solver
    from ds_smoke import abi
    from ds_smoke.solver import Context

    other = Context(sc.config(tolerance=1e-6, cg_variant=OTHER[variant]), abi.MODE_FORWARD)
    _, i1, _ = ctx.pcg_solve(ctx.to_device(b))
    _, i2, _ = other.pcg_solve(other.to_device(b))
    print(f"{scene}: {variant} {i1.iterations} iterations, other {i2.iterations}")
    if variant == "mg":
        assert 0 < i1.iterations <= 0.5 * i2.iterations
    else:
        assert abs(int(i1.iterations) - int(i2.iterations)) <= max(3, 0.05 * i2.iterations)


def test_guess_equal_to_solution_stops_at_once(solver):
    scene, variant, sc, ctx, orc, code, b = solver
    x, info, _ = ctx.pcg_solve(ctx.to_device(b))
    x2, info2, st = ctx.pcg_solve(ctx.to_device(b), x0=x)
    assert st == 0 and info2.converged and info2.warm_started and not info2.guess_rejected
    # x is the solution pair rounded to FP32: its residual can sit just above the tolerance
    assert info2.iterations <= 3


def test_non_finite_guess_is_rejected(solver):
    scene, variant, sc, ctx, orc, code, b = solver
    bad = np.full(code.shape, np.nan, dtype=np.float32)
    x, info, st = ctx.pcg_solve(ctx.to_device(b), x0=ctx.to_device(bad))
    assert st == 0 and info.converged and info.guess_rejected
    assert np.all(np.isfinite(ctx.to_host(x)))


def test_zero_rhs_gives_zero(solver):
    scene, variant, sc, ctx, orc, code, b = solver
    x, info, st = ctx.pcg_solve(ctx.to_device(np.zeros_like(b)))
    assert st == 0 and info.converged and info.iterations == 0
    assert not np.any(ctx.to_host(x))


def test_iteration_cap_reports_not_converged(solver):
    scene, variant, sc, ctx, orc, code, b = solver
    from ds_smoke import abi

    cap = 1 if variant == "mg" else 3
    x, info, st = ctx.pcg_solve(ctx.to_device(b), max_iterations=cap, check=False)
    assert st == abi.DS_ERROR_NOT_CONVERGED
    assert not info.converged and info.iterations == cap
    assert np.isfinite(info.final_residual) and info.final_residual > 1e-6 * info.rhs_norm
