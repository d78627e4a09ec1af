"""Inverse control (ds_smoke.inverse_control: Adam on the initial velocity through the
checkpointed fwd+VJP, the optimisation loop of the 128^3 inverse-control workload) at a small size:
the loss falls, and starting each pressure solve from the previous solution (warm start) changes
the iteration counts, not the optimisation path."""
import pytest

pytestmark = pytest.mark.gpu


def test_inverse_control_reduces_the_loss_and_warm_start_keeps_the_path():
    from ds_smoke import inverse_control

    print("inverse control 32^3, 12 Adam iterations of 8 steps", flush=True)
    plain = inverse_control.run(grid=32, iterations=12, steps=8, k=4, lr=1e-2, warm_start="none", report=4)
    warm = inverse_control.run(grid=32, iterations=12, steps=8, k=4, lr=1e-2, warm_start="previous", report=4)
    print(f"  plain: {plain['initial_loss']:.6e} -> {plain['final_loss']:.6e}, iterations {plain['cg_iterations_per_solve']}")
    print(f"  warm:  {warm['initial_loss']:.6e} -> {warm['final_loss']:.6e}, iterations {warm['cg_iterations_per_solve']}")
    hist = plain["loss_every_10"]
    assert plain["final_loss"] < 0.98 * plain["initial_loss"]
    assert all(b < a for a, b in zip(hist, hist[1:]))
    assert abs(warm["final_loss"] - plain["final_loss"]) < 1e-3 * plain["final_loss"]
    assert warm["cg_iterations_per_solve"]["forward"] < plain["cg_iterations_per_solve"]["forward"]


def test_inverse_control_128_full_workload_three_iterations():
    """The 128^3 inverse-control workload (64 steps + VJP per iteration, checkpoint 8, Adam lr 1e-2)
    for 3 iterations: the loss falls at every iteration, and the warm-start network (correcting the
    sequence guess) follows the same optimisation path with fewer CG iterations than x0 = 0."""
    import os

    from ds_smoke import inverse_control

    weights = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "diff-solver_amd", "data",
                           "warm_start_plume.npz")
    print("inverse control 128^3, 3 Adam iterations of 64 steps + VJP", flush=True)
    plain = inverse_control.run(grid=128, iterations=3, steps=64, k=8, lr=1e-2, warm_start="none", report=1)
    cnn = inverse_control.run(grid=128, iterations=3, steps=64, k=8, lr=1e-2, warm_start="cnn", weights=weights, report=1)
    print(f"  plain: losses {plain['losses']}, CG per solve {plain['cg_iterations_per_solve']}, {plain['wall_seconds']:.1f} s")
    print(f"  cnn:   losses {cnn['losses']}, CG per solve {cnn['cg_iterations_per_solve']}, {cnn['wall_seconds']:.1f} s")
    assert all(b < a for a, b in zip(plain["losses"], plain["losses"][1:]))
    # Adam's first update is lr * sign(g): entries of g near 0 can take either sign under the two
    # solve sequences, so the paths agree to a small relative tolerance, not bit for bit
    for a, b in zip(plain["losses"], cnn["losses"]):
        assert abs(a - b) < 1e-4 * a
    total = lambda r: r["cg_iterations_per_solve"]["forward"] + r["cg_iterations_per_solve"]["adjoint"]
    assert total(cnn) < total(plain)
