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
