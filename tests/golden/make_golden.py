"""Writes the golden fixtures under tests/golden/ from the FP64 oracle (oracle/smoke_oracle.cpp).

The reference ships no tests, fixtures or golden vectors for the smoke solver, so these vectors
come from the oracle, whose own correctness is pinned by tests/test_oracle.py. They let GPU tests
check long runs that the oracle cannot repeat within a test's time budget.

plume64_100.npz -- the 64^3 plume (h = 1/64, dt = h/2, no-slip box, source sphere of radius
6 cells at (32, 8, 32) cells adding 1 per step, velocity noise U(-0.01, 0.01) from splitmix64
seed 1) after 100 forward steps, pressure solved to a relative residual of 1e-10:
density (FP32) and per-step CG iterations. Temperature is not stored: with the same source, the
same zero-gradient boundaries and no colliders it equals the density exactly.

Usage: python tests/golden/make_golden.py [--threads T]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "diff-solver_amd")]

import oracle as O  # noqa: E402
from ds_smoke import scenes  # noqa: E402


def plume64_100(threads):
    sc = scenes.plume(64, with_target=False)
    O.set_threads(threads)
    o = O.Oracle(sc, tolerance=1e-10)
    state, ctl = scenes.initial_state(sc), scenes.control(sc)
    iters = []
    assert np.array_equal(ctl["density_source"], ctl["temperature_source"])
    t0 = time.time()
    for step in range(100):
        state, _, info = o.forward_step(state, ctl, sc.params)
        iters.append(info.iterations)
        if step % 10 == 9:
            print(f"step {step + 1}: {info.iterations} CG iterations, {time.time() - t0:.1f} s", flush=True)
    assert np.array_equal(state["density"], state["temperature"])
    np.savez_compressed(os.path.join(HERE, "plume64_100.npz"),
                        density=state["density"].astype(np.float32),
                        cg_iterations=np.array(iters, dtype=np.int32),
                        tolerance=np.float64(1e-10))


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--threads", type=int, default=os.cpu_count() or 1)
    args = ap.parse_args()
    plume64_100(args.threads)
