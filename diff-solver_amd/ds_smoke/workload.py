"""Timed fwd+VJP runs of a plume workload for each pressure-solve warm start, with checks that hold
at any size.

  python -m ds_smoke.workload --grid 256 --solids 8 --steps 32 --k 8 --modes none,previous,cnn \\
         --weights warm_start_256.npz --out result.json

For each mode: one untimed run, then `--repeats` timed runs of `steps` forward steps plus the
reverse sweep of 0.5 ||rho_T - target||^2 (ds_loss_gradient, checkpoint every k steps). Reports
fwd+VJP steps/s and the pressure-solve iterations per solve (forward, recompute, adjoint). Checks:
every solve converged; the gradients of all modes agree (each solve runs to the same tolerance);
and forward mode along a random velocity direction v gives the directional derivative that the
gradient gives (<rho_T - target, d rho_T> = <grad, v>), a size-independent check of the reverse
sweep against the separately written tangent.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np


def run(grid=256, solids=8, steps=32, k=8, modes=("none", "previous", "cnn"), weights=None, repeats=1, tolerance=1e-6, device=0,
        log=print):
    import torch

    from . import abi, scenes
    from .solver import Context

    sc = scenes.plume(grid, solids=solids)
    warm = {"none": abi.WARM_START_NONE, "previous": abi.WARM_START_PREVIOUS, "extrapolate": abi.WARM_START_EXTRAPOLATE, "projection": abi.WARM_START_PROJECTION,
            "cnn_alone": abi.WARM_START_CNN_ALONE, "cnn": abi.WARM_START_CNN}
    s0h, kh = scenes.initial_state(sc), scenes.control(sc)
    results, grads = {}, {}
    for mode in modes:
        ctx = Context(sc.config(tolerance=tolerance, warm_start=warm[mode]), abi.MODE_DIFFERENTIABLE, device)
        if mode in ("cnn", "cnn_alone"):
            from . import warm_start as ws

            ws.load_into(ctx, weights)
        s0, ctl = ctx.state_from_host(s0h), ctx.control_from_host(kh)
        params = ctx.allocate_parameters(sc.params)
        target = ctx.to_device(sc.target)
        log(f"{mode}: untimed run")
        ctx.loss_gradient(s0, ctl, params, steps, k, target)
        torch.cuda.synchronize(device)
        t0 = time.perf_counter()
        for _ in range(repeats):
            loss, grad, info = ctx.loss_gradient(s0, ctl, params, steps, k, target)
        torch.cuda.synchronize(device)
        elapsed = (time.perf_counter() - t0) / repeats
        per_solve = {"forward": info.forward_iterations / max(1, info.forward_solves),
                     "recompute": info.recompute_iterations / max(1, info.recompute_solves),
                     "adjoint": info.adjoint_iterations / max(1, info.adjoint_solves)}
        results[mode] = {"steps_per_s": steps / elapsed, "seconds": elapsed, "loss": loss, "cg_iterations_per_solve": per_solve,
                         "not_converged": info.not_converged, "guesses_rejected": info.guesses_rejected,
                         "max_relative_residual": info.max_relative_residual}
        grads[mode] = [v.double().cpu().numpy() for v in grad["velocity"]]
        log(json.dumps({"mode": mode, **results[mode]}))
        if mode == modes[0]:
            # forward mode along a random velocity direction against the gradient
            rng = np.random.default_rng(7)
            v = tuple(rng.standard_normal(x.shape).astype(np.float32) for x in s0h["velocity"])
            zero = lambda a: np.zeros_like(a)
            st_t = ctx.state_from_host({"density": zero(s0h["density"]), "temperature": zero(s0h["temperature"]), "velocity": v})
            ct_t = ctx.control_from_host({key: (tuple(zero(x) for x in val) if isinstance(val, tuple) else zero(val))
                                          for key, val in kh.items()})
            fin, fin_t, _ = ctx.trajectory_jvp(s0, ctl, params, steps, st_t, ct_t, ctx.allocate_parameters((0.0, 0.0, 0.0, 0.0)))
            rho = ctx.to_host(fin)["density"].astype(np.float64)
            fwd = float(np.sum((rho - sc.target) * ctx.to_host(fin_t)["density"].astype(np.float64)))
            rev = sum(float(np.sum(g * x.astype(np.float64))) for g, x in zip(grads[mode], v))
            results["forward_vs_reverse"] = {"forward_mode": fwd, "reverse_mode": rev, "relative_difference": abs(fwd - rev) / abs(rev)}
            log(json.dumps(results["forward_vs_reverse"]))
        ctx.close()
    ref = modes[0]
    for mode in modes[1:]:
        num = sum(np.sum((a - b) ** 2) for a, b in zip(grads[mode], grads[ref]))
        den = sum(np.sum(b ** 2) for b in grads[ref])
        results[mode]["gradient_relative_l2_vs_" + ref] = float(np.sqrt(num / den))
    domain = {"grid": grid, "solids": solids, "steps": steps, "checkpoint_interval": k, "cg_tolerance": tolerance}
    return {"workload": f"plume{grid}" + (f"_solids{solids}" if solids else "") + f"_fwd{steps}_vjp_checkpoint{k}", **domain,
            "modes": results}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--solids", type=int, default=8)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--k", type=int, default=8)
    ap.add_argument("--modes", default="none,previous,cnn")
    ap.add_argument("--weights", default=None)
    ap.add_argument("--repeats", type=int, default=1)
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    modes = tuple(m for m in args.modes.split(",") if m)
    if "cnn" in modes and not (args.weights and os.path.exists(args.weights)):
        modes = tuple(m for m in modes if m != "cnn")
        print("no warm-start weights: the cnn mode is skipped", file=sys.stderr, flush=True)
    res = run(args.grid, args.solids, args.steps, args.k, modes, args.weights, args.repeats,
              log=lambda s: print(s, file=sys.stderr, flush=True))
    line = json.dumps(res)
    print(line, flush=True)
    if args.out:
        with open(args.out, "w") as f:
            f.write(line + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
